"""Seeded synthetic molecular systems for the nonbonded + PME hot path.

NAMD reads PSF/PDB/parameter files (src/Molecule.C, src/Parameters.C); this
package has no file front-end, so every workload is generated here from a
seed.  A system is TIP3P water plus random-walk bead chains in an orthorhombic
periodic box:

* water: rOH 0.9572 A, HOH 104.52 deg, q_O -0.834, q_H +0.417, CHARMM TIP3P
  LJ (O: eps 0.1521, Rmin/2 1.7682; H: eps 0.046, Rmin/2 0.2245), random
  orientation from uniform quaternions, oxygen on a jittered simple-cubic
  lattice;
* chains: persistent random walks of 1.5 A bonds, bond angle >= 120 deg,
  soft self-avoidance, seeded LJ types (sigma U[2.5,4.0] A, eps U[0.02,0.2]
  kcal/mol), charges +-0.8 alternating so every chain is neutral;
* exclusions: all intra-water pairs, chain 1-2 and 1-3 pairs (NAMD
  ``exclude 1-3``, src/Molecule.C:3860-3900 builds the same full lists).
  1-4 pairs are ordinary nonbonded pairs, which equals ``scaled1-4`` with
  1-4scaling 1.0 and identical 1-4 LJ parameters.

Water sites that would overlap a chain bead are removed before waters are
placed; the removal radius shrinks until enough sites remain, so the result
is deterministic for a seed.
"""
from __future__ import annotations

import math
from typing import NamedTuple
from dataclasses import dataclass, field

import numpy as np

# TIP3P geometry / charges (CHARMM flavour, H carries LJ)
TIP3P_ROH = 0.9572
TIP3P_THETA = math.radians(104.52)
TIP3P_QO = -0.834
TIP3P_QH = 0.417
TIP3P_O_EPS, TIP3P_O_RMIN2 = 0.1521, 1.7682
TIP3P_H_EPS, TIP3P_H_RMIN2 = 0.046, 0.2245
TIP3P_RHH = 2.0 * TIP3P_ROH * math.sin(0.5 * TIP3P_THETA)
MASS_O, MASS_H = 15.9994, 1.008
# chain beads (a choice of this build, the configs leave them open): united
# CH2-like mass, harmonic bonds at the random walk's 1.5 A step, angles held
# at their built value
MASS_BEAD = 14.027
BEAD_BOND_K, BEAD_BOND_R0 = 300.0, 1.5       # kcal/mol/A^2, A
BEAD_ANGLE_K = 40.0                           # kcal/mol/rad^2
RMIN_TO_SIGMA = 2.0 / 2.0 ** (1.0 / 6.0)  # sigma = Rmin / 2^(1/6) = 2*(Rmin/2)/2^(1/6)


@dataclass
class NbParams:
    """Nonbonded / PME parameters (the SimParameters keys the path reads).

    Keys: cutoff (SimParameters.C:548), switchdist (:578), pairlistdist (:589),
    PMETolerance (:1054), PMEInterpOrder (:1056), PMEGridSize (:1058).
    """
    cutoff: float = 12.0
    switchdist: float = 10.0
    pairlistdist: float = 13.5
    pme_grid: tuple = (32, 32, 32)
    pme_order: int = 4
    ewald_tol: float = 1e-6
    full_elect_frequency: int = 1
    nonbonded_frequency: int = 1
    margin: float = 0.0


@dataclass
class System:
    name: str
    pos: np.ndarray          # (N,3) float64, Angstrom
    charge: np.ndarray       # (N,) float64, e
    vdw_type: np.ndarray     # (N,) int32
    lj_sigma: np.ndarray     # (ntypes,) float64
    lj_eps: np.ndarray       # (ntypes,) float64
    excl_ptr: np.ndarray     # (N+1,) int64, CSR of full exclusions (no self)
    excl_idx: np.ndarray     # (nexcl,) int32
    box: np.ndarray          # (3,) float64 orthorhombic edge lengths
    params: NbParams = field(default_factory=NbParams)
    seed: int = 0
    partition: np.ndarray | None = None  # (N,) int8 alchemical partition (0 = environment)
    n_waters: int = 0
    n_beads: int = 0
    chain_len: int = 0       # beads per chain (chains are stored first, chain by chain)
    lj_AB: tuple | None = None   # ordered (A, B) ntypes x ntypes tables given directly (e.g. read
                                 # from a dumpbench file) instead of mixing lj_sigma / lj_eps
    cell: np.ndarray | None = None   # (3,3) lattice rows a, b, c of a general periodic cell
                                     # (None: the orthorhombic box)

    @property
    def natoms(self) -> int:
        return int(self.pos.shape[0])

    def exclusion_pairs(self) -> np.ndarray:
        """Unique excluded pairs (i<j) as (M,2) int64."""
        i = np.repeat(np.arange(self.natoms, dtype=np.int64), np.diff(self.excl_ptr))
        j = self.excl_idx.astype(np.int64)
        m = i < j
        return np.stack([i[m], j[m]], axis=1)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from uniform unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random((3, n))
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    qw = a * np.sin(2 * np.pi * u2)
    qx = a * np.cos(2 * np.pi * u2)
    qy = b * np.sin(2 * np.pi * u3)
    qz = b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    R[:, 0, 1] = 2 * (qx * qy - qz * qw)
    R[:, 0, 2] = 2 * (qx * qz + qy * qw)
    R[:, 1, 0] = 2 * (qx * qy + qz * qw)
    R[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    R[:, 1, 2] = 2 * (qy * qz - qx * qw)
    R[:, 2, 0] = 2 * (qx * qz - qy * qw)
    R[:, 2, 1] = 2 * (qy * qz + qx * qw)
    R[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    return R


def _water_template() -> np.ndarray:
    h = 0.5 * TIP3P_THETA
    return np.array([
        [0.0, 0.0, 0.0],
        [TIP3P_ROH * math.sin(h), 0.0, TIP3P_ROH * math.cos(h)],
        [-TIP3P_ROH * math.sin(h), 0.0, TIP3P_ROH * math.cos(h)],
    ])


def _place_waters(rng, oxygen: np.ndarray) -> np.ndarray:
    R = _random_rotations(rng, oxygen.shape[0])
    tmpl = _water_template()
    return oxygen[:, None, :] + np.einsum("nij,aj->nai", R, tmpl)


def _random_unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


class _BeadGrid:
    """Spatial hash of placed beads (cell = avoidance radius) for self- and
    cross-chain avoidance."""

    def __init__(self, box, cell):
        self.box = box
        self.n = np.maximum(1, np.floor(box / cell).astype(int))
        self.cell = box / self.n
        self.cells = {}

    def _key(self, p):
        g = np.floor(np.mod(p, self.box) / self.cell).astype(int) % self.n
        return int(g[0]), int(g[1]), int(g[2])

    def add(self, p, tag):
        self.cells.setdefault(self._key(p), []).append((p.copy(), tag))

    def near(self, p, skip):
        gx, gy, gz = self._key(p)
        out = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    k = ((gx + dx) % self.n[0], (gy + dy) % self.n[1], (gz + dz) % self.n[2])
                    for q, tag in self.cells.get(k, ()):
                        if tag not in skip:
                            out.append(q)
        return np.array(out) if out else np.zeros((0, 3))


def _chain_walk(rng, start, nbeads, box, grid, chain_id, bond=1.5, min_cos=-0.5, avoid=2.2, tries=64):
    """One persistent random walk (bond angle >= 120 deg) avoiding earlier beads
    of this and previous chains (soft: the farthest candidate wins if none is clear)."""
    pts = np.empty((nbeads, 3))
    pts[0] = start
    prev_dir = _random_unit(rng, 1)[0]
    grid.add(pts[0], (chain_id, 0))
    for k in range(1, nbeads):
        cand = _random_unit(rng, tries)
        # bond angle >= 120 deg  <=>  dot(new, prev) >= cos(60 deg) = 0.5
        ok = cand @ prev_dir >= -min_cos if k > 1 else np.ones(tries, bool)
        cand = cand[ok] if ok.any() else cand
        newp = pts[k - 1] + bond * cand
        near = grid.near(pts[k - 1], skip={(chain_id, k - 1), (chain_id, k - 2)})
        if near.shape[0]:
            d = newp[:, None, :] - near[None, :, :]
            d -= box * np.round(d / box)
            dmin = np.sqrt((d * d).sum(-1)).min(axis=1)
            good = np.nonzero(dmin >= avoid)[0]
            pick = good[0] if good.size else int(np.argmax(dmin))
        else:
            pick = 0
        pts[k] = newp[pick]
        prev_dir = cand[pick]
        grid.add(pts[k], (chain_id, k))
    return pts


def _relax_contacts(rng, wat, base, bead_pos, boxv, jitter, dmin, iters=60):
    """Re-draw the orientation and jitter of waters that have an intermolecular
    contact closer than dmin (to another water or a bead), a few rounds; the
    lattice + jitter scheme is kept, only the random draws are repeated."""
    if dmin <= 0 or wat.shape[0] == 0:
        return wat
    from scipy.spatial import cKDTree
    nw = wat.shape[0]
    nb = bead_pos.shape[0]
    for _ in range(iters):
        pts = np.mod(np.concatenate([bead_pos, wat.reshape(-1, 3)], 0), boxv)
        tree = cKDTree(pts, boxsize=boxv)
        pairs = tree.query_pairs(dmin, output_type="ndarray")
        if pairs.size == 0:
            break
        mol = np.concatenate([-1 - np.arange(nb), np.repeat(np.arange(nw), 3)])
        a, b = mol[pairs[:, 0]], mol[pairs[:, 1]]
        inter = a != b
        bad = np.unique(np.concatenate([a[inter & (a >= 0)], b[inter & (b >= 0)]]))
        if bad.size == 0:
            break
        oxy = base[bad] + rng.uniform(-jitter, jitter, (bad.size, 3))
        wat[bad] = _place_waters(rng, oxy)
    return wat


def make_system(name: str, n_waters: int, n_chains: int, chain_len: int, box: float,
                n_bead_types: int, seed: int, params: NbParams | None = None,
                lattice_jitter: float = 0.3, lattice_shape=None, min_contact: float = 1.2,
                type_seed: int | None = None) -> System:
    """Water + chain box.  Atom order: chains first (bead 0..L-1 per chain), then waters O,H1,H2.
    type_seed: draw the bead LJ parameters from their own seed (systems of a
    batch share one parameter table)."""
    rng = np.random.default_rng(seed)
    boxv = np.array([box, box, box], dtype=np.float64)

    # ---- LJ types: 0..n_bead_types-1 beads, then water O, water H
    trng = rng if type_seed is None else np.random.default_rng(type_seed)
    sig = trng.uniform(2.5, 4.0, n_bead_types)
    eps = trng.uniform(0.02, 0.2, n_bead_types)
    lj_sigma = np.concatenate([sig, [TIP3P_O_RMIN2 * RMIN_TO_SIGMA, TIP3P_H_RMIN2 * RMIN_TO_SIGMA]])
    lj_eps = np.concatenate([eps, [TIP3P_O_EPS, TIP3P_H_EPS]])
    tO, tH = n_bead_types, n_bead_types + 1

    # ---- chains
    beads = []
    grid = _BeadGrid(boxv, 3.7)
    for c in range(n_chains):
        # a chain starts at least 2.2 A from every earlier bead (redrawn only on a clash)
        start = rng.random(3) * boxv
        for _ in range(200):
            near = grid.near(start, skip=set())
            if near.shape[0] == 0:
                break
            d = near - start
            d -= boxv * np.round(d / boxv)
            if np.sqrt((d * d).sum(-1)).min() >= 2.2:
                break
            start = rng.random(3) * boxv
        beads.append(_chain_walk(rng, start, chain_len, boxv, grid, c))
    bead_pos = np.concatenate(beads, 0) if beads else np.zeros((0, 3))
    n_beads = bead_pos.shape[0]
    bead_type = rng.integers(0, n_bead_types, n_beads).astype(np.int32) if n_beads else np.zeros(0, np.int32)
    bead_q = np.tile(np.where(np.arange(chain_len) % 2 == 0, 0.8, -0.8), n_chains) if n_beads else np.zeros(0)
    if chain_len % 2 == 1 and n_beads:
        # keep every chain neutral: odd length gets a zero-charge last bead
        bead_q = bead_q.reshape(n_chains, chain_len)
        bead_q[:, -1] = 0.0
        bead_q = bead_q.reshape(-1)

    # ---- water oxygen sites on a lattice, minus sites overlapping beads
    if lattice_shape is None:
        n = int(math.ceil((n_waters * 1.25) ** (1.0 / 3.0))) if n_beads else int(math.ceil(n_waters ** (1.0 / 3.0)))
        lattice_shape = (n, n, n)
    nx, ny, nz = lattice_shape
    a = boxv / np.array([max(nx, 1), max(ny, 1), max(nz, 1)])
    gi = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    sites = (gi + 0.5) * a
    if n_beads:
        from scipy.spatial import cKDTree
        tree = cKDTree(np.mod(bead_pos, boxv), boxsize=boxv)
        radius = 2.4
        while True:
            dist, _ = tree.query(sites, k=1)
            keep = dist >= radius
            if keep.sum() >= n_waters or radius < 0.5:
                break
            radius -= 0.1
        sites = sites[keep]
    if sites.shape[0] < n_waters:
        raise ValueError(f"{name}: only {sites.shape[0]} lattice sites for {n_waters} waters")
    choose = np.sort(rng.choice(sites.shape[0], n_waters, replace=False))
    base = sites[choose]
    oxy = base + rng.uniform(-lattice_jitter, lattice_jitter, (n_waters, 3))
    wat = _place_waters(rng, oxy)
    wat = _relax_contacts(rng, wat, base, bead_pos, boxv, lattice_jitter, min_contact)
    wat = wat.reshape(-1, 3)

    pos = np.concatenate([bead_pos, wat], 0)
    pos = np.mod(pos, boxv)
    N = pos.shape[0]
    charge = np.concatenate([bead_q, np.tile([TIP3P_QO, TIP3P_QH, TIP3P_QH], n_waters)])
    vdw_type = np.concatenate([bead_type, np.tile(np.array([tO, tH, tH], np.int32), n_waters)]).astype(np.int32)

    # ---- exclusions (full lists, no self)
    lists = [[] for _ in range(N)]
    for c in range(n_chains):
        base = c * chain_len
        for k in range(chain_len):
            i = base + k
            for d in (1, 2):
                if k + d < chain_len:
                    lists[i].append(i + d)
                    lists[i + d].append(i)
    w0 = n_beads
    for w in range(n_waters):
        o = w0 + 3 * w
        for (p, q) in ((o, o + 1), (o, o + 2), (o + 1, o + 2)):
            lists[p].append(q)
            lists[q].append(p)
    excl_ptr = np.zeros(N + 1, np.int64)
    excl_ptr[1:] = np.cumsum([len(l) for l in lists])
    excl_idx = np.array([j for l in lists for j in sorted(l)], dtype=np.int32)

    return System(name=name, pos=pos, charge=charge.astype(np.float64), vdw_type=vdw_type,
                  lj_sigma=lj_sigma, lj_eps=lj_eps, excl_ptr=excl_ptr, excl_idx=excl_idx,
                  box=boxv, params=params or NbParams(), seed=seed,
                  n_waters=n_waters, n_beads=n_beads, chain_len=chain_len if n_chains else 0)


# ---------------------------------------------------------------------------
# The named workloads
# ---------------------------------------------------------------------------

def water_box_9k(seed: int = 1) -> System:
    """9,000-atom TIP3P box: 3,000 waters on a 15x15x14 lattice subset, 31.0 A cube.

    cutoff 12, switchdist 10, pairlistdist 13.5, PME 32^3 order 4, tol 1e-6.
    The density is about three times liquid water, so close contacts are hard.
    """
    p = NbParams(cutoff=12.0, switchdist=10.0, pairlistdist=13.5, pme_grid=(32, 32, 32))
    return make_system("water9k", 3000, 0, 0, 31.0, 0, seed, p, lattice_shape=(15, 15, 14), min_contact=1.0)


def water_box_liquid(n_side=(15, 15, 14), seed: int = 1, pme_grid=(48, 48, 48)) -> System:
    """The 9,000-atom TIP3P box at liquid density: the same 3,000-site lattice
    as water_box_9k, box scaled by 3^(1/3) to 44.71 A (0.1007 atoms/A^3)."""
    box = round(31.0 * 3.0 ** (1.0 / 3.0), 2)
    p = NbParams(cutoff=12.0, switchdist=10.0, pairlistdist=13.5, pme_grid=tuple(pme_grid))
    return make_system("water9k_liquid", 3000, 0, 0, box, 0, seed, p, lattice_shape=tuple(n_side), min_contact=1.6)


def solvated_96k(seed: int = 2) -> System:
    """96,000 atoms: 26,000 TIP3P + 60 chains x 300 beads, 98 A cube, PME 108^3."""
    p = NbParams(cutoff=12.0, switchdist=10.0, pairlistdist=13.5, pme_grid=(108, 108, 108))
    return make_system("solvated96k", 26000, 60, 300, 98.0, 32, seed, p)


def solvated_1m(seed: int = 3) -> System:
    """1.065M atoms: 340,000 TIP3P + 150 chains x 300 beads, 216 A cube, PME 216^3."""
    p = NbParams(cutoff=12.0, switchdist=10.0, pairlistdist=13.5, pme_grid=(216, 216, 216))
    return make_system("solvated1m", 340000, 150, 300, 216.0, 32, seed, p)


def fit_23k(seed: int = 5) -> System:
    """23,000 atoms: 7,000 TIP3P + 20 chains x 100 beads (16 LJ types), 61.3 A cube."""
    p = NbParams(cutoff=12.0, switchdist=10.0, pairlistdist=13.5, pme_grid=(64, 64, 64))
    return make_system("fit23k", 7000, 20, 100, 61.3, 16, seed, p)


class BatchMember(NamedTuple):
    system: System
    partition: np.ndarray   # int8 per atom: 1 for the alchemical solute, 0 otherwise
    lam: float              # alchemical lambda of this system


def batch_systems(n_systems: int = 64, seed: int = 7, n_min: int = 3000, n_max: int = 40000,
                  density: float = 0.102, solute: int = 50, chain_len: int = 50, bead_fraction: float = 0.19,
                  n_bead_types: int = 32) -> list:
    """Independent systems of uneven size for one batched launch: atom counts
    round(exp(U(ln n_min, ln n_max))), each a water + chain box at the given
    atom density (cutoff 12, switch 10, pairlist 13.5) sharing one LJ table;
    the first `solute` beads (the first chain) form alchemical partition 1
    with a lambda drawn from U[0, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_systems):
        n_target = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        n_chains = max(1, int(round(bead_fraction * n_target / chain_len)))
        n_waters = max(1, (n_target - n_chains * chain_len) // 3)
        natoms = n_chains * chain_len + 3 * n_waters
        box = (natoms / density) ** (1.0 / 3.0)
        p = NbParams(cutoff=12.0, switchdist=10.0, pairlistdist=13.5,
                     pme_grid=tuple([int(2 * math.ceil(box / 2.0))] * 3))
        s = make_system(f"batch{k}", n_waters, n_chains, chain_len, box, n_bead_types, seed * 1000 + k, p,
                        type_seed=seed)
        part = np.zeros(s.natoms, np.int8)
        part[: min(solute, s.n_beads)] = 1
        out.append(BatchMember(s, part, float(rng.uniform(0.0, 1.0))))
    return out


def small_box(n_waters: int = 216, n_chains: int = 2, chain_len: int = 20, box: float = 20.0,
              n_bead_types: int = 4, seed: int = 11, cutoff: float = 8.0, switchdist: float = 7.0,
              pairlistdist: float = 9.0, pme_grid=(20, 20, 20)) -> System:
    """Small test box; the cutoff is scaled so that box >= 2 * pairlistdist."""
    p = NbParams(cutoff=cutoff, switchdist=switchdist, pairlistdist=pairlistdist, pme_grid=tuple(pme_grid))
    return make_system(f"small{n_waters}w{n_chains}c", n_waters, n_chains, chain_len, box, n_bead_types, seed, p)


MADELUNG_NACL = 1.747564594633182  # rock salt, per ion pair, in units of q^2 / nearest-neighbour distance


def nacl_crystal(nside: int = 8, spacing: float = 4.0, ewald_tol: float = 1e-8, pme_grid=(48, 48, 48),
                 pme_order: int = 8) -> System:
    """Rock-salt lattice of +-1 e ions (nside^3 sites, nside even, nearest
    neighbour `spacing` A), no Lennard-Jones, no exclusions: its total Ewald
    energy is -(N/2) * MADELUNG_NACL * COULOMB / spacing."""
    assert nside % 2 == 0
    idx = np.stack(np.meshgrid(*[np.arange(nside)] * 3, indexing="ij"), -1).reshape(-1, 3)
    n = idx.shape[0]
    L = nside * spacing
    p = NbParams(pme_grid=tuple(pme_grid), pme_order=pme_order, ewald_tol=ewald_tol)
    return System(name=f"nacl{nside}", pos=idx * spacing, charge=np.where(idx.sum(1) % 2 == 0, 1.0, -1.0),
                  vdw_type=np.zeros(n, np.int32), lj_sigma=np.array([1.0]), lj_eps=np.array([0.0]),
                  excl_ptr=np.zeros(n + 1, np.int64), excl_idx=np.zeros(0, np.int32), box=np.array([L, L, L]),
                  params=p)


def chain_bonded_terms(s: System):
    """Harmonic bonds (k, r0) and angles (k, theta0 = the built angle) of the
    chains: (bonds (nb, 2), bond_k, bond_r0, angles (na, 3), angle_k,
    angle_theta0)."""
    L = s.chain_len
    nch = s.n_beads // L if L else 0
    b, a = [], []
    for c in range(nch):
        o = c * L
        b += [(o + k, o + k + 1) for k in range(L - 1)]
        a += [(o + k, o + k + 1, o + k + 2) for k in range(L - 2)]
    bonds = np.array(b, np.int32).reshape(-1, 2)
    angles = np.array(a, np.int32).reshape(-1, 3)
    u = s.pos[angles[:, 0]] - s.pos[angles[:, 1]]
    v = s.pos[angles[:, 2]] - s.pos[angles[:, 1]]
    u -= s.box * np.round(u / s.box)
    v -= s.box * np.round(v / s.box)
    cs = (u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
    theta0 = np.arccos(np.clip(cs, -1.0, 1.0))
    return (bonds, np.full(len(bonds), BEAD_BOND_K), np.full(len(bonds), BEAD_BOND_R0), angles,
            np.full(len(angles), BEAD_ANGLE_K), theta0)


def sheared(s: System, cell) -> System:
    """The same molecules in a general periodic cell: fractional coordinates
    of the orthorhombic box carried onto lattice rows a, b, c (cell, 3 x 3),
    x = s_a a + s_b b + s_c c.  Bond geometry follows the shear; the
    exclusion lists are unchanged."""
    import dataclasses
    cell = np.asarray(cell, np.float64).reshape(3, 3)
    frac = s.pos / s.box
    return dataclasses.replace(s, name=s.name + "-triclinic", pos=frac @ cell, cell=cell)


def chain_pairs_14(s: System):
    """The chains' 1-4 pairs (bead k, bead k + 3), which the full exclusion
    lists (1-2, 1-3) leave as ordinary pairs, as an atom CSR (mod_ptr int64
    (N+1), mod_idx int32) listing each pair once from its lower atom: NAMD's
    modified exclusions under 'exclude scaled1-4'."""
    L = s.chain_len
    nch = s.n_beads // L if L else 0
    cnt = np.zeros(s.natoms + 1, np.int64)
    idx = []
    for c in range(nch):
        o = c * L
        for k in range(L - 3):
            cnt[o + k + 1] += 1
            idx.append(o + k + 3)
    return np.cumsum(cnt), np.array(idx, np.int32)


def lj14_tables(s: System, scale_a: float = 0.5, scale_b: float = 0.7):
    """1-4 LJ tables (A14, B14): the ordinary tables scaled (a stand-in for
    CHARMM's separate 1-4 parameters, LJTable's second entry)."""
    A, B = lj_tables(s)
    return A * scale_a, B * scale_b


def lj_tables(s: System):
    """Ordered LJ tables (A, B), ntypes x ntypes: the System's own when it
    carries them, else Lorentz-Berthelot mixing of lj_sigma / lj_eps."""
    if s.lj_AB is not None:
        return np.asarray(s.lj_AB[0], np.float64), np.asarray(s.lj_AB[1], np.float64)
    from . import lib as L
    return L.lj_mix(s.lj_sigma, s.lj_eps)


def masses(s: System) -> np.ndarray:
    return np.concatenate([np.full(s.n_beads, MASS_BEAD), np.tile([MASS_O, MASS_H, MASS_H], s.n_waters)])


def ewald_coefficient(cutoff: float, tol: float) -> float:
    """PMEEwaldCoefficient: bisection on erfc(beta*rc)/rc < tol (SimParameters.C:4243-4256)."""
    from math import erfc
    b = 1.0
    while erfc(b * cutoff) / cutoff >= tol:
        b *= 2.0
    lo, hi = 0.0, b
    for _ in range(100):
        b = 0.5 * (lo + hi)
        if erfc(b * cutoff) / cutoff >= tol:
            lo = b
        else:
            hi = b
    return b
