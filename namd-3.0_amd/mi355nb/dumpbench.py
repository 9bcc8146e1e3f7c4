"""NAMD's dumpbench text format (src/DumpBench.C:26-311; parameter list
src/DumpBenchParams.h): the input of the nonbonded path as NAMD's
`dumpbench <file>` Tcl command writes it on a serial run (src/ScriptTcl.C:
1901-1925), read into a System and written from one, so that the oracle and
the GPU path consume one file and a NAMD user can run a real system through
the path.

Sections, in order:
  SIMPARAMETERS   one `name value` line per parameter of DumpBenchParams.h
                  (int as %d, BigReal as %g, Vector as %f %f %f)
  LJTABLE         table_dim, then for i <= j: A B A14 B14 (%g)
  MOLECULE        numAtoms numCalcExclusions, then per atom
                  `vdw min max flags...`: min/max the absolute index range
                  of the atom's exclusion partners (numAtoms -1 if none) and
                  one flag per index of the range, 0 none, 1 full exclusion
                  (EXCHCK_FULL), 2 modified (EXCHCK_MOD); rows whose
                  exclusions fill the range (the atom itself aside) share an
                  all-1 array, so the atom's own index then reads 1
                  (src/Molecule.C:4063-4121)
  PATCHLIST       numPatches, then per patch PATCH_BEGIN, numAtoms and
                  `x y z q id hydrogenGroupSize ngia atomFixed groupFixed
                  partition` per atom (%f for reals), PATCH_END
  COMPUTEPAIR     number of pair computes, then `pid1 trans1 pid2 trans2`,
                  trans = Lattice::index(i, j, k) = 9(k+1) + 3(j+1) + (i+1)

What this build reads: any right-handed cell (a general one comes back as
System.cell, which ComputeNonbonded lays out with assign_patches_lattice),
exclude <= 1-3 or scaled1-4
(the modified pairs, scale14 and the 1-4 LJ table come back in info as
mod_ptr / mod_idx / scale14 / A14 / B14, the arguments of
ComputeNonbonded.set_modified_exclusions), no fixed atoms, no LES or
pair-interaction modes.  Anything else raises ValueError.
"""
from __future__ import annotations

import io
import math

import numpy as np

from . import systems as S

# src/DumpBenchParams.h, in order, with the printf kind of each
PARAMS = [
    ("cellBasisVector1", "v"), ("cellBasisVector2", "v"), ("cellBasisVector3", "v"), ("cellOrigin", "v"),
    ("nonbondedFrequency", "i"), ("fullElectFrequency", "i"), ("dielectric", "g"), ("exclude", "i"),
    ("scale14", "g"), ("nonbondedScaling", "g"), ("cutoff", "g"), ("switchingActive", "i"),
    ("switchingDist", "g"), ("pairlistDist", "g"), ("alchOn", "i"), ("alchLambda", "g"), ("alchLambda2", "g"),
    ("alchOutFreq", "i"), ("alchEquilSteps", "i"), ("alchVdwShiftCoeff", "g"), ("pressureProfileOn", "i"),
    ("pressureProfileSlabs", "i"), ("pressureProfileFreq", "i"), ("lesOn", "i"), ("lesFactor", "i"),
    ("pairInteractionOn", "i"), ("pairInteractionGroup1", "i"), ("pairInteractionGroup2", "i"),
    ("pairInteractionSelf", "i"), ("fixedAtomsOn", "i"), ("fixedAtomsForces", "i"), ("FMAOn", "i"),
    ("fullDirectOn", "i"), ("PMEOn", "i"), ("PMETolerance", "g"), ("PMEEwaldCoefficient", "g"),
    ("longSplitting", "i"), ("hgroupCutoff", "g"), ("mollyOn", "i"), ("commOnly", "i"),
]
EXCLUDE_ONETHREE = 2     # src/SimParameters.h:40-44
EXCLUDE_SCALED14 = 4
EXCHCK_FULL = 1          # src/Molecule.h:86-87
EXCHCK_MOD = 2


def lattice_index(i: int, j: int, k: int) -> int:
    """Lattice::index (src/Lattice.h:25-28)."""
    return 9 * (k + 1) + 3 * (j + 1) + (i + 1)


def _fmt_g(x: float) -> str:
    return "%g" % x


def _patch_grid(s: S.System):
    """NAMD-style patch grid of the build's layout: floor(box / (pairlistdist
    + margin)) patches per edge (a general cell: per lattice direction, from
    its perpendicular width), atoms binned by their wrapped position."""
    p = s.params
    if s.cell is not None:
        cell = np.asarray(s.cell, np.float64)
        inv = np.linalg.inv(cell)
        widths = 1.0 / np.linalg.norm(inv, axis=0)
        dims = np.maximum(1, np.floor(widths / (p.pairlistdist + p.margin)).astype(int))
        frac = np.mod(s.pos @ inv, 1.0)
        g = np.minimum(dims - 1, np.floor(frac * dims).astype(int))
        pid = (g[:, 0] * dims[1] + g[:, 1]) * dims[2] + g[:, 2]
        return dims, pid
    box = np.asarray(s.box, np.float64)
    dims = np.maximum(1, np.floor(box / (p.pairlistdist + p.margin)).astype(int))
    w = np.mod(s.pos, box)
    g = np.minimum(dims - 1, np.floor(w / (box / dims)).astype(int))
    pid = (g[:, 0] * dims[1] + g[:, 1]) * dims[2] + g[:, 2]
    return dims, pid


def _exclusion_rows(s: S.System, mod_ptr=None, mod_idx=None):
    """Per atom: (min, max, flags) as NAMD's ExclusionCheck holds them;
    modified partners (atom CSR mod_ptr / mod_idx) flagged EXCHCK_MOD."""
    n = s.natoms
    partners = [dict() for _ in range(n)]
    for a in range(n):
        for b in s.excl_idx[s.excl_ptr[a]:s.excl_ptr[a + 1]]:
            partners[a][int(b)] = EXCHCK_FULL
            partners[int(b)][a] = EXCHCK_FULL
    if mod_ptr is not None:
        for a in range(n):
            for b in mod_idx[mod_ptr[a]:mod_ptr[a + 1]]:
                partners[a][int(b)] = EXCHCK_MOD
                partners[int(b)][a] = EXCHCK_MOD
    rows = []
    for a in range(n):
        if not partners[a]:
            rows.append((n, -1, []))
            continue
        lo, hi = min(partners[a]), max(partners[a])
        size = hi - lo + 1
        in_middle = 1 if lo < a < hi else 0
        all_full = all(v == EXCHCK_FULL for v in partners[a].values())
        if all_full and size == len(partners[a]) + in_middle:
            flags = [EXCHCK_FULL] * size          # the shared all-full array
        else:
            flags = [partners[a].get(lo + k, 0) for k in range(size)]
        rows.append((lo, hi, flags))
    return rows


def write(f, s: S.System, A=None, B=None, alch=None, modified=None):
    """Write System s in dumpbench format to a path or text file object.
    A, B: ordered ntypes x ntypes LJ tables (default: the System's own);
    alch: optional dict alchLambda, alchLambda2, alchVdwShiftCoeff;
    modified: optional dict mod_ptr, mod_idx, scale14, A14, B14 (scaled 1-4
    exclusions; written as exclude scaled1-4)."""
    if isinstance(f, (str, bytes)) or hasattr(f, "__fspath__"):
        with open(f, "w") as fh:
            return write(fh, s, A, B, alch, modified)
    p = s.params
    if A is None or B is None:
        A, B = S.lj_tables(s)
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    nt = A.shape[0]
    box = np.asarray(s.box, np.float64)
    beta = S.ewald_coefficient(p.cutoff, p.ewald_tol)
    alch = alch or {}
    if modified is not None:
        A14 = np.asarray(modified["A14"], np.float64)
        B14 = np.asarray(modified["B14"], np.float64)
    else:
        A14, B14 = A, B
    cellm = np.asarray(s.cell, np.float64) if s.cell is not None else np.diag(box)
    vals = {
        "cellBasisVector1": tuple(cellm[0]), "cellBasisVector2": tuple(cellm[1]),
        "cellBasisVector3": tuple(cellm[2]), "cellOrigin": tuple(0.5 * cellm.sum(0)),
        "nonbondedFrequency": p.nonbonded_frequency, "fullElectFrequency": p.full_elect_frequency,
        "dielectric": 1.0, "exclude": EXCLUDE_SCALED14 if modified is not None else EXCLUDE_ONETHREE,
        "scale14": float(modified["scale14"]) if modified is not None else 1.0, "nonbondedScaling": 1.0,
        "cutoff": p.cutoff, "switchingActive": 1, "switchingDist": p.switchdist, "pairlistDist": p.pairlistdist,
        "alchOn": 1 if s.partition is not None else 0, "alchLambda": alch.get("alchLambda", 0.0),
        "alchLambda2": alch.get("alchLambda2", 0.0), "alchOutFreq": 0, "alchEquilSteps": 0,
        "alchVdwShiftCoeff": alch.get("alchVdwShiftCoeff", 0.0), "pressureProfileOn": 0,
        "pressureProfileSlabs": 0, "pressureProfileFreq": 0, "lesOn": 0, "lesFactor": 0, "pairInteractionOn": 0,
        "pairInteractionGroup1": 0, "pairInteractionGroup2": 0, "pairInteractionSelf": 0, "fixedAtomsOn": 0,
        "fixedAtomsForces": 0, "FMAOn": 0, "fullDirectOn": 0, "PMEOn": 1, "PMETolerance": p.ewald_tol,
        "PMEEwaldCoefficient": beta, "longSplitting": 0, "hgroupCutoff": 2.5, "mollyOn": 0, "commOnly": 0,
    }
    out = ["SIMPARAMETERS_BEGIN"]
    for name, kind in PARAMS:
        v = vals[name]
        if kind == "v":
            out.append("%s %f %f %f" % (name, v[0], v[1], v[2]))
        elif kind == "i":
            out.append("%s %d" % (name, int(v)))
        else:
            out.append("%s %s" % (name, _fmt_g(float(v))))
    out.append("SIMPARAMETERS_END")
    out.append("LJTABLE_BEGIN")
    out.append("%d" % nt)
    for i in range(nt):
        for j in range(i, nt):
            out.append("%s %s %s %s" % (_fmt_g(A[i, j]), _fmt_g(B[i, j]), _fmt_g(A14[i, j]), _fmt_g(B14[i, j])))
    out.append("LJTABLE_END")
    rows = _exclusion_rows(s, *((modified["mod_ptr"], modified["mod_idx"]) if modified is not None else ()))
    # numCalcExclusions: the unique excluded pairs, full and modified
    npairs = sum(sum(1 for k, fl in enumerate(fr) if fl and lo + k > a) for a, (lo, hi, fr) in enumerate(rows))
    out.append("MOLECULE_BEGIN")
    out.append("%d %d" % (s.natoms, npairs))
    for a, (lo, hi, flags) in enumerate(rows):
        out.append(" ".join(["%d %d %d" % (int(s.vdw_type[a]), lo, hi)] + ["%d" % x for x in flags]))
    out.append("MOLECULE_END")
    dims, pid = _patch_grid(s)
    np_ = int(np.prod(dims))
    part = s.partition if s.partition is not None else np.zeros(s.natoms, np.int8)
    hgs = np.ones(s.natoms, int)
    first_water = s.n_beads
    if s.n_waters:
        hgs[first_water:] = 0
        hgs[first_water::3] = 3
    out.append("PATCHLIST_BEGIN")
    out.append("%d" % np_)
    for p_ in range(np_):
        members = np.nonzero(pid == p_)[0]
        out.append("PATCH_BEGIN")
        out.append("%d" % len(members))
        for a in members:
            x = s.pos[a]
            out.append("%f %f %f %f %d %d %d %d %d %d" % (x[0], x[1], x[2], s.charge[a], a, hgs[a], 0, 0, 0,
                                                          int(part[a])))
        out.append("PATCH_END")
    out.append("PATCHLIST_END")
    pairs = []
    for gx in range(dims[0]):
        for gy in range(dims[1]):
            for gz in range(dims[2]):
                pa = (gx * dims[1] + gy) * dims[2] + gz
                for d in [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]:
                    if (d[0] * 3 + d[1]) * 3 + d[2] <= 0:
                        continue
                    g1 = (gx + d[0], gy + d[1], gz + d[2])
                    q = [math.floor(g1[k] / dims[k]) for k in range(3)]
                    g2 = [g1[k] - q[k] * dims[k] for k in range(3)]
                    pb = (g2[0] * dims[1] + g2[1]) * dims[2] + g2[2]
                    pairs.append((pa, lattice_index(0, 0, 0), pb, lattice_index(*q)))
    out.append("COMPUTEPAIR_BEGIN")
    out.append("%d" % len(pairs))
    out.extend("%d %d %d %d" % t for t in pairs)
    out.append("COMPUTEPAIR_END")
    f.write("\n".join(out) + "\n")


def _smooth_at_least(n: int) -> int:
    """Smallest m >= n whose prime factors are 2, 3 and 5."""
    m = max(1, n)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def read(f, name="dumpbench", pme_grid=None, pme_order=4):
    """Parse a dumpbench file (path or text) into (System, info).  info holds
    the parameters, the LJ table (ordered A, B and the 1-4 A14, B14), the
    patch list (atom ids per patch) and the pair computes.  pme_grid is not
    in the format: by default the smallest 2,3,5-smooth size of at least one
    point per Angstrom along each edge."""
    if isinstance(f, (str, bytes)) or hasattr(f, "__fspath__"):
        text = open(f).read() if not (isinstance(f, str) and "\n" in f) else f
    else:
        text = f.read()
    lines = [ln.strip() for ln in io.StringIO(text) if ln.strip()]
    pos_ = 0

    def expect(tag):
        nonlocal pos_
        if lines[pos_] != tag:
            raise ValueError(f"dumpbench: expected {tag} at line {pos_ + 1}, found {lines[pos_]!r}")
        pos_ += 1

    def nxt():
        nonlocal pos_
        ln = lines[pos_]
        pos_ += 1
        return ln

    expect("SIMPARAMETERS_BEGIN")
    prm = {}
    while lines[pos_] != "SIMPARAMETERS_END":
        parts = nxt().split()
        vals = [float(v) for v in parts[1:]]
        prm[parts[0]] = tuple(vals) if len(vals) > 1 else vals[0]
    expect("SIMPARAMETERS_END")
    for key in ("cellBasisVector1", "cellBasisVector2", "cellBasisVector3", "cutoff", "pairlistDist"):
        if key not in prm:
            raise ValueError(f"dumpbench: parameter {key} missing")
    basis = np.array([prm["cellBasisVector1"], prm["cellBasisVector2"], prm["cellBasisVector3"]], np.float64)
    if not np.linalg.det(basis) > 0:
        raise ValueError("dumpbench: the cell basis vectors must be right-handed with positive volume")
    general = np.abs(basis - np.diag(np.diag(basis))).max() > 0
    for key in ("fixedAtomsOn", "lesOn", "pairInteractionOn", "FMAOn", "fullDirectOn", "mollyOn"):
        if prm.get(key, 0):
            raise ValueError(f"dumpbench: {key} is not supported")
    if prm.get("dielectric", 1.0) != 1.0 or prm.get("nonbondedScaling", 1.0) != 1.0:
        raise ValueError("dumpbench: dielectric and nonbondedScaling must be 1")
    exclude = int(prm.get("exclude", EXCLUDE_ONETHREE))
    scale14 = float(prm.get("scale14", 1.0))

    expect("LJTABLE_BEGIN")
    nt = int(nxt())
    A = np.zeros((nt, nt))
    B = np.zeros((nt, nt))
    A14 = np.zeros((nt, nt))
    B14 = np.zeros((nt, nt))
    for i in range(nt):
        for j in range(i, nt):
            a, b, a14, b14 = (float(v) for v in nxt().split())
            A[i, j] = A[j, i] = a
            B[i, j] = B[j, i] = b
            A14[i, j] = A14[j, i] = a14
            B14[i, j] = B14[j, i] = b14
    expect("LJTABLE_END")

    expect("MOLECULE_BEGIN")
    natoms, _ = (int(v) for v in nxt().split())
    vdw = np.zeros(natoms, np.int32)
    ptr = np.zeros(natoms + 1, np.int64)
    idx = []
    modified = 0
    mptr = np.zeros(natoms + 1, np.int64)
    midx = []
    for a in range(natoms):
        parts = [int(v) for v in nxt().split()]
        vdw[a], lo, hi = parts[0], parts[1], parts[2]
        flags = parts[3:]
        if lo <= hi and len(flags) != hi - lo + 1:
            raise ValueError(f"dumpbench: atom {a}: {len(flags)} exclusion flags for range {lo}..{hi}")
        row = []
        for k, fl in enumerate(flags):
            b = lo + k
            if b == a or fl == 0:
                continue
            if fl == EXCHCK_FULL:
                row.append(b)
            elif fl == EXCHCK_MOD:
                modified += 1
                midx.append(b)
            else:
                raise ValueError(f"dumpbench: atom {a}: unknown exclusion flag {fl}")
        idx.extend(sorted(row))
        ptr[a + 1] = len(idx)
        mptr[a + 1] = len(midx)
    expect("MOLECULE_END")
    if modified and exclude != EXCLUDE_SCALED14:
        raise ValueError("dumpbench: modified (1-4) exclusions in a file whose exclude is not scaled1-4")
    if int(vdw.max(initial=0)) >= nt or int(vdw.min(initial=0)) < 0:
        raise ValueError("dumpbench: vdw type outside the LJ table")

    expect("PATCHLIST_BEGIN")
    npatch = int(nxt())
    pos = np.full((natoms, 3), np.nan)
    charge = np.zeros(natoms)
    hgs = np.zeros(natoms, np.int32)
    part = np.zeros(natoms, np.int8)
    seen = np.zeros(natoms, bool)
    patches = []
    for _ in range(npatch):
        expect("PATCH_BEGIN")
        na = int(nxt())
        ids = []
        for _ in range(na):
            t = nxt().split()
            aid = int(t[4])
            if not 0 <= aid < natoms or seen[aid]:
                raise ValueError(f"dumpbench: bad or repeated atom id {aid}")
            seen[aid] = True
            pos[aid] = (float(t[0]), float(t[1]), float(t[2]))
            charge[aid] = float(t[3])
            hgs[aid] = int(t[5])
            if int(t[7]) or int(t[8]):
                raise ValueError("dumpbench: fixed atoms are not supported")
            part[aid] = int(t[9])
            ids.append(aid)
        patches.append(np.array(ids, np.int64))
        expect("PATCH_END")
    expect("PATCHLIST_END")
    if not seen.all():
        raise ValueError(f"dumpbench: {int((~seen).sum())} atoms missing from the patch list")
    computes = []
    if pos_ < len(lines) and lines[pos_] == "COMPUTEPAIR_BEGIN":
        pos_ += 1
        for _ in range(int(nxt())):
            computes.append(tuple(int(v) for v in nxt().split()))
        expect("COMPUTEPAIR_END")

    # single atoms first, then three-atom hydrogen groups (O, H, H): the
    # build's chain + water order, which the integrator's SETTLE relies on
    n_single = int(np.argmax(hgs != 1)) if (hgs != 1).any() else natoms
    rest = hgs[n_single:]
    waters = rest.size % 3 == 0 and np.array_equal(rest, np.tile([3, 0, 0], rest.size // 3))
    n_beads, n_waters = (n_single, rest.size // 3) if waters else (0, 0)
    box = np.linalg.norm(basis, axis=1) if general else np.diag(basis).copy()
    cutoff = float(prm["cutoff"])
    switchdist = float(prm.get("switchingDist", cutoff)) if prm.get("switchingActive", 1) else cutoff
    grid = tuple(pme_grid) if pme_grid is not None else tuple(_smooth_at_least(int(math.ceil(L))) for L in box)
    params = S.NbParams(cutoff=cutoff, switchdist=switchdist, pairlistdist=float(prm["pairlistDist"]),
                        pme_grid=grid, pme_order=pme_order, ewald_tol=float(prm.get("PMETolerance", 1e-6)),
                        full_elect_frequency=int(prm.get("fullElectFrequency", 1)),
                        nonbonded_frequency=int(prm.get("nonbondedFrequency", 1)))
    s = S.System(name=name, pos=pos, charge=charge, vdw_type=vdw, lj_sigma=np.zeros(nt), lj_eps=np.zeros(nt),
                 excl_ptr=ptr, excl_idx=np.asarray(idx, np.int32), box=box, params=params,
                 partition=part if prm.get("alchOn", 0) else None, n_waters=n_waters, n_beads=n_beads,
                 lj_AB=(A, B), cell=basis if general else None)
    info = dict(params=prm, A=A, B=B, A14=A14, B14=B14, patches=patches, computes=computes,
                modified_pairs=modified // 2, hydrogen_group_size=hgs, mod_ptr=mptr,
                mod_idx=np.asarray(midx, np.int32), scale14=scale14)
    return s, info
