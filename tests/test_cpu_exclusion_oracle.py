"""Pins of the oracle's PME exclusion correction (oracle.pme_exclusion_correction):
with it, real space over non-excluded pairs + PME reciprocal + self energy
equals the brute-force Ewald sum of all pairs (real space over every pair
and image, reciprocal as an explicit k-sum, self) minus the bare Coulomb
energy of the excluded pairs, for energies and forces of a molecular box;
its virial is the strain derivative of its energy; an isolated neutral
molecule in a large box has only a small image energy left."""
import math

import numpy as np
import pytest
from scipy.special import erfc

from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


def _molecular_box(seed=4):
    # a few waters and a short charged chain in a 30 A box
    s = systems.small_box(n_waters=40, n_chains=1, chain_len=10, box=30.0, seed=seed, cutoff=12.0,
                          switchdist=10.0, pairlistdist=13.5, pme_grid=(64, 64, 64))
    return s


def _brute_ewald_with_exclusions(pos, q, box, beta, excl_pairs, kmax):
    n = len(q)
    # real space over every pair, minimum image (erfc(beta L/2) is negligible)
    E = 0.0
    F = np.zeros((n, 3))
    for i in range(n):
        d = pos[i] - pos[i + 1:]
        d -= box * np.rint(d / box)
        r = np.sqrt((d * d).sum(1))
        qq = q[i] * q[i + 1:]
        E += (qq * erfc(beta * r) / r).sum()
        dedr = -qq * (erfc(beta * r) / r ** 2 + 2 * beta / math.sqrt(math.pi) * np.exp(-(beta * r) ** 2) / r)
        f = (-dedr / r)[:, None] * d
        F[i] += f.sum(0)
        F[i + 1:] -= f
    Er, Fr = orc.ewald_recip_direct(pos, q, np.diag(box), beta, kmax)
    E += Er - beta / math.sqrt(math.pi) * (q ** 2).sum()
    F += Fr
    # remove the bare Coulomb interaction of the excluded pairs
    for a, b in excl_pairs:
        d = pos[a] - pos[b]
        d -= box * np.rint(d / box)
        r = math.sqrt((d * d).sum())
        E -= q[a] * q[b] / r
        f = q[a] * q[b] / r ** 3 * d
        F[a] -= f
        F[b] += f
    return E, F


def _excl_pairs(s):
    out = set()
    for a in range(len(s.charge)):
        for k in range(s.excl_ptr[a], s.excl_ptr[a + 1]):
            b = int(s.excl_idx[k])
            out.add((min(a, b), max(a, b)))
    return sorted(out)


def test_ewald_identity_with_exclusions():
    s = _molecular_box()
    n = len(s.charge)
    q = s.charge * math.sqrt(COULOMB)
    box = s.box
    # beta large enough that the 12 A real-space cutoff drops < 1e-12 of the
    # energy, and the explicit k-sum runs to exp(-pi^2 m^2/beta^2) < 1e-13
    beta = 0.45
    K, order = (96, 96, 96), 8
    A, B = np.zeros((1, 1)), np.zeros((1, 1))
    atoms = np.zeros((n, 4), np.float32)
    # fractional coordinates on multiples of 2^-20, so that the spread's
    # float32 K*x is exact and every term sees the same positions
    frac = np.round(((s.pos / box) % 1.0) * 2.0 ** 20) / 2.0 ** 20 % 1.0
    atoms[:, :3] = frac
    atoms[:, 3] = q
    # positions the float32 fractional inputs represent exactly
    pos = atoms[:, :3].astype(np.float64) * box
    r = orc.nb_fp64(pos, q.astype(np.float32).astype(np.float64), np.zeros(n, np.int32), A, B, s.excl_ptr,
                    s.excl_idx, box, 12.0, 10.0, beta)
    qf = atoms[:, 3].astype(np.float64)
    recip = np.diag(1.0 / box)
    g = orc.pme_spread_gpu(atoms, K, order)
    phi, Erec, _ = orc.pme_recip(g, K, recip, float(np.prod(box)), beta, order)
    Frec = orc.pme_gather_gpu(atoms, K, order, phi, recip)
    Ecor, Fcor, _ = orc.pme_exclusion_correction(pos, qf, s.excl_ptr, s.excl_idx, box, beta, 12.0)
    Eself = -beta / math.sqrt(math.pi) * (qf ** 2).sum()
    E = r["e_elec"] + r["e_slow"] + Erec + Eself + Ecor
    F = r["f"] + r["fs"] + Frec + Fcor
    Eb, Fb = _brute_ewald_with_exclusions(pos, qf, box, beta, _excl_pairs(s), kmax=24)
    assert abs(E - Eb) / abs(Eb) < 1e-8
    assert np.sqrt(((F - Fb) ** 2).sum() / (Fb ** 2).sum()) < 1e-6
    # the correction is not small: leaving it out breaks the identity
    assert abs(Ecor) > 1e3 * abs(E - Eb)


def _whole_molecules(s):
    """Positions with every molecule (connected component of the exclusion
    graph) unwrapped by minimum-image steps from its first atom."""
    n = len(s.charge)
    pos = s.pos.copy()
    seen = np.zeros(n, bool)
    for root in range(n):
        if seen[root]:
            continue
        seen[root] = True
        stack = [root]
        while stack:
            a = stack.pop()
            for k in range(s.excl_ptr[a], s.excl_ptr[a + 1]):
                b = int(s.excl_idx[k])
                if not seen[b]:
                    d = s.pos[b] - pos[a]
                    pos[b] = pos[a] + d - s.box * np.rint(d / s.box)
                    seen[b] = True
                    stack.append(b)
    return pos


def test_exclusion_virial_is_strain_derivative():
    s = _molecular_box(seed=6)
    q = s.charge * math.sqrt(COULOMB)
    beta = 0.3
    # the correction depends on excluded-pair distances only: in a box far
    # larger than the system no pair wraps, so straining the coordinates
    # strains every pair vector
    box = np.array([500.0, 500.0, 500.0])
    pos = _whole_molecules(s) + 100.0
    E, F, W = orc.pme_exclusion_correction(pos, q, s.excl_ptr, s.excl_idx, box, beta, 12.0)
    E0 = orc.pme_exclusion_correction(s.pos, q, s.excl_ptr, s.excl_idx, s.box, beta, 12.0)[0]
    assert abs(E - E0) < 1e-9 * abs(E0)
    h = 1e-6
    Wfd = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            es = []
            for sgn in (1.0, -1.0):
                eps = np.zeros((3, 3))
                eps[a, b] = sgn * h
                M = np.eye(3) + eps
                es.append(orc.pme_exclusion_correction(pos @ M.T, q, s.excl_ptr, s.excl_idx, box, beta, 12.0)[0])
            Wfd[a, b] = -(es[0] - es[1]) / (2 * h)
    assert np.abs(W - Wfd).max() < 1e-6 * np.abs(W).max()
    # forces are the negative gradient
    for i in (0, 3, 7):
        for d in range(3):
            es = []
            for sgn in (1.0, -1.0):
                p = pos.copy()
                p[i, d] += sgn * 1e-5
                es.append(orc.pme_exclusion_correction(p, q, s.excl_ptr, s.excl_idx, box, beta, 12.0)[0])
            assert abs(-(es[0] - es[1]) / 2e-5 - F[i, d]) < 1e-6 * np.abs(F).max() + 1e-9


def test_isolated_molecule_leaves_image_energy_only():
    # one TIP3P-like water in a 60 A box: recip + self + correction is the
    # interaction with the periodic images only, O(mu^2 / L^3)
    pos = np.array([[30.0, 30.0, 30.0], [30.9572, 30.0, 30.0], [29.76, 30.927, 30.0]])
    q = np.array([-0.834, 0.417, 0.417]) * math.sqrt(COULOMB)
    box = np.array([60.0, 60.0, 60.0])
    beta, K, order = 0.3, (64, 64, 64), 8
    atoms = np.zeros((3, 4), np.float32)
    atoms[:, :3] = np.round(pos / box * 2.0 ** 20) / 2.0 ** 20
    atoms[:, 3] = q
    pos = atoms[:, :3].astype(np.float64) * box
    qf = atoms[:, 3].astype(np.float64)
    g = orc.pme_spread_gpu(atoms, K, order)
    _, Erec, _ = orc.pme_recip(g, K, np.diag(1.0 / box), float(np.prod(box)), beta, order)
    ptr = np.array([0, 2, 4, 6])
    idx = np.array([1, 2, 0, 2, 0, 1], np.int32)
    Ecor, _, _ = orc.pme_exclusion_correction(pos, qf, ptr, idx, box, beta, 12.0)
    Eself = -beta / math.sqrt(math.pi) * (qf ** 2).sum()
    mu2 = ((qf[:, None] * pos).sum(0) ** 2).sum()
    total = Erec + Eself + Ecor
    assert abs(Eself) > 50.0 and abs(Ecor) > 1.0
    assert abs(total) < 10.0 * mu2 / 60.0 ** 3
