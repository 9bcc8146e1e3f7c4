"""Pins of the oracle's PME reciprocal space (oracle.pme_b_moduli,
pme_scalar_sum, pme_recip) by what the mathematics fixes: the order-4 closed
form of the B-spline moduli and the explicit B-spline sum for higher orders;
the brute-force Ewald reciprocal sum (no grid) for energy and forces in
orthorhombic and triclinic cells; the virial as the strain derivative of the
energy; and the rock-salt Madelung constant for the full Ewald sum (real
space from the FP64 pair oracle + reciprocal + self).  Also checks the
library's host-side B-spline moduli against the oracle's (no GPU needed)."""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


def bspline_explicit(n, u):
    """M_n(u) = 1/(n-1)! sum_k (-1)^k C(n,k) (u-k)_+^(n-1)."""
    s = 0.0
    for k in range(n + 1):
        if u - k > 0:
            s += (-1) ** k * math.comb(n, k) * (u - k) ** (n - 1)
    return s / math.factorial(n - 1)


@pytest.mark.parametrize("order", [4, 6, 8])
def test_spline_values_at_integers(order):
    M = orc.pme_spline_at_integers(order)
    want = [bspline_explicit(order, j + 1) for j in range(order)]
    assert np.allclose(M, want, rtol=0, atol=1e-15)
    assert abs(M.sum() - 1.0) < 1e-14


@pytest.mark.parametrize("K", [20, 32, 45, 108])
def test_b_moduli_closed_form_order4(K):
    th = 2 * np.pi * np.arange(K) / K
    assert np.allclose(orc.pme_b_moduli(K, 4), 9.0 / (2.0 + np.cos(th)) ** 2, rtol=1e-13, atol=0)


@pytest.mark.parametrize("K,order", [(32, 4), (108, 4), (30, 6), (44, 8), (28, 4)])
def test_library_b_moduli_match_oracle(K, order):
    assert np.allclose(L.pme_b_moduli(K, order), orc.pme_b_moduli(K, order), rtol=1e-13, atol=0)


def _random_cell_system(cell, n, seed):
    rng = np.random.default_rng(seed)
    frac = rng.random((n, 3)).astype(np.float32)
    q = rng.choice([-1.0, 1.0], n)
    q -= q.mean()
    atoms = np.zeros((n, 4), np.float32)
    atoms[:, :3] = frac
    atoms[:, 3] = q * math.sqrt(COULOMB)
    pos = frac.astype(np.float64) @ cell
    return atoms, pos, atoms[:, 3].astype(np.float64)


def _oracle_pme(atoms, K, order, cell, beta):
    recip = np.linalg.inv(cell).T
    vol = abs(np.linalg.det(cell))
    g = orc.pme_spread_gpu(atoms, K, order)
    phi, E, vir = orc.pme_recip(g, K, recip, vol, beta, order)
    F = orc.pme_gather_gpu(atoms, K, order, phi, recip)
    return E, vir, F, g


CELLS = {
    "ortho": np.diag([20.0, 22.0, 24.0]),
    "triclinic": np.array([[20.0, 0.0, 0.0], [3.0, 21.0, 0.0], [-2.0, 4.0, 23.0]]),
}


@pytest.mark.parametrize("cell", ["ortho", "triclinic"])
def test_recip_matches_brute_force_ewald(cell):
    cellm = CELLS[cell]
    atoms, pos, q = _random_cell_system(cellm, 40, 3)
    beta = 0.35
    Ed, Fd = orc.ewald_recip_direct(pos, q, cellm, beta, kmax=14)
    # order 8 on a ~0.5 A grid: PME converges to the Ewald sum
    E, vir, F, _ = _oracle_pme(atoms, (40, 44, 48), 8, cellm, beta)
    assert abs(E - Ed) / abs(Ed) < 1e-7
    assert np.sqrt(((F - Fd) ** 2).sum() / (Fd ** 2).sum()) < 1e-5
    # order 4 on the same grid: the usual PME accuracy, and it converges with the grid
    E4, _, F4, _ = _oracle_pme(atoms, (40, 44, 48), 4, cellm, beta)
    E4c, _, F4c, _ = _oracle_pme(atoms, (20, 22, 24), 4, cellm, beta)
    e_fine = abs(E4 - Ed) / abs(Ed)
    e_coarse = abs(E4c - Ed) / abs(Ed)
    assert e_fine < 1e-4 and e_fine < 0.1 * e_coarse


@pytest.mark.parametrize("cell", ["ortho", "triclinic"])
def test_recip_virial_is_strain_derivative(cell):
    cellm = CELLS[cell]
    atoms, _, _ = _random_cell_system(cellm, 30, 5)
    beta, K, order = 0.33, (20, 22, 24), 6
    E, vir, _, g = _oracle_pme(atoms, K, order, cellm, beta)
    # fractional coordinates fixed (the charge grid does not change); the
    # cell is strained, cell' = cell (I + eps)^T
    h = 1e-6
    W = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            es = []
            for sgn in (1.0, -1.0):
                eps = np.zeros((3, 3))
                eps[a, b] = sgn * h
                c2 = cellm @ (np.eye(3) + eps).T
                _, e2, _ = orc.pme_recip(g, K, np.linalg.inv(c2).T, abs(np.linalg.det(c2)), beta, order)
                es.append(e2)
            W[a, b] = -(es[0] - es[1]) / (2 * h)
    want = np.array([W[0, 0], W[0, 1], W[0, 2], W[1, 1], W[1, 2], W[2, 2]])
    assert np.abs(vir - want).max() / np.abs(want).max() < 1e-7
    # the strain derivative is symmetric
    assert abs(W[0, 1] - W[1, 0]) < 1e-6 * abs(W).max()


def test_forces_are_energy_gradient():
    cellm = CELLS["ortho"]
    atoms, _, _ = _random_cell_system(cellm, 12, 9)
    K, order, beta = (20, 22, 24), 4, 0.3
    _, _, F, _ = _oracle_pme(atoms, K, order, cellm, beta)
    recip = np.linalg.inv(cellm).T
    vol = abs(np.linalg.det(cellm))
    h = 1e-4
    for i in (0, 5, 11):
        for d in range(3):
            es = []
            for sgn in (1.0, -1.0):
                a2 = atoms.astype(np.float64)
                pos = a2[:, :3] @ cellm
                pos[i, d] += sgn * h
                # float32 fractional inputs would lose the step: spread in FP64
                es.append(_energy_fp64_frac(pos @ recip.T, a2[:, 3], K, order, recip, vol, beta))
            assert abs(-(es[0] - es[1]) / (2 * h) - F[i, d]) < 2e-4 * np.abs(F).max()


def _energy_fp64_frac(frac, q, K, order, recip, vol, beta):
    # GPU-convention spread in FP64 from unrounded fractional coordinates,
    # with the oracle's spline weights (same grid convention)
    K1, K2, K3 = K
    g = np.zeros((K3, K2, K1))
    for (x, y, z), qq in zip(frac, q):
        fr = [K1 * (x % 1.0), K2 * (y % 1.0), K3 * (z % 1.0)]
        fi = [int(f) for f in fr]
        th = [orc.bspline_gpu(f - i, order)[0] for f, i in zip(fr, fi)]
        for c in range(order):
            for b in range(order):
                for a in range(order):
                    g[(fi[2] + c) % K3, (fi[1] + b) % K2, (fi[0] + a) % K1] += qq * th[0][a] * th[1][b] * th[2][c]
    _, E, _ = orc.pme_recip(g, K, recip, vol, beta, order)
    return E


@pytest.mark.parametrize("TOL,bound", [(1e-12, 1e-8), (1e-8, 2e-7)])
def test_madelung_constant_nacl(TOL, bound):
    s = systems.nacl_crystal(nside=8, spacing=4.0, ewald_tol=TOL, pme_grid=(48, 48, 48), pme_order=8)
    n = len(s.charge)
    q = s.charge * math.sqrt(COULOMB)
    beta = systems.ewald_coefficient(s.params.cutoff, s.params.ewald_tol)
    r = orc.nb_fp64(s.pos, q, s.vdw_type, np.zeros((1, 1)), np.zeros((1, 1)), s.excl_ptr, s.excl_idx, s.box,
                    s.params.cutoff, s.params.switchdist, beta)
    cell = np.diag(s.box)
    atoms = np.zeros((n, 4), np.float32)
    atoms[:, :3] = s.pos / s.box
    atoms[:, 3] = q
    E_rec, _, _, _ = _oracle_pme(atoms, s.params.pme_grid, s.params.pme_order, cell, beta)
    E_self = -beta / math.sqrt(math.pi) * (q ** 2).sum()
    E = r["e_elec"] + r["e_slow"] + E_rec + E_self
    want = -(n / 2) * systems.MADELUNG_NACL * COULOMB / 4.0
    assert abs(E - want) / abs(want) < bound
