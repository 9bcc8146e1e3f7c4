"""Pins of the FP64 force-matching gradient (oracle_nb_adjoint_fp64) and of
the mixing-rule chain rule in mi355nb.fitting, by central finite differences
of the FP64 loss L = sum |F - T|^2 (F = fast + slow force of oracle_nb_fp64)."""
import math

import numpy as np
import pytest

from mi355nb import fitting
from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


@pytest.fixture(scope="module")
def setup():
    s = systems.small_box(n_waters=150, n_chains=2, chain_len=20, box=20.0, n_bead_types=4, seed=31)
    beta = systems.ewald_coefficient(s.params.cutoff, s.params.ewald_tol)
    sig_t, eps_t, q_t = fitting.perturbed_parameters(s, seed=4)
    A_t, B_t = fitting.lj_tables(sig_t, eps_t)
    T = _forces(s, A_t, B_t, q_t * math.sqrt(COULOMB), beta)
    return s, beta, T


def _forces(s, A, B, qs, beta):
    p = s.params
    o = orc.nb_fp64(s.pos, qs, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, beta)
    return o["f"] + o["fs"]


def _loss(s, A, B, qs, beta, T):
    return float(((_forces(s, A, B, qs, beta) - T) ** 2).sum())


def test_lj_tables_match_library_mixing():
    sig = np.array([2.6, 3.1, 3.9, 3.15])
    eps = np.array([0.05, 0.12, 0.2, 0.1521])
    A, B = fitting.lj_tables(sig, eps)
    A2, B2 = L.lj_mix(sig, eps)
    assert np.allclose(A, A2, rtol=1e-14) and np.allclose(B, B2, rtol=1e-14)


def test_table_and_charge_gradient_matches_finite_differences(setup):
    s, beta, T = setup
    p = s.params
    nt = s.lj_sigma.size
    A, B = fitting.lj_tables(s.lj_sigma, s.lj_eps)
    qs = s.charge * math.sqrt(COULOMB)
    F = _forces(s, A, B, qs, beta)
    g = 2.0 * (F - T)
    gA, gB, gq = orc.nb_adjoint_fp64(s.pos, qs, s.vdw_type, nt, s.excl_ptr, s.excl_idx, s.box, p.cutoff,
                                     p.switchdist, beta, g)
    for (a, b) in [(0, 0), (1, 3), (0, nt - 2), (nt - 2, nt - 2), (nt - 2, nt - 1)]:
        for name, tab, grad in (("A", A, gA), ("B", B, gB)):
            h = 1e-4 * tab[a, b]
            tp = tab.copy()
            tm = tab.copy()
            tp[a, b] += h
            tm[a, b] -= h
            if a != b:
                tp[b, a] += h
                tm[b, a] -= h
            Lp = _loss(s, tp if name == "A" else A, tp if name == "B" else B, qs, beta, T)
            Lm = _loss(s, tm if name == "A" else A, tm if name == "B" else B, qs, beta, T)
            fd = (Lp - Lm) / (2 * h)
            an = grad[a, b] + (grad[b, a] if a != b else 0.0)
            assert abs(fd - an) <= 1e-6 * max(abs(fd), abs(an)) + 1e-9, (name, a, b, fd, an)
    for i in (0, 7, 41, 300, 400):
        h = 1e-4 * abs(qs[i])
        qp = qs.copy()
        qm = qs.copy()
        qp[i] += h
        qm[i] -= h
        fd = (_loss(s, A, B, qp, beta, T) - _loss(s, A, B, qm, beta, T)) / (2 * h)
        assert abs(fd - gq[i]) <= 1e-6 * max(abs(fd), abs(gq[i])) + 1e-9, (i, fd, gq[i])


def test_mixing_chain_rule(setup):
    """lj_param_grad against finite differences of a function that is linear
    in the tables, so only the chain rule is tested (diagonal included)."""
    rng = np.random.default_rng(0)
    nt = 5
    sig = rng.uniform(2.5, 4.0, nt)
    eps = rng.uniform(0.02, 0.2, nt)
    gA = rng.standard_normal((nt, nt))
    gB = rng.standard_normal((nt, nt))

    def f(sg, ep):
        A, B = fitting.lj_tables(sg, ep)
        return float((gA * A).sum() + (gB * B).sum())

    ds, de = fitting.lj_param_grad(sig, eps, gA, gB)
    for k in range(nt):
        h = 1e-6 * sig[k]
        sp, sm = sig.copy(), sig.copy()
        sp[k] += h
        sm[k] -= h
        fd = (f(sp, eps) - f(sm, eps)) / (2 * h)
        assert abs(fd - ds[k]) <= 1e-6 * abs(fd), (k, fd, ds[k])
        h = 1e-6 * eps[k]
        ep, em = eps.copy(), eps.copy()
        ep[k] += h
        em[k] -= h
        fd = (f(sig, ep) - f(sig, em)) / (2 * h)
        assert abs(fd - de[k]) <= 1e-6 * abs(fd), (k, fd, de[k])


def test_parameter_gradient_matches_finite_differences_h1e4(setup):
    """dL/d(sigma, eps, q) from the FP64 adjoint + chain rule against central
    differences with h = 1e-4 (relative) of the FP64 loss."""
    s, beta, T = setup
    p = s.params
    nt = s.lj_sigma.size
    sig, eps, q = s.lj_sigma.copy(), s.lj_eps.copy(), s.charge.copy()

    def loss(sg, ep, qq):
        A, B = fitting.lj_tables(sg, ep)
        return _loss(s, A, B, qq * math.sqrt(COULOMB), beta, T)

    A, B = fitting.lj_tables(sig, eps)
    qs = q * math.sqrt(COULOMB)
    g = 2.0 * (_forces(s, A, B, qs, beta) - T)
    gA, gB, gq = orc.nb_adjoint_fp64(s.pos, qs, s.vdw_type, nt, s.excl_ptr, s.excl_idx, s.box, p.cutoff,
                                     p.switchdist, beta, g)
    ds, de = fitting.lj_param_grad(sig, eps, gA, gB)
    dq = gq * math.sqrt(COULOMB)
    for k in range(nt):
        for arr, grad, name in ((sig, ds, "sigma"), (eps, de, "eps")):
            h = 1e-4 * arr[k]
            ap, am = arr.copy(), arr.copy()
            ap[k] += h
            am[k] -= h
            if name == "sigma":
                fd = (loss(ap, eps, q) - loss(am, eps, q)) / (2 * h)
            else:
                fd = (loss(sig, ap, q) - loss(sig, am, q)) / (2 * h)
            assert abs(fd - grad[k]) <= 1e-5 * max(abs(fd), abs(grad[k])) + 1e-6, (name, k, fd, grad[k])
    for i in (3, 50, 333):
        h = 1e-4 * abs(q[i])
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fd = (loss(sig, eps, qp) - loss(sig, eps, qm)) / (2 * h)
        assert abs(fd - dq[i]) <= 1e-5 * max(abs(fd), abs(dq[i])) + 1e-6, (i, fd, dq[i])
