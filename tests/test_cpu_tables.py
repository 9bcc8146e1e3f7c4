"""NAMD's 1/r force/energy tables (table mode): the library's tables
(mi355nb_build_force_energy_tables) against NAMD's spline-sampled tables
restated in the oracle, closed forms at the sample points, and the table-mode
pair evaluation of the oracle converging to the analytic oracle at second
order in the table spacing (which pins the index/interpolation convention)."""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

RC, RS = 12.0, 10.0
N = 4096


@pytest.fixture(scope="module")
def tables():
    beta = systems.ewald_coefficient(RC, 1e-6)
    f, e = L.force_energy_tables(N, RC, RS, beta)
    return beta, f, e


def test_get_do_table_follows_reference_heuristic():
    lib = L.lib()
    assert lib.mi355nb_get_do_table(0, 0, 0) == 0
    assert lib.mi355nb_get_do_table(1, 0, 0) == 1
    assert lib.mi355nb_get_do_table(0, 1, 0) == 1
    assert lib.mi355nb_get_do_table(0, 1, 1) == 1


def test_tables_match_namd_spline_samples(tables):
    beta, f, e = tables
    fo, eo = orc.namd_gpu_tables(N, RC, RS, beta)
    i = np.arange(N)
    m = (N / (i + 0.5) <= RC) & (i > 0)
    for col in range(4):
        for a, b in ((f, fo), (e, eo)):
            scale = np.abs(b[m, col]).max()
            assert np.abs(a[m, col] - b[m, col]).max() <= 2e-5 * scale, col
    assert np.all(f[~m] == 0.0) and np.all(fo[~m] == 0.0)
    assert np.array_equal(e[0], e[1])


def test_closed_forms_at_sample_points(tables):
    beta, f, e = tables
    for i in (400, 1000, 2000, 4000):
        r = N / (i + 0.5)
        # fast + slow correction = erfc(beta r)/r
        assert abs((e[i, 0] + e[i, 3]) - math.erfc(beta * r) / r) <= 1e-6 / r
        g = -(2 * beta / math.sqrt(math.pi)) * math.exp(-(beta * r) ** 2) / r - math.erfc(beta * r) / r ** 2
        # (float32 columns: tolerance on the size of the two terms, which nearly cancel at long range)
        assert abs((f[i, 0] + f[i, 3]) - g / r) <= 1e-6 * (abs(f[i, 0]) + abs(f[i, 3]))
        S = 1.0
        if r > RS:
            S = (RC ** 2 - r * r) ** 2 * (RC ** 2 + 2 * r * r - 3 * RS ** 2) / (RC ** 2 - RS ** 2) ** 3
        assert abs(e[i, 2] - S * r ** -12) <= 1e-6 * S * r ** -12
        assert abs(e[i, 1] + S * r ** -6) <= 1e-6 * S * r ** -6
        if r < RS:
            assert abs(f[i, 2] - (-12.0 * r ** -14)) <= 1e-6 * 12.0 * r ** -14
            assert abs(f[i, 1] - 6.0 * r ** -8) <= 1e-6 * 6.0 * r ** -8


def test_force_columns_are_twice_the_r2_derivative(tables):
    beta, f, e = tables
    for i in (500, 1500, 3000):
        r0, r1 = N / (i + 0.5), N / (i + 1.5)
        for col in range(4):
            slope = (e[i + 1, col] - e[i, col]) / (r1 * r1 - r0 * r0)
            mid = 0.25 * (f[i, col] + f[i + 1, col])
            assert abs(slope - mid) <= 2e-3 * abs(mid) + 1e-12, (i, col, slope, mid)


def test_table_mode_converges_at_second_order():
    s = systems.small_box(n_waters=120, n_chains=2, chain_len=10, box=20.0, seed=41)
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    q = s.charge * math.sqrt(332.0636)
    o = orc.nb_fp64(s.pos, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, beta)
    errs = []
    for n in (2048, 8192):
        ft, et = L.force_energy_tables(n, p.cutoff, p.switchdist, beta)
        t = orc.nb_table(s.pos, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, ft, et)
        assert t["npairs"] == o["npairs"]
        errs.append(np.sqrt(((t["f"] - o["f"]) ** 2).sum() / (o["f"] ** 2).sum()))
    # linear interpolation in 1/r: error ~ spacing^2, so 4x the entries -> ~16x smaller
    assert errs[0] / errs[1] > 8.0, errs
