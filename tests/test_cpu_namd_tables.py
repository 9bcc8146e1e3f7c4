"""Pins of the NAMD CPU-path restatement (oracle_namd_tables / oracle_namd_cpu).

  * table size and r2 spacing for rc = 12 A: 769 points, r2_delta = 1/16
    (ComputeNonbondedUtil.C:607-626);
  * every bin's cubic reproduces the next node's value and gradient (the
    Hermite property the refinement steps enforce);
  * node values equal the closed-form energies (fast = 1/r - slow_energy,
    vdwA = S(r^2)/r^12, ...);
  * the table-based pair path agrees with the analytic FP64 oracle to its
    interpolation error, pair counts equal.
"""
import math

import numpy as np

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc


def test_table_shape_rc12():
    r2, t = orc.namd_tables(12.0, 10.0, 0.257952)
    assert r2.size == 769 and t.shape == (5, 769, 4)
    assert r2[0] == 1 / 16 and r2[64] == 2 / 16
    assert np.all(np.diff(r2[:-1]) > 0)


def test_cubics_interpolate_nodes():
    r2, t = orc.namd_tables(12.0, 10.0, 0.257952)
    n = r2.size
    for k in range(5):
        for i in range(1, n - 2):
            x = r2[i + 1] - r2[i]
            a, b, c, d = t[k, i]
            v = ((d * x + c) * x + b) * x + a
            g = (3 * d * x + 2 * c) * x + b
            scale_v = max(abs(t[k, i + 1, 0]), 1e-30)
            scale_g = max(abs(t[k, i + 1, 1]), 1e-30)
            assert abs(v - t[k, i + 1, 0]) <= 1e-9 * scale_v + 1e-14, (k, i)
            assert abs(g - t[k, i + 1, 1]) <= 1e-7 * scale_g + 1e-12, (k, i)


def test_node_values_are_closed_forms():
    rc, rs, beta = 12.0, 10.0, 0.257952
    r2t, t = orc.namd_tables(rc, rs, beta)
    c1 = 1.0 / (rc * rc - rs * rs) ** 3
    for i in (100, 300, 500, 700):
        r2 = r2t[i] - 1 / 16
        r = math.sqrt(r2)
        slow = 0.5 / rc * (3 - r2 / rc ** 2)
        assert abs(t[0, i, 0] - (1 / r - slow)) < 1e-12
        assert abs(t[1, i, 0] - (slow + (math.erfc(beta * r) - 1) / r)) < 1e-12
        S = 1.0 if r2 <= rs * rs else (rc * rc - r2) ** 2 * (rc * rc + 2 * r2 - 3 * rs * rs) * c1
        assert abs(t[3, i, 0] - S / r2 ** 6) < 1e-12 * max(1.0, 1 / r2 ** 6)
        assert abs(t[4, i, 0] - S / r2 ** 3) < 1e-12 * max(1.0, 1 / r2 ** 3)


def test_cpu_path_matches_analytic_to_table_error():
    s = systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7)
    p = s.params
    q = 2.0 ** -17
    pos = np.round(np.mod(s.pos, s.box) / q) * q
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    qs = s.charge * math.sqrt(332.0636)
    a = orc.nb_fp64(pos, qs, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, beta)
    b = orc.namd_cpu(pos, qs, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, beta)
    assert a["npairs"] == b["npairs"]

    def rel(x, y):
        return float(np.sqrt(((x - y) ** 2).sum() / (y ** 2).sum()))

    assert rel(b["f"], a["f"]) < 2e-5
    assert rel(b["fs"], a["fs"]) < 2e-6
    assert abs(b["e_vdw"] / a["e_vdw"] - 1) < 1e-6
    assert abs(b["e_elec"] + b["e_slow"] - a["e_elec"] - a["e_slow"]) < 1e-7 * abs(a["e_elec"] + a["e_slow"])
    # exact rule: pair counts within the cutoff equal the analytic oracle's
    e = orc.pairs_exact(pos, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.pairlistdist)
    assert e["n_cut"] == b["npairs"]
