"""Pins of the oracle's vdW force switching (NAMD's vdwForceSwitching,
src/ComputeNonbondedUtil.C:523-533, :900-919; Steinbach & Brooks, J. Comput.
Chem. 15, 667 (1994)) by what the construction fixes: the LJ force is the
plain one below the switching distance and falls smoothly to zero at the
cutoff, the energy is continuous with its derivative at the switching
distance and zero at the cutoff, below the switching distance the energy is
the plain LJ energy plus one constant per pair (so it is a shift, not a
change of the force), and the forces of a molecular box are minus the
gradient of its energy (central differences)."""
import math

import numpy as np
import pytest

from mi355nb import systems
from oracle import oracle as orc

RC, RS = 12.0, 10.0


def _pair(r, A=4.0e5, B=6.0e2, force_switch=True):
    """Energy and radial force (-dE/dr) of one LJ pair (no charge) at distance r."""
    pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
    o = orc.nb_fp64(pos, np.zeros(2), np.zeros(2, np.int32), np.array([[A]]), np.array([[B]]),
                    np.zeros(3, np.int64), np.zeros(0, np.int32), np.array([60.0, 60.0, 60.0]), RC, RS, 0.3,
                    vdw_force_switching=force_switch)
    return o["e_vdw"], o["f"][1, 0]


def test_plain_force_below_the_switching_distance():
    A, B = 4.0e5, 6.0e2
    for r in (3.0, 5.0, 8.0, 9.99):
        _, f = _pair(r)
        plain = 12 * A / r ** 13 - 6 * B / r ** 7
        assert abs(f - plain) <= 1e-12 * abs(plain), r


def test_energy_is_plain_plus_a_constant_below_switch():
    A, B = 4.0e5, 6.0e2
    shift = A * (-1.0 / (RS ** 6 * RC ** 6)) - B * (-1.0 / (RS ** 3 * RC ** 3))
    for r in (3.0, 6.0, 9.5):
        e, _ = _pair(r)
        assert abs(e - (A / r ** 12 - B / r ** 6 + shift)) <= 1e-12 * abs(A / r ** 12) + 1e-15, r


def test_continuous_at_switch_and_zero_at_cutoff():
    h = 1e-7
    e_in, f_in = _pair(RS - h)
    e_out, f_out = _pair(RS + h)
    assert abs(e_in - e_out) < 1e-9 and abs(f_in - f_out) < 1e-7
    e_c, f_c = _pair(RC - 1e-9)
    assert abs(e_c) < 1e-12 and abs(f_c) < 1e-10
    # beyond the cutoff nothing
    e_x, f_x = _pair(RC + 0.01)
    assert e_x == 0.0 and f_x == 0.0
    # force switching changes the force only between the switching distance and the cutoff
    _, f_es = _pair(11.0, force_switch=False)
    _, f_fs = _pair(11.0)
    assert f_es != f_fs


def test_force_is_minus_energy_gradient_in_a_box():
    s = systems.small_box(n_waters=120, n_chains=1, chain_len=20, box=20.0)
    p = s.params
    A, B = systems.lj_tables(s)
    q = s.charge * math.sqrt(332.0636)
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)

    def run(pos):
        return orc.nb_fp64(pos, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, beta,
                           vdw_force_switching=True)

    o = run(s.pos)
    rng = np.random.default_rng(2)
    for a in rng.choice(s.natoms, 6, replace=False):
        for d in range(3):
            h = 1e-5
            xp = s.pos.copy(); xp[a, d] += h
            xm = s.pos.copy(); xm[a, d] -= h
            ep, em = run(xp), run(xm)
            g = (ep["e_vdw"] + ep["e_elec"] + ep["e_slow"] - em["e_vdw"] - em["e_elec"] - em["e_slow"]) / (2 * h)
            f = o["f"][a, d] + o["fs"][a, d]
            assert abs(f + g) <= 2e-5 * max(1.0, abs(f)), (a, d, f, -g)
