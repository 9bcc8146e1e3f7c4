"""Pins of the oracle's modified (scaled 1-4) exclusion correction
(modifiedExclusionForce, src/ComputeBondedCUDAKernel.cu:316-632, C1 forms of
src/CudaComputeNonbondedInteractions.h:170-254), by what NAMD's scaled 1-4
rule fixes: a listed pair keeps its ordinary tile-kernel interaction, and the
correction turns it into (a) LJ with the 1-4 coefficients and (b) Coulomb
scaled by scale14 minus the Ewald smooth part:
  * scale14 = 1 with A14 = A, B14 = B: the correction is zero;
  * one pair: ordinary pair (plain oracle) + correction has total
    electrostatics s14 qq/r - qq erf(beta r)/r (closed form) and the LJ of the
    plain oracle run with the 1-4 tables;
  * its fast part is -(1 - s14) times the ordinary pair's fast part;
  * forces are minus the gradient of the correction's energy (central
    differences) on a box with its chains' 1-4 pairs listed;
  * a pair listed from both ends counts once;
  * with alchemical partitions: unit lambdas give the plain correction; on one
    soft-core pair, ordinary alchemical pair + correction equals the
    alchemical pair with the 1-4 tables (vdW energies at lambda and lambda2,
    the TI sum, forces) and lambda_elec (s14 qq/r - qq erf(beta r)/r) for the
    electrostatics; cross-partition pairs get nothing, decoupled ones the
    plain correction; forces are minus the energy gradient.
"""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


def _pair_system(r, qi=0.6, qj=-0.45):
    pos = np.array([[20.0, 20.0, 20.0], [20.0 + r * 0.6, 20.0 + r * 0.8, 20.0]])
    q = np.array([qi, qj]) * math.sqrt(COULOMB)
    t = np.array([0, 1], np.int32)
    A = np.array([[4.0e5, 6.0e5], [6.0e5, 9.0e5]])
    B = np.array([[5.0e2, 7.0e2], [7.0e2, 9.0e2]])
    return pos, q, t, A, B


def _mod(pos, q, t, A, B, A14, B14, s14, fs=False, both_ends=False):
    ptr = np.array([0, 1, 2 if both_ends else 1], np.int64)
    idx = np.array([1, 0] if both_ends else [1], np.int32)
    return orc.mod_excl_fp64(pos, q, t, A, B, A14, B14, ptr, idx, np.array([60.0, 60.0, 60.0]), 12.0, 10.0, s14,
                             vdw_force_switching=fs)


@pytest.mark.parametrize("r", [1.6, 3.9, 10.7])
def test_unit_scale_and_same_tables_give_zero(r):
    pos, q, t, A, B = _pair_system(r)
    o = _mod(pos, q, t, A, B, A, B, 1.0)
    assert o["npairs"] == 1
    assert o["e_vdw"] == 0.0 and o["e_elec"] == 0.0 and o["e_slow"] == 0.0
    assert not o["f"].any() and not o["fs"].any()


@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("r", [1.6, 3.9, 10.7])
def test_one_pair_closed_forms(r, fs):
    pos, q, t, A, B = _pair_system(r)
    A14, B14 = 0.5 * A, 0.7 * B
    s14, beta = 0.8333, 0.32
    box = np.array([60.0, 60.0, 60.0])
    none_p, none_i = np.zeros(3, np.int64), np.zeros(0, np.int32)
    plain = orc.nb_fp64(pos, q, t, A, B, none_p, none_i, box, 12.0, 10.0, beta, vdw_force_switching=fs)
    plain14 = orc.nb_fp64(pos, q * 0, t, A14, B14, none_p, none_i, box, 12.0, 10.0, beta, vdw_force_switching=fs)
    corr = _mod(pos, q, t, A, B, A14, B14, s14, fs=fs)
    qq = q[0] * q[1]
    total_elec = plain["e_elec"] + plain["e_slow"] + corr["e_elec"] + corr["e_slow"]
    want = s14 * qq / r - qq * math.erf(beta * r) / r
    assert abs(total_elec - want) <= 1e-12 * abs(qq / r)
    assert abs(plain["e_vdw"] + corr["e_vdw"] - plain14["e_vdw"]) <= 1e-12 * max(1.0, abs(plain14["e_vdw"]))
    assert abs(corr["e_elec"] + (1 - s14) * plain["e_elec"]) <= 1e-12 * abs(qq / r)
    # forces: the LJ part turns the plain LJ force into the 1-4 one
    plain_lj = orc.nb_fp64(pos, q * 0, t, A, B, none_p, none_i, box, 12.0, 10.0, beta, vdw_force_switching=fs)
    elec_fast = plain["f"] - plain_lj["f"]
    want_f = plain14["f"] - plain_lj["f"] - (1 - s14) * elec_fast
    assert np.abs(corr["f"] - want_f).max() <= 1e-10 * np.abs(want_f).max()


def test_pair_listed_from_both_ends_counts_once():
    pos, q, t, A, B = _pair_system(3.0)
    one = _mod(pos, q, t, A, B, 0.5 * A, 0.7 * B, 0.5)
    two = _mod(pos, q, t, A, B, 0.5 * A, 0.7 * B, 0.5, both_ends=True)
    assert two["npairs"] == 1
    assert two["e_vdw"] == one["e_vdw"] and two["e_elec"] == one["e_elec"]


@pytest.mark.parametrize("fs", [False, True])
def test_forces_are_energy_gradient(fs):
    s = systems.small_box(n_waters=60, n_chains=3, chain_len=20, box=20.0, seed=5)
    p = s.params
    A, B = systems.lj_tables(s)
    A14, B14 = systems.lj14_tables(s)
    mptr, midx = systems.chain_pairs_14(s)
    q = s.charge * math.sqrt(COULOMB)

    def run(pos):
        return orc.mod_excl_fp64(pos, q, s.vdw_type, A, B, A14, B14, mptr, midx, s.box, p.cutoff, p.switchdist,
                                 0.8333, vdw_force_switching=fs)

    o = run(s.pos)
    assert o["npairs"] == 3 * 17
    for a in (0, 5, 21, 47):
        for d in range(3):
            h = 1e-5
            xp = s.pos.copy(); xp[a, d] += h
            xm = s.pos.copy(); xm[a, d] -= h
            ep, em = run(xp), run(xm)
            g_fast = (ep["e_vdw"] + ep["e_elec"] - em["e_vdw"] - em["e_elec"]) / (2 * h)
            g_slow = (ep["e_slow"] - em["e_slow"]) / (2 * h)
            assert abs(o["f"][a, d] + g_fast) <= 1e-5 * max(1.0, abs(g_fast)), (a, d)
            assert abs(o["fs"][a, d] + g_slow) <= 1e-5 * max(1.0, abs(g_slow)), (a, d)
    # pair forces are central, so the virial is symmetric
    assert np.allclose(o["virial"], o["virial"].T, rtol=0, atol=1e-9 * np.abs(o["virial"]).max())


# ---- alchemical correction (modifiedExclusionForce's pswitch route,
# src/ComputeBondedCUDAKernel.cu:379-475, :575-586; pair classes
# src/ComputeBondedCUDA.C:97-99, :3126-3129)

def _alch(up=1.0, down=1.0, eup=None, edown=None, up2=None, down2=None, shup=0.0, shdown=0.0, coeff=0.0,
          decouple=False):
    eup = up if eup is None else eup
    edown = down if edown is None else edown
    up2 = up if up2 is None else up2
    down2 = down if down2 is None else down2
    return dict(elecLambdaUp=eup, elecLambdaDown=edown, vdwLambdaUp=up, vdwLambdaDown=down,
                elecLambda2Up=up2, elecLambda2Down=down2, vdwLambda2Up=up2, vdwLambda2Down=down2,
                vdwShiftUp=shup, vdwShiftDown=shdown, vdwShift2Up=shup, vdwShift2Down=shdown,
                alchDecouple=int(decouple), alchVdwShiftCoeff=coeff)


def _mod_alch(pos, q, t, A, B, A14, B14, s14, part, al):
    ptr = np.array([0, 1, 1], np.int64)
    idx = np.array([1], np.int32)
    return orc.mod_excl_fp64(pos, q, t, A, B, A14, B14, ptr, idx, np.array([60.0, 60.0, 60.0]), 12.0, 10.0, s14,
                             partition=np.asarray(part, np.int8), alch=al)


def test_alch_unit_lambdas_give_the_plain_correction():
    s = systems.small_box(n_waters=60, n_chains=3, chain_len=20, box=20.0, seed=5)
    p = s.params
    A, B = systems.lj_tables(s)
    A14, B14 = systems.lj14_tables(s)
    mptr, midx = systems.chain_pairs_14(s)
    q = s.charge * math.sqrt(COULOMB)
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    plain = orc.mod_excl_fp64(s.pos, q, s.vdw_type, A, B, A14, B14, mptr, midx, s.box, p.cutoff, p.switchdist,
                              0.8333)
    al = orc.mod_excl_fp64(s.pos, q, s.vdw_type, A, B, A14, B14, mptr, midx, s.box, p.cutoff, p.switchdist, 0.8333,
                           partition=part, alch=_alch())
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(al[k] - plain[k]) <= 1e-12 * abs(plain[k]), k
        assert abs(al[k + "_s"] - plain[k]) <= 1e-12 * abs(plain[k]), k
    assert np.abs(al["f"] - plain["f"]).max() <= 1e-10 * np.abs(plain["f"]).max()


@pytest.mark.parametrize("r", [2.2, 4.5, 10.9])
def test_alch_one_pair_with_softcore(r):
    """Ordinary alchemical pair (A, B) + correction = the alchemical pair with
    the 1-4 tables for vdW (energies at lambda and lambda2, TI sum, forces)
    and lambda_elec (s14 qq/r - qq erf(beta r)/r) for the electrostatics."""
    pos, q, t, A, B = _pair_system(r)
    A14, B14 = 0.5 * A, 0.7 * B
    s14, beta = 0.8333, 0.32
    part = [1, 0]
    al = dict(_alch(up=0.4, eup=0.7, up2=0.5, shup=5.0 * 0.6, coeff=5.0), vdwShift2Up=5.0 * 0.5)
    box = np.array([60.0, 60.0, 60.0])
    none_p, none_i = np.zeros(3, np.int64), np.zeros(0, np.int32)
    nb = orc.nb_fp64_alch(pos, q, t, A, B, none_p, none_i, box, 12.0, 10.0, beta, np.array(part, np.int8), al)
    nb14 = orc.nb_fp64_alch(pos, q * 0, t, A14, B14, none_p, none_i, box, 12.0, 10.0, beta, np.array(part, np.int8),
                            al)
    nb_lj = orc.nb_fp64_alch(pos, q * 0, t, A, B, none_p, none_i, box, 12.0, 10.0, beta, np.array(part, np.int8), al)
    m = _mod_alch(pos, q, t, A, B, A14, B14, s14, part, al)
    for k in ("e_vdw", "e_vdw_s", "e_vdw_ti_1"):
        assert abs(nb[k] + m[k] - nb14[k]) <= 1e-11 * max(1.0, abs(nb14[k])), k
    qq = q[0] * q[1]
    closed = s14 * qq / r - qq * math.erf(beta * r) / r
    assert abs(nb["e_elec"] + nb["e_slow"] + m["e_elec"] + m["e_slow"] - 0.7 * closed) <= 1e-11 * abs(qq / r)
    assert abs(nb["e_elec_s"] + nb["e_slow_s"] + m["e_elec_s"] + m["e_slow_s"] - 0.5 * closed) <= 1e-11 * abs(qq / r)
    assert abs(nb["e_elec_ti_1"] + nb["e_slow_ti_1"] + m["e_elec_ti_1"] + m["e_slow_ti_1"] - closed) <= \
        1e-11 * abs(qq / r)
    # forces: vdW part as the 1-4 alchemical pair, elec part lambda_elec-scaled
    elec_nb = nb["f"] - nb_lj["f"]
    want = nb14["f"] - nb_lj["f"] - (1 - s14) * elec_nb
    assert np.abs(m["f"] - want).max() <= 1e-10 * np.abs(want).max()
    assert m["e_vdw_ti_2"] == 0.0


def test_alch_cross_pairs_vanish_and_decoupled_pairs_are_plain():
    pos, q, t, A, B = _pair_system(3.5)
    A14, B14 = 0.5 * A, 0.7 * B
    cross = _mod_alch(pos, q, t, A, B, A14, B14, 0.5, [1, 2], _alch(0.3, 0.6))
    assert cross["npairs"] == 0 and cross["e_vdw"] == 0.0 and not cross["f"].any()
    dec = _mod_alch(pos, q, t, A, B, A14, B14, 0.5, [2, 2], _alch(0.3, 0.6, decouple=True))
    plain = _mod(pos, q, t, A, B, A14, B14, 0.5)
    assert dec["e_vdw"] == plain["e_vdw"] and dec["e_elec"] == plain["e_elec"] and dec["e_vdw_ti_2"] == 0.0


def test_alch_forces_are_energy_gradient():
    s = systems.small_box(n_waters=60, n_chains=3, chain_len=20, box=20.0, seed=5)
    p = s.params
    A, B = systems.lj_tables(s)
    A14, B14 = systems.lj14_tables(s)
    mptr, midx = systems.chain_pairs_14(s)
    q = s.charge * math.sqrt(COULOMB)
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    part[20:40] = 2
    al = _alch(0.35, 0.8, eup=0.6, edown=0.9, shup=5.0 * 0.65, shdown=5.0 * 0.2, coeff=5.0)

    def run(pos):
        return orc.mod_excl_fp64(pos, q, s.vdw_type, A, B, A14, B14, mptr, midx, s.box, p.cutoff, p.switchdist,
                                 0.8333, partition=part, alch=al)

    o = run(s.pos)
    for a in (0, 5, 21, 47):
        for d in range(3):
            h = 1e-5
            xp = s.pos.copy(); xp[a, d] += h
            xm = s.pos.copy(); xm[a, d] -= h
            ep, em = run(xp), run(xm)
            g_fast = (ep["e_vdw"] + ep["e_elec"] - em["e_vdw"] - em["e_elec"]) / (2 * h)
            g_slow = (ep["e_slow"] - em["e_slow"]) / (2 * h)
            assert abs(o["f"][a, d] + g_fast) <= 1e-5 * max(1.0, abs(g_fast)), (a, d)
            assert abs(o["fs"][a, d] + g_slow) <= 1e-5 * max(1.0, abs(g_slow)), (a, d)


def test_general_cell_variant_matches_the_box_in_an_orthorhombic_cell():
    s = systems.small_box(n_waters=60, n_chains=3, chain_len=20, box=20.0, seed=5)
    p = s.params
    A, B = systems.lj_tables(s)
    A14, B14 = systems.lj14_tables(s)
    mptr, midx = systems.chain_pairs_14(s)
    q = s.charge * math.sqrt(COULOMB)
    o = orc.mod_excl_fp64(s.pos, q, s.vdw_type, A, B, A14, B14, mptr, midx, s.box, p.cutoff, p.switchdist, 0.8333)
    c = orc.mod_excl_fp64(s.pos, q, s.vdw_type, A, B, A14, B14, mptr, midx, s.box, p.cutoff, p.switchdist, 0.8333,
                          cell=np.diag(s.box))
    assert c["npairs"] == o["npairs"]
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(c[k] - o[k]) <= 1e-12 * abs(o[k]), k
    assert np.abs(c["f"] - o["f"]).max() <= 1e-10 * np.abs(o["f"]).max()
