"""Pins of the oracle's alchemical force-switched LJ (the vdwForceSwitching
branch of fep_vdw_forceandenergies and ti_vdw_force_energy_dUdl,
src/ComputeNonbondedFEP.C:136-159, src/ComputeNonbondedTI.C:113-128, with
vdw_fswitch_forceandenergy / vdw_fswitch_shift of
src/ComputeNonbondedAlch.h:56-104).  In that form the soft-core shift sh
moves both the pair's r^2 and the switching distance (sd = switchdist^2 +
sh) but not the cutoff, and which branch a pair takes is decided on the
unshifted r^2.  Checks:
  * all factors 1, no shift, no cross-partition pairs -> the plain
    force-switched oracle;
  * forces are minus the finite-difference gradient of the energy at an
    intermediate lambda with a shift;
  * the lambda2 energies are the energies of a run at lambda2;
  * TI dU/dlambda is the lambda derivative of the energy with the shift
    coefficient zero (finite differences over lambda), and with a shift
    coefficient NAMD's formula U + lambda coeff F/2 (+ the constant below
    switchdist) is the derivative at a fixed shifted switching distance,
    checked on one pair by moving switchdist to hold sd fixed.
"""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


def _factors(up=1.0, down=1.0, up2=None, down2=None, shup=0.0, shdown=0.0, coeff=0.0, decouple=False):
    up2 = up if up2 is None else up2
    down2 = down if down2 is None else down2
    return dict(elecLambdaUp=up, elecLambdaDown=down, vdwLambdaUp=up, vdwLambdaDown=down, elecLambda2Up=up2,
                elecLambda2Down=down2, vdwLambda2Up=up2, vdwLambda2Down=down2, vdwShiftUp=shup,
                vdwShiftDown=shdown, vdwShift2Up=shup, vdwShift2Down=shdown, alchDecouple=int(decouple),
                alchVdwShiftCoeff=coeff)


@pytest.fixture(scope="module")
def box():
    s = systems.small_box(n_waters=180, n_chains=2, chain_len=20, box=20.0, seed=21)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    beta = systems.ewald_coefficient(s.params.cutoff, s.params.ewald_tol)
    return s, A, B, beta


def _alch(s, part, al, A, B, beta, pos=None, switchdist=None):
    p = s.params
    return orc.nb_fp64_alch(s.pos if pos is None else pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B,
                            s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist if switchdist is None else switchdist,
                            beta, part, al, vdw_force_switching=True)


def _part(s, with_down=True):
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    if with_down:
        part[20:40] = 2
    return part


def test_unit_factors_recover_plain_force_switch(box):
    s, A, B, beta = box
    r = _alch(s, _part(s, with_down=False), _factors(1.0, 1.0), A, B, beta)
    p = s.params
    o = orc.nb_fp64(s.pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box, p.cutoff,
                    p.switchdist, beta, vdw_force_switching=True)
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k] - o[k]) <= 1e-11 * max(1.0, abs(o[k])), k
    assert np.abs(r["f"] - o["f"]).max() <= 1e-10 * np.abs(o["f"]).max()
    # and differs from the energy-switched alchemical oracle
    p = s.params
    e = orc.nb_fp64_alch(s.pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr, s.excl_idx, s.box,
                         p.cutoff, p.switchdist, beta, _part(s, with_down=False), _factors(1.0, 1.0))
    assert abs(e["e_vdw"] - r["e_vdw"]) > 1e-6 * abs(r["e_vdw"])


def test_forces_are_energy_gradient(box):
    s, A, B, beta = box
    part = _part(s)
    al = _factors(0.4, 0.7, shup=5.0 * 0.6, shdown=5.0 * 0.3)
    r = _alch(s, part, al, A, B, beta)
    rng = np.random.default_rng(3)
    for a in list(rng.choice(40, 4, replace=False)) + list(rng.choice(np.arange(40, s.natoms), 3, replace=False)):
        for d in range(3):
            h = 1e-5
            xp = s.pos.copy(); xp[a, d] += h
            xm = s.pos.copy(); xm[a, d] -= h
            ep, em = _alch(s, part, al, A, B, beta, pos=xp), _alch(s, part, al, A, B, beta, pos=xm)
            g = sum(ep[k] - em[k] for k in ("e_vdw", "e_elec", "e_slow")) / (2 * h)
            f = r["f"][a, d] + r["fs"][a, d]
            assert abs(f + g) <= 2e-5 * max(1.0, abs(f)), (a, d, f, -g)


def test_lambda2_energies_are_a_run_at_lambda2(box):
    s, A, B, beta = box
    part = _part(s)
    r = _alch(s, part, dict(_factors(0.4, 0.7, up2=0.5, down2=0.6, shup=3.0, shdown=1.5),
                            vdwShift2Up=2.5, vdwShift2Down=2.0), A, B, beta)
    r2 = _alch(s, part, _factors(0.5, 0.6, shup=2.5, shdown=2.0), A, B, beta)
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k + "_s"] - r2[k]) <= 1e-11 * max(1.0, abs(r2[k])), k


def test_ti_is_lambda_derivative_without_shift(box):
    s, A, B, beta = box
    part = _part(s)
    lam, h = 0.45, 1e-4
    r = _alch(s, part, _factors(lam, lam), A, B, beta)
    for side, key in ((1, "Up"), (2, "Down")):
        ep = _alch(s, part, dict(_factors(lam, lam), **{"vdwLambda" + key: lam + h}), A, B, beta)
        em = _alch(s, part, dict(_factors(lam, lam), **{"vdwLambda" + key: lam - h}), A, B, beta)
        fd = (ep["e_vdw"] - em["e_vdw"]) / (2 * h)
        got = r[f"e_vdw_ti_{side}"]
        assert abs(got - fd) <= 1e-7 * max(1.0, abs(fd)), (side, got, fd)


@pytest.mark.parametrize("dist", [3.6, 11.0])
def test_ti_softcore_term_at_fixed_shifted_switch(dist):
    """One LJ pair (atom 0 appearing, atom 1 environment) at a distance
    below (3.6 A) or above (11 A) switchdist 10 A: NAMD's dU/dlambda equals
    the finite difference of the energy over lambda when switchdist is moved
    with lambda so that switchdist^2 + shift(lambda) stays fixed."""
    cutoff, sw0, coeff, lam, h = 12.0, 10.0, 5.0, 0.35, 1e-5
    A = np.array([[6.0e5]]); B = np.array([[6.0e2]])
    pos = np.array([[20.0, 20.0, 20.0], [20.0 + dist, 20.0, 20.0]])
    part = np.array([1, 0], np.int8)
    box = np.array([60.0, 60.0, 60.0])
    sd = sw0 * sw0 + coeff * (1 - lam)

    def run(l, sw):
        al = _factors(l, 1.0, shup=coeff * (1 - l), coeff=coeff)
        return orc.nb_fp64_alch(pos, np.zeros(2), np.zeros(2, np.int32), A, B, np.zeros(3, np.int64),
                                np.zeros(0, np.int32), box, cutoff, sw, 0.3, part, al, vdw_force_switching=True)

    r = run(lam, sw0)
    sw = lambda l: math.sqrt(sd - coeff * (1 - l))
    fd = (run(lam + h, sw(lam + h))["e_vdw"] - run(lam - h, sw(lam - h))["e_vdw"]) / (2 * h)
    assert abs(r["e_vdw_ti_1"] - fd) <= 1e-6 * max(1.0, abs(fd)), (r["e_vdw_ti_1"], fd)
    assert r["e_vdw_ti_2"] == 0.0
