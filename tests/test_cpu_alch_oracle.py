"""Pins of the FP64 alchemical (FEP) oracle, oracle_nb_fp64_alch.

Each test checks a property that NAMD's alchemical rules fix independently
of the soft-core formula itself:
  * all factors 1 and no shift -> the plain nonbonded oracle;
  * lambda = 0 without decoupling -> the solute is invisible: the
    environment alone (solute atoms removed) gives the same energies and
    forces, and the solute feels nothing;
  * lambda = 0 with decoupling -> environment + isolated solute;
  * (1, 2) pairs vanish: E = E_full - E_cross(P, Q), where the cross term is
    built from plain-oracle runs on sub-systems;
  * forces are minus the finite-difference gradient of the energy at an
    intermediate lambda with a soft-core shift;
  * the soft-core energy stays finite at contact.
"""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


def _subsystem(s, keep):
    keep = np.asarray(sorted(keep))
    new = -np.ones(s.natoms, np.int64)
    new[keep] = np.arange(keep.size)
    ptr = [0]
    idx = []
    for a in keep:
        js = [new[j] for j in s.excl_idx[s.excl_ptr[a]:s.excl_ptr[a + 1]] if new[j] >= 0]
        idx.extend(sorted(js))
        ptr.append(len(idx))
    return dict(pos=s.pos[keep], q=s.charge[keep] * math.sqrt(COULOMB), t=s.vdw_type[keep],
                ptr=np.array(ptr, np.int64), idx=np.array(idx, np.int32))


def _plain(s, sub, A, B, beta):
    p = s.params
    return orc.nb_fp64(sub["pos"], sub["q"], sub["t"], A, B, sub["ptr"], sub["idx"], s.box, p.cutoff, p.switchdist,
                       beta)


def _alch(s, part, al, A, B, beta, pos=None):
    p = s.params
    return orc.nb_fp64_alch(s.pos if pos is None else pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B,
                            s.excl_ptr, s.excl_idx, s.box, p.cutoff, p.switchdist, beta, part, al)


def _factors(up=1.0, down=1.0, up2=None, down2=None, shup=0.0, shdown=0.0, decouple=False):
    up2 = up if up2 is None else up2
    down2 = down if down2 is None else down2
    return dict(elecLambdaUp=up, elecLambdaDown=down, vdwLambdaUp=up, vdwLambdaDown=down, elecLambda2Up=up2,
                elecLambda2Down=down2, vdwLambda2Up=up2, vdwLambda2Down=down2, vdwShiftUp=shup,
                vdwShiftDown=shdown, vdwShift2Up=shup, vdwShift2Down=shdown, alchDecouple=int(decouple))


@pytest.fixture(scope="module")
def box():
    s = systems.small_box(n_waters=180, n_chains=2, chain_len=20, box=20.0, seed=21)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    beta = systems.ewald_coefficient(s.params.cutoff, s.params.ewald_tol)
    return s, A, B, beta


def _etot(r, lam2=False):
    k = "_s" if lam2 else ""
    return r["e_vdw" + k] + r["e_elec" + k] + r["e_slow" + k]


def test_unit_factors_recover_plain_oracle(box):
    s, A, B, beta = box
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    r = _alch(s, part, _factors(1.0, 1.0), A, B, beta)
    o = _plain(s, _subsystem(s, range(s.natoms)), A, B, beta)
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k] - o[k]) <= 1e-12 * abs(o[k]), k
        assert abs(r[k + "_s"] - o[k]) <= 1e-12 * abs(o[k]), k
    assert np.abs(r["f"] - o["f"]).max() <= 1e-10 * np.abs(o["f"]).max()
    assert np.abs(r["fs"] - o["fs"]).max() <= 1e-10 * np.abs(o["fs"]).max()
    assert np.abs(r["virial"] - o["virial"]).max() <= 1e-10 * np.abs(o["virial"]).max()


def test_lambda_zero_hides_the_solute(box):
    s, A, B, beta = box
    part = np.zeros(s.natoms, np.int8)
    sol = np.arange(20)
    part[sol] = 1
    r = _alch(s, part, _factors(0.0, 1.0, shup=5.0), A, B, beta)
    env = np.arange(20, s.natoms)
    o = _plain(s, _subsystem(s, env), A, B, beta)
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k] - o[k]) <= 1e-11 * abs(o[k]), k
    assert np.abs(r["f"][env] - o["f"]).max() <= 1e-10 * np.abs(o["f"]).max()
    assert np.abs(r["f"][sol]).max() == 0.0 and np.abs(r["fs"][sol]).max() == 0.0


def test_lambda_zero_decoupled_keeps_solute_internal(box):
    s, A, B, beta = box
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    r = _alch(s, part, _factors(0.0, 1.0, shup=5.0, decouple=True), A, B, beta)
    e = _plain(s, _subsystem(s, range(20, s.natoms)), A, B, beta)
    q = _plain(s, _subsystem(s, range(20)), A, B, beta)
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k] - (e[k] + q[k])) <= 1e-11 * (abs(e[k]) + abs(q[k])), k


def test_cross_partition_pairs_vanish(box):
    s, A, B, beta = box
    P = np.arange(0, 20)
    Q = np.arange(20, 40)
    part = np.zeros(s.natoms, np.int8)
    part[P] = 1
    part[Q] = 2
    r = _alch(s, part, _factors(1.0, 1.0), A, B, beta)
    full = _plain(s, _subsystem(s, range(s.natoms)), A, B, beta)
    pq = _plain(s, _subsystem(s, np.concatenate([P, Q])), A, B, beta)
    p = _plain(s, _subsystem(s, P), A, B, beta)
    q = _plain(s, _subsystem(s, Q), A, B, beta)
    for k in ("e_vdw", "e_elec", "e_slow"):
        cross = pq[k] - p[k] - q[k]
        assert abs(cross) > 1e-6 * abs(full[k]) or k == "e_vdw"
        assert abs(r[k] - (full[k] - cross)) <= 1e-10 * abs(full[k]), k


def test_forces_are_energy_gradient_with_softcore(box):
    s, A, B, beta = box
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    part[20:30] = 2
    al = _factors(0.35, 0.8, 0.35, 0.8, shup=5.0 * 0.65, shdown=5.0 * 0.2)
    r = _alch(s, part, al, A, B, beta)
    # lambda2 == lambda -> the "_s" sums equal the plain sums
    for k in ("e_vdw", "e_elec", "e_slow"):
        assert abs(r[k + "_s"] - r[k]) <= 1e-12 * abs(r[k])
    ftot = r["f"] + r["fs"]
    h = 1e-5
    for a in (0, 5, 19, 22, 31, 60, 200):
        for d in range(3):
            pp = s.pos.copy()
            pm = s.pos.copy()
            pp[a, d] += h
            pm[a, d] -= h
            g = (_etot(_alch(s, part, al, A, B, beta, pp)) - _etot(_alch(s, part, al, A, B, beta, pm))) / (2 * h)
            assert abs(-g - ftot[a, d]) <= 2e-5 * max(1.0, abs(ftot[a, d])), (a, d, -g, ftot[a, d])


def test_softcore_finite_at_contact():
    """Two atoms 0.05 A apart, solute at vdw lambda 0.5 with shift 2.5:
    E = lambda S U(r^2 + shift) + scaled electrostatics, all finite."""
    sig = np.array([3.0])
    eps = np.array([0.15])
    A, B = L.lj_mix(sig, eps)
    pos = np.array([[5.0, 5.0, 5.0], [5.05, 5.0, 5.0]])
    part = np.array([1, 0], np.int8)
    al = _factors(0.5, 1.0, shup=2.5)
    al["elecLambdaUp"] = al["elecLambda2Up"] = 0.0
    r = orc.nb_fp64_alch(pos, np.zeros(2), np.zeros(2, np.int32), A, B, np.zeros(3, np.int64), np.zeros(1, np.int32),
                         np.array([30.0, 30.0, 30.0]), 12.0, 10.0, 0.26, part, al)
    x = 0.05 ** 2 + 2.5
    U = A[0, 0] / x ** 6 - B[0, 0] / x ** 3
    assert math.isfinite(r["e_vdw"]) and abs(r["e_vdw"] - 0.5 * U) <= 1e-12 * abs(U)
    assert np.all(np.isfinite(r["f"]))


def test_ti_sums_are_lambda_derivatives(box):
    """The TI sums are the derivatives of the FEP energies with respect to
    each partition's lambda, the soft-core shift moving with it
    (shift = coeff (1 - lambda)): central differences of the oracle's own FEP
    energies, with the vdw and elec lambdas varied separately."""
    s, A, B, beta = box
    part = np.zeros(s.natoms, np.int8)
    part[:20] = 1
    part[20:30] = 2
    coeff = 5.0

    def factors(vu, vd, eu, ed):
        f = dict(elecLambdaUp=eu, elecLambdaDown=ed, vdwLambdaUp=vu, vdwLambdaDown=vd, elecLambda2Up=eu,
                 elecLambda2Down=ed, vdwLambda2Up=vu, vdwLambda2Down=vd, vdwShiftUp=coeff * (1 - vu),
                 vdwShiftDown=coeff * (1 - vd), vdwShift2Up=coeff * (1 - vu), vdwShift2Down=coeff * (1 - vd),
                 alchDecouple=0, alchVdwShiftCoeff=coeff)
        return f

    lam = dict(vu=0.4, vd=0.7, eu=0.3, ed=0.9)
    r = _alch(s, part, factors(**lam), A, B, beta)
    h = 1e-5
    checks = [("vu", "e_vdw", "e_vdw_ti_1"), ("vd", "e_vdw", "e_vdw_ti_2"), ("eu", "e_elec", "e_elec_ti_1"),
              ("ed", "e_elec", "e_elec_ti_2"), ("eu", "e_slow", "e_slow_ti_1"), ("ed", "e_slow", "e_slow_ti_2")]
    for key, energy, ti in checks:
        lp, lm = dict(lam), dict(lam)
        lp[key] += h
        lm[key] -= h
        g = (_alch(s, part, factors(**lp), A, B, beta)[energy] -
             _alch(s, part, factors(**lm), A, B, beta)[energy]) / (2 * h)
        assert abs(r[ti]) > 1e-3, ti
        assert abs(g - r[ti]) <= 1e-6 * max(1.0, abs(r[ti])), (ti, g, r[ti])
    # reference-only systems have no TI terms
    r0 = _alch(s, np.zeros(s.natoms, np.int8), factors(**lam), A, B, beta)
    assert all(r0[k] == 0.0 for k in ("e_vdw_ti_1", "e_vdw_ti_2", "e_elec_ti_1", "e_elec_ti_2"))
