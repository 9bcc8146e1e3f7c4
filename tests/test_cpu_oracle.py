"""Pins of the oracle against what the mathematics fixes (no GPU).

Each check is chosen so a plausible mistake (dropped term, wrong sign or
index, transposed operand) fails it:
  * two-body closed forms: fast + slow electrostatics = qq erfc(beta r)/r, LJ
    = A/r^12 - B/r^6 below the switching distance, zero at the cutoff;
  * energy/force consistency by central finite differences on a small box;
  * virial trace = -dE/ds under uniform scaling s of positions and box;
  * sum of forces = 0; pair count vs a k-d tree;
  * canonical FP32 pair sets vs scipy's pair enumeration.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

COULOMB = 332.0636


def _two_body(r, qq=1.0, A=0.0, B=0.0, rc=12.0, rs=10.0, beta=0.3):
    pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
    box = np.array([40.0, 40.0, 40.0])
    q = np.array([1.0, qq])
    Am = np.array([[A]]) if A or B else np.zeros((1, 1))
    Bm = np.array([[B]]) if A or B else np.zeros((1, 1))
    return orc.nb_fp64(pos, q, np.zeros(2, np.int32), Am, Bm, np.zeros(3, np.int64), np.zeros(1, np.int32), box,
                       rc, rs, beta)


@pytest.mark.parametrize("r", [0.9, 2.5, 6.0, 9.99, 11.5])
def test_two_body_electrostatics_closed_form(r):
    beta = 0.3
    o = _two_body(r, qq=-0.7, beta=beta)
    want = -0.7 * math.erfc(beta * r) / r
    assert abs(o["e_elec"] + o["e_slow"] - want) < 1e-12 * max(1.0, abs(want))
    # fast part vanishes at the cutoff with zero slope; check it is 1/r - 3/(2rc) + r^2/(2rc^3)
    assert abs(o["e_elec"] - (-0.7) * (1 / r - 1.5 / 12 + 0.5 * r * r / 12 ** 3)) < 1e-12
    # force = -dE/dr along x on atom 1
    h = 1e-6
    ep = _two_body(r + h, qq=-0.7, beta=beta)
    em = _two_body(r - h, qq=-0.7, beta=beta)
    dE = ((ep["e_elec"] + ep["e_slow"]) - (em["e_elec"] + em["e_slow"])) / (2 * h)
    assert abs((o["f"][1, 0] + o["fs"][1, 0]) + dE) < 1e-6 * max(1.0, abs(dE))


@pytest.mark.parametrize("r", [3.0, 9.5, 10.5, 11.9, 12.5])
def test_two_body_lj_closed_form(r):
    A, B = 5.0e5, 6.0e2
    o = _two_body(r, qq=0.0, A=A, B=B)
    E = A / r ** 12 - B / r ** 6
    if r <= 10.0:
        want = E
    elif r < 12.0:
        S = (144 - r * r) ** 2 * (144 + 2 * r * r - 300) / (144 - 100) ** 3
        want = E * S
    else:
        want = 0.0
    assert abs(o["e_vdw"] - want) < 1e-12 * max(1.0, abs(E))
    if r < 12.0:
        h = 1e-6
        dE = (_two_body(r + h, 0.0, A, B)["e_vdw"] - _two_body(r - h, 0.0, A, B)["e_vdw"]) / (2 * h)
        assert abs(o["f"][1, 0] + dE) < 1e-5 * max(1.0, abs(dE))


def _small_system():
    return systems.small_box(n_waters=60, n_chains=2, chain_len=12, box=16.0, cutoff=6.0, switchdist=5.0,
                             pairlistdist=7.0, seed=5)


def _oracle(s, pos=None, box=None):
    beta = systems.ewald_coefficient(s.params.cutoff, s.params.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    return orc.nb_fp64(s.pos if pos is None else pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr,
                       s.excl_idx, s.box if box is None else box, s.params.cutoff, s.params.switchdist, beta)


def _etot(o):
    return o["e_vdw"] + o["e_elec"] + o["e_slow"]


def test_oracle_finite_difference_forces():
    s = _small_system()
    # pull atoms apart slightly so no pair sits exactly at the cutoff band
    o = _oracle(s)
    F = o["f"] + o["fs"]
    rng = np.random.default_rng(0)
    h = 1e-5
    for a in rng.choice(s.natoms, 12, replace=False):
        for d in range(3):
            p = s.pos.copy(); p[a, d] += h
            m = s.pos.copy(); m[a, d] -= h
            dE = (_etot(_oracle(s, p)) - _etot(_oracle(s, m))) / (2 * h)
            assert abs(F[a, d] + dE) < 2e-4 * max(1.0, abs(dE)), (a, d)


def test_oracle_virial_trace_is_scaling_derivative():
    s = _small_system()
    o = _oracle(s)
    h = 1e-6
    ep = _etot(_oracle(s, s.pos * (1 + h), s.box * (1 + h)))
    em = _etot(_oracle(s, s.pos * (1 - h), s.box * (1 - h)))
    dEds = (ep - em) / (2 * h)
    tr = np.trace(o["virial"] + o["virial_slow"])
    assert abs(tr + dEds) < 1e-4 * max(1.0, abs(dEds))
    # symmetric for central pair forces
    assert np.allclose(o["virial"], o["virial"].T, rtol=1e-12, atol=1e-9)


def test_oracle_momentum_and_pair_count():
    s = _small_system()
    o = _oracle(s)
    assert np.abs(o["f"].sum(0)).max() < 1e-9 * np.abs(o["f"]).sum()
    tree = cKDTree(np.mod(s.pos, s.box), boxsize=s.box)
    pairs = tree.query_pairs(s.params.cutoff, output_type="ndarray")
    excl = set(map(tuple, s.exclusion_pairs().tolist()))
    n = sum(1 for a, b in pairs.tolist() if (min(a, b), max(a, b)) not in excl)
    assert o["npairs"] == n


def test_canonical_pair_sets_match_kdtree():
    s = _small_system()
    q = 2.0 ** -17
    pos = np.round(np.mod(s.pos, s.box) / q) * q
    p = s.params
    o = orc.pairs_fp32(pos, s.excl_ptr, s.excl_idx, s.box, np.float32(p.cutoff ** 2),
                       np.float32(p.pairlistdist ** 2), want_pairs=True)
    tree = cKDTree(pos, boxsize=s.box)
    pairs = tree.query_pairs(p.pairlistdist, output_type="ndarray")
    excl = set(map(tuple, s.exclusion_pairs().tolist()))
    kd = {(min(a, b), max(a, b)) for a, b in pairs.tolist()} - excl
    got = set(map(tuple, o["pairs"].tolist()))
    # the FP32 rule may differ from FP64 only for pairs within a few ulp of the cutoff
    diff = got ^ kd
    for a, b in diff:
        d = pos[b] - pos[a]
        d -= s.box * np.round(d / s.box)
        assert abs(np.sqrt((d * d).sum()) - p.pairlistdist) < 1e-5
    assert o["n_list"] == len(got)
    # checksums agree with the numpy hash of the same set
    h = orc.pair_hash(o["pairs"][:, 0], o["pairs"][:, 1], s.natoms)
    assert int(np.bitwise_xor.reduce(h)) == o["xor_list"]


def test_ewald_coefficient():
    b = systems.ewald_coefficient(12.0, 1e-6)
    assert abs(b - 0.257952) < 5e-6
    # the bisection converges onto erfc(beta rc)/rc = tol
    assert abs(math.erfc(b * 12.0) / 12.0 - 1e-6) < 1e-12
    assert abs(L.ewald_coefficient(12.0, 1e-6) - b) < 1e-15
