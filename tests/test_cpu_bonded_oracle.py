"""Pins of the oracle's harmonic bond and angle terms (oracle.bonded_fp64):
closed forms for a single bond and a single angle, forces = -gradient of the
energy by central differences, zero net force and torque, and the build's
chain terms (built angles give zero angle energy)."""
import math

import numpy as np

from mi355nb import systems
from oracle import oracle as orc


def _one(pos, bonds=((0, 1),), angles=(), bk=300.0, r0=1.5, ak=40.0, th0=2.0):
    b = np.array(bonds, np.int32).reshape(-1, 2)
    a = np.array(angles, np.int32).reshape(-1, 3)
    return orc.bonded_fp64(np.asarray(pos, float), np.array([100.0, 100.0, 100.0]), b, np.full(len(b), bk),
                           np.full(len(b), r0), a, np.full(len(a), ak), np.full(len(a), th0))


def test_single_bond_closed_form():
    Eb, Ea, F = _one([[0, 0, 0], [1.7, 0, 0]])
    assert abs(Eb - 300.0 * 0.2 ** 2) < 1e-12 and Ea == 0.0
    # stretched bond pulls the atoms together with 2 k (r - r0)
    assert np.allclose(F, [[120.0, 0, 0], [-120.0, 0, 0]], atol=1e-10)


def test_single_angle_closed_form():
    th = 1.8
    pos = [[1.2, 0, 0], [0, 0, 0], [1.4 * math.cos(th), 1.4 * math.sin(th), 0]]
    Eb, Ea, F = _one(pos, bonds=(), angles=((0, 1, 2),), th0=2.0)
    assert abs(Ea - 40.0 * 0.2 ** 2) < 1e-12
    # the end force is tangential with magnitude 2 k (theta - theta0) / r
    assert abs(np.linalg.norm(F[0]) - 2 * 40.0 * 0.2 / 1.2) < 1e-10
    assert abs(np.dot(F[0], pos[0])) < 1e-10


def test_gradient_and_conservation():
    s = systems.small_box(n_waters=0, n_chains=3, chain_len=12, seed=3)
    terms = systems.chain_bonded_terms(s)
    rng = np.random.default_rng(1)
    pos = s.pos + rng.normal(scale=0.05, size=s.pos.shape)
    Eb, Ea, F = orc.bonded_fp64(pos, s.box, *terms)
    assert np.abs(F.sum(0)).max() < 1e-9
    # torque about the origin vanishes for an unwrapped configuration in a large box
    big = np.array([1000.0, 1000.0, 1000.0])
    p0 = pos.copy()
    for c in range(3):
        o = c * 12
        for k in range(1, 12):
            d = p0[o + k] - p0[o + k - 1]
            p0[o + k] = p0[o + k - 1] + d - s.box * np.round(d / s.box)
    _, _, F0 = orc.bonded_fp64(p0, big, *terms)
    assert np.abs(np.cross(p0, F0).sum(0)).max() < 1e-8
    h = 1e-6
    for i in (0, 5, 17, 30):
        for d in range(3):
            e = []
            for sg in (1, -1):
                q = pos.copy()
                q[i, d] += sg * h
                eb, ea, _ = orc.bonded_fp64(q, s.box, *terms)
                e.append(eb + ea)
            assert abs(-(e[0] - e[1]) / (2 * h) - F[i, d]) < 1e-5 * max(1.0, np.abs(F).max())


def test_built_angles_are_at_rest():
    s = systems.small_box(n_waters=10, n_chains=2, chain_len=20, seed=4)
    Eb, Ea, _ = orc.bonded_fp64(s.pos, s.box, *systems.chain_bonded_terms(s))
    assert Ea < 1e-20
    assert Eb < 1e-6 * len(systems.chain_bonded_terms(s)[0])
