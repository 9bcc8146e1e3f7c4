"""Pins of the oracle's rigid-water constraints: SETTLE (closed form) equals
an independent iterative SHAKE on the same displaced waters; the result has
exact TIP3P distances and keeps the centre of mass of the unconstrained
positions; the velocity constraint makes every distance stationary and
conserves momentum."""
import numpy as np

from mi355nb import systems
from oracle import oracle as orc

MO, MH = systems.MASS_O, systems.MASS_H
ROH, RHH = systems.TIP3P_ROH, systems.TIP3P_RHH


def _waters(nw, seed, step=0.05):
    rng = np.random.default_rng(seed)
    s = systems.water_box_liquid()
    ref = s.pos[: 3 * nw].reshape(nw, 3, 3)
    pos = ref + rng.normal(scale=step, size=ref.shape)
    return ref, pos


def test_settle_equals_shake():
    ref, pos = _waters(200, 1)
    a = orc.settle_positions(ref, pos, MO, MH, ROH, RHH)
    b = orc.shake_positions(ref, pos, MO, MH, ROH, RHH)
    assert np.abs(a - b).max() < 1e-10


def test_settle_geometry_and_centre_of_mass():
    ref, pos = _waters(500, 2, step=0.1)
    a = orc.settle_positions(ref, pos, MO, MH, ROH, RHH)
    d = lambda x, i, j: np.linalg.norm(x[:, i] - x[:, j], axis=1)
    assert np.abs(d(a, 0, 1) - ROH).max() < 1e-12
    assert np.abs(d(a, 0, 2) - ROH).max() < 1e-12
    assert np.abs(d(a, 1, 2) - RHH).max() < 1e-12
    m = np.array([MO, MH, MH])
    com = lambda x: (x * m[None, :, None]).sum(1) / m.sum()
    assert np.abs(com(a) - com(pos)).max() < 1e-12
    # already constrained waters are left where they are
    assert np.abs(orc.settle_positions(ref, ref, MO, MH, ROH, RHH) - ref).max() < 1e-12


def test_velocity_constraint():
    ref, _ = _waters(100, 3)
    rng = np.random.default_rng(4)
    v = rng.normal(size=ref.shape)
    u = orc.settle_velocities(ref, v, MO, MH)
    m = np.array([MO, MH, MH])
    for i, j in ((0, 1), (0, 2), (1, 2)):
        assert np.abs(((u[:, i] - u[:, j]) * (ref[:, i] - ref[:, j])).sum(1)).max() < 1e-12
    assert np.abs((u * m[None, :, None]).sum(1) - (v * m[None, :, None]).sum(1)).max() < 1e-12
    # a rigid-body motion is left unchanged
    w = rng.normal(size=(100, 3))
    rigid = np.cross(w[:, None, :], ref - ref.mean(1, keepdims=True)) + rng.normal(size=(100, 1, 3))
    assert np.abs(orc.settle_velocities(ref, rigid, MO, MH) - rigid).max() < 1e-12
