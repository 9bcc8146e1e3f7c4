"""Pins of the PME real-space oracle (no GPU).

  * B-spline weights: partition of unity (sum theta = 1), sum dtheta = 0,
    order-4 weights equal the closed-form cubics of PmeRealSpace.C:272-295;
  * NAMD's CPU convention (PmeBase.inl scale_coordinates + fill_charges)
    and the GPU convention (CudaPmeSolverUtilKernel.cu) give the same grid up
    to the cyclic index shift (K+order-1)/2 - K/2 - (order-1) for even K, and
    the same fractional forces;
  * gather is minus the gradient of E = sum_i q_i sum_p theta_p(x_i) G_p
    (central finite differences in the fractional coordinates).
"""
import numpy as np
import pytest

from oracle import oracle as orc


@pytest.mark.parametrize("order", [4, 6, 8])
def test_bspline_partition_of_unity(order):
    for w in (0.0, 0.13, 0.5, 0.77, 0.999):
        th, dth = orc.bspline_gpu(w, order)
        assert abs(th.sum() - 1.0) < 1e-14
        assert abs(dth.sum()) < 1e-14
        assert (th >= 0).all()


def test_order4_matches_closed_form_cubics():
    for f in np.linspace(0, 1, 17, endpoint=False):
        th, dth = orc.bspline_gpu(f, 4)
        M = [((-1 / 6 * f + 0.5) * f - 0.5) * f + 1 / 6, ((0.5 * f - 1.0) * f) * f + 2 / 3,
             ((-0.5 * f + 0.5) * f + 0.5) * f + 1 / 6, 1 / 6 * f ** 3]
        dM = [(-0.5 * f + 1.0) * f - 0.5, (1.5 * f - 2.0) * f, (-1.5 * f + 1.0) * f + 0.5, 0.5 * f * f]
        assert np.allclose(th, M, atol=1e-15) and np.allclose(dth, dM, atol=1e-15)


def _random_system(n, L, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L
    q = rng.uniform(-1, 1, n)
    return pos, q


@pytest.mark.parametrize("order", [4, 6, 8])
@pytest.mark.parametrize("K", [(32, 32, 32), (24, 28, 30)])
def test_cpu_and_gpu_conventions_agree(order, K):
    L = np.array([31.0, 33.0, 29.0])
    pos, q = _random_system(200, L, order)
    recip = np.diag(1.0 / L)
    # CPU convention (FP64 inputs)
    p = orc.pme_scale_coordinates(pos, np.zeros(3), recip, K, order)
    g_cpu = orc.pme_spread_cpu(p, q, K, 6 if order == 6 else order)
    # GPU convention: w - (floor(w+0.5) - 0.5) in [0,1); fed as float32 to the oracle
    w = pos @ recip.T
    w = w - (np.floor(w + 0.5) - 0.5)
    at = np.zeros((len(q), 4), np.float32)
    at[:, :3] = w
    at[:, 3] = q
    g_gpu = orc.pme_spread_gpu(at, K, order)
    g = g_gpu
    for d, Kd in enumerate(K):
        shift = (Kd + order - 1) // 2 - Kd // 2 - (order - 1)
        axis = 2 - d  # grid arrays are (z, y, x)
        g = np.roll(g, shift, axis=axis)
    # float32 fractional inputs on the GPU side: agreement to single precision
    assert np.abs(g - g_cpu).max() < 2e-5 * np.abs(g_cpu).max()
    # same physics for forces on an arbitrary potential grid
    rng = np.random.default_rng(1)
    G = rng.normal(size=g_cpu.shape)
    f_cpu = orc.pme_gather_cpu(p, q, K, order, G)
    Gs = G
    for d, Kd in enumerate(K):
        shift = (Kd + order - 1) // 2 - Kd // 2 - (order - 1)
        Gs = np.roll(Gs, -shift, axis=2 - d)
    f_gpu = orc.pme_gather_gpu(at, K, order, Gs)
    assert np.abs(f_gpu - f_cpu).max() < 1e-4 * np.abs(f_cpu).max()


def test_gather_is_minus_gradient():
    K = (20, 22, 24)
    order = 4
    rng = np.random.default_rng(7)
    G = rng.normal(size=(K[2], K[1], K[0]))
    at = np.zeros((5, 4), np.float32)
    at[:, :3] = rng.uniform(0.05, 0.95, (5, 3))
    at[:, 3] = rng.uniform(-1, 1, 5)
    f = orc.pme_gather_gpu(at, K, order, G)

    def energy(a):
        g = orc.pme_spread_gpu(a, K, order)
        return float((g * G).sum())

    for i in range(5):
        for d in range(3):
            h = 3e-4
            ap = at.copy(); ap[i, d] += h
            am = at.copy(); am[i, d] -= h
            dE = (energy(ap) - energy(am)) / (np.float64(ap[i, d]) - np.float64(am[i, d]))
            assert abs(f[i, d] + dE) < 2e-3 * max(1.0, abs(dE)), (i, d, f[i, d], dE)
