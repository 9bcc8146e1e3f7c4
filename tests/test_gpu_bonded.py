"""GPU harmonic bonds and angles (FP64, through the C ABI) against the
oracle (pinned in test_cpu_bonded_oracle.py), and NVE of a small solvated
box with chains: minimization, velocity-rescaled equilibration, NVE.

Tolerances: bonded forces and energies within 1e-10 relative of the FP64
oracle on the same positions; NVE over 300 steps at 2 fs: total-energy RMS
below 5% of the kinetic-energy RMS, water geometry exact to 1e-9 A.
"""
import ctypes as C

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import md
from mi355nb import systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_bonded_forces_match_oracle(torch_cuda):
    torch = torch_cuda
    s = systems.small_box(n_waters=50, n_chains=4, chain_len=25, seed=5)
    terms = systems.chain_bonded_terms(s)
    bonds, bk, br0, angles, ak, at0 = terms
    rng = np.random.default_rng(2)
    pos = s.pos + rng.normal(scale=0.05, size=s.pos.shape)
    h = C.c_void_p()
    L.check(L.lib().mi355nb_bonded_create(s.natoms, len(bonds), bonds.ctypes.data, bk.ctypes.data, br0.ctypes.data,
                                          len(angles), angles.ctypes.data, ak.ctypes.data, at0.ctypes.data,
                                          C.byref(h)), "bonded_create")
    dpos = torch.tensor(pos, dtype=torch.float64, device="cuda")
    f = torch.zeros_like(dpos)
    e = torch.zeros(2, dtype=torch.float64, device="cuda")
    print("[bonded] forces", flush=True)
    L.check(L.lib().mi355nb_bonded_forces(h, dpos.data_ptr(), np.ascontiguousarray(s.box), f.data_ptr(),
                                          e.data_ptr(), None), "bonded_forces")
    torch.cuda.synchronize()
    Eb, Ea, F = orc.bonded_fp64(pos, s.box, *terms)
    got = f.cpu().numpy()
    ee = e.cpu().numpy()
    assert np.abs(got - F).max() < 1e-10 * np.abs(F).max()
    assert abs(ee[0] - Eb) < 1e-10 * Eb and abs(ee[1] - Ea) < 1e-10 * Ea
    L.lib().mi355nb_bonded_destroy(h)


def test_nve_small_solvated_box(torch_cuda):
    s = systems.small_box(n_waters=216, n_chains=2, chain_len=20, seed=11)
    sim = md.NVE(s, dt_fs=2.0, temperature=300.0, seed=2)
    print("[nve chains] minimize 200", flush=True)
    e0 = sim.minimize(200)
    print(f"    after minimization: potential {e0['potential']:.2f}", flush=True)
    sim._init_velocities(300.0, 3)
    k = sim.equilibrate(400, 300.0, every=10)
    print("[nve chains] 300 NVE steps", flush=True)
    rec = sim.run(300, energy_every=10, start=k)
    fit = md.drift_fit(rec, 2.0, sim.dof)
    print(f"    T {fit['mean_temperature']:.1f} K; total rms {fit['total_rms']:.3f} kinetic rms "
          f"{fit['kinetic_rms']:.3f}; bond {rec[-1]['bond']:.2f} angle {rec[-1]['angle']:.2f}")
    pos = sim.positions()
    w = pos[s.n_beads:].reshape(-1, 3, 3)
    assert np.abs(np.linalg.norm(w[:, 0] - w[:, 1], axis=1) - systems.TIP3P_ROH).max() < 1e-9
    assert all(np.isfinite(r["total"]) for r in rec)
    assert fit["total_rms"] < 0.05 * fit["kinetic_rms"]
    sim.close()
