"""Every force term of the MD step is the gradient of its energy: along a
short NVE trajectory of a water + chain box, the change of each term's
energy equals minus the work of its force (trapezoid rule over each step),
term by term: nonbonded fast (LJ + fast electrostatics), slow (real-space
slow part + exclusion correction), PME reciprocal and bonded.

The trapezoid rule leaves an error of order dt^2 per unit time; at 0.5 fs
over 300 steps the accumulated mismatch of each term must stay below 0.1% of
the energy that term exchanged (measured: 2e-6 to 7e-5) (the sum of |dE| over the steps).  A force
with a wrong factor or a missing piece exchanges energy without matching work
and fails by orders of magnitude more.
"""
import numpy as np
import pytest

from mi355nb import md
from mi355nb import systems

pytestmark = pytest.mark.gpu

TERMS = ("fast", "slow", "recip", "bonded")


def test_each_force_term_does_the_work_of_its_energy(torch_cuda):
    s = systems.small_box(n_waters=216, n_chains=2, chain_len=20)
    sim = md.NVE(s, dt_fs=1.0, temperature=300.0, seed=3)
    print(f"[work-energy] {s.natoms} atoms: minimise, equilibrate at 1 fs", flush=True)
    sim.minimize(200)
    sim._init_velocities(300.0, 4)
    k = sim.equilibrate(200, 300.0, every=10)
    dt = 0.5
    sim.dt = dt / md.TIMEFACTOR
    sim.dt_fs = dt
    sim._forces(energy=True)
    prev = sim.components()
    mismatch = dict.fromkeys(TERMS, 0.0)
    exchanged = dict.fromkeys(TERMS, 0.0)
    for i in range(300):
        sim.step(k + i, energy=True)
        cur = sim.components()
        dx = cur["pos"] - prev["pos"]
        for t in TERMS:
            de = cur["energy"][t] - prev["energy"][t]
            work = -0.5 * float(((cur["force"][t] + prev["force"][t]) * dx).sum())
            mismatch[t] += de - work
            exchanged[t] += abs(de)
        prev = cur
    sim.close()
    for t in TERMS:
        rel = abs(mismatch[t]) / exchanged[t]
        print(f"    {t}: dE - W {mismatch[t]:.3e} kcal/mol, exchanged {exchanged[t]:.3e}, rel {rel:.2e}")
        assert exchanged[t] > 0.0
        assert rel < 1e-3, t
