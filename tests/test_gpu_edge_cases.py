"""Edge cases of the nonbonded path against the oracle:

* pairs exactly at the cutoff: a rock-salt lattice of spacing 4 A with
  rc = 12 A has every pair with a^2 + b^2 + c^2 = 9 lattice steps at exactly
  12.000 A; NAMD's CPU rule r^2 <= rc^2 includes them, and the pair counts
  and pair sets must match the oracle's exact rule bit for bit;
* a box of exactly twice the pairlist distance (two patches per edge, each
  patch its own neighbour's image in both directions);
* a sparse box (60 atoms in 27 patches, most tiles padding, empty patches);
* an isolated system with no exclusions and charge-only atoms.
Forces, energies and virial within 1e-5 of the FP64 oracle (absolute 2e-4
kcal/mol/A where the forces vanish by symmetry)."""
import numpy as np
import pytest

from mi355nb import systems
from nbcheck import gpu_compute, oracle_for, rel_rms

import test_gpu_nonbonded as T

pytestmark = pytest.mark.gpu


def _random_system(n, box, seed, n_excl=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, box, (n, 3))
    charge = rng.choice([-0.5, 0.5], n)
    vt = rng.integers(0, 2, n).astype(np.int32)
    pairs = set()
    while len(pairs) < n_excl:
        a, b = sorted(rng.choice(n, 2, replace=False).tolist())
        pairs.add((a, b))
    nb = [[] for _ in range(n)]
    for a, b in pairs:
        nb[a].append(b)
        nb[b].append(a)
    ptr = np.zeros(n + 1, np.int64)
    for a in range(n):
        ptr[a + 1] = ptr[a] + len(nb[a])
    idx = np.array([b for a in range(n) for b in sorted(nb[a])], np.int32)
    return systems.System(name=f"random{n}", pos=pos, charge=charge, vdw_type=vt,
                          lj_sigma=np.array([3.0, 3.5]), lj_eps=np.array([0.1, 0.15]), excl_ptr=ptr, excl_idx=idx,
                          box=np.array([box, box, box], float), params=systems.NbParams())


def test_pairs_exactly_at_the_cutoff(torch_cuda):
    s = systems.nacl_crystal(nside=8, spacing=4.0)
    print(f"[edge nacl] {s.natoms} ions, box {s.box[0]}", flush=True)
    c, r = gpu_compute(s)
    o = T._check_pairs(s, c, r, want_pairs=True)
    # the lattice has pairs at exactly rc (r^2 = 144 in exact arithmetic):
    # the count includes them, as r^2 <= rc^2 asks
    x = s.pos
    dx = x[:, None, :] - x[None, :, :]
    dx -= s.box * np.round(dx / s.box)
    r2 = np.triu((dx * dx).sum(-1), 1) + np.tril(np.full((s.natoms, s.natoms), np.inf))
    at_rc = int((r2 == 144.0).sum())
    print(f"    pairs at exactly rc: {at_rc}; below: {int((r2 < 144.0).sum())}")
    assert at_rc > 0 and o["n_cut"] == int((r2 <= 144.0).sum()) == r["stat"]["numPairsCutoff"]
    of = oracle_for(s, c.effective_positions(), c.info()["box"])
    # forces vanish by symmetry: compare absolutely
    assert np.abs(r["f"] - of["f"]).max() < 2e-4 and np.abs(r["fs"] - of["fs"]).max() < 2e-4
    coul = of["e_elec"] + of["e_slow"]
    assert abs(r["e_elec"] + r["e_slow"] - coul) <= 1e-5 * abs(coul)
    c.close()


def test_box_of_twice_the_pairlist_distance(torch_cuda):
    s = systems.small_box(n_waters=160, n_chains=1, chain_len=20, box=18.0)
    assert s.box[0] == 2 * s.params.pairlistdist
    print(f"[edge 2x pairlist] {s.natoms} atoms", flush=True)
    c, r = gpu_compute(s)
    assert tuple(c.info()["patchDims"]) == (2, 2, 2)
    T._check_against_oracle(s, c, r)
    T._check_pairs(s, c, r, want_pairs=True)
    c.close()


def test_sparse_box_with_empty_patches(torch_cuda):
    s = _random_system(60, 45.0, seed=4, n_excl=10)
    print(f"[edge sparse] {s.natoms} atoms", flush=True)
    c, r = gpu_compute(s)
    assert tuple(c.info()["patchDims"]) == (3, 3, 3)
    o = T._check_pairs(s, c, r, want_pairs=True)
    of = oracle_for(s, c.effective_positions(), c.info()["box"])
    if o["n_cut"]:
        assert rel_rms(r["f"], of["f"]) < 1e-5 and rel_rms(r["fs"], of["fs"]) < 1e-5
    c.close()


def test_no_exclusions_charge_only(torch_cuda):
    s = _random_system(3000, 32.0, seed=9)
    s.lj_eps = np.zeros(2)
    print(f"[edge charge-only] {s.natoms} atoms", flush=True)
    c, r = gpu_compute(s)
    T._check_against_oracle(s, c, r)
    T._check_pairs(s, c, r)
    assert r["e_vdw"] == 0.0
    c.close()


def test_switching_off(torch_cuda):
    """switching off (switchdist = cutoff): truncated LJ, no switch region
    (ComputeNonbondedUtil.C:504-515 leaves the switch inactive); forces,
    energies and virial against the oracle, and the pair sets bit-exact
    (the LJ energy jumps at the cutoff, so the inclusive cutoff rule shows)."""
    import dataclasses

    s = systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7)
    s.params = dataclasses.replace(s.params, switchdist=s.params.cutoff)
    print(f"[switching off] {s.natoms} atoms", flush=True)
    c, r = gpu_compute(s)
    T._check_against_oracle(s, c, r)
    T._check_pairs(s, c, r)
    c.close()
