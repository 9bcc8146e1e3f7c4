"""GPU parity of the modified (scaled 1-4) exclusion correction (FP64,
through the C ABI) against the oracle's restatement of
modifiedExclusionForce (src/ComputeBondedCUDAKernel.cu:316-632), pinned in
test_cpu_mod_excl_oracle.py.  The chains' 1-4 pairs are the modified
exclusions, scale14 0.8333, 1-4 LJ tables (0.5 A, 0.7 B).

Tolerances: the correction's energies and virials within 1e-9 relative (FP64
from the same float32 inputs); the fast and slow forces of pair kernel +
correction within 1e-5 relative RMS of the FP64 oracles' sum (the pair
kernel's bar); the correction's own forces, read as the difference of the
launch's forces after and before it, elementwise within the two FP32
roundings that adding them makes, |dF - F| <= 2^-23 (|F| + |F_after|) (a
factor 2 over the bound 2^-24 per rounding; the FP64 values themselves agree
to ~1e-15).
"""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from mi355nb.compute import ComputeNonbonded
from nbcheck import COULOMB, rel_rms
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

S14 = 0.8333


def _q32(s):
    return (s.charge * math.sqrt(COULOMB)).astype(np.float32).astype(np.float64)


def _run(torch, c):
    """Forces of one launch read before and after the correction (finish
    only copies them out), and the correction's per-system energy/virial."""
    c.launch_work(do_energy=False, do_virial=False, do_slow=True, do_pairlist=True)
    r0 = c.finish_batch()
    ev = c.modified_exclusion_correction()
    r1 = c.finish_batch()
    torch.cuda.synchronize()
    return r0, r1, ev.cpu().numpy()


def _check(tag, s, pos, box, f0, fs0, f1, fs1, ev, force_switch=False):
    p = s.params
    A, B = systems.lj_tables(s)
    A14, B14 = systems.lj14_tables(s)
    mp, mi = systems.chain_pairs_14(s)
    m = orc.mod_excl_fp64(pos, _q32(s), s.vdw_type, A, B, A14, B14, mp, mi, box, p.cutoff, p.switchdist, S14,
                          vdw_force_switching=force_switch)
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    o = orc.nb_fp64(pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr, s.excl_idx, box, p.cutoff,
                    p.switchdist, beta, vdw_force_switching=force_switch)
    assert m["npairs"] > 0
    for k, key in enumerate(("e_vdw", "e_elec", "e_slow")):
        err = abs(ev[k] - m[key]) / abs(m[key])
        print(f"    [{tag}] {key} gpu {ev[k]:.12e} oracle {m[key]:.12e} rel {err:.2e}", flush=True)
        assert err < 1e-9, key
    for off, key in ((3, "virial"), (12, "virial_slow")):
        err = np.abs(ev[off:off + 9] - m[key].reshape(-1)).max() / np.abs(m[key]).max()
        assert err < 1e-9, key
    ef_tot = rel_rms(f1, o["f"] + m["f"])
    es_tot = rel_rms(fs1, o["fs"] + m["fs"])
    # the correction alone: worst ratio of its error to the FP32 rounding bound
    rf = (np.abs((f1 - f0) - m["f"]) / (2.0 ** -23 * (np.abs(m["f"]) + np.abs(f1)) + 1e-30)).max()
    rs = (np.abs((fs1 - fs0) - m["fs"]) / (2.0 ** -23 * (np.abs(m["fs"]) + np.abs(fs1)) + 1e-30)).max()
    print(f"    [{tag}] totals fast {ef_tot:.2e} slow {es_tot:.2e}; correction alone / FP32 bound: "
          f"fast {rf:.2f} slow {rs:.2f}", flush=True)
    assert ef_tot < 1e-5 and es_tot < 1e-5
    assert rf <= 1.0 and rs <= 1.0


@pytest.mark.parametrize("name,force_switch", [("small", False), ("small", True), ("solvated96k", False)])
def test_modified_exclusions_match_oracle(torch_cuda, name, force_switch):
    torch = torch_cuda
    s = systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7) if name == "small" \
        else systems.solvated_96k()
    c = ComputeNonbonded.from_system(s, stream=torch.cuda.current_stream().cuda_stream)
    if force_switch:
        c.set_vdw_force_switching(True)
    A14, B14 = systems.lj14_tables(s)
    mp, mi = systems.chain_pairs_14(s)
    c.set_modified_exclusions(mp, mi, S14, A14, B14)
    print(f"[modified exclusions {name} fs={force_switch}] {s.natoms} atoms, {mi.size} pairs", flush=True)
    r0, r1, ev = _run(torch, c)
    _check(name, s, c.effective_positions(), c.system_boxes()[0], r0[2], r0[3], r1[2], r1[3], ev[0], force_switch)
    c.close()


def test_modified_exclusions_batch(torch_cuda):
    torch = torch_cuda
    members = systems.batch_systems(n_systems=4, seed=9, n_min=3000, n_max=8000)
    c = ComputeNonbonded.from_systems([m.system for m in members], stream=torch.cuda.current_stream().cuda_stream)
    ptrs, idxs, off = [np.zeros(1, np.int64)], [], 0
    for m in members:
        mp, mi = systems.chain_pairs_14(m.system)
        ptrs.append(mp[1:] + ptrs[-1][-1])
        idxs.append(mi + off)
        off += m.system.natoms
    A14, B14 = systems.lj14_tables(members[0].system)
    c.set_modified_exclusions(np.concatenate(ptrs), np.concatenate(idxs), S14, A14, B14)
    print(f"[modified exclusions batch] {c.natoms} atoms", flush=True)
    r0, r1, ev = _run(torch, c)
    pos = c.effective_positions()
    boxes = c.system_boxes()
    st = c.sys_atom_start
    for k, m in enumerate(members):
        a, b = st[k], st[k + 1]
        _check(f"sys {k}", m.system, pos[a:b], boxes[k], r0[2][a:b], r0[3][a:b], r1[2][a:b], r1[3][a:b], ev[k])
    c.close()


def test_modified_exclusions_argument_and_state_errors(torch_cuda):
    s = systems.small_box()
    c = ComputeNonbonded.from_system(s)
    A14, B14 = systems.lj14_tables(s)
    mp, mi = systems.chain_pairs_14(s)
    with pytest.raises(L.Mi355nbError):   # no modified exclusions set
        c.modified_exclusion_correction()
    # a pair that is also a full exclusion (1-2 bond) is refused
    bad = mi.copy()
    bad[0] = 1
    with pytest.raises(L.Mi355nbError):
        c.set_modified_exclusions(mp, bad, S14, A14, B14)
    c.set_modified_exclusions(mp, mi, S14, A14, B14)
    with pytest.raises(L.Mi355nbError):   # needs a launch with slow forces first
        c.modified_exclusion_correction()
    c.set_modified_exclusions(None, None, 1.0, None, None)
    with pytest.raises(L.Mi355nbError):
        c.modified_exclusion_correction()
    c.close()
