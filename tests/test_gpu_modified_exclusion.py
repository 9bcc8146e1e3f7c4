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


@pytest.mark.parametrize("method,decouple", [("fep", False), ("ti", False), ("fep", True)])
def test_modified_exclusions_alchemical_batch(torch_cuda, method, decouple):
    """Alchemical route (pswitch): a four-system batch whose first chain is
    partition 1 and second chain partition 2, the chains' 1-4 pairs plus one
    cross-partition pair per system listed as modified; the correction's
    energies at lambda, the FEP sums at lambda2 or the TI sums, and virials
    within 1e-9 of the alchemical oracle system by system (FP64 from the same
    float32 inputs), and pair kernel + correction forces within 1e-5 of the
    oracles' sum."""
    torch = torch_cuda
    members = systems.batch_systems(n_systems=4, seed=9, n_min=3000, n_max=8000)
    c = ComputeNonbonded.from_systems([m.system for m in members], stream=torch.cuda.current_stream().cuda_stream)
    ptrs, idxs, parts, per_sys, off = [np.zeros(1, np.int64)], [], [], [], 0
    for m in members:
        s = m.system
        L_ = s.chain_len
        pk = m.partition.copy()
        pk[:L_] = 1
        pk[L_:2 * L_] = 2
        mp, mi = systems.chain_pairs_14(s)
        # one cross-partition pair (first bead of chain 1 with first bead of chain 2)
        lists = [list(mi[mp[a]:mp[a + 1]]) for a in range(s.natoms)]
        lists[0].append(L_)
        mp = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int64)
        mi = np.array([b for x in lists for b in x], np.int32)
        per_sys.append((mp, mi, pk))
        ptrs.append(mp[1:] + ptrs[-1][-1])
        idxs.append(mi + off)
        parts.append(pk)
        off += s.natoms
    A14, B14 = systems.lj14_tables(members[0].system)
    alch = [L.alch_data(c.params, m.lam, min(1.0, m.lam + 0.05), elec_lambda_start=0.5, vdw_lambda_end=1.0,
                        vdw_shift_coeff=5.0, decouple=decouple) for m in members]
    c.set_alch(np.concatenate(parts), alch, method=method)
    c.set_modified_exclusions(np.concatenate(ptrs), np.concatenate(idxs), S14, A14, B14)
    print(f"[modified exclusions alch {method} decouple={decouple}] {c.natoms} atoms", flush=True)
    c.launch_work(do_energy=False, do_virial=False, do_slow=True, do_pairlist=True)
    r0 = c.finish_batch()
    ev = c.modified_exclusion_correction()
    r1 = c.finish_batch()
    torch.cuda.synchronize()
    ev = ev.cpu().numpy()
    pos, boxes, st = c.effective_positions(), c.system_boxes(), c.sys_atom_start
    keys = ("e_vdw", "e_elec", "e_slow") + (("e_vdw_s", "e_elec_s", "e_slow_s") if method == "fep" else ())
    ti_keys = ("e_vdw_ti_1", "e_vdw_ti_2", "e_elec_ti_1", "e_elec_ti_2", "e_slow_ti_1", "e_slow_ti_2")
    col = dict(e_vdw=0, e_elec=1, e_slow=2, e_vdw_s=21, e_elec_s=22, e_slow_s=23,
               **{k: 24 + i for i, k in enumerate(ti_keys)})
    for k, m in enumerate(members):
        s = m.system
        a, b = st[k], st[k + 1]
        mp, mi, pk = per_sys[k]
        p = s.params
        A, B = systems.lj_tables(s)
        o = orc.mod_excl_fp64(pos[a:b], _q32(s), s.vdw_type, A, B, A14, B14, mp, mi, boxes[k], p.cutoff,
                              p.switchdist, S14, partition=pk, alch=alch[k])
        scale = max(abs(o["e_vdw"]), abs(o["e_elec"]), abs(o["e_slow"]))
        for key in keys + (ti_keys if method == "ti" else ()):
            err = abs(ev[k, col[key]] - o[key])
            print(f"    [sys {k}] {key} gpu {ev[k, col[key]]:.12e} oracle {o[key]:.12e}", flush=True)
            assert err <= 1e-9 * max(abs(o[key]), 1e-3 * scale), key
        if method == "ti":
            assert not ev[k, 21:24].any()
        else:
            assert not ev[k, 24:30].any()
        for off_, key in ((3, "virial"), (12, "virial_slow")):
            assert np.abs(ev[k, off_:off_ + 9] - o[key].reshape(-1)).max() <= 1e-9 * np.abs(o[key]).max(), key
        beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
        nb = orc.nb_fp64_alch(pos[a:b], s.charge * math.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr, s.excl_idx,
                              boxes[k], p.cutoff, p.switchdist, beta, pk, alch[k])
        ef, es = rel_rms(r1[2][a:b], nb["f"] + o["f"]), rel_rms(r1[3][a:b], nb["fs"] + o["fs"])
        print(f"    [sys {k}] totals fast {ef:.2e} slow {es:.2e}", flush=True)
        assert ef < 1e-5 and es < 1e-5
    c.close()
