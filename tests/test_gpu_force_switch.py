"""GPU parity of vdW force switching (NAMD's vdwForceSwitching,
src/ComputeNonbondedUtil.C:523-533, :900-919) through the C ABI against the
FP64 oracle's force-switched LJ at the same effective positions.  Bar: fast
and slow forces within 1e-5 relative RMS, e_vdw within 1e-5 relative, the
electrostatic energies within 1e-5 of the real-space Coulomb energy, the
virials within 1e-5 of their norm (the tolerances of the energy-switched
parity tests).  Switching it on and off between launches must give each
form's own result, and the combinations the kernel does not provide must be
refused.  With alchemical partitions the soft-core form of NAMD's CPU FEP/TI
code is checked against the alchemical oracle with force switching."""
import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from mi355nb.compute import ComputeNonbonded
from nbcheck import COULOMB, rel_rms
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle(s, c, force_switch):
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = systems.lj_tables(s)
    return orc.nb_fp64(c.effective_positions(), s.charge * np.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr,
                       s.excl_idx, np.asarray(c.info()["box"]), p.cutoff, p.switchdist, beta,
                       vdw_force_switching=force_switch)


def _compare(tag, r, o):
    ef, es = rel_rms(r["f"], o["f"]), rel_rms(r["fs"], o["fs"])
    print(f"[{tag}] force rel-RMS fast {ef:.2e} slow {es:.2e}", flush=True)
    assert ef < TOL and es < TOL
    coul = abs(o["e_elec"] + o["e_slow"])
    for k in ("e_vdw", "e_elec", "e_slow"):
        scale = abs(o[k]) if k == "e_vdw" else max(abs(o[k]), coul)
        err = abs(r[k] - o[k]) / max(scale, 1e-12)
        print(f"[{tag}] {k} gpu {r[k]:.10e} oracle {o[k]:.10e} rel {err:.2e}", flush=True)
        assert err < TOL, k
    for k in ("virial", "virial_slow"):
        err = np.linalg.norm(r[k] - o[k]) / np.linalg.norm(o[k])
        assert err < TOL, k


@pytest.mark.parametrize("name", ["small", "water9k", "solvated96k"])
def test_force_switch_matches_oracle(torch_cuda, name):
    s = {"small": lambda: systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7),
         "water9k": systems.water_box_9k,
         "solvated96k": systems.solvated_96k}[name]()
    c = ComputeNonbonded.from_system(s)
    c.set_vdw_force_switching(True)
    r = c.compute(True, True, True, do_pairlist=True)
    o = _oracle(s, c, True)
    _compare(s.name + "/force-switch", r, o)
    # force-only replay on the same lists gives the same forces
    r2 = c.compute(False, False, True, do_pairlist=False)
    assert rel_rms(r2["f"], r["f"]) < 1e-6
    c.close()


def test_force_switch_toggle(torch_cuda):
    s = systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7)
    c = ComputeNonbonded.from_system(s)
    r_es = c.compute(True, True, True, do_pairlist=True)
    c.set_vdw_force_switching(True)
    r_fs = c.compute(True, True, True, do_pairlist=False)
    c.set_vdw_force_switching(False)
    r_es2 = c.compute(True, True, True, do_pairlist=False)
    o_es, o_fs = _oracle(s, c, False), _oracle(s, c, True)
    _compare("toggle/energy-switch", r_es, o_es)
    _compare("toggle/force-switch", r_fs, o_fs)
    _compare("toggle/energy-switch-again", r_es2, o_es)
    # the two forms differ by more than the tolerance on this box
    assert rel_rms(o_fs["f"], o_es["f"]) > 10 * TOL
    c.close()


def test_force_switch_refused_combinations(torch_cuda):
    s = systems.small_box()
    c = ComputeNonbonded.from_system(s)
    c.compute(False, False, True, do_pairlist=True)
    c.set_force_table(True)
    with pytest.raises(L.Mi355nbError):
        c.set_vdw_force_switching(True)
    c.set_force_table(False)
    c.set_vdw_force_switching(True)
    with pytest.raises(L.Mi355nbError):
        c.set_force_table(True)
    with pytest.raises(L.Mi355nbError):
        c.force_matching_grad(np.zeros((s.natoms, 3)))
    c.set_vdw_force_switching(False)
    c.close()


TI_KEYS = ("e_vdw_ti_1", "e_vdw_ti_2", "e_elec_ti_1", "e_elec_ti_2", "e_slow_ti_1", "e_slow_ti_2")
FEP_KEYS = ("e_vdw_s", "e_elec_s", "e_slow_s")


@pytest.mark.parametrize("method", ["fep", "ti"])
def test_alchemical_force_switch_batch(torch_cuda, method):
    """Force-switched soft-core LJ (doAlchVdwForceSwitching) on a batch of six
    uneven systems, partitions 1 and 2 on two chains of each, soft-core shift
    coefficient 5, against the alchemical oracle with force switching, system
    by system: forces, energies at lambda, the lambda2 sums (FEP) or the
    dU/dlambda sums of each partition (TI), each sum to 1e-5 of its own size."""
    torch = torch_cuda
    members = systems.batch_systems(n_systems=6, seed=3, n_min=3000, n_max=9000)
    c = ComputeNonbonded.from_systems([m.system for m in members], stream=torch.cuda.current_stream().cuda_stream)
    parts = []
    for m in members:
        pk = m.partition.copy()
        pk[m.system.chain_len:2 * m.system.chain_len] = 2
        parts.append(pk)
    alch = [L.alch_data(c.params, m.lam, min(1.0, m.lam + 0.05), elec_lambda_start=0.5, vdw_lambda_end=1.0,
                        vdw_shift_coeff=5.0) for m in members]
    c.set_vdw_force_switching(True)
    c.set_alch(np.concatenate(parts), alch, method=method)
    print(f"[alch force switch {method}] {len(members)} systems, {c.natoms} atoms, launching", flush=True)
    c.launch_work(do_energy=True, do_virial=True, do_slow=True, do_pairlist=True)
    per, _, f, fs = c.finish_batch()
    pos = c.effective_positions()
    boxes = c.system_boxes()
    starts = c.sys_atom_start
    for k, m in enumerate(members):
        a, b = starts[k], starts[k + 1]
        s = m.system
        p = s.params
        A, B = systems.lj_tables(s)
        beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
        o = orc.nb_fp64_alch(pos[a:b], s.charge * np.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr, s.excl_idx,
                             boxes[k], p.cutoff, p.switchdist, beta, parts[k], alch[k], vdw_force_switching=True)
        r = per[k]
        ef, es = rel_rms(f[a:b], o["f"]), rel_rms(fs[a:b], o["fs"])
        print(f"[{method} sys {k} n={s.natoms} lam={m.lam:.3f}] force fast {ef:.2e} slow {es:.2e}", flush=True)
        assert ef < TOL and es < TOL
        elec_scale = abs(o["e_elec"] + o["e_slow"])
        for key in ("e_vdw", "e_elec", "e_slow") + (TI_KEYS if method == "ti" else FEP_KEYS):
            scale = abs(o[key]) if ("vdw" in key or "ti" in key) else max(abs(o[key]), elec_scale)
            err = abs(r[key] - o[key]) / max(scale, 1e-30)
            print(f"    {key} gpu {r[key]:.10e} oracle {o[key]:.10e} rel {err:.2e}", flush=True)
            assert err < TOL, key
        for key in ("virial", "virial_slow"):
            assert np.linalg.norm(r[key] - o[key]) / np.linalg.norm(o[key]) < 2e-5, key
    c.close()
