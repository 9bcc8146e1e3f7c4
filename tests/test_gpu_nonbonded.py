"""GPU parity of the nonbonded path (through the C ABI) against the oracle.

Tolerances: forces (fast and slow separately) within 1e-5 relative RMS of
the FP64 oracle evaluated at the same effective positions; e_vdw within 1e-5
relative; e_elec and e_slow within 1e-5 of the real-space Coulomb energy
(e_elec + e_slow); the virial within 1e-5 of its norm.  Pair sets, pair counts
and exclusion counts are bit-exact against NAMD's CPU membership rule
(r^2 <= rc^2 in FP64, exact on the effective positions).  Against the
restatement of NAMD's CPU spline-table path the bar is the same 1e-5 for
forces and energies; that path's own table error versus exact physics
(about 5e-6 in fast forces and 1.1e-5 in the slow virial on the 96k box) is
stated in the test that measures it.
"""
import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from mi355nb.compute import decode_pairs
from nbcheck import gpu_compute, oracle_for, pair_set_checksums, rel_rms
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _small():
    return systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7)


def _check_against_oracle(s, c, r, tol=TOL):
    pos = c.effective_positions()
    box = c.info()["box"]
    o = oracle_for(s, pos, box)
    _compare(s.name + "/fp64", r, o, tol, tol)
    return o


def _compare(tag, r, o, tol, vir_tol):
    ef = rel_rms(r["f"], o["f"])
    es = rel_rms(r["fs"], o["fs"])
    print(f"[{tag}] force rel-RMS fast {ef:.2e} slow {es:.2e}")
    assert ef < tol and es < tol
    # The C1 split makes e_elec and e_slow a near-cancelling pair in dense neutral
    # boxes (9k water: e_slow = +83 from terms of magnitude ~1e2 each over 1e7
    # pairs), so each electrostatic part is held to 1e-5 of the real-space
    # Coulomb energy e_elec + e_slow (the physical quantity); e_vdw to itself.
    coul = abs(o["e_elec"] + o["e_slow"])
    for k in ("e_vdw", "e_elec", "e_slow"):
        scale = abs(o[k]) if k == "e_vdw" else max(abs(o[k]), coul)
        err = abs(r[k] - o[k]) / max(scale, 1e-12)
        print(f"[{tag}] {k} gpu {r[k]:.10e} oracle {o[k]:.10e} rel {err:.2e}")
        assert err < tol, k
    for k in ("virial", "virial_slow"):
        err = np.linalg.norm(r[k] - o[k]) / np.linalg.norm(o[k])
        print(f"[{tag}] {k} rel {err:.2e}")
        assert err < vir_tol, k


def _check_pairs(s, c, r, want_pairs=False):
    pos = c.effective_positions()
    box = np.asarray(c.info()["box"])
    p = s.params
    o = orc.pairs_exact(pos, s.excl_ptr, s.excl_idx, box, p.cutoff, p.pairlistdist, want_pairs=want_pairs)
    st = r["stat"]
    got = (st["numPairsCutoff"], st["numPairsList"], st["numExcluded"])
    want = (o["n_cut"], o["n_list"], o["n_excl_list"])
    print(f"[{s.name}] pairs (cutoff, list, excluded) gpu {got} oracle {want}")
    assert got == want
    i, j = decode_pairs(c.tile_lists())
    assert i.size == o["n_list"]
    assert pair_set_checksums(i, j, s.natoms) == (o["sum_list"], o["xor_list"])
    if want_pairs:
        assert set(zip(i.tolist(), j.tolist())) == set(map(tuple, o["pairs"].tolist()))
    return o


def test_small_box_forces_energies_virial(torch_cuda):
    s = _small()
    c, r = gpu_compute(s)
    o = _check_against_oracle(s, c, r)
    assert r["stat"]["numPairsCutoff"] == o["npairs"]
    c.close()


def test_small_box_pair_sets_bit_exact(torch_cuda):
    s = _small()
    c, r = gpu_compute(s)
    _check_pairs(s, c, r, want_pairs=True)
    c.close()


def test_small_box_vs_namd_cpu_path(torch_cuda):
    s = _small()
    c, r = gpu_compute(s)
    pos = c.effective_positions()
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    o = orc.namd_cpu(pos, s.charge * np.sqrt(332.0636), s.vdw_type, A, B, s.excl_ptr, s.excl_idx,
                     np.asarray(c.info()["box"]), p.cutoff, p.switchdist, beta)
    assert o["npairs"] == r["stat"]["numPairsCutoff"]
    _compare(s.name + "/namd-cpu", r, o, TOL, TOL)
    c.close()


def test_pairlist_replay_matches_pairlist_step(torch_cuda):
    s = _small()
    c, r0 = gpu_compute(s)
    c.launch_work(True, True, True, do_pairlist=False)
    r1 = c.finish()
    assert rel_rms(r1["f"], r0["f"]) < 1e-6
    assert rel_rms(r1["fs"], r0["fs"]) < 1e-6
    assert abs(r1["e_elec"] - r0["e_elec"]) <= 1e-6 * abs(r0["e_elec"])
    c.close()


def test_flag_combinations(torch_cuda):
    s = _small()
    c, ref = gpu_compute(s)
    for e in (0, 1):
        for v in (0, 1):
            for sl in (0, 1):
                c.launch_work(e, v, sl, do_pairlist=False)
                r = c.finish()
                assert rel_rms(r["f"], ref["f"]) < 1e-6, (e, v, sl)
                if sl:
                    assert rel_rms(r["fs"], ref["fs"]) < 1e-6
                if e:
                    assert abs(r["e_vdw"] - ref["e_vdw"]) <= 1e-6 * abs(ref["e_vdw"])
                if v:
                    assert np.linalg.norm(r["virial"] - ref["virial"]) <= 1e-6 * np.linalg.norm(ref["virial"])
    c.close()


def test_momentum_conservation(torch_cuda):
    s = _small()
    c, r = gpu_compute(s)
    scale = np.abs(r["f"]).sum()
    assert np.abs(r["f"].sum(0)).max() < 1e-5 * scale
    assert np.abs(r["fs"].sum(0)).max() < 1e-5 * np.abs(r["fs"]).sum()
    c.close()


def test_water_box_9k(torch_cuda):
    """3x liquid density (contacts below 0.2 A): forces vs FP64, pair sets bit-exact."""
    s = systems.water_box_9k()
    c, r = gpu_compute(s)
    _check_against_oracle(s, c, r)
    _check_pairs(s, c, r)
    c.close()


def test_update_positions_then_rebuild(torch_cuda):
    s = _small()
    c, r0 = gpu_compute(s)
    rng = np.random.default_rng(3)
    pos2 = s.pos + rng.uniform(-0.2, 0.2, s.pos.shape)
    c.update_positions(pos2)
    c.launch_work(True, True, True, do_pairlist=True)
    r1 = c.finish()
    s2 = systems.System(**{**s.__dict__, "pos": pos2})
    _check_against_oracle(s2, c, r1)
    _check_pairs(s2, c, r1)
    c.close()


def test_solvated_96k_full_size(torch_cuda):
    """The benchmark workload itself: bit-exact pair/exclusion counts and set
    checksums; forces, energies, virial vs the FP64 oracle and vs the NAMD
    CPU-path restatement."""
    s = systems.solvated_96k()
    c, r = gpu_compute(s)
    _check_against_oracle(s, c, r)
    _check_pairs(s, c, r)
    pos = c.effective_positions()
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    o = orc.namd_cpu(pos, s.charge * np.sqrt(332.0636), s.vdw_type, A, B, s.excl_ptr, s.excl_idx,
                     np.asarray(c.info()["box"]), p.cutoff, p.switchdist, beta)
    # NAMD's slow-virial table error vs exact physics is 1.1e-5 on this box
    _compare(s.name + "/namd-cpu", r, o, TOL, 2e-5)
    c.close()
