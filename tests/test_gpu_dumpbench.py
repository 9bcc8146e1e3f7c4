"""The GPU path on a system read from NAMD's dumpbench format: forces,
energies and virial within 1e-5 of the FP64 oracle on the same parsed data,
and the pair sets bit-exact, so a dumped system runs through the path as
the build's own systems do."""
import io

import pytest

from mi355nb import dumpbench as D
from mi355nb import systems
from nbcheck import gpu_compute

import test_gpu_nonbonded as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["small", "water9k"])
def test_gpu_on_dumped_system(torch_cuda, which):
    s0 = {"small": lambda: systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7),
          "water9k": systems.water_box_9k}[which]()
    buf = io.StringIO()
    D.write(buf, s0)
    s, _ = D.read(buf.getvalue(), name=f"dump-{which}", pme_grid=s0.params.pme_grid)
    print(f"[dumpbench {which}] {s.natoms} atoms", flush=True)
    c, r = gpu_compute(s)
    T._check_against_oracle(s, c, r)
    T._check_pairs(s, c, r)
    c.close()


def test_gpu_modified_exclusions_from_dumped_file(torch_cuda):
    """A dumped system with scaled 1-4 exclusions (exclude scaled1-4,
    scale14 0.8333, separate 1-4 LJ columns): the lists and tables the reader
    returns drive the modified-exclusion correction, whose energies match the
    oracle's on the same parsed data to 1e-9 and whose forces, added to the
    pair kernel's, match the oracles' sum to 1e-5."""
    import math

    import numpy as np

    from mi355nb.compute import ComputeNonbonded
    from nbcheck import rel_rms
    from oracle import oracle as orc

    torch = torch_cuda
    s0 = systems.small_box(n_waters=300, n_chains=3, chain_len=20, box=22.0, seed=7)
    mp, mi = systems.chain_pairs_14(s0)
    A14, B14 = systems.lj14_tables(s0)
    buf = io.StringIO()
    D.write(buf, s0, modified=dict(mod_ptr=mp, mod_idx=mi, scale14=0.8333, A14=A14, B14=B14))
    s, info = D.read(buf.getvalue(), name="dump-mod", pme_grid=s0.params.pme_grid)
    print(f"[dumpbench modified] {s.natoms} atoms, {info['modified_pairs']} modified pairs", flush=True)
    c = ComputeNonbonded.from_system(s, stream=torch.cuda.current_stream().cuda_stream)
    c.set_modified_exclusions(info["mod_ptr"], info["mod_idx"], info["scale14"], info["A14"], info["B14"])
    c.launch_work(do_energy=False, do_virial=False, do_slow=True, do_pairlist=True)
    r0 = c.finish_batch()
    ev = c.modified_exclusion_correction()
    r1 = c.finish_batch()
    torch.cuda.synchronize()
    ev = ev.cpu().numpy()[0]
    pos, box = c.effective_positions(), c.system_boxes()[0]
    p = s.params
    A, B = systems.lj_tables(s)
    q32 = (s.charge * math.sqrt(332.0636)).astype(np.float32).astype(np.float64)
    m = orc.mod_excl_fp64(pos, q32, s.vdw_type, A, B, info["A14"], info["B14"], info["mod_ptr"], info["mod_idx"],
                          box, p.cutoff, p.switchdist, info["scale14"])
    assert m["npairs"] == info["modified_pairs"]
    for k, key in enumerate(("e_vdw", "e_elec", "e_slow")):
        assert abs(ev[k] - m[key]) <= 1e-9 * abs(m[key]), key
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    o = orc.nb_fp64(pos, s.charge * math.sqrt(332.0636), s.vdw_type, A, B, s.excl_ptr, s.excl_idx, box, p.cutoff,
                    p.switchdist, beta)
    ef, es = rel_rms(r1[2], o["f"] + m["f"]), rel_rms(r1[3], o["fs"] + m["fs"])
    print(f"    totals fast {ef:.2e} slow {es:.2e}", flush=True)
    assert ef < 1e-5 and es < 1e-5
    c.close()
