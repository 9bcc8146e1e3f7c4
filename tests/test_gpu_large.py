"""Full-size parity on the 1.07M-atom box (340,000 TIP3P + 150 chains x 300
beads, 216 A cube): the exact pair count, forces, energies and virial of one
pairlist + energy evaluation against the FP64 oracle, and the slot-order PME
spread against the oracle's grid."""
import math

import numpy as np
import pytest

from mi355nb import compute as mc
from mi355nb import lib as L
from mi355nb import systems
from oracle import oracle as orc

from nbcheck import rel_rms

pytestmark = pytest.mark.gpu


def test_solvated_1m_matches_fp64_oracle(torch_cuda):
    torch = torch_cuda
    s = systems.solvated_1m()
    p = s.params
    print(f"[1m] {s.natoms} atoms, building compute", flush=True)
    c = mc.ComputeNonbonded.from_system(s, stream=torch.cuda.current_stream().cuda_stream)
    r = c.compute(do_energy=True, do_virial=True, do_slow=True, do_pairlist=True)
    pos = c.effective_positions()
    box = c.system_boxes()[0]
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    q = s.charge * math.sqrt(332.0636)
    print("[1m] oracle", flush=True)
    o = orc.nb_fp64(pos, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, box, p.cutoff, p.switchdist, beta)
    ef, es = rel_rms(r["f"], o["f"]), rel_rms(r["fs"], o["fs"])
    print(f"[1m] pairs gpu {r['stat']['numPairsCutoff']} oracle {o['npairs']}; force fast {ef:.2e} slow {es:.2e}")
    assert r["stat"]["numPairsCutoff"] == o["npairs"]
    assert ef < 1e-5 and es < 1e-5
    scale = abs(o["e_elec"] + o["e_slow"])
    for key in ("e_vdw", "e_elec", "e_slow"):
        sc = abs(o[key]) if key == "e_vdw" else scale
        err = abs(r[key] - o[key]) / sc
        print(f"    {key} gpu {r[key]:.10e} oracle {o[key]:.10e} rel {err:.2e}")
        assert err < 1e-5, key
    for key in ("virial", "virial_slow"):
        err = np.linalg.norm(r[key] - o[key]) / np.linalg.norm(o[key])
        print(f"    {key} rel {err:.2e}")
        assert err < 2e-5, key
    # PME spread on the 216^3 grid from the slot-ordered atoms
    K = p.pme_grid
    recip = np.diag(1.0 / box)
    at = c.pme_atoms(recip)
    grid = torch.zeros(K[0] * K[1] * K[2], dtype=torch.float32, device="cuda")
    mc.spread_charge(at, grid, K, p.pme_order)
    torch.cuda.synchronize()
    g = grid.cpu().numpy().reshape(K[2], K[1], K[0])
    og = orc.pme_spread_gpu(at.cpu().numpy(), K, p.pme_order)
    eg = float(np.sqrt(((g - og) ** 2).sum() / (og ** 2).sum()))
    print(f"[1m] PME grid rel {eg:.2e}")
    assert eg < 1e-5
    c.close()
