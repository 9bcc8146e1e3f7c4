"""Table mode of the pair kernel (NAMD's 1/r force/energy tables with linear
interpolation) through the C ABI, against the oracle's table-mode pair
evaluation with the same tables (oracle_nb_table, pinned in
test_cpu_tables.py), and its distance from the analytic FP64 oracle."""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from mi355nb.compute import ComputeNonbonded
from oracle import oracle as orc

from nbcheck import rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5      # vs the same table evaluated in FP64
ENERGY_TOL = 1e-5
TABLE_VS_EXACT = 2e-3  # documented accuracy of 4096-entry linear tables vs the analytic forces


@pytest.mark.parametrize("name", ["small", "water9k"])
def test_table_mode_matches_table_oracle(torch_cuda, name):
    torch = torch_cuda
    s = systems.small_box() if name == "small" else systems.water_box_9k()
    p = s.params
    c = ComputeNonbonded.from_system(s, stream=torch.cuda.current_stream().cuda_stream)
    c.set_force_table(True)
    print(f"[table {name}] launching", flush=True)
    r = c.compute(do_energy=True, do_virial=True, do_slow=True, do_pairlist=True)
    pos = c.effective_positions()
    box = c.system_boxes()[0]
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = L.lj_mix(s.lj_sigma, s.lj_eps)
    q = s.charge * math.sqrt(332.0636)
    ft, et = L.force_energy_tables(4096, p.cutoff, p.switchdist, beta)
    o = orc.nb_table(pos, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, box, p.cutoff, p.switchdist, ft, et)
    x = orc.nb_fp64(pos, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, box, p.cutoff, p.switchdist, beta)
    ef, es = rel_rms(r["f"], o["f"]), rel_rms(r["fs"], o["fs"])
    xf, xs = rel_rms(r["f"], x["f"]), rel_rms(r["fs"], x["fs"])
    print(f"[table {name}] vs table oracle: fast {ef:.2e} slow {es:.2e}; vs analytic: fast {xf:.2e} slow {xs:.2e}")
    assert r["stat"]["numPairsCutoff"] == o["npairs"]
    assert ef < FORCE_TOL and es < FORCE_TOL
    if name == "small":
        # the 9,000-atom water box at its specified density has intermolecular
        # contacts below 1 A, inside the table's first entry (r = 4096/4095.5 A),
        # where the table (like NAMD's) holds the last entry's value
        assert xf < TABLE_VS_EXACT and xs < TABLE_VS_EXACT
    scale = abs(o["e_elec"] + o["e_slow"])
    for key in ("e_vdw", "e_elec", "e_slow"):
        sc = abs(o[key]) if key == "e_vdw" else scale
        err = abs(r[key] - o[key]) / sc
        print(f"    {key} gpu {r[key]:.10e} table-oracle {o[key]:.10e} rel {err:.2e}")
        assert err < ENERGY_TOL, key
    # switching back restores direct math
    c.set_force_table(False)
    d = c.compute(do_energy=False, do_virial=False, do_slow=True, do_pairlist=False)
    assert rel_rms(d["f"], x["f"]) < 1e-5
    c.close()
