"""GPU parity in a general periodic cell (assign_patches_lattice; NAMD's
lata / latb / latc) against the brute-force general-cell oracle
(oracle_nb_fp64_cell, pinned in test_cpu_cell_oracle.py) at the same
effective positions and rounded lattice: forces within 1e-5 relative RMS,
energies and virial within 1e-5 (the orthorhombic bar), pairs within the
cutoff counted exactly (the exact inclusive rule on exact displacements), and
the FP64 exclusion correction with the general minimum image within 1e-9."""
import math

import numpy as np
import pytest

from mi355nb import lib as L
from mi355nb import systems
from mi355nb.compute import ComputeNonbonded
from nbcheck import COULOMB, rel_rms
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-5
# perpendicular widths 32.3, 32.3, 34.0 A (at least 2 x pairlistdist 13.5)
CELL = np.array([[34.0, 0.0, 0.0], [8.0, 33.0, 0.0], [-6.0, 7.0, 34.0]])


@pytest.mark.parametrize("cell", ["skewed", "orthorhombic-as-lattice"])
def test_triclinic_matches_oracle(torch_cuda, cell):
    torch = torch_cuda
    base = systems.small_box(n_waters=1000, n_chains=3, chain_len=20, box=34.0, seed=3)
    s = systems.sheared(base, CELL if cell == "skewed" else np.diag(base.box))
    c = ComputeNonbonded.from_system(s, stream=torch.cuda.current_stream().cuda_stream)
    print(f"[triclinic {cell}] {s.natoms} atoms", flush=True)
    r = c.compute(True, True, True, do_pairlist=True)
    pos = c.effective_positions()
    lat = c.system_cells()[0]
    p = s.params
    A, B = systems.lj_tables(s)
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    o = orc.nb_fp64_cell(pos, s.charge * math.sqrt(COULOMB), s.vdw_type, A, B, s.excl_ptr, s.excl_idx, lat,
                         p.cutoff, p.switchdist, beta)
    print(f"    pairs within the cutoff gpu {r['stat']['numPairsCutoff']} oracle {o['npairs']}", flush=True)
    assert r["stat"]["numPairsCutoff"] == o["npairs"]
    ef, es = rel_rms(r["f"], o["f"]), rel_rms(r["fs"], o["fs"])
    print(f"    force rel-RMS fast {ef:.2e} slow {es:.2e}", flush=True)
    assert ef < TOL and es < TOL
    coul = abs(o["e_elec"] + o["e_slow"])
    for k in ("e_vdw", "e_elec", "e_slow"):
        scale = abs(o[k]) if k == "e_vdw" else max(abs(o[k]), coul)
        err = abs(r[k] - o[k]) / scale
        print(f"    {k} gpu {r[k]:.10e} oracle {o[k]:.10e} rel {err:.2e}", flush=True)
        assert err < TOL, k
    for k in ("virial", "virial_slow"):
        assert np.linalg.norm(r[k] - o[k]) / np.linalg.norm(o[k]) < TOL, k
    # FP64 exclusion correction with the general minimum image
    c.launch_work(do_energy=False, do_virial=False, do_slow=True, do_pairlist=False)
    ev = c.exclusion_correction().cpu().numpy()[0]
    q32 = (s.charge * math.sqrt(COULOMB)).astype(np.float32).astype(np.float64)
    inv = np.linalg.inv(lat)
    E = 0.0
    for i, j in s.exclusion_pairs():
        d = pos[j] - pos[i]
        d = d - np.rint(d @ inv) @ lat
        rr = np.linalg.norm(d)
        if rr < p.cutoff:
            E -= q32[i] * q32[j] * math.erf(beta * rr) / rr
    print(f"    exclusion correction gpu {ev[0]:.12e} numpy {E:.12e}", flush=True)
    assert abs(ev[0] - E) <= 1e-9 * abs(E)
    c.close()


def test_lattice_argument_errors(torch_cuda):
    base = systems.small_box(n_waters=300, n_chains=1, chain_len=20, box=28.0, seed=3)
    c = ComputeNonbonded.from_system(base)
    with pytest.raises(L.Mi355nbError):   # left-handed
        c.assign_patches_lattice(np.array([[28.0, 0, 0], [0, 0, 28.0], [0, 28.0, 0]]), base.pos)
    # a perpendicular width (13.7 A) below 2 x pairlistdist (this box: 2 x 9 A)
    assert 2 * base.params.pairlistdist == 18.0
    with pytest.raises(L.Mi355nbError):
        c.assign_patches_lattice(np.array([[28.0, 0, 0], [20.0, 28.0, 0], [0, 0, 28.0]]) * 0.6, base.pos)
    c.close()
