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
