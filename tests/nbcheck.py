"""Shared helpers for the nonbonded parity tests: run the HIP path through the
C ABI and the oracle on the same effective positions."""
from __future__ import annotations

import numpy as np

from mi355nb import lib as L
from mi355nb import systems
from mi355nb.compute import ComputeNonbonded, decode_pairs
from oracle import oracle as orc

COULOMB = 332.0636


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def oracle_for(s: systems.System, pos_eff, box, nthreads=0):
    p = s.params
    beta = systems.ewald_coefficient(p.cutoff, p.ewald_tol)
    A, B = systems.lj_tables(s)
    q = s.charge * np.sqrt(COULOMB)
    return orc.nb_fp64(pos_eff, q, s.vdw_type, A, B, s.excl_ptr, s.excl_idx, np.asarray(box), p.cutoff,
                       p.switchdist, beta, nthreads=nthreads)


def gpu_compute(s: systems.System, do_energy=True, do_virial=True, do_slow=True):
    c = ComputeNonbonded.from_system(s)
    r = c.compute(do_energy, do_virial, do_slow, do_pairlist=True)
    return c, r


def pair_set_checksums(i, j, n):
    h = orc.pair_hash(i, j, n)
    with np.errstate(over="ignore"):
        s = int(np.sum(h, dtype=np.uint64))
    x = int(np.bitwise_xor.reduce(h)) if h.size else 0
    return s, x
