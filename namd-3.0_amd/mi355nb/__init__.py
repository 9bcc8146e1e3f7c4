"""mi355nb: MI355X-native NAMD nonbonded + PME real-space hot path.

The product is libmi355nb.so (C ABI in include/mi355nb.h, HIP kernels for
gfx950).  This package holds the Python driver of that ABI and the seeded
synthetic workloads used by the tests and the benchmark.
"""
from . import systems  # noqa: F401

__all__ = ["systems", "lib", "compute"]
