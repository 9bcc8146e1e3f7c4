"""Launch-shape autotuner for the pair kernel.

Times the force-only pair kernel (HIP events on the compute's stream) for
each candidate number of persistent workgroups per CU (0 = one wave per tile
list, no persistence) and keeps the fastest.
The compute object must have built its pairlists already.
"""
from __future__ import annotations

DEFAULT_CANDIDATES = (0, 4, 8, 16, 32)


def tune_pair_kernel(compute, candidates=DEFAULT_CANDIDATES, reps=8, do_slow=True):
    """Returns (best_workgroups_per_cu, {candidate: ms_per_launch}) and leaves the best one set."""
    results = {}
    compute.set_timing(True)
    try:
        for wg in candidates:
            compute.set_launch_config(wg)
            compute.launch_work(do_energy=False, do_virial=False, do_slow=do_slow, do_pairlist=False)
            compute.kernel_time()  # discard the first launch of each shape
            for _ in range(reps):
                compute.launch_work(do_energy=False, do_virial=False, do_slow=do_slow, do_pairlist=False)
            _, _, ms, n = compute.kernel_time()
            results[wg] = ms / max(n, 1)
    finally:
        compute.set_timing(False)
    best = min(results, key=results.get)
    compute.set_launch_config(best)
    return best, results
