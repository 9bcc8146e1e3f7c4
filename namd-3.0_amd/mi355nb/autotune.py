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


STEP_CANDIDATES = (0, 4, 6, 8, 12, 16, 32)


def tune_step(compute, step, sync, clock, candidates=STEP_CANDIDATES, warmup=3, reps=10):
    """Launch shape chosen on the whole step: step(k) enqueues step k (the pair
    kernel runs on non-pairlist steps k = 1 .. reps), sync() waits for the
    device, clock() is a host timer.  With other work running concurrently
    (PME on its own stream) a smaller persistent grid can leave room for it,
    so the pair kernel alone does not decide.  Returns (best, {candidate: ms
    per step}) and leaves the best one set."""
    results = {}
    for wg in candidates:
        compute.set_launch_config(wg)
        for k in range(1, warmup + 1):
            step(k)
        sync()
        t0 = clock()
        for k in range(1, reps + 1):
            step(k)
        sync()
        results[wg] = (clock() - t0) * 1e3 / reps
    best = min(results, key=results.get)
    compute.set_launch_config(best)
    return best, results
