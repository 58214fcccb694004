"""Autotuning of the ensemble runner on the current GPU.

A universe's pack is a chain of dependent launches with a host round trip per sub-batch, so
`ug_ensemble_run` keeps several universes in flight, one worker thread and HIP stream each.
How many pay off depends on the device and on the mix of models (small T universes are
launch-latency bound, 3+1 T=128 ones fill the GPU by themselves), so the count is measured
rather than fixed: every candidate count packs the same sample of the ensemble and the
fastest wins.  Results never depend on the count (each universe is the deterministic pack of
its config), only the time does.
"""

from __future__ import annotations

import time

from .abi import ensemble

DEFAULT_CHOICES = (4, 8, 16, 24, 32)


def ensemble_threads(configs, choices=DEFAULT_CHOICES, repeat: int = 1):
    """Time `ensemble(configs, threads=n)` for every n in `choices` (after one untimed pass
    that warms the buffer and table caches); return (fastest n, {n: seconds})."""
    if not configs:
        raise ValueError("ensemble_threads: empty sample")
    ensemble(configs, threads=max(choices))
    times = {}
    for n in choices:
        best = float("inf")
        for _ in range(max(repeat, 1)):
            t0 = time.perf_counter()
            ensemble(configs, threads=n)  # returns once every universe's stream has finished
            best = min(best, time.perf_counter() - t0)
        times[n] = best
    return min(times, key=times.get), times
