"""Autotuning of the ensemble runner on the current GPU.

A universe's pack is a chain of dependent launches with a host round trip per sub-batch, so
`ug_ensemble_run` keeps several universes in flight, one worker thread and HIP stream each.
How many pay off, and how large a universe's sub-batches should grow (the survivors a
sub-batch aims at, its largest size), depend on the device and on the mix of models (small T
universes are launch-latency bound, 3+1 T=128 ones fill the GPU by themselves), so they are
measured rather than fixed: every candidate setting packs the same sample of the ensemble
and the fastest wins.  Results never depend on these settings (each universe is the
deterministic pack of its config), only the time does.
"""

from __future__ import annotations

import time

from .abi import ensemble

DEFAULT_CHOICES = (1, 2, 4, 8, 16, 32)
DEFAULT_TARGETS = (512, 1024, 2048)      # survivors a sub-batch aims at
DEFAULT_SUB_MAX = (1 << 18, 1 << 20)     # largest sub-batch


def _time(configs, threads: int, repeat: int) -> float:
    best = float("inf")
    for _ in range(max(repeat, 1)):
        t0 = time.perf_counter()
        ensemble(configs, threads=threads)  # returns once every universe's stream has finished
        best = min(best, time.perf_counter() - t0)
    return best


def ensemble_threads(configs, choices=DEFAULT_CHOICES, repeat: int = 1):
    """Time `ensemble(configs, threads=n)` for every n in `choices` (after one untimed pass
    that warms the buffer and table caches); return (fastest n, {n: seconds})."""
    if not configs:
        raise ValueError("ensemble_threads: empty sample")
    ensemble(configs, threads=max(choices))
    times = {n: _time(configs, n, repeat) for n in choices}
    return min(times, key=times.get), times


def with_sizing(configs, sub_target: int, sub_max: int):
    """Copies of the configs with the given sub-batch sizing."""
    out = []
    for c in configs:
        c2 = type(c).from_buffer_copy(c)
        c2.sub_target, c2.sub_max = sub_target, sub_max
        out.append(c2)
    return out


def ensemble_plan(configs, choices=DEFAULT_CHOICES, targets=DEFAULT_TARGETS, sub_max=DEFAULT_SUB_MAX,
                  repeat: int = 1):
    """The worker count first (default sizing), then the sub-batch sizing at that count:
    returns ({"threads", "sub_target", "sub_max"}, {setting: seconds})."""
    threads, times = ensemble_threads(configs, choices, repeat)
    sweep = {f"threads={n}": t for n, t in times.items()}
    best = {"threads": threads, "sub_target": 1024, "sub_max": 1 << 20}
    best_t = times[threads]
    for tg in targets:
        for sm in sub_max:
            if (tg, sm) == (1024, 1 << 20):
                continue
            t = _time(with_sizing(configs, tg, sm), threads, repeat)
            sweep[f"threads={threads},sub_target={tg},sub_max={sm}"] = t
            if t < best_t:
                best, best_t = {"threads": threads, "sub_target": tg, "sub_max": sm}, t
    return best, sweep
