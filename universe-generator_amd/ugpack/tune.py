"""Autotuning of the ensemble runners on the current GPU.

Two engines pack an ensemble: `ug_ensemble_run` keeps several universes in flight on worker
threads (one HIP stream each; a universe's pack is a chain of dependent launches with a host
round trip per sub-batch), and `ug_ensemble_run_batched` advances a window of universes in
lock-step on one stream, each stage one launch for all of them.  Which engine wins, how many
universes it keeps in flight, and how large a universe's sub-batches grow (the survivors a
sub-batch aims at, its largest size) depend on the device and on the mix of models, so they
are measured rather than fixed: every candidate setting packs the same sample of the ensemble
and the fastest wins.  Results never depend on these settings (each universe is the
deterministic pack of its config), only the time does.
"""

from __future__ import annotations

import time

from .abi import ensemble

DEFAULT_CHOICES = (1, 2, 4, 8, 16, 32)
# batched engine: (lanes, universes in flight per lane)
DEFAULT_BATCHED = ((4, 8), (4, 16), (8, 4), (8, 8), (16, 4), (16, 8))
DEFAULT_TARGETS = (512, 1024, 2048)      # survivors a sub-batch aims at
DEFAULT_SUB_TARGET = 0                    # the library's per-dimension default
# largest sub-batch: the tuning sample (2*10^6 candidates a universe) favours small caps that
# the 10^7-candidate workload does not (2^18: 1.84 s against 1.75 s for the ensemble)
DEFAULT_SUB_MAX = (1 << 20,)


def _time(configs, threads: int, repeat: int, batched: bool = False, window: int = 0, lanes: int = 1) -> float:
    best = float("inf")
    for _ in range(max(repeat, 1)):
        t0 = time.perf_counter()
        # returns once every universe's stream has finished
        ensemble(configs, threads=threads, batched=batched, window=window, lanes=lanes)
        best = min(best, time.perf_counter() - t0)
    return best


def ensemble_threads(configs, choices=DEFAULT_CHOICES, repeat: int = 1):
    """Time the worker-thread runner at every worker count n in `choices` (after one untimed
    pass that warms the buffer and table caches); return (fastest n, {n: seconds})."""
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
                  repeat: int = 1, batched=DEFAULT_BATCHED):
    """The engine and its universes in flight first (worker counts `choices` of the thread
    runner, (lanes, window) settings `batched` of the batched engine; default sizing), then the
    sub-batch sizing with that setting.  Returns ({"engine": "threads" | "batched", "threads",
    "lanes", "window", "sub_target", "sub_max"}, {setting: seconds})."""
    threads, times = ensemble_threads(configs, choices, repeat)
    sweep = {f"threads={n}": t for n, t in times.items()}
    best = {"engine": "threads", "threads": threads, "lanes": 0, "window": 0, "sub_target": DEFAULT_SUB_TARGET,
            "sub_max": 1 << 20}
    best_t = times[threads]
    if batched:
        ensemble(configs, batched=True, window=max(w for _, w in batched))  # warm the caches
    for lanes, w in batched:
        t = _time(configs, 0, repeat, batched=True, window=w, lanes=lanes)
        sweep[f"batched,lanes={lanes},window={w}"] = t
        if t < best_t:
            best, best_t = dict(best, engine="batched", lanes=lanes, window=w), t
    is_b = best["engine"] == "batched"
    key = f"batched,lanes={best['lanes']},window={best['window']}" if is_b else f"threads={best['threads']}"
    for tg in targets:
        for sm in sub_max:
            if (tg, sm) == (DEFAULT_SUB_TARGET, 1 << 20):
                continue
            t = _time(with_sizing(configs, tg, sm), best["threads"], repeat, batched=is_b, window=best["window"],
                      lanes=best["lanes"])
            sweep[f"{key},sub_target={tg},sub_max={sm}"] = t
            if t < best_t:
                best, best_t = dict(best, sub_target=tg, sub_max=sm), t
    return best, sweep
