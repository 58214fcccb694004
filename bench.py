"""Benchmark: candidates tested/s and seconds to jam, 2+1 F=64 (one MI355X).

One step = one complete pack of a 2+1, maxFreq F = T = 64 universe to jamming: Philox
candidates in batches of 2^20, stopping at the first 2^20 boundary where the accepts
among the last 2^22 candidates are fewer than 1e-4 * 2^22.  Step k packs seed k+1
(warmup packs use seeds 1001...).  value = candidates consumed by the timed packs /
their summed step times (a step: handle creation, the run and a device synchronisation);
ms_per_step = mean time to jam.

    python bench.py [--gpus 1] [--steps 8] [--warmup 2] [--precision split|fp64|both]
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "universe-generator_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0


WORKLOADS = {
    # headline: 2+1, F = T = 64, packed to jamming
    "cfg2": dict(dim=2, T=64, max_attempts=0.0,
                 text="2+1 pack to jamming, F=T=64, batches of 2^20, stop when accepts in the "
                      "last 2^22 candidates < 1e-4*2^22"),
    # 3+1, F = T = 128, packed to jamming
    "cfg3": dict(dim=3, T=128, max_attempts=0.0,
                 text="3+1 pack to jamming, F=T=128, batches of 2^20, stop when accepts in the "
                      "last 2^22 candidates < 1e-4*2^22"),
    # 512 independent universes: T uniform in {16, 32, 64, 128}, dim 2 or 3 (p = 1/2), seeds
    # 1..512, each packed to 10^7 candidates (engine and universes in flight tuned on the GPU)
    "cfg4": dict(dim=0, T=0, max_attempts=1e7, universes=512,
                 text="ensemble of 512 universes, T uniform in {16,32,64,128}, dim 2|3 with "
                      "p=1/2, seeds 1..512, each packed to 10^7 candidates"),
    # 3+1, F = T = 256, packed to jamming, then 512 equal-time frames with 2048^2 rasters
    "cfg5": dict(dim=3, T=256, max_attempts=0.0, frames=512, raster=2048,
                 text="3+1 pack to jamming, F=T=256 (same stop rule), then 512 frames at "
                      "z_j=(j+1/2)pi/512: per-frame positions of every worldline and 2048^2 "
                      "float32 density rasters of the (x,y) torus projection"),
    # knot regime (the reference's braid_knot tool): 2+1, F = T = 32, sin1 pinned, full band;
    # one step = 2^30 primary-species candidates tested against the resident strands, no grid
    "knot": dict(dim=2, T=32, max_attempts=float(1 << 30),
                 text="knot regime: 2+1, F=T=32, sin1 pinned, full band, 2^30 candidates against "
                      "the resident strands (brute force, no grid)"),
    # the reference's published light-cone bake: 480 observation instants z in [0.05, 0.95] pi
    # of a 19,413-path 2+1 universe (full spectrum T = 100, uniques then subpaths), one
    # comoving observer (path 0), fit front, square metric
    "lightcone": dict(dim=2, T=100, paths=19413, frames=480,
                      text="light-cone bake: 480 instants z in [0.05pi, 0.95pi] of a 19,413-path 2+1 "
                           "F=T=100 universe (uniques then subpaths), observer path 0, fit front, "
                           "square metric"),
}


SPARSE = False  # --sparse: the packs of configs 2, 3 and 5 use the sparse occupied-cell grid


def pack_once(ugpack, seed: int, precision: int, profile: int = 0, workload: str = "cfg2"):
    """One pack of the workload; `profile` = the kernels timed with HIP events (PROFILE_*
    bits; counters run whenever it is nonzero)."""
    w = WORKLOADS[workload]
    p = ugpack.Packer(w["dim"], w["T"], seed=seed, precision=precision,
                      stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4,
                      stop_batch=1 << 20, stop_window=1 << 22, max_attempts=w["max_attempts"],
                      sparse=SPARSE)
    if profile:
        p.profile(True, kernels=profile)
    st = p.run()
    return p, st


def measured_traffic(workload: str, split: bool):
    """HBM bytes per prefilter launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of the same workload (None when no summary exists)."""
    f = os.path.join(ROOT, "profiles", f"{workload}_prefilter_hbm_traffic.md")
    if not split or not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def measured_valu_busy(workload: str, split: bool):
    """VALU-busy percentage of the prefilter from the committed rocprofv3 PMC summary of the
    same workload (None when no summary exists)."""
    f = os.path.join(ROOT, "profiles", f"{workload}_prefilter_valu.md")
    if not split or not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("valu_busy_pct:"):
            return float(line.split(":")[1])
    return None


def mfma_roofline(stats: dict, dim: int, T: int) -> dict:
    """Algorithmic flops of the prefilter's evaluation per launch over its average duration:
    2*d*K*32 per candidate tested (K = 2(F-1) coefficient rows, the 32 timesteps of the first
    column block).  The split's three products, the padding column of each frequency pair and
    the later column blocks (T > 160) are not counted."""
    launches = max(stats["launches"], 1)
    flops = stats["items"] * 2 * dim * 2 * (T - 1) * min(T, 32)
    avg_s = stats["ms"] / launches * 1e-3
    achieved = (flops / launches) / avg_s / 1e12 if avg_s > 0 else 0.0
    return {"bound": "mfma", "achieved": round(achieved, 3), "peak": MFMA_BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 6),
            "flops_per_launch": round(flops / launches, 1)}


def sum_stats(stats) -> dict:
    out = {"launches": 0, "items": 0, "ms": 0.0, "visits": 0, "points": 0}
    for st in stats:
        for k in out:
            out[k] += st[k]
    return out


def roofline_of(stats: dict, dim: int, T: int, precision_split: bool, traffic=None) -> dict:
    """Algorithmic bytes of the candidate-test kernel per launch over its average duration.

    Per visited (candidate, timestep): 3^(d-1) rows of 3 int32 cell heads (12 B each);
    per grid point compared: its coordinates and chain link (d*8 + 4 B in the FP64 form,
    16 B FP32 record in the split-BF16 form)."""
    launches = max(stats["launches"], 1)
    rows = 3 ** (dim - 1)
    per_point = 16 if precision_split else 8 * dim + 4
    bytes_total = stats["visits"] * rows * 12 + stats["points"] * per_point
    avg_ms = stats["ms"] / launches
    achieved = (bytes_total / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {
        "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
        "kernel": "k_fast_prefilter" if precision_split else "k_cand64<test>",
        "avg_launch_ms": round(avg_ms, 5), "launches": stats["launches"],
        "algorithmic_bytes_per_launch": round(bytes_total / launches, 1),
        "visits_per_test": round(stats["visits"] / max(stats["items"], 1), 4),
    }


def split_error(ugpack, torch, workload: str) -> dict:
    """Max |X_split-bf16 - X_fp64| over 65,536 candidates, on the timesteps the prefilter
    examines and on all timesteps, beside the certainty margin in use."""
    w = WORKLOADS[workload]
    idx = torch.arange(0, 1 << 16, dtype=torch.int64) * 104729 + 7
    with ugpack.Packer(w["dim"], w["T"], seed=1, precision=ugpack.PREC_SPLIT_BF16) as p:
        margin = p.status().margin
        x16 = p.eval_candidates(idx, ugpack.PREC_SPLIT_BF16)
        x64 = p.eval_candidates(idx, ugpack.PREC_FP64)
    d = (x16 - x64).abs()
    d = torch.minimum(d, 2.0 - d)
    T = w["T"]
    order = [i // 2 if i % 2 == 0 else T - 1 - i // 2 for i in range(T)][: min(T, 32)]
    return {"max_err_examined": float(d[:, order, :].max()), "max_err_all_t": float(d.max()),
            "margin_max": margin, "candidates": int(idx.numel())}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bake_frames(ugpack, torch, packer, nframes: int, px: int) -> dict:
    """Per-frame positions and density rasters of every accepted worldline (64 frames per
    call), timed from the device-built worldline set to the last raster."""
    import math

    import numpy as np
    from ugpack import frames

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paths = frames.Paths.from_packer(packer)
    z = (np.arange(nframes) + 0.5) * math.pi / nframes
    mass = 0.0
    for b in range(0, nframes, 64):
        pos, ras = frames.slices(paths, z[b:b + 64], raster_px=px)
        mass += float(ras.sum())
        del pos, ras
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    paths.close()
    return {"frames": nframes, "raster_px": px, "paths": paths.n, "seconds": round(dt, 4),
            "ms_per_frame": round(1e3 * dt / nframes, 4),
            "raster_mass_ok": abs(mass - nframes * paths.n) < 0.5}


def progress(msg: str) -> None:
    """A step-done line on stderr (the JSON result line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensemble_configs(ugpack, n: int, cands: float, precision: int):
    """The ensemble's universes: T and dim drawn from a fixed-seed generator."""
    import numpy as np

    rng = np.random.default_rng(20260709)
    Ts = rng.choice([16, 32, 64, 128], size=n)
    ds = rng.choice([2, 3], size=n)
    return [ugpack.pack_config(int(d), int(T), seed=i + 1, max_attempts=cands, precision=precision)
            for i, (d, T) in enumerate(zip(ds, Ts))]


def run_ensemble(args, ugpack, torch, forms, prec_main) -> int:
    """One step = the whole ensemble; value = all universes' candidates / wall time."""
    from ugpack import tune

    w = WORKLOADS["cfg4"]
    n, cands, threads = w["universes"], w["max_attempts"], args.threads
    sweep = None
    plan = {"engine": args.engine, "threads": threads, "lanes": args.lanes, "window": args.window,
            "sub_target": tune.DEFAULT_SUB_TARGET, "sub_max": 1 << 20}
    if args.engine == "tuned":
        # the engine, its universes in flight and the sub-batch sizing are tuned on this GPU:
        # a 128-universe sample of the same mix (seeds 1..128) at 2*10^6 candidates each,
        # outside the timed region
        plan, sweep = tune.ensemble_plan(ensemble_configs(ugpack, 128, 2e6, prec_main))
        progress(f"ensemble tuned: {plan}")
    threads = plan["threads"] if plan["threads"] > 0 else 4
    batched = plan["engine"] == "batched"
    run_kw = dict(threads=threads, batched=batched, window=plan["window"], lanes=max(plan["lanes"], 1))
    res = {}
    for prec in forms:
        cfgs = tune.with_sizing(ensemble_configs(ugpack, n, cands, prec), plan["sub_target"], plan["sub_max"])
        ugpack.ensemble(ensemble_configs(ugpack, 32, 1e6, prec), **run_kw)  # warmup
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = ugpack.ensemble(cfgs, **run_kw)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        res[prec] = dict(wall=wall, value=n * cands / wall, n=[s.n for s in st],
                         attempts=sum(s.attempts for s in st))
        progress(f"ensemble precision {prec}: {wall:.3f} s per step")
    # roofline of the dominant universe class (T=128, 3+1) from a profiled pack at 10^7
    p = ugpack.Packer(3, 128, seed=1, precision=prec_main, max_attempts=cands)
    p.profile(True)
    p.run()
    stats = p.kernel_stats(ugpack.abi.KERNEL_PREFILTER)
    p.close()
    main_r = res[prec_main]
    cfgs = ensemble_configs(ugpack, n, cands, prec_main)
    split = prec_main == ugpack.PREC_SPLIT_BF16
    line = {
        "metric": "candidates tested/s & s to jam, 2+1 F=64; split-BF16 vs FP64 GPU, same accepts",
        "value": round(main_r["value"], 1), "unit": "candidates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(1e3 * main_r["wall"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f32acc, f64 recheck" if split else "f64",
        "data": "synthetic (Philox candidate streams, seeds 1..512)",
        "config": {"workload": w["text"], "universes": n, "candidates_per_universe": cands,
                   "engine": plan["engine"], "threads": threads if not batched else None,
                   "lanes": plan["lanes"] if batched else None, "window": plan["window"] if batched else None,
                   "sub_target": plan["sub_target"], "sub_max": plan["sub_max"],
                   "tuning_sweep_s": {str(k): round(v, 4) for k, v in sweep.items()} if sweep else None,
                   "T_counts": {str(T): sum(1 for c in cfgs if c.T == T) for T in (16, 32, 64, 128)},
                   "dim3_count": sum(1 for c in cfgs if c.dim == 3),
                   "precision": "split-bf16" if split else "fp64"},
        "accepted_total": sum(main_r["n"]),
        "roofline": dict(roofline_of(stats, 3, 128, split), class_profiled="T=128 3+1 universe, 10^7 candidates"),
    }
    if split:
        line["roofline"]["mfma"] = mfma_roofline(stats, 3, 128)
    if ugpack.PREC_FP64 in res and split:
        f = res[ugpack.PREC_FP64]
        line["fp64_form"] = {"value": round(f["value"], 1), "ms_per_step": round(1e3 * f["wall"], 3),
                             "same_accepts_as_split": f["n"] == main_r["n"]}
        line["speedup_split_vs_fp64"] = round(main_r["value"] / f["value"], 3)
    line["cpu_baseline"] = None if args.no_cpu else cpu_sample(ugpack, torch, "cfg4")
    print(json.dumps(line))
    return 0


def _omp_threads() -> int:
    """Threads the C oracle's OpenMP loops use on this host."""
    try:
        return int(os.environ.get("OMP_NUM_THREADS", "")) or (os.cpu_count() or 1)
    except ValueError:
        return os.cpu_count() or 1


def run_knot(args, ugpack, torch) -> int:
    """One step = the knot workload's primary phase (2^30 candidates from an empty strand set,
    seed k+1); value = candidates / summed step time.  The CPU oracle runs the first 2^23
    candidates of seed 1 on the host cores, and the GPU the same sample, lists compared."""
    w = WORKLOADS["knot"]

    def one(seed, budget, prec):
        k = ugpack.Knot(w["dim"], w["T"], seed=seed, pin_sin1=True, max_attempts=budget, precision=prec)
        st = k.run()
        idx, sp = k.strands()
        k.close()
        return st, idx

    forms = [ugpack.PREC_SPLIT_BF16, ugpack.PREC_FP64] if args.precision == "both" else \
        [ugpack.PREC_SPLIT_BF16 if args.precision == "split" else ugpack.PREC_FP64]
    res = {}
    for prec in forms:
        for i in range(args.warmup):
            one(1001 + i, w["max_attempts"], prec)
        per, lists = [], []
        for k in range(args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, idx = one(k + 1, w["max_attempts"], prec)
            torch.cuda.synchronize()
            per.append(time.perf_counter() - t0)
            lists.append(idx.tolist())
        res[prec] = (sum(per), lists)
    main = forms[0]
    wall, lists = res[main]
    split = main == ugpack.PREC_SPLIT_BF16
    line = {
        "metric": "candidates tested/s, knot regime (braid_knot), 2+1 F=32 pinned full band",
        "value": round(args.steps * w["max_attempts"] / wall, 1), "unit": "candidates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f32acc, f64 test" if split else "f64",
        "data": "synthetic (Philox candidate stream, seeds 1..steps)",
        "config": {"workload": w["text"], "dim": w["dim"], "T": w["T"], "candidates_per_step": w["max_attempts"],
                   "precision": "split-bf16" if split else "fp64"},
        "strands_per_step": [len(x) for x in lists],
    }
    if len(forms) == 2:
        fw, fl = res[ugpack.PREC_FP64]
        line["fp64_form"] = {"value": round(args.steps * w["max_attempts"] / fw, 1),
                             "ms_per_step": round(1e3 * fw / args.steps, 3), "same_strands_as_split": fl == lists}
        line["speedup_split_vs_fp64"] = round(fw / wall, 3)
    if args.no_cpu:
        line["cpu_baseline"] = None
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        M = 1 << 23
        t0 = time.perf_counter()
        ref = pyoracle.Model(w["dim"], w["T"], seed=1, pin=True).pack_knot(M)
        dt = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, idx = one(1, M, main)
        torch.cuda.synchronize()
        gdt = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": round(M / dt, 1), "unit": "candidates/s", "cores": _omp_threads(), "kind": "port",
            "sample": f"the first {M} candidates of seed 1 (primary phase; parallel batch test, serial "
                      "in-order admission)", "seconds": round(dt, 3),
            "gpu_same_sample": {"value": round(M / gdt, 1), "seconds": round(gdt, 4),
                                "same_strands": list(idx) == list(ref["idx"])},
            "nproc": os.cpu_count(), "cpu_model": _cpu_model()}
    print(json.dumps(line))
    return 0


def lightcone_paths(ugpack, n_paths: int, T: int):
    """The light-cone workload's universe: 2+1 F=T full spectrum, seed 1, packed on the
    engine schedule to its --until-accept-rate 1e-5 window, then its subpaths until the
    worldlines number n_paths; the first n_paths of them (uniques, then subpaths) as a
    device set."""
    from ugpack import frames

    with ugpack.Packer(2, T, seed=1, precision=ugpack.PREC_SPLIT_BF16, stop_kind=ugpack.STOP_REF_WINDOW,
                       max_attempts=1e12, stop_rate=1e-5) as p:
        st = p.run()
        for _ in range(64):  # subpath rounds of ~2^28 candidates, never split
            st = p.run_sub(max_attempts=1e12, max_candidates=1 << 28)
            if st.n + st.nsub >= n_paths:
                break
        total = st.n + st.nsub
        k = min(n_paths, total)
        par = p.params(0, k)
    paths = frames.Paths(par[:, :, 0], par[:, :, 1::3], par[:, :, 2::3], par[:, :, 3::3])
    axes = [{"a2": par[:, a, 0], "a": par[:, a, 1::3], "b": par[:, a, 2::3], "f": par[:, a, 3::3]}
            for a in range(2)]
    return paths, axes, {"uniques": st.n, "subpaths": st.nsub, "used": k}


def run_lightcone(args, ugpack, torch) -> int:
    """One step = all 480 instants of the light-cone workload (count pass, scan and hit pass
    per instant, hits left in HBM); value = instants/s."""
    import ctypes as C
    import math

    from ugpack import frames
    from ugpack.abi import Front, FrameInfo, Observer, lib

    w = WORKLOADS["lightcone"]
    paths, axes, made = lightcone_paths(ugpack, w["paths"], w["T"])
    z_min = max(2e-3, 0.02 * math.pi)
    fr = Front(1, 1, 0.33, z_min, -1, 0)
    obs = Observer(0, 0, 0.0, 0.0)
    zs = [(0.05 + 0.90 * i / (w["frames"] - 1)) * math.pi for i in range(w["frames"])]
    cap = 1 << 25
    hits = torch.empty(cap * 20, dtype=torch.uint8, device="cuda")
    info = FrameInfo()

    def one_pass():
        total = 0
        for z in zs:
            rc = lib().ug_frame_bake_lightcone(paths._h, C.byref(fr), float(z), C.byref(obs), hits.data_ptr(),
                                               cap, C.byref(info))
            ugpack.abi.check(rc)
            total += int(info.n_hits)
        return total

    for _ in range(args.warmup):
        one_pass()
    per = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nh = one_pass()
        torch.cuda.synchronize()
        per.append(time.perf_counter() - t0)
    wall = sum(per)
    line = {
        "metric": "instants baked/s, light-cone frame bake (frame_bake) of a 19,413-path 2+1 universe",
        "value": round(args.steps * w["frames"] / wall, 2), "unit": "instants/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (a packed universe of the Philox candidate stream, seed 1)",
        "config": {"workload": w["text"], "paths": made["used"], "universe": made, "frames": w["frames"]},
        "hits_per_step": nh,
        # the reference's published figure for the same sizes, on other hardware: ~7 min on an
        # RTX 4090, i.e. 480/420 instants/s
        "reference_published": {"seconds": 420, "instants_per_s": round(480 / 420, 3), "hardware": "RTX 4090",
                                "source": "docs/lab-notes/2026-07-27/index.html:97-98"},
        "ratio_vs_published": round(args.steps * w["frames"] / wall / (480 / 420), 1),
    }
    if args.no_cpu:
        line["cpu_baseline"] = None
    else:
        # the baker oracle (numpy restatement of braidlab.frames.bake_reference) on one instant of
        # the first 8192 paths, the GPU on the same sample, hit counts compared
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import bake_oracle

        k = 8192
        sub = [{key: v[:k] for key, v in ax.items()} for ax in axes]
        z = zs[len(zs) // 2]
        t0 = time.perf_counter()
        ref = bake_oracle.bake(sub, z, z_min, observer_path=0, front="fit")
        dt = time.perf_counter() - t0
        got = frames.bake(sub, z, z_min, observer_path=0, front="fit")
        line["cpu_baseline"] = {
            "value": round(k / w["paths"] / dt, 5), "unit": "instants/s", "cores": 1, "kind": "port",
            "sample": f"one instant (z={z:.4f}) of the first {k} paths, scaled by {k}/{w['paths']} paths",
            "seconds": round(dt, 3), "hits": len(ref), "gpu_same_sample_hits": len(got),
            "nproc": os.cpu_count(), "cpu_model": _cpu_model()}
    paths.close()
    print(json.dumps(line))
    return 0


# CPU samples of the larger workloads (about 10-30 s of oracle work on 16 host cores): the
# first candidates of seed 1, the early, unjammed part of a pack, where every candidate is
# tested at many timesteps and most are admitted -- the CPU's slowest regime.
CPU_PREFIX = {"cfg3": 1 << 20, "cfg5": 1 << 15}
CPU_ENSEMBLE = (32, 1 << 16)  # the ensemble's first 32 universes, 2^16 candidates each


def cpu_sample(ugpack, torch, workload: str, threads: int = 0) -> dict:
    """The CPU oracle on a bounded sample of configs 3-5, with the GPU timed on the same
    sample and the accepted lists compared."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    nproc = os.cpu_count() or 1
    threads = threads or min(nproc, 16)
    split = ugpack.PREC_SPLIT_BF16
    if workload == "cfg4":
        nu, M = CPU_ENSEMBLE
        cfgs = ensemble_configs(ugpack, nu, float(M), split)

        def one(c):
            idx, _, att = pyoracle.Model(c.dim, c.T, seed=c.seed).pack(M)
            return idx, att

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:  # one universe per core (ctypes drops the GIL)
            ref = list(ex.map(one, cfgs))
        dt = time.perf_counter() - t0
        att = sum(a for _, a in ref)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, packers = ugpack.ensemble(cfgs, threads=16, keep=True)
        torch.cuda.synchronize()
        gdt = time.perf_counter() - t0
        same = all(list(p.accepted()[0]) == list(r[0]) for p, r in zip(packers, ref))
        for p in packers:
            p.close()
        sample = f"the ensemble's first {nu} universes, {M} candidates each, one universe per core"
    else:
        w = WORKLOADS[workload]
        M = CPU_PREFIX[workload]
        t0 = time.perf_counter()
        idx, att = pyoracle.Model(w["dim"], w["T"], seed=1).pack_mt(M, threads)
        dt = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with ugpack.Packer(w["dim"], w["T"], seed=1, precision=split, max_attempts=M) as p:
            p.run()
            torch.cuda.synchronize()
            gdt = time.perf_counter() - t0
            same = list(p.accepted()[0]) == list(idx)
        sample = (f"the first {M} candidates of seed 1 (N={len(idx)}), parallel prefilter per "
                  "block and serial in-order admission")
    return {"value": round(att / dt, 1), "unit": "candidates/s", "cores": threads, "kind": "port",
            "sample": sample, "seconds": round(dt, 3),
            "gpu_same_sample": {"value": round(att / gdt, 1), "seconds": round(gdt, 4),
                                "same_accepts": same},
            "nproc": nproc, "cpu_model": _cpu_model()}


def cpu_baseline(threads: int = 0) -> dict:
    """The CPU oracle (serial RSA restatement, oracle/) packing config 2 seed 1 to jam with
    the same fixed-window rule: once on all host cores (parallel prefilter per block, serial
    in-order admission, the reference's multi-core pattern) and once single-threaded."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    nproc = os.cpu_count() or 1
    threads = threads or min(nproc, 16)
    m = pyoracle.Model(2, 64, seed=1)
    t0 = time.perf_counter()
    idx_mt, att_mt = m.pack_mt(1 << 40, threads, stop_kind=1)
    dt_mt = time.perf_counter() - t0
    t0 = time.perf_counter()
    idx_1, _, att_1 = m.pack(1 << 40, stop_kind=1)
    dt_1 = time.perf_counter() - t0
    assert att_1 == att_mt and list(idx_1) == list(idx_mt)
    return {"value": round(att_mt / dt_mt, 1), "unit": "candidates/s", "cores": threads,
            "kind": "port",
            "sample": f"config 2 seed 1 packed to jam ({att_mt} candidates, N={len(idx_mt)}), "
                      "same stop rule as the GPU",
            "s_to_jam": round(dt_mt, 3),
            "single_thread": {"value": round(att_1 / dt_1, 1), "s_to_jam": round(dt_1, 3)},
            "nproc": nproc, "cpu_model": _cpu_model()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="both", choices=["split", "fp64", "both"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--engine", default="tuned", choices=["tuned", "batched", "threads"],
                    help="ensemble engine (cfg4): tuned on this GPU, the batched lock-step engine, or "
                         "worker threads")
    ap.add_argument("--threads", type=int, default=0, help="ensemble workers (cfg4, --engine threads)")
    ap.add_argument("--window", type=int, default=0,
                    help="universes in flight per lane (cfg4, --engine batched; 0 = 32)")
    ap.add_argument("--lanes", type=int, default=1, help="lock-step lanes (cfg4, --engine batched)")
    ap.add_argument("--calib", action="store_true",
                    help="first copy 256 MiB device-to-device (counter calibration for rocprof --pmc)")
    ap.add_argument("--sparse", action="store_true",
                    help="pack on the sparse occupied-cell grid instead of the dense cell table")
    args = ap.parse_args()
    global SPARSE
    SPARSE = args.sparse
    if args.gpus != 1:
        print("bench.py: only one GPU is measured here (a universe packs on one GPU; "
              "ensembles shard by universe)", file=sys.stderr)
        return 2

    import torch

    import ugpack

    torch.cuda.set_device(0)
    if args.calib:
        src = torch.ones(1 << 26, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    prec_main = ugpack.PREC_SPLIT_BF16 if args.precision in ("split", "both") else ugpack.PREC_FP64
    forms = [prec_main] + ([ugpack.PREC_FP64] if args.precision == "both" else [])
    if args.workload == "cfg4":
        return run_ensemble(args, ugpack, torch, forms, prec_main)
    if args.workload == "knot":
        return run_knot(args, ugpack, torch)
    if args.workload == "lightcone":
        return run_lightcone(args, ugpack, torch)

    golden = {}
    for s in range(1, 9):
        f = os.path.join(ROOT, "tests", "golden", f"cfg2_d2_T64_s{s}_jam.json")
        if args.workload != "cfg2":
            break
        if os.path.exists(f):
            golden[s] = json.load(open(f))["idx"]

    results = {}
    frame_stats = None
    cold = {}
    for prec in forms:
        timed_kernels = ugpack.abi.PROFILE_PREFILTER
        # cold start: the first pack of the process (model tables built, buffers allocated)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p, _ = pack_once(ugpack, 2001, prec, workload=args.workload)
        torch.cuda.synchronize()
        cold[prec] = time.perf_counter() - t0
        p.close()
        for w in range(args.warmup):  # profiled (warms the event pool)
            p, _ = pack_once(ugpack, 1001 + w, prec, profile=timed_kernels, workload=args.workload)
            p.close()
        torch.cuda.synchronize()
        cands, accepted, per_pack, parity, digests = 0, [], [], [], []
        # Every timed pack is profiled: each prefilter launch is bracketed by its own
        # timing-only events and the kernels count visits and points in device counters (the
        # event pairs cost about 4% of a config-2 pack; timing the FP64 recheck as well would
        # double that, so its time comes from one separate pack below).  A step is one pack --
        # handle creation, the run to jam and a device synchronisation -- bracketed by
        # synchronisations; between steps the stats are read, the list is checked against the
        # golden fixture and the handle is released (its buffers serve the next pack),
        # outside the step's time.
        pstats, rstats_l = [], []
        for k in range(args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            p, st = pack_once(ugpack, k + 1, prec, profile=timed_kernels, workload=args.workload)
            torch.cuda.synchronize()
            per_pack.append(time.perf_counter() - t0)
            cands += st.attempts
            accepted.append(st.n)
            pstats.append(p.kernel_stats(ugpack.abi.KERNEL_PREFILTER))
            rstats_l.append(p.kernel_stats(ugpack.abi.KERNEL_RECHECK))
            idx, _ = p.accepted()
            digests.append(hashlib.sha256(idx.astype("<i8").tobytes()).hexdigest())
            if (k + 1) in golden:
                parity.append(list(idx) == golden[k + 1])
            p.close()
        wall = sum(per_pack)
        stats, rstats = sum_stats(pstats), sum_stats(rstats_l)
        if prec == ugpack.PREC_SPLIT_BF16:  # the FP64 recheck timed in one untimed pack (seed 1)
            p, _ = pack_once(ugpack, 1, prec, profile=ugpack.abi.PROFILE_ALL | ugpack.abi.PROFILE_REASONS,
                             workload=args.workload)
            one = p.kernel_stats(ugpack.abi.KERNEL_RECHECK)
            pre1 = p.kernel_stats(ugpack.abi.KERNEL_PREFILTER)
            p.close()
            rstats = {"items": rstats["items"], "visits": rstats["visits"], "points": rstats["points"],
                      "per_candidate_tested": round(rstats["items"] / max(stats["items"], 1), 6),
                      # why stage 1 left candidates to the FP64 test, in that pack
                      "left_by_reason_seed1": {"spacing_sum": pre1["left_wsum"],
                                               "possible_contact": pre1["left_possible"],
                                               "no_possible_contact": pre1["left_none"]},
                      "timed_pack_seed1": {k: one[k] for k in ("launches", "items", "ms", "visits", "points")}}
        wl = WORKLOADS[args.workload]
        if wl.get("frames") and prec == prec_main:
            p, _ = pack_once(ugpack, 1, prec, workload=args.workload)
            frame_stats = bake_frames(ugpack, torch, p, wl["frames"], wl["raster"])
            p.close()
        results[prec] = dict(value=cands / wall, wall=wall, per_pack=per_pack, accepted=accepted,
                             parity=parity, stats=stats, rstats=rstats, cands=cands, digests=digests)
        progress(f"{args.workload} precision {prec}: {args.steps} packs, {cands / wall:.4g} candidates/s")

    main_r = results[prec_main]
    split = prec_main == ugpack.PREC_SPLIT_BF16
    line = {
        "metric": "candidates tested/s & s to jam, 2+1 F=64; split-BF16 vs FP64 GPU, same accepts",
        "value": round(main_r["value"], 1),
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * main_r["wall"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+f32acc, f64 recheck" if split else "f64",
        "data": "synthetic (Philox candidate stream, seeds 1..steps)",
        "config": {
            "workload": WORKLOADS[args.workload]["text"],
            "dim": WORKLOADS[args.workload]["dim"], "T": WORKLOADS[args.workload]["T"],
            "seeds": list(range(1, args.steps + 1)),
            "precision": "split-bf16" if split else "fp64",
            "grid": "sparse occupied cells" if SPARSE else "dense cell table",
        },
        "s_to_jam_mean": round(main_r["wall"] / args.steps, 5),
        "cold_first_pack_s": round(cold[prec_main], 5),
        "candidates_per_pack_mean": main_r["cands"] / args.steps,
        "accepted_per_pack": main_r["accepted"],
        "parity_vs_oracle_fixtures": main_r["parity"],
        "roofline": roofline_of(main_r["stats"], WORKLOADS[args.workload]["dim"],
                                WORKLOADS[args.workload]["T"], split,
                                measured_traffic(args.workload, split)),
    }
    if ugpack.PREC_FP64 in results and split:
        f = results[ugpack.PREC_FP64]
        line["fp64_form"] = {"value": round(f["value"], 1), "s_to_jam_mean": round(f["wall"] / args.steps, 5),
                             "accepted_per_pack": f["accepted"], "parity_vs_oracle_fixtures": f["parity"],
                             "cold_first_pack_s": round(cold[ugpack.PREC_FP64], 5),
                             # full accepted lists of every timed pack (SHA-256 digests), not counts
                             "same_accepts_as_split": f["digests"] == main_r["digests"]}
        line["speedup_split_vs_fp64"] = round(main_r["value"] / f["value"], 3)
    if split:
        line["roofline"]["mfma"] = mfma_roofline(main_r["stats"], WORKLOADS[args.workload]["dim"],
                                                 WORKLOADS[args.workload]["T"])
        line["roofline"]["valu_busy_pct"] = measured_valu_busy(args.workload, split)
        # these two come from committed rocprofv3 counter summaries, not from this run
        line["roofline"]["traffic_source"] = f"profiles/{args.workload}_prefilter_hbm_traffic.md (committed)"
        line["roofline"]["valu_busy_source"] = f"profiles/{args.workload}_prefilter_valu.md (committed)"
        line["recheck"] = main_r["rstats"]
        line["split_error"] = split_error(ugpack, torch, args.workload)
    if frame_stats:
        line["frames"] = frame_stats
    if args.no_cpu:
        line["cpu_baseline"] = None
    elif args.workload == "cfg2":
        line["cpu_baseline"] = cpu_baseline()
    else:
        line["cpu_baseline"] = cpu_sample(ugpack, torch, args.workload)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
