"""Ensemble runs (ug_ensemble_run): independent universes of uneven T and dimension packed
concurrently, one stream per worker.  Bar: every universe's accepted list equals its own
single-universe pack (and the CPU oracle's serial RSA), for any worker count."""

from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

MIX = [(2, 16, 1, 20000), (3, 12, 2, 15000), (2, 32, 3, 30000), (3, 16, 4, 20000),
       (2, 24, 5, 25000), (2, 16, 6, 20000), (3, 12, 7, 15000), (2, 32, 8, 30000),
       (3, 16, 9, 10000), (2, 64, 10, 40000), (2, 16, 11, 5000), (3, 24, 12, 12000)]


def _configs(ugpack, precision):
    return [ugpack.pack_config(d, T, seed=s, max_attempts=m, precision=precision) for d, T, s, m in MIX]


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_ensemble_equals_single_universe_packs(gpu, precision):
    import ugpack

    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    print(f"ensemble {precision}", flush=True)
    status, packers = ugpack.ensemble(_configs(ugpack, prec), threads=4, keep=True)
    for (d, T, s, m), st, p in zip(MIX, status, packers):
        idx, gid = p.accepted()
        p.close()
        assert st.attempts == m and st.done and st.n == len(idx)
        with ugpack.Packer(d, T, seed=s, max_attempts=m, precision=prec) as q:
            q.run()
            want, _ = q.accepted()
        assert list(idx) == list(want), (d, T, s)
        assert list(gid) == list(range(len(idx)))
        ref, _, _ = po.Model(d, T, seed=s).pack(m)  # about 1 s of oracle work for the mix
        assert list(idx) == list(ref), (d, T, s)


def test_ensemble_is_independent_of_worker_count(gpu):
    # destroyed handles hand their buffers to the next universes of other workers' streams;
    # the accepted lists (digests) equal the oracle's for 1 and 12 workers
    import ugpack

    cfgs = _configs(ugpack, ugpack.PREC_SPLIT_BF16)
    a, da = ugpack.ensemble(cfgs, threads=1, digests=True)
    b, db = ugpack.ensemble(cfgs, threads=12, digests=True)
    assert [(s.n, s.attempts, s.filled) for s in a] == [(s.n, s.attempts, s.filled) for s in b]
    want = [ugpack.abi.list_digest(po.Model(d, T, seed=s).pack(m)[0]) for d, T, s, m in MIX]
    assert da == want and db == want


def _mix_slice(ugpack, first: int, count: int, cands: float, precision: int):
    """Universes first..first+count-1 of the 512-universe ensemble workload (bench.py
    ensemble_configs): T from {16, 32, 64, 128} and dim from {2, 3} by a fixed-seed draw,
    seed = index + 1."""
    rng = np.random.default_rng(20260709)
    Ts = rng.choice([16, 32, 64, 128], size=512)
    ds = rng.choice([2, 3], size=512)
    return [ugpack.pack_config(int(ds[i]), int(Ts[i]), seed=i + 1, max_attempts=cands, precision=precision)
            for i in range(first, first + count)]


def test_ensemble_workload_slice_at_full_size(gpu):
    # the first 20 universes of the ensemble workload at its 10^7 candidates each (three of
    # them 3+1 T=128), plus a 3+1 T=128 seed-1 universe whose first 2^20 candidates have an
    # oracle golden list: every universe satisfies the serial-RSA characterization over its
    # whole range, kept and discarded runs report the same digests
    import json
    import os

    import ugpack

    cfgs = _mix_slice(ugpack, 0, 20, 1e7, ugpack.PREC_SPLIT_BF16)
    cfgs.append(ugpack.pack_config(3, 128, seed=1, max_attempts=1e7, precision=ugpack.PREC_SPLIT_BF16))
    assert sum(1 for c in cfgs if (c.dim, c.T) == (3, 128)) == 4
    print("ensemble slice: 21 universes x 1e7", flush=True)
    status, packers, dig = ugpack.ensemble(cfgs, threads=8, keep=True, digests=True)
    _, dig2 = ugpack.ensemble(cfgs, threads=8, digests=True)
    assert dig == dig2
    try:
        for c, st, p, dg in zip(cfgs, status, packers, dig):
            assert st.done and st.attempts == 10 ** 7
            chk = p.check(0, st.attempts)
            assert chk == {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": st.n,
                           "rejects": st.attempts - st.n}, (c.dim, c.T, c.seed)
            idx, _ = p.accepted()
            assert ugpack.abi.list_digest(idx) == dg
        g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cfg3_d3_T128_s1_prefix1M.json")))
        idx, _ = packers[-1].accepted()
        assert list(idx[idx < (1 << 20)]) == g["idx"]
    finally:
        for p in packers:
            p.close()


def test_ensemble_reports_a_failing_universe(gpu):
    import ugpack

    cfgs = _configs(ugpack, ugpack.PREC_FP64)[:3]
    cfgs.append(ugpack.pack_config(2, 16, terms=40, seed=1, max_attempts=100))  # terms > pool
    with pytest.raises(ugpack.UgError, match="universe 3"):
        ugpack.ensemble(cfgs, threads=2)
    # with keep, the universes finished before the failure are released by the library
    with pytest.raises(ugpack.UgError, match="universe 3"):
        ugpack.ensemble(cfgs, threads=2, keep=True)


def test_tuned_worker_count_is_one_of_the_choices(gpu):
    import ugpack
    from ugpack import tune

    cfgs = _configs(ugpack, ugpack.PREC_SPLIT_BF16)
    best, times = tune.ensemble_threads(cfgs, choices=(1, 4, 12))
    assert best in (1, 4, 12) and sorted(times) == [1, 4, 12]
    assert times[best] == min(times.values()) and all(t > 0 for t in times.values())
    # the tuned count packs the same universes as any other
    a = ugpack.ensemble(cfgs, threads=best)
    b = ugpack.ensemble(cfgs, threads=1)
    assert [(s.n, s.attempts) for s in a] == [(s.n, s.attempts) for s in b]


def test_tuned_plan_and_sizing_leave_the_lists_unchanged(gpu):
    # the tuner's sub-batch sizing settings change speed only: every universe's accepted list
    # (digest) is the oracle's for each setting
    import ugpack
    from ugpack import tune

    cfgs = _configs(ugpack, ugpack.PREC_SPLIT_BF16)
    plan, sweep = tune.ensemble_plan(cfgs, choices=(1, 2, 4), targets=(512, 1024), sub_max=(1 << 16, 1 << 20),
                                     batched=((1, 2), (2, 6)))
    assert plan["engine"] in ("threads", "batched") and plan["sub_target"] in (0, 512, 1024)
    assert plan["threads"] in (1, 2, 4)
    if plan["engine"] == "batched":
        assert (plan["lanes"], plan["window"]) in ((1, 2), (2, 6))
    assert len(sweep) == 3 + 2 + 4 and all(t > 0 for t in sweep.values())
    want = [ugpack.abi.list_digest(po.Model(d, T, seed=s).pack(m)[0]) for d, T, s, m in MIX]
    for tg, sm in ((512, 1 << 16), (2048, 1 << 20), (plan["sub_target"], plan["sub_max"])):
        for batched in (False, True):
            _, dg = ugpack.ensemble(tune.with_sizing(cfgs, tg, sm), threads=plan["threads"], digests=True,
                                    batched=batched, window=plan["window"], lanes=max(plan["lanes"], 1))
            assert dg == want, (tg, sm, batched)
