"""The batched ensemble engine (ug_ensemble_run_batched): universes of uneven T and dimension
packed in lock-step on one stream, every stage one launch for all of them.  Bar: every
universe's accepted list, attempts and status equal its own single-universe pack (and the
CPU oracle's serial RSA), whatever the window of universes in flight."""

from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

MIX = [(2, 16, 1, 20000), (3, 12, 2, 15000), (2, 32, 3, 30000), (3, 16, 4, 20000),
       (2, 24, 5, 25000), (2, 16, 6, 20000), (3, 12, 7, 15000), (2, 32, 8, 30000),
       (3, 16, 9, 10000), (2, 64, 10, 40000), (2, 16, 11, 5000), (3, 24, 12, 12000)]


def _configs(ugpack, precision, mix=MIX):
    return [ugpack.pack_config(d, T, seed=s, max_attempts=m, precision=precision) for d, T, s, m in mix]


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_batched_equals_single_universe_packs(gpu, precision):
    import ugpack

    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    print(f"batched ensemble {precision}", flush=True)
    status, packers = ugpack.ensemble(_configs(ugpack, prec), keep=True, batched=True)
    try:
        for (d, T, s, m), st, p in zip(MIX, status, packers):
            idx, gid = p.accepted()
            assert st.attempts == m and st.done and st.n == len(idx)
            with ugpack.Packer(d, T, seed=s, max_attempts=m, precision=prec) as q:
                qs = q.run()
                want, _ = q.accepted()
            assert list(idx) == list(want), (d, T, s)
            assert list(gid) == list(range(len(idx)))
            assert (st.survivors, st.filled) == (qs.survivors, qs.filled), (d, T, s)
            ref, _, _ = po.Model(d, T, seed=s).pack(m)
            assert list(idx) == list(ref), (d, T, s)
            # the kept handle is a usable pack: its curve and positions are there
            assert len(p.curve()) > 0
    finally:
        for p in packers:
            p.close()


@pytest.mark.parametrize("lanes,window", [(1, 1), (1, 3), (3, 2), (4, 64)])
def test_batched_window_leaves_the_lists_unchanged(gpu, lanes, window):
    # universes retire and the next ones take their place, on one lane or several; a config
    # outside the batched kernels' cover (the sparse grid) runs on the single-universe loop
    import ugpack

    cfgs = _configs(ugpack, ugpack.PREC_SPLIT_BF16)
    cfgs.insert(4, ugpack.pack_config(2, 24, seed=13, max_attempts=20000, precision=ugpack.PREC_SPLIT_BF16,
                                      sparse=True))
    print(f"batched lanes {lanes} window {window}", flush=True)
    st, dg = ugpack.ensemble(cfgs, digests=True, batched=True, window=window, lanes=lanes)
    mix = MIX[:4] + [(2, 24, 13, 20000)] + MIX[4:]
    want = [ugpack.abi.list_digest(po.Model(d, T, seed=s).pack(m)[0]) for d, T, s, m in mix]
    assert dg == want
    assert [s.attempts for s in st] == [m for _, _, _, m in mix]


def test_batched_stop_rules_equal_single_packs(gpu):
    # fixed-window stop (the headline rule at a small size) and max_n, in both forms
    import ugpack

    cfgs, single = [], []
    for prec in (ugpack.PREC_FP64, ugpack.PREC_SPLIT_BF16):
        for d, T, s in ((2, 16, 1), (3, 12, 2), (2, 32, 3)):
            kw = dict(seed=s, precision=prec, stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=2e-3,
                      stop_batch=1 << 14, stop_window=1 << 16, max_attempts=1 << 24)
            cfgs.append(ugpack.pack_config(d, T, **kw))
            single.append((d, T, kw))
        cfgs.append(ugpack.pack_config(2, 16, seed=9, precision=prec, max_n=30, max_attempts=1e6))
        single.append((2, 16, dict(seed=9, precision=prec, max_n=30, max_attempts=1e6)))
    print("batched stop rules", flush=True)
    st, dg = ugpack.ensemble(cfgs, digests=True, batched=True, window=4)
    for (d, T, kw), s, g in zip(single, st, dg):
        with ugpack.Packer(d, T, **kw) as q:
            qs = q.run()
            idx, _ = q.accepted()
        assert (s.n, s.attempts, s.done) == (qs.n, qs.attempts, qs.done), (d, T, kw)
        assert g == ugpack.abi.list_digest(idx), (d, T, kw)


def test_batched_workload_slice_at_full_size(gpu):
    # the first 24 universes of the 512-universe ensemble workload at 10^7 candidates each:
    # the batched engine's digests equal the worker-thread runner's, and every universe
    # satisfies the serial-RSA characterization over its whole range
    import ugpack

    rng = np.random.default_rng(20260709)
    Ts = rng.choice([16, 32, 64, 128], size=512)
    ds = rng.choice([2, 3], size=512)
    cfgs = [ugpack.pack_config(int(ds[i]), int(Ts[i]), seed=i + 1, max_attempts=1e7,
                               precision=ugpack.PREC_SPLIT_BF16) for i in range(24)]
    print("batched slice: 24 universes x 1e7", flush=True)
    status, packers, dig = ugpack.ensemble(cfgs, keep=True, digests=True, batched=True, lanes=2, window=8)
    try:
        _, dig2 = ugpack.ensemble(cfgs, threads=8, digests=True)
        assert dig == dig2
        for c, st, p in zip(cfgs, status, packers):
            assert st.done and st.attempts == 10 ** 7
            chk = p.check(0, st.attempts)
            assert chk == {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": st.n,
                           "rejects": st.attempts - st.n}, (c.dim, c.T, c.seed)
    finally:
        for p in packers:
            p.close()


def test_batched_reports_a_failing_universe(gpu):
    import ugpack

    cfgs = _configs(ugpack, ugpack.PREC_FP64)[:3]
    cfgs.append(ugpack.pack_config(2, 16, terms=40, seed=1, max_attempts=100))  # terms > pool
    with pytest.raises(ugpack.UgError, match="universe 3"):
        ugpack.ensemble(cfgs, batched=True)
    with pytest.raises(ugpack.UgError, match="universe 3"):
        ugpack.ensemble(cfgs, keep=True, batched=True, lanes=2, window=1)
