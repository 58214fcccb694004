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
    import ugpack

    cfgs = _configs(ugpack, ugpack.PREC_SPLIT_BF16)
    a = ugpack.ensemble(cfgs, threads=1)
    b = ugpack.ensemble(cfgs, threads=12)
    assert [(s.n, s.attempts, s.filled) for s in a] == [(s.n, s.attempts, s.filled) for s in b]


def test_ensemble_reports_a_failing_universe(gpu):
    import ugpack

    cfgs = _configs(ugpack, ugpack.PREC_FP64)[:3]
    cfgs.append(ugpack.pack_config(2, 16, terms=40, seed=1, max_attempts=100))  # terms > pool
    with pytest.raises(ugpack.UgError, match="universe 3"):
        ugpack.ensemble(cfgs, threads=2)


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
