"""The reference engines' round schedule and --until-accept-rate window on the GPU.

`braid_ug` / `braid_ug3d` consume candidates in the engines' rounds (cuda/braid_cuda.cu:885-944,
cuda/braid_cuda3d.cu:1446-1566): the round sizes, each round's survivor count against the
table at its start, its admissions and newly occupied cells, the accepted list and the
`done:` line must equal the oracle's (oracle/ug_oracle.c or_pack_ref, pinned by hand in
tests/test_oracle.py) in both precision forms.  The golden cases stop on the window.
"""

from __future__ import annotations

import json
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
DONE_RE = re.compile(r"done:\s*N=(\d+)\s+in\s+(\d+)\s+attempts")
CASES = ["ref_d2_T16_s1_rate2e-6", "ref_d3_T8_s1_rate1.5e-6", "ref_d2_T64_s1_rate1e-4",
         "ref_d3_T32_s2_rate1e-4"]


def _gold(name):
    return json.load(open(os.path.join(GOLD, name + ".json")))


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("name", CASES)
def test_reference_schedule_equals_oracle(gpu, name, precision):
    import ugpack

    g = _gold(name)
    print(f"ref schedule {name} {precision}", flush=True)
    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    with ugpack.Packer(g["dim"], g["T"], seed=g["seed"], precision=prec, stop_kind=ugpack.STOP_REF_WINDOW,
                       max_attempts=g["budget"], stop_rate=g["rate"]) as p:
        st = p.run()
        rounds = p.rounds()
        idx, gid = p.accepted()
    assert st.done and st.attempts == g["attempts"] and st.n == g["n"]
    assert rounds.tolist() == g["rounds"]
    assert list(idx) == g["idx"] and np.array_equal(gid, np.arange(g["n"]))
    assert st.filled == g["filled"]


@pytest.mark.parametrize("name", CASES)
def test_engine_argv_window_fires_in_both_forms(gpu, tmp_path, name):
    # the argv braidlab/engine.py:50-101 builds, run in both precision forms: identical done:
    # lines, equal to the oracle's N and attempts; the curve ends on them
    import ugpack

    g = _gold(name)
    lines = []
    for precision in ("fp64", "split"):
        curve = str(tmp_path / f"c_{precision}.csv")
        cmd = ugpack.command(g["dim"], g["T"], attempts=g["budget"], seed=g["seed"], accept_rate=g["rate"],
                             terms=g["T"], curve_path=curve) + ["--precision", precision]
        print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr
        done = [ln for ln in r.stderr.splitlines() if ln.startswith("done:")]
        assert len(done) == 1
        lines.append(done[0])
        n, att = map(int, DONE_RE.search(done[0]).groups())
        assert (n, att) == (g["n"], g["attempts"])
        rows = np.loadtxt(curve, delimiter=",", skiprows=1, dtype=np.int64, ndmin=2)
        assert tuple(rows[-1, :2]) == (g["attempts"], g["n"])
    assert lines[0] == lines[1]
    if g["dim"] == 2:
        assert re.search(rf"filled {g['filled']}/{g['T'] ** 3} cells", lines[0])


@pytest.mark.parametrize("d,T,budget", [(2, 16, 20000), (3, 12, 20000), (2, 24, 300000), (3, 12, 300000)])
def test_budget_rounds_equal_oracle(gpu, d, T, budget):
    # no window: rounds run while the budget lasts (2+1) or while the sized rounds stay below
    # it (3+1), the last one passing it
    import ugpack

    r = po.Model(d, T, seed=2).pack_ref(budget)
    for prec in (ugpack.PREC_FP64, ugpack.PREC_SPLIT_BF16):
        with ugpack.Packer(d, T, seed=2, precision=prec, stop_kind=ugpack.STOP_REF_WINDOW,
                           max_attempts=budget) as p:
            st = p.run()
            idx, _ = p.accepted()
            rounds = p.rounds()
        assert st.attempts == r["attempts"] >= budget
        assert np.array_equal(idx, r["idx"]) and np.array_equal(rounds, r["rounds"])


def test_max_n_mid_round_and_the_decided_range(gpu):
    import ugpack

    r = po.Model(2, 16, seed=2).pack_ref(10 ** 7, max_n=40)
    with ugpack.Packer(2, 16, seed=2, stop_kind=ugpack.STOP_REF_WINDOW, max_attempts=1e7, max_n=40,
                       precision=ugpack.PREC_SPLIT_BF16) as p:
        st = p.run()
        idx, _ = p.accepted()
        assert st.n == 40 and st.attempts == r["attempts"] and np.array_equal(idx, r["idx"])
        ok = p.check(0, r["decided"])
        assert ok["colliding_uniques"] == 0 and ok["unexplained_rejects"] == 0
        with pytest.raises(ugpack.UgError):
            p.check(0, r["decided"] + 1)


def test_rounds_are_never_split_between_calls(gpu):
    import ugpack

    g = _gold("ref_d2_T16_s1_rate2e-6")
    with ugpack.Packer(2, 16, seed=1, stop_kind=ugpack.STOP_REF_WINDOW, max_attempts=g["budget"],
                       stop_rate=g["rate"]) as p:
        calls = 0
        while not p.run(max_candidates=100000).done:
            calls += 1
        st = p.status()
        assert calls > 3 and st.attempts == g["attempts"]
        assert p.rounds().tolist() == g["rounds"]
