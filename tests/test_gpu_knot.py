"""The knot-regime engine (cuda/braid_knot.cu) on the GPU: strands, species, attempts and
batches equal the oracle's (oracle/ug_oracle.c or_pack_knot, pinned in
tests/test_oracle.py); the primary species equals the grid packer's serial RSA; the
braid_knot_ug front end writes the reference's lines, curve and dump."""

from __future__ import annotations

import math
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [  # d, T, terms, seed, pin, budget, nu, nu_div, max_strands
    (2, 16, 0, 1, True, 2e6, 5e5, 2.0, 0),
    (3, 12, 0, 2, True, 1e6, 3e5, 4.0, 0),
    (2, 32, 0, 3, False, 3e6, 0.0, 1.0, 0),
    (2, 24, 5, 4, True, 2e6, 1e6, 3.0, 0),
    (3, 16, 0, 5, False, 1e6, 2e5, 2.0, 40),
    (2, 64, 0, 6, True, 4e6, 1e6, 2.0, 0),
]


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("case", CASES)
def test_knot_equals_oracle(gpu, case, precision):
    import ugpack

    d, T, terms, seed, pin, budget, nu, div, cap = case
    print(f"knot {case} {precision}", flush=True)
    want = po.Model(d, T, terms=terms or None, seed=seed, pin=pin).pack_knot(
        budget, nu_attempts=nu, nu_div=div, max_strands=cap or 16384)
    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    with ugpack.Knot(d, T, terms=terms, seed=seed, pin_sin1=pin, max_attempts=budget, nu_attempts=nu,
                     nu_div=div, max_strands=cap, precision=prec) as k:
        st = k.run()
        idx, sp = k.strands()
        rows = k.batches()
    assert st.done and st.attempts == want["attempts"]
    assert st.n_primary == want["n_primary"] and st.primary_attempts == want["primary_attempts"]
    assert np.array_equal(idx, want["idx"]) and np.array_equal(sp, want["species"])
    assert np.array_equal(rows, want["batches"])
    if cap:
        assert st.n == cap


def test_knot_primary_equals_the_grid_packer(gpu):
    # brute force against every strand and the cell table give the same serial RSA
    import ugpack

    for d, T, pin in ((2, 32, True), (3, 16, True), (2, 48, False)):
        with ugpack.Knot(d, T, seed=7, pin_sin1=pin, max_attempts=2e6, precision=ugpack.PREC_SPLIT_BF16) as k:
            k.run()
            kidx, _ = k.strands()
        with ugpack.Packer(d, T, seed=7, pin_sin1=pin, max_attempts=2e6, precision=ugpack.PREC_SPLIT_BF16) as p:
            p.run()
            pidx, _ = p.accepted()
        assert len(kidx) > 5 and np.array_equal(kidx, pidx), (d, T)


def test_knot_in_pieces_equals_one_run(gpu):
    import ugpack

    with ugpack.Knot(2, 16, seed=1, pin_sin1=True, max_attempts=2e6, nu_attempts=5e5, nu_div=2.0) as k:
        k.run()
        a = k.strands()
    with ugpack.Knot(2, 16, seed=1, pin_sin1=True, max_attempts=2e6, nu_attempts=5e5, nu_div=2.0) as k:
        calls = 0
        while not k.run(max_candidates=300000).done:
            calls += 1
        b = k.strands()
    assert calls > 3 and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_braid_knot_ug_outputs(gpu, tmp_path):
    import ugpack

    d, T, seed, budget, nu, div = 2, 16, 1, 2e6, 5e5, 2.0
    want = po.Model(d, T, seed=seed, pin=True).pack_knot(budget, nu_attempts=nu, nu_div=div)
    curve, dump = tmp_path / "k.csv", tmp_path / "k.dump"
    cmd = [os.path.join(ugpack.BIN_DIR, "braid_knot_ug"), "-t", str(T), "--dim", str(d), "--pin-sin1",
           "--seed", str(seed), "--attempts", repr(budget), "--nu-attempts", repr(nu), "--nu-div", repr(div),
           "--curve", str(curve), "--dump", str(dump)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n, npri = len(want["idx"]), want["n_primary"]
    assert f"phase 0 done: N={npri} at {want['primary_attempts']} attempts (cell=0.125)" in r.stderr
    assert f"phase 1 done: N={n} at {want['attempts']} attempts (cell=0.0625)" in r.stderr
    assert re.search(rf"done: N={n} \(primary {npri}, nu {n - npri}\) in {want['attempts']} attempts", r.stderr)
    rows = np.loadtxt(curve, delimiter=",", skiprows=1, dtype=np.int64, ndmin=2)
    assert np.array_equal(rows[:, 0], want["idx"] + 1) and np.array_equal(rows[:, 1], np.arange(1, n + 1))
    assert np.array_equal(rows[:, 2], want["species"])
    lines = dump.read_text().splitlines()
    assert lines[0] == f"# species,attempt, then per axis: a2, then nw x (b,a,f); dim={d} T={T} nw={T - 1}"
    m = po.Model(d, T, seed=seed, pin=True)
    z = (np.arange(T) + 1.0) * math.pi / (T + 1.0)
    for line, c, s in zip(lines[1:], want["idx"], want["species"]):
        v = line.split(",")
        assert int(v[0]) == s and int(v[1]) == c + 1
        X = np.zeros((T, d))
        off = 2
        for a in range(d):
            a2 = float(v[off])
            trip = np.array(v[off + 1: off + 1 + 3 * (T - 1)], dtype=float).reshape(-1, 3)
            off += 1 + 3 * (T - 1)
            b, amp, f = trip[:, 0], trip[:, 1], trip[:, 2]
            w = (amp[None] * (np.sin(b[None] * z[:, None] + f[None]) - np.sin(f[None]))).sum(1)
            X[:, a] = a2 + w / np.sin(z)
        X = X - 2.0 * np.floor((X + 1.0) * 0.5)
        dd = np.abs(X - m.positions(int(c)))
        assert np.minimum(dd, 2.0 - dd).max() < 1e-12
