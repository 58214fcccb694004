"""The command-line front ends on the GPU, driven with the argv braidlab/engine.py builds.

Checks: the `done:` line (parsed with the engine's own pattern, braidlab/engine.py:14) gives
the oracle's N and attempts; the curve CSV ends on (attempts, N); the dump CSV's parameters
rebuild each accepted worldline (the reference's evaluator, braidlab/corrdim.py:116-137)
to within the dump's 10 significant digits of the oracle's canonical positions; subpath
runs report the fixture's counts and groups; frame_bake_ug's BRF1 file carries the
reference's hit lists.
"""

from __future__ import annotations

import csv
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po
from test_frames import load_case, match_hits

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
DONE_RE = re.compile(r"done:\s*N=(\d+)\s+in\s+(\d+)\s+attempts")


def _gold(name):
    return json.load(open(os.path.join(GOLD, name + ".json")))


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r.stderr


def _read_dump(path, dim):
    rows = list(csv.DictReader(open(path)))
    names = "xyw"[:dim]
    nw = 1
    while f"ax_{nw + 1}" in rows[0]:
        nw += 1
    out = []
    for r in rows:
        a2 = np.array([float(r[f"a{n}2"]) for n in names])
        amp = np.array([[float(r[f"a{n}_{j}"]) for j in range(1, nw + 1)] for n in names])
        frq = np.array([[float(r[f"b{n}_{j}"]) for j in range(1, nw + 1)] for n in names])
        phs = np.array([[float(r[f"f{n}_{j}"]) for j in range(1, nw + 1)] for n in names])
        out.append((a2, amp, frq, phs, int(r["gid"]) if "gid" in r else None))
    return out


def _positions_from_dump(row, T):
    """X_t = [a2 sin z + sum_j a_j (sin(b_j z + f_j) - sin f_j)] / sin z on the engine grid."""
    a2, amp, frq, phs, _ = row
    z = (np.arange(T) + 1.0) * math.pi / (T + 1.0)
    s = np.sin(z)
    terms = amp[:, None, :] * (np.sin(frq[:, None, :] * z[None, :, None] + phs[:, None, :])
                               - np.sin(phs[:, None, :]))
    return ((a2[:, None] * s[None, :] + terms.sum(-1)) / s[None, :]).T     # (T, dim)


def _check_dump(path, dim, T, model, idx, k=40):
    rows = _read_dump(path, dim)
    assert len(rows) == len(idx)
    for i in list(range(min(k, len(idx)))) + [len(idx) - 1]:
        X = _positions_from_dump(rows[i], T)
        d = np.abs(X - model.positions(int(idx[i])))
        d = np.minimum(d % 2.0, 2.0 - d % 2.0)
        assert d.max() < 1e-7, (i, d.max())
    return rows


def _check_curve(path, attempts, n, four=False):
    rows = list(csv.reader(open(path)))
    assert rows[0] == (["attempts", "n", "nsub", "filled"] if four else ["attempts", "n"])
    vals = np.array(rows[1:], dtype=np.int64)
    assert np.all(np.diff(vals[:, 0]) >= 0) and np.all(np.diff(vals[:, 1]) >= 0)
    assert vals[-1, 0] == attempts and vals[-1, 1] == n
    return vals


@pytest.mark.parametrize("seed", [1, 3])
def test_braid_ug_engine_argv_equals_oracle(gpu, tmp_path, seed):
    # --attempts 20000 without a window: one engine round of 2^16 candidates, as the
    # reference's loop runs it (cuda/braid_cuda.cu:885,900)
    import ugpack

    m = po.Model(2, 16, seed=seed)
    r = m.pack_ref(20000)
    curve, dump = str(tmp_path / "c.csv"), str(tmp_path / "d.csv")
    cmd = ugpack.command(2, 16, attempts=20000.0, seed=seed, accept_rate=0.0, terms=16,
                         curve_path=curve) + ["--dump-params", dump]
    err = _run(cmd)
    n, att = map(int, DONE_RE.search(err).groups())
    assert n == len(r["idx"]) and att == r["attempts"] == 1 << 16
    assert re.search(rf"filled {r['rounds'][:, 4].sum()}/4096 cells", err)
    _check_curve(curve, att, n)
    _check_dump(dump, 2, 16, m, r["idx"])


def test_braid_ug_terms2_legacy_dump(gpu, tmp_path):
    import ugpack

    m = po.Model(2, 32, terms=2, seed=1)
    r = m.pack_ref(50000)
    dump = str(tmp_path / "d.csv")
    err = _run(ugpack.command(2, 32, attempts=50000.0, seed=1, accept_rate=0.0,
                              curve_path=str(tmp_path / "c.csv")) + ["--dump-params", dump])
    n, att = map(int, DONE_RE.search(err).groups())
    assert (n, att) == (len(r["idx"]), r["attempts"])
    rows = list(csv.DictReader(open(dump)))
    assert list(rows[0].keys()) == ["ax", "ay", "bx", "by", "ax2", "ay2", "fx", "fy"]
    for i in (0, 1, len(rows) - 1):
        cd = m.candidate(int(r["idx"][i]))
        assert float(rows[i]["ax2"]) == pytest.approx(cd.a1[0], abs=1e-9)
        assert float(rows[i]["bx"]) == cd.b[0][0] and float(rows[i]["by"]) == cd.b[1][0]


def test_braid_ug_subpaths_equals_golden(gpu, tmp_path):
    import ugpack

    g = _gold("subref_d2_T16_s1_accept2e-3")
    curve, dump = str(tmp_path / "c.csv"), str(tmp_path / "d.csv")
    cmd = ugpack.command(2, 16, attempts=20000.0, seed=1, accept_rate=g["accept_rate"], terms=16,
                         subpaths=True, sub_attempts=1e15, curve_path=curve)
    err = _run(cmd + ["--dump-params", dump])
    n, att = map(int, DONE_RE.search(err).groups())
    assert (n, att) == (g["n"], g["unique_attempts"])
    m = re.search(r"sub done \((.+?)\): Nsub=(\d+) in (\d+) attempts", err)
    assert m.group(1) == "accept-rate"
    assert int(m.group(2)) == g["nsub"] and int(m.group(3)) == g["sub_attempts"]
    vals = _check_curve(curve, g["attempts"], g["n"], four=True)
    assert vals[-1, 2] == g["nsub"]
    gids = [int(r["gid"]) for r in csv.DictReader(open(dump))]
    assert len(gids) == g["n"] + g["nsub"]
    assert gids[:g["n"]] == list(range(g["n"]))
    assert gids[g["n"]:g["n"] + 64] == g["head_gid"] and gids[-64:] == g["tail_gid"]


@pytest.mark.parametrize("euclid", [False, True])
def test_braid_ug3d_engine_argv(gpu, tmp_path, euclid):
    import ugpack

    m = po.Model(3, 12, seed=1, euclid=euclid)
    r = m.pack_ref(20000)
    want, att = r["idx"], r["attempts"]
    curve, dump = str(tmp_path / "c.csv"), str(tmp_path / "d.csv")
    cmd = ugpack.command(3, 12, attempts=20000.0, seed=1, accept_rate=0.0, terms=12,
                         euclid=euclid, curve_path=curve) + ["--dump-params", dump]
    err = _run(cmd)
    n, a = map(int, DONE_RE.search(err).groups())
    assert (n, a) == (len(want), att)
    rows = list(csv.reader(open(dump)))
    assert rows[0][:6] == ["ax2", "ay2", "aw2", "ax_1", "bx_1", "fx_1"] and "gid" not in rows[0]
    _check_curve(curve, att, n)
    _check_dump(dump, 3, 12, m, want)


def test_pack_api_equals_cli(gpu, tmp_path):
    import ugpack

    r = ugpack.pack(2, 24, terms=5, seed=3, attempts=30000)
    want = po.Model(2, 24, terms=5, seed=3).pack_ref(30000)
    assert list(r.idx) == list(want["idx"]) and r.attempts == want["attempts"]
    assert r.params.shape == (r.n, 2, 1 + 3 * 4)


@pytest.mark.parametrize("name", ["dump3_fit_path0", "dump3_speed_point", "demo50_budget_point"])
def test_frame_bake_ug_equals_reference_hits(gpu, tmp_path, name):
    import ugpack
    from ugpack import frames

    g, axes, kw = load_case(name)
    nw = axes[0]["a"].shape[1]
    dump = tmp_path / "d.csv"
    with open(dump, "w") as f:
        f.write("ax2,ay2" + "".join(f",ax_{j},bx_{j},fx_{j},ay_{j},by_{j},fy_{j}"
                                    for j in range(1, nw + 1)) + "\n")
        for p in range(len(axes[0]["a2"])):
            f.write(f"{float(axes[0]['a2'][p])!r},{float(axes[1]['a2'][p])!r}")
            for j in range(nw):
                for ax in axes:
                    f.write(",%r,%r,%r" % (float(ax["a"][p, j]), float(ax["b"][p, j]),
                                           float(ax["f"][p, j])))
            f.write("\n")
    out = tmp_path / "f.bin"
    argv = [os.path.join(ugpack.BIN_DIR, "frame_bake_ug"), "--dump", str(dump), "--out", str(out),
            "--from", repr(g["z_obs"] / math.pi), "--to", repr(g["z_obs"] / math.pi),
            "--zmin", repr(g["z_min"] / math.pi), "--front", kw.get("front", "fit")]
    if "observer_path" in kw:
        argv += ["--observer-path", str(kw["observer_path"])]
    else:
        argv += ["--observer-point", "%r,%r" % tuple(float(v) for v in kw["observer_point"])]
    if kw.get("cheb") is False:
        argv += ["--metric", "circle"]
    if "speed" in kw:
        argv += ["--speed", repr(kw["speed"])]
    if kw.get("wraps") is not None:
        argv += ["--wraps", str(kw["wraps"])]
    err = _run(argv)
    assert "done: wrote" in err
    fs = frames.load_frames(str(out))
    hits = fs["frames"][0][0]["hits"]
    got = [(int(h["src"]), float(h["dx"]), float(h["dy"]), float(h["chi"]), float(h["z_emit"]))
           for h in hits]
    match_hits(got, [tuple(h) for h in g["hits"]], 2e-6)
    # the same file as the in-process writer
    out2 = tmp_path / "f2.bin"
    obs = [("path", kw["observer_path"])] if "observer_path" in kw else [("point", kw["observer_point"])]
    frames.bake_file(str(dump), str(out2), obs, from_pi=g["z_obs"] / math.pi, to_pi=g["z_obs"] / math.pi,
                     frames=1, front=kw.get("front", "fit"),
                     metric="circle" if kw.get("cheb") is False else "square",
                     speed=kw.get("speed", 0.33), zmin_pi=g["z_min"] / math.pi,
                     wraps=-1 if kw.get("wraps") is None else kw["wraps"])
    assert out.read_bytes() == out2.read_bytes()
