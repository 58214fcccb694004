"""Frame baker: the CPU oracle (oracle/bake_oracle.py) against the reference's own
bake_reference hit lists (tests/golden/frames_*.json, written by make_frames_golden.py),
and the reference's frame properties (tests/test_frames.py:34-68 of the reference)."""

from __future__ import annotations

import glob
import json
import math
import os

import numpy as np
import pytest

import bake_oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(os.path.basename(p)[7:-5] for p in glob.glob(os.path.join(GOLD, "frames_*.json")))


def load_case(name):
    g = json.load(open(os.path.join(GOLD, f"frames_{name}.json")))
    axes = [{k: np.array(v) for k, v in ax.items()} for ax in g["axes"]]
    kw = dict(g["settings"])
    if "observer_point" in kw:
        kw["observer_point"] = tuple(kw["observer_point"])
    return g, axes, kw


def match_hits(got, want, tol):
    """Same number of hits and, sorted by (src, chi), equal within tol."""
    a = sorted(got, key=lambda h: (h[0], h[3]))
    b = sorted(want, key=lambda h: (h[0], h[3]))
    assert len(a) == len(b), (len(a), len(b))
    for x, y in zip(a, b):
        assert x[0] == y[0]
        assert np.allclose(x[1:], y[1:], rtol=0, atol=tol), (x, y)


@pytest.mark.parametrize("name", CASES)
def test_oracle_equals_reference_hits(name):
    g, axes, kw = load_case(name)
    got = bake_oracle.bake(axes, g["z_obs"], g["z_min"], **kw)
    match_hits(got, [tuple(h) for h in g["hits"]], 1e-9)


@pytest.mark.parametrize("name", ["dump3_fit_path0", "dump3_budget_path0", "demo50_fit_path0"])
def test_oracle_hits_lie_on_the_front(name):
    g, axes, kw = load_case(name)
    cheb = kw.get("cheb", True) or kw.get("front") == "budget"
    for src, dx, dy, chi, z_emit in bake_oracle.bake(axes, g["z_obs"], g["z_min"], **kw):
        assert g["z_min"] <= z_emit <= g["z_obs"]
        assert chi >= 0
        d = max(abs(dx), abs(dy)) if cheb else math.hypot(dx, dy)
        assert d == pytest.approx(chi, abs=1e-6)


def test_oracle_point_observer_matches_path_observer():
    g, axes, _ = load_case("dump3_fit_path0")
    z_obs, z_min = 0.5 * np.pi, 0.05 * np.pi
    ox, oy = (float(v) for v in bake_oracle._comov(axes, 1, z_obs)[:2])
    a = bake_oracle.bake(axes, z_obs, z_min, observer_path=1)
    b = bake_oracle.bake(axes, z_obs, z_min, observer_point=(ox, oy))

    def keys(hits):
        return {(h[0], round(h[3], 8)) for h in hits if not (h[0] == 1 and h[3] < 1e-6)}

    assert keys(a) == keys(b)
