"""Writes tests/golden/frames_*.json: frame-bake hit lists from the reference's own
Python oracle, braidlab.frames.bake_reference (braidlab/frames.py:246-362).

Nothing else pins the exact crossing lists of the past-light-cone baker, so this script
imports the reference in the build container (it is not available, and not needed, where
the tests run).  Each fixture holds the input worldlines as per-axis term arrays and the
settings, and the hits (src, dx, dy, chi, z_emit) sorted by (src, chi).

    python tests/golden/make_frames_golden.py
"""

from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

# the 3-worldline multi-term dump of the reference's tests/test_frames.py:19-24
DUMP3 = (
    "ax2,ay2,ax_1,bx_1,fx_1,ay_1,by_1,fy_1,ax_2,bx_2,fx_2,ay_2,by_2,fy_2\n"
    "0.31,-0.62,0.20,3,0,-0.13,2,0.8,0.10,4,1.1,0.24,5,0\n"
    "-0.55,0.12,-0.33,2,0.4,0.25,4,2.1,0.08,5,0,-0.15,3,0\n"
    "0.05,0.77,0.5,2,2.2,-0.5,2,0.3,0.0,3,0,0.0,4,0\n"
)

CASES = [
    # name, dump, z_obs/pi, z_min/pi, kwargs
    ("dump3_fit_path0", "dump3", 0.5, 0.02, dict(observer_path=0, front="fit")),
    ("dump3_budget_path0", "dump3", 0.5, 0.02, dict(observer_path=0, front="budget")),
    ("dump3_speed_point", "dump3", 0.6, 0.05, dict(observer_point=(0.1, 0.3), front="speed", speed=0.33)),
    ("dump3_fit_circle_path2", "dump3", 0.4, 0.03, dict(observer_path=2, front="fit", cheb=False)),
    ("dump3_fit_wraps1_path1", "dump3", 0.7, 0.02, dict(observer_path=1, front="fit", wraps=1)),
    ("demo50_fit_path0", "demo50", 0.5, 0.05, dict(observer_path=0, front="fit")),
    ("demo50_budget_point", "demo50", 0.35, 0.1, dict(observer_point=(-0.2, 0.4), front="budget")),
]


def main() -> None:
    sys.path.insert(0, REF)
    from braidlab.corrdim import load_axis_terms
    from braidlab.frames import bake_reference

    dumps = {}
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "dump3.csv")
        open(p, "w").write(DUMP3)
        dumps["dump3"] = load_axis_terms(p)
    dumps["demo50"] = load_axis_terms(os.path.join(REF, "examples", "dumps", "d2_demo_phase_T100_s0.csv"))
    for name, dump, zo, zm, kw in CASES:
        axes = dumps[dump]
        hits = bake_reference(axes, zo * np.pi, zm * np.pi, **kw)
        hits = sorted((int(s), float(dx), float(dy), float(c), float(z)) for s, dx, dy, c, z in hits)
        rec = {
            "name": name, "z_obs": zo * np.pi, "z_min": zm * np.pi,
            "settings": {k: (list(v) if isinstance(v, tuple) else v) for k, v in kw.items()},
            "axes": [{"a": ax.a.tolist(), "b": ax.b.tolist(), "f": ax.f.tolist(), "a2": ax.a2.tolist()}
                     for ax in axes],
            "hits": hits,
        }
        with open(os.path.join(HERE, f"frames_{name}.json"), "w") as f:
            json.dump(rec, f)
        print(f"{name}: {len(hits)} hits")


if __name__ == "__main__":
    main()
