"""CPU oracle of the frame baker (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the reference's past-light-cone crossing finder,
braidlab/frames.py:246-362 (bake_reference), itself the validation oracle of
cuda/frame_bake.cu:281-371.  Only tests/ load it.  Worldlines are given per axis as
dicts (or objects) with a, b, f of shape (n, nw) and a2 of shape (n,).

Steps (as in the reference): rewind grid of M + 1 rows -- chi-uniform for the conformal
fronts (z = 2 atan(exp(eta_obs - chi/c)), eta = ln tan(z/2); c = speed or -1/eta(z_min) for
"fit"), z-uniform with the wiggle-budget reach for "budget" (M = min(4000, max(420, 2 bmax)));
torus images (2 nx - ox, 2 ny - oy), |n| <= K = ceil((chi_max + 4)/2); per (path, image) a
sign change of dist - chi between rows is refined by 24 bisections with fresh evaluations.
"""

from __future__ import annotations

import math

import numpy as np

FRAME_GRID = 420
FRAME_BISECT = 24
BUDGET_GRID_CAP = 4000


def _get(ax, k):
    return np.asarray(ax[k] if isinstance(ax, dict) else getattr(ax, k), dtype=np.float64)


def _comov(axes, p, z):
    """Comoving position(s) of path p at z (scalar or 1-D array) -> tuple per axis."""
    z = np.asarray(z, dtype=np.float64)
    s = np.sin(z)
    out = []
    for ax in axes:
        a, b, f = _get(ax, "a")[p], _get(ax, "b")[p], _get(ax, "f")[p]
        a2 = _get(ax, "a2")[p]
        terms = a * (np.sin(np.multiply.outer(z, b) + f) - np.sin(f))
        out.append((a2 * s + terms.sum(axis=-1)) / s)
    return tuple(out)


def reach(z, z_obs, b_max):
    s, so = math.sin(z), math.sin(z_obs)
    b = np.arange(2, b_max + 1)
    da = (np.cos(b * z) - 1.0) / (b * s) - (np.cos(b * z_obs) - 1.0) / (b * so)
    db = np.sin(b * z) / (b * s) - np.sin(b * z_obs) / (b * so)
    return float(np.where(b % 2 == 0, np.hypot(da, db), np.abs(db)).max())


def bake(axes, z_obs, z_min, *, observer_path=None, observer_point=None, front="fit", cheb=True,
         speed=0.33, wraps=None):
    """Hits as (src, dx, dy, chi, z_emit) tuples, unordered."""
    n = len(_get(axes[0], "a2"))
    b_max = max(2, int(max(_get(ax, "b").max() for ax in axes)))
    eta_obs = math.log(math.tan(z_obs / 2.0))
    budget = front == "budget"
    cheb = True if budget else cheb
    spd = max(1e-3, speed) if front == "speed" else -1.0 / math.log(math.tan(z_min / 2.0))

    def dist(dx, dy):
        return np.maximum(np.abs(dx), np.abs(dy)) if cheb else np.hypot(dx, dy)

    def z_at_chi(chi):
        return 2.0 * math.atan(math.exp(eta_obs - chi / spd))

    m = min(BUDGET_GRID_CAP, max(FRAME_GRID, 2 * b_max)) if budget else FRAME_GRID
    if budget:
        zg = z_obs - (z_obs - z_min) * np.arange(m + 1) / m
        chig = np.array([reach(float(z), z_obs, b_max) for z in zg])
        chi_max = float(chig.max())
    else:
        chi_max = spd * (eta_obs - math.log(math.tan(z_min / 2.0)))
        chig = chi_max * np.arange(m + 1) / m
        zg = np.array([z_at_chi(float(c)) for c in chig])
    if observer_path is not None:
        ox, oy = (float(v) for v in _comov(axes, observer_path, z_obs)[:2])
        obs_idx = observer_path
    else:
        ox, oy = observer_point
        obs_idx = -1
    k = int(math.ceil((chi_max + 4.0) / 2.0)) if wraps is None else wraps

    hits = []
    for p in range(n):
        px, py = _comov(axes, p, zg)[:2]
        for nx in range(-k, k + 1):
            for ny in range(-k, k + 1):
                bx, by = 2.0 * nx - ox, 2.0 * ny - oy
                dmin = max(0.0, 2.0 * max(abs(nx), abs(ny)) - 4.0)
                if dmin >= chi_max:
                    continue
                i0 = 1 if budget else max(1, int(math.ceil(dmin * m / chi_max)))
                if i0 > m:
                    continue
                g = dist(px + bx, py + by) - chig
                prev0 = float(g[i0 - 1])
                if p == obs_idx and nx == 0 and ny == 0 and i0 == 1 and abs(prev0) < 1e-9:
                    prev0 = -1e-12
                elif prev0 == 0.0:
                    prev0 = 1e-12
                prevs = np.concatenate(([prev0], g[i0:m]))
                cur = g[i0:m + 1]
                for j in np.nonzero((prevs < 0.0) != (cur < 0.0))[0]:
                    i = i0 + int(j)
                    inside_prev = prevs[j] < 0.0
                    if budget:
                        lo, hi = float(zg[i - 1]), float(zg[i])
                        for _ in range(FRAME_BISECT):
                            mid = 0.5 * (lo + hi)
                            qx, qy = (float(v) for v in _comov(axes, p, mid)[:2])
                            gm = float(dist(qx + bx, qy + by)) - reach(mid, z_obs, b_max)
                            if (gm < 0.0) == inside_prev:
                                lo = mid
                            else:
                                hi = mid
                        z_emit = 0.5 * (lo + hi)
                        chi = reach(z_emit, z_obs, b_max)
                    else:
                        lo, hi = float(chig[i - 1]), float(chig[i])
                        for _ in range(FRAME_BISECT):
                            mid = 0.5 * (lo + hi)
                            qx, qy = (float(v) for v in _comov(axes, p, z_at_chi(mid))[:2])
                            gm = float(dist(qx + bx, qy + by)) - mid
                            if (gm < 0.0) == inside_prev:
                                lo = mid
                            else:
                                hi = mid
                        chi = 0.5 * (lo + hi)
                        z_emit = z_at_chi(chi)
                    qx, qy = (float(v) for v in _comov(axes, p, z_emit)[:2])
                    hits.append((p, qx + bx, qy + by, chi, z_emit))
    return hits
