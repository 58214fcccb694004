"""GPU frame baker against the reference's bake_reference hit lists (golden fixtures) and
the CPU oracle; BRF1 file layout; equal-time slices and rasters.

Tolerance: hits are float32 records (cuda/frame_bake.cu:261-264); the GPU evaluates the
worldlines with a rotation recurrence instead of per-term libm sin, so values agree to
float32 resolution: atol 2e-6 on dx, dy, chi, z_emit, identical hit counts and sources."""

from __future__ import annotations

import math
import os
import struct

import numpy as np
import pytest

import bake_oracle
from test_frames import CASES, load_case, match_hits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CASES)
def test_gpu_bake_equals_reference_hits(gpu, name):
    from ugpack import frames

    print(f"frames {name}", flush=True)
    g, axes, kw = load_case(name)
    got = frames.bake(axes, g["z_obs"], g["z_min"], **kw)
    match_hits(got, [tuple(h) for h in g["hits"]], 2e-6)


def test_gpu_bake_matches_oracle_on_a_packing(gpu):
    """Worldlines of a real packing (2+1, T=32, 60 paths), fit and budget fronts."""
    import ugpack
    from ugpack import frames

    with ugpack.Packer(2, 32, seed=3, max_n=60) as p:
        p.run()
        paths = frames.Paths.from_packer(p)
        par = p.params()
    axes = [{"a": par[:, a, 1::3], "b": par[:, a, 2::3], "f": par[:, a, 3::3], "a2": par[:, a, 0]}
            for a in range(2)]
    for front, kw in (("fit", dict(observer_path=5)), ("budget", dict(observer_point=(0.2, -0.4)))):
        got = frames.bake(paths, 0.55 * math.pi, 0.1 * math.pi, front=front, **kw)
        want = bake_oracle.bake(axes, 0.55 * math.pi, 0.1 * math.pi, front=front, **kw)
        assert len(want) > 20
        match_hits(got, want, 2e-6)


def test_gpu_hits_are_ordered_deterministically(gpu):
    from ugpack import frames

    g, axes, kw = load_case("demo50_fit_path0")
    a = frames.bake(axes, g["z_obs"], g["z_min"], **kw)
    b = frames.bake(axes, g["z_obs"], g["z_min"], **kw)
    assert a == b
    assert [h[0] for h in a] == sorted(h[0] for h in a)


def test_bake_file_writes_brf1(gpu, tmp_path):
    from ugpack import frames

    dump = tmp_path / "d.csv"
    from test_frames import GOLD  # noqa: F401
    dump.write_text(
        "ax2,ay2,ax_1,bx_1,fx_1,ay_1,by_1,fy_1,ax_2,bx_2,fx_2,ay_2,by_2,fy_2,gid\n"
        "0.31,-0.62,0.20,3,0,-0.13,2,0.8,0.10,4,1.1,0.24,5,0,0\n"
        "-0.55,0.12,-0.33,2,0.4,0.25,4,2.1,0.08,5,0,-0.15,3,0,1\n"
        "0.05,0.77,0.5,2,2.2,-0.5,2,0.3,0.0,3,0,0.0,4,0,1\n")
    out = tmp_path / "f.frames"
    n = frames.bake_file(str(dump), str(out), [("path", 0), ("point", (0.1, 0.3))],
                         from_pi=0.3, to_pi=0.6, frames=3, front="fit")
    raw = out.read_bytes()
    assert raw[:4] == b"BRF1"
    n_paths, n_obs, n_inst, front, cheb = struct.unpack_from("<5I", raw, 4)
    assert (n_paths, n_obs, n_inst, front, cheb) == (3, 2, 3, 1, 1)
    fs = frames.load_frames(str(out))
    assert len(fs["frames"]) == 3 and sum(len(o["hits"]) for r in fs["frames"] for o in r) == n
    assert fs["observers"][1][:2] == (1, -1)
    # every instant matches a direct bake with the same settings
    z0 = fs["frames"][0][0]["z_obs"]
    assert z0 == pytest.approx(0.3 * math.pi)
    hits = frames.bake(frames.Paths.from_dump(str(dump)), z0, 0.02 * math.pi, observer_path=0)
    assert len(hits) == len(fs["frames"][0][0]["hits"])
    # file size equals the layout's sum
    size = 4 + 20 + 16 + 4 + 24 * 2 + 3 * (20 + 2 * 40) + 20 * n
    assert len(raw) == size


@pytest.mark.parametrize("d,T,seed,nz", [(3, 24, 2, 16), (2, 32, 1, 70), (3, 16, 3, 5)])
def test_slices_and_rasters(gpu, d, T, seed, nz):
    """FP64-MFMA slices against the oracle's per-term evaluation (frames beyond one 64-frame
    tile and fewer than one tile included)."""
    import torch
    import ugpack
    from ugpack import frames

    with ugpack.Packer(d, T, seed=seed, max_attempts=20000) as p:
        p.run()
        paths = frames.Paths.from_packer(p)
        par = p.params()
    n = par.shape[0]
    z = (np.arange(nz) + 0.5) * math.pi / nz
    pos, ras = frames.slices(paths, z, raster_px=256)
    pos, ras = pos.cpu().numpy(), ras.cpu().numpy()
    axes = [{"a": par[:, a, 1::3], "b": par[:, a, 2::3], "f": par[:, a, 3::3], "a2": par[:, a, 0]}
            for a in range(d)]
    for j, zj in enumerate(z):
        want = np.stack([np.asarray(bake_oracle._comov(axes, i, zj)) for i in range(n)])
        want = want - 2.0 * np.floor((want + 1.0) * 0.5)
        d = np.abs(pos[j] - want)
        assert np.minimum(d, 2.0 - d).max() < 1e-5
        assert ras[j].sum() == n
        # raster = histogram of the GPU's own positions with the same bin rule
        cx = np.clip(np.floor((pos[j][:, 0].astype(np.float64) + 1.0) * 128).astype(int), 0, 255)
        cy = np.clip(np.floor((pos[j][:, 1].astype(np.float64) + 1.0) * 128).astype(int), 0, 255)
        h = np.zeros((256, 256))
        np.add.at(h, (cy, cx), 1.0)
        assert np.abs(h - ras[j]).sum() <= 2 * max(1, int(0.002 * n))


@pytest.mark.parametrize("d,T,terms,seed", [(2, 32, 0, 3), (3, 16, 0, 1), (2, 24, 5, 2)])
def test_paths_from_pack_equals_dump_parameters(gpu, d, T, terms, seed):
    """The device-built worldline set equals the one built from the dump parameters."""
    import torch

    import ugpack
    from ugpack import frames

    with ugpack.Packer(d, T, terms=terms, seed=seed, max_attempts=20000) as p:
        p.run()
        dev = frames.Paths.from_packer(p)
        host = frames.Paths.from_params(p)
        z = np.linspace(0.05, math.pi - 0.05, 17)
        a, _ = frames.slices(dev, z, raster_px=64)
        b, _ = frames.slices(host, z, raster_px=64)
    assert dev.n == host.n > 5
    diff = (a - b).abs()
    diff = torch.minimum(diff, 2.0 - diff)
    assert float(diff.max()) < 1e-6
    if d == 2:
        got = frames.bake(dev, 0.5 * math.pi, 0.1 * math.pi, observer_path=1, front="fit")
        want = frames.bake(host, 0.5 * math.pi, 0.1 * math.pi, observer_path=1, front="fit")
        match_hits(got, want, 1e-5)


def test_config5_frames_equal_numpy_histograms(gpu):
    """3+1 F=256 packed to jamming (~3.8e5 worldlines), then three of the workload's 512
    instants z_j = (j + 1/2) pi / 512 at 2048^2: each raster is numpy.histogram2d of the
    worldlines' positions evaluated from their dump parameters with the reference's direct
    expression (braidlab/corrdim.py:116-137), and holds exactly N."""
    import torch

    import ugpack
    from ugpack import frames

    print("cfg5 pack + frames", flush=True)
    js = np.array([0, 255, 511])
    z = (js + 0.5) * math.pi / 512
    px = 2048
    with ugpack.Packer(3, 256, seed=1, precision=ugpack.PREC_SPLIT_BF16,
                       stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4) as p:
        st = p.run()
        n = st.n
        paths = frames.Paths.from_packer(p)
        pos, ras = frames.slices(paths, z, raster_px=px)
        paths.close()
        ras = ras.cpu().numpy()
        pos = pos.cpu().numpy()
        want = np.zeros((len(z), n, 3))
        sz = np.sin(z)
        for first in range(0, n, 16384):
            par = p.params(first, min(16384, n - first))          # (k, 3, 1 + 3 nw)
            a2, amp, frq, phs = par[:, :, 0], par[:, :, 1::3], par[:, :, 2::3], par[:, :, 3::3]
            for j, zj in enumerate(z):
                s = (amp * (np.sin(frq * zj + phs) - np.sin(phs))).sum(-1)
                want[j, first:first + len(par)] = (a2 * sz[j] + s) / sz[j]
    assert n > 300000
    want = want - 2.0 * np.floor((want + 1.0) * 0.5)
    for j in range(len(z)):
        d = np.abs(pos[j] - want[j])
        assert np.minimum(d, 2.0 - d).max() < 2e-6
        h, _, _ = np.histogram2d(want[j][:, 1], want[j][:, 0], bins=px, range=[[-1, 1], [-1, 1]])
        assert ras[j].sum() == n
        # positions within float32 rounding of a bin edge may land in the neighbour bin
        assert np.abs(h - ras[j]).sum() <= 2 * int(0.001 * n)
