"""ug_frame_bake_slices_noalloc (no allocation; the frame basis in a fixed device array)
against ug_frame_bake_slices, both called through the C ABI: the same positions and
rasters bit for bit, on two calls in a row, the second with a smaller input."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _slices(fn_name, paths, z, px):
    import torch

    from ugpack.abi import lib

    z = np.ascontiguousarray(z, np.float64)
    # outputs start as garbage: the entry point must write every position and clear the raster
    pos = torch.full((len(z), paths.n, paths.dim), float("nan"), dtype=torch.float32, device="cuda")
    ras = torch.full((len(z), px, px), 7.0, dtype=torch.float32, device="cuda")
    rc = getattr(lib(), fn_name)(paths._h, z.ctypes.data, len(z), pos.data_ptr(), ras.data_ptr(), px)
    assert rc == 0, lib().ug_last_error().decode()
    torch.cuda.synchronize()
    return pos.cpu().numpy(), ras.cpu().numpy()


def test_noalloc_slices_equal_the_allocating_entry_point(gpu):
    import ugpack
    from ugpack import frames

    # first call: 3+1 T=64 (largest frequency 64), 600 frames (two basis batches) at 256^2;
    # second: 2+1 T=16, 7 frames at 128^2 -- stale scratch contents of the first would show
    cases = []
    for d, T, seed, nz, px in ((3, 64, 2, 600, 256), (2, 16, 1, 7, 128)):
        with ugpack.Packer(d, T, seed=seed, max_attempts=30000) as p:
            p.run()
            paths = frames.Paths.from_packer(p)
        cases.append((paths, (np.arange(nz) + 0.5) * math.pi / nz, px))
    got = [_slices("ug_frame_bake_slices_noalloc", *c) for c in cases]
    want = [_slices("ug_frame_bake_slices", *c) for c in cases]
    for (paths, z, px), (gp, gr), (wp, wr) in zip(cases, got, want):
        assert paths.n > 10
        assert np.array_equal(gp, wp) and np.array_equal(gr, wr)
        assert np.all(gr.sum(axis=(1, 2)) == paths.n)
        paths.close()


def test_noalloc_rejects_frequencies_beyond_its_scratch(gpu):
    import torch

    from ugpack import frames
    from ugpack.abi import UG_EOVERFLOW, lib

    a2 = np.zeros((2, 2))
    amp = np.full((2, 2, 2), 1e-3)
    freq = np.array([[[2.0, 5000.0], [3.0, 4.0]], [[2.0, 6.0], [3.0, 4.0]]])
    paths = frames.Paths(a2, amp, freq)
    z = np.array([0.5, 1.5])
    pos = torch.empty((2, 2, 2), dtype=torch.float32, device="cuda")
    rc = lib().ug_frame_bake_slices_noalloc(paths._h, z.ctypes.data, 2, pos.data_ptr(), None, 0)
    assert rc == UG_EOVERFLOW and b"scratch" in lib().ug_last_error()
    assert lib().ug_frame_bake_slices(paths._h, z.ctypes.data, 2, pos.data_ptr(), None, 0) == 0
    paths.close()
