"""Frame baker front end: the host side of cuda/frame_bake.cu over libugpack.so.

- :func:`load_dump_terms` reads a parameter dump by header names, like the engine's
  load_dump (cuda/frame_bake.cu:109-190) -- legacy single-wiggle and multi-term layouts,
  2+1 and 3+1, a trailing gid column ignored;
- :class:`Paths` is a device-resident worldline set (from a dump, from per-axis term arrays
  like braidlab.corrdim.AxisTerms, or from a packing);
- :func:`bake` mirrors braidlab.frames.bake_reference (braidlab/frames.py:246-257) and returns
  the same (src, dx, dy, chi, z_emit) tuples, computed on the GPU;
- :func:`bake_file` is the frame_bake command line (cuda/frame_bake.cu:398-635): it writes the
  BRF1 file that braidlab.frames.load_frames reads;
- :func:`slices` gives equal-time positions and density rasters.
"""

from __future__ import annotations

import csv
import ctypes as C
import math
import struct

import numpy as np

from .abi import Front, FrameInfo, Observer, UgError, check, lib

FRONTS = {"speed": 0, "fit": 1, "budget": 2}
HIT_DTYPE = np.dtype([("src", "<u4"), ("dx", "<f4"), ("dy", "<f4"), ("chi", "<f4"),
                      ("z_emit", "<f4")])


def load_dump_terms(path: str):
    """(dim, a2 (n, dim), amp/freq/phase (n, dim, nw)) from a parameter dump CSV."""
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    if not rows:
        raise ValueError(f"{path}: empty dump")
    fields = set(rows[0].keys())
    names = ("x", "y", "w") if ("aw_1" in fields or "aw" in fields) else ("x", "y")
    dim = len(names)

    def col(k):
        return np.array([float(r[k]) for r in rows]) if k in fields else np.zeros(len(rows))

    if "ax_1" in fields:
        nw = 1
        while f"ax_{nw + 1}" in fields:
            nw += 1
        amp = np.stack([np.stack([col(f"a{n}_{j}") for j in range(1, nw + 1)], -1) for n in names], 1)
        freq = np.stack([np.stack([col(f"b{n}_{j}") for j in range(1, nw + 1)], -1) for n in names], 1)
        phase = np.stack([np.stack([col(f"f{n}_{j}") for j in range(1, nw + 1)], -1) for n in names], 1)
    else:
        amp = np.stack([col(f"a{n}")[:, None] for n in names], 1)
        freq = np.stack([col(f"b{n}")[:, None] for n in names], 1)
        phase = np.stack([col(f"f{n}")[:, None] for n in names], 1)
    a2 = np.stack([col(f"a{n}2") for n in names], 1)
    return dim, a2, amp, freq, phase


class Paths:
    """A worldline set on the device (dense per-frequency coefficients)."""

    def __init__(self, a2, amp, freq, phase=None, stream: int = 0):
        a2 = np.ascontiguousarray(a2, np.float64)
        amp = np.ascontiguousarray(amp, np.float64)
        freq = np.ascontiguousarray(freq, np.float64)
        phase = np.zeros_like(amp) if phase is None else np.ascontiguousarray(phase, np.float64)
        n, dim, nw = amp.shape
        h = C.c_void_p()
        check(lib().ug_paths_create(dim, n, nw, a2.ctypes.data, amp.ctypes.data, freq.ctypes.data,
                                    phase.ctypes.data, C.c_void_p(stream), C.byref(h)))
        self._h, self.n, self.dim, self.nw = h, n, dim, nw

    @classmethod
    def from_dump(cls, path: str) -> "Paths":
        _, a2, amp, freq, phase = load_dump_terms(path)
        return cls(a2, amp, freq, phase)

    @classmethod
    def from_axes(cls, axes) -> "Paths":
        """From per-axis term arrays a/b/f (n, nw) and a2 (n,): braidlab AxisTerms objects
        (braidlab/corrdim.py:49-63) or dicts with those keys."""

        def get(ax, k):
            return np.asarray(ax[k] if isinstance(ax, dict) else getattr(ax, k), float)

        a2 = np.stack([get(ax, "a2") for ax in axes], 1)
        amp = np.stack([get(ax, "a") for ax in axes], 1)
        freq = np.stack([get(ax, "b") for ax in axes], 1)
        phase = np.stack([get(ax, "f") for ax in axes], 1)
        return cls(a2, amp, freq, phase)

    @classmethod
    def from_packer(cls, packer, stream: int = 0) -> "Paths":
        """The accepted worldlines of a ugpack.Packer, in admission order, built on the
        device from the accepted candidate indices (no host copy of the parameters)."""
        h = C.c_void_p()
        check(lib().ug_paths_from_pack(packer._h, C.c_void_p(stream), C.byref(h)))
        self = cls.__new__(cls)
        st = packer.status()
        self._h, self.n, self.dim, self.nw = h, st.n + st.nsub, packer.dim, packer.nw
        return self

    @classmethod
    def from_params(cls, packer) -> "Paths":
        """The same set through the host dump parameters (ug_pack_params)."""
        par = packer.params()  # (n, dim, 1 + 3 nw)
        return cls(par[:, :, 0], par[:, :, 1::3], par[:, :, 2::3], par[:, :, 3::3])

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().ug_paths_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _observer(observer_path, observer_point) -> Observer:
    if observer_path is not None:
        return Observer(0, int(observer_path), 0.0, 0.0)
    if observer_point is None:
        raise ValueError("need observer_path or observer_point")
    return Observer(1, -1, float(observer_point[0]), float(observer_point[1]))


def bake_frame(paths: Paths, z_obs: float, z_min: float, observer: Observer, front: str = "fit",
               cheb: bool = True, speed: float = 0.33, wraps: int | None = None):
    """Hits (HIT_DTYPE array, (path, image, crossing) order) and the frame header fields."""
    import torch

    if front not in FRONTS:
        raise ValueError("front must be speed|fit|budget")
    fr = Front(FRONTS[front], int(cheb), float(speed), float(z_min), -1 if wraps is None else int(wraps), 0)
    info = FrameInfo()
    rc = lib().ug_frame_bake_lightcone(paths._h, C.byref(fr), float(z_obs), C.byref(observer), None, 0,
                                       C.byref(info))
    if rc not in (0, -4):
        check(rc)
    n = int(info.n_hits)
    buf = torch.empty(max(n, 1) * HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    if n:
        check(lib().ug_frame_bake_lightcone(paths._h, C.byref(fr), float(z_obs), C.byref(observer),
                                            buf.data_ptr(), n, C.byref(info)))
    hits = np.frombuffer(buf[: n * HIT_DTYPE.itemsize].cpu().numpy().tobytes(), dtype=HIT_DTYPE)
    return hits, info


def bake(axes, z_obs: float, z_min: float, *, observer_path: int | None = None,
         observer_point: tuple[float, float] | None = None, front: str = "fit", cheb: bool = True,
         speed: float = 0.33, wraps: int | None = None):
    """GPU counterpart of braidlab.frames.bake_reference: list of (src, dx, dy, chi, z_emit)."""
    paths = axes if isinstance(axes, Paths) else Paths.from_axes(axes)
    hits, _ = bake_frame(paths, z_obs, z_min, _observer(observer_path, observer_point), front,
                         cheb, speed, wraps)
    return [(int(h["src"]), float(h["dx"]), float(h["dy"]), float(h["chi"]), float(h["z_emit"]))
            for h in hits]


def bake_file(dump: str, out: str, observers, from_pi: float = 0.25, to_pi: float = 0.75,
              frames: int = 1, front: str = "fit", metric: str = "square", speed: float = 0.33,
              zmin_pi: float = 0.02, wraps: int = -1) -> int:
    """The frame_bake command (cuda/frame_bake.cu:398-635): writes a BRF1 file.

    observers: list of ("path", idx) or ("point", (x, y)).  Returns the total hit count."""
    if front not in FRONTS:
        raise ValueError("front must be speed|fit|budget")
    paths = Paths.from_dump(dump)
    if paths.dim != 2:
        raise ValueError(f"{dump} is a 3+1 dump; the light-cone baker is 2+1 only")
    obs = []
    for kind, v in observers:
        if kind == "path":
            if not (0 <= int(v) < paths.n):
                raise ValueError(f"observer path {v} out of range")
            obs.append(Observer(0, int(v), 0.0, 0.0))
        else:
            obs.append(Observer(1, -1, float(v[0]), float(v[1])))
    z_min = max(2e-3, zmin_pi * math.pi)
    cheb = front == "budget" or metric == "square"
    total = 0
    with open(out, "wb") as f:
        f.write(b"BRF1" + struct.pack("<5I", paths.n, len(obs), frames, FRONTS[front], int(cheb)))
        f.write(struct.pack("<2d", speed, z_min) + struct.pack("<i", wraps))
        for o in obs:
            f.write(struct.pack("<Ii2d", o.type, o.path if o.type == 0 else -1, o.px, o.py))
        for fi in range(frames):
            frac = 0.0 if frames == 1 else fi / (frames - 1)
            z_obs = min((from_pi + (to_pi - from_pi) * frac) * math.pi, math.pi - 1e-3)
            if z_min >= z_obs:
                raise ValueError(f"rewind floor {z_min:.4f} >= z_obs {z_obs:.4f}")
            head = None
            for o in obs:
                hits, info = bake_frame(paths, z_obs, z_min, o, front, cheb, speed,
                                        None if wraps < 0 else wraps)
                if head is None:
                    head = struct.pack("<2d", info.z_obs, info.chi_max) + struct.pack("<i", info.K)
                    f.write(head)
                f.write(struct.pack("<4d", info.ox, info.oy, info.betax, info.betay))
                f.write(struct.pack("<Q", len(hits)) + hits.tobytes())
                total += len(hits)
    paths.close()
    return total


def load_frames(path: str) -> dict:
    """Minimal BRF1 reader (the layout braidlab/frames.py:88-144 reads)."""
    raw = open(path, "rb").read()
    if raw[:4] != b"BRF1":
        raise ValueError("not a BRF1 file")
    off = 4
    n_paths, n_obs, n_inst, front, cheb = struct.unpack_from("<5I", raw, off)
    off += 20
    speed, z_min = struct.unpack_from("<2d", raw, off)
    off += 16
    (wraps,) = struct.unpack_from("<i", raw, off)
    off += 4
    observers = []
    for _ in range(n_obs):
        t, pidx, px, py = struct.unpack_from("<Ii2d", raw, off)
        off += 24
        observers.append((t, pidx, px, py))
    frames = []
    for _ in range(n_inst):
        z_obs, chi_max = struct.unpack_from("<2d", raw, off)
        off += 16
        (k,) = struct.unpack_from("<i", raw, off)
        off += 4
        row = []
        for _ in range(n_obs):
            ox, oy, bx, by = struct.unpack_from("<4d", raw, off)
            off += 32
            (nh,) = struct.unpack_from("<Q", raw, off)
            off += 8
            hits = np.frombuffer(raw, dtype=HIT_DTYPE, count=nh, offset=off)
            off += nh * HIT_DTYPE.itemsize
            row.append(dict(z_obs=z_obs, chi_max=chi_max, K=k, ox=ox, oy=oy, beta=(bx, by), hits=hits))
        frames.append(row)
    return dict(n_paths=n_paths, front=front, cheb=cheb, speed=speed, z_min=z_min, wraps=wraps,
                observers=observers, frames=frames)


def slices(paths: Paths, z, raster_px: int = 0, want_pos: bool = True):
    """Wrapped positions (nz, n, dim) float32 and/or (nz, px, px) density rasters (torch)."""
    import torch

    z = np.ascontiguousarray(z, np.float64)
    nz = len(z)
    pos = torch.empty((nz, paths.n, paths.dim), dtype=torch.float32, device="cuda") if want_pos else None
    ras = (torch.empty((nz, raster_px, raster_px), dtype=torch.float32, device="cuda")
           if raster_px else None)
    check(lib().ug_frame_bake_slices(paths._h, z.ctypes.data, nz, pos.data_ptr() if want_pos else None,
                                     ras.data_ptr() if raster_px else None, raster_px))
    return pos, ras


__all__ = ["Paths", "bake", "bake_file", "bake_frame", "load_dump_terms", "load_frames", "slices",
           "HIT_DTYPE", "UgError"]
