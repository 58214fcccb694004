"""ctypes binding of libugpack.so (the C ABI declared in include/ugpack.h).

This is the binding a maintainer of the reference would add in place of spawning the
engine binaries from braidlab/engine.py:50-112.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._build import LIB_PATH

UG_OK, UG_EINVAL, UG_ENOMEM, UG_EHIP, UG_EOVERFLOW, UG_ESTATE = 0, -1, -2, -3, -4, -5
PREC_FP64, PREC_SPLIT_BF16 = 0, 1
STOP_BUDGET, STOP_FIXED_WINDOW, STOP_REF_WINDOW = 0, 1, 2
RNG_PHILOX, RNG_XOSHIRO_REF = 0, 1
ABI_VERSION = 5


class UgError(RuntimeError):
    """A libugpack call returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class PackConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("T", C.c_int32), ("terms", C.c_int32), ("precision", C.c_int32),
        ("seed", C.c_uint64), ("pin_sin1", C.c_int32), ("stop_kind", C.c_int32),
        ("max_attempts", C.c_double), ("stop_rate", C.c_double), ("stop_batch", C.c_int64),
        ("stop_window", C.c_int64), ("max_n", C.c_int64), ("margin", C.c_double),
        ("max_paths", C.c_int64), ("euclid", C.c_int32), ("sparse", C.c_int32),
        ("sub_target", C.c_int32), ("sub_max", C.c_int32), ("rng_kind", C.c_int32),
    ]


class PackStatus(C.Structure):
    _fields_ = [
        ("attempts", C.c_int64), ("n", C.c_int64), ("nsub", C.c_int64), ("done", C.c_int32),
        ("sub_done", C.c_int32), ("survivors", C.c_int64), ("rechecked", C.c_int64),
        ("margin", C.c_double), ("seconds", C.c_double), ("filled", C.c_int64),
        ("sub_attempts", C.c_int64),
    ]


class SubConfig(C.Structure):
    _fields_ = [("sub_terms", C.c_int32), ("reserved", C.c_int32), ("max_attempts", C.c_double),
                ("accept_rate", C.c_double), ("fill_rate", C.c_double)]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("items", C.c_int64), ("ms", C.c_double),
                ("visits", C.c_int64), ("points", C.c_int64), ("left_wsum", C.c_int64),
                ("left_possible", C.c_int64), ("left_none", C.c_int64)]


class KnotConfig(C.Structure):
    _fields_ = [("dim", C.c_int32), ("T", C.c_int32), ("terms", C.c_int32), ("pin_sin1", C.c_int32),
                ("seed", C.c_uint64), ("max_attempts", C.c_double), ("nu_attempts", C.c_double),
                ("nu_div", C.c_double), ("max_strands", C.c_int32), ("precision", C.c_int32),
                ("rng_kind", C.c_int32)]


class KnotStatus(C.Structure):
    _fields_ = [("attempts", C.c_int64), ("n", C.c_int64), ("n_primary", C.c_int64),
                ("primary_attempts", C.c_int64), ("phase", C.c_int32), ("done", C.c_int32),
                ("seconds", C.c_double)]


KERNEL_PREFILTER, KERNEL_RECHECK = 0, 1
PROFILE_PREFILTER, PROFILE_RECHECK = 1 << KERNEL_PREFILTER, 1 << KERNEL_RECHECK
PROFILE_ALL = PROFILE_PREFILTER | PROFILE_RECHECK
PROFILE_REASONS = 1 << 2  # also count why the prefilter leaves candidates to the FP64 test
SUB_BUDGET, SUB_ACCEPT_RATE, SUB_FILL_RATE = 1, 2, 3


class Front(C.Structure):
    _fields_ = [("kind", C.c_int32), ("cheb", C.c_int32), ("speed", C.c_double),
                ("z_min", C.c_double), ("wraps", C.c_int32), ("reserved", C.c_int32)]


class Observer(C.Structure):
    _fields_ = [("type", C.c_int32), ("path", C.c_int32), ("px", C.c_double), ("py", C.c_double)]


class FrameInfo(C.Structure):
    _fields_ = [("z_obs", C.c_double), ("chi_max", C.c_double), ("K", C.c_int32),
                ("reserved", C.c_int32), ("ox", C.c_double), ("oy", C.c_double),
                ("betax", C.c_double), ("betay", C.c_double), ("n_hits", C.c_uint64)]

_lib = None


def lib() -> C.CDLL:
    """Load libugpack.so once; raise if it is missing (build with ugpack.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libugpack.so not built ({LIB_PATH}); run ugpack.build() first")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    sigs = {
        "ug_abi_version": ([], i32),
        "ug_last_error": ([], C.c_char_p),
        "ug_pack_create": ([C.POINTER(PackConfig), vp, C.POINTER(vp)], i32),
        "ug_pack_run": ([vp, i64, C.POINTER(PackStatus)], i32),
        "ug_pack_run_sub": ([vp, C.POINTER(SubConfig), i64, C.POINTER(PackStatus)], i32),
        "ug_pack_get_status": ([vp, C.POINTER(PackStatus)], i32),
        "ug_pack_accepted": ([vp, vp, vp, i64, C.POINTER(i64)], i32),
        "ug_pack_params": ([vp, i64, i64, vp], i32),
        "ug_pack_positions": ([vp, vp, i64], i32),
        "ug_pack_curve": ([vp, vp, i64, C.POINTER(i64)], i32),
        "ug_pack_rounds": ([vp, vp, i64, C.POINTER(i64)], i32),
        "ug_eval_candidates": ([vp, vp, i64, i32, vp], i32),
        "ug_pack_profile": ([vp, i32], i32),
        "ug_pack_kernel_stats": ([vp, i32, C.POINTER(KernelStats)], i32),
        "ug_pack_destroy": ([vp], None),
        "ug_cache_trim": ([], i32),
        "ug_pack_check": ([vp, C.c_uint64, i64, vp], i32),
        "ug_ensemble_run": ([C.POINTER(PackConfig), i32, i32, C.POINTER(vp), C.POINTER(PackStatus), vp],
                            i32),
        "ug_paths_create": ([i32, i64, i32, vp, vp, vp, vp, vp, C.POINTER(vp)], i32),
        "ug_paths_destroy": ([vp], None),
        "ug_paths_from_pack": ([vp, vp, C.POINTER(vp)], i32),
        "ug_frame_bake_lightcone": ([vp, C.POINTER(Front), C.c_double, C.POINTER(Observer), vp,
                                     C.c_uint64, C.POINTER(FrameInfo)], i32),
        "ug_frame_bake_slices": ([vp, vp, i32, vp, vp, i32], i32),
        "ug_knot_create": ([C.POINTER(KnotConfig), vp, C.POINTER(vp)], i32),
        "ug_knot_run": ([vp, i64, C.POINTER(KnotStatus)], i32),
        "ug_knot_get_status": ([vp, C.POINTER(KnotStatus)], i32),
        "ug_knot_strands": ([vp, vp, vp, i64, C.POINTER(i64)], i32),
        "ug_knot_params": ([vp, i64, i64, vp], i32),
        "ug_knot_batches": ([vp, vp, i64, C.POINTER(i64)], i32),
        "ug_knot_destroy": ([vp], None),
        "ug_frame_bake_slices_noalloc": ([vp, vp, i32, vp, vp, i32], i32),
        "ug_split_generation_error": ([vp, vp], i32),
        "ug_ensemble_run_batched": ([C.POINTER(PackConfig), i32, i32, i32, C.POINTER(vp), C.POINTER(PackStatus),
                                     vp], i32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.ug_abi_version() != ABI_VERSION:
        raise RuntimeError("libugpack ABI version mismatch")
    _lib = L
    return L


def cache_trim() -> None:
    """Free the cached device buffers of destroyed handles (model tables stay)."""
    check(lib().ug_cache_trim())


def split_generation_error() -> tuple[float, float]:
    """Exhaustive accuracy of the split-BF16 prefilter's FP32 generation on the current device:
    (max |(cos, sin) of a phase - FP64| over all 2^32 phase words,
     max |exponential spacing - FP64| / (1 + value) over all 2^31 Dirichlet words)."""
    import torch

    out = (C.c_double * 2)()
    check(lib().ug_split_generation_error(C.c_void_p(torch.cuda.current_stream().cuda_stream), out))
    return float(out[0]), float(out[1])


def exported_symbols() -> list[str]:
    return [
        "ug_abi_version", "ug_last_error", "ug_pack_create", "ug_pack_run", "ug_pack_run_sub",
        "ug_pack_get_status", "ug_pack_accepted", "ug_pack_params", "ug_pack_positions",
        "ug_pack_curve", "ug_pack_rounds", "ug_eval_candidates", "ug_pack_profile", "ug_pack_kernel_stats",
        "ug_pack_destroy", "ug_cache_trim", "ug_pack_check", "ug_ensemble_run", "ug_paths_create",
        "ug_paths_destroy", "ug_paths_from_pack",
        "ug_frame_bake_lightcone",
        "ug_frame_bake_slices", "ug_frame_bake_slices_noalloc", "ug_knot_create", "ug_knot_run",
        "ug_knot_get_status", "ug_knot_strands", "ug_knot_params", "ug_knot_batches", "ug_knot_destroy",
        "ug_split_generation_error", "ug_ensemble_run_batched",
    ]


def check(rc: int) -> None:
    if rc != UG_OK:
        raise UgError(rc, lib().ug_last_error().decode())


@dataclass
class Status:
    attempts: int
    n: int
    nsub: int
    done: bool
    survivors: int
    rechecked: int
    margin: float
    seconds: float
    sub_done: int = 0       # 0 running, else SUB_BUDGET / SUB_ACCEPT_RATE / SUB_FILL_RATE
    filled: int = 0
    sub_attempts: int = 0

    @staticmethod
    def of(st: PackStatus) -> "Status":
        return Status(st.attempts, st.n, st.nsub, bool(st.done), st.survivors, st.rechecked,
                      st.margin, st.seconds, st.sub_done, st.filled, st.sub_attempts)


class Packer:
    """One packing run on the current HIP device (a handle of the C ABI)."""

    def __init__(self, dim: int, T: int, terms: int = 0, seed: int = 1,
                 precision: int = PREC_FP64, pin_sin1: bool = False,
                 stop_kind: int = STOP_BUDGET, max_attempts: float = 0.0,
                 stop_rate: float = 0.0, stop_batch: int = 1 << 20, stop_window: int = 1 << 22,
                 max_n: int = 0, margin: float = 0.0, max_paths: int = 0, stream: int = 0,
                 euclid: bool = False, sparse: bool = False, sub_target: int = 0, sub_max: int = 0,
                 rng_kind: int = RNG_PHILOX):
        self._h = None
        cfg = PackConfig(dim, T, terms, precision, seed, int(pin_sin1), stop_kind,
                         float(max_attempts), float(stop_rate), stop_batch, stop_window, max_n,
                         float(margin), max_paths, int(euclid), int(sparse), sub_target, sub_max,
                         rng_kind)
        h = C.c_void_p()
        check(lib().ug_pack_create(C.byref(cfg), C.c_void_p(stream), C.byref(h)))
        self._h = h
        self.dim, self.T = dim, T
        self.terms = terms if terms else T
        self.nw = self.terms - 1

    @classmethod
    def _adopt(cls, h, dim: int, T: int, terms: int) -> "Packer":
        """Wrap a handle the library created (ensemble runs); the wrapper destroys it."""
        self = cls.__new__(cls)
        self._h = h
        self.dim, self.T = dim, T
        self.terms = terms if terms else T
        self.nw = self.terms - 1
        return self

    def close(self) -> None:
        if self._h is not None:
            lib().ug_pack_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, max_candidates: int = 0) -> Status:
        st = PackStatus()
        check(lib().ug_pack_run(self._h, max_candidates, C.byref(st)))
        return Status.of(st)

    def run_sub(self, max_attempts: float = 0.0, accept_rate: float = 0.0, fill_rate: float = 0.0,
                max_candidates: int = 0, sub_terms: int = 0) -> Status:
        """Subpath phase (--subpaths): candidates continue the index stream."""
        sc = SubConfig(sub_terms, 0, float(max_attempts), float(accept_rate), float(fill_rate))
        st = PackStatus()
        check(lib().ug_pack_run_sub(self._h, C.byref(sc), max_candidates, C.byref(st)))
        return Status.of(st)

    def check(self, first: int = 0, count: int = 0) -> dict:
        """Serial-RSA characterization on the device: collisions among stored uniques and
        rejected candidates of [first, first+count) without an earlier contact (both 0)."""
        out = np.zeros(4, np.int64)
        check(lib().ug_pack_check(self._h, first, count, out.ctypes.data))
        return {"colliding_uniques": int(out[0]), "unexplained_rejects": int(out[1]),
                "uniques": int(out[2]), "rejects": int(out[3])}

    def status(self) -> Status:
        st = PackStatus()
        check(lib().ug_pack_get_status(self._h, C.byref(st)))
        return Status.of(st)

    def accepted(self) -> tuple[np.ndarray, np.ndarray]:
        n = C.c_int64()
        check(lib().ug_pack_accepted(self._h, None, None, 0, C.byref(n)))
        idx = np.zeros(n.value, np.int64)
        gid = np.zeros(n.value, np.int32)
        if n.value:
            check(lib().ug_pack_accepted(self._h, idx.ctypes.data, gid.ctypes.data, n.value,
                                         C.byref(n)))
        return idx, gid

    def params(self, first: int = 0, count: int | None = None) -> np.ndarray:
        """(count, dim, 1 + 3 nw): per axis a1 then (amplitude, frequency, phase) per slot."""
        if count is None:
            count = self.status().n + self.status().nsub - first
        out = np.zeros((count, self.dim, 1 + 3 * self.nw))
        check(lib().ug_pack_params(self._h, first, count, out.ctypes.data))
        return out

    def positions(self):
        """Canonical positions of the accepted paths as a torch float64 tensor (n, T, dim)."""
        import torch

        st = self.status()
        n = st.n + st.nsub
        out = torch.empty((max(n, 1), self.T, self.dim), dtype=torch.float64, device="cuda")
        check(lib().ug_pack_positions(self._h, out.data_ptr(), out.shape[0]))
        return out[:n]

    def profile(self, enable: bool = True, kernels: int = PROFILE_ALL) -> None:
        """Count visits and points in every candidate-test kernel and time the launches of
        `kernels` (PROFILE_* bits) with HIP events; enable=False turns both off."""
        check(lib().ug_pack_profile(self._h, kernels if enable else 0))

    def kernel_stats(self, which: int = KERNEL_PREFILTER) -> dict:
        ks = KernelStats()
        check(lib().ug_pack_kernel_stats(self._h, which, C.byref(ks)))
        return {"launches": ks.launches, "items": ks.items, "ms": ks.ms, "visits": ks.visits,
                "points": ks.points, "left_wsum": ks.left_wsum, "left_possible": ks.left_possible,
                "left_none": ks.left_none}

    def curve(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().ug_pack_curve(self._h, None, 0, C.byref(n)))
        rows = np.zeros((n.value, 4), np.int64)
        if n.value:
            check(lib().ug_pack_curve(self._h, rows.ctypes.data, n.value, C.byref(n)))
        return rows

    def rounds(self) -> np.ndarray:
        """Rounds of the reference schedule: (rows, 5) = first candidate, candidates,
        survivors at the round's start, admitted, newly occupied cells."""
        n = C.c_int64()
        check(lib().ug_pack_rounds(self._h, None, 0, C.byref(n)))
        rows = np.zeros((n.value, 5), np.int64)
        if n.value:
            check(lib().ug_pack_rounds(self._h, rows.ctypes.data, n.value, C.byref(n)))
        return rows

    def eval_candidates(self, idx, precision: int = PREC_FP64):
        """Positions (count, T, dim) float64 of candidate indices (torch cuda int64 tensor)."""
        import torch

        idx = idx.to(device="cuda", dtype=torch.int64).contiguous()
        out = torch.empty((idx.numel(), self.T, self.dim), dtype=torch.float64, device="cuda")
        check(lib().ug_eval_candidates(self._h, idx.data_ptr(), idx.numel(), precision,
                                       out.data_ptr()))
        return out


@dataclass
class KnotRun:
    attempts: int
    n: int
    n_primary: int
    primary_attempts: int
    phase: int
    done: bool
    seconds: float


class Knot:
    """The knot-regime packer (ug_knot_*): the primary species over [0, max_attempts), then
    the nu species over the next nu_attempts candidates at the exclusion (2/T)/nu_div."""

    def __init__(self, dim: int, T: int, terms: int = 0, seed: int = 1, pin_sin1: bool = False,
                 max_attempts: float = 1e10, nu_attempts: float = 0.0, nu_div: float = 1.0,
                 max_strands: int = 0, stream: int = 0, precision: int = PREC_FP64,
                 rng_kind: int = RNG_PHILOX):
        self._h = None
        cfg = KnotConfig(dim, T, terms, int(pin_sin1), seed, float(max_attempts), float(nu_attempts),
                         float(nu_div), max_strands, precision, rng_kind)
        h = C.c_void_p()
        check(lib().ug_knot_create(C.byref(cfg), C.c_void_p(stream), C.byref(h)))
        self._h = h
        self.dim, self.T = dim, T
        self.nw = (terms if terms else T) - 1

    def close(self) -> None:
        if self._h is not None:
            lib().ug_knot_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, max_candidates: int = 0) -> KnotRun:
        st = KnotStatus()
        check(lib().ug_knot_run(self._h, max_candidates, C.byref(st)))
        return KnotRun(st.attempts, st.n, st.n_primary, st.primary_attempts, st.phase, bool(st.done),
                       st.seconds)

    def strands(self) -> tuple[np.ndarray, np.ndarray]:
        n = C.c_int64()
        check(lib().ug_knot_strands(self._h, None, None, 0, C.byref(n)))
        idx = np.zeros(n.value, np.int64)
        sp = np.zeros(n.value, np.int32)
        if n.value:
            check(lib().ug_knot_strands(self._h, idx.ctypes.data, sp.ctypes.data, n.value, C.byref(n)))
        return idx, sp

    def batches(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().ug_knot_batches(self._h, None, 0, C.byref(n)))
        rows = np.zeros((n.value, 5), np.int64)
        if n.value:
            check(lib().ug_knot_batches(self._h, rows.ctypes.data, n.value, C.byref(n)))
        return rows

    def params(self, first: int = 0, count: int | None = None) -> np.ndarray:
        if count is None:
            count = len(self.strands()[0]) - first
        out = np.zeros((count, self.dim, 1 + 3 * self.nw))
        check(lib().ug_knot_params(self._h, first, count, out.ctypes.data))
        return out


def pack_config(dim: int, T: int, terms: int = 0, seed: int = 1, precision: int = PREC_FP64,
                pin_sin1: bool = False, stop_kind: int = STOP_BUDGET, max_attempts: float = 0.0,
                stop_rate: float = 0.0, stop_batch: int = 1 << 20, stop_window: int = 1 << 22,
                max_n: int = 0, margin: float = 0.0, max_paths: int = 0,
                euclid: bool = False, sparse: bool = False, sub_target: int = 0,
                sub_max: int = 0, rng_kind: int = RNG_PHILOX) -> PackConfig:
    return PackConfig(dim, T, terms, precision, seed, int(pin_sin1), stop_kind, float(max_attempts),
                      float(stop_rate), stop_batch, stop_window, max_n, float(margin), max_paths,
                      int(euclid), int(sparse), sub_target, sub_max, rng_kind)


def list_digest(idx) -> int:
    """FNV-1a (64-bit) of an accepted index list's int64 little-endian bytes (the digest
    ug_ensemble_run reports per universe)."""
    h = 1469598103934665603
    for byte in np.asarray(idx, "<i8").tobytes():
        h = ((h ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def ensemble(configs: list[PackConfig], threads: int = 16, keep: bool = False, digests: bool = False,
             batched: bool = False, window: int = 0, lanes: int = 1):
    """Pack independent universes: `threads` at a time on worker threads (ug_ensemble_run),
    or with batched=True on the batched engine, `lanes` lock-step lanes of at most `window`
    universes each (0 = 32, at most 64; ug_ensemble_run_batched).  Returns the list of final Status; with
    keep=True also the list of Packer handles, with digests=True also each universe's
    accepted-list digest (list_digest)."""
    n = len(configs)
    arr = (PackConfig * n)(*configs)
    st = (PackStatus * n)()
    handles = (C.c_void_p * n)() if keep else None
    dig = np.zeros(n, np.uint64)
    dptr = dig.ctypes.data if digests else None
    if batched:
        check(lib().ug_ensemble_run_batched(arr, n, window, lanes, handles, st, dptr))
    else:
        check(lib().ug_ensemble_run(arr, n, threads, handles, st, dptr))
    out = [[Status.of(st[i]) for i in range(n)]]
    if keep:
        out.append([Packer._adopt(C.c_void_p(handles[i]), c.dim, c.T, c.terms) for i, c in enumerate(configs)])
    if digests:
        out.append([int(v) for v in dig])
    return out[0] if len(out) == 1 else tuple(out)
