"""ctypes view of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Imported by tests/, the smoke check and bench.py's cpu_baseline leg as the
checker; never by the product path.  See ug_oracle.h for what each function
restates from the reference.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        L.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.or_philox4x32_10.restype = None
        L.or_sincospi_rat.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.or_sincospi_rat.restype = None
        L.or_log.argtypes = [C.c_double]
        L.or_log.restype = C.c_double
        L.or_model_init.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int]
        L.or_model_init.restype = C.c_int
        L.or_model_set_euclid.argtypes = [C.c_void_p, C.c_int]
        L.or_model_set_euclid.restype = None
        L.or_model_set_rng.argtypes = [C.c_void_p, C.c_int]
        L.or_model_set_rng.restype = None
        L.or_xoshiro_seed.argtypes = [C.c_uint64, C.c_void_p]
        L.or_xoshiro_seed.restype = None
        L.or_xoshiro_next.argtypes = [C.c_void_p]
        L.or_xoshiro_next.restype = C.c_uint64
        L.or_touches.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_touches.restype = C.c_int
        L.or_model_free.argtypes = [C.c_void_p]
        L.or_model_free.restype = None
        L.or_candidate.argtypes = [C.c_void_p, C.c_uint64] + [C.c_void_p] * 6
        L.or_candidate.restype = None
        L.or_eval_t.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 5
        L.or_eval_t.restype = None
        L.or_cell.argtypes = [C.c_void_p, C.c_double]
        L.or_cell.restype = C.c_int
        L.or_pack.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int64, C.c_int64, C.c_double,
                              C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                              C.POINTER(C.c_int64)]
        L.or_pack.restype = C.c_int64
        L.or_pack_mt.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int64, C.c_int64, C.c_double,
                                 C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        L.or_pack_mt.restype = C.c_int64
        L.or_pack_sub.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64,
                                  C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                  C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        L.or_pack_sub.restype = C.c_int64
        L.or_pack_ref.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int64,
                                  C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.c_void_p, C.c_int64,
                                  C.POINTER(C.c_int64)]
        L.or_pack_ref.restype = C.c_int64
        L.or_positions.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.or_positions.restype = None
        L.or_count_clear.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64]
        L.or_count_clear.restype = C.c_int64
        L.or_pack_knot.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int64,
                                   C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p,
                                   C.c_int64, C.POINTER(C.c_int64)]
        L.or_pack_knot.restype = C.c_int64
        L.or_check_rsa.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
        L.or_check_rsa.restype = C.c_int64
        _lib = L
    return _lib


class _Model(C.Structure):
    _fields_ = [
        ("d", C.c_int), ("T", C.c_int), ("nw", C.c_int), ("modmax", C.c_int), ("pin", C.c_int),
        ("seed", C.c_uint64), ("r", C.c_double), ("halfT", C.c_double),
        ("S", C.POINTER(C.c_double)), ("C", C.POINTER(C.c_double)),
        ("sz", C.POINTER(C.c_double)), ("order", C.POINTER(C.c_int)), ("euclid", C.c_int),
        ("rng", C.c_int),
    ]


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.ascontiguousarray(ctr, np.uint32),
                           np.ascontiguousarray(key, np.uint32), out)
    return out


def sincospi_rat(n: int, m: int) -> tuple[float, float]:
    s, c = C.c_double(), C.c_double()
    lib().or_sincospi_rat(n, m, C.byref(s), C.byref(c))
    return s.value, c.value


def log(v: float) -> float:
    return lib().or_log(v)


def xoshiro_seed(seed: int) -> np.ndarray:
    """The xoshiro256** state seeded from `seed` (four splitmix64 outputs)."""
    st = np.zeros(4, np.uint64)
    lib().or_xoshiro_seed(seed, st.ctypes.data)
    return st


def xoshiro_next(st: np.ndarray) -> int:
    """One xoshiro256** output; advances the state array in place."""
    assert st.dtype == np.uint64 and st.flags.c_contiguous and st.size == 4
    return int(lib().or_xoshiro_next(st.ctypes.data))


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class Candidate:
    a1: np.ndarray      # (d,)
    b: np.ndarray       # (d, nw) int
    c: np.ndarray       # (d, nw) a_j cos f_j
    s: np.ndarray       # (d, nw) a_j sin f_j
    amp: np.ndarray     # (d, nw) signed amplitudes a_j
    phase: np.ndarray   # (d, nw) phases f_j


class Model:
    """Oracle model for (d, T, terms, seed, pin)."""

    def __init__(self, d: int, T: int, terms: int | None = None, seed: int = 1, pin: bool = False,
                 euclid: bool = False, rng: int = 0):
        self.d, self.T = d, T
        self.terms = T if terms is None else terms
        self._m = _Model()
        rc = lib().or_model_init(C.byref(self._m), d, T, self.terms, seed, int(pin))
        if rc != 0:
            raise ValueError(f"or_model_init failed ({rc}) for d={d} T={T} terms={self.terms}")
        if euclid:
            lib().or_model_set_euclid(C.byref(self._m), 1)
        if rng:
            lib().or_model_set_rng(C.byref(self._m), int(rng))
        self.euclid = bool(euclid)
        self.rng = int(rng)
        self.nw = self._m.nw
        self.r = self._m.r

    def __del__(self):
        try:
            lib().or_model_free(C.byref(self._m))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self._m)

    @property
    def S(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._m.S, shape=(self.T, self.T + 1)).copy()

    @property
    def C(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._m.C, shape=(self.T, self.T + 1)).copy()

    @property
    def sz(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._m.sz, shape=(self.T,)).copy()

    @property
    def order(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._m.order, shape=(self.T,)).copy()

    def candidate(self, idx: int) -> Candidate:
        d, nw = self.d, self.nw
        a1 = np.zeros(d)
        b = np.zeros((d, nw), np.int32)
        c = np.zeros((d, nw))
        s = np.zeros((d, nw))
        amp = np.zeros((d, nw))
        ph = np.zeros((d, nw))
        lib().or_candidate(self.ref, idx, _p(a1), _p(b), _p(c), _p(s), _p(amp), _p(ph))
        return Candidate(a1, b, c, s, amp, ph)

    def positions(self, idx: int) -> np.ndarray:
        """Canonical wrapped positions (T, d) of candidate idx."""
        cd = self.candidate(idx)
        X = np.zeros((self.T, self.d))
        for t in range(self.T):
            lib().or_eval_t(self.ref, t, _p(cd.a1), _p(cd.b), _p(cd.c), _p(cd.s), _p(X[t]))
        return X

    def cell(self, x: float) -> int:
        return lib().or_cell(self.ref, x)

    def pack(self, max_candidates: int, stop_kind: int = 0, stop_batch: int = 1 << 20,
             window: int = 1 << 22, stop_rate: float = 1e-4, max_n: int = 0,
             cap: int = 1 << 22, want_pos: bool = False):
        """Serial RSA; returns (idx int64[N], gid int32[N], attempts[, pos (N,T,d)])."""
        idx = np.zeros(cap, np.int64)
        gid = np.zeros(cap, np.int32)
        pos = np.zeros((cap, self.T, self.d)) if want_pos else None
        att = C.c_int64(0)
        n = lib().or_pack(self.ref, max_candidates, stop_kind, stop_batch, window, stop_rate,
                          max_n, _p(idx), _p(gid), _p(pos) if want_pos else None, cap,
                          C.byref(att))
        if n < 0:
            raise RuntimeError(f"or_pack failed ({n})")
        n = min(n, cap)
        out = (idx[:n].copy(), gid[:n].copy(), att.value)
        if want_pos:
            out = out + (pos[:n].copy(),)
        return out

    def pack_mt(self, max_candidates: int, threads: int, stop_kind: int = 0,
                stop_batch: int = 1 << 20, window: int = 1 << 22, stop_rate: float = 1e-4,
                cap: int = 1 << 22):
        """or_pack with a parallel per-block prefilter: returns (idx int64[N], attempts)."""
        idx = np.zeros(cap, np.int64)
        att = C.c_int64(0)
        n = lib().or_pack_mt(self.ref, max_candidates, stop_kind, stop_batch, window, stop_rate,
                             threads, _p(idx), cap, C.byref(att))
        if n < 0:
            raise RuntimeError(f"or_pack_mt failed ({n})")
        return idx[:min(n, cap)].copy(), att.value

    def pack_sub(self, pos0: np.ndarray, gid0: np.ndarray, first: int, budget: float,
                 accept_rate: float = 0.0, fill_rate: float = 0.0, cap: int = 1 << 20,
                 rounds: bool = False):
        """Subpath phase after the uniques pos0 (n0,T,d) with groups gid0: returns
        (idx int64[Nsub], gid int32[Nsub], subpath candidates drawn, newly filled cells)
        and, with rounds=True, the round records (start, size, survivors, admitted, fresh)."""
        pos0 = np.ascontiguousarray(pos0, np.float64)
        gid0 = np.ascontiguousarray(gid0, np.int32)
        idx = np.zeros(cap, np.int64)
        gid = np.zeros(cap, np.int32)
        att = C.c_int64(0)
        filled = C.c_int64(0)
        rcap = 1 << 16
        rec = np.zeros((rcap, 5), np.int64)
        nr = C.c_int64(0)
        n = lib().or_pack_sub(self.ref, _p(pos0), _p(gid0), len(gid0), first, float(budget),
                              float(accept_rate), float(fill_rate), _p(idx), _p(gid), cap,
                              C.byref(att), C.byref(filled), _p(rec), rcap, C.byref(nr))
        if n < 0:
            raise RuntimeError(f"or_pack_sub failed ({n})")
        n = min(n, cap)
        out = (idx[:n].copy(), gid[:n].copy(), att.value, filled.value)
        if rounds:
            out = out + (rec[:min(nr.value, rcap)].copy(),)
        return out

    def pack_ref(self, budget: float, rule: int | None = None, stop_rate: float = 0.0,
                 max_n: int = 0, cap: int = 1 << 22):
        """Serial RSA on the reference engine's round schedule (rule 2 = braid_cuda,
        3 = braid_cuda3d; default: the model's dimension).  Returns a dict with idx, gid,
        attempts, decided and rounds (start, size, survivors, admitted, fresh)."""
        rule = self.d if rule is None else rule
        idx = np.zeros(cap, np.int64)
        gid = np.zeros(cap, np.int32)
        att, dec, nr = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        rcap = 1 << 16
        rec = np.zeros((rcap, 5), np.int64)
        n = lib().or_pack_ref(self.ref, float(budget), rule, float(stop_rate), max_n, _p(idx),
                              _p(gid), cap, C.byref(att), C.byref(dec), _p(rec), rcap,
                              C.byref(nr))
        if n < 0:
            raise RuntimeError(f"or_pack_ref failed ({n})")
        n = min(n, cap)
        return {"idx": idx[:n].copy(), "gid": gid[:n].copy(), "attempts": att.value,
                "decided": dec.value, "rounds": rec[:min(nr.value, rcap)].copy()}

    def count_clear(self, first: int, count: int, acc) -> int:
        """Candidates in [first, first + count) touching none of acc (pairwise, no grid)."""
        acc = np.ascontiguousarray(acc, np.int64)
        n = lib().or_count_clear(self.ref, first, count, _p(acc), len(acc))
        if n < 0:
            raise RuntimeError(f"or_count_clear failed ({n})")
        return n

    def pack_knot(self, budget: float, nu_attempts: float = 0.0, nu_div: float = 1.0,
                  max_strands: int = 16384):
        """The knot-regime tool's two phases: dict with idx, species, attempts,
        primary_attempts, n_primary and batches (start, size, survivors, admitted, species)."""
        idx = np.zeros(max_strands, np.int64)
        sp = np.zeros(max_strands, np.int32)
        att, patt, npri, nr = C.c_int64(0), C.c_int64(0), C.c_int64(0), C.c_int64(0)
        rcap = 1 << 16
        rec = np.zeros((rcap, 5), np.int64)
        n = lib().or_pack_knot(self.ref, float(budget), float(nu_attempts), float(nu_div), max_strands,
                               _p(idx), _p(sp), max_strands, C.byref(att), C.byref(patt),
                               C.byref(npri), _p(rec), rcap, C.byref(nr))
        if n < 0:
            raise RuntimeError(f"or_pack_knot failed ({n})")
        return {"idx": idx[:n].copy(), "species": sp[:n].copy(), "attempts": att.value,
                "primary_attempts": patt.value, "n_primary": npri.value,
                "batches": rec[:min(nr.value, rcap)].copy()}

    def positions_of(self, idx) -> np.ndarray:
        """Canonical wrapped positions (n, T, d) of candidates idx (evaluated in C)."""
        idx = np.ascontiguousarray(idx, np.int64)
        out = np.zeros((len(idx), self.T, self.d))
        if len(idx):
            lib().or_positions(self.ref, _p(idx), len(idx), _p(out))
        return out

    def touches(self, x, y) -> bool:
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        return bool(lib().or_touches(self.ref, _p(x), _p(y)))

    def check_rsa(self, M: int, acc: np.ndarray) -> int:
        acc = np.ascontiguousarray(acc, np.int64)
        return lib().or_check_rsa(self.ref, M, _p(acc), len(acc))
