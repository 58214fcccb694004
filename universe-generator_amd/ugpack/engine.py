"""In-process packing runs and the command line of the engine binaries.

`pack()` runs one universe through the C ABI and returns what the reference's engine
writes (the `done:` numbers, the dump parameters with group labels, the curve).
`command()` builds the argv of the `braid_ug` / `braid_ug3d` front ends from the same
fields as braidlab.engine.build_command (braidlab/engine.py:50-101), so an orchestrator can
point its binary path at ugpack/bin/ unchanged.
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

BIN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin")


@dataclass
class PackResult:
    n: int                    # accepted uniques (done: N=...)
    attempts: int             # candidates consumed, both phases (done: ... in <attempts>)
    nsub: int                 # accepted subpaths
    sub_attempts: int         # candidates consumed by the subpath phase
    filled: int               # occupied (timestep, cell) pairs
    idx: np.ndarray           # candidate index of each accepted path, admission order
    gid: np.ndarray           # group of each path (a unique's is its own row)
    params: np.ndarray        # (rows, dim, 1 + 3 nw): a1, then (amplitude, frequency, phase)
    curve: np.ndarray         # (rows, 4): attempts, n, nsub, filled
    seconds: float
    status: abi.Status = field(repr=False, default=None)


def pack(dim: int, T: int, *, terms: int = 2, seed: int = 12345, attempts: float = 1e8,
         accept_rate: float = 0.0, fixed_window: bool = False, max_n: int = 0,
         pin_sin1: bool = False, euclid: bool = False, sparse: bool = False, subpaths: bool = False,
         sub_attempts: float = 0.0, sub_accept_rate: float = 0.0, sub_fill_rate: float = 0.0,
         precision: str = "split", want_params: bool = True) -> PackResult:
    """One run with the engine's flag semantics (defaults are the engine's: T=120 aside,
    --terms 2, --seed 12345, --attempts 1e8): the engine's round schedule, whose last round
    may pass `attempts`, and with accept_rate > 0 the windowed stop of
    cuda/braid_cuda.cu:930-943 (fixed_window: the fixed 2^20/2^22 rule instead)."""
    prec = abi.PREC_SPLIT_BF16 if precision == "split" else abi.PREC_FP64
    kind = abi.STOP_FIXED_WINDOW if fixed_window and accept_rate > 0 else abi.STOP_REF_WINDOW
    with abi.Packer(dim, T, terms=max(2, terms), seed=seed, precision=prec, pin_sin1=pin_sin1,
                    stop_kind=kind, max_attempts=attempts, stop_rate=accept_rate, max_n=max_n,
                    euclid=euclid, sparse=sparse) as p:
        st = p.run()
        if subpaths:
            st = p.run_sub(max_attempts=sub_attempts or attempts, accept_rate=sub_accept_rate,
                           fill_rate=sub_fill_rate)
        idx, gid = p.accepted()
        params = p.params() if want_params else np.zeros((0, dim, 0))
        curve = p.curve()
    return PackResult(st.n, st.attempts, st.nsub, st.sub_attempts, st.filled, idx, gid, params,
                      curve, st.seconds, st)


def command(dim: int, T: int, *, attempts: float, seed: int, accept_rate: float,
            terms: int = 2, subpaths: bool = False, sub_attempts: float = 0.0,
            euclid: bool = False, sparse: bool = False, curve_path: str = "curve.csv",
            binary: str | None = None) -> list[str]:
    """The argv braidlab.engine.build_command would produce, aimed at this package's binary."""
    exe = binary or os.path.join(BIN_DIR, "braid_ug" if dim == 2 else "braid_ug3d")
    cmd = [exe, "-t", str(T), "--attempts", repr(attempts), "--uniform", "--torus", "--phase",
           "--seed", str(seed), "--until-accept-rate", repr(accept_rate)]
    if euclid:
        cmd += ["--euclid-collision"]
    if sparse:
        cmd += ["--sparse"]
    if terms != 2:
        cmd += ["--terms", str(terms)]
    if subpaths:
        cmd += ["--subpaths", "--sub-until-accept-rate", repr(accept_rate)]
        if sub_attempts:
            cmd += ["--sub-attempts", repr(sub_attempts)]
    cmd += ["--maxfreq", str(T), "--curve", curve_path]
    return cmd
