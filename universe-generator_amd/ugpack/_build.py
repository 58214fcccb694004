"""Build libugpack.so in-tree with hipcc for gfx950."""

from __future__ import annotations

import os
import shutil
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG_DIR), "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libugpack.so")
SOURCES = ["ug_pack.hip", "ug_fast.hip", "ug_api.hip", "ug_cache.hip", "ug_frames.hip", "ug_ensemble.hip",
           "ug_knot.hip", "ug_batch.hip"]

# -ffp-contract=off: the canonical FP64 arithmetic uses explicit fma only, so host and
# device produce identical bits.  Packed FP32 instructions are disabled for all HIP code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the ugpack HIP library cannot be built")


BIN_DIR = os.path.join(PKG_DIR, "bin")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(PKG_DIR)), "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# Command-line front ends over the C ABI: the packer engines (2+1 and 3+1 builds of one
# source) and the frame baker.  Host-only C++; they link libugpack.so by rpath.
CLIS = {
    "braid_ug": ("ug_cli_pack.cpp", ["-DUG_CLI_DIM=2"]),
    "braid_ug3d": ("ug_cli_pack.cpp", ["-DUG_CLI_DIM=3"]),
    "frame_bake_ug": ("ug_cli_bake.cpp", ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
                                          f"-L{ROCM}/lib", "-lamdhip64"]),
    "braid_knot_ug": ("ug_cli_knot.cpp", []),
}


def _run(cmd: list[str], what: str, verbose: bool) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)


def build_clis(verbose: bool = False) -> list[str]:
    """Compile the command-line front ends into ugpack/bin/ (always rebuilds)."""
    os.makedirs(BIN_DIR, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    out = []
    for name, (src, extra) in CLIS.items():
        exe = os.path.join(BIN_DIR, name)
        srcp = os.path.join(CSRC, src)
        libs = [x for x in extra if x.startswith("-l") or x.startswith("-L")]
        flags = [x for x in extra if x not in libs]
        _run([cxx, "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", *flags, "-o", exe + ".tmp", srcp,
              f"-L{PKG_DIR}", "-lugpack", *libs, "-Wl,-rpath,$ORIGIN/..",
              f"-Wl,-rpath-link,{ROCM}/lib"], f"{name} build", verbose)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build(verbose: bool = False) -> str:
    """Compile every HIP source into ugpack/libugpack.so, then the command-line front ends
    (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), *HIPCC_FLAGS, "-o", tmp, *srcs], "hipcc", verbose)
    os.replace(tmp, LIB_PATH)
    build_clis(verbose)
    return LIB_PATH
