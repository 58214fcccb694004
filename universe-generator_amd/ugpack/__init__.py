"""ugpack -- MI355X-native RSA worldline packer (drop-in for braidlab's engine path).

The compute runs in libugpack.so (hand-written gfx950 HIP kernels behind a C ABI,
include/ugpack.h); this package is the host-side mirror of the reference's
braidlab/engine.py call surface.
"""

from ._build import build, LIB_PATH  # noqa: F401
from .abi import (  # noqa: F401
    PREC_FP64, PREC_SPLIT_BF16, RNG_PHILOX, RNG_XOSHIRO_REF, STOP_BUDGET, STOP_FIXED_WINDOW, STOP_REF_WINDOW,
    SUB_ACCEPT_RATE, SUB_BUDGET, SUB_FILL_RATE, Knot, Packer, Status, UgError, ensemble, lib,
    pack_config,
)
from .engine import BIN_DIR, PackResult, command, pack  # noqa: F401
