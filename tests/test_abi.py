"""The C-ABI library loads and exports every entry point include/ugpack.h declares
(no compute calls: these run without a GPU)."""

from __future__ import annotations

import ctypes
import os
import re

import pytest

import ugpack
from ugpack import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols() -> set[str]:
    text = open(os.path.join(ROOT, "include", "ugpack.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(ug_[a-z_0-9]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ugpack.LIB_PATH)
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(lib, name), f"{name} not exported"
    assert decl == set(abi.exported_symbols())


def test_abi_version_and_error_string():
    L = ugpack.lib()
    assert L.ug_abi_version() == abi.ABI_VERSION
    assert isinstance(L.ug_last_error(), bytes)


def test_config_struct_layout_matches_header():
    # ug_pack_config: 4 i32, u64, 2 i32, 2 f64, 3 i64, f64, i64, 2 i32
    assert ctypes.sizeof(abi.PackConfig) == 4 * 4 + 8 + 2 * 4 + 2 * 8 + 3 * 8 + 8 + 8 + 2 * 4
    assert ctypes.sizeof(abi.PackStatus) == 3 * 8 + 2 * 4 + 2 * 8 + 2 * 8 + 2 * 8


def test_profile_and_kernel_constants_match_header():
    text = open(os.path.join(ROOT, "include", "ugpack.h")).read()
    assert re.search(r"UG_KERNEL_PREFILTER\s*=\s*0", text) and re.search(r"UG_KERNEL_RECHECK\s*=\s*1", text)
    assert (abi.KERNEL_PREFILTER, abi.KERNEL_RECHECK) == (0, 1)
    assert "UG_PROFILE_PREFILTER = 1 << UG_KERNEL_PREFILTER" in text
    assert "UG_PROFILE_RECHECK = 1 << UG_KERNEL_RECHECK" in text
    assert (abi.PROFILE_PREFILTER, abi.PROFILE_RECHECK, abi.PROFILE_ALL) == (1, 2, 3)


def test_tuner_rejects_an_empty_sample():
    from ugpack import tune

    with pytest.raises(ValueError):
        tune.ensemble_threads([])
