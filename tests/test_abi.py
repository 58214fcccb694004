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


STRUCTS = {  # C struct name -> ctypes mirror
    "ug_pack_config": abi.PackConfig, "ug_pack_status": abi.PackStatus, "ug_sub_config": abi.SubConfig,
    "ug_kernel_stats": abi.KernelStats, "ug_front": abi.Front, "ug_observer": abi.Observer,
    "ug_frame_info": abi.FrameInfo, "ug_knot_config": abi.KnotConfig, "ug_knot_status": abi.KnotStatus,
}


def test_struct_layouts_match_the_header(tmp_path):
    # every field's offset and every struct's size, as the C compiler lays them out from
    # include/ugpack.h, against the ctypes mirrors
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ugpack.h"', "int main(void) {"]
    for name, st in STRUCTS.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in st._fields_:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out if ln}
    for name, st in STRUCTS.items():
        assert got[(name, "size")] == ctypes.sizeof(st), name
        for f, _ in st._fields_:
            assert got[(name, f)] == getattr(st, f).offset, (name, f)


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
