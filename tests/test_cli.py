"""Command-line front ends (braid_ug, braid_ug3d, frame_bake_ug): presence and the argv
checks that end before any GPU work (cuda/braid_cuda.cu:534-636, cuda/braid_cuda3d.cu:862-964,
cuda/frame_bake.cu:409-470)."""

from __future__ import annotations

import os
import subprocess

import pytest

import ugpack

BIN = ugpack.BIN_DIR


def run(*argv):
    r = subprocess.run([*argv], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stderr


@pytest.mark.parametrize("name", ["braid_ug", "braid_ug3d", "frame_bake_ug", "braid_knot_ug"])
def test_binaries_are_built(name):
    exe = os.path.join(BIN, name)
    assert os.access(exe, os.X_OK), f"{exe} missing: run __graft_entry__.build()"


@pytest.mark.parametrize("name", ["braid_ug", "braid_ug3d"])
def test_stale_flags_fail_loudly(name):
    exe = os.path.join(BIN, name)
    rc, err = run(exe, "-t", "16", "--smart")
    assert rc == 1 and "--smart" in err
    rc, err = run(exe, "-t", "20", "--maxfreq", "10")       # maxfreq must equal T
    assert rc == 1 and "--maxfreq 10" in err
    rc, err = run(exe, "-t", "16", "--diag", "x") if name == "braid_ug3d" else (1, "--diag")
    assert rc == 1 and "--diag" in err


def test_angle_sample_is_fatal_in_3d():
    rc, err = run(os.path.join(BIN, "braid_ug3d"), "--angle-sample")
    assert rc == 1 and "--angle-sample" in err


def test_frame_bake_argv_errors(tmp_path):
    exe = os.path.join(BIN, "frame_bake_ug")
    rc, err = run(exe)
    assert rc == 1 and err.startswith("usage:")
    rc, err = run(exe, "--bogus")
    assert rc == 1 and "unknown flag --bogus" in err
    dump = tmp_path / "d.csv"
    dump.write_text("ax2,ay2,ax_1,bx_1,fx_1,ay_1,by_1,fy_1\n0.1,0.2,0.3,2,0,0.1,3,0\n")
    rc, err = run(exe, "--dump", str(dump), "--observer-path", "0", "--front", "slow")
    assert rc == 1 and "--front" in err
    rc, err = run(exe, "--dump", str(dump), "--observer-path", "4")
    assert rc == 1 and "observer path 4 out of range" in err
    d3 = tmp_path / "d3.csv"
    d3.write_text("ax2,ay2,aw2,ax_1,bx_1,fx_1,ay_1,by_1,fy_1,aw_1,bw_1,fw_1\n"
                  "0.1,0.2,0.3,0.1,2,0,0.1,3,0,0.1,2,0\n")
    rc, err = run(exe, "--dump", str(d3), "--observer-path", "0")
    assert rc == 1 and "3+1 dump" in err
    rc, err = run(exe, "--dump", str(tmp_path / "missing.csv"), "--observer-path", "0")
    assert rc == 1 and "cannot open dump" in err


def test_engine_argv_matches_build_command_layout():
    # the flag order of braidlab/engine.py:50-101
    cmd = ugpack.command(2, 64, attempts=1e8, seed=3, accept_rate=1e-6, terms=64, subpaths=True,
                         sub_attempts=1e9, curve_path="c.csv")
    assert cmd[0].endswith("braid_ug")
    assert cmd[1:] == ["-t", "64", "--attempts", "100000000.0", "--uniform", "--torus", "--phase",
                       "--seed", "3", "--until-accept-rate", "1e-06", "--terms", "64",
                       "--subpaths", "--sub-until-accept-rate", "1e-06", "--sub-attempts",
                       "1000000000.0", "--maxfreq", "64", "--curve", "c.csv"]
    assert ugpack.command(3, 12, attempts=1e6, seed=1, accept_rate=0.0)[0].endswith("braid_ug3d")


def test_knot_argv_errors():
    # cuda/braid_knot.cu:284-306: unknown flags, --nu-div < 1 and a bad --dim end the run
    exe = os.path.join(BIN, "braid_knot_ug")
    rc, err = run(exe, "--bogus")
    assert rc == 1 and "unknown flag --bogus" in err
    rc, err = run(exe, "--nu-div", "0.5")
    assert rc == 1 and "--nu-div must be >= 1" in err
    rc, err = run(exe, "--dim", "4")
    assert rc == 1 and "--dim must be 2 or 3" in err
    rc, err = run(exe, "-t", "16", "--terms", "40")
    assert rc == 1 and "terms 40" in err
    rc, err = run(exe, "--rng", "mt19937")  # the one flag of its own besides --precision
    assert rc == 1 and "--rng must be philox|xoshiro-ref" in err
