"""The sparse occupied-cell grid (--sparse, cuda/braid_cuda3d.cu:82-91,1262-1353): the same
accepted lists and group labels as the dense cell table and the oracle, in both precision
forms, through table growth; the serial-RSA characterization at full size."""

from __future__ import annotations

import json
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gold(name):
    return json.load(open(os.path.join(GOLD, name + ".json")))


def _pack(ugpack, d, T, seed, precision, sparse, **kw):
    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    with ugpack.Packer(d, T, seed=seed, precision=prec, sparse=sparse, **kw) as p:
        st = p.run()
        idx, gid = p.accepted()
    return st, idx, gid


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("name", ["cfg1_d2_T16_s2", "d3_T12_s1", "d2_T512_s2", "d3_T256_s1", "d3_T16_terms4_s5"])
def test_sparse_equals_golden(gpu, name, precision):
    import ugpack

    g = _gold(name)
    print(f"sparse {name} {precision}", flush=True)
    st, idx, gid = _pack(ugpack, g["dim"], g["T"], g["seed"], precision, True, terms=g["terms"],
                         max_attempts=g["candidates"], max_paths=16)
    assert st.attempts == g["candidates"] and list(idx) == g["idx"] and list(gid) == g["gid"]


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_sparse_config3_prefix_through_table_growth(gpu, precision):
    # 39,810 paths x 128 timesteps from a table sized for 1024 paths: five rebuilds
    import ugpack

    g = _gold("cfg3_d3_T128_s1_prefix1M")
    st, idx, _ = _pack(ugpack, 3, 128, 1, precision, True, max_attempts=g["candidates"], max_paths=16)
    assert st.n == g["n"] and list(idx) == g["idx"]


@pytest.mark.parametrize("euclid", [False, True])
def test_sparse_equals_dense_and_oracle(gpu, euclid):
    import ugpack

    for d, T, seed, M in ((3, 16, 2, 30000), (2, 24, 3, 30000), (3, 33, 4, 8000)):
        want, _, _ = po.Model(d, T, seed=seed, euclid=euclid).pack(M)
        for precision in ("fp64", "split"):
            _, a, _ = _pack(ugpack, d, T, seed, precision, True, max_attempts=M, euclid=euclid)
            _, b, _ = _pack(ugpack, d, T, seed, precision, False, max_attempts=M, euclid=euclid)
            assert list(a) == list(want) == list(b), (d, T, precision)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_sparse_subpaths_equal_golden(gpu, precision):
    import ugpack

    g = _gold("sub_d2_T16_s1_accept2e-3")
    z = np.load(os.path.join(GOLD, "sub_d2_T16_s1_accept2e-3.npz"))
    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    with ugpack.Packer(2, 16, seed=1, precision=prec, sparse=True, max_attempts=g["unique_candidates"]) as p:
        p.run()
        st = p.run_sub(max_attempts=1e15, accept_rate=g["accept_rate"])
        idx, gid = p.accepted()
    assert st.nsub == g["nsub"] and st.attempts == g["attempts"]
    assert np.array_equal(idx, np.concatenate([z["uidx"], z["sidx"]]))
    assert np.array_equal(gid, np.concatenate([z["ugid"], z["sgid"]]))
    with ugpack.Packer(2, 16, seed=1, precision=prec, max_attempts=g["unique_candidates"]) as q:
        q.run()
        sd = q.run_sub(max_attempts=1e15, accept_rate=g["accept_rate"])
    assert sd.filled == st.filled


def test_sparse_config3_to_jam_at_full_size(gpu):
    # 3+1 T=128 to jamming on the sparse grid: the dense form's list; every consumed candidate
    # and every accepted path checked on the sparse grid
    import ugpack

    print("cfg3 sparse jam", flush=True)
    lists = {}
    for sparse in (True, False):
        with ugpack.Packer(3, 128, seed=1, precision=ugpack.PREC_SPLIT_BF16, sparse=sparse,
                           stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4) as p:
            st = p.run()
            lists[sparse] = (p.accepted()[0], st.filled)
            if sparse:
                tot = {"unexplained_rejects": 0, "rejects": 0}
                for first in range(0, st.attempts, 1 << 26):
                    c = p.check(first, min(1 << 26, st.attempts - first))
                    assert c["colliding_uniques"] == 0 and c["uniques"] == st.n
                    tot["unexplained_rejects"] += c["unexplained_rejects"]
                    tot["rejects"] += c["rejects"]
                assert tot == {"unexplained_rejects": 0, "rejects": st.attempts - st.n}
    assert np.array_equal(lists[True][0], lists[False][0]) and lists[True][1] == lists[False][1]


def test_sparse_config5_to_jam_equals_dense(gpu):
    # 3+1 T=256 to jamming: the sparse grid's table is rebuilt as it grows past 65,536 paths x
    # 256 timesteps (16.7M entries, more than one entry per thread of a capped grid); the list
    # must stay the dense table's (whose whole range test_gpu_pack checks)
    import ugpack

    out = {}
    for sparse in (True, False):
        print(f"cfg5 jam sparse={sparse}", flush=True)
        with ugpack.Packer(3, 256, seed=1, precision=ugpack.PREC_SPLIT_BF16, sparse=sparse,
                           stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4) as p:
            st = p.run()
            out[sparse] = (st.n, st.attempts, st.filled, p.accepted()[0])
    assert out[True][0] > 65536
    assert out[True][:3] == out[False][:3]
    assert np.array_equal(out[True][3], out[False][3])


def test_braid_ug3d_sparse_flag(gpu, tmp_path):
    import ugpack

    lines = []
    for sparse in (False, True):
        cmd = ugpack.command(3, 24, attempts=300000.0, seed=5, accept_rate=0.0, terms=24, sparse=sparse,
                             curve_path=str(tmp_path / "c.csv"))
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert ("(sparse grid)" in r.stderr) == sparse
        lines.append(re.search(r"done: .*", r.stderr).group(0))
    assert lines[0] == lines[1]
