"""The reference's own candidate stream on the GPU (rng_kind UG_RNG_XOSHIRO_REF): candidate
idx draws xoshiro256** seeded as cuda/braid_knot.cu:207, in the draw order of its proposal
(:112-166).  Bar: bit-exact against the oracle's restatement (oracle/ug_oracle.c
candidate_xref, pinned in tests/test_oracle.py by the generator's published outputs and the
stream-order and distribution checks) -- positions, accepted lists, dump parameters, the
subpath phase, the knot engine and the baker's device-built worldlines."""

from __future__ import annotations

import math

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d,T,terms,seed,pin", [
    (2, 16, 16, 1, False), (3, 12, 12, 2, False), (2, 24, 5, 3, False), (3, 16, 4, 5, False),
    (2, 16, 16, 2, True), (3, 128, 128, 1, False), (2, 200, 200, 9, False), (2, 512, 512, 2, False),
])
def test_xref_positions_bit_exact(gpu, d, T, terms, seed, pin):
    import torch
    import ugpack

    print(f"xref eval d={d} T={T} terms={terms} seed={seed} pin={pin}", flush=True)
    idx = np.array([0, 1, 2, 3, 17, 999, 123456, 2**33 + 5] + list(range(40, 60)), np.int64)
    with ugpack.Packer(d, T, terms=terms, seed=seed, pin_sin1=pin, rng_kind=ugpack.RNG_XOSHIRO_REF) as p:
        got = p.eval_candidates(torch.from_numpy(idx)).cpu().numpy()
    m = po.Model(d, T, terms=terms, seed=seed, pin=pin, rng=1)
    want = np.stack([m.positions(int(i)) for i in idx])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), \
        f"max diff {np.max(np.abs(got - want))}"
    # and the stream is not the Philox one
    assert not np.array_equal(want, np.stack([po.Model(d, T, terms=terms, seed=seed, pin=pin).positions(int(i))
                                              for i in idx]))


@pytest.mark.parametrize("d,T,terms,seed,pin,sparse,cands", [
    (2, 16, 0, 1, False, False, 20000), (3, 12, 0, 2, False, False, 8000), (2, 24, 5, 3, False, False, 20000),
    (2, 16, 0, 4, True, False, 20000), (3, 16, 0, 5, False, True, 10000), (2, 32, 0, 6, False, True, 30000),
])
def test_xref_pack_equals_oracle(gpu, d, T, terms, seed, pin, sparse, cands):
    import ugpack

    print(f"xref pack d={d} T={T} terms={terms} seed={seed} pin={pin} sparse={sparse}", flush=True)
    m = po.Model(d, T, terms=terms or None, seed=seed, pin=pin, rng=1)
    want, wgid, att = m.pack(cands)
    with ugpack.Packer(d, T, terms=terms, seed=seed, pin_sin1=pin, sparse=sparse, max_attempts=cands,
                       rng_kind=ugpack.RNG_XOSHIRO_REF) as p:
        st = p.run()
        idx, gid = p.accepted()
        pos = p.positions().cpu().numpy()
        par = p.params()
        chk = p.check(0, cands)
    assert chk == {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": len(want),
                   "rejects": cands - len(want)}
    assert st.done and st.attempts == att
    assert list(idx) == list(want) and list(gid) == list(wgid)
    assert np.array_equal(pos.view(np.uint64), m.positions_of(idx).view(np.uint64))
    for r, i in enumerate(idx[:20]):
        cd = m.candidate(int(i))
        for a in range(d):
            assert par[r, a, 0] == cd.a1[a]
            assert np.array_equal(par[r, a, 1::3], cd.amp[a])
            assert np.array_equal(par[r, a, 2::3], cd.b[a].astype(float))
            assert np.array_equal(par[r, a, 3::3], cd.phase[a])


def test_xref_subpaths_equal_oracle(gpu):
    import ugpack

    d, T, seed, unique, sub = 2, 16, 4, 3000, 300000
    m = po.Model(d, T, seed=seed, rng=1)
    idx0, gid0, att = m.pack(unique)
    sidx, sgid, satt, _ = m.pack_sub(m.positions_of(idx0), gid0, first=att, budget=sub)
    with ugpack.Packer(d, T, seed=seed, max_attempts=unique, rng_kind=ugpack.RNG_XOSHIRO_REF) as p:
        st = p.run()
        assert st.attempts == att
        st = p.run_sub(max_attempts=sub)
        idx, gid = p.accepted()
    assert st.sub_done and st.attempts == att + satt and st.nsub == len(sidx) > 0
    assert list(idx) == list(idx0) + list(sidx) and list(gid) == list(gid0) + list(sgid)


@pytest.mark.parametrize("case", [(2, 16, 0, 1, True, 1e6, 3e5, 2.0), (3, 12, 0, 2, False, 5e5, 0.0, 1.0),
                                  (2, 24, 5, 4, True, 1e6, 5e5, 3.0)])
def test_xref_knot_equals_oracle(gpu, case):
    import ugpack

    d, T, terms, seed, pin, budget, nu, div = case
    print(f"xref knot {case}", flush=True)
    want = po.Model(d, T, terms=terms or None, seed=seed, pin=pin, rng=1).pack_knot(
        budget, nu_attempts=nu, nu_div=div, max_strands=16384)
    with ugpack.Knot(d, T, terms=terms, seed=seed, pin_sin1=pin, max_attempts=budget, nu_attempts=nu,
                     nu_div=div, rng_kind=ugpack.RNG_XOSHIRO_REF) as k:
        st = k.run()
        idx, sp = k.strands()
        rows = k.batches()
    assert st.done and st.attempts == want["attempts"] and st.n_primary == want["n_primary"]
    assert np.array_equal(idx, want["idx"]) and np.array_equal(sp, want["species"])
    assert np.array_equal(rows, want["batches"])


def test_xref_paths_from_pack_equal_dump_parameters(gpu):
    import torch

    import ugpack
    from ugpack import frames

    with ugpack.Packer(2, 32, seed=3, max_attempts=20000, rng_kind=ugpack.RNG_XOSHIRO_REF) as p:
        p.run()
        dev = frames.Paths.from_packer(p)
        host = frames.Paths.from_params(p)
        z = np.linspace(0.05, math.pi - 0.05, 9)
        a, _ = frames.slices(dev, z, raster_px=64)
        b, _ = frames.slices(host, z, raster_px=64)
    assert dev.n == host.n > 5
    diff = (a - b).abs()
    diff = torch.minimum(diff, 2.0 - diff)
    assert float(diff.max()) < 1e-6


def test_xref_needs_the_fp64_form(gpu):
    import ugpack

    with pytest.raises(ugpack.UgError, match="FP64 form only"):
        ugpack.Packer(2, 16, precision=ugpack.PREC_SPLIT_BF16, rng_kind=ugpack.RNG_XOSHIRO_REF)
    with pytest.raises(ugpack.UgError, match="FP64 form only"):
        ugpack.Knot(2, 16, precision=ugpack.PREC_SPLIT_BF16, rng_kind=ugpack.RNG_XOSHIRO_REF)
    with pytest.raises(ugpack.UgError, match="unknown rng_kind"):
        ugpack.Packer(2, 16, rng_kind=7)
    import torch

    with ugpack.Packer(2, 16, rng_kind=ugpack.RNG_XOSHIRO_REF) as p:
        with pytest.raises(ugpack.UgError, match="FP64 form only"):
            p.eval_candidates(torch.arange(4, dtype=torch.int64), precision=ugpack.PREC_SPLIT_BF16)


@pytest.mark.parametrize("batched", [False, True])
def test_xref_universes_in_an_ensemble(gpu, batched):
    # the batched engine hands reference-stream universes to its single-universe loop; either
    # runner gives each universe the oracle's list on its own stream
    import ugpack

    mix = [(2, 16, 1, 20000, 1), (3, 12, 2, 15000, 0), (2, 32, 3, 20000, 1), (3, 16, 4, 10000, 1),
           (2, 16, 5, 20000, 0)]
    print(f"xref ensemble batched={batched}", flush=True)
    cfgs = [ugpack.pack_config(d, T, seed=s, max_attempts=m, rng_kind=x) for d, T, s, m, x in mix]
    status, digests = ugpack.ensemble(cfgs, threads=3, digests=True, batched=batched, window=4)
    for (d, T, s, m, x), st, dg in zip(mix, status, digests):
        ref, _, _ = po.Model(d, T, seed=s, rng=x).pack(m)
        assert st.attempts == m and st.done and st.n == len(ref)
        assert dg == ugpack.abi.list_digest(ref), (d, T, s, x)


def test_braid_knot_ug_draws_the_reference_stream(gpu):
    # --rng xoshiro-ref: the knot front end's phase lines and curve follow the oracle's packing
    # of the reference's own candidates (the FP64 form chosen by the flag)
    import os
    import re
    import subprocess

    import ugpack

    d, T, seed, budget, nu, div = 2, 16, 3, 1e6, 3e5, 2.0
    want = po.Model(d, T, seed=seed, pin=True, rng=1).pack_knot(budget, nu_attempts=nu, nu_div=div)
    philox = po.Model(d, T, seed=seed, pin=True).pack_knot(budget, nu_attempts=nu, nu_div=div)
    assert not np.array_equal(want["idx"], philox["idx"])
    cmd = [os.path.join(ugpack.BIN_DIR, "braid_knot_ug"), "-t", str(T), "--dim", str(d), "--pin-sin1",
           "--seed", str(seed), "--attempts", repr(budget), "--nu-attempts", repr(nu), "--nu-div", repr(div),
           "--rng", "xoshiro-ref", "--curve"]
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        curve = os.path.join(tmp, "k.csv")
        r = subprocess.run(cmd + [curve], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        rows = np.loadtxt(curve, delimiter=",", skiprows=1, dtype=np.int64, ndmin=2)
    n, npri = len(want["idx"]), want["n_primary"]
    assert re.search(rf"done: N={n} \(primary {npri}, nu {n - npri}\) in {want['attempts']} attempts", r.stderr)
    assert np.array_equal(rows[:, 0], want["idx"] + 1) and np.array_equal(rows[:, 2], want["species"])
    bad = subprocess.run(cmd[:-1] + ["--rng", "mt19937"], capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "--rng must be" in bad.stderr
