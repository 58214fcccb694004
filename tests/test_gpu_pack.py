"""GPU parity of the HIP packer against the CPU oracle, through the C ABI.

Bar: bit-exact.  Canonical FP64 positions must equal the oracle's bits; the accepted
index lists and group labels must equal serial RSA (oracle or committed golden fixture).
"""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def _gold(name):
    return json.load(open(os.path.join(GOLD, name + ".json")))


@pytest.mark.parametrize("d,T,terms,seed,pin", [
    (2, 16, 16, 1, False), (2, 64, 64, 2, False), (3, 12, 12, 1, False), (2, 24, 5, 3, False),
    (3, 16, 4, 5, False), (2, 16, 16, 2, True), (3, 128, 128, 1, False), (2, 200, 200, 9, False),
    (3, 256, 256, 1, False), (2, 512, 512, 2, False),
])
def test_fp64_positions_bit_exact(gpu, d, T, terms, seed, pin):
    import torch
    import ugpack

    print(f"eval d={d} T={T} terms={terms} seed={seed}", flush=True)
    idx = np.array([0, 1, 2, 3, 17, 999, 123456, 2**33 + 5] + list(range(40, 80)), np.int64)
    with ugpack.Packer(d, T, terms=terms, seed=seed, pin_sin1=pin) as p:
        got = p.eval_candidates(torch.from_numpy(idx)).cpu().numpy()
    m = po.Model(d, T, terms=terms, seed=seed, pin=pin)
    want = np.stack([m.positions(int(i)) for i in idx])
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), \
        f"max diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_config1_accepted_list_equals_golden(gpu, seed):
    import ugpack

    print(f"cfg1 seed={seed}", flush=True)
    g = _gold(f"cfg1_d2_T16_s{seed}")
    with ugpack.Packer(2, 16, seed=seed, max_attempts=20000) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.attempts == 20000 and st.done
    assert list(idx) == g["idx"] and list(gid) == g["gid"]


@pytest.mark.parametrize("name", ["d2_T32_terms2_s1", "d2_T24_terms5_s3", "d2_T16_pin_s2",
                                  "d3_T12_s1", "d3_T16_terms4_s5"])
def test_variant_accepted_lists_equal_golden(gpu, name):
    import ugpack

    print(f"variant {name}", flush=True)
    g = _gold(name)
    with ugpack.Packer(g["dim"], g["T"], terms=g["terms"], seed=g["seed"],
                       pin_sin1=g["pin_sin1"], max_attempts=g["candidates"]) as p:
        p.run()
        idx, gid = p.accepted()
    assert list(idx) == g["idx"] and list(gid) == g["gid"]


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("name", ["d3_T256_s1", "d2_T512_s2"])
def test_long_scan_accepted_lists_equal_golden(gpu, name, precision):
    """3+1 T=256 and 2+1 T=512: the FP64 test's first pass walks all T timesteps in head
    rounds (37 of them at 3+1 T=256, where a doubling round size once overflowed and the
    last timesteps went unchecked), with two and four slot pairs per lane."""
    import ugpack

    print(f"long scan {name} {precision}", flush=True)
    g = _gold(name)
    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    with ugpack.Packer(g["dim"], g["T"], terms=g["terms"], seed=g["seed"],
                       max_attempts=g["candidates"], precision=prec) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.attempts == g["candidates"]
    assert list(idx) == g["idx"] and list(gid) == g["gid"]


def test_run_in_pieces_equals_one_run(gpu):
    import ugpack

    g = _gold("cfg1_d2_T16_s3")
    with ugpack.Packer(2, 16, seed=3, max_attempts=20000) as p:
        for _ in range(7):
            p.run(max_candidates=3001)
        p.run()
        idx, _ = p.accepted()
    assert list(idx) == g["idx"]


@pytest.mark.parametrize("seed", [1, 2])
def test_config2_pack_to_jam_equals_golden(gpu, seed):
    import ugpack

    print(f"cfg2 jam seed={seed}", flush=True)
    g = _gold(f"cfg2_d2_T64_s{seed}_jam")
    with ugpack.Packer(2, 64, seed=seed, stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.done and st.attempts == g["candidates"]
    assert list(idx) == g["idx"] and list(gid) == g["gid"]


def test_positions_and_params_of_accepted_paths(gpu):
    import ugpack

    m = po.Model(3, 12, seed=1)
    with ugpack.Packer(3, 12, seed=1, max_attempts=20000) as p:
        p.run()
        idx, _ = p.accepted()
        pos = p.positions().cpu().numpy()
        par = p.params()
    want = np.stack([m.positions(int(i)) for i in idx])
    assert np.array_equal(pos.view(np.uint64), want.view(np.uint64))
    for r, i in enumerate(idx[:20]):
        cd = m.candidate(int(i))
        for a in range(3):
            assert par[r, a, 0] == cd.a1[a]
            assert np.array_equal(par[r, a, 1::3], cd.amp[a])
            assert np.array_equal(par[r, a, 2::3], cd.b[a].astype(float))
            assert np.array_equal(par[r, a, 3::3], cd.phase[a])


def test_max_n(gpu):
    import ugpack

    ref_idx, _, att = po.Model(2, 16, seed=2).pack(20000, max_n=10)
    with ugpack.Packer(2, 16, seed=2, max_n=10) as p:
        st = p.run()
        idx, _ = p.accepted()
    assert st.n == 10 and list(idx) == list(ref_idx) and st.attempts == att


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_max_n_in_the_pipelined_steady_state(gpu, precision):
    """max_n reached in a full-size sub-batch, while the next sub-batch's filter runs ahead:
    the run stops at the max_n-th accept of the golden jam list, as serial RSA does."""
    import ugpack

    g = _gold("cfg2_d2_T64_s1_jam")
    k = len(g["idx"]) - 5
    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    with ugpack.Packer(2, 64, seed=1, stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4, max_n=k,
                       precision=prec) as p:
        st = p.run()
        idx, _ = p.accepted()
    assert g["idx"][k - 1] >= 1 << 20  # the cut falls after the first full-size sub-batch
    assert st.done and st.n == k and list(idx) == g["idx"][:k]
    assert st.attempts == g["idx"][k - 1] + 1


def test_recycled_buffers_and_cache_trim(gpu):
    """Handles reuse the buffers of destroyed ones (other models in between, the cache freed
    midway): every pack still equals its golden list."""
    import ugpack

    g1, g2 = _gold("cfg1_d2_T16_s1"), _gold("d3_T12_s1")
    for rep in range(3):
        with ugpack.Packer(2, 16, seed=1, max_attempts=20000) as p:
            p.run()
            assert list(p.accepted()[0]) == g1["idx"]
        with ugpack.Packer(3, 12, seed=1, max_attempts=g2["candidates"]) as p:
            p.run()
            assert list(p.accepted()[0]) == g2["idx"]
        if rep == 1:
            ugpack.abi.cache_trim()


def test_profiling_selects_the_timed_kernels(gpu):
    """ug_pack_profile times only the selected kernels, counts in all of them, never changes
    the accepted list, and rejects values outside the UG_PROFILE_* bits."""
    import ugpack
    from ugpack import abi

    g = _gold("cfg1_d2_T16_s1")
    seen = {}
    for kernels in (abi.PROFILE_PREFILTER, abi.PROFILE_RECHECK, abi.PROFILE_ALL):
        with ugpack.Packer(2, 16, seed=1, max_attempts=20000, precision=ugpack.PREC_SPLIT_BF16) as p:
            p.profile(True, kernels=kernels)
            p.run()
            assert list(p.accepted()[0]) == g["idx"]
            pre, re = p.kernel_stats(abi.KERNEL_PREFILTER), p.kernel_stats(abi.KERNEL_RECHECK)
            seen[kernels] = (pre, re)
            assert (pre["launches"] > 0) == bool(kernels & abi.PROFILE_PREFILTER)
            assert (re["launches"] > 0) == bool(kernels & abi.PROFILE_RECHECK)
            assert pre["visits"] > 0 and re["items"] > 0  # the counters run either way
    # the counters do not depend on which launches are timed
    counts = {k: (v[0]["visits"], v[0]["points"], v[1]["items"], v[1]["visits"]) for k, v in seen.items()}
    assert len(set(counts.values())) == 1
    with ugpack.Packer(2, 16, seed=1, max_attempts=100) as p:
        for bad in (-1, (abi.PROFILE_ALL | abi.PROFILE_REASONS) + 1):
            with pytest.raises(ugpack.UgError, match="UG_PROFILE"):
                abi.check(abi.lib().ug_pack_profile(p._h, bad))


@pytest.mark.parametrize("d,T,cands", [(2, 64, 3 << 20), (3, 32, 1 << 20)])
def test_prefilter_reasons_account_for_every_recheck(gpu, d, T, cands):
    """With UG_PROFILE_REASONS the prefilter sorts every candidate it leaves to the FP64 test
    by reason; with one prefilter stage (T <= 160) they are exactly the recheck's items, and
    counting them changes neither the lists nor the other counters."""
    import ugpack
    from ugpack import abi

    print(f"reasons d={d} T={T}", flush=True)
    out = {}
    for kernels in (abi.PROFILE_ALL, abi.PROFILE_ALL | abi.PROFILE_REASONS):
        with ugpack.Packer(d, T, seed=3, max_attempts=cands, precision=ugpack.PREC_SPLIT_BF16) as p:
            p.profile(True, kernels=kernels)
            p.run()
            pre, re = p.kernel_stats(abi.KERNEL_PREFILTER), p.kernel_stats(abi.KERNEL_RECHECK)
            out[kernels] = (list(p.accepted()[0]), pre, re)
    (la, pa, ra), (lb, pb, rb) = out[abi.PROFILE_ALL], out[abi.PROFILE_ALL | abi.PROFILE_REASONS]
    assert la == lb
    assert (pa["visits"], pa["points"], ra["items"]) == (pb["visits"], pb["points"], rb["items"])
    assert pa["left_wsum"] == pa["left_possible"] == pa["left_none"] == 0
    assert pb["left_wsum"] + pb["left_possible"] + pb["left_none"] == rb["items"] > 0


def test_invalid_config_is_an_error_not_an_exit(gpu):
    import ugpack

    with pytest.raises(ugpack.UgError):
        ugpack.Packer(2, 16, terms=40)  # exceeds the frequency pool, like the engine's guard
    with pytest.raises(ugpack.UgError):
        ugpack.Packer(4, 16)


# ------------------------------------------------------------------ split-BF16 form
@pytest.mark.parametrize("d,T", [(2, 16), (2, 64), (3, 12), (3, 128)])
def test_split_positions_within_margin(gpu, d, T):
    """The MFMA split-BF16 positions stay within the certainty margin of the canonical
    FP64 positions on the timesteps the prefilter examines (its first 32 in scan order)."""
    import torch
    import ugpack

    print(f"split eval d={d} T={T}", flush=True)
    idx = torch.arange(0, 4096, dtype=torch.int64) * 7919 + 3
    with ugpack.Packer(d, T, seed=4, precision=ugpack.PREC_SPLIT_BF16) as p:
        margin = p.status().margin
        x16 = p.eval_candidates(idx, ugpack.PREC_SPLIT_BF16).cpu().numpy()
        x64 = p.eval_candidates(idx, ugpack.PREC_FP64).cpu().numpy()
    order = po.Model(d, T).order[: min(T, 32)]
    diff = np.abs(x16 - x64)
    diff = np.minimum(diff, 2.0 - diff)
    err = diff[:, order, :].max()
    print(f"max |X_splitbf16 - X_fp64| = {err:.3e} (all t: {diff.max():.3e}), margin {margin:.3e}")
    assert 0 < margin < 2e-3
    assert err < margin


def test_split_generation_error_inside_margin_budget(gpu):
    """The prefilter's FP32 generation (hardware sin/cos of the phases, v_log of the Dirichlet
    words) measured exhaustively on the device: with W >= nw/4 (smaller sums go to the FP64
    test) a coefficient's error is at most (e_phase + 5 e_expo) times the unit budget, and the
    certainty margin allots 2^-20 to it."""
    import ugpack

    print("split generation error (all 2^32 phase words)", flush=True)
    e_phase, e_expo = ugpack.abi.split_generation_error()
    print(f"phase {e_phase:.3e}, expo {e_expo:.3e}: {(e_phase + 5 * e_expo) / 2**-20:.3f} of 2^-20")
    assert 0 < e_phase < 2.5e-7
    assert 0 < e_expo
    assert e_phase + 5 * e_expo <= 2**-20


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_split_config1_equals_golden(gpu, seed):
    import ugpack

    g = _gold(f"cfg1_d2_T16_s{seed}")
    with ugpack.Packer(2, 16, seed=seed, max_attempts=20000, precision=ugpack.PREC_SPLIT_BF16) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert list(idx) == g["idx"] and list(gid) == g["gid"] and st.attempts == 20000


@pytest.mark.parametrize("name", ["d3_T12_s1", "d2_T24_terms5_s3", "d2_T16_pin_s2"])
def test_split_variants_equal_golden(gpu, name):
    import ugpack

    g = _gold(name)
    with ugpack.Packer(g["dim"], g["T"], terms=g["terms"], seed=g["seed"], pin_sin1=g["pin_sin1"],
                       max_attempts=g["candidates"], precision=ugpack.PREC_SPLIT_BF16) as p:
        p.run()
        idx, gid = p.accepted()
    assert list(idx) == g["idx"] and list(gid) == g["gid"]


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_split_config2_pack_to_jam_equals_golden(gpu, seed):
    import ugpack

    print(f"split cfg2 jam seed={seed}", flush=True)
    g = _gold(f"cfg2_d2_T64_s{seed}_jam")
    with ugpack.Packer(2, 64, seed=seed, precision=ugpack.PREC_SPLIT_BF16,
                       stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.done and st.attempts == g["candidates"]
    assert list(idx) == g["idx"] and list(gid) == g["gid"]
    assert st.rechecked < st.attempts


# ------------------------------------------------------------------ 3+1 at config-3 scale
@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_config3_prefix_equals_golden(gpu, precision):
    """3+1, F = T = 128: the first 2^20 candidates of seed 1 (N = 39,810 accepts)."""
    import ugpack

    print(f"cfg3 prefix {precision}", flush=True)
    g = _gold("cfg3_d3_T128_s1_prefix1M")
    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    with ugpack.Packer(3, 128, seed=1, max_attempts=g["candidates"], precision=prec) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.attempts == g["candidates"] and st.n == g["n"]
    assert list(idx) == g["idx"] and list(gid) == g["gid"]


def _sub_reference(d, T, seed, unique_cands, sub_budget, terms=0, accept_rate=0.0, fill_rate=0.0):
    """Unique phase then subpath phase on the oracle."""
    m = po.Model(d, T, terms=terms or None, seed=seed)
    idx, gid, att = m.pack(unique_cands)
    pos = np.stack([m.positions(int(i)) for i in idx])
    sidx, sgid, satt, fresh = m.pack_sub(pos, gid, first=att, budget=sub_budget,
                                         accept_rate=accept_rate, fill_rate=fill_rate)
    return list(idx) + list(sidx), list(gid) + list(sgid), att, satt


@pytest.mark.parametrize("d,T,seed,unique,sub,precision", [
    (2, 16, 1, 20000, 200000, "fp64"), (2, 16, 4, 3000, 400000, "fp64"), (2, 32, 2, 30000, 300000, "fp64"),
    (3, 12, 1, 5000, 100000, "fp64"), (2, 64, 3, 200000, 200000, "fp64"),
    (2, 16, 4, 3000, 400000, "split"), (2, 64, 3, 200000, 200000, "split"), (3, 16, 2, 8000, 150000, "split"),
])
def test_subpath_phase_equals_oracle(gpu, d, T, seed, unique, sub, precision):
    import ugpack

    print(f"subpaths d={d} T={T} seed={seed} unique={unique} sub={sub} {precision}", flush=True)
    want_idx, want_gid, att, satt = _sub_reference(d, T, seed, unique, sub)
    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    with ugpack.Packer(d, T, seed=seed, max_attempts=unique, precision=prec) as p:
        st = p.run()
        assert st.attempts == att
        st = p.run_sub(max_attempts=sub)
        idx, gid = p.accepted()
    assert st.sub_done and st.attempts == att + satt
    assert st.n + st.nsub == len(want_idx)
    assert list(idx) == want_idx and list(gid) == want_gid


def test_subpath_phase_in_pieces(gpu):
    import ugpack

    want_idx, want_gid, att, satt = _sub_reference(2, 16, 3, 8000, 500000)
    with ugpack.Packer(2, 16, seed=3, max_attempts=8000) as p:
        p.run()
        for _ in range(5):
            p.run_sub(max_attempts=500000, max_candidates=40000)
        st = p.run_sub(max_attempts=500000)
        idx, gid = p.accepted()
    assert st.sub_done and st.attempts == att + satt
    assert list(idx) == want_idx and list(gid) == want_gid


def test_subpath_phase_split_precision(gpu):
    import ugpack

    want_idx, want_gid, _, _ = _sub_reference(2, 32, 5, 20000, 200000)
    with ugpack.Packer(2, 32, seed=5, max_attempts=20000, precision=ugpack.PREC_SPLIT_BF16) as p:
        p.run()
        p.run_sub(max_attempts=200000)
        idx, gid = p.accepted()
    assert list(idx) == want_idx and list(gid) == want_gid


def test_subpath_phase_rules(gpu):
    import ugpack

    with ugpack.Packer(2, 16, seed=1, max_attempts=2000) as p:
        p.run()
        with pytest.raises(ugpack.UgError):
            p.run_sub(sub_terms=3)  # --sub-terms other than --terms is unsupported
        st = p.run_sub(max_attempts=1000)
        assert st.sub_done and st.attempts == 2000 + (1 << 16)
        with pytest.raises(ugpack.UgError):
            p.run()  # the unique phase ended where the subpath phase started


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("name", ["sub_d2_T16_s1_accept2e-3", "sub_d2_T16_s2_fill3e-5"])
def test_subpath_windowed_stop_equals_golden(gpu, name, precision):
    import ugpack

    print(f"subpath window {name} {precision}", flush=True)
    g = _gold(name)
    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    with ugpack.Packer(g["dim"], g["T"], seed=g["seed"], max_attempts=g["unique_candidates"],
                       precision=prec) as p:
        p.run()
        st = p.run_sub(max_attempts=1e15, accept_rate=g["accept_rate"], fill_rate=g["fill_rate"])
        idx, gid = p.accepted()
    assert st.sub_done and st.n == g["n"] and st.nsub == g["nsub"]
    assert st.attempts == g["attempts"]
    sidx, sgid = list(idx[st.n:]), list(gid[st.n:])
    assert sidx[:64] == g["head_idx"] and sgid[:64] == g["head_gid"]
    assert sidx[-64:] == g["tail_idx"] and sgid[-64:] == g["tail_gid"]
    h = hashlib.sha256()
    h.update(np.asarray(idx, "<i8").tobytes())
    h.update(np.asarray(gid, "<i4").tobytes())
    assert h.hexdigest() == g["digest"]


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("d,T,seed,cands", [(3, 12, 1, 20000), (2, 16, 3, 20000), (3, 32, 2, 60000),
                                            (2, 64, 4, 400000)])
def test_euclid_collision_equals_oracle(gpu, precision, d, T, seed, cands):
    import ugpack

    print(f"euclid {precision} d={d} T={T} seed={seed}", flush=True)
    want_idx, _, att = po.Model(d, T, seed=seed, euclid=True).pack(cands)
    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    with ugpack.Packer(d, T, seed=seed, max_attempts=cands, precision=prec, euclid=True) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.attempts == att
    assert list(idx) == list(want_idx) and list(gid) == list(range(len(idx)))


_UNEVEN_REF = {}


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("d,T,seed,cands,euclid", [
    (3, 161, 7, 12000, False),  # split form: stage 2 on (6 column blocks, the last one column wide)
    (2, 161, 8, 20000, False),
    (3, 170, 3, 6000, True),    # stage 2 with the L2 ball
    (2, 33, 4, 20000, False),   # two column blocks, the second one column wide
    (3, 33, 5, 8000, False),
    (2, 100, 6, 15000, False),
    (3, 65, 2, 6000, False),
])
def test_uneven_column_blocks_equal_oracle(gpu, precision, d, T, seed, cands, euclid):
    """T not a multiple of 32: the prefilter's last block of 32 timesteps is partial, and past
    T = 160 the split form's second stage walks it.  The accepted lists equal the oracle's
    serial RSA (about 10 s of oracle work for all cases, shared by both forms)."""
    import ugpack

    print(f"uneven {precision} d={d} T={T} seed={seed} euclid={euclid}", flush=True)
    key = (d, T, seed, cands, euclid)
    if key not in _UNEVEN_REF:
        _UNEVEN_REF[key] = po.Model(d, T, seed=seed, euclid=euclid).pack(cands)
    want_idx, _, att = _UNEVEN_REF[key]
    prec = ugpack.PREC_FP64 if precision == "fp64" else ugpack.PREC_SPLIT_BF16
    with ugpack.Packer(d, T, seed=seed, max_attempts=cands, precision=prec, euclid=euclid) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.attempts == att
    assert list(idx) == list(want_idx) and list(gid) == list(range(len(idx)))


def test_characterization_check_on_config1(gpu):
    """ug_pack_check restates the oracle's brute-force characterization on the device."""
    import ugpack

    g = _gold("cfg1_d2_T16_s1")
    with ugpack.Packer(2, 16, seed=1, max_attempts=20000) as p:
        p.run()
        c = p.check(0, 20000)
    n = len(g["idx"])
    assert c == {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": n, "rejects": 20000 - n}


def _check_whole_range(p, attempts, chunk=1 << 26):
    """ug_pack_check over every consumed candidate, in chunks; the summed counts."""
    tot = {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": 0, "rejects": 0}
    for first in range(0, attempts, chunk):
        c = p.check(first, min(chunk, attempts - first))
        for k in ("unexplained_rejects", "rejects"):
            tot[k] += c[k]
        tot["colliding_uniques"] = c["colliding_uniques"]
        tot["uniques"] = c["uniques"]
    return tot


@pytest.mark.parametrize("precision", ["split", "fp64"])
def test_config3_to_jam_satisfies_serial_rsa_at_full_size(gpu, precision):
    """3+1 T=128 packed to jamming (~5.6e7 candidates): no two of the ~6.4e4 accepted paths
    collide, and every rejected candidate of the whole consumed range touches a path
    accepted before it."""
    import ugpack

    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    print(f"cfg3 jam check {precision}", flush=True)
    with ugpack.Packer(3, 128, seed=1, precision=prec, stop_kind=ugpack.STOP_FIXED_WINDOW,
                       stop_rate=1e-4) as p:
        st = p.run()
        c = _check_whole_range(p, st.attempts)
        idx, _ = p.accepted()
    assert st.done and st.n > 50000
    g = _gold("cfg3_d3_T128_s1_prefix1M")
    assert list(idx[idx < (1 << 20)]) == g["idx"]
    assert c == {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": st.n,
                 "rejects": st.attempts - st.n}


def test_config5_to_jam_satisfies_serial_rsa_at_full_size(gpu):
    """3+1 T=256 packed to jamming (~3.3e8 candidates, ~3.8e5 paths) in both forms: the same
    accepted list, and both properties over every accepted path and every consumed
    candidate."""
    import ugpack

    lists = {}
    for precision, prec in (("split", ugpack.PREC_SPLIT_BF16), ("fp64", ugpack.PREC_FP64)):
        print(f"cfg5 jam check {precision}", flush=True)
        with ugpack.Packer(3, 256, seed=1, precision=prec,
                           stop_kind=ugpack.STOP_FIXED_WINDOW, stop_rate=1e-4) as p:
            st = p.run()
            c = _check_whole_range(p, st.attempts)
            lists[precision], _ = p.accepted()
        assert st.done and st.n > 300000
        assert c == {"colliding_uniques": 0, "unexplained_rejects": 0, "uniques": st.n,
                     "rejects": st.attempts - st.n}
    assert np.array_equal(lists["split"], lists["fp64"]), "the two forms accepted different lists"
