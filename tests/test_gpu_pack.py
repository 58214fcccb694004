"""GPU parity of the HIP packer against the CPU oracle, through the C ABI.

Bar: bit-exact.  Canonical FP64 positions must equal the oracle's bits; the accepted
index lists and group labels must equal serial RSA (oracle or committed golden fixture).
"""

from __future__ import annotations

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
    """3+1, F = T = 128: the first 2^20 candidates of seed 1 (N = 39,818 accepts)."""
    import ugpack

    print(f"cfg3 prefix {precision}", flush=True)
    g = _gold("cfg3_d3_T128_s1_prefix1M")
    prec = ugpack.PREC_SPLIT_BF16 if precision == "split" else ugpack.PREC_FP64
    with ugpack.Packer(3, 128, seed=1, max_attempts=g["candidates"], precision=prec) as p:
        st = p.run()
        idx, gid = p.accepted()
    assert st.attempts == g["candidates"] and st.n == g["n"]
    assert list(idx) == g["idx"] and list(gid) == g["gid"]
