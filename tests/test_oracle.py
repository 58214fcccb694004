"""Pins the CPU oracle (oracle/) against things it did not compute itself.

- Philox4x32-10: Random123's published known-answer vectors and rocRAND's Philox.
- log / sin / cos: the C library (math.log, math.sin, math.cos) to a few ulp.
- basis tables and evaluation: the reference's direct expression
  X = [a1 sin z + sum a_j (sin(b_j z + f_j) - sin f_j)] / sin z (cuda/braid_cuda.cu:418-441,
  braidlab/corrdim.py:116-137) evaluated with math.sin, and its closed forms at z = pi/2
  (tests/test_corrdim.py:154-171 of the reference).
- proposal measure: exact invariants and distribution tests (cuda/braid_cuda.cu:295-373).
- torus wrap / min-image: the reference's own examples (tests/test_corrdim.py:138-145,
  tests/test_group_partners.py:65-82).
- serial RSA: the brute-force characterization (no two accepted collide; every rejected
  collides with an earlier accepted), which planted mistakes must fail.
"""

from __future__ import annotations

import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import stats

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ Philox4x32-10
def test_philox_random123_kat():
    # Random123 kat_vectors, philox4x32 10 rounds.
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(po.philox(ctr, key)) == want


_ROCRAND_PROBE = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct E : rocrand_device::philox4x32_10_engine {
    E() : rocrand_device::philox4x32_10_engine(0, 0, 0) {}
    uint4 go(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    E e;
    unsigned s = 12345u;
    for (int i = 0; i < 200; i++) {
        unsigned v[6];
        for (int j = 0; j < 6; j++) { s = s * 1664525u + 1013904223u; v[j] = s; }
        uint4 r = e.go(uint4{v[0], v[1], v[2], v[3]}, uint2{v[4], v[5]});
        printf("%u %u %u %u %u %u %u %u %u %u\n", v[0], v[1], v[2], v[3], v[4], v[5], r.x, r.y, r.z, r.w);
    }
}
"""


def test_philox_matches_rocrand(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "probe.cpp"
    src.write_text(_ROCRAND_PROBE)
    exe = tmp_path / "probe"
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-o", str(exe), "-x", "hip", str(src)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("rocRAND host probe did not compile: " + r.stderr[-300:])
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows = [list(map(int, ln.split())) for ln in out.strip().splitlines()]
    assert len(rows) == 200
    for row in rows:
        assert list(po.philox(row[0:4], row[4:6])) == row[6:10]


# ------------------------------------------------------------------ log, sinpi, cospi
def _ulp_diff(a: float, b: float) -> float:
    return abs(a - b) / math.ulp(max(abs(a), abs(b), 1e-300))


def test_log_against_libm():
    rng = np.random.default_rng(1)
    vals = list(rng.uniform(2.0**-31, 1.0, 3000)) + list(2.0 ** rng.uniform(-31, 0, 3000))
    vals += [1.0, 0.5, 2.0**-31, 1 - 2.0**-31, 0.7071067811865476, 0.7071067811865475]
    for v in vals:
        got, want = po.log(v), math.log(v)
        if want == 0.0:
            assert got == 0.0
        else:
            assert _ulp_diff(got, want) <= 2.0, (v, got, want)


def test_sincospi_against_libm_and_exact_points():
    import mpmath

    mpmath.mp.dps = 40
    rng = np.random.default_rng(2)
    for _ in range(1500):
        m = int(rng.integers(1, 2**32))
        n = int(rng.integers(0, 4 * m))
        s, c = po.sincospi_rat(n, m)
        x = mpmath.pi * mpmath.mpf(n) / m
        ws, wc = float(mpmath.sin(x)), float(mpmath.cos(x))
        assert abs(s - ws) <= 2 * math.ulp(max(abs(ws), 2**-60)) + 1e-300, (n, m)
        assert abs(c - wc) <= 2 * math.ulp(max(abs(wc), 2**-60)) + 1e-300, (n, m)
    # exact angles
    assert po.sincospi_rat(0, 7) == (0.0, 1.0)
    assert po.sincospi_rat(1, 2) == (1.0, -0.0) or po.sincospi_rat(1, 2)[0] == 1.0
    s, c = po.sincospi_rat(1, 6)
    assert abs(s - 0.5) <= 1.2e-16 and abs(c - math.sqrt(3) / 2) <= 2.3e-16
    s, c = po.sincospi_rat(1, 4)
    assert abs(s - math.sqrt(0.5)) <= 1.2e-16 and abs(c - math.sqrt(0.5)) <= 1.2e-16


# ------------------------------------------------------------------ basis and evaluation
@pytest.mark.parametrize("T", [16, 64, 101])
def test_basis_tables_match_direct_trig(T):
    m = po.Model(2, T)
    z = (np.arange(T) + 1) * np.pi / (T + 1)
    assert np.max(np.abs(m.sz - np.sin(z))) <= 4.5e-16
    import mpmath

    mpmath.mp.dps = 30
    S, Cc = m.S, m.C
    for t in range(T):
        zt = mpmath.pi * (t + 1) / (T + 1)
        for k in range(T + 1):
            s_ref = float(mpmath.sin(k * zt) / mpmath.sin(zt))
            c_ref = float((mpmath.cos(k * zt) - 1) / mpmath.sin(zt))
            assert abs(S[t, k] - s_ref) <= 4e-15 * max(1.0, abs(s_ref)), (t, k)
            assert abs(Cc[t, k] - c_ref) <= 4e-15 * max(1.0, abs(c_ref)), (t, k)
    o = m.order
    assert sorted(o) == list(range(T))
    assert list(o[:4]) == [0, T - 1, 1, T - 2]
    keys = np.minimum(o, T - 1 - o)
    assert np.all(np.diff(keys) >= 0)


def _direct_positions(cd, z, wrap=True):
    """The reference's per-term expression with libm sin (cuda/braid_cuda.cu:422-441)."""
    out = []
    for a in range(cd.a1.shape[0]):
        x = cd.a1[a] * np.sin(z)
        for j in range(cd.amp.shape[1]):
            x = x + cd.amp[a, j] * (np.sin(cd.b[a, j] * z + cd.phase[a, j]) - np.sin(cd.phase[a, j]))
        X = x / np.sin(z)
        out.append(X - 2.0 * np.floor((X + 1.0) * 0.5) if wrap else X)
    return np.stack(out, axis=-1)


@pytest.mark.parametrize("d,T,terms", [(2, 16, 16), (2, 64, 64), (3, 32, 32), (2, 40, 6), (3, 20, 2)])
def test_eval_matches_reference_expression(d, T, terms):
    m = po.Model(d, T, terms=terms, seed=7)
    z = (np.arange(T) + 1) * np.pi / (T + 1)
    worst = 0.0
    for idx in range(0, 4000, 97):
        cd = m.candidate(idx)
        X = m.positions(idx)
        ref = _direct_positions(cd, z)
        dX = np.abs(X - ref)
        dX = np.minimum(dX, 2.0 - dX)  # same point across the seam
        worst = max(worst, float(dX.max()))
    assert worst < 1e-11, worst


def test_eval_closed_form_at_half_pi():
    # T odd: t = (T-1)/2 sits at z = pi/2 where sin z = 1 (reference tests/test_corrdim.py:154-171)
    T = 31
    m = po.Model(2, T, seed=3)
    t = (T - 1) // 2
    for idx in (0, 5, 99):
        cd = m.candidate(idx)
        X = m.positions(idx)[t]
        for a in range(2):
            x = cd.a1[a] + sum(cd.amp[a, j] * (math.sin(cd.b[a, j] * math.pi / 2 + cd.phase[a, j])
                                               - math.sin(cd.phase[a, j])) for j in range(m.nw))
            want = x - 2.0 * math.floor((x + 1.0) * 0.5)
            assert abs(X[a] - want) < 1e-12


# ------------------------------------------------------------------ proposal measure
def test_candidate_invariants_full_spectrum():
    for d, T in ((2, 16), (2, 64), (3, 128)):
        m = po.Model(d, T, seed=5)
        for idx in range(0, 300, 7):
            cd = m.candidate(idx)
            assert np.all(np.abs(cd.a1) < 1.0)
            for a in range(d):
                assert sorted(cd.b[a]) == list(range(2, T + 1))
                assert abs(np.sum(np.abs(cd.amp[a]) * cd.b[a]) - 1.0) < 1e-13
                odd = cd.b[a] % 2 == 1
                assert np.all(cd.phase[a][odd] == 0.0) and np.all(cd.s[a][odd] == 0.0)
                assert np.all((cd.phase[a][~odd] >= 0) & (cd.phase[a][~odd] < np.pi))
                assert np.allclose(cd.c[a], cd.amp[a] * np.cos(cd.phase[a]), rtol=0, atol=1e-16)
                assert np.allclose(cd.s[a], cd.amp[a] * np.sin(cd.phase[a]), rtol=0, atol=1e-16)


def test_candidate_invariants_partial_and_pin():
    m = po.Model(2, 50, terms=7, seed=9)
    for idx in range(200):
        cd = m.candidate(idx)
        for a in range(2):
            b = cd.b[a]
            assert len(set(b)) == 6 and b.min() >= 2 and b.max() <= 50
            assert abs(np.sum(np.abs(cd.amp[a]) * b) - 1.0) < 1e-14
    legacy = po.Model(3, 30, terms=2, seed=1)
    for idx in range(100):
        cd = legacy.candidate(idx)
        assert np.allclose(np.abs(cd.amp[:, 0]) * cd.b[:, 0], 1.0)
    pinned = po.Model(2, 16, seed=4, pin=True)
    assert np.all(pinned.candidate(3).a1 == 0.0)


def _words(seed, idx, a, n):
    """The first n stage-1 words of axis a: components of consecutive Philox blocks."""
    key = [seed & 0xFFFFFFFF, seed >> 32]
    out = []
    for k in range((n + 3) // 4):
        out.extend(int(x) for x in po.philox([idx & 0xFFFFFFFF, idx >> 32, k, (1 << 8) | a], key))
    return out[:n]


def test_slot_word_layout():
    # full pool: pair p reads words 3p, 3p+1 (Dirichlet words of slots 2p, 2p+1) and 3p+2 (the
    # even slot's phase) of the axis' stream; partial spectra: block p whole, phases in z, w.
    # A layout slip moves the spacing shares and phases, checked here against the words.
    def shares(words):
        e = np.array([-np.log((0x80000000 - (w >> 1)) * 2.0**-31) for w in words])
        return e / e.sum(), np.array([w & 1 for w in words])
    for d, T, seed, idx in ((2, 16, 5, 777), (3, 12, 2, 31), (2, 64, 1, 123456)):
        m = po.Model(d, T, seed=seed)
        cd = m.candidate(idx)
        nw = m.nw
        assert nw == T - 1
        npairs = (nw + 1) // 2
        for a in range(d):
            w = _words(seed, idx, a, 3 * npairs)
            dw = [w[3 * (j // 2) + (j % 2)] for j in range(nw)]
            sh, neg = shares(dw)
            assert np.allclose(np.abs(cd.amp[a]) * cd.b[a], sh, rtol=1e-12, atol=0)
            assert np.array_equal(cd.amp[a] < 0, neg == 1)
            for j in range(0, nw, 2):
                assert cd.phase[a][j] == w[3 * (j // 2) + 2] * 2.0**-32 * np.pi
    m = po.Model(2, 24, terms=5, seed=3)
    cd = m.candidate(99)
    nw = m.nw
    for a in range(2):
        w = _words(3, 99, a, 4 * ((nw + 1) // 2))
        dw = [w[4 * (j // 2) + (j % 2)] for j in range(nw)]
        sh, neg = shares(dw)
        assert np.allclose(np.abs(cd.amp[a]) * cd.b[a], sh, rtol=1e-12, atol=0)
        for j in range(nw):
            if cd.b[a][j] % 2 == 0:
                assert cd.phase[a][j] == w[4 * (j // 2) + 2 + (j % 2)] * 2.0**-32 * np.pi


def test_candidate_distributions():
    d, T, n = 2, 32, 3000
    m = po.Model(d, T, seed=21)
    a1, share, ph, neg = [], [], [], []
    for idx in range(n):
        cd = m.candidate(idx)
        a1.extend(cd.a1)
        w = np.abs(cd.amp[0]) * cd.b[0]
        share.append(w[5])
        even = cd.b[0] % 2 == 0
        ph.extend(cd.phase[0][even][:3])
        neg.extend(cd.amp[0] < 0)
    assert stats.kstest(a1, "uniform", args=(-1, 2)).pvalue > 1e-3
    nw = T - 1
    assert stats.kstest(share, "beta", args=(1, nw - 1)).pvalue > 1e-3
    assert stats.kstest(np.array(ph) / np.pi, "uniform").pvalue > 1e-3
    assert abs(np.mean(neg) - 0.5) < 0.01


def test_candidates_are_counter_addressed():
    m = po.Model(2, 24, seed=3)
    a = m.candidate(12345)
    b = m.candidate(12345)
    c = po.Model(2, 24, seed=4).candidate(12345)
    assert np.array_equal(a.c, b.c) and not np.array_equal(a.c, c.c)


# ------------------------------------------------------------------ the reference's stream
def test_xoshiro_published_vectors():
    # splitmix64 from 0 (its published first outputs) seeds the state; xoshiro256** from the
    # state {1, 2, 3, 4} (the generator's published first ten outputs)
    assert [int(x) for x in po.xoshiro_seed(0)[:3]] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                       0x06C45D188009454F]
    st = np.array([1, 2, 3, 4], np.uint64)
    want = [11520, 0, 1509978240, 1215971899390074240, 1216172134540287360, 607988272756665600,
            16172922978634559625, 8476171486693032832, 10595114339597558777, 2904607092377533576]
    assert [po.xoshiro_next(st) for _ in range(10)] == want


def _xref_draws(seed, idx, d, nw, modmax, pin):
    """The knot tool's proposal (cuda/braid_knot.cu:124-166) as raw draws, in stream order,
    from the generator pinned above: per axis the frequencies, the spacing uniforms, per slot
    a sign and (even frequency) a phase uniform, then the sin1 uniform."""
    M64 = (1 << 64) - 1
    st = po.xoshiro_seed(seed ^ ((idx * 0x9E3779B97F4A7C15 + 0x1234567) & M64))
    nxt = lambda: po.xoshiro_next(st)
    axes = []
    for _ in range(d):
        if nw == modmax:
            b = list(range(2, nw + 2))
            for j in range(nw - 1, 0, -1):
                k = nxt() % (j + 1)
                b[j], b[k] = b[k], b[j]
        else:
            b = []
            for j in range(nw):
                for _g in range(64):
                    c = nxt() % modmax + 2
                    if c not in b:
                        break
                b.append(c)
        u = [(nxt() >> 11) * 2.0**-53 for _ in range(nw)]
        sign, ph = [], []
        for j in range(nw):
            sign.append(nxt() & 1)
            ph.append((nxt() >> 11) * 2.0**-53 if b[j] % 2 == 0 else 0.0)
        a1 = 0.0 if pin else (nxt() >> 11) * 2.0**-53 * 2 - 1
        axes.append((np.array(b), np.array(u), np.array(sign), np.array(ph), a1))
    return axes


@pytest.mark.parametrize("d,T,terms,pin,idx", [(2, 16, None, False, 5), (3, 12, None, True, 77),
                                               (2, 24, 5, False, 123456), (3, 64, None, False, 9)])
def test_xref_candidate_follows_the_stream_order(d, T, terms, pin, idx):
    # which draw feeds which quantity: a slip in the order (signs before spacings, the sin1
    # draw first as in the engines' proposal, phases for odd frequencies) moves these
    m = po.Model(d, T, terms=terms, seed=11, pin=pin, rng=1)
    cd = m.candidate(idx)
    for a, (b, u, sign, ph, a1) in enumerate(_xref_draws(11, idx, d, m.nw, max(T - 1, 2), pin)):
        assert np.array_equal(cd.b[a], b)
        e = -np.log1p(-u)
        assert np.allclose(np.abs(cd.amp[a]) * b, e / e.sum(), rtol=1e-13, atol=0)
        assert np.array_equal(cd.amp[a] < 0, sign == 1)
        assert np.allclose(cd.phase[a], ph * np.pi, rtol=1e-15, atol=0)
        assert cd.a1[a] == a1


def test_xref_candidate_invariants_and_distributions():
    for d, T in ((2, 16), (3, 32)):
        m = po.Model(d, T, seed=5, rng=1)
        for idx in range(0, 200, 9):
            cd = m.candidate(idx)
            assert np.all((cd.a1 >= -1.0) & (cd.a1 < 1.0))
            for a in range(d):
                assert sorted(cd.b[a]) == list(range(2, T + 1))
                assert abs(np.sum(np.abs(cd.amp[a]) * cd.b[a]) - 1.0) < 1e-13
                odd = cd.b[a] % 2 == 1
                assert np.all(cd.phase[a][odd] == 0.0) and np.all(cd.s[a][odd] == 0.0)
                assert np.allclose(cd.c[a], cd.amp[a] * np.cos(cd.phase[a]), rtol=0, atol=1e-16)
                assert np.allclose(cd.s[a], cd.amp[a] * np.sin(cd.phase[a]), rtol=0, atol=1e-16)
    assert np.all(po.Model(2, 16, seed=4, pin=True, rng=1).candidate(3).a1 == 0.0)
    # the measure: sin1 uniform on [-1, 1), Dirichlet shares, uniform phases, fair signs, and
    # the Fisher-Yates slot of a frequency uniform over the slots
    T, n = 32, 3000
    m = po.Model(2, T, seed=21, rng=1)
    a1, share, ph, neg, where2 = [], [], [], [], []
    for idx in range(n):
        cd = m.candidate(idx)
        a1.extend(cd.a1)
        share.append((np.abs(cd.amp[0]) * cd.b[0])[5])
        even = cd.b[0] % 2 == 0
        ph.extend(cd.phase[0][even][:3])
        neg.extend(cd.amp[0] < 0)
        where2.append(int(np.flatnonzero(cd.b[1] == 2)[0]))
    assert stats.kstest(a1, "uniform", args=(-1, 2)).pvalue > 1e-3
    assert stats.kstest(share, "beta", args=(1, T - 2)).pvalue > 1e-3
    assert stats.kstest(np.array(ph) / np.pi, "uniform").pvalue > 1e-3
    assert abs(np.mean(neg) - 0.5) < 0.01
    assert stats.chisquare(np.bincount(where2, minlength=T - 1)).pvalue > 1e-3


def test_xref_pack_characterization():
    # serial RSA on the reference's stream: no two accepted collide, every rejected candidate
    # collides with an earlier accepted one, and the list differs from the Philox stream's
    for d, T, terms, cands in ((2, 16, None, 6000), (3, 12, None, 5000), (2, 24, 5, 6000)):
        m = po.Model(d, T, terms=terms, seed=3, rng=1)
        idx, _, att = m.pack(cands)
        assert att == cands and len(idx) > 20
        assert m.check_rsa(cands, idx) == 0
        assert m.check_rsa(cands, idx[:-1]) != 0
        philox, _, _ = po.Model(d, T, terms=terms, seed=3).pack(cands)
        assert not np.array_equal(idx, philox)


# ------------------------------------------------------------------ wrap and predicate
def test_wrap_reference_examples():
    m = po.Model(2, 16)
    x = np.array([-2.0, -1.2, -1.0, -0.3, 0.0, 0.99, 1.0, 1.2, 1.9])
    w = x - 2.0 * np.floor((x + 1.0) * 0.5)
    assert np.all(w >= -1.0) and np.all(w < 1.0)
    assert np.allclose(w[3:6], x[3:6])
    # cell index wraps modulo T at both ends (cell width 2/T)
    assert m.cell(-1.0) == 0 and m.cell(0.999999) == 15 and m.cell(-1.0 - 1e-17) in (0, 15)
    assert m.cell(np.nextafter(-1.0, -2.0)) == 15


def test_rsa_predicate_reference_cases():
    # The reference's contact cases (tests/test_group_partners.py:65-82, T = 50, exclusion
    # 2/T), restated on the oracle's predicate with constant comoving positions.
    m = po.Model(2, 50)
    T = 50
    a = np.array([0.3, 0.3])
    # identical strands touch (from the first timestep)
    assert m.touches(a, a)
    # parked 1.0 apart on both axes: no contact, also with a small wiggle on each
    b = np.array([-0.7, -0.7])
    assert not m.touches(a, b)
    assert not m.touches(a + 0.01, b - 0.01) and not m.touches(a - 0.01, b + 0.01)
    # X = -0.999 against +0.999: the unwrapped gap 1.998 exceeds the cell, the wrapped gap
    # 0.002 does not -- contact, and the two cells are neighbours across the seam, so the
    # 3x3 cell scan finds it
    x, y = np.array([-0.999, 0.0]), np.array([0.999, 0.0])
    assert m.touches(x, y) and m.touches(y, x)
    assert (m.cell(-0.999), m.cell(0.999)) == (0, T - 1)
    # the boundary of the inclusive cube (T = 16, r = 1/8, exact in binary): exactly r apart on
    # both axes touches, a step beyond does not
    m16 = po.Model(2, 16)
    p0 = np.array([0.25, 0.5])
    assert m16.touches(p0, np.array([0.375, 0.375]))
    assert not m16.touches(p0, np.array([0.25, np.nextafter(0.625, 1.0)]))
    # and on a real packing no two accepted paths touch at any timestep
    mp = po.Model(2, 16, seed=1)
    idx, _, _ = mp.pack(5000)
    X = mp.positions_of(idx[:12])
    for i in range(len(X)):
        for j in range(i):
            assert not any(mp.touches(X[i][t], X[j][t]) for t in range(16))


# ------------------------------------------------------------------ serial RSA
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_config1_golden_and_characterization(seed):
    g = json.load(open(os.path.join(GOLD, f"cfg1_d2_T16_s{seed}.json")))
    m = po.Model(2, 16, seed=seed)
    idx, gid, att = m.pack(20000)
    assert att == 20000
    assert list(idx) == g["idx"] and list(gid) == g["gid"]
    assert list(gid) == list(range(len(idx)))
    assert m.check_rsa(20000, idx) == 0


@pytest.mark.parametrize("name", ["d2_T32_terms2_s1", "d2_T24_terms5_s3", "d2_T16_pin_s2",
                                  "d3_T12_s1", "d3_T16_terms4_s5"])
def test_golden_variants(name):
    g = json.load(open(os.path.join(GOLD, name + ".json")))
    m = po.Model(g["dim"], g["T"], terms=g["terms"], seed=g["seed"], pin=g["pin_sin1"])
    idx, gid, att = m.pack(g["candidates"])
    assert list(idx) == g["idx"]
    assert m.check_rsa(min(g["candidates"], 8000), idx[idx < 8000]) == 0


def test_characterization_rejects_planted_mistakes():
    m = po.Model(2, 16, seed=1)
    idx, _, _ = m.pack(6000)
    assert m.check_rsa(6000, idx) == 0
    assert m.check_rsa(6000, idx[:-1]) != 0            # a missing accept
    assert m.check_rsa(6000, np.delete(idx, 3)) != 0   # a missing early accept
    rejected = np.setdiff1d(np.arange(6000), idx)[5]
    bad = np.sort(np.append(idx, rejected))
    assert m.check_rsa(6000, bad) != 0                 # an extra (colliding) accept


@pytest.mark.parametrize("d,T,seed,cands,threads", [(2, 16, 1, 20000, 3), (3, 12, 1, 20000, 8),
                                                    (2, 24, 3, 30000, 2)])
def test_multicore_pack_equals_serial(d, T, seed, cands, threads):
    m = po.Model(d, T, seed=seed)
    want, _, att = m.pack(cands)
    got, att2 = m.pack_mt(cands, threads)
    assert list(got) == list(want) and att2 == att == cands


def test_multicore_pack_to_jam_equals_golden():
    g = json.load(open(os.path.join(GOLD, "cfg2_d2_T64_s2_jam.json")))
    got, att = po.Model(2, 64, seed=2).pack_mt(1 << 40, 8, stop_kind=1)
    assert att == g["candidates"] and list(got) == g["idx"]


def test_euclid_predicate_cases():
    # --euclid-collision: the L2 ball of radius r replaces the Chebyshev cube
    for d in (2, 3):
        cube, ball = po.Model(d, 16), po.Model(d, 16, euclid=True)
        r = cube.r
        o = np.zeros(d)
        on_axis = np.zeros(d)
        on_axis[0] = r                      # distance exactly r: contact in both
        diag = np.full(d, 0.9 * r)          # inside the cube, outside the ball
        seam = np.zeros(d)
        seam[0], o2 = 0.999, np.zeros(d)
        o2[0] = -0.999                      # 0.002 apart across the seam
        for m in (cube, ball):
            assert m.touches(o, on_axis) and m.touches(seam, o2)
        assert cube.touches(o, diag) and not ball.touches(o, diag)
        inner = np.full(d, 0.99 * r / np.sqrt(d))
        assert ball.touches(o, inner)


@pytest.mark.parametrize("d,T,seed", [(3, 12, 1), (2, 16, 3)])
def test_euclid_rsa_by_numpy_brute_force(d, T, seed):
    # serial RSA under the L2 ball, restated with numpy norms over all candidate pairs
    m = po.Model(d, T, seed=seed, euclid=True)
    M = 2500
    idx, _, _ = m.pack(M)
    X = np.stack([m.positions(c) for c in range(M)])       # (M, T, d)
    acc = []
    for c in range(M):
        if acc:
            dd = np.abs(X[c][None] - X[acc])
            dd = np.minimum(dd, 2.0 - dd)
            if np.any(np.sum(dd * dd, axis=2) <= m.r * m.r):
                continue
        acc.append(c)
    assert list(idx) == acc
    cube_idx, _, _ = po.Model(d, T, seed=seed).pack(M)
    assert list(cube_idx) != acc


def test_fixed_window_stop_rule_by_hand():
    # tiny batches: the rule fires at the first boundary whose window holds too few accepts
    m = po.Model(2, 16, seed=1)
    idx, _, att = m.pack(10**7, stop_kind=1, stop_batch=1024, window=4096, stop_rate=2e-3)
    all_idx, _, _ = m.pack(att + 8192)
    assert list(all_idx[all_idx < att]) == list(idx)
    for c in range(4096, att + 1, 1024):
        cnt = np.sum((all_idx >= c - 4096) & (all_idx < c))
        if c < att:
            assert cnt >= 2e-3 * 4096
        else:
            assert cnt < 2e-3 * 4096


def test_max_n_stops_at_the_nth_accept():
    m = po.Model(2, 16, seed=2)
    idx, _, att = m.pack(20000, max_n=10)
    assert len(idx) == 10 and att == idx[-1] + 1


def _touch_groups(X, others, groups, r):
    """Groups of the paths in others (n, T, d) that X (T, d) touches at some timestep."""
    if len(others) == 0:
        return set()
    d = np.abs(X[None] - others)
    d = np.minimum(d, 2.0 - d)
    hit = np.any(np.all(d <= r, axis=2), axis=1)
    return {int(g) for g in np.asarray(groups)[hit]}


def test_subpath_first_round_by_brute_force():
    # Round 1 of the subpath phase (2^16 candidates) restated with dense numpy distances:
    # survivors touch exactly one group among the uniques; then, in index order, a survivor
    # joins iff it touches exactly one group among uniques + subpaths admitted so far.
    m = po.Model(2, 16, seed=1)
    idx, gid, att = m.pack(20000)
    pos = np.stack([m.positions(int(i)) for i in idx])
    sidx, sgid, satt, _ = m.pack_sub(pos, gid, first=att, budget=1)
    assert satt == 1 << 16  # one full round however small the budget
    r = m.r
    surv = []
    for c in range(att, att + (1 << 16)):
        X = m.positions(c)
        g = _touch_groups(X, pos, gid, r)
        if len(g) == 1:
            surv.append((c, X))
    want_idx, want_gid, apos, agid = [], [], list(pos), list(gid)
    for c, X in surv:
        g = _touch_groups(X, np.array(apos), agid, r)
        if len(g) == 1:
            (gg,) = g
            want_idx.append(c)
            want_gid.append(gg)
            apos.append(X)
            agid.append(gg)
    assert len(want_idx) > 10
    assert list(sidx) == want_idx and list(sgid) == want_gid


def test_subpath_round_start_rule_differs_from_serial():
    # A candidate whose only contact is a subpath admitted earlier in the same round is
    # dropped (the round-start test sees only the uniques), so the subpath list is not a
    # serial RSA over single-group contacts; check that at least one such candidate exists
    # in the first round and that it is absent from the output.
    m = po.Model(2, 16, seed=1)
    idx, gid, att = m.pack(20000)
    pos = np.stack([m.positions(int(i)) for i in idx])
    sidx, sgid, _, _ = m.pack_sub(pos, gid, first=att, budget=1)
    spos = np.stack([m.positions(int(i)) for i in sidx])
    first_round = set(range(att, att + (1 << 16)))
    found = 0
    for c in range(att, att + (1 << 16)):
        if c in set(sidx.tolist()):
            continue
        X = m.positions(c)
        if _touch_groups(X, pos, gid, m.r):
            continue
        earlier = sidx < c
        if len(_touch_groups(X, spos[earlier], sgid[earlier], m.r)) == 1:
            found += 1
            assert c in first_round
            if found >= 3:
                break
    assert found >= 1


def test_subpath_round_schedule_and_budget():
    # rounds run while fewer than budget candidates were drawn; round sizes follow the
    # survivor-driven halving/doubling from 2^16, so attempts land on a sum of round sizes
    m = po.Model(2, 16, seed=2)
    idx, gid, att = m.pack(5000)
    pos = np.stack([m.positions(int(i)) for i in idx])
    _, _, a1, _ = m.pack_sub(pos, gid, first=att, budget=1)
    _, _, a2, _ = m.pack_sub(pos, gid, first=att, budget=(1 << 16) + 1)
    assert a1 == 1 << 16
    assert a2 - a1 in (1 << 15, 1 << 16, 1 << 17)
    sidx, sgid, a3, filled = m.pack_sub(pos, gid, first=att, budget=a2)
    assert a3 == a2 and filled > 0 and np.all(np.diff(sidx) > 0)
    assert np.all(sidx >= att) and np.all(sidx < att + a3)



# ---------------------------------------------------------------- reference round schedule
def _halve_double(batch, surv):
    """Next round size from a round's survivors (cuda/braid_cuda.cu:920-924)."""
    got = min(surv, 1 << 16)
    if got > 512 and batch > 4096:
        return batch // 2
    if got < 64 and batch < (1 << 26):
        return batch * 2
    return batch


def _window_verdicts(entries, attempts_after, rate):
    """Per round: does the --until-accept-rate window fire (cuda/braid_cuda.cu:585-590,
    930-943)?  The window after entry k is the shortest suffix of entries[..k] whose
    attempts reach the target (all of them while the total is below it); it fires when it
    reaches the target, more than 10^6 candidates were consumed, and its acceptance rate
    is below `rate`."""
    target = int(min(5e9, max(2e7, 30.0 / rate)))
    out = []
    for k in range(len(entries)):
        att = adm = 0
        for a, n in reversed(entries[:k + 1]):
            att += a
            adm += n
            if att >= target:
                break
        out.append(att >= target and attempts_after[k] > 1_000_000 and adm / att < rate)
    return out


def _ref_cells(X, T):
    """(t, cell) keys of positions X (T, d) with the reference's host cell index
    floor((v + 1)/cell), cell = 2/T (cuda/braid_cuda.cu:699,868-870)."""
    c = np.floor((X + 1.0) / (2.0 / T)).astype(np.int64)
    key = np.arange(T, dtype=np.int64)
    for a in range(X.shape[1]):
        key = key * (T + 1) + c[:, a]
    return key


@pytest.mark.parametrize("d,T,rate", [(2, 16, 2e-6), (3, 8, 1.5e-6)])
def test_reference_round_schedule_by_hand(d, T, rate):
    # The unique phase consumed in the reference engine's rounds with --until-accept-rate:
    # every field of every round record is recomputed here from outside the schedule code.
    m = po.Model(d, T, seed=1)
    r = m.pack_ref(1e12, stop_rate=rate)
    rec, acc = r["rounds"], r["idx"]
    start, size, surv, adm, fresh = rec.T
    # contiguous rounds from 0; attempts = their sum; nothing left undecided
    assert start[0] == 0 and np.array_equal(start[1:], np.cumsum(size)[:-1])
    assert r["attempts"] == size.sum() == r["decided"]
    # the accepted list is serial RSA over the consumed range, whatever the round sizes
    plain, _ = m.pack_mt(r["attempts"], 8)
    assert np.array_equal(plain, acc) and np.array_equal(r["gid"], np.arange(len(acc)))
    # admitted per round from the accepted list
    assert np.array_equal(adm, [np.sum((acc >= s) & (acc < s + z)) for s, z in zip(start, size)])
    # survivors: candidates touching none of the paths accepted before the round, by pairwise
    # tests without a grid (rounds up to 2^21 candidates; later ones have doubled from them)
    for k in range(len(rec)):
        if size[k] <= 1 << 21:
            assert surv[k] == m.count_clear(int(start[k]), int(size[k]), acc[acc < start[k]]), k
        assert surv[k] >= adm[k]
    # sizes: braid_cuda adapts the batch after each round; braid_cuda3d sizes round k+1
    # before round k is drained, so its sizes lag one round (cuda/braid_cuda3d.cu:1446-1452)
    b, want = 1 << 16, []
    if d == 2:
        for k in range(len(rec)):
            want.append(b)
            b = _halve_double(b, int(surv[k]))
    else:
        want.append(b)
        for k in range(len(rec) - 1):
            want.append(b)
            b = _halve_double(b, int(surv[k]))
    assert list(size) == want
    # stop: braid_cuda's window records the adapted batch and ends after the firing round;
    # braid_cuda3d's records the drained round and drains the one already sized after it
    if d == 2:
        entries = [(_halve_double(int(want[k]), int(surv[k])), int(adm[k])) for k in range(len(rec))]
    else:
        entries = [(int(size[k]), int(adm[k])) for k in range(len(rec))]
    fires = _window_verdicts(entries, np.cumsum(size), rate)
    last = len(rec) - 1 if d == 2 else len(rec) - 2
    assert fires[last] and not any(fires[:last])
    # fresh cells per round: (t, cell) buckets first occupied by the round's accepts, in
    # admission order (insertPath, cuda/braid_cuda.cu:865-879)
    seen, per_round = set(), []
    X = m.positions_of(acc)
    for k in range(len(rec)):
        new = 0
        for i in np.nonzero((acc >= start[k]) & (acc < start[k] + size[k]))[0]:
            keys = set(_ref_cells(X[i], T).tolist())
            new += len(keys - seen)
            seen |= keys
        per_round.append(new)
    assert list(fresh) == per_round and fresh.sum() == len(seen)


def test_reference_schedule_budget_and_max_n():
    # without a window the last round overshoots the budget (the loop tests attempts <
    # budget before each round); max_n ends the admissions mid-round
    m = po.Model(2, 16, seed=2)
    r = m.pack_ref(90000)
    assert list(r["rounds"][:, 1]) == [1 << 16, 1 << 15] and r["attempts"] == 3 << 15
    assert np.array_equal(r["idx"], m.pack(r["attempts"])[0])
    r3 = m.pack_ref(90000, rule=3)
    assert list(r3["rounds"][:, 1]) == [1 << 16, 1 << 16]
    assert r3["attempts"] == 1 << 17 and np.array_equal(r3["idx"], m.pack(1 << 17)[0])
    r = m.pack_ref(10**7, max_n=40)
    assert len(r["idx"]) == 40 and r["decided"] == r["idx"][-1] + 1
    assert r["attempts"] == r["rounds"][:, 1].sum() > r["decided"]


@pytest.mark.parametrize("name", ["sub_d2_T16_s1_accept2e-3", "sub_d2_T16_s2_fill3e-5",
                                  "subref_d2_T16_s1_accept2e-3"])
def test_subpath_windows_and_fresh_cells_by_hand(name):
    # The subpath fixtures (the oracle's windowed stops, minutes of oracle time to rerun) carry
    # their full lists and round records; every record is recomputed here: admissions from the
    # list, fresh cells from the positions, the sizes from the survivor counts and both
    # windows (cuda/braid_cuda.cu:1000-1037) from the records.
    import hashlib
    import json
    import os

    gdir = os.path.join(os.path.dirname(__file__), "golden")
    g = json.load(open(os.path.join(gdir, name + ".json")))
    z = np.load(os.path.join(gdir, name + ".npz"))
    uidx, ugid, sidx, sgid, rec = z["uidx"], z["ugid"], z["sidx"], z["sgid"], z["rounds"]
    m = po.Model(g["dim"], g["T"], seed=g["seed"])
    # the unique phase is the oracle's (budget or engine schedule over unique_candidates)
    if g["unique_rule"] == "ref":
        r = m.pack_ref(g["unique_candidates"])
        want_u, att = r["idx"], r["attempts"]
    else:
        want_u, _, att = m.pack(g["unique_candidates"])
    assert np.array_equal(uidx, want_u) and att == g["unique_attempts"]
    h = hashlib.sha256()
    h.update(np.concatenate([uidx, sidx]).astype("<i8").tobytes())
    h.update(np.concatenate([ugid, sgid]).astype("<i4").tobytes())
    assert h.hexdigest() == g["digest"] and len(sidx) == g["nsub"]
    start, size, surv, adm, fresh = rec.T
    assert start[0] == 0 and np.array_equal(start[1:], np.cumsum(size)[:-1])
    assert size.sum() == g["sub_attempts"]
    first = g["unique_attempts"]
    assert np.array_equal(adm, [np.sum((sidx >= first + s) & (sidx < first + s + k))
                                for s, k in zip(start, size)])
    b, want = 1 << 16, []
    for k in range(len(rec)):
        want.append(b)
        b = _halve_double(b, int(surv[k]))
    assert list(size) == want
    adapted = [_halve_double(int(size[k]), int(surv[k])) for k in range(len(rec))]
    att_after = np.cumsum(size)
    acc_f = _window_verdicts(list(zip(adapted, adm.tolist())), att_after, g["accept_rate"]) \
        if g["accept_rate"] > 0 else [False] * len(rec)
    fill_f = _window_verdicts(list(zip(adapted, fresh.tolist())), att_after, g["fill_rate"]) \
        if g["fill_rate"] > 0 else [False] * len(rec)
    fired = [a or f for a, f in zip(acc_f, fill_f)]
    assert fired[-1] and not any(fired[:-1])
    # fresh cells: buckets first occupied, uniques first, then the subpaths in admission order
    X = m.positions_of(np.concatenate([uidx, sidx]))
    seen = set()
    for i in range(len(uidx)):
        seen |= set(_ref_cells(X[i], g["T"]).tolist())
    per_round = np.zeros(len(rec), np.int64)
    rnd = np.searchsorted(first + start, sidx, side="right") - 1
    for j in range(len(sidx)):
        keys = set(_ref_cells(X[len(uidx) + j], g["T"]).tolist())
        per_round[rnd[j]] += len(keys - seen)
        seen |= keys
    assert np.array_equal(per_round, fresh) and fresh.sum() == g["filled_new"]


# ---------------------------------------------------------------- knot regime
def _touch_any(X, others, r, T):
    """Does X (T, d) touch any of others (n, T, d) at some timestep?  The canonical pair
    predicate: neighbouring cells on every axis and |torus_delta| <= r on every axis."""
    if len(others) == 0:
        return False
    cx = np.floor((X + 1.0) * (T / 2.0)).astype(np.int64) % T
    co = np.floor((others + 1.0) * (T / 2.0)).astype(np.int64) % T
    dc = np.abs(cx[None] - co)
    adj = np.all((dc <= 1) | (dc == T - 1), axis=2)
    d = X[None] - others
    d = np.where(d > 1.0, d - 2.0, np.where(d < -1.0, d + 2.0, d))
    return bool(np.any(adj & np.all(np.abs(d) <= r, axis=2)))


@pytest.mark.parametrize("d,T,seed,budget,nu,div", [(2, 16, 1, 2e6, 1e5, 2.0), (3, 8, 2, 4e5, 1e5, 4.0)])
def test_knot_oracle_by_hand(d, T, seed, budget, nu, div):
    # cuda/braid_knot.cu's two phases: the primary species is serial RSA over [0, budget)
    # (the packer's own list); the nu species, checked here by dense numpy tests, admits a
    # candidate iff it touches no earlier strand of either species at the exclusion r/div;
    # the batches follow the ramp of :464-468 on survivor counts recounted pairwise
    m = po.Model(d, T, seed=seed, pin=True)
    r = m.pack_knot(budget, nu_attempts=nu, nu_div=div)
    idx, sp, rec = r["idx"], r["species"], r["batches"]
    assert r["primary_attempts"] == budget and r["attempts"] == budget + nu
    prim, _, _ = m.pack(int(budget))
    assert np.array_equal(idx[sp == 0], prim) and r["n_primary"] == len(prim)
    assert np.all(sp[:len(prim)] == 0) and np.all(sp[len(prim):] == 1) and len(idx) > len(prim)
    rnu = m.r / div
    cand = np.arange(int(budget), int(budget + nu))
    Xc = m.positions_of(cand)
    Xs = m.positions_of(idx)
    member = set(idx[sp == 1].tolist())
    for c, X in zip(cand, Xc):
        earlier = Xs[(sp == 0) | (idx < c)]
        assert _touch_any(X, earlier, rnu, T) == (int(c) not in member), c
    # batches: contiguous per phase, cut at the phase budget, sizes from the ramp
    start, size, surv, adm, spc = rec.T
    assert start[0] == 0 and np.array_equal(start[1:], np.cumsum(size)[:-1])
    for ph, end in ((0, budget), (1, budget + nu)):
        b = 4096
        for k in np.nonzero(spc == ph)[0]:
            assert size[k] == min(b, int(end - start[k]))
            if surv[k] > size[k] // 8 and b > 4096:
                b = max(4096, b // 4)
            elif surv[k] < size[k] // 64 and b < 1 << 24:
                b *= 4
        assert adm[spc == ph].sum() == np.sum(sp == ph)
    # survivors of the nu batches: pairwise count against the strands of each batch's start
    mn = po.Model(d, T, seed=seed, pin=True)
    mn._m.r = rnu
    for k in np.nonzero(spc == 1)[0]:
        before = idx[(sp == 0) | (idx < start[k])]
        assert surv[k] == mn.count_clear(int(start[k]), int(size[k]), before)


def test_knot_strand_capacity_ends_at_the_batch():
    # the capacity stops admissions inside a batch; the batch is still counted
    m = po.Model(2, 16, seed=3)
    r = m.pack_knot(1e6, max_strands=5)
    full = m.pack_knot(1e6)
    assert len(r["idx"]) == 5 and np.array_equal(r["idx"], full["idx"][:5])
    last = r["batches"][-1]
    assert r["attempts"] == last[0] + last[1] and r["idx"][-1] < r["attempts"] < 1e6
