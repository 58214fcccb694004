// ug_arith.h -- canonical arithmetic of the packer, compiled identically for the
// host (basis tables) and for gfx950 (kernels).  Every decision the FP64 path makes
// is a function of the bits produced here, so host and device must agree exactly:
// the translation units that include this header are built with -ffp-contract=off
// and every fused multiply-add below is written as an explicit fma.
//
// Reference semantics (universe-analysis/universe-generator):
//   torus_wrap / torus_delta             cuda/braid_cuda.cu:158-168
//   z_i = (i+1) pi/(T+1)                 cuda/braid_cuda.cu:655-660
//   sin1 amplitude U(-1,1)               cuda/braid_knot.cu:164
//   Dirichlet(1) by exponential spacings cuda/braid_cuda.cu:250-264
//   a_j = w_j / b_j, random signs        cuda/braid_cuda.cu:348-359
//   phases on even frequencies only      cuda/braid_cuda.cu:366-371
// The default RNG is new: Philox4x32-10 keyed by the seed, one counter per (candidate,
// stage, axis, slot pair), so every draw is addressable without a serial stream.  The
// reference's own serial xoshiro256** stream and draw order (its knot tool's proposal) are
// available as a cross-check mode in the FP64 forms (xo_* below).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define UG_HD __host__ __device__ __forceinline__

namespace ug {

// ----------------------------------------------------------------------------------
// Philox4x32-10.  Counter words: (idx lo, idx hi, pair or draw block, stage<<8 | axis).
// Stage 0: sin1 amplitudes (word a = axis a).  Stage 1: slot words; bit 0 of a Dirichlet word
// is that slot's sign.  Partial spectra: block p holds slot pair p's words {Dirichlet(2p),
// Dirichlet(2p+1), phase(2p), phase(2p+1)}.  Full-pool spectra (slot j has frequency j + 2, so
// only even slots have a phase): the axis' blocks form one word stream, word i = component
// i % 4 of block i / 4, and pair p takes words {3p, 3p+1, 3p+2} = {Dirichlet(2p),
// Dirichlet(2p+1), phase(2p)} (pair_words_full).  Stage 2: frequency draws for partial
// spectra, 4 per block.
struct u32x4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c: one three-input bitwise instruction on gfx950 (truth table 0x96), where the
// round function would otherwise spend two XORs per output word.
UG_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

UG_HD u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                   uint32_t k1) {
#pragma unroll
    for (int rnd = 0; rnd < 10; rnd++) {
        const uint64_t m0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3((uint32_t)(m1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(m0 >> 32), c3, k1);
        c1 = (uint32_t)m1;
        c3 = (uint32_t)m0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return u32x4{c0, c1, c2, c3};
}

// Stage-1 words {Dirichlet(2p), Dirichlet(2p+1), phase(2p), 0} of slot pair p of axis a in a
// full-pool spectrum: words 3p .. 3p+2 of the axis' stream (one block, or the end of one and
// the start of the next).
UG_HD u32x4 pair_words_full(uint32_t lo, uint32_t hi, int p, int a, uint32_t k0, uint32_t k1) {
    const uint32_t i = 3u * (uint32_t)p, b = i >> 2, o = i & 3u, c3 = (1u << 8) | (uint32_t)a;
    const u32x4 r = philox(lo, hi, b, c3, k0, k1);
    if (o == 0)
        return u32x4{r.x, r.y, r.z, 0u};
    if (o == 1)
        return u32x4{r.y, r.z, r.w, 0u};
    const u32x4 n = philox(lo, hi, b + 1, c3, k0, k1);
    if (o == 2)
        return u32x4{r.z, r.w, n.x, 0u};
    return u32x4{r.w, n.x, n.y, 0u};
}

// ----------------------------------------------------------------------------------
// sin(pi n/m), cos(pi n/m) for integers n >= 0, 0 < m <= 2^53: exact integer reduction
// to pi (r + q/2), |r| <= 1/4, then Taylor polynomials of sin(pi r) / cos(pi r) in r^2.
struct sc2 {
    double s, c;
};

UG_HD double horner_sinpi(double x) {
    double p = -0x1.8a404211f9547p-26;
    p = fma(p, x, 0x1.aaec32af93359p-21);
    p = fma(p, x, -0x1.6fadb9f155744p-16);
    p = fma(p, x, 0x1.e8f434d018d63p-12);
    p = fma(p, x, -0x1.e3074fde8871fp-8);
    p = fma(p, x, 0x1.50783487ee782p-4);
    p = fma(p, x, -0x1.32d2cce62bd86p-1);
    p = fma(p, x, 0x1.466bc6775aae2p+1);
    p = fma(p, x, -0x1.4abbce625be53p+2);
    p = fma(p, x, 0x1.921fb54442d18p+1);
    return p;
}
UG_HD double horner_cospi(double x) {
    double p = -0x1.2a0c591af8314p-23;
    p = fma(p, x, 0x1.20c62c2f2d7f5p-18);
    p = fma(p, x, -0x1.b6e24f44b128fp-14);
    p = fma(p, x, 0x1.f9d38a3763cc3p-10);
    p = fma(p, x, -0x1.a6d1f2a204a8cp-6);
    p = fma(p, x, 0x1.e1f506891babbp-3);
    p = fma(p, x, -0x1.55d3c7e3cbffap+0);
    p = fma(p, x, 0x1.03c1f081b5ac4p+2);
    p = fma(p, x, -0x1.3bd3cc9be45dep+2);
    p = fma(p, x, 0x1.0000000000000p+0);
    return p;
}

UG_HD sc2 sincospi_rat(uint64_t n, uint64_t m) {
    const uint64_t m2 = m + m;
    const uint64_t nn = n % m2;
    const uint64_t q = (4 * nn + m) / m2;
    const int64_t num = (int64_t)(nn + nn) - (int64_t)(q * m);
    const double r = (double)num / (double)m2;
    const double x = r * r;
    const double sn = r * horner_sinpi(x);
    const double cs = horner_cospi(x);
    switch ((unsigned)q & 3u) {
        case 0: return sc2{sn, cs};
        case 1: return sc2{cs, -sn};
        case 2: return sc2{-sn, -cs};
        default: return sc2{-cs, sn};
    }
}

// ----------------------------------------------------------------------------------
// log(v), v normal and positive: v = 2^e m with m in [sqrt(1/2), sqrt(2)],
// log m = 2 atanh(s) with s = (m-1)/(m+1) summed as 2s + s z P(z), z = s^2,
// P(z) = sum_{i=1..10} 2/(2i+1) z^(i-1); then + e ln2 with ln2 split in hi/lo.
UG_HD double log_pos(double v) {
    uint64_t bits = (uint64_t)__builtin_bit_cast(uint64_t, v);
    int e = (int)((bits >> 52) & 0x7ffu) - 1023;
    uint64_t mb = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    if (mb > 0x3ff6a09e667f3bcdull) {
        mb -= 0x0010000000000000ull;
        e += 1;
    }
    const double m = __builtin_bit_cast(double, mb);
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 0x1.8618618618618p-4;
    p = fma(p, z, 0x1.af286bca1af28p-4);
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-4);
    p = fma(p, z, 0x1.1111111111111p-3);
    p = fma(p, z, 0x1.3b13b13b13b14p-3);
    p = fma(p, z, 0x1.745d1745d1746p-3);
    p = fma(p, z, 0x1.c71c71c71c71cp-3);
    p = fma(p, z, 0x1.2492492492492p-2);
    p = fma(p, z, 0x1.999999999999ap-2);
    p = fma(p, z, 0x1.5555555555555p-1);
    const double lm = fma(s, z * p, s + s);
    const double de = (double)e;
    return fma(de, 0x1.62e42fee00000p-1, fma(de, 0x1.a39ef35793c76p-33, lm));
}

// ----------------------------------------------------------------------------------
// Uniform variates of the candidate stream.
UG_HD double a1_of_word(uint32_t w) {  // (2w+1) 2^-32 - 1, in (-1, 1), exact
    return (double)(2 * (uint64_t)w + 1) * 0x1p-32 - 1.0;
}
UG_HD double expo_of_word(uint32_t w) {  // -log(1 - u), u = (w >> 1) 2^-31
    const double v = (double)(0x80000000u - (w >> 1)) * 0x1p-31;
    return -log_pos(v);
}

// ----------------------------------------------------------------------------------
// The reference's candidate-indexed proposal stream (rng_kind UG_RNG_XOSHIRO_REF), the one its
// knot tool draws (cuda/braid_knot.cu:56-90 Rng, :207 seed): xoshiro256** (Blackman and
// Vigna) whose four state words are successive splitmix64 outputs from the candidate's seed
// seed ^ (idx * 0x9E3779B97F4A7C15 + 0x1234567).
struct xo256 {
    uint64_t s0, s1, s2, s3;
};

UG_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

UG_HD uint64_t splitmix64_next(uint64_t& x) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

UG_HD xo256 xo_seeded(uint64_t seed) {
    xo256 r;
    r.s0 = splitmix64_next(seed);
    r.s1 = splitmix64_next(seed);
    r.s2 = splitmix64_next(seed);
    r.s3 = splitmix64_next(seed);
    return r;
}

UG_HD uint64_t xo_next(xo256& r) {
    const uint64_t out = rotl64(r.s1 * 5u, 7) * 9u;
    const uint64_t sh = r.s1 << 17;
    r.s2 ^= r.s0;
    r.s3 ^= r.s1;
    r.s1 ^= r.s2;
    r.s0 ^= r.s3;
    r.s2 ^= sh;
    r.s3 = rotl64(r.s3, 45);
    return out;
}

UG_HD xo256 xo_candidate(uint64_t seed, uint64_t idx) {
    return xo_seeded(seed ^ (idx * 0x9E3779B97F4A7C15ull + 0x1234567ull));
}

// One axis of the knot tool's proposal (cuda/braid_knot.cu:124-166) as raw draws, in stream
// order: the nw frequencies (Fisher-Yates over 2..modmax+1 for the full pool, else up to 64
// rejection draws each, a duplicate kept after the 64th), then nw 53-bit uniforms of the
// exponential spacings (ek[j * st]), then per slot a sign draw and, for an even frequency, a
// 53-bit phase uniform (pk[j * st] = sign << 63 | phase uniform, the uniform 0 for odd
// frequencies), then the sin1 uniform unless pinned (returned; 0 when pinned).
UG_HD uint64_t xo_axis(xo256& r, int modmax, int nw, int pin, int* b, uint64_t* ek, uint64_t* pk, int st) {
    if (nw == modmax) {
        for (int j = 0; j < nw; j++)
            b[j] = j + 2;
        for (int j = nw - 1; j > 0; j--) {
            const int k = (int)(xo_next(r) % (uint64_t)(j + 1));
            const int t = b[j];
            b[j] = b[k];
            b[k] = t;
        }
    } else {
        for (int j = 0; j < nw; j++) {
            int bn = 0;
            bool dup = true;
            for (int guard = 0; dup && guard < 64; guard++) {
                bn = (int)(xo_next(r) % (uint64_t)modmax) + 2;
                dup = false;
                for (int k = 0; k < j; k++)
                    dup = dup || b[k] == bn;
            }
            b[j] = bn;
        }
    }
    for (int j = 0; j < nw; j++)
        ek[j * st] = xo_next(r) >> 11;
    for (int j = 0; j < nw; j++) {
        uint64_t v = (xo_next(r) & 1u) << 63;
        if ((b[j] & 1) == 0)
            v |= xo_next(r) >> 11;
        pk[j * st] = v;
    }
    return pin ? 0u : xo_next(r) >> 11;
}

// The arithmetic on those draws (cuda/braid_knot.cu:112-120,155-164), canonical here as for the
// Philox stream: spacing e = -log(1 - u), u = k 2^-53 (1 - u exact); the weights e_j / sum with
// the sum taken serially in slot order; amplitude (e_j / sum) / b_j, both divisions IEEE;
// phase angle pi k 2^-53 through sincospi_rat (the reference rounds u pi first); sin1
// amplitude 2u - 1 (exact).
UG_HD double xo_spacing(uint64_t k) { return -log_pos(1.0 - (double)k * 0x1p-53); }
UG_HD double xo_a1(uint64_t k) { return (double)k * 0x1p-52 - 1.0; }

// torus_wrap / torus_delta (cuda/braid_cuda.cu:158-168).
UG_HD double torus_wrap(double x) { return x - 2.0 * floor((x + 1.0) * 0.5); }
UG_HD double torus_delta(double d) {
    if (d > 1.0)
        return d - 2.0;
    if (d < -1.0)
        return d + 2.0;
    return d;
}
// Canonical cell index of a wrapped coordinate, wrapped modulo T.
UG_HD int cell_of(double x, double halfT, int T) {
    int c = (int)floor((x + 1.0) * halfT);
    if (c < 0)
        c += T;
    if (c >= T)
        c -= T;
    return c;
}

}  // namespace ug
