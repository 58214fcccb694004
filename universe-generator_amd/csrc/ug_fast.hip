// ug_fast.hip -- split-BF16 prefilter of the packer (UG_PREC_SPLIT_BF16).
//
// Evaluation is the dense contraction X[c][t] = a1[c] + (1/W_c) sum_r A[c][r] B[r][t]:
// A holds, per frequency pair (k even, k+1 odd), the unnormalized coefficients
//   (e~_k cos f_k, e~_k sin f_k, e~_{k+1}, 0),   e~ = +-e  (e the exponential spacing),
// and B the canonical basis rows divided by their frequency, (S_k/k, C_k/k, S_{k+1}/(k+1), 0),
// at the timesteps.  The reference
// evaluates the same sum term by term with FP64 sin() (cuda/braid_cuda.cu:418-441) and
// its --sparse mode from FP32 tables (cuda/braid_cuda3d.cu:630-717).
//
// One wave owns a tile of 32 candidates: lanes generate their A fragments directly in
// registers (Philox, FP32 log, hardware sin/cos of the phases), split each value into bf16 hi + lo, and run
// the 3-term product hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation against the basis of the first 32 endpoint-first timesteps, staged in LDS.
// The lookup then rejects a candidate only on a CERTAIN contact (every axis within
// r - m_t, m_t a bound on the FP32/BF16 position error, like the one-sided margin of
// cuda/braid_cuda3d.cu:938-948); everything else -- gray contacts, candidates without a
// certain contact among those timesteps -- is flagged for the exact FP64 test.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ug_arith.h"
#include "ug_internal.h"

using namespace ug;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace {

constexpr int TILE = 32;        // candidates per wave tile (MFMA M)
// A lookup round's split of the wave among U unresolved candidates: per = 64 / U lanes each,
// and M = ceil(2^16 / per), with lane / per = (lane * M) >> 16 exactly for every lane < 64;
// entry U = per | M << 8.
__constant__ uint32_t k_round_split[33] = {0, 262208, 524320, 798997, 1048592, 1398284, 1677834, 1864201, 2097160, 2396935, 2796294, 3355653, 3355653, 4194308, 4194308, 4194308, 4194308, 5592579, 5592579, 5592579, 5592579, 5592579, 8388610, 8388610, 8388610, 8388610, 8388610, 8388610, 8388610, 8388610, 8388610, 8388610, 8388610};
constexpr int NCOL = 32;        // timesteps per N block (MFMA N)
constexpr int FAST_WAVES = 4;   // waves per workgroup

// cos(pi u), sin(pi u) for u = w 2^-32 in [0, 1), on the transcendental unit: v_cos_f32 and
// v_sin_f32 take revolutions, and pi u = 2 pi (w 2^-33).  Over all 2^32 phase words the
// absolute error against FP64 is at most 1.9e-7 on gfx950 (measured exhaustively on one
// MI355X; ug_split_generation_error repeats the measurement), inside the 2^-20 generation term
// of the certainty margin.
__device__ __forceinline__ void phase_cs(uint32_t w, float& cf, float& sf) {
    const float x = (float)w * 0x1p-33f;
    cf = __builtin_amdgcn_cosf(x);
    sf = __builtin_amdgcn_sinf(x);
}

// An exponential spacing in units of ln 2: -log2(1 - u), u = (w >> 1) 2^-31.  All spacings of
// an axis share the scale, which cancels in the amplitudes e_j / W.
__device__ __forceinline__ float expo2_f32(uint32_t w) {
    const float v = (float)(0x80000000u - (w >> 1)) * 0x1p-31f;
    return -__log2f(v);
}

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(4))) uint32_t u32v4;

// Two values split into bf16 hi (round to nearest even) and the bf16 of the exact FP32
// remainder, packed as (hi of x, hi of y) and (lo of x, lo of y): one packed conversion per
// word, the hi halves widened back by shifts.
__device__ __forceinline__ void split_bf16x2(float x, float y, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x, y}, bf16x2));
    const float xl = x - __uint_as_float(hi << 16), yl = y - __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){xl, yl}, bf16x2));
}
// The same for (x, 0).
__device__ __forceinline__ void split_bf16x1(float x, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x, 0.f}, bf16x2));
    const float xl = x - __uint_as_float(hi << 16);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){xl, 0.f}, bf16x2));
}

// A fragments of one K step from the words of its two slot pairs p0, p0 + 1 (u0, u1:
// {Dirichlet(2p), Dirichlet(2p+1), phase(2p)} of a full-pool spectrum).  The 1/k factors live in
// the basis rows (B = S_k/k, C_k/k, S_{k+1}/(k+1)), so a slot's value is its signed exponential
// spacing, times cos f / sin f for even frequencies.  Missing slots (past the last frequency)
// and padding candidates contribute zeros, selected without branches; their basis rows are
// zero as well.
__device__ __forceinline__ void gen_kstep(const DevModel& m, bool valid, int p0, const u32x4& u0, const u32x4& u1,
                                          bf16x8& Ahi, bf16x8& Alo, float& Wpart) {
    uint32_t hw[4], lw[4];
#pragma unroll
    for (int pp = 0; pp < 2; pp++) {
        const int p = p0 + pp;
        const u32x4 r = pp ? u1 : u0;
        const int j0 = 2 * p;
        // slot 2p exists iff p < npairs, slot 2p + 1 iff 2p + 1 < nw
        const float e0 = valid && p < m.npairs ? expo2_f32(r.x) : 0.f;
        const float e1 = valid && j0 + 1 < m.nw ? expo2_f32(r.y) : 0.f;
        // bit 0 of the Dirichlet word is the slot's sign (e0, e1 >= 0: OR sets it)
        const float ae = __uint_as_float(__float_as_uint(e0) | (r.x << 31));
        float cf, sf;
        phase_cs(r.z, cf, sf);
        const float v0 = ae * cf, v1 = ae * sf;
        const float v2 = __uint_as_float(__float_as_uint(e1) | (r.y << 31));
        Wpart += e0;
        Wpart += e1;
        // slots (v0, v1) and (v2, 0) of the fragment, hi and lo words
        split_bf16x2(v0, v1, hw[2 * pp], lw[2 * pp]);
        split_bf16x1(v2, hw[2 * pp + 1], lw[2 * pp + 1]);
    }
    Ahi = __builtin_bit_cast(bf16x8, (u32v4){hw[0], hw[1], hw[2], hw[3]});
    Alo = __builtin_bit_cast(bf16x8, (u32v4){lw[0], lw[1], lw[2], lw[3]});
}

struct FastArgs {
    DevModel m;
    int ks;                 // K steps (4 pairs each)
    int ncols;              // ordered timesteps with a certainty margin (T, in N blocks of 32)
    const bf16x8* bfrag;    // [nb][ks][2][64] basis fragments, columns in endpoint-first order
    const int* colt;        // [nb*32] timestep of each column (-1 = padding)
    // knot regime: the resident strands' FP32 positions [T][ks] (x, y, w, unused), nk of them
    const float4* kpos32;
    int64_t kstride;
    int nk;
};

// Positions of the wave's 32 candidates at the 32 columns of N block nb into
// pos[D][32][32] (LDS), FP32, wrapped.  Returns per-candidate W-check flags in lane bits.
template <int D>
__device__ __forceinline__ bool tile_positions(const FastArgs& f, uint64_t idx, bool valid,
                                               int lane, int nb, const bf16x8* bsrc,
                                               float* pos, float* sa1, float* sinvw) {
    const DevModel& m = f.m;
    const int mrow = lane & 31, h = lane >> 5;
    const u32x4 w0 = philox((uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u, m.key0, m.key1);
    const uint32_t wa[3] = {w0.x, w0.y, w0.z};
    bool wbad = false;
    // K steps outer, axes inner: each basis fragment is loaded once for all axes; every
    // axis accumulates its own K steps in the same order as before
    f32x16 acc[D];
    float Wpart[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        Wpart[a] = 0.f;
#pragma unroll
        for (int v = 0; v < 16; v++)
            acc[a][v] = 0.f;
    }
    // K steps Q, Q + 1 (Q even) of lane half h take the four slot pairs 4g .. 4g + 3, g = Q + h,
    // whose words are blocks 3g .. 3g + 2 of the axis' stream: two blocks at step Q (its two
    // pairs and the first two words of the next pair, kept), one at step Q + 1
    const uint32_t ilo = (uint32_t)idx, ihi = (uint32_t)(idx >> 32);
    for (int Q = 0; Q < f.ks; Q += 2) {
        const int g = Q + h;
        uint32_t keep[D][2];
        {
            const bf16x8 Bhi = bsrc[((size_t)(nb * f.ks + Q) * 2 + 0) * 64 + lane];
            const bf16x8 Blo = bsrc[((size_t)(nb * f.ks + Q) * 2 + 1) * 64 + lane];
#pragma unroll
            for (int a = 0; a < D; a++) {
                const uint32_t c3 = (1u << 8) | (uint32_t)a;
                const u32x4 b0 = philox(ilo, ihi, 3u * g, c3, m.key0, m.key1);
                const u32x4 b1 = philox(ilo, ihi, 3u * g + 1u, c3, m.key0, m.key1);
                keep[a][0] = b1.z;
                keep[a][1] = b1.w;
                bf16x8 Ahi, Alo;
                gen_kstep(m, valid, 4 * g, u32x4{b0.x, b0.y, b0.z, 0u}, u32x4{b0.w, b1.x, b1.y, 0u}, Ahi, Alo,
                          Wpart[a]);
                acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi, Bhi, acc[a], 0, 0, 0);
                acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi, Blo, acc[a], 0, 0, 0);
                acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Alo, Bhi, acc[a], 0, 0, 0);
            }
        }
        {
            const bf16x8 Bhi = bsrc[((size_t)(nb * f.ks + Q + 1) * 2 + 0) * 64 + lane];
            const bf16x8 Blo = bsrc[((size_t)(nb * f.ks + Q + 1) * 2 + 1) * 64 + lane];
#pragma unroll
            for (int a = 0; a < D; a++) {
                const uint32_t c3 = (1u << 8) | (uint32_t)a;
                const u32x4 b2 = philox(ilo, ihi, 3u * g + 2u, c3, m.key0, m.key1);
                bf16x8 Ahi, Alo;
                gen_kstep(m, valid, 4 * g + 2, u32x4{keep[a][0], keep[a][1], b2.x, 0u}, u32x4{b2.y, b2.z, b2.w, 0u},
                          Ahi, Alo, Wpart[a]);
                acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi, Bhi, acc[a], 0, 0, 0);
                acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi, Blo, acc[a], 0, 0, 0);
                acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Alo, Bhi, acc[a], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float W = Wpart[a] + __shfl_xor(Wpart[a], 32, 64);
        // W sums spacings in units of ln 2: W ln2 < nw/4
        wbad = wbad || (valid && W * 0.693147182f < 0.25f * (float)m.nw);
        if (h == 0) {
            sa1[mrow] = m.pin ? 0.f : (float)a1_of_word(wa[a]);
            sinvw[mrow] = 1.0f / W;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // D layout of 32x32 MFMA: lane holds column n = lane & 31, rows 8(v/4) + 4h + (v%4)
        // per-lane offsets kept opaque to the compiler here, so that the 16 stores use one
        // base register with immediate offsets instead of 16 hoisted addresses.  The rows'
        // 1/W and a1 are read first, four consecutive rows per LDS read.
        int poff = (a * TILE + 4 * h) * NCOL + mrow, roff = 4 * h;
        asm volatile("" : "+v"(poff), "+v"(roff));
        float4 rw[4], r1[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            rw[j] = *(const float4*)&sinvw[roff + 8 * j];
            r1[j] = *(const float4*)&sa1[roff + 8 * j];
        }
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int ro = 8 * (v >> 2) + (v & 3);
            const float w = v % 4 == 0 ? rw[v >> 2].x : v % 4 == 1 ? rw[v >> 2].y : v % 4 == 2 ? rw[v >> 2].z : rw[v >> 2].w;
            const float c = v % 4 == 0 ? r1[v >> 2].x : v % 4 == 1 ? r1[v >> 2].y : v % 4 == 2 ? r1[v >> 2].z : r1[v >> 2].w;
            float X = fmaf(acc[a][v], w, c);
            // X - 2 floor((X + 1) / 2), each fma rounding exactly as the separate operations
            // (halving and doubling are exact)
            X = fmaf(-2.0f, floorf(fmaf(X, 0.5f, 0.5f)), X);
            pos[poff + ro * NCOL] = X;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    return wbad;
}

// One work item of the prefilter: the wave's tile `tile` of 32 candidates over the N blocks
// [nbs, nbe) of 32 ordered timesteps (count: the launch's candidates, already bounded).
// RSN (measurement): lane 0's cls[0..2] count the candidates left to the FP64 test by reason:
// spacing sum below the bound's premise, a possible contact (within r + m_t) but no certain one
// among the examined timesteps, no possible contact there at all.
template <int D, bool SUB, bool SPARSE, bool KNOT, bool RSN = false>
__device__ __forceinline__ void prefilter_item(const FastArgs& f, int64_t tile, int nbs, int nbe, int par,
                                               uint64_t base, int64_t count, const int64_t* list,
                                               uint8_t* flags, const int* acc_gid, int64_t nlimit,
                                               unsigned long long& nvis, unsigned long long& npts,
                                               unsigned long long* cls, uint32_t* skip) {
    extern __shared__ __align__(16) unsigned char lds[];
    const DevModel& m = f.m;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // basis fragments of N block 0 are read from global memory (a few tens of KB, cache
    // resident); LDS holds only each wave's position tile, which keeps occupancy up
    const bf16x8* bl = f.bfrag;
    float* pos = (float*)lds + (size_t)wv * (D * TILE * NCOL + 5 * TILE);
    float* sa1 = pos + D * TILE * NCOL;
    float* sinvw = sa1 + TILE;
    int* smin = (int*)(sinvw + 2 * TILE);  // after ulist: per-candidate group range of a round
    int* smax = smin + TILE;
    const int mrow = lane & 31, h = lane >> 5;
    const float r = (float)m.r;
    {
        const int64_t i = tile * TILE + mrow;
        const bool valid = i < count;
        const uint64_t idx = list ? (valid ? (uint64_t)list[i] : 0ull) : base + (uint64_t)(valid ? i : 0);
        uint32_t rejm = 0, graym = 0, wbadm = 0;
        uint32_t known = 0;  // lanes < 32: block-0 columns proven free of any FP64 contact
        int g1 = -1;  // SUB, lanes < 32: the one group candidate `lane` certainly touches so far
        // N blocks of 32 ordered timesteps each; the coefficients are generated again per
        // block (a later block only sees the candidates no earlier one rejected)
        for (int nb = nbs; nb < nbe; nb++) {
        const bool wbad = tile_positions<D>(f, idx, valid, lane, nb, bl, pos, sa1, sinvw);
        // Adaptive cooperative lookup: every round hands all 64 lanes to the still
        // unresolved candidates of the tile (64/U consecutive columns each), so the tile
        // finishes in a few dependent memory round trips.  A candidate whose exponential
        // spacings sum below nw/4 (wbad) is outside the error bound's premise and goes
        // straight to the FP64 test.
        uint32_t unres = (uint32_t)__ballot(lane < 32 && valid && !wbad && !((rejm >> lane) & 1u));
        if (RSN)
            wbadm |= (uint32_t)__ballot(lane < 32 && valid && wbad);
        if (!unres)
            break;
        const int col0 = NCOL * nb, bcols = min(NCOL, f.ncols - col0);  // this block's columns
        int ncol = 0;  // lanes < 32: next column (within the block) of candidate `lane`
        int* ulist = (int*)(sinvw + TILE);
        while (unres) {
            const int U = __popc(unres);
            const uint32_t split = k_round_split[U];
            const int per = (int)(split & 255u);
            if (lane < 32 && ((unres >> lane) & 1u))
                ulist[__popc(unres & ((1u << lane) - 1u))] = lane;
            if (SUB && lane < 32) {
                smin[lane] = INT32_MAX;
                smax[lane] = -1;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int rk = (int)(((uint32_t)lane * (split >> 8)) >> 16), off = lane - rk * per;
            const int cm = rk < U ? ulist[rk] : 0;
            const int col = __shfl(ncol, cm, 64) + off;
            // cert: a certain contact; maybe: a possible one (within r + m_t); clr: this
            // column provably has no FP64 contact (see the interior test below)
            bool cert = false, maybe = false, clr = false;
            int gmn = INT32_MAX, gmx = -1;  // SUB: groups of this lane's certain contacts
            if (rk < U && col < bcols) {
                const int t = f.colt[col0 + col];
                const float mt = m.marg[col0 + col];
                // L2 ball: the FP64 distance exceeds the FP32 one by at most sqrt(D)*mt, and
                // the FP32 sum of squares is within (D+1)*2^-24 of its exact value
                const float rb = r - sqrtf((float)D) * mt;
                const float rb2 = rb > 0.f ? rb * rb * (1.0f - 0x1p-19f) : -1.0f;
                const float rp = r + sqrtf((float)D) * mt;  // a possible ball contact
                const float rp2 = rp * rp * (1.0f + 0x1p-19f);
                float X[D];
                int c[D];
#pragma unroll
                for (int a = 0; a < D; a++) {
                    X[a] = pos[(a * TILE + cm) * NCOL + col];
                    int cc = (int)floorf((X[a] + 1.0f) * (float)m.halfT);
                    cc = cc < 0 ? cc + m.T : (cc >= m.T ? cc - m.T : cc);
                    c[a] = cc;
                }
                nvis++;
                constexpr int ROWS = D == 2 ? 3 : 9;
                // neighbour row rr: x offset rr/3 - 1 and y offset rr%3 - 1 (3+1), x offset rr - 1 (2+1)
                auto gx_of = [&](int rr) { return wrap_cell(c[0], (D == 2 ? rr : rr / 3) - 1, m.T); };
                auto gy_of = [&](int rr) { return D == 2 ? 0 : wrap_cell(c[1], rr % 3 - 1, m.T); };
                auto row_of = [&](int rr) -> int64_t {
                    return D == 2 ? (int64_t)t * m.T + gx_of(rr) : ((int64_t)t * m.T + gx_of(rr)) * m.T + gy_of(rr);
                };
                // head-table slot of neighbour bit b (row b / 3, padded cell c + b % 3): below
                // 2^31 in 2+1 (T <= 512 rows of T + 2 slots per timestep), 32-bit math there
                auto tab_at = [&](int b) -> int64_t {
                    if constexpr (D == 2)
                        return (int64_t)(((uint32_t)t * (uint32_t)m.T + (uint32_t)gx_of(b / 3)) * (uint32_t)m.rowlen +
                                         (uint32_t)(c[1] + b % 3));
                    return row_of(b / 3) * m.rowlen + c[D - 1] + b % 3;
                };
                // one grid point of this lane's visit: certain contact (or, SUB, its group)
                auto point = [&](int node, const float4& P) {
                    const float p3[3] = {P.x, P.y, P.z};
                    float dmax = 0.f, q = 0.f;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        float d = X[a] - p3[a];
                        d = d > 1.f ? d - 2.f : (d < -1.f ? d + 2.f : d);
                        dmax = fmaxf(dmax, fabsf(d));
                        q = fmaf(d, d, q);
                    }
                    npts++;
                    maybe = maybe || (m.euclid ? q <= rp2 : dmax <= r + mt);
                    const bool sure = m.euclid ? q <= rb2 : dmax <= r - mt;
                    if (!SUB) {
                        cert = cert || sure;
                    } else if (sure && node < nlimit) {
                        const int gid = acc_gid[node];
                        gmn = min(gmn, gid);
                        gmx = max(gmx, gid);
                        cert = gmn != gmx;  // two groups: nothing more to learn here
                    }
                };
                if constexpr (KNOT) {
                    // the strands of this timestep are contiguous: two loads in flight at a time
                    for (int sidx = 0; sidx < f.nk && !cert; sidx += 2) {
                        const float4 P0 = f.kpos32[(size_t)t * f.kstride + sidx];
                        const float4 P1 = f.kpos32[(size_t)t * f.kstride + min(sidx + 1, f.nk - 1)];
                        point(sidx, P0);
                        point(sidx + 1, P1);
                    }
                } else {
                // occupancy of all neighbour cells first (independent loads, issued together)
                // as one mask; then up to 2 occupied cells at a time load their heads together
                // (from the dense table, or by probing the sparse grid) and walk their chains
                // together, so a visit costs about three dependent memory steps (bits, heads,
                // points)
                uint32_t cells = 0;
                if constexpr (D == 2) {
                    // the three rows' word pairs loaded before any is used
                    const int wq = c[1] >> 5, sh = c[1] & 31;
                    uint64_t raw[3];
#pragma unroll
                    for (int rr = 0; rr < 3; rr++) {
                        const uint32_t* wp = m.occ + occ_index<2>(m, t, gx_of(rr), 0, wq);
                        raw[rr] = (uint64_t)wp[0] | ((uint64_t)wp[1] << 32);
                    }
#pragma unroll
                    for (int rr = 0; rr < 3; rr++)
                        cells |= ((uint32_t)(raw[rr] >> sh) & 7u) << (3 * rr);
                } else {
                    // the nine rows' words loaded before any is used; the next word plane only
                    // where the window crosses into it (the same for all nine rows)
                    const int wq = c[2] >> 5, sh = c[2] & 31;
                    uint32_t lo[ROWS];
#pragma unroll
                    for (int rr = 0; rr < ROWS; rr++)
                        lo[rr] = m.occ[occ_index<3>(m, t, gx_of(rr), gy_of(rr), wq)];
                    if (sh > 29) {
                        uint32_t hi[ROWS];
#pragma unroll
                        for (int rr = 0; rr < ROWS; rr++)
                            hi[rr] = m.occ[occ_index<3>(m, t, gx_of(rr), gy_of(rr), wq + 1)];
#pragma unroll
                        for (int rr = 0; rr < ROWS; rr++)
                            cells |= ((uint32_t)((((uint64_t)hi[rr] << 32) | lo[rr]) >> sh) & 7u) << (3 * rr);
                    } else {
#pragma unroll
                        for (int rr = 0; rr < ROWS; rr++)
                            cells |= ((lo[rr] >> sh) & 7u) << (3 * rr);
                    }
                }
                while (cells && !cert) {
                    constexpr int NG = 2;
                    int node[NG];
#pragma unroll
                    for (int g = 0; g < NG; g++) {
                        node[g] = -1;
                        if (cells) {
                            const int b = __builtin_ctz(cells);
                            cells &= cells - 1u;
                            node[g] = SPARSE ? hash_head(m, cell_hkey(row_of(b / 3), wrap_cell(c[D - 1], b % 3 - 1, m.T), m.T))
                                             : m.tab[tab_at(b)];
                        }
                    }
                    while (!cert && (node[0] >= 0 || node[1] >= 0)) {
                        float4 P[NG];
#pragma unroll
                        for (int g = 0; g < NG; g++)
                            if (node[g] >= 0)
                                P[g] = m.pos32[(uint64_t)(uint32_t)node[g] * (uint32_t)m.T + (uint32_t)t];
#pragma unroll
                        for (int g = 0; g < NG; g++) {
                            if (node[g] < 0)
                                continue;
                            point(node[g], P[g]);
                            node[g] = __float_as_int(P[g].w);
                        }
                    }
                }
                // the column provably holds no FP64 contact: no point of the scanned cells
                // within r + m_t, and the FP32 position more than m_t (plus its rounding) inside
                // its cell on every axis -- the FP64 position then lies in the same cell, the
                // FP64 test scans the same neighbour cells, and every point there is farther
                // than r in FP64
                if (!cert && !maybe) {
                    bool inner = true;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        const float u = (X[a] + 1.0f) * (float)m.halfT;
                        const float fr = u - floorf(u);
                        inner = inner && fminf(fr, 1.0f - fr) * r > mt + 0x1p-20f;
                    }
                    clr = inner;
                }
                }
            }
            if (!SUB) {
                const uint64_t hits = __ballot(cert);
                const uint64_t maybes = RSN ? __ballot(maybe) : 0ull;
                const uint64_t clrs = KNOT ? 0ull : __ballot(clr);
                if (lane < 32 && ((unres >> lane) & 1u)) {
                    const int myrk = __popc(unres & ((1u << lane) - 1u));
                    const uint64_t segmask = per >= 64 ? ~0ull : ((1ull << per) - 1ull);
                    if ((hits >> (myrk * per)) & segmask)
                        rejm |= 1u << lane;
                    if ((maybes >> (myrk * per)) & segmask)
                        graym |= 1u << lane;
                    known |= (uint32_t)((((clrs >> (myrk * per)) & segmask) << ncol) & 0xffffffffull);
                    ncol += per;
                }
            } else {
                // merge the segment's certain groups with the candidate's earlier one
                if (rk < U && gmx >= 0) {
                    atomicMin(&smin[cm], gmn);
                    atomicMax(&smax[cm], gmx);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (lane < 32 && ((unres >> lane) & 1u)) {
                    const int mn = smin[lane], mx = smax[lane];
                    if (mx >= 0) {
                        if (mn != mx || (g1 >= 0 && g1 != mn))
                            rejm |= 1u << lane;
                        else
                            g1 = mn;
                    }
                    ncol += per;
                }
            }
            const uint32_t rej_all = (uint32_t)__ballot(lane < 32 && ((rejm >> lane) & 1u));
            const uint32_t done = (uint32_t)__ballot(lane < 32 && ncol >= bcols);
            unres &= ~(rej_all | done);
            rejm = rej_all;
        }
        }
        if (valid && h == 0) {
            if (!par)
                flags[i] = ((rejm >> lane) & 1u) ? 0 : 1;
            else if ((rejm >> lane) & 1u)
                flags[i] = 0;
            if (skip && !par && nbs == 0 && !((rejm >> lane) & 1u))
                skip[i] = known;
        }
        if (RSN && !par && !SUB) {
            const uint32_t left = (uint32_t)__ballot(valid && h == 0 && !((rejm >> lane) & 1u));
            if (lane == 0) {
                cls[0] += __popc(left & wbadm);
                cls[1] += __popc(left & ~wbadm & graym);
                cls[2] += __popc(left & ~wbadm & ~graym);
            }
        }
    }
}

// Prefilter: flags[i] = 0 when candidate i certainly collides, 1 when the FP64 test must decide.
// SUB (subpath phase): flags[i] = 0 only when candidate i certainly touches two different
// groups among the paths stored before row nlimit (the round-start packing) -- such a
// candidate fails both the round-start test and the adoption test.
// Occupancy: 2+1 runs 4 waves per SIMD (at most 128 VGPRs, no spills; 4 blocks of 35 KB LDS
// fit a CU), 2-4% faster per launch than 3 on one MI355X; 3+1 stays at 3, which is also what
// its 52 KB of LDS per block allows.
// KNOT: no grid -- a visit compares the candidate with every resident strand (the knot
// regime's brute force) and rejects it on a certain contact.
// RSN: also count why candidates are left to the FP64 test (measurement; ctr must be set).
template <int D, bool SUB, bool SPARSE, bool KNOT = false, bool RSN = false>
__global__ __launch_bounds__(256, (D == 2 && !SUB) ? 4 : 3) void k_fast_prefilter(FastArgs f, uint64_t base, int64_t count,
                                                        const int64_t* list, const int* dcount,
                                                        int cb0, int cb1, int par, uint8_t* flags,
                                                        unsigned long long* ctr, const int* acc_gid,
                                                        int64_t nlimit, uint32_t* skip) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (dcount)
        count = min(count, (int64_t)*dcount);
    const int64_t ntiles = (count + TILE - 1) / TILE;
    // par: one work item per (tile, N block), flags preset to 1 and cleared by whichever block
    // rejects; otherwise one item per tile walking its N blocks in order
    const int nbk = par ? cb1 - cb0 : 1;
    unsigned long long nvis = 0, npts = 0;
    unsigned long long cls[3] = {0, 0, 0};  // lane 0: the wave's FP64-test candidates by reason
    for (int64_t item = (int64_t)blockIdx.x * FAST_WAVES + wv; item < ntiles * nbk;
         item += (int64_t)gridDim.x * FAST_WAVES) {
        const int64_t tile = item / nbk;
        const int nbs = par ? cb0 + (int)(item % nbk) : cb0, nbe = par ? nbs + 1 : cb1;
        prefilter_item<D, SUB, SPARSE, KNOT, RSN>(f, tile, nbs, nbe, par, base, count, list, flags, acc_gid,
                                                  nlimit, nvis, npts, cls, skip);
    }
    if (ctr) {
        unsigned long long a = nvis, b = npts;
        for (int s = 32; s >= 1; s >>= 1) {
            a += __shfl_xor(a, s, 64);
            b += __shfl_xor(b, s, 64);
        }
        unsigned long long* c = ctr + (blockIdx.x % UG_CTR_SLOTS) * UG_CTR_STRIDE;
        if (lane == 0 && a) {
            atomicAdd(&c[0], a);
            atomicAdd(&c[1], b);
        }
        if (RSN && lane == 0)
            for (int k = 0; k < 3; k++)
                if (cls[k])
                    atomicAdd(&c[3 + k], cls[k]);
    }
}

// flags[0 .. min(count, *dcount)) = 1 (the parallel-block mode's starting value).
__global__ void k_flags_one(uint8_t* flags, int64_t count, const int* dcount) {
    if (dcount)
        count = min(count, (int64_t)*dcount);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256)
        flags[i] = 1;
}

// One universe's share of a batched stage-1 prefilter launch (the batched engine).
struct MFast {
    FastArgs f;
    uint64_t base;
    int64_t count;
    const int64_t* list;
    uint8_t* flags;
    uint32_t* skip;  // per flagged candidate: its block-0 columns proven clear
};

// Stage-1 prefilter over the candidates of many universes (at most 64; one dimension): the
// tiles of all of them as one flattened work queue, one tile per wave item, the same work
// item as the single-universe kernel.
template <int D>
__global__ __launch_bounds__(256, D == 2 ? 4 : 3) void k_fast_prefilter_multi(const MFast* __restrict__ A, int U) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    BatchItems q;
    q.scan(lane < U ? (A[lane].count + TILE - 1) / TILE : 0, U);
    unsigned long long nvis = 0, npts = 0;
    for (int64_t item = (int64_t)blockIdx.x * FAST_WAVES + wv; item < q.total;
         item += (int64_t)gridDim.x * FAST_WAVES) {
        int64_t tile = 0;
        const int u = q.owner(item, &tile);
        // the universe's record in registers, read once per item, its pointers global
        const MFast* r = A + u;
        MFast a = *r;
        a.f.m = load_model(&r->f.m);
        a.f.bfrag = ug_gptr(&r->f.bfrag);
        a.f.colt = ug_gptr(&r->f.colt);
        prefilter_item<D, false, false, false>(a.f, tile, 0, 1, 0, a.base, a.count, ug_gptr(&r->list),
                                               ug_gptr(&r->flags), nullptr, 0, nvis, npts, nullptr,
                                               ug_gptr(&r->skip));
    }
}

// Split-BF16 positions of listed candidates at every timestep (error reporting, tests).
template <int D>
__global__ __launch_bounds__(256) void k_fast_eval(FastArgs f, int nb_count, const int64_t* list,
                                                   int64_t count, double* out) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float* pos = (float*)lds + (size_t)wv * (D * TILE * NCOL + 2 * TILE);
    float* sa1 = pos + D * TILE * NCOL;
    float* sinvw = sa1 + TILE;
    const int mrow = lane & 31, h = lane >> 5;
    const int64_t ntiles = (count + TILE - 1) / TILE;
    for (int64_t tile = (int64_t)blockIdx.x * FAST_WAVES + wv; tile < ntiles;
         tile += (int64_t)gridDim.x * FAST_WAVES) {
        const int64_t i = tile * TILE + mrow;
        const bool valid = i < count;
        const uint64_t idx = valid ? (uint64_t)list[i] : 0ull;
        for (int nb = 0; nb < nb_count; nb++) {
            (void)tile_positions<D>(f, idx, valid, lane, nb, f.bfrag, pos, sa1, sinvw);
            for (int col = h; col < NCOL; col += 2) {
                const int t = f.colt[nb * NCOL + col];
                if (valid && t >= 0)
                    for (int a = 0; a < D; a++)
                        out[((size_t)i * f.m.T + t) * D + a] = (double)pos[(a * TILE + mrow) * NCOL + col];
            }
        }
    }
}

// ------------------------------------------------------------------ host side
uint16_t bf16_bits_rne(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf16_to_float(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

struct FastState {
    bf16x8* d_bfrag = nullptr;
    int* d_colt = nullptr;
    int ks = 0, nb = 0, ncols = 0;
    size_t lds_prefilter = 0;
};

__global__ void k_fill_u32x2_zero(uint32_t* p) {
    if (threadIdx.x < 2)
        p[threadIdx.x] = 0u;
}

// Exhaustive accuracy of the FP32 generation (ug_split_generation_error): the phases over all
// 2^32 words, the exponential spacings over all 2^31 distinct Dirichlet words (bit 0 is the
// sign); err[0] = max |(cos, sin) - FP64|, err[1] = max |e - FP64| / (1 + FP64) with e in
// units of ln 2 (the scale the prefilter uses), as float bits.
__global__ __launch_bounds__(256) void k_generation_error(uint32_t* err) {
    float e0 = 0.f, e1 = 0.f;
    for (uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x; w < (1ull << 32); w += (uint64_t)gridDim.x * 256) {
        const double ang = 3.14159265358979323846 * (double)w * 0x1p-32;
        float c, s;
        phase_cs((uint32_t)w, c, s);
        e0 = fmaxf(e0, (float)fmax(fabs((double)c - cos(ang)), fabs((double)s - sin(ang))));
        if ((w & 1u) == 0) {
            // in units of ln 2, as the prefilter computes it
            const double ex = -log2((double)(0x80000000u - ((uint32_t)w >> 1)) * 0x1p-31);
            e1 = fmaxf(e1, (float)(fabs((double)expo2_f32((uint32_t)w) - ex) / (1.0 + ex)));
        }
    }
    for (int sh = 32; sh >= 1; sh >>= 1) {
        e0 = fmaxf(e0, __shfl_xor(e0, sh, 64));
        e1 = fmaxf(e1, __shfl_xor(e1, sh, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(err, __float_as_uint(e0));  // non-negative floats order like their bits
        atomicMax(err + 1, __float_as_uint(e1));
    }
}

}  // namespace

namespace ug {

// Split-BF16 tables of one full-spectrum model, from the host copies of its canonical basis
// (built once per model, see ug_cache.hip).
int32_t fast_build(ModelTables& tb, const std::vector<double>& S, const std::vector<double>& C,
                   const std::vector<int>& order, double margin_req) {
    const int T = tb.T, W = T + 1, npairs = (tb.nw + 1) / 2;
    FastState* fs = new FastState();
    fs->ks = 2 * ((npairs + 7) / 8);  // even: a lane's four pairs span two K steps
    fs->nb = (T + NCOL - 1) / NCOL;
    fs->ncols = T;
    std::vector<int> colt((size_t)fs->nb * NCOL, -1);
    for (int oi = 0; oi < T; oi++)
        colt[oi] = order[oi];
    // basis fragments: row 8hh + i of K step q holds slot s = i % 4 of pair
    // p = 8 (q / 2) + 4 hh + 2 (q % 2) + i / 4 (the generation's pair order), i.e. basis row
    // r = 4 p + s: S_k/k, C_k/k, S_{k+1}/(k+1), 0 (the amplitudes' 1/k folded in)
    std::vector<uint16_t> frag((size_t)fs->nb * fs->ks * 2 * 64 * 8, 0);
    auto basis = [&](int r, int t) -> double {
        const int p = r / 4, s = r % 4;
        const int ke = 2 * p + 2, ko = 2 * p + 3;
        if (t < 0 || p >= npairs)
            return 0.0;
        if (s == 0)
            return S[(size_t)t * W + ke] / ke;
        if (s == 1)
            return C[(size_t)t * W + ke] / ke;
        if (s == 2)
            return ko <= T ? S[(size_t)t * W + ko] / ko : 0.0;
        return 0.0;
    };
    for (int nb = 0; nb < fs->nb; nb++)
        for (int q = 0; q < fs->ks; q++)
            for (int lane = 0; lane < 64; lane++)
                for (int i = 0; i < 8; i++) {
                    const int r = 4 * (8 * (q >> 1) + 4 * (lane >> 5) + 2 * (q & 1) + (i >> 2)) + (i & 3);
                    const int t = colt[(size_t)nb * NCOL + (lane & 31)];
                    const double b = basis(r, t);
                    const uint16_t hb = bf16_bits_rne((float)b);
                    const double rem = b - (double)bf16_to_float(hb);
                    const uint16_t lb = bf16_bits_rne((float)rem);
                    const size_t base = (((size_t)(nb * fs->ks + q) * 2) * 64 + lane) * 8 + i;
                    frag[base] = hb;
                    frag[base + (size_t)64 * 8] = lb;
                }
    // certainty margin per examined column: B_t bounds sum |coef * basis| for unit budget
    const double n_add = 48.0 * fs->ks;
    const double eps = 3.1 * 0x1p-16 + n_add * 0x1p-24 + 0x1p-20;
    std::vector<float> marg((size_t)fs->nb * NCOL, 0.f);  // padding columns are never examined
    double mmax = 0.0;
    for (int col = 0; col < fs->ncols; col++) {
        const int t = colt[col];
        double Bt = 0.0;
        for (int k = 2; k <= T; k++) {
            double v = std::fabs(S[(size_t)t * W + k]);
            if ((k & 1) == 0)
                v += std::fabs(C[(size_t)t * W + k]);
            Bt = std::max(Bt, v / k);
        }
        double mt = Bt * eps + 2e-6;
        mt = std::max(mt, margin_req);
        marg[col] = (float)mt;
        mmax = std::max(mmax, mt);
    }
    tb.margin = mmax;
    tb.fast = fs;  // owned by the tables from here on (kept for the process)
    if (hipMalloc((void**)&fs->d_bfrag, frag.size() * 2) != hipSuccess ||
        hipMalloc((void**)&fs->d_colt, colt.size() * 4) != hipSuccess ||
        hipMalloc((void**)&tb.d_marg, marg.size() * 4) != hipSuccess)
        return fail(UG_ENOMEM, "split-BF16 table allocation failed");
    int32_t rc = hip_check(hipMemcpy(fs->d_bfrag, frag.data(), frag.size() * 2, hipMemcpyHostToDevice), "frag upload");
    if (!rc)
        rc = hip_check(hipMemcpy(fs->d_colt, colt.data(), colt.size() * 4, hipMemcpyHostToDevice), "colt upload");
    if (!rc)
        rc = hip_check(hipMemcpy(tb.d_marg, marg.data(), marg.size() * 4, hipMemcpyHostToDevice), "margin upload");
    fs->lds_prefilter = (size_t)FAST_WAVES * (tb.d * TILE * NCOL + 5 * TILE) * 4;
    if (!rc && fs->lds_prefilter > 160 * 1024)
        rc = fail(UG_EINVAL, "split-BF16 tile needs %zu bytes of LDS", fs->lds_prefilter);
    if (!rc) {
        const void* kf[6] = {
            tb.d == 2 ? (const void*)k_fast_prefilter<2, false, false> : (const void*)k_fast_prefilter<3, false, false>,
            tb.d == 2 ? (const void*)k_fast_prefilter<2, true, false> : (const void*)k_fast_prefilter<3, true, false>,
            tb.d == 2 ? (const void*)k_fast_prefilter<2, false, true> : (const void*)k_fast_prefilter<3, false, true>,
            tb.d == 2 ? (const void*)k_fast_prefilter<2, true, true> : (const void*)k_fast_prefilter<3, true, true>,
            tb.d == 2 ? (const void*)k_fast_prefilter<2, false, false, true>
                      : (const void*)k_fast_prefilter<3, false, false, true>,
            tb.d == 2 ? (const void*)k_fast_prefilter<2, false, false, false, true>
                      : (const void*)k_fast_prefilter<3, false, false, false, true>};
        for (int v = 0; v < 6 && !rc; v++)
            rc = hip_check(hipFuncSetAttribute(kf[v], hipFuncAttributeMaxDynamicSharedMemorySize, (int)fs->lds_prefilter),
                           "lds attribute");
    }
    return rc;
}

// A handle's view of its model's split-BF16 tables.
int32_t fast_setup(Pack* h) {
    if (!h->full)
        return UG_OK;  // partial spectra run the FP64 candidate test (see launch_fast)
    const ModelTables& tb = *h->tabs;
    if (!tb.fast)
        return fail(UG_ESTATE, "split-BF16 tables missing for this model");
    h->d_basis16 = tb.fast;
    h->d_marg = tb.d_marg;
    h->dm.marg = tb.d_marg;
    h->margin = tb.margin;
    return UG_OK;
}

static FastArgs fast_args(Pack* h) {
    FastState* fs = (FastState*)h->d_basis16;
    FastArgs f;
    f.m = h->dm;
    f.ks = fs->ks;
    f.ncols = fs->ncols;
    f.bfrag = fs->d_bfrag;
    f.colt = fs->d_colt;
    f.kpos32 = nullptr;
    f.kstride = 0;
    f.nk = 0;
    return f;
}

size_t fast_multi_arg_bytes() { return sizeof(MFast); }

int32_t fast_multi_arg(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags, uint32_t* skip,
                       void* out) {
    if (!h->full || !h->d_basis16)
        return fail(UG_ESTATE, "split-BF16 prefilter not set up for this model");
    if (h->dm.sparse)
        return fail(UG_EINVAL, "the batched prefilter covers the dense cell table only");
    MFast a;
    a.f = fast_args(h);
    a.base = base;
    a.count = count;
    a.list = list;
    a.flags = flags;
    a.skip = skip;
    memcpy(out, &a, sizeof a);
    return UG_OK;
}

int32_t launch_fast_multi(hipStream_t st, int d, const void* dA, int U, int64_t tiles) {
    if (U <= 0 || tiles <= 0)
        return UG_OK;
    if (U > 64)
        return fail(UG_EINVAL, "batched prefilter of %d universes (at most 64 a launch)", U);
    const size_t lds = (size_t)FAST_WAVES * (d * TILE * NCOL + 5 * TILE) * 4;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + FAST_WAVES - 1) / FAST_WAVES, 16384));
    if (d == 2)
        hipLaunchKernelGGL(k_fast_prefilter_multi<2>, dim3(blocks), dim3(256), lds, st, (const MFast*)dA, U);
    else
        hipLaunchKernelGGL(k_fast_prefilter_multi<3>, dim3(blocks), dim3(256), lds, st, (const MFast*)dA, U);
    return hip_check(hipGetLastError(), "k_fast_prefilter_multi");
}

int fast_blocks(const Pack* h) {
    return h->d_basis16 ? ((const FastState*)h->d_basis16)->nb : 0;
}

int32_t launch_fast(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags,
                    const int* dcount, int cb0, int cb1, uint32_t* skip) {
    if (count <= 0)
        return UG_OK;
    if (!h->full || !h->d_basis16)
        return fail(UG_ESTATE, "split-BF16 prefilter not set up for this model");
    FastState* fs = (FastState*)h->d_basis16;
    const int64_t ntiles = (count + TILE - 1) / TILE;
    // a device-counted list is short (what an earlier stage left): fewer persistent blocks.
    // Otherwise one tile per wave up to 2^20 candidates: tiles finish at uneven times (the
    // lookup stops at the first certain contact), and finer work balances the tail better
    // (one MI355X, 2^20-candidate launches: 2+1 T=64 142 -> 138 us, 3+1 T=128 436 -> 413 us
    // against a 2048-block cap; a 1024-block cap was slower still)
    const int cap = (dcount && list) ? 512 : 8192;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((ntiles + FAST_WAVES - 1) / FAST_WAVES, cap));
    const FastArgs f = fast_args(h);
    unsigned long long* ctr = h->profile ? h->d_ctr : nullptr;
    cb1 = std::min(cb1, fs->nb);
    if (cb0 >= cb1)
        return fail(UG_EINVAL, "prefilter: empty column-block range [%d, %d)", cb0, cb1);
    // several N blocks of a short list run in parallel (one work item per tile and block)
    const int par = (dcount && list && cb1 - cb0 > 1) ? 1 : 0;
    const int nblk = par ? std::min(2048, blocks * (cb1 - cb0)) : blocks;
    if (par)
        hipLaunchKernelGGL(k_flags_one, dim3(grid_for(count, 256 * 16)), dim3(256), 0, h->stream, flags, count, dcount);
#define UG_FAST(DD, SP)                                                                                     \
    UG_LAUNCH(h, (k_fast_prefilter<DD, false, SP>), dim3(nblk), dim3(256), fs->lds_prefilter, f, base, count, \
              list, dcount, cb0, cb1, par, flags, ctr, nullptr, 0, skip)
    // the reason counts (UG_PROFILE_REASONS) in their own instantiation, off the timed path
    const bool rsn = ctr && ((h->profile_mask >> 2) & 1) && !par && !f.m.sparse;
#define UG_FASTR(DD)                                                                                      \
    UG_LAUNCH(h, (k_fast_prefilter<DD, false, false, false, true>), dim3(nblk), dim3(256), fs->lds_prefilter, \
              f, base, count, list, dcount, cb0, cb1, par, flags, ctr, nullptr, 0, skip)
    if (rsn && h->d == 2)
        UG_FASTR(2);
    else if (rsn)
        UG_FASTR(3);
    else if (h->d == 2 && !f.m.sparse)
        UG_FAST(2, false);
    else if (h->d == 2)
        UG_FAST(2, true);
    else if (!f.m.sparse)
        UG_FAST(3, false);
    else
        UG_FAST(3, true);
#undef UG_FASTR
#undef UG_FAST
    return hip_check(hipGetLastError(), "k_fast_prefilter");
}

int32_t launch_fast_knot(Pack* h, uint64_t base, int64_t count, uint8_t* flags, const float4* kpos32, int64_t ks,
                         int nk) {
    if (count <= 0)
        return UG_OK;
    if (!h->full || !h->d_basis16)
        return fail(UG_ESTATE, "split-BF16 prefilter not set up for this model");
    FastState* fs = (FastState*)h->d_basis16;
    const int64_t ntiles = (count + TILE - 1) / TILE;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((ntiles + FAST_WAVES - 1) / FAST_WAVES, 8192));
    FastArgs f = fast_args(h);
    f.kpos32 = kpos32;
    f.kstride = ks;
    f.nk = nk;
    if (h->d == 2)
        UG_LAUNCH(h, (k_fast_prefilter<2, false, false, true>), dim3(blocks), dim3(256), fs->lds_prefilter, f, base,
                  count, nullptr, nullptr, 0, 1, 0, flags, nullptr, nullptr, 0, nullptr);
    else
        UG_LAUNCH(h, (k_fast_prefilter<3, false, false, true>), dim3(blocks), dim3(256), fs->lds_prefilter, f, base,
                  count, nullptr, nullptr, 0, 1, 0, flags, nullptr, nullptr, 0, nullptr);
    return hip_check(hipGetLastError(), "k_fast_prefilter (knot)");
}

int32_t launch_fast_sub(Pack* h, uint64_t base, int64_t count, uint8_t* flags, int64_t nlimit) {
    if (count <= 0)
        return UG_OK;
    if (!h->full || !h->d_basis16)
        return fail(UG_ESTATE, "split-BF16 prefilter not set up for this model");
    FastState* fs = (FastState*)h->d_basis16;
    const int64_t ntiles = (count + TILE - 1) / TILE;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((ntiles + FAST_WAVES - 1) / FAST_WAVES, 2048));
    const FastArgs f = fast_args(h);
#define UG_FASTSUB(DD, SP)                                                                                         \
    hipLaunchKernelGGL((k_fast_prefilter<DD, true, SP>), dim3(blocks), dim3(256), fs->lds_prefilter, h->stream, f, \
                       base, count, nullptr, nullptr, 0, 1, 0, flags, nullptr, h->d_acc_gid, nlimit, nullptr)
    if (h->d == 2 && !f.m.sparse)
        UG_FASTSUB(2, false);
    else if (h->d == 2)
        UG_FASTSUB(2, true);
    else if (!f.m.sparse)
        UG_FASTSUB(3, false);
    else
        UG_FASTSUB(3, true);
#undef UG_FASTSUB
    return hip_check(hipGetLastError(), "k_fast_prefilter (subpaths)");
}

int32_t launch_fast_eval(Pack* h, const int64_t* list, int64_t count, double* out) {
    if (!h->full)
        return fail(UG_EINVAL, "split-BF16 evaluation covers full-spectrum models only (terms = T)");
    if (!h->d_basis16) {
        int32_t rc = fast_setup(h);
        if (rc)
            return rc;
    }
    if (count <= 0)
        return UG_OK;
    FastState* fs = (FastState*)h->d_basis16;
    const int64_t ntiles = (count + TILE - 1) / TILE;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((ntiles + FAST_WAVES - 1) / FAST_WAVES, 1024));
    const size_t lds = (size_t)FAST_WAVES * (h->d * TILE * NCOL + 2 * TILE) * 4;
    const FastArgs f = fast_args(h);
    if (h->d == 2)
        hipLaunchKernelGGL(k_fast_eval<2>, dim3(blocks), dim3(256), lds, h->stream, f, fs->nb, list, count, out);
    else
        hipLaunchKernelGGL(k_fast_eval<3>, dim3(blocks), dim3(256), lds, h->stream, f, fs->nb, list, count, out);
    return hip_check(hipGetLastError(), "k_fast_eval");
}

}  // namespace ug

extern "C" int32_t ug_split_generation_error(void* hip_stream, double* out2) {
    if (!out2)
        return ug::fail(UG_EINVAL, "ug_split_generation_error: NULL output");
    hipStream_t st = (hipStream_t)hip_stream;
    uint32_t* d = nullptr;
    int32_t rc = ug::hip_check(hipMalloc((void**)&d, 8), "generation error buffer");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_fill_u32x2_zero, dim3(1), dim3(64), 0, st, d);
    hipLaunchKernelGGL(k_generation_error, dim3(16384), dim3(256), 0, st, d);
    uint32_t h[2] = {0, 0};
    rc = ug::hip_check(hipGetLastError(), "k_generation_error");
    if (!rc)
        rc = ug::hip_check(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, st), "generation error copy");
    if (!rc)
        rc = ug::hip_check(hipStreamSynchronize(st), "generation error sync");
    (void)hipFree(d);
    float f[2];
    memcpy(f, h, 8);
    out2[0] = f[0];
    out2[1] = f[1];
    return rc;
}
