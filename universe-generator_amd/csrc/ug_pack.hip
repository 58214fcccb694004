// ug_pack.hip -- MI355X RSA worldline packer: kernels, host driver and C ABI.
//
// Path (reference universe-analysis/universe-generator):
//   braidlab/engine.py:50-112 -> cuda/braid_cuda.cu:519-1106 (2+1) / cuda/braid_cuda3d.cu:849-1672 (3+1)
// The reference prefilters a batch on the GPU with atomicAdd survivor slots
// (cuda/braid_cuda.cu:475-510) and admits survivors serially on the host
// (:903-919).  Here every stage stays on the device and is deterministic:
//   1. candidate test   one wave per candidate: Philox draws -> FP64 coefficients
//                       -> lazy evaluation in the endpoint-first order -> 3^d cell
//                       lookup per timestep with early exit        (k_cand64)
//      [split-BF16 form: tile generation in FP32 + MFMA evaluation + certain-hit
//       classification, FP64 recheck of everything uncertain       (ug_fast.hip)]
//   2. ordered compaction of survivors by candidate index          (k_compact_*)
//   3. full FP64 trajectories of the survivors                     (k_cand64<FULL>)
//   4. pairwise conflicts among survivors + greedy in index order  (k_conflicts_cells, k_greedy)
//   5. insertion into the per-timestep cell table                  (k_insert)
// so the accepted list equals serial RSA in candidate-index order (cuda/braid_knot.cu:433-462).
#include <stdarg.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ug_arith.h"
#include "ug_internal.h"

// Second split-BF16 stage from this many N blocks of 32 ordered timesteps (T > 160).
#ifndef UG_STAGE2_MIN_BLOCKS
#define UG_STAGE2_MIN_BLOCKS 6
#endif

using namespace ug;

// =====================================================================================
// errors
static thread_local std::string g_err;

int32_t ug::fail(int32_t code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char* ug_last_error(void) { return g_err.c_str(); }
extern "C" int32_t ug_abi_version(void) { return UG_ABI_VERSION; }

// =====================================================================================
// wave helpers
__device__ __forceinline__ double wave_sum_canonical(double v) {
    // Halving tree over the 64 lane partials: step h adds lane l^h.  Lanes l and l^h
    // compute the same commutative sum, so every lane ends with the canonical total.
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1)
        v = v + __shfl_xor(v, h, 64);
    return v;
}

// =====================================================================================
// per-wave candidate coefficients (FP64, canonical)
template <int D, int PL>
struct WaveCoef {
    double a1[D];
    double c[D][PL][2];
    double s[D][PL][2];
    int b[D][PL][2];  // frequency of the slot, 0 = no slot
};

// Frequencies of a partial spectrum (nw < modmax): lane 0 replays the unique-frequency
// rejection draws of cuda/braid_cuda.cu:220-243 from stage-2 words into LDS.
template <int D>
__device__ void draw_freqs_partial(const DevModel& m, uint64_t idx, int* sb, uint32_t* used) {
    const uint32_t lo = (uint32_t)idx, hi = (uint32_t)(idx >> 32);
    for (int a = 0; a < D; a++) {
        const int nwords = (m.modmax + 2 + 31) / 32;
        for (int i = 0; i < nwords; i++)
            used[i] = 0u;
        uint32_t g = 0;
        u32x4 wb{0, 0, 0, 0};
        for (int j = 0; j < m.nw; j++) {
            int bb = 0;
            bool dup = true;
            for (int guard = 0; dup && guard < 64; guard++) {
                if ((g & 3u) == 0)
                    wb = philox(lo, hi, g >> 2, (2u << 8) | (uint32_t)a, m.key0, m.key1);
                const uint32_t wv = (g & 3u) == 0 ? wb.x : (g & 3u) == 1 ? wb.y : (g & 3u) == 2 ? wb.z : wb.w;
                bb = (int)(wv % (uint32_t)m.modmax) + 2;
                g++;
                dup = (used[bb >> 5] >> (bb & 31)) & 1u;
            }
            if (dup) {
                for (int cc = 2; cc <= m.modmax + 1; cc++)
                    if (!((used[cc >> 5] >> (cc & 31)) & 1u)) {
                        bb = cc;
                        break;
                    }
            }
            used[bb >> 5] |= 1u << (bb & 31);
            sb[a * m.nw + j] = bb;
        }
    }
}

// The slot arrays of the lane-per-timestep evaluation use a padded stride of 128 PL slots per
// axis (PL = slot pairs per lane: 1 up to 64 pairs, 2 up to 128, else 4).
__host__ __device__ inline int slot_pl(int npairs) { return npairs <= 64 ? 1 : (npairs <= 128 ? 2 : 4); }

template <int D>
__host__ __device__ inline int cand64_lds_head(const DevModel& m, int pl) {
    const int draws = D * m.nw + (m.modmax + 2 + 31) / 32;
    const int slots = D * 128 * pl;
    return ((draws > slots ? draws : slots) + 3) / 4 * 4;
}

template <int D>
__host__ __device__ inline int cand64_lds_words(const DevModel& m) {
    const int pl = slot_pl(m.npairs);
    return cand64_lds_head<D>(m, pl) + D * 128 * pl * 4 + 64 * D * 2;
}

// The reference's serial stream (rng_kind UG_RNG_XOSHIRO_REF; the XR instantiations of the
// kernels, so the Philox ones keep their register allocation): lane 0 replays the
// candidate's draws (xo_axis, every axis in stream order) into LDS -- the frequencies at
// sb[a nw + j], the spacing and sign/phase uniforms of slot j of axis a at
// xr[2 (a NWP + j)] and xr[2 (a NWP + j) + 1], past the frequency area -- then the lanes turn
// their own slots into coefficients.  The spacings are summed serially in slot order (as
// cuda/braid_knot.cu:114-117), every lane the same sum.  The callers' dynamic LDS holds
// cand64_lds_words words.
template <int D, int PL>
__device__ __forceinline__ void gen_wave64_xref(const DevModel& m, uint64_t idx, int lane, WaveCoef<D, PL>& w,
                                                int* sb) {
    constexpr int NWP = 128 * PL;
    const int nw = m.nw;
    uint64_t* xr = (uint64_t*)(sb + cand64_lds_head<D>(m, PL));
    __shared__ uint64_t xa[3];
    if (lane == 0) {
        xo256 r = xo_candidate((uint64_t)m.key0 | (uint64_t)m.key1 << 32, idx);
        for (int a = 0; a < D; a++)
            xa[a] = xo_axis(r, m.modmax, nw, m.pin, sb + a * nw, xr + 2 * a * NWP, xr + 2 * a * NWP + 1, 2);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < D; a++) {
        w.a1[a] = m.pin ? 0.0 : xo_a1(xa[a]);
        double e[PL][2];
#pragma unroll
        for (int i = 0; i < PL; i++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int j = 2 * (lane + 64 * i) + k;
                e[i][k] = 0.0;
                if (j < nw) {
                    // each lane rewrites only its own slots' entries
                    e[i][k] = xo_spacing(xr[2 * (a * NWP + j)]);
                    xr[2 * (a * NWP + j)] = (uint64_t)__double_as_longlong(e[i][k]);
                }
            }
        __syncthreads();
        double sum = 0.0;
        for (int j = 0; j < nw; j++)
            sum = sum + __longlong_as_double((long long)xr[2 * (a * NWP + j)]);
#pragma unroll
        for (int i = 0; i < PL; i++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int j = 2 * (lane + 64 * i) + k;
                double cj = 0.0, sj = 0.0;
                int bj = 2;  // unused slots: zero coefficients on a valid basis row
                if (j < nw) {
                    bj = sb[a * nw + j];
                    const uint64_t pk = xr[2 * (a * NWP + j) + 1];
                    double aj = (e[i][k] / sum) / (double)bj;
                    if (pk >> 63)
                        aj = -aj;
                    if ((bj & 1) == 0) {
                        const sc2 f = sincospi_rat(pk & ((1ull << 53) - 1ull), 1ull << 53);
                        cj = aj * f.c;
                        sj = aj * f.s;
                    } else {
                        cj = aj;
                    }
                }
                w.c[a][i][k] = cj;
                w.s[a][i][k] = sj;
                w.b[a][i][k] = bj;
            }
    }
    __syncthreads();  // the caller reuses both areas
}

template <int D, int PL, bool XR = false>
__device__ void gen_wave64(const DevModel& m, uint64_t idx, int lane, WaveCoef<D, PL>& w, int* sb,
                           uint32_t* used) {
    if constexpr (XR) {
        gen_wave64_xref<D, PL>(m, idx, lane, w, sb);
        return;
    }
    const uint32_t lo = (uint32_t)idx, hi = (uint32_t)(idx >> 32);
    const u32x4 w0 = philox(lo, hi, 0u, 0u, m.key0, m.key1);
    const uint32_t wa[3] = {w0.x, w0.y, w0.z};
#pragma unroll
    for (int a = 0; a < D; a++)
        w.a1[a] = m.pin ? 0.0 : a1_of_word(wa[a]);
    if (!m.full) {
        if (lane == 0)
            draw_freqs_partial<D>(m, idx, sb, used);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        double e[PL][2];
        uint32_t neg[PL][2], ph[PL][2];
        double q = 0.0;
#pragma unroll
        for (int i = 0; i < PL; i++) {
            const int p = lane + 64 * i;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                e[i][k] = 0.0;
                neg[i][k] = 0;
                ph[i][k] = 0;
                w.b[a][i][k] = 0;
            }
            if (p < m.npairs) {
                const u32x4 r = m.full ? pair_words_full(lo, hi, p, a, m.key0, m.key1)
                                       : philox(lo, hi, (uint32_t)p, (1u << 8) | (uint32_t)a, m.key0, m.key1);
                const uint32_t dw[2] = {r.x, r.y}, pw[2] = {r.z, r.w};
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int j = 2 * p + k;
                    if (j < m.nw) {
                        e[i][k] = expo_of_word(dw[k]);
                        neg[i][k] = dw[k] & 1u;
                        ph[i][k] = pw[k];
                        w.b[a][i][k] = m.full ? j + 2 : sb[a * m.nw + j];  // read before the caller reuses sb
                        q = q + e[i][k];
                    }
                }
            }
        }
        const double W = wave_sum_canonical(q);
        const double invW = 1.0 / W;
#pragma unroll
        for (int i = 0; i < PL; i++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int bj = w.b[a][i][k];
                double cj = 0.0, sj = 0.0;
                if (bj) {
                    const double invb = 1.0 / (double)bj;
                    double aj = (m.nw == 1) ? invb : (e[i][k] * invW) * invb;
                    if (neg[i][k])
                        aj = -aj;
                    if ((bj & 1) == 0) {
                        const sc2 f = sincospi_rat((uint64_t)ph[i][k], 0x100000000ull);
                        cj = aj * f.c;
                        sj = aj * f.s;
                    } else {
                        cj = aj;
                    }
                }
                w.c[a][i][k] = cj;
                w.s[a][i][k] = sj;
                // unused slots: zero coefficients on a valid basis row, so evaluation needs no
                // branches (an accumulator that starts at +0 is never -0, hence
                // fma(0, v, acc) == acc exactly)
                if (!bj)
                    w.b[a][i][k] = 2;
            }
    }
}

// Canonical wrapped position of the wave's candidate at timestep t (all lanes): lane
// partials over its slot pairs, then the halving tree across lanes.
template <int D, int PL>
__device__ __forceinline__ void eval_wave64(const DevModel& m, const WaveCoef<D, PL>& w, int t,
                                            double* X) {
    const double* St = m.S + (size_t)t * m.ldb;
    const double* Ct = m.C + (size_t)t * m.ldb;
#pragma unroll
    for (int a = 0; a < D; a++) {
        double part = 0.0;
#pragma unroll
        for (int i = 0; i < PL; i++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                // odd frequencies and unused slots carry s = 0 (and unused c = 0): adding
                // their zero products leaves the sum unchanged, bit for bit
                const int bj = w.b[a][i][k];
                part = fma(w.c[a][i][k], St[bj], part);
                part = fma(w.s[a][i][k], Ct[bj], part);
            }
        X[a] = torus_wrap(w.a1[a] + wave_sum_canonical(part));
    }
}

__host__ __device__ constexpr int bitrev3(int i) {
    return ((i & 1) << 2) | (i & 2) | ((i & 4) >> 2);
}

// The same canonical position, evaluated by ONE lane (lanes take different timesteps).
// Slot partial P_s is the fma chain over the pairs p = s + 64 i (ascending), each adding
// c S_b then s C_b (s = 0 for odd frequencies and padding slots, whose zero products leave
// the chain unchanged bit for bit, so no branches are needed).
// The halving tree over the 64 slot partials equals the adjacent pairwise tree over the
// slots in bit-reversed order; element 8g + k of that order is slot 8 bitrev3(k) + bitrev3(g),
// so each group of 8 is a fixed 3-level tree and the 8 group sums merge as a binary counter.
template <int D, int PL>
__device__ __forceinline__ void eval_lane(const DevModel& m, const double2* cs, const int* bb,
                                          const double* a1, int t, double* X) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double2* ca = cs + a * (128 * PL);
        const int* ba = bb + a * (128 * PL);
        double L0 = 0.0, L1 = 0.0, L2 = 0.0, total = 0.0;
#pragma unroll 1
        for (int g = 0; g < 8; g++) {
            const int rg = bitrev3(g);
            double e[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                e[k] = 0.0;
#pragma unroll
            for (int ii = 0; ii < PL; ii++) {
                // the 32 basis values of this group's slot pairs are loaded in one batch (the
                // scheduling barrier keeps the loads ahead of the fma chains that use them)
                double sv[8][2], cv[8][2];
                double2 cc[8][2];
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const int j = 2 * (8 * bitrev3(k) + rg + 64 * ii) + kk;
                        const int b = ba[j];
                        cc[k][kk] = ca[j];
                        sv[k][kk] = m.ST[(size_t)b * m.T + t];
                        cv[k][kk] = m.CT[(size_t)b * m.T + t];
                    }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        e[k] = fma(cc[k][kk].x, sv[k][kk], e[k]);
                        e[k] = fma(cc[k][kk].y, cv[k][kk], e[k]);
                    }
            }
            double G = ((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7]));
            if ((g & 1) == 0) {
                L0 = G;
                continue;
            }
            G = L0 + G;
            if ((g & 2) == 0) {
                L1 = G;
                continue;
            }
            G = L1 + G;
            if ((g & 4) == 0) {
                L2 = G;
                continue;
            }
            total = L2 + G;
        }
        X[a] = torus_wrap(a1[a] + total);
    }
}

// eval_lane for full spectra: slot j carries frequency j + 2 on every axis (2 on padding
// slots), so the 32 basis values of a slot group are loaded once for all axes; each axis's
// fma chains and tree see exactly the operations, in the same order, as in eval_lane.
template <int D, int PL>
__device__ __forceinline__ void eval_lane_full(const DevModel& m, const double2* cs, const int* bb,
                                               const double* a1, int t, double* X) {
    double L0[D], L1[D], L2[D], total[D];
#pragma unroll
    for (int a = 0; a < D; a++)
        L0[a] = L1[a] = L2[a] = total[a] = 0.0;
#pragma unroll 1
    for (int g = 0; g < 8; g++) {
        const int rg = bitrev3(g);
        double e[D][8];
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int k = 0; k < 8; k++)
                e[a][k] = 0.0;
#pragma unroll
        for (int ii = 0; ii < PL; ii++) {
            double sv[8][2], cv[8][2];
#pragma unroll
            for (int k = 0; k < 8; k++)
#pragma unroll
                for (int kk = 0; kk < 2; kk++) {
                    const int j = 2 * (8 * bitrev3(k) + rg + 64 * ii) + kk;
                    const int b = bb[j];
                    sv[k][kk] = m.ST[(size_t)b * m.T + t];
                    cv[k][kk] = m.CT[(size_t)b * m.T + t];
                }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int a = 0; a < D; a++) {
                const double2* ca = cs + a * (128 * PL);
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const double2 c = ca[2 * (8 * bitrev3(k) + rg + 64 * ii) + kk];
                        e[a][k] = fma(c.x, sv[k][kk], e[a][k]);
                        e[a][k] = fma(c.y, cv[k][kk], e[a][k]);
                    }
            }
        }
#pragma unroll
        for (int a = 0; a < D; a++) {
            double G = ((e[a][0] + e[a][1]) + (e[a][2] + e[a][3])) + ((e[a][4] + e[a][5]) + (e[a][6] + e[a][7]));
            if ((g & 1) == 0) {
                L0[a] = G;
                continue;
            }
            G = L0[a] + G;
            if ((g & 2) == 0) {
                L1[a] = G;
                continue;
            }
            G = L1[a] + G;
            if ((g & 4) == 0) {
                L2[a] = G;
                continue;
            }
            total[a] = L2[a] + G;
        }
    }
#pragma unroll
    for (int a = 0; a < D; a++)
        X[a] = torus_wrap(a1[a] + total[a]);
}

// Lookup round: lanes take (timestep slot sl, 3-cell row rr) items of the timesteps whose
// positions are in pos[]; returns true when any lane finds a contact (canonical predicate).
template <int D>
__device__ __forceinline__ bool lookup_round(const DevModel& m, const double* pos, const int* tsl,
                                             int G, int lane, unsigned long long& npts) {
    constexpr int R = D == 2 ? 3 : 9;
    const int T = m.T;
    const int sl = lane / R, rr = lane % R;
    bool h = false;
    if (sl < G) {
        const int t = tsl[sl];
        double x[D];
        int c[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            x[a] = pos[sl * D + a];
            c[a] = cell_of(x[a], m.halfT, T);
        }
        const int gx = wrap_cell(c[0], (D == 2 ? rr : rr / 3) - 1, T);
        const int gy = D == 2 ? 0 : wrap_cell(c[1], rr % 3 - 1, T);
        const int64_t rid = D == 2 ? (int64_t)t * T + gx : ((int64_t)t * T + gx) * T + gy;
        // the occupied cells' heads load together and their chains are walked together
        int node[3];
        row_heads3<D>(m, t, gx, gy, rid, c[D - 1], node);
        while (!h && (node[0] >= 0 || node[1] >= 0 || node[2] >= 0)) {
            double P[3][D];
            int nxt[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                nxt[k] = -1;
                if (node[k] >= 0) {
                    const size_t q = (size_t)node[k] * T + t;
#pragma unroll
                    for (int a = 0; a < D; a++)
                        P[k][a] = m.pos64[q * D + a];
                    nxt[k] = __float_as_int(m.pos32[q].w);
                }
            }
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (node[k] >= 0) {
                    h = h || contact<D>(m, x, P[k]);
                    npts++;
                    node[k] = nxt[k];
                }
        }
    }
    return __ballot(h) != 0;
}

// One position's whole neighbourhood by one lane (the bulk phase, one lane per timestep): the
// occupancy words of all its rows first, then the heads of the occupied cells NG at a time,
// their chains walked together -- about three dependent memory steps for 64 timesteps at
// once, where lookup_round spreads one timestep's rows over lanes and needs a round per
// 64 / 3^(D-1) timesteps.  The same cells as row_heads3; true on a contact.
template <int D>
__device__ __forceinline__ bool lookup_lane(const DevModel& m, const double* x, int t, unsigned long long& npts) {
    constexpr int R = D == 2 ? 3 : 9;
    const int T = m.T;
    int c[D];
#pragma unroll
    for (int a = 0; a < D; a++)
        c[a] = cell_of(x[a], m.halfT, T);
    auto gx_of = [&](int rr) { return wrap_cell(c[0], (D == 2 ? rr : rr / 3) - 1, T); };
    auto gy_of = [&](int rr) { return D == 2 ? 0 : wrap_cell(c[1], rr % 3 - 1, T); };
    uint32_t cells = 0;
#pragma unroll
    for (int rr = 0; rr < R; rr++)
        cells |= occ_bits3<D>(m, t, gx_of(rr), gy_of(rr), c[D - 1]) << (3 * rr);
    bool h = false;
    constexpr int NG = 4;
    while (cells && !h) {
        int node[NG];
#pragma unroll
        for (int g = 0; g < NG; g++) {
            node[g] = -1;
            if (cells) {
                const int b = __builtin_ctz(cells);
                cells &= cells - 1u;
                const int rr = b / 3, k = b % 3;
                const int64_t rid =
                    D == 2 ? (int64_t)t * T + gx_of(rr) : ((int64_t)t * T + gx_of(rr)) * T + gy_of(rr);
                node[g] = m.sparse ? hash_head(m, cell_hkey(rid, wrap_cell(c[D - 1], k - 1, T), T))
                                   : m.tab[rid * m.rowlen + c[D - 1] + k];
            }
        }
        bool any = false;
#pragma unroll
        for (int g = 0; g < NG; g++)
            any = any || node[g] >= 0;
        while (!h && any) {
            double P[NG][D];
            int nx[NG];
#pragma unroll
            for (int g = 0; g < NG; g++) {
                nx[g] = -1;
                if (node[g] >= 0) {
                    const size_t q = (size_t)node[g] * T + t;
#pragma unroll
                    for (int a = 0; a < D; a++)
                        P[g][a] = m.pos64[q * D + a];
                    nx[g] = __float_as_int(m.pos32[q].w);
                }
            }
            any = false;
#pragma unroll
            for (int g = 0; g < NG; g++)
                if (node[g] >= 0) {
                    h = h || contact<D>(m, x, P[g]);
                    npts++;
                    node[g] = nx[g];
                    any = any || node[g] >= 0;
                }
        }
    }
    return h;
}

// One wave per candidate.  Lanes generate the coefficients over slot pairs.  The first
// `head` endpoint-first timesteps (where most candidates meet a contact) are evaluated across
// the wave (canonical lane tree) in chunks of 2, 4, 8, ... (at most 64 / 3^(D-1)), each
// chunk looked up in one round; a candidate still clear after them has every remaining
// timestep evaluated in parallel, one lane per timestep, each lane looking up its own
// timestep's neighbourhood (lookup_lane; cuda/braid_cuda.cu:418-462).  Bulk tests of fresh
// candidates use
// head = T; lists of likely survivors use a short head.
// FULL = false: flags[i] = 1 iff candidate i survives the cell table.
// FULL = true: out[i][T][D] = canonical trajectory (no lookup).
// SB: the lane-per-timestep evaluation shares each slot group's basis loads across axes
// (full spectra, one slot pair per lane).  A separate instantiation, because the shared
// evaluation's extra registers would lower the occupancy of launches that do not use it.
// One candidate (item i of the launch, candidate index idx) of k_cand64; the wave is the
// workgroup.  Shared by the single-universe kernel and the batched engine's launches.
// known >= 0: the candidate's scan positions < min(T, 32) were examined by the split-BF16
// prefilter and bit c of known marks position c proven to hold no contact; only the others
// (usually a handful) are evaluated across the wave before the bulk scan starts at 32.
template <int D, bool FULL, int PL, bool SB, bool XR = false>
__device__ __forceinline__ void cand64_item(const DevModel& m, int64_t i, uint64_t idx, uint8_t* flags,
                                            double* out, int head, int* cells, uint64_t* mask,
                                            unsigned long long& nvis, unsigned long long& npts,
                                            int64_t known = -1) {
    extern __shared__ __align__(16) int dyn_lds[];
    const int lane = threadIdx.x;
    const int T = m.T, nw = m.nw;
    int* sb = dyn_lds;
    uint32_t* used = (uint32_t*)(sb + D * nw);
    constexpr int NWP = 128 * PL;  // padded slots per axis
    const int lds_head = cand64_lds_head<D>(m, PL);
    double2* cs = (double2*)(dyn_lds + lds_head);
    double* pos = (double*)(cs + D * NWP);
    __shared__ int tsl[64];
    constexpr int R = D == 2 ? 3 : 9;
    constexpr int PER_ROUND = 64 / R;
    {
        if (FULL && mask && lane < UG_MASK_WORDS)
            mask[(size_t)i * UG_MASK_WORDS + lane] = 0ull;  // the conflict row, cleared
        WaveCoef<D, PL> w;
        gen_wave64<D, PL, XR>(m, idx, lane, w, sb, used);
        bool hit = false;
        int o0 = 0;
        const int kn = min(T, 32);
        uint32_t rest = known >= 0 ? ~(uint32_t)known & (kn >= 32 ? ~0u : ((1u << kn) - 1u)) : 0u;
        if (!FULL && known >= 0 && __popc(rest) <= head) {
            // positions the prefilter could not clear, PER_ROUND per lookup round, then the
            // bulk scan from position kn
            while (!hit && rest) {
                int G = 0;
                while (rest && G < PER_ROUND) {
                    const int pidx = __builtin_ctz(rest);
                    rest &= rest - 1u;
                    double X[D];
                    const int t = m.order[pidx];
                    eval_wave64<D, PL>(m, w, t, X);
                    if (lane == G) {
                        tsl[G] = t;
#pragma unroll
                        for (int a = 0; a < D; a++)
                            pos[G * D + a] = X[a];
                    }
                    G++;
                }
                __syncthreads();
                hit = lookup_round<D>(m, pos, tsl, G, lane, npts);
                nvis += (unsigned long long)G;
                __syncthreads();
            }
            o0 = kn;
        } else if (!FULL) {
            // head: chunks of 2, 4, 8 timesteps evaluated across the wave, one lookup round
            // each (the chunk stops growing at PER_ROUND: a head of T = 256 in 3+1 takes 37
            // rounds, and an ever-doubling int would overflow)
            for (int chunk = 2; !hit && o0 < min(head, T); chunk = min(2 * chunk, PER_ROUND)) {
                const int G = min(min(chunk, PER_ROUND), T - o0);
                for (int g = 0; g < G; g++) {
                    double X[D];
                    const int t = m.order[o0 + g];
                    eval_wave64<D, PL>(m, w, t, X);
                    if (lane == g) {
                        tsl[g] = t;
#pragma unroll
                        for (int a = 0; a < D; a++)
                            pos[g * D + a] = X[a];
                    }
                }
                __syncthreads();
                hit = lookup_round<D>(m, pos, tsl, G, lane, npts);
                nvis += (unsigned long long)G;
                o0 += G;
                __syncthreads();
            }
        }
        if (!hit && o0 < T) {
            // remaining timesteps: coefficients to LDS, one lane per timestep
#pragma unroll
            for (int a = 0; a < D; a++)
#pragma unroll
                for (int ii = 0; ii < PL; ii++)
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        const int j = 2 * (lane + 64 * ii) + k;
                        cs[a * NWP + j] = make_double2(w.c[a][ii][k], w.s[a][ii][k]);
                        sb[a * NWP + j] = w.b[a][ii][k];
                    }
            __syncthreads();
            for (int ps = o0; ps < T && !hit; ps += 64) {
                const int oi = ps + lane;
                const int pend = min(64, T - ps);
                bool hl = false;
                if (oi < T) {
                    const int t = FULL ? oi : m.order[oi];
                    double X[D];
                    if constexpr (SB)
                        eval_lane_full<D, PL>(m, cs, sb, w.a1, t, X);
                    else
                        eval_lane<D, PL>(m, cs, sb, w.a1, t, X);
                    if (FULL) {
                        // trajectories of a resolve chunk; with cells, also their packed
                        // cells (the conflict test's first filter)
                        int v = 0;
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            out[((size_t)i * T + t) * D + a] = X[a];
                            v |= cell_of(X[a], m.halfT, T) << (10 * a);
                        }
                        if (cells)
                            cells[(size_t)i * T + t] = v;
                    } else {
                        hl = lookup_lane<D>(m, X, t, npts);
                    }
                }
                if (FULL)
                    continue;
                hit = __ballot(hl) != 0;
                nvis += (unsigned long long)pend;
            }
        }
        if (!FULL && lane == 0)
            flags[i] = hit ? 0 : 1;
        __syncthreads();
    }
}

// Measurement counters of a k_cand64 wave (one atomic per wave and counter).
__device__ __forceinline__ void cand64_counters(unsigned long long* ctr, unsigned long long nvis,
                                                unsigned long long npts) {
    unsigned long long* c = ctr + (blockIdx.x % UG_CTR_SLOTS) * UG_CTR_STRIDE;
    unsigned long long pw = npts;
    for (int h = 32; h >= 1; h >>= 1)
        pw += __shfl_xor(pw, h, 64);
    if (threadIdx.x == 0 && nvis) {
        atomicAdd(&c[0], nvis);
        atomicAdd(&c[1], pw);
    }
}

template <int D, bool FULL, int PL, bool SB = false, bool XR = false>
__global__ __launch_bounds__(64) void k_cand64(DevModel m, uint64_t base, int64_t count,
                                               const int64_t* list, const int* dcount,
                                               uint8_t* flags, double* out, int head,
                                               unsigned long long* ctr, int* cells,
                                               uint64_t* mask, const uint32_t* skip) {
    if (dcount)
        count = min(count, (int64_t)*dcount);
    unsigned long long nvis = 0, npts = 0;
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const uint64_t idx = list ? (uint64_t)list[i] : base + (uint64_t)i;
        cand64_item<D, FULL, PL, SB, XR>(m, i, idx, flags, out, head, cells, mask, nvis, npts,
                                         skip ? (int64_t)skip[i] : -1);
    }
    if (!FULL && ctr) {
        // exact counts for the algorithmic-bytes figure
        if (dcount && blockIdx.x == 0 && threadIdx.x == 0)
            atomicAdd(&ctr[(blockIdx.x % UG_CTR_SLOTS) * UG_CTR_STRIDE + 2], (unsigned long long)count);
        cand64_counters(ctr, nvis, npts);
    }
}

// k_cand64 over the candidates of many universes (the batched engine, at most 64 a launch):
// universe u's min(count, *dcount) items as one flattened work queue, one wave per item.
template <int D, bool FULL, int PL, bool SB>
__global__ __launch_bounds__(64) void k_cand64_multi(const MCand* __restrict__ A, int U) {
    const int lane = threadIdx.x;
    int64_t c = 0;
    if (lane < U) {
        c = A[lane].count;
        const int* dc = ug_gptr(&A[lane].dcount);
        if (dc)
            c = min(c, (int64_t)*dc);
    }
    BatchItems q;
    q.scan(c, U);
    unsigned long long nvis = 0, npts = 0;
    for (int64_t item = blockIdx.x; item < q.total; item += gridDim.x) {
        int64_t i = 0;
        const int u = q.owner(item, &i);
        // the universe's record in registers, read once per item, its pointers global
        const MCand* r = A + u;
        const DevModel m = load_model(&r->m);
        const int64_t* list = ug_gptr(&r->list);
        const uint64_t idx = list ? (uint64_t)list[i] : r->base + (uint64_t)i;
        const uint32_t* skip = ug_gptr(&r->skip);
        cand64_item<D, FULL, PL, SB>(m, i, idx, ug_gptr(&r->flags), ug_gptr(&r->out), r->head, ug_gptr(&r->cells),
                                     ug_gptr(&r->mask), nvis, npts, skip ? (int64_t)skip[i] : -1);
    }
}

// Full-size check of the serial-RSA characterization (one wave per item, every timestep
// evaluated and looked up).  mode 0, item p < n: stored unique path p, regenerated from its
// candidate index -- a contact with any other stored unique is a violation (no two accepted
// paths collide).  mode 1, candidate first + i: skipped when accepted; otherwise it must touch
// a unique accepted before it (rows < #accepted with smaller index) -- no contact is a
// violation.  out[0..1] = violations of modes 0/1, out[2..3] = items checked.
template <int D, int PL, bool XR = false>
__global__ __launch_bounds__(64) void k_check(DevModel m, int mode, uint64_t first, int64_t count,
                                              const int64_t* acc_idx, int64_t n,
                                              unsigned long long* out) {
    extern __shared__ __align__(16) int dyn_lds[];
    const int lane = threadIdx.x;
    const int T = m.T;
    constexpr int NWP = 128 * PL;
    int* sb = dyn_lds;
    uint32_t* used = (uint32_t*)(sb + D * m.nw);
    const int lds_head = cand64_lds_head<D>(m, PL);
    double2* cs = (double2*)(dyn_lds + lds_head);
    double* pos = (double*)(cs + D * NWP);
    __shared__ int tsl[64];
    constexpr int R = D == 2 ? 3 : 9;
    constexpr int PER_ROUND = 64 / R;
    unsigned long long nviol = 0, nitems = 0;  // this block's counts (one atomic each at the end)
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        uint64_t idx;
        int64_t self = -1, limit = n;
        if (mode == 0) {
            idx = (uint64_t)acc_idx[i];
            self = i;
        } else {
            idx = first + (uint64_t)i;
            int64_t lo = 0, hi = n;  // first accepted index >= idx
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((uint64_t)acc_idx[mid] < idx)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < n && (uint64_t)acc_idx[lo] == idx)
                continue;  // accepted: mode 0 covers it (uniform across the wave)
            limit = lo;
        }
        WaveCoef<D, PL> w;
        gen_wave64<D, PL, XR>(m, idx, lane, w, sb, used);
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int ii = 0; ii < PL; ii++)
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int j = 2 * (lane + 64 * ii) + k;
                    cs[a * NWP + j] = make_double2(w.c[a][ii][k], w.s[a][ii][k]);
                    sb[a * NWP + j] = w.b[a][ii][k];
                }
        __syncthreads();
        bool hit = false;
        for (int ps = 0; ps < T && !hit; ps += 64) {
            const int oi = ps + lane;
            const int pend = min(64, T - ps);
            if (oi < T) {
                const int t = m.order[oi];
                double X[D];
                // (more slot pairs per lane would spill registers; the reference's stream
                // shuffles the full pool over the slots)
                if (!XR && PL == 1 && m.full)
                    eval_lane_full<D, PL>(m, cs, sb, w.a1, t, X);
                else
                    eval_lane<D, PL>(m, cs, sb, w.a1, t, X);
                tsl[lane] = t;
#pragma unroll
                for (int a = 0; a < D; a++)
                    pos[lane * D + a] = X[a];
            }
            __syncthreads();
            for (int r0 = 0; r0 < pend && !hit; r0 += PER_ROUND) {
                const int G = min(PER_ROUND, pend - r0);
                const int sl = lane / R, rr = lane % R;
                bool h = false;
                if (sl < G) {
                    const int t = tsl[r0 + sl];
                    double x[D];
                    int c[D];
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        x[a] = pos[(r0 + sl) * D + a];
                        c[a] = cell_of(x[a], m.halfT, T);
                    }
                    const int gx = wrap_cell(c[0], (D == 2 ? rr : rr / 3) - 1, T);
                    const int gy = D == 2 ? 0 : wrap_cell(c[1], rr % 3 - 1, T);
                    const int64_t rid = D == 2 ? (int64_t)t * T + gx : ((int64_t)t * T + gx) * T + gy;
                    int tr[3];
                    row_heads3<D>(m, t, gx, gy, rid, c[D - 1], tr);
                    for (int k = 0; k < 3 && !h; k++)
                        for (int nn = tr[k]; nn >= 0 && !h;) {
                            const size_t q = (size_t)nn * T + t;
                            if (nn != self && nn < limit)
                                h = contact<D>(m, x, m.pos64 + q * D);
                            nn = __float_as_int(m.pos32[q].w);
                        }
                }
                hit = __ballot(h) != 0;
            }
            __syncthreads();
        }
        if (mode == 0 ? hit : !hit)
            nviol++;
        nitems++;
        __syncthreads();
    }
    if (lane == 0) {
        if (nviol)
            atomicAdd(&out[mode], nviol);
        if (nitems)
            atomicAdd(&out[2 + mode], nitems);
    }
}

// =====================================================================================
// knot regime (cuda/braid_knot.cu): no grid, every candidate is compared with every resident
// strand.  Strands are stored timestep-major -- packed cells kcell[t][ks] and FP64 positions
// kpos[t][ks][D] -- so the strands of one timestep are contiguous; the FP64 positions are
// read only where the packed cells neighbour (the canonical pair predicate: neighbouring
// cells on every axis and |torus_delta| <= r, r = the phase's exclusion).

template <int D>
__device__ __forceinline__ bool cells_adjacent(int u, int v, int T);

// Any contact between the positions pos[g] (timesteps tsl[g], g < G) and the nk strands:
// lanes take (g, strand) items 64 at a time.
template <int D>
__device__ __forceinline__ bool knot_round(const DevModel& m, const double* pos, const int* tsl, int G, int lane,
                                           const int* kcell, const double* kpos, int64_t ks, int nk) {
    const int items = G * nk;
    for (int it0 = 0; it0 < items; it0 += 64) {
        const int it = it0 + lane;
        bool h = false;
        if (it < items) {
            const int g = it / nk, sidx = it - g * nk;
            const int t = tsl[g];
            double x[D];
            int cx = 0;
#pragma unroll
            for (int a = 0; a < D; a++) {
                x[a] = pos[g * D + a];
                cx |= cell_of(x[a], m.halfT, m.T) << (10 * a);
            }
            const size_t q = (size_t)t * ks + sidx;
            if (cells_adjacent<D>(cx, kcell[q], m.T))
                h = contact<D>(m, x, kpos + q * D);
        }
        if (__ballot(h))
            return true;
    }
    return false;
}

// One wave per candidate, as k_cand64's test (head chunks of 2, 4, ... 64 timesteps across
// the wave, then one lane per timestep) with the brute-force lookup: flags[i] = 1 iff
// candidate i touches none of the nk strands.
template <int D, int PL, bool SB, bool XR = false>
__global__ __launch_bounds__(64) void k_knot_test(DevModel m, uint64_t base, int64_t count, const int64_t* list,
                                                  const int* dcount, uint8_t* flags, int head, const int* kcell,
                                                  const double* kpos, int64_t ks, int nk) {
    if (dcount)
        count = min(count, (int64_t)*dcount);
    extern __shared__ __align__(16) int dyn_lds[];
    const int lane = threadIdx.x;
    const int T = m.T;
    int* sb = dyn_lds;
    uint32_t* used = (uint32_t*)(sb + D * m.nw);
    constexpr int NWP = 128 * PL;
    const int lds_head = cand64_lds_head<D>(m, PL);
    double2* cs = (double2*)(dyn_lds + lds_head);
    double* pos = (double*)(cs + D * NWP);
    __shared__ int tsl[64];
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const uint64_t idx = list ? (uint64_t)list[i] : base + (uint64_t)i;
        WaveCoef<D, PL> w;
        gen_wave64<D, PL, XR>(m, idx, lane, w, sb, used);
        bool hit = false;
        int o0 = 0;
        for (int chunk = 2; !hit && o0 < min(head, T); chunk = min(2 * chunk, 64)) {
            const int G = min(chunk, T - o0);
            for (int g = 0; g < G; g++) {
                double X[D];
                const int t = m.order[o0 + g];
                eval_wave64<D, PL>(m, w, t, X);
                if (lane == g) {
                    tsl[g] = t;
#pragma unroll
                    for (int a = 0; a < D; a++)
                        pos[g * D + a] = X[a];
                }
            }
            __syncthreads();
            hit = knot_round<D>(m, pos, tsl, G, lane, kcell, kpos, ks, nk);
            o0 += G;
            __syncthreads();
        }
        if (!hit && o0 < T) {
#pragma unroll
            for (int a = 0; a < D; a++)
#pragma unroll
                for (int ii = 0; ii < PL; ii++)
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        const int j = 2 * (lane + 64 * ii) + k;
                        cs[a * NWP + j] = make_double2(w.c[a][ii][k], w.s[a][ii][k]);
                        sb[a * NWP + j] = w.b[a][ii][k];
                    }
            __syncthreads();
            for (int ps = o0; ps < T && !hit; ps += 64) {
                const int oi = ps + lane;
                const int pend = min(64, T - ps);
                if (oi < T) {
                    const int t = m.order[oi];
                    double X[D];
                    if constexpr (SB)
                        eval_lane_full<D, PL>(m, cs, sb, w.a1, t, X);
                    else
                        eval_lane<D, PL>(m, cs, sb, w.a1, t, X);
                    tsl[lane] = t;
#pragma unroll
                    for (int a = 0; a < D; a++)
                        pos[lane * D + a] = X[a];
                }
                __syncthreads();
                hit = knot_round<D>(m, pos, tsl, pend, lane, kcell, kpos, ks, nk);
                __syncthreads();
            }
        }
        if (lane == 0)
            flags[i] = hit ? 0 : 1;
        __syncthreads();
    }
}

// Appends a resolve chunk's accepted survivors to the strands (one thread per (strand,
// timestep)): trajectory rows of pos (k_cand64 FULL order [s][t][D]) and packed cells go to
// timestep-major rows n0 + k; t == 0 also records the candidate index.
template <int D>
__global__ __launch_bounds__(256) void k_knot_append(DevModel m, const double* pos, const int* cells,
                                                     const int* acc_list, const int* nacc, int limit,
                                                     const int64_t* surv_idx, int64_t n0, int* kcell, double* kpos,
                                                     float4* kpos32, int64_t ks, int64_t* kidx) {
    const int64_t item = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int na = min(*nacc, limit);
    const int k = (int)(item / m.T), t = (int)(item % m.T);
    if (k >= na)
        return;
    const int sidx = acc_list[k];
    const size_t src = (size_t)sidx * m.T + t, dst = (size_t)t * ks + (size_t)(n0 + k);
    float f[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; a++) {
        kpos[dst * D + a] = pos[src * D + a];
        f[a] = (float)pos[src * D + a];
    }
    kpos32[dst] = make_float4(f[0], f[1], f[2], 0.f);
    kcell[dst] = cells[src];
    if (t == 0)
        kidx[n0 + k] = surv_idx[sidx];
}

// Baker coefficients of stored path p (one wave per path): the canonical per-slot
// (a cos f, a sin f) of its candidate, scattered by frequency into dense rows.
template <int D, int PL, bool XR = false>
__global__ __launch_bounds__(64) void k_paths_from_pack(DevModel m, const int64_t* acc_idx, int64_t n,
                                                        double* a2, double* A, double* B, int kmax,
                                                        int* bmax) {
    extern __shared__ __align__(16) int dyn_lds[];
    const int lane = threadIdx.x;
    int* sb = dyn_lds;
    uint32_t* used = (uint32_t*)(sb + D * m.nw);
    const size_t W = (size_t)kmax + 1;
    int bm = 0;
    for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
        WaveCoef<D, PL> w;
        gen_wave64<D, PL, XR>(m, (uint64_t)acc_idx[p], lane, w, sb, used);
#pragma unroll
        for (int a = 0; a < D; a++) {
            if (lane == 0)
                a2[(size_t)p * D + a] = w.a1[a];
#pragma unroll
            for (int ii = 0; ii < PL; ii++)
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int j = 2 * (lane + 64 * ii) + k;
                    if (j < m.nw) {
                        const int b = w.b[a][ii][k];
                        A[((size_t)p * D + a) * W + b] = w.c[a][ii][k];
                        B[((size_t)p * D + a) * W + b] = w.s[a][ii][k];
                        bm = max(bm, b);
                    }
                }
        }
        __syncthreads();
    }
    for (int s = 32; s >= 1; s >>= 1)
        bm = max(bm, __shfl_xor(bm, s, 64));
    if (lane == 0)
        atomicMax(bmax, bm);
}

// Subpath test (cuda/braid_cuda.cu:452-465 sub branch, host_sub_gid :726-750): every
// timestep is scanned; the contacted groups are reduced across lanes with min/max, and the
// candidate fails as soon as two distinct groups appear.  Paths stored at rows < nlimit are
// the packing at the start of the current round.  flags[i] = 1 iff the candidate touches
// exactly one group in the table and at least one of its contacts is a round-start path
// (the reference's device test sees only the round-start grid and drops candidates that
// touch nothing there, :463-464); gid_out[i] = the touched group.  When rs_count is given,
// it counts the candidates the round-start grid alone lets through (the reference's
// survivor count, which drives its batch size).
template <int D, int PL, bool XR = false>
__global__ __launch_bounds__(64) void k_sub_test(DevModel m, uint64_t base, int64_t count,
                                                 const int64_t* list, const int* dcount,
                                                 const int32_t* acc_gid, int64_t nlimit, int* rs_count,
                                                 uint8_t* flags, int* gid_out) {
    if (dcount)
        count = min(count, (int64_t)*dcount);
    extern __shared__ __align__(16) int dyn_lds[];
    const int lane = threadIdx.x;
    const int T = m.T, nw = m.nw;
    int* sb = dyn_lds;
    uint32_t* used = (uint32_t*)(sb + D * nw);
    constexpr int NWP = 128 * PL;  // padded slots per axis
    const int lds_head = cand64_lds_head<D>(m, PL);
    double2* cs = (double2*)(dyn_lds + lds_head);
    double* pos = (double*)(cs + D * NWP);
    __shared__ int tsl[64];
    constexpr int R = D == 2 ? 3 : 9;
    constexpr int PER_ROUND = 64 / R;
    for (int64_t i = blockIdx.x; i < count; i += gridDim.x) {
        const uint64_t idx = list ? (uint64_t)list[i] : base + (uint64_t)i;
        WaveCoef<D, PL> w;
        gen_wave64<D, PL, XR>(m, idx, lane, w, sb, used);
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int ii = 0; ii < PL; ii++)
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int j = 2 * (lane + 64 * ii) + k;
                    cs[a * NWP + j] = make_double2(w.c[a][ii][k], w.s[a][ii][k]);
                    sb[a * NWP + j] = w.b[a][ii][k];
                }
        __syncthreads();
        // all contacts: (gcur, bad); contacts with round-start paths: (gcur0, bad0)
        int gcur = -1, gcur0 = -1;
        bool bad = false, bad0 = false;
        for (int ps = 0; ps < T && !bad0 && !(bad && !rs_count); ps += 64) {
            const int oi = ps + lane;
            const int pend = min(64, T - ps);
            if (oi < T) {
                const int t = m.order[oi];
                double X[D];
                // (more slot pairs per lane would spill registers; the reference's stream
                // shuffles the full pool over the slots)
                if (!XR && PL == 1 && m.full)
                    eval_lane_full<D, PL>(m, cs, sb, w.a1, t, X);
                else
                    eval_lane<D, PL>(m, cs, sb, w.a1, t, X);
                tsl[lane] = t;
#pragma unroll
                for (int a = 0; a < D; a++)
                    pos[lane * D + a] = X[a];
            }
            __syncthreads();
            for (int r0 = 0; r0 < pend && !bad0 && !(bad && !rs_count); r0 += PER_ROUND) {
                const int G = min(PER_ROUND, pend - r0);
                const int sl = lane / R, rr = lane % R;
                int gmin = INT32_MAX, gmax = -1, gmin0 = INT32_MAX, gmax0 = -1;
                if (sl < G) {
                    const int t = tsl[r0 + sl];
                    double x[D];
                    int c[D];
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        x[a] = pos[(r0 + sl) * D + a];
                        c[a] = cell_of(x[a], m.halfT, T);
                    }
                    const int gx = wrap_cell(c[0], (D == 2 ? rr : rr / 3) - 1, T);
                    const int gy = D == 2 ? 0 : wrap_cell(c[1], rr % 3 - 1, T);
                    const int64_t rid = D == 2 ? (int64_t)t * T + gx : ((int64_t)t * T + gx) * T + gy;
                    int tr[3];
                    row_heads3<D>(m, t, gx, gy, rid, c[D - 1], tr);
                    for (int k = 0; k < 3; k++)
                        for (int n = tr[k]; n >= 0;) {
                            const size_t q = (size_t)n * T + t;
                            const double* P = m.pos64 + q * D;
                            const bool hh = contact<D>(m, x, P);
                            if (hh) {
                                const int g = acc_gid[n];
                                gmin = min(gmin, g);
                                gmax = max(gmax, g);
                                if (n < nlimit) {
                                    gmin0 = min(gmin0, g);
                                    gmax0 = max(gmax0, g);
                                }
                            }
                            n = __float_as_int(m.pos32[q].w);
                        }
                }
                for (int s = 32; s >= 1; s >>= 1) {
                    gmin = min(gmin, __shfl_xor(gmin, s, 64));
                    gmax = max(gmax, __shfl_xor(gmax, s, 64));
                    gmin0 = min(gmin0, __shfl_xor(gmin0, s, 64));
                    gmax0 = max(gmax0, __shfl_xor(gmax0, s, 64));
                }
                if (gmax >= 0) {
                    if (gmin != gmax || (gcur >= 0 && gcur != gmin))
                        bad = true;
                    else
                        gcur = gmin;
                }
                if (gmax0 >= 0) {
                    if (gmin0 != gmax0 || (gcur0 >= 0 && gcur0 != gmin0))
                        bad0 = true;
                    else
                        gcur0 = gmin0;
                }
            }
            __syncthreads();
        }
        if (lane == 0) {
            flags[i] = (bad || gcur0 < 0) ? 0 : 1;
            gid_out[i] = gcur;
            if (rs_count && !bad0 && gcur0 >= 0)
                atomicAdd(rs_count, 1);
        }
        __syncthreads();
    }
}

// Sequential adoption over the chunk's survivors in index order (one wave): survivor j's
// groups are its table group (gin[j], -1 if none) plus the groups of earlier accepted
// survivors it collides with (conflict bits); it is accepted iff exactly one group results,
// and carries that group (cuda/braid_cuda.cu:984-999).
__global__ __launch_bounds__(64) void k_greedy_sub(const uint64_t* mask, int S, const int* gin,
                                                   int* gsurv, int* acc_list, int* nacc) {
    __shared__ uint64_t accw[UG_MASK_WORDS];
    __shared__ int gacc[UG_S_MAX];
    const int lane = threadIdx.x;
    if (lane < UG_MASK_WORDS)
        accw[lane] = 0ull;
    __syncthreads();
    int na = 0;
    for (int j = 0; j < S; j++) {
        int gmin = INT32_MAX, gmax = -1;
        const int nwd = (j + 63) / 64;
        if (lane < nwd) {
            uint64_t bits = mask[(size_t)j * UG_MASK_WORDS + lane] & accw[lane];
            while (bits) {
                const int b = __ffsll((long long)bits) - 1;
                bits &= bits - 1;
                const int g = gacc[lane * 64 + b];
                gmin = min(gmin, g);
                gmax = max(gmax, g);
            }
        }
        for (int s = 32; s >= 1; s >>= 1) {
            gmin = min(gmin, __shfl_xor(gmin, s, 64));
            gmax = max(gmax, __shfl_xor(gmax, s, 64));
        }
        const int g0 = gin[j];
        if (g0 >= 0) {
            gmin = min(gmin, g0);
            gmax = max(gmax, g0);
        }
        const bool ok = gmax >= 0 && gmin == gmax;
        if (lane == 0) {
            gacc[j] = ok ? gmax : -1;
            gsurv[j] = ok ? gmax : -1;
            if (ok) {
                accw[j >> 6] |= 1ull << (j & 63);
                acc_list[na] = j;
            }
        }
        na += ok ? 1 : 0;
        __syncthreads();
    }
    if (lane == 0)
        *nacc = na;
}

// =====================================================================================
// ordered compaction: out[k] = candidate index of the k-th flagged item (stable)
constexpr int CMP_ITEMS = 4096;  // items per block (256 threads x 16)

__device__ __forceinline__ void compact_count_block(const uint8_t* flags, int64_t count, const int* dcount,
                                                    int* bcount, int bx) {
    __shared__ int red[256];
    if (dcount)
        count = min(count, (int64_t)*dcount);
    const int64_t b0 = (int64_t)bx * CMP_ITEMS;
    int n = 0;
    for (int k = threadIdx.x; k < CMP_ITEMS; k += 256) {
        const int64_t i = b0 + k;
        if (i < count)
            n += flags[i];
    }
    red[threadIdx.x] = n;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        bcount[bx] = red[0];
}

__global__ __launch_bounds__(256) void k_compact_count(const uint8_t* flags, int64_t count,
                                                       const int* dcount, int* bcount) {
    compact_count_block(flags, count, dcount, bcount, blockIdx.x);
}

__device__ __forceinline__ void compact_scatter_block(const uint8_t* flags, int64_t count, const int* dcount,
                                                      uint64_t base, const int64_t* list, const int* bcount,
                                                      int nblocks, int64_t* out, int* total, const int* vin,
                                                      int* vout, int bx) {
    __shared__ int red[256];
    if (dcount)
        count = min(count, (int64_t)*dcount);
    __shared__ int scan[256];
    // offset of this block = sum of earlier block counts
    int s0 = 0;
    for (int b = threadIdx.x; b < bx; b += 256)
        s0 += bcount[b];
    red[threadIdx.x] = s0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const int boff = red[0];
    // each thread owns 16 consecutive items
    const int64_t i0 = (int64_t)bx * CMP_ITEMS + (int64_t)threadIdx.x * 16;
    int n = 0;
    for (int k = 0; k < 16; k++)
        if (i0 + k < count)
            n += flags[i0 + k];
    scan[threadIdx.x] = n;
    __syncthreads();
    for (int s = 1; s < 256; s <<= 1) {  // inclusive Hillis-Steele scan
        const int v = threadIdx.x >= s ? scan[threadIdx.x - s] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    int o = boff + scan[threadIdx.x] - n;
    for (int k = 0; k < 16; k++) {
        const int64_t i = i0 + k;
        if (i < count && flags[i]) {
            if (vin)
                vout[o] = vin[i];
            out[o++] = list ? list[i] : (int64_t)(base + (uint64_t)i);
        }
    }
    if (bx == nblocks - 1 && threadIdx.x == 255)
        *total = boff + scan[255];
}

__global__ __launch_bounds__(256) void k_compact_scatter(const uint8_t* flags, int64_t count,
                                                         const int* dcount, uint64_t base,
                                                         const int64_t* list, const int* bcount,
                                                         int nblocks, int64_t* out, int* total,
                                                         const int* vin, int* vout) {
    compact_scatter_block(flags, count, dcount, base, list, bcount, nblocks, out, total, vin, vout, blockIdx.x);
}

// Batched compaction: blockIdx.y = universe, blockIdx.x = its block (ceil(count / 4096) of
// them, at least one, so a universe with nothing to compact still gets *total = 0).
__global__ __launch_bounds__(256) void k_compact_count_multi(const MCompact* A) {
    const MCompact* r = A + blockIdx.y;
    const int64_t count = r->count;
    const int nb = (int)max((int64_t)1, (count + CMP_ITEMS - 1) / CMP_ITEMS);
    if ((int)blockIdx.x < nb)
        compact_count_block(ug_gptr(&r->flags), count, ug_gptr(&r->dcount), ug_gptr(&r->bcount), blockIdx.x);
}
__global__ __launch_bounds__(256) void k_compact_scatter_multi(const MCompact* A) {
    const MCompact* r = A + blockIdx.y;
    const int64_t count = r->count;
    const int nb = (int)max((int64_t)1, (count + CMP_ITEMS - 1) / CMP_ITEMS);
    if ((int)blockIdx.x < nb)
        compact_scatter_block(ug_gptr(&r->flags), count, ug_gptr(&r->dcount), r->base, ug_gptr(&r->list),
                              ug_gptr(&r->bcount), nb, ug_gptr(&r->out), ug_gptr(&r->total), ug_gptr(&r->vin),
                              ug_gptr(&r->vout), blockIdx.x);
}

// =====================================================================================
// in-batch accept

// Neighbouring cells on every axis (torus), from packed cells (10 bits per axis, written
// with a resolve chunk's trajectories).
template <int D>
__device__ __forceinline__ bool cells_adjacent(int u, int v, int T) {
    bool adj = true;
#pragma unroll
    for (int a = 0; a < D; a++) {
        int dc = ((u >> (10 * a)) & 1023) - ((v >> (10 * a)) & 1023);
        dc = dc < 0 ? -dc : dc;
        adj = adj && (dc <= 1 || dc == T - 1);
    }
    return adj;
}

// Conflict bits among a resolve chunk's first S survivors: mask[i][j/64] bit j%64 set iff
// i > j collide at some timestep -- their cells neighbour on every axis (torus) and the
// canonical FP64 contact predicate holds, what a cell-table lookup would find.  Workgroup
// (t, g) hashes the whole chunk by packed cell at timestep t in LDS (open addressing, 2048
// slots for at most 1024 survivors, a chain per cell, positions beside), then survivor
// CONF_WG g + thread probes its 3^D neighbouring cells (wrapped) and tests the earlier survivors
// found there: O(S 3^D) per timestep instead of O(S^2) pair tests (a pair-per-thread kernel
// and a 64x64-tile kernel over all pairs took 9-45 us where this takes a few).
constexpr int CONF_HBITS = 12;
constexpr int CONF_HSLOTS = 1 << CONF_HBITS;  // at least twice the survivors of a chunk
static_assert(CONF_HSLOTS >= 2 * UG_S_MAX, "conflict hash load above one half");
constexpr size_t conf_lds_bytes(int D) { return (size_t)CONF_HSLOTS * 8 + (size_t)UG_S_MAX * (2 + 8 * D); }
static_assert((CONF_HSLOTS * 8 + UG_S_MAX * 2) % 8 == 0, "positions after the chains stay aligned");
constexpr int CONF_WG = 1024;  // threads per workgroup: one survivor each, one group per timestep

__device__ __forceinline__ int conf_hslot(int key) {
    return (int)(((uint32_t)key * 0x9E3779B1u) >> (32 - CONF_HBITS));
}

template <int D>
__device__ __forceinline__ void conflicts_cells_body(const DevModel& m, const double* pos, const int* cells, int S,
                                                     uint64_t* mask, int t, int g) {
    // (dynamic LDS, conf_lds_bytes: the hash, the chains, the positions)
    extern __shared__ __align__(16) int dyn_lds[];
    int* hkey = dyn_lds;
    int* hhead = hkey + CONF_HSLOTS;
    short* nxt = (short*)(hhead + CONF_HSLOTS);
    double* sp = (double*)(nxt + UG_S_MAX);
    // workgroup (t, g): the whole chunk hashed at timestep t, probes for survivors CONF_WG g ..
    const int T = m.T, s0 = g * CONF_WG;
    if (t >= T || s0 >= S)
        return;  // uniform over the workgroup
    for (int k = threadIdx.x; k < CONF_HSLOTS; k += CONF_WG) {
        hkey[k] = -1;
        hhead[k] = -1;
    }
    __syncthreads();
    for (int s = threadIdx.x; s < S; s += CONF_WG) {
        const int key = cells[(size_t)s * T + t];
#pragma unroll
        for (int a = 0; a < D; a++)
            sp[s * D + a] = pos[((size_t)s * T + t) * D + a];
        int h = conf_hslot(key);
        for (;;) {
            const int prev = atomicCAS(&hkey[h], -1, key);
            if (prev == -1 || prev == key)
                break;
            h = (h + 1) & (CONF_HSLOTS - 1);
        }
        nxt[s] = (short)atomicExch(&hhead[h], s);
    }
    __syncthreads();
    const int s = s0 + (int)threadIdx.x;
    if (s >= S)
        return;
    const int key = cells[(size_t)s * T + t];
    int c[D];
#pragma unroll
    for (int a = 0; a < D; a++)
        c[a] = (key >> (10 * a)) & 1023;
    double x[D];
#pragma unroll
    for (int a = 0; a < D; a++)
        x[a] = sp[s * D + a];
    constexpr int NB = D == 2 ? 9 : 27;
    for (int nb = 0; nb < NB; nb++) {
        int nk = 0, q = nb;
#pragma unroll
        for (int a = 0; a < D; a++) {
            int v = c[a] + q % 3 - 1;
            q /= 3;
            v = v < 0 ? v + T : (v >= T ? v - T : v);
            nk |= v << (10 * a);
        }
        int h = conf_hslot(nk);
        int kk;
        while ((kk = hkey[h]) != -1 && kk != nk)
            h = (h + 1) & (CONF_HSLOTS - 1);
        if (kk != nk)
            continue;
        for (int j = hhead[h]; j >= 0; j = nxt[j])
            if (j < s && contact<D>(m, x, &sp[j * D]))
                atomicOr((unsigned long long*)&mask[(size_t)s * UG_MASK_WORDS + (j >> 6)], 1ull << (j & 63));
    }
}

template <int D>
__global__ __launch_bounds__(CONF_WG) void k_conflicts_cells(DevModel m, const double* pos, const int* cells, int S,
                                                         const int* dS, uint64_t* mask) {
    if (dS)
        S = min(S, *dS);
    conflicts_cells_body<D>(m, pos, cells, S, mask, blockIdx.x, blockIdx.y);
}

// A resolve chunk's results, stored straight into the pinned staging block (one dispatch in
// place of four copies): [8] accepted count, [16] occupied (t, cell) buckets, [32] the
// filter's four counts when they were still on the device, [64 ..] the accepted candidate
// indices (at most s).  The host reads the block after synchronising the stream.
__device__ __forceinline__ void report_block(const int* nacc, const unsigned long long* filled,
                                             const int64_t* acc_idx, int s, const int* counts, char* pin,
                                             int* spec_count, int spec_n, int64_t stop_lim, int64_t n_room) {
    const int na = min(*nacc, s);
    if (threadIdx.x == 0) {
        *(int*)(pin + 8) = *nacc;
        *(unsigned long long*)(pin + 16) = *filled;
        if (spec_count) {
            // the count of the filter run ahead: none when the run stops after this sub-batch
            const int64_t a = *nacc;
            const int v = (a < stop_lim || a >= n_room) ? 0 : spec_n;
            *spec_count = v;
            *(int*)(pin + 48) = v;
        }
    }
    if (counts && threadIdx.x < 4)
        ((int*)(pin + 32))[threadIdx.x] = counts[threadIdx.x];
    int64_t* out = (int64_t*)(pin + 64);
    for (int k = threadIdx.x; k < na; k += 256)
        out[k] = acc_idx[k];
}

// Greedy in index order over the conflict bits (one wave): survivor i is accepted iff it
// conflicts with no earlier accepted survivor -- serial RSA restricted to the chunk.  Lane
// i of block blk needs words 0..blk of its mask row; they are loaded together, and the next
// two blocks' while this one is decided, so the pass pays about one load latency in all.
__device__ __forceinline__ void greedy_body(const uint64_t* mask, int S, int* acc_list, int* nacc) {
    __shared__ uint64_t accw[UG_MASK_WORDS];
    const int lane = threadIdx.x;
    uint64_t cur[UG_MASK_WORDS], nxt[UG_MASK_WORDS], nx2[UG_MASK_WORDS];
    auto load = [&](uint64_t* w, int blk) {
        const int i = blk * 64 + lane;
#pragma unroll
        for (int wd = 0; wd < UG_MASK_WORDS; wd++)
            w[wd] = (i < S && wd <= blk) ? mask[(size_t)i * UG_MASK_WORDS + wd] : 0ull;
    };
    load(cur, 0);
    if (64 < S)
        load(nxt, 1);
    int na = 0;
    for (int blk = 0; blk * 64 < S; blk++) {
        if ((blk + 2) * 64 < S)
            load(nx2, blk + 2);
        const int i = blk * 64 + lane;
        const bool valid = i < S;
        uint64_t killed = 0ull, inrow = 0ull;
#pragma unroll
        for (int wd = 0; wd < UG_MASK_WORDS; wd++) {
            if (wd < blk)
                killed |= cur[wd] & accw[wd];
            if (wd == blk)
                inrow = cur[wd];
        }
        const uint64_t deadm = __ballot(!valid || killed != 0ull);
        // Index-order greedy inside the block, decided in waves of lanes: a lane is accepted
        // once every earlier lane it conflicts with is decided and none of them accepted, and
        // dropped once one of them is accepted.  The lowest undecided lane is always decidable,
        // so the loop ends; it takes as many steps as the longest chain of conflicts (usually
        // a few) instead of 64.
        uint64_t accb = 0, undec = ~deadm;
        const uint64_t earlier = inrow & ((1ull << lane) - 1ull);  // (bits >= lane are zero)
        while (undec) {
            const bool mine = (undec >> lane) & 1ull;
            const bool drop = mine && (earlier & accb) != 0ull;
            const bool take = mine && !drop && (earlier & undec) == 0ull;
            accb |= __ballot(take);
            undec &= ~(__ballot(take) | __ballot(drop));
        }
        if (lane == 0)
            accw[blk] = accb;
        __syncthreads();
        if (valid && ((accb >> lane) & 1ull)) {
            const int pos = na + __popcll(accb & ((1ull << lane) - 1ull));
            acc_list[pos] = i;
        }
        na += __popcll(accb);
#pragma unroll
        for (int wd = 0; wd < UG_MASK_WORDS; wd++) {
            cur[wd] = nxt[wd];
            nxt[wd] = nx2[wd];
        }
    }
    if (lane == 0)
        *nacc = na;
}

__global__ __launch_bounds__(64) void k_greedy(const uint64_t* mask, int S, const int* dS, int* acc_list,
                                               int* nacc) {
    if (dS)
        S = min(S, *dS);
    greedy_body(mask, S, acc_list, nacc);
}

// Sparse grid: path n becomes the head of the cell `key` (the slot claimed by compare-and-swap
// when the cell was empty, *fresh set then); returns the previous head, -1 if none.  Slots of an
// empty table hold {0, 0xffffffff}; a table is never more than half full.
__device__ __forceinline__ int hash_insert(const DevModel& m, uint64_t key, int n, bool* fresh) {
    uint64_t h = hmix64(key) & m.hmask;
    for (;;) {
        const unsigned long long old = atomicCAS(&m.hslots[h].x, 0ull, (unsigned long long)(key + 1));
        if (old == 0ull || old == key + 1) {
            *fresh = old == 0ull;
            return atomicExch((int*)&m.hslots[h].y, n);
        }
        h = (h + 1) & m.hmask;
    }
}

__global__ void k_hash_clear(ulonglong2* slots, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        slots[i] = make_ulonglong2(0ull, 0xffffffffull);
}

// Rebuild of the sparse grid into a larger (cleared) table: every stored (path, timestep)
// is inserted again from its canonical position and its chain link rewritten.
template <int D>
// Grid-stride: a table of 3+1 T=256 outgrows one item per thread of a capped grid (65,535
// blocks of 256 cover 16.7M entries, about 65k paths).
__global__ __launch_bounds__(256) void k_hash_rebuild(DevModel m, int64_t rows, float4* pos32) {
    for (int64_t item = (int64_t)blockIdx.x * 256 + threadIdx.x; item < rows * m.T;
         item += (int64_t)gridDim.x * 256) {
        const int64_t n = item / m.T;
        const int t = (int)(item % m.T);
        const size_t q = (size_t)n * m.T + t;
        int c[D];
#pragma unroll
        for (int a = 0; a < D; a++)
            c[a] = cell_of(m.pos64[q * D + a], m.halfT, m.T);
        int64_t rid = (int64_t)t * m.T + c[0];
        if (D == 3)
            rid = rid * m.T + c[1];
        bool fresh = false;
        const int prev = hash_insert(m, cell_hkey(rid, c[D - 1], m.T), (int)n, &fresh);
        float4 r = pos32[q];
        r.w = __int_as_float(prev);
        pos32[q] = r;
    }
}

// Insert the chunk's accepted survivors into the cell table, one thread per (path, timestep):
// path n becomes the head of its (t, cell) chain and links to the previous head
// (cuda/braid_cuda.cu:865-879).  Accepted paths of one chunk never collide, so each cell
// receives at most one of them except at rounding ties, where the atomic exchange keeps the
// chain intact.  Row-edge cells are mirrored into the halo cells afterwards (k_halo).
template <int D>
__device__ __forceinline__ void insert_item(const DevModel& m, const double* pos, const int* acc_list,
                                            const int* nacc, int limit, const int64_t* surv_idx,
                                            const int* surv_gid, int64_t n0, int64_t* acc_idx,
                                            int32_t* acc_gid, double* pos64, float4* pos32,
                                            unsigned long long* filled, int64_t item) {
    const int na = min(*nacc, limit);
    const int k = (int)(item / m.T), t = (int)(item % m.T);
    bool fresh = false;
    if (k < na) {
        const int s = acc_list[k];
        const int64_t n = n0 + k;
        const double* X = pos + ((size_t)s * m.T + t) * D;
        int c[D];
        float f[3] = {0.f, 0.f, 0.f};
        const size_t o = (size_t)n * m.T + t;
#pragma unroll
        for (int a = 0; a < D; a++) {
            c[a] = cell_of(X[a], m.halfT, m.T);
            pos64[o * D + a] = X[a];
            f[a] = (float)X[a];
        }
        int64_t rid = (int64_t)t * m.T + c[0];
        if (D == 3)
            rid = rid * m.T + c[1];
        const int cl = c[D - 1];
        int prev;
        if (m.sparse) {
            prev = hash_insert(m, cell_hkey(rid, cl, m.T), (int)n, &fresh);
        } else {
            prev = atomicExch(&m.tab[rid * m.rowlen + cl + 1], (int)n);
            fresh = prev < 0;
        }
        // occupancy bit of the padded cell, and of its halo copy at the row ends
        const int oy = D == 3 ? c[1] : 0;
        atomicOr(&m.occ[occ_index<D>(m, t, c[0], oy, (cl + 1) >> 5)], 1u << ((cl + 1) & 31));
        if (cl == 0)
            atomicOr(&m.occ[occ_index<D>(m, t, c[0], oy, (m.T + 1) >> 5)], 1u << ((m.T + 1) & 31));
        if (cl == m.T - 1)
            atomicOr(&m.occ[occ_index<D>(m, t, c[0], oy, 0)], 1u);
        pos32[o] = make_float4(f[0], f[1], f[2], __int_as_float(prev));
        if (t == 0) {
            acc_idx[n] = surv_idx[s];
            acc_gid[n] = surv_gid ? surv_gid[s] : (int32_t)n;  // a unique's group is its admission index
        }
    }
    const uint64_t fm = __ballot(fresh);
    if ((threadIdx.x & 63) == 0 && fm)
        atomicAdd(filled, (unsigned long long)__popcll(fm));
}

template <int D>
__global__ __launch_bounds__(256) void k_insert(DevModel m, const double* pos, const int* acc_list,
                                                const int* nacc, int limit, const int64_t* surv_idx,
                                                const int* surv_gid, int64_t n0, int64_t* acc_idx,
                                                int32_t* acc_gid, double* pos64, float4* pos32,
                                                unsigned long long* filled) {
    insert_item<D>(m, pos, acc_list, nacc, limit, surv_idx, surv_gid, n0, acc_idx, acc_gid, pos64, pos32, filled,
                   (int64_t)blockIdx.x * 256 + threadIdx.x);
}

template <int D>
__device__ __forceinline__ void halo_item(const DevModel& m, const double* pos, const int* acc_list,
                                          const int* nacc, int limit, int64_t item) {
    const int na = min(*nacc, limit);
    const int k = (int)(item / m.T), t = (int)(item % m.T);
    if (k >= na || m.sparse)  // the sparse grid has no halo cells
        return;
    const double* X = pos + ((size_t)acc_list[k] * m.T + t) * D;
    int c[D];
#pragma unroll
    for (int a = 0; a < D; a++)
        c[a] = cell_of(X[a], m.halfT, m.T);
    int64_t row = (int64_t)t * m.T + c[0];
    if (D == 3)
        row = row * m.T + c[1];
    row *= m.rowlen;
    const int cl = c[D - 1];
    if (cl == 0)
        m.tab[row + m.T + 1] = m.tab[row + 1];
    if (cl == m.T - 1)
        m.tab[row] = m.tab[row + m.T];
}

template <int D>
__global__ __launch_bounds__(256) void k_halo(DevModel m, const double* pos, const int* acc_list,
                                              const int* nacc, int limit) {
    halo_item<D>(m, pos, acc_list, nacc, limit, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

// The halo copies in blocks [0, gridDim.x - 1) and the chunk's report in the last block: the
// report reads only what k_insert and the filter wrote, never the halo cells, so it needs no
// order against the copies and the chunk ends with one dispatch instead of two.
template <int D>
__global__ __launch_bounds__(256) void k_halo_report(DevModel m, const double* pos, const int* acc_list,
                                                     const int* nacc, int limit, const unsigned long long* filled,
                                                     const int64_t* acc_idx, int s, const int* counts, char* pin,
                                                     int* spec_count, int spec_n, int64_t stop_lim,
                                                     int64_t n_room) {
    if (blockIdx.x == gridDim.x - 1)
        report_block(nacc, filled, acc_idx, s, counts, pin, spec_count, spec_n, stop_lim, n_room);
    else
        halo_item<D>(m, pos, acc_list, nacc, limit, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

// ------------------------------------------------------------------ batched resolve
// A resolve chunk of every universe of a batched step (MResolve; blockIdx.y = universe for
// the per-item kernels, blockIdx.x = universe for the per-universe ones).  The same bodies as
// the single-universe kernels, so each universe gets the bits it would get alone.
// the chunk's survivors: at most S, and the filter's count when it is on the device
__device__ __forceinline__ int resolve_count(const MResolve& a) {
    return a.counts ? min(a.S, a.counts[1]) : a.S;
}
// A universe's resolve record, its pointers global (see ug_gptr).
__device__ __forceinline__ MResolve resolve_rec(const MResolve* A, int u) {
    const MResolve* r = A + u;
    MResolve a = *r;
    a.m = load_model(&r->m);
    a.pos = ug_gptr(&r->pos);
    a.cells = ug_gptr(&r->cells);
    a.mask = ug_gptr(&r->mask);
    a.counts = ug_gptr(&r->counts);
    a.acc_list = ug_gptr(&r->acc_list);
    a.nacc = ug_gptr(&r->nacc);
    a.surv_idx = ug_gptr(&r->surv_idx);
    a.acc_idx = ug_gptr(&r->acc_idx);
    a.acc_gid = ug_gptr(&r->acc_gid);
    a.pos64 = ug_gptr(&r->pos64);
    a.pos32 = ug_gptr(&r->pos32);
    a.filled = ug_gptr(&r->filled);
    a.rep = ug_gptr(&r->rep);
    return a;
}
template <int D>
__global__ __launch_bounds__(CONF_WG) void k_conflicts_multi(const MResolve* A) {
    const MResolve a = resolve_rec(A, blockIdx.z);
    const int S = resolve_count(a);
    if (S > 1)
        conflicts_cells_body<D>(a.m, a.pos, a.cells, S, a.mask, blockIdx.x, blockIdx.y);
}
__global__ __launch_bounds__(64) void k_greedy_multi(const MResolve* A) {
    const MResolve a = resolve_rec(A, blockIdx.x);
    greedy_body(a.mask, resolve_count(a), a.acc_list, a.nacc);
}
template <int D>
__global__ __launch_bounds__(256) void k_insert_multi(const MResolve* A) {
    const MResolve a = resolve_rec(A, blockIdx.y);
    if ((int64_t)blockIdx.x * 256 >= (int64_t)a.S * a.m.T)
        return;  // the whole workgroup is past this universe's items
    insert_item<D>(a.m, a.pos, a.acc_list, a.nacc, a.limit, a.surv_idx, nullptr, a.n0, a.acc_idx, a.acc_gid,
                   a.pos64, a.pos32, a.filled, (int64_t)blockIdx.x * 256 + threadIdx.x);
}
template <int D>
__global__ __launch_bounds__(256) void k_halo_multi(const MResolve* A) {
    const MResolve a = resolve_rec(A, blockIdx.y);
    const int64_t item = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (item < (int64_t)a.S * a.m.T)
        halo_item<D>(a.m, a.pos, a.acc_list, a.nacc, a.limit, item);
}
// rep[0] = accepted (at most limit), rep[1] = the filter's survivors, rep[2..3] = occupied
// buckets, rep[4] = the filter's FP64 rechecks, rep + 6: the accepted candidate indices
// (int64) -- written into pinned host memory, read after the stream is synchronised.
__global__ __launch_bounds__(256) void k_report_multi(const MResolve* A) {
    const MResolve a = resolve_rec(A, blockIdx.x);
    const int na = min(*a.nacc, a.limit);
    if (threadIdx.x == 0) {
        a.rep[0] = na;
        a.rep[1] = a.counts ? a.counts[1] : a.S;
        *(unsigned long long*)(a.rep + 2) = *a.filled;
        a.rep[4] = a.counts ? a.counts[0] : 0;
    }
    int64_t* out = (int64_t*)(a.rep + 6);
    for (int k = threadIdx.x; k < na; k += 256)
        out[k] = a.acc_idx[a.n0 + k];
}

__global__ void k_fill_i32(int* p, int64_t n, int v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = v;
}
// =====================================================================================
// host driver
namespace ug {

int grid_for(int64_t items, int per_block) {
    int64_t b = (items + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 65535));
}

int32_t launch_cand64(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags,
                      double* out, bool full, unsigned long long* ctr, const int* dcount, int head,
                      int* cells, uint64_t* mask, const uint32_t* skip) {
    if (count <= 0)
        return UG_OK;
    if (head < 0)
        head = list ? (h->d == 2 ? 14 : 13) : h->T;
    const int blocks = (int)std::min<int64_t>(count, 256 * 32);
    const size_t lds = (size_t)4 * (h->d == 2 ? cand64_lds_words<2>(h->dm) : cand64_lds_words<3>(h->dm));
    // slot pairs per lane: 1 up to 64 pairs (T <= 129), 2 up to 128, else 4
    const int pl = slot_pl(h->npairs);
    // the reference's stream (rng_kind UG_RNG_XOSHIRO_REF): its own instantiations, one lane
    // per timestep without the shared basis loads
#define UG_CAND64(DD, FF, PP, SS)                                                                              \
    do {                                                                                                       \
        if (h->dm.xref)                                                                                        \
            UG_LAUNCH(h, (k_cand64<DD, FF, PP, false, true>), dim3(blocks), dim3(64), lds, h->dm, base, count, \
                      list, dcount, flags, out, head, ctr, cells, mask, skip);                                 \
        else                                                                                                   \
            UG_LAUNCH(h, (k_cand64<DD, FF, PP, SS>), dim3(blocks), dim3(64), lds, h->dm, base, count, list,    \
                      dcount, flags, out, head, ctr, cells, mask, skip);                                       \
    } while (0)
    // shared basis loads where the lane-per-timestep evaluation does the work: rechecks and
    // retests of a list, and trajectories (a fresh first pass scans its head across the wave)
    const bool sb = pl <= 2 && h->full && (full || list);
#define UG_CAND64_PL(DD, FF)              \
    do {                                  \
        if (pl == 1 && sb)                \
            UG_CAND64(DD, FF, 1, true);   \
        else if (pl == 1)                 \
            UG_CAND64(DD, FF, 1, false);  \
        else if (pl == 2 && sb)           \
            UG_CAND64(DD, FF, 2, true);   \
        else if (pl == 2)                 \
            UG_CAND64(DD, FF, 2, false);  \
        else                              \
            UG_CAND64(DD, FF, 4, false);  \
    } while (0)
    if (h->d == 2) {
        if (full)
            UG_CAND64_PL(2, true);
        else
            UG_CAND64_PL(2, false);
    } else {
        if (full)
            UG_CAND64_PL(3, true);
        else
            UG_CAND64_PL(3, false);
    }
#undef UG_CAND64_PL
#undef UG_CAND64
    return hip_check(hipGetLastError(), "k_cand64");
}

// ------------------------------------------------------------------ batched launches
int slot_pairs_per_lane(int npairs) { return slot_pl(npairs); }

size_t cand64_lds_bytes(const Pack* h) {
    return (size_t)4 * (h->d == 2 ? cand64_lds_words<2>(h->dm) : cand64_lds_words<3>(h->dm));
}

int32_t launch_cand64_multi(hipStream_t st, int d, int pl, bool full, bool sb, const MCand* dA, int U,
                            int64_t items, size_t lds) {
    if (U <= 0 || items <= 0)
        return UG_OK;
    if (U > 64)
        return fail(UG_EINVAL, "batched candidate test of %d universes (at most 64 a launch)", U);
    if (pl != 1 && pl != 2)
        return fail(UG_EINVAL, "batched candidate test: %d slot pairs per lane (T <= 256 only)", pl);
    const int blocks = (int)std::min<int64_t>(items, 256 * 32);
#define UG_CM(DD, FF, PP, SS) \
    hipLaunchKernelGGL((k_cand64_multi<DD, FF, PP, SS>), dim3(blocks), dim3(64), lds, st, dA, U)
#define UG_CM_PL(DD, FF, SS)          \
    do {                              \
        if (pl == 1)                  \
            UG_CM(DD, FF, 1, SS);     \
        else                          \
            UG_CM(DD, FF, 2, SS);     \
    } while (0)
    if (d == 2) {
        if (full)
            UG_CM_PL(2, true, true);
        else if (sb)
            UG_CM_PL(2, false, true);
        else
            UG_CM_PL(2, false, false);
    } else {
        if (full)
            UG_CM_PL(3, true, true);
        else if (sb)
            UG_CM_PL(3, false, true);
        else
            UG_CM_PL(3, false, false);
    }
#undef UG_CM_PL
#undef UG_CM
    return hip_check(hipGetLastError(), "k_cand64_multi");
}

int32_t launch_compact_multi(hipStream_t st, const MCompact* dA, int U, int64_t max_count) {
    if (U <= 0)
        return UG_OK;
    const int nb = (int)std::max<int64_t>(1, (max_count + CMP_ITEMS - 1) / CMP_ITEMS);
    hipLaunchKernelGGL(k_compact_count_multi, dim3(nb, U), dim3(256), 0, st, dA);
    hipLaunchKernelGGL(k_compact_scatter_multi, dim3(nb, U), dim3(256), 0, st, dA);
    return hip_check(hipGetLastError(), "batched compaction");
}

// The conflict kernels' LDS (above the 64 KB default for D = 3), allowed once per process.
void conf_lds_attributes() {
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void*)k_conflicts_cells<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)conf_lds_bytes(2));
        (void)hipFuncSetAttribute((const void*)k_conflicts_cells<3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)conf_lds_bytes(3));
        (void)hipFuncSetAttribute((const void*)k_conflicts_multi<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)conf_lds_bytes(2));
        (void)hipFuncSetAttribute((const void*)k_conflicts_multi<3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)conf_lds_bytes(3));
    });
}

int32_t launch_resolve_multi(hipStream_t st, int d, const MResolve* dA, int U, int maxS, int maxT) {
    if (U <= 0)
        return UG_OK;
    conf_lds_attributes();
    const dim3 cg(maxT, (maxS + CONF_WG - 1) / CONF_WG, U);
    const int ib = grid_for((int64_t)maxS * maxT, 256);
    if (d == 2) {
        if (maxS > 1)
            hipLaunchKernelGGL(k_conflicts_multi<2>, cg, dim3(CONF_WG), conf_lds_bytes(2), st, dA);
        hipLaunchKernelGGL(k_greedy_multi, dim3(U), dim3(64), 0, st, dA);
        hipLaunchKernelGGL(k_insert_multi<2>, dim3(ib, U), dim3(256), 0, st, dA);
        hipLaunchKernelGGL(k_halo_multi<2>, dim3(ib, U), dim3(256), 0, st, dA);
    } else {
        if (maxS > 1)
            hipLaunchKernelGGL(k_conflicts_multi<3>, cg, dim3(CONF_WG), conf_lds_bytes(3), st, dA);
        hipLaunchKernelGGL(k_greedy_multi, dim3(U), dim3(64), 0, st, dA);
        hipLaunchKernelGGL(k_insert_multi<3>, dim3(ib, U), dim3(256), 0, st, dA);
        hipLaunchKernelGGL(k_halo_multi<3>, dim3(ib, U), dim3(256), 0, st, dA);
    }
    hipLaunchKernelGGL(k_report_multi, dim3(U), dim3(256), 0, st, dA);
    return hip_check(hipGetLastError(), "batched resolve");
}

// Event pairs of profiled launches outlive the handles that used them: a process-wide free
// list per device, so a handle created inside a timed region creates no events once the
// list is warm.
static std::mutex g_ev_mu;
static std::map<int, std::vector<std::array<hipEvent_t, 2>>> g_ev_free;

void release_events(Pack* h) {
    for (const Pack::TimedLaunch& t : h->ev_pending)  // no launch in flight still signals them
        (void)hipEventSynchronize(h->ev_pool[(size_t)t.pair][1]);
    std::lock_guard<std::mutex> g(g_ev_mu);
    auto& f = g_ev_free[h->device];
    f.insert(f.end(), h->ev_pool.begin(), h->ev_pool.end());
    h->ev_pool.clear();
    h->ev_pending.clear();
    h->ev0 = h->ev1 = nullptr;
}

static int32_t take_events(Pack* h) {
    {
        std::lock_guard<std::mutex> g(g_ev_mu);
        auto& f = g_ev_free[h->device];
        if (!f.empty()) {
            h->ev_pool.push_back(f.back());
            f.pop_back();
            return UG_OK;
        }
    }
    // timing only: no system-scope fence (cache writeback) when the events are reached
    std::array<hipEvent_t, 2> e{nullptr, nullptr};
    const unsigned fl = hipEventDisableSystemFence;
    if (hipEventCreateWithFlags(&e[0], fl) != hipSuccess || hipEventCreateWithFlags(&e[1], fl) != hipSuccess) {
        if (e[0])
            (void)hipEventDestroy(e[0]);
        return fail(UG_EHIP, "timed launch: event creation failed");
    }
    h->ev_pool.push_back(e);
    return UG_OK;
}

int32_t timed_begin(Pack* h, int which) {
    h->timing_armed = h->timing_fired = false;
    if (!h->profile || !((h->profile_mask >> which) & 1))
        return UG_OK;
    if (h->ev_pending.size() >= 512) {  // bound the pool: fold what has finished
        const int32_t rc = timed_collect(h);
        if (rc)
            return rc;
    }
    const size_t k = h->ev_pending.size();
    if (k == h->ev_pool.size()) {
        const int32_t rc = take_events(h);
        if (rc)
            return rc;
    }
    h->ev0 = h->ev_pool[k][0];
    h->ev1 = h->ev_pool[k][1];
    h->timing_armed = true;
    return UG_OK;
}

// Books the armed launch (when it was made) under kstats[which].
int32_t timed_end(Pack* h, int which, int64_t items) {
    if (h->profile && h->timing_fired)
        h->ev_pending.push_back({which, (int)h->ev_pending.size(), items});
    h->timing_armed = h->timing_fired = false;
    return UG_OK;
}

int32_t timed_collect(Pack* h) {
    if (!h->profile && h->ev_pending.empty())
        return UG_OK;
    for (const Pack::TimedLaunch& t : h->ev_pending) {
        const auto& e = h->ev_pool[(size_t)t.pair];
        float ms = 0.f;
        int32_t rc = hip_check(hipEventSynchronize(e[1]), "event sync");
        if (!rc)
            rc = hip_check(hipEventElapsedTime(&ms, e[0], e[1]), "event time");
        if (rc)
            return rc;
        ug_kernel_stats& k = h->kstats[t.which];
        k.launches++;
        k.items += t.items;
        k.ms += ms;
    }
    h->ev_pending.clear();
    if (h->d_ctr) {
        std::vector<unsigned long long> c(UG_CTR_WORDS);
        int32_t rc = hip_check(hipStreamSynchronize(h->stream), "counter sync");
        if (!rc)
            rc = hip_check(hipMemcpy(c.data(), h->d_ctr, UG_CTR_WORDS * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost),
                           "counter copy");
        if (rc)
            return rc;
        for (int w = 0; w < 2; w++) {
            int64_t sum[6] = {0, 0, 0, 0, 0, 0};
            for (int k = 0; k < UG_CTR_SLOTS; k++)
                for (int j = 0; j < 6; j++)
                    sum[j] += (int64_t)c[(size_t)(w * UG_CTR_SLOTS + k) * UG_CTR_STRIDE + j];
            h->kstats[w].visits = sum[0];
            h->kstats[w].points = sum[1];
            if (w == UG_KERNEL_RECHECK)
                h->kstats[w].items = sum[2];
            h->kstats[w].left_wsum = sum[3];
            h->kstats[w].left_possible = sum[4];
            h->kstats[w].left_none = sum[5];
        }
    }
    return UG_OK;
}

int32_t hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        return fail(UG_EHIP, "%s: %s", what, hipGetErrorString(e));
    return UG_OK;
}

// Ordered compaction of flags[count] (count bounded by *dcount when given) into out; the
// number kept lands in h->d_count[slot] and, when sync is set, in *S.
int32_t compact(Pack* h, const uint8_t* flags, int64_t count, uint64_t base, const int64_t* list,
                int64_t* out, int64_t* S, const int* dcount, int slot, bool sync, const int* vin,
                int* vout) {
    const int nb = (int)((count + CMP_ITEMS - 1) / CMP_ITEMS);
    if (nb > h->cmp_blocks_cap)
        return fail(UG_EOVERFLOW, "compaction of %lld items exceeds the batch capacity",
                    (long long)count);
    if (count <= 0) {
        if (S)
            *S = 0;
        return hip_check(hipMemsetAsync(h->d_count + slot, 0, 4, h->stream), "count clear");
    }
    hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(256), 0, h->stream, flags, count, dcount, h->d_bcount);
    hipLaunchKernelGGL(k_compact_scatter, dim3(nb), dim3(256), 0, h->stream, flags, count, dcount, base, list,
                       h->d_bcount, nb, out, h->d_count + slot, vin, vout);
    int32_t rc = hip_check(hipGetLastError(), "compact");
    if (rc || !sync)
        return rc;
    rc = hip_check(hipMemcpyAsync(h->h_pinned, h->d_count, 16, hipMemcpyDeviceToHost, h->stream),
                   "compact count copy");
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "compact sync");
    if (!rc)
        *S = ((int*)h->h_pinned)[slot];
    return rc;
}

// out[i] = base + i for i < count; counts[1] = count, counts[0] = counts[2] = 0 (no rechecks).
__global__ void k_all_survive(uint64_t base, int64_t count, int64_t* out, int* counts) {
    const int64_t i0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int64_t i = i0; i < count; i += (int64_t)gridDim.x * 256)
        out[i] = (int64_t)(base + (uint64_t)i);
    if (i0 == 0) {
        counts[0] = 0;
        counts[1] = (int)count;
        counts[2] = 0;
    }
}

// The filter of a sub-batch against an empty table: nothing can touch a path, so every
// candidate survives (survivor count left in d_count[1] as filter_candidates leaves it).
int32_t all_survive(Pack* h, uint64_t base, int64_t count, int64_t* out) {
    hipLaunchKernelGGL(k_all_survive, dim3(grid_for(count, 256 * 4)), dim3(256), 0, h->stream, base, count, out,
                       h->d_count);
    return hip_check(hipGetLastError(), "k_all_survive");
}

// Candidates [base, base+count) or list[0..count) -> survivors of the exact FP64 test
// against the cell table, in index order, into out (*S of them; without sync the count
// stays in d_count[1] and *S is not set).  The split-BF16 form
// prefilters with k_fast_prefilter and rechecks only the candidates it could not reject,
// chained through a device-side count.  Profiled runs time both stages.
int32_t filter_candidates(Pack* h, uint64_t base, int64_t count, const int64_t* list, int64_t* out,
                          int64_t* S, bool sync, const int* dfresh) {
    int32_t rc;
    unsigned long long* ctr = h->profile ? h->d_ctr : nullptr;
    unsigned long long* rctr = h->profile ? h->d_ctr + UG_CTR_WORDS / 2 : nullptr;
    const bool split = h->cfg.precision == UG_PREC_SPLIT_BF16 && h->full;
    // every pass is measured, retests of a survivor list included
    if (!split) {
        rc = timed_begin(h, UG_KERNEL_PREFILTER);
        if (!rc)
            rc = launch_cand64(h, base, count, list, h->d_flags, nullptr, false, ctr, dfresh, -1);
        if (!rc)
            rc = timed_end(h, UG_KERNEL_PREFILTER, count);
        if (!rc)
            rc = compact(h, h->d_flags, count, base, list, out, S, dfresh, 1, sync);
        return rc;
    }
    // stage 1: the first 32 ordered timesteps of every candidate -> survivors in d_list3
    // (count in d_count[0])
    rc = timed_begin(h, UG_KERNEL_PREFILTER);
    if (!rc)
        rc = launch_fast(h, base, count, list, h->d_flags, dfresh, 0, 1, h->d_skip);
    if (!rc)
        rc = timed_end(h, UG_KERNEL_PREFILTER, count);
    if (!rc)
        rc = compact(h, h->d_flags, count, base, list, h->d_list3, nullptr, dfresh, 0, false, (const int*)h->d_skip,
                     (int*)h->d_skip3);
    const int64_t* rlist = h->d_list3;
    const uint32_t* rskip = h->d_skip3;
    int rslot = 0;
    if (!rc && fast_blocks(h) >= UG_STAGE2_MIN_BLOCKS) {
        // stage 2 (T > 160): the remaining ordered timesteps, split-BF16 as well, for the few
        // that stage 1 could not reject -> d_list4 (count in d_count[2]); the FP64 recheck
        // then sees only candidates no timestep certainly rejects.  Measured on one MI355X:
        // 3+1 T=256 packs 1.23x faster with it, 3+1 T=128 about 4% slower (its FP64 recheck
        // is cheap enough), so shorter models skip it.
        rc = timed_begin(h, UG_KERNEL_PREFILTER);
        if (!rc)
            rc = launch_fast(h, 0, count, h->d_list3, h->d_flags, h->d_count, 1, fast_blocks(h));
        if (!rc)
            rc = timed_end(h, UG_KERNEL_PREFILTER, 0);  // items: stage-1 survivors, on the device
        if (!rc)
            rc = compact(h, h->d_flags, count, 0, h->d_list3, h->d_list4, nullptr, h->d_count, 2, false,
                         (const int*)h->d_skip3, (int*)h->d_skip4);
        rlist = h->d_list4;
        rskip = h->d_skip4;
        rslot = 2;
    }
    h->recheck_slot = rslot;
    if (!rc)
        rc = timed_begin(h, UG_KERNEL_RECHECK);
    if (!rc)
        rc = launch_cand64(h, 0, count, rlist, h->d_flags, nullptr, false, rctr, h->d_count + rslot, -1, nullptr,
                           nullptr, rskip);
    if (!rc)
        rc = timed_end(h, UG_KERNEL_RECHECK, 0);  // items: counted on the device
    if (!rc)
        rc = compact(h, h->d_flags, count, 0, rlist, out, S, h->d_count + rslot, 1, sync);
    if (!rc && sync)
        h->rechecked += ((int*)h->h_pinned)[rslot];
    return rc;
}

int32_t ensure_paths(Pack* h, int64_t need) {
    if (need <= h->cap)
        return UG_OK;
    int64_t nc = std::max<int64_t>(need, h->cap * 2);
    double* p64 = nullptr;
    float4* p32 = nullptr;
    int64_t* ai = nullptr;
    int32_t* ag = nullptr;
    const size_t T = (size_t)h->T, D = (size_t)h->d;
    if (!(p64 = (double*)cache_alloc((size_t)nc * T * D * 8)) || !(p32 = (float4*)cache_alloc((size_t)nc * T * 16)) ||
        !(ai = (int64_t*)cache_alloc((size_t)nc * 8)) || !(ag = (int32_t*)cache_alloc((size_t)nc * 4))) {
        cache_free(p64);
        cache_free(p32);
        cache_free(ai);
        cache_free(ag);
        return fail(UG_ENOMEM, "cannot grow path storage to %lld paths", (long long)nc);
    }
    const size_t rows = (size_t)(h->n + h->nsub);  // uniques, then subpaths
    if (rows > 0) {
        (void)hipMemcpyAsync(p64, h->d_pos64, rows * T * D * 8, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(p32, h->d_pos32, rows * T * 16, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(ai, h->d_acc_idx, rows * 8, hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(ag, h->d_acc_gid, rows * 4, hipMemcpyDeviceToDevice, h->stream);
    }
    // the copies have finished before the old buffers can serve another handle
    int32_t rc = hip_check(hipStreamSynchronize(h->stream), "grow copy");
    cache_free(h->d_pos64);
    cache_free(h->d_pos32);
    cache_free(h->d_acc_idx);
    cache_free(h->d_acc_gid);
    h->d_pos64 = p64;
    h->d_pos32 = p32;
    h->d_acc_idx = ai;
    h->d_acc_gid = ag;
    h->cap = nc;
    h->dm.pos64 = p64;
    h->dm.pos32 = p32;
    return rc;
}

int32_t ensure_hash(Pack* h, int64_t points) {
    if (!h->dm.sparse || (uint64_t)points * 2 <= h->hash_cap)
        return UG_OK;
    // load at most 1/2 before the next insert; a new table starts at most 1/3 full
    uint64_t cap = 1u << 16;
    while (cap < (uint64_t)points * 3)
        cap *= 2;
    ulonglong2* slots = (ulonglong2*)cache_alloc((size_t)cap * sizeof(ulonglong2));
    if (!slots)
        return fail(UG_ENOMEM, "sparse grid of %llu slots", (unsigned long long)cap);
    hipLaunchKernelGGL(k_hash_clear, dim3(grid_for((int64_t)cap, 256 * 8)), dim3(256), 0, h->stream, slots, cap);
    ulonglong2* old = h->d_hash;
    h->d_hash = slots;
    h->hash_cap = cap;
    h->dm.hslots = slots;
    h->dm.hmask = cap - 1;
    const int64_t rows = h->n + h->nsub;
    if (rows > 0) {
        const int nb = grid_for(rows * h->T, 256);
        if (h->d == 2)
            hipLaunchKernelGGL(k_hash_rebuild<2>, dim3(nb), dim3(256), 0, h->stream, h->dm, rows, h->d_pos32);
        else
            hipLaunchKernelGGL(k_hash_rebuild<3>, dim3(nb), dim3(256), 0, h->stream, h->dm, rows, h->d_pos32);
    }
    int32_t rc = hip_check(hipGetLastError(), "sparse grid rebuild");
    // the old table may serve another handle only once nothing reads it
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "sparse grid sync");
    cache_free(old);
    return rc;
}

// Conflict bits among the chunk's s survivors (trajectories in d_spos, packed cells in
// d_scell, mask rows cleared: all from its trajectory launch); with dS the count is
// min(s, *dS), known only on the device.  The cell-hash kernel, one workgroup per timestep.
void launch_conflicts(Pack* h, int s, const int* dS) {
    if ((int64_t)s * (s - 1) / 2 <= 0)
        return;
    conf_lds_attributes();
    const dim3 grid(h->T, (s + CONF_WG - 1) / CONF_WG);
    if (h->d == 2)
        hipLaunchKernelGGL(k_conflicts_cells<2>, grid, dim3(CONF_WG), conf_lds_bytes(2), h->stream, h->dm, h->d_spos,
                           h->d_scell, s, dS, h->d_mask);
    else
        hipLaunchKernelGGL(k_conflicts_cells<3>, grid, dim3(CONF_WG), conf_lds_bytes(3), h->stream, h->dm, h->d_spos,
                           h->d_scell, s, dS, h->d_mask);
}

// Resolve the survivors list[0..S) (already clear of the cell table) in index order and
// insert the accepted ones.  Survivors beyond one chunk are retested against the updated
// table before their turn.  Returns the number accepted via *acc_total.
int32_t resolve_and_insert(Pack* h, int64_t* list, int64_t S, int64_t* acc_total, const int* dS,
                           SpecFilter* spec) {
    int32_t rc;
    *acc_total = 0;
    // dS: the survivor count is still on the device (filter without a host round trip); the
    // first chunk runs on min(*dS, chunk_max) and the count comes back with its results
    bool dev = dS != nullptr;
    while (dev || S > 0) {
        const int s = dev ? h->chunk_max : (int)std::min<int64_t>(S, h->chunk_max);
        const int* cs = dev ? dS : nullptr;
        rc = ensure_paths(h, h->n + s);
        if (!rc)
            rc = ensure_hash(h, (h->n + h->nsub + s) * h->T);
        if (rc)
            return rc;
        // trajectories, packed cells and cleared conflict rows of the chunk in one launch
        rc = launch_cand64(h, 0, s, list, nullptr, h->d_spos, true, nullptr, cs, 0, h->d_scell, h->d_mask);
        if (rc)
            return rc;
        // with the count still on the device, the last filter's survivors are the estimate
        launch_conflicts(h, s, cs);
        hipLaunchKernelGGL(k_greedy, dim3(1), dim3(64), 0, h->stream, h->d_mask, s, cs, h->d_acc_list,
                           h->d_nacc);
        const int limit = h->max_n > 0 ? (int)std::min<int64_t>(h->max_n - h->n, INT32_MAX) : INT32_MAX;
        const int ib = grid_for((int64_t)s * h->T, 256);
        // one round trip: accepted count, their candidate indices and the filled-cell count,
        // written into the pinned block by the last block of the halo launch
        char* pin = (char*)h->h_pinned;
        const bool ahead = dev && spec && spec->count > 0;
        const int* rcounts = dev ? h->d_count : nullptr;
        int* spec_count = ahead ? spec->dcount : nullptr;
        const int spec_n = ahead ? (int)spec->count : 0;
        const int64_t stop_lim = ahead ? spec->stop_lim : 0, n_room = ahead ? spec->n_room : 0;
        if (h->d == 2) {
            hipLaunchKernelGGL(k_insert<2>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_acc_list,
                               h->d_nacc, limit, list, nullptr, h->n, h->d_acc_idx, h->d_acc_gid, h->d_pos64,
                               h->d_pos32, h->d_filled);
            hipLaunchKernelGGL(k_halo_report<2>, dim3(ib + 1), dim3(256), 0, h->stream, h->dm, h->d_spos,
                               h->d_acc_list, h->d_nacc, limit, h->d_filled, h->d_acc_idx + h->n, s, rcounts, pin,
                               spec_count, spec_n, stop_lim, n_room);
        } else {
            hipLaunchKernelGGL(k_insert<3>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_acc_list,
                               h->d_nacc, limit, list, nullptr, h->n, h->d_acc_idx, h->d_acc_gid, h->d_pos64,
                               h->d_pos32, h->d_filled);
            hipLaunchKernelGGL(k_halo_report<3>, dim3(ib + 1), dim3(256), 0, h->stream, h->dm, h->d_spos,
                               h->d_acc_list, h->d_nacc, limit, h->d_filled, h->d_acc_idx + h->n, s, rcounts, pin,
                               spec_count, spec_n, stop_lim, n_room);
        }
        rc = hip_check(hipGetLastError(), "resolve kernels");
        if (!rc && ahead) {
            // the next sub-batch's filter goes in behind the report, so the device works on
            // it while the host reads this chunk's results (it only reads the table)
            rc = hip_check(hipEventRecord(h->ev_report, h->stream), "report event");
            if (!rc)
                rc = filter_candidates(h, spec->base, spec->count, nullptr, spec->out, nullptr, false, spec->dcount);
            if (!rc) {
                rc = hip_check(hipEventSynchronize(h->ev_report), "report wait");
                // usable only with the full count (a stop decided on the device leaves 0)
                spec->launched = !rc && *(const int*)(pin + 48) == (int)spec->count;
            }
        } else if (!rc) {
            rc = hip_check(hipStreamSynchronize(h->stream), "resolve sync");
        }
        if (rc)
            return rc;
        if (dev) {
            // the filter's counts: slot 1 = survivors; FP64 rechecks (split form) in slot 0,
            // or in slot 2 after the second split-BF16 stage
            S = ((const int*)(pin + 32))[dS - h->d_count];
            if (h->cfg.precision == UG_PREC_SPLIT_BF16 && h->full)
                h->rechecked += ((const int*)(pin + 32))[h->recheck_slot];
            h->last_S = S;
            dev = false;
        }
        const int na = std::min(*(int*)(pin + 8), limit);
        h->filled = *(unsigned long long*)(pin + 16);
        const int64_t* ai = (const int64_t*)(pin + 64);
        h->acc_host.insert(h->acc_host.end(), ai, ai + na);
        h->n += na;
        *acc_total += na;
        if (h->max_n > 0 && h->n >= h->max_n)
            return UG_OK;
        if (S > h->chunk_max) {
            // retest the rest against the updated table (a filter run ahead saw the table
            // before these inserts: it is dropped)
            if (spec)
                spec->launched = false;
            int64_t S2 = 0;
            rc = filter_candidates(h, 0, S - s, list + s, h->d_list2, &S2);
            if (rc)
                return rc;
            if (h->sub_from_empty) {
                // the sub-batch met an empty table, so its list is the whole range: the retest
                // of candidates sub_base + s .. is the table's first survival-rate observation
                pack_note_rate(h, h->sub_base + s, S - s, S2);
                h->sub_from_empty = false;
            }
            rc = hip_check(hipMemcpyAsync(list, h->d_list2, (size_t)S2 * 8, hipMemcpyDeviceToDevice, h->stream),
                           "list copy");
            if (rc)
                return rc;
            S = S2;
        } else {
            S = 0;
        }
    }
    return UG_OK;
}

// Candidates -> survivors of the subpath test, with their table groups (gout).  In the
// split-BF16 form a fresh range first goes through the prefilter's subpath mode (certain
// two-group contacts among the round-start rows), and only the rest gets the FP64 scan.
int32_t sub_filter(Pack* h, uint64_t base, int64_t count, const int64_t* list, int64_t* out, int* gout,
                   int64_t* S, int* rs_count) {
    if (count <= 0) {
        *S = 0;
        return UG_OK;
    }
    int32_t rc = UG_OK;
    const int* dcount = nullptr;
    if (!list && h->cfg.precision == UG_PREC_SPLIT_BF16 && h->full) {
        rc = launch_fast_sub(h, base, count, h->d_flags, h->sub_nlimit);
        if (!rc)
            rc = compact(h, h->d_flags, count, base, nullptr, h->d_list3, nullptr, nullptr, 0, false);
        list = h->d_list3;
        dcount = h->d_count;
    }
    if (rc)
        return rc;
    const int blocks = (int)std::min<int64_t>(count, 256 * 32);
    const size_t lds = (size_t)4 * (h->d == 2 ? cand64_lds_words<2>(h->dm) : cand64_lds_words<3>(h->dm));
    const int pl = slot_pl(h->npairs);
#define UG_SUBT_X(DD, PP, XX)                                                                                    \
    hipLaunchKernelGGL((k_sub_test<DD, PP, XX>), dim3(blocks), dim3(64), lds, h->stream, h->dm, base, count, list, \
                       dcount, h->d_acc_gid, h->sub_nlimit, rs_count, h->d_flags, h->d_gflag)
#define UG_SUBT(DD, PP)               \
    do {                              \
        if (h->dm.xref)               \
            UG_SUBT_X(DD, PP, true);  \
        else                          \
            UG_SUBT_X(DD, PP, false); \
    } while (0)
    if (h->d == 2) {
        if (pl == 1)
            UG_SUBT(2, 1);
        else if (pl == 2)
            UG_SUBT(2, 2);
        else
            UG_SUBT(2, 4);
    } else {
        if (pl == 1)
            UG_SUBT(3, 1);
        else if (pl == 2)
            UG_SUBT(3, 2);
        else
            UG_SUBT(3, 4);
    }
#undef UG_SUBT
#undef UG_SUBT_X
    rc = hip_check(hipGetLastError(), "k_sub_test");
    if (!rc)
        rc = compact(h, h->d_flags, count, base, list, out, S, dcount, 1, true, h->d_gflag, gout);
    return rc;
}

// Subpath survivors list[0..S) with table groups gl[]: adoption in index order, chunk by
// chunk, later survivors retested against the table after each chunk's insertion.
int32_t resolve_sub(Pack* h, int64_t* list, int* gl, int64_t S, int64_t* acc_total) {
    int32_t rc;
    *acc_total = 0;
    while (S > 0) {
        const int s = (int)std::min<int64_t>(S, UG_S_MAX);
        const int64_t n0 = h->n + h->nsub;
        rc = ensure_paths(h, n0 + s);
        if (!rc)
            rc = ensure_hash(h, (n0 + s) * h->T);
        if (rc)
            return rc;
        rc = launch_cand64(h, 0, s, list, nullptr, h->d_spos, true, nullptr, nullptr, 0, h->d_scell, h->d_mask);
        if (rc)
            return rc;
        launch_conflicts(h, s);
        hipLaunchKernelGGL(k_greedy_sub, dim3(1), dim3(64), 0, h->stream, h->d_mask, s, gl, h->d_gsurv,
                           h->d_acc_list, h->d_nacc);
        const int ib = grid_for((int64_t)s * h->T, 256);
        if (h->d == 2) {
            hipLaunchKernelGGL(k_insert<2>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_acc_list,
                               h->d_nacc, INT32_MAX, list, h->d_gsurv, n0, h->d_acc_idx, h->d_acc_gid, h->d_pos64,
                               h->d_pos32, h->d_filled);
            hipLaunchKernelGGL(k_halo<2>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_acc_list,
                               h->d_nacc, INT32_MAX);
        } else {
            hipLaunchKernelGGL(k_insert<3>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_acc_list,
                               h->d_nacc, INT32_MAX, list, h->d_gsurv, n0, h->d_acc_idx, h->d_acc_gid, h->d_pos64,
                               h->d_pos32, h->d_filled);
            hipLaunchKernelGGL(k_halo<3>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_acc_list,
                               h->d_nacc, INT32_MAX);
        }
        rc = hip_check(hipGetLastError(), "sub resolve kernels");
        char* pin = (char*)h->h_pinned;
        if (!rc)
            rc = hip_check(hipMemcpyAsync(pin + 8, h->d_nacc, 4, hipMemcpyDeviceToHost, h->stream), "nacc copy");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(pin + 16, h->d_filled, 8, hipMemcpyDeviceToHost, h->stream), "filled copy");
        if (!rc)
            rc = hip_check(hipStreamSynchronize(h->stream), "sub resolve sync");
        if (rc)
            return rc;
        const int na = *(int*)(pin + 8);
        h->filled = *(unsigned long long*)(pin + 16);
        h->nsub += na;
        *acc_total += na;
        if (S > s) {
            int64_t S2 = 0;
            rc = sub_filter(h, 0, S - s, list + s, h->d_list2, h->d_glist2, &S2, nullptr);
            if (!rc && S2 > 0)
                rc = hip_check(hipMemcpyAsync(list, h->d_list2, (size_t)S2 * 8, hipMemcpyDeviceToDevice, h->stream),
                               "list copy");
            if (!rc && S2 > 0)
                rc = hip_check(hipMemcpyAsync(gl, h->d_glist2, (size_t)S2 * 4, hipMemcpyDeviceToDevice, h->stream),
                               "gid copy");
            if (rc)
                return rc;
            S = S2;
        } else {
            S = 0;
        }
    }
    return UG_OK;
}

int32_t launch_paths_from_pack(Pack* h, hipStream_t stream, double* a2, double* A, double* B, int kmax,
                               int* d_bmax) {
    const int64_t n = h->n + h->nsub;
    const size_t W = (size_t)kmax + 1;
    // the coefficient rows are filled after the handle's queued work, on the given stream
    int32_t rc = hip_check(hipStreamSynchronize(h->stream), "paths: handle sync");
    if (!rc)
        rc = hip_check(hipMemsetAsync(A, 0, (size_t)n * h->d * W * 8, stream), "paths A clear");
    if (!rc)
        rc = hip_check(hipMemsetAsync(B, 0, (size_t)n * h->d * W * 8, stream), "paths B clear");
    if (!rc)
        rc = hip_check(hipMemsetAsync(d_bmax, 0, 4, stream), "paths bmax clear");
    if (rc)
        return rc;
    const int blocks = (int)std::min<int64_t>(n, 256 * 32);
    // the frequency draws; the whole candidate area with the reference's stream (gen_wave64_xref)
    const size_t lds = h->dm.xref ? cand64_lds_bytes(h) : (size_t)4 * (h->d * h->nw + (h->modmax + 2 + 31) / 32);
    const int pl = slot_pl(h->npairs);
#define UG_PFP_X(DD, PP, XX)                                                                                 \
    hipLaunchKernelGGL((k_paths_from_pack<DD, PP, XX>), dim3(blocks), dim3(64), lds, stream, h->dm, h->d_acc_idx, \
                       n, a2, A, B, kmax, d_bmax)
#define UG_PFP(DD, PP)               \
    do {                             \
        if (h->dm.xref)              \
            UG_PFP_X(DD, PP, true);  \
        else                         \
            UG_PFP_X(DD, PP, false); \
    } while (0)
    if (h->d == 2) {
        if (pl == 1)
            UG_PFP(2, 1);
        else if (pl == 2)
            UG_PFP(2, 2);
        else
            UG_PFP(2, 4);
    } else {
        if (pl == 1)
            UG_PFP(3, 1);
        else if (pl == 2)
            UG_PFP(3, 2);
        else
            UG_PFP(3, 4);
    }
#undef UG_PFP
#undef UG_PFP_X
    return hip_check(hipGetLastError(), "k_paths_from_pack");
}

int32_t launch_check(Pack* h, int mode, uint64_t first, int64_t count, unsigned long long* d_out) {
    if (count <= 0)
        return UG_OK;
    const int blocks = (int)std::min<int64_t>(count, 256 * 32);
    const size_t lds = (size_t)4 * (h->d == 2 ? cand64_lds_words<2>(h->dm) : cand64_lds_words<3>(h->dm));
    const int pl = slot_pl(h->npairs);
#define UG_CHK_X(DD, PP, XX)                                                                                     \
    hipLaunchKernelGGL((k_check<DD, PP, XX>), dim3(blocks), dim3(64), lds, h->stream, h->dm, mode, first, count, \
                       h->d_acc_idx, h->n, d_out)
#define UG_CHK(DD, PP)               \
    do {                             \
        if (h->dm.xref)              \
            UG_CHK_X(DD, PP, true);  \
        else                         \
            UG_CHK_X(DD, PP, false); \
    } while (0)
    if (h->d == 2) {
        if (pl == 1)
            UG_CHK(2, 1);
        else if (pl == 2)
            UG_CHK(2, 2);
        else
            UG_CHK(2, 4);
    } else {
        if (pl == 1)
            UG_CHK(3, 1);
        else if (pl == 2)
            UG_CHK(3, 2);
        else
            UG_CHK(3, 4);
    }
#undef UG_CHK
#undef UG_CHK_X
    return hip_check(hipGetLastError(), "k_check");
}

int32_t launch_knot_test(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags,
                         const int* dcount, const int* kcell, const double* kpos, int64_t ks, int nk) {
    if (count <= 0)
        return UG_OK;
    const int head = list ? (h->d == 2 ? 14 : 13) : h->T;
    const int blocks = (int)std::min<int64_t>(count, 256 * 32);
    const size_t lds = (size_t)4 * (h->d == 2 ? cand64_lds_words<2>(h->dm) : cand64_lds_words<3>(h->dm));
    const int pl = slot_pl(h->npairs);
    const bool sb = pl <= 2 && h->full && list;
#define UG_KNOT(DD, PP, SS)                                                                                      \
    do {                                                                                                         \
        if (h->dm.xref)                                                                                          \
            hipLaunchKernelGGL((k_knot_test<DD, PP, false, true>), dim3(blocks), dim3(64), lds, h->stream, h->dm, \
                               base, count, list, dcount, flags, head, kcell, kpos, ks, nk);                     \
        else                                                                                                     \
            hipLaunchKernelGGL((k_knot_test<DD, PP, SS>), dim3(blocks), dim3(64), lds, h->stream, h->dm, base,    \
                               count, list, dcount, flags, head, kcell, kpos, ks, nk);                           \
    } while (0)
#define UG_KNOT_PL(DD)                \
    do {                              \
        if (pl == 1 && sb)            \
            UG_KNOT(DD, 1, true);     \
        else if (pl == 1)             \
            UG_KNOT(DD, 1, false);    \
        else if (pl == 2 && sb)       \
            UG_KNOT(DD, 2, true);     \
        else if (pl == 2)             \
            UG_KNOT(DD, 2, false);    \
        else                          \
            UG_KNOT(DD, 4, false);    \
    } while (0)
    if (h->d == 2)
        UG_KNOT_PL(2);
    else
        UG_KNOT_PL(3);
#undef UG_KNOT_PL
#undef UG_KNOT
    return hip_check(hipGetLastError(), "k_knot_test");
}

int32_t launch_knot_append(Pack* h, const int64_t* list, int s, int64_t n0, int* kcell, double* kpos, float4* kpos32,
                           int64_t ks, int64_t* kidx, int limit) {
    const int ib = grid_for((int64_t)s * h->T, 256);
    if (h->d == 2)
        hipLaunchKernelGGL(k_knot_append<2>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_scell,
                           h->d_acc_list, h->d_nacc, limit, list, n0, kcell, kpos, kpos32, ks, kidx);
    else
        hipLaunchKernelGGL(k_knot_append<3>, dim3(ib), dim3(256), 0, h->stream, h->dm, h->d_spos, h->d_scell,
                           h->d_acc_list, h->d_nacc, limit, list, n0, kcell, kpos, kpos32, ks, kidx);
    return hip_check(hipGetLastError(), "k_knot_append");
}

int32_t launch_greedy(Pack* h, int s) {
    hipLaunchKernelGGL(k_greedy, dim3(1), dim3(64), 0, h->stream, h->d_mask, s, nullptr, h->d_acc_list, h->d_nacc);
    return hip_check(hipGetLastError(), "k_greedy");
}

}  // namespace ug
