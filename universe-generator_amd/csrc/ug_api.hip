// ug_api.hip -- C ABI entry points and the packing loop of libugpack.so.
//
// The loop replaces the reference engine's main() (cuda/braid_cuda.cu:885-944,
// cuda/braid_cuda3d.cu:1446-1566): candidates are consumed in index order in
// sub-batches; each sub-batch is prefiltered against the device cell table, its
// survivors are resolved in index order and inserted.  Because the accepted list is
// serial RSA in index order, the sub-batch size only affects speed, never results.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ug_arith.h"
#include "ug_internal.h"

using namespace ug;

namespace {

}  // namespace

// Host copy of the canonical generator (same arithmetic as the device's gen_wave64):
// dump parameters of accepted paths, per axis a1 then (amplitude, frequency, phase).
void ug::gen_params_host(const Pack* h, uint64_t idx, double* out_row) {
    const int D = h->d, nw = h->nw, npairs = h->npairs;
    const uint32_t lo = (uint32_t)idx, hi = (uint32_t)(idx >> 32);
    const uint32_t k0 = h->dm.key0, k1 = h->dm.key1;
    if (h->dm.xref) {
        // the reference's serial stream: every axis' draws, then the arithmetic of gen_wave64
        xo256 r = xo_candidate((uint64_t)k0 | (uint64_t)k1 << 32, idx);
        std::vector<int> b(nw);
        std::vector<uint64_t> ek(nw), pk(nw);
        for (int a = 0; a < D; a++) {
            double* row = out_row + (size_t)a * (1 + 3 * nw);
            const uint64_t ua = xo_axis(r, h->modmax, nw, h->dm.pin, b.data(), ek.data(), pk.data(), 1);
            row[0] = h->dm.pin ? 0.0 : xo_a1(ua);
            std::vector<double> e(nw);
            double sum = 0.0;
            for (int j = 0; j < nw; j++) {
                e[j] = xo_spacing(ek[j]);
                sum = sum + e[j];
            }
            for (int j = 0; j < nw; j++) {
                double aj = (e[j] / sum) / (double)b[j];
                if (pk[j] >> 63)
                    aj = -aj;
                row[1 + 3 * j] = aj;
                row[2 + 3 * j] = (double)b[j];
                row[3 + 3 * j] = (b[j] & 1) ? 0.0 : (double)(pk[j] & ((1ull << 53) - 1)) * 0x1p-53 * 3.141592653589793;
            }
        }
        return;
    }
    const u32x4 w0 = philox(lo, hi, 0u, 0u, k0, k1);
    const uint32_t wa[3] = {w0.x, w0.y, w0.z};
    std::vector<int> b(nw);
    std::vector<double> e(nw);
    std::vector<uint32_t> neg(nw), ph(nw);
    for (int a = 0; a < D; a++) {
        double* row = out_row + (size_t)a * (1 + 3 * nw);
        row[0] = h->dm.pin ? 0.0 : a1_of_word(wa[a]);
        if (h->full) {
            for (int j = 0; j < nw; j++)
                b[j] = j + 2;
        } else {
            std::vector<char> used(h->modmax + 2, 0);
            uint32_t g = 0;
            u32x4 wb{0, 0, 0, 0};
            for (int j = 0; j < nw; j++) {
                int bb = 0;
                bool dup = true;
                for (int guard = 0; dup && guard < 64; guard++) {
                    if ((g & 3u) == 0)
                        wb = philox(lo, hi, g >> 2, (2u << 8) | (uint32_t)a, k0, k1);
                    const uint32_t wv = (g & 3u) == 0 ? wb.x : (g & 3u) == 1 ? wb.y : (g & 3u) == 2 ? wb.z : wb.w;
                    bb = (int)(wv % (uint32_t)h->modmax) + 2;
                    g++;
                    dup = used[bb] != 0;
                }
                if (dup)
                    for (int cc = 2; cc <= h->modmax + 1; cc++)
                        if (!used[cc]) {
                            bb = cc;
                            break;
                        }
                used[bb] = 1;
                b[j] = bb;
            }
        }
        double Q[64];
        for (int i = 0; i < 64; i++)
            Q[i] = 0.0;
        for (int p = 0; p < npairs; p++) {
            const u32x4 r = h->full ? pair_words_full(lo, hi, p, a, k0, k1)
                                    : philox(lo, hi, (uint32_t)p, (1u << 8) | (uint32_t)a, k0, k1);
            const uint32_t dw[2] = {r.x, r.y}, pw[2] = {r.z, r.w};
            for (int k = 0; k < 2; k++) {
                const int j = 2 * p + k;
                if (j >= nw)
                    break;
                e[j] = expo_of_word(dw[k]);
                neg[j] = dw[k] & 1u;
                ph[j] = pw[k];
                Q[p & 63] = Q[p & 63] + e[j];
            }
        }
        for (int s = 32; s >= 1; s >>= 1)
            for (int i = 0; i < s; i++)
                Q[i] = Q[i] + Q[i + s];
        const double invW = 1.0 / Q[0];
        for (int j = 0; j < nw; j++) {
            const double invb = 1.0 / (double)b[j];
            double aj = (nw == 1) ? invb : (e[j] * invW) * invb;
            if (neg[j])
                aj = -aj;
            row[1 + 3 * j] = aj;
            row[2 + 3 * j] = (double)b[j];
            row[3 + 3 * j] = (b[j] & 1) ? 0.0 : (double)ph[j] * 0x1p-32 * 3.141592653589793;
        }
    }
}

namespace {

// Buffers go back to the process-wide cache for the next handle (ug_cache.hip).
template <typename P>
void free_dev(P*& p) {
    cache_free((void*)p);
    p = nullptr;
}

void release(Pack* h) {
    // model tables are shared: drop the views and this handle's reference
    h->d_basis16 = nullptr;
    h->d_S = h->d_C = h->d_ST = h->d_CT = nullptr;
    h->d_order = nullptr;
    h->d_marg = nullptr;
    h->tabs.reset();
    free_dev(h->d_tab);
    free_dev(h->d_occ);
    free_dev(h->d_pos64);
    free_dev(h->d_pos32);
    free_dev(h->d_acc_idx);
    free_dev(h->d_acc_gid);
    free_dev(h->d_filled);
    free_dev(h->d_flags);
    free_dev(h->d_list);
    free_dev(h->d_list2);
    free_dev(h->d_list3);
    free_dev(h->d_list4);
    free_dev(h->d_listb);
    free_dev(h->d_skip);
    free_dev(h->d_skip3);
    free_dev(h->d_skip4);
    free_dev(h->d_spec_count);
    if (h->ev_report)
        (void)hipEventDestroy(h->ev_report);
    h->ev_report = nullptr;
    free_dev(h->d_bcount);
    free_dev(h->d_count);
    free_dev(h->d_spos);
    free_dev(h->d_scell);
    free_dev(h->d_mask);
    free_dev(h->d_acc_list);
    free_dev(h->d_nacc);
    free_dev(h->d_ctr);
    free_dev(h->d_gflag);
    free_dev(h->d_glist);
    free_dev(h->d_glist2);
    free_dev(h->d_gsurv);
    free_dev(h->d_rscount);
    free_dev(h->d_rlist);
    h->rlist_cap = 0;
    free_dev(h->d_hash);
    h->hash_cap = 0;
    release_events(h);
    cache_host_free(h->h_pinned);
    h->h_pinned = nullptr;
}

__global__ void k_fill_tab(int* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = -1;
}

int32_t record_curve(Pack* h) {
    while (h->attempts >= h->next_ms) {
        h->curve.push_back({h->attempts, h->n, h->nsub, (int64_t)h->filled});
        h->next_ms = (int64_t)((double)h->next_ms * 1.15) + 1;
    }
    return UG_OK;
}

}  // namespace

extern "C" int32_t ug_pack_create(const ug_pack_config* cfg_in, void* hip_stream, ug_pack** out) {
    if (!cfg_in || !out)
        return fail(UG_EINVAL, "ug_pack_create: NULL argument");
    *out = nullptr;
    Pack* h = nullptr;
    const int32_t rc = create_handle(cfg_in, hip_stream, true, &h);
    if (rc == UG_OK)
        *out = (ug_pack*)h;
    return rc;
}

int32_t ug::create_handle(const ug_pack_config* cfg_in, void* hip_stream, bool with_table, Pack** out) {
    *out = nullptr;
    ug_pack_config cfg = *cfg_in;
    if (cfg.dim != 2 && cfg.dim != 3)
        return fail(UG_EINVAL, "error: dim must be 2 or 3, got %d", cfg.dim);
    if (cfg.T < 4 || cfg.T > UG_MAX_T)
        return fail(UG_EINVAL, "error: T=%d outside the supported range [4, %d]", cfg.T, UG_MAX_T);
    const int modmax = cfg.T - 1 > 2 ? cfg.T - 1 : 2;
    if (cfg.terms == 0)
        cfg.terms = cfg.T;
    if (cfg.terms < 2)
        cfg.terms = 2;
    if ((uint32_t)(cfg.terms - 1) > (uint32_t)modmax)
        return fail(UG_EINVAL, "error: --terms %d exceeds unique-frequency pool %d (max terms %d)",
                    cfg.terms, modmax, modmax + 1);
    if (cfg.precision != UG_PREC_FP64 && cfg.precision != UG_PREC_SPLIT_BF16)
        return fail(UG_EINVAL, "error: unknown precision %d", cfg.precision);
    if (cfg.stop_kind < UG_STOP_BUDGET || cfg.stop_kind > UG_STOP_REF_WINDOW)
        return fail(UG_EINVAL, "error: unknown stop rule %d", cfg.stop_kind);
    if (cfg.rng_kind != UG_RNG_PHILOX && cfg.rng_kind != UG_RNG_XOSHIRO_REF)
        return fail(UG_EINVAL, "error: unknown rng_kind %d", cfg.rng_kind);
    if (cfg.rng_kind == UG_RNG_XOSHIRO_REF && cfg.precision != UG_PREC_FP64)
        return fail(UG_EINVAL, "error: rng_kind UG_RNG_XOSHIRO_REF runs in the FP64 form only (the split-BF16 "
                               "prefilter draws Philox words inside its MFMA loop)");
    if (cfg.stop_kind == UG_STOP_FIXED_WINDOW && !(cfg.stop_rate > 0))
        return fail(UG_EINVAL, "error: the stop rule needs stop_rate > 0");
    if (cfg.stop_kind == UG_STOP_REF_WINDOW && cfg.stop_rate < 0)
        return fail(UG_EINVAL, "error: --until-accept-rate must not be negative");
    if (cfg.stop_batch <= 0)
        cfg.stop_batch = 1 << 20;
    if (cfg.stop_window <= 0)
        cfg.stop_window = 1 << 22;
    if (cfg.stop_kind == UG_STOP_FIXED_WINDOW && (cfg.stop_batch & (cfg.stop_batch - 1)))
        return fail(UG_EINVAL, "error: stop_batch must be a power of two");
    // survivors aimed at per sub-batch: one resolve chunk -- 2048 in 3+1 (fewer chunks of fixed
    // launch costs: 3+1 T=128 packs in 57.3 ms against 60.2 ms with 1024), 1024 in 2+1 (its
    // early sub-batches then stay small: 2+1 T=64 1.715 against 1.743 ms with 2048)
    const int chunk_max = cfg.dim == 3 ? UG_S_MAX : UG_S_MAX / 2;
    if (cfg.sub_target <= 0)
        cfg.sub_target = chunk_max;
    if (cfg.sub_max <= 0)
        cfg.sub_max = UG_SUB_MAX;
    if (cfg.sub_max < 4096 || cfg.sub_max > UG_SUB_MAX || (cfg.sub_max & (cfg.sub_max - 1)))
        return fail(UG_EINVAL, "error: sub_max must be a power of two in [4096, %d]", UG_SUB_MAX);
    if (!(cfg.max_attempts > 0))
        cfg.max_attempts = 9.2e18;

    Pack* h = new Pack();
    h->cfg = cfg;
    h->stream = (hipStream_t)hip_stream;
    (void)hipGetDevice(&h->device);
    h->d = cfg.dim;
    h->T = cfg.T;
    h->nw = cfg.terms - 1;
    h->npairs = (h->nw + 1) / 2;
    h->modmax = modmax;
    h->chunk_max = chunk_max;
    h->full = h->nw == modmax;
    h->max_n = cfg.max_n;
    const int T = h->T;

    // canonical basis tables (and the split-BF16 ones), shared by every handle of the model
    int32_t rc = model_tables(h->d, T, h->nw, cfg.precision == UG_PREC_SPLIT_BF16 && h->full, cfg.margin, &h->tabs);
    if (rc != UG_OK) {
        delete h;
        return rc;
    }
    h->d_S = h->tabs->d_S;
    h->d_C = h->tabs->d_C;
    h->d_ST = h->tabs->d_ST;
    h->d_CT = h->tabs->d_CT;
    h->d_order = h->tabs->d_order;
    int64_t tab_len = (int64_t)T * T * (T + 2);
    if (h->d == 3)
        tab_len *= T;
    const int occw = (T + 2 + 31) / 32 + 1;  // 2+1: one spare word, a 3-bit window may cross a word
    // 3+1: word planes of 4 x 8 row tiles (occ_index)
    const int occ_nwq = h->d == 3 ? (T + 2 + 31) / 32 : 0, occ_nxb = (T + 3) / 4, occ_nyb = (T + 7) / 8;
    const bool sparse = with_table && cfg.sparse;
    // occupancy bits with either grid (a 32nd of the dense table), the dense head table only
    // without --sparse
    h->occ_len = !with_table ? 0
                 : h->d == 3 ? (int64_t)T * occ_nwq * occ_nxb * occ_nyb * 32
                             : tab_len / (T + 2) * occw;
    if (!with_table || sparse)
        tab_len = 0;
    h->tab_len = tab_len;
    int64_t cap0 = cfg.max_paths > 0 ? cfg.max_paths : 4096;

#define UG_ALLOC(ptr, bytes)                                                                 \
    do {                                                                                     \
        if (rc == UG_OK && !((ptr) = (decltype(ptr))cache_alloc(bytes)))                     \
            rc = fail(UG_ENOMEM, "device allocation of %zu bytes failed", (size_t)(bytes));     \
    } while (0)
    if (tab_len > 0)
        UG_ALLOC(h->d_tab, (size_t)tab_len * 4);
    if (h->occ_len > 0)
        UG_ALLOC(h->d_occ, (size_t)h->occ_len * 4);
    UG_ALLOC(h->d_filled, 8);
    UG_ALLOC(h->d_flags, (size_t)UG_SUB_MAX);
    UG_ALLOC(h->d_list, (size_t)UG_SUB_MAX * 8);
    UG_ALLOC(h->d_list2, (size_t)UG_SUB_MAX * 8);
    UG_ALLOC(h->d_list3, (size_t)UG_SUB_MAX * 8);
    UG_ALLOC(h->d_list4, (size_t)UG_SUB_MAX * 8);
    UG_ALLOC(h->d_listb, (size_t)UG_SUB_MAX * 8);
    UG_ALLOC(h->d_skip, (size_t)UG_SUB_MAX * 4);
    UG_ALLOC(h->d_skip3, (size_t)UG_SUB_MAX * 4);
    UG_ALLOC(h->d_skip4, (size_t)UG_SUB_MAX * 4);
    UG_ALLOC(h->d_spec_count, 4);
    h->cmp_blocks_cap = UG_SUB_MAX / 4096 + 1;
    UG_ALLOC(h->d_bcount, (size_t)h->cmp_blocks_cap * 4);
    UG_ALLOC(h->d_count, 16);
    UG_ALLOC(h->d_spos, (size_t)UG_S_MAX * T * h->d * 8);
    UG_ALLOC(h->d_scell, (size_t)UG_S_MAX * T * 4);
    UG_ALLOC(h->d_mask, (size_t)UG_S_MAX * UG_MASK_WORDS * 8);
    UG_ALLOC(h->d_acc_list, (size_t)UG_S_MAX * 4);
    UG_ALLOC(h->d_nacc, 4);
#undef UG_ALLOC
    if (rc == UG_OK && !(h->h_pinned = cache_host_alloc(64 + 8 * UG_S_MAX)))
        rc = fail(UG_ENOMEM, "pinned host allocation failed");
    if (rc == UG_OK && hipEventCreateWithFlags(&h->ev_report, hipEventDisableTiming) != hipSuccess)
        rc = fail(UG_EHIP, "event creation failed");
    if (rc == UG_OK) {
        if (tab_len > 0)
            hipLaunchKernelGGL(k_fill_tab, dim3(grid_for(tab_len, 256 * 8)), dim3(256), 0, h->stream, h->d_tab,
                               tab_len);
        rc = hip_check(hipMemsetAsync(h->d_filled, 0, 8, h->stream), "clear filled");
        if (!rc && h->occ_len > 0)
            rc = hip_check(hipMemsetAsync(h->d_occ, 0, (size_t)h->occ_len * 4, h->stream), "clear occupancy");
    }
    DevModel& m = h->dm;
    m.T = T;
    m.nw = h->nw;
    m.npairs = h->npairs;
    m.modmax = modmax;
    m.pin = cfg.pin_sin1 ? 1 : 0;
    m.full = h->full ? 1 : 0;
    m.key0 = (uint32_t)cfg.seed;
    m.key1 = (uint32_t)(cfg.seed >> 32);
    m.xref = cfg.rng_kind == UG_RNG_XOSHIRO_REF ? 1 : 0;
    m.r = 2.0 / (double)T;
    m.euclid = cfg.euclid ? 1 : 0;
    m.r2 = m.r * m.r;
    m.halfT = 0.5 * (double)T;
    m.ldb = T + 1;
    m.rowlen = T + 2;
    m.S = h->d_S;
    m.C = h->d_C;
    m.ST = h->d_ST;
    m.CT = h->d_CT;
    m.order = h->d_order;
    m.tab = h->d_tab;
    m.occ = h->d_occ;
    m.occw = occw;
    m.occ_nwq = occ_nwq;
    m.occ_nxb = occ_nxb;
    m.occ_nyb = occ_nyb;
    m.pos64 = nullptr;
    m.pos32 = nullptr;
    m.marg = nullptr;
    m.sparse = sparse ? 1 : 0;
    m.hslots = nullptr;
    m.hmask = 0;
    if (rc == UG_OK && sparse)
        rc = ensure_hash(h, std::max<int64_t>(cap0, 1024) * T);
    if (rc == UG_OK && with_table)
        rc = ensure_paths(h, cap0);
    if (rc == UG_OK && cfg.precision == UG_PREC_SPLIT_BF16)
        rc = fast_setup(h);
    if (rc == UG_OK)
        rc = hip_check(hipStreamSynchronize(h->stream), "create sync");
    if (rc != UG_OK) {
        (void)hipStreamSynchronize(h->stream);  // nothing may still write a recycled buffer
        release(h);
        delete h;
        return rc;
    }
    *out = h;
    return UG_OK;
}

void ug::destroy_handle(Pack* h) {
    (void)hipStreamSynchronize(h->stream);
    release(h);
    if (h->own_stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" void ug_pack_destroy(ug_pack* hp) {
    if (hp)
        destroy_handle((Pack*)hp);
}

static void fill_status(const Pack* h, ug_pack_status* st) {
    if (!st)
        return;
    st->attempts = h->attempts;
    st->n = h->n;
    st->nsub = h->nsub;
    st->done = h->done ? 1 : 0;
    st->sub_done = h->sub_done;
    st->survivors = h->survivors;
    st->rechecked = h->rechecked;
    st->margin = h->margin;
    st->seconds = h->seconds;
    st->filled = (int64_t)h->filled;
    st->sub_attempts = h->sub_attempts;
}

extern "C" int32_t ug_pack_get_status(const ug_pack* hp, ug_pack_status* st) {
    if (!hp || !st)
        return fail(UG_EINVAL, "ug_pack_get_status: NULL argument");
    fill_status((const Pack*)hp, st);
    return UG_OK;
}

// Fixed-window rule: at a boundary c (multiple of stop_batch, c >= window) stop when
// the accepts among [c - window, c) are fewer than stop_rate * window.
static bool fixed_window_stop(const Pack* h) {
    const ug_pack_config& c = h->cfg;
    const int64_t at = h->attempts;
    if (at == 0 || at % c.stop_batch != 0 || at < c.stop_window)
        return false;
    const int64_t lo = at - c.stop_window;
    const auto it = std::lower_bound(h->acc_host.begin(), h->acc_host.end(), lo);
    const int64_t cnt = (int64_t)(h->acc_host.end() - it);
    return (double)cnt < c.stop_rate * (double)c.stop_window;
}

void ug::pack_fill_status(const Pack* h, ug_pack_status* st) { fill_status(h, st); }
bool ug::pack_fixed_window_stop(const Pack* h) { return fixed_window_stop(h); }
void ug::pack_record_curve(Pack* h) { (void)record_curve(h); }

static int64_t sub_cap_of(const ug_pack_config& cfg) {
    return std::min<int64_t>(cfg.sub_max, cfg.stop_kind == UG_STOP_FIXED_WINDOW ? cfg.stop_batch : UG_SUB_MAX);
}

// h->sub, cut at the budget, the call's end and (fixed window) the next stop-rule boundary;
// when the rest up to that boundary is within the size the last sizing aimed at, the rest
// whole (one sub-batch instead of a power of two and its remainder).  0 when none may start.
int64_t ug::pack_size_at(const Pack* h, int64_t from, int64_t call_end) {
    const ug_pack_config& cfg = h->cfg;
    const int64_t budget = (int64_t)std::min(cfg.max_attempts, 9.2e18);
    if (from >= budget || from >= call_end)
        return 0;
    const int64_t cap = sub_cap_of(cfg);
    int64_t sub = std::min(h->sub, cap);
    if (cfg.stop_kind == UG_STOP_FIXED_WINDOW) {
        const int64_t rest = (from / cfg.stop_batch + 1) * cfg.stop_batch - from;
        sub = rest <= std::min<double>(h->sub_want, (double)cap) ? rest : std::min(sub, rest);
    }
    sub = std::min(sub, budget - from);
    return std::min(sub, call_end - from);
}

void ug::pack_note_rate(Pack* h, int64_t at, int64_t count, int64_t S) {
    if (at > 0 && count > 0) {
        h->rate_at = at;
        h->rate_p = (double)S / (double)count;
    }
}

// Aims at ~sub_target survivors (UG_S_MAX by default: one in-batch resolve chunk) and at most
// ~16k candidates for the FP64 recheck per sub-batch (the reference adapts toward ~256
// survivors, cuda/braid_cuda.cu:920-924).  The survival rate of the next sub-batch is the last
// observed one scaled by (observed at) / (next start): while the table fills, the rate of a
// random sequential packing falls about as the inverse of the candidates drawn (2+1 T=64
// from 1024 to 2^20 candidates: the rate times the candidate index stays within 16..53, while
// the rate falls 1700-fold), so sizing by the last rate alone lags by the growth in between.
// Every survivor is rechecked in FP64 (split form), so the recheck rate is the survival rate
// plus that of the candidates the prefilter could not decide but which collide; only the
// survival part is scaled.
void ug::pack_size_next(Pack* h, int64_t sub, int64_t S, int64_t rech) {
    const ug_pack_config& cfg = h->cfg;
    const int64_t cap = sub_cap_of(cfg);
    const double L = (double)std::max<int64_t>(1, sub);
    double p = (double)std::max<int64_t>(1, S) / L;
    if (h->rate_at > 0 && h->attempts > 0)
        p = std::max(h->rate_p * (double)h->rate_at / (double)std::max(h->rate_at, h->attempts), 1e-12);
    const double gray = rech > 0 ? (double)std::max<int64_t>(0, rech - S) / L : 0.0;
    double want = (double)cfg.sub_target / p;
    if (rech > 0)
        want = std::min(want, 16384.0 / std::max(p + gray, 1.0 / L));
    want = std::min(want, 64.0 * L);
    int64_t next = 4096;
    while (next * 2 <= (int64_t)want && next * 2 <= cap)
        next *= 2;
    h->sub = next;
    h->sub_want = want;
}

// The smallest accepted count of the next sub-batch (sub candidates) for which the run goes
// on: the loop stops after it iff its uniques number fewer (0: it cannot stop there).  The
// fixed window counts accepts with index >= end - window, all of the sub-batch's among them.
static int64_t stop_limit(const Pack* h, int64_t sub) {
    const ug_pack_config& c = h->cfg;
    const int64_t at = h->attempts + sub;
    if (c.stop_kind == UG_STOP_FIXED_WINDOW) {
        if (at % c.stop_batch != 0 || at < c.stop_window)
            return 0;
        const int64_t lo = at - c.stop_window;
        const int64_t prev = (int64_t)(h->acc_host.end() - std::lower_bound(h->acc_host.begin(), h->acc_host.end(), lo));
        const double thr = c.stop_rate * (double)c.stop_window;
        int64_t x = std::max<int64_t>(0, (int64_t)std::ceil(thr) - prev);
        while (x > 0 && !((double)(prev + x - 1) < thr))
            x--;
        while ((double)(prev + x) < thr)
            x++;
        return x;
    }
    return 0;
}

// One round of the reference schedule: candidates [base, base + size) are tested against
// the table as the round found it (pieces of at most UG_SUB_MAX, no inserts in between), so
// *S is the round's survivor count the engines adapt on (the `got` of
// cuda/braid_cuda.cu:894-897); then the survivors are admitted in index order.
static int32_t ref_round(Pack* h, uint64_t base, int64_t size, int64_t* S_out) {
    const bool empty = h->n + h->nsub == 0;
    int64_t S = 0;
    int32_t rc = UG_OK;
    for (int64_t off = 0; off < size && rc == UG_OK; off += UG_SUB_MAX) {
        const int64_t piece = std::min<int64_t>(UG_SUB_MAX, size - off);
        if (S > UG_SUB_MAX)
            return fail(UG_EOVERFLOW, "a round of %lld candidates left more than %d survivors",
                        (long long)size, UG_SUB_MAX);
        if (S + piece > h->rlist_cap) {
            const int64_t nc = std::max<int64_t>(2 * UG_SUB_MAX, S + piece);
            int64_t* nl = (int64_t*)cache_alloc((size_t)nc * 8);
            if (!nl)
                return fail(UG_ENOMEM, "round survivor list of %lld entries", (long long)nc);
            if (S > 0)
                rc = hip_check(hipMemcpyAsync(nl, h->d_rlist, (size_t)S * 8, hipMemcpyDeviceToDevice, h->stream),
                               "round list copy");
            if (!rc)
                rc = hip_check(hipStreamSynchronize(h->stream), "round list sync");
            cache_free(h->d_rlist);
            h->d_rlist = nl;
            h->rlist_cap = nc;
            if (rc)
                return rc;
        }
        int64_t Sp = piece;
        if (empty)
            rc = all_survive(h, base + (uint64_t)off, piece, h->d_rlist + S);
        else
            rc = filter_candidates(h, base + (uint64_t)off, piece, nullptr, h->d_rlist + S, &Sp, true);
        S += Sp;
    }
    if (rc)
        return rc;
    if (S > UG_SUB_MAX)
        return fail(UG_EOVERFLOW, "a round of %lld candidates left more than %d survivors", (long long)size,
                    UG_SUB_MAX);
    *S_out = S;
    int64_t acc = 0;
    return S > 0 ? resolve_and_insert(h, h->d_rlist, S, &acc) : UG_OK;
}

// The unique phase on the reference engines' round schedule (cuda/braid_cuda.cu:885-944 for
// 2+1, cuda/braid_cuda3d.cu:1446-1566 for 3+1): rounds start at 2^16 candidates and adapt
// toward a few hundred survivors (halve above 512 down to 4096, double below 64 up to 2^26);
// the --until-accept-rate window (clamp(30/rate, 2e7, 5e9) attempts, attempts > 1e6) stops
// after a round.  The 2+1 engine's window records the adapted batch and its loop runs while
// attempts < budget; the 3+1 engine sizes round k+1 before round k is drained (while the
// sized candidates stay below the budget and the window has not fired), records the drained
// round, and drains the sized round after a stop.  Every round here sees the table as it
// stood at its start (the 3+1 engine's kernel sees a grid that lags by at least a round,
// :1247-1249, which can only raise its survivor count).  Survivors are admitted in index
// order, so the accepted list is serial RSA over the consumed candidates and the schedule
// depends only on it: both precision forms run the same rounds.
static int32_t run_ref(Pack* h, int64_t call_end) {
    const ug_pack_config& cfg = h->cfg;
    const double budget = cfg.max_attempts;
    const int64_t target =
        cfg.stop_rate > 0 ? (int64_t)std::min(5e9, std::max(2e7, 30.0 / cfg.stop_rate)) : 0;
    const bool lag = h->d == 3;
    if (h->ref_next < 0) {
        h->ref_batch = 1 << 16;
        h->ref_next = h->ref_batch;
        h->ref_enq = h->ref_batch;
    }
    int32_t rc = UG_OK;
    while (!h->done && rc == UG_OK) {
        if (h->max_n > 0 && h->n >= h->max_n) {
            h->done = true;
            break;
        }
        if (!lag && (double)h->attempts >= budget) {
            h->done = true;
            break;
        }
        if (lag && h->ref_next == 0) {
            h->done = true;
            break;
        }
        if (h->attempts >= call_end)
            break;  // rounds are never split between calls
        int64_t size = h->ref_batch;
        if (lag) {
            size = h->ref_next;
            h->ref_next = 0;
            if ((double)h->ref_enq < budget && !h->ref_stop) {
                h->ref_next = h->ref_batch;
                h->ref_enq += h->ref_next;
            }
        }
        const int64_t start = h->attempts, n0 = h->n;
        const unsigned long long f0 = h->filled;
        int64_t S = 0;
        rc = ref_round(h, (uint64_t)start, size, &S);
        if (rc)
            break;
        const int64_t adm = h->n - n0;
        h->rounds.push_back({start, size, S, adm, (int64_t)(h->filled - f0)});
        h->survivors += S;
        if (h->max_n > 0 && h->n >= h->max_n)
            h->undecided_from = h->acc_host[(size_t)h->max_n - 1] + 1;
        h->attempts += size;
        const int64_t got = std::min<int64_t>(S, 1 << 16);  // survivor buffer, cuda/braid_cuda.cu:760
        if (got > 512 && h->ref_batch > 4096)
            h->ref_batch /= 2;
        else if (got < 64 && h->ref_batch < (int64_t(1) << 26))
            h->ref_batch *= 2;
        rc = record_curve(h);
        if (target) {
            const int64_t e = lag ? size : h->ref_batch;
            h->win.push_back({e, adm});
            h->win_att += e;
            h->win_adm += adm;
            while (h->win.size() > 1 && h->win_att - h->win.front().first >= target) {
                h->win_att -= h->win.front().first;
                h->win_adm -= h->win.front().second;
                h->win.pop_front();
            }
            if (h->win_att >= target && h->attempts > 1000000 &&
                (double)h->win_adm / (double)h->win_att < cfg.stop_rate) {
                if (lag)
                    h->ref_stop = true;
                else
                    h->done = true;
            }
        }
    }
    return rc;
}

extern "C" int32_t ug_pack_run(ug_pack* hp, int64_t max_candidates, ug_pack_status* st) {
    if (!hp)
        return fail(UG_EINVAL, "ug_pack_run: NULL handle");
    Pack* h = (Pack*)hp;
    if (h->d_gflag)
        return fail(UG_ESTATE, "ug_pack_run: the subpath phase has started");
    const auto t0 = std::chrono::steady_clock::now();
    const ug_pack_config& cfg = h->cfg;
    const int64_t budget = (int64_t)std::min(cfg.max_attempts, 9.2e18);
    const int64_t call_end = max_candidates > 0 ? h->attempts + max_candidates : INT64_MAX;
    if (cfg.stop_kind == UG_STOP_REF_WINDOW) {
        int32_t rc = run_ref(h, call_end);
        if (rc == UG_OK)
            rc = hip_check(hipStreamSynchronize(h->stream), "run sync");
        h->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fill_status(h, st);
        return rc;
    }
    const int64_t sub_cap = sub_cap_of(cfg);
    int32_t rc = UG_OK;
    auto size_at = [&](int64_t from) -> int64_t { return pack_size_at(h, from, call_end); };
    // Sub-batches are pipelined one deep: while the host reads sub-batch k's results, the
    // device already filters sub-batch k+1 against the table as it stands after k's first
    // resolve chunk (all of k's inserts when that chunk was the only one; otherwise the
    // filter is redone).  A stop decided after k drops the filter run ahead -- it changed
    // nothing but scratch buffers -- and the device already gave it no candidates when the
    // stop rule or max_n fires after k.
    SpecFilter ahead;
    int64_t* cur_list = h->d_list;
    while (!h->done && rc == UG_OK) {
        if (cfg.stop_kind == UG_STOP_FIXED_WINDOW && fixed_window_stop(h)) {
            h->done = true;
            break;
        }
        if (h->attempts >= budget || (h->max_n > 0 && h->n >= h->max_n)) {
            h->done = true;
            break;
        }
        if (h->attempts >= call_end)
            break;
        const uint64_t base = (uint64_t)h->attempts;
        int64_t sub = 0;
        if (ahead.launched && ahead.base == base) {
            // survivors already filtered, their count still on the device
            sub = ahead.count;
            cur_list = ahead.out;
        } else {
            sub = size_at(h->attempts);
            cur_list = h->d_list;
            int64_t S0 = 0;
            h->sub_base = h->attempts;
            h->sub_from_empty = h->n + h->nsub == 0;
            // Survivors stay on the device: the resolve reads their count there, and it comes
            // back with the resolve's results (one host round trip per sub-batch).  An empty
            // table rejects nothing, so the first sub-batch goes straight to the resolve.
            rc = h->n + h->nsub == 0 ? all_survive(h, base, sub, cur_list)
                                     : filter_candidates(h, base, sub, nullptr, cur_list, &S0, false);
            if (rc)
                break;
        }
        ahead = SpecFilter();
        ahead.base = base + (uint64_t)sub;
        // Only in the steady state (full-size sub-batches, few survivors): earlier, sizes still
        // grow from the measured rates and a sub-batch often needs several resolve chunks,
        // which would void the run ahead.
        // (survivors of this sub-batch expected from the last survival-rate observation, else
        // the last sub-batch's count)
        double expect_S = (double)h->last_S;
        if (h->rate_at > 0)
            expect_S = h->rate_p * (double)h->rate_at / (double)std::max<int64_t>(h->rate_at, (int64_t)base) * (double)sub;
        const bool steady = h->sub >= sub_cap && expect_S <= h->chunk_max / 4;
        ahead.count = steady ? size_at(h->attempts + sub) : 0;
        ahead.out = cur_list == h->d_list ? h->d_listb : h->d_list;
        ahead.dcount = h->d_spec_count;
        ahead.stop_lim = stop_limit(h, sub);
        ahead.n_room = h->max_n > 0 ? h->max_n - h->n : INT64_MAX;
        int64_t S = 0;
        const int64_t rechecked_before = h->rechecked;
        int64_t acc = 0;
        rc = resolve_and_insert(h, cur_list, 0, &acc, h->d_count + 1, &ahead);
        if (rc)
            break;
        S = h->last_S;
        h->survivors += S;
        if (h->max_n > 0 && h->n >= h->max_n) {
            h->attempts = h->acc_host[(size_t)h->max_n - 1] + 1;
            h->done = true;
        } else {
            h->attempts += sub;
        }
        // next sub-batch size (results do not depend on it); a sub-batch filtered against an
        // empty table observes nothing (its retest after the first chunk does)
        if (!h->sub_from_empty)
            pack_note_rate(h, (int64_t)base, sub, S);
        h->sub_from_empty = false;
        pack_size_next(h, sub, S, h->rechecked - rechecked_before);
        rc = record_curve(h);
    }
    if (rc == UG_OK)
        rc = hip_check(hipStreamSynchronize(h->stream), "run sync");
    h->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_status(h, st);
    return rc;
}

extern "C" int32_t ug_pack_accepted(const ug_pack* hp, int64_t* idx, int32_t* gid, int64_t cap, int64_t* n) {
    if (!hp)
        return fail(UG_EINVAL, "ug_pack_accepted: NULL handle");
    const Pack* h = (const Pack*)hp;
    const int64_t tot = h->n + h->nsub;
    if (n)
        *n = tot;
    const int64_t k = std::min(cap, tot);
    if (k <= 0)
        return UG_OK;
    int32_t rc = UG_OK;
    if (idx)
        rc = hip_check(hipMemcpyAsync(idx, h->d_acc_idx, (size_t)k * 8, hipMemcpyDeviceToHost, h->stream), "idx copy");
    if (!rc && gid)
        rc = hip_check(hipMemcpyAsync(gid, h->d_acc_gid, (size_t)k * 4, hipMemcpyDeviceToHost, h->stream), "gid copy");
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "accepted sync");
    return rc;
}

extern "C" int32_t ug_pack_params(const ug_pack* hp, int64_t first, int64_t count, double* out) {
    if (!hp || (!out && count > 0))
        return fail(UG_EINVAL, "ug_pack_params: NULL argument");
    const Pack* h = (const Pack*)hp;
    const int64_t tot = h->n + h->nsub;
    if (first < 0 || count < 0 || first + count > tot)
        return fail(UG_EINVAL, "ug_pack_params: range [%lld, %lld) outside %lld accepted paths",
                    (long long)first, (long long)(first + count), (long long)tot);
    if (count == 0)
        return UG_OK;
    std::vector<int64_t> idx(count);
    int32_t rc = hip_check(hipMemcpyAsync(idx.data(), h->d_acc_idx + first, (size_t)count * 8, hipMemcpyDeviceToHost,
                                          h->stream),
                           "params idx copy");
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "params sync");
    if (rc)
        return rc;
    const size_t row = (size_t)h->d * (1 + 3 * h->nw);
    for (int64_t i = 0; i < count; i++)
        gen_params_host(h, (uint64_t)idx[i], out + (size_t)i * row);
    return UG_OK;
}

extern "C" int32_t ug_pack_positions(const ug_pack* hp, double* dev_out, int64_t cap) {
    if (!hp || !dev_out)
        return fail(UG_EINVAL, "ug_pack_positions: NULL argument");
    const Pack* h = (const Pack*)hp;
    const int64_t tot = h->n + h->nsub;
    if (cap < tot)
        return fail(UG_EOVERFLOW, "ug_pack_positions: capacity %lld < %lld paths", (long long)cap, (long long)tot);
    if (tot == 0)
        return UG_OK;
    int32_t rc = hip_check(hipMemcpyAsync(dev_out, h->d_pos64, (size_t)tot * h->T * h->d * 8, hipMemcpyDeviceToDevice,
                                          h->stream),
                           "positions copy");
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "positions sync");
    return rc;
}

extern "C" int32_t ug_pack_curve(const ug_pack* hp, int64_t* rows4, int64_t cap, int64_t* n) {
    if (!hp)
        return fail(UG_EINVAL, "ug_pack_curve: NULL handle");
    const Pack* h = (const Pack*)hp;
    if (n)
        *n = (int64_t)h->curve.size();
    const int64_t k = std::min<int64_t>(cap, (int64_t)h->curve.size());
    for (int64_t i = 0; i < k && rows4; i++)
        for (int j = 0; j < 4; j++)
            rows4[i * 4 + j] = h->curve[(size_t)i][(size_t)j];
    return UG_OK;
}

extern "C" int32_t ug_pack_rounds(const ug_pack* hp, int64_t* rows5, int64_t cap, int64_t* n) {
    if (!hp)
        return fail(UG_EINVAL, "ug_pack_rounds: NULL handle");
    const Pack* h = (const Pack*)hp;
    if (n)
        *n = (int64_t)h->rounds.size();
    const int64_t k = std::min<int64_t>(cap, (int64_t)h->rounds.size());
    for (int64_t i = 0; i < k && rows5; i++)
        for (int j = 0; j < 5; j++)
            rows5[i * 5 + j] = h->rounds[(size_t)i][(size_t)j];
    return UG_OK;
}

extern "C" int32_t ug_eval_candidates(const ug_pack* hp, const int64_t* dev_idx, int64_t count, int32_t precision,
                                      double* dev_out) {
    if (!hp || (count > 0 && (!dev_idx || !dev_out)))
        return fail(UG_EINVAL, "ug_eval_candidates: NULL argument");
    Pack* h = (Pack*)hp;
    int32_t rc;
    if (precision == UG_PREC_FP64)
        rc = launch_cand64(h, 0, count, dev_idx, nullptr, dev_out, true);
    else if (precision == UG_PREC_SPLIT_BF16 && h->dm.xref)
        return fail(UG_EINVAL, "ug_eval_candidates: the split-BF16 evaluation draws the Philox stream; this "
                               "handle draws the reference's (rng_kind UG_RNG_XOSHIRO_REF, FP64 form only)");
    else if (precision == UG_PREC_SPLIT_BF16)
        rc = launch_fast_eval(h, dev_idx, count, dev_out);
    else
        return fail(UG_EINVAL, "ug_eval_candidates: unknown precision %d", precision);
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "eval sync");
    return rc;
}

extern "C" int32_t ug_pack_check(ug_pack* hp, uint64_t first, int64_t count, int64_t* out) {
    if (!hp || !out)
        return fail(UG_EINVAL, "ug_pack_check: NULL argument");
    Pack* h = (Pack*)hp;
    const int64_t decided = h->undecided_from >= 0 ? h->undecided_from : h->attempts;
    if (count < 0 || first + (uint64_t)count > (uint64_t)decided)
        return fail(UG_EINVAL, "ug_pack_check: candidates [%llu, %llu) beyond the %lld decided",
                    (unsigned long long)first, (unsigned long long)(first + (uint64_t)count),
                    (long long)decided);
    unsigned long long* d_out = nullptr;
    int32_t rc = hip_check(hipMalloc((void**)&d_out, 32), "check counters");
    if (!rc)
        rc = hip_check(hipMemsetAsync(d_out, 0, 32, h->stream), "check clear");
    if (!rc)
        rc = launch_check(h, 0, 0, h->n, d_out);
    if (!rc)
        rc = launch_check(h, 1, first, count, d_out);
    unsigned long long c[4] = {0, 0, 0, 0};
    if (!rc)
        rc = hip_check(hipMemcpyAsync(c, d_out, 32, hipMemcpyDeviceToHost, h->stream), "check copy");
    if (!rc)
        rc = hip_check(hipStreamSynchronize(h->stream), "check sync");
    if (d_out)
        (void)hipFree(d_out);
    for (int i = 0; i < 4; i++)
        out[i] = (int64_t)c[i];
    return rc;
}

extern "C" int32_t ug_pack_profile(ug_pack* hp, int32_t kernels) {
    if (!hp)
        return fail(UG_EINVAL, "ug_pack_profile: NULL handle");
    if (kernels < 0 || (kernels & ~(UG_PROFILE_ALL | UG_PROFILE_REASONS)))
        return fail(UG_EINVAL, "ug_pack_profile: kernels=%d is not a set of UG_PROFILE_* bits", kernels);
    Pack* h = (Pack*)hp;
    if (kernels && !h->d_ctr) {
        if (!(h->d_ctr = (unsigned long long*)cache_alloc(UG_CTR_WORDS * sizeof(unsigned long long))))
            return fail(UG_ENOMEM, "ug_pack_profile: counter allocation failed");
        if (hipMemset(h->d_ctr, 0, UG_CTR_WORDS * sizeof(unsigned long long)) != hipSuccess)
            return fail(UG_EHIP, "ug_pack_profile: counter clear failed");
    }
    h->profile = kernels != 0;
    h->profile_mask = kernels;
    return UG_OK;
}

extern "C" int32_t ug_pack_kernel_stats(ug_pack* hp, int32_t which, ug_kernel_stats* out) {
    if (!hp || !out || which < 0 || which > 1)
        return fail(UG_EINVAL, "ug_pack_kernel_stats: invalid argument");
    Pack* h = (Pack*)hp;
    const int32_t rc = timed_collect(h);
    if (rc)
        return rc;
    *out = h->kstats[which];
    return UG_OK;
}

extern "C" int32_t ug_pack_run_sub(ug_pack* hp, const ug_sub_config* sc, int64_t max_candidates,
                                   ug_pack_status* st) {
    if (!hp || !sc)
        return fail(UG_EINVAL, "ug_pack_run_sub: NULL argument");
    Pack* h = (Pack*)hp;
    if (sc->sub_terms != 0 && sc->sub_terms != h->cfg.terms)
        return fail(UG_EINVAL, "error: --sub-terms %d differs from --terms %d (not supported)", sc->sub_terms,
                    h->cfg.terms);
    if (h->d_gflag == nullptr) {
        // all or nothing: d_gflag marks the phase as started, so it is set last
        int *gflag = (int*)cache_alloc((size_t)UG_SUB_MAX * 4), *glist = (int*)cache_alloc((size_t)UG_SUB_MAX * 4);
        int *glist2 = (int*)cache_alloc((size_t)UG_SUB_MAX * 4), *gsurv = (int*)cache_alloc((size_t)UG_S_MAX * 4);
        int* rscount = (int*)cache_alloc(16);
        if (!gflag || !glist || !glist2 || !gsurv || !rscount) {
            cache_free(gflag);
            cache_free(glist);
            cache_free(glist2);
            cache_free(gsurv);
            cache_free(rscount);
            return fail(UG_ENOMEM, "ug_pack_run_sub: allocation failed");
        }
        h->d_glist = glist;
        h->d_glist2 = glist2;
        h->d_gsurv = gsurv;
        h->d_rscount = rscount;
        h->d_gflag = gflag;
        h->done = true;  // the unique phase ends where the subpath phase starts
    }
    const auto t0 = std::chrono::steady_clock::now();
    const double budget = sc->max_attempts > 0 ? sc->max_attempts : h->cfg.max_attempts;
    const int64_t call_end = max_candidates > 0 ? h->sub_attempts + max_candidates : INT64_MAX;
    // windowed stops of cuda/braid_cuda.cu:956-1037 (30 events, clamped to [2e7, 5e9] attempts)
    const int64_t a_target = sc->accept_rate > 0 ? (int64_t)std::min(5e9, std::max(2e7, 30.0 / sc->accept_rate)) : 0;
    const int64_t f_target = sc->fill_rate > 0 ? (int64_t)std::min(5e9, std::max(2e7, 30.0 / sc->fill_rate)) : 0;
    int32_t rc = UG_OK;
    int* rs_host = (int*)((char*)h->h_pinned + 24);
    // One round = sub_batch candidates tested against the packing as it stood when the round
    // began (rows < sub_nlimit), survivors admitted in index order.  Rounds are never split
    // between calls, so results do not depend on max_candidates.
    while (!h->sub_done && rc == UG_OK) {
        if ((double)h->sub_attempts >= budget) {
            h->sub_done = UG_SUB_BUDGET;
            break;
        }
        if (h->sub_attempts >= call_end)
            break;
        const int64_t round = h->sub_batch;
        h->sub_nlimit = h->n + h->nsub;
        const unsigned long long filled0 = h->filled;
        rc = hip_check(hipMemsetAsync(h->d_rscount, 0, 4, h->stream), "rscount clear");
        int64_t adm = 0;
        for (int64_t off = 0; off < round && rc == UG_OK; off += UG_SUB_MAX) {
            const int64_t piece = std::min<int64_t>(UG_SUB_MAX, round - off);
            int64_t S = 0, a = 0;
            rc = sub_filter(h, (uint64_t)(h->attempts + off), piece, nullptr, h->d_list, h->d_glist, &S,
                            h->d_rscount);
            if (rc == UG_OK)
                rc = resolve_sub(h, h->d_list, h->d_glist, S, &a);
            adm += a;
        }
        if (rc == UG_OK)
            rc = hip_check(hipMemcpyAsync(rs_host, h->d_rscount, 4, hipMemcpyDeviceToHost, h->stream),
                           "rscount copy");
        if (rc == UG_OK)
            rc = hip_check(hipStreamSynchronize(h->stream), "round sync");
        if (rc)
            break;
        const int64_t got = std::min<int64_t>(*rs_host, 1 << 16);  // survCap at KMAX_WIGGLE > 16
        h->survivors += *rs_host;
        h->attempts += round;
        h->sub_attempts += round;
        const int64_t fresh = (int64_t)(h->filled - filled0);
        if (got > 512 && h->sub_batch > 4096)
            h->sub_batch /= 2;
        else if (got < 64 && h->sub_batch < (int64_t(1) << 26))
            h->sub_batch *= 2;
        // the windows record the adapted round size, as cuda/braid_cuda.cu:1009-1010 does
        if (a_target) {
            h->awin.push_back({h->sub_batch, adm});
            h->a_att += h->sub_batch;
            h->a_adm += adm;
            while (h->awin.size() > 1 && h->a_att - h->awin.front().first >= a_target) {
                h->a_att -= h->awin.front().first;
                h->a_adm -= h->awin.front().second;
                h->awin.pop_front();
            }
            if (h->a_att >= a_target && h->sub_attempts > 1000000 &&
                (double)h->a_adm / (double)h->a_att < sc->accept_rate)
                h->sub_done = UG_SUB_ACCEPT_RATE;
        }
        if (f_target && !h->sub_done) {
            h->fwin.push_back({h->sub_batch, fresh});
            h->f_att += h->sub_batch;
            h->f_new += fresh;
            while (h->fwin.size() > 1 && h->f_att - h->fwin.front().first >= f_target) {
                h->f_att -= h->fwin.front().first;
                h->f_new -= h->fwin.front().second;
                h->fwin.pop_front();
            }
            if (h->f_att >= f_target && h->sub_attempts > 1000000 &&
                (double)h->f_new / (double)h->f_att < sc->fill_rate)
                h->sub_done = UG_SUB_FILL_RATE;
        }
        rc = record_curve(h);
    }
    if (rc == UG_OK)
        rc = hip_check(hipStreamSynchronize(h->stream), "run_sub sync");
    h->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_status(h, st);
    return rc;
}
