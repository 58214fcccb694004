// ug_knot.hip -- the knot-regime packer (the reference's cuda/braid_knot.cu): few strands,
// very many candidates, no spatial grid.
//
// Candidates are counter-indexed on the packer's Philox stream and admitted in index order,
// so a run is serial RSA per species and does not depend on batch sizes.  Each batch is
// tested against the strands resident at its start by k_knot_test (one wave per candidate,
// canonical FP64 evaluation, every strand compared at every visited timestep), in the
// split-BF16 form after the packer's prefilter in its knot mode has removed the candidates
// with a certain contact in the first 32 ordered timesteps; the survivors
// are resolved in index order by the packer's own resolve kernels (trajectories, pairwise
// conflict bits, greedy) and appended to the timestep-major strand arrays.  Two phases, as
// cuda/braid_knot.cu:387-399: the primary species at the exclusion 2/T over candidates
// [0, --attempts), then, with --nu-attempts, the nu species at (2/T)/nu_div over the next
// nu_attempts candidates, tested against every strand of both species.  The batch ramp of
// :404-469 (4096 per phase, /4 above 1/8 survivors, x4 below 1/64, up to 2^24) is replayed on
// the survivor counts, so a run that reaches the strand capacity stops at the same attempt
// count as the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <vector>

#include "ug_internal.h"

using namespace ug;

struct ug_knot {
    Pack* p = nullptr;        // model, stream and resolve scratch (no cell table)
    ug_knot_config cfg{};
    int* d_kcell = nullptr;   // [T][ks] packed cells of the strands
    double* d_kpos = nullptr; // [T][ks][D] canonical positions
    float4* d_kpos32 = nullptr;  // [T][ks] the same in FP32 (split-BF16 prefilter)
    bool split = false;
    int64_t* d_kidx = nullptr;
    int64_t ks = 0, nk = 0, n_primary = 0, primary_attempts = 0;
    int64_t attempts = 0, batch = 4096;
    int phase = 0;            // 0 primary, 1 nu, 2 done
    double r0 = 0.0;          // primary exclusion 2/T
    std::vector<int64_t> idx;
    std::vector<int32_t> species;
    std::vector<std::array<int64_t, 5>> batches;  // {start, size, survivors, admitted, species}
    int64_t* d_blist = nullptr;  // a batch's survivors
    int64_t blist_cap = 0;
    double seconds = 0.0;
};

namespace {

template <typename P>
void free_cached(P*& ptr) {
    cache_free((void*)ptr);
    ptr = nullptr;
}

void knot_release(ug_knot* K) {
    if (K->p)
        (void)hipStreamSynchronize(K->p->stream);
    free_cached(K->d_kcell);
    free_cached(K->d_kpos);
    free_cached(K->d_kpos32);
    free_cached(K->d_kidx);
    free_cached(K->d_blist);
    if (K->p)
        destroy_handle(K->p);
    K->p = nullptr;
}

// Enter phase ph: its exclusion and a fresh batch ramp.
void set_phase(ug_knot* K, int ph) {
    K->phase = ph;
    K->batch = 4096;
    if (ph == 1) {
        K->p->dm.r = K->r0 / K->cfg.nu_div;
        K->p->dm.r2 = K->p->dm.r * K->p->dm.r;
    }
}

// Resolve the survivors list[0..S) (clear of the strands the batch started with) in index
// order: chunks of at most UG_S_MAX by the packer's resolve kernels, each chunk's accepts
// appended as strands; later survivors are retested against the grown strand set.
int32_t knot_resolve(ug_knot* K, int64_t* list, int64_t S, int64_t* admitted) {
    Pack* h = K->p;
    int32_t rc = UG_OK;
    *admitted = 0;
    int* pin = (int*)h->h_pinned;
    while (S > 0 && K->nk < K->ks && rc == UG_OK) {
        const int s = (int)std::min<int64_t>(S, UG_S_MAX);
        rc = launch_cand64(h, 0, s, list, nullptr, h->d_spos, true, nullptr, nullptr, 0, h->d_scell, h->d_mask);
        if (rc)
            break;
        launch_conflicts(h, s, nullptr);
        rc = launch_greedy(h, s);
        const int limit = (int)std::min<int64_t>(K->ks - K->nk, INT32_MAX);
        if (!rc)
            rc = launch_knot_append(h, list, s, K->nk, K->d_kcell, K->d_kpos, K->d_kpos32, K->ks, K->d_kidx, limit);
        if (!rc)
            rc = hip_check(hipMemcpyAsync(pin, h->d_nacc, 4, hipMemcpyDeviceToHost, h->stream), "knot count");
        if (!rc)
            rc = hip_check(hipStreamSynchronize(h->stream), "knot resolve sync");
        if (rc)
            break;
        const int na = std::min(pin[0], limit);
        if (na > 0) {
            const size_t at = K->idx.size();
            K->idx.resize(at + (size_t)na);
            rc = hip_check(hipMemcpy(K->idx.data() + at, K->d_kidx + K->nk, (size_t)na * 8, hipMemcpyDeviceToHost),
                           "knot indices");
            K->species.insert(K->species.end(), (size_t)na, K->phase);
        }
        K->nk += na;
        *admitted += na;
        if (rc || S <= s || K->nk >= K->ks)
            break;
        // the rest of the batch against the strands as they stand now
        int64_t S2 = 0;
        rc = launch_knot_test(h, 0, S - s, list + s, h->d_flags, nullptr, K->d_kcell, K->d_kpos, K->ks, (int)K->nk);
        if (!rc)
            rc = compact(h, h->d_flags, S - s, 0, list + s, h->d_list2, &S2);
        if (!rc && S2 > 0)
            rc = hip_check(hipMemcpyAsync(list, h->d_list2, (size_t)S2 * 8, hipMemcpyDeviceToDevice, h->stream),
                           "knot list copy");
        S = S2;
    }
    return rc;
}

// One batch [base, base + b): survivors against the strands of its start (pieces of at most
// UG_SUB_MAX candidates, no admission in between), then the index-order resolve.
int32_t knot_batch(ug_knot* K, uint64_t base, int64_t b, int64_t* S_out, int64_t* admitted) {
    Pack* h = K->p;
    int64_t S = 0;
    int32_t rc = UG_OK;
    const int nk0 = (int)K->nk;
    for (int64_t off = 0; off < b && rc == UG_OK; off += UG_SUB_MAX) {
        const int64_t piece = std::min<int64_t>(UG_SUB_MAX, b - off);
        if (S > UG_SUB_MAX)
            return fail(UG_EOVERFLOW, "a knot batch of %lld candidates left more than %d survivors", (long long)b,
                        UG_SUB_MAX);
        if (S + piece > K->blist_cap) {
            const int64_t nc = std::max<int64_t>(2 * UG_SUB_MAX, S + piece);
            int64_t* nl = (int64_t*)cache_alloc((size_t)nc * 8);
            if (!nl)
                return fail(UG_ENOMEM, "knot survivor list of %lld entries", (long long)nc);
            if (S > 0)
                rc = hip_check(hipMemcpyAsync(nl, K->d_blist, (size_t)S * 8, hipMemcpyDeviceToDevice, h->stream),
                               "knot list grow");
            if (!rc)
                rc = hip_check(hipStreamSynchronize(h->stream), "knot list sync");
            cache_free(K->d_blist);
            K->d_blist = nl;
            K->blist_cap = nc;
            if (rc)
                return rc;
        }
        int64_t Sp = piece;
        if (nk0 == 0) {
            rc = all_survive(h, base + (uint64_t)off, piece, K->d_blist + S);
        } else if (K->split) {
            // certain contacts in the first 32 ordered timesteps on the split-BF16 prefilter,
            // then the exact FP64 test of the rest
            int64_t S1 = 0;
            rc = launch_fast_knot(h, base + (uint64_t)off, piece, h->d_flags, K->d_kpos32, K->ks, nk0);
            if (!rc)
                rc = compact(h, h->d_flags, piece, base + (uint64_t)off, nullptr, h->d_list3, &S1);
            if (!rc)
                rc = launch_knot_test(h, 0, S1, h->d_list3, h->d_flags, nullptr, K->d_kcell, K->d_kpos, K->ks, nk0);
            if (!rc)
                rc = compact(h, h->d_flags, S1, 0, h->d_list3, K->d_blist + S, &Sp);
        } else {
            rc = launch_knot_test(h, base + (uint64_t)off, piece, nullptr, h->d_flags, nullptr, K->d_kcell, K->d_kpos,
                                  K->ks, nk0);
            if (!rc)
                rc = compact(h, h->d_flags, piece, base + (uint64_t)off, nullptr, K->d_blist + S, &Sp);
        }
        S += Sp;
    }
    if (rc)
        return rc;
    if (S > UG_SUB_MAX)
        return fail(UG_EOVERFLOW, "a knot batch of %lld candidates left more than %d survivors", (long long)b,
                    UG_SUB_MAX);
    *S_out = S;
    return knot_resolve(K, K->d_blist, S, admitted);
}

void fill_status(const ug_knot* K, ug_knot_status* st) {
    if (!st)
        return;
    st->attempts = K->attempts;
    st->n = K->nk;
    st->n_primary = K->phase >= 1 ? K->n_primary : K->nk;
    st->primary_attempts = K->phase >= 1 ? K->primary_attempts : K->attempts;
    st->phase = K->phase;
    st->done = K->phase >= 2 ? 1 : 0;
    st->seconds = K->seconds;
}

}  // namespace

extern "C" int32_t ug_knot_create(const ug_knot_config* cfg_in, void* hip_stream, ug_knot** out) {
    if (!cfg_in || !out)
        return fail(UG_EINVAL, "ug_knot_create: NULL argument");
    *out = nullptr;
    ug_knot_config cfg = *cfg_in;
    if (cfg.dim != 2 && cfg.dim != 3)
        return fail(UG_EINVAL, "error: --dim must be 2 or 3");
    if (!(cfg.nu_div >= 1.0))
        return fail(UG_EINVAL, "error: --nu-div must be >= 1");
    if (cfg.nu_attempts < 0 || cfg.max_attempts < 0)
        return fail(UG_EINVAL, "error: negative attempt budget");
    if (cfg.max_strands <= 0)
        cfg.max_strands = 16384;
    ug_pack_config pc;
    std::memset(&pc, 0, sizeof pc);
    pc.dim = cfg.dim;
    pc.T = cfg.T;
    pc.terms = cfg.terms;
    if (cfg.precision != UG_PREC_FP64 && cfg.precision != UG_PREC_SPLIT_BF16)
        return fail(UG_EINVAL, "error: unknown precision %d", cfg.precision);
    pc.precision = cfg.precision;
    pc.seed = cfg.seed;
    pc.pin_sin1 = cfg.pin_sin1;
    pc.rng_kind = cfg.rng_kind;
    pc.stop_kind = UG_STOP_BUDGET;
    pc.max_attempts = cfg.max_attempts;
    Pack* h = nullptr;
    int32_t rc = create_handle(&pc, hip_stream, false, &h);
    if (rc)
        return rc;
    ug_knot* K = new ug_knot();
    K->p = h;
    K->cfg = cfg;
    K->ks = cfg.max_strands;
    K->r0 = h->dm.r;
    const size_t cells = (size_t)h->T * (size_t)K->ks;
    K->split = cfg.precision == UG_PREC_SPLIT_BF16 && h->full;
    if (!(K->d_kcell = (int*)cache_alloc(cells * 4)) || !(K->d_kpos = (double*)cache_alloc(cells * h->d * 8)) ||
        !(K->d_kpos32 = (float4*)cache_alloc(cells * 16)) || !(K->d_kidx = (int64_t*)cache_alloc((size_t)K->ks * 8)))
        rc = fail(UG_ENOMEM, "ug_knot_create: strand storage for %lld strands", (long long)K->ks);
    if (rc) {
        knot_release(K);
        delete K;
        return rc;
    }
    set_phase(K, 0);
    *out = K;
    return UG_OK;
}

extern "C" int32_t ug_knot_run(ug_knot* K, int64_t max_candidates, ug_knot_status* st) {
    if (!K)
        return fail(UG_EINVAL, "ug_knot_run: NULL handle");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t call_end = max_candidates > 0 ? K->attempts + max_candidates : INT64_MAX;
    int32_t rc = UG_OK;
    while (K->phase < 2 && rc == UG_OK) {
        const double budget_end = K->phase == 0 ? K->cfg.max_attempts : K->cfg.max_attempts + K->cfg.nu_attempts;
        int64_t b = 0;
        if ((double)K->attempts < budget_end && K->nk < K->ks) {
            b = K->batch;
            if ((double)(K->attempts + b) > budget_end)
                b = (int64_t)(budget_end - (double)K->attempts);
        }
        if (b <= 0) {
            // phase over (cuda/braid_knot.cu:470-475)
            if (K->phase == 0) {
                K->n_primary = K->nk;
                K->primary_attempts = K->attempts;
            }
            if (K->phase == 0 && K->cfg.nu_attempts > 0)
                set_phase(K, 1);
            else
                K->phase = 2;
            continue;
        }
        if (K->attempts >= call_end)
            break;  // batches are never split between calls
        int64_t S = 0, adm = 0;
        const int64_t start = K->attempts;
        rc = knot_batch(K, (uint64_t)start, b, &S, &adm);
        if (rc)
            break;
        K->batches.push_back({start, b, S, adm, K->phase});
        K->attempts += b;
        if (S > b / 8 && K->batch > 4096)
            K->batch = std::max<int64_t>(4096, K->batch / 4);
        else if (S < b / 64 && K->batch < (int64_t(1) << 24))
            K->batch *= 4;
    }
    K->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_status(K, st);
    return rc;
}

extern "C" int32_t ug_knot_get_status(const ug_knot* K, ug_knot_status* st) {
    if (!K || !st)
        return fail(UG_EINVAL, "ug_knot_get_status: NULL argument");
    fill_status(K, st);
    return UG_OK;
}

extern "C" int32_t ug_knot_strands(const ug_knot* K, int64_t* idx, int32_t* species, int64_t cap, int64_t* n) {
    if (!K)
        return fail(UG_EINVAL, "ug_knot_strands: NULL handle");
    if (n)
        *n = K->nk;
    const int64_t k = std::min(cap, K->nk);
    for (int64_t i = 0; i < k; i++) {
        if (idx)
            idx[i] = K->idx[(size_t)i];
        if (species)
            species[i] = K->species[(size_t)i];
    }
    return UG_OK;
}

extern "C" int32_t ug_knot_params(const ug_knot* K, int64_t first, int64_t count, double* out) {
    if (!K || (!out && count > 0))
        return fail(UG_EINVAL, "ug_knot_params: NULL argument");
    if (first < 0 || count < 0 || first + count > K->nk)
        return fail(UG_EINVAL, "ug_knot_params: range [%lld, %lld) outside %lld strands", (long long)first,
                    (long long)(first + count), (long long)K->nk);
    const size_t row = (size_t)K->p->d * (1 + 3 * K->p->nw);
    for (int64_t i = 0; i < count; i++)
        gen_params_host(K->p, (uint64_t)K->idx[(size_t)(first + i)], out + (size_t)i * row);
    return UG_OK;
}

extern "C" int32_t ug_knot_batches(const ug_knot* K, int64_t* rows5, int64_t cap, int64_t* n) {
    if (!K)
        return fail(UG_EINVAL, "ug_knot_batches: NULL handle");
    if (n)
        *n = (int64_t)K->batches.size();
    const int64_t k = std::min<int64_t>(cap, (int64_t)K->batches.size());
    for (int64_t i = 0; i < k && rows5; i++)
        for (int j = 0; j < 5; j++)
            rows5[i * 5 + j] = K->batches[(size_t)i][(size_t)j];
    return UG_OK;
}

extern "C" void ug_knot_destroy(ug_knot* K) {
    if (!K)
        return;
    knot_release(K);
    delete K;
}
