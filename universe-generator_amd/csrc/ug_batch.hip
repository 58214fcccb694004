// ug_batch.hip -- the batched ensemble engine: many independent universes (uneven T and
// dimension, distinct seeds) packed in lock-step on one stream, each stage of a step one
// launch for all of them.
//
// A universe alone is a chain of small dependent launches with a host round trip per
// sub-batch (ug_pack_run); thousands of such chains are bound by launch and round-trip
// latency, not by the device.  Here a step advances every universe in flight by one
// sub-batch (or one retest of survivors left over from its last resolve chunk):
//   filter:  split-BF16 prefilter of all universes' tiles -> compaction -> FP64 recheck of
//            every universe's flagged candidates -> compaction (or the FP64 test alone);
//            launches over a flattened (universe, tile / candidate) work queue;
//   resolve: the first <= 1024 survivors of each universe: trajectories, conflict bits,
//            index-order greedy, insertion, halo cells and the report, blockIdx.y (or x) =
//            universe;
// with one host synchronisation after each half.  The kernels are the single-universe
// kernels' own bodies (cand64_item, prefilter_item, ...), each universe sees exactly the
// work it would see alone, and a universe's accepted list is the serial RSA of its config
// whatever the others do.  The reference runs one engine process per universe
// (braidlab/orchestrator.py:70-82,106-107).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ug_internal.h"

using namespace ug;

namespace {

// Argument arrays of a step's launches: written into a pinned host block, copied to a
// device block in one transfer per phase on the engine's stream, reused only after the
// stream has been synchronised.
struct ArgStage {
    char* host = nullptr;
    char* dev = nullptr;
    size_t cap = 0, used = 0, flushed = 0;
    // device address of n bytes copied from src (nullptr when the block is full)
    void* put(const void* src, size_t n) {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + n > cap)
            return nullptr;
        memcpy(host + off, src, n);
        used = off + n;
        return dev + off;
    }
    int32_t flush(hipStream_t st) {
        if (used <= flushed)
            return UG_OK;
        const int32_t rc = hip_check(
            hipMemcpyAsync(dev + flushed, host + flushed, used - flushed, hipMemcpyHostToDevice, st), "batch args");
        flushed = used;
        return rc;
    }
    void reset() { used = flushed = 0; }
};

// out[i] = base + i, i < count; counts[0] = 0 (no rechecks), counts[1] = count: the filter of
// a sub-batch against an empty table (blockIdx.y = universe).
struct MAll {
    uint64_t base;
    int64_t count;
    int64_t* out;
    int* counts;
};
__global__ __launch_bounds__(256) void k_all_survive_multi(const MAll* A) {
    const MAll& a = A[blockIdx.y];
    const int64_t i0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int64_t i = i0; i < a.count; i += (int64_t)gridDim.x * 256)
        a.out[i] = (int64_t)(a.base + (uint64_t)i);
    if (i0 == 0) {
        a.counts[0] = 0;
        a.counts[1] = (int)a.count;
    }
}

constexpr size_t REP_BYTES = 24 + 8 * UG_S_MAX;  // one universe's resolve report (k_report_multi)

struct Uni {
    int idx = -1;  // position in the caller's arrays
    Pack* h = nullptr;
    bool fresh = true;          // the step filters a fresh sub-batch (else retests pend)
    int64_t sub = 0;            // the current sub-batch's size
    int64_t* list = nullptr;    // the current survivors (first resolve chunk at list[0..))
    int64_t* pend = nullptr;    // survivors after the last resolve chunk, to retest
    int64_t npend = 0;
    int64_t S = 0;              // this step's survivors
    int64_t S_first = 0;        // survivors of the sub-batch's fresh filter (sizing)
    int64_t rech_sub = 0;       // FP64 rechecks during the sub-batch (sizing)
    int chunk = 0;              // this step's resolve chunk
    int64_t bound = 0;          // upper bound on this step's chunk (sizes the resolve grids)
    int slot = 0;               // position among the step's universes (report block)
    bool in_b = false;          // list lives in the handle's d_listb (else d_list)
    bool all = false;           // filter against an empty table
    std::chrono::steady_clock::time_point t0;
};

int64_t budget_of(const Pack* h) { return (int64_t)std::min(h->cfg.max_attempts, 9.2e18); }

// Size of the sub-batch that starts at h->attempts (ug_pack_run's rule).
int64_t size_at(const Pack* h) { return pack_size_at(h, h->attempts, INT64_MAX); }

// The run has ended before the next sub-batch (stop rule, budget or max_n).
bool finished(Pack* h) {
    if (h->cfg.stop_kind == UG_STOP_FIXED_WINDOW && pack_fixed_window_stop(h))
        return true;
    return h->attempts >= budget_of(h) || (h->max_n > 0 && h->n >= h->max_n);
}

// A config the batched kernels cover: dense cell table, full spectrum, T <= 160 (one
// prefilter stage), budget or fixed-window stop, unique mode, the Philox stream.
bool batchable(const ug_pack_config& c) {
    const int terms = c.terms == 0 ? c.T : c.terms;
    return (c.dim == 2 || c.dim == 3) && c.T >= 4 && c.T <= 160 && terms == c.T && !c.sparse &&
           (c.stop_kind == UG_STOP_BUDGET || c.stop_kind == UG_STOP_FIXED_WINDOW) && c.rng_kind == UG_RNG_PHILOX;
}

struct Engine {
    hipStream_t st = nullptr;
    ArgStage stage;
    char* h_rep = nullptr;  // pinned: per universe resolve report
    int cap = 0;            // universes in flight at most
    std::vector<std::function<int32_t()>> launches;  // the step's launches, after the upload

    // stage n bytes for a launch; the launch reads them once the step's records are uploaded
    const void* put(const void* src, size_t n) { return stage.put(src, n); }
    int32_t step(std::vector<Uni*>& act);
    int32_t stage_filter(std::vector<Uni*>& act);
    int32_t stage_resolve(std::vector<Uni*>& act);
};

MCand cand_arg(Pack* h) {
    MCand a;
    memset(&a, 0, sizeof a);
    a.m = h->dm;
    return a;
}

#define UG_STAGE(ptr, vec)                                                             \
    const auto* ptr = (decltype((vec).data()))put((vec).data(), (vec).size() * sizeof((vec)[0])); \
    if (!ptr)                                                                          \
        return fail(UG_EOVERFLOW, "batched engine: argument staging full")

// The filter of every active universe: survivors of its fresh sub-batch or of its retest
// list into u->list, counts {rechecks, survivors} in its d_count (left on the device).
int32_t Engine::stage_filter(std::vector<Uni*>& act) {
    int32_t rc = UG_OK;
    std::vector<Uni*> split, fp64, all;
    for (int i = 0; i < (int)act.size(); i++) {
        Uni* u = act[(size_t)i];
        Pack* h = u->h;
        // a retest reads the rest of the last list and writes the other buffer
        u->in_b = u->fresh ? false : !u->in_b;
        u->list = u->in_b ? h->d_listb : h->d_list;
        u->slot = i;
        u->all = u->fresh && h->n == 0;
        const int64_t cnt = u->fresh ? u->sub : u->npend;
        u->bound = std::min<int64_t>(cnt, h->chunk_max);
        if (u->all)
            all.push_back(u);
        else if (h->cfg.precision == UG_PREC_SPLIT_BF16)
            split.push_back(u);
        else
            fp64.push_back(u);
    }
    auto cnt_of = [](const Uni* u) { return u->fresh ? u->sub : u->npend; };
    // split-BF16 stage 1, one launch per dimension over the universes' tiles
    for (int d = 2; d <= 3; d++) {
        std::vector<char> args;
        int64_t tiles = 0;
        int n = 0;
        for (Uni* u : split) {
            Pack* h = u->h;
            if (h->d != d)
                continue;
            args.resize(args.size() + fast_multi_arg_bytes());
            rc = fast_multi_arg(h, (uint64_t)h->attempts, cnt_of(u), u->fresh ? nullptr : u->pend, h->d_flags,
                                h->d_skip, args.data() + args.size() - fast_multi_arg_bytes());
            if (rc)
                return rc;
            tiles += (cnt_of(u) + 31) / 32;
            n++;
        }
        if (!n)
            continue;
        UG_STAGE(dA, args);
        launches.push_back([=]() { return launch_fast_multi(st, d, dA, n, tiles); });
    }
    // stage-1 survivors into d_list3 (count in d_count[0])
    if (!split.empty()) {
        std::vector<MCompact> cs;
        int64_t mx = 0;
        for (Uni* u : split) {
            Pack* h = u->h;
            cs.push_back(MCompact{h->d_flags, cnt_of(u), nullptr, (uint64_t)h->attempts, u->fresh ? nullptr : u->pend,
                                  h->d_bcount, h->d_list3, h->d_count + 0, (const int*)h->d_skip,
                                  (int*)h->d_skip3});
            mx = std::max(mx, cnt_of(u));
        }
        UG_STAGE(dA, cs);
        const int n = (int)cs.size();
        launches.push_back([=]() { return launch_compact_multi(st, dA, n, mx); });
    }
    // FP64 tests: the split universes' rechecks (device-counted lists) and the FP64 form's
    // candidates, one launch per (dimension, slot pairs per lane, kind)
    struct Group {
        int d, pl;
        bool sb;
        std::vector<MCand> a;
        size_t lds = 0;
        int64_t items = 0;
    };
    std::vector<Group> groups;
    auto add = [&](int d, int pl, bool sb, const MCand& a, size_t lds) {
        for (Group& g : groups)
            if (g.d == d && g.pl == pl && g.sb == sb) {
                g.a.push_back(a);
                g.lds = std::max(g.lds, lds);
                g.items += a.count;
                return;
            }
        groups.push_back(Group{d, pl, sb, {a}, lds, a.count});
    };
    for (Uni* u : split) {
        Pack* h = u->h;
        MCand a = cand_arg(h);
        a.count = cnt_of(u);
        a.list = h->d_list3;
        a.dcount = h->d_count + 0;
        a.flags = h->d_flags;
        a.skip = h->d_skip3;
        a.head = h->d == 2 ? 14 : 13;
        add(h->d, slot_pairs_per_lane(h->npairs), true, a, cand64_lds_bytes(h));
    }
    for (Uni* u : fp64) {
        Pack* h = u->h;
        MCand a = cand_arg(h);
        a.base = (uint64_t)h->attempts;
        a.count = cnt_of(u);
        a.list = u->fresh ? nullptr : u->pend;
        a.flags = h->d_flags;
        a.head = u->fresh ? h->T : (h->d == 2 ? 14 : 13);
        add(h->d, slot_pairs_per_lane(h->npairs), !u->fresh, a, cand64_lds_bytes(h));
    }
    for (Group& g : groups) {
        UG_STAGE(dA, g.a);
        const int n = (int)g.a.size(), d = g.d, pl = g.pl;
        const bool sb = g.sb;
        const int64_t items = g.items;
        const size_t lds = g.lds;
        launches.push_back([=]() { return launch_cand64_multi(st, d, pl, false, sb, dA, n, items, lds); });
    }
    // FP64 survivors into the universes' lists (count in d_count[1])
    if (!split.empty() || !fp64.empty()) {
        std::vector<MCompact> cs;
        int64_t mx = 0;
        for (Uni* u : split) {
            Pack* h = u->h;
            cs.push_back(MCompact{h->d_flags, cnt_of(u), h->d_count + 0, 0, h->d_list3, h->d_bcount, u->list,
                                  h->d_count + 1, nullptr, nullptr});
            mx = std::max(mx, cnt_of(u));
        }
        for (Uni* u : fp64) {
            Pack* h = u->h;
            cs.push_back(MCompact{h->d_flags, cnt_of(u), nullptr, (uint64_t)h->attempts, u->fresh ? nullptr : u->pend,
                                  h->d_bcount, u->list, h->d_count + 1, nullptr, nullptr});
            mx = std::max(mx, cnt_of(u));
        }
        UG_STAGE(dA, cs);
        const int n = (int)cs.size();
        launches.push_back([=]() { return launch_compact_multi(st, dA, n, mx); });
    }
    if (!all.empty()) {
        std::vector<MAll> as;
        int64_t mx = 0;
        for (Uni* u : all) {
            Pack* h = u->h;
            as.push_back(MAll{(uint64_t)h->attempts, u->sub, u->list, h->d_count});
            mx = std::max(mx, u->sub);
        }
        UG_STAGE(dA, as);
        const int n = (int)as.size();
        launches.push_back([=]() {
            hipLaunchKernelGGL(k_all_survive_multi, dim3(grid_for(mx, 256 * 4), (unsigned)n), dim3(256), 0, st, dA);
            return hip_check(hipGetLastError(), "k_all_survive_multi");
        });
    }
    return UG_OK;
}

// The first resolve chunk of every active universe (at most 1024 of the filter's survivors,
// whose count stays on the device): trajectories, conflicts, greedy, insertion and report.
int32_t Engine::stage_resolve(std::vector<Uni*>& act) {
    // trajectories, packed cells and cleared conflict rows (one launch per dimension and
    // slot pairs per lane; item counts on the device)
    for (int d = 2; d <= 3; d++)
        for (int pl = 1; pl <= 2; pl++) {
            std::vector<MCand> a;
            int64_t items = 0;
            size_t lds = 0;
            for (Uni* u : act) {
                Pack* h = u->h;
                if (h->d != d || slot_pairs_per_lane(h->npairs) != pl)
                    continue;
                MCand x = cand_arg(h);
                x.count = u->bound;
                x.dcount = h->d_count + 1;
                x.list = u->list;
                x.out = h->d_spos;
                x.cells = h->d_scell;
                x.mask = h->d_mask;
                a.push_back(x);
                items += u->bound;
                lds = std::max(lds, cand64_lds_bytes(h));
            }
            if (a.empty())
                continue;
            UG_STAGE(dA, a);
            const int n = (int)a.size();
            launches.push_back([=]() { return launch_cand64_multi(st, d, pl, true, true, dA, n, items, lds); });
        }
    // conflicts, greedy, insertion, halo cells and report, one set of launches per dimension;
    // grids cover each universe's bound on the chunk, the kernels stop at the device count
    for (int d = 2; d <= 3; d++) {
        std::vector<MResolve> a;
        int maxS = 0, maxT = 0;
        for (Uni* u : act) {
            Pack* h = u->h;
            if (h->d != d)
                continue;
            MResolve x;
            memset(&x, 0, sizeof x);
            x.m = h->dm;
            x.pos = h->d_spos;
            x.cells = h->d_scell;
            x.mask = h->d_mask;
            x.counts = h->d_count;
            x.S = (int)u->bound;
            x.limit = h->max_n > 0 ? (int)std::min<int64_t>(h->max_n - h->n, INT32_MAX) : INT32_MAX;
            x.acc_list = h->d_acc_list;
            x.nacc = h->d_nacc;
            x.surv_idx = u->list;
            x.n0 = h->n;
            x.acc_idx = h->d_acc_idx;
            x.acc_gid = h->d_acc_gid;
            x.pos64 = h->d_pos64;
            x.pos32 = h->d_pos32;
            x.filled = h->d_filled;
            x.rep = (int*)(h_rep + REP_BYTES * (size_t)u->slot);
            a.push_back(x);
            maxS = std::max(maxS, x.S);
            maxT = std::max(maxT, h->T);
        }
        if (a.empty())
            continue;
        UG_STAGE(dA, a);
        const int n = (int)a.size();
        launches.push_back([=]() {
            return launch_resolve_multi(st, d, dA, n, maxS, maxT);
        });
    }
    return UG_OK;
}

// One step of a lane: room for each universe's largest chunk, every record staged and
// uploaded in one copy, the launches, one synchronisation; then the reports.
int32_t Engine::step(std::vector<Uni*>& act) {
    int32_t rc = UG_OK;
    for (Uni* u : act) {
        const int64_t cnt = u->fresh ? u->sub : u->npend;
        rc = ensure_paths(u->h, u->h->n + std::min<int64_t>(cnt, u->h->chunk_max));
        if (rc)
            return rc;
    }
    launches.clear();
    stage.reset();
    rc = stage_filter(act);
    if (!rc)
        rc = stage_resolve(act);
    if (!rc)
        rc = stage.flush(st);
    for (size_t k = 0; k < launches.size() && !rc; k++)
        rc = launches[k]();
    if (!rc)
        rc = hip_check(hipStreamSynchronize(st), "batched step sync");
    if (rc)
        return rc;
    for (Uni* u : act) {
        Pack* h = u->h;
        const int* rep = (const int*)(h_rep + REP_BYTES * (size_t)u->slot);
        const int na = rep[0];
        u->S = rep[1];
        u->chunk = (int)std::min<int64_t>(u->S, h->chunk_max);
        const int64_t rech = h->cfg.precision == UG_PREC_SPLIT_BF16 && !u->all ? rep[4] : 0;
        h->rechecked += rech;
        u->rech_sub += rech;
        if (u->fresh) {
            u->S_first = u->S;
            h->survivors += u->S;
        } else if (h->sub_from_empty) {
            // the sub-batch met an empty table, so its list was the whole range: this retest
            // of its last npend candidates is the table's first survival-rate observation
            pack_note_rate(h, h->attempts + u->sub - u->npend, u->npend, u->S);
            h->sub_from_empty = false;
        }
        h->filled = *(const unsigned long long*)(rep + 2);
        const int64_t* ai = (const int64_t*)(rep + 6);
        h->acc_host.insert(h->acc_host.end(), ai, ai + na);
        h->n += na;
    }
    return UG_OK;
}
#undef UG_STAGE

}  // namespace

namespace {

// One lane of the batched engine: its own stream and staging, universes admitted from the
// shared queue up to `window` in flight, one step (filter + resolve, one synchronisation) at
// a time.
struct Lane {
    Engine E;
    std::vector<int> other;  // configs outside the batched kernels' cover, packed at the end
};

struct Shared {
    const ug_pack_config* cfgs;
    int n;
    ug_pack** keep;
    ug_pack_status* status;
    uint64_t* digests;
    std::vector<Uni> unis;
    std::atomic<int> next{0};
    std::atomic<bool> stop{false};
    std::mutex mu;
    int32_t err = UG_OK;
    std::string msg;
    void record(int32_t rc, const std::string& m) {
        std::lock_guard<std::mutex> g(mu);
        if (err == UG_OK) {
            err = rc;
            msg = m;
        }
        stop = true;
    }
};

void retire(Shared& S, Uni* u) {
    Pack* h = u->h;
    h->done = true;
    h->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - u->t0).count();
    pack_fill_status(h, &S.status[u->idx]);
    if (S.digests)
        S.digests[u->idx] = list_digest(h->acc_host);
    if (S.keep)
        S.keep[u->idx] = (ug_pack*)h;
    else
        destroy_handle(h);
    u->h = nullptr;
}

// Finishes u's sub-batch (no survivors left to resolve): attempts, sizing, curve; true when
// max_n ended the run.
bool end_sub(Uni* u) {
    Pack* h = u->h;
    if (h->max_n > 0 && h->n >= h->max_n) {
        h->attempts = h->acc_host[(size_t)h->max_n - 1] + 1;
        return true;
    }
    if (!h->sub_from_empty)
        pack_note_rate(h, h->attempts, u->sub, u->S_first);
    h->sub_from_empty = false;
    h->attempts += u->sub;
    pack_size_next(h, u->sub, u->S_first, u->rech_sub);
    pack_record_curve(h);
    return false;
}

int32_t run_lane(Shared& S, Lane& L) {
    Engine& E = L.E;
    std::vector<Uni*> act;
    int32_t rc = UG_OK;
    while (!rc && !S.stop) {
        while ((int)act.size() < E.cap && !rc) {
            const int i = S.next.fetch_add(1);
            if (i >= S.n)
                break;
            if (!batchable(S.cfgs[i])) {
                L.other.push_back(i);
                continue;
            }
            Uni& u = S.unis[(size_t)i];
            u.idx = i;
            u.t0 = std::chrono::steady_clock::now();
            rc = create_handle(&S.cfgs[i], (void*)E.st, true, &u.h);
            if (rc) {
                rc = fail(rc, "universe %d: %s", i, ug_last_error());
                break;
            }
            act.push_back(&u);
        }
        if (rc || act.empty())
            break;
        // retire finished universes; the rest pick this step's work
        std::vector<Uni*> live;
        for (Uni* u : act) {
            if (u->fresh && finished(u->h)) {
                retire(S, u);
                continue;
            }
            if (u->fresh) {
                u->sub = size_at(u->h);
                u->S_first = 0;
                u->rech_sub = 0;
                u->h->sub_from_empty = u->h->n == 0;
            }
            live.push_back(u);
        }
        act.swap(live);
        if (act.empty())
            continue;
        rc = E.step(act);
        if (rc)
            break;
        std::vector<Uni*> live2;
        for (Uni* u : act) {
            Pack* h = u->h;
            const bool maxed = h->max_n > 0 && h->n >= h->max_n;
            if (u->S > u->chunk && !maxed) {
                // retest the rest against the updated table in the next step
                u->pend = u->list + u->chunk;
                u->npend = u->S - u->chunk;
                u->fresh = false;
            } else {
                u->fresh = true;
                u->npend = 0;
                if (end_sub(u))
                    h->done = true;
            }
            if (h->done)
                retire(S, u);  // max_n reached
            else
                live2.push_back(u);
        }
        act.swap(live2);
    }
    // configs the batched kernels do not cover run the single-universe loop on this lane
    for (size_t k = 0; k < L.other.size() && !rc && !S.stop; k++) {
        const int i = L.other[k];
        Pack* h = nullptr;
        rc = create_handle(&S.cfgs[i], (void*)E.st, true, &h);
        if (!rc)
            rc = ug_pack_run((ug_pack*)h, 0, &S.status[i]);
        if (rc) {
            rc = fail(rc, "universe %d: %s", i, ug_last_error());
            if (h)
                destroy_handle(h);
            break;
        }
        if (S.digests)
            S.digests[i] = list_digest(h->acc_host);
        if (S.keep)
            S.keep[i] = (ug_pack*)h;
        else
            destroy_handle(h);
    }
    if (rc)
        S.record(rc, ug_last_error());
    (void)hipStreamSynchronize(E.st);
    for (Uni* u : act)
        if (u->h) {
            destroy_handle(u->h);
            u->h = nullptr;
        }
    return rc;
}

}  // namespace

extern "C" int32_t ug_ensemble_run_batched(const ug_pack_config* cfgs, int32_t n, int32_t window, int32_t lanes,
                                           ug_pack** keep, ug_pack_status* status, uint64_t* digests) {
    if (!cfgs || !status || n < 0)
        return fail(UG_EINVAL, "ug_ensemble_run_batched: NULL argument or negative count");
    if (window < 0 || window > 64)
        return fail(UG_EINVAL, "ug_ensemble_run_batched: window %d outside [0, 64]", window);
    if (lanes < 0 || lanes > 32)
        return fail(UG_EINVAL, "ug_ensemble_run_batched: %d lanes outside [0, 32]", lanes);
    if (n == 0)
        return UG_OK;
    if (lanes == 0)
        lanes = 1;
    for (int i = 0; keep && i < n; i++)
        keep[i] = nullptr;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
        return fail(UG_EHIP, "ug_ensemble_run_batched: no current device");
    Shared S;
    S.cfgs = cfgs;
    S.n = n;
    S.keep = keep;
    S.status = status;
    S.digests = digests;
    S.unis.resize((size_t)n);
    std::vector<Lane> L((size_t)lanes);
    auto lane_main = [&](int k) {
        Engine& E = L[(size_t)k].E;
        if (hipSetDevice(device) != hipSuccess) {
            S.record(UG_EHIP, "ug_ensemble_run_batched: hipSetDevice failed in a lane");
            return;
        }
        E.cap = window == 0 ? 32 : window;
        E.stage.cap = (size_t)E.cap * 4096 + (1 << 20);
        int32_t rc = hip_check(hipStreamCreate(&E.st), "batched engine stream");
        if (!rc && !(E.stage.host = (char*)cache_host_alloc(E.stage.cap)))
            rc = fail(UG_ENOMEM, "batched engine: pinned staging");
        if (!rc && !(E.stage.dev = (char*)cache_alloc(E.stage.cap)))
            rc = fail(UG_ENOMEM, "batched engine: device staging");
        if (!rc && !(E.h_rep = (char*)cache_host_alloc(REP_BYTES * (size_t)E.cap)))
            rc = fail(UG_ENOMEM, "batched engine: pinned reports");
        if (rc)
            S.record(rc, ug_last_error());
        else
            (void)run_lane(S, L[(size_t)k]);
        if (E.st)
            (void)hipStreamSynchronize(E.st);
        cache_host_free(E.stage.host);
        cache_free(E.stage.dev);
        cache_host_free(E.h_rep);
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < lanes; k++)
        pool.emplace_back(lane_main, k);
    lane_main(0);
    for (auto& th : pool)
        th.join();
    int32_t rc = S.err;
    // kept handles outlive the lanes' streams: each gets a stream of its own
    for (int i = 0; keep && i < n && rc == UG_OK; i++) {
        Pack* h = (Pack*)keep[i];
        if (!h)
            continue;
        hipStream_t s = nullptr;
        if (hipStreamCreate(&s) != hipSuccess) {
            rc = fail(UG_EHIP, "ug_ensemble_run_batched: hipStreamCreate failed");
            break;
        }
        h->stream = s;
        h->own_stream = true;
    }
    for (Lane& l : L)
        if (l.E.st)
            (void)hipStreamDestroy(l.E.st);
    if (rc != UG_OK) {
        // the universes finished before the failure are not handed out either
        for (int i = 0; keep && i < n; i++)
            if (keep[i]) {
                destroy_handle((Pack*)keep[i]);
                keep[i] = nullptr;
            }
        return fail(rc, "%s", S.msg.empty() ? ug_last_error() : S.msg.c_str());
    }
    return UG_OK;
}
