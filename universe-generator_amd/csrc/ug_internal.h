// ug_internal.h -- shared types of the packer's translation units (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <deque>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/ugpack.h"
#include "ug_arith.h"

#define UG_MAX_T 512      // basis pairs per lane per axis <= 4 (T <= 512)
#define MAXPL 4
#define UG_S_MAX 2048     // survivors resolved per in-batch chunk
#define UG_MASK_WORDS (UG_S_MAX / 64)
#define UG_SUB_MAX (1 << 20)  // largest candidate sub-batch
// Measurement counters: a wave adds into slot blockIdx % UG_CTR_SLOTS (UG_CTR_STRIDE words
// apart), so thousands of waves ending together do not serialise on one address.
#define UG_CTR_SLOTS 64
#define UG_CTR_STRIDE 16
#define UG_CTR_WORDS (2 * UG_CTR_SLOTS * UG_CTR_STRIDE)

namespace ug {

// Everything a kernel needs about the model and the packing, by value.
struct DevModel {
    int T, nw, npairs, modmax, pin, full;
    uint32_t key0, key1;
    double r, halfT;
    int euclid;       // contact predicate: 0 = Chebyshev cube |d_a| <= r, 1 = L2 ball
    double r2;        // r * r (L2 ball)
    int ldb;          // basis row length (T + 1)
    int rowlen;       // cell-table row length along the last axis (T + 2: one halo cell each side)
    const double* S;  // [T][T+1]  sin(k z_t)/sin z_t
    const double* C;  // [T][T+1]  (cos(k z_t) - 1)/sin z_t
    const double* ST; // [T+1][T]  S transposed (lanes over timesteps read contiguously)
    const double* CT; // [T+1][T]  C transposed
    const int* order; // [T] endpoint-first scan order
    int* tab;         // cell table: head path of each (t, cell) chain, -1 = empty
    uint32_t* occ;    // occupancy bits of the same padded rows (bit i <-> head i), see occ_index
    int occw;         // 2+1: words a row (one spare: a 3-bit window may cross a word)
    int occ_nwq, occ_nxb, occ_nyb;  // 3+1 tiles: words along a padded row, x tiles of 4, y tiles of 8
    const double* pos64;  // [n][T][D] canonical positions of accepted paths
    const float4* pos32;  // [n][T] (x, y, w, next path in the same cell as int bits)
    // split-BF16 form
    const float* marg;    // [T] certainty margin per timestep
    // sparse occupied-cell grid (--sparse): the occupancy bits as above, and an
    // open-addressing hash of the occupied cells' heads in place of tab; slot = {key + 1
    // (0 = empty), head path in the low 32 bits}
    int sparse;
    ulonglong2* hslots;
    uint64_t hmask;       // slots - 1 (a power of two)
    // candidate stream: 0 = Philox (key0, key1), 1 = the reference's xoshiro256** stream
    // seeded from key0 | key1 << 32 (FP64 forms only)
    int xref;
};

// Canonical FP64 contact predicate between two points at one timestep (torus min-image
// per axis).  Cube: every |d_a| <= r.  Ball: d0*d0, then fma(d1, d1, .), fma(d2, d2, .)
// <= r*r, the order fixed so host and device round alike.
template <int D>
__host__ __device__ inline bool contact(const DevModel& m, const double* x, const double* p) {
    double dd[D];
#pragma unroll
    for (int a = 0; a < D; a++)
        dd[a] = torus_delta(x[a] - p[a]);
    if (m.euclid) {
        double q = dd[0] * dd[0];
#pragma unroll
        for (int a = 1; a < D; a++)
            q = fma(dd[a], dd[a], q);
        return q <= m.r2;
    }
    bool h = true;
#pragma unroll
    for (int a = 0; a < D; a++)
        h = h && fabs(dd[a]) <= m.r;
    return h;
}

// Cell index c + d (d in {-1, 0, 1}) wrapped onto [0, T).
__host__ __device__ inline int wrap_cell(int c, int d, int T) {
    const int v = c + d;
    return v < 0 ? v + T : (v >= T ? v - T : v);
}

// Word wq (padded bits 32 wq .. 32 wq + 31) of the occupancy row of cells (t, x[, y], *).
// 2+1: rows [t][x] of occw words, each row contiguous.  3+1: per timestep and word plane wq,
// tiles of 4 x 8 rows ((x mod 4, y mod 8), one word each: 128 bytes, one cache line), so the
// 3 x 3 rows around a cell usually lie in one or two lines instead of three rows a whole
// x-plane apart -- the lookup of the 17 GB table at 3+1 T=256 then fetches about 2 lines of
// bits per visited timestep instead of about 5.  The bits are a 32nd of the head table (42 MB
// at 3+1 T=128, inside the last-level cache), so empty neighbour cells cost no head traffic.
// (The layout follows the dimension: D = 2 rows, D = 3 tiles.)
template <int D>
__host__ __device__ __forceinline__ int64_t occ_index(const DevModel& m, int t, int x, int y, int wq) {
    // below 2^32 words for T <= 512 (3+1 T=512: 2.3G), so 32-bit unsigned math
    if constexpr (D == 2)
        return (int64_t)(((uint32_t)t * (uint32_t)m.T + (uint32_t)x) * (uint32_t)m.occw + (uint32_t)wq);
    return (int64_t)((((((uint32_t)t * (uint32_t)m.occ_nwq + (uint32_t)wq) * (uint32_t)m.occ_nxb + (uint32_t)(x >> 2)) *
                          (uint32_t)m.occ_nyb +
                      (uint32_t)(y >> 3)) *
                         32u +
                     (uint32_t)((x & 3) * 8 + (y & 7))));
}

// Occupancy of the padded cells c..c+2 (cells c-1..c+1 of the unpadded row, wrapped by the
// halo bits at the row ends) of row (t, x[, y]) as bits 0..2.  3+1 reads the next word plane
// only where the window crosses into it (it exists then: the padded row has T + 2 bits); a
// 2+1 row's next word (or its spare word) is always read, it shares the row's line.
template <int D>
__device__ __forceinline__ uint32_t occ_bits3(const DevModel& m, int t, int x, int y, int c) {
    const int wq = c >> 5, sh = c & 31;
    const uint64_t lo = m.occ[occ_index<D>(m, t, x, y, wq)];
    uint64_t hi = 0;
    if constexpr (D == 2)
        hi = m.occ[occ_index<D>(m, t, x, y, wq + 1)];  // the row's next word, in the same line
    else if (sh > 29)
        hi = m.occ[occ_index<D>(m, t, x, y, wq + 1)];
    return (uint32_t)(((hi << 32) | lo) >> sh) & 7u;
}

// Sparse grid: the key of cell c (0 <= c < T) of table row `row`, and the slot probe sequence
// (linear from a 64-bit finaliser of the key).
__host__ __device__ __forceinline__ uint64_t cell_hkey(int64_t row, int cell, int T) {
    return (uint64_t)row * (uint64_t)T + (uint64_t)cell;
}
__host__ __device__ __forceinline__ uint64_t hmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}
// Head path of an occupied cell, -1 when the cell is empty.
__device__ __forceinline__ int hash_head(const DevModel& m, uint64_t key) {
    uint64_t h = hmix64(key) & m.hmask;
    for (;;) {
        const ulonglong2 s = m.hslots[h];
        if (s.x == key + 1)
            return (int)(uint32_t)s.y;
        if (s.x == 0)
            return -1;
        h = (h + 1) & m.hmask;
    }
}

// Heads of the cells c-1, c, c+1 (wrapped) of table row (t, x[, y]) (row = t*T + x in 2+1,
// (t*T + x)*T + y in 3+1), -1 for empty cells: the occupancy bits first, then the occupied
// cells' heads from the dense table's padded row or from the sparse grid.
template <int D>
__device__ __forceinline__ void row_heads3(const DevModel& m, int t, int x, int y, int64_t row, int c, int node[3]) {
    const uint32_t bits = occ_bits3<D>(m, t, x, y, c);
    if (m.sparse) {
#pragma unroll
        for (int k = 0; k < 3; k++)
            node[k] = ((bits >> k) & 1u) ? hash_head(m, cell_hkey(row, wrap_cell(c, k - 1, m.T), m.T)) : -1;
        return;
    }
    const int* tr = m.tab + row * m.rowlen + c;  // padded cells c-1..c+1 at c..c+2
#pragma unroll
    for (int k = 0; k < 3; k++)
        node[k] = ((bits >> k) & 1u) ? tr[k] : -1;
}

// Work items of a batched launch over at most 64 universes: universe u owns the flattened
// items [excl(u), excl(u) + n(u)).  Each wave scans the counts once (lane u holds universe u)
// and then finds the owner of an item with one ballot: the universes whose items all lie
// before it are a prefix of the lanes, because the running sums never decrease.
struct BatchItems {
    int64_t incl, excl, total;
    int U;
    __device__ __forceinline__ void scan(int64_t n_lane, int nU) {
        const int lane = threadIdx.x & 63;
        U = nU;
        const int64_t c = lane < nU ? max(n_lane, (int64_t)0) : 0;
        incl = c;
        for (int s = 1; s < 64; s <<= 1) {
            const int64_t v = __shfl_up(incl, s, 64);
            if (lane >= s)
                incl += v;
        }
        excl = incl - c;
        total = __shfl(incl, 63, 64);
    }
    // universe of item (uniform over the wave), and the item's index within it
    __device__ __forceinline__ int owner(int64_t item, int64_t* i) const {
        const int lane = threadIdx.x & 63;
        const int u = __popcll(__ballot(lane < U && incl <= item));
        *i = item - __shfl(excl, u, 64);
        return u;
    }
};

// Batched launches read each universe's record (pointers included) from device memory; a
// pointer loaded as a plain pointer is generic to the compiler and every access through it
// becomes a flat access.  ug_gptr loads a pointer field of a record in device memory as a
// global-address-space pointer, so accesses through it are global loads and stores, as for
// pointers passed as kernel arguments.  Every pointer in these records is device memory (or
// pinned host memory, which lies in the same aperture).
#define UG_GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T* ug_gptr(T* const* field) {
    return (T*)*(T UG_GAS* const*)(const void*)field;
}
// A DevModel from one in device memory, its pointers global.
__device__ __forceinline__ DevModel load_model(const DevModel* src) {
    DevModel m = *src;
    m.S = ug_gptr(&src->S);
    m.C = ug_gptr(&src->C);
    m.ST = ug_gptr(&src->ST);
    m.CT = ug_gptr(&src->CT);
    m.order = ug_gptr(&src->order);
    m.tab = ug_gptr(&src->tab);
    m.occ = ug_gptr(&src->occ);
    m.pos64 = ug_gptr(&src->pos64);
    m.pos32 = ug_gptr(&src->pos32);
    m.marg = ug_gptr(&src->marg);
    m.hslots = ug_gptr(&src->hslots);
    return m;
}

// ---- batched ensemble engine (ug_batch.hip): one universe's share of a launch that serves
// many universes; the launches read these from a device array indexed by universe.
// k_cand64 work: fresh candidates base + i or list[i], i < min(count, *dcount).
struct MCand {
    DevModel m;
    uint64_t base;
    int64_t count;
    const int64_t* list;
    const int* dcount;
    uint8_t* flags;
    double* out;     // trajectories (full)
    int* cells;      // their packed cells
    uint64_t* mask;  // conflict rows cleared with them
    const uint32_t* skip;  // per listed candidate: scan positions < 32 proven clear (or null)
    int head;
};
// Ordered compaction of flags[0 .. min(count, *dcount)) into out (count into *total); bcount
// holds the per-block counts (ceil(count / 4096) of them).
struct MCompact {
    const uint8_t* flags;
    int64_t count;
    const int* dcount;
    uint64_t base;
    const int64_t* list;
    int* bcount;
    int64_t* out;
    int* total;
    const int* vin;  // with vout: a value per item, compacted alongside (or null)
    int* vout;
};
// A resolve chunk of min(S, counts[1]) survivors (trajectories pos, packed cells, cleared
// conflict rows; counts = the filter's device counts {rechecks, survivors}): conflict bits,
// index-order greedy,
// insertion of at most limit accepted paths as rows n0, ..., and the report: rep[0] =
// accepted, rep[1] = survivors, rep[2..3] = occupied buckets (64 bits), rep[4] = rechecks,
// then the accepted candidate indices from rep + 6 (as int64).
struct MResolve {
    DevModel m;
    const double* pos;
    const int* cells;
    uint64_t* mask;
    const int* counts;
    int S, limit;
    int* acc_list;
    int* nacc;
    const int64_t* surv_idx;
    int64_t n0;
    int64_t* acc_idx;
    int32_t* acc_gid;
    double* pos64;
    float4* pos32;
    unsigned long long* filled;
    int* rep;
};

// Read-only tables of one model, built once per (device, dim, T, terms, form, margin) and
// shared by its handles (ug_cache.hip).
struct ModelTables {
    int device = 0, d = 0, T = 0, nw = 0;
    bool split = false;
    double *d_S = nullptr, *d_C = nullptr, *d_ST = nullptr, *d_CT = nullptr;
    int* d_order = nullptr;
    // split-BF16 form of full spectra (ug_fast.hip)
    void* fast = nullptr;     // basis fragments and launch shape
    float* d_marg = nullptr;  // certainty margin per examined column
    double margin = 0.0;      // its maximum
};

struct Pack {
    ug_pack_config cfg;
    std::shared_ptr<const ModelTables> tabs;
    hipStream_t stream;
    bool own_stream = false;  // the handle destroys its stream (ensemble handles kept for the caller)
    int device;
    int d, T, nw, npairs, modmax;
    bool full;
    DevModel dm;
    // tables (d_S .. d_order, d_marg: views of tabs)
    double *d_S = nullptr, *d_C = nullptr, *d_ST = nullptr, *d_CT = nullptr;
    int* d_order = nullptr;
    int* d_tab = nullptr;
    int64_t tab_len = 0;
    uint32_t* d_occ = nullptr;
    int64_t occ_len = 0;
    float* d_marg = nullptr;
    double margin = 0.0;
    // accepted paths
    double* d_pos64 = nullptr;
    float4* d_pos32 = nullptr;
    int64_t* d_acc_idx = nullptr;
    int32_t* d_acc_gid = nullptr;
    unsigned long long* d_filled = nullptr;
    int64_t cap = 0, n = 0, nsub = 0;
    unsigned long long filled = 0;  // occupied (t, cell) buckets, as of the last insert
    // batch scratch
    uint8_t* d_flags = nullptr;
    int64_t* d_list = nullptr;
    int64_t* d_list2 = nullptr;
    int64_t* d_list3 = nullptr;
    int64_t* d_list4 = nullptr;
    int64_t* d_listb = nullptr;     // survivors of a filter run ahead (alternates with d_list)
    uint32_t* d_skip = nullptr;     // per candidate of a prefilter launch: block-0 columns proven clear
    uint32_t* d_skip3 = nullptr;    // the same for the stage-1 survivors (d_list3)
    uint32_t* d_skip4 = nullptr;    // and for the stage-2 survivors (d_list4)
    int* d_spec_count = nullptr;    // that filter's candidate count, set on the device
    hipEvent_t ev_report = nullptr;  // after a resolve chunk's report (timing disabled)
    int recheck_slot = 0;  // d_count slot holding the last filter's FP64 recheck count
    int* d_bcount = nullptr;
    int cmp_blocks_cap = 0;
    int* d_count = nullptr;
    double* d_spos = nullptr;
    int* d_scell = nullptr;  // packed cells of a resolve chunk's survivors [UG_S_MAX][T]
    uint64_t* d_mask = nullptr;
    int* d_acc_list = nullptr;
    int* d_nacc = nullptr;
    // subpath phase
    int* d_gflag = nullptr;   // per-candidate table group of a sub-batch
    int* d_glist = nullptr;   // groups of the survivors list
    int* d_glist2 = nullptr;
    int* d_gsurv = nullptr;     // adopted groups of a resolve chunk
    int* d_rscount = nullptr;   // survivors of a round against its round-start packing
    int sub_done = 0;  // UG_SUB_* once the subpath phase stopped
    int64_t last_S = UG_S_MAX;  // survivors of the last filter whose count stayed on the device
    // survivors per resolve chunk: UG_S_MAX in 3+1, half of it in 2+1 (whose early chunks of
    // candidates from an empty table are cheaper small)
    int chunk_max = UG_S_MAX;
                                // (a fresh table lets a whole first sub-batch survive)
    // sub_batch: the round size, adapted as cuda/braid_cuda.cu:952,1000-1003;
    // sub_nlimit: rows stored before the current round started
    int64_t sub_attempts = 0, sub_batch = 1 << 16, sub_nlimit = 0;
    std::deque<std::pair<int64_t, int64_t>> awin, fwin;
    int64_t a_att = 0, a_adm = 0, f_att = 0, f_new = 0;
    void* h_pinned = nullptr;
    // split-BF16 scratch
    void* d_basis16 = nullptr;
    // host state
    std::vector<int64_t> acc_host;
    std::vector<std::array<int64_t, 4>> curve;
    int64_t next_ms = 1000;
    int64_t attempts = 0, max_n = 0, sub = 4096;
    // sub-batch sizing: the last survival-rate observation (rate_p survivors per candidate
    // filtered against the table as it stood at candidate rate_at; rate_at 0 = none), the
    // size the last sizing aimed at, and the current sub-batch's start and whether the table
    // was empty there
    int64_t rate_at = 0;
    double rate_p = 0.0, sub_want = 0.0;
    int64_t sub_base = 0;
    bool sub_from_empty = false;
    bool done = false;
    int64_t survivors = 0, rechecked = 0;
    double seconds = 0.0;
    // measurement (ug_pack_profile)
    // Profiled launches carry start/stop events of their own (one pair per launch, read back
    // only when the stats are asked for, so profiling adds no host wait inside a pack).
    bool profile = false;
    int profile_mask = 0;  // kernels whose launches are timed (bit UG_KERNEL_*)
    std::vector<std::array<hipEvent_t, 2>> ev_pool;  // event pairs, reused after a collect
    struct TimedLaunch {
        int which;
        int pair;
        int64_t items;
    };
    std::vector<TimedLaunch> ev_pending;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // the armed pair
    bool timing_armed = false;                 // the next UG_LAUNCH carries ev0/ev1
    bool timing_fired = false;                 // that launch happened
    // cumulative device counters per kernel kind: UG_CTR_SLOTS slots of (visits, points,
    // items of the recheck), one 128-byte line each
    unsigned long long* d_ctr = nullptr;
    ug_kernel_stats kstats[2] = {};
    std::deque<std::pair<int64_t, int64_t>> win;
    int64_t win_att = 0, win_adm = 0;
    // reference round schedule (UG_STOP_REF_WINDOW): the engines' adaptive batch, the 3+1
    // engine's already-sized next round (-1 before the first round, 0 = none) and the
    // candidates it has sized, and whether its window fired (the sized round still runs)
    int64_t ref_batch = 1 << 16, ref_next = -1, ref_enq = 0;
    bool ref_stop = false;
    int64_t* d_rlist = nullptr;  // a round's survivors (all pieces of the round)
    int64_t rlist_cap = 0;
    std::vector<std::array<int64_t, 5>> rounds;  // {start, size, survivors, admitted, fresh}
    int64_t undecided_from = -1;  // max_n reached mid-round: candidates from here are untested
    // sparse occupied-cell grid
    ulonglong2* d_hash = nullptr;
    uint64_t hash_cap = 0;
};

int32_t fail(int32_t code, const char* fmt, ...);
// recycled memory and shared model tables (ug_cache.hip)
void* cache_alloc(size_t bytes);  // device memory of the current device; nullptr on failure
void cache_free(void* p);
void* cache_host_alloc(size_t bytes);  // pinned host memory
void cache_host_free(void* p);
int32_t model_tables(int d, int T, int nw, bool split_full, double margin_req,
                     std::shared_ptr<const ModelTables>* out);
int32_t hip_check(hipError_t e, const char* what);
int grid_for(int64_t items, int per_block);
// FP64 candidate kernel: the exact test (full = false: flags) or the trajectories of a
// resolve chunk (full = true: out, and with cells / mask also their packed cells and cleared
// conflict rows).
int32_t launch_cand64(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags,
                      double* out, bool full, unsigned long long* ctr = nullptr,
                      const int* dcount = nullptr, int head = -1, int* cells = nullptr,
                      uint64_t* mask = nullptr, const uint32_t* skip = nullptr);
// Profiled timing of one launch: timed_begin arms the next UG_LAUNCH on h's stream, which
// then records an event pair directly before and after the dispatch (timing-only events
// without the system-scope fence, never waited on inside a pack); timed_end books the pair
// for timed_collect.
int32_t timed_begin(Pack* h, int which);
int32_t timed_end(Pack* h, int which, int64_t items);
int32_t timed_collect(Pack* h);  // waits for pending timed launches and folds them into kstats
void release_events(Pack* h);    // hands h's event pairs to the process-wide free list
#define UG_LAUNCH(h, K, grid, block, lds, ...)                                                 \
    do {                                                                                     \
        if ((h)->timing_armed) {                                                             \
            (h)->timing_armed = false;                                                       \
            (h)->timing_fired = true;                                                        \
            (void)hipEventRecord((h)->ev0, (h)->stream);                                     \
            hipLaunchKernelGGL(K, grid, block, lds, (h)->stream, __VA_ARGS__);               \
            (void)hipEventRecord((h)->ev1, (h)->stream);                                     \
        } else {                                                                             \
            hipLaunchKernelGGL(K, grid, block, lds, (h)->stream, __VA_ARGS__);               \
        }                                                                                    \
    } while (0)
int32_t compact(Pack* h, const uint8_t* flags, int64_t count, uint64_t base, const int64_t* list,
                int64_t* out, int64_t* S, const int* dcount = nullptr, int slot = 1, bool sync = true,
                const int* vin = nullptr, int* vout = nullptr);
// subpath phase (ug_pack.hip)
// Baker coefficients of every stored path (uniques, then subpaths) from its candidate index:
// a2[p][d], A/B[p][d][kmax+1] (A_k = a_k cos f_k, B_k = a_k sin f_k), largest frequency into
// *d_bmax; enqueued on `stream` after the handle's own work.
int32_t launch_paths_from_pack(Pack* h, hipStream_t stream, double* a2, double* A, double* B, int kmax,
                               int* d_bmax);
// Serial-RSA characterization checks (k_check): mode 0 over the stored uniques, mode 1 over
// candidates [first, first + count); counters into d_out[4].
int32_t launch_check(Pack* h, int mode, uint64_t first, int64_t count, unsigned long long* d_out);
int32_t sub_filter(Pack* h, uint64_t base, int64_t count, const int64_t* list, int64_t* out, int* gout,
                   int64_t* S, int* rs_count);
int32_t resolve_sub(Pack* h, int64_t* list, int* gl, int64_t S, int64_t* acc_total);
int32_t filter_candidates(Pack* h, uint64_t base, int64_t count, const int64_t* list, int64_t* out,
                          int64_t* S, bool sync = true, const int* dfresh = nullptr);
int32_t ensure_paths(Pack* h, int64_t need);
// Sparse grid: room for `points` (path, timestep) entries at load <= 1/2; a larger table is
// rebuilt from the stored paths.  No-op for the dense table.
int32_t ensure_hash(Pack* h, int64_t points);
int32_t all_survive(Pack* h, uint64_t base, int64_t count, int64_t* out);
// A filter run ahead: fresh candidates [base, base + count) into out, enqueued behind the
// current sub-batch's first resolve chunk; launched stays true only if no later chunk of the
// current sub-batch changed the table first (then its survivors in out and the counts in
// d_count are the ones a filter run now would give).
// The device sets the filter's count (*dcount) from the chunk's accepted count: 0 when the
// stop rule fires after the current sub-batch (nacc < stop_lim) or max_n is reached
// (nacc >= n_room), so a filter nobody will use costs a few empty launches.
struct SpecFilter {
    uint64_t base = 0;
    int64_t count = 0;
    int64_t* out = nullptr;
    int* dcount = nullptr;
    int64_t stop_lim = 0;
    int64_t n_room = INT64_MAX;
    bool launched = false;
};
int32_t resolve_and_insert(Pack* h, int64_t* list, int64_t S, int64_t* acc_total, const int* dS = nullptr,
                           SpecFilter* spec = nullptr);
// split-BF16 prefilter (ug_fast.hip)
int32_t fast_build(ModelTables& t, const std::vector<double>& S, const std::vector<double>& C,
                   const std::vector<int>& order, double margin_req);
int32_t fast_setup(Pack* h);
// Split-BF16 certain-contact test of fresh candidates base + i (list == nullptr) or of
// list[i], i < count (bounded by *dcount when given), over the N blocks [cb0, cb1) of 32
// ordered timesteps: flags[i] = 0 when candidate i certainly collides.  With skip (first
// stage, cb0 = 0, sequential blocks), skip[i] of each candidate left to the FP64 test = the
// block-0 columns (bit c: scan position c) proven to hold no FP64 contact.
int32_t launch_fast(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags,
                    const int* dcount = nullptr, int cb0 = 0, int cb1 = 1, uint32_t* skip = nullptr);
int fast_blocks(const Pack* h);  // N blocks of the model's ordered timesteps (0: no split form)
// Subpath-phase prefilter: flags[i] = 0 when candidate base + i certainly touches two groups
// among the rows < nlimit.
int32_t launch_fast_sub(Pack* h, uint64_t base, int64_t count, uint8_t* flags, int64_t nlimit);
int32_t launch_fast_eval(Pack* h, const int64_t* list, int64_t count, double* out);
// Knot regime: flags[i] = 0 when candidate base + i certainly touches one of the nk strands
// (FP32 positions kpos32[t][ks]) in the first 32 ordered timesteps.
int32_t launch_fast_knot(Pack* h, uint64_t base, int64_t count, uint8_t* flags, const float4* kpos32, int64_t ks,
                         int nk);
// Conflict bits among a resolve chunk's s survivors (trajectories in d_spos, packed cells in
// d_scell, mask rows cleared); dS: count bounded on the device; expect: survivors expected.
void launch_conflicts(Pack* h, int s, const int* dS = nullptr);
// Index-order greedy over the chunk's conflict bits -> d_acc_list, d_nacc.
int32_t launch_greedy(Pack* h, int s);
// Knot regime: brute-force test of fresh candidates base + i (list == nullptr) or list[i]
// against the nk strands (kcell[t][ks], kpos[t][ks][D]): flags[i] = 1 iff clear.
int32_t launch_knot_test(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags,
                         const int* dcount, const int* kcell, const double* kpos, int64_t ks, int nk);
// Appends the chunk's accepted survivors (d_acc_list, at most limit) of list as strands n0, ...
int32_t launch_knot_append(Pack* h, const int64_t* list, int s, int64_t n0, int* kcell, double* kpos, float4* kpos32,
                           int64_t ks, int64_t* kidx, int limit);
// Batched engine (ug_batch.hip): launches that serve many universes at once (at most 64 for
// the candidate-test kernels), per-universe arguments in device arrays (MCand / MCompact /
// MResolve, and the prefilter's opaque records of fast_multi_arg_bytes() bytes written by
// fast_multi_arg).  items / tiles: an upper bound on the launch's work (sizes its grid).
size_t cand64_lds_bytes(const Pack* h);
int32_t launch_cand64_multi(hipStream_t st, int d, int pl, bool full, bool sb, const MCand* dA, int U,
                            int64_t items, size_t lds);
int32_t launch_compact_multi(hipStream_t st, const MCompact* dA, int U, int64_t max_count);
int32_t launch_resolve_multi(hipStream_t st, int d, const MResolve* dA, int U, int maxS, int maxT);
size_t fast_multi_arg_bytes();
int32_t fast_multi_arg(Pack* h, uint64_t base, int64_t count, const int64_t* list, uint8_t* flags, uint32_t* skip,
                       void* out);
int32_t launch_fast_multi(hipStream_t st, int d, const void* dA, int U, int64_t tiles);
int slot_pairs_per_lane(int npairs);
// Pieces of the single-universe loop (ug_api.hip) that the batched engine shares.
void pack_fill_status(const Pack* h, ug_pack_status* st);
bool pack_fixed_window_stop(const Pack* h);  // the fixed-window rule fires at h->attempts
void pack_record_curve(Pack* h);             // log-spaced curve rows up to h->attempts
// Sub-batch sizing shared by ug_pack_run and the batched engine (results do not depend on it):
// the size of the sub-batch starting at `from` (call_end: the call's last candidate), a
// survival-rate observation, and the next size after a sub-batch of `sub` candidates with S
// survivors of its fresh filter and `rech` FP64 rechecks.
int64_t pack_size_at(const Pack* h, int64_t from, int64_t call_end);
void pack_note_rate(Pack* h, int64_t at, int64_t count, int64_t S);
void pack_size_next(Pack* h, int64_t sub, int64_t S, int64_t rech);
uint64_t list_digest(const std::vector<int64_t>& idx);  // FNV-1a of an accepted list (ug_ensemble.hip)
// Handle set-up shared by the packer and the knot engine: with_table = false builds no cell
// table, occupancy bits or path records (the knot engine keeps its own strands).
int32_t create_handle(const ug_pack_config* cfg, void* hip_stream, bool with_table, Pack** out);
void destroy_handle(Pack* h);
// Dump parameters of candidate idx: per axis a1 then nw (amplitude, frequency, phase).
void gen_params_host(const Pack* h, uint64_t idx, double* out_row);

}  // namespace ug
