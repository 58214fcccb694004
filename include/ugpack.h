/*
 * ugpack.h -- C ABI of the MI355X RSA worldline packer (libugpack.so).
 *
 * The reference exposes this path as a process boundary, not an FFI: the engine
 * binaries cuda/braid_cuda (2+1) and cuda/braid_cuda3d (3+1) take the argv that
 * braidlab/engine.py:50-101 (build_command) assembles, print the
 * "done: N=<n> in <attempts> attempts" line that braidlab/engine.py:104-112
 * (parse_done) reads, and write the curve / parameter-dump CSVs
 * (cuda/braid_cuda.cu:1045-1104).  The functions below are what a binding of that
 * path binds instead: one handle per packing run, all device work on the caller's
 * HIP stream, plain pointers and sizes, no exit() on error.
 *
 * Conventions
 *   - Return 0 on success, a negative UG_E* code on failure; ug_last_error() holds
 *     the message for the calling thread.
 *   - A handle is used by one host thread at a time; handles are independent.
 *   - Host output buffers are caller-owned; a cap of 0 with a non-NULL count
 *     pointer is a size query.  Device outputs (dev_*) are caller-owned device
 *     memory on the handle's device.
 */
#ifndef UGPACK_H
#define UGPACK_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UG_ABI_VERSION 5

enum {
    UG_OK = 0,
    UG_EINVAL = -1,    /* invalid argument or configuration (engine "error:" exits) */
    UG_ENOMEM = -2,    /* device or host allocation failed */
    UG_EHIP = -3,      /* HIP runtime error */
    UG_EOVERFLOW = -4, /* a fixed capacity was exceeded (never silently truncated) */
    UG_ESTATE = -5     /* call not valid in the handle's current state */
};

/* Precision of the prefilter.  Both give the identical accepted set. */
enum {
    UG_PREC_FP64 = 0,      /* plain all-FP64 GPU form */
    UG_PREC_SPLIT_BF16 = 1 /* split-BF16 MFMA evaluation + FP64 recheck near the radius */
};

/* Candidate stream (ug_pack_config.rng_kind, ug_knot_config.rng_kind). */
enum {
    UG_RNG_PHILOX = 0,      /* Philox4x32-10 counters per (candidate, stage, axis, slot pair) */
    UG_RNG_XOSHIRO_REF = 1  /* cross-check mode: candidate idx draws the reference's own serial
                               xoshiro256** stream, seeded as cuda/braid_knot.cu:207, in the
                               draw order of its proposal (cuda/braid_knot.cu:112-166); FP64
                               precision only (UG_EINVAL with UG_PREC_SPLIT_BF16) */
};

/* Stop rules. */
enum {
    UG_STOP_BUDGET = 0,       /* candidates [0, max_attempts) */
    UG_STOP_FIXED_WINDOW = 1, /* at each multiple c of stop_batch with c >= stop_window:
                                 stop when accepts in [c-stop_window, c) < stop_rate*stop_window */
    UG_STOP_REF_WINDOW = 2    /* the reference engines' round schedule and --until-accept-rate
                                 window (cuda/braid_cuda.cu:585-590,885-944 for dim 2,
                                 cuda/braid_cuda3d.cu:1446-1566 for dim 3): rounds of 2^16
                                 candidates adapted on their survivors (4096..2^26), the
                                 last round may pass max_attempts; stop_rate 0 = no window */
};

/* One packing run: the flags of cuda/braid_cuda.cu:534-584 / cuda/braid_cuda3d.cu:862-913. */
typedef struct ug_pack_config {
    int32_t dim;         /* 2 = braid_cuda, 3 = braid_cuda3d */
    int32_t T;           /* -t: timesteps; the max frequency is T (--maxfreq must equal T) */
    int32_t terms;       /* --terms incl. sin1; 0 = T (full spectrum) */
    int32_t precision;   /* UG_PREC_* */
    uint64_t seed;       /* --seed: Philox key */
    int32_t pin_sin1;    /* --pin-sin1 */
    int32_t stop_kind;   /* UG_STOP_* */
    double max_attempts; /* --attempts */
    double stop_rate;    /* --until-accept-rate */
    int64_t stop_batch;  /* fixed-window rule: boundary spacing (default 2^20) */
    int64_t stop_window; /* fixed-window rule: window length (default 2^22) */
    int64_t max_n;       /* --max-n: stop after this many uniques (0 = off) */
    double margin;       /* split-BF16 certainty margin; 0 = derived error bound */
    int64_t max_paths;   /* initial path capacity (0 = automatic; grows on demand) */
    int32_t euclid;      /* --euclid-collision: L2 ball sum(d_a^2) <= r^2 instead of the
                            Chebyshev cube (cuda/braid_cuda3d.cu:605-608) */
    int32_t sparse;      /* --sparse: the occupied cells only (O(N T) memory, a hash of
                            (timestep, cell) buckets) instead of the dense T^(d+1) cell table
                            (cuda/braid_cuda3d.cu:82-91,1262-1353); same accepted lists */
    int32_t sub_target;  /* sub-batch sizing (budget and fixed-window rules): survivors aimed
                            at per sub-batch, 0 = 2048 in 3+1 (one resolve chunk), 1024 in
                            2+1; speed only, never results */
    int32_t sub_max;     /* largest sub-batch, a power of two <= 2^20 (0 = 2^20) */
    int32_t rng_kind;    /* UG_RNG_* (0 = Philox) */
} ug_pack_config;

typedef struct ug_pack_status {
    int64_t attempts;  /* candidates consumed (the done: line's attempts) */
    int64_t n;         /* accepted unique worldlines (the done: line's N) */
    int64_t nsub;      /* accepted subpaths */
    int32_t done;      /* 1 once the stop rule (or budget / max_n) has fired */
    int32_t sub_done;  /* subpath phase: 0 running, else its stop rule: UG_SUB_BUDGET,
                          UG_SUB_ACCEPT_RATE or UG_SUB_FILL_RATE */
    int64_t survivors; /* candidates that reached the in-batch accept */
    int64_t rechecked; /* candidates settled by the FP64 recheck (split-BF16 form) */
    double margin;     /* certainty margin in use (split-BF16 form) */
    double seconds;    /* wall time spent inside ug_pack_run / ug_pack_run_sub */
    int64_t filled;    /* occupied (timestep, cell) pairs (the done: line's filled) */
    int64_t sub_attempts; /* candidates consumed by the subpath phase */
} ug_pack_status;

/* ug_pack_status.sub_done: why the subpath phase stopped (cuda/braid_cuda.cu:968,1020,1035). */
#define UG_SUB_BUDGET 1
#define UG_SUB_ACCEPT_RATE 2
#define UG_SUB_FILL_RATE 3

typedef struct ug_pack ug_pack;

/* ABI version of the loaded library (UG_ABI_VERSION). */
int32_t ug_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* ug_last_error(void);

/* Validates cfg like the engine's argv checks (cuda/braid_cuda.cu:597-636), builds the
 * basis tables and the device grid, and binds the handle to hip_stream (NULL = the
 * default stream) on the current device.  Replaces launching braid_cuda[3d]. */
int32_t ug_pack_create(const ug_pack_config* cfg, void* hip_stream, ug_pack** out);

/* Runs the packing loop (cuda/braid_cuda.cu:885-944) until the stop rule fires or
 * max_candidates more candidates have been consumed (0 = no limit).  st may be NULL. */
int32_t ug_pack_run(ug_pack* h, int64_t max_candidates, ug_pack_status* st);

/* Subpath phase flags (cuda/braid_cuda.cu:572-579). */
typedef struct ug_sub_config {
    int32_t sub_terms;   /* --sub-terms (0 = --terms; other values are rejected) */
    int32_t reserved;
    double max_attempts; /* --sub-attempts (0 = the --attempts budget again) */
    double accept_rate;  /* --sub-until-accept-rate (0 = off) */
    double fill_rate;    /* --sub-until-fill-rate (0 = off) */
} ug_sub_config;

/* Subpath phase (cuda/braid_cuda.cu:949-1043, --subpaths): continues the candidate index
 * stream after the unique phase; a candidate joins the single group it touches (contacts
 * with the table and with earlier subpaths, host_sub_gid :726-750), touching none or two
 * groups rejects.  Ends the unique phase.  Stops on the budget or the accept-rate /
 * fill-rate windows (:956-1037), or after max_candidates more candidates (0 = no limit). */
int32_t ug_pack_run_sub(ug_pack* h, const ug_sub_config* cfg, int64_t max_candidates,
                        ug_pack_status* st);

/* Current status without running. */
int32_t ug_pack_get_status(const ug_pack* h, ug_pack_status* st);

/* Accepted candidate indices (admission order) and group ids; replaces the dump's row
 * order and gid column (cuda/braid_cuda.cu:1062-1104).  *n receives the total. */
int32_t ug_pack_accepted(const ug_pack* h, int64_t* idx, int32_t* gid, int64_t cap, int64_t* n);

/* Dump parameters of accepted paths [first, first+count): per path and axis,
 * a1 then nw triples (amplitude, frequency, phase) -- the values the reference
 * prints per row (cuda/braid_cuda.cu:1086-1099).  out holds count*dim*(1+3*nw) doubles. */
int32_t ug_pack_params(const ug_pack* h, int64_t first, int64_t count, double* out);

/* Canonical wrapped comoving positions of all accepted paths, [n][T][dim] FP64, into
 * caller device memory of cap paths; the reference keeps them in px/py
 * (cuda/braid_cuda.cu:685-694). */
int32_t ug_pack_positions(const ug_pack* h, double* dev_out, int64_t cap);

/* Log-spaced growth curve (attempts, n, nsub, filled) as cuda/braid_cuda.cu:925-929. */
int32_t ug_pack_curve(const ug_pack* h, int64_t* rows4, int64_t cap, int64_t* n);

/* Rounds run under UG_STOP_REF_WINDOW, 5 values each: first candidate, candidates, survivors
 * against the table at the round's start (the engines' `got`, cuda/braid_cuda.cu:894-897),
 * uniques admitted, (timestep, cell) buckets first occupied.  The reference keeps these in
 * locals of its loop; exposed so a caller can audit the schedule.  *n receives the total. */
int32_t ug_pack_rounds(const ug_pack* h, int64_t* rows5, int64_t cap, int64_t* n);

/* Positions of arbitrary candidates of the handle's model: dev_idx[count] -> dev_out
 * [count][T][dim] (FP64).  precision UG_PREC_FP64 gives the canonical values
 * (collides_dev's evaluation, cuda/braid_cuda.cu:418-441); UG_PREC_SPLIT_BF16 gives the
 * prefilter's split-BF16 MFMA values (for error reporting). */
int32_t ug_eval_candidates(const ug_pack* h, const int64_t* dev_idx, int64_t count,
                           int32_t precision, double* dev_out);

/* Accuracy of the split-BF16 prefilter's FP32 generation on the current device, measured
 * exhaustively against FP64: out2[0] = the largest absolute error of (cos, sin) of a phase
 * over all 2^32 phase words; out2[1] = the largest |e - e_FP64| / (1 + e_FP64) of an
 * exponential spacing e over all 2^31 Dirichlet words.  With spacing sums W >= nw/4 (smaller
 * ones go to the FP64 test) a coefficient is then within (out2[0] + 5 out2[1]) of the unit
 * slope budget, and the certainty margin allots 2^-20 to that.  Runs on the given stream and
 * synchronizes it.
 * No counterpart in the reference (its FP32 prefilter uses a fixed one-sided margin,
 * cuda/braid_cuda3d.cu:938-948). */
int32_t ug_split_generation_error(void* hip_stream, double* out2);

/* Per-kernel measurement of the candidate-test kernels.  When enabled, every launch of the
 * selected kernels is bracketed by a pair of timing-only HIP events on the handle's stream
 * (read back by ug_pack_kernel_stats, so the pack waits on nothing extra; each event still
 * holds the next launch back by a few microseconds) and the kernels count the timesteps they
 * visit and the grid points they compare (exactly, in device counters), so algorithmic
 * bytes per launch can be derived.
 * No counterpart in the reference (which prints only attempts/s, src/main.rs:150-159). */
typedef struct ug_kernel_stats {
    int64_t launches; /* kernel launches measured */
    int64_t items;    /* candidate tests (a survivor retested after an insert counts again) */
    double ms;        /* summed event time */
    int64_t visits;   /* (candidate, timestep) pairs visited */
    int64_t points;   /* grid points compared */
    /* split-BF16 prefilter, stage 1 (counted in every launch with UG_PROFILE_REASONS): the
     * candidates it leaves to the FP64 test, by reason -- spacing sum below the error bound's
     * premise; a possible contact (within r + margin) but no certain one among the examined
     * timesteps; no possible contact there at all.  Zero for the other kernels. */
    int64_t left_wsum, left_possible, left_none;
} ug_kernel_stats;

enum {
    UG_KERNEL_PREFILTER = 0, /* first test of fresh candidates (FP64 or split-BF16 form) */
    UG_KERNEL_RECHECK = 1    /* FP64 recheck of the split-BF16 form's uncertain candidates */
};

/* ug_pack_profile's `kernels`: which launches are timed (bit UG_KERNEL_* set); any nonzero
 * value turns the counters on, 0 turns measurement off.  UG_PROFILE_REASONS also counts why
 * the split-BF16 prefilter leaves candidates to the FP64 test (ug_kernel_stats left_*; it
 * slows the prefilter, so timed packs leave it off).  Other bits are rejected.  In ug_kernel_stats, launches, ms and the prefilter's items cover the timed
 * launches only; visits, points and the recheck's items are counted in every launch while
 * measurement is on. */
enum {
    UG_PROFILE_PREFILTER = 1 << UG_KERNEL_PREFILTER,
    UG_PROFILE_RECHECK = 1 << UG_KERNEL_RECHECK,
    UG_PROFILE_ALL = UG_PROFILE_PREFILTER | UG_PROFILE_RECHECK,
    UG_PROFILE_REASONS = 1 << 2
};

int32_t ug_pack_profile(ug_pack* h, int32_t kernels);
/* waits for the handle's profiled launches still in flight */
int32_t ug_pack_kernel_stats(ug_pack* h, int32_t which, ug_kernel_stats* out);

/* Full-size check of the serial-RSA characterization of the unique phase (no reference
 * counterpart; the oracle's brute force, done on the device at any size):
 *   out[0] = stored uniques that touch another stored unique (must be 0),
 *   out[1] = rejected candidates in [first, first + count) that touch no unique accepted
 *            before them (must be 0),
 *   out[2] = uniques checked, out[3] = rejected candidates checked.
 * Candidates must lie within the decided range (the consumed one, except that max_n
 * reached inside a UG_STOP_REF_WINDOW round leaves the rest of the round untested).
 * Subpaths are not included. */
int32_t ug_pack_check(ug_pack* h, uint64_t first, int64_t count, int64_t* out);

/* Releases the handle.  Its device buffers and pinned staging block are kept in a
 * process-wide cache (at most 16 GiB per device) for the next handle of the same model,
 * and the read-only basis tables of each (device, dim, T, terms, precision, margin) are
 * built once and shared; a run takes nothing from either that it does not reinitialise. */
void ug_pack_destroy(ug_pack* h);

/* Frees the cached buffers of destroyed handles (the shared model tables stay).  No
 * counterpart in the reference, whose engine is one process per universe; for callers that
 * need the memory back between batches of packs. */
int32_t ug_cache_trim(void);

/* Packs n independent universes (cfgs[i], each run as ug_pack_run with max_candidates 0)
 * on the current device, `threads` at a time: one host worker thread and one HIP stream
 * each, universes taken in index order from a shared counter (no reference counterpart:
 * the reference runs one universe per engine process, braidlab/orchestrator.py).
 * status[i] receives universe i's final status; digests (may be NULL) receives the 64-bit
 * FNV-1a hash of universe i's accepted candidate indices (int64 little-endian bytes in
 * admission order).  keep == NULL: every handle is destroyed after its run; else keep[i]
 * receives universe i's handle (on a stream of its own, released with it) for the caller
 * to query and destroy.  On error the remaining universes are not started, no handle is
 * kept, and the first failure's message is in ug_last_error(). */
int32_t ug_ensemble_run(const ug_pack_config* cfgs, int32_t n, int32_t threads, ug_pack** keep,
                        ug_pack_status* status, uint64_t* digests);

/* Packs n independent universes like ug_ensemble_run with the batched engine: `lanes` host
 * threads (0 = 1), each with a HIP stream of its own, advance up to `window` universes each
 * (0 = 32, at most 64) in lock-step; finished universes are replaced from a shared queue in
 * index order.  A step moves every universe of a lane by one sub-batch (or one retest of survivors left
 * from its last resolve chunk), each stage one launch for all of them -- the split-BF16
 * prefilter and the FP64 tests over a flattened (universe, tile / candidate) work queue,
 * compaction and the resolve chunk one workgroup row per universe -- with survivor counts
 * kept on the device and one host synchronisation per step.  Each universe's accepted
 * list, status and digest are those of its own ug_pack_run.  The batched kernels cover
 * full-spectrum universes with the dense cell table, T <= 160 and the budget or fixed-window
 * stop; other configs run one at a time on a lane's stream after its batched universes.
 * keep: as ug_ensemble_run (each kept handle gets a stream of its own).  No counterpart in
 * the reference (one engine process per universe, braidlab/orchestrator.py:70-82,106-107). */
int32_t ug_ensemble_run_batched(const ug_pack_config* cfgs, int32_t n, int32_t window, int32_t lanes,
                                ug_pack** keep, ug_pack_status* status, uint64_t* digests);

/* ------------------------------------------------------------------ knot regime
 * Replaces the braid_knot executable (cuda/braid_knot.cu): few strands, very many
 * candidates, no spatial grid.  Candidates are the packer's counter-indexed Philox
 * candidates, admitted in index order against every resident strand (the canonical FP64
 * evaluation and pair predicate), so a run is serial RSA per species and independent of
 * batch sizes (the reference's own determinism model, :14-20,433-462). */

/* The flags of cuda/braid_knot.cu:255-289. */
typedef struct ug_knot_config {
    int32_t dim;         /* --dim: 2 or 3 */
    int32_t T;           /* -t */
    int32_t terms;       /* --terms (0 = T, the full band) */
    int32_t pin_sin1;    /* --pin-sin1 */
    uint64_t seed;       /* --seed */
    double max_attempts; /* --attempts: primary-species candidates [0, max_attempts) */
    double nu_attempts;  /* --nu-attempts: nu-species candidates after them (0 = off) */
    double nu_div;       /* --nu-div k >= 1: the nu exclusion is (2/T)/k */
    int32_t max_strands; /* strand capacity, 0 = 16384 (kMaxStrands, :94-97) */
    int32_t precision;   /* UG_PREC_FP64, or UG_PREC_SPLIT_BF16: full-band models test the first
                            32 ordered timesteps on the split-BF16 prefilter first; same strands */
    int32_t rng_kind;    /* UG_RNG_* (0 = Philox; UG_RNG_XOSHIRO_REF draws the reference knot
                            tool's own candidates, cuda/braid_knot.cu:151-166,207) */
} ug_knot_config;

typedef struct ug_knot_status {
    int64_t attempts;         /* candidates consumed, both phases */
    int64_t n;                /* strands admitted, both species */
    int64_t n_primary;        /* primary-species strands */
    int64_t primary_attempts; /* attempts when the primary phase ended */
    int32_t phase;            /* 0 primary, 1 nu, 2 finished */
    int32_t done;
    double seconds;
} ug_knot_status;

typedef struct ug_knot ug_knot;

/* Validates cfg like the tool's argv checks (:290-306) and binds the handle to hip_stream on
 * the current device. */
int32_t ug_knot_create(const ug_knot_config* cfg, void* hip_stream, ug_knot** out);
/* Runs the phases (:404-476) until both budgets are spent or the strand capacity is reached,
 * or until max_candidates more candidates were consumed (0 = no limit; batches are never
 * split between calls).  st may be NULL. */
int32_t ug_knot_run(ug_knot* h, int64_t max_candidates, ug_knot_status* st);
int32_t ug_knot_get_status(const ug_knot* h, ug_knot_status* st);
/* Admitted strands in admission order: candidate index (the reference logs index + 1) and
 * species (0 primary, 1 nu); *n receives the total. */
int32_t ug_knot_strands(const ug_knot* h, int64_t* idx, int32_t* species, int64_t cap, int64_t* n);
/* Dump parameters of strands [first, first + count): per strand and axis a1 then nw
 * (amplitude, frequency, phase), as ug_pack_params (the tool's --dump rows, :442-452). */
int32_t ug_knot_params(const ug_knot* h, int64_t first, int64_t count, double* out);
/* Batches run, 5 values each: first candidate, candidates, survivors against the strands of
 * the batch's start (the ramp's `sc`), admitted, species. */
int32_t ug_knot_batches(const ug_knot* h, int64_t* rows5, int64_t cap, int64_t* n);
void ug_knot_destroy(ug_knot* h);

/* ------------------------------------------------------------------ frame baker
 * Replaces the frame_bake executable (cuda/frame_bake.cu:398-635), whose BRF1 output
 * braidlab/frames.py:88-144 reads. */

/* A device-resident worldline set.  Host arrays in dump order: a2[n][dim] (sin1
 * amplitudes), amp/freq/phase[n][dim][nw] (the dump's a_j, b_j, f_j columns; freq must be
 * integers; phase may be NULL for phase-free dumps), as load_dump reads them
 * (cuda/frame_bake.cu:109-190) or ug_pack_params returns them. */
typedef struct ug_paths ug_paths;
int32_t ug_paths_create(int32_t dim, int64_t n, int32_t nw, const double* a2, const double* amp,
                        const double* freq, const double* phase, void* hip_stream, ug_paths** out);
void ug_paths_destroy(ug_paths* p);

/* The worldline set of a packing (uniques, then subpaths), built on the device from the
 * accepted candidate indices: no host round trip of the dump parameters (no reference
 * counterpart; the reference bakes from the dump CSV, cuda/frame_bake.cu:109-190). */
int32_t ug_paths_from_pack(const ug_pack* h, void* hip_stream, ug_paths** out);

/* Front of one observation (the --front/--metric/--speed/--zmin/--wraps flags,
 * cuda/frame_bake.cu:409-453). */
typedef struct ug_front {
    int32_t kind;   /* 0 speed, 1 fit, 2 budget */
    int32_t cheb;   /* 1 square (Chebyshev) metric, 0 circle; budget is always square */
    double speed;   /* conformal front speed for kind 0 */
    double z_min;   /* rewind floor (radians) */
    int32_t wraps;  /* torus image cap K, -1 = auto ceil((chi_max + 4)/2) */
    int32_t reserved;
} ug_front;

typedef struct ug_observer {
    int32_t type;  /* 0 path, 1 comoving point */
    int32_t path;  /* path index for type 0 */
    double px, py; /* point for type 1 */
} ug_observer;

/* One image: the BRF1 hit record (u32 src, f32 dx, f32 dy, f32 chi, f32 z_emit). */
typedef struct ug_hit {
    uint32_t src;
    float dx, dy, chi, z_emit;
} ug_hit;

/* Per (instant, observer) header fields of BRF1. */
typedef struct ug_frame_info {
    double z_obs, chi_max;
    int32_t K, reserved;
    double ox, oy, betax, betay;
    uint64_t n_hits;
} ug_frame_info;

/* Past-light-cone hits of every worldline (and torus image) for one observation, as
 * hits_kernel + its host driver (cuda/frame_bake.cu:281-371,522-623), into caller device
 * memory dev_hits[cap], ordered by (path, image, crossing).  When the hits exceed cap,
 * returns UG_EOVERFLOW with info->n_hits set (cap = 0 is a size query).  2+1 only. */
int32_t ug_frame_bake_lightcone(ug_paths* p, const ug_front* front, double z_obs,
                                const ug_observer* obs, ug_hit* dev_hits, uint64_t cap,
                                ug_frame_info* info);

/* Equal-time slices (no reference counterpart; the frame step of the large 3+1 workload):
 * wrapped positions of every path at z[0..nz) into dev_pos[nz][n][dim] (float, may be
 * NULL) and raster_px^2 float density rasters of the (x, y) projection over [-1, 1)^2
 * into dev_raster[nz][raster_px][raster_px] (may be NULL; row = y bin, column = x bin). */
int32_t ug_frame_bake_slices(ug_paths* p, const double* z, int32_t nz, float* dev_pos,
                             float* dev_raster, int32_t raster_px);

/* ug_frame_bake_slices without any allocation: no reference counterpart.  For callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc): the same slices and rasters, bit for bit, with the frame basis
 * kept in a fixed 4 MiB device array (one per GPU, held until the process ends) and a host
 * staging array of the same size instead of per-call buffers.  Limit: 2 * kq * 64 <= 2^19
 * doubles, kq = the paths' largest frequency rounded up to a multiple of 4 (frequencies up
 * to 4096); larger returns UG_EOVERFLOW.  Runs on the current device, which must be the
 * paths' device.  Calls share the arrays: a lock serialises them within a process, so
 * calls that overlap (several threads, several devices) wait for each other. */
int32_t ug_frame_bake_slices_noalloc(ug_paths* p, const double* z, int32_t nz, float* dev_pos,
                                     float* dev_raster, int32_t raster_px);

#ifdef __cplusplus
}
#endif
#endif
