/*
 * ug_oracle.h -- CPU oracle for the RSA worldline packer (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the package's
 * smoke check and bench.py's cpu_baseline leg load it.  It restates, in plain
 * C99, the serial random-sequential-adsorption admission of the reference
 * (cuda/braid_cuda.cu:703-722 host_collides, :831-879 evalPath/insertPath,
 * cuda/braid_knot.cu:433-462 index-ordered admission) on the Philox candidate
 * stream: Philox4x32-10 keyed by the seed, one counter per (candidate, stage,
 * axis, slot pair), evaluated with explicit fma in a fixed summation order.
 * Every function below is written independently of the HIP implementation; the
 * two must agree bit for bit.
 */
#ifndef UG_ORACLE_H
#define UG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon et al., SC'11), one block. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* sin(pi*n/m), cos(pi*n/m) for integers n >= 0, 0 < m <= 2^32. */
void or_sincospi_rat(uint64_t n, uint64_t m, double* s, double* c);

/* Natural log of v in [2^-1022, +inf) (own atanh-series method). */
double or_log(double v);

/* Model geometry: basis tables and scan order for T timesteps, max frequency F = T. */
typedef struct or_model {
    int d;          /* spatial axes: 2 or 3 */
    int T;          /* timesteps (= max frequency) */
    int nw;         /* wiggle terms per axis (terms - 1) */
    int modmax;     /* frequency pool size: frequencies 2..modmax+1 */
    int pin;        /* --pin-sin1 */
    uint64_t seed;  /* Philox key */
    double r;       /* inclusive Chebyshev radius 2/T */
    double halfT;   /* T/2: cell index multiplier */
    double* S;      /* [T][T+1]: sin(k z_t)/sin z_t */
    double* C;      /* [T][T+1]: (cos(k z_t) - 1)/sin z_t */
    double* sz;     /* [T]: sin z_t */
    int* order;     /* [T]: endpoint-first scan order */
    int euclid;     /* contact: 0 = Chebyshev cube, 1 = L2 ball (--euclid-collision) */
    int rng;        /* candidate stream: 0 = Philox, 1 = the reference's xoshiro256** stream */
} or_model;

int or_model_init(or_model* m, int d, int T, int terms, uint64_t seed, int pin);
void or_model_free(or_model* m);

/* --euclid-collision (cuda/braid_cuda3d.cu:605-608): contact iff
 * dx*dx + dy*dy [+ dw*dw] <= r*r, summed in that order with fma. */
void or_model_set_euclid(or_model* m, int euclid);

/* Candidate stream (rng_kind 1 = UG_RNG_XOSHIRO_REF): candidate idx draws xoshiro256**
 * (cuda/braid_knot.cu:56-90) seeded with splitmix64 from seed ^ (idx*0x9E3779B97F4A7C15 +
 * 0x1234567) (:207), in the draw order of the knot tool's proposal (:112-166). */
void or_model_set_rng(or_model* m, int rng);

/* The four splitmix64 outputs that seed xoshiro256** from `seed` (cuda/braid_knot.cu:62-70),
 * and one xoshiro256** step on that state (:71-82). */
void or_xoshiro_seed(uint64_t seed, uint64_t s[4]);
uint64_t or_xoshiro_next(uint64_t s[4]);

/* The contact predicate of the model at one timestep (points x[d], y[d]). */
int or_touches(const or_model* m, const double* x, const double* y);

/* Candidate idx: a1[d], b[d*nw], c[d*nw], s[d*nw] (per-slot coefficients),
 * and optionally the dump form amp[d*nw], phase[d*nw] (may be NULL). */
void or_candidate(const or_model* m, uint64_t idx, double* a1, int* b, double* c, double* s,
                  double* amp, double* phase);

/* Canonical wrapped comoving position of a candidate at timestep t (X[d]). */
void or_eval_t(const or_model* m, int t, const double* a1, const int* b, const double* c,
               const double* s, double* X);

/* Canonical torus cell of a wrapped coordinate. */
int or_cell(const or_model* m, double x);

/* Serial RSA in candidate-index order.
 *   stop_kind 0: run candidates [0, max_candidates)
 *   stop_kind 1: as 0, but also stop at the first boundary c = k*stop_batch (c >= window)
 *                where accepts among [c-window, c) < stop_rate*window
 *   max_n > 0 stops right after the max_n-th accept.
 * Outputs (caller-allocated, capacity cap): idx[], gid[]; positions pos[n][t][d]
 * (may be NULL, capacity cap*T*d).  Returns N (>= 0) or a negative error;
 * *attempts = candidates consumed. */
int64_t or_pack(const or_model* m, int64_t max_candidates, int stop_kind, int64_t stop_batch,
                int64_t window, double stop_rate, int64_t max_n, int64_t* idx, int32_t* gid,
                double* pos, int64_t cap, int64_t* attempts);

/* or_pack (stop_kind 0/1, no max_n) with a parallel prefilter per block of candidates and
 * serial in-order admission on `threads` OpenMP threads; same output as or_pack. */
int64_t or_pack_mt(const or_model* m, int64_t max_candidates, int stop_kind, int64_t stop_batch,
                   int64_t window, double stop_rate, int threads, int64_t* idx, int64_t cap,
                   int64_t* attempts);

/* Subpath phase continuing an existing packing (cuda/braid_cuda.cu:949-1043).  Rounds of
 * `batch` candidates from index `first` on: a candidate survives the round when, against
 * the packing as it stood at the round's start, it touches exactly one group (collides_dev
 * sub mode, :463-464); survivors are then taken in index order and admitted into the
 * single group they touch in the up-to-date packing (host_sub_gid, :726-750).  batch
 * starts at 2^16 and adapts on min(survivors, 2^16) (:1000-1003); rounds run while fewer
 * than `budget` subpath candidates were drawn; accept_rate / fill_rate > 0 arm the
 * windowed stops of :1008-1037.  Uniques are pos0[n0][T][d] with groups gid0[n0].
 * rounds (may be NULL) receives up to rounds_cap records {start, size, survivors,
 * admitted, fresh cells} (start counted from `first`), *nrounds the number of rounds.
 * Returns the number of admitted subpaths (idx/gid hold the first cap), or < 0. */
int64_t or_pack_sub(const or_model* m, const double* pos0, const int32_t* gid0, int64_t n0,
                    uint64_t first, double budget, double accept_rate, double fill_rate,
                    int64_t* idx, int32_t* gid, int64_t cap, int64_t* attempts,
                    int64_t* filled_new, int64_t* rounds, int64_t rounds_cap, int64_t* nrounds);

/* Serial RSA consumed on the reference engines' round schedule (the unique phase of
 * cuda/braid_cuda.cu:885-944 for rule 2, cuda/braid_cuda3d.cu:1446-1566 for rule 3).
 * Round k holds candidates [start_k, start_k + size_k); its survivors are the candidates
 * that touch no path accepted before start_k (the grid the round's kernel sees), admitted
 * in index order against the up-to-date packing.  Sizes start at 2^16 and adapt on the
 * survivor count (halve above 512 down to 4096, double below 64 up to 2^26).  Rule 2:
 * rounds run while attempts < budget; the --until-accept-rate window (stop_rate > 0,
 * clamp(30/rate, 2e7, 5e9) attempts, attempts > 1e6) records the adapted size and stops
 * after the round.  Rule 3: the next round is sized before the current one is drained
 * (while enqueued attempts < budget and no stop), the window records the drained round,
 * and a stop still drains the round already sized.  max_n > 0 stops at the max_n-th
 * accept; *decided = the first candidate whose fate was not decided (= *attempts unless
 * max_n stopped mid-round).  rounds/nrounds as in or_pack_sub (fresh = newly occupied
 * cells).  Deviation (documented): the reference's 3+1 kernel sees a grid that lags by at
 * least one round; here every round sees the packing at its start.  Returns N or < 0. */
int64_t or_pack_ref(const or_model* m, double budget, int rule, double stop_rate, int64_t max_n,
                    int64_t* idx, int32_t* gid, int64_t cap, int64_t* attempts,
                    int64_t* decided, int64_t* rounds, int64_t rounds_cap, int64_t* nrounds);

/* Candidates in [first, first + count) that touch none of the candidates acc[0..na), by
 * direct pairwise tests at every timestep (no grid): an independent count of a round's
 * survivors against the paths accepted before it.  Returns the count or < 0. */
int64_t or_count_clear(const or_model* m, int64_t first, int64_t count, const int64_t* acc,
                       int64_t na);

/* The knot-regime tool (cuda/braid_knot.cu:387-476) on the Philox stream: the primary
 * species over candidates [0, budget) at the exclusion r = 2/T, then, when nu_attempts > 0,
 * the nu species over the next nu_attempts candidates at r/nu_div, against every strand of
 * both species.  Each batch's survivors are the candidates touching no strand of the batch's
 * start; they are admitted in index order against the up-to-date strands, at most
 * max_strands in all.  Batches start at 4096 per phase, shrink /4 (not below 4096) when
 * survivors exceed size/8 and grow x4 (up to 2^24) below size/64, and are cut at the phase's
 * budget.  Outputs: idx/species of the strands (admission order), *attempts,
 * *primary_attempts and *n_primary (at the end of the primary phase), batch records
 * {start, size, survivors, admitted, species}.  Returns the strand count or < 0. */
int64_t or_pack_knot(const or_model* m, double budget, double nu_attempts, double nu_div,
                     int64_t max_strands, int64_t* idx, int32_t* species, int64_t cap,
                     int64_t* attempts, int64_t* primary_attempts, int64_t* n_primary,
                     int64_t* rounds, int64_t rounds_cap, int64_t* nrounds);

/* Canonical wrapped positions out[n][T][d] of candidates idx[0..n). */
void or_positions(const or_model* m, const int64_t* idx, int64_t n, double* out);

/* Brute-force check that acc[0..na) is exactly the serial RSA output over candidate
 * indices [0, M): (i) no two members collide, (ii) every non-member collides with an
 * earlier member.  Returns 0 when both hold, else 1 + the first offending index. */
int64_t or_check_rsa(const or_model* m, int64_t M, const int64_t* acc, int64_t na);

#ifdef __cplusplus
}
#endif
#endif
