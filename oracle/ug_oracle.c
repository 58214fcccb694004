/*
 * ug_oracle.c -- CPU oracle for the RSA worldline packer.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, the smoke check and bench.py's cpu_baseline leg, never by the
 * product path.
 *
 * What it restates (reference files under universe-analysis/universe-generator):
 *   - proposal measure: cuda/braid_cuda.cu:295-373 (multi-term branch), with the
 *     draw stages keyed by Philox counters instead of a serial xoshiro stream;
 *     full-pool frequency set cuda/braid_cuda.cu:203-219, unique-frequency
 *     rejection draws :220-243, Dirichlet(1) by exponential spacings :250-264,
 *     a_j = w_j/b_j with random signs :348-359, phases on even b only :366-371,
 *     sin1 amplitude U(-1,1) (cuda/braid_knot.cu:164);
 *   - time grid z_i = (i+1)*pi/(T+1), endpoint-first scan: cuda/braid_cuda.cu:647-664;
 *   - evaluation X = [a1 sin z + sum a_j(sin(b_j z + f_j) - sin f_j)]/sin z in the
 *     basis-table form of cuda/braid_knot.cu:219-230,341-354;
 *   - torus_wrap / torus_delta: cuda/braid_cuda.cu:158-168;
 *   - Chebyshev exclusion of half-width 2/T, inclusive: cuda/braid_cuda.cu:449-451,604;
 *   - per-timestep hash grid, 3^d neighbour scan: cuda/braid_cuda.cu:697-722,865-879;
 *   - index-ordered serial admission: cuda/braid_knot.cu:433-462;
 *   - subpath adoption rule: cuda/braid_cuda.cu:726-750.
 *
 * The arithmetic is the canonical FP64 definition shared with the GPU kernels
 * (explicit fma, fixed summation tree, own sin/cos/log); this file is compiled
 * with -ffp-contract=off so no other contraction happens.
 */
#include "ug_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- Philox4x32-10 */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0;
        x1 = y1;
        x2 = y2;
        x3 = y3;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

/* ------------------------------------------------------------ sin(pi x), cos(pi x)
 * Taylor coefficients of sin(pi r) and cos(pi r) in r (|r| <= 1/4):
 * (-1)^i pi^(2i+1)/(2i+1)! and (-1)^i pi^(2i)/(2i)!, rounded to double. */
static const double kSinPi[10] = {
    0x1.921fb54442d18p+1,  -0x1.4abbce625be53p+2, 0x1.466bc6775aae2p+1,  -0x1.32d2cce62bd86p-1,
    0x1.50783487ee782p-4,  -0x1.e3074fde8871fp-8, 0x1.e8f434d018d63p-12, -0x1.6fadb9f155744p-16,
    0x1.aaec32af93359p-21, -0x1.8a404211f9547p-26};
static const double kCosPi[10] = {
    0x1.0000000000000p+0,  -0x1.3bd3cc9be45dep+2, 0x1.03c1f081b5ac4p+2,  -0x1.55d3c7e3cbffap+0,
    0x1.e1f506891babbp-3,  -0x1.a6d1f2a204a8cp-6, 0x1.f9d38a3763cc3p-10, -0x1.b6e24f44b128fp-14,
    0x1.20c62c2f2d7f5p-18, -0x1.2a0c591af8314p-23};

void or_sincospi_rat(uint64_t n, uint64_t m, double* s, double* c) {
    /* Exact reduction of pi*n/m to pi*(r + q/2), |r| <= 1/4, q in 0..4. */
    uint64_t two_m = 2 * m;
    uint64_t nn = n % two_m;
    uint64_t q = (4 * nn + m) / two_m;
    int64_t num = (int64_t)(2 * nn) - (int64_t)(q * m);
    double r = (double)num / (double)two_m;
    double x = r * r;
    double ps = kSinPi[9];
    for (int i = 8; i >= 0; i--)
        ps = fma(ps, x, kSinPi[i]);
    double pc = kCosPi[9];
    for (int i = 8; i >= 0; i--)
        pc = fma(pc, x, kCosPi[i]);
    double sn = r * ps, cs = pc;
    switch (q & 3) {
        case 0: *s = sn; *c = cs; break;
        case 1: *s = cs; *c = -sn; break;
        case 2: *s = -sn; *c = -cs; break;
        default: *s = -cs; *c = sn; break;
    }
}

/* ----------------------------------------------------------------------- log(v)
 * v = 2^e * m, m in [sqrt(1/2), sqrt(2)];  log m = 2 atanh(s), s = (m-1)/(m+1),
 * 2 atanh(s) = 2s + s * z * sum_{i=1..10} (2/(2i+1)) z^(i-1),  z = s^2 <= 0.0295. */
static const double kLogAt[10] = {
    0x1.5555555555555p-1, 0x1.999999999999ap-2, 0x1.2492492492492p-2, 0x1.c71c71c71c71cp-3,
    0x1.745d1745d1746p-3, 0x1.3b13b13b13b14p-3, 0x1.1111111111111p-3, 0x1.e1e1e1e1e1e1ep-4,
    0x1.af286bca1af28p-4, 0x1.8618618618618p-4};
static const double kLn2Hi = 0x1.62e42fee00000p-1;
static const double kLn2Lo = 0x1.a39ef35793c76p-33;

double or_log(double v) {
    uint64_t bits;
    memcpy(&bits, &v, 8);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    uint64_t mb = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    if (mb > 0x3ff6a09e667f3bcdULL) { /* m > sqrt(2): use m/2 */
        mb -= 0x0010000000000000ULL;
        e += 1;
    }
    double m;
    memcpy(&m, &mb, 8);
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = kLogAt[9];
    for (int i = 8; i >= 0; i--)
        p = fma(p, z, kLogAt[i]);
    double lm = fma(s, z * p, 2.0 * s);
    double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}

/* ------------------------------------------------------------------------ model */
static double wrap1(double x) { /* torus_wrap, cuda/braid_cuda.cu:158-160 */
    return x - 2.0 * floor((x + 1.0) * 0.5);
}
static double tdelta(double d) { /* torus_delta, cuda/braid_cuda.cu:161-168 */
    if (d > 1.0)
        return d - 2.0;
    if (d < -1.0)
        return d + 2.0;
    return d;
}

int or_model_init(or_model* m, int d, int T, int terms, uint64_t seed, int pin) {
    memset(m, 0, sizeof(*m));
    if ((d != 2 && d != 3) || T < 4 || T > 4096)
        return -1;
    int modmax = T - 1 > 2 ? T - 1 : 2;
    if (terms < 2)
        terms = 2;
    if (terms - 1 > modmax)
        return -1;
    m->d = d;
    m->T = T;
    m->nw = terms - 1;
    m->modmax = modmax;
    m->pin = pin;
    m->seed = seed;
    m->r = 2.0 / (double)T;
    m->halfT = 0.5 * (double)T;
    int W = T + 1;
    m->S = (double*)calloc((size_t)T * W, sizeof(double));
    m->C = (double*)calloc((size_t)T * W, sizeof(double));
    m->sz = (double*)calloc((size_t)T, sizeof(double));
    m->order = (int*)calloc((size_t)T, sizeof(int));
    if (!m->S || !m->C || !m->sz || !m->order)
        return -2;
    /* z_t = (t+1) pi/(T+1) (cuda/braid_cuda.cu:655-660). */
    uint64_t D = (uint64_t)T + 1;
    for (int t = 0; t < T; t++) {
        double s1, c1;
        or_sincospi_rat((uint64_t)(t + 1), D, &s1, &c1);
        m->sz[t] = s1;
        for (int k = 0; k <= T; k++) {
            double sk, ck, hk, hc;
            or_sincospi_rat((uint64_t)k * (uint64_t)(t + 1), D, &sk, &ck);
            or_sincospi_rat((uint64_t)k * (uint64_t)(t + 1), 2 * D, &hk, &hc);
            m->S[(size_t)t * W + k] = sk / s1;
            m->C[(size_t)t * W + k] = (-2.0 * (hk * hk)) / s1;
        }
    }
    /* Endpoint-first order (cuda/braid_cuda.cu:663-664): 0, T-1, 1, T-2, ... */
    int o = 0;
    for (int i = 0; o < T; i++) {
        m->order[o++] = i;
        if (o < T)
            m->order[o++] = T - 1 - i;
    }
    return 0;
}

static int touches(const or_model* m, const double* x, const double* y);

void or_model_set_euclid(or_model* m, int euclid) { m->euclid = euclid ? 1 : 0; }

void or_model_set_rng(or_model* m, int rng) { m->rng = rng ? 1 : 0; }

/* splitmix64 (Steele, Lea and Flood): the state advances by the golden-ratio increment and
 * each output is the state passed through two xor-shift-multiply rounds and a final
 * xor-shift; the four outputs from `seed` are xoshiro's state (cuda/braid_knot.cu:62-70). */
void or_xoshiro_seed(uint64_t seed, uint64_t s[4]) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) {
        x += 0x9E3779B97F4A7C15ULL;
        uint64_t z = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
}

static uint64_t rol64(uint64_t v, unsigned k) { return (v << k) | (v >> (64u - k)); }

/* xoshiro256** (Blackman and Vigna): output = rotl(s1 * 5, 7) * 9, then the linear state
 * update (cuda/braid_knot.cu:71-82). */
uint64_t or_xoshiro_next(uint64_t s[4]) {
    const uint64_t result = rol64(s[1] * 5u, 7) * 9u;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rol64(s[3], 45);
    return result;
}

/* 53-bit uniform k of one draw: u = k 2^-53 in [0, 1) (cuda/braid_knot.cu:84-86). */
static uint64_t xo_u53(uint64_t s[4]) { return or_xoshiro_next(s) >> 11; }

/* Candidate idx on the reference's stream: per axis, in stream order (cuda/braid_knot.cu:
 * 151-166), the frequencies (draw_freqs :124-149: Fisher-Yates over 2..modmax+1 for the full
 * pool, else up to 64 rejection draws per slot, the last kept even when a duplicate), the nw
 * exponential spacings -log(1 - u) summed serially and divided by their sum (split_unit
 * :112-120), per slot the amplitude w_j / b_j, its sign draw and, for an even frequency, its
 * phase draw f = pi u, then the sin1 amplitude 2u - 1 unless pinned.  The phase's sine and
 * cosine are taken of the exact pi k 2^-53 (sincospi of the rational k / 2^53). */
static void candidate_xref(const or_model* m, uint64_t idx, double* a1, int* b, double* c, double* s,
                           double* amp, double* phase) {
    const int d = m->d, nw = m->nw;
    uint64_t st[4];
    or_xoshiro_seed(m->seed ^ (idx * 0x9E3779B97F4A7C15ULL + 0x1234567ULL), st);
    double* e = (double*)malloc((size_t)nw * sizeof(double));
    for (int a = 0; a < d; a++) {
        int* ba = b + (size_t)a * nw;
        if (nw == m->modmax) {
            for (int j = 0; j < nw; j++)
                ba[j] = j + 2;
            for (int j = nw - 1; j >= 1; j--) {
                const int k = (int)(or_xoshiro_next(st) % (uint64_t)(j + 1));
                const int tmp = ba[k];
                ba[k] = ba[j];
                ba[j] = tmp;
            }
        } else {
            for (int j = 0; j < nw; j++) {
                int cand = 0, clash = 1;
                for (int tries = 0; clash && tries < 64; tries++) {
                    cand = 2 + (int)(or_xoshiro_next(st) % (uint64_t)m->modmax);
                    clash = 0;
                    for (int k = 0; k < j && !clash; k++)
                        clash = ba[k] == cand;
                }
                ba[j] = cand;
            }
        }
        double sum = 0.0;
        for (int j = 0; j < nw; j++) {
            const uint64_t k = xo_u53(st);
            e[j] = -or_log(1.0 - (double)k * 0x1p-53);
            sum += e[j];
        }
        for (int j = 0; j < nw; j++) {
            double aj = (e[j] / sum) / (double)ba[j];
            if (or_xoshiro_next(st) & 1u)
                aj = -aj;
            double cj = aj, sj = 0.0, fj = 0.0;
            if (ba[j] % 2 == 0) {
                const uint64_t k = xo_u53(st);
                double sf, cf;
                or_sincospi_rat(k, 1ULL << 53, &sf, &cf);
                cj = aj * cf;
                sj = aj * sf;
                fj = (double)k * 0x1p-53 * 3.141592653589793;
            }
            c[(size_t)a * nw + j] = cj;
            s[(size_t)a * nw + j] = sj;
            if (amp)
                amp[(size_t)a * nw + j] = aj;
            if (phase)
                phase[(size_t)a * nw + j] = fj;
        }
        a1[a] = m->pin ? 0.0 : (double)xo_u53(st) * 0x1p-53 * 2.0 - 1.0;
    }
    free(e);
}

int or_touches(const or_model* m, const double* x, const double* y) { return touches(m, x, y); }

void or_model_free(or_model* m) {
    free(m->S);
    free(m->C);
    free(m->sz);
    free(m->order);
    memset(m, 0, sizeof(*m));
}

static void philox_at(const or_model* m, uint64_t idx, uint32_t c2, uint32_t c3, uint32_t w[4]) {
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), c2, c3};
    uint32_t key[2] = {(uint32_t)m->seed, (uint32_t)(m->seed >> 32)};
    or_philox4x32_10(ctr, key, w);
}

void or_candidate(const or_model* m, uint64_t idx, double* a1, int* b, double* c, double* s,
                  double* amp, double* phase) {
    if (m->rng) {
        candidate_xref(m, idx, a1, b, c, s, amp, phase);
        return;
    }
    const int d = m->d, nw = m->nw;
    const int npairs = (nw + 1) / 2;
    uint32_t w0[4];
    philox_at(m, idx, 0u, 0u, w0);
    for (int a = 0; a < d; a++) {
        /* sin1 amplitude: (2w+1) 2^-32 - 1, uniform on (-1, 1). */
        a1[a] = m->pin ? 0.0 : (double)(2 * (uint64_t)w0[a] + 1) * 0x1p-32 - 1.0;
        int* ba = b + (size_t)a * nw;
        /* Frequencies: the full pool {2..modmax+1} in ascending slot order, or nw
         * unique rejection draws (cuda/braid_cuda.cu:220-243) from stream 2. */
        if (nw == m->modmax) {
            for (int j = 0; j < nw; j++)
                ba[j] = j + 2;
        } else {
            uint32_t g = 0, wb[4] = {0, 0, 0, 0};
            for (int j = 0; j < nw; j++) {
                int bb = 0, dup = 1;
                for (int guard = 0; dup && guard < 64; guard++) {
                    if ((g & 3) == 0)
                        philox_at(m, idx, g >> 2, (2u << 8) | (uint32_t)a, wb);
                    bb = (int)(wb[g & 3] % (uint32_t)m->modmax) + 2;
                    g++;
                    dup = 0;
                    for (int k = 0; k < j; k++)
                        if (ba[k] == bb)
                            dup = 1;
                }
                if (dup) {
                    for (int cc = 2; cc <= m->modmax + 1; cc++) {
                        int used = 0;
                        for (int k = 0; k < j; k++)
                            if (ba[k] == cc)
                                used = 1;
                        if (!used) {
                            bb = cc;
                            break;
                        }
                    }
                }
                ba[j] = bb;
            }
        }
        /* Per slot pair p: Dirichlet words (bit 0 = sign), phase words.  Partial spectra: block
         * p = {Dirichlet(2p), Dirichlet(2p+1), phase(2p), phase(2p+1)}.  Full pool (only even
         * slots have a phase): the axis' blocks are one word stream (word i = component i % 4
         * of block i / 4) and pair p takes words 3p, 3p+1, 3p+2. */
        double e[4096];
        uint32_t neg[4096], ph[4096];
        double Q[64];
        for (int i = 0; i < 64; i++)
            Q[i] = 0.0;
        for (int p = 0; p < npairs; p++) {
            uint32_t w[4];
            if (nw == m->modmax) {
                uint32_t b0[4], b1[4];
                const uint32_t i0 = 3u * (uint32_t)p;
                philox_at(m, idx, i0 >> 2, (1u << 8) | (uint32_t)a, b0);
                philox_at(m, idx, (i0 >> 2) + 1u, (1u << 8) | (uint32_t)a, b1);
                for (int k = 0; k < 3; k++) {
                    const uint32_t wi = (i0 & 3u) + (uint32_t)k;
                    w[k] = wi < 4 ? b0[wi] : b1[wi - 4];
                }
                w[3] = 0u;
            } else {
                philox_at(m, idx, (uint32_t)p, (1u << 8) | (uint32_t)a, w);
            }
            for (int q = 0; q < 2; q++) {
                int j = 2 * p + q;
                if (j >= nw)
                    break;
                uint32_t mm = w[q] >> 1;
                double v = (double)(0x80000000u - mm) * 0x1p-31;
                e[j] = -or_log(v);
                neg[j] = w[q] & 1u;
                ph[j] = w[2 + q];
                Q[p & 63] = Q[p & 63] + e[j];
            }
        }
        for (int h = 32; h >= 1; h >>= 1)
            for (int i = 0; i < h; i++)
                Q[i] = Q[i] + Q[i + h];
        double invW = 1.0 / Q[0];
        for (int j = 0; j < nw; j++) {
            double invb = 1.0 / (double)ba[j];
            double aj = (nw == 1) ? invb : (e[j] * invW) * invb;
            if (neg[j])
                aj = -aj;
            double cj, sj, fj = 0.0;
            if ((ba[j] & 1) == 0) {
                double sf, cf;
                or_sincospi_rat((uint64_t)ph[j], 0x100000000ULL, &sf, &cf);
                cj = aj * cf;
                sj = aj * sf;
                fj = (double)ph[j] * 0x1p-32 * 3.141592653589793;
            } else {
                cj = aj;
                sj = 0.0;
            }
            c[(size_t)a * nw + j] = cj;
            s[(size_t)a * nw + j] = sj;
            if (amp)
                amp[(size_t)a * nw + j] = aj;
            if (phase)
                phase[(size_t)a * nw + j] = fj;
        }
    }
}

void or_eval_t(const or_model* m, int t, const double* a1, const int* b, const double* c,
               const double* s, double* X) {
    const int nw = m->nw, npairs = (nw + 1) / 2, W = m->T + 1;
    const double* St = m->S + (size_t)t * W;
    const double* Ct = m->C + (size_t)t * W;
    for (int a = 0; a < m->d; a++) {
        double P[64];
        for (int i = 0; i < 64; i++)
            P[i] = 0.0;
        for (int p = 0; p < npairs; p++) {
            for (int q = 0; q < 2; q++) {
                int j = 2 * p + q;
                if (j >= nw)
                    break;
                size_t o = (size_t)a * nw + j;
                int k = b[o];
                P[p & 63] = fma(c[o], St[k], P[p & 63]);
                if ((k & 1) == 0)
                    P[p & 63] = fma(s[o], Ct[k], P[p & 63]);
            }
        }
        for (int h = 32; h >= 1; h >>= 1)
            for (int i = 0; i < h; i++)
                P[i] = P[i] + P[i + h];
        X[a] = wrap1(a1[a] + P[0]);
    }
}

int or_cell(const or_model* m, double x) {
    int c = (int)floor((x + 1.0) * m->halfT);
    if (c < 0)
        c += m->T;
    if (c >= m->T)
        c -= m->T;
    return c;
}

/* ----------------------------------------------------------- per-timestep hash grid
 * Open addressing keyed by (t, cell); value = head path of a chain through nxt[n*T+t]
 * (the role of the reference's unordered_map<long, vector<HPt>> per timestep). */
typedef struct {
    uint64_t* keys; /* key+1, 0 = empty */
    int32_t* vals;
    uint64_t cap, used;
} hgrid;

static uint64_t hmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}
static int hg_init(hgrid* h, uint64_t cap) {
    h->cap = cap;
    h->used = 0;
    h->keys = (uint64_t*)calloc(cap, 8);
    h->vals = (int32_t*)calloc(cap, 4);
    return (h->keys && h->vals) ? 0 : -2;
}
static void hg_free(hgrid* h) {
    free(h->keys);
    free(h->vals);
}
static int32_t* hg_slot(hgrid* h, uint64_t key, int create);
static int hg_grow(hgrid* h) {
    hgrid n;
    if (hg_init(&n, h->cap * 2))
        return -2;
    for (uint64_t i = 0; i < h->cap; i++)
        if (h->keys[i])
            *hg_slot(&n, h->keys[i] - 1, 1) = h->vals[i];
    hg_free(h);
    *h = n;
    return 0;
}
static int32_t* hg_slot(hgrid* h, uint64_t key, int create) {
    uint64_t i = hmix(key) & (h->cap - 1);
    for (;;) {
        if (h->keys[i] == key + 1)
            return &h->vals[i];
        if (h->keys[i] == 0) {
            if (!create)
                return NULL;
            h->keys[i] = key + 1;
            h->vals[i] = -1;
            h->used++;
            return &h->vals[i];
        }
        i = (i + 1) & (h->cap - 1);
    }
}

typedef struct {
    const or_model* m;
    hgrid h;
    int64_t n, cap;
    double* pos;   /* [cap][T][d] */
    int32_t* nxt;  /* [cap][T] */
    int32_t* gid;  /* [cap] */
    uint64_t ncell;
} packing;

static int pk_init(packing* P, const or_model* m) {
    memset(P, 0, sizeof(*P));
    P->m = m;
    P->ncell = 1;
    for (int a = 0; a < m->d; a++)
        P->ncell *= (uint64_t)m->T;
    P->cap = 1024;
    P->pos = (double*)malloc((size_t)P->cap * m->T * m->d * 8);
    P->nxt = (int32_t*)malloc((size_t)P->cap * m->T * 4);
    P->gid = (int32_t*)malloc((size_t)P->cap * 4);
    if (!P->pos || !P->nxt || !P->gid)
        return -2;
    return hg_init(&P->h, 1u << 16);
}
static void pk_free(packing* P) {
    free(P->pos);
    free(P->nxt);
    free(P->gid);
    hg_free(&P->h);
}
static uint64_t cell_key(const packing* P, int t, const int* cx) {
    uint64_t id = 0;
    for (int a = 0; a < P->m->d; a++)
        id = id * (uint64_t)P->m->T + (uint64_t)cx[a];
    return (uint64_t)t * P->ncell + id;
}
/* Inserts trajectory X[T][d] under group g; returns newly occupied cells or <0. */
static int64_t pk_insert(packing* P, const double* X, int32_t g) {
    const or_model* m = P->m;
    const int T = m->T, d = m->d;
    if (P->n == P->cap) {
        int64_t nc = P->cap * 2;
        double* np = (double*)realloc(P->pos, (size_t)nc * T * d * 8);
        if (!np)
            return -2;
        P->pos = np;
        int32_t* nn = (int32_t*)realloc(P->nxt, (size_t)nc * T * 4);
        if (!nn)
            return -2;
        P->nxt = nn;
        int32_t* ng = (int32_t*)realloc(P->gid, (size_t)nc * 4);
        if (!ng)
            return -2;
        P->gid = ng;
        P->cap = nc;
    }
    int64_t n = P->n, fresh = 0;
    memcpy(P->pos + (size_t)n * T * d, X, (size_t)T * d * 8);
    for (int t = 0; t < T; t++) {
        if (2 * (P->h.used + 1) > P->h.cap && hg_grow(&P->h))
            return -2;
        int cx[3];
        for (int a = 0; a < d; a++)
            cx[a] = or_cell(m, X[t * d + a]);
        int32_t* head = hg_slot(&P->h, cell_key(P, t, cx), 1);
        if (*head < 0)
            fresh++;
        P->nxt[(size_t)n * T + t] = *head;
        *head = (int32_t)n;
    }
    P->gid[n] = g;
    P->n++;
    return fresh;
}
/* Contact of two points at one timestep (torus min-image per axis): Chebyshev cube
 * (every |d| <= r, cuda/braid_cuda.cu:451) or the L2 ball of --euclid-collision. */
static int touches(const or_model* m, const double* x, const double* y) {
    double dd[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < m->d; a++)
        dd[a] = tdelta(x[a] - y[a]);
    if (m->euclid) {
        double q = dd[0] * dd[0];
        for (int a = 1; a < m->d; a++)
            q = fma(dd[a], dd[a], q);
        return q <= m->r * m->r;
    }
    for (int a = 0; a < m->d; a++)
        if (fabs(dd[a]) > m->r)
            return 0;
    return 1;
}

/* Scans the 3^d neighbour cells of x at timestep t.  mode 0: returns 1 on the first
 * contact.  mode 1 (subpath): tracks the single contacted group in *g, returns 1 on a
 * second distinct group. */
static int pk_scan(const packing* P, int t, const double* x, int mode, int32_t* g) {
    const or_model* m = P->m;
    const int T = m->T, d = m->d;
    int c0[3], cx[3];
    for (int a = 0; a < d; a++)
        c0[a] = or_cell(m, x[a]);
    int nb = d == 2 ? 9 : 27;
    for (int q = 0; q < nb; q++) {
        int qq = q;
        for (int a = d - 1; a >= 0; a--) {
            int off = qq % 3 - 1;
            qq /= 3;
            cx[a] = (c0[a] + off + T) % T;
        }
        int32_t* head = hg_slot((hgrid*)&P->h, cell_key(P, t, cx), 0);
        if (!head)
            continue;
        for (int32_t n = *head; n >= 0; n = P->nxt[(size_t)n * T + t]) {
            const double* p = P->pos + ((size_t)n * T + t) * d;
            if (!touches(m, x, p))
                continue;
            if (mode == 0)
                return 1;
            if (*g == -1)
                *g = P->gid[n];
            else if (P->gid[n] != *g)
                return 1;
        }
    }
    return 0;
}

typedef struct {
    double a1[3];
    int* b;
    double *c, *s, *X;
} cand_buf;

static int cb_init(cand_buf* cb, const or_model* m) {
    size_t n = (size_t)m->d * m->nw;
    cb->b = (int*)malloc(n * sizeof(int));
    cb->c = (double*)malloc(n * 8);
    cb->s = (double*)malloc(n * 8);
    cb->X = (double*)malloc((size_t)m->T * m->d * 8);
    return (cb->b && cb->c && cb->s && cb->X) ? 0 : -2;
}
static void cb_free(cand_buf* cb) {
    free(cb->b);
    free(cb->c);
    free(cb->s);
    free(cb->X);
}

int64_t or_pack(const or_model* m, int64_t max_candidates, int stop_kind, int64_t stop_batch,
                int64_t window, double stop_rate, int64_t max_n, int64_t* idx, int32_t* gid,
                double* pos, int64_t cap, int64_t* attempts) {
    packing P;
    cand_buf cb;
    if (pk_init(&P, m) || cb_init(&cb, m)) {
        pk_free(&P);
        cb_free(&cb);
        return -2;
    }
    const int T = m->T, d = m->d;
    int64_t nacc_cap = 1024, nacc = 0, lo = 0;
    int64_t* acc = (int64_t*)malloc((size_t)nacc_cap * 8);
    int64_t c = 0, ret = 0;
    for (;; c++) {
        if (stop_kind == 1 && stop_batch > 0 && c > 0 && c % stop_batch == 0 && c >= window) {
            while (lo < nacc && acc[lo] < c - window)
                lo++;
            if ((double)(nacc - lo) < stop_rate * (double)window)
                break;
        }
        if (c >= max_candidates)
            break;
        or_candidate(m, (uint64_t)c, cb.a1, cb.b, cb.c, cb.s, NULL, NULL);
        int hit = 0;
        for (int oi = 0; oi < T && !hit; oi++) {
            int t = m->order[oi];
            or_eval_t(m, t, cb.a1, cb.b, cb.c, cb.s, cb.X + (size_t)t * d);
            hit = pk_scan(&P, t, cb.X + (size_t)t * d, 0, NULL);
        }
        if (hit)
            continue;
        if (pk_insert(&P, cb.X, (int32_t)P.n) < 0) {
            ret = -2;
            break;
        }
        if (nacc == nacc_cap) {
            nacc_cap *= 2;
            acc = (int64_t*)realloc(acc, (size_t)nacc_cap * 8);
        }
        acc[nacc++] = c;
        if (max_n > 0 && nacc >= max_n) {
            c++;
            break;
        }
    }
    if (ret == 0) {
        for (int64_t i = 0; i < nacc && i < cap; i++) {
            if (idx)
                idx[i] = acc[i];
            if (gid)
                gid[i] = P.gid[i];
            if (pos)
                memcpy(pos + (size_t)i * T * d, P.pos + (size_t)i * T * d, (size_t)T * d * 8);
        }
        ret = nacc;
        if (attempts)
            *attempts = c;
    }
    free(acc);
    pk_free(&P);
    cb_free(&cb);
    return ret;
}

/* Serial RSA with the reference's multi-core pattern (cuda/braid_cuda3d.cu:190-212,
 * 1472-1535): blocks of candidates are first tested in parallel against the packing as it
 * stands (read-only), then the survivors are re-tested and admitted serially in index order.
 * A candidate rejected in the parallel pass collides with a path the serial run also holds,
 * so the output equals or_pack's for every block size and thread count. */
int64_t or_pack_mt(const or_model* m, int64_t max_candidates, int stop_kind, int64_t stop_batch,
                   int64_t window, double stop_rate, int threads, int64_t* idx, int64_t cap,
                   int64_t* attempts) {
    packing P;
    cand_buf cb;
    if (pk_init(&P, m) || cb_init(&cb, m)) {
        pk_free(&P);
        cb_free(&cb);
        return -2;
    }
    const int T = m->T, d = m->d;
    const int64_t blk = 8192;
    char* flag = (char*)malloc((size_t)blk);
    int64_t nacc_cap = 1024, nacc = 0, lo = 0;
    int64_t* acc = (int64_t*)malloc((size_t)nacc_cap * 8);
    int64_t c = 0, ret = 0;
    int stop = 0;
    while (!stop && ret == 0) {
        if (stop_kind == 1 && stop_batch > 0 && c > 0 && c % stop_batch == 0 && c >= window) {
            while (lo < nacc && acc[lo] < c - window)
                lo++;
            if ((double)(nacc - lo) < stop_rate * (double)window)
                break;
        }
        if (c >= max_candidates)
            break;
        int64_t e = c + blk < max_candidates ? c + blk : max_candidates;
        if (stop_kind == 1 && stop_batch > 0) {
            int64_t nb = (c / stop_batch + 1) * stop_batch;
            e = e < nb ? e : nb;
        }
        const int64_t nblk = e - c;
        int par_err = 0;
#pragma omp parallel num_threads(threads) reduction(| : par_err)
        {
            cand_buf tb;
            if (cb_init(&tb, m)) {
                par_err = 1;
            } else {
#pragma omp for schedule(dynamic, 64)
                for (int64_t k = 0; k < nblk; k++) {
                    or_candidate(m, (uint64_t)(c + k), tb.a1, tb.b, tb.c, tb.s, NULL, NULL);
                    int hit = 0;
                    for (int oi = 0; oi < T && !hit; oi++) {
                        int t = m->order[oi];
                        or_eval_t(m, t, tb.a1, tb.b, tb.c, tb.s, tb.X + (size_t)t * d);
                        hit = pk_scan(&P, t, tb.X + (size_t)t * d, 0, NULL);
                    }
                    flag[k] = (char)!hit;
                }
                cb_free(&tb);
            }
        }
        if (par_err) {
            ret = -2;
            break;
        }
        for (int64_t k = 0; k < nblk; k++) {
            if (!flag[k])
                continue;
            const int64_t cc = c + k;
            or_candidate(m, (uint64_t)cc, cb.a1, cb.b, cb.c, cb.s, NULL, NULL);
            int hit = 0;
            for (int oi = 0; oi < T && !hit; oi++) {
                int t = m->order[oi];
                or_eval_t(m, t, cb.a1, cb.b, cb.c, cb.s, cb.X + (size_t)t * d);
                hit = pk_scan(&P, t, cb.X + (size_t)t * d, 0, NULL);
            }
            if (hit)
                continue;
            if (pk_insert(&P, cb.X, (int32_t)P.n) < 0) {
                ret = -2;
                break;
            }
            if (nacc == nacc_cap) {
                nacc_cap *= 2;
                acc = (int64_t*)realloc(acc, (size_t)nacc_cap * 8);
            }
            acc[nacc++] = cc;
        }
        c = e;
    }
    if (ret == 0) {
        for (int64_t i = 0; i < nacc && i < cap; i++)
            if (idx)
                idx[i] = acc[i];
        ret = nacc;
        if (attempts)
            *attempts = c;
    }
    free(flag);
    free(acc);
    pk_free(&P);
    cb_free(&cb);
    return ret;
}

/* Touched group of candidate c against the packing P: -1 none, -2 two or more. */
static int32_t sub_group(const or_model* m, const packing* P, cand_buf* cb, uint64_t c) {
    or_candidate(m, c, cb->a1, cb->b, cb->c, cb->s, NULL, NULL);
    int32_t g = -1;
    for (int oi = 0; oi < m->T; oi++) {
        int t = m->order[oi];
        or_eval_t(m, t, cb->a1, cb->b, cb->c, cb->s, cb->X + (size_t)t * m->d);
        if (pk_scan(P, t, cb->X + (size_t)t * m->d, 1, &g))
            return -2;
    }
    return g;
}

typedef struct {
    int64_t* att;
    int64_t* val;
    int64_t n, head, cap, sum_att, sum_val, target;
} rate_window;

/* One windowed-rate stop (cuda/braid_cuda.cu:1008-1037); returns 1 when it fires. */
static int window_push(rate_window* w, int64_t att, int64_t val, int64_t sub_att, double rate) {
    if (!w->target)
        return 0;
    if (w->n == w->cap) {
        int64_t nc = w->cap ? 2 * w->cap : 64;
        int64_t* a = (int64_t*)malloc((size_t)nc * 8);
        int64_t* v = (int64_t*)malloc((size_t)nc * 8);
        for (int64_t k = 0; k < w->n; k++) {
            a[k] = w->att[(w->head + k) % w->cap];
            v[k] = w->val[(w->head + k) % w->cap];
        }
        free(w->att);
        free(w->val);
        w->att = a;
        w->val = v;
        w->head = 0;
        w->cap = nc;
    }
    int64_t tail = (w->head + w->n) % w->cap;
    w->att[tail] = att;
    w->val[tail] = val;
    w->n++;
    w->sum_att += att;
    w->sum_val += val;
    while (w->n > 1 && w->sum_att - w->att[w->head] >= w->target) {
        w->sum_att -= w->att[w->head];
        w->sum_val -= w->val[w->head];
        w->head = (w->head + 1) % w->cap;
        w->n--;
    }
    return w->sum_att >= w->target && sub_att > 1000000 &&
           (double)w->sum_val / (double)w->sum_att < rate;
}

static int64_t window_target(double rate) {
    if (rate <= 0)
        return 0;
    double t = 30.0 / rate;
    t = t < 2e7 ? 2e7 : (t > 5e9 ? 5e9 : t);
    return (int64_t)t;
}

/* Appends one round record {start, size, survivors, admitted, fresh} when there is room. */
static void round_log(int64_t* rounds, int64_t rounds_cap, int64_t* nrounds, int64_t start,
                      int64_t size, int64_t surv, int64_t adm, int64_t fresh) {
    if (!nrounds)
        return;
    if (rounds && *nrounds < rounds_cap) {
        int64_t* r = rounds + 5 * *nrounds;
        r[0] = start;
        r[1] = size;
        r[2] = surv;
        r[3] = adm;
        r[4] = fresh;
    }
    (*nrounds)++;
}

/* Candidates [start, start + size) that touch no path of P (the packing as the round found
 * it; read-only, so tested in parallel) -> *surv (ascending), their number returned. */
static int64_t round_survivors(const or_model* m, const packing* P, int64_t start, int64_t size,
                               char** flag, int64_t* flag_alloc, int64_t** surv,
                               int64_t* surv_alloc) {
    const int T = m->T, d = m->d;
    if (size > *flag_alloc) {
        free(*flag);
        *flag = (char*)malloc((size_t)size);
        *flag_alloc = *flag ? size : 0;
        if (!*flag)
            return -2;
    }
    char* fl = *flag;
    int par_err = 0;
#pragma omp parallel reduction(| : par_err)
    {
        cand_buf tb;
        if (cb_init(&tb, m)) {
            par_err = 1;
        } else {
#pragma omp for schedule(dynamic, 256)
            for (int64_t k = 0; k < size; k++) {
                or_candidate(m, (uint64_t)(start + k), tb.a1, tb.b, tb.c, tb.s, NULL, NULL);
                int hit = 0;
                for (int oi = 0; oi < T && !hit; oi++) {
                    int t = m->order[oi];
                    or_eval_t(m, t, tb.a1, tb.b, tb.c, tb.s, tb.X + (size_t)t * d);
                    hit = pk_scan(P, t, tb.X + (size_t)t * d, 0, NULL);
                }
                fl[k] = (char)!hit;
            }
            cb_free(&tb);
        }
    }
    if (par_err)
        return -2;
    int64_t n = 0;
    for (int64_t k = 0; k < size; k++) {
        if (!fl[k])
            continue;
        if (n == *surv_alloc) {
            int64_t na = *surv_alloc ? 2 * *surv_alloc : 1024;
            int64_t* ns = (int64_t*)realloc(*surv, (size_t)na * 8);
            if (!ns)
                return -2;
            *surv = ns;
            *surv_alloc = na;
        }
        (*surv)[n++] = start + k;
    }
    return n;
}

/* The reference engines' batch adaptation toward ~256 survivors a round
 * (cuda/braid_cuda.cu:920-924, cuda/braid_cuda3d.cu:1540-1543); got is capped at the
 * survivor buffer (2^16 at KMAX_WIGGLE > 16, :760), which never changes the outcome. */
static int64_t adapt_batch(int64_t batch, int64_t surv) {
    const int64_t got = surv < ((int64_t)1 << 16) ? surv : ((int64_t)1 << 16);
    if (got > 512 && batch > 4096)
        return batch / 2;
    if (got < 64 && batch < ((int64_t)1 << 26))
        return batch * 2;
    return batch;
}

int64_t or_pack_ref(const or_model* m, double budget, int rule, double stop_rate, int64_t max_n,
                    int64_t* idx, int32_t* gid, int64_t cap, int64_t* attempts,
                    int64_t* decided, int64_t* rounds, int64_t rounds_cap, int64_t* nrounds) {
    packing P;
    cand_buf cb;
    if ((rule != 2 && rule != 3) || pk_init(&P, m) || cb_init(&cb, m)) {
        pk_free(&P);
        cb_free(&cb);
        return rule != 2 && rule != 3 ? -1 : -2;
    }
    const int T = m->T, d = m->d;
    char* flag = NULL;
    int64_t* surv = NULL;
    int64_t flag_alloc = 0, surv_alloc = 0, nacc = 0, nacc_cap = 1024, ret = 0;
    int64_t* acc = (int64_t*)malloc((size_t)nacc_cap * 8);
    rate_window w = {0};
    w.target = window_target(stop_rate);
    int64_t att = 0, dec = 0, batch = (int64_t)1 << 16;
    /* rule 3: the round after the current one is sized (enqueued) before the current one is
     * drained, and a stop still drains it (cuda/braid_cuda3d.cu:1446-1566) */
    int64_t next = batch, enq = batch;
    int stop = 0;
    if (nrounds)
        *nrounds = 0;
    for (;;) {
        int64_t size;
        if (rule == 2) {
            if ((double)att >= budget || (max_n > 0 && nacc >= max_n))
                break;
            size = batch;
        } else {
            if (next == 0 || (max_n > 0 && nacc >= max_n))
                break;
            size = next;
            next = 0;
            if ((double)enq < budget && !stop) {
                next = batch;
                enq += next;
            }
        }
        const int64_t ns = round_survivors(m, &P, att, size, &flag, &flag_alloc, &surv,
                                           &surv_alloc);
        if (ns < 0) {
            ret = -2;
            break;
        }
        /* admission in index order against the up-to-date packing (host_collides, :703-722) */
        int64_t adm = 0, fresh = 0;
        dec = att + size;
        for (int64_t s = 0; s < ns && ret == 0; s++) {
            or_candidate(m, (uint64_t)surv[s], cb.a1, cb.b, cb.c, cb.s, NULL, NULL);
            int hit = 0;
            for (int oi = 0; oi < T && !hit; oi++) {
                int t = m->order[oi];
                or_eval_t(m, t, cb.a1, cb.b, cb.c, cb.s, cb.X + (size_t)t * d);
                hit = pk_scan(&P, t, cb.X + (size_t)t * d, 0, NULL);
            }
            if (hit)
                continue;
            int64_t f = pk_insert(&P, cb.X, (int32_t)P.n);
            if (f < 0) {
                ret = -2;
                break;
            }
            fresh += f;
            if (nacc == nacc_cap) {
                nacc_cap *= 2;
                acc = (int64_t*)realloc(acc, (size_t)nacc_cap * 8);
            }
            acc[nacc++] = surv[s];
            adm++;
            if (max_n > 0 && nacc >= max_n) {
                dec = surv[s] + 1;
                break;
            }
        }
        if (ret)
            break;
        round_log(rounds, rounds_cap, nrounds, att, size, ns, adm, fresh);
        att += size;
        batch = adapt_batch(batch, ns);
        /* the 2+1 engine's window records the adapted batch (cuda/braid_cuda.cu:930-936), the
         * 3+1 engine's the round just drained (cuda/braid_cuda3d.cu:1549-1551) */
        if (window_push(&w, rule == 2 ? batch : size, adm, att, stop_rate)) {
            if (rule == 2)
                break;
            stop = 1;
        }
    }
    if (ret == 0) {
        for (int64_t i = 0; i < nacc && i < cap; i++) {
            if (idx)
                idx[i] = acc[i];
            if (gid)
                gid[i] = P.gid[i];
        }
        ret = nacc;
        if (attempts)
            *attempts = att;
        if (decided)
            *decided = dec;
    }
    free(acc);
    free(flag);
    free(surv);
    free(w.att);
    free(w.val);
    pk_free(&P);
    cb_free(&cb);
    return ret;
}

void or_positions(const or_model* m, const int64_t* idx, int64_t n, double* out) {
    const int T = m->T, d = m->d;
#pragma omp parallel
    {
        cand_buf tb;
        if (!cb_init(&tb, m)) {
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; i++) {
                or_candidate(m, (uint64_t)idx[i], tb.a1, tb.b, tb.c, tb.s, NULL, NULL);
                for (int t = 0; t < T; t++)
                    or_eval_t(m, t, tb.a1, tb.b, tb.c, tb.s, out + ((size_t)i * T + t) * d);
            }
            cb_free(&tb);
        }
    }
}

int64_t or_pack_sub(const or_model* m, const double* pos0, const int32_t* gid0, int64_t n0,
                    uint64_t first, double budget, double accept_rate, double fill_rate,
                    int64_t* idx, int32_t* gid, int64_t cap, int64_t* attempts,
                    int64_t* filled_new, int64_t* rounds, int64_t rounds_cap, int64_t* nrounds) {
    packing P;
    cand_buf cb;
    if (pk_init(&P, m) || cb_init(&cb, m)) {
        pk_free(&P);
        cb_free(&cb);
        return -2;
    }
    int64_t ret = 0, nsub = 0, fresh = 0, att = 0, batch = 1 << 16;
    uint64_t* surv = NULL;
    int64_t surv_n = 0, surv_alloc = 0, flag_alloc = 0;
    char* flag = NULL;
    rate_window aw = {0}, fw = {0};
    aw.target = window_target(accept_rate);
    fw.target = window_target(fill_rate);
    if (nrounds)
        *nrounds = 0;
    for (int64_t i = 0; i < n0; i++)
        if (pk_insert(&P, pos0 + (size_t)i * m->T * m->d, gid0[i]) < 0)
            ret = -2;
    while (ret == 0 && (double)att < budget) {
        /* round-start test: exactly one group touched in the packing as it stands (the
         * packing is read-only here, so candidates are tested in parallel) */
        if (batch > flag_alloc) {
            free(flag);
            flag = (char*)malloc((size_t)batch);
            flag_alloc = batch;
            if (!flag) {
                ret = -2;
                break;
            }
        }
        int par_err = 0;
#pragma omp parallel reduction(| : par_err)
        {
            cand_buf tb;
            if (cb_init(&tb, m)) {
                par_err = 1;
            } else {
#pragma omp for schedule(dynamic, 256)
                for (int64_t k = 0; k < batch; k++)
                    flag[k] = sub_group(m, &P, &tb, first + (uint64_t)(att + k)) >= 0;
                cb_free(&tb);
            }
        }
        if (par_err) {
            ret = -2;
            break;
        }
        surv_n = 0;
        for (int64_t k = 0; k < batch; k++) {
            if (!flag[k])
                continue;
            if (surv_n == surv_alloc) {
                surv_alloc = surv_alloc ? 2 * surv_alloc : 1024;
                surv = (uint64_t*)realloc(surv, (size_t)surv_alloc * 8);
            }
            surv[surv_n++] = first + (uint64_t)(att + k);
        }
        att += batch;
        /* admission in index order against the up-to-date packing */
        int64_t admitted = 0, round_new = 0;
        for (int64_t s = 0; s < surv_n; s++) {
            int32_t g = sub_group(m, &P, &cb, surv[s]);
            if (g < 0)
                continue;
            int64_t f = pk_insert(&P, cb.X, g);
            if (f < 0) {
                ret = -2;
                break;
            }
            round_new += f;
            if (nsub < cap) {
                if (idx)
                    idx[nsub] = (int64_t)surv[s];
                if (gid)
                    gid[nsub] = g;
            }
            nsub++;
            admitted++;
        }
        fresh += round_new;
        round_log(rounds, rounds_cap, nrounds, att - batch, batch, surv_n, admitted, round_new);
        batch = adapt_batch(batch, surv_n);
        /* the windows record the adapted batch, as the reference does */
        if (window_push(&aw, batch, admitted, att, accept_rate))
            break;
        if (window_push(&fw, batch, round_new, att, fill_rate))
            break;
    }
    if (attempts)
        *attempts = att;
    if (filled_new)
        *filled_new = fresh;
    free(surv);
    free(flag);
    free(aw.att);
    free(aw.val);
    free(fw.att);
    free(fw.val);
    pk_free(&P);
    cb_free(&cb);
    return ret ? ret : nsub;
}

/* Canonical pair predicate at one timestep: adjacent cells on every axis and
 * |torus_delta| <= r on every axis. */
static int pair_hit(const or_model* m, const double* x, const double* y) {
    for (int a = 0; a < m->d; a++) {
        int dc = or_cell(m, x[a]) - or_cell(m, y[a]);
        if (dc < 0)
            dc = -dc;
        if (dc > 1 && dc != m->T - 1)
            return 0;
    }
    return touches(m, x, y);
}

int64_t or_check_rsa(const or_model* m, int64_t M, const int64_t* acc, int64_t na) {
    const int T = m->T, d = m->d;
    cand_buf cb;
    if (cb_init(&cb, m))
        return -2;
    double* all = (double*)malloc((size_t)M * T * d * 8);
    char* member = (char*)calloc((size_t)M, 1);
    if (!all || !member) {
        free(all);
        free(member);
        cb_free(&cb);
        return -2;
    }
    for (int64_t c = 0; c < M; c++) {
        or_candidate(m, (uint64_t)c, cb.a1, cb.b, cb.c, cb.s, NULL, NULL);
        for (int t = 0; t < T; t++)
            or_eval_t(m, t, cb.a1, cb.b, cb.c, cb.s, all + ((size_t)c * T + t) * d);
    }
    int64_t bad = 0;
    for (int64_t i = 0; i < na && !bad; i++) {
        if (acc[i] < 0 || acc[i] >= M || (i > 0 && acc[i] <= acc[i - 1]))
            bad = 1 + (acc[i] >= 0 ? acc[i] : 0);
        else
            member[acc[i]] = 1;
    }
    /* (i) members pairwise collision-free */
    for (int64_t i = 0; i < na && !bad; i++)
        for (int64_t j = 0; j < i && !bad; j++)
            for (int t = 0; t < T; t++)
                if (pair_hit(m, all + ((size_t)acc[i] * T + t) * d,
                             all + ((size_t)acc[j] * T + t) * d)) {
                    bad = 1 + acc[i];
                    break;
                }
    /* (ii) every non-member collides with an earlier member */
    for (int64_t c = 0; c < M && !bad; c++) {
        if (member[c])
            continue;
        int found = 0;
        for (int64_t j = 0; j < na && acc[j] < c && !found; j++)
            for (int t = 0; t < T; t++)
                if (pair_hit(m, all + ((size_t)c * T + t) * d,
                             all + ((size_t)acc[j] * T + t) * d)) {
                    found = 1;
                    break;
                }
        if (!found)
            bad = 1 + c;
    }
    free(all);
    free(member);
    cb_free(&cb);
    return bad;
}

int64_t or_count_clear(const or_model* m, int64_t first, int64_t count, const int64_t* acc,
                       int64_t na) {
    const int T = m->T, d = m->d;
    double* A = (double*)malloc((size_t)(na > 0 ? na : 1) * T * d * 8);
    if (!A)
        return -2;
    or_positions(m, acc, na, A);
    int64_t clear = 0;
    int par_err = 0;
#pragma omp parallel reduction(+ : clear) reduction(| : par_err)
    {
        double* X = (double*)malloc((size_t)T * d * 8);
        cand_buf tb;
        if (!X || cb_init(&tb, m)) {
            par_err = 1;
        } else {
#pragma omp for schedule(dynamic, 1024)
            for (int64_t k = 0; k < count; k++) {
                or_candidate(m, (uint64_t)(first + k), tb.a1, tb.b, tb.c, tb.s, NULL, NULL);
                for (int t = 0; t < T; t++)
                    or_eval_t(m, t, tb.a1, tb.b, tb.c, tb.s, X + (size_t)t * d);
                int hit = 0;
                for (int64_t j = 0; j < na && !hit; j++)
                    for (int t = 0; t < T && !hit; t++)
                        hit = pair_hit(m, X + (size_t)t * d, A + ((size_t)j * T + t) * d);
                clear += !hit;
            }
            cb_free(&tb);
        }
        free(X);
    }
    free(A);
    return par_err ? -2 : clear;
}

int64_t or_pack_knot(const or_model* m, double budget, double nu_attempts, double nu_div,
                     int64_t max_strands, int64_t* idx, int32_t* species, int64_t cap,
                     int64_t* attempts, int64_t* primary_attempts, int64_t* n_primary,
                     int64_t* rounds, int64_t rounds_cap, int64_t* nrounds) {
    packing P;
    cand_buf cb;
    if (nu_div < 1.0 || pk_init(&P, m) || cb_init(&cb, m)) {
        pk_free(&P);
        cb_free(&cb);
        return nu_div < 1.0 ? -1 : -2;
    }
    const int T = m->T, d = m->d;
    or_model mnu = *m; /* the nu species: same model, exclusion (2/T)/nu_div */
    mnu.r = m->r / nu_div;
    char* flag = NULL;
    int64_t *surv = NULL, *acc = NULL;
    int32_t* sp = NULL;
    int64_t flag_alloc = 0, surv_alloc = 0, nacc = 0, nacc_cap = 0, ret = 0, att = 0;
    if (nrounds)
        *nrounds = 0;
    if (primary_attempts)
        *primary_attempts = 0;
    if (n_primary)
        *n_primary = 0;
    for (int ph = 0; ph < 2 && ret == 0; ph++) {
        if (ph == 1 && !(nu_attempts > 0))
            break;
        const or_model* mp = ph == 0 ? m : &mnu;
        P.m = mp;
        const double end = ph == 0 ? budget : budget + nu_attempts;
        int64_t batch = 4096;
        while ((double)att < end && nacc < max_strands) {
            int64_t b = batch;
            if ((double)(att + b) > end)
                b = (int64_t)(end - (double)att);
            if (b == 0)
                break;
            /* survivors against the strands of the batch's start, then index-order admission
             * against the up-to-date strand set (cuda/braid_knot.cu:421-462) */
            const int64_t ns = round_survivors(mp, &P, att, b, &flag, &flag_alloc, &surv, &surv_alloc);
            if (ns < 0) {
                ret = -2;
                break;
            }
            int64_t adm = 0;
            for (int64_t k = 0; k < ns && nacc < max_strands; k++) {
                or_candidate(mp, (uint64_t)surv[k], cb.a1, cb.b, cb.c, cb.s, NULL, NULL);
                int hit = 0;
                for (int oi = 0; oi < T && !hit; oi++) {
                    int t = mp->order[oi];
                    or_eval_t(mp, t, cb.a1, cb.b, cb.c, cb.s, cb.X + (size_t)t * d);
                    hit = pk_scan(&P, t, cb.X + (size_t)t * d, 0, NULL);
                }
                if (hit)
                    continue;
                if (pk_insert(&P, cb.X, (int32_t)P.n) < 0) {
                    ret = -2;
                    break;
                }
                if (nacc == nacc_cap) {
                    nacc_cap = nacc_cap ? 2 * nacc_cap : 1024;
                    acc = (int64_t*)realloc(acc, (size_t)nacc_cap * 8);
                    sp = (int32_t*)realloc(sp, (size_t)nacc_cap * 4);
                }
                acc[nacc] = surv[k];
                sp[nacc] = ph;
                nacc++;
                adm++;
            }
            if (ret)
                break;
            round_log(rounds, rounds_cap, nrounds, att, b, ns, adm, ph);
            att += b;
            if (ns > b / 8 && batch > 4096)
                batch = batch / 4 > 4096 ? batch / 4 : 4096;
            else if (ns < b / 64 && batch < ((int64_t)1 << 24))
                batch *= 4;
        }
        if (ph == 0) {
            if (primary_attempts)
                *primary_attempts = att;
            if (n_primary)
                *n_primary = nacc;
        }
    }
    if (ret == 0) {
        for (int64_t i = 0; i < nacc && i < cap; i++) {
            if (idx)
                idx[i] = acc[i];
            if (species)
                species[i] = sp[i];
        }
        ret = nacc;
        if (attempts)
            *attempts = att;
    }
    P.m = m;
    free(flag);
    free(surv);
    free(acc);
    free(sp);
    pk_free(&P);
    cb_free(&cb);
    return ret;
}
