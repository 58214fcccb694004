// ug_cli_knot.cpp -- `braid_knot_ug`: a drop-in for the reference's braid_knot tool.
//
// The argv of cuda/braid_knot.cu:255-306 (-t, --dim, --terms, --pin-sin1, --seed,
// --attempts, --nu-attempts, --nu-div, --curve, --dump; an unknown flag is an error) and its
// outputs: the stderr banner, one `phase <s> done: N=<n> at <attempts> attempts (cell=<c>)`
// line per phase and `done: N=<n> (primary <p>, nu <q>) in <attempts> attempts`; the curve
// CSV `attempt,n,species` with one row per admission at its exact attempt (index + 1); the
// dump with one row per strand: species, attempt, then per axis a2 and the nw (b, a, f)
// triples.  Candidates come from this packer's Philox stream (its canonical draw, so slot j
// carries frequency j + 2 in the full band), evaluated and compared in FP64 through
// libugpack.so (include/ugpack.h).  Two flags beyond the reference's: --precision fp64|split
// (default split: full-band models go through the split-BF16 prefilter first; the same
// strands either way), and --rng philox|xoshiro-ref: xoshiro-ref draws the reference tool's
// own candidates (its xoshiro256** stream and draw order, cuda/braid_knot.cu:151-166,207),
// so the strands can be compared with a reference run of the same seed; it runs in the FP64
// form, which it selects unless --precision says otherwise.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "ugpack.h"

namespace {

int die(const char* what) {
    fprintf(stderr, "error: %s: %s\n", what, ug_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    int T = 32, dim = 2, terms = 0, pin = 0;
    unsigned long long seed = 1;
    double budget = 1e10, nu_attempts = 0.0, nu_div = 1.0;
    const char* curve_path = nullptr;
    const char* dump_path = nullptr;
    int precision = UG_PREC_SPLIT_BF16, rng = UG_RNG_PHILOX;
    bool precision_set = false;
    for (int i = 1; i < argc; i++) {
        const char* f = argv[i];
        const bool has = i + 1 < argc;
        if (!strcmp(f, "-t") && has)
            T = atoi(argv[++i]);
        else if (!strcmp(f, "--dim") && has)
            dim = atoi(argv[++i]);
        else if (!strcmp(f, "--terms") && has)
            terms = atoi(argv[++i]);
        else if (!strcmp(f, "--pin-sin1"))
            pin = 1;
        else if (!strcmp(f, "--seed") && has)
            seed = strtoull(argv[++i], nullptr, 10);
        else if (!strcmp(f, "--attempts") && has)
            budget = atof(argv[++i]);
        else if (!strcmp(f, "--nu-attempts") && has)
            nu_attempts = atof(argv[++i]);
        else if (!strcmp(f, "--nu-div") && has)
            nu_div = atof(argv[++i]);
        else if (!strcmp(f, "--curve") && has)
            curve_path = argv[++i];
        else if (!strcmp(f, "--dump") && has)
            dump_path = argv[++i];
        else if (!strcmp(f, "--precision") && has) {
            const char* p = argv[++i];
            if (!strcmp(p, "fp64"))
                precision = UG_PREC_FP64;
            else if (!strcmp(p, "split"))
                precision = UG_PREC_SPLIT_BF16;
            else {
                fprintf(stderr, "error: --precision must be fp64|split\n");
                return 1;
            }
            precision_set = true;
        }
        else if (!strcmp(f, "--rng") && has) {
            const char* r = argv[++i];
            if (!strcmp(r, "philox"))
                rng = UG_RNG_PHILOX;
            else if (!strcmp(r, "xoshiro-ref"))
                rng = UG_RNG_XOSHIRO_REF;
            else {
                fprintf(stderr, "error: --rng must be philox|xoshiro-ref\n");
                return 1;
            }
        }
        else {
            fprintf(stderr, "unknown flag %s\n", f);
            return 1;
        }
    }
    if (rng == UG_RNG_XOSHIRO_REF && !precision_set)
        precision = UG_PREC_FP64;
    if (nu_div < 1.0) {
        fprintf(stderr, "error: --nu-div must be >= 1\n");
        return 1;
    }
    if (dim != 2 && dim != 3) {
        fprintf(stderr, "error: --dim must be 2 or 3\n");
        return 1;
    }
    if (terms == 0)
        terms = T;  // the full band by default
    const unsigned modmax = (unsigned)(T - 1 > 2 ? T - 1 : 2);
    const int nw = terms - 1;
    if (nw < 1 || (unsigned)nw > modmax) {
        fprintf(stderr, "error: terms %d exceeds the pool %u\n", terms, modmax);
        return 1;
    }
    ug_knot_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.dim = dim;
    cfg.T = T;
    cfg.terms = terms;
    cfg.pin_sin1 = pin;
    cfg.seed = seed;
    cfg.max_attempts = budget;
    cfg.nu_attempts = nu_attempts;
    cfg.nu_div = nu_div;
    cfg.precision = precision;
    cfg.rng_kind = rng;
    ug_knot* h = nullptr;
    if (ug_knot_create(&cfg, nullptr, &h) != UG_OK) {
        fprintf(stderr, "%s\n", ug_last_error());
        return 1;
    }
    fprintf(stderr,
            "braid_knot: dim=%d T=%d terms=%d pin=%d seed=%llu budget=%.3g nu=%.3g nu-div=%.3g "
            "(exact-index admission, deterministic)\n",
            dim, T, terms, pin, seed, budget, nu_attempts, nu_div);
    ug_knot_status st;
    if (ug_knot_run(h, 0, &st) != UG_OK)
        return die("ug_knot_run");
    const double cell = 2.0 / T;
    fprintf(stderr, "phase 0 done: N=%lld at %llu attempts (cell=%.6g)\n", (long long)st.n_primary,
            (unsigned long long)st.primary_attempts, cell);
    if (nu_attempts > 0)
        fprintf(stderr, "phase 1 done: N=%lld at %llu attempts (cell=%.6g)\n", (long long)st.n,
                (unsigned long long)st.attempts, cell / nu_div);
    int64_t n = 0;
    if (ug_knot_strands(h, nullptr, nullptr, 0, &n) != UG_OK)
        return die("ug_knot_strands");
    std::vector<int64_t> idx((size_t)n);
    std::vector<int32_t> species((size_t)n);
    if (n > 0 && ug_knot_strands(h, idx.data(), species.data(), n, &n) != UG_OK)
        return die("ug_knot_strands");
    int rc = 0;
    if (curve_path) {
        FILE* f = fopen(curve_path, "w");
        if (!f) {
            fprintf(stderr, "error: cannot write %s\n", curve_path);
            rc = 1;
        } else {
            fprintf(f, "attempt,n,species\n");
            for (int64_t i = 0; i < n; i++)
                fprintf(f, "%llu,%lld,%d\n", (unsigned long long)idx[(size_t)i] + 1, (long long)(i + 1),
                        species[(size_t)i]);
            fclose(f);
        }
    }
    if (!rc && dump_path) {
        FILE* f = fopen(dump_path, "w");
        if (!f) {
            fprintf(stderr, "error: cannot write %s\n", dump_path);
            rc = 1;
        } else {
            fprintf(f, "# species,attempt, then per axis: a2, then nw x (b,a,f); dim=%d T=%d nw=%d\n", dim, T, nw);
            const size_t row = (size_t)dim * (1 + 3 * nw);
            std::vector<double> par(row);
            for (int64_t i = 0; i < n && !rc; i++) {
                if (ug_knot_params(h, i, 1, par.data()) != UG_OK) {
                    rc = die("ug_knot_params");
                    break;
                }
                fprintf(f, "%d,%llu", species[(size_t)i], (unsigned long long)idx[(size_t)i] + 1);
                for (int a = 0; a < dim; a++) {
                    const double* r = par.data() + (size_t)a * (1 + 3 * nw);
                    fprintf(f, ",%.17g", r[0]);
                    for (int j = 0; j < nw; j++)
                        fprintf(f, ",%d,%.17g,%.17g", (int)r[2 + 3 * j], r[1 + 3 * j], r[3 + 3 * j]);
                }
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }
    fprintf(stderr, "done: N=%lld (primary %lld, nu %lld) in %llu attempts\n", (long long)st.n,
            (long long)st.n_primary, (long long)(st.n - st.n_primary), (unsigned long long)st.attempts);
    ug_knot_destroy(h);
    return rc;
}
