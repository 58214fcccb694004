// ug_cli_pack.cpp -- the packer's command line: a drop-in for the reference's engine binaries.
//
// Built twice: UG_CLI_DIM=2 gives `braid_ug` (the argv of cuda/braid_cuda.cu:519-646), and
// UG_CLI_DIM=3 gives `braid_ug3d` (cuda/braid_cuda3d.cu:849-1050).  The argv that
// braidlab/engine.py:50-101 builds runs unchanged.  The outputs match the reference's:
//   stderr   the banner, `done: N=<n> in <attempts> attempts[, filled ...]` (the line that
//            braidlab/engine.py:104-112 parses), `sub done (...)`, `dump-params: wrote ...`;
//   --curve  `attempts,n` (or `attempts,n,nsub,filled` with --subpaths), log-spaced rows
//            plus a final one (cuda/braid_cuda.cu:1045-1060);
//   --dump-params  the per-worldline parameter CSV (:1062-1104; 3+1 :1578-1613).
// Everything runs through the C ABI of libugpack.so (include/ugpack.h); a failing call
// prints ug_last_error() and exits 1, like the reference's CK().
//
// Candidates are consumed on the engines' round schedule (rounds of 2^16 adapted on their
// survivors; the last round may pass --attempts) with their --until-accept-rate window, so
// the done: line's attempts count the same rounds (UG_STOP_REF_WINDOW).
//
// Flags beyond the reference's: --precision fp64|split (default split: split-BF16 MFMA
// prefilter + FP64 recheck; both give the same accepts and rounds), --fixed-window (the
// fixed rule: stop at the first 2^20 boundary where the last 2^22 candidates hold fewer
// than rate * 2^22 accepts, rate from --until-accept-rate).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ugpack.h"

#ifndef UG_CLI_DIM
#define UG_CLI_DIM 2
#endif

namespace {

constexpr int kDim = UG_CLI_DIM;

int die(const char* what) {
    fprintf(stderr, "error: %s: %s\n", what, ug_last_error());
    return 1;
}

struct Args {
    int T = 120;
    double budget = 1e8;
    uint64_t seed = 12345;
    const char* curve = nullptr;
    int maxfreq = 0;
    double accept = 0;
    int terms = 2;
    bool pin = false;
    bool subpaths = false;
    double sub_budget = 0, sub_accept = 0, sub_fill = 0;
    int sub_terms = 0;
    long long max_n = 0;
    const char* dump = nullptr;
    bool euclid = false, sparse = false;
    const char* diag = nullptr;
    int precision = UG_PREC_SPLIT_BF16;
    bool fixed_window = false;
};

// Returns -1 to continue, else the exit code.  Unknown flags are skipped, as the reference's
// parsers do.
int parse(int argc, char** argv, Args& a) {
    for (int i = 1; i < argc; i++) {
        const char* f = argv[i];
        auto val = [&]() -> const char* { return i + 1 < argc ? argv[++i] : "0"; };
        if (!strcmp(f, "-t") || !strcmp(f, "--timesteps"))
            a.T = atoi(val());
        else if (!strcmp(f, "--attempts"))
            a.budget = atof(val());
        else if (!strcmp(f, "--seed"))
            a.seed = strtoull(val(), nullptr, 10);
        else if (!strcmp(f, "--uniform") || !strcmp(f, "--torus") || !strcmp(f, "--phase")) {
            // the only sampler / geometry / schema: accepted as no-op guards
        } else if (!strcmp(f, "--smart")) {
            fprintf(stderr,
                    "error: --smart was removed (frequencies are always drawn uniformly now).\n");
            return 1;
        } else if (!strcmp(f, "--curve"))
            a.curve = val();
        else if (!strcmp(f, "--maxfreq"))
            a.maxfreq = atoi(val());
        else if (!strcmp(f, "--until-accept-rate"))
            a.accept = atof(val());
        else if (!strcmp(f, "--terms"))
            a.terms = atoi(val());
        else if (!strcmp(f, "--pin-sin1"))
            a.pin = true;
        else if (!strcmp(f, "--max-n"))
            a.max_n = atoll(val());
        else if (!strcmp(f, "--dump-params"))
            a.dump = val();
        else if (!strcmp(f, "--precision")) {
            const char* p = val();
            if (!strcmp(p, "fp64"))
                a.precision = UG_PREC_FP64;
            else if (!strcmp(p, "split"))
                a.precision = UG_PREC_SPLIT_BF16;
            else {
                fprintf(stderr, "error: --precision must be fp64|split\n");
                return 1;
            }
        } else if (!strcmp(f, "--fixed-window"))
            a.fixed_window = true;
        else if (kDim == 2 && !strcmp(f, "--subpaths"))
            a.subpaths = true;
        else if (kDim == 2 && !strcmp(f, "--sub-attempts"))
            a.sub_budget = atof(val());
        else if (kDim == 2 && !strcmp(f, "--sub-until-accept-rate"))
            a.sub_accept = atof(val());
        else if (kDim == 2 && !strcmp(f, "--sub-until-fill-rate"))
            a.sub_fill = atof(val());
        else if (kDim == 2 && !strcmp(f, "--sub-terms"))
            a.sub_terms = atoi(val());
        else if (kDim == 3 && !strcmp(f, "--angle-sample")) {
            fprintf(stderr, "error: --angle-sample was removed along with the hard-wall "
                            "geometry it compensated for; the sin1 offset is always uniform.\n");
            return 1;
        } else if (kDim == 3 && !strcmp(f, "--euclid-collision"))
            a.euclid = true;
        else if (kDim == 3 && !strcmp(f, "--sparse"))
            a.sparse = true;
        else if (kDim == 3 && !strcmp(f, "--diag"))
            a.diag = val();
        else if (kDim == 3 && !strcmp(f, "--probe-n"))
            (void)val();
    }
    return -1;
}

// One dump row (cuda/braid_cuda.cu:1062-1104; 3+1 cuda/braid_cuda3d.cu:1578-1613) from the
// parameter record [axis][a1, (amplitude, frequency, phase) x nw].
void dump_row(FILE* f, const double* rec, int nw, int gid, bool with_gid) {
    const int stride = 1 + 3 * nw;
    auto amp = [&](int a, int j) { return rec[a * stride + 1 + 3 * j]; };
    auto frq = [&](int a, int j) { return rec[a * stride + 2 + 3 * j]; };
    auto phs = [&](int a, int j) { return rec[a * stride + 3 + 3 * j]; };
    auto a1 = [&](int a) { return rec[a * stride]; };
    if (nw == 1) {
        if (kDim == 2)
            fprintf(f, "%.10g,%.10g,%.0f,%.0f,%.10g,%.10g,%.10g,%.10g", amp(0, 0), amp(1, 0), frq(0, 0),
                    frq(1, 0), a1(0), a1(1), phs(0, 0), phs(1, 0));
        else
            fprintf(f, "%.10g,%.10g,%.10g,%.0f,%.0f,%.0f,%.10g,%.10g,%.10g,%.10g,%.10g,%.10g", amp(0, 0),
                    amp(1, 0), amp(2, 0), frq(0, 0), frq(1, 0), frq(2, 0), a1(0), a1(1), a1(2), phs(0, 0),
                    phs(1, 0), phs(2, 0));
    } else {
        fprintf(f, "%.10g", a1(0));
        for (int a = 1; a < kDim; a++)
            fprintf(f, ",%.10g", a1(a));
        for (int j = 0; j < nw; j++)
            for (int a = 0; a < kDim; a++)
                fprintf(f, ",%.10g,%.0f,%.10g", amp(a, j), frq(a, j), phs(a, j));
    }
    if (with_gid)
        fprintf(f, ",%d", gid);
    fprintf(f, "\n");
}

int write_dump(ug_pack* h, const char* path, int nw, bool with_gid) {
    int64_t n = 0;
    if (ug_pack_accepted(h, nullptr, nullptr, 0, &n) != UG_OK)
        return die("ug_pack_accepted");
    std::vector<int64_t> idx((size_t)n);
    std::vector<int32_t> gid((size_t)n);
    if (n > 0 && ug_pack_accepted(h, idx.data(), gid.data(), n, &n) != UG_OK)
        return die("ug_pack_accepted");
    FILE* f = fopen(path, "w");
    if (!f) {
        fprintf(stderr, "error: cannot write %s\n", path);
        return 1;
    }
    static const char* ax[3] = {"x", "y", "w"};
    if (nw == 1) {
        fprintf(f, kDim == 2 ? "ax,ay,bx,by,ax2,ay2,fx,fy" : "ax,ay,aw,bx,by,bw,ax2,ay2,aw2,fx,fy,fw");
    } else {
        fprintf(f, "ax2,ay2%s", kDim == 3 ? ",aw2" : "");
        for (int j = 1; j <= nw; j++)
            for (int a = 0; a < kDim; a++)
                fprintf(f, ",a%s_%d,b%s_%d,f%s_%d", ax[a], j, ax[a], j, ax[a], j);
    }
    fprintf(f, "%s\n", with_gid ? ",gid" : "");
    const size_t rec = (size_t)kDim * (1 + 3 * nw);
    const int64_t chunk = 4096;
    std::vector<double> buf((size_t)chunk * rec);
    for (int64_t i0 = 0; i0 < n; i0 += chunk) {
        const int64_t k = std::min(chunk, n - i0);
        if (ug_pack_params(h, i0, k, buf.data()) != UG_OK) {
            fclose(f);
            return die("ug_pack_params");
        }
        for (int64_t i = 0; i < k; i++)
            dump_row(f, buf.data() + (size_t)i * rec, nw, gid[(size_t)(i0 + i)], with_gid);
    }
    fclose(f);
    fprintf(stderr, "dump-params: wrote %lld worldlines to %s\n", (long long)n, path);
    return 0;
}

int write_curve(ug_pack* h, const char* path, const ug_pack_status& st, bool four) {
    int64_t rows = 0;
    if (ug_pack_curve(h, nullptr, 0, &rows) != UG_OK)
        return die("ug_pack_curve");
    std::vector<int64_t> c((size_t)rows * 4);
    if (rows > 0 && ug_pack_curve(h, c.data(), rows, &rows) != UG_OK)
        return die("ug_pack_curve");
    c.push_back(st.attempts);
    c.push_back(st.n);
    c.push_back(st.nsub);
    c.push_back(st.filled);
    rows++;
    FILE* f = fopen(path, "w");
    if (!f) {
        fprintf(stderr, "error: cannot write %s\n", path);
        return 1;
    }
    fprintf(f, four ? "attempts,n,nsub,filled\n" : "attempts,n\n");
    for (int64_t r = 0; r < rows; r++) {
        const int64_t* v = c.data() + 4 * r;
        if (four)
            fprintf(f, "%lld,%lld,%lld,%lld\n", (long long)v[0], (long long)v[1], (long long)v[2],
                    (long long)v[3]);
        else
            fprintf(f, "%lld,%lld\n", (long long)v[0], (long long)v[1]);
    }
    fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    const int pr = parse(argc, argv, a);
    if (pr >= 0)
        return pr;
    if (a.maxfreq != 0 && a.maxfreq != a.T) {
        fprintf(stderr,
                "error: --maxfreq %d: the max frequency is hard-coded to T=%d "
                "(the full Nyquist band).\n",
                a.maxfreq, a.T);
        return 1;
    }
    if (a.diag) {
        fprintf(stderr, "error: --diag is not supported by this engine\n");
        return 1;
    }
    if (a.terms < 2)
        a.terms = 2;
    const unsigned modmax = (unsigned)std::max(a.T - 1, 2);
    ug_pack_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.dim = kDim;
    cfg.T = a.T;
    cfg.terms = a.terms;
    cfg.precision = a.precision;
    cfg.seed = a.seed;
    cfg.pin_sin1 = a.pin ? 1 : 0;
    cfg.max_attempts = a.budget;
    cfg.stop_rate = a.accept;
    // the engines' round schedule (a last round past --attempts included), and with
    // --until-accept-rate their window; --fixed-window swaps in the fixed 2^20/2^22 rule
    cfg.stop_kind = a.fixed_window && a.accept > 0 ? UG_STOP_FIXED_WINDOW : UG_STOP_REF_WINDOW;
    cfg.stop_batch = 1 << 20;
    cfg.stop_window = 1 << 22;
    cfg.max_n = a.max_n;
    cfg.euclid = a.euclid ? 1 : 0;
    cfg.sparse = a.sparse ? 1 : 0;
    ug_pack* h = nullptr;
    if (ug_pack_create(&cfg, nullptr, &h) != UG_OK) {
        fprintf(stderr, "%s\n", ug_last_error());
        return 1;
    }
    if (a.pin)
        fprintf(stderr, "pin-sin1: comoving offsets pinned to zero\n");
    const double T = a.T;
    const double ncell = kDim == 2 ? T * T * T : T * T * T * T;
    if (kDim == 2) {
        fprintf(stderr, "2+1 (torus, phase): T=%d  modmax=%u (maxfreq=%u)  terms=%d\n", a.T, modmax,
                modmax + 1, a.terms);
        if (a.subpaths)
            fprintf(stderr, "subpaths: terms=%d  max-n=%lld\n", a.sub_terms > 0 ? a.sub_terms : a.terms,
                    a.max_n);
    } else {
        // the dense cell table holds T^3 rows of T + 2 heads (one halo cell each side); the
        // sparse grid holds the occupied cells only
        if (a.sparse)
            fprintf(stderr, "3+1 (torus, phase) (sparse grid): T=%d  modmax=%u (maxfreq=%u)  terms=%d\n", a.T,
                    modmax, modmax + 1, a.terms);
        else
            fprintf(stderr,
                    "3+1 (torus, phase): T=%d  modmax=%u (maxfreq=%u)  terms=%d  grid cells=%.2e  (~%.1f GB)\n",
                    a.T, modmax, modmax + 1, a.terms, ncell, T * T * T * (T + 2) * 4.0 / 1e9);
    }
    ug_pack_status st;
    if (ug_pack_run(h, 0, &st) != UG_OK)
        return die("ug_pack_run");
    if (kDim == 2)
        fprintf(stderr, "done: N=%lld in %lld attempts, filled %lld/%.0f cells (%.3f%%)\n", (long long)st.n,
                (long long)st.attempts, (long long)st.filled, ncell, 100.0 * (double)st.filled / ncell);
    else
        fprintf(stderr, "done: N=%lld in %lld attempts\n", (long long)st.n, (long long)st.attempts);
    if (a.subpaths) {
        ug_sub_config sc;
        memset(&sc, 0, sizeof sc);
        sc.sub_terms = a.sub_terms > 0 ? std::max(a.sub_terms, 2) : 0;
        sc.max_attempts = a.sub_budget > 0 ? a.sub_budget : a.budget;
        sc.accept_rate = a.sub_accept;
        sc.fill_rate = a.sub_fill;
        if (ug_pack_run_sub(h, &sc, 0, &st) != UG_OK)
            return die("ug_pack_run_sub");
        const char* why = st.sub_done == UG_SUB_ACCEPT_RATE ? "accept-rate"
                          : st.sub_done == UG_SUB_FILL_RATE ? "fill-rate"
                                                             : "attempt budget";
        fprintf(stderr, "sub done (%s): Nsub=%lld in %lld attempts, filled %lld/%.0f cells (%.3f%%)\n", why,
                (long long)st.nsub, (long long)st.sub_attempts, (long long)st.filled, ncell,
                100.0 * (double)st.filled / ncell);
    }
    int rc = 0;
    if (a.curve)
        rc = write_curve(h, a.curve, st, a.subpaths);
    if (!rc && a.dump)
        rc = write_dump(h, a.dump, a.terms - 1, a.subpaths);
    ug_pack_destroy(h);
    return rc;
}
