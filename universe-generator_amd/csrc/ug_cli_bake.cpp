// ug_cli_bake.cpp -- `frame_bake_ug`, a drop-in for the reference's frame_bake executable
// (cuda/frame_bake.cu:398-635): same flags, same stderr lines, same BRF1 bytes layout
// (read by braidlab/frames.py:88-144).  The hit search runs through the C ABI of
// libugpack.so (ug_paths_create, ug_frame_bake_lightcone).
//
//   frame_bake_ug --dump D.csv --observer-path N | --observer-point x,y [...]
//                 [--from 0.25] [--to 0.75] [--frames 1] [--front speed|fit|budget]
//                 [--metric square|circle] [--speed 0.33] [--zmin 0.02] [--wraps K]
//                 [--out frames.bin]
//
// The reference stages hits through a fixed 2^24-record buffer and fails beyond it; here
// the device buffer grows to the exact count (a size query first), so no cap applies.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ugpack.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Dump {
    int n = 0, nw = 1;
    std::vector<double> a2, amp, freq, phase;  // a2[n][2], others [n][2][nw]
};

std::vector<std::string> split_csv(const std::string& line) {
    std::vector<std::string> out;
    std::string cell;
    std::istringstream ss(line);
    while (std::getline(ss, cell, ','))
        out.push_back(cell);
    if (!line.empty() && line.back() == ',')
        out.push_back("");
    return out;
}

int col_of(const std::vector<std::string>& h, const std::string& name) {
    for (size_t i = 0; i < h.size(); i++)
        if (h[i] == name)
            return (int)i;
    return -1;
}

void chomp(std::string& s) {
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r'))
        s.pop_back();
}

// Parameter dump -> per-path terms, the column mapping of load_dump
// (cuda/frame_bake.cu:109-190): multi-term "ax2,ay2,ax_j,bx_j,fx_j,ay_j,by_j,fy_j..." or the
// legacy single-term "ax,ay,bx,by,ax2,ay2[,fx,fy]" (missing phases read as 0).
bool load_dump(const char* path, Dump& P) {
    std::ifstream f(path);
    if (!f) {
        fprintf(stderr, "error: cannot open dump %s\n", path);
        return false;
    }
    std::string line;
    if (!std::getline(f, line)) {
        fprintf(stderr, "error: empty dump %s\n", path);
        return false;
    }
    chomp(line);
    const std::vector<std::string> h = split_csv(line);
    if (col_of(h, "aw") >= 0 || col_of(h, "aw_1") >= 0) {
        fprintf(stderr, "error: %s is a 3+1 dump; frame_bake is 2+1 only\n", path);
        return false;
    }
    const bool multi = col_of(h, "ax_1") >= 0;
    int nw = 1;
    while (multi && col_of(h, "ax_" + std::to_string(nw + 1)) >= 0)
        nw++;
    P.nw = nw;
    static const char* axis[2] = {"x", "y"};
    std::vector<int> ca(2 * nw), cb(2 * nw), cf(2 * nw);
    int c2[2];
    for (int a = 0; a < 2; a++) {
        c2[a] = col_of(h, std::string("a") + axis[a] + "2");
        for (int j = 0; j < nw; j++) {
            const std::string s = multi ? "_" + std::to_string(j + 1) : "";
            ca[a * nw + j] = col_of(h, std::string("a") + axis[a] + s);
            cb[a * nw + j] = col_of(h, std::string("b") + axis[a] + s);
            cf[a * nw + j] = col_of(h, std::string("f") + axis[a] + s);
        }
    }
    if (c2[0] < 0 || c2[1] < 0 || ca[0] < 0) {
        fprintf(stderr, "error: unrecognized dump header in %s\n", path);
        return false;
    }
    while (std::getline(f, line)) {
        chomp(line);
        if (line.empty())
            continue;
        const std::vector<std::string> cells = split_csv(line);
        auto val = [&](int c) { return c >= 0 && c < (int)cells.size() ? atof(cells[c].c_str()) : 0.0; };
        for (int a = 0; a < 2; a++)
            P.a2.push_back(val(c2[a]));
        for (int a = 0; a < 2; a++)
            for (int j = 0; j < nw; j++) {
                P.amp.push_back(val(ca[a * nw + j]));
                P.freq.push_back(val(cb[a * nw + j]));
                P.phase.push_back(val(cf[a * nw + j]));
            }
        P.n++;
    }
    return true;
}

template <typename V>
void put(FILE* f, const V& v) {
    fwrite(&v, sizeof v, 1, f);
}

}  // namespace

int main(int argc, char** argv) {
    const char* dump_path = nullptr;
    const char* out_path = "frames.bin";
    double from_pi = 0.25, to_pi = 0.75, zmin_pi = 0.02, speed = 0.33;
    int n_frames = 1, wraps = -1;
    std::string front = "fit", metric = "square";
    std::vector<ug_observer> observers;
    for (int i = 1; i < argc; i++) {
        const char* f = argv[i];
        auto val = [&]() -> const char* { return i + 1 < argc ? argv[++i] : "0"; };
        if (!strcmp(f, "--dump"))
            dump_path = val();
        else if (!strcmp(f, "--out"))
            out_path = val();
        else if (!strcmp(f, "--from"))
            from_pi = atof(val());
        else if (!strcmp(f, "--to"))
            to_pi = atof(val());
        else if (!strcmp(f, "--frames"))
            n_frames = atoi(val());
        else if (!strcmp(f, "--zmin"))
            zmin_pi = atof(val());
        else if (!strcmp(f, "--front"))
            front = val();
        else if (!strcmp(f, "--metric"))
            metric = val();
        else if (!strcmp(f, "--speed"))
            speed = atof(val());
        else if (!strcmp(f, "--wraps"))
            wraps = atoi(val());
        else if (!strcmp(f, "--observer-path")) {
            observers.push_back(ug_observer{0, atoi(val()), 0.0, 0.0});
        } else if (!strcmp(f, "--observer-point")) {
            double x = 0.0, y = 0.0;
            if (sscanf(val(), "%lf,%lf", &x, &y) != 2) {
                fprintf(stderr, "error: --observer-point wants x,y\n");
                return 1;
            }
            observers.push_back(ug_observer{1, -1, x, y});
        } else {
            fprintf(stderr, "error: unknown flag %s\n", f);
            return 1;
        }
    }
    if (!dump_path || observers.empty() || n_frames < 1) {
        fprintf(stderr,
                "usage: frame_bake_ug --dump D.csv --observer-path N|--observer-point x,y [...] "
                "--from 0.25 --to 0.75 --frames 60 [--front speed|fit|budget] [--metric "
                "square|circle] [--speed c] [--zmin 0.02] [--wraps K] --out F.bin\n");
        return 1;
    }
    if (front != "speed" && front != "fit" && front != "budget") {
        fprintf(stderr, "error: --front must be speed|fit|budget\n");
        return 1;
    }
    const int front_code = front == "speed" ? 0 : front == "fit" ? 1 : 2;
    Dump P;
    if (!load_dump(dump_path, P))
        return 1;
    fprintf(stderr, "loaded %d paths, %d terms/axis from %s\n", P.n, P.nw, dump_path);
    for (const ug_observer& o : observers)
        if (o.type == 0 && (o.path < 0 || o.path >= P.n)) {
            fprintf(stderr, "error: observer path %d out of range\n", o.path);
            return 1;
        }
    ug_paths* paths = nullptr;
    if (ug_paths_create(2, P.n, P.nw, P.a2.data(), P.amp.data(), P.freq.data(), P.phase.data(), nullptr,
                        &paths) != UG_OK) {
        fprintf(stderr, "error: %s\n", ug_last_error());
        return 1;
    }
    FILE* out = fopen(out_path, "wb");
    if (!out) {
        fprintf(stderr, "error: cannot write %s\n", out_path);
        return 1;
    }
    const double z_min = fmax(2e-3, zmin_pi * kPi);
    const bool cheb = front_code == 2 || metric == "square";
    fwrite("BRF1", 1, 4, out);
    put(out, (uint32_t)P.n);
    put(out, (uint32_t)observers.size());
    put(out, (uint32_t)n_frames);
    put(out, (uint32_t)front_code);
    put(out, (uint32_t)(cheb ? 1 : 0));
    put(out, speed);
    put(out, z_min);
    put(out, (int32_t)wraps);
    for (const ug_observer& o : observers) {
        put(out, (uint32_t)o.type);
        put(out, (int32_t)(o.type == 0 ? o.path : -1));
        put(out, o.px);
        put(out, o.py);
    }
    ug_front fr;
    memset(&fr, 0, sizeof fr);
    fr.kind = front_code;
    fr.cheb = cheb ? 1 : 0;
    fr.speed = speed;
    fr.z_min = z_min;
    fr.wraps = wraps;
    ug_hit* d_hits = nullptr;
    uint64_t cap = 0;
    std::vector<ug_hit> hits;
    int rc = 0;
    for (int fi = 0; fi < n_frames && !rc; fi++) {
        const double z_obs =
            fmin((from_pi + (to_pi - from_pi) * (n_frames == 1 ? 0.0 : (double)fi / (n_frames - 1))) * kPi,
                 kPi - 1e-3);
        if (z_min >= z_obs) {
            fprintf(stderr, "error: rewind floor %.4f >= z_obs %.4f\n", z_min, z_obs);
            rc = 1;
            break;
        }
        ug_frame_info info;
        for (size_t oi = 0; oi < observers.size() && !rc; oi++) {
            int32_t r = ug_frame_bake_lightcone(paths, &fr, z_obs, &observers[oi], d_hits, cap, &info);
            if (r == UG_EOVERFLOW) {  // grow the device buffer to the exact count and rerun
                if (d_hits)
                    (void)hipFree(d_hits);
                d_hits = nullptr;
                cap = info.n_hits + info.n_hits / 4 + 1024;
                if (hipMalloc((void**)&d_hits, cap * sizeof(ug_hit)) != hipSuccess) {
                    fprintf(stderr, "error: cannot allocate %llu hits\n", (unsigned long long)cap);
                    rc = 1;
                    break;
                }
                r = ug_frame_bake_lightcone(paths, &fr, z_obs, &observers[oi], d_hits, cap, &info);
            }
            if (r != UG_OK) {
                fprintf(stderr, "error: %s\n", ug_last_error());
                rc = 1;
                break;
            }
            if (oi == 0) {
                put(out, info.z_obs);
                put(out, info.chi_max);
                put(out, (int32_t)info.K);
            }
            hits.resize((size_t)info.n_hits);
            if (info.n_hits > 0 &&
                hipMemcpy(hits.data(), d_hits, info.n_hits * sizeof(ug_hit), hipMemcpyDeviceToHost) != hipSuccess) {
                fprintf(stderr, "error: hit copy failed\n");
                rc = 1;
                break;
            }
            put(out, info.ox);
            put(out, info.oy);
            put(out, info.betax);
            put(out, info.betay);
            put(out, (uint64_t)info.n_hits);
            fwrite(hits.data(), sizeof(ug_hit), hits.size(), out);
        }
        if (!rc)
            fprintf(stderr, "frame %d/%d: z_obs=%.4f chiMax=%.3f K=%d\n", fi + 1, n_frames, z_obs, info.chi_max,
                    info.K);
    }
    fclose(out);
    if (d_hits)
        (void)hipFree(d_hits);
    ug_paths_destroy(paths);
    if (!rc)
        fprintf(stderr, "done: wrote %s\n", out_path);
    return rc;
}
