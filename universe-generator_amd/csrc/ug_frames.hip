// ug_frames.hip -- frame baker of the packer's worldlines (2+1 past light cone, BRF1 hits,
// and equal-time slices with density rasters).
//
// Reference: cuda/frame_bake.cu (grid_kernel :376-386, hits_kernel :281-371, host driver
// :398-635) and its Python oracle braidlab/frames.py:246-362 (bake_reference).  An
// observation (observer, z_obs) freezes a pulse centre at the observer's comoving position,
// rewinds to z_min while a front expands (conformal chi = c (eta(z_obs) - eta(z)),
// eta = ln tan(z/2), or the wiggle-budget reach), and images every worldline (and every
// torus copy shifted by (2nx, 2ny), |n| <= K) where it crosses the front: a sign change of
// dist - chi on the rewind grid refined by 24 bisection steps with fresh evaluations.
//
// Here worldlines are held once as dense per-frequency coefficients,
//   X(z) = a2 + [sum_k A_k sin(kz) + B_k (cos(kz) - 1)] / sin z,
//   A_k = sum_{b_j = k} a_j cos f_j,  B_k = sum_{b_j = k} a_j sin f_j,
// evaluated with a half-angle rotation recurrence (sin kz = 2 s_k c_k, cos kz - 1 = -2 s_k^2,
// (s_k, c_k) = (sin ku, cos ku), u = z/2), so no per-term sin() is needed.  The hit search
// runs one wave per worldline with its rewind rows staged in LDS and one lane per torus
// image; a count pass and a prefix scan give every worldline a fixed output range, so hits
// come out in (path, image, row) order rather than atomic order.
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ug_internal.h"

using namespace ug;

namespace {

struct DevPaths {
    int n, dim, kmax;    // kmax: largest frequency; coefficient rows k = 0..kmax
    const double* a2;    // [n][dim]
    const double* A;     // [n][dim][kmax+1]
    const double* B;     // [n][dim][kmax+1]
};

// Unwrapped comoving position of path p at z (x[0..dim)).
__device__ __forceinline__ void comov(const DevPaths& P, int p, double z, double* x) {
    const double u = 0.5 * z;
    double s1, c1;
    sincos(u, &s1, &c1);
    const double sz = 2.0 * s1 * c1;
    double acc[3] = {0.0, 0.0, 0.0};
    double sk = 0.0, ck = 1.0;
    const size_t base = (size_t)p * P.dim * (P.kmax + 1);
    for (int k = 1; k <= P.kmax; k++) {
        const double sn = sk * c1 + ck * s1;
        const double cn = ck * c1 - sk * s1;
        sk = sn;
        ck = cn;
        const double sinkz = 2.0 * sk * ck;
        const double cosm1 = -2.0 * sk * sk;
        for (int a = 0; a < P.dim; a++) {
            const size_t o = base + (size_t)a * (P.kmax + 1) + k;
            acc[a] = fma(P.A[o], sinkz, fma(P.B[o], cosm1, acc[a]));
        }
    }
    for (int a = 0; a < P.dim; a++)
        x[a] = P.a2[(size_t)p * P.dim + a] + acc[a] / sz;
}

struct FrontDev {
    int budget, cheb, M, bmax;
    double spd, zobs, etaobs, chimax;
    const double* zg;     // [M+1]
    const double* chig;   // [M+1]
    const double* robs;   // [2 (bmax+1)]: (cos(b zobs)-1)/(b sobs), sin(b zobs)/(b sobs)
};

__device__ __forceinline__ double dist_of(double dx, double dy, int cheb) {
    return cheb ? fmax(fabs(dx), fabs(dy)) : sqrt(dx * dx + dy * dy);
}

__device__ __forceinline__ double z_at_chi(const FrontDev& F, double chi) {
    return 2.0 * atan(exp(F.etaobs - chi / F.spd));
}

// Wiggle-budget reach (cuda/frame_bake.cu:240-251): the largest two-time displacement of a
// unit-budget term over frequencies 2..bmax.
__device__ double reach_dev(const FrontDev& F, double z) {
    double s1, c1;
    sincos(z, &s1, &c1);
    double sk = s1, ck = c1, best = 0.0;
    for (int b = 2; b <= F.bmax; b++) {
        const double sn = sk * c1 + ck * s1;
        const double cn = ck * c1 - sk * s1;
        sk = sn;
        ck = cn;
        const double da = (ck - 1.0) / (b * s1) - F.robs[2 * b];
        const double db = sk / (b * s1) - F.robs[2 * b + 1];
        const double d = (b % 2 == 0) ? sqrt(da * da + db * db) : fabs(db);
        best = fmax(best, d);
    }
    return best;
}

// Comoving x, y of every path at every rewind row (cuda/frame_bake.cu:376-386).
__global__ __launch_bounds__(256) void k_frame_grid(DevPaths P, int M, const double* zg,
                                                    double2* rows) {
    const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (tid >= (int64_t)P.n * (M + 1))
        return;
    const int p = (int)(tid / (M + 1)), i = (int)(tid % (M + 1));
    double x[3];
    comov(P, p, zg[i], x);
    rows[tid] = make_double2(x[0], x[1]);
}

// One wave per path; lanes over torus images (nx, ny) in [-K, K]^2.  COUNT: per-path
// number of crossings into counts[p].  Otherwise: the crossings' descriptors (path, image,
// row, side) at offs[p] in (image, row) order, refined into hits by k_frame_bisect.
template <bool COUNT>
__global__ __launch_bounds__(64) void k_frame_hits(DevPaths P, FrontDev F, double ox, double oy,
                                                   int K, int obs_idx, const double2* rows,
                                                   uint32_t* counts, const uint64_t* offs,
                                                   ug_hit* hits) {
    extern __shared__ double2 lrows[];
    const int p = blockIdx.x, lane = threadIdx.x;
    const int M = F.M;
    for (int i = lane; i <= M; i += 64)
        lrows[i] = rows[(size_t)p * (M + 1) + i];
    __syncthreads();
    const int side = 2 * K + 1, nimg = side * side;
    uint32_t cnt = 0;
    uint64_t out = COUNT ? 0 : offs[p];
    // lanes write in image order: each image's output position is the prefix of earlier
    // images' counts, so the kernel loops images in rounds and scans per round
    for (int r0 = 0; r0 < nimg; r0 += 64) {
        const int im = r0 + lane;
        uint32_t mine = 0;
        uint64_t wpos = 0;
        if (!COUNT) {
            // count this lane's crossings first (cheap walk) to place its output
            if (im < nimg) {
                const int nx = im % side - K, ny = im / side - K;
                const double bx = 2.0 * nx - ox, by = 2.0 * ny - oy;
                const double dmin = fmax(0.0, 2.0 * fmax(fabs((double)nx), fabs((double)ny)) - 4.0);
                const int i0 = F.budget ? 1 : max(1, (int)ceil(dmin * M / F.chimax));
                if (dmin < F.chimax && i0 <= M) {
                    double prev = dist_of(lrows[i0 - 1].x + bx, lrows[i0 - 1].y + by, F.cheb) - F.chig[i0 - 1];
                    if (p == obs_idx && nx == 0 && ny == 0 && i0 == 1 && fabs(prev) < 1e-9)
                        prev = -1e-12;
                    else if (prev == 0.0)
                        prev = 1e-12;
                    for (int i = i0; i <= M; i++) {
                        const double g = dist_of(lrows[i].x + bx, lrows[i].y + by, F.cheb) - F.chig[i];
                        mine += ((prev < 0.0) != (g < 0.0)) ? 1u : 0u;
                        prev = g;
                    }
                }
            }
            // exclusive prefix over lanes
            uint32_t incl = mine;
            for (int s = 1; s < 64; s <<= 1) {
                const uint32_t v = __shfl_up(incl, s, 64);
                if (lane >= s)
                    incl += v;
            }
            wpos = out + (incl - mine);
            out += __shfl(incl, 63, 64);
        }
        if (im < nimg) {
            const int nx = im % side - K, ny = im / side - K;
            const double bx = 2.0 * nx - ox, by = 2.0 * ny - oy;
            const double dmin = fmax(0.0, 2.0 * fmax(fabs((double)nx), fabs((double)ny)) - 4.0);
            const int i0 = F.budget ? 1 : max(1, (int)ceil(dmin * M / F.chimax));
            if (dmin < F.chimax && i0 <= M) {
                double prev = dist_of(lrows[i0 - 1].x + bx, lrows[i0 - 1].y + by, F.cheb) - F.chig[i0 - 1];
                if (p == obs_idx && nx == 0 && ny == 0 && i0 == 1 && fabs(prev) < 1e-9)
                    prev = -1e-12;
                else if (prev == 0.0)
                    prev = 1e-12;
                for (int i = i0; i <= M; i++) {
                    const double g = dist_of(lrows[i].x + bx, lrows[i].y + by, F.cheb) - F.chig[i];
                    if ((prev < 0.0) != (g < 0.0)) {
                        if (COUNT) {
                            cnt++;
                        } else {
                            // the crossing's descriptor; k_frame_bisect refines it in place
                            ug_hit hrec;
                            hrec.src = (uint32_t)p;
                            hrec.dx = __int_as_float(nx);
                            hrec.dy = __int_as_float(ny);
                            hrec.chi = __int_as_float(i);
                            hrec.z_emit = __int_as_float(prev < 0.0 ? 1 : 0);
                            hits[wpos++] = hrec;
                        }
                    }
                    prev = g;
                }
            }
        }
    }
    if (COUNT) {
        for (int s = 32; s >= 1; s >>= 1)
            cnt += __shfl_xor(cnt, s, 64);
        if (lane == 0)
            counts[p] = cnt;
    }
}

// One thread per crossing found by k_frame_hits<false>: 24 bisections between rewind rows
// i - 1 and i (in z for the budget front, in chi otherwise) with fresh evaluations, then the
// hit record (cuda/frame_bake.cu:326-369).  Separate from the walk so that every lane works
// on a crossing: in the walk only the few images with crossings of a path would be busy.
__global__ __launch_bounds__(256) void k_frame_bisect(DevPaths P, FrontDev F, double ox, double oy, ug_hit* hits,
                                                      uint64_t n) {
    const uint64_t h = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (h >= n)
        return;
    const ug_hit d = hits[h];
    const int p = (int)d.src, nx = __float_as_int(d.dx), ny = __float_as_int(d.dy);
    const int i = __float_as_int(d.chi);
    const bool inside_prev = __float_as_int(d.z_emit) != 0;
    const double bx = 2.0 * nx - ox, by = 2.0 * ny - oy;
    double chi, zemit, q[3];
    if (F.budget) {
        double lo = F.zg[i - 1], hi = F.zg[i];
        for (int it = 0; it < 24; it++) {
            const double mid = 0.5 * (lo + hi);
            comov(P, p, mid, q);
            const double gm = dist_of(q[0] + bx, q[1] + by, F.cheb) - reach_dev(F, mid);
            if ((gm < 0.0) == inside_prev)
                lo = mid;
            else
                hi = mid;
        }
        zemit = 0.5 * (lo + hi);
        chi = reach_dev(F, zemit);
    } else {
        double lo = F.chig[i - 1], hi = F.chig[i];
        for (int it = 0; it < 24; it++) {
            const double mid = 0.5 * (lo + hi);
            comov(P, p, z_at_chi(F, mid), q);
            const double gm = dist_of(q[0] + bx, q[1] + by, F.cheb) - mid;
            if ((gm < 0.0) == inside_prev)
                lo = mid;
            else
                hi = mid;
        }
        chi = 0.5 * (lo + hi);
        zemit = z_at_chi(F, chi);
    }
    comov(P, p, zemit, q);
    ug_hit hrec;
    hrec.src = (uint32_t)p;
    hrec.dx = (float)(q[0] + bx);
    hrec.dy = (float)(q[1] + by);
    hrec.chi = (float)chi;
    hrec.z_emit = (float)zemit;
    hits[h] = hrec;
}

// Equal-time slices as one dense contraction on FP64 MFMA: the unwrapped positions of
// coefficient row r = p*dim + a at frame f are
//   X[r][f] = a2[r] + sum_k A[r][k] sin(k z_f)/sin z_f + B[r][k] (cos(k z_f) - 1)/sin z_f,
// i.e. Coef (rows x 2kq) times Basis (2kq x frames).  A block computes a 48-row x 64-frame
// tile with v_mfma_f64_16x16x4f64 (each wave 16 frames x 3 row tiles), stages it in LDS, and
// the epilogue wraps, writes float positions [frame][path][dim] and bins (x, y) into the
// frame's raster.  48 rows hold whole paths in 2+1 (24) and 3+1 (16).
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int FR_ROWS = 48, FR_COLS = 64;

__global__ __launch_bounds__(256) void k_frames_mfma(DevPaths P, int kq, const double* basis, int nzp,
                                                     int nz, int fbase, int nz_total, float* pos,
                                                     float* raster, int rpx) {
    __shared__ double tile[FR_ROWS][FR_COLS + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t nrows = (int64_t)P.n * P.dim;
    const int64_t r0 = (int64_t)blockIdx.x * FR_ROWS;
    const int f0 = blockIdx.y * FR_COLS;
    const size_t W = (size_t)P.kmax + 1;
    // operand layouts (v_mfma_f64_16x16x4f64): A lane l = A[l % 16][l / 16],
    // B lane l = B[l / 16][l % 16], D[v] = D[l / 16 + 4 v][l % 16]
    const int ii = lane & 15, kk = lane >> 4;
    int64_t row[3];
    bool rok[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        row[t] = r0 + 16 * t + ii;
        rok[t] = row[t] < nrows;
    }
    const int fcol = f0 + 16 * wv + ii;
    dbl4 acc[3];
#pragma unroll
    for (int t = 0; t < 3; t++)
        acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int ksteps = kq / 2;  // 2 kq contraction rows, 4 per step
    for (int ks = 0; ks < ksteps; ks++) {
        const int r = 4 * ks + kk;
        const bool sinrow = r < kq;  // uniform within a step (kq is a multiple of 4)
        const int k = sinrow ? r + 1 : r - kq + 1;
        const double* src = sinrow ? P.A : P.B;
        const double b = basis[(size_t)r * nzp + fcol];
        double a[3];
#pragma unroll
        for (int t = 0; t < 3; t++)
            a[t] = (rok[t] && k <= P.kmax) ? src[(size_t)row[t] * W + k] : 0.0;
#pragma unroll
        for (int t = 0; t < 3; t++)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b, acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 3; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int rr = 16 * t + kk + 4 * v;
            const int64_t rg = r0 + rr;
            tile[rr][16 * wv + ii] = (rg < nrows ? P.a2[rg] : 0.0) + acc[t][v];
        }
    __syncthreads();
    const int npb = FR_ROWS / P.dim;
    const int64_t p0 = r0 / P.dim;
    for (int q = threadIdx.x; q < npb * FR_COLS; q += 256) {
        const int pp = q / FR_COLS, f = q % FR_COLS;
        const int64_t p = p0 + pp;
        if (p >= P.n || f0 + f >= nz)
            continue;
        const int fg = fbase + f0 + f;
        double x[3];
        for (int a = 0; a < P.dim; a++) {
            const double v = tile[pp * P.dim + a][f];
            x[a] = v - 2.0 * floor((v + 1.0) * 0.5);
        }
        if (pos)
            for (int a = 0; a < P.dim; a++)
                pos[((size_t)fg * P.n + p) * P.dim + a] = (float)x[a];
        if (raster) {
            int cx = (int)floor((x[0] + 1.0) * 0.5 * rpx), cy = (int)floor((x[1] + 1.0) * 0.5 * rpx);
            cx = cx < 0 ? 0 : (cx >= rpx ? rpx - 1 : cx);
            cy = cy < 0 ? 0 : (cy >= rpx ? rpx - 1 : cy);
            atomicAdd(&raster[((size_t)fg * rpx + cy) * rpx + cx], 1.0f);
        }
    }
    (void)nz_total;
}

__global__ void k_scan_u32(const uint32_t* counts, int n, uint64_t* offs, uint64_t* total) {
    // single block exclusive scan (n up to a few million; sequential chunks per thread)
    __shared__ uint64_t part[256];
    const int per = (n + 255) / 256;
    const int b = threadIdx.x * per, e = min(n, b + per);
    uint64_t s = 0;
    for (int i = b; i < e; i++)
        s += counts[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int i = 0; i < 256; i++) {
            const uint64_t v = part[i];
            part[i] = acc;
            acc += v;
        }
        *total = acc;
    }
    __syncthreads();
    uint64_t acc = part[threadIdx.x];
    for (int i = b; i < e; i++) {
        offs[i] = acc;
        acc += counts[i];
    }
}

// host-side reference formulas (cuda/frame_bake.cu:201-251), libm
double eta_host(double z) { return std::log(std::tan(z / 2.0)); }

}  // namespace

struct ug_paths {
    int dim, kmax;
    int64_t n;
    hipStream_t stream;
    double *d_a2 = nullptr, *d_A = nullptr, *d_B = nullptr;
    std::vector<double> a2, A, B;  // host copies (observer evaluation)
    int bmax_terms = 2;            // largest frequency carried by a term (reach band)
};

namespace {

int32_t host_comov(const ug_paths* P, int p, double z, double* x, double* v) {
    // Same dense form as the device, with libm sin/cos per frequency.  Sets built on the
    // device (ug_paths_from_pack) keep no host copy: the one path is fetched.
    const size_t W = (size_t)P->kmax + 1, row = (size_t)P->dim * W;
    std::vector<double> a2v, Av, Bv;
    const double *A = P->A.data(), *B = P->B.data(), *A2 = P->a2.data();
    if (P->A.empty()) {
        a2v.resize(P->dim);
        Av.resize(row);
        Bv.resize(row);
        int32_t rc = hip_check(hipMemcpy(a2v.data(), P->d_a2 + (size_t)p * P->dim, P->dim * 8, hipMemcpyDeviceToHost),
                               "observer a2");
        if (!rc)
            rc = hip_check(hipMemcpy(Av.data(), P->d_A + (size_t)p * row, row * 8, hipMemcpyDeviceToHost), "observer A");
        if (!rc)
            rc = hip_check(hipMemcpy(Bv.data(), P->d_B + (size_t)p * row, row * 8, hipMemcpyDeviceToHost), "observer B");
        if (rc)
            return rc;
        A = Av.data() - (size_t)p * row;
        B = Bv.data() - (size_t)p * row;
        A2 = a2v.data() - (size_t)p * P->dim;
    }
    const double s = std::sin(z), c = std::cos(z);
    for (int a = 0; a < P->dim; a++) {
        double xx = 0.0, dd = 0.0;
        for (int k = 1; k <= P->kmax; k++) {
            const size_t o = ((size_t)p * P->dim + a) * W + k;
            const double skz = std::sin(k * z), ckz = std::cos(k * z);
            xx += A[o] * skz + B[o] * (ckz - 1.0);
            dd += k * (A[o] * ckz - B[o] * skz);
        }
        const double a2 = A2[(size_t)p * P->dim + a];
        x[a] = a2 + xx / s;
        // d/dz [a2 sin z + xx] / sin z, as comov_vel_xy (cuda/frame_bake.cu:215-228)
        const double num = a2 * s + xx, dnum = a2 * c + dd;
        if (v)
            v[a] = (dnum * s - num * c) / (s * s);
    }
    return UG_OK;
}

}  // namespace

extern "C" int32_t ug_paths_create(int32_t dim, int64_t n, int32_t nw, const double* a2, const double* amp,
                                   const double* freq, const double* phase, void* hip_stream, ug_paths** out) {
    if (!out || (n > 0 && (!a2 || !amp || !freq)))
        return fail(UG_EINVAL, "ug_paths_create: NULL argument");
    if (dim < 2 || dim > 3 || n < 0 || nw < 1)
        return fail(UG_EINVAL, "ug_paths_create: invalid dim %d / n %lld / nw %d", dim, (long long)n, nw);
    int kmax = 1;
    for (int64_t i = 0; i < n * dim * nw; i++) {
        const double b = freq[i];
        if (!(b >= 0.0 && b <= 8192.0) || b != std::floor(b))
            return fail(UG_EINVAL, "ug_paths_create: frequency %g is not an integer in [0, 8192]", b);
        kmax = std::max(kmax, (int)b);
    }
    ug_paths* P = new ug_paths();
    P->dim = dim;
    P->n = n;
    P->kmax = kmax;
    P->stream = (hipStream_t)hip_stream;
    const size_t W = (size_t)kmax + 1;
    P->a2.assign(a2, a2 + (size_t)n * dim);
    P->A.assign((size_t)n * dim * W, 0.0);
    P->B.assign((size_t)n * dim * W, 0.0);
    for (int64_t p = 0; p < n; p++)
        for (int a = 0; a < dim; a++)
            for (int j = 0; j < nw; j++) {
                const size_t t = ((size_t)p * dim + a) * nw + j;
                const int b = (int)freq[t];
                const double f = phase ? phase[t] : 0.0;
                if (b == 0 || amp[t] == 0.0)
                    continue;
                P->bmax_terms = std::max(P->bmax_terms, b);
                P->A[((size_t)p * dim + a) * W + b] += amp[t] * std::cos(f);
                P->B[((size_t)p * dim + a) * W + b] += amp[t] * std::sin(f);
            }
    int32_t rc = UG_OK;
    if (hipMalloc((void**)&P->d_a2, std::max<size_t>(8, P->a2.size() * 8)) != hipSuccess ||
        hipMalloc((void**)&P->d_A, std::max<size_t>(8, P->A.size() * 8)) != hipSuccess ||
        hipMalloc((void**)&P->d_B, std::max<size_t>(8, P->B.size() * 8)) != hipSuccess)
        rc = fail(UG_ENOMEM, "ug_paths_create: device allocation failed");
    if (!rc && n > 0) {
        rc = hip_check(hipMemcpyAsync(P->d_a2, P->a2.data(), P->a2.size() * 8, hipMemcpyHostToDevice, P->stream), "a2");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(P->d_A, P->A.data(), P->A.size() * 8, hipMemcpyHostToDevice, P->stream), "A");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(P->d_B, P->B.data(), P->B.size() * 8, hipMemcpyHostToDevice, P->stream), "B");
        if (!rc)
            rc = hip_check(hipStreamSynchronize(P->stream), "paths sync");
    }
    if (rc) {
        ug_paths_destroy(P);
        return rc;
    }
    *out = P;
    return UG_OK;
}

extern "C" void ug_paths_destroy(ug_paths* P) {
    if (!P)
        return;
    if (P->d_a2)
        (void)hipFree(P->d_a2);
    if (P->d_A)
        (void)hipFree(P->d_A);
    if (P->d_B)
        (void)hipFree(P->d_B);
    delete P;
}

extern "C" int32_t ug_paths_from_pack(const ug_pack* hp, void* hip_stream, ug_paths** out) {
    if (!hp || !out)
        return fail(UG_EINVAL, "ug_paths_from_pack: NULL argument");
    const ug::Pack* h = (const ug::Pack*)hp;
    const int64_t n = h->n + h->nsub;
    ug_paths* P = new ug_paths();
    P->dim = h->d;
    P->n = n;
    P->kmax = h->T;
    P->stream = (hipStream_t)hip_stream;
    const size_t W = (size_t)P->kmax + 1;
    int32_t rc = UG_OK;
    int* d_bmax = nullptr;
    if (hipMalloc((void**)&P->d_a2, std::max<size_t>(8, (size_t)n * P->dim * 8)) != hipSuccess ||
        hipMalloc((void**)&P->d_A, std::max<size_t>(8, (size_t)n * P->dim * W * 8)) != hipSuccess ||
        hipMalloc((void**)&P->d_B, std::max<size_t>(8, (size_t)n * P->dim * W * 8)) != hipSuccess ||
        hipMalloc((void**)&d_bmax, 4) != hipSuccess)
        rc = fail(UG_ENOMEM, "ug_paths_from_pack: device allocation failed");
    if (!rc && n > 0) {
        rc = ug::launch_paths_from_pack(const_cast<ug::Pack*>(h), P->stream, P->d_a2, P->d_A, P->d_B, P->kmax, d_bmax);
        int bmax = 2;
        if (!rc)
            rc = hip_check(hipMemcpyAsync(&bmax, d_bmax, 4, hipMemcpyDeviceToHost, P->stream), "bmax copy");
        if (!rc)
            rc = hip_check(hipStreamSynchronize(P->stream), "paths sync");
        P->bmax_terms = std::max(2, bmax);
    }
    if (d_bmax)
        (void)hipFree(d_bmax);
    if (rc) {
        ug_paths_destroy(P);
        return rc;
    }
    *out = P;
    return UG_OK;
}

static DevPaths dev_of(const ug_paths* P) {
    DevPaths D;
    D.n = (int)P->n;
    D.dim = P->dim;
    D.kmax = P->kmax;
    D.a2 = P->d_a2;
    D.A = P->d_A;
    D.B = P->d_B;
    return D;
}

extern "C" int32_t ug_frame_bake_lightcone(ug_paths* P, const ug_front* fr, double z_obs, const ug_observer* obs,
                                           ug_hit* dev_hits, uint64_t cap, ug_frame_info* info) {
    if (!P || !fr || !obs || !info)
        return fail(UG_EINVAL, "ug_frame_bake_lightcone: NULL argument");
    if (P->dim != 2)
        return fail(UG_EINVAL, "error: the light-cone baker is 2+1 only (cuda/frame_bake.cu:124-127)");
    if (fr->kind < 0 || fr->kind > 2)
        return fail(UG_EINVAL, "error: front must be speed|fit|budget");
    if (obs->type == 0 && (obs->path < 0 || obs->path >= P->n))
        return fail(UG_EINVAL, "error: observer path %d out of range", obs->path);
    const double z_min = fr->z_min;
    if (!(z_min < z_obs))
        return fail(UG_EINVAL, "error: rewind floor %.4f >= z_obs %.4f", z_min, z_obs);
    const bool budget = fr->kind == 2;
    const int bmax = P->bmax_terms;
    FrontDev F;
    F.budget = budget ? 1 : 0;
    F.cheb = (budget || fr->cheb) ? 1 : 0;
    F.zobs = z_obs;
    F.etaobs = eta_host(z_obs);
    F.spd = fr->kind == 0 ? std::max(1e-3, fr->speed) : -1.0 / eta_host(z_min);
    F.bmax = bmax;
    F.M = budget ? std::min(4000, std::max(420, 2 * bmax)) : 420;
    const int M = F.M;
    std::vector<double> zg(M + 1), chig(M + 1), robs(2 * (size_t)(bmax + 1), 0.0);
    const double sobs = std::sin(z_obs);
    for (int b = 2; b <= bmax; b++) {
        robs[2 * b] = (std::cos(b * z_obs) - 1.0) / (b * sobs);
        robs[2 * b + 1] = std::sin(b * z_obs) / (b * sobs);
    }
    auto reach_host = [&](double z) {
        const double s = std::sin(z);
        double best = 0.0;
        for (int b = 2; b <= bmax; b++) {
            const double da = (std::cos(b * z) - 1.0) / (b * s) - robs[2 * b];
            const double db = std::sin(b * z) / (b * s) - robs[2 * b + 1];
            const double d = (b % 2 == 0) ? std::sqrt(da * da + db * db) : std::fabs(db);
            best = std::max(best, d);
        }
        return best;
    };
    if (budget) {
        F.chimax = 0.0;
        for (int i = 0; i <= M; i++) {
            zg[i] = z_obs - (z_obs - z_min) * i / M;
            chig[i] = reach_host(zg[i]);
            F.chimax = std::max(F.chimax, chig[i]);
        }
    } else {
        F.chimax = F.spd * (F.etaobs - eta_host(z_min));
        for (int i = 0; i <= M; i++) {
            chig[i] = F.chimax * i / M;
            zg[i] = 2.0 * std::atan(std::exp(F.etaobs - chig[i] / F.spd));
        }
    }
    const int K = fr->wraps >= 0 ? fr->wraps : (int)std::ceil((F.chimax + 4.0) / 2.0);
    double ox, oy, bx = 0.0, by = 0.0;
    if (obs->type == 0) {
        double x[3], v[3];
        {
            const int32_t hc = host_comov(P, obs->path, z_obs, x, v);
            if (hc)
                return hc;
        }
        ox = x[0];
        oy = x[1];
        bx = std::sin(z_obs) * v[0] / F.spd;
        by = std::sin(z_obs) * v[1] / F.spd;
    } else {
        ox = obs->px;
        oy = obs->py;
    }
    info->z_obs = z_obs;
    info->chi_max = F.chimax;
    info->K = K;
    info->ox = ox;
    info->oy = oy;
    info->betax = bx;
    info->betay = by;
    info->n_hits = 0;
    if (P->n == 0)
        return UG_OK;
    hipStream_t st = P->stream;
    double *dzg = nullptr, *dchig = nullptr, *drobs = nullptr;
    double2* drows = nullptr;
    uint32_t* dcnt = nullptr;
    uint64_t *doff = nullptr, *dtot = nullptr;
    int32_t rc = UG_OK;
    if (hipMalloc((void**)&dzg, (M + 1) * 8) != hipSuccess || hipMalloc((void**)&dchig, (M + 1) * 8) != hipSuccess ||
        hipMalloc((void**)&drobs, robs.size() * 8) != hipSuccess ||
        hipMalloc((void**)&drows, (size_t)P->n * (M + 1) * 16) != hipSuccess ||
        hipMalloc((void**)&dcnt, (size_t)P->n * 4) != hipSuccess ||
        hipMalloc((void**)&doff, (size_t)P->n * 8) != hipSuccess || hipMalloc((void**)&dtot, 8) != hipSuccess)
        rc = fail(UG_ENOMEM, "ug_frame_bake_lightcone: device allocation failed");
    if (!rc)
        rc = hip_check(hipMemcpyAsync(dzg, zg.data(), (M + 1) * 8, hipMemcpyHostToDevice, st), "zg");
    if (!rc)
        rc = hip_check(hipMemcpyAsync(dchig, chig.data(), (M + 1) * 8, hipMemcpyHostToDevice, st), "chig");
    if (!rc)
        rc = hip_check(hipMemcpyAsync(drobs, robs.data(), robs.size() * 8, hipMemcpyHostToDevice, st), "robs");
    F.zg = dzg;
    F.chig = dchig;
    F.robs = drobs;
    const DevPaths D = dev_of(P);
    const int obs_idx = obs->type == 0 ? obs->path : -1;
    uint64_t total = 0;
    if (!rc) {
        const int64_t nr = P->n * (int64_t)(M + 1);
        hipLaunchKernelGGL(k_frame_grid, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, st, D, M, dzg, drows);
        const size_t lds = (size_t)(M + 1) * 16;
        hipLaunchKernelGGL(k_frame_hits<true>, dim3((unsigned)P->n), dim3(64), lds, st, D, F, ox, oy, K, obs_idx,
                           drows, dcnt, doff, dev_hits);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(256), 0, st, dcnt, (int)P->n, doff, dtot);
        rc = hip_check(hipGetLastError(), "frame count");
        if (!rc)
            rc = hip_check(hipMemcpyAsync(&total, dtot, 8, hipMemcpyDeviceToHost, st), "total");
        if (!rc)
            rc = hip_check(hipStreamSynchronize(st), "frame sync");
        info->n_hits = total;
        if (!rc && total > cap)
            rc = fail(UG_EOVERFLOW, "ug_frame_bake_lightcone: %llu hits exceed the capacity %llu",
                      (unsigned long long)total, (unsigned long long)cap);
        if (!rc && total > 0) {
            hipLaunchKernelGGL(k_frame_hits<false>, dim3((unsigned)P->n), dim3(64), lds, st, D, F, ox, oy, K,
                               obs_idx, drows, dcnt, doff, dev_hits);
            hipLaunchKernelGGL(k_frame_bisect, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, D, F, ox, oy,
                               dev_hits, total);
            rc = hip_check(hipGetLastError(), "frame hits");
            if (!rc)
                rc = hip_check(hipStreamSynchronize(st), "frame sync");
        }
    }
    (void)hipFree(dzg);
    (void)hipFree(dchig);
    (void)hipFree(drobs);
    (void)hipFree(drows);
    (void)hipFree(dcnt);
    (void)hipFree(doff);
    (void)hipFree(dtot);
    return rc;
}

namespace {

__global__ void k_fill_f32(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = 0.0f;
}

// Slices of frames z[0..nz) in batches of at most FB frames.  The basis of a batch (rows
// r < kq: sin(k z)/sin z with k = r + 1; rows kq + r: (cos(k z) - 1)/sin z; FP64 libm on the
// host) is built in hb and uploaded to dbasis, both holding 2 kq nzp doubles.  The raster is
// cleared by a fill kernel.
int32_t slices_run(ug_paths* P, const double* z, int nz, float* dev_pos, float* dev_raster, int raster_px,
                   double* dbasis, double* hb, int FB) {
    hipStream_t st = P->stream;
    const int kq = (P->kmax + 3) / 4 * 4;
    const int nzp = (std::min(nz, FB) + FR_COLS - 1) / FR_COLS * FR_COLS;
    const size_t nb_words = (size_t)2 * kq * nzp;
    int32_t rc = UG_OK;
    if (dev_raster) {
        const size_t nr = (size_t)nz * raster_px * raster_px;
        hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<size_t>((nr + 255) / 256, 8192)), dim3(256), 0, st,
                           dev_raster, nr);
        rc = hip_check(hipGetLastError(), "raster clear");
    }
    const int64_t nrows = P->n * P->dim;
    for (int fb = 0; fb < nz && !rc; fb += FB) {
        const int nb = std::min(FB, nz - fb);
        // the previous batch's upload and kernel must be done with both buffers first
        rc = hip_check(hipStreamSynchronize(st), "slices sync");
        if (rc)
            break;
        std::fill(hb, hb + nb_words, 0.0);
        for (int f = 0; f < nb; f++) {
            const double zf = z[fb + f], sz = std::sin(zf);
            for (int k = 1; k <= kq; k++) {
                hb[(size_t)(k - 1) * nzp + f] = std::sin(k * zf) / sz;
                hb[(size_t)(kq + k - 1) * nzp + f] = (std::cos(k * zf) - 1.0) / sz;
            }
        }
        rc = hip_check(hipMemcpyAsync(dbasis, hb, nb_words * 8, hipMemcpyHostToDevice, st), "basis upload");
        if (!rc) {
            const dim3 grid((unsigned)((nrows + FR_ROWS - 1) / FR_ROWS), (unsigned)((nb + FR_COLS - 1) / FR_COLS));
            hipLaunchKernelGGL(k_frames_mfma, grid, dim3(256), 0, st, dev_of(P), kq, dbasis, nzp, nb, fb, nz, dev_pos,
                               dev_raster, raster_px);
            rc = hip_check(hipGetLastError(), "k_frames_mfma");
        }
    }
    if (!rc)
        rc = hip_check(hipStreamSynchronize(st), "slices sync");
    return rc;
}

int32_t slices_check(const ug_paths* P, const double* z, int32_t nz, const float* dev_raster, int32_t raster_px,
                     const char* who) {
    if (!P || (!z && nz > 0) || nz < 0 || (dev_raster && raster_px <= 0))
        return fail(UG_EINVAL, "%s: invalid argument", who);
    for (int j = 0; j < nz; j++)
        if (!(z[j] > 0.0 && z[j] < 3.141592653589793))
            return fail(UG_EINVAL, "%s: z[%d] = %g outside (0, pi)", who, j, z[j]);
    return UG_OK;
}

// The frame basis of ug_frame_bake_slices_noalloc: a file-scope device array (each GPU holds
// its own copy until the process ends) and a host staging array of the same size.  Host code
// never uses the device array's name as an address: its per-device address comes from
// hipGetSymbolAddress on the first call on that device.  One lock serialises the calls of a
// process, since they share both arrays.
constexpr size_t NOALLOC_BASIS_DOUBLES = (size_t)2 * 512 * 512;  // 4 MiB
double g_noalloc_host[NOALLOC_BASIS_DOUBLES];
constexpr int NOALLOC_MAX_DEVICES = 64;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_noalloc_mu;

}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_BASIS_DOUBLES];

extern "C" int32_t ug_frame_bake_slices(ug_paths* P, const double* z, int32_t nz, float* dev_pos, float* dev_raster,
                                        int32_t raster_px) {
    int32_t rc = slices_check(P, z, nz, dev_raster, raster_px, "ug_frame_bake_slices");
    if (rc || nz == 0 || P->n == 0)
        return rc;
    const int kq = (P->kmax + 3) / 4 * 4;
    const int FB = 512;
    const int nzp = (std::min(nz, FB) + FR_COLS - 1) / FR_COLS * FR_COLS;
    double* dbasis = nullptr;
    if (hipMalloc((void**)&dbasis, (size_t)2 * kq * nzp * 8) != hipSuccess)
        return fail(UG_ENOMEM, "ug_frame_bake_slices: allocation failed");
    std::vector<double> hb((size_t)2 * kq * nzp);
    rc = slices_run(P, z, nz, dev_pos, dev_raster, raster_px, dbasis, hb.data(), FB);
    (void)hipFree(dbasis);
    return rc;
}

extern "C" int32_t ug_frame_bake_slices_noalloc(ug_paths* P, const double* z, int32_t nz, float* dev_pos,
                                                float* dev_raster, int32_t raster_px) {
    int32_t rc = slices_check(P, z, nz, dev_raster, raster_px, "ug_frame_bake_slices_noalloc");
    if (rc || nz == 0 || P->n == 0)
        return rc;
    const int kq = (P->kmax + 3) / 4 * 4;
    // frames per batch: as many 64-frame columns as the scratch holds, at most 512
    const int FB = (int)std::min<size_t>(512, NOALLOC_BASIS_DOUBLES / ((size_t)2 * kq) / FR_COLS * FR_COLS);
    if (FB < FR_COLS)
        return fail(UG_EOVERFLOW,
                    "ug_frame_bake_slices_noalloc: largest frequency %d needs %zu basis doubles per 64 frames, "
                    "above the %zu of the scratch",
                    P->kmax, (size_t)2 * kq * FR_COLS, NOALLOC_BASIS_DOUBLES);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(UG_EHIP, "ug_frame_bake_slices_noalloc: no usable current device");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch address");
        if (rc)
            return rc;
        g_scratch_ptr[dev] = (double*)ptr;
    }
    return slices_run(P, z, nz, dev_pos, dev_raster, raster_px, g_scratch_ptr[dev], g_noalloc_host, FB);
}
