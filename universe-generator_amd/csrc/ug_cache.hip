// ug_cache.hip -- recycled memory and per-model constant tables of packing handles.
//
// A pack of a small universe takes a few milliseconds, so the ~20 device allocations, the
// pinned staging block and the basis tables of a new handle (host-side sincospi of every
// (t, k) pair, the split-BF16 fragments and margins, their uploads) would otherwise cost a
// sixth of it.  Two process-wide caches remove that cost for the second and later handles:
//
//  - buffers: a destroyed handle's device buffers and pinned block go to a free list per
//    device keyed by exact byte size (at most UG_CACHE_BYTES cached per device; beyond it,
//    and on ug_cache_trim, they are freed); a new handle of the same model finds every size
//    it needs there.  Handles clear what must start empty themselves, as before.
//  - model tables: the read-only tables of one (device, dim, T, terms, form, margin) are
//    built once and shared by every handle of that model.
//
// Both are guarded by one mutex (the ensemble creates handles from several threads) and are
// never destroyed at process exit (their storage is deliberately leaked, so no HIP call runs
// after the runtime has shut down).
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ug_arith.h"
#include "ug_internal.h"

#define UG_CACHE_BYTES ((size_t)16 << 30)

namespace ug {
namespace {

constexpr int kHost = -1;  // the pinned host blocks' free list

struct FreeList {
    std::multimap<size_t, void*> blocks;
    size_t bytes = 0;
};

struct Caches {
    std::mutex mu;
    std::map<int, FreeList> free;                              // device (or kHost) -> blocks
    std::unordered_map<void*, std::pair<int, size_t>> live;    // handed out: owner, size
    std::map<std::tuple<int, int, int, int, int, double>, std::shared_ptr<ModelTables>> tables;
};

Caches& caches() {
    static Caches* c = new Caches();  // never destroyed (see above)
    return *c;
}

void* take(int owner, size_t bytes) {
    Caches& c = caches();
    std::lock_guard<std::mutex> g(c.mu);
    FreeList& f = c.free[owner];
    auto it = f.blocks.find(bytes);
    if (it == f.blocks.end())
        return nullptr;
    void* p = it->second;
    f.blocks.erase(it);
    f.bytes -= bytes;
    c.live[p] = {owner, bytes};
    return p;
}

void trim_owner(int owner) {
    std::vector<void*> drop;
    {
        Caches& c = caches();
        std::lock_guard<std::mutex> g(c.mu);
        FreeList& f = c.free[owner];
        for (auto& b : f.blocks)
            drop.push_back(b.second);
        f.blocks.clear();
        f.bytes = 0;
    }
    for (void* p : drop)
        (void)(owner == kHost ? hipHostFree(p) : hipFree(p));
}

void* fresh(int owner, size_t bytes) {
    void* p = nullptr;
    hipError_t e = owner == kHost ? hipHostMalloc(&p, bytes) : hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim_owner(owner);  // cached blocks of other sizes may be what is missing
        e = owner == kHost ? hipHostMalloc(&p, bytes) : hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    Caches& c = caches();
    std::lock_guard<std::mutex> g(c.mu);
    c.live[p] = {owner, bytes};
    return p;
}

void give_back(void* p) {
    if (!p)
        return;
    int owner = 0;
    size_t bytes = 0;
    {
        Caches& c = caches();
        std::lock_guard<std::mutex> g(c.mu);
        auto it = c.live.find(p);
        if (it == c.live.end())
            return;  // not ours: never happens for handle buffers
        owner = it->second.first;
        bytes = it->second.second;
        c.live.erase(it);
        FreeList& f = c.free[owner];
        if (f.bytes + bytes <= UG_CACHE_BYTES) {
            f.blocks.emplace(bytes, p);
            f.bytes += bytes;
            return;
        }
    }
    (void)(owner == kHost ? hipHostFree(p) : hipFree(p));
}

}  // namespace

void* cache_alloc(size_t bytes) {
    if (bytes == 0)
        bytes = 1;
    int dev = 0;
    (void)hipGetDevice(&dev);
    void* p = take(dev, bytes);
    return p ? p : fresh(dev, bytes);
}

void cache_free(void* p) {
    give_back(p);
}

void* cache_host_alloc(size_t bytes) {
    void* p = take(kHost, bytes);
    return p ? p : fresh(kHost, bytes);
}

void cache_host_free(void* p) {
    give_back(p);
}

// The canonical basis tables (ug_arith.h arithmetic): S[t][k] = sin(k z_t)/sin z_t and
// C[t][k] = (cos(k z_t) - 1)/sin z_t = -2 sin^2(k z_t / 2)/sin z_t at z_t = (t+1) pi/(T+1),
// their transposes, and the endpoint-first scan order 0, T-1, 1, T-2, ...
int32_t model_tables(int d, int T, int nw, bool split_full, double margin_req,
                     std::shared_ptr<const ModelTables>* out) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple(dev, d, T, nw, split_full ? 1 : 0, margin_req);
    Caches& c = caches();
    std::lock_guard<std::mutex> g(c.mu);  // held while building: concurrent creators wait
    auto it = c.tables.find(key);
    if (it != c.tables.end()) {
        *out = it->second;
        return UG_OK;
    }
    auto t = std::make_shared<ModelTables>();
    t->device = dev;
    t->d = d;
    t->T = T;
    t->nw = nw;
    t->split = split_full;
    const int W = T + 1;
    std::vector<double> S((size_t)T * W), C((size_t)T * W);
    const uint64_t Dn = (uint64_t)T + 1;
    for (int ti = 0; ti < T; ti++) {
        const double s1 = sincospi_rat((uint64_t)(ti + 1), Dn).s;
        for (int k = 0; k <= T; k++) {
            const uint64_t nk = (uint64_t)k * (uint64_t)(ti + 1);
            const double sk = sincospi_rat(nk, Dn).s;
            const double hk = sincospi_rat(nk, 2 * Dn).s;
            S[(size_t)ti * W + k] = sk / s1;
            C[(size_t)ti * W + k] = (-2.0 * (hk * hk)) / s1;
        }
    }
    std::vector<double> ST((size_t)W * T), CT((size_t)W * T);
    for (int ti = 0; ti < T; ti++)
        for (int k = 0; k <= T; k++) {
            ST[(size_t)k * T + ti] = S[(size_t)ti * W + k];
            CT[(size_t)k * T + ti] = C[(size_t)ti * W + k];
        }
    std::vector<int> order(T);
    for (int i = 0, o = 0; o < T; i++) {
        order[o++] = i;
        if (o < T)
            order[o++] = T - 1 - i;
    }
    if (hipMalloc((void**)&t->d_S, S.size() * 8) != hipSuccess ||
        hipMalloc((void**)&t->d_C, C.size() * 8) != hipSuccess ||
        hipMalloc((void**)&t->d_ST, ST.size() * 8) != hipSuccess ||
        hipMalloc((void**)&t->d_CT, CT.size() * 8) != hipSuccess ||
        hipMalloc((void**)&t->d_order, (size_t)T * 4) != hipSuccess)
        return fail(UG_ENOMEM, "model table allocation failed (T=%d)", T);  // kept for the process
    int32_t rc = hip_check(hipMemcpy(t->d_S, S.data(), S.size() * 8, hipMemcpyHostToDevice), "upload S");
    if (!rc)
        rc = hip_check(hipMemcpy(t->d_C, C.data(), C.size() * 8, hipMemcpyHostToDevice), "upload C");
    if (!rc)
        rc = hip_check(hipMemcpy(t->d_ST, ST.data(), ST.size() * 8, hipMemcpyHostToDevice), "upload ST");
    if (!rc)
        rc = hip_check(hipMemcpy(t->d_CT, CT.data(), CT.size() * 8, hipMemcpyHostToDevice), "upload CT");
    if (!rc)
        rc = hip_check(hipMemcpy(t->d_order, order.data(), (size_t)T * 4, hipMemcpyHostToDevice), "upload order");
    if (!rc && split_full)
        rc = fast_build(*t, S, C, order, margin_req);
    if (rc)
        return rc;
    c.tables.emplace(key, t);
    *out = t;
    return UG_OK;
}

}  // namespace ug

extern "C" int32_t ug_cache_trim(void) {
    std::vector<int> owners;
    {
        ug::Caches& c = ug::caches();
        std::lock_guard<std::mutex> g(c.mu);
        for (auto& f : c.free)
            owners.push_back(f.first);
    }
    for (int o : owners)
        ug::trim_owner(o);
    return UG_OK;
}
