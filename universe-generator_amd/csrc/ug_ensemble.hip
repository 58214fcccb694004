// ug_ensemble.hip -- many independent universes on one GPU (the batched-ensemble workload:
// uneven T and dimension, distinct seeds).
//
// A universe's pack is a chain of dependent launches with a host round trip per
// sub-batch, so one universe leaves most of the GPU idle.  The ensemble runner keeps
// `threads` universes in flight: each worker thread owns one HIP stream (default priority;
// a kept handle gets one of its own) and takes the next universe from a shared counter, so
// the launches of different universes overlap on the device while each worker waits on its
// own stream.  Results do not depend on the thread count: every universe is the
// deterministic pack of its config.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ug_internal.h"

// FNV-1a (64-bit) over the little-endian bytes of an accepted index list.
uint64_t ug::list_digest(const std::vector<int64_t>& idx) {
    uint64_t h = 1469598103934665603ull;
    for (const int64_t v : idx)
        for (int b = 0; b < 8; b++) {
            h ^= (uint64_t)((v >> (8 * b)) & 0xff);
            h *= 1099511628211ull;
        }
    return h;
}

extern "C" int32_t ug_ensemble_run(const ug_pack_config* cfgs, int32_t n, int32_t threads,
                                   ug_pack** keep, ug_pack_status* status, uint64_t* digests) {
    if (!cfgs || !status || n < 0)
        return ug::fail(UG_EINVAL, "ug_ensemble_run: NULL argument or negative count");
    if (n == 0)
        return UG_OK;
    if (threads < 1)
        threads = 1;
    threads = std::min(threads, n);
    for (int i = 0; keep && i < n; i++)
        keep[i] = nullptr;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
        return ug::fail(UG_EHIP, "ug_ensemble_run: no current device");
    std::atomic<int> next{0};
    std::atomic<bool> stop{false};
    std::mutex err_mu;
    int32_t err_rc = UG_OK;
    std::string err_msg;
    auto record = [&](int32_t rc, const std::string& msg) {
        std::lock_guard<std::mutex> g(err_mu);
        if (err_rc == UG_OK) {
            err_rc = rc;
            err_msg = msg;
        }
        stop = true;
    };
    auto worker = [&]() {
        if (hipSetDevice(device) != hipSuccess) {
            record(UG_EHIP, "ug_ensemble_run: hipSetDevice failed in a worker");
            return;
        }
        // one stream per worker; a kept handle outlives the call, so it gets a stream of its
        // own that it destroys with itself
        hipStream_t ws = nullptr;
        if (!keep && hipStreamCreate(&ws) != hipSuccess) {
            record(UG_EHIP, "ug_ensemble_run: hipStreamCreate failed");
            return;
        }
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n || stop)
                break;
            hipStream_t s = ws;
            if (keep && hipStreamCreate(&s) != hipSuccess) {
                record(UG_EHIP, "ug_ensemble_run: hipStreamCreate failed");
                break;
            }
            ug_pack* h = nullptr;
            int32_t rc = ug_pack_create(&cfgs[i], (void*)s, &h);
            if (h && keep)
                ((ug::Pack*)h)->own_stream = true;
            else if (keep)
                (void)hipStreamDestroy(s);
            if (rc == UG_OK)
                rc = ug_pack_run(h, 0, &status[i]);
            if (rc != UG_OK) {
                record(rc, std::string("universe ") + std::to_string(i) + ": " + ug_last_error());
                if (h)
                    ug_pack_destroy(h);
                break;
            }
            if (digests)
                digests[i] = ug::list_digest(((const ug::Pack*)h)->acc_host);
            if (keep)
                keep[i] = h;
            else
                ug_pack_destroy(h);
        }
        if (ws)
            (void)hipStreamDestroy(ws);
    };
    std::vector<std::thread> pool;
    pool.reserve((size_t)threads);
    for (int t = 0; t < threads; t++)
        pool.emplace_back(worker);
    for (auto& th : pool)
        th.join();
    if (err_rc != UG_OK) {
        // the universes finished before the failure are not handed out either
        for (int i = 0; keep && i < n; i++)
            if (keep[i]) {
                ug_pack_destroy(keep[i]);
                keep[i] = nullptr;
            }
        return ug::fail(err_rc, "%s", err_msg.c_str());
    }
    return UG_OK;
}
