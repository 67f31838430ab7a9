// Cache of device blocks for library-allocated outputs (see outmem.hpp).
//
// No counterpart in the reference, which calls cudaMalloc for C on every
// product (src/opsparse/core_algo/spgemm/spgemm.cu:96, 109, 114) and cudaFree
// in CSR::release (src/opsparse/core_ds/csr.cu:40-46).
//
// Policy: sizes are rounded up (512 B below 1 MiB, 2 MiB above) and a
// request is served by the smallest cached block of at least the rounded
// size and at most 2x it (small) or 1.25x it (64 MiB and above).  Cached
// bytes per device are capped (default: half of the device's memory,
// ops_mem_cache_set_limit changes it, 0 turns caching off).  A released
// block within the cap is kept and the least recently released cached
// blocks are freed to make room for it (so a caller alternating between two
// products that do not fit together keeps the one it is using); a block
// larger than the cap is freed at once.
#include "outmem.hpp"

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>

#include "opsparse_amd.h"

namespace opsamd {

namespace {

constexpr int kMaxDev = 64;

struct Rec {
    size_t bytes;
    int dev;
};

struct State {
    std::mutex mu;
    std::unordered_map<void *, Rec> live;        // handed out by out_malloc
    std::multimap<size_t, void *> pool[kMaxDev];  // cached blocks by size
    // cached blocks by release order (oldest first) and each block's key there
    std::map<unsigned long long, std::pair<size_t, void *>> age[kMaxDev];
    std::unordered_map<void *, unsigned long long> age_of;
    unsigned long long tick = 0;
    size_t cached[kMaxDev] = {};
    size_t live_bytes = 0;
    long long limit = -1;  // < 0: half of the device's memory
    size_t dev_total[kMaxDev] = {};
    long long hits = 0, misses = 0;
};

State &state()
{
    static State *s = new State();  // never destroyed: blocks may be released during process exit
    return *s;
}

size_t round_size(size_t b)
{
    b = std::max<size_t>(b, 1);
    constexpr size_t kSmall = 512, kLarge = size_t(2) << 20;
    return b < (size_t(1) << 20) ? (b + kSmall - 1) / kSmall * kSmall : (b + kLarge - 1) / kLarge * kLarge;
}

size_t max_fit(size_t need) { return need < (size_t(64) << 20) ? 2 * need : need + need / 4; }

// Caller holds the lock; the device is current.
size_t limit_for(State &s, int dev)
{
    if (s.limit >= 0) return static_cast<size_t>(s.limit);
    if (!s.dev_total[dev]) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        s.dev_total[dev] = tot;
    }
    return s.dev_total[dev] / 2;
}

// Caller holds the lock.  Adds a released block to the cache of `dev`.
void pool_put(State &s, int dev, size_t bytes, void *p)
{
    s.pool[dev].emplace(bytes, p);
    const unsigned long long t = ++s.tick;
    s.age[dev].emplace(t, std::make_pair(bytes, p));
    s.age_of[p] = t;
    s.cached[dev] += bytes;
}

// Caller holds the lock.  Removes cached block `it` of `dev` (not freed).
void pool_take(State &s, int dev, std::multimap<size_t, void *>::iterator it)
{
    auto a = s.age_of.find(it->second);
    if (a != s.age_of.end()) {
        s.age[dev].erase(a->second);
        s.age_of.erase(a);
    }
    s.cached[dev] -= it->first;
    s.pool[dev].erase(it);
}

// Caller holds the lock; the device is current.  Frees the least recently
// released cached block of `dev`.
void evict_oldest(State &s, int dev)
{
    auto o = s.age[dev].begin();
    const size_t bytes = o->second.first;
    void *p = o->second.second;
    auto range = s.pool[dev].equal_range(bytes);
    for (auto it = range.first; it != range.second; ++it)
        if (it->second == p) {
            pool_take(s, dev, it);
            break;
        }
    (void)hipFree(p);
}

// Frees every cached block of `dev` (the device must be current).
void drain(State &s, int dev)
{
    for (auto &kv : s.pool[dev]) {
        s.age_of.erase(kv.second);
        (void)hipFree(kv.second);
    }
    s.pool[dev].clear();
    s.age[dev].clear();
    s.cached[dev] = 0;
}

}  // namespace

hipError_t out_malloc(void **p, size_t bytes)
{
    State &s = state();
    *p = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDev) return hipMalloc(p, bytes);
    const size_t need = round_size(bytes);
    {
        std::lock_guard<std::mutex> g(s.mu);
        auto &pl = s.pool[dev];
        auto it = pl.lower_bound(need);
        if (it != pl.end() && it->first <= max_fit(need)) {
            *p = it->second;
            const size_t sz = it->first;
            pool_take(s, dev, it);
            s.live[*p] = Rec{sz, dev};
            s.live_bytes += sz;
            s.hits++;
            return hipSuccess;
        }
    }
    e = hipMalloc(p, need);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> g(s.mu);
            drain(s, dev);
        }
        e = hipMalloc(p, need);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> g(s.mu);
    s.live[*p] = Rec{need, dev};
    s.live_bytes += need;
    s.misses++;
    return hipSuccess;
}

hipError_t out_free(void *p)
{
    if (!p) return hipSuccess;
    State &s = state();
    Rec r{};
    bool owned = false;
    {
        std::lock_guard<std::mutex> g(s.mu);
        auto it = s.live.find(p);
        if (it != s.live.end()) {
            r = it->second;
            s.live.erase(it);
            s.live_bytes -= r.bytes;
            owned = true;
        }
    }
    if (!owned) return hipFree(p);
    // A block released behind the library's back (plain hipFree) and handed
    // out again by the runtime would no longer span the recorded size.
    void *base = nullptr;
    size_t sz = 0;
    if (hipMemGetAddressRange(&base, &sz, p) != hipSuccess || base != p || sz < r.bytes) {
        (void)hipGetLastError();
        return hipFree(p);
    }
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != r.dev && (e = hipSetDevice(r.dev)) != hipSuccess) return e;
    // Like hipFree: work still reading or writing the block finishes first.
    e = hipDeviceSynchronize();
    bool keep = false;
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> g(s.mu);
        const size_t cap = limit_for(s, r.dev);
        if (r.bytes <= cap) {
            while (s.cached[r.dev] + r.bytes > cap && !s.age[r.dev].empty()) evict_oldest(s, r.dev);
            pool_put(s, r.dev, r.bytes, p);
            keep = true;
        }
    }
    if (!keep) {
        const hipError_t f = hipFree(p);
        if (e == hipSuccess) e = f;
    }
    if (cur != r.dev) (void)hipSetDevice(cur);
    return e;
}

}  // namespace opsamd

using namespace opsamd;

extern "C" {

int ops_free(void *p)
{
    const hipError_t e = out_free(p);
    if (e == hipSuccess) return OPS_OK;
    (void)hipGetLastError();
    return OPS_HIP_ERROR;
}

int ops_mem_cache_set_limit(int64_t bytes)
{
    State &s = state();
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) {
        (void)hipGetLastError();
        return OPS_HIP_ERROR;
    }
    std::lock_guard<std::mutex> g(s.mu);
    s.limit = bytes < 0 ? -1 : bytes;
    for (int d = 0; d < kMaxDev; d++) {
        if (s.pool[d].empty()) continue;
        if (d != cur) (void)hipSetDevice(d);
        const size_t cap = limit_for(s, d);
        // Free the least recently released blocks until the device is under its cap.
        while (s.cached[d] > cap && !s.age[d].empty()) evict_oldest(s, d);
        if (d != cur) (void)hipSetDevice(cur);
    }
    return OPS_OK;
}

int ops_mem_cache_release(void)
{
    State &s = state();
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) {
        (void)hipGetLastError();
        return OPS_HIP_ERROR;
    }
    std::lock_guard<std::mutex> g(s.mu);
    for (int d = 0; d < kMaxDev; d++) {
        if (s.pool[d].empty()) continue;
        if (d != cur) (void)hipSetDevice(d);
        drain(s, d);
        if (d != cur) (void)hipSetDevice(cur);
    }
    return OPS_OK;
}

int ops_mem_cache_stats(ops_mem_stats *out)
{
    if (!out) return OPS_INVALID;
    State &s = state();
    std::lock_guard<std::mutex> g(s.mu);
    *out = ops_mem_stats{};
    for (int d = 0; d < kMaxDev; d++) {
        out->cached_bytes += static_cast<int64_t>(s.cached[d]);
        out->cached_blocks += static_cast<int64_t>(s.pool[d].size());
    }
    out->live_bytes = static_cast<int64_t>(s.live_bytes);
    out->live_blocks = static_cast<int64_t>(s.live.size());
    out->hits = s.hits;
    out->misses = s.misses;
    out->limit = s.limit;
    return OPS_OK;
}

}  // extern "C"
