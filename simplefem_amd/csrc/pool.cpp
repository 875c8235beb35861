// pool.cpp -- cached device and pinned host blocks for the library's own
// buffers.
//
// hipFree and hipHostFree wait until the whole device is idle before they
// return.  A process that runs several solves side by side (the solve
// server's dedicated systems, each on its own stream, next to the batch
// engine) then stalls every time one of them frees its buffers: freeing a
// system took longer than solving it.  So a freed block goes back to a cache
// instead, per device and kind, and the next allocation of about its size
// takes it.  Whoever frees a block has first made sure that no queued GPU
// work still reads or writes it (every owner synchronises its own stream
// before releasing its buffers), so a block can go to any stream.
//
// The cache holds at most kCacheBytes of device blocks per device and
// kPinnedCacheBytes of pinned blocks; beyond that a free is a real hipFree.
// When hipMalloc runs out of memory, the cached blocks of that device are
// released and the allocation is tried once more.  Cached blocks are held
// until the process ends.
#include <cstddef>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "host_common.hpp"

namespace sfa {
namespace {

constexpr size_t kCacheBytes = size_t(16) << 30;        // per device
constexpr size_t kPinnedCacheBytes = size_t(256) << 20;

struct Block {
    size_t bytes;
    int device;  // -1: pinned host memory
};

struct Pool {
    std::mutex mu;
    std::unordered_map<void*, Block> live;               // handed out
    std::map<int, std::multimap<size_t, void*>> cached;  // device (-1 pinned) -> size -> block
    std::map<int, size_t> cached_bytes;
};

Pool& pool()
{
    // never destroyed: blocks stay with the process, and no hipFree runs
    // after the HIP runtime has been torn down at exit
    static Pool* p = new Pool;
    return *p;
}

// Sizes are rounded so that blocks of nearby requests can be reused: powers
// of two below 1 MiB (at least 256 B), above that steps of 1/8 of the
// enclosing power of two.
size_t block_size(size_t bytes)
{
    if (bytes <= 256) return 256;
    size_t p2 = 1;
    while (p2 < bytes) p2 <<= 1;
    if (p2 <= (size_t(1) << 20)) return p2;
    const size_t step = p2 / 16;  // p2/2 < bytes <= p2: steps of (p2/2)/8
    return (bytes + step - 1) / step * step;
}

hipError_t raw_alloc(void** p, size_t bytes, int device)
{
    return device < 0 ? hipHostMalloc(p, bytes, hipHostMallocDefault) : hipMalloc(p, bytes);
}

void raw_free(void* p, int device)
{
    if (device < 0)
        (void)hipHostFree(p);
    else
        (void)hipFree(p);
}

// Takes the smallest cached block of at least `size` and at most 1.5x it.
void* take_cached(Pool& P, int device, size_t size, size_t* got)
{
    auto f = P.cached.find(device);
    if (f == P.cached.end()) return nullptr;
    auto it = f->second.lower_bound(size);
    if (it == f->second.end() || it->first > size + size / 2) return nullptr;
    void* p = it->second;
    *got = it->first;
    P.cached_bytes[device] -= it->first;
    f->second.erase(it);
    return p;
}

// Releases every cached block of `device` (after an out-of-memory).
void drop_cached(int device)
{
    std::vector<void*> drop;
    {
        Pool& P = pool();
        std::lock_guard<std::mutex> g(P.mu);
        auto f = P.cached.find(device);
        if (f == P.cached.end()) return;
        for (auto& kv : f->second) drop.push_back(kv.second);
        f->second.clear();
        P.cached_bytes[device] = 0;
    }
    for (void* p : drop) raw_free(p, device);
}

void* alloc_block(size_t bytes, int device)
{
    const size_t size = block_size(bytes);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> g(P.mu);
        size_t got = 0;
        if (void* p = take_cached(P, device, size, &got)) {
            P.live[p] = Block{got, device};
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = raw_alloc(&p, size, device);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();  // the failed allocation's error is not sticky for later calls
        drop_cached(device);
        e = raw_alloc(&p, size, device);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        const bool oom = e == hipErrorOutOfMemory;
        throw Error{oom ? SFA_ERR_OUT_OF_MEMORY : SFA_ERR_HIP,
                    std::string(device < 0 ? "hipHostMalloc: " : "hipMalloc: ") +
                        hipGetErrorString(e)};
    }
    std::lock_guard<std::mutex> g(P.mu);
    P.live[p] = Block{size, device};
    return p;
}

void free_block(void* p)
{
    if (!p) return;
    Pool& P = pool();
    Block b{};
    {
        std::lock_guard<std::mutex> g(P.mu);
        auto it = P.live.find(p);
        if (it == P.live.end()) return;  // not ours: nothing to do
        b = it->second;
        P.live.erase(it);
        const size_t cap = b.device < 0 ? kPinnedCacheBytes : kCacheBytes;
        size_t& held = P.cached_bytes[b.device];
        if (held + b.bytes <= cap) {
            held += b.bytes;
            P.cached[b.device].emplace(b.bytes, p);
            return;
        }
    }
    raw_free(p, b.device);
}

}  // namespace

void* device_block(size_t bytes)
{
    int device = 0;
    SFA_HIP(hipGetDevice(&device));
    return alloc_block(bytes, device);
}

void* pinned_block(size_t bytes) { return alloc_block(bytes, -1); }

void release_block(void* p) { free_block(p); }

}  // namespace sfa
