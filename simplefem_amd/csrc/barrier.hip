// barrier.hip -- what a grid-wide barrier costs inside one resident launch,
// against a kernel boundary, on this GPU.
//
// A multicolour SSOR sweep needs every colour finished, and its values
// visible to every workgroup, before the next colour starts.  The solver
// ends a kernel there (one launch per colour).  The alternative is one
// resident grid (a cooperative launch: every workgroup is running at once)
// that waits at a grid barrier between colours.  sfa_grid_barrier_cost times
// both on the same grid:
//   mode 0  one counter: every workgroup adds 1 (release, agent scope) and
//           waits until the count reaches round * workgroups (acquire);
//   mode 1  grouped: workgroups add to one of 8 counters by blockIdx % 8
//           (the XCD a workgroup is dispatched to), the last arrival of a
//           group adds 1 to the top counter, everyone waits on the top one;
//   mode 2  a kernel boundary: one launch per round, replayed from a graph.
// Every round each workgroup stores a value and, after the barrier, reads
// the value a neighbouring workgroup stored; a wrong value counts as an
// error, so the fences are checked as they are timed.  Every wait is
// bounded: a workgroup that has spun too long records a time-out and
// returns, and so does every other, so the kernel always ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>

#include "../../include/simplefem_amd.h"
#include "host_common.hpp"
#include "launch.hpp"

namespace sfa {
namespace barrier {

constexpr int kGroups = 8;
constexpr int kStride = 32;  // counters 128 B apart
constexpr long long kSpinLimit = 1LL << 20;

struct Args {
    unsigned* counters;  // [0]: top; [kStride * (1 + g)]: group g
    double* data;        // 2 x blocks
    int* flags;          // [0]: time-outs, [1]: wrong values seen
    int rounds;
};

__device__ inline bool wait_until(unsigned* p, unsigned target, int* flags)
{
    long long n = 0;
    while (__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++n > kSpinLimit) {
            __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Barrier number r (1, 2, ...) of a grid of G workgroups; false on a time-out.
template <int MODE>
__device__ bool grid_barrier(const Args& a, unsigned r, unsigned G)
{
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        bool good;
        if constexpr (MODE == 0) {
            __hip_atomic_fetch_add(&a.counters[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            good = wait_until(&a.counters[0], r * G, a.flags);
        } else {
            const unsigned ng = G < kGroups ? G : kGroups;
            const unsigned g = blockIdx.x % kGroups;
            const unsigned members = G / kGroups + (g < G % kGroups ? 1u : 0u);
            unsigned* mine = &a.counters[kStride * (1 + g)];
            const unsigned prev =
                __hip_atomic_fetch_add(mine, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (prev + 1 == r * members)
                __hip_atomic_fetch_add(&a.counters[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            good = wait_until(&a.counters[0], r * ng, a.flags);
        }
        ok = good ? 1 : 0;
    }
    __syncthreads();
    return ok != 0;
}

template <int MODE>
__global__ void k_barrier_rounds(Args a)
{
    const unsigned G = gridDim.x;
    const unsigned b = blockIdx.x;
    for (int r = 1; r <= a.rounds; r++) {
        double* buf = a.data + (size_t)(r & 1) * G;
        if (threadIdx.x == 0) buf[b] = (double)r;
        if (!grid_barrier<MODE>(a, (unsigned)r, G)) return;
        if (threadIdx.x == 0 && buf[(b + 1) % G] != (double)r)
            __hip_atomic_store(&a.flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One round per launch: read what the neighbour stored in the previous
// launch, store this round's value.
__global__ void k_round(Args a, int r)
{
    const unsigned G = gridDim.x;
    const unsigned b = blockIdx.x;
    if (threadIdx.x != 0) return;
    if (r > 1 && a.data[(size_t)((r - 1) & 1) * G + (b + 1) % G] != (double)(r - 1))
        __hip_atomic_store(&a.flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a.data[(size_t)(r & 1) * G + b] = (double)r;
}

__global__ void k_clear(Args a, int blocks)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kStride * (1 + kGroups)) a.counters[i] = 0u;
    if (i < 2 * blocks) a.data[i] = 0.;
    if (i < 2) a.flags[i] = 0;
}

double run(int blocks, int threads, int rounds, int mode, int* errors)
{
    ensure_device();
    if (mode < 0 || mode > 2) throw Error{SFA_ERR_INVALID, "mode must be 0, 1 or 2"};
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512 && threads != 1024)
        throw Error{SFA_ERR_INVALID, "threads must be 64, 128, 256, 512 or 1024"};
    if (rounds < 1 || rounds > 100000) throw Error{SFA_ERR_INVALID, "rounds must be 1..100000"};
    int device = 0, cus = 0;
    SFA_HIP(hipGetDevice(&device));
    SFA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (mode != 2) {
        int per_cu = 0;
        const void* fn = mode == 0 ? reinterpret_cast<const void*>(&k_barrier_rounds<0>)
                                   : reinterpret_cast<const void*>(&k_barrier_rounds<1>);
        SFA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
        if (blocks < 1 || blocks > per_cu * cus)
            throw Error{SFA_ERR_INVALID, "blocks must be 1.." + std::to_string(per_cu * cus) +
                                             " (all resident at once)"};
    } else if (blocks < 1 || blocks > 65536) {
        throw Error{SFA_ERR_INVALID, "blocks must be 1..65536"};
    }
    Args a{};
    a.counters = dev_alloc<unsigned>(kStride * (1 + kGroups));
    a.data = dev_alloc<double>(2 * (size_t)blocks);
    a.flags = dev_alloc<int>(2);
    a.rounds = rounds;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&] {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamSynchronize(s);  // before the buffers go back to the cache
        if (s) (void)hipStreamDestroy(s);
        dev_free(a.counters);
        dev_free(a.data);
        dev_free(a.flags);
    };
    try {
        SFA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        SFA_HIP(hipEventCreate(&e0));
        SFA_HIP(hipEventCreate(&e1));
        const int clear_n = std::max(kStride * (1 + kGroups), 2 * blocks);
        auto clear = [&] {
            hipLaunchKernelGGL(k_clear, dim3((clear_n + 255) / 256), dim3(256), 0, s, a, blocks);
        };
        float ms = 0.f;
        if (mode == 2) {
            std::unique_ptr<GraphRecorder> rec(new GraphRecorder());
            for (int r = 1; r <= rounds; r++) launch(k_round, dim3(blocks), dim3(threads), s, a, r);
            const hipError_t e = rec->instantiate(&exec);
            if (e != hipSuccess) throw Error{SFA_ERR_HIP, std::string("graph: ") + hipGetErrorString(e)};
            clear();
            SFA_HIP(hipGraphLaunch(exec, s));  // warm-up
            clear();
            SFA_HIP(hipEventRecord(e0, s));
            SFA_HIP(hipGraphLaunch(exec, s));
            SFA_HIP(hipEventRecord(e1, s));
        } else {
            void* params[] = {&a};
            const void* fn = mode == 0 ? reinterpret_cast<const void*>(&k_barrier_rounds<0>)
                                       : reinterpret_cast<const void*>(&k_barrier_rounds<1>);
            clear();  // warm-up launch
            SFA_HIP(hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(threads), params, 0, s));
            clear();
            SFA_HIP(hipEventRecord(e0, s));
            SFA_HIP(hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(threads), params, 0, s));
            SFA_HIP(hipEventRecord(e1, s));
        }
        SFA_HIP(hipEventSynchronize(e1));
        SFA_HIP(hipEventElapsedTime(&ms, e0, e1));
        int flags[2] = {0, 0};
        SFA_HIP(hipMemcpy(flags, a.flags, sizeof(flags), hipMemcpyDeviceToHost));
        *errors = (flags[0] ? 1 : 0) | (flags[1] ? 2 : 0);
        cleanup();
        return ms * 1e3 / rounds;
    } catch (...) {
        (void)hipDeviceSynchronize();
        cleanup();
        throw;
    }
}

}  // namespace barrier
}  // namespace sfa

extern "C" int sfa_grid_barrier_cost(int blocks, int threads, int rounds, int mode,
                                     double* us_per_round, int* errors)
{
    return sfa::guarded([&] {
        if (!us_per_round || !errors) throw sfa::Error{SFA_ERR_INVALID, "null output"};
        *errors = 0;
        *us_per_round = sfa::barrier::run(blocks, threads, rounds, mode, errors);
    });
}
