// device_common.hpp -- device helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/simplefem_amd.h"
#include "kernels.hpp"

namespace sfa {
namespace dev {

constexpr int kWave = 64;
constexpr int kSellBlock = 64;     // one wavefront = one slice per block
constexpr int kVecBlock = 256;
constexpr int kVecMaxBlocks = 2048;
constexpr double kDesiredPrecision = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON) = 2^-26
constexpr double kMachineEpsilon = 2.220446049250313e-16;     // DBL_EPSILON

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block-wide deterministic reduction across the whole grid.  Returns true in
// every thread of the last block; *total is then valid in all its threads.
template <int BS>
__device__ inline bool grid_reduce_last(double v, double* partials, unsigned* ticket, double* total)
{
    constexpr int NW = BS / kWave;
    __shared__ double ws[NW];
    __shared__ int is_last;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    double s = wave_sum(v);
    if (NW > 1) {
        if (lane == 0) ws[wave] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.;
            for (int w = 0; w < NW; w++) t += ws[w];
            s = t;
        }
    }
    if (threadIdx.x == 0) {
        __hip_atomic_store(&partials[blockIdx.x], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!is_last) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double acc = 0.;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += BS)
        acc += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    acc = wave_sum(acc);
    if (NW > 1) {
        __syncthreads();
        if (lane == 0) ws[wave] = acc;
        __syncthreads();
        double t = 0.;
        for (int w = 0; w < NW; w++) t += ws[w];
        acc = t;
    }
    *total = acc;
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// base/base.hpp:229-237
__device__ inline double ref_absolute(double x)
{
    if (x >= 0) return x;
    if (x <= 0) return -x;
    return x;
}

// Scalar step after d.Ad (solver/sparsesolver.cpp:629-645): breakdown tests
// and alpha.  One thread.
__device__ inline void after_dAd(SolveState* st, double dAd)
{
    st->dAd = dAd;
    const double zr = st->zr;
    const bool num_bad = !isfinite(zr) || zr < 0.;
    const bool den_bad = !isfinite(dAd) || dAd < 0.;
    const bool small = sqrt(ref_absolute(dAd)) < kMachineEpsilon;
    if (num_bad || den_bad) {
        st->done = 1;
        st->exit_reason = SFA_EXIT_BREAKDOWN;
    } else if (small && st->small_breaks) {
        st->done = 1;
        st->exit_reason = SFA_EXIT_SMALL;
    } else {
        st->alpha = zr / dAd;
    }
}

// Scalar step after the new z.r (solver/sparsesolver.cpp:694-708 and the
// loop head :504,601-604): beta, K++, then the loop's exit tests.  One thread.
__device__ inline void after_zr_new(SolveState* st, double zr_new)
{
    st->beta = zr_new / st->zr;
    st->zr = zr_new;
    const int K = st->K + 1;
    st->K = K;
    st->hist[K] = zr_new;
    if (K >= st->maxK) {
        st->done = 1;
        st->exit_reason = SFA_EXIT_CAP;
    } else if (ref_absolute(zr_new) < st->tol2) {
        st->done = 1;
        st->exit_reason = SFA_EXIT_CONVERGED;
    }
}

// z.r of the restart (K = 0) and the first loop-head test.
__device__ inline void after_zr_restart(SolveState* st, double zr)
{
    st->zr = zr;
    st->hist[0] = zr;
    if (st->rel2 > 0.) {  // inner solves: stop at a fixed reduction of z.r
        st->tol2 = st->rel2 * ref_absolute(zr);
        st->tol = sqrt(st->tol2);
    }
    if (ref_absolute(zr) < st->tol2) {
        st->done = 1;
        st->exit_reason = SFA_EXIT_CONVERGED;
    }
}

// acc (+|-)= sum_k v_k * vec[c_k] over one SELL row, entries in order.
// Loads of a chunk are issued before its arithmetic so the gathers overlap.
template <bool SUB, typename T>
__device__ inline T sell_row(const SellViewT<T>& A, int slice, int lane, const T* __restrict__ vec,
                             T acc)
{
    const int beg = A.off[slice] + lane;
    const int end = A.off[slice + 1];
    constexpr int U = 8;
    int e = beg;
    for (; e + (U - 1) * kWave < end; e += U * kWave) {
        T v[U], g[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = __builtin_nontemporal_load(&A.cols[e + u * kWave]);
            v[u] = __builtin_nontemporal_load(&A.vals[e + u * kWave]);
            g[u] = vec[c];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (SUB) acc -= v[u] * g[u];
            else     acc += v[u] * g[u];
        }
    }
    for (; e < end; e += kWave) {
        const int c = __builtin_nontemporal_load(&A.cols[e]);
        const T v = __builtin_nontemporal_load(&A.vals[e]);
        if (SUB) acc -= v * vec[c];
        else     acc += v * vec[c];
    }
    return acc;
}

// Device row of a lane: computed when the group's rows are consecutive
// (row0 >= 0), read from the slice's row list otherwise.
template <typename T>
__device__ inline int group_row(const SellViewT<T>& A, int slice, int local, int lane, int row0,
                                int nrows)
{
    if (row0 >= 0) {
        const int i = local * kWave + lane;
        return i < nrows ? row0 + i : -1;
    }
    return A.rows[slice * kWave + lane];
}


}  // namespace dev
}  // namespace sfa
