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

// Stores the block's sum of v at *out (thread 0, plain store).  Every thread
// of the block must call it.  Waves are added in wave order.
template <int BS>
__device__ inline void block_partial(double v, double* out)
{
    constexpr int NW = BS / kWave;
    const double s = wave_sum(v);
    if constexpr (NW == 1) {
        if (threadIdx.x == 0) *out = s;
    } else {
        __shared__ double ws[NW];
        if (threadIdx.x % kWave == 0) ws[threadIdx.x / kWave] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.;
            for (int w = 0; w < NW; w++) t += ws[w];
            *out = t;
        }
    }
}

// Sum of v[0..count) in a fixed order (lane l adds v[l], v[l+64], ..., then
// the xor butterfly), formed by the calling wavefront; every lane gets the
// same bits, and so does every block that sums the same values.
__device__ inline double wave_fold(const double* v, int count)
{
    const int lane = threadIdx.x % kWave;
    constexpr int B = 16;  // loads in flight per lane before they are added
    double a = 0.;
    for (int i0 = lane; i0 < count; i0 += B * kWave) {
        double t[B];
#pragma unroll
        for (int u = 0; u < B; u++) {
            const int i = i0 + u * kWave;
            t[u] = i < count ? v[i] : 0.;
        }
#pragma unroll
        for (int u = 0; u < B; u++)
            if (i0 + u * kWave < count) a += t[u];
    }
    return wave_sum(a);
}

// wave_fold for a block of BS threads: the first wavefront sums, the block
// shares the result.  Every thread of the block must call it.
template <int BS>
__device__ inline double block_fold(const double* v, int count)
{
    if constexpr (BS == kWave) {
        return wave_fold(v, count);
    } else {
        __shared__ double sh;
        if (threadIdx.x < kWave) {
            const double t = wave_fold(v, count);
            if (threadIdx.x == 0) sh = t;
        }
        __syncthreads();
        return sh;
    }
}

// base/base.hpp:229-237
__device__ inline double ref_absolute(double x)
{
    if (x >= 0) return x;
    if (x <= 0) return -x;
    return x;
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

// The first U entries of a lane's SELL row, loaded ahead of whatever the
// kernel reads next so that the loads overlap.  All lanes of a slice hold
// the same number of entries w (the slice is padded to its longest row), and
// w comes from the slice offsets, which are wave-uniform, so the guards below
// are scalar branches that never split a wavefront.
template <typename T, int U = 16>
struct SellChunk {
    int e = 0;  // the lane's first entry
    int w = 0;  // entries per lane
    // unsigned: widening an index for the address is then free and does not
    // make the wavefront wait for the column load
    unsigned c[U];
    T v[U];
};

// Slice `local` of a group whose slices all have gw entries per lane starts
// at gbase + local*64*gw; with gw < 0 the offsets are read from the part.
template <typename T, int U>
__device__ inline void sell_prefetch(const SellViewT<T>& A, int slice, int local, int gbase,
                                     int gw, int lane, SellChunk<T, U>& R)
{
    int o0;
    if (gw >= 0) {
        o0 = gbase + local * (gw * kWave);
        R.w = gw;
    } else {
        o0 = A.off[slice];
        R.w = (A.off[slice + 1] - o0) / kWave;
    }
    R.e = o0 + lane;
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (u < R.w) {
            R.c[u] = (unsigned)__builtin_nontemporal_load(&A.cols[R.e + u * kWave]);
            R.v[u] = __builtin_nontemporal_load(&A.vals[R.e + u * kWave]);
        } else {
            R.c[u] = 0u;
            R.v[u] = T(0);
        }
    }
}

// acc (+|-)= sum_k v_k * get(c_k) over the row of a prefetched chunk, entries
// in order; entries past the chunk are loaded here.
template <bool SUB, typename T, int U, class Get>
__device__ inline T sell_finish(const SellViewT<T>& A, const SellChunk<T, U>& R, Get get, T acc)
{
    T g[U];
#pragma unroll
    for (int u = 0; u < U; u++) g[u] = u < R.w ? get(R.c[u]) : T(0);
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (u < R.w) {
            if (SUB) acc -= R.v[u] * g[u];
            else     acc += R.v[u] * g[u];
        }
    }
    // longer rows: further chunks of 8, loads issued before the arithmetic
    constexpr int V = 8;
    for (int k0 = U; k0 < R.w; k0 += V) {
        unsigned c[V];
        T v[V], h[V];
#pragma unroll
        for (int u = 0; u < V; u++) {
            if (k0 + u < R.w) {
                c[u] = (unsigned)__builtin_nontemporal_load(&A.cols[R.e + (k0 + u) * kWave]);
                v[u] = __builtin_nontemporal_load(&A.vals[R.e + (k0 + u) * kWave]);
            }
        }
#pragma unroll
        for (int u = 0; u < V; u++) h[u] = k0 + u < R.w ? get(c[u]) : T(0);
#pragma unroll
        for (int u = 0; u < V; u++) {
            if (k0 + u < R.w) {
                if (SUB) acc -= v[u] * h[u];
                else     acc += v[u] * h[u];
            }
        }
    }
    return acc;
}

// ---- K lanes per row (K = 2, 4, 8) over the same SELL-64 storage.
// A wavefront covers rows q*64/K .. (q+1)*64/K - 1 of a slice (its part q of
// K).  A row's entries are taken in chunks of K*U: in chunk c, lane t of the
// row holds entries c*K*U + t*U + u, u < U, so one load instruction reads
// 64/K consecutive rows at K entry positions (K contiguous runs of 64/K
// elements).  Inside a chunk the lanes add their products into the row's sum
// in turn (lane 0's U products, then lane 1's, ...), the running sum passed
// from lane to lane, so the sum runs over the row's entries in order: every
// lane of the row ends with the same bits, the sum one lane forms over the
// row.  Per chunk: U loads per lane, K passes of the sum.

// Row inside the slice of a lane of part q.
template <int K>
__device__ inline int sellk_slice_row(int q, int lane)
{
    return q * (kWave / K) + lane / K;
}

// Loads chunk 0 (R.e: the lane's first entry of chunk 0).
template <int K, typename T, int U>
__device__ inline void sellk_prefetch(const SellViewT<T>& A, int slice, int local, int gbase,
                                      int gw, int q, int lane, SellChunk<T, U>& R)
{
    int o0;
    if (gw >= 0) {
        o0 = gbase + local * (gw * kWave);
        R.w = gw;
    } else {
        o0 = A.off[slice];
        R.w = (A.off[slice + 1] - o0) / kWave;
    }
    const int t = lane % K;
    R.e = o0 + sellk_slice_row<K>(q, lane) + t * U * kWave;
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (t * U + u < R.w) {
            R.c[u] = (unsigned)__builtin_nontemporal_load(&A.cols[R.e + u * kWave]);
            R.v[u] = __builtin_nontemporal_load(&A.vals[R.e + u * kWave]);
        } else {
            R.c[u] = 0u;
            R.v[u] = T(0);
        }
    }
}

template <bool SUB, typename T>
__device__ inline T sell_acc(T acc, T x)
{
    return SUB ? acc - x : acc + x;
}

// Adds the products pr[u] of chunk c (entries c*K*U + t*U + u) into acc in
// entry order, lane by lane.
template <int K, bool SUB, typename T, int U>
__device__ inline T sellk_chain(const T (&pr)[U], int c, int w, int t, int base, T acc)
{
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int e0 = (c * K + r) * U;  // first entry of lane r in this chunk
        if (e0 >= w) break;               // wave-uniform
        if (t == r) {
#pragma unroll
            for (int u = 0; u < U; u++)
                if (e0 + u < w) acc = sell_acc<SUB>(acc, pr[u]);
        }
        acc = __shfl(acc, base + r, kWave);
    }
    return acc;
}

// acc (+|-)= prod(v, c) over the row's entries, in order (see above); chunk 0
// comes prefetched in R.
template <int K, bool SUB, typename T, int U, class Prod>
__device__ inline T sellk_finish_with(const SellViewT<T>& A, const SellChunk<T, U>& R, Prod prod,
                                      T acc, int lane)
{
    const int t = lane % K;
    const int base = lane - t;
    {
        T pr[U];
#pragma unroll
        for (int u = 0; u < U; u++) pr[u] = t * U + u < R.w ? prod(R.v[u], R.c[u]) : T(0);
        acc = sellk_chain<K, SUB>(pr, 0, R.w, t, base, acc);
    }
    for (int c = 1; c * K * U < R.w; c++) {
        const int e = R.e + c * K * U * kWave;
        const int first = (c * K + t) * U;
        unsigned cc[U];
        T vv[U], pr[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (first + u < R.w) {
                cc[u] = (unsigned)__builtin_nontemporal_load(&A.cols[e + u * kWave]);
                vv[u] = __builtin_nontemporal_load(&A.vals[e + u * kWave]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) pr[u] = first + u < R.w ? prod(vv[u], cc[u]) : T(0);
        acc = sellk_chain<K, SUB>(pr, c, R.w, t, base, acc);
    }
    return acc;
}

template <int K, bool SUB, typename T, int U, class Get>
__device__ inline T sellk_finish(const SellViewT<T>& A, const SellChunk<T, U>& R, Get get, T acc,
                                 int lane)
{
    return sellk_finish_with<K, SUB>(A, R, [&](T v, unsigned c) { return v * get(c); }, acc, lane);
}

// One row of a part with K lanes per row (K = 1: sell_row), offsets read
// from the part.
template <int K, bool SUB, typename T>
__device__ inline T sellk_row(const SellViewT<T>& A, int slice, int q, int lane,
                              const T* __restrict__ vec, T acc)
{
    if constexpr (K == 1) {
        return sell_row<SUB>(A, slice, lane, vec, acc);
    } else {
        SellChunk<T, 8> R;
        sellk_prefetch<K>(A, slice, slice, 0, -1, q, lane, R);
        return sellk_finish<K, SUB>(A, R, [&](unsigned c) { return vec[c]; }, acc, lane);
    }
}

// Workgroups are dispatched to the GPU's 8 XCDs round robin (workgroup i on
// XCD i % 8), and each XCD has its own L2.  xcd_block maps workgroup i of n
// to a logical block so that every XCD gets a contiguous range of logical
// blocks: rows that a wavefront gathers are then mostly rows of its own
// XCD's range, already in that L2.  A bijection of [0, n), so only which
// hardware runs a block changes, never a result.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int i, int n)
{
    const int per = n / kXcds, rem = n % kXcds;
    const int x = i % kXcds, k = i / kXcds;
    return x * per + (x < rem ? x : rem) + k;
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

// Device row of row i (0..63) of a slice.
template <typename T>
__device__ inline int group_row_at(const SellViewT<T>& A, int slice, int local, int i, int row0,
                                   int nrows)
{
    if (row0 >= 0) {
        const int j = local * kWave + i;
        return j < nrows ? row0 + j : -1;
    }
    return A.rows[slice * kWave + i];
}


}  // namespace dev
}  // namespace sfa
