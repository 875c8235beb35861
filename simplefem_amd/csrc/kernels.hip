// kernels.hip -- gfx950 kernels of the CSR CG / SSOR / Rainbow-SSOR path.
//
// Roofline: every kernel here is HBM-bound (about 0.1 flop per byte); no MFMA.
// Matrix parts are SELL-64: one 64-lane wavefront owns 64 rows, lane l reads
// entry k of its row at off + 64*k + l, so each entry load of the wavefront
// is one contiguous 512-byte (values) / 256-byte (columns) segment.  Each
// row's sum runs over its entries in the reference's CSR order, with products
// rounded before the add (built with -ffp-contract=off), so a row's value
// matches the reference's loop bit for bit given the same inputs.
//
// Reductions (d.Ad, z.r) are deterministic and deferred: each block stores
// one partial, and the next kernel on the stream sums them in a fixed order
// (wave_fold) in every block that needs the total.  The scalar steps of the
// recurrence (alpha, beta, K, the exit tests) are computed identically by
// every such block; thread 0 of block 0 records them in the state for the
// kernels that follow and for the host.  Sweep kernels issue their matrix
// and row loads before they read the state, so the loads overlap it.
#include "kernels.hpp"

#include <cmath>

#include "../../include/simplefem_amd.h"
#include "device_common.hpp"
#include "launch.hpp"

namespace sfa {

GraphRecorder*& current_recorder()
{
    thread_local GraphRecorder* rec = nullptr;
    return rec;
}

int spmv_blocks(int nslices)
{
    const int b = (nslices + kSpmvWaves - 1) / kSpmvWaves;
    return b < 1 ? 1 : (b > kSpmvMaxBlocks ? kSpmvMaxBlocks : b);
}

int vec_reduce_blocks(int n)
{
    const int b = (n + dev::kVecBlock - 1) / dev::kVecBlock;
    return b < 1 ? 1 : (b > kVecReduceBlocks ? kVecReduceBlocks : b);
}

namespace {

using namespace dev;

constexpr int kSpmvBlock = kSpmvWaves * kWave;

// ------------------------------------------------------------------ scalar steps

// z.r of the restart and the first loop-head test (solver/sparsesolver.cpp:
// 586-596 and siblings).  One thread.
__device__ inline void restart_step(SolveState* st, double zr)
{
    st->zr = zr;
    st->hist[0] = zr;
    st->zrp[0] = zr;
    st->Kp[0] = 0;
    st->fold[0] = 0;
    st->K = 0;
    if (st->rel2 > 0.) {  // inner solves: stop at a fixed reduction of z.r
        st->tol2 = st->rel2 * ref_absolute(zr);
        st->tol = sqrt(st->tol2);
    }
    if (ref_absolute(zr) < st->tol2) {
        st->exit_reason = SFA_EXIT_CONVERGED;
        st->done = 1;
    }
}

// Opens the iteration of parity p.  With fold, zr_new is the z.r of the new
// iterate: beta = zr_new / zr, K+1, then the loop's exit tests
// (solver/sparsesolver.cpp:694-708 and the loop head :504,601-604).  Every
// calling block computes the same values; thread 0 of block 0 records them.
// Returns true when the loop has ended.
__device__ inline bool open_step(SolveState* st, int p, int fold, double zr_old, int K,
                                 double zr_new, double* beta_out)
{
    bool stop = false;
    int reason = SFA_EXIT_CAP;
    int Kn = K;
    double zr = zr_old;
    double beta = 0.;
    if (fold) {
        beta = zr_new / zr_old;
        zr = zr_new;
        Kn = K + 1;
        if (Kn >= st->maxK) {
            stop = true;
            reason = SFA_EXIT_CAP;
        } else if (ref_absolute(zr_new) < st->tol2) {
            stop = true;
            reason = SFA_EXIT_CONVERGED;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->Kp[1 - p] = Kn;
        st->fold[1 - p] = 1;
        st->zrp[1 - p] = zr;
        st->K = Kn;
        st->zr = zr;
        if (fold) {
            st->hist[Kn] = zr;
            st->beta = beta;
        }
        if (stop) {
            st->exit_reason = reason;
            st->done = 1;
        }
    }
    *beta_out = beta;
    return stop;
}

// After d.Ad (solver/sparsesolver.cpp:629-645): breakdown tests and
// alpha = z.r / d.Ad.  Every calling block computes the same values; thread 0
// of block 0 records them.  Returns true when the loop has ended.
__device__ inline bool dad_step(SolveState* st, double zr, double dAd, double* alpha_out)
{
    const bool num_bad = !isfinite(zr) || zr < 0.;
    const bool den_bad = !isfinite(dAd) || dAd < 0.;
    const bool small = sqrt(ref_absolute(dAd)) < kMachineEpsilon;
    int reason = -1;
    if (num_bad || den_bad) reason = SFA_EXIT_BREAKDOWN;
    else if (small && st->small_breaks) reason = SFA_EXIT_SMALL;
    const double alpha = zr / dAd;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->dAd = dAd;
        if (reason >= 0) {
            st->exit_reason = reason;
            st->done = 1;
        } else {
            st->alpha = alpha;
        }
    }
    *alpha_out = alpha;
    return reason >= 0;
}

// ------------------------------------------------------------------ vectors

template <typename T>
__global__ void k_fill(T* v, T value, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = value;
}

template <typename T>
__global__ void k_gather(T* dst, const T* src, const int* perm, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

template <typename T>
__global__ void k_scatter(T* dst, const T* src, const int* perm, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}

// Partials of sum b^2 (solver/sparsesolver.cpp:469-472).
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_sumsq(const T* b, int n, double* partials)
{
    double s = 0.;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += (double)b[i] * (double)b[i];
    block_partial<kVecBlock>(s, &partials[blockIdx.x]);
}

// sum b^2 over b[0..n) in row order (solver/sparsesolver.cpp:469-472 exactly,
// so the tolerance has the reference's bits): the block stages 2048 squares
// in LDS, then one lane adds them in order.  One block; for callers whose
// vectors are short (the inner solves of the block solvers).
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_sumsq_rows(const T* __restrict__ b, int n,
                                                         double* out)
{
    constexpr int kChunk = 8 * kVecBlock;
    __shared__ double sq[kChunk];
    double t = 0.;
    for (int i0 = 0; i0 < n; i0 += kChunk) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = u * kVecBlock + threadIdx.x;
            const double v = i0 + k < n ? (double)b[i0 + k] : 0.;
            sq[k] = v * v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = min(kChunk, n - i0);
            for (int k = 0; k < cnt; k++) t += sq[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = t;
}

// State reset and tolerance: tol = max(precision * sqrt(sum b^2), desired_precision).
__global__ void __launch_bounds__(kWave) k_init_state(SolveState* st, const double* partials,
                                                      int np, double precision, double host_tol,
                                                      int use_host_tol, int maxK,
                                                      int small_breaks, double rel2)
{
    const double total = use_host_tol ? 0. : wave_fold(partials, np);
    if (threadIdx.x != 0) return;
    double tol = host_tol;
    if (!use_host_tol) {
        tol = precision * sqrt(total);
        // base/base.hpp:275-283 maximum()
        if (!(tol >= kDesiredPrecision) && (tol <= kDesiredPrecision)) tol = kDesiredPrecision;
    }
    st->tol = tol;
    st->tol2 = tol * tol;
    st->rel2 = rel2;
    st->maxK = maxK;
    st->small_breaks = small_breaks;
    st->Kp[0] = st->Kp[1] = 0;
    st->fold[0] = st->fold[1] = 0;
    st->zrp[0] = st->zrp[1] = __builtin_nan("");
    st->alpha = 0.;
    st->beta = 0.;
    st->zr = __builtin_nan("");
    st->dAd = 0.;
    st->result = 0.;
    st->K = 0;
    st->done = 0;
    st->exit_reason = SFA_EXIT_CAP;
}

// Restart's scalar step from the restart kernel's partials.
__global__ void __launch_bounds__(kWave) k_restart_finish(SolveState* st, const double* partials,
                                                          int np)
{
    const double total = wave_fold(partials, np);
    if (threadIdx.x == 0) restart_step(st, total);
}

__global__ void __launch_bounds__(kWave) k_result_finish(SolveState* st, const double* partials,
                                                         int np)
{
    const double total = wave_fold(partials, np);
    if (threadIdx.x == 0) st->result = total;
}

// ------------------------------------------------------------------ apply

// MatrixCSR::apply (sparse/matcsr.cpp:216-244): dest = 0; dest += scaling*V*x,
// entries in order.  K lanes per row: one wavefront per task (slice
// task/K, part task%K of 64/K rows), the row's sum passed lane to lane in
// entry order, so every K gives the same bits.
template <typename T, int K>
__global__ void __launch_bounds__(kSellBlock) k_apply(SellViewT<T> A, int nslices, int row0,
                                                      int nrows, const T* src, T* dest,
                                                      T scaling)
{
    const int task = blockIdx.x;
    const int lane = threadIdx.x;
    if (task >= nslices * K) return;
    const int slice = task / K, part = task % K;
    const int i = sellk_slice_row<K>(part, lane);
    const int row = group_row_at(A, slice, slice, i, row0, nrows);
    if (row < 0) return;  // every lane of a row agrees
    T acc = 0;
    if constexpr (K == 1) {
        const int beg = A.off[slice] + lane, end = A.off[slice + 1];
        for (int e = beg; e < end; e += kWave) acc += scaling * A.vals[e] * src[A.cols[e]];
    } else {
        SellChunk<T, 8> R;
        sellk_prefetch<K>(A, slice, slice, 0, -1, part, lane, R);
        acc = sellk_finish_with<K, false>(
            A, R, [&](T v, unsigned c) { return scaling * v * src[c]; }, acc, lane);
        if (lane % K != 0) return;
    }
    dest[row] = acc;
}

// ------------------------------------------------------------------ restart

// solver/sparsesolver.cpp:80-98 (plain), :281-306 (Jacobi), :512-524 (SSOR),
// :797-812 (Eisenstat, columns already filtered into the part).
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_residual(int mode, SellViewT<T> A, int nslices,
                                                         int row0, int nrows, const T* b,
                                                         const T* x, T* out, T* d, T* z,
                                                         const T* precon, double* partials)
{
    const int slice = blockIdx.x;
    const int lane = threadIdx.x;
    double part = 0.;
    const int row = slice < nslices ? group_row(A, slice, slice, lane, row0, nrows) : -1;
    if (row >= 0) {
        const T r = sell_row<true>(A, slice, lane, x, b[row]);
        out[row] = r;
        if (mode == RES_PLAIN) {
            d[row] = r;
            part = r * r;
        } else if (mode == RES_JACOBI) {
            const T zz = precon[row] * r;
            z[row] = zz;
            d[row] = zz;
            part = zz * r;
        }
    }
    if (mode != RES_ONLY) block_partial<kSellBlock>(part, &partials[blockIdx.x]);
}

// solver/sparsesolver.cpp:570-583: d = z, partials of z.r.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_direction_init(int n, const T* z, const T* r, T* d,
                                                              double* partials)
{
    double part = 0.;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T zz = z[i];
        d[i] = zz;
        part += (double)(zz * r[i]);
    }
    block_partial<kVecBlock>(part, &partials[blockIdx.x]);
}

// Continues an earlier CG's direction after its residual was replaced (the
// restart has just made d = z and z.r of the new residual, st->zr):
// d = z + beta d_prev with beta = z.r / zr_prev, zr_prev the z.r of the
// iterate that formed d_prev -- the iteration's own direction update with
// the replaced residual in place of the recursive one.  The loop's cap
// becomes maxK.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_direction_continue(SolveState* st, int n, T* d,
                                                                  const T* dprev, double zr_prev,
                                                                  int maxK)
{
    const T beta = (T)(st->zr / zr_prev);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->maxK = maxK;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        d[i] = d[i] + beta * dprev[i];
}

// solver/sparsesolver.cpp:830-844: z = r*diag*(2-w)/w, d = z, partials of z.r.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_eisenstat_init(int n, const T* r, const T* diag,
                                                              T omega, T* z, T* d,
                                                              double* partials)
{
    double part = 0.;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T rr = r[i];
        const T zz = rr * diag[i] * (T(2) - omega) / omega;
        z[i] = zz;
        d[i] = zz;
        part += (double)(zz * rr);
    }
    block_partial<kVecBlock>(part, &partials[blockIdx.x]);
}

// ------------------------------------------------------------------ sweeps

// Forward sweep of one colour / level, solver/sparsesolver.cpp:528-541 (restart)
// and :629-665, :1234-1261 (iteration: d.Ad step in the first group, then the
// x and r updates of the group's rows).  With zb set, the group is also the
// first group of the backward sweep and has no upper entries (the last colour
// of a Rainbow schedule): its backward rows z = (m*diag*c2)*w/diag (:667-692)
// are formed here, with their z.r partials at zpart[block] when zpart is set.
// With m0_nrows > 0 the kernel also runs the sweep's first group, rows m0_row0
// .. m0_row0+m0_nrows-1, which has no lower entries: extra blocks write its m
// = (r - a q) * w / diag (its x and r are updated later, by its backward
// kernel), and this group's rows form those m values themselves, with the
// same arithmetic, where they gather them.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_forward(SellViewT<T> L, int slice0, int nslices,
                                                        int row0, int nrows, SolveState* st,
                                                        int mode, int p, T* x, T* r,
                                                        const T* d, const T* q, T* m,
                                                        const T* diag, T omega,
                                                        const double* spmv_partials, int nspmv,
                                                        T* zb, T c2, double* zpart, int gbase,
                                                        int gw, int m0_row0, int m0_nrows)
{
    const int lane = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int slice = slice0 + b;
    const bool has = b < nslices;
    const bool first_group = b >= nslices && m0_nrows > 0;  // a block of the merged group
    int row = has ? group_row(L, slice, b, lane, row0, nrows) : -1;
    if (first_group) {
        const int i = (b - nslices) * kWave + lane;
        row = i < m0_nrows ? m0_row0 + i : -1;
    }
    const int ri = row >= 0 ? row : L.zero;  // padding lanes read the zero slot
    T rr = r[ri];
    const T dg = diag[ri];
    T xx = T(0), dd = T(0), qq = T(0);
    if (mode != FWD_RESTART) {
        xx = x[ri];
        dd = d[ri];
        qq = q[ri];
    }
    SellChunk<T> R;
    if (has) sell_prefetch(L, slice, b, gbase, gw, lane, R);
    if (st->done) return;
    T alpha = T(0);
    if (mode == FWD_FIRST) {
        const double dAd = wave_fold(spmv_partials, nspmv);
        double a;
        if (dad_step(st, st->zrp[1 - p], dAd, &a)) return;
        alpha = (T)a;
    } else if (mode == FWD_UPDATE) {
        alpha = (T)st->alpha;
    }
    if (first_group) {  // no lower entries: m = (r - a q) * w / diag
        if (row >= 0) m[row] = (mode != FWD_RESTART ? rr - alpha * qq : rr) * omega / dg;
        return;
    }
    // m of a row of the merged first group, formed as its own block forms it
    auto get_m = [&](unsigned c) -> T {
        if (c - (unsigned)m0_row0 < (unsigned)m0_nrows) {
            const T rc = r[c];
            return (mode != FWD_RESTART ? rc - alpha * q[c] : rc) * omega / diag[c];
        }
        return m[c];
    };
    double part = 0.;
    if (row >= 0) {
        if (mode != FWD_RESTART) {
            x[row] = xx + alpha * dd;
            rr = rr - alpha * qq;
            r[row] = rr;
        }
        const T acc = sell_finish<true>(L, R, get_m, rr);
        const T mm = acc * omega / dg;
        m[row] = mm;
        if (zb) {
            T t = mm;
            t *= dg;
            t *= c2;
            const T zz = t * omega / dg;
            zb[row] = zz;
            part = (double)(zz * rr);
        }
    }
    if (zpart && has) block_partial<kSellBlock>(part, &zpart[b]);
}

// Backward sweep of one colour / level, solver/sparsesolver.cpp:543-568
// (restart) and :667-692, :1263-1303 (iteration): the scaling of m fused in.
// With reduce, the z.r partial of each block goes to gpart[slice], and extra
// blocks (logical blocks nslices, nslices+1) sum earlier groups' partials
// into *prev_sum (and *prev_sumb), so z.r is summed per group and over the
// groups in launch order.
// With defer_update the group's x and r updates of the iteration (x += a d,
// r -= a q) are made here, first (its forward rows were run merged into the
// next group's kernel, which leaves them).
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_backward(SellViewT<T> U, int slice0, int nslices,
                                                         int row0, int nrows, SolveState* st,
                                                         int reduce, T* r, const T* m,
                                                         T* z, const T* diag, T omega, T c2,
                                                         double* gpart, int prev0, int prevn,
                                                         double* prev_sum, int prev0b,
                                                         int prevnb, double* prev_sumb,
                                                         int gbase, int gw,
                                                         int defer_update, T* x, const T* d,
                                                         const T* q)
{
    const int lane = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int slice = slice0 + b;
    const bool has = b < nslices;
    const int row = has ? group_row(U, slice, b, lane, row0, nrows) : -1;
    const int ri = row >= 0 ? row : U.zero;  // padding lanes read the zero slot
    const T dg = diag[ri];
    const T mm = m[ri];
    T rr = reduce ? r[ri] : T(0);
    T xx = T(0), dd = T(0), qq = T(0);
    if (defer_update) {
        xx = x[ri];
        dd = d[ri];
        qq = q[ri];
    }
    SellChunk<T> R;
    if (has) sell_prefetch(U, slice, b, gbase, gw, lane, R);
    if (st->done) return;
    if (defer_update && row >= 0) {
        const T alpha = (T)st->alpha;
        x[row] = xx + alpha * dd;
        rr = rr - alpha * qq;
        r[row] = rr;
    }
    if (reduce && prev_sum && b == nslices) {
        const double t = wave_fold(gpart + prev0, prevn);
        if (lane == 0) *prev_sum = t;
        return;
    }
    if (reduce && prev_sumb && b == nslices + 1) {
        const double t = wave_fold(gpart + prev0b, prevnb);
        if (lane == 0) *prev_sumb = t;
        return;
    }
    double part = 0.;
    if (row >= 0) {
        T t = mm;
        t *= dg;
        t *= c2;
        const T acc = sell_finish<true>(U, R, [&](unsigned c) { return z[c]; }, t);
        const T zz = acc * omega / dg;
        z[row] = zz;
        part = (double)(zz * rr);
    }
    if (reduce && has) block_partial<kSellBlock>(part, &gpart[slice]);
}

// ------------------------------------------------------------------ sweep pairs

constexpr int kPairBlock = kPairWaves * kWave;

// m of row c of group A as A's wavefront forms it in the iteration
// (k_forward, FWD_UPDATE; k_forward_pair): (r - a q - L m) * w / diag, the
// entries in order, from the m of earlier groups.
template <typename T>
__device__ inline T pair_forward_m(const SellViewT<T>& L, const SweepGroup& A, unsigned c,
                                   T alpha, const T* r, const T* q, const T* m, const T* diag,
                                   T omega)
{
    const int pos = (int)(c - (unsigned)A.row0);
    const int local = pos / kWave;
    SellChunk<T> R;
    sell_prefetch(L, A.slice0 + local, local, A.gbase, A.gw, pos % kWave, R);
    const T rc = r[c] - alpha * q[c];
    const T acc = sell_finish<true>(L, R, [&](unsigned k) { return m[k]; }, rc);
    return acc * omega / diag[c];
}

// z of row c of group A as A's wavefront forms it (k_backward): (m*diag*c2
// - U z) * w / diag, from the z of earlier groups.
template <typename T>
__device__ inline T pair_backward_z(const SellViewT<T>& U, const SweepGroup& A, unsigned c,
                                    const T* m, const T* z, const T* diag, T omega, T c2)
{
    const int pos = (int)(c - (unsigned)A.row0);
    const int local = pos / kWave;
    SellChunk<T> R;
    sell_prefetch(U, A.slice0 + local, local, A.gbase, A.gw, pos % kWave, R);
    const T dg = diag[c];
    T t = m[c];
    t *= dg;
    t *= c2;
    const T acc = sell_finish<true>(U, R, [&](unsigned k) { return z[k]; }, t);
    return acc * omega / dg;
}

// Wavefront 0 of a B block: the entries of its row that read A rows.  Their
// columns go to acol[k][lane] (k in entry order), their count to cnt[lane];
// returns the entries' bit mask.
template <typename T, int U>
__device__ inline unsigned pair_collect(const SellChunk<T, U>& R, const SweepGroup& A, int row,
                                        int lane, unsigned (*acol)[kWave], int* cnt)
{
    unsigned mask = 0u;
    int k = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
        if (u < R.w && R.c[u] - (unsigned)A.row0 < (unsigned)A.nrows) {
            mask |= 1u << u;
            if (k < kPairRefs) acol[k][lane] = R.c[u];
            k++;
        }
    }
    cnt[lane] = row >= 0 ? min(k, kPairRefs) : 0;
    return mask;
}

// Forward sweep of groups A and B in one launch (see launch_forward_pair).
// Blocks [0, ceil(A.nslices/4)): one A slice per wavefront.  Then one block
// per B slice: wavefront 0 owns the slice's rows and lists the A rows they
// read; all four wavefronts form those rows' m into LDS; wavefront 0 then
// runs the rows, taking A's m from LDS and the other m from memory, in entry
// order.  A's x and r are left for A's backward kernel: B's blocks read A's r
// here, before any update.
template <typename T>
__global__ void __launch_bounds__(kPairBlock) k_forward_pair(SellViewT<T> L, SweepGroup A,
                                                             SweepGroup B, SolveState* st, T* x,
                                                             T* r, const T* d, const T* q, T* m,
                                                             const T* diag, T omega, T* zb, T c2,
                                                             double* zpart)
{
    __shared__ unsigned acol[kPairRefs][kWave];
    __shared__ T aval[kPairRefs][kWave];
    __shared__ int acnt[kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nab = (A.nslices + kPairWaves - 1) / kPairWaves;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    if (b < nab) {
        const int local = b * kPairWaves + wave;
        if (local >= A.nslices) return;
        const int slice = A.slice0 + local;
        const int row = group_row(L, slice, local, lane, A.row0, A.nrows);
        const int ri = row >= 0 ? row : L.zero;
        const T rr = r[ri];
        const T qq = q[ri];
        const T dg = diag[ri];
        SellChunk<T> R;
        sell_prefetch(L, slice, local, A.gbase, A.gw, lane, R);
        if (st->done) return;
        const T alpha = (T)st->alpha;
        if (row < 0) return;
        const T rn = rr - alpha * qq;
        const T acc = sell_finish<true>(L, R, [&](unsigned k) { return m[k]; }, rn);
        m[row] = acc * omega / dg;
        return;
    }
    const int local = b - nab;
    const int slice = B.slice0 + local;
    int row = -1;
    T rr = T(0), dg = T(0), xx = T(0), dd = T(0), qq = T(0);
    SellChunk<T> R;
    if (wave == 0) {
        row = group_row(L, slice, local, lane, B.row0, B.nrows);
        const int ri = row >= 0 ? row : L.zero;
        rr = r[ri];
        dg = diag[ri];
        xx = x[ri];
        dd = d[ri];
        qq = q[ri];
        sell_prefetch(L, slice, local, B.gbase, B.gw, lane, R);
    }
    if (__syncthreads_or(st->done)) return;
    const T alpha = (T)st->alpha;
    unsigned mask = 0u;
    if (wave == 0) mask = pair_collect(R, A, row, lane, acol, acnt);
    __syncthreads();
    T g[kPairWidth];
    if (wave == 0) {
#pragma unroll
        for (int u = 0; u < kPairWidth; u++)
            g[u] = u < R.w && !(mask >> u & 1u) ? m[R.c[u]] : T(0);
    }
    const int cnt = acnt[lane];
#pragma unroll
    for (int k = wave; k < kPairRefs; k += kPairWaves)
        if (k < cnt) aval[k][lane] = pair_forward_m(L, A, acol[k][lane], alpha, r, q, m, diag, omega);
    __syncthreads();
    if (wave != 0) return;
    double part = 0.;
    if (row >= 0) {
        x[row] = xx + alpha * dd;
        const T rn = rr - alpha * qq;
        r[row] = rn;
        int k = 0;
#pragma unroll
        for (int u = 0; u < kPairWidth; u++)
            if (mask >> u & 1u) g[u] = aval[k++][lane];
        T acc = rn;
#pragma unroll
        for (int u = 0; u < kPairWidth; u++)
            if (u < R.w) acc -= R.v[u] * g[u];
        const T mm = acc * omega / dg;
        m[row] = mm;
        if (zb) {
            T t = mm;
            t *= dg;
            t *= c2;
            const T zz = t * omega / dg;
            zb[row] = zz;
            part = (double)(zz * rn);
        }
    }
    if (zpart) block_partial<kSellBlock>(part, &zpart[local]);
}

// Backward sweep of groups A and B in one launch, with the z.r partials
// (see launch_backward_pair); the block layout of k_forward_pair, plus one
// block that sums earlier groups' partials.
template <typename T>
__global__ void __launch_bounds__(kPairBlock) k_backward_pair(
    SellViewT<T> U, SweepGroup A, SweepGroup B, SolveState* st, T* r, const T* m, T* z,
    const T* diag, T omega, T c2, double* gpart, int nfold, int fold0a, int foldna,
    double* fsuma, int fold0b, int foldnb, double* fsumb, int defer_a, int defer_b, T* x,
    const T* d, const T* q)
{
    __shared__ unsigned acol[kPairRefs][kWave];
    __shared__ T aval[kPairRefs][kWave];
    __shared__ int acnt[kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nab = (A.nslices + kPairWaves - 1) / kPairWaves;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    if (b < nab) {
        const int local = b * kPairWaves + wave;
        if (local >= A.nslices) return;
        const int slice = A.slice0 + local;
        const int row = group_row(U, slice, local, lane, A.row0, A.nrows);
        const int ri = row >= 0 ? row : U.zero;
        const T dg = diag[ri];
        const T mm = m[ri];
        T rr = r[ri];
        T xx = T(0), dd = T(0), qq = T(0);
        if (defer_a) {
            xx = x[ri];
            dd = d[ri];
            qq = q[ri];
        }
        SellChunk<T> R;
        sell_prefetch(U, slice, local, A.gbase, A.gw, lane, R);
        if (st->done) return;
        double part = 0.;
        if (row >= 0) {
            if (defer_a) {
                const T alpha = (T)st->alpha;
                x[row] = xx + alpha * dd;
                rr = rr - alpha * qq;
                r[row] = rr;
            }
            T t = mm;
            t *= dg;
            t *= c2;
            const T acc = sell_finish<true>(U, R, [&](unsigned k) { return z[k]; }, t);
            const T zz = acc * omega / dg;
            z[row] = zz;
            part = (double)(zz * rr);
        }
        const double ws = wave_sum(part);  // the slice's partial, as k_backward's block forms it
        if (lane == 0) gpart[slice] = ws;
        return;
    }
    if (b >= nab + B.nslices) {  // sums of earlier groups
        if (st->done) return;
        if (wave == 0 && nfold > 0) {
            const double t = wave_fold(gpart + fold0a, foldna);
            if (lane == 0) *fsuma = t;
        } else if (wave == 1 && nfold > 1) {
            const double t = wave_fold(gpart + fold0b, foldnb);
            if (lane == 0) *fsumb = t;
        }
        return;
    }
    const int local = b - nab;
    const int slice = B.slice0 + local;
    int row = -1;
    T dg = T(0), mm = T(0), rr = T(0), xx = T(0), dd = T(0), qq = T(0);
    SellChunk<T> R;
    if (wave == 0) {
        row = group_row(U, slice, local, lane, B.row0, B.nrows);
        const int ri = row >= 0 ? row : U.zero;
        dg = diag[ri];
        mm = m[ri];
        rr = r[ri];
        if (defer_b) {
            xx = x[ri];
            dd = d[ri];
            qq = q[ri];
        }
        sell_prefetch(U, slice, local, B.gbase, B.gw, lane, R);
    }
    if (__syncthreads_or(st->done)) return;
    unsigned mask = 0u;
    if (wave == 0) mask = pair_collect(R, A, row, lane, acol, acnt);
    __syncthreads();
    T g[kPairWidth];
    if (wave == 0) {
#pragma unroll
        for (int u = 0; u < kPairWidth; u++)
            g[u] = u < R.w && !(mask >> u & 1u) ? z[R.c[u]] : T(0);
    }
    const int cnt = acnt[lane];
#pragma unroll
    for (int k = wave; k < kPairRefs; k += kPairWaves)
        if (k < cnt) aval[k][lane] = pair_backward_z(U, A, acol[k][lane], m, z, diag, omega, c2);
    __syncthreads();
    if (wave != 0) return;
    double part = 0.;
    if (row >= 0) {
        if (defer_b) {
            const T alpha = (T)st->alpha;
            x[row] = xx + alpha * dd;
            rr = rr - alpha * qq;
            r[row] = rr;
        }
        int k = 0;
#pragma unroll
        for (int u = 0; u < kPairWidth; u++)
            if (mask >> u & 1u) g[u] = aval[k++][lane];
        T t = mm;
        t *= dg;
        t *= c2;
        T acc = t;
#pragma unroll
        for (int u = 0; u < kPairWidth; u++)
            if (u < R.w) acc -= R.v[u] * g[u];
        const T zz = acc * omega / dg;
        z[row] = zz;
        part = (double)(zz * rr);
    }
    block_partial<kSellBlock>(part, &gpart[slice]);
}

// Eisenstat triangular solves (solver/sparsesolver.cpp:814-828, 876-908,
// 1429-1455, 1501-1560): out = (in - [c3*diag*t] - P out) * w / diag.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_eisenstat_solve(SellViewT<T> P, int slice0,
                                                                int nslices, int row0, int nrows,
                                                                const SolveState* st, int with_t,
                                                                const T* in, const T* t,
                                                                T* out, const T* diag,
                                                                T omega, T c3, int gbase, int gw)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int slice = slice0 + b;
    const int lane = threadIdx.x;
    if (b >= nslices) return;
    const int row = group_row(P, slice, b, lane, row0, nrows);
    const int ri = row >= 0 ? row : P.zero;  // padding lanes read the zero slot
    const T dg = diag[ri];
    T a = in[ri];
    const T tt = with_t ? t[ri] : T(0);
    SellChunk<T> R;
    sell_prefetch(P, slice, b, gbase, gw, lane, R);
    if (st->done) return;
    if (row < 0) return;
    if (with_t) a = a - c3 * dg * tt;
    const T acc = sell_finish<true>(P, R, [&](unsigned c) { return out[c]; }, a);
    out[row] = acc * omega / dg;
}

// Eisenstat upper solve (solver/sparsesolver.cpp:876-892, 1429-1455) with the
// previous iteration's direction update (:960-966) fused in.  The first
// group opens the iteration from the update kernel's z.r partials.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_eisenstat_upper(SellViewT<T> P, int slice0,
                                                                int nslices, int row0, int nrows,
                                                                SolveState* st, int first, int p,
                                                                const T* z, const T* dold,
                                                                T* dnew, T* out, const T* diag,
                                                                T omega, const double* fpart,
                                                                int nfold, int gbase, int gw)
{
    const int lane = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int slice = slice0 + b;
    const bool has = b < nslices;
    const int row = has ? group_row(P, slice, b, lane, row0, nrows) : -1;
    const int ri = row >= 0 ? row : P.zero;  // padding lanes read the zero slot
    const T dg = diag[ri];
    const T zz = z[ri];
    const T dd = dold[ri];
    SellChunk<T> R;
    if (has) sell_prefetch(P, slice, b, gbase, gw, lane, R);
    if (st->done) return;
    const int fold = st->fold[p];
    T beta;
    if (first) {
        const double zr_new = fold ? wave_fold(fpart, nfold) : 0.;
        double b;
        if (open_step(st, p, fold, st->zrp[p], st->Kp[p], zr_new, &b)) return;
        beta = (T)b;
    } else {
        beta = (T)st->beta;
    }
    if (row < 0) return;
    const T dn = fold ? zz + beta * dd : dd;
    dnew[row] = dn;
    const T acc = sell_finish<true>(P, R, [&](unsigned c) { return out[c]; }, dn);
    out[row] = acc * omega / dg;
}

// ------------------------------------------------------------------ iteration

// Opens the iteration (z.r of the previous sweep, beta, K, exit tests), then
// dn = src + beta*dold (solver/sparsesolver.cpp:698-706 and its siblings),
// q = A dn and the d.Ad partials (:611-626, 1197-1212).  With K lanes per row
// a slice is K tasks of 64/K rows (task i: slice i/K, part i%K); each
// wavefront owns tasks b*4 + wave + k*4*gridDim.x (b: the block's logical
// index, see xcd_block).  Neighbours' dn are formed from src and dold where
// they are gathered, with the same arithmetic as the owner's, so every copy
// has the same bits.
template <typename T, int K>
__global__ void __launch_bounds__(kSpmvBlock) k_spmv_dir(SellViewT<T> A, int nslices, int row0,
                                                         int nrows, int n_active,
                                                         SolveState* st, int p,
                                                         const T* src, const T* dold, T* dnew,
                                                         T* q, double* spart,
                                                         const double* fpart, int nfold,
                                                         const double* fpartb, int nfoldb,
                                                         const double* gsum, int ngroups,
                                                         int gw)
{
    // the wavefront index is uniform: readfirstlane lets the slice offsets be
    // scalar loads
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int ntasks = nslices * K;
    const int first_task = xcd_block(blockIdx.x, gridDim.x) * kSpmvWaves + wave;
    const int stride = gridDim.x * kSpmvWaves;
    // the lane that stores its row's results (lane 0 of the row's K lanes)
    const bool owner = lane % K == 0;
    // the first task's own-row values and entries, loaded ahead of the state
    int row = -1;
    T s_own = T(0), d_own = T(0);
    SellChunk<T, 8> R;  // 8: keeps every block of the grid resident
    if (first_task < ntasks) {
        const int slice = first_task / K, part = first_task % K;
        row = group_row_at(A, slice, slice, sellk_slice_row<K>(part, lane), row0, nrows);
        const int ri = row >= 0 ? row : A.zero;
        s_own = src[ri];
        d_own = dold[ri];
        if constexpr (K == 1) sell_prefetch(A, slice, slice, 0, gw, lane, R);
        else sellk_prefetch<K>(A, slice, slice, 0, gw, part, lane, R);
    }
    const int fold = st->fold[p];
    const int Kit = st->Kp[p];
    const double zr_old = st->zrp[p];
    // Block 0 may set done while other blocks start; the vote keeps the
    // waves of a block together (a set flag means the loop has ended).
    if (__syncthreads_or(st->done)) return;
    double zr_new = 0.;
    if (fold) {
        __shared__ double sh;
        if (wave == 0) {
            double t = 0.;
            for (int g = 0; g < ngroups; g++) t += gsum[g];
            t += wave_fold(fpart, nfold);
            if (fpartb) t += wave_fold(fpartb, nfoldb);
            if (lane == 0) sh = t;
        }
        __syncthreads();
        zr_new = sh;
    }
    double beta_d;
    if (open_step(st, p, fold, zr_old, Kit, zr_new, &beta_d)) return;
    const T beta = (T)beta_d;
    // rows past n_active (inactive rows, the zero slot) keep their d
    auto dn = [&](unsigned c) -> T {
        return fold && c < (unsigned)n_active ? src[c] + beta * dold[c] : dold[c];
    };
    double part_sum = 0.;
    for (int task = first_task; task < ntasks; task += stride) {
        if (task != first_task) {
            const int slice = task / K, part = task % K;
            row = group_row_at(A, slice, slice, sellk_slice_row<K>(part, lane), row0, nrows);
            const int ri = row >= 0 ? row : A.zero;
            s_own = src[ri];
            d_own = dold[ri];
            if constexpr (K == 1) sell_prefetch(A, slice, slice, 0, gw, lane, R);
            else sellk_prefetch<K>(A, slice, slice, 0, gw, part, lane, R);
        }
        T acc;
        if constexpr (K == 1) acc = sell_finish<false>(A, R, dn, T(0));
        else acc = sellk_finish<K, false>(A, R, dn, T(0), lane);
        if (row >= 0 && owner) {
            const T dr = fold && row < n_active ? s_own + beta * d_own : d_own;  // = dn(row)
            dnew[row] = dr;
            q[row] = acc;
            part_sum += (double)(dr * acc);
        }
    }
    block_partial<kSpmvBlock>(part_sum, &spart[blockIdx.x]);
}

// d.Ad step from dpart, then x += a dx; r -= a q; z.r by mode
// (solver/sparsesolver.cpp:160-174, 372-389, 952-968) into zpart.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_update(int mode, SolveState* st, int p, int n, T* x,
                                                      T* r, const T* dx, const T* q, T* z,
                                                      const T* dp, T omega,
                                                      const double* dpart, int ndad,
                                                      double* zpart)
{
    if (__syncthreads_or(st->done)) return;  // see k_spmv_dir
    const double zr = st->zrp[1 - p];
    const double dAd = block_fold<kVecBlock>(dpart, ndad);
    double a;
    if (dad_step(st, zr, dAd, &a)) return;
    const T alpha = (T)a;
    double part = 0.;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * dx[i];
        const T rr = r[i] - alpha * q[i];
        r[i] = rr;
        if (mode == UPD_PLAIN) {
            part += (double)(rr * rr);
        } else if (mode == UPD_JACOBI) {
            const T zz = dp[i] * rr;
            z[i] = zz;
            part += (double)(zz * rr);
        } else {
            const T zz = rr * (dp[i] * (T(2) - omega) / omega);
            z[i] = zz;
            part += (double)(zz * rr);
        }
    }
    block_partial<kVecBlock>(part, &zpart[blockIdx.x]);
}

// solver/sparsesolver.cpp:912-922: q += t, partials of d.Ad = sum q*d.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_eisenstat_combine(const SolveState* st, int n, T* q,
                                                                 const T* t, const T* d,
                                                                 double* partials)
{
    if (st->done) return;
    double part = 0.;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T qq = q[i] + t[i];
        q[i] = qq;
        part += (double)(qq * d[i]);
    }
    block_partial<kVecBlock>(part, &partials[blockIdx.x]);
}

// ------------------------------------------------------------------ mixed precision

// out = b - A x in the part's rows, partials of the sum of squares.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_residual_norm(SellViewT<T> A, int nslices, int row0,
                                                              int nrows, const T* b, const T* x,
                                                              T* out, double* partials)
{
    const int slice = blockIdx.x;
    const int lane = threadIdx.x;
    double part = 0.;
    const int row = slice < nslices ? group_row(A, slice, slice, lane, row0, nrows) : -1;
    if (row >= 0) {
        const T r = sell_row<true>(A, slice, lane, x, b[row]);
        out[row] = r;
        part = (double)(r * r);
    }
    block_partial<kSellBlock>(part, &partials[blockIdx.x]);
}

template <typename TD, typename TS>
__global__ void k_convert(TD* dst, const TS* src, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (TD)src[i];
}

// x += (double)e over the first n rows
__global__ void k_add_correction(double* x, const float* e, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)e[i];
}

int flat_blocks(int n)
{
    int b = (n + kVecBlock - 1) / kVecBlock;
    return b < 1 ? 1 : b;
}

}  // namespace

template <typename T>
void launch_fill(hipStream_t s, T* v, T value, int n)
{
    if (n > 0) launch(k_fill<T>, dim3(flat_blocks(n)), dim3(kVecBlock), s, v, value, n);
}

template <typename T>
void launch_gather(hipStream_t s, T* dst, const T* src, const int* perm, int n)
{
    if (n > 0) launch(k_gather<T>, dim3(flat_blocks(n)), dim3(kVecBlock), s, dst, src, perm, n);
}

template <typename T>
void launch_scatter(hipStream_t s, T* dst, const T* src, const int* perm, int n)
{
    if (n > 0) launch(k_scatter<T>, dim3(flat_blocks(n)), dim3(kVecBlock), s, dst, src, perm, n);
}

template <typename T>
void launch_init_state(hipStream_t s, SolveState* st, const T* b, int n, double precision,
                       double host_tol, int use_host_tol, int maxK, int small_breaks, double rel2,
                       double* partials, const T* rows_b)
{
    int np = vec_reduce_blocks(n);
    if (!use_host_tol && rows_b) {  // the reference's row order
        launch(k_sumsq_rows<T>, dim3(1), dim3(kVecBlock), s, rows_b, n, partials);
        np = 1;
    } else if (!use_host_tol) {
        launch(k_sumsq<T>, dim3(np), dim3(kVecBlock), s, b, n, partials);
    }
    launch(k_init_state, dim3(1), dim3(kWave), s, st, partials, np, precision, host_tol, use_host_tol, maxK,
                                     small_breaks, rel2);
}

template <typename T>
void launch_apply(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows, const T* src,
                  T* dest, T scaling, int lanes)
{
    if (nslices <= 0) return;
    const dim3 grid(nslices * lanes), block(kSellBlock);
    switch (lanes) {
    case 1: launch(k_apply<T, 1>, grid, block, s, A, nslices, row0, nrows, src, dest, scaling); break;
    case 2: launch(k_apply<T, 2>, grid, block, s, A, nslices, row0, nrows, src, dest, scaling); break;
    case 4: launch(k_apply<T, 4>, grid, block, s, A, nslices, row0, nrows, src, dest, scaling); break;
    case 8: launch(k_apply<T, 8>, grid, block, s, A, nslices, row0, nrows, src, dest, scaling); break;
    default: break;
    }
}

template <typename T>
void launch_residual(hipStream_t s, ResidualMode mode, SellViewT<T> A, int nslices, int row0,
                     int nrows, SolveState* st, const T* b, const T* x, T* out, T* d, T* z,
                     const T* precon, double* partials)
{
    const int grid = nslices > 0 ? nslices : 1;
    launch(k_residual<T>, dim3(grid), dim3(kSellBlock), s, (int)mode, A, nslices, row0, nrows, b, x, out, d, z,
                                              precon, partials);
    if (mode != RES_ONLY) launch(k_restart_finish, dim3(1), dim3(kWave), s, st, partials, grid);
}

template <typename T>
void launch_residual_norm(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                          SolveState* st, const T* b, const T* x, T* out, double* partials)
{
    const int grid = nslices > 0 ? nslices : 1;
    launch(k_residual_norm<T>, dim3(grid), dim3(kSellBlock), s, A, nslices, row0, nrows, b, x, out, partials);
    launch(k_result_finish, dim3(1), dim3(kWave), s, st, partials, grid);
}

template <typename T>
void launch_direction_init(hipStream_t s, SolveState* st, int n_active, const T* z, const T* r,
                           T* d, double* partials)
{
    const int grid = vec_reduce_blocks(n_active);
    launch(k_direction_init<T>, dim3(grid), dim3(kVecBlock), s, n_active, z, r, d, partials);
    launch(k_restart_finish, dim3(1), dim3(kWave), s, st, partials, grid);
}

template <typename T>
void launch_direction_continue(hipStream_t s, SolveState* st, int n_active, T* d,
                               const T* dprev, double zr_prev, int maxK)
{
    const int grid = vec_reduce_blocks(n_active);
    launch(k_direction_continue<T>, dim3(grid), dim3(kVecBlock), s, st, n_active, d, dprev, zr_prev,
           maxK);
}

template <typename T>
void launch_eisenstat_init(hipStream_t s, SolveState* st, int n_active, const T* r,
                           const T* diag, T omega, T* z, T* d, double* partials)
{
    const int grid = vec_reduce_blocks(n_active);
    launch(k_eisenstat_init<T>, dim3(grid), dim3(kVecBlock), s, n_active, r, diag, omega, z, d, partials);
    launch(k_restart_finish, dim3(1), dim3(kWave), s, st, partials, grid);
}

template <typename T>
void launch_forward(hipStream_t s, SellViewT<T> L, int slice0, int nslices, int row0, int nrows,
                    SolveState* st, int mode, int p, T* x, T* r, const T* d, const T* q, T* m,
                    const T* diag, T omega, const double* spmv_partials, int nspmv, T* zb, T c2,
                    double* zpart, int gbase, int gw, int m0_row0, int m0_nrows)
{
    // The first group of an iteration runs one block even when empty: it
    // forms alpha for the groups after it.
    const int extra = (m0_nrows + kWave - 1) / kWave;  // blocks of a merged first group
    int grid = nslices + extra;
    if (grid == 0 && mode == FWD_FIRST) grid = 1;
    if (grid == 0) return;
    launch(k_forward<T>, dim3(grid), dim3(kSellBlock), s, L, slice0, nslices, row0, nrows, st, mode, p, x, r, d,
                                             q, m, diag, omega, spmv_partials, nspmv, zb, c2,
                                             zpart, gbase, gw, m0_row0, m0_nrows);
}

template <typename T>
void launch_backward(hipStream_t s, SellViewT<T> U, int slice0, int nslices, int row0, int nrows,
                     SolveState* st, int reduce, T* r, const T* m, T* z, const T* diag,
                     T omega, T c2, double* group_partials, int prev0, int prevn,
                     double* prev_sum, int prev0b, int prevnb, double* prev_sumb, int gbase,
                     int gw, int defer_update, T* x, const T* d, const T* q)
{
    // one more block per earlier group to sum
    const int grid = nslices + (reduce && prev_sum ? 1 : 0) + (reduce && prev_sumb ? 1 : 0);
    if (grid == 0) return;
    launch(k_backward<T>, dim3(grid), dim3(kSellBlock), s, U, slice0, nslices, row0, nrows, st,
           reduce, r, m, z, diag, omega, c2, group_partials, prev0, prevn, prev_sum, prev0b,
           prevnb, prev_sumb, gbase, gw, defer_update, x, d, q);
}

template <typename T>
void launch_forward_pair(hipStream_t s, SellViewT<T> L, SweepGroup A, SweepGroup B,
                         SolveState* st, T* x, T* r, const T* d, const T* q, T* m,
                         const T* diag, T omega, T* zb, T c2, double* zpart)
{
    const int grid = (A.nslices + kPairWaves - 1) / kPairWaves + B.nslices;
    if (grid == 0) return;
    launch(k_forward_pair<T>, dim3(grid), dim3(kPairBlock), s, L, A, B, st, x, r, d, q, m, diag,
           omega, zb, c2, zpart);
}

template <typename T>
void launch_backward_pair(hipStream_t s, SellViewT<T> U, SweepGroup A, SweepGroup B,
                          SolveState* st, T* r, const T* m, T* z, const T* diag, T omega, T c2,
                          double* group_partials, int nfold, int fold0a, int foldna,
                          double* fold_suma, int fold0b, int foldnb, double* fold_sumb,
                          int defer_a, int defer_b, T* x, const T* d, const T* q)
{
    const int grid = (A.nslices + kPairWaves - 1) / kPairWaves + B.nslices + (nfold > 0 ? 1 : 0);
    if (grid == 0) return;
    launch(k_backward_pair<T>, dim3(grid), dim3(kPairBlock), s, U, A, B, st, r, m, z, diag, omega,
           c2, group_partials, nfold, fold0a, foldna, fold_suma, fold0b, foldnb, fold_sumb,
           defer_a, defer_b, x, d, q);
}

template <typename T>
void launch_eisenstat_solve(hipStream_t s, SellViewT<T> P, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int with_t, const T* in, const T* t,
                            T* out, const T* diag, T omega, T c3, int gbase, int gw)
{
    if (nslices <= 0) return;
    launch(k_eisenstat_solve<T>, dim3(nslices), dim3(kSellBlock), s, P, slice0, nslices, row0, nrows, st,
                                                        with_t, in, t, out, diag, omega, c3,
                                                        gbase, gw);
}

template <typename T>
void launch_eisenstat_upper(hipStream_t s, SellViewT<T> P, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int first, int p, const T* z,
                            const T* dold, T* dnew, T* out, const T* diag, T omega,
                            const double* fold_partials, int nfold, int gbase, int gw)
{
    const int grid = nslices > 0 ? nslices : (first ? 1 : 0);
    if (grid == 0) return;
    launch(k_eisenstat_upper<T>, dim3(grid), dim3(kSellBlock), s, P, slice0, nslices, row0, nrows, st, first, p,
                                                     z, dold, dnew, out, diag, omega,
                                                     fold_partials, nfold, gbase, gw);
}

template <typename T>
void launch_spmv_dir(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                     int n_active, SolveState* st, int p, const T* src, const T* dold, T* dnew,
                     T* q, double* spmv_partials, const double* fold_partials, int nfold,
                     const double* fold_partials_b, int nfold_b, const double* group_sums,
                     int ngroups, int gw, int lanes)
{
    const dim3 grid(spmv_blocks(nslices * lanes)), block(kSpmvBlock);
#define SFA_SPMV_DIR(KK)                                                                       \
    launch(k_spmv_dir<T, KK>, grid, block, s, A, nslices, row0, nrows, n_active, st, p, src,  \
           dold, dnew, q, spmv_partials, fold_partials, nfold, fold_partials_b, nfold_b,      \
           group_sums, ngroups, gw)
    switch (lanes) {
    case 1: SFA_SPMV_DIR(1); break;
    case 2: SFA_SPMV_DIR(2); break;
    case 4: SFA_SPMV_DIR(4); break;
    case 8: SFA_SPMV_DIR(8); break;
    default: break;
    }
#undef SFA_SPMV_DIR
}

template <typename T>
void launch_update(hipStream_t s, UpdateMode mode, SolveState* st, int p, int n_active, T* x,
                   T* r, const T* dx, const T* q, T* z, const T* diag_or_precon, T omega,
                   const double* dad_partials, int ndad, double* zr_partials)
{
    launch(k_update<T>, dim3(vec_reduce_blocks(n_active)), dim3(kVecBlock), s, 
        (int)mode, st, p, n_active, x, r, dx, q, z, diag_or_precon, omega, dad_partials, ndad,
        zr_partials);
}

template <typename T>
void launch_eisenstat_combine(hipStream_t s, SolveState* st, int n_active, T* q, const T* t,
                              const T* d, double* partials)
{
    launch(k_eisenstat_combine<T>, dim3(vec_reduce_blocks(n_active)), dim3(kVecBlock), s, st, n_active, q, t, d,
                                                                            partials);
}

void launch_to_float(hipStream_t s, float* dst, const double* src, int n)
{
    if (n > 0) launch(k_convert<float, double>, dim3(flat_blocks(n)), dim3(kVecBlock), s, dst, src, n);
}

void launch_add_correction(hipStream_t s, double* x, const float* e, int n)
{
    if (n > 0) launch(k_add_correction, dim3(flat_blocks(n)), dim3(kVecBlock), s, x, e, n);
}

#define SFA_INSTANTIATE(T)                                                                     \
    template void launch_fill<T>(hipStream_t, T*, T, int);                                     \
    template void launch_gather<T>(hipStream_t, T*, const T*, const int*, int);                \
    template void launch_scatter<T>(hipStream_t, T*, const T*, const int*, int);               \
    template void launch_init_state<T>(hipStream_t, SolveState*, const T*, int, double, double, \
                                       int, int, int, double, double*, const T*);              \
    template void launch_apply<T>(hipStream_t, SellViewT<T>, int, int, int, const T*, T*, T,   \
                                  int);                                                        \
    template void launch_residual<T>(hipStream_t, ResidualMode, SellViewT<T>, int, int, int,   \
                                     SolveState*, const T*, const T*, T*, T*, T*, const T*,    \
                                     double*);                                                 \
    template void launch_residual_norm<T>(hipStream_t, SellViewT<T>, int, int, int,            \
                                          SolveState*, const T*, const T*, T*, double*);       \
    template void launch_direction_continue<T>(hipStream_t, SolveState*, int, T*, const T*,   \
                                               double, int);                                   \
    template void launch_direction_init<T>(hipStream_t, SolveState*, int, const T*, const T*,  \
                                           T*, double*);                                       \
    template void launch_eisenstat_init<T>(hipStream_t, SolveState*, int, const T*, const T*,  \
                                           T, T*, T*, double*);                                \
    template void launch_forward<T>(hipStream_t, SellViewT<T>, int, int, int, int, SolveState*, \
                                    int, int, T*, T*, const T*, const T*, T*, const T*, T,     \
                                    const double*, int, T*, T, double*, int, int, int, int);   \
    template void launch_backward<T>(hipStream_t, SellViewT<T>, int, int, int, int,            \
                                     SolveState*, int, T*, const T*, T*, const T*, T, T,       \
                                     double*, int, int, double*, int, int, double*, int, int,  \
                                     int, T*, const T*, const T*);                             \
    template void launch_forward_pair<T>(hipStream_t, SellViewT<T>, SweepGroup, SweepGroup,    \
                                         SolveState*, T*, T*, const T*, const T*, T*,          \
                                         const T*, T, T*, T, double*);                         \
    template void launch_backward_pair<T>(hipStream_t, SellViewT<T>, SweepGroup, SweepGroup,   \
                                          SolveState*, T*, const T*, T*, const T*, T, T,       \
                                          double*, int, int, int, double*, int, int, double*,  \
                                          int, int, T*, const T*, const T*);                   \
    template void launch_eisenstat_solve<T>(hipStream_t, SellViewT<T>, int, int, int, int,     \
                                            SolveState*, int, const T*, const T*, T*,          \
                                            const T*, T, T, int, int);                         \
    template void launch_eisenstat_upper<T>(hipStream_t, SellViewT<T>, int, int, int, int,     \
                                            SolveState*, int, int, const T*, const T*, T*, T*, \
                                            const T*, T, const double*, int, int, int);        \
    template void launch_spmv_dir<T>(hipStream_t, SellViewT<T>, int, int, int, int,            \
                                     SolveState*, int, const T*, const T*, T*, T*, double*,    \
                                     const double*, int, const double*, int, const double*,    \
                                     int, int, int);                                           \
    template void launch_update<T>(hipStream_t, UpdateMode, SolveState*, int, int, T*, T*,     \
                                   const T*, const T*, T*, const T*, T, const double*, int,    \
                                   double*);                                                   \
    template void launch_eisenstat_combine<T>(hipStream_t, SolveState*, int, T*, const T*,     \
                                              const T*, double*);

SFA_INSTANTIATE(double)
SFA_INSTANTIATE(float)

}  // namespace sfa
