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
// Reductions (d.Ad, z.r) are deterministic: each block writes one partial,
// the last block to finish (agent-scope ticket) sums the partials in block
// order and runs the scalar part of the CG step on the device.
#include "kernels.hpp"

#include <cmath>

#include "../../include/simplefem_amd.h"
#include "device_common.hpp"

namespace sfa {

namespace {

using namespace dev;

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

// State reset and tolerance (solver/sparsesolver.cpp:469-472):
// tol = max(precision * sqrt(sum b^2), desired_precision).
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_init_state(SolveState* st, const T* b, int n,
                                                          double precision, double host_tol,
                                                          int use_host_tol, int maxK,
                                                          int small_breaks, double rel2,
                                                          double* partials)
{
    double s = 0.;
    if (!use_host_tol)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            s += (double)b[i] * (double)b[i];
    double total;
    if (!grid_reduce_last<kVecBlock>(s, partials, &st->ticket, &total)) return;
    if (threadIdx.x != 0) return;
    double tol = host_tol;
    if (!use_host_tol) {
        tol = precision * sqrt(total);
        // base/base.hpp:275-283 maximum()
        if (!(tol >= kDesiredPrecision) && (tol <= kDesiredPrecision)) tol = kDesiredPrecision;
    }
    st->tol = tol;
    st->tol2 = tol * tol;
    st->zr = __builtin_nan("");
    st->dAd = 0.;
    st->alpha = 0.;
    st->beta = 0.;
    st->K = 0;
    st->maxK = maxK;
    st->done = 0;
    st->exit_reason = SFA_EXIT_CAP;
    st->small_breaks = small_breaks;
    st->rel2 = rel2;
    st->result = 0.;
}

// ------------------------------------------------------------------ apply

// MatrixCSR::apply (sparse/matcsr.cpp:216-244): dest = 0; dest += scaling*V*x.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_apply(SellViewT<T> A, int nslices, int row0,
                                                      int nrows, const T* src, T* dest,
                                                      T scaling)
{
    const int slice = blockIdx.x;
    const int lane = threadIdx.x;
    if (slice >= nslices) return;
    const int row = group_row(A, slice, slice, lane, row0, nrows);
    if (row < 0) return;
    const int beg = A.off[slice] + lane, end = A.off[slice + 1];
    T acc = 0;
    for (int e = beg; e < end; e += kWave) acc += scaling * A.vals[e] * src[A.cols[e]];
    dest[row] = acc;
}

// ------------------------------------------------------------------ restart

// solver/sparsesolver.cpp:80-98 (plain), :281-306 (Jacobi), :512-524 (SSOR),
// :797-812 (Eisenstat, columns already filtered into the part).
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_residual(int mode, SellViewT<T> A, int nslices,
                                                         int row0, int nrows,
                                                         SolveState* st, const T* b,
                                                         const T* x, T* out, T* d,
                                                         T* z, const T* precon,
                                                         double* partials)
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
    if (mode == RES_ONLY) return;
    double total;
    if (!grid_reduce_last<kSellBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) after_zr_restart(st, total);
}

// solver/sparsesolver.cpp:570-583: d = z, z.r = sum z*r.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_direction_init(SolveState* st, int n, const T* z,
                                                              const T* r, T* d,
                                                              double* partials)
{
    double part = 0.;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T zz = z[i];
        d[i] = zz;
        part += (double)(zz * r[i]);
    }
    double total;
    if (!grid_reduce_last<kVecBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) after_zr_restart(st, total);
}

// solver/sparsesolver.cpp:830-844: z = r*diag*(2-w)/w, d = z, z.r.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_eisenstat_init(SolveState* st, int n, const T* r,
                                                              const T* diag, T omega,
                                                              T* z, T* d,
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
    double total;
    if (!grid_reduce_last<kVecBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) after_zr_restart(st, total);
}

// ------------------------------------------------------------------ sweeps

// Forward sweep of one colour / level, solver/sparsesolver.cpp:528-541 (restart)
// and :647-665, :1234-1261 (iteration, with the x and r updates).
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_forward(SellViewT<T> L, int slice0, int nslices,
                                                        int row0, int nrows,
                                                        const SolveState* st, int update_xr,
                                                        T* x, T* r, const T* d,
                                                        const T* q, T* m,
                                                        const T* diag, T omega)
{
    if (st->done) return;
    const int slice = slice0 + blockIdx.x;
    const int lane = threadIdx.x;
    if (blockIdx.x >= nslices) return;
    const int row = group_row(L, slice, blockIdx.x, lane, row0, nrows);
    if (row < 0) return;
    T rr;
    if (update_xr) {
        const T alpha = (T)st->alpha;
        x[row] += alpha * d[row];
        rr = r[row] - alpha * q[row];
        r[row] = rr;
    } else {
        rr = r[row];
    }
    const T acc = sell_row<true>(L, slice, lane, m, rr);
    m[row] = acc * omega / diag[row];
}

// Backward sweep of one colour / level, solver/sparsesolver.cpp:543-568
// (restart) and :667-692, :1263-1303 (iteration): the scaling of m fused in,
// z.r summed per group and over groups in launch order.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_backward(SellViewT<T> U, int slice0, int nslices,
                                                         int row0, int nrows,
                                                         SolveState* st, int reduce, int group,
                                                         int last_group, const T* r,
                                                         const T* m, T* z,
                                                         const T* diag, T omega,
                                                         T c2, double* grp, double* partials)
{
    if (st->done) return;
    const int slice = slice0 + blockIdx.x;
    const int lane = threadIdx.x;
    double part = 0.;
    const int row = blockIdx.x < nslices ? group_row(U, slice, blockIdx.x, lane, row0, nrows) : -1;
    if (row >= 0) {
        const T dg = diag[row];
        T t = m[row];
        t *= dg;
        t *= c2;
        const T acc = sell_row<true>(U, slice, lane, z, t);
        const T zz = acc * omega / dg;
        z[row] = zz;
        part = (double)(zz * r[row]);
    }
    if (!reduce) return;
    double total;
    if (!grid_reduce_last<kSellBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x != 0) return;
    grp[group] = total;
    if (!last_group) return;
    double zr_new = 0.;
    for (int g = 0; g <= group; g++) zr_new += grp[g];
    after_zr_new(st, zr_new);
}

// Eisenstat triangular solves (solver/sparsesolver.cpp:814-828, 876-908,
// 1429-1455, 1501-1560): out = (in - [c3*diag*t] - T out) * w / diag.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_eisenstat_solve(SellViewT<T> P, int slice0,
                                                                int nslices, int row0, int nrows,
                                                                const SolveState* st, int with_t,
                                                                const T* in, const T* t,
                                                                T* out, const T* diag,
                                                                T omega, T c3)
{
    if (st->done) return;
    const int slice = slice0 + blockIdx.x;
    const int lane = threadIdx.x;
    if (blockIdx.x >= nslices) return;
    const int row = group_row(P, slice, blockIdx.x, lane, row0, nrows);
    if (row < 0) return;
    const T dg = diag[row];
    T a = in[row];
    if (with_t) a = a - c3 * dg * t[row];
    const T acc = sell_row<true>(P, slice, lane, out, a);
    out[row] = acc * omega / dg;
}

// ------------------------------------------------------------------ iteration

// q = A d, d.Ad (solver/sparsesolver.cpp:611-626, 1197-1212), then the
// breakdown tests and alpha.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_spmv_dot(SellViewT<T> A, int nslices, int row0,
                                                         int nrows, SolveState* st,
                                                         const T* d, T* q,
                                                         double* partials)
{
    if (st->done) return;
    const int slice = blockIdx.x;
    const int lane = threadIdx.x;
    double part = 0.;
    const int row = slice < nslices ? group_row(A, slice, slice, lane, row0, nrows) : -1;
    if (row >= 0) {
        const T acc = sell_row<false>(A, slice, lane, d, T(0));
        q[row] = acc;
        part = (double)(d[row] * acc);
    }
    double total;
    if (!grid_reduce_last<kSellBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) after_dAd(st, total);
}

// x += a dx; r -= a q; then z.r by mode (solver/sparsesolver.cpp:160-174,
// 372-389, 952-968); beta and the loop-head tests.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_update(int mode, SolveState* st, int n, T* x,
                                                      T* r, const T* dx,
                                                      const T* q, T* z,
                                                      const T* dp, T omega,
                                                      double* partials)
{
    if (st->done) return;
    const T alpha = (T)st->alpha;
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
    double total;
    if (!grid_reduce_last<kVecBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) after_zr_new(st, total);
}

// solver/sparsesolver.cpp:912-922: q += t, d.Ad = sum q*d; breakdown tests.
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_eisenstat_combine(SolveState* st, int n, T* q,
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
    double total;
    if (!grid_reduce_last<kVecBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) after_dAd(st, total);
}

// d = src + beta*d (solver/sparsesolver.cpp:698-706 and its siblings).
template <typename T>
__global__ void __launch_bounds__(kVecBlock) k_direction(const SolveState* st, int n,
                                                         const T* src, T* d)
{
    if (st->done) return;
    const T beta = (T)st->beta;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = src[i] + beta * d[i];
}

// ------------------------------------------------------------------ mixed precision

// out = b - A x in the part's rows, sum of squares into st->result.
template <typename T>
__global__ void __launch_bounds__(kSellBlock) k_residual_norm(SellViewT<T> A, int nslices, int row0,
                                                              int nrows, SolveState* st,
                                                              const T* b, const T* x, T* out,
                                                              double* partials)
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
    double total;
    if (!grid_reduce_last<kSellBlock>(part, partials, &st->ticket, &total)) return;
    if (threadIdx.x == 0) st->result = total;
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

int vec_blocks(int n)
{
    int b = (n + kVecBlock - 1) / kVecBlock;
    if (b < 1) b = 1;
    return b > kVecMaxBlocks ? kVecMaxBlocks : b;
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
    if (n > 0) k_fill<T><<<flat_blocks(n), kVecBlock, 0, s>>>(v, value, n);
}

template <typename T>
void launch_gather(hipStream_t s, T* dst, const T* src, const int* perm, int n)
{
    if (n > 0) k_gather<T><<<flat_blocks(n), kVecBlock, 0, s>>>(dst, src, perm, n);
}

template <typename T>
void launch_scatter(hipStream_t s, T* dst, const T* src, const int* perm, int n)
{
    if (n > 0) k_scatter<T><<<flat_blocks(n), kVecBlock, 0, s>>>(dst, src, perm, n);
}

template <typename T>
void launch_init_state(hipStream_t s, SolveState* st, const T* b, int n, double precision,
                       double host_tol, int use_host_tol, int maxK, int small_breaks, double rel2,
                       double* partials)
{
    k_init_state<T><<<use_host_tol ? 1 : vec_blocks(n), kVecBlock, 0, s>>>(
        st, b, n, precision, host_tol, use_host_tol, maxK, small_breaks, rel2, partials);
}

template <typename T>
void launch_apply(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows, const T* src,
                  T* dest, T scaling)
{
    if (nslices > 0)
        k_apply<T><<<nslices, kSellBlock, 0, s>>>(A, nslices, row0, nrows, src, dest, scaling);
}

template <typename T>
void launch_residual(hipStream_t s, ResidualMode mode, SellViewT<T> A, int nslices, int row0,
                     int nrows, SolveState* st, const T* b, const T* x, T* out, T* d, T* z,
                     const T* precon, double* partials)
{
    const int grid = nslices > 0 ? nslices : 1;
    k_residual<T><<<grid, kSellBlock, 0, s>>>((int)mode, A, nslices, row0, nrows, st, b, x, out,
                                              d, z, precon, partials);
}

template <typename T>
void launch_residual_norm(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                          SolveState* st, const T* b, const T* x, T* out, double* partials)
{
    const int grid = nslices > 0 ? nslices : 1;
    k_residual_norm<T><<<grid, kSellBlock, 0, s>>>(A, nslices, row0, nrows, st, b, x, out,
                                                   partials);
}

template <typename T>
void launch_direction_init(hipStream_t s, SolveState* st, int n_active, const T* z, const T* r,
                           T* d, double* partials)
{
    k_direction_init<T><<<vec_blocks(n_active), kVecBlock, 0, s>>>(st, n_active, z, r, d,
                                                                   partials);
}

template <typename T>
void launch_eisenstat_init(hipStream_t s, SolveState* st, int n_active, const T* r,
                           const T* diag, T omega, T* z, T* d, double* partials)
{
    k_eisenstat_init<T><<<vec_blocks(n_active), kVecBlock, 0, s>>>(st, n_active, r, diag, omega,
                                                                   z, d, partials);
}

template <typename T>
void launch_forward(hipStream_t s, SellViewT<T> L, int slice0, int nslices, int row0, int nrows,
                    SolveState* st, int update_xr, T* x, T* r, const T* d, const T* q, T* m,
                    const T* diag, T omega)
{
    if (nslices <= 0) return;
    k_forward<T><<<nslices, kSellBlock, 0, s>>>(L, slice0, nslices, row0, nrows, st, update_xr, x,
                                                r, d, q, m, diag, omega);
}

template <typename T>
void launch_backward(hipStream_t s, SellViewT<T> U, int slice0, int nslices, int row0, int nrows,
                     SolveState* st, int reduce, int group, int last_group, const T* r,
                     const T* m, T* z, const T* diag, T omega, T c2, double* grp,
                     double* partials)
{
    // A group with no rows still runs one block when it has to publish its
    // (zero) partial or finish the z.r sum.
    const int grid = nslices > 0 ? nslices : (reduce ? 1 : 0);
    if (grid == 0) return;
    k_backward<T><<<grid, kSellBlock, 0, s>>>(U, slice0, nslices, row0, nrows, st, reduce, group,
                                              last_group, r, m, z, diag, omega, c2, grp,
                                              partials);
}

template <typename T>
void launch_eisenstat_solve(hipStream_t s, SellViewT<T> P, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int with_t, const T* in, const T* t,
                            T* out, const T* diag, T omega, T c3)
{
    if (nslices <= 0) return;
    k_eisenstat_solve<T><<<nslices, kSellBlock, 0, s>>>(P, slice0, nslices, row0, nrows, st,
                                                        with_t, in, t, out, diag, omega, c3);
}

template <typename T>
void launch_spmv_dot(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                     SolveState* st, const T* d, T* q, double* partials)
{
    const int grid = nslices > 0 ? nslices : 1;
    k_spmv_dot<T><<<grid, kSellBlock, 0, s>>>(A, nslices, row0, nrows, st, d, q, partials);
}

template <typename T>
void launch_update(hipStream_t s, UpdateMode mode, SolveState* st, int n_active, T* x, T* r,
                   const T* dx, const T* q, T* z, const T* diag_or_precon, T omega,
                   double* partials)
{
    k_update<T><<<vec_blocks(n_active), kVecBlock, 0, s>>>((int)mode, st, n_active, x, r, dx, q,
                                                           z, diag_or_precon, omega, partials);
}

template <typename T>
void launch_eisenstat_combine(hipStream_t s, SolveState* st, int n_active, T* q, const T* t,
                              const T* d, double* partials)
{
    k_eisenstat_combine<T><<<vec_blocks(n_active), kVecBlock, 0, s>>>(st, n_active, q, t, d,
                                                                      partials);
}

template <typename T>
void launch_direction(hipStream_t s, SolveState* st, int n_active, const T* src, T* d)
{
    if (n_active > 0)
        k_direction<T><<<flat_blocks(n_active), kVecBlock, 0, s>>>(st, n_active, src, d);
}

void launch_to_float(hipStream_t s, float* dst, const double* src, int n)
{
    if (n > 0) k_convert<float, double><<<flat_blocks(n), kVecBlock, 0, s>>>(dst, src, n);
}

void launch_add_correction(hipStream_t s, double* x, const float* e, int n)
{
    if (n > 0) k_add_correction<<<flat_blocks(n), kVecBlock, 0, s>>>(x, e, n);
}

#define SFA_INSTANTIATE(T)                                                                     \
    template void launch_fill<T>(hipStream_t, T*, T, int);                                     \
    template void launch_gather<T>(hipStream_t, T*, const T*, const int*, int);                \
    template void launch_scatter<T>(hipStream_t, T*, const T*, const int*, int);               \
    template void launch_init_state<T>(hipStream_t, SolveState*, const T*, int, double, double, \
                                       int, int, int, double, double*);                        \
    template void launch_apply<T>(hipStream_t, SellViewT<T>, int, int, int, const T*, T*, T);  \
    template void launch_residual<T>(hipStream_t, ResidualMode, SellViewT<T>, int, int, int,   \
                                     SolveState*, const T*, const T*, T*, T*, T*, const T*,    \
                                     double*);                                                 \
    template void launch_residual_norm<T>(hipStream_t, SellViewT<T>, int, int, int,            \
                                          SolveState*, const T*, const T*, T*, double*);       \
    template void launch_direction_init<T>(hipStream_t, SolveState*, int, const T*, const T*,  \
                                           T*, double*);                                       \
    template void launch_eisenstat_init<T>(hipStream_t, SolveState*, int, const T*, const T*,  \
                                           T, T*, T*, double*);                                \
    template void launch_forward<T>(hipStream_t, SellViewT<T>, int, int, int, int, SolveState*, \
                                    int, T*, T*, const T*, const T*, T*, const T*, T);         \
    template void launch_backward<T>(hipStream_t, SellViewT<T>, int, int, int, int,            \
                                     SolveState*, int, int, int, const T*, const T*, T*,       \
                                     const T*, T, T, double*, double*);                        \
    template void launch_eisenstat_solve<T>(hipStream_t, SellViewT<T>, int, int, int, int,     \
                                            SolveState*, int, const T*, const T*, T*,          \
                                            const T*, T, T);                                   \
    template void launch_spmv_dot<T>(hipStream_t, SellViewT<T>, int, int, int, SolveState*,    \
                                     const T*, T*, double*);                                   \
    template void launch_update<T>(hipStream_t, UpdateMode, SolveState*, int, T*, T*,          \
                                   const T*, const T*, T*, const T*, T, double*);              \
    template void launch_eisenstat_combine<T>(hipStream_t, SolveState*, int, T*, const T*,     \
                                              const T*, double*);                              \
    template void launch_direction<T>(hipStream_t, SolveState*, int, const T*, T*);

SFA_INSTANTIATE(double)
SFA_INSTANTIATE(float)

}  // namespace sfa
