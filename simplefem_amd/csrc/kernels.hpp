// kernels.hpp -- launch interface of the gfx950 CG / SSOR kernels.
//
// All vectors live in device order (see plan.hpp) with n+1 entries; slot n
// holds 0 and is never written, so SELL padding entries (value 0, column n)
// add exactly nothing.  Scalars of the CG recurrence stay on the device in a
// SolveState; every kernel of an iteration first reads state->done and
// returns at once when the solve has finished, so a captured chunk of
// iterations can be replayed past convergence without changing the result.
#pragma once

#include <hip/hip_runtime.h>

namespace sfa {

template <typename T>
struct SellViewT {
    const int* off;      // nslices+1
    const int* rows;     // nslices*64, -1 = padding lane
    const int* cols;
    const T* vals;
};
using SellView = SellViewT<double>;

struct SolveState {
    double zr;          // z.r of the current iterate
    double dAd;
    double alpha;
    double beta;
    double tol;         // max(precision*||b||, sqrt(eps))
    double tol2;        // tol*tol
    double rel2;        // > 0: tol2 = rel2 * |z.r of the restart| (inner solves)
    double result;      // output of k_residual_norm
    int K;              // iterations done
    int maxK;           // N
    int done;           // 1 once the loop has ended
    int exit_reason;    // SFA_EXIT_*
    int small_breaks;   // 0: a small d.Ad does not end the loop (natural Eisenstat, quiet)
    unsigned ticket;    // last-block counter of the running reduction
    double* hist;       // z.r at the top of iteration k, k = 0..K
};

// Vector kernels
template <typename T> void launch_fill(hipStream_t s, T* v, T value, int n);
template <typename T> void launch_gather(hipStream_t s, T* dst, const T* src, const int* perm, int n);
template <typename T> void launch_scatter(hipStream_t s, T* dst, const T* src, const int* perm, int n);
template <typename T>
void launch_init_state(hipStream_t s, SolveState* st, const T* b, int n, double precision,
                       double host_tol, int use_host_tol, int maxK, int small_breaks, double rel2,
                       double* partials);

// MatrixCSR::apply: dest[row] = sum (scaling*v)*src[col], over a part's rows.
// (row0, nrows): a group whose rows are row0 .. row0+nrows-1 in slice order;
// row0 = -1 makes the kernels read row ids from the part's row list.
template <typename T>
void launch_apply(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows, const T* src,
                  T* dest, T scaling);

// --- restart (K = 0) ---
enum ResidualMode { RES_ONLY = 0, RES_PLAIN = 1, RES_JACOBI = 2 };
// r = b - A x on the part's rows; RES_PLAIN also sets d = r and z.r = r.r,
// RES_JACOBI z = p*r, d = z, z.r; both then test convergence.  With
// out != r the result goes to out (the Eisenstat restart writes q).
template <typename T>
void launch_residual(hipStream_t s, ResidualMode mode, SellViewT<T> A, int nslices, int row0,
                     int nrows, SolveState* st, const T* b, const T* x, T* out, T* d, T* z,
                     const T* precon, double* partials);
// out = b - A x, ||out||^2 into st->result (outer loop of the mixed-precision form)
template <typename T>
void launch_residual_norm(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                          SolveState* st, const T* b, const T* x, T* out, double* partials);
// d = z, z.r, convergence test (SSOR family restart).
template <typename T>
void launch_direction_init(hipStream_t s, SolveState* st, int n_active, const T* z, const T* r,
                           T* d, double* partials);
// Eisenstat restart: z = r*diag*(2-w)/w, d = z, z.r, convergence test.
template <typename T>
void launch_eisenstat_init(hipStream_t s, SolveState* st, int n_active, const T* r,
                           const T* diag, T omega, T* z, T* d, double* partials);

// --- sweeps over one group (colour or level) of a triangular part ---
// forward: [x += a d; r -= a q;]  m = (r - L m) * w / diag
template <typename T>
void launch_forward(hipStream_t s, SellViewT<T> L, int slice0, int nslices, int row0, int nrows,
                    SolveState* st, int update_xr, T* x, T* r, const T* d, const T* q, T* m,
                    const T* diag, T omega);
// backward: t = m*diag*c2; z = (t - U z) * w / diag; [group z.r partial]
template <typename T>
void launch_backward(hipStream_t s, SellViewT<T> U, int slice0, int nslices, int row0, int nrows,
                     SolveState* st, int reduce, int group, int last_group, const T* r,
                     const T* m, T* z, const T* diag, T omega, T c2, double* grp,
                     double* partials);
// Eisenstat: out = (in - [c3*diag*t] - P out) * w / diag  (P = L or U part)
template <typename T>
void launch_eisenstat_solve(hipStream_t s, SellViewT<T> P, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int with_t, const T* in, const T* t,
                            T* out, const T* diag, T omega, T c3);

// --- iteration ---
// q = A d, d.Ad, breakdown tests, alpha
template <typename T>
void launch_spmv_dot(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                     SolveState* st, const T* d, T* q, double* partials);
enum UpdateMode { UPD_PLAIN = 0, UPD_JACOBI = 1, UPD_EISENSTAT = 2 };
// x += a*dx; r -= a*q; z.r per mode; beta; K++; convergence test
template <typename T>
void launch_update(hipStream_t s, UpdateMode mode, SolveState* st, int n_active, T* x, T* r,
                   const T* dx, const T* q, T* z, const T* diag_or_precon, T omega,
                   double* partials);
// Eisenstat: q += t, d.Ad, breakdown tests, alpha
template <typename T>
void launch_eisenstat_combine(hipStream_t s, SolveState* st, int n_active, T* q, const T* t,
                              const T* d, double* partials);
// d = src + beta*d
template <typename T>
void launch_direction(hipStream_t s, SolveState* st, int n_active, const T* src, T* d);

// mixed precision helpers
void launch_to_float(hipStream_t s, float* dst, const double* src, int n);
void launch_add_correction(hipStream_t s, double* x, const float* e, int n);

}  // namespace sfa
