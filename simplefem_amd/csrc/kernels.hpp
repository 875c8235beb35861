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

struct SellView {
    const int* off;      // nslices+1
    const int* rows;     // nslices*64, -1 = padding lane
    const int* cols;
    const double* vals;
};

struct SolveState {
    double zr;          // z.r of the current iterate
    double dAd;
    double alpha;
    double beta;
    double tol;         // max(precision*||b||, sqrt(eps))
    double tol2;        // tol*tol
    int K;              // iterations done
    int maxK;           // N
    int done;           // 1 once the loop has ended
    int exit_reason;    // SFA_EXIT_*
    int small_breaks;   // 0: a small d.Ad does not end the loop (natural Eisenstat, quiet)
    unsigned ticket;    // last-block counter of the running reduction
    double* hist;       // z.r at the top of iteration k, k = 0..K
};

// Vector kernels
void launch_fill(hipStream_t s, double* v, double value, int n);
void launch_gather(hipStream_t s, double* dst, const double* src, const int* perm, int n);
void launch_scatter(hipStream_t s, double* dst, const double* src, const int* perm, int n);
void launch_init_state(hipStream_t s, SolveState* st, const double* b, int n, double precision,
                       double host_tol, int use_host_tol, int maxK, int small_breaks,
                       double* partials);

// MatrixCSR::apply: dest[row] = sum (scaling*v)*src[col], over a part's rows.
// (row0, nrows): a group whose rows are row0 .. row0+nrows-1 in slice order;
// row0 = -1 makes the kernels read row ids from the part's row list.
void launch_apply(hipStream_t s, SellView A, int nslices, int row0, int nrows, const double* src,
                  double* dest, double scaling);

// --- restart (K = 0) ---
enum ResidualMode { RES_ONLY = 0, RES_PLAIN = 1, RES_JACOBI = 2 };
// r = b - A x on the part's rows; RES_PLAIN also sets d = r and z.r = r.r,
// RES_JACOBI z = p*r, d = z, z.r; both then test convergence.  With
// out != r the result goes to out (the Eisenstat restart writes q).
void launch_residual(hipStream_t s, ResidualMode mode, SellView A, int nslices, int row0,
                     int nrows, SolveState* st, const double* b, const double* x, double* out,
                     double* d, double* z, const double* precon, double* partials);
// d = z, z.r, convergence test (SSOR family restart).
void launch_direction_init(hipStream_t s, SolveState* st, int n_active, const double* z,
                           const double* r, double* d, double* partials);
// Eisenstat restart: z = r*diag*(2-w)/w, d = z, z.r, convergence test.
void launch_eisenstat_init(hipStream_t s, SolveState* st, int n_active, const double* r,
                           const double* diag, double omega, double* z, double* d,
                           double* partials);

// --- sweeps over one group (colour or level) of a triangular part ---
// forward: [x += a d; r -= a q;]  m = (r - L m) * w / diag
void launch_forward(hipStream_t s, SellView L, int slice0, int nslices, int row0, int nrows,
                    SolveState* st, int update_xr, double* x, double* r, const double* d,
                    const double* q, double* m, const double* diag, double omega);
// backward: t = m*diag*c2; z = (t - U z) * w / diag; [group z.r partial]
void launch_backward(hipStream_t s, SellView U, int slice0, int nslices, int row0, int nrows,
                     SolveState* st, int reduce, int group, int last_group, const double* r,
                     const double* m, double* z, const double* diag, double omega, double c2,
                     double* grp, double* partials);
// Eisenstat: out = (in - [c3*diag*t] - T out) * w / diag  (T = L or U part)
void launch_eisenstat_solve(hipStream_t s, SellView T, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int with_t, const double* in,
                            const double* t, double* out, const double* diag, double omega,
                            double c3);

// --- iteration ---
// q = A d, d.Ad, breakdown tests, alpha
void launch_spmv_dot(hipStream_t s, SellView A, int nslices, int row0, int nrows, SolveState* st,
                     const double* d, double* q, double* partials);
enum UpdateMode { UPD_PLAIN = 0, UPD_JACOBI = 1, UPD_EISENSTAT = 2 };
// x += a*dx; r -= a*q; z.r per mode; beta; K++; convergence test
void launch_update(hipStream_t s, UpdateMode mode, SolveState* st, int n_active, double* x,
                   double* r, const double* dx, const double* q, double* z,
                   const double* diag_or_precon, double omega, double* partials);
// Eisenstat: q += t, d.Ad, breakdown tests, alpha
void launch_eisenstat_combine(hipStream_t s, SolveState* st, int n_active, double* q,
                              const double* t, const double* d, double* partials);
// d = src + beta*d
void launch_direction(hipStream_t s, SolveState* st, int n_active, const double* src, double* d);

}  // namespace sfa
