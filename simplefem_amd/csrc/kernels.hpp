// kernels.hpp -- launch interface of the gfx950 CG / SSOR kernels.
//
// All vectors live in device order (see plan.hpp) with n+1 entries; slot n
// holds 0 and is never written, so SELL padding entries (value 0, column n)
// add exactly nothing.  Scalars of the CG recurrence stay on the device in a
// SolveState; every kernel of an iteration first reads state->done and
// returns at once when the solve has finished, so a captured chunk of
// iterations can be replayed past convergence without changing the result.
//
// Reductions are deferred across a kernel boundary: the producing kernel
// stores one partial sum per block (plain stores), and the kernel that runs
// next on the stream -- which sees them because kernels on one stream are
// ordered -- sums them in a fixed order in every block that needs the total
// and runs the scalar part of the CG step there.  No block waits for another
// and no grid-wide atomics or cache-maintenance fences are needed.  Values
// that a kernel both reads and updates (iteration count, z.r, the direction
// d) are double-buffered by the parity p of the iteration: the iteration of
// parity p reads slot p and writes slot 1-p.
#pragma once

#include <hip/hip_runtime.h>

namespace sfa {

template <typename T>
struct SellViewT {
    const int* off;      // nslices+1
    const int* rows;     // nslices*64, -1 = padding lane
    const int* cols;
    const T* vals;
    int zero;            // the vectors' zero slot (n): padding lanes read there
};
using SellView = SellViewT<double>;

struct SolveState {
    // fixed during the iteration loop
    double tol;         // max(precision*||b||, sqrt(eps))
    double tol2;        // tol*tol
    double rel2;        // > 0: tol2 = rel2 * |z.r of the restart| (inner solves)
    int maxK;           // N
    int small_breaks;   // 0: a small d.Ad does not end the loop (natural Eisenstat, quiet)
    // double-buffered by iteration parity
    int Kp[2];          // iterations done when the iteration of parity p starts
    int fold[2];        // 1: that iteration first folds the z.r of the previous one
    double zrp[2];      // z.r when that iteration's d.Ad step runs
    // written by one thread of the kernel that forms them, read by later kernels
    double alpha;
    double beta;
    // host-facing record
    double zr;          // z.r of the current iterate
    double dAd;
    double result;      // output of the residual norm
    int K;              // iterations done
    int done;           // 1 once the loop has ended
    int exit_reason;    // SFA_EXIT_*
    int pad_;
    double* hist;       // z.r at the top of iteration k, k = 0..K
};

// Block sizes and grid limits of the reducing kernels (their partial counts).
constexpr int kSpmvWaves = 4;          // slices per block of the SpMV kernel
constexpr int kSpmvMaxBlocks = 1024;
constexpr int kVecReduceBlocks = 512;  // vector kernels with a reduction

int spmv_blocks(int nslices);
int vec_reduce_blocks(int n);

// Vector kernels
template <typename T> void launch_fill(hipStream_t s, T* v, T value, int n);
template <typename T> void launch_gather(hipStream_t s, T* dst, const T* src, const int* perm, int n);
template <typename T> void launch_scatter(hipStream_t s, T* dst, const T* src, const int* perm, int n);
// Resets the state for a solve: tol = max(precision*||b||, 2^-26) (or host_tol),
// Kp = {0, 0}, fold = {0, 0}.  ||b|| from block partials over b, or, when
// rows_b is given, summed in row order over rows_b (b in its caller's order).
template <typename T>
void launch_init_state(hipStream_t s, SolveState* st, const T* b, int n, double precision,
                       double host_tol, int use_host_tol, int maxK, int small_breaks, double rel2,
                       double* partials, const T* rows_b = nullptr);

// MatrixCSR::apply: dest[row] = sum (scaling*v)*src[col], over a part's rows.
// (row0, nrows): a group whose rows are row0 .. row0+nrows-1 in slice order;
// row0 = -1 makes the kernels read row ids from the part's row list.
// lanes: lanes per row (1, 2, 4, 8); every value gives the same bits.
template <typename T>
void launch_apply(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows, const T* src,
                  T* dest, T scaling, int lanes = 1);

// --- restart (K = 0) ---
enum ResidualMode { RES_ONLY = 0, RES_PLAIN = 1, RES_JACOBI = 2 };
// r = b - A x on the part's rows; RES_PLAIN also sets d = r and z.r = r.r,
// RES_JACOBI z = p*r, d = z, z.r; both then run the restart's scalar step
// (z.r, loop-head test).  With out != r the result goes to out.
template <typename T>
void launch_residual(hipStream_t s, ResidualMode mode, SellViewT<T> A, int nslices, int row0,
                     int nrows, SolveState* st, const T* b, const T* x, T* out, T* d, T* z,
                     const T* precon, double* partials);
// out = b - A x, ||out||^2 into st->result (outer loop of the mixed-precision form)
template <typename T>
void launch_residual_norm(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                          SolveState* st, const T* b, const T* x, T* out, double* partials);
// d = z, z.r, restart scalar step (SSOR family).
template <typename T>
void launch_direction_init(hipStream_t s, SolveState* st, int n_active, const T* z, const T* r,
                           T* d, double* partials);
// After a restart: d = z + (st->zr / zr_prev) dprev, the direction of an
// earlier CG continued across a replaced residual; the loop's cap becomes maxK.
template <typename T>
void launch_direction_continue(hipStream_t s, SolveState* st, int n_active, T* d,
                               const T* dprev, double zr_prev, int maxK);
// Eisenstat restart: z = r*diag*(2-w)/w, d = z, z.r, restart scalar step.
template <typename T>
void launch_eisenstat_init(hipStream_t s, SolveState* st, int n_active, const T* r,
                           const T* diag, T omega, T* z, T* d, double* partials);

// --- sweeps over one group (colour or level) of a triangular part ---
// (gbase, gw): the group's first entry and, when all its slices have the same
// number of entries per lane, that number (the kernels then compute slice
// offsets instead of loading them); gw = -1 otherwise.
// mode FWD_RESTART: m = (r - L m) * w / diag.
// mode FWD_UPDATE:  x += a d; r -= a q; then the same, a = st->alpha.
// mode FWD_FIRST:   as FWD_UPDATE, but first forms d.Ad from the SpMV's
//                   partials (spmv_partials, nspmv), runs the breakdown tests
//                   and computes a (first group of the iteration, parity p).
// zb != null: the group is also the backward sweep's first group and has no
// upper entries; its z = (m*diag*c2)*w/diag goes to zb, with z.r partials at
// zpart[block] when zpart != null.  m0_nrows > 0: the kernel also runs the
// sweep's first group (rows m0_row0.., no lower entries), writing only its m;
// that group's x and r updates are left to its backward kernel (defer_update).
enum ForwardMode { FWD_RESTART = 0, FWD_UPDATE = 1, FWD_FIRST = 2 };
template <typename T>
void launch_forward(hipStream_t s, SellViewT<T> L, int slice0, int nslices, int row0, int nrows,
                    SolveState* st, int mode, int p, T* x, T* r, const T* d, const T* q, T* m,
                    const T* diag, T omega, const double* spmv_partials, int nspmv, T* zb, T c2,
                    double* zpart, int gbase, int gw, int m0_row0, int m0_nrows);
// backward: t = m*diag*c2; z = (t - U z) * w / diag.  With reduce, each
// block stores its z.r partial at group_partials[slice0 + block], and extra
// blocks sum earlier groups' partials (prev0, prevn) into *prev_sum and, when
// prev_sumb != null, (prev0b, prevnb) into *prev_sumb.
template <typename T>
void launch_backward(hipStream_t s, SellViewT<T> U, int slice0, int nslices, int row0, int nrows,
                     SolveState* st, int reduce, T* r, const T* m, T* z, const T* diag,
                     T omega, T c2, double* group_partials, int prev0, int prevn,
                     double* prev_sum, int prev0b, int prevnb, double* prev_sumb, int gbase,
                     int gw, int defer_update, T* x, const T* d, const T* q);
// --- two consecutive groups of a sweep in one launch (Rainbow iteration) ---
// A group as the pair kernels see it: slices [slice0, slice0+nslices) of the
// part, rows row0.. (consecutive; row0 < 0: read from the part's row list),
// and (gbase, gw) as above.
struct SweepGroup {
    int slice0, nslices, row0, nrows, gbase, gw;
};
constexpr int kPairWaves = 4;  // wavefronts per block of the pair kernels
constexpr int kPairRefs = 8;   // most entries of a B row whose column is an A row
constexpr int kPairWidth = 16; // most entries per lane of a B slice
// Forward sweep of groups A then B (B's rows read A's m): A's blocks run A as
// the one-group kernel does (FWD_UPDATE) but leave A's x and r updates to
// A's backward kernel; B's blocks form, for each A row a B row reads, that
// row's m with the same arithmetic as A's wavefront (from r - a q and the m of
// earlier groups), then run B (x, r updates; with zb the fused first
// backward group as in launch_forward).  A's rows must be consecutive; every
// B row may read at most kPairRefs A rows and B's slices at most kPairWidth
// entries per lane.
template <typename T>
void launch_forward_pair(hipStream_t s, SellViewT<T> L, SweepGroup A, SweepGroup B,
                         SolveState* st, T* x, T* r, const T* d, const T* q, T* m,
                         const T* diag, T omega, T* zb, T c2, double* zpart);
// Backward sweep of groups A then B (B's rows read A's z), with reduce: the
// z.r partial of every slice at group_partials[slice]; B's blocks form the z
// of the A rows they read as A's wavefront does.  defer_a / defer_b: that
// group's x and r updates (x += a d, r -= a q) are made here, first.  nfold
// (0..2) extra blocks sum earlier groups' partials: fold k sums
// group_partials[fold0[k] .. fold0[k]+foldn[k]) into *fold_sum[k].
template <typename T>
void launch_backward_pair(hipStream_t s, SellViewT<T> U, SweepGroup A, SweepGroup B,
                          SolveState* st, T* r, const T* m, T* z, const T* diag, T omega, T c2,
                          double* group_partials, int nfold, int fold0a, int foldna,
                          double* fold_suma, int fold0b, int foldnb, double* fold_sumb,
                          int defer_a, int defer_b, T* x, const T* d, const T* q);

// Eisenstat: out = (in - [c3*diag*t] - P out) * w / diag  (P = L or U part)
template <typename T>
void launch_eisenstat_solve(hipStream_t s, SellViewT<T> P, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int with_t, const T* in, const T* t,
                            T* out, const T* diag, T omega, T c3, int gbase, int gw);
// Eisenstat upper solve with the direction update of the previous
// iteration fused in: dn = fold ? z + beta*dold : dold on the group's rows,
// dnew = dn, out = (dn - U out) * w / diag.  first: this group opens the
// iteration and forms z.r, beta, K from the update's partials (fold_partials,
// nfold); later groups read beta from the state.
template <typename T>
void launch_eisenstat_upper(hipStream_t s, SellViewT<T> P, int slice0, int nslices, int row0,
                            int nrows, SolveState* st, int first, int p, const T* z,
                            const T* dold, T* dnew, T* out, const T* diag, T omega,
                            const double* fold_partials, int nfold, int gbase, int gw);

// --- iteration ---
// Opens an iteration of parity p for the SpMV-based variants: when
// state.fold[p], forms the new z.r from the previous iteration's partials
// (sum of group_sums[0..ngroups), then of fold_partials[0..nfold), then, when
// fold_partials_b != null, of fold_partials_b[0..nfold_b)), beta, K
// and the loop-head tests; then dn = fold ? src + beta*dold : dold on the
// active rows (dold elsewhere), dnew = dn, q = A dn and the d.Ad partials
// (spmv_partials, one per block: spmv_blocks(nslices * lanes) of them).
// lanes (1, 2, 4, 8): lanes per row of the SpMV mapping.
template <typename T>
void launch_spmv_dir(hipStream_t s, SellViewT<T> A, int nslices, int row0, int nrows,
                     int n_active, SolveState* st, int p, const T* src, const T* dold, T* dnew,
                     T* q, double* spmv_partials, const double* fold_partials, int nfold,
                     const double* fold_partials_b, int nfold_b, const double* group_sums,
                     int ngroups, int gw, int lanes);
enum UpdateMode { UPD_PLAIN = 0, UPD_JACOBI = 1, UPD_EISENSTAT = 2 };
// Forms d.Ad from dad_partials (ndad), breakdown tests, alpha; then
// x += a*dx; r -= a*q; z.r per mode into zr_partials (one per block).
template <typename T>
void launch_update(hipStream_t s, UpdateMode mode, SolveState* st, int p, int n_active, T* x,
                   T* r, const T* dx, const T* q, T* z, const T* diag_or_precon, T omega,
                   const double* dad_partials, int ndad, double* zr_partials);
// Eisenstat: q += t, d.Ad partials (one per block).
template <typename T>
void launch_eisenstat_combine(hipStream_t s, SolveState* st, int n_active, T* q, const T* t,
                              const T* d, double* partials);

// mixed precision helpers
void launch_to_float(hipStream_t s, float* dst, const double* src, int n);
void launch_add_correction(hipStream_t s, double* x, const float* e, int n);

}  // namespace sfa
