/*
 * simplefem_amd.h -- C-ABI of the MI355X (gfx950) sparse CG / SSOR solve path.
 *
 * The entry points replace FEEC++'s raw-pointer CSR solvers and the CSR
 * operations they rely on.  Each declaration names the reference interface it
 * stands in for (paths relative to the martinlicht/simplefem repository).
 *
 * Conventions
 *   - Plain C types only: int, double, pointers and sizes.
 *   - Every function returns an int status: SFA_OK (0) or a negative SFA_ERR_*;
 *     sfa_last_error() gives the message of the calling thread's last failure.
 *   - The reference returns the iteration count K from its solvers; here K and
 *     the convergence data come back in an sfa_solve_info.
 *   - Matrices are square CSR with int32 row offsets (N+1), int32 column
 *     indices and double values, exactly the arrays MatrixCSR::getA/getC/getV
 *     hand to the reference solvers (sparse/matcsr.hpp:89-93).
 *   - "Host" entry points take host arrays, copy them to the GPU, solve there
 *     and copy x and residual back: the semantics of the reference call.
 *     The sfa_system_* handle API keeps the matrix and preconditioner
 *     resident in device memory and can take device pointers.
 */
#ifndef SIMPLEFEM_AMD_H
#define SIMPLEFEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SFA_API_VERSION 2

/* status codes */
#define SFA_OK                 0
#define SFA_ERR_INVALID       -1  /* bad argument (sizes, pointers, CSR)       */
#define SFA_ERR_HIP           -2  /* a HIP runtime call failed                 */
#define SFA_ERR_NO_DEVICE     -3  /* no usable gfx950 device                   */
#define SFA_ERR_OUT_OF_MEMORY -4  /* device or host allocation failed          */
#define SFA_ERR_COLOURING     -5  /* F/B/R inconsistent, or a colour holds an  */
                                  /* edge (the sweep would race)               */

/* exit reasons of the CG loop; the reference logs these
 * (solver/sparsesolver.cpp:601-604,629-642,713-714) */
#define SFA_EXIT_CONVERGED 0  /* |z.r| < tolerance^2                          */
#define SFA_EXIT_CAP       1  /* K reached N                                  */
#define SFA_EXIT_BREAKDOWN 2  /* z.r or d.Ad non-finite or negative           */
#define SFA_EXIT_SMALL     3  /* sqrt|d.Ad| < machine epsilon                 */

/* OR into the method of sfa_system_create: also keep an FP32 copy of the
 * matrix and schedules for sfa_system_solve_mixed */
#define SFA_FLAG_FP32_INNER 0x100
/* Or-ed into the method of sfa_system_create: two consecutive colours of a
 * Rainbow sweep share one launch where the second reads few rows of the
 * first (the second colour's blocks form those rows' values themselves, with
 * the same arithmetic), fewer launches per iteration; results are the same
 * bit for bit as with one colour per launch (the default).  Measured on
 * MI355X: +1% on the 64^3 P1 Poisson system, -13% on the 48^3 H(curl)
 * system, so off unless asked for. */
#define SFA_FLAG_PAIRED_SWEEPS 0x200
/* Or-ed into the method of sfa_system_create with SFA_FLAG_FP32_INNER: every
 * inner solve of sfa_system_solve_mixed starts its CG afresh (d = z), plain
 * iterative refinement.  By default an inner solve after the first continues
 * the previous one's search direction (residual replacement: the FP64
 * residual replaces the FP32 recursive one, the Krylov space is kept), which
 * needs fewer inner iterations in total.  CG, Jacobi, SSOR and Rainbow
 * variants; the Eisenstat variants always restart. */
#define SFA_FLAG_INNER_RESTARTS 0x400
/* Or-ed into the method of sfa_system_create: the system keeps pointers to
 * csrrows, csrcolumns and csrvalues instead of a host copy of them (12 B per
 * entry, read only by sfa_system_apply); the caller keeps the three arrays
 * alive and unchanged until sfa_system_destroy. */
#define SFA_FLAG_BORROW_CSR 0x800

/* solver variants (one per function of solver/sparsesolver.hpp:11-84) */
#define SFA_METHOD_CG                  0
#define SFA_METHOD_DIAGONAL            1
#define SFA_METHOD_SSOR                2
#define SFA_METHOD_SSOR_EISENSTAT      3
#define SFA_METHOD_RAINBOW             4
#define SFA_METHOD_EISENSTAT_RAINBOW   5

typedef struct sfa_solve_info {
    int    iterations;   /* K, what the reference solver returns             */
    int    success;      /* 1 iff sqrt(z.r) < tolerance at exit ("SUCCESS")  */
    int    exit_reason;  /* SFA_EXIT_*                                       */
    int    reserved;
    double zr;           /* final z.r (r.r for unpreconditioned CG)          */
    double tolerance;    /* max(precision*||b||_2, sqrt(DBL_EPSILON))        */
} sfa_solve_info;

/* Message of the calling thread's last failed call ("" if none). */
const char* sfa_last_error(void);

/* SFA_API_VERSION of the built library. */
int sfa_api_version(void);

/* ---------------------------------------------------------------------------
 * Drop-in solvers on host arrays.  Same arguments, same meaning and the same
 * in/out contract for x (initial guess in, solution out) and residual (rows
 * that a variant skips keep their input) as the reference.  print_modulo
 * selects the reference's log lines (>= -1 begin/finish, >= 0 restart,
 * >= 1 every k-th iteration, >= -2 breakdown/warning), printed to stdout.
 * ------------------------------------------------------------------------- */

/* ConjugateGradientSolverCSR, solver/sparsesolver.hpp:11-19 */
int sfa_ConjugateGradientSolverCSR(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    sfa_solve_info* info);

/* ConjugateGradientSolverCSR_DiagonalPreconditioner, solver/sparsesolver.hpp:21-30 */
int sfa_ConjugateGradientSolverCSR_DiagonalPreconditioner(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    const double* precon, sfa_solve_info* info);

/* ConjugateGradientSolverCSR_SSOR, solver/sparsesolver.hpp:33-43 */
int sfa_ConjugateGradientSolverCSR_SSOR(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    const double* diagonal, double omega, sfa_solve_info* info);

/* ConjugateGradientSolverCSR_SSOR_Eisenstat, solver/sparsesolver.hpp:46-56 */
int sfa_ConjugateGradientSolverCSR_SSOR_Eisenstat(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    const double* diagonal, double omega, sfa_solve_info* info);

/* ConjugateGradientSolverCSR_Rainbow, solver/sparsesolver.hpp:59-70 */
int sfa_ConjugateGradientSolverCSR_Rainbow(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    const double* diagonal, double omega,
    int num_colors, const int* F, const int* B, const int* R,
    sfa_solve_info* info);

/* ConjugateGradientSolverCSR_Eisenstat_Rainbow, solver/sparsesolver.hpp:73-84 */
int sfa_ConjugateGradientSolverCSR_Eisenstat_Rainbow(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    const double* diagonal, double omega,
    int num_colors, const int* F, const int* B, const int* R,
    sfa_solve_info* info);

/* ---------------------------------------------------------------------------
 * No-allocation solve.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The arguments of
 * ConjugateGradientSolverCSR_Rainbow, all arrays DEVICE pointers on the
 * current GPU (x, residual in/out as in the reference; B is also read by the
 * host).  The solve runs on the caller's CSR arrays in their own row order
 * on the null stream and returns when it has finished.  Its Krylov vectors,
 * z.r history, reduction partials and scalars live in one static device
 * array of SFA_NOALLOC_SCRATCH_BYTES that each GPU holds from the first call
 * until the process ends; no hipMalloc, hipFree or other allocating routine
 * is called.  Calls on one GPU share that array: calls that overlap on one
 * GPU race and must not be made.  N > SFA_NOALLOC_MAX_ROWS or num_colors >
 * SFA_NOALLOC_MAX_COLORS is rejected with SFA_ERR_INVALID; the CSR structure
 * and the colouring are checked on the device before any value is read
 * (SFA_ERR_INVALID / SFA_ERR_COLOURING).  Same K, x and residual as
 * sfa_ConjugateGradientSolverCSR_Rainbow up to the summation order of the
 * dot products.
 * ------------------------------------------------------------------------- */
#define SFA_NOALLOC_MAX_ROWS    1048576
#define SFA_NOALLOC_MAX_COLORS  1024
#define SFA_NOALLOC_SCRATCH_BYTES (42u << 20)  /* 42 MiB per GPU */

int sfa_ConjugateGradientSolverCSR_Rainbow_noalloc(
    int N, double* x, const double* b,
    const int* csrrows, const int* csrcolumns, const double* csrvalues,
    double* residual, double precision, int print_modulo,
    const double* diagonal, double omega,
    int num_colors, const int* F, const int* B, const int* R,
    sfa_solve_info* info);

/* ---------------------------------------------------------------------------
 * CSR operations
 * ------------------------------------------------------------------------- */

/* MatrixCSR::apply, sparse/matcsr.hpp:73 / sparse/matcsr.cpp:216-244:
 * dest = scaling * A * src, with dest zeroed first.  Host arrays; the product
 * runs on the GPU. */
int sfa_MatrixCSR_apply(int rows, int columns,
                        const int* A, const int* C, const double* V,
                        double* dest, const double* src, double scaling);

/* MatrixCSR::getDiagonal, sparse/matcsr.hpp:82 / sparse/matcsr.cpp:266-286:
 * diag[r] = sum of the entries of row r in column r (host). */
int sfa_MatrixCSR_getDiagonal(int n, const int* A, const int* C,
                              const double* V, double* diag);

/* Rainbow::Rainbow, sparse/rainbow.hpp:15 / sparse/rainbow.cpp:14-130
 * (do_shuffle = false): the reference's greedy colouring, bit for bit.
 * F and R need n ints, B needs n+1 ints.  Host. */
int sfa_Rainbow(int n, const int* A, const int* C,
                int* F, int* B, int* R, int* num_colors);

/* ---------------------------------------------------------------------------
 * CSR assembly operations on the GPU, bit for bit with the reference's host
 * code: the results are held in an sfa_csr handle (host arrays) until copied
 * out.  All end with the reference's row-wise sort-and-compress
 * (sparse/matcsr.cpp:585-687: entries sorted by column, duplicates summed in
 * the reference's order, zero entries dropped), which runs one row per GPU
 * thread in the reference's loop order.
 * ------------------------------------------------------------------------- */
typedef struct sfa_csr_s* sfa_csr;

/* MatrixCSRMultiplication, sparse/matcsr.hpp:126 / sparse/matcsr.cpp:767-849:
 * m1 (rows1 x cols1) times m2 (rows2 x cols2, rows2 == cols1). */
int sfa_MatrixCSRMultiplication(int rows1, int cols1, const int* A1, const int* C1,
                                const double* V1, int rows2, int cols2, const int* A2,
                                const int* C2, const double* V2, sfa_csr* out);

/* MatrixCSRMultiplication_reduced, sparse/matcsr.hpp:128 /
 * sparse/matcsr.cpp:857-962: only pairs of nonzero values form products;
 * each row's distinct columns in increasing order, each value 0 plus its
 * products in the order they are formed, then the same compression. */
int sfa_MatrixCSRMultiplication_reduced(int rows1, int cols1, const int* A1, const int* C1,
                                        const double* V1, int rows2, int cols2, const int* A2,
                                        const int* C2, const double* V2, sfa_csr* out);

/* MatrixCSR::getTranspose, sparse/matcsr.hpp:84 / sparse/matcsr.cpp:291-350. */
int sfa_MatrixCSR_getTranspose(int rows, int cols, const int* A, const int* C, const double* V,
                               sfa_csr* out);

/* Shape and entries of a result: A needs rows + 1 ints, C and V nnz each. */
int sfa_csr_sizes(sfa_csr m, int* rows, int* cols, int* nnz);
int sfa_csr_copy(sfa_csr m, int* A, int* C, double* V);
int sfa_csr_destroy(sfa_csr m);

/* ---------------------------------------------------------------------------
 * Device-resident systems.  No reference counterpart: the reference
 * rebuilds nothing between calls because it keeps everything in host memory;
 * these keep the matrix, the colour/level schedules and the Krylov
 * workspace in HBM across solves.  sfa_system_create takes host CSR arrays
 * plus the per-variant data of the matching drop-in solver (aux = precon
 * for SFA_METHOD_DIAGONAL, diagonal otherwise; omega, num_colors/F/B/R
 * only where the variant has them -- pass 0 / NULL elsewhere).
 * ------------------------------------------------------------------------- */

typedef struct sfa_system_s* sfa_system;

int sfa_system_create(int method, int N,
                      const int* csrrows, const int* csrcolumns,
                      const double* csrvalues, const double* aux,
                      double omega, int num_colors,
                      const int* F, const int* B, const int* R,
                      sfa_system* out);

/* Solve with the system's matrix.  x, b, residual are N-vectors with the
 * same meaning as in the drop-in call; device_pointers = 1 means they are
 * device (HBM) addresses on the system's GPU, 0 that they are host arrays.
 * Device inputs are read after all work queued on the null stream before the
 * call (the system's own stream waits for it; on a stream set with
 * sfa_system_set_stream the caller orders them).  The same holds for
 * sfa_system_solve_mixed, sfa_system_apply and the Hodge solve with device
 * pointers.  Returns once the solve has finished. */
int sfa_system_solve(sfa_system sys, double* x, const double* b,
                     double* residual, double precision, int print_modulo,
                     int device_pointers, sfa_solve_info* info);

/* Mixed precision.  No reference counterpart (the reference's single-precision
 * build, common.compile.mk:162-172, makes everything FP32).  FP64 iterative
 * refinement: r = b - A x in FP64, the system's variant solves A e = r in
 * FP32 (FP32 matrix, vectors and sweeps; FP64 dot products) until its
 * preconditioned residual has dropped by inner_rtol, x += e in FP64; stops
 * when ||b - A x||_2 <= max(rtol*||b||_2, sqrt(DBL_EPSILON)) or after
 * max_outer corrections.  Each inner solve after the first continues the
 * previous one's search direction d = z + (z.r / z.r_prev) d_prev unless the
 * system was created with SFA_FLAG_INNER_RESTARTS, or the FP64 residual's
 * z.r exceeds 4x the z.r the previous inner solve ended with (its FP32
 * residual has drifted: the old direction no longer fits), or the previous
 * inner solve did not converge; a continued inner solve stops after
 * max(100, 3 K_prev) iterations (the outer loop then corrects and the next
 * inner solve restarts).  Needs a system created with SFA_FLAG_FP32_INNER;
 * rows the variant skips keep their x. */
typedef struct sfa_refine_info {
    int    outer;             /* FP64 corrections applied                    */
    int    inner_iterations;  /* FP32 CG iterations, summed                  */
    int    success;           /* final FP64 residual met the target          */
    int    continued;         /* inner solves that continued the previous
                                 direction (the others restarted)            */
    double residual_norm;     /* ||b - A x||_2 of the returned x (FP64)      */
    double relative_residual; /* residual_norm / ||b||_2                     */
} sfa_refine_info;

int sfa_system_solve_mixed(sfa_system sys, double* x, const double* b,
                           double rtol, double inner_rtol, int max_outer,
                           int device_pointers, sfa_refine_info* info);

/* dest = scaling * A * src on the device (the system's SpMV kernel). */
int sfa_system_apply(sfa_system sys, double* dest, const double* src,
                     double scaling, int device_pointers);

/* Run the solve on this HIP stream (a hipStream_t); NULL restores the
 * system's own stream.  Switching waits for the work queued on the previous
 * stream, and sfa_system_destroy waits for the work on the current one, so a
 * stream set here must stay valid until it is replaced or the system is
 * destroyed. */
int sfa_system_set_stream(sfa_system sys, void* hip_stream);

/* Plan facts: groups per sweep, stored entries, algorithmic bytes. */
typedef struct sfa_system_stats {
    int    n;                 /* rows                                        */
    int    n_active;          /* rows the variant does not skip              */
    long long nnz;            /* CSR entries                                 */
    int    groups_forward;    /* colours / levels of the forward sweep       */
    int    groups_backward;   /* colours / levels of the backward sweep      */
    long long stored_entries; /* entries incl. SELL padding, all parts       */
    double bytes_per_iteration; /* algorithmic HBM bytes of one CG step      */
    double bytes_spmv;        /* algorithmic bytes of one SpMV launch        */
    int    kernels_per_iteration;
    int    spmv_lanes_per_row; /* SpMV mapping in use (sfa_system_tune_spmv) */
    double setup_ms[4];       /* sfa_system_create's wall time (ms): checks and
                                 host plan; the host copy of the CSR arrays
                                 kept for sfa_system_apply; FP64 device
                                 set-up (allocation, uploads); the FP32 copy
                                 (0 without one)                             */
} sfa_system_stats;

int sfa_system_query(sfa_system sys, sfa_system_stats* stats);

/* Kernel classes of a solve, for sfa_system_profile. */
#define SFA_KERNEL_SPMV        0  /* q = A d and d.Ad                         */
#define SFA_KERNEL_FORWARD     1  /* forward sweep of one colour / level      */
#define SFA_KERNEL_BACKWARD    2  /* backward sweep of one colour / level     */
#define SFA_KERNEL_DIRECTION   3  /* d = z + beta d                           */
#define SFA_KERNEL_UPDATE      4  /* x, r update and z.r (CG, Jacobi, Eisenstat) */
#define SFA_KERNEL_TRIANGULAR  5  /* Eisenstat triangular solve of one group  */
#define SFA_KERNEL_COMBINE     6  /* Eisenstat q += t and d.Ad                */
#define SFA_KERNEL_RESTART     7  /* K = 0 residual and set-up kernels        */
#define SFA_KERNEL_CLASSES     8

typedef struct sfa_kernel_profile {
    int    launches;
    int    reserved;
    double milliseconds;  /* sum over launches of the HIP-event interval    */
    double bytes;         /* algorithmic HBM bytes of those launches        */
} sfa_kernel_profile;

/* Runs one solve exactly as sfa_system_solve does, but launches every kernel
 * on its own (no graph) between two HIP events on the system's stream, and
 * returns per-class launch counts, summed event time and algorithmic bytes
 * (out must hold SFA_KERNEL_CLASSES entries).  No reference counterpart;
 * used to price each kernel against the HBM roofline. */
int sfa_system_profile(sfa_system sys, double* x, const double* b,
                       double* residual, double precision,
                       int device_pointers, sfa_kernel_profile* out,
                       sfa_solve_info* info);

/* SpMV mapping.  Matrix parts are SELL-64 slices (64 rows, entries stored
 * column-major); a mapping is the number of lanes that share a row, 1, 2, 4
 * or 8: with K lanes per row a wavefront covers 64/K rows of a slice and
 * lane t of a row reads the row's entries t, t+K, t+2K, ...  Every mapping
 * adds a row's products in the row's entry order, so all give the same
 * q = A d bit for bit; only the block partials of d.Ad follow the mapping.
 * Row-length buckets: the widest row of a slice has <= 8, <= 16, <= 32 or
 * more entries.  No reference counterpart (the reference's SpMV is a serial
 * row loop, sparse/matcsr.cpp:216-244). */
#define SFA_SPMV_BUCKETS  4
#define SFA_SPMV_MAPPINGS 4   /* candidates: 1, 2, 4, 8 lanes per row */
typedef struct sfa_spmv_tuning {
    int       lanes_per_row[SFA_SPMV_BUCKETS];  /* chosen mapping; 0: no rows   */
    long long rows[SFA_SPMV_BUCKETS];           /* rows per bucket              */
    /* mean time of one launch per candidate, in microseconds (0: not timed) */
    double    us[SFA_SPMV_BUCKETS][SFA_SPMV_MAPPINGS];
} sfa_spmv_tuning;

/* Autotunes the system's iteration SpMV: times every candidate mapping on
 * the system's own SpMV kernel (on its device, its matrix, its stream) and
 * installs the fastest.  The SpMV of one system is one launch and runs one
 * mapping, so the candidates are timed over the whole matrix and reported
 * under the bucket that holds most of its entries; every non-empty bucket
 * reports the same choice.  fp32_inner != 0 tunes the FP32 inner system of
 * a mixed-precision system instead.  Eisenstat variants have no SpMV: the
 * call then reports lanes_per_row = 1 and times nothing. */
int sfa_system_tune_spmv(sfa_system sys, int fp32_inner, sfa_spmv_tuning* out);
/* Installs a mapping (1, 2, 4 or 8 lanes per row) without timing, for the
 * system and its FP32 inner system. */
int sfa_system_set_spmv(sfa_system sys, int lanes_per_row);

/* Microseconds per launch of the iteration's forward (SFA_KERNEL_FORWARD) or
 * backward (SFA_KERNEL_BACKWARD) sweep kernels as they run back to back in a
 * recorded graph: the last right-hand side's solve is restarted from x = 0
 * and run one iteration, then reps copies of one iteration's launches of
 * that class are replayed between two HIP events.  SSOR and Rainbow systems,
 * after at least one solve.  No reference counterpart: the launch time a
 * roofline is priced with, free of the per-launch events of
 * sfa_system_profile.  The next solve restarts, so nothing of it remains. */
int sfa_system_time_sweeps(sfa_system sys, int kernel_class, int reps,
                           double* us_per_launch);

int sfa_system_destroy(sfa_system sys);

/* Microseconds per round of a grid-wide barrier inside one resident launch
 * of `blocks` workgroups of `threads` (a cooperative launch; mode 0: one
 * counter, mode 1: a counter per group of workgroups blockIdx % 8, then one
 * top counter), or of a kernel boundary (mode 2: one launch per round,
 * replayed from a graph), timed over `rounds` rounds between two HIP
 * events.  Each round every workgroup stores a value and reads its
 * neighbour's after the barrier; *errors gets bit 0 for a wait that gave
 * up (every wait is bounded) and bit 1 for a stale value seen.  No reference
 * counterpart: the measurement behind one launch per colour in the sweeps. */
int sfa_grid_barrier_cost(int blocks, int threads, int rounds, int mode,
                          double* us_per_round, int* errors);

/* ---------------------------------------------------------------------------
 * Batched and served solves.  No reference counterpart (the reference solves
 * one system per call).  A batch holds up to max_slots (<= 1024) solves of
 * Rainbow CG-SSOR (the iteration of ConjugateGradientSolverCSR_Rainbow,
 * solver/sparsesolver.cpp:1002-1335) over any number of device-resident
 * matrices; each step advances every running solve by one phase in one set of
 * three kernel launches (a captured graph replays 32 steps at a time), so
 * solves of any size share each launch.  Solves can be submitted and
 * collected between sfa_batch_run calls (continuous batching).
 *   SFA_BATCH_FP64:  FP64 throughout; stopping rule of the reference
 *                    (sqrt(z.r) < max(precision*||b||, 2^-26)).
 *   SFA_BATCH_MIXED: FP32 matrix/vectors/sweeps inside FP64 iterative
 *                    refinement, stopping on ||b - A x||_2 <=
 *                    max(precision*||b||, 2^-26) in FP64; each inner solve
 *                    after the first continues the previous one's search
 *                    direction (as sfa_system_solve_mixed).
 *   SFA_BATCH_MIXED_RESTARTS: the same with every inner CG started afresh
 *                    (d = z), plain iterative refinement.
 * ------------------------------------------------------------------------- */
#define SFA_BATCH_FP64  0
#define SFA_BATCH_MIXED 1
#define SFA_BATCH_MIXED_RESTARTS 2

typedef struct sfa_batch_s* sfa_batch;

typedef struct sfa_batch_info {
    int    iterations;     /* CG iterations (FP32 inner iterations, summed, for MIXED) */
    int    outer;          /* FP64 corrections (MIXED)                       */
    int    success;        /* the stopping rule was met                      */
    int    exit_reason;    /* SFA_EXIT_* of the last CG loop                 */
    double zr;             /* final z.r of the last CG loop                  */
    double residual_norm;  /* sqrt(z.r) (FP64) or ||b - A x||_2 (MIXED)      */
} sfa_batch_info;

int sfa_batch_create(int precision_mode, int max_slots, sfa_batch* out);

/* Uploads a matrix with its Rainbow colouring (arguments of
 * ConjugateGradientSolverCSR_Rainbow); returns its id. */
int sfa_batch_add_matrix(sfa_batch b, int N, const int* csrrows,
                         const int* csrcolumns, const double* csrvalues,
                         const double* diagonal, double omega, int num_colors,
                         const int* F, const int* B, const int* R, int* matrix_id);

/* The host half of sfa_batch_add_matrix: builds the matrix's device-order
 * plan (permutation, colour groups, SELL parts) without touching the device
 * or any batch, so any thread can run it while a batch runs; the plan is then
 * uploaded with sfa_batch_add_prepared (on the thread that drives the batch)
 * and freed with sfa_prepared_destroy. */
typedef struct sfa_prepared_s* sfa_prepared;
int sfa_batch_prepare_matrix(int N, const int* csrrows, const int* csrcolumns,
                             const double* csrvalues, const double* diagonal, double omega,
                             int num_colors, const int* F, const int* B, const int* R,
                             sfa_prepared* out);
int sfa_batch_add_prepared(sfa_batch b, sfa_prepared prepared, int* matrix_id);
int sfa_prepared_destroy(sfa_prepared prepared);

/* Frees a matrix's device memory; no slot may still use it. */
int sfa_batch_remove_matrix(sfa_batch b, int matrix_id);

/* Queues a solve of matrix_id for the host right-hand side rhs (N entries,
 * original row order) from x0 (NULL: 0); returns its slot.  inner_rtol and
 * max_outer apply to SFA_BATCH_MIXED only. */
int sfa_batch_submit(sfa_batch b, int matrix_id, const double* rhs,
                     const double* x0, double precision, double inner_rtol,
                     int max_outer, int* slot);

/* Queues nrhs (1..4) solves of matrix_id at once: right-hand side k at
 * rhs + k N, initial guess at x0 + k N (x0 NULL: all 0).  They take nrhs
 * consecutive slots (slots[0..nrhs)) whose vectors are stored interleaved,
 * row i of solve k at i nrhs + k, so the solves riding together gather a
 * column's values from one cache line (four of them with one vector load);
 * each solve's arithmetic, and so its result, is that of sfa_batch_submit.
 * The block is freed with the last of its slots.  No reference counterpart
 * (the reference solves one right-hand side per call). */
int sfa_batch_submit_group(sfa_batch b, int matrix_id, int nrhs, const double* rhs,
                           const double* x0, double precision, double inner_rtol,
                           int max_outer, int* slots);

/* Runs max_steps steps rounded down to whole graph replays of 32 steps, at
 * least one replay (0: until no solve is running); *running receives the
 * number of solves still running. */
int sfa_batch_run(sfa_batch b, int max_steps, int* running);

/* Lists finished, not yet collected slots. */
int sfa_batch_finished(sfa_batch b, int* slots, int capacity, int* count);

/* Copies a finished solve's x (original row order; NULL to skip) and frees
 * its slot. */
int sfa_batch_collect(sfa_batch b, int slot, double* x, sfa_batch_info* info);

/* Device memory held by matrices and slots. */
int sfa_batch_memory(sfa_batch b, long long* bytes);

/* Steps the engine has executed since the batch was created (each step runs
 * the current phase of every running solve: four kernels). */
int sfa_batch_steps(sfa_batch b, long long* steps);
/* Runs `steps` engine steps with the kernels launched on their own between
 * HIP events and adds their summed times, in milliseconds, to ms[0] (the two
 * step kernels), ms[1] (per-slot fold) and ms[2] (schedule).  The
 * solves advance exactly as under sfa_batch_run.  No reference counterpart. */
int sfa_batch_profile(sfa_batch b, int steps, double* ms);

/* Autotunes the batch's SpMV mapping per row-length bucket: for each bucket
 * that holds matrices (a matrix's bucket is the one holding most of its
 * entries), times the SpMV phase of all those matrices under every candidate
 * mapping with the step kernel's SpMV code and grid, and installs the
 * fastest for the bucket; matrices added later take their bucket's choice.
 * Only between solves (SFA_ERR_INVALID while solves are running).  No
 * reference counterpart. */
int sfa_batch_tune_spmv(sfa_batch b, sfa_spmv_tuning* out);
/* Installs lanes per row (1, 2, 4, 8) per bucket without timing; between
 * solves only. */
int sfa_batch_set_spmv(sfa_batch b, const int lanes_per_row[SFA_SPMV_BUCKETS]);

int sfa_batch_destroy(sfa_batch b);

/* ---------------------------------------------------------------------------
 * Block (Hodge) systems.  HodgeConjugateResidualSolverCSR_SSOR,
 * solver/systemsparsesolver.hpp:58-72 (solver/systemsparsesolver.cpp:426-797):
 * conjugate residuals on the Schur complement (B A^-1 Bt + C) x = b, every
 * A^-1 an inner natural-order CG-SSOR solve (omega 1, the diagonal of A) of
 * precision inneriteration_precision.  A is L x L, B is N x L, Bt is L x N,
 * C is N x N, all CSR as MatrixCSR::getA/getC/getV hand them over (C may
 * have no entries).  K and the final Md.r come back in the sfa_solve_info
 * (zr = Md.r; success iff sqrt(Md.r) < tolerance).
 * ------------------------------------------------------------------------- */
int sfa_HodgeConjugateResidualSolverCSR_SSOR(
    int N, int L, double* x, const double* b,
    const int* Arows, const int* Acolumns, const double* Avalues,
    const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues,
    const int* Crows, const int* Ccolumns, const double* Cvalues,
    double* residual, double precision, int print_modulo,
    double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info);

/* HodgeConjugateResidualSolverCSR, solver/systemsparsesolver.hpp:10-24: the
 * reference's plain variant calls the SSOR one (solver/systemsparsesolver.cpp:18-49),
 * and so does this. */
int sfa_HodgeConjugateResidualSolverCSR(
    int N, int L, double* x, const double* b,
    const int* Arows, const int* Acolumns, const double* Avalues,
    const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues,
    const int* Crows, const int* Ccolumns, const double* Cvalues,
    double* residual, double precision, int print_modulo,
    double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info);

/* HodgeConjugateResidualSolverCSR_diagonal, solver/systemsparsesolver.hpp:26-40
 * (solver/systemsparsesolver.cpp:52-424): the same conjugate residual loop
 * with every A^-1 an inner Jacobi CG solve (precon = 1 / the diagonal of A,
 * not guarded against a zero diagonal, as in the reference's NDEBUG build). */
int sfa_HodgeConjugateResidualSolverCSR_diagonal(
    int N, int L, double* x, const double* b,
    const int* Arows, const int* Acolumns, const double* Avalues,
    const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues,
    const int* Crows, const int* Ccolumns, const double* Cvalues,
    double* residual, double precision, int print_modulo,
    double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info);

/* HodgeConjugateResidualSolverCSR_textbook, solver/systemsparsesolver.hpp:42-56
 * (solver/systemsparsesolver.cpp:798-1121): the same loop with every A^-1 an
 * inner plain CG solve. */
int sfa_HodgeConjugateResidualSolverCSR_textbook(
    int N, int L, double* x, const double* b,
    const int* Arows, const int* Acolumns, const double* Avalues,
    const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues,
    const int* Crows, const int* Ccolumns, const double* Cvalues,
    double* residual, double precision, int print_modulo,
    double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info);

/* The same solvers with the four matrices resident on the device between
 * solves (A as an sfa_system with its SSOR schedules, its Jacobi
 * preconditioner or none): for callers that solve the same block system for
 * many right-hand sides.  No reference counterpart.  sfa_hodge_create keeps
 * the SSOR inner solver; sfa_hodge_create_inner takes SFA_HODGE_INNER_SSOR,
 * _DIAGONAL or _TEXTBOOK.  device_pointers = 1: x, b and residual are
 * device buffers. */
#define SFA_HODGE_INNER_SSOR 0
#define SFA_HODGE_INNER_DIAGONAL 1
#define SFA_HODGE_INNER_TEXTBOOK 2
typedef struct sfa_hodge_s* sfa_hodge;
int sfa_hodge_create(int N, int L,
                     const int* Arows, const int* Acolumns, const double* Avalues,
                     const int* Brows, const int* Bcolumns, const double* Bvalues,
                     const int* Btrows, const int* Btcolumns, const double* Btvalues,
                     const int* Crows, const int* Ccolumns, const double* Cvalues,
                     sfa_hodge* out);
int sfa_hodge_create_inner(int N, int L,
                           const int* Arows, const int* Acolumns, const double* Avalues,
                           const int* Brows, const int* Bcolumns, const double* Bvalues,
                           const int* Btrows, const int* Btcolumns, const double* Btvalues,
                           const int* Crows, const int* Ccolumns, const double* Cvalues,
                           int inner, sfa_hodge* out);
int sfa_hodge_solve(sfa_hodge h, double* x, const double* b, double* residual,
                    double precision, int print_modulo,
                    double inneriteration_precision, int inneriteration_print_modulo,
                    int device_pointers, sfa_solve_info* info);
int sfa_hodge_destroy(sfa_hodge h);

#ifdef __cplusplus
}
#endif
#endif /* SIMPLEFEM_AMD_H */
