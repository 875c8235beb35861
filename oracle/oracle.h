/*
 * oracle.h -- CPU restatement of FEEC++'s sparse CG/SSOR solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the HIP path with this
 * library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in simplefem_amd/
 * links, loads or calls it.
 *
 * Each function restates one routine of the reference (martinlicht/simplefem)
 * loop by loop, in the reference's floating-point operation order, with
 * reductions in sequential row order (the order the reference produces when
 * it runs with one OpenMP thread).  Built with -ffp-contract=off so that no
 * multiply-add is fused; the reference build in oracle/_ref uses the same
 * flag, which makes the two bit-identical at one thread.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   orc_csr_apply, orc_csr_diagonal  -- dense brute force; reference KATs
 *                                       (tests/sparse/sparsematrix.cpp:13-69)
 *   orc_rainbow                       -- hand KAT, mex closed form, validity
 *                                       invariant, oracle/_ref bit-exact
 *   orc_cg_*                          -- diagonal-matrix KAT (K=1, x=b/d),
 *                                       dense Cholesky solve, natural SSOR ==
 *                                       Rainbow SSOR with one colour per row,
 *                                       oracle/_ref bit-exact (K, x, residual)
 *   orc_hodge_cr_ssor/_diagonal/_textbook -- dense Schur-complement solve,
 *                                       B = identity KAT (x = (A^-1 + C)^-1 b),
 *                                       oracle/_ref bit-exact (K, x, residual)
 */
#ifndef SFA_ORACLE_H
#define SFA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Exit reasons of the CG loop (the reference only logs these). */
enum {
    ORC_EXIT_CONVERGED = 0,   /* |zr| < tol^2 at the top of the loop       */
    ORC_EXIT_CAP       = 1,   /* K reached N                               */
    ORC_EXIT_BREAKDOWN = 2,   /* non-finite or negative zr / dAd           */
    ORC_EXIT_SMALL     = 3    /* sqrt|dAd| < machine epsilon               */
};

typedef struct orc_info {
    int    iterations;   /* K, the reference's return value              */
    int    success;      /* 1 iff sqrt(zr) < tolerance at exit (SUCCESS)  */
    int    exit_reason;  /* ORC_EXIT_*                                    */
    int    pad_;
    double zr;           /* final z^T r (r^T r for plain CG)              */
    double tolerance;    /* max(precision*||b||, sqrt(DBL_EPSILON))       */
} orc_info;

void orc_csr_apply(int nrows, const int* A, const int* C, const double* V,
                   const double* src, double* dest, double scaling);
void orc_csr_diagonal(int n, const int* A, const int* C, const double* V,
                      double* diag);
int  orc_rainbow(int n, const int* A, const int* C, int* F, int* B, int* R);

int orc_cg(int N, double* x, const double* b, const int* A, const int* C,
           const double* V, double* residual, double precision,
           int print_modulo, orc_info* info);
int orc_cg_jacobi(int N, double* x, const double* b, const int* A,
                  const int* C, const double* V, double* residual,
                  double precision, int print_modulo, const double* precon,
                  orc_info* info);
int orc_cg_ssor(int N, double* x, const double* b, const int* A, const int* C,
                const double* V, double* residual, double precision,
                int print_modulo, const double* diagonal, double omega,
                orc_info* info);
int orc_cg_ssor_eisenstat(int N, double* x, const double* b, const int* A,
                          const int* C, const double* V, double* residual,
                          double precision, int print_modulo,
                          const double* diagonal, double omega,
                          orc_info* info);
int orc_cg_rainbow(int N, double* x, const double* b, const int* A,
                   const int* C, const double* V, double* residual,
                   double precision, int print_modulo, const double* diagonal,
                   double omega, int num_colors, const int* F, const int* B,
                   const int* R, orc_info* info);
int orc_cg_eisenstat_rainbow(int N, double* x, const double* b, const int* A,
                             const int* C, const double* V, double* residual,
                             double precision, int print_modulo,
                             const double* diagonal, double omega,
                             int num_colors, const int* F, const int* B,
                             const int* R, orc_info* info);

/* HodgeConjugateResidualSolverCSR_SSOR (solver/systemsparsesolver.cpp:426-797):
 * B is N x L, Bt is L x N, C is N x N, A is L x L. */
int orc_hodge_cr_ssor(int N, int L, double* x, const double* b,
                      const int* Arows, const int* Acols, const double* Avals,
                      const int* Brows, const int* Bcols, const double* Bvals,
                      const int* Btrows, const int* Btcols, const double* Btvals,
                      const int* Crows, const int* Ccols, const double* Cvals,
                      double* res, double precision, int print_modulo,
                      double inner_precision, int inner_print_modulo,
                      orc_info* info);

/* HodgeConjugateResidualSolverCSR_textbook (solver/systemsparsesolver.cpp:798-1121):
 * the same loop with inner plain CG solves. */
int orc_hodge_cr_textbook(int N, int L, double* x, const double* b,
                          const int* Arows, const int* Acols, const double* Avals,
                          const int* Brows, const int* Bcols, const double* Bvals,
                          const int* Btrows, const int* Btcols, const double* Btvals,
                          const int* Crows, const int* Ccols, const double* Cvals,
                          double* res, double precision, int print_modulo,
                          double inner_precision, int inner_print_modulo,
                          orc_info* info);

/* HodgeConjugateResidualSolverCSR_diagonal (solver/systemsparsesolver.cpp:52-424):
 * the same loop with inner Jacobi CG solves, precon = 1 / diag(A). */
int orc_hodge_cr_diagonal(int N, int L, double* x, const double* b,
                          const int* Arows, const int* Acols, const double* Avals,
                          const int* Brows, const int* Bcols, const double* Bvals,
                          const int* Btrows, const int* Btcols, const double* Btvals,
                          const int* Crows, const int* Ccols, const double* Cvals,
                          double* res, double precision, int print_modulo,
                          double inner_precision, int inner_print_modulo,
                          orc_info* info);

#ifdef __cplusplus
}
#endif
#endif
