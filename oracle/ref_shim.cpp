// ref_shim.cpp -- extern "C" entry points into the reference's own solver,
// colouring and CSR code, compiled from the reference sources by
// oracle/Makefile into oracle/_ref/libfeecref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle and by bench.py's
// cpu_baseline leg.  The shim adds nothing to the algorithms; it only turns
// raw arrays into the reference's MatrixCSR / Rainbow objects and times the
// solver call.
#include <chrono>
#include <cstring>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#endif

#include "base/include.hpp"
#include "solver/sparsesolver.hpp"
#include "solver/systemsparsesolver.hpp"
#include "sparse/matcsr.hpp"
#include "sparse/rainbow.hpp"
#include "sparse/sparsematrix.hpp"

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

MatrixCSR make_csr(int n, const int* A, const int* C, const double* V)
{
    const int nnz = A[n];
    return MatrixCSR(n, n, std::vector<int>(A, A + n + 1), std::vector<int>(C, C + nnz),
                     std::vector<Float>(V, V + nnz));
}

// The result of the last feecref_coo_to_csr / feecref_csr_op call, until
// feecref_result copies it out.
struct Result {
    int rows = 0, cols = 0;
    std::vector<int> A, C;
    std::vector<double> V;
} g_result;

int keep(const MatrixCSR& m)
{
    g_result.rows = m.getdimout();
    g_result.cols = m.getdimin();
    const int nnz = m.getA()[m.getdimout()];
    g_result.A.assign(m.getA(), m.getA() + m.getdimout() + 1);
    g_result.C.assign(m.getC(), m.getC() + nnz);
    g_result.V.assign(m.getV(), m.getV() + nnz);
    return nnz;
}

MatrixCSR make_rect(int rows, int cols, const int* A, const int* C, const double* V)
{
    const int nnz = A[rows];
    return MatrixCSR(rows, cols, std::vector<int>(A, A + rows + 1), std::vector<int>(C, C + nnz),
                     std::vector<Float>(V, V + nnz));
}

}  // namespace

extern "C" {

// MatrixCSR(const SparseMatrix&) (sparse/matcsr.cpp:43-89) of a list of entries.
int feecref_coo_to_csr(int rows, int cols, int n, const int* er, const int* ec, const double* ev)
{
    std::vector<SparseMatrix::MatrixEntry> list(n);
    for (int i = 0; i < n; i++) {
        list[i].row = er[i];
        list[i].column = ec[i];
        list[i].value = ev[i];
    }
    SparseMatrix sm(rows, cols, list);
    return keep(MatrixCSR(sm));
}

// 0 MatrixCSRMultiplication, 1 MatrixCSRMultiplication_reduced,
// 2 Conjugation(center = m1, factor = m2), 3 m1.getTranspose(),
// 4 MatrixCSRAddition(m1, m2, s1, s2), 5 m1.compressentries().
int feecref_csr_op(int op, int r1, int c1, const int* A1, const int* C1, const double* V1, int r2,
                   int c2, const int* A2, const int* C2, const double* V2, double s1, double s2)
{
    MatrixCSR m1 = make_rect(r1, c1, A1, C1, V1);
    switch (op) {
    case 0: return keep(MatrixCSRMultiplication(m1, make_rect(r2, c2, A2, C2, V2)));
    case 1: return keep(MatrixCSRMultiplication_reduced(m1, make_rect(r2, c2, A2, C2, V2)));
    case 2: return keep(Conjugation(m1, make_rect(r2, c2, A2, C2, V2)));
    case 3: return keep(m1.getTranspose());
    case 4: return keep(MatrixCSRAddition(m1, make_rect(r2, c2, A2, C2, V2), s1, s2));
    case 5: m1.compressentries(); return keep(m1);
    default: return -1;
    }
}

void feecref_result(int* dims, int* A, int* C, double* V)
{
    dims[0] = g_result.rows;
    dims[1] = g_result.cols;
    std::memcpy(A, g_result.A.data(), sizeof(int) * g_result.A.size());
    std::memcpy(C, g_result.C.data(), sizeof(int) * g_result.C.size());
    std::memcpy(V, g_result.V.data(), sizeof(double) * g_result.V.size());
}

double feecref_desired_precision() { return desired_precision; }

int feecref_omp_threads()
{
#if defined(_OPENMP)
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Rainbow::Rainbow (sparse/rainbow.cpp:14-130).  B needs room for n+1 ints.
int feecref_rainbow(int n, const int* A, const int* C, const double* V, int* F, int* B, int* R)
{
    MatrixCSR mat = make_csr(n, A, C, V);
    Rainbow rb(mat);
    std::memcpy(F, rb.F.data(), sizeof(int) * n);
    std::memcpy(B, rb.B.data(), sizeof(int) * (rb.num_colors + 1));
    std::memcpy(R, rb.R.data(), sizeof(int) * n);
    return rb.num_colors;
}

// MatrixCSR::apply (sparse/matcsr.cpp:216-244) through LinearOperator.
void feecref_csr_apply(int n, const int* A, const int* C, const double* V,
                       const double* src, double* dest, double scaling)
{
    MatrixCSR mat = make_csr(n, A, C, V);
    FloatVector s(n, 0.), d(n, 0.);
    for (int i = 0; i < n; i++) s[i] = src[i];
    mat.apply(d, s, scaling);
    for (int i = 0; i < n; i++) dest[i] = d[i];
}

// MatrixCSR::getDiagonal (sparse/matcsr.cpp:266-286).
void feecref_csr_diagonal(int n, const int* A, const int* C, const double* V, double* diag)
{
    MatrixCSR mat = make_csr(n, A, C, V);
    FloatVector d = mat.getDiagonal();
    for (int i = 0; i < n; i++) diag[i] = d[i];
}

// The six CG entry points of solver/sparsesolver.hpp:11-84.  `variant`:
// 0 plain, 1 diagonal (aux = precon), 2 SSOR, 3 SSOR-Eisenstat, 4 Rainbow,
// 5 Eisenstat-Rainbow.  Wall time of the solver call alone goes to *seconds.
int feecref_cg(int variant, int N, double* x, const double* b, const int* A, const int* C,
               const double* V, double* residual, double precision, int print_modulo,
               const double* aux, double omega, int num_colors, const int* F, const int* B,
               const int* R, double* seconds)
{
    int K = -1;
    auto t0 = std::chrono::steady_clock::now();
    switch (variant) {
    case 0:
        K = ConjugateGradientSolverCSR(N, x, b, A, C, V, residual, precision, print_modulo);
        break;
    case 1:
        K = ConjugateGradientSolverCSR_DiagonalPreconditioner(N, x, b, A, C, V, residual,
                                                              precision, print_modulo, aux);
        break;
    case 2:
        K = ConjugateGradientSolverCSR_SSOR(N, x, b, A, C, V, residual, precision, print_modulo,
                                            aux, omega);
        break;
    case 3:
        K = ConjugateGradientSolverCSR_SSOR_Eisenstat(N, x, b, A, C, V, residual, precision,
                                                      print_modulo, aux, omega);
        break;
    case 4:
        K = ConjugateGradientSolverCSR_Rainbow(N, x, b, A, C, V, residual, precision,
                                               print_modulo, aux, omega, num_colors, F, B, R);
        break;
    case 5:
        K = ConjugateGradientSolverCSR_Eisenstat_Rainbow(N, x, b, A, C, V, residual, precision,
                                                         print_modulo, aux, omega, num_colors, F,
                                                         B, R);
        break;
    default:
        return -1;
    }
    if (seconds) *seconds = seconds_since(t0);
    return K;
}

// HodgeConjugateResidualSolverCSR_SSOR (solver/systemsparsesolver.cpp:426-797).
int feecref_hodge_cr_ssor(int N, int L, double* x, const double* b, const int* Ar, const int* Ac,
                          const double* Av, const int* Br, const int* Bc, const double* Bv,
                          const int* Btr, const int* Btc, const double* Btv, const int* Cr,
                          const int* Cc, const double* Cv, double* res, double precision,
                          int print_modulo, double inner_precision, int inner_print_modulo)
{
    return HodgeConjugateResidualSolverCSR_SSOR(N, L, x, b, Ar, Ac, Av, Br, Bc, Bv, Btr, Btc, Btv,
                                                Cr, Cc, Cv, res, precision, print_modulo,
                                                inner_precision, inner_print_modulo);
}

// HodgeConjugateResidualSolverCSR_textbook (solver/systemsparsesolver.cpp:798-1121).
int feecref_hodge_cr_textbook(int N, int L, double* x, const double* b, const int* Ar,
                              const int* Ac, const double* Av, const int* Br, const int* Bc,
                              const double* Bv, const int* Btr, const int* Btc, const double* Btv,
                              const int* Cr, const int* Cc, const double* Cv, double* res,
                              double precision, int print_modulo, double inner_precision,
                              int inner_print_modulo)
{
    return HodgeConjugateResidualSolverCSR_textbook(N, L, x, b, Ar, Ac, Av, Br, Bc, Bv, Btr, Btc,
                                                    Btv, Cr, Cc, Cv, res, precision, print_modulo,
                                                    inner_precision, inner_print_modulo);
}

// HodgeConjugateResidualSolverCSR_diagonal (solver/systemsparsesolver.cpp:52-424).
int feecref_hodge_cr_diagonal(int N, int L, double* x, const double* b, const int* Ar,
                              const int* Ac, const double* Av, const int* Br, const int* Bc,
                              const double* Bv, const int* Btr, const int* Btc, const double* Btv,
                              const int* Cr, const int* Cc, const double* Cv, double* res,
                              double precision, int print_modulo, double inner_precision,
                              int inner_print_modulo)
{
    return HodgeConjugateResidualSolverCSR_diagonal(N, L, x, b, Ar, Ac, Av, Br, Bc, Bv, Btr, Btc,
                                                    Btv, Cr, Cc, Cv, res, precision, print_modulo,
                                                    inner_precision, inner_print_modulo);
}

}  // extern "C"
