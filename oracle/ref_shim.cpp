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
#include "sparse/matcsr.hpp"
#include "sparse/rainbow.hpp"

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

}  // namespace

extern "C" {

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

}  // extern "C"
