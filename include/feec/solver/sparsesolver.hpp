// feec/solver/sparsesolver.hpp -- FEEC++'s raw-pointer CSR CG solvers
// (solver/sparsesolver.hpp:11-84) with their signatures unchanged.  Each
// call copies the host arrays to the GPU, solves there with the HIP kernels
// of libsimplefem_amd and copies x and residual back; the return value is
// the iteration count K, and the log lines follow print_modulo as in the
// reference.  A failed argument check aborts, as the reference's asserts do.
#pragma once

#include "../base/include.hpp"

int ConjugateGradientSolverCSR(const int N, Float* x, const Float* b, const int* csrrows,
                               const int* csrcolumns, const Float* csrvalues, Float* residual,
                               Float precision, int print_modulo);

int ConjugateGradientSolverCSR_DiagonalPreconditioner(const int N, Float* x, const Float* b,
                                                      const int* csrrows, const int* csrcolumns,
                                                      const Float* csrvalues, Float* residual,
                                                      Float precision, int print_modulo,
                                                      const Float* precon);

int ConjugateGradientSolverCSR_SSOR(const int N, Float* x, const Float* b, const int* csrrows,
                                    const int* csrcolumns, const Float* csrvalues, Float* residual,
                                    Float precision, int print_modulo, const Float* diagonal,
                                    Float omega);

int ConjugateGradientSolverCSR_SSOR_Eisenstat(const int N, Float* x, const Float* b,
                                              const int* csrrows, const int* csrcolumns,
                                              const Float* csrvalues, Float* residual,
                                              Float precision, int print_modulo,
                                              const Float* diagonal, Float omega);

int ConjugateGradientSolverCSR_Rainbow(const int N, Float* x, const Float* b, const int* csrrows,
                                       const int* csrcolumns, const Float* csrvalues,
                                       Float* residual, Float precision, int print_modulo,
                                       const Float* diagonal, Float omega, int num_colors,
                                       const int* F, const int* B, const int* R);

int ConjugateGradientSolverCSR_Eisenstat_Rainbow(const int N, Float* x, const Float* b,
                                                 const int* csrrows, const int* csrcolumns,
                                                 const Float* csrvalues, Float* residual,
                                                 Float precision, int print_modulo,
                                                 const Float* diagonal, Float omega,
                                                 int num_colors, const int* F, const int* B,
                                                 const int* R);
