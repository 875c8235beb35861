// feec/solver/systemsparsesolver.hpp -- FEEC++'s block (Hodge) solver
// HodgeConjugateResidualSolverCSR_SSOR (solver/systemsparsesolver.hpp:58-72)
// with its signature unchanged: conjugate residuals on the Schur complement
// (B A^-1 Bt + C) x = b, every A^-1 an inner CG-SSOR solve.  A (L x L) is
// kept on the GPU with its SSOR schedules for the whole call; every inner
// solve and product runs there.  Returns the iteration count K; the log
// lines follow print_modulo as in the reference; a failed argument check
// aborts, as the reference's asserts do.  The reference's other block
// solvers (the plain, diagonal and textbook conjugate residual variants and
// HodgeHerzogSoodhalterMethod) are not part of this library.
#pragma once

#include "../base/include.hpp"

const Float expected_sign_of_A = 1.;

int HodgeConjugateResidualSolverCSR_SSOR(
    const int N, const int L, Float* x, const Float* b,
    const int* Arows, const int* Acolumns, const Float* Avalues,
    const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues,
    const int* Crows, const int* Ccolumns, const Float* Cvalues,
    Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo);
