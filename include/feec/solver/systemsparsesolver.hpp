// feec/solver/systemsparsesolver.hpp -- FEEC++'s block (Hodge) solvers
// HodgeConjugateResidualSolverCSR_SSOR (solver/systemsparsesolver.hpp:58-72),
// HodgeConjugateResidualSolverCSR (:10-24, which calls the SSOR one, as in the
// reference), HodgeConjugateResidualSolverCSR_diagonal (:26-40, inner Jacobi
// CG) and HodgeConjugateResidualSolverCSR_textbook (:42-56, inner plain CG)
// with their signatures unchanged: conjugate residuals on the Schur
// complement (B A^-1 Bt + C) x = b, every A^-1 an inner CG solve.  A (L x L) is
// kept on the GPU with its preconditioner for the whole call; every inner
// solve and product runs there.  Returns the iteration count K; the log
// lines follow print_modulo as in the reference; a failed argument check
// aborts, as the reference's asserts do.  The reference's
// HodgeHerzogSoodhalterMethod is not part of this library.
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

int HodgeConjugateResidualSolverCSR(
    const int N, const int L, Float* x, const Float* b,
    const int* Arows, const int* Acolumns, const Float* Avalues,
    const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues,
    const int* Crows, const int* Ccolumns, const Float* Cvalues,
    Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo);

int HodgeConjugateResidualSolverCSR_diagonal(
    const int N, const int L, Float* x, const Float* b,
    const int* Arows, const int* Acolumns, const Float* Avalues,
    const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues,
    const int* Crows, const int* Ccolumns, const Float* Cvalues,
    Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo);

int HodgeConjugateResidualSolverCSR_textbook(
    const int N, const int L, Float* x, const Float* b,
    const int* Arows, const int* Acolumns, const Float* Avalues,
    const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues,
    const int* Crows, const int* Ccolumns, const Float* Cvalues,
    Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo);
