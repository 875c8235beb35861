/*
 * oracle.c -- CPU restatement of FEEC++'s CSR CG / SSOR / Rainbow-SSOR path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests compare the HIP path with it; never the product.
 * Every loop below follows the reference routine named in its comment, with
 * the same operation order; reductions run in sequential row order.
 * Paths are relative to the reference repository (martinlicht/simplefem).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* base/base.hpp:121-128: desired_precision = Sqrt(machine_epsilon), which the
 * reference evaluates by a 40-step Newton iteration that lands exactly on
 * 2^-26 for IEEE double (checked in tests/test_oracle_solvers.py). */
static const double ORC_DESIRED_PRECISION = 1.4901161193847656e-08;
static const double ORC_MACHINE_EPSILON   = DBL_EPSILON;

/* base/base.hpp:229-237 */
static double orc_absolute(double x)
{
    if (x >= 0) return x;
    if (x <= 0) return -x;
    return x; /* NaN passes through */
}

/* base/base.hpp:275-283 (two-argument case) */
static double orc_maximum(double a, double b)
{
    if (a >= b) return a;
    if (a <= b) return b;
    return a;
}

/* solver/sparsesolver.cpp:46-49 (identical in every CG variant, e.g. :469-472) */
static double orc_tolerance(int N, const double* b, double precision)
{
    double t = 0.;
    for (int i = 0; i < N; i++) t += b[i] * b[i];
    t = precision * sqrt(t);
    return orc_maximum(t, ORC_DESIRED_PRECISION);
}

static double* orc_fill(int N, double v)
{
    double* p = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; i++) p[i] = v;
    return p;
}

static void orc_finish(orc_info* info, int K, double zr, double tol, int reason)
{
    if (!info) return;
    info->iterations  = K;
    info->zr          = zr;
    info->tolerance   = tol;
    info->exit_reason = reason;
    info->success     = sqrt(zr) < tol ? 1 : 0; /* :713-714 */
}

/* The three breakdown tests shared by all variants, e.g.
 * solver/sparsesolver.cpp:629-642.  Returns an ORC_EXIT_* or -1. */
static int orc_breakdown(double zr, double dAd, int print_modulo, int warning_needs_print)
{
    int numerator_bad   = !isfinite(zr) || zr < 0.;
    int denominator_bad = !isfinite(dAd) || dAd < 0.;
    int denominator_small = sqrt(orc_absolute(dAd)) < ORC_MACHINE_EPSILON;
    if (numerator_bad || denominator_bad) return ORC_EXIT_BREAKDOWN;
    /* The natural-order Eisenstat variant only breaks on a small energy when
     * warnings are printed (solver/sparsesolver.cpp:936). */
    if (denominator_small && (!warning_needs_print || print_modulo >= -2))
        return ORC_EXIT_SMALL;
    return -1;
}

/* MatrixCSR::apply, sparse/matcsr.cpp:216-244 */
void orc_csr_apply(int nrows, const int* A, const int* C, const double* V,
                   const double* src, double* dest, double scaling)
{
    for (int i = 0; i < nrows; i++) dest[i] = 0.;
    for (int i = 0; i < nrows; i++)
        for (int j = A[i]; j < A[i + 1]; j++)
            dest[i] += scaling * V[j] * src[C[j]];
}

/* MatrixCSR::getDiagonal, sparse/matcsr.cpp:266-286 */
void orc_csr_diagonal(int n, const int* A, const int* C, const double* V, double* diag)
{
    for (int r = 0; r < n; r++) {
        diag[r] = 0.;
        for (int i = A[r]; i < A[r + 1]; i++)
            if (C[i] == r) diag[r] += V[i];
    }
}

/* Rainbow::Rainbow, sparse/rainbow.cpp:14-130 (do_shuffle = false).
 * Greedy colouring in row order; a row takes the smallest colour not used by
 * any other column of its row (unvisited rows still hold colour 0).  B needs
 * room for n+1 entries.  Returns num_colors. */
int orc_rainbow(int n, const int* A, const int* C, int* F, int* B, int* R)
{
    int num_colors = 1;                              /* :36 */
    for (int r = 0; r < n; r++) F[r] = 0;            /* :46 */
    for (int r = 0; r < n; r++) {                    /* :48-76 */
        int again = 1;
        while (again) {
            again = 0;
            for (int c = A[r]; c < A[r + 1]; c++) {
                int j = C[c];
                if (j == r) continue;
                if (F[j] == F[r]) { F[r]++; again = 1; break; }
            }
        }
        if (F[r] == num_colors) num_colors++;        /* :67 */
    }
    int* counter = (int*)calloc((size_t)num_colors, sizeof(int));
    for (int r = 0; r < n; r++) counter[F[r]]++;     /* :84 */
    B[0] = 0;                                        /* :92-97 */
    for (int c = 1; c < num_colors + 1; c++) B[c] = counter[c - 1];
    for (int c = 1; c < num_colors + 1; c++) B[c] += B[c - 1];
    for (int r = 0; r < n; r++) R[r] = -1;           /* :105-117 */
    for (int r = 0; r < n; r++) {
        int color = F[r];
        counter[color]--;
        R[B[color] + counter[color]] = r;
    }
    free(counter);
    return num_colors;
}

/* ConjugateGradientSolverCSR, solver/sparsesolver.cpp:26-201 */
int orc_cg(int N, double* x, const double* b, const int* A, const int* C,
           const double* V, double* residual, double precision,
           int print_modulo, orc_info* info)
{
    (void)print_modulo;
    const double tol = orc_tolerance(N, b, precision);
    double* direction = orc_fill(N, NAN);            /* :62-63 */
    double* auxiliary = orc_fill(N, NAN);
    double r_r = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {                                /* :80-98 */
            r_r = 0.;
            for (int c = 0; c < N; c++) {
                residual[c] = b[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    residual[c] -= V[d] * x[C[d]];
                direction[c] = residual[c];
                r_r += residual[c] * residual[c];
            }
        }
        if (orc_absolute(r_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; } /* :116-119 */
        double d_Ad = 0.;                            /* :127-139 */
        for (int c = 0; c < N; c++) {
            auxiliary[c] = 0.;
            for (int d = A[c]; d < A[c + 1]; d++)
                auxiliary[c] += V[d] * direction[C[d]];
            d_Ad += direction[c] * auxiliary[c];
        }
        int br = orc_breakdown(r_r, d_Ad, print_modulo, 0);
        if (br >= 0) { reason = br; break; }
        double alpha = r_r / d_Ad;                   /* :158 */
        double r_r_new = 0.;                         /* :160-174 */
        for (int c = 0; c < N; c++) {
            x[c] += alpha * direction[c];
            residual[c] -= alpha * auxiliary[c];
            r_r_new += residual[c] * residual[c];
        }
        double beta = r_r_new / r_r;                 /* :175-177 */
        r_r = r_r_new;
        for (int c = 0; c < N; c++)                  /* :180-183 */
            direction[c] = residual[c] + beta * direction[c];
        K++;
    }
    orc_finish(info, K, r_r, tol, reason);
    free(direction); free(auxiliary);
    return K;
}

/* ConjugateGradientSolverCSR_DiagonalPreconditioner, solver/sparsesolver.cpp:221-416 */
int orc_cg_jacobi(int N, double* x, const double* b, const int* A,
                  const int* C, const double* V, double* residual,
                  double precision, int print_modulo, const double* precon,
                  orc_info* info)
{
    const double tol = orc_tolerance(N, b, precision);
    double* direction = orc_fill(N, NAN);            /* :258-260 */
    double* zirconium = orc_fill(N, NAN);
    double* auxiliary = orc_fill(N, NAN);
    double z_r = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {                                /* :281-306 */
            z_r = 0.;
            for (int c = 0; c < N; c++) {
                if (precon[c] == 0.) continue;
                residual[c] = b[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    residual[c] -= V[d] * x[C[d]];
                zirconium[c] = precon[c] * residual[c];
                direction[c] = zirconium[c];
                z_r += zirconium[c] * residual[c];
            }
        }
        if (orc_absolute(z_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; }
        double d_Ad = 0.;                            /* :339-351 */
        for (int c = 0; c < N; c++) {
            if (precon[c] == 0.) continue;
            auxiliary[c] = 0.;
            for (int d = A[c]; d < A[c + 1]; d++)
                auxiliary[c] += V[d] * direction[C[d]];
            d_Ad += direction[c] * auxiliary[c];
        }
        int br = orc_breakdown(z_r, d_Ad, print_modulo, 0);
        if (br >= 0) { reason = br; break; }
        double alpha = z_r / d_Ad;                   /* :370 */
        double z_r_new = 0.;                         /* :372-389 */
        for (int c = 0; c < N; c++) {
            if (precon[c] == 0.) continue;
            x[c] += alpha * direction[c];
            residual[c] -= alpha * auxiliary[c];
            zirconium[c] = precon[c] * residual[c];
            z_r_new += zirconium[c] * residual[c];
        }
        double beta = z_r_new / z_r;
        z_r = z_r_new;
        for (int c = 0; c < N; c++) {                /* :395-402 */
            if (precon[c] == 0.) continue;
            direction[c] = zirconium[c] + beta * direction[c];
        }
        K++;
    }
    orc_finish(info, K, z_r, tol, reason);
    free(direction); free(zirconium); free(auxiliary);
    return K;
}

/* ConjugateGradientSolverCSR_SSOR, solver/sparsesolver.cpp:446-724 */
int orc_cg_ssor(int N, double* x, const double* b, const int* A, const int* C,
                const double* V, double* residual, double precision,
                int print_modulo, const double* diagonal, double omega,
                orc_info* info)
{
    const double tol = orc_tolerance(N, b, precision);
    double* direction = orc_fill(N, 0.);             /* :485-495 */
    double* zirconium = orc_fill(N, 0.);
    double* auxiliary = orc_fill(N, 0.);
    double* mittlerer = orc_fill(N, 0.);
    double z_r = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {
            for (int c = 0; c < N; c++) {            /* :512-524 */
                if (diagonal[c] == 0.) continue;
                residual[c] = b[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    residual[c] -= V[d] * x[C[d]];
            }
            for (int c = 0; c < N; c++) {            /* :528-541 */
                if (diagonal[c] == 0.) continue;
                double aux = residual[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (C[d] < c) aux -= V[d] * mittlerer[C[d]];
                mittlerer[c] = aux * omega / diagonal[c];
            }
            for (int c = 0; c < N; c++) {            /* :543-553 */
                if (diagonal[c] == 0.) continue;
                mittlerer[c] *= diagonal[c];
                mittlerer[c] *= (2. - omega) / omega;
            }
            for (int c = N - 1; c >= 0; c--) {       /* :555-568 */
                if (diagonal[c] == 0.) continue;
                double aux = mittlerer[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (C[d] > c) aux -= V[d] * zirconium[C[d]];
                zirconium[c] = aux * omega / diagonal[c];
            }
            z_r = 0.;                                /* :570-583 */
            for (int c = 0; c < N; c++) {
                if (diagonal[c] == 0.) continue;
                direction[c] = zirconium[c];
                z_r += zirconium[c] * residual[c];
            }
        }
        if (orc_absolute(z_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; } /* :601-604 */
        double d_Ad = 0.;                            /* :611-626 */
        for (int c = 0; c < N; c++) {
            auxiliary[c] = 0.;
            if (diagonal[c] == 0.) continue;
            for (int d = A[c]; d < A[c + 1]; d++)
                auxiliary[c] += V[d] * direction[C[d]];
            d_Ad += direction[c] * auxiliary[c];
        }
        int br = orc_breakdown(z_r, d_Ad, print_modulo, 0);
        if (br >= 0) { reason = br; break; }
        const double alpha = z_r / d_Ad;             /* :645 */
        for (int c = 0; c < N; c++) {                /* :647-665 */
            if (diagonal[c] == 0.) continue;
            x[c] += alpha * direction[c];
            residual[c] -= alpha * auxiliary[c];
            double aux = residual[c];
            for (int d = A[c]; d < A[c + 1]; d++)
                if (C[d] < c) aux -= V[d] * mittlerer[C[d]];
            mittlerer[c] = aux * omega / diagonal[c];
        }
        double z_r_new = 0.;                         /* :667-692 */
        for (int c = N - 1; c >= 0; c--) {
            if (diagonal[c] == 0.) continue;
            mittlerer[c] *= diagonal[c];
            mittlerer[c] *= (2. - omega) / omega;
            double aux = mittlerer[c];
            for (int d = A[c]; d < A[c + 1]; d++)
                if (C[d] > c) aux -= V[d] * zirconium[C[d]];
            zirconium[c] = aux * omega / diagonal[c];
            z_r_new += zirconium[c] * residual[c];
        }
        const double beta = z_r_new / z_r;           /* :694-696 */
        z_r = z_r_new;
        for (int c = 0; c < N; c++) {                /* :698-706 */
            if (diagonal[c] == 0.) continue;
            direction[c] = zirconium[c] + beta * direction[c];
        }
        K++;
    }
    orc_finish(info, K, z_r, tol, reason);
    free(direction); free(zirconium); free(auxiliary); free(mittlerer);
    return K;
}

/* ConjugateGradientSolverCSR_SSOR_Eisenstat, solver/sparsesolver.cpp:730-995 */
int orc_cg_ssor_eisenstat(int N, double* x, const double* b, const int* A,
                          const int* C, const double* V, double* residual,
                          double precision, int print_modulo,
                          const double* diagonal, double omega,
                          orc_info* info)
{
    const double tol = orc_tolerance(N, b, precision);
    double* direction = orc_fill(N, NAN);            /* :769-776 */
    double* zirconium = orc_fill(N, NAN);
    double* technical = orc_fill(N, NAN);
    double* auxiliary = orc_fill(N, NAN);
    double z_r = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {
            for (int c = 0; c < N; c++) {            /* :797-812 */
                if (diagonal[c] == 0.) continue;
                double aux = b[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (diagonal[C[d]] != 0.) aux -= V[d] * x[C[d]];
                auxiliary[c] = aux;
            }
            for (int c = 0; c < N; c++) {            /* :814-828 */
                if (diagonal[c] == 0.) continue;
                double aux = auxiliary[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (C[d] < c && diagonal[C[d]] != 0.) aux -= V[d] * residual[C[d]];
                residual[c] = aux * omega / diagonal[c];
            }
            z_r = 0.;                                /* :830-844 */
            for (int c = 0; c < N; c++) {
                if (diagonal[c] == 0.) continue;
                zirconium[c] = residual[c] * diagonal[c] * (2. - omega) / omega;
                direction[c] = zirconium[c];
                z_r += zirconium[c] * residual[c];
            }
        }
        if (orc_absolute(z_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; } /* :861-864 */
        double d_Ad = 0.;
        for (int c = N - 1; c >= 0; c--) {           /* :876-890 */
            if (diagonal[c] == 0.) continue;
            double aux = direction[c];
            for (int d = A[c]; d < A[c + 1]; d++)
                if (C[d] > c && diagonal[C[d]] != 0.) aux -= V[d] * technical[C[d]];
            technical[c] = aux * omega / diagonal[c];
        }
        for (int c = 0; c < N; c++) {                /* :892-908 */
            if (diagonal[c] == 0.) continue;
            double aux = direction[c] - (2. / omega - 1.) * diagonal[c] * technical[c];
            for (int d = A[c]; d < A[c + 1]; d++)
                if (C[d] < c && diagonal[C[d]] != 0.) aux -= V[d] * auxiliary[C[d]];
            auxiliary[c] = aux * omega / diagonal[c];
        }
        for (int c = 0; c < N; c++) {                /* :912-922 */
            if (diagonal[c] == 0.) continue;
            auxiliary[c] += technical[c];
            d_Ad += auxiliary[c] * direction[c];
        }
        int br = orc_breakdown(z_r, d_Ad, print_modulo, 1); /* :926-945 */
        if (br >= 0) { reason = br; break; }
        const double alpha = z_r / d_Ad;
        double z_r_new = 0.;                         /* :952-968 */
        for (int c = 0; c < N; c++) {
            if (diagonal[c] == 0.) continue;
            x[c] += alpha * technical[c];
            residual[c] -= alpha * auxiliary[c];
            zirconium[c] = residual[c] * (diagonal[c] * (2. - omega) / omega);
            z_r_new += zirconium[c] * residual[c];
        }
        const double beta = z_r_new / z_r;
        z_r = z_r_new;
        for (int c = 0; c < N; c++) {                /* :974-982 */
            if (diagonal[c] == 0.) continue;
            direction[c] = zirconium[c] + beta * direction[c];
        }
        K++;
    }
    orc_finish(info, K, z_r, tol, reason);
    free(direction); free(zirconium); free(technical); free(auxiliary);
    return K;
}

/* ConjugateGradientSolverCSR_Rainbow, solver/sparsesolver.cpp:1002-1335 */
int orc_cg_rainbow(int N, double* x, const double* b, const int* A,
                   const int* C, const double* V, double* residual,
                   double precision, int print_modulo, const double* diagonal,
                   double omega, int num_colors, const int* F, const int* B,
                   const int* R, orc_info* info)
{
    const double tol = orc_tolerance(N, b, precision);
    double* direction = orc_fill(N, 0.);             /* :1046-1057 */
    double* zirconium = orc_fill(N, 0.);
    double* auxiliary = orc_fill(N, 0.);
    double* mittlerer = orc_fill(N, 0.);
    double z_r = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {
            for (int c = 0; c < N; c++) {            /* :1075-1087 */
                if (diagonal[c] == 0.) continue;
                residual[c] = b[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    residual[c] -= V[d] * x[C[d]];
            }
            for (int color = 0; color < num_colors; color++)   /* :1091-1116 */
                for (int i = B[color]; i < B[color + 1]; i++) {
                    int c = R[i];
                    if (diagonal[c] == 0.) continue;
                    double aux = residual[c];
                    for (int d = A[c]; d < A[c + 1]; d++)
                        if (F[C[d]] < F[c]) aux -= V[d] * mittlerer[C[d]];
                    mittlerer[c] = aux * omega / diagonal[c];
                }
            for (int c = 0; c < N; c++) {            /* :1118-1128 */
                if (diagonal[c] == 0.) continue;
                mittlerer[c] *= diagonal[c];
                mittlerer[c] *= (2. - omega) / omega;
            }
            for (int color = num_colors - 1; color >= 0; color--)  /* :1130-1154 */
                for (int i = B[color]; i < B[color + 1]; i++) {
                    int c = R[i];
                    if (diagonal[c] == 0.) continue;
                    double aux = mittlerer[c];
                    for (int d = A[c]; d < A[c + 1]; d++)
                        if (F[C[d]] > F[c]) aux -= V[d] * zirconium[C[d]];
                    zirconium[c] = aux * omega / diagonal[c];
                }
            z_r = 0.;                                /* :1156-1169 */
            for (int c = 0; c < N; c++) {
                if (diagonal[c] == 0.) continue;
                direction[c] = zirconium[c];
                z_r += zirconium[c] * residual[c];
            }
        }
        if (orc_absolute(z_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; } /* :1187-1190 */
        double d_Ad = 0.;                            /* :1197-1212 */
        for (int c = 0; c < N; c++) {
            auxiliary[c] = 0.;
            if (diagonal[c] == 0.) continue;
            for (int d = A[c]; d < A[c + 1]; d++)
                auxiliary[c] += V[d] * direction[C[d]];
            d_Ad += direction[c] * auxiliary[c];
        }
        int br = orc_breakdown(z_r, d_Ad, print_modulo, 0);  /* :1215-1228 */
        if (br >= 0) { reason = br; break; }
        const double alpha = z_r / d_Ad;             /* :1231 */
        for (int color = 0; color < num_colors; color++)   /* :1234-1261 */
            for (int i = B[color]; i < B[color + 1]; i++) {
                int c = R[i];
                if (diagonal[c] == 0.) continue;
                x[c] += alpha * direction[c];
                residual[c] -= alpha * auxiliary[c];
                double aux = residual[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (F[C[d]] < F[c]) aux -= V[d] * mittlerer[C[d]];
                mittlerer[c] = aux * omega / diagonal[c];
            }
        double z_r_new = 0.;                         /* :1263-1303 */
        for (int color = num_colors - 1; color >= 0; color--) {
            double z_r_new_local = 0.;
            for (int i = B[color]; i < B[color + 1]; i++) {
                int c = R[i];
                if (diagonal[c] == 0.) continue;
                mittlerer[c] *= diagonal[c];
                mittlerer[c] *= (2. - omega) / omega;
                double aux = mittlerer[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (F[C[d]] > F[c]) aux -= V[d] * zirconium[C[d]];
                zirconium[c] = aux * omega / diagonal[c];
                z_r_new_local += zirconium[c] * residual[c];
            }
            z_r_new += z_r_new_local;
        }
        const double beta = z_r_new / z_r;           /* :1305-1307 */
        z_r = z_r_new;
        for (int c = 0; c < N; c++) {                /* :1309-1318 */
            if (diagonal[c] == 0.) continue;
            direction[c] = zirconium[c] + beta * direction[c];
        }
        K++;
    }
    orc_finish(info, K, z_r, tol, reason);
    free(direction); free(zirconium); free(auxiliary); free(mittlerer);
    return K;
}

/* ConjugateGradientSolverCSR_Eisenstat_Rainbow, solver/sparsesolver.cpp:1342-1645.
 * Unlike the natural-order Eisenstat variant, the iteration's triangular
 * solves do not skip columns whose diagonal is zero (:1522-1524,1550-1552), so
 * entries that couple to such rows multiply the NaN-initialised scratch. */
int orc_cg_eisenstat_rainbow(int N, double* x, const double* b, const int* A,
                             const int* C, const double* V, double* residual,
                             double precision, int print_modulo,
                             const double* diagonal, double omega,
                             int num_colors, const int* F, const int* B,
                             const int* R, orc_info* info)
{
    const double tol = orc_tolerance(N, b, precision);
    double* direction = orc_fill(N, NAN);            /* :1386-1398 */
    double* zirconium = orc_fill(N, NAN);
    double* technical = orc_fill(N, NAN);
    double* auxiliary = orc_fill(N, NAN);
    double z_r = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {
            for (int c = 0; c < N; c++) {            /* :1412-1427 */
                if (diagonal[c] == 0.) continue;
                double aux = b[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (diagonal[C[d]] != 0.) aux -= V[d] * x[C[d]];
                auxiliary[c] = aux;
            }
            for (int color = 0; color < num_colors; color++)   /* :1429-1455 */
                for (int i = B[color]; i < B[color + 1]; i++) {
                    const int c = R[i];
                    if (diagonal[c] == 0.) continue;
                    double aux = auxiliary[c];
                    for (int d = A[c]; d < A[c + 1]; d++)
                        if (F[C[d]] < F[c] && diagonal[C[d]] != 0.)
                            aux -= V[d] * residual[C[d]];
                    residual[c] = aux * omega / diagonal[c];
                }
            z_r = 0.;                                /* :1457-1471 */
            for (int c = 0; c < N; c++) {
                if (diagonal[c] == 0.) continue;
                zirconium[c] = residual[c] * diagonal[c] * (2. - omega) / omega;
                direction[c] = zirconium[c];
                z_r += zirconium[c] * residual[c];
            }
        }
        if (orc_absolute(z_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; } /* :1488-1491 */
        double d_Ad = 0.;
        for (int color = num_colors - 1; color >= 0; color--)  /* :1501-1533 */
            for (int i = B[color]; i < B[color + 1]; i++) {
                int c = R[i];
                if (diagonal[c] == 0.) continue;
                double aux = direction[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (F[C[d]] > F[c]) aux -= V[d] * technical[C[d]];
                technical[c] = aux * omega / diagonal[c];
            }
        for (int color = 0; color < num_colors; color++)   /* :1536-1560 */
            for (int i = B[color]; i < B[color + 1]; i++) {
                int c = R[i];
                if (diagonal[c] == 0.) continue;
                double aux = direction[c] - (2. / omega - 1.) * diagonal[c] * technical[c];
                for (int d = A[c]; d < A[c + 1]; d++)
                    if (F[C[d]] < F[c]) aux -= V[d] * auxiliary[C[d]];
                auxiliary[c] = aux * omega / diagonal[c];
            }
        for (int c = 0; c < N; c++) {                /* :1564-1574 */
            if (diagonal[c] == 0.) continue;
            auxiliary[c] += technical[c];
            d_Ad += auxiliary[c] * direction[c];
        }
        int br = orc_breakdown(z_r, d_Ad, print_modulo, 0);  /* :1578-1591 */
        if (br >= 0) { reason = br; break; }
        const double alpha = z_r / d_Ad;
        double z_r_new = 0.;                         /* :1598-1614 */
        for (int c = 0; c < N; c++) {
            if (diagonal[c] == 0.) continue;
            x[c] += alpha * technical[c];
            residual[c] -= alpha * auxiliary[c];
            zirconium[c] = residual[c] * (diagonal[c] * (2. - omega) / omega);
            z_r_new += zirconium[c] * residual[c];
        }
        const double beta = z_r_new / z_r;
        z_r = z_r_new;
        for (int c = 0; c < N; c++) {                /* :1620-1628 */
            if (diagonal[c] == 0.) continue;
            direction[c] = zirconium[c] + beta * direction[c];
        }
        K++;
    }
    orc_finish(info, K, z_r, tol, reason);
    free(direction); free(zirconium); free(technical); free(auxiliary);
    return K;
}

/* HodgeConjugateResidualSolverCSR_SSOR, solver/systemsparsesolver.cpp:426-797:
 * conjugate residuals on the Schur complement (B A^-1 Bt + C) x = b (with
 * expected_sign_of_A = 1, :8), every A^-1 an inner ConjugateGradientSolverCSR_SSOR
 * (omega 1, the diagonal of A).  The two inner solves of the restart start
 * from 0, the iteration's from the previous inner solution.  info->zr holds
 * the final Md_r (the reference's restart flags are both false, :11-12, so
 * only K == 0 restarts).  HodgeConjugateResidualSolverCSR_diagonal
 * (:52-424) is the same loop with every inner solve a
 * ConjugateGradientSolverCSR_DiagonalPreconditioner with precon = 1 / the
 * diagonal of A (:76-89, :123-131, :167-175, :259-267): inner = 1.
 * HodgeConjugateResidualSolverCSR_textbook (:798-1121) is the same loop with
 * every inner solve a plain ConjugateGradientSolverCSR (:895-903, :938-946,
 * :1030-1038): inner = 2 (its Mr_r is the Md_r here: at the restart
 * Mres = Mdir, :975-977, and both are updated alike after). */
static void orc_hodge_inner(int inner, int L, double* aux2, const double* aux1,
                            const int* Arows, const int* Acols, const double* Avals,
                            double* auxR, double inner_precision, int inner_print_modulo,
                            const double* diag_or_precon)
{
    if (inner == 2)
        orc_cg(L, aux2, aux1, Arows, Acols, Avals, auxR, inner_precision, inner_print_modulo,
               NULL);
    else if (inner == 1)
        orc_cg_jacobi(L, aux2, aux1, Arows, Acols, Avals, auxR, inner_precision,
                      inner_print_modulo, diag_or_precon, NULL);
    else
        orc_cg_ssor(L, aux2, aux1, Arows, Acols, Avals, auxR, inner_precision,
                    inner_print_modulo, diag_or_precon, 1., NULL);
}

static int orc_hodge_cr(int inner, int N, int L, double* x, const double* b,
                        const int* Arows, const int* Acols, const double* Avals,
                        const int* Brows, const int* Bcols, const double* Bvals,
                        const int* Btrows, const int* Btcols, const double* Btvals,
                        const int* Crows, const int* Ccols, const double* Cvals,
                        double* res, double precision, int print_modulo,
                        double inner_precision, int inner_print_modulo, orc_info* info)
{
    const double tol = orc_tolerance(N, b, precision);           /* :457-461 */
    double* dir  = orc_fill(N, 0.);
    double* Mdir = orc_fill(N, 0.);
    double* Mres = orc_fill(N, 0.);
    double* vil  = orc_fill(N, 0.);
    double* aux1 = orc_fill(L, 0.);
    double* aux2 = orc_fill(L, 0.);
    double* auxR = orc_fill(L, 0.);
    double* diagonal = orc_fill(L, 0.);
    for (int c = 0; c < L; c++)                                   /* :498-511 */
        for (int d = Arows[c]; d < Arows[c + 1]; d++)
            if (Acols[d] == c) diagonal[c] += Avals[d];
    if (inner == 1)                                               /* :85 */
        for (int c = 0; c < L; c++) diagonal[c] = 1. / diagonal[c];
    double Md_r = NAN, Md_Md = NAN;
    int K = 0, reason = ORC_EXIT_CAP;
    while (K < N) {
        if (K == 0) {
            for (int c = 0; c < L; c++) {                         /* :533-542 */
                aux1[c] = 0.;
                for (int d = Btrows[c]; d < Btrows[c + 1]; d++) aux1[c] += Btvals[d] * x[Btcols[d]];
            }
            for (int c = 0; c < L; c++) aux2[c] = 0.;            /* :544-554 */
            orc_hodge_inner(inner, L, aux2, aux1, Arows, Acols, Avals, auxR, inner_precision,
                            inner_print_modulo, diagonal);
            for (int c = 0; c < N; c++) {                         /* :556-570 */
                res[c] = b[c];
                for (int d = Brows[c]; d < Brows[c + 1]; d++) res[c] -= 1. * Bvals[d] * aux2[Bcols[d]];
                for (int d = Crows[c]; d < Crows[c + 1]; d++) res[c] -= Cvals[d] * x[Ccols[d]];
                dir[c] = res[c];
            }
            Md_r = 0.;
            Md_Md = 0.;
            for (int c = 0; c < L; c++) {                         /* :576-586 */
                aux1[c] = 0.;
                for (int d = Btrows[c]; d < Btrows[c + 1]; d++) aux1[c] += Btvals[d] * dir[Btcols[d]];
            }
            for (int c = 0; c < L; c++) aux2[c] = 0.;            /* :588-598 */
            orc_hodge_inner(inner, L, aux2, aux1, Arows, Acols, Avals, auxR, inner_precision,
                            inner_print_modulo, diagonal);
            for (int c = 0; c < N; c++) {                         /* :600-619 */
                Mdir[c] = 0.;
                for (int d = Brows[c]; d < Brows[c + 1]; d++) Mdir[c] += 1. * Bvals[d] * aux2[Bcols[d]];
                for (int d = Crows[c]; d < Crows[c + 1]; d++) Mdir[c] += Cvals[d] * dir[Ccols[d]];
                Mres[c] = Mdir[c];
                Md_r += Mdir[c] * res[c];
                Md_Md += Mdir[c] * Mdir[c];
            }
        }
        if (!isfinite(Md_r) || Md_r < 0.) { reason = ORC_EXIT_BREAKDOWN; break; }  /* :632-637 */
        if (orc_absolute(Md_r) < tol * tol) { reason = ORC_EXIT_CONVERGED; break; } /* :639-642 */
        if (!isfinite(Md_Md) || Md_Md < 0.) { reason = ORC_EXIT_BREAKDOWN; break; } /* :646-652 */
        if (sqrt(orc_absolute(Md_Md)) < ORC_MACHINE_EPSILON) { reason = ORC_EXIT_SMALL; break; } /* :654-658 */
        const double alpha = Md_r / Md_Md;                        /* :663 */
        double new_Mr_r = 0.;
        for (int c = 0; c < L; c++) {                             /* :667-677 */
            aux1[c] = 0.;
            for (int d = Btrows[c]; d < Btrows[c + 1]; d++) aux1[c] += Btvals[d] * Mdir[Btcols[d]];
        }
        orc_hodge_inner(inner, L, aux2, aux1, Arows, Acols, Avals, auxR,  /* :679-689 */
                        inner_precision, inner_print_modulo, diagonal);
        for (int c = 0; c < N; c++) {                             /* :691-721 */
            vil[c] = 0.;
            for (int d = Brows[c]; d < Brows[c + 1]; d++) vil[c] += 1. * Bvals[d] * aux2[Bcols[d]];
            for (int d = Crows[c]; d < Crows[c + 1]; d++) vil[c] += Cvals[d] * Mdir[Ccols[d]];
            x[c] = x[c] + alpha * dir[c];
            res[c] = res[c] - alpha * Mdir[c];
            Mres[c] = Mres[c] - alpha * vil[c];
            new_Mr_r += Mres[c] * res[c];
        }
        const double beta = new_Mr_r / Md_r;                      /* :723 */
        Md_Md = 0.;
        for (int c = 0; c < N; c++) {                             /* :725-739 */
            dir[c] = res[c] + beta * dir[c];
            Mdir[c] = Mres[c] + beta * Mdir[c];
            Md_Md += Mdir[c] * Mdir[c];
        }
        Md_r = new_Mr_r;
        K++;
    }
    orc_finish(info, K, Md_r, tol, reason);
    free(dir); free(Mdir); free(Mres); free(vil);
    free(aux1); free(aux2); free(auxR); free(diagonal);
    return K;
}

int orc_hodge_cr_ssor(int N, int L, double* x, const double* b,
                      const int* Arows, const int* Acols, const double* Avals,
                      const int* Brows, const int* Bcols, const double* Bvals,
                      const int* Btrows, const int* Btcols, const double* Btvals,
                      const int* Crows, const int* Ccols, const double* Cvals,
                      double* res, double precision, int print_modulo,
                      double inner_precision, int inner_print_modulo, orc_info* info)
{
    return orc_hodge_cr(0, N, L, x, b, Arows, Acols, Avals, Brows, Bcols, Bvals, Btrows, Btcols,
                        Btvals, Crows, Ccols, Cvals, res, precision, print_modulo, inner_precision,
                        inner_print_modulo, info);
}

int orc_hodge_cr_textbook(int N, int L, double* x, const double* b,
                          const int* Arows, const int* Acols, const double* Avals,
                          const int* Brows, const int* Bcols, const double* Bvals,
                          const int* Btrows, const int* Btcols, const double* Btvals,
                          const int* Crows, const int* Ccols, const double* Cvals,
                          double* res, double precision, int print_modulo,
                          double inner_precision, int inner_print_modulo, orc_info* info)
{
    return orc_hodge_cr(2, N, L, x, b, Arows, Acols, Avals, Brows, Bcols, Bvals, Btrows, Btcols,
                        Btvals, Crows, Ccols, Cvals, res, precision, print_modulo, inner_precision,
                        inner_print_modulo, info);
}

int orc_hodge_cr_diagonal(int N, int L, double* x, const double* b,
                          const int* Arows, const int* Acols, const double* Avals,
                          const int* Brows, const int* Bcols, const double* Bvals,
                          const int* Btrows, const int* Btcols, const double* Btvals,
                          const int* Crows, const int* Ccols, const double* Cvals,
                          double* res, double precision, int print_modulo,
                          double inner_precision, int inner_print_modulo, orc_info* info)
{
    return orc_hodge_cr(1, N, L, x, b, Arows, Acols, Avals, Brows, Bcols, Bvals, Btrows, Btcols,
                        Btvals, Crows, Ccols, Cvals, res, precision, print_modulo, inner_precision,
                        inner_print_modulo, info);
}
