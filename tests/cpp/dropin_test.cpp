// dropin_test.cpp -- a caller written against the drop-in headers the way the
// reference's own tests are written (tests/solver/solvespd_csr.cpp,
// tests/operators/floatvector.cpp, tests/sparse/sparsematrix.cpp): only the
// include paths differ from a reference test.  Exit status 0 = all checks hold.
//
//   dropin_test cpu   host-side classes only (no GPU needed)
//   dropin_test gpu   products, all six CG variants, InverseOperator and the
//                     Hodge block solver on the GPU
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "feec/base/include.hpp"
#include "feec/operators/floatvector.hpp"
#include "feec/solver/inv.hpp"
#include "feec/solver/sparsesolver.hpp"
#include "feec/solver/systemsparsesolver.hpp"
#include "feec/sparse/matcsr.hpp"
#include "feec/sparse/rainbow.hpp"

static int failures = 0;
#define CHECK(cond)                                                            \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);      \
            failures++;                                                        \
        }                                                                      \
    } while (0)

// 5-point Laplacian of tests/solver/solvespd_csr.cpp:157-177, rows sorted
static MatrixCSR laplacian(int N)
{
    const Float h = Float(1.) / (N + 1), h2 = h * h;
    std::vector<int> A(1, 0), C;
    std::vector<Float> V;
    for (int e = 0; e < N * N; e++) {
        const int x = e / N, y = e % N;
        auto push = [&](int col, Float v) { C.push_back(col); V.push_back(v); };
        if (x != 0) push((x - 1) * N + y, -1. / h2);
        if (y != 0) push(x * N + y - 1, -1. / h2);
        push(e, 4. / h2);
        if (y != N - 1) push(x * N + y + 1, -1. / h2);
        if (x != N - 1) push((x + 1) * N + y, -1. / h2);
        A.push_back((int)C.size());
    }
    return MatrixCSR(N * N, N * N, A, C, V);
}

static void cpu_checks()
{
    // FloatVector (facts checked by tests/operators/floatvector.cpp)
    FloatVector v1(5);
    for (int i = 0; i < 5; i++) CHECK(std::isnan(v1.getentry(i)));
    FloatVector v2(5, 3.0);
    for (int i = 0; i < 5; i++) CHECK(v2[i] == 3.0);
    FloatVector a(5);
    a.zero();
    for (int i = 0; i < 5; i++) CHECK(a[i] == 0.);
    for (int i = 0; i < 5; i++) a.setentry(i, i + 1);
    const FloatVector am = a.getslice(1, 3);
    CHECK(am[0] == a[1] && am[1] == a[2] && am[2] == a[3]);
    const FloatVector api = Constants::pi * a;
    for (int i = 0; i < 5; i++) CHECK(api[i] == Constants::pi * a[i]);
    const FloatVector nega = -a;
    for (int i = 0; i < 5; i++) CHECK(nega[i] == -a[i]);
    FloatVector t(5);
    for (int i = 0; i < 5; i++) t.setentry(i, 3. * i + 1);
    const FloatVector apt = a + t, amt = a - t;
    for (int i = 0; i < 5; i++) CHECK(apt[i] == a[i] + t[i] && amt[i] == a[i] - t[i]);
    CHECK(a * a == 1 + 4 + 9 + 16 + 25);
    a.setslice(1, t.getslice(1, 3));
    CHECK(a[0] == 1 && a[1] == 4 && a[2] == 7 && a[3] == 10 && a[4] == 5);
    a.addslice(1, t.getslice(1, 3), 1000.);
    CHECK(a[0] == 1 && a[1] == 4004 && a[2] == 7007 && a[3] == 10010 && a[4] == 5);
    FloatVector e(0);
    CHECK(e.getdimension() == 0);
    CHECK(a == a && !(a != a) && t != a && e == e);
    FloatVector p{1, 3, 0};
    CHECK(!p.is_positive() && !p.is_negative() && p.is_nonnegative() && !p.is_nonpositive() &&
          !p.is_zero());
    FloatVector z{0, 0, 0};
    CHECK(z.is_zero() && z.is_nonnegative() && z.is_nonpositive());

    // MatrixCSR host methods
    MatrixCSR L = laplacian(6);
    const Float h = Float(1.) / 7, h2 = h * h;
    FloatVector d = L.getDiagonal();
    for (int i = 0; i < d.getdimension(); i++) CHECK(d[i] == 4. / h2);
    MatrixCSR Lt = L.getTranspose();
    CHECK(Lt.getnumberofentries() == L.getnumberofentries());
    for (int k = 0; k < L.getnumberofentries(); k++)
        CHECK(Lt.getC()[k] == L.getC()[k] && Lt.getV()[k] == L.getV()[k]);
    CHECK(L.getmaxrowwidth() == 5);
    CHECK(L.eigenvalueupperbound() == 4. / h2 + 4. * (1. / h2));
    // duplicates merge, zeros drop, rows sort (sparse/matcsr.cpp:585-703)
    MatrixCSR M(2, 3, std::vector<int>{0, 4, 6}, std::vector<int>{2, 0, 2, 1, 1, 1},
                std::vector<Float>{1., 5., 2., 0., 3., -3.});
    M.compressentries();
    CHECK(M.getnumberofentries() == 2);
    CHECK(M.getC()[0] == 0 && M.getV()[0] == 5. && M.getC()[1] == 2 && M.getV()[1] == 3.);
    MatrixCSR S = L - L;
    CHECK(S.getnumberofentries() == 0);

    // Rainbow: tridiagonal 4x4 traced by hand through sparse/rainbow.cpp:14-130
    MatrixCSR T(4, 4, std::vector<int>{0, 2, 5, 8, 10}, std::vector<int>{0, 1, 0, 1, 2, 1, 2, 3, 2, 3},
                std::vector<Float>(10, 1.));
    Rainbow rb(T);
    rb.check();
    CHECK(rb.num_colors == 3);
    CHECK(rb.F == (std::vector<int>{1, 2, 1, 0}));
    CHECK(rb.B == (std::vector<int>{0, 1, 3, 4}));
    CHECK(rb.R == (std::vector<int>{3, 2, 0, 1}));
}

static void gpu_checks()
{
    // tests/sparse/sparsematrix.cpp:51-67
    MatrixCSR M(2, 3, std::vector<int>{0, 3, 6}, std::vector<int>{0, 1, 2, 0, 1, 2},
                std::vector<Float>{1., 2., 3., 4., 5., 6.});
    FloatVector vec{13., 17., 19.};
    FloatVector Mv = M * vec;
    CHECK(Mv[0] == 104. && Mv[1] == 251.);

    // tests/solver/solvespd_csr.cpp flow: random solutions, rhs = A sol, all solvers
    for (int N : {4, 8, 16, 32}) {
        MatrixCSR system = laplacian(N);
        FloatVector diagonal = system.getDiagonal();
        CHECK(diagonal.is_finite() && diagonal.is_nonnegative());
        FloatVector precon = diagonal;
        precon.reciprocal();
        Rainbow rainbow(system);
        for (int trial = 0; trial < 5; trial++) {
            FloatVector sol(N * N);
            sol.random();
            sol.normalize();
            const FloatVector rhs = system * sol;
            for (int variant = 0; variant < 6; variant++) {
                FloatVector mysol(N * N, 0.);
                FloatVector residual(rhs);
                const int n = mysol.getdimension();
                int K = -1;
                const Float omega = 0.9123456789;
                switch (variant) {
                case 0: K = ConjugateGradientSolverCSR(n, mysol.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), residual.raw(), desired_precision, -3); break;
                case 1: K = ConjugateGradientSolverCSR_DiagonalPreconditioner(n, mysol.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), residual.raw(), desired_precision, -3, precon.raw()); break;
                case 2: K = ConjugateGradientSolverCSR_SSOR(n, mysol.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), residual.raw(), desired_precision, -3, diagonal.raw(), omega); break;
                case 3: K = ConjugateGradientSolverCSR_SSOR_Eisenstat(n, mysol.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), residual.raw(), desired_precision, -3, diagonal.raw(), omega); break;
                case 4: K = ConjugateGradientSolverCSR_Rainbow(n, mysol.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), residual.raw(), desired_precision, -3, diagonal.raw(), omega, rainbow.num_colors, rainbow.F.data(), rainbow.B.data(), rainbow.R.data()); break;
                default: K = ConjugateGradientSolverCSR_Eisenstat_Rainbow(n, mysol.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), residual.raw(), desired_precision, -3, diagonal.raw(), omega, rainbow.num_colors, rainbow.F.data(), rainbow.B.data(), rainbow.R.data()); break;
                }
                const Float err = (sol - mysol).norm();
                const Float res = (system * mysol - rhs).norm();
                CHECK(K > 0 && K <= n);
                CHECK(err < 1e-4);
                CHECK(res < 1e-3 * rhs.norm());
                if (trial == 0)
                    std::printf("N=%2d variant %d: K=%4d |sol-x|=%.3e |Ax-b|=%.3e\n", N, variant,
                                K, err, res);
            }
        }
    }
}

// InverseOperator<MatrixCSR> (solver/inv.hpp:151-201): each apply is the
// Eisenstat-Rainbow solve (omega 1) from the previous solution, bit for bit
// the raw call's; a changed operator is seen at the next apply.
static void inverse_operator_checks()
{
    MatrixCSR system = laplacian(16);
    const int n = system.getdimin();
    InverseOperator<MatrixCSR> inv(system, desired_precision);
    const FloatVector diagonal = system.getDiagonal();
    const Rainbow rainbow(system);
    FloatVector prev(n, 0.);
    for (int trial = 0; trial < 3; trial++) {
        FloatVector rhs(n);
        rhs.random();
        FloatVector x(n, 0.);
        inv.apply(x, rhs, 1.);
        FloatVector want(prev), res(prev);
        const int K = ConjugateGradientSolverCSR_Eisenstat_Rainbow(
            n, want.raw(), rhs.raw(), system.getA(), system.getC(), system.getV(), res.raw(),
            desired_precision, -1, diagonal.raw(), 1.0, rainbow.num_colors, rainbow.F.data(),
            rainbow.B.data(), rainbow.R.data());
        bool same = true;
        for (int i = 0; i < n; i++) same = same && x[i] == want[i];
        CHECK(same);
        CHECK(inv.last_iterations() == K);
        std::printf("InverseOperator apply %d: K=%d, same as the raw call: %s\n", trial, K,
                    same ? "yes" : "no");
        prev.copydatafrom(x);
    }
    FloatVector half(n);
    half.random();
    FloatVector x1(n, 0.);
    inv.apply(x1, half, 0.5);  // dest = 0.5 * op^-1 src
    CHECK((2. * (system * x1) - half).norm() < 1e-6 * half.norm());
    system.scale(2.);
    inv.use_previous_sol = false;
    FloatVector rhs(n);
    rhs.random();
    FloatVector x(n, 0.);
    inv.apply(x, rhs, 1.);
    CHECK((system * x - rhs).norm() < 1e-6 * rhs.norm());
}

// The call shape of tests/sullivan2D/darcyessential2D.cpp:207-224 with B = Bt =
// the identity and C = the zero matrix: the Schur complement is A^-1 (A the
// Laplacian plus the identity), so x = A b.
static void hodge_checks()
{
    MatrixCSR lap = laplacian(12);
    const int n = lap.getdimin();
    const Float h2 = Float(1.) / (13 * 13);
    std::vector<int> Ar(1, 0), Ac, Ir(1, 0), Ic, Zr(n + 1, 0);
    std::vector<Float> Av, Iv;
    for (int r = 0; r < n; r++) {
        for (int k = lap.getA()[r]; k < lap.getA()[r + 1]; k++) {
            Ac.push_back(lap.getC()[k]);
            Av.push_back(lap.getV()[k] * h2 + (lap.getC()[k] == r ? 1. : 0.));
        }
        Ar.push_back((int)Ac.size());
        Ic.push_back(r);
        Iv.push_back(1.);
        Ir.push_back(r + 1);
    }
    const MatrixCSR A(n, n, Ar, Ac, Av), I(n, n, Ir, Ic, Iv);
    const MatrixCSR Z(n, n, Zr, std::vector<int>(), std::vector<Float>());
    FloatVector b(n), x(n, 0.), res(n, 0.);
    b.random();
    const int K = HodgeConjugateResidualSolverCSR_SSOR(
        n, n, x.raw(), b.raw(), A.getA(), A.getC(), A.getV(), I.getA(), I.getC(), I.getV(),
        I.getA(), I.getC(), I.getV(), Z.getA(), Z.getC(), Z.getV(), res.raw(), desired_precision,
        -1, desired_precision, -3);
    const FloatVector want = A * b;
    CHECK((x - want).norm() < 1e-6 * want.norm());
    std::printf("HodgeConjugateResidualSolverCSR_SSOR: K=%d, |x - A b| / |A b| = %.2e\n", K,
                (x - want).norm() / want.norm());
    // the plain variant is the SSOR one; the diagonal one has inner Jacobi CG
    FloatVector xp(n, 0.), xd(n, 0.), rp(n, 0.), rd(n, 0.);
    const int Kp = HodgeConjugateResidualSolverCSR(
        n, n, xp.raw(), b.raw(), A.getA(), A.getC(), A.getV(), I.getA(), I.getC(), I.getV(),
        I.getA(), I.getC(), I.getV(), Z.getA(), Z.getC(), Z.getV(), rp.raw(), desired_precision,
        -3, desired_precision, -3);
    CHECK(Kp == K && (xp - x).norm() == 0.);
    const int Kd = HodgeConjugateResidualSolverCSR_diagonal(
        n, n, xd.raw(), b.raw(), A.getA(), A.getC(), A.getV(), I.getA(), I.getC(), I.getV(),
        I.getA(), I.getC(), I.getV(), Z.getA(), Z.getC(), Z.getV(), rd.raw(), desired_precision,
        -1, desired_precision, -3);
    CHECK((xd - want).norm() < 1e-6 * want.norm());
    std::printf("HodgeConjugateResidualSolverCSR_diagonal: K=%d, |x - A b| / |A b| = %.2e\n", Kd,
                (xd - want).norm() / want.norm());
    FloatVector xt(n, 0.), rt(n, 0.);
    const int Kt = HodgeConjugateResidualSolverCSR_textbook(
        n, n, xt.raw(), b.raw(), A.getA(), A.getC(), A.getV(), I.getA(), I.getC(), I.getV(),
        I.getA(), I.getC(), I.getV(), Z.getA(), Z.getC(), Z.getV(), rt.raw(), desired_precision,
        -1, desired_precision, -3);
    CHECK((xt - want).norm() < 1e-6 * want.norm());
    std::printf("HodgeConjugateResidualSolverCSR_textbook: K=%d, |x - A b| / |A b| = %.2e\n", Kt,
                (xt - want).norm() / want.norm());
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    cpu_checks();
    if (gpu) gpu_checks();
    if (gpu) inverse_operator_checks();
    if (gpu) hodge_checks();
    std::printf("%s checks: %d failure(s)\n", gpu ? "cpu+gpu" : "cpu", failures);
    return failures == 0 ? 0 : 1;
}
