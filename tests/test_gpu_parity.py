"""GPU parity of the HIP path against the oracle, through the C-ABI.

Bars (stated per test):
  * MatrixCSR::apply, restart residuals: bit-exact (same per-row order, no
    fused multiply-add on either side);
  * CG solves: K within max(1, 2% of K) of the oracle, exit reason equal,
    x and residual within 1e-10 relative (||.||_inf against the oracle's),
    NaN patterns equal where the reference breaks down.
"""
import numpy as np
import pytest

import oracle
from simplefem_amd import capi
from simplefem_amd import systems as S

pytestmark = pytest.mark.gpu

REL = 1e-10


def _aux_for(variant, system):
    d = system.diagonal()
    if variant == "cg":
        return None
    if variant == "jacobi":
        safe = np.where(d != 0, d, 1.0)
        return np.where(d != 0, 1.0 / safe, 0.0)
    return d


def _gpu_solve(variant, system, b, aux, omega, col, x0=None, r0=None, precision=1e-10,
               print_modulo=-3):
    n = system.n
    x = np.zeros(n) if x0 is None else x0.copy()
    r = np.zeros(n) if r0 is None else r0.copy()
    info = {}
    args = (n, x, b, system.A, system.C, system.V, r, precision, print_modulo)
    if variant == "cg":
        capi.ConjugateGradientSolverCSR(*args, info=info)
    elif variant == "jacobi":
        capi.ConjugateGradientSolverCSR_DiagonalPreconditioner(*args, aux, info=info)
    elif variant == "ssor":
        capi.ConjugateGradientSolverCSR_SSOR(*args, aux, omega, info=info)
    elif variant == "ssor_eisenstat":
        capi.ConjugateGradientSolverCSR_SSOR_Eisenstat(*args, aux, omega, info=info)
    elif variant == "rainbow":
        capi.ConjugateGradientSolverCSR_Rainbow(*args, aux, omega, *col, info=info)
    else:
        capi.ConjugateGradientSolverCSR_Eisenstat_Rainbow(*args, aux, omega, *col, info=info)
    return x, r, info


def _close(a, b, rel=REL, scale=None):
    """Same NaN pattern and |a-b|_inf <= rel * max(|b|_inf, scale, tiny)."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), "NaN patterns differ"
    fa, fb = a[~na], b[~nb]
    if fb.size == 0:
        return
    scale = max(np.abs(fb).max(), scale or 0.0, 1e-300)
    err = np.abs(fa - fb).max() / scale
    assert err <= rel, f"relative error {err:.3e} > {rel:.1e}"


def _check_parity(variant, system, b, omega=1.0, precision=1e-10, x0=None, r0=None):
    aux = _aux_for(variant, system)
    col = oracle.rainbow(system.n, system.A, system.C)
    xo, ro, io = oracle.solve(variant, system.A, system.C, system.V, b, x0=x0, residual0=r0,
                              precision=precision, aux=aux, omega=omega, colouring=col)
    xg, rg, ig = _gpu_solve(variant, system, b, aux, omega, col, x0=x0, r0=r0,
                            precision=precision)
    Ko, Kg = io["iterations"], ig["iterations"]
    assert abs(Kg - Ko) <= max(1, int(0.02 * Ko)), (Ko, Kg)
    assert ig["exit_reason"] == io["exit_reason"], (io, ig)
    assert ig["success"] == io["success"]
    # host pointers: the tolerance is summed in row order on both sides
    assert ig["tolerance"] == io["tolerance"] or (np.isnan(ig["tolerance"])
                                                  and np.isnan(io["tolerance"]))
    _close(xg, xo)
    # The final residual is a difference of O(||b||) terms: compare on that scale.
    _close(rg, ro, scale=np.nanmax(np.abs(b)) if np.isfinite(b).any() else 1.0)
    return io, ig


SYSTEMS = {
    "fd2d_12": lambda: S.fd_laplacian_2d(12),
    "p1_2d_L4": lambda: S.poisson2d_p1(4),
    "kuhn_6": lambda: S.poisson3d_p1_kuhn(6),
    "spd_400": lambda: S.random_spd(400, 0.02, 7),
}


@pytest.mark.parametrize("variant", oracle.VARIANTS)
@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_solver_parity(variant, name):
    system = SYSTEMS[name]()
    b = S.rhs_uniform(system, seed=11)
    print(f"[parity] {variant} on {name} n={system.n} nnz={system.nnz}", flush=True)
    _check_parity(variant, system, b, omega=1.0)


@pytest.mark.parametrize("variant", ["ssor", "rainbow", "ssor_eisenstat"])
def test_omega_and_initial_guess(variant):
    """solvespd_csr.cpp:530-566 uses omega = 0.9123456789; nonzero x0 and residual0."""
    system = S.fd_laplacian_2d(16)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, system.n)
    x0 = rng.uniform(-1, 1, system.n)
    r0 = rng.uniform(-1, 1, system.n)
    _check_parity(variant, system, b, omega=0.9123456789, x0=x0, r0=r0)


def test_apply_bitexact():
    for system in (S.fd_laplacian_2d(9), S.poisson2d_p1(5), S.poisson3d_p1_kuhn(7),
                   S.random_spd(257, 0.05, 1)):
        src = S.uniform_pm1(5, system.n)
        for scaling in (1.0, -0.37):
            want = oracle.csr_apply(system.n, system.A, system.C, system.V, src, scaling)
            got = capi.csr_apply(system.n, system.n, system.A, system.C, system.V, src, scaling)
            assert np.array_equal(got, want), system.name


def test_apply_keeps_parts_and_sees_changes():
    """sfa_MatrixCSR_apply keeps the device parts of recent matrices: applies
    that alternate between matrices give each one's product with the oracle's
    bits, and a matrix changed in place (one value near the end of 13.7 MB of
    values, compared by several host threads) gives the new matrix's."""
    small, large = S.random_spd(300, 0.05, 3), S.poisson3d_p1_kuhn(48)
    xs, xl = S.uniform_pm1(1, small.n), S.uniform_pm1(2, large.n)
    print(f"case: apply cache, n={small.n} and n={large.n}", flush=True)
    for _ in range(3):
        for s, x in ((small, xs), (large, xl)):
            got = capi.csr_apply(s.n, s.n, s.A, s.C, s.V, x)
            assert np.array_equal(got, oracle.csr_apply(s.n, s.A, s.C, s.V, x)), s.name
    for s, x in ((small, xs), (large, xl)):
        s.V[-3] *= 2.0
        try:
            got = capi.csr_apply(s.n, s.n, s.A, s.C, s.V, x)
            assert np.array_equal(got, oracle.csr_apply(s.n, s.A, s.C, s.V, x)), s.name
        finally:
            s.V[-3] /= 2.0
        assert np.array_equal(capi.csr_apply(s.n, s.n, s.A, s.C, s.V, x),
                              oracle.csr_apply(s.n, s.A, s.C, s.V, x)), s.name


def test_apply_rectangular_kat():
    """tests/sparse/sparsematrix.cpp:51-67: [[1,2,3],[4,5,6]] x [13,17,19] = [104, 251]."""
    A = np.array([0, 3, 6], dtype=np.int32)
    C = np.array([0, 1, 2, 0, 1, 2], dtype=np.int32)
    V = np.array([1., 2., 3., 4., 5., 6.])
    got = capi.csr_apply(2, 3, A, C, V, np.array([13., 17., 19.]))
    assert got.tolist() == [104.0, 251.0]


def test_identity_kat():
    """tests/sparse/sparsematrix.cpp:13-28: identity action is exact."""
    n = 5
    A = np.arange(n + 1, dtype=np.int32)
    C = np.arange(n, dtype=np.int32)
    V = np.ones(n)
    vec = np.arange(n) * 1.2345
    assert np.array_equal(capi.csr_apply(n, n, A, C, V, vec), vec)


@pytest.mark.parametrize("variant", ["ssor", "rainbow", "ssor_eisenstat", "eisenstat_rainbow"])
def test_diagonal_matrix_kat(variant):
    """A = D: the SSOR preconditioner is omega(2-omega) D^-1, so K = 1 and x = b/d."""
    n = 50
    d = np.linspace(1.0, 3.0, n)
    A = np.arange(n + 1, dtype=np.int32)
    C = np.arange(n, dtype=np.int32)
    system = S.CSRSystem("diag", n, A, C, d.copy(), np.zeros(n, dtype=bool))
    b = S.uniform_pm1(9, n)
    io, ig = _check_parity(variant, system, b, omega=1.3)
    assert ig["iterations"] == 1


def test_zero_rhs_and_single_row():
    system = S.poisson2d_p1(3)
    b = np.zeros(system.n)
    for variant in oracle.VARIANTS:
        io, ig = _check_parity(variant, system, b)
    one = S.CSRSystem("one", 1, np.array([0, 1], dtype=np.int32), np.array([0], dtype=np.int32),
                      np.array([2.0]), np.zeros(1, dtype=bool))
    for variant in oracle.VARIANTS:
        _check_parity(variant, one, np.array([0.7]))


def test_breakdown_on_nan_rhs():
    system = S.fd_laplacian_2d(6)
    b = S.uniform_pm1(1, system.n)
    b[5] = np.nan
    for variant in oracle.VARIANTS:
        io, ig = _check_parity(variant, system, b)
        assert ig["exit_reason"] == capi.EXIT_BREAKDOWN


def test_indefinite_breakdown():
    """d.Ad < 0 on a negative definite matrix ends the loop with BREAKDOWN."""
    system = S.fd_laplacian_2d(5)
    neg = S.CSRSystem("neg", system.n, system.A, system.C, -system.V, system.dirichlet)
    b = S.uniform_pm1(2, system.n)
    io, ig = _check_parity("cg", neg, b)
    assert ig["exit_reason"] == capi.EXIT_BREAKDOWN


def _ill_conditioned(n, seed):
    """Dense SPD matrix with eigenvalues log-spaced over 12 decades, as CSR."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    M = (Q * np.logspace(0, 12, n)) @ Q.T
    M = 0.5 * (M + M.T)
    A = np.arange(0, n * n + 1, n, dtype=np.int32)
    C = np.tile(np.arange(n, dtype=np.int32), n)
    return S.CSRSystem("illcond", n, A, C, M.ravel().copy(), np.zeros(n, dtype=bool))


def test_iteration_cap():
    """CG on an ill-conditioned system with a large right-hand side runs into
    K = N (the loop bound solver/sparsesolver.cpp:75) on both sides."""
    system = _ill_conditioned(40, 4)
    b = 1e10 * S.uniform_pm1(3, system.n)
    xo, ro, io = oracle.solve("cg", system.A, system.C, system.V, b, precision=1e-300)
    xg, rg, ig = _gpu_solve("cg", system, b, None, 1.0, None, precision=1e-300)
    assert io["iterations"] == system.n and io["exit_reason"] == capi.EXIT_CAP
    assert ig["iterations"] == system.n and ig["exit_reason"] == capi.EXIT_CAP
    assert ig["success"] == io["success"] == 0


def test_colouring_is_validated():
    system = S.poisson2d_p1(3)
    nc, F, B, R = oracle.rainbow(system.n, system.A, system.C)
    bad = R.copy()
    bad[0] = bad[1]
    n = system.n
    with pytest.raises(capi.SfaError) as e:
        capi.ConjugateGradientSolverCSR_Rainbow(n, np.zeros(n), np.ones(n), system.A, system.C,
                                               system.V, np.zeros(n), 1e-10, -3,
                                               system.diagonal(), 1.0, nc, F, B, bad)
    assert e.value.code == -5


def test_log_lines(capfd):
    system = S.poisson2d_p1(4)
    b = S.rhs_uniform(system, 1)
    col = oracle.rainbow(system.n, system.A, system.C)
    x, r, info = _gpu_solve("rainbow", system, b, system.diagonal(), 1.2, col, print_modulo=5)
    out = capfd.readouterr().out.splitlines()
    K = info["iterations"]
    assert out[0].startswith(f"(0/{system.n})     BEGIN: Conjugate Gradient CSR (SSOR-Rainbow)")
    assert out[1].startswith(f"(0/{system.n})   RESTART: Residual norm is ")
    interim = [ln for ln in out if "INTERIM" in ln]
    assert len(interim) == K // 5 + 1
    assert out[-1].startswith(f"({K}/{system.n})   SUCCESS: Residual norm is ")


def test_device_resident_system_matches_host_call():
    from simplefem_amd.hipmem import DeviceArray

    system = S.poisson3d_p1_kuhn(10)
    b = S.rhs_uniform(system, 4)
    d = system.diagonal()
    col = oracle.rainbow(system.n, system.A, system.C)
    xh, rh, ih = _gpu_solve("rainbow", system, b, d, 1.0, col)
    with capi.System(capi.METHOD_RAINBOW, system.n, system.A, system.C, system.V, d, 1.0,
                     *col) as sysd:
        bt = DeviceArray.from_host(b)
        x = DeviceArray(system.n)
        r = DeviceArray(system.n)
        info = {}
        for _ in range(2):  # a second solve on the same system replays the same graph
            x.zero()
            r.zero()
            sysd.solve(x, bt, r, 1e-10, info=info)
            assert info["iterations"] == ih["iterations"]
            _close(x.download(), xh, 1e-13)
            _close(r.download(), rh, 1e-13, scale=np.abs(b).max())
        q = DeviceArray(system.n)
        sysd.apply(q, bt)
        want = oracle.csr_apply(system.n, system.A, system.C, system.V, b)
        assert np.array_equal(q.download(), want)
        # the profiled solve gives the same answer as the graph-replayed one
        x.zero()
        r.zero()
        prof, pinfo = sysd.profile(x, bt, r, 1e-10)
        assert pinfo["iterations"] == ih["iterations"]
        _close(x.download(), xh, 1e-13)
        K = pinfo["iterations"]
        st = sysd.stats()
        assert prof["spmv"]["launches"] == K + 1 or prof["spmv"]["launches"] == K
        # every iteration launches kernels_per_iteration kernels (groups merged or
        # fused into one launch count once); the restart adds at most one sweep
        total = sum(prof[k]["launches"] for k in ("spmv", "forward", "backward"))
        per_iter = st["kernels_per_iteration"]
        assert K * per_iter <= total <= (K + 1) * per_iter + st["groups_forward"] + st["groups_backward"]


def test_borrowed_csr_system_matches_copied():
    """A System that borrows the caller's CSR arrays (SFA_FLAG_BORROW_CSR)
    solves, solves mixed and applies bit for bit as one that copies them;
    the mixed System's FP32 copy shares the FP64 system's structure."""
    system = S.poisson3d_p1_kuhn(9)
    b = S.rhs_uniform(system, 12)
    d = system.diagonal()
    col = oracle.rainbow(system.n, system.A, system.C)
    out = []
    for borrow in (False, True):
        A, C, V = system.A.copy(), system.C.copy(), system.V.copy()
        print(f"case: borrow_csr={borrow}, n={system.n}", flush=True)
        with capi.System(capi.METHOD_RAINBOW, system.n, A, C, V, d, 1.0, *col, fp32_inner=True,
                         borrow_csr=borrow) as sysd:
            x, r, info = np.zeros(system.n), np.zeros(system.n), {}
            sysd.solve(x, b, r, 1e-10, info=info)
            xm = np.zeros(system.n)
            mi = sysd.solve_mixed(xm, b, 1e-10)
            q = np.zeros(system.n)
            sysd.apply(q, b)
            st = sysd.stats()
            assert len(st["setup_ms"]) == 4 and min(st["setup_ms"]) >= 0
            assert mi["success"]
            out.append((x, r, info["iterations"], xm, mi["inner_iterations"], q))
    for mine, theirs in zip(*out):
        assert np.array_equal(np.asarray(mine), np.asarray(theirs))
    assert np.array_equal(out[1][5], oracle.csr_apply(system.n, system.A, system.C, system.V, b))


@pytest.mark.parametrize("numbering", ["reference", "lexicographic"])
def test_full_size_3d_rainbow(numbering):
    """The 64^3 Kuhn P1 Poisson system (250,047 active rows), rtol 1e-10,
    omega 1, in the reference's vertex numbering (its CSR structure and
    10-colour Rainbow colouring) and in x-fastest numbering (9 colours)."""
    system = S.poisson3d_p1_kuhn(64, numbering=numbering)
    b = S.rhs_uniform(system, seed=2026)
    io, ig = _check_parity("rainbow", system, b, omega=1.0)
    print("[full]", numbering, "K oracle", io["iterations"], "K gpu", ig["iterations"], flush=True)


@pytest.mark.parametrize("numbering", ["reference", "lexicographic"])
@pytest.mark.parametrize("omega", [1.0, 1.2])
def test_full_size_2d_ssor(omega, numbering):
    """The 129^2 P1 Poisson system (16,129 active rows) with natural-order SSOR,
    the reference's own CPU configuration: ~250 dependency levels per sweep."""
    system = S.poisson2d_p1(7, numbering=numbering)
    b = S.rhs_uniform(system, seed=2026)
    io, ig = _check_parity("ssor", system, b, omega=omega)
    print("[full 2d]", numbering, "omega", omega, "K oracle", io["iterations"], "K gpu",
          ig["iterations"], flush=True)


def test_cpp_dropin_callers():
    """A C++ caller of the drop-in headers, in the shape of the reference's
    tests/solver/solvespd_csr.cpp: every variant converges on the GPU."""
    import os
    import subprocess

    binary = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "simplefem_amd", "_build", "dropin_test")
    proc = subprocess.run([binary, "gpu"], capture_output=True, text=True, timeout=240)
    print(proc.stdout[-3000:])
    assert proc.returncode == 0, proc.stdout[-3000:] + proc.stderr[-2000:]


@pytest.mark.parametrize("inner_restarts", [False, True], ids=["continued", "restarts"])
@pytest.mark.parametrize("variant,method", [("rainbow", capi.METHOD_RAINBOW),
                                            ("ssor", capi.METHOD_SSOR)])
def test_mixed_precision_refinement(variant, method, inner_restarts):
    """FP32 inner solves + FP64 refinement reach ||b-Ax|| <= 1e-10 ||b|| in FP64
    and agree with the FP64 solve of the same system to the accuracy that
    residual allows (relative error bound 1e-10 * condition estimate); inner
    solves that continue the previous direction and inner solves that
    restart alike."""
    system = S.poisson3d_p1_kuhn(16)
    b = S.rhs_uniform(system, 8)
    d = system.diagonal()
    col = oracle.rainbow(system.n, system.A, system.C)
    colour_args = col if variant == "rainbow" else (0, None, None, None)
    with capi.System(method, system.n, system.A, system.C, system.V, d, 1.0, *colour_args,
                     fp32_inner=True, inner_restarts=inner_restarts) as sysm:
        x = np.zeros(system.n)
        info = sysm.solve_mixed(x, b, 1e-10, inner_rtol=1e-4)
        assert info["success"] == 1 and info["relative_residual"] <= 1e-10, info
        assert info["outer"] >= 1
        res = b - oracle.csr_apply(system.n, system.A, system.C, system.V, x)
        res[system.dirichlet] = 0.0
        assert np.linalg.norm(res) <= 1.0000001e-10 * np.linalg.norm(b)
        assert np.array_equal(x[system.dirichlet], np.zeros(system.dirichlet.sum()))
        x64, r64, i64 = oracle.solve(variant, system.A, system.C, system.V, b, aux=d,
                                     colouring=col, precision=1e-12)
        err = np.abs(x - x64).max() / np.abs(x64).max()
        assert err <= 1e-8, err
        print(f"[mixed] {variant}: outer {info['outer']} inner {info['inner_iterations']} "
              f"rel.res {info['relative_residual']:.2e} err vs FP64 {err:.2e}")


@pytest.mark.parametrize("variant,method", [("rainbow", capi.METHOD_RAINBOW),
                                            ("cg", capi.METHOD_CG)])
def test_mixed_continued_direction_keeps_the_krylov_space(variant, method):
    """With many corrections (inner_rtol 1e-2), inner solves that continue the
    previous one's direction need far fewer FP32 iterations than inner solves
    that restart, and about as many as one FP64 solve of the same system;
    both reach the FP64 residual target."""
    system = S.poisson3d_p1_kuhn(16)
    b = S.rhs_uniform(system, 9)
    d = system.diagonal()
    col = oracle.rainbow(system.n, system.A, system.C) if variant == "rainbow" else (0, None, None, None)
    aux = d if variant == "rainbow" else None
    runs = {}
    for restarts in (False, True):
        with capi.System(method, system.n, system.A, system.C, system.V, aux, 1.0, *col,
                         fp32_inner=True, inner_restarts=restarts) as sysm:
            x = np.zeros(system.n)
            info = sysm.solve_mixed(x, b, 1e-10, inner_rtol=1e-2)
            assert info["success"] == 1 and info["relative_residual"] <= 1e-10, (restarts, info)
            res = b - oracle.csr_apply(system.n, system.A, system.C, system.V, x)
            res[system.dirichlet] = 0.0
            assert np.linalg.norm(res) <= 1.0000001e-10 * np.linalg.norm(b)
            runs[restarts] = info
    _, _, i64 = oracle.solve(variant, system.A, system.C, system.V, b, aux=aux,
                             colouring=col if variant == "rainbow" else None)
    cont, rest = runs[False], runs[True]
    print(f"[continued] {variant}: continued {cont['outer']} outer / {cont['inner_iterations']} "
          f"inner; restarts {rest['outer']} / {rest['inner_iterations']}; FP64 K "
          f"{i64['iterations']}", flush=True)
    assert cont["outer"] >= 3 and cont["continued"] >= 1 and rest["continued"] == 0
    assert cont["inner_iterations"] < 0.8 * rest["inner_iterations"]
    assert cont["inner_iterations"] <= 1.3 * i64["iterations"] + 10


@pytest.mark.parametrize("maker", [lambda: S.poisson3d_p1_kuhn(12), lambda: S.poisson2d_p1(6),
                                   lambda: S.hcurl3d_whitney_kuhn(5),
                                   lambda: S.lagrange_pr(2, 6, 3)],
                         ids=["kuhn12", "p1_2d_L6", "hcurl5", "p3_2d"])
def test_paired_sweeps_are_bit_identical(maker):
    """Two colours per launch (the second forms the first's values it reads)
    give the same x, residual and iteration count, bit for bit, as one colour
    per launch, with fewer launches per iteration; mixed precision too."""
    system = maker()
    b = S.rhs_uniform(system, 21)
    d = system.diagonal()
    col = oracle.rainbow(system.n, system.A, system.C)
    out = []
    for paired in (True, False):
        with capi.System(capi.METHOD_RAINBOW, system.n, system.A, system.C, system.V, d, 1.0,
                         *col, fp32_inner=True, paired_sweeps=paired) as sysd:
            x = np.zeros(system.n)
            r = np.zeros(system.n)
            info = {}
            sysd.solve(x, b, r, 1e-10, info=info)
            xm = np.zeros(system.n)
            mi = sysd.solve_mixed(xm, b, 1e-10)
            out.append((x, r, info["iterations"], xm, mi["inner_iterations"],
                        sysd.stats()["kernels_per_iteration"]))
    (xp, rp, kp, xmp, kmp, lp), (xs, rs, ks, xms, kms, ls) = out
    print(f"[pairs] {system.name}: K={kp} launches/iteration {lp} paired vs {ls} single", flush=True)
    assert kp == ks and np.array_equal(xp, xs) and np.array_equal(rp, rs)
    assert kmp == kms and np.array_equal(xmp, xms)
    assert lp <= ls
    xo, _, io = oracle.solve("rainbow", system.A, system.C, system.V, b, aux=d, colouring=col)
    assert abs(kp - io["iterations"]) <= max(1, int(0.02 * io["iterations"]))
    _close(xp, xo, 1e-10)


def test_sweep_timing_leaves_later_solves_intact():
    """sfa_system_time_sweeps times the forward and backward sweep launches in
    a recorded graph (a few microseconds each); the next solve restarts and
    still matches the oracle; plain CG systems have no sweeps to time."""
    s = S.poisson3d_p1_kuhn(12)
    b = S.rhs_uniform(s, 4)
    d = s.diagonal()
    col = oracle.rainbow(s.n, s.A, s.C)
    with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0, *col) as sysd:
        x = np.zeros(s.n)
        sysd.solve(x, b, np.zeros(s.n), 1e-10)
        for klass in ("forward", "backward"):
            us = sysd.time_sweeps(klass, 50)
            print(f"[sweep timing] {klass} {us:.2f} us per launch", flush=True)
            assert 0.5 < us < 200.0, (klass, us)
        x = np.zeros(s.n)
        info = {}
        sysd.solve(x, b, np.zeros(s.n), 1e-10, info=info)
        _check_against_oracle = oracle.solve("rainbow", s.A, s.C, s.V, b, aux=d, colouring=col)
        xo, _, io = _check_against_oracle
        assert abs(info["iterations"] - io["iterations"]) <= 1, (info, io)
        assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
    sr = S.random_spd(100, 0.1, 2)
    with capi.System(capi.METHOD_CG, sr.n, sr.A, sr.C, sr.V) as plain:
        plain.solve(np.zeros(sr.n), S.uniform_pm1(1, sr.n), np.zeros(sr.n), 1e-10)
        with pytest.raises(capi.SfaError):
            plain.time_sweeps("forward", 10)


def test_mixed_refinement_edge_cases():
    """max_outer = 0 returns x0 untouched with its residual; b = 0 needs no
    correction; a nonzero x0 close to the solution needs fewer corrections."""
    s = S.poisson3d_p1_kuhn(10)
    d = s.diagonal()
    col = oracle.rainbow(s.n, s.A, s.C)
    b = S.rhs_uniform(s, 12)
    with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0, *col,
                     fp32_inner=True) as sysm:
        x = np.zeros(s.n)
        info = sysm.solve_mixed(x, b, 1e-10, max_outer=0)
        assert info["outer"] == 0 and info["inner_iterations"] == 0 and not x.any()
        assert info["success"] == 0
        x = np.zeros(s.n)
        info = sysm.solve_mixed(x, np.zeros(s.n), 1e-10)
        assert info["outer"] == 0 and info["success"] == 1 and not x.any()
        x = np.zeros(s.n)
        full = sysm.solve_mixed(x, b, 1e-10)
        x2 = x + 1e-7 * S.uniform_pm1(3, s.n) * np.abs(x).max()
        x2[s.dirichlet] = 0.0
        near = sysm.solve_mixed(x2, b, 1e-10)
        assert near["success"] == 1 and near["outer"] < full["outer"], (near, full)


def test_repeated_dropin_calls_reuse_the_plan_and_see_changes():
    """Drop-in calls on host arrays keep the last few Systems: a repeated call
    with the same arrays reuses one (same bits as the first call and the
    oracle's bar), a call whose values changed in place -- the same arrays,
    new contents -- is solved with the new values, and calls that alternate
    between matrices give each matrix's own result."""
    s1, s2 = S.poisson3d_p1_kuhn(8), S.poisson2d_p1(5)
    b1, b2 = S.rhs_uniform(s1, 3), S.rhs_uniform(s2, 4)
    first = _gpu_solve("rainbow", s1, b1, s1.diagonal(), 1.0,
                       oracle.rainbow(s1.n, s1.A, s1.C))
    for _ in range(2):
        again = _gpu_solve("rainbow", s1, b1, s1.diagonal(), 1.0,
                           oracle.rainbow(s1.n, s1.A, s1.C))
        assert again[2]["iterations"] == first[2]["iterations"]
        assert np.array_equal(again[0], first[0]) and np.array_equal(again[1], first[1])
        _check_parity("ssor", s2, b2)  # another matrix in between
    _check_parity("rainbow", s1, b1)
    # the same arrays, scaled in place: x halves
    V = s1.V
    V *= 2.0
    d = s1.diagonal()
    x2, _, i2 = _gpu_solve("rainbow", s1, b1, d, 1.0, oracle.rainbow(s1.n, s1.A, s1.C))
    V *= 0.5
    xo, _, io = oracle.solve("rainbow", s1.A, s1.C, s1.V * 2.0, b1, aux=d, omega=1.0,
                             colouring=oracle.rainbow(s1.n, s1.A, s1.C))
    assert abs(i2["iterations"] - io["iterations"]) <= 1
    _close(x2, xo)


def test_dropin_sees_a_change_in_the_last_entries_of_large_arrays():
    """The drop-in's comparison of large arrays with a kept plan (above 4 MB
    it compares chunks on several host threads) finds a single changed value
    near the end: the diagonal of the last active row of the 48^3 system,
    scaled in place, gives the modified matrix's solution."""
    s = S.poisson3d_p1_kuhn(48)  # 117,649 rows, ~1.7 M entries: 13.7 MB of values
    b = S.rhs_uniform(s, 21)
    col = oracle.rainbow(s.n, s.A, s.C)
    print(f"case: drop-in change detection, n={s.n} nnz={s.nnz}", flush=True)
    x1, _, i1 = _gpu_solve("rainbow", s, b, s.diagonal(), 1.0, col)
    r = int(np.nonzero(s.diagonal())[0][-1])
    k = int(s.A[r] + np.nonzero(s.C[s.A[r]:s.A[r + 1]] == r)[0][0])
    assert k > 0.98 * s.nnz  # the change lies among the last chunks compared
    s.V[k] *= 1.5
    d = s.diagonal()
    x2, _, i2 = _gpu_solve("rainbow", s, b, d, 1.0, col)
    xo, _, io = oracle.solve("rainbow", s.A, s.C, s.V, b, aux=d, omega=1.0, colouring=col)
    s.V[k] /= 1.5
    assert not np.array_equal(x1, x2)
    assert abs(i2["iterations"] - io["iterations"]) <= 1
    _close(x2, xo)
