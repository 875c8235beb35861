"""HodgeConjugateResidualSolverCSR_SSOR on the GPU (sfa_HodgeConjugateResidualSolverCSR_SSOR,
sfa_hodge_*) against the oracle.

Bars: K and the exit reason equal to the oracle's; x within X_REL of the
oracle's, relative (inf-norm): the outer dot products are summed in block
order here and in row order there, and so are the inner solves' dot products
(their tolerances are summed in row order on both sides), so the iterates
differ in the last digits of every inner solve, which stop at 1e-12 relative;
the Schur residual of the GPU x within 2e-6 relative.  The handle API gives
the same x from host and device buffers and on repeated solves.
"""
import numpy as np
import pytest

import oracle
from simplefem_amd import capi
from simplefem_amd import systems as S
from simplefem_amd.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

CASES = [(16, 2, 0.0, 3), (16, 2, 0.5, 3), (32, 2, 0.0, 3), (40, 2, 1.0, 3), (64, 4, 0.1, 7)]
X_REL = 1e-12  # measured: at most 2.1e-14 over every case


def _schur(bs):
    A, B, Bt, C = bs.dense()
    return B @ np.linalg.solve(A, Bt) + C


def _check(bs, b, x, info, xo, io):
    assert info["iterations"] == io["iterations"], (bs.name, info, io)
    assert info["exit_reason"] == io["exit_reason"], (bs.name, info, io)
    err = np.abs(x - xo).max() / np.abs(xo).max()
    print(f"[hodge] {bs.name}: K {info['iterations']} x rel. difference {err:.2e}", flush=True)
    assert err <= X_REL, bs.name
    schur = _schur(bs)
    assert np.linalg.norm(schur @ x - b) <= 2e-6 * np.linalg.norm(b), bs.name


@pytest.mark.parametrize("case", CASES, ids=[f"m{c[0]}_c{c[1]}_w{c[2]}" for c in CASES])
def test_hodge_one_shot_matches_oracle(case):
    bs = S.schur_blocks(*case)
    b = S.uniform_pm1(5, bs.N)
    print(f"[hodge] {bs.name} N={bs.N} L={bs.L}", flush=True)
    xo, ro, io = oracle.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b, precision=1e-10,
                                      inner_precision=1e-12)
    x = np.zeros(bs.N)
    r = np.zeros(bs.N)
    info = {}
    K = capi.HodgeConjugateResidualSolverCSR_SSOR(bs.N, bs.L, x, b, bs.A, bs.B, bs.Bt, bs.C, r,
                                                  1e-10, -3, 1e-12, -3, info=info)
    assert K == info["iterations"]
    print(f"[hodge] K gpu {K} oracle {io['iterations']}", flush=True)
    _check(bs, b, x, info, xo, io)


def test_hodge_handle_host_device_and_repeated_solves():
    bs = S.schur_blocks(32, 2, 0.25, 11)
    b = S.uniform_pm1(8, bs.N)
    xo, _, io = oracle.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b, precision=1e-10,
                                     inner_precision=1e-12)
    with capi.Hodge(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C) as h:
        runs = []
        for _ in range(2):
            x = np.zeros(bs.N)
            r = np.zeros(bs.N)
            info = h.solve(x, b, r, 1e-10, -3, 1e-12, -3)
            _check(bs, b, x, info, xo, io)
            runs.append(x)
        assert np.array_equal(runs[0], runs[1])  # the same work in the same order
        xd, bd, rd = DeviceArray(bs.N), DeviceArray.from_host(b), DeviceArray(bs.N)
        xd.zero()  # the same initial guess as the host solves
        info = h.solve(xd, bd, rd, 1e-10, -3, 1e-12, -3)
        assert np.array_equal(xd.download(), runs[0]), info
        # a second right-hand side on the same resident blocks
        b2 = S.uniform_pm1(9, bs.N)
        x2o, _, i2o = oracle.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b2,
                                           precision=1e-10, inner_precision=1e-12)
        x2 = np.zeros(bs.N)
        info2 = h.solve(x2, b2, np.zeros(bs.N), 1e-10, -3, 1e-12, -3)
        _check(bs, b2, x2, info2, x2o, i2o)


def test_hodge_identity_coupling_gives_a_times_b():
    bs = S.schur_blocks(12, 2, 0.0, 1)
    L = bs.L
    eye = (np.arange(L + 1, dtype=np.int32), np.arange(L, dtype=np.int32), np.ones(L))
    empty = (np.zeros(L + 1, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    b = S.uniform_pm1(2, L)
    x = np.zeros(L)
    capi.HodgeConjugateResidualSolverCSR_SSOR(L, L, x, b, bs.A, eye, eye, empty, np.zeros(L),
                                              1e-12, -3, 1e-14, -3)
    want = oracle.csr_apply(L, bs.A[0], bs.A[1], bs.A[2], b)
    # each inner solve stops at the reference's absolute floor sqrt(z.r) < 2^-26,
    # which bounds the accuracy of every A^-1 application at ~1e-8
    assert np.abs(x - want).max() <= 1e-6 * np.abs(want).max()


def test_hodge_edge_cases_match_oracle():
    """b = 0 (the restart meets the tolerance floor: K = 0, x unchanged), a
    one-unknown Schur complement, and a nonzero initial guess: K, exit and x
    as the oracle's."""
    bs = S.schur_blocks(8, 2, 0.5, 4)
    cases = [("zero rhs", bs, np.zeros(bs.N), np.zeros(bs.N)),
             ("x0", bs, S.uniform_pm1(3, bs.N), S.uniform_pm1(4, bs.N) * 0.1)]
    one = S.schur_blocks(2, 2, 0.0, 1)  # N = 1, L = 4
    cases.append(("N=1", one, np.array([0.7]), np.zeros(1)))
    for name, blk, b, x0 in cases:
        xo, ro, io = oracle.hodge_cr_ssor(blk.N, blk.L, blk.A, blk.B, blk.Bt, blk.C, b, x0=x0,
                                          precision=1e-10, inner_precision=1e-12)
        x = x0.copy()
        info = {}
        K = capi.HodgeConjugateResidualSolverCSR_SSOR(blk.N, blk.L, x, b, blk.A, blk.B, blk.Bt,
                                                      blk.C, np.zeros(blk.N), 1e-10, -3, 1e-12,
                                                      -3, info=info)
        print(f"[hodge edge] {name}: K gpu {K} oracle {io['iterations']}", flush=True)
        assert K == io["iterations"] and info["exit_reason"] == io["exit_reason"], name
        assert np.abs(x - xo).max() <= X_REL * max(np.abs(xo).max(), 1e-300) or \
            np.array_equal(x, xo), name


# HodgeConjugateResidualSolverCSR_diagonal (solver/systemsparsesolver.cpp:52-424, inner
# Jacobi CG), HodgeConjugateResidualSolverCSR_textbook (:798-1121, inner plain CG) and
# HodgeConjugateResidualSolverCSR (:18-49, the SSOR variant), same bars.

@pytest.mark.parametrize("inner", ["diagonal", "textbook"])
@pytest.mark.parametrize("case", CASES, ids=[f"m{c[0]}_c{c[1]}_w{c[2]}" for c in CASES])
def test_hodge_variant_one_shot_matches_oracle(case, inner):
    bs = S.schur_blocks(*case)
    b = S.uniform_pm1(5, bs.N)
    print(f"[hodge {inner}] {bs.name} N={bs.N} L={bs.L}", flush=True)
    xo, ro, io = oracle.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b, precision=1e-10,
                                      inner_precision=1e-12, inner=inner)
    x = np.zeros(bs.N)
    r = np.zeros(bs.N)
    info = {}
    fn = getattr(capi, "HodgeConjugateResidualSolverCSR_" + inner)
    K = fn(bs.N, bs.L, x, b, bs.A, bs.B, bs.Bt, bs.C, r, 1e-10, -3, 1e-12, -3, info=info)
    assert K == info["iterations"]
    print(f"[hodge {inner}] K gpu {K} oracle {io['iterations']}", flush=True)
    _check(bs, b, x, info, xo, io)


def test_hodge_plain_variant_is_the_ssor_one():
    bs = S.schur_blocks(16, 2, 0.5, 3)
    b = S.uniform_pm1(5, bs.N)
    xs, xp = np.zeros(bs.N), np.zeros(bs.N)
    rs, rp = np.zeros(bs.N), np.zeros(bs.N)
    Ks = capi.HodgeConjugateResidualSolverCSR_SSOR(bs.N, bs.L, xs, b, bs.A, bs.B, bs.Bt, bs.C, rs,
                                                   1e-10, -3, 1e-12, -3)
    Kp = capi.HodgeConjugateResidualSolverCSR(bs.N, bs.L, xp, b, bs.A, bs.B, bs.Bt, bs.C, rp,
                                              1e-10, -3, 1e-12, -3)
    assert Ks == Kp and np.array_equal(xs, xp) and np.array_equal(rs, rp)


@pytest.mark.parametrize("inner", ["diagonal", "textbook"])
def test_hodge_variant_handle_matches_one_shot(inner):
    bs = S.schur_blocks(32, 2, 0.25, 11)
    b = S.uniform_pm1(8, bs.N)
    x1, r1 = np.zeros(bs.N), np.zeros(bs.N)
    K1 = getattr(capi, "HodgeConjugateResidualSolverCSR_" + inner)(
        bs.N, bs.L, x1, b, bs.A, bs.B, bs.Bt, bs.C, r1, 1e-10, -3, 1e-12, -3)
    xo, _, io = oracle.hodge_cr_ssor(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, b, precision=1e-10,
                                     inner_precision=1e-12, inner=inner)
    with capi.Hodge(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, inner=inner) as h:
        for _ in range(2):
            x = np.zeros(bs.N)
            r = np.zeros(bs.N)
            info = h.solve(x, b, r, 1e-10, -3, 1e-12, -3)
            assert info["iterations"] == K1 and np.array_equal(x, x1) and np.array_equal(r, r1)
            _check(bs, b, x, info, xo, io)
