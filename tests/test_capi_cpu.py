"""C-ABI library checks that need no GPU: it loads, exports every declared
symbol, and its host-side logic (greedy colouring, diagonal extraction,
argument validation) agrees with the oracle.  No kernel is launched here."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from simplefem_amd import capi
from simplefem_amd import systems as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_library_exports_every_header_symbol():
    lib = capi.load()
    names = capi.header_symbols()
    assert len(names) >= 17
    for name in names:
        assert hasattr(lib, name), name
    assert lib.sfa_api_version() == capi.API_VERSION == 2


def test_python_signatures_cover_header():
    assert set(capi.header_symbols()) == set(capi._SIGNATURES)


@pytest.mark.parametrize("make", [lambda: S.poisson2d_p1(7), lambda: S.poisson3d_p1_kuhn(64),
                                  lambda: S.random_spd(500, 0.01, 2), lambda: S.fd_laplacian_2d(20)])
def test_rainbow_bit_exact(make):
    s = make()
    want = oracle.rainbow(s.n, s.A, s.C)
    got = capi.Rainbow(s.n, s.A, s.C)
    assert got.num_colors == want[0]
    assert np.array_equal(got.F, want[1])
    assert np.array_equal(got.B, want[2])
    assert np.array_equal(got.R, want[3])


@pytest.mark.parametrize("numbering", ["lexicographic", "reference"])
def test_rainbow_matches_reference_digest(numbering):
    name = "kuhn_64" + ("_refnum" if numbering == "reference" else "")
    meta = json.load(open(os.path.join(GOLDEN, "golden.json")))["full"][name]
    s = S.poisson3d_p1_kuhn(64, numbering=numbering)
    rb = capi.Rainbow(s.n, s.A, s.C)
    import hashlib
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()
    assert rb.num_colors == meta["num_colors"]
    assert (sha(rb.F), sha(rb.B), sha(rb.R)) == (meta["sha256_F"], meta["sha256_B"],
                                                 meta["sha256_R"])


def test_get_diagonal():
    for s in (S.poisson3d_p1_kuhn(5), S.random_spd(200, 0.05, 1)):
        assert np.array_equal(capi.get_diagonal(s.n, s.A, s.C, s.V),
                              oracle.csr_diagonal(s.n, s.A, s.C, s.V))


def test_invalid_csr_is_rejected_before_the_device():
    n = 4
    A = np.array([0, 2, 1, 3, 4], dtype=np.int32)  # decreasing offsets
    C = np.array([0, 1, 2, 3], dtype=np.int32)
    V = np.ones(4)
    with pytest.raises(capi.SfaError) as e:
        capi.ConjugateGradientSolverCSR(n, np.zeros(n), np.ones(n), A, C, V, np.zeros(n), 1e-10, -3)
    assert e.value.code == -1
    A = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    C = np.array([0, 1, 2, 7], dtype=np.int32)  # column out of range
    with pytest.raises(capi.SfaError) as e:
        capi.ConjugateGradientSolverCSR_SSOR(n, np.zeros(n), np.ones(n), A, C, V, np.zeros(n),
                                             1e-10, -3, np.ones(n), 1.0)
    assert e.value.code == -1
    assert "column" in str(e.value)


def test_invalid_blocks_are_rejected_before_the_device():
    """sfa_HodgeConjugateResidualSolverCSR_SSOR checks all four blocks on the
    host: B's columns must lie in [0, L), Bt's in [0, N)."""
    bs = S.schur_blocks(8, 2, 0.0, 1)
    bad_B = (bs.B[0], bs.B[1].copy(), bs.B[2])
    bad_B[1][3] = bs.L + 5
    for B, Bt, word in ((bad_B, bs.Bt, "B"), (bs.B, (bs.Bt[0], bs.Bt[1] + bs.N, bs.Bt[2]), "Bt")):
        with pytest.raises(capi.SfaError) as e:
            capi.HodgeConjugateResidualSolverCSR_SSOR(bs.N, bs.L, np.zeros(bs.N), np.ones(bs.N),
                                                      bs.A, B, Bt, bs.C, np.zeros(bs.N), 1e-10,
                                                      -3, 1e-10, -3)
        assert e.value.code == -1 and f": {word}: column" in str(e.value), str(e.value)
    # the plain and diagonal variants run the same host checks
    for fn in (capi.HodgeConjugateResidualSolverCSR, capi.HodgeConjugateResidualSolverCSR_diagonal,
               capi.HodgeConjugateResidualSolverCSR_textbook):
        with pytest.raises(capi.SfaError) as e:
            fn(bs.N, bs.L, np.zeros(bs.N), np.ones(bs.N), bs.A, bad_B, bs.Bt, bs.C,
               np.zeros(bs.N), 1e-10, -3, 1e-10, -3)
        assert e.value.code == -1 and ": B: column" in str(e.value), str(e.value)


def test_unknown_hodge_inner_solver_is_rejected_before_the_device():
    bs = S.schur_blocks(8, 2, 0.0, 1)
    with pytest.raises(ValueError):
        capi.Hodge(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C, inner="jacobi")
    mats = capi._blocks(bs.N, bs.L, bs.A, bs.B, bs.Bt, bs.C)
    ptrs = [(capi._i if k % 3 < 2 else capi._d)(m) for k, m in enumerate(mats)]
    h = ctypes.c_void_p()
    rc = capi.load().sfa_hodge_create_inner(bs.N, bs.L, *ptrs, 7, ctypes.byref(h))
    assert rc == -1 and not h.value
    assert "inner must be" in capi.load().sfa_last_error().decode()


def test_inconsistent_colouring_is_rejected_before_the_device():
    s = S.poisson2d_p1(3)
    nc, F, B, R = oracle.rainbow(s.n, s.A, s.C)
    F2 = F.copy()
    F2[3] = (F2[3] + 1) % nc
    n = s.n
    with pytest.raises(capi.SfaError) as e:
        capi.ConjugateGradientSolverCSR_Rainbow(n, np.zeros(n), np.ones(n), s.A, s.C, s.V,
                                               np.zeros(n), 1e-10, -3, s.diagonal(), 1.0, nc, F2,
                                               B, R)
    assert e.value.code == -5


def test_large_inputs_are_checked_with_the_same_messages():
    """At 2^18 rows and more the host checks run a parallel pass first and,
    when it finds a problem, the serial pass for its message: the first
    problem in row order, as below that size."""
    s = S.poisson3d_p1_kuhn(64)  # 274,625 rows
    n = s.n
    rb = capi.Rainbow(n, s.A, s.C)
    d = s.diagonal()

    def call(A=s.A, C=s.C, F=rb.F, R=rb.R):
        return capi.ConjugateGradientSolverCSR_Rainbow(n, np.zeros(n), np.ones(n), A, C, s.V,
                                                      np.zeros(n), 1e-10, -3, d, 1.0,
                                                      rb.num_colors, F, rb.B, R)

    c0, c1 = int(rb.B[3]), int(rb.B[4])  # colour 3's positions in R
    # a row listed twice early in colour 3, and a wrong colour for a row later on
    R2 = rb.R.copy()
    R2[c0 + 10] = R2[c0 + 9]
    F2 = rb.F.copy()
    r_late = int(rb.R[c1 - 5])
    F2[r_late] = (F2[r_late] + 1) % rb.num_colors
    for F, R, code, text in ((F2, R2, -5, "R must be a permutation of the rows"),
                             (F2, rb.R, -5, "F disagrees with B/R")):
        with pytest.raises(capi.SfaError) as e:
            call(F=F, R=R)
        assert e.value.code == code and text in str(e.value), str(e.value)
    C2 = s.C.copy()
    C2[-1] = n
    with pytest.raises(capi.SfaError) as e:
        call(C=C2)
    assert e.value.code == -1 and "column index out of range" in str(e.value)
    A2 = s.A.copy()
    A2[n - 7] = A2[n - 6] + 1
    with pytest.raises(capi.SfaError) as e:
        call(A=A2)
    assert e.value.code == -1 and "csrrows must be non-decreasing" in str(e.value)
    # valid arrays pass every check (here without a GPU the call then stops there)
    try:
        call()
    except capi.SfaError as e:
        assert e.code == -3, str(e)


def test_null_arguments():
    lib = capi.load()
    info = capi.SolveInfo()
    rc = lib.sfa_ConjugateGradientSolverCSR(4, None, None, None, None, None, None, 1e-10, -3,
                                            ctypes.byref(info))
    assert rc == -1
    assert lib.sfa_last_error().decode()
    assert lib.sfa_system_destroy(None) == 0
    assert lib.sfa_system_query(None, None) == -1
