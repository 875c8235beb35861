"""CSR products and transpose on the GPU (sfa_MatrixCSRMultiplication,
sfa_MatrixCSRMultiplication_reduced, sfa_MatrixCSR_getTranspose) against the
compiled reference's MatrixCSRMultiplication, MatrixCSRMultiplication_reduced
and MatrixCSR::getTranspose (sparse/matcsr.cpp), bit for bit: the A, C and V
arrays are equal.  The inputs hold what makes the
reference's compression order matter -- repeated columns within a row,
unsorted rows, exact zeros, small integer values whose products cancel to
exactly zero (dropped), irregular values whose duplicate sums round by
order -- and empty rows.
"""
import time

import numpy as np
import pytest

import oracle
from simplefem_amd import capi
from simplefem_amd import systems as S

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.ref_available(),
                                 reason="the reference library (oracle/_ref) is not built")]


def _random_csr(rng, rows, cols, per_row, unsorted):
    """Rows of up to per_row entries (some empty); unsorted rows repeat
    columns; values irregular, small integers (products cancel) or 0."""
    A, C, V = [0], [], []
    for _ in range(rows):
        w = int(rng.integers(0, per_row + 1))
        if unsorted:
            cs = rng.integers(0, cols, w)
        else:
            cs = np.sort(rng.choice(cols, min(w, cols), replace=False))
        vs = rng.standard_normal(len(cs))
        small = rng.random(len(cs)) < 0.3
        vs[small] = rng.integers(-2, 3, int(small.sum()))
        vs[rng.random(len(cs)) < 0.1] = 0.0
        C += list(cs)
        V += list(vs)
        A.append(len(C))
    return rows, cols, np.array(A, np.int32), np.array(C, np.int32), np.array(V)


def _same(mine, theirs, what):
    assert mine[0] == theirs[0] and mine[1] == theirs[1], what
    for x, y in zip(mine[2:], theirs[2:]):
        assert np.array_equal(np.asarray(x), np.asarray(y)), what


def test_product_and_transpose_match_reference_bitwise():
    from oracle.ref import RefRunner

    rng = np.random.default_rng(11)
    with RefRunner(1) as ref:
        for unsorted in (False, True):
            for rows, inner, cols, per_row in [(1, 1, 1, 3), (30, 25, 20, 6), (200, 150, 180, 12),
                                               (500, 40, 700, 9)]:
                m1 = _random_csr(rng, rows, inner, per_row, unsorted)
                m2 = _random_csr(rng, inner, cols, per_row, unsorted)
                what = ("mul", rows, inner, cols, unsorted)
                print(f"case: {what}", flush=True)
                _same(capi.csr_multiply(m1, m2), ref.csr_op("mul", m1, m2), what)
                what = ("mul_reduced", rows, inner, cols, unsorted)
                _same(capi.csr_multiply(m1, m2, reduced=True), ref.csr_op("mul_reduced", m1, m2),
                      what)
                what = ("transpose", rows, inner, unsorted)
                _same(capi.csr_transpose(m1), ref.csr_op("transpose", m1), what)
        # one long row among short ones: too uneven to share a wavefront's
        # interleaved scratch, so its rows are compressed in place
        cols_l = rng.integers(0, 150, 600)
        A_l = np.concatenate([[0], 600 + np.arange(200)]).astype(np.int32)
        C_l = np.concatenate([cols_l, rng.integers(0, 150, 199)]).astype(np.int32)
        V_l = rng.standard_normal(C_l.size)
        V_l[rng.random(C_l.size) < 0.3] = rng.integers(-2, 3, 1)[0]
        uneven = (200, 150, A_l, C_l, V_l)
        m2 = _random_csr(rng, 150, 180, 12, True)
        for reduced, op in ((False, "mul"), (True, "mul_reduced")):
            _same(capi.csr_multiply(uneven, m2, reduced=reduced), ref.csr_op(op, uneven, m2),
                  ("uneven rows", op))
        _same(capi.csr_transpose(capi.csr_multiply(uneven, m2)),
              ref.csr_op("transpose", ref.csr_op("mul", uneven, m2)), "uneven rows, transpose")
        # no entries at all, and a product with no surviving entry
        empty = (4, 3, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0))
        m = _random_csr(rng, 3, 5, 4, True)
        _same(capi.csr_multiply(empty, m), ref.csr_op("mul", empty, m), "empty product")
        _same(capi.csr_multiply(empty, m, reduced=True), ref.csr_op("mul_reduced", empty, m),
              "empty reduced product")
        _same(capi.csr_transpose(empty), ref.csr_op("transpose", empty), "empty transpose")
        zero = (3, 5, m[2], m[3], np.zeros_like(m[4]))
        m2 = _random_csr(rng, 5, 4, 3, False)
        _same(capi.csr_multiply(zero, m2), ref.csr_op("mul", zero, m2), "all-zero values")
        _same(capi.csr_multiply(zero, m2, reduced=True), ref.csr_op("mul_reduced", zero, m2),
              "all-zero values, reduced")


def test_stiffness_products_match_reference_and_time():
    """Products and transposes of the P1 stiffness on Kuhn cubes (up to 32^3
    cubes: 35,937 rows, ~15 entries per row, ~225 products per output row
    before compression); prints the GPU and reference times."""
    from oracle.ref import RefRunner

    with RefRunner(1) as ref:
        for n in (2, 8, 16, 32):
            s = S.poisson3d_p1_kuhn(n)
            m = (s.n, s.n, s.A, s.C, s.V)
            t0 = time.perf_counter()
            mine = capi.csr_multiply(m, m)
            t1 = time.perf_counter()
            theirs = ref.csr_op("mul", m, m)
            t2 = time.perf_counter()
            print(f"stiffness 3D n={n}: {s.n} rows, product {len(mine[3])} entries; GPU call "
                  f"{(t1 - t0) * 1e3:.1f} ms, reference (one thread, with its process round "
                  f"trip) {(t2 - t1) * 1e3:.1f} ms", flush=True)
            _same(mine, theirs, ("stiffness product", n))
            _same(capi.csr_transpose(m), ref.csr_op("transpose", m), ("stiffness transpose", n))
            if n <= 16:
                _same(capi.csr_multiply(m, m, reduced=True), ref.csr_op("mul_reduced", m, m),
                      ("stiffness reduced product", n))


def test_dropin_product_takes_the_gpu_for_large_products():
    """The C++ drop-in's MatrixCSRMultiplication and
    MatrixCSRMultiplication_reduced hand products of at least 2^19 pairs to
    the GPU: on the 16^3 stiffness squared (~1.1 M products) their arrays are
    still the reference's, bit for bit, as they are below the threshold
    (8^3)."""
    from oracle.ref import RefRunner
    from test_dropin_cpu import dropin_op

    with RefRunner(1) as ref:
        for n in (8, 16):
            s = S.poisson3d_p1_kuhn(n)
            m = (s.n, s.n, s.A, s.C, s.V)
            products = int(sum(np.diff(s.A)[s.C]))
            print(f"case: drop-in product 3D n={n}, {products} products", flush=True)
            _same(dropin_op("mul", m, m), ref.csr_op("mul", m, m), ("drop-in product", n))
            _same(dropin_op("mul_reduced", m, m), ref.csr_op("mul_reduced", m, m),
                  ("drop-in reduced product", n))
