"""The C++ drop-in (include/feec/..., libsimplefem_amd_cxx.so): host-side
classes checked by a caller written like the reference's own tests."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "simplefem_amd", "_build", "dropin_test")


def test_dropin_host_classes():
    assert os.path.exists(BIN), "build first: python -m simplefem_amd.build"
    proc = subprocess.run([BIN, "cpu"], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "0 failure(s)" in proc.stdout


# ---------------------------------------------------------------- sparse host operations
import numpy as np
import pytest

import oracle

SPARSE_OPS = os.path.join(ROOT, "simplefem_amd", "_build", "sparse_ops_test")
needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="the reference library (oracle/_ref) is not built")
OPS = {"mul": 1, "mul_reduced": 2, "conjugation": 3, "transpose": 4, "add": 5, "compress": 6}


def _csr_bytes(m):
    rows, cols, A, C, V = m
    return (np.array([rows, cols, len(C)], np.int32).tobytes() + np.asarray(A, np.int32).tobytes()
            + np.asarray(C, np.int32).tobytes() + np.asarray(V, np.float64).tobytes())


def _parse(out):
    rows, cols, nnz = np.frombuffer(out[:12], np.int32)
    A = np.frombuffer(out, np.int32, rows + 1, 12)
    C = np.frombuffer(out, np.int32, nnz, 12 + 4 * (rows + 1))
    V = np.frombuffer(out, np.float64, nnz, 12 + 4 * (rows + 1 + nnz))
    return int(rows), int(cols), A.copy(), C.copy(), V.copy()


def _dropin(payload):
    assert os.path.exists(SPARSE_OPS), "build first: python -m simplefem_amd.build"
    proc = subprocess.run([SPARSE_OPS], input=payload, capture_output=True, timeout=60)
    assert proc.returncode == 0, proc.stderr
    return _parse(proc.stdout)


def dropin_coo_to_csr(rows, cols, r, c, v):
    return _dropin(np.array([0, rows, cols, len(v)], np.int32).tobytes()
                   + np.asarray(r, np.int32).tobytes() + np.asarray(c, np.int32).tobytes()
                   + np.asarray(v, np.float64).tobytes())


def dropin_op(op, m1, m2=None, s1=1.0, s2=1.0):
    m2 = m1 if m2 is None else m2
    return _dropin(np.array([OPS[op]], np.int32).tobytes() + _csr_bytes(m1) + _csr_bytes(m2)
                   + np.array([s1, s2], np.float64).tobytes())


def _dense(m):
    rows, cols, A, C, V = m
    D = np.zeros((rows, cols))
    for r in range(rows):
        for k in range(A[r], A[r + 1]):
            D[r, C[k]] += V[k]
    return D


def _random_coo(rng, rows, cols, n):
    """Entries in random order with repeated (row, column) pairs, exact zeros,
    and small integer values next to irregular ones."""
    r = rng.integers(0, rows, n)
    c = rng.integers(0, cols, n)
    v = rng.standard_normal(n)
    v[rng.random(n) < 0.2] = rng.integers(-3, 4, n)[: 1]  # small integers (sums may cancel)
    v[rng.random(n) < 0.15] = 0.0
    half = n // 8  # repeat some pairs
    r[:half], c[:half] = r[half:2 * half], c[half:2 * half]
    perm = rng.permutation(n)
    return r[perm], c[perm], v[perm]


def _random_csr(rng, rows, cols, per_row, unsorted=False):
    """A CSR matrix, rows possibly unsorted with repeated columns and zeros."""
    A = [0]
    C, V = [], []
    for _ in range(rows):
        w = int(rng.integers(0, per_row + 1))
        cs = rng.integers(0, cols, w) if unsorted else np.sort(rng.choice(cols, min(w, cols), replace=False))
        vs = rng.standard_normal(len(cs))
        vs[rng.random(len(cs)) < 0.1] = 0.0
        C += list(cs)
        V += list(vs)
        A.append(len(C))
    return rows, cols, np.array(A, np.int32), np.array(C, np.int32), np.array(V)


def test_coo_to_csr_closed_form():
    rng = np.random.default_rng(5)
    for rows, cols, n in [(7, 5, 40), (30, 40, 300), (1, 1, 5)]:
        r, c, v = _random_coo(rng, rows, cols, n)
        out = dropin_coo_to_csr(rows, cols, r, c, v)
        R, Cc, A, C, V = out
        assert (R, Cc) == (rows, cols)
        key = np.asarray(r, np.int64) * cols + np.asarray(c)
        got = [(row, int(C[k])) for row in range(rows) for k in range(A[row], A[row + 1])]
        if np.all(np.diff(key) >= 0):  # a sorted list is taken as it is
            assert got == [(int(a), int(b)) for a, b in zip(r, c)]
        else:  # the (row, column) pairs with at least one nonzero entry, once each
            assert got == sorted({(int(a), int(b)) for a, b, x in zip(r, c, v) if x != 0.0})
        D = np.zeros((rows, cols))
        np.add.at(D, (r, c), v)
        assert np.allclose(_dense(out), D, rtol=0, atol=1e-13)
    # an already sorted list is taken as it is: duplicates and zeros stay
    r, c, v = np.array([0, 0, 1, 1]), np.array([1, 1, 0, 2]), np.array([1.0, 2.0, 0.0, 4.0])
    rows, cols, A, C, V = dropin_coo_to_csr(2, 3, r, c, v)
    assert list(A) == [0, 2, 4] and list(C) == [1, 1, 0, 2] and list(V) == [1.0, 2.0, 0.0, 4.0]
    # an unsorted list of zeros compacts to one (0, 0, 0.) entry, as the reference's does
    rows, cols, A, C, V = dropin_coo_to_csr(2, 2, [1, 0], [0, 1], [0.0, 0.0])
    assert list(A) == [0, 1, 1] and list(C) == [0] and list(V) == [0.0]


def test_products_closed_form():
    rng = np.random.default_rng(6)
    for unsorted in (False, True):
        m1 = _random_csr(rng, 20, 15, 5, unsorted)
        m2 = _random_csr(rng, 15, 12, 4, unsorted)
        p = dropin_op("mul", m1, m2)
        assert np.allclose(_dense(p), _dense(m1) @ _dense(m2), atol=1e-12)
        q = dropin_op("mul_reduced", m1, m2)
        assert np.allclose(_dense(q), _dense(m1) @ _dense(m2), atol=1e-12)
        for prod in (p, q):  # rows sorted, no repeated columns, no zeros
            _, _, A, C, V = prod
            for row in range(20):
                assert np.all(np.diff(C[A[row]:A[row + 1]]) > 0)
            assert np.all(V != 0.0)
        center = _random_csr(rng, 15, 15, 4, unsorted)
        factor = _random_csr(rng, 15, 9, 3, unsorted)
        cj = dropin_op("conjugation", center, factor)
        F = _dense(factor)
        assert np.allclose(_dense(cj), F.T @ _dense(center) @ F, atol=1e-12)
        # one entry per (r, i, j, k) product, nothing merged
        _, _, cA, cC, _ = center
        _, _, fA, fC, _ = factor
        count = sum((fA[r + 1] - fA[r]) * (fA[cC[i] + 1] - fA[cC[i]])
                    for r in range(15) for i in range(cA[r], cA[r + 1]))
        assert cj[2][-1] == count
        t = dropin_op("transpose", m1)
        assert np.allclose(_dense(t), _dense(m1).T, atol=1e-13)


@needs_ref
def test_sparse_ops_match_reference_bitwise():
    """COO -> CSR, products, Conjugation, transpose, addition and compression
    give the reference's arrays bit for bit (the compiled reference,
    sparse/sparsematrix.cpp and sparse/matcsr.cpp)."""
    from oracle.ref import RefRunner

    rng = np.random.default_rng(7)
    with RefRunner(1) as ref:
        for rows, cols, n in [(5, 4, 30), (40, 35, 500), (200, 150, 4000)]:
            r, c, v = _random_coo(rng, rows, cols, n)
            mine = dropin_coo_to_csr(rows, cols, r, c, v)
            theirs = ref.coo_to_csr(rows, cols, r, c, v)
            for a, b in zip(mine, theirs):
                assert np.array_equal(np.asarray(a), np.asarray(b))
        for unsorted in (False, True):
            for k in range(4):
                m1 = _random_csr(rng, 30 + k, 25, 6, unsorted)
                m2 = _random_csr(rng, 25, 20 + k, 5, unsorted)
                sq = _random_csr(rng, 30 + k, 30 + k, 5, unsorted)
                f = _random_csr(rng, 30 + k, 18, 4, unsorted)
                m1b = _random_csr(rng, 30 + k, 25, 6, unsorted)
                cases = [("mul", m1, m2, {}), ("mul_reduced", m1, m2, {}),
                         ("conjugation", sq, f, {}), ("transpose", m1, None, {}),
                         ("add", m1, m1b, {"s1": 0.75, "s2": -1.25}), ("compress", m1, None, {})]
                for op, a, b, kw in cases:
                    mine = dropin_op(op, a, b, **kw)
                    theirs = ref.csr_op(op, a, b, **kw)
                    for x, y in zip(mine, theirs):
                        assert np.array_equal(np.asarray(x), np.asarray(y)), (op, unsorted, k)
