"""Calls into the reference's own compiled code (oracle/_ref/libfeecref.so).

TEST INFRASTRUCTURE ONLY (test comparisons and the CPU baseline).  Each RefRunner owns
one child process (oracle/_refworker.py) that loads the library with
OMP_NUM_THREADS fixed, so the reference's start-up code (banner, flush-to-zero
on every OpenMP thread, base/base.cpp:38-292) stays out of the calling
process.
"""
from __future__ import annotations

import ctypes
import os
import pickle
import struct
import subprocess
import sys
from ctypes import POINTER, c_double, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB_PATH = os.path.join(HERE, "_ref", "libfeecref.so")

_P_I = POINTER(c_int)
_P_D = POINTER(c_double)
_VARIANT = {"cg": 0, "jacobi": 1, "ssor": 2, "ssor_eisenstat": 3, "rainbow": 4,
            "eisenstat_rainbow": 5}
_lib = None


def _init_library():
    """Runs in the child process only."""
    global _lib
    _lib = ctypes.CDLL(REF_LIB_PATH)
    _lib.feecref_desired_precision.restype = c_double
    _lib.feecref_omp_threads.restype = c_int
    _lib.feecref_rainbow.argtypes = [c_int, _P_I, _P_I, _P_D, _P_I, _P_I, _P_I]
    _lib.feecref_rainbow.restype = c_int
    _lib.feecref_csr_apply.argtypes = [c_int, _P_I, _P_I, _P_D, _P_D, _P_D, c_double]
    _lib.feecref_csr_diagonal.argtypes = [c_int, _P_I, _P_I, _P_D, _P_D]
    _lib.feecref_cg.argtypes = [c_int, c_int, _P_D, _P_D, _P_I, _P_I, _P_D, _P_D, c_double, c_int,
                                _P_D, c_double, c_int, _P_I, _P_I, _P_I, _P_D]
    _lib.feecref_cg.restype = c_int
    _lib.feecref_coo_to_csr.argtypes = [c_int, c_int, c_int, _P_I, _P_I, _P_D]
    _lib.feecref_coo_to_csr.restype = c_int
    _lib.feecref_csr_op.argtypes = [c_int, c_int, c_int, _P_I, _P_I, _P_D, c_int, c_int, _P_I,
                                    _P_I, _P_D, c_double, c_double]
    _lib.feecref_csr_op.restype = c_int
    _lib.feecref_result.argtypes = [_P_I, _P_I, _P_I, _P_D]
    _lib.feecref_hodge_cr_ssor.argtypes = [c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 + \
        [_P_D, c_double, c_int, c_double, c_int]
    _lib.feecref_hodge_cr_ssor.restype = c_int
    _lib.feecref_hodge_cr_diagonal.argtypes = _lib.feecref_hodge_cr_ssor.argtypes
    _lib.feecref_hodge_cr_diagonal.restype = c_int
    _lib.feecref_hodge_cr_textbook.argtypes = _lib.feecref_hodge_cr_ssor.argtypes
    _lib.feecref_hodge_cr_textbook.restype = c_int
    _lib.feecref_mesh.argtypes = [c_int, c_int, c_int]
    _lib.feecref_mesh.restype = c_int
    _lib.feecref_mesh_sizes.argtypes = [_P_I]
    _lib.feecref_mesh_copy.argtypes = [_P_D, _P_I, _P_I, _P_I, _P_I, _P_D, _P_I, _P_I, _P_D]
    _lib.feecref_whitney1.argtypes = [c_int]
    _lib.feecref_whitney1.restype = c_int
    _lib.feecref_whitney1_sizes.argtypes = [_P_I]
    _lib.feecref_whitney1_copy.argtypes = [_P_I, _P_I, _P_I, _P_I, _P_D, _P_I, _P_I, _P_D]
    _lib.feecref_sullivan2d.argtypes = [c_int, c_int]
    _lib.feecref_sullivan2d.restype = c_int
    _lib.feecref_sullivan2d_sizes.argtypes = [_P_I]
    _lib.feecref_sullivan2d_copy.argtypes = [_P_I] * 7 + [_P_D]


def _i(a):
    return None if a is None else a.ctypes.data_as(_P_I)


def _d(a):
    return None if a is None else a.ctypes.data_as(_P_D)


def _result(nnz, rows):
    dims = np.zeros(2, dtype=np.int32)
    A = np.empty(rows + 1, dtype=np.int32)
    C = np.empty(max(nnz, 1), dtype=np.int32)
    V = np.empty(max(nnz, 1))
    _lib.feecref_result(_i(dims), _i(A), _i(C), _d(V))
    return int(dims[0]), int(dims[1]), A, C[:nnz].copy(), V[:nnz].copy()


def _task(kind, kw):
    L = _lib
    if kind == "info":
        return {"desired_precision": L.feecref_desired_precision(),
                "threads": L.feecref_omp_threads()}
    if kind == "coo_to_csr":
        er = np.ascontiguousarray(kw["rows_of"], dtype=np.int32)
        ec = np.ascontiguousarray(kw["cols_of"], dtype=np.int32)
        ev = np.ascontiguousarray(kw["vals_of"], dtype=np.float64)
        nnz = L.feecref_coo_to_csr(int(kw["rows"]), int(kw["cols"]), len(ev), _i(er), _i(ec),
                                   _d(ev))
        return _result(nnz, int(kw["rows"]))
    if kind == "csr_op":
        m1, m2 = kw["m1"], kw["m2"] if kw.get("m2") is not None else kw["m1"]
        a1 = [np.ascontiguousarray(m1[k], dtype=t) for k, t in ((2, np.int32), (3, np.int32), (4, np.float64))]
        a2 = [np.ascontiguousarray(m2[k], dtype=t) for k, t in ((2, np.int32), (3, np.int32), (4, np.float64))]
        nnz = L.feecref_csr_op(int(kw["op"]), int(m1[0]), int(m1[1]), _i(a1[0]), _i(a1[1]),
                               _d(a1[2]), int(m2[0]), int(m2[1]), _i(a2[0]), _i(a2[1]), _d(a2[2]),
                               float(kw.get("s1", 1.0)), float(kw.get("s2", 1.0)))
        if nnz < 0:
            raise ValueError("unknown csr op")
        out_rows = {2: m2[1], 3: m1[1]}.get(int(kw["op"]), m1[0])
        return _result(nnz, int(out_rows))
    if kind == "mesh":
        dim, levels, mats = int(kw["dim"]), int(kw["levels"]), int(kw["matrices"])
        L.feecref_mesh(dim, levels, mats)
        sz = np.zeros(4, dtype=np.int32)
        L.feecref_mesh_sizes(_i(sz))
        nv, ns, nb, na = (int(t) for t in sz)
        coords = np.empty(nv * dim)
        simp = np.empty(ns * (dim + 1), dtype=np.int32)
        dirichlet = np.empty(nv, dtype=np.int32)
        bA, aA = np.empty(nv + 1, dtype=np.int32), np.empty(nv + 1, dtype=np.int32)
        bC, bV = np.empty(max(nb, 1), dtype=np.int32), np.empty(max(nb, 1))
        aC, aV = np.empty(max(na, 1), dtype=np.int32), np.empty(max(na, 1))
        L.feecref_mesh_copy(_d(coords), _i(simp), _i(dirichlet), _i(bA), _i(bC), _d(bV), _i(aA),
                            _i(aC), _d(aV))
        out = {"coords": coords.reshape(nv, dim), "simplices": simp.reshape(ns, dim + 1),
               "dirichlet": dirichlet.astype(bool)}
        if mats:
            out["structural"] = (bA, bC[:nb].copy(), bV[:nb].copy())
            out["compressed"] = (aA, aC[:na].copy(), aV[:na].copy())
        return out
    if kind == "whitney1":
        L.feecref_whitney1(int(kw["levels"]))
        sz = np.zeros(3, dtype=np.int32)
        L.feecref_whitney1_sizes(_i(sz))
        ne, nk, nm = (int(t) for t in sz)
        edges = np.empty(2 * ne, dtype=np.int32)
        dirichlet = np.empty(ne, dtype=np.int32)
        kA, mA = np.empty(ne + 1, dtype=np.int32), np.empty(ne + 1, dtype=np.int32)
        kC, kV = np.empty(max(nk, 1), dtype=np.int32), np.empty(max(nk, 1))
        mC, mV = np.empty(max(nm, 1), dtype=np.int32), np.empty(max(nm, 1))
        L.feecref_whitney1_copy(_i(edges), _i(dirichlet), _i(kA), _i(kC), _d(kV), _i(mA), _i(mC),
                                _d(mV))
        return {"edges": edges.reshape(ne, 2), "dirichlet": dirichlet.astype(bool),
                "curlcurl": (kA, kC[:nk].copy(), kV[:nk].copy()),
                "mass": (mA, mC[:nm].copy(), mV[:nm].copy())}
    if kind == "sullivan2d":
        L.feecref_sullivan2d(int(kw["levels"]), int(kw["r"]))
        sz = np.zeros(5, dtype=np.int32)
        L.feecref_sullivan2d_sizes(_i(sz))
        nv, ne, nt, n, nnz = (int(t) for t in sz)
        edges = np.empty(2 * ne, dtype=np.int32)
        vflag, eflag = np.empty(nv, dtype=np.int32), np.empty(ne, dtype=np.int32)
        tri, tri_e = np.empty(3 * nt, dtype=np.int32), np.empty(3 * nt, dtype=np.int32)
        A, C, V = np.empty(n + 1, dtype=np.int32), np.empty(max(nnz, 1), dtype=np.int32), \
            np.empty(max(nnz, 1))
        L.feecref_sullivan2d_copy(_i(edges), _i(vflag), _i(eflag), _i(tri), _i(tri_e), _i(A),
                                  _i(C), _d(V))
        return {"edges": edges.reshape(ne, 2), "vertex_dirichlet": vflag.astype(bool),
                "edge_dirichlet": eflag.astype(bool), "triangles": tri.reshape(nt, 3),
                "triangle_edges": tri_e.reshape(nt, 3),
                "stiffness": (A, C[:nnz].copy(), V[:nnz].copy())}
    if kind == "hodge":
        N, Ldim = int(kw["N"]), int(kw["L"])
        mats = []
        for m in ("A", "B", "Bt", "C"):
            a, c, v = kw[m]
            mats += [np.ascontiguousarray(a, dtype=np.int32), np.ascontiguousarray(c, dtype=np.int32),
                     np.ascontiguousarray(v, dtype=np.float64)]
        x = np.array(kw["x0"], dtype=np.float64)
        b = np.ascontiguousarray(kw["b"], dtype=np.float64)
        res = np.zeros(N)
        ptrs = [(_i if k % 3 < 2 else _d)(m) for k, m in enumerate(mats)]
        fn = getattr(L, "feecref_hodge_cr_" + kw.get("inner", "ssor"))
        K = fn(N, Ldim, _d(x), _d(b), *ptrs, _d(res), float(kw["precision"]),
               int(kw["print_modulo"]), float(kw["inner_precision"]), int(kw["inner_print_modulo"]))
        return x, res, K
    A = np.ascontiguousarray(kw["A"], dtype=np.int32)
    C = np.ascontiguousarray(kw["C"], dtype=np.int32)
    V = np.ascontiguousarray(kw["V"], dtype=np.float64)
    n = A.shape[0] - 1
    if kind == "rainbow":
        F = np.empty(n, dtype=np.int32)
        B = np.empty(n + 1, dtype=np.int32)
        R = np.empty(n, dtype=np.int32)
        nc = L.feecref_rainbow(n, _i(A), _i(C), _d(V), _i(F), _i(B), _i(R))
        return nc, F, B[: nc + 1].copy(), R
    if kind == "apply":
        src = np.ascontiguousarray(kw["src"], dtype=np.float64)
        dest = np.empty(n)
        L.feecref_csr_apply(n, _i(A), _i(C), _d(V), _d(src), _d(dest), float(kw["scaling"]))
        return dest
    if kind == "diagonal":
        d = np.empty(n)
        L.feecref_csr_diagonal(n, _i(A), _i(C), _d(V), _d(d))
        return d
    if kind == "solve":
        b = np.ascontiguousarray(kw["b"], dtype=np.float64)
        x = np.array(kw["x0"], dtype=np.float64)
        res = np.array(kw["residual0"], dtype=np.float64)
        aux = kw.get("aux")
        aux = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
        col = kw.get("colouring")
        nc, F, B, R = (0, None, None, None) if col is None else col
        F = None if F is None else np.ascontiguousarray(F, dtype=np.int32)
        B = None if B is None else np.ascontiguousarray(B, dtype=np.int32)
        R = None if R is None else np.ascontiguousarray(R, dtype=np.int32)
        secs = c_double(0.0)
        times = []
        K = -1
        for _ in range(kw.get("repeat", 1)):
            xx, rr = x.copy(), res.copy()
            K = L.feecref_cg(_VARIANT[kw["variant"]], n, _d(xx), _d(b), _i(A), _i(C), _d(V),
                             _d(rr), float(kw["precision"]), int(kw["print_modulo"]), _d(aux),
                             float(kw["omega"]), int(nc), _i(F), _i(B), _i(R), ctypes.byref(secs))
            times.append(secs.value)
        return xx, rr, K, times
    raise ValueError(kind)


class RefRunner:
    """One child process running the reference library at `threads` threads."""

    def __init__(self, threads: int = 1, timeout: float = 600.0, bind: bool = False):
        """bind: OMP_PROC_BIND=close, OMP_PLACES=cores (one thread per core of
        the process's CPU set), as the timed CPU baseline runs."""
        if not os.path.exists(REF_LIB_PATH):
            raise FileNotFoundError(REF_LIB_PATH)
        self.threads = threads
        self.timeout = timeout
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        if bind:
            env.update(OMP_PROC_BIND="close", OMP_PLACES="cores")
        self._proc = subprocess.Popen([sys.executable, os.path.join(HERE, "_refworker.py"),
                                       str(threads)], stdin=subprocess.PIPE,
                                      stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                                      env=env)

    def _call(self, kind, **kw):
        data = pickle.dumps((kind, kw), protocol=4)
        self._proc.stdin.write(struct.pack("<Q", len(data)) + data)
        self._proc.stdin.flush()
        head = self._proc.stdout.read(8)
        if len(head) < 8:
            raise RuntimeError("reference worker exited (status %s)" % self._proc.poll())
        (size,) = struct.unpack("<Q", head)
        ok, out = pickle.loads(self._proc.stdout.read(size))
        if not ok:
            raise RuntimeError("reference worker: " + out)
        return out

    def info(self):
        return self._call("info")

    def coo_to_csr(self, rows, cols, rows_of, cols_of, vals_of):
        """MatrixCSR(SparseMatrix(rows, cols, entries)): (rows, cols, A, C, V)."""
        return self._call("coo_to_csr", rows=rows, cols=cols, rows_of=rows_of, cols_of=cols_of,
                          vals_of=vals_of)

    def csr_op(self, op, m1, m2=None, s1=1.0, s2=1.0):
        """op in mul, mul_reduced, conjugation, transpose, add, compress; a
        matrix is (rows, cols, A, C, V).  Returns (rows, cols, A, C, V)."""
        ops = {"mul": 0, "mul_reduced": 1, "conjugation": 2, "transpose": 3, "add": 4,
               "compress": 5}
        return self._call("csr_op", op=ops[op], m1=m1, m2=m2, s1=s1, s2=s2)

    def rainbow(self, A, C, V):
        return self._call("rainbow", A=A, C=C, V=V)

    def mesh(self, dim, levels, matrices=True):
        """UnitSquare2D_simple (dim 2) or UnitCube3D (dim 3) refined `levels`
        times: {"coords", "simplices", "dirichlet"} and, with matrices, the
        structural P1 stiffness (LagrangeCoefficientStiffnessMatrix, identity
        weight) and the compressed one (MatrixCSR(LagrangeStiffnessMatrix) +
        compressentries) as (A, C, V)."""
        return self._call("mesh", dim=dim, levels=levels, matrices=1 if matrices else 0)

    def whitney1(self, levels):
        """UnitCube3D refined `levels` times: {"edges" (vertex pairs, the
        mesh's order), "dirichlet" (per edge), "curlcurl", "mass"} with the
        Whitney 1-form matrices as (A, C, V), duplicates kept as
        Conjugation leaves them."""
        return self._call("whitney1", levels=levels)

    def sullivan2d(self, levels, r):
        """UnitSquare2D_simple refined `levels` times: {"edges", "vertex_dirichlet",
        "edge_dirichlet", "triangles", "triangle_edges", "stiffness" (A, C, V)}
        with the degree-r stiffness in the Sullivan basis, composed as
        tests/sullivan2D/poissondirichlet2D.cpp composes it."""
        return self._call("sullivan2d", levels=levels, r=r)

    def csr_apply(self, A, C, V, src, scaling=1.0):
        return self._call("apply", A=A, C=C, V=V, src=src, scaling=scaling)

    def csr_diagonal(self, A, C, V):
        return self._call("diagonal", A=A, C=C, V=V)

    def solve(self, variant, A, C, V, b, x0=None, residual0=None, precision=1e-10,
              print_modulo=-3, aux=None, omega=1.0, colouring=None, repeat=1):
        """Returns (x, residual, K, [seconds per solve])."""
        n = len(A) - 1
        x0 = np.zeros(n) if x0 is None else x0
        residual0 = np.zeros(n) if residual0 is None else residual0
        return self._call("solve", variant=variant, A=A, C=C, V=V, b=b, x0=x0,
                          residual0=residual0, precision=precision, print_modulo=print_modulo,
                          aux=aux, omega=omega, colouring=colouring, repeat=repeat)

    def hodge_cr_ssor(self, N, L, A, B, Bt, C, b, x0=None, precision=1e-10, print_modulo=-3,
                      inner_precision=1e-10, inner_print_modulo=-3, inner="ssor"):
        """HodgeConjugateResidualSolverCSR_SSOR (inner="diagonal":
        HodgeConjugateResidualSolverCSR_diagonal, inner="textbook":
        HodgeConjugateResidualSolverCSR_textbook); matrices are (rows, cols,
        vals) CSR triples.  Returns (x, residual, K)."""
        if inner not in ("ssor", "diagonal", "textbook"):
            raise ValueError(inner)
        x0 = np.zeros(N) if x0 is None else x0
        return self._call("hodge", N=N, L=L, A=A, B=B, Bt=Bt, C=C, b=b, x0=x0,
                          precision=precision, print_modulo=print_modulo,
                          inner_precision=inner_precision, inner_print_modulo=inner_print_modulo,
                          inner=inner)

    def close(self):
        if self._proc.poll() is None:
            self._proc.stdin.close()
            try:
                self._proc.wait(timeout=10)
            except subprocess.TimeoutExpired:
                self._proc.kill()
                self._proc.wait()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
