"""Python binding of the oracle (CPU restatement of FEEC++'s CG/SSOR path).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of bench.py may import this package: it is what the tests
compare the HIP path against, never part of the product.

* ``liboracle.so`` (oracle.c) -- the restatement, in-process via ctypes;
* ``oracle.ref`` -- the reference's own code compiled from its sources into
  ``_ref/libfeecref.so`` (present only where the reference was available at
  build time), run in a child process because loading it switches the
  calling threads to flush-to-zero (base/base.cpp:255-288).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libfeecref.so")

EXIT_CONVERGED, EXIT_CAP, EXIT_BREAKDOWN, EXIT_SMALL = 0, 1, 2, 3


class OrcInfo(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("success", c_int), ("exit_reason", c_int),
                ("pad_", c_int), ("zr", c_double), ("tolerance", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


_P_I = POINTER(c_int)
_P_D = POINTER(c_double)
_SOLVER = [c_int, _P_D, _P_D, _P_I, _P_I, _P_D, _P_D, c_double, c_int]
_SIG = {
    "orc_csr_apply": ([c_int, _P_I, _P_I, _P_D, _P_D, _P_D, c_double], None),
    "orc_csr_diagonal": ([c_int, _P_I, _P_I, _P_D, _P_D], None),
    "orc_rainbow": ([c_int, _P_I, _P_I, _P_I, _P_I, _P_I], c_int),
    "orc_cg": (_SOLVER + [POINTER(OrcInfo)], c_int),
    "orc_cg_jacobi": (_SOLVER + [_P_D, POINTER(OrcInfo)], c_int),
    "orc_cg_ssor": (_SOLVER + [_P_D, c_double, POINTER(OrcInfo)], c_int),
    "orc_cg_ssor_eisenstat": (_SOLVER + [_P_D, c_double, POINTER(OrcInfo)], c_int),
    "orc_cg_rainbow": (_SOLVER + [_P_D, c_double, c_int, _P_I, _P_I, _P_I, POINTER(OrcInfo)], c_int),
    "orc_cg_eisenstat_rainbow":
        (_SOLVER + [_P_D, c_double, c_int, _P_I, _P_I, _P_I, POINTER(OrcInfo)], c_int),
    "orc_hodge_cr_ssor": ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
                          [_P_D, c_double, c_int, c_double, c_int, POINTER(OrcInfo)], c_int),
    "orc_hodge_cr_diagonal": ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
                              [_P_D, c_double, c_int, c_double, c_int, POINTER(OrcInfo)], c_int),
    "orc_hodge_cr_textbook": ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
                              [_P_D, c_double, c_int, c_double, c_int, POINTER(OrcInfo)], c_int),
}

_lib = None


def build():
    """Compiles liboracle.so and, where the reference sources exist, _ref/."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIG.items():
            fn = getattr(_lib, name)
            fn.argtypes = args
            fn.restype = res
    return _lib


def _i(a):
    return None if a is None else a.ctypes.data_as(_P_I)


def _d(a):
    return None if a is None else a.ctypes.data_as(_P_D)


def _arr_i(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _arr_d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def csr_apply(n, A, C, V, src, scaling=1.0):
    A, C, V, src = _arr_i(A), _arr_i(C), _arr_d(V), _arr_d(src)
    dest = np.empty(n)
    lib().orc_csr_apply(n, _i(A), _i(C), _d(V), _d(src), _d(dest), float(scaling))
    return dest


def csr_diagonal(n, A, C, V):
    A, C, V = _arr_i(A), _arr_i(C), _arr_d(V)
    d = np.empty(n)
    lib().orc_csr_diagonal(n, _i(A), _i(C), _d(V), _d(d))
    return d


def rainbow(n, A, C):
    """Returns (num_colors, F, B, R)."""
    A, C = _arr_i(A), _arr_i(C)
    F = np.empty(n, dtype=np.int32)
    B = np.empty(n + 1, dtype=np.int32)
    R = np.empty(n, dtype=np.int32)
    nc = lib().orc_rainbow(n, _i(A), _i(C), _i(F), _i(B), _i(R))
    return nc, F, B[: nc + 1].copy(), R


VARIANTS = ("cg", "jacobi", "ssor", "ssor_eisenstat", "rainbow", "eisenstat_rainbow")


def solve(variant, A, C, V, b, x0=None, residual0=None, precision=1e-10, print_modulo=-3,
          aux=None, omega=1.0, colouring=None):
    """Runs one oracle solver.  Returns (x, residual, info dict)."""
    A, C, V, b = _arr_i(A), _arr_i(C), _arr_d(V), _arr_d(b)
    n = A.shape[0] - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    res = np.zeros(n) if residual0 is None else np.array(residual0, dtype=np.float64)
    info = OrcInfo()
    head = [n, _d(x), _d(b), _i(A), _i(C), _d(V), _d(res), float(precision), int(print_modulo)]
    aux = _arr_d(aux)
    L = lib()
    if variant == "cg":
        L.orc_cg(*head, ctypes.byref(info))
    elif variant == "jacobi":
        L.orc_cg_jacobi(*head, _d(aux), ctypes.byref(info))
    elif variant in ("ssor", "ssor_eisenstat"):
        getattr(L, "orc_cg_" + variant)(*head, _d(aux), float(omega), ctypes.byref(info))
    elif variant in ("rainbow", "eisenstat_rainbow"):
        nc, F, B, R = colouring
        F, B, R = _arr_i(F), _arr_i(B), _arr_i(R)
        getattr(L, "orc_cg_" + variant)(*head, _d(aux), float(omega), int(nc), _i(F), _i(B),
                                        _i(R), ctypes.byref(info))
    else:
        raise ValueError(variant)
    return x, res, info.as_dict()


def hodge_cr_ssor(N, L, A, B, Bt, C, b, x0=None, precision=1e-10, print_modulo=-3,
                  inner_precision=1e-10, inner_print_modulo=-3, inner="ssor"):
    """HodgeConjugateResidualSolverCSR_SSOR (solver/systemsparsesolver.cpp:426-797)
    on (B A^-1 Bt + C) x = b; A is L x L, B is N x L, Bt is L x N, C is N x N,
    each a (rows, cols, vals) CSR triple.  inner="diagonal" runs
    HodgeConjugateResidualSolverCSR_diagonal (:52-424, inner Jacobi CG),
    inner="textbook" HodgeConjugateResidualSolverCSR_textbook (:798-1121, inner
    plain CG).  Returns (x, residual, info dict)."""
    if inner not in ("ssor", "diagonal", "textbook"):
        raise ValueError(inner)
    mats = []
    for m in (A, B, Bt, C):
        mats += [_arr_i(m[0]), _arr_i(m[1]), _arr_d(m[2])]
    x = np.zeros(N) if x0 is None else np.array(x0, dtype=np.float64)
    b = _arr_d(b)
    res = np.zeros(N)
    info = OrcInfo()
    ptrs = [(_i if k % 3 < 2 else _d)(m) for k, m in enumerate(mats)]
    fn = getattr(lib(), "orc_hodge_cr_" + inner)
    fn(int(N), int(L), _d(x), _d(b), *ptrs, _d(res), float(precision), int(print_modulo),
       float(inner_precision), int(inner_print_modulo), ctypes.byref(info))
    return x, res, info.as_dict()


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH)
