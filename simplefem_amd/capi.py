"""ctypes binding of libsimplefem_amd.so and a Python mirror of the reference API.

The functions below keep the names, argument order and in/out meaning of
FEEC++'s raw-pointer CSR solvers (solver/sparsesolver.hpp:11-84): ``x`` and
``residual`` are float64 numpy arrays updated in place and the return value is
the iteration count K.  Where the reference would abort on a failed assert,
these raise :class:`SfaError`.  Every call goes to the HIP library; nothing
here computes a result on the CPU, and a missing library is an ImportError.
"""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import POINTER, c_double, c_int, c_longlong, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsimplefem_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "simplefem_amd.h")

SFA_OK = 0
ERRORS = {-1: "SFA_ERR_INVALID", -2: "SFA_ERR_HIP", -3: "SFA_ERR_NO_DEVICE",
          -4: "SFA_ERR_OUT_OF_MEMORY", -5: "SFA_ERR_COLOURING"}
EXIT_CONVERGED, EXIT_CAP, EXIT_BREAKDOWN, EXIT_SMALL = 0, 1, 2, 3
METHOD_CG, METHOD_DIAGONAL, METHOD_SSOR, METHOD_SSOR_EISENSTAT, METHOD_RAINBOW, \
    METHOD_EISENSTAT_RAINBOW = range(6)


class SfaError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("success", c_int), ("exit_reason", c_int),
                ("reserved", c_int), ("zr", c_double), ("tolerance", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SystemStats(ctypes.Structure):
    _fields_ = [("n", c_int), ("n_active", c_int), ("nnz", c_longlong),
                ("groups_forward", c_int), ("groups_backward", c_int),
                ("stored_entries", c_longlong), ("bytes_per_iteration", c_double),
                ("bytes_spmv", c_double), ("kernels_per_iteration", c_int),
                ("spmv_lanes_per_row", c_int), ("setup_ms", c_double * 4)]


class RefineInfo(ctypes.Structure):
    _fields_ = [("outer", c_int), ("inner_iterations", c_int), ("success", c_int),
                ("continued", c_int), ("residual_norm", c_double), ("relative_residual", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


SPMV_BUCKETS = 4
SPMV_MAPPINGS = (1, 2, 4, 8)
SPMV_BUCKET_NAMES = ("<=8", "<=16", "<=32", ">32")


class SpmvTuning(ctypes.Structure):
    _fields_ = [("lanes_per_row", c_int * SPMV_BUCKETS), ("rows", c_longlong * SPMV_BUCKETS),
                ("us", (c_double * len(SPMV_MAPPINGS)) * SPMV_BUCKETS)]

    def as_dict(self):
        """{bucket: {"rows", "lanes_per_row", "us": {lanes: microseconds}}} for non-empty buckets."""
        out = {}
        for b, name in enumerate(SPMV_BUCKET_NAMES):
            if self.rows[b] == 0:
                continue
            out[name] = {"rows": int(self.rows[b]), "lanes_per_row": int(self.lanes_per_row[b]),
                         "us": {k: round(self.us[b][i], 3) for i, k in enumerate(SPMV_MAPPINGS)
                                if self.us[b][i] > 0}}
        return out


FLAG_FP32_INNER = 0x100
FLAG_PAIRED_SWEEPS = 0x200
FLAG_INNER_RESTARTS = 0x400
FLAG_BORROW_CSR = 0x800
BATCH_FP64, BATCH_MIXED, BATCH_MIXED_RESTARTS = 0, 1, 2


class PreparedMatrix:
    """A matrix's device-order plan built on the host (sfa_batch_prepare_matrix):
    no device or batch is touched, so any thread may build one (the call
    releases the GIL) while another drives a batch; Batch.add_prepared uploads it."""

    def __init__(self, N, A, C, V, diagonal, omega, num_colors, F, B, R):
        A, C, V = _csr(A, C, V)
        diagonal = _f64(diagonal, "diagonal", N)
        F, B, R = _i32(F, "F"), _i32(B, "B"), _i32(R, "R")
        self.N = N
        self._h = c_void_p()
        _check(load().sfa_batch_prepare_matrix(N, _i(A), _i(C), _d(V), _d(diagonal),
                                               float(omega), int(num_colors), _i(F), _i(B),
                                               _i(R), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            load().sfa_prepared_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchInfo(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("outer", c_int), ("success", c_int),
                ("exit_reason", c_int), ("zr", c_double), ("residual_norm", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelProfile(ctypes.Structure):
    _fields_ = [("launches", c_int), ("reserved", c_int), ("milliseconds", c_double),
                ("bytes", c_double)]


KERNEL_CLASSES = 8
KERNEL_NAMES = ["spmv", "forward", "backward", "direction", "update", "triangular", "combine",
                "restart"]

_P_I = POINTER(c_int)
_P_D = POINTER(c_double)
_SOLVER_ARGS = [c_int, _P_D, _P_D, _P_I, _P_I, _P_D, _P_D, c_double, c_int]
_SIGNATURES = {
    "sfa_last_error": ([], ctypes.c_char_p),
    "sfa_api_version": ([], c_int),
    "sfa_ConjugateGradientSolverCSR": (_SOLVER_ARGS + [POINTER(SolveInfo)], c_int),
    "sfa_ConjugateGradientSolverCSR_DiagonalPreconditioner":
        (_SOLVER_ARGS + [_P_D, POINTER(SolveInfo)], c_int),
    "sfa_ConjugateGradientSolverCSR_SSOR": (_SOLVER_ARGS + [_P_D, c_double, POINTER(SolveInfo)], c_int),
    "sfa_ConjugateGradientSolverCSR_SSOR_Eisenstat":
        (_SOLVER_ARGS + [_P_D, c_double, POINTER(SolveInfo)], c_int),
    "sfa_ConjugateGradientSolverCSR_Rainbow":
        (_SOLVER_ARGS + [_P_D, c_double, c_int, _P_I, _P_I, _P_I, POINTER(SolveInfo)], c_int),
    "sfa_ConjugateGradientSolverCSR_Eisenstat_Rainbow":
        (_SOLVER_ARGS + [_P_D, c_double, c_int, _P_I, _P_I, _P_I, POINTER(SolveInfo)], c_int),
    "sfa_ConjugateGradientSolverCSR_Rainbow_noalloc":
        ([c_int] + [c_void_p] * 5 + [c_void_p, c_double, c_int, c_void_p, c_double, c_int]
         + [c_void_p] * 3 + [POINTER(SolveInfo)], c_int),
    "sfa_MatrixCSR_apply": ([c_int, c_int, _P_I, _P_I, _P_D, _P_D, _P_D, c_double], c_int),
    "sfa_MatrixCSR_getDiagonal": ([c_int, _P_I, _P_I, _P_D, _P_D], c_int),
    "sfa_Rainbow": ([c_int, _P_I, _P_I, _P_I, _P_I, _P_I, _P_I], c_int),
    "sfa_system_create": ([c_int, c_int, _P_I, _P_I, _P_D, _P_D, c_double, c_int, _P_I, _P_I,
                           _P_I, POINTER(c_void_p)], c_int),
    "sfa_system_solve": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, c_int,
                          POINTER(SolveInfo)], c_int),
    "sfa_system_apply": ([c_void_p, c_void_p, c_void_p, c_double, c_int], c_int),
    "sfa_system_solve_mixed": ([c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, c_int,
                                POINTER(RefineInfo)], c_int),
    "sfa_system_set_stream": ([c_void_p, c_void_p], c_int),
    "sfa_system_query": ([c_void_p, POINTER(SystemStats)], c_int),
    "sfa_system_profile": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int,
                            POINTER(KernelProfile), POINTER(SolveInfo)], c_int),
    "sfa_system_tune_spmv": ([c_void_p, c_int, POINTER(SpmvTuning)], c_int),
    "sfa_system_time_sweeps": ([c_void_p, c_int, c_int, POINTER(c_double)], c_int),
    "sfa_system_set_spmv": ([c_void_p, c_int], c_int),
    "sfa_system_destroy": ([c_void_p], c_int),
    "sfa_grid_barrier_cost": ([c_int, c_int, c_int, c_int, POINTER(c_double), POINTER(c_int)],
                              c_int),
    "sfa_batch_create": ([c_int, c_int, POINTER(c_void_p)], c_int),
    "sfa_batch_add_matrix": ([c_void_p, c_int, _P_I, _P_I, _P_D, _P_D, c_double, c_int, _P_I,
                              _P_I, _P_I, POINTER(c_int)], c_int),
    "sfa_batch_remove_matrix": ([c_void_p, c_int], c_int),
    "sfa_batch_prepare_matrix": ([c_int, _P_I, _P_I, _P_D, _P_D, c_double, c_int, _P_I, _P_I,
                                  _P_I, POINTER(c_void_p)], c_int),
    "sfa_batch_add_prepared": ([c_void_p, c_void_p, POINTER(c_int)], c_int),
    "sfa_prepared_destroy": ([c_void_p], c_int),
    "sfa_batch_submit": ([c_void_p, c_int, _P_D, _P_D, c_double, c_double, c_int,
                          POINTER(c_int)], c_int),
    "sfa_batch_submit_group": ([c_void_p, c_int, c_int, _P_D, _P_D, c_double, c_double, c_int,
                                _P_I], c_int),
    "sfa_batch_run": ([c_void_p, c_int, POINTER(c_int)], c_int),
    "sfa_MatrixCSRMultiplication": ([c_int, c_int, _P_I, _P_I, _P_D, c_int, c_int, _P_I, _P_I,
                                     _P_D, POINTER(c_void_p)], c_int),
    "sfa_MatrixCSRMultiplication_reduced": ([c_int, c_int, _P_I, _P_I, _P_D, c_int, c_int, _P_I,
                                             _P_I, _P_D, POINTER(c_void_p)], c_int),
    "sfa_MatrixCSR_getTranspose": ([c_int, c_int, _P_I, _P_I, _P_D, POINTER(c_void_p)], c_int),
    "sfa_csr_sizes": ([c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)], c_int),
    "sfa_csr_copy": ([c_void_p, _P_I, _P_I, _P_D], c_int),
    "sfa_csr_destroy": ([c_void_p], c_int),
    "sfa_batch_finished": ([c_void_p, _P_I, c_int, POINTER(c_int)], c_int),
    "sfa_batch_collect": ([c_void_p, c_int, _P_D, POINTER(BatchInfo)], c_int),
    "sfa_batch_memory": ([c_void_p, POINTER(c_longlong)], c_int),
    "sfa_batch_steps": ([c_void_p, POINTER(c_longlong)], c_int),
    "sfa_batch_profile": ([c_void_p, c_int, _P_D], c_int),
    "sfa_batch_tune_spmv": ([c_void_p, POINTER(SpmvTuning)], c_int),
    "sfa_batch_set_spmv": ([c_void_p, _P_I], c_int),
    "sfa_batch_destroy": ([c_void_p], c_int),
    "sfa_HodgeConjugateResidualSolverCSR_SSOR":
        ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
         [_P_D, c_double, c_int, c_double, c_int, POINTER(SolveInfo)], c_int),
    "sfa_HodgeConjugateResidualSolverCSR":
        ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
         [_P_D, c_double, c_int, c_double, c_int, POINTER(SolveInfo)], c_int),
    "sfa_HodgeConjugateResidualSolverCSR_diagonal":
        ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
         [_P_D, c_double, c_int, c_double, c_int, POINTER(SolveInfo)], c_int),
    "sfa_HodgeConjugateResidualSolverCSR_textbook":
        ([c_int, c_int, _P_D, _P_D] + [_P_I, _P_I, _P_D] * 4 +
         [_P_D, c_double, c_int, c_double, c_int, POINTER(SolveInfo)], c_int),
    "sfa_hodge_create": ([c_int, c_int] + [_P_I, _P_I, _P_D] * 4 + [POINTER(c_void_p)], c_int),
    "sfa_hodge_create_inner": ([c_int, c_int] + [_P_I, _P_I, _P_D] * 4 + [c_int, POINTER(c_void_p)],
                               c_int),
    "sfa_hodge_solve": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, c_double, c_int,
                         c_int, POINTER(SolveInfo)], c_int),
    "sfa_hodge_destroy": ([c_void_p], c_int),
}

_lib = None


def header_symbols(path: str = HEADER):
    """Function names the C header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfa_\w+)\s*\(", text)))


# SFA_API_VERSION of include/simplefem_amd.h that the structures here follow
API_VERSION = 2


def load(path: str = LIB_PATH):
    """Loads the HIP library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m simplefem_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.sfa_api_version() != API_VERSION:
        raise ImportError(f"{path} has API version {lib.sfa_api_version()}, this binding "
                          f"needs {API_VERSION}: rebuild it with `python -m simplefem_amd.build`")
    if path == LIB_PATH:
        _lib = lib
    return lib


def _check(status: int):
    if status != SFA_OK:
        raise SfaError(status, load().sfa_last_error().decode())


def _i(a):
    return a.ctypes.data_as(_P_I) if a is not None else None


def _d(a):
    return a.ctypes.data_as(_P_D) if a is not None else None


def _f64(a, name, n=None, writable=False):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        if writable:
            raise TypeError(f"{name} must be a C-contiguous float64 numpy array (updated in place)")
        a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.shape[0] < n:
        raise ValueError(f"{name} has {a.shape[0]} entries, needs {n}")
    return a


def _i32(a, name):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.int32)


def _csr(csrrows, csrcolumns, csrvalues):
    return _i32(csrrows, "csrrows"), _i32(csrcolumns, "csrcolumns"), _f64(csrvalues, "csrvalues")


# --------------------------------------------------------------------------- solvers
# Each returns K and fills `info` (a dict) when one is passed.

def _solve(fn_name, N, x, b, A, C, V, residual, precision, print_modulo, extra, info):
    lib = load()
    x = _f64(x, "x", N, writable=True)
    residual = _f64(residual, "residual", N, writable=True)
    b = _f64(b, "b", N)
    A, C, V = _csr(A, C, V)
    out = SolveInfo()
    _check(getattr(lib, fn_name)(N, _d(x), _d(b), _i(A), _i(C), _d(V), _d(residual),
                                 float(precision), int(print_modulo), *extra, ctypes.byref(out)))
    if info is not None:
        info.update(out.as_dict())
    return out.iterations


def ConjugateGradientSolverCSR(N, x, b, csrrows, csrcolumns, csrvalues, residual, precision,
                               print_modulo, info=None):
    """solver/sparsesolver.hpp:11-19"""
    return _solve("sfa_ConjugateGradientSolverCSR", N, x, b, csrrows, csrcolumns, csrvalues,
                  residual, precision, print_modulo, [], info)


def ConjugateGradientSolverCSR_DiagonalPreconditioner(N, x, b, csrrows, csrcolumns, csrvalues,
                                                      residual, precision, print_modulo, precon,
                                                      info=None):
    """solver/sparsesolver.hpp:21-30"""
    precon = _f64(precon, "precon", N)
    return _solve("sfa_ConjugateGradientSolverCSR_DiagonalPreconditioner", N, x, b, csrrows,
                  csrcolumns, csrvalues, residual, precision, print_modulo, [_d(precon)], info)


def ConjugateGradientSolverCSR_SSOR(N, x, b, csrrows, csrcolumns, csrvalues, residual, precision,
                                    print_modulo, diagonal, omega, info=None):
    """solver/sparsesolver.hpp:33-43"""
    diagonal = _f64(diagonal, "diagonal", N)
    return _solve("sfa_ConjugateGradientSolverCSR_SSOR", N, x, b, csrrows, csrcolumns, csrvalues,
                  residual, precision, print_modulo, [_d(diagonal), float(omega)], info)


def ConjugateGradientSolverCSR_SSOR_Eisenstat(N, x, b, csrrows, csrcolumns, csrvalues, residual,
                                              precision, print_modulo, diagonal, omega,
                                              info=None):
    """solver/sparsesolver.hpp:46-56"""
    diagonal = _f64(diagonal, "diagonal", N)
    return _solve("sfa_ConjugateGradientSolverCSR_SSOR_Eisenstat", N, x, b, csrrows, csrcolumns,
                  csrvalues, residual, precision, print_modulo, [_d(diagonal), float(omega)], info)


def ConjugateGradientSolverCSR_Rainbow(N, x, b, csrrows, csrcolumns, csrvalues, residual,
                                       precision, print_modulo, diagonal, omega, num_colors, F, B,
                                       R, info=None):
    """solver/sparsesolver.hpp:59-70"""
    diagonal = _f64(diagonal, "diagonal", N)
    F, B, R = _i32(F, "F"), _i32(B, "B"), _i32(R, "R")
    return _solve("sfa_ConjugateGradientSolverCSR_Rainbow", N, x, b, csrrows, csrcolumns,
                  csrvalues, residual, precision, print_modulo,
                  [_d(diagonal), float(omega), int(num_colors), _i(F), _i(B), _i(R)], info)


def ConjugateGradientSolverCSR_Eisenstat_Rainbow(N, x, b, csrrows, csrcolumns, csrvalues,
                                                 residual, precision, print_modulo, diagonal,
                                                 omega, num_colors, F, B, R, info=None):
    """solver/sparsesolver.hpp:73-84"""
    diagonal = _f64(diagonal, "diagonal", N)
    F, B, R = _i32(F, "F"), _i32(B, "B"), _i32(R, "R")
    return _solve("sfa_ConjugateGradientSolverCSR_Eisenstat_Rainbow", N, x, b, csrrows,
                  csrcolumns, csrvalues, residual, precision, print_modulo,
                  [_d(diagonal), float(omega), int(num_colors), _i(F), _i(B), _i(R)], info)


NOALLOC_MAX_ROWS = 1048576
NOALLOC_MAX_COLORS = 1024


def ConjugateGradientSolverCSR_Rainbow_noalloc(N, x, b, csrrows, csrcolumns, csrvalues, residual,
                                               precision, print_modulo, diagonal, omega,
                                               num_colors, F, B, R, info=None):
    """sfa_ConjugateGradientSolverCSR_Rainbow_noalloc: the arguments of
    ConjugateGradientSolverCSR_Rainbow (solver/sparsesolver.hpp:59-70) as device
    buffers (objects with a ``ptr`` attribute, e.g. hipmem.DeviceArray: float64
    x, b, csrvalues, residual, diagonal; int32 csrrows, csrcolumns, F, B, R).
    The library allocates no device memory for it (its workspace is a static
    device array).  Returns K; x and residual are updated in place."""
    out = SolveInfo()
    lib = load()
    p = lambda a: a.ptr if a is not None else None  # noqa: E731
    _check(lib.sfa_ConjugateGradientSolverCSR_Rainbow_noalloc(
        int(N), p(x), p(b), p(csrrows), p(csrcolumns), p(csrvalues), p(residual),
        float(precision), int(print_modulo), p(diagonal), float(omega), int(num_colors),
        p(F), p(B), p(R), ctypes.byref(out)))
    if info is not None:
        info.update(out.as_dict())
    return out.iterations


# --------------------------------------------------------------------------- CSR operations

def csr_apply(rows, columns, A, C, V, src, scaling=1.0):
    """MatrixCSR::apply (sparse/matcsr.cpp:216-244) on the GPU; returns dest."""
    A, C, V = _csr(A, C, V)
    src = _f64(src, "src", columns)
    dest = np.empty(rows, dtype=np.float64)
    _check(load().sfa_MatrixCSR_apply(rows, columns, _i(A), _i(C), _d(V), _d(dest), _d(src),
                                      float(scaling)))
    return dest


def _take_csr(h):
    """(rows, cols, A, C, V) of an sfa_csr result; frees it."""
    try:
        rows, cols, nnz = c_int(0), c_int(0), c_int(0)
        _check(load().sfa_csr_sizes(h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nnz)))
        A = np.empty(rows.value + 1, dtype=np.int32)
        C = np.empty(nnz.value, dtype=np.int32)
        V = np.empty(nnz.value, dtype=np.float64)
        _check(load().sfa_csr_copy(h, _i(A), _i(C), _d(V)))
        return rows.value, cols.value, A, C, V
    finally:
        load().sfa_csr_destroy(h)


def csr_multiply(m1, m2, reduced=False):
    """MatrixCSRMultiplication (sparse/matcsr.cpp:767-849), or with reduced=True
    MatrixCSRMultiplication_reduced (:857-962), on the GPU, bit for bit with
    the reference; m1, m2 and the result are (rows, cols, A, C, V)."""
    r1, c1, A1, C1, V1 = m1
    r2, c2, A2, C2, V2 = m2
    A1, C1, V1 = _csr(A1, C1, V1)
    A2, C2, V2 = _csr(A2, C2, V2)
    h = c_void_p()
    fn = (load().sfa_MatrixCSRMultiplication_reduced if reduced
          else load().sfa_MatrixCSRMultiplication)
    _check(fn(int(r1), int(c1), _i(A1), _i(C1), _d(V1), int(r2), int(c2), _i(A2), _i(C2), _d(V2),
              ctypes.byref(h)))
    return _take_csr(h)


def csr_transpose(m):
    """MatrixCSR::getTranspose (sparse/matcsr.cpp:291-350) with its compression
    on the GPU, bit for bit with the reference; (rows, cols, A, C, V)."""
    r, c, A, C, V = m
    A, C, V = _csr(A, C, V)
    h = c_void_p()
    _check(load().sfa_MatrixCSR_getTranspose(int(r), int(c), _i(A), _i(C), _d(V), ctypes.byref(h)))
    return _take_csr(h)


def get_diagonal(n, A, C, V):
    """MatrixCSR::getDiagonal (sparse/matcsr.cpp:266-286)."""
    A, C, V = _csr(A, C, V)
    diag = np.empty(n, dtype=np.float64)
    _check(load().sfa_MatrixCSR_getDiagonal(n, _i(A), _i(C), _d(V), _d(diag)))
    return diag


def grid_barrier_cost(blocks, threads=256, rounds=2000, mode=1):
    """sfa_grid_barrier_cost: (microseconds per round, error bits)."""
    us = c_double(0.0)
    err = c_int(0)
    _check(load().sfa_grid_barrier_cost(int(blocks), int(threads), int(rounds), int(mode),
                                        ctypes.byref(us), ctypes.byref(err)))
    return us.value, err.value


class Rainbow:
    """Rainbow::Rainbow (sparse/rainbow.hpp:13-31): num_rows, num_colors, F, B, R."""

    def __init__(self, n, A, C):
        A, C = _i32(A, "A"), _i32(C, "C")
        self.num_rows = n
        self.F = np.empty(n, dtype=np.int32)
        B = np.empty(n + 1, dtype=np.int32)
        self.R = np.empty(n, dtype=np.int32)
        nc = c_int(0)
        _check(load().sfa_Rainbow(n, _i(A), _i(C), _i(self.F), _i(B), _i(self.R), ctypes.byref(nc)))
        self.num_colors = nc.value
        self.B = B[: self.num_colors + 1].copy()


# --------------------------------------------------------------------------- device systems

class System:
    """A matrix + preconditioner schedule resident in HBM (sfa_system_*).

    ``solve`` accepts numpy arrays (host memory) or device buffers
    (``simplefem_amd.hipmem.DeviceArray``, float64).
    """

    def __init__(self, method, N, A, C, V, aux=None, omega=1.0, num_colors=0, F=None, B=None,
                 R=None, fp32_inner=False, paired_sweeps=False, inner_restarts=False,
                 borrow_csr=False):
        lib = load()
        if fp32_inner:
            method |= FLAG_FP32_INNER
        if paired_sweeps:
            method |= FLAG_PAIRED_SWEEPS
        if inner_restarts:
            method |= FLAG_INNER_RESTARTS
        A, C, V = _csr(A, C, V)
        # borrow_csr: the library keeps pointers to the CSR arrays instead of
        # a copy; this object holds them (as passed to it, or their converted
        # copies) until it is closed, and they must not be changed meanwhile
        self._csr = None
        if borrow_csr:
            method |= FLAG_BORROW_CSR
            self._csr = (A, C, V)
        aux = _f64(aux, "aux", N)
        F, B, R = _i32(F, "F"), _i32(B, "B"), _i32(R, "R")
        self._h = c_void_p()
        self.N = N
        _check(lib.sfa_system_create(method, N, _i(A), _i(C), _d(V), _d(aux), float(omega),
                                     int(num_colors), _i(F), _i(B), _i(R), ctypes.byref(self._h)))

    @staticmethod
    def _ptr(a, N, writable):
        if isinstance(a, np.ndarray):
            a = _f64(a, "vector", N, writable)
            return a, a.ctypes.data, 0
        # a device buffer: anything with .address (simplefem_amd.hipmem.DeviceArray)
        if getattr(a, "n", N) < N or getattr(a, "dtype", np.float64) != np.float64:
            raise TypeError("device vectors must be float64 with at least N entries")
        return a, a.address, 1

    def solve(self, x, b, residual, precision, print_modulo=-3, info=None):
        x, px, dx = self._ptr(x, self.N, True)
        b, pb, db = self._ptr(b, self.N, False)
        residual, pr, dr = self._ptr(residual, self.N, True)
        if not (dx == db == dr):
            raise TypeError("x, b and residual must all be host arrays or all device tensors")
        out = SolveInfo()
        _check(load().sfa_system_solve(self._h, px, pb, pr, float(precision), int(print_modulo), dx,
                                       ctypes.byref(out)))
        if info is not None:
            info.update(out.as_dict())
        return out.iterations

    def solve_mixed(self, x, b, rtol, inner_rtol=1e-1, max_outer=50):
        """FP64 refinement around FP32 inner solves (system created with
        fp32_inner=True).  Returns a dict (outer, inner_iterations, success,
        continued, residual_norm, relative_residual); `continued` counts the
        inner solves that continued the previous one's direction."""
        x, px, dx = self._ptr(x, self.N, True)
        b, pb, db = self._ptr(b, self.N, False)
        if dx != db:
            raise TypeError("x and b must both be host arrays or both device buffers")
        out = RefineInfo()
        _check(load().sfa_system_solve_mixed(self._h, px, pb, float(rtol), float(inner_rtol),
                                             int(max_outer), dx, ctypes.byref(out)))
        return out.as_dict()

    def time_sweeps(self, kernel_class, reps=200):
        """Microseconds per launch of the forward (KERNEL_CLASS "forward") or
        backward sweep kernels, back to back in a recorded graph
        (sfa_system_time_sweeps); after a solve."""
        klass = KERNEL_NAMES.index(kernel_class) if isinstance(kernel_class, str) else kernel_class
        us = c_double(0.0)
        _check(load().sfa_system_time_sweeps(self._h, int(klass), int(reps), ctypes.byref(us)))
        return us.value

    def profile(self, x, b, residual, precision):
        """One solve with every kernel timed by HIP events; per-class totals."""
        x, px, dx = self._ptr(x, self.N, True)
        b, pb, db = self._ptr(b, self.N, False)
        residual, pr, dr = self._ptr(residual, self.N, True)
        if not (dx == db == dr):
            raise TypeError("x, b and residual must all be host arrays or all device buffers")
        out = (KernelProfile * KERNEL_CLASSES)()
        info = SolveInfo()
        _check(load().sfa_system_profile(self._h, px, pb, pr, float(precision), dx, out,
                                         ctypes.byref(info)))
        prof = {KERNEL_NAMES[c]: {"launches": out[c].launches, "ms": out[c].milliseconds,
                                  "bytes": out[c].bytes} for c in range(KERNEL_CLASSES)}
        return prof, info.as_dict()

    def apply(self, dest, src, scaling=1.0):
        dest, pd, dd = self._ptr(dest, self.N, True)
        src, ps, ds = self._ptr(src, self.N, False)
        if dd != ds:
            raise TypeError("dest and src must both be host arrays or both device tensors")
        _check(load().sfa_system_apply(self._h, pd, ps, float(scaling), dd))

    def set_stream(self, stream_handle):
        _check(load().sfa_system_set_stream(self._h, c_void_p(stream_handle)))

    def tune_spmv(self, fp32_inner=False):
        """Times every SpMV mapping (lanes per row) on this system and installs
        the fastest; returns the per-bucket report (SpmvTuning.as_dict)."""
        t = SpmvTuning()
        _check(load().sfa_system_tune_spmv(self._h, 1 if fp32_inner else 0, ctypes.byref(t)))
        return t.as_dict()

    def set_spmv(self, lanes_per_row):
        _check(load().sfa_system_set_spmv(self._h, int(lanes_per_row)))

    def stats(self):
        st = SystemStats()
        _check(load().sfa_system_query(self._h, ctypes.byref(st)))
        out = {k: getattr(st, k) for k, _ in st._fields_ if k != "reserved"}
        out["setup_ms"] = list(out["setup_ms"])
        return out

    def close(self):
        if self._h:
            load().sfa_system_destroy(self._h)
            self._h = c_void_p()
        self._csr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many Rainbow CG-SSOR solves advanced together by one persistent kernel
    (sfa_batch_*): add matrices, submit right-hand sides, run, collect."""

    def __init__(self, mixed=False, max_slots=256, inner_restarts=False):
        self._h = c_void_p()
        self.mixed = mixed
        mode = (BATCH_MIXED_RESTARTS if inner_restarts else BATCH_MIXED) if mixed else BATCH_FP64
        _check(load().sfa_batch_create(mode, int(max_slots), ctypes.byref(self._h)))
        self.n_of = {}

    def add_matrix(self, N, A, C, V, diagonal, omega, num_colors, F, B, R):
        A, C, V = _csr(A, C, V)
        diagonal = _f64(diagonal, "diagonal", N)
        F, B, R = _i32(F, "F"), _i32(B, "B"), _i32(R, "R")
        mid = c_int(-1)
        _check(load().sfa_batch_add_matrix(self._h, N, _i(A), _i(C), _d(V), _d(diagonal),
                                           float(omega), int(num_colors), _i(F), _i(B), _i(R),
                                           ctypes.byref(mid)))
        self.n_of[mid.value] = N
        return mid.value

    def add_prepared(self, prepared):
        """Uploads a PreparedMatrix; returns its matrix id."""
        mid = c_int(-1)
        _check(load().sfa_batch_add_prepared(self._h, prepared._h, ctypes.byref(mid)))
        self.n_of[mid.value] = prepared.N
        return mid.value

    def remove_matrix(self, mid):
        _check(load().sfa_batch_remove_matrix(self._h, int(mid)))
        self.n_of.pop(mid, None)

    def submit(self, mid, b, x0=None, precision=1e-10, inner_rtol=1e-1, max_outer=50):
        N = self.n_of[mid]
        b = _f64(b, "b", N)
        x0 = _f64(x0, "x0", N)
        slot = c_int(-1)
        _check(load().sfa_batch_submit(self._h, int(mid), _d(b), _d(x0), float(precision),
                                       float(inner_rtol), int(max_outer), ctypes.byref(slot)))
        return slot.value

    def submit_group(self, mid, bs, x0s=None, precision=1e-10, inner_rtol=1e-1, max_outer=50):
        """Up to four right-hand sides of one matrix in consecutive slots
        whose vectors are interleaved (sfa_batch_submit_group); returns their
        slots in order.  Each result equals that of submit()."""
        N = self.n_of[mid]
        bs = [_f64(b, "b", N)[:N] for b in bs]
        k = len(bs)
        if not 1 <= k <= 4:
            raise ValueError("a group holds 1 to 4 right-hand sides")
        flat = np.ascontiguousarray(np.concatenate(bs))
        x0 = None
        if x0s is not None:
            if len(x0s) != k:
                raise ValueError("one initial guess per right-hand side")
            x0 = np.ascontiguousarray(np.concatenate([_f64(x, "x0", N)[:N] for x in x0s]))
        out = np.full(k, -1, dtype=np.int32)
        _check(load().sfa_batch_submit_group(self._h, int(mid), k, _d(flat), _d(x0),
                                             float(precision), float(inner_rtol),
                                             int(max_outer), _i(out)))
        return out.tolist()

    def run(self, max_steps=0):
        running = c_int(0)
        _check(load().sfa_batch_run(self._h, int(max_steps), ctypes.byref(running)))
        return running.value

    def finished(self, capacity=1024):
        buf = np.empty(capacity, dtype=np.int32)
        cnt = c_int(0)
        _check(load().sfa_batch_finished(self._h, _i(buf), capacity, ctypes.byref(cnt)))
        return buf[: cnt.value].tolist()

    def collect(self, slot, N):
        x = np.empty(N)
        info = BatchInfo()
        _check(load().sfa_batch_collect(self._h, int(slot), _d(x), ctypes.byref(info)))
        return x, info.as_dict()

    def profile(self, steps):
        """Runs `steps` steps kernel by kernel between HIP events:
        {"step_ms", "fold_ms", "prefix_ms"} summed over the steps."""
        ms = np.zeros(3)
        _check(load().sfa_batch_profile(self._h, int(steps), _d(ms)))
        return {"step_ms": ms[0], "fold_ms": ms[1], "prefix_ms": ms[2]}

    def steps(self):
        """Engine steps executed so far."""
        n = c_longlong(0)
        _check(load().sfa_batch_steps(self._h, ctypes.byref(n)))
        return n.value

    def tune_spmv(self):
        """Per-bucket SpMV autotuning over the batch's matrices (between solves)."""
        t = SpmvTuning()
        _check(load().sfa_batch_tune_spmv(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_spmv(self, lanes_per_row):
        lanes = np.ascontiguousarray(lanes_per_row, dtype=np.int32)
        if lanes.shape != (SPMV_BUCKETS,):
            raise ValueError("one mapping per bucket (%d)" % SPMV_BUCKETS)
        _check(load().sfa_batch_set_spmv(self._h, _i(lanes)))

    def memory(self):
        b = c_longlong(0)
        _check(load().sfa_batch_memory(self._h, ctypes.byref(b)))
        return b.value

    def close(self):
        if self._h:
            load().sfa_batch_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --------------------------------------------------------------------------- block systems

def _blocks(N, L, A, B, Bt, C):
    out = []
    for (r, c, v), rows in ((A, L), (B, N), (Bt, L), (C, N)):
        r, c, v = _csr(r, c, v)
        if r.shape[0] != rows + 1:
            raise ValueError(f"a block has {r.shape[0] - 1} rows, needs {rows}")
        out += [r, c, v]
    return out


HODGE_INNER_SSOR = 0
HODGE_INNER_DIAGONAL = 1
HODGE_INNER_TEXTBOOK = 2


def _hodge_once(entry, N, L, x, b, A, B, Bt, C, residual, precision, print_modulo,
                inneriteration_precision, inneriteration_print_modulo, info):
    lib = load()
    x = _f64(x, "x", N, writable=True)
    residual = _f64(residual, "residual", N, writable=True)
    b = _f64(b, "b", N)
    mats = _blocks(N, L, A, B, Bt, C)
    ptrs = [(_i if k % 3 < 2 else _d)(m) for k, m in enumerate(mats)]
    out = SolveInfo()
    _check(getattr(lib, entry)(
        int(N), int(L), _d(x), _d(b), *ptrs, _d(residual), float(precision), int(print_modulo),
        float(inneriteration_precision), int(inneriteration_print_modulo), ctypes.byref(out)))
    if info is not None:
        info.update(out.as_dict())
    return out.iterations


def HodgeConjugateResidualSolverCSR_SSOR(N, L, x, b, A, B, Bt, C, residual, precision,
                                         print_modulo, inneriteration_precision,
                                         inneriteration_print_modulo, info=None):
    """solver/systemsparsesolver.hpp:58-72: (B A^-1 Bt + C) x = b by conjugate
    residuals with inner CG-SSOR solves of A.  A, B, Bt, C are
    (rows, cols, vals) CSR triples.  Returns K."""
    return _hodge_once("sfa_HodgeConjugateResidualSolverCSR_SSOR", N, L, x, b, A, B, Bt, C,
                       residual, precision, print_modulo, inneriteration_precision,
                       inneriteration_print_modulo, info)


def HodgeConjugateResidualSolverCSR(N, L, x, b, A, B, Bt, C, residual, precision, print_modulo,
                                    inneriteration_precision, inneriteration_print_modulo,
                                    info=None):
    """solver/systemsparsesolver.hpp:10-24: the SSOR variant, as the reference's
    plain variant calls it (solver/systemsparsesolver.cpp:18-49).  Returns K."""
    return _hodge_once("sfa_HodgeConjugateResidualSolverCSR", N, L, x, b, A, B, Bt, C, residual,
                       precision, print_modulo, inneriteration_precision,
                       inneriteration_print_modulo, info)


def HodgeConjugateResidualSolverCSR_diagonal(N, L, x, b, A, B, Bt, C, residual, precision,
                                             print_modulo, inneriteration_precision,
                                             inneriteration_print_modulo, info=None):
    """solver/systemsparsesolver.hpp:26-40: the same conjugate residual loop
    with inner Jacobi CG solves of A (precon = 1 / diag A).  Returns K."""
    return _hodge_once("sfa_HodgeConjugateResidualSolverCSR_diagonal", N, L, x, b, A, B, Bt, C,
                       residual, precision, print_modulo, inneriteration_precision,
                       inneriteration_print_modulo, info)


def HodgeConjugateResidualSolverCSR_textbook(N, L, x, b, A, B, Bt, C, residual, precision,
                                             print_modulo, inneriteration_precision,
                                             inneriteration_print_modulo, info=None):
    """solver/systemsparsesolver.hpp:42-56: the same conjugate residual loop
    with inner plain CG solves of A.  Returns K."""
    return _hodge_once("sfa_HodgeConjugateResidualSolverCSR_textbook", N, L, x, b, A, B, Bt, C,
                       residual, precision, print_modulo, inneriteration_precision,
                       inneriteration_print_modulo, info)


class Hodge:
    """A block system kept on the device between solves (sfa_hodge_*);
    inner="diagonal" solves with inner Jacobi CG as
    HodgeConjugateResidualSolverCSR_diagonal does, inner="textbook" with inner
    plain CG as HodgeConjugateResidualSolverCSR_textbook does."""

    _INNER = {"ssor": 0, "diagonal": 1, "textbook": 2}

    def __init__(self, N, L, A, B, Bt, C, inner="ssor"):
        if inner not in self._INNER:
            raise ValueError(f"inner must be 'ssor', 'diagonal' or 'textbook', not {inner!r}")
        self._h = c_void_p()
        self.N, self.L = int(N), int(L)
        mats = _blocks(self.N, self.L, A, B, Bt, C)
        ptrs = [(_i if k % 3 < 2 else _d)(m) for k, m in enumerate(mats)]
        kind = self._INNER[inner]
        _check(load().sfa_hodge_create_inner(self.N, self.L, *ptrs, kind, ctypes.byref(self._h)))

    def solve(self, x, b, residual, precision=1e-10, print_modulo=-3,
              inneriteration_precision=1e-10, inneriteration_print_modulo=-3):
        """x, b, residual: all host float64 arrays or all device buffers
        (simplefem_amd.hipmem.DeviceArray).  Returns the info dict."""
        x, px, dx = System._ptr(x, self.N, True)
        b, pb, db = System._ptr(b, self.N, False)
        residual, pr, dr = System._ptr(residual, self.N, True)
        if not (dx == db == dr):
            raise TypeError("x, b and residual must all be host arrays or all device buffers")
        out = SolveInfo()
        _check(load().sfa_hodge_solve(self._h, px, pb, pr, float(precision), int(print_modulo),
                                      float(inneriteration_precision),
                                      int(inneriteration_print_modulo), dx, ctypes.byref(out)))
        return out.as_dict()

    def close(self):
        if self._h:
            load().sfa_hodge_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
