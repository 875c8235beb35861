"""Builds libsimplefem_amd.so (and the C++ drop-in library) for gfx950 in-tree.

    python -m simplefem_amd.build          # or simplefem_amd.build.build()

hipcc cross-compiles for gfx950 without a GPU present.  Flags:
  -ffp-contract=off        products are rounded before they are added, as in
                           the reference's loops, so per-row sums agree bit
                           for bit with the CPU path;
  -packed-fp32-ops off     no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsimplefem_amd.so")
CXXLIB = os.path.join(HERE, "libsimplefem_amd_cxx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SFA_OFFLOAD_ARCH", "gfx950")

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [  # (source, compile as HIP)
    ("kernels.hip", True),
    ("batch.hip", True),
    ("hodge.hip", True),
    ("noalloc.hip", True),
    ("barrier.hip", True),
    ("assembly.hip", True),
    ("system.cpp", True),
    ("plan.cpp", False),
    ("pool.cpp", True),
]
CXX_SOURCES = ["dropin.cpp"]
DROPIN_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
DROPIN_TEST_BIN = os.path.join(OBJ, "dropin_test")
SPARSE_OPS_SRC = os.path.join(ROOT, "tests", "cpp", "sparse_ops_test.cpp")
SPARSE_OPS_BIN = os.path.join(OBJ, "sparse_ops_test")


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + proc.stdout + proc.stderr)
    return proc


def _compile(src, hip):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if hip:
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", path, "-o", out]
    else:  # host-only sources; the plan builder runs its passes with OpenMP
        cmd = ["g++"] + COMMON + ["-fopenmp", "-c", path, "-o", out]
    _run(cmd)
    return out


def build(verbose: bool = False) -> str:
    """Compiles every source and links the shared libraries; returns LIB."""
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as pool:
        objs = list(pool.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs
         + ["-Wl,-soname,libsimplefem_amd.so", "-lgomp"])
    os.replace(tmp, LIB)
    cxx_objs = []
    for src in CXX_SOURCES:
        if os.path.exists(os.path.join(CSRC, src)):
            cxx_objs.append(_compile(src, False))
    if cxx_objs:
        tmp = CXXLIB + ".tmp"
        _run(["g++", "-shared", "-o", tmp] + cxx_objs
             + ["-L" + HERE, "-lsimplefem_amd", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, CXXLIB)
        # a caller written like the reference's tests, against the drop-in headers
        for src, binary in ((DROPIN_TEST_SRC, DROPIN_TEST_BIN), (SPARSE_OPS_SRC, SPARSE_OPS_BIN)):
            _run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), src,
                  "-o", binary, "-L" + HERE, "-lsimplefem_amd_cxx", "-lsimplefem_amd",
                  "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
