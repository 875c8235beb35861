"""MI355X (gfx950) sparse CG / SSOR / Rainbow-SSOR solve path of FEEC++.

The compute path is the HIP library libsimplefem_amd.so (csrc/), reached
through its C-ABI (include/simplefem_amd.h).  ``capi`` binds it for Python
with the reference's function names; ``systems`` builds the synthetic CSR
systems used by the tests and the benchmark.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
