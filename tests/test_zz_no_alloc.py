"""The no-allocation Rainbow solve (sfa_ConjugateGradientSolverCSR_Rainbow_noalloc)
against the drop-in sfa_ConjugateGradientSolverCSR_Rainbow, through the C-ABI.

Bar (the GPU parity bar): K within max(1, 2% of K), the same exit reason,
x and residual within 1e-10 relative (inf-norm; the residual on the scale of
b).  The two calls in a row share the GPU's static scratch array; the second
is smaller, so stale contents of the first would show.
"""
import numpy as np
import pytest

from simplefem_amd import capi
from simplefem_amd import systems as S
from simplefem_amd.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

REL = 1e-10


def _close(a, b, rel=REL, scale=None):
    scale = max(np.abs(b).max(), scale or 0.0, 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rel, f"relative error {err:.3e} > {rel:.1e}"


def _inputs(system, seed, x0_scale=0.0):
    b = S.rhs_uniform(system, seed)
    rng = np.random.default_rng(seed)
    x0 = x0_scale * rng.uniform(-1, 1, system.n)
    r0 = rng.uniform(-1, 1, system.n)  # rows the solver skips keep these
    return b, x0, r0


def _both(system, b, x0, r0, omega):
    n = system.n
    d = system.diagonal()
    rb = capi.Rainbow(n, system.A, system.C)
    col = (rb.num_colors, rb.F, rb.B, rb.R)
    x_ref, r_ref, i_ref = x0.copy(), r0.copy(), {}
    capi.ConjugateGradientSolverCSR_Rainbow(n, x_ref, b, system.A, system.C, system.V, r_ref,
                                            1e-10, -3, d, omega, *col, info=i_ref)
    dev = {k: DeviceArray.from_host(v) for k, v in {
        "x": x0, "b": b, "A": system.A.astype(np.int32), "C": system.C.astype(np.int32),
        "V": system.V, "r": r0, "d": d, "F": rb.F.astype(np.int32),
        "B": rb.B.astype(np.int32), "R": rb.R.astype(np.int32)}.items()}
    synchronize()
    i_na = {}
    K = capi.ConjugateGradientSolverCSR_Rainbow_noalloc(
        n, dev["x"], dev["b"], dev["A"], dev["C"], dev["V"], dev["r"], 1e-10, -3, dev["d"],
        omega, rb.num_colors, dev["F"], dev["B"], dev["R"], info=i_na)
    x_na, r_na = dev["x"].download(), dev["r"].download()
    for a in dev.values():
        a.free()
    assert K == i_na["iterations"]
    return (x_ref, r_ref, i_ref), (x_na, r_na, i_na)


def _check(system, seed, omega=1.0, x0_scale=0.0):
    b, x0, r0 = _inputs(system, seed, x0_scale)
    (xr, rr, ir), (xn, rn, inn) = _both(system, b, x0, r0, omega)
    print(f"{system.name}: K drop-in {ir['iterations']} no-alloc {inn['iterations']}", flush=True)
    Kr = ir["iterations"]
    assert abs(inn["iterations"] - Kr) <= max(1, int(0.02 * Kr)), (ir, inn)
    assert inn["exit_reason"] == ir["exit_reason"], (ir, inn)
    assert inn["success"] == ir["success"]
    _close(xn, xr)
    _close(rn, rr, scale=np.abs(b).max())


def test_two_calls_in_a_row_second_smaller():
    print("case: no-alloc kuhn_10 then p1_2d_L4", flush=True)
    _check(S.poisson3d_p1_kuhn(10), seed=11)
    _check(S.poisson2d_p1(4), seed=12, omega=1.2, x0_scale=0.5)


def test_full_size_64_cube():
    print("case: no-alloc 64^3, the reference's vertex numbering", flush=True)
    _check(S.poisson3d_p1_kuhn(64, numbering="reference"), seed=2026)


def test_rejects_sizes_beyond_the_maximum():
    # rejected on the host before anything touches the device
    n = capi.NOALLOC_MAX_ROWS + 1
    with pytest.raises(capi.SfaError, match="SFA_NOALLOC_MAX_ROWS"):
        capi.ConjugateGradientSolverCSR_Rainbow_noalloc(
            n, *([DeviceArray(1)] * 6), 1e-10, -3, DeviceArray(1), 1.0, 1,
            *([DeviceArray(1, np.int32)] * 3))


def test_rejects_an_invalid_colouring():
    system = S.poisson2d_p1(3)
    rb = capi.Rainbow(system.n, system.A, system.C)
    R = rb.R.copy()
    R[1] = R[0]  # not a permutation
    b, x0, r0 = _inputs(system, 3)
    dev = [DeviceArray.from_host(v) for v in (x0, b, system.A.astype(np.int32),
                                              system.C.astype(np.int32), system.V, r0,
                                              system.diagonal(), rb.F.astype(np.int32),
                                              rb.B.astype(np.int32), R.astype(np.int32))]
    with pytest.raises(capi.SfaError, match="SFA_ERR_COLOURING"):
        capi.ConjugateGradientSolverCSR_Rainbow_noalloc(
            system.n, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], 1e-10, -3, dev[6], 1.0,
            rb.num_colors, dev[7], dev[8], dev[9])
    # a valid call afterwards still matches the drop-in
    _check(system, seed=4)
