"""SpMV mappings (K = 1, 2, 4, 8 lanes per row over SELL-64 slices) and the
autotuner that picks them.

Bars: every mapping gives MatrixCSR::apply bit for bit (a row's products are
added in entry order whatever lane formed them); solves with any mapping
match the oracle to the single-system bar (K within max(1, 2%), x within
1e-10 relative); the tuners time every candidate and install the fastest,
and the results after tuning meet the same bars.  Matrices cover the four
row-length buckets (<= 8, <= 16, <= 32, > 32 entries) and rows longer than
one chunk of every mapping (> 64 entries).
"""
import numpy as np
import pytest

import oracle
from simplefem_amd import capi
from simplefem_amd import systems as S
from simplefem_amd.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

MAPPINGS = (1, 2, 4, 8)


def _bucket_matrices():
    return [S.poisson2d_p1(4),          # <= 7 entries per row
            S.poisson3d_p1_kuhn(6),     # <= 15
            S.random_spd(400, 0.04, 3),  # about 17 on average, up to 32
            S.random_spd(300, 0.5, 4)]   # about 150: several chunks per lane


def _iter_bar(k, ko):
    return abs(k - ko) <= max(1, int(0.02 * ko))


@pytest.mark.parametrize("lanes", MAPPINGS)
def test_apply_bitexact_every_mapping(lanes):
    for s in _bucket_matrices():
        print(f"[apply] {s.name} lanes={lanes}", flush=True)
        src = S.uniform_pm1(11, s.n)
        want = oracle.csr_apply(s.n, s.A, s.C, s.V, src, -0.75)
        with capi.System(capi.METHOD_CG, s.n, s.A, s.C, s.V) as sysd:
            sysd.set_spmv(lanes)
            assert sysd.stats()["spmv_lanes_per_row"] == lanes
            q = DeviceArray(s.n)
            sysd.apply(q, DeviceArray.from_host(src), -0.75)
            assert np.array_equal(q.download(), want), s.name


def _solve(method, s, b, aux=None, col=None, lanes=None, tune=False):
    args = (method, s.n, s.A, s.C, s.V, aux, 1.0) + (tuple(col) if col else ())
    with capi.System(*args) as sysd:
        report = None
        if lanes is not None:
            sysd.set_spmv(lanes)
        if tune:
            report = sysd.tune_spmv()
        x = np.zeros(s.n)
        r = np.zeros(s.n)
        info = {}
        sysd.solve(x, b, r, 1e-10, info=info)
        return x, info, report, sysd.stats()


def _check(variant, s, b, x, info, aux=None, col=None):
    xo, _, io = oracle.solve(variant, s.A, s.C, s.V, b, aux=aux, colouring=col)
    assert _iter_bar(info["iterations"], io["iterations"]), (s.name, variant, info, io)
    assert np.abs(x - xo).max() <= 1e-10 * max(np.abs(xo).max(), 1e-300), (s.name, variant)


@pytest.mark.parametrize("lanes", MAPPINGS)
def test_solvers_every_mapping(lanes):
    s3 = S.poisson3d_p1_kuhn(8)
    b3 = S.rhs_uniform(s3, 5)
    d3 = s3.diagonal()
    col = oracle.rainbow(s3.n, s3.A, s3.C)
    print(f"[solve] rainbow lanes={lanes}", flush=True)
    x, info, _, _ = _solve(capi.METHOD_RAINBOW, s3, b3, d3, col, lanes)
    _check("rainbow", s3, b3, x, info, d3, col)
    s2 = S.poisson2d_p1(5)
    b2 = S.rhs_uniform(s2, 6)
    d2 = s2.diagonal()
    print(f"[solve] ssor lanes={lanes}", flush=True)
    x, info, _, _ = _solve(capi.METHOD_SSOR, s2, b2, d2, None, lanes)
    _check("ssor", s2, b2, x, info, d2)
    sr = S.random_spd(300, 0.3, 8)
    br = S.uniform_pm1(9, sr.n)
    print(f"[solve] cg lanes={lanes}", flush=True)
    x, info, _, _ = _solve(capi.METHOD_CG, sr, br, None, None, lanes)
    _check("cg", sr, br, x, info)
    dr = sr.diagonal()
    pre = np.where(dr != 0, 1.0 / np.where(dr != 0, dr, 1.0), 0.0)
    print(f"[solve] jacobi lanes={lanes}", flush=True)
    x, info, _, _ = _solve(capi.METHOD_DIAGONAL, sr, br, pre, None, lanes)
    _check("jacobi", sr, br, x, info, pre)


def test_system_tuner_installs_the_fastest():
    s = S.poisson3d_p1_kuhn(12)
    b = S.rhs_uniform(s, 2)
    d = s.diagonal()
    col = oracle.rainbow(s.n, s.A, s.C)
    x, info, report, stats = _solve(capi.METHOD_RAINBOW, s, b, d, col, tune=True)
    print(f"[tune] {report}", flush=True)
    assert list(report) == ["<=16"], report  # 15-entry rows
    entry = report["<=16"]
    assert sorted(entry["us"]) == list(MAPPINGS) and min(entry["us"].values()) > 0
    best = min(entry["us"], key=entry["us"].get)
    assert entry["lanes_per_row"] == best == stats["spmv_lanes_per_row"]
    _check("rainbow", s, b, x, info, d, col)


def test_system_tuner_fp32_inner():
    s = S.poisson3d_p1_kuhn(8)
    d = s.diagonal()
    col = oracle.rainbow(s.n, s.A, s.C)
    b = S.rhs_uniform(s, 3)
    with capi.System(capi.METHOD_RAINBOW | capi.FLAG_FP32_INNER, s.n, s.A, s.C, s.V, d, 1.0,
                     *col) as sysd:
        rep = sysd.tune_spmv(fp32_inner=True)
        assert rep["<=16"]["lanes_per_row"] in MAPPINGS
        x = np.zeros(s.n)
        info = sysd.solve_mixed(x, b, 1e-10)
        res = b - oracle.csr_apply(s.n, s.A, s.C, s.V, x)
        assert np.linalg.norm(res) <= max(1e-10 * np.linalg.norm(b), 2.0 ** -26), info
    with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0, *col) as plain:
        with pytest.raises(capi.SfaError):
            plain.tune_spmv(fp32_inner=True)
        with pytest.raises(capi.SfaError):
            plain.set_spmv(3)


def _batch_jobs(batch, systems):
    jobs = []
    for k, s in enumerate(systems):
        col = oracle.rainbow(s.n, s.A, s.C)
        mid = batch.add_matrix(s.n, s.A, s.C, s.V, s.diagonal(), 1.0, *col)
        b = S.rhs_uniform(s, 40 + k)
        jobs.append((mid, s, b, col))
    return jobs


def _batch_run_check(batch, jobs):
    slots = [(batch.submit(mid, b), s, b, col) for mid, s, b, col in jobs]
    while batch.run(0) > 0:
        pass
    assert set(batch.finished()) == {sl for sl, *_ in slots}
    for sl, s, b, col in slots:
        x, info = batch.collect(sl, s.n)
        xo, _, io = oracle.solve("rainbow", s.A, s.C, s.V, b, aux=s.diagonal(), colouring=col)
        assert _iter_bar(info["iterations"], io["iterations"]), (s.name, info, io)
        assert np.abs(x - xo).max() <= 1e-10 * max(np.abs(xo).max(), 1e-300), s.name


@pytest.mark.parametrize("table", [(1, 1, 1, 1), (2, 2, 2, 2), (4, 4, 4, 4), (8, 8, 8, 8),
                                   (8, 4, 2, 1)])
def test_batch_every_mapping(table):
    print(f"[batch] table={table}", flush=True)
    with capi.Batch(mixed=False, max_slots=16) as batch:
        jobs = _batch_jobs(batch, _bucket_matrices())
        batch.set_spmv(table)
        _batch_run_check(batch, jobs)


def test_batch_tuner_per_bucket():
    with capi.Batch(mixed=False, max_slots=16) as batch:
        systems = _bucket_matrices()
        jobs = _batch_jobs(batch, systems)
        report = batch.tune_spmv()
        print(f"[batch tune] {report}", flush=True)
        for name, entry in report.items():
            assert entry["lanes_per_row"] in MAPPINGS
            if entry["us"]:  # buckets that hold a matrix were timed
                assert sorted(entry["us"]) == list(MAPPINGS)
                assert entry["lanes_per_row"] == min(entry["us"], key=entry["us"].get)
        assert sum(1 for e in report.values() if e["us"]) == 4  # one matrix per bucket
        _batch_run_check(batch, jobs)
        # not while solves are running
        s = systems[0]
        batch.submit(jobs[0][0], S.rhs_uniform(s, 1))
        with pytest.raises(capi.SfaError):
            batch.tune_spmv()
        with pytest.raises(capi.SfaError):
            batch.set_spmv((1, 1, 1, 1))
        while batch.run(0) > 0:
            pass
