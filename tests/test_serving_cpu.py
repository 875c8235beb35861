"""Host-side logic of the solve server (simplefem_amd.serving) that needs no
GPU: routing, the byte estimates the admission budget charges, the residual
target and the status rule, and the latency/throughput summary."""
import numpy as np
import pytest

from simplefem_amd import serving
from simplefem_amd import systems as S


def _req(system, arrival=0.0, rid=-1):
    return serving.Request(system, S.rhs_uniform(system, 1), arrival, rid)


def test_routes_by_rows_and_workers():
    small, large = S.poisson2d_p1(4), S.poisson2d_p1(7)  # 289 and 16,641 rows
    srv = serving.SolveServer(route_rows=10000)
    assert srv.route_of(_req(small)) == "batch"
    assert srv.route_of(_req(large)) == "dedicated"
    assert serving.SolveServer(route_rows=large.n).route_of(_req(large)) == "dedicated"
    assert serving.SolveServer(route_rows=large.n + 1).route_of(_req(large)) == "batch"
    # without dedicated workers every request is batched
    assert serving.SolveServer(route_rows=10, dedicated_workers=0).route_of(_req(large)) == "batch"


def test_byte_estimates_follow_the_route():
    small, large = S.poisson2d_p1(4), S.poisson2d_p1(7)
    srv = serving.SolveServer(route_rows=10000)
    assert srv.bytes_of(_req(small)) == serving.batch_bytes(small.n, small.nnz)
    assert srv.bytes_of(_req(large)) == serving.dedicated_bytes(large.n, large.nnz)
    # a dedicated System holds more than a batch slot of the same matrix
    assert serving.dedicated_bytes(large.n, large.nnz) > serving.batch_bytes(large.n, large.nnz)
    assert serving.batch_bytes(2 * small.n, 2 * small.nnz) == 2 * serving.batch_bytes(small.n, small.nnz)


def test_target_and_status_rule():
    b = np.full(100, 0.5)
    assert serving._target(b, 1e-10) == serving.FLOOR  # 1e-10 * 5 is below 2^-26
    big = np.full(100, 1e3)
    assert serving._target(big, 1e-10) == pytest.approx(1e-10 * 1e4)
    assert serving._status(1, 1e-9, 2e-9) == "ok"
    assert serving._status(1, 3e-9, 2e-9) == "failed"
    assert serving._status(0, 1e-12, 2e-9) == "failed"
    assert serving._status(1, float("nan"), 2e-9) == "failed"


def test_summary_counts_and_percentiles():
    res = [serving.Result(k, "batch", "ok", arrival=0.1 * k, finished=0.1 * k + 0.5 + 0.01 * k)
           for k in range(10)]
    res.append(serving.Result(10, "none", "rejected", arrival=1.0, finished=1.0))
    res.append(serving.Result(11, "dedicated", "failed", arrival=1.1, finished=2.6))
    summ = serving.summary(res, 3.0)
    assert (summ["ok"], summ["failed"], summ["rejected"]) == (10, 1, 1)
    assert summ["solves_per_s"] == pytest.approx(11 / 3.0)  # rejected requests excluded
    lat = np.array([0.5 + 0.01 * k for k in range(10)] + [1.5])
    assert summ["p50_latency_ms"] == pytest.approx(np.percentile(lat, 50) * 1e3)
    assert summ["p99_latency_ms"] == pytest.approx(np.percentile(lat, 99) * 1e3)


def test_server_rejects_a_nonpositive_budget():
    with pytest.raises(ValueError):
        serving.SolveServer(budget_bytes=0)
