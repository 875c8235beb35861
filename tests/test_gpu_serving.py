"""The served form (simplefem_amd.serving.SolveServer) on the GPU.

Bars: every served result is checked by the server itself (success and an
FP64 residual within max(1e-10 ||b||, 2^-26)) and reported "ok"; each served
x equals, bit for bit, the same request solved alone by the same route (a
batch engine holding only it, or its own System) -- every solve's reductions
involve only its own rows, so the mix cannot change it; a request with a NaN
in its right-hand side ends "failed" and leaves every other result
bit-identical to a run of the stream without it; admission never holds more
device bytes than the budget.
"""
import numpy as np
import pytest

from simplefem_amd import serving
from simplefem_amd import workloads as W

pytestmark = pytest.mark.gpu

ROUTE_ROWS = 15000


def _stream(seed=11):
    jobs = W.served_requests(16, rate=400.0, seed=seed, dofs=(1e3, 3e4))
    return [serving.Request(j.system, j.rhs[0], j.arrival, k) for k, j in enumerate(jobs)]


def _server(**kw):
    args = dict(route_rows=ROUTE_ROWS, dedicated_workers=4, prep_workers=4, max_slots=64)
    args.update(kw)
    return serving.SolveServer(**args)


def test_served_stream_matches_alone_and_isolates_a_nan_request():
    reqs = _stream()
    srv = _server()
    routes = [srv.route_of(r) for r in reqs]
    assert "batch" in routes and "dedicated" in routes
    print(f"case: served stream of {len(reqs)} ({routes.count('dedicated')} dedicated)", flush=True)
    clean = srv.serve(reqs)
    for res, req in zip(clean, reqs):
        assert res.status == "ok", (req.system.name, res.route, res.residual_norm, res.target)
        assert res.route == srv.route_of(req)
    print(f"served: {serving.summary(clean, srv.stats.makespan)}", flush=True)

    # each result equals the same request alone, by the same route
    for res, req in zip(clean, reqs):
        alone = srv.solve_alone(req)
        assert alone.status == "ok"
        assert np.array_equal(res.x, alone.x), (req.system.name, res.route)
        assert res.iterations == alone.iterations and res.outer == alone.outer

    # the same stream with one request's right-hand side poisoned by a NaN:
    # one batched and one dedicated request in turn
    for route in ("batch", "dedicated"):
        k = next(i for i, r in enumerate(reqs) if srv.route_of(r) == route)
        bad = [serving.Request(r.system, r.rhs.copy(), r.arrival, r.id) for r in reqs]
        act = np.flatnonzero(bad[k].system.diagonal() != 0)
        bad[k].rhs[act[len(act) // 2]] = np.nan
        print(f"case: NaN in request {k} ({route})", flush=True)
        mixed = _server().serve(bad)
        assert mixed[k].status == "failed", mixed[k]
        for i, (a, b) in enumerate(zip(mixed, clean)):
            if i == k:
                continue
            assert a.status == "ok"
            assert np.array_equal(a.x, b.x), (i, reqs[i].system.name)


def test_admission_stays_within_the_budget():
    reqs = _stream(seed=5)
    srv = _server()
    needs = [srv.bytes_of(r) for r in reqs]
    budget = 2.5 * max(needs)  # a few requests at a time
    srv = _server(budget_bytes=budget)
    print(f"case: budget {budget / 1e6:.1f} MB for {len(reqs)} requests", flush=True)
    res = srv.serve(reqs)
    assert all(r.status == "ok" for r in res)
    assert srv.stats.peak_bytes <= budget
    assert srv.stats.peak_bytes > 0

    # a request larger than the whole budget is rejected, the rest are served
    small = _server(budget_bytes=sorted(needs)[-2])
    res = small.serve(reqs)
    big = int(np.argmax(needs))
    assert res[big].status == "rejected"
    assert all(r.status == "ok" for i, r in enumerate(res) if i != big)
    assert small.stats.peak_bytes <= sorted(needs)[-2]
