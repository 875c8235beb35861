"""Continuous-batching solve server on one GPU.

A request is a square CSR system with one right-hand side, solved by Rainbow
CG-SSOR (the iteration of ConjugateGradientSolverCSR_Rainbow,
solver/sparsesolver.cpp:1002-1335, omega 1, the reference's greedy colouring)
in FP32 inner solves with FP64 iterative refinement to
||b - A x||_2 <= max(precision ||b||_2, 2^-26).  The reference solves one
system per call and has no server; this is the request loop around that call.

* A request is seen only once it has arrived.
* Routes: requests with fewer than ``route_rows`` rows join the shared batch
  engine (``capi.Batch``: every running solve advances in the same launches,
  and new solves join between engine runs -- continuous batching); larger
  ones get a device-resident ``capi.System`` of their own on one of
  ``dedicated_workers`` host threads (its own stream, overlapping the batch
  on the GPU).  The colouring and device-order plan of a batched request are
  built on ``prep_workers`` host threads right after it arrives.
* Admission: requests reach the device under ONE budget of device bytes
  covering both routes -- the batch engine's resident bytes
  (``sfa_batch_memory``) plus an estimate for every dedicated System in
  flight (``dedicated_bytes``) -- in arrival order among the requests that
  are ready (a batched request whose host plan is still being built lets
  later ones pass).  A request that would exceed the budget waits, and so do
  all after it, until earlier ones finish; one larger than the whole budget
  is rejected.
* Matrices, colourings and Krylov state stay on the device while a request
  runs and are freed when its result is collected.
* Every result is checked: status "ok" only if the solve reports success and
  its FP64 residual norm (formed on the device in FP64) is within the target;
  otherwise "failed".  A request whose solve breaks down or meets a NaN ends
  "failed" on its own: each solve's reductions involve only its own rows, so
  the other requests' results are the same, bit for bit, as without it.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import capi

FLOOR = 2.0 ** -26  # desired_precision, base/base.hpp:121-128


def batch_bytes(n: int, nnz: int) -> float:
    """Estimated device bytes of a batched request: the FP64 SpMV part of the
    refinement residual (12 B per entry), the FP32 SpMV and triangular parts
    (16 B), SELL padding; seven FP32 and three FP64 slot vectors, diagonal and
    permutation."""
    return 48.0 * nnz + 96.0 * n


def dedicated_bytes(n: int, nnz: int) -> float:
    """Estimated device bytes of a dedicated System with its FP32 inner system:
    FP64 SpMV and triangular parts (24 B per entry), their values in FP32 (8 B;
    the FP32 system shares their columns), SELL padding, ~15 FP64 and ~10 FP32
    vectors."""
    return 40.0 * nnz + 192.0 * n


@dataclass
class Request:
    """One solve: a square CSR system (``system.n``, ``.A``, ``.C``, ``.V``)
    and a right-hand side; ``arrival`` in seconds after the stream starts.
    The SSOR diagonal is the matrix's own (MatrixCSR::getDiagonal, formed by
    the library on the serving side)."""
    system: object
    rhs: np.ndarray
    arrival: float = 0.0
    id: int = -1


@dataclass
class Result:
    id: int
    route: str                    # "batch", "dedicated" or "none" (rejected)
    status: str                   # "ok", "failed" or "rejected"
    x: Optional[np.ndarray] = None
    iterations: int = 0           # FP32 inner iterations, summed
    outer: int = 0                # FP64 corrections
    residual_norm: float = float("nan")  # ||b - A x||_2, FP64, on the device
    target: float = float("nan")  # max(precision ||b||_2, 2^-26)
    arrival: float = 0.0
    admitted: float = float("nan")
    finished: float = float("nan")
    bytes: float = 0.0            # device bytes charged to the budget

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _target(b, precision):
    nb = float(np.sqrt(np.dot(b, b)))
    t = precision * nb
    return t if t >= FLOOR else FLOOR


def _status(success, residual_norm, target):
    ok = bool(success) and np.isfinite(residual_norm) and residual_norm <= target
    return "ok" if ok else "failed"


def _diagonal(s):
    """MatrixCSR::getDiagonal of the request's matrix (sfa_MatrixCSR_getDiagonal:
    C++, without the GIL, so worker threads form theirs side by side)."""
    return capi.get_diagonal(s.n, s.A, s.C, s.V)


def _prepare(req: Request):
    """Colouring and device-order plan of a batched request (host only)."""
    s = req.system
    rb = capi.Rainbow(s.n, s.A, s.C)
    return capi.PreparedMatrix(s.n, s.A, s.C, s.V, _diagonal(s), 1.0, rb.num_colors, rb.F, rb.B,
                               rb.R)


@dataclass
class ServerStats:
    peak_bytes: float = 0.0       # most device bytes held at once: the engine's
                                  # actual bytes plus the dedicated estimates
    batched: int = 0
    dedicated: int = 0
    rejected: int = 0
    makespan: float = 0.0
    # per dedicated request, seconds: (wait for a worker, colouring + System
    # creation, solve + download, freeing the System and its buffers)
    dedicated_phases: list = field(default_factory=list)


class SolveServer:
    # route_rows: on the 48-request stream (1e3..2e6 unknowns, 20/s, MI355X),
    # 30,000 served ~6% more solves/s than 10,000 (three runs each: 1.92-2.05x
    # against 1.85-1.89x the plain FP64 path's rate), at a p50 of 1.13-1.38 s
    # instead of 0.92-1.07 s; 100,000 served less.
    def __init__(self, budget_bytes=200e9, max_slots=256, route_rows=30000, dedicated_workers=12,
                 prep_workers=8, precision=1e-10, inner_rtol=1e-1, max_outer=50,
                 inner_restarts=False, engine_steps=128):
        if budget_bytes <= 0:
            raise ValueError("budget_bytes must be positive")
        self.budget = float(budget_bytes)
        self.max_slots = int(max_slots)
        self.route_rows = int(route_rows)
        self.dedicated_workers = int(dedicated_workers)
        self.prep_workers = int(prep_workers)
        self.precision = float(precision)
        self.inner_rtol = float(inner_rtol)
        self.max_outer = int(max_outer)
        self.inner_restarts = bool(inner_restarts)
        self.engine_steps = int(engine_steps)
        self.stats = ServerStats()

    def route_of(self, req: Request) -> str:
        if self.dedicated_workers > 0 and req.system.n >= self.route_rows:
            return "dedicated"
        return "batch"

    def bytes_of(self, req: Request) -> float:
        s = req.system
        if self.route_of(req) == "dedicated":
            return dedicated_bytes(s.n, s.nnz)
        return batch_bytes(s.n, s.nnz)

    # ------------------------------------------------------------------ routes

    def _solve_dedicated(self, req: Request, start: float):
        """Colouring, plan, upload and mixed-precision solve of one request on
        its own System (runs on a worker thread).  Returns (x, refine info,
        time taken, time ready, time solved, time done) in seconds after
        start."""
        t_take = time.perf_counter() - start
        s = req.system
        rb = capi.Rainbow(s.n, s.A, s.C)
        with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, _diagonal(s), 1.0,
                         rb.num_colors, rb.F, rb.B, rb.R, fp32_inner=True,
                         inner_restarts=self.inner_restarts, borrow_csr=True) as sysd:
            t_ready = time.perf_counter() - start
            # host arrays: the System copies them through its own buffers,
            # where a hipMalloc'd vector would end in a hipFree that waits
            # for every other stream's work
            xh = np.zeros(s.n)
            info = sysd.solve_mixed(xh, np.ascontiguousarray(req.rhs, dtype=np.float64),
                                    self.precision, inner_rtol=self.inner_rtol,
                                    max_outer=self.max_outer)
            t_solved = time.perf_counter() - start
        return xh, info, t_take, t_ready, t_solved, time.perf_counter() - start

    def solve_alone(self, req: Request) -> Result:
        """The request on its own, by the route the server gives it: a batch
        engine holding only this solve, or its own System."""
        b = np.ascontiguousarray(req.rhs, dtype=np.float64)
        target = _target(b, self.precision)
        s = req.system
        if self.route_of(req) == "dedicated":
            x, info, *_ = self._solve_dedicated(req, time.perf_counter())
            return Result(req.id, "dedicated", _status(info["success"], info["residual_norm"], target),
                          x, info["inner_iterations"], info["outer"], info["residual_norm"], target,
                          req.arrival)
        with capi.Batch(mixed=True, max_slots=1, inner_restarts=self.inner_restarts) as eng:
            prep = _prepare(req)
            mid = eng.add_prepared(prep)
            prep.close()
            slot = eng.submit(mid, b, precision=self.precision, inner_rtol=self.inner_rtol,
                              max_outer=self.max_outer)
            while eng.run(0) > 0:
                pass
            x, info = eng.collect(slot, s.n)
        return Result(req.id, "batch", _status(info["success"], info["residual_norm"], target), x,
                      info["iterations"], info["outer"], info["residual_norm"], target, req.arrival)

    # ------------------------------------------------------------------ loop

    def serve(self, requests, log=None):
        """Serves a stream of requests, replaying their arrival times in real
        time from now.  Returns one Result per request, in request order."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for k, r in enumerate(reqs):
            if r.id < 0:
                r.id = k
        results = {}
        self.stats = st = ServerStats()
        with capi.Batch(mixed=True, max_slots=self.max_slots,
                        inner_restarts=self.inner_restarts) as eng, \
                ThreadPoolExecutor(max_workers=max(1, self.prep_workers)) as prep_pool, \
                ThreadPoolExecutor(max_workers=max(1, self.dedicated_workers)) as big_pool:
            arrivals = list(reqs)
            pending = []       # arrived, not yet on the device, arrival order: (req, prep future)
            in_batch = {}      # slot -> (req, matrix id, bytes, target, admitted)
            in_dedicated = []  # (req, future, bytes, target, admitted)
            ded_bytes = 0.0
            start = time.perf_counter()
            last_log = 0.0
            while arrivals or pending or in_batch or in_dedicated:
                now = time.perf_counter() - start
                while arrivals and arrivals[0].arrival <= now:
                    r = arrivals.pop(0)
                    fut = prep_pool.submit(_prepare, r) if self.route_of(r) == "batch" else None
                    pending.append((r, fut))
                # admission under one budget for both routes, in arrival order
                # among the requests that are ready (a batched request is
                # ready once its host plan is built; one still being planned
                # lets later ones pass rather than hold them back)
                i = 0
                while i < len(pending):
                    r, fut = pending[i]
                    need = self.bytes_of(r)
                    b = np.ascontiguousarray(r.rhs, dtype=np.float64)
                    target = _target(b, self.precision)
                    if need > self.budget:
                        if fut is not None:
                            fut.result().close()
                        results[r.id] = Result(r.id, "none", "rejected", target=target,
                                               arrival=r.arrival, finished=now)
                        st.rejected += 1
                        pending.pop(i)
                        continue
                    if fut is not None and not fut.done():
                        i += 1
                        continue
                    used = (eng.memory() if in_batch else 0.0) + ded_bytes
                    if used + need > self.budget and (in_batch or in_dedicated):
                        break
                    t_adm = time.perf_counter() - start
                    if fut is None:
                        in_dedicated.append((r, big_pool.submit(self._solve_dedicated, r, start),
                                             need, target, t_adm))
                        ded_bytes += need
                        st.dedicated += 1
                    else:
                        prep = fut.result()
                        mid = eng.add_prepared(prep)
                        prep.close()
                        slot = eng.submit(mid, b, precision=self.precision,
                                          inner_rtol=self.inner_rtol, max_outer=self.max_outer)
                        in_batch[slot] = (r, mid, need, target, t_adm)
                        st.batched += 1
                    # what is now held: the engine's actual bytes, the dedicated estimates
                    st.peak_bytes = max(st.peak_bytes,
                                        (eng.memory() if in_batch else 0.0) + ded_bytes)
                    pending.pop(i)
                if in_batch:
                    eng.run(self.engine_steps)
                for slot in eng.finished():
                    r, mid, need, target, t_adm = in_batch.pop(slot)
                    x, info = eng.collect(slot, r.system.n)
                    eng.remove_matrix(mid)
                    results[r.id] = Result(r.id, "batch",
                                           _status(info["success"], info["residual_norm"], target),
                                           x, info["iterations"], info["outer"],
                                           info["residual_norm"], target, r.arrival, t_adm,
                                           time.perf_counter() - start, need)
                still = []
                for item in in_dedicated:
                    r, fut, need, target, t_adm = item
                    if not fut.done():
                        still.append(item)
                        continue
                    x, info, t_take, t_ready, t_solved, t_done = fut.result()
                    ded_bytes -= need
                    st.dedicated_phases.append((t_take - r.arrival, t_ready - t_take,
                                                t_solved - t_ready, t_done - t_solved))
                    results[r.id] = Result(r.id, "dedicated",
                                           _status(info["success"], info["residual_norm"], target),
                                           x, info["inner_iterations"], info["outer"],
                                           info["residual_norm"], target, r.arrival, t_adm, t_done,
                                           need)
                in_dedicated = still
                if not in_batch:
                    if pending or in_dedicated:
                        time.sleep(0.0005)
                    elif arrivals:
                        time.sleep(max(0.0, arrivals[0].arrival - (time.perf_counter() - start)))
                if log is not None and time.perf_counter() - start - last_log > 5.0:
                    last_log = time.perf_counter() - start
                    log(f"served {len(results)} of {len(reqs)} after {last_log:.1f}s, "
                        f"{len(in_batch)} batched and {len(in_dedicated)} dedicated in flight, "
                        f"{len(pending)} waiting")
            st.makespan = time.perf_counter() - start
        return [results[r.id] for r in sorted(reqs, key=lambda r: r.id)]


def summary(results, makespan):
    """p50/p99 latency (ms) and solves/s over the results that were served."""
    served = [r for r in results if r.status != "rejected"]
    lat = np.array([r.latency for r in served]) if served else np.array([np.nan])
    return {"requests": len(results), "ok": sum(r.status == "ok" for r in results),
            "failed": sum(r.status == "failed" for r in results),
            "rejected": sum(r.status == "rejected" for r in results),
            "solves_per_s": len(served) / makespan if makespan > 0 else float("nan"),
            "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_latency_ms": float(np.percentile(lat, 99) * 1e3), "makespan_s": makespan}
