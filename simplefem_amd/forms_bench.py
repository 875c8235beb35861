"""Measurements of the mixed-precision, batched and served forms (bench.py --config ...).

hcurl-mixed  curl-curl + mass of Whitney 1-forms on 48^3 Kuhn cubes (795,024
             edges), Rainbow CG-SSOR: FP64 solve vs FP32-inner/FP64-refined
             solve to ||b - A x|| <= 1e-10 ||b||; error of the mixed solution
             against the FP64 one.
batched      256 2D Poisson systems (L in 4..8, degree 1..3, 4 right-hand sides
             each) in the persistent batch engine, mixed precision, against the
             FP64 batch engine and against one-at-a-time FP64 solves.
served       a Poisson stream of requests (2D/3D, log-uniform size, degree 1..3)
             admitted under an HBM budget into the running batch engine;
             p50/p99 latency from arrival to result, solves/s.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np
import scipy.sparse as sp

from . import capi
from . import systems as S
from . import workloads as W
from .hipmem import DeviceArray, synchronize

METRIC = "solves/s + HBM GB/s (3D P1 Poisson, rtol 1e-10); p50/p99 served latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _base(args, value, unit, ms, config, dtype):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": config}


def _true_relres(s, x, b):
    r = b - capi.csr_apply(s.n, s.n, s.A, s.C, s.V, x)
    r[s.diagonal() == 0] = 0.0
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def hcurl_mixed(args):
    t0 = time.time()
    s = S.hcurl3d_whitney_kuhn(48)
    b = S.rhs_uniform(s, 2026)
    d = capi.get_diagonal(s.n, s.A, s.C, s.V)
    rb = capi.Rainbow(s.n, s.A, s.C)
    col = (rb.num_colors, rb.F, rb.B, rb.R)
    log(f"hcurl system n={s.n} nnz={s.nnz} colours={rb.num_colors} ({time.time() - t0:.1f}s)")
    sysd = capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0, *col, fp32_inner=True,
                       paired_sweeps=args.paired_sweeps, inner_restarts=args.inner_restarts)
    if args.spmv_lanes:
        sysd.set_spmv(args.spmv_lanes)
        tuning = {"lanes_per_row": args.spmv_lanes, "tuned": False}
    else:
        tuning = {"fp64": sysd.tune_spmv(), "fp32_inner": sysd.tune_spmv(fp32_inner=True)}
        log(f"spmv tuning: {tuning}")
    bt = DeviceArray.from_host(b)
    x = DeviceArray(s.n)
    r = DeviceArray(s.n)
    info = {}
    steps = max(1, args.steps // 4)
    if args.engine_profile:
        # one FP64 solve with every kernel launched on its own between two
        # events (no graphs): per-class launches, event time and bytes
        x.zero()
        r.zero()
        prof, pinfo = sysd.profile(x, bt, r, 1e-10)
        sysd.close()
        return {"hcurl_kernel_profile_fp64": {k: v for k, v in prof.items() if v["launches"]},
                "iterations": pinfo["iterations"], "rows": s.n, "nnz": s.nnz,
                "colours": rb.num_colors}

    def fp64(precision=1e-10):
        x.zero()
        r.zero()
        sysd.solve(x, bt, r, precision, info=info)

    def mixed(rtol=1e-10):
        x.zero()
        return sysd.solve_mixed(x, bt, rtol, inner_rtol=args.inner_rtol, max_outer=50)

    fp64()
    synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        fp64()
    synchronize()
    t64 = (time.perf_counter() - t1) / steps
    x64 = x.download()
    res64 = _true_relres(s, x64, b)
    K64 = info["iterations"]
    log(f"fp64: {1.0 / t64:.2f} solves/s, K = {K64}")
    # The FP64 solve stops on the reference's rule, sqrt(z.r) below the
    # tolerance, where its true residual ||b - A x|| can stay above the
    # mixed solve's target (max(1e-10 ||b||, 2^-26) on the true residual).
    # The FP64 solve asked for more: precision halved until its true residual
    # meets that target (or 12 times), then timed.
    target_rel = max(1e-10 * float(np.linalg.norm(b)), 2.0 ** -26) / float(np.linalg.norm(b))
    p_eq, res_eq, K_eq, exit_eq = 1e-10, res64, K64, info.get("exit_reason")
    for _ in range(12):
        if res_eq <= target_rel:
            break
        p_eq *= 0.5
        fp64(p_eq)
        synchronize()
        res_eq, K_eq = _true_relres(s, x.download(), b), info["iterations"]
        exit_eq = info.get("exit_reason")
    t1 = time.perf_counter()
    for _ in range(steps):
        fp64(p_eq)
    synchronize()
    t_eq = (time.perf_counter() - t1) / steps
    log(f"fp64 tightened: precision {p_eq:.3g}, K = {K_eq}, exit {exit_eq}, "
        f"relres {res_eq:.3g} (target {target_rel:.3g}), {1.0 / t_eq:.2f} solves/s")
    mi = mixed()
    synchronize()
    log(f"mixed: {mi}")
    t1 = time.perf_counter()
    for _ in range(steps):
        mi = mixed()
    synchronize()
    tmx = (time.perf_counter() - t1) / steps
    xm = x.download()
    # the mixed solve to the accuracy the FP64 solve reaches with the
    # reference's rule (its true relative residual), for a like-for-like time
    mi_eq = mixed(res64)
    synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        mi_eq = mixed(res64)
    synchronize()
    tmx_eq = (time.perf_counter() - t1) / steps
    log(f"mixed to the FP64 solve's accuracy {res64:.3g}: {mi_eq}, {1.0 / tmx_eq:.2f} solves/s")
    resm = _true_relres(s, xm, b)
    err = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
    sysd.close()
    cfg = {"workload": "H(curl) curl-curl + mass, Whitney 1-forms, 48^3 Kuhn cubes, Rainbow "
                       "CG-SSOR omega=1, FP32 inner / FP64 refinement to rtol 1e-10",
           "rows": s.n, "nnz": s.nnz, "colours": rb.num_colors,
           "fp64": {"solves_per_s": 1.0 / t64, "iterations": K64, "true_relative_residual": res64},
           "fp64_tightened": {"what": "the FP64 solve with its precision halved until its "
                                      "true residual meets the mixed solve's target, at most "
                                      "12 times",
                              "precision": p_eq, "solves_per_s": 1.0 / t_eq,
                              "iterations": K_eq,
                              "exit_reason": {0: "converged", 1: "cap", 2: "breakdown",
                                              3: "small d.Ad"}.get(exit_eq, exit_eq),
                              "true_relative_residual": res_eq,
                              "meets_true_residual_target": bool(res_eq <= target_rel)},
           "mixed": {"solves_per_s": 1.0 / tmx, "outer": mi["outer"],
                     "inner_iterations": mi["inner_iterations"],
                     "continued_inner_solves": mi["continued"],
                     "inner_rtol": args.inner_rtol,
                     "inner_restarts": bool(args.inner_restarts),
                     "true_relative_residual": resm},
           "mixed_to_fp64_accuracy": {"rtol": res64, "solves_per_s": 1.0 / tmx_eq,
                                      "outer": mi_eq["outer"],
                                      "inner_iterations": mi_eq["inner_iterations"],
                                      "speedup_over_fp64": t64 / tmx_eq},
           "mixed_vs_fp64_solution_rel_error": err,
           "spmv_mapping": tuning,
           "speedup_mixed_over_fp64": t64 / tmx}
    return _base(args, round(1.0 / tmx, 3), "solves/s", round(tmx * 1e3, 3), cfg, "f32+f64")


def _submit_all(eng, jobs, mids, grouped, **kw):
    """Every right-hand side of every job: one interleaved group per system
    (sfa_batch_submit_group) or one slot at a time; [(slot, job, b)]."""
    out = []
    for j, mid in zip(jobs, mids):
        if grouped:
            for k in range(0, len(j.rhs), 4):
                bs = j.rhs[k:k + 4]
                out += [(sl, j, b) for sl, b in zip(eng.submit_group(mid, bs, **kw), bs)]
        else:
            out += [(eng.submit(mid, b, **kw), j, b) for b in j.rhs]
    return out


def engine_profile(jobs, cols, mixed, chunk=2000, grouped=True, max_steps=400000):
    """The batched workload on one engine with every kernel of every step
    launched between HIP events (sfa_batch_profile), in chunks of `chunk`
    steps (up to max_steps): per chunk the solves still running and the mean
    step, fold and schedule kernel times."""
    rows = []
    with capi.Batch(mixed=mixed, max_slots=min(1024, sum(len(j.rhs) for j in jobs))) as eng:
        mids = [eng.add_matrix(j.system.n, j.system.A, j.system.C, j.system.V, j.system.diagonal(),
                               1.0, rb.num_colors, rb.F, rb.B, rb.R) for j, rb in zip(jobs, cols)]
        nslots = len(_submit_all(eng, jobs, mids, grouped, precision=1e-10))
        total = 0
        while True:
            chunk = min(chunk, max_steps - total)
            p = eng.profile(chunk)
            total += chunk
            running = nslots - len(eng.finished(2048))
            rows.append({"steps": total, "running": running,
                         "step_us": round(p["step_ms"] / chunk * 1e3, 1),
                         "fold_us": round(p["fold_ms"] / chunk * 1e3, 1),
                         "schedule_us": round(p["prefix_ms"] / chunk * 1e3, 1)})
            log(f"engine profile: {rows[-1]}")
            if running == 0 or total >= max_steps:
                break
    return rows


def batched(args):
    t0 = time.time()
    jobs = W.batched_systems(args.systems, 4)
    if args.only_system:  # the workload's systems of that name only (profiling)
        jobs = [j for j in jobs if j.system.name == args.only_system]
        if not jobs:
            raise SystemExit(f"no system named {args.only_system} in the batched workload")
    nrhs = sum(len(j.rhs) for j in jobs)
    log(f"{len(jobs)} systems, {nrhs} right-hand sides, "
        f"{sum(j.system.n for j in jobs)} rows in total ({time.time() - t0:.1f}s)")
    cols = [capi.Rainbow(j.system.n, j.system.A, j.system.C) for j in jobs]

    def run_engine(mixed):
        with capi.Batch(mixed=mixed, max_slots=min(1024, nrhs),
                        inner_restarts=args.inner_restarts) as eng:
            mids = []
            for j, rb in zip(jobs, cols):
                s = j.system
                mids.append(eng.add_matrix(s.n, s.A, s.C, s.V, s.diagonal(), 1.0, rb.num_colors,
                                           rb.F, rb.B, rb.R))
            if args.spmv_lanes:
                eng.set_spmv([args.spmv_lanes] * capi.SPMV_BUCKETS)
                tuning = {"lanes_per_row": args.spmv_lanes, "tuned": False}
            else:
                tuning = eng.tune_spmv()
                log(f"{'mixed' if mixed else 'fp64'} engine spmv tuning: {tuning}")
            synchronize()
            steps0 = eng.steps()
            t1 = time.perf_counter()
            slots = _submit_all(eng, jobs, mids, args.rhs_layout == "grouped", precision=1e-10,
                                inner_rtol=args.inner_rtol)
            t_sub = time.perf_counter() - t1
            last = t1
            while True:
                left = eng.run(args.chunk_steps)
                if left == 0:
                    break
                now = time.perf_counter()
                if now - last > 5.0:
                    log(f"{'mixed' if mixed else 'fp64'} engine: {left} of {len(slots)} solves "
                        f"running after {now - t1:.1f}s")
                    last = now
            synchronize()
            elapsed = time.perf_counter() - t1
            steps = eng.steps() - steps0
            worst = 0.0  # largest true relative residual over a sample of the solves
            above = sampled = 0  # sampled solves whose true residual misses max(1e-10 ||b||, 2^-26)
            iters = 0
            ok = 0
            check_every = max(1, len(slots) // 16)
            for k, (slot, j, b) in enumerate(slots):
                x, info = eng.collect(slot, j.system.n)
                iters += info["iterations"]
                ok += info["success"]
                if k % check_every == 0:  # on the host, in FP64 (scipy)
                    s = j.system
                    res = b - sp.csr_matrix((s.V, s.C, s.A), shape=(s.n, s.n)) @ x
                    res[s.diagonal() == 0] = 0.0  # rows the solver skips
                    bn = float(np.linalg.norm(b))
                    rel = float(np.linalg.norm(res)) / bn
                    worst = max(worst, rel if np.isfinite(rel) else float("inf"))
                    sampled += 1
                    above += not (rel * bn <= max(1e-10 * bn, 2.0 ** -26))
            return (elapsed, t_sub, iters, ok, eng.memory(), tuning, worst, steps,
                    f"{above} of {sampled}")

    if args.engine_profile:
        return {"engine_profile_fp64": engine_profile(jobs, cols, False,
                                                      grouped=args.rhs_layout == "grouped",
                                                      max_steps=args.profile_steps)}
    e_mx, sub_mx, it_mx, ok_mx, mem, tune_mx, res_mx, st_mx, ab_mx = run_engine(True)
    log(f"batched mixed: {nrhs} solves in {e_mx:.3f}s ({ok_mx} converged, {st_mx} steps)")
    e_64, sub_64, it_64, ok_64, _, tune_64, res_64, st_64, ab_64 = run_engine(False)
    log(f"batched fp64: {nrhs} solves in {e_64:.3f}s ({ok_64} converged)")
    # one-at-a-time FP64 path (device-resident System per matrix), on a sample
    sample = jobs[:: max(1, len(jobs) // 32)]
    t_seq = 0.0
    n_seq = 0
    for j in sample:
        s = j.system
        rb = capi.Rainbow(s.n, s.A, s.C)
        with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, s.diagonal(), 1.0,
                         rb.num_colors, rb.F, rb.B, rb.R) as sysd:
            if args.spmv_lanes:
                sysd.set_spmv(args.spmv_lanes)
            else:
                sysd.tune_spmv()
            bt = DeviceArray.from_host(j.rhs[0])
            x = DeviceArray(s.n)
            r = DeviceArray(s.n)
            sysd.solve(x, bt, r, 1e-10)
            synchronize()
            t1 = time.perf_counter()
            for b in j.rhs:
                bt.upload(b)
                x.zero()
                r.zero()
                sysd.solve(x, bt, r, 1e-10)
            synchronize()
            t_seq += time.perf_counter() - t1
            n_seq += len(j.rhs)
    seq_rate = n_seq / t_seq
    log(f"one-at-a-time fp64: {n_seq} solves in {t_seq:.3f}s")
    cfg = {"workload": f"{len(jobs)} 2D Poisson systems (L 4..8, Lagrange degree 1..3), 4 RHS "
                       "each, Rainbow CG-SSOR, persistent batch engine, FP32 inner / FP64 "
                       "refinement to rtol 1e-10",
           "rhs_layout": args.rhs_layout,
           "systems": len(jobs), "solves": nrhs, "rows_total": int(sum(j.system.n for j in jobs)),
           "mixed": {"solves_per_s": nrhs / e_mx, "converged": ok_mx, "inner_iterations": it_mx,
                     "submit_s": sub_mx, "max_sampled_true_relative_residual": res_mx,
                     "sampled_above_true_residual_target": ab_mx,
                     "engine_steps": st_mx, "us_per_step": e_mx / max(st_mx, 1) * 1e6,
                     "spmv_mapping": tune_mx},
           "fp64_batched": {"solves_per_s": nrhs / e_64, "converged": ok_64, "iterations": it_64,
                            "max_sampled_true_relative_residual": res_64,
                            "sampled_above_true_residual_target": ab_64,
                            "engine_steps": st_64, "us_per_step": e_64 / max(st_64, 1) * 1e6,
                            "spmv_mapping": tune_64},
           "fp64_one_at_a_time": {"solves_per_s": seq_rate, "sample_solves": n_seq},
           "speedup_mixed_batched_over_fp64_one_at_a_time": (nrhs / e_mx) / seq_rate,
           "device_bytes": mem}
    return _base(args, round(nrhs / e_mx, 3), "solves/s", round(e_mx / nrhs * 1e3, 4), cfg,
                 "f32+f64")


def _served_target(b):
    """The served solves' target on the true residual ||b - A x||_2."""
    return max(1e-10 * float(np.linalg.norm(b)), 2.0 ** -26)


def _fp64_same_residual_precisions(jobs, tries=12):
    """Untimed: per request, the loosest precision 1e-10 * 2^-k (k <= tries)
    at which the plain FP64 solve from x0 = 0 leaves a true residual
    ||b - A x||_2 (rows with diagonal 0 left out) within the served target.
    The reference's rule stops on sqrt(z.r), which can leave the true
    residual above it.  Returns the precisions and how many requests needed
    one below 1e-10."""
    precs, tightened = [], 0
    for j in jobs:
        s, b = j.system, j.rhs[0]
        d = capi.get_diagonal(s.n, s.A, s.C, s.V)
        rb = capi.Rainbow(s.n, s.A, s.C)
        mat = sp.csr_matrix((s.V, s.C, s.A), shape=(s.n, s.n))
        target = _served_target(b)
        p = 1e-10
        with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0,
                         rb.num_colors, rb.F, rb.B, rb.R) as sysd:
            bt, x, r = DeviceArray.from_host(b), DeviceArray(s.n), DeviceArray(s.n)
            for k in range(tries + 1):
                x.zero()
                r.zero()
                sysd.solve(x, bt, r, p)
                res = b - mat @ x.download()
                res[d == 0] = 0.0
                if float(np.linalg.norm(res)) <= target or k == tries:
                    break
                p *= 0.5
        precs.append(p)
        tightened += p < 1e-10
    return precs, tightened


def _served_fp64_one_at_a_time(jobs, precisions=None):
    """The plain FP64 path on the same stream: requests in arrival order, one
    device-resident System each (colouring, plan, upload, solve from x0 = 0),
    each at precision 1e-10 or at precisions[k]."""
    lat = []
    split = np.zeros(3)  # seconds: colouring + diagonal, System creation, upload + solve
    start = time.perf_counter()
    last = start
    for k, j in enumerate(jobs):
        now = time.perf_counter() - start
        if j.arrival > now:
            time.sleep(j.arrival - now)
        s = j.system
        t0 = time.perf_counter()
        rb = capi.Rainbow(s.n, s.A, s.C)
        # the diagonal formed as the server forms it (MatrixCSR::getDiagonal)
        d = capi.get_diagonal(s.n, s.A, s.C, s.V)
        t1 = time.perf_counter()
        # the request's arrays outlive the System, as on the served route
        with capi.System(capi.METHOD_RAINBOW, s.n, s.A, s.C, s.V, d, 1.0,
                         rb.num_colors, rb.F, rb.B, rb.R, borrow_csr=True) as sysd:
            t2 = time.perf_counter()
            bt = DeviceArray.from_host(j.rhs[0])
            x = DeviceArray(s.n)
            r = DeviceArray(s.n)
            sysd.solve(x, bt, r, 1e-10 if precisions is None else precisions[k])
            synchronize()
            split += (t1 - t0, t2 - t1, time.perf_counter() - t2)
        lat.append(time.perf_counter() - start - j.arrival)
        if time.perf_counter() - last > 5.0:
            last = time.perf_counter()
            log(f"fp64 one at a time: {k + 1} of {len(jobs)} after {last - start:.1f}s")
    return time.perf_counter() - start, np.array(lat), split


def served(args, budget=200e9):
    """The served stream through serving.SolveServer (admission under one HBM
    budget for both routes, continuous batching, every result checked),
    against the plain FP64 path on the same stream."""
    from . import serving

    t0 = time.time()

    def progress(k, job):
        if k % 8 == 7:
            log(f"generated {k + 1} requests ({time.time() - t0:.1f}s, last n={job.system.n})")

    jobs = W.served_requests(args.requests, args.rate, dofs=(1e3, 2e6), progress=progress)
    log(f"{len(jobs)} requests generated ({time.time() - t0:.1f}s)")
    reqs = [serving.Request(j.system, j.rhs[0], j.arrival, k) for k, j in enumerate(jobs)]
    srv = serving.SolveServer(budget_bytes=budget, max_slots=256, route_rows=args.route_rows,
                              dedicated_workers=args.dedicated_workers,
                              prep_workers=args.prep_workers, precision=1e-10,
                              inner_rtol=args.inner_rtol, inner_restarts=args.inner_restarts)
    results = srv.serve(reqs, log=log)
    summ = serving.summary(results, srv.stats.makespan)
    log(f"served: {summ}")
    # the same requests all arriving at once: the server's capacity, which the
    # arrival rate above bounds (the stream arrives over its whole span)
    sat = serving.SolveServer(budget_bytes=budget, max_slots=256, route_rows=args.route_rows,
                              dedicated_workers=args.dedicated_workers,
                              prep_workers=args.prep_workers, precision=1e-10,
                              inner_rtol=args.inner_rtol, inner_restarts=args.inner_restarts)
    sat_results = sat.serve([serving.Request(r.system, r.rhs, 0.0, r.id) for r in reqs], log=log)
    summ_sat = serving.summary(sat_results, sat.stats.makespan)
    log(f"saturated: {summ_sat}")
    if args.no_served_baseline:
        total64, lat64, split64 = float("nan"), np.array([np.nan]), np.full(3, np.nan)
        total_eq, lat_eq, split_eq, tightened = total64, lat64, split64, 0
    else:
        log("plain FP64 path on the same stream ...")
        total64, lat64, split64 = _served_fp64_one_at_a_time(jobs)
        precs, tightened = _fp64_same_residual_precisions(jobs)
        log(f"plain FP64 path at the served true-residual target: {tightened} of {len(jobs)} "
            "requests need a precision below 1e-10")
        total_eq, lat_eq, split_eq = _served_fp64_one_at_a_time(jobs, precs)
    phases = np.array(srv.stats.dedicated_phases) if srv.stats.dedicated_phases else None
    worst = max((r.residual_norm / r.target for r in results if r.status == "ok"), default=None)
    cfg = {"workload": f"{len(jobs)} requests, Poisson arrivals at {args.rate}/s, 2D/3D, "
                       "log-uniform 1e3..2e6 unknowns, Lagrange degree 1..3, one RHS each, "
                       "mixed precision to rtol 1e-10, 200 GB HBM admission budget over both "
                       "routes (simplefem_amd.serving.SolveServer); colouring and plan built on "
                       "host worker threads after arrival; requests of at least "
                       f"{args.route_rows} rows on their own System and stream",
           "requests": len(jobs), "ok": summ["ok"], "failed": summ["failed"],
           "rejected": summ["rejected"],
           "max_residual_over_target": worst,
           "p50_latency_ms": summ["p50_latency_ms"], "p99_latency_ms": summ["p99_latency_ms"],
           "makespan_s": summ["makespan_s"], "prep_workers": args.prep_workers,
           "peak_admitted_bytes": srv.stats.peak_bytes,
           "dedicated": {"rows_at_least": args.route_rows, "workers": args.dedicated_workers,
                         "requests": srv.stats.dedicated,
                         "mean_s": dict(zip(("wait", "create", "solve", "release"),
                                            phases.mean(axis=0).round(3).tolist()))
                         if phases is not None else {},
                         "max_s": dict(zip(("wait", "create", "solve", "release"),
                                           phases.max(axis=0).round(3).tolist()))
                         if phases is not None else {}},
           "fp64_one_at_a_time": {"solves_per_s": len(jobs) / total64,
                                  "p50_latency_ms": float(np.percentile(lat64, 50) * 1e3),
                                  "p99_latency_ms": float(np.percentile(lat64, 99) * 1e3),
                                  "makespan_s": total64,
                                  "summed_s": dict(zip(("colouring_and_diagonal", "create",
                                                        "upload_and_solve"),
                                                       split64.round(3).tolist()))},
           "fp64_one_at_a_time_same_residual": {
               "what": "the same path, each request at the loosest precision 1e-10 * 2^-k "
                       "whose FP64 solve meets the served target on the true residual "
                       "(found beforehand, untimed)",
               "requests_below_1e-10": tightened,
               "solves_per_s": len(jobs) / total_eq,
               "p50_latency_ms": float(np.percentile(lat_eq, 50) * 1e3),
               "p99_latency_ms": float(np.percentile(lat_eq, 99) * 1e3),
               "makespan_s": total_eq,
               "summed_s": dict(zip(("colouring_and_diagonal", "create", "upload_and_solve"),
                                    split_eq.round(3).tolist()))},
           "speedup_over_fp64_one_at_a_time": summ["solves_per_s"] / (len(jobs) / total64),
           "speedup_over_fp64_at_same_residual": summ["solves_per_s"] / (len(jobs) / total_eq),
           "saturated": {"what": "the same requests all arriving at t = 0; the FP64 path's "
                                 "rate above is service-bound already (it never waits for "
                                 "an arrival after the first)",
                         "solves_per_s": summ_sat["solves_per_s"], "ok": summ_sat["ok"],
                         "p50_latency_ms": summ_sat["p50_latency_ms"],
                         "p99_latency_ms": summ_sat["p99_latency_ms"],
                         "makespan_s": summ_sat["makespan_s"],
                         "speedup_over_fp64_one_at_a_time":
                             summ_sat["solves_per_s"] / (len(jobs) / total64),
                         "speedup_over_fp64_at_same_residual":
                             summ_sat["solves_per_s"] / (len(jobs) / total_eq)}}
    return _base(args, round(summ["solves_per_s"], 3), "solves/s",
                 round(summ["makespan_s"] / max(1, len(jobs)) * 1e3, 3), cfg, "f32+f64")


# Reduction of the preconditioned residual per FP32 inner solve, per form
# (measured on MI355X; the batched form's lockstep right-hand-side groups wait
# for each other at every correction, so fewer, longer inner solves pay there:
# 1e-1 4.8 s, 1e-3 4.24 s, 1e-4 4.01 s, 1e-5 4.26 s; on H(curl) 1e-1 is best:
# 1e-3 and 1e-2 take 0.85x and 0.68x its solves/s; 0.05, 0.2 and 0.3 gave
# 5.13, 5.10 and 5.15 solves/s against 5.41).
INNER_RTOL = {"hcurl-mixed": 1e-1, "batched": 1e-4, "served": 1e-1}


def run(args):
    if args.inner_rtol is None:
        args = argparse.Namespace(**vars(args))
        args.inner_rtol = INNER_RTOL[args.config]
    if args.config == "hcurl-mixed":
        return hcurl_mixed(args)
    if args.config == "batched":
        return batched(args)
    return served(args)
