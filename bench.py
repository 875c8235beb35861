#!/usr/bin/env python3
"""Benchmark of the MI355X sparse solve path.

Workload: 3D P1 Lagrange Poisson on the unit cube, 64^3 cubes of six Kuhn
tetrahedra each (274,625 vertex rows, 250,047 active, 4,018,753 CSR entries
with the structural zeros of the assembly), FP64, Rainbow CG-SSOR
(ConjugateGradientSolverCSR_Rainbow) with the reference's greedy colouring,
omega = 1.0, precision 1e-10 (tolerance max(1e-10*||b||, 2^-26)), x0 = 0,
seeded U[-1,1) right-hand side on the active rows.

A step is one complete solve with the matrix, colouring and right-hand side
already resident in HBM.  Prints one JSON line with solves/s, the roofline of
the dominant kernel class (chosen by per-launch HIP events in a profiled solve
of the same system, its launch time measured by replaying 200 iterations'
launches of that class in a recorded graph between two HIP events) and the CPU
baseline (the reference's own solver, compiled from its sources, timed on this
host on a bounded sample).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + HBM GB/s (3D P1 Poisson, rtol 1e-10); p50/p99 served latency"
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of each kernel class on the default workload (64^3 in
# the reference's vertex numbering, omega 1, one colour per sweep launch): rocprofv3
# FETCH_SIZE and WRITE_SIZE runs, each scaled by what the
# counter reads for a device-to-device copy of known size (bench.py
# --copy-calibration); commands and per-launch figures are kept in
# profiles/rainbow_cg_ssor_64cube_fp64_hbm_traffic.md.
MEASURED_TRAFFIC = {"spmv": 56382191, "forward": 6474410, "backward": 5118184}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(system, b, d, colouring, precision, omega, budget_s=25.0):
    """The reference's ConjugateGradientSolverCSR_Rainbow on the host cores."""
    import oracle
    from oracle.ref import RefRunner

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    if oracle.ref_available():
        with RefRunner(cores) as ref:
            t0 = time.time()
            x, r, K, times = ref.solve("rainbow", system.A, system.C, system.V, b, aux=d,
                                       omega=omega, colouring=colouring, precision=precision)
            reps = max(0, min(4, int((budget_s - (time.time() - t0)) / max(times[0], 1e-3)) - 1))
            if reps:
                times += ref.solve("rainbow", system.A, system.C, system.V, b, aux=d,
                                   omega=omega, colouring=colouring, precision=precision,
                                   repeat=reps)[3]
        med = float(np.median(times))
        return {"value": 1.0 / med, "unit": "solves/s", "cores": cores, "kind": "reference",
                "sample": f"{len(times)} full solves of the same system (K={K}), median "
                          f"{med:.3f} s, OMP_NUM_THREADS={cores}"}
    # no reference build: the oracle's C restatement, one thread
    t0 = time.time()
    x, r, info = oracle.solve("rainbow", system.A, system.C, system.V, b, aux=d, omega=omega,
                              colouring=colouring, precision=precision)
    dt = time.time() - t0
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"1 full solve (K={info['iterations']}), {dt:.3f} s"}


def forms_cpu_baseline(args, budget_s=25.0):
    """The reference's FP64 ConjugateGradientSolverCSR_Rainbow (it has no
    mixed precision, batching or serving) on the host cores, over a bounded
    sample of the form's own systems and right-hand sides, taken in workload
    order until budget_s of solving has passed."""
    import oracle
    from oracle.ref import RefRunner
    from simplefem_amd import capi
    from simplefem_amd import systems as S
    from simplefem_amd import workloads as W

    if not oracle.ref_available():
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    if args.config == "hcurl-mixed":
        s = S.hcurl3d_whitney_kuhn(48)
        sample = iter([(s, S.rhs_uniform(s, 2026))])
        what = "the same H(curl) system and right-hand side"
    elif args.config == "batched":
        sample = ((j.system, j.rhs[0]) for j in W.batched_systems(args.systems, 4))
        what = f"the first systems of the {args.systems}-system workload, first right-hand side each"
    else:
        sample = ((j.system, j.rhs[0]) for j in
                  W.served_requests(args.requests, args.rate, dofs=(1e3, 2e6)))
        what = "the first requests of the served stream, in arrival order"
    spent, count, rows, iters = 0.0, 0, [], []
    with RefRunner(cores) as ref:
        for s, b in sample:
            d = capi.get_diagonal(s.n, s.A, s.C, s.V)
            rb = capi.Rainbow(s.n, s.A, s.C)
            _, _, K, times = ref.solve("rainbow", s.A, s.C, s.V, b, aux=d, omega=1.0,
                                       colouring=(rb.num_colors, rb.F, rb.B, rb.R),
                                       precision=1e-10)
            spent += times[0]
            count += 1
            rows.append(s.n)
            iters.append(K)
            log(f"cpu baseline: {s.name} K={K} {times[0]:.2f}s")
            if spent >= budget_s:
                break
    return {"value": count / spent, "unit": "solves/s", "cores": cores, "kind": "reference",
            "sample": f"{count} FP64 solves at rtol 1e-10 of {what} (rows {rows}, K {iters}), "
                      f"{spent:.1f} s of solver time, OMP_NUM_THREADS={cores}"}


def copy_calibration(mib):
    """A plain copy of known size: its counters read (bytes read, bytes written)."""
    from simplefem_amd.hipmem import DeviceArray, synchronize

    n = int(mib * 2 ** 20) // 8
    src = DeviceArray.from_host(np.arange(n, dtype=np.float64))
    dst = DeviceArray(n)
    for _ in range(5):
        dst.copy_from(src)
    synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 5}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cube", type=int, default=64, help="cubes per axis")
    ap.add_argument("--omega", type=float, default=1.0)
    ap.add_argument("--precision", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--numbering", default="reference", choices=["reference", "lexicographic"],
                    help="vertex numbering of the 64^3 system: the reference's uniform-refinement "
                         "order (its CSR structure and colouring; default) or x-fastest")
    ap.add_argument("--paired-sweeps", action="store_true",
                    help="two colours per sweep launch where that is cheap (SFA_FLAG_PAIRED_SWEEPS)")
    ap.add_argument("--spmv-lanes", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="SpMV lanes per row; 0: autotune on the system (default)")
    ap.add_argument("--config", default="poisson3d",
                    choices=["poisson3d", "hcurl-mixed", "batched", "served"],
                    help="poisson3d: the headline FP64 Rainbow solve (default); the others "
                         "run the mixed-precision, batched and served forms")
    ap.add_argument("--systems", type=int, default=256, help="batched: number of systems")
    ap.add_argument("--requests", type=int, default=48, help="served: number of requests")
    ap.add_argument("--chunk-steps", type=int, default=2000,
                    help="batched: engine steps per persistent launch between progress checks")
    ap.add_argument("--rate", type=float, default=20.0, help="served: arrivals per second")
    ap.add_argument("--inner-rtol", type=float, default=1e-1,
                    help="mixed precision: reduction of the preconditioned residual per FP32 "
                         "inner solve")
    ap.add_argument("--engine-profile", action="store_true",
                    help="batched: only time the FP64 engine's kernels step by step (HIP events, "
                         "chunks of 2000 steps) and print that table")
    ap.add_argument("--inner-restarts", action="store_true",
                    help="mixed forms: restart every inner CG (plain iterative refinement) "
                         "instead of continuing the previous inner solve's direction")
    ap.add_argument("--route-rows", type=int, default=10000,
                    help="served: requests with at least this many rows get their own System")
    ap.add_argument("--dedicated-workers", type=int, default=12,
                    help="served: concurrent dedicated Systems (0: everything in the batch)")
    ap.add_argument("--no-served-baseline", action="store_true",
                    help="served: skip the plain FP64 pass over the same stream")
    ap.add_argument("--prep-workers", type=int, default=8,
                    help="served: host threads that colour and plan arriving matrices")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (hipMemcpy), to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters of a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.config != "poisson3d":
        from simplefem_amd import forms_bench
        out = forms_bench.run(args)
        if "engine_profile_fp64" in out:
            print(json.dumps(out), flush=True)
            return 0
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = forms_cpu_baseline(args)
        print(json.dumps(out), flush=True)
        return 0

    from simplefem_amd import capi
    from simplefem_amd import systems as S
    from simplefem_amd.hipmem import DeviceArray, synchronize

    t0 = time.time()
    system = S.poisson3d_p1_kuhn(args.cube, numbering=args.numbering)
    b = S.rhs_uniform(system, seed=2026)
    d = capi.get_diagonal(system.n, system.A, system.C, system.V)
    rb = capi.Rainbow(system.n, system.A, system.C)
    colouring = (rb.num_colors, rb.F, rb.B, rb.R)
    log(f"system n={system.n} nnz={system.nnz} colours={rb.num_colors} "
        f"built in {time.time() - t0:.1f}s")

    sysd = capi.System(capi.METHOD_RAINBOW, system.n, system.A, system.C, system.V, d,
                       args.omega, *colouring, paired_sweeps=args.paired_sweeps)
    if args.spmv_lanes:
        sysd.set_spmv(args.spmv_lanes)
        tuning = {"lanes_per_row": args.spmv_lanes, "tuned": False}
    else:
        rep = sysd.tune_spmv()
        tuning = {"lanes_per_row": sysd.stats()["spmv_lanes_per_row"], "tuned": True,
                  "buckets": rep}
        log(f"spmv tuning: {rep}")
    stats = sysd.stats()
    bt = DeviceArray.from_host(b)
    x = DeviceArray(system.n)
    r = DeviceArray(system.n)
    info = {}

    def step():
        x.zero()
        r.zero()
        sysd.solve(x, bt, r, args.precision, info=info)

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup}: K={info['iterations']}")
    synchronize()
    t1 = time.perf_counter()
    iters = 0
    for i in range(args.steps):
        step()
        iters += info["iterations"]
    synchronize()
    elapsed = time.perf_counter() - t1
    log(f"timed {args.steps} solves in {elapsed:.3f}s, K={info['iterations']}")
    solves_per_s = args.steps / elapsed
    hbm_gbs = iters * stats["bytes_per_iteration"] / elapsed / 1e9

    # residual check of the last solve against the tolerance it claims
    xs = x.download()
    res = b - capi.csr_apply(system.n, system.n, system.A, system.C, system.V, xs)
    res[system.dirichlet] = 0.0
    rel_res = float(np.linalg.norm(res) / np.linalg.norm(b))

    roofline = None
    prof = None
    if not args.no_profile:
        x.zero()
        r.zero()
        prof, pinfo = sysd.profile(x, bt, r, args.precision)
        dom = max(prof, key=lambda k: prof[k]["ms"])
        p = prof[dom]
        bytes_per_launch = p["bytes"] / max(p["launches"], 1)
        event_us = p["ms"] * 1e3 / max(p["launches"], 1)  # each launch between two events
        # the class's launches back to back in a recorded graph, timed by two
        # events around the replay: the duration the kernel trace reports
        launch_us = sysd.time_sweeps(dom, 200) if dom in ("forward", "backward") else event_us
        achieved = bytes_per_launch / (launch_us * 1e-6) / 1e9
        default_workload = (args.cube == 64 and args.omega == 1.0 and args.numbering == "reference"
                            and not args.paired_sweeps)
        traffic = MEASURED_TRAFFIC.get(dom) if default_workload else None
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                    "kernel": dom, "launches_per_solve": p["launches"],
                    "avg_launch_us": round(launch_us, 3),
                    "avg_launch_us_timing": "graph replay of 200 iterations' launches of the class"
                                            if dom in ("forward", "backward") else "per-launch events",
                    "event_bracketed_launch_us": round(event_us, 3),
                    "bytes_per_launch": round(bytes_per_launch)}
        log("profile: " + ", ".join(f"{k} {v['launches']}x {v['ms']:.2f}ms" for k, v in
                                     prof.items() if v["launches"]))

    base = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        base = cpu_baseline(system, b, d, colouring, args.precision, args.omega)
        log(f"cpu baseline {base['value']:.3f} solves/s ({base['kind']}, {base['cores']} cores)")

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D P1 Poisson, {args.cube}^3 Kuhn cubes (UnitCube3D refined, "
                        f"{args.numbering} vertex numbering), Rainbow CG-SSOR "
                        f"omega={args.omega}, precision {args.precision:g}, x0=0",
            "numbering": args.numbering,
            "rows": system.n, "active_rows": stats["n_active"], "nnz": system.nnz,
            "colours": rb.num_colors, "iterations": info["iterations"],
            "kernels_per_iteration": stats["kernels_per_iteration"],
            "bytes_per_iteration": stats["bytes_per_iteration"],
            "spmv_mapping": tuning,
            "true_relative_residual": rel_res,
        },
        "hbm_gbs_algorithmic": round(hbm_gbs, 1),
        "roofline": roofline,
        "cpu_baseline": base,
    }
    if prof is not None:
        out["kernel_profile"] = {k: v for k, v in prof.items() if v["launches"]}
    print(json.dumps(out), flush=True)
    sysd.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
