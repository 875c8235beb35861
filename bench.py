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
of the same system; its launch duration is the kernel trace's, from the
committed profile whose source digest matches the sources built, with the
live graph-replay time beside it) and the CPU baseline (the reference's own
solver, compiled from its sources, timed on this host on a bounded sample, at
all threads of the host's share and at one).  The same line carries the
host-array drop-in call's cost per call, the 2D natural-order SSOR
configuration (GPU and the reference at one and all threads) and, under
"forms", the mixed-precision, batched and served configurations.
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
# The kernel trace and the HBM counters of the default workload (64^3 in the
# reference's vertex numbering, omega 1) are taken by rocprofv3 runs of this
# script and summarised in a profile whose "roofline-stamp:" line carries, per
# kernel class, the traced mean duration of a launch and its HBM bytes, plus
# a digest of the sources that make those kernels and launches.  The
# roofline uses them only while the digest matches the sources built now.
STAMPED_PROFILE = os.path.join(ROOT, "profiles", "rainbow_cg_ssor_64cube_fp64_kernels.md")
KERNEL_SOURCES = ["kernels.hip", "kernels.hpp", "device_common.hpp", "launch.hpp",
                  "system.cpp", "plan.cpp", "plan.hpp", "host_common.hpp"]


def kernel_source_digest():
    """sha256 over the sources of the single-system kernels and their launches."""
    import hashlib

    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "simplefem_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def stamped_profile():
    """The profile's stamp if it was taken from the sources built now, else None."""
    try:
        with open(STAMPED_PROFILE) as f:
            for line in f:
                if line.startswith("roofline-stamp:"):
                    stamp = json.loads(line.split(":", 1)[1])
                    return stamp if stamp.get("source_digest") == kernel_source_digest() else None
    except (OSError, ValueError):
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU (lscpu) and the threads a CPU baseline runs with: the
    CPUs this process may use, capped by OMP_NUM_THREADS where the
    environment sets it (the CPU share of one GPU on a shared node)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(key):
        try:
            return int(info.get(key, "0").split()[0])
        except ValueError:
            return 0

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail
    return {"model": info.get("Model name"), "sockets": num("Socket(s)"),
            "physical_cores": num("Socket(s)") * num("Core(s) per socket"),
            "threads_per_core": num("Thread(s) per core"), "cpus_available": avail,
            "omp_num_threads_env": env or None, "threads": max(1, threads),
            "binding": "OMP_PROC_BIND=close OMP_PLACES=cores"}


def _ref_median(threads, variant, system, b, d, colouring, precision, omega, budget_s):
    """Median seconds of the reference's solver over up to 5 solves within
    budget_s, at `threads` OpenMP threads, bound one per core."""
    from oracle.ref import RefRunner

    with RefRunner(threads, bind=True) as ref:
        t0 = time.time()
        _, _, K, times = ref.solve(variant, system.A, system.C, system.V, b, aux=d, omega=omega,
                                   colouring=colouring, precision=precision)
        reps = max(0, min(4, int((budget_s - (time.time() - t0)) / max(times[0], 1e-3)) - 1))
        if reps:
            times += ref.solve(variant, system.A, system.C, system.V, b, aux=d, omega=omega,
                               colouring=colouring, precision=precision, repeat=reps)[3]
    return float(np.median(times)), K, len(times)


def cpu_baseline(system, b, d, colouring, precision, omega, budget_s=25.0):
    """The reference's ConjugateGradientSolverCSR_Rainbow on the host cores
    (all threads of this process's CPU share, and one thread)."""
    import oracle

    cpu = host_cpu()
    cores = cpu["threads"]
    if oracle.ref_available():
        med, K, n = _ref_median(cores, "rainbow", system, b, d, colouring, precision, omega,
                                0.4 * budget_s)
        med1, K1, n1 = _ref_median(1, "rainbow", system, b, d, colouring, precision, omega,
                                   0.6 * budget_s)
        return {"value": 1.0 / med, "unit": "solves/s", "cores": cores, "kind": "reference",
                "sample": f"{n} full solves of the same system (K={K}), median {med:.3f} s, "
                          f"OMP_NUM_THREADS={cores}, OMP_PROC_BIND=close, OMP_PLACES=cores",
                "one_thread": {"value": 1.0 / med1, "unit": "solves/s", "cores": 1,
                               "sample": f"{n1} full solves (K={K1}), median {med1:.3f} s"},
                "host": cpu,
                "build": "oracle/Makefile: g++ -O3 -march=x86-64-v3 -fopenmp -DNDEBUG "
                         "-ffp-contract=off (built with the library and shipped with it, so "
                         "for a portable x86-64 level rather than -march=native)"}
    # no reference build: the oracle's C restatement, one thread
    t0 = time.time()
    x, r, info = oracle.solve("rainbow", system.A, system.C, system.V, b, aux=d, omega=omega,
                              colouring=colouring, precision=precision)
    dt = time.time() - t0
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"1 full solve (K={info['iterations']}), {dt:.3f} s"}


def _per_call(fn, reps):
    """Seconds per call of fn(): the median of reps calls after one warm-up."""
    fn()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def dropin_per_call(system, b, d, colouring, omega, precision, reps=3):
    """The raw drop-in call on host arrays (sfa_ConjugateGradientSolverCSR_Rainbow:
    argument checks, device-order plan, upload, solve, download), as the
    reference's callers make it, seconds per call."""
    from simplefem_amd import capi

    info = {}

    def call():
        x = np.zeros(system.n)
        r = np.zeros(system.n)
        capi.ConjugateGradientSolverCSR_Rainbow(system.n, x, b, system.A, system.C, system.V, r,
                                                precision, -3, d, omega, *colouring, info=info)

    s = _per_call(call, reps)
    return {"seconds_per_call": round(s, 5), "calls_per_s": round(1.0 / s, 3),
            "iterations": info["iterations"], "calls": reps}


def c1_natural_ssor(precision=1e-10, omega=1.0, budget_s=12.0, with_cpu=True):
    """2D P1 Poisson, 129^2 vertices in the reference's numbering,
    ConjugateGradientSolverCSR_SSOR (natural-order SSOR) omega 1: the GPU
    with the system resident, the raw drop-in call on host arrays, and the
    reference's solver at one and all threads of the host."""
    import oracle
    from simplefem_amd import capi
    from simplefem_amd import systems as S
    from simplefem_amd.hipmem import DeviceArray, synchronize

    s = S.poisson2d_p1(7, numbering="reference")
    b = S.rhs_uniform(s, seed=2026)
    d = capi.get_diagonal(s.n, s.A, s.C, s.V)
    out = {"workload": "2D P1 Poisson, 129^2 vertices (UnitSquare2D_simple refined 7 times, "
                       "reference numbering), ConjugateGradientSolverCSR_SSOR omega 1, "
                       f"precision {precision:g}, x0 = 0", "rows": s.n, "nnz": s.nnz}
    with capi.System(capi.METHOD_SSOR, s.n, s.A, s.C, s.V, d, omega) as sysd:
        st = sysd.stats()
        bt, x, r = DeviceArray.from_host(b), DeviceArray(s.n), DeviceArray(s.n)
        info = {}

        def solve():
            x.zero()
            r.zero()
            sysd.solve(x, bt, r, precision, info=info)
            synchronize()

        t = _per_call(solve, 3)
        out["gpu_resident"] = {"solves_per_s": round(1.0 / t, 3), "seconds": round(t, 5),
                               "iterations": info["iterations"],
                               "levels_per_sweep": st["groups_forward"],
                               "kernels_per_iteration": st["kernels_per_iteration"]}
    info = {}

    def call():
        xh, rh = np.zeros(s.n), np.zeros(s.n)
        capi.ConjugateGradientSolverCSR_SSOR(s.n, xh, b, s.A, s.C, s.V, rh, precision, -3, d,
                                             omega, info=info)

    t = _per_call(call, 3)
    out["gpu_dropin_call"] = {"seconds_per_call": round(t, 5), "calls_per_s": round(1.0 / t, 3),
                              "iterations": info["iterations"]}
    if with_cpu and oracle.ref_available():
        cpu = host_cpu()
        for name, threads in (("cpu_reference_all_threads", cpu["threads"]),
                              ("cpu_reference_one_thread", 1)):
            med, K, n = _ref_median(threads, "ssor", s, b, d, None, precision, omega,
                                    budget_s / 2)
            out[name] = {"solves_per_s": round(1.0 / med, 3), "seconds": round(med, 5),
                         "iterations": K, "threads": threads, "solves": n}
    return out


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


def barrier_cost(rounds=2000):
    """Grid-wide barrier inside one resident launch against a kernel boundary
    (sfa_grid_barrier_cost) on grids of 256-1024 workgroups of 256 threads:
    one JSON line per grid and mode."""
    from simplefem_amd import capi

    names = {0: "one counter", 1: "8 group counters + top", 2: "kernel boundary (graph replay)"}
    for blocks in (256, 512, 1024):
        for mode in (0, 1, 2):
            try:
                us, err = capi.grid_barrier_cost(blocks, 256, rounds, mode)
                row = {"blocks": blocks, "threads": 256, "mode": names[mode], "rounds": rounds,
                       "us_per_round": round(us, 3), "errors": err}
            except capi.SfaError as e:
                row = {"blocks": blocks, "threads": 256, "mode": names[mode], "error": str(e)}
            print(json.dumps(row), flush=True)
            if row.get("errors"):
                return 1
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
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the per-call drop-in timing and the 2D natural-order SSOR lines")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the precision, batched and served configurations (the 'forms' "
                         "key of the default line)")
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
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="mixed precision: reduction of the preconditioned residual per FP32 "
                         "inner solve (default: the form's own, measured best: hcurl-mixed "
                         "1e-1, batched 1e-4, served 1e-1)")
    ap.add_argument("--engine-profile", action="store_true",
                    help="batched: only time the FP64 engine's kernels step by step (HIP events, "
                         "chunks of 2000 steps) and print that table; hcurl-mixed: only run one "
                         "FP64 solve with every kernel launched on its own (no graphs) between "
                         "HIP events and print the per-class table")
    ap.add_argument("--only-system", default="",
                    help="batched: keep only the workload's systems of this name (e.g. "
                         "sullivan2d_L8_r3, the long tail), for profiling")
    ap.add_argument("--profile-steps", type=int, default=400000,
                    help="batched --engine-profile: stop after this many steps")
    ap.add_argument("--rhs-layout", default="grouped", choices=["grouped", "separate"],
                    help="batched: a system's right-hand sides as one interleaved slot group "
                         "(sfa_batch_submit_group; default) or as separate slots")
    ap.add_argument("--inner-restarts", action="store_true",
                    help="mixed forms: restart every inner CG (plain iterative refinement) "
                         "instead of continuing the previous inner solve's direction")
    ap.add_argument("--route-rows", type=int, default=30000,
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
    ap.add_argument("--barrier-cost", action="store_true",
                    help="only time a grid barrier inside one resident launch against a kernel "
                         "boundary (sfa_grid_barrier_cost) and print one line per grid")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.barrier_cost:
        return barrier_cost()
    if args.config != "poisson3d":
        from simplefem_amd import forms_bench
        out = forms_bench.run(args)
        if "engine_profile_fp64" in out or "hcurl_kernel_profile_fp64" in out:
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
        # events around the replay (live, this run)
        replay_us = sysd.time_sweeps(dom, 200) if dom in ("forward", "backward") else event_us
        default_workload = (args.cube == 64 and args.omega == 1.0 and args.numbering == "reference"
                            and not args.paired_sweeps)
        stamp = stamped_profile() if default_workload else None
        traced = stamp["classes"].get(dom) if stamp else None
        if traced:  # the kernel trace's mean duration of a launch of the class
            launch_us, source = traced["traced_us"], "rocprofv3 kernel trace (" + stamp["profile"] + ")"
        else:
            launch_us, source = replay_us, "graph replay of 200 iterations' launches (no current trace)"
        achieved = bytes_per_launch / (launch_us * 1e-6) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                    "traffic": traced["hbm_bytes"] if traced else None,
                    "kernel": dom, "launches_per_solve": p["launches"],
                    "bytes_per_launch": round(bytes_per_launch),
                    "avg_launch_us": round(launch_us, 3), "avg_launch_us_source": source,
                    "replay_launch_us": round(replay_us, 3),
                    "replay_frac": round(bytes_per_launch / (replay_us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
                    "event_bracketed_launch_us": round(event_us, 3),
                    "stamp_source_digest": kernel_source_digest()}
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
            # the System's creation (a drop-in call's first-call cost), ms
            "setup_ms": dict(zip(("host_plan", "host_copy", "device", "fp32_copy"),
                                 [round(v, 1) for v in stats["setup_ms"]])),
            "spmv_mapping": tuning,
            "true_relative_residual": rel_res,
        },
        "hbm_gbs_algorithmic": round(hbm_gbs, 1),
        "roofline": roofline,
        "cpu_baseline": base,
    }
    if prof is not None:
        out["kernel_profile"] = {k: v for k, v in prof.items() if v["launches"]}
    sysd.close()
    if not args.no_extras:
        # what the reference's callers see: one raw call on host arrays
        log("drop-in call on host arrays ...")
        out["dropin_per_call"] = dropin_per_call(system, b, d, colouring, args.omega,
                                                 args.precision)
        log(f"drop-in call: {out['dropin_per_call']}")
        log("2D natural-order SSOR (129^2) ...")
        out["c1_ssor_2d"] = c1_natural_ssor(args.precision, args.omega,
                                            with_cpu=not args.no_cpu_baseline)
        log(f"c1: {out['c1_ssor_2d']}")
    if not args.no_forms:
        out["forms"] = forms_summary(args)
    print(json.dumps(out), flush=True)
    return 0


def forms_summary(args):
    """The precision, batched and served configurations, each run as
    `bench.py --config <name> --no-cpu-baseline` runs it, in one line each
    (per-mapping tuning tables left out)."""
    from simplefem_amd import forms_bench

    def strip(d):
        if isinstance(d, dict):
            return {k: strip(v) for k, v in d.items() if k != "spmv_mapping"}
        return d

    out = {}
    for name in ("hcurl-mixed", "batched", "served"):
        fargs = argparse.Namespace(**vars(args))
        fargs.config = name
        fargs.no_cpu_baseline = True
        log(f"form {name} ...")
        t0 = time.perf_counter()
        res = forms_bench.run(fargs)
        out[name] = {"value": res["value"], "unit": res["unit"], "dtype": res["dtype"],
                     "ms_per_solve": res["ms_per_step"], "seconds": round(time.perf_counter() - t0, 1),
                     "config": strip(res["config"])}
        log(f"form {name}: {res['value']} {res['unit']}")
    return out


if __name__ == "__main__":
    sys.exit(main())
