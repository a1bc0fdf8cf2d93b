"""Benchmark of the Bacchus hot path on one MI355X.

Workload: the 3 proton + 3 neutron system at Nmax 6 (M = 0, positive parity)
with seeded Gaussian 2-body (sigma 1 MeV) and 3-body (sigma 0.1 MeV)
elements and the one-body term 20(N+1.5) MeV: D = 197,822.

A step is one solve of the lowest 5 eigenvalues to 1e-6 relative accuracy
against FP64 reference eigenvalues (the FP64 GPU path run to 300 Lanczos
iterations with full reorthogonalization) with the fast path: FP32-stored
matrix elements with FP32 products and accumulation, FP64 Lanczos
recurrence, full FP64 reorthogonalization and the FP64 tridiagonal solve on
the GPU.  The plain FP64 GPU path is timed to the same criterion beside it.

The line also carries the roofline of the dominant kernel (the FP64 matvec,
HBM-bound: algorithmic bytes = 8 sum|V| + 12 sum|L| + 16 D per launch over
its HIP-event time on the launch stream) and the CPU oracle timed on the
host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-5-eigvals @1e-6 rel vs FP64 GPU path; matvec HBM GB/s % of peak"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="li6_nmax6_3b")
    ap.add_argument("--iterations", type=int, default=300)
    ap.add_argument("--eigenvalues", type=int, default=5)
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--matvecs", type=int, default=20, help="timed matvecs for the roofline")
    ap.add_argument("--cpu-matvecs", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--matvec-only", type=int, default=0,
                    help="only run this many FP64 matvecs (for profilers) and exit")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only copy this many MiB device-to-device 3 times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(workload):
    """HBM bytes per matvec launch measured by rocprofv3 --pmc
    (profiles/matvec_<workload>_fp64_traffic.txt, 'traffic_bytes: N')."""
    path = os.path.join(ROOT, "profiles", f"matvec_{workload}_fp64_traffic.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding is a design, not a build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from jupiterncsm_amd import bacchus
    from jupiterncsm_amd.inputs import SYSTEMS, build

    assert torch.cuda.is_available(), "no GPU"
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB x 3")
        return
    t0 = time.time()
    inputs = build(SYSTEMS[args.workload])
    log(f"inputs {args.workload}: {inputs.summary()} in {time.time() - t0:.1f}s")
    D = inputs.dimension
    op64 = bacchus.Operator.from_inputs(inputs, bacchus.FP64)
    stats = op64.stats()
    x = torch.as_tensor(np.random.default_rng(1).standard_normal(D), device="cuda")
    y = torch.empty_like(x)

    # --- roofline of the FP64 matvec ------------------------------------
    stream = torch.cuda.current_stream()
    for _ in range(3):
        op64.matvec(x, y, stream=stream)
    torch.cuda.synchronize()
    n_mv = args.matvec_only or args.matvecs
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(n_mv):
        op64.matvec(x, y, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    mv_ms = e0.elapsed_time(e1) / n_mv
    achieved = stats["algorithmic_bytes"] / (mv_ms * 1e-3) / 1e9
    log(f"fp64 matvec {mv_ms:.3f} ms, {achieved:.0f} GB/s algorithmic, "
        f"{stats['flops'] / mv_ms / 1e9:.1f} TFLOP/s")
    if args.matvec_only:
        return

    op32 = bacchus.Operator.from_inputs(inputs, bacchus.FP32)

    # --- FP64 reference eigenvalues (GPU FP64 path, full length) --------
    t0 = time.time()
    ref = bacchus.diagonalize(op64, args.iterations, num_eigenvalues=args.eigenvalues,
                              start=bacchus.START_RANDOM, seed=1,
                              reorthogonalization=bacchus.REORTH_FULL,
                              convergence=bacchus.CONVERGE_NONE)
    targets = ref.eigenvalues
    log(f"reference eigenvalues ({ref.iterations} its, {time.time() - t0:.2f}s): {targets}")

    def solve(op, seed):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = bacchus.diagonalize(op, args.iterations, num_eigenvalues=args.eigenvalues,
                                start=bacchus.START_RANDOM, seed=seed,
                                reorthogonalization=bacchus.REORTH_FULL,
                                convergence=bacchus.CONVERGE_TARGETS,
                                tolerance=args.tolerance, target_eigenvalues=targets,
                                check_every=1)
        torch.cuda.synchronize()
        return time.perf_counter() - t, r

    # plain FP64 GPU path to the same criterion
    t64, r64 = solve(op64, 2)
    t64, r64 = solve(op64, 2)
    log(f"fp64 path: {t64:.3f}s, {r64.iterations} its, errors "
        f"{np.abs(r64.eigenvalues - targets) / np.abs(targets)}")

    for _ in range(args.warmup):
        solve(op32, 2)
    times, last = [], None
    for _ in range(args.steps):
        dt, last = solve(op32, 2)
        times.append(dt)
    t_fast = float(np.mean(times))
    errs = np.abs(last.eigenvalues - targets) / np.abs(targets)
    converged = last.iterations < args.iterations or bool(np.all(errs <= args.tolerance))
    log(f"fast path: {t_fast:.3f}s, {last.iterations} its, errors {errs}")

    # --- CPU oracle on the host cores (bounded sample) -------------------
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        cores = len(os.sched_getaffinity(0))
        prob = O.Problem.from_inputs(inputs)  # views into `inputs`, kept alive
        xv = x.cpu().numpy()
        O.matvec(prob, xv, threads=cores)
        t = time.perf_counter()
        for _ in range(args.cpu_matvecs):
            O.matvec(prob, xv, threads=cores)
        cpu_mv = (time.perf_counter() - t) / args.cpu_matvecs
        cpu = {"value": cpu_mv * r64.iterations, "unit": "s", "cores": cores,
               "kind": "port",
               "sample": f"{args.cpu_matvecs} oracle FP64 matvecs of the workload "
                         f"({cpu_mv:.2f} s each, OpenMP, per-thread replicas) x "
                         f"{r64.iterations} Lanczos iterations of the FP64 solve; "
                         "Lanczos vector work not included"}
        log(f"cpu oracle matvec {cpu_mv:.2f}s on {cores} threads")

    line = {
        "metric": METRIC,
        "value": round(t_fast, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_fast * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: 3p+3n Nmax6 M=0 +parity, 2+3-body, "
                               f"D={D}", "eigenvalues": args.eigenvalues,
                   "tolerance": args.tolerance, "max_iterations": args.iterations,
                   "fast_path": "fp32 elements + fp64 lanczos, full reorth, random start"},
        "fp64_gpu_seconds": round(t64, 4),
        "fp64_gpu_iterations": r64.iterations,
        "fast_iterations": last.iterations,
        "speedup_vs_fp64_gpu": round(t64 / t_fast, 3),
        "converged": converged,
        "eigenvalues_fp64_reference": [float(v) for v in targets],
        "relative_errors_fast": [float(v) for v in errs],
        "relative_errors_fp64": [float(v) for v in
                                 np.abs(r64.eigenvalues - targets) / np.abs(targets)],
        "matvec_fp64_ms": round(mv_ms, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(args.workload)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
