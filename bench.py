"""Benchmark of the Bacchus hot path on one MI355X.

Workload: the 3 proton + 3 neutron system at Nmax 6 (M = 0, positive parity)
with seeded Gaussian 2-body (sigma 1 MeV) and 3-body (sigma 0.1 MeV)
elements and the one-body term 20(N+1.5) MeV: D = 197,822.

A step is one solve of the lowest 5 eigenvalues to 1e-6 relative accuracy
with the fast path; the reference eigenvalues are the oracle's 300-iteration
FP64 Lanczos from a seeded start (tests/golden/lanczos300_<workload>.npz),
which the FP64 GPU path run the same way is checked against:

- FP32-stored matrix elements with FP32 products and accumulation;
- block Lanczos, 5 vectors per block (one per eigenvalue; 8 took up to three
  block steps at held-out points where 5 takes one), each block step one
  batched launch,
  FP64 inner products and block tridiagonal solve, stopping on its own
  residual bound |r_k|^2 / gap_k <= 1e-6 |theta_k| (gap_k corrected by the
  other Ritz values' residuals);
- started from a projected family: the band Ritz vectors of batched solves
  of the family H(c1, c2) = H1 + c1 V2 + c2 V3 (--projected-basis of them,
  orthonormalized to their numerical rank) with H1, V2, V3 projected onto
  them once; the start block is the Ritz vectors of H(c1, c2) in that span
  (solves and projections untimed, reported as setup_seconds and amortized
  per solve).

The plain FP64 GPU path (single-vector Lanczos, seeded start, full
reorthogonalization) is timed to the target criterion beside it, and so are
two starts the MLPs (2-256x3-64, GELU, BF16 weights) predict over a stored
basis of 64 vectors: the single-vector FP32 Lanczos from one network's
start vector (mlp_start), and the same block Lanczos from the block of
--block networks, each predicting the projection of its own fixed vector
onto the band (mlp_block_start; its own residual bound, with the target
criterion reported only as a check).  The fast path and the MLP block start
are also run at --heldout further (c1, c2) points, each against its own FP64
reference.  The FP32- and BF16-stored operators (and a mixed store) go
through the reference's own solve (300 iterations, same start) as well, and
the line states their error per eigenvalue (relative_errors_by_storage).

Other modes (one JSON line each):
  --batched   64 Hamiltonians H(c1, c2) over the 2+2 systems at Nmax 4, 6,
              8 in one batched Lanczos (persistent batched matvec, captured
              graph) against 64 sequential plain FP64 solves;
  --learned   the learned start of the 3+3 Nmax 6 ground state: 512 batched
              training solves, 64 held-out points, iterations and time to
              1e-6 of the lowest eigenvalue against a random start.
The default line carries both under "configs", with C1/C2.

The line also carries the roofline of the metric's plain matvec (the FP64
tile kernel, HBM-bound: algorithmic bytes = 8 sum|V| + 12 sum|L| + 16 D per
launch over its HIP-event time on the launch stream, the HBM traffic from
the rocprofv3 summary in profiles/), the same figures for the fast path's
batched matvec of one block (block_matvec), and the CPU baseline: the oracle's
FP64 Lanczos of the workload on the host's cores for the iteration count the
FP64 GPU path took.
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
    ap.add_argument("--train-points", type=int, default=128,
                    help="family solves the learned start is trained on")
    ap.add_argument("--train-iterations", type=int, default=80)
    ap.add_argument("--no-learned", action="store_true",
                    help="no family solves, MLPs or projected start: the fast path is the "
                         "single-vector FP32 Lanczos from the seeded start")
    ap.add_argument("--start-mix", type=float, default=1e-4,
                    help="relative size of the random component mixed into the MLP start")
    ap.add_argument("--block", type=int, default=5,
                    help="vectors per block of the fast path (at least --eigenvalues)")
    ap.add_argument("--projected-basis", type=int, default=128,
                    help="stored band Ritz vectors the projected start's basis is made of")
    ap.add_argument("--block-mlp-steps", type=int, default=8000,
                    help="training steps of each MLP of the learned start block (0: none)")
    ap.add_argument("--max-blocks", type=int, default=16)
    ap.add_argument("--heldout", type=int, default=8,
                    help="further (c1, c2) points the fast path is checked on, each against "
                         "its own FP64 reference")
    ap.add_argument("--batched", action="store_true", help="batched 64-Hamiltonian sweep")
    ap.add_argument("--batched-matvec-only", type=int, default=0,
                    help="only time this many batched matvecs of the sweep's batch (profiling)")
    ap.add_argument("--learned", action="store_true", help="learned-start measurement")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the C2 / C4 / C5 sections of the default line")
    ap.add_argument("--c5-train", type=int, default=512, help="C5 training points")
    ap.add_argument("--c5-test", type=int, default=64, help="C5 held-out points")
    ap.add_argument("--matvec-only", type=int, default=0,
                    help="only run this many FP64 matvecs (for profilers) and exit")
    ap.add_argument("--block-only", type=int, default=0,
                    help="only run this many batched FP32 matvecs of one fast-path block "
                         "(--block vectors, autotuned) and exit (for profilers)")
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
    if args.batched or args.batched_matvec_only:
        return batched_sweep(args)
    if args.learned:
        return learned_start(args)
    t0 = time.time()
    inputs = build(SYSTEMS[args.workload], device=True)
    gen_device = time.time() - t0
    log(f"inputs {args.workload}: {inputs.summary()} in {gen_device:.2f}s (index lists and "
        f"elements on the GPU)")
    gen_host = None
    if not args.matvec_only and not args.block_only:
        t0 = time.time()
        build(SYSTEMS[args.workload])
        gen_host = time.time() - t0
        log(f"the same inputs from the host generator: {gen_host:.2f}s")
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
    if args.block_only:
        with bacchus.Batch([op32], [args.block]) as blk:
            blk.set_scales(0, np.ones((args.block, 3)))
            tuned = blk.autotune(3)
            xs = [torch.randn(D, args.block, dtype=torch.float32, device="cuda")]
            ys = [torch.empty_like(xs[0])]
            e0.record(stream)
            for _ in range(args.block_only):
                blk.matvec(xs, ys, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            log(f"block matvec ({args.block} vectors, {tuned}): "
                f"{e0.elapsed_time(e1) / args.block_only:.3f} ms")
        return
    for _ in range(3):
        op32.matvec(x, y, stream=stream)
    e0.record(stream)
    for _ in range(n_mv):
        op32.matvec(x, y, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    mv32_ms = e0.elapsed_time(e1) / n_mv
    log(f"fp32 matvec {mv32_ms:.3f} ms")

    # --- reference eigenvalues: the oracle's 300-iteration Lanczos from the
    # fixture's seeded start (tests/golden/lanczos300_<workload>.npz); the
    # FP64 GPU path's own 300-iteration run from the same start is checked
    # against it ----------------------------------------------------------
    fx = fixture(args.workload)
    sv_host = start_vector(D, fx["seed"] if fx is not None else FIXTURE_SEED)
    sv = torch.as_tensor(sv_host, device="cuda")
    t0 = time.time()
    ref = bacchus.diagonalize(op64, args.iterations, num_eigenvalues=args.eigenvalues,
                              start=bacchus.START_GIVEN, start_vector=sv,
                              reorthogonalization=bacchus.REORTH_FULL,
                              convergence=bacchus.CONVERGE_NONE)
    if fx is not None:
        targets = fx["eigenvalues"][:args.eigenvalues]
        ref_vs_fixture = [float(v) for v in np.abs(ref.eigenvalues - targets) / np.abs(targets)]
        reference = {"source": f"oracle fixture tests/golden/lanczos300_{args.workload}.npz "
                               f"({int(fx['iterations'])} iterations, seed {int(fx['seed'])})",
                     "fp64_gpu_300_relative_errors": ref_vs_fixture}
    else:
        targets = ref.eigenvalues
        reference = {"source": "FP64 GPU path, 300 iterations (no oracle fixture)"}
    log(f"reference eigenvalues {targets}; FP64 GPU 300 its ({time.time() - t0:.2f}s): "
        f"{reference}")

    # --- every storage precision through the same solve (300 iterations,
    # full reorthogonalization, same start, FP64 Lanczos): relative error of
    # each eigenvalue against the reference ------------------------------
    by_precision = {}
    for pname, prec in (("fp32", bacchus.FP32), ("bf16", bacchus.BF16)):
        opx = op32 if prec == bacchus.FP32 else bacchus.Operator.from_inputs(inputs, prec)
        rx = bacchus.diagonalize(opx, args.iterations, num_eigenvalues=args.eigenvalues,
                                 start=bacchus.START_GIVEN, start_vector=sv,
                                 reorthogonalization=bacchus.REORTH_FULL,
                                 convergence=bacchus.CONVERGE_NONE)
        if opx is not op32:
            opx.close()
        by_precision[pname] = [float(v) for v in np.abs(rx.eigenvalues - targets)
                               / np.abs(targets)]
        log(f"{pname} storage: relative eigenvalue errors {by_precision[pname]}")
    # a mixed store: the 3-body elements (sigma 0.1 MeV) as BF16 values, the 1-
    # and 2-body ones as FP32 (held here in an FP32 operator, so the numerics
    # are those of the mixed store, not its bytes)
    opm = mixed_operator(inputs, (3,))
    rx = bacchus.diagonalize(opm, args.iterations, num_eigenvalues=args.eigenvalues,
                             start=bacchus.START_GIVEN, start_vector=sv,
                             reorthogonalization=bacchus.REORTH_FULL,
                             convergence=bacchus.CONVERGE_NONE)
    opm.close()
    by_precision["bf16_3body_fp32_rest"] = [float(v) for v in np.abs(rx.eigenvalues - targets)
                                            / np.abs(targets)]
    log(f"BF16 3-body / FP32 1- and 2-body elements: relative eigenvalue errors "
        f"{by_precision['bf16_3body_fp32_rest']}")

    # --- learned start: trained on batched solves of the family (untimed) -
    model = fam = bt = mblock = block_mv = None
    train_s = 0.0
    if not args.no_learned:
        from jupiterncsm_amd import learned as L
        t0 = time.time()
        pts = L.family_points(args.train_points, 1)
        sol = L.solve_family(op32, pts, iterations=args.train_iterations,
                             num_eigenvalues=args.eigenvalues, target="band",
                             band=args.eigenvalues,
                             keep_ritz=-(-args.projected_basis // args.eigenvalues),
                             band_references=args.block if args.block_mlp_steps else 0)
        model = L.fit(sol, basis_size=64, steps=3000, target="band", precision=MLP_PRECISION)
        if args.block_mlp_steps:
            mblock = L.fit_block(sol, basis_size=64, steps=args.block_mlp_steps,
                                 precision=MLP_PRECISION)
            log(f"MLP start block: {args.block} networks, losses "
                f"{min(mblock.train_losses):.2e}..{max(mblock.train_losses):.2e}")
        t1 = time.time()
        # the projected family's own stored basis: the band Ritz vectors of
        # the first points (more of them than the networks' 64 fit the band
        # at a new point better at the same numerical rank)
        fam_basis, fam_rank = L.orthonormal_basis(sol.ritz[:args.projected_basis].contiguous())
        fam = L.ProjectedFamily.build(op64, fam_basis, fam_rank)
        train_s = time.time() - t0
        log(f"learned start: {args.train_points} batched family solves "
            f"({sol.seconds:.1f}s), MLP loss {model.train_loss:.2e}, MLP basis rank {model.rank}, "
            f"projected basis {args.projected_basis} vectors of rank {fam_rank}, "
            f"projections {time.time() - t1:.1f}s, total {train_s:.1f}s")
        bt = bacchus.Batch([op32], [args.block])
        bt.set_scales(0, np.ones((args.block, 3)))
        tuned = bt.autotune(3)
        bst = bt.stats()
        block_mv = {"vectors": args.block, "ms": round(tuned["ms"], 4),
                    # its algorithmic bytes (elements and lists once, the 8
                    # vectors in and out) over the launch time
                    "algorithmic_gb_s": round(bst["algorithmic_bytes"] / (tuned["ms"] * 1e-3) / 1e9, 1),
                    "hbm_frac": round(bst["algorithmic_bytes"] / (tuned["ms"] * 1e-3) / 1e9
                                      / HBM_PEAK_GBS, 4),
                    "tflop_s": round(bst["flops"] / (tuned["ms"] * 1e-3) / 1e12, 2),
                    "mfma": bt.mfma_stats()}
        log(f"block matvec ({args.block} vectors) autotuned: {tuned['ms']:.2f} ms, "
            f"matrix cores {block_mv['mfma']}")

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t, r

    def timed_median(fn, n=3):
        """Median time of n runs (after the caller's warm-up) and the last
        result: one stall of the host does not make a point's figure."""
        runs = [timed(fn) for _ in range(n)]
        return float(np.median([dt for dt, _ in runs])), runs[-1][1]

    def single(op, seed, start=None):
        kw = dict(num_eigenvalues=args.eigenvalues, reorthogonalization=bacchus.REORTH_FULL,
                  convergence=bacchus.CONVERGE_TARGETS, tolerance=args.tolerance,
                  target_eigenvalues=targets, check_every=1)
        if start is not None:
            return bacchus.diagonalize(op, args.iterations, start=bacchus.START_GIVEN,
                                       start_vector=start, **kw)
        return bacchus.diagonalize(op, args.iterations, start=bacchus.START_RANDOM, seed=seed,
                                   **kw)

    def mlp_solve():
        sv = model.start_vector(1.0, 1.0, mix=args.start_mix)
        return single(op32, 2, sv)

    def block_solve(c1=1.0, c2=1.0, start=None, targets_at=None):
        X = (start or fam).start_block(c1, c2, args.block)
        if targets_at is not None:
            return bt.block_lanczos(args.max_blocks, args.eigenvalues, start_block=X,
                                    convergence=bacchus.CONVERGE_TARGETS, targets=targets_at,
                                    tolerance=args.tolerance)
        return bt.block_lanczos(args.max_blocks, args.eigenvalues, start_block=X,
                                convergence=bacchus.CONVERGE_RESIDUALS,
                                tolerance=args.tolerance)

    def mlp_block_runs(c1, c2, ref_ev):
        """(seconds, blocks, max relative error, converged) of the block solve
        from the MLP start block: on its residual bound, then on the target
        criterion (converged: the criterion, not --max-blocks, stopped it)."""
        out = []
        for tg in (None, ref_ev):
            f = lambda: block_solve(c1, c2, mblock, tg)  # noqa: E731
            timed(f)
            dt, r = timed_median(f)
            out.append((dt, r.blocks,
                        float((np.abs(r.eigenvalues - ref_ev) / np.abs(ref_ev)).max()),
                        bool(r.converged)))
        return out

    def rel(ev):
        return np.abs(np.asarray(ev) - targets) / np.abs(targets)

    # plain FP64 GPU path (fixture start) to the target criterion
    t64, r64 = timed(lambda: single(op64, 2, sv))
    t64, r64 = timed(lambda: single(op64, 2, sv))
    log(f"fp64 path: {t64:.3f}s, {r64.iterations} its, errors {rel(r64.eigenvalues)}")

    mlp = None
    if model is not None:
        timed(mlp_solve)
        tm, rm = timed(mlp_solve)
        mlp = {"seconds": round(tm, 4), "iterations": rm.iterations,
               "speedup_vs_fp64_gpu": round(t64 / tm, 3),
               "relative_errors": [float(v) for v in rel(rm.eigenvalues)]}
        log(f"fp32 single-vector Lanczos from the MLP start: {tm:.3f}s, {rm.iterations} its, "
            f"errors {rel(rm.eigenvalues)}")

    mlp_block = None
    if mblock is not None:
        (tr, br, er, cr), (tt, btg, et, _) = mlp_block_runs(1.0, 1.0, targets)
        mlp_block = {"seconds": round(tr, 4), "blocks": br, "max_relative_error": er,
                     "converged": cr,
                     "speedup_vs_fp64_gpu": round(t64 / tr, 3),
                     # stops on the known answer: a convergence check, not a speed
                     "target_criterion_check": {"seconds": round(tt, 4), "blocks": btg,
                                                "max_relative_error": et}}
        log(f"block Lanczos from the MLP start block: {tr:.4f}s, {br} blocks (residual "
            f"bound), {tt:.4f}s, {btg} blocks (target criterion)")

    fast = block_solve if fam is not None else (lambda: single(op32, 2))
    for _ in range(args.warmup):
        timed(fast)
    times, last = [], None
    for _ in range(args.steps):
        dt, last = timed(fast)
        times.append(dt)
    t_fast = float(np.mean(times))
    errs = rel(last.eigenvalues)
    converged = bool(np.all(errs <= args.tolerance))
    # |H y - theta y| / |theta| of the fast path's Ritz pairs, from its recurrence
    residuals = ([float(r / abs(t)) for r, t in zip(last.residuals, last.eigenvalues)]
                 if fam is not None else None)
    steps_taken = last.blocks if fam is not None else last.iterations
    log(f"fast path: {t_fast:.4f}s, {steps_taken} {'blocks' if fam is not None else 'its'}, "
        f"errors {errs}")

    # --- the fast path at held-out points of the family ---------------------
    heldout = None
    if fam is not None and args.heldout > 0:
        rows = []
        for c1, c2 in L.family_points(args.heldout, 7):
            op64.set_scales(1.0, c1, c2)
            ref_h = bacchus.diagonalize(op64, args.iterations, num_eigenvalues=args.eigenvalues,
                                        start=bacchus.START_RANDOM, seed=1,
                                        reorthogonalization=bacchus.REORTH_FULL,
                                        convergence=bacchus.CONVERGE_NONE).eigenvalues
            bt.set_scales(0, np.tile((1.0, c1, c2), (args.block, 1)))

            def solve_at(c1=c1, c2=c2):
                X = fam.start_block(c1, c2, args.block)
                return bt.block_lanczos(args.max_blocks, args.eigenvalues, start_block=X,
                                        convergence=bacchus.CONVERGE_RESIDUALS,
                                        tolerance=args.tolerance)
            timed(solve_at)
            dt, r = timed_median(solve_at)
            err = float((np.abs(r.eigenvalues - ref_h) / np.abs(ref_h)).max())
            row = [dt, r.blocks, err]
            if mblock is not None:
                for m in mlp_block_runs(c1, c2, ref_h):
                    row += list(m)
            rows.append(row)
            log(f"  held-out ({c1:.4f}, {c2:.4f}): {dt * 1e3:.1f} ms, {r.blocks} blocks, "
                f"max error {err:.2e}" + (f"; MLP block {row[3] * 1e3:.1f} ms, {row[4]} blocks"
                                          if mblock is not None else ""))
        op64.set_scales(1.0, 1.0, 1.0)
        bt.set_scales(0, np.ones((args.block, 3)))
        a = np.array(rows)
        heldout = {"points": len(rows), "seed": 7, "mean_seconds": round(float(a[:, 0].mean()), 4),
                   "max_seconds": round(float(a[:, 0].max()), 4),
                   "max_blocks": int(a[:, 1].max()),
                   "max_relative_error": float(a[:, 2].max()),
                   "all_within_tolerance": bool((a[:, 2] <= args.tolerance).all())}
        if mblock is not None:
            heldout["mlp_block_start"] = {
                "mean_seconds": round(float(a[:, 3].mean()), 4),
                "max_seconds": round(float(a[:, 3].max()), 4),
                "max_blocks": int(a[:, 4].max()),
                "max_relative_error": float(a[:, 5].max()),
                "converged": int(a[:, 6].sum()),
                "target_criterion_mean_seconds": round(float(a[:, 7].mean()), 4),
                "target_criterion_max_blocks": int(a[:, 8].max()),
                "target_criterion_max_relative_error": float(a[:, 9].max())}
        log(f"held-out: {heldout}")

    # --- CPU oracle on the host cores: the whole FP64 solve -------------
    cpu = None
    if not args.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0))
        c = cpu_solve(inputs, sv_host, r64.iterations, cores, args.eigenvalues)
        cpu = {"value": round(c["seconds"], 3), "unit": "s", "cores": cores, "kind": "port",
               "sample": f"the oracle's FP64 Lanczos (oracle/oracle.c, OpenMP matvec with "
                         f"per-thread replicas, full reorthogonalization) of the workload "
                         f"from the same start for the {r64.iterations} iterations the FP64 "
                         f"GPU path took to the same criterion; lowest "
                         f"{args.eigenvalues} relative errors "
                         f"{max(np.abs(c['eigenvalues'] - targets) / np.abs(targets)):.1e}",
               "fp64_gpu_seconds": round(t64, 4),
               "speedup_fp64_gpu_vs_cpu": round(c["seconds"] / t64, 1)}
        log(f"cpu oracle solve {c['seconds']:.2f}s on {cores} threads")

    # --- the other configurations --------------------------------------
    configs = {}
    if not args.no_configs:
        configs["c1_c2"] = small_systems(args, cores=len(os.sched_getaffinity(0)),
                                         cpu=not args.no_cpu_baseline)
        configs["c4_batched_sweep"] = batched_sweep(args, emit=False)
        configs["c5_learned_start"] = learned_start(args, emit=False)

    if fam is not None:
        fast_desc = (f"fp32 elements; block Lanczos ({args.block} vectors per batched launch, "
                     "FP64 inner products and block tridiagonal solve, stops on its residual "
                     "bound; P-form products on the matrix cores); start block = Ritz vectors "
                     f"of H(c1,c2) in a stored basis of {args.projected_basis} band Ritz vectors "
                     f"(rank {fam.rank}) from {args.train_points} batched family solves")
    else:
        fast_desc = "fp32 elements + fp64 single-vector lanczos, full reorth, random start"
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
                   "fast_path": fast_desc},
        "fp64_gpu_seconds": round(t64, 4),
        "fp64_gpu_iterations": r64.iterations,
        "fast_steps": steps_taken,
        "fast_matvecs": steps_taken * (args.block if fam is not None else 1),
        "speedup_vs_fp64_gpu": round(t64 / t_fast, 3),
        "converged": converged,
        "training_seconds": round(train_s, 2),
        "mlp_start": mlp,
        "mlp_block_start": mlp_block,
        "fast_path_heldout": heldout,
        "eigenvalues_fp64_reference": [float(v) for v in targets],
        "relative_errors_fast": [float(v) for v in errs],
        "relative_residuals_fast": residuals,
        "relative_errors_fp64": [float(v) for v in rel(r64.eigenvalues)],
        "relative_errors_by_storage": by_precision,
        "reference": reference,
        "setup_seconds": round(train_s, 2),
        "amortized_seconds_per_solve": (
            {str(n): round(t_fast + train_s / n, 4) for n in (64, 512, 4096)}
            if fam is not None else None),
        "input_generation_seconds": {"device": round(gen_device, 3),
                                     "host": round(gen_host, 3) if gen_host else None},
        "matvec_fp64_ms": round(mv_ms, 4),
        "matvec_fp32_ms": round(mv32_ms, 4),
        "block_matvec": block_mv,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(args.workload)},
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(line), flush=True)


FIXTURE_SEED = 20261204
# the learned-start networks: BF16 weights, activations carried as two BF16
# parts (bacchus_hip_mlp_set_precision)
MLP_PRECISION = 1


def start_vector(dimension, seed):
    from tests.golden.fixture_start import start_vector as sv
    return sv(dimension, int(seed))


def fixture(workload):
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", f"lanczos300_{workload}.npz")
    return np.load(path) if os.path.exists(path) else None


def cpu_solve(inputs, start, iterations, cores, num_eigenvalues):
    """The oracle's FP64 Lanczos (full reorthogonalization) from `start`
    for `iterations` iterations on `cores` host threads: the CPU baseline."""
    from oracle import oracle as O
    prob = O.Problem.from_inputs(inputs)  # views into `inputs`, kept alive
    t = time.perf_counter()
    ev, _, _, k = O.lanczos(inputs.dimension, lambda v: O.matvec(prob, v, threads=cores),
                            iterations, start=O.START_GIVEN, start_vector=start,
                            reorth=O.REORTH_FULL, num_eigenvalues=num_eigenvalues)
    return {"seconds": time.perf_counter() - t, "eigenvalues": ev, "iterations": k}


def small_systems(args, cores, cpu=True):
    """C1 (2+2 Nmax 4: 200 iterations, lowest 5, FP64 GPU against the CPU
    oracle) and C2 (2+2 Nmax 8: FP64, FP32 and BF16 element storage through
    the same 300-iteration solve, every eigenvalue against the oracle's
    fixture; the FP64 GPU and CPU solves to 1e-6 of it, timed)."""
    import numpy as np
    import torch

    from jupiterncsm_amd import bacchus
    from jupiterncsm_amd.inputs import SYSTEMS, build

    out = {}
    for name, iters in (("he4_nmax4", 200), ("he4_nmax8", 300)):
        inp = build(SYSTEMS[name])
        D = inp.dimension
        fx = fixture(name)
        sv_host = start_vector(D, fx["seed"] if fx is not None else FIXTURE_SEED)
        sv = torch.as_tensor(sv_host, device="cuda")
        ops = {p: bacchus.Operator.from_inputs(inp, v) for p, v in
               (("fp64", bacchus.FP64), ("fp32", bacchus.FP32), ("bf16", bacchus.BF16))}
        full = {p: bacchus.diagonalize(o, iters, num_eigenvalues=args.eigenvalues,
                                       start=bacchus.START_GIVEN, start_vector=sv,
                                       reorthogonalization=bacchus.REORTH_FULL,
                                       convergence=bacchus.CONVERGE_NONE).eigenvalues
                for p, o in ops.items()}
        ref = fx["eigenvalues"][:args.eigenvalues] if fx is not None else full["fp64"]
        rec = {"dimension": D, "iterations": iters,
               "reference": "oracle fixture" if fx is not None else "FP64 GPU",
               "relative_errors_by_storage": {
                   p: [float(v) for v in np.abs(e - ref) / np.abs(ref)] for p, e in full.items()}}
        # FP64 GPU and CPU solves to the 1e-6 target criterion
        kw = dict(num_eigenvalues=args.eigenvalues, reorthogonalization=bacchus.REORTH_FULL,
                  convergence=bacchus.CONVERGE_TARGETS, tolerance=args.tolerance,
                  target_eigenvalues=ref, check_every=1, start=bacchus.START_GIVEN,
                  start_vector=sv)
        bacchus.diagonalize(ops["fp64"], iters, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = bacchus.diagonalize(ops["fp64"], iters, **kw)
        torch.cuda.synchronize()
        rec["fp64_gpu_seconds"] = round(time.perf_counter() - t, 4)
        rec["fp64_gpu_iterations"] = r.iterations
        if cpu:
            c = cpu_solve(inp, sv_host, r.iterations, cores, args.eigenvalues)
            rec["cpu_seconds"] = round(c["seconds"], 4)
            rec["cpu_cores"] = cores
            rec["cpu_relative_errors"] = [float(v) for v in np.abs(c["eigenvalues"] - ref)
                                          / np.abs(ref)]
        for o in ops.values():
            o.close()
        out[name] = rec
        log(f"{name}: {rec}")
    return out


def bf16_values(a):
    """FP64 array -> the values of its BF16 rounding (round to nearest even
    of the FP32 value), as the library stores BF16 elements."""
    import numpy as np
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    u = np.where(finite, u + 0x7FFF + ((u >> 16) & 1), u)
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def mixed_operator(inputs, bf16_forces):
    """An FP32 operator whose elements of the given force classes are BF16
    values (the numerics of a store with those classes in BF16)."""
    import numpy as np
    from jupiterncsm_amd import bacchus
    elems = np.array(inputs.matrix_elements, dtype=np.float64, copy=True)
    offs = inputs.matrix_offsets
    for k, force in enumerate(inputs.matrix_shapes[:, 2]):
        if int(force) in bf16_forces:
            elems[offs[k]:offs[k + 1]] = bf16_values(elems[offs[k]:offs[k + 1]])
    ins = inputs.instructions
    raw = np.stack([ins[f] for f in ins.dtype.names], axis=1).astype(np.int32)
    op = bacchus.Operator.from_arrays(inputs.blocks[:, 4], inputs.blocks[:, 5],
                                      inputs.list_offsets, inputs.list_entries, offs,
                                      inputs.matrix_shapes[:, 0], inputs.matrix_shapes[:, 2],
                                      elems, raw, bacchus.FP32)
    del elems
    return op


def batch_without_mfma(ops, counts):
    """A batch whose P-form instructions run on the vector path (the
    matrix-core path is chosen when the batch is created)."""
    from jupiterncsm_amd import bacchus
    saved = os.environ.get("BACCHUS_HIP_BATCH_MFMA")
    os.environ["BACCHUS_HIP_BATCH_MFMA"] = "0"
    try:
        return bacchus.Batch(ops, counts)
    finally:
        if saved is None:
            del os.environ["BACCHUS_HIP_BATCH_MFMA"]
        else:
            os.environ["BACCHUS_HIP_BATCH_MFMA"] = saved


def batched_sweep(args, emit=True):
    """64 Hamiltonians H(c1, c2) spread over the 2+2 systems at Nmax 4, 6,
    8: one batched Lanczos (all matvecs of an iteration in one persistent
    launch, each iteration a replayed graph) against sequential plain FP64
    single-vector solves, same Krylov length, lowest 3 eigenvalues."""
    import numpy as np
    import torch

    from jupiterncsm_amd import bacchus
    from jupiterncsm_amd import learned as L
    from jupiterncsm_amd.inputs import SYSTEMS, build

    names = ["he4_nmax4", "he4_nmax6", "he4_nmax8"]
    inps = [build(SYSTEMS[n]) for n in names]
    pts = L.family_points(args.instances, 4)
    split = np.array_split(np.arange(args.instances), len(names))
    counts = [len(s) for s in split]
    K, nev = 120, 3
    ops32 = [bacchus.Operator.from_inputs(i, bacchus.FP32) for i in inps]
    bt = bacchus.Batch(ops32, counts)
    for k, s in enumerate(split):
        bt.set_scales(k, L.scales_of(pts[s]))
    tune = bt.autotune(3)
    st = bt.stats()
    log(f"batch: {counts} instances of D={[i.dimension for i in inps]}, {st['work_items']} "
        f"work items, autotuned {tune}")
    # batched matvec rate
    xs = [torch.randn(i.dimension, b, dtype=torch.float32, device="cuda") for i, b in
          zip(inps, counts)]
    ys = [torch.empty_like(x) for x in xs]
    for _ in range(3):
        bt.matvec(xs, ys)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        bt.matvec(xs, ys)
    e1.record()
    torch.cuda.synchronize()
    bmv = e0.elapsed_time(e1) / 20
    mfma = bt.mfma_stats()
    if args.batched_matvec_only:
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.batched_matvec_only):
            bt.matvec(xs, ys)
        e1.record()
        torch.cuda.synchronize()
        log(f"batched matvec {e0.elapsed_time(e1) / args.batched_matvec_only:.4f} ms "
            f"(P-form on the matrix cores: {mfma})")
        for o in ops32:
            o.close()
        bt.close()
        return None
    # the same batch with the P-form instructions on the vector path
    bt_v = batch_without_mfma(ops32, counts)
    for k, s in enumerate(split):
        bt_v.set_scales(k, L.scales_of(pts[s]))
    bt_v.autotune(3)
    for _ in range(3):
        bt_v.matvec(xs, ys)
    e0.record()
    for _ in range(20):
        bt_v.matvec(xs, ys)
    e1.record()
    torch.cuda.synchronize()
    bmv_vector = e0.elapsed_time(e1) / 20
    bt_v.close()
    log(f"batched matvec {bmv:.4f} ms (P-form on the matrix cores: {mfma}), "
        f"{bmv_vector:.4f} ms on the vector path")
    for _ in range(args.warmup):
        bt.lanczos(K, num_eigenvalues=nev, use_graph=True, seed=5)
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        ev, its, _ = bt.lanczos(K, num_eigenvalues=nev, use_graph=True, seed=5)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    t_batch = float(np.mean(times))
    ev_e, _, t_eager = bt.lanczos(K, num_eigenvalues=nev, use_graph=False, seed=5)
    log(f"batched: {t_batch:.3f}s for {args.instances} solves ({its} its each), eager "
        f"{t_eager:.3f}s")
    # sequential plain FP64 solves
    ops64 = [bacchus.Operator.from_inputs(i, bacchus.FP64) for i in inps]
    errs = []
    torch.cuda.synchronize()
    t = time.perf_counter()
    seq = []
    for k, s in enumerate(split):
        for b, p in enumerate(pts[s]):
            ops64[k].set_scales(1.0, p[0], p[1])
            r = bacchus.diagonalize(ops64[k], K, num_eigenvalues=nev,
                                    start=bacchus.START_RANDOM, seed=5,
                                    reorthogonalization=bacchus.REORTH_FULL,
                                    convergence=bacchus.CONVERGE_NONE)
            seq.append((k, b, r.eigenvalues))
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t
    for k, b, e in seq:
        errs.append(float(np.max(np.abs(ev[k][b] - e) / np.abs(e))))
    log(f"sequential fp64: {t_seq:.3f}s; max rel diff batched vs fp64 {max(errs):.2e}")
    line = {
        "metric": "batched sweep: 64 H(c1,c2) lowest 3 eigenvalues, solves/s",
        "value": round(args.instances / t_batch, 2), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 2),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 Hamiltonians H1 + c1 V2 + c2 V3, c ~ U[0.8,1.2], over "
                               "2+2 Nmax 4/6/8", "instances": counts,
                   "dimensions": [i.dimension for i in inps], "krylov": K, "eigenvalues": nev},
        "sequential_fp64_seconds": round(t_seq, 3),
        "sequential_solves_per_s": round(args.instances / t_seq, 2),
        "speedup_vs_sequential": round(t_seq / t_batch, 2),
        "eager_seconds": round(t_eager, 3),
        "batched_matvec_ms": round(bmv, 4),
        "batched_matvec_gflops": round(st["flops"] / bmv / 1e6, 1),
        "batched_matvec_algorithmic_gbs": round(st["algorithmic_bytes"] / bmv / 1e6, 1),
        "autotuned": tune,
        "max_rel_diff_vs_fp64_solves": max(errs),
        "mfma": {"form": "P-form instructions (P, and N in the transposed space) as 16x4 "
                         "tiles of their signed element matrix on v_mfma_f32_16x16x4f32 over "
                         "(state x vector) panels; NP instructions on the vector path",
                 **mfma, "batched_matvec_ms_vector_path": round(bmv_vector, 4)},
    }
    for o in ops32 + ops64:
        o.close()
    bt.close()
    if emit:
        print(json.dumps(line), flush=True)
    return {k: line[k] for k in ("value", "unit", "ms_per_step", "config",
                                 "sequential_fp64_seconds", "speedup_vs_sequential",
                                 "batched_matvec_ms", "batched_matvec_gflops",
                                 "max_rel_diff_vs_fp64_solves", "mfma")}


def learned_start(args, emit=True):
    """Learned start of the 3+3 Nmax 6 ground state: trained on 512 batched
    family solves, tested on 64 held-out (c1, c2) points."""
    import numpy as np
    import torch

    from jupiterncsm_amd import bacchus
    from jupiterncsm_amd import learned as L
    from jupiterncsm_amd.inputs import SYSTEMS, build

    inputs = build(SYSTEMS["li6_nmax6_3b"])
    op32 = bacchus.Operator.from_inputs(inputs, bacchus.FP32)
    train = L.family_points(args.c5_train, 1)
    test = L.family_points(args.c5_test, 2)
    t0 = time.time()
    ref = L.solve_family(op32, test, iterations=300, num_eigenvalues=1, keep_targets=False)
    log(f"held-out references: {ref.seconds:.1f}s")
    sol = L.solve_family(op32, train, iterations=60, num_eigenvalues=1, target="ground")
    log(f"training solves: {sol.seconds:.1f}s")
    model = L.fit(sol, basis_size=64, steps=6000, target="ground", precision=MLP_PRECISION)
    log(f"MLP fit: loss {model.train_loss:.2e}, total {time.time() - t0:.1f}s")
    rows = []
    fallbacks = 0
    for i, c in enumerate(test):
        torch.cuda.synchronize()
        t = time.perf_counter()
        sv = model.start_vector(*c, mix=0.01)
        fallbacks += sv is None
        rl = L.solve_from(op32, c, sv, num_eigenvalues=1, targets=ref.eigenvalues[i][:1])
        torch.cuda.synchronize()
        tl = time.perf_counter() - t
        t = time.perf_counter()
        rr = L.solve_from(op32, c, None, num_eigenvalues=1, targets=ref.eigenvalues[i][:1])
        torch.cuda.synchronize()
        tr = time.perf_counter() - t
        rows.append((rl.iterations, tl, rr.iterations, tr))
    a = np.array(rows, dtype=float)
    log(f"learned: {a[:, 0].mean():.1f} its {a[:, 1].mean() * 1e3:.1f} ms; random: "
        f"{a[:, 2].mean():.1f} its {a[:, 3].mean() * 1e3:.1f} ms")
    for i, r in enumerate(rows):
        log(f"  c=({test[i][0]:.4f},{test[i][1]:.4f}) learned {r[0]} its {r[1] * 1e3:.1f} ms, "
            f"random {r[2]} its {r[3] * 1e3:.1f} ms")
    line = {
        "metric": "learned start: Lanczos iterations to 1e-6 of the lowest eigenvalue",
        "value": round(float(a[:, 0].mean()), 2), "unit": "iterations", "n_gpus": 1,
        "steps": len(test), "warmup": 0, "ms_per_step": round(float(a[:, 1].mean()) * 1e3, 2),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "li6_nmax6_3b family H1 + c1 V2 + c2 V3, c ~ U[0.8,1.2]",
                   "train_points": len(train), "test_points": len(test), "basis": 64,
                   "mlp": "2-256-256-256-64 GELU, BF16 weights, inputs and activations as "
                          "BF16 hi+lo pairs on the MFMA, FP32 accumulate"},
        "random_start_iterations": round(float(a[:, 2].mean()), 2),
        "learned_ms": round(float(a[:, 1].mean()) * 1e3, 2),
        "random_ms": round(float(a[:, 3].mean()) * 1e3, 2),
        "iterations_ratio": round(float(a[:, 2].mean() / a[:, 0].mean()), 2),
        "time_ratio": round(float(a[:, 3].mean() / a[:, 1].mean()), 2),
        "max_learned_iterations": int(a[:, 0].max()),
        "median_learned_iterations": float(np.median(a[:, 0])),
        "median_random_iterations": float(np.median(a[:, 2])),
        "points_not_converged_learned": int((a[:, 0] >= 300).sum()),
        "points_not_converged_random": int((a[:, 2] >= 300).sum()),
        "fallbacks": int(fallbacks),
        "train_loss": model.train_loss,
        "training_seconds": round(sol.seconds, 1),
    }
    op32.close()
    if emit:
        print(json.dumps(line), flush=True)
    return {k: line[k] for k in ("value", "unit", "config", "random_start_iterations",
                                 "learned_ms", "random_ms", "iterations_ratio", "time_ratio",
                                 "max_learned_iterations", "points_not_converged_learned",
                                 "points_not_converged_random", "fallbacks",
                                 "training_seconds")}


if __name__ == "__main__":
    main()
