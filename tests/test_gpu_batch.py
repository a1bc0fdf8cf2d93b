"""Parity of the batched form (several operators, several vectors each, own
scales per vector, one persistent launch) with the CPU oracle, and of the
batched Lanczos with dense eigenvalues."""
import numpy as np
import pytest

from jupiterncsm_amd import bacchus
from jupiterncsm_amd.inputs import SYSTEMS, SystemConfig, build
from oracle import oracle as O
from tests import skip_cases as SK

pytestmark = pytest.mark.gpu

SMALL_3B = SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1)


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    return torch


def scales_for(B, seed):
    rng = np.random.default_rng(seed)
    s = np.ones((B, 3))
    s[:, 1:] = rng.uniform(0.8, 1.2, size=(B, 2))
    return s


def batch_vs_oracle(systems, counts, precision, seed=0, **tuning):
    torch = torch_dev()
    inps = [build(s) for s in systems]
    ops = [bacchus.Operator.from_inputs(i, precision=precision) for i in inps]
    errs = []
    try:
        with bacchus.Batch(ops, counts, **tuning) as bt:
            xs, ys, refs = [], [], []
            for k, (inp, B) in enumerate(zip(inps, counts)):
                sc = scales_for(B, seed + k)
                bt.set_scales(k, sc)
                x = np.random.default_rng(seed + 10 + k).standard_normal((inp.dimension, B))
                p = O.Problem.from_inputs(inp)
                refs.append(np.stack([O.matvec(p, x[:, b], scales=sc[b]) for b in range(B)], 1))
                xs.append(torch.as_tensor(x, dtype=torch.float32, device="cuda").contiguous())
                ys.append(torch.full((inp.dimension, B), 3.0, dtype=torch.float32, device="cuda"))
            bt.matvec(xs, ys)
            torch.cuda.synchronize()
            for y, r in zip(ys, refs):
                yh = y.cpu().numpy().astype(np.float64)
                errs.append(np.abs(yh - r).max() / np.abs(r).max())
    finally:
        for o in ops:
            o.close()
    return errs


@pytest.mark.parametrize("precision,tol", [(bacchus.FP32, 2e-6), (bacchus.BF16, 2e-2)])
def test_batched_matvec_uneven_operators_match_oracle(precision, tol):
    """Three operators of different dimension (2-body and 3-body), vector
    counts that are not powers of two, per-vector scales."""
    print(f"case batch precision {precision}", flush=True)
    errs = batch_vs_oracle([SYSTEMS["he4_nmax4"], SYSTEMS["he4_nmax6"], SMALL_3B], [3, 5, 2],
                           precision, seed=3)
    print("errors", errs, flush=True)
    assert max(errs) <= tol
    assert min(errs) > 0.0


@pytest.mark.parametrize("env", [{"BACCHUS_HIP_BATCH_GROUPS": "16"},
                                 {"BACCHUS_HIP_BATCH_ORDER": "1"},
                                 {"BACCHUS_HIP_BATCH_GROUPS": "1", "BACCHUS_HIP_BATCH_MFMA": "0"}])
def test_batched_queue_switches_keep_results(env):
    """The batched plan's switches (input groups of the work queues, the
    order of a queue's items, the matrix-core path off) change the schedule,
    not the results: FP32 against the oracle within 2e-6."""
    import os
    print(f"case batch switches {env}", flush=True)
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        errs = batch_vs_oracle([SYSTEMS["he4_nmax4"], SYSTEMS["he4_nmax6"], SMALL_3B], [3, 5, 2],
                               bacchus.FP32, seed=4)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    print("errors", errs, flush=True)
    assert max(errs) <= 2e-6


@pytest.mark.parametrize("counts,vectors", [([1, 1], 0), ([64, 17], 0), ([33, 7], 16)])
def test_batched_matvec_vector_widths(counts, vectors):
    """One vector, a full 64-wide tile and a tile width cap."""
    print(f"case widths {counts} {vectors}", flush=True)
    errs = batch_vs_oracle([SYSTEMS["he4_nmax4"], SMALL_3B], counts, bacchus.FP32, seed=5,
                           max_tile_vectors=vectors)
    assert max(errs) <= 2e-6


@pytest.mark.parametrize("precision,tol", [(bacchus.FP32, 2e-6), (bacchus.BF16, 2e-2)])
def test_batched_matvec_four_waves_per_simd(precision, tol):
    """The kernel variant held to 128 VGPRs (waves_per_simd = 4, one of the
    autotuner's choices) on the uneven three-operator batch."""
    print(f"case batch four waves {precision}", flush=True)
    errs = batch_vs_oracle([SYSTEMS["he4_nmax4"], SYSTEMS["he4_nmax6"], SMALL_3B], [3, 5, 2],
                           precision, seed=9, waves_per_simd=4)
    print("errors", errs, flush=True)
    assert max(errs) <= tol


def test_batched_matvec_c2_size():
    """2+2 Nmax 8 at 16 vectors against the FP64 single-vector GPU path
    (pinned to the oracle elsewhere): linear in the scales."""
    print("case c2 batch", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax8"])
    B = 16
    sc = scales_for(B, 7)
    x = np.random.default_rng(8).standard_normal((inp.dimension, B))
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op32, \
            bacchus.Operator.from_inputs(inp) as op64:
        with bacchus.Batch([op32], [B]) as bt:
            bt.set_scales(0, sc)
            xd = torch.as_tensor(x, dtype=torch.float32, device="cuda").contiguous()
            yd = torch.empty_like(xd)
            bt.matvec([xd], [yd])
            torch.cuda.synchronize()
            y = yd.cpu().numpy().astype(np.float64)
        for b in (0, 5, 15):
            op64.set_scales(*sc[b])
            xb = torch.as_tensor(x[:, b].copy(), dtype=torch.float64, device="cuda")
            yb = torch.empty_like(xb)
            op64.matvec(xb, yb)
            torch.cuda.synchronize()
            r = yb.cpu().numpy()
            assert np.abs(y[:, b] - r).max() / np.abs(r).max() <= 2e-6


def dense_h(inp, scales):
    p = O.Problem.from_inputs(inp)
    D = inp.dimension
    H = np.zeros((D, D))
    for j in range(D):
        e = np.zeros(D)
        e[j] = 1.0
        H[:, j] = O.matvec(p, e, scales=scales, threads=8)
    return 0.5 * (H + H.T)


def test_batched_lanczos_matches_dense_eigenvalues():
    """Every instance's lowest 3 eigenvalues against numpy's dense eigh of
    its own H(c1, c2) (2+2 Nmax 4, D = 952), graph and eager runs equal."""
    print("case batch lanczos", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax4"])
    B = 4
    sc = scales_for(B, 11)
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op, \
            bacchus.Batch([op], [B]) as bt:
        bt.set_scales(0, sc)
        ritz = torch.zeros((inp.dimension, B), dtype=torch.float64, device="cuda")
        ev_g, k_g, _ = bt.lanczos(150, num_eigenvalues=3, use_graph=True, seed=3,
                                  ritz_vectors=[ritz])
        ev_e, k_e, _ = bt.lanczos(150, num_eigenvalues=3, use_graph=False, seed=3)
        torch.cuda.synchronize()
        psi = ritz.cpu().numpy()
    assert k_g == k_e == 150
    for b in range(B):
        H = dense_h(inp, sc[b])
        w = np.linalg.eigvalsh(H)[:3]
        err = np.abs(ev_g[0][b] - w) / np.abs(w)
        print("instance", b, ev_g[0][b], w, err, flush=True)
        assert err.max() <= 1e-6
        # FP32 LDS atomics sum in a run-dependent order: the two runs agree
        # to the FP32 matvec's rounding, not bitwise
        assert np.abs(ev_g[0][b] - ev_e[0][b]).max() <= 1e-6 * np.abs(w).max()
        v = psi[:, b]
        assert abs(np.linalg.norm(v) - 1.0) <= 1e-6
        res = np.linalg.norm(H @ v - ev_g[0][b][0] * v)
        assert res <= 1e-3 * abs(w[0])


def test_batched_lanczos_uneven_batch_converges():
    """Two systems of different dimension in one batch, stopping on the
    eigenvalue-change criterion; every instance against its own
    single-vector FP64 solve."""
    print("case batch lanczos uneven", flush=True)
    inps = [build(SYSTEMS["he4_nmax4"]), build(SMALL_3B)]
    counts = [3, 2]
    ops = [bacchus.Operator.from_inputs(i, precision=bacchus.FP32) for i in inps]
    try:
        with bacchus.Batch(ops, counts) as bt:
            scs = [scales_for(B, 20 + k) for k, B in enumerate(counts)]
            for k, sc in enumerate(scs):
                bt.set_scales(k, sc)
            ev, iters, _ = bt.lanczos(200, num_eigenvalues=3, tolerance=1e-7, check_every=5,
                                      seed=9)
        assert iters < 200
        for k, inp in enumerate(inps):
            with bacchus.Operator.from_inputs(inp) as op64:
                for b in range(counts[k]):
                    op64.set_scales(*scs[k][b])
                    r = bacchus.diagonalize(op64, min(200, inp.dimension), num_eigenvalues=3,
                                            start=bacchus.START_RANDOM,
                                            reorthogonalization=bacchus.REORTH_FULL,
                                            convergence=bacchus.CONVERGE_NONE, seed=1)
                    err = np.abs(ev[k][b] - r.eigenvalues) / np.abs(r.eigenvalues)
                    print(k, b, ev[k][b], r.eigenvalues, err, flush=True)
                    assert err.max() <= 1e-6
    finally:
        for o in ops:
            o.close()


def test_batched_lanczos_operator_streams_keep_results():
    """Three operators: their vector work on streams of their own (the
    default) or one after the other on one stream
    (BACCHUS_HIP_LANCZOS_STREAMS=0), captured or eager, gives the same
    eigenvalues to the FP32 matvec's rounding."""
    import os
    print("case batch lanczos operator streams", flush=True)
    inps = [build(SYSTEMS["he4_nmax4"]), build(SYSTEMS["he4_nmax6"]), build(SMALL_3B)]
    counts = [3, 2, 2]
    ops = [bacchus.Operator.from_inputs(i, precision=bacchus.FP32) for i in inps]
    saved = os.environ.get("BACCHUS_HIP_LANCZOS_STREAMS")
    runs = {}
    try:
        with bacchus.Batch(ops, counts) as bt:
            for k, B in enumerate(counts):
                bt.set_scales(k, scales_for(B, 40 + k))
            for flag in ("1", "0"):
                os.environ["BACCHUS_HIP_LANCZOS_STREAMS"] = flag
                for graph in (True, False):
                    ev, iters, _ = bt.lanczos(80, num_eigenvalues=3, use_graph=graph, seed=6)
                    assert iters == 80
                    runs[(flag, graph)] = ev
    finally:
        if saved is None:
            os.environ.pop("BACCHUS_HIP_LANCZOS_STREAMS", None)
        else:
            os.environ["BACCHUS_HIP_LANCZOS_STREAMS"] = saved
        for o in ops:
            o.close()
    ref = runs[("0", True)]
    for key, ev in runs.items():
        for k in range(len(counts)):
            err = np.abs(ev[k] - ref[k]) / np.abs(ref[k])
            print(key, k, err.max(), flush=True)
            assert err.max() <= 1e-6


def test_autotune_keeps_results():
    print("case autotune", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax6"])
    B = 8
    x = np.random.default_rng(2).standard_normal((inp.dimension, B))
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op, \
            bacchus.Batch([op], [B]) as bt:
        xd = torch.as_tensor(x, dtype=torch.float32, device="cuda").contiguous()
        y0, y1 = torch.empty_like(xd), torch.empty_like(xd)
        bt.matvec([xd], [y0])
        best = bt.autotune(repetitions=2)
        bt.matvec([xd], [y1])
        torch.cuda.synchronize()
    print(best, flush=True)
    assert best["ms"] > 0 and best["tiles_per_cu"] > 0
    a, b = y0.cpu().numpy(), y1.cpu().numpy()
    assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


def test_batch_rejects_fp64_operator():
    inp = build(SYSTEMS["he4_nmax4"])
    with bacchus.Operator.from_inputs(inp) as op:
        with pytest.raises(bacchus.BacchusError, match="FP64"):
            bacchus.Batch([op], [2])


@pytest.mark.parametrize("case", ["N", "P narrow", "P wide", "NP"])
def test_batched_tiny_element_skip(case):
    """The batched kernel skips N and P elements per vector when |s V| <
    1e-12, s that vector's 2-body scale (matrix_vector_multiplication.c:
    50-53, :105-108 on the Vulcan-combined element), and nothing for NP, on
    FP32-stored elements: 4 vectors with scales 1, 0.8, 1.25, 1 and x of
    order 1e12, so each skipped product is 0.5 or more; against the
    products written by hand and the oracle in FP64 (FP32 bar 2e-6)."""
    print("case batched tiny element skip", case, flush=True)
    torch = torch_dev()
    kind = case.split()[0]
    arrays, np_, nn = SK.problem(kind, wide_p=case == "P wide")
    D = np_ * nn
    scales = [1.0, 0.8, 1.25, 1.0]
    B = len(scales)
    X = 1e12 * (1.0 + np.random.default_rng(6).random((D, B)))
    prob = O.Problem(*arrays)
    ref = np.stack([O.matvec(prob, X[:, b], scales=[1.0, scales[b], 1.0]) for b in range(B)],
                   axis=1)
    want = np.stack([SK.expected(kind, X[:, b].reshape(np_, nn), scales[b]) for b in range(B)],
                    axis=1)
    assert np.abs(ref - want).max() <= 1e-12 * np.abs(want).max()
    with bacchus.Operator.from_arrays(*arrays, precision=bacchus.FP32) as op, \
            bacchus.Batch([op], [B]) as bt:
        bt.set_scales(0, np.array([[1.0, s, 1.0] for s in scales]))
        xs = [torch.as_tensor(X, dtype=torch.float32, device="cuda").contiguous()]
        ys = [torch.zeros_like(xs[0])]
        bt.matvec(xs, ys)
        torch.cuda.synchronize()
        Y = ys[0].cpu().numpy().astype(np.float64)
    for b in range(B):
        err = np.abs(Y[:, b] - want[:, b]).max() / np.abs(want[:, b]).max()
        print("batched vs hand-computed, vector", b, err, flush=True)
        assert err <= 2e-6


def batched_y(inps, ops, counts, scales, xs_host, mfma):
    """y of one batched matvec, the batch created with the matrix-core path
    of the P-form instructions on or off; returns (ys, mfma_stats)."""
    import os
    torch = torch_dev()
    saved = os.environ.get("BACCHUS_HIP_BATCH_MFMA")
    os.environ["BACCHUS_HIP_BATCH_MFMA"] = "1" if mfma else "0"
    try:
        bt = bacchus.Batch(ops, counts)
    finally:
        if saved is None:
            del os.environ["BACCHUS_HIP_BATCH_MFMA"]
        else:
            os.environ["BACCHUS_HIP_BATCH_MFMA"] = saved
    with bt:
        for k, sc in enumerate(scales):
            bt.set_scales(k, sc)
        xs = [torch.as_tensor(x, dtype=torch.float32, device="cuda").contiguous() for x in xs_host]
        ys = [torch.full_like(x, 7.0) for x in xs]
        bt.matvec(xs, ys)
        torch.cuda.synchronize()
        stats = bt.mfma_stats()
        return [y.cpu().numpy().astype(np.float64) for y in ys], stats


@pytest.mark.parametrize("precision,tol", [(bacchus.FP32, 2e-6), (bacchus.BF16, 2e-2)])
def test_batched_mfma_path_matches_oracle_and_vector_path(precision, tol):
    """The P-form instructions (P, and N in the transposed space) on
    v_mfma_f32_16x16x4f32: three operators of uneven dimension (2-body and
    3-body), vector counts 3, 5, 2 (padded to 4, 8, 4), per-vector scales;
    against the FP64 oracle and against the same batch with the matrix-core
    path off (both FP32 products and sums, so within 1e-5)."""
    print(f"case batch mfma precision {precision}", flush=True)
    systems = [SYSTEMS["he4_nmax4"], SYSTEMS["he4_nmax6"], SMALL_3B]
    counts = [3, 5, 2]
    inps = [build(s) for s in systems]
    scales = [scales_for(B, 20 + k) for k, B in enumerate(counts)]
    xs = [np.random.default_rng(30 + k).standard_normal((i.dimension, B))
          for k, (i, B) in enumerate(zip(inps, counts))]
    ops = [bacchus.Operator.from_inputs(i, precision=precision) for i in inps]
    try:
        y_on, st_on = batched_y(inps, ops, counts, scales, xs, True)
        y_off, st_off = batched_y(inps, ops, counts, scales, xs, False)
    finally:
        for o in ops:
            o.close()
    print("mfma stats", st_on, st_off, flush=True)
    assert st_on["operators_on"] == 3 and st_on["kgroups"] > 0
    assert 0.0 < st_on["fill"] <= 1.0 and st_on["mfma_per_matvec"] >= st_on["kgroups"]
    assert st_off["operators_on"] == 0 and st_off["kgroups"] == 0
    for k, (inp, B) in enumerate(zip(inps, counts)):
        p = O.Problem.from_inputs(inp)
        ref = np.stack([O.matvec(p, xs[k][:, b], scales=scales[k][b]) for b in range(B)], 1)
        e_on = np.abs(y_on[k] - ref).max() / np.abs(ref).max()
        e_ab = np.abs(y_on[k] - y_off[k]).max() / np.abs(y_off[k]).max()
        print(f"operator {k}: vs oracle {e_on:.2e}, vs vector path {e_ab:.2e}", flush=True)
        assert e_on <= tol
        assert e_ab <= 1e-5


def test_batched_mfma_path_off_for_small_scales():
    """A scale that is nonzero and below 1e-3 in magnitude turns the
    matrix-core path off for that operator (the 1e-12 skip on s V could then
    act on a product), and the results stay those of the oracle; normal
    scales turn it on again."""
    print("case batch mfma scale gate", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax6"])
    B = 4
    x = np.random.default_rng(41).standard_normal((inp.dimension, B))
    p = O.Problem.from_inputs(inp)
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op, \
            bacchus.Batch([op], [B]) as bt:
        xd = torch.as_tensor(x, dtype=torch.float32, device="cuda").contiguous()
        for tiny, on in ((5e-4, 0), (1.0, 1), (0.0, 1)):
            sc = scales_for(B, 42)
            sc[2, 0] = tiny
            bt.set_scales(0, sc)
            assert bt.mfma_stats()["operators_on"] == on
            yd = torch.zeros_like(xd)
            bt.matvec([xd], [yd])
            torch.cuda.synchronize()
            y = yd.cpu().numpy().astype(np.float64)
            ref = np.stack([O.matvec(p, x[:, b], scales=sc[b]) for b in range(B)], 1)
            err = np.abs(y - ref).max() / np.abs(ref).max()
            print(f"s = {tiny}: path on {on}, error {err:.2e}", flush=True)
            assert err <= 2e-6
