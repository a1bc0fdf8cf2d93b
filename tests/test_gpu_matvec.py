"""Parity of the gfx950 matvec (through the C ABI) with the CPU oracle and
the compiled reference's golden outputs."""
import os
import shutil
import tempfile

import numpy as np
import pytest

from jupiterncsm_amd import bacchus
from jupiterncsm_amd.inputs import SYSTEMS, SystemConfig, build
from oracle import oracle as O
from tests import skip_cases as SK

pytestmark = pytest.mark.gpu

CONFIGS = {
    "he4_nmax4": SystemConfig(2, 2, 4, max_body=2, seed=1),
    "he4_nmax2_3b": SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3),
    "li6_nmax2_3b": SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1),
}
# FP64 bar: the oracle's own summation order differs from the kernel's, so
# agreement is to rounding: 1e-12 of max|y|.
FP64_TOL = 1e-12


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    return torch


def gpu_matvec(op, x, accumulate=None):
    torch = torch_dev()
    xd = torch.as_tensor(x, dtype=torch.float64, device="cuda")
    yd = (torch.as_tensor(accumulate, dtype=torch.float64, device="cuda").clone()
          if accumulate is not None else torch.empty_like(xd))
    op.matvec(xd, yd, accumulate=accumulate is not None)
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_fp64_matvec_matches_compiled_reference(name, golden_dir):
    print(f"case {name}", flush=True)
    g = np.load(os.path.join(golden_dir, f"{name}.npz"))
    inp = build(CONFIGS[name])
    with bacchus.Operator.from_inputs(inp) as op:
        y = gpu_matvec(op, g["x"])
    assert rel_err(y, g["y_ref"]) <= FP64_TOL


@pytest.mark.parametrize("name,seed", [("he4_nmax4", 1), ("he4_nmax6", 2), ("he4_nmax8", 3)])
def test_fp64_matvec_matches_oracle(name, seed):
    print(f"case {name}", flush=True)
    inp = build(SYSTEMS[name])
    p = O.Problem.from_inputs(inp)
    x = np.random.default_rng(seed).standard_normal(inp.dimension)
    with bacchus.Operator.from_inputs(inp) as op:
        y = gpu_matvec(op, x)
        # accumulate: y0 + H x
        y0 = np.random.default_rng(seed + 100).standard_normal(inp.dimension)
        y_acc = gpu_matvec(op, x, accumulate=y0)
    y_o = O.matvec(p, x)
    assert rel_err(y, y_o) <= FP64_TOL
    assert rel_err(y_acc, y_o + y0) <= FP64_TOL


def test_force_scales_match_oracle():
    print("case scales li6_nmax2_3b", flush=True)
    inp = build(CONFIGS["li6_nmax2_3b"])
    p = O.Problem.from_inputs(inp)
    x = np.random.default_rng(9).standard_normal(inp.dimension)
    with bacchus.Operator.from_inputs(inp) as op:
        op.set_scales(1.0, 0.85, 1.15)
        y = gpu_matvec(op, x)
    y_o = O.matvec(p, x, scales=[1.0, 0.85, 1.15])
    assert rel_err(y, y_o) <= FP64_TOL
    y_plain = O.matvec(p, x)
    assert rel_err(y_plain, y_o) > 1e-6  # the scales did act


@pytest.mark.parametrize("precision,tol", [(bacchus.FP32, 2e-6), (bacchus.BF16, 2e-2)])
def test_reduced_precision_storage_error_bounds(precision, tol):
    """FP32/BF16-stored elements with FP32 accumulation: relative error of y
    against the FP64 oracle stays within the storage precision's bound."""
    print(f"case precision {precision}", flush=True)
    inp = build(SYSTEMS["he4_nmax6"])
    x = np.random.default_rng(4).standard_normal(inp.dimension)
    y_o = O.matvec(O.Problem.from_inputs(inp), x)
    with bacchus.Operator.from_inputs(inp, precision=precision) as op:
        y = gpu_matvec(op, x)
    err = rel_err(y, y_o)
    assert err <= tol
    assert err > 0.0  # it really is a reduced-precision path


def test_file_drop_in_matches_compiled_reference(golden_dir):
    """comb.txt + evaluation order + index_list_<id> + <id>_matrix_elements
    + vec_<id> directories in, vec_<id> out (accumulating into the existing
    output files like the reference)."""
    print("case file drop-in he4_nmax4", flush=True)
    g = np.load(os.path.join(golden_dir, "he4_nmax4.npz"))
    inp = build(CONFIGS["he4_nmax4"])
    d = tempfile.mkdtemp(prefix="bacchus_dropin_")
    try:
        inp.write(d)
        off = inp.block_offsets
        y0 = np.random.default_rng(1).standard_normal(inp.dimension)
        for name, v in (("x", g["x"]), ("y", y0)):
            os.makedirs(os.path.join(d, name))
            for b in range(len(inp.blocks)):
                v[off[b]:off[b + 1]].tofile(os.path.join(d, name, f"vec_{b + 1}"))
        with bacchus.Operator.load(os.path.join(d, "comb.txt"), 2, 2,
                                   os.path.join(d, "evaluation_order"),
                                   os.path.join(d, "index_lists"),
                                   os.path.join(d, "matrix_elements")) as op:
            bacchus.run_matrix_vector_multiplication(os.path.join(d, "y"), os.path.join(d, "x"),
                                                     op)
        y = np.concatenate([np.fromfile(os.path.join(d, "y", f"vec_{b + 1}"))
                            for b in range(len(inp.blocks))])
        assert rel_err(y - y0, g["y_ref"]) <= FP64_TOL
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_c2_size_properties():
    """At the 2+2 Nmax 8 size: linearity and symmetry (x.Hy = y.Hx)."""
    print("case c2 properties", flush=True)
    inp = build(SYSTEMS["he4_nmax8"])
    rng = np.random.default_rng(8)
    x, z = rng.standard_normal(inp.dimension), rng.standard_normal(inp.dimension)
    with bacchus.Operator.from_inputs(inp) as op:
        hx, hz = gpu_matvec(op, x), gpu_matvec(op, z)
        hxz = gpu_matvec(op, 2.0 * x - 0.5 * z)
    assert rel_err(hxz, 2.0 * hx - 0.5 * hz) <= 1e-13
    assert abs(x @ hz - z @ hx) <= 1e-12 * np.abs(hx).max() * np.abs(z).sum()


def test_c3_full_size_matches_oracle():
    """The headline workload itself (3+3 Nmax 6, 2+3-body, D = 197,822):
    the FP64 matvec against the oracle to 1e-12 of max|y| (every kernel
    path at its real shape: paired and unpaired NP lanes, narrow and wide
    P rows, both spaces), FP32 storage within 2e-6, and x.Hz = z.Hx."""
    print("case c3 full size", flush=True)
    import os
    inp = build(SYSTEMS["li6_nmax6_3b"])
    rng = np.random.default_rng(12)
    x, z = rng.standard_normal(inp.dimension), rng.standard_normal(inp.dimension)
    ref = O.matvec(O.Problem.from_inputs(inp), x,
                   threads=min(16, len(os.sched_getaffinity(0))))
    with bacchus.Operator.from_inputs(inp) as op:
        hx, hz = gpu_matvec(op, x), gpu_matvec(op, z)
    print("fp64 rel err", rel_err(hx, ref), flush=True)
    assert rel_err(hx, ref) <= FP64_TOL
    assert abs(x @ hz - z @ hx) <= 1e-12 * np.abs(hx).max() * np.abs(z).sum()
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op32:
        h32 = gpu_matvec(op32, x)
    print("fp32 rel err", rel_err(h32, ref), flush=True)
    assert rel_err(h32, ref) <= 2e-6


@pytest.mark.parametrize("env", [{"BACCHUS_HIP_IN_GROUPS": "1"}, {"BACCHUS_HIP_XCDS": "1"},
                                 {"BACCHUS_HIP_TILE_ORDER": "0"}, {"BACCHUS_HIP_EXPAND": "0"},
                                 {"BACCHUS_HIP_EXPAND": "1000:1000"},
                                 {"BACCHUS_HIP_TILE_ELEMS": "8192", "BACCHUS_HIP_TILE_MULTS": "4000"}])
def test_plan_switches_keep_results(env):
    """The single-vector plan's switches (input groups, XCD grouping, tile
    order, which lane lists get expanded element rows, tile size and work)
    change the schedule and layout, not the results: FP64 against the oracle
    to 1e-12 of max|y| on a 3-body system."""
    print(f"case plan switches {env}", flush=True)
    inp = build(CONFIGS["li6_nmax2_3b"])
    x = np.random.default_rng(23).standard_normal(inp.dimension)
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        op = bacchus.Operator.from_inputs(inp)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    with op:
        y = gpu_matvec(op, x)
    assert rel_err(y, O.matvec(O.Problem.from_inputs(inp), x)) <= FP64_TOL


@pytest.mark.parametrize("name", ["he4_nmax6", "li6_nmax2_3b"])
def test_element_row_major_np_kernel_matches_oracle(name):
    """The element-row-major NP kernel (np.hip, selected by BACCHUS_HIP_NP=1
    when the operator is built; off by default, where the tile kernel runs
    the NP instructions): FP64 against the oracle to 1e-12 of max|y| with
    force scales other than 1, FP32 storage within 2e-6."""
    print(f"case element-row-major np {name}", flush=True)
    inp = build(SYSTEMS[name] if name in SYSTEMS else CONFIGS[name])
    p = O.Problem.from_inputs(inp)
    x = np.random.default_rng(21).standard_normal(inp.dimension)
    sc = [1.0, 0.9, 1.1]
    saved = os.environ.get("BACCHUS_HIP_NP")
    os.environ["BACCHUS_HIP_NP"] = "1"
    try:
        op = bacchus.Operator.from_inputs(inp)
        op32 = bacchus.Operator.from_inputs(inp, precision=bacchus.FP32)
    finally:
        if saved is None:
            del os.environ["BACCHUS_HIP_NP"]
        else:
            os.environ["BACCHUS_HIP_NP"] = saved
    with op, op32:
        op.set_scales(*sc)
        op32.set_scales(*sc)
        y, y32 = gpu_matvec(op, x), gpu_matvec(op32, x)
    y_o = O.matvec(p, x, scales=sc)
    assert rel_err(y, y_o) <= FP64_TOL
    assert rel_err(y32, y_o) <= 2e-6


@pytest.mark.parametrize("scale", SK.SCALES)
@pytest.mark.parametrize("case", ["N", "P narrow", "P wide", "NP"])
def test_tiny_element_skip(case, scale):
    """N and P instructions skip elements with |s V| < 1e-12 (s the force
    scale: the reference tests the Vulcan-combined element it reads,
    matrix_vector_multiplication.c:50-53, :105-108), NP instructions skip
    nothing (:134-210); exactly 1e-12 counts.  Elements straddle 1e-12 / s
    and x is of order 1e12, so each skipped product is 0.5 or more, far
    above FP64 rounding: the device result must equal the products computed
    by hand, and the oracle's."""
    print("case tiny element skip", case, "scale", scale, flush=True)
    arrays, np_, nn = SK.problem(case.split()[0], wide_p=case == "P wide")
    rng = np.random.default_rng(5)
    X = 1e12 * (1.0 + rng.random((np_, nn)))
    expect = SK.expected(case.split()[0], X, scale)
    with bacchus.Operator.from_arrays(*arrays) as op:
        op.set_scales(1.0, scale, 1.0)
        y = gpu_matvec(op, X.ravel())
    ref = O.matvec(O.Problem(*arrays), X.ravel(), scales=[1.0, scale, 1.0])
    print("device", y, "expected", expect, flush=True)
    assert rel_err(y, expect) <= FP64_TOL
    assert rel_err(ref, expect) <= FP64_TOL


def test_empty_instruction_list_gives_zero():
    print("case empty instructions", flush=True)
    inp = build(CONFIGS["he4_nmax2_3b"])
    torch = torch_dev()
    with bacchus.Operator.from_arrays(inp.blocks[:, 4], inp.blocks[:, 5], [0], np.zeros(0),
                                      [0], np.zeros(0), np.zeros(0), np.zeros(0),
                                      np.zeros((0, 8))) as op:
        x = torch.ones(inp.dimension, dtype=torch.float64, device="cuda")
        y = torch.full_like(x, 7.0)
        op.matvec(x, y)
        torch.cuda.synchronize()
        assert float(y.abs().max()) == 0.0
