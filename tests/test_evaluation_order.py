"""The evaluation-order generator (Mars's job, src/Mars/programs/mars.c)
against the orders the compiled reference Mars wrote for the same comb.txt
(tests/golden/*.mars_order), and, on the GPU, an operator loaded with such
an order against the oracle."""
import os

import numpy as np
import pytest

from jupiterncsm_amd import bacchus
from jupiterncsm_amd.inputs import SystemConfig, build

CASES = {
    "he4_nmax4": SystemConfig(2, 2, 4, max_body=2, seed=1),
    "he4_nmax2_3b": SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3),
    "li6_nmax2_3b": SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("two_nf", [False, True])
def test_order_matches_compiled_mars(name, two_nf, golden_dir, tmp_path):
    fixture = os.path.join(golden_dir, f"{name}.mars_order" + ("_2nf" if two_nf else ""))
    if not os.path.exists(fixture):
        pytest.skip("no 3-body blocks: the reference Mars has no two-nucleon-only order here")
    inp = build(CASES[name])
    inp.write(str(tmp_path))
    out = tmp_path / "order"
    bacchus.write_evaluation_order(tmp_path / "comb.txt", out, two_nucleon_forces_only=two_nf)
    assert out.read_text() == open(fixture).read()


def test_order_drops_one_body_blocks(tmp_path):
    inp = build(CASES["he4_nmax4"])
    inp.write(str(tmp_path))
    bacchus.write_evaluation_order(tmp_path / "comb.txt", tmp_path / "order")
    rows = (tmp_path / "order").read_text().splitlines()
    full = (tmp_path / "evaluation_order").read_text().splitlines()
    assert 0 < len(rows) < len(full)
    assert all(r.startswith("BLOCK: ") for r in rows)


def test_missing_comb_file_is_an_error(tmp_path):
    with pytest.raises(bacchus.BacchusError, match="Could not open"):
        bacchus.write_evaluation_order(tmp_path / "nope.txt", tmp_path / "order")


@pytest.mark.gpu
@pytest.mark.parametrize("two_nf", [False, True])
def test_operator_from_mars_order_matches_oracle(two_nf, tmp_path):
    """Mars drops the 1-body blocks (and the 3-body ones in two-nucleon
    mode): the operator it drives equals the oracle's H with those force
    classes scaled to zero."""
    print(f"case mars order li6_nmax2_3b two_nf={two_nf}", flush=True)
    import torch
    from oracle import oracle as O
    cfg = CASES["li6_nmax2_3b"]
    inp = build(cfg)
    inp.write(str(tmp_path))
    bacchus.write_evaluation_order(tmp_path / "comb.txt", tmp_path / "order",
                                   two_nucleon_forces_only=two_nf)
    x = np.random.default_rng(4).standard_normal(inp.dimension)
    with bacchus.Operator.load(tmp_path / "comb.txt", 3, 3, tmp_path / "order",
                               tmp_path / "index_lists", tmp_path / "matrix_elements") as op:
        xd = torch.as_tensor(x, device="cuda")
        yd = torch.empty_like(xd)
        op.matvec(xd, yd)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
    y_ref = O.matvec(O.Problem.from_inputs(inp), x, scales=[0.0, 1.0, 0.0 if two_nf else 1.0])
    assert np.abs(y - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
