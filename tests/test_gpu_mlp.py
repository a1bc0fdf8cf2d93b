"""The learned-start MLP's device kernels (BF16 MFMA, FP32 accumulate)
against the CPU oracle with the same BF16 rounding."""
import numpy as np
import pytest

from jupiterncsm_amd import bacchus
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def test_forward_and_gradients_match_oracle():
    import torch
    print("case mlp 2-256x3-64", flush=True)
    mlp = bacchus.MLP(2, 256, 3, 64, seed=5)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.8, 1.2, (37, 2)).astype(np.float32)
    t = (rng.standard_normal((37, 64)) * 0.1).astype(np.float32)
    p = mlp.params().astype(np.float64)
    xd = torch.as_tensor(x, device="cuda")
    td = torch.as_tensor(t, device="cuda")
    y = mlp.forward(xd).cpu().numpy()
    y_ref, *_ = M.forward(p, mlp.shape, x, bf16=True)
    assert np.abs(y - y_ref).max() <= 1e-3 * np.abs(y_ref).max()
    loss, g = mlp.gradients(xd, td)
    loss_ref, g_ref = M.loss_and_gradients(p, mlp.shape, x, t, bf16=True)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    # BF16 operands (including dZ in the backward GEMMs), FP32 accumulation
    assert np.abs(g - g_ref).max() <= 2e-2 * np.abs(g_ref).max()
    cos = g @ g_ref / np.linalg.norm(g) / np.linalg.norm(g_ref)
    assert cos > 0.9999
    mlp.close()


def test_mfma_gemm_tile_edges():
    """Ragged shapes (not multiples of 16) through the same kernels."""
    import torch
    print("case mlp ragged 3-20x2-7", flush=True)
    mlp = bacchus.MLP(3, 20, 2, 7, seed=2)
    x = np.random.default_rng(2).standard_normal((5, 3)).astype(np.float32)
    y = mlp.forward(torch.as_tensor(x, device="cuda")).cpu().numpy()
    y_ref, *_ = M.forward(mlp.params().astype(np.float64), mlp.shape, x, bf16=True)
    assert np.abs(y - y_ref).max() <= 1e-3 * max(1.0, np.abs(y_ref).max())
    mlp.close()


def test_training_reduces_loss():
    import torch
    print("case mlp training", flush=True)
    mlp = bacchus.MLP(2, 64, 2, 4, seed=3)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.8, 1.2, (256, 2)).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, 0]), np.cos(2 * x[:, 1]), x[:, 0] * x[:, 1],
                  x[:, 0] - x[:, 1]], axis=1).astype(np.float32)
    xd, td = torch.as_tensor(x, device="cuda"), torch.as_tensor(t, device="cuda")
    first = mlp.train_step(xd, td, lr=3e-3, want_loss=True)
    for _ in range(400):
        mlp.train_step(xd, td, lr=3e-3)
    last = mlp.train_step(xd, td, lr=3e-3, want_loss=True)
    assert last < 0.05 * first
    mlp.close()


@pytest.mark.parametrize("precision", [bacchus.MLP_SPLIT_ACTIVATIONS,
                                       bacchus.MLP_SPLIT_ACTIVATIONS | bacchus.MLP_FP32_LAST])
def test_split_precision_matches_oracle(precision):
    """Activations (and the output layer's weights) carried as two BF16
    parts: forward and gradients against the oracle emulating the same
    parts; with weights that are BF16 values already, the forward is then
    the FP64 network to FP32-level accuracy, where plain BF16 operands are
    off by ~1e-3."""
    import torch
    print(f"case mlp split precision {precision}", flush=True)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1.0, 1.0, (37, 2)).astype(np.float32)
    t = (rng.standard_normal((37, 64)) * 0.1).astype(np.float32)
    xd, td = torch.as_tensor(x, device="cuda"), torch.as_tensor(t, device="cuda")
    mlp = bacchus.MLP(2, 256, 3, 64, seed=6, precision=precision)
    p = mlp.params().astype(np.float64)
    y = mlp.forward(xd).cpu().numpy()
    y_ref, *_ = M.forward(p, mlp.shape, x, bf16=True, precision=precision)
    assert np.abs(y - y_ref).max() <= 1e-4 * np.abs(y_ref).max()
    loss, g = mlp.gradients(xd, td)
    loss_ref, g_ref = M.loss_and_gradients(p, mlp.shape, x, t, bf16=True, precision=precision)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    assert np.abs(g - g_ref).max() <= 2e-3 * np.abs(g_ref).max()
    # weights on the BF16 grid: split activations leave only FP32 rounding
    pb = M.bf16_round(p.astype(np.float32))
    mlp.set_params(pb.astype(np.float32))
    plain = bacchus.MLP(2, 256, 3, 64, seed=6)
    plain.set_params(pb.astype(np.float32))
    y_exact, *_ = M.forward(pb, mlp.shape, x)
    e_split = np.abs(mlp.forward(xd).cpu().numpy() - y_exact).max() / np.abs(y_exact).max()
    e_plain = np.abs(plain.forward(xd).cpu().numpy() - y_exact).max() / np.abs(y_exact).max()
    print(f"vs FP64 network: split {e_split:.2e}, plain BF16 {e_plain:.2e}", flush=True)
    assert e_split <= 1e-4 and e_split * 20 <= e_plain
    mlp.close()
    plain.close()


def test_unknown_precision_flags_rejected():
    with pytest.raises(bacchus.BacchusError, match="precision"):
        bacchus.MLP(2, 8, 1, 2, precision=4)
