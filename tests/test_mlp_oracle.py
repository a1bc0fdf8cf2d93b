"""The MLP oracle's analytic gradients pinned by central finite differences
(FP64, no BF16 rounding)."""
import numpy as np

from oracle import mlp_oracle as M


def test_gradients_match_finite_differences():
    shape = (2, 8, 2, 3)
    n = 2 * 8 + 8 * 8 + 8 * 3 + 8 + 8 + 3
    rng = np.random.default_rng(0)
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((5, 2))
    t = rng.standard_normal((5, 3))
    _, g = M.loss_and_gradients(p, shape, x, t)
    fd = np.zeros(n)
    h = 1e-6
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (M.loss_and_gradients(p + e, shape, x, t)[0]
                 - M.loss_and_gradients(p - e, shape, x, t)[0]) / (2 * h)
    assert np.abs(g - fd).max() < 1e-7 * max(1.0, np.abs(g).max())


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 3.0])
    # GELU(0) = 0, GELU'(0) = 1/2, GELU(x) - GELU(-x) = x
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    assert np.allclose(M.gelu(z) - M.gelu(-z), z, atol=1e-15)


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0], dtype=np.float32)
    got = M.bf16_round(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0          # tie -> even
    assert got[2] == 1.0 + 2 ** -7
    assert got[3] == -2.5 and got[4] == 0.0
