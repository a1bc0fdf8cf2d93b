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


def test_bf16_parts_carry_sixteen_bits():
    """hi + lo of an FP32 value keeps 16 significant bits: the error is below
    2^-16 relative, where one BF16 part (8 significant bits) leaves up to
    2^-8; lo is exactly the BF16 rounding of the remainder."""
    v = np.random.default_rng(3).standard_normal(1000).astype(np.float32)
    hi, lo = M.bf16_parts(v)
    assert np.all(hi == M.bf16_round(v))
    rel_one = np.abs(hi - v) / np.abs(v)
    rel_two = np.abs(hi + lo - v) / np.abs(v)
    assert rel_one.max() <= 2.0 ** -8 and rel_one.max() > 2.0 ** -10
    assert rel_two.max() <= 2.0 ** -16
    assert np.all(lo == M.bf16_round((v.astype(np.float64) - hi).astype(np.float32)))


def test_split_product_terms():
    """The split product is hi.hi + lo.hi + hi.lo: against a direct sum of
    those terms, and with both operands exactly BF16 values (lo = 0) equal
    to the plain BF16 product."""
    rng = np.random.default_rng(4)
    a = rng.standard_normal((6, 5)).astype(np.float32)
    w = rng.standard_normal((3, 5)).astype(np.float32)
    ah, al = M.bf16_parts(a)
    wh, wl = M.bf16_parts(w)
    want = ah @ wh.T + al @ wh.T + ah @ wl.T
    assert np.allclose(M.product(a, w, True, True, True), want, rtol=0, atol=1e-12)
    assert np.allclose(M.product(a, w, True, True, False), ah @ wh.T + al @ wh.T,
                       rtol=0, atol=1e-12)
    ab, wb = M.bf16_round(a), M.bf16_round(w)
    assert np.array_equal(M.product(ab, wb, True, True, True), M.product(ab, wb, True))


def test_split_activations_reach_the_fp64_network():
    """With weights on the BF16 grid, carrying inputs and activations as two
    BF16 parts brings the forward pass to the FP64 network within ~1e-5
    relative; plain BF16 operands stay ~1e-3 off."""
    shape = (2, 64, 2, 8)
    n = 2 * 64 + 64 * 64 + 64 * 8 + 64 + 64 + 8
    rng = np.random.default_rng(5)
    p = M.bf16_round((rng.standard_normal(n) * 0.3).astype(np.float32))
    x = rng.uniform(-1, 1, (20, 2))
    exact, *_ = M.forward(p, shape, x)
    split, *_ = M.forward(p, shape, x, bf16=True, precision=M.SPLIT_ACTIVATIONS)
    plain, *_ = M.forward(p, shape, x, bf16=True)
    e_split = np.abs(split - exact).max() / np.abs(exact).max()
    e_plain = np.abs(plain - exact).max() / np.abs(exact).max()
    assert e_split < 2e-5 and e_plain > 50 * e_split
