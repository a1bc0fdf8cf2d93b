"""CPU oracle of the learned-start MLP (numpy, FP64).

TEST INFRASTRUCTURE ONLY.  Mirrors the device MLP: layers in -> hidden
(x depth, GELU, exact erf form) -> out, parameters laid out as all weight
matrices (out x in, row-major) then all biases; loss = mean squared error.
With ``bf16=True`` every GEMM operand is rounded to BF16 (round to nearest
even) before the FP64 product, as the device does before MFMA.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf)


def bf16_round(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    u = np.where(finite, u + 0x7FFF + ((u >> 16) & 1), u)
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, shape):
    inp, hidden, depth, out = shape
    din = [inp] + [hidden] * depth
    dout = [hidden] * depth + [out]
    W, b, o = [], [], 0
    for i, j in zip(din, dout):
        W.append(np.asarray(params[o:o + i * j], dtype=np.float64).reshape(j, i))
        o += i * j
    for j in dout:
        b.append(np.asarray(params[o:o + j], dtype=np.float64))
        o += j
    return W, b


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def bf16_parts(a):
    """(hi, lo) with hi = bf16(a), lo = bf16(a - hi), a taken in FP32: the
    two BF16 parts an operand is carried as under split precision."""
    a32 = np.asarray(a, dtype=np.float32)
    hi = bf16_round(a32)
    lo = bf16_round((a32.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


def product(a, w, bf16=False, split_a=False, split_w=False):
    """a @ w.T with the device's operand rounding: plain FP64 (bf16=False),
    both operands rounded to BF16, or an operand carried as two BF16 parts,
    the product then summing hi.hi + lo.hi + hi.lo (no lo.lo term)."""
    if not bf16:
        return np.asarray(a, dtype=np.float64) @ np.asarray(w, dtype=np.float64).T
    ah, al = bf16_parts(a) if split_a else (bf16_round(a), None)
    wh, wl = bf16_parts(w) if split_w else (bf16_round(w), None)
    z = ah @ wh.T
    if split_a:
        z = z + al @ wh.T
    if split_w:
        z = z + ah @ wl.T
    return z


SPLIT_ACTIVATIONS = 1   # the device's BACCHUS_HIP_MLP_SPLIT_ACTIVATIONS
FP32_LAST = 2           # BACCHUS_HIP_MLP_FP32_LAST


def forward(params, shape, x, bf16=False, precision=0):
    W, b = unpack(params, shape)
    sa = bool(precision & SPLIT_ACTIVATIONS)
    a = np.asarray(x, dtype=np.float64)
    Zs, As = [], [a]
    for l, (w, bb) in enumerate(zip(W, b)):
        last = l == len(W) - 1
        z = product(a, w, bf16, sa, last and bool(precision & FP32_LAST)) + bb
        Zs.append(z)
        a = z if last else gelu(z)
        As.append(a)
    return a, Zs, As


def loss_and_gradients(params, shape, x, t, bf16=False, precision=0):
    W, b = unpack(params, shape)
    sa = bool(precision & SPLIT_ACTIVATIONS)
    y, Zs, As = forward(params, shape, x, bf16, precision)
    n = y.size
    loss = float(((y - t) ** 2).sum() / n)
    dz = 2.0 * (y - t) / n
    gW, gb = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        # dW = dz^T a: both operands split under split activations
        gW[l] = product(np.asarray(dz).T, np.asarray(As[l]).T, bf16, sa, sa)
        gb[l] = dz.sum(axis=0)
        if l == 0:
            break
        last = l == len(W) - 1
        da = product(dz, np.asarray(W[l]).T, bf16, sa, last and bool(precision & FP32_LAST))
        dz = da * gelu_grad(Zs[l - 1])
    grads = np.concatenate([g.ravel() for g in gW] + [g.ravel() for g in gb])
    return loss, grads
