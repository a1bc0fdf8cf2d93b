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


def forward(params, shape, x, bf16=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W, b = unpack(params, shape)
    a = np.asarray(x, dtype=np.float64)
    Zs, As = [], [a]
    for l, (w, bb) in enumerate(zip(W, b)):
        z = r(a) @ r(w).T + bb
        Zs.append(z)
        a = z if l == len(W) - 1 else gelu(z)
        As.append(a)
    return a, Zs, As


def loss_and_gradients(params, shape, x, t, bf16=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W, b = unpack(params, shape)
    y, Zs, As = forward(params, shape, x, bf16)
    n = y.size
    loss = float(((y - t) ** 2).sum() / n)
    dz = 2.0 * (y - t) / n
    gW, gb = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        gW[l] = r(dz).T @ r(As[l])
        gb[l] = dz.sum(axis=0)
        if l == 0:
            break
        da = r(dz) @ r(W[l])
        dz = da * gelu_grad(Zs[l - 1])
    grads = np.concatenate([g.ravel() for g in gW] + [g.ravel() for g in gb])
    return loss, grads
