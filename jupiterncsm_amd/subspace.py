"""Subspace projection of the operator onto stored vectors (the Ceres step
that consumes stored eigenvectors, src/Ceres/subspace_operator/
subspace_operator.c:7-61 and src/Ceres/norm_matrix/norm_matrix.c:4-28).

The reference multiplies every training vector by H with one Bacchus matvec
each and forms v_j . (H v_i) and v_i . v_j with disk-backed dot products.
Here all n vectors go through one batched matvec (FP32/BF16 operator, FP32
vectors, FP64 sums in LDS) and the n x n dot products run in the stored-basis
kernels with FP64 accumulation.  The matrices are written in NumPy's .npy
format like save_as_numpy_matrix (src/Ceres/numpy_matrix/numpy_matrix.c:
7-36).
"""
from __future__ import annotations

import numpy as np

from . import bacchus


def norm_matrix(vectors):
    """N[i][j] = v_i . v_j for float32 device vectors (n, D)."""
    return bacchus.basis_gram(vectors.contiguous())


def subspace_operator(operator, vectors, scales=(1.0, 1.0, 1.0)):
    """M[i][j] = v_j . (H v_i), mirrored from j >= i as the reference does,
    for an FP32 or BF16 operator and float32 device vectors (n, D); H is
    s1 H1 + s2 V2 + s3 V3."""
    import torch
    U = vectors.contiguous()
    n, D = U.shape
    if D != operator.dimension:
        raise bacchus.BacchusError(f"vectors have {D} components, the operator {operator.dimension}")
    with bacchus.Batch([operator], [n]) as bt:
        bt.set_scales(0, np.tile(np.asarray(scales, dtype=np.float64), (n, 1)))
        X = U.T.contiguous()
        Y = torch.empty_like(X)
        bt.matvec([X], [Y])
        # P[j][i] = v_j . (H v_i)
        P = bacchus.basis_project(U, Y.to(torch.float64).contiguous())
    M = np.empty((n, n))
    for i in range(n):
        for j in range(i, n):
            M[i, j] = M[j, i] = P[j, i]
    return M


def save_numpy_matrix(path, matrix):
    """FP64 row-major .npy file (version 1.0 header), the format
    save_as_numpy_matrix writes."""
    np.save(path, np.ascontiguousarray(matrix, dtype="<f8"), allow_pickle=False)
