"""Subspace projection of the operator onto stored vectors (the Ceres step
that consumes stored eigenvectors, src/Ceres/subspace_operator/
subspace_operator.c:7-61 and src/Ceres/norm_matrix/norm_matrix.c:4-28).

The reference multiplies every training vector by H with one Bacchus matvec
each and forms v_j . (H v_i) and v_i . v_j with disk-backed dot products, in
FP64.  Float64 vectors take that form on the device: one FP64-interface
matvec per vector and FP64 dot products (bacchus_hip_subspace_operator,
bacchus_hip_norm_matrix).  Float32 vectors with an FP32/BF16 operator can
instead go through one batched matvec for all n vectors (FP32 vectors, FP64
sums in LDS) and the stored-basis kernels (FP64 accumulation).  The matrices
are written in NumPy's .npy format like save_as_numpy_matrix
(src/Ceres/numpy_matrix/numpy_matrix.c:7-36).
"""
from __future__ import annotations

import numpy as np

from . import bacchus


def norm_matrix(vectors):
    """N[i][j] = v_i . v_j for device vectors (n, D), float64 or float32."""
    import torch
    if vectors.dtype == torch.float64:
        return bacchus.norm_matrix(vectors.contiguous())
    return bacchus.basis_gram(vectors.contiguous())


def subspace_operator(operator, vectors, scales=(1.0, 1.0, 1.0)):
    """M[i][j] = v_j . (H v_i), mirrored from j >= i as the reference does,
    for device vectors (n, D); H is s1 H1 + s2 V2 + s3 V3.  Float64 vectors:
    one matvec of `operator` per vector (any storage precision; its scales
    are set to `scales`).  Float32 vectors: one batched matvec (FP32 or BF16
    operator)."""
    import torch
    U = vectors.contiguous()
    n, D = U.shape
    if D != operator.dimension:
        raise bacchus.BacchusError(f"vectors have {D} components, the operator {operator.dimension}")
    if U.dtype == torch.float64:
        operator.set_scales(*scales)
        return bacchus.subspace_operator(operator, U)
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
