"""bacchus_hip_basis_project_static, the projection entry point that
allocates no device memory (its partial sums live in a static device
array), against bacchus_hip_basis_project through the C ABI: the same
kernels in the same order, so equal bit for bit, on two calls in a row (the
second smaller, over the scratch the first one left) and with sizes beyond
the static array rejected."""
import ctypes

import numpy as np
import pytest

from jupiterncsm_amd import bacchus

pytestmark = pytest.mark.gpu


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    return torch


def project(fn, U, V):
    n, D = U.shape
    B = V.shape[1]
    out = np.full((n, B), np.nan)
    rc = fn(ctypes.c_void_p(U.data_ptr()), n, D, ctypes.c_void_p(V.data_ptr()), B,
            out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


def test_static_projection_equals_the_allocating_one():
    print("case static projection", flush=True)
    torch = torch_dev()
    lib = bacchus.library()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    for n, D, B in ((48, 20011, 16), (5, 3001, 3)):
        U = torch.randn(n, D, dtype=torch.float32, device="cuda", generator=gen)
        V = torch.randn(D, B, dtype=torch.float64, device="cuda", generator=gen)
        rc1, a = project(lib.bacchus_hip_basis_project, U, V)
        rc2, b = project(lib.bacchus_hip_basis_project_static, U, V)
        torch.cuda.synchronize()
        assert rc1 == 0 and rc2 == 0, bacchus.library().bacchus_hip_last_error()
        ref = U.double().cpu().numpy() @ V.cpu().numpy()
        print(n, D, B, "static vs allocating max diff", np.abs(a - b).max(),
              "vs numpy", np.abs(b - ref).max(), flush=True)
        assert np.array_equal(a, b)
        assert np.abs(b - ref).max() <= 1e-10 * np.abs(ref).max()


def test_static_projection_rejects_sizes_beyond_its_scratch():
    torch = torch_dev()
    lib = bacchus.library()
    U = torch.zeros(255, 64, dtype=torch.float32, device="cuda")
    V = torch.zeros(64, 32, dtype=torch.float64, device="cuda")
    rc, _ = project(lib.bacchus_hip_basis_project_static, U, V)
    assert rc != 0
    assert b"static scratch" in lib.bacchus_hip_last_error()
