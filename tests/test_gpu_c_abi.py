"""The library called from plain C (tests/c_abi/bacchus_driver.c: no Python
and no torch in the process), linked as a Bacchus build would link it: the
file-based matvec against the compiled reference's y, and the reference-mode
Lanczos against the compiled reference's converged eigenvalues (he4_nmax4)."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from jupiterncsm_amd.inputs import SystemConfig, build

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c_abi", "bacchus_driver")


def test_c_caller_file_matvec_and_reference_lanczos(golden_dir):
    print("case c abi he4_nmax4", flush=True)
    assert os.path.exists(DRIVER), "build it with __graft_entry__.build()"
    g = np.load(os.path.join(golden_dir, "he4_nmax4.npz"))
    inp = build(SystemConfig(2, 2, 4, max_body=2, seed=1))
    d = tempfile.mkdtemp(prefix="bacchus_c_abi_")
    try:
        inp.write(d)
        off = inp.block_offsets
        y0 = np.random.default_rng(2).standard_normal(inp.dimension)
        for name, v in (("x", g["x"]), ("y", y0)):
            os.makedirs(os.path.join(d, name))
            for b in range(len(inp.blocks)):
                v[off[b]:off[b + 1]].tofile(os.path.join(d, name, f"vec_{b + 1}"))
        r = subprocess.run([DRIVER, os.path.join(d, "comb.txt"), "2", "2",
                            os.path.join(d, "evaluation_order"), os.path.join(d, "index_lists"),
                            os.path.join(d, "matrix_elements"), os.path.join(d, "x"),
                            os.path.join(d, "y"), "200"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        y = np.concatenate([np.fromfile(os.path.join(d, "y", f"vec_{b + 1}"))
                            for b in range(len(inp.blocks))])
    finally:
        shutil.rmtree(d, ignore_errors=True)
    ref = g["y_ref"]
    assert np.abs((y - y0) - ref).max() <= 1e-12 * np.abs(ref).max()
    lines = r.stdout.split("\n")
    assert lines[0].startswith("iterations ")
    k = int(lines[0].split()[1])
    ev = np.array([float(v) for v in lines[1:1 + k]])
    conv = g["eig_ref_conv"]
    assert k == int(g["krylov_conv"])
    assert np.abs(ev[:len(conv)] - conv).max() <= 1e-8 * np.abs(conv).max()
