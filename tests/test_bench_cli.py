"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message, and the JSON line's keys are those the contract
names (checked on the source, since producing the line needs the GPU)."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_headline_line_carries_the_contract_keys():
    src = open(os.path.join(ROOT, "bench.py")).read()
    keys = set(re.findall(r'^\s+"([a-z_0-9]+)":', src, flags=re.M))
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in keys, k
    for k in ("bound", "achieved", "peak", "frac", "traffic"):
        assert f'"{k}"' in src, k


def test_projected_family_ritz_is_the_generalized_eigenproblem():
    """ProjectedFamily.ritz solves (P1 + c1 P2 + c2 P3) s = theta N s on the
    host: against scipy's generalized symmetric solver, with N-normalized
    coefficient vectors (no GPU involved)."""
    import numpy as np
    import scipy.linalg
    from jupiterncsm_amd.learned import ProjectedFamily
    rng = np.random.default_rng(4)
    r = 12
    parts = rng.standard_normal((3, r, r))
    parts = parts + parts.transpose(0, 2, 1)
    A = rng.standard_normal((r, r))
    N = A @ A.T + r * np.eye(r)
    fam = ProjectedFamily(basis=None, rank=r, parts=parts, gram=N)
    c1, c2 = 0.93, 1.12
    th, S = fam.ritz(c1, c2, 5)
    M = parts[0] + c1 * parts[1] + c2 * parts[2]
    ref = scipy.linalg.eigh(M, N, eigvals_only=True)[:5]
    assert np.abs(th - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.allclose(S.T @ N @ S, np.eye(5), atol=1e-12)
    assert np.allclose(M @ S, N @ S * th, atol=1e-10)
