"""The drop-in at the reference's own C interfaces: jupiterncsm_amd/dropin/
replaces src/Bacchus/matrix/matrix.c, src/Minerva/scheduler/scheduler.c and
src/Bacchus/lanczos/lanczos.c, and the rest of the reference runs on top of
it unchanged.

oracle/Makefile links one driver (oracle/ref_driver.c, a main() around the
reference's library calls) twice: with the reference's three files
(oracle/_ref/bacchus_ref) and with the drop-in's (oracle/_ref/bacchus_dropin,
calling libbacchus_hip).  The GPU tests run the drop-in binary on the golden
systems and compare with what the reference binary wrote into
tests/golden/*.npz: y = H x, the eigenvalues of three Lanczos runs (fixed
iterations, the eigenvalue and the eigenvector-amplitude criteria) with
their Krylov dimensions, and eigenvector_1/2 written as bacchus.c:70-87
writes them.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from jupiterncsm_amd.inputs import SystemConfig, build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "bacchus_dropin")
REFERENCE_BIN = os.path.join(ROOT, "oracle", "_ref", "bacchus_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = {
    "he4_nmax4": SystemConfig(2, 2, 4, max_body=2, seed=1),
    "he4_nmax2_3b": SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3),
    "li6_nmax2_3b": SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1),
}
ITERATIONS_FIXED = {"he4_nmax4": 60, "he4_nmax2_3b": 40, "li6_nmax2_3b": 60}
SHIM_FUNCTIONS = ["new_scheduler", "run_matrix_vector_multiplication", "free_scheduler",
                  "new_generative_matrix", "matrix_vector_multiplication", "new_zero_matrix",
                  "free_matrix", "new_lanczos_environment", "diagonalize", "get_eigensystem",
                  "free_lanczos_environment"]

needs_binary = pytest.mark.skipif(not os.path.exists(DROPIN),
                                  reason="oracle/_ref/bacchus_dropin is built only where the "
                                         "reference sources are (make -C oracle dropin)")


class System:
    """A golden system written in the reference's file formats."""

    def __init__(self, name):
        self.name = name
        self.cfg = CASES[name]
        self.inp = build(self.cfg)
        self.dir = tempfile.mkdtemp(prefix=f"dropin_{name}_")
        self.inp.write(self.dir)
        self.golden = np.load(os.path.join(GOLDEN, f"{name}.npz"))

    def args(self):
        d = self.dir
        return [os.path.join(d, "comb.txt"), str(self.cfg.num_protons),
                str(self.cfg.num_neutrons), os.path.join(d, "evaluation_order"),
                os.path.join(d, "index_lists"), os.path.join(d, "matrix_elements")]

    def write_vector(self, sub, v):
        path = os.path.join(self.dir, sub)
        os.makedirs(path, exist_ok=True)
        off = self.inp.block_offsets
        for b in range(len(self.inp.blocks)):
            v[off[b]:off[b + 1]].tofile(os.path.join(path, f"vec_{b + 1}"))
        return path

    def read_vector(self, path):
        return np.concatenate([np.fromfile(os.path.join(path, f"vec_{b + 1}"))
                               for b in range(len(self.inp.blocks))])

    def close(self):
        shutil.rmtree(self.dir, ignore_errors=True)


def run(binary, args, **kw):
    print("run", os.path.basename(binary), args[0], flush=True)
    return subprocess.run([binary, *args], capture_output=True, text=True, timeout=300,
                          env=dict(os.environ, OMP_NUM_THREADS="4"), **kw)


@needs_binary
def test_dropin_binary_defines_the_reference_interfaces_and_calls_the_library():
    """The drop-in binary defines the scheduler.h / matrix.h / lanczos.h
    functions itself and imports the GPU library's entry points."""
    defined = subprocess.run(["nm", "--defined-only", DROPIN], capture_output=True,
                             text=True, check=True).stdout.split()
    undefined = subprocess.run(["nm", "-D", "--undefined-only", DROPIN], capture_output=True,
                               text=True, check=True).stdout.split()
    for f in SHIM_FUNCTIONS:
        assert f in defined, f
    for f in ["bacchus_hip_operator_load_blocks", "bacchus_hip_run_matrix_vector_multiplication",
              "bacchus_hip_lanczos", "bacchus_hip_operator_create"]:
        assert f in undefined, f


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"),
                    reason="compiles against the reference's headers, present where the "
                           "reference tree is")
def test_dropin_sources_compile_against_the_reference_headers():
    """The three files compile warning-free against the reference's own
    headers: the same prototypes as the translation units they replace."""
    src = "/root/reference/src"
    for f in ["dropin_matrix.c", "dropin_scheduler.c", "dropin_lanczos.c"]:
        r = subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            f"-I{src}", f"-I{src}/Utilities", f"-I{src}/Bacchus",
                            f"-I{src}/Minerva", f"-I{ROOT}/include",
                            os.path.join(ROOT, "jupiterncsm_amd", "dropin", f)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@needs_binary
def test_dropin_reports_errors_the_reference_way_without_gpu():
    """Without a GPU (or with a broken input) the drop-in exits with
    EXIT_FAILURE and an "Error in <function> @ <file>:<line>" message, the
    reference's error() convention (src/Utilities/error/error.h:7-13)."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present: the missing-file case covers the error path")
    s = System("he4_nmax2_3b")
    try:
        x = s.write_vector("x", np.zeros(s.inp.dimension))
        y = s.write_vector("y", np.zeros(s.inp.dimension))
        r = run(DROPIN, ["matvec", *s.args(), x, y])
        assert r.returncode == 1
        assert "Error in" in r.stderr
    finally:
        s.close()


@pytest.mark.gpu
@needs_binary
def test_dropin_reports_a_missing_index_list():
    s = System("he4_nmax2_3b")
    try:
        os.remove(os.path.join(s.dir, "index_lists", sorted(os.listdir(
            os.path.join(s.dir, "index_lists")))[0]))
        x = s.write_vector("x", np.zeros(s.inp.dimension))
        y = s.write_vector("y", np.zeros(s.inp.dimension))
        r = run(DROPIN, ["matvec", *s.args(), x, y])
        print(r.stderr, flush=True)
        assert r.returncode == 1
        assert "Error in" in r.stderr and "index_list_" in r.stderr
    finally:
        s.close()


@pytest.mark.gpu
@needs_binary
@pytest.mark.parametrize("name", list(CASES))
def test_dropin_matvec_matches_reference(name):
    s = System(name)
    try:
        g = s.golden
        x = s.write_vector("x", g["x"])
        y = s.write_vector("y", np.zeros(s.inp.dimension))
        r = run(DROPIN, ["matvec", *s.args(), x, y])
        assert r.returncode == 0, r.stderr
        got = s.read_vector(y)
        err = np.abs(got - g["y_ref"]).max() / np.abs(g["y_ref"]).max()
        print(name, "drop-in matvec vs reference", err, flush=True)
        assert err <= 1e-12
    finally:
        s.close()


@pytest.mark.gpu
@needs_binary
@pytest.mark.parametrize("name", list(CASES))
def test_dropin_lanczos_matches_reference(name):
    """Three runs of the reference's Lanczos interface on the GPU: fixed
    iterations, the eigenvalue criterion at 1e-7 (with eigenvectors 1, 2
    written through eigensystem.c as bacchus.c does) and the
    eigenvector-amplitude criterion at 1e-5.  Same Krylov dimensions as the
    reference, eigenvalues within 1e-8 relative, eigenvectors within 1e-8."""
    s = System(name)
    try:
        g = s.golden
        a = s.args()
        d = s.dir
        r = run(DROPIN, ["lanczos", *a, os.path.join(d, "k1"), str(ITERATIONS_FIXED[name]),
                         "1e-12", "2", os.path.join(d, "e1")])
        assert r.returncode == 0, r.stderr
        e1 = np.loadtxt(os.path.join(d, "e1"), ndmin=1)
        assert len(e1) == len(g["eig_ref_fixed"])
        scale = np.abs(g["eig_ref_fixed"]).max()
        print(name, "fixed", np.abs(e1 - g["eig_ref_fixed"]).max() / scale, flush=True)
        assert np.abs(e1 - g["eig_ref_fixed"]).max() <= 1e-8 * scale

        os.makedirs(os.path.join(d, "ev"))
        r = run(DROPIN, ["lanczos", *a, os.path.join(d, "k2"), "200", "1e-7", "0",
                         os.path.join(d, "e2"), os.path.join(d, "ev"), "1"])
        assert r.returncode == 0, r.stderr
        e2 = np.loadtxt(os.path.join(d, "e2"), ndmin=1)
        assert len(e2) == int(g["krylov_conv"])
        assert np.abs(e2 - g["eig_ref_conv"]).max() <= 1e-8 * np.abs(g["eig_ref_conv"]).max()
        for i in range(2):
            v = s.read_vector(os.path.join(d, "ev", f"eigenvector_{i + 1}"))
            dv = np.abs(v - g["eigvec_ref_conv"][i]).max()
            print(name, "eigenvector", i + 1, "max diff", dv, flush=True)
            assert dv <= 1e-8

        r = run(DROPIN, ["lanczos", *a, os.path.join(d, "k3"), "200", "1e-5", "1",
                         os.path.join(d, "e3")])
        assert r.returncode == 0, r.stderr
        e3 = np.loadtxt(os.path.join(d, "e3"), ndmin=1)
        print(name, "eigenvector criterion K", len(e3), "reference", int(g["krylov_vec"]),
              flush=True)
        assert len(e3) == int(g["krylov_vec"])
        assert np.abs(e3 - g["eig_ref_vec"]).max() <= 1e-8 * np.abs(g["eig_ref_vec"]).max()
    finally:
        s.close()


@pytest.mark.gpu
@needs_binary
@pytest.mark.parametrize("case", ["3x3", "50x50"])
def test_dropin_explicit_matrix_lanczos(case):
    """The EXPLICIT backend through the drop-in: the reference's own Lanczos
    test matrices (lanczos.c:295-352, the 3x3 closed form; :354-404, a random
    symmetric 50x50 against a dense solver) on the GPU, and against the
    reference binary on the same matrix (same Krylov dimension)."""
    work = tempfile.mkdtemp(prefix="dropin_explicit_")
    try:
        if case == "3x3":
            M = np.array([[1.0, 2, 3], [2, 1, 2], [3, 2, 1]])
            want = np.array([-2.0, 2.5 - np.sqrt(41) / 2, 2.5 + np.sqrt(41) / 2])
            iters = 3
        else:
            A = np.random.default_rng(50).uniform(-1, 1, (50, 50))
            M = A + A.T
            want = np.linalg.eigvalsh(M)
            iters = 50
        path = os.path.join(work, "m.bin")
        np.ascontiguousarray(M).tofile(path)
        outs = {}
        for tag, binary in [("dropin", DROPIN), ("reference", REFERENCE_BIN)]:
            res = os.path.join(work, f"e_{tag}")
            r = run(binary, ["explicit", path, str(len(M)), os.path.join(work, f"k_{tag}"),
                             str(iters), "1e-12", "2", res])
            assert r.returncode == 0, r.stderr
            outs[tag] = np.loadtxt(res, ndmin=1)
        got = outs["dropin"]
        print(case, "drop-in", got[:3], "reference", outs["reference"][:3], flush=True)
        assert len(got) == len(outs["reference"]) == iters
        # the 3x3 test's own tolerance is 1e-5 (lanczos.c:295-352); the
        # full-dimension run reproduces the spectrum to rounding
        assert np.abs(got - want).max() <= 1e-8 * np.abs(want).max()
        assert np.abs(got - outs["reference"]).max() <= 1e-8 * np.abs(want).max()
    finally:
        shutil.rmtree(work, ignore_errors=True)
