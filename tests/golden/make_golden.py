"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs the compiled reference (oracle/_ref/bacchus_ref, built by
`make -C oracle ref` from the reference sources) on generated inputs written
in the reference's file formats, and stores:

  x, y_ref          one matvec y = H x through run_matrix_vector_multiplication
                    (src/Minerva/scheduler/scheduler.c:80-145)
  eig_ref_fixed     eigenvalues of T after a fixed number of Lanczos iterations
                    (no convergence test, src/Bacchus/lanczos/lanczos.c:171-253)
  eig_ref_conv      eigenvalues after converging lambda_0 to 1e-7
  krylov_conv       Krylov dimension reached by that run
  eigvec_ref_conv   eigenvector_1 and eigenvector_2 written by that run the way
                    src/Bacchus/programs/bacchus.c:70-87 writes them
  eig_ref_vec       eigenvalues after the eigenvector-amplitude criterion
                    (lanczos.c:276-287) at 1e-5, krylov_vec its Krylov dimension
  digest            checksums of the generated arrays (pins the generator)

and, from the compiled Mars program (oracle/_ref/mars_ref,
src/Mars/programs/mars.c), the evaluation order it writes for each case's
comb.txt: <name>.mars_order (default mode) and, for the 3-body cases,
<name>.mars_order_2nf (--only-two-nucleon-forces).

Only runs where the reference tree is present; the fixtures are data.
`--mars-only` rewrites just the evaluation-order fixtures.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from jupiterncsm_amd.inputs import SystemConfig, build  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "bacchus_ref")
MARS = os.path.join(ROOT, "oracle", "_ref", "mars_ref")

CASES = {
    "he4_nmax4": (SystemConfig(2, 2, 4, max_body=2, seed=1), 60),
    "he4_nmax2_3b": (SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3), 40),
    "li6_nmax2_3b": (SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1), 60),
}


def digest(inp):
    return np.array([
        inp.dimension, len(inp.blocks), len(inp.list_offsets) - 1, int(inp.list_offsets[-1]),
        int(inp.list_entries.astype(np.int64).sum()),
        int((inp.list_entries.astype(np.int64) * np.arange(1, inp.list_entries.size + 1).reshape(
            inp.list_entries.shape) % 1000003).sum()),
        len(inp.matrix_offsets) - 1, int(inp.matrix_offsets[-1]), len(inp.instructions),
    ], dtype=np.int64), np.array([inp.matrix_elements.sum(), np.abs(inp.matrix_elements).sum(),
                                  (inp.matrix_elements ** 2).sum()])


def write_vector(inp, directory, v):
    os.makedirs(directory, exist_ok=True)
    off = inp.block_offsets
    for b in range(len(inp.blocks)):
        v[off[b]:off[b + 1]].tofile(os.path.join(directory, f"vec_{b + 1}"))


def read_vector(inp, directory):
    return np.concatenate([np.fromfile(os.path.join(directory, f"vec_{b + 1}"))
                           for b in range(len(inp.blocks))])


def run(args):
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([REF, *args], capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])


def main():
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name, (cfg, iters) in CASES.items():
        inp = build(cfg)
        work = tempfile.mkdtemp(prefix=f"golden_{name}_")
        try:
            inp.write(work)
            comb = [os.path.join(work, "comb.txt"), str(cfg.num_protons),
                    str(cfg.num_neutrons), os.path.join(work, "evaluation_order"),
                    os.path.join(work, "index_lists"), os.path.join(work, "matrix_elements")]
            x = np.random.default_rng(2024).standard_normal(inp.dimension)
            write_vector(inp, os.path.join(work, "x"), x)
            write_vector(inp, os.path.join(work, "y"), np.zeros(inp.dimension))
            run(["matvec", *comb, os.path.join(work, "x"), os.path.join(work, "y")])
            y = read_vector(inp, os.path.join(work, "y"))
            run(["lanczos", *comb, os.path.join(work, "k1"), str(iters), "1e-12", "2",
                 os.path.join(work, "e1")])
            eig_fixed = np.loadtxt(os.path.join(work, "e1"), ndmin=1)
            os.makedirs(os.path.join(work, "ev"))
            run(["lanczos", *comb, os.path.join(work, "k2"), "200", "1e-7", "0",
                 os.path.join(work, "e2"), os.path.join(work, "ev"), "1"])
            eig_conv = np.loadtxt(os.path.join(work, "e2"), ndmin=1)
            # eigenvector_1, eigenvector_2 of that run (bacchus.c:70-87)
            eigvec = np.stack([read_vector(inp, os.path.join(work, "ev", f"eigenvector_{i + 1}"))
                               for i in range(2)])
            # the eigenvector-amplitude criterion (lanczos.c:276-287)
            run(["lanczos", *comb, os.path.join(work, "k3"), "200", "1e-5", "1",
                 os.path.join(work, "e3")])
            eig_vec = np.loadtxt(os.path.join(work, "e3"), ndmin=1)
            ints, floats = digest(inp)
            np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), x=x, y_ref=y,
                                eig_ref_fixed=eig_fixed, iterations_fixed=iters,
                                eig_ref_conv=eig_conv, krylov_conv=len(eig_conv),
                                eigvec_ref_conv=eigvec, eig_ref_vec=eig_vec,
                                krylov_vec=len(eig_vec),
                                digest_int=ints, digest_float=floats)
            print(name, "D", inp.dimension, "|y|", np.abs(y).max(), "lowest", eig_fixed[:3],
                  "conv K", len(eig_conv))
        finally:
            shutil.rmtree(work, ignore_errors=True)


def mars_orders():
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name, (cfg, _) in CASES.items():
        inp = build(cfg)
        work = tempfile.mkdtemp(prefix=f"mars_{name}_")
        try:
            inp.write(work)
            comb = os.path.join(work, "comb.txt")
            modes = [("mars_order", [])]
            if cfg.max_body >= 3:
                modes.append(("mars_order_2nf", ["--only-two-nucleon-forces"]))
            for suffix, extra in modes:
                target = os.path.join(out_dir, f"{name}.{suffix}")
                r = subprocess.run([MARS, comb, target, *extra], capture_output=True, text=True,
                                   cwd=work)
                if r.returncode != 0:
                    raise RuntimeError(r.stderr[-2000:])
                print(name, suffix, sum(1 for _ in open(target)), "blocks")
        finally:
            shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    if "--mars-only" not in sys.argv:
        main()
    mars_orders()
