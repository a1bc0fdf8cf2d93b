"""Writes the full-length Lanczos fixtures of tests/golden/ with the oracle.

For the 2+2 Nmax 8 and the 3+3 Nmax 6 (2+3-body) workloads: 300 iterations of
the oracle's Lanczos (oracle/oracle.c, the reference's recurrence of
src/Bacchus/lanczos/lanczos.c:60-253 with classical Gram-Schmidt twice
against every Krylov vector) from the seeded start vector start_vector(D),
and its lowest 5 Ritz values:

  lanczos300_<name>.npz: seed, dimension, iterations, eigenvalues

The GPU path and every reduced-precision mode are measured against these
eigenvalues (tests/test_gpu_configs.py, bench.py).  The fixtures are data;
this script runs the oracle in this container (several minutes at 3+3 Nmax 6).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from jupiterncsm_amd.inputs import SYSTEMS, build  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.golden.fixture_start import SEED, start_vector  # noqa: E402

ITERATIONS = 300
NAMES = ["he4_nmax8", "li6_nmax6_3b"]


def main(names):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name in names:
        inp = build(SYSTEMS[name])
        p = O.Problem.from_inputs(inp)
        D = inp.dimension
        t = time.time()
        ev, _, _, k = O.lanczos(D, lambda v: O.matvec(p, v), ITERATIONS,
                                start=O.START_GIVEN, start_vector=start_vector(D),
                                reorth=O.REORTH_FULL, num_eigenvalues=5)
        np.savez(os.path.join(out_dir, f"lanczos300_{name}.npz"), seed=SEED, dimension=D,
                 iterations=k, eigenvalues=ev)
        print(name, "D", D, "iterations", k, "eigenvalues", ev, f"{time.time() - t:.0f}s",
              flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or NAMES)
