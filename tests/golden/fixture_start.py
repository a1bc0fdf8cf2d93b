"""The seeded start vector of the full-length Lanczos fixtures
(tests/golden/lanczos300_<name>.npz, written by make_lanczos_fixtures.py):
a standard normal vector of numpy's default generator, normalized."""
import numpy as np

SEED = 20261204


def start_vector(dimension, seed=SEED):
    v = np.random.default_rng(seed).standard_normal(dimension)
    return v / np.linalg.norm(v)
