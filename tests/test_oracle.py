"""The CPU oracle (oracle/liboracle.so) pinned against the reference's own
known answers, the compiled reference's outputs (tests/golden) and what the
mathematics fixes (closed forms, brute force on tiny bases)."""
import os

import numpy as np
import pytest

from jupiterncsm_amd.inputs import SYSTEMS, SystemConfig, build
from oracle import oracle as O
from tests import skip_cases as SK

GOLDEN = ["he4_nmax4", "he4_nmax2_3b", "li6_nmax2_3b"]
CONFIGS = {
    "he4_nmax4": SystemConfig(2, 2, 4, max_body=2, seed=1),
    "he4_nmax2_3b": SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3),
    "li6_nmax2_3b": SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1),
}


def dense(problem):
    D = problem.dimension
    H = np.zeros((D, D))
    for i in range(D):
        e = np.zeros(D)
        e[i] = 1.0
        H[:, i] = O.matvec(problem, e)
    return H


# --- tridiagonal eigenvalues (the dsteqr_ boundary) ---------------------------

@pytest.mark.parametrize("n", [1, 2, 7, 50, 300])
def test_tridiagonal_toeplitz_closed_form(n):
    a, b = 3.0, -1.25
    got = O.tridiagonal_eigenvalues(np.full(n, a), np.full(max(n - 1, 1), b))
    j = np.arange(1, n + 1)
    want = np.sort(a + 2 * b * np.cos(j * np.pi / (n + 1)))
    assert np.allclose(got, want, rtol=0, atol=1e-12 * (abs(a) + 2 * abs(b)))


def test_tridiagonal_matches_dense_eigensolver():
    rng = np.random.default_rng(3)
    for n in (5, 40, 200):
        d, e = rng.standard_normal(n) * 10, rng.standard_normal(n - 1)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.allclose(O.tridiagonal_eigenvalues(d, e), np.linalg.eigvalsh(T),
                           rtol=0, atol=1e-11)
    # a dropped off-diagonal term changes the answer (the check is sensitive)
    d, e = rng.standard_normal(20), rng.standard_normal(19)
    e2 = e.copy()
    e2[7] = 0
    assert not np.allclose(O.tridiagonal_eigenvalues(d, e2), O.tridiagonal_eigenvalues(d, e))


# --- Lanczos known answers of the reference's tests ---------------------------

def test_lanczos_3x3_known_answer():
    # src/Bacchus/lanczos/lanczos.c:295-352
    M = np.array([[1.0, 2, 3], [2, 1, 2], [3, 2, 1]])
    ev, alpha, beta, k = O.lanczos(3, lambda x: M @ x, 3, convergence=O.CONVERGE_NONE,
                                   num_eigenvalues=3)
    assert k == 3
    want = [-2.0, 2.5 - np.sqrt(41) / 2, 2.5 + np.sqrt(41) / 2]
    assert np.allclose(ev, want, atol=1e-5)
    assert np.allclose(ev, want, atol=1e-12)


def test_lanczos_50x50_random_against_dense():
    # src/Bacchus/lanczos/lanczos.c:354-404 (seeded here)
    rng = np.random.default_rng(50)
    A = rng.uniform(-1, 1, (50, 50))
    M = (A + A.T) / 2
    ev, alpha, beta, k = O.lanczos(50, lambda x: M @ x, 50, reorth=O.REORTH_FULL,
                                   convergence=O.CONVERGE_NONE, num_eigenvalues=50)
    assert k == 50
    assert np.abs(ev - np.linalg.eigvalsh(M)).max() < 1e-10


# --- brute force: the Hamiltonian the index lists must encode ------------------

@pytest.mark.parametrize("cfg", [
    SystemConfig(2, 2, 2, max_body=2, seed=3),
    SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3),
    SystemConfig(3, 3, 0, max_body=3, seed=7, sigma3=0.5),
    SystemConfig(3, 2, 1, max_body=3, seed=9, sigma3=0.5, twice_m=1),
    SystemConfig(2, 3, 1, max_body=3, seed=13, sigma3=0.5, twice_m=-1),
])
def test_index_lists_encode_brute_force_hamiltonian(cfg):
    inp = build(cfg)
    order = O.basis_order_of(inp)
    # basis enumeration: bit for bit the set an independent enumeration gives
    assert len(order) == len(set(order)) == inp.dimension
    assert set(order) == O.enumerate_basis(inp.orbitals, cfg.num_protons, cfg.num_neutrons,
                                           cfg.nmax, cfg.twice_m)
    H_bf = O.brute_force_hamiltonian(inp, order)
    H = dense(O.Problem.from_inputs(inp))
    assert np.abs(H_bf - H_bf.T).max() == 0.0
    assert np.abs(H - H_bf).max() < 1e-12 * np.abs(H_bf).max()
    assert np.count_nonzero(H_bf) > inp.dimension  # off-diagonal structure exercised


def test_planted_phase_error_is_caught():
    """A wrong phase in one index-list entry must break agreement with the
    brute force (the mistake is planted in a copy of the oracle's input)."""
    inp = build(SystemConfig(2, 2, 2, max_body=2, seed=3))
    order = O.basis_order_of(inp)
    H_bf = O.brute_force_hamiltonian(inp, order)
    p = O.Problem.from_inputs(inp)
    ins = p.instructions
    k = int(np.nonzero(ins[:, 0] == 3)[0][0])
    entry = int(p.list_offsets[ins[k, 4]])
    p.list_entries = p.list_entries.copy()
    p.list_entries[3 * entry + 2] ^= np.int32(-2 ** 31)
    assert np.abs(dense(p) - H_bf).max() > 1e-3


# --- the compiled reference's outputs ------------------------------------------

@pytest.mark.parametrize("name", GOLDEN)
def test_generator_reproduces_golden_inputs(name, golden_dir):
    g = np.load(os.path.join(golden_dir, f"{name}.npz"))
    inp = build(CONFIGS[name])
    assert inp.dimension == len(g["x"])
    from tests.golden.make_golden import digest
    ints, floats = digest(inp)
    assert np.array_equal(ints, g["digest_int"])
    assert np.allclose(floats, g["digest_float"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_matvec_matches_compiled_reference(name, golden_dir):
    g = np.load(os.path.join(golden_dir, f"{name}.npz"))
    inp = build(CONFIGS[name])
    y = O.matvec(O.Problem.from_inputs(inp), g["x"])
    assert np.abs(y - g["y_ref"]).max() <= 1e-12 * np.abs(g["y_ref"]).max()


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_lanczos_matches_compiled_reference(name, golden_dir):
    g = np.load(os.path.join(golden_dir, f"{name}.npz"))
    inp = build(CONFIGS[name])
    p = O.Problem.from_inputs(inp)
    iters = int(g["iterations_fixed"])
    ev, *_ , k = O.lanczos(inp.dimension, lambda v: O.matvec(p, v), iters,
                           convergence=O.CONVERGE_NONE, num_eigenvalues=iters)
    ref = g["eig_ref_fixed"]
    assert k == len(ref)
    assert np.abs(ev - ref).max() <= 1e-8 * np.abs(ref).max()
    ev2, *_, k2 = O.lanczos(inp.dimension, lambda v: O.matvec(p, v), 200,
                            convergence=O.CONVERGE_EIGENVALUES, tolerance=1e-7,
                            num_eigenvalues=200)
    assert k2 == int(g["krylov_conv"])
    assert np.abs(ev2 - g["eig_ref_conv"]).max() <= 1e-8 * np.abs(ref).max()


def test_c1_lowest_eigenvalues_against_dense(c1_inputs):
    p = O.Problem.from_inputs(c1_inputs)
    H = dense(p)
    w = np.linalg.eigvalsh(H)
    ev, *_ = O.lanczos(c1_inputs.dimension, lambda v: O.matvec(p, v), 200,
                       start=O.START_GIVEN,
                       start_vector=np.ones(c1_inputs.dimension) / np.sqrt(c1_inputs.dimension),
                       reorth=O.REORTH_FULL, num_eigenvalues=5)
    assert np.abs(ev - w[:5]).max() < 1e-9 * abs(w[0])


@pytest.mark.parametrize("scale", SK.SCALES)
@pytest.mark.parametrize("case", ["N", "P", "NP"])
def test_oracle_tiny_element_skip(case, scale):
    """The oracle skips N and P elements with |s V| < 1e-12 and keeps
    exactly 1e-12 (matrix_vector_multiplication.c:50-53, :105-108, on the
    Vulcan-combined element s V); NP skips nothing (:134-210).  Against
    products written out by hand, elements straddling 1e-12 / s."""
    arrays, np_, nn = SK.problem(case)
    prob = O.Problem(*arrays)
    X = 1e12 * (1.0 + np.random.default_rng(5).random((np_, nn)))
    want = SK.expected(case, X, scale)
    y = O.matvec(prob, X.ravel(), scales=[1.0, scale, 1.0])
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    if case != "NP":
        # the scale decides: some element is kept at one scale and skipped
        # at another
        assert len(SK.kept(case, scale)) != len(SK.kept(case, 1.0)) or scale == 1.0
        assert len(SK.kept(case, scale)) < len(SK.ELEMENTS)
