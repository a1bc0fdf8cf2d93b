"""The learned Lanczos start on the GPU: stored-basis kernels against numpy,
and what the trained network gains over a random start."""
import numpy as np
import pytest

from jupiterncsm_amd import bacchus
from jupiterncsm_amd import learned as L
from jupiterncsm_amd.inputs import SYSTEMS, build

pytestmark = pytest.mark.gpu


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    return torch


def test_basis_kernels_match_numpy():
    print("case basis kernels", flush=True)
    torch = torch_dev()
    rng = np.random.default_rng(1)
    n, D, B = 7, 5003, 5
    U = rng.standard_normal((n, D)).astype(np.float32)
    V = rng.standard_normal((D, B))
    c = rng.standard_normal(n).astype(np.float32)
    Ud = torch.as_tensor(U, device="cuda")
    G = bacchus.basis_gram(Ud)
    P = bacchus.basis_project(Ud, torch.as_tensor(V, device="cuda"))
    s = bacchus.basis_combine(Ud, torch.as_tensor(c, device="cuda"))
    torch.cuda.synchronize()
    U64 = U.astype(np.float64)
    assert np.allclose(G, U64 @ U64.T, rtol=1e-12, atol=1e-9)
    assert np.allclose(P, U64 @ V, rtol=1e-12, atol=1e-9)
    ref = c.astype(np.float64) @ U64
    ref /= np.linalg.norm(ref)
    assert np.abs(s.cpu().numpy() - ref).max() <= 1e-12


def test_basis_transform_few_and_many_outputs():
    """out = C U on the device: 20 outputs (one pass over U per output) and
    the first 8 of them (one pass over U for all outputs) against numpy in
    FP64, and bit-equal to each other (the same FP64 sums in the same order,
    rounded once to FP32)."""
    print("case basis transform", flush=True)
    torch = torch_dev()
    rng = np.random.default_rng(3)
    n, D = 52, 10007
    U = rng.standard_normal((n, D)).astype(np.float32)
    C = rng.standard_normal((20, n))
    Ud = torch.as_tensor(U, device="cuda")
    many = bacchus.basis_transform(Ud, C).cpu().numpy()
    few = bacchus.basis_transform(Ud, C[:8]).cpu().numpy()
    ref = C @ U.astype(np.float64)
    tol = 2.0 ** -23 * np.abs(ref).max()
    assert np.abs(many - ref).max() <= tol
    assert np.array_equal(few, many[:8])


def test_learned_start_beats_random_start():
    """2+2 Nmax 6 family: 64 batched training solves, 16-vector basis, a
    small MLP trained on the GPU; on held-out points the lowest eigenvalue
    converges to 1e-6 in far fewer iterations than from a random start."""
    print("case learned start he4_nmax6", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax6"])
    op = bacchus.Operator.from_inputs(inp, precision=bacchus.FP32)
    try:
        train = L.family_points(64, 1)
        test = L.family_points(6, 2)
        ref = L.solve_family(op, test, iterations=120, num_eigenvalues=1, keep_targets=False)
        sol = L.solve_family(op, train, iterations=60, num_eigenvalues=1, target="ground")
        model = L.fit(sol, basis_size=16, hidden=64, depth=3, steps=1500, target="ground")
        print("train loss", model.train_loss, flush=True)
        assert model.train_loss < 1e-2
        learned, random = [], []
        for i, c in enumerate(test):
            sv = model.start_vector(*c, mix=0.01)
            assert sv is not None
            assert abs(float(sv.norm()) - 1.0) < 1e-12
            rl = L.solve_from(op, c, sv, num_eigenvalues=1, targets=ref.eigenvalues[i][:1])
            rr = L.solve_from(op, c, None, num_eigenvalues=1, targets=ref.eigenvalues[i][:1])
            learned.append(rl.iterations)
            random.append(rr.iterations)
            err = abs(rl.eigenvalues[0] - ref.eigenvalues[i][0]) / abs(ref.eigenvalues[i][0])
            assert err <= 1e-6
        print("learned", learned, "random", random, flush=True)
        assert max(random) < 300
        assert np.mean(learned) * 2 <= np.mean(random)
    finally:
        op.close()


def test_learned_start_block_beats_random_block():
    """2+2 Nmax 6 family, 3 eigenvalues: band targets of 4 fixed vectors kept
    from 48 batched training solves, one small MLP per vector; at held-out
    points the block Lanczos from the predicted block stops on its residual
    bound, within 1e-6 of the FP64 references, in at most half the block
    steps a random start block takes (60 at most)."""
    print("case learned start block he4_nmax6", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax6"])
    b = 4
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op, \
            bacchus.Operator.from_inputs(inp) as op64, bacchus.Batch([op], [b]) as bt:
        train = L.family_points(48, 1)
        sol = L.solve_family(op, train, iterations=120, num_eigenvalues=3, target="band",
                             band=3, keep_ritz=6, band_references=b)
        assert len(sol.band_targets) == b
        for T in sol.band_targets:
            assert tuple(T.shape) == (48, inp.dimension)
            assert torch.allclose(T.norm(dim=1), torch.ones(48, device="cuda"), atol=1e-5)
        mb = L.fit_block(sol, basis_size=16, hidden=64, depth=3, steps=1500)
        print("losses", mb.train_losses, flush=True)
        assert max(mb.train_losses) < 1e-2
        learned, random = [], []
        for c1, c2 in L.family_points(4, 2):
            op64.set_scales(1.0, c1, c2)
            ref = bacchus.diagonalize(op64, 400, num_eigenvalues=3, start=bacchus.START_RANDOM,
                                      seed=1, reorthogonalization=bacchus.REORTH_FULL,
                                      convergence=bacchus.CONVERGE_NONE).eigenvalues
            bt.set_scales(0, np.tile((1.0, c1, c2), (b, 1)))
            X = mb.start_block(c1, c2)
            assert X is not None and tuple(X.shape) == (inp.dimension, b)
            for start, out in ((X, learned), (None, random)):
                r = bt.block_lanczos(60, 3, start_block=start,
                                     convergence=bacchus.CONVERGE_RESIDUALS, tolerance=1e-6)
                if start is not None:
                    err = np.abs(r.eigenvalues - ref) / np.abs(ref)
                    assert err.max() <= 1e-6, (c1, c2, err)
                out.append(r.blocks)
        print("learned", learned, "random", random, flush=True)
        assert np.mean(learned) * 2 <= np.mean(random)


def test_band_target_projects_onto_lowest_states():
    """The band target of a point is the projection of a fixed vector onto
    its lowest eigenvectors: its Rayleigh quotient lies inside the band."""
    print("case band target", flush=True)
    torch = torch_dev()
    inp = build(SYSTEMS["he4_nmax4"])
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op, \
            bacchus.Operator.from_inputs(inp) as op64:
        pts = L.family_points(3, 5)
        sol = L.solve_family(op, pts, iterations=150, num_eigenvalues=3, target="band", band=3)
        for i, c in enumerate(pts):
            t = sol.targets[i].to(torch.float64).contiguous()
            op64.set_scales(1.0, c[0], c[1])
            ht = torch.empty_like(t)
            op64.matvec(t, ht)
            torch.cuda.synchronize()
            rq = float(t @ ht)
            ev = sol.eigenvalues[i]
            assert ev[0] - 1e-6 * abs(ev[0]) <= rq <= ev[2] + 1e-6 * abs(ev[2])
            # residual lies in the band: |H t|^2 - rq^2 small against the spread
            assert float(ht @ ht) - rq * rq <= (ev[2] - ev[0]) ** 2 + 1e-6 * rq * rq


def test_subspace_matrices_match_dense_projection(tmp_path):
    """Operator and norm matrices over stored vectors (the Ceres step) against
    the oracle's dense H (2+2 Nmax 4), and the .npy file they are saved in."""
    print("case subspace matrices", flush=True)
    torch = torch_dev()
    from jupiterncsm_amd import subspace
    from oracle import oracle as O
    inp = build(SYSTEMS["he4_nmax4"])
    rng = np.random.default_rng(3)
    n = 6
    V = rng.standard_normal((n, inp.dimension)).astype(np.float32)
    Vd = torch.as_tensor(V, device="cuda")
    scales = (1.0, 0.9, 1.1)
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP32) as op:
        M = subspace.subspace_operator(op, Vd, scales)
    N = subspace.norm_matrix(Vd)
    p = O.Problem.from_inputs(inp)
    HV = np.stack([O.matvec(p, v.astype(np.float64), scales=scales) for v in V], 0)
    M_ref = V.astype(np.float64) @ HV.T
    M_ref = 0.5 * (M_ref + M_ref.T)
    assert np.abs(M - M_ref).max() <= 2e-6 * np.abs(M_ref).max()
    assert np.allclose(M, M.T)
    assert np.allclose(N, V.astype(np.float64) @ V.astype(np.float64).T, rtol=1e-12, atol=1e-9)
    path = tmp_path / "subspace_operator.npy"
    subspace.save_numpy_matrix(str(path), M)
    raw = open(path, "rb").read()
    assert raw[:6] == b"\x93NUMPY" and raw[6] == 1 and b"'<f8'" in raw[:128]
    assert np.array_equal(np.load(path, allow_pickle=False), M)


@pytest.mark.parametrize("name", ["he4_nmax4", "he4_nmax2_3b"])
def test_fp64_subspace_matrices_match_oracle(name):
    """The FP64 form of the Ceres step (bacchus_hip_subspace_operator,
    bacchus_hip_norm_matrix): float64 vectors, an FP64 operator with scales
    other than 1, against the oracle's matvecs and numpy's dot products to
    1e-12; the matrices exactly symmetric as the reference fills them."""
    print(f"case FP64 subspace matrices {name}", flush=True)
    torch = torch_dev()
    from jupiterncsm_amd import subspace
    from jupiterncsm_amd.inputs import SystemConfig
    from oracle import oracle as O
    cfg = SYSTEMS.get(name) or SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3)
    inp = build(cfg)
    rng = np.random.default_rng(5)
    n = 7
    V = rng.standard_normal((n, inp.dimension))
    Vd = torch.as_tensor(V, device="cuda")
    scales = (1.0, 0.85, 1.15)
    with bacchus.Operator.from_inputs(inp, precision=bacchus.FP64) as op:
        M = subspace.subspace_operator(op, Vd, scales)
    N = subspace.norm_matrix(Vd)
    p = O.Problem.from_inputs(inp)
    HV = np.stack([O.matvec(p, v, scales=scales) for v in V], 0)
    M_ref = np.empty((n, n))
    for i in range(n):
        for j in range(i, n):
            M_ref[i, j] = M_ref[j, i] = V[j] @ HV[i]
    assert np.abs(M - M_ref).max() <= 1e-12 * np.abs(M_ref).max()
    assert np.array_equal(M, M.T) and np.array_equal(N, N.T)
    assert np.abs(N - V @ V.T).max() <= 1e-12 * np.abs(V @ V.T).max()

