"""Learned Lanczos start for the family H(c1, c2) = H1 + c1 V2 + c2 V3.

Pipeline (no counterpart in the reference, whose Lanczos starts from e_0,
src/Bacchus/lanczos/lanczos.c:60-72):

1. :func:`solve_family` runs the batched device Lanczos over many (c1, c2)
   points at once (one batched matvec per iteration for all of them) and
   returns, per point, its lowest eigenvalues and a *target vector*: the
   ground state (``target="ground"``, sign fixed against a reference vector)
   or the projection of a fixed vector onto the lowest ``m`` eigenvectors
   (``target="band"``, free of the eigenvectors' sign and of their order
   inside the band).
2. :func:`fit` keeps ``n`` of the targets as the stored FP32 basis U, turns
   every target into least-squares coefficients over U, and trains the MLP
   (our BF16-MFMA forward/backward and Adam kernels) from (c1, c2) to those
   coefficients.
3. :meth:`LearnedStart.start_vector` evaluates the MLP and combines the
   basis into a normalized FP64 start vector on the device; the Lanczos
   solve then uses it with ``START_GIVEN``.  A random start is the fallback
   when the prediction is not usable (non-finite or zero).

Device tensors come from torch (memory and streams only); every arithmetic
step on the path runs in the library's HIP kernels.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import bacchus


def family_points(count, seed, low=0.8, high=1.2):
    """(count, 2) uniform (c1, c2) draws."""
    return np.random.default_rng(seed).uniform(low, high, size=(count, 2))


def scales_of(points):
    s = np.ones((len(points), 3))
    s[:, 1:] = points
    return s


@dataclass
class FamilySolves:
    points: np.ndarray          # (P, 2)
    eigenvalues: np.ndarray     # (P, nev)
    targets: object             # torch float32 (P, D), unit norm
    iterations: int
    seconds: float
    ritz: object = None         # torch float32 (keep_ritz * band, D): the
                                # band's Ritz vectors of the first points
    band_targets: list = None   # band targets of further fixed vectors,
                                # one torch float32 (P, D) per vector


def solve_family(operator, points, *, iterations=80, num_eigenvalues=5, batch=64, target="ground",
                 band=5, reference=None, seed=1, keep_targets=True, tolerance=0.0, keep_ritz=0,
                 band_references=0):
    """Batched Lanczos over the points with an FP32/BF16 operator.

    reference: float64 device vector (D,) that fixes the ground state's
    sign / is projected onto the band.  Default: the ground state of the
    first point (ground target; its overlap with every other ground state of
    a smooth family stays far from zero, so the sign is continuous in c),
    or a seeded random vector (band target).  band_references > 0 also
    keeps the band targets of that many further seeded random vectors
    (band_targets): together they span the band at every point."""
    import torch
    D = operator.dimension
    if reference is None and target != "ground":
        g = torch.Generator(device="cuda")
        g.manual_seed(int(seed) + 12345)
        reference = torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
        reference /= reference.norm()
    nritz = 1 if target == "ground" else int(band)
    evs, tg, kept = [], [], []
    refs = []
    if band_references > 0:
        g = torch.Generator(device="cuda")
        g.manual_seed(int(seed) + 54321)
        for _ in range(int(band_references)):
            r = torch.randn(D, dtype=torch.float64, device="cuda", generator=g)
            refs.append(r / r.norm())
    band_tg = [[] for _ in refs]
    t0 = time.time()
    total_it = 0
    for start in range(0, len(points), batch):
        pts = points[start:start + batch]
        B = len(pts)
        with bacchus.Batch([operator], [B]) as bt:
            bt.set_scales(0, scales_of(pts))
            ritz = torch.empty((nritz, D, B), dtype=torch.float64, device="cuda")
            ev, it, _ = bt.lanczos(iterations, num_eigenvalues=num_eigenvalues,
                                   tolerance=tolerance, seed=seed + start, ritz_vectors=[ritz],
                                   num_ritz=nritz)
        total_it += it
        evs.append(ev[0])
        take = min(B, keep_ritz - start) if keep_ritz > start else 0
        if take > 0:
            # (r, D, b) -> (b, r, D): the Ritz vectors of each kept point
            kept.append(ritz[:, :, :take].permute(2, 0, 1).reshape(-1, D)
                        .to(torch.float32).contiguous())
        if keep_targets:
            if reference is None:
                reference = ritz[0][:, 0].clone()
            # overlaps of every Ritz vector with the reference, per instance
            ov = torch.einsum("rdb,d->rb", ritz, reference)
            if target == "ground":
                t = ritz[0] * torch.sign(ov[0])[None, :]
            else:
                t = torch.einsum("rdb,rb->db", ritz, ov)
            t = t / t.norm(dim=0, keepdim=True)
            tg.append(t.T.to(torch.float32).contiguous())
        for k, r in enumerate(refs):
            ov = torch.einsum("rdb,d->rb", ritz, r)
            t = torch.einsum("rdb,rb->db", ritz, ov)
            t = t / t.norm(dim=0, keepdim=True)
            band_tg[k].append(t.T.to(torch.float32).contiguous())
        del ritz
    targets = torch.cat(tg, 0) if keep_targets else None
    ritz_kept = torch.cat(kept, 0) if kept else None
    torch.cuda.synchronize()
    return FamilySolves(np.asarray(points), np.concatenate(evs, 0), targets, total_it,
                        time.time() - t0, ritz_kept,
                        [torch.cat(b, 0) for b in band_tg] if refs else None)


@dataclass
class LearnedStart:
    basis: object               # torch float32 (n, D)
    mlp: bacchus.MLP
    out_mean: np.ndarray        # (n,)
    out_scale: float
    train_loss: float
    target: str
    rank: int = 0               # leading basis rows that are used

    def coefficients(self, points):
        import torch
        x = torch.as_tensor((np.asarray(points, dtype=np.float64).reshape(-1, 2) - 1.0) / 0.2,
                            dtype=torch.float32, device="cuda").contiguous()
        y = self.mlp.forward(x)
        m = torch.as_tensor(self.out_mean, dtype=torch.float32, device="cuda")
        return (y * self.out_scale + m).contiguous()

    def start_vector(self, c1, c2, out=None, mix=0.05, seed=11):
        """Normalized FP64 start vector for H(c1, c2) on the device, or None
        when the prediction is unusable (the caller then starts randomly).
        A random component of relative size `mix` keeps every eigenvector
        in the Krylov space, so a prediction that lands on the wrong state
        (near a level crossing) still converges to the right one."""
        import torch
        c = self.coefficients([[c1, c2]])[0].contiguous()
        if not bool(torch.isfinite(c).all()) or float(c.abs().max()) == 0.0:
            return None
        v = bacchus.basis_combine(self.basis, c, out)
        if mix > 0:
            g = torch.Generator(device="cuda")
            g.manual_seed(seed)
            r = torch.randn(v.shape, dtype=torch.float64, device="cuda", generator=g)
            v += (mix / float(r.norm())) * r
            v /= v.norm()
        return v


def orthonormal_basis(U, threshold=1e-12):
    """Orthonormal combinations of the stored vectors U (n, D): from the
    Gram matrix G = V diag(w) V^T, rows w_j^-1/2 V_j^T U for the w_j above
    threshold * max(w) (directions below it are noise in FP32 storage);
    padded with zero rows to n rows.  Returns (Q (n, D), rank)."""
    n = U.shape[0]
    G = bacchus.basis_gram(U)
    w, V = np.linalg.eigh(0.5 * (G + G.T))
    keep = w > threshold * w.max()
    C = np.zeros((n, n))
    C[:keep.sum()] = (V[:, keep] / np.sqrt(w[keep])).T[::-1]
    return bacchus.basis_transform(U, C), int(keep.sum())


def fit(solves: FamilySolves, *, basis_size=64, hidden=256, depth=3, steps=3000, lr=2e-3,
        seed=1, target="ground", log=None, precision=0):
    """Stored basis: the orthonormalized span of the first basis_size
    targets (a fixed set of basis_size FP32 vectors, zero rows past the
    numerical rank).  The MLP is trained on all points to the coordinates
    of their targets in that basis: bounded, smooth in (c1, c2), and an
    error in them is the same error in the start vector."""
    import torch
    T = solves.targets
    # eigenvector continuation: the individual Ritz vectors of the first
    # points span a band's eigenvectors at nearby (c1, c2) far better than
    # the band targets themselves do
    src = solves.ritz if solves.ritz is not None else T
    U, rank = orthonormal_basis(src[:basis_size].contiguous())
    n = U.shape[0]
    if log:
        log(f"basis rank {rank} of {n}")
    mlp, mean, scale, loss = _train_coefficients(U, T, solves.points, hidden, depth, steps, lr,
                                                 seed, log, precision)
    return LearnedStart(U, mlp, mean, scale, loss, target, rank)


def _train_coefficients(U, T, points, hidden, depth, steps, lr, seed, log, precision=0):
    """An MLP (c1, c2) -> coordinates of the targets T (P, D) over the
    stored basis U, trained on the device; returns (mlp, mean, scale,
    final loss) with the outputs standardized as (Z - mean) / scale.
    precision: bacchus.MLP precision flags of the network's GEMMs."""
    import torch
    n = U.shape[0]
    Z = []
    for s in range(0, T.shape[0], 64):
        V = T[s:s + 64].to(torch.float64).T.contiguous()       # (D, b)
        Z.append(bacchus.basis_project(U, V).T)
    Z = np.concatenate(Z, 0)                                     # (P, n)
    mean = Z.mean(0)
    scale = float(np.abs(Z - mean).max()) or 1.0
    X = (np.asarray(points) - 1.0) / 0.2
    xd = torch.as_tensor(X, dtype=torch.float32, device="cuda").contiguous()
    td = torch.as_tensor((Z - mean) / scale, dtype=torch.float32, device="cuda").contiguous()
    mlp = bacchus.MLP(2, hidden, depth, n, seed=seed, precision=precision)
    loss = None
    for k in range(steps):
        want = (k + 1) % 500 == 0 or k + 1 == steps
        # cosine decay to 1% of lr
        step_lr = lr * (0.01 + 0.99 * 0.5 * (1.0 + np.cos(np.pi * k / steps)))
        loss = mlp.train_step(xd, td, lr=step_lr, want_loss=want) if want else \
            mlp.train_step(xd, td, lr=step_lr)
        if want and log:
            log(f"step {k + 1} loss {loss:.3e}")
    torch.cuda.synchronize()
    return mlp, mean, scale, float(loss)


@dataclass
class LearnedBlockStart:
    """One MLP per fixed reference vector, each predicting the coordinates
    of that vector's band target over the stored basis: at a new (c1, c2)
    the predictions together span the band and start the block solve."""
    basis: object               # torch float32 (n, D)
    rank: int
    models: list                # (mlp, mean, scale) per reference vector
    train_losses: list

    def coefficients(self, c1, c2):
        """(k, n) FP64 host array of predicted coordinates."""
        import torch
        x = torch.as_tensor([[(c1 - 1.0) / 0.2, (c2 - 1.0) / 0.2]], dtype=torch.float32,
                            device="cuda").contiguous()
        rows = []
        for mlp, mean, scale in self.models:
            y = mlp.forward(x)[0].to(torch.float64).cpu().numpy()
            rows.append(y * scale + mean)
        return np.stack(rows, 0)

    def start_block(self, c1, c2, b=None):
        """Float32 device block (D, b): the predicted band targets,
        combined on the device (b = number of models by default; extra
        columns repeat the predictions in order)."""
        C = self.coefficients(c1, c2)
        b = C.shape[0] if b is None else int(b)
        C = C[np.arange(b) % C.shape[0]]
        if not np.all(np.isfinite(C)):
            return None
        return bacchus.basis_transform(self.basis, C).T.contiguous()


def fit_block(solves: FamilySolves, *, basis_size=64, hidden=256, depth=3, steps=3000, lr=2e-3,
              seed=1, log=None, precision=0):
    """Stored basis as in fit(); one MLP per band_targets entry of the
    solves, each of the shape fit() trains."""
    import torch
    if not solves.band_targets:
        raise ValueError("solve_family(..., band_references=k) keeps the targets this needs")
    src = solves.ritz if solves.ritz is not None else solves.targets
    U, rank = orthonormal_basis(src[:basis_size].contiguous())
    models, losses = [], []
    for k, T in enumerate(solves.band_targets):
        mlp, mean, scale, loss = _train_coefficients(U, T, solves.points, hidden, depth, steps,
                                                     lr, seed + 101 * k, log, precision)
        models.append((mlp, mean, scale))
        losses.append(loss)
    torch.cuda.synchronize()
    return LearnedBlockStart(U, rank, models, losses)


@dataclass
class ProjectedFamily:
    """H(c1, c2) = H1 + c1 V2 + c2 V3 restricted to the span of the stored
    basis U (rows, FP32): the three parts are projected once,
    P_k = U H_k U^T, and for any (c1, c2) the Rayleigh-Ritz problem
    (P_1 + c1 P_2 + c2 P_3) s = theta N s, N = U U^T, is an n x n FP64
    eigenproblem on the host.  Because H is linear in (c1, c2) this is
    exact within the span; its Ritz vectors (eigenvector continuation over
    the stored solutions) start the block solve."""
    basis: object               # torch float32 (n, D)
    rank: int                   # leading rows of basis that are used
    parts: np.ndarray           # (3, rank, rank)
    gram: np.ndarray            # (rank, rank)

    @classmethod
    def build(cls, operator, basis, rank=None):
        """Projects H1, V2, V3 of operator (any precision; one FP64
        single-vector matvec per basis row and part) onto basis[:rank];
        the operator's scales are restored afterwards."""
        import torch
        n, D = basis.shape
        rank = n if rank is None else int(rank)
        if D != operator.dimension:
            raise bacchus.BacchusError(f"basis has {D} components, the operator {operator.dimension}")
        U = basis[:rank].contiguous()
        parts = np.zeros((3, rank, rank))
        x = torch.empty(D, dtype=torch.float64, device=basis.device)
        Y = torch.empty((D, rank), dtype=torch.float64, device=basis.device)
        y = torch.empty(D, dtype=torch.float64, device=basis.device)
        previous = operator.scales
        try:
            for k, sc in enumerate(((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))):
                operator.set_scales(*sc)
                for i in range(rank):
                    x.copy_(U[i])
                    operator.matvec(x, y)
                    Y[:, i].copy_(y)
                P = bacchus.basis_project(U, Y)
                parts[k] = 0.5 * (P + P.T)
        finally:
            operator.set_scales(*previous)
        return cls(basis, rank, parts, bacchus.basis_gram(U))

    def ritz(self, c1, c2, k):
        """Lowest k Ritz values and coefficient vectors (rank, k) over the
        basis rows, normalized in the N inner product."""
        # N = L L^T; eigen of L^-1 M L^-T with M = P_1 + c1 P_2 + c2 P_3, so
        # the three L^-1 P_k L^-T are formed once for every point
        cache = getattr(self, "_whitened_parts", None)
        if cache is None:
            Li = np.linalg.inv(np.linalg.cholesky(self.gram))
            cache = (Li, [Li @ P @ Li.T for P in self.parts])
            self._whitened_parts = cache
        Li, (A1, A2, A3) = cache
        A = A1 + c1 * A2 + c2 * A3
        w, Z = np.linalg.eigh(0.5 * (A + A.T))
        return w[:k], Li.T @ Z[:, :k]

    def start_block(self, c1, c2, b):
        """Float32 device block (D, b): the b lowest Ritz vectors of
        H(c1, c2) in the basis span."""
        _, S = self.ritz(c1, c2, b)
        return bacchus.basis_transform(self.basis[:self.rank], S.T).T.contiguous()


def solve_from(operator, c, start_vector, *, num_eigenvalues, targets, tolerance=1e-6,
               max_iterations=300, seed=7):
    """Single-vector Lanczos for H(c) from the given start (random if None)
    until the lowest num_eigenvalues are within tolerance of targets."""
    operator.set_scales(1.0, float(c[0]), float(c[1]))
    kw = dict(num_eigenvalues=num_eigenvalues, reorthogonalization=bacchus.REORTH_FULL,
              convergence=bacchus.CONVERGE_TARGETS, target_eigenvalues=targets,
              tolerance=tolerance, check_every=1)
    if start_vector is None:
        return bacchus.diagonalize(operator, max_iterations, start=bacchus.START_RANDOM,
                                   seed=seed, **kw)
    return bacchus.diagonalize(operator, max_iterations, start=bacchus.START_GIVEN,
                               start_vector=start_vector, **kw)
