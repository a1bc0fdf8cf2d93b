"""Python bindings of the CPU oracle (``liboracle.so``, a C restatement of the
reference's Minerva matvec and Bacchus Lanczos) and of the brute-force
second-quantized Hamiltonian used to pin it.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker, never by the
product path.
"""
from __future__ import annotations

import ctypes
import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_MATVEC_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double))


def library() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle` or __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    lib.oracle_matvec.restype = ctypes.c_int
    lib.oracle_matvec.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp,
                                  ctypes.c_int]
    lib.oracle_tridiagonal_eigenvalues.restype = None
    lib.oracle_tridiagonal_eigenvalues.argtypes = [i64, vp, vp, i64, vp]
    lib.oracle_lanczos.restype = i64
    lib.oracle_lanczos.argtypes = [i64, _MATVEC_FN, vp, i64, ctypes.c_int, vp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, i64, vp, vp, vp]
    _LIB = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Problem:
    """Contiguous numpy arrays describing one operator (the in-memory form of
    comb.txt + evaluation order + index lists + element files)."""

    def __init__(self, block_np, block_nn, list_offsets, list_entries, matrix_offsets,
                 matrix_dn, matrix_force, matrix_elements, instructions):
        c = np.ascontiguousarray
        self.block_np = c(block_np, dtype=np.int64)
        self.block_nn = c(block_nn, dtype=np.int64)
        self.list_offsets = c(list_offsets, dtype=np.int64)
        self.list_entries = c(list_entries, dtype=np.int32).reshape(-1)
        self.matrix_offsets = c(matrix_offsets, dtype=np.int64)
        self.matrix_dn = c(matrix_dn, dtype=np.int64)
        self.matrix_force = c(matrix_force, dtype=np.int32)
        self.matrix_elements = c(matrix_elements, dtype=np.float64)
        self.instructions = c(instructions, dtype=np.int32).reshape(-1, 8)
        self.dimension = int((self.block_np * self.block_nn).sum())

    @classmethod
    def from_inputs(cls, inputs):
        ins = inputs.instructions
        raw = np.stack([ins[f] for f in ins.dtype.names], axis=1).astype(np.int32)
        return cls(inputs.blocks[:, 4], inputs.blocks[:, 5], inputs.list_offsets,
                   inputs.list_entries, inputs.matrix_offsets, inputs.matrix_shapes[:, 0],
                   inputs.matrix_shapes[:, 2], inputs.matrix_elements, raw)


def matvec(problem: Problem, x, y=None, scales=None, threads=None) -> np.ndarray:
    """y += H x with the oracle (y defaults to zeros)."""
    lib = library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    if y is None:
        y = np.zeros(problem.dimension)
    sc = None if scales is None else np.ascontiguousarray(scales, dtype=np.float64)
    threads = threads or min(8, os.cpu_count() or 1)
    rc = lib.oracle_matvec(len(problem.block_np), _p(problem.block_np), _p(problem.block_nn),
                           _p(problem.list_offsets), _p(problem.list_entries),
                           _p(problem.matrix_offsets), _p(problem.matrix_dn),
                           _p(problem.matrix_force), _p(sc), _p(problem.matrix_elements),
                           len(problem.instructions), _p(problem.instructions), _p(x), _p(y),
                           threads)
    if rc != 0:
        raise MemoryError("oracle_matvec failed")
    return y


def tridiagonal_eigenvalues(diag, off, count=None) -> np.ndarray:
    lib = library()
    d = np.ascontiguousarray(diag, dtype=np.float64)
    e = np.ascontiguousarray(off, dtype=np.float64)
    if len(e) < max(len(d) - 1, 1):
        e = np.concatenate([e, np.zeros(max(len(d) - 1, 1) - len(e))])
    count = len(d) if count is None else count
    out = np.zeros(count)
    lib.oracle_tridiagonal_eigenvalues(len(d), _p(d), _p(e), count, _p(out))
    return out


START_E0, START_GIVEN = 0, 2
REORTH_REFERENCE, REORTH_FULL, REORTH_NONE = 0, 1, 2
CONVERGE_EIGENVALUES, CONVERGE_NONE = 0, 2


def lanczos(dimension, apply, max_iterations, *, start=START_E0, start_vector=None,
            reorth=REORTH_REFERENCE, convergence=CONVERGE_NONE, tolerance=1e-7,
            num_eigenvalues=5):
    """Oracle Lanczos.  ``apply(x) -> H x`` (numpy).  Returns (eigenvalues,
    alpha, beta, krylov_dimension)."""
    lib = library()

    def cb(_ctx, xp, yp):
        x = np.ctypeslib.as_array(xp, shape=(dimension,))
        y = np.ctypeslib.as_array(yp, shape=(dimension,))
        y[:] = apply(x.copy())

    fn = _MATVEC_FN(cb)
    kmax = min(dimension, max_iterations)
    alpha = np.zeros(kmax)
    beta = np.zeros(kmax)
    ev = np.zeros(num_eigenvalues)
    sv = None if start_vector is None else np.ascontiguousarray(start_vector, dtype=np.float64)
    k = lib.oracle_lanczos(dimension, fn, None, max_iterations, start, _p(sv), reorth,
                           convergence, tolerance, num_eigenvalues, _p(ev), _p(alpha), _p(beta))
    if k < 0:
        raise MemoryError("oracle_lanczos failed")
    return ev[:min(k, num_eigenvalues)], alpha[:k], beta[:k], int(k)


# ---------------------------------------------------------------------------
# Brute force: the synthetic Hamiltonian built directly in second
# quantization on bit-string determinants (protons in bits 0..n-1, neutrons
# in bits n..2n-1), independently of index lists and element blocks.
# ---------------------------------------------------------------------------

def enumerate_basis(orbitals, num_protons, num_neutrons, nmax, twice_m=0):
    """All (proton tuple, neutron tuple) with total 2M, the parity of the
    lowest filling and total HO quanta <= lowest + nmax, as a set."""
    e = orbitals[:, 0]
    m2 = orbitals[:, 3]
    order = np.argsort(e, kind="stable")
    emin_p = int(e[order[:num_protons]].sum())
    emin_n = int(e[order[:num_neutrons]].sum())
    emax = emin_p + emin_n + nmax
    parity = (emin_p + emin_n) & 1
    norb = len(e)

    def species(k, bound):
        out = {}
        for t in itertools.combinations(range(norb), k):
            en = int(e[list(t)].sum())
            if en <= bound:
                out.setdefault((en, int(m2[list(t)].sum())), []).append(t)
        return out

    sp = species(num_protons, emax - emin_n)
    sn = species(num_neutrons, emax - emin_p)
    basis = set()
    for (ep, mp), ps in sp.items():
        for (en, mn), ns in sn.items():
            if ep + en <= emax and ((ep + en) & 1) == parity and mp + mn == twice_m:
                for p in ps:
                    for n in ns:
                        basis.add((p, n))
    return basis


def _apply_ops(bits, ops):
    """Apply a list of (kind, bit) operators, rightmost first ('c' = create,
    'a' = annihilate).  Returns (new_bits, sign) or (None, 0)."""
    sign = 1
    for kind, b in reversed(ops):
        below = bin(bits & ((1 << b) - 1)).count("1")
        if kind == "a":
            if not bits >> b & 1:
                return None, 0
            bits &= ~(1 << b)
        else:
            if bits >> b & 1:
                return None, 0
            bits |= 1 << b
        if below & 1:
            sign = -sign
    return bits, sign


def brute_force_hamiltonian(inputs, basis_order):
    """Dense H over ``basis_order`` (list of (proton tuple, neutron tuple)).

    H = sum of k-body terms a+_{A} a_{C} with species-blocked operator order
    a+_{Ap..} a+_{An..} a_{Cn..} a_{Cp..} (protons outermost), ascending
    tuples, annihilation applied lowest first and creation highest first;
    elements from the seeded generator keyed by global orbital indices."""
    cfg = inputs.config
    norb = len(inputs.orbitals)
    index = {s: i for i, s in enumerate(basis_order)}
    D = len(basis_order)
    H = np.zeros((D, D))
    zp, zn = cfg.num_protons, cfg.num_neutrons
    ranks = [(1, 0), (0, 1)] if cfg.include_one_body else []
    if cfg.max_body >= 2:
        ranks += [(2, 0), (0, 2), (1, 1)]
    if cfg.max_body >= 3:
        ranks += [(3, 0), (0, 3), (2, 1), (1, 2)]
    eorb = inputs.orbitals[:, 0]
    m2 = inputs.orbitals[:, 3]
    elem_cache = {}
    # k-orbital tuples grouped by total 2M, with their HO quanta
    by_m = {}
    for k in range(0, 4):
        groups = {}
        for t in itertools.combinations(range(norb), k):
            groups.setdefault(int(m2[list(t)].sum()), []).append(
                (t, int(eorb[list(t)].sum())))
        by_m[k] = groups
    emax = max(int(sum(eorb[list(p)]) + sum(eorb[list(n)])) for p, n in basis_order)

    def element(bra, ket):
        key = (bra, ket)
        v = elem_cache.get(key)
        if v is None:
            v = inputs.element(bra, ket)
            elem_cache[key] = v
        return v

    for col, (pt, nt) in enumerate(basis_order):
        bits = sum(1 << p for p in pt) | sum(1 << (norb + n) for n in nt)
        for kp, kn in ranks:
            if kp > zp or kn > zn:
                continue
            for cp in itertools.combinations(pt, kp):
                for cn in itertools.combinations(nt, kn):
                    ec = int(eorb[list(cp)].sum() + eorb[list(cn)].sum())
                    mc = int(m2[list(cp)].sum() + m2[list(cn)].sum())
                    ket = tuple(2 * p for p in cp) + tuple(2 * n + 1 for n in cn)
                    rest_p = [p for p in pt if p not in cp]
                    rest_n = [n for n in nt if n not in cn]
                    e_rest = int(eorb[rest_p].sum() + eorb[rest_n].sum()) if (rest_p or rest_n) else 0
                    for mp_, plist in by_m[kp].items():
                        nlist = by_m[kn].get(mc - mp_)
                        if not nlist:
                            continue
                        for ap, eap in plist:
                            if e_rest + eap > emax or set(ap) & set(rest_p):
                                continue
                            for an, ean in nlist:
                                ea = eap + ean
                                if e_rest + ea > emax or (ea - ec) & 1:
                                    continue
                                if set(an) & set(rest_n):
                                    continue
                                out_state = (tuple(sorted(rest_p + list(ap))),
                                             tuple(sorted(rest_n + list(an))))
                                row = index.get(out_state)
                                if row is None:
                                    continue
                                bra = tuple(2 * p for p in ap) + tuple(2 * n + 1 for n in an)
                                v = element(bra, ket)
                                if v == 0.0:
                                    continue
                                ops = ([("c", p) for p in ap] + [("c", norb + n) for n in an]
                                       + [("a", norb + n) for n in reversed(cn)]
                                       + [("a", p) for p in reversed(cp)])
                                nb, s = _apply_ops(bits, ops)
                                if nb is None:
                                    continue
                                H[row, col] += s * v
    return H


def basis_order_of(inputs):
    """(proton tuple, neutron tuple) of every vector element in storage
    order: block by block, element n + Nn*p."""
    order = []
    for b in range(len(inputs.blocks)):
        P = [tuple(int(v) for v in r) for r in inputs.subbasis(b, 0)]
        N = [tuple(int(v) for v in r) for r in inputs.subbasis(b, 1)]
        for p in P:
            for n in N:
                order.append((p, n))
    return order
