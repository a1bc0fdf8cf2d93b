"""One-block problems for the 1e-12 element skip of N and P instructions
(matrix_vector_multiplication.c:50-53, :105-108; none for NP, :134-210),
with the products written out by hand.

The reference skips an element when |V| < 1e-12 for the element it reads;
for H(c1, c2) that is the Vulcan-combined element s V (vulcan.c:114-207), so
the test is on the scaled value.  The elements straddle 1e-12 / s for the
scales used: at s = 0.8 only 2e-12 is kept, at s = 1 also 1e-12 and
1.2e-12, at s = 1.25 also -0.9e-12.  x of order 1e12 makes every product,
kept or skipped, of order 1.
"""
import numpy as np

ELEMENTS = [5e-13, -5e-13, 2e-12, 1e-12, 1.2e-12, -0.9e-12]
LIST = [(0, 1, 0), (1, 0, 1), (2, 2, 2), (0, 0, 3), (1, 2, 4), (2, 1, 5)]
NP_NEUTRONS = [(0, 1, 0), (1, 0, 1), (2, 2, 0)]
NP_PROTONS = [(0, 0, 0), (1, 1, 0), (0, 1, 1)]
NP_ELEMENTS = [5e-13, -5e-13, 2e-13, 1e-12]   # m_n + 2 m_p
SCALES = [1.0, 0.8, 1.25]


def problem(case, wide_p=False):
    """(arrays for Operator.from_arrays / oracle.Problem, Np, Nn); the only
    element block is 2-body (force class 2)."""
    if case == "N":
        np_, nn, lists, ins, dn, V = 2, 3, [LIST], [1, 1, 0, 0, 0, -1, 0, 0], len(ELEMENTS), ELEMENTS
    elif case == "P":
        np_, nn, lists, ins, dn, V = 3, (40 if wide_p else 5), [LIST], [2, 1, 0, 0, -1, 0, 0, 0], 1, ELEMENTS
    elif case == "NP":
        np_, nn, dn, V = 2, 3, 2, NP_ELEMENTS
        lists, ins = [NP_NEUTRONS, NP_PROTONS], [3, 2, 0, 0, 0, 1, 0, 0]
    else:
        raise ValueError(case)
    offsets = np.cumsum([0] + [len(e) for e in lists])
    entries = np.array([t for e in lists for t in e], dtype=np.int32).reshape(-1, 3)
    arrays = ([np_], [nn], offsets, entries, [0, len(V)], [dn], [2],
              np.asarray(V, dtype=np.float64), np.array([ins]))
    return arrays, np_, nn


def expected(case, X, scale):
    """H x by hand for one block X[p][n] with the 2-body scale `scale`."""
    np_, nn = X.shape
    Y = np.zeros((np_, nn))
    if case == "N":
        for i, o, m in LIST:
            if abs(ELEMENTS[m] * scale) >= 1e-12:
                Y[:, o] += scale * ELEMENTS[m] * X[:, i]
    elif case == "P":
        for i, o, m in LIST:
            if abs(ELEMENTS[m] * scale) >= 1e-12:
                Y[o, :] += scale * ELEMENTS[m] * X[i, :]
    else:
        for ni, no, mn in NP_NEUTRONS:
            for pi, po, mp in NP_PROTONS:
                Y[po, no] += scale * NP_ELEMENTS[mn + 2 * mp] * X[pi, ni]
    return Y.ravel()


def kept(case, scale):
    """Indices of the elements that survive the skip at this scale."""
    if case == "NP":
        return list(range(len(NP_ELEMENTS)))
    return [m for m, v in enumerate(ELEMENTS) if abs(v * scale) >= 1e-12]
