"""Host-side checks: the generator's sizes and file formats, and that the
native libraries load and export every symbol their headers declare (no GPU
calls)."""
import os
import re
import tempfile

import numpy as np
import pytest

from jupiterncsm_amd import bacchus
from jupiterncsm_amd.inputs import SYSTEMS, SystemConfig, build, library as inputs_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:bacchus_hip|ncsm)_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [
    ("bacchus_hip.h", "libbacchus_hip.so"),
    ("ncsm_inputs.h", "libncsm_inputs.so"),
])
def test_library_exports_every_declared_symbol(header, lib):
    import ctypes
    path = os.path.join(ROOT, "jupiterncsm_amd", lib)
    handle = ctypes.CDLL(path)
    names = declared_functions(header)
    assert names
    for name in names:
        assert hasattr(handle, name), f"{lib} does not export {name}"
    if header == "bacchus_hip.h":
        assert set(names) == set(bacchus.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("name,dim", [("he4_nmax4", 952), ("he4_nmax6", 7916),
                                      ("he4_nmax8", 44838)])
def test_generated_dimensions(name, dim):
    inp = build(SYSTEMS[name])
    assert inp.dimension == dim == int((inp.blocks[:, 4] * inp.blocks[:, 5]).sum())
    # every instruction's operator conserves 2M: blocks have total 2M = 0
    assert np.all(inp.blocks[:, 1] + inp.blocks[:, 3] == 0)


def test_written_files_follow_reference_layouts():
    inp = build(SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3))
    with tempfile.TemporaryDirectory() as d:
        inp.write(d)
        comb = open(os.path.join(d, "comb.txt")).read().splitlines()
        titles = [r for r in comb if "***" in r]
        for t in ("mp-states", "Conn lists", "Matrix-elements V (cross p-n)",
                  "Matrix-elements V (same p/n)"):
            assert sum(t in r for r in titles) == 1
        nb, nl = len(inp.blocks), len(inp.list_offsets) - 1
        mp_rows = [r.split() for r in comb if "ARRAYMP:" in r]
        assert [int(w[9].rstrip("=")) for w in mp_rows] == list(range(1, nb + 1))
        assert [(int(w[4]), int(w[5])) for w in mp_rows] == [tuple(b) for b in inp.blocks[:, 4:6]]
        calc = [r for r in comb if "CALCBLOCK:" in r]
        assert len(calc) == len(inp.instructions)
        # the multiplication count sits two words before CALCBLOCK:
        mults = [int(r.split("#")[0].split()[-1]) for r in calc]
        assert mults == list(inp.multiplications())
        order = open(os.path.join(d, "evaluation_order")).read().splitlines()
        assert len(order) == len(inp.instructions)
        for row, ins in zip(order, inp.instructions):
            w = row.split()
            assert w[0] == "BLOCK:" and len(w) == (6 if ins["type"] == 3 else 5)
            assert int(w[1]) == ins["block_in"] + 1 and int(w[2]) == ins["block_out"] + 1
        for l in range(nl):
            raw = np.fromfile(os.path.join(d, "index_lists", f"index_list_{nb + 1 + l}"),
                              dtype=np.int32)
            a, b = inp.list_offsets[l], inp.list_offsets[l + 1]
            assert np.array_equal(raw.reshape(-1, 3), inp.list_entries[a:b])
        for v in range(len(inp.matrix_offsets) - 1):
            raw = open(os.path.join(d, "matrix_elements",
                                    f"{nb + nl + 1 + v}_matrix_elements"), "rb").read()
            dn, dp = np.frombuffer(raw[:16], dtype=np.uint64)
            n = dp if dn == 0 else (dn if dp == 0 else dn * dp)
            els = np.frombuffer(raw[16:], dtype=np.float64)
            assert len(els) == n
            assert np.array_equal(els, inp.matrix_elements[inp.matrix_offsets[v]:
                                                           inp.matrix_offsets[v + 1]])


def test_file_loader_reports_missing_files_without_gpu():
    with pytest.raises(bacchus.BacchusError, match=r"Error in .*Could not open"):
        bacchus.Operator.load("/nonexistent/comb.txt", 2, 2, "/nonexistent/order",
                              "/nonexistent", "/nonexistent")


def test_invalid_problem_is_rejected_on_the_host():
    inp = build(SystemConfig(2, 2, 2, max_body=2, seed=3))
    ins = inp.instructions
    raw = np.stack([ins[f] for f in ins.dtype.names], axis=1).astype(np.int32)
    entries = inp.list_entries.copy()
    entries[0, 1] = 10 ** 6  # out index beyond its sub-basis
    with pytest.raises(bacchus.BacchusError, match="indexes beyond"):
        bacchus.Operator.from_arrays(inp.blocks[:, 4], inp.blocks[:, 5], inp.list_offsets,
                                     entries, inp.matrix_offsets, inp.matrix_shapes[:, 0],
                                     inp.matrix_shapes[:, 2], inp.matrix_elements, raw)
    bad = raw.copy()
    bad[0, 6] = 10 ** 6  # missing matrix block
    with pytest.raises(bacchus.BacchusError, match="missing block"):
        bacchus.Operator.from_arrays(inp.blocks[:, 4], inp.blocks[:, 5], inp.list_offsets,
                                     inp.list_entries, inp.matrix_offsets,
                                     inp.matrix_shapes[:, 0], inp.matrix_shapes[:, 2],
                                     inp.matrix_elements, bad)


def test_element_generator_symmetry_and_selection_rules():
    inp = build(SystemConfig(2, 2, 4, max_body=3, seed=1))
    orb = inp.orbitals
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(2000):
        bra = tuple(int(v) for v in rng.choice(2 * len(orb), 2, replace=False))
        ket = tuple(int(v) for v in rng.choice(2 * len(orb), 2, replace=False))
        v = inp.element(bra, ket)
        assert v == inp.element(ket, bra)
        m = lambda t: sum(orb[i >> 1, 3] for i in t)
        e = lambda t: sum(orb[i >> 1, 0] for i in t)
        p = lambda t: sum(1 - (i & 1) for i in t)
        if m(bra) != m(ket) or (e(bra) - e(ket)) % 2 or p(bra) != p(ket):
            assert v == 0.0
        else:
            checked += 1
            assert v != 0.0
    assert checked > 10


@pytest.mark.parametrize("n", [1, 2, 7, 40, 96])
def test_symmetric_eigen_matches_lapack(n):
    """The block solver's host Jacobi eigensolver against numpy's LAPACK
    eigh, plus the closed form of the tridiagonal Toeplitz matrix
    (a on the diagonal, b beside it: a + 2 b cos(j pi / (n + 1)))."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, V = bacchus.symmetric_eigen(A)
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(w - ref).max() <= 1e-13 * scale
    assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-13
    assert np.abs(A @ V - V * w).max() <= 1e-12 * scale
    a, b = 2.5, -0.75
    T = np.diag(np.full(n, a)) + np.diag(np.full(n - 1, b), 1) + np.diag(np.full(n - 1, b), -1)
    wt, _ = bacchus.symmetric_eigen(T)
    closed = np.sort(a + 2 * b * np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))
    assert np.abs(wt - closed).max() <= 1e-13


def test_symmetric_eigen_orders_and_symmetrizes():
    """Ascending order with repeated eigenvalues, and only the symmetric
    part of a non-symmetric input is used."""
    A = np.diag([3.0, 1.0, 1.0, -2.0])
    A[0, 1] = 1e-3            # the solver uses (A + A^T) / 2
    w, V = bacchus.symmetric_eigen(A)
    ref = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - ref).max() <= 1e-14
    with pytest.raises(bacchus.BacchusError):
        bacchus.symmetric_eigen(np.zeros((2, 3)))


def test_c_driver_links_and_reports_errors_without_gpu():
    """The plain C caller (tests/c_abi/bacchus_driver.c, built with the
    library) links against libbacchus_hip and the HIP runtime, prints its
    usage on bad arguments (exit 2) and passes the library's error message
    through for a missing comb.txt (exit 1) - no GPU call on either path."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "c_abi", "bacchus_driver")
    assert os.path.exists(exe), "build it with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    with tempfile.TemporaryDirectory() as d:
        missing = os.path.join(d, "comb.txt")
        r = subprocess.run([exe, missing, "2", "2", os.path.join(d, "order"), d, d, d, d, "10"],
                           capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Error in" in r.stderr and "comb.txt" in r.stderr


@pytest.mark.parametrize("cfg", [SystemConfig(2, 2, 2, max_body=3, seed=5, sigma3=0.3),
                                 SystemConfig(3, 3, 2, max_body=3, seed=11, sigma3=0.1)])
def test_three_body_parity_systems_exercise_paired_lanes(cfg):
    """The 3-body systems of the FP64 parity tests hold NP lane lists that the
    FP64 operator stores as expanded element rows (list length at most twice
    its distinct elements, three times for lane lists of 3-body
    instructions; the lane side is the longer list) with input runs of odd
    length, so those tests run the paired-lane loop and its zero padding
    entries, in both directions of off-diagonal instructions."""
    inp = build(cfg)
    ins = inp.instructions
    lo = inp.list_offsets
    le = inp.list_entries.reshape(-1, 3)
    lens = np.diff(lo)
    expanded_np = odd_runs = 0
    np_ins = [I for I in ins[ins["type"] == 3]
              if lens[I["neutron_list"]] > 0 and lens[I["proton_list"]] > 0]

    def lane_of(I):
        ln, lp = lens[I["neutron_list"]], lens[I["proton_list"]]
        return I["proton_list"] if lp > ln else I["neutron_list"]
    three_body = {lane_of(I) for I in np_ins if inp.matrix_shapes[I["matrix_block"], 2] == 3}
    for I in np_ins:
        lane = lane_of(I)
        E = le[lo[lane]:lo[lane + 1]]
        if len(E) > (3 if lane in three_body else 2) * len(np.unique(E[:, 2] & 0x7FFFFFFF)):
            continue
        expanded_np += 1
        dirs = (0,) if I["block_in"] == I["block_out"] else (0, 1)
        for d in dirs:
            _, runs = np.unique(E[:, d], return_counts=True)
            odd_runs += int((runs % 2).sum())
    assert expanded_np > 0 and odd_runs > 0, (expanded_np, odd_runs)
