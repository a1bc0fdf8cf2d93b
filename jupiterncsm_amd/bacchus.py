"""Host-side mirror of the reference's Bacchus/Minerva interface over the
gfx950 library ``libbacchus_hip.so`` (C ABI in ``include/bacchus_hip.h``).

Reference interface                                   here
----------------------------------------------------  ---------------------------------
new_generative_matrix / new_scheduler                 :meth:`Operator.load`,
  (src/Bacchus/matrix/matrix.c:52-66,                  :meth:`Operator.from_inputs`
   src/Minerva/scheduler/scheduler.c:62-78)
matrix_vector_multiplication                          :func:`matrix_vector_multiplication`
  (src/Bacchus/matrix/matrix.c:68-96)
run_matrix_vector_multiplication                      :func:`run_matrix_vector_multiplication`
  (src/Minerva/scheduler/scheduler.c:80-145)
free_scheduler / free_matrix                          :meth:`Operator.close`
new_lanczos_environment + diagonalize +               :func:`diagonalize`
  get_eigensystem (src/Bacchus/lanczos/lanczos.c:
  75-96,209-303)

Errors surface as :class:`BacchusError` carrying the library's
"Error in <func> @ <file>:<line>" message (the reference prints the same
shape and exits, src/Utilities/error/error.h:7-13).  There is no CPU
fallback: without the built library every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FP64, FP32, BF16 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16": BF16}
START_E0, START_RANDOM, START_GIVEN = 0, 1, 2
REORTH_REFERENCE, REORTH_FULL, REORTH_NONE = 0, 1, 2
CONVERGE_EIGENVALUES, CONVERGE_EIGENVECTORS, CONVERGE_NONE, CONVERGE_TARGETS, CONVERGE_RESIDUALS = \
    0, 1, 2, 3, 4

EXPORTED_SYMBOLS = (
    "bacchus_hip_operator_create", "bacchus_hip_operator_load", "bacchus_hip_operator_free",
    "bacchus_hip_operator_dimension", "bacchus_hip_operator_set_scales",
    "bacchus_hip_operator_matvec", "bacchus_hip_operator_stats",
    "bacchus_hip_run_matrix_vector_multiplication", "bacchus_hip_lanczos",
    "bacchus_hip_last_error", "bacchus_hip_mlp_create", "bacchus_hip_mlp_free",
    "bacchus_hip_mlp_num_params", "bacchus_hip_mlp_get_params", "bacchus_hip_mlp_set_params",
    "bacchus_hip_mlp_forward", "bacchus_hip_mlp_gradients", "bacchus_hip_mlp_train_step",
    "bacchus_hip_batch_create", "bacchus_hip_batch_free", "bacchus_hip_batch_set_scales",
    "bacchus_hip_batch_matvec", "bacchus_hip_batch_autotune", "bacchus_hip_batch_stats",
    "bacchus_hip_batch_lanczos", "bacchus_hip_basis_gram", "bacchus_hip_basis_project",
    "bacchus_hip_basis_combine", "bacchus_hip_basis_transform",
    "bacchus_hip_write_evaluation_order", "bacchus_hip_block_lanczos",
    "bacchus_hip_symmetric_eigen", "bacchus_hip_operator_load_blocks",
    "bacchus_hip_basis_project_static", "bacchus_hip_norm_matrix",
    "bacchus_hip_subspace_operator", "bacchus_hip_batch_mfma_stats",
    "bacchus_hip_mlp_set_precision",
)


class BacchusError(RuntimeError):
    pass


class _Problem(ctypes.Structure):
    _fields_ = [
        ("num_blocks", ctypes.c_int64), ("block_np", ctypes.c_void_p),
        ("block_nn", ctypes.c_void_p), ("num_lists", ctypes.c_int64),
        ("list_offsets", ctypes.c_void_p), ("list_entries", ctypes.c_void_p),
        ("num_matrix_blocks", ctypes.c_int64), ("matrix_offsets", ctypes.c_void_p),
        ("matrix_dn", ctypes.c_void_p), ("matrix_force", ctypes.c_void_p),
        ("matrix_elements", ctypes.c_void_p), ("num_instructions", ctypes.c_int64),
        ("instructions", ctypes.c_void_p),
    ]


class _LanczosSettings(ctypes.Structure):
    _fields_ = [
        ("max_num_iterations", ctypes.c_int64), ("start", ctypes.c_int32),
        ("reorthogonalization", ctypes.c_int32), ("convergence", ctypes.c_int32),
        ("num_eigenvalues", ctypes.c_int32), ("eigenvalue_tolerance", ctypes.c_double),
        ("seed", ctypes.c_uint64), ("start_vector", ctypes.c_void_p),
        ("target_eigenvalues", ctypes.c_void_p), ("check_every", ctypes.c_int32),
        ("target_eigenvalue", ctypes.c_int32), ("num_ritz_vectors", ctypes.c_int32),
        ("pad", ctypes.c_int32), ("krylov_basis", ctypes.c_void_p),
    ]


class _BatchTuning(ctypes.Structure):
    _fields_ = [("tiles_per_cu", ctypes.c_int32), ("max_tile_vectors", ctypes.c_int32),
                ("workgroups_per_cu", ctypes.c_int32), ("tile_words", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32)]


class _BatchLanczosSettings(ctypes.Structure):
    _fields_ = [("max_num_iterations", ctypes.c_int64), ("num_eigenvalues", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("eigenvalue_tolerance", ctypes.c_double),
                ("check_every", ctypes.c_int32), ("num_ritz_vectors", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("start_vectors", ctypes.c_void_p)]


class _BlockLanczosSettings(ctypes.Structure):
    _fields_ = [("max_num_blocks", ctypes.c_int32), ("num_eigenvalues", ctypes.c_int32),
                ("convergence", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("target_eigenvalues", ctypes.c_void_p),
                ("start_block", ctypes.c_void_p), ("seed", ctypes.c_uint64)]


class _LanczosResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("matvec_seconds", ctypes.c_double),
                ("total_seconds", ctypes.c_double), ("converged", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


def library_path() -> str:
    return os.path.join(_HERE, "libbacchus_hip.so")


def library() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise BacchusError(f"{path} is missing: build it with __graft_entry__.build()")
    # torch bundles its own HIP runtime under the same SONAME
    # (libamdhip64.so.7); loading it first makes this library bind to that
    # one, so device pointers and streams are shared by a single runtime.
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    vp, i64, c_int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.bacchus_hip_operator_create.restype = c_int
    lib.bacchus_hip_operator_create.argtypes = [ctypes.POINTER(_Problem), c_int,
                                                ctypes.POINTER(vp)]
    lib.bacchus_hip_operator_load.restype = c_int
    lib.bacchus_hip_operator_load.argtypes = [ctypes.c_char_p, c_int, c_int, ctypes.c_char_p,
                                              ctypes.c_char_p, ctypes.c_char_p, c_int,
                                              ctypes.POINTER(vp)]
    lib.bacchus_hip_operator_free.restype = None
    lib.bacchus_hip_operator_free.argtypes = [vp]
    lib.bacchus_hip_operator_dimension.restype = i64
    lib.bacchus_hip_operator_dimension.argtypes = [vp]
    lib.bacchus_hip_operator_set_scales.restype = c_int
    lib.bacchus_hip_operator_set_scales.argtypes = [vp, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double]
    lib.bacchus_hip_operator_matvec.restype = c_int
    lib.bacchus_hip_operator_matvec.argtypes = [vp, vp, vp, c_int, vp]
    lib.bacchus_hip_operator_stats.restype = c_int
    lib.bacchus_hip_operator_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(i64)]
    lib.bacchus_hip_run_matrix_vector_multiplication.restype = c_int
    lib.bacchus_hip_run_matrix_vector_multiplication.argtypes = [ctypes.c_char_p,
                                                                 ctypes.c_char_p, vp]
    lib.bacchus_hip_lanczos.restype = c_int
    lib.bacchus_hip_lanczos.argtypes = [vp, ctypes.POINTER(_LanczosSettings), vp, vp, vp, vp,
                                        ctypes.POINTER(_LanczosResult)]
    lib.bacchus_hip_last_error.restype = ctypes.c_char_p
    lib.bacchus_hip_last_error.argtypes = []
    lib.bacchus_hip_batch_create.restype = c_int
    lib.bacchus_hip_batch_create.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(_BatchTuning),
                                             ctypes.POINTER(vp)]
    lib.bacchus_hip_batch_free.restype = None
    lib.bacchus_hip_batch_free.argtypes = [vp]
    lib.bacchus_hip_batch_set_scales.restype = c_int
    lib.bacchus_hip_batch_set_scales.argtypes = [vp, ctypes.c_int32, vp]
    lib.bacchus_hip_batch_matvec.restype = c_int
    lib.bacchus_hip_batch_matvec.argtypes = [vp, vp, vp, vp]
    lib.bacchus_hip_batch_autotune.restype = c_int
    lib.bacchus_hip_batch_autotune.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(_BatchTuning),
                                               ctypes.POINTER(ctypes.c_double)]
    lib.bacchus_hip_batch_stats.restype = c_int
    lib.bacchus_hip_batch_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]
    lib.bacchus_hip_batch_mfma_stats.restype = c_int
    lib.bacchus_hip_batch_mfma_stats.argtypes = [vp, ctypes.POINTER(i64),
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_int32)]
    lib.bacchus_hip_batch_lanczos.restype = c_int
    lib.bacchus_hip_batch_lanczos.argtypes = [vp, ctypes.POINTER(_BatchLanczosSettings), vp, vp,
                                              ctypes.POINTER(_LanczosResult)]
    lib.bacchus_hip_basis_gram.restype = c_int
    lib.bacchus_hip_basis_gram.argtypes = [vp, i64, i64, vp]
    lib.bacchus_hip_basis_project.restype = c_int
    lib.bacchus_hip_basis_project.argtypes = [vp, i64, i64, vp, i64, vp]
    lib.bacchus_hip_basis_project_static.restype = c_int
    lib.bacchus_hip_basis_project_static.argtypes = [vp, i64, i64, vp, i64, vp]
    lib.bacchus_hip_basis_combine.restype = c_int
    lib.bacchus_hip_basis_combine.argtypes = [vp, i64, i64, vp, vp, vp]
    lib.bacchus_hip_write_evaluation_order.restype = c_int
    lib.bacchus_hip_write_evaluation_order.argtypes = [ctypes.c_char_p, c_int, ctypes.c_char_p]
    lib.bacchus_hip_block_lanczos.restype = c_int
    lib.bacchus_hip_block_lanczos.argtypes = [vp, ctypes.POINTER(_BlockLanczosSettings), vp, vp,
                                              vp, ctypes.POINTER(_LanczosResult)]
    lib.bacchus_hip_symmetric_eigen.restype = c_int
    lib.bacchus_hip_symmetric_eigen.argtypes = [i64, vp, vp, vp]
    lib.bacchus_hip_basis_transform.restype = c_int
    lib.bacchus_hip_basis_transform.argtypes = [vp, i64, i64, vp, i64, vp]
    lib.bacchus_hip_norm_matrix.restype = c_int
    lib.bacchus_hip_norm_matrix.argtypes = [vp, i64, i64, vp, vp]
    lib.bacchus_hip_subspace_operator.restype = c_int
    lib.bacchus_hip_subspace_operator.argtypes = [vp, vp, i64, vp, vp]
    _LIB = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise BacchusError(library().bacchus_hip_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Operator:
    """Device-resident Hamiltonian: index lists, element blocks and the
    instruction tables live in HBM from construction until :meth:`close`."""

    def __init__(self, handle: ctypes.c_void_p, precision: int):
        self._h = handle
        self.precision = precision
        self.scales = (1.0, 1.0, 1.0)
        self.dimension = int(library().bacchus_hip_operator_dimension(handle))

    # -- construction ------------------------------------------------------
    @classmethod
    def from_arrays(cls, block_np, block_nn, list_offsets, list_entries, matrix_offsets,
                    matrix_dn, matrix_force, matrix_elements, instructions, precision=FP64):
        lib = library()
        keep = [np.ascontiguousarray(block_np, dtype=np.int64),
                np.ascontiguousarray(block_nn, dtype=np.int64),
                np.ascontiguousarray(list_offsets, dtype=np.int64),
                np.ascontiguousarray(list_entries, dtype=np.int32).reshape(-1),
                np.ascontiguousarray(matrix_offsets, dtype=np.int64),
                np.ascontiguousarray(matrix_dn, dtype=np.int64),
                np.ascontiguousarray(matrix_force, dtype=np.int32),
                np.ascontiguousarray(matrix_elements, dtype=np.float64),
                np.ascontiguousarray(instructions, dtype=np.int32).reshape(-1, 8)]
        p = _Problem(len(keep[0]), _ptr(keep[0]), _ptr(keep[1]), len(keep[2]) - 1,
                     _ptr(keep[2]), _ptr(keep[3]), len(keep[4]) - 1, _ptr(keep[4]),
                     _ptr(keep[5]), _ptr(keep[6]), _ptr(keep[7]), len(keep[8]), _ptr(keep[8]))
        h = ctypes.c_void_p()
        _check(lib.bacchus_hip_operator_create(ctypes.byref(p), precision, ctypes.byref(h)))
        return cls(h, precision)

    @classmethod
    def from_inputs(cls, inputs, precision=FP64):
        """Upload a generated system (:mod:`jupiterncsm_amd.inputs`)."""
        ins = inputs.instructions
        raw = np.stack([ins[f] for f in ins.dtype.names], axis=1).astype(np.int32)
        return cls.from_arrays(inputs.blocks[:, 4], inputs.blocks[:, 5], inputs.list_offsets,
                               inputs.list_entries, inputs.matrix_offsets,
                               inputs.matrix_shapes[:, 0], inputs.matrix_shapes[:, 2],
                               inputs.matrix_elements, raw, precision)

    @classmethod
    def load(cls, combination_table_file, num_protons, num_neutrons, evaluation_order_file,
             index_lists_base_directory, matrix_file_base_directory, precision=FP64):
        """From the reference's files (comb.txt, evaluation order,
        index_list_<id>, <id>_matrix_elements)."""
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.bacchus_hip_operator_load(
            str(combination_table_file).encode(), num_protons, num_neutrons,
            str(evaluation_order_file).encode(), str(index_lists_base_directory).encode(),
            str(matrix_file_base_directory).encode(), precision, ctypes.byref(h)))
        return cls(h, precision)

    # -- use ----------------------------------------------------------------
    @property
    def handle(self):
        if self._h is None or not self._h.value:
            raise BacchusError("operator is closed")
        return self._h

    def set_scales(self, s1=1.0, s2=1.0, s3=1.0):
        """Scale the 1-, 2- and 3-body parts: H1*s1 + V2*s2 + V3*s3."""
        _check(library().bacchus_hip_operator_set_scales(self.handle, s1, s2, s3))
        self.scales = (float(s1), float(s2), float(s3))

    def matvec(self, x, y, accumulate=False, stream=None):
        """y = H x (or y += H x) on FP64 torch tensors on the GPU."""
        import torch
        if x.dtype != torch.float64 or y.dtype != torch.float64:
            raise BacchusError("x and y must be float64")
        if not (x.is_cuda and y.is_cuda and x.is_contiguous() and y.is_contiguous()):
            raise BacchusError("x and y must be contiguous device tensors")
        if x.numel() != self.dimension or y.numel() != self.dimension:
            raise BacchusError(f"vectors must have {self.dimension} elements")
        _check(library().bacchus_hip_operator_matvec(
            self.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
            int(bool(accumulate)), _stream_handle(stream)))
        return y

    def stats(self) -> dict:
        b, f = ctypes.c_double(), ctypes.c_double()
        w = ctypes.c_int64()
        _check(library().bacchus_hip_operator_stats(self.handle, ctypes.byref(b),
                                                    ctypes.byref(f), ctypes.byref(w)))
        return {"algorithmic_bytes": b.value, "flops": f.value, "work_items": w.value}

    def close(self):
        if self._h is not None and self._h.value:
            library().bacchus_hip_operator_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def matrix_vector_multiplication(result, matrix: Operator, vector, stream=None):
    """result = H vector (src/Bacchus/matrix/matrix.c:68-96)."""
    return matrix.matvec(vector, result, accumulate=False, stream=stream)


def run_matrix_vector_multiplication(output_vector_base_directory,
                                     input_vector_base_directory, scheduler: Operator):
    """Add H x to the vec_<id> files of the output directory, x read from
    the input directory (src/Minerva/scheduler/scheduler.c:80-145)."""
    _check(library().bacchus_hip_run_matrix_vector_multiplication(
        str(output_vector_base_directory).encode(), str(input_vector_base_directory).encode(),
        scheduler.handle))


def write_evaluation_order(combination_table_file, output_path, two_nucleon_forces_only=False):
    """Write the BLOCK: evaluation order Mars would write for comb.txt
    (src/Mars/programs/mars.c:64-199); host only, no GPU needed."""
    _check(library().bacchus_hip_write_evaluation_order(
        str(combination_table_file).encode(), int(bool(two_nucleon_forces_only)),
        str(output_path).encode()))


@dataclass
class LanczosResult:
    eigenvalues: np.ndarray
    alpha: np.ndarray
    beta: np.ndarray
    iterations: int
    matvec_seconds: float
    total_seconds: float
    converged: bool = False     # the criterion (not the iteration limit) stopped it


@dataclass
class BlockLanczosResult:
    eigenvalues: np.ndarray     # lowest Ritz values, ascending
    residuals: np.ndarray       # |H y - theta y| of those Ritz pairs
    blocks: int                 # block steps (batched matvecs) taken
    matvec_seconds: float
    total_seconds: float
    converged: bool = False     # the criterion (not max_num_blocks) stopped it


def symmetric_eigen(A):
    """Eigenvalues (ascending) and eigenvectors (columns) of the symmetric
    matrix A with the library's host Jacobi solver (no GPU needed)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    if A.shape != (n, n):
        raise BacchusError("A must be square")
    w = np.zeros(n)
    V = np.zeros((n, n))
    _check(library().bacchus_hip_symmetric_eigen(n, _ptr(A), _ptr(w), _ptr(V)))
    return w, V


def diagonalize(matrix: Operator, max_num_iterations: int, *, num_eigenvalues=5,
                start=START_E0, reorthogonalization=REORTH_REFERENCE,
                convergence=CONVERGE_NONE, tolerance=1e-7, seed=1, start_vector=None,
                target_eigenvalues=None, check_every=1, ritz_vector=None,
                target_eigenvalue=0, num_ritz_vectors=1) -> LanczosResult:
    """Lanczos on the device.  Defaults reproduce the reference
    (e_0 start, MGS against u_0..u_{k-3}, lanczos.c:60-253).  ritz_vector:
    float64 device tensor of num_ritz_vectors x D (or D for one) receiving
    the Ritz vectors of the lowest eigenvalues; target_eigenvalue: index
    the EIGENVALUES / EIGENVECTORS criteria watch (lanczos.h:23)."""
    lib = library()
    targets = None
    if target_eigenvalues is not None:
        targets = np.ascontiguousarray(target_eigenvalues, dtype=np.float64)
    st = _LanczosSettings(
        int(max_num_iterations), start, reorthogonalization, convergence,
        int(num_eigenvalues), float(tolerance), int(seed),
        ctypes.c_void_p(start_vector.data_ptr()) if start_vector is not None else None,
        _ptr(targets) if targets is not None else None, int(check_every),
        int(target_eigenvalue), int(num_ritz_vectors), 0, None)
    if ritz_vector is not None and ritz_vector.numel() < int(num_ritz_vectors) * matrix.dimension:
        raise BacchusError(f"ritz_vector must hold {num_ritz_vectors} x {matrix.dimension} doubles")
    kmax = max(1, min(matrix.dimension, int(max_num_iterations)))
    ev = np.zeros(max(1, num_eigenvalues))
    alpha = np.zeros(kmax)
    beta = np.zeros(kmax)
    res = _LanczosResult()
    rv = ctypes.c_void_p(ritz_vector.data_ptr()) if ritz_vector is not None else None
    _check(lib.bacchus_hip_lanczos(matrix.handle, ctypes.byref(st), _ptr(ev), _ptr(alpha),
                                   _ptr(beta), rv, ctypes.byref(res)))
    k = int(res.iterations)
    return LanczosResult(ev[:min(k, num_eigenvalues)], alpha[:k], beta[:k], k,
                         res.matvec_seconds, res.total_seconds, bool(res.converged))


class Batch:
    """Several operators (FP32 or BF16 elements) with B_i vectors each, every
    vector with its own (s1, s2, s3) scales, multiplied in one persistent
    launch.  Vectors are float32 device tensors of shape (D_i, B_i)."""

    def __init__(self, operators, num_vectors, tiles_per_cu=0, max_tile_vectors=0,
                 workgroups_per_cu=0, tile_words=0, waves_per_simd=0):
        lib = library()
        self.operators = list(operators)
        self.num_vectors = [int(b) for b in num_vectors]
        if len(self.operators) != len(self.num_vectors):
            raise BacchusError("one vector count per operator")
        hs = (ctypes.c_void_p * len(self.operators))(*[o.handle.value for o in self.operators])
        nv = np.ascontiguousarray(self.num_vectors, dtype=np.int32)
        tu = _BatchTuning(tiles_per_cu, max_tile_vectors, workgroups_per_cu, tile_words,
                          waves_per_simd)
        h = ctypes.c_void_p()
        _check(lib.bacchus_hip_batch_create(ctypes.cast(hs, ctypes.c_void_p), _ptr(nv),
                                            len(self.operators), ctypes.byref(tu), ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None or not self._h.value:
            raise BacchusError("batch is closed")
        return self._h

    def set_scales(self, index, scales):
        """scales: (B_index, 3) array of (s1, s2, s3) per vector."""
        sc = np.ascontiguousarray(scales, dtype=np.float64).reshape(self.num_vectors[index], 3)
        _check(library().bacchus_hip_batch_set_scales(self.handle, int(index), _ptr(sc)))

    def matvec(self, xs, ys, stream=None):
        import torch
        for i, (x, y) in enumerate(zip(xs, ys)):
            shape = (self.operators[i].dimension, self.num_vectors[i])
            for t in (x, y):
                if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() \
                        or tuple(t.shape) != shape:
                    raise BacchusError(f"operator {i}: vectors must be contiguous float32 "
                                       f"device tensors of shape {shape}")
        n = len(self.operators)
        xp = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        yp = (ctypes.c_void_p * n)(*[y.data_ptr() for y in ys])
        _check(library().bacchus_hip_batch_matvec(self.handle, ctypes.cast(xp, ctypes.c_void_p),
                                                  ctypes.cast(yp, ctypes.c_void_p),
                                                  _stream_handle(stream)))
        return ys

    def autotune(self, repetitions=3):
        best = _BatchTuning()
        ms = ctypes.c_double()
        _check(library().bacchus_hip_batch_autotune(self.handle, int(repetitions),
                                                    ctypes.byref(best), ctypes.byref(ms)))
        return {"tiles_per_cu": best.tiles_per_cu, "max_tile_vectors": best.max_tile_vectors,
                "workgroups_per_cu": best.workgroups_per_cu, "tile_words": best.tile_words,
                "waves_per_simd": best.waves_per_simd,
                "ms": ms.value}

    def stats(self) -> dict:
        b, f = ctypes.c_double(), ctypes.c_double()
        w = ctypes.c_int64()
        _check(library().bacchus_hip_batch_stats(self.handle, ctypes.byref(b), ctypes.byref(f),
                                                 ctypes.byref(w)))
        return {"algorithmic_bytes": b.value, "flops": f.value, "work_items": w.value}

    def mfma_stats(self) -> dict:
        """The matrix-core form of the P-form instructions: its 16 x 4 A
        tiles, MFMA instructions per batched matvec, the share of tile
        slots that hold a list entry, and the operators it runs for now."""
        k, on = ctypes.c_int64(), ctypes.c_int32()
        m, fill = ctypes.c_double(), ctypes.c_double()
        _check(library().bacchus_hip_batch_mfma_stats(self.handle, ctypes.byref(k), ctypes.byref(m),
                                                      ctypes.byref(fill), ctypes.byref(on)))
        return {"kgroups": k.value, "mfma_per_matvec": m.value, "fill": fill.value,
                "operators_on": on.value}

    def lanczos(self, max_num_iterations, num_eigenvalues=3, tolerance=0.0, check_every=5,
                use_graph=True, seed=1, start_vectors=None, ritz_vectors=None, num_ritz=1):
        """Independent Lanczos recurrences for every (operator, vector).
        Returns (eigenvalues, iterations, seconds); eigenvalues is a list of
        (B_i, num_eigenvalues) arrays.  start_vectors: list of float64
        device tensors (D_i, B_i) or None; ritz_vectors: list of float64
        device tensors (num_ritz, D_i, B_i) or None receiving the lowest
        num_ritz Ritz vectors of every instance."""
        n = len(self.operators)
        sv = None
        if start_vectors is not None:
            sv = (ctypes.c_void_p * n)(*[t.data_ptr() if t is not None else None
                                         for t in start_vectors])
        st = _BatchLanczosSettings(int(max_num_iterations), int(num_eigenvalues),
                                   int(bool(use_graph)), float(tolerance), int(check_every),
                                   int(num_ritz),
                                   int(seed), ctypes.cast(sv, ctypes.c_void_p) if sv else None)
        total = sum(self.num_vectors) * num_eigenvalues
        ev = np.zeros(total)
        rv = None
        if ritz_vectors is not None:
            rv = (ctypes.c_void_p * n)(*[t.data_ptr() if t is not None else None
                                         for t in ritz_vectors])
        res = _LanczosResult()
        _check(library().bacchus_hip_batch_lanczos(self.handle, ctypes.byref(st), _ptr(ev),
                                                   ctypes.cast(rv, ctypes.c_void_p) if rv else None,
                                                   ctypes.byref(res)))
        out, pos = [], 0
        for b in self.num_vectors:
            out.append(ev[pos:pos + b * num_eigenvalues].reshape(b, num_eigenvalues))
            pos += b * num_eigenvalues
        return out, int(res.iterations), float(res.total_seconds)

    def block_lanczos(self, max_num_blocks, num_eigenvalues, *, start_block=None,
                      convergence=CONVERGE_RESIDUALS, tolerance=1e-6, targets=None, seed=1,
                      ritz_vectors=None):
        """Block Lanczos over this batch's single operator, the block being
        its b vectors (all with the same scales).  start_block: float32
        device tensor (D, b) or None (seeded random); ritz_vectors: float64
        device tensor (num_eigenvalues, D) or None.  Returns
        BlockLanczosResult."""
        import torch
        if len(self.operators) != 1:
            raise BacchusError("block Lanczos needs a batch of one operator")
        D, b = self.operators[0].dimension, self.num_vectors[0]
        sb = rv = None
        if start_block is not None:
            sb = _tensor(start_block, torch.float32, "start_block")
            if tuple(start_block.shape) != (D, b):
                raise BacchusError(f"start_block must be ({D}, {b})")
        if ritz_vectors is not None:
            rv = _tensor(ritz_vectors, torch.float64, "ritz_vectors")
            if tuple(ritz_vectors.shape) != (num_eigenvalues, D):
                raise BacchusError(f"ritz_vectors must be ({num_eigenvalues}, {D})")
        tg = None
        if targets is not None:
            tg = np.ascontiguousarray(targets, dtype=np.float64)
            if tg.size < num_eigenvalues:
                raise BacchusError(f"need {num_eigenvalues} target eigenvalues")
        st = _BlockLanczosSettings(int(max_num_blocks), int(num_eigenvalues), int(convergence), 0,
                                   float(tolerance), _ptr(tg) if tg is not None else None,
                                   sb, int(seed))
        ev = np.zeros(num_eigenvalues)
        rs = np.zeros(num_eigenvalues)
        res = _LanczosResult()
        _check(library().bacchus_hip_block_lanczos(
            self.handle, ctypes.byref(st), _ptr(ev), _ptr(rs), rv, ctypes.byref(res)))
        return BlockLanczosResult(ev, rs, int(res.iterations), float(res.matvec_seconds),
                                  float(res.total_seconds), bool(res.converged))

    def close(self):
        if self._h is not None and self._h.value:
            library().bacchus_hip_batch_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _tensor(t, dtype, what):
    import torch
    if t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
        raise BacchusError(f"{what} must be a contiguous {dtype} device tensor")
    return ctypes.c_void_p(t.data_ptr())


def norm_matrix(V, stream=None):
    """N[i][j] = v_i . v_j (FP64, host (n, n)) for n float64 device vectors
    V of shape (n, D), as create_norm_matrix (src/Ceres/norm_matrix/
    norm_matrix.c:4-28)."""
    import torch
    n, D = V.shape
    N = np.zeros((n, n))
    _check(library().bacchus_hip_norm_matrix(_tensor(V, torch.float64, "V"), n, D, _ptr(N),
                                              _stream_handle(stream)))
    return N


def subspace_operator(operator, V, stream=None):
    """M[i][j] = v_j . (H v_i) (FP64, host (n, n)) for n float64 device
    vectors V of shape (n, D), one FP64-interface matvec of `operator` per
    vector, as create_subspace_operator (src/Ceres/subspace_operator/
    subspace_operator.c:7-61)."""
    import torch
    n, D = V.shape
    if D != operator.dimension:
        raise BacchusError(f"vectors have {D} components, the operator {operator.dimension}")
    M = np.zeros((n, n))
    _check(library().bacchus_hip_subspace_operator(operator.handle, _tensor(V, torch.float64, "V"),
                                                    n, _ptr(M), _stream_handle(stream)))
    return M


def basis_gram(U):
    """G = U U^T (FP64, host) for a float32 device tensor U of shape (n, D)."""
    import torch
    n, D = U.shape
    G = np.zeros((n, n))
    _check(library().bacchus_hip_basis_gram(_tensor(U, torch.float32, "U"), n, D, _ptr(G)))
    return G


def basis_project(U, V):
    """U V (FP64, host (n, B)) for U float32 (n, D) and V float64 (D, B)."""
    import torch
    n, D = U.shape
    V2 = V.reshape(D, -1)
    out = np.zeros((n, V2.shape[1]))
    _check(library().bacchus_hip_basis_project(_tensor(U, torch.float32, "U"), n, D,
                                               _tensor(V2, torch.float64, "V"), V2.shape[1],
                                               _ptr(out)))
    return out


def basis_combine(U, c, out=None, stream=None):
    """out = normalize(sum_i c_i U_i) on the device; c float32 (n,)."""
    import torch
    n, D = U.shape
    if out is None:
        out = torch.empty(D, dtype=torch.float64, device=U.device)
    if c.numel() != n:
        raise BacchusError(f"need {n} coefficients")
    _check(library().bacchus_hip_basis_combine(_tensor(U, torch.float32, "U"), n, D,
                                               _tensor(c, torch.float32, "c"),
                                               _tensor(out, torch.float64, "out"),
                                               _stream_handle(stream)))
    return out


def basis_transform(U, C):
    """out[j] = sum_i C[j, i] U[i] on the device (float32 (r, D)); C host."""
    import torch
    n, D = U.shape
    C = np.ascontiguousarray(C, dtype=np.float64)
    r = C.shape[0]
    assert C.shape[1] == n
    out = torch.empty((r, D), dtype=torch.float32, device=U.device)
    _check(library().bacchus_hip_basis_transform(_tensor(U, torch.float32, "U"), n, D, _ptr(C), r,
                                                 _tensor(out, torch.float32, "out")))
    return out


# bacchus_hip_mlp_set_precision flags
MLP_SPLIT_ACTIVATIONS = 1
MLP_FP32_LAST = 2


class MLP:
    """Learned-start MLP on the device (BF16 MFMA forward/backward, FP32
    master weights, Adam).  Inputs and targets are float32 torch tensors on
    the GPU, row-major (M rows)."""

    def __init__(self, inputs=2, hidden=256, depth=3, outputs=64, seed=1, precision=0):
        """precision: MLP_SPLIT_ACTIVATIONS | MLP_FP32_LAST flags (0: every
        GEMM operand rounded to BF16)."""
        lib = library()
        _bind_mlp(lib)
        h = ctypes.c_void_p()
        _check(lib.bacchus_hip_mlp_create(inputs, hidden, depth, outputs, seed, ctypes.byref(h)))
        self._h = h
        self.shape = (inputs, hidden, depth, outputs)
        self.num_params = int(lib.bacchus_hip_mlp_num_params(h))
        self.precision = int(precision)
        if self.precision:
            _check(lib.bacchus_hip_mlp_set_precision(h, self.precision))

    def close(self):
        if self._h is not None and self._h.value:
            library().bacchus_hip_mlp_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        _check(library().bacchus_hip_mlp_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(library().bacchus_hip_mlp_set_params(self._h, _ptr(p)))

    def forward(self, x, stream=None):
        import torch
        x = x.contiguous()
        y = torch.empty(x.shape[0], self.shape[3], dtype=torch.float32, device=x.device)
        _check(library().bacchus_hip_mlp_forward(self._h, ctypes.c_void_p(x.data_ptr()),
                                                 x.shape[0], ctypes.c_void_p(y.data_ptr()),
                                                 _stream_handle(stream)))
        return y

    def gradients(self, x, t, stream=None):
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_float()
        _check(library().bacchus_hip_mlp_gradients(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()), x.shape[0],
            _ptr(g), ctypes.byref(loss), _stream_handle(stream)))
        return float(loss.value), g

    def train_step(self, x, t, lr=1e-3, stream=None, want_loss=False):
        loss = ctypes.c_float()
        _check(library().bacchus_hip_mlp_train_step(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()), x.shape[0],
            float(lr), ctypes.byref(loss) if want_loss else None, _stream_handle(stream)))
        return float(loss.value) if want_loss else None


_MLP_BOUND = False


def _bind_mlp(lib):
    global _MLP_BOUND
    if _MLP_BOUND:
        return
    vp, c_int = ctypes.c_void_p, ctypes.c_int
    lib.bacchus_hip_mlp_create.restype = c_int
    lib.bacchus_hip_mlp_create.argtypes = [c_int, c_int, c_int, c_int, ctypes.c_uint64,
                                           ctypes.POINTER(vp)]
    lib.bacchus_hip_mlp_free.restype = None
    lib.bacchus_hip_mlp_free.argtypes = [vp]
    lib.bacchus_hip_mlp_num_params.restype = ctypes.c_int64
    lib.bacchus_hip_mlp_num_params.argtypes = [vp]
    lib.bacchus_hip_mlp_get_params.restype = c_int
    lib.bacchus_hip_mlp_get_params.argtypes = [vp, vp]
    lib.bacchus_hip_mlp_set_params.restype = c_int
    lib.bacchus_hip_mlp_set_params.argtypes = [vp, vp]
    lib.bacchus_hip_mlp_forward.restype = c_int
    lib.bacchus_hip_mlp_forward.argtypes = [vp, vp, c_int, vp, vp]
    lib.bacchus_hip_mlp_gradients.restype = c_int
    lib.bacchus_hip_mlp_gradients.argtypes = [vp, vp, vp, c_int, vp, ctypes.POINTER(ctypes.c_float), vp]
    lib.bacchus_hip_mlp_train_step.restype = c_int
    lib.bacchus_hip_mlp_train_step.argtypes = [vp, vp, vp, c_int, ctypes.c_float,
                                               ctypes.POINTER(ctypes.c_float), vp]
    lib.bacchus_hip_mlp_set_precision.restype = c_int
    lib.bacchus_hip_mlp_set_precision.argtypes = [vp, ctypes.c_int32]
    _MLP_BOUND = True
