"""Synthetic M-scheme inputs (basis blocks, index lists, matrix-element blocks,
instructions) from the native generator ``libncsm_inputs.so``.

The reference reads these from files produced by external programs (Anicre,
Mercury, Mars); :func:`build` makes the same artefacts in memory and
:meth:`Inputs.write` writes them in the formats parsed by
``src/Utilities/combination_table/combination_table.c:83-137``,
``src/Minerva/evaluation_order/evaluation_order.c:175-243``,
``src/Minerva/index_list/index_list.c:69-97`` and
``src/Minerva/matrix_block/matrix_block.c:22-62``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Config(ctypes.Structure):
    _fields_ = [
        ("num_protons", ctypes.c_int),
        ("num_neutrons", ctypes.c_int),
        ("nmax", ctypes.c_int),
        ("twice_m", ctypes.c_int),
        ("max_body", ctypes.c_int),
        ("include_one_body", ctypes.c_int),
        ("seed", ctypes.c_uint64),
        ("hw", ctypes.c_double),
        ("sigma2", ctypes.c_double),
        ("sigma3", ctypes.c_double),
    ]


INSTRUCTION_DTYPE = np.dtype([
    ("type", np.int32), ("num_particles", np.int32), ("block_in", np.int32),
    ("block_out", np.int32), ("neutron_list", np.int32),
    ("proton_list", np.int32), ("matrix_block", np.int32), ("pad", np.int32),
])


def library() -> ctypes.CDLL:
    """Load ``libncsm_inputs.so`` (built by ``__graft_entry__.build``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "libncsm_inputs.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
    # the library links the HIP runtime for its device stages; torch bundles
    # its own under the same SONAME, and loading torch first binds this
    # library to that one (one runtime per process)
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    lib.ncsm_inputs_build.restype = vp
    lib.ncsm_inputs_build.argtypes = [ctypes.POINTER(_Config)]
    lib.ncsm_inputs_build_gpu.restype = vp
    lib.ncsm_inputs_build_gpu.argtypes = [ctypes.POINTER(_Config)]
    lib.ncsm_inputs_free.argtypes = [vp]
    lib.ncsm_inputs_write.restype = ctypes.c_int
    lib.ncsm_inputs_write.argtypes = [vp, ctypes.c_char_p]
    for name in ("ncsm_dimension", "ncsm_num_blocks", "ncsm_num_lists",
                 "ncsm_num_matrix_blocks", "ncsm_num_instructions"):
        getattr(lib, name).restype = i64
        getattr(lib, name).argtypes = [vp]
    for name in ("ncsm_blocks", "ncsm_list_offsets", "ncsm_matrix_offsets",
                 "ncsm_matrix_shapes"):
        getattr(lib, name).restype = ctypes.POINTER(i64)
        getattr(lib, name).argtypes = [vp]
    for name in ("ncsm_list_entries", "ncsm_list_settings", "ncsm_orbitals"):
        getattr(lib, name).restype = ctypes.POINTER(i32)
        getattr(lib, name).argtypes = [vp]
    lib.ncsm_matrix_elements.restype = ctypes.POINTER(ctypes.c_double)
    lib.ncsm_matrix_elements.argtypes = [vp]
    lib.ncsm_instructions.restype = vp
    lib.ncsm_instructions.argtypes = [vp]
    lib.ncsm_num_orbitals.restype = i32
    lib.ncsm_num_orbitals.argtypes = [vp]
    lib.ncsm_subbasis_states.restype = i64
    lib.ncsm_subbasis_states.argtypes = [vp, i64, ctypes.c_int, ctypes.POINTER(i32)]
    lib.ncsm_element_value.restype = ctypes.c_double
    lib.ncsm_element_value.argtypes = [vp, ctypes.c_int, ctypes.POINTER(i32),
                                       ctypes.POINTER(i32)]
    _LIB = lib
    return lib


def _view(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)


@dataclass(frozen=True)
class SystemConfig:
    """A synthetic no-core system: Z+N fermions in HO shells, Nmax-truncated,
    total M and the parity of the lowest filling, seeded Gaussian 2- and
    3-body elements and the one-body term hw*(N+1.5)."""

    num_protons: int = 2
    num_neutrons: int = 2
    nmax: int = 4
    twice_m: int = 0
    max_body: int = 2
    include_one_body: bool = True
    seed: int = 1
    hw: float = 20.0
    sigma2: float = 1.0
    sigma3: float = 0.1


class Inputs:
    """Owner of one generated system; arrays are read-only numpy views into
    the native object and live as long as this instance."""

    def __init__(self, config: SystemConfig, device: bool = False):
        lib = library()
        c = _Config(config.num_protons, config.num_neutrons, config.nmax,
                    config.twice_m, config.max_body, int(config.include_one_body),
                    config.seed, config.hw, config.sigma2, config.sigma3)
        build_fn = lib.ncsm_inputs_build_gpu if device else lib.ncsm_inputs_build
        handle = build_fn(ctypes.byref(c))
        if not handle:
            raise ValueError(f"invalid system configuration {config}"
                             + (" or device failure" if device else ""))
        self.config = config
        self._lib = lib
        self._h = ctypes.c_void_p(handle)
        h = self._h
        self.dimension = int(lib.ncsm_dimension(h))
        nb = int(lib.ncsm_num_blocks(h))
        self.blocks = _view(lib.ncsm_blocks(h), 6 * nb, np.int64).reshape(nb, 6)
        nl = int(lib.ncsm_num_lists(h))
        self.list_offsets = _view(lib.ncsm_list_offsets(h), nl + 1, np.int64)
        self.list_entries = _view(lib.ncsm_list_entries(h),
                                  3 * int(self.list_offsets[-1]), np.int32).reshape(-1, 3)
        self.list_settings = _view(lib.ncsm_list_settings(h), 7 * nl, np.int32).reshape(nl, 7)
        nv = int(lib.ncsm_num_matrix_blocks(h))
        self.matrix_offsets = _view(lib.ncsm_matrix_offsets(h), nv + 1, np.int64)
        self.matrix_elements = _view(lib.ncsm_matrix_elements(h),
                                     int(self.matrix_offsets[-1]), np.float64)
        self.matrix_shapes = _view(lib.ncsm_matrix_shapes(h), 3 * nv, np.int64).reshape(nv, 3)
        ni = int(lib.ncsm_num_instructions(h))
        if ni:
            buf = (ctypes.c_char * (ni * INSTRUCTION_DTYPE.itemsize)).from_address(
                lib.ncsm_instructions(h))
            self.instructions = np.frombuffer(buf, dtype=INSTRUCTION_DTYPE)
        else:
            self.instructions = np.zeros(0, dtype=INSTRUCTION_DTYPE)
        no = int(lib.ncsm_num_orbitals(h))
        self.orbitals = _view(lib.ncsm_orbitals(h), 4 * no, np.int32).reshape(no, 4)

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.ncsm_inputs_free(h)
            self._h = None

    @property
    def block_offsets(self) -> np.ndarray:
        sizes = self.blocks[:, 4] * self.blocks[:, 5]
        return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

    def subbasis(self, block: int, species: int) -> np.ndarray:
        """Determinants (rows of orbital indices) of a block's proton
        (species 0) or neutron (species 1) sub-basis in storage order."""
        k = self.config.num_protons if species == 0 else self.config.num_neutrons
        n = int(self._lib.ncsm_subbasis_states(self._h, block, species, None))
        out = np.zeros(n * k, dtype=np.int32)
        self._lib.ncsm_subbasis_states(
            self._h, block, species, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out.reshape(n, k)

    def element(self, bra, ket) -> float:
        """Seeded matrix element between global orbital tuples (proton orbital
        s -> 2s, neutron s -> 2s+1)."""
        rank = len(bra)
        a = (ctypes.c_int32 * rank)(*bra)
        b = (ctypes.c_int32 * rank)(*ket)
        return float(self._lib.ncsm_element_value(self._h, rank, a, b))

    def write(self, directory: str) -> None:
        """Write comb.txt, evaluation_order, index_lists/ and matrix_elements/."""
        if self._lib.ncsm_inputs_write(self._h, directory.encode()) != 0:
            raise OSError(f"writing inputs to {directory} failed")

    def multiplications(self) -> np.ndarray:
        """Per-instruction multiplication count (the CALCBLOCK size field)."""
        ins = self.instructions
        lens = np.diff(self.list_offsets)
        np_in = self.blocks[ins["block_in"], 4]
        nn_in = self.blocks[ins["block_in"], 5]
        out = np.zeros(len(ins), dtype=np.int64)
        t = ins["type"]
        out[t == 1] = lens[ins["neutron_list"][t == 1]] * np_in[t == 1]
        out[t == 2] = lens[ins["proton_list"][t == 2]] * nn_in[t == 2]
        m = t == 3
        out[m] = lens[ins["neutron_list"][m]] * lens[ins["proton_list"][m]]
        return out

    def summary(self) -> dict:
        ins = self.instructions
        mults = self.multiplications()
        offdiag = ins["block_in"] != ins["block_out"]
        return {
            "dimension": self.dimension,
            "blocks": int(len(self.blocks)),
            "max_block": int((self.blocks[:, 4] * self.blocks[:, 5]).max()),
            "lists": int(len(self.list_offsets) - 1),
            "list_entries": int(self.list_offsets[-1]),
            "matrix_blocks": int(len(self.matrix_offsets) - 1),
            "matrix_elements": int(self.matrix_offsets[-1]),
            "instructions": int(len(ins)),
            "multiplications": int(mults.sum() + mults[offdiag].sum()),
        }


def build(config: SystemConfig, device: bool = False) -> Inputs:
    """The system's inputs; device=True runs the index-list and
    matrix-element stages on the GPU (ncsm_inputs_build_gpu)."""
    return Inputs(config, device)


# The workloads the benchmarks and tests use (sizes are the generator's).
SYSTEMS = {
    "he4_nmax4": SystemConfig(2, 2, 4, max_body=2, seed=1),
    "he4_nmax6": SystemConfig(2, 2, 6, max_body=2, seed=1),
    "he4_nmax8": SystemConfig(2, 2, 8, max_body=2, seed=1),
    "li6_nmax6_3b": SystemConfig(3, 3, 6, max_body=3, seed=1),
}
