"""Build the native libraries in-tree.

* ``jupiterncsm_amd/libbacchus_hip.so`` - the gfx950 HIP kernels and the C ABI
  declared in ``include/bacchus_hip.h``;
* ``jupiterncsm_amd/libncsm_inputs.so`` - the synthetic input generator
  declared in ``include/ncsm_inputs.h``;
* ``tests/c_abi/bacchus_driver`` - a plain C caller of the library (test).

Run ``python -m jupiterncsm_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["operator.hip", "np.hip", "files.hip", "lanczos.hip", "batch.hip", "block.hip", "learned.hip", "mlp.hip", "subspace.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 atomic add as the hardware instruction, not a CAS loop
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return proc


def build_inputs(verbose=False):
    """libncsm_inputs.so: the host generator (g++, OpenMP) and its device
    stages (ncsm_inputs_gpu.hip, gfx950), linked by hipcc."""
    out = os.path.join(HERE, "libncsm_inputs.so")
    bdir = os.path.join(CSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    host_obj = os.path.join(bdir, "ncsm_inputs.o")
    dev_obj = os.path.join(bdir, "ncsm_inputs_gpu.o")
    _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-c",
          os.path.join(CSRC, "ncsm_inputs.cpp"), "-o", host_obj], verbose)
    _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c",
          os.path.join(CSRC, "ncsm_inputs_gpu.hip"), "-o", dev_obj], verbose)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", host_obj, dev_obj, "-lgomp",
          "-o", out], verbose)
    return out


def build_hip(verbose=False):
    """Compiles the HIP sources in parallel (one hipcc per source, at most
    8 at a time) and links them."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(HERE, "libbacchus_hip.so")
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    objs = [os.path.join(CSRC, "_build", src.replace(".hip", ".o")) for src in HIP_SOURCES]
    cmds = [[HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src),
             "-o", obj] for src, obj in zip(HIP_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out], verbose)
    return out


def build_c_driver(verbose=False):
    """The plain C caller of the library used by the C-ABI test
    (tests/c_abi/bacchus_driver.c), linked as a C build would link it."""
    src = os.path.join(ROOT, "tests", "c_abi", "bacchus_driver.c")
    out = os.path.join(ROOT, "tests", "c_abi", "bacchus_driver")
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "lib")
    _run(["gcc", "-std=c99", "-O2", "-Wall", f"-I{INCLUDE}", src, "-o", out, f"-L{HERE}",
          "-lbacchus_hip", f"-L{rocm_lib}", f"-Wl,-rpath-link,{rocm_lib}",
          "-Wl,-rpath,$ORIGIN/../../jupiterncsm_amd", f"-Wl,-rpath,{rocm_lib}"], verbose)
    return out


def build_all(verbose=False):
    return [build_inputs(verbose), build_hip(verbose), build_c_driver(verbose)]


if __name__ == "__main__":
    for path in build_all(verbose=True):
        print("built", path)
