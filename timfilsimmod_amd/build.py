"""Builds the gfx950 shared library timfilsimmod_amd/libtfs_rpy.so with hipcc.

Run ``python -m timfilsimmod_amd.build`` (or ``__graft_entry__.build()``).  The
library is built in-tree so it travels with the source tree; objects go to
timfilsimmod_amd/_build/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtfs_rpy.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = [
    "rpy_matvec.hip",
    "flow_field.hip",
    "solver_kernels.hip",
    "dense_kernels.hip",
    "tfs_rpy_api.hip",
    "tfs_solver_api.hip",
    "tfs_flow_api.hip",
    "tfs_broyden_api.hip",
    "tfs_velocities_api.hip",
    "tfs_filament_api.hip",
    "tfs_real32.hip",
    "tfs_io.cpp",
]

# Packed FP32 VALU forms are switched off: only scalar FP32 and FP64 VALU.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


# Host-only sources (no device code) go through the host compiler.
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE, "-I", CSRC]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".cpp"):
        cmd = ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    build_abi_consumer()
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_abi_consumer() -> str:
    """The C++ caller used by tests/test_capi_cpp_gpu.py (plain g++ against the C-ABI)."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "abi_consumer.cpp")
    out = os.path.join(root, "tests", "cpp", "abi_consumer")
    if not os.path.exists(src):
        return ""
    cmd = ["g++", "-O2", "-std=c++17", "-I", INCLUDE, src, "-o", out, "-L", HERE, "-ltfs_rpy",
           f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../timfilsimmod_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed on abi_consumer.cpp:\n{res.stderr}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
