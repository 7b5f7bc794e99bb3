"""Build the in-tree shared library aeroflex_amd/libafx_rans.so for gfx950.

hipcc compiles the HIP translation units (device code for gfx950 only) and
g++ the plain C++ host mesh code; outputs go to aeroflex_amd/_build/ and the
library next to this file.  Run: python -m aeroflex_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libafx_rans.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# packed FP32 instructions are disabled for all device code
HIP_COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wno-unused-result"]
UNITS = [
    # (source, extra flags)
    ("afx_k64.hip", ["-ffp-contract=off"]),     # plain FP64 form: reference rounding order
    # FP32 throughput form: FMA contraction, and min/max without the NaN
    # quieting of IEEE mode (no code in this unit tests for NaN; a NaN
    # residual stays NaN through the sums and the host's norm check)
    ("afx_k32.hip", ["-ffp-contract=fast", "-fno-honor-nans"]),
    ("afx_implicit.hip", ["-ffp-contract=off"]),
    ("afx_adjoint.hip", ["-ffp-contract=off"]),
    ("afx_adjoint_sa.hip", ["-ffp-contract=off"]),
    ("afx_mapper.hip", ["-ffp-contract=off"]),   # multigrid level prolongation
    ("afx_capi.hip", ["-ffp-contract=off"]),
    ("afx_vlm.hip", ["-ffp-contract=fast"]),    # vortex lattice, FP32 LU, FP64 refinement
]
CXX_UNITS = [("afx_mesh.cpp", ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]),
             ("afx_tiles.cpp", ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]),
             ("afx_structure.cpp", ["-O2", "-std=c++17", "-fPIC"]),
             ("afx_post.cpp", ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread"])]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, flags in UNITS:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + HIP_COMMON + flags + ["-c", os.path.join(SRC, src), "-o", obj])
        objs.append(obj)
    for src, flags in CXX_UNITS:
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++"] + flags + ["-c", os.path.join(SRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-pthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    # C++ driver over the C ABI (rans_gpu.hpp), linked against the library
    drv = os.path.join(HERE, "rans_gpu_driver")
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"),
          os.path.join(HERE, "cpp", "rans_gpu_driver.cpp"), "-L", HERE, "-lafx_rans",
          "-Wl,-rpath,$ORIGIN", "-o", drv + ".tmp"])
    os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
