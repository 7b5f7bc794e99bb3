"""aeroflex_amd -- MI355X-native residual of AeroFLEX's unstructured FV RANS solver.

The compute path is libafx_rans.so (HIP kernels for gfx950 behind the C ABI in
include/afx_rans.h); `aeroflex_amd.rans` mirrors the reference's rans::mesh /
rans::solver interface on top of it.
"""
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libafx_rans.so")


def lib_built():
    return os.path.exists(LIB_PATH)
