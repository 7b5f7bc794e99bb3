"""afx_rans_residual_norm_dev (the allocation-free residual norm) against
afx_rans_residual's norm2, through the C ABI.

The device output is one double from hipMalloc (the caller's buffer, through
the HIP runtime the library is linked with); the entry enqueues on the
handle's stream, so the test synchronises that stream before reading it.
Two solvers in a row share the library's static partial-sum array: first a
Spalart-Allmaras solver on a 48k-cell jittered mesh (FP64, FP32 with the
float4 + rho nu~ rows), then the reference's 4096-cell quad mesh (smaller,
inviscid; FP64 and FP32), so the second call reads partial sums the first
call's larger input left behind.  Bar: equal to 1e-15 relative (the same
reduction kernels and the same final arithmetic).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver_sa(rans):
    m = rans.Mesh.jittered(384, 129, first_h=1e-3, seed=3)
    s = rans.Solver(m, gas=rans.Gas(mu_L=1e-6), viscosity_model="spalart-allmaras")
    far = rans.BoundaryVariables(mach=0.7, angle=np.deg2rad(1.5))
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("wall")})
    s.init()
    s.refill_bcs()
    rng = np.random.default_rng(5)
    q = s.get_q()
    q[: 4 * s.n_real] *= 1.0 + 0.01 * rng.standard_normal(4 * s.n_real)
    s.set_q(q)
    nu = s.get_nu()
    nu[: s.n_real] *= 1.0 + 0.1 * rng.standard_normal(s.n_real)
    s.set_nu(nu)
    return m, s


def _solver_quad(rans):
    from cases import QUAD
    m = rans.Mesh.read_file(QUAD)
    s = rans.Solver(m, viscosity_model="inviscid")
    far = rans.BoundaryVariables(mach=0.5, angle=2 * 0.01745)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
    s.init()
    s.refill_bcs()
    rng = np.random.default_rng(6)
    q = s.get_q()
    q[: 4 * s.n_real] *= 1.0 + 0.01 * rng.standard_normal(4 * s.n_real)
    s.set_q(q)
    return m, s


class _DeviceDouble:
    """One double of device memory allocated by the caller (hipMalloc).  The
    HIP runtime functions are looked up through the library's own handle
    (its dependency scope), so they are those of the runtime the library is
    linked against even when another copy of libamdhip64 is loaded in the
    process."""

    def __init__(self):
        from aeroflex_amd.rans import LIB_PATH
        self.hip = C.CDLL(LIB_PATH)
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), 8) == 0

    def write(self, v):
        x = C.c_double(v)
        assert self.hip.hipMemcpy(self.p, C.byref(x), 8, 1) == 0   # host to device

    def read(self):
        x = C.c_double()
        assert self.hip.hipMemcpy(C.byref(x), self.p, 8, 2) == 0   # device to host
        return x.value

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _norm_dev(rans, s, precision, out):
    out.write(-1.0)
    rc = rans.lib.afx_rans_residual_norm_dev(s.h, precision, out.p)
    if rc != 0:
        raise rans.RansError(rans.lib.afx_last_error().decode())
    s.synchronize()
    return out.read()


def _norm_ref(rans, s, precision):
    nrm = C.c_double()
    rc = rans.lib.afx_rans_residual(s.h, precision, None, C.byref(nrm))
    assert rc == 0, rans.lib.afx_last_error().decode()
    return nrm.value


def test_norm_dev_matches_residual_norm_two_calls():
    from aeroflex_amd import rans
    out = _DeviceDouble()
    m1, s1 = _solver_sa(rans)
    m2, s2 = _solver_quad(rans)
    assert s2.n_real < s1.n_real
    try:
        for s, label in ((s1, "sa-48k"), (s2, "quad-4k")):
            for prec in (rans.AFX_FP64, rans.AFX_FP32):
                print(f"case {label} precision {prec}", flush=True)
                ref = _norm_ref(rans, s, prec)
                got = _norm_dev(rans, s, prec, out)
                assert np.isfinite(ref) and ref > 0
                assert abs(got - ref) <= 1e-15 * ref, (label, prec, got, ref)
                # a second call on the same handle gives the same value
                assert _norm_dev(rans, s, prec, out) == got
    finally:
        for h in (s2, m2, s1, m1):
            h.close()
        out.free()


def test_norm_dev_rejects_null_output():
    from aeroflex_amd import rans
    m, s = _solver_quad(rans)
    try:
        rc = rans.lib.afx_rans_residual_norm_dev(s.h, rans.AFX_FP64, None)
        assert rc == -1
        assert "null" in rans.lib.afx_last_error().decode()
    finally:
        s.close()
        m.close()
