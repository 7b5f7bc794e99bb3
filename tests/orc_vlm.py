"""ctypes access to the oracle's vortex-lattice / beam restatement
(oracle/oracle_vlm.cpp inside oracle/liboracle_rans.so).  Test
infrastructure only."""
import ctypes as C

import numpy as np

from orc import ORACLE_LIB, _load as _load_rans  # noqa: F401  (builds the library when missing)

_load_rans()
_lib = C.CDLL(ORACLE_LIB)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i, _f = C.c_int, C.c_double
for _name, _args in {
    "ovl_segment": [_dp, _dp, _dp, _f, _dp],
    "ovl_ring": [_dp, _dp, _f, _dp],
    "ovl_geometry": [_i, _dp, _i, _ip, _ip, _f, _dp, _dp, _dp],
    "ovl_build": [_i, _dp, _i, _ip, _ip, _f, _f, _f, _f, _dp, _dp],
    "ovl_solve": [_i, _dp, _dp, _dp],
    "ovl_forces": [_i, _dp, _i, _ip, _ip, _dp, _f, _f, _f, _f, _dp, _dp],
    "ovl_beam_solve": [_i, _dp, _i, _ip, _dp, _i, _ip, _dp, _dp],
    "ovl_transfer_loads": [_i, _dp, _dp, _i, _dp, _i, _ip, _dp],
    "ovl_transfer_disp": [_i, _dp, _i, _dp, _i, _ip, _dp, _dp],
    "ovl_nonlinear": [_i, _dp, _i, _ip, _ip, _f, _f, _f, _f, _f, _i, _ip, _dp, _dp, _dp, _i, _ip, _dp, _f, _f, _i,
                      _dp, _dp, _dp],
    "ovl_beam_nonlinear": [_i, _dp, _i, _ip, _dp, _i, _ip, _dp, _i, _f, _f, _i, _dp, _dp],
    "ovl_aeroelastic": [_i, _dp, _i, _ip, _ip, _f, _f, _f, _f, _f, _i, _dp, _i, _ip, _dp, _i, _ip, _f, _i, _f,
                        _dp, _dp],
}.items():
    getattr(_lib, _name).argtypes = _args
    getattr(_lib, _name).restype = C.c_int if _name == "ovl_solve" else None


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_dp)


def _n(a):
    return np.ascontiguousarray(a, np.int32).ctypes.data_as(_ip)


def segment(P, a, b, rc=0.0):
    v = np.zeros(3)
    _lib.ovl_segment(_d(np.asarray(P, float)), _d(np.asarray(a, float)), _d(np.asarray(b, float)), rc, _d(v))
    return v


def ring(P, q, rc=0.0):
    v = np.zeros(3)
    q = np.ascontiguousarray(q, np.float64).reshape(12)
    _lib.ovl_ring(_d(np.asarray(P, float)), _d(q), rc, _d(v))
    return v


def geometry(xyz, rings, soff, alpha):
    xyz = np.ascontiguousarray(xyz, np.float64)
    rings = np.ascontiguousarray(rings, np.int32)
    soff = np.ascontiguousarray(soff, np.int32)
    n = len(rings)
    c, nn_, a = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    _lib.ovl_geometry(len(xyz), _d(xyz), len(soff) - 1, _n(soff), _n(rings), alpha, _d(c), _d(nn_), _d(a))
    return c, nn_, a


def build(xyz, rings, soff, alpha, vinf=1.0, wake_length=100.0, rc=0.0):
    xyz = np.ascontiguousarray(xyz, np.float64)
    rings = np.ascontiguousarray(rings, np.int32)
    soff = np.ascontiguousarray(soff, np.int32)
    n = len(rings)
    A, rhs = np.zeros((n, n)), np.zeros(n)
    _lib.ovl_build(len(xyz), _d(xyz), len(soff) - 1, _n(soff), _n(rings), alpha, vinf, wake_length, rc, _d(A),
                   _d(rhs))
    return A, rhs


def solve(A, b):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b))
    assert _lib.ovl_solve(len(b), _d(A), _d(b), _d(x)) == 0
    return x


def forces(xyz, rings, soff, gamma, alpha, vinf=1.0, rho=1.0, sref=1.0):
    xyz = np.ascontiguousarray(xyz, np.float64)
    rings = np.ascontiguousarray(rings, np.int32)
    soff = np.ascontiguousarray(soff, np.int32)
    gamma = np.ascontiguousarray(gamma, np.float64)
    rf, c = np.zeros((len(rings), 3)), np.zeros(3)
    _lib.ovl_forces(len(xyz), _d(xyz), len(soff) - 1, _n(soff), _n(rings), _d(gamma), alpha, vinf, rho, sref,
                    _d(rf), _d(c))
    return rf, {"cl": c[0], "cy": c[1], "cd_kj": c[2]}


def beam_solve(xyz, en, props, clamped, load):
    xyz = np.ascontiguousarray(xyz, np.float64)
    en = np.ascontiguousarray(en, np.int32)
    props = np.ascontiguousarray(props, np.float64)
    clamped = np.ascontiguousarray(clamped, np.int32)
    f = np.ascontiguousarray(load, np.float64).reshape(-1)
    u = np.zeros(6 * len(xyz))
    _lib.ovl_beam_solve(len(xyz), _d(xyz), len(en), _n(en), _d(props), len(clamped), _n(clamped), _d(f), _d(u))
    return u.reshape(-1, 6)


def transfer_loads(fpoints, rforce, bxyz, en):
    fpoints = np.ascontiguousarray(fpoints, np.float64)
    rforce = np.ascontiguousarray(rforce, np.float64)
    bxyz = np.ascontiguousarray(bxyz, np.float64)
    en = np.ascontiguousarray(en, np.int32)
    load = np.zeros(6 * len(bxyz))
    _lib.ovl_transfer_loads(len(fpoints), _d(fpoints), _d(rforce), len(bxyz), _d(bxyz), len(en), _n(en), _d(load))
    return load.reshape(-1, 6)


def transfer_disp(x0, bxyz, en, disp):
    x0 = np.ascontiguousarray(x0, np.float64)
    bxyz = np.ascontiguousarray(bxyz, np.float64)
    en = np.ascontiguousarray(en, np.int32)
    disp = np.ascontiguousarray(disp, np.float64)
    out = np.zeros_like(x0)
    _lib.ovl_transfer_disp(len(x0), _d(x0), len(bxyz), _d(bxyz), len(en), _n(en), _d(disp), _d(out))
    return out


def aeroelastic(xyz, rings, soff, alpha, beam, vinf=1.0, rho=1.0, sref=1.0, cref=1.0, tol=1e-10, max_iter=60,
                relax=1.0):
    bxyz, en, props, clamped = beam
    xyz = np.ascontiguousarray(xyz, np.float64)
    rings = np.ascontiguousarray(rings, np.int32)
    soff = np.ascontiguousarray(soff, np.int32)
    bxyz = np.ascontiguousarray(bxyz, np.float64)
    en = np.ascontiguousarray(en, np.int32)
    props = np.ascontiguousarray(props, np.float64)
    clamped = np.ascontiguousarray(clamped, np.int32)
    disp, out = np.zeros(6 * len(bxyz)), np.zeros(3)
    _lib.ovl_aeroelastic(len(xyz), _d(xyz), len(soff) - 1, _n(soff), _n(rings), alpha, vinf, rho, sref, cref,
                         len(bxyz), _d(bxyz), len(en), _n(en), _d(props), len(clamped), _n(clamped), tol, max_iter,
                         relax, _d(disp), _d(out))
    return {"iterations": int(out[0]), "tip": out[1], "cl": out[2], "disp": disp.reshape(-1, 6)}


def nonlinear(xyz, rings, soff, alpha, polar, vinf=1.0, rho=1.0, sref=1.0, cref=1.0, relax=1.0, tol=1e-6,
              max_iter=100):
    xyz = np.ascontiguousarray(xyz, np.float64)
    rings = np.ascontiguousarray(rings, np.int32)
    soff = np.ascontiguousarray(soff, np.int32)
    p = {k: np.ascontiguousarray(v) for k, v in polar.items()}
    ns = len(soff) - 1
    a, c, out = np.zeros(ns), np.zeros(ns), np.zeros(5)
    _lib.ovl_nonlinear(len(xyz), _d(xyz), ns, _n(soff), _n(rings), alpha, vinf, rho, sref, cref,
                       len(p["offsets"]) - 1, _n(p["offsets"]), _d(p["alpha_deg"]), _d(p["cl"]), _d(p["cd"]),
                       len(p["section_airfoil"]), _n(p["section_airfoil"]), _d(p["section_span"]), relax, tol,
                       max_iter, _d(a), _d(c), _d(out))
    return {"iterations": int(out[0]), "residual": out[1], "cl": out[2], "cd": out[3], "cy": out[4],
            "alpha": a, "station_cl": c}


def beam_nonlinear(xyz, en, props, clamped, load, n_steps=10, tol=1e-6, damping=1.0, max_iter=1000):
    """MODEL::get_NonLin_Solve (model.hpp:240-339): (disp (n, 6), info dict)."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    en = np.ascontiguousarray(en, np.int32)
    props = np.ascontiguousarray(props, np.float64)
    clamped = np.ascontiguousarray(clamped, np.int32)
    f = np.ascontiguousarray(load, np.float64).reshape(-1)
    u, out = np.zeros(6 * len(xyz)), np.zeros(3)
    _lib.ovl_beam_nonlinear(len(xyz), _d(xyz), len(en), _n(en), _d(props), len(clamped), _n(clamped), _d(f),
                            n_steps, tol, damping, max_iter, _d(u), _d(out))
    return u.reshape(-1, 6), {"iterations": int(out[0]), "converged": bool(out[1]), "residual": out[2]}

