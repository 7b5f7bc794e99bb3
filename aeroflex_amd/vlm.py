"""Python mirror of the vortex-lattice / beam coupling path over the C ABI in
include/afx_vlm.h (libafx_rans.so).

Names follow the reference where it has them:
  wing_lattice(...)                   lattice of a trapezoidal wing (vortex rings)
  VLM(nodes, rings, stations)         vlm::model + linear::steady (vlm/src/model.cpp, solver.cpp)
  VLM.set_flow(alpha_deg, ...)        simParam (vlm/src/input.cpp:25-52, 200-226)
  VLM.build()                         buildLHS + buildRHS (solver.cpp:139-290)
  VLM.factor() / VLM.solve()          system.compute / system.solve (solver.cpp:41-50), FP32 LU + FP64 refinement
  VLM.forces()                        wingStation::computeForces (surface.cpp:164-201)
  Beam(...) / Beam.solve(load)        structure::MODEL linear solve (model.hpp:103-131, 225-238)
  Aeroelastic(vlm, beam).solve(...)   coupling loop (intent of aeroelasticite, aerotools.cpp:25-217)
Errors raise VlmError.  There is no CPU fallback: when the HIP library is
missing, importing this module fails.
"""
import ctypes as C
import os

import numpy as np

from ._handle import Handle
from .rans import LIB_PATH

_P = C.c_void_p
_i32, _f64 = C.c_int32, C.c_double
_dp = C.POINTER(C.c_double)


class VlmError(RuntimeError):
    pass


class WingSpec(C.Structure):
    _fields_ = [("n_chord", _i32), ("n_span", _i32), ("aspect_ratio", _f64), ("taper", _f64),
                ("sweep_deg", _f64), ("root_chord", _f64), ("cosine", _i32)]


class Flow(C.Structure):
    _fields_ = [("alpha_deg", _f64), ("vinf", _f64), ("rho", _f64), ("sref", _f64), ("cref", _f64),
                ("wake_length", _f64), ("core_radius", _f64)]


class Polar(C.Structure):
    _fields_ = [("n_airfoils", _i32), ("offsets", C.POINTER(_i32)), ("alpha_deg", _dp), ("cl", _dp), ("cd", _dp),
                ("cmy", _dp), ("n_sections", _i32), ("section_airfoil", C.POINTER(_i32)), ("section_span", _dp)]


class NonlinearInfo(C.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32), ("residual", _f64), ("cl", _f64), ("cd", _f64),
                ("cy", _f64)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32), ("backward_error", _f64), ("contraction", _f64)]


class CouplingOpts(C.Structure):
    _fields_ = [("tol", _f64), ("max_iter", _i32), ("relax", _f64), ("solve_tol", _f64),
                ("solve_max_iter", _i32), ("refactor_above", _f64), ("warm_start", _i32), ("aitken", _i32)]


class CouplingInfo(C.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32), ("refactors", _i32), ("refine_steps", _i32),
                ("tip_deflection", _f64), ("cl", _f64), ("backward_error", _f64)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m aeroflex_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    ip = C.POINTER(_i32)
    sig = {
        "afx_last_error": (C.c_char_p, []),
        "afx_wing_lattice": (C.c_int, [C.POINTER(WingSpec), ip, ip, ip, _dp, ip, ip, _dp]),
        "afx_vlm_create": (C.c_int, [_i32, _dp, _i32, ip, ip, C.POINTER(_P)]),
        "afx_vlm_destroy": (C.c_int, [_P]),
        "afx_vlm_counts": (C.c_int, [_P, ip, ip, ip]),
        "afx_vlm_set_flow": (C.c_int, [_P, C.POINTER(Flow)]),
        "afx_vlm_set_nodes": (C.c_int, [_P, _dp]),
        "afx_vlm_get_nodes": (C.c_int, [_P, _dp]),
        "afx_vlm_get_rings": (C.c_int, [_P, ip, ip]),
        "afx_vlm_build": (C.c_int, [_P]),
        "afx_vlm_factor": (C.c_int, [_P]),
        "afx_vlm_solve": (C.c_int, [_P, _f64, _i32, _i32, _dp, C.POINTER(SolveInfo)]),
        "afx_vlm_forces": (C.c_int, [_P, _dp, _dp, _dp]),
        "afx_vlm_get_aic": (C.c_int, [_P, _dp]),
        "afx_vlm_nonlinear_solve": (C.c_int, [_P, C.POINTER(Polar), _f64, _f64, _i32, C.POINTER(NonlinearInfo), _dp,
                                              _dp]),
        "afx_vlm_get_system": (C.c_int, [_P, _dp, _dp, _dp, _dp]),
        "afx_dense_create": (C.c_int, [_i32, C.POINTER(_P)]),
        "afx_dense_destroy": (C.c_int, [_P]),
        "afx_dense_set": (C.c_int, [_P, _dp]),
        "afx_dense_factor": (C.c_int, [_P, ip]),
        "afx_dense_solve": (C.c_int, [_P, _dp, _f64, _i32, _dp, C.POINTER(SolveInfo)]),
        "afx_dense_get_lu": (C.c_int, [_P, C.POINTER(C.c_float), ip]),
        "afx_dense_raise_abort": (C.c_int, [_P]),
        "afx_beam_create": (C.c_int, [_i32, _dp, _i32, ip, _dp, _i32, ip, C.POINTER(_P)]),
        "afx_beam_destroy": (C.c_int, [_P]),
        "afx_beam_solve": (C.c_int, [_P, _dp, _dp]),
        "afx_beam_solve_nonlinear": (C.c_int, [_P, _dp, C.c_int32, C.c_double, C.c_double, C.c_int32, _dp,
                                               C.c_void_p]),
        "afx_aeroelastic_create": (C.c_int, [_P, _P, C.POINTER(_P)]),
        "afx_aeroelastic_destroy": (C.c_int, [_P]),
        "afx_default_coupling_opts": (None, [C.POINTER(CouplingOpts)]),
        "afx_aeroelastic_solve": (C.c_int, [_P, C.POINTER(CouplingOpts), C.POINTER(CouplingInfo)]),
        "afx_aeroelastic_get_disp": (C.c_int, [_P, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = _load()


def _chk(rc):
    if rc != 0:
        raise VlmError(f"afx error {rc}: {_lib.afx_last_error().decode()}")


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(C.POINTER(_i32))


def wing_lattice(n_chord, n_span, aspect_ratio=10.0, taper=0.4, sweep_deg=25.0, root_chord=1.0, cosine=True):
    """Vortex-ring lattice of a trapezoidal wing, both halves.  Returns
    (xyz [nn,3], ring_nodes [nr,4], station_offsets [ns+1], ref {sref, cref, span})."""
    sp = WingSpec(n_chord, n_span, aspect_ratio, taper, sweep_deg, root_chord, int(bool(cosine)))
    nn, nr, ns = _i32(), _i32(), _i32()
    ref = np.zeros(3)
    _chk(_lib.afx_wing_lattice(C.byref(sp), C.byref(nn), C.byref(nr), C.byref(ns), None, None, None, _d(ref)))
    xyz = np.zeros((nn.value, 3))
    rn = np.zeros((nr.value, 4), np.int32)
    so = np.zeros(ns.value + 1, np.int32)
    _chk(_lib.afx_wing_lattice(C.byref(sp), C.byref(nn), C.byref(nr), C.byref(ns), _d(xyz), _i(rn), _i(so), _d(ref)))
    return xyz, rn, so, {"sref": ref[0], "cref": ref[1], "span": ref[2]}


class VLM(Handle):
    """Linear steady vortex-lattice solver on the device."""

    def __init__(self, xyz, ring_nodes, station_offsets):
        self.xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        self.rings = np.ascontiguousarray(ring_nodes, np.int32).reshape(-1, 4)
        self.soff = np.ascontiguousarray(station_offsets, np.int32)
        h = _P()
        _chk(_lib.afx_vlm_create(len(self.xyz), _d(self.xyz), len(self.soff) - 1, _i(self.soff), _i(self.rings),
                                 C.byref(h)))
        self._own(h)
        self.n = len(self.rings)

    def _destroy(self, h):
        _chk(_lib.afx_vlm_destroy(h))

    def set_flow(self, alpha_deg, vinf=1.0, rho=1.0, sref=1.0, cref=1.0, wake_length=0.0, core_radius=0.0):
        self.flow = Flow(alpha_deg, vinf, rho, sref, cref, wake_length, core_radius)
        _chk(_lib.afx_vlm_set_flow(self.h, C.byref(self.flow)))

    def set_nodes(self, xyz):
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        _chk(_lib.afx_vlm_set_nodes(self.h, _d(x)))

    def get_nodes(self):
        x = np.zeros_like(self.xyz)
        _chk(_lib.afx_vlm_get_nodes(self.h, _d(x)))
        return x

    def build(self):
        _chk(_lib.afx_vlm_build(self.h))

    def factor(self):
        _chk(_lib.afx_vlm_factor(self.h))

    def solve(self, tol=0.0, max_iter=30, warm=False):
        g = np.zeros(self.n)
        info = SolveInfo()
        _chk(_lib.afx_vlm_solve(self.h, tol, max_iter, int(bool(warm)), _d(g), C.byref(info)))
        return g, info

    def forces(self):
        rf = np.zeros((self.n, 3))
        sf = np.zeros((len(self.soff) - 1, 3))
        c = np.zeros(3)
        _chk(_lib.afx_vlm_forces(self.h, _d(rf), _d(sf), _d(c)))
        return rf, sf, {"cl": c[0], "cy": c[1], "cd_kj": c[2]}

    def nonlinear_solve(self, polar, relaxation=1.0, tol=1e-6, max_iter=100):
        """nonlinear::steady::solve with a polar table (read_polar).  Returns
        (info, station alpha [deg], station viscous cl)."""
        p = _polar_struct(polar)
        ns = len(self.soff) - 1
        a, c = np.zeros(ns), np.zeros(ns)
        info = NonlinearInfo()
        _chk(_lib.afx_vlm_nonlinear_solve(self.h, C.byref(p), relaxation, tol, max_iter, C.byref(info), _d(a), _d(c)))
        return info, a, c

    def aic(self):
        a = np.zeros((self.n, self.n))
        _chk(_lib.afx_vlm_get_aic(self.h, _d(a)))
        return a

    def system(self):
        rhs, g, col, nor = np.zeros(self.n), np.zeros(self.n), np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _chk(_lib.afx_vlm_get_system(self.h, _d(rhs), _d(g), _d(col), _d(nor)))
        return rhs, g, col, nor


class Dense(Handle):
    """FP32 LU (matrix cores) + FP64 iterative refinement for a dense matrix."""

    def __init__(self, n):
        h = _P()
        _chk(_lib.afx_dense_create(int(n), C.byref(h)))
        self._own(h)
        self.n = int(n)

    def _destroy(self, h):
        _chk(_lib.afx_dense_destroy(h))

    def set(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.shape == (self.n, self.n)
        _chk(_lib.afx_dense_set(self.h, _d(a)))

    def factor(self):
        info = _i32()
        _chk(_lib.afx_dense_factor(self.h, C.byref(info)))
        return info.value

    def solve(self, b, tol=0.0, max_iter=30):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n)
        info = SolveInfo()
        _chk(_lib.afx_dense_solve(self.h, _d(b), tol, max_iter, _d(x), C.byref(info)))
        return x, info

    def raise_abort(self):
        """Test hook: raise the triangular solves' abort flag, as a timed-out
        wait does."""
        _chk(_lib.afx_dense_raise_abort(self.h))

    def lu(self):
        lu = np.zeros((self.n, self.n), np.float32, order="F")
        piv = np.zeros(self.n, np.int32)
        _chk(_lib.afx_dense_get_lu(self.h, lu.ctypes.data_as(C.POINTER(C.c_float)), _i(piv)))
        return lu, piv


class BeamNonlinearInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("residual", C.c_double)]


class Beam(Handle):
    """Euler-Bernoulli frame (CBAR elements, SPC1 clamps): linear solve and
    the corotational load-stepping solve."""

    def __init__(self, xyz, elem_nodes, props, clamped):
        self.xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        self.en = np.ascontiguousarray(elem_nodes, np.int32).reshape(-1, 2)
        self.props = np.ascontiguousarray(props, np.float64).reshape(-1, 7)
        self.clamped = np.ascontiguousarray(clamped, np.int32).reshape(-1)
        h = _P()
        _chk(_lib.afx_beam_create(len(self.xyz), _d(self.xyz), len(self.en), _i(self.en), _d(self.props),
                                  len(self.clamped), _i(self.clamped), C.byref(h)))
        self._own(h)

    def _destroy(self, h):
        _chk(_lib.afx_beam_destroy(h))

    def solve(self, load):
        f = np.ascontiguousarray(load, np.float64).reshape(-1)
        u = np.zeros(6 * len(self.xyz))
        _chk(_lib.afx_beam_solve(self.h, _d(f), _d(u)))
        return u.reshape(-1, 6)

    def solve_nonlinear(self, load, n_steps=10, tol=1e-6, damping=1.0, max_iter=1000):
        """MODEL::get_NonLin_Solve (model.hpp:240-339); defaults of the
        structure settings (structure.hpp:21-23).  Returns (disp, info)."""
        f = np.ascontiguousarray(load, np.float64).reshape(-1)
        u = np.zeros(6 * len(self.xyz))
        info = BeamNonlinearInfo()
        _chk(_lib.afx_beam_solve_nonlinear(self.h, _d(f), n_steps, tol, damping, max_iter, _d(u),
                                           C.cast(C.byref(info), C.c_void_p)))
        return u.reshape(-1, 6), info


def read_structure(path):
    """The reference's structure input (MODEL::read_data_file, model.hpp:368-617):
    comma-separated GRID, MAT1, PBAR, CBAR, SPC1, FORCE and MOMENT cards, other
    lines ignored.  Nodes are numbered in GRID order, elements taken in CBAR id
    order (the reference's std::map); SPC1 must clamp all six dofs (code
    123456).  Returns (xyz, elem_nodes, props, clamped, load) for Beam and its
    solves: props per element {E A, E Iy, E Iz, G J, V}, with PBAR's fields
    A, I1 (= Iz), I2 (= Iy), J; load 6 per node (FORCE / MOMENT = scale times
    direction)."""
    grids, mats, pbars, cbars, spcs, loads = {}, {}, {}, {}, [], []
    order = []
    for raw in open(path):
        line = raw.replace(" ", "").rstrip("\r\n")
        if not line:
            continue
        f = line.split(",")
        card = f[0]
        if card == "GRID":
            gid = int(f[1])
            grids[gid] = (float(f[3]), float(f[4]), float(f[5]))
            order.append(gid)
        elif card == "MAT1":
            mats[int(f[1])] = (float(f[2]), float(f[3]))
        elif card == "PBAR":
            pbars[int(f[1])] = (int(f[2]), float(f[3]), float(f[4]), float(f[5]), float(f[6]))
        elif card == "CBAR":
            cbars[int(f[1])] = (int(f[2]), int(f[3]), int(f[4]), float(f[5]), float(f[6]), float(f[7]))
        elif card == "SPC1":
            code = f[2]
            for tok in f[3:]:
                if tok:
                    spcs.append((int(tok), code))
        elif card in ("FORCE", "MOMENT"):
            loads.append((card, int(f[2]), float(f[4]), float(f[5]), float(f[6]), float(f[7])))
    index = {g: i for i, g in enumerate(order)}
    xyz = np.array([grids[g] for g in order], np.float64)
    en, props = [], []
    for cid in sorted(cbars):
        pid, a, b, vx, vy, vz = cbars[cid]
        mid, A, Iz, Iy, J = pbars[pid]
        E, G = mats[mid]
        en.append((index[a], index[b]))
        props.append((E * A, E * Iy, E * Iz, G * J, vx, vy, vz))
    clamped = []
    for g, code in spcs:
        if sorted(code) != list("123456"):
            raise ValueError(f"SPC1 code {code} on node {g}: only full clamps (123456) are supported")
        clamped.append(index[g])
    load = np.zeros((len(order), 6))
    for card, g, scale, dx, dy, dz in loads:
        off = 0 if card == "FORCE" else 3
        load[index[g], off:off + 3] += scale * np.array([dx, dy, dz])
    return xyz, np.array(en, np.int32), np.array(props), np.array(clamped, np.int32), load


def default_coupling_opts():
    o = CouplingOpts()
    _lib.afx_default_coupling_opts(C.byref(o))
    return o


class Aeroelastic(Handle):
    def __init__(self, vlm, beam):
        h = _P()
        _chk(_lib.afx_aeroelastic_create(vlm.h, beam.h, C.byref(h)))
        self.vlm, self.beam = vlm, beam
        self._own(h)

    def _destroy(self, h):
        _chk(_lib.afx_aeroelastic_destroy(h))

    def solve(self, **kw):
        o = default_coupling_opts()
        for k, v in kw.items():
            setattr(o, k, v)
        info = CouplingInfo()
        _chk(_lib.afx_aeroelastic_solve(self.h, C.byref(o), C.byref(info)))
        return info

    def disp(self):
        u = np.zeros(6 * len(self.beam.xyz))
        _chk(_lib.afx_aeroelastic_get_disp(self.h, _d(u)))
        return u.reshape(-1, 6)


def wing_beam(xyz, station_offsets, n_chord, elastic_axis=0.4, ea=1e5, ei_flap=(40.0, 4.0),
              ei_chord=(400.0, 40.0), gj=(30.0, 3.0)):
    """Beam along a wing lattice from wing_lattice(): one node per station
    line at elastic_axis of the local chord (leading-edge ring line to the
    trailing-edge ring line), elements between neighbouring lines, EI and GJ
    linear in |y| from the root value to the tip value, orientation vector
    (-1, 0, 0) as the reference's wing structure (examples/structure/
    Point_maillage_structure.txt) so that flap bending uses EIy, clamped at
    the line on y = 0.  Returns (xyz, elem_nodes, props, clamped)."""
    nl = (len(xyz)) // (n_chord + 1)
    pts = np.asarray(xyz).reshape(nl, n_chord + 1, 3)
    le, te = pts[:, 0, :], pts[:, -1, :]
    bx = le + elastic_axis * (te - le)
    semi = np.abs(bx[:, 1]).max()
    en = np.stack([np.arange(nl - 1), np.arange(1, nl)], 1).astype(np.int32)
    ym = 0.5 * np.abs(bx[en[:, 0], 1] + bx[en[:, 1], 1])
    f = ym / semi

    def lin(p):
        return p[0] + (p[1] - p[0]) * f

    props = np.zeros((nl - 1, 7))
    props[:, 0] = ea
    props[:, 1] = lin(ei_flap)     # EIy: bending in the local x-z plane (z up)
    props[:, 2] = lin(ei_chord)    # EIz: chordwise bending
    props[:, 3] = lin(gj)
    props[:, 4] = -1.0
    clamped = np.array([int(np.argmin(np.abs(bx[:, 1])))], np.int32)
    return bx, en, props, clamped


def read_polar(database_path, locations_path, surface=0):
    """database::table::importAirfoils / importLocations (database.cpp): the
    AIRFOIL blocks (alpha CL CD CMy rows) of one file and the SURFACE block of
    the locations file; returns the arrays of afx_polar."""
    foils, order, cur = {}, [], None
    for line in open(database_path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "AIRFOIL":
            cur = t[1]
            foils[cur] = []
            order.append(cur)
        elif t[0] == "END":
            cur = None
        elif cur is not None:
            foils[cur].append([float(v) for v in t[:4]])
    sections, spans, inside = [], [], False
    for line in open(locations_path):
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        if t[0] == "SURFACE":
            inside = int(t[1]) == surface
        elif t[0] == "END":
            inside = False
        elif inside:
            sections.append(order.index(t[0]))
            spans.append(float(t[1]))
    rows = [np.asarray(foils[n]) for n in order]
    off = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    tab = np.concatenate(rows)
    return {"offsets": off, "alpha_deg": tab[:, 0].copy(), "cl": tab[:, 1].copy(), "cd": tab[:, 2].copy(),
            "cmy": tab[:, 3].copy(), "section_airfoil": np.asarray(sections, np.int32),
            "section_span": np.asarray(spans, np.float64)}


def polar_from_airfoils(airfoils, sections):
    """The polar table of database::table (common/database/database.cpp:192-242)
    from airfoil polars computed by the RANS alpha sweep (rans.Airfoil, as
    app.cpp:168-176 fills vlm.database.airfoils from Rans::solve_airfoil)
    instead of a database file.  airfoils: {name: Airfoil} (alpha in degrees,
    increasing); sections: [(airfoil name, span location in [0, 1])] along
    the wing, as a locations file's SURFACE block lists them.  Returns the
    arrays of afx_polar (the format of read_polar)."""
    order = list(airfoils)
    rows = []
    for n in order:
        a = airfoils[n]
        if not (len(a.alpha) == len(a.cl) == len(a.cd) == len(a.cmy)) or not a.alpha:
            raise VlmError(f"airfoil '{n}': the polar needs alpha, cl, cd and cmy for every angle")
        t = np.column_stack([a.alpha, a.cl, a.cd, a.cmy]).astype(np.float64)
        if np.any(np.diff(t[:, 0]) <= 0):
            raise VlmError(f"airfoil '{n}': the angles must increase")
        rows.append(t)
    off = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    tab = np.concatenate(rows)
    sec = np.asarray([order.index(name) for name, _ in sections], np.int32)
    span = np.asarray([float(x) for _, x in sections], np.float64)
    return {"offsets": off, "alpha_deg": tab[:, 0].copy(), "cl": tab[:, 1].copy(), "cd": tab[:, 2].copy(),
            "cmy": tab[:, 3].copy(), "section_airfoil": sec, "section_span": span}


def _polar_struct(p):
    keep = {k: np.ascontiguousarray(v) for k, v in p.items()}
    st = Polar(len(keep["offsets"]) - 1, _i(keep["offsets"]), _d(keep["alpha_deg"]), _d(keep["cl"]), _d(keep["cd"]),
               _d(keep["cmy"]), len(keep["section_airfoil"]), _i(keep["section_airfoil"]), _d(keep["section_span"]))
    st._keep = keep
    return st


def read_lattice(path):
    """The reference's vortex-lattice geometry file (NODE / VORTEX /
    WINGSTATION / WING records, vlm/src/input.cpp): returns (xyz, rings,
    station_offsets) with the stations in wing order and each station's rings
    as listed (leading to trailing edge)."""
    nodes, vort, stations, wings = {}, {}, {}, []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "NODE":
            nodes[int(t[1])] = [float(v) for v in t[2:5]]
        elif t[0] == "VORTEX":
            vort[int(t[1])] = [int(v) for v in t[2:6]]
        elif t[0] == "WINGSTATION":
            stations[int(t[1])] = [int(v) for v in t[2:]]
        elif t[0] == "WING":
            wings.append([int(v) for v in t[2:]])
    ids = sorted(nodes)
    pos = {n: i for i, n in enumerate(ids)}
    xyz = np.array([nodes[n] for n in ids])
    rings, soff = [], [0]
    for w in wings:
        for st in w:
            for v in stations[st]:
                rings.append([pos[n] for n in vort[v]])
            soff.append(len(rings))
    return xyz, np.array(rings, np.int32), np.array(soff, np.int32)
