"""Python mirror of the reference's rans::mesh / rans::solver interface over the
C ABI in include/afx_rans.h (libafx_rans.so).

Names and argument meanings follow src/rans/include/rans:
  Mesh.read_file(path)                       mesh.h:834-884
  Solver(mesh, gas, viscosity_model)         solver.h:108-110
  Solver.set_bcs({name: BoundaryCondition})  solver.h:200-247
  set_second_order / set_gradient_scheme / set_limiter_k / set_cfl
  init / refill_bcs / bcs_from_internal / get_q / set_q
  calc_dt                                    solver.h:308-356
  fillRhoRHS                                 solver.h:1079-1152 (returns RhoVector)
  explicit_solve(relaxation)                 solver.h:802-828
  get_uniform_residual                       solver.h:636-690
  get_wall_profile(patch)                    post.h:301-387
Vectors use the reference AoS layout q(4*i+k) over real + ghost cells.

Errors raise RansError (the reference raises std::runtime_error /
std::invalid_argument / std::out_of_range at the same places).  There is no
CPU fallback: when the HIP library is missing, importing this module fails.
"""
import ctypes as C
import os

import numpy as np

from ._handle import Handle, close_all, live_handles  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafx_rans.so")

AFX_FP64 = 0
AFX_FP32 = 1
BC_TYPES = {"farfield": 1, "slip-wall": 2, "wall": 3}


class RansError(RuntimeError):
    pass


class AfxGas(C.Structure):
    _fields_ = [("R", C.c_double), ("mu_L", C.c_double), ("Pr_L", C.c_double),
                ("Pr_T", C.c_double), ("gamma", C.c_double)]


class AfxOpts(C.Structure):
    _fields_ = [("viscosity", C.c_int32), ("second_order", C.c_int32), ("gradient", C.c_int32),
                ("limiter_k", C.c_double), ("renumber", C.c_int32), ("tile_cells", C.c_int32),
                ("fused", C.c_int32), ("limiter", C.c_int32), ("reserved", C.c_int32 * 4)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m aeroflex_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i32, i64, f64 = C.c_int32, C.c_int64, C.c_double
    dp = C.POINTER(C.c_double)
    sig = {
        "afx_last_error": (C.c_char_p, []),
        "afx_default_gas": (None, [C.POINTER(AfxGas)]),
        "afx_default_opts": (None, [C.POINTER(AfxOpts)]),
        "afx_mesh_read_msh": (C.c_int, [C.c_char_p, C.POINTER(P)]),
        "afx_mesh_from_elements": (C.c_int, [i64, dp, dp, i64, P, P, i64, P, P, i32, P, P, C.POINTER(P)]),
        "afx_mesh_gen_naca_ogrid": (C.c_int, [i32, i32, f64, f64, f64, C.POINTER(P)]),
        "afx_mesh_gen_jittered": (C.c_int, [i32, i32, f64, f64, f64, f64, f64, C.c_uint64, C.POINTER(P)]),
        "afx_mesh_write_msh": (C.c_int, [P, C.c_char_p]),
        "afx_mesh_counts": (C.c_int, [P, C.POINTER(i64)]),
        "afx_mesh_get": (C.c_int, [P, C.c_char_p, P, C.POINTER(i64)]),
        "afx_mesh_boundary_name": (C.c_int, [P, i64, C.c_char_p, i32]),
        "afx_mesh_rcm": (C.c_int, [P, P]),
        "afx_mesh_tile_stats": (C.c_int, [P, i32, C.POINTER(i64)]),
        "afx_mesh_destroy": (None, [P]),
        "afx_rans_create": (C.c_int, [P, C.POINTER(AfxGas), C.POINTER(AfxOpts), C.POINTER(P)]),
        "afx_rans_set_bcs": (C.c_int, [P, i32, P, P, dp]),
        "afx_rans_set_options": (C.c_int, [P, C.POINTER(AfxOpts)]),
        "afx_rans_init": (C.c_int, [P]),
        "afx_rans_refill_bcs": (C.c_int, [P]),
        "afx_rans_bcs_from_internal": (C.c_int, [P]),
        "afx_rans_set_walls_from_internal": (C.c_int, [P]),
        "afx_rans_set_state": (C.c_int, [P, dp]),
        "afx_rans_get_state": (C.c_int, [P, dp]),
        "afx_rans_set_cfl": (C.c_int, [P, f64]),
        "afx_rans_calc_dt": (C.c_int, [P, dp]),
        "afx_rans_gradients": (C.c_int, [P, i32, dp, dp, dp]),
        "afx_rans_residual": (C.c_int, [P, i32, dp, dp]),
        "afx_rans_residual_norm_dev": (C.c_int, [P, i32, P]),
        "afx_rans_explicit_residual": (C.c_int, [P, dp, dp]),
        "afx_rans_explicit_step": (C.c_int, [P, f64, dp]),
        "afx_rans_uniform_residual": (C.c_int, [P, dp]),
        "afx_rans_set_linear_solver": (C.c_int, [P, i32, f64, i32]),
        "afx_rans_set_preconditioner": (C.c_int, [P, i32]),
        "afx_rans_precond_apply": (C.c_int, [P, P, P]),
        "afx_rans_cell_colours": (C.c_int, [P, P, P]),
        "afx_rans_implicit_fill": (C.c_int, [P]),
        "afx_rans_implicit_compute": (C.c_int, [P, C.POINTER(i32)]),
        "afx_rans_implicit_solve": (C.c_int, [P, f64, f64, i32, i32, dp]),
        "afx_rans_lhs_apply": (C.c_int, [P, dp, dp]),
        "afx_rans_lhs_apply_nu": (C.c_int, [P, dp, dp, dp, dp]),
        "afx_rans_precond_apply_nu": (C.c_int, [P, dp, dp, dp, dp]),
        "afx_rans_set_matrix_free": (C.c_int, [P, C.c_int32]),
        "afx_mapper_create": (C.c_int, [P, P, C.POINTER(P)]),
        "afx_mapper_destroy": (C.c_int, [P]),
        "afx_mapper_get": (C.c_int, [P, C.POINTER(C.c_int64), P, P, P]),
        "afx_mapper_prolongate": (C.c_int, [P, P, P]),
        "afx_rans_matrix_free_apply": (C.c_int, [P, dp, dp]),
        "afx_rans_linear_iterations": (C.c_int, [P, C.POINTER(i64)]),
        "afx_rans_wall_forces": (C.c_int, [P, C.c_char_p, dp, dp, dp]),
        "afx_rans_wall_dist": (C.c_int, [P, dp]),
        "afx_rans_sa_wall_dist": (C.c_int, [P, dp]),
        "afx_rans_save_vtu": (C.c_int, [P, C.c_char_p]),
        "afx_rans_cp_profile": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_int64), dp, dp, dp]),
        "afx_rans_counts": (C.c_int, [P, C.POINTER(i64)]),
        "afx_rans_stream": (C.c_int, [P, C.POINTER(P)]),
        "afx_rans_launch_stage": (C.c_int, [P, i32, i32, i32]),
        "afx_rans_synchronize": (C.c_int, [P]),
        "afx_rans_destroy": (C.c_int, [P]),
        "afx_rans_set_turbulence": (C.c_int, [P, f64]),
        "afx_rans_set_nu": (C.c_int, [P, dp]),
        "afx_rans_get_nu": (C.c_int, [P, dp]),
        "afx_rans_nu_residual": (C.c_int, [P, dp]),
        "afx_rans_create_like": (C.c_int, [P, C.POINTER(AfxGas), C.POINTER(P)]),
        "afx_rans_time_residual": (C.c_int, [P, i32, i32, dp]),
        "afx_rans_adjoint": (C.c_int, [P, dp, dp]),
        "afx_rans_adjoint_nu": (C.c_int, [P, dp, dp, dp, dp]),
        "afx_rans_tangent": (C.c_int, [P, dp, dp, dp, dp]),
        "afx_rans_time_tangent": (C.c_int, [P, i32, dp]),
        "afx_rans_adjoint_nodes": (C.c_int, [P, dp, dp, dp]),
        "afx_rans_dcl_dalpha": (C.c_int, [P, C.c_char_p, i32, i32, f64, dp, dp]),
        "afx_rans_dcl_dnodes": (C.c_int, [P, C.c_char_p, i32, i32, f64, dp, dp]),
        "afx_rans_time_adjoint": (C.c_int, [P, i32, dp]),
        "afx_batch_create": (C.c_int, [i32, P, f64, C.POINTER(P)]),
        "afx_batch_run": (C.c_int, [P, i32, i32]),
        "afx_batch_residual_norms": (C.c_int, [P, i32, dp]),
        "afx_batch_stream": (C.c_int, [P, C.POINTER(P)]),
        "afx_batch_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc):
    if rc != 0:
        msg = lib.afx_last_error().decode(errors="replace")
        raise RansError(f"afx error {rc}: {msg}")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Gas:
    """rans::gas (core.h:27-46)."""

    def __init__(self, R=0.71428571428, mu_L=1e-5, Pr_L=0.72, Pr_T=0.9, gamma=1.4):
        self.R, self.mu_L, self.Pr_L, self.Pr_T, self.gamma = R, mu_L, Pr_L, Pr_T, gamma

    def c_struct(self):
        return AfxGas(self.R, self.mu_L, self.Pr_L, self.Pr_T, self.gamma)


class BoundaryVariables:
    """rans::boundary_variables (core.h:61-84); angle in radians."""

    def __init__(self, mach=0.2, angle=0.0, T=1.0, p=1.0):
        self.mach, self.angle, self.T, self.p = mach, angle, T, p

    def get_conservative(self, g):
        c = np.sqrt(g.gamma * g.R * self.T)
        u = self.mach * c * np.cos(self.angle)
        v = self.mach * c * np.sin(self.angle)
        rho = self.p / (g.R * self.T)
        rhoE = self.p / (g.gamma - 1) + 0.5 * rho * (u * u + v * v)
        return np.array([rho, rho * u, rho * v, rhoE])


class BoundaryCondition:
    """rans::boundary_condition (core.h:112-115)."""

    def __init__(self, bc_type, vars_far=None):
        if bc_type not in BC_TYPES:
            raise RansError(f"unsupported boundary type '{bc_type}'")
        self.bc_type = bc_type
        self.vars_far = vars_far if vars_far is not None else BoundaryVariables()


class Mesh(Handle):
    """rans::mesh (mesh.h:198-273), host side, reference order."""

    _U32 = {"edgesNodes", "edgesCells", "cellsNodes", "cellsEdges", "boundaryEdges", "boundaryEdgesIntTag"}

    def __init__(self, handle):
        self._own(C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle)

    @classmethod
    def read_file(cls, path):
        h = C.c_void_p()
        _check(lib.afx_mesh_read_msh(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def naca_ogrid(cls, ni=128, nj=65, r_far=20.0, thickness=0.12, first_h=2e-3):
        h = C.c_void_p()
        _check(lib.afx_mesh_gen_naca_ogrid(ni, nj, r_far, thickness, first_h, C.byref(h)))
        return cls(h)

    @classmethod
    def jittered(cls, ni, nj, r_far=20.0, thickness=0.12, first_h=1e-3, jitter=0.1, quad_frac=0.25, seed=1):
        h = C.c_void_p()
        _check(lib.afx_mesh_gen_jittered(ni, nj, r_far, thickness, first_h, jitter, quad_frac, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def from_elements(cls, x, y, cells, nverts, bsegs, bkeys, names):
        """names: dict physical key -> name."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        nverts = np.ascontiguousarray(nverts, dtype=np.int32)
        bsegs = np.ascontiguousarray(bsegs, dtype=np.int32).reshape(-1, 2)
        bkeys = np.ascontiguousarray(bkeys, dtype=np.int32)
        keys = np.array(list(names.keys()), dtype=np.int32)
        cn = (C.c_char_p * len(names))(*[v.encode() for v in names.values()])
        h = C.c_void_p()
        _check(lib.afx_mesh_from_elements(len(x), _dptr(x), _dptr(y), len(cells), _vptr(cells), _vptr(nverts),
                                          len(bsegs), _vptr(bsegs), _vptr(bkeys), len(keys), _vptr(keys),
                                          C.cast(cn, C.c_void_p), C.byref(h)))
        return cls(h)

    def write_msh(self, path):
        _check(lib.afx_mesh_write_msh(self.h, str(path).encode()))

    def counts(self):
        c = (C.c_int64 * 5)()
        _check(lib.afx_mesh_counts(self.h, c))
        return dict(zip(("nodes", "real_cells", "cells", "faces", "bfaces"), list(c)))

    def get(self, field):
        n = C.c_int64()
        _check(lib.afx_mesh_get(self.h, field.encode(), None, C.byref(n)))
        dt = np.uint32 if field in self._U32 else (np.int32 if field == "cellsIsTriangle" else np.float64)
        out = np.empty(n.value, dtype=dt)
        _check(lib.afx_mesh_get(self.h, field.encode(), _vptr(out), C.byref(n)))
        return out

    def boundary_names(self):
        nb = self.counts()["bfaces"]
        buf = C.create_string_buffer(256)
        out = []
        for b in range(nb):
            _check(lib.afx_mesh_boundary_name(self.h, b, buf, 256))
            out.append(buf.value.decode())
        return out

    def tile_stats(self, tile_cells):
        st = (C.c_int64 * 8)()
        _check(lib.afx_mesh_tile_stats(self.h, tile_cells, st))
        keys = ("tiles", "lmax", "gmax", "fmax", "ftmax", "sum_local", "sum_grad", "sum_faces")
        return dict(zip(keys, list(st)))

    def rcm(self):
        p = np.empty(self.counts()["real_cells"], dtype=np.int32)
        _check(lib.afx_mesh_rcm(self.h, _vptr(p)))
        return p

    def _destroy(self, h):
        lib.afx_mesh_destroy(h)


# "spalart-allmaras" is the name solver::set_bcs tests (solver.h:206);
# "spallart-allmaras" the one rans::Settings offers (core.h:176,224)
VISCOSITY = {"inviscid": 0, "laminar": 1, "spalart-allmaras": 2, "spallart-allmaras": 2}
GRADIENT = {"least-squares": 0, "green-gauss": 1}
LIMITERS = {"default": 0, "michalak": 1}


class Solver(Handle):
    """Device residual behind the rans::solver interface (solver.h:40-175)."""

    def __init__(self, mesh, gas=None, viscosity_model="inviscid", second_order=True,
                 gradient_scheme="green-gauss", limiter_k=5.0, renumber=True, tile_cells=0, fused=True,
                 limiter="default"):
        if viscosity_model not in VISCOSITY:
            raise RansError(f"unsupported viscosity model '{viscosity_model}'")
        self.gas = gas if gas is not None else Gas()
        self.mesh = mesh
        o = AfxOpts()
        lib.afx_default_opts(C.byref(o))
        o.viscosity = VISCOSITY[viscosity_model]
        o.second_order = int(second_order)
        o.gradient = GRADIENT[gradient_scheme]
        o.limiter_k = limiter_k
        o.renumber = int(renumber)
        o.tile_cells = int(tile_cells)
        o.fused = int(fused)
        if limiter not in LIMITERS:
            raise RansError(f"unknown limiter '{limiter}' (default, michalak)")
        o.limiter = LIMITERS[limiter]
        self._opts = o
        g = self.gas.c_struct()
        h = C.c_void_p()
        _check(lib.afx_rans_create(mesh.h, C.byref(g), C.byref(o), C.byref(h)))
        self._own(h)
        c = (C.c_int64 * 5)()
        _check(lib.afx_rans_counts(self.h, c))
        self.n_real, self.n_cells, self.n_faces, self.n_bfaces = c[1], c[2], c[3], c[4]
        self.bcs = {}

    # -- options (solver.h:135,149-152,162,250-252)
    def _push_opts(self):
        _check(lib.afx_rans_set_options(self.h, C.byref(self._opts)))

    def set_second_order(self, x=True):
        self._opts.second_order = int(x)
        self._push_opts()

    def set_gradient_scheme(self, s):
        self._opts.gradient = GRADIENT[s]
        self._push_opts()

    def set_limiter(self, name):
        """The reference's limiter builds: "default" (solver.h:575-583) or
        "michalak" (RANS_MICHALAK_LIMITER, solver.h:557-573)."""
        if name not in LIMITERS:
            raise RansError(f"unknown limiter '{name}' (default, michalak)")
        self._opts.limiter = LIMITERS[name]
        self._push_opts()

    def set_limiter_k(self, k):
        self._opts.limiter_k = k
        self._push_opts()

    def set_cfl(self, cfl):
        self.cfl = cfl
        _check(lib.afx_rans_set_cfl(self.h, cfl))

    def set_bcs(self, bcs):
        self.bcs = dict(bcs)
        names = list(bcs.keys())
        cn = (C.c_char_p * len(names))(*[n.encode() for n in names])
        types = np.array([BC_TYPES[bcs[n].bc_type] for n in names], dtype=np.int32)
        vars_ = np.array([[bcs[n].vars_far.mach, bcs[n].vars_far.angle, bcs[n].vars_far.T, bcs[n].vars_far.p]
                          for n in names], dtype=np.float64).reshape(-1)
        _check(lib.afx_rans_set_bcs(self.h, len(names), C.cast(cn, C.c_void_p), _vptr(types), _dptr(vars_)))

    def init(self):
        _check(lib.afx_rans_init(self.h))

    def refill_bcs(self):
        _check(lib.afx_rans_refill_bcs(self.h))

    def bcs_from_internal(self):
        _check(lib.afx_rans_bcs_from_internal(self.h))

    def set_walls_from_internal(self):
        _check(lib.afx_rans_set_walls_from_internal(self.h))

    def get_q(self):
        q = np.empty(4 * self.n_cells)
        _check(lib.afx_rans_get_state(self.h, _dptr(q)))
        return q

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.size != 4 * self.n_cells:
            raise RansError("state has the wrong length")
        _check(lib.afx_rans_set_state(self.h, _dptr(q)))

    def calc_dt(self):
        dt = np.empty(self.n_cells)
        _check(lib.afx_rans_calc_dt(self.h, _dptr(dt)))
        return dt

    def gradients(self, precision=AFX_FP64):
        gx, gy, lim = (np.empty(4 * self.n_cells) for _ in range(3))
        _check(lib.afx_rans_gradients(self.h, precision, _dptr(gx), _dptr(gy), _dptr(lim)))
        return gx, gy, lim

    def fillRhoRHS(self, precision=AFX_FP64):
        R = np.empty(4 * self.n_cells)
        nrm = C.c_double()
        _check(lib.afx_rans_residual(self.h, precision, _dptr(R), C.byref(nrm)))
        return R, nrm.value

    def explicit_residual(self, qk):
        qk = np.ascontiguousarray(qk, dtype=np.float64)
        qW = np.empty(4 * self.n_cells)
        _check(lib.afx_rans_explicit_residual(self.h, _dptr(qk), _dptr(qW)))
        return qW

    def explicit_solve(self, relaxation=1.0):
        e = C.c_double()
        _check(lib.afx_rans_explicit_step(self.h, relaxation, C.byref(e)))
        return e.value

    # -- implicit pseudo-time step (implicitSolver, solver.h:852-1213)
    def set_linear_solver(self, restart=30, tol=1e-2, maxit=500):
        _check(lib.afx_rans_set_linear_solver(self.h, restart, tol, maxit))

    PRECONDITIONERS = {"block-jacobi": 0, "dilu": 1}

    def set_preconditioner(self, kind="dilu"):
        """The GMRES preconditioner standing in for the reference's
        IncompleteLUT (solver.h:886): "block-jacobi" or "dilu" (multicolour
        block DILU, the default)."""
        if kind not in self.PRECONDITIONERS:
            raise RansError(f"unknown preconditioner {kind!r}")
        _check(lib.afx_rans_set_preconditioner(self.h, self.PRECONDITIONERS[kind]))

    def fill(self):
        """fillRhoLHS (solver.h:973-1075); with the SA model the 5x5-block
        matrix of the model (Green-Gauss gradients)."""
        _check(lib.afx_rans_implicit_fill(self.h))

    def compute(self):
        """Preconditioner set-up; 0 = ok, -1 = failure (solver.h:1160-1167)."""
        st = C.c_int32()
        _check(lib.afx_rans_implicit_compute(self.h, C.byref(st)))
        return st.value

    def solve(self, relaxation=1.0, tol=0.0, rhs_iterations=5, precision=AFX_FP64):
        """implicitSolver::solve (solver.h:1170-1213); -1 when the linear solve fails."""
        e = C.c_double()
        _check(lib.afx_rans_implicit_solve(self.h, relaxation, tol, rhs_iterations, precision, C.byref(e)))
        return e.value

    def _operator(self, fn, x, x_nu):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (4 * self.n_cells,):
            raise RansError("operand has the wrong length")
        y = np.empty_like(x)
        if self._opts.viscosity == 2:
            if x_nu is None:
                raise RansError("the Spalart-Allmaras operator needs the rho nu~ rows")
            xn = np.ascontiguousarray(x_nu, dtype=np.float64)
            if xn.shape != (self.n_cells,):
                raise RansError("rho nu~ rows have the wrong length")
            yn = np.empty_like(xn)
            _check(fn(self.h, _dptr(x), _dptr(xn), _dptr(y), _dptr(yn)))
            return y, yn
        if x_nu is not None:
            raise RansError("rho nu~ rows only with the Spalart-Allmaras model")
        _check(fn(self.h, _dptr(x), None, _dptr(y), None))
        return y

    def lhs_apply(self, x, x_nu=None):
        """LHS x with the matrix of the last fill (test hook); with the SA
        model also the rho nu~ rows: returns (y, y_nu)."""
        return self._operator(lib.afx_rans_lhs_apply_nu, x, x_nu)

    def precond_apply(self, x, x_nu=None):
        """M^-1 x with the preconditioner of the last fill (test hook); with
        the SA model as lhs_apply."""
        return self._operator(lib.afx_rans_precond_apply_nu, x, x_nu)

    def cell_colours(self):
        """(colour per real cell in reference order, number of colours) of the
        DILU sweeps (test hook)."""
        col = np.empty(self.n_real, dtype=np.int32)
        n = C.c_int32()
        _check(lib.afx_rans_cell_colours(self.h, col.ctypes.data_as(C.c_void_p), C.byref(n)))
        return col, n.value

    def set_matrix_free(self, enable=True):
        """The RANS_MATRIX_FREE build of implicitSolver (solver.h:863-925):
        RansMatrix operator, unpreconditioned GMRES (tol 1e-3, 1000 its)."""
        _check(lib.afx_rans_set_matrix_free(self.h, 1 if enable else 0))

    def matrix_free_apply(self, x):
        """RansMatrix * x (ransMatrix.h:150-178) at the current q and dt."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib.afx_rans_matrix_free_apply(self.h, _dptr(x), _dptr(y)))
        return y

    def linear_iterations(self):
        n = C.c_int64()
        _check(lib.afx_rans_linear_iterations(self.h, C.byref(n)))
        return n.value

    def get_uniform_residual(self):
        e = C.c_double()
        _check(lib.afx_rans_uniform_residual(self.h, C.byref(e)))
        return e.value

    def get_wall_profile(self, patch="wall"):
        cl, cd, cm = C.c_double(), C.c_double(), C.c_double()
        _check(lib.afx_rans_wall_forces(self.h, patch.encode(), C.byref(cl), C.byref(cd), C.byref(cm)))
        return cl.value, cd.value, cm.value

    def cp_profile(self, patch="wall"):
        """CpProfile::calc_cp (post.h:246-290): face centres x, y and cp of
        the patch's boundary faces, in the reference's boundary order."""
        n = C.c_int64(0)
        _check(lib.afx_rans_cp_profile(self.h, patch.encode(), C.byref(n), None, None, None))
        x, y, cp = np.empty(n.value), np.empty(n.value), np.empty(n.value)
        _check(lib.afx_rans_cp_profile(self.h, patch.encode(), C.byref(n), _dptr(x), _dptr(y), _dptr(cp)))
        return x, y, cp

    def save_vtu(self, path):
        """save(filename, solver) (post.h:58-180): the reference's ASCII VTU
        of the current state, the same file byte for byte."""
        _check(lib.afx_rans_save_vtu(self.h, os.fsencode(path)))

    def wall_dist(self):
        """mesh::compute_wall_dist (mesh.h:794-830): distance to the nearest
        wall / slip-wall face centre."""
        d = np.empty(self.n_cells)
        _check(lib.afx_rans_wall_dist(self.h, _dptr(d)))
        return d

    def sa_wall_dist(self):
        """The SA model's wall distance: to the nearest point of the wall /
        slip-wall face segments."""
        d = np.empty(self.n_cells)
        _check(lib.afx_rans_sa_wall_dist(self.h, _dptr(d)))
        return d

    # -- Spalart-Allmaras: rho nu~ per cell (the reference's nu, solver.h:55)
    def set_turbulence(self, nu_ratio=3.0):
        _check(lib.afx_rans_set_turbulence(self.h, nu_ratio))

    def set_nu(self, nu):
        nu = np.ascontiguousarray(nu, np.float64)
        if nu.shape != (self.n_cells,):
            raise RansError("rho nu~ vector has the wrong length")
        _check(lib.afx_rans_set_nu(self.h, _dptr(nu)))

    def get_nu(self):
        nu = np.empty(self.n_cells)
        _check(lib.afx_rans_get_nu(self.h, _dptr(nu)))
        return nu

    def nu_residual(self):
        """rho nu~ row of the last fillRhoRHS."""
        r = np.empty(self.n_cells)
        _check(lib.afx_rans_nu_residual(self.h, _dptr(r)))
        return r

    def stream(self):
        s = C.c_void_p()
        _check(lib.afx_rans_stream(self.h, C.byref(s)))
        return s.value

    def launch_stage(self, precision, stage, n):
        _check(lib.afx_rans_launch_stage(self.h, precision, stage, n))

    def synchronize(self):
        _check(lib.afx_rans_synchronize(self.h))

    def like(self, gas=None):
        """A second flow condition on this solver's mesh and tiles (no
        re-tiling); BCs and state are set on the new solver as usual."""
        other = Solver.__new__(Solver)
        other.gas = gas if gas is not None else self.gas
        other.mesh = self.mesh
        other._opts = AfxOpts()
        C.memmove(C.byref(other._opts), C.byref(self._opts), C.sizeof(AfxOpts))
        g = other.gas.c_struct()
        h = C.c_void_p()
        _check(lib.afx_rans_create_like(self.h, C.byref(g), C.byref(h)))
        other._own(h)
        other.n_real, other.n_cells, other.n_faces, other.n_bfaces = (self.n_real, self.n_cells, self.n_faces,
                                                                      self.n_bfaces)
        other.bcs = {}
        return other

    def clone(self):
        """Same mesh, tiles, gas, BCs, CFL and state (like() + copies)."""
        c = self.like(self.gas)
        c.set_bcs(self.bcs)
        c.set_q(self.get_q())
        if self._opts.viscosity == 2:
            c.set_nu(self.get_nu())
        c.set_cfl(getattr(self, "cfl", 1.0))
        return c

    def tangent(self, v, v_nu=None):
        """(dR/dq) v of fillRhoRHS at the current state (every model), in the
        q(4*i+k) layout; with the SA model also the rho nu~ direction v_nu and
        row.  Returns jv, or (jv, jv_nu) with the SA model."""
        v = np.ascontiguousarray(v, np.float64)
        if v.shape != (4 * self.n_cells,):
            raise RansError("tangent direction has the wrong length")
        jv = np.empty(4 * self.n_cells)
        sa = self._opts.viscosity == 2
        if sa:
            vn = None if v_nu is None else np.ascontiguousarray(v_nu, np.float64)
            if vn is not None and vn.shape != (self.n_cells,):
                raise RansError("rho nu~ direction has the wrong length")
            jn = np.empty(self.n_cells)
            _check(lib.afx_rans_tangent(self.h, _dptr(v), None if vn is None else _dptr(vn), _dptr(jv), _dptr(jn)))
            return jv, jn
        if v_nu is not None:
            raise RansError("rho nu~ direction only with the Spalart-Allmaras model")
        _check(lib.afx_rans_tangent(self.h, _dptr(v), None, _dptr(jv), None))
        return jv

    def adjoint(self, w, w_nu=None):
        """(dR/dq)^T w of fillRhoRHS at the current state (every model,
        Green-Gauss), in the q(4*i+k) layout; with the SA model also the
        rho nu~ seed w_nu and row.  Returns lambda, or (lambda, lambda_nu) with
        the SA model."""
        w = np.ascontiguousarray(w, np.float64)
        if w.shape != (4 * self.n_cells,):
            raise RansError("adjoint seed has the wrong length")
        lam = np.empty(4 * self.n_cells)
        if self._opts.viscosity == 2:
            wn = None if w_nu is None else np.ascontiguousarray(w_nu, np.float64)
            if wn is not None and wn.shape != (self.n_cells,):
                raise RansError("rho nu~ seed has the wrong length")
            ln = np.empty(self.n_cells)
            _check(lib.afx_rans_adjoint_nu(self.h, _dptr(w), None if wn is None else _dptr(wn), _dptr(lam),
                                           _dptr(ln)))
            return lam, ln
        if w_nu is not None:
            raise RansError("rho nu~ seed only with the Spalart-Allmaras model")
        _check(lib.afx_rans_adjoint(self.h, _dptr(w), _dptr(lam)))
        return lam

    def adjoint_nodes(self, w):
        """((dR/dq)^T w, (dR/dX)^T w): the state adjoint and the adjoint with
        respect to the node coordinates, shape (n_nodes, 2)."""
        w = np.ascontiguousarray(w, np.float64)
        if w.shape != (4 * self.n_cells,):
            raise RansError("adjoint seed has the wrong length")
        lam = np.empty(4 * self.n_cells)
        nn = self.mesh.counts()["nodes"]
        xb = np.empty(2 * nn)
        _check(lib.afx_rans_adjoint_nodes(self.h, _dptr(w), _dptr(lam), _dptr(xb)))
        return lam, xb.reshape(-1, 2)

    def dcl_dalpha(self, steps=300, checkpoint_every=25, relaxation=1.0, patch="wall"):
        """Run `steps` explicit steps from the current state and return (CL,
        dCL/dangle) by reverse accumulation with checkpoints."""
        cl, d = C.c_double(), C.c_double()
        _check(lib.afx_rans_dcl_dalpha(self.h, patch.encode(), steps, checkpoint_every, relaxation, C.byref(cl),
                                       C.byref(d)))
        return cl.value, d.value

    def dcl_dnodes(self, steps=300, checkpoint_every=25, relaxation=1.0, patch="wall"):
        """Run `steps` explicit steps from the current state and return (CL,
        dCL/dX) by reverse accumulation with checkpoints; dCL/dX has shape
        (n_nodes, 2) in the mesh's node order."""
        cl = C.c_double()
        nn = int(self.mesh.counts()["nodes"])
        xb = np.empty(2 * nn)
        _check(lib.afx_rans_dcl_dnodes(self.h, patch.encode(), steps, checkpoint_every, relaxation, C.byref(cl),
                                       _dptr(xb)))
        return cl.value, xb.reshape(-1, 2)

    def time_tangent(self, reps=20):
        """Average ms of one tangent evaluation (HIP events)."""
        ms = C.c_double()
        _check(lib.afx_rans_time_tangent(self.h, reps, C.byref(ms)))
        return ms.value

    def time_adjoint(self, reps=20):
        ms = C.c_double()
        _check(lib.afx_rans_time_adjoint(self.h, reps, C.byref(ms)))
        return ms.value

    def time_residual(self, precision=AFX_FP32, reps=20):
        """Average ms of one residual evaluation (HIP events)."""
        ms = C.c_double()
        _check(lib.afx_rans_time_residual(self.h, precision, reps, C.byref(ms)))
        return ms.value

    def _destroy(self, h):
        _check(lib.afx_rans_destroy(h))


class Batch(Handle):
    """Explicit pseudo-time steps of many solvers at once (FP32 fused
    residuals of all members in one launch per stage, FP64 updates,
    hipGraph replay).  Keeps references to its members."""

    def __init__(self, solvers, relaxation=1.0):
        self.members = list(solvers)
        arr = (C.c_void_p * len(self.members))(*[m.h.value if isinstance(m.h, C.c_void_p) else m.h
                                                 for m in self.members])
        h = C.c_void_p()
        _check(lib.afx_batch_create(len(self.members), arr, relaxation, C.byref(h)))
        self._own(h)

    def run(self, steps, graph_steps=25):
        _check(lib.afx_batch_run(self.h, steps, graph_steps))

    def residual_norms(self, precision=AFX_FP32):
        out = np.empty(len(self.members))
        _check(lib.afx_batch_residual_norms(self.h, precision, _dptr(out)))
        return out

    def stream(self):
        s = C.c_void_p()
        _check(lib.afx_batch_stream(self.h, C.byref(s)))
        return s.value

    def _destroy(self, h):
        _check(lib.afx_batch_destroy(h))


def autotune_tile(make_solver, candidates=(128, 160, 192, 224, 256), precision=AFX_FP32, reps=20, replicas=1):
    """Tile-size autotuner: build the solver at each candidate tile size with
    make_solver(tile_cells) (a fully set-up solver: BCs and state) and time
    it, returning (best tile, {tile: ms}).  replicas == 1: the fused residual
    (HIP events); replicas > 1: one batched explicit step of that many copies
    of the solver (the work of a multi-condition sweep, where the GPU is
    filled by all members together).  Candidates whose tiles do not fit the
    kernel are skipped."""
    import time as _time
    times = {}
    for t in candidates:
        s = None
        try:
            s = make_solver(t)
            if replicas <= 1:
                times[t] = s.time_residual(precision, reps)
            else:
                b = Batch([s] + [s.clone() for _ in range(replicas - 1)])
                b.run(2, 0)
                t0 = _time.perf_counter()
                b.run(reps, 0)
                times[t] = (_time.perf_counter() - t0) * 1e3 / reps
                b = None
        except RansError:
            continue
        finally:
            s = None
    if not times:
        raise RansError("no candidate tile size fits the fused kernel")
    return min(times, key=times.get), times


class Settings:
    """Solver settings of rans::Settings (core.h:166-232) that the iteration
    loop reads."""

    def __init__(self, start_cfl=40.0, slope_cfl=50.0, max_cfl=100.0, relaxation=0.9, tolerance=1e-4,
                 rhs_iterations=5, max_iterations=300, airfoil_name="wall"):
        self.start_cfl, self.slope_cfl, self.max_cfl = start_cfl, slope_cfl, max_cfl
        self.relaxation, self.tolerance = relaxation, tolerance
        self.rhs_iterations, self.max_iterations = rhs_iterations, max_iterations
        self.airfoil_name = airfoil_name


class CpProfile:
    """CpProfile (post.h:182-290) without the GUI's plotting arrays: the wall
    patch's face centres x, y and cp, refreshed by calc_cp from the device
    state (a gather on the device, then one copy of three short vectors).
    `updates` counts the refreshes."""

    def __init__(self):
        self.x = self.y = self.cp = np.empty(0)
        self.updates = 0

    def calc_cp(self, s, patch="wall"):
        self.x, self.y, self.cp = s.cp_profile(patch)
        self.updates += 1


class Mapper(Handle):
    """multigrid<T>::gen_mapper (multigrid.h:96-163), built on the device:
    the inverse-distance map from the coarse solver's cells to the fine
    solver's cells.  Both solvers must outlive it."""

    def __init__(self, coarse, fine):
        self.coarse, self.fine = coarse, fine
        h = C.c_void_p()
        _check(lib.afx_mapper_create(coarse.h, fine.h, C.byref(h)))
        self._own(h)

    def _destroy(self, h):
        _check(lib.afx_mapper_destroy(h))

    def csr(self):
        """(row_ptr, cols, weights) in reference cell order."""
        nnz = C.c_int64()
        _check(lib.afx_mapper_get(self.h, C.byref(nnz), None, None, None))
        rp = np.empty(self.fine.n_cells + 1, np.int64)
        cols = np.empty(nnz.value, np.int32)
        w = np.empty(nnz.value)
        _check(lib.afx_mapper_get(self.h, C.byref(nnz), rp.ctypes.data, cols.ctypes.data, w.ctypes.data))
        return rp, cols, w

    def prolongate(self):
        """q_fine = M q_coarse (multigrid.h:334), every variable."""
        _check(lib.afx_mapper_prolongate(self.h, self.coarse.h, self.fine.h))


DEG_TO_RAD = 0.01745   # the degree factor of Rans::run_airfoil (rans.h:88,94)
SOLVER_TYPES = ("explicit", "implicit")   # rans::Settings::solver_options (core.h:174)


class Airfoil:
    """database::airfoil (common/database/database.hpp:17-52): the polar of
    one section, alpha in degrees with cl, cd, cmy per angle."""

    def __init__(self, alpha=()):
        self.alpha = list(alpha)
        self.cl, self.cd, self.cmy = [], [], []


def compute_alphas(alpha_start, alpha_end, alpha_step):
    """Rans::compute_alphas (rans.h:53-57): alpha_start, alpha_start + step,
    ... while <= alpha_end, accumulated as the reference does."""
    out = []
    a = alpha_start
    while a <= alpha_end:
        out.append(a)
        a += alpha_step
    return out


def solve_airfoil(levels, settings, db, solver_type="implicit", patch="wall", save_prefix=None, stop=None,
                  residuals=None):
    """Rans::solve_airfoil / run_airfoil (rans.h:78-123) over device solvers
    built on the airfoil's meshes (coarse to fine, BCs set): the farfield
    angle is db.alpha[0] * 0.01745 when level 0 is initialised, then for
    each alpha (warm-started from the previous angle's state) every level
    gets the farfield at alpha * 0.01745, the multigrid loop runs without
    re-initialising, and the wall profile's CL, CD, CM are appended to db
    (with the solution file <save_prefix>_<alpha>.vtu when save_prefix is
    given).  solver_type: "implicit" or "explicit" (settings.solver_type(),
    core.h:174,202).  Returns db."""
    if solver_type not in SOLVER_TYPES:
        raise RansError(f"unknown solver type '{solver_type}' ({', '.join(SOLVER_TYPES)})")
    loop = run_solver if solver_type == "implicit" else run_solver_explicit

    def at_angle(alpha):
        for s in levels:
            bcs = dict(s.bcs)
            far = bcs["farfield"].vars_far
            bcs["farfield"] = BoundaryCondition("farfield", BoundaryVariables(far.mach, alpha * DEG_TO_RAD,
                                                                              far.T, far.p))
            s.set_bcs(bcs)

    if not db.alpha:
        return db
    mappers = [Mapper(levels[i], levels[i + 1]) for i in range(len(levels) - 1)]
    try:
        at_angle(db.alpha[0])
        levels[0].init()
        for alpha in db.alpha:
            at_angle(alpha)
            s, _ = multigrid_run(levels, settings, reinit=False, mappers=mappers, residuals=residuals, stop=stop,
                                 loop=loop)
            cl, cd, cm = s.get_wall_profile(patch)
            db.cl.append(cl)
            db.cd.append(cd)
            db.cmy.append(cm)
            if save_prefix is not None:
                s.save_vtu(f"{save_prefix}_{alpha}.vtu")
            if stop and stop():
                break
    finally:
        for mp in reversed(mappers):
            mp.close()
    return db


def run_solver_explicit(s, settings, residuals=None, stop=None, profile=None, profile_every=1):
    """multigrid<explicitSolver>::run_solver (multigrid.h:183-237): the CFL
    stays at its start value, the residual is normalised only when positive,
    the loop stops on tolerance / iteration count / stop flag.  With a
    CpProfile, calc_cp on settings.airfoil_name after every profile_every-th
    iteration, as the implicit loop does (an addition here: the reference's
    explicit loop refreshes no profile).  Returns (state, residuals): state 0
    ok, 1 failure or a non-finite residual."""
    residuals = [] if residuals is None else residuals
    i = 0
    err_0 = s.get_uniform_residual()
    while True:
        s.set_cfl(settings.start_cfl)
        err = s.explicit_solve(settings.relaxation)
        if i == 0 and err > 2 * err_0:
            err_0 = err
        if err > 0:
            err /= err_0
        if err < 0 or not np.isfinite(err):
            return 1, residuals
        residuals.append(err)
        i += 1
        if profile is not None and i % profile_every == 0:
            profile.calc_cp(s, settings.airfoil_name)
        if not (err > settings.tolerance and i < settings.max_iterations and not (stop and stop())):
            return 0, residuals


def multigrid_run(solvers, settings, reinit=True, mappers=None, residuals=None, stop=None, loop=None, profile=None,
                  profile_every=1):
    """multigrid<implicitSolver>::run (multigrid.h:317-363): level 0 from
    freestream (reinit) and its boundary states, then for each finer level
    the coarse solution mapped over (bcs_from_internal, q_fine = M q_coarse,
    refill_bcs) and solved by run_solver (or `loop`, e.g. run_solver_explicit
    for multigrid<explicitSolver>::run).  Returns (solver, state) of the last
    level run; state 1 when a level's solver failed."""
    if mappers is None:
        mappers = [Mapper(solvers[i], solvers[i + 1]) for i in range(len(solvers) - 1)]
    residuals = [] if residuals is None else residuals
    if reinit:
        solvers[0].init()
    solvers[0].refill_bcs()
    for i, s in enumerate(solvers):
        if i > 0:
            solvers[i - 1].bcs_from_internal()
            mappers[i - 1].prolongate()
            s.refill_bcs()
        state, _ = (loop or run_solver)(s, settings, residuals, stop, profile, profile_every)
        if state or (stop and stop()):
            return s, state
    return solvers[-1], 0


def run_solver(s, settings, residuals=None, stop=None, profile=None, profile_every=1):
    """multigrid<implicitSolver>::run_solver (multigrid.h:240-293) over the
    device solver: CFL ramp, first-iteration normalisation rebase, stop on
    tolerance / iteration count / stop flag; with a CpProfile, calc_cp on
    settings.airfoil_name after every profile_every-th iteration (the
    reference: every iteration, multigrid.h:286).  Returns (state,
    residuals): state 0 ok, 1 linear-solver failure or a non-finite
    residual."""
    residuals = [] if residuals is None else residuals
    cfl = settings.start_cfl
    i = 0
    err = 0.0
    err_0 = s.get_uniform_residual()
    while True:
        s.set_cfl(cfl)
        s.fill()
        ok = s.compute()
        if ok == 0:
            err = s.solve(settings.relaxation, err_0 * settings.tolerance, settings.rhs_iterations)
            if i == 0 and err > 2 * err_0:
                err_0 = err
        else:
            err = -1
        err /= err_0
        cfl = min(settings.start_cfl + (i + 1) * settings.slope_cfl, settings.max_cfl)
        i += 1
        # a non-finite normalised residual is a failure (the reference's
        # loop would read NaN as converged)
        if err < 0 or not np.isfinite(err):
            return 1, residuals
        if profile is not None and i % profile_every == 0:
            profile.calc_cp(s, settings.airfoil_name)
        residuals.append(err)
        if not (err > settings.tolerance and i < settings.max_iterations and not (stop and stop())):
            return 0, residuals
