"""ctypes access to the oracle (oracle/liboracle_rans.so): the CPU FP64
restatement of the reference residual.  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# AFX_ORACLE_LIB selects another build of the oracle (a build with a planted
# mistake, to show that a pin fails on it)
ORACLE_LIB = os.environ.get("AFX_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle_rans.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

BC = {"farfield": 1, "slip-wall": 2, "wall": 3}


def _load():
    if not os.path.exists(ORACLE_LIB):
        import subprocess
        subprocess.run(["python", os.path.join(ROOT, "oracle", "build_oracle.py")], check=True)
    lib = C.CDLL(ORACLE_LIB)
    P, dp = C.c_void_p, C.POINTER(C.c_double)
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_mesh_read": (P, [C.c_char_p]),
        "orc_mesh_build": (P, [C.c_int, dp, dp, C.c_int, P, P, C.c_int, P, P, C.c_char_p]),
        "orc_mesh_free": (None, [P]),
        "orc_mesh_counts": (None, [P, P]),
        "orc_mesh_get": (C.c_long, [P, C.c_char_p, P]),
        "orc_mesh_bname": (C.c_char_p, [P, C.c_int]),
        "orc_solver_create": (P, [P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
        "orc_solver_free": (None, [P]),
        "orc_solver_set_bcs": (C.c_int, [P, C.c_int, C.c_char_p, P, dp]),
        "orc_solver_options": (None, [P, C.c_int, C.c_int, C.c_double, C.c_double]),
        "orc_solver_init": (None, [P]),
        "orc_solver_refill_bcs": (None, [P]),
        "orc_solver_bcs_from_internal": (None, [P]),
        "orc_solver_set_walls": (None, [P]),
        "orc_solver_calc_dt": (None, [P]),
        "orc_solver_calc_gradients": (None, [P]),
        "orc_solver_calc_limiters": (None, [P]),
        "orc_solver_fill_rhs": (None, [P]),
        "orc_solver_calc_residual": (None, [P, dp]),
        "orc_solver_explicit_step": (C.c_double, [P, C.c_double]),
        "orc_solver_uniform_residual": (C.c_double, [P]),
        "orc_solver_wall_profile": (None, [P, C.c_char_p, dp]),
        "orc_solver_face_jacobian": (None, [P, C.c_int, dp]),
        "orc_solver_mf_apply": (None, [P, dp, dp]),
        "orc_solver_set_michalak": (None, [P, C.c_int]),
        "orc_gen_mapper": (C.c_long, [C.c_int, dp, dp, C.c_int, dp, dp, dp, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_solver_fill_lhs": (C.c_long, [P, P, P, P]),
        "orc_solver_wall_dist": (None, [P]),
        "orc_save_vtu": (C.c_int, [P, C.c_char_p]),
        "orc_cp_profile": (C.c_int, [P, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_solver_vec": (C.c_long, [P, C.c_char_p, dp, dp]),
        "orc_face_flux": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, dp, dp, dp, dp, dp]),
        "orc_face_vars": (None, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp]),
        "orc_freestream": (None, [C.c_double] * 6 + [dp]),
        "orc_solver_set_prt": (None, [P, C.c_double]),
        "orc_solver_set_a1": (None, [P, C.c_int]),
        "orc_solver_face_gradient": (None, [P, C.c_int, dp]),
        "orc_sa_create": (P, [P]),
        "orc_sa_free": (None, [P]),
        "orc_sa_set_terms": (None, [P, C.c_int]),
        "orc_sa_calc_dt": (None, [P]),
        "orc_sa_fill_rhs": (None, [P]),
        "orc_sa_vec": (C.c_long, [P, C.c_char_p, dp, dp]),
        "orc_sa_source": (C.c_double, [C.c_double] * 7),
        "orc_sa_fw": (C.c_double, [C.c_double]),
        "orc_sa_fv1": (C.c_double, [C.c_double]),
        "orc_sa_prepare": (None, [P]),
        "orc_sa_face_jacobian": (None, [P, C.c_int, dp]),
        "orc_sa_source_row": (None, [P, C.c_int, dp]),
        "orc_sa_cell_source": (C.c_double, [P, C.c_int, dp]),
        "orc_sa_face_flux": (None, [P, C.c_int, dp, dp, dp, dp, dp]),
        "orc_sa_fill_lhs": (C.c_long, [P, P, P, P]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _v(a):
    return C.c_void_p(a.ctypes.data)


U32 = {"edgesNodes", "edgesCells", "cellsNodes", "cellsEdges", "boundaryEdges", "boundaryEdgesIntTag"}


class OMesh:
    def __init__(self, h):
        if not h:
            raise RuntimeError(lib.orc_last_error().decode())
        self.h = h

    @classmethod
    def read(cls, path):
        return cls(lib.orc_mesh_read(str(path).encode()))

    @classmethod
    def build(cls, x, y, cells, nverts, bsegs, bkeys, names):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        cells = np.ascontiguousarray(cells, np.int32).reshape(-1, 4)
        is_tri = np.ascontiguousarray(np.asarray(nverts) == 3, np.int32)
        bsegs = np.ascontiguousarray(bsegs, np.int32).reshape(-1, 2)
        bkeys = np.ascontiguousarray(bkeys, np.int32)
        nm = "\n".join(f"{k}={v}" for k, v in names.items()).encode()
        return cls(lib.orc_mesh_build(len(x), _d(x), _d(y), len(cells), _v(cells), _v(is_tri), len(bsegs),
                                      _v(bsegs), _v(bkeys), nm))

    def counts(self):
        c = np.zeros(5, np.int32)
        lib.orc_mesh_counts(self.h, _v(c))
        return dict(zip(("nodes", "real_cells", "cells", "faces", "bfaces"), c.tolist()))

    def get(self, f):
        n = lib.orc_mesh_get(self.h, f.encode(), None)
        dt = np.uint32 if f in U32 else (np.int32 if f == "cellsIsTriangle" else np.float64)
        out = np.empty(n, dt)
        lib.orc_mesh_get(self.h, f.encode(), _v(out))
        return out

    def boundary_names(self):
        return [lib.orc_mesh_bname(self.h, b).decode() for b in range(self.counts()["bfaces"])]

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_mesh_free(self.h)
            self.h = None


SA_CONV, SA_DIFF, SA_SOURCE, SA_MUT, SA_WALL, SA_ALL = 1, 2, 4, 8, 16, 31


class OSolver:
    """Oracle solver.  viscous: 0 inviscid, 1 laminar, 2 Spalart-Allmaras."""

    def __init__(self, mesh, R=0.71428571428, mu_L=1e-5, Pr_L=0.72, gamma=1.4, viscous=0, Pr_T=0.9):
        self.mesh = mesh
        self.h = lib.orc_solver_create(mesh.h, R, mu_L, Pr_L, gamma, viscous)
        lib.orc_solver_set_prt(self.h, Pr_T)
        self.sa = lib.orc_sa_create(self.h) if viscous == 2 else None
        c = mesh.counts()
        self.n_cells, self.n_real, self.n_faces = c["cells"], c["real_cells"], c["faces"]

    def sa_vec(self, name, value=None):
        n = lib.orc_sa_vec(self.sa, name.encode(), None, None)
        out = np.empty(n)
        lib.orc_sa_vec(self.sa, name.encode(), _d(out),
                       None if value is None else _d(np.ascontiguousarray(value, np.float64)))
        return out

    def sa_calc_dt(self):
        """Local time step of the SA explicit step (calc_dt + viscous radius)."""
        lib.orc_sa_calc_dt(self.sa)
        return self.vec("dt")

    def sa_terms(self, terms):
        lib.orc_sa_set_terms(self.sa, terms)

    def sa_fill_rhs(self):
        """fillRhoRHS with the SA equation: (RhoVector (4N), TurbVector (N))."""
        lib.orc_sa_fill_rhs(self.sa)
        return self.vec("RhoVector"), self.sa_vec("TurbVector")

    def sa_prepare(self):
        """Wall and slip ghosts from their owners, unlimited gradients of the
        five variables (and the wall distance once)."""
        lib.orc_sa_prepare(self.sa)

    def sa_face_jacobian(self, e):
        """The SA implicit step's 10 x 10 face block of face e (times its length)."""
        J = np.zeros(100)
        lib.orc_sa_face_jacobian(self.sa, e, _d(J))
        return J.reshape(10, 10)

    def sa_source_row(self, i):
        """-A dS/dv of cell i (diagonal entry clipped at 0)."""
        out = np.zeros(5)
        lib.orc_sa_source_row(self.sa, i, _d(out))
        return out

    def sa_cell_source(self, i, q5=None):
        """Per-volume source of cell i at its state (or at q5) with its gradients."""
        return lib.orc_sa_cell_source(self.sa, i, None if q5 is None else _d(np.ascontiguousarray(q5, np.float64)))

    def sa_face_flux(self, e, qL5, qR5, zL5=None, zR5=None):
        """Five-variable face flux (without the length) at face states qL5, qR5
        with centre states zL5, zR5 (None: the member states)."""
        f = np.zeros(5)
        a = [np.ascontiguousarray(v, np.float64) if v is not None else None for v in (qL5, qR5, zL5, zR5)]
        lib.orc_sa_face_flux(self.sa, e, _d(a[0]), _d(a[1]), None if a[2] is None else _d(a[2]),
                             None if a[3] is None else _d(a[3]), _d(f))
        return f

    def sa_fill_lhs(self):
        """The SA implicit step's matrix as a scipy CSR matrix (5N x 5N, AoS [5 i + k])."""
        import scipy.sparse as sp
        n = lib.orc_sa_fill_lhs(self.sa, None, None, None)
        r = np.empty(n, np.int64)
        c = np.empty(n, np.int64)
        v = np.empty(n)
        lib.orc_sa_fill_lhs(self.sa, _v(r), _v(c), _v(v))
        N = 5 * self.n_cells
        return sp.coo_matrix((v, (r, c)), shape=(N, N)).tocsr()

    def set_bcs(self, bcs):
        """bcs: {name: (type_str, (mach, angle, T, p))}"""
        names = list(bcs.keys())
        types = np.array([BC[bcs[n][0]] for n in names], np.int32)
        vars_ = np.array([bcs[n][1] for n in names], np.float64).reshape(-1)
        rc = lib.orc_solver_set_bcs(self.h, len(names), "\n".join(names).encode(), _v(types), _d(vars_))
        if rc:
            raise RuntimeError(lib.orc_last_error().decode())

    def options(self, second_order=True, green_gauss=True, limiter_k=5.0, cfl=1.0):
        lib.orc_solver_options(self.h, int(second_order), int(green_gauss), limiter_k, cfl)

    def vec(self, name, value=None):
        n = lib.orc_solver_vec(self.h, name.encode(), None, None)
        out = np.empty(n)
        if value is not None:
            value = np.ascontiguousarray(value, np.float64)
            lib.orc_solver_vec(self.h, name.encode(), _d(out), _d(value))
        else:
            lib.orc_solver_vec(self.h, name.encode(), _d(out), None)
        return out

    def __getattr__(self, name):
        fn = getattr(lib, "orc_solver_" + name)
        return lambda *a: fn(self.h, *a)

    def cp_profile(self, patch="wall"):
        """CpProfile::calc_cp (post.h:246-290): x, y, cp per patch face."""
        n = lib.orc_cp_profile(self.h, patch.encode(), None, None, None)
        x, y, cp = np.empty(n), np.empty(n), np.empty(n)
        lib.orc_cp_profile(self.h, patch.encode(), x.ctypes.data, y.ctypes.data, cp.ctypes.data)
        return x, y, cp

    def save_vtu(self, path):
        """save (post.h:58-180) of the current state and wall distance."""
        if lib.orc_save_vtu(self.h, os.fsencode(str(path))) != 0:
            raise OSError(f"cannot write {path}")

    def wall_profile(self, patch="wall"):
        out = np.zeros(3)
        lib.orc_solver_wall_profile(self.h, patch.encode(), _d(out))
        return out

    def face_jacobian(self, e):
        J = np.zeros(64)
        lib.orc_solver_face_jacobian(self.h, e, _d(J))
        return J.reshape(8, 8)

    def mf_apply(self, x):
        """RansMatrix * x (ransMatrix.h:150-178) at the member q and dt."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        lib.orc_solver_mf_apply(self.h, _d(x), _d(y))
        return y

    def fill_lhs(self):
        """fillRhoLHS (solver.h:979-1075) as a scipy CSR matrix (4N x 4N)."""
        import scipy.sparse as sp
        n = lib.orc_solver_fill_lhs(self.h, None, None, None)
        r = np.empty(n, np.int64)
        c = np.empty(n, np.int64)
        v = np.empty(n)
        lib.orc_solver_fill_lhs(self.h, _v(r), _v(c), _v(v))
        N = 4 * self.n_cells
        return sp.coo_matrix((v, (r, c)), shape=(N, N)).tocsr()

    def calc_residual(self, qk):
        qk = np.ascontiguousarray(qk, np.float64)
        lib.orc_solver_calc_residual(self.h, _d(qk))
        return self.vec("qW")

    def __del__(self):
        if getattr(self, "sa", None):
            lib.orc_sa_free(self.sa)
            self.sa = None
        if getattr(self, "h", None):
            lib.orc_solver_free(self.h)
            self.h = None


def face_flux(kind, qL, qR, nx, ny, viscous=0, gx=None, gy=None, gamma=1.4, R=0.71428571428, mu=1e-5, Pr=0.72):
    qL = np.ascontiguousarray(qL, np.float64)
    qR = np.ascontiguousarray(qR, np.float64)
    gx = np.zeros(4) if gx is None else np.ascontiguousarray(gx, np.float64)
    gy = np.zeros(4) if gy is None else np.ascontiguousarray(gy, np.float64)
    f = np.zeros(4)
    lib.orc_face_flux(kind, viscous, nx, ny, gamma, R, mu, Pr, _d(qL), _d(qR), _d(gx), _d(gy), _d(f))
    return f


def face_vars(kind, qL, qbc, nx, ny, gamma=1.4, R=0.71428571428):
    qL = np.ascontiguousarray(qL, np.float64)
    qbc = np.ascontiguousarray(qbc, np.float64)
    out = np.zeros(4)
    lib.orc_face_vars(kind, nx, ny, gamma, R, _d(qL), _d(qbc), _d(out))
    return out


def freestream(mach, angle, T=1.0, p=1.0, gamma=1.4, R=0.71428571428):
    out = np.zeros(4)
    lib.orc_freestream(gamma, R, mach, angle, T, p, _d(out))
    return out


def gen_mapper(fx, fy, cx, cy, ca):
    """multigrid::gen_mapper (multigrid.h:96-163) as a scipy CSR matrix
    (n_fine x n_coarse, one weight per cell pair; the reference repeats it
    on the 4 variables)."""
    import scipy.sparse as sp
    fx, fy, cx, cy, ca = (np.ascontiguousarray(a, np.float64) for a in (fx, fy, cx, cy, ca))
    nnz = lib.orc_gen_mapper(len(fx), _d(fx), _d(fy), len(cx), _d(cx), _d(cy), _d(ca), None, None, None)
    rp = np.empty(len(fx) + 1, np.int64)
    cols = np.empty(nnz, np.int32)
    vals = np.empty(nnz)
    lib.orc_gen_mapper(len(fx), _d(fx), _d(fy), len(cx), _d(cx), _d(cy), _d(ca), rp.ctypes.data, cols.ctypes.data,
                       vals.ctypes.data)
    return sp.csr_matrix((vals, cols, rp), shape=(len(fx), len(cx)))

