"""Generate the converged-flow fixtures for the implicit-step parity tests.

The oracle (CPU FP64 restatement of src/rans) is iterated with the reference's
implicit pseudo-time loop (multigrid.h:240-293, solver.h:1170-1213): the
matrix is the oracle's fillRhoLHS (solver.h:979-1075) and each linear system
is solved exactly by sparse LU (the limit of the reference's GMRES at zero
tolerance).  The steady state does not depend on the linear solver, so the
device's GMRES path must reach the same state.

Writes tests/golden/converged_<case>.npz with the state (AoS q(4*i+k)), CL,
CD, CM and the residual history.  Run from the repo root:
    python tests/golden/gen_converged.py [case ...]
"""
import os
import sys
import time

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import orc  # noqa: E402

CASES = {
    # NACA 0012 O-grid 128 x 65 nodes (16 384 triangles), M 0.5, alpha = 2 deg
    # in the reference's degree factor (rans.h:88,94), inviscid, slip wall
    "naca_ogrid_m05_a2_inviscid": dict(mesh=("ogrid", 128, 65, 20.0, 2e-3), mach=0.5, angle=2 * 0.01745,
                                       viscous=0, wall="slip-wall", iters=400),
    # the reference's quad fixture, M 0.5, alpha 1.25 deg, inviscid
    "naca0012q_coarse_m05_inviscid": dict(mesh=("msh", "naca0012q_coarse.msh"), mach=0.5, angle=1.25 * 0.01745,
                                          viscous=0, wall="slip-wall", iters=400),
    # the same O-grid with the Spalart-Allmaras model at Re 1e6 (mu_L 7e-7),
    # no-slip wall, rho nu~ = 3 mu_L at the start; the model's implicit matrix
    # (SASolver::fill_lhs), CFL 5 + i up to 20, relaxation 0.5 (the
    # reference's CFL 40-100 with relaxation 0.9 goes non-finite at the
    # trailing edge within three iterations); each factorisation serves up to
    # ten updates (the sparse LU is the cost; the steady state does not depend
    # on the matrix)
    "naca_ogrid_m05_a2_sa": dict(mesh=("ogrid", 128, 65, 20.0, 2e-3), mach=0.5, angle=2 * 0.01745, viscous=2,
                                 mu_L=7e-7, wall="wall", iters=20000, start_cfl=5.0, slope_cfl=1.0, max_cfl=20.0,
                                 relax=0.5, rhs_iterations=9, floor=1e-10),
}


def build_mesh(spec):
    if spec[0] == "msh":
        return orc.OMesh.read(os.path.join(HERE, spec[1]))
    from aeroflex_amd import rans
    _, ni, nj, rfar, h = spec
    m = rans.Mesh.naca_ogrid(ni, nj, r_far=rfar, first_h=h)
    c = m.counts()
    nr = c["real_cells"]
    cells = m.get("cellsNodes").reshape(-1, 4)[:nr].astype(np.int32)
    nv = np.where(m.get("cellsIsTriangle")[:nr] == 1, 3, 4)
    en = m.get("edgesNodes").reshape(-1, 2)
    bsegs = en[m.get("boundaryEdges")].astype(np.int32)
    return orc.OMesh.build(m.get("nodesX"), m.get("nodesY"), cells, nv, bsegs,
                           m.get("boundaryEdgesIntTag").astype(np.int32),
                           {1001: "farfield", 1002: "wall", 2003: "internal"})


def aos5(R, T):
    out = np.empty(5 * T.shape[0])
    for k in range(4):
        out[k::5] = R[k::4]
    out[4::5] = T
    return out


def run_sa(name, cfg, save=True):
    """The SA model: per iteration the model's matrix at the current state,
    factorised once, then up to rhs_iterations + 1 updates v += relax A^-1
    [R; T] with the reference step's early exits (solver.h:1170-1213)."""
    m = build_mesh(cfg["mesh"])
    mu = cfg["mu_L"]
    o = orc.OSolver(m, mu_L=mu, viscous=2)
    o.set_bcs({"farfield": ("farfield", (cfg["mach"], cfg["angle"], 1.0, 1.0)),
               "wall": (cfg["wall"], (0.2, 0.0, 1.0, 1.0))})
    cfl = cfg["start_cfl"]
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=cfl)
    o.init()
    o.refill_bcs()
    o.sa_vec("nu", np.full(o.n_cells, 3.0 * mu))
    err0 = o.uniform_residual()
    hist = []
    t0 = time.time()
    rhs_iterations = cfg.get("rhs_iterations", 0)
    for i in range(cfg["iters"]):
        o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=cfl)
        lu = spla.splu(o.sa_fill_lhs().tocsc())
        for k in range(rhs_iterations + 1):
            R, T = o.sa_fill_rhs()
            b = aos5(R, T)
            err = np.linalg.norm(b)
            if k == 0:
                err_ini = err
            elif err > 10 * err_ini:
                break
            dv = lu.solve(b)
            q, nu = o.vec("q"), o.sa_vec("nu")
            for j in range(4):
                q[j::4] += cfg["relax"] * dv[j::5]
            nu += cfg["relax"] * dv[4::5]
            o.vec("q", q)
            o.sa_vec("nu", nu)
        R, T = o.sa_fill_rhs()
        err = np.sqrt(np.sum(R ** 2) + np.sum(T ** 2))
        if not np.isfinite(err):
            raise RuntimeError(f"{name}: non-finite residual at iteration {i}")
        if i == 0 and err > 2 * err0:
            err0 = err
        hist.append(err / err0)
        cfl = min(cfg["start_cfl"] + (i + 1) * cfg["slope_cfl"], cfg["max_cfl"])
        if i % 50 == 0:
            print(f"{name} it {i} res {hist[-1]:.3e} ({time.time() - t0:.0f}s)", flush=True)
        if hist[-1] < cfg["floor"]:
            break
    cl, cd, cm = o.wall_profile("wall")
    if save:
        np.savez_compressed(os.path.join(HERE, f"converged_{name}.npz"), q=o.vec("q"), nu=o.sa_vec("nu"), cl=cl,
                            cd=cd, cm=cm, residuals=np.array(hist))
    print(f"{name}: {len(hist)} iterations, residual {hist[-1]:.3e}, CL {cl:.12f} CD {cd:.12e} CM {cm:.12e}")
    return cl, cd, cm, hist[-1]


def run(name, cfg, start_cfl=40.0, slope_cfl=50.0, max_cfl=100.0, relax=0.9, rhs_iterations=5, floor=1e-11,
        save=True):
    if cfg["viscous"] == 2:
        return run_sa(name, cfg, save)
    m = build_mesh(cfg["mesh"])
    o = orc.OSolver(m, viscous=cfg["viscous"])
    o.set_bcs({"farfield": ("farfield", (cfg["mach"], cfg["angle"], 1.0, 1.0)),
               "wall": (cfg["wall"], (0.2, 0.0, 1.0, 1.0))})
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=start_cfl)
    o.init()
    o.refill_bcs()
    err0 = o.uniform_residual()
    hist = []
    cfl = start_cfl
    t0 = time.time()
    for i in range(cfg["iters"]):
        o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=cfl)
        A = o.fill_lhs()
        lu = spla.splu(A.tocsc())
        q = o.vec("q")
        err = None
        for k in range(rhs_iterations + 1):
            o.fill_rhs()
            R = o.vec("RhoVector")
            err = np.linalg.norm(R)
            if k == 0:
                err_ini = err
            elif err > 10 * err_ini:
                break
            q = o.vec("q") + relax * lu.solve(R)
            o.vec("q", q)
        o.fill_rhs()
        err = np.linalg.norm(o.vec("RhoVector"))
        if i == 0 and err > 2 * err0:
            err0 = err
        hist.append(err / err0)
        cfl = min(start_cfl + (i + 1) * slope_cfl, max_cfl)
        if i % 20 == 0:
            print(f"{name} it {i} res {hist[-1]:.3e} ({time.time() - t0:.0f}s)", flush=True)
        if hist[-1] < floor:
            break
    cl, cd, cm = o.wall_profile("wall")
    if save:
        np.savez_compressed(os.path.join(HERE, f"converged_{name}.npz"), q=o.vec("q"), cl=cl, cd=cd, cm=cm,
                            residuals=np.array(hist))
    print(f"{name}: {len(hist)} iterations, residual {hist[-1]:.3e}, CL {cl:.12f} CD {cd:.12e} CM {cm:.12e}")
    return cl, cd, cm, hist[-1]


# Polars of converged states (the alpha sweep of Rans::run_airfoil, rans.h:78-106,
# in the reference's degree factor 0.01745): the step of the reference's
# examples/conf.ini (3 deg) from -2 deg, so that the polar also covers the
# low effective angles of a finite wing's tip stations
POLARS = {
    "naca0012q_coarse_m05_inviscid": dict(mesh=("msh", "naca0012q_coarse.msh"), mach=0.5,
                                          alphas=(-2.0, 1.0, 4.0, 7.0), viscous=0, wall="slip-wall", iters=400),
}


def run_polar(name, cfg):
    rows = []
    for a in cfg["alphas"]:
        c = dict(cfg, angle=a * 0.01745)
        rows.append(run(f"{name}_a{a:g}", c, save=False))
    rows = np.array(rows)
    np.savez_compressed(os.path.join(HERE, f"polar_{name}.npz"), alpha=np.array(cfg["alphas"]), cl=rows[:, 0],
                        cd=rows[:, 1], cm=rows[:, 2], final_residual=rows[:, 3])


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES) + [f"polar:{n}" for n in POLARS]
    for n in names:
        if n.startswith("polar:"):
            run_polar(n[6:], POLARS[n[6:]])
        else:
            run(n, CASES[n])
