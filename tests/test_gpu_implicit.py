"""Implicit pseudo-time step on the device (implicitSolver, solver.h:852-1213)
against the oracle: the operator of fillRhoLHS, the two preconditioners
(block Jacobi; multicolour block DILU against its definition built in numpy
from the oracle's matrix), the device-resident GMRES against its stopping
rule with the oracle's operator, and the converged steady states / wall
forces of the oracle iterated with exact linear solves
(tests/golden/gen_converged.py).  The pseudo-time trajectories differ (GMRES
vs exact solves); the steady state does not."""
import json
import os
import subprocess

import numpy as np
import pytest

from cases import QUAD, TRI, make_pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("path,visc,wall", [(TRI, "inviscid", "slip-wall"), (QUAD, "inviscid", "slip-wall"),
                                            (TRI, "laminar", "wall")])
def test_lhs_operator_matches_oracle(path, visc, wall):
    s, o, q = make_pair(path, viscosity=visc, wall=wall, mu_L=7e-4 if visc == "laminar" else 1e-5)
    s.set_cfl(40.0)
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=40.0)
    s.fill()
    assert s.compute() == 0
    A = o.fill_lhs()
    rng = np.random.default_rng(11)
    for _ in range(3):
        x = rng.standard_normal(4 * s.n_cells)
        x[4 * s.n_real:] = 0.0
        y, yo = s.lhs_apply(x), A @ x
        assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))


def converge(s, iters, tol=1e-13):
    from aeroflex_amd import rans
    st = rans.Settings(tolerance=tol, max_iterations=iters)
    state, res = rans.run_solver(s, st)
    return state, res


def case_solver(mesh, mach, angle):
    from aeroflex_amd import rans
    s = rans.Solver(mesh, viscosity_model="inviscid")
    far = rans.BoundaryVariables(mach=mach, angle=angle)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
    s.init()
    s.refill_bcs()
    return s


def check_against_fixture(s, name, cl_tol=1e-6):
    ref = np.load(os.path.join(GOLDEN, f"converged_{name}.npz"), allow_pickle=False)
    cl, cd, cm = s.get_wall_profile("wall")
    assert abs(cl - float(ref["cl"])) <= cl_tol
    assert abs(cd - float(ref["cd"])) <= cl_tol
    q = s.get_q()
    nr = 4 * s.n_real
    assert np.max(np.abs(q[:nr] - ref["q"][:nr])) <= 1e-6 * np.max(np.abs(ref["q"][:nr]))


def real_blocks(A, nr):
    """The 4x4 blocks of the real-cell rows and columns of a CSR matrix in the
    reference's AoS order: dict (i, j) -> block."""
    B = A.tocsr()[:4 * nr, :4 * nr].tobsr(blocksize=(4, 4))
    out = {}
    for i in range(nr):
        for p in range(B.indptr[i], B.indptr[i + 1]):
            out[(i, int(B.indices[p]))] = B.data[p]
    return out


@pytest.mark.parametrize("path,visc,wall", [(TRI, "inviscid", "slip-wall"), (QUAD, "inviscid", "slip-wall"),
                                            (TRI, "laminar", "wall")])
def test_preconditioners_match_their_definitions(path, visc, wall):
    """Block DILU in the device's colour order, built from the oracle's
    matrix: Dt_i = D_i - sum_{j before i} A_ij Dt_j^-1 A_ji, then
    z = (Dt + U)^-1 Dt (Dt + L)^-1 x by forward and backward substitution;
    block Jacobi z_i = D_i^-1 x_i.  Ghost rows are identity.  The colouring is
    proper (no two face neighbours share a colour)."""
    s, o, q = make_pair(path, viscosity=visc, wall=wall, mu_L=7e-4 if visc == "laminar" else 1e-5)
    s.set_cfl(40.0)
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=40.0)
    s.fill()
    assert s.compute() == 0
    nr, nt = s.n_real, s.n_cells
    blk = real_blocks(o.fill_lhs(), nr)
    col, ncol = s.cell_colours()
    assert 2 <= ncol <= 5 and col.min() == 0 and col.max() == ncol - 1
    nbrs = [[] for _ in range(nr)]
    for (i, j) in blk:
        if i != j:
            nbrs[i].append(j)
            assert col[i] != col[j]
    order = np.argsort(col, kind="stable")
    Dt = {}
    for i in order:
        B = blk[(i, i)].copy()
        for j in nbrs[i]:
            if col[j] < col[i]:
                B -= blk[(i, j)] @ np.linalg.solve(Dt[j], blk[(j, i)])
        Dt[i] = B
    rng = np.random.default_rng(31)
    x = rng.standard_normal(4 * nt)
    xr = x[:4 * nr].reshape(nr, 4)
    y = np.zeros((nr, 4))
    for i in order:
        acc = xr[i].copy()
        for j in nbrs[i]:
            if col[j] < col[i]:
                acc -= blk[(i, j)] @ y[j]
        y[i] = np.linalg.solve(Dt[i], acc)
    for i in order[::-1]:
        acc = np.zeros(4)
        for j in nbrs[i]:
            if col[j] > col[i]:
                acc += blk[(i, j)] @ y[j]
        y[i] -= np.linalg.solve(Dt[i], acc)
    z = s.precond_apply(x)
    assert np.array_equal(z[4 * nr:], x[4 * nr:])
    assert np.max(np.abs(z[:4 * nr] - y.ravel())) <= 1e-10 * np.max(np.abs(y))
    s.set_preconditioner("block-jacobi")
    s.fill()
    zj = s.precond_apply(x)
    yj = np.stack([np.linalg.solve(blk[(i, i)], xr[i]) for i in range(nr)])
    assert np.max(np.abs(zj[:4 * nr] - yj.ravel())) <= 1e-12 * np.max(np.abs(yj))


@pytest.mark.parametrize("precond", ["dilu", "block-jacobi"])
@pytest.mark.parametrize("path", [TRI, QUAD])
def test_gmres_step_meets_its_tolerance(path, precond):
    """One implicit step (solver.h:1170-1213 with rhs_iterations 0) through the
    device-resident GMRES: dq = (q_new - q)/relaxation satisfies
    ||A dq - R|| <= tol ||R|| for the oracle's matrix A and residual R at the
    starting state, at the reference's tolerance 1e-2 and at 1e-10 (several
    restart cycles)."""
    for tol in (1e-2, 1e-10):
        s, o, q = make_pair(path)
        s.set_preconditioner(precond)
        s.set_linear_solver(restart=30, tol=tol, maxit=3000)
        s.set_cfl(40.0)
        o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=40.0)
        s.fill()
        assert s.compute() == 0
        relax = 0.9
        err = s.solve(relax, 0.0, 0)
        assert err > 0 and s.linear_iterations() > 0
        A = o.fill_lhs()
        o.fill_rhs()
        R = o.vec("RhoVector")
        nr = 4 * s.n_real
        dq = (s.get_q() - q) / relax
        dq[nr:] = 0.0
        assert np.linalg.norm(A @ dq - R) <= 1.001 * tol * np.linalg.norm(R)
        s.close()


def test_dilu_needs_fewer_iterations_than_block_jacobi():
    """On the 129 x 65 O-grid at CFL 40, the DILU-preconditioned GMRES takes
    fewer iterations than block Jacobi for the same tolerance."""
    from aeroflex_amd import rans
    its = {}
    for precond in ("block-jacobi", "dilu"):
        s = case_solver(rans.Mesh.naca_ogrid(128, 65, r_far=20.0, first_h=2e-3), 0.5, 2 * 0.01745)
        s.set_preconditioner(precond)
        s.set_cfl(40.0)
        s.fill()
        assert s.compute() == 0
        assert s.solve(0.9, 0.0, 0) > 0
        its[precond] = s.linear_iterations()
        s.close()
    print("linear iterations", its)
    assert its["dilu"] < its["block-jacobi"]


@pytest.mark.parametrize("precond", ["dilu", "block-jacobi"])
def test_converged_quad_fixture_matches_oracle(precond):
    from aeroflex_amd import rans
    s = case_solver(rans.Mesh.read_file(QUAD), 0.5, 1.25 * 0.01745)
    s.set_preconditioner(precond)
    state, res = converge(s, 300)
    assert state == 0 and res[-1] < 1e-9
    check_against_fixture(s, "naca0012q_coarse_m05_inviscid")


def test_config1_ogrid_500_iterations_matches_oracle():
    """NACA 0012 O-grid 129x65 (16 384 triangles), M 0.5, alpha 2 deg, 500
    pseudo-time iterations of the reference loop; CL/CD within 1e-6 of the
    oracle's converged values."""
    from aeroflex_amd import rans
    s = case_solver(rans.Mesh.naca_ogrid(128, 65, r_far=20.0, first_h=2e-3), 0.5, 2 * 0.01745)
    state, res = converge(s, 500)
    assert state == 0 and len(res) <= 500
    check_against_fixture(s, "naca_ogrid_m05_a2_inviscid")


def test_linear_solver_failure_returns_sentinel():
    """A linear solve that does not converge makes solve() return -1 and the
    iteration loop report failure (solver.h:1184,1198; multigrid.h:285)."""
    from aeroflex_amd import rans
    s = case_solver(rans.Mesh.read_file(QUAD), 0.5, 0.0)
    s.set_linear_solver(restart=5, tol=1e-14, maxit=2)
    state, res = converge(s, 5)
    assert state == 1 and res == []


def test_cpp_driver_matches_python_mirror():
    drv = os.path.join(ROOT, "aeroflex_amd", "rans_gpu_driver")
    out = subprocess.run([drv, QUAD, "0.5", "1.25", "40", "implicit", "inviscid", "slip-wall", "1e-13"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    from aeroflex_amd import rans
    s = case_solver(rans.Mesh.read_file(QUAD), 0.5, 1.25 * 0.01745)
    state, res = converge(s, 40)
    assert r["iterations"] == len(res)
    assert np.allclose(r["residuals"], res, rtol=1e-12, atol=0)
    assert abs(r["cl"] - s.get_wall_profile("wall")[0]) <= 1e-14
    # the driver refreshes the wall's Cp profile every iteration (multigrid.h:286)
    x, y, cp = s.cp_profile("wall")
    assert r["cp_faces"] == len(cp) and abs(r["cp_min"] - min(0.0, cp.min())) <= 1e-12


@pytest.mark.parametrize("every", [1, 4])
def test_loop_refreshes_the_cp_profile(every):
    """The iteration loop refreshes the CpProfile after every (or every
    n-th) iteration, as multigrid.h:286 does each iteration; the last refresh
    after the last iteration equals the final state's profile."""
    from aeroflex_amd import rans
    s = case_solver(rans.Mesh.read_file(QUAD), 0.5, 1.25 * 0.01745)
    prof = rans.CpProfile()
    st = rans.Settings(tolerance=1e-13, max_iterations=12)
    state, res = rans.run_solver(s, st, profile=prof, profile_every=every)
    assert state == 0 and len(res) == 12
    assert prof.updates == len(res) // every
    x, y, cp = s.cp_profile("wall")
    assert np.array_equal(prof.cp, cp) and np.array_equal(prof.x, x)


def test_cpp_driver_errors_are_exceptions():
    drv = os.path.join(ROOT, "aeroflex_amd", "rans_gpu_driver")
    out = subprocess.run([drv, "/nonexistent.msh", "0.5", "1", "2"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 3 and "cannot open" in out.stderr


@pytest.mark.parametrize("path,visc,wall", [(TRI, "inviscid", "slip-wall"), (QUAD, "inviscid", "slip-wall"),
                                            (TRI, "laminar", "wall")])
def test_matrix_free_operator_matches_oracle(path, visc, wall):
    """RansMatrix * x (ransMatrix.h:150-178, the RANS_MATRIX_FREE build) on the
    device against the oracle at the same q and dt, ghost rows included.  The
    forward difference divides flux rounding by h = 1e-6, hence 1e-9."""
    s, o, q = make_pair(path, viscosity=visc, wall=wall, mu_L=7e-4 if visc == "laminar" else 1e-5)
    s.set_matrix_free(True)
    s.set_cfl(3.0)
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=3.0)
    s.fill()
    assert s.compute() == 0
    o.calc_dt()
    rng = np.random.default_rng(13)
    for _ in range(3):
        x = rng.standard_normal(4 * s.n_cells)
        y, yo = s.matrix_free_apply(x), o.mf_apply(x)
        assert np.max(np.abs(y - yo)) <= 1e-9 * np.max(np.abs(yo))
        assert np.array_equal(y[4 * s.n_real:], x[4 * s.n_real:])


@pytest.mark.parametrize("path", [TRI, QUAD])
def test_matrix_free_step_meets_the_gmres_tolerance(path):
    """One implicit step of the matrix-free build (solver.h:1170-1213 with
    GMRES(tol 1e-3) on RansMatrix, identity preconditioner): the update dq =
    (q_new - q)/relaxation satisfies ||B dq - R|| <= 1e-3 ||R|| for the
    oracle's operator B and residual R at the starting state, and the returned
    error is the oracle's residual norm at the new state."""
    s, o, q = make_pair(path)
    s.set_matrix_free(True)
    s.set_cfl(2.0)
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=2.0)
    s.fill()
    assert s.compute() == 0
    relax = 0.9
    err = s.solve(relax, 0.0, 0)
    assert err > 0 and s.linear_iterations() > 0
    o.fill_rhs()
    R = o.vec("RhoVector")
    dq = (s.get_q() - q) / relax
    nr = 4 * s.n_real
    dq[nr:] = 0.0      # ghost rows: refreshed from the interior by the last residual, not by the solve
    r = o.mf_apply(dq) - R
    assert np.linalg.norm(r) <= 1.001e-3 * np.linalg.norm(R)
    o.vec("q", s.get_q())
    o.fill_rhs()
    assert abs(err - np.linalg.norm(o.vec("RhoVector"))) <= 1e-12 * err


def test_matrix_free_mode_switches():
    """lhs_apply has no matrix in matrix-free mode, matrix_free_apply needs it;
    switching back restores the assembled operator."""
    from aeroflex_amd import rans
    s, o, q = make_pair(QUAD)
    s.set_cfl(3.0)
    x = np.random.default_rng(1).standard_normal(4 * s.n_cells)
    x[4 * s.n_real:] = 0.0
    s.fill()
    with pytest.raises(rans.RansError):
        s.matrix_free_apply(x)
    y0 = s.lhs_apply(x)
    s.set_matrix_free(True)
    s.fill()
    with pytest.raises(rans.RansError):
        s.lhs_apply(x)
    s.matrix_free_apply(x)
    s.set_matrix_free(False)
    s.fill()
    assert np.array_equal(s.lhs_apply(x), y0)
