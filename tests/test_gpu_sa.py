"""Spalart-Allmaras residual on the device against the oracle's SA terms.

FP64 fused form: both rows (mean flow and rho nu~) within 1e-12 of the
largest entry of the oracle's row (the only differences are the last-bit
roundings of cbrt in f_w).  FP32 form: mean-flow rows within 1e-4 relative L2,
rho nu~ row within 1e-3 relative L2 of the FP64 oracle.  With rho nu~ = 0
the device SA residual is the oracle's laminar residual with the model's
(textbook) face-gradient direction.
"""
import numpy as np
import pytest

import orc
from cases import QUAD, TRI, perturbed_state, rel_max

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def nu_field(n_cells, mu_L, rho, kind, seed=7):
    rng = np.random.default_rng(seed)
    if kind == "freestream":   # 3 mu_L with 1% seeded perturbation
        return 3 * mu_L * (1 + 0.01 * rng.standard_normal(n_cells))
    # chi = rho nu~ / mu_L spread over 1e-2 .. 3e2, 5% negative cells
    nu = mu_L * 10 ** rng.uniform(-2, 2.5, n_cells)
    neg = rng.random(n_cells) < 0.05
    nu[neg] = -mu_L * rng.uniform(0, 1, neg.sum())
    return nu


def sa_pair(path=TRI, wall="wall", second_order=True, gradient="green-gauss", mu_L=5e-4, nu_kind="wide",
            tile_cells=0, mesh=None, omesh=None, seed=2):
    from aeroflex_amd import rans
    m = mesh if mesh is not None else rans.Mesh.read_file(path)
    om = omesh if omesh is not None else orc.OMesh.read(path)
    s = rans.Solver(m, gas=rans.Gas(mu_L=mu_L), viscosity_model="spalart-allmaras", second_order=second_order,
                    gradient_scheme=gradient, tile_cells=tile_cells)
    far = rans.BoundaryVariables(mach=0.5, angle=2 * 0.01745)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition(wall)})
    s.init()
    s.refill_bcs()
    q = perturbed_state(s.n_real, s.n_cells, s.get_q(), orc.freestream(0.5, 2 * 0.01745), seed=seed)
    s.set_q(q)
    nu = s.get_nu()
    nu[:s.n_real] = nu_field(s.n_real, mu_L, q[0:4 * s.n_real:4], nu_kind)
    s.set_nu(nu)
    o = orc.OSolver(om, mu_L=mu_L, viscous=2)
    o.set_bcs({"farfield": ("farfield", (0.5, 2 * 0.01745, 1.0, 1.0)), "wall": (wall, (0.2, 0.0, 1.0, 1.0))})
    o.options(second_order=second_order, green_gauss=gradient == "green-gauss", limiter_k=5.0, cfl=1.0)
    o.vec("q", q)
    o.sa_vec("nu", nu)
    return s, o


@pytest.mark.parametrize("path", [TRI, QUAD])
@pytest.mark.parametrize("wall", ["wall", "slip-wall"])
@pytest.mark.parametrize("nu_kind", ["wide", "freestream"])
def test_sa_residual_fp64(path, wall, nu_kind):
    from aeroflex_amd import rans
    s, o = sa_pair(path, wall=wall, nu_kind=nu_kind)
    R, nrm = s.fillRhoRHS(rans.AFX_FP64)
    Rn = s.nu_residual()
    Ro, To = o.sa_fill_rhs()
    assert rel_max(R, Ro) <= FP64_TOL
    assert rel_max(Rn, To) <= FP64_TOL
    assert nrm == pytest.approx(np.sqrt(np.sum(Ro ** 2) + np.sum(To ** 2)), rel=1e-12)
    # set_walls_from_internal on rho nu~: wall ghosts hold the owner's value
    assert np.array_equal(s.get_nu(), o.sa_vec("nu"))


@pytest.mark.parametrize("second_order,gradient", [(True, "least-squares"), (False, "green-gauss")])
def test_sa_residual_fp64_schemes(second_order, gradient):
    from aeroflex_amd import rans
    s, o = sa_pair(QUAD, second_order=second_order, gradient=gradient)
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    Ro, To = o.sa_fill_rhs()
    tol = 1e-10 if gradient == "least-squares" else FP64_TOL   # 2x2 inverse rounding (least squares)
    assert rel_max(R, Ro) <= tol
    assert rel_max(s.nu_residual(), To) <= tol


@pytest.mark.parametrize("tile_cells", [64, 128])
def test_sa_tile_sizes(tile_cells):
    from aeroflex_amd import rans
    s, o = sa_pair(TRI, tile_cells=tile_cells)
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    Ro, To = o.sa_fill_rhs()
    assert rel_max(R, Ro) <= FP64_TOL
    assert rel_max(s.nu_residual(), To) <= FP64_TOL


def test_sa_jittered_mixed_mesh():
    from aeroflex_amd import rans
    m = rans.Mesh.jittered(96, 40, seed=3, quad_frac=0.3)
    path = "/tmp/afx_sa_jitter.msh"
    m.write_msh(path)
    s, o = sa_pair(path, mesh=m, omesh=orc.OMesh.read(path), wall="wall")
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    Ro, To = o.sa_fill_rhs()
    assert rel_max(R, Ro) <= FP64_TOL
    assert rel_max(s.nu_residual(), To) <= FP64_TOL
    R32, _ = s.fillRhoRHS(rans.AFX_FP32)
    assert np.linalg.norm(R32 - Ro) / np.linalg.norm(Ro) <= 1e-4
    assert np.linalg.norm(s.nu_residual() - To) / np.linalg.norm(To) <= 1e-3


@pytest.mark.parametrize("path", [TRI, QUAD])
@pytest.mark.parametrize("nu_kind", ["wide", "freestream"])
def test_sa_residual_fp32(path, nu_kind):
    from aeroflex_amd import rans
    s, o = sa_pair(path, nu_kind=nu_kind)
    R, _ = s.fillRhoRHS(rans.AFX_FP32)
    Rn = s.nu_residual()
    Ro, To = o.sa_fill_rhs()
    assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= 1e-4
    assert np.linalg.norm(Rn - To) / np.linalg.norm(To) <= 1e-3


def test_sa_zero_nu_equals_laminar_with_textbook_face_gradients():
    """rho nu~ = 0: the device SA residual is the oracle's laminar residual
    with the model's face-gradient direction, and its fifth row vanishes."""
    from aeroflex_amd import rans
    s, _ = sa_pair(TRI)
    s.set_nu(np.zeros(s.n_cells))
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    assert np.all(s.nu_residual() == 0.0)
    lam = orc.OSolver(orc.OMesh.read(TRI), mu_L=5e-4, viscous=1)
    lam.set_bcs({"farfield": ("farfield", (0.5, 2 * 0.01745, 1.0, 1.0)), "wall": ("wall", (0.2, 0.0, 1.0, 1.0))})
    lam.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=1.0)
    orc.lib.orc_solver_set_a1(lam.h, 0)
    lam.vec("q", s.get_q())
    lam.fill_rhs()
    assert rel_max(R, lam.vec("RhoVector")) <= FP64_TOL


def test_sa_init_and_refill_freestream_nu():
    from aeroflex_amd import rans
    s = rans.Solver(rans.Mesh.read_file(QUAD), gas=rans.Gas(mu_L=2e-5), viscosity_model="spallart-allmaras")
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", rans.BoundaryVariables(0.5, 0.0)),
               "wall": rans.BoundaryCondition("wall")})
    s.set_turbulence(4.0)
    s.init()
    s.refill_bcs()
    assert np.allclose(s.get_nu(), 4.0 * 2e-5, rtol=0, atol=0)
    with pytest.raises(rans.RansError):   # the matrix-free operator is the four-variable one
        s.set_matrix_free(True)


@pytest.mark.parametrize("path,wall", [(TRI, "wall"), (QUAD, "slip-wall")])
def test_sa_wall_distance_matches_oracle(path, wall):
    """The model's wall distance (to the nearest point of the wall face
    segments) on the device equals the oracle's bit for bit; the reference's
    distance to the face centres stays what afx_rans_wall_dist returns."""
    s, o = sa_pair(path, wall=wall)
    d = s.sa_wall_dist()
    assert np.array_equal(d, o.sa_vec("dist"))
    o.wall_dist()
    assert np.array_equal(s.wall_dist(), o.vec("wall_dist"))
