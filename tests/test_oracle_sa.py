"""Pins of the oracle's Spalart-Allmaras terms.

The reference has no turbulence model (its SA branch is empty,
physics.h:259-261), so these checks come from the model's mathematics and
from the reference's laminar residual:
  * f_w(1) = 1 and the r -> large limit (1 + c_w3^6)^(1/6);
  * the log-layer balance that defines c_w1 (production + cross diffusion +
    diffusion = destruction for nu~ = kappa u_tau d, S = u_tau/(kappa d));
  * the small-chi limit of the source (pure production c_b1 Omega nu~);
  * continuity of the S-tilde limiter at its switch point and the sign of
    the negative-nu~ branch;
  * rho nu~ = 0 reduces the mean-flow rows to the laminar residual bit for
    bit (the reference restatement, with the face-gradient direction of the
    cell connection instead of A1's sum of centroids) and the fifth row to
    zero; that direction is exact for linear fields, A1's is not;
  * first-order convection of a passive nu~ = c is c times the mass residual;
  * a uniform rho nu~ = C with Pr_T = Pr_L gives the laminar residual of
    mu = mu_L + C f_v1(C/mu_L);
  * a uniform rho nu~ on uniform density diffuses nothing.
"""
import numpy as np
import pytest

import orc
from cases import QUAD, TRI, perturbed_state

KAPPA, CB1, CB2, SIGMA, CW3 = 0.41, 0.1355, 0.622, 2.0 / 3.0, 2.0


def test_fw_fixed_points():
    assert orc.lib.orc_sa_fw(1.0) == pytest.approx(1.0, abs=1e-15)
    big = orc.lib.orc_sa_fw(10.0)
    assert big == pytest.approx((1 + CW3 ** 6) ** (1 / 6), rel=1e-12)
    assert orc.lib.orc_sa_fv1(0.0) == 0.0 and orc.lib.orc_sa_fv1(-3.0) == 0.0
    assert orc.lib.orc_sa_fv1(7.1) == pytest.approx(0.5, rel=1e-15)   # chi = c_v1


@pytest.mark.parametrize("utau,d", [(0.05, 0.01), (0.2, 1e-3), (1.0, 0.3)])
def test_log_layer_balance(utau, d):
    """nu~ = kappa u_tau d, Omega = u_tau/(kappa d), grad nu~ = kappa u_tau:
    c_b1 u^2 - c_w1 kappa^2 u^2 + (1 + c_b2)/sigma kappa^2 u^2 = 0."""
    nut = KAPPA * utau * d
    omega = utau / (KAPPA * d)
    src = orc.lib.orc_sa_source(1.0, nut, 1e-14, omega, d, KAPPA * utau, 0.0)
    diffusion = (KAPPA * utau) ** 2 / SIGMA        # div(nu~ grad nu~) / sigma, linear nu~
    production = CB1 * utau ** 2
    assert abs(src + diffusion) <= 1e-6 * production


def test_small_chi_is_pure_production():
    nut, omega = 1e-9, 3.0
    src = orc.lib.orc_sa_source(1.2, nut, 1e-5, omega, 1e6, 0.0, 0.0)
    assert src == pytest.approx(1.2 * CB1 * omega * nut, rel=1e-6)


def test_stilde_limiter_continuity_and_negative_branch():
    # the switch S_bar = -c_v2 Omega: choose nu~, d with f_v2 < 0 and scan Omega
    mu, rho, d = 1e-5, 1.0, 0.05
    nut = 5 * mu
    vals = [orc.lib.orc_sa_source(rho, nut, mu, w, d, 0.0, 0.0) for w in np.linspace(1e-4, 2e-2, 4001)]
    jumps = np.abs(np.diff(vals))
    assert jumps.max() <= 20 * np.median(jumps)
    # negative nu~: destruction changes sign and damps nu~ toward zero
    assert orc.lib.orc_sa_source(rho, -1e-4, mu, 0.0, d, 0.0, 0.0) > 0.0
    assert orc.lib.orc_sa_source(rho, 1e-4, mu, 0.0, d, 0.0, 0.0) < 0.0


def sa_pair(path, second_order=True, gg=True, mu_L=5e-4, wall="wall", seed=2, Pr_T=0.9, viscous=2):
    om = orc.OMesh.read(path)
    o = orc.OSolver(om, mu_L=mu_L, viscous=viscous, Pr_T=Pr_T)
    o.set_bcs({"farfield": ("farfield", (0.5, 2 * 0.01745, 1.0, 1.0)), "wall": (wall, (0.2, 0.0, 1.0, 1.0))})
    o.options(second_order=second_order, green_gauss=gg, limiter_k=5.0, cfl=1.0)
    o.init()
    o.refill_bcs()
    q0 = o.vec("q")
    q = perturbed_state(o.n_real, o.n_cells, q0, orc.freestream(0.5, 2 * 0.01745), seed=seed)
    o.vec("q", q)
    return o, q


@pytest.mark.parametrize("path", [TRI, QUAD])
@pytest.mark.parametrize("second_order,gg", [(True, True), (True, False), (False, True)])
def test_zero_nu_is_laminar(path, second_order, gg):
    o, q = sa_pair(path, second_order, gg)
    o.sa_vec("nu", np.zeros(o.n_cells))
    R, T = o.sa_fill_rhs()
    lam, _ = sa_pair(path, second_order, gg, viscous=1)
    orc.lib.orc_solver_set_a1(lam.h, 0)   # the model's face gradients (textbook direction)
    lam.vec("q", q)
    lam.fill_rhs()
    assert np.array_equal(R, lam.vec("RhoVector"))
    assert np.all(T == 0.0)


@pytest.mark.parametrize("path", [TRI, QUAD])
def test_face_gradient_direction(path):
    """For a linear field the corrected face gradient's derivative along the
    cell connection is exact, whatever the cell gradients; the reference's
    sum-of-centroids direction (A1) is not."""
    o, _ = sa_pair(path, viscous=1)
    m = o.mesh
    xc, yc = m.get("cellsCentersX"), m.get("cellsCentersY")
    a, b, c = np.array([1.2, 0.3, -0.2, 2.5]), np.array([0.05, -0.02, 0.01, 0.03]), np.array([-0.01, 0.04, 0.02, -0.05])
    q = (a[None, :] + np.outer(xc, b) + np.outer(yc, c)).reshape(-1)
    o.vec("q", q)
    o.calc_gradients()
    ec = m.get("edgesCells").reshape(-1, 2)
    errs = {}
    for a1 in (0, 1):
        orc.lib.orc_solver_set_a1(o.h, a1)
        worst = 0.0
        for e in range(0, len(ec), 7):
            i, j = ec[e]
            if j >= o.n_real:
                continue
            t = np.array([xc[j] - xc[i], yc[j] - yc[i]])
            t /= np.linalg.norm(t)
            out = np.zeros(8)
            orc.lib.orc_solver_face_gradient(o.h, int(e), orc._d(out))
            got = out[:4] * t[0] + out[4:] * t[1]
            worst = max(worst, np.max(np.abs(got - (b * t[0] + c * t[1]))))
        errs[a1] = worst
    assert errs[0] <= 1e-10 * np.max(np.abs(b)) and errs[1] > 1e-3 * np.max(np.abs(b))


def test_passive_scalar_convection_first_order():
    o, q = sa_pair(QUAD, second_order=False, wall="slip-wall")
    c = 3e-4
    o.sa_vec("nu", c * q[0::4])
    o.sa_terms(orc.SA_CONV)
    R, T = o.sa_fill_rhs()
    n = o.n_real
    assert np.max(np.abs(T[:n] - c * R[0:4 * n:4])) <= 1e-13 * np.max(np.abs(c * R[0:4 * n:4]))


def test_uniform_mut_is_laminar_with_effective_viscosity():
    mu_L, C = 2e-4, 3e-3
    o, q = sa_pair(TRI, second_order=False, mu_L=mu_L, wall="slip-wall", Pr_T=0.72)
    o.sa_vec("nu", np.full(o.n_cells, C))
    o.sa_terms(orc.SA_MUT)
    R, _ = o.sa_fill_rhs()
    chi = C / mu_L
    mu_eff = mu_L + C * chi ** 3 / (chi ** 3 + 7.1 ** 3)
    lam, _ = sa_pair(TRI, second_order=False, mu_L=mu_eff, wall="slip-wall", viscous=1)
    orc.lib.orc_solver_set_a1(lam.h, 0)
    lam.vec("q", q)
    lam.fill_rhs()
    ref = lam.vec("RhoVector")
    assert np.max(np.abs(R - ref)) <= 1e-12 * np.max(np.abs(ref))
    # and the turbulent viscosity really enters: mu_L alone differs
    plain, _ = sa_pair(TRI, second_order=False, mu_L=mu_L, wall="slip-wall", viscous=1)
    orc.lib.orc_solver_set_a1(plain.h, 0)
    plain.vec("q", q)
    plain.fill_rhs()
    assert np.max(np.abs(R - plain.vec("RhoVector"))) > 1e-6 * np.max(np.abs(ref))


def test_uniform_nu_does_not_diffuse():
    o, q = sa_pair(QUAD, second_order=True, wall="slip-wall")
    qq = q.copy()
    qq[0::4] = 1.3   # uniform density
    o.vec("q", qq)
    o.sa_vec("nu", np.full(o.n_cells, 2e-3))
    o.sa_terms(orc.SA_DIFF)
    _, T = o.sa_fill_rhs()
    assert np.max(np.abs(T)) <= 1e-15 * 2e-3 * 1e3


def test_sa_residual_finite_and_sourced():
    """All terms on a perturbed turbulent state: finite, ghost rows zero, and
    the source is what separates the fifth row from its flux part."""
    o, q = sa_pair(TRI)
    rng = np.random.default_rng(5)
    nu = 3 * 5e-4 * q[0::4] * (1 + 0.01 * rng.standard_normal(o.n_cells))
    o.sa_vec("nu", nu)
    R, T = o.sa_fill_rhs()
    assert np.all(np.isfinite(R)) and np.all(np.isfinite(T))
    assert np.all(T[o.n_real:] == 0)
    o.sa_terms(orc.SA_ALL & ~orc.SA_SOURCE)
    _, T2 = o.sa_fill_rhs()
    assert np.max(np.abs(T - T2)) > 0


def test_sa_time_step_pins():
    """The SA local time step reduces to the reference calc_dt when mu_L and
    nu~ vanish, and equals the hand-evaluated cfl A / (sum lambda len + 4 coef
    sum len^2 / A) on a uniform state."""
    o, q = sa_pair(TRI, mu_L=1e-300, wall="slip-wall")
    o.sa_vec("nu", np.zeros(o.n_cells))
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=0.7)
    dt_sa = o.sa_calc_dt().copy()
    o.calc_dt()
    assert np.array_equal(dt_sa, o.vec("dt"))
    # uniform state: coef = max(4/3, gamma) (mu/Pr_L + mu_t/Pr_T) / rho on every face
    mu = 2e-3
    o2, q2 = sa_pair(QUAD, mu_L=mu, wall="slip-wall")
    qinf = orc.freestream(0.5, 2 * 0.01745)
    qq = np.tile(qinf, o2.n_cells)
    o2.vec("q", qq)
    C = 0.05
    o2.sa_vec("nu", np.full(o2.n_cells, C))
    o2.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=0.7)
    dt = o2.sa_calc_dt()
    m = o2.mesh
    chi = C / mu
    mut = C * chi ** 3 / (chi ** 3 + 7.1 ** 3)
    coef = 1.4 * (mu / 0.72 + mut / 0.9) / qinf[0]
    p = 0.4 * (qinf[3] - 0.5 * (qinf[1] ** 2 + qinf[2] ** 2) / qinf[0])
    c = np.sqrt(1.4 * p / qinf[0])
    ec = m.get("edgesCells").reshape(-1, 2)
    L, nx, ny = m.get("edgesLengths"), m.get("edgesNormalsX"), m.get("edgesNormalsY")
    lam = c + np.abs((qinf[1] * nx + qinf[2] * ny) / qinf[0])
    A = m.get("cellsAreas")
    sc = np.zeros(o2.n_cells)
    sv = np.zeros(o2.n_cells)
    np.add.at(sc, ec[:, 0], lam * L)
    np.add.at(sv, ec[:, 0], coef * L * L)
    two = ec[:, 1] < o2.n_real
    np.add.at(sc, ec[two, 1], (lam * L)[two])
    np.add.at(sv, ec[two, 1], (coef * L * L)[two])
    nr = o2.n_real
    ref = 0.7 * A[:nr] / (sc[:nr] + 4 * sv[:nr] / A[:nr])
    assert np.max(np.abs(dt[:nr] - ref) / ref) <= 1e-12


def test_wall_diffusion_is_the_closed_form_sink():
    """The nu~ = 0 condition of a no-slip wall face removes
    (mu_L/sigma) nu~_c / d_n * len from its cell (d_n = (x_f - x_c).n), and
    nothing else changes."""
    mu = 3e-4
    o, q = sa_pair(TRI, mu_L=mu, wall="wall")
    rng = np.random.default_rng(12)
    nu = 2e-3 * (1 + 0.1 * rng.random(o.n_cells))
    o.sa_vec("nu", nu)
    o.sa_terms(orc.SA_ALL)
    _, T_all = o.sa_fill_rhs()
    o.sa_terms(orc.SA_ALL & ~orc.SA_WALL)
    _, T_nowall = o.sa_fill_rhs()
    m = o.mesh
    ec = m.get("edgesCells").reshape(-1, 2)
    be = m.get("boundaryEdges")
    names = m.boundary_names()
    xc, yc = m.get("cellsCentersX"), m.get("cellsCentersY")
    xf, yf = m.get("edgesCentersX"), m.get("edgesCentersY")
    nx, ny, L = m.get("edgesNormalsX"), m.get("edgesNormalsY"), m.get("edgesLengths")
    qq = o.vec("q")
    expect = np.zeros(o.n_cells)
    for b, e in enumerate(be):
        if names[b] != "wall":
            continue
        c = ec[e, 0]
        dn = (xf[e] - xc[c]) * nx[e] + (yf[e] - yc[c]) * ny[e]
        assert dn > 0
        expect[c] -= mu / (2.0 / 3.0) * (nu[c] / qq[4 * c]) / dn * L[e]
    diff = T_all - T_nowall
    assert np.count_nonzero(expect) > 0
    assert np.max(np.abs(diff - expect)) <= 1e-12 * np.max(np.abs(expect))


@pytest.mark.parametrize("path,wall", [(TRI, "wall"), (QUAD, "slip-wall")])
def test_model_wall_distance_brute_force(path, wall):
    """The model's wall distance (centroid to the nearest point of the wall and
    slip-wall face segments) against a numpy brute force over all face
    segments; never above the reference's distance to the face centres
    (mesh.h:794-830), and below it on first wall cells whose wall face centre
    is off the foot of the normal (on these near-isotropic meshes by little:
    the anisotropic O-grid's wall cells are where it matters)."""
    o, q = sa_pair(path, wall=wall)
    d = o.sa_vec("dist")
    om = o.mesh
    be = om.get("boundaryEdges")
    on_wall = np.array(om.boundary_names()) == "wall"
    en = om.get("edgesNodes").reshape(-1, 2)[be[on_wall]]
    x, y = om.get("nodesX"), om.get("nodesY")
    ax, ay, bx, by = x[en[:, 0]], y[en[:, 0]], x[en[:, 1]], y[en[:, 1]]
    cx, cy = om.get("cellsCentersX"), om.get("cellsCentersY")
    ref = np.empty_like(cx)
    for i in range(len(cx)):
        # parametric foot of the perpendicular, clamped to the segment
        ux, uy = bx - ax, by - ay
        t = np.clip(((cx[i] - ax) * ux + (cy[i] - ay) * uy) / (ux * ux + uy * uy), 0.0, 1.0)
        ref[i] = np.min(np.hypot(cx[i] - ax - t * ux, cy[i] - ay - t * uy))
    assert np.allclose(d, ref, rtol=1e-13, atol=0)
    o.wall_dist()
    dc = o.vec("wall_dist")
    assert np.all(d <= dc * (1 + 1e-14))
    first = np.unique(om.get("edgesCells").reshape(-1, 2)[be[on_wall], 0])
    assert np.sum(d[first] < dc[first] * (1 - 1e-6)) > len(first) // 2
