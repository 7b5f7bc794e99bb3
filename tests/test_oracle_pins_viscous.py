"""Pins of the oracle's laminar viscous flux, its averaged face gradient and
its wall distance, each against an independent formulation:

- the laminar terms of internal_flux (physics.h:230-258) with the pressure,
  temperature and velocity gradient helpers (physics.h:33-82): with
  q_L = q_R the Roe dissipation vanishes, so flux(laminar) - flux(inviscid)
  is the Navier-Stokes viscous flux, here written from the primitive
  variables (rho, u, v, p) and their gradients, with the conservative
  gradients the oracle receives formed by the product rule;
- average_gradients (solver.h:359-398, quirk A1: the direction is the
  normalised SUM of the two centroids) on a two-triangle mesh whose numbers
  are worked by hand;
- compute_wall_dist (mesh.h:794-830) against a k-d tree search of the wall
  and slip-wall face centres (scipy), on both reference meshes.
"""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

import orc

GOLDEN = orc.GOLDEN
GAMMA, RGAS = 1.4, 0.71428571428
BC_INTERNAL = 0


def conservative(rho, u, v, p):
    return np.array([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])


def conservative_gradient(rho, u, v, p, d_rho, d_u, d_v, d_p):
    """d(rho, rho u, rho v, rho E) from the primitive derivatives (one direction)."""
    return np.array([d_rho,
                     d_rho * u + rho * d_u,
                     d_rho * v + rho * d_v,
                     d_p / (GAMMA - 1) + 0.5 * d_rho * (u * u + v * v) + rho * (u * d_u + v * d_v)])


def ns_viscous_flux(rho, u, v, p, gx_prim, gy_prim, nx, ny, mu, Pr, cp=1.0):
    """Viscous part of the face flux, F_v . n, with the sign of physics.h:254-256
    (the flux is F_inviscid - F_viscous).  gx_prim / gy_prim: x / y
    derivatives of (rho, u, v, p).  Stokes' hypothesis, Fourier's law with
    k = cp mu / Pr, T = p / (rho R)."""
    rx, ux, vx, px = gx_prim
    ry, uy, vy, py = gy_prim
    div = ux + vy
    txx = 2 * mu * ux - 2.0 / 3.0 * mu * div
    tyy = 2 * mu * vy - 2.0 / 3.0 * mu * div
    txy = mu * (uy + vx)
    Tx = (px * rho - p * rx) / (rho * rho * RGAS)
    Ty = (py * rho - p * ry) / (rho * rho * RGAS)
    k = cp * mu / Pr
    return -np.array([0.0,
                      nx * txx + ny * txy,
                      nx * txy + ny * tyy,
                      nx * (u * txx + v * txy + k * Tx) + ny * (u * txy + v * tyy + k * Ty)])


def oracle_viscous_part(prim, gx_prim, gy_prim, nx, ny, mu, Pr):
    q = conservative(*prim)
    gx = conservative_gradient(*prim, *gx_prim)
    gy = conservative_gradient(*prim, *gy_prim)
    f_lam = orc.face_flux(BC_INTERNAL, q, q, nx, ny, viscous=1, gx=gx, gy=gy, mu=mu, Pr=Pr)
    f_inv = orc.face_flux(BC_INTERNAL, q, q, nx, ny, viscous=0, gx=gx, gy=gy, mu=mu, Pr=Pr)
    return f_lam - f_inv, f_inv


@pytest.mark.parametrize("angle", [0.0, 0.7, 2.3, -1.9])
def test_laminar_flux_is_navier_stokes(angle):
    """General state and gradients: the oracle's laminar terms equal the
    primitive-variable Navier-Stokes flux to rounding, on several normals."""
    nx, ny = np.cos(angle), np.sin(angle)
    prim = (1.3, 0.4, -0.25, 0.9)
    gxp = (0.2, 0.3, -0.5, 0.05)
    gyp = (-0.1, 0.7, 0.15, -0.2)
    mu, Pr = 0.37, 0.72
    visc, f_inv = oracle_viscous_part(prim, gxp, gyp, nx, ny, mu, Pr)
    want = ns_viscous_flux(*prim, gxp, gyp, nx, ny, mu, Pr)
    assert visc[0] == 0.0
    np.testing.assert_allclose(visc, want, rtol=1e-12, atol=1e-14)
    # with q_L = q_R the inviscid part is the exact Euler flux
    rho, u, v, p = prim
    Vn = u * nx + v * ny
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    np.testing.assert_allclose(f_inv, [rho * Vn, rho * u * Vn + p * nx, rho * v * Vn + p * ny, (E + p) * Vn],
                               rtol=1e-13, atol=1e-15)


def test_laminar_flux_rigid_rotation_has_no_stress():
    """u = -w y, v = w x at uniform p and rho: the strain rate is zero, so the
    momentum rows carry no viscous flux and the energy row none either (T is
    uniform).  A sign error in tau_xy or a missing symmetrisation fails it."""
    w = 0.8
    prim = (1.1, 0.3, -0.2, 1.0)
    gxp = (0.0, 0.0, w, 0.0)      # d/dx of (rho, u, v, p)
    gyp = (0.0, -w, 0.0, 0.0)     # d/dy
    for angle in (0.0, 0.4, 1.3):
        visc, _ = oracle_viscous_part(prim, gxp, gyp, np.cos(angle), np.sin(angle), 0.5, 0.72)
        np.testing.assert_allclose(visc, 0.0, atol=1e-14)


def test_laminar_flux_simple_shear_and_conduction():
    """Closed forms worked by hand.  Couette shear u = s y on a face with
    normal (0, 1): the x-momentum row loses mu s and the energy row u mu s.
    Conduction alone (p_x != 0 at rest): the energy row loses
    (mu / Pr) p_x / (rho R) n_x, because c_p = 1 (core.h:32)."""
    mu, Pr, s = 0.25, 0.72, 2.0
    rho, u, v, p = 1.0, 0.5, 0.0, 1.0
    visc, _ = oracle_viscous_part((rho, u, v, p), (0, 0, 0, 0), (0, s, 0, 0), 0.0, 1.0, mu, Pr)
    np.testing.assert_allclose(visc, [0.0, -mu * s, 0.0, -u * mu * s], rtol=1e-14, atol=1e-16)
    px = 0.3
    visc, _ = oracle_viscous_part((rho, 0.0, 0.0, p), (0, 0, 0, px), (0, 0, 0, 0), 1.0, 0.0, mu, Pr)
    np.testing.assert_allclose(visc, [0.0, 0.0, 0.0, -(mu / Pr) * px / (rho * RGAS)], rtol=1e-13, atol=1e-16)
    # the normal stress of a uniaxial stretch u = s x: tau_xx = 4/3 mu s
    visc, _ = oracle_viscous_part((rho, 0.0, 0.0, p), (0, s, 0, 0), (0, 0, 0, 0), 1.0, 0.0, mu, Pr)
    np.testing.assert_allclose(visc, [0.0, -4.0 / 3.0 * mu * s, 0.0, 0.0], rtol=1e-14, atol=1e-16)


def two_triangles():
    """Triangle A = (0,0), (3,0), (0,3) (centroid (1,1)) and triangle B =
    (3,0), (3,6), (0,3) (centroid (2,3)) sharing the edge (3,0)-(0,3); A comes
    first in the file, so A owns the shared face."""
    X = np.array([0.0, 3.0, 0.0, 3.0])
    Y = np.array([0.0, 0.0, 3.0, 6.0])
    cells = np.array([[0, 1, 2, 0], [1, 3, 2, 0]])
    bsegs = np.array([[0, 1], [2, 0], [1, 3], [3, 2]])
    bkeys = np.array([1001, 1001, 1001, 1001])
    return orc.OMesh.build(X, Y, cells, [3, 3], bsegs, bkeys, {1001: "farfield", 2002: "internal"})


def test_average_gradients_two_cell_hand_case():
    """solver.h:359-398 with quirk A1.  Centroids (1,1) and (2,3): their SUM
    is (3,4), length 5, t = (0.6, 0.8).  q_A - q_B = 5 gives the directional
    derivative 1; mean gradient (2, 1) has t-component 2.  The corrected
    gradient is (2,1) - (2 - 1)(0.6, 0.8) = (1.4, 0.2).  Second variable:
    q_A - q_B = -2.5 (derivative -0.5), mean gradient (-1, 3) with
    t-component 1.8: (-1, 3) - 2.3 (0.6, 0.8) = (-2.38, 1.16)."""
    m = two_triangles()
    np.testing.assert_allclose(m.get("cellsCentersX")[:2], [1.0, 2.0])
    np.testing.assert_allclose(m.get("cellsCentersY")[:2], [1.0, 3.0])
    ec = m.get("edgesCells").reshape(-1, 2)
    nr = m.counts()["real_cells"]
    shared = [e for e in range(len(ec)) if ec[e, 0] < nr and ec[e, 1] < nr]
    assert len(shared) == 1 and tuple(ec[shared[0]]) == (0, 1)
    o = orc.OSolver(m, viscous=1)
    o.set_bcs({"farfield": ("farfield", (0.5, 0.0, 1.0, 1.0))})
    nt = m.counts()["cells"]
    q = np.zeros(4 * nt)
    gx = np.zeros(4 * nt)
    gy = np.zeros(4 * nt)
    # variable 0: q_A = 7, q_B = 2; gradients (1, 0) and (3, 2): mean (2, 1)
    q[0], q[4] = 7.0, 2.0
    gx[0], gy[0], gx[4], gy[4] = 1.0, 0.0, 3.0, 2.0
    # variable 1: q_A = 1, q_B = 3.5; gradients (-2, 3) and (0, 3): mean (-1, 3)
    q[1], q[5] = 1.0, 3.5
    gx[1], gy[1], gx[5], gy[5] = -2.0, 3.0, 0.0, 3.0
    o.vec("q", q)
    o.vec("gx", gx)
    o.vec("gy", gy)
    out = np.zeros(8)
    orc.lib.orc_solver_face_gradient(o.h, shared[0], orc._d(out))
    np.testing.assert_allclose([out[0], out[4]], [1.4, 0.2], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose([out[1], out[5]], [-2.38, 1.16], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(out[[2, 3, 6, 7]], 0.0, atol=0.0)


def test_average_gradients_keeps_normal_part():
    """Property of the A1 correction: along t the corrected gradient equals
    the difference quotient, across t it keeps the mean gradient.  Random
    gradients on the two-triangle mesh."""
    m = two_triangles()
    o = orc.OSolver(m, viscous=1)
    o.set_bcs({"farfield": ("farfield", (0.5, 0.0, 1.0, 1.0))})
    nt = m.counts()["cells"]
    rng = np.random.default_rng(11)
    q = rng.normal(size=4 * nt)
    gx = rng.normal(size=4 * nt)
    gy = rng.normal(size=4 * nt)
    o.vec("q", q)
    o.vec("gx", gx)
    o.vec("gy", gy)
    out = np.zeros(8)
    ec = m.get("edgesCells").reshape(-1, 2)
    e = [i for i in range(len(ec)) if tuple(ec[i]) == (0, 1)][0]
    orc.lib.orc_solver_face_gradient(o.h, e, orc._d(out))
    t = np.array([3.0, 4.0]) / 5.0
    for k in range(4):
        g = np.array([out[k], out[4 + k]])
        gm = 0.5 * np.array([gx[k] + gx[4 + k], gy[k] + gy[4 + k]])
        assert abs(g @ t - (q[k] - q[4 + k]) / 5.0) < 1e-14
        perp = np.array([-t[1], t[0]])
        assert abs(g @ perp - gm @ perp) < 1e-14


def wall_dist_case(path, walls):
    m = orc.OMesh.read(path)
    names = sorted(set(m.boundary_names()))
    bcs = {}
    for n in names:
        bcs[n] = (walls, (0.2, 0.0, 1.0, 1.0)) if n == "wall" else ("farfield", (0.5, 0.0, 1.0, 1.0))
    o = orc.OSolver(m)
    o.set_bcs(bcs)
    o.wall_dist()
    return m, o, o.vec("wall_dist")


@pytest.mark.parametrize("fixture", ["naca0012q_coarse.msh", "naca0012_coarse.msh"])
@pytest.mark.parametrize("walls", ["wall", "slip-wall"])
def test_wall_distance_against_kd_tree(fixture, walls):
    """mesh.h:794-830: every cell (ghosts included) gets the distance from its
    centroid to the nearest centre of a wall or slip-wall face; farfield
    faces do not count."""
    m, o, d = wall_dist_case(os.path.join(GOLDEN, fixture), walls)
    cx, cy = m.get("cellsCentersX"), m.get("cellsCentersY")
    ex, ey = m.get("edgesCentersX"), m.get("edgesCentersY")
    be = m.get("boundaryEdges")
    bn = m.boundary_names()
    wf = np.array([be[b] for b in range(len(be)) if bn[b] == "wall"])
    assert 0 < len(wf) < len(be)
    tree = cKDTree(np.c_[ex[wf], ey[wf]])
    want, _ = tree.query(np.c_[cx, cy])
    assert len(d) == len(cx)
    np.testing.assert_allclose(d, want, rtol=1e-15, atol=0.0)
    # the wall ghosts sit at half their owner's distance mirrored: a wall
    # ghost's nearest wall face centre is its own face centre
    gh = np.array([m.counts()["real_cells"] + b for b in range(len(be)) if bn[b] == "wall"])
    gdist = np.hypot(cx[gh] - ex[wf], cy[gh] - ey[wf])
    assert np.all(d[gh] <= gdist * (1 + 1e-15))


def test_wall_distance_without_walls_is_one():
    """With no wall or slip-wall patch the distance keeps its initial value 1
    (mesh.h:799)."""
    path = os.path.join(GOLDEN, "naca0012q_coarse.msh")
    m = orc.OMesh.read(path)
    o = orc.OSolver(m)
    o.set_bcs({n: ("farfield", (0.5, 0.0, 1.0, 1.0)) for n in set(m.boundary_names())})
    o.wall_dist()
    d = o.vec("wall_dist")
    assert np.all(d == 1.0)
