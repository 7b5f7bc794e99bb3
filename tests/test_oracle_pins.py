"""Pins of the oracle (CPU FP64 restatement of src/rans) against closed forms,
invariants and independent formulations.  No golden vectors exist in the
reference (it has no tests), so every check below derives its expected value
from the mathematics or from the reference's mesh fixtures, never by retyping
the oracle's formula."""
import math
import os

import numpy as np
import pytest

import orc
from cases import QUAD, TRI

G = 1.4
R = 0.71428571428


# ----------------------------------------------------------------- helpers --
def euler_flux(q, nx, ny, g=G):
    rho, ru, rv, re = q
    u, v = ru / rho, rv / rho
    p = (g - 1) * (re - 0.5 * rho * (u * u + v * v))
    V = u * nx + v * ny
    return np.array([rho * V, ru * V + p * nx, rv * V + p * ny, (re + p) * V])


def prim_to_cons(rho, u, v, p, g=G):
    return np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])


def random_state(rng):
    return prim_to_cons(rng.uniform(0.5, 2), rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), rng.uniform(0.5, 2))


def strip_mesh(xs, height=1.0, reverse=False):
    """One row of rectangles between the x positions xs (non-uniform), walls on
    top and bottom, farfield at both ends.  Cells are listed left to right
    (reverse=True: right to left)."""
    n = len(xs) - 1
    X = np.concatenate([xs, xs])
    Y = np.concatenate([np.zeros(n + 1), np.full(n + 1, height)])
    cells = [[k, k + 1, n + 2 + k, n + 1 + k] for k in range(n)]
    if reverse:
        cells = cells[::-1]
    bsegs, bkeys = [], []
    for k in range(n):
        bsegs += [[k, k + 1], [n + 1 + k, n + 2 + k]]
        bkeys += [1002, 1002]
    bsegs += [[0, n + 1], [n, 2 * n + 1]]
    bkeys += [1001, 1001]
    names = {1001: "farfield", 1002: "wall", 2003: "internal"}
    return X, Y, np.array(cells), np.full(n, 4), np.array(bsegs), np.array(bkeys), names


def solver_on(mesh, wall="slip-wall", mach=0.5, angle=0.0, viscous=0, **opts):
    o = orc.OSolver(mesh, viscous=viscous)
    o.set_bcs({"farfield": ("farfield", (mach, angle, 1.0, 1.0)), "wall": (wall, (0.2, 0.0, 1.0, 1.0))})
    o.options(**opts)
    return o


def msh_header_counts(path):
    """Node and element counts straight from the $Nodes / $Elements headers."""
    with open(path) as f:
        lines = f.read().splitlines()
    n_nodes = int(lines[lines.index("$Nodes") + 1].split()[1])
    i = lines.index("$Elements")
    nblocks = int(lines[i + 1].split()[0])
    j = i + 2
    n_lines = n_cells = 0
    for _ in range(nblocks):
        dim, _, etype, n = map(int, lines[j].split())
        if dim == 1:
            n_lines += n
        elif dim == 2:
            n_cells += n
        j += n + 1
    return n_nodes, n_lines, n_cells


# ------------------------------------------------------------------- mesh --
@pytest.mark.parametrize("path", [QUAD, TRI])
def test_mesh_counts_match_file(path):
    m = orc.OMesh.read(path)
    c = m.counts()
    nn, nl, nc = msh_header_counts(path)
    assert c["nodes"] == nn and c["real_cells"] == nc and c["bfaces"] == nl
    assert c["cells"] == nc + nl
    tri = m.get("cellsIsTriangle")[:nc]
    # every edge is shared by two cells or is a boundary segment
    assert c["faces"] == (3 * tri.sum() + 4 * (nc - tri.sum()) + nl) // 2
    # annulus: Euler characteristic V - E + F = 0
    assert c["nodes"] - c["faces"] + c["real_cells"] == 0


def test_quad_fixture_counts():
    c = orc.OMesh.read(QUAD).counts()
    assert (c["nodes"], c["real_cells"], c["bfaces"]) == (4160, 4096, 128)


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_cells_are_closed_and_normals_point_away_from_owner(path):
    m = orc.OMesh.read(path)
    c = m.counts()
    ec = m.get("edgesCells").reshape(-1, 2).astype(np.int64)
    nx, ny, L = m.get("edgesNormalsX"), m.get("edgesNormalsY"), m.get("edgesLengths")
    xf, yf = m.get("edgesCentersX"), m.get("edgesCentersY")
    xc, yc = m.get("cellsCentersX"), m.get("cellsCentersY")
    sx = np.zeros(c["cells"])
    sy = np.zeros(c["cells"])
    np.add.at(sx, ec[:, 0], nx * L)
    np.add.at(sy, ec[:, 0], ny * L)
    np.add.at(sx, ec[:, 1], -nx * L)
    np.add.at(sy, ec[:, 1], -ny * L)
    nr = c["real_cells"]
    per = np.zeros(c["cells"])
    np.add.at(per, ec[:, 0], L)
    np.add.at(per, ec[:, 1], L)
    assert np.max(np.abs(sx[:nr]) / per[:nr]) < 1e-12
    assert np.max(np.abs(sy[:nr]) / per[:nr]) < 1e-12
    assert np.all(nx * (xf - xc[ec[:, 0]]) + ny * (yf - yc[ec[:, 0]]) > 0)
    assert np.allclose(nx ** 2 + ny ** 2, 1.0, atol=1e-14)


def polygon_area(x, y):
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def boundary_loop(m, name):
    """Ordered node loop of one boundary patch (from the face node pairs)."""
    names = m.boundary_names()
    en = m.get("edgesNodes").reshape(-1, 2)
    be = m.get("boundaryEdges")
    segs = [tuple(en[be[b]]) for b in range(len(be)) if names[b] == name]
    nxt = {}
    for a, b in segs:
        nxt.setdefault(a, []).append(b)
        nxt.setdefault(b, []).append(a)
    start = segs[0][0]
    loop, prev, cur = [start], None, start
    while True:
        cand = [v for v in nxt[cur] if v != prev]
        prev, cur = cur, cand[0]
        if cur == start:
            break
        loop.append(cur)
    return np.array(loop)


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_areas_sum_to_domain_area(path):
    m = orc.OMesh.read(path)
    X, Y = m.get("nodesX"), m.get("nodesY")
    outer = boundary_loop(m, "farfield")
    inner = boundary_loop(m, "wall")
    dom = polygon_area(X[outer], Y[outer]) - polygon_area(X[inner], Y[inner])
    nr = m.counts()["real_cells"]
    assert abs(m.get("cellsAreas")[:nr].sum() - dom) <= 1e-10 * dom


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_face_order_owner_rule(path):
    """Face f is owned by the lowest-index cell containing its edge, and faces
    are numbered in order of (owner, position of the edge in the owner)."""
    m = orc.OMesh.read(path)
    nr = m.counts()["real_cells"]
    cn = m.get("cellsNodes").reshape(-1, 4)[:nr]
    tri = m.get("cellsIsTriangle")[:nr]
    first = {}
    for c in range(nr):
        k = 3 if tri[c] else 4
        for i in range(k):
            a, b = cn[c, i], cn[c, (i + 1) % k]
            first.setdefault((min(a, b), max(a, b)), (c, i, a, b))
    expect = sorted(first.values())
    en = m.get("edgesNodes").reshape(-1, 2)
    ec = m.get("edgesCells").reshape(-1, 2)
    assert len(expect) == len(en)
    assert np.array_equal(en, np.array([[e[2], e[3]] for e in expect]))
    assert np.array_equal(ec[:, 0], np.array([e[0] for e in expect]))


def test_ghost_cells_mirror_owner():
    m = orc.OMesh.read(TRI)
    c = m.counts()
    nr = c["real_cells"]
    be = m.get("boundaryEdges")
    ec = m.get("edgesCells").reshape(-1, 2)
    xc, yc, A = m.get("cellsCentersX"), m.get("cellsCentersY"), m.get("cellsAreas")
    xf, yf = m.get("edgesCentersX")[be], m.get("edgesCentersY")[be]
    nx, ny = m.get("edgesNormalsX")[be], m.get("edgesNormalsY")[be]
    own, gh = ec[be, 0], ec[be, 1]
    assert np.array_equal(gh, nr + np.arange(len(be)))
    d_own = np.hypot(xf - xc[own], yf - yc[own])
    assert np.allclose(xc[gh] - xf, d_own * nx, atol=1e-14) and np.allclose(yc[gh] - yf, d_own * ny, atol=1e-14)
    assert np.array_equal(A[gh], A[own])


# ------------------------------------------------------------------- flux ---
def test_roe_consistency_gives_euler_flux():
    rng = np.random.default_rng(0)
    for _ in range(50):
        q = random_state(rng)
        t = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(t), np.sin(t)
        f = orc.face_flux(0, q, q, nx, ny)
        assert np.allclose(f, euler_flux(q, nx, ny), rtol=1e-12, atol=1e-12)


def rot(q, t):
    c, s = np.cos(t), np.sin(t)
    return np.array([q[0], c * q[1] - s * q[2], s * q[1] + c * q[2], q[3]])


def test_roe_rotation_invariance_and_conservation():
    rng = np.random.default_rng(1)
    for _ in range(50):
        qL, qR = random_state(rng), random_state(rng)
        t, a = rng.uniform(0, 2 * np.pi, 2)
        nx, ny = np.cos(t), np.sin(t)
        f = orc.face_flux(0, qL, qR, nx, ny)
        fr = orc.face_flux(0, rot(qL, a), rot(qR, a), np.cos(t + a), np.sin(t + a))
        assert np.allclose(fr, rot(f, a), rtol=1e-11, atol=1e-12)
        # the flux seen from the other side is the negative
        fb = orc.face_flux(0, qR, qL, -nx, -ny)
        assert np.allclose(fb, -f, rtol=1e-11, atol=1e-12)


def test_roe_upwinding_supersonic():
    """For a supersonic flow in +n on both sides, every wave speed is positive
    and above the entropy-fix threshold, so by the Roe property
    A(q_R - q_L) = F(q_R) - F(q_L) the flux is the left state's flux, up to the
    smoothing |x| -> sqrt(x^2 + 1e-4) of physics.h:84-86 (relative 1e-4/lambda^2)."""
    qL = prim_to_cons(1.0, 3.0, 0.1, 1.0)
    qR = prim_to_cons(1.1, 3.1, 0.0, 1.05)
    f = orc.face_flux(0, qL, qR, 1.0, 0.0)
    fl, fr = euler_flux(qL, 1.0, 0.0), euler_flux(qR, 1.0, 0.0)
    assert np.max(np.abs(f - fl)) < 1e-4 * np.max(np.abs(fr - fl))
    # with the normal reversed the flow runs from R to L: the right state's flux
    f2 = orc.face_flux(0, qL, qR, -1.0, 0.0)
    assert np.max(np.abs(f2 - euler_flux(qR, -1.0, 0.0))) < 1e-4 * np.max(np.abs(fr - fl))


def test_wall_state_functors():
    rng = np.random.default_rng(3)
    q = random_state(rng)
    t = rng.uniform(0, 2 * np.pi)
    nx, ny = np.cos(t), np.sin(t)
    qs = orc.face_vars(2, q, q, nx, ny)    # slip: mirrored normal momentum
    assert np.isclose(qs[1] * nx + qs[2] * ny, -(q[1] * nx + q[2] * ny))
    assert np.isclose(-qs[1] * ny + qs[2] * nx, -q[1] * ny + q[2] * nx)
    qw = orc.face_vars(3, q, q, nx, ny)    # no-slip: reversed momentum
    assert np.array_equal(qw, [q[0], -q[1], -q[2], q[3]])
    # slip wall flux carries no mass and no energy, only pressure
    f = orc.face_flux(2, q, q, nx, ny)
    p = (G - 1) * (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0])
    assert abs(f[0]) < 1e-12 and abs(f[3]) < 1e-12
    assert np.allclose([f[1] / nx, f[2] / ny], 0, atol=10) and f[1] * nx + f[2] * ny > 0.5 * p


def test_farfield_freestream_is_fixed_point():
    for mach, ang in [(0.5, 0.03), (0.3, -0.5), (1.5, 0.2)]:
        qinf = orc.freestream(mach, ang)
        for t in np.linspace(0, 2 * np.pi, 9):
            qb = orc.face_vars(1, qinf, qinf, np.cos(t), np.sin(t))
            assert np.allclose(qb, qinf, rtol=1e-13, atol=1e-13)


def test_freestream_matches_closed_form():
    q = orc.freestream(0.5, 0.0349)
    c = math.sqrt(G * R * 1.0)
    rho = 1.0 / R
    assert np.allclose(q, [rho, rho * 0.5 * c * math.cos(0.0349), rho * 0.5 * c * math.sin(0.0349),
                           1.0 / (G - 1) + 0.5 * rho * (0.5 * c) ** 2], rtol=1e-15)


# ---------------------------------------------------------------- residual --
@pytest.mark.parametrize("second_order", [True, False])
def test_freestream_preservation(second_order):
    """A uniform freestream gives R_i = 0 in every cell away from the slip
    wall: first order, every cell not touching it; second order, every cell
    not touching a wall cell either (the mirrored momentum gives the wall
    cells a gradient that their neighbours reconstruct with)."""
    m = orc.OMesh.read(QUAD)
    o = solver_on(m, mach=0.5, angle=0.0349, second_order=second_order)
    o.init()
    o.refill_bcs()
    o.fill_rhs()
    Rv = o.vec("RhoVector").reshape(-1, 4)
    names = m.boundary_names()
    be = m.get("boundaryEdges")
    ec = m.get("edgesCells").reshape(-1, 2)
    wall_cells = {int(ec[be[b], 0]) for b in range(len(be)) if names[b] == "wall"}
    if second_order:
        near = set(wall_cells)
        for f in range(len(ec)):
            a, b = int(ec[f, 0]), int(ec[f, 1])
            if a in wall_cells or b in wall_cells:
                near.update((a, b))
        wall_cells = near
    keep = np.array([i not in wall_cells for i in range(m.counts()["real_cells"])])
    scale = np.abs(orc.freestream(0.5, 0.0349)).max() * m.get("edgesLengths").max()
    assert np.max(np.abs(Rv[: len(keep)][keep])) <= 1e-13 * scale


def test_green_gauss_closed_form_on_stretched_strip():
    """Face value q_o(1-g) + (q_o+q_n)/2 g (solver.h:450, physics.h:275) on a
    strip of rectangles of widths h_i with owners on the left: for q = b x the
    interior gradient is b (h_{i-1} + 3 h_i) / (4 h_i)."""
    rng = np.random.default_rng(4)
    h = rng.uniform(0.5, 2.0, 12)
    xs = np.concatenate([[0.0], np.cumsum(h)])
    m = orc.OMesh.build(*strip_mesh(xs))
    o = solver_on(m)
    nt = m.counts()["cells"]
    nr = m.counts()["real_cells"]
    xc = m.get("cellsCentersX")
    b = 0.3
    q = np.tile([1.0, 0.1, 0.0, 2.5], nt)
    q[0::4] = 1.0 + b * xc
    o.vec("q", q)
    o.calc_gradients()
    gx = o.vec("gx").reshape(-1, 4)[:, 0]
    gy = o.vec("gy").reshape(-1, 4)[:, 0]
    i = np.arange(1, nr - 1)
    assert np.allclose(gx[i], b * (h[i - 1] + 3 * h[i]) / (4 * h[i]), rtol=1e-13)
    assert np.allclose(gy[:nr], 0, atol=1e-13)
    # ownership reversed (cells listed right to left): owners are on the right
    # and the same rule gives b (3 h_i + h_{i+1}) / (4 h_i)
    m2 = orc.OMesh.build(*strip_mesh(xs, reverse=True))
    o2 = solver_on(m2)
    xc2 = m2.get("cellsCentersX")
    q2 = np.tile([1.0, 0.1, 0.0, 2.5], nt)
    q2[0::4] = 1.0 + b * xc2
    o2.vec("q", q2)
    o2.calc_gradients()
    gx2 = o2.vec("gx").reshape(-1, 4)[:, 0]
    order = np.argsort(xc2[:nr])
    g_sorted = gx2[:nr][order]
    assert np.allclose(g_sorted[i], b * (3 * h[i] + h[i + 1]) / (4 * h[i]), rtol=1e-13)


def test_least_squares_is_minus_half_gradient():
    """Delta = q_i - (q_i+q_n)/2 (solver.h:496-500): for a linear field, cells
    whose neighbours are all real cells get exactly -1/2 of the gradient."""
    m = orc.OMesh.read(TRI)
    o = solver_on(m, green_gauss=False)
    c = m.counts()
    nr, nt = c["real_cells"], c["cells"]
    xc, yc = m.get("cellsCentersX"), m.get("cellsCentersY")
    bx, by = 0.7, -0.4
    q = np.tile([1.0, 0.2, 0.1, 2.5], nt)
    q[0::4] = 2.0 + bx * xc + by * yc
    o.vec("q", q)
    o.calc_gradients()
    ce = m.get("cellsEdges").reshape(-1, 4)[:nr]
    ec = m.get("edgesCells").reshape(-1, 2)
    interior = []
    for i in range(nr):
        fs = ce[i][ce[i] != 0xFFFFFFFF]
        if np.all(ec[fs, 1] < nr):
            interior.append(i)
    gx = o.vec("gx").reshape(-1, 4)[interior, 0]
    gy = o.vec("gy").reshape(-1, 4)[interior, 0]
    assert len(interior) > 3000
    assert np.allclose(gx, -0.5 * bx, rtol=1e-9) and np.allclose(gy, -0.5 * by, rtol=1e-9)


def test_limiter_reduces_to_venkatakrishnan_at_k0():
    """With limiter_k = 0 the limiter of solver.h:578-584 is
    min over faces of V(y) = (y^2+2y)/(y^2+y+2), y = Delta/dqg (physics.h:605)."""
    m = orc.OMesh.read(TRI)
    o = solver_on(m, limiter_k=0.0)
    c = m.counts()
    nr, nt = c["real_cells"], c["cells"]
    o.init()
    o.refill_bcs()
    rng = np.random.default_rng(5)
    q = o.vec("q")
    q[: 4 * nr] *= 1 + 0.02 * rng.standard_normal(4 * nr)
    o.vec("q", q)
    o.set_walls()
    o.calc_gradients()
    o.calc_limiters()
    q = o.vec("q").reshape(-1, 4)
    gx, gy = o.vec("gx").reshape(-1, 4), o.vec("gy").reshape(-1, 4)
    lim = o.vec("limiters").reshape(-1, 4)
    ce = m.get("cellsEdges").reshape(-1, 4)[:nr]
    ec = m.get("edgesCells").reshape(-1, 2)
    xf, yf = m.get("edgesCentersX"), m.get("edgesCentersY")
    xc, yc = m.get("cellsCentersX"), m.get("cellsCentersY")
    V = lambda y: (y * y + 2 * y) / (y * y + y + 2)
    for i in rng.choice(nr, 400, replace=False):
        fs = ce[i][ce[i] != 0xFFFFFFFF]
        nb = np.where(ec[fs, 0] == i, ec[fs, 1], ec[fs, 0])
        qmax = np.maximum(q[i], q[nb].max(0))
        qmin = np.minimum(q[i], q[nb].min(0))
        expect = np.ones(4)
        for f in fs:
            d = gx[i] * (xf[f] - xc[i]) + gy[i] * (yf[f] - yc[i])
            for k in range(4):
                if abs(d[k]) > 1e-16:
                    D = (qmax[k] if d[k] > 0 else qmin[k]) - q[i, k]
                    expect[k] = min(expect[k], V(D / d[k]))
        assert np.allclose(lim[i], expect, rtol=1e-10, atol=1e-13)


def test_dt_closed_form_uniform_state():
    m = orc.OMesh.read(TRI)
    o = solver_on(m, cfl=7.0)
    o.init()
    o.refill_bcs()
    o.calc_dt()
    dt = o.vec("dt")
    q = orc.freestream(0.5, 0.0)
    c = math.sqrt(G * 1.0 * R)   # T = 1
    u, v = q[1] / q[0], q[2] / q[0]
    nx, ny, L = m.get("edgesNormalsX"), m.get("edgesNormalsY"), m.get("edgesLengths")
    ec = m.get("edgesCells").reshape(-1, 2)
    eig = (c + np.abs(u * nx + v * ny)) * L
    nr = m.counts()["real_cells"]
    s = np.zeros(m.counts()["cells"])
    np.add.at(s, ec[:, 0], eig)
    interior = ec[:, 1] < nr
    np.add.at(s, ec[interior, 1], eig[interior])
    A = m.get("cellsAreas")
    assert np.allclose(dt[:nr], 7.0 * A[:nr] / s[:nr], rtol=1e-13)
    assert np.all(np.isinf(dt[nr:]))


def test_explicit_step_is_rk3_of_the_residual():
    """explicitSolver::solve (solver.h:802-828) at first order equals the
    three-stage update q + R(q_k)/A dt alpha_s relax, alpha = 1/4, 1/2, 1,
    composed from fillRhoRHS (solver.h:1079-1152)."""
    m = orc.OMesh.read(TRI)
    nr, nt = m.counts()["real_cells"], m.counts()["cells"]
    rng = np.random.default_rng(6)
    a = solver_on(m, second_order=False, cfl=0.8)
    a.init()
    a.refill_bcs()
    q0 = a.vec("q")
    q0[: 4 * nr] *= 1 + 0.01 * rng.standard_normal(4 * nr)
    a.vec("q", q0)
    a.calc_dt()
    dt = a.vec("dt")
    A = m.get("cellsAreas")
    b = solver_on(m, second_order=False, cfl=0.8)
    qk = q0.copy()
    relax = 0.9
    for alpha in (0.25, 0.5, 1.0):
        b.vec("q", qk)
        b.fill_rhs()
        Rk = b.vec("RhoVector").reshape(-1, 4)
        W = Rk[:nr] / A[:nr, None]
        new = qk.reshape(-1, 4).copy()
        new[:nr] = q0.reshape(-1, 4)[:nr] + W * dt[:nr, None] * alpha * relax
        qk = new.reshape(-1)
    a.vec("q", q0)
    a.explicit_step(relax)
    assert np.allclose(a.vec("q")[: 4 * nr], qk[: 4 * nr], rtol=1e-14, atol=1e-15)


def test_uniform_residual_and_stale_accumulation():
    """get_uniform_residual (solver.h:636-690) is the first-order residual of
    the uniform freestream; a second call accumulates onto the stale qW (A8)
    and returns exactly twice the first value."""
    m = orc.OMesh.read(QUAD)
    o = solver_on(m, mach=0.5, angle=0.0349, second_order=False)
    e1 = o.uniform_residual()
    e2 = o.uniform_residual()
    p = solver_on(m, mach=0.5, angle=0.0349, second_order=False)
    p.init()
    p.refill_bcs()
    p.fill_rhs()
    Rv = p.vec("RhoVector")
    assert e1 > 0 and abs(e1 - np.linalg.norm(Rv)) <= 1e-13 * e1
    assert e2 == 2 * e1


def test_uniform_residual_slip_wall_closed_form():
    """get_uniform_residual on a channel of rectangles (slip walls on top and
    bottom, farfield at both ends) with the freestream at angle a: exact
    freestream fluxes cancel around each closed cell, so only the wall faces
    leave a residual.  At a slip wall the ghost mirrors the normal velocity
    (u_n -> -u_n), the Roe average has u~_n = 0, H~ = H and sound speed
    c~^2 = c^2 + (gamma - 1) u_n^2 / 2, and only the two acoustic waves carry
    the jump, with eigenvalue magnitude lam = sabs(c~) = sqrt(c~^2 + 1e-4)
    (the smoothed |.| of physics.h:84, no Harten correction at that size):
    F_wall = (0, (p + rho u_n^2 + rho u_n lam) n, 0).  Top and bottom walls
    together leave -2 rho v lam L on the y-momentum row of each cell of width
    L, so |R| = 2 rho v lam |L| with v = V sin(a)."""
    xs = np.cumsum(np.r_[0.0, np.random.default_rng(5).uniform(0.3, 1.7, 9)])
    m = orc.OMesh.build(*strip_mesh(xs, height=0.8))
    mach, a, gamma = 0.6, 0.3, 1.4
    o = solver_on(m, mach=mach, angle=a, second_order=False)
    e1 = o.uniform_residual()
    q = orc.freestream(mach, a)
    rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
    p = (gamma - 1) * (q[3] - 0.5 * rho * (u * u + v * v))
    c2 = gamma * p / rho
    ct2 = c2 + 0.5 * (gamma - 1) * v * v
    lam = np.sqrt(ct2 + 1e-4)
    L = np.diff(xs)
    expect = 2 * rho * v * lam * np.linalg.norm(L)
    assert abs(e1 - expect) <= 1e-12 * expect, (e1, expect)
    # the plausible slips (the cell sound speed c for c~, |c~| unsmoothed)
    # are far outside that tolerance at this angle
    for wrong in (np.sqrt(c2 + 1e-4), np.sqrt(ct2)):
        assert abs(2 * rho * v * wrong * np.linalg.norm(L) - expect) > 1e-5 * expect


def test_wall_profile_lower_surface_pressure():
    """cp = 1 on the lower surface, 0 elsewhere, alpha = 0: the lift is the
    lower surface's x-extent over the chord, the drag its (closed) y-extent."""
    m = orc.OMesh.read(TRI)
    o = solver_on(m, mach=0.5, angle=0.0)
    o.init()
    o.refill_bcs()
    q = o.vec("q").reshape(-1, 4)
    names = m.boundary_names()
    be = m.get("boundaryEdges")
    ec = m.get("edgesCells").reshape(-1, 2)
    ny = m.get("edgesNormalsY")
    en = m.get("edgesNodes").reshape(-1, 2)
    X = m.get("nodesX")
    qinf = orc.freestream(0.5, 0.0)
    cp1 = prim_to_cons(qinf[0], qinf[1] / qinf[0], qinf[2] / qinf[0], 1.0 + 0.5 * G * 0.25)   # p/pinf - 1 = gM^2/2
    wall = [b for b in range(len(be)) if names[b] == "wall"]
    lower = [b for b in wall if ny[be[b]] > 0]
    for b in lower:
        q[ec[be[b], 0]] = cp1
    o.vec("q", q.reshape(-1))
    cl, cd, cm = o.wall_profile("wall")
    xf = m.get("edgesCentersX")[be[wall]]
    chord = xf.max() - xf.min()
    extent = sum(abs(X[en[be[b], 1]] - X[en[be[b], 0]]) for b in lower)
    assert abs(cl - extent / chord) < 1e-9
    assert abs(cd) < 1e-6   # the lower surface runs from y ~ 0 at the LE to y = 0 at the TE


def test_wall_profile_uniform_pressure_is_zero():
    m = orc.OMesh.read(QUAD)
    o = solver_on(m, mach=0.5, angle=0.1)
    o.init()
    o.refill_bcs()
    q = o.vec("q").reshape(-1, 4)
    q[:, 3] *= 1.3   # raises p uniformly
    o.vec("q", q.reshape(-1))
    cl, cd, cm = o.wall_profile("wall")
    assert abs(cl) < 1e-12 and abs(cd) < 1e-12


def test_face_jacobian_matches_central_differences():
    """calc_convective_jacobian (physics.h:533-578): forward differences with
    h = max(1e-6, 1e-6|q|) agree with central differences to O(h)."""
    m = orc.OMesh.read(TRI)
    o = solver_on(m)
    o.init()
    o.refill_bcs()
    nt = m.counts()["cells"]
    rng = np.random.default_rng(7)
    q = o.vec("q")
    q[: 4 * m.counts()["real_cells"]] *= 1 + 0.05 * rng.standard_normal(4 * m.counts()["real_cells"])
    o.vec("q", q)
    ec = m.get("edgesCells").reshape(-1, 2)
    nx, ny, L = m.get("edgesNormalsX"), m.get("edgesNormalsY"), m.get("edgesLengths")
    qq = q.reshape(-1, 4)
    for e in [0, 10, 500, 3000]:
        J = o.face_jacobian(e)
        c0, c1 = ec[e]
        for i in range(4):
            hh = 1e-5 * max(1.0, abs(qq[c0, i]))
            qp, qm = qq[c0].copy(), qq[c0].copy()
            qp[i] += hh
            qm[i] -= hh
            d = (orc.face_flux(0, qp, qq[c1], nx[e], ny[e]) - orc.face_flux(0, qm, qq[c1], nx[e], ny[e])) / (2 * hh)
            assert np.allclose(J[:4, i], d * L[e], rtol=1e-4, atol=1e-6 * L[e])
            assert np.array_equal(J[4:, i], -J[:4, i])


def test_matrix_free_operator_is_block_diagonal_own_state_derivative():
    """RansMatrix * x (ransMatrix.h:150-178): a row sees only its own block of
    x (the neighbour's perturbation enters the other row only), starts from
    x A / dt, and adds the derivative of each face flux along x_c without the
    face length: +dF/dq_L x on the owner row, -dF/dq_R x on the neighbour row.
    Checked against central differences of the face flux on an interior cell,
    and ghost rows against the identity."""
    m = orc.OMesh.read(TRI)
    o = solver_on(m, cfl=3.0)
    o.init()
    o.refill_bcs()
    cnt = m.counts()
    nt, nr = cnt["cells"], cnt["real_cells"]
    rng = np.random.default_rng(5)
    q = o.vec("q")
    q[: 4 * nr] *= 1 + 0.05 * rng.standard_normal(4 * nr)
    o.vec("q", q)
    o.calc_dt()
    dt = o.vec("dt")
    area = m.get("cellsAreas")
    ec = m.get("edgesCells").reshape(-1, 2)
    nx, ny = m.get("edgesNormalsX"), m.get("edgesNormalsY")
    qq = q.reshape(-1, 4)
    ghost_faces = set(np.nonzero(ec[:, 1] >= nr)[0])
    interior = [c for c in range(nr) if not any(f in ghost_faces for f in np.nonzero((ec == c).any(axis=1))[0])]
    for c in interior[:: max(1, len(interior) // 6)][:6]:
        xc = rng.standard_normal(4)
        x = np.zeros(4 * nt)
        x[4 * c: 4 * c + 4] = xc
        y = o.mf_apply(x).reshape(-1, 4)
        others = np.delete(y, c, axis=0)
        assert np.array_equal(others, np.zeros_like(others))
        expect = xc * area[c] / dt[c]
        eps = 1e-5
        for f in np.nonzero((ec == c).any(axis=1))[0]:
            a, b = ec[f]
            if a == c:
                d = (orc.face_flux(0, qq[a] + eps * xc, qq[b], nx[f], ny[f])
                     - orc.face_flux(0, qq[a] - eps * xc, qq[b], nx[f], ny[f])) / (2 * eps)
                expect = expect + d
            else:
                d = (orc.face_flux(0, qq[a], qq[b] + eps * xc, nx[f], ny[f])
                     - orc.face_flux(0, qq[a], qq[b] - eps * xc, nx[f], ny[f])) / (2 * eps)
                expect = expect - d
        assert np.allclose(y[c], expect, rtol=1e-4, atol=1e-6 * np.max(np.abs(expect)))
    # ghost rows are identity rows and receive nothing from one-sided faces
    x = np.zeros(4 * nt)
    x[4 * nr:] = rng.standard_normal(4 * (nt - nr))
    y = o.mf_apply(x)
    assert np.array_equal(y[4 * nr:], x[4 * nr:])
    assert np.array_equal(y[: 4 * nr], np.zeros(4 * nr))


def test_matrix_free_operator_ignores_viscosity():
    """The operator calls the two-argument flux (zero gradients), so the
    laminar stress vanishes and the laminar operator equals the inviscid one."""
    m = orc.OMesh.read(QUAD)
    ops = []
    rng = np.random.default_rng(9)
    x = rng.standard_normal(4 * m.counts()["cells"])
    for visc in (0, 1):
        o = orc.OSolver(m, mu_L=7e-4, viscous=visc)
        o.set_bcs({"farfield": ("farfield", (0.5, 0.02, 1.0, 1.0)), "wall": ("wall", (0.2, 0.0, 1.0, 1.0))})
        o.options(cfl=2.0)
        o.init()
        o.refill_bcs()
        q = o.vec("q")
        nr = 4 * m.counts()["real_cells"]
        q[:nr] *= 1 + 0.05 * np.random.default_rng(3).standard_normal(nr)
        o.vec("q", q)
        o.calc_dt()
        ops.append(o.mf_apply(x))
    assert np.array_equal(ops[0], ops[1])


def test_gen_mapper_hand_computed_weights():
    """multigrid::gen_mapper (multigrid.h:96-163) on three coarse cells, each
    row worked by hand: a coarse cell j is taken when d2 < 2 A_j (strict), its
    weight is 1/max(0.1 sqrt(A_j), sqrt(d2)) over the row's sum; a fine cell
    out of reach of every coarse cell gets an empty row."""
    cx, cy, ca = [0.0, 1.0, 5.0], [0.0, 0.0, 5.0], [1.0, 1.0, 0.01]
    fx = [0.5, 0.0, 5.1, 3.0, 1.0]
    fy = [0.0, 0.0, 5.0, 3.0, 1.0]
    M = orc.gen_mapper(fx, fy, cx, cy, ca).toarray()
    expect = np.array([
        [0.5, 0.5, 0.0],            # d = 0.5 to both unit cells: 1/0.5 each
        [10 / 11, 1 / 11, 0.0],     # d = 0 -> 1/(0.1*1) = 10; d = 1 -> 1
        [0.0, 0.0, 1.0],            # only the small cell (d2 = 0.01 < 0.02)
        [0.0, 0.0, 0.0],            # nothing within reach
        [0.0, 1.0, 0.0],            # d2 = 2 to cell 0 is not < 2; d2 = 1 to cell 1
    ])
    assert np.allclose(M, expect, rtol=1e-15, atol=0)


@pytest.mark.parametrize("coarse,fine", [(TRI, QUAD), (QUAD, TRI)])
def test_gen_mapper_rows_are_partitions_of_unity(coarse, fine):
    """Between the reference's two coarse airfoil meshes every row's weights
    are positive and sum to one, so a uniform state maps to itself.  The
    triangle mesh's farfield is a circle of 5 chords about mid-chord, the quad mesh's much further
    out: fine cells beyond the coarse domain get empty rows (state 0, as in the
    reference), all others are covered."""
    mc, mf = orc.OMesh.read(coarse), orc.OMesh.read(fine)
    M = orc.gen_mapper(mf.get("cellsCentersX"), mf.get("cellsCentersY"), mc.get("cellsCentersX"),
                       mc.get("cellsCentersY"), mc.get("cellsAreas"))
    assert np.all(M.data > 0)
    sums = np.asarray(M.sum(axis=1)).ravel()
    nonempty = np.diff(M.indptr) > 0
    assert np.allclose(sums[nonempty], 1.0, rtol=0, atol=4e-15)
    # distances from mid-chord, the centre of the triangle mesh's farfield circle
    rf = np.hypot(mf.get("cellsCentersX") - 0.5, mf.get("cellsCentersY"))
    nrc = mc.counts()["real_cells"]
    rc = np.hypot(mc.get("cellsCentersX")[:nrc] - 0.5, mc.get("cellsCentersY")[:nrc])
    inside = rf < 0.98 * rc.max()
    assert nonempty[inside].all()
    assert not nonempty[rf > 1.1 * rc.max()].any()


def michalak_setup(limiter_k):
    m = orc.OMesh.read(TRI)
    o = solver_on(m, limiter_k=limiter_k)
    o.set_michalak(1)
    c = m.counts()
    nr = c["real_cells"]
    o.init()
    o.refill_bcs()
    rng = np.random.default_rng(5)
    q = o.vec("q")
    q[: 4 * nr] *= 1 + 0.02 * rng.standard_normal(4 * nr)
    o.vec("q", q)
    o.set_walls()
    o.calc_gradients()
    o.calc_limiters()
    return m, o, nr, rng


def test_michalak_limiter_reduces_to_the_michalak_gooch_cubic_at_k0():
    """RANS_MICHALAK_LIMITER (solver.h:557-573): with limiter_k = 0 the switch
    sig is 0 wherever the neighbours' values differ, so each limiter is the
    min over faces of the Michalak & Ollivier-Gooch (2008) cubic with y_t = 2:
    the cubic through P(0) = 0 with P'(0) = 1 that meets 1 with zero slope at
    y_t, worked out here from those four conditions, then 1 beyond y_t."""
    m, o, nr, rng = michalak_setup(0.0)
    # P(y) = y + b y^2 + a y^3, P(2) = 1, P'(2) = 0
    a, b = np.linalg.solve([[8.0, 4.0], [12.0, 4.0]], [1.0 - 2.0, -1.0])
    P = lambda y: 1.0 if y >= 2 else y + b * y * y + a * y ** 3
    q = o.vec("q").reshape(-1, 4)
    gx, gy = o.vec("gx").reshape(-1, 4), o.vec("gy").reshape(-1, 4)
    lim = o.vec("limiters").reshape(-1, 4)
    ce = m.get("cellsEdges").reshape(-1, 4)[:nr]
    ec = m.get("edgesCells").reshape(-1, 2)
    xf, yf = m.get("edgesCentersX"), m.get("edgesCentersY")
    xc, yc = m.get("cellsCentersX"), m.get("cellsCentersY")
    for i in rng.choice(nr, 400, replace=False):
        fs = ce[i][ce[i] != 0xFFFFFFFF]
        nb = np.where(ec[fs, 0] == i, ec[fs, 1], ec[fs, 0])
        qmax = np.maximum(q[i], q[nb].max(0))
        qmin = np.minimum(q[i], q[nb].min(0))
        expect = np.ones(4)
        for f in fs:
            d = gx[i] * (xf[f] - xc[i]) + gy[i] * (yf[f] - yc[i])
            for k in range(4):
                if qmax[k] == qmin[k] or abs(d[k]) <= 1e-14:
                    continue
                D = (qmax[k] if d[k] > 0 else qmin[k]) - q[i, k]
                expect[k] = min(expect[k], P(D / d[k]))
        assert np.allclose(lim[i], expect, rtol=1e-12, atol=1e-14)


def test_michalak_limiter_switch_is_off_for_smooth_ranges():
    """With a limiter_k so large that (q_max - q_min)^2 <= (K sqrt(A))^3 in
    every cell, sig = 1 and the build applies no limiting at all, unlike the
    default limiter at the same K (which stays below 1 wherever dqg != 0)."""
    m, o, nr, rng = michalak_setup(1e4)
    assert np.array_equal(o.vec("limiters")[: 4 * nr], np.ones(4 * nr))
    o.set_michalak(0)
    o.calc_limiters()
    assert np.min(o.vec("limiters")[: 4 * nr]) < 1.0


def test_vtu_round_trip_and_format(tmp_path):
    """save (post.h:58-180) on a channel of rectangles whose cells carry
    chosen primitives (rho, u, v, p): the file gives back rho, (u, v), p, the
    ideal-gas temperature p / (R rho) and the Mach number |V| / sqrt(gamma p /
    rho) of those primitives; node lines are std::to_string's "%f" and cell
    lines std::fixed precision 16, character for character; the connectivity,
    offsets and types blocks follow the reference's layout (two spaces after
    every entry, quads as type 9)."""
    import xml.etree.ElementTree as ET
    xs = np.cumsum(np.r_[0.0, np.random.default_rng(8).uniform(0.3, 1.7, 6)])
    mesh = orc.OMesh.build(*strip_mesh(xs, height=0.7))
    o = solver_on(mesh)
    o.init()
    rng = np.random.default_rng(9)
    nc = mesh.counts()["real_cells"]
    prim = np.column_stack([rng.uniform(0.5, 2, nc), rng.uniform(-0.6, 0.6, nc), rng.uniform(-0.6, 0.6, nc),
                            rng.uniform(0.5, 2, nc)])
    q = o.vec("q").reshape(-1, 4)
    q[:nc] = np.array([prim_to_cons(*row) for row in prim])
    o.vec("q", q.reshape(-1))
    o.wall_dist()
    path = tmp_path / "out.vtu"
    o.save_vtu(path)
    text = path.read_text()
    root = ET.fromstring(text)
    piece = root.find("UnstructuredGrid/Piece")
    nodes_x, nodes_y = mesh.get("nodesX"), mesh.get("nodesY")
    assert int(piece.get("NumberOfPoints")) == len(nodes_x) and int(piece.get("NumberOfCells")) == nc
    arrays = {a.get("Name"): a.text for a in piece.iter("DataArray") if a.get("Name")}
    assert piece.find("Points/DataArray").text == "\n" + "".join(
        f"          {x:f} {y:f} 0.0\n" for x, y in zip(nodes_x, nodes_y)) + "        "
    cells = mesh.get("cellsNodes").reshape(-1, 4)[:nc]
    assert arrays["connectivity"] == "\n          " + "".join(
        "".join(f"{n}  " for n in row) + "\n          " for row in cells) + "        "
    assert arrays["offsets"] == "\n          " + "".join(f"{4 * (i + 1)}  " for i in range(nc)) + "\n        "
    assert arrays["types"] == "\n          " + "9  " * nc + "\n        "

    def col(name, k=0):
        rows = arrays[name].strip().split("\n")
        return np.array([float(r.split()[k]) for r in rows]), [r.strip().split()[k] for r in rows]

    gamma, R = 1.4, 0.71428571428
    rho, u, v, p = prim.T
    for name, want, tol in (("Density", rho, 1e-16), ("Pressure", p, 1e-14), ("Temperature", p / (R * rho), 1e-14),
                            ("Mach", np.hypot(u, v) / np.sqrt(gamma * p / rho), 1e-14)):
        got, _ = col(name)
        assert np.max(np.abs(got - want) / np.abs(want)) <= tol, name
    ux, _ = col("Velocity", 0)
    vy, _ = col("Velocity", 1)
    assert np.max(np.abs(ux - u)) <= 1e-15 and np.max(np.abs(vy - v)) <= 1e-15
    _, dens_txt = col("Density")
    assert dens_txt == [f"{x:.16f}" for x in q[:nc, 0]]
    wd, _ = col("Wall Distance")
    assert np.array_equal(wd, np.array([float(f"{d:.16f}") for d in o.vec("wall_dist")[:nc]]))
    assert text.endswith("    </Piece>\n  </UnstructuredGrid>\n</VTKFile>")


def test_cp_profile_closed_forms():
    """CpProfile::calc_cp (post.h:246-290): cp = 2 / (gamma M^2) (p / p_inf -
    1) of the owner cell.  The freestream state gives cp = 0 on every wall
    face; the freestream velocity at pressure p_inf (1 + k) gives
    2 k / (gamma M^2); the positions are the wall faces' centres, in the
    mesh's boundary order."""
    m = orc.OMesh.read(TRI)
    mach = 0.5
    o = solver_on(m, mach=mach, angle=0.03)
    o.init()
    o.refill_bcs()
    x, y, cp = o.cp_profile("wall")
    names = m.boundary_names()
    be = m.get("boundaryEdges")
    walls = [be[b] for b in range(len(names)) if names[b] == "wall"]
    assert len(cp) == len(walls) > 0
    assert np.array_equal(x, m.get("edgesCentersX")[walls]) and np.array_equal(y, m.get("edgesCentersY")[walls])
    assert np.max(np.abs(cp)) <= 1e-14
    q = o.vec("q").reshape(-1, 4)
    k = 0.37
    nc = m.counts()["real_cells"]
    # raise the pressure by (1 + k) at fixed density and velocity: rhoE += k p_inf / (gamma - 1)
    q[:nc, 3] += k * 1.0 / 0.4
    o.vec("q", q.reshape(-1))
    _, _, cp2 = o.cp_profile("wall")
    want = 2 * k / (1.4 * mach * mach)
    assert np.max(np.abs(cp2 - want)) <= 1e-13 * want
