"""State adjoint of the FP64 residual (inviscid, Green-Gauss): lambda =
(dR/dq)^T w against central finite differences w . (R(q + h v) - R(q - h v))
/ 2h of the device's FP64 residual (which matches the oracle to 1e-12) with
random unit-norm seeds w and directions v on the real cells, h = 1e-7
(relative error <= 1e-6), and against the same differences of the oracle's
CPU residual on the quadrilateral fixture."""
import os

import numpy as np
import pytest

import orc
from cases import QUAD, TRI, make_pair

pytestmark = pytest.mark.gpu


def unit_on_real(rng, n_real, n_cells):
    x = np.zeros(4 * n_cells)
    x[:4 * n_real] = rng.standard_normal(4 * n_real)
    return x / np.linalg.norm(x)


def fd_check(s, q, w, v, h=1e-7, residual=None):
    from aeroflex_amd import rans
    if residual is None:
        def residual(x):
            s.set_q(x)
            return s.fillRhoRHS(rans.AFX_FP64)[0]
    scale = max(1.0, np.max(np.abs(q)))
    rp, rm = residual(q + h * scale * v), residual(q - h * scale * v)
    return float(w @ (rp - rm)) / (2 * h * scale)


@pytest.mark.parametrize("path,wall,second_order", [(QUAD, "slip-wall", True), (TRI, "slip-wall", True),
                                                    (TRI, "wall", True), (QUAD, "slip-wall", False)])
def test_adjoint_matches_central_differences(path, wall, second_order):
    s, o, q = make_pair(path, wall=wall, second_order=second_order)
    rng = np.random.default_rng(3)
    for _ in range(4):
        w = unit_on_real(rng, s.n_real, s.n_cells)
        s.set_q(q)
        lam = s.adjoint(w)
        assert np.all(lam[4 * s.n_real:] == 0)
        for _ in range(4):
            v = unit_on_real(rng, s.n_real, s.n_cells)
            fd = fd_check(s, q, w, v)
            ad = float(lam @ v)
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(lam)), (ad, fd)


def test_adjoint_on_jittered_mesh_farfield_and_quads():
    from aeroflex_amd import rans
    m = rans.Mesh.jittered(96, 40, first_h=2e-3, seed=5, quad_frac=0.3)
    path = "/tmp/afx_adj_jitter.msh"
    m.write_msh(path)
    s, o, q = make_pair(path, wall="slip-wall", mesh=m, omesh=orc.OMesh.read(path), mach=0.6)
    rng = np.random.default_rng(4)
    w = unit_on_real(rng, s.n_real, s.n_cells)
    s.set_q(q)
    lam = s.adjoint(w)
    for _ in range(8):
        v = unit_on_real(rng, s.n_real, s.n_cells)
        fd, ad = fd_check(s, q, w, v), float(lam @ v)
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(lam)), (ad, fd)


def test_adjoint_against_oracle_differences():
    s, o, q = make_pair(QUAD, wall="slip-wall")
    rng = np.random.default_rng(8)
    w = unit_on_real(rng, s.n_real, s.n_cells)
    s.set_q(q)
    lam = s.adjoint(w)

    def oracle_residual(x):
        o.vec("q", x)
        o.fill_rhs()
        return o.vec("RhoVector")

    for _ in range(3):
        v = unit_on_real(rng, s.n_real, s.n_cells)
        fd = fd_check(s, q, w, v, residual=oracle_residual)
        assert abs(float(lam @ v) - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(lam))


def test_adjoint_rejects_other_models():
    """The state adjoint covers every model; the node / pseudo-time adjoints
    cover the inviscid residual, so the laminar and SA ones are refused, and a
    rho nu~ seed is refused without the SA model."""
    from aeroflex_amd import rans
    import test_gpu_sa
    s, _ = test_gpu_sa.sa_pair(TRI)
    s.adjoint(np.zeros(4 * s.n_cells))
    with pytest.raises(rans.RansError):
        s.adjoint_nodes(np.zeros(4 * s.n_cells))
    s.close()
    s, _, _ = make_pair(TRI, viscosity="laminar", wall="wall", mu_L=1e-3)
    s.adjoint(np.zeros(4 * s.n_cells))
    with pytest.raises(rans.RansError):
        s.adjoint_nodes(np.zeros(4 * s.n_cells))
    with pytest.raises(rans.RansError):
        s.adjoint(np.zeros(4 * s.n_cells), w_nu=np.zeros(s.n_cells))


# ------------------------------------------------------ node coordinates --
def elements_of(m):
    """(x, y, cells, nverts, bsegs, bkeys, names) of a mesh, in its own order,
    so that a mesh rebuilt from moved nodes has the same cells, faces and
    ghosts."""
    c = m.counts()
    nr = c["real_cells"]
    x, y = m.get("nodesX"), m.get("nodesY")
    cells = m.get("cellsNodes").reshape(-1, 4)[:nr].astype(np.int32)
    nverts = np.where(m.get("cellsIsTriangle")[:nr] != 0, 3, 4).astype(np.int32)
    en = m.get("edgesNodes").reshape(-1, 2)
    be = m.get("boundaryEdges")
    bsegs = en[be].astype(np.int32)
    bkeys = m.get("boundaryEdgesIntTag").astype(np.int32)
    bnames = m.boundary_names()
    names = {}
    for k, nme in zip(bkeys, bnames):
        names[int(k)] = nme
    return x, y, cells, nverts, bsegs, bkeys, names


def moved_residual(el, dxy, q, wall, second_order, mach=0.5, angle=2 * 0.01745):
    from aeroflex_amd import rans
    x, y, cells, nverts, bsegs, bkeys, names = el
    m = rans.Mesh.from_elements(x + dxy[:, 0], y + dxy[:, 1], cells, nverts, bsegs, bkeys, names)
    s = rans.Solver(m, second_order=second_order, gradient_scheme="green-gauss", limiter_k=5.0)
    far = rans.BoundaryVariables(mach=mach, angle=angle)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition(wall)})
    s.init()
    s.set_q(q)
    return s.fillRhoRHS(rans.AFX_FP64)[0]


@pytest.mark.parametrize("path,wall,second_order", [(QUAD, "slip-wall", True), (TRI, "slip-wall", True),
                                                    (TRI, "wall", True), (QUAD, "slip-wall", False)])
def test_node_adjoint_matches_central_differences(path, wall, second_order):
    """xbar = (dR/dX)^T w against w . (R(X + h V) - R(X - h V)) / 2h of the
    device FP64 residual on meshes rebuilt from moved nodes (h = 1e-7,
    random unit directions V over all nodes, farfield and wall nodes included)."""
    print(f"[adjoint] nodes {os.path.basename(path)} {wall} second_order={second_order}", flush=True)
    s, o, q = make_pair(path, wall=wall, second_order=second_order)
    el = elements_of(s.mesh)
    rng = np.random.default_rng(11)
    w = unit_on_real(rng, s.n_real, s.n_cells)
    s.set_q(q)
    lam0 = s.adjoint(w)
    s.set_q(q)
    lam, xb = s.adjoint_nodes(w)
    np.testing.assert_array_equal(lam, lam0)
    h = 1e-7
    for _ in range(3):
        # every node moves by about h: the difference quotient then carries
        # the signal of all nodes above the residual's rounding
        V = rng.standard_normal(xb.shape)
        rp = moved_residual(el, h * V, q, wall, second_order)
        rm = moved_residual(el, -h * V, q, wall, second_order)
        fd = float(w @ (rp - rm)) / (2 * h)
        ad = float(np.sum(xb * V))
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(xb) * np.linalg.norm(V)), (ad, fd)


def test_node_adjoint_single_nodes_on_jittered_mesh():
    """Per-node check: central differences (h = 1e-5) for individual interior,
    wall and farfield nodes of a jittered mixed mesh, to 1e-5 relative plus
    the quotient's rounding bound 50 eps sum|w R| / 2h."""
    from aeroflex_amd import rans
    m = rans.Mesh.jittered(64, 32, first_h=2e-3, seed=7, quad_frac=0.3)
    path = "/tmp/afx_adj_nodes.msh"
    m.write_msh(path)
    s, o, q = make_pair(path, wall="slip-wall", mesh=m, omesh=orc.OMesh.read(path), mach=0.6)
    el = elements_of(m)
    rng = np.random.default_rng(12)
    w = unit_on_real(rng, s.n_real, s.n_cells)
    s.set_q(q)
    _, xb = s.adjoint_nodes(w)
    nn = len(xb)
    h = 1e-5
    for node in [0, 5, nn // 2, nn - 3] + list(rng.integers(0, nn, 4)):
        for d in (0, 1):
            V = np.zeros_like(xb)
            V[node, d] = 1.0
            rp = moved_residual(el, h * V, q, "slip-wall", True, mach=0.6)
            rm = moved_residual(el, -h * V, q, "slip-wall", True, mach=0.6)
            fd = float(w @ (rp - rm)) / (2 * h)
            noise = 50 * 2.2e-16 * float(np.abs(w) @ np.abs(rp)) / (2 * h)
            assert abs(xb[node, d] - fd) <= 1e-5 * abs(fd) + noise, (node, d, xb[node, d], fd, noise)


# ------------------------------------------------ dCL/dalpha through steps --
def cl_after_steps(s, q, angle, steps, mach=0.5):
    from aeroflex_amd import rans
    far = rans.BoundaryVariables(mach=mach, angle=angle)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
    s.set_q(q)
    s.refill_bcs()
    for _ in range(steps):
        s.explicit_solve(1.0)
    return s.get_wall_profile("wall")[0]


@pytest.mark.parametrize("path,second_order", [(QUAD, False), (QUAD, True), (TRI, True)])
def test_dcl_dalpha_matches_central_differences(path, second_order):
    """dCL/dangle by reverse accumulation through 60 explicit steps with
    checkpoints every 7 steps (a ragged last segment) against central
    differences of CL after the same 60 steps (h = 1e-6 rad); the
    checkpoint interval does not change the result."""
    from aeroflex_amd import rans
    print(f"[adjoint] dcl/dalpha {os.path.basename(path)} second_order={second_order}", flush=True)
    angle = 2 * 0.01745
    s, o, q = make_pair(path, wall="slip-wall", second_order=second_order, angle=angle)
    s.set_cfl(1.0)
    steps = 60
    s.set_q(q)
    cl, d = s.dcl_dalpha(steps=steps, checkpoint_every=7)
    s.set_q(q)
    cl1, d1 = s.dcl_dalpha(steps=steps, checkpoint_every=steps)
    assert cl1 == cl and d1 == d
    cl0 = cl_after_steps(s, q, angle, steps)
    assert abs(cl0 - cl) <= 1e-12 * abs(cl)
    h = 1e-6
    fd = (cl_after_steps(s, q, angle + h, steps) - cl_after_steps(s, q, angle - h, steps)) / (2 * h)
    assert abs(d - fd) <= 1e-6 * abs(fd), (d, fd)


# ----------------------------------------------------- dCL/dX through steps --
def cl_after_steps_moved(el, dxy, q, steps, second_order, angle=2 * 0.01745, mach=0.5):
    from aeroflex_amd import rans
    x, y, cells, nverts, bsegs, bkeys, names = el
    m = rans.Mesh.from_elements(x + dxy[:, 0], y + dxy[:, 1], cells, nverts, bsegs, bkeys, names)
    s = rans.Solver(m, second_order=second_order, gradient_scheme="green-gauss", limiter_k=5.0)
    far = rans.BoundaryVariables(mach=mach, angle=angle)
    s.set_bcs({"farfield": rans.BoundaryCondition("farfield", far), "wall": rans.BoundaryCondition("slip-wall")})
    s.init()
    s.set_cfl(1.0)
    s.set_q(q)
    s.refill_bcs()
    for _ in range(steps):
        s.explicit_solve(1.0)
    return s.get_wall_profile("wall")[0]


@pytest.mark.parametrize("path,second_order", [(QUAD, False), (QUAD, True), (TRI, True)])
def test_dcl_dnodes_matches_central_differences(path, second_order):
    """dCL/dX by reverse accumulation through 40 explicit steps with
    checkpoints every 7 steps against central differences of CL after the
    same 40 steps on meshes rebuilt from moved nodes (h = 1e-8 along a
    random unit direction over all nodes, wall and farfield included, and
    along the wall nodes only); the checkpoint interval does not change it.
    At second order the limiter's min / max switches make larger steps
    straddle kinks (on the triangle mesh along wall displacements: 2e-3
    relative at h = 1e-6, 7e-4 at 1e-7, 7e-8 at 1e-8), hence the small h;
    the bound adds the quotient's rounding, 50 eps |CL| / 2h."""
    print(f"[adjoint] dcl/dX {os.path.basename(path)} second_order={second_order}", flush=True)
    angle = 2 * 0.01745
    s, o, q = make_pair(path, wall="slip-wall", second_order=second_order, angle=angle)
    s.set_cfl(1.0)
    steps = 40
    s.set_q(q)
    cl, xb = s.dcl_dnodes(steps=steps, checkpoint_every=7)
    s.set_q(q)
    cl1, xb1 = s.dcl_dnodes(steps=steps, checkpoint_every=steps)
    assert cl1 == cl and np.array_equal(xb1, xb)
    el = elements_of(s.mesh)
    assert abs(cl_after_steps_moved(el, np.zeros_like(xb), q, steps, second_order) - cl) <= 1e-12 * abs(cl)
    rng = np.random.default_rng(23)
    on_wall = np.array(s.mesh.boundary_names()) == "wall"
    wall_nodes = np.unique(el[4][on_wall].ravel())
    # CL divides by the chord extent max - min of the wall-face midpoints; where
    # a second face lies within 1e-6 of an extreme without tying it exactly (the
    # triangle mesh: 7e-11 and 2e-9) the extent is not differentiable at the
    # scale of h, so those faces' nodes stay put (exact ties are kept: the
    # derivative splits evenly between them)
    xc = s.mesh.get("edgesCentersX")[s.mesh.get("boundaryEdges")[on_wall]]
    near = [(0 < abs(x - xc.max()) < 1e-6) or (0 < abs(x - xc.min()) < 1e-6) for x in xc]
    extreme = (xc == xc.max()) | (xc == xc.min())
    frozen = np.unique(el[4][on_wall][np.array(near) | (extreme & np.any(near))].ravel()) if any(near) else []
    dirs = [rng.standard_normal(xb.shape)]
    vw = np.zeros_like(xb)
    vw[wall_nodes] = rng.standard_normal((len(wall_nodes), 2))
    dirs.append(vw)
    for V in dirs:
        V[frozen] = 0.0
    h = 1e-8
    noise = 50 * 2.2e-16 * abs(cl) / (2 * h)
    for V in dirs:
        V = V / np.linalg.norm(V)
        fd = (cl_after_steps_moved(el, h * V, q, steps, second_order)
              - cl_after_steps_moved(el, -h * V, q, steps, second_order)) / (2 * h)
        ad = float(np.sum(xb * V))
        assert abs(ad - fd) <= 1e-6 * abs(fd) + noise, (ad, fd)
