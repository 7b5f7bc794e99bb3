"""Pins of the oracle's implicit-step matrix for the Spalart-Allmaras model.

The reference has no SA Jacobian (its SA branch is empty, physics.h:259-261);
the oracle's is this build's own, made by the reference's rule for the mean
flow (calc_convective_jacobian, physics.h:533-578) applied to the
five-variable face flux, plus the source row.  What fixes it:
  * the face function the blocks differentiate is the residual's: the
    first-order SA residual re-assembled in numpy from that function (member
    centre states) and the cell source equals fill_rhs;
  * each face block column equals central differences of that function with
    the centre states following the perturbed face states (rtol 1e-4: the
    forward difference is O(h)), and rows 5-9 are the negatives of rows 0-4;
  * the source row equals -A times central differences of the cell source
    at frozen gradients, its (4,4) entry clipped at zero from below;
  * the assembled 5N x 5N matrix equals a numpy assembly of A/dt (the
    model's time step), the face blocks and the source rows.
"""
import numpy as np
import pytest

import orc
from cases import QUAD, TRI, perturbed_state

MU = 5e-4


def sa_case(path, wall="wall", second_order=False, seed=3):
    om = orc.OMesh.read(path)
    o = orc.OSolver(om, mu_L=MU, viscous=2)
    o.set_bcs({"farfield": ("farfield", (0.5, 2 * 0.01745, 1.0, 1.0)), "wall": (wall, (0.2, 0.0, 1.0, 1.0))})
    o.options(second_order=second_order, green_gauss=True, limiter_k=5.0, cfl=7.0)
    o.init()
    o.refill_bcs()
    q = perturbed_state(o.n_real, o.n_cells, o.vec("q"), orc.freestream(0.5, 2 * 0.01745), seed=seed, amp=0.03)
    o.vec("q", q)
    rng = np.random.default_rng(seed)
    chi = np.exp(rng.uniform(np.log(0.3), np.log(300.0), o.n_cells))
    nu = MU * chi
    nu[: o.n_real][rng.random(o.n_real) < 0.05] *= -0.2     # a few negative nu~ cells (SA-neg branch)
    o.sa_vec("nu", nu)
    o.sa_prepare()
    return o, om


def cell_states(o):
    q = o.vec("q").reshape(-1, 4)
    nu = o.sa_vec("nu")
    return np.hstack([q, nu[:, None]])


def face_sets(o, om):
    ec = om.get("edgesCells").reshape(-1, 2)
    internal = np.flatnonzero(ec[:, 1] < o.n_real)
    boundary = np.flatnonzero(ec[:, 1] >= o.n_real)
    return ec, internal, boundary


@pytest.mark.parametrize("path,wall", [(TRI, "wall"), (QUAD, "slip-wall")])
def test_face_function_is_the_residual_flux(path, wall):
    """First order: R_k(i) = -sum over owned faces L F_k + sum over
    neighbour faces L F_k (+ A S on the fifth row), with F the function the
    blocks differentiate at the member states."""
    o, om = sa_case(path, wall)
    R, T = o.sa_fill_rhs()
    v = cell_states(o)
    ec, _, _ = face_sets(o, om)
    L = om.get("edgesLengths")
    area = om.get("cellsAreas")
    nr = o.n_real
    acc = np.zeros((o.n_cells, 5))
    for e in range(len(ec)):
        c0, c1 = ec[e]
        f = o.sa_face_flux(e, v[c0], v[c1]) * L[e]
        acc[c0] -= f
        if c1 < nr:
            acc[c1] += f
    src = np.array([o.sa_cell_source(i) for i in range(nr)])
    acc[:nr, 4] += area[:nr] * src
    Rr = R.reshape(-1, 4)[:nr]
    assert np.max(np.abs(acc[:nr, :4] - Rr)) <= 1e-12 * np.max(np.abs(Rr))
    assert np.max(np.abs(acc[:nr, 4] - T[:nr])) <= 1e-12 * np.max(np.abs(T[:nr]))


@pytest.mark.parametrize("path,wall", [(TRI, "wall"), (QUAD, "slip-wall")])
def test_face_blocks_match_central_differences(path, wall):
    o, om = sa_case(path, wall)
    v = cell_states(o)
    ec, internal, boundary = face_sets(o, om)
    L = om.get("edgesLengths")
    rng = np.random.default_rng(4)
    faces = list(rng.choice(internal, 6, replace=False)) + list(rng.choice(boundary, 4, replace=False))
    for e in faces:
        J = o.sa_face_jacobian(int(e))
        c0, c1 = ec[e]
        assert np.array_equal(J[5:], -J[:5])
        for side, c in ((0, c0), (1, c1)):
            for i in range(5):
                h = 1e-6 * max(1.0, abs(v[c, i])) if i < 4 else 1e-4 * max(abs(v[c, 4]), MU)
                a, b = v[c0].copy(), v[c1].copy()
                (b if side else a)[i] += h
                fp = o.sa_face_flux(int(e), a, b, a, b)
                (b if side else a)[i] -= 2 * h
                fm = o.sa_face_flux(int(e), a, b, a, b)
                d = (fp - fm) / (2 * h) * L[e]
                col = J[:5, 5 * side + i]
                scale = np.max(np.abs(J[:5, 5 * side:5 * side + 5])) + 1e-30
                assert np.allclose(col, d, rtol=1e-4, atol=1e-5 * scale), (e, side, i, col, d)


def test_source_rows_match_central_differences():
    o, om = sa_case(TRI, "wall")
    v = cell_states(o)
    area = om.get("cellsAreas")
    rng = np.random.default_rng(5)
    clipped = 0
    for i in rng.choice(o.n_real, 40, replace=False):
        row = o.sa_source_row(int(i))
        d = np.zeros(5)
        for k in range(5):
            h = 1e-6 * max(1.0, abs(v[i, k])) if k < 4 else 1e-4 * max(abs(v[i, 4]), MU)
            a = v[i].copy()
            a[k] += h
            sp = o.sa_cell_source(int(i), a)
            a[k] -= 2 * h
            sm = o.sa_cell_source(int(i), a)
            d[k] = -area[i] * (sp - sm) / (2 * h)
        if d[4] < 0:
            clipped += 1
            assert row[4] == 0.0
            d[4] = 0.0
        assert np.allclose(row, d, rtol=1e-4, atol=1e-5 * np.max(np.abs(d))), (i, row, d)
    assert clipped > 0     # production-dominated cells exist in this state


def test_matrix_assembly():
    """The COO matrix against a numpy assembly of its parts (face blocks at
    both ends of a two-sided face, owner rows only on boundary faces, A/dt
    of the model's time step, source rows, identity ghost rows)."""
    import scipy.sparse as sp
    o, om = sa_case(TRI, "wall")
    A = o.sa_fill_lhs()
    dt = o.vec("dt")
    ec, _, _ = face_sets(o, om)
    area = om.get("cellsAreas")
    nt, nr = o.n_cells, o.n_real
    rows, cols, vals = [], [], []
    blk_r, blk_c = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")

    def put(r0, c0, B):
        rows.append(r0 + blk_r.ravel())
        cols.append(c0 + blk_c.ravel())
        vals.append(np.asarray(B).ravel())

    for i in range(nr):
        B = np.eye(5) * area[i] / dt[i]
        B[4] += o.sa_source_row(i)
        put(5 * i, 5 * i, B)
    for i in range(nr, nt):
        put(5 * i, 5 * i, np.eye(5))
    for e in range(len(ec)):
        J = o.sa_face_jacobian(e)
        c0, c1 = ec[e]
        put(5 * c0, 5 * c0, J[:5, :5])
        put(5 * c0, 5 * c1, J[:5, 5:])
        if c1 < nr:
            put(5 * c1, 5 * c0, J[5:, :5])
            put(5 * c1, 5 * c1, J[5:, 5:])
    M = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(5 * nt, 5 * nt)).tocsr()
    assert abs(A - M).max() <= 1e-13 * abs(M).max()
    # the model's time step: dt = cfl A / (sum |lambda| L + 4 sum coef L^2 / A)
    assert np.array_equal(dt, o.sa_calc_dt())
