"""Host-side parts of the vortex-lattice / beam path that run without a GPU:
the wing lattice generator and the banded-Cholesky frame model, against the
oracle and the cantilever closed forms."""
import math
import os

import numpy as np
import pytest

import orc_vlm as ov
from aeroflex_amd import vlm as V


def test_wing_lattice_counts_and_reference_values():
    xyz, rn, so, ref = V.wing_lattice(64, 256)
    assert xyz.shape == (65 * 257, 3) and rn.shape == (16384, 4) and so[-1] == 16384 and len(so) == 257
    cr, lam, ar = 1.0, 0.4, 10.0
    span = ar * cr * (1 + lam) / 2
    assert abs(ref["span"] - span) < 1e-15 and abs(ref["sref"] - span * cr * (1 + lam) / 2) < 1e-15
    assert abs(ref["span"] ** 2 / ref["sref"] - ar) < 1e-12
    assert abs(ref["cref"] - 2 / 3 * cr * (1 + lam + lam * lam) / (1 + lam)) < 1e-15
    y = xyz[:, 1].reshape(257, 65)[:, 0]
    np.testing.assert_allclose(y, -y[::-1], atol=1e-15)
    assert y[128] == 0.0 and np.all(np.diff(y) > 0)
    # quarter-chord sweep 25 deg: the ring line at the panels' quarter chord
    # of the leading panel row moves aft by |y| tan 25 plus the taper shift
    _, nor, area = ov.geometry(xyz, rn, so, 0.0)
    np.testing.assert_allclose(nor[:, 2], 1.0, atol=1e-15)
    assert np.all(area > 0)


def test_wing_lattice_rejects_bad_specs():
    with pytest.raises(V.VlmError):
        V.wing_lattice(4, 7)
    with pytest.raises(V.VlmError):
        V.wing_lattice(0, 8)


def test_beam_tip_moment_closed_form():
    E, G, A, I, J = 140e3, 70e3, 100.0, 120.0, 120.0
    x = np.zeros((13, 3))
    x[:, 0] = np.linspace(0, 300, 13)
    en = np.stack([np.arange(12), np.arange(1, 13)], 1)
    props = np.tile([E * A, E * I, E * I, G * J, 0.0, 1.0, 0.0], (12, 1))
    b = V.Beam(x, en, props, [0])
    f = np.zeros((13, 6))
    f[12, 5] = 35e3
    u = b.solve(f)
    assert abs(u[12, 5] - 0.625) < 1e-12 and abs(u[12, 1] - 93.75) < 1e-9


def test_beam_matches_oracle_on_a_swept_wing_beam():
    xyz, rn, so, ref = V.wing_lattice(8, 32)
    bx, en, props, cl = V.wing_beam(xyz, so, 8)
    rng = np.random.default_rng(5)
    f = rng.normal(size=(len(bx), 6))
    u = V.Beam(bx, en, props, cl).solve(f)
    uo = ov.beam_solve(bx, en, props, cl, f)
    # banded Cholesky against the oracle's dense LU: K's condition number
    # (EA / EI ~ 1e4 over short elements) amplifies rounding to ~1e-9
    np.testing.assert_allclose(u, uo, rtol=0, atol=1e-7 * np.abs(uo).max())
    assert np.abs(u[cl[0]]).max() == 0.0


def test_beam_rejects_mechanism_and_bad_input():
    x = np.zeros((3, 3))
    x[:, 1] = [0, 1, 2]
    en = [[0, 1], [1, 2]]
    props = np.tile([1.0, 1.0, 1.0, 1.0, -1.0, 0.0, 0.0], (2, 1))
    with pytest.raises(V.VlmError, match="positive definite"):
        V.Beam(x, en, props, np.zeros(0, np.int32))
    with pytest.raises(V.VlmError, match="orientation"):
        V.Beam(x, en, np.tile([1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 0.0], (2, 1)), [0])
    with pytest.raises(V.VlmError, match="invalid node"):
        V.Beam(x, [[0, 1], [1, 1]], props, [0])


GOLDEN_STRUCT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "structure")


def test_read_structure_moment_file():
    """MODEL::read_data_file (model.hpp:368-617) on examples/structure/Moment.txt:
    13 GRIDs along x, 12 CBARs with PBAR 10 / MAT1 11, SPC1 on node 100,
    MOMENT 35e3 about z on node 112."""
    xyz, en, props, cl, load = V.read_structure(os.path.join(GOLDEN_STRUCT, "Moment.txt"))
    assert xyz.shape == (13, 3) and np.allclose(xyz[:, 0], np.arange(13) * 25.0)
    assert en.tolist() == [[k, k + 1] for k in range(12)]
    assert np.allclose(props, [140e3 * 100, 140e3 * 120, 140e3 * 120, 70e3 * 120, 0, 1, 0])
    assert cl.tolist() == [0]
    expect = np.zeros((13, 6))
    expect[12, 5] = 35e3
    assert np.array_equal(load, expect)


@pytest.mark.parametrize("case", ["moment_txt", "inclined_3d"])
def test_nonlinear_beam_matches_oracle(case):
    """afx_beam_solve_nonlinear against the oracle's restatement of
    get_NonLin_Solve: the reference's tip-moment example, and a cantilever
    inclined in 3D under a follower tip force, torque and moment."""
    if case == "moment_txt":
        xyz, en, props, cl, load = V.read_structure(os.path.join(GOLDEN_STRUCT, "Moment.txt"))
        steps = 10
    else:
        ne = 20
        d = np.array([0.3, 1.0, 0.2]) / np.linalg.norm([0.3, 1.0, 0.2])
        xyz = np.outer(np.linspace(0.0, 10.0, ne + 1), d)
        en = np.c_[np.arange(ne), np.arange(1, ne + 1)]
        props = np.tile([1e6, 2e3, 5e3, 1.5e3, 1.0, 0.0, 0.0], (ne, 1))
        cl = np.array([0])
        load = np.zeros((ne + 1, 6))
        load[-1] = [3.0, 1.0, 20.0, 15.0, 5.0, 2.0]
        steps = 5
    b = V.Beam(xyz, en, props, cl)
    u, info = b.solve_nonlinear(load, n_steps=steps, tol=1e-8)
    uo, io = ov.beam_nonlinear(xyz, en, props, cl, load, n_steps=steps, tol=1e-8)
    assert info.converged and io["converged"]
    assert np.max(np.abs(u - uo)) <= 1e-7 * np.max(np.abs(uo))
    assert abs(info.iterations - io["iterations"]) <= 1


def test_nonlinear_beam_rejects_bad_settings():
    xyz, en, props, cl, load = V.read_structure(os.path.join(GOLDEN_STRUCT, "Moment.txt"))
    b = V.Beam(xyz, en, props, cl)
    for kw in ({"n_steps": 0}, {"tol": 0.0}, {"damping": 0.0}, {"max_iter": 0}):
        with pytest.raises(V.VlmError):
            b.solve_nonlinear(load, **kw)
    u, info = b.solve_nonlinear(load, n_steps=1, tol=1e-14, max_iter=3)
    assert not info.converged and info.iterations == 3
