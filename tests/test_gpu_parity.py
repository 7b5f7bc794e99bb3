"""Device (HIP) residual vs the oracle, through the C ABI.

Plain FP64 form: bit-for-bit or within 1e-12 relative (max-norm relative to
the largest residual entry).  FP32 form: within 1e-4 relative L2 of the FP64
oracle.  LSQ gradients: 1e-12 (the 2x2 inverse and the d^T delta product are
summed in a different order than Eigen's, so not bit-exact)."""
import numpy as np
import pytest

from cases import QUAD, TRI, make_pair, rel_max

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("path", [QUAD, TRI])
@pytest.mark.parametrize("second_order", [True, False])
def test_residual_fp64_inviscid(path, second_order, fused):
    from aeroflex_amd import rans
    s, o, q = make_pair(path, second_order=second_order, fused=fused)
    R, nrm = s.fillRhoRHS(rans.AFX_FP64)
    o.fill_rhs()
    Ro = o.vec("RhoVector")
    assert rel_max(R, Ro) <= FP64_TOL
    assert abs(nrm - np.linalg.norm(Ro)) <= 1e-12 * np.linalg.norm(Ro)


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("path", [QUAD, TRI])
def test_residual_fp64_laminar_wall(path, fused):
    from aeroflex_amd import rans
    s, o, q = make_pair(path, viscosity="laminar", wall="wall", mu_L=7e-4, fused=fused)
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    o.fill_rhs()
    assert rel_max(R, o.vec("RhoVector")) <= FP64_TOL


@pytest.mark.parametrize("fused", [True, False])
def test_residual_fp64_least_squares(fused):
    from aeroflex_amd import rans
    s, o, q = make_pair(QUAD, gradient="least-squares", fused=fused)
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    o.fill_rhs()
    assert rel_max(R, o.vec("RhoVector")) <= 1e-10


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_gradients_limiters_fp64(path):
    from aeroflex_amd import rans
    s, o, q = make_pair(path)
    gx, gy, lim = s.gradients(rans.AFX_FP64)
    o.set_walls()
    o.calc_gradients()
    o.calc_limiters()
    nr = 4 * s.n_real
    assert rel_max(gx[:nr], o.vec("gx")[:nr]) <= FP64_TOL
    assert rel_max(gy[:nr], o.vec("gy")[:nr]) <= FP64_TOL
    assert np.max(np.abs(lim[:nr] - o.vec("limiters")[:nr])) <= 1e-12


def test_dt_fp64():
    s, o, q = make_pair(QUAD)
    s.set_cfl(40.0)
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=40.0)
    dt = s.calc_dt()
    o.calc_dt()
    dto = o.vec("dt")
    nr = s.n_real
    assert rel_max(dt[:nr], dto[:nr]) <= FP64_TOL
    assert np.all(np.isinf(dt[nr:])) and np.all(np.isinf(dto[nr:]))


def test_explicit_step_fp64():
    s, o, q = make_pair(TRI)
    s.set_cfl(0.5)
    o.options(second_order=True, green_gauss=True, limiter_k=5.0, cfl=0.5)
    for _ in range(3):
        e = s.explicit_solve(0.9)
        eo = o.explicit_step(0.9)
        assert abs(e - eo) <= 1e-12 * eo
    assert rel_max(s.get_q(), o.vec("q")) <= FP64_TOL


def test_uniform_residual_and_stale_qw():
    s, o, q = make_pair(QUAD)
    e1 = s.get_uniform_residual()
    e1o = o.uniform_residual()
    assert abs(e1 - e1o) <= 1e-12 * max(e1o, 1e-300)
    # second call accumulates onto the stale qW (reference quirk)
    e2 = s.get_uniform_residual()
    e2o = o.uniform_residual()
    assert abs(e2 - e2o) <= 1e-12 * max(e2o, 1e-300)


def test_wall_forces_fp64():
    s, o, q = make_pair(TRI)
    cl, cd, cm = s.get_wall_profile("wall")
    ref = o.wall_profile("wall")
    assert np.allclose([cl, cd, cm], ref, rtol=0, atol=1e-14)


def test_wall_distance():
    s, o, q = make_pair(QUAD)
    d = s.wall_dist()
    o.wall_dist()
    assert np.array_equal(d, o.vec("wall_dist"))


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("path", [QUAD, TRI])
def test_residual_fp32(path, fused):
    from aeroflex_amd import rans
    s, o, q = make_pair(path, fused=fused)
    R, _ = s.fillRhoRHS(rans.AFX_FP32)
    o.fill_rhs()
    Ro = o.vec("RhoVector")
    assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= 1e-4


def test_renumbering_invariance():
    from aeroflex_amd import rans
    s1, o, q = make_pair(TRI, renumber=True)
    s2, _, _ = make_pair(TRI, renumber=False)
    R1, _ = s1.fillRhoRHS(rans.AFX_FP64)
    R2, _ = s2.fillRhoRHS(rans.AFX_FP64)
    assert np.array_equal(R1, R2)


@pytest.mark.parametrize("tile_cells", [16, 48, 128, 512])
def test_fused_tile_sizes_bitexact(tile_cells):
    """Small tiles put most faces on tile boundaries (halo gradients and cut
    faces computed by two workgroups); the result must not change."""
    from aeroflex_amd import rans
    s, o, q = make_pair(TRI, tile_cells=tile_cells)
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    o.fill_rhs()
    Ro = o.vec("RhoVector")
    assert rel_max(R, Ro) <= FP64_TOL
    # wall ghosts are refreshed from their owners as set_walls_from_internal does
    assert np.array_equal(s.get_q(), o.vec("q"))


def test_fused_matches_three_kernel_path_fp32():
    from aeroflex_amd import rans
    s1, o, q = make_pair(QUAD, fused=True)
    s2, _, _ = make_pair(QUAD, fused=False)
    R1, _ = s1.fillRhoRHS(rans.AFX_FP32)
    R2, _ = s2.fillRhoRHS(rans.AFX_FP32)
    assert np.linalg.norm(R1 - R2) / np.linalg.norm(R2) <= 1e-5


# ------------------------------------------ RANS_MICHALAK_LIMITER build --
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("path,limiter_k", [(QUAD, 5.0), (TRI, 5.0), (TRI, 0.3)])
def test_residual_fp64_michalak(path, limiter_k, fused):
    """The Michalak limiter build (solver.h:557-573): FP64 residual, fused and
    three-kernel paths, against the oracle; limiter_k 0.3 puts many cells in
    the blended range of the switch."""
    from aeroflex_amd import rans
    s, o, q = make_pair(path, limiter_k=limiter_k, fused=fused, limiter="michalak")
    R, _ = s.fillRhoRHS(rans.AFX_FP64)
    o.fill_rhs()
    Ro = o.vec("RhoVector")
    assert rel_max(R, Ro) <= FP64_TOL
    # and it differs from the default build
    s.set_limiter("default")
    R0, _ = s.fillRhoRHS(rans.AFX_FP64)
    assert rel_max(R0, Ro) > 1e-6


@pytest.mark.parametrize("path", [QUAD, TRI])
def test_limiters_michalak_fp64(path):
    from aeroflex_amd import rans
    s, o, q = make_pair(path, limiter_k=0.3, limiter="michalak")
    gx, gy, lim = s.gradients(rans.AFX_FP64)
    o.set_walls()
    o.calc_gradients()
    o.calc_limiters()
    nr = 4 * s.n_real
    ol = o.vec("limiters")[:nr]
    assert np.max(np.abs(lim[:nr] - ol)) <= 1e-12
    assert np.min(ol) < 0.9 and np.mean((ol > 0) & (ol < 1)) > 0.05


def test_residual_fp32_michalak_uses_the_three_kernel_path():
    from aeroflex_amd import rans
    s, o, q = make_pair(TRI, limiter_k=0.3, limiter="michalak")
    R, _ = s.fillRhoRHS(rans.AFX_FP32)
    o.fill_rhs()
    Ro = o.vec("RhoVector")
    assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= 1e-4


def test_michalak_build_refuses_adjoint_and_batch():
    from aeroflex_amd import rans
    s, o, q = make_pair(QUAD, limiter="michalak")
    w = np.zeros(4 * s.n_cells)
    w[: 4 * s.n_real] = 1.0
    with pytest.raises(rans.RansError, match="Michalak"):
        s.adjoint(w)
    with pytest.raises(rans.RansError, match="Michalak"):
        rans.Batch([s])
