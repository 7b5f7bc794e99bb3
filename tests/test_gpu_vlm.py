"""Vortex-lattice / beam coupling path on the GPU against the oracle
(oracle/oracle_vlm.cpp) and against what the mathematics fixes.

Tolerances: the influence matrix, right-hand side and geometry are FP64 on
both sides but evaluated in a different algebraic form (the classical
Katz-Plotkin segment formula on the device, the reference's h / direction
form in the oracle): 1e-12 of the largest entry.  Gamma, forces and CL come
from FP32 LU + FP64 refinement to a backward error of sqrt(N) eps64, so
they agree with the oracle's exact FP64 solve to 1e-10 relative.  The
coupled tip deflection is converged to 1e-11 on both sides and compared at
1e-8."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vlm():
    from aeroflex_amd import vlm as V
    return V


@pytest.fixture(scope="module")
def ov():
    import orc_vlm
    return orc_vlm


def small_wing(V, nc=6, ns=20, **kw):
    xyz, rn, so, ref = V.wing_lattice(nc, ns, **kw)
    return xyz, rn, so, ref


@pytest.mark.parametrize("alpha,rc", [(0.0, 0.0), (3.0, 0.0), (-2.0, 0.0), (6.0, 0.05)])
def test_aic_matches_oracle(vlm, ov, alpha, rc):
    print(f"[vlm] aic alpha={alpha} rc={rc}", flush=True)
    xyz, rn, so, ref = small_wing(vlm)
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(alpha, sref=ref["sref"], cref=ref["cref"], core_radius=rc)
    h.build()
    A = h.aic()
    rhs, _, col, nor = h.system()
    Ao, rhso = ov.build(xyz, rn, so, alpha, wake_length=100 * ref["cref"], rc=rc)
    co, no, _ = ov.geometry(xyz, rn, so, alpha)
    assert np.abs(A - Ao).max() <= 1e-12 * np.abs(Ao).max()
    np.testing.assert_allclose(rhs, rhso, rtol=0, atol=1e-15)
    np.testing.assert_allclose(col, co, rtol=0, atol=1e-15)
    np.testing.assert_allclose(nor, no, rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [1, 5, 119, 120, 128, 129, 250, 1000, 1537])
def test_lu_factor_backward_error_and_pivoting(vlm, n):
    print(f"[vlm] lu n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    d = vlm.Dense(n)
    d.set(A)
    assert d.factor() == 0
    lu, piv = d.lu()
    L = np.tril(lu.astype(np.float64), -1) + np.eye(n)
    U = np.triu(lu.astype(np.float64))
    perm = np.arange(n)
    for k in range(n):
        assert k <= piv[k] < n
        perm[[k, piv[k]]] = perm[[piv[k], k]]
    A32 = A.astype(np.float32).astype(np.float64)
    err = np.linalg.norm(A32[perm] - L @ U) / np.linalg.norm(A32)
    assert err < 2e-6 * math.sqrt(n), err
    assert np.abs(L).max() <= 1.0                  # partial pivoting: |multipliers| <= 1
    assert piv[0] == np.argmax(np.abs(A32[:, 0]))  # first pivot is the column maximum


@pytest.mark.parametrize("n,cond", [(300, 1e2), (700, 1e6), (1537, 1e3)])
def test_refined_solve_matches_fp64(vlm, n, cond):
    print(f"[vlm] refine n={n} cond={cond:g}", flush=True)
    rng = np.random.default_rng(7)
    Q1, _ = np.linalg.qr(rng.normal(size=(n, n)))
    Q2, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = (Q1 * np.logspace(0, -math.log10(cond), n)) @ Q2
    b = rng.normal(size=n)
    d = vlm.Dense(n)
    d.set(A)
    d.factor()
    x, info = d.solve(b)
    assert info.converged and info.backward_error <= math.sqrt(n) * 1.12e-16
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 50 * cond * 1.12e-16
    # a second right-hand side reuses the factor
    x2, info2 = d.solve(2 * b)
    np.testing.assert_allclose(x2, 2 * xr, rtol=0, atol=50 * cond * 1.12e-16 * np.abs(xr).max() * 2)


def test_singular_matrix_is_reported(vlm):
    A = np.eye(40)
    A[:, 17] = 0.0
    d = vlm.Dense(40)
    d.set(A)
    with pytest.raises(vlm.VlmError, match="zero pivot at column 17"):
        d.factor()
    with pytest.raises(vlm.VlmError, match="no factorisation"):
        d.solve(np.ones(40))


def test_solve_after_a_timed_out_wait(vlm):
    """The host side of a triangular solve whose bounded wait timed out,
    forced from the host by raising the kernels' abort flag as the wait would:
    the solve fails with a clear error and discards the factorisation, a
    second solve is refused until the next factor, which clears the flag, and
    the solve is then right again."""
    n = 300
    rng = np.random.default_rng(11)
    A = rng.normal(size=(n, n)) + n * np.eye(n)
    b = rng.normal(size=n)
    xr = np.linalg.solve(A, b)
    d = vlm.Dense(n)
    d.set(A)
    assert d.factor() == 0
    x, info = d.solve(b)
    assert info.converged
    d.raise_abort()
    with pytest.raises(vlm.VlmError, match="timed out"):
        d.solve(b)
    with pytest.raises(vlm.VlmError, match="no factorisation"):
        d.solve(b)
    assert d.factor() == 0
    x2, info2 = d.solve(b)
    assert info2.converged
    np.testing.assert_allclose(x2, xr, rtol=0, atol=1e-12 * np.abs(xr).max())
    np.testing.assert_allclose(x2, x, rtol=0, atol=1e-13 * np.abs(xr).max())


@pytest.mark.parametrize("alpha", [2.0, 8.0])
def test_gamma_forces_cl_match_oracle(vlm, ov, alpha):
    print(f"[vlm] solve alpha={alpha}", flush=True)
    xyz, rn, so, ref = small_wing(vlm, 8, 24)
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(alpha, vinf=1.3, rho=0.9, sref=ref["sref"], cref=ref["cref"])
    h.build()
    h.factor()
    g, info = h.solve()
    assert info.converged
    rf, sf, c = h.forces()
    Ao, rhso = ov.build(xyz, rn, so, alpha, vinf=1.3, wake_length=100 * ref["cref"])
    go = ov.solve(Ao, rhso)
    rfo, co = ov.forces(xyz, rn, so, go, alpha, vinf=1.3, rho=0.9, sref=ref["sref"])
    assert np.abs(g - go).max() <= 1e-10 * np.abs(go).max()
    assert np.abs(rf - rfo).max() <= 1e-10 * np.abs(rfo).max()
    assert abs(c["cl"] - co["cl"]) <= 1e-10 * abs(co["cl"])
    np.testing.assert_allclose(sf, rfo.reshape(len(so) - 1, -1, 3).sum(1), rtol=0, atol=1e-10 * np.abs(rfo).max())


def test_independent_lattice_matches_oracle(vlm, ov):
    import test_oracle_vlm as t
    xyz, rn, so, sref = t.lattice(5, 14, 7.0, 1.0, 30.0, 0.5)
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(4.0, sref=sref, cref=1.0)
    h.build()
    h.factor()
    g, info = h.solve()
    Ao, rhso = ov.build(xyz, rn, so, 4.0, wake_length=100.0)
    go = ov.solve(Ao, rhso)
    assert info.converged and np.abs(g - go).max() <= 1e-10 * np.abs(go).max()


def test_stale_factor_still_converges(vlm, ov):
    # the factor of the alpha = 0 matrix preconditions the alpha = 8 solve
    xyz, rn, so, ref = small_wing(vlm, 8, 24)
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(0.0, sref=ref["sref"], cref=ref["cref"])
    h.build()
    h.factor()
    h.set_flow(8.0, sref=ref["sref"], cref=ref["cref"])
    h.build()
    g, info = h.solve(max_iter=60)
    assert info.converged
    Ao, rhso = ov.build(xyz, rn, so, 8.0, wake_length=100 * ref["cref"])
    assert np.abs(g - ov.solve(Ao, rhso)).max() <= 1e-10 * np.abs(g).max()


# (20, 2): tip deflection ~4 % of the half span, plain fixed point; (4, 0.4):
# ~14 %, where the undamped iteration oscillates and relaxation 0.7 is needed
# (the oracle runs the fixed relaxation; Aitken's must reach the same point)
@pytest.mark.parametrize("ei,relax,aitken", [((20.0, 2.0), 1.0, 0), ((20.0, 2.0), 1.0, 1), ((4.0, 0.4), 0.7, 0),
                                             ((4.0, 0.4), 0.7, 1)])
def test_aeroelastic_matches_oracle(vlm, ov, ei, relax, aitken):
    print(f"[vlm] aeroelastic small ei={ei} relax={relax} aitken={aitken}", flush=True)
    nc, ns = 4, 16
    xyz, rn, so, ref = vlm.wing_lattice(nc, ns)
    beam = vlm.wing_beam(xyz, so, nc, ei_flap=ei, ei_chord=(10 * ei[0], 10 * ei[1]),
                         gj=(0.75 * ei[0], 0.75 * ei[1]))
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(5.0, sref=ref["sref"], cref=ref["cref"])
    b = vlm.Beam(*beam)
    ae = vlm.Aeroelastic(h, b)
    info = ae.solve(tol=1e-11, max_iter=80, relax=relax, aitken=aitken)
    r = ov.aeroelastic(xyz, rn, so, 5.0, beam, sref=ref["sref"], cref=ref["cref"], tol=1e-11, max_iter=80,
                       relax=relax)
    assert info.converged and info.tip_deflection > 0
    assert abs(info.tip_deflection - r["tip"]) <= 1e-8 * abs(r["tip"])
    assert abs(info.cl - r["cl"]) <= 1e-8 * abs(r["cl"])
    np.testing.assert_allclose(ae.disp(), r["disp"], rtol=0, atol=1e-8 * np.abs(r["disp"]).max())


def test_full_size_wing_solve():
    """64 x 256 rings (N = 16384), AR 10, taper 0.4, sweep 25, cosine spacing:
    refinement reaches sqrt(N) eps64, the loading is mirror-symmetric and the
    lift slope is within 3 % of Helmbold's."""
    from aeroflex_amd import vlm as V
    print("[vlm] full size 16384", flush=True)
    xyz, rn, so, ref = V.wing_lattice(64, 256)
    h = V.VLM(xyz, rn, so)
    h.set_flow(2.0, sref=ref["sref"], cref=ref["cref"])
    h.build()
    h.factor()
    g, info = h.solve()
    assert info.converged and info.backward_error <= math.sqrt(16384) * 1.12e-16
    G = g.reshape(256, 64)
    assert np.abs(G - G[::-1]).max() <= 1e-8 * np.abs(G).max()
    _, _, c = h.forces()
    slope = c["cl"] / math.radians(2.0)
    tan_half = math.tan(math.radians(25.0)) - 4 / 10 * 0.25 * 0.6 / 1.4
    helm = 2 * math.pi * 10 / (2 + math.sqrt(100 * (1 + tan_half ** 2) + 4))
    assert abs(slope - helm) / helm < 0.03, (slope, helm)


# ------------------------------------------------ nonlinear (viscous polar) --
GOLD_VLM = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "vlm")


@pytest.mark.parametrize("alpha,relax", [(13.0, 0.5), (16.0, 0.3)])
def test_nonlinear_stall_matches_oracle(vlm, ov, alpha, relax):
    print(f"[vlm] nonlinear alpha={alpha}", flush=True)
    polar = vlm.read_polar(f"{GOLD_VLM}/database_stall.dat", f"{GOLD_VLM}/locations.dat")
    xyz, rn, so, ref = small_wing(vlm, 6, 20)
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(alpha, sref=ref["sref"], cref=ref["cref"])
    info, a, c = h.nonlinear_solve(polar, relaxation=relax, tol=1e-9, max_iter=300)
    r = ov.nonlinear(xyz, rn, so, alpha, polar, sref=ref["sref"], cref=ref["cref"], relax=relax, tol=1e-9,
                     max_iter=300)
    assert info.converged and info.iterations == r["iterations"]
    np.testing.assert_allclose(a, r["alpha"], rtol=0, atol=1e-8)
    np.testing.assert_allclose(c, r["station_cl"], rtol=0, atol=1e-8)
    assert abs(info.cl - r["cl"]) <= 1e-9 and abs(info.cy - r["cy"]) <= 1e-9


def test_nonlinear_on_the_reference_lattice_file(vlm, ov):
    """examples/vlm/wing_0.dat (two wings of 4 rings) with the stall polar."""
    polar = vlm.read_polar(f"{GOLD_VLM}/database_stall.dat", f"{GOLD_VLM}/locations.dat")
    xyz, rn, so = vlm.read_lattice(f"{GOLD_VLM}/wing_0.dat")
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(15.0, sref=4.0, cref=1.0)
    info, a, c = h.nonlinear_solve(polar, relaxation=0.5, tol=1e-10, max_iter=500)
    r = ov.nonlinear(xyz, rn, so, 15.0, polar, sref=4.0, cref=1.0, relax=0.5, tol=1e-10, max_iter=500)
    assert info.converged
    np.testing.assert_allclose(a, r["alpha"], rtol=0, atol=1e-8)
    assert abs(info.cl - r["cl"]) <= 1e-9 * max(1.0, abs(r["cl"]))


def test_nonlinear_thin_airfoil_polar_reproduces_the_linear_cl(vlm):
    polar = vlm.read_polar(f"{GOLD_VLM}/database.dat", f"{GOLD_VLM}/locations.dat")
    import test_oracle_vlm as t
    xyz, rn, so, sref = t.lattice(4, 16, 8.0, 1.0)
    h = vlm.VLM(xyz, rn, so)
    h.set_flow(4.0, sref=sref, cref=1.0)
    info, a, c = h.nonlinear_solve(polar)
    h.build()
    h.factor()
    h.solve()
    _, _, lin = h.forces()
    assert info.iterations == 1 and abs(info.cl - lin["cl"]) <= 1e-10 * abs(lin["cl"])
