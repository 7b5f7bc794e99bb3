"""Tangent of the FP64 residual, jv = (dR/dq) v (afx_rans_tangent), through
the C ABI:

- against central differences w . (R(q + h v) - R(q - h v)) / 2h of the
  device's FP64 residual (which matches the oracle to 1e-12) for every model
  (inviscid, laminar, Spalart-Allmaras with a rho nu~ direction), both
  gradient schemes and both orders, h = 1e-7 relative to max |q|, random
  unit seeds on the real cells (relative error <= 1e-6);
- the tangent/adjoint dot-product test <w, J v> = <J^T w, v> against the
  hand-written reverse sweep (afx_rans_adjoint: inviscid and laminar,
  Green-Gauss): |<w, Jv> - <J^T w, v>| <= 1e-12 |w| |Jv|, and the laminar
  reverse sweep against central differences;
- a torch.autograd.Function over the C ABI (test harness only; CPU float64
  tensors: forward = fillRhoRHS of the real rows, backward = the adjoint,
  jvp = the tangent), checked by torch.autograd.gradcheck in fast mode (a
  central-difference check of a random projection) in reverse and forward
  mode.
"""
import numpy as np
import pytest

from cases import QUAD, TRI, make_pair

pytestmark = pytest.mark.gpu


def unit_on_real(rng, n_real, n_cells, nv=4):
    x = np.zeros(nv * n_cells)
    x[:nv * n_real] = rng.standard_normal(nv * n_real)
    return x


def sa_case(path, wall="wall", gradient="green-gauss", second_order=True):
    import test_gpu_sa
    s, o = test_gpu_sa.sa_pair(path, wall=wall, gradient=gradient, second_order=second_order)
    return s


def residual_fn(s):
    from aeroflex_amd import rans

    def R(q, nu=None):
        s.set_q(q)
        if nu is not None:
            s.set_nu(nu)
        r, _ = s.fillRhoRHS(rans.AFX_FP64)
        return (r, s.nu_residual()) if nu is not None else r
    return R


CASES = [("inviscid", QUAD, "slip-wall", "green-gauss", True),
         ("inviscid", TRI, "wall", "least-squares", True),
         ("inviscid", QUAD, "slip-wall", "green-gauss", False),
         ("laminar", TRI, "wall", "green-gauss", True),
         ("laminar", QUAD, "slip-wall", "least-squares", True)]


@pytest.mark.parametrize("model,path,wall,gradient,second_order", CASES)
def test_tangent_matches_central_differences(model, path, wall, gradient, second_order):
    s, o, q = make_pair(path, viscosity=model, wall=wall, gradient=gradient, second_order=second_order,
                        mu_L=5e-3 if model == "laminar" else 1e-5)
    R = residual_fn(s)
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(21)
    scale = max(1.0, np.max(np.abs(q)))
    h = 1e-7
    for _ in range(4):
        v = unit_on_real(rng, nr, nt)
        v /= np.linalg.norm(v)
        s.set_q(q)
        jv = s.tangent(v)
        assert np.all(jv[4 * nr:] == 0)
        d = (R(q + h * scale * v) - R(q - h * scale * v)) / (2 * h * scale)
        for _ in range(3):
            w = unit_on_real(rng, nr, nt)
            ad, fd = float(w @ jv), float(w @ d)
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(w) * np.linalg.norm(jv)), (ad, fd)


# Slip walls only: on a no-slip wall face the ghost mirrors the velocity, so
# the mass flux f0 is nearly zero and the rho nu~ convection, upwinded by its
# sign with the wall ghost's -nu~, is |f0| nu~ / rho - a kink at the operating
# point, which central differences straddle while the tangent takes the
# active side.
@pytest.mark.parametrize("path,wall,gradient", [(TRI, "slip-wall", "green-gauss"),
                                                (QUAD, "slip-wall", "least-squares")])
def test_tangent_spalart_allmaras_matches_central_differences(path, wall, gradient):
    s = sa_case(path, wall=wall, gradient=gradient)
    R = residual_fn(s)
    q, nu = s.get_q(), s.get_nu()
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(22)
    for _ in range(3):
        v = unit_on_real(rng, nr, nt)
        vn = np.zeros(nt)
        vn[:nr] = rng.standard_normal(nr) * np.abs(nu[:nr]).mean() / max(np.abs(q).mean(), 1e-30)
        nv = np.sqrt(v @ v + vn @ vn)
        v, vn = v / nv, vn / nv
        sq = max(1.0, np.max(np.abs(q)))
        s.set_q(q)
        s.set_nu(nu)
        jv, jn = s.tangent(v, vn)
        h = 1e-7 * sq
        rp, tp = R(q + h * v, nu + h * vn)
        rm, tm = R(q - h * v, nu - h * vn)
        d, dn = (rp - rm) / (2 * h), (tp - tm) / (2 * h)
        for _ in range(3):
            w = unit_on_real(rng, nr, nt)
            wn = np.zeros(nt)
            wn[:nr] = rng.standard_normal(nr)
            ad = float(w @ jv) + float(wn @ jn)
            fd = float(w @ d) + float(wn @ dn)
            wj = np.sqrt(w @ w + wn @ wn) * np.sqrt(jv @ jv + jn @ jn)
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * wj), (ad, fd)


@pytest.mark.parametrize("model,path,wall,second_order", [("inviscid", QUAD, "slip-wall", True),
                                                          ("inviscid", TRI, "wall", True),
                                                          ("inviscid", TRI, "slip-wall", False),
                                                          ("laminar", TRI, "wall", True),
                                                          ("laminar", QUAD, "slip-wall", True),
                                                          ("laminar", TRI, "wall", False)])
def test_tangent_adjoint_dot_product(model, path, wall, second_order):
    s, o, q = make_pair(path, viscosity=model, wall=wall, second_order=second_order,
                        mu_L=5e-3 if model == "laminar" else 1e-5)
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(23)
    for _ in range(4):
        v = rng.standard_normal(4 * nt)    # ghost rows are ignored by both
        w = rng.standard_normal(4 * nt)
        s.set_q(q)
        jv = s.tangent(v)
        s.set_q(q)
        lam = s.adjoint(w)
        lhs, rhs = float(w[:4 * nr] @ jv[:4 * nr]), float(lam[:4 * nr] @ v[:4 * nr])
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(jv), (lhs, rhs)


@pytest.mark.parametrize("path,wall,second_order", [(TRI, "wall", True), (QUAD, "slip-wall", False)])
def test_laminar_adjoint_matches_central_differences(path, wall, second_order):
    """The laminar terms in the reverse passes: lambda . v against central
    differences of w . R, as for the inviscid adjoint."""
    s, o, q = make_pair(path, viscosity="laminar", wall=wall, second_order=second_order, mu_L=5e-3)
    R = residual_fn(s)
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(25)
    scale = max(1.0, np.max(np.abs(q)))
    h = 1e-7
    for _ in range(3):
        w = unit_on_real(rng, nr, nt)
        w /= np.linalg.norm(w)
        s.set_q(q)
        lam = s.adjoint(w)
        assert np.all(lam[4 * nr:] == 0)
        for _ in range(3):
            v = unit_on_real(rng, nr, nt)
            v /= np.linalg.norm(v)
            fd = float(w @ (R(q + h * scale * v) - R(q - h * scale * v))) / (2 * h * scale)
            ad = float(lam @ v)
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(lam)), (ad, fd)


def make_residual_function(s, q_ghost):
    """torch.autograd.Function over the C ABI: input the real-cell rows of q
    (AoS, float64, CPU), output those rows of fillRhoRHS; the ghost rows are
    boundary data held fixed."""
    import torch
    nr = s.n_real
    from aeroflex_amd import rans

    def full(qr):
        q = q_ghost.copy()
        q[:4 * nr] = qr
        return q

    class Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, qr):
            x = qr.detach().numpy()
            ctx.q = x.copy()
            s.set_q(full(x))
            r, _ = s.fillRhoRHS(rans.AFX_FP64)
            return torch.from_numpy(r[:4 * nr].copy())

        @staticmethod
        def backward(ctx, g):
            s.set_q(full(ctx.q))
            w = np.zeros(4 * s.n_cells)
            w[:4 * nr] = g.detach().numpy()
            return torch.from_numpy(s.adjoint(w)[:4 * nr].copy())

        @staticmethod
        def jvp(ctx, dq):
            s.set_q(full(ctx.q))
            v = np.zeros(4 * s.n_cells)
            v[:4 * nr] = dq.detach().numpy()
            return torch.from_numpy(s.tangent(v)[:4 * nr].copy())

    return Residual


def test_autograd_function_gradcheck():
    import torch
    s, o, q = make_pair(QUAD, wall="slip-wall", second_order=True)
    nr = s.n_real
    Residual = make_residual_function(s, q.copy())
    x = torch.tensor(q[:4 * nr], dtype=torch.float64, requires_grad=True)
    # central differences of a random projection (fast mode), reverse and forward mode
    assert torch.autograd.gradcheck(Residual.apply, (x,), eps=1e-7, atol=1e-6, rtol=1e-5, fast_mode=True,
                                    check_forward_ad=True, check_backward_ad=True, check_undefined_grad=False,
                                    check_batched_grad=False)
    # backward() through the wrapper is the adjoint
    rng = np.random.default_rng(24)
    wt = torch.tensor(rng.standard_normal(4 * nr))
    (Residual.apply(x) * wt).sum().backward()
    w = np.zeros(4 * s.n_cells)
    w[:4 * nr] = wt.numpy()
    s.set_q(q)
    np.testing.assert_allclose(x.grad.numpy(), s.adjoint(w)[:4 * nr], rtol=1e-14, atol=1e-14)
