"""State adjoint of the FP64 Spalart-Allmaras residual (afx_rans_adjoint_nu,
the five reverse passes of afx_adjoint_sa.hip), through the C ABI:

- the tangent/adjoint dot-product test <w, J v> = <J^T w, v> over all five
  rows (rho nu~ seed and output included) against the forward-mode tangent
  (afx_rans_tangent, itself checked against central differences), on the
  reference's triangle and quad meshes, no-slip and slip walls, first and
  second order: |<w, Jv> - <J^T w, v>| <= 1e-11 |w| |Jv|;
- lambda . v against central differences of w . R along random unit
  directions (slip walls: on a no-slip wall the rho nu~ convection is |f0|
  nu~ / rho with a nearly zero mass flux f0, a kink that central differences
  straddle), relative error <= 1e-6;
- the ghost rows of lambda are zero.
"""
import numpy as np
import pytest

from cases import QUAD, TRI

pytestmark = pytest.mark.gpu


def sa_case(path, wall, second_order):
    import test_gpu_sa
    s, _ = test_gpu_sa.sa_pair(path, wall=wall, second_order=second_order)
    return s


def real_seed(rng, nr, nt, nv=4):
    x = np.zeros(nv * nt)
    x[:nv * nr] = rng.standard_normal(nv * nr)
    return x


@pytest.mark.parametrize("path,wall,second_order", [(TRI, "wall", True), (QUAD, "slip-wall", True),
                                                    (TRI, "slip-wall", False), (QUAD, "wall", False),
                                                    (QUAD, "wall", True)])
def test_sa_adjoint_dot_product(path, wall, second_order):
    s = sa_case(path, wall, second_order)
    q, nu = s.get_q(), s.get_nu()
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(41)
    for _ in range(3):
        v, vn = real_seed(rng, nr, nt), real_seed(rng, nr, nt, 1)
        w, wn = real_seed(rng, nr, nt), real_seed(rng, nr, nt, 1)
        s.set_q(q)
        s.set_nu(nu)
        jv, jn = s.tangent(v, vn)
        s.set_q(q)
        s.set_nu(nu)
        lam, ln = s.adjoint(w, w_nu=wn)
        assert np.all(lam[4 * nr:] == 0) and np.all(ln[nr:] == 0)
        lhs = float(w @ jv) + float(wn @ jn)
        rhs = float(lam @ v) + float(ln @ vn)
        scale = np.sqrt(w @ w + wn @ wn) * np.sqrt(jv @ jv + jn @ jn)
        assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs)
    s.close()


@pytest.mark.parametrize("path,second_order", [(TRI, True), (QUAD, True), (QUAD, False)])
def test_sa_adjoint_matches_central_differences(path, second_order):
    from aeroflex_amd import rans
    s = sa_case(path, "slip-wall", second_order)
    q, nu = s.get_q(), s.get_nu()
    nr, nt = s.n_real, s.n_cells
    rng = np.random.default_rng(43)

    def wR(qq, nn, w, wn):
        s.set_q(qq)
        s.set_nu(nn)
        r, _ = s.fillRhoRHS(rans.AFX_FP64)
        return float(w @ r) + float(wn @ s.nu_residual())

    sq = max(1.0, float(np.max(np.abs(q))))
    h = 1e-7 * sq
    for _ in range(2):
        w, wn = real_seed(rng, nr, nt), real_seed(rng, nr, nt, 1)
        s.set_q(q)
        s.set_nu(nu)
        lam, ln = s.adjoint(w, w_nu=wn)
        for _ in range(3):
            v = real_seed(rng, nr, nt)
            vn = real_seed(rng, nr, nt, 1) * np.abs(nu[:nr]).mean() / max(np.abs(q).mean(), 1e-30)
            nv = np.sqrt(v @ v + vn @ vn)
            v, vn = v / nv, vn / nv
            fd = (wR(q + h * v, nu + h * vn, w, wn) - wR(q - h * v, nu - h * vn, w, wn)) / (2 * h)
            ad = float(lam @ v) + float(ln @ vn)
            floor = 1e-3 * np.sqrt(lam @ lam + ln @ ln)
            assert abs(ad - fd) <= 1e-6 * max(abs(fd), floor), (ad, fd)
    s.close()
