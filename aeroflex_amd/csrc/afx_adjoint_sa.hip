// State adjoint of the FP64 Spalart-Allmaras residual: lambda = (dR/dq)^T w
// over the five rows (rho, rho u, rho v, rho E, rho nu~) of the fused tile
// residual with the SA model (k_residual_tile<double, 5>), Green-Gauss
// gradients (solver.h:428-469), the default limiter (solver.h:517-593),
// first or second order.
//
// The same five atomics-free passes as the inviscid and laminar adjoint
// (afx_adjoint.hip), over five variables:
//   1 cells  gradients, limiters and the limiter's active branch
//   2 faces  reconstruction, then the adjoint of the five-variable face flux
//            (sa_face_flux: Roe, upwind rho nu~ convection, viscous terms with
//            mu_L + mu_t at the face gradients, nu~ diffusion, the no-slip
//            wall's nu~ diffusion) with respect to the reconstructed states,
//            the face gradients and the owner's cell-centre nu~
//   3 cells  gather: direct state adjoint, gradient adjoint (reconstruction,
//            face gradients, and the cell source sa_cell_source with its
//            unlimited gradients), limiter adjoint on the active branch
//   4 faces  transpose of the Green-Gauss face values (boundary faces: the
//            face states of face_vars and sa_bnd_nu)
//   5 cells  gather of pass 4 and of the neighbours' limiter parts
// The Roe part of the flux goes through the hand-written reverse sweep of the
// inviscid adjoint; the diffusive part (sa_face_diffusion), the cell source,
// the boundary faces and their face values through forward-mode sweeps (one
// per input) of the same device functions the residual evaluates, so every
// local Jacobian is exact; the structure around them is transposed by hand.
// Wall / slip ghost states enter nothing (their face states are functions of
// the owner's); farfield ghost states are boundary data and get no adjoint.
#include "afx_device.hpp"
#include "afx_adjoint.hpp"
#include "afx_adjoint_kernels.hpp"

namespace afx {

template <> __device__ __forceinline__ Dual1 ddiv<Dual1>(Dual1 a, Dual1 b) { return a / b; }
template <> __device__ __forceinline__ Dual1 dsqrt<Dual1>(Dual1 x) { return dual_sqrt(x); }
template <> __device__ __forceinline__ Dual1 dcbrt<Dual1>(Dual1 x) { return dual_cbrt(x); }

namespace {

constexpr int NS = 5;

__device__ __forceinline__ GasT<Dual1> sa_dual_gas(const GasT<double>& g) {
    return GasT<Dual1>{Dual1(g.gamma), Dual1(g.R), Dual1(g.mu), Dual1(g.k), Dual1(g.cp), Dual1(g.PrL), Dual1(g.PrT)};
}

// Green-Gauss face value of the five variables on a boundary face from the
// owner's state qo (5) and the ghost's qn (5): face_vars for the mean flow,
// sa_bnd_nu for rho nu~ (as k_residual_tile stages them)
template <typename T>
__device__ __forceinline__ void bnd_face_value(const GasT<T>& g, int kind, T nx, T ny, const T* qo, const T* qn,
                                               T* qv) {
    face_vars<T>(g, kind, nx, ny, qo, qn, qv);
    qv[4] = sa_bnd_nu<T>(kind, nx, ny, qo, qo[4], qn, qn[4], qv);
}

// Reverse sweep of sa_face_diffusion (FP64): given the adjoints bF[1..4] of
// the four flux rows it changes (the rows it subtracts from), adds the
// adjoints of its inputs to bL, bR (5 each) and writes those of the face
// gradients to bgx, bgy (5 each).  Every output depends on q_L and q_R only
// through the face averages qc = (q_L + q_R) / 2, so both sides receive half
// of qc's adjoint.  The derivatives are those of the mathematical functions
// (the forward's algebraic form for the pressure gradient differs in
// rounding only).
__device__ __forceinline__ void sa_face_diffusion_vjp(const GasT<double>& g, double nx, double ny, const double* qL,
                                                      const double* qR, const double* gx, const double* gy,
                                                      const double* bF, double* bL, double* bR, double* bgx,
                                                      double* bgy) {
    double qc[5];
    for (int k = 0; k < 5; ++k) qc[k] = 0.5 * (qL[k] + qR[k]);
    const double gm1 = g.gamma - 1.0, iR = 1.0 / g.R;
    // forward values
    const double nc = qc[4], chi = nc / g.mu, c3 = chi * chi * chi;
    const double den_v = c3 + sa::cv1_3;
    const double fv1 = chi <= 0.0 ? 0.0 : c3 / den_v;
    const double dfv1 = chi <= 0.0 ? 0.0 : 3.0 * chi * chi * sa::cv1_3 / (den_v * den_v);
    const double mut = nc * fv1, mue = g.mu + mut;
    const double kc = g.cp * (g.mu / g.PrL + mut / g.PrT);
    const double r = qc[0], ir = 1.0 / r, r2 = r * r;
    const double u = qc[1] * ir, v = qc[2] * ir;
    const double ke2 = qc[1] * qc[1] + qc[2] * qc[2];
    const double p = gm1 * (qc[3] - 0.5 * ir * ke2);
    const double X0 = 2.0 * gx[3] - 2.0 * gx[1] * u - 2.0 * gx[2] * v + gx[0] * ke2 / r2;
    const double X1 = 2.0 * gy[3] - 2.0 * gy[1] * u - 2.0 * gy[2] * v + gy[0] * ke2 / r2;
    const double gp0 = 0.5 * gm1 * X0, gp1 = 0.5 * gm1 * X1;
    const double gT0 = iR * (gp0 * r - gx[0] * p) / r2, gT1 = iR * (gp1 * r - gy[0] * p) / r2;
    const double gu0 = (r * gx[1] - qc[1] * gx[0]) / r2, gu1 = (r * gy[1] - qc[1] * gy[0]) / r2;
    const double gv0 = (r * gx[2] - qc[2] * gx[0]) / r2, gv1 = (r * gy[2] - qc[2] * gy[0]) / r2;
    const double div3 = (gu0 + gv1) / 3.0;
    const double txx = 2.0 * mue * (gu0 - div3), tyy = 2.0 * mue * (gv1 - div3), txy = mue * (gu1 + gv0);
    const double nut = nc * ir;
    const double gvx = (gx[4] - nut * gx[0]) * ir, gvy = (gy[4] - nut * gy[0]) * ir;
    const double den_n = sa::cn1 - c3;
    const double fn = chi >= 0.0 ? 1.0 : (sa::cn1 + c3) / den_n;
    const double dfn = chi >= 0.0 ? 0.0 : 6.0 * sa::cn1 * chi * chi / (den_n * den_n);
    const double dcoef = (g.mu + nc * fn) / sa::sigma;
    const double S = nx * gvx + ny * gvy;
    // reverse: the rows subtract their terms, so the terms' adjoints are -bF
    const double B1 = -bF[1], B2 = -bF[2], B3 = -bF[3], B4 = -bF[4];
    const double bphi0 = B3 * nx, bphi1 = B3 * ny;
    const double btxx = B1 * nx + bphi0 * u;
    const double btxy = B1 * ny + B2 * nx + bphi0 * v + bphi1 * u;
    const double btyy = B2 * ny + bphi1 * v;
    const double bkc = bphi0 * gT0 + bphi1 * gT1;
    const double bgT0 = bphi0 * kc, bgT1 = bphi1 * kc;
    double bu = bphi0 * txx + bphi1 * txy;
    double bv = bphi0 * txy + bphi1 * tyy;
    const double bmue = btxx * 2.0 * (gu0 - div3) + btyy * 2.0 * (gv1 - div3) + btxy * (gu1 + gv0);
    const double bgu0 = btxx * mue * (4.0 / 3.0) - btyy * mue * (2.0 / 3.0);
    const double bgv1 = btyy * mue * (4.0 / 3.0) - btxx * mue * (2.0 / 3.0);
    const double bgu1 = btxy * mue, bgv0 = btxy * mue;
    double bq[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 5; ++k) { bgx[k] = 0.0; bgy[k] = 0.0; }
    // velocity gradients g = (r gk - qk g0) / r^2
    bq[0] += bgu0 * (gx[1] / r2 - 2.0 * gu0 * ir) + bgu1 * (gy[1] / r2 - 2.0 * gu1 * ir) +
             bgv0 * (gx[2] / r2 - 2.0 * gv0 * ir) + bgv1 * (gy[2] / r2 - 2.0 * gv1 * ir);
    bq[1] -= (bgu0 * gx[0] + bgu1 * gy[0]) / r2;
    bq[2] -= (bgv0 * gx[0] + bgv1 * gy[0]) / r2;
    bgx[1] += bgu0 * ir;
    bgy[1] += bgu1 * ir;
    bgx[2] += bgv0 * ir;
    bgy[2] += bgv1 * ir;
    bgx[0] -= (bgu0 * qc[1] + bgv0 * qc[2]) / r2;
    bgy[0] -= (bgu1 * qc[1] + bgv1 * qc[2]) / r2;
    // temperature gradients gT = iR (gp r - g0 p) / r^2
    const double bgp0 = bgT0 * iR * ir, bgp1 = bgT1 * iR * ir;
    bq[0] += bgT0 * (iR * gp0 / r2 - 2.0 * gT0 * ir) + bgT1 * (iR * gp1 / r2 - 2.0 * gT1 * ir);
    bgx[0] -= bgT0 * iR * p / r2;
    bgy[0] -= bgT1 * iR * p / r2;
    const double bp = -(bgT0 * gx[0] + bgT1 * gy[0]) * iR / r2;
    // pressure gradients gp = gm1/2 (2 g3 - 2 u g1 - 2 v g2 + |qc12|^2/r^2 g0)
    const double bX0 = 0.5 * gm1 * bgp0, bX1 = 0.5 * gm1 * bgp1;
    bgx[3] += 2.0 * bX0;
    bgy[3] += 2.0 * bX1;
    bgx[1] -= 2.0 * bX0 * u;
    bgy[1] -= 2.0 * bX1 * u;
    bgx[2] -= 2.0 * bX0 * v;
    bgy[2] -= 2.0 * bX1 * v;
    bgx[0] += bX0 * ke2 / r2;
    bgy[0] += bX1 * ke2 / r2;
    bu -= 2.0 * (bX0 * gx[1] + bX1 * gy[1]);
    bv -= 2.0 * (bX0 * gx[2] + bX1 * gy[2]);
    const double bke2 = (bX0 * gx[0] + bX1 * gy[0]) / r2;
    bq[0] -= 2.0 * (bX0 * gx[0] + bX1 * gy[0]) * ke2 / (r2 * r);
    // p = gm1 (qc3 - ke2 / (2 r))
    bq[3] += bp * gm1;
    const double bke2p = bke2 - bp * gm1 * 0.5 * ir;
    bq[0] += bp * gm1 * 0.5 * ke2 / r2;
    bq[1] += 2.0 * bke2p * qc[1];
    bq[2] += 2.0 * bke2p * qc[2];
    // u = qc1 / r, v = qc2 / r
    bq[1] += bu * ir;
    bq[2] += bv * ir;
    bq[0] -= (bu * u + bv * v) * ir;
    // mu_t through mu_e and kc, then nc = qc4
    const double bmut = bmue + bkc * g.cp / g.PrT;
    bq[4] += bmut * (fv1 + nc * dfv1 / g.mu);
    // nu~ diffusion: B4 dcoef (n . gv)
    const double bdc = B4 * S, bS = B4 * dcoef;
    bq[4] += bdc * (fn + nc * dfn / g.mu) / sa::sigma;
    const double bgvx = bS * nx, bgvy = bS * ny;
    bgx[4] += bgvx * ir;
    bgy[4] += bgvy * ir;
    const double bnut = -(bgvx * gx[0] + bgvy * gy[0]) * ir;
    bgx[0] -= bgvx * nut * ir;
    bgy[0] -= bgvy * nut * ir;
    bq[0] -= (bgvx * gvx + bgvy * gvy) * ir;
    bq[4] += bnut * ir;
    bq[0] -= bnut * nut * ir;
    for (int k = 0; k < 5; ++k) {
        bL[k] += 0.5 * bq[k];
        bR[k] += 0.5 * bq[k];
    }
}

// Reverse sweep of sa_cell_source (FP64): adds ws times the derivatives of
// the per-volume source S(q, grad q, d) to lq (the state, 5), lgx (x
// gradients of rho, rho v, rho nu~) and lgy (y gradients of rho, rho u,
// rho nu~), the only gradients S reads.  The branches are the forward's
// (negative nu~, the S-tilde limiter, the cap r <= 10, |omega| with slope +1
// at zero); f_w'(r) = g'(r) c / (g^6 + c) ((1 + c) / (g^6 + c))^(1/6),
// c = c_w3^6.
__device__ __forceinline__ void sa_cell_source_vjp(const GasT<double>& g, const double* q, const double* gx,
                                                   const double* gy, double d, double ws, double* lq, double* lgx,
                                                   double* lgy) {
    const double rho = q[0], ir = 1.0 / rho, nut = q[4] * ir;
    const double dudy = (gy[1] - q[1] * ir * gy[0]) * ir;
    const double dvdx = (gx[2] - q[2] * ir * gx[0]) * ir;
    const double om_raw = dvdx - dudy;
    const double omega = om_raw < 0.0 ? -om_raw : om_raw, som = om_raw < 0.0 ? -1.0 : 1.0;
    const double gvx = (gx[4] - nut * gx[0]) * ir, gvy = (gy[4] - nut * gy[0]) * ir;
    const double cc = sa::cb2 / sa::sigma;
    const double cross = cc * (gvx * gvx + gvy * gvy);
    const double nd = nut / d;
    double b_rho = 0.0, b_nut = 0.0, b_om = 0.0;
    const double b_gvx = ws * rho * cc * 2.0 * gvx, b_gvy = ws * rho * cc * 2.0 * gvy;
    if (nut < 0.0) {
        const double a1 = sa::cb1 * (1.0 - sa::ct3);
        const double E = a1 * omega * nut + sa::cw1 * nd * nd + cross;
        b_rho = ws * E;
        b_nut = ws * rho * (a1 * omega + 2.0 * sa::cw1 * nd / d);
        b_om = ws * rho * a1 * nut;
    } else {
        const double chi = rho * nut / g.mu, c3 = chi * chi * chi;
        const double dv = c3 + sa::cv1_3;
        const double f1 = chi <= 0.0 ? 0.0 : c3 / dv;
        const double df1 = chi <= 0.0 ? 0.0 : 3.0 * chi * chi * sa::cv1_3 / (dv * dv);
        const double den2 = 1.0 + chi * f1;
        const double f2 = 1.0 - chi / den2;
        const double df2 = -(1.0 - chi * chi * df1) / (den2 * den2);
        const double kd2 = sa::kappa * sa::kappa * d * d;
        const double sbar = nut * f2 / kd2;
        double st, dst_dom, dst_dsb;
        if (sbar >= -sa::cv2 * omega) {
            st = omega + sbar;
            dst_dom = 1.0;
            dst_dsb = 1.0;
        } else {
            const double N = omega * (sa::cv2 * sa::cv2 * omega + sa::cv3 * sbar);
            const double Dn = (sa::cv3 - 2.0 * sa::cv2) * omega - sbar;
            st = omega + N / Dn;
            const double dN_dom = 2.0 * sa::cv2 * sa::cv2 * omega + sa::cv3 * sbar, dN_dsb = sa::cv3 * omega;
            dst_dom = 1.0 + (dN_dom * Dn - N * (sa::cv3 - 2.0 * sa::cv2)) / (Dn * Dn);
            dst_dsb = (dN_dsb * Dn + N) / (Dn * Dn);
        }
        double r = 10.0, dr_dnut = 0.0, dr_dst = 0.0;
        if (st > 0.0) {
            const double r0 = nut / (st * kd2);
            if (r0 < 10.0) {
                r = r0;
                dr_dnut = 1.0 / (st * kd2);
                dr_dst = -r0 / st;
            }
        }
        const double r2 = r * r, gg = r + sa::cw2 * (r2 * r2 * r2 - r);
        const double g2 = gg * gg, g6 = g2 * g2 * g2;
        const double h6 = cbrt(sqrt((1.0 + sa::cw3_6) / (g6 + sa::cw3_6)));
        const double fw = gg * h6;
        const double dfw = (1.0 + sa::cw2 * (6.0 * r2 * r2 * r - 1.0)) * h6 * sa::cw3_6 / (g6 + sa::cw3_6);
        const double P = sa::cb1 * st * nut, D = sa::cw1 * fw * nd * nd;
        const double E = P - D + cross;
        const double bE = ws * rho;
        b_rho = ws * E;
        const double b_r = -bE * sa::cw1 * dfw * nd * nd;
        const double b_st = bE * sa::cb1 * nut + b_r * dr_dst;
        b_nut = bE * (sa::cb1 * st - 2.0 * sa::cw1 * fw * nd / d) + b_r * dr_dnut;
        const double b_sb = b_st * dst_dsb;
        b_om = b_st * dst_dom;
        b_nut += b_sb * f2 / kd2;
        const double b_chi = b_sb * nut / kd2 * df2;
        b_rho += b_chi * nut / g.mu;
        b_nut += b_chi * rho / g.mu;
    }
    // omega = |dvdx - dudy|, the velocity derivatives and grad nu~ = (grad
    // rho nu~ - nu~ grad rho) / rho, then nu~ = (rho nu~) / rho
    const double b_dv = b_om * som, b_du = -b_om * som;
    lgx[2] += b_dv * ir;
    lq[2] -= b_dv * gx[0] * ir * ir;
    lgx[0] -= b_dv * q[2] * ir * ir;
    b_rho += b_dv * (-gx[2] * ir * ir + 2.0 * q[2] * gx[0] * ir * ir * ir);
    lgy[1] += b_du * ir;
    lq[1] -= b_du * gy[0] * ir * ir;
    lgy[0] -= b_du * q[1] * ir * ir;
    b_rho += b_du * (-gy[1] * ir * ir + 2.0 * q[1] * gy[0] * ir * ir * ir);
    lgx[4] += b_gvx * ir;
    lgy[4] += b_gvy * ir;
    lgx[0] -= b_gvx * nut * ir;
    lgy[0] -= b_gvy * nut * ir;
    b_nut -= (b_gvx * gx[0] + b_gvy * gy[0]) * ir;
    b_rho -= (b_gvx * gvx + b_gvy * gvy) * ir;
    lq[4] += b_nut * ir;
    lq[0] += b_rho - b_nut * nut * ir;
}

// pass 1: gradients, limiters and the active branch of every (cell, variable)
__global__ void __launch_bounds__(kBlock) k_sadj_forward(CellsT<double> C, FacesT<double> F,
                                                         const double* __restrict__ q, GasT<double> g,
                                                         int second_order, double limiter_k, AdjBufs A) {
    const int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC;
    double qc[NS], sx[NS], sy[NS], mn[NS], mx[NS];
    int jmn[NS], jmx[NS];
    for (int k = 0; k < NS; ++k) {
        qc[k] = q[k * nT + c];
        mn[k] = mx[k] = qc[k];
        sx[k] = sy[k] = 0.0;
        jmn[k] = jmx[k] = -1;
    }
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
        const double nx = F.nx[f], ny = F.ny[f];
        double qo[NS], qn[NS], qv[NS];
        for (int k = 0; k < NS; ++k) { qo[k] = q[k * nT + o]; qn[k] = q[k * nT + n]; }
        face_vars<double>(g, kind, nx, ny, qo, qn, qv);
        qv[4] = kind == K_INTERNAL ? (qo[4] + qn[4]) * 0.5 : sa_bnd_nu<double>(kind, nx, ny, qo, qo[4], qn, qn[4], qv);
        const double gw = F.gw[f], L = F.len[f];
        for (int k = 0; k < NS; ++k) {
            const double fk = (qo[k] * (1.0 - gw) + qv[k] * gw) * L;
            if (e >> 31) { sx[k] -= fk * nx; sy[k] -= fk * ny; }
            else { sx[k] += fk * nx; sy[k] += fk * ny; }
        }
        if (second_order && kind != K_WALL && kind != K_SLIP) {   // wall ghost = owner copy
            const int other = (e >> 31) ? o : n;
            for (int k = 0; k < NS; ++k) {
                const double v = (e >> 31) ? qo[k] : qn[k];
                if (v < mn[k]) { mn[k] = v; jmn[k] = other; }
                if (mx[k] < v) { mx[k] = v; jmx[k] = other; }
            }
        }
    }
    const double Ar = C.area[c];
    for (int k = 0; k < NS; ++k) {
        sx[k] /= Ar;
        sy[k] /= Ar;
        A.gx[k * nT + c] = sx[k];
        A.gy[k * nT + c] = sy[k];
    }
    if (!second_order) return;
    const double Ka = limiter_k * sqrt(Ar), K3a = Ka * Ka * Ka;
    double lm[NS];
    int act[NS];
    for (int k = 0; k < NS; ++k) { lm[k] = 1.0; act[k] = -1; }
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const double dx = F.xf[f] - C.xc[c], dy = F.yf[f] - C.yc[c];
        for (int k = 0; k < NS; ++k) {
            const double dqg = sx[k] * dx + sy[k] * dy;
            const double l = limiter_phi<double>(dqg, mx[k] - qc[k], mn[k] - qc[k], K3a);
            if (l < lm[k]) { lm[k] = l; act[k] = s; }
        }
    }
    for (int k = 0; k < NS; ++k) {
        A.lim[k * nT + c] = lm[k];
        A.act[k * nC + c] = (int8_t)act[k];
        A.jmax[k * nC + c] = jmx[k];
        A.jmin[k * nC + c] = jmn[k];
    }
}

// pass 2: per face, adjoints of the reconstructed states (bL, bR), of the
// face gradients (vgx, vgy: half of them, owed to both cells' gradients) and
// of the direct cell-state dependences (vdq: the face gradient's difference
// quotient, owner sign, and the wall diffusion's cell-centre nu~).  This
// kernel takes the interior faces; k_sadj_face_bnd the boundary faces.
__global__ void __launch_bounds__(kBlock) k_sadj_face(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                      const double* __restrict__ w, GasT<double> g, int second_order,
                                                      AdjBufs A) {
    const int f = xcd_index();
    if (f >= F.nF || F.kind[f] != K_INTERNAL) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f];
    double qL[NS], qR[NS];
    for (int k = 0; k < NS; ++k) { qL[k] = q[k * nT + o]; qR[k] = q[k * nT + n]; }
    if (second_order) {
        const double d0x = F.xf[f] - C.xc[o], d0y = F.yf[f] - C.yc[o];
        const double d1x = F.xf[f] - C.xc[n], d1y = F.yf[f] - C.yc[n];
        for (int k = 0; k < NS; ++k) {
            qL[k] += (A.gx[k * nT + o] * d0x + A.gy[k * nT + o] * d0y) * A.lim[k * nT + o];
            qR[k] += (A.gx[k * nT + n] * d1x + A.gy[k * nT + n] * d1y) * A.lim[k * nT + n];
        }
    }
    double bF[NS], bL[NS], bR[NS], vgx[NS], vgy[NS], vdq[NS];
    for (int k = 0; k < NS; ++k) {
        bF[k] = L * (w[k * nT + n] - w[k * nT + o]);
        bL[k] = bR[k] = 0.0;
    }
    // face gradients of the SA model: textbook direction t = (x_n - x_o) /
    // l and derivative (q_n - q_o) / l, g = gbar - (gbar . t - dq / l) t
    double t0 = C.xc[n] - C.xc[o], t1 = C.yc[n] - C.yc[o];
    const double l = sqrt(t0 * t0 + t1 * t1);
    t0 /= l;
    t1 /= l;
    double gxa[NS], gya[NS];
    for (int k = 0; k < NS; ++k) {
        const double dir = (q[k * nT + n] - q[k * nT + o]) / l;
        const double gxb = (A.gx[k * nT + o] + A.gx[k * nT + n]) * 0.5;
        const double gyb = (A.gy[k * nT + o] + A.gy[k * nT + n]) * 0.5;
        const double dot = gxb * t0 + gyb * t1;
        gxa[k] = gxb - (dot - dir) * t0;
        gya[k] = gyb - (dot - dir) * t1;
    }
    // Roe and the rho nu~ convection f4 = f0 s (s = nu~ of the upwind
    // side): the hand-written reverse sweep of the Roe flux with the
    // mass-flux adjoint bF0 + bF4 s, then s's own dependence
    double fr[4];
    roe_flux<double>(g, nx, ny, qL, qR, false, nullptr, nullptr, fr);
    const bool up = fr[0] >= 0.0;
    const double s = up ? qL[4] / qL[0] : qR[4] / qR[0];
    const double bRoe[4] = {bF[0] + bF[4] * s, bF[1], bF[2], bF[3]};
    roe_flux_vjp<double>(g.gamma, nx, ny, qL, qR, bRoe, bL, bR);
    // s's dependence, on the upwind side (selects, not a pointer to one of
    // the two arrays, which would move both to scratch memory)
    const double qu0 = up ? qL[0] : qR[0], qu4 = up ? qL[4] : qR[4];
    const double d4 = bF[4] * fr[0] / qu0, d0 = bF[4] * fr[0] * qu4 / (qu0 * qu0);
    bL[4] = up ? bL[4] + d4 : bL[4];
    bR[4] = up ? bR[4] : bR[4] + d4;
    bL[0] = up ? bL[0] - d0 : bL[0];
    bR[0] = up ? bR[0] : bR[0] - d0;
    // the diffusive part (sa_face_diffusion): its hand-written reverse
    double bgx[NS], bgy[NS];
    sa_face_diffusion_vjp(g, nx, ny, qL, qR, gxa, gya, bF, bL, bR, bgx, bgy);
    for (int k = 0; k < NS; ++k) {
        const double bt = bgx[k] * t0 + bgy[k] * t1;   // the adjoint of dq / l
        vgx[k] = 0.5 * (bgx[k] - t0 * bt);
        vgy[k] = 0.5 * (bgy[k] - t1 * bt);
        vdq[k] = -bt / l;                              // owner's share; the neighbour takes -vdq
    }
    for (int k = 0; k < NS; ++k) {
        A.bL[k * nF + f] = bL[k];
        A.bR[k * nF + f] = bR[k];
        A.vgx[k * nF + f] = vgx[k];
        A.vgy[k * nF + f] = vgy[k];
        A.vdq[k * nF + f] = vdq[k];
    }
}

// pass 2, boundary faces: the flux of the reconstructed owner state and the
// ghost's boundary data, swept forward once per input (the owner state and
// the owner's cell-centre nu~ = (rho nu~) / rho that the wall diffusion reads
// with its normal distance d0 . n).  One thread per (face, input): the six
// sweeps of a face run side by side instead of one after another, since the
// few boundary faces leave the launch bound by one thread's FP64 chain.
__global__ void __launch_bounds__(kBlock) k_sadj_face_bnd(CellsT<double> C, FacesT<double> F,
                                                          const double* __restrict__ q, const double* __restrict__ w,
                                                          GasT<double> g, int second_order, AdjBufs A) {
    constexpr int NJ = NS + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.nbf * NJ) return;
    const int j = i % NJ, f = A.bface[i / NJ];
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f];
    const double d0x = F.xf[f] - C.xc[o], d0y = F.yf[f] - C.yc[o];
    double qL[NS], qR[NS];
    for (int k = 0; k < NS; ++k) { qL[k] = q[k * nT + o]; qR[k] = q[k * nT + n]; }
    if (second_order)
        for (int k = 0; k < NS; ++k)
            qL[k] += (A.gx[k * nT + o] * d0x + A.gy[k * nT + o] * d0y) * A.lim[k * nT + o];
    const double qo0 = q[o], qo4 = q[4 * nT + o];
    const double nut_c = qo4 / qo0, dn = d0x * nx + d0y * ny;
    const GasT<Dual1> gd = sa_dual_gas(g);
    Dual1 a[NS], b[NS], z[NS], fo[NS];
    for (int k = 0; k < NS; ++k) {
        a[k] = Dual1(qL[k], j == k ? 1.0 : 0.0);
        b[k] = Dual1(qR[k]);
        z[k] = Dual1(0.0);
    }
    sa_face_flux<Dual1>(gd, kind, Dual1(nx), Dual1(ny), a, b, z, z, Dual1(nut_c, j == NS ? 1.0 : 0.0), Dual1(dn), fo);
    double v = 0.0;
    for (int k = 0; k < NS; ++k) v += -L * w[k * nT + o] * fo[k].d;
    if (j < NS) {
        A.bL[j * nF + f] = v;
        A.bR[j * nF + f] = 0.0;
        A.vgx[j * nF + f] = 0.0;
        A.vgy[j * nF + f] = 0.0;
        if (j != 0 && j != 4) A.vdq[j * nF + f] = 0.0;
    } else {   // the cell-centre nu~ part, owed to the owner's rho and rho nu~
        A.vdq[f] = -v * qo4 / (qo0 * qo0);
        A.vdq[4 * nF + f] = v / qo0;
    }
}

// pass 3: per cell, state / gradient / limiter adjoints, with the cell
// source's adjoint (w_rho nu~ A dS/dq, dS/d grad)
__global__ void __launch_bounds__(kBlock) k_sadj_cell1(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                       const double* __restrict__ w, const double* __restrict__ wd,
                                                       GasT<double> g, int second_order, double limiter_k, AdjBufs A,
                                                       double* __restrict__ lam) {
    const int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC, nF = (size_t)F.nF;
    double lq[NS], lgx[NS], lgy[NS], lphi[NS], gx[NS], gy[NS], ph[NS], qc[NS];
    for (int k = 0; k < NS; ++k) {
        lq[k] = lgx[k] = lgy[k] = lphi[k] = 0.0;
        gx[k] = A.gx[k * nT + c];
        gy[k] = A.gy[k * nT + c];
        ph[k] = second_order ? A.lim[k * nT + c] : 1.0;
        qc[k] = q[k * nT + c];
    }
    double dxs[4] = {0, 0, 0, 0}, dys[4] = {0, 0, 0, 0};   // the slots' offsets, for the limiter's active slot
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const bool nb = (e >> 31) != 0;
        const double dx = F.xf[f] - C.xc[c], dy = F.yf[f] - C.yc[c];
        dxs[s] = dx; dys[s] = dy;
        for (int k = 0; k < NS; ++k) {
            const double b = nb ? A.bR[k * nF + f] : A.bL[k * nF + f];
            lq[k] += b;
            if (second_order) {
                lphi[k] += b * (gx[k] * dx + gy[k] * dy);
                lgx[k] += b * ph[k] * dx;
                lgy[k] += b * ph[k] * dy;
            }
            lgx[k] += A.vgx[k * nF + f];
            lgy[k] += A.vgy[k * nF + f];
            lq[k] += nb ? -A.vdq[k * nF + f] : A.vdq[k * nF + f];
        }
    }
    // the source term A S(q_c, grad q_c, d_c) of the rho nu~ row: its
    // hand-written reverse (the state and the gradients of rho, rho v and
    // rho nu~ in x and of rho, rho u and rho nu~ in y)
    const double ws = w[4 * nT + c] * C.area[c];
    if (ws != 0.0) sa_cell_source_vjp(g, qc, gx, gy, wd[c], ws, lq, lgx, lgy);
    for (int k = 0; k < NS; ++k) {
        double nbv = 0.0;
        int nbj = -1;
        const int s = second_order ? (int)A.act[k * nC + c] : -1;
        if (s >= 0) {
            const double dx = s == 0 ? dxs[0] : s == 1 ? dxs[1] : s == 2 ? dxs[2] : dxs[3];
            const double dy = s == 0 ? dys[0] : s == 1 ? dys[1] : s == 2 ? dys[2] : dys[3];
            const double dqg = gx[k] * dx + gy[k] * dy;
            const int j = dqg > 1e-16 ? A.jmax[k * nC + c] : A.jmin[k * nC + c];
            const double D = (j >= 0 ? q[k * nT + j] : qc[k]) - qc[k];
            const double Ka = limiter_k * sqrt(C.area[c]), K3a = Ka * Ka * Ka;
            double d_dqg, d_D;
            (void)limiter_phi_d<double>(dqg, D, K3a, &d_dqg, &d_D, nullptr);
            lgx[k] += lphi[k] * d_dqg * dx;
            lgy[k] += lphi[k] * d_dqg * dy;
            if (j >= 0) {   // D = q_j - q_c
                lq[k] -= lphi[k] * d_D;
                if (j < C.nC) { nbv = lphi[k] * d_D; nbj = j; }   // a farfield ghost is boundary data
            }
        }
        lam[k * nT + c] = lq[k];
        A.lgx[k * nT + c] = lgx[k];
        A.lgy[k * nT + c] = lgy[k];
        A.nbv[k * nC + c] = nbv;
        A.nbj[k * nC + c] = nbj;
    }
}

// pass 4: transpose of the Green-Gauss face values; interior faces here
// (face value q_o (1 - gw) + (q_o + q_n) gw / 2, every variable), boundary
// faces in k_sadj_face2_bnd
__global__ void __launch_bounds__(kBlock) k_sadj_face2(CellsT<double> C, FacesT<double> F, AdjBufs A) {
    const int f = xcd_index();
    if (f >= F.nF || F.kind[f] != K_INTERNAL) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f], gw = F.gw[f];
    const double so = L / C.area[o], sn = L / C.area[n];
    for (int k = 0; k < NS; ++k) {
        const double v = (A.lgx[k * nT + o] * nx + A.lgy[k * nT + o] * ny) * so -
                         (A.lgx[k * nT + n] * nx + A.lgy[k * nT + n] * ny) * sn;
        A.co[k * nF + f] = v * (1.0 - 0.5 * gw);
        A.cn[k * nF + f] = v * (0.5 * gw);
    }
}

// pass 4, boundary faces: q_o (1 - gw) + face value (q_o; ghost) gw, the face
// value swept forward once per owner-state input, one thread per (face,
// input)
__global__ void __launch_bounds__(kBlock) k_sadj_face2_bnd(CellsT<double> C, FacesT<double> F,
                                                           const double* __restrict__ q, GasT<double> g, AdjBufs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.nbf * NS) return;
    const int j = i % NS, f = A.bface[i / NS];
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f], gw = F.gw[f];
    const double so = L / C.area[o];
    double lfv[NS];
    Dual1 a[NS], qn[NS], v[NS];
    for (int k = 0; k < NS; ++k) {
        lfv[k] = (A.lgx[k * nT + o] * nx + A.lgy[k * nT + o] * ny) * so;
        a[k] = Dual1(q[k * nT + o], j == k ? 1.0 : 0.0);
        qn[k] = Dual1(q[k * nT + n]);
    }
    bnd_face_value<Dual1>(sa_dual_gas(g), kind, Dual1(nx), Dual1(ny), a, qn, v);
    double s = 0.0;
    for (int k = 0; k < NS; ++k) s += lfv[k] * gw * v[k].d;
    A.co[j * nF + f] = lfv[j] * (1.0 - gw) + s;
    A.cn[j * nF + f] = 0.0;
}

// pass 5: gather of the Green-Gauss transpose and the neighbours' limiter parts
__global__ void __launch_bounds__(kBlock) k_sadj_cell2(CellsT<double> C, FacesT<double> F, AdjBufs A,
                                                       double* __restrict__ lam) {
    const int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC, nF = (size_t)F.nF;
    double acc[NS];
    for (int k = 0; k < NS; ++k) acc[k] = lam[k * nT + c];
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const bool nb = (e >> 31) != 0;
        const int other = nb ? F.c0[f] : F.c1[f];
        for (int k = 0; k < NS; ++k) {
            acc[k] += nb ? A.cn[k * nF + f] : A.co[k * nF + f];
            if (other < C.nC && A.nbj[k * nC + other] == c) acc[k] += A.nbv[k * nC + other];
        }
    }
    for (int k = 0; k < NS; ++k) lam[k * nT + c] = acc[k];
}

}  // namespace

hipError_t launch_sa_gradients(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                              double* gx, double* gy) {
    AdjBufs A{};
    A.gx = gx;
    A.gy = gy;
    hipLaunchKernelGGL(k_sadj_forward, dim3(xcd_grid(C.nC)), dim3(kBlock), 0, s, C, F, q, g, 0, 0.0, A);
    return hipGetLastError();
}

hipError_t launch_adjoint_state_sa(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q,
                                   const double* w, const double* wd, GasT<double> g, int second_order,
                                   double limiter_k, AdjBufs A, double* lam) {
    if (!A.vgx || !A.vgy || !A.vdq) return hipErrorInvalidValue;
    const int nbc = xcd_grid(C.nC), nbf = xcd_grid(F.nF);
    hipLaunchKernelGGL(k_sadj_forward, dim3(nbc), dim3(kBlock), 0, s, C, F, q, g, second_order, limiter_k, A);
    hipLaunchKernelGGL(k_sadj_face, dim3(nbf), dim3(kBlock), 0, s, C, F, q, w, g, second_order, A);
    if (A.nbf > 0)
        hipLaunchKernelGGL(k_sadj_face_bnd, dim3((A.nbf * (NS + 1) + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F,
                           q, w, g, second_order, A);
    hipLaunchKernelGGL(k_sadj_cell1, dim3(nbc), dim3(kBlock), 0, s, C, F, q, w, wd, g, second_order, limiter_k, A,
                       lam);
    // the gradients always enter (face gradients of the viscous terms, the source)
    hipLaunchKernelGGL(k_sadj_face2, dim3(nbf), dim3(kBlock), 0, s, C, F, A);
    if (A.nbf > 0)
        hipLaunchKernelGGL(k_sadj_face2_bnd, dim3((A.nbf * NS + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, q, g,
                           A);
    hipLaunchKernelGGL(k_sadj_cell2, dim3(nbc), dim3(kBlock), 0, s, C, F, A, lam);
    return hipGetLastError();
}

}  // namespace afx
