// Adjoint of the FP64 residual with respect to the state: lambda = (dR/dq)^T w
// for the inviscid residual of fillRhoRHS (solver.h:1079-1152) with
// Green-Gauss gradients (solver.h:428-469), the limiter of solver.h:517-593
// and MUSCL reconstruction, first or second order.
//
// Five atomics-free passes over the device layout (cells SoA q[k*nT+i]):
//   1 cells  gradients, limiters and the limiter's active branch (the face
//            that gives the minimum, the neighbour that gives q_max / q_min)
//   2 faces  reconstruction, adjoint of the Roe flux (hand-written reverse
//            sweep) and of the boundary states (forward-mode 4x4 Jacobian of
//            face_vars): adjoints of the reconstructed left/right states
//   3 cells  gather: direct state adjoint, gradient adjoint, limiter adjoint
//            (active branch; the part owed to the q_max/q_min neighbour is
//            kept for pass 5)
//   4 faces  transpose of the Green-Gauss face values
//   5 cells  gather of pass 4 and of the neighbours' limiter parts
// The wall/slip ghost states are copies of their owners that no flux reads;
// the farfield ghost states are boundary data and get no adjoint.
#include "afx_device.hpp"
#include "afx_adjoint.hpp"
#include "afx_adjoint_kernels.hpp"

namespace afx {

template <> __device__ __forceinline__ Dual1 ddiv<Dual1>(Dual1 a, Dual1 b) { return a / b; }
template <> __device__ __forceinline__ Dual1 dsqrt<Dual1>(Dual1 x) { return dual_sqrt(x); }

namespace {

__device__ __forceinline__ GasT<Dual1> dual_gas(const GasT<double>& g) {
    return GasT<Dual1>{Dual1(g.gamma), Dual1(g.R), Dual1(g.mu), Dual1(g.k), Dual1(g.cp), Dual1(g.PrL), Dual1(g.PrT)};
}

// b += (d face_vars(kind, qL, qbc) / d qL)^T a  (qbc is boundary data)
__device__ __forceinline__ void face_vars_vjp(const GasT<double>& g, int kind, double nx, double ny, const double* qL,
                                              const double* qbc, const double* a, double* b) {
    const GasT<Dual1> gd = dual_gas(g);
    Dual1 bc[4], out[4];
    for (int k = 0; k < 4; ++k) bc[k] = Dual1(qbc[k]);
    for (int j = 0; j < 4; ++j) {
        Dual1 in[4];
        for (int k = 0; k < 4; ++k) in[k] = Dual1(qL[k], k == j ? 1.0 : 0.0);
        face_vars<Dual1>(gd, kind, Dual1(nx), Dual1(ny), in, bc, out);
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += out[k].d * a[k];
        b[j] += s;
    }
}

// pass 1: gradients, limiters and the active branch of every (cell, variable)
__global__ void __launch_bounds__(kBlock) k_adj_forward(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                        GasT<double> g, int second_order, double limiter_k,
                                                        AdjBufs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC;
    double qc[4], sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) qc[k] = q[k * nT + c];
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const int o = F.c0[f], n = F.c1[f];
        double qo[4], qn[4], qv[4];
        for (int k = 0; k < 4; ++k) { qo[k] = q[k * nT + o]; qn[k] = q[k * nT + n]; }
        face_vars<double>(g, F.kind[f], F.nx[f], F.ny[f], qo, qn, qv);
        const double gw = F.gw[f], L = F.len[f];
        for (int k = 0; k < 4; ++k) {
            const double fk = (qo[k] * (1.0 - gw) + qv[k] * gw) * L;
            if (e >> 31) { sx[k] -= fk * F.nx[f]; sy[k] -= fk * F.ny[f]; }
            else { sx[k] += fk * F.nx[f]; sy[k] += fk * F.ny[f]; }
        }
    }
    const double Ar = C.area[c];
    for (int k = 0; k < 4; ++k) {
        A.gx[k * nT + c] = sx[k] / Ar;
        A.gy[k * nT + c] = sy[k] / Ar;
    }
    if (!second_order) return;
    double mn[4], mx[4];
    int jmn[4] = {-1, -1, -1, -1}, jmx[4] = {-1, -1, -1, -1};
    for (int k = 0; k < 4; ++k) { mn[k] = qc[k]; mx[k] = qc[k]; }
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const int kind = F.kind[f];
        if (kind == K_WALL || kind == K_SLIP) continue;   // ghost = owner copy
        const int other = (e >> 31) ? F.c0[f] : F.c1[f];
        for (int k = 0; k < 4; ++k) {
            const double v = q[k * nT + other];
            if (v < mn[k]) { mn[k] = v; jmn[k] = other; }
            if (mx[k] < v) { mx[k] = v; jmx[k] = other; }
        }
    }
    const double Ka = limiter_k * sqrt(Ar), K3a = Ka * Ka * Ka;
    double lm[4] = {1, 1, 1, 1};
    int act[4] = {-1, -1, -1, -1};
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const double dx = F.xf[f] - C.xc[c], dy = F.yf[f] - C.yc[c];
        for (int k = 0; k < 4; ++k) {
            const double dqg = sx[k] / Ar * dx + sy[k] / Ar * dy;
            const double l = limiter_phi<double>(dqg, mx[k] - qc[k], mn[k] - qc[k], K3a);
            if (l < lm[k]) { lm[k] = l; act[k] = s; }
        }
    }
    for (int k = 0; k < 4; ++k) {
        A.lim[k * nT + c] = lm[k];
        A.act[k * nC + c] = (int8_t)act[k];
        A.jmax[k * nC + c] = jmx[k];
        A.jmin[k * nC + c] = jmn[k];
    }
}

// pass 2: per face, adjoints of the reconstructed states
__global__ void __launch_bounds__(kBlock) k_adj_face(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                     const double* __restrict__ w, GasT<double> g, int second_order,
                                                     AdjBufs A) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f];
    double qL[4], qR[4];
    for (int k = 0; k < 4; ++k) { qL[k] = q[k * nT + o]; qR[k] = q[k * nT + n]; }
    if (second_order) {
        const double d0x = F.xf[f] - C.xc[o], d0y = F.yf[f] - C.yc[o];
        for (int k = 0; k < 4; ++k)
            qL[k] += (A.gx[k * nT + o] * d0x + A.gy[k * nT + o] * d0y) * A.lim[k * nT + o];
        if (kind == K_INTERNAL) {
            const double d1x = F.xf[f] - C.xc[n], d1y = F.yf[f] - C.yc[n];
            for (int k = 0; k < 4; ++k)
                qR[k] += (A.gx[k * nT + n] * d1x + A.gy[k * nT + n] * d1y) * A.lim[k * nT + n];
        }
    }
    double bF[4], bL[4], bR[4];
    for (int k = 0; k < 4; ++k)
        bF[k] = kind == K_INTERNAL ? L * (w[k * nT + n] - w[k * nT + o]) : -L * w[k * nT + o];
    if (kind == K_INTERNAL) {
        roe_flux_vjp<double>(g.gamma, nx, ny, qL, qR, bF, bL, bR);
    } else {
        double qb[4], bB[4];
        face_vars<double>(g, kind, nx, ny, qL, qR, qb);
        roe_flux_vjp<double>(g.gamma, nx, ny, qL, qb, bF, bL, bB);
        face_vars_vjp(g, kind, nx, ny, qL, qR, bB, bL);
        for (int k = 0; k < 4; ++k) bR[k] = 0.0;
    }
    for (int k = 0; k < 4; ++k) {
        A.bL[k * nF + f] = bL[k];
        A.bR[k * nF + f] = bR[k];
    }
}

// pass 3: per cell, state / gradient / limiter adjoints
__global__ void __launch_bounds__(kBlock) k_adj_cell1(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                      int second_order, double limiter_k, AdjBufs A,
                                                      double* __restrict__ lam) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC, nF = (size_t)F.nF;
    double lq[4] = {0, 0, 0, 0}, lgx[4] = {0, 0, 0, 0}, lgy[4] = {0, 0, 0, 0}, lphi[4] = {0, 0, 0, 0};
    double gx[4], gy[4], ph[4];
    for (int k = 0; k < 4; ++k) { gx[k] = A.gx[k * nT + c]; gy[k] = A.gy[k * nT + c]; ph[k] = A.lim[k * nT + c]; }
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const bool nb = (e >> 31) != 0;
        const double dx = F.xf[f] - C.xc[c], dy = F.yf[f] - C.yc[c];
        for (int k = 0; k < 4; ++k) {
            const double b = nb ? A.bR[k * nF + f] : A.bL[k * nF + f];
            lq[k] += b;
            if (second_order) {
                lphi[k] += b * (gx[k] * dx + gy[k] * dy);
                lgx[k] += b * ph[k] * dx;
                lgy[k] += b * ph[k] * dy;
            }
        }
    }
    for (int k = 0; k < 4; ++k) {
        double nbv = 0.0;
        int nbj = -1;
        const int s = second_order ? (int)A.act[k * nC + c] : -1;
        if (s >= 0) {
            const int f = (int)(C.cface[(size_t)s * nC + c] & 0x7FFFFFFFu);
            const double dx = F.xf[f] - C.xc[c], dy = F.yf[f] - C.yc[c];
            const double qc = q[k * nT + c];
            const double dqg = gx[k] * dx + gy[k] * dy;
            const int j = dqg > 1e-16 ? A.jmax[k * nC + c] : A.jmin[k * nC + c];
            const double D = (j >= 0 ? q[k * nT + j] : qc) - qc;
            const double Ka = limiter_k * sqrt(C.area[c]), K3a = Ka * Ka * Ka;
            double d_dqg, d_D;
            (void)limiter_phi_d<double>(dqg, D, K3a, &d_dqg, &d_D);
            lgx[k] += lphi[k] * d_dqg * dx;
            lgy[k] += lphi[k] * d_dqg * dy;
            if (j >= 0) {   // D = q_j - q_c (a limiter on q_c itself has D = 0 and no net dependence)
                lq[k] -= lphi[k] * d_D;
                if (j < C.nC) { nbv = lphi[k] * d_D; nbj = j; }
            }
        }
        lam[k * nT + c] = lq[k];
        A.lgx[k * nT + c] = lgx[k];
        A.lgy[k * nT + c] = lgy[k];
        A.nbv[k * nC + c] = nbv;
        A.nbj[k * nC + c] = nbj;
    }
}

// pass 4: transpose of the Green-Gauss face values
__global__ void __launch_bounds__(kBlock) k_adj_face2(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                      GasT<double> g, AdjBufs A) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f], gw = F.gw[f];
    const bool two = kind == K_INTERNAL;
    double lfv[4];
    for (int k = 0; k < 4; ++k) {
        double v = (A.lgx[k * nT + o] * nx + A.lgy[k * nT + o] * ny) * L / C.area[o];
        if (two) v -= (A.lgx[k * nT + n] * nx + A.lgy[k * nT + n] * ny) * L / C.area[n];
        lfv[k] = v;
    }
    double co[4], cn[4];
    if (two) {   // face value q_o (1 - gw) + (q_o + q_n) gw / 2
        for (int k = 0; k < 4; ++k) { co[k] = lfv[k] * (1.0 - 0.5 * gw); cn[k] = lfv[k] * (0.5 * gw); }
    } else {     // q_o (1 - gw) + face_vars(q_o, q_ghost) gw
        double a[4], qo[4], qn[4];
        for (int k = 0; k < 4; ++k) {
            a[k] = lfv[k] * gw;
            co[k] = lfv[k] * (1.0 - gw);
            cn[k] = 0.0;
            qo[k] = q[k * nT + o];
            qn[k] = q[k * nT + n];
        }
        face_vars_vjp(g, kind, nx, ny, qo, qn, a, co);
    }
    for (int k = 0; k < 4; ++k) {
        A.co[k * nF + f] = co[k];
        A.cn[k * nF + f] = cn[k];
    }
}

// pass 5: gather of the Green-Gauss transpose and the neighbours' limiter parts
__global__ void __launch_bounds__(kBlock) k_adj_cell2(CellsT<double> C, FacesT<double> F, AdjBufs A,
                                                      double* __restrict__ lam) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC, nF = (size_t)F.nF;
    double acc[4];
    for (int k = 0; k < 4; ++k) acc[k] = lam[k * nT + c];
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const bool nb = (e >> 31) != 0;
        const int other = nb ? F.c0[f] : F.c1[f];
        for (int k = 0; k < 4; ++k) {
            acc[k] += nb ? A.cn[k * nF + f] : A.co[k * nF + f];
            if (other < C.nC && A.nbj[k * nC + other] == c) acc[k] += A.nbv[k * nC + other];
        }
    }
    for (int k = 0; k < 4; ++k) lam[k * nT + c] = acc[k];
}

}  // namespace

hipError_t launch_adjoint_state(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* w,
                                GasT<double> g, int second_order, double limiter_k, AdjBufs A, double* lam) {
    const int nbc = (C.nC + kBlock - 1) / kBlock, nbf = (F.nF + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_adj_forward, dim3(nbc), dim3(kBlock), 0, s, C, F, q, g, second_order, limiter_k, A);
    hipLaunchKernelGGL(k_adj_face, dim3(nbf), dim3(kBlock), 0, s, C, F, q, w, g, second_order, A);
    hipLaunchKernelGGL(k_adj_cell1, dim3(nbc), dim3(kBlock), 0, s, C, F, q, second_order, limiter_k, A, lam);
    if (second_order) {
        hipLaunchKernelGGL(k_adj_face2, dim3(nbf), dim3(kBlock), 0, s, C, F, q, g, A);
        hipLaunchKernelGGL(k_adj_cell2, dim3(nbc), dim3(kBlock), 0, s, C, F, A, lam);
    }
    return hipGetLastError();
}

}  // namespace afx
