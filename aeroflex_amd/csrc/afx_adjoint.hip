// Adjoint of the FP64 residual with respect to the state: lambda = (dR/dq)^T w
// for the inviscid and laminar residual of fillRhoRHS (solver.h:1079-1152)
// with Green-Gauss gradients (solver.h:428-469), the limiter of
// solver.h:517-593 and MUSCL reconstruction, first or second order.  The
// laminar terms (physics.h:230-258 at the face gradient of
// average_gradients, solver.h:359-398) add their adjoints in passes 2 and 3.
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
//
// With geometry on, passes 2-4 also accumulate the adjoints of the face and
// cell metrics the residual reads (normal, length, Green-Gauss weight,
// reconstruction offsets x_f - x_c, cell area), and three more passes carry
// them back to the node coordinates through the metric definitions of
// mesh.h:378-453, 744-787 and solver.h:436-444:
//   6 faces  face ends and the two centroids (normal, length, midpoint,
//            weight |x_f - x_0| / |x_0 - x_1|, ghost centroid x_f + |x_f - x_0| n)
//   7 cells  centroid (mean of the nodes) and area (two triangles) to vertices
//   8 nodes  gather over each node's face ends and cell vertices (CSR)
#include "afx_device.hpp"
#include "afx_adjoint.hpp"
#include "afx_adjoint_kernels.hpp"

namespace afx {

template <> __device__ __forceinline__ Dual1 ddiv<Dual1>(Dual1 a, Dual1 b) { return a / b; }
template <> __device__ __forceinline__ Dual1 dsqrt<Dual1>(Dual1 x) { return dual_sqrt(x); }
template <> __device__ __forceinline__ Dual1 dcbrt<Dual1>(Dual1 x) { return dual_cbrt(x); }

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

// b += (d face_vars(kind, qL, qbc) / d qbc)^T a  (farfield boundary data)
__device__ __forceinline__ void face_vars_vjp_bc(const GasT<double>& g, int kind, double nx, double ny,
                                                 const double* qL, const double* qbc, const double* a, double* b) {
    const GasT<Dual1> gd = dual_gas(g);
    Dual1 in[4], out[4];
    for (int k = 0; k < 4; ++k) in[k] = Dual1(qL[k]);
    for (int j = 0; j < 4; ++j) {
        Dual1 bc[4];
        for (int k = 0; k < 4; ++k) bc[k] = Dual1(qbc[k], k == j ? 1.0 : 0.0);
        face_vars<Dual1>(gd, kind, Dual1(nx), Dual1(ny), in, bc, out);
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += out[k].d * a[k];
        b[j] += s;
    }
}

// d face_vars(kind, qL, qbc; n) / d n_dir, dir 0 = x, 1 = y
__device__ __forceinline__ void face_vars_dn(const GasT<double>& g, int kind, double nx, double ny, const double* qL,
                                             const double* qbc, int dir, double* out_d) {
    const GasT<Dual1> gd = dual_gas(g);
    Dual1 a[4], bc[4], o[4];
    for (int k = 0; k < 4; ++k) { a[k] = Dual1(qL[k]); bc[k] = Dual1(qbc[k]); }
    face_vars<Dual1>(gd, kind, Dual1(nx, dir == 0 ? 1.0 : 0.0), Dual1(ny, dir == 1 ? 1.0 : 0.0), a, bc, o);
    for (int k = 0; k < 4; ++k) out_d[k] = o[k].d;
}

// bF . d roe(qL, qR(n), n) / d n_dir with qR = face_vars(qL, qbc; n) on boundary faces
__device__ __forceinline__ double roe_dn(const GasT<double>& g, int kind, double nx, double ny, const double* qL,
                                         const double* qR, const double* bF, int dir) {
    const GasT<Dual1> gd = dual_gas(g);
    const Dual1 nxd(nx, dir == 0 ? 1.0 : 0.0), nyd(ny, dir == 1 ? 1.0 : 0.0);
    Dual1 a[4], b[4], f[4];
    for (int k = 0; k < 4; ++k) a[k] = Dual1(qL[k]);
    if (kind == K_INTERNAL) {
        for (int k = 0; k < 4; ++k) b[k] = Dual1(qR[k]);
    } else {
        Dual1 bc[4];
        for (int k = 0; k < 4; ++k) bc[k] = Dual1(qR[k]);
        face_vars<Dual1>(gd, kind, nxd, nyd, a, bc, b);
    }
    roe_flux<Dual1>(gd, nxd, nyd, a, b, false, nullptr, nullptr, f);
    double s = 0.0;
    for (int k = 0; k < 4; ++k) s += bF[k] * f[k].d;
    return s;
}

// VJP of the laminar terms (viscous_terms, physics.h:230-258) of a face with
// respect to its reconstructed states and its face gradients: bL, bR +=
// (d F_v / d q)^T bF and bgx, bgy = (d F_v / d grad)^T bF, by sixteen
// forward-mode sweeps of the same device code
__device__ void viscous_vjp(const GasT<double>& g, double nx, double ny, const double* qL, const double* qR,
                            const double* gxa, const double* gya, const double* bF, double* bL, double* bR,
                            double* bgx, double* bgy) {
    const GasT<Dual1> gd = dual_gas(g);
    for (int j = 0; j < 16; ++j) {
        Dual1 a[4], b[4], gx[4], gy[4];
        for (int k = 0; k < 4; ++k) {
            a[k] = Dual1(qL[k], j == k ? 1.0 : 0.0);
            b[k] = Dual1(qR[k], j == 4 + k ? 1.0 : 0.0);
            gx[k] = Dual1(gxa[k], j == 8 + k ? 1.0 : 0.0);
            gy[k] = Dual1(gya[k], j == 12 + k ? 1.0 : 0.0);
        }
        Dual1 f1(0.0), f2(0.0), f3(0.0);
        viscous_terms<Dual1>(gd, Dual1(nx), Dual1(ny), a, b, gx, gy, gd.mu, gd.k, f1, f2, f3);
        const double v = bF[1] * f1.d + bF[2] * f2.d + bF[3] * f3.d;
        if (j < 4) bL[j] += v;
        else if (j < 8) bR[j - 4] += v;
        else if (j < 12) bgx[j - 8] = v;
        else bgy[j - 12] = v;
    }
}

// pass 1: gradients, limiters and the active branch of every (cell, variable)
__global__ void __launch_bounds__(kBlock)
k_adj_forward(CellsT<double> C, FacesT<double> F, const double* qg, const double* q, GasT<double> g, int second_order,
              double limiter_k, AdjBufs A) {
    const int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC;
    // gradients and limiters of one state (qg == q): the neighbour extrema
    // come from the values the gradient loop already holds
    const bool same = qg == q;
    double qc[4], sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0};
    double mn[4], mx[4];
    int jmn[4] = {-1, -1, -1, -1}, jmx[4] = {-1, -1, -1, -1};
    for (int k = 0; k < 4; ++k) { qc[k] = q[k * nT + c]; mn[k] = qc[k]; mx[k] = qc[k]; }
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
        double qo[4], qn[4], qv[4];
        for (int k = 0; k < 4; ++k) { qo[k] = qg[k * nT + o]; qn[k] = qg[k * nT + n]; }
        face_vars<double>(g, kind, F.nx[f], F.ny[f], qo, qn, qv);
        const double gw = F.gw[f], L = F.len[f];
        for (int k = 0; k < 4; ++k) {
            const double fk = (qo[k] * (1.0 - gw) + qv[k] * gw) * L;
            if (e >> 31) { sx[k] -= fk * F.nx[f]; sy[k] -= fk * F.ny[f]; }
            else { sx[k] += fk * F.nx[f]; sy[k] += fk * F.ny[f]; }
        }
        if (same && second_order && kind != K_WALL && kind != K_SLIP) {   // wall ghost = owner copy
            const int other = (e >> 31) ? o : n;
            for (int k = 0; k < 4; ++k) {
                const double v = (e >> 31) ? qo[k] : qn[k];
                if (v < mn[k]) { mn[k] = v; jmn[k] = other; }
                if (mx[k] < v) { mx[k] = v; jmx[k] = other; }
            }
        }
    }
    const double Ar = C.area[c];
    for (int k = 0; k < 4; ++k) {
        A.gx[k * nT + c] = sx[k] / Ar;
        A.gy[k * nT + c] = sy[k] / Ar;
    }
    if (!second_order) return;
    if (!same)
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

// pass 2: per face, adjoints of the reconstructed states.  PART 1 takes the
// interior faces (boundary code compiled out, so the kernel needs far fewer
// registers), PART 2 the boundary faces from the list A.bface.  LAM compiles
// the laminar terms in (interior faces of the laminar residual, A.vgx set);
// without them the inviscid instances need far fewer registers.
template <bool GEO, bool GHOST, int PART, bool LAM>
__global__ void __launch_bounds__(kBlock) k_adj_face(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                     const double* __restrict__ w, GasT<double> g, int second_order,
                                                     AdjBufs A, GeoBufs G, double* __restrict__ ghost) {
    const int i = PART == 1 ? xcd_index() : (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int f;
    if constexpr (PART == 1) {
        f = i;
        if (f >= F.nF || F.kind[f] != K_INTERNAL) return;
    } else {
        if (i >= A.nbf) return;
        f = A.bface[i];
    }
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f];
    const int kind = PART == 1 ? (int)K_INTERNAL : (int)F.kind[f];
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
        if (LAM && A.vgx) {
            // laminar terms at the face gradient of average_gradients
            // (solver.h:359-398, A1): g = gbar - (gbar . t - dir) t with t the
            // normalised sum of the centroids and dir = (q_0 - q_1) / |x_0 + x_1|
            double t0 = C.xc[o] + C.xc[n], t1 = C.yc[o] + C.yc[n];
            const double l = sqrt(t0 * t0 + t1 * t1);
            t0 /= l;
            t1 /= l;
            double gxa[4], gya[4], bgx[4], bgy[4];
            for (int k = 0; k < 4; ++k) {
                const double dir = (q[k * nT + o] - q[k * nT + n]) / l;
                const double gxb = (A.gx[k * nT + o] + A.gx[k * nT + n]) * 0.5;
                const double gyb = (A.gy[k * nT + o] + A.gy[k * nT + n]) * 0.5;
                const double dot = gxb * t0 + gyb * t1;
                gxa[k] = gxb - (dot - dir) * t0;
                gya[k] = gyb - (dot - dir) * t1;
            }
            viscous_vjp(g, nx, ny, qL, qR, gxa, gya, bF, bL, bR, bgx, bgy);
            for (int k = 0; k < 4; ++k) {
                const double bt = bgx[k] * t0 + bgy[k] * t1;   // the adjoint of dir
                A.vgx[k * nF + f] = 0.5 * (bgx[k] - t0 * bt);
                A.vgy[k * nF + f] = 0.5 * (bgy[k] - t1 * bt);
                A.vdq[k * nF + f] = bt / l;
            }
        }
    } else {
        double qb[4], bB[4];
        face_vars<double>(g, kind, nx, ny, qL, qR, qb);
        roe_flux_vjp<double>(g.gamma, nx, ny, qL, qb, bF, bL, bB);
        face_vars_vjp(g, kind, nx, ny, qL, qR, bB, bL);
        for (int k = 0; k < 4; ++k) bR[k] = 0.0;
        if (GHOST && kind == K_FARFIELD) {   // the ghost's first adjoint write of this evaluation
            double gb[4] = {0.0, 0.0, 0.0, 0.0};
            face_vars_vjp_bc(g, kind, nx, ny, qL, qR, bB, gb);
            for (int k = 0; k < 4; ++k) ghost[k * nT + n] = gb[k];
        }
    }
    for (int k = 0; k < 4; ++k) {
        A.bL[k * nF + f] = bL[k];
        A.bR[k * nF + f] = bR[k];
    }
    if (!GEO) return;
    // metric adjoints of F(qL, qR; n) len: length, normal, reconstruction offsets
    double fl[4], qRuse[4];
    if (kind == K_INTERNAL) {
        for (int k = 0; k < 4; ++k) qRuse[k] = qR[k];
    } else {
        face_vars<double>(g, kind, nx, ny, qL, qR, qRuse);
    }
    roe_flux<double>(g, nx, ny, qL, qRuse, false, nullptr, nullptr, fl);
    double lb = 0.0;
    for (int k = 0; k < 4; ++k) lb += (kind == K_INTERNAL ? w[k * nT + n] - w[k * nT + o] : -w[k * nT + o]) * fl[k];
    const double nbx = roe_dn(g, kind, nx, ny, qL, qR, bF, 0), nby = roe_dn(g, kind, nx, ny, qL, qR, bF, 1);
    double d0b[2] = {0.0, 0.0}, d1b[2] = {0.0, 0.0};
    if (second_order) {
        for (int k = 0; k < 4; ++k) {
            const double ph = A.lim[k * nT + o] * bL[k];
            d0b[0] += ph * A.gx[k * nT + o];
            d0b[1] += ph * A.gy[k * nT + o];
        }
        if (kind == K_INTERNAL)
            for (int k = 0; k < 4; ++k) {
                const double ph = A.lim[k * nT + n] * bR[k];
                d1b[0] += ph * A.gx[k * nT + n];
                d1b[1] += ph * A.gy[k * nT + n];
            }
    }
    G.fbar[0 * nF + f] = nbx;
    G.fbar[1 * nF + f] = nby;
    G.fbar[2 * nF + f] = lb;
    G.fbar[3 * nF + f] = 0.0;
    G.fbar[4 * nF + f] = d0b[0];
    G.fbar[5 * nF + f] = d0b[1];
    G.fbar[6 * nF + f] = d1b[0];
    G.fbar[7 * nF + f] = d1b[1];
}

// pass 3: per cell, state / gradient / limiter adjoints
template <bool GEO, bool GHOST>
__global__ void __launch_bounds__(kBlock) k_adj_cell1(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                      int second_order, double limiter_k, AdjBufs A,
                                                      double* __restrict__ lam, GeoBufs G,
                                                      double* __restrict__ ghost) {
    const int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC, nF = (size_t)F.nF;
    double lq[4] = {0, 0, 0, 0}, lgx[4] = {0, 0, 0, 0}, lgy[4] = {0, 0, 0, 0}, lphi[4] = {0, 0, 0, 0};
    double gx[4], gy[4], ph[4];
    for (int k = 0; k < 4; ++k) { gx[k] = A.gx[k * nT + c]; gy[k] = A.gy[k * nT + c]; ph[k] = A.lim[k * nT + c]; }
    uint32_t es[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};   // the slots' words and offsets, kept
    double dxs[4] = {0, 0, 0, 0}, dys[4] = {0, 0, 0, 0};                   // for the limiter's active slot
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const bool nb = (e >> 31) != 0;
        const double dx = F.xf[f] - C.xc[c], dy = F.yf[f] - C.yc[c];
        es[s] = e; dxs[s] = dx; dys[s] = dy;
        for (int k = 0; k < 4; ++k) {
            const double b = nb ? A.bR[k * nF + f] : A.bL[k * nF + f];
            lq[k] += b;
            if (second_order) {
                lphi[k] += b * (gx[k] * dx + gy[k] * dy);
                lgx[k] += b * ph[k] * dx;
                lgy[k] += b * ph[k] * dy;
            }
        }
        if (A.vgx && F.kind[f] == K_INTERNAL)   // laminar terms: face gradient and difference quotient
            for (int k = 0; k < 4; ++k) {
                lgx[k] += A.vgx[k * nF + f];
                lgy[k] += A.vgy[k * nF + f];
                lq[k] += nb ? -A.vdq[k * nF + f] : A.vdq[k * nF + f];
            }
    }
    double abar = 0.0;
    for (int k = 0; k < 4; ++k) {
        double nbv = 0.0;
        int nbj = -1;
        const int s = second_order ? (int)A.act[k * nC + c] : -1;
        if (s >= 0) {
            const uint32_t e = s == 0 ? es[0] : s == 1 ? es[1] : s == 2 ? es[2] : es[3];
            const int f = (int)(e & 0x7FFFFFFFu);
            const double dx = s == 0 ? dxs[0] : s == 1 ? dxs[1] : s == 2 ? dxs[2] : dxs[3];
            const double dy = s == 0 ? dys[0] : s == 1 ? dys[1] : s == 2 ? dys[2] : dys[3];
            const double qc = q[k * nT + c];
            const double dqg = gx[k] * dx + gy[k] * dy;
            const int j = dqg > 1e-16 ? A.jmax[k * nC + c] : A.jmin[k * nC + c];
            const double D = (j >= 0 ? q[k * nT + j] : qc) - qc;
            const double Ka = limiter_k * sqrt(C.area[c]), K3a = Ka * Ka * Ka;
            double d_dqg, d_D, d_K = 0.0;
            (void)limiter_phi_d<double>(dqg, D, K3a, &d_dqg, &d_D, GEO ? &d_K : nullptr);
            lgx[k] += lphi[k] * d_dqg * dx;
            lgy[k] += lphi[k] * d_dqg * dy;
            if (GEO) {
                // K3a = (K sqrt(A))^3, and the offset of the limiting face
                abar += lphi[k] * d_K * 1.5 * K3a / C.area[c];
                const size_t side = (e >> 31) ? 6 : 4;
                G.fbar[side * nF + f] += lphi[k] * d_dqg * gx[k];
                G.fbar[(side + 1) * nF + f] += lphi[k] * d_dqg * gy[k];
            }
            if (j >= 0) {   // D = q_j - q_c (a limiter on q_c itself has D = 0 and no net dependence)
                lq[k] -= lphi[k] * d_D;
                if (j < C.nC) { nbv = lphi[k] * d_D; nbj = j; }
                else if (GHOST) ghost[k * nT + j] += lphi[k] * d_D;   // farfield ghost (its owner is this cell)
            }
        }
        lam[k * nT + c] = lq[k];
        A.lgx[k * nT + c] = lgx[k];
        A.lgy[k * nT + c] = lgy[k];
        A.nbv[k * nC + c] = nbv;
        A.nbj[k * nC + c] = nbj;
    }
    if (GEO) {
        // gradient = S / A
        if (second_order)
            for (int k = 0; k < 4; ++k) abar -= (lgx[k] * gx[k] + lgy[k] * gy[k]) / C.area[c];
        G.abar[c] = abar;
    }
}

// pass 4: transpose of the Green-Gauss face values.  PART 1 takes the
// interior faces (the boundary face values' reverse compiled out, so the
// kernel needs far fewer registers), PART 2 the boundary faces of A.bface.
template <bool GEO, bool GHOST, int PART>
__global__ void __launch_bounds__(kBlock) k_adj_face2(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                      GasT<double> g, AdjBufs A, GeoBufs G,
                                                      double* __restrict__ ghost) {
    int f;
    if constexpr (PART == 1) {
        f = xcd_index();
        if (f >= F.nF || F.kind[f] != K_INTERNAL) return;
    } else {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= A.nbf) return;
        f = A.bface[i];
    }
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f];
    const int kind = PART == 1 ? (int)K_INTERNAL : (int)F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f], gw = F.gw[f];
    constexpr bool two = PART == 1;
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
        if (GHOST && kind == K_FARFIELD) {
            double gb[4] = {0.0, 0.0, 0.0, 0.0};
            face_vars_vjp_bc(g, kind, nx, ny, qo, qn, a, gb);
            for (int k = 0; k < 4; ++k) ghost[k * nT + n] += gb[k];
        }
    }
    for (int k = 0; k < 4; ++k) {
        A.co[k * nF + f] = co[k];
        A.cn[k * nF + f] = cn[k];
    }
    if (!GEO) return;
    // face value fv (as in pass 1) and the metric adjoints of fv n len / A
    double qo[4], qn[4], qv[4], fv[4];
    for (int k = 0; k < 4; ++k) { qo[k] = q[k * nT + o]; qn[k] = q[k * nT + n]; }
    face_vars<double>(g, kind, nx, ny, qo, qn, qv);
    for (int k = 0; k < 4; ++k) fv[k] = qo[k] * (1.0 - gw) + qv[k] * gw;
    const double Ao = C.area[o], An = two ? C.area[n] : 1.0;
    double nbx = 0.0, nby = 0.0, lb = 0.0, gb = 0.0;
    for (int k = 0; k < 4; ++k) {
        double ax = A.lgx[k * nT + o] / Ao, ay = A.lgy[k * nT + o] / Ao;
        if (two) { ax -= A.lgx[k * nT + n] / An; ay -= A.lgy[k * nT + n] / An; }
        nbx += fv[k] * L * ax;
        nby += fv[k] * L * ay;
        lb += fv[k] * (ax * nx + ay * ny);
        gb += lfv[k] * (qv[k] - qo[k]);
    }
    if (!two) {
        double dvx[4], dvy[4];
        face_vars_dn(g, kind, nx, ny, qo, qn, 0, dvx);
        face_vars_dn(g, kind, nx, ny, qo, qn, 1, dvy);
        for (int k = 0; k < 4; ++k) {
            nbx += lfv[k] * gw * dvx[k];
            nby += lfv[k] * gw * dvy[k];
        }
    }
    G.fbar[0 * nF + f] += nbx;
    G.fbar[1 * nF + f] += nby;
    G.fbar[2 * nF + f] += lb;
    G.fbar[3 * nF + f] += gb;
}

__global__ void k_fill0(double* __restrict__ x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// pass 6: face metrics -> face ends and centroids (mesh.h:378-453, 744-787;
// solver.h:436-444).  n = s (-dy, dx) / L with the sign s fixed by the mesh,
// L = |b - a|, x_f = (a + b) / 2, gw = |x_f - x_o| / |x_o - x_n| with x_n the
// ghost centroid x_f + |x_f - x_o| n on boundary faces.
__global__ void __launch_bounds__(kBlock) k_geo_face(CellsT<double> C, FacesT<double> F, GeoBufs G) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nF = (size_t)F.nF;
    const int a = G.fnode[f], b = G.fnode[nF + f];
    const int o = F.c0[f], n = F.c1[f];
    const bool two = F.kind[f] == K_INTERNAL;
    const double xa = G.nodex[a], ya = G.nodey[a], xb = G.nodex[b], yb = G.nodey[b];
    const double dx = xb - xa, dy = yb - ya, L = sqrt(dx * dx + dy * dy);
    const double nx = F.nx[f], ny = F.ny[f];
    const double xf = 0.5 * (xa + xb), yf = 0.5 * (ya + yb);
    double nbx = G.fbar[0 * nF + f], nby = G.fbar[1 * nF + f], lb = G.fbar[2 * nF + f], gb = G.fbar[3 * nF + f];
    const double d0x = G.fbar[4 * nF + f], d0y = G.fbar[5 * nF + f], d1x = G.fbar[6 * nF + f], d1y = G.fbar[7 * nF + f];
    double xfb = d0x + d1x, yfb = d0y + d1y;
    double xob = -d0x, yob = -d0y, xnb = -d1x, ynb = -d1y;
    // Green-Gauss weight
    const double e1x = xf - C.xc[o], e1y = yf - C.yc[o], r1 = sqrt(e1x * e1x + e1y * e1y);
    const double e2x = C.xc[o] - C.xc[n], e2y = C.yc[o] - C.yc[n], r2 = sqrt(e2x * e2x + e2y * e2y);
    if (r1 > 0.0 && r2 > 0.0) {
        const double u1x = e1x / r1, u1y = e1y / r1, u2x = e2x / r2, u2y = e2y / r2;
        xfb += gb * u1x / r2;
        yfb += gb * u1y / r2;
        xob += gb * (-u1x / r2 - r1 / (r2 * r2) * u2x);
        yob += gb * (-u1y / r2 - r1 / (r2 * r2) * u2y);
        xnb += gb * (r1 / (r2 * r2) * u2x);
        ynb += gb * (r1 / (r2 * r2) * u2y);
        if (!two) {   // ghost centroid x_f + r1 n
            const double sg = nx * xnb + ny * ynb;
            xfb += xnb + sg * u1x;
            yfb += ynb + sg * u1y;
            xob -= sg * u1x;
            yob -= sg * u1y;
            nbx += r1 * xnb;
            nby += r1 * ynb;
            xnb = 0.0;
            ynb = 0.0;
        }
    }
    // normal and length -> edge vector (dx, dy)
    const double sgn = (nx * (-dy) + ny * dx) >= 0.0 ? 1.0 : -1.0;
    const double nn = nx * nbx + ny * nby;
    const double ubx = sgn * (nbx - nx * nn) / L, uby = sgn * (nby - ny * nn) / L;
    const double dxb = uby + lb * dx / L, dyb = -ubx + lb * dy / L;
    G.fend[0 * nF + f] = 0.5 * xfb - dxb;
    G.fend[1 * nF + f] = 0.5 * yfb - dyb;
    G.fend[2 * nF + f] = 0.5 * xfb + dxb;
    G.fend[3 * nF + f] = 0.5 * yfb + dyb;
    G.fcen[0 * nF + f] = xob;
    G.fcen[1 * nF + f] = yob;
    G.fcen[2 * nF + f] = xnb;
    G.fcen[3 * nF + f] = ynb;
}

// pass 7: centroid (node mean) and area (|triangle| sums) to cell vertices
__global__ void __launch_bounds__(kBlock) k_geo_cell(CellsT<double> C, FacesT<double> F, GeoBufs G) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nC = (size_t)C.nC, nF = (size_t)F.nF;
    double xcb = 0.0, ycb = 0.0;
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const size_t f = e & 0x7FFFFFFFu;
        const size_t side = (e >> 31) ? 2 : 0;
        xcb += G.fcen[side * nF + f];
        ycb += G.fcen[(side + 1) * nF + f];
    }
    int v[4];
    for (int s = 0; s < 4; ++s) v[s] = G.cnode[(size_t)s * nC + c];
    const int nv = v[3] < 0 ? 3 : 4;
    double bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
    for (int s = 0; s < nv; ++s) { bx[s] = xcb / nv; by[s] = ycb / nv; }
    const double Ab = G.abar[c];
    auto tri = [&](int i1, int i2, int i3) {
        const double x1 = G.nodex[v[i1]], y1 = G.nodey[v[i1]], x2 = G.nodex[v[i2]], y2 = G.nodey[v[i2]];
        const double x3 = G.nodex[v[i3]], y3 = G.nodey[v[i3]];
        const double S = x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2);
        const double h = 0.5 * Ab * (S >= 0.0 ? 1.0 : -1.0);
        bx[i1] += h * (y2 - y3); by[i1] += h * (x3 - x2);
        bx[i2] += h * (y3 - y1); by[i2] += h * (x1 - x3);
        bx[i3] += h * (y1 - y2); by[i3] += h * (x2 - x1);
    };
    tri(0, 1, 2);
    if (nv == 4) tri(0, 2, 3);
    for (int s = 0; s < 4; ++s) {
        G.cvert[(size_t)(2 * s) * nC + c] = bx[s];
        G.cvert[(size_t)(2 * s + 1) * nC + c] = by[s];
    }
}

// pass 8: per node, the sum over its face ends and cell vertices
__global__ void __launch_bounds__(kBlock) k_geo_node(CellsT<double> C, FacesT<double> F, GeoBufs G) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= G.nN) return;
    const size_t nF = (size_t)F.nF, nC = (size_t)C.nC;
    double x = 0.0, y = 0.0;
    for (int i = G.n_off[v]; i < G.n_off[v + 1]; ++i) {
        const int e = G.n_ent[i];
        if (e >= 0) {
            const size_t f = (size_t)(e >> 1), end = (size_t)(e & 1);
            x += G.fend[(2 * end) * nF + f];
            y += G.fend[(2 * end + 1) * nF + f];
        } else {
            const int idx = -(e + 1);
            const size_t c = (size_t)(idx >> 2), slot = (size_t)(idx & 3);
            x += G.cvert[(2 * slot) * nC + c];
            y += G.cvert[(2 * slot + 1) * nC + c];
        }
    }
    G.xbar[2 * (size_t)v] = x;
    G.xbar[2 * (size_t)v + 1] = y;
}

// pass 5: gather of the Green-Gauss transpose and the neighbours' limiter parts
// lam gets the state part (pass 3 + neighbours' limiter parts), lam_g the
// Green-Gauss part (the gradient's source state); with lam_g == lam both add
__global__ void __launch_bounds__(kBlock) k_adj_cell2(CellsT<double> C, FacesT<double> F, AdjBufs A,
                                                      double* __restrict__ lam, double* lam_g) {
    const int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC, nF = (size_t)F.nF;
    double acc[4], accg[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < 4; ++k) acc[k] = lam[k * nT + c];
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const bool nb = (e >> 31) != 0;
        const int other = nb ? F.c0[f] : F.c1[f];
        for (int k = 0; k < 4; ++k) {
            accg[k] += nb ? A.cn[k * nF + f] : A.co[k * nF + f];
            if (other < C.nC && A.nbj[k * nC + other] == c) acc[k] += A.nbv[k * nC + other];
        }
    }
    if (lam_g == lam) {
        for (int k = 0; k < 4; ++k) lam[k * nT + c] = acc[k] + accg[k];
    } else {
        for (int k = 0; k < 4; ++k) {
            lam[k * nT + c] = acc[k];
            lam_g[k * nT + c] = accg[k];
        }
    }
}

}  // namespace

namespace {

// spectral radius sqrt(gamma p / rho) + |V| of one state on a face and its
// gradient with respect to the state (k_dt, solver.h:308-356)
__device__ inline double eig_and_grad(const GasT<double>& g, double nx, double ny, const double* qc, double* d) {
    const double gm1 = g.gamma - 1.0;
    const double r = qc[0], ru = qc[1], rv = qc[2], re = qc[3];
    const double V = (ru * nx + rv * ny) / r;
    const double p = gm1 * (re - 0.5 / r * (ru * ru + rv * rv));
    const double a = sqrt(p * g.gamma / r);
    const double dp[4] = {gm1 * 0.5 * (ru * ru + rv * rv) / (r * r), -gm1 * ru / r, -gm1 * rv / r, gm1};
    const double sv = V < 0.0 ? -1.0 : 1.0;
    const double dV[4] = {-V / r, nx / r, ny / r, 0.0};
    for (int k = 0; k < 4; ++k) {
        const double da = g.gamma / (2.0 * a) * (dp[k] / r - (k == 0 ? p / (r * r) : 0.0));
        d[k] = da + sv * dV[k];
    }
    return a + (V < 0.0 ? -V : V);
}

// ebar_f = len_f (sbar_o + sbar_n), sbar_c = d dt_c / d s_c dtbar_c = -dtbar_c dt_c^2 / (cfl A_c)
__global__ void __launch_bounds__(kBlock) k_dt_bar_face(CellsT<double> C, FacesT<double> F,
                                                        const double* __restrict__ dt, const double* __restrict__ dtbar,
                                                        double cfl, double* __restrict__ ebar) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const int o = F.c0[f], n = F.c1[f];
    double e = -dtbar[o] * dt[o] * dt[o] / (cfl * C.area[o]);
    if (F.kind[f] == K_INTERNAL) e += -dtbar[n] * dt[n] * dt[n] / (cfl * C.area[n]);
    ebar[f] = e * F.len[f];
}

// per cell: the faces on which this cell's state gives max(eig_o, eig_n)
__global__ void __launch_bounds__(kBlock) k_dt_bar_cell(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                        GasT<double> g, const double* __restrict__ ebar,
                                                        double* __restrict__ qbar) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT, nC = (size_t)C.nC;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const int o = F.c0[f], n = F.c1[f];
        const double nx = F.nx[f], ny = F.ny[f];
        double qo[4], d[4];
        for (int k = 0; k < 4; ++k) qo[k] = q[k * nT + o];
        const double eo = eig_and_grad(g, nx, ny, qo, d);
        int active = o;
        if (F.kind[f] == K_INTERNAL) {
            double qn[4], dn[4];
            for (int k = 0; k < 4; ++k) qn[k] = q[k * nT + n];
            const double en = eig_and_grad(g, nx, ny, qn, dn);
            if (eo < en) { active = n; for (int k = 0; k < 4; ++k) d[k] = dn[k]; }
        }
        if (active != c) continue;
        for (int k = 0; k < 4; ++k) acc[k] += ebar[f] * d[k];
    }
    for (int k = 0; k < 4; ++k) qbar[k * nT + c] += acc[k];
}

// reverse of qk_s+1 = q + W_s dt coef for the real cells: qbar += g,
// wseed = g dt coef / A (the seed of the undivided residual), dtbar += g . W coef
__global__ void __launch_bounds__(kBlock) k_rev_stage(const double* __restrict__ gk, const double* __restrict__ W,
                                                      const double* __restrict__ dt, const double* __restrict__ area,
                                                      double coef, int32_t nC, int32_t nT, double* __restrict__ wseed,
                                                      double* __restrict__ dtbar, double* __restrict__ qbar) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    double sdt = 0.0;
    for (int k = 0; k < 4; ++k) {
        const size_t i = (size_t)k * nT + c;
        qbar[i] += gk[i];
        wseed[i] = gk[i] * dt[c] * coef / area[c];
        sdt += gk[i] * W[i] * coef;
    }
    dtbar[c] += sdt;
}

// y (+)= a x over the real-cell rows of 4 planes
__global__ void __launch_bounds__(kBlock) k_axpy4(const double* __restrict__ x, double a, int32_t nC, int32_t nT,
                                                  int overwrite, double* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    for (int k = 0; k < 4; ++k) {
        const size_t i = (size_t)k * nT + c;
        y[i] = overwrite ? a * x[i] : y[i] + a * x[i];
    }
}

// Facc += G.fbar (8 planes over faces), Aacc += G.abar (real cells)
__global__ void __launch_bounds__(kBlock) k_acc_metrics(const double* __restrict__ fbar, const double* __restrict__ abar,
                                                        int32_t nF, int32_t nC, double* __restrict__ facc,
                                                        double* __restrict__ aacc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nF)
        for (int k = 0; k < 8; ++k) facc[(size_t)k * nF + i] += fbar[(size_t)k * nF + i];
    if (i < nC) aacc[i] += abar[i];
}

// metric part of the local time step's VJP: dt_c = cfl A_c / S_c with
// S_c = sum_f lambda_f len_f and lambda_f = max over the face's sides of
// c + |u n_x + v n_y| (the owner only on boundary faces).  With
// sbar_c = -dtbar_c dt_c^2 / (cfl A_c): lenbar_f += lambda_f (sbar_o + sbar_n),
// (nbar_x, nbar_y)_f += len_f (sbar_o + sbar_n) sign(V) (u, v) of the active
// side.  (The area term dtbar dt / A cancels the residual's division by the
// area in the stage update, so neither is added.)
__global__ void __launch_bounds__(kBlock) k_dt_bar_geo(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                       GasT<double> g, const double* __restrict__ dt,
                                                       const double* __restrict__ dtbar, double cfl,
                                                       double* __restrict__ facc) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f];
    const double nx = F.nx[f], ny = F.ny[f];
    double sb = -dtbar[o] * dt[o] * dt[o] / (cfl * C.area[o]);
    double qa[4], d[4];
    for (int k = 0; k < 4; ++k) qa[k] = q[k * nT + o];
    double lam = eig_and_grad(g, nx, ny, qa, d);
    if (F.kind[f] == K_INTERNAL) {
        sb += -dtbar[n] * dt[n] * dt[n] / (cfl * C.area[n]);
        double qn[4];
        for (int k = 0; k < 4; ++k) qn[k] = q[k * nT + n];
        const double en = eig_and_grad(g, nx, ny, qn, d);
        if (lam < en) { lam = en; for (int k = 0; k < 4; ++k) qa[k] = qn[k]; }
    }
    const double u = qa[1] / qa[0], v = qa[2] / qa[0];
    const double sv = (u * nx + v * ny) < 0.0 ? -1.0 : 1.0;
    const double e = F.len[f] * sb;
    facc[0 * nF + f] += e * sv * u;
    facc[1 * nF + f] += e * sv * v;
    facc[2 * nF + f] += lam * sb;
}

// acc[k][j] += ghost[k*nT + cell_j] over the farfield ghost cells
__global__ void __launch_bounds__(kBlock) k_ghost_acc(const double* __restrict__ ghost, const int32_t* __restrict__ cells,
                                                      int32_t n, int32_t nT, double* __restrict__ acc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int k = 0; k < 4; ++k) acc[(size_t)k * n + j] += ghost[(size_t)k * nT + cells[j]];
}

}  // namespace

hipError_t launch_acc_metrics(hipStream_t s, const GeoBufs& G, int32_t nF, int32_t nC, double* facc, double* aacc) {
    const int n = nF > nC ? nF : nC;
    hipLaunchKernelGGL(k_acc_metrics, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, G.fbar, G.abar, nF, nC,
                       facc, aacc);
    return hipGetLastError();
}

hipError_t launch_dt_vjp_geo(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                             const double* dt, const double* dtbar, double cfl, double* facc) {
    hipLaunchKernelGGL(k_dt_bar_geo, dim3((F.nF + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, q, g, dt, dtbar,
                       cfl, facc);
    return hipGetLastError();
}

hipError_t launch_geo_to_nodes(hipStream_t s, CellsT<double> C, FacesT<double> F, const GeoBufs& G) {
    hipLaunchKernelGGL(k_geo_face, dim3((F.nF + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, G);
    hipLaunchKernelGGL(k_geo_cell, dim3((C.nC + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, G);
    hipLaunchKernelGGL(k_geo_node, dim3((G.nN + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, G);
    return hipGetLastError();
}

hipError_t launch_rev_stage(hipStream_t s, const double* gk, const double* W, const double* dt, const double* area,
                            double coef, int32_t nC, int32_t nT, double* wseed, double* dtbar, double* qbar) {
    hipLaunchKernelGGL(k_rev_stage, dim3((nC + kBlock - 1) / kBlock), dim3(kBlock), 0, s, gk, W, dt, area, coef, nC,
                       nT, wseed, dtbar, qbar);
    return hipGetLastError();
}

hipError_t launch_axpy4(hipStream_t s, const double* x, double a, int32_t nC, int32_t nT, int overwrite, double* y) {
    hipLaunchKernelGGL(k_axpy4, dim3((nC + kBlock - 1) / kBlock), dim3(kBlock), 0, s, x, a, nC, nT, overwrite, y);
    return hipGetLastError();
}

hipError_t launch_ghost_acc(hipStream_t s, const double* ghost, const int32_t* cells, int32_t n, int32_t nT,
                            double* acc) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ghost_acc, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ghost, cells, n, nT, acc);
    return hipGetLastError();
}

hipError_t launch_dt_vjp(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                         const double* dt, const double* dtbar, double cfl, double* ebar, double* qbar) {
    hipLaunchKernelGGL(k_dt_bar_face, dim3((F.nF + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, dt, dtbar, cfl,
                       ebar);
    hipLaunchKernelGGL(k_dt_bar_cell, dim3((C.nC + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, q, g, ebar, qbar);
    return hipGetLastError();
}

namespace {
template <bool GEO, bool GHOST>
void adjoint_passes(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* qg, const double* q,
                    const double* w, GasT<double> g, int second_order, double limiter_k, AdjBufs A, double* lam,
                    double* lam_g, double* ghost, const GeoBufs& G, bool to_nodes = true) {
    // the pass kernels index cells and faces in XCD-aware block order
    const int nbc = xcd_grid(C.nC), nbf = xcd_grid(F.nF);
    hipLaunchKernelGGL(k_adj_forward, dim3(nbc), dim3(kBlock), 0, s, C, F, qg, q, g, second_order, limiter_k, A);
    if (A.vgx)
        hipLaunchKernelGGL((k_adj_face<GEO, GHOST, 1, true>), dim3(nbf), dim3(kBlock), 0, s, C, F, q, w, g,
                           second_order, A, G, ghost);
    else
        hipLaunchKernelGGL((k_adj_face<GEO, GHOST, 1, false>), dim3(nbf), dim3(kBlock), 0, s, C, F, q, w, g,
                           second_order, A, G, ghost);
    if (A.nbf > 0)
        hipLaunchKernelGGL((k_adj_face<GEO, GHOST, 2, false>), dim3((A.nbf + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F,
                           q, w, g, second_order, A, G, ghost);
    hipLaunchKernelGGL((k_adj_cell1<GEO, GHOST>), dim3(nbc), dim3(kBlock), 0, s, C, F, q, second_order, limiter_k, A,
                       lam, G, ghost);
    if (second_order || A.vgx) {   // gradients enter the residual (reconstruction or laminar terms)
        hipLaunchKernelGGL((k_adj_face2<GEO, GHOST, 1>), dim3(nbf), dim3(kBlock), 0, s, C, F, qg, g, A, G, ghost);
        if (A.nbf > 0)
            hipLaunchKernelGGL((k_adj_face2<GEO, GHOST, 2>), dim3((A.nbf + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               C, F, qg, g, A, G, ghost);
        hipLaunchKernelGGL(k_adj_cell2, dim3(nbc), dim3(kBlock), 0, s, C, F, A, lam, lam_g);
    } else if (lam_g != lam) {
        hipLaunchKernelGGL(k_fill0, dim3((4 * C.nT + kBlock - 1) / kBlock), dim3(kBlock), 0, s, lam_g, 4 * C.nT);
    }
    if (GEO && to_nodes) {
        hipLaunchKernelGGL(k_geo_face, dim3(nbf), dim3(kBlock), 0, s, C, F, G);
        hipLaunchKernelGGL(k_geo_cell, dim3(nbc), dim3(kBlock), 0, s, C, F, G);
        hipLaunchKernelGGL(k_geo_node, dim3((G.nN + kBlock - 1) / kBlock), dim3(kBlock), 0, s, C, F, G);
    }
}
}  // namespace

hipError_t launch_adjoint_stage(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* qg, const double* q,
                                const double* w, GasT<double> g, int second_order, double limiter_k, AdjBufs A,
                                double* lam, double* lam_g, double* ghost, const GeoBufs* Gp) {
    const GeoBufs G = Gp ? *Gp : GeoBufs{};
    if (Gp && ghost) return hipErrorInvalidValue;   // no caller needs both
    if (Gp) adjoint_passes<true, false>(s, C, F, qg, q, w, g, second_order, limiter_k, A, lam, lam_g, nullptr, G);
    else if (ghost) adjoint_passes<false, true>(s, C, F, qg, q, w, g, second_order, limiter_k, A, lam, lam_g, ghost, G);
    else adjoint_passes<false, false>(s, C, F, qg, q, w, g, second_order, limiter_k, A, lam, lam_g, nullptr, G);
    return hipGetLastError();
}

hipError_t launch_adjoint_stage_metrics(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* qg,
                                        const double* q, const double* w, GasT<double> g, int second_order,
                                        double limiter_k, AdjBufs A, double* lam, double* lam_g, const GeoBufs& G) {
    adjoint_passes<true, false>(s, C, F, qg, q, w, g, second_order, limiter_k, A, lam, lam_g, nullptr, G, false);
    return hipGetLastError();
}

hipError_t launch_adjoint_state(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* w,
                                GasT<double> g, int second_order, double limiter_k, AdjBufs A, double* lam,
                                const GeoBufs* Gp) {
    return launch_adjoint_stage(s, C, F, q, q, w, g, second_order, limiter_k, A, lam, lam, nullptr, Gp);
}

// ------------------------------------------------------------------ tangent --
// Tangent of the FP64 residual: the fused tile kernel k_residual_tile
// instantiated with Dual1, so R and (dR/dq) v come out of the same operations
// as the plain residual (gradients, limiter on its active branch, Roe, laminar
// and Spalart-Allmaras terms, cell sums).  The value parts reproduce the FP64
// residual bit for bit; the derivative parts are the exact directional
// derivative of the piecewise-smooth residual.
namespace {
__global__ void __launch_bounds__(kBlock) k_to_dual(const double* __restrict__ x, int64_t n, Dual1* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = Dual1(x[i]);
}
// q + eps v over nv planes of nT cells; ghost cells (i >= nC) keep a zero
// derivative: their states are boundary data
__global__ void __launch_bounds__(kBlock) k_dual_state(const double* __restrict__ q, const double* __restrict__ v,
                                                       int32_t nT, int32_t nC, int nv, Dual1* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nv * nT) return;
    const int32_t c = (int32_t)(i % nT);
    out[i] = Dual1(q[i], c < nC ? v[i] : 0.0);
}
__global__ void __launch_bounds__(kBlock) k_dual_parts(const Dual1* __restrict__ x, int64_t n, double* __restrict__ val,
                                                       double* __restrict__ der) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (val) val[i] = x[i].v;
    der[i] = x[i].d;
}
inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
}  // namespace

hipError_t launch_to_dual(hipStream_t s, const double* x, int64_t n, Dual1* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_dual, dim3(nblk(n)), dim3(kBlock), 0, s, x, n, out);
    return hipGetLastError();
}
hipError_t launch_dual_state(hipStream_t s, const double* q, const double* v, int32_t nT, int32_t nC, int nv,
                             Dual1* out) {
    hipLaunchKernelGGL(k_dual_state, dim3(nblk((int64_t)nv * nT)), dim3(kBlock), 0, s, q, v, nT, nC, nv, out);
    return hipGetLastError();
}
hipError_t launch_dual_parts(hipStream_t s, const Dual1* x, int64_t n, double* val, double* der) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dual_parts, dim3(nblk(n)), dim3(kBlock), 0, s, x, n, val, der);
    return hipGetLastError();
}
size_t tile_lds_bytes_tangent(int32_t lmax, int32_t gmax, int32_t ftmax, int nv) {
    return tile_lds_bytes<Dual1>(lmax, gmax, ftmax, nv);
}
hipError_t launch_residual_tile_tangent(hipStream_t s, TilesT<Dual1> S, const Dual1* q, int32_t nT, int32_t nC,
                                        const Dual1* area, const Dual1* lsq, const Dual1* xc, const Dual1* yc,
                                        const Dual1* wd, GasT<double> g64, int gg, int so, double limiter_k,
                                        int viscous, Dual1* R) {
    const int nb = ((S.ntiles + 7) / 8) * 8;
    const GasT<Dual1> g{Dual1(g64.gamma), Dual1(g64.R), Dual1(g64.mu), Dual1(g64.k), Dual1(g64.cp), Dual1(g64.PrL),
                        Dual1(g64.PrT)};
    const int nv = viscous == 2 ? 5 : 4;
    const size_t lds = tile_lds_bytes<Dual1>(S.lmax, S.gmax, S.ftmax, nv);
    hipError_t e;
    // the state array is also the ghost array: the kernel refreshes the
    // wall/slip ghosts of its tiles from their owners (set_walls_from_internal)
    Dual1* qg = const_cast<Dual1*>(q);
    if (nv == 5) {
        if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_residual_tile<Dual1, 5>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess)
            return e;
        hipLaunchKernelGGL((k_residual_tile<Dual1, 5>), dim3(nb), dim3(kBlock), lds, s, S, q, nT, nC, area, lsq, xc,
                           yc, wd, g, gg, so, Dual1(limiter_k), viscous, R, qg);
    } else {
        if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_residual_tile<Dual1, 4>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess)
            return e;
        hipLaunchKernelGGL((k_residual_tile<Dual1, 4>), dim3(nb), dim3(kBlock), lds, s, S, q, nT, nC, area, lsq, xc,
                           yc, wd, g, gg, so, Dual1(limiter_k), viscous, R, qg);
    }
    return hipGetLastError();
}

}  // namespace afx
