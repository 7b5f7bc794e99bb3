// Device physics and residual kernels, templated on the arithmetic type.
//
// The FP64 instantiation is compiled with -ffp-contract=off and keeps the
// reference's operation order term by term, so that with the per-cell face
// lists sorted by reference face index the plain form reproduces the serial
// CPU residual of src/rans (solver.h:1079-1152) to the last bit.  The FP32
// instantiation (contraction allowed) is the throughput form.
#pragma once
#include "afx_kernels.hpp"

namespace afx {

template <typename T> __device__ __forceinline__ T dsqrt(T x);
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }

template <typename T> __device__ __forceinline__ T dabs(T x) { return x < T(0) ? -x : x; }

// physics.h:84-86
template <typename T> __device__ __forceinline__ T sabs(T x) { return dsqrt<T>(x * x + T(1e-4)); }
// physics.h:133-135
template <typename T> __device__ __forceinline__ T entropy_fix(T l, T d) { return l > d ? l : (l * l + d * d) / (T(2) * d); }

// physics.h:160-228 (Roe + Harten) and 230-258 (laminar), internal_flux::operator()
template <typename T>
__device__ __forceinline__ void roe_flux(const GasT<T>& g, T nx, T ny, const T* q_L, const T* q_R, bool laminar,
                                         const T* gx, const T* gy, T* f) {
    const T rho_L = q_L[0], rho_u_L = q_L[1], rho_v_L = q_L[2], rho_e_L = q_L[3];
    const T rho_R = q_R[0], rho_u_R = q_R[1], rho_v_R = q_R[2], rho_e_R = q_R[3];
    const T gm1 = g.gamma - T(1);
    const T V_L = (rho_u_L * nx + rho_v_L * ny) / rho_L;
    const T V_R = (rho_u_R * nx + rho_v_R * ny) / rho_R;
    const T p_L = gm1 * (rho_e_L - T(0.5) / rho_L * (rho_u_L * rho_u_L + rho_v_L * rho_v_L));
    const T p_R = gm1 * (rho_e_R - T(0.5) / rho_R * (rho_u_R * rho_u_R + rho_v_R * rho_v_R));
    T f0 = (V_L * rho_L + V_R * rho_R) * T(0.5);
    T f1 = (V_L * rho_u_L + p_L * nx + V_R * rho_u_R + p_R * nx) * T(0.5);
    T f2 = (V_L * rho_v_L + p_L * ny + V_R * rho_v_R + p_R * ny) * T(0.5);
    T f3 = (V_L * (rho_e_L + p_L) + V_R * (rho_e_R + p_R)) * T(0.5);
    const T pL = p_L, pR = p_R;
    const T uL = rho_u_L / rho_L;
    const T uR = rho_u_R / rho_R;
    const T vL = rho_v_L / rho_L;
    const T vR = rho_v_R / rho_R;
    const T srhoL = dsqrt<T>(rho_L);
    const T srhoR = dsqrt<T>(rho_R);
    const T rho = srhoR * srhoL;
    const T u = (uL * srhoL + uR * srhoR) / (srhoL + srhoR);
    const T v = (vL * srhoL + vR * srhoR) / (srhoL + srhoR);
    const T h = ((rho_e_L + pL) / rho_L * srhoL + (rho_e_R + pR) / rho_R * srhoR) / (srhoL + srhoR);
    const T q2 = u * u + v * v;
    const T c = dsqrt<T>(gm1 * (h - T(0.5) * q2));
    const T V = u * nx + v * ny;
    const T VR = uR * nx + vR * ny;
    const T VL = uL * nx + vL * ny;
    const T d = T(0.05) * c;
    const T lambda_cm = entropy_fix<T>(sabs<T>(V - c), d);
    const T lambda_c = entropy_fix<T>(sabs<T>(V), d);
    const T lambda_cp = entropy_fix<T>(sabs<T>(V + c), d);
    const T dp = pR - pL, dV = VR - VL;
    const T kF1 = lambda_cm * (dp - rho * c * dV) / (T(2) * c * c);
    const T kF234_0 = lambda_c * ((rho_R - rho_L) - dp / (c * c));
    const T kF234_1 = lambda_c * rho;
    const T kF5 = lambda_cp * (dp + rho * c * dV) / (T(2) * c * c);
    f0 -= T(0.5) * (kF1 + kF234_0 + kF5);
    f1 -= T(0.5) * (kF1 * (u - c * nx) + kF234_0 * u + kF234_1 * (uR - uL - dV * nx) + kF5 * (u + c * nx));
    f2 -= T(0.5) * (kF1 * (v - c * ny) + kF234_0 * v + kF234_1 * (vR - vL - dV * ny) + kF5 * (v + c * ny));
    f3 -= T(0.5) * (kF1 * (h - c * V) + kF234_0 * q2 * T(0.5) + kF234_1 * (u * (uR - uL) + v * (vR - vL) - V * dV) + kF5 * (h + c * V));
    if (laminar) {
        T qc[4];
        for (int k = 0; k < 4; ++k) qc[k] = T(0.5) * (q_L[k] + q_R[k]);
        const T r2 = qc[0] * qc[0];
        // physics.h:33-68 temperature gradient through the pressure gradient
        const T p = gm1 * (qc[3] - T(0.5) / qc[0] * (qc[1] * qc[1] + qc[2] * qc[2]));
        T gp0 = T(2.) * gx[3];
        gp0 -= gx[1] * qc[1] / qc[0] + qc[1] * (gx[1] * qc[0] - gx[0] * qc[1]) / r2;
        gp0 -= gx[2] * qc[2] / qc[0] + qc[2] * (gx[2] * qc[0] - gx[0] * qc[2]) / r2;
        gp0 *= T(0.5) * gm1;
        T gp1 = T(2.) * gy[3];
        gp1 -= gy[1] * qc[1] / qc[0] + qc[1] * (gy[1] * qc[0] - gy[0] * qc[1]) / r2;
        gp1 -= gy[2] * qc[2] / qc[0] + qc[2] * (gy[2] * qc[0] - gy[0] * qc[2]) / r2;
        gp1 *= T(0.5) * gm1;
        const T gT0 = (T(1.) / g.R) * ((gp0 * qc[0] - gx[0] * p) / r2);
        const T gT1 = (T(1.) / g.R) * ((gp1 * qc[0] - gy[0] * p) / r2);
        const T gu0 = (qc[0] * gx[1] - qc[1] * gx[0]) / r2;
        const T gu1 = (qc[0] * gy[1] - qc[1] * gy[0]) / r2;
        const T gv0 = (qc[0] * gx[2] - qc[2] * gx[0]) / r2;
        const T gv1 = (qc[0] * gy[2] - qc[2] * gy[0]) / r2;
        const T div_v = gu0 + gv1;
        const T tau_xx = T(2.) * g.mu * (gu0 - div_v / T(3.));
        const T tau_yy = T(2.) * g.mu * (gv1 - div_v / T(3.));
        const T tau_xy = g.mu * (gu1 + gv0);
        const T phi0 = qc[1] / qc[0] * tau_xx + qc[2] / qc[0] * tau_xy + g.k * gT0;
        const T phi1 = qc[1] / qc[0] * tau_xy + qc[2] / qc[0] * tau_yy + g.k * gT1;
        f1 -= nx * tau_xx + ny * tau_xy;
        f2 -= nx * tau_xy + ny * tau_yy;
        f3 -= nx * phi0 + ny * phi1;
    }
    f[0] = f0; f[1] = f1; f[2] = f2; f[3] = f3;
}

// vars() of each functor (physics.h:267-276, 311-339, 377-405, 446-530)
template <typename T>
__device__ __forceinline__ void face_vars(const GasT<T>& g, int kind, T nx, T ny, const T* q_L, const T* q_bc, T* q_R) {
    if (kind == K_INTERNAL) {
        for (int k = 0; k < 4; ++k) q_R[k] = (q_L[k] + q_bc[k]) * T(0.5);
    } else if (kind == K_SLIP) {
        const T rhoV = q_L[1] * nx + q_L[2] * ny;
        q_R[0] = q_L[0];
        q_R[1] = q_L[1] - T(2.) * rhoV * nx;
        q_R[2] = q_L[2] - T(2.) * rhoV * ny;
        q_R[3] = q_L[3];
    } else if (kind == K_WALL) {
        q_R[0] = q_L[0]; q_R[1] = -q_L[1]; q_R[2] = -q_L[2]; q_R[3] = q_L[3];
    } else {
        const T gm1 = g.gamma - T(1);
        const T rho = q_L[0], rho_u = q_L[1], rho_v = q_L[2], rho_e = q_L[3];
        const T bc_rho = q_bc[0], bc_rhou = q_bc[1], bc_rhov = q_bc[2], bc_rhoe = q_bc[3];
        const T bc_u = bc_rhou / bc_rho;
        const T bc_v = bc_rhov / bc_rho;
        const T bc_p = gm1 * (bc_rhoe - T(0.5) / bc_rho * (bc_rhou * bc_rhou + bc_rhov * bc_rhov));
        const T p = gm1 * (rho_e - T(0.5) / rho * (rho_u * rho_u + rho_v * rho_v));
        const T c = dsqrt<T>(g.gamma * p / rho);
        const T mach = dsqrt<T>(rho_u * rho_u + rho_v * rho_v) / (rho * c);
        const T io = rho_u * nx + rho_v * ny;
        if (mach > T(1)) {
            if (io < T(0)) {
                q_R[0] = bc_rho;
                q_R[1] = bc_rho * bc_u;
                q_R[2] = bc_rho * bc_v;
                q_R[3] = bc_p / gm1 + T(0.5) * bc_rho * (bc_u * bc_u + bc_v * bc_v);
            } else {
                q_R[0] = rho; q_R[1] = rho_u; q_R[2] = rho_v; q_R[3] = rho_e;
            }
        } else {
            const T pa = bc_p, rhoa = bc_rho, ua = bc_u, va = bc_v;
            const T pd = p, rhod = rho, ud = rho_u / rho, vd = rho_v / rho;
            const T rho0 = rho, c0 = c;
            if (io < T(0)) {
                const T pb = T(0.5) * (pa + pd - rho0 * c0 * (nx * (ua - ud) + ny * (va - vd)));
                const T r0 = rhoa + (pb - pa) / (c0 * c0);
                const T r1 = r0 * (ua - nx * (pa - pb) / (rho0 * c0));
                const T r2 = r0 * (va - ny * (pa - pb) / (rho0 * c0));
                q_R[0] = r0; q_R[1] = r1; q_R[2] = r2;
                q_R[3] = pb / gm1 + T(0.5) / r0 * (r1 * r1 + r2 * r2);
            } else {
                const T pb = pa;
                const T r0 = rhod + (pb - pd) / (c0 * c0);
                const T r1 = r0 * (ud + nx * (pd - pb) / (rho0 * c0));
                const T r2 = r0 * (va + ny * (pd - pb) / (rho0 * c0));   // farfield v, physics.h:524
                q_R[0] = r0; q_R[1] = r1; q_R[2] = r2;
                q_R[3] = pb / gm1 + T(0.5) / r0 * (r1 * r1 + r2 * r2);
            }
        }
    }
}

// operator() of the four functors: boundary functors evaluate the Roe kernel on
// (q_L, vars(q_L, q_ghost)) without gradients (physics.h:287-299,352-364,417-432)
template <typename T>
__device__ __forceinline__ void face_flux(const GasT<T>& g, int kind, T nx, T ny, const T* q_L, const T* q_R,
                                          bool laminar, const T* gx, const T* gy, T* f) {
    if (kind == K_INTERNAL) {
        roe_flux<T>(g, nx, ny, q_L, q_R, laminar, gx, gy, f);
    } else {
        T qb[4];
        face_vars<T>(g, kind, nx, ny, q_L, q_R, qb);
        roe_flux<T>(g, nx, ny, q_L, qb, false, gx, gy, f);
    }
}

// ---------------------------------------------------------------- kernels --
constexpr int kBlock = 256;

template <typename T>
__global__ void __launch_bounds__(kBlock) k_set_walls(const int32_t* __restrict__ owner, const int32_t* __restrict__ ghost,
                                                      int32_t n, T* __restrict__ q, int32_t nT, int nv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = owner[i], gh = ghost[i];
    for (int k = 0; k < nv; ++k) q[(size_t)k * nT + gh] = q[(size_t)k * nT + o];
}

// solver.h:308-356: dt_i = cfl*A_i / sum_e eig_e*len_e, ghost cells get cfl*A/0
template <typename T>
__global__ void __launch_bounds__(kBlock) k_dt(CellsT<T> C, FacesT<T> F, const T* __restrict__ q, GasT<T> g, T cfl,
                                               T* __restrict__ dt) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nT) return;
    const size_t nT = (size_t)C.nT;
    T s = T(0);
    if (c < C.nC) {
        for (int slot = 0; slot < 4; ++slot) {
            uint32_t e = C.cface[(size_t)slot * C.nC + c];
            if (e == 0xFFFFFFFFu) break;
            int f = (int)(e & 0x7FFFFFFFu);
            int o = F.c0[f], n = F.c1[f];
            T nx = F.nx[f], ny = F.ny[f];
            const T gm1 = g.gamma - T(1);
            T q0 = q[o], q1 = q[nT + o], q2 = q[2 * nT + o], q3 = q[3 * nT + o];
            T V_L = (q1 * nx + q2 * ny) / q0;
            T p_L = gm1 * (q3 - T(0.5) / q0 * (q1 * q1 + q2 * q2));
            T eig = dsqrt<T>(p_L * g.gamma / q0) + dabs<T>(V_L);
            if (F.kind[f] == K_INTERNAL) {
                T r0 = q[n], r1 = q[nT + n], r2 = q[2 * nT + n], r3 = q[3 * nT + n];
                T V_R = (r1 * nx + r2 * ny) / r0;
                T p_R = gm1 * (r3 - T(0.5) / r0 * (r1 * r1 + r2 * r2));
                T eig_R = dsqrt<T>(p_R * g.gamma / r0) + dabs<T>(V_R);
                eig = (eig < eig_R) ? eig_R : eig;   // std::max(eig_L, eig_R)
            }
            s += eig * F.len[f];
        }
    }
    dt[c] = cfl * C.area[c] / s;
}

// solver.h:425-514 (gradients from qg) fused with solver.h:517-593 (limiters from ql)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_grad_lim(CellsT<T> C, FacesT<T> F, const T* __restrict__ qg,
                                                     const T* __restrict__ ql, GasT<T> g, int gg, int second_order,
                                                     T limiter_k, T* __restrict__ gx, T* __restrict__ gy,
                                                     T* __restrict__ lim) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT;
    uint32_t ef[4];
    for (int s = 0; s < 4; ++s) ef[s] = C.cface[(size_t)s * C.nC + c];
    T sx[4] = {T(0), T(0), T(0), T(0)}, sy[4] = {T(0), T(0), T(0), T(0)};
    T qcg[4];
    for (int k = 0; k < 4; ++k) qcg[k] = qg[k * nT + c];
    if (gg) {
        for (int s = 0; s < 4; ++s) {
            uint32_t e = ef[s];
            if (e == 0xFFFFFFFFu) break;
            int f = (int)(e & 0x7FFFFFFFu);
            bool nb_side = (e >> 31) != 0;
            int o = F.c0[f], n = F.c1[f];
            int other = nb_side ? o : n;
            T qo_[4], qother[4];
            for (int k = 0; k < 4; ++k) qother[k] = qg[k * nT + other];
            const T* qo = nb_side ? qother : qcg;
            const T* qn = nb_side ? qcg : qother;
            for (int k = 0; k < 4; ++k) qo_[k] = qo[k];
            T nx = F.nx[f], ny = F.ny[f];
            T qv[4];
            face_vars<T>(g, F.kind[f], nx, ny, qo_, qn, qv);
            T gf = F.gw[f], L = F.len[f];
            for (int k = 0; k < 4; ++k) {
                T fk = (qo_[k] * (T(1.0) - gf) + qv[k] * gf) * L;
                if (!nb_side) { sx[k] += fk * nx; sy[k] += fk * ny; }
                else { sx[k] -= fk * nx; sy[k] -= fk * ny; }
            }
        }
        T A = C.area[c];
        for (int k = 0; k < 4; ++k) { sx[k] /= A; sy[k] /= A; }
    } else {
        T rx[4] = {T(0), T(0), T(0), T(0)}, ry[4] = {T(0), T(0), T(0), T(0)};
        T xc = C.xc[c], yc = C.yc[c];
        for (int s = 0; s < 4; ++s) {
            uint32_t e = ef[s];
            if (e == 0xFFFFFFFFu) break;
            int f = (int)(e & 0x7FFFFFFFu);
            bool nb_side = (e >> 31) != 0;
            int other = nb_side ? F.c0[f] : F.c1[f];
            T dx = C.xc[other] - xc, dy = C.yc[other] - yc;
            T qn[4], qv[4];
            for (int k = 0; k < 4; ++k) qn[k] = qg[k * nT + other];
            face_vars<T>(g, F.kind[f], F.nx[f], F.ny[f], qcg, qn, qv);
            for (int k = 0; k < 4; ++k) {
                T del = qcg[k] - qv[k];
                rx[k] += dx * del;
                ry[k] += dy * del;
            }
        }
        T m0 = C.lsq[c], m1 = C.lsq[(size_t)C.nC + c], m2 = C.lsq[2 * (size_t)C.nC + c], m3 = C.lsq[3 * (size_t)C.nC + c];
        for (int k = 0; k < 4; ++k) {
            sx[k] = m0 * rx[k] + m1 * ry[k];
            sy[k] = m2 * rx[k] + m3 * ry[k];
        }
    }
    for (int k = 0; k < 4; ++k) { gx[k * nT + c] = sx[k]; gy[k * nT + c] = sy[k]; }
    if (!second_order) return;
    T qc[4], qmin[4], qmax[4];
    for (int k = 0; k < 4; ++k) { qc[k] = ql[k * nT + c]; qmin[k] = qc[k]; qmax[k] = qc[k]; }
    for (int s = 0; s < 4; ++s) {
        uint32_t e = ef[s];
        if (e == 0xFFFFFFFFu) break;
        int f = (int)(e & 0x7FFFFFFFu);
        int other = (e >> 31) ? F.c0[f] : F.c1[f];
        for (int k = 0; k < 4; ++k) {
            T v = ql[k * nT + other];
            qmin[k] = (v < qmin[k]) ? v : qmin[k];
            qmax[k] = (qmax[k] < v) ? v : qmax[k];
        }
    }
    const T xc = C.xc[c], yc = C.yc[c];
    const T Ka = limiter_k * dsqrt<T>(C.area[c]);
    const T K3a = Ka * Ka * Ka;
    T lm[4] = {T(1), T(1), T(1), T(1)};
    for (int s = 0; s < 4; ++s) {
        uint32_t e = ef[s];
        if (e == 0xFFFFFFFFu) break;
        int f = (int)(e & 0x7FFFFFFFu);
        const T dx = F.xf[f] - xc, dy = F.yf[f] - yc;
        for (int k = 0; k < 4; ++k) {
            T dqg = sx[k] * dx + sy[k] * dy;
            T dmax = qmax[k] - qc[k];
            T dmin = qmin[k] - qc[k];
            T l = T(1.0);
            if (dqg > T(1e-16))
                l = T(1) / dqg * ((dmax * dmax + K3a) * dqg + T(2) * dqg * dqg * dmax) / (dmax * dmax + T(2) * dqg * dqg + dmax * dqg + K3a);
            else if (dqg < T(-1e-16))
                l = T(1) / dqg * ((dmin * dmin + K3a) * dqg + T(2) * dqg * dqg * dmin) / (dmin * dmin + T(2) * dqg * dqg + dmin * dqg + K3a);
            lm[k] = (l < lm[k]) ? l : lm[k];
        }
    }
    for (int k = 0; k < 4; ++k) lim[k * nT + c] = lm[k];
}

// flux of face f times its length, reconstructed states (solver.h:1106-1134)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_face_flux(CellsT<T> C, FacesT<T> F, const T* __restrict__ qr,
                                                      const T* __restrict__ qa, const T* __restrict__ gx,
                                                      const T* __restrict__ gy, const T* __restrict__ lim, GasT<T> g,
                                                      int viscous, int second_order, const T* __restrict__ uniform,
                                                      T* __restrict__ flux) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT;
    const int o = F.c0[f], n = F.c1[f];
    const int kind = F.kind[f];
    const T nx = F.nx[f], ny = F.ny[f];
    T qL[4], qR[4];
    const bool laminar = (viscous == 1) && (kind == K_INTERNAL);
    T gxa[4] = {T(0), T(0), T(0), T(0)}, gya[4] = {T(0), T(0), T(0), T(0)};
    if (uniform) {
        for (int k = 0; k < 4; ++k) { qL[k] = uniform[k]; qR[k] = uniform[k]; }
    } else {
        for (int k = 0; k < 4; ++k) { qL[k] = qr[k * nT + o]; qR[k] = qr[k * nT + n]; }
        if (second_order) {
            const T xfv = F.xf[f], yfv = F.yf[f];
            const T d0x = xfv - C.xc[o], d0y = yfv - C.yc[o];
            const T d1x = xfv - C.xc[n], d1y = yfv - C.yc[n];
            for (int k = 0; k < 4; ++k) {
                qL[k] = qL[k] + (gx[k * nT + o] * d0x + gy[k * nT + o] * d0y) * lim[k * nT + o];
                qR[k] = qR[k] + (gx[k * nT + n] * d1x + gy[k * nT + n] * d1y) * lim[k * nT + n];
            }
        }
    }
    if (laminar) {
        // solver.h:359-398 (direction from the SUM of the two centroids)
        T t0 = C.xc[o] + C.xc[n];
        T t1 = C.yc[o] + C.yc[n];
        const T l = dsqrt<T>(t0 * t0 + t1 * t1);
        t0 /= l; t1 /= l;
        for (int k = 0; k < 4; ++k) {
            T dir = (qa[k * nT + o] - qa[k * nT + n]) / l;
            T gxb = (gx[k * nT + o] + gx[k * nT + n]) * T(0.5);
            T gyb = (gy[k * nT + o] + gy[k * nT + n]) * T(0.5);
            const T dot = gxb * t0 + gyb * t1;
            gxa[k] = gxb - (dot - dir) * t0;
            gya[k] = gyb - (dot - dir) * t1;
        }
    }
    T fl[4];
    face_flux<T>(g, kind, nx, ny, qL, qR, laminar, gxa, gya, fl);
    const T L = F.len[f];
    for (int k = 0; k < 4; ++k) flux[(size_t)k * F.nF + f] = fl[k] * L;
}

template <typename T>
struct Acc {
    T s = T(0);
    __device__ __forceinline__ void add(T x) { s = s + x; }
    __device__ __forceinline__ void sub(T x) { s = s - x; }
    __device__ __forceinline__ T value() const { return s; }
};
// Neumaier-compensated FP32 accumulation
template <>
struct Acc<float> {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ void sub(float x) { add(-x); }
    __device__ __forceinline__ float value() const { return s + c; }
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_gather(CellsT<T> C, const T* __restrict__ flux, int32_t nF, int mode,
                                                   T* __restrict__ R) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    const size_t nT = (size_t)C.nT;
    Acc<T> acc[4];
    if (mode == 2)
        for (int k = 0; k < 4; ++k) acc[k].s = R[k * nT + c];
    for (int s = 0; s < 4; ++s) {
        uint32_t e = C.cface[(size_t)s * C.nC + c];
        if (e == 0xFFFFFFFFu) break;
        int f = (int)(e & 0x7FFFFFFFu);
        if (e >> 31) for (int k = 0; k < 4; ++k) acc[k].add(flux[(size_t)k * nF + f]);
        else for (int k = 0; k < 4; ++k) acc[k].sub(flux[(size_t)k * nF + f]);
    }
    if (mode == 1) {
        T A = C.area[c];
        for (int k = 0; k < 4; ++k) R[k * nT + c] = acc[k].value() / A;
    } else {
        for (int k = 0; k < 4; ++k) R[k * nT + c] = acc[k].value();
    }
}

// ---------------------------------------------------------------- launchers --
inline int nblocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

#define AFX_INSTANTIATE_LAUNCHERS(T)                                                                              \
    template <>                                                                                                   \
    hipError_t launch_set_walls<T>(hipStream_t s, const int32_t* owner, const int32_t* ghost, int32_t n, T* q,     \
                                   int32_t nT, int nv) {                                                          \
        if (n <= 0) return hipSuccess;                                                                            \
        hipLaunchKernelGGL(k_set_walls<T>, dim3(nblocks(n)), dim3(kBlock), 0, s, owner, ghost, n, q, nT, nv);    \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_dt<T>(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* q, GasT<T> g, T cfl, T* dt) {       \
        hipLaunchKernelGGL(k_dt<T>, dim3(nblocks(C.nT)), dim3(kBlock), 0, s, C, F, q, g, cfl, dt);                \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_grad_lim<T>(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* qg, const T* ql, GasT<T> g,   \
                                  int gg, int so, T lk, T* gx, T* gy, T* lim) {                                   \
        hipLaunchKernelGGL(k_grad_lim<T>, dim3(nblocks(C.nC)), dim3(kBlock), 0, s, C, F, qg, ql, g, gg, so, lk,   \
                           gx, gy, lim);                                                                          \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_face_flux<T>(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* qr, const T* qa, const T* gx, \
                                   const T* gy, const T* lim, GasT<T> g, int viscous, int so, const T* uniform,   \
                                   T* flux) {                                                                     \
        hipLaunchKernelGGL(k_face_flux<T>, dim3(nblocks(F.nF)), dim3(kBlock), 0, s, C, F, qr, qa, gx, gy, lim, g, \
                           viscous, so, uniform, flux);                                                           \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_gather<T>(hipStream_t s, CellsT<T> C, const T* flux, int32_t nF, int mode, T* R) {          \
        hipLaunchKernelGGL(k_gather<T>, dim3(nblocks(C.nC)), dim3(kBlock), 0, s, C, flux, nF, mode, R);           \
        return hipGetLastError();                                                                                 \
    }

}  // namespace afx
