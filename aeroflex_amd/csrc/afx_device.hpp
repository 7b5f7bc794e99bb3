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
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename T> __device__ __forceinline__ T dabs(T x) { return x < T(0) ? -x : x; }

// Division: IEEE-exact in FP64 (reference rounding); hardware reciprocal
// (v_rcp_f32, about 1 ulp) times the numerator in the FP32 form.
template <typename T> __device__ __forceinline__ T ddiv(T a, T b);
template <> __device__ __forceinline__ double ddiv<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float ddiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// physics.h:84-86
template <typename T> __device__ __forceinline__ T sabs(T x) { return dsqrt<T>(x * x + T(1e-4)); }
// physics.h:133-135
template <typename T> __device__ __forceinline__ T entropy_fix(T l, T d) { return l > d ? l : ddiv<T>(l * l + d * d, T(2) * d); }

// solver.h:578-584: limiter value of one (cell, face, variable).  The
// reference's two branches share one expression in the selected bound, so the
// select below returns the same bits as the branches.
template <typename T>
__device__ __forceinline__ T limiter_phi(T dqg, T dmax, T dmin, T K3a) {
    const bool pos = dqg > T(1e-16);
    const bool neg = dqg < T(-1e-16);
    const T d = pos ? dmax : dmin;
    const T l = ddiv<T>(ddiv<T>(T(1), dqg) * ((d * d + K3a) * dqg + T(2) * dqg * dqg * d), d * d + T(2) * dqg * dqg + d * dqg + K3a);
    return (pos || neg) ? l : T(1.0);
}

// physics.h:160-228 (Roe + Harten) and 230-258 (laminar), internal_flux::operator()
template <typename T>
__device__ __forceinline__ void roe_flux(const GasT<T>& g, T nx, T ny, const T* q_L, const T* q_R, bool laminar,
                                         const T* gx, const T* gy, T* f) {
    const T rho_L = q_L[0], rho_u_L = q_L[1], rho_v_L = q_L[2], rho_e_L = q_L[3];
    const T rho_R = q_R[0], rho_u_R = q_R[1], rho_v_R = q_R[2], rho_e_R = q_R[3];
    const T gm1 = g.gamma - T(1);
    const T V_L = ddiv<T>(rho_u_L * nx + rho_v_L * ny, rho_L);
    const T V_R = ddiv<T>(rho_u_R * nx + rho_v_R * ny, rho_R);
    const T p_L = gm1 * (rho_e_L - ddiv<T>(T(0.5), rho_L) * (rho_u_L * rho_u_L + rho_v_L * rho_v_L));
    const T p_R = gm1 * (rho_e_R - ddiv<T>(T(0.5), rho_R) * (rho_u_R * rho_u_R + rho_v_R * rho_v_R));
    T f0 = (V_L * rho_L + V_R * rho_R) * T(0.5);
    T f1 = (V_L * rho_u_L + p_L * nx + V_R * rho_u_R + p_R * nx) * T(0.5);
    T f2 = (V_L * rho_v_L + p_L * ny + V_R * rho_v_R + p_R * ny) * T(0.5);
    T f3 = (V_L * (rho_e_L + p_L) + V_R * (rho_e_R + p_R)) * T(0.5);
    const T pL = p_L, pR = p_R;
    const T uL = ddiv<T>(rho_u_L, rho_L);
    const T uR = ddiv<T>(rho_u_R, rho_R);
    const T vL = ddiv<T>(rho_v_L, rho_L);
    const T vR = ddiv<T>(rho_v_R, rho_R);
    const T srhoL = dsqrt<T>(rho_L);
    const T srhoR = dsqrt<T>(rho_R);
    const T rho = srhoR * srhoL;
    const T u = ddiv<T>(uL * srhoL + uR * srhoR, srhoL + srhoR);
    const T v = ddiv<T>(vL * srhoL + vR * srhoR, srhoL + srhoR);
    const T h = ddiv<T>(ddiv<T>(rho_e_L + pL, rho_L) * srhoL + ddiv<T>(rho_e_R + pR, rho_R) * srhoR, srhoL + srhoR);
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
    const T kF1 = ddiv<T>(lambda_cm * (dp - rho * c * dV), T(2) * c * c);
    const T kF234_0 = lambda_c * ((rho_R - rho_L) - ddiv<T>(dp, c * c));
    const T kF234_1 = lambda_c * rho;
    const T kF5 = ddiv<T>(lambda_cp * (dp + rho * c * dV), T(2) * c * c);
    f0 -= T(0.5) * (kF1 + kF234_0 + kF5);
    f1 -= T(0.5) * (kF1 * (u - c * nx) + kF234_0 * u + kF234_1 * (uR - uL - dV * nx) + kF5 * (u + c * nx));
    f2 -= T(0.5) * (kF1 * (v - c * ny) + kF234_0 * v + kF234_1 * (vR - vL - dV * ny) + kF5 * (v + c * ny));
    f3 -= T(0.5) * (kF1 * (h - c * V) + kF234_0 * q2 * T(0.5) + kF234_1 * (u * (uR - uL) + v * (vR - vL) - V * dV) + kF5 * (h + c * V));
    if (laminar) {
        T qc[4];
        for (int k = 0; k < 4; ++k) qc[k] = T(0.5) * (q_L[k] + q_R[k]);
        const T r2 = qc[0] * qc[0];
        // physics.h:33-68 temperature gradient through the pressure gradient
        const T p = gm1 * (qc[3] - ddiv<T>(T(0.5), qc[0]) * (qc[1] * qc[1] + qc[2] * qc[2]));
        T gp0 = T(2.) * gx[3];
        gp0 -= ddiv<T>(gx[1] * qc[1], qc[0]) + ddiv<T>(qc[1] * (gx[1] * qc[0] - gx[0] * qc[1]), r2);
        gp0 -= ddiv<T>(gx[2] * qc[2], qc[0]) + ddiv<T>(qc[2] * (gx[2] * qc[0] - gx[0] * qc[2]), r2);
        gp0 *= T(0.5) * gm1;
        T gp1 = T(2.) * gy[3];
        gp1 -= ddiv<T>(gy[1] * qc[1], qc[0]) + ddiv<T>(qc[1] * (gy[1] * qc[0] - gy[0] * qc[1]), r2);
        gp1 -= ddiv<T>(gy[2] * qc[2], qc[0]) + ddiv<T>(qc[2] * (gy[2] * qc[0] - gy[0] * qc[2]), r2);
        gp1 *= T(0.5) * gm1;
        const T gT0 = ddiv<T>(T(1.), g.R) * ddiv<T>(gp0 * qc[0] - gx[0] * p, r2);
        const T gT1 = ddiv<T>(T(1.), g.R) * ddiv<T>(gp1 * qc[0] - gy[0] * p, r2);
        const T gu0 = ddiv<T>(qc[0] * gx[1] - qc[1] * gx[0], r2);
        const T gu1 = ddiv<T>(qc[0] * gy[1] - qc[1] * gy[0], r2);
        const T gv0 = ddiv<T>(qc[0] * gx[2] - qc[2] * gx[0], r2);
        const T gv1 = ddiv<T>(qc[0] * gy[2] - qc[2] * gy[0], r2);
        const T div_v = gu0 + gv1;
        const T tau_xx = T(2.) * g.mu * (gu0 - ddiv<T>(div_v, T(3.)));
        const T tau_yy = T(2.) * g.mu * (gv1 - ddiv<T>(div_v, T(3.)));
        const T tau_xy = g.mu * (gu1 + gv0);
        const T phi0 = ddiv<T>(qc[1], qc[0]) * tau_xx + ddiv<T>(qc[2], qc[0]) * tau_xy + g.k * gT0;
        const T phi1 = ddiv<T>(qc[1], qc[0]) * tau_xy + ddiv<T>(qc[2], qc[0]) * tau_yy + g.k * gT1;
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
        const T bc_u = ddiv<T>(bc_rhou, bc_rho);
        const T bc_v = ddiv<T>(bc_rhov, bc_rho);
        const T bc_p = gm1 * (bc_rhoe - ddiv<T>(T(0.5), bc_rho) * (bc_rhou * bc_rhou + bc_rhov * bc_rhov));
        const T p = gm1 * (rho_e - ddiv<T>(T(0.5), rho) * (rho_u * rho_u + rho_v * rho_v));
        const T c = dsqrt<T>(ddiv<T>(g.gamma * p, rho));
        const T mach = ddiv<T>(dsqrt<T>(rho_u * rho_u + rho_v * rho_v), rho * c);
        const T io = rho_u * nx + rho_v * ny;
        if (mach > T(1)) {
            if (io < T(0)) {
                q_R[0] = bc_rho;
                q_R[1] = bc_rho * bc_u;
                q_R[2] = bc_rho * bc_v;
                q_R[3] = ddiv<T>(bc_p, gm1) + T(0.5) * bc_rho * (bc_u * bc_u + bc_v * bc_v);
            } else {
                q_R[0] = rho; q_R[1] = rho_u; q_R[2] = rho_v; q_R[3] = rho_e;
            }
        } else {
            const T pa = bc_p, rhoa = bc_rho, ua = bc_u, va = bc_v;
            const T pd = p, rhod = rho, ud = ddiv<T>(rho_u, rho), vd = ddiv<T>(rho_v, rho);
            const T rho0 = rho, c0 = c;
            if (io < T(0)) {
                const T pb = T(0.5) * (pa + pd - rho0 * c0 * (nx * (ua - ud) + ny * (va - vd)));
                const T r0 = rhoa + ddiv<T>(pb - pa, c0 * c0);
                const T r1 = r0 * (ua - ddiv<T>(nx * (pa - pb), rho0 * c0));
                const T r2 = r0 * (va - ddiv<T>(ny * (pa - pb), rho0 * c0));
                q_R[0] = r0; q_R[1] = r1; q_R[2] = r2;
                q_R[3] = ddiv<T>(pb, gm1) + ddiv<T>(T(0.5), r0) * (r1 * r1 + r2 * r2);
            } else {
                const T pb = pa;
                const T r0 = rhod + ddiv<T>(pb - pd, c0 * c0);
                const T r1 = r0 * (ud + ddiv<T>(nx * (pd - pb), rho0 * c0));
                const T r2 = r0 * (va + ddiv<T>(ny * (pd - pb), rho0 * c0));   // farfield v, physics.h:524
                q_R[0] = r0; q_R[1] = r1; q_R[2] = r2;
                q_R[3] = ddiv<T>(pb, gm1) + ddiv<T>(T(0.5), r0) * (r1 * r1 + r2 * r2);
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
            T V_L = ddiv<T>(q1 * nx + q2 * ny, q0);
            T p_L = gm1 * (q3 - ddiv<T>(T(0.5), q0) * (q1 * q1 + q2 * q2));
            T eig = dsqrt<T>(ddiv<T>(p_L * g.gamma, q0)) + dabs<T>(V_L);
            if (F.kind[f] == K_INTERNAL) {
                T r0 = q[n], r1 = q[nT + n], r2 = q[2 * nT + n], r3 = q[3 * nT + n];
                T V_R = ddiv<T>(r1 * nx + r2 * ny, r0);
                T p_R = gm1 * (r3 - ddiv<T>(T(0.5), r0) * (r1 * r1 + r2 * r2));
                T eig_R = dsqrt<T>(ddiv<T>(p_R * g.gamma, r0)) + dabs<T>(V_R);
                eig = (eig < eig_R) ? eig_R : eig;   // std::max(eig_L, eig_R)
            }
            s += eig * F.len[f];
        }
    }
    dt[c] = ddiv<T>(cfl * C.area[c], s);
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
        for (int k = 0; k < 4; ++k) { sx[k] = ddiv<T>(sx[k], A); sy[k] = ddiv<T>(sy[k], A); }
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
            const T l = limiter_phi<T>(dqg, dmax, dmin, K3a);
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
        t0 = ddiv<T>(t0, l); t1 = ddiv<T>(t1, l);
        for (int k = 0; k < 4; ++k) {
            T dir = ddiv<T>(qa[k * nT + o] - qa[k * nT + n], l);
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
        for (int k = 0; k < 4; ++k) R[k * nT + c] = ddiv<T>(acc[k].value(), A);
    } else {
        for (int k = 0; k < 4; ++k) R[k * nT + c] = acc[k].value();
    }
}


// XCD-aware block -> tile map: the 8 XCDs (block b runs on XCD b % 8) each
// take a contiguous range of tiles so that halo cells shared by neighbouring
// tiles are found in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int per = nblocks >> 3;   // nblocks is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_residual_tile(TilesT<T> S, const T* __restrict__ q, int32_t nT, int32_t nC,
                                                          const T* __restrict__ area, const T* __restrict__ lsq,
                                                          const T* __restrict__ xc, const T* __restrict__ yc,
                                                          GasT<T> g, int gg, int second_order, T limiter_k,
                                                          int viscous, T* __restrict__ R, T* __restrict__ qghost) {
    extern __shared__ unsigned char smem_raw[];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= S.ntiles) return;
    const int c0g = S.cell0[t], nTt = S.cell0[t + 1] - c0g;
    const int lo = S.loff[t], nL = S.loff[t + 1] - lo;
    const int nG = S.ng[t];
    const int fo = S.foff[t], nFT = S.nft[t];
    const int go = S.goff[t];
    const int LM = S.lmax, GM = S.gmax, FM = S.ftmax;
    const int64_t NF = S.nFtot;
    T* sq = reinterpret_cast<T*>(smem_raw);   // [4][LM] states of L
    T* sgx = sq + 4 * LM;                      // [4][GM]
    T* sgy = sgx + 4 * GM;
    T* slm = sgy + 4 * GM;
    T* sfl = slm + 4 * GM;                     // [4][FM] face fluxes x length
    const int* Lg = S.Lglob + lo;

    // 1. stage the states of the local cells
    for (int i = threadIdx.x; i < nL; i += blockDim.x) {
        const int gi = Lg[i];
        if (gi < 0) continue;   // wall/slip ghost: never read
        for (int k = 0; k < 4; ++k) sq[k * LM + i] = q[(size_t)k * nT + gi];
    }
    __syncthreads();

    // 2. gradients (solver.h:425-514) and limiters (solver.h:517-593) of T u H1
    for (int c = threadIdx.x; c < nG; c += blockDim.x) {
        const uint64_t sw = S.slots[go + c];
        T qc[4];
        for (int k = 0; k < 4; ++k) qc[k] = sq[k * LM + c];
        T sx[4] = {T(0), T(0), T(0), T(0)}, sy[4] = {T(0), T(0), T(0), T(0)};
        const int gc = Lg[c];
        if (gg) {
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const bool nb_side = (e >> 15) != 0;
                const uint32_t fc = S.fcell[f];
                const int o = (int)(fc & 0xFFFFu), n = (int)(fc >> 16);
                T qo[4], qn[4], qv[4];
                for (int k = 0; k < 4; ++k) { qo[k] = sq[k * LM + o]; qn[k] = sq[k * LM + n]; }
                const T nx = S.geo[f], ny = S.geo[NF + f], L = S.geo[2 * NF + f], gf = S.geo[3 * NF + f];
                face_vars<T>(g, S.fkind[f], nx, ny, qo, qn, qv);
                for (int k = 0; k < 4; ++k) {
                    const T fk = (qo[k] * (T(1.0) - gf) + qv[k] * gf) * L;
                    if (!nb_side) { sx[k] += fk * nx; sy[k] += fk * ny; }
                    else { sx[k] -= fk * nx; sy[k] -= fk * ny; }
                }
            }
            const T A = area[gc];
            for (int k = 0; k < 4; ++k) { sx[k] = ddiv<T>(sx[k], A); sy[k] = ddiv<T>(sy[k], A); }
        } else {
            T rx[4] = {T(0), T(0), T(0), T(0)}, ry[4] = {T(0), T(0), T(0), T(0)};
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const bool nb_side = (e >> 15) != 0;
                const uint32_t fc = S.fcell[f];
                const int other = nb_side ? (int)(fc & 0xFFFFu) : (int)(fc >> 16);
                const T dcx = nb_side ? S.geo[6 * NF + f] : S.geo[4 * NF + f];
                const T dcy = nb_side ? S.geo[7 * NF + f] : S.geo[5 * NF + f];
                const T dox = nb_side ? S.geo[4 * NF + f] : S.geo[6 * NF + f];
                const T doy = nb_side ? S.geo[5 * NF + f] : S.geo[7 * NF + f];
                const T dx = dcx - dox, dy = dcy - doy;   // x_other - x_cell
                T qn[4], qv[4];
                for (int k = 0; k < 4; ++k) qn[k] = sq[k * LM + other];
                face_vars<T>(g, S.fkind[f], S.geo[f], S.geo[NF + f], qc, qn, qv);
                for (int k = 0; k < 4; ++k) {
                    const T del = qc[k] - qv[k];
                    rx[k] += dx * del;
                    ry[k] += dy * del;
                }
            }
            const T m0 = lsq[gc], m1 = lsq[(size_t)nC + gc], m2 = lsq[2 * (size_t)nC + gc], m3 = lsq[3 * (size_t)nC + gc];
            for (int k = 0; k < 4; ++k) {
                sx[k] = m0 * rx[k] + m1 * ry[k];
                sy[k] = m2 * rx[k] + m3 * ry[k];
            }
        }
        for (int k = 0; k < 4; ++k) { sgx[k * GM + c] = sx[k]; sgy[k * GM + c] = sy[k]; }
        if (second_order) {
            T qmin[4], qmax[4];
            for (int k = 0; k < 4; ++k) { qmin[k] = qc[k]; qmax[k] = qc[k]; }
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const int kind = S.fkind[f];
                // a wall/slip ghost holds the owner's state (set_walls_from_internal)
                if (kind == K_WALL || kind == K_SLIP) continue;
                const uint32_t fc = S.fcell[f];
                const int other = (e >> 15) ? (int)(fc & 0xFFFFu) : (int)(fc >> 16);
                for (int k = 0; k < 4; ++k) {
                    const T v = sq[k * LM + other];
                    qmin[k] = (v < qmin[k]) ? v : qmin[k];
                    qmax[k] = (qmax[k] < v) ? v : qmax[k];
                }
            }
            const T Ka = limiter_k * dsqrt<T>(area[gc]);
            const T K3a = Ka * Ka * Ka;
            T lm[4] = {T(1), T(1), T(1), T(1)};
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const bool nb_side = (e >> 15) != 0;
                const T dx = nb_side ? S.geo[6 * NF + f] : S.geo[4 * NF + f];
                const T dy = nb_side ? S.geo[7 * NF + f] : S.geo[5 * NF + f];
                for (int k = 0; k < 4; ++k) {
                    const T dqg = sx[k] * dx + sy[k] * dy;
                    const T dmax = qmax[k] - qc[k];
                    const T dmin = qmin[k] - qc[k];
                    const T l = limiter_phi<T>(dqg, dmax, dmin, K3a);
                    lm[k] = (l < lm[k]) ? l : lm[k];
                }
            }
            for (int k = 0; k < 4; ++k) slm[k * GM + c] = lm[k];
        }
    }
    __syncthreads();

    // 3. fluxes of the faces touching T (solver.h:1097-1134)
    for (int fl = threadIdx.x; fl < nFT; fl += blockDim.x) {
        const int f = fo + fl;
        const uint32_t fc = S.fcell[f];
        const int o = (int)(fc & 0xFFFFu), n = (int)(fc >> 16);
        const int kind = S.fkind[f];
        const T nx = S.geo[f], ny = S.geo[NF + f], L = S.geo[2 * NF + f];
        T qL[4], qR[4];
        for (int k = 0; k < 4; ++k) { qL[k] = sq[k * LM + o]; qR[k] = sq[k * LM + n]; }
        if (second_order) {
            const T d0x = S.geo[4 * NF + f], d0y = S.geo[5 * NF + f];
            for (int k = 0; k < 4; ++k)
                qL[k] = qL[k] + (sgx[k * GM + o] * d0x + sgy[k * GM + o] * d0y) * slm[k * GM + o];
            if (n < nG) {
                const T d1x = S.geo[6 * NF + f], d1y = S.geo[7 * NF + f];
                for (int k = 0; k < 4; ++k)
                    qR[k] = qR[k] + (sgx[k * GM + n] * d1x + sgy[k * GM + n] * d1y) * slm[k * GM + n];
            }
        }
        const bool laminar = (viscous == 1) && (kind == K_INTERNAL);
        T gxa[4] = {T(0), T(0), T(0), T(0)}, gya[4] = {T(0), T(0), T(0), T(0)};
        if (laminar) {
            // solver.h:359-398 (direction from the SUM of the two centroids)
            const int go_ = Lg[o], gn_ = Lg[n];
            T t0 = xc[go_] + xc[gn_];
            T t1 = yc[go_] + yc[gn_];
            const T l = dsqrt<T>(t0 * t0 + t1 * t1);
            t0 = ddiv<T>(t0, l); t1 = ddiv<T>(t1, l);
            for (int k = 0; k < 4; ++k) {
                const T dir = ddiv<T>(sq[k * LM + o] - sq[k * LM + n], l);
                const T gxb = (sgx[k * GM + o] + sgx[k * GM + n]) * T(0.5);
                const T gyb = (sgy[k * GM + o] + sgy[k * GM + n]) * T(0.5);
                const T dot = gxb * t0 + gyb * t1;
                gxa[k] = gxb - (dot - dir) * t0;
                gya[k] = gyb - (dot - dir) * t1;
            }
        }
        T F[4];
        face_flux<T>(g, kind, nx, ny, qL, qR, laminar, gxa, gya, F);
        for (int k = 0; k < 4; ++k) sfl[k * FM + fl] = F[k] * L;
    }
    __syncthreads();

    // 4. signed sums in reference face order; one coalesced write per cell
    for (int c = threadIdx.x; c < nTt; c += blockDim.x) {
        const uint64_t sw = S.slots[go + c];
        Acc<T> acc[4];
        for (int s = 0; s < 4; ++s) {
            const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
            if (e == 0xFFFFu) break;
            const int fl = (int)(e & 0x7FFFu);
            if (e >> 15) for (int k = 0; k < 4; ++k) acc[k].add(sfl[k * FM + fl]);
            else for (int k = 0; k < 4; ++k) acc[k].sub(sfl[k * FM + fl]);
            if (qghost) {
                const int kind = S.fkind[fo + fl];
                if (kind == K_WALL || kind == K_SLIP) {
                    const int gh = -1 - Lg[(int)(S.fcell[fo + fl] >> 16)];
                    for (int k = 0; k < 4; ++k) qghost[(size_t)k * nT + gh] = sq[k * LM + c];
                }
            }
        }
        for (int k = 0; k < 4; ++k) R[(size_t)k * nT + c0g + c] = acc[k].value();
    }
}


// ---------------------------------------------------------------------------
// FP32 throughput form of the fused tile residual.  Same algorithm as
// k_residual_tile, restated with reciprocals instead of divisions and the
// limiter of solver.h:578-584 in the algebraically equal form
//   phi = (D^2 + K3a + 2 dqg D) / (D^2 + 2 dqg^2 + D dqg + K3a),
// the Green-Gauss face value of an interior face as
//   (1 - g/2) q_owner + (g/2) q_neighbour          (solver.h:450, physics.h:275),
// and the cell sums accumulated exactly in FP64 from the FP32 face fluxes
// (at most four terms, rounded once to FP32).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Roe flux of physics.h:160-228 with shared reciprocals (FP32)
__device__ __forceinline__ void roe_flux_f32(float gamma, float nx, float ny, const float* qL, const float* qR,
                                             float* f) {
    const float gm1 = gamma - 1.f;
    const float rL = qL[0], ruL = qL[1], rvL = qL[2], reL = qL[3];
    const float rR = qR[0], ruR = qR[1], rvR = qR[2], reR = qR[3];
    const float irL = frcp(rL), irR = frcp(rR);
    const float uL = ruL * irL, vL = rvL * irL, uR = ruR * irR, vR = rvR * irR;
    const float VL = uL * nx + vL * ny, VR = uR * nx + vR * ny;
    const float pL = gm1 * (reL - 0.5f * (ruL * uL + rvL * vL));
    const float pR = gm1 * (reR - 0.5f * (ruR * uR + rvR * vR));
    float f0 = 0.5f * (VL * rL + VR * rR);
    float f1 = 0.5f * (VL * ruL + VR * ruR + (pL + pR) * nx);
    float f2 = 0.5f * (VL * rvL + VR * rvR + (pL + pR) * ny);
    float f3 = 0.5f * (VL * (reL + pL) + VR * (reR + pR));
    const float srL = __builtin_amdgcn_sqrtf(rL), srR = __builtin_amdgcn_sqrtf(rR);
    const float is = frcp(srL + srR);
    const float wL = srL * is, wR = srR * is;
    const float rho = srL * srR;
    const float u = uL * wL + uR * wR;
    const float v = vL * wL + vR * wR;
    const float h = (reL + pL) * irL * wL + (reR + pR) * irR * wR;
    const float q2 = u * u + v * v;
    const float c2 = gm1 * (h - 0.5f * q2);
    const float c = __builtin_amdgcn_sqrtf(c2);
    const float V = u * nx + v * ny;
    const float d = 0.05f * c;
    const float i2d = frcp(2.f * d), dd = d * d;
    float lcm = __builtin_amdgcn_sqrtf((V - c) * (V - c) + 1e-4f);
    float lc = __builtin_amdgcn_sqrtf(V * V + 1e-4f);
    float lcp = __builtin_amdgcn_sqrtf((V + c) * (V + c) + 1e-4f);
    lcm = lcm > d ? lcm : (lcm * lcm + dd) * i2d;
    lc = lc > d ? lc : (lc * lc + dd) * i2d;
    lcp = lcp > d ? lcp : (lcp * lcp + dd) * i2d;
    const float dp = pR - pL, dV = VR - VL;
    const float ic2 = frcp(c2);
    const float rcdV = rho * c * dV;
    const float kF1 = lcm * (dp - rcdV) * (0.5f * ic2);
    const float kF5 = lcp * (dp + rcdV) * (0.5f * ic2);
    const float k0 = lc * ((rR - rL) - dp * ic2);
    const float k1 = lc * rho;
    const float du = uR - uL, dv = vR - vL;
    f0 -= 0.5f * (kF1 + k0 + kF5);
    f1 -= 0.5f * (kF1 * (u - c * nx) + k0 * u + k1 * (du - dV * nx) + kF5 * (u + c * nx));
    f2 -= 0.5f * (kF1 * (v - c * ny) + k0 * v + k1 * (dv - dV * ny) + kF5 * (v + c * ny));
    f3 -= 0.5f * (kF1 * (h - c * V) + k0 * q2 * 0.5f + k1 * (u * du + v * dv - V * dV) + kF5 * (h + c * V));
    f[0] = f0; f[1] = f1; f[2] = f2; f[3] = f3;
}

#ifdef AFX_FP32_UNIT   // defined in afx_k32.hip only
__device__ __forceinline__ float4 f4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
__device__ __forceinline__ float get(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
// slot s of a cell's packed slot word (4 x u16, see TilesF32)
__device__ __forceinline__ uint32_t slot_of(const uint2& w, int s) {
    return s == 0 ? (w.x & 0xFFFFu) : s == 1 ? (w.x >> 16) : s == 2 ? (w.y & 0xFFFFu) : (w.y >> 16);
}
#define F4_LOOP(expr) { expr(x); expr(y); expr(z); expr(w); }

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) k_residual_tile_f32(TilesF32 S, const float4* __restrict__ q, int32_t nC,
                                                              const float* __restrict__ iarea,
                                                              const float* __restrict__ a15,
                                                              const float* __restrict__ lsq,
                                                              const float* __restrict__ xc,
                                                              const float* __restrict__ yc, GasT<float> g, int gg,
                                                              int second_order, float K3, int viscous,
                                                              float4* __restrict__ R, float4* __restrict__ qghost) {
    extern __shared__ float4 smem4[];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= S.ntiles) return;
    const int c0g = S.cell0[t], nTt = S.cell0[t + 1] - c0g;
    const int lo = S.loff[t], nL = S.loff[t + 1] - lo;
    const int nG = S.ng[t];
    const int fo = S.foff[t], nFT = S.nft[t];
    const int go = S.goff[t];
    float4* sq = smem4;              // [LM] states of the local cells
    float4* sgx = sq + S.lmax;       // [GM]
    float4* sgy = sgx + S.gmax;
    float4* slm = sgy + S.gmax;
    float4* sfl = slm + S.gmax;      // [FM] face fluxes x length
    const int* Lg = S.Lglob + lo;
    const uint32_t* fck = S.fck + fo;
    const float4* geo = S.geo + 2 * (int64_t)fo;

    // 1. stage states
    for (int i = threadIdx.x; i < nL; i += blockDim.x) {
        const int gi = Lg[i];
        if (gi >= 0) sq[i] = q[gi];   // wall/slip ghosts (flagged < 0) are never read
    }
    __syncthreads();

    // 2. gradients and limiters of T u H1
    const uint2* slots2 = reinterpret_cast<const uint2*>(S.slots) + go;
    for (int c = threadIdx.x; c < nG; c += blockDim.x) {
        const uint2 sw = slots2[c];
        const int gc = Lg[c];
        const float4 qc = sq[c];
        float4 qmin = qc, qmax = qc;
        float4 sx = f4(0.f, 0.f, 0.f, 0.f), sy = f4(0.f, 0.f, 0.f, 0.f);
        float dxs[4], dys[4];
        int ns = 0;
        if (gg) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = slot_of(sw, s);
                if (e == 0xFFFFu) break;
                const int f = (int)(e & 0x7FFFu);
                const bool nb = (e >> 15) != 0;
                const uint32_t fc = fck[f];
                const int kind = (int)(fc >> 24);
                const int other = nb ? (int)(fc & 0xFFFu) : (int)((fc >> 12) & 0xFFFu);
                const float4 ga = geo[2 * f], gb = geo[2 * f + 1];   // (nx, ny, len, g), (d0x, d0y, d1x, d1y)
                dxs[s] = nb ? gb.z : gb.x;
                dys[s] = nb ? gb.w : gb.y;
                ns = s + 1;
                const float sl = nb ? -ga.z : ga.z;
                const float nlx = ga.x * sl, nly = ga.y * sl;
                if (kind == K_INTERNAL) {
                    const float4 qn = sq[other];
                    const float wn = 0.5f * ga.w, wo = 1.f - wn;
                    const float a = nb ? wn : wo, b = nb ? wo : wn;   // weights of this cell and the other
#define GG_INT(m) { const float fv = a * qc.m + b * qn.m; sx.m += fv * nlx; sy.m += fv * nly; \
                    qmin.m = fminf(qmin.m, qn.m); qmax.m = fmaxf(qmax.m, qn.m); }
                    F4_LOOP(GG_INT)
#undef GG_INT
                } else {
                    float qb[4] = {0.f, 0.f, 0.f, 0.f}, qv[4];
                    const float qa[4] = {qc.x, qc.y, qc.z, qc.w};
                    if (kind == K_FARFIELD) {
                        const float4 t4 = sq[other];
                        qb[0] = t4.x; qb[1] = t4.y; qb[2] = t4.z; qb[3] = t4.w;
                        qmin = f4(fminf(qmin.x, t4.x), fminf(qmin.y, t4.y), fminf(qmin.z, t4.z), fminf(qmin.w, t4.w));
                        qmax = f4(fmaxf(qmax.x, t4.x), fmaxf(qmax.y, t4.y), fmaxf(qmax.z, t4.z), fmaxf(qmax.w, t4.w));
                    }
                    face_vars<float>(g, kind, ga.x, ga.y, qa, qb, qv);
                    const float w1 = 1.f - ga.w;
                    const float fv0 = qa[0] * w1 + qv[0] * ga.w, fv1 = qa[1] * w1 + qv[1] * ga.w;
                    const float fv2 = qa[2] * w1 + qv[2] * ga.w, fv3 = qa[3] * w1 + qv[3] * ga.w;
                    sx.x += fv0 * nlx; sx.y += fv1 * nlx; sx.z += fv2 * nlx; sx.w += fv3 * nlx;
                    sy.x += fv0 * nly; sy.y += fv1 * nly; sy.z += fv2 * nly; sy.w += fv3 * nly;
                }
            }
            const float ia = iarea[gc];
            sx = f4(sx.x * ia, sx.y * ia, sx.z * ia, sx.w * ia);
            sy = f4(sy.x * ia, sy.y * ia, sy.z * ia, sy.w * ia);
        } else {
            float rx[4] = {0.f, 0.f, 0.f, 0.f}, ry[4] = {0.f, 0.f, 0.f, 0.f};
            const float qa[4] = {qc.x, qc.y, qc.z, qc.w};
            float mn[4] = {qc.x, qc.y, qc.z, qc.w}, mx[4] = {qc.x, qc.y, qc.z, qc.w};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = slot_of(sw, s);
                if (e == 0xFFFFu) break;
                const int f = (int)(e & 0x7FFFu);
                const bool nb = (e >> 15) != 0;
                const uint32_t fc = fck[f];
                const int kind = (int)(fc >> 24);
                const int other = nb ? (int)(fc & 0xFFFu) : (int)((fc >> 12) & 0xFFFu);
                const float4 ga = geo[2 * f], gb = geo[2 * f + 1];
                const float dcx = nb ? gb.z : gb.x, dcy = nb ? gb.w : gb.y;
                const float dox = nb ? gb.x : gb.z, doy = nb ? gb.y : gb.w;
                dxs[s] = dcx; dys[s] = dcy;
                ns = s + 1;
                float qb[4] = {0.f, 0.f, 0.f, 0.f}, qv[4];
                const bool walls = (kind == K_WALL || kind == K_SLIP);
                if (!walls) { const float4 t4 = sq[other]; qb[0] = t4.x; qb[1] = t4.y; qb[2] = t4.z; qb[3] = t4.w; }
                face_vars<float>(g, kind, ga.x, ga.y, qa, qb, qv);
                const float dx = dcx - dox, dy = dcy - doy;
                for (int k = 0; k < 4; ++k) {
                    const float del = qa[k] - qv[k];
                    rx[k] += dx * del;
                    ry[k] += dy * del;
                    if (!walls) { mn[k] = fminf(mn[k], qb[k]); mx[k] = fmaxf(mx[k], qb[k]); }
                }
            }
            const float m0 = lsq[gc], m1 = lsq[(size_t)nC + gc], m2 = lsq[2 * (size_t)nC + gc], m3 = lsq[3 * (size_t)nC + gc];
            sx = f4(m0 * rx[0] + m1 * ry[0], m0 * rx[1] + m1 * ry[1], m0 * rx[2] + m1 * ry[2], m0 * rx[3] + m1 * ry[3]);
            sy = f4(m2 * rx[0] + m3 * ry[0], m2 * rx[1] + m3 * ry[1], m2 * rx[2] + m3 * ry[2], m2 * rx[3] + m3 * ry[3]);
            qmin = f4(mn[0], mn[1], mn[2], mn[3]);
            qmax = f4(mx[0], mx[1], mx[2], mx[3]);
        }
        sgx[c] = sx;
        sgy[c] = sy;
        if (second_order) {
            const float K3a = K3 * a15[gc];
            float4 lm = f4(1.f, 1.f, 1.f, 1.f);
            const float4 dM = f4(qmax.x - qc.x, qmax.y - qc.y, qmax.z - qc.z, qmax.w - qc.w);
            const float4 dm = f4(qmin.x - qc.x, qmin.y - qc.y, qmin.z - qc.z, qmin.w - qc.w);
            const float4 ddM = f4(dM.x * dM.x + K3a, dM.y * dM.y + K3a, dM.z * dM.z + K3a, dM.w * dM.w + K3a);
            const float4 ddm = f4(dm.x * dm.x + K3a, dm.y * dm.y + K3a, dm.z * dm.z + K3a, dm.w * dm.w + K3a);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s >= ns) break;
                const float dx = dxs[s], dy = dys[s];
                // D has the sign of dqg, so the denominator D^2 + K3a + dqg D
                // + 2 dqg^2 is positive whenever dqg != 0 (no NaN reaches fminf)
#define LIM(m) { const float dqg = sx.m * dx + sy.m * dy; const bool pos = dqg > 1e-16f, neg = dqg < -1e-16f; \
                 const float D = pos ? dM.m : dm.m, DD = pos ? ddM.m : ddm.m; const float tq = dqg * D; \
                 float phi = (DD + 2.f * tq) * frcp(DD + tq + 2.f * dqg * dqg); phi = (pos || neg) ? phi : 1.f; \
                 lm.m = fminf(lm.m, phi); }
                F4_LOOP(LIM)
#undef LIM
            }
            slm[c] = lm;
        }
    }
    __syncthreads();

    // 3. fluxes of the faces touching T
    for (int fl = threadIdx.x; fl < nFT; fl += blockDim.x) {
        const uint32_t fc = fck[fl];
        const int o = (int)(fc & 0xFFFu), n = (int)((fc >> 12) & 0xFFFu);
        const int kind = (int)(fc >> 24);
        const float4 ga = geo[2 * fl], gb = geo[2 * fl + 1];
        float4 qL = sq[o];
        float4 qR = (kind == K_INTERNAL || kind == K_FARFIELD) ? sq[n] : f4(0.f, 0.f, 0.f, 0.f);
        if (second_order) {
            const float4 ax = sgx[o], ay = sgy[o], al = slm[o];
#define RECL(m) qL.m += (ax.m * gb.x + ay.m * gb.y) * al.m;
            F4_LOOP(RECL)
#undef RECL
            if (n < nG) {
                const float4 bx = sgx[n], by = sgy[n], bl = slm[n];
#define RECR(m) qR.m += (bx.m * gb.z + by.m * gb.w) * bl.m;
                F4_LOOP(RECR)
#undef RECR
            }
        }
        float F[4];
        const float qLa[4] = {qL.x, qL.y, qL.z, qL.w};
        float qRa[4] = {qR.x, qR.y, qR.z, qR.w};
        if (kind == K_INTERNAL) {
            if (viscous == 1) {
                float gxa[4], gya[4];
                const int go_ = Lg[o], gn_ = Lg[n];
                float t0 = xc[go_] + xc[gn_], t1 = yc[go_] + yc[gn_];
                const float il = frcp(__builtin_amdgcn_sqrtf(t0 * t0 + t1 * t1));
                t0 *= il; t1 *= il;
                const float4 so = sq[o], sn = sq[n], xo = sgx[o], xn = sgx[n], yo = sgy[o], yn = sgy[n];
                for (int k = 0; k < 4; ++k) {
                    const float dir = (get(so, k) - get(sn, k)) * il;
                    const float gxb = 0.5f * (get(xo, k) + get(xn, k));
                    const float gyb = 0.5f * (get(yo, k) + get(yn, k));
                    const float dot = gxb * t0 + gyb * t1;
                    gxa[k] = gxb - (dot - dir) * t0;
                    gya[k] = gyb - (dot - dir) * t1;
                }
                roe_flux<float>(g, ga.x, ga.y, qLa, qRa, true, gxa, gya, F);
            } else {
                roe_flux_f32(g.gamma, ga.x, ga.y, qLa, qRa, F);
            }
        } else {
            float qb[4];
            face_vars<float>(g, kind, ga.x, ga.y, qLa, qRa, qb);
            roe_flux_f32(g.gamma, ga.x, ga.y, qLa, qb, F);
        }
        sfl[fl] = f4(F[0] * ga.z, F[1] * ga.z, F[2] * ga.z, F[3] * ga.z);
    }
    __syncthreads();

    // 4. signed sums in reference face order (exact FP64 accumulation of the FP32 fluxes)
    for (int c = threadIdx.x; c < nTt; c += blockDim.x) {
        const uint2 sw = slots2[c];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t e = slot_of(sw, s);
            if (e == 0xFFFFu) break;
            const int fl = (int)(e & 0x7FFFu);
            const float4 v = sfl[fl];
            const bool nb = (e >> 15) != 0;
            a0 += (double)(nb ? v.x : -v.x);
            a1 += (double)(nb ? v.y : -v.y);
            a2 += (double)(nb ? v.z : -v.z);
            a3 += (double)(nb ? v.w : -v.w);
            if (qghost) {
                const uint32_t fc = fck[fl];
                const int kind = (int)(fc >> 24);
                if (kind == K_WALL || kind == K_SLIP) qghost[-1 - Lg[(int)((fc >> 12) & 0xFFFu)]] = sq[c];
            }
        }
        R[c0g + c] = f4((float)a0, (float)a1, (float)a2, (float)a3);
    }
}

hipError_t launch_residual_tile_f32(hipStream_t s, TilesF32 S, const float4* q, int32_t nC, const float* iarea,
                                    const float* a15, const float* lsq, const float* xc, const float* yc,
                                    GasT<float> g, int gg, int second_order, float K3, int viscous, float4* R,
                                    float4* qghost) {
    const int nb = ((S.ntiles + 7) / 8) * 8;
    const size_t lds = tile_lds_bytes<float>(S.lmax, S.gmax, S.ftmax);
    hipLaunchKernelGGL(k_residual_tile_f32, dim3(nb), dim3(kBlock), lds, s, S, q, nC, iarea, a15, lsq, xc, yc, g, gg,
                       second_order, K3, viscous, R, qghost);
    return hipGetLastError();
}
#endif  // AFX_FP32_UNIT

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
    hipError_t launch_residual_tile<T>(hipStream_t s, TilesT<T> S, const T* q, int32_t nT, int32_t nC,             \
                                       const T* area, const T* lsq, const T* xc, const T* yc, GasT<T> g, int gg,   \
                                       int so, T lk, int viscous, T* R, T* qghost) {                               \
        const int nb = ((S.ntiles + 7) / 8) * 8;                                                                  \
        const size_t lds = tile_lds_bytes<T>(S.lmax, S.gmax, S.ftmax);                                            \
        hipLaunchKernelGGL(k_residual_tile<T>, dim3(nb), dim3(kBlock), lds, s, S, q, nT, nC, area, lsq, xc, yc, g, \
                           gg, so, lk, viscous, R, qghost);                                                       \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_gather<T>(hipStream_t s, CellsT<T> C, const T* flux, int32_t nF, int mode, T* R) {          \
        hipLaunchKernelGGL(k_gather<T>, dim3(nblocks(C.nC)), dim3(kBlock), 0, s, C, flux, nF, mode, R);           \
        return hipGetLastError();                                                                                 \
    }

}  // namespace afx
