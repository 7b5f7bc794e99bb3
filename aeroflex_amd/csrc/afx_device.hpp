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

// physics.h:583-592: michalak_limiter with RANS_YT = 2, the cubic
// a y^3 + b y^2 + y joining phi(y_t) = 1 with zero slope (a = 0, b = -1/4)
template <typename T>
__device__ __forceinline__ T michalak(T y) {
    constexpr double yt = 2.0;
    constexpr double a = 1.0 / (yt * yt) - 2.0 / (yt * yt * yt);
    constexpr double b = -3.0 / 2.0 * a * yt - 0.5 / yt;
    return y >= T(yt) ? T(1.0) : T(a) * y * y * y + T(b) * y * y + y;
}

// solver.h:557-573, the RANS_MICHALAK_LIMITER build: a smooth switch sig
// from 1 (range of the neighbours' values below K3a) to 0 (above 2 K3a)
// blends the unlimited value with the Michalak limiter of the selected bound
template <typename T>
__device__ __forceinline__ T limiter_michalak(T dqg, T dmax, T dmin, T K3a) {
    const T dd = (dmax - dmin) * (dmax - dmin);
    T sig;
    if (dd <= K3a) {
        sig = T(1.0);
    } else if (dd <= 2 * K3a) {
        const T y = ddiv<T>(dd, K3a) - T(1.0);
        sig = T(2.0) * y * y * y - T(3.0) * y * y + T(1.0);
    } else {
        sig = T(0.0);
    }
    T lim = T(1.0);
    if (sig < T(1.0)) {
        if (dqg > T(1e-14)) lim = michalak<T>(ddiv<T>(dmax, dqg));
        else if (dqg < T(-1e-14)) lim = michalak<T>(ddiv<T>(dmin, dqg));
    }
    return sig + (T(1.0) - sig) * lim;
}

// the limiter of the selected build: second_order bit 1 set = Michalak
template <typename T>
__device__ __forceinline__ T limiter_value(int second_order, T dqg, T dmax, T dmin, T K3a) {
    return (second_order & 2) ? limiter_michalak<T>(dqg, dmax, dmin, K3a) : limiter_phi<T>(dqg, dmax, dmin, K3a);
}

// physics.h:230-257: stresses and heat flux at the face from the face-averaged
// state and the face gradients, with viscosity mu and conductivity kc as
// arguments (laminar: mu_L, cp mu_L/Pr_L; SA: mu_L + mu_t, cp (mu_L/Pr_L + mu_t/Pr_T))
template <typename T>
__device__ __forceinline__ void viscous_terms(const GasT<T>& g, T nx, T ny, const T* q_L, const T* q_R, const T* gx,
                                              const T* gy, T mu, T kc, T& f1, T& f2, T& f3) {
    const T gm1 = g.gamma - T(1);
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
    const T tau_xx = T(2.) * mu * (gu0 - ddiv<T>(div_v, T(3.)));
    const T tau_yy = T(2.) * mu * (gv1 - ddiv<T>(div_v, T(3.)));
    const T tau_xy = mu * (gu1 + gv0);
    const T phi0 = ddiv<T>(qc[1], qc[0]) * tau_xx + ddiv<T>(qc[2], qc[0]) * tau_xy + kc * gT0;
    const T phi1 = ddiv<T>(qc[1], qc[0]) * tau_xy + ddiv<T>(qc[2], qc[0]) * tau_yy + kc * gT1;
    f1 -= nx * tau_xx + ny * tau_xy;
    f2 -= nx * tau_xy + ny * tau_yy;
    f3 -= nx * phi0 + ny * phi1;
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
    if (laminar) viscous_terms<T>(g, nx, ny, q_L, q_R, gx, gy, g.mu, g.k, f1, f2, f3);
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

// ------------------------------------------------------ Spalart-Allmaras --
// The reference's SA branch is empty (physics.h:259-261).  Model: SA-noft2 of
// Spalart & Allmaras (1994) with the S-tilde limiter and negative-nu~ branch
// of SA-neg (Allmaras, Johnson & Spalart 2012), rho nu~ as fifth conservative
// variable.  Every expression keeps the operation order of the FP64 oracle so
// the plain form agrees with it to rounding (cbrt aside, bit for bit).
namespace sa {
constexpr double cb1 = 0.1355, cb2 = 0.622, sigma = 2.0 / 3.0, kappa = 0.41, cv1 = 7.1, cw2 = 0.3, cw3 = 2.0;
constexpr double cv2 = 0.7, cv3 = 0.9, ct3 = 1.2, cn1 = 16.0;
constexpr double cw1 = cb1 / (kappa * kappa) + (1.0 + cb2) / sigma;
constexpr double cv1_3 = cv1 * cv1 * cv1;
constexpr double cw3_6 = (cw3 * cw3) * (cw3 * cw3) * (cw3 * cw3);
}  // namespace sa

template <typename T> __device__ __forceinline__ T dcbrt(T x);
template <> __device__ __forceinline__ double dcbrt<double>(double x) { return cbrt(x); }
template <> __device__ __forceinline__ float dcbrt<float>(float x) { return cbrtf(x); }

// f_v1 (zero for chi <= 0: no eddy viscosity from negative nu~)
template <typename T>
__device__ __forceinline__ T sa_fv1(T chi) {
    const T c3 = chi * chi * chi;
    return chi <= T(0) ? T(0) : ddiv<T>(c3, c3 + T(sa::cv1_3));
}
// f_n of the diffusion coefficient (1 for chi >= 0)
template <typename T>
__device__ __forceinline__ T sa_fn(T chi) {
    const T c3 = chi * chi * chi;
    return chi >= T(0) ? T(1) : ddiv<T>(T(sa::cn1) + c3, T(sa::cn1) - c3);
}
template <typename T>
__device__ __forceinline__ T sa_fw(T r) {
    const T r2 = r * r;
    const T gg = r + T(sa::cw2) * (r2 * r2 * r2 - r);
    const T g2 = gg * gg;
    return gg * dcbrt<T>(dsqrt<T>(ddiv<T>(T(1) + T(sa::cw3_6), g2 * g2 * g2 + T(sa::cw3_6))));
}
// per-volume source of rho nu~: rho (P - D) + c_b2/sigma rho |grad nu~|^2
template <typename T>
__device__ __forceinline__ T sa_source(T rho, T nut, T mu, T omega, T d, T gvx, T gvy) {
    const T cross = T(sa::cb2 / sa::sigma) * (gvx * gvx + gvy * gvy);
    const T nd = ddiv<T>(nut, d);
    if (nut < T(0)) return rho * (T(sa::cb1 * (1.0 - sa::ct3)) * omega * nut + T(sa::cw1) * nd * nd + cross);
    const T chi = ddiv<T>(rho * nut, mu);
    const T f1 = sa_fv1<T>(chi);
    const T f2 = T(1) - ddiv<T>(chi, T(1) + chi * f1);
    const T kd2 = T(sa::kappa * sa::kappa) * d * d;
    const T sbar = ddiv<T>(nut * f2, kd2);
    const T st = (sbar >= -T(sa::cv2) * omega)
                     ? omega + sbar
                     : omega + ddiv<T>(omega * (T(sa::cv2 * sa::cv2) * omega + T(sa::cv3) * sbar),
                                       T(sa::cv3 - 2.0 * sa::cv2) * omega - sbar);
    T r = T(10);
    if (st > T(0)) {
        r = ddiv<T>(nut, st * kd2);
        r = r < T(10) ? r : T(10);
    }
    const T P = T(sa::cb1) * st * nut;
    const T D = T(sa::cw1) * sa_fw<T>(r) * nd * nd;
    return rho * (P - D + cross);
}
// rho nu~ on the far side of a boundary face: no-slip wall -rho nu~ (nu~ = 0
// at the wall), slip wall the interior value, farfield the ghost nu~ on inflow
// and the interior nu~ on outflow, times the boundary density
template <typename T>
__device__ __forceinline__ T sa_bnd_nu(int kind, T nx, T ny, const T* qL, T nL, const T* qR, T nR, const T* qb) {
    if (kind == K_WALL) return -nL;
    if (kind == K_SLIP) return nL;
    const T io = qL[1] * nx + qL[2] * ny;
    return qb[0] * (io < T(0) ? ddiv<T>(nR, qR[0]) : ddiv<T>(nL, qL[0]));
}
// the diffusive part of an interior face's five-variable flux, subtracted
// from f[1..4]: the mean-flow viscous terms with mu_L + mu_t and
// k = cp (mu_L/Pr_L + mu_t/Pr_T), and the nu~ diffusion (mu_L + rho nu~ f_n) /
// sigma grad nu~ . n at the face gradients gx, gy of the 5 variables
template <typename T>
__device__ __forceinline__ void sa_face_diffusion(const GasT<T>& g, T nx, T ny, const T* qL, const T* qR, const T* gx,
                                                  const T* gy, T* f) {
    const T nc = T(0.5) * (qL[4] + qR[4]);
    const T chi = ddiv<T>(nc, g.mu);
    const T mut = nc * sa_fv1<T>(chi);
    const T mu = g.mu + mut;
    const T kc = g.cp * (ddiv<T>(g.mu, g.PrL) + ddiv<T>(mut, g.PrT));
    T f1 = f[1], f2 = f[2], f3 = f[3];
    viscous_terms<T>(g, nx, ny, qL, qR, gx, gy, mu, kc, f1, f2, f3);
    f[1] = f1; f[2] = f2; f[3] = f3;
    const T qc0 = T(0.5) * (qL[0] + qR[0]);
    const T nut = ddiv<T>(nc, qc0);
    const T gvx = ddiv<T>(gx[4] - nut * gx[0], qc0), gvy = ddiv<T>(gy[4] - nut * gy[0], qc0);
    const T dcoef = ddiv<T>(g.mu + nc * sa_fn<T>(chi), T(sa::sigma));
    f[4] -= dcoef * (nx * gvx + ny * gvy);
}

// five-variable face flux: Roe on the mean flow, upwind convection of rho nu~
// by the sign of the mass flux; internal faces add the viscous terms with
// mu_L + mu_t and the nu~ diffusion; a no-slip wall face adds the diffusion
// (mu_L/sigma) (0 - nu~_c) / d_n of the nu~ = 0 wall condition (nut_c: the
// owner's cell-centre nu~, dn: its normal distance to the face).  gx/gy: face
// gradients of the 5 variables.
template <typename T>
__device__ __forceinline__ void sa_face_flux(const GasT<T>& g, int kind, T nx, T ny, const T* qL, const T* qR,
                                             const T* gx, const T* gy, T nut_c, T dn, T* f) {
    if (kind != K_INTERNAL) {
        T qb[4];
        face_vars<T>(g, kind, nx, ny, qL, qR, qb);
        const T nb = sa_bnd_nu<T>(kind, nx, ny, qL, qL[4], qR, qR[4], qb);
        roe_flux<T>(g, nx, ny, qL, qb, false, gx, gy, f);
        f[4] = f[0] * (f[0] >= T(0) ? ddiv<T>(qL[4], qL[0]) : ddiv<T>(nb, qb[0]));
        if (kind == K_WALL) f[4] -= ddiv<T>(g.mu, T(sa::sigma)) * ddiv<T>(T(0) - nut_c, dn);
        return;
    }
    roe_flux<T>(g, nx, ny, qL, qR, false, gx, gy, f);
    f[4] = f[0] * (f[0] >= T(0) ? ddiv<T>(qL[4], qL[0]) : ddiv<T>(qR[4], qR[0]));
    sa_face_diffusion<T>(g, nx, ny, qL, qR, gx, gy, f);
}
// source of cell with state q (5), unlimited gradients gx/gy (5) and wall
// distance d: vorticity from the momentum gradients, grad nu~ from grad(rho nu~)
template <typename T>
__device__ __forceinline__ T sa_cell_source(const GasT<T>& g, const T* q, const T* gx, const T* gy, T d) {
    const T rho = q[0], nut = ddiv<T>(q[4], rho);
    const T dudy = ddiv<T>(gy[1] - ddiv<T>(q[1], rho) * gy[0], rho);
    const T dvdx = ddiv<T>(gx[2] - ddiv<T>(q[2], rho) * gx[0], rho);
    const T omega = dabs<T>(dvdx - dudy);
    const T gvx = ddiv<T>(gx[4] - nut * gx[0], rho), gvy = ddiv<T>(gy[4] - nut * gy[0], rho);
    return sa_source<T>(rho, nut, g.mu, omega, d, gvx, gvy);
}

// viscous spectral-radius coefficient of a face for the explicit local time
// step of the SA model, max(4/3, gamma) (mu_L/Pr_L + mu_t/Pr_T) / rho (Blazek,
// Computational Fluid Dynamics, eq. 6.22) from the face-averaged cell states
// (owner state on boundary faces); the cell's viscous radius is the sum of
// coef len^2 / A and enters with weight 4.  q: planes of nT cells.
template <typename T>
__device__ __forceinline__ T sa_viscous_radius(const GasT<T>& g, const T* q, size_t nT, int o, int n, bool internal) {
    T rho = q[o], nuf = q[4 * nT + o];
    if (internal) {
        rho = T(0.5) * (rho + q[n]);
        nuf = T(0.5) * (nuf + q[4 * nT + n]);
    }
    const T mut = nuf * sa_fv1<T>(ddiv<T>(nuf, g.mu));
    const T kmax = g.gamma > T(4.0 / 3.0) ? g.gamma : T(4.0 / 3.0);
    return ddiv<T>(kmax * (ddiv<T>(g.mu, g.PrL) + ddiv<T>(mut, g.PrT)), rho);
}

// ---------------------------------------------------------------- kernels --
constexpr int kBlock = 256;

// XCD-aware block -> tile map: the 8 XCDs (block b runs on XCD b % 8) each
// take a contiguous range of tiles so that halo cells shared by neighbouring
// tiles are found in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int per = nblocks >> 3;   // nblocks is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

// The same order for one-thread-per-cell (or per-face) loops: launched with
// xcd_grid(n) blocks of kBlock threads, the thread's item is xcd_index().
// Each XCD then sweeps a contiguous range of cells or faces, so the
// neighbours' data that a block gathers was mostly fetched into that XCD's L2
// by the blocks before it, instead of by another XCD.
__device__ __forceinline__ int xcd_index() {
    return xcd_tile((int)blockIdx.x, (int)gridDim.x) * (int)blockDim.x + (int)threadIdx.x;
}
inline int xcd_grid(int n) { return (((n + kBlock - 1) / kBlock) + 7) & ~7; }

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
                                               int sa, T* __restrict__ dt) {
    int c = xcd_index();
    if (c >= C.nT) return;
    const size_t nT = (size_t)C.nT;
    T s = T(0), sv = T(0);
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
            if (sa) sv += sa_viscous_radius<T>(g, q, nT, o, n, F.kind[f] == K_INTERNAL) * F.len[f] * F.len[f];
        }
    }
    dt[c] = sa ? ddiv<T>(cfl * C.area[c], s + T(4) * ddiv<T>(sv, C.area[c])) : ddiv<T>(cfl * C.area[c], s);
}

// solver.h:425-514 (gradients from qg) fused with solver.h:517-593 (limiters from ql)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_grad_lim(CellsT<T> C, FacesT<T> F, const T* __restrict__ qg,
                                                     const T* __restrict__ ql, GasT<T> g, int gg, int second_order,
                                                     T limiter_k, T* __restrict__ gx, T* __restrict__ gy,
                                                     T* __restrict__ lim) {
    int c = xcd_index();
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
            const T l = limiter_value<T>(second_order, dqg, dmax, dmin, K3a);
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
    int f = xcd_index();
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
    int c = xcd_index();
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



// BND: the tile's face list holds boundary faces; without, every face is
// taken as interior at compile time (no kind tests, no boundary states, no
// ghost writes) - the same operations on the same values, so the same bits
template <typename T, int NV, bool BND>
__device__ __forceinline__ void residual_tile_body(const TilesT<T>& S, const T* __restrict__ q, int32_t nT, int32_t nC,
                                                   const T* __restrict__ area, const T* __restrict__ lsq,
                                                   const T* __restrict__ xc, const T* __restrict__ yc,
                                                   const T* __restrict__ wd, const GasT<T>& g, int gg,
                                                   int second_order, T limiter_k, int viscous,
                                                   T* __restrict__ R, T* __restrict__ qghost, const int t,
                                                   unsigned char* smem_raw) {
    constexpr bool SA = (NV == 5);
    const int c0g = S.cell0[t], nTt = S.cell0[t + 1] - c0g;
    const int lo = S.loff[t], nL = S.loff[t + 1] - lo;
    const int nG = S.ng[t];
    const int fo = S.foff[t], nFT = S.nft[t];
    const int go = S.goff[t];
    const int LM = S.lmax, GM = S.gmax, FM = S.ftmax;
    const int64_t NF = S.nFtot;
    T* sq = reinterpret_cast<T*>(smem_raw);   // [NV][LM] states of L
    T* sgx = sq + NV * LM;                     // [NV][GM]
    T* sgy = sgx + NV * GM;
    T* slm = sgy + NV * GM;
    T* sfl = slm + NV * GM;                    // [NV][FM] face fluxes x length
    const int* Lg = S.Lglob + lo;

    // 1. stage the states of the local cells
    for (int i = threadIdx.x; i < nL; i += blockDim.x) {
        const int gi = Lg[i];
        if (gi < 0) continue;   // wall/slip ghost: never read
        for (int k = 0; k < NV; ++k) sq[k * LM + i] = q[(size_t)k * nT + gi];
    }
    __syncthreads();

    // 2. gradients (solver.h:425-514) and limiters (solver.h:517-593) of T u H1
    for (int c = threadIdx.x; c < nG; c += blockDim.x) {
        const uint64_t sw = S.slots[go + c];
        T qc[NV];
        for (int k = 0; k < NV; ++k) qc[k] = sq[k * LM + c];
        T sx[NV], sy[NV];
        for (int k = 0; k < NV; ++k) { sx[k] = T(0); sy[k] = T(0); }
        const int gc = Lg[c];
        if (gg) {
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const bool nb_side = (e >> 15) != 0;
                const uint32_t fc = S.fcell[f];
                const int o = (int)(fc & 0xFFFFu), n = (int)(fc >> 16);
                T qo[4], qn[4], qv[NV];
                for (int k = 0; k < 4; ++k) { qo[k] = sq[k * LM + o]; qn[k] = sq[k * LM + n]; }
                const T nx = S.geo[f], ny = S.geo[NF + f], L = S.geo[2 * NF + f], gf = S.geo[3 * NF + f];
                const int kind = BND ? (int)S.fkind[f] : (int)K_INTERNAL;
                face_vars<T>(g, kind, nx, ny, qo, qn, qv);
                T vo[NV];
                for (int k = 0; k < 4; ++k) vo[k] = qo[k];
                if constexpr (SA) {
                    const T no = sq[4 * LM + o], nn = sq[4 * LM + n];
                    vo[4] = no;
                    qv[4] = (kind == K_INTERNAL) ? (no + nn) * T(0.5) : sa_bnd_nu<T>(kind, nx, ny, qo, no, qn, nn, qv);
                }
                for (int k = 0; k < NV; ++k) {
                    const T fk = (vo[k] * (T(1.0) - gf) + qv[k] * gf) * L;
                    if (!nb_side) { sx[k] += fk * nx; sy[k] += fk * ny; }
                    else { sx[k] -= fk * nx; sy[k] -= fk * ny; }
                }
            }
            const T A = area[gc];
            for (int k = 0; k < NV; ++k) { sx[k] = ddiv<T>(sx[k], A); sy[k] = ddiv<T>(sy[k], A); }
        } else {
            T rx[NV], ry[NV];
            for (int k = 0; k < NV; ++k) { rx[k] = T(0); ry[k] = T(0); }
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
                const int kind = BND ? (int)S.fkind[f] : (int)K_INTERNAL;
                T qn[4], qv[NV];
                for (int k = 0; k < 4; ++k) qn[k] = sq[k * LM + other];
                face_vars<T>(g, kind, S.geo[f], S.geo[NF + f], qc, qn, qv);
                if constexpr (SA) {
                    const T nn = sq[4 * LM + other];
                    qv[4] = (kind == K_INTERNAL) ? (qc[4] + nn) * T(0.5)
                                                 : sa_bnd_nu<T>(kind, S.geo[f], S.geo[NF + f], qc, qc[4], qn, nn, qv);
                }
                for (int k = 0; k < NV; ++k) {
                    const T del = qc[k] - qv[k];
                    rx[k] += dx * del;
                    ry[k] += dy * del;
                }
            }
            const T m0 = lsq[gc], m1 = lsq[(size_t)nC + gc], m2 = lsq[2 * (size_t)nC + gc], m3 = lsq[3 * (size_t)nC + gc];
            for (int k = 0; k < NV; ++k) {
                sx[k] = m0 * rx[k] + m1 * ry[k];
                sy[k] = m2 * rx[k] + m3 * ry[k];
            }
        }
        for (int k = 0; k < NV; ++k) { sgx[k * GM + c] = sx[k]; sgy[k * GM + c] = sy[k]; }
        if (second_order) {
            T qmin[NV], qmax[NV];
            for (int k = 0; k < NV; ++k) { qmin[k] = qc[k]; qmax[k] = qc[k]; }
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const int kind = BND ? (int)S.fkind[f] : (int)K_INTERNAL;
                // a wall/slip ghost holds the owner's state (set_walls_from_internal)
                if (kind == K_WALL || kind == K_SLIP) continue;
                const uint32_t fc = S.fcell[f];
                const int other = (e >> 15) ? (int)(fc & 0xFFFFu) : (int)(fc >> 16);
                for (int k = 0; k < NV; ++k) {
                    const T v = sq[k * LM + other];
                    qmin[k] = (v < qmin[k]) ? v : qmin[k];
                    qmax[k] = (qmax[k] < v) ? v : qmax[k];
                }
            }
            const T Ka = limiter_k * dsqrt<T>(area[gc]);
            const T K3a = Ka * Ka * Ka;
            T lm[NV];
            for (int k = 0; k < NV; ++k) lm[k] = T(1);
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
                if (e == 0xFFFFu) break;
                const int f = fo + (int)(e & 0x7FFFu);
                const bool nb_side = (e >> 15) != 0;
                const T dx = nb_side ? S.geo[6 * NF + f] : S.geo[4 * NF + f];
                const T dy = nb_side ? S.geo[7 * NF + f] : S.geo[5 * NF + f];
                for (int k = 0; k < NV; ++k) {
                    const T dqg = sx[k] * dx + sy[k] * dy;
                    const T dmax = qmax[k] - qc[k];
                    const T dmin = qmin[k] - qc[k];
                    const T l = limiter_value<T>(second_order, dqg, dmax, dmin, K3a);
                    lm[k] = (l < lm[k]) ? l : lm[k];
                }
            }
            for (int k = 0; k < NV; ++k) slm[k * GM + c] = lm[k];
        }
    }
    __syncthreads();

    // 3. fluxes of the faces touching T (solver.h:1097-1134)
    for (int fl = threadIdx.x; fl < nFT; fl += blockDim.x) {
        const int f = fo + fl;
        const uint32_t fc = S.fcell[f];
        const int o = (int)(fc & 0xFFFFu), n = (int)(fc >> 16);
        const int kind = BND ? (int)S.fkind[f] : (int)K_INTERNAL;
        const T nx = S.geo[f], ny = S.geo[NF + f], L = S.geo[2 * NF + f];
        T qL[NV], qR[NV];
        for (int k = 0; k < NV; ++k) { qL[k] = sq[k * LM + o]; qR[k] = sq[k * LM + n]; }
        if (second_order) {
            const T d0x = S.geo[4 * NF + f], d0y = S.geo[5 * NF + f];
            for (int k = 0; k < NV; ++k)
                qL[k] = qL[k] + (sgx[k * GM + o] * d0x + sgy[k * GM + o] * d0y) * slm[k * GM + o];
            if (!BND || n < nG) {
                const T d1x = S.geo[6 * NF + f], d1y = S.geo[7 * NF + f];
                for (int k = 0; k < NV; ++k)
                    qR[k] = qR[k] + (sgx[k * GM + n] * d1x + sgy[k * GM + n] * d1y) * slm[k * GM + n];
            }
        }
        const bool visc = (viscous >= 1) && (kind == K_INTERNAL);
        T gxa[NV], gya[NV];
        for (int k = 0; k < NV; ++k) { gxa[k] = T(0); gya[k] = T(0); }
        if (visc) {
            // solver.h:359-398: direction from the SUM of the two centroids
            // (reference, A1); the SA model takes the textbook direction
            // x_n - x_o and derivative (q_n - q_o) / |x_n - x_o|
            const int go_ = Lg[o], gn_ = Lg[n];
            T t0 = SA ? xc[gn_] - xc[go_] : xc[go_] + xc[gn_];
            T t1 = SA ? yc[gn_] - yc[go_] : yc[go_] + yc[gn_];
            const T l = dsqrt<T>(t0 * t0 + t1 * t1);
            t0 = ddiv<T>(t0, l); t1 = ddiv<T>(t1, l);
            for (int k = 0; k < NV; ++k) {
                const T dir = SA ? ddiv<T>(sq[k * LM + n] - sq[k * LM + o], l) : ddiv<T>(sq[k * LM + o] - sq[k * LM + n], l);
                const T gxb = (sgx[k * GM + o] + sgx[k * GM + n]) * T(0.5);
                const T gyb = (sgy[k * GM + o] + sgy[k * GM + n]) * T(0.5);
                const T dot = gxb * t0 + gyb * t1;
                gxa[k] = gxb - (dot - dir) * t0;
                gya[k] = gyb - (dot - dir) * t1;
            }
        }
        T F[NV];
        if constexpr (SA) {
            const T dn = S.geo[4 * NF + f] * nx + S.geo[5 * NF + f] * ny;
            const T nut_c = ddiv<T>(sq[4 * LM + o], sq[o]);
            sa_face_flux<T>(g, kind, nx, ny, qL, qR, gxa, gya, nut_c, dn, F);
        } else {
            face_flux<T>(g, kind, nx, ny, qL, qR, visc, gxa, gya, F);
        }
        for (int k = 0; k < NV; ++k) sfl[k * FM + fl] = F[k] * L;
    }
    __syncthreads();

    // 4. signed sums in reference face order; one coalesced write per cell
    for (int c = threadIdx.x; c < nTt; c += blockDim.x) {
        const uint64_t sw = S.slots[go + c];
        Acc<T> acc[NV];
        for (int s = 0; s < 4; ++s) {
            const uint32_t e = (uint32_t)(sw >> (16 * s)) & 0xFFFFu;
            if (e == 0xFFFFu) break;
            const int fl = (int)(e & 0x7FFFu);
            if (e >> 15) for (int k = 0; k < NV; ++k) acc[k].add(sfl[k * FM + fl]);
            else for (int k = 0; k < NV; ++k) acc[k].sub(sfl[k * FM + fl]);
            if (BND && qghost) {
                const int kind = S.fkind[fo + fl];
                if (kind == K_WALL || kind == K_SLIP) {
                    const int gh = -1 - Lg[(int)(S.fcell[fo + fl] >> 16)];
                    for (int k = 0; k < NV; ++k) qghost[(size_t)k * nT + gh] = sq[k * LM + c];
                }
            }
        }
        for (int k = 0; k < 4; ++k) R[(size_t)k * nT + c0g + c] = acc[k].value();
        if constexpr (SA) {
            T qs[NV], gxs[NV], gys[NV];
            for (int k = 0; k < NV; ++k) { qs[k] = sq[k * LM + c]; gxs[k] = sgx[k * GM + c]; gys[k] = sgy[k * GM + c]; }
            const int gc = c0g + c;
            R[(size_t)4 * nT + gc] = acc[4].value() + area[gc] * sa_cell_source<T>(g, qs, gxs, gys, wd[gc]);
        }
    }
}

template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_residual_tile(TilesT<T> S, const T* __restrict__ q, int32_t nT, int32_t nC,
                                                          const T* __restrict__ area, const T* __restrict__ lsq,
                                                          const T* __restrict__ xc, const T* __restrict__ yc,
                                                          const T* __restrict__ wd, GasT<T> g, int gg,
                                                          int second_order, T limiter_k, int viscous,
                                                          T* __restrict__ R, T* __restrict__ qghost) {
    extern __shared__ unsigned char smem_raw[];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= S.ntiles) return;
    if (S.nbf && S.nbf[t] == 0)
        residual_tile_body<T, NV, false>(S, q, nT, nC, area, lsq, xc, yc, wd, g, gg, second_order, limiter_k, viscous, R,
                                         qghost, t, smem_raw);
    else
        residual_tile_body<T, NV, true>(S, q, nT, nC, area, lsq, xc, yc, wd, g, gg, second_order, limiter_k, viscous, R,
                                        qghost, t, smem_raw);
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

// Roe flux of physics.h:160-228 (FP32), times 2 sc: sc = 1/2 gives the flux,
// sc = len/2 the flux times the face length.  Reciprocals and roots share
// reciprocal square roots: 1/rho = rsq(rho)^2, sqrt(rho) = rho rsq(rho),
// c = c^2 rsq(c^2), 1/c^2 = rsq(c^2)^2, 1/(2 d) = 10 rsq(c^2) (d = c/20).
__device__ __forceinline__ void roe_flux_f32(float gamma, float nx, float ny, const float* qL, const float* qR,
                                             float* f, float sc = 0.5f) {
    const float gm1 = gamma - 1.f;
    const float rL = qL[0], ruL = qL[1], rvL = qL[2], reL = qL[3];
    const float rR = qR[0], ruR = qR[1], rvR = qR[2], reR = qR[3];
    const float rsL = __builtin_amdgcn_rsqf(rL), rsR = __builtin_amdgcn_rsqf(rR);
    const float irL = rsL * rsL, irR = rsR * rsR;
    const float uL = ruL * irL, vL = rvL * irL, uR = ruR * irR, vR = rvR * irR;
    const float VL = uL * nx + vL * ny, VR = uR * nx + vR * ny;
    const float pL = gm1 * (reL - 0.5f * (ruL * uL + rvL * vL));
    const float pR = gm1 * (reR - 0.5f * (ruR * uR + rvR * vR));
    const float hL = reL + pL, hR = reR + pR;   // rho H per side
    // central part (twice the average of the two sides' fluxes)
    const float ps = pL + pR;
    float f0 = VL * rL + VR * rR;
    float f1 = VL * ruL + VR * ruR + ps * nx;
    float f2 = VL * rvL + VR * rvR + ps * ny;
    float f3 = VL * hL + VR * hR;
    const float srL = rL * rsL, srR = rR * rsR;
    const float is = frcp(srL + srR);
    const float wL = srL * is, wR = srR * is;
    const float rho = srL * srR;
    const float u = uL * wL + uR * wR;
    const float v = vL * wL + vR * wR;
    const float h = (hL * irL) * wL + (hR * irR) * wR;
    const float q2 = u * u + v * v;
    const float c2 = gm1 * (h - 0.5f * q2);
    const float rc = __builtin_amdgcn_rsqf(c2);
    const float c = c2 * rc, ic2 = rc * rc;
    const float V = u * nx + v * ny;
    const float d = 0.05f * c;
    const float i2d = 10.f * rc, dd = d * d;
    float lcm = __builtin_amdgcn_sqrtf((V - c) * (V - c) + 1e-4f);
    float lc = __builtin_amdgcn_sqrtf(V * V + 1e-4f);
    float lcp = __builtin_amdgcn_sqrtf((V + c) * (V + c) + 1e-4f);
    lcm = lcm > d ? lcm : (lcm * lcm + dd) * i2d;
    lc = lc > d ? lc : (lc * lc + dd) * i2d;
    lcp = lcp > d ? lcp : (lcp * lcp + dd) * i2d;
    const float dp = pR - pL, dV = VR - VL;
    const float rcdV = rho * c * dV;
    const float hic2 = 0.5f * ic2;
    const float kF1 = lcm * (dp - rcdV) * hic2;
    const float kF5 = lcp * (dp + rcdV) * hic2;
    const float k0 = lc * ((rR - rL) - dp * ic2);
    const float k1 = lc * rho;
    const float du = uR - uL, dv = vR - vL;
    const float cnx = c * nx, cny = c * ny, cV = c * V;
    f0 -= kF1 + k0 + kF5;
    f1 -= kF1 * (u - cnx) + k0 * u + k1 * (du - dV * nx) + kF5 * (u + cnx);
    f2 -= kF1 * (v - cny) + k0 * v + k1 * (dv - dV * ny) + kF5 * (v + cny);
    f3 -= kF1 * (h - cV) + k0 * q2 * 0.5f + k1 * (u * du + v * dv - V * dV) + kF5 * (h + cV);
    f[0] = sc * f0; f[1] = sc * f1; f[2] = sc * f2; f[3] = sc * f3;
}

// physics.h:230-257 in FP32 with one reciprocal of the face density: the
// reference's pressure / temperature / velocity gradients (physics.h:33-82)
// in the algebraically equal forms
//   grad p = (gamma-1) (grad E - u grad(rho u) - v grad(rho v) + |u|^2/2 grad rho)
//   grad T = (grad p - (p/rho) grad rho) / (rho R),  grad u = (grad(rho u) - u grad rho) / rho
__device__ __forceinline__ void viscous_terms_f32(float gm1, float iR, float nx, float ny, const float* qL,
                                                  const float* qR, const float* gx, const float* gy, float mu, float kc,
                                                  float& f1, float& f2, float& f3) {
    const float r = 0.5f * (qL[0] + qR[0]), ru = 0.5f * (qL[1] + qR[1]);
    const float rv = 0.5f * (qL[2] + qR[2]), re = 0.5f * (qL[3] + qR[3]);
    const float ir = frcp(r);
    const float u = ru * ir, v = rv * ir;
    const float ke = 0.5f * (u * u + v * v);
    const float pr = gm1 * (re * ir - ke);   // p / rho
    const float gpx = gm1 * (gx[3] - u * gx[1] - v * gx[2] + ke * gx[0]);
    const float gpy = gm1 * (gy[3] - u * gy[1] - v * gy[2] + ke * gy[0]);
    const float cT = ir * iR;
    const float gTx = (gpx - pr * gx[0]) * cT, gTy = (gpy - pr * gy[0]) * cT;
    const float gu0 = (gx[1] - u * gx[0]) * ir, gu1 = (gy[1] - u * gy[0]) * ir;
    const float gv0 = (gx[2] - v * gx[0]) * ir, gv1 = (gy[2] - v * gy[0]) * ir;
    const float div3 = (gu0 + gv1) * (1.f / 3.f);
    const float m2 = 2.f * mu;
    const float txx = m2 * (gu0 - div3), tyy = m2 * (gv1 - div3), txy = mu * (gu1 + gv0);
    f1 -= nx * txx + ny * txy;
    f2 -= nx * txy + ny * tyy;
    f3 -= nx * (u * txx + v * txy + kc * gTx) + ny * (u * txy + v * tyy + kc * gTy);
}

// x^(1/6) for x > 0 (f_w of the SA model) through the hardware log2 / exp2
__device__ __forceinline__ float pow_sixth_f32(float x) {
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.f / 6.f));
}

// sa_source in FP32 with shared reciprocals (imu = 1/mu_L)
__device__ __forceinline__ float sa_source_f32(float rho, float nut, float imu, float omega, float d, float gvx,
                                               float gvy) {
    const float cross = float(sa::cb2 / sa::sigma) * (gvx * gvx + gvy * gvy);
    const float id = frcp(d);
    const float nd = nut * id;
    const float neg = float(sa::cb1 * (1.0 - sa::ct3)) * omega * nut + float(sa::cw1) * nd * nd + cross;
    const float chi = rho * nut * imu;
    const float c3 = chi * chi * chi;
    const float f1 = c3 * frcp(c3 + float(sa::cv1_3));
    const float f2 = 1.f - chi * frcp(1.f + chi * f1);
    const float ikd2 = float(1.0 / (sa::kappa * sa::kappa)) * id * id;
    const float sbar = nut * f2 * ikd2;
    const float st = (sbar >= -float(sa::cv2) * omega)
                         ? omega + sbar
                         : omega + omega * (float(sa::cv2 * sa::cv2) * omega + float(sa::cv3) * sbar) *
                                       frcp(float(sa::cv3 - 2.0 * sa::cv2) * omega - sbar);
    const float r = st > 0.f ? fminf(nut * frcp(st) * ikd2, 10.f) : 10.f;
    const float r2 = r * r;
    const float gg = r + float(sa::cw2) * (r2 * r2 * r2 - r);
    const float g2 = gg * gg;
    const float fw = gg * pow_sixth_f32(float(1.0 + sa::cw3_6) * frcp(g2 * g2 * g2 + float(sa::cw3_6)));
    const float pos = float(sa::cb1) * st * nut - float(sa::cw1) * fw * nd * nd + cross;
    return rho * (nut < 0.f ? neg : pos);
}

// five-variable face flux of sa_face_flux with the FP32 Roe kernel above,
// times the face length L.  gx2, gy2: twice the face gradients (interior
// faces; the halves are taken into the viscosities).
__device__ __forceinline__ void sa_face_flux_f32(const GasT<float>& g, int kind, float nx, float ny, float L,
                                                 const float* qL, const float* qR, const float* gx2, const float* gy2,
                                                 float nut_c, float dn, float* f) {
    if (kind != K_INTERNAL) {
        float qb[4];
        face_vars<float>(g, kind, nx, ny, qL, qR, qb);
        const float nb = sa_bnd_nu<float>(kind, nx, ny, qL, qL[4], qR, qR[4], qb);
        roe_flux_f32(g.gamma, nx, ny, qL, qb, f, 0.5f * L);
        const bool up = f[0] >= 0.f;
        f[4] = f[0] * (up ? qL[4] : nb) * frcp(up ? qL[0] : qb[0]);
        if (kind == K_WALL) f[4] += L * g.mu * float(1.0 / sa::sigma) * nut_c * frcp(dn);
        return;
    }
    roe_flux_f32(g.gamma, nx, ny, qL, qR, f, 0.5f * L);
    const bool up = f[0] >= 0.f;
    f[4] = f[0] * (up ? qL[4] : qR[4]) * frcp(up ? qL[0] : qR[0]);
    const float nc = 0.5f * (qL[4] + qR[4]);
    const float chi = nc * frcp(g.mu);
    const float c3 = chi * chi * chi;
    // mu_t = nc chi^3 / (chi^3 + cv1^3) for chi > 0, f_n = (cn1 + chi^3) /
    // (cn1 - chi^3) for chi < 0: one reciprocal serves the side taken
    const bool pos = chi > 0.f;
    const float rden = frcp(pos ? c3 + float(sa::cv1_3) : float(sa::cn1) - c3);
    const float mut = pos ? nc * c3 * rden : 0.f;
    const float fn = chi >= 0.f ? 1.f : (float(sa::cn1) + c3) * rden;
    const float kc = g.k + mut * (g.cp * frcp(g.PrT));
    const float nxl = nx * L, nyl = ny * L;
    float f1 = f[1], f2 = f[2], f3 = f[3];
    viscous_terms_f32(g.gamma - 1.f, frcp(g.R), nxl, nyl, qL, qR, gx2, gy2, 0.5f * (g.mu + mut), 0.5f * kc, f1, f2,
                      f3);
    f[1] = f1; f[2] = f2; f[3] = f3;
    const float iqc0 = frcp(0.5f * (qL[0] + qR[0]));
    const float nut = nc * iqc0;
    const float gvx = (gx2[4] - nut * gx2[0]) * iqc0, gvy = (gy2[4] - nut * gy2[0]) * iqc0;
    f[4] -= (g.mu + nc * fn) * float(0.5 / sa::sigma) * (nxl * gvx + nyl * gvy);
}

// sa_cell_source in FP32
__device__ __forceinline__ float sa_cell_source_f32(const GasT<float>& g, const float* q, const float* gx,
                                                    const float* gy, float d) {
    const float ir = frcp(q[0]);
    const float nut = q[4] * ir, u = q[1] * ir, v = q[2] * ir;
    const float omega = fabsf((gx[2] - v * gx[0]) * ir - (gy[1] - u * gy[0]) * ir);
    const float gvx = (gx[4] - nut * gx[0]) * ir, gvy = (gy[4] - nut * gy[0]) * ir;
    return sa_source_f32(q[0], nut, frcp(g.mu), omega, d, gvx, gvy);
}

#ifdef AFX_FP32_UNIT   // defined in afx_k32.hip only
__device__ __forceinline__ float4 f4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
__device__ __forceinline__ float get(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
// slot s of a cell's packed slot word (4 x u16, see TilesF32)
__device__ __forceinline__ uint32_t slot_of(const uint2& w, int s) {
    return s == 0 ? (w.x & 0xFFFFu) : s == 1 ? (w.x >> 16) : s == 2 ? (w.y & 0xFFFFu) : (w.y >> 16);
}
#define F4_LOOP(expr) { expr(x); expr(y); expr(z); expr(w); }

// Limiter of one variable of one cell (solver.h:536-591) from the extreme
// face projections.  Per face the reference takes, for dqg = grad q . d_f,
// phi(dqg, D) = (D^2 + K3a + 2 dqg D) / (D^2 + K3a + dqg D + 2 dqg^2) with
// D = q_max - q when dqg > 1e-16, D = q_min - q when dqg < -1e-16 (else 1), and
// the cell's limiter is the minimum over its faces, at most 1.  phi is even
// in (dqg, D) -> (-dqg, -D).  For dqg > 0 and D >= 0, phi >= 1 exactly when
// dqg <= D/2, and for dqg >= D/4 its derivative is proportional to
// (D - 4 dqg)(D^2 + K3a) - 4 D dqg^2 < 0, so min(1, phi) over the faces is
// min(1, phi(ymax, D+)) with ymax the largest projection, and likewise
// phi(ymin, D-) for the negative side: two evaluations per variable instead
// of one per face.  ymax, ymin start at 0 (faces at or below the 1e-16
// threshold give 1 and are skipped).
__device__ __forceinline__ float limiter_extremes(float ymax, float ymin, float dM, float dm, float K3a) {
    const float DDp = fmaf(dM, dM, K3a), DDm = fmaf(dm, dm, K3a);
    const float tp = ymax * dM, tm = ymin * dm;
    const float pp = fmaf(2.f, tp, DDp) * frcp(fmaf(2.f * ymax, ymax, DDp + tp));
    const float pm = fmaf(2.f, tm, DDm) * frcp(fmaf(2.f * ymin, ymin, DDm + tm));
    float l = 1.f;
    l = ymax > 1e-16f ? fminf(l, pp) : l;
    l = ymin < -1e-16f ? fminf(l, pm) : l;
    return l;
}

// one tile of the FP32 residual (the body of both the single-condition and
// the batched kernels).  MODEL: 0 inviscid, 1 laminar, 2 Spalart-Allmaras;
// GG: Green-Gauss (else least squares) gradients; SO: second order; BND: the
// tile's face list holds boundary faces (without, every face is taken as
// interior at compile time: no kind tests, no boundary states, no ghost
// writes).
template <int MODEL, bool STAGE, bool GG, bool SO, bool BND>
__device__ __forceinline__ void tile_residual_f32(const TileArgsF32& A, const int t, float4* smem4) {
    constexpr bool SA = MODEL == 2;
    const TilesF32& S = A.S;
    const float4* __restrict__ q = A.q;
    const int32_t nC = A.nC;
    const float* __restrict__ iarea = A.iarea;
    const float* __restrict__ a15 = A.a15;
    const float* __restrict__ lsq = A.lsq;
    const float* __restrict__ xc = A.xc;
    const float* __restrict__ yc = A.yc;
    const GasT<float> g = A.g;
    const float K3 = A.K3;
    float4* __restrict__ R = A.R;
    float4* __restrict__ qghost = A.qghost;
    const SAFields32 sa = A.sa;
    float* __restrict__ dt_out = STAGE ? A.dt : nullptr;   // the local time step belongs to the stage form
    // wavefront roles rotate with the tile: the waves of a workgroup sit on
    // the CU's SIMDs in order, so the partial last chunk of each phase falls
    // on a different SIMD for consecutive tiles instead of always the same
    constexpr int kWaves = kBlock / 64;
    static_assert((kWaves & (kWaves - 1)) == 0, "power-of-two wavefronts per workgroup");
    const int tid = ((((int)threadIdx.x >> 6) + t) & (kWaves - 1)) * 64 + ((int)threadIdx.x & 63);
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
    // SA: rho nu~ planes after the float4 arrays
    float2* sgn = reinterpret_cast<float2*>(sfl + S.ftmax);   // [GM] gradient of rho nu~
    float* sqn = reinterpret_cast<float*>(sgn + S.gmax);      // [LM]
    float* slmn = sqn + S.lmax;                                 // [GM]
    float* sfln = slmn + S.gmax;                                // [FM]
    // local time step: lambda len per face touching T (SA: and the viscous
    // coefficient times len^2)
    float* sev = SA ? sfln + S.ftmax : reinterpret_cast<float*>(sfl + S.ftmax);   // [FM]
    float* sev2 = sev + S.ftmax;                                                   // [FM], SA only
    const int* Lg = S.Lglob + lo;
    const uint32_t* fck = S.fck + fo;
    const float4* geo = S.geo + 2 * (int64_t)fo;

    // 1. stage states
    for (int i = tid; i < nL; i += blockDim.x) {
        const int gi = Lg[i];
        if (gi >= 0) {   // wall/slip ghosts (flagged < 0) are never read
            sq[i] = q[gi];
            if constexpr (SA) sqn[i] = sa.qn[gi];
        }
    }
    __syncthreads();

    // 2. gradients and limiters of T u H1
    const uint2* slots2 = reinterpret_cast<const uint2*>(S.slots) + go;
    for (int c = tid; c < nG; c += blockDim.x) {
        const uint2 sw = slots2[c];
        const int gc = Lg[c];
        const float4 qc = sq[c];
        float4 qmin = qc, qmax = qc;
        float4 sx, sy;
        const float nuc = SA ? sqn[c] : 0.f;
        float sxn = 0.f, syn = 0.f, mnn = nuc, mxn = nuc;
        float dxs[4] = {0.f, 0.f, 0.f, 0.f}, dys[4] = {0.f, 0.f, 0.f, 0.f};   // 0 past the last slot
        if constexpr (GG) {
            // Green-Gauss (solver.h:428-469): face value a q_c + b q_other
            // (interior: the weights 1 - g/2 and g/2 of physics.h:275, the
            // same face value for both cells; boundary: 1 - g and g with the
            // BC state), times n len / A.  Summed as
            //   q_c sum_s a_s n_s len_s / A + sum_s b_s q_other,s n_s len_s / A.
            const float ia = iarea[gc];
            float ax = 0.f, ay = 0.f, bxn = 0.f, byn = 0.f;
            float4 bx = f4(0.f, 0.f, 0.f, 0.f), by = f4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const uint32_t e = slot_of(sw, s);
                if (e == 0xFFFFu) break;
                const int f = (int)(e & 0x7FFFu);
                const bool nb = (e >> 15) != 0;
                const uint32_t fc = fck[f];
                const int kind = BND ? (int)(fc >> 24) : (int)K_INTERNAL;
                const int other = nb ? (int)(fc & 0xFFFu) : (int)((fc >> 12) & 0xFFFu);
                const float4 ga = geo[2 * f];   // (nx, ny, len, g)
                if constexpr (SO) {
                    const float4 gb = geo[2 * f + 1];   // (d0x, d0y, d1x, d1y)
                    dxs[s] = nb ? gb.z : gb.x;
                    dys[s] = nb ? gb.w : gb.y;
                }
                const float sl = (nb ? -ga.z : ga.z) * ia;
                const float nlx = ga.x * sl, nly = ga.y * sl;
                float wa, wb, nuo = 0.f;
                float4 qo;
                if (kind == K_INTERNAL) {
                    qo = sq[other];
                    const float wn = 0.5f * ga.w, wo = 1.f - wn;
                    wa = nb ? wn : wo;
                    wb = nb ? wo : wn;
#define MINMAX(m) { qmin.m = fminf(qmin.m, qo.m); qmax.m = fmaxf(qmax.m, qo.m); }
                    F4_LOOP(MINMAX)
#undef MINMAX
                    if constexpr (SA) {
                        nuo = sqn[other];
                        mnn = fminf(mnn, nuo);
                        mxn = fmaxf(mxn, nuo);
                    }
                } else {
                    float qb[4] = {0.f, 0.f, 0.f, 0.f}, qv[4];
                    const float qa[4] = {qc.x, qc.y, qc.z, qc.w};
                    float nug = 0.f;
                    if (kind == K_FARFIELD) {
                        const float4 t4 = sq[other];
                        qb[0] = t4.x; qb[1] = t4.y; qb[2] = t4.z; qb[3] = t4.w;
                        qmin = f4(fminf(qmin.x, t4.x), fminf(qmin.y, t4.y), fminf(qmin.z, t4.z), fminf(qmin.w, t4.w));
                        qmax = f4(fmaxf(qmax.x, t4.x), fmaxf(qmax.y, t4.y), fmaxf(qmax.z, t4.z), fmaxf(qmax.w, t4.w));
                        if constexpr (SA) { nug = sqn[other]; mnn = fminf(mnn, nug); mxn = fmaxf(mxn, nug); }
                    }
                    face_vars<float>(g, kind, ga.x, ga.y, qa, qb, qv);
                    qo = f4(qv[0], qv[1], qv[2], qv[3]);
                    wa = 1.f - ga.w;
                    wb = ga.w;
                    if constexpr (SA) nuo = sa_bnd_nu<float>(kind, ga.x, ga.y, qa, nuc, qb, nug, qv);
                }
                ax = fmaf(wa, nlx, ax);
                ay = fmaf(wa, nly, ay);
                const float bnx = wb * nlx, bny = wb * nly;
#define BSUM(m) { bx.m = fmaf(bnx, qo.m, bx.m); by.m = fmaf(bny, qo.m, by.m); }
                F4_LOOP(BSUM)
#undef BSUM
                if constexpr (SA) {
                    bxn = fmaf(bnx, nuo, bxn);
                    byn = fmaf(bny, nuo, byn);
                }
            }
#define GSUM(m) { sx.m = fmaf(qc.m, ax, bx.m); sy.m = fmaf(qc.m, ay, by.m); }
            F4_LOOP(GSUM)
#undef GSUM
            if constexpr (SA) {
                sxn = fmaf(nuc, ax, bxn);
                syn = fmaf(nuc, ay, byn);
            }
        } else {
            // least squares (solver.h:470-513, quirk A2)
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
                const int kind = BND ? (int)(fc >> 24) : (int)K_INTERNAL;
                const int other = nb ? (int)(fc & 0xFFFu) : (int)((fc >> 12) & 0xFFFu);
                const float4 ga = geo[2 * f], gb = geo[2 * f + 1];
                const float dcx = nb ? gb.z : gb.x, dcy = nb ? gb.w : gb.y;
                const float dox = nb ? gb.x : gb.z, doy = nb ? gb.y : gb.w;
                dxs[s] = dcx; dys[s] = dcy;
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
                if constexpr (SA) {
                    const float nuo = walls ? 0.f : sqn[other];
                    const float vr = (kind == K_INTERNAL) ? (nuc + nuo) * 0.5f
                                                          : sa_bnd_nu<float>(kind, ga.x, ga.y, qa, nuc, qb, nuo, qv);
                    const float del = nuc - vr;
                    sxn += dx * del;
                    syn += dy * del;
                    if (!walls) { mnn = fminf(mnn, nuo); mxn = fmaxf(mxn, nuo); }
                }
            }
            const float m0 = lsq[gc], m1 = lsq[(size_t)nC + gc], m2 = lsq[2 * (size_t)nC + gc], m3 = lsq[3 * (size_t)nC + gc];
            sx = f4(m0 * rx[0] + m1 * ry[0], m0 * rx[1] + m1 * ry[1], m0 * rx[2] + m1 * ry[2], m0 * rx[3] + m1 * ry[3]);
            sy = f4(m2 * rx[0] + m3 * ry[0], m2 * rx[1] + m3 * ry[1], m2 * rx[2] + m3 * ry[2], m2 * rx[3] + m3 * ry[3]);
            qmin = f4(mn[0], mn[1], mn[2], mn[3]);
            qmax = f4(mx[0], mx[1], mx[2], mx[3]);
            if constexpr (SA) {
                const float rxn = sxn, ryn = syn;
                sxn = m0 * rxn + m1 * ryn;
                syn = m2 * rxn + m3 * ryn;
            }
        }
        sgx[c] = sx;
        sgy[c] = sy;
        if constexpr (SA) sgn[c] = make_float2(sxn, syn);
        if constexpr (SO) {
            const float K3a = K3 * a15[gc];
            float4 lm;
#define LIM(m) { float ymax = 0.f, ymin = 0.f;                                         \
                 _Pragma("unroll") for (int s = 0; s < 4; ++s) {                       \
                     const float y = fmaf(sx.m, dxs[s], sy.m * dys[s]);                \
                     ymax = fmaxf(ymax, y); ymin = fminf(ymin, y); }                   \
                 lm.m = limiter_extremes(ymax, ymin, qmax.m - qc.m, qmin.m - qc.m, K3a); }
            F4_LOOP(LIM)
#undef LIM
            slm[c] = lm;
            if constexpr (SA) {
                float ymax = 0.f, ymin = 0.f;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const float y = fmaf(sxn, dxs[s], syn * dys[s]);
                    ymax = fmaxf(ymax, y);
                    ymin = fminf(ymin, y);
                }
                slmn[c] = limiter_extremes(ymax, ymin, mxn - nuc, mnn - nuc, K3a);
            }
        }
    }
    __syncthreads();

    // 3. fluxes of the faces touching T
    for (int fl = tid; fl < nFT; fl += blockDim.x) {
        const uint32_t fc = fck[fl];
        const int o = (int)(fc & 0xFFFu), n = (int)((fc >> 12) & 0xFFFu);
        const int kind = BND ? (int)(fc >> 24) : (int)K_INTERNAL;
        const float4 ga = geo[2 * fl], gb = geo[2 * fl + 1];
        float4 qL = sq[o];
        float4 qR = (kind == K_INTERNAL || kind == K_FARFIELD) ? sq[n] : f4(0.f, 0.f, 0.f, 0.f);
        if (dt_out) {
            // solver.h:313-350: max of the two sides' c + |V_n| (owner side
            // only on boundary faces), from the cell states
            const float gm1 = g.gamma - 1.f;
            const float irL = frcp(qL.x);
            const float VL = (qL.y * ga.x + qL.z * ga.y) * irL;
            const float pL = gm1 * (qL.w - 0.5f * (qL.y * qL.y + qL.z * qL.z) * irL);
            float eig = __builtin_amdgcn_sqrtf(g.gamma * pL * irL) + fabsf(VL);
            if (kind == K_INTERNAL) {
                const float irR = frcp(qR.x);
                const float VR = (qR.y * ga.x + qR.z * ga.y) * irR;
                const float pR = gm1 * (qR.w - 0.5f * (qR.y * qR.y + qR.z * qR.z) * irR);
                eig = fmaxf(eig, __builtin_amdgcn_sqrtf(g.gamma * pR * irR) + fabsf(VR));
            }
            sev[fl] = eig * ga.z;
            if constexpr (SA) {
                const bool in = kind == K_INTERNAL;
                const float rho = in ? 0.5f * (qL.x + qR.x) : qL.x;
                const float nuf = in ? 0.5f * (sqn[o] + sqn[n]) : sqn[o];
                const float chi = nuf * frcp(g.mu), c3 = chi * chi * chi;
                const float mut = chi > 0.f ? nuf * c3 * frcp(c3 + float(sa::cv1_3)) : 0.f;
                const float kmax = fmaxf(g.gamma, 4.f / 3.f);
                sev2[fl] = kmax * (g.k + mut * frcp(g.PrT)) * frcp(rho) * ga.z * ga.z;
            }
        }
        if constexpr (SO) {
            const float4 ax = sgx[o], ay = sgy[o], al = slm[o];
#define RECL(m) qL.m = fmaf(fmaf(ax.m, gb.x, ay.m * gb.y), al.m, qL.m);
            F4_LOOP(RECL)
#undef RECL
            if (!BND || n < nG) {
                const float4 bx = sgx[n], by = sgy[n], bl = slm[n];
#define RECR(m) qR.m = fmaf(fmaf(bx.m, gb.z, by.m * gb.w), bl.m, qR.m);
                F4_LOOP(RECR)
#undef RECR
            }
        }
        float F[5];
        if constexpr (SA) {
            const bool staged = (kind == K_INTERNAL || kind == K_FARFIELD);
            const float nuo_c = sqn[o];
            float nL = nuo_c, nR = staged ? sqn[n] : 0.f;
            if constexpr (SO) {
                const float2 ga_ = sgn[o];
                nL = fmaf(fmaf(ga_.x, gb.x, ga_.y * gb.y), slmn[o], nL);
                if (!BND || n < nG) { const float2 gn2 = sgn[n]; nR = fmaf(fmaf(gn2.x, gb.z, gn2.y * gb.w), slmn[n], nR); }
            }
            const float qL5[5] = {qL.x, qL.y, qL.z, qL.w, nL};
            const float qR5[5] = {qR.x, qR.y, qR.z, qR.w, nR};
            float gxa[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, gya[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
            const float4 so_ = sq[o];
            if (kind == K_INTERNAL) {
                // twice the face gradients along the cell connection
                // (textbook direction): x_n - x_o = d_0 - d_1 from the face
                // offsets; 2 g = S - (S . t - 2 dq/l) t with S the sum of
                // the two cells' gradients
                float t0 = gb.x - gb.z, t1 = gb.y - gb.w;
                const float il = __builtin_amdgcn_rsqf(fmaf(t0, t0, t1 * t1));
                t0 *= il; t1 *= il;
                const float il2 = 2.f * il;
                const float4 sn = sq[n], xo = sgx[o], xn = sgx[n], yo = sgy[o], yn = sgy[n];
                const float2 go2 = sgn[o], gn2 = sgn[n];
                const float so5[5] = {so_.x, so_.y, so_.z, so_.w, nuo_c}, sn5[5] = {sn.x, sn.y, sn.z, sn.w, sqn[n]};
                const float xo5[5] = {xo.x, xo.y, xo.z, xo.w, go2.x}, xn5[5] = {xn.x, xn.y, xn.z, xn.w, gn2.x};
                const float yo5[5] = {yo.x, yo.y, yo.z, yo.w, go2.y}, yn5[5] = {yn.x, yn.y, yn.z, yn.w, gn2.y};
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const float dir2 = (sn5[k] - so5[k]) * il2;
                    const float Sx = xo5[k] + xn5[k], Sy = yo5[k] + yn5[k];
                    const float corr2 = fmaf(Sx, t0, fmaf(Sy, t1, -dir2));
                    gxa[k] = fmaf(-corr2, t0, Sx);
                    gya[k] = fmaf(-corr2, t1, Sy);
                }
            }
            sa_face_flux_f32(g, kind, ga.x, ga.y, ga.z, qL5, qR5, gxa, gya, nuo_c * frcp(so_.x),
                             gb.x * ga.x + gb.y * ga.y, F);
            sfl[fl] = f4(F[0], F[1], F[2], F[3]);
            sfln[fl] = F[4];
            continue;
        }
        const float qLa[4] = {qL.x, qL.y, qL.z, qL.w};
        float qRa[4] = {qR.x, qR.y, qR.z, qR.w};
        if (kind == K_INTERNAL) {
            if constexpr (MODEL == 1) {
                // twice the averaged face gradient with the reference's
                // direction (A1: the normalised sum of the two centroids,
                // solver.h:359-398)
                float gxa[4], gya[4];
                const int go_ = Lg[o], gn_ = Lg[n];
                float t0 = xc[go_] + xc[gn_], t1 = yc[go_] + yc[gn_];
                const float il = __builtin_amdgcn_rsqf(t0 * t0 + t1 * t1);
                t0 *= il; t1 *= il;
                const float il2 = 2.f * il;
                const float4 so = sq[o], sn = sq[n], xo = sgx[o], xn = sgx[n], yo = sgy[o], yn = sgy[n];
                for (int k = 0; k < 4; ++k) {
                    const float dir2 = (get(so, k) - get(sn, k)) * il2;
                    const float Sx = get(xo, k) + get(xn, k), Sy = get(yo, k) + get(yn, k);
                    const float corr2 = fmaf(Sx, t0, fmaf(Sy, t1, -dir2));
                    gxa[k] = fmaf(-corr2, t0, Sx);
                    gya[k] = fmaf(-corr2, t1, Sy);
                }
                roe_flux_f32(g.gamma, ga.x, ga.y, qLa, qRa, F, 0.5f * ga.z);
                viscous_terms_f32(g.gamma - 1.f, frcp(g.R), ga.x * ga.z, ga.y * ga.z, qLa, qRa, gxa, gya, 0.5f * g.mu,
                                  0.5f * g.k, F[1], F[2], F[3]);
            } else {
                roe_flux_f32(g.gamma, ga.x, ga.y, qLa, qRa, F, 0.5f * ga.z);
            }
        } else {
            float qb[4];
            face_vars<float>(g, kind, ga.x, ga.y, qLa, qRa, qb);
            roe_flux_f32(g.gamma, ga.x, ga.y, qLa, qb, F, 0.5f * ga.z);
        }
        sfl[fl] = f4(F[0], F[1], F[2], F[3]);
    }
    __syncthreads();

    // 4. signed sums in reference face order (exact FP64 accumulation of the FP32 fluxes)
    for (int c = tid; c < nTt; c += blockDim.x) {
        const uint2 sw = slots2[c];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
        float lsum = 0.f, vsum = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t e = slot_of(sw, s);
            if (e == 0xFFFFu) break;
            const int fl = (int)(e & 0x7FFFu);
            if (dt_out) {
                lsum += sev[fl];
                if constexpr (SA) vsum += sev2[fl];
            }
            const float4 v = sfl[fl];
            const bool nb = (e >> 15) != 0;
            a0 += (double)(nb ? v.x : -v.x);
            a1 += (double)(nb ? v.y : -v.y);
            a2 += (double)(nb ? v.z : -v.z);
            a3 += (double)(nb ? v.w : -v.w);
            if constexpr (SA) { const float vn = sfln[fl]; a4 += (double)(nb ? vn : -vn); }
            if (BND && qghost) {
                const uint32_t fc = fck[fl];
                const int kind = (int)(fc >> 24);
                if (kind == K_WALL || kind == K_SLIP) {
                    const int gh = -1 - Lg[(int)((fc >> 12) & 0xFFFu)];
                    qghost[gh] = sq[c];
                    if constexpr (SA) sa.qghost_n[gh] = sqn[c];
                }
            }
        }
        const int gc = c0g + c;
        float dtc = 0.f;
        if (dt_out) {
            const float ia = iarea[gc];
            dtc = A.cfl * frcp(ia * lsum + (SA ? 4.f * vsum * ia * ia : 0.f));
            dt_out[gc] = dtc;
        }
        if constexpr (SA) {
            const float4 qs = sq[c], xs = sgx[c], ys = sgy[c];
            const float2 gs = sgn[c];
            const float q5[5] = {qs.x, qs.y, qs.z, qs.w, sqn[c]};
            const float gx5[5] = {xs.x, xs.y, xs.z, xs.w, gs.x}, gy5[5] = {ys.x, ys.y, ys.z, ys.w, gs.y};
            const float2 ad = sa.sa_ad[gc];
            a4 += (double)(ad.x * sa_cell_source_f32(g, q5, gx5, gy5, ad.y));
        }
        if (STAGE) {
            if (!dt_out) dtc = A.dt_in[gc];
            const double w = A.coef * (double)dtc / A.area64[gc];
            const size_t nT = (size_t)A.nT;
            const double v0 = A.q0[gc] + w * a0, v1 = A.q0[nT + gc] + w * a1;
            const double v2 = A.q0[2 * nT + gc] + w * a2, v3 = A.q0[3 * nT + gc] + w * a3;
            A.qout[gc] = v0;
            A.qout[nT + gc] = v1;
            A.qout[2 * nT + gc] = v2;
            A.qout[3 * nT + gc] = v3;
            A.q32out[gc] = f4((float)v0, (float)v1, (float)v2, (float)v3);
            if constexpr (SA) {
                const double v4 = A.q0[4 * nT + gc] + w * a4;
                A.qout[4 * nT + gc] = v4;
                A.q32nout[gc] = (float)v4;
            }
        } else {
            R[gc] = f4((float)a0, (float)a1, (float)a2, (float)a3);
            if constexpr (SA) sa.Rn[gc] = (float)a4;
        }
    }
}

template <int MODEL, bool GG, bool SO>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MODEL == 2 ? 5 : 6)))
k_residual_tile_f32(TileArgsF32 A) {
    extern __shared__ float4 smem4[];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= A.S.ntiles) return;
    if (A.S.nbf[t]) tile_residual_f32<MODEL, false, GG, SO, true>(A, t, smem4);
    else tile_residual_f32<MODEL, false, GG, SO, false>(A, t, smem4);
}

// batched explicit stage: residual + stage update of every (member, tile).
// Five waves per SIMD: for the SA form that costs a few spilled VGPRs, and
// runs faster than four waves without them.
template <int MODEL, bool GG, bool SO>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
k_residual_tile_f32_batch(const TileArgsF32* __restrict__ members, const int2* __restrict__ items) {
    extern __shared__ float4 smem4[];
    const int2 it = items[xcd_tile(blockIdx.x, gridDim.x)];
    if (it.x < 0) return;
    const TileArgsF32& M = members[it.x];
    if (M.S.nbf[it.y]) tile_residual_f32<MODEL, true, GG, SO, true>(M, it.y, smem4);
    else tile_residual_f32<MODEL, true, GG, SO, false>(M, it.y, smem4);
}

// the (model, gradient scheme, order) instantiation for the run-time options
#define AFX_TILE32_DISPATCH(KERNEL, MODEL_, GG_, SO_, ...)                                                 \
    do {                                                                                                   \
        const int sel_ = (MODEL_) * 4 + ((GG_) ? 2 : 0) + ((SO_) ? 1 : 0);                                 \
        switch (sel_) {                                                                                    \
            case 0: hipLaunchKernelGGL((KERNEL<0, false, false>), __VA_ARGS__); break;                     \
            case 1: hipLaunchKernelGGL((KERNEL<0, false, true>), __VA_ARGS__); break;                      \
            case 2: hipLaunchKernelGGL((KERNEL<0, true, false>), __VA_ARGS__); break;                      \
            case 3: hipLaunchKernelGGL((KERNEL<0, true, true>), __VA_ARGS__); break;                       \
            case 4: hipLaunchKernelGGL((KERNEL<1, false, false>), __VA_ARGS__); break;                     \
            case 5: hipLaunchKernelGGL((KERNEL<1, false, true>), __VA_ARGS__); break;                      \
            case 6: hipLaunchKernelGGL((KERNEL<1, true, false>), __VA_ARGS__); break;                      \
            case 7: hipLaunchKernelGGL((KERNEL<1, true, true>), __VA_ARGS__); break;                       \
            case 8: hipLaunchKernelGGL((KERNEL<2, false, false>), __VA_ARGS__); break;                     \
            case 9: hipLaunchKernelGGL((KERNEL<2, false, true>), __VA_ARGS__); break;                      \
            case 10: hipLaunchKernelGGL((KERNEL<2, true, false>), __VA_ARGS__); break;                     \
            default: hipLaunchKernelGGL((KERNEL<2, true, true>), __VA_ARGS__); break;                      \
        }                                                                                                  \
    } while (0)

hipError_t launch_residual_tile_f32(hipStream_t s, TilesF32 S, const float4* q, int32_t nC, const float* iarea,
                                    const float* a15, const float* lsq, const float* xc, const float* yc,
                                    GasT<float> g, int gg, int second_order, float K3, int viscous, float4* R,
                                    float4* qghost, SAFields32 sa) {
    if (viscous < 0 || viscous > 2) return hipErrorInvalidValue;
    const int nb = ((S.ntiles + 7) / 8) * 8;
    TileArgsF32 A{S, q, iarea, a15, lsq, xc, yc, R, qghost, sa, nullptr, g, K3, 0.f, nC, gg, second_order, viscous,
                  nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, 0};
    const size_t lds = tile_lds_bytes_f32(S, viscous == 2 ? 5 : 4, false);
    AFX_TILE32_DISPATCH(k_residual_tile_f32, viscous, gg != 0, second_order != 0, dim3(nb), dim3(kBlock), lds, s, A);
    return hipGetLastError();
}

hipError_t launch_residual_tile_f32_batch(hipStream_t s, const TileArgsF32* members, const int2* items, int32_t nitems,
                                          size_t lds, int viscous, int gg, int second_order) {
    if (nitems <= 0 || (nitems & 7) || viscous < 0 || viscous > 2) return hipErrorInvalidValue;
    AFX_TILE32_DISPATCH(k_residual_tile_f32_batch, viscous, gg != 0, second_order != 0, dim3(nitems), dim3(kBlock),
                        lds, s, members, items);
    return hipGetLastError();
}
#undef AFX_TILE32_DISPATCH
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
    hipError_t launch_dt<T>(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* q, GasT<T> g, T cfl, T* dt, int sa) { \
        hipLaunchKernelGGL(k_dt<T>, dim3(xcd_grid(C.nT)), dim3(kBlock), 0, s, C, F, q, g, cfl, sa, dt);            \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_grad_lim<T>(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* qg, const T* ql, GasT<T> g,   \
                                  int gg, int so, T lk, T* gx, T* gy, T* lim) {                                   \
        hipLaunchKernelGGL(k_grad_lim<T>, dim3(xcd_grid(C.nC)), dim3(kBlock), 0, s, C, F, qg, ql, g, gg, so, lk,   \
                           gx, gy, lim);                                                                          \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_face_flux<T>(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* qr, const T* qa, const T* gx, \
                                   const T* gy, const T* lim, GasT<T> g, int viscous, int so, const T* uniform,   \
                                   T* flux) {                                                                     \
        hipLaunchKernelGGL(k_face_flux<T>, dim3(xcd_grid(F.nF)), dim3(kBlock), 0, s, C, F, qr, qa, gx, gy, lim, g, \
                           viscous, so, uniform, flux);                                                           \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_residual_tile<T>(hipStream_t s, TilesT<T> S, const T* q, int32_t nT, int32_t nC,             \
                                       const T* area, const T* lsq, const T* xc, const T* yc, const T* wd,         \
                                       GasT<T> g, int gg, int so, T lk, int viscous, T* R, T* qghost) {            \
        const int nb = ((S.ntiles + 7) / 8) * 8;                                                                  \
        if (viscous == 2) {                                                                                       \
            const size_t lds = tile_lds_bytes<T>(S.lmax, S.gmax, S.ftmax, 5);                                     \
            hipLaunchKernelGGL((k_residual_tile<T, 5>), dim3(nb), dim3(kBlock), lds, s, S, q, nT, nC, area, lsq,  \
                               xc, yc, wd, g, gg, so, lk, viscous, R, qghost);                                    \
        } else {                                                                                                  \
            const size_t lds = tile_lds_bytes<T>(S.lmax, S.gmax, S.ftmax, 4);                                     \
            hipLaunchKernelGGL((k_residual_tile<T, 4>), dim3(nb), dim3(kBlock), lds, s, S, q, nT, nC, area, lsq,  \
                               xc, yc, wd, g, gg, so, lk, viscous, R, qghost);                                    \
        }                                                                                                         \
        return hipGetLastError();                                                                                 \
    }                                                                                                             \
    template <>                                                                                                   \
    hipError_t launch_gather<T>(hipStream_t s, CellsT<T> C, const T* flux, int32_t nF, int mode, T* R) {          \
        hipLaunchKernelGGL(k_gather<T>, dim3(xcd_grid(C.nC)), dim3(kBlock), 0, s, C, flux, nF, mode, R);           \
        return hipGetLastError();                                                                                 \
    }

}  // namespace afx
