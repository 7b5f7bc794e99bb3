// Reverse-mode (adjoint) pieces of the residual, FP64.
//
//   roe_flux_vjp   given the face states (qL, qR), the normal and the adjoint
//                  bF of the inviscid Roe flux (physics.h:160-228), returns
//                  bqL = (dF/dqL)^T bF and bqR = (dF/dqR)^T bF: the flux is
//                  re-evaluated forward, then swept backwards operation by
//                  operation (derivatives written out by hand);
//   Dual1          forward-mode number: the 4x4 Jacobians of the boundary
//                  states (face_vars, physics.h:267-530) and the tangent of
//                  the whole residual (dR/dq) v, evaluated by the same
//                  templated device code as the residual;
//   limiter_phi_d  value and partial derivatives of the limiter function of
//                  solver.h:578-584 in its selected branch.
#pragma once
#include <cmath>

namespace afx {

template <typename T>
__host__ __device__ inline void roe_flux_vjp(T gamma, T nx, T ny, const T* qL, const T* qR, const T* bF, T* bqL,
                                             T* bqR) {
    using std::sqrt;
    const T gm1 = gamma - T(1);
    // ---- forward
    const T rL = qL[0], ruL = qL[1], rvL = qL[2], reL = qL[3];
    const T rR = qR[0], ruR = qR[1], rvR = qR[2], reR = qR[3];
    const T VLn = (ruL * nx + rvL * ny) / rL, VRn = (ruR * nx + rvR * ny) / rR;
    const T kL = ruL * ruL + rvL * rvL, kR = ruR * ruR + rvR * rvR;
    const T pL = gm1 * (reL - T(0.5) / rL * kL), pR = gm1 * (reR - T(0.5) / rR * kR);
    const T uL = ruL / rL, vL = rvL / rL, uR = ruR / rR, vR = rvR / rR;
    const T sL = sqrt(rL), sR = sqrt(rR), S = sL + sR;
    const T rho = sR * sL;
    const T u = (uL * sL + uR * sR) / S, v = (vL * sL + vR * sR) / S;
    const T HL = (reL + pL) / rL, HR = (reR + pR) / rR;
    const T h = (HL * sL + HR * sR) / S;
    const T q2 = u * u + v * v;
    const T c2 = gm1 * (h - T(0.5) * q2);
    const T c = sqrt(c2);
    const T V = u * nx + v * ny, VR = uR * nx + vR * ny, VL = uL * nx + vL * ny;
    const T d = T(0.05) * c;
    const T a1 = V - c, a3 = V + c;
    const T s1 = sqrt(a1 * a1 + T(1e-4)), s2 = sqrt(V * V + T(1e-4)), s3 = sqrt(a3 * a3 + T(1e-4));
    // entropy fix l > d ? l : (l^2 + d^2) / (2 d), physics.h:133-135
    const bool p1 = s1 > d, p2 = s2 > d, p3 = s3 > d;
    const T lcm = p1 ? s1 : (s1 * s1 + d * d) / (T(2) * d);
    const T lc = p2 ? s2 : (s2 * s2 + d * d) / (T(2) * d);
    const T lcp = p3 ? s3 : (s3 * s3 + d * d) / (T(2) * d);
    const T dp = pR - pL, dV = VR - VL;
    const T cc2 = T(2) * c * c;
    const T A1 = dp - rho * c * dV, A5 = dp + rho * c * dV;
    const T kF1 = lcm * A1 / cc2, kF5 = lcp * A5 / cc2;
    const T B0 = (rR - rL) - dp / (c * c);
    const T k0 = lc * B0, k1 = lc * rho;
    const T du = uR - uL, dv = vR - vL;
    // ---- reverse of F_k = central_k - 0.5 D_k
    const T b0 = bF[0], b1 = bF[1], b2 = bF[2], b3 = bF[3];
    const T e0 = T(-0.5) * b0, e1 = T(-0.5) * b1, e2 = T(-0.5) * b2, e3 = T(-0.5) * b3;
    T b_kF1 = e0 + e1 * (u - c * nx) + e2 * (v - c * ny) + e3 * (h - c * V);
    T b_k0 = e0 + e1 * u + e2 * v + e3 * q2 * T(0.5);
    T b_k1 = e1 * (du - dV * nx) + e2 * (dv - dV * ny) + e3 * (u * du + v * dv - V * dV);
    T b_kF5 = e0 + e1 * (u + c * nx) + e2 * (v + c * ny) + e3 * (h + c * V);
    const T ksum = kF1 + k0 + kF5;
    T b_u = e1 * ksum + e3 * k1 * du;
    T b_v = e2 * ksum + e3 * k1 * dv;
    T b_c = (e1 * nx + e2 * ny + e3 * V) * (kF5 - kF1);
    T b_h = e3 * (kF1 + kF5);
    T b_V = e3 * c * (kF5 - kF1) - e3 * k1 * dV;
    T b_q2 = e3 * k0 * T(0.5);
    T b_du = e1 * k1 + e3 * k1 * u, b_dv = e2 * k1 + e3 * k1 * v;
    T b_dV = -(e1 * nx + e2 * ny + e3 * V) * k1;
    // kF1, kF5, k0, k1
    T b_lcm = b_kF1 * A1 / cc2, b_A1 = b_kF1 * lcm / cc2;
    T b_lcp = b_kF5 * A5 / cc2, b_A5 = b_kF5 * lcp / cc2;
    const T b_cc2 = -(b_kF1 * kF1 + b_kF5 * kF5) / cc2;
    T b_lc = b_k0 * B0 + b_k1 * rho;
    const T b_B0 = b_k0 * lc;
    T b_rho = b_k1 * lc;
    // A1 = dp - rho c dV, A5 = dp + rho c dV, B0 = (rR - rL) - dp / c^2
    T b_dp = b_A1 + b_A5 - b_B0 / (c * c);
    const T b_rcd = b_A5 - b_A1;
    b_rho += b_rcd * c * dV;
    b_c += b_rcd * rho * dV + b_B0 * T(2) * dp / (c * c * c) + b_cc2 * T(4) * c;
    b_dV += b_rcd * rho * c;
    T b_rR = b_B0, b_rL = -b_B0;
    // eigenvalues with the entropy fix; d = 0.05 c
    T b_d = T(0);
    const T l1 = p1 ? T(1) : s1 / d, l2 = p2 ? T(1) : s2 / d, l3 = p3 ? T(1) : s3 / d;
    const T m1 = p1 ? T(0) : (d * d - s1 * s1) / (T(2) * d * d);
    const T m2 = p2 ? T(0) : (d * d - s2 * s2) / (T(2) * d * d);
    const T m3 = p3 ? T(0) : (d * d - s3 * s3) / (T(2) * d * d);
    const T b_s1 = b_lcm * l1, b_s2 = b_lc * l2, b_s3 = b_lcp * l3;
    b_d += b_lcm * m1 + b_lc * m2 + b_lcp * m3;
    const T b_a1 = b_s1 * a1 / s1, b_a3 = b_s3 * a3 / s3;
    b_V += b_a1 + b_s2 * V / s2 + b_a3;
    b_c += b_a3 - b_a1 + T(0.05) * b_d;
    // dV = VR - VL, dp = pR - pL, du, dv
    const T b_VR = b_dV, b_VL = -b_dV;
    T b_pR = b_dp, b_pL = -b_dp;
    T b_uR = b_du + b_VR * nx, b_uL = -b_du + b_VL * nx;
    T b_vR = b_dv + b_VR * ny, b_vL = -b_dv + b_VL * ny;
    // V = u nx + v ny
    b_u += b_V * nx;
    b_v += b_V * ny;
    // c = sqrt(c2), c2 = gm1 (h - q2 / 2)
    const T b_c2 = b_c / (T(2) * c);
    b_h += gm1 * b_c2;
    b_q2 += T(-0.5) * gm1 * b_c2;
    b_u += T(2) * u * b_q2;
    b_v += T(2) * v * b_q2;
    // Roe averages
    T b_sL = b_rho * sR, b_sR = b_rho * sL;
    b_uL += b_u * sL / S; b_uR += b_u * sR / S;
    b_sL += b_u * (uL - u) / S; b_sR += b_u * (uR - u) / S;
    b_vL += b_v * sL / S; b_vR += b_v * sR / S;
    b_sL += b_v * (vL - v) / S; b_sR += b_v * (vR - v) / S;
    const T b_HL = b_h * sL / S, b_HR = b_h * sR / S;
    b_sL += b_h * (HL - h) / S; b_sR += b_h * (HR - h) / S;
    b_rL += b_sL / (T(2) * sL);
    b_rR += b_sR / (T(2) * sR);
    // HL = (reL + pL) / rL
    T b_reL = b_HL / rL, b_reR = b_HR / rR;
    b_pL += b_HL / rL; b_pR += b_HR / rR;
    b_rL -= b_HL * HL / rL; b_rR -= b_HR * HR / rR;
    // uL = ruL / rL ...
    T b_ruL = b_uL / rL, b_rvL = b_vL / rL, b_ruR = b_uR / rR, b_rvR = b_vR / rR;
    b_rL -= (b_uL * uL + b_vL * vL) / rL;
    b_rR -= (b_uR * uR + b_vR * vR) / rR;
    // central part: 0.5 (VLn qL-terms + VRn qR-terms) with the pressures
    T b_VLn = T(0.5) * (b0 * rL + b1 * ruL + b2 * rvL + b3 * (reL + pL));
    T b_VRn = T(0.5) * (b0 * rR + b1 * ruR + b2 * rvR + b3 * (reR + pR));
    b_rL += T(0.5) * b0 * VLn; b_rR += T(0.5) * b0 * VRn;
    b_ruL += T(0.5) * b1 * VLn; b_ruR += T(0.5) * b1 * VRn;
    b_rvL += T(0.5) * b2 * VLn; b_rvR += T(0.5) * b2 * VRn;
    b_reL += T(0.5) * b3 * VLn; b_reR += T(0.5) * b3 * VRn;
    b_pL += T(0.5) * (b1 * nx + b2 * ny + b3 * VLn);
    b_pR += T(0.5) * (b1 * nx + b2 * ny + b3 * VRn);
    // p = gm1 (re - 0.5 (ru^2 + rv^2) / r)
    b_reL += gm1 * b_pL; b_reR += gm1 * b_pR;
    b_ruL -= gm1 * b_pL * ruL / rL; b_ruR -= gm1 * b_pR * ruR / rR;
    b_rvL -= gm1 * b_pL * rvL / rL; b_rvR -= gm1 * b_pR * rvR / rR;
    b_rL += gm1 * b_pL * T(0.5) * kL / (rL * rL);
    b_rR += gm1 * b_pR * T(0.5) * kR / (rR * rR);
    // V_n = (ru nx + rv ny) / r
    b_ruL += b_VLn * nx / rL; b_rvL += b_VLn * ny / rL; b_rL -= b_VLn * VLn / rL;
    b_ruR += b_VRn * nx / rR; b_rvR += b_VRn * ny / rR; b_rR -= b_VRn * VRn / rR;
    bqL[0] = b_rL; bqL[1] = b_ruL; bqL[2] = b_rvL; bqL[3] = b_reL;
    bqR[0] = b_rR; bqR[1] = b_ruR; bqR[2] = b_rvR; bqR[3] = b_reR;
}

// forward-mode number with one derivative
struct Dual1 {
    double v, d;
    __host__ __device__ Dual1() : v(0), d(0) {}
    __host__ __device__ Dual1(double x) : v(x), d(0) {}
    __host__ __device__ Dual1(double x, double dx) : v(x), d(dx) {}
};
__host__ __device__ inline Dual1 operator+(Dual1 a, Dual1 b) { return Dual1(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual1 operator-(Dual1 a, Dual1 b) { return Dual1(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual1 operator-(Dual1 a) { return Dual1(-a.v, -a.d); }
__host__ __device__ inline Dual1 operator*(Dual1 a, Dual1 b) { return Dual1(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual1 operator/(Dual1 a, Dual1 b) {
    const double q = a.v / b.v;
    return Dual1(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual1& operator+=(Dual1& a, Dual1 b) { return a = a + b; }
__host__ __device__ inline Dual1& operator-=(Dual1& a, Dual1 b) { return a = a - b; }
__host__ __device__ inline Dual1& operator*=(Dual1& a, Dual1 b) { return a = a * b; }
__host__ __device__ inline bool operator<(Dual1 a, Dual1 b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(Dual1 a, Dual1 b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(Dual1 a, Dual1 b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(Dual1 a, Dual1 b) { return a.v >= b.v; }
__host__ __device__ inline Dual1 dual_sqrt(Dual1 a) {
    const double s = sqrt(a.v);
    return Dual1(s, a.d / (2.0 * s));
}
// x^(1/3) and its derivative x^(1/3) / (3 x) (x != 0)
__host__ __device__ inline Dual1 dual_cbrt(Dual1 a) {
    const double c = cbrt(a.v);
    return Dual1(c, a.d * c / (3.0 * a.v));
}

// limiter function of solver.h:578-584 in its selected branch (dqg != 0):
// phi = (D^2 + K3a + 2 dqg D) / (D^2 + 2 dqg^2 + D dqg + K3a) and its
// partial derivatives with respect to dqg and D
template <typename T>
__host__ __device__ inline T limiter_phi_d(T dqg, T D, T K3a, T* d_dqg, T* d_D, T* d_K3a = nullptr) {
    const T N = D * D + K3a + T(2) * dqg * D;
    const T M = D * D + T(2) * dqg * dqg + D * dqg + K3a;
    *d_dqg = (T(2) * D * M - N * (T(4) * dqg + D)) / (M * M);
    *d_D = ((T(2) * D + T(2) * dqg) * M - N * (T(2) * D + dqg)) / (M * M);
    if (d_K3a) *d_K3a = (M - N) / (M * M);
    return N / M;
}

}  // namespace afx
