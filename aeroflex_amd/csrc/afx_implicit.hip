// Implicit pseudo-time step on the device (FP64): the finite-difference flux
// Jacobian of implicitSolver::fillRhoLHS (solver.h:979-1075,
// physics.h:533-578) in 4x4 blocks, its block-Jacobi preconditioner, the
// operator product and the Krylov vector kernels used by the restarted
// GMRES driver in afx_capi.hip.
//
// Layout: vectors are SoA [k*nT + i] over real and ghost cells (ghost rows of
// the reference's matrix are identity rows with zero right-hand side, so their
// unknowns stay zero and the ghost columns of boundary faces are dropped).
// Per face Jf[32*f + ...]: J00 = dF/dq_owner * len (row-major 4x4), then
// J01 = dF/dq_neighbour * len.  Per real cell D[16*c + ...] = A/dt I + sum of
// the cell's own blocks, Dinv its inverse.
#include "afx_device.hpp"
#include "afx_implicit.hpp"

namespace afx {

// physics.h:533-578: forward differences with h = max(1e-6, |q|*1e-6), the
// perturbed component restored by subtraction (as the reference does, so a
// rounding difference persists into the later columns)
__global__ void __launch_bounds__(kBlock) k_face_jacobian(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                          const double* __restrict__ gx, const double* __restrict__ gy,
                                                          GasT<double> g, int viscous, double* __restrict__ Jf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT;
    const int o = F.c0[f], n = F.c1[f];
    const int kind = F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f];
    double qL[4], qR[4];
    for (int k = 0; k < 4; ++k) { qL[k] = q[k * nT + o]; qR[k] = q[k * nT + n]; }
    const bool laminar = (viscous == 1) && (kind == K_INTERNAL);
    double gxa[4] = {0, 0, 0, 0}, gya[4] = {0, 0, 0, 0};
    if (laminar) {
        // solver.h:359-398, from the member state q and its gradients
        double t0 = C.xc[o] + C.xc[n];
        double t1 = C.yc[o] + C.yc[n];
        const double l = __builtin_sqrt(t0 * t0 + t1 * t1);
        t0 /= l; t1 /= l;
        for (int k = 0; k < 4; ++k) {
            double dir = (q[k * nT + o] - q[k * nT + n]) / l;
            double gxb = (gx[k * nT + o] + gx[k * nT + n]) * 0.5;
            double gyb = (gy[k * nT + o] + gy[k * nT + n]) * 0.5;
            const double dot = gxb * t0 + gyb * t1;
            gxa[k] = gxb - (dot - dir) * t0;
            gya[k] = gyb - (dot - dir) * t1;
        }
    }
    double f0[4], fp[4];
    face_flux<double>(g, kind, nx, ny, qL, qR, laminar, gxa, gya, f0);
    double* J = Jf + (size_t)32 * f;
    for (int i = 0; i < 4; ++i) {
        {
            const double a = 1e-6, b = dabs<double>(qL[i]) * 1e-6;
            const double up = (a < b) ? b : a;   // std::max(1e-6, |q|*1e-6)
            qL[i] += up;
            face_flux<double>(g, kind, nx, ny, qL, qR, laminar, gxa, gya, fp);
            qL[i] -= up;
            for (int r = 0; r < 4; ++r) J[4 * r + i] = (fp[r] - f0[r]) / up * L;
        }
        {
            const double a = 1e-6, b = dabs<double>(qR[i]) * 1e-6;
            const double up = (a < b) ? b : a;
            qR[i] += up;
            face_flux<double>(g, kind, nx, ny, qL, qR, laminar, gxa, gya, fp);
            qR[i] -= up;
            for (int r = 0; r < 4; ++r) J[16 + 4 * r + i] = (fp[r] - f0[r]) / up * L;
        }
    }
}

// 4x4 inverse by Gauss-Jordan with partial pivoting; returns false if singular
__device__ bool inv4(const double* A, double* X) {
    double M[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) { M[r][c] = A[4 * r + c]; M[r][c + 4] = (r == c) ? 1.0 : 0.0; }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        double best = dabs<double>(M[c][c]);
        for (int r = c + 1; r < 4; ++r)
            if (dabs<double>(M[r][c]) > best) { best = dabs<double>(M[r][c]); p = r; }
        if (!(best > 0.0) || !isfinite(best)) return false;
        if (p != c)
            for (int k = 0; k < 8; ++k) { double t = M[c][k]; M[c][k] = M[p][k]; M[p][k] = t; }
        const double ip = 1.0 / M[c][c];
        for (int k = 0; k < 8; ++k) M[c][k] *= ip;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double fct = M[r][c];
            for (int k = 0; k < 8; ++k) M[r][k] -= fct * M[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) X[4 * r + c] = M[r][c + 4];
    return true;
}

// diagonal blocks (solver.h:1013-1056): A/dt on the diagonal, then the
// cell's own-state blocks of each of its faces in reference face order
// (owner: J00, neighbour: the lower block -J01); then their inverses
__global__ void __launch_bounds__(kBlock) k_assemble_diag(CellsT<double> C, FacesT<double> F, const double* __restrict__ dt,
                                                          const double* __restrict__ Jf, double* __restrict__ D,
                                                          double* __restrict__ Dinv, int* __restrict__ bad) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nC) return;
    double B[16];
    const double ad = C.area[c] / dt[c];
    for (int k = 0; k < 16; ++k) B[k] = 0.0;
    for (int k = 0; k < 4; ++k) B[5 * k] = ad;
    for (int s = 0; s < 4; ++s) {
        uint32_t e = C.cface[(size_t)s * C.nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const double* J = Jf + (size_t)32 * f;
        if (e >> 31) for (int k = 0; k < 16; ++k) B[k] += -J[16 + k];
        else for (int k = 0; k < 16; ++k) B[k] += J[k];
    }
    for (int k = 0; k < 16; ++k) D[(size_t)16 * c + k] = B[k];
    double X[16];
    if (!inv4(B, X)) {
        atomicAdd(bad, 1);
        for (int k = 0; k < 16; ++k) X[k] = 0.0;
    }
    for (int k = 0; k < 16; ++k) Dinv[(size_t)16 * c + k] = X[k];
}

// y = LHS x (ghost rows identity; ghost columns dropped since x_ghost = 0)
__global__ void __launch_bounds__(kBlock) k_lhs_apply(CellsT<double> C, FacesT<double> F, const double* __restrict__ D,
                                                      const double* __restrict__ Jf, const double* __restrict__ x,
                                                      double* __restrict__ y) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nT) return;
    const size_t nT = (size_t)C.nT;
    if (c >= C.nC) {
        for (int k = 0; k < 4; ++k) y[k * nT + c] = x[k * nT + c];
        return;
    }
    double xc[4], acc[4];
    for (int k = 0; k < 4; ++k) xc[k] = x[k * nT + c];
    const double* B = D + (size_t)16 * c;
    for (int r = 0; r < 4; ++r) acc[r] = B[4 * r] * xc[0] + B[4 * r + 1] * xc[1] + B[4 * r + 2] * xc[2] + B[4 * r + 3] * xc[3];
    for (int s = 0; s < 4; ++s) {
        uint32_t e = C.cface[(size_t)s * C.nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        if (F.kind[f] != K_INTERNAL) continue;
        const double* J = Jf + (size_t)32 * f;
        if (e >> 31) {   // this cell is the neighbour: block (c1, c0) = -J00
            const int o = F.c0[f];
            double xo[4];
            for (int k = 0; k < 4; ++k) xo[k] = x[k * nT + o];
            for (int r = 0; r < 4; ++r) acc[r] -= J[4 * r] * xo[0] + J[4 * r + 1] * xo[1] + J[4 * r + 2] * xo[2] + J[4 * r + 3] * xo[3];
        } else {         // owner: block (c0, c1) = J01
            const int n = F.c1[f];
            double xn[4];
            for (int k = 0; k < 4; ++k) xn[k] = x[k * nT + n];
            for (int r = 0; r < 4; ++r)
                acc[r] += J[16 + 4 * r] * xn[0] + J[16 + 4 * r + 1] * xn[1] + J[16 + 4 * r + 2] * xn[2] + J[16 + 4 * r + 3] * xn[3];
        }
    }
    for (int k = 0; k < 4; ++k) y[k * nT + c] = acc[k];
}

// z = Dinv x (ghost rows identity)
__global__ void __launch_bounds__(kBlock) k_block_jacobi(int32_t nC, int32_t nT, const double* __restrict__ Dinv,
                                                         const double* __restrict__ x, double* __restrict__ z) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nT) return;
    const size_t n = (size_t)nT;
    if (c >= nC) {
        for (int k = 0; k < 4; ++k) z[k * n + c] = x[k * n + c];
        return;
    }
    double xc[4];
    for (int k = 0; k < 4; ++k) xc[k] = x[k * n + c];
    const double* B = Dinv + (size_t)16 * c;
    for (int r = 0; r < 4; ++r) z[r * n + c] = B[4 * r] * xc[0] + B[4 * r + 1] * xc[1] + B[4 * r + 2] * xc[2] + B[4 * r + 3] * xc[3];
}

// partial[j*nblk + b] = sum over block b of V_j . w, j < m (V_j contiguous, stride ld)
__global__ void __launch_bounds__(kBlock) k_multidot(const double* __restrict__ V, int64_t ld, int m, const double* __restrict__ w,
                                                     int64_t n, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    for (int j = 0; j < m; ++j) {
        const double* v = V + (size_t)j * ld;
        double s = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
            s += v[i] * w[i];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < kBlock / 64; ++k) t += red[k];
            partial[(size_t)j * gridDim.x + blockIdx.x] = t;
        }
        __syncthreads();
    }
}

// out[j] = sum_b partial[j*nblk + b]
__global__ void k_multidot_final(const double* __restrict__ partial, int nblk, int m, double* __restrict__ out) {
    __shared__ double red[kBlock / 64];
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += partial[(size_t)j * nblk + i];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < kBlock / 64; ++k) t += red[k];
            out[j] = t;
        }
        __syncthreads();
    }
}

// w -= sum_j h[j] V_j   (h on the device)
__global__ void __launch_bounds__(kBlock) k_multiaxpy(const double* __restrict__ V, int64_t ld, int m,
                                                      const double* __restrict__ h, double* __restrict__ w, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = w[i];
    for (int j = 0; j < m; ++j) acc -= h[j] * V[(size_t)j * ld + i];
    w[i] = acc;
}

// x += sum_j y[j] V_j   (y on the device)
__global__ void __launch_bounds__(kBlock) k_lincomb(const double* __restrict__ V, int64_t ld, int m,
                                                    const double* __restrict__ y, double* __restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = x[i];
    for (int j = 0; j < m; ++j) acc += y[j] * V[(size_t)j * ld + i];
    x[i] = acc;
}

// y = a * x
__global__ void __launch_bounds__(kBlock) k_scale(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

// y += a * x
__global__ void __launch_bounds__(kBlock) k_axpy(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// ---------------------------------------------- matrix-free operator (a18) --
// RansMatrix * x (ransMatrix.h:138-178, the RANS_MATRIX_FREE build): the
// operator calls each face's flux with two arguments, i.e. at cell values and
// zero gradients, with one step h = 1e-6 along the whole 4-vector of the
// perturbed side and no face length.  Per face it adds
// (F(q_L + h x_L, q_R) - F)/h to the owner row and, for two-sided faces,
// subtracts (F(q_L, q_R + h x_R) - F)/h from the neighbour row: every row sees
// only its own block of x.  Real rows start from x A / dt, ghost rows are x.

// F(q_L, q_R) per face at the current state (SoA Fb[k * nF + f])
__global__ void __launch_bounds__(kBlock) k_mf_base(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                    GasT<double> g, int viscous, double* __restrict__ Fb) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    const int o = F.c0[f], n = F.c1[f], kind = F.kind[f];
    double qL[4], qR[4], f0[4];
    const double z[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) { qL[k] = q[k * nT + o]; qR[k] = q[k * nT + n]; }
    face_flux<double>(g, kind, F.nx[f], F.ny[f], qL, qR, (viscous == 1) && (kind == K_INTERNAL), z, z, f0);
    for (int k = 0; k < 4; ++k) Fb[k * nF + f] = f0[k];
}

// one thread per cell; the cell's faces in reference face order, as the
// reference's edge loop reaches them
__global__ void __launch_bounds__(kBlock) k_mf_apply(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                     const double* __restrict__ dt, const double* __restrict__ Fb,
                                                     GasT<double> g, int viscous, const double* __restrict__ x,
                                                     double* __restrict__ y) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C.nT) return;
    const size_t nT = (size_t)C.nT, nF = (size_t)F.nF;
    double xc[4];
    for (int k = 0; k < 4; ++k) xc[k] = x[k * nT + c];
    if (c >= C.nC) {
        for (int k = 0; k < 4; ++k) y[k * nT + c] = xc[k];
        return;
    }
    constexpr double h = 1e-6;
    const double z[4] = {0, 0, 0, 0};
    double acc[4], qc[4], qp[4], qo[4], fp[4];
    for (int k = 0; k < 4; ++k) {
        acc[k] = xc[k] * C.area[c] / dt[c];
        qc[k] = q[k * nT + c];
        qp[k] = qc[k] + h * xc[k];
    }
    for (int s = 0; s < 4; ++s) {
        uint32_t e = C.cface[(size_t)s * C.nC + c];
        if (e == 0xFFFFFFFFu) break;
        const int f = (int)(e & 0x7FFFFFFFu);
        const int kind = F.kind[f];
        const bool lam = (viscous == 1) && (kind == K_INTERNAL);
        if (e >> 31) {   // neighbour row of a two-sided face
            if (kind != K_INTERNAL) continue;
            const int o = F.c0[f];
            for (int k = 0; k < 4; ++k) qo[k] = q[k * nT + o];
            face_flux<double>(g, kind, F.nx[f], F.ny[f], qo, qp, lam, z, z, fp);
            for (int k = 0; k < 4; ++k) acc[k] -= (fp[k] - Fb[k * nF + f]) / h;
        } else {         // owner row
            const int n = F.c1[f];
            for (int k = 0; k < 4; ++k) qo[k] = q[k * nT + n];
            face_flux<double>(g, kind, F.nx[f], F.ny[f], qp, qo, lam, z, z, fp);
            for (int k = 0; k < 4; ++k) acc[k] += (fp[k] - Fb[k * nF + f]) / h;
        }
    }
    for (int k = 0; k < 4; ++k) y[k * nT + c] = acc[k];
}

// ---------------------------------------------------------------- launchers --
hipError_t launch_face_jacobian(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* gx,
                                const double* gy, GasT<double> g, int viscous, double* Jf) {
    hipLaunchKernelGGL(k_face_jacobian, dim3(nblocks(F.nF)), dim3(kBlock), 0, s, C, F, q, gx, gy, g, viscous, Jf);
    return hipGetLastError();
}
hipError_t launch_assemble_diag(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* dt, const double* Jf,
                                double* D, double* Dinv, int* bad) {
    hipLaunchKernelGGL(k_assemble_diag, dim3(nblocks(C.nC)), dim3(kBlock), 0, s, C, F, dt, Jf, D, Dinv, bad);
    return hipGetLastError();
}
hipError_t launch_lhs_apply(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* D, const double* Jf,
                            const double* x, double* y) {
    hipLaunchKernelGGL(k_lhs_apply, dim3(nblocks(C.nT)), dim3(kBlock), 0, s, C, F, D, Jf, x, y);
    return hipGetLastError();
}
hipError_t launch_block_jacobi(hipStream_t s, int32_t nC, int32_t nT, const double* Dinv, const double* x, double* z) {
    hipLaunchKernelGGL(k_block_jacobi, dim3(nblocks(nT)), dim3(kBlock), 0, s, nC, nT, Dinv, x, z);
    return hipGetLastError();
}
hipError_t launch_multidot(hipStream_t s, const double* V, int64_t ld, int m, const double* w, int64_t n,
                           double* partial, int nblk, double* out) {
    hipLaunchKernelGGL(k_multidot, dim3(nblk), dim3(kBlock), 0, s, V, ld, m, w, n, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_multidot_final, dim3(1), dim3(kBlock), 0, s, partial, nblk, m, out);
    return hipGetLastError();
}
hipError_t launch_multiaxpy(hipStream_t s, const double* V, int64_t ld, int m, const double* h, double* w, int64_t n) {
    hipLaunchKernelGGL(k_multiaxpy, dim3(nblocks(n)), dim3(kBlock), 0, s, V, ld, m, h, w, n);
    return hipGetLastError();
}
hipError_t launch_lincomb(hipStream_t s, const double* V, int64_t ld, int m, const double* y, double* x, int64_t n) {
    hipLaunchKernelGGL(k_lincomb, dim3(nblocks(n)), dim3(kBlock), 0, s, V, ld, m, y, x, n);
    return hipGetLastError();
}
hipError_t launch_scale(hipStream_t s, double a, const double* x, double* y, int64_t n) {
    hipLaunchKernelGGL(k_scale, dim3(nblocks(n)), dim3(kBlock), 0, s, a, x, y, n);
    return hipGetLastError();
}
hipError_t launch_axpy(hipStream_t s, double a, const double* x, double* y, int64_t n) {
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), dim3(kBlock), 0, s, a, x, y, n);
    return hipGetLastError();
}

hipError_t launch_mf_base(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                          int viscous, double* Fb) {
    hipLaunchKernelGGL(k_mf_base, dim3(nblocks(F.nF)), dim3(kBlock), 0, s, C, F, q, g, viscous, Fb);
    return hipGetLastError();
}
hipError_t launch_mf_apply(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* dt,
                           const double* Fb, GasT<double> g, int viscous, const double* x, double* y) {
    hipLaunchKernelGGL(k_mf_apply, dim3(nblocks(C.nT)), dim3(kBlock), 0, s, C, F, q, dt, Fb, g, viscous, x, y);
    return hipGetLastError();
}

}  // namespace afx
