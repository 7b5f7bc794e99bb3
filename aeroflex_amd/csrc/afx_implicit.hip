// Implicit pseudo-time step on the device (FP64): the finite-difference flux
// Jacobian of implicitSolver::fillRhoLHS (solver.h:979-1075,
// physics.h:533-578) in 4x4 blocks, its block-Jacobi and multicolour block
// DILU preconditioners, the operator product and the kernels of the
// device-resident restarted GMRES driven from afx_capi.hip (the Arnoldi
// coefficients, Givens rotations and convergence flags live in device
// memory; the host reads the flags between chunks of Arnoldi steps).
//
// Blocks are N x N with N = 4 (the reference's four variables) or N = 5 (the
// Spalart-Allmaras model, whose Jacobian is this build's own: the reference's
// SA branch is empty, physics.h:259-261).
//
// Layout: vectors are SoA [k*nT + i] over real and ghost cells (ghost rows of
// the reference's matrix are identity rows with zero right-hand side, so their
// unknowns stay zero and the ghost columns of boundary faces are dropped).
// Per face Jf[2NN*f + ...]: J00 = dF/dq_owner * len (row-major N x N), then
// J01 = dF/dq_neighbour * len.  Per real cell D[NN*c + ...] = A/dt I + sum of
// the cell's own blocks (+ the SA source row), Dinv its inverse.
#include "afx_device.hpp"
#include "afx_implicit.hpp"

namespace afx {

// physics.h:533-578: forward differences with h = max(1e-6, |q|*1e-6), the
// perturbed component restored by subtraction (as the reference does, so a
// rounding difference persists into the later columns)
__global__ void __launch_bounds__(kBlock) k_face_jacobian(CellsT<double> C, FacesT<double> F, const double* __restrict__ q,
                                                          const double* __restrict__ gx, const double* __restrict__ gy,
                                                          GasT<double> g, int viscous, double* __restrict__ Jf) {
    int f = xcd_index();
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

// Spalart-Allmaras: 5x5 blocks by the same rule (forward differences,
// h = max(1e-6, |v|*1e-6), the perturbed component restored by subtraction,
// first-order cell states, times the face length) of the five-variable face
// flux sa_face_flux.  The cell gradients gx, gy (5 planes) stay frozen; the
// face gradients' directional parts (v_n - v_o)/l, l along the cell
// connection, and the no-slip wall's cell-centre nu~ follow the perturbed
// states, so the viscous and nu~ diffusion couplings enter the blocks.  The
// rho nu~ column's step floor is 1e-6 mu_L instead of 1e-6 (rho nu~ is of the
// order mu_L chi; the absolute floor would perturb it by O(1)).
__global__ void __launch_bounds__(kBlock) k_face_jacobian_sa(CellsT<double> C, FacesT<double> F,
                                                             const double* __restrict__ q,
                                                             const double* __restrict__ gx,
                                                             const double* __restrict__ gy, GasT<double> g,
                                                             double* __restrict__ Jf) {
    const int f = xcd_index();
    if (f >= F.nF) return;
    const size_t nT = (size_t)C.nT;
    const int o = F.c0[f], n = F.c1[f];
    const int kind = F.kind[f];
    const double nx = F.nx[f], ny = F.ny[f], L = F.len[f];
    const bool internal = kind == K_INTERNAL;
    double a[5], b[5], gxb[5], gyb[5];
    for (int k = 0; k < 5; ++k) {
        a[k] = q[k * nT + o];
        b[k] = q[k * nT + n];
        gxb[k] = gyb[k] = 0.0;
    }
    double t0 = 0.0, t1 = 0.0, l = 1.0;
    if (internal) {
        t0 = C.xc[n] - C.xc[o];
        t1 = C.yc[n] - C.yc[o];
        l = __builtin_sqrt(t0 * t0 + t1 * t1);
        t0 /= l; t1 /= l;
        for (int k = 0; k < 5; ++k) {
            gxb[k] = (gx[k * nT + o] + gx[k * nT + n]) * 0.5;
            gyb[k] = (gy[k * nT + o] + gy[k * nT + n]) * 0.5;
        }
    }
    const double dn = (F.xf[f] - C.xc[o]) * nx + (F.yf[f] - C.yc[o]) * ny;
    auto flux = [&](double* out) {
        double gxa[5] = {0, 0, 0, 0, 0}, gya[5] = {0, 0, 0, 0, 0};
        if (internal)
            for (int k = 0; k < 5; ++k) {
                const double dir = (b[k] - a[k]) / l;
                const double dot = gxb[k] * t0 + gyb[k] * t1;
                gxa[k] = gxb[k] - (dot - dir) * t0;
                gya[k] = gyb[k] - (dot - dir) * t1;
            }
        sa_face_flux<double>(g, kind, nx, ny, a, b, gxa, gya, a[4] / a[0], dn, out);
    };
    double f0[5], fp[5];
    flux(f0);
    double* J = Jf + (size_t)50 * f;
    for (int i = 0; i < 5; ++i) {
        const double lo = i == 4 ? 1e-6 * g.mu : 1e-6;
        {
            const double hi = dabs<double>(a[i]) * 1e-6;
            const double up = (lo < hi) ? hi : lo;
            a[i] += up;
            flux(fp);
            a[i] -= up;
            for (int r = 0; r < 5; ++r) J[5 * r + i] = (fp[r] - f0[r]) / up * L;
        }
        {
            const double hi = dabs<double>(b[i]) * 1e-6;
            const double up = (lo < hi) ? hi : lo;
            b[i] += up;
            flux(fp);
            b[i] -= up;
            for (int r = 0; r < 5; ++r) J[25 + 5 * r + i] = (fp[r] - f0[r]) / up * L;
        }
    }
}

// N x N inverse by Gauss-Jordan with partial pivoting; returns false if
// singular.  Every loop is unrolled and the row exchange is a select over the
// candidate rows, so M is indexed at compile time only and stays in
// registers (a runtime row index would put it in scratch memory).
template <int N>
__device__ __forceinline__ bool invN(const double* A, double* X) {
    double M[N][2 * N];
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) { M[r][c] = A[N * r + c]; M[r][c + N] = (r == c) ? 1.0 : 0.0; }
#pragma unroll
    for (int c = 0; c < N; ++c) {
        int p = c;
        double best = dabs<double>(M[c][c]);
#pragma unroll
        for (int r = c + 1; r < N; ++r) {
            const double a = dabs<double>(M[r][c]);
            if (a > best) { best = a; p = r; }
        }
        if (!(best > 0.0) || !isfinite(best)) return false;
#pragma unroll
        for (int r = c + 1; r < N; ++r) {   // exchange rows c and p
            const bool sw = r == p;
#pragma unroll
            for (int k = 0; k < 2 * N; ++k) {
                const double a = M[c][k], b = M[r][k];
                M[c][k] = sw ? b : a;
                M[r][k] = sw ? a : b;
            }
        }
        const double ip = 1.0 / M[c][c];
#pragma unroll
        for (int k = 0; k < 2 * N; ++k) M[c][k] *= ip;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            if (r == c) continue;
            const double fct = M[r][c];
#pragma unroll
            for (int k = 0; k < 2 * N; ++k) M[r][k] -= fct * M[c][k];
        }
    }
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) X[N * r + c] = M[r][c + N];
    return true;
}

// diagonal blocks (solver.h:1013-1056): A/dt on the diagonal, then the
// cell's own-state blocks of each of its faces in reference face order
// (owner: J00, neighbour: the lower block -J01); then their inverses.
// N = 5 adds the source row: -A dS/dv at the frozen gradients (same step
// rule and rho nu~ floor), its diagonal entry only where >= 0 (destruction),
// so production stays explicit and never lowers the diagonal of the nu~ row.
// The same pass over the cell's faces writes the ELL copy of its row for the
// product (k_lhs_apply): per slot s the off-diagonal block (-J00 where the
// cell is the face's neighbour, J01 where it is its owner; zero on empty
// slots and boundary faces) at E[(s NN + k) nC + c], the other cell at
// nbr[s nC + c] (-1: none), the diagonal block at Ds[k nC + c].  Every ELL
// array is read with consecutive lanes on consecutive addresses.
template <int N>
__global__ void __launch_bounds__(kBlock) k_assemble_diag(CellsT<double> C, FacesT<double> F, const double* __restrict__ dt,
                                                          const double* __restrict__ Jf, SaSource src,
                                                          double* __restrict__ D, double* __restrict__ Dinv,
                                                          double* __restrict__ E, double* __restrict__ Ds,
                                                          int32_t* __restrict__ nbr, int* __restrict__ bad) {
    constexpr int NN = N * N;
    int c = xcd_index();
    if (c >= C.nC) return;
    const size_t nC = (size_t)C.nC;
    double B[NN];
    const double ad = C.area[c] / dt[c];
    for (int k = 0; k < NN; ++k) B[k] = 0.0;
    for (int k = 0; k < N; ++k) B[(N + 1) * k] = ad;
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * nC + c];
        int other = -1;
        if (e != 0xFFFFFFFFu) {
            const int f = (int)(e & 0x7FFFFFFFu);
            const double* J = Jf + (size_t)2 * NN * f;
            const bool nb = (e >> 31) != 0;
            if (nb) for (int k = 0; k < NN; ++k) B[k] += -J[NN + k];
            else for (int k = 0; k < NN; ++k) B[k] += J[k];
            if (F.kind[f] == K_INTERNAL) {
                other = nb ? F.c0[f] : F.c1[f];
                for (int k = 0; k < NN; ++k) E[((size_t)s * NN + k) * nC + c] = nb ? -J[k] : J[NN + k];
            }
        }
        nbr[s * nC + c] = other;
        if (other < 0)
            for (int k = 0; k < NN; ++k) E[((size_t)s * NN + k) * nC + c] = 0.0;
    }
    if constexpr (N == 5) {
        if (src.q) {
            const size_t nT = (size_t)C.nT;
            double v[5], sx[5], sy[5];
            for (int k = 0; k < 5; ++k) {
                v[k] = src.q[k * nT + c];
                sx[k] = src.gx[k * nT + c];
                sy[k] = src.gy[k * nT + c];
            }
            const double d = src.wd[c], A = C.area[c];
            const double s0 = sa_cell_source<double>(src.g, v, sx, sy, d);
            for (int k = 0; k < 5; ++k) {
                const double lo = k == 4 ? 1e-6 * src.g.mu : 1e-6, hi = dabs<double>(v[k]) * 1e-6;
                const double up = (lo < hi) ? hi : lo;
                v[k] += up;
                const double sp = sa_cell_source<double>(src.g, v, sx, sy, d);
                v[k] -= up;
                double val = -A * (sp - s0) / up;
                if (k == 4 && val < 0) val = 0;
                B[20 + k] += val;
            }
        }
    }
    for (int k = 0; k < NN; ++k) {
        D[(size_t)NN * c + k] = B[k];
        Ds[k * nC + c] = B[k];
    }
    double X[NN];
    if (!invN<N>(B, X)) {
        atomicAdd(bad, 1);
        for (int k = 0; k < NN; ++k) X[k] = 0.0;
    }
    for (int k = 0; k < NN; ++k) Dinv[(size_t)NN * c + k] = X[k];
}

// y = LHS x (ghost rows identity; ghost columns dropped since x_ghost = 0)
// from the ELL copy written by k_assemble_diag.  Per row: the diagonal block
// times x_c, then each slot's block times the neighbour's x in slot order
// (the sign sits in the block instead of in x, exactly).
template <int N>
__global__ void __launch_bounds__(kBlock) k_lhs_apply(int32_t nC, int32_t nT, const double* __restrict__ Ds,
                                                      const double* __restrict__ E, const int32_t* __restrict__ nbr,
                                                      const double* __restrict__ x, double* __restrict__ y,
                                                      const int* __restrict__ active) {
    constexpr int NN = N * N;
    if (active && *active == 0) return;
    const int c = xcd_index();
    if (c >= nT) return;
    const size_t n = (size_t)nT, m = (size_t)nC;
    if (c >= nC) {
        for (int k = 0; k < N; ++k) y[k * n + c] = x[k * n + c];
        return;
    }
    int other[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) other[s] = nbr[s * m + c];
    double xc[N], acc[N];
    for (int k = 0; k < N; ++k) xc[k] = x[k * n + c];
    for (int r = 0; r < N; ++r) {
        double t = 0.0;
        for (int k = 0; k < N; ++k) t += Ds[(r * N + k) * m + c] * xc[k];
        acc[r] = t;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (other[s] < 0) continue;
        double xo[N];
        for (int k = 0; k < N; ++k) xo[k] = x[k * n + other[s]];
        const double* B = E + (size_t)s * NN * m + c;
        for (int r = 0; r < N; ++r) {
            double t = 0.0;
            for (int k = 0; k < N; ++k) t += B[(r * N + k) * m] * xo[k];
            acc[r] += t;
        }
    }
    for (int k = 0; k < N; ++k) y[k * n + c] = acc[k];
}

// z = Dinv x (ghost rows identity)
template <int N>
__global__ void __launch_bounds__(kBlock) k_block_jacobi(int32_t nC, int32_t nT, const double* __restrict__ Dinv,
                                                         const double* __restrict__ x, double* __restrict__ z,
                                                         const int* __restrict__ active) {
    if (active && *active == 0) return;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nT) return;
    const size_t n = (size_t)nT;
    if (c >= nC) {
        for (int k = 0; k < N; ++k) z[k * n + c] = x[k * n + c];
        return;
    }
    double xc[N];
    for (int k = 0; k < N; ++k) xc[k] = x[k * n + c];
    const double* B = Dinv + (size_t)N * N * c;
    for (int r = 0; r < N; ++r) {
        double t = 0.0;
        for (int k = 0; k < N; ++k) t += B[N * r + k] * xc[k];
        z[r * n + c] = t;
    }
}

// y += a * x
__global__ void __launch_bounds__(kBlock) k_axpy(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// ------------------------------------------- multicolour block DILU (f1) --
// The reference preconditions GMRES with an incomplete LU (IncompleteLUT,
// solver.h:868-910).  On the device: block DILU in a colour order of the
// real cells (no two face neighbours share a colour), so each colour is one
// parallel sweep.  With A_ij the off-diagonal block of row i, column j:
//   Dt_i = D_i - sum_{j before i} A_ij Dt_j^-1 A_ji          (set-up, by colour)
//   M = (Dt + L) Dt^-1 (Dt + U),  L / U: the blocks A_ij with j before / after i.
// Off-diagonal blocks from the face Jacobians: for face f with owner o and
// neighbour n, A_on = J01(f) and A_no = -J00(f); boundary faces have none.

// the block A_ij (sgn, J) of cell i across its slot entry e, and the other cell
template <int N>
__device__ __forceinline__ const double* offdiag(const FacesT<double>& F, const double* Jf, uint32_t e, int* other,
                                                  double* sgn) {
    const int f = (int)(e & 0x7FFFFFFFu);
    const double* J = Jf + (size_t)2 * N * N * f;
    if (e >> 31) { *other = F.c0[f]; *sgn = -1.0; return J; }        // i is the neighbour: -J00
    *other = F.c1[f]; *sgn = 1.0; return J + N * N;                   // i is the owner: J01
}
// the transposed-position block A_ji of the same face
template <int N>
__device__ __forceinline__ const double* offdiag_t(const double* Jf, uint32_t e, double* sgn) {
    const int f = (int)(e & 0x7FFFFFFFu);
    const double* J = Jf + (size_t)2 * N * N * f;
    if (e >> 31) { *sgn = 1.0; return J + N * N; }
    *sgn = -1.0; return J;
}

// Set-up of colour c over positions p0 .. p0 + ncells of the colour order.
// Besides Dt_i^-1 (AoS, read back by the later colours' set-up) it writes the
// sweeps' colour-ordered ELL copy of row i at its position p: the
// off-diagonal blocks A_ij by slot at Ec[(s NN + k) nC + p], the neighbour
// code nc[s nC + p] = pos_j + 1 for a neighbour of an earlier colour,
// -(pos_j + 1) for a later one, 0 for none (pos_j: the neighbour's position in
// the colour order), and Dt_i^-1 at Dc[k nC + p], so the threads of one
// colour's sweep read consecutive addresses.
template <int N>
__global__ void __launch_bounds__(kBlock) k_dilu_setup(CellsT<double> C, FacesT<double> F, const uint8_t* __restrict__ colour,
                                                       int c, const int32_t* __restrict__ cells,
                                                       const int32_t* __restrict__ pos, int32_t p0,
                                                       int32_t ncells, const double* __restrict__ D,
                                                       const double* __restrict__ Jf, double* __restrict__ Dti,
                                                       double* __restrict__ Ec, int32_t* __restrict__ nc,
                                                       double* __restrict__ Dc, int* __restrict__ bad) {
    constexpr int NN = N * N;
    const int k = xcd_index();
    if (k >= ncells) return;
    const size_t mc = (size_t)C.nC;
    const int p = p0 + k;
    const int i = cells[p];
    double B[NN];
    for (int r = 0; r < NN; ++r) B[r] = D[(size_t)NN * i + r];
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * C.nC + i];
        if (e == 0xFFFFFFFFu) break;
        if (F.kind[e & 0x7FFFFFFFu] != K_INTERNAL) continue;
        int j;
        double sa, sb;
        const double* Aij = offdiag<N>(F, Jf, e, &j, &sa);
        if (colour[j] >= c) continue;
        const double* Aji = offdiag_t<N>(Jf, e, &sb);
        const double* Pj = Dti + (size_t)NN * j;
        double T[NN];   // Dt_j^-1 A_ji
        for (int r = 0; r < N; ++r)
            for (int q = 0; q < N; ++q) {
                double t = 0.0;
                for (int m = 0; m < N; ++m) t += Pj[N * r + m] * Aji[N * m + q];
                T[N * r + q] = sb * t;
            }
        for (int r = 0; r < N; ++r)
            for (int q = 0; q < N; ++q) {
                double t = 0.0;
                for (int m = 0; m < N; ++m) t += Aij[N * r + m] * T[N * m + q];
                B[N * r + q] -= sa * t;
            }
    }
    double X[NN];
    if (!invN<N>(B, X)) {
        atomicAdd(bad, 1);
        for (int r = 0; r < NN; ++r) X[r] = 0.0;
    }
    for (int r = 0; r < NN; ++r) {
        Dti[(size_t)NN * i + r] = X[r];
        Dc[r * mc + p] = X[r];
    }
    for (int s = 0; s < 4; ++s) {
        const uint32_t e = C.cface[(size_t)s * mc + i];
        int code = 0;
        if (e != 0xFFFFFFFFu && F.kind[e & 0x7FFFFFFFu] == K_INTERNAL) {
            int j;
            double sa;
            const double* Aij = offdiag<N>(F, Jf, e, &j, &sa);
            code = colour[j] < c ? pos[j] + 1 : -(pos[j] + 1);
            for (int r = 0; r < NN; ++r) Ec[((size_t)s * NN + r) * mc + p] = sa * Aij[r];
        } else {
            for (int r = 0; r < NN; ++r) Ec[((size_t)s * NN + r) * mc + p] = 0.0;
        }
        nc[s * mc + p] = code;
    }
}

// forward sweep of colour c: y_i = Dt_i^-1 (b_i - sum_{j before i} A_ij y_j)
// backward sweep of colour c (in place): y_i -= Dt_i^-1 sum_{j after i} A_ij y_j
// over positions p0 .. p0 + ncells of the colour order, from the set-up's
// colour-ordered copy (the sign of A_ij sits in the block, exactly).  y is
// held by position (yp, nv planes of nC), so each sweep reads and writes it
// with consecutive lanes on consecutive addresses; b is in cell order.
template <int N, bool FORWARD>
__global__ void __launch_bounds__(kBlock) k_dilu_sweep(int32_t nC, int32_t nT, int32_t p0, int32_t ncells,
                                                       const int32_t* __restrict__ cells,
                                                       const double* __restrict__ Ec, const int32_t* __restrict__ nc,
                                                       const double* __restrict__ Dc, const double* __restrict__ b,
                                                       double* __restrict__ yp, const int* __restrict__ active) {
    constexpr int NN = N * N;
    if (active && *active == 0) return;
    const int k = xcd_index();
    if (k >= ncells) return;
    const size_t n = (size_t)nT, m = (size_t)nC;
    const int p = p0 + k;
    int code[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) code[s] = nc[s * m + p];
    double acc[N];
    if (FORWARD) {
        const int i = cells[p];
        for (int r = 0; r < N; ++r) acc[r] = b[r * n + i];
    } else {
        for (int r = 0; r < N; ++r) acc[r] = 0.0;
    }
    bool any = FORWARD;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (FORWARD ? code[s] <= 0 : code[s] >= 0) continue;
        any = true;
        const int pj = (FORWARD ? code[s] : -code[s]) - 1;   // the neighbour's position
        double yj[N];
        for (int r = 0; r < N; ++r) yj[r] = yp[r * m + pj];
        const double* A = Ec + (size_t)s * NN * m + p;
        for (int r = 0; r < N; ++r) {
            double t = 0.0;
            for (int q = 0; q < N; ++q) t += A[(r * N + q) * m] * yj[q];
            acc[r] += (FORWARD ? -1.0 : 1.0) * t;
        }
    }
    if (!any) return;
    for (int r = 0; r < N; ++r) {
        double t = 0.0;
        for (int q = 0; q < N; ++q) t += Dc[(r * N + q) * m + p] * acc[q];
        if (FORWARD) yp[r * m + p] = t;
        else yp[r * m + p] -= t;
    }
}

// z_i = yp at the position of real cell i (the sweeps' result in cell order)
__global__ void __launch_bounds__(kBlock) k_dilu_unpermute(int32_t nC, int32_t nT, int nv,
                                                           const int32_t* __restrict__ pos,
                                                           const double* __restrict__ yp, double* __restrict__ z,
                                                           const int* __restrict__ active) {
    if (active && *active == 0) return;
    const int i = xcd_index();
    if (i >= nC) return;
    const int p = pos[i];
    for (int r = 0; r < nv; ++r) z[(size_t)r * nT + i] = yp[(size_t)r * nC + p];
}

// ghost rows are identity rows: z = x there
__global__ void __launch_bounds__(kBlock) k_copy_ghost_rows(int32_t nC, int32_t nT, int nv, const double* __restrict__ x,
                                                            double* __restrict__ z, const int* __restrict__ active) {
    if (active && *active == 0) return;
    const int c = nC + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nT) return;
    for (int k = 0; k < nv; ++k) z[(size_t)k * nT + c] = x[(size_t)k * nT + c];
}

// --------------------------------------------- device-resident GMRES (f1) --
// One Arnoldi step j with classical Gram-Schmidt applied twice, as three
// passes over the basis: (1) h1 = V^T w; (2) w -= V h1 and h2 = V^T w in the
// same pass; (3) w -= V h2 and |w|^2.  Each pass writes per-block partial
// sums that k_gm_reduce adds up.
constexpr int kGsMax = 32;

template <bool UPDATE, bool DOTS, int MM>
__global__ void __launch_bounds__(kBlock) k_gs_pass(const double* __restrict__ V, int64_t ld, int m,
                                                    const double* __restrict__ h, double* __restrict__ w, int64_t n,
                                                    double* __restrict__ partial, const int* __restrict__ active) {
    if (active && *active == 0) return;
    __shared__ double red[kBlock / 64][MM + 1];
    double acc[MM + 1];
#pragma unroll
    for (int k = 0; k <= MM; ++k) acc[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v[MM];
        double wi = w[i];
#pragma unroll
        for (int k = 0; k < MM; ++k) {
            if (k < m && (UPDATE || DOTS)) {
                v[k] = V[(size_t)k * ld + i];
                if (UPDATE) wi -= h[k] * v[k];
            }
        }
        if (UPDATE) w[i] = wi;
#pragma unroll
        for (int k = 0; k < MM; ++k)
            if (DOTS && k < m) acc[k] += v[k] * wi;
        acc[MM] += wi * wi;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= MM; ++k) {
        if (k == MM || (DOTS && k < m)) {
            double t = acc[k];
            for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
            if (lane == 0) red[wave][k] = t;
        }
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k <= m && (DOTS || k == m)) {
        const int src = k == m ? MM : k;
        double t = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) t += red[q][src];
        partial[(size_t)k * gridDim.x + blockIdx.x] = t;
    }
}

// out[k] = sum_b partial[(first + k) * nblk + b], k < cnt; one block per k
__global__ void __launch_bounds__(kBlock) k_gm_reduce(const double* __restrict__ partial, int nblk, int first, int cnt,
                                                      double* __restrict__ out, const int* __restrict__ active) {
    if (active && *active == 0) return;
    // one block per k: every thread sums a strided share of the nblk partials
    __shared__ double red[kBlock / 64];
    const int k = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (k >= cnt) return;
    const double* p = partial + (size_t)(first + k) * nblk;
    double t = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) t += p[b];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (lane == 0) red[wave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) r += red[q];
        out[k] = r;
    }
}

// start of a cycle from the residual norm^2 in hdr[3]: the first cycle sets
// beta0 (zero right-hand side: converged at once); a restart tests the true
// residual against tol beta0
__global__ void k_gm_begin(GmDev G, int first) {
    if (threadIdx.x != 0) return;
    const double beta = __builtin_sqrt(G.hdr[3]);
    if (first) { G.hdr[0] = beta; G.flags[1] = 0; G.flags[2] = 0; }
    G.hdr[1] = beta;
    G.flags[3] = 0;
    for (int k = 0; k <= G.m; ++k) G.g[k] = 0.0;
    G.g[0] = beta;
    const bool done = first ? !(beta > 0.0) : (beta <= G.tol * G.hdr[0]);
    if (done) { G.flags[1] = 1; G.flags[0] = 0; return; }
    G.flags[0] = 1;
    G.hdr[2] = 1.0 / beta;
}

// Hessenberg column j from the two passes and |w|, the previous Givens
// rotations and a new one, the residual estimate |g_{j+1}|, the flags and the
// scale of the next basis vector
__global__ void k_gm_step(GmDev G, int j) {
    if (threadIdx.x != 0 || G.flags[0] == 0) return;
    const int m = G.m;
    double* H = G.H;
    for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = G.h1[i] + G.h2[i];
    const double nrm = __builtin_sqrt(G.hdr[3] > 0.0 ? G.hdr[3] : 0.0);
    H[(size_t)(j + 1) * m + j] = nrm;
    for (int i = 0; i < j; ++i) {
        const double a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = G.cs[i] * a + G.sn[i] * bb;
        H[(size_t)(i + 1) * m + j] = -G.sn[i] * a + G.cs[i] * bb;
    }
    const double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
    const double r = hypot(a, bb);
    G.cs[j] = r > 0 ? a / r : 1.0;
    G.sn[j] = r > 0 ? bb / r : 0.0;
    H[(size_t)j * m + j] = r;
    H[(size_t)(j + 1) * m + j] = 0.0;
    G.g[j + 1] = -G.sn[j] * G.g[j];
    G.g[j] = G.cs[j] * G.g[j];
    G.flags[2] += 1;
    G.flags[3] = j + 1;
    if (fabs(G.g[j + 1]) <= G.tol * G.hdr[0]) { G.flags[1] = 1; G.flags[0] = 0; return; }
    if (!(nrm > 0) || G.flags[2] >= G.maxit) { G.flags[0] = 0; return; }
    G.hdr[2] = 1.0 / nrm;
}

// y = H^-1 g over the cycle's jend columns (upper triangular)
__global__ void k_gm_solve(GmDev G) {
    if (threadIdx.x != 0) return;
    const int m = G.m, jend = G.flags[3];
    for (int i = jend - 1; i >= 0; --i) {
        double t = G.g[i];
        for (int k = i + 1; k < jend; ++k) t -= G.H[(size_t)i * m + k] * G.y[k];
        G.y[i] = t / G.H[(size_t)i * m + i];
    }
}

// y = a x with the scale a on the device
__global__ void __launch_bounds__(kBlock) k_scale_dev(const double* __restrict__ a, const double* __restrict__ x,
                                                      double* __restrict__ y, int64_t n, const int* __restrict__ active) {
    if (active && *active == 0) return;
    const double s = *a;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = s * x[i];
}

// x = sum_{k < jend} y_k V_k, jend on the device
__global__ void __launch_bounds__(kBlock) k_lincomb_dev(const double* __restrict__ V, int64_t ld, const int* __restrict__ jend,
                                                        const double* __restrict__ y, double* __restrict__ x, int64_t n) {
    const int m = *jend;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < m; ++k) acc += y[k] * V[(size_t)k * ld + i];
    x[i] = acc;
}

// r = b - w
__global__ void __launch_bounds__(kBlock) k_sub(const double* __restrict__ b, const double* __restrict__ w,
                                                double* __restrict__ r, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - w[i];
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
                                                     double* __restrict__ y, const int* __restrict__ active) {
    if (active && *active == 0) return;
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
    hipLaunchKernelGGL(k_face_jacobian, dim3(xcd_grid(F.nF)), dim3(kBlock), 0, s, C, F, q, gx, gy, g, viscous, Jf);
    return hipGetLastError();
}
hipError_t launch_face_jacobian_sa(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* gx,
                                   const double* gy, GasT<double> g, double* Jf) {
    hipLaunchKernelGGL(k_face_jacobian_sa, dim3(xcd_grid(F.nF)), dim3(kBlock), 0, s, C, F, q, gx, gy, g, Jf);
    return hipGetLastError();
}
hipError_t launch_assemble_diag(hipStream_t s, int nv, CellsT<double> C, FacesT<double> F, const double* dt,
                                const double* Jf, SaSource src, double* D, double* Dinv, double* E, double* Ds,
                                int32_t* nbr, int* bad) {
    if (nv == 5)
        hipLaunchKernelGGL(k_assemble_diag<5>, dim3(xcd_grid(C.nC)), dim3(kBlock), 0, s, C, F, dt, Jf, src, D, Dinv, E,
                           Ds, nbr, bad);
    else if (nv == 4)
        hipLaunchKernelGGL(k_assemble_diag<4>, dim3(xcd_grid(C.nC)), dim3(kBlock), 0, s, C, F, dt, Jf, src, D, Dinv, E,
                           Ds, nbr, bad);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}
hipError_t launch_lhs_apply(hipStream_t s, int nv, int32_t nC, int32_t nT, const double* Ds, const double* E,
                            const int32_t* nbr, const double* x, double* y, const int* active) {
    if (nv == 5)
        hipLaunchKernelGGL(k_lhs_apply<5>, dim3(xcd_grid(nT)), dim3(kBlock), 0, s, nC, nT, Ds, E, nbr, x, y, active);
    else if (nv == 4)
        hipLaunchKernelGGL(k_lhs_apply<4>, dim3(xcd_grid(nT)), dim3(kBlock), 0, s, nC, nT, Ds, E, nbr, x, y, active);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}
hipError_t launch_block_jacobi(hipStream_t s, int nv, int32_t nC, int32_t nT, const double* Dinv, const double* x,
                               double* z, const int* active) {
    if (nv == 5)
        hipLaunchKernelGGL(k_block_jacobi<5>, dim3(nblocks(nT)), dim3(kBlock), 0, s, nC, nT, Dinv, x, z, active);
    else if (nv == 4)
        hipLaunchKernelGGL(k_block_jacobi<4>, dim3(nblocks(nT)), dim3(kBlock), 0, s, nC, nT, Dinv, x, z, active);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}
template <int N>
static hipError_t dilu_setup_n(hipStream_t s, CellsT<double> C, FacesT<double> F, const uint8_t* colour,
                               const int32_t* cells, const int32_t* pos, const int32_t* coff, int ncolours,
                               const double* D, const double* Jf, double* Dti, double* Ec, int32_t* nc, double* Dc,
                               int* bad) {
    for (int c = 0; c < ncolours; ++c) {
        const int32_t n = coff[c + 1] - coff[c];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_dilu_setup<N>, dim3(xcd_grid(n)), dim3(kBlock), 0, s, C, F, colour, c, cells, pos,
                           coff[c], n, D, Jf, Dti, Ec, nc, Dc, bad);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t launch_dilu_setup(hipStream_t s, int nv, CellsT<double> C, FacesT<double> F, const uint8_t* colour,
                             const int32_t* cells, const int32_t* pos, const int32_t* coff, int ncolours,
                             const double* D, const double* Jf, double* Dti, double* Ec, int32_t* nc, double* Dc,
                             int* bad) {
    if (nv == 5) return dilu_setup_n<5>(s, C, F, colour, cells, pos, coff, ncolours, D, Jf, Dti, Ec, nc, Dc, bad);
    if (nv == 4) return dilu_setup_n<4>(s, C, F, colour, cells, pos, coff, ncolours, D, Jf, Dti, Ec, nc, Dc, bad);
    return hipErrorInvalidValue;
}
template <int N>
static hipError_t dilu_apply_n(hipStream_t s, int32_t nC, int32_t nT, const int32_t* cells, const int32_t* pos,
                               const int32_t* coff, int ncolours, const double* Ec, const int32_t* nc, const double* Dc,
                               const double* x, double* z, double* yp, const int* active) {
    hipError_t e;
    if (nT > nC) {
        hipLaunchKernelGGL(k_copy_ghost_rows, dim3(nblocks(nT - nC)), dim3(kBlock), 0, s, nC, nT, N, x, z, active);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (int c = 0; c < ncolours; ++c) {
        const int32_t n = coff[c + 1] - coff[c];
        if (n == 0) continue;
        hipLaunchKernelGGL((k_dilu_sweep<N, true>), dim3(xcd_grid(n)), dim3(kBlock), 0, s, nC, nT, coff[c], n, cells,
                           Ec, nc, Dc, x, yp, active);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (int c = ncolours - 2; c >= 0; --c) {
        const int32_t n = coff[c + 1] - coff[c];
        if (n == 0) continue;
        hipLaunchKernelGGL((k_dilu_sweep<N, false>), dim3(xcd_grid(n)), dim3(kBlock), 0, s, nC, nT, coff[c], n, cells,
                           Ec, nc, Dc, x, yp, active);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (nC > 0) {
        hipLaunchKernelGGL(k_dilu_unpermute, dim3(xcd_grid(nC)), dim3(kBlock), 0, s, nC, nT, N, pos, yp, z, active);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t launch_dilu_apply(hipStream_t s, int nv, int32_t nC, int32_t nT, const int32_t* cells, const int32_t* pos,
                             const int32_t* coff, int ncolours, const double* Ec, const int32_t* nc, const double* Dc,
                             const double* x, double* z, double* yp, const int* active) {
    if (nv == 5) return dilu_apply_n<5>(s, nC, nT, cells, pos, coff, ncolours, Ec, nc, Dc, x, z, yp, active);
    if (nv == 4) return dilu_apply_n<4>(s, nC, nT, cells, pos, coff, ncolours, Ec, nc, Dc, x, z, yp, active);
    return hipErrorInvalidValue;
}
template <int MM>
hipError_t gs_pass_variant(hipStream_t s, const double* V, int64_t ld, int m, const double* h, double* w, int64_t n,
                           bool dots, double* partial, int nblk, const int* active) {
    if (h && dots)
        hipLaunchKernelGGL((k_gs_pass<true, true, MM>), dim3(nblk), dim3(kBlock), 0, s, V, ld, m, h, w, n, partial,
                           active);
    else if (h)
        hipLaunchKernelGGL((k_gs_pass<true, false, MM>), dim3(nblk), dim3(kBlock), 0, s, V, ld, m, h, w, n, partial,
                           active);
    else if (dots)
        hipLaunchKernelGGL((k_gs_pass<false, true, MM>), dim3(nblk), dim3(kBlock), 0, s, V, ld, m, h, w, n, partial,
                           active);
    else
        hipLaunchKernelGGL((k_gs_pass<false, false, MM>), dim3(nblk), dim3(kBlock), 0, s, V, ld, m, h, w, n, partial,
                           active);
    return hipGetLastError();
}
hipError_t launch_gs_pass(hipStream_t s, const double* V, int64_t ld, int m, const double* h, double* w, int64_t n,
                          bool dots, double* partial, int nblk, const int* active) {
    if (m > kGsMax) return hipErrorInvalidValue;
    // the variant with the fewest registers that holds m columns: the small
    // bases of short restart cycles keep more waves in flight per SIMD
    if (m <= 8) return gs_pass_variant<8>(s, V, ld, m, h, w, n, dots, partial, nblk, active);
    if (m <= 16) return gs_pass_variant<16>(s, V, ld, m, h, w, n, dots, partial, nblk, active);
    return gs_pass_variant<kGsMax>(s, V, ld, m, h, w, n, dots, partial, nblk, active);
}
hipError_t launch_gm_reduce(hipStream_t s, const double* partial, int nblk, int first, int cnt, double* out,
                            const int* active) {
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gm_reduce, dim3(cnt), dim3(kBlock), 0, s, partial, nblk, first, cnt, out, active);
    return hipGetLastError();
}
hipError_t launch_gm_begin(hipStream_t s, GmDev G, int first) {
    hipLaunchKernelGGL(k_gm_begin, dim3(1), dim3(64), 0, s, G, first);
    return hipGetLastError();
}
hipError_t launch_gm_step(hipStream_t s, GmDev G, int j) {
    hipLaunchKernelGGL(k_gm_step, dim3(1), dim3(64), 0, s, G, j);
    return hipGetLastError();
}
hipError_t launch_gm_solve(hipStream_t s, GmDev G) {
    hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, s, G);
    return hipGetLastError();
}
hipError_t launch_scale_dev(hipStream_t s, const double* a, const double* x, double* y, int64_t n, const int* active) {
    hipLaunchKernelGGL(k_scale_dev, dim3(nblocks(n)), dim3(kBlock), 0, s, a, x, y, n, active);
    return hipGetLastError();
}
hipError_t launch_lincomb_dev(hipStream_t s, const double* V, int64_t ld, const int* jend, const double* y, double* x,
                              int64_t n) {
    hipLaunchKernelGGL(k_lincomb_dev, dim3(nblocks(n)), dim3(kBlock), 0, s, V, ld, jend, y, x, n);
    return hipGetLastError();
}
hipError_t launch_sub(hipStream_t s, const double* b, const double* w, double* r, int64_t n) {
    hipLaunchKernelGGL(k_sub, dim3(nblocks(n)), dim3(kBlock), 0, s, b, w, r, n);
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
                           const double* Fb, GasT<double> g, int viscous, const double* x, double* y, const int* active) {
    hipLaunchKernelGGL(k_mf_apply, dim3(nblocks(C.nT)), dim3(kBlock), 0, s, C, F, q, dt, Fb, g, viscous, x, y, active);
    return hipGetLastError();
}

}  // namespace afx
