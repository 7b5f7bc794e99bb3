// FP64 instantiation of the residual kernels (plain form; built with
// -ffp-contract=off so every operation rounds as in the reference's serial
// CPU code) and the FP64 pseudo-time / reduction helpers.
#include "afx_device.hpp"

namespace afx {
AFX_INSTANTIATE_LAUNCHERS(double)

// sum of squares over nv rows of length n (row stride ld); one partial per block
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sumsq(const T* __restrict__ x, int64_t n, int64_t ld, int nv,
                                                  double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    double s = 0.0;
    const int64_t total = n * nv;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = i / n, j = i - k * n;
        double v = (double)x[k * ld + j];
        s += v * v;
    }
    // wavefront reduction (64 lanes) then across the block's wavefronts
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        partial[blockIdx.x] = t;
    }
}

__global__ void k_final_sum(const double* __restrict__ partial, int nblk, double* __restrict__ out) {
    __shared__ double red[kBlock / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += partial[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        out[0] = t;
    }
}

hipError_t launch_sumsq(hipStream_t s, const double* x, int64_t n, int64_t ld, int nv, double* partial, int nblk) {
    hipLaunchKernelGGL(k_sumsq<double>, dim3(nblk), dim3(kBlock), 0, s, x, n, ld, nv, partial);
    return hipGetLastError();
}
hipError_t launch_sumsq_f32(hipStream_t s, const float* x, int64_t n, int64_t ld, int nv, double* partial, int nblk) {
    hipLaunchKernelGGL(k_sumsq<float>, dim3(nblk), dim3(kBlock), 0, s, x, n, ld, nv, partial);
    return hipGetLastError();
}
hipError_t launch_final_sum(hipStream_t s, const double* partial, int nblk, double* out) {
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, s, partial, nblk, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_copy(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}
hipError_t launch_copy(hipStream_t s, const double* src, double* dst, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, n);
    return hipGetLastError();
}

// explicitSolver::solve stage update (solver.h:818-821)
__global__ void __launch_bounds__(kBlock) k_rk_update(const double* __restrict__ q, const double* __restrict__ qW,
                                                      const double* __restrict__ dt, double ai, double relax, int32_t nC,
                                                      int32_t nT, int nv, double* __restrict__ qk) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    for (int k = 0; k < nv; ++k) {
        size_t i = (size_t)k * nT + c;
        qk[i] = q[i] + qW[i] * dt[c] * ai * relax;
    }
}
hipError_t launch_rk_update(hipStream_t s, const double* q, const double* qW, const double* dt, double ai, double relax,
                            int32_t nC, int32_t nT, int nv, double* qk) {
    hipLaunchKernelGGL(k_rk_update, dim3(nblocks(nC)), dim3(kBlock), 0, s, q, qW, dt, ai, relax, nC, nT, nv, qk);
    return hipGetLastError();
}

// explicit stage update from an undivided residual (R / A as qW of
// solver.h:790-792), for the Spalart-Allmaras explicit step
__global__ void __launch_bounds__(kBlock) k_rk_update_area(const double* __restrict__ q, const double* __restrict__ R,
                                                           const double* __restrict__ area,
                                                           const double* __restrict__ dt, double coef, int32_t nC,
                                                           int32_t nT, int nv, double* __restrict__ qk,
                                                           double* __restrict__ qW) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    const double ia = 1.0 / area[c];
    for (int k = 0; k < nv; ++k) {
        size_t i = (size_t)k * nT + c;
        const double w = R[i] * ia;
        qW[i] = w;
        qk[i] = q[i] + w * dt[c] * coef;
    }
}
hipError_t launch_rk_update_area(hipStream_t s, const double* q, const double* R, const double* area, const double* dt,
                                 double coef, int32_t nC, int32_t nT, int nv, double* qk, double* qW) {
    hipLaunchKernelGGL(k_rk_update_area, dim3(nblocks(nC)), dim3(kBlock), 0, s, q, R, area, dt, coef, nC, nT, nv, qk,
                       qW);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
hipError_t launch_to_f32(hipStream_t s, const double* src, float* dst, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, n);
    return hipGetLastError();
}

// SoA FP64 state -> AoS float4 copy used by the FP32 fused residual
__global__ void __launch_bounds__(kBlock) k_to_f32x4(const double* __restrict__ a, int32_t nT, float4* __restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nT) b[i] = make_float4((float)a[i], (float)a[(size_t)nT + i], (float)a[2 * (size_t)nT + i], (float)a[3 * (size_t)nT + i]);
}
hipError_t launch_to_f32x4(hipStream_t s, const double* src, int32_t nT, float4* dst) {
    hipLaunchKernelGGL(k_to_f32x4, dim3(nblocks(nT)), dim3(kBlock), 0, s, src, nT, dst);
    return hipGetLastError();
}

// mesh.h:794-830: minimum distance to the wall/slip-wall face centres (LDS tiles)
__global__ void __launch_bounds__(kBlock) k_wall_dist(const double* __restrict__ xc, const double* __restrict__ yc,
                                                      int32_t nT, const double* __restrict__ wx,
                                                      const double* __restrict__ wy, int32_t nw, double* __restrict__ d) {
    __shared__ double sx[kBlock], sy[kBlock];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double cx = c < nT ? xc[c] : 0.0, cy = c < nT ? yc[c] : 0.0;
    double mind = 1.0;
    bool first = true;
    for (int base = 0; base < nw; base += kBlock) {
        int j = base + threadIdx.x;
        if (j < nw) { sx[threadIdx.x] = wx[j]; sy[threadIdx.x] = wy[j]; }
        __syncthreads();
        int m = nw - base < kBlock ? nw - base : kBlock;
        for (int t = 0; t < m; ++t) {
            double dx = cx - sx[t], dy = cy - sy[t];
            double dd = __builtin_sqrt(dx * dx + dy * dy);
            if (first) { mind = dd; first = false; }
            else mind = (dd < mind) ? dd : mind;
        }
        __syncthreads();
    }
    if (c < nT) d[c] = mind;
}
hipError_t launch_wall_dist(hipStream_t s, const double* xc, const double* yc, int32_t nT, const double* wx,
                            const double* wy, int32_t nw, double* d) {
    hipLaunchKernelGGL(k_wall_dist, dim3(nblocks(nT)), dim3(kBlock), 0, s, xc, yc, nT, wx, wy, nw, d);
    return hipGetLastError();
}

// The Spalart-Allmaras model's wall distance: from each cell centroid to the
// nearest point of the wall / slip-wall face segments (a + t (b - a), t
// clamped to [0, 1]; seg = {ax, ay, bx - ax, by - ay} per face), 1 without
// such faces.  Segments staged through LDS tiles as in k_wall_dist.
__global__ void __launch_bounds__(kBlock) k_wall_dist_seg(const double* __restrict__ xc, const double* __restrict__ yc,
                                                          int32_t nT, const double4* __restrict__ seg, int32_t nw,
                                                          double* __restrict__ d) {
    __shared__ double4 ss[kBlock];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const double px = c < nT ? xc[c] : 0.0, py = c < nT ? yc[c] : 0.0;
    double mind = 1.0;
    bool first = true;
    for (int base = 0; base < nw; base += kBlock) {
        const int j = base + threadIdx.x;
        if (j < nw) ss[threadIdx.x] = seg[j];
        __syncthreads();
        const int m = nw - base < kBlock ? nw - base : kBlock;
        for (int k = 0; k < m; ++k) {
            const double4 g = ss[k];
            const double ll = g.z * g.z + g.w * g.w;
            double t = ((px - g.x) * g.z + (py - g.y) * g.w) / ll;
            t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
            const double ex = px - (g.x + t * g.z), ey = py - (g.y + t * g.w);
            const double dd = __builtin_sqrt(ex * ex + ey * ey);
            if (first) { mind = dd; first = false; }
            else mind = (dd < mind) ? dd : mind;
        }
        __syncthreads();
    }
    if (c < nT) d[c] = mind;
}
hipError_t launch_wall_dist_seg(hipStream_t s, const double* xc, const double* yc, int32_t nT, const double4* seg,
                                int32_t nw, double* d) {
    hipLaunchKernelGGL(k_wall_dist_seg, dim3(nblocks(nT)), dim3(kBlock), 0, s, xc, yc, nT, seg, nw, d);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_gather_states(const double* __restrict__ q, int32_t nT, int nv,
                                                          const int32_t* __restrict__ cells, int32_t n,
                                                          double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = cells[i];
    for (int k = 0; k < nv; ++k) out[(size_t)i * nv + k] = q[(size_t)k * nT + c];
}
hipError_t launch_gather_states(hipStream_t s, const double* q, int32_t nT, int nv, const int32_t* cells, int32_t n,
                                double* out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_states, dim3(nblocks(n)), dim3(kBlock), 0, s, q, nT, nv, cells, n, out);
    return hipGetLastError();
}

}  // namespace afx
