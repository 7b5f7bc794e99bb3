// Level prolongation of the full-multigrid start (multigrid<T>::gen_mapper,
// multigrid.h:96-163, and the product of multigrid.h:334):
//
//   fine cell i (real or ghost) takes every coarse cell j (real or ghost) whose
//   centroid lies within d2 < 2 A_j, with weight s_ij / sum_j s_ij,
//   s_ij = 1 / max(0.1 sqrt(A_j), sqrt(d2)); q_fine = M q_coarse.
//
// The reference tests all N_fine x N_coarse pairs on one core.  Here one
// workgroup takes 256 consecutive fine cells, forms their bounding box, and
// walks the coarse cells in tiles of 256 in ascending order, skipping any tile
// whose box (centroids grown by sqrt(2 A_j)) misses it; surviving tiles are
// staged in LDS and tested by every thread.  Matches therefore come out in
// ascending coarse index, the order in which the reference sums the scales and
// Eigen's column-major product sums each row.  A count pass sizes the CSR
// rows, a fill pass writes (j, s_ij) and normalises its row.
#include "afx_device.hpp"
#include "afx_mapper.hpp"

namespace afx {

namespace {
constexpr int kMapBlock = 256;

// conservative reach of coarse cell j: d2 < 2 A implies |dx|, |dy| below it
__device__ __forceinline__ double reach(double a) { return __builtin_sqrt(2.0 * a) * (1.0 + 1e-9) + 1e-300; }
}  // namespace

__global__ void __launch_bounds__(kMapBlock) k_map_tile_boxes(const double* __restrict__ cx, const double* __restrict__ cy,
                                                              const double* __restrict__ ca, int32_t nc,
                                                              double* __restrict__ box) {
    __shared__ double red[4][kMapBlock];
    const int t = blockIdx.x, j = t * kMapBlock + threadIdx.x;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    if (j < nc) {
        const double r = reach(ca[j]);
        x0 = cx[j] - r; x1 = cx[j] + r; y0 = cy[j] - r; y1 = cy[j] + r;
    }
    red[0][threadIdx.x] = x0; red[1][threadIdx.x] = -x1; red[2][threadIdx.x] = y0; red[3][threadIdx.x] = -y1;
    __syncthreads();
    for (int w = kMapBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        box[4 * t + 0] = red[0][0]; box[4 * t + 1] = -red[1][0];
        box[4 * t + 2] = red[2][0]; box[4 * t + 3] = -red[3][0];
    }
}

// FILL = false: count[i] = number of coarse cells taken by fine cell i.
// FILL = true: cols/vals of row i at rowptr[i], weights normalised.
template <bool FILL>
__global__ void __launch_bounds__(kMapBlock) k_map_search(const double* __restrict__ fx, const double* __restrict__ fy,
                                                          int32_t nf, const double* __restrict__ cx,
                                                          const double* __restrict__ cy, const double* __restrict__ ca,
                                                          int32_t nc, const double* __restrict__ box, int32_t* __restrict__ count,
                                                          const int64_t* __restrict__ rowptr, int32_t* __restrict__ cols,
                                                          double* __restrict__ vals) {
    __shared__ double sx[kMapBlock], sy[kMapBlock], sa[kMapBlock];
    __shared__ double red[4][kMapBlock];
    const int i = blockIdx.x * kMapBlock + threadIdx.x;
    const bool live = i < nf;
    const double xi = live ? fx[i] : 0.0, yi = live ? fy[i] : 0.0;
    red[0][threadIdx.x] = live ? xi : INFINITY;
    red[1][threadIdx.x] = live ? -xi : INFINITY;
    red[2][threadIdx.x] = live ? yi : INFINITY;
    red[3][threadIdx.x] = live ? -yi : INFINITY;
    __syncthreads();
    for (int w = kMapBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        __syncthreads();
    }
    const double bx0 = red[0][0], bx1 = -red[1][0], by0 = red[2][0], by1 = -red[3][0];
    int64_t pos = 0, p0 = 0;
    if (FILL && live) pos = p0 = rowptr[i];
    int32_t cnt = 0;
    const int ntiles = (nc + kMapBlock - 1) / kMapBlock;
    for (int t = 0; t < ntiles; ++t) {
        // uniform across the workgroup: every thread reads the same box
        if (box[4 * t + 1] < bx0 || box[4 * t] > bx1 || box[4 * t + 3] < by0 || box[4 * t + 2] > by1) continue;
        const int j0 = t * kMapBlock;
        const int jn = min(kMapBlock, nc - j0);
        __syncthreads();
        if ((int)threadIdx.x < jn) {
            sx[threadIdx.x] = cx[j0 + threadIdx.x];
            sy[threadIdx.x] = cy[j0 + threadIdx.x];
            sa[threadIdx.x] = ca[j0 + threadIdx.x];
        }
        __syncthreads();
        if (!live) continue;
        for (int k = 0; k < jn; ++k) {
            const double dx = xi - sx[k], dy = yi - sy[k];
            const double d2 = dx * dx + dy * dy;
            const double r2 = sa[k];
            if (d2 < 2 * r2) {
                if (FILL) {
                    const double a = 0.1 * __builtin_sqrt(r2), b = __builtin_sqrt(d2);
                    cols[pos] = j0 + k;
                    vals[pos] = 1 / ((a < b) ? b : a);   // 1 / std::max(a, b)
                    ++pos;
                } else {
                    ++cnt;
                }
            }
        }
    }
    if (!live) return;
    if (!FILL) {
        count[i] = cnt;
        return;
    }
    double scale = 0;
    for (int64_t p = p0; p < pos; ++p) scale += vals[p];
    for (int64_t p = p0; p < pos; ++p) vals[p] = vals[p] / scale;
}

// coarse column index in reference order -> the coarse handle's device order
__global__ void __launch_bounds__(kBlock) k_map_cols(const int32_t* __restrict__ cols, const int32_t* __restrict__ old2new,
                                                     int64_t nnz, int32_t* __restrict__ cols_dev) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) cols_dev[p] = old2new[cols[p]];
}

// q_fine[k][i] = sum_p vals[p] q_coarse[k][cols_dev[p]] over the reference row
// of fine device cell i, in ascending coarse order (Eigen's sparse product)
__global__ void __launch_bounds__(kBlock) k_map_prolong(const int64_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ fnew2old,
                                                        const int32_t* __restrict__ cols_dev,
                                                        const double* __restrict__ vals, const double* __restrict__ qc,
                                                        int64_t ntc, double* __restrict__ qf, int32_t ntf, int nv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntf) return;
    const int r = fnew2old[i];
    const int64_t s = rowptr[r], e = rowptr[r + 1];
    for (int k = 0; k < nv; ++k) {
        double acc = 0;
        for (int64_t p = s; p < e; ++p) acc += vals[p] * qc[k * ntc + cols_dev[p]];
        qf[(int64_t)k * ntf + i] = acc;
    }
}

hipError_t launch_map_tile_boxes(hipStream_t s, const double* cx, const double* cy, const double* ca, int32_t nc,
                                 double* box) {
    const int nt = (nc + kMapBlock - 1) / kMapBlock;
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_map_tile_boxes, dim3(nt), dim3(kMapBlock), 0, s, cx, cy, ca, nc, box);
    return hipGetLastError();
}
hipError_t launch_map_count(hipStream_t s, const double* fx, const double* fy, int32_t nf, const double* cx,
                            const double* cy, const double* ca, int32_t nc, const double* box, int32_t* count) {
    const int nb = (nf + kMapBlock - 1) / kMapBlock;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_map_search<false>, dim3(nb), dim3(kMapBlock), 0, s, fx, fy, nf, cx, cy, ca, nc, box, count,
                       (const int64_t*)nullptr, (int32_t*)nullptr, (double*)nullptr);
    return hipGetLastError();
}
hipError_t launch_map_fill(hipStream_t s, const double* fx, const double* fy, int32_t nf, const double* cx,
                           const double* cy, const double* ca, int32_t nc, const double* box, const int64_t* rowptr,
                           int32_t* cols, double* vals) {
    const int nb = (nf + kMapBlock - 1) / kMapBlock;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_map_search<true>, dim3(nb), dim3(kMapBlock), 0, s, fx, fy, nf, cx, cy, ca, nc, box,
                       (int32_t*)nullptr, rowptr, cols, vals);
    return hipGetLastError();
}
hipError_t launch_map_cols(hipStream_t s, const int32_t* cols, const int32_t* old2new, int64_t nnz, int32_t* cols_dev) {
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(k_map_cols, dim3(nblocks(nnz)), dim3(kBlock), 0, s, cols, old2new, nnz, cols_dev);
    return hipGetLastError();
}
hipError_t launch_map_prolong(hipStream_t s, const int64_t* rowptr, const int32_t* fnew2old, const int32_t* cols_dev,
                              const double* vals, const double* qc, int64_t ntc, double* qf, int32_t ntf, int nv) {
    hipLaunchKernelGGL(k_map_prolong, dim3(nblocks(ntf)), dim3(kBlock), 0, s, rowptr, fnew2old, cols_dev, vals, qc,
                       ntc, qf, ntf, nv);
    return hipGetLastError();
}

}  // namespace afx
