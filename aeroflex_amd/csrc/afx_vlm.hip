// Vortex-lattice aerodynamic influence matrix and its dense solve on MI355X
// (include/afx_vlm.h).
//
// Device side of the coupling path of AeroFLEX's src/vlm:
//   - panel geometry, collocation points, right-hand side and wake rings
//     (panel.cpp:46-128, surface.cpp:35-58, solver.cpp:256-290);
//   - the FP64 influence matrix, one thread per matrix entry, four Biot-Savart
//     segments per ring (potential.cpp:36-58, solver.cpp:139-254);
//   - FP32 LU with partial pivoting: a cooperative panel kernel (rows of the
//     panel in LDS, pivot keys exchanged between workgroups), row
//     interchanges, a wavefront triangular solve for the block row and the
//     trailing update as a hand-written FP32 GEMM on the matrix cores
//     (k_gemm_sub_f32, v_mfma_f32_32x32x2f32);
//   - FP64 iterative refinement on the FP64 matrix with the FP32 factor as
//     the preconditioner (replaces Eigen's BiCGSTAB, solver.cpp:41-50);
//   - Kutta-Joukowski ring forces (surface.cpp:164-201).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/afx_vlm.h"
#include "afx_vlm.hpp"

using afx::api_fail;

namespace {

#define VCHK(x)                                                                                     \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return api_fail(AFX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int LU_NB = 128;          // panel width = trailing-update depth = solve block
constexpr int PANEL_R = 120;        // panel rows per workgroup held in LDS (120 x 128 FP32 = 60 KiB)
constexpr int MAX_PANEL_WG = 512;   // -> n <= 61440
constexpr unsigned SPIN_LIMIT = 1u << 21;
constexpr double FOUR_PI = 12.566370614359172;

// ------------------------------------------------- panel pivot exchange --
// Candidate key of one workgroup for one column: |a| bits (31) | row (17) |
// sequence (16).  The sequence is the 1-based column of the whole matrix, so
// a key left from an earlier column (or an earlier factorisation: the keys
// are cleared before each one) never matches.  Row 0x1ffff: no candidate.
constexpr unsigned NO_ROW = 0x1ffffu;
__device__ inline unsigned long long make_key(float absval, unsigned row, unsigned seq) {
    const unsigned long long ab = (unsigned long long)(__float_as_uint(absval) & 0x7fffffffu);
    return (ab << 33) | ((unsigned long long)(row & NO_ROW) << 16) | (unsigned long long)(seq & 0xffffu);
}

__device__ inline void argmax_combine(float& v, int& i, float ov, int oi) {
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}

// (|a|, row) as one ordered 64-bit key: larger |a| first, then the lower
// row; |a| < 0 (no candidate) maps to 0.  |a| >= 0 (or +inf) orders as its bits.
__device__ inline unsigned long long argmax_key(float v, int i) {
    return v < 0.f ? 0ull : ((unsigned long long)__float_as_uint(v) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)i);
}
__device__ inline void argmax_unkey(unsigned long long k, float& v, int& i) {
    v = k ? __uint_as_float((unsigned)(k >> 32)) : -1.f;
    i = k ? (int)(0xFFFFFFFFu - (unsigned)(k & 0xFFFFFFFFull)) : INT_MAX;
}
// one step of a wavefront max of 64-bit keys through DPP lane moves
template <int CTRL, int ROW_MASK>
__device__ inline unsigned long long dpp_max_step(unsigned long long k) {
    const int lo = (int)(unsigned)k, hi = (int)(unsigned)(k >> 32);
    const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
    const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
    const unsigned long long o = ((unsigned long long)(unsigned)ohi << 32) | (unsigned long long)(unsigned)olo;
    return o > k ? o : k;
}
// maximum over the wavefront's 64 lanes, returned in every lane: pairs and
// quads (quad_perm), half rows and rows (mirrors), then rows 1 and 3 from
// lane 15 of the row below (row_bcast15) and rows 2 and 3 from lane 31
// (row_bcast31); lane 63 then holds the maximum
__device__ inline unsigned long long wave_max_key(unsigned long long k) {
    k = dpp_max_step<0xB1, 0xF>(k);    // quad_perm [1,0,3,2]
    k = dpp_max_step<0x4E, 0xF>(k);    // quad_perm [2,3,0,1]
    k = dpp_max_step<0x141, 0xF>(k);   // row_half_mirror
    k = dpp_max_step<0x140, 0xF>(k);   // row_mirror
    k = dpp_max_step<0x142, 0xA>(k);   // row_bcast15
    k = dpp_max_step<0x143, 0xC>(k);   // row_bcast31
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)k, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(k >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ inline void wave_argmax(float& v, int& i) { argmax_unkey(wave_max_key(argmax_key(v, i)), v, i); }

// --------------------------------------------------- panel factorisation --
// Unblocked LU with partial pivoting of the m x nbc panel whose top-left
// entry is A(j0, j0) (column-major, leading dimension lda).  Workgroup w holds
// panel rows [w R, w R + R) of all nbc columns in LDS.  Per column every
// workgroup posts its largest |a| (ties: lowest row) as one 64-bit key after
// the row's values (and the owner of the diagonal row posts that row), then
// polls the keys of all workgroups: once every key carries this column's
// sequence, all workgroups know the same pivot without any counter, read the
// pivot row, perform the interchange on the rows they own, scale their part
// of the column and update their part of the panel.  The update is ordered so
// that the next column's key leaves early: column j + 1 is updated first, its
// candidate found, the two rows to be posted (candidate and diagonal) finish
// step j, the key is posted, and only then does the rest of the rank-1 update
// run, while the other workgroups' keys are in flight.  Every element still
// takes the same operations as in the plain right-looking order.  Keys and
// rows alternate between two buffers, so posting for column j + 1 cannot
// overwrite what a slower workgroup still reads for column j (it cannot post
// j + 2 before every workgroup has posted j + 1, which each does only after
// reading column j's pivot row).  Every wait is bounded: a workgroup that
// polls longer than SPIN_LIMIT raises `abort` and every workgroup returns.
__device__ inline unsigned long long tagged(unsigned seq, float v) {
    return ((unsigned long long)seq << 32) | (unsigned long long)__float_as_uint(v);
}
// wait (bounded) for a value | sequence word of this sequence; false on timeout
__device__ inline bool wait_tagged(const unsigned long long* p, unsigned seq, float& v) {
    unsigned spins = 0;
    unsigned long long w;
    while (((w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) != seq) {
        if (++spins > SPIN_LIMIT) { v = 0.f; return false; }
        __builtin_amdgcn_s_sleep(1);
    }
    v = __uint_as_float((unsigned)(w & 0xFFFFFFFFull));
    return true;
}

// the workgroup's candidate (largest |a|, lowest row) from one (|a|, row) per
// thread; leaves it in rv[0], ri[0] and returns its workgroup-local row, or -1
__device__ inline int block_candidate(float best, int bi, float* rv, int* ri, int* s_lrow, int rbase) {
    const int t = threadIdx.x;
    wave_argmax(best, bi);
    if ((t & 63) == 0) { rv[t >> 6] = best; ri[t >> 6] = bi; }
    __syncthreads();
    if (t == 0) {
        for (int q = 1; q < 4; ++q) argmax_combine(rv[0], ri[0], rv[q], ri[q]);
        *s_lrow = ri[0] != INT_MAX ? ri[0] - rbase : -1;
    }
    __syncthreads();
    return *s_lrow;
}

__global__ __launch_bounds__(256) void k_panel_lu(float* __restrict__ A, int lda, int m, int nbc, int j0,
                                                  int* __restrict__ ipiv, unsigned long long* keys,
                                                  unsigned long long* __restrict__ s_vals,
                                                  unsigned long long* __restrict__ s_diag,
                                                  unsigned* abort_flag, int* __restrict__ info) {
    __shared__ float slab[LU_NB * PANEL_R];   // panel column c at slab[c * PANEL_R]
    __shared__ float urow[LU_NB];
    __shared__ float rv[4];
    __shared__ int ri[4];
    __shared__ int s_lrow, s_win;
    const int t = threadIdx.x, w = blockIdx.x, nwg = gridDim.x;
    const int rbase = w * PANEL_R;
    const int rows = min(PANEL_R, m - rbase);
    float* P = A + (size_t)j0 * lda + j0;

    for (int idx = t; idx < nbc * PANEL_R; idx += 256) {
        const int c = idx / PANEL_R, r = idx - c * PANEL_R;
        slab[idx] = r < rows ? P[(size_t)c * lda + rbase + r] : 0.f;
    }
    __syncthreads();

    // column j's candidate row (local row lrow, or none) and, from the
    // workgroup holding panel row j, the diagonal row, as value | sequence
    // words (a reader checks each word's sequence, so no fence orders them
    // before the key), then the key
    auto post = [&](int j, int lrow) {
        const int par = j & 1;
        const unsigned seq = (unsigned)(j0 + j + 1);
        if (lrow >= 0)
            for (int c = t; c < nbc; c += 256)
                __hip_atomic_store(&s_vals[((size_t)par * MAX_PANEL_WG + w) * LU_NB + c],
                                   tagged(seq, slab[c * PANEL_R + lrow]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (w == j / PANEL_R)
            for (int c = t; c < nbc; c += 256)
                __hip_atomic_store(&s_diag[par * LU_NB + c], tagged(seq, slab[c * PANEL_R + (j - rbase)]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == 0) {
            const unsigned long long key =
                lrow >= 0 ? make_key(rv[0], (unsigned)ri[0], seq) : make_key(0.f, NO_ROW, seq);
            __hip_atomic_store(&keys[par * MAX_PANEL_WG + w], key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    // |a| of a candidate entry; a NaN is taken as the largest
    auto mag = [](float a) { const float v = fabsf(a); return v == v ? v : INFINITY; };

    {
        const bool live = t < rows;
        post(0, block_candidate(live ? mag(slab[t]) : -1.f, live ? rbase + t : INT_MAX, rv, ri, &s_lrow, rbase));
    }

    for (int j = 0; j < nbc; ++j) {
        const int par = j & 1;
        const unsigned seq = (unsigned)(j0 + j + 1);

        // gather: every key of this column's sequence
        float best = -1.f;
        int bi = INT_MAX;
        unsigned spins = 0;
        for (;;) {
            int ok = 1;
            best = -1.f;
            bi = INT_MAX;
            for (int q = t; q < nwg; q += 256) {
                const unsigned long long k =
                    __hip_atomic_load(&keys[par * MAX_PANEL_WG + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((unsigned)(k & 0xffffu) != (seq & 0xffffu)) {
                    ok = 0;
                } else {
                    const unsigned row = (unsigned)(k >> 16) & NO_ROW;
                    if (row != NO_ROW) argmax_combine(best, bi, __uint_as_float((unsigned)(k >> 33)), (int)row);
                }
            }
            if (__syncthreads_and(ok)) break;
            const int ab = __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
            if (__syncthreads_or(ab || ++spins > SPIN_LIMIT)) {
                if (t == 0) __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        wave_argmax(best, bi);
        if ((t & 63) == 0) { rv[t >> 6] = best; ri[t >> 6] = bi; }
        __syncthreads();
        if (t == 0) {
            for (int q = 1; q < 4; ++q) argmax_combine(rv[0], ri[0], rv[q], ri[q]);
            s_win = ri[0] == INT_MAX ? j : ri[0];   // every posted row is valid; keep j otherwise
        }
        __syncthreads();
        const int p = s_win;                 // panel-relative pivot row, >= j
        const int pw = p / PANEL_R, jw = j / PANEL_R;
        int late = 0;
        for (int c = t; c < nbc; c += 256) {
            float v;
            late |= !wait_tagged(&s_vals[((size_t)par * MAX_PANEL_WG + pw) * LU_NB + c], seq, v);
            urow[c] = v;
        }
        if (w == 0 && t == 0) ipiv[j0 + j] = j0 + p;
        if (p != j && w == pw)
            for (int c = t; c < nbc; c += 256) {
                float v;
                late |= !wait_tagged(&s_diag[par * LU_NB + c], seq, v);
                slab[c * PANEL_R + (p - rbase)] = v;
            }
        if (__syncthreads_or(late)) {
            if (t == 0) __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        if (p != j && w == jw)
            for (int c = t; c < nbc; c += 256) slab[c * PANEL_R + (j - rbase)] = urow[c];
        __syncthreads();
        const float piv = urow[j];
        const bool upd = piv != 0.f;
        if (!upd && w == 0 && t == 0 && *info == 0) *info = j0 + j + 1;
        const bool below = t < rows && rbase + t > j;   // rows that step j updates
        if (j + 1 == nbc) {
            if (upd && below) slab[j * PANEL_R + t] /= piv;
            break;
        }
        // one thread per row: scale column j, update column j + 1, and offer
        // the result as column j + 1's candidate
        float c1 = -1.f;
        if (below) {
            float a1 = slab[(j + 1) * PANEL_R + t];
            if (upd) {
                const float l = slab[j * PANEL_R + t] / piv;
                slab[j * PANEL_R + t] = l;
                a1 -= l * urow[j + 1];
                slab[(j + 1) * PANEL_R + t] = a1;
            }
            c1 = mag(a1);
        }
        const int lrow1 = block_candidate(c1, below ? rbase + t : INT_MAX, rv, ri, &s_lrow, rbase);
        // the rows posted for column j + 1 finish step j first
        const int drow1 = w == (j + 1) / PANEL_R ? j + 1 - rbase : -1;
        if (upd)
            for (int c = j + 2 + t; c < nbc; c += 256) {
                if (lrow1 >= 0) slab[c * PANEL_R + lrow1] -= slab[j * PANEL_R + lrow1] * urow[c];
                if (drow1 >= 0 && drow1 != lrow1) slab[c * PANEL_R + drow1] -= slab[j * PANEL_R + drow1] * urow[c];
            }
        __syncthreads();
        post(j + 1, lrow1);
        // the rest of the rank-1 update: two threads per row, the multiplier
        // in a register and the pivot row broadcast from LDS, columns interleaved
        if (upd && t < 2 * PANEL_R) {
            const int r = t < PANEL_R ? t : t - PANEL_R, h2 = t < PANEL_R ? 0 : 1;
            if (r < rows && rbase + r > j && r != lrow1 && r != drow1) {
                const float l = slab[j * PANEL_R + r];
#pragma unroll 4
                for (int c = j + 2 + h2; c < nbc; c += 2) slab[c * PANEL_R + r] -= l * urow[c];
            }
        }
        __syncthreads();
    }
    __syncthreads();

    for (int idx = t; idx < nbc * PANEL_R; idx += 256) {
        const int c = idx / PANEL_R, r = idx - c * PANEL_R;
        if (r < rows) P[(size_t)c * lda + rbase + r] = slab[idx];
    }
}

// row interchanges of the panel [j0, j0 + nbc) applied to the columns
// [cbeg, cend) (LAPACK slaswp), one thread per column
__global__ __launch_bounds__(256) void k_laswp(float* __restrict__ A, int lda, int cbeg, int cend, int j0, int nbc,
                                               const int* __restrict__ ipiv) {
    const int col = cbeg + blockIdx.x * 256 + threadIdx.x;
    if (col >= cend) return;
    float* cp = A + (size_t)col * lda;
    for (int k = 0; k < nbc; ++k) {
        const int p = ipiv[j0 + k];
        if (p != j0 + k) {
            const float a = cp[j0 + k];
            cp[j0 + k] = cp[p];
            cp[p] = a;
        }
    }
}

__device__ inline float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// A(j0:j0+nbc, cbeg:cend) <- L11^-1 A(j0:j0+nbc, cbeg:cend), L11 unit lower
// (strsm L,L,N,U).  Each wavefront solves TRSM_CPW columns together, two rows
// per lane, L11 in LDS: the columns' substitution chains are independent, so
// interleaving them inside the k loop hides the latency of each step
// (readlane of x_k after its last update) behind the other columns' work.
constexpr int TRSM_CPW = 8;   // columns per wavefront
__global__ __launch_bounds__(256) void k_trsm_llnu(float* __restrict__ A, int lda, int cbeg, int cend, int j0,
                                                   int nbc) {
    __shared__ float Ls[LU_NB * LU_NB];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int idx = t; idx < nbc * nbc; idx += 256) {
        const int k = idx / nbc, r = idx - k * nbc;
        Ls[k * LU_NB + r] = A[(size_t)(j0 + k) * lda + j0 + r];
    }
    __syncthreads();
    const int c0 = cbeg + (blockIdx.x * 4 + wv) * TRSM_CPW;
    float b0[TRSM_CPW], b1[TRSM_CPW];
#pragma unroll
    for (int q = 0; q < TRSM_CPW; ++q) {
        const float* cp = A + (size_t)(c0 + q) * lda + j0;
        const bool live = c0 + q < cend;
        b0[q] = live && lane < nbc ? cp[lane] : 0.f;
        b1[q] = live && lane + 64 < nbc ? cp[lane + 64] : 0.f;
    }
    for (int k = 0; k < nbc; ++k) {
        // only rows r > k (and r < nbc) take -L(r, k) x_k
        const bool u0 = lane > k && lane < nbc, u1 = lane + 64 > k && lane + 64 < nbc;
        const float l0 = Ls[k * LU_NB + lane], l1 = Ls[k * LU_NB + lane + 64];
        if (k < 64) {
#pragma unroll
            for (int q = 0; q < TRSM_CPW; ++q) {
                const float xk = readlane_f(b0[q], k);
                b0[q] = u0 ? b0[q] - l0 * xk : b0[q];
                b1[q] = u1 ? b1[q] - l1 * xk : b1[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < TRSM_CPW; ++q) {
                const float xk = readlane_f(b1[q], k - 64);
                b1[q] = u1 ? b1[q] - l1 * xk : b1[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < TRSM_CPW; ++q) {
        if (c0 + q >= cend) break;
        float* cp = A + (size_t)(c0 + q) * lda + j0;
        if (lane < nbc) cp[lane] = b0[q];
        if (lane + 64 < nbc) cp[lane + 64] = b1[q];
    }
}

// FP64 row-major -> FP32 column-major (the factor's working copy)
__global__ __launch_bounds__(256) void k_to_cm32(const double* __restrict__ A64, int n, float* __restrict__ A32,
                                                 int lda) {
    __shared__ float tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    for (int k = ty; k < 64; k += 4) {
        const int i = i0 + k, j = j0 + tx;
        tile[k][tx] = (i < n && j < n) ? (float)A64[(size_t)i * n + j] : 0.f;
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const int j = j0 + k, i = i0 + tx;
        if (i < n && j < n) A32[(size_t)j * lda + i] = tile[tx][k];
    }
}

// ------------------------------------------------- triangular solves ----
// y = (float) r[perm]  (P applied to the FP64 residual)
__global__ void k_gather_perm(const double* __restrict__ r, const int* __restrict__ perm, float* __restrict__ y,
                              int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)r[perm[i]];
}

// Inverse of every 128 x 128 diagonal block of the factor, computed in
// place in LDS column by column (LAPACK strti2): unit lower L -> L^-1 and
// upper U -> U^-1, blocks past n padded with the identity.  One workgroup per
// block; the triangular solves apply these with a matrix-vector product
// instead of a 128-step substitution.
template <bool LOWER>
__global__ __launch_bounds__(128) void k_tri_inv(const float* __restrict__ A, int lda, int n,
                                                 float* __restrict__ Tinv) {
    __shared__ float X[LU_NB * LU_NB];   // column k at X[k * LU_NB]
    const int i = threadIdx.x, b0 = blockIdx.x * LU_NB, nb = min(LU_NB, n - b0);
    for (int k = 0; k < LU_NB; ++k) {
        float v = (k == i) ? 1.f : 0.f;
        if (k < nb && i < nb) {
            const bool keep = LOWER ? (i > k) : (i <= k);
            v = keep ? A[(size_t)(b0 + k) * lda + b0 + i] : ((k == i) ? 1.f : 0.f);
        }
        X[k * LU_NB + i] = v;
    }
    __syncthreads();
    if (LOWER) {
        for (int j = LU_NB - 2; j >= 0; --j) {
            float s = 0.f;
            if (i > j) {
                s = X[j * LU_NB + i];
                for (int k = j + 1; k < i; ++k) s += X[k * LU_NB + i] * X[j * LU_NB + k];
            }
            __syncthreads();
            if (i > j) X[j * LU_NB + i] = -s;
            __syncthreads();
        }
    } else {
        for (int j = 0; j < LU_NB; ++j) {
            const float d = 1.f / X[j * LU_NB + j];
            float s = 0.f;
            if (i < j)
                for (int k = i; k < j; ++k) s += X[k * LU_NB + i] * X[j * LU_NB + k];
            __syncthreads();
            if (i < j) X[j * LU_NB + i] = -d * s;
            if (i == j) X[j * LU_NB + j] = d;
            __syncthreads();
        }
    }
    float* out = Tinv + (size_t)blockIdx.x * LU_NB * LU_NB;
    for (int k = 0; k < LU_NB; ++k) out[k * LU_NB + i] = X[k * LU_NB + i];
}

// Blocked triangular solve in one launch (y <- L^-1 y or U^-1 y in place).
// Each workgroup takes a ticket; ticket t owns block row t (forward) or
// nblk - 1 - t (backward), so a workgroup only waits on blocks owned by
// workgroups that started before it.  It accumulates T(b, c) x_c over the
// blocks c it depends on (the block of T is loaded before the wait), applies
// the inverse of its diagonal block and publishes x_b.  Publication is one
// 64-bit word per element, the value in the low half and the call's epoch in
// the high half: a consumer thread polls the element it needs, so the value
// arrives with its own flag in a single round trip (no separate flag, fence
// or reload).  Every wait is bounded (abort flag, the grid still drains).
template <bool LOWER>
__global__ __launch_bounds__(256) void k_trsv(const float* __restrict__ A, int lda, const float* __restrict__ Tinv,
                                              float* __restrict__ y, int n, int nblk, unsigned* ticket,
                                              unsigned base, unsigned long long* yx, unsigned epoch,
                                              unsigned* abort_flag) {
    __shared__ float xs[LU_NB];
    __shared__ float part[2][LU_NB];
    __shared__ int s_b;
    const int t = threadIdx.x, r = t & (LU_NB - 1), half = t >> 7;
    if (t == 0) s_b = (int)(atomicAdd(ticket, 1u) - base);
    __syncthreads();
    const int b = LOWER ? s_b : nblk - 1 - s_b;
    const int b0 = b * LU_NB, nb = min(LU_NB, n - b0);
    float tinv[LU_NB / 2];
    const float* Tb = Tinv + (size_t)b * LU_NB * LU_NB;
#pragma unroll
    for (int q = 0; q < LU_NB / 2; ++q) tinv[q] = Tb[(2 * q + half) * LU_NB + r];
    float acc = 0.f;
    const int ndep = LOWER ? b : nblk - 1 - b;
    for (int d = 0; d < ndep; ++d) {
        const int c = LOWER ? d : nblk - 1 - d;
        const int c0 = c * LU_NB, nc = min(LU_NB, n - c0);
        float a[LU_NB / 2];
#pragma unroll
        for (int q = 0; q < LU_NB / 2; ++q) {
            const int k = 2 * q + half;
            a[q] = (k < nc && r < nb) ? A[(size_t)(c0 + k) * lda + b0 + r] : 0.f;
        }
        if (t < LU_NB) {
            float v = 0.f;
            if (t < nc) {
                unsigned spins = 0;
                unsigned long long w = 0ull;
                while (true) {
                    w = __hip_atomic_load(&yx[c0 + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((unsigned)(w >> 32) == epoch) break;
                    if ((spins & 31u) == 0u &&
                        __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                        break;
                    if (++spins > SPIN_LIMIT) {
                        __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                v = __uint_as_float((unsigned)(w & 0xFFFFFFFFull));
            }
            xs[t] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LU_NB / 2; ++q) acc += a[q] * xs[2 * q + half];
        __syncthreads();
    }
    part[half][r] = acc;
    __syncthreads();
    if (t < LU_NB) xs[t] = t < nb ? y[b0 + t] - (part[0][t] + part[1][t]) : 0.f;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < LU_NB / 2; ++q) s += tinv[q] * xs[2 * q + half];
    part[half][r] = s;
    __syncthreads();
    if (t < nb) {
        const float xv = part[0][t] + part[1][t];
        y[b0 + t] = xv;
        __hip_atomic_store(&yx[b0 + t], ((unsigned long long)epoch << 32) | (unsigned long long)__float_as_uint(xv),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// x += (double) y
__global__ void k_add_f32(double* __restrict__ x, const float* __restrict__ y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)y[i];
}

// ------------------------------------------------------ FP64 residual ----
constexpr int RES_ROWS = 4;   // matrix rows per wavefront (x is read once per 4 rows)
__global__ __launch_bounds__(256) void k_resid64(const double* __restrict__ A, int n, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ r,
                                                 double* __restrict__ rowabs) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + wv) * RES_ROWS;
    if (row0 >= n) return;
    const int nrow = min(RES_ROWS, n - row0);
    const double* Ar[RES_ROWS];
#pragma unroll
    for (int q = 0; q < RES_ROWS; ++q) Ar[q] = A + (size_t)(row0 + min(q, nrow - 1)) * n;
    double acc[RES_ROWS], aab[RES_ROWS];
#pragma unroll
    for (int q = 0; q < RES_ROWS; ++q) { acc[q] = 0.0; aab[q] = 0.0; }
    int j = lane;
    for (; j + 192 < n; j += 256) {
        double xv[4], av[RES_ROWS][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = x[j + 64 * u];
#pragma unroll
        for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) av[q][u] = __builtin_nontemporal_load(Ar[q] + j + 64 * u);
#pragma unroll
        for (int q = 0; q < RES_ROWS; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[q] += av[q][u] * xv[u];
                aab[q] += fabs(av[q][u]);
            }
    }
    for (; j < n; j += 64) {
        const double xj = x[j];
#pragma unroll
        for (int q = 0; q < RES_ROWS; ++q) {
            const double a = Ar[q][j];
            acc[q] += a * xj;
            aab[q] += fabs(a);
        }
    }
#pragma unroll
    for (int q = 0; q < RES_ROWS; ++q) {
        for (int off = 32; off > 0; off >>= 1) {
            acc[q] += __shfl_xor(acc[q], off);
            aab[q] += __shfl_xor(aab[q], off);
        }
    }
    if (lane < nrow) {
        double a = 0.0, s = 0.0;
#pragma unroll
        for (int q = 0; q < RES_ROWS; ++q) if (q == lane) { a = acc[q]; s = aab[q]; }
        r[row0 + lane] = b[row0 + lane] - a;
        if (rowabs) rowabs[row0 + lane] = s;
    }
}

// out = {max|r|, max|x|, max|b|, max rowabs}: one workgroup
__global__ __launch_bounds__(1024) void k_norms(const double* __restrict__ r, const double* __restrict__ x,
                                                const double* __restrict__ b, const double* __restrict__ rowabs,
                                                int n, double* __restrict__ out) {
    __shared__ double red[4][16];
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += 1024) {
        m[0] = fmax(m[0], fabs(r[i]));
        m[1] = fmax(m[1], fabs(x[i]));
        m[2] = fmax(m[2], fabs(b[i]));
        if (rowabs) m[3] = fmax(m[3], rowabs[i]);
    }
    for (int q = 0; q < 4; ++q)
        for (int off = 32; off > 0; off >>= 1) m[q] = fmax(m[q], __shfl_xor(m[q], off));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) for (int q = 0; q < 4; ++q) red[q][wv] = m[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int k = 0; k < 16; ++k) v = fmax(v, red[threadIdx.x][k]);
        out[threadIdx.x] = v;
    }
}

// ---------------------------------------------------------- VLM kernels --
// Ring geometry (panel.cpp:46-72, 103-120): corners copied per ring, centre,
// unit normal -sum dl_i x dl_i+1, collocation point k centre + (1 - k) leading-
// edge midpoint, and rhs_i = -Vinf . n_i (solver.cpp:256-268).
__global__ void k_ring_geom(const double* __restrict__ nodes, const int* __restrict__ rn, int nr, double kfac,
                            double vx, double vy, double vz, double* __restrict__ corner,
                            double* __restrict__ colloc, double* __restrict__ normal, double* __restrict__ rhs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    double p[4][3];
    for (int a = 0; a < 4; ++a) {
        const int id = rn[4 * k + a];
        for (int d = 0; d < 3; ++d) {
            p[a][d] = nodes[3 * id + d];
            corner[12 * k + 3 * a + d] = p[a][d];
        }
    }
    double c[3], nrm[3] = {0.0, 0.0, 0.0}, dl[4][3];
    for (int d = 0; d < 3; ++d) c[d] = (((p[0][d] + p[1][d]) + p[2][d]) + p[3][d]) / 4.0;
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) dl[a][d] = p[(a + 1) & 3][d] - p[a][d];
    for (int a = 0; a < 4; ++a) {
        const double* u = dl[a];
        const double* v = dl[(a + 1) & 3];
        nrm[0] -= u[1] * v[2] - u[2] * v[1];
        nrm[1] -= u[2] * v[0] - u[0] * v[2];
        nrm[2] -= u[0] * v[1] - u[1] * v[0];
    }
    const double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int d = 0; d < 3; ++d) {
        nrm[d] /= len;
        normal[3 * k + d] = nrm[d];
        colloc[3 * k + d] = kfac * c[d] + (1.0 - kfac) * (0.5 * (p[0][d] + p[1][d]));
    }
    rhs[k] = -(vx * nrm[0] + vy * nrm[1] + vz * nrm[2]);
}

// wake ring of every station (surface.cpp:35-58): trailing-edge nodes 3 and
// 2 of the station's last ring, then the same points moved wake_length along
// the unit freestream
__global__ void k_wake(const double* __restrict__ corner, const int* __restrict__ soff, int ns, double L,
                       double wx, double wy, double wz, double* __restrict__ wcorner) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int k = soff[s + 1] - 1;
    const double* p = corner + 12 * k;
    const double w[3] = {wx, wy, wz};
    for (int d = 0; d < 3; ++d) {
        const double p1 = p[9 + d], p2 = p[6 + d];
        wcorner[12 * s + 0 + d] = p1;
        wcorner[12 * s + 3 + d] = p2;
        wcorner[12 * s + 6 + d] = p2 + L * w[d];
        wcorner[12 * s + 9 + d] = p1 + L * w[d];
    }
}

// velocity induced at P by a unit-strength ring with corners q[0..3]
// (4 pi times vortexRing::influence, panel.cpp:122-128).  Per segment a->b
// with r1 = P - a, r2 = P - b the reference's h and direction
// (potential.cpp:36-58) give v = (dl.r1/|r1| - dl.r2/|r2|) (r1 x r2) / |r1 x r2|^2
// since dl x r1 = r1 x r2 and |dl x r1| = |r1 x r2|; the Lamb-Oseen factor
// 1 - exp(-(h/rc)^2) multiplies it when rc > 0.
__device__ inline void ring_velocity(const double P[3], const double q[12], double rc, double v[3]) {
    double r[4][3], ir[4];
    for (int a = 0; a < 4; ++a) {
        for (int d = 0; d < 3; ++d) r[a][d] = P[d] - q[3 * a + d];
        ir[a] = rsqrt(r[a][0] * r[a][0] + r[a][1] * r[a][1] + r[a][2] * r[a][2]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int b = (a + 1) & 3;
        const double* r1 = r[a];
        const double* r2 = r[b];
        const double c0 = r1[1] * r2[2] - r1[2] * r2[1];
        const double c1 = r1[2] * r2[0] - r1[0] * r2[2];
        const double c2 = r1[0] * r2[1] - r1[1] * r2[0];
        const double cc = c0 * c0 + c1 * c1 + c2 * c2;
        const double d0 = q[3 * b] - q[3 * a], d1 = q[3 * b + 1] - q[3 * a + 1], d2 = q[3 * b + 2] - q[3 * a + 2];
        const double tt = (d0 * r1[0] + d1 * r1[1] + d2 * r1[2]) * ir[a] - (d0 * r2[0] + d1 * r2[1] + d2 * r2[2]) * ir[b];
        double s = tt / cc;
        if (rc > 0.0) {
            const double h = sqrt(cc) / tt;
            s *= 1.0 - exp(-(h / rc) * (h / rc));
        }
        v[0] += s * c0;
        v[1] += s * c1;
        v[2] += s * c2;
    }
}

// AIC ring part: A(i, j) = v_j(P_i) . n_i.  Block = 64 columns x 64 rows;
// a wavefront holds 64 consecutive columns of one row, so its stores are
// contiguous in the row-major matrix and its row data is an LDS broadcast.
constexpr int AIC_ROWS_PER_WAVE = 16;
__global__ __launch_bounds__(256) void k_aic(const double* __restrict__ colloc, const double* __restrict__ normal,
                                             const double* __restrict__ corner, int n, double rc,
                                             double* __restrict__ A) {
    __shared__ double sP[64][6];
    const int t = threadIdx.x;
    const int rbase = blockIdx.y * 64;
    if (t < 64) {
        const int i = rbase + t;
        for (int d = 0; d < 3; ++d) {
            sP[t][d] = i < n ? colloc[3 * i + d] : 0.0;
            sP[t][3 + d] = i < n ? normal[3 * i + d] : 0.0;
        }
    }
    __syncthreads();
    const int j = blockIdx.x * 64 + (t & 63);
    if (j >= n) return;
    double q[12];
    for (int a = 0; a < 12; ++a) q[a] = corner[12 * j + a];
    const int r0 = (t >> 6) * AIC_ROWS_PER_WAVE;
    for (int rr = 0; rr < AIC_ROWS_PER_WAVE; ++rr) {
        const int il = r0 + rr, i = rbase + il;
        if (i >= n) break;
        const double P[3] = {sP[il][0], sP[il][1], sP[il][2]};
        double v[3] = {0.0, 0.0, 0.0};
        ring_velocity(P, q, rc, v);
        A[(size_t)i * n + j] = (v[0] * sP[il][3] + v[1] * sP[il][4] + v[2] * sP[il][5]) / FOUR_PI;
    }
}

// wake rings added to their trailing-edge ring's column (solver.cpp:221-236)
__global__ __launch_bounds__(256) void k_aic_wake(const double* __restrict__ colloc, const double* __restrict__ normal,
                                                  const double* __restrict__ wcorner, const int* __restrict__ soff,
                                                  int ns, int n, double rc, double* __restrict__ A) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (long long)n * ns) return;
    const int s = (int)(idx % ns), i = (int)(idx / ns);
    double q[12];
    for (int a = 0; a < 12; ++a) q[a] = wcorner[12 * s + a];
    const double P[3] = {colloc[3 * i], colloc[3 * i + 1], colloc[3 * i + 2]};
    double v[3] = {0.0, 0.0, 0.0};
    ring_velocity(P, q, rc, v);
    const int te = soff[s + 1] - 1;
    A[(size_t)i * n + te] += (v[0] * normal[3 * i] + v[1] * normal[3 * i + 1] + v[2] * normal[3 * i + 2]) / FOUR_PI;
}

// Kutta-Joukowski per ring, leading to trailing edge of each station
// (wingStation::computeForces, surface.cpp:164-201, with the local stream
// equal to the freestream of the linear solver)
__global__ void k_forces(const double* __restrict__ corner, const double* __restrict__ gamma,
                         const int* __restrict__ soff, int ns, double rho, double vx, double vy, double vz,
                         double* __restrict__ rforce, double* __restrict__ sforce) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    double prev = 0.0, F[3] = {0.0, 0.0, 0.0};
    for (int k = soff[s]; k < soff[s + 1]; ++k) {
        const double* p = corner + 12 * k;
        const double dl0 = p[3] - p[0], dl1 = p[4] - p[1], dl2 = p[5] - p[2];
        const double g = rho * (gamma[k] - prev);
        const double f0 = (vy * dl2 - vz * dl1) * g;
        const double f1 = (vz * dl0 - vx * dl2) * g;
        const double f2 = (vx * dl1 - vy * dl0) * g;
        rforce[3 * k] = f0;
        rforce[3 * k + 1] = f1;
        rforce[3 * k + 2] = f2;
        F[0] += f0;
        F[1] += f1;
        F[2] += f2;
        prev = gamma[k];
    }
    for (int d = 0; d < 3; ++d) sforce[3 * s + d] = F[d];
}

// nonlinear::steady::buildRHS (solver.cpp:303-316): rhs_i = -V(alpha_s).n_i
// with the local angle of the ring's station
__global__ void k_rhs_stations(const double* __restrict__ normal, const int* __restrict__ rst, int nr, double vinf,
                               const double* __restrict__ alpha_deg, double* __restrict__ rhs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    const double a = alpha_deg[rst[k]] * M_PI / 180.0;
    rhs[k] = -(vinf * cos(a) * normal[3 * k] + vinf * sin(a) * normal[3 * k + 2]);
}

// Kutta-Joukowski with each station's local stream (wingStation::computeForces
// with inertial_stream() = V(alpha_s))
__global__ void k_forces_stations(const double* __restrict__ corner, const double* __restrict__ gamma,
                                  const int* __restrict__ soff, int ns, double rho, double vinf,
                                  const double* __restrict__ alpha_deg, double* __restrict__ sforce) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const double a = alpha_deg[s] * M_PI / 180.0;
    const double vx = vinf * cos(a), vz = vinf * sin(a);
    double prev = 0.0, F[3] = {0.0, 0.0, 0.0};
    for (int k = soff[s]; k < soff[s + 1]; ++k) {
        const double* p = corner + 12 * k;
        const double dl0 = p[3] - p[0], dl1 = p[4] - p[1], dl2 = p[5] - p[2];
        const double g = rho * (gamma[k] - prev);
        F[0] += (0.0 * dl2 - vz * dl1) * g;
        F[1] += (vz * dl0 - vx * dl2) * g;
        F[2] += (vx * dl1 - 0.0 * dl0) * g;
        prev = gamma[k];
    }
    for (int d = 0; d < 3; ++d) sforce[3 * s + d] = F[d];
}

// ------------------------------------------------- trailing update (MFMA) --
// C -= A B in FP32 on the matrix cores, all three column-major with one
// leading dimension ld (the LU's A22 -= A21 U12: A is m x k, B is k x n).
// One 256-thread workgroup computes a 128 x 128 tile of C: each wavefront a
// 64 x 64 block as 2 x 2 v_mfma_f32_32x32x2f32 accumulators (16 floats per
// lane each).  K advances in chunks of 16 staged through two LDS buffers: the
// A chunk k-major as in memory (rows contiguous), the B chunk transposed to
// k-major rows padded by 4 words so its column-wise stores spread over the
// banks; the next chunk's global loads are issued before the current chunk's
// MFMAs.  Operand layout of the 32x32x2 instruction: lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32] of the 32 x 2 and 2 x 32 slices;
// accumulator element v of lane l is C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
// Blocks map to tiles XCD-aware (consecutive tiles of one XCD share A rows).
constexpr int GM_TM = 128, GM_TN = 128, GM_TK = 16;
constexpr int GM_LDB = GM_TN + 4;
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(256) k_gemm_sub_f32(const float* __restrict__ A, const float* __restrict__ B,
                                                       float* __restrict__ C, int ld, int m, int n, int k) {
    __shared__ float As[2][GM_TK][GM_TM];
    __shared__ float Bs[2][GM_TK][GM_LDB];
    const int tiles_m = (m + GM_TM - 1) / GM_TM, tiles_n = (n + GM_TN - 1) / GM_TN;
    const int ntiles = tiles_m * tiles_n;
    // XCD-aware: block b runs on XCD b % 8; XCD x takes tiles [x * per, (x + 1) * per)
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int tile = (b % 8) * per + b / 8;
    if (tile >= ntiles) return;
    const int tm = tile % tiles_m, tn = tile / tiles_m;
    const int row0 = tm * GM_TM, col0 = tn * GM_TN;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    const bool full = row0 + GM_TM <= m && col0 + GM_TN <= n;
    // global -> register staging: A chunk (128 rows x 16 k): column kk = t / 16,
    // rows 8 (t % 16) .. +8; B chunk (16 k x 128 columns): column t / 2, k
    // 8 (t % 2) .. +8
    const int a_k = t >> 4, a_r = (t & 15) * 8;
    const int b_c = t >> 1, b_k = (t & 1) * 8;
    float ra[8], rb[8];
    auto load = [&](int kb) {
        const int ka = kb + a_k, kbb = kb + b_k;
        if (full && kb + GM_TK <= k) {
            const float4* pa = reinterpret_cast<const float4*>(A + (size_t)ka * ld + row0 + a_r);
            const float4 a0 = pa[0], a1 = pa[1];
            ra[0] = a0.x; ra[1] = a0.y; ra[2] = a0.z; ra[3] = a0.w; ra[4] = a1.x; ra[5] = a1.y; ra[6] = a1.z; ra[7] = a1.w;
            const float4* pb = reinterpret_cast<const float4*>(B + (size_t)(col0 + b_c) * ld + kbb);
            const float4 b0 = pb[0], b1 = pb[1];
            rb[0] = b0.x; rb[1] = b0.y; rb[2] = b0.z; rb[3] = b0.w; rb[4] = b1.x; rb[5] = b1.y; rb[6] = b1.z; rb[7] = b1.w;
        } else {
            for (int i = 0; i < 8; ++i) {
                const int r = row0 + a_r + i;
                ra[i] = (r < m && ka < k) ? A[(size_t)ka * ld + r] : 0.f;
                const int c = col0 + b_c, kk = kbb + i;
                rb[i] = (c < n && kk < k) ? B[(size_t)c * ld + kk] : 0.f;
            }
        }
    };
    auto store = [&](int buf) {
        float4* da = reinterpret_cast<float4*>(&As[buf][a_k][a_r]);
        da[0] = make_float4(ra[0], ra[1], ra[2], ra[3]);
        da[1] = make_float4(ra[4], ra[5], ra[6], ra[7]);
        for (int i = 0; i < 8; ++i) Bs[buf][b_k + i][b_c] = rb[i];
    };
    f32x16 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    const int nk = (k + GM_TK - 1) / GM_TK;
    load(0);
    store(0);
    __syncthreads();
    const int lr = lane & 31, lk = lane >> 5;
    for (int c = 0; c < nk; ++c) {
        const int buf = c & 1;
        if (c + 1 < nk) load((c + 1) * GM_TK);
#pragma unroll
        for (int kk = 0; kk < GM_TK; kk += 2) {
            const float a0 = As[buf][kk + lk][wm + lr], a1 = As[buf][kk + lk][wm + 32 + lr];
            const float b0 = Bs[buf][kk + lk][wn + lr], b1 = Bs[buf][kk + lk][wn + 32 + lr];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (c + 1 < nk) {
            store(buf ^ 1);   // the other buffer: its last readers finished before the previous barrier
        }
        __syncthreads();
    }
    // C -= acc
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            const int col = col0 + wn + j * 32 + lr;
            for (int g = 0; g < 4; ++g) {
                const int row = row0 + wm + i * 32 + 8 * g + 4 * lk;   // rows row .. row + 3
                if (full) {
                    float4* pc = reinterpret_cast<float4*>(C + (size_t)col * ld + row);
                    float4 cv = *pc;
                    cv.x -= acc[i][j][4 * g + 0];
                    cv.y -= acc[i][j][4 * g + 1];
                    cv.z -= acc[i][j][4 * g + 2];
                    cv.w -= acc[i][j][4 * g + 3];
                    *pc = cv;
                } else if (col < n) {
                    for (int v = 0; v < 4; ++v)
                        if (row + v < m) C[(size_t)col * ld + row + v] -= acc[i][j][4 * g + v];
                }
            }
        }
}

hipError_t launch_gemm_sub_f32(hipStream_t s, const float* A, const float* B, float* C, int ld, int m, int n, int k) {
    if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
    const int tiles = ((m + GM_TM - 1) / GM_TM) * ((n + GM_TN - 1) / GM_TN);
    const int nb = ((tiles + 7) / 8) * 8;
    hipLaunchKernelGGL(k_gemm_sub_f32, dim3(nb), dim3(256), 0, s, A, B, C, ld, m, n, k);
    return hipGetLastError();
}

__global__ void k_panel_reset(unsigned long long* keys, unsigned long long* s_vals, unsigned long long* s_diag,
                              unsigned* abort_flag, unsigned* trsv_abort, int* info) {
    const int stride = gridDim.x * blockDim.x, i0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int q = i0; q < 2 * MAX_PANEL_WG; q += stride) keys[q] = 0ull;
    // the row words carry the column sequence j0 + j + 1, which repeats from one
    // factorisation to the next: clear them with the keys
    for (int q = i0; q < 2 * MAX_PANEL_WG * LU_NB; q += stride) s_vals[q] = 0ull;
    for (int q = i0; q < 2 * LU_NB; q += stride) s_diag[q] = 0ull;
    if (i0 == 0) {
        *abort_flag = 0u;
        *trsv_abort = 0u;   // a new factorisation also re-arms the triangular solves
        *info = 0;
    }
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace

// ======================================================= dense solver ====
struct afx_dense {
    int n = 0, lda = 0;
    hipStream_t s = nullptr;
    // look-ahead of the factorisation: the update of the columns past the
    // next panel runs on s2 while s factors that panel
    hipStream_t s2 = nullptr;
    hipEvent_t ev_panel = nullptr, ev_far = nullptr;
    double* A64 = nullptr;     // row-major n x n
    float* A32 = nullptr;      // column-major lda x n, LU in place
    int* ipiv = nullptr;
    int* perm = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr, *rowabs = nullptr, *norms = nullptr;
    float* y = nullptr;
    unsigned long long* keys = nullptr;
    unsigned long long* s_vals = nullptr;   // candidate rows, value | column sequence
    unsigned long long* s_diag = nullptr;
    unsigned* pabort = nullptr;
    int* info = nullptr;
    float *Linv = nullptr, *Uinv = nullptr;   // inverses of the 128 x 128 diagonal blocks
    unsigned *tk = nullptr, *abort = nullptr;   // triangular-solve tickets and abort flag
    unsigned long long* yx = nullptr;            // published solution elements: value | epoch << 32
    unsigned tk_base = 0, epoch = 0;
    int nblk = 0;
    double anorm = -1.0;       // |A|inf of the current FP64 matrix (-1: not yet known)
    bool factored = false;
    ~afx_dense() {
        for (void* p : {(void*)A64, (void*)A32, (void*)ipiv, (void*)perm, (void*)b, (void*)x, (void*)r,
                        (void*)rowabs, (void*)norms, (void*)y, (void*)keys, (void*)s_vals,
                        (void*)s_diag, (void*)pabort, (void*)info, (void*)Linv, (void*)Uinv, (void*)tk,
                        (void*)yx, (void*)abort})
            if (p) (void)hipFree(p);
        if (ev_panel) (void)hipEventDestroy(ev_panel);
        if (ev_far) (void)hipEventDestroy(ev_far);
        if (s) (void)hipStreamDestroy(s);
        if (s2) (void)hipStreamDestroy(s2);
    }
};

namespace {

int dense_init(afx_dense* h, int n) {
    if (n <= 0) return api_fail(AFX_EARG, "matrix order must be positive");
    if ((n + PANEL_R - 1) / PANEL_R > MAX_PANEL_WG)
        return api_fail(AFX_EUNSUP, "matrix order " + std::to_string(n) + " exceeds the panel kernel's " +
                                        std::to_string(MAX_PANEL_WG * PANEL_R));
    int dev = 0, coop = 0;
    VCHK(hipGetDevice(&dev));
    VCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) return api_fail(AFX_EUNSUP, "device does not support cooperative launches");
    h->n = n;
    h->lda = ((n + 31) / 32) * 32 + 32;   // padded so that columns do not alias HBM channels
    VCHK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    VCHK(hipStreamCreateWithFlags(&h->s2, hipStreamNonBlocking));
    VCHK(hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming));
    VCHK(hipEventCreateWithFlags(&h->ev_far, hipEventDisableTiming));
    const size_t nn = (size_t)n;
    VCHK(dalloc(&h->A64, nn * nn));
    VCHK(dalloc(&h->A32, (size_t)h->lda * nn));
    VCHK(dalloc(&h->ipiv, nn));
    VCHK(dalloc(&h->perm, nn));
    VCHK(dalloc(&h->b, nn));
    VCHK(dalloc(&h->x, nn));
    VCHK(dalloc(&h->r, nn));
    VCHK(dalloc(&h->rowabs, nn));
    VCHK(dalloc(&h->norms, 4));
    VCHK(dalloc(&h->y, nn));
    VCHK(dalloc(&h->keys, 2 * MAX_PANEL_WG));
    VCHK(dalloc(&h->s_vals, (size_t)2 * MAX_PANEL_WG * LU_NB));
    VCHK(dalloc(&h->s_diag, 2 * LU_NB));
    VCHK(dalloc(&h->pabort, 1));
    VCHK(dalloc(&h->info, 1));
    h->nblk = (n + LU_NB - 1) / LU_NB;
    VCHK(dalloc(&h->Linv, (size_t)h->nblk * LU_NB * LU_NB));
    VCHK(dalloc(&h->Uinv, (size_t)h->nblk * LU_NB * LU_NB));
    VCHK(dalloc(&h->tk, 1));
    VCHK(dalloc(&h->yx, h->n));
    VCHK(dalloc(&h->abort, 1));
    VCHK(hipMemsetAsync(h->tk, 0, sizeof(unsigned), h->s));
    VCHK(hipMemsetAsync(h->yx, 0, (size_t)h->n * sizeof(unsigned long long), h->s));
    VCHK(hipMemsetAsync(h->abort, 0, sizeof(unsigned), h->s));
    VCHK(hipStreamSynchronize(h->s));
    return AFX_OK;
}

int dense_factor(afx_dense* h, int32_t* info_out) {
    const int n = h->n, lda = h->lda;
    hipStream_t s = h->s;
    h->factored = false;
    dim3 tg((n + 63) / 64, (n + 63) / 64);
    k_to_cm32<<<tg, 256, 0, s>>>(h->A64, n, h->A32, lda);
    VCHK(hipGetLastError());
    k_panel_reset<<<64, 256, 0, s>>>(h->keys, h->s_vals, h->s_diag, h->pabort, h->abort, h->info);
    VCHK(hipGetLastError());
    // Right-looking blocked LU with a look-ahead of one panel.  After panel
    // k (columns [j0, c1)) is factored on s, s2 swaps the rows of the columns
    // left of it and past the next panel ([0, j0) and [c2, n)) and applies
    // step k's update to [c2, n), while s applies it to the next panel's
    // columns [c1, c2) alone and factors that panel.  s waits for the far
    // update of step k - 1 before its own update of step k, since that one
    // brought [c1, c2) up to step k - 1; the column sets the two streams touch
    // at the same time are disjoint.
    float* A = h->A32;
    for (int j0 = 0; j0 < n; j0 += LU_NB) {
        const int nbc = std::min(LU_NB, n - j0);
        const int m = n - j0;
        const int nwg = (m + PANEL_R - 1) / PANEL_R;
        int ldai = lda, mi = m, nbci = nbc, j0i = j0;
        void* args[] = {&A, &ldai, &mi, &nbci, &j0i, &h->ipiv, &h->keys, &h->s_vals, &h->s_diag, &h->pabort,
                        &h->info};
        VCHK(hipLaunchCooperativeKernel((const void*)k_panel_lu, dim3(nwg), dim3(256), args, 0, s));
        VCHK(hipEventRecord(h->ev_panel, s));
        const int c1 = j0 + nbc, c2 = std::min(n, c1 + LU_NB), rest = n - c1;
        const float* A21 = A + (size_t)j0 * lda + c1;
        auto update = [&](hipStream_t st, int cb, int ce) -> int {
            k_laswp<<<(ce - cb + 255) / 256, 256, 0, st>>>(A, lda, cb, ce, j0, nbc, h->ipiv);
            k_trsm_llnu<<<(ce - cb + 4 * TRSM_CPW - 1) / (4 * TRSM_CPW), 256, 0, st>>>(A, lda, cb, ce, j0, nbc);
            VCHK(hipGetLastError());
            VCHK(launch_gemm_sub_f32(st, A21, A + (size_t)cb * lda + j0, A + (size_t)cb * lda + c1, lda, rest, ce - cb,
                                     nbc));
            return AFX_OK;
        };
        VCHK(hipStreamWaitEvent(h->s2, h->ev_panel, 0));
        if (j0 > 0) {
            k_laswp<<<(j0 + 255) / 256, 256, 0, h->s2>>>(A, lda, 0, j0, j0, nbc, h->ipiv);
            VCHK(hipGetLastError());
        }
        if (c2 < n) {
            const int rc = update(h->s2, c2, n);
            if (rc != AFX_OK) return rc;
        }
        if (c1 < n) {
            if (j0 > 0) VCHK(hipStreamWaitEvent(s, h->ev_far, 0));   // step k - 1 on [c1, c2)
            const int rc = update(s, c1, c2);
            if (rc != AFX_OK) return rc;
        }
        VCHK(hipEventRecord(h->ev_far, h->s2));
    }
    VCHK(hipStreamWaitEvent(s, h->ev_far, 0));
    unsigned aborted = 0;
    int info = 0;
    VCHK(hipMemcpyAsync(&aborted, h->pabort, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    VCHK(hipMemcpyAsync(&info, h->info, sizeof(int), hipMemcpyDeviceToHost, s));
    std::vector<int> ipiv(n);
    VCHK(hipMemcpyAsync(ipiv.data(), h->ipiv, n * sizeof(int), hipMemcpyDeviceToHost, s));
    VCHK(hipStreamSynchronize(s));
    if (aborted) return api_fail(AFX_EHIP, "LU panel: pivot exchange timed out (workgroups not co-resident)");
    // P b = b[perm]: the interchanges applied in order to the identity
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        const int p = ipiv[k];
        if (p < k || p >= n) return api_fail(AFX_EHIP, "LU panel produced an invalid pivot");
        std::swap(perm[k], perm[p]);
    }
    VCHK(hipMemcpyAsync(h->perm, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    if (!info) {
        k_tri_inv<true><<<h->nblk, LU_NB, 0, s>>>(h->A32, lda, n, h->Linv);
        k_tri_inv<false><<<h->nblk, LU_NB, 0, s>>>(h->A32, lda, n, h->Uinv);
        VCHK(hipGetLastError());
    }
    VCHK(hipStreamSynchronize(s));
    if (info_out) *info_out = info;
    h->factored = info == 0;
    if (info) return api_fail(AFX_EARG, "matrix is singular in FP32: zero pivot at column " + std::to_string(info - 1));
    return AFX_OK;
}

// y <- U^-1 L^-1 y (y already permuted): one launch per triangle
int dense_lu_solve(afx_dense* h) {
    hipStream_t s = h->s;
    const int nb = h->nblk;
    k_trsv<true><<<nb, 256, 0, s>>>(h->A32, h->lda, h->Linv, h->y, h->n, nb, h->tk, h->tk_base, h->yx,
                                    ++h->epoch, h->abort);
    h->tk_base += (unsigned)nb;
    k_trsv<false><<<nb, 256, 0, s>>>(h->A32, h->lda, h->Uinv, h->y, h->n, nb, h->tk, h->tk_base, h->yx,
                                     ++h->epoch, h->abort);
    h->tk_base += (unsigned)nb;
    VCHK(hipGetLastError());
    return AFX_OK;
}

// FP64 iterative refinement: x = 0 (or the caller's x when warm); repeat r = b - A x (FP64), stop when the
// normwise backward error |r| / (|A| |x| + |b|) <= tol, else x += LU^-1 P r.
int dense_solve_dev(afx_dense* h, const double* b_dev, double* x_dev, double tol, int max_iter, bool warm,
                    afx_solve_info* info) {
    if (!h->factored) return api_fail(AFX_ESTATE, "no factorisation: call factor first");
    const int n = h->n;
    hipStream_t s = h->s;
    if (tol <= 0.0) tol = std::sqrt((double)n) * 1.1102230246251565e-16;
    if (max_iter <= 0) max_iter = 30;
    if (!warm) VCHK(hipMemsetAsync(x_dev, 0, n * sizeof(double), s));
    afx_solve_info inf{};
    double prev = -1.0;
    const int tb = 256;
    for (int it = 0;; ++it) {
        const bool need_norm = h->anorm < 0.0;
        k_resid64<<<(n + 4 * RES_ROWS - 1) / (4 * RES_ROWS), 256, 0, s>>>(h->A64, n, x_dev, b_dev, h->r,
                                                                         need_norm ? h->rowabs : nullptr);
        k_norms<<<1, 1024, 0, s>>>(h->r, x_dev, b_dev, need_norm ? h->rowabs : nullptr, n, h->norms);
        VCHK(hipGetLastError());
        double nrm[4];
        unsigned aborted = 0;
        VCHK(hipMemcpyAsync(nrm, h->norms, sizeof(nrm), hipMemcpyDeviceToHost, s));
        VCHK(hipMemcpyAsync(&aborted, h->abort, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        VCHK(hipStreamSynchronize(s));
        if (aborted) {
            // the timed-out launch left the solution words incomplete: the
            // factorisation is discarded, and the next factor call clears the
            // flag (k_panel_reset)
            h->factored = false;
            return api_fail(AFX_EHIP, "triangular solve: a block wait timed out (factor again before solving)");
        }
        if (need_norm) h->anorm = nrm[3];
        const double denom = h->anorm * nrm[1] + nrm[2];
        inf.backward_error = denom > 0.0 ? nrm[0] / denom : 0.0;
        if (prev > 0.0) inf.contraction = nrm[0] / prev;
        prev = nrm[0];
        inf.iterations = it;
        if (!std::isfinite(inf.backward_error)) break;
        if (inf.backward_error <= tol) { inf.converged = 1; break; }
        if (it >= max_iter) break;
        k_gather_perm<<<(n + tb - 1) / tb, tb, 0, s>>>(h->r, h->perm, h->y, n);
        VCHK(hipGetLastError());
        int rc = dense_lu_solve(h);
        if (rc) return rc;
        k_add_f32<<<(n + tb - 1) / tb, tb, 0, s>>>(x_dev, h->y, n);
        VCHK(hipGetLastError());
    }
    if (info) *info = inf;
    return AFX_OK;
}

}  // namespace

// ================================================================ VLM =====
struct afx_vlm {
    int nn = 0, nr = 0, ns = 0;
    std::vector<double> xyz;
    std::vector<int> rn, soff;
    afx_vlm_flow flow{};
    bool flow_set = false, built = false, solved = false, have_gamma = false;
    afx_dense* dense = nullptr;
    double *d_nodes = nullptr, *d_corner = nullptr, *d_wcorner = nullptr, *d_colloc = nullptr,
           *d_normal = nullptr, *d_rhs = nullptr, *d_gamma = nullptr, *d_rforce = nullptr, *d_sforce = nullptr;
    int *d_rn = nullptr, *d_soff = nullptr, *d_rst = nullptr;
    double* d_salpha = nullptr;
    ~afx_vlm() {
        for (void* p : {(void*)d_nodes, (void*)d_corner, (void*)d_wcorner, (void*)d_colloc, (void*)d_normal,
                        (void*)d_rhs, (void*)d_gamma, (void*)d_rforce, (void*)d_sforce, (void*)d_rn, (void*)d_soff,
                        (void*)d_rst, (void*)d_salpha})
            if (p) (void)hipFree(p);
        delete dense;
    }
};

namespace {
double flow_kfac(double alpha_deg) {
    // vortexRing::computeCollocationPoint (panel.cpp:111-120): the branch
    // condition keeps k = 1 for every angle below 1e-6 degrees
    if (alpha_deg < 1e-6 && alpha_deg > -1e6) return 1.0;
    const double a = alpha_deg * M_PI / 180.0;
    return alpha_deg * M_PI / (180.0 * std::sin(a));
}
}  // namespace

extern "C" {

int afx_dense_create(int32_t n, afx_dense** out) {
    if (!out) return api_fail(AFX_EARG, "null argument");
    auto* h = new afx_dense();
    int rc = dense_init(h, n);
    if (rc) { delete h; return rc; }
    *out = h;
    return AFX_OK;
}

int afx_dense_destroy(afx_dense* h) {
    delete h;
    return AFX_OK;
}

int afx_dense_set(afx_dense* h, const double* a) {
    if (!h || !a) return api_fail(AFX_EARG, "null argument");
    VCHK(hipMemcpyAsync(h->A64, a, (size_t)h->n * h->n * sizeof(double), hipMemcpyHostToDevice, h->s));
    VCHK(hipStreamSynchronize(h->s));
    h->anorm = -1.0;
    h->factored = false;
    return AFX_OK;
}

int afx_dense_factor(afx_dense* h, int32_t* info) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    return dense_factor(h, info);
}

int afx_dense_solve(afx_dense* h, const double* b, double tol, int32_t max_iter, double* x, afx_solve_info* info) {
    if (!h || !b || !x) return api_fail(AFX_EARG, "null argument");
    const int n = h->n;
    VCHK(hipMemcpyAsync(h->b, b, n * sizeof(double), hipMemcpyHostToDevice, h->s));
    int rc = dense_solve_dev(h, h->b, h->x, tol, max_iter, false, info);
    if (rc) return rc;
    VCHK(hipMemcpyAsync(x, h->x, n * sizeof(double), hipMemcpyDeviceToHost, h->s));
    VCHK(hipStreamSynchronize(h->s));
    return AFX_OK;
}

// Test hook: raise the triangular solves' abort flag as a wait that timed
// out raises it.  The host side of that case follows: the next solve fails
// and discards the factorisation, and the next factor clears the flag
// (k_panel_reset).  No kernel runs here.
int afx_dense_raise_abort(afx_dense* h) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    const unsigned one = 1u;
    VCHK(hipMemcpyAsync(h->abort, &one, sizeof(unsigned), hipMemcpyHostToDevice, h->s));
    VCHK(hipStreamSynchronize(h->s));
    return AFX_OK;
}

int afx_dense_get_lu(afx_dense* h, float* lu, int32_t* ipiv) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    const int n = h->n;
    if (lu) VCHK(hipMemcpy2DAsync(lu, n * sizeof(float), h->A32, h->lda * sizeof(float), n * sizeof(float), n,
                                  hipMemcpyDeviceToHost, h->s));
    if (ipiv) VCHK(hipMemcpyAsync(ipiv, h->ipiv, n * sizeof(int), hipMemcpyDeviceToHost, h->s));
    VCHK(hipStreamSynchronize(h->s));
    return AFX_OK;
}

int afx_vlm_create(int32_t n_nodes, const double* xyz, int32_t n_stations, const int32_t* soff,
                   const int32_t* ring_nodes, afx_vlm** out) {
    if (!xyz || !soff || !ring_nodes || !out) return api_fail(AFX_EARG, "null argument");
    if (n_nodes <= 0 || n_stations <= 0) return api_fail(AFX_EARG, "empty lattice");
    if (soff[0] != 0) return api_fail(AFX_EARG, "station_offsets[0] must be 0");
    for (int s = 0; s < n_stations; ++s)
        if (soff[s + 1] <= soff[s]) return api_fail(AFX_EARG, "every station needs at least one ring");
    const int nr = soff[n_stations];
    for (int k = 0; k < 4 * nr; ++k)
        if (ring_nodes[k] < 0 || ring_nodes[k] >= n_nodes) return api_fail(AFX_EARG, "ring node id out of range");
    auto* h = new afx_vlm();
    auto bail = [&](int rc) { delete h; return rc; };
    h->nn = n_nodes;
    h->nr = nr;
    h->ns = n_stations;
    h->xyz.assign(xyz, xyz + 3 * (size_t)n_nodes);
    h->rn.assign(ring_nodes, ring_nodes + 4 * (size_t)nr);
    h->soff.assign(soff, soff + n_stations + 1);
    h->dense = new afx_dense();
    int rc = dense_init(h->dense, nr);
    if (rc) return bail(rc);
#define VCHK_C(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return bail(api_fail(AFX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_))); \
    } while (0)
    VCHK_C(dalloc(&h->d_nodes, 3 * (size_t)n_nodes));
    VCHK_C(dalloc(&h->d_corner, 12 * (size_t)nr));
    VCHK_C(dalloc(&h->d_wcorner, 12 * (size_t)n_stations));
    VCHK_C(dalloc(&h->d_colloc, 3 * (size_t)nr));
    VCHK_C(dalloc(&h->d_normal, 3 * (size_t)nr));
    VCHK_C(dalloc(&h->d_rhs, nr));
    VCHK_C(dalloc(&h->d_gamma, nr));
    VCHK_C(dalloc(&h->d_rforce, 3 * (size_t)nr));
    VCHK_C(dalloc(&h->d_sforce, 3 * (size_t)n_stations));
    VCHK_C(dalloc(&h->d_rn, 4 * (size_t)nr));
    VCHK_C(dalloc(&h->d_soff, n_stations + 1));
    VCHK_C(dalloc(&h->d_rst, nr));
    VCHK_C(dalloc(&h->d_salpha, n_stations));
    std::vector<int> rst(nr);
    for (int st = 0; st < n_stations; ++st)
        for (int k = soff[st]; k < soff[st + 1]; ++k) rst[k] = st;
    hipStream_t s = h->dense->s;
    VCHK_C(hipMemcpyAsync(h->d_nodes, xyz, 3 * (size_t)n_nodes * sizeof(double), hipMemcpyHostToDevice, s));
    VCHK_C(hipMemcpyAsync(h->d_rn, ring_nodes, 4 * (size_t)nr * sizeof(int), hipMemcpyHostToDevice, s));
    VCHK_C(hipMemcpyAsync(h->d_soff, soff, (n_stations + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    VCHK_C(hipMemcpyAsync(h->d_rst, rst.data(), nr * sizeof(int), hipMemcpyHostToDevice, s));
    VCHK_C(hipStreamSynchronize(s));
#undef VCHK_C
    *out = h;
    return AFX_OK;
}

int afx_vlm_destroy(afx_vlm* h) {
    delete h;
    return AFX_OK;
}

int afx_vlm_counts(const afx_vlm* h, int32_t* nn, int32_t* nr, int32_t* ns) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    if (nn) *nn = h->nn;
    if (nr) *nr = h->nr;
    if (ns) *ns = h->ns;
    return AFX_OK;
}

int afx_vlm_set_flow(afx_vlm* h, const afx_vlm_flow* f) {
    if (!h || !f) return api_fail(AFX_EARG, "null argument");
    if (!(f->vinf > 0.0) || !(f->rho > 0.0) || !(f->sref > 0.0) || !(f->cref > 0.0))
        return api_fail(AFX_EARG, "vinf, rho, sref and cref must be positive");
    if (f->core_radius < 0.0) return api_fail(AFX_EARG, "core radius must be >= 0");
    h->flow = *f;
    if (!(h->flow.wake_length > 0.0)) h->flow.wake_length = 100.0 * f->cref;   // solver.cpp:36
    h->flow_set = true;
    h->built = h->solved = false;
    return AFX_OK;
}

int afx_vlm_set_nodes(afx_vlm* h, const double* xyz) {
    if (!h || !xyz) return api_fail(AFX_EARG, "null argument");
    std::copy(xyz, xyz + 3 * (size_t)h->nn, h->xyz.begin());
    VCHK(hipMemcpyAsync(h->d_nodes, xyz, 3 * (size_t)h->nn * sizeof(double), hipMemcpyHostToDevice, h->dense->s));
    VCHK(hipStreamSynchronize(h->dense->s));
    h->built = h->solved = false;
    return AFX_OK;
}

int afx_vlm_get_nodes(const afx_vlm* h, double* xyz) {
    if (!h || !xyz) return api_fail(AFX_EARG, "null argument");
    std::copy(h->xyz.begin(), h->xyz.end(), xyz);
    return AFX_OK;
}

int afx_vlm_get_rings(const afx_vlm* h, int32_t* ring_nodes, int32_t* soff) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    if (ring_nodes) std::copy(h->rn.begin(), h->rn.end(), ring_nodes);
    if (soff) std::copy(h->soff.begin(), h->soff.end(), soff);
    return AFX_OK;
}

int afx_vlm_build(afx_vlm* h) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    if (!h->flow_set) return api_fail(AFX_ESTATE, "flow not set");
    const afx_vlm_flow& f = h->flow;
    const double a = f.alpha_deg * M_PI / 180.0;
    const double vx = f.vinf * std::cos(a), vy = 0.0, vz = f.vinf * std::sin(a);
    const double kf = flow_kfac(f.alpha_deg);
    const int nr = h->nr, ns = h->ns;
    hipStream_t s = h->dense->s;
    k_ring_geom<<<(nr + 127) / 128, 128, 0, s>>>(h->d_nodes, h->d_rn, nr, kf, vx, vy, vz, h->d_corner, h->d_colloc,
                                                 h->d_normal, h->d_rhs);
    k_wake<<<(ns + 127) / 128, 128, 0, s>>>(h->d_corner, h->d_soff, ns, f.wake_length, std::cos(a), 0.0, std::sin(a),
                                            h->d_wcorner);
    dim3 g((nr + 63) / 64, (nr + 63) / 64);
    k_aic<<<g, 256, 0, s>>>(h->d_colloc, h->d_normal, h->d_corner, nr, f.core_radius, h->dense->A64);
    const long long nw = (long long)nr * ns;
    k_aic_wake<<<(unsigned)((nw + 255) / 256), 256, 0, s>>>(h->d_colloc, h->d_normal, h->d_wcorner, h->d_soff, ns,
                                                            nr, f.core_radius, h->dense->A64);
    VCHK(hipGetLastError());
    VCHK(hipStreamSynchronize(s));
    h->dense->anorm = -1.0;
    h->built = true;
    h->solved = false;
    return AFX_OK;
}

int afx_vlm_factor(afx_vlm* h) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    if (!h->built) return api_fail(AFX_ESTATE, "matrix not built");
    int32_t info = 0;
    return dense_factor(h->dense, &info);
}

int afx_vlm_solve(afx_vlm* h, double tol, int32_t max_iter, int32_t warm, double* gamma, afx_solve_info* info) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    if (!h->built) return api_fail(AFX_ESTATE, "matrix not built");
    afx_solve_info inf{};
    int rc = dense_solve_dev(h->dense, h->d_rhs, h->d_gamma, tol, max_iter, warm != 0 && h->have_gamma, &inf);
    if (rc) return rc;
    if (info) *info = inf;
    if (gamma) {
        VCHK(hipMemcpyAsync(gamma, h->d_gamma, h->nr * sizeof(double), hipMemcpyDeviceToHost, h->dense->s));
        VCHK(hipStreamSynchronize(h->dense->s));
    }
    h->solved = h->have_gamma = true;
    return AFX_OK;
}

int afx_vlm_forces(afx_vlm* h, double* ring_force, double* station_force, double* coeffs) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    if (!h->solved) return api_fail(AFX_ESTATE, "no solution: call afx_vlm_solve first");
    const afx_vlm_flow& f = h->flow;
    const double a = f.alpha_deg * M_PI / 180.0;
    const double vx = f.vinf * std::cos(a), vz = f.vinf * std::sin(a);
    hipStream_t s = h->dense->s;
    k_forces<<<(h->ns + 127) / 128, 128, 0, s>>>(h->d_corner, h->d_gamma, h->d_soff, h->ns, f.rho, vx, 0.0, vz,
                                                 h->d_rforce, h->d_sforce);
    VCHK(hipGetLastError());
    std::vector<double> sf(3 * (size_t)h->ns);
    VCHK(hipMemcpyAsync(sf.data(), h->d_sforce, sf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    if (ring_force)
        VCHK(hipMemcpyAsync(ring_force, h->d_rforce, 3 * (size_t)h->nr * sizeof(double), hipMemcpyDeviceToHost, s));
    VCHK(hipStreamSynchronize(s));
    if (station_force) std::copy(sf.begin(), sf.end(), station_force);
    if (coeffs) {
        // lift axis = freestream x unit-y (input.cpp:46-52), stream axis = freestream
        const double q = 0.5 * f.rho * f.vinf * f.vinf;
        const double lx = -std::sin(a), lz = std::cos(a);
        double cl = 0.0, cy = 0.0, cd = 0.0;
        for (int st = 0; st < h->ns; ++st) {
            const double* F = &sf[3 * st];
            cl += (F[0] * lx + F[2] * lz) / (q * f.sref);
            cy += F[1] / (q * f.sref);
            cd += (F[0] * std::cos(a) + F[2] * std::sin(a)) / (q * f.sref);
        }
        coeffs[0] = cl;
        coeffs[1] = cy;
        coeffs[2] = cd;
    }
    return AFX_OK;
}

}  // extern "C"

namespace {
constexpr double VLM_CLA_DEG = 0.109662271123215082635482531259185634553432464599609375;   // 2 pi per radian

// mlinterp 1D (libs/headeronly/mlinterp): linear inside, end values outside
double interp1(const double* xd, const double* yd, int n, double x) {
    if (n == 1 || x <= xd[0]) return yd[0];
    if (x >= xd[n - 1]) return yd[n - 1];
    int lo = 0, hi = n - 2;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (x < xd[mid]) hi = mid - 1;
        else if (x >= xd[mid + 1]) lo = mid + 1;
        else {
            const double w = (xd[mid + 1] - x) / (xd[mid + 1] - xd[mid]);
            return w * yd[mid] + (1.0 - w) * yd[mid + 1];
        }
    }
    return yd[n - 1];
}

// database::table::coefficients (database.cpp:192-242): each section's
// airfoil at alpha, then across the span
void polar_coeffs(const afx_polar& P, double alpha, double span, double* out3) {
    std::vector<double> c[3];
    for (int q = 0; q < P.n_sections; ++q) {
        const int a = P.section_airfoil[q];
        const int o = P.offsets[a], n = P.offsets[a + 1] - o;
        c[0].push_back(interp1(P.alpha_deg + o, P.cl + o, n, alpha));
        c[1].push_back(interp1(P.alpha_deg + o, P.cd + o, n, alpha));
        c[2].push_back(interp1(P.alpha_deg + o, P.cmy + o, n, alpha));
    }
    for (int k = 0; k < 3; ++k) out3[k] = interp1(P.section_span, c[k].data(), P.n_sections, span);
}
}  // namespace

extern "C" {

int afx_vlm_nonlinear_solve(afx_vlm* h, const afx_polar* P, double relaxation, double tol, int32_t max_iter,
                            afx_nonlinear_info* info, double* station_alpha, double* station_cl) {
    if (!h || !P) return api_fail(AFX_EARG, "null argument");
    if (!h->flow_set) return api_fail(AFX_ESTATE, "flow not set");
    if (P->n_airfoils < 1 || P->n_sections < 1 || !P->offsets || !P->alpha_deg || !P->cl || !P->cd || !P->cmy ||
        !P->section_airfoil || !P->section_span)
        return api_fail(AFX_EARG, "polar table is empty");
    for (int a = 0; a < P->n_airfoils; ++a)
        if (P->offsets[a + 1] <= P->offsets[a]) return api_fail(AFX_EARG, "an airfoil polar has no rows");
    for (int q = 0; q < P->n_sections; ++q)
        if (P->section_airfoil[q] < 0 || P->section_airfoil[q] >= P->n_airfoils)
            return api_fail(AFX_EARG, "section airfoil index out of range");
    const afx_vlm_flow& f = h->flow;
    const int ns = h->ns;
    int rc = afx_vlm_build(h);
    if (rc) return rc;
    if ((rc = afx_vlm_factor(h))) return rc;
    // station geometry (wingStation / wing, surface.cpp): area = sum of ring
    // areas (panel.cpp:64-72), leading-edge segment of the first ring, span
    // location (wing::computeSpan) normalised from the first station
    std::vector<double> area(ns, 0.0), le(3 * (size_t)ns), spanloc(ns);
    auto node = [&](int id, int d) { return h->xyz[3 * (size_t)id + d]; };
    double span = 0.0;
    for (int st = 0; st < ns; ++st) {
        for (int k = h->soff[st]; k < h->soff[st + 1]; ++k) {
            double c[3] = {0, 0, 0};
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) c[d] += node(h->rn[4 * k + a], d) / 4.0;
            for (int a = 0; a < 4; ++a) {
                const int i0 = h->rn[4 * k + a], i1 = h->rn[4 * k + ((a + 1) & 3)];
                const double dl[3] = {node(i1, 0) - node(i0, 0), node(i1, 1) - node(i0, 1), node(i1, 2) - node(i0, 2)};
                const double ce[3] = {c[0] - node(i0, 0), c[1] - node(i0, 1), c[2] - node(i0, 2)};
                const double x = dl[1] * ce[2] - dl[2] * ce[1], y = dl[2] * ce[0] - dl[0] * ce[2],
                             z = dl[0] * ce[1] - dl[1] * ce[0];
                area[st] += 0.5 * std::sqrt(x * x + y * y + z * z);
            }
        }
        const int k0 = h->soff[st];
        for (int d = 0; d < 3; ++d) le[3 * st + d] = node(h->rn[4 * k0 + 1], d) - node(h->rn[4 * k0], d);
        const double len = std::sqrt(le[3 * st] * le[3 * st] + le[3 * st + 1] * le[3 * st + 1] +
                                     le[3 * st + 2] * le[3 * st + 2]);
        span += len;
        spanloc[st] = span - 0.5 * len;
    }
    for (int st = ns - 1; st >= 0; --st) spanloc[st] = (spanloc[st] - spanloc[0]) / span;
    const double q = 0.5 * f.rho * f.vinf * f.vinf;
    std::vector<double> alpha(ns, f.alpha_deg), sf(3 * (size_t)ns), clinv(ns);
    hipStream_t s = h->dense->s;
    auto lift_axis = [&](int st, double* ax) {   // unit(V(alpha_s) x dl_LE)
        const double a = alpha[st] * M_PI / 180.0, vx = std::cos(a), vz = std::sin(a);
        const double* d = &le[3 * st];
        ax[0] = 0.0 * d[2] - vz * d[1];
        ax[1] = vz * d[0] - vx * d[2];
        ax[2] = vx * d[1] - 0.0 * d[0];
        const double n = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        for (int k = 0; k < 3; ++k) ax[k] /= n;
    };
    afx_nonlinear_info inf{};
    int iter = 0;
    double residual = 0.0;
    do {
        VCHK(hipMemcpyAsync(h->d_salpha, alpha.data(), ns * sizeof(double), hipMemcpyHostToDevice, s));
        k_rhs_stations<<<(h->nr + 127) / 128, 128, 0, s>>>(h->d_normal, h->d_rst, h->nr, f.vinf, h->d_salpha, h->d_rhs);
        VCHK(hipGetLastError());
        afx_solve_info si{};
        if ((rc = dense_solve_dev(h->dense, h->d_rhs, h->d_gamma, 0.0, 30, iter > 0, &si))) return rc;
        if (!si.converged) return api_fail(AFX_EARG, "vortex-lattice refinement did not converge");
        h->solved = h->have_gamma = true;
        k_forces_stations<<<(ns + 127) / 128, 128, 0, s>>>(h->d_corner, h->d_gamma, h->d_soff, ns, f.rho, f.vinf,
                                                           h->d_salpha, h->d_sforce);
        VCHK(hipGetLastError());
        VCHK(hipMemcpyAsync(sf.data(), h->d_sforce, sf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        VCHK(hipStreamSynchronize(s));
        residual = 0.0;
        for (int st = 0; st < ns; ++st) {
            double ax[3];
            lift_axis(st, ax);
            clinv[st] = (sf[3 * st] * ax[0] + sf[3 * st + 1] * ax[1] + sf[3 * st + 2] * ax[2]) / (q * area[st]);
            const double aoa_eff = clinv[st] / VLM_CLA_DEG - alpha[st] + f.alpha_deg;
            double c3[3];
            polar_coeffs(*P, aoa_eff, spanloc[st], c3);
            alpha[st] += relaxation * (c3[0] - clinv[st]) / VLM_CLA_DEG;
            const double r = (c3[0] - clinv[st]) / c3[0];
            residual += r * r;
        }
        residual = std::sqrt(residual);
        ++iter;
    } while (residual > tol && iter <= max_iter);
    // nonlinear::steady::computeForces (solver.cpp:434-476): viscous section
    // coefficients at the final effective angles
    const double A = f.alpha_deg * M_PI / 180.0;
    const double lx = -std::sin(A), lz = std::cos(A);
    double cl = 0.0, cd = 0.0, cy = 0.0;
    for (int st = 0; st < ns; ++st) {
        const double aoa_eff = clinv[st] / VLM_CLA_DEG - alpha[st] + f.alpha_deg;
        double c3[3], ax[3];
        polar_coeffs(*P, aoa_eff, spanloc[st], c3);
        lift_axis(st, ax);
        const double scl = c3[0] * (ax[0] * lx + ax[2] * lz);
        cl += scl * area[st] / f.sref;
        cy += c3[0] * ax[1] * area[st] / f.sref;
        cd += c3[1] * area[st] / f.sref;
        if (station_cl) station_cl[st] = scl;
    }
    if (station_alpha) std::copy(alpha.begin(), alpha.end(), station_alpha);
    inf.iterations = iter;
    inf.converged = residual <= tol;
    inf.residual = residual;
    inf.cl = cl;
    inf.cd = cd;
    inf.cy = cy;
    if (info) *info = inf;
    return AFX_OK;
}

int afx_vlm_get_aic(afx_vlm* h, double* aic) {
    if (!h || !aic) return api_fail(AFX_EARG, "null argument");
    if (!h->built) return api_fail(AFX_ESTATE, "matrix not built");
    VCHK(hipMemcpyAsync(aic, h->dense->A64, (size_t)h->nr * h->nr * sizeof(double), hipMemcpyDeviceToHost,
                        h->dense->s));
    VCHK(hipStreamSynchronize(h->dense->s));
    return AFX_OK;
}

int afx_vlm_get_system(afx_vlm* h, double* rhs, double* gamma, double* colloc, double* normal) {
    if (!h) return api_fail(AFX_EARG, "null argument");
    hipStream_t s = h->dense->s;
    const size_t nr = h->nr;
    if (rhs) VCHK(hipMemcpyAsync(rhs, h->d_rhs, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    if (gamma) VCHK(hipMemcpyAsync(gamma, h->d_gamma, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    if (colloc) VCHK(hipMemcpyAsync(colloc, h->d_colloc, 3 * nr * sizeof(double), hipMemcpyDeviceToHost, s));
    if (normal) VCHK(hipMemcpyAsync(normal, h->d_normal, 3 * nr * sizeof(double), hipMemcpyDeviceToHost, s));
    VCHK(hipStreamSynchronize(s));
    return AFX_OK;
}

}  // extern "C"
