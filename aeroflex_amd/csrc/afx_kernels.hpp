// Launcher interface between the C-ABI layer and the HIP kernels.
//
// Device layout (all structure-of-arrays, GPU cell numbering):
//   cells  0..nC-1 are real cells in the device order (RCM), nC..nT-1 ghost
//          cells in the reference's boundary-segment order;
//   state  q[k*nT + i], k < NV;
//   faces  in the device face order; c0 = reference owner, c1 = neighbour or
//          ghost (device numbers); kind = face functor type;
//   cface  [slot*nC + i]: the faces of real cell i sorted by their reference
//          face index (so every per-cell sum runs in the reference's serial
//          edge-loop order), bit 31 set when i is the face's neighbour,
//          0xFFFFFFFF padding for triangles.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace afx {

enum FaceKind : int8_t { K_INTERNAL = 0, K_FARFIELD = 1, K_SLIP = 2, K_WALL = 3 };

template <typename T>
struct GasT {
    T gamma, R, mu, k;   // k = cp*mu_L/Pr_L with cp = 1 (core.h:32,43-45)
    T cp, PrL, PrT;      // turbulent conductivity cp (mu_L/Pr_L + mu_t/Pr_T) of the SA model
};

template <typename T>
struct FacesT {
    const int32_t* c0;
    const int32_t* c1;
    const T* nx;
    const T* ny;
    const T* len;
    const T* xf;
    const T* yf;
    const T* gw;         // Green-Gauss weight |x_f - x_0| / |x_0 - x_1| (solver.h:436-444)
    const int8_t* kind;
    int32_t nF;
};

template <typename T>
struct CellsT {
    const T* xc;
    const T* yc;
    const T* area;
    const uint32_t* cface;   // 4*nC
    const T* lsq;            // 4*nC, (d^T d)^-1 row-major (least-squares only)
    int32_t nC;
    int32_t nT;
};

// Per-tile tables of the fused residual (afx_tiles.hpp), device side.
template <typename T>
struct TilesT {
    const int32_t* cell0;    // ntiles+1: first device cell of each tile
    const int32_t* loff;     // ntiles+1: local-cell table offsets
    const int32_t* ng;       // ntiles: gradient cells (T u H1)
    const int32_t* foff;     // ntiles+1: local-face table offsets
    const int32_t* nft;      // ntiles: faces touching T
    const int32_t* goff;     // ntiles+1: slot table offsets
    const int32_t* Lglob;    // device cell index of each local cell (-1 - index: wall/slip ghost, not staged)
    const uint64_t* slots;   // per G cell: 4 x u16 local face (bit 15: cell is the neighbour), 0xFFFF pad
    const uint32_t* fcell;   // per local face: c0 | c1 << 16 (local cells)
    const int8_t* fkind;     // per local face
    const T* geo;            // 8 x nFtot: nx, ny, len, gw, d0x, d0y, d1x, d1y (d = x_f - x_cell)
    const int32_t* nbf;      // ntiles: boundary faces in the tile's list (0: interior tile); nullable
    int64_t nFtot;
    int32_t ntiles;
    int32_t lmax, gmax, ftmax;
};

// nv = 4 (Euler / laminar) or 5 (Spalart-Allmaras: rho nu~ as fifth variable)
template <typename T>
inline size_t tile_lds_bytes(int32_t lmax, int32_t gmax, int32_t ftmax, int nv = 4) {
    return sizeof(T) * (size_t)nv * (size_t)(lmax + 3 * gmax + ftmax);
}

// fused residual of every tile: gradients, limiters, fluxes and cell sums in
// LDS; R written once per cell; wall/slip ghost states of the tile's cells
// written to qghost (set_walls_from_internal) when non-null.  viscous: 0
// inviscid, 1 laminar, 2 Spalart-Allmaras (nv = 5 planes of q and R; wd = wall
// distance per device cell)
template <typename T>
hipError_t launch_residual_tile(hipStream_t s, TilesT<T> S, const T* q, int32_t nT, int32_t nC, const T* area,
                                const T* lsq, const T* xc, const T* yc, const T* wd, GasT<T> g, int gg,
                                int second_order, T limiter_k, int viscous, T* R, T* qghost);

// FP32 throughput form of the same (afx_k32.hip): cell states, gradients,
// limiters, fluxes and residuals are float4 per cell / face (rho, rho u,
// rho v, rho E); per tile face two float4 (nx, ny, len, g) and (d0x, d0y,
// d1x, d1y) and one word c0 | c1 << 12 | kind << 24.
struct TilesF32 {
    const int32_t* cell0;
    const int32_t* loff;
    const int32_t* ng;
    const int32_t* foff;
    const int32_t* nft;
    const int32_t* goff;
    const int32_t* Lglob;
    const uint64_t* slots;
    const uint32_t* fck;
    const float4* geo;
    const int32_t* nbf;      // per tile: boundary faces in its face list (0: interior tile)
    int32_t ntiles;
    int32_t lmax, gmax, ftmax;
};
// iarea = 1/A, a15 = A^1.5, K3 = limiter_k^3 (K3a of solver.h:551-552).
// Spalart-Allmaras (viscous == 2): rho nu~ in qn / Rn / qghost_n (one float
// per cell) and (area, wall distance) per cell in sa_ad.
struct SAFields32 {
    const float* qn;
    const float2* sa_ad;
    float* Rn;
    float* qghost_n;
};
hipError_t launch_residual_tile_f32(hipStream_t s, TilesF32 S, const float4* q, int32_t nC, const float* iarea,
                                    const float* a15, const float* lsq, const float* xc, const float* yc,
                                    GasT<float> g, int gg, int second_order, float K3, int viscous, float4* R,
                                    float4* qghost, SAFields32 sa);

// Everything one FP32 tile residual reads and writes (one mesh, one flow
// condition).  dt != nullptr: also the local time step of calc_dt
// (solver.h:308-356), cfl A / sum_e lambda_e len_e, from the unreconstructed
// cell states.
// qout != nullptr: an explicit stage (solver.h:818-823) instead of writing R:
// qout = q0 + coef dt R / A in FP64 from the FP64 cell sums, with the FP32
// copy of qout in q32out / q32nout (a different buffer from the one read, so
// no block reads a state another block of the same launch rewrites); dt from
// this launch (dt != nullptr) or from dt_in.
struct TileArgsF32 {
    TilesF32 S;
    const float4* q;
    const float* iarea;
    const float* a15;
    const float* lsq;
    const float* xc;
    const float* yc;
    float4* R;
    float4* qghost;
    SAFields32 sa;
    float* dt;
    GasT<float> g;
    float K3, cfl;
    int32_t nC, gg, second_order, viscous;
    // explicit stage (batched form)
    const double* q0;
    double* qout;
    float4* q32out;
    float* q32nout;
    const float* dt_in;
    const double* area64;
    double coef;
    int32_t nT;
};
inline size_t tile_lds_bytes_f32(const TilesF32& S, int nv, bool dt) {
    return tile_lds_bytes<float>(S.lmax, S.gmax, S.ftmax, nv) +
           (dt ? sizeof(float) * (size_t)S.ftmax * (nv == 5 ? 2 : 1) : 0);
}
// one launch over the (member, tile) work items of many meshes / conditions;
// items padded with member -1 to a multiple of 8 blocks; lds = the largest
// member's tile_lds_bytes_f32; every member runs the same viscosity model
// (0/1/2), gradient scheme (gg) and order
hipError_t launch_residual_tile_f32_batch(hipStream_t s, const TileArgsF32* members, const int2* items, int32_t nitems,
                                          size_t lds, int viscous, int gg, int second_order);

// SA explicit step of one solver: qk = q + R / A dt coef (real cells, nv planes), qW = R / A
hipError_t launch_rk_update_area(hipStream_t s, const double* q, const double* R, const double* area, const double* dt,
                                 double coef, int32_t nC, int32_t nT, int nv, double* qk, double* qW);
hipError_t launch_to_f32x4(hipStream_t s, const double* src, int32_t nT, float4* dst);

// ---- launchers (instantiated for double in afx_k64.hip, float in afx_k32.hip)
template <typename T>
hipError_t launch_set_walls(hipStream_t s, const int32_t* owner, const int32_t* ghost, int32_t n, T* q, int32_t nT, int nv);

// calc_dt (solver.h:308-356); sa = 1: plus the viscous spectral radius of
// the SA model (q with five planes)
template <typename T>
hipError_t launch_dt(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* q, GasT<T> g, T cfl, T* dt, int sa = 0);

// gradients from qg (solver.h:425-514) and limiters from ql (solver.h:517-593)
template <typename T>
hipError_t launch_grad_lim(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* qg, const T* ql, GasT<T> g,
                           int gg, int second_order, T limiter_k, T* gx, T* gy, T* lim);

// per-face flux times length (solver.h:1097-1134); uniform != nullptr -> both
// sides take that constant state with no reconstruction (solver.h:659-679)
template <typename T>
hipError_t launch_face_flux(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* qr, const T* qa, const T* gx,
                            const T* gy, const T* lim, GasT<T> g, int viscous, int second_order,
                            const T* uniform, T* flux);

// per-cell signed sum of face fluxes in reference face order.
// mode 0: R = sum            (fillRhoRHS)
// mode 1: R = sum / area     (explicitSolver::calc_residual)
// mode 2: R = R_in + sum     (get_uniform_residual on a stale qW)
// FP32 instantiation accumulates with Neumaier compensation.
template <typename T>
hipError_t launch_gather(hipStream_t s, CellsT<T> C, const T* flux, int32_t nF, int mode, T* R);

// ---- FP64-only helpers (afx_k64.hip)
hipError_t launch_sumsq(hipStream_t s, const double* x, int64_t n, int64_t ld, int nv, double* partial, int nblk);
hipError_t launch_sumsq_f32(hipStream_t s, const float* x, int64_t n, int64_t ld, int nv, double* partial, int nblk);
hipError_t launch_final_sum(hipStream_t s, const double* partial, int nblk, double* out);
hipError_t launch_copy(hipStream_t s, const double* src, double* dst, int64_t n);
hipError_t launch_rk_update(hipStream_t s, const double* q, const double* qW, const double* dt, double ai, double relax,
                            int32_t nC, int32_t nT, int nv, double* qk);
hipError_t launch_to_f32(hipStream_t s, const double* src, float* dst, int64_t n);
hipError_t launch_wall_dist(hipStream_t s, const double* xc, const double* yc, int32_t nT, const double* wx,
                            const double* wy, int32_t nw, double* d);
// the Spalart-Allmaras model's distance to the wall / slip-wall face segments
// (seg: {ax, ay, bx - ax, by - ay} per face)
hipError_t launch_wall_dist_seg(hipStream_t s, const double* xc, const double* yc, int32_t nT, const double4* seg,
                                int32_t nw, double* d);
hipError_t launch_gather_states(hipStream_t s, const double* q, int32_t nT, int nv, const int32_t* cells, int32_t n,
                                double* out);

}  // namespace afx
