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

struct ResidualArgs64;  // fwd

// ---- launchers (instantiated for double in afx_k64.hip, float in afx_k32.hip)
template <typename T>
hipError_t launch_set_walls(hipStream_t s, const int32_t* owner, const int32_t* ghost, int32_t n, T* q, int32_t nT, int nv);

template <typename T>
hipError_t launch_dt(hipStream_t s, CellsT<T> C, FacesT<T> F, const T* q, GasT<T> g, T cfl, T* dt);

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
hipError_t launch_gather_states(hipStream_t s, const double* q, int32_t nT, int nv, const int32_t* cells, int32_t n,
                                double* out);

}  // namespace afx
