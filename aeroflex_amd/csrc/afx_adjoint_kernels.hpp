// Launcher of the state adjoint of the FP64 residual (afx_adjoint.hip).
#pragma once
#include "afx_kernels.hpp"

namespace afx {

// work arrays of one adjoint evaluation ([4][nT] / [4][nC] / [4][nF] planes)
struct AdjBufs {
    double* gx;      // [4][nT] cell gradients
    double* gy;
    double* lim;     // [4][nT] limiters
    int8_t* act;     // [4][nC] face slot that gives the limiter (-1: none)
    int32_t* jmax;   // [4][nC] neighbour giving q_max / q_min (-1: the cell itself)
    int32_t* jmin;
    double* bL;      // [4][nF] adjoints of the reconstructed owner / neighbour states
    double* bR;
    double* lgx;     // [4][nT] gradient adjoints
    double* lgy;
    double* nbv;     // [4][nC] limiter adjoint owed to neighbour nbj
    int32_t* nbj;
    double* co;      // [4][nF] Green-Gauss transpose: owner / neighbour parts
    double* cn;
};

// lam[k*nT + i] (real cells) = ((dR/dq)^T w)[i, k] for the inviscid residual
// with Green-Gauss gradients; w in the same layout (ghost rows ignored)
hipError_t launch_adjoint_state(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* w,
                                GasT<double> g, int second_order, double limiter_k, AdjBufs A, double* lam);

}  // namespace afx
