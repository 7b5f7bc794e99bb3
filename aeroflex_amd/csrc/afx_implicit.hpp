// Launchers of the implicit-step kernels (afx_implicit.hip), FP64.
#pragma once
#include "afx_kernels.hpp"

namespace afx {

hipError_t launch_face_jacobian(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* gx,
                                const double* gy, GasT<double> g, int viscous, double* Jf);
hipError_t launch_assemble_diag(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* dt, const double* Jf,
                                double* D, double* Dinv, int* bad);
hipError_t launch_lhs_apply(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* D, const double* Jf,
                            const double* x, double* y);
hipError_t launch_block_jacobi(hipStream_t s, int32_t nC, int32_t nT, const double* Dinv, const double* x, double* z);
hipError_t launch_multidot(hipStream_t s, const double* V, int64_t ld, int m, const double* w, int64_t n,
                           double* partial, int nblk, double* out);
hipError_t launch_multiaxpy(hipStream_t s, const double* V, int64_t ld, int m, const double* h, double* w, int64_t n);
hipError_t launch_lincomb(hipStream_t s, const double* V, int64_t ld, int m, const double* y, double* x, int64_t n);
hipError_t launch_scale(hipStream_t s, double a, const double* x, double* y, int64_t n);
hipError_t launch_axpy(hipStream_t s, double a, const double* x, double* y, int64_t n);
// matrix-free operator of the RANS_MATRIX_FREE build (ransMatrix.h:138-178)
hipError_t launch_mf_base(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                          int viscous, double* Fb);
hipError_t launch_mf_apply(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* dt,
                           const double* Fb, GasT<double> g, int viscous, const double* x, double* y);

}  // namespace afx
