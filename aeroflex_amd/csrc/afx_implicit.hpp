// Launchers of the implicit-step kernels (afx_implicit.hip), FP64.
#pragma once
#include "afx_kernels.hpp"

namespace afx {

hipError_t launch_face_jacobian(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* gx,
                                const double* gy, GasT<double> g, int viscous, double* Jf);
// the Spalart-Allmaras model's 5x5 face blocks (q, gx, gy: 5 planes)
hipError_t launch_face_jacobian_sa(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* gx,
                                   const double* gy, GasT<double> g, double* Jf);
// the source row of the Spalart-Allmaras diagonal blocks (q null: none)
struct SaSource {
    const double* q;    // 5 planes
    const double* gx;   // unlimited cell gradients, 5 planes
    const double* gy;
    const double* wd;   // wall distance
    GasT<double> g;
};
// nv: block size, 4 or 5 (Spalart-Allmaras).  Also writes the ELL copy of
// the matrix for launch_lhs_apply: E 4 nv^2 nC off-diagonal blocks by slot,
// Ds nv^2 nC diagonal blocks, nbr 4 nC neighbour indices (-1: none), all
// slot/entry-major over the real cells
hipError_t launch_assemble_diag(hipStream_t s, int nv, CellsT<double> C, FacesT<double> F, const double* dt,
                                const double* Jf, SaSource src, double* D, double* Dinv, double* E, double* Ds,
                                int32_t* nbr, int* bad);
// y = LHS x from the ELL copy.  active: a device flag (nullable); the kernel
// does nothing while it reads 0
hipError_t launch_lhs_apply(hipStream_t s, int nv, int32_t nC, int32_t nT, const double* Ds, const double* E,
                            const int32_t* nbr, const double* x, double* y, const int* active = nullptr);
hipError_t launch_block_jacobi(hipStream_t s, int nv, int32_t nC, int32_t nT, const double* Dinv, const double* x,
                               double* z, const int* active = nullptr);
// multicolour block DILU: colour per real cell, cells listed colour by colour
// (cells[coff[c] .. coff[c+1]), coff on the host), pos the inverse list (the
// position of each real cell).  The set-up also writes the sweeps'
// colour-ordered copy by position p: Ec 4 nv^2 nC off-diagonal blocks by
// slot, nc 4 nC neighbour codes (q + 1 for the neighbour at position q of an
// earlier colour, -(q + 1) of a later one, 0 none), Dc nv^2 nC inverted
// diagonal blocks.  The sweeps work on yp (nv planes of nC, by position) and
// a last pass gathers it into z in cell order.
hipError_t launch_dilu_setup(hipStream_t s, int nv, CellsT<double> C, FacesT<double> F, const uint8_t* colour,
                             const int32_t* cells, const int32_t* pos, const int32_t* coff, int ncolours,
                             const double* D, const double* Jf, double* Dti, double* Ec, int32_t* nc, double* Dc,
                             int* bad);
hipError_t launch_dilu_apply(hipStream_t s, int nv, int32_t nC, int32_t nT, const int32_t* cells, const int32_t* pos,
                             const int32_t* coff, int ncolours, const double* Ec, const int32_t* nc, const double* Dc,
                             const double* x, double* z, double* yp, const int* active = nullptr);

// device-resident restarted GMRES state (m <= 32)
struct GmDev {
    double* hdr;    // [0] beta0, [1] beta of the cycle, [2] scale of the next basis vector, [3] a reduced |w|^2
    int* flags;     // [0] Arnoldi active, [1] converged, [2] total iterations, [3] columns of the cycle (jend)
    double* H;      // (m + 1) x m, row-major
    double *cs, *sn, *g, *y;
    double *h1, *h2;   // Gram-Schmidt coefficients of the two passes (m + 1 each)
    double tol;     // converged when |g_{j+1}| <= tol beta0
    int m, maxit;
};
// one Gram-Schmidt pass over the first m basis vectors (see k_gs_pass);
// partial has (m + 1) rows of nblk sums, the last |w|^2
hipError_t launch_gs_pass(hipStream_t s, const double* V, int64_t ld, int m, const double* h, double* w, int64_t n,
                          bool dots, double* partial, int nblk, const int* active);
hipError_t launch_gm_reduce(hipStream_t s, const double* partial, int nblk, int first, int cnt, double* out,
                            const int* active);
hipError_t launch_gm_begin(hipStream_t s, GmDev G, int first);
hipError_t launch_gm_step(hipStream_t s, GmDev G, int j);
hipError_t launch_gm_solve(hipStream_t s, GmDev G);
hipError_t launch_scale_dev(hipStream_t s, const double* a, const double* x, double* y, int64_t n, const int* active);
hipError_t launch_lincomb_dev(hipStream_t s, const double* V, int64_t ld, const int* jend, const double* y, double* x,
                              int64_t n);
hipError_t launch_sub(hipStream_t s, const double* b, const double* w, double* r, int64_t n);
hipError_t launch_axpy(hipStream_t s, double a, const double* x, double* y, int64_t n);
// matrix-free operator of the RANS_MATRIX_FREE build (ransMatrix.h:138-178)
hipError_t launch_mf_base(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                          int viscous, double* Fb);
hipError_t launch_mf_apply(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* dt,
                           const double* Fb, GasT<double> g, int viscous, const double* x, double* y,
                           const int* active = nullptr);

}  // namespace afx
