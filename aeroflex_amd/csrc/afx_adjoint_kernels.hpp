// Launcher of the state adjoint of the FP64 residual (afx_adjoint.hip).
#pragma once
#include "afx_kernels.hpp"
#include "afx_adjoint.hpp"

namespace afx {

// Tangent of the FP64 residual (forward mode, Dual1 instantiation of the fused
// tile kernel).  launch_to_dual: Dual1 copies of geometry arrays (zero
// derivative); launch_dual_state: (q, v) -> q + eps v over nv planes, ghost
// cells with a zero derivative; launch_dual_parts: value and derivative
// planes of a Dual1 array (val may be null).
hipError_t launch_to_dual(hipStream_t s, const double* x, int64_t n, Dual1* out);
hipError_t launch_dual_state(hipStream_t s, const double* q, const double* v, int32_t nT, int32_t nC, int nv,
                             Dual1* out);
hipError_t launch_dual_parts(hipStream_t s, const Dual1* x, int64_t n, double* val, double* der);
size_t tile_lds_bytes_tangent(int32_t lmax, int32_t gmax, int32_t ftmax, int nv);
hipError_t launch_residual_tile_tangent(hipStream_t s, TilesT<Dual1> S, const Dual1* q, int32_t nT, int32_t nC,
                                        const Dual1* area, const Dual1* lsq, const Dual1* xc, const Dual1* yc,
                                        const Dual1* wd, GasT<double> g64, int gg, int so, double limiter_k,
                                        int viscous, Dual1* R);

// work arrays of one adjoint evaluation ([4][nT] / [4][nC] / [4][nF] planes;
// five planes each with the Spalart-Allmaras model)
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
    const int32_t* bface;   // [nbf] the boundary faces (device index, ascending)
    int32_t nbf;
    // laminar residual (null otherwise): per interior face, half the adjoint of
    // the averaged face gradient (owed to both cells' gradients) and the
    // adjoint of the face's difference quotient divided by |x_0 + x_1| (owed to
    // the owner's state, with the opposite sign to the neighbour's)
    double* vgx;     // [4][nF]
    double* vgy;
    double* vdq;
};

// work arrays of the adjoint with respect to the mesh nodes
struct GeoBufs {
    double* fbar;            // [8][nF] adjoints of nx, ny, len, gw, d0x, d0y, d1x, d1y (d = x_f - x_cell)
    double* abar;            // [nC] adjoint of the cell area
    const int32_t* fnode;    // [2][nF] end nodes of each face (device face order)
    const int32_t* cnode;    // [4][nC] nodes of each real cell (device cell order); slot 3 = -1 for triangles
    const double* nodex;     // [nN] node coordinates
    const double* nodey;
    double* fend;            // [4][nF] (x, y) adjoints of face end a, then end b
    double* fcen;            // [4][nF] (x, y) adjoints of the owner, then the neighbour centroid
    double* cvert;           // [8][nC] (x, y) adjoints of each vertex slot
    const int32_t* n_off;    // [nN + 1] incidences per node (CSR)
    const int32_t* n_ent;    // 2 f + end (face ends) or -(1 + 4 c + slot) (cell vertices)
    int32_t nN;
    double* xbar;            // [nN][2] output: (x, y) adjoint per node
};

// lam[k*nT + i] (real cells) = ((dR/dq)^T w)[i, k] for the inviscid residual
// with Green-Gauss gradients; w in the same layout (ghost rows ignored).
// With G non-null also G->xbar = (dR/dX)^T w over the node coordinates X.
hipError_t launch_adjoint_state(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, const double* w,
                                GasT<double> g, int second_order, double limiter_k, AdjBufs A, double* lam,
                                const GeoBufs* G = nullptr);
// Green-Gauss gradients of the five variables of the Spalart-Allmaras model
// (pass 1 of its adjoint without the limiters): gx, gy 5 planes, real cells
hipError_t launch_sa_gradients(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                              double* gx, double* gy);
// state adjoint of the Spalart-Allmaras residual (five planes per array of A,
// vgx / vgy / vdq required; wd: wall distance per device cell)
hipError_t launch_adjoint_state_sa(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q,
                                   const double* w, const double* wd, GasT<double> g, int second_order,
                                   double limiter_k, AdjBufs A, double* lam);

// The explicit stage residual of explicitSolver::solve (solver.h:802-828)
// takes its gradients from the step's start state qg and everything else
// from the stage state q (A3).  lam = (dR/dq)^T w, lam_g = (dR/dqg)^T w; with
// ghost non-null the adjoints of the farfield ghost states (boundary data)
// are written into its ghost rows [k*nT + ghost cell].
hipError_t launch_adjoint_stage(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* qg, const double* q,
                                const double* w, GasT<double> g, int second_order, double limiter_k, AdjBufs A,
                                double* lam, double* lam_g, double* ghost, const GeoBufs* G = nullptr);

// launch_adjoint_stage with the metric adjoints of the stage left in G.fbar /
// G.abar (overwritten) and not converted to node adjoints
hipError_t launch_adjoint_stage_metrics(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* qg,
                                        const double* q, const double* w, GasT<double> g, int second_order,
                                        double limiter_k, AdjBufs A, double* lam, double* lam_g, const GeoBufs& G);
// facc += G.fbar (8 x nF), aacc += G.abar (nC)
hipError_t launch_acc_metrics(hipStream_t s, const GeoBufs& G, int32_t nF, int32_t nC, double* facc, double* aacc);
// facc += metric adjoints (n, len) of the local time step for dtbar
hipError_t launch_dt_vjp_geo(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                             const double* dt, const double* dtbar, double cfl, double* facc);
// G.xbar from the metric adjoints in G.fbar / G.abar
hipError_t launch_geo_to_nodes(hipStream_t s, CellsT<double> C, FacesT<double> F, const GeoBufs& G);

// reverse of one explicit stage update qk_s+1 = q + W_s dt coef (real cells)
hipError_t launch_rev_stage(hipStream_t s, const double* gk, const double* W, const double* dt, const double* area,
                            double coef, int32_t nC, int32_t nT, double* wseed, double* dtbar, double* qbar);
// y = a x (overwrite) or y += a x, over the real-cell rows of the 4 planes
hipError_t launch_axpy4(hipStream_t s, const double* x, double a, int32_t nC, int32_t nT, int overwrite, double* y);
// acc[k][j] += ghost[k*nT + cells[j]]
hipError_t launch_ghost_acc(hipStream_t s, const double* ghost, const int32_t* cells, int32_t n, int32_t nT,
                            double* acc);

// dt = cfl A / sum_f max(eig_o, eig_n) len (solver.h:308-356): qbar += (d dt / d q)^T dtbar
hipError_t launch_dt_vjp(hipStream_t s, CellsT<double> C, FacesT<double> F, const double* q, GasT<double> g,
                         const double* dt, const double* dtbar, double cfl, double* ebar, double* qbar);

}  // namespace afx
