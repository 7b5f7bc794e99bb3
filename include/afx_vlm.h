/*
 * afx_vlm.h -- C ABI of the MI355X vortex-lattice / beam coupling path of
 * AeroFLEX (reference: src/vlm, src/structure, src/aeroelasticite).
 *
 * Conventions are those of afx_rans.h: every function returns AFX_OK (0) or
 * a negative AFX_E* code, afx_last_error() describes the last failure, host
 * arrays are caller-owned and copied, one HIP stream per handle, handles are
 * thread-compatible, not thread-safe.
 *
 * The lattice is the reference's vortex-ring model (vlm/src/panel.cpp): each
 * ring has four corner nodes in the order leading-edge j, leading-edge j+1,
 * trailing-edge j+1, trailing-edge j; rings are grouped in spanwise stations
 * (wingStation, surface.cpp), each station listing its rings from the leading
 * to the trailing edge.  The last ring of every station sheds one wake ring
 * (wingStation::generateWake, surface.cpp:35-58).
 */
#ifndef AFX_VLM_H
#define AFX_VLM_H

#include <stdint.h>

#include "afx_rans.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afx_vlm afx_vlm;
typedef struct afx_dense afx_dense;
typedef struct afx_beam afx_beam;
typedef struct afx_aeroelastic afx_aeroelastic;

/* ------------------------------------------------------------ lattice --- */

/* Trapezoidal wing, both halves (span -b/2 .. b/2), in the x-y plane with x
 * downstream.  Panels of n_chord x n_span; vortex rings on the panels'
 * quarter-chord lines (ring leading edge at the panel quarter chord, ring
 * trailing edge a quarter panel behind the panel), cosine spacing in both
 * directions when cosine != 0.  sweep_deg is the quarter-chord sweep. */
typedef struct afx_wing_spec {
    int32_t n_chord;
    int32_t n_span;          /* even: a station line lies on y = 0 */
    double aspect_ratio;
    double taper;            /* tip chord / root chord */
    double sweep_deg;
    double root_chord;
    int32_t cosine;
} afx_wing_spec;

/* Two-call pattern: with null arrays only the counts are returned.
 * xyz: 3 per node; ring_nodes: 4 per ring; station_offsets: n_stations + 1
 * (rings of station s are [off[s], off[s+1]), leading to trailing edge);
 * ref: {sref, cref (mean aerodynamic chord), span}. */
int afx_wing_lattice(const afx_wing_spec* spec, int32_t* n_nodes, int32_t* n_rings, int32_t* n_stations,
                     double* xyz, int32_t* ring_nodes, int32_t* station_offsets, double* ref);

/* ---------------------------------------------------------------- VLM --- */

/* simParam (vlm/src/input.cpp:25-52, 200-226) as used by the linear steady
 * solver: freestream vinf*(cos a, 0, sin a), lift axis freestream x unit-y. */
typedef struct afx_vlm_flow {
    double alpha_deg;
    double vinf;
    double rho;
    double sref;
    double cref;
    double wake_length;      /* length of the wake rings along the freestream */
    double core_radius;      /* Lamb-Oseen core (0: none, the reference default) */
} afx_vlm_flow;

/* model + linear::steady::initialize (vlm/src/model.cpp, solver.cpp:22-40):
 * copies the lattice to the device. */
int afx_vlm_create(int32_t n_nodes, const double* xyz, int32_t n_stations, const int32_t* station_offsets,
                   const int32_t* ring_nodes, afx_vlm** out);
int afx_vlm_destroy(afx_vlm* h);
int afx_vlm_counts(const afx_vlm* h, int32_t* n_nodes, int32_t* n_rings, int32_t* n_stations);
int afx_vlm_set_flow(afx_vlm* h, const afx_vlm_flow* flow);
/* moved nodes (model::updateGeometry, model.cpp); marks the matrix stale */
int afx_vlm_set_nodes(afx_vlm* h, const double* xyz);
int afx_vlm_get_nodes(const afx_vlm* h, double* xyz);
/* the lattice as created: 4 node ids per ring, n_stations + 1 offsets (nullable) */
int afx_vlm_get_rings(const afx_vlm* h, int32_t* ring_nodes, int32_t* station_offsets);

/* linear::steady::buildLHS + buildRHS (solver.cpp:139-254, 256-290):
 * panel geometry (panel.cpp:46-72), collocation points (panel.cpp:111-120),
 * wake rings, AIC_ij = v_j(P_i).n_i with the wake ring added to the column of
 * its trailing-edge ring, rhs_i = -Vinf.n_i.  The FP64 matrix stays on the
 * device, row-major as lhs(i, j). */
int afx_vlm_build(afx_vlm* h);
/* FP32 LU with partial pivoting of the current matrix (replaces the
 * BiCGSTAB compute of solver.cpp:41-50 together with afx_vlm_solve) */
int afx_vlm_factor(afx_vlm* h);

typedef struct afx_solve_info {
    int32_t iterations;      /* refinement steps taken (each: FP64 residual + FP32 LU solve) */
    int32_t converged;
    double backward_error;   /* |b - A x|inf / (|A|inf |x|inf + |b|inf) in FP64 */
    double contraction;      /* |r_k| / |r_k-1| of the last step */
} afx_solve_info;

/* Gamma = AIC^-1 rhs (solver.cpp:343-361): FP32 LU of the last factored
 * matrix as the preconditioner of FP64 iterative refinement on the current
 * FP64 matrix.  tol <= 0 selects sqrt(N) * eps64; warm != 0 starts from the
 * handle's previous Gamma (zero before the first solve).  gamma: N doubles
 * or null. */
int afx_vlm_solve(afx_vlm* h, double tol, int32_t max_iter, int32_t warm, double* gamma, afx_solve_info* info);

/* wingStation::computeForces (surface.cpp:164-201), linear form: Kutta-
 * Joukowski force rho (G_k - G_k-1)(Vinf x dl) of every ring at its leading-
 * edge midpoint.  ring_force: 3 per ring, station_force: 3 per station (both
 * nullable); coeffs (nullable): {CL, CY, CD_kj} = total force on the lift
 * axis, unit-y and stream axis over (q sref). */
int afx_vlm_forces(afx_vlm* h, double* ring_force, double* station_force, double* coeffs);
/* host copies for checks: matrix (N*N, row-major lhs(i,j)), rhs, gamma,
 * collocation points and normals (3 per ring each; nullable) */
int afx_vlm_get_aic(afx_vlm* h, double* aic);
int afx_vlm_get_system(afx_vlm* h, double* rhs, double* gamma, double* colloc, double* normal);

/* ------------------------------------------------ nonlinear (viscous) --- */

/* database::table (common/database/database.cpp:192-242): airfoil polars
 * (alpha in degrees, CL, CD, CMy; linear interpolation clamped at the ends,
 * mlinterp) placed at normalised span locations of the wing. */
typedef struct afx_polar {
    int32_t n_airfoils;
    const int32_t* offsets;       /* n_airfoils + 1: rows of airfoil a are [offsets[a], offsets[a+1]) */
    const double* alpha_deg;      /* increasing within each airfoil */
    const double* cl;
    const double* cd;
    const double* cmy;
    int32_t n_sections;
    const int32_t* section_airfoil;
    const double* section_span;   /* increasing, 0 at the wing's first station, 1 at its end */
} afx_polar;

typedef struct afx_nonlinear_info {
    int32_t iterations;
    int32_t converged;
    double residual;              /* sqrt(sum_s ((cl_visc - cl_inv) / cl_visc)^2) of the last iteration */
    double cl, cd, cy;            /* viscous coefficients (nonlinear::steady::computeForces) */
} afx_nonlinear_info;

/* nonlinear::steady::solve (vlm/src/solver.cpp:330-388, iterate :390-423):
 * the matrix is built and factored once at the flow's alpha; each iteration
 * solves Gamma for rhs_i = -V(alpha_s).n_i with the station's local angle
 * alpha_s, takes each station's inviscid cl, interpolates the viscous cl at
 * alpha_eff = cl_inv / CLa - alpha_s + alpha (CLa = 2 pi per radian in
 * degrees, VLM_CL_ALPHA_DEG) and moves alpha_s by relaxation (cl_visc -
 * cl_inv) / CLa, until the residual is <= tol or max_iter is passed.  The
 * handle's stations form one wing, in their given order.  station_alpha,
 * station_cl (nullable): n_stations each, final local angles (degrees) and
 * viscous section lift coefficients. */
int afx_vlm_nonlinear_solve(afx_vlm* h, const afx_polar* polar, double relaxation, double tol, int32_t max_iter,
                            afx_nonlinear_info* info, double* station_alpha, double* station_cl);

/* ------------------------------------------- dense FP32-LU / FP64 solve --- */

/* The solver behind afx_vlm_solve for any dense matrix (no reference
 * counterpart: the reference uses Eigen's BiCGSTAB on the AIC). */
int afx_dense_create(int32_t n, afx_dense** out);
int afx_dense_destroy(afx_dense* h);
int afx_dense_set(afx_dense* h, const double* a_rowmajor);
int afx_dense_factor(afx_dense* h, int32_t* info);   /* info: 0 or 1 + first zero pivot */
int afx_dense_solve(afx_dense* h, const double* b, double tol, int32_t max_iter, double* x, afx_solve_info* info);
/* factor in LAPACK getrf form: column-major n x n (unit L below, U on and
 * above the diagonal) and 0-based pivots ipiv[k] = row swapped with k */
int afx_dense_get_lu(afx_dense* h, float* lu_colmajor, int32_t* ipiv);
/* Test hook (no counterpart in the reference): raises the triangular solves'
 * abort flag, as one of their bounded waits does when it times out.  The next
 * afx_dense_solve then fails (AFX_EHIP) and discards the factorisation; the
 * next afx_dense_factor clears the flag. */
int afx_dense_raise_abort(afx_dense* h);

/* --------------------------------------------------------------- beam --- */

/* structure::MODEL with CBAR elements (structure/include/structure/
 * model.hpp:103-131, element.hpp:77-109): 12x12 Euler-Bernoulli frame
 * stiffness per element, T^T K T into global axes, SPC1 on the clamped nodes
 * (all six dofs).  elem_nodes: 2 per element; props: 7 per element
 * {EA, EIy, EIz, GJ, vx, vy, vz} with (vx, vy, vz) the orientation vector of
 * the local y axis (CBAR's V). */
int afx_beam_create(int32_t n_nodes, const double* xyz, int32_t n_elems, const int32_t* elem_nodes,
                    const double* props, int32_t n_clamped, const int32_t* clamped, afx_beam** out);
int afx_beam_destroy(afx_beam* h);
/* MODEL::get_Lin_Solve (model.hpp:225-238): K u = f, 6 dofs per node
 * (Fx, Fy, Fz, Mx, My, Mz) -> (ux, uy, uz, rx, ry, rz) */
int afx_beam_solve(afx_beam* h, const double* load, double* disp);
/* MODEL::get_NonLin_Solve (model.hpp:240-339) with the corotational CBAR
 * (element.hpp:111-370): n_steps load steps of (step / n_steps) load; each
 * solves with the stiffness re-assembled from every element's current
 * mid-node rotation (set_K_global, model.hpp:103-131), updates node rotations
 * by quaternions built from the rotation increments read as roll, pitch, yaw,
 * recomputes internal forces in each element's corotated frame, turns the
 * loads with their node's increment (follower loads, Rotate_Ext_Loads), and
 * iterates until ||du|| <= tol; once ||du|| < 100 tol the increment is scaled
 * by damping.  As in the reference the iteration count is cumulative over the
 * steps; max_iter bounds it (the reference has no bound).  disp: accumulated
 * increments, 6 per node (rotations as summed increments, like the
 * reference's Dep). */
typedef struct afx_beam_nl_info {
    int32_t iterations;
    int32_t converged;
    double residual;               /* ||du|| of the last iteration */
} afx_beam_nl_info;
int afx_beam_solve_nonlinear(afx_beam* h, const double* load, int32_t n_steps, double tol, double damping,
                             int32_t max_iter, double* disp, afx_beam_nl_info* info);

/* -------------------------------------------------------- aeroelastic --- */

/* Coupling of one afx_vlm and one afx_beam handle, following the intent of
 * aeroelasticite (aerotools.cpp:25-217): lattice nodes follow the beam through the two
 * nearest beam nodes along y (u = sum w_a (u_a + r_a x (x - x_a))); ring
 * forces go to the same two nodes as forces and lever-arm moments, which
 * conserves total force and moment. */
int afx_aeroelastic_create(afx_vlm* vlm, afx_beam* beam, afx_aeroelastic** out);
int afx_aeroelastic_destroy(afx_aeroelastic* h);

typedef struct afx_coupling_opts {
    double tol;              /* on the change of tip deflection, relative */
    int32_t max_iter;
    double relax;            /* relaxation of the beam displacement (first Aitken factor) */
    double solve_tol;        /* refinement tolerance; <= 0: 1 % of the last relative tip change,
                                clamped to [sqrt(N) eps64, 1e-6] */
    int32_t solve_max_iter;
    double refactor_above;   /* refactor when refinement contracts worse than this */
    int32_t warm_start;      /* start from the previous solve's displacement */
    int32_t aitken;          /* Aitken dynamic relaxation, starting from relax (default 1) */
} afx_coupling_opts;

typedef struct afx_coupling_info {
    int32_t iterations;
    int32_t converged;
    int32_t refactors;
    int32_t refine_steps;    /* total refinement steps over the iterations */
    double tip_deflection;   /* z displacement of the beam node with the largest y */
    double cl;
    double backward_error;   /* of the last VLM solve */
} afx_coupling_info;

void afx_default_coupling_opts(afx_coupling_opts* o);
/* one load case at flow->alpha_deg; the VLM flow must have been set */
int afx_aeroelastic_solve(afx_aeroelastic* h, const afx_coupling_opts* opts, afx_coupling_info* info);
/* beam displacement of the last solve (6 per beam node) */
int afx_aeroelastic_get_disp(afx_aeroelastic* h, double* disp);

#ifdef __cplusplus
}
#endif

#endif
