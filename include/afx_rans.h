/*
 * afx_rans.h -- C ABI of the MI355X-native residual of AeroFLEX's
 * unstructured finite-volume RANS solver (reference: src/rans/include/rans).
 *
 * Conventions
 *   - Every function returns int: AFX_OK (0) or a negative AFX_E* code;
 *     afx_last_error() returns a thread-local description.  No C++ exception
 *     crosses this boundary (the reference throws std::runtime_error /
 *     std::invalid_argument / std::out_of_range at the corresponding places).
 *   - Host arrays are caller-owned and copied; all vectors exchanged with the
 *     caller are in the reference's order and layout: state is the
 *     Eigen::VectorXd layout q(4*i+k) over all cells (real cells, then ghost
 *     cells in boundary-segment order, solver.h:54,184), per-face arrays are in
 *     face-creation order (mesh.h:317-341).
 *   - One HIP stream per solver handle; calls are synchronous unless stated.
 *     Handles are thread-compatible, not thread-safe.
 */
#ifndef AFX_RANS_H
#define AFX_RANS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFX_OK 0
#define AFX_EARG (-1)        /* invalid argument / unknown name            */
#define AFX_EHIP (-2)        /* HIP runtime error                          */
#define AFX_EIO (-3)         /* file / parse error                         */
#define AFX_ESTATE (-4)      /* call out of order (e.g. no BCs set)        */
#define AFX_EUNSUP (-5)      /* unsupported option combination            */

/* boundary-condition types (solver.h:216-230 strings) */
#define AFX_BC_OTHER 0       /* any other name in the reference (interior flux with the
                                ghost); rejected by afx_rans_set_bcs */
#define AFX_BC_FARFIELD 1    /* "farfield"  -> farfield_flux  (physics.h:409-530) */
#define AFX_BC_SLIP 2        /* "slip-wall" -> slip_wall_flux (physics.h:279-339) */
#define AFX_BC_WALL 3        /* "wall"      -> wall_flux      (physics.h:344-405) */

/* residual precision forms */
#define AFX_FP64 0           /* plain form: FP64 everywhere, reference operation order */
#define AFX_FP32 1           /* fluxes/gradients/limiters FP32, compensated FP32 cell sums */

typedef struct afx_mesh afx_mesh;
typedef struct afx_rans afx_rans;
typedef struct afx_batch afx_batch;
typedef struct afx_mapper afx_mapper;

/* core.h:27-46 (rans::gas; cp is fixed to 1 as in the reference) */
typedef struct afx_gas {
    double R;        /* 0.71428571428 */
    double mu_L;     /* 1e-5          */
    double Pr_L;     /* 0.72          */
    double Pr_T;     /* 0.9           */
    double gamma;    /* 1.4           */
} afx_gas;

/* solver options: core.h:166-232 (rans::Settings) as used by
 * multigrid.h:72-76 (solver construction + set_second_order /
 * set_gradient_scheme / set_limiter_k) */
typedef struct afx_opts {
    int32_t viscosity;     /* 0 "inviscid", 1 "laminar", 2 "spalart-allmaras": SA-noft2
                              with rho nu~ as fifth variable (the reference's SA
                              branch is empty, physics.h:259-261); residual only,
                              fused tile kernels */
    int32_t second_order;  /* Settings::second_order (default 1) */
    int32_t gradient;      /* 0 "least-squares", 1 "green-gauss" (default) */
    double limiter_k;      /* Settings::limiter_k (default 5) */
    int32_t renumber;      /* 1: tiled RCM cell order on the device (default), 0: file order */
    int32_t tile_cells;    /* cells per tile of the fused residual kernel (0: automatic, the
                              largest of 256, 224, ... whose FP64 tile fits 64 KiB of LDS) */
    int32_t fused;         /* 1: fused tile kernel for the residual (default), 0: three-kernel path */
    int32_t limiter;       /* 0: the reference's default limiter (solver.h:575-583, default);
                              1: its RANS_MICHALAK_LIMITER build (solver.h:557-573, physics.h:583-592).
                              With 1 the FP32 residual runs the three-kernel path; the adjoint,
                              dCL entries and batched stepping return AFX_EUNSUP */
    int32_t reserved[4];
} afx_opts;

const char* afx_last_error(void);
void afx_default_gas(afx_gas* g);
void afx_default_opts(afx_opts* o);

/* ------------------------------------------------------------------ mesh --
 * rans::mesh (mesh.h:198-273).  Host object in reference order. */

/* mesh::read_file (mesh.h:834-884): gmsh MSH 4.1 ASCII */
int afx_mesh_read_msh(const char* path, afx_mesh** out);

/* Build from elements, i.e. what mesh::read_nodes/read_boundaries/read_elements
 * (mesh.h:524-738) leave before add_cell_edges: node coordinates; cells with
 * 3 or 4 node indices (cell_nodes: 4 per cell, 4th ignored for triangles);
 * boundary segments (2 node indices each) with their physical key
 * (tag + 1000*dim); physical names as parallel arrays of key and name. */
int afx_mesh_from_elements(int64_t n_nodes, const double* x, const double* y,
                           int64_t n_cells, const int32_t* cell_nodes, const int32_t* cell_nverts,
                           int64_t n_bsegs, const int32_t* bseg_nodes, const int32_t* bseg_keys,
                           int32_t n_names, const int32_t* name_keys, const char* const* names,
                           afx_mesh** out);

/* Analytic NACA 4-digit O-grid: ni points around, nj from wall to a circle of
 * r_far chords, quads split into triangles (physical names "farfield",
 * "wall"). */
int afx_mesh_gen_naca_ogrid(int32_t ni, int32_t nj, double r_far, double thickness,
                            double first_h, afx_mesh** out);
/* Seeded mixed tri/quad mesh of jittered points around a NACA 4-digit section. */
int afx_mesh_gen_jittered(int32_t ni, int32_t nj, double r_far, double thickness, double first_h,
                          double jitter, double quad_frac, uint64_t seed, afx_mesh** out);
int afx_mesh_write_msh(const afx_mesh* m, const char* path);
/* counts[5] = {n_nodes, n_real_cells, n_cells (real + ghost), n_faces, n_boundary_faces} */
int afx_mesh_counts(const afx_mesh* m, int64_t counts[5]);
/* Copy one reference array: nodesX nodesY edgesNodes edgesCells edgesLengths
 * edgesNormalsX edgesNormalsY edgesCentersX edgesCentersY cellsNodes
 * cellsEdges cellsIsTriangle cellsAreas cellsCentersX cellsCentersY
 * boundaryEdges boundaryEdgesIntTag.  Doubles are f64, indices u32 (i32 for
 * cellsIsTriangle).  out may be NULL to query the length. */
int afx_mesh_get(const afx_mesh* m, const char* field, void* out, int64_t* n_out);
/* boundaryEdgesPhysicals[b] (mesh.h:225) */
int afx_mesh_boundary_name(const afx_mesh* m, int64_t b, char* buf, int32_t buflen);
/* Reverse Cuthill-McKee order of the real cells (perm[new] = old). */
int afx_mesh_rcm(const afx_mesh* m, int32_t* perm);
/* Tile statistics for a tile size: stats[8] = {n_tiles, max local cells,
 * max gradient cells, max local faces, max faces touching a tile,
 * total local cells, total gradient cells, total local faces}. */
int afx_mesh_tile_stats(const afx_mesh* m, int32_t tile_cells, int64_t stats[8]);
void afx_mesh_destroy(afx_mesh* m);

/* ---------------------------------------------------------------- solver --
 * rans::solver / explicitSolver / implicitSolver (solver.h:40-1213). */

/* solver(const mesh&, const gas&, std::string viscosity_model) (solver.h:108-110)
 * + set_second_order / set_gradient_scheme / set_limiter_k (multigrid.h:73-76) */
int afx_rans_create(const afx_mesh* m, const afx_gas* g, const afx_opts* o, afx_rans** out);
/* solver::set_bcs (solver.h:200-247): per physical name its type and, for
 * farfield, boundary_variables {mach, angle[rad], T, p} (core.h:61-84).
 * Names absent from the list make the call fail (reference: std::out_of_range
 * from bcs.at(), solver.h:220). */
int afx_rans_set_bcs(afx_rans* h, int32_t n, const char* const* names, const int32_t* types,
                     const double* farfield_vars /* 4*n: mach, angle, T, p */);
/* set_second_order / set_gradient_scheme / set_limiter_k (solver.h:135,149-152,162)
 * and the limiter build; viscosity and renumber are fixed at creation and
 * ignored here */
int afx_rans_set_options(afx_rans* h, const afx_opts* o);
int afx_rans_init(afx_rans* h);                 /* solver::init (solver.h:615-631) */
int afx_rans_refill_bcs(afx_rans* h);           /* solver::refill_bcs (solver.h:259-274) */
int afx_rans_bcs_from_internal(afx_rans* h);    /* solver::bcs_from_internal (solver.h:276-287) */
int afx_rans_set_walls_from_internal(afx_rans* h); /* solver.h:289-305 on q */
/* get_q() (solver.h:255-257): AoS q(4*i+k), 4*n_cells doubles */
int afx_rans_set_state(afx_rans* h, const double* q);
int afx_rans_get_state(afx_rans* h, double* q);
int afx_rans_set_cfl(afx_rans* h, double cfl);  /* solver::set_cfl (solver.h:250-252) */
/* solver::calc_dt (solver.h:308-356); dt_out (n_cells, nullable) */
int afx_rans_calc_dt(afx_rans* h, double* dt_out);
/* set_walls_from_internal + calc_gradients + calc_limiters on q
 * (solver.h:1083-1087); outputs AoS 4*n_cells each, nullable */
int afx_rans_gradients(afx_rans* h, int32_t precision, double* gx, double* gy, double* limiters);
/* implicitSolver::fillRhoRHS (solver.h:1079-1152) without the calc_dt it
 * performs first (call afx_rans_calc_dt for that): R = RhoVector (AoS,
 * nullable), norm2 = ||R||_2 (nullable).  With both outputs NULL the call
 * only enqueues the kernels on the handle's stream and returns without
 * synchronising (inputs and outputs stay resident in HBM). */
int afx_rans_residual(afx_rans* h, int32_t precision, double* R, double* norm2);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory or synchronise with the host (code under HIP graph capture, serving
 * loops): afx_rans_residual(h, precision, NULL, norm2) with ||R||_2 written to
 * the device double *norm2_dev, enqueued on the handle's stream without any
 * allocation or host synchronisation (R stays in the handle).  The partial
 * sums of the norm live in one static 16 KiB array per GPU, shared by every
 * call on that GPU: calls that overlap on one GPU (different handles or
 * streams) race.  Limit: 1024 partial sums per reduction (AFX_EUNSUP beyond).
 * The result equals afx_rans_residual's norm2. */
int afx_rans_residual_norm_dev(afx_rans* h, int32_t precision, double* norm2_dev);
/* explicitSolver::calc_residual (solver.h:745-799) for a stage state qk
 * (AoS), with gradients from the current q (A3); qW out (AoS, nullable) */
int afx_rans_explicit_residual(afx_rans* h, const double* qk, double* qW);
/* explicitSolver::solve (solver.h:802-828): 3-stage pseudo-time step on the
 * device, err = ||qW||_2 of the last stage */
int afx_rans_explicit_step(afx_rans* h, double relaxation, double* err);
/* ---- implicit pseudo-time step (implicitSolver, solver.h:852-1213) ----
 * The linear solver replacing Eigen's GMRES + IncompleteLUT (solver.h:886,
 * 907-910): restarted GMRES(restart) with a block right preconditioner
 * (4x4 blocks; 5x5 with the Spalart-Allmaras model),
 * FP64, stopping at ||b - A x|| <= tol ||b|| or after maxit iterations
 * (defaults 30, 1e-2, 500 as the reference sets them; restart at most 32).
 * The Arnoldi state stays on the device; the host synchronises once per
 * restart cycle. */
int afx_rans_set_linear_solver(afx_rans* h, int32_t restart, double tol, int32_t maxit);
/* The preconditioner standing in for IncompleteLUT (solver.h:886, fill 2,
 * drop tolerance 1e-3): kind 0 block Jacobi (the inverse 4x4 diagonal
 * blocks), kind 1 (default) block DILU in a greedy colour order of the real
 * cells, M = (Dt + L) Dt^-1 (Dt + U) with Dt_i = D_i - sum over earlier
 * neighbours j of A_ij Dt_j^-1 A_ji; one parallel sweep per colour.  The next
 * implicit step needs a new afx_rans_implicit_fill. */
int afx_rans_set_preconditioner(afx_rans* h, int32_t kind);
/* implicitSolver::fill -> fillRhoLHS (solver.h:973-1075): calc_dt, gradients
 * when viscous, FD flux Jacobian blocks (physics.h:533-578), A/dt diagonal.
 * With the Spalart-Allmaras model (no counterpart: the reference's SA branch
 * is empty, physics.h:259-261) the same rule gives 5x5 blocks of the
 * five-variable face flux, with the cell gradients frozen and the face
 * gradients' directional parts following the perturbed states; the diagonal
 * blocks add the source row -A dS/dv (its diagonal entry only where >= 0);
 * dt is the model's time step (with the viscous radius).  It needs
 * Green-Gauss gradients and the fused residual; no matrix-free form. */
int afx_rans_implicit_fill(afx_rans* h);
/* implicitSolver::compute (solver.h:1160-1167): factorises the
 * preconditioner; *status = 0 on success, -1 when a diagonal block is
 * singular (the reference's info() != Success) */
int afx_rans_implicit_compute(afx_rans* h, int32_t* status);
/* implicitSolver::solve (solver.h:1170-1213): RHS, linear solve,
 * q += relaxation*dq, repeated rhs_iterations times with the frozen matrix
 * and the same early exits; *err = final ||R||_2, or -1 when the linear
 * solver did not converge (the reference's sentinel).  precision selects
 * the residual form (AFX_FP64 or AFX_FP32); the update and solve are FP64. */
int afx_rans_implicit_solve(afx_rans* h, double relaxation, double tol, int32_t rhs_iterations, int32_t precision,
                            double* err);
/* y = LHS x with the matrix of the last fill (AoS, 4*n_cells); for tests.
 * Not with the Spalart-Allmaras model (use afx_rans_lhs_apply_nu). */
int afx_rans_lhs_apply(afx_rans* h, const double* x, double* y);
/* The same with the rho nu~ rows of the Spalart-Allmaras model: x_nu, y_nu
 * (n_cells each) are required with that model and must be null otherwise. */
int afx_rans_lhs_apply_nu(afx_rans* h, const double* x, const double* x_nu, double* y, double* y_nu);
/* y = M^-1 x with the preconditioner of the last fill (AoS, 4*n_cells; ghost
 * rows identity); for tests.  Not with the Spalart-Allmaras model. */
int afx_rans_precond_apply(afx_rans* h, const double* x, double* y);
/* The same with the rho nu~ rows (as afx_rans_lhs_apply_nu). */
int afx_rans_precond_apply_nu(afx_rans* h, const double* x, const double* x_nu, double* y, double* y_nu);
/* colour of each real cell (reference order, n_real entries, nullable) in the
 * DILU sweeps and the number of colours; for tests */
int afx_rans_cell_colours(afx_rans* h, int32_t* colours, int32_t* ncolours);
/* The RANS_MATRIX_FREE build of implicitSolver (solver.h:863-868, 911-914,
 * 920-925): enable = 1 replaces the assembled matrix by the matrix-free
 * operator RansMatrix (ransMatrix.h:138-178) and the preconditioned GMRES by
 * unpreconditioned GMRES with that build's tolerance 1e-3 and 1000 iterations
 * (afx_rans_set_linear_solver afterwards overrides them); enable = 0 returns
 * to the assembled form (1e-2, 500).  Either way the next implicit step needs
 * a new afx_rans_implicit_fill.  As in that build, the operator reads the
 * current q and the dt of the latest residual, so it follows the state
 * through the rhs_iterations of afx_rans_implicit_solve. */
int afx_rans_set_matrix_free(afx_rans* h, int32_t enable);
/* y = RansMatrix x (ransMatrix.h:150-178) at the current q and dt (AoS,
 * 4*n_cells): x A/dt on real rows and x on ghost rows, plus per face
 * (F(q_L + 1e-6 x_L, q_R) - F(q_L, q_R))/1e-6 on the owner row and, for
 * interior faces, minus (F(q_L, q_R + 1e-6 x_R) - F)/1e-6 on the neighbour
 * row, fluxes at cell values without gradients or face lengths.  Needs
 * matrix-free mode and a fill. */
int afx_rans_matrix_free_apply(afx_rans* h, const double* x, double* y);
/* iterations of the last linear solves (sum over the last implicit_solve) */
int afx_rans_linear_iterations(afx_rans* h, int64_t* iters);

/* solver::get_uniform_residual (solver.h:636-690), with the stale-qW rule */
int afx_rans_uniform_residual(afx_rans* h, double* err);
/* get_wall_profile (post.h:301-387): CL, CD, CM on the named patch */
int afx_rans_wall_forces(afx_rans* h, const char* patch, double* cl, double* cd, double* cm);
/* CpProfile::calc_chord_coords / calc_cp (post.h:200-290): per boundary face
 * of the named patch, in the reference's boundary order, the face centre
 * (x, y) and cp = 2 / (gamma M_inf^2) (p / p_inf - 1) of the owner cell's
 * state.  *n: in, the arrays' capacity; out, the patch's face count.  With x,
 * y and cp all null only *n is returned. */
int afx_rans_cp_profile(afx_rans* h, const char* patch, int64_t* n, double* x, double* y, double* cp);
/* mesh::compute_wall_dist (mesh.h:794-830) on the device; wall_dist out
 * (n_cells, nullable).  Requires BCs. */
int afx_rans_wall_dist(afx_rans* h, double* wall_dist);
/* The wall distance the Spalart-Allmaras model uses (no counterpart: the
 * reference's SA branch is empty): from each cell centroid to the nearest
 * point of the wall / slip-wall face segments, 1 without such faces.  The
 * face centres of mesh::compute_wall_dist overestimate it up to a/2 times on
 * wall cells of aspect ratio a.  dist out (n_cells, nullable); SA only. */
int afx_rans_sa_wall_dist(afx_rans* h, double* dist);
/* counts[5] as afx_mesh_counts */
int afx_rans_counts(afx_rans* h, int64_t counts[5]);
/* save (post.h:58-180): the reference's ASCII VTU of the current state (real
 * cells: wall distance, Mach, density, pressure, temperature, velocity),
 * byte for byte, formatted on the host's cores.  Computes the wall distance
 * first when it is not yet known, as multigrid<T>::run does (multigrid.h:71).
 * Requires BCs. */
int afx_rans_save_vtu(afx_rans* h, const char* path);

/* ---------------------------------------------------- Spalart-Allmaras --
 * Replace the reference's unused SA members nu / nuW (solver.h:55,59,187-188):
 * rho nu~ per cell (n_cells, reference order, ghosts included).  Valid only
 * for a solver created with viscosity 2 (AFX_ESTATE otherwise).
 * set_turbulence: freestream rho nu~ = nu_ratio * mu_L used by init and
 * refill_bcs (default 3).  nu_residual: the rho nu~ row of the last
 * afx_rans_residual (its norm2 covers all five rows). */
int afx_rans_set_turbulence(afx_rans* h, double nu_ratio);
int afx_rans_set_nu(afx_rans* h, const double* nu);
int afx_rans_get_nu(afx_rans* h, double* nu);
int afx_rans_nu_residual(afx_rans* h, double* Rnu);

/* ------------------------------------------- conditions, batching, tuning --
 * No counterpart in the reference, whose Rans facade runs one condition at a
 * time (rans.h:78-106 sweeps alpha by re-solving).
 * create_like: a second flow condition on the mesh of `base` (same options,
 * tiles and cell order; no re-tiling), with its own gas (e.g. mu_L from Re);
 * set its BCs / state as for afx_rans_create.
 * time_residual: average duration (ms, HIP events on the handle's stream) of
 * `reps` back-to-back residual evaluations - the measurement the tile
 * autotuner compares across tile sizes. */
int afx_rans_create_like(const afx_rans* base, const afx_gas* g, afx_rans** out);
/* Adjoint of the residual (no counterpart in the reference): lambda =
 * (dR/dq)^T w for R = fillRhoRHS (solver.h:1079-1152) at the current state,
 * for every viscosity model with Green-Gauss gradients, first or second order
 * (the limiter differentiated on its active branch; the laminar terms of
 * physics.h:230-258 at the face gradient of solver.h:359-398; with the
 * Spalart-Allmaras model the five-variable face flux and the cell source).
 * w, lambda: reference AoS layout 4 * n_cells; ghost rows of w are ignored
 * and of lambda zero.  With the Spalart-Allmaras model this is
 * afx_rans_adjoint_nu with a zero rho nu~ seed.
 * time_adjoint: average ms of one adjoint evaluation (HIP events). */
int afx_rans_adjoint(afx_rans* h, const double* w, double* lambda);
/* The same with the rho nu~ row of the Spalart-Allmaras model: seed w_nu and
 * output lambda_nu (n_cells each, reference order, nullable; only with that
 * model). */
int afx_rans_adjoint_nu(afx_rans* h, const double* w, const double* w_nu, double* lambda, double* lambda_nu);
/* Tangent of the residual (no counterpart in the reference): jv = (dR/dq) v
 * for R = fillRhoRHS (solver.h:1079-1152) at the current state, for every
 * model (inviscid, laminar, Spalart-Allmaras), gradient scheme and order, the
 * limiter differentiated on its active branch.  v, jv: reference AoS layout
 * 4 * n_cells (ghost rows of v are ignored: ghost states are boundary data;
 * ghost rows of jv are zero); with the Spalart-Allmaras model v_nu / jv_nu
 * (n_cells each, nullable) carry the rho nu~ direction and row.  Forward-mode
 * instantiation of the fused FP64 tile kernel; needs the fused kernel and
 * twice its LDS per workgroup (AFX_EUNSUP otherwise). */
int afx_rans_tangent(afx_rans* h, const double* v, const double* v_nu, double* jv, double* jv_nu);
/* Average ms of one tangent evaluation on the device (HIP events; the last
 * direction, no host transfers). */
int afx_rans_time_tangent(afx_rans* h, int32_t reps, double* ms);
/* The same evaluation plus the adjoint with respect to the mesh nodes:
 * xbar[2 i + d] = ((dR/dX)^T w) for coordinate d of node i (reference node
 * order, mesh.h nodesX / nodesY), through the metrics of mesh.h:378-453 and
 * 744-787 (normals, lengths, midpoints, centroids, areas, ghost centroids)
 * and the Green-Gauss weight of solver.h:436-444.  lambda is nullable.
 * Inviscid residual (AFX_EUNSUP for the laminar and SA models). */
int afx_rans_adjoint_nodes(afx_rans* h, const double* w, double* lambda, double* xbar);
/* dCL/dalpha through pseudo-time (no counterpart in the reference): runs
 * `steps` explicit steps (explicitSolver::solve, solver.h:802-828, with the
 * handle's CFL) from the current state, which becomes the result as after
 * the same number of afx_rans_explicit_step calls, returns CL of the patch
 * (get_wall_profile, post.h:301-387) and its derivative with respect to the
 * farfield angle in radians (boundary_variables::angle, core.h:61-84), the
 * starting state held fixed.  Reverse accumulation through every stage
 * (A3 gradients, local time step), states stored every `checkpoint_every`
 * steps and recomputed per segment.  Inviscid, Green-Gauss, four variables. */
int afx_rans_dcl_dalpha(afx_rans* h, const char* patch, int32_t steps, int32_t checkpoint_every, double relaxation,
                        double* cl, double* dcl_dalpha);
/* dCL/dX through pseudo-time (no counterpart in the reference): as
 * afx_rans_dcl_dalpha, the derivative of CL after `steps` explicit steps
 * with respect to the mesh-node coordinates, the starting state held fixed:
 * xbar[2 i + d] for coordinate d of node i (reference node order).  It
 * covers every metric of the residual (as afx_rans_adjoint_nodes), the
 * local time step's face lengths and normals, and CL's own wall normals,
 * lengths and chord extent (post.h:301-387); where several wall faces
 * share the extreme midpoint x (the two surfaces of a symmetric section)
 * the extent's derivative is split evenly among them. */
int afx_rans_dcl_dnodes(afx_rans* h, const char* patch, int32_t steps, int32_t checkpoint_every, double relaxation,
                        double* cl, double* xbar);
int afx_rans_time_adjoint(afx_rans* h, int32_t reps, double* ms);
int afx_rans_time_residual(afx_rans* h, int32_t precision, int32_t reps, double* ms);
/* Batched explicit pseudo-time stepping of many solvers (conditions and/or
 * meshes of any size; all with the same number of variables): every step is
 * the three-stage scheme of explicitSolver::solve (solver.h:802-828) with
 * each member's CFL (afx_rans_set_cfl) and the given relaxation; the FP32
 * fused residuals of ALL members run in one kernel launch per stage and the
 * FP64 stage updates in one more.  run: `steps` pseudo-time steps, replaying
 * a hipGraph of graph_steps captured steps (0: plain launches).  Members'
 * FP64 states (afx_rans_get_state / get_nu) hold the result; their ghost
 * rows are refreshed by the next residual.  residual_norms: one residual per
 * member in the given precision, its 2-norm in norms[n]. */
int afx_batch_create(int32_t n, afx_rans* const* members, double relaxation, afx_batch** out);
int afx_batch_run(afx_batch* b, int32_t steps, int32_t graph_steps);
int afx_batch_residual_norms(afx_batch* b, int32_t precision, double* norms);
int afx_batch_stream(afx_batch* b, void** stream);

/* ------------------------------------------------ multigrid prolongation --
 * multigrid<T>::gen_mapper (multigrid.h:96-163): the sparse map from a coarse
 * solver's state to a fine one's.  Fine cell i (ghosts included) takes every
 * coarse cell j (ghosts included) whose centroid distance satisfies
 * d2 < 2 A_j, weighted s_ij / sum_j s_ij with s_ij = 1 / max(0.1 sqrt(A_j),
 * sqrt(d2)); a fine cell that takes none maps to zero, as in the reference.
 * Built on the device from the two handles' meshes (the handles must outlive
 * the mapper).  get: the map in reference cell order as CSR, row_ptr
 * (n_fine_cells + 1), cols (coarse cell), weights; each output nullable, nnz
 * first to size them.  prolongate: q_fine = M q_coarse for every variable the
 * handles carry (multigrid.h:334); the caller runs bcs_from_internal on the
 * coarse solver before and refill_bcs on the fine one after, as
 * multigrid<T>::run does (multigrid.h:330-336). */
int afx_mapper_create(const afx_rans* coarse, const afx_rans* fine, afx_mapper** out);
int afx_mapper_destroy(afx_mapper* m);
int afx_mapper_get(const afx_mapper* m, int64_t* nnz, int64_t* row_ptr, int32_t* cols, double* weights);
int afx_mapper_prolongate(const afx_mapper* m, afx_rans* coarse, afx_rans* fine);
int afx_batch_destroy(afx_batch* b);
/* The handle's hipStream_t (as void*), for event timing by the caller. */
int afx_rans_stream(afx_rans* h, void** stream);
/* Enqueue n back-to-back launches of one residual stage without
 * synchronising (stage: 0 gradients+limiters, 1 face flux, 2 cell gather,
 * 3 whole residual).  For timing the individual kernels. */
int afx_rans_launch_stage(afx_rans* h, int32_t precision, int32_t stage, int32_t n);
int afx_rans_synchronize(afx_rans* h);
int afx_rans_destroy(afx_rans* h);

#ifdef __cplusplus
}
#endif
#endif /* AFX_RANS_H */
