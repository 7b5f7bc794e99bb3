// C++ interface of the MI355X residual, shaped like the reference's
// src/rans/include/rans headers so that a caller of rans::implicitSolver /
// rans::explicitSolver can switch to the device solver:
//
//   reference                                   here
//   rans::gas (core.h:27-46)                    rans::gpu::gas
//   rans::boundary_variables (core.h:61-84)     rans::gpu::boundary_variables
//   rans::boundary_condition (core.h:112-115)   rans::gpu::boundary_condition
//   rans::mesh(filename) (mesh.h:198-273)       rans::gpu::mesh
//   rans::implicitSolver (solver.h:852-970)     rans::gpu::implicitSolver
//   rans::explicitSolver (solver.h:721-742)     rans::gpu::explicitSolver
//   multigrid<T>::run_solver (multigrid.h:183-293) rans::gpu::run_solver
//   multigrid<T>::gen_mapper (multigrid.h:96-163)  rans::gpu::mapper
//   multigrid<T>::run (multigrid.h:296-363)        rans::gpu::multigrid_run
//   get_wall_profile (post.h:301-387)           rans::gpu::get_wall_profile
//   save(filename, solver) (post.h:58-180)      rans::gpu::save
//   CpProfile (post.h:182-290)                  rans::gpu::CpProfile
//   database::airfoil (database.hpp:17-52)      rans::gpu::airfoil
//   Rans::compute_alphas (rans.h:53-57)         rans::gpu::compute_alphas
//   Rans::run_airfoil / solve_airfoil           rans::gpu::run_airfoil / solve_airfoil
//     (rans.h:78-123)
//
// Every call goes through the C ABI in include/afx_rans.h; errors come back
// as the exceptions the reference throws at the same places
// (std::runtime_error for I/O, std::invalid_argument for bad meshes,
// std::out_of_range for unknown boundary names).  Vectors use the reference's
// Eigen layout q(4*i+k) in a std::vector<double>.
#pragma once

#include <afx_rans.h>

#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rans {
namespace gpu {

inline void check(int rc) {
    if (rc == AFX_OK) return;
    std::string msg = afx_last_error();
    if (rc == AFX_EIO) throw std::runtime_error(msg);
    if (rc == AFX_EARG && msg.find("no boundary condition") != std::string::npos) throw std::out_of_range(msg);
    if (rc == AFX_EARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

struct gas {
    double R = 0.71428571428, mu_L = 1e-5, Pr_L = 0.72, Pr_T = 0.9, cp = 1., gamma = 1.4;
    gas(double R_ = 0.71428571428, double mu_L_ = 1e-5, double Pr_L_ = 0.72, double Pr_T_ = 0.9,
        double gamma_ = 1.4)
        : R(R_), mu_L(mu_L_), Pr_L(Pr_L_), Pr_T(Pr_T_), gamma(gamma_) {}
    double k() const { return cp * mu_L / Pr_L; }
};

struct boundary_variables {
    double mach = 0.2, angle = 0, T = 1, p = 1.;
    boundary_variables() {}
    boundary_variables(double m, double a = 0, double t = 1, double pp = 1.) : mach(m), angle(a), T(t), p(pp) {}
};

struct boundary_condition {
    std::string bc_type;
    boundary_variables vars_far;
};

class mesh {
    afx_mesh* h_ = nullptr;

public:
    mesh() {}
    explicit mesh(const std::string& filename) { read_file(filename); }
    mesh(const mesh&) = delete;
    mesh& operator=(const mesh&) = delete;
    mesh(mesh&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    mesh& operator=(mesh&& o) noexcept {
        if (this != &o) {
            if (h_) afx_mesh_destroy(h_);
            h_ = o.h_;
            o.h_ = nullptr;
        }
        return *this;
    }
    ~mesh() { if (h_) afx_mesh_destroy(h_); }
    void read_file(const std::string& filename) {
        afx_mesh* m = nullptr;
        check(afx_mesh_read_msh(filename.c_str(), &m));
        if (h_) afx_mesh_destroy(h_);
        h_ = m;
    }
    static mesh naca_ogrid(int ni, int nj, double r_far = 20.0, double thickness = 0.12, double first_h = 2e-3) {
        mesh out;
        check(afx_mesh_gen_naca_ogrid(ni, nj, r_far, thickness, first_h, &out.h_));
        return out;
    }
    const afx_mesh* handle() const { return h_; }
    int64_t count(int k) const {
        int64_t c[5];
        check(afx_mesh_counts(h_, c));
        return c[k];
    }
    int64_t nRealCells() const { return count(1); }
    int64_t nCells() const { return count(2); }
};

// Common part of rans::solver (solver.h:40-175)
class solver {
protected:
    afx_rans* h_ = nullptr;
    afx_opts opts_{};
    std::vector<double> q_;
    int64_t nT_ = 0, nR_ = 0;
    double cfl_ = 1;
    std::string viscosity_model_;
    std::map<std::string, boundary_condition> bcs_;

    // "spalart-allmaras" is the name solver::set_bcs tests (solver.h:206),
    // "spallart-allmaras" the one rans::Settings offers (core.h:176,224)
    static int viscosity_code(const std::string& v) {
        if (v == "inviscid") return 0;
        if (v == "laminar") return 1;
        if (v == "spalart-allmaras" || v == "spallart-allmaras") return 2;
        throw std::invalid_argument("unsupported viscosity model: " + v);
    }

public:
    solver(const mesh& m, const gas& g, std::string viscosity_model) : viscosity_model_(viscosity_model) {
        afx_default_opts(&opts_);
        opts_.viscosity = viscosity_code(viscosity_model);
#ifdef RANS_MICHALAK_LIMITER
        opts_.limiter = 1;   // the same compile-time switch as solver.h:557
#endif
        afx_gas ag{g.R, g.mu_L, g.Pr_L, g.Pr_T, g.gamma};
        check(afx_rans_create(m.handle(), &ag, &opts_, &h_));
        nR_ = m.nRealCells();
        nT_ = m.nCells();
        q_.assign(4 * nT_, 0.0);
    }
    solver(const solver&) = delete;
    solver& operator=(const solver&) = delete;
    virtual ~solver() { if (h_) afx_rans_destroy(h_); }

    void set_bcs(std::map<std::string, boundary_condition> bcs_in) {
        bcs_ = bcs_in;
        std::vector<const char*> names;
        std::vector<int32_t> types;
        std::vector<double> vars;
        for (auto& kv : bcs_) {
            const std::string& t = kv.second.bc_type;
            int32_t code = t == "farfield" ? AFX_BC_FARFIELD : t == "slip-wall" ? AFX_BC_SLIP : t == "wall" ? AFX_BC_WALL : -1;
            if (code < 0) throw std::invalid_argument("unsupported boundary type: " + t);
            names.push_back(kv.first.c_str());
            types.push_back(code);
            const boundary_variables& v = kv.second.vars_far;
            vars.insert(vars.end(), {v.mach, v.angle, v.T, v.p});
        }
        check(afx_rans_set_bcs(h_, (int32_t)names.size(), names.data(), types.data(), vars.data()));
    }
    void set_cfl(const double& cfl_in) { cfl_ = cfl_in; check(afx_rans_set_cfl(h_, cfl_in)); }
    void set_second_order(const bool x = true) { opts_.second_order = x; check(afx_rans_set_options(h_, &opts_)); }
    void set_gradient_scheme(std::string grads) {
        if (grads == "least-squares") opts_.gradient = 0;
        else if (grads == "green-gauss") opts_.gradient = 1;
        else throw std::invalid_argument("unknown gradient scheme: " + grads);
        check(afx_rans_set_options(h_, &opts_));
    }
    void set_limiter_k(const double& x) { opts_.limiter_k = x; check(afx_rans_set_options(h_, &opts_)); }
    void init() { check(afx_rans_init(h_)); }
    void refill_bcs() { check(afx_rans_refill_bcs(h_)); }
    void bcs_from_internal() { check(afx_rans_bcs_from_internal(h_)); }
    double get_uniform_residual() {
        double e = 0;
        check(afx_rans_uniform_residual(h_, &e));
        return e;
    }
    // get_q() (solver.h:255-257): a host copy; write back with set_q()
    std::vector<double>& get_q() {
        check(afx_rans_get_state(h_, q_.data()));
        return q_;
    }
    void set_q(const std::vector<double>& q) {
        if ((int64_t)q.size() != 4 * nT_) throw std::invalid_argument("state has the wrong length");
        check(afx_rans_set_state(h_, q.data()));
    }
    // rho nu~ of the Spalart-Allmaras model (the reference's nu, solver.h:55)
    std::vector<double> get_nu() {
        std::vector<double> nu(nT_);
        check(afx_rans_get_nu(h_, nu.data()));
        return nu;
    }
    void set_nu(const std::vector<double>& nu) {
        if ((int64_t)nu.size() != nT_) throw std::invalid_argument("rho nu~ vector has the wrong length");
        check(afx_rans_set_nu(h_, nu.data()));
    }
    std::map<std::string, boundary_condition> get_bcs() const { return bcs_; }
    std::string get_viscosity_model() const { return viscosity_model_; }
    afx_rans* handle() const { return h_; }
    unsigned get_print_interval() const { return 1; }

    virtual void fill() {}
    virtual int compute() { return -1; }
    virtual double solve(const double relaxation = 1, const double tol = 0, const int rhs_iterations = 5) = 0;
};

// rans::implicitSolver (solver.h:852-1213)
class implicitSolver : public solver {
public:
    implicitSolver(const mesh& m, const gas& g, std::string viscosity_model) : solver(m, g, viscosity_model) {
#ifdef RANS_MATRIX_FREE
        // the same compile-time switch as solver.h:863-868
        check(afx_rans_set_matrix_free(h_, 1));
#endif
    }
    void fill() override { check(afx_rans_implicit_fill(h_)); }
    int compute() override {
        int32_t st = -1;
        check(afx_rans_implicit_compute(h_, &st));
        return st;
    }
    double solve(const double relaxation = 1, const double tol = 0, const int rhs_iterations = 5) override {
        double e = 0;
        check(afx_rans_implicit_solve(h_, relaxation, tol, rhs_iterations, AFX_FP64, &e));
        return e;
    }
};

// rans::explicitSolver (solver.h:721-828)
class explicitSolver : public solver {
public:
    explicitSolver(const mesh& m, const gas& g, std::string viscosity_model) : solver(m, g, viscosity_model) {}
    void fill() override {}
    int compute() override { return 0; }
    double solve(const double relaxation = 1, const double = 0, const int = 5) override {
        double e = 0;
        check(afx_rans_explicit_step(h_, relaxation, &e));
        return e;
    }
};

struct Settings {
    double start_cfl = 40., slope_cfl = 50., max_cfl = 100., relaxation = 0.9, tolerance = 1e-4;
    int rhs_iterations = 5, max_iterations = 300;
    std::string airfoil_name = "wall";   // the patch of the loop's Cp profile (core.h:196)
};

struct wallProfile {
    double cd = 0., cl = 0., cm = 0.;
};

// get_wall_profile (post.h:301-387)
inline wallProfile get_wall_profile(solver& s, const std::string& patch) {
    wallProfile w;
    check(afx_rans_wall_forces(s.handle(), patch.c_str(), &w.cl, &w.cd, &w.cm));
    return w;
}

// CpProfile (post.h:182-290) without its plotting arrays: face centres and
// cp of a patch's faces in boundary order, and the moment point of
// calc_chord_coords (a quarter of the way from the leftmost face centre to
// the rightmost)
struct CpProfile {
    double x_moment = 0., y_moment = 0.;
    std::vector<double> x, y, cp;
    void calc_chord_coords(solver& s, const std::string& af) {
        calc_cp(s, af);
        size_t l = 0, r = 0;
        for (size_t k = 0; k < x.size(); ++k) {
            if (x[k] < x[l]) l = k;
            if (x[k] > x[r]) r = k;
        }
        x_moment = (x[r] - x[l]) * 0.25 + x[l];
        y_moment = (y[r] - y[l]) * 0.25 + y[l];
    }
    void calc_cp(solver& s, const std::string& af) {
        int64_t n = 0;
        check(afx_rans_cp_profile(s.handle(), af.c_str(), &n, nullptr, nullptr, nullptr));
        x.resize((size_t)n); y.resize((size_t)n); cp.resize((size_t)n);
        check(afx_rans_cp_profile(s.handle(), af.c_str(), &n, x.data(), y.data(), cp.data()));
    }
};

// save(filename, solver) (post.h:58-180): the reference's ASCII VTU of the
// current state, the same bytes
inline void save(const std::string& filename, solver& s) { check(afx_rans_save_vtu(s.handle(), filename.c_str())); }

// multigrid<implicitSolver>::run_solver (multigrid.h:240-293); returns 0 on
// success, 1 when the solver reports failure or the residual is not finite;
// residuals receives the normalised residual of every iteration.  With a
// profile, profile->calc_cp(s, st.airfoil_name) after every
// profile_every-th iteration (the reference refreshes it every iteration,
// multigrid.h:286; the gather runs on the device).
inline int run_solver(solver& s, const Settings& st, std::vector<double>& residuals,
                      const std::function<bool()>& stop = nullptr, CpProfile* profile = nullptr,
                      int profile_every = 1) {
    double cfl = st.start_cfl;
    int i = 0;
    double err = 0;
    double err_0 = s.get_uniform_residual();
    do {
        s.set_cfl(cfl);
        s.fill();
        int ok = s.compute();
        if (ok == 0) {
            err = s.solve(st.relaxation, err_0 * st.tolerance, st.rhs_iterations);
            if ((i == 0) && (err > 2 * err_0)) err_0 = err;
        } else {
            err = -1;
        }
        err /= err_0;
        cfl = std::min(st.start_cfl + (i + 1) * st.slope_cfl, st.max_cfl);
        i++;
        // a non-finite normalised residual is a failure (the reference's
        // loop would read NaN as converged)
        if (err < 0 || !std::isfinite(err)) return 1;
        if (profile && i % profile_every == 0) profile->calc_cp(s, st.airfoil_name);
        residuals.push_back(err);
    } while ((err > st.tolerance) && (i < st.max_iterations) && !(stop && stop()));
    return 0;
}

// multigrid<explicitSolver>::run_solver (multigrid.h:183-237): the CFL stays
// at its start value and the residual is normalised only when positive;
// returns 0 on success, 1 on failure or a non-finite residual.  A profile is
// refreshed as in the implicit loop (an addition: the reference's explicit
// loop refreshes none).
inline int run_solver(explicitSolver& s, const Settings& st, std::vector<double>& residuals,
                      const std::function<bool()>& stop = nullptr, CpProfile* profile = nullptr,
                      int profile_every = 1) {
    int i = 0;
    double err = 0;
    double err_0 = s.get_uniform_residual();
    do {
        s.set_cfl(st.start_cfl);
        s.fill();
        int ok = s.compute();
        if (ok == 0) {
            err = s.solve(st.relaxation);
            if ((i == 0) && (err > 2 * err_0)) err_0 = err;
        } else {
            err = -1;
        }
        if (err > 0) err /= err_0;
        if (err < 0 || !std::isfinite(err)) return 1;
        residuals.push_back(err);
        i++;
        if (profile && i % profile_every == 0) profile->calc_cp(s, st.airfoil_name);
    } while ((err > st.tolerance) && (i < st.max_iterations) && !(stop && stop()));
    return 0;
}

// multigrid<T>::gen_mapper (multigrid.h:96-163), built on the device: the
// inverse-distance map from a coarse solver's cells to a fine one's.  The two
// solvers must outlive the mapper.
class mapper {
    afx_mapper* h_ = nullptr;
    solver* coarse_ = nullptr;
    solver* fine_ = nullptr;

public:
    mapper(solver& coarse, solver& fine) : coarse_(&coarse), fine_(&fine) {
        check(afx_mapper_create(coarse.handle(), fine.handle(), &h_));
    }
    mapper(const mapper&) = delete;
    mapper& operator=(const mapper&) = delete;
    mapper(mapper&& o) noexcept : h_(o.h_), coarse_(o.coarse_), fine_(o.fine_) { o.h_ = nullptr; }
    ~mapper() { if (h_) afx_mapper_destroy(h_); }
    // solvers[i].get_q() = mappers[i-1] * solvers[i-1].get_q() (multigrid.h:334)
    void prolongate() { check(afx_mapper_prolongate(h_, coarse_->handle(), fine_->handle())); }
};

// multigrid<T>::run (multigrid.h:296-363) over levels listed coarse to fine,
// each constructed with its boundary conditions and options as the
// multigrid constructor does (multigrid.h:70-77): level 0 from freestream
// (reinit) and its boundary states, then each finer level from the mapped
// solution of the previous one.  Returns the level the loop stopped on;
// state receives run_solver's result for it; profile / profile_every are
// handed to run_solver.
template <class S>
S& multigrid_run(std::vector<S*>& levels, const Settings& st, std::vector<double>& residuals, int* state = nullptr,
                 bool reinit = true, const std::function<bool()>& stop = nullptr,
                 std::vector<mapper>* prebuilt = nullptr, CpProfile* profile = nullptr, int profile_every = 1) {
    std::vector<mapper> own;
    if (!prebuilt) {
        own.reserve(levels.size());
        for (size_t i = 0; i + 1 < levels.size(); ++i) own.emplace_back(*levels[i], *levels[i + 1]);
    }
    std::vector<mapper>& mappers = prebuilt ? *prebuilt : own;
    if (reinit) levels[0]->init();
    levels[0]->refill_bcs();
    for (size_t i = 0; i < levels.size(); ++i) {
        if (i > 0) {
            levels[i - 1]->bcs_from_internal();
            mappers[i - 1].prolongate();
            levels[i]->refill_bcs();
        }
        int rc = run_solver(*levels[i], st, residuals, stop, profile, profile_every);
        if (state) *state = rc;
        if (rc || (stop && stop())) return *levels[i];
    }
    return *levels.back();
}

// database::airfoil (common/database/database.hpp:17-52): the polar of one
// section, alpha in degrees
struct airfoil {
    std::vector<double> alpha, cl, cd, cmy;
};

// Rans::compute_alphas (rans.h:53-57)
inline std::vector<double> compute_alphas(double alpha_start, double alpha_end, double alpha_step) {
    std::vector<double> out;
    for (double a = alpha_start; a <= alpha_end; a += alpha_step) out.push_back(a);
    return out;
}

// Rans::run_airfoil<T> (rans.h:78-106) over solvers built on the airfoil's
// meshes (coarse to fine), each with its options set: the farfield angle is
// db.alpha[0] * 0.01745 when level 0 is initialised, then for every alpha the
// levels get the farfield at alpha * 0.01745, the multigrid loop runs without
// re-initialising (warm start from the previous angle) and the wall profile
// is appended to db; with a non-empty save_prefix the solution file
// <save_prefix>_<alpha>.vtu is written.  Returns the last loop state.
template <class S>
int run_airfoil(std::vector<S*>& levels, const Settings& st, std::map<std::string, boundary_condition> bcs,
                airfoil& db, const std::string& patch = "wall", const std::string& save_prefix = "",
                std::vector<double>* residuals = nullptr, const std::function<bool()>& stop = nullptr) {
    if (db.alpha.empty() || levels.empty()) return 0;
    std::vector<double> own_res;
    std::vector<double>& res = residuals ? *residuals : own_res;
    std::vector<mapper> mappers;
    mappers.reserve(levels.size());
    for (size_t i = 0; i + 1 < levels.size(); ++i) mappers.emplace_back(*levels[i], *levels[i + 1]);
    bcs.at("farfield").vars_far.angle = db.alpha[0] * 0.01745;
    for (S* s : levels) s->set_bcs(bcs);
    levels[0]->init();
    int state = 0;
    for (double alpha : db.alpha) {
        bcs.at("farfield").vars_far.angle = alpha * 0.01745;   // deg to rad, rans.h:94
        for (S* s : levels) s->set_bcs(bcs);
        S& s = multigrid_run(levels, st, res, &state, false, stop, &mappers);
        wallProfile wp = get_wall_profile(s, patch);
        db.cl.push_back(wp.cl);
        db.cd.push_back(wp.cd);
        db.cmy.push_back(wp.cm);
        if (!save_prefix.empty()) save(save_prefix + "_" + std::to_string(alpha) + ".vtu", s);
        if (stop && stop()) break;
    }
    return state;
}

// Rans::solve_airfoil (rans.h:116-123): the solver type of the settings
// (solver_options "explicit" / "implicit", core.h:174,202) picks the device
// solver and loop; the levels are set up as the multigrid constructor does
// (multigrid.h:70-77).  An unknown type is rejected (the reference returns
// without solving).
inline int solve_airfoil(const std::string& solver_type, const std::vector<const mesh*>& meshes, const gas& g,
                         const std::string& viscosity_model, const std::map<std::string, boundary_condition>& bcs,
                         const Settings& st, airfoil& db, bool second_order = true,
                         const std::string& gradient_scheme = "green-gauss", double limiter_k = 5.0,
                         const std::string& patch = "wall", const std::string& save_prefix = "") {
    auto setup = [&](solver& s) {
        s.set_bcs(bcs);
        s.set_second_order(second_order);
        s.set_gradient_scheme(gradient_scheme);
        s.set_limiter_k(limiter_k);
    };
    if (solver_type == "implicit") {
        std::vector<std::unique_ptr<implicitSolver>> own;
        std::vector<implicitSolver*> levels;
        for (const mesh* m : meshes) {
            own.emplace_back(new implicitSolver(*m, g, viscosity_model));
            setup(*own.back());
            levels.push_back(own.back().get());
        }
        return run_airfoil(levels, st, bcs, db, patch, save_prefix);
    }
    if (solver_type == "explicit") {
        std::vector<std::unique_ptr<explicitSolver>> own;
        std::vector<explicitSolver*> levels;
        for (const mesh* m : meshes) {
            own.emplace_back(new explicitSolver(*m, g, viscosity_model));
            setup(*own.back());
            levels.push_back(own.back().get());
        }
        return run_airfoil(levels, st, bcs, db, patch, save_prefix);
    }
    throw std::invalid_argument("unknown solver type: " + solver_type + " (explicit, implicit)");
}

}  // namespace gpu
}  // namespace rans
