// Command-line driver of the device solver through the C++ interface in
// rans_gpu.hpp: the steps of Rans::run_airfoil (rans.h:78-106) for one angle
// of attack on one mesh, with the implicit or explicit pseudo-time loop.
//
//   rans_gpu_driver <mesh.msh | ogrid:NI:NJ> <mach> <alpha_deg> <iterations>
//                   [implicit|explicit] [inviscid|laminar] [slip-wall|wall] [tolerance]
//
// A comma-separated list of meshes (coarse to fine) runs the multigrid loop
// of multigrid<T>::run (multigrid.h:296-363) over them.
// Prints one JSON object: iterations, residual history, CL, CD, CM (of the
// last level run), and the size and minimum of the wall's Cp profile, which
// the loop refreshes every iteration (multigrid.h:286).
//
//   rans_gpu_driver polar <meshes> <mach> <alpha_start> <alpha_end> <alpha_step>
//                   <iterations> [implicit|explicit] [wing_alpha_deg]
//
// The chain of app.cpp:168-184: Rans::compute_alphas, Rans::solve_airfoil
// (the alpha sweep with warm starts, rans.h:78-123) into a database::airfoil,
// then that polar as the section table of the nonlinear viscous vortex-lattice
// solve (nonlinear::steady::solve, vlm/src/solver.cpp:330-388) of a
// rectangular AR 8 wing (16 x 32 rings) at wing_alpha_deg (default 3).  Each
// angle runs to a normalised residual of 1e-12 or the iteration count.
// Prints the polar and the wing's iterations, convergence and CL.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <afx_vlm.h>

#include "rans_gpu.hpp"

namespace {

std::vector<std::string> split_list(const std::string& src) {
    std::vector<std::string> out;
    for (size_t a = 0, b; a <= src.size(); a = b + 1) {
        b = src.find(',', a);
        if (b == std::string::npos) b = src.size();
        out.push_back(src.substr(a, b - a));
    }
    return out;
}

void load_mesh(rans::gpu::mesh& m, const std::string& src) {
    if (src.rfind("ogrid:", 0) == 0) {
        int ni = 128, nj = 65;
        std::sscanf(src.c_str() + 6, "%d:%d", &ni, &nj);
        m = rans::gpu::mesh::naca_ogrid(ni, nj);
    } else {
        m.read_file(src);
    }
}

void vchk(int rc) {
    if (rc != AFX_OK) throw std::runtime_error(afx_last_error());
}

int polar_chain(int argc, char** argv) {
    if (argc < 9) {
        std::fprintf(stderr, "usage: %s polar <meshes> <mach> <alpha_start> <alpha_end> <alpha_step> <iterations> "
                             "[implicit|explicit] [wing_alpha_deg]\n", argv[0]);
        return 2;
    }
    const std::vector<std::string> srcs = split_list(argv[2]);
    const double mach = std::atof(argv[3]);
    rans::gpu::Settings st;
    st.max_iterations = std::atoi(argv[7]);
    st.tolerance = 1e-12;   // each angle to its converged state
    const std::string kind = argc > 8 ? argv[8] : "implicit";
    const double wing_alpha = argc > 9 ? std::atof(argv[9]) : 3.0;
    if (kind == "explicit") {
        st.start_cfl = st.max_cfl = 0.8;
        st.slope_cfl = 0.0;
    }
    std::vector<rans::gpu::mesh> meshes(srcs.size());
    std::vector<const rans::gpu::mesh*> mp;
    for (size_t l = 0; l < srcs.size(); ++l) {
        load_mesh(meshes[l], srcs[l]);
        mp.push_back(&meshes[l]);
    }
    std::map<std::string, rans::gpu::boundary_condition> bcs;
    bcs["farfield"].bc_type = "farfield";
    bcs["farfield"].vars_far = rans::gpu::boundary_variables(mach, 0.0, 1.0, 1.0);
    bcs["wall"].bc_type = "slip-wall";
    rans::gpu::airfoil db;
    db.alpha = rans::gpu::compute_alphas(std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6]));
    const int state = rans::gpu::solve_airfoil(kind, mp, rans::gpu::gas(), "inviscid", bcs, st, db);

    // the polar as the wing's section table: one airfoil at both span ends
    afx_wing_spec ws{16, 32, 8.0, 1.0, 0.0, 1.0, 1};
    int32_t nn = 0, nr = 0, ns = 0;
    double ref[3];
    vchk(afx_wing_lattice(&ws, &nn, &nr, &ns, nullptr, nullptr, nullptr, ref));
    std::vector<double> xyz(3 * (size_t)nn);
    std::vector<int32_t> rings(4 * (size_t)nr), soff((size_t)ns + 1);
    vchk(afx_wing_lattice(&ws, &nn, &nr, &ns, xyz.data(), rings.data(), soff.data(), ref));
    afx_vlm* v = nullptr;
    vchk(afx_vlm_create(nn, xyz.data(), ns, soff.data(), rings.data(), &v));
    const int32_t offs[2] = {0, (int32_t)db.alpha.size()};
    const int32_t sec[2] = {0, 0};
    const double span[2] = {0.0, 1.0};
    afx_polar pol{1, offs, db.alpha.data(), db.cl.data(), db.cd.data(), db.cmy.data(), 2, sec, span};
    afx_vlm_flow fl{wing_alpha, 1.0, 1.0, ref[0], ref[1], 100.0 * ref[1], 0.0};
    afx_nonlinear_info info{};
    int rc = afx_vlm_set_flow(v, &fl);
    if (rc == AFX_OK) rc = afx_vlm_nonlinear_solve(v, &pol, 1.0, 1e-10, 200, &info, nullptr, nullptr);
    afx_vlm_destroy(v);
    vchk(rc);
    std::printf("{\"state\": %d, \"alpha\": [", state);
    for (size_t i = 0; i < db.alpha.size(); ++i) std::printf("%s%.17g", i ? ", " : "", db.alpha[i]);
    std::printf("], \"cl\": [");
    for (size_t i = 0; i < db.cl.size(); ++i) std::printf("%s%.17g", i ? ", " : "", db.cl[i]);
    std::printf("], \"cd\": [");
    for (size_t i = 0; i < db.cd.size(); ++i) std::printf("%s%.17g", i ? ", " : "", db.cd[i]);
    std::printf("], \"cm\": [");
    for (size_t i = 0; i < db.cmy.size(); ++i) std::printf("%s%.17g", i ? ", " : "", db.cmy[i]);
    std::printf("], \"wing\": {\"alpha\": %.17g, \"iterations\": %d, \"converged\": %d, \"cl\": %.17g, "
                "\"cd\": %.17g}}\n", wing_alpha, info.iterations, info.converged, info.cl, info.cd);
    return state;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "polar") == 0) {
        try {
            return polar_chain(argc, argv);
        } catch (std::exception& e) {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 3;
        }
    }
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <mesh.msh|ogrid:NI:NJ> <mach> <alpha_deg> <iterations> "
                             "[implicit|explicit] [inviscid|laminar|spalart-allmaras] [slip-wall|wall] [tolerance] "
                             "[mu_L] [start_cfl:slope_cfl:max_cfl] [relaxation] [rhs_iterations]\n", argv[0]);
        return 2;
    }
    try {
        const std::string src = argv[1];
        const double mach = std::atof(argv[2]), alpha_deg = std::atof(argv[3]);
        rans::gpu::Settings st;
        st.max_iterations = std::atoi(argv[4]);
        const std::string kind = argc > 5 ? argv[5] : "implicit";
        const std::string visc = argc > 6 ? argv[6] : "inviscid";
        const std::string wall = argc > 7 ? argv[7] : "slip-wall";
        st.tolerance = argc > 8 ? std::atof(argv[8]) : 1e-4;
        const std::vector<std::string> srcs = split_list(src);
        rans::gpu::gas g;
        if (argc > 9) g.mu_L = std::atof(argv[9]);
        if (kind == "explicit") {
            st.start_cfl = st.max_cfl = 0.8;
            st.slope_cfl = 0.0;
        }
        if (argc > 10 && std::sscanf(argv[10], "%lf:%lf:%lf", &st.start_cfl, &st.slope_cfl, &st.max_cfl) != 3)
            throw std::runtime_error("CFL schedule must read start:slope:max");
        if (argc > 11) st.relaxation = std::atof(argv[11]);
        if (argc > 12) st.rhs_iterations = std::atoi(argv[12]);
        std::map<std::string, rans::gpu::boundary_condition> bcs;
        bcs["farfield"].bc_type = "farfield";
        bcs["farfield"].vars_far = rans::gpu::boundary_variables(mach, alpha_deg * 0.01745, 1.0, 1.0);   // rans.h:94
        bcs["wall"].bc_type = wall;
        // one solver per mesh, set up as the multigrid constructor does (multigrid.h:70-77)
        std::vector<rans::gpu::mesh> meshes(srcs.size());
        std::vector<std::unique_ptr<rans::gpu::solver>> owned;
        std::vector<rans::gpu::solver*> levels;
        for (size_t l = 0; l < srcs.size(); ++l) {
            load_mesh(meshes[l], srcs[l]);
            if (kind == "explicit") owned.emplace_back(new rans::gpu::explicitSolver(meshes[l], g, visc));
            else owned.emplace_back(new rans::gpu::implicitSolver(meshes[l], g, visc));
            rans::gpu::solver* s = owned.back().get();
            s->set_bcs(bcs);
            s->set_second_order(true);
            s->set_gradient_scheme("green-gauss");
            s->set_limiter_k(5.0);
            levels.push_back(s);
        }
        std::vector<double> res;
        int state = 0;
        // the wall's Cp profile refreshed every iteration, as multigrid.h:286 does
        rans::gpu::CpProfile prof;
        rans::gpu::solver& last = rans::gpu::multigrid_run(levels, st, res, &state, true, nullptr, nullptr, &prof);
        rans::gpu::wallProfile w = rans::gpu::get_wall_profile(last, "wall");
        double cp_min = 0.0;
        for (double c : prof.cp) cp_min = std::min(cp_min, c);
        std::printf("{\"state\": %d, \"iterations\": %zu, \"residuals\": [", state, res.size());
        for (size_t i = 0; i < res.size(); ++i) std::printf("%s%.17g", i ? ", " : "", res[i]);
        std::printf("], \"cl\": %.17g, \"cd\": %.17g, \"cm\": %.17g, \"cp_faces\": %zu, \"cp_min\": %.17g}\n",
                    w.cl, w.cd, w.cm, prof.cp.size(), cp_min);
        return state;
    } catch (std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
}
