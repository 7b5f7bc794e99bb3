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
// last level run).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "rans_gpu.hpp"

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <mesh.msh|ogrid:NI:NJ> <mach> <alpha_deg> <iterations> "
                             "[implicit|explicit] [inviscid|laminar] [slip-wall|wall] [tolerance]\n", argv[0]);
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
        std::vector<std::string> srcs;
        for (size_t a = 0, b; a <= src.size(); a = b + 1) {
            b = src.find(',', a);
            if (b == std::string::npos) b = src.size();
            srcs.push_back(src.substr(a, b - a));
        }
        rans::gpu::gas g;
        if (kind == "explicit") {
            st.start_cfl = st.max_cfl = 0.8;
            st.slope_cfl = 0.0;
        }
        std::map<std::string, rans::gpu::boundary_condition> bcs;
        bcs["farfield"].bc_type = "farfield";
        bcs["farfield"].vars_far = rans::gpu::boundary_variables(mach, alpha_deg * 0.01745, 1.0, 1.0);   // rans.h:94
        bcs["wall"].bc_type = wall;
        // one solver per mesh, set up as the multigrid constructor does (multigrid.h:70-77)
        std::vector<rans::gpu::mesh> meshes(srcs.size());
        std::vector<std::unique_ptr<rans::gpu::solver>> owned;
        std::vector<rans::gpu::solver*> levels;
        for (size_t l = 0; l < srcs.size(); ++l) {
            if (srcs[l].rfind("ogrid:", 0) == 0) {
                int ni = 128, nj = 65;
                std::sscanf(srcs[l].c_str() + 6, "%d:%d", &ni, &nj);
                meshes[l] = rans::gpu::mesh::naca_ogrid(ni, nj);
            } else {
                meshes[l].read_file(srcs[l]);
            }
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
        rans::gpu::solver& last = rans::gpu::multigrid_run(levels, st, res, &state);
        rans::gpu::wallProfile w = rans::gpu::get_wall_profile(last, "wall");
        std::printf("{\"state\": %d, \"iterations\": %zu, \"residuals\": [", state, res.size());
        for (size_t i = 0; i < res.size(); ++i) std::printf("%s%.17g", i ? ", " : "", res[i]);
        std::printf("], \"cl\": %.17g, \"cd\": %.17g, \"cm\": %.17g}\n", w.cl, w.cd, w.cm);
        return state;
    } catch (std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
}
