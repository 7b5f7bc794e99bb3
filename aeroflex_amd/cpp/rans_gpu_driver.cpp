// Command-line driver of the device solver through the C++ interface in
// rans_gpu.hpp: the steps of Rans::run_airfoil (rans.h:78-106) for one angle
// of attack on one mesh, with the implicit or explicit pseudo-time loop.
//
//   rans_gpu_driver <mesh.msh | ogrid:NI:NJ> <mach> <alpha_deg> <iterations>
//                   [implicit|explicit] [inviscid|laminar] [slip-wall|wall] [tolerance]
//
// Prints one JSON object: iterations, residual history, CL, CD, CM.
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
        rans::gpu::mesh m;
        if (src.rfind("ogrid:", 0) == 0) {
            int ni = 128, nj = 65;
            std::sscanf(src.c_str() + 6, "%d:%d", &ni, &nj);
            m = rans::gpu::mesh::naca_ogrid(ni, nj);
        } else {
            m.read_file(src);
        }
        rans::gpu::gas g;
        std::unique_ptr<rans::gpu::solver> s;
        if (kind == "explicit") {
            s.reset(new rans::gpu::explicitSolver(m, g, visc));
            st.start_cfl = st.max_cfl = 0.8;
            st.slope_cfl = 0.0;
        } else {
            s.reset(new rans::gpu::implicitSolver(m, g, visc));
        }
        std::map<std::string, rans::gpu::boundary_condition> bcs;
        bcs["farfield"].bc_type = "farfield";
        bcs["farfield"].vars_far = rans::gpu::boundary_variables(mach, alpha_deg * 0.01745, 1.0, 1.0);   // rans.h:94
        bcs["wall"].bc_type = wall;
        s->set_bcs(bcs);
        s->set_second_order(true);
        s->set_gradient_scheme("green-gauss");
        s->set_limiter_k(5.0);
        s->init();
        s->refill_bcs();
        std::vector<double> res;
        int state = rans::gpu::run_solver(*s, st, res);
        rans::gpu::wallProfile w = rans::gpu::get_wall_profile(*s, "wall");
        std::printf("{\"state\": %d, \"iterations\": %zu, \"residuals\": [", state, res.size());
        for (size_t i = 0; i < res.size(); ++i) std::printf("%s%.17g", i ? ", " : "", res[i]);
        std::printf("], \"cl\": %.17g, \"cd\": %.17g, \"cm\": %.17g}\n", w.cl, w.cd, w.cm);
        return state;
    } catch (std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
}
