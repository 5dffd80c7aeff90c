// The run of fvmr_c_caller.cpp written against include/fvmr.hpp, the C++ layer with the
// reference's names (GPU::GPUSolver2, CIteration::iteration_device): the same SU2 mesh,
// shock-tube field and N SSP-RK3 steps; prints t, the density residual, the state fold and the
// mean node density.  Errors arrive as fvmr::Error.
//
//   fvmr_cpp_mirror MESH.su2 STEPS
#include "fvmr.hpp"

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static uint64_t fold(const std::vector<double> U[4]) {
    uint64_t h = 0;
    for (int k = 0; k < 4; k++)
        for (double v : U[k]) {
            uint64_t b;
            std::memcpy(&b, &v, sizeof b);
            h ^= b + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        }
    return h;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s MESH.su2 STEPS\n", argv[0]);
        return 2;
    }
    try {
        fvmr_mesh *mesh = nullptr;
        fvmr::check(fvmr_mesh_read_su2(argv[1], FVMR_EQ_EULER, 1.0, &mesh), "fvmr_mesh_read_su2");
        fvmr_mesh_soa soa;
        fvmr::check(fvmr_mesh_soa_view(mesh, &soa), "fvmr_mesh_soa_view");
        fvmr_params p;
        fvmr::check(fvmr_params_default(&p), "fvmr_params_default");
        const double in[4] = {1.0, 0.0, 0.0, 1.0}, out[4] = {0.125, 0.0, 0.0, 0.1};
        for (int k = 0; k < 4; k++) {
            p.inlet_ruvp[k] = in[k];
            p.outlet_ruvp[k] = out[k];
            p.inf_ruvp[k] = out[k];
        }
        p.CFL = 0.5;
        p.flux_scheme = FVMR_SOL_ROE;
        p.reconstruct = FVMR_REC_LINEAR;
        p.gradient = FVMR_GRA_GREENGAUSS;
        p.limiter = FVMR_LIM_BARTH;
        p.time_advance = FVMR_EVO_RK3;
        {
            GPU::GPUSolver2 solver(soa, p);
            solver.allocateMemory();
            double *rows[4] = {solver.elementField_U[0].data(), solver.elementField_U[1].data(),
                               solver.elementField_U[2].data(), solver.elementField_U[3].data()};
            const double shock[4] = {0.0, 0.0, 1.0, 0.0};
            fvmr::check(fvmr_field_init(&soa, &p, FVMR_INIT_SHOCK_TUBE, shock, rows), "fvmr_field_init");
            solver.initializeData();
            double t = 0.0;
            const int steps = std::atoi(argv[2]);
            for (int s = 0; s < steps; s++) CIteration::iteration_device(t, 1e9, &solver);
            solver.device_to_host();
            solver.updateResidualVector();
            solver.updateOutputNodeField();
            double rho = 0.0;
            for (double r : solver.outputNodeField.ruvp[0]) rho += r;
            std::printf("t %.17g residual %.17g fold %016" PRIx64 " node_rho_mean %.17g\n", t, solver.residualVector[0],
                        fold(solver.elementField_U), rho / soa.num_node);
            solver.freeMemory();
        }
        fvmr_mesh_free(mesh);
    } catch (const fvmr::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
