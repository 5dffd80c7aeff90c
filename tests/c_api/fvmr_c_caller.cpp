// A C++ caller of libfvmr through include/fvmr.h alone, in the order the reference's host code
// drives its GPU solver (SU2MeshReader + FieldInitializer, GPUSolver2::allocateMemory /
// initializeData, CIteration::iteration_device per step, updateResidualVector, freeMemory):
// reads an SU2 mesh, sets a shock-tube field (inlet 1, 0, 0, 1 | outlet 0.125, 0, 0, 0.1 across
// x = 0), runs N SSP-RK3 steps and prints t, the infinity norm of the density residual and an
// XOR fold of the state's bit patterns.
//
//   fvmr_c_caller MESH.su2 STEPS            run on the device
//   fvmr_c_caller MESH.su2 0 --host-only    mesh and field only (no device)
#include "fvmr.h"

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static int fail(const char *what) {
    std::fprintf(stderr, "%s: %s\n", what, fvmr_last_error());
    return 1;
}

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
        std::fprintf(stderr, "usage: %s MESH.su2 STEPS [--host-only]\n", argv[0]);
        return 2;
    }
    const int steps = std::atoi(argv[2]);
    const bool host_only = argc > 3 && std::strcmp(argv[3], "--host-only") == 0;
    fvmr_mesh *mesh = nullptr;
    if (fvmr_mesh_read_su2(argv[1], FVMR_EQ_EULER, 1.0, &mesh)) return fail("fvmr_mesh_read_su2");
    fvmr_mesh_soa soa;
    if (fvmr_mesh_soa_view(mesh, &soa)) return fail("fvmr_mesh_soa_view");
    fvmr_params p;
    if (fvmr_params_default(&p)) return fail("fvmr_params_default");
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
    std::vector<double> U[4];
    double *rows[4];
    for (int k = 0; k < 4; k++) {
        U[k].assign(soa.num_element, 0.0);
        rows[k] = U[k].data();
    }
    const double shock[4] = {0.0, 0.0, 1.0, 0.0};   // shock_x, shock_y, normal
    if (fvmr_field_init(&soa, &p, FVMR_INIT_SHOCK_TUBE, shock, rows)) return fail("fvmr_field_init");
    if (host_only) {
        std::printf("cells %d edges %d nodes %d sets %d fold %016" PRIx64 "\n", soa.num_element, soa.num_edge,
                    soa.num_node, soa.num_boundary_set, fold(U));
        fvmr_mesh_free(mesh);
        return 0;
    }
    fvmr_handle *h = nullptr;
    if (fvmr_create(&soa, &p, &h)) return fail("fvmr_create");
    const double *crows[4] = {rows[0], rows[1], rows[2], rows[3]};
    if (fvmr_set_state(h, crows)) return fail("fvmr_set_state");
    double t = 0.0;
    fvmr_step_info info;
    std::memset(&info, 0, sizeof info);
    for (int s = 0; s < steps; s++)
        if (fvmr_step(h, &t, 1e9, &info)) return fail("fvmr_step");
    double res[4];
    if (fvmr_residual(h, FVMR_NORM_INF, res)) return fail("fvmr_residual");
    if (fvmr_get_state(h, rows)) return fail("fvmr_get_state");
    std::printf("t %.17g residual %.17g fold %016" PRIx64 "\n", t, res[0], fold(U));
    if (fvmr_destroy(h)) return fail("fvmr_destroy");
    fvmr_mesh_free(mesh);
    return 0;
}
