/* C++ host layer over the C ABI of fvmr.h with the reference's class and method names, for C++
 * callers that drive the solver as FVMR's own host code does:
 *   GPU::GPUSolver2            FVMR/src/solvers/GPUSolver2.h:32-89 (allocateMemory, initializeData,
 *                              host_to_device / device_to_host, updateResidualVector,
 *                              updateOutputNodeField, freeMemory; residualVector[4] starts at 1)
 *   CIteration::iteration_device  FVMR/src/time/CIteration.h:9 (one step: dt, evolve, restrict)
 * Header only; the device arrays live in libfvmr.  Where the reference logs and exits, these
 * functions throw fvmr::Error with fvmr_last_error(). */
#ifndef FVMR_HPP
#define FVMR_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "fvmr.h"

namespace fvmr {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char *what) {
    if (rc != FVMR_OK) throw Error(rc, std::string(what) + ": " + fvmr_last_error());
}

}  // namespace fvmr

namespace GPU {

/* The device solver of one mesh.  elementField_U is the host copy of U (caller's element order,
 * the reference's elementField_host.U); outputNodeField the node quantities of the writers. */
class GPUSolver2 {
public:
    struct OutputNodeField {
        std::vector<double> ruvp[4], Cp, Ma;
    };

    double residualVector[4] = {1.0, 1.0, 1.0, 1.0};
    std::vector<double> elementField_U[4];
    OutputNodeField outputNodeField;

    GPUSolver2(const fvmr_mesh_soa &mesh, const fvmr_params &params) : mesh_(mesh), params_(params) {}
    ~GPUSolver2() { release(); }
    GPUSolver2(const GPUSolver2 &) = delete;
    GPUSolver2 &operator=(const GPUSolver2 &) = delete;

    /* device arrays (renumbered mesh, state buffers) and the host field, sized from the mesh */
    void allocateMemory() {
        if (h_) return;
        fvmr::check(fvmr_create(&mesh_, &params_, &h_), "fvmr_create");
        for (auto &u : elementField_U) u.assign(mesh_.num_element, 0.0);
    }
    /* the host field (set by the caller, e.g. from fvmr_field_init) to the device */
    void initializeData() { host_to_device(); }
    void host_to_device() {
        const double *rows[4] = {elementField_U[0].data(), elementField_U[1].data(), elementField_U[2].data(),
                                 elementField_U[3].data()};
        fvmr::check(fvmr_set_state(handle(), rows), "fvmr_set_state");
    }
    void device_to_host() {
        double *rows[4] = {elementField_U[0].data(), elementField_U[1].data(), elementField_U[2].data(),
                           elementField_U[3].data()};
        fvmr::check(fvmr_get_state(handle(), rows), "fvmr_get_state");
    }
    /* infinity norm of the last step's dU/dt per variable (get_residual_host's default) */
    void updateResidualVector() { fvmr::check(fvmr_residual(handle(), FVMR_NORM_INF, residualVector), "fvmr_residual"); }
    /* ruvp, Cp and Ma per node of the current state */
    void updateOutputNodeField() {
        OutputNodeField &f = outputNodeField;
        for (auto &r : f.ruvp) r.assign(mesh_.num_node, 0.0);
        f.Cp.assign(mesh_.num_node, 0.0);
        f.Ma.assign(mesh_.num_node, 0.0);
        double *rows[4] = {f.ruvp[0].data(), f.ruvp[1].data(), f.ruvp[2].data(), f.ruvp[3].data()};
        fvmr::check(fvmr_node_field(handle(), rows, f.Cp.data(), f.Ma.data()), "fvmr_node_field");
    }
    void freeMemory() {
        if (h_) fvmr::check(fvmr_destroy(h_), "fvmr_destroy");
        h_ = nullptr;
    }

    void setIterationStarted(bool b) { iterationStarted_ = b; }
    bool isIterationStarted() const { return iterationStarted_; }
    bool isDeviceMemoryAllocated() const { return h_ != nullptr; }
    fvmr_handle *handle() {
        if (!h_) throw fvmr::Error(FVMR_ERR_INVALID_ARG, "GPUSolver2: allocateMemory() has not been called");
        return h_;
    }

private:
    void release() noexcept {
        if (h_) fvmr_destroy(h_);
        h_ = nullptr;
    }
    fvmr_mesh_soa mesh_;
    fvmr_params params_;
    fvmr_handle *h_ = nullptr;
    bool iterationStarted_ = false;
};

}  // namespace GPU

class CIteration {
public:
    /* one step: global dt clamped to T - t, evolve (forward Euler or SSP-RK3), restriction; t advanced */
    static void iteration_device(double &t, double T, GPU::GPUSolver2 *pSolver) {
        fvmr::check(fvmr_step(pSolver->handle(), &t, T, nullptr), "fvmr_step");
        pSolver->setIterationStarted(true);
    }
};

#endif  // FVMR_HPP
