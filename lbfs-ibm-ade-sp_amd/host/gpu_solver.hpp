// C++ host mirror of the reference's GPU solver over the C ABI:
// gpu_solver::General_Purpose_Solver_mk_i (Solver_gpu.h:28-31) and the kernel
// groups its time loop launches (Solver_gpu.h:168-174, Solver_gpu.cu:749-1246),
// under the reference's names, so a caller can run whole steps or drive and
// check one group at a time.  Errors become lbfs_amd::lbfs_failure exceptions
// carrying the ABI status and message.
#pragma once

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbfs_abi.h"

namespace lbfs_amd {

class lbfs_failure : public std::runtime_error {
public:
    lbfs_failure(int code, const std::string &what) : std::runtime_error(what), code(code) {}
    int code;
};

class gpu_solver {
public:
    gpu_solver(const lbfs_mesh_desc &mesh, const lbfs_params &params) : n_cells_(mesh.n_cells), n_faces_(mesh.n_faces)
    {
        check(lbfs_create(&mesh, &params, &ctx_), "lbfs_create");
    }
    ~gpu_solver() { lbfs_destroy(ctx_); }
    gpu_solver(const gpu_solver &) = delete;
    gpu_solver &operator=(const gpu_solver &) = delete;

    // lagrangian objects (Solver_gpu.cu:659-732)
    void add_objects(const std::vector<lbfs_membrane_desc> &objects, int ibm_kernel = LBFS_IBM_COS4)
    {
        check(lbfs_add_membranes(ctx_, objects.data(), (int)objects.size(), ibm_kernel), "lbfs_add_membranes");
        n_objects_ = (int)objects.size();
        n_nodes_ = 0;
        for (auto &o : objects) n_nodes_ += o.n_nodes;
    }

    // Solution in and out (Solution::clone, Solution.cpp:308-319)
    void set_solution(const std::vector<double> &rho, const std::vector<double> &u, const std::vector<double> &v,
                      const std::vector<double> &w)
    {
        check(lbfs_set_state(ctx_, rho.data(), u.data(), v.data(), w.data()), "lbfs_set_state");
    }
    void get_solution(std::vector<double> &rho, std::vector<double> &u, std::vector<double> &v, std::vector<double> &w)
    {
        rho.resize(n_cells_); u.resize(n_cells_); v.resize(n_cells_); w.resize(n_cells_);
        check(lbfs_get_state(ctx_, rho.data(), u.data(), v.data(), w.data()), "lbfs_get_state");
    }

    // The time loop: `timesteps` steps; every `output_step` steps the output
    // callback receives the step count and the residual/mass diagnostics
    // (the reference writes its output there, Solver_gpu.cu:1143-1246).
    void General_Purpose_Solver_mk_i(int timesteps, int output_step = 0,
                                     const std::function<void(int, const lbfs_diag &)> &output = {},
                                     double ref_rho = 1.0, double ref_u = 1.0)
    {
        const int chunk = output_step > 0 ? output_step : timesteps;
        for (int t = 0; t < timesteps; t += chunk) {
            const int n = timesteps - t < chunk ? timesteps - t : chunk;
            check(lbfs_step(ctx_, n), "lbfs_step");
            if (output) {
                lbfs_diag d;
                check(lbfs_reduce_diagnostics(ctx_, ref_rho, ref_u, &d), "lbfs_reduce_diagnostics");
                output(t + n, d);
            }
        }
    }

    // ---- kernel groups of one step, in the order of the time loop ----
    // interpolate_velocities_on_nodes + update_node_positions
    void interpolate_velocities_on_nodes() { stage("ibm_interp"); }
    // run_spring_network_functions_atomics
    void run_spring_network_functions_atomics()
    {
        stage("mem_volume");
        stage("mem_edges");
        stage_if_present("mem_nodes");  // the binned spreading gathers the node forces itself
    }
    // update_interior_viscosities_plucker, when the step is due (t = 1 mod 10000)
    void update_interior_viscosities() { stage_if_present("internal_visc"); }
    // spread_forces_on_structured_grid
    void spread_forces_on_structured_grid()
    {
        stage("ibm_spread");
    }
    // get_interior_gradients + get_bc_gradients
    void get_interior_gradients() { stage("gradients"); }
    // get_cfl_device + min (time_stepping 3)
    void get_cfl() { stage_if_present("dt_min"); }
    // calc_face_flux + calc_face_flux_bc + add_face_flux_to_cell + add + time_integration
    // (with membranes the spread force is added by a second, tile-sparse pass)
    void time_integration()
    {
        stage("flux_update");
        stage_if_present("ibm_apply");  // with membranes: the spread force added
    }
    void end_step() { check(lbfs_stage_end_step(ctx_), "lbfs_stage_end_step"); }

    // one step group by group; equals General_Purpose_Solver_mk_i(1)
    void step_by_groups()
    {
        if (n_objects_) {
            interpolate_velocities_on_nodes();
            run_spring_network_functions_atomics();
            update_interior_viscosities();
            spread_forces_on_structured_grid();
        }
        get_interior_gradients();
        get_cfl();
        time_integration();
        end_step();
    }

    // ---- stage outputs for checking (do not advance the state) ----
    std::vector<double> calc_face_flux()
    {
        std::vector<double> f(4 * (size_t)n_faces_);
        check(lbfs_stage_face_flux(ctx_, f.data()), "lbfs_stage_face_flux");
        return f;
    }
    std::vector<double> add_face_flux_to_cell()
    {
        std::vector<double> r(4 * (size_t)n_cells_);
        check(lbfs_stage_residual(ctx_, r.data()), "lbfs_stage_residual");
        return r;
    }
    std::vector<double> object_nodes()
    {
        std::vector<double> x(3 * (size_t)n_nodes_);
        check(lbfs_get_membrane_nodes(ctx_, x.data(), nullptr, nullptr), "lbfs_get_membrane_nodes");
        return x;
    }
    lbfs_ctx *context() { return ctx_; }

private:
    void check(int rc, const char *what)
    {
        if (rc != LBFS_OK) {
            const char *m = ctx_ ? lbfs_last_error(ctx_) : "";
            throw lbfs_failure(rc, std::string(what) + ": " + (m ? m : ""));
        }
    }
    void stage(const char *name) { check(lbfs_stage_run(ctx_, name), name); }
    void stage_if_present(const char *name)
    {
        const char *names[32];
        int n = 0;
        check(lbfs_stage_names(ctx_, 32, names, &n), "lbfs_stage_names");
        for (int k = 0; k < n; k++)
            if (std::string(names[k]) == name) return stage(name);
    }

    lbfs_ctx *ctx_ = nullptr;
    int n_cells_ = 0, n_faces_ = 0, n_objects_ = 0, n_nodes_ = 0;
};

}  // namespace lbfs_amd
