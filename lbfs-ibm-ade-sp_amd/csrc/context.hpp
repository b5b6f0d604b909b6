// The opaque context behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "coupling.hpp"
#include "fixed_point.hpp"
#include "fluid_kernels.hpp"
#include "lbfs_abi.h"
#include "mesh_host.hpp"

namespace lbfs {
struct AdjointState;  // adjoint.cpp
void adjoint_release(lbfs_ctx *ctx);
}  // namespace lbfs

struct lbfs_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;       // second stream: the coupling chain beside the fluid passes (or the reverse)
    hipStream_t side2 = nullptr;      // third stream: membrane forces (step schedule 2)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_fork2 = nullptr, ev_join2 = nullptr;
    lbfs_params params{};
    std::vector<lbfs::HostBlock> blocks;   // logical boxes of the mesh (empty: none)

    lbfs::Geometry geo;
    lbfs::DeviceLayout lay;
    lbfs::FluidArgs args{};
    std::vector<int> bc_vel_type;   // caller velocity BC type per boundary face

    float4 *st[2] = {nullptr, nullptr};
    int cur = 0;
    float4 *grad = nullptr, *res = nullptr, *face_scratch = nullptr;
    double *diag_part = nullptr;
    int *dev_err = nullptr;         // device error word: ERR_FIX_* bits (fixed_point.hpp), sticky until
                                    // the next lbfs_add_membranes
    float *static_force = nullptr;  // [3 n] body force + closure term
    float *cell_force = nullptr;    // [3 n] static force + spread IBM force
    std::vector<void *> allocs;

    // membranes and IBM coupling (lbfs_add_membranes)
    bool have_membranes = false;
    lbfs::MemArgs mem{};
    lbfs::IbmArgs ibm{};
    std::vector<int> mem_node_start;
    double *mem_energy = nullptr;   // [7 n_mem] output of lbfs_get_membrane_energies
    double *energy_scratch = nullptr; // this device's static staging of lbfs_get_membrane_energies_noalloc
    int ibm_binned_option = -1;     // LBFS_OPT_IBM_BINNED: -1 by node density, else IBM_SPREAD_*
    std::vector<double> mem_origin;
    // host copies behind lbfs_set_membrane_moduli: per-membrane parameters,
    // per-edge spring constant per unit shear modulus (FP64), edge ranges
    std::vector<lbfs::MemParamDev> mem_params;
    std::vector<float4> mem_edge_const;
    std::vector<double> mem_ks_unit;
    std::vector<int> mem_edge_start;
    std::vector<double> mem_shear;
    std::vector<void *> coupling_allocs;
    float *visc_factor = nullptr;   // [n] internal viscosity field, null while every ratio is 1
    lbfs::ViscGrid visc_grid{};     // its search structure (allocated with it)
    int visc_search = 1;            // LBFS_OPT_VISC_SEARCH
    float delta_h = 0.1f;           // smoothing width of the field (unstructured_mesh.cpp:263-264)
    long long t_step = 0;           // steps since the membranes were attached

    // checkpoints and buffers of the discrete adjoint (lbfs_adjoint_*)
    lbfs::AdjointState *adj = nullptr;

    // hipGraph of two consecutive steps starting from cur == 0
    int schedule_option = 0;        // LBFS_OPT_STEP_SCHEDULE
    bool fused_ok = false;          // the tiles fit the fused fluid pass's LDS
    bool fused = false;             // LBFS_OPT_FUSED_FLUID: whole steps run the fused fluid pass
    bool use_graph = true;
    hipGraphExec_t graph_exec = nullptr;
    void drop_graph();
    // arguments of the apply pass that ends a coupled step
    lbfs::ApplyArgs apply_args(float4 *st_new, float *cell_force_out, int all) const
    {
        lbfs::ApplyArgs p{};
        p.tile_cell_start = args.tile_cell_start;
        p.n_tiles = lay.n_tiles;
        p.dt = args.dt;
        p.rho0 = args.rho0;
        p.st = st_new;
        p.cell_force_out = cell_force_out;
        p.all = all;
        p.mem_acc = mem.mem_acc;
        p.glob = mem.glob;
        p.n_mem = mem.n_mem;
        return p;
    }

    // lbfs_step_residual: the face pass of an eager step also writes the flux
    // residual (never set while a graph is captured)
    bool res_track = false;
    float4 *res_out() const { return res_track ? res : nullptr; }

    int fail(int code, const char *fmt, ...);
    // waits for the stream and turns a set device error word into LBFS_E_NONFINITE
    int check_device_error();
    int enqueue_step();
    int enqueue_steps(int nsteps);
    void release_extensions();

    // The step as an ordered list of kernel stages (each enqueues one launch
    // on `stream`), with the modelled device bytes of that launch.
    struct Stage {
        const char *name;
        std::function<int()> run;
        double model_bytes;
    };
    // fused: the fluid passes as the one fused launch "fluid" (the step's own
    // launches); otherwise "gradients" and "flux_update" (the stage API)
    std::vector<Stage> step_stages(bool with_visc = false, bool fused = false);
    // the internal viscosity field is rebuilt at t = 1 (mod 10000) (Solver_gpu.cu:884-889)
    bool visc_due(long long t) const { return visc_factor && (t % 10000) == 1; }
    int run_stages(bool with_visc);
};
