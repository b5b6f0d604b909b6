// The reference's step entry with its own parameter list,
// gpu_solver::General_Purpose_Solver_mk_i (Solver_gpu.h:28-31,
// Solver_gpu.cu:178-1275), over the C ABI.
//
// Templated on the caller's classes, it reads them only through the members
// the reference declares, so it compiles inside the reference against its own
// headers (the member function's body becomes one call of this template) and
// against stand-ins in a test:
//   global_variables  import_file, gradient_calc_type ("hybrid" | "LS" | else
//                     Green-Gauss, gradients.cpp:255-265), visc,
//                     pre_conditioned_gamma, fneq_min, time_marching_step,
//                     gpu_time_stepping, max_velocity, simulation_length,
//                     output_step, tolerance, ref_rho
//   domain_geometry   X, Y, Z, dx, dy, dz, origin_x, origin_y, origin_z
//   initial_conditions force_x, force_y, force_z (pressure_gradient along
//                     pressure_direction, preprocessor.cpp:714-736), average_rho
//   Boundary_Conditions get_name, get_rho_type, get_vel_type, get_rho, get_u,
//                     get_v, get_w per boundary face
//   Solution          get_rho, get_u, get_v, get_w, set_rho, set_u, set_v, set_w
//   lagrangian_object type (2 = spring network), num_nodes, num_tets, node_x,
//                     node_y, node_z (the stress-free shape at entry),
//                     tet_connectivity, mesh_file, mz_importer, centre_x/y/z
//                     (the initial shape, read_initial_mesh), the moduli,
//                     ratios, viscosity, sphere_reference
// The other parameters (mesh, source, quad_bcs_orig, mg, residual, fmg, pp)
// are the reference's and are not read: the library reads the polyMesh
// directory itself.
//
// The loop is the reference's: timesteps = ceil(simulation_length)
// (Solver_gpu.cu:373), each step's residual RMS checked against tolerance
// (max_error(), Solver_gpu.cu:1206), the solution copied back into soln at the
// end.  Errors throw lbfs_failure with the ABI status and message (the
// reference exits instead).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "gpu_solver.hpp"
#include "lbfs_abi.h"

namespace lbfs_amd {

// choices of the build with no counterpart in the reference's arguments
struct step_entry_options {
    int lattice = LBFS_D3Q15;                // the reference lattice (Solver_gpu.cu:1407-1445)
    int flags = LBFS_FLAGS_REFERENCE_GPU;    // the reference GPU solver's face semantics
    int ibm_kernel = LBFS_IBM_LINEAR2;       // the active IBM kernel (immersed_boundary_method.cu:14-97)
    int check_every = 1;                     // residual check interval: 1 = every step, as the reference
    // called every output_step steps and at the end with the step count and
    // diagnostics: where the reference writes its output (Solver_gpu.cu:1143-1246)
    std::function<void(int, const lbfs_diag &)> output;
};

namespace detail {
inline void check_rc(lbfs_ctx *ctx, int rc, const char *what)
{
    if (rc != LBFS_OK) {
        std::string msg = std::string(what) + ": " + lbfs_last_error(ctx);
        if (ctx) lbfs_destroy(ctx);
        throw lbfs_failure(rc, msg);
    }
}
inline int gradient_type(const std::string &s)
{
    if (s == "hybrid") return LBFS_GRAD_HYBRID;
    if (s == "LS") return LBFS_GRAD_LS;
    return LBFS_GRAD_GG;
}
}  // namespace detail

template <class unstructured_mesh, class Solution, class Boundary_Conditions, class external_forces,
          class global_variables, class domain_geometry, class initial_conditions, class unstructured_bcs,
          class post_processing, class lagrangian_object>
void General_Purpose_Solver_mk_i(unstructured_mesh &Mesh, Solution &soln, Boundary_Conditions &bcs,
                                 external_forces &source, global_variables &globals, domain_geometry &domain,
                                 initial_conditions &init_conds, unstructured_bcs &quad_bcs_orig, int mg,
                                 Solution &residual, int fmg, post_processing &pp,
                                 std::vector<lagrangian_object> &object_vec,
                                 const step_entry_options &opt = step_entry_options())
{
    (void)Mesh; (void)source; (void)quad_bcs_orig; (void)mg; (void)residual; (void)fmg; (void)pp;
    lbfs_ctx *ctx = nullptr;
    // mesh: the polyMesh directory the reference imports (unstructured_mesh.cpp:670-941)
    const std::string dir = globals.import_file;
    int np = 0, nc = 0, nf = 0, ni = 0, npatch = 0;
    detail::check_rc(nullptr, lbfs_polymesh_counts(dir.c_str(), &np, &nc, &nf, &ni, &npatch), "lbfs_polymesh_counts");
    std::vector<double> pts(3 * (size_t)np), shift(3 * (size_t)nf);
    std::vector<int> fv(4 * (size_t)nf), own(nf), nb(ni), patch(nf);
    detail::check_rc(nullptr, lbfs_polymesh_read(dir.c_str(), pts.data(), fv.data(), own.data(), nb.data(),
                                                 patch.data(), shift.data()),
                     "lbfs_polymesh_read");
    // boundary conditions per boundary face (preprocessor.cpp:469-491); the
    // CPU path zeroes "empty" patches (Solver.cpp:308-310)
    const int nbc = nf - ni;
    std::vector<int> rt(nbc), vt(nbc);
    std::vector<double> bv(4 * (size_t)nbc);
    for (int b = 0; b < nbc; b++) {
        const bool empty = std::string(bcs.get_name(b)) == "empty";
        rt[b] = empty ? LBFS_BC_EMPTY : bcs.get_rho_type(b);
        vt[b] = empty ? LBFS_BC_EMPTY : bcs.get_vel_type(b);
        bv[4 * (size_t)b] = bcs.get_rho(b);
        bv[4 * (size_t)b + 1] = bcs.get_u(b);
        bv[4 * (size_t)b + 2] = bcs.get_v(b);
        bv[4 * (size_t)b + 3] = bcs.get_w(b);
    }
    lbfs_mesh_desc md;
    std::memset(&md, 0, sizeof(md));
    md.n_points = np; md.points = pts.data(); md.n_cells = nc; md.n_faces = nf; md.n_internal = ni;
    md.face_vertices = fv.data(); md.owner = own.data(); md.neighbour = nb.data(); md.face_shift = shift.data();
    md.bc_rho_type = rt.data(); md.bc_vel_type = vt.data(); md.bc_values = bv.data();
    // logical numbering of generated boxes (cell i + j nx + k nx ny); left 0
    // when the box does not match the cell count
    if (domain.dx > 0 && domain.dy > 0 && domain.dz > 0) {
        const long lx = std::lround(domain.X / domain.dx), ly = std::lround(domain.Y / domain.dy),
                   lz = std::lround(domain.Z / domain.dz);
        if (lx * ly * lz == (long)nc) {
            md.nx = (int)lx; md.ny = (int)ly; md.nz = (int)lz;
            md.origin[0] = domain.origin_x; md.origin[1] = domain.origin_y; md.origin[2] = domain.origin_z;
            md.spacing = domain.dx;
        }
    }
    lbfs_params p;
    std::memset(&p, 0, sizeof(p));
    p.lattice = opt.lattice;
    p.gradient_type = detail::gradient_type(globals.gradient_calc_type);
    p.flags = opt.flags;
    p.visc = globals.visc;
    p.gamma = globals.pre_conditioned_gamma;
    p.fneq_min = globals.fneq_min;
    p.dt = globals.time_marching_step;
    p.rho0 = init_conds.average_rho > 0.0 ? init_conds.average_rho : 1.0;
    p.time_stepping = globals.gpu_time_stepping;
    p.max_velocity = globals.max_velocity;
    p.channel_diameter = domain.Y;
    // the force vector the GPU solver fills each step (Solver_gpu.cu:834-839)
    p.body_force[0] = init_conds.force_x;
    p.body_force[1] = init_conds.force_y;
    p.body_force[2] = init_conds.force_z;
    detail::check_rc(nullptr, lbfs_create(&md, &p, &ctx), "lbfs_create");

    std::vector<double> r(nc), u(nc), v(nc), w(nc);
    for (int i = 0; i < nc; i++) {
        r[i] = soln.get_rho(i); u[i] = soln.get_u(i); v[i] = soln.get_v(i); w[i] = soln.get_w(i);
    }
    detail::check_rc(ctx, lbfs_set_state(ctx, r.data(), u.data(), v.data(), w.data()), "lbfs_set_state");

    // spring-network membranes (Solver_gpu.cu:659-732): contour lengths from
    // the shape the object holds, positions from its mesh file
    std::vector<lbfs_membrane_desc> mds;
    std::vector<std::vector<double>> sf(object_vec.size()), init(object_vec.size());
    for (size_t k = 0; k < object_vec.size(); k++) {
        auto &o = object_vec[k];
        if (o.type != 2) continue;
        sf[k].resize(3 * (size_t)o.num_nodes);
        for (int i = 0; i < o.num_nodes; i++) {
            sf[k][3 * (size_t)i] = o.node_x[i]; sf[k][3 * (size_t)i + 1] = o.node_y[i]; sf[k][3 * (size_t)i + 2] = o.node_z[i];
        }
        init[k] = sf[k];
        if (!std::string(o.mesh_file).empty()) {
            const int fmt = o.mz_importer ? LBFS_MESH_MZ : LBFS_MESH_OSAY;
            int n = 0, t = 0, s = 0;
            detail::check_rc(ctx, lbfs_membrane_mesh_counts(o.mesh_file.c_str(), fmt, &n, &t, &s), "lbfs_membrane_mesh_counts");
            if (n != o.num_nodes) detail::check_rc(ctx, LBFS_E_INVALID_ARG, "mesh_file node count");
            std::vector<int> tri(3 * (size_t)t);
            const double centre[3] = {o.centre_x, o.centre_y, o.centre_z};
            detail::check_rc(ctx, lbfs_membrane_mesh_read(o.mesh_file.c_str(), fmt, centre, init[k].data(), tri.data(),
                                                          nullptr),
                             "lbfs_membrane_mesh_read");
        }
        lbfs_membrane_desc d;
        std::memset(&d, 0, sizeof(d));
        d.n_nodes = o.num_nodes;
        d.n_tri = o.num_tets;
        d.nodes = init[k].data();
        d.nodes_stress_free = sf[k].data();
        d.tri = o.tet_connectivity;
        d.shear_modulus = o.shear_modulus; d.wlc_ratio = o.wlc_ratio; d.volume_modulus = o.volume_modulus;
        d.local_area_modulus = o.local_area_modulus; d.global_area_modulus = o.global_area_modulus;
        d.global_bending_modulus = o.global_bending_modulus; d.local_bending_modulus = o.local_bending_modulus;
        d.viscosity = o.viscosity; d.curvature_ratio = o.curvature_ratio; d.volume_ratio = o.volume_ratio;
        d.sphere_reference = o.sphere_reference ? 1 : 0;
        d.mad_compat = 1;  // the MAD partial sum of Solver_gpu.cu:2355
        d.internal_viscosity_ratio = o.internal_viscosity_ratio;
        d.spring_law = LBFS_SPRING_WLC;
        mds.push_back(d);
    }
    if (!mds.empty())
        detail::check_rc(ctx, lbfs_add_membranes(ctx, mds.data(), (int)mds.size(), opt.ibm_kernel), "lbfs_add_membranes");

    // the time loop: stop at ceil(simulation_length) steps or when a step's
    // momentum residual RMS falls below tolerance
    const int timesteps = (int)std::ceil(globals.simulation_length);
    const int chunk = globals.output_step > 0 ? globals.output_step : timesteps;
    int t = 0;
    lbfs_diag dg;
    std::memset(&dg, 0, sizeof(dg));
    bool converged = false;
    while (t < timesteps && !converged) {
        const int n = std::min(chunk, timesteps - t);
        int done = 0, conv = 0;
        detail::check_rc(ctx, lbfs_run_to_tolerance(ctx, n, globals.tolerance, opt.check_every, globals.ref_rho,
                                                    globals.max_velocity, &done, &conv, &dg),
                         "lbfs_run_to_tolerance");
        t += done;
        converged = conv != 0;
        if (opt.output) opt.output(t, dg);
    }
    detail::check_rc(ctx, lbfs_get_state(ctx, r.data(), u.data(), v.data(), w.data()), "lbfs_get_state");
    for (int i = 0; i < nc; i++) {
        soln.set_rho(i, r[i]); soln.set_u(i, u[i]); soln.set_v(i, v[i]); soln.set_w(i, w[i]);
    }
    lbfs_destroy(ctx);
}

}  // namespace lbfs_amd
