// Check of the reference-signature step entry (reference_entry.hpp): minimal
// stand-ins carrying the members the reference's classes declare (the names
// of Solution.h, Boundary_Conditions.h, global_variables.h,
// domain_geometry.h, initial_conditions.h, lagrangian_object.h), a jittered
// channel written as a polyMesh directory and one biconcave cell.  The entry
// must give the solution and nodes of the same run made through the C ABI
// directly, bit for bit; a loose tolerance must stop it after one step.
// Exit status: 0 agree, 1 disagree, 2 other failure, 3 no HIP device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "reference_entry.hpp"

namespace ref {
struct unstructured_mesh {};
struct external_forces {};
struct unstructured_bcs {};
struct post_processing {};
struct Solution {
    std::vector<double> rho, u, v, w;
    explicit Solution(int n) : rho(n, 1.0), u(n), v(n), w(n) {}
    double get_rho(int i) { return rho[i]; }
    double get_u(int i) { return u[i]; }
    double get_v(int i) { return v[i]; }
    double get_w(int i) { return w[i]; }
    void set_rho(int i, double a) { rho[i] = a; }
    void set_u(int i, double a) { u[i] = a; }
    void set_v(int i, double a) { v[i] = a; }
    void set_w(int i, double a) { w[i] = a; }
};
struct Boundary_Conditions {
    std::vector<std::string> name;
    std::vector<int> type_rho, type_vel;
    std::vector<double> rho, u, v, w;
    std::string get_name(int i) { return name[i]; }
    int get_rho_type(int i) { return type_rho[i]; }
    int get_vel_type(int i) { return type_vel[i]; }
    double get_rho(int i) { return rho[i]; }
    double get_u(int i) { return u[i]; }
    double get_v(int i) { return v[i]; }
    double get_w(int i) { return w[i]; }
};
struct global_variables {
    double tolerance = 0.0, pre_conditioned_gamma = 1, simulation_length = 2000, time_marching_step = 0.25;
    double max_velocity = 0.01, ref_rho = 1.0, visc = 1.0 / 6.0, fneq_min = 0.0;
    int output_step = 5, gpu_time_stepping = 2;
    std::string import_file, gradient_calc_type = "hybrid";
};
struct domain_geometry {
    double X, Y, Z, dx, dy, dz, origin_x = 0, origin_y = 0, origin_z = 0;
};
struct initial_conditions {
    double average_rho = 1.0, pressure_gradient = 0.0, force_x = 0, force_y = 0, force_z = 0;
};
struct lagrangian_object {
    std::string name, mesh_file, stress_free_mesh, curvature_mesh;
    int num_nodes = 0, num_springs = 0, num_tets = 0, type = 2;
    double centre_x = 0, centre_y = 0, centre_z = 0;
    double *node_x = nullptr, *node_y = nullptr, *node_z = nullptr;
    int *tet_connectivity = nullptr;
    bool sphere_reference = true, mz_importer = false;
    double volume_ratio = 1.0, volume_modulus = 0, local_area_modulus = 0, global_area_modulus = 0, shear_modulus = 0;
    double global_bending_modulus = 0, local_bending_modulus = 0, viscosity = 0, wlc_ratio = 2.2;
    double internal_viscosity_ratio = 1.0, curvature_ratio = 1.0;
};
}  // namespace ref

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "reference_entry_mesh";
    const int nx = 16, ny = 12, nz = 12, per[3] = {1, 0, 0};
    int np, nf, ni;
    if (lbfs_mesh_hex_counts(nx, ny, nz, per, &np, &nf, &ni) != LBFS_OK) return 2;
    std::vector<double> pts(3 * (size_t)np), shift(3 * (size_t)nf);
    std::vector<int> fv(4 * (size_t)nf), own(nf), nb(ni), patch(nf);
    const double L[3] = {(double)nx, (double)ny, (double)nz}, O[3] = {0, 0, 0};
    if (lbfs_mesh_hex_generate(nx, ny, nz, L, O, per, 0.1, 12, pts.data(), fv.data(), own.data(), nb.data(),
                               shift.data(), patch.data()) != LBFS_OK)
        return 2;
    const int nc = nx * ny * nz, nbc = nf - ni;
    // boundary faces come patch after patch (bottom, top, back, front walls)
    const char *names[4] = {"bottom", "top", "back", "front"}, *types[4] = {"wall", "wall", "wall", "wall"};
    int start[4], size[4];
    for (int k = 0; k < 4; k++) { start[k] = -1; size[k] = 0; }
    for (int f = ni; f < nf; f++) {
        const int k = patch[f] - 2;
        if (k < 0 || k > 3) return 2;
        if (start[k] < 0) start[k] = f;
        if (f != start[k] + size[k]) return 2;  // each patch's faces must be contiguous
        size[k]++;
    }
    mkdir(dir.c_str(), 0755);
    if (lbfs_polymesh_write(dir.c_str(), np, pts.data(), nc, nf, ni, fv.data(), own.data(), nb.data(), 4, names, types,
                            start, size, shift.data()) != LBFS_OK)
        return 2;

    ref::Boundary_Conditions bcs;
    for (int b = 0; b < nbc; b++) {
        bcs.name.push_back(names[patch[ni + b] - 2]);
        bcs.type_rho.push_back(LBFS_BC_NEUMANN);
        bcs.type_vel.push_back(LBFS_BC_DIRICHLET);
        bcs.rho.push_back(1.0); bcs.u.push_back(0.0); bcs.v.push_back(0.0); bcs.w.push_back(0.0);
    }
    ref::global_variables globals;
    globals.import_file = dir;
    globals.simulation_length = 12;
    globals.output_step = 5;
    globals.time_marching_step = 0.1;  // the reference GPU wall form is unstable at 0.25 with nu = 1/6
    ref::domain_geometry domain{(double)nx, (double)ny, (double)nz, 1.0, 1.0, 1.0};
    ref::initial_conditions init;
    init.pressure_gradient = 2e-5;
    init.force_x = init.pressure_gradient;  // pressure_direction "x" (preprocessor.cpp:722-726)
    ref::Solution soln(nc), residual(nc);
    for (int i = 0; i < nc; i++) {
        soln.u[i] = 0.004 * std::sin(0.3 * (i % nx)) * std::sin(0.5 * ((i / nx) % ny));
        soln.rho[i] = 1.0 + 1e-4 * std::cos(0.2 * i);
    }
    // one frequency-4 Evans-Fung cell (162 nodes), 7 cells across, its
    // stress-free shape 4 % larger than the initial one
    const int fq = 4, nn = 10 * fq * fq + 2, nt = 20 * fq * fq;
    std::vector<double> unit(3 * (size_t)nn), Xi(3 * (size_t)nn), Xs(3 * (size_t)nn);
    std::vector<int> tri(3 * (size_t)nt);
    if (lbfs_mesh_icosphere(fq, unit.data(), tri.data()) != LBFS_OK) return 2;
    if (lbfs_mesh_evans_fung(nn, unit.data(), 7.0, Xi.data()) != LBFS_OK) return 2;
    const double c[3] = {8.0, 6.0, 6.0};
    for (int i = 0; i < nn; i++)
        for (int a = 0; a < 3; a++) {
            Xs[3 * i + a] = c[a] + 1.04 * Xi[3 * i + a];
            Xi[3 * i + a] += c[a];
        }
    // the initial shape comes from the object's mesh file (Osay format)
    const std::string mesh_file = dir + "/cell.dat";
    if (lbfs_membrane_mesh_write(mesh_file.c_str(), LBFS_MESH_OSAY, nn, Xi.data(), nt, tri.data(), 0, nullptr) != LBFS_OK)
        return 2;
    std::vector<double> nx_(nn), ny_(nn), nz_(nn);
    for (int i = 0; i < nn; i++) { nx_[i] = Xs[3 * i]; ny_[i] = Xs[3 * i + 1]; nz_[i] = Xs[3 * i + 2]; }
    std::vector<ref::lagrangian_object> objects(1);
    ref::lagrangian_object &o = objects[0];
    o.mesh_file = mesh_file;
    o.num_nodes = nn; o.num_tets = nt; o.node_x = nx_.data(); o.node_y = ny_.data(); o.node_z = nz_.data();
    o.tet_connectivity = tri.data();
    o.shear_modulus = 5e-3; o.volume_modulus = 0.1; o.local_area_modulus = 5e-3; o.global_area_modulus = 0.1;
    o.global_bending_modulus = 1e-3; o.local_bending_modulus = 1e-3; o.viscosity = 0.02; o.volume_ratio = 0.95;
    ref::unstructured_mesh mesh;
    ref::external_forces source;
    ref::unstructured_bcs qbc;
    ref::post_processing pp;

    try {
        // 1. the reference's call, with its parameter list
        ref::Solution a = soln;
        int outputs = 0;
        lbfs_amd::step_entry_options opt;
        opt.output = [&](int, const lbfs_diag &d) { outputs += d.finite ? 1 : 0; };
        lbfs_amd::General_Purpose_Solver_mk_i(mesh, a, bcs, source, globals, domain, init, qbc, 0, residual, 0, pp,
                                              objects, opt);
        // 2. the same run through the ABI directly
        lbfs_mesh_desc m = {};
        std::vector<int> rt(nbc, LBFS_BC_NEUMANN), vt(nbc, LBFS_BC_DIRICHLET);
        std::vector<double> bv(4 * (size_t)nbc, 0.0);
        for (int b = 0; b < nbc; b++) bv[4 * (size_t)b] = 1.0;
        m.n_points = np; m.points = pts.data(); m.n_cells = nc; m.n_faces = nf; m.n_internal = ni;
        m.face_vertices = fv.data(); m.owner = own.data(); m.neighbour = nb.data(); m.face_shift = shift.data();
        m.bc_rho_type = rt.data(); m.bc_vel_type = vt.data(); m.bc_values = bv.data();
        m.nx = nx; m.ny = ny; m.nz = nz; m.spacing = 1.0;
        lbfs_params p = {};
        p.lattice = LBFS_D3Q15; p.flags = LBFS_FLAGS_REFERENCE_GPU; p.visc = 1.0 / 6.0; p.gamma = 1.0;
        p.dt = 0.1; p.rho0 = 1.0; p.time_stepping = 2; p.max_velocity = 0.01; p.channel_diameter = ny;
        p.body_force[0] = 2e-5;
        lbfs_ctx *ctx = nullptr;
        if (lbfs_create(&m, &p, &ctx) != LBFS_OK) {
            std::printf("reference_entry_check: %s\n", lbfs_last_error(nullptr));
            return 2;
        }
        lbfs_membrane_desc d = {};
        d.n_nodes = nn; d.n_tri = nt; d.nodes = Xi.data(); d.nodes_stress_free = Xs.data(); d.tri = tri.data();
        d.shear_modulus = o.shear_modulus; d.wlc_ratio = o.wlc_ratio; d.volume_modulus = o.volume_modulus;
        d.local_area_modulus = o.local_area_modulus; d.global_area_modulus = o.global_area_modulus;
        d.global_bending_modulus = o.global_bending_modulus; d.local_bending_modulus = o.local_bending_modulus;
        d.viscosity = o.viscosity; d.curvature_ratio = 1.0; d.volume_ratio = o.volume_ratio; d.sphere_reference = 1;
        d.mad_compat = 1; d.internal_viscosity_ratio = 1.0;
        int ok = lbfs_set_state(ctx, soln.rho.data(), soln.u.data(), soln.v.data(), soln.w.data()) == LBFS_OK &&
                 lbfs_add_membranes(ctx, &d, 1, LBFS_IBM_LINEAR2) == LBFS_OK && lbfs_step(ctx, 12) == LBFS_OK;
        std::vector<double> r(nc), u(nc), v(nc), w(nc);
        ok = ok && lbfs_get_state(ctx, r.data(), u.data(), v.data(), w.data()) == LBFS_OK;
        lbfs_destroy(ctx);
        if (!ok) return 2;
        const bool same = a.rho == r && a.u == u && a.v == v && a.w == w;
        double du = 0.0;
        for (int i = 0; i < nc; i++) du = std::fmax(du, std::fabs(a.u[i] - soln.u[i]));
        // 3. a tolerance above the first step's residual stops after one step
        ref::Solution b = soln;
        int last = -1;
        globals.tolerance = 1e6;
        opt.output = [&](int t, const lbfs_diag &) { last = t; };
        lbfs_amd::General_Purpose_Solver_mk_i(mesh, b, bcs, source, globals, domain, init, qbc, 0, residual, 0, pp,
                                              objects, opt);
        std::printf("reference_entry_check: outputs %d, |u change| %.3e, direct ABI run %s, stop after %d step(s)\n",
                    outputs, du, same ? "equal" : "DIFFERS", last);
        return same && outputs == 3 && du > 0.0 && last == 1 ? 0 : 1;
    } catch (const lbfs_amd::lbfs_failure &e) {
        std::printf("reference_entry_check: %s\n", e.what());
        return e.code == LBFS_E_NO_DEVICE ? 3 : 2;
    }
}
