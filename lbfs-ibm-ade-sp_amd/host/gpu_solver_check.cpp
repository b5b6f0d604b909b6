// Check of the C++ host mirror: a jittered channel with one biconcave cell,
// advanced by General_Purpose_Solver_mk_i and, in a second solver, group by
// group under the reference's names.  Both must agree bit for bit.
// Exit status: 0 agree, 1 disagree, 2 other failure, 3 no HIP device.
#include <cmath>
#include <cstdio>
#include <vector>

#include "gpu_solver.hpp"

using lbfs_amd::gpu_solver;

int main()
{
    const int nx = 24, ny = 16, nz = 16, per[3] = {1, 0, 0};
    int np, nf, ni;
    if (lbfs_mesh_hex_counts(nx, ny, nz, per, &np, &nf, &ni) != LBFS_OK) return 2;
    std::vector<double> pts(3 * (size_t)np), shift(3 * (size_t)nf);
    std::vector<int> fv(4 * (size_t)nf), own(nf), nb(ni), patch(nf);
    const double L[3] = {(double)nx, (double)ny, (double)nz}, O[3] = {0, 0, 0};
    if (lbfs_mesh_hex_generate(nx, ny, nz, L, O, per, 0.1, 11, pts.data(), fv.data(), own.data(), nb.data(),
                               shift.data(), patch.data()) != LBFS_OK)
        return 2;
    const int nbc = nf - ni, nc = nx * ny * nz;
    std::vector<int> rt(nbc, LBFS_BC_NEUMANN), vt(nbc, LBFS_BC_DIRICHLET);
    std::vector<double> bv(4 * (size_t)nbc, 0.0);
    for (int b = 0; b < nbc; b++) bv[4 * (size_t)b] = 1.0;
    lbfs_mesh_desc m = {};
    m.n_points = np; m.points = pts.data(); m.n_cells = nc; m.n_faces = nf; m.n_internal = ni;
    m.face_vertices = fv.data(); m.owner = own.data(); m.neighbour = nb.data(); m.face_shift = shift.data();
    m.bc_rho_type = rt.data(); m.bc_vel_type = vt.data(); m.bc_values = bv.data();
    m.nx = nx; m.ny = ny; m.nz = nz; m.spacing = 1.0;
    lbfs_params p = {};
    p.lattice = LBFS_D3Q19; p.visc = 1.0 / 6.0; p.gamma = 1.0; p.dt = 0.25; p.rho0 = 1.0;
    p.body_force[0] = 2e-5;

    // frequency-4 sphere mapped to the Evans-Fung shape, 8 cells across
    const int f = 4, nn = 10 * f * f + 2, nt = 20 * f * f;
    std::vector<double> unit(3 * (size_t)nn), X(3 * (size_t)nn);
    std::vector<int> tri(3 * (size_t)nt);
    if (lbfs_mesh_icosphere(f, unit.data(), tri.data()) != LBFS_OK) return 2;
    if (lbfs_mesh_evans_fung(nn, unit.data(), 8.0, X.data()) != LBFS_OK) return 2;
    for (int i = 0; i < nn; i++) { X[3 * i] += 12.0; X[3 * i + 1] += 8.0; X[3 * i + 2] += 8.0; }
    lbfs_membrane_desc md = {};
    md.n_nodes = nn; md.n_tri = nt; md.nodes = X.data(); md.tri = tri.data();
    md.shear_modulus = 5e-3; md.wlc_ratio = 2.2; md.volume_modulus = 0.1; md.local_area_modulus = 5e-3;
    md.global_area_modulus = 0.1; md.global_bending_modulus = 1e-3; md.local_bending_modulus = 1e-3;
    md.viscosity = 0.02; md.curvature_ratio = 1.0; md.volume_ratio = 0.95; md.internal_viscosity_ratio = 1.0;

    try {
        gpu_solver a(m, p), b(m, p);
        a.add_objects({md});
        b.add_objects({md});
        int outputs = 0;
        a.General_Purpose_Solver_mk_i(12, 4, [&](int, const lbfs_diag &d) { outputs += d.finite ? 1 : 0; }, 1.0, 0.01);
        for (int s = 0; s < 12; s++) b.step_by_groups();
        std::vector<double> r1, u1, v1, w1, r2, u2, v2, w2;
        a.get_solution(r1, u1, v1, w1);
        b.get_solution(r2, u2, v2, w2);
        const bool same = r1 == r2 && u1 == u2 && v1 == v2 && w1 == w2 && a.object_nodes() == b.object_nodes();
        double umax = 0.0;
        for (double x : u1) umax = std::fmax(umax, std::fabs(x));
        std::printf("gpu_solver_check: outputs %d, max |u| %.3e, groups %s whole steps\n", outputs, umax,
                    same ? "equal" : "DIFFER from");
        return same && outputs == 3 && umax > 0.0 ? 0 : 1;
    } catch (const lbfs_amd::lbfs_failure &e) {
        std::printf("gpu_solver_check: %s\n", e.what());
        return e.code == LBFS_E_NO_DEVICE ? 3 : 2;
    }
}
