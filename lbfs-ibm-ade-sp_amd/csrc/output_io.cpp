// Restart and output files of the reference, written and read without TECIO:
//   state files  <dir>/rho.txt, u.txt, v.txt, w.txt: one line "i ,value" per
//                cell in the caller's order (Solution::output, Solution.cpp:221-251;
//                read back by Solution::import, Solution.cpp:987-1064)
//   membrane     the FEPOINT triangle zone of
//                tecplot_output::spring_network_output_gpu (tecplot_output.cpp:844-915):
//                per node x y z, force, velocity, nodal area, index, curvature;
//                then the 1-based vertex ids of each triangle, one per line
// The reference writes with the stream's default format (6 significant
// digits); precision <= 0 reproduces that, a positive precision gives that
// many significant digits (17 round-trips a double).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "lbfs_abi.h"

namespace {

std::string fmt(double v, int precision)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.*g", precision > 0 ? precision : 6, v);
    return b;
}

const char *const STATE_FILES[4] = {"rho.txt", "u.txt", "v.txt", "w.txt"};

struct V3 {
    double x, y, z;
};
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

}  // namespace

extern "C" int lbfs_state_txt_write(const char *dir, int n_cells, const double *rho, const double *u, const double *v,
                                    const double *w, int precision)
{
    if (!dir || n_cells <= 0 || !rho || !u || !v || !w) return LBFS_E_INVALID_ARG;
    const double *f[4] = {rho, u, v, w};
    for (int k = 0; k < 4; k++) {
        std::ofstream out(std::string(dir) + "/" + STATE_FILES[k]);
        if (!out) return LBFS_E_INVALID_ARG;
        for (int i = 0; i < n_cells; i++) out << i << " ," << fmt(f[k][i], precision) << '\n';
        if (!out) return LBFS_E_INVALID_ARG;
    }
    return LBFS_OK;
}

extern "C" int lbfs_state_txt_read(const char *dir, int n_cells, double *rho, double *u, double *v, double *w)
{
    if (!dir || n_cells <= 0 || !rho || !u || !v || !w) return LBFS_E_INVALID_ARG;
    double *f[4] = {rho, u, v, w};
    for (int k = 0; k < 4; k++) {
        std::ifstream in(std::string(dir) + "/" + STATE_FILES[k]);
        if (!in) return LBFS_E_INVALID_ARG;
        std::string line;
        int i = 0;
        while (std::getline(in, line)) {
            // "i ,value": the value follows the comma (as import reads it)
            const size_t c = line.find(',');
            if (c == std::string::npos) {
                if (line.find_first_not_of(" \t\r") == std::string::npos) continue;  // blank line
                return LBFS_E_INVALID_ARG;
            }
            if (i >= n_cells) return LBFS_E_INVALID_ARG;
            std::istringstream vs(line.substr(c + 1));
            if (!(vs >> f[k][i])) return LBFS_E_INVALID_ARG;
            i++;
        }
        if (i != n_cells) return LBFS_E_INVALID_ARG;
    }
    return LBFS_OK;
}

extern "C" int lbfs_membrane_dat_write(const char *path, int n_nodes, const double *X, const double *F, const double *U,
                                       int n_tri, const int *tri, int time, int precision, double *nodal_area,
                                       double *curvature)
{
    if (!path || n_nodes <= 0 || n_tri <= 0 || !X || !F || !U || !tri) return LBFS_E_INVALID_ARG;
    for (int t = 0; t < 3 * n_tri; t++)
        if (tri[t] < 0 || tri[t] >= n_nodes) return LBFS_E_INVALID_ARG;
    auto P = [&](int i) { return V3{X[3 * (size_t)i], X[3 * (size_t)i + 1], X[3 * (size_t)i + 2]}; };
    // nodal area: a third of each incident triangle (node_area of spring_network.cu:1486);
    // curvature: 0.5 theta L summed over the node's springs (spring_network.cu:1532),
    // theta the angle between the normals of the spring's two triangles, positive
    // where the surface is convex (seen from the side the triangles' orientation
    // points to)
    std::vector<double> area(n_nodes, 0.0), curv(n_nodes, 0.0);
    std::vector<V3> N(n_tri), cen(n_tri);
    std::map<std::pair<int, int>, std::vector<int>> edge_tris;
    for (int t = 0; t < n_tri; t++) {
        const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        N[t] = cross(P(b) - P(a), P(c) - P(a));
        cen[t] = scale(P(a) + P(b) + P(c), 1.0 / 3.0);
        const double A = 0.5 * std::sqrt(dot(N[t], N[t]));
        area[a] += A / 3.0;
        area[b] += A / 3.0;
        area[c] += A / 3.0;
        const int v[3] = {a, b, c};
        for (int k = 0; k < 3; k++) {
            const int p = v[k], q = v[(k + 1) % 3];
            edge_tris[{std::min(p, q), std::max(p, q)}].push_back(t);
        }
    }
    for (const auto &e : edge_tris) {
        if (e.second.size() != 2) return LBFS_E_INVALID_ARG;  // not a closed surface
        const int t1 = e.second[0], t2 = e.second[1];
        const double n1 = std::sqrt(dot(N[t1], N[t1])), n2 = std::sqrt(dot(N[t2], N[t2]));
        double cs = dot(N[t1], N[t2]) / (n1 * n2);
        cs = std::fmin(1.0, std::fmax(-1.0, cs));
        const double dir = dot(N[t1], cen[t2] - cen[t1]);
        const double theta = std::acos(cs) * (double)((dir < 0.0) - (dir > 0.0));
        const V3 l = P(e.first.second) - P(e.first.first);
        const double L = std::sqrt(dot(l, l));
        curv[e.first.first] += 0.5 * theta * L;
        curv[e.first.second] += 0.5 * theta * L;
    }
    std::ofstream out(path);
    if (!out) return LBFS_E_INVALID_ARG;
    out << "VARIABLES = \"nodx\" \"nody\" \"nodz\" "
        << "\"i_force\" \"j_force\" \"k_force\" \"vel_x\" \"vel_y\" \"vel_z\" \"nodal_area\" \"index\" \"curvature\"\n";
    out << "ZONE Nodes= " << n_nodes << " Elements= " << n_tri << " F=FEPOINT ET=TRIANGLE \n";
    out << "SOLUTIONTIME = " << time << " \n";
    out << "DT=(DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE DOUBLE)\n";
    for (int j = 0; j < n_nodes; j++) {
        for (int k = 0; k < 3; k++) out << fmt(X[3 * (size_t)j + k], precision) << '\t';
        for (int k = 0; k < 3; k++) out << fmt(F[3 * (size_t)j + k], precision) << '\t';
        for (int k = 0; k < 3; k++) out << fmt(U[3 * (size_t)j + k], precision) << '\t';
        out << fmt(area[j], precision) << '\t' << j << '\t' << fmt(curv[j], precision) << "\t\n";
    }
    for (int t = 0; t < 3 * n_tri; t++) out << tri[t] + 1 << "\t\n";
    if (!out) return LBFS_E_INVALID_ARG;
    if (nodal_area)
        for (int j = 0; j < n_nodes; j++) nodal_area[j] = area[j];
    if (curvature)
        for (int j = 0; j < n_nodes; j++) curvature[j] = curv[j];
    return LBFS_OK;
}
