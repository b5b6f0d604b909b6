// Membrane mesh files of the reference (lagrangian_object.cpp):
//   LBFS_MESH_OSAY  "N T", N lines x y z, T lines of three 1-based node ids
//                   (osay_mesh_read, 238-302)
//   LBFS_MESH_MZ    "N S T", N lines x y z, 3T 1-based node ids, 4S 1-based
//                   spring ids (a b c d), then per node its degree and
//                   1-based neighbour ids (mesh_read, 495-587)
// Positions are shifted by the object's centre as the readers do.
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "lbfs_abi.h"

namespace {

bool read_header(std::ifstream &f, int format, long &n, long &s, long &t)
{
    s = 0;
    if (format == LBFS_MESH_OSAY) return (bool)(f >> n >> t) && n > 0 && t > 0;
    if (format == LBFS_MESH_MZ) return (bool)(f >> n >> s >> t) && n > 0 && s >= 0 && t > 0;
    return false;
}

}  // namespace

extern "C" int lbfs_membrane_mesh_counts(const char *path, int format, int *n_nodes, int *n_tri, int *n_springs)
{
    if (!path || !n_nodes || !n_tri || !n_springs) return LBFS_E_INVALID_ARG;
    std::ifstream f(path);
    long n, s, t;
    if (!f || !read_header(f, format, n, s, t)) return LBFS_E_INVALID_ARG;
    *n_nodes = (int)n;
    *n_tri = (int)t;
    *n_springs = (int)s;
    return LBFS_OK;
}

extern "C" int lbfs_membrane_mesh_read(const char *path, int format, const double centre[3], double *nodes, int *tri,
                                       int *springs)
{
    if (!path || !nodes || !tri) return LBFS_E_INVALID_ARG;
    std::ifstream f(path);
    long n, s, t;
    if (!f || !read_header(f, format, n, s, t)) return LBFS_E_INVALID_ARG;
    const double c[3] = {centre ? centre[0] : 0.0, centre ? centre[1] : 0.0, centre ? centre[2] : 0.0};
    for (long i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            double v;
            if (!(f >> v)) return LBFS_E_INVALID_ARG;
            nodes[3 * i + k] = v + c[k];
        }
    for (long i = 0; i < 3 * t; i++) {
        long k;
        if (!(f >> k) || k < 1 || k > n) return LBFS_E_INVALID_ARG;
        tri[i] = (int)(k - 1);
    }
    if (format == LBFS_MESH_MZ) {
        for (long i = 0; i < 4 * s; i++) {
            long k;
            if (!(f >> k) || k < 1 || k > n) return LBFS_E_INVALID_ARG;
            if (springs) springs[i] = (int)(k - 1);
        }
        // per-node neighbour lists: read to check the file is whole (the
        // library derives its node incidences from the springs)
        for (long i = 0; i < n; i++) {
            long d;
            if (!(f >> d) || d < 0) return LBFS_E_INVALID_ARG;
            for (long j = 0; j < d; j++) {
                long k;
                if (!(f >> k) || k < 1 || k > n) return LBFS_E_INVALID_ARG;
            }
        }
    }
    return LBFS_OK;
}

extern "C" int lbfs_membrane_mesh_write(const char *path, int format, int n_nodes, const double *nodes, int n_tri,
                                        const int *tri, int n_springs, const int *springs)
{
    if (!path || !nodes || !tri || n_nodes <= 0 || n_tri <= 0) return LBFS_E_INVALID_ARG;
    if (format != LBFS_MESH_OSAY && format != LBFS_MESH_MZ) return LBFS_E_INVALID_ARG;
    if (format == LBFS_MESH_MZ && (n_springs <= 0 || !springs)) return LBFS_E_INVALID_ARG;
    std::ofstream f(path);
    if (!f) return LBFS_E_INVALID_ARG;
    f.precision(17);
    if (format == LBFS_MESH_OSAY) f << n_nodes << " " << n_tri << "\n";
    else f << n_nodes << " " << n_springs << " " << n_tri << "\n";
    for (int i = 0; i < n_nodes; i++) f << nodes[3 * i] << " " << nodes[3 * i + 1] << " " << nodes[3 * i + 2] << "\n";
    for (int i = 0; i < n_tri; i++) f << tri[3 * i] + 1 << " " << tri[3 * i + 1] + 1 << " " << tri[3 * i + 2] + 1 << "\n";
    if (format == LBFS_MESH_MZ) {
        std::vector<std::set<int>> nb(n_nodes);
        for (int i = 0; i < n_springs; i++) {
            const int *q = springs + 4 * i;
            f << q[0] + 1 << " " << q[1] + 1 << " " << q[2] + 1 << " " << q[3] + 1 << "\n";
            nb[q[0]].insert(q[2]);
            nb[q[2]].insert(q[0]);
        }
        for (int i = 0; i < n_nodes; i++) {
            f << nb[i].size();
            for (int k : nb[i]) f << " " << k + 1;
            f << "\n";
        }
    }
    return f ? LBFS_OK : LBFS_E_INVALID_ARG;
}
