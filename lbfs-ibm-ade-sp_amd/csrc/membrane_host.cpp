// Host-side membrane set-up (FP64).
#include "membrane_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <array>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <unordered_set>

namespace lbfs {

namespace {
struct P3 {
    double x, y, z;
};
inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 mul(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline P3 at(const double *X, int i) { return {X[3 * i], X[3 * i + 1], X[3 * i + 2]}; }
}  // namespace

void icosphere(int n, std::vector<double> &nodes, std::vector<int> &tri)
{
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    const P3 v[12] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                      {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    const int f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    nodes.clear();
    tri.clear();
    // nodes shared between faces are merged on their position (rounded)
    std::map<std::tuple<long long, long long, long long>, int> index;
    auto node = [&](P3 q) -> int {
        const double r = std::sqrt(dot(q, q));
        q = mul(q, 1.0 / r);
        auto key = std::make_tuple((long long)std::llround(q.x * 1e9), (long long)std::llround(q.y * 1e9),
                                   (long long)std::llround(q.z * 1e9));
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        const int id = (int)(nodes.size() / 3);
        nodes.push_back(q.x);
        nodes.push_back(q.y);
        nodes.push_back(q.z);
        index.emplace(key, id);
        return id;
    };
    for (int t = 0; t < 20; t++) {
        P3 A = v[f[t][0]], B = v[f[t][1]], C = v[f[t][2]];
        if (dot(cross(sub(B, A), sub(C, A)), add(add(A, B), C)) < 0.0) std::swap(B, C);  // outward winding
        std::vector<int> id((size_t)(n + 1) * (n + 1), -1);
        auto P = [&](int i, int j) -> int {
            int &r = id[(size_t)i * (n + 1) + j];
            if (r < 0) r = node(add(A, add(mul(sub(B, A), (double)i / n), mul(sub(C, A), (double)j / n))));
            return r;
        };
        for (int i = 0; i < n; i++)
            for (int j = 0; i + j < n; j++) {
                tri.push_back(P(i, j)); tri.push_back(P(i + 1, j)); tri.push_back(P(i, j + 1));
                if (i + j + 1 < n) {
                    tri.push_back(P(i + 1, j)); tri.push_back(P(i + 1, j + 1)); tri.push_back(P(i, j + 1));
                }
            }
    }
}

void evans_fung(const std::vector<double> &unit, double d0, std::vector<double> &out)
{
    // z = +-D0 sqrt(1 - rho^2) (a0 + a1 rho^2 / 4 + a2 rho^4 / 16), rho = 2 r / D0
    const double a0 = 0.0518, a1 = 2.0026, a2 = -4.491;
    out.resize(unit.size());
    for (size_t i = 0; i < unit.size() / 3; i++) {
        const double x = unit[3 * i], y = unit[3 * i + 1], z = unit[3 * i + 2];
        const double r2 = std::min(1.0, x * x + y * y);
        const double h = d0 * std::sqrt(std::max(0.0, 1.0 - r2)) * (a0 + a1 * r2 / 4.0 + a2 * r2 * r2 / 16.0);
        out[3 * i] = 0.5 * d0 * x;
        out[3 * i + 1] = 0.5 * d0 * y;
        out[3 * i + 2] = z > 0.0 ? h : (z < 0.0 ? -h : 0.0);
    }
}

int build_springs(int n_nodes, const std::vector<int> &tri, std::vector<int> &springs, std::string &err)
{
    const int nt = (int)tri.size() / 3;
    // (next, prev) pair of every triangle around each node, in triangle order
    std::vector<std::vector<std::pair<int, int>>> pairs(n_nodes);
    for (int t = 0; t < nt; t++)
        for (int i = 0; i < 3; i++) {
            const int c = tri[3 * t + i];
            if (c < 0 || c >= n_nodes) { err = "triangle node out of range"; return LBFS_E_INVALID_ARG; }
            pairs[c].push_back({tri[3 * t + (i + 1) % 3], tri[3 * t + (i + 2) % 3]});
        }
    springs.clear();
    std::unordered_set<uint64_t> seen;
    std::vector<int> ring;
    for (int i = 0; i < n_nodes; i++) {
        const auto &pr = pairs[i];
        const int d = (int)pr.size();
        if (d < 3) { err = "node " + std::to_string(i) + " has fewer than 3 triangles"; return LBFS_E_INVALID_ARG; }
        // one-ring following the winding: (p1, p2), then the pair that starts with the last node
        ring.assign(1, pr[0].first);
        ring.push_back(pr[0].second);
        while ((int)ring.size() < d) {
            const int last = ring.back();
            int nxt = -1;
            for (const auto &q : pr)
                if (q.first == last) { nxt = q.second; break; }
            if (nxt < 0 || nxt == ring[0]) { err = "open or non-manifold one-ring at node " + std::to_string(i); return LBFS_E_INVALID_ARG; }
            ring.push_back(nxt);
        }
        for (int j = 0; j < d; j++) {
            int q[4] = {i, ring[j], ring[(j + 1) % d], ring[(j + 2) % d]};
            // a spring is its edge a-c (each edge of a closed surface once)
            const uint64_t lo = (uint64_t)std::min(q[0], q[2]), hi = (uint64_t)std::max(q[0], q[2]);
            if (!seen.insert(lo * (uint64_t)n_nodes + hi).second) continue;
            springs.insert(springs.end(), q, q + 4);
        }
    }
    if ((int)springs.size() / 4 * 2 != 3 * nt) { err = "spring count is not 3/2 of the triangle count"; return LBFS_E_INVALID_ARG; }
    return LBFS_OK;
}

int membrane_setup(const lbfs_membrane_desc &m, MembraneSetup &s, std::string &err)
{
    if (m.n_nodes < 4 || m.n_tri < 4 || !m.nodes || !m.tri) { err = "membrane description incomplete"; return LBFS_E_INVALID_ARG; }
    if (!(m.wlc_ratio > 1.0)) { err = "wlc_ratio must exceed 1"; return LBFS_E_INVALID_ARG; }
    s.n_nodes = m.n_nodes;
    s.n_tri = m.n_tri;
    s.tri.assign(m.tri, m.tri + 3 * (size_t)m.n_tri);
    int rc = LBFS_OK;
    if (m.n_springs > 0 && m.springs) {
        // springs given with the mesh (MZ files): each must be an edge a-c
        // shared by the triangles (a, b, c) and (a, c, d)
        if (2 * m.n_springs != 3 * m.n_tri) { err = "spring count must be 3/2 of the triangle count"; return LBFS_E_INVALID_ARG; }
        std::set<std::array<int, 3>> tris;
        for (int t = 0; t < m.n_tri; t++) {
            const int *q = &s.tri[3 * (size_t)t];
            for (int r = 0; r < 3; r++) tris.insert({q[r], q[(r + 1) % 3], q[(r + 2) % 3]});
        }
        s.springs.assign(m.springs, m.springs + 4 * (size_t)m.n_springs);
        for (int e = 0; e < m.n_springs; e++) {
            const int *q = &s.springs[4 * (size_t)e];
            for (int r = 0; r < 4; r++)
                if (q[r] < 0 || q[r] >= m.n_nodes) { err = "spring node out of range"; return LBFS_E_INVALID_ARG; }
            if (!tris.count({q[0], q[1], q[2]}) || !tris.count({q[0], q[2], q[3]})) {
                err = "spring " + std::to_string(e) + " is not an edge between triangles abc and acd";
                return LBFS_E_INVALID_ARG;
            }
        }
    } else {
        rc = build_springs(m.n_nodes, s.tri, s.springs, err);
        if (rc) return rc;
    }
    const int ne = (int)s.springs.size() / 4;
    s.n_edges = ne;
    s.X.assign(m.nodes, m.nodes + 3 * (size_t)m.n_nodes);
    const double *Xsf = m.nodes_stress_free ? m.nodes_stress_free : m.nodes;
    const double *Xc = m.nodes_curvature ? m.nodes_curvature : Xsf;
    const double *Xi = m.nodes;
    // contour lengths and spring constants (get_contour_lengths, get_spring_constant)
    const double r = m.wlc_ratio;
    const double cons = 0.5 * std::pow(1.0 - 1.0 / r, -3.0) + 1.0;
    if (m.spring_law != LBFS_SPRING_WLC && m.spring_law != LBFS_SPRING_WLC_POW) {
        err = "unknown spring law " + std::to_string(m.spring_law);
        return LBFS_E_INVALID_ARG;
    }
    s.contour_length.resize(ne);
    s.spring_constant.resize(ne);
    s.pow_ratio.assign(ne, 0.0);
    for (int e = 0; e < ne; e++) {
        const int *q = &s.springs[4 * (size_t)e];
        const P3 l = sub(at(Xsf, q[2]), at(Xsf, q[0]));
        const double L = std::sqrt(dot(l, l));
        s.spring_constant[e] = 4.0 * r * L * m.shear_modulus / (std::sqrt(3.0) * cons);
        s.contour_length[e] = L * r;
        if (m.spring_law == LBFS_SPRING_WLC_POW) {
            // get_spring_pow_constants (spring_network.cu:1655-1682) on the rest length
            const double x = 1.0 / r;
            const double alpha = std::sqrt(3.0) / L;
            const double beta = 0.5 * x / std::pow(1.0 - x, 3.0) - 0.25 / std::pow(1.0 - x, 2.0) + 0.25;
            const double gamma = 0.75 * std::sqrt(3.0) / (L * L * L);
            const double beta2 = 1.0 / std::pow(1.0 - x, 2.0) - 1.0 + 4.0 * x;
            s.spring_constant[e] = m.shear_modulus / (alpha * beta + gamma * L * L * beta2);
            s.pow_ratio[e] = -L * L * beta2;  // k_p = -L^2 k_s beta2: no force at the rest length
        }
    }
    // reference curvature (get_reference_curvatures)
    s.curvature_0.assign(ne, m.sphere_reference ? m.curvature_ratio / 3.265483301 : 0.0);
    (void)Xc;  // reference angles of the curvature shape are overwritten by the initial call below
    // initial call on the initial shape: totals, reference areas, per-edge state, MAD
    s.spring_area_0.resize(2 * (size_t)ne);
    s.spring_area.resize(2 * (size_t)ne);
    s.curvature.resize(ne);
    double V = 0.0, A = 0.0;
    for (int t = 0; t < m.n_tri; t++) {
        const P3 a = at(Xi, s.tri[3 * t]), b = at(Xi, s.tri[3 * t + 1]), c = at(Xi, s.tri[3 * t + 2]);
        const P3 N = cross(sub(b, a), sub(c, a));
        A += 0.5 * std::sqrt(dot(N, N));
        V += dot(N, mul(add(add(a, b), c), 1.0 / 3.0)) / 6.0;
    }
    s.V = V;
    s.A = A;
    s.V0 = V * m.volume_ratio;
    s.A0 = A;
    double mad = 0.0;
    const int n_mad = m.mad_compat ? std::min(m.n_nodes, ne) : ne;
    for (int e = 0; e < ne; e++) {
        const int *q = &s.springs[4 * (size_t)e];
        const P3 a = at(Xi, q[0]), b = at(Xi, q[1]), c = at(Xi, q[2]), d = at(Xi, q[3]);
        const P3 N1 = cross(sub(b, a), sub(c, a)), N2 = cross(sub(c, a), sub(d, a));
        const double n1 = std::sqrt(dot(N1, N1)), n2 = std::sqrt(dot(N2, N2));
        const double cs = std::max(-1.0, std::min(1.0, dot(N1, N2) / (n1 * n2)));
        const P3 c1 = mul(add(add(a, b), c), 1.0 / 3.0), c2 = mul(add(add(a, c), d), 1.0 / 3.0);
        const double dir = dot(N1, sub(c2, c1));
        const double ang = std::acos(cs) * (double)((dir < 0) - (dir > 0));
        const P3 l = sub(c, a);
        const double L = std::sqrt(dot(l, l));
        s.spring_area_0[2 * (size_t)e] = n1;
        s.spring_area_0[2 * (size_t)e + 1] = n2;
        s.spring_area[2 * (size_t)e] = n1;
        s.spring_area[2 * (size_t)e + 1] = n2;
        s.curvature[e] = ang * L / ((n1 + n2) / 6.0);
        if (e < n_mad) mad += s.curvature[e] * (n1 + n2) / 3.0;
    }
    s.MAD = mad / A;
    s.MAD0 = s.MAD;
    return LBFS_OK;
}

}  // namespace lbfs
