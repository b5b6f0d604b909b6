// Host-side mesh set-up (FP64): generated hex boxes, geometry, LS weights and
// the tile-blocked device layout.
#include "mesh_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace lbfs {

namespace {

// Runs body(begin, end) over [0, n) in contiguous ranges of at least
// `grain` items on up to 16 host threads (LBFS_HOST_THREADS overrides the
// count), the calling thread taking the first range.  The ranges must not
// write to the same elements; the result then does not depend on the split.
template <class F>
void parallel_for(int n, F body, int grain = 1024)
{
    int nth = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
    if (const char *e = std::getenv("LBFS_HOST_THREADS")) nth = std::max(1, std::atoi(e));
    nth = std::min(nth, std::max(1, n / std::max(1, grain)));
    if (nth <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (n + nth - 1) / nth;
    for (int t = 1; t < nth; t++) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back([&body, b, e]() { body(b, e); });
    }
    body(0, std::min(n, chunk));
    for (std::thread &th : pool) th.join();
}

struct V3 {
    double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 load3(const double *p) { return {p[0], p[1], p[2]}; }
inline void store3(double *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

}  // namespace

// ---------------------------------------------------------------------------
// generated hex boxes (face order and vertex winding of unstructured_mesh.cpp:
// 405-664: x, y, z internal faces, then left/right/bottom/top/back/front)
// ---------------------------------------------------------------------------
HexCounts hex_counts(int nx, int ny, int nz, const int periodic[3])
{
    HexCounts c;
    c.n_points = (nx + 1) * (ny + 1) * (nz + 1);
    int ix = (nx - 1 + (periodic[0] ? 1 : 0)) * ny * nz;
    int iy = (ny - 1 + (periodic[1] ? 1 : 0)) * nx * nz;
    int iz = (nz - 1 + (periodic[2] ? 1 : 0)) * nx * ny;
    int bx = periodic[0] ? 0 : 2 * ny * nz;
    int by = periodic[1] ? 0 : 2 * nx * nz;
    int bz = periodic[2] ? 0 : 2 * nx * ny;
    c.n_internal = ix + iy + iz;
    c.n_faces = c.n_internal + bx + by + bz;
    return c;
}

void hex_generate(int nx, int ny, int nz, const double L[3], const double origin[3], const int periodic[3],
                  double jitter, uint64_t seed, double *points, int *fv, int *owner, int *neighbour,
                  double *shift, int *patch)
{
    const int n[3] = {nx, ny, nz};
    const double h[3] = {L[0] / nx, L[1] / ny, L[2] / nz};
    const int vx = nx + 1, vy = ny + 1;
    auto V = [&](int i, int j, int k) { return i + j * vx + k * vx * vy; };
    auto C = [&](int i, int j, int k) { return i + j * nx + k * nx * ny; };
    // vertex positions: jitter drawn per logical vertex (periodic images share
    // it), quantised to 2^-16 of the spacing, wall vertices keep their normal
    // coordinate, stored at FP32 precision like the reference's float vertex
    // arrays (unstructured_mesh.cpp:284-289).
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++) {
                int idx[3] = {i, j, k};
                int key[3];
                for (int a = 0; a < 3; a++) key[a] = (periodic[a] && idx[a] == n[a]) ? 0 : idx[a];
                uint64_t lin = (uint64_t)key[0] + (uint64_t)(nx + 1) * ((uint64_t)key[1] + (uint64_t)(ny + 1) * key[2]);
                SplitMix64 rng(seed * 0x100000001B3ull + lin * 0x9E3779B97F4A7C15ull + 1);
                for (int a = 0; a < 3; a++) {
                    double r = (2.0 * rng.uniform() - 1.0) * jitter;
                    bool wall = !periodic[a] && (idx[a] == 0 || idx[a] == n[a]);
                    // a single-cell axis is the extrusion direction of a 2D mesh: no jitter
                    if (wall || n[a] == 1) r = 0.0;
                    r = std::nearbyint(r * 65536.0) / 65536.0;
                    double x = origin[a] + (idx[a] + r) * h[a];
                    points[3 * V(i, j, k) + a] = (double)(float)x;
                }
            }
    int f = 0;
    auto put = [&](int a, int b, int c, int d, int o, int nb, double sx, double sy, double sz, int p) {
        fv[4 * f] = a; fv[4 * f + 1] = b; fv[4 * f + 2] = c; fv[4 * f + 3] = d;
        owner[f] = o;
        if (nb >= 0) neighbour[f] = nb;
        shift[3 * f] = sx; shift[3 * f + 1] = sy; shift[3 * f + 2] = sz;
        if (p >= 0) patch[f] = p;
        f++;
    };
    // internal x faces (+ periodic wraps at i = nx)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 1; i < nx; i++)
                put(V(i, j, k), V(i, j + 1, k), V(i, j + 1, k + 1), V(i, j, k + 1), C(i - 1, j, k), C(i, j, k), 0, 0, 0, -1);
    if (periodic[0])
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                put(V(nx, j, k), V(nx, j + 1, k), V(nx, j + 1, k + 1), V(nx, j, k + 1), C(nx - 1, j, k), C(0, j, k), L[0], 0, 0, -1);
    for (int k = 0; k < nz; k++)
        for (int j = 1; j < ny; j++)
            for (int i = 0; i < nx; i++)
                put(V(i, j, k), V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j, k), C(i, j - 1, k), C(i, j, k), 0, 0, 0, -1);
    if (periodic[1])
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                put(V(i, ny, k), V(i, ny, k + 1), V(i + 1, ny, k + 1), V(i + 1, ny, k), C(i, ny - 1, k), C(i, 0, k), 0, L[1], 0, -1);
    for (int k = 1; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                put(V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k), C(i, j, k - 1), C(i, j, k), 0, 0, 0, -1);
    if (periodic[2])
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                put(V(i, j, nz), V(i + 1, j, nz), V(i + 1, j + 1, nz), V(i, j + 1, nz), C(i, j, nz - 1), C(i, j, 0), 0, 0, L[2], -1);
    // boundary patches
    if (!periodic[0]) {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                put(V(0, j, k), V(0, j, k + 1), V(0, j + 1, k + 1), V(0, j + 1, k), C(0, j, k), -1, 0, 0, 0, 0);
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                put(V(nx, j, k), V(nx, j + 1, k), V(nx, j + 1, k + 1), V(nx, j, k + 1), C(nx - 1, j, k), -1, 0, 0, 0, 1);
    }
    if (!periodic[1]) {
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                put(V(i, 0, k), V(i + 1, 0, k), V(i + 1, 0, k + 1), V(i, 0, k + 1), C(i, 0, k), -1, 0, 0, 0, 2);
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                put(V(i, ny, k), V(i, ny, k + 1), V(i + 1, ny, k + 1), V(i + 1, ny, k), C(i, ny - 1, k), -1, 0, 0, 0, 3);
    }
    if (!periodic[2]) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                put(V(i, j, 0), V(i, j + 1, 0), V(i + 1, j + 1, 0), V(i + 1, j, 0), C(i, j, 0), -1, 0, 0, 0, 4);
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                put(V(i, j, nz), V(i + 1, j, nz), V(i + 1, j + 1, nz), V(i, j + 1, nz), C(i, j, nz - 1), -1, 0, 0, 0, 5);
    }
}

// ---------------------------------------------------------------------------
// geometry: fan triangulation from vertex 0 for area / centroid / normal
// (unstructured_mesh.cpp:1170-1204), origin tetrahedra for volume
// (unstructured_mesh.cpp:1015-1045), area-weighted face-centroid mean for the
// cell centroid and the stencil in face order (unstructured_mesh.cpp:1592-1692),
// face dt min(0.5 sqrt(A), 3 nu / (gamma fneq_min)) (unstructured_mesh.cpp:
// 1435-1443), ghost centroid halfway to the face (unstructured_mesh.cpp:1963-1996).
// Periodic wrap faces are seen by the neighbour displaced by -shift.
// ---------------------------------------------------------------------------
int compute_geometry(const lbfs_mesh_desc &m, double visc, double gamma, double fneq_min, bool closed, Geometry &g,
                     std::string &err)
{
    const int nc = m.n_cells, nf = m.n_faces, ni = m.n_internal, nb = nf - ni;
    if (nc <= 0 || nf <= 0 || ni < 0 || ni > nf || !m.points || !m.face_vertices || !m.owner ||
        (ni > 0 && !m.neighbour)) {
        err = "mesh description incomplete";
        return LBFS_E_INVALID_ARG;
    }
    g.n_cells = nc; g.n_faces = nf; g.n_int = ni; g.n_bc = nb;
    g.owner.assign(m.owner, m.owner + nf);
    g.nbr.resize(nf);
    g.shift.assign(3 * (size_t)nf, 0.0);
    if (m.face_shift) std::copy(m.face_shift, m.face_shift + 3 * (size_t)nf, g.shift.begin());
    g.area.assign(nf, 0.0); g.normal.assign(3 * (size_t)nf, 0.0); g.fc.assign(3 * (size_t)nf, 0.0);
    g.fdt.assign(nf, 0.0);
    g.cc.assign(3 * (size_t)(nc + nb), 0.0);
    g.vol.assign(nc, 0.0);
    g.gcell.assign(6 * (size_t)nc, -1); g.gface.assign(6 * (size_t)nc, -1); g.grole.assign(6 * (size_t)nc, 0);
    std::vector<double> tot_area(nc, 0.0);
    std::vector<int> cnt(nc, 0);
    auto push = [&](int cell, int other, int face, int role) -> bool {
        if (cnt[cell] >= 6) return false;
        g.gcell[6 * (size_t)cell + cnt[cell]] = other;
        g.gface[6 * (size_t)cell + cnt[cell]] = face;
        g.grole[6 * (size_t)cell + cnt[cell]] = role;
        cnt[cell]++;
        return true;
    };
    for (int f = 0; f < nf; f++) {
        const int o = m.owner[f];
        const int n = f < ni ? m.neighbour[f] : nc + f - ni;
        if (o < 0 || o >= nc || (f < ni && (n < 0 || n >= nc))) {
            err = "face " + std::to_string(f) + " has an invalid owner/neighbour";
            return LBFS_E_INVALID_ARG;
        }
        g.nbr[f] = n;
        const V3 s = load3(&g.shift[3 * (size_t)f]);
        V3 P[4];
        for (int t = 0; t < 4; t++) P[t] = load3(m.points + 3 * (size_t)m.face_vertices[4 * (size_t)f + t]);
        V3 csum{0, 0, 0}, nsum{0, 0, 0};
        double asum = 0.0;
        for (int t = 0; t < 2; t++) {
            const V3 a = P[0], b = P[t + 1], c = P[t + 2];
            const V3 N = cross(b - a, c - a);
            const double tv = dot(a, N) / 6.0;
            g.vol[o] += tv;
            if (f < ni) g.vol[n] -= tv - dot(s, N) / 6.0;
            const double ta = 0.5 * norm(N);
            const V3 cen = (1.0 / 3.0) * (a + b + c);
            csum = csum + ta * cen;
            nsum = nsum + N;
            asum += ta;
        }
        // closed: |sum N| / 2, the exact area vector length, so every cell
        // closes; otherwise the reference's sum of triangle areas
        g.area[f] = closed ? 0.5 * norm(nsum) : asum;
        store3(&g.fc[3 * (size_t)f], (1.0 / asum) * csum);
        store3(&g.normal[3 * (size_t)f], (1.0 / norm(nsum)) * nsum);
        double fd = 0.5 * std::sqrt(asum);
        if (fneq_min > 0.0) fd = std::min(fd, visc * 3.0 / gamma / fneq_min);
        g.fdt[f] = fd;
        tot_area[o] += asum;
        store3(&g.cc[3 * (size_t)o], load3(&g.cc[3 * (size_t)o]) + csum);
        if (!push(o, n, f, 0)) { err = "cell " + std::to_string(o) + " has more than 6 faces"; return LBFS_E_UNSUPPORTED; }
        if (f < ni) {
            tot_area[n] += asum;
            store3(&g.cc[3 * (size_t)n], load3(&g.cc[3 * (size_t)n]) + csum - asum * s);
            if (!push(n, o, f, 1)) { err = "cell " + std::to_string(n) + " has more than 6 faces"; return LBFS_E_UNSUPPORTED; }
        }
    }
    for (int i = 0; i < nc; i++) {
        if (cnt[i] != 6) {
            err = "cell " + std::to_string(i) + " is not a hexahedron (" + std::to_string(cnt[i]) + " faces)";
            return LBFS_E_UNSUPPORTED;
        }
        store3(&g.cc[3 * (size_t)i], (1.0 / tot_area[i]) * load3(&g.cc[3 * (size_t)i]));
        if (!(g.vol[i] > 0.0)) { err = "cell " + std::to_string(i) + " has non-positive volume"; return LBFS_E_INVALID_ARG; }
    }
    for (int f = ni; f < nf; f++) {
        const int gh = nc + f - ni, o = m.owner[f];
        const V3 co = load3(&g.cc[3 * (size_t)o]), xf = load3(&g.fc[3 * (size_t)f]);
        store3(&g.cc[3 * (size_t)gh], co + 0.5 * (xf - co));
    }
    return LBFS_OK;
}

// ---------------------------------------------------------------------------
// least-squares / Green-Gauss hybrid weights (gradients.cpp:149-210, 214-313)
// ---------------------------------------------------------------------------
int compute_ls(const Geometry &g, int grad_type, bool det_compat, std::vector<double> &lhs,
               std::vector<double> &rhs, std::string &err)
{
    const int nc = g.n_cells;
    lhs.assign(9 * (size_t)nc, 0.0);
    rhs.assign(18 * (size_t)nc, 0.0);
    for (int i = 0; i < nc; i++) {
        V3 d[6], dxf[6];
        const V3 ci = load3(&g.cc[3 * (size_t)i]);
        double max_area = 0.0, max_disp = 0.0;
        for (int k = 0; k < 6; k++) {
            const int i1 = g.gcell[6 * (size_t)i + k], f = g.gface[6 * (size_t)i + k];
            const V3 s = load3(&g.shift[3 * (size_t)f]);
            const V3 xf = load3(&g.fc[3 * (size_t)f]);
            const V3 c1 = load3(&g.cc[3 * (size_t)i1]);
            if (g.grole[6 * (size_t)i + k] == 0) { d[k] = c1 + s - ci; dxf[k] = xf - ci; }
            else { d[k] = c1 - s - ci; dxf[k] = xf - s - ci; }
            max_area = std::max(max_area, g.area[f]);
            max_disp = std::max(max_disp, norm(d[k]));
        }
        const double aspect = 2.0 * std::fabs(max_disp) * max_area / g.vol[i];
        double beta = grad_type == LBFS_GRAD_HYBRID ? std::min(1.0, 2.0 / aspect) : (grad_type == LBFS_GRAD_LS ? 1.0 : 0.0);
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        M[0] = M[4] = M[8] = 2.0 * (1.0 - beta) * g.vol[i];
        for (int k = 0; k < 6; k++) {
            const int f = g.gface[6 * (size_t)i + k];
            const V3 n = load3(&g.normal[3 * (size_t)f]);
            const double sgn = g.grole[6 * (size_t)i + k] == 0 ? 1.0 : -1.0;
            const double ln = dot(dxf[k], n), Ln = dot(d[k], n), Lc = norm(d[k]);
            const double r = 2.0 * ln / Ln;
            const double w = r * r * g.area[f] / Lc;
            const double dv[3] = {d[k].x, d[k].y, d[k].z}, nv[3] = {n.x, n.y, n.z};
            for (int p = 0; p < 3; p++) {
                for (int q = 0; q < 3; q++) M[3 * p + q] += w * dv[p] * dv[q] * beta;
                rhs[18 * (size_t)i + 3 * k + p] = beta * w * dv[p] + (1.0 - beta) * nv[p] * sgn * g.area[f];
            }
        }
        const double a = M[0], b = M[1], c = M[2], dd = M[3], e = M[4], ff = M[5], gg = M[6], h = M[7], kk = M[8];
        const double det = det_compat ? a * (e * kk - ff * h) + b * (dd * kk - ff - gg) + c * (dd * h - e * gg)
                                      : a * (e * kk - ff * h) - b * (dd * kk - ff * gg) + c * (dd * h - e * gg);
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
            err = "singular least-squares matrix at cell " + std::to_string(i);
            return LBFS_E_INVALID_ARG;
        }
        double *L = &lhs[9 * (size_t)i];
        L[0] = (e * kk - ff * h) / det; L[1] = (c * h - b * kk) / det; L[2] = (b * ff - c * e) / det;
        L[3] = (ff * gg - dd * kk) / det; L[4] = (a * kk - c * gg) / det; L[5] = (c * dd - a * ff) / det;
        L[6] = (dd * h - e * gg) / det; L[7] = (b * gg - a * h) / det; L[8] = (a * e - b * dd) / det;
    }
    return LBFS_OK;
}

// ---------------------------------------------------------------------------
// tile-blocked device layout
// ---------------------------------------------------------------------------
int mesh_blocks(const lbfs_mesh_desc &m, const double body_force[3], std::vector<HostBlock> &blocks,
                std::string &err)
{
    blocks.clear();
    if (m.block_body_force && m.n_blocks <= 1) {
        err = "block_body_force needs n_blocks > 1";
        return LBFS_E_INVALID_ARG;
    }
    if (m.n_blocks > 1) {
        if (!m.block_dims || !m.block_origin || !m.block_spacing) { err = "block arrays missing"; return LBFS_E_INVALID_ARG; }
        long long c0 = 0;
        for (int b = 0; b < m.n_blocks; b++) {
            HostBlock h;
            for (int t = 0; t < 3; t++) {
                h.n[t] = m.block_dims[3 * b + t];
                h.origin[t] = m.block_origin[3 * b + t];
            }
            h.spacing = m.block_spacing[b];
            for (int t = 0; t < 3; t++) {
                h.force[t] = m.block_body_force ? m.block_body_force[3 * b + t] : body_force[t];
                if (!std::isfinite(h.force[t])) {
                    err = "block " + std::to_string(b) + ": non-finite body force";
                    return LBFS_E_INVALID_ARG;
                }
            }
            if (h.n[0] < 1 || h.n[1] < 1 || h.n[2] < 1 || !(h.spacing > 0.0)) {
                err = "block " + std::to_string(b) + ": dims and spacing must be positive";
                return LBFS_E_INVALID_ARG;
            }
            h.cell0 = (int)c0;
            c0 += (long long)h.n[0] * h.n[1] * h.n[2];
            blocks.push_back(h);
        }
        if (c0 != m.n_cells) { err = "block cells do not add up to n_cells"; return LBFS_E_INVALID_ARG; }
        return LBFS_OK;
    }
    if (m.nx > 0 && m.ny > 0 && m.nz > 0 && (long long)m.nx * m.ny * m.nz == m.n_cells) {
        HostBlock h;
        h.n[0] = m.nx; h.n[1] = m.ny; h.n[2] = m.nz;
        for (int t = 0; t < 3; t++) h.origin[t] = m.origin[t];
        h.spacing = m.spacing;
        h.cell0 = 0;
        for (int t = 0; t < 3; t++) h.force[t] = body_force[t];
        blocks.push_back(h);
    }
    return LBFS_OK;
}

const double *cell_body_force(const std::vector<HostBlock> &blocks, int c, const double *dflt)
{
    // boxes are numbered in cell order: the last box starting at or before c
    int lo = 0, hi = (int)blocks.size() - 1, b = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        if (blocks[mid].cell0 <= c) { b = mid; lo = mid + 1; } else hi = mid - 1;
    }
    if (b < 0) return dflt;
    const HostBlock &h = blocks[b];
    return c < h.cell0 + h.n[0] * h.n[1] * h.n[2] ? h.force : dflt;
}

int build_layout(const lbfs_mesh_desc &m, const Geometry &g, const std::vector<double> &lhs,
                 const std::vector<double> &rhs, double rho0, const double body_force[3], int dims, const int tile[3],
                 const std::vector<HostBlock> &blocks, DeviceLayout &d, std::string &err)
{
    const int nc = g.n_cells;
    d.n_cells = nc;
    d.n_bc = g.n_bc;
    d.perm.clear();
    d.tile_cell_start.assign(1, 0);
    // cell order: TX x TY x TZ blocks of the logical grid when the mesh has
    // one, else consecutive chunks of TX TY TZ caller ids
    const int TX = tile[0], TY = tile[1], TZ = tile[2], TN = TX * TY * TZ;
    for (int a = 0; a < 3; a++) d.tile[a] = tile[a];
    if (!blocks.empty()) {
        for (const HostBlock &b : blocks) {
            const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
            for (int tz = 0; tz < nz; tz += TZ)
                for (int ty = 0; ty < ny; ty += TY)
                    for (int tx = 0; tx < nx; tx += TX) {
                        for (int k = tz; k < std::min(tz + TZ, nz); k++)
                            for (int j = ty; j < std::min(ty + TY, ny); j++)
                                for (int i = tx; i < std::min(tx + TX, nx); i++)
                                    d.perm.push_back(b.cell0 + i + j * nx + k * nx * ny);
                        d.tile_cell_start.push_back((int)d.perm.size());
                    }
        }
    } else {
        for (int i = 0; i < nc; i++) {
            d.perm.push_back(i);
            if ((i + 1) % TN == 0 || i + 1 == nc) d.tile_cell_start.push_back(i + 1);
        }
    }
    d.n_tiles = (int)d.tile_cell_start.size() - 1;
    d.inv.assign(nc, -1);
    parallel_for(nc, [&](int i0, int i1) {
        for (int i = i0; i < i1; i++) d.inv[d.perm[i]] = i;
    });
    // per-cell arrays
    d.stencil.assign(6 * (size_t)nc, 0);
    d.gw.assign(18 * (size_t)nc, 0.0f);
    d.inv_vol.assign(nc, 0.0f);
    d.cell_slots.assign(6 * (size_t)nc, 0);
    parallel_for(nc, [&](int i0, int i1) {
        for (int i = i0; i < i1; i++) {
            const int c = d.perm[i];
            d.inv_vol[i] = (float)(1.0 / g.vol[c]);
            const double *L = &lhs[9 * (size_t)c];
            for (int k = 0; k < 6; k++) {
                const int nb = g.gcell[6 * (size_t)c + k];
                d.stencil[(size_t)k * nc + i] = nb < nc ? d.inv[nb] : -1 - (nb - nc);
                const double *R = &rhs[18 * (size_t)c + 3 * k];
                for (int t = 0; t < 3; t++)
                    d.gw[(size_t)(k * 3 + t) * nc + i] =
                        (float)(L[3 * t] * R[0] + L[3 * t + 1] * R[1] + L[3 * t + 2] * R[2]);
            }
        }
    });
    // Tile face lists: every face touching a cell of the tile, in caller face
    // order; and the tile's halo, the cells across its border in the order
    // its faces first reach them.  Tiles are independent: first each tile's
    // lists (on worker threads), then their offsets, then the records.
    const int nt = d.n_tiles;
    std::vector<std::vector<int>> tfaces(nt), thalo(nt);
    std::string terr;
    std::mutex terr_lock;
    parallel_for(nt, [&](int t0, int t1) {
        std::unordered_map<int, int> slot;
        for (int t = t0; t < t1; t++) {
            std::vector<int> &faces = tfaces[t];
            for (int i = d.tile_cell_start[t]; i < d.tile_cell_start[t + 1]; i++)
                for (int k = 0; k < 6; k++) faces.push_back(g.gface[6 * (size_t)d.perm[i] + k]);
            std::sort(faces.begin(), faces.end());
            faces.erase(std::unique(faces.begin(), faces.end()), faces.end());
            const int c0 = d.tile_cell_start[t], c1 = d.tile_cell_start[t + 1], ntc = c1 - c0;
            slot.clear();
            std::vector<int> &halo = thalo[t];
            auto visit = [&](int ic) {
                if (ic >= c0 && ic < c1) return;
                if (slot.emplace(ic, ntc + (int)halo.size()).second) halo.push_back(ic);
            };
            for (int f : faces) {
                visit(d.inv[g.owner[f]]);
                if (g.nbr[f] < nc) visit(d.inv[g.nbr[f]]);
            }
            if ((int)faces.size() > 32767 || ntc + (int)halo.size() >= 0xffff) {
                std::lock_guard<std::mutex> lk(terr_lock);
                if (terr.empty())
                    terr = (int)faces.size() > 32767 ? "tile has too many faces" : "tile and its halo exceed 65534 cells";
            }
        }
    }, 4);
    if (!terr.empty()) { err = terr; return LBFS_E_UNSUPPORTED; }
    d.tile_face_start.assign(nt + 1, 0);
    d.tile_halo_start.assign(nt + 1, 0);
    d.tile_bface_start.assign(nt, 0);
    d.max_tile_faces = 0;
    d.max_grad_slots = 0;
    for (int t = 0; t < nt; t++) {
        const int nft = (int)tfaces[t].size(), ntc = d.tile_cell_start[t + 1] - d.tile_cell_start[t];
        d.tile_face_start[t + 1] = d.tile_face_start[t] + nft;
        d.tile_halo_start[t + 1] = d.tile_halo_start[t] + (int)thalo[t].size();
        // caller order puts internal faces (ids < n_int) before boundary faces
        d.tile_bface_start[t] = d.tile_face_start[t] +
            (int)(std::lower_bound(tfaces[t].begin(), tfaces[t].end(), g.n_int) - tfaces[t].begin());
        d.max_tile_faces = std::max(d.max_tile_faces, nft);
        d.max_grad_slots = std::max(d.max_grad_slots, ntc + (int)thalo[t].size());
    }
    const size_t ntf = (size_t)d.tile_face_start[nt];
    d.face_gslot.assign(ntf, 0u);
    d.halo_ids.assign((size_t)d.tile_halo_start[nt], 0);
    d.fA.assign(4 * ntf, 0.0f); d.fB.assign(4 * ntf, 0.0f); d.fC.assign(4 * ntf, 0.0f);
    d.face_ref.assign(ntf, 0);
    parallel_for(nt, [&](int t0, int t1) {
        std::unordered_map<int, int> slot;
        for (int t = t0; t < t1; t++) {
            const std::vector<int> &faces = tfaces[t], &halo = thalo[t];
            const int base = d.tile_face_start[t], c0 = d.tile_cell_start[t], c1 = d.tile_cell_start[t + 1];
            const int ntc = c1 - c0;
            slot.clear();
            for (size_t h = 0; h < halo.size(); h++) slot.emplace(halo[h], ntc + (int)h);
            std::copy(halo.begin(), halo.end(), d.halo_ids.begin() + d.tile_halo_start[t]);
            auto gslot = [&](int ic) -> uint32_t {
                if (ic >= c0 && ic < c1) return (uint32_t)(ic - c0);
                return (uint32_t)slot.at(ic);
            };
            for (size_t q = 0; q < faces.size(); q++) {
                const int f = faces[q];
                const int o = g.owner[f], n = g.nbr[f];
                const uint32_t so = gslot(d.inv[o]), sn = n < nc ? gslot(d.inv[n]) : 0xffffu;
                const size_t r = (size_t)base + q;
                d.face_gslot[r] = so | (sn << 16);
                const V3 xf = load3(&g.fc[3 * (size_t)f]);
                const V3 d1 = load3(&g.cc[3 * (size_t)o]) - xf;
                const V3 d2 = load3(&g.cc[3 * (size_t)n]) + load3(&g.shift[3 * (size_t)f]) - xf;
                const double A = g.area[f];
                const int on[2] = {d.inv[o], n < nc ? d.inv[n] : n};
                float ob, nbb;
                std::memcpy(&ob, &on[0], 4);
                std::memcpy(&nbb, &on[1], 4);
                const float a4[4] = {(float)(g.normal[3 * (size_t)f] * A), (float)(g.normal[3 * (size_t)f + 1] * A),
                                     (float)(g.normal[3 * (size_t)f + 2] * A), (float)g.fdt[f]};
                const float b4[4] = {(float)d1.x, (float)d1.y, (float)d1.z, ob};
                const float c4[4] = {(float)d2.x, (float)d2.y, (float)d2.z, nbb};
                std::copy(a4, a4 + 4, d.fA.begin() + 4 * r);
                std::copy(b4, b4 + 4, d.fB.begin() + 4 * r);
                std::copy(c4, c4 + 4, d.fC.begin() + 4 * r);
                d.face_ref[r] = f;
            }
            for (int i = c0; i < c1; i++) {
                const int c = d.perm[i];
                for (int k = 0; k < 6; k++) {
                    const int f = g.gface[6 * (size_t)c + k];
                    const int sl = (int)(std::lower_bound(faces.begin(), faces.end(), f) - faces.begin());
                    const int own = g.grole[6 * (size_t)c + k] == 0 ? 0x8000 : 0;
                    d.cell_slots[6 * (size_t)i + k] = (int16_t)(uint16_t)(sl | own);
                }
            }
        }
    }, 4);
    d.face_first.assign(g.n_faces, -1);
    for (size_t r = 0; r < ntf; r++)
        if (d.face_first[d.face_ref[r]] < 0) d.face_first[d.face_ref[r]] = (int)r;
    // Static cell force: body force plus the pressure term rho0/3 sum_f n_f A_f
    // that the deviation-form face flux leaves out.  It vanishes on closed
    // cells; with the reference's face areas (sum of triangle areas) warped
    // faces leave a defect, kept here exactly in FP64.
    {
        // per cell over its faces in face order (an owner's term before a
        // neighbour's on a face the cell is both of): the sums of a face loop
        std::vector<double> &S = d.closure;
        S.assign(3 * (size_t)nc, 0.0);
        parallel_for(nc, [&](int c0, int c1) {
            for (int c = c0; c < c1; c++)
                for (int k = 0; k < 6; k++) {
                    const int f = g.gface[6 * (size_t)c + k];
                    // faces of empty patches carry no flux at all
                    if (f >= g.n_int && m.bc_vel_type && m.bc_vel_type[f - g.n_int] == LBFS_BC_EMPTY) continue;
                    const double A = g.area[f];
                    for (int t = 0; t < 3; t++) {
                        const double v = g.normal[3 * (size_t)f + t] * A;
                        if (g.grole[6 * (size_t)c + k] == 0) S[3 * (size_t)c + t] += v;
                        else S[3 * (size_t)c + t] -= v;
                    }
                }
        });
        d.static_force.assign(3 * (size_t)nc, 0.0f);
        parallel_for(nc, [&](int i0, int i1) {
            for (int i = i0; i < i1; i++) {
                const int c = d.perm[i];
                const double *bf = cell_body_force(blocks, c, body_force);
                for (int t = 0; t < 3; t++)
                    d.static_force[3 * (size_t)i + t] =
                        (float)(bf[t] - (t < dims ? rho0 / 3.0 * S[3 * (size_t)c + t] / g.vol[c] : 0.0));
            }
        });
    }
    d.bc_owner.assign(g.n_bc, 0);
    for (int b = 0; b < g.n_bc; b++) d.bc_owner[b] = d.inv[g.owner[g.n_int + b]];
    return LBFS_OK;
}

}  // namespace lbfs
