// Mesh geometry and least-squares set-up on the GPU, in FP64 (gfx950).
//
//   k_geo_faces   per face: the two triangles' normals, areas, centroids and
//                 origin-tetrahedron volumes (unstructured_mesh.cpp:1015-1045,
//                 1170-1204), the face area, unit normal, centroid and
//                 streaming step (unstructured_mesh.cpp:1435-1443); the face
//                 joins its owner's and neighbour's face lists
//   k_geo_cells   per cell: its six faces in face order, the volume, the
//                 area-weighted centroid and the stencil (unstructured_mesh.cpp:
//                 1592-1692)
//   k_geo_ghosts  per boundary face: the ghost centroid halfway to the face
//                 (unstructured_mesh.cpp:1963-1996)
//   k_geo_ls      per cell: the hybrid least-squares / Green-Gauss weights and
//                 the stored inverse (gradients.cpp:149-210, 214-313)
//
// Every sum runs in the order of the host set-up in mesh_host.cpp (faces in
// increasing index, the two triangles of a face in turn) and no product is
// fused into an add, so the results are the host's bit for bit: the host
// path stays the definition, this one the fast way to the same numbers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "geometry_gpu.hpp"

#pragma clang fp contract(off)

namespace lbfs {

namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double norm(D3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ D3 ld(const double *p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ void st(double *p, D3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// error slots: the smallest offending face or cell of each kind
enum { GE_FACE = 0, GE_MORE = 1, GE_FEWER = 2, GE_VOL = 3, GE_LS = 4, GE_N = 5 };

struct GeoArgs {
    int nc, nf, ni;
    const double *pts;
    const int *fv, *own, *nbr;
    const double *shift;       // 3 per face, or null
    double visc, gamma, fneq_min;
    int closed;
    double *area, *normal, *fc, *fdt;
    double *tvo, *tvn;         // 2 per face: the triangles' volume terms for the owner and the neighbour
    double *csum, *asum;       // 3 + 1 per face: area-weighted centroid sum and area sum
    double *cc, *vol;          // 3 per cell and ghost; per cell
    int *cnt;                  // per cell: faces listed
    int2 *ent;                 // 6 per cell: (face, role) as listed
    int *gcell, *gface, *grole;
    double *lhs, *rhs;         // 9, 18 per cell
    int grad_type, det_compat;
    int *err;                  // GE_N slots (initialised to INT_MAX)
};

__device__ __forceinline__ D3 face_shift(const GeoArgs &a, int f)
{
    return a.shift ? ld(a.shift + 3 * (size_t)f) : D3{0.0, 0.0, 0.0};
}

__global__ __launch_bounds__(256) void k_geo_faces(GeoArgs a)
{
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= a.nf) return;
    const int o = a.own[f];
    const int n = f < a.ni ? a.nbr[f] : a.nc + f - a.ni;
    if (o < 0 || o >= a.nc || (f < a.ni && (n < 0 || n >= a.nc))) {
        atomicMin(a.err + GE_FACE, f);
        return;
    }
    const D3 s = face_shift(a, f);
    D3 P[4];
#pragma unroll
    for (int t = 0; t < 4; t++) P[t] = ld(a.pts + 3 * (size_t)a.fv[4 * (size_t)f + t]);
    D3 csum{0.0, 0.0, 0.0}, nsum{0.0, 0.0, 0.0};
    double asum = 0.0;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const D3 pa = P[0], pb = P[t + 1], pc = P[t + 2];
        const D3 N = cross(pb - pa, pc - pa);
        const double tv = dot(pa, N) / 6.0;
        a.tvo[2 * (size_t)f + t] = tv;
        if (f < a.ni) a.tvn[2 * (size_t)f + t] = tv - dot(s, N) / 6.0;
        const double ta = 0.5 * norm(N);
        const D3 cen = (1.0 / 3.0) * (pa + pb + pc);
        csum = csum + ta * cen;
        nsum = nsum + N;
        asum += ta;
    }
    a.area[f] = a.closed ? 0.5 * norm(nsum) : asum;
    st(a.fc + 3 * (size_t)f, (1.0 / asum) * csum);
    st(a.normal + 3 * (size_t)f, (1.0 / norm(nsum)) * nsum);
    double fd = 0.5 * sqrt(asum);
    if (a.fneq_min > 0.0) {
        const double cap = a.visc * 3.0 / a.gamma / a.fneq_min;
        fd = cap < fd ? cap : fd;
    }
    a.fdt[f] = fd;
    st(a.csum + 3 * (size_t)f, csum);
    a.asum[f] = asum;
    // the face joins its cells' lists (order restored per cell below)
    const int ko = atomicAdd(a.cnt + o, 1);
    if (ko < 6) a.ent[6 * (size_t)o + ko] = make_int2(f, 0);
    if (f < a.ni) {
        const int kn = atomicAdd(a.cnt + n, 1);
        if (kn < 6) a.ent[6 * (size_t)n + kn] = make_int2(f, 1);
    }
}

__global__ __launch_bounds__(256) void k_geo_cells(GeoArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.nc) return;
    const int c = a.cnt[i];
    if (c != 6) {
        atomicMin(a.err + (c > 6 ? GE_MORE : GE_FEWER), i);
        return;
    }
    int2 e[6];
#pragma unroll
    for (int k = 0; k < 6; k++) e[k] = a.ent[6 * (size_t)i + k];
    // face order, the owner's entry before the neighbour's of the same face
    for (int k = 1; k < 6; k++) {
        const int2 v = e[k];
        int q = k - 1;
        while (q >= 0 && (e[q].x > v.x || (e[q].x == v.x && e[q].y > v.y))) {
            e[q + 1] = e[q];
            q--;
        }
        e[q + 1] = v;
    }
    double vol = 0.0, tot = 0.0;
    D3 cc{0.0, 0.0, 0.0};
    for (int k = 0; k < 6; k++) {
        const int f = e[k].x;
        const double as = a.asum[f];
        const D3 cs = ld(a.csum + 3 * (size_t)f);
        if (e[k].y == 0 && k + 1 < 6 && e[k + 1].x == f) {
            // the cell is its own neighbour across a one-cell periodic axis:
            // the two roles' volume terms interleave triangle by triangle
            vol += a.tvo[2 * (size_t)f];
            vol -= a.tvn[2 * (size_t)f];
            vol += a.tvo[2 * (size_t)f + 1];
            vol -= a.tvn[2 * (size_t)f + 1];
            tot += as;
            cc = cc + cs;
            tot += as;
            cc = cc + cs - as * face_shift(a, f);
            k++;
            continue;
        }
        if (e[k].y == 0) {
            vol += a.tvo[2 * (size_t)f];
            vol += a.tvo[2 * (size_t)f + 1];
            tot += as;
            cc = cc + cs;
        } else {
            vol -= a.tvn[2 * (size_t)f];
            vol -= a.tvn[2 * (size_t)f + 1];
            tot += as;
            cc = cc + cs - as * face_shift(a, f);
        }
    }
    st(a.cc + 3 * (size_t)i, (1.0 / tot) * cc);
    a.vol[i] = vol;
    if (!(vol > 0.0)) atomicMin(a.err + GE_VOL, i);
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int f = e[k].x;
        const int n = f < a.ni ? a.nbr[f] : a.nc + f - a.ni;
        a.gcell[6 * (size_t)i + k] = e[k].y == 0 ? n : a.own[f];
        a.gface[6 * (size_t)i + k] = f;
        a.grole[6 * (size_t)i + k] = e[k].y;
    }
}

__global__ __launch_bounds__(256) void k_geo_ghosts(GeoArgs a)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int f = a.ni + b;
    if (f >= a.nf) return;
    const D3 co = ld(a.cc + 3 * (size_t)a.own[f]), xf = ld(a.fc + 3 * (size_t)f);
    st(a.cc + 3 * (size_t)(a.nc + b), co + 0.5 * (xf - co));
}

__global__ __launch_bounds__(256) void k_geo_ls(GeoArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.nc) return;
    D3 d[6], dxf[6];
    const D3 ci = ld(a.cc + 3 * (size_t)i);
    double max_area = 0.0, max_disp = 0.0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int i1 = a.gcell[6 * (size_t)i + k], f = a.gface[6 * (size_t)i + k];
        const D3 s = face_shift(a, f);
        const D3 xf = ld(a.fc + 3 * (size_t)f);
        const D3 c1 = ld(a.cc + 3 * (size_t)i1);
        if (a.grole[6 * (size_t)i + k] == 0) {
            d[k] = c1 + s - ci;
            dxf[k] = xf - ci;
        } else {
            d[k] = c1 - s - ci;
            dxf[k] = xf - s - ci;
        }
        const double ar = a.area[f], dn = norm(d[k]);
        max_area = max_area < ar ? ar : max_area;
        max_disp = max_disp < dn ? dn : max_disp;
    }
    const double aspect = 2.0 * fabs(max_disp) * max_area / a.vol[i];
    double beta = 0.0;
    if (a.grad_type == LBFS_GRAD_HYBRID) {
        const double b2 = 2.0 / aspect;
        beta = b2 < 1.0 ? b2 : 1.0;
    } else if (a.grad_type == LBFS_GRAD_LS) {
        beta = 1.0;
    }
    double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    M[0] = M[4] = M[8] = 2.0 * (1.0 - beta) * a.vol[i];
    double *R = a.rhs + 18 * (size_t)i;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int f = a.gface[6 * (size_t)i + k];
        const D3 n = ld(a.normal + 3 * (size_t)f);
        const double sgn = a.grole[6 * (size_t)i + k] == 0 ? 1.0 : -1.0;
        const double ln = dot(dxf[k], n), Ln = dot(d[k], n), Lc = norm(d[k]);
        const double r = 2.0 * ln / Ln;
        const double w = r * r * a.area[f] / Lc;
        const double dv[3] = {d[k].x, d[k].y, d[k].z}, nv[3] = {n.x, n.y, n.z};
#pragma unroll
        for (int p = 0; p < 3; p++) {
#pragma unroll
            for (int q = 0; q < 3; q++) M[3 * p + q] += w * dv[p] * dv[q] * beta;
            R[3 * k + p] = beta * w * dv[p] + (1.0 - beta) * nv[p] * sgn * a.area[f];
        }
    }
    const double ma = M[0], mb = M[1], mc = M[2], md = M[3], me = M[4], mf = M[5], mg = M[6], mh = M[7], mk = M[8];
    const double det = a.det_compat ? ma * (me * mk - mf * mh) + mb * (md * mk - mf - mg) + mc * (md * mh - me * mg)
                                    : ma * (me * mk - mf * mh) - mb * (md * mk - mf * mg) + mc * (md * mh - me * mg);
    if (!(fabs(det) > 0.0) || !isfinite(det)) atomicMin(a.err + GE_LS, i);
    double *L = a.lhs + 9 * (size_t)i;
    L[0] = (me * mk - mf * mh) / det; L[1] = (mc * mh - mb * mk) / det; L[2] = (mb * mf - mc * me) / det;
    L[3] = (mf * mg - md * mk) / det; L[4] = (ma * mk - mc * mg) / det; L[5] = (mc * md - ma * mf) / det;
    L[6] = (md * mh - me * mg) / det; L[7] = (mb * mg - ma * mh) / det; L[8] = (ma * me - mb * md) / det;
}

__global__ void k_geo_err_init(int *err)
{
    if (threadIdx.x < GE_N) err[threadIdx.x] = 0x7fffffff;
}

// device buffers of the set-up, freed on every exit path
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <class T>
    hipError_t alloc(T **out, size_t n)
    {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        p.push_back(q);
        *out = (T *)q;
        return hipSuccess;
    }
    template <class T>
    hipError_t upload(T **out, const T *src, size_t n, hipStream_t s)
    {
        const hipError_t e = alloc(out, n);
        if (e != hipSuccess || !n) return e;
        return hipMemcpyAsync(*out, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

template <class T>
hipError_t download(std::vector<T> &dst, const T *src, size_t n, hipStream_t s)
{
    dst.resize(n);
    return n ? hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, s) : hipSuccess;
}

}  // namespace

int gpu_geometry(const lbfs_mesh_desc &m, double visc, double gamma, double fneq_min, bool closed, int grad_type,
                 bool det_compat, hipStream_t s, Geometry &g, std::vector<double> &lhs, std::vector<double> &rhs,
                 std::string &err)
{
    const int nc = m.n_cells, nf = m.n_faces, ni = m.n_internal, nb = nf - ni;
    if (nc <= 0 || nf <= 0 || ni < 0 || ni > nf || !m.points || !m.face_vertices || !m.owner ||
        (ni > 0 && !m.neighbour) || m.n_points <= 0) {
        err = "mesh description incomplete";
        return LBFS_E_INVALID_ARG;
    }
    for (size_t q = 0; q < 4 * (size_t)nf; q++)
        if (m.face_vertices[q] < 0 || m.face_vertices[q] >= m.n_points) {
            err = "face " + std::to_string(q / 4) + " has an invalid vertex";
            return LBFS_E_INVALID_ARG;
        }
#define GCHK(x)                                                              \
    do {                                                                     \
        const hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) {                                              \
            err = std::string("geometry on the GPU: ") + hipGetErrorString(e_); \
            return LBFS_E_HIP;                                               \
        }                                                                    \
    } while (0)
    DevBufs B;
    GeoArgs a{};
    a.nc = nc; a.nf = nf; a.ni = ni;
    a.visc = visc; a.gamma = gamma; a.fneq_min = fneq_min; a.closed = closed ? 1 : 0;
    a.grad_type = grad_type; a.det_compat = det_compat ? 1 : 0;
    double *pts, *shift = nullptr;
    int *fv, *own, *nbr;
    GCHK(B.upload(&pts, m.points, 3 * (size_t)m.n_points, s));
    GCHK(B.upload(&fv, m.face_vertices, 4 * (size_t)nf, s));
    GCHK(B.upload(&own, m.owner, (size_t)nf, s));
    GCHK(B.upload(&nbr, m.neighbour, (size_t)ni, s));
    if (m.face_shift) GCHK(B.upload(&shift, m.face_shift, 3 * (size_t)nf, s));
    a.pts = pts; a.fv = fv; a.own = own; a.nbr = nbr; a.shift = shift;
    GCHK(B.alloc(&a.area, nf)); GCHK(B.alloc(&a.normal, 3 * (size_t)nf)); GCHK(B.alloc(&a.fc, 3 * (size_t)nf));
    GCHK(B.alloc(&a.fdt, nf)); GCHK(B.alloc(&a.tvo, 2 * (size_t)nf)); GCHK(B.alloc(&a.tvn, 2 * (size_t)nf));
    GCHK(B.alloc(&a.csum, 3 * (size_t)nf)); GCHK(B.alloc(&a.asum, nf));
    GCHK(B.alloc(&a.cc, 3 * (size_t)(nc + nb))); GCHK(B.alloc(&a.vol, nc));
    GCHK(B.alloc(&a.cnt, nc)); GCHK(B.alloc(&a.ent, 6 * (size_t)nc));
    GCHK(B.alloc(&a.gcell, 6 * (size_t)nc)); GCHK(B.alloc(&a.gface, 6 * (size_t)nc));
    GCHK(B.alloc(&a.grole, 6 * (size_t)nc));
    GCHK(B.alloc(&a.lhs, 9 * (size_t)nc)); GCHK(B.alloc(&a.rhs, 18 * (size_t)nc));
    GCHK(B.alloc(&a.err, GE_N));
    GCHK(hipMemsetAsync(a.cnt, 0, sizeof(int) * (size_t)nc, s));
    hipLaunchKernelGGL(k_geo_err_init, dim3(1), dim3(64), 0, s, a.err);
    hipLaunchKernelGGL(k_geo_faces, dim3((nf + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_geo_cells, dim3((nc + 255) / 256), dim3(256), 0, s, a);
    GCHK(hipGetLastError());
    int he[GE_N];
    GCHK(hipMemcpyAsync(he, a.err, sizeof(he), hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    if (he[GE_FACE] != 0x7fffffff) {
        err = "face " + std::to_string(he[GE_FACE]) + " has an invalid owner/neighbour";
        return LBFS_E_INVALID_ARG;
    }
    if (he[GE_MORE] != 0x7fffffff) {
        err = "cell " + std::to_string(he[GE_MORE]) + " has more than 6 faces";
        return LBFS_E_UNSUPPORTED;
    }
    if (he[GE_FEWER] != 0x7fffffff) {
        err = "cell " + std::to_string(he[GE_FEWER]) + " is not a hexahedron";
        return LBFS_E_UNSUPPORTED;
    }
    if (he[GE_VOL] != 0x7fffffff) {
        err = "cell " + std::to_string(he[GE_VOL]) + " has non-positive volume";
        return LBFS_E_INVALID_ARG;
    }
    if (nb > 0) hipLaunchKernelGGL(k_geo_ghosts, dim3((nb + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_geo_ls, dim3((nc + 255) / 256), dim3(256), 0, s, a);
    GCHK(hipGetLastError());
    GCHK(hipMemcpyAsync(he, a.err, sizeof(he), hipMemcpyDeviceToHost, s));
    g.n_cells = nc; g.n_faces = nf; g.n_int = ni; g.n_bc = nb;
    g.owner.assign(m.owner, m.owner + nf);
    g.nbr.resize(nf);
    for (int f = 0; f < nf; f++) g.nbr[f] = f < ni ? m.neighbour[f] : nc + f - ni;
    g.shift.assign(3 * (size_t)nf, 0.0);
    if (m.face_shift) std::copy(m.face_shift, m.face_shift + 3 * (size_t)nf, g.shift.begin());
    GCHK(download(g.area, a.area, nf, s));
    GCHK(download(g.normal, a.normal, 3 * (size_t)nf, s));
    GCHK(download(g.fc, a.fc, 3 * (size_t)nf, s));
    GCHK(download(g.fdt, a.fdt, nf, s));
    GCHK(download(g.cc, a.cc, 3 * (size_t)(nc + nb), s));
    GCHK(download(g.vol, a.vol, nc, s));
    GCHK(download(g.gcell, a.gcell, 6 * (size_t)nc, s));
    GCHK(download(g.gface, a.gface, 6 * (size_t)nc, s));
    GCHK(download(g.grole, a.grole, 6 * (size_t)nc, s));
    GCHK(download(lhs, a.lhs, 9 * (size_t)nc, s));
    GCHK(download(rhs, a.rhs, 18 * (size_t)nc, s));
    GCHK(hipStreamSynchronize(s));
#undef GCHK
    if (he[GE_LS] != 0x7fffffff) {
        err = "singular least-squares matrix at cell " + std::to_string(he[GE_LS]);
        return LBFS_E_INVALID_ARG;
    }
    return LBFS_OK;
}

}  // namespace lbfs
