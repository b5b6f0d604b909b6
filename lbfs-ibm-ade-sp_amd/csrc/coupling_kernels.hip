// Immersed-boundary coupling and spring-network membrane kernels for gfx950.
//
//   k_ibm_interp   per node: fluid velocity at the node (immersed_boundary_method.cu:14-97
//                  or the 4-point cosine kernel), explicit advection X += U dt
//                  (immersed_boundary_method.cu:234-264), bin key of the advected node
//   k_ibm_scan     exclusive scan of the bin counts (one workgroup), counts cleared
//   k_ibm_place    nodes placed in bin order
//   k_mem_globals  per membrane (one workgroup): volume, area (FP64) and the MAD of
//                  the previous force call (spring_network.cu:1186-1238,
//                  lagrangian_object.cpp:601-650)
//   k_mem_edges    per spring: all membrane force terms (spring_network.cu:613-971),
//                  written per edge and slot (no atomics)
//   k_mem_nodes    per node: sum of its edge slots in a fixed order
//   k_ibm_spread   per 4x4x4 bin tile: each cell gathers the forces of the nodes
//                  binned around it (immersed_boundary_method.cu:947-1024); the sum
//                  is accumulated in 64-bit fixed point so its value does not depend
//                  on the order the nodes are visited in
#include <hip/hip_runtime.h>

#include "coupling.hpp"

namespace lbfs {

namespace {

constexpr double FIX_SCALE = 1099511627776.0;  // 2^40
constexpr float PI_F = 3.14159265358979f;

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ f3 ld3(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float cmp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
// IBM stencil helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int stencil_size(const IbmArgs &g) { return g.kernel == LBFS_IBM_LINEAR2 ? 2 : 4; }

// nearest vertex plane, ties to the larger index (find_index_closest,
// immersed_boundary_method.cu:588-636), by bisection
__device__ __forceinline__ int closest_vertex(const double *v, int n, double t)
{
    if (t <= v[0]) return 0;
    if (t >= v[n - 1]) return n - 1;
    int lo = 0, hi = n - 1;  // v[lo] < t <= v[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < t) lo = mid; else hi = mid;
    }
    return (t - v[lo] >= v[hi] - t) ? hi : lo;
}

// cell-line coordinate at unwrapped index c (periodic image or wall continuation)
__device__ __forceinline__ double line_coord(const IbmArgs &g, const IbmBlock &B, int a, int c)
{
    const int n = B.n[a];
    const double *l = g.axis + B.ax_off[a] + n + 1;
    if (c >= 0 && c < n) return l[c];
    if (B.per[a]) {
        const int k = c >= 0 ? c / n : -((-c + n - 1) / n);
        return l[c - k * n] + (double)k * (double)B.L[a];
    }
    const double d0 = n > 1 ? l[1] - l[0] : (double)B.h, d1 = n > 1 ? l[n - 1] - l[n - 2] : (double)B.h;
    return c < 0 ? l[0] + c * d0 : l[n - 1] + (c - (n - 1)) * d1;
}

// cubic Lagrange basis i of the 4 line points from `base` at x (the first
// column of the inverse Vandermonde matrix, vandemonde_inverse :919-941)
__device__ __forceinline__ double lagrange4(const IbmArgs &g, const IbmBlock &B, int a, int base, int i, double x)
{
    double p[4];
#pragma unroll
    for (int k = 0; k < 4; k++) p[k] = line_coord(g, B, a, base + k);
    double v = 1.0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (k != i) v *= (x - p[k]) / (p[i] - p[k]);
    return v;
}

__device__ __forceinline__ void node_base(const IbmArgs &g, const IbmBlock &B, const float P[3], float xr[3], int base[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        xr[a] = (P[a] - B.origin[a]) * B.inv_h;
        if (g.kernel == LBFS_IBM_LAGRANGE4)
            base[a] = closest_vertex(g.axis + B.ax_off[a], B.n[a] + 1, (double)P[a]) - 2;
        else
            base[a] = g.kernel == LBFS_IBM_LINEAR2 ? (int)roundf(xr[a]) - 1 : (int)floorf(xr[a] - 1.5f);
    }
}

// wrap a logical index into the grid; false if it lies outside a wall
__device__ __forceinline__ bool wrap_index(const IbmBlock &B, int a, int i, int &w)
{
    const int n = B.n[a];
    if (i >= 0 && i < n) { w = i; return true; }
    if (!B.per[a]) return false;
    w = ((i % n) + n) % n;
    return true;
}

// unnormalised weight of the cell with unwrapped index idx / wrapped index w
// phi(r) = (1 + cos(pi r / 2)) / 4 for |r| < 2, written as (1 - sin^2(pi r / 4)) / 2
// with sin from its Taylor series to x^11 (|x| <= pi/2: truncation below 6e-8)
__device__ __forceinline__ float cos4_phi(float r)
{
    const float x = r * (PI_F * 0.25f), x2 = x * x;
    float p = -2.5052108385e-8f;                 // -1/11!
    p = fmaf(p, x2, 2.7557319224e-6f);           //  1/9!
    p = fmaf(p, x2, -1.9841269841e-4f);          // -1/7!
    p = fmaf(p, x2, 8.3333333333e-3f);           //  1/5!
    p = fmaf(p, x2, -1.6666666667e-1f);          // -1/3!
    const float sn = fmaf(p * x2, x, x);
    return 0.5f * fmaf(-sn, sn, 1.0f);
}

__device__ __forceinline__ float cell_weight(const IbmArgs &g, const IbmBlock &B, const int idx[3], const int w[3],
                                             int internal, const float P[3], const float xr[3])
{
    float wt = 1.0f;
    if (g.kernel == LBFS_IBM_LINEAR2) {
#pragma unroll
        for (int a = 0; a < 3; a++) wt *= 1.0f - fabsf((float)idx[a] + 0.5f - xr[a]);
    } else {
        const float4 c = g.cen[internal];
        const float cc[3] = {c.x, c.y, c.z};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float xc = cc[a] + (float)(idx[a] - w[a]) * (B.L[a] / (float)B.n[a]);
            const float r = (xc - P[a]) * B.inv_h;
            wt *= fabsf(r) < 2.0f ? cos4_phi(r) : 0.0f;
        }
    }
    return wt;
}

__device__ __forceinline__ int cell_internal(const IbmArgs &g, const IbmBlock &B, const int w[3])
{
    return g.inv[B.cell0 + w[0] + B.n[0] * (w[1] + B.n[1] * w[2])];
}

__device__ __forceinline__ int bin_coord(const IbmBlock &B, int a, int base)
{
    const int n = B.n[a];
    int b = base;
    if (B.per[a]) b = ((b % n) + n) % n;
    else b = b < 0 ? 0 : (b >= n ? n - 1 : b);
    return b >> 2;
}

}  // namespace

// ---------------------------------------------------------------------------
// IBM: interpolation, advection, weights, binning.  One wavefront per node,
// one lane per stencil cell (64 for the 4-point kernel, 8 for the 2-point
// one); sums are butterfly reductions, so their order is fixed.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void lane_offset(int S, int lane, int off[3])
{
    if (S == 4) { off[0] = lane & 3; off[1] = (lane >> 2) & 3; off[2] = lane >> 4; }
    else { off[0] = lane & 1; off[1] = (lane >> 1) & 1; off[2] = lane >> 2; }
}

// INTERP: interpolate the cell velocity with the node's cached weights (those
// of its current position), advect the node and count it into its bin.
// Always: weights of the (new) position, normalised for the cosine kernel,
// cached for spreading and the next interpolation.
template <bool INTERP>
__global__ __launch_bounds__(256) void k_ibm_interp(MemArgs m, IbmArgs g, const float4 *__restrict__ st)
{
    const int lane = threadIdx.x & 63;
    const int nd = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (nd >= m.n_nodes) return;  // uniform over the wavefront
    const int S = stencil_size(g);
    const bool act = lane < S * S * S;
    int off[3];
    lane_offset(S, lane, off);
    const IbmBlock B = g.blk[g.node_block[nd]];
    const float4 o = m.origin[m.node_mem[nd]];
    float4 xl = m.X[nd];
    if (INTERP) {
        const int4 nb = g.nbase[nd];
        float ux = 0.0f, uy = 0.0f, uz = 0.0f;
        if (act) {
            const int idx[3] = {nb.x + off[0], nb.y + off[1], nb.z + off[2]};
            int w[3];
            if (wrap_index(B, 0, idx[0], w[0]) && wrap_index(B, 1, idx[1], w[1]) && wrap_index(B, 2, idx[2], w[2])) {
                const float wt = g.wnode[64 * (size_t)nd + lane];
                const float4 s = st[cell_internal(g, B, w)];
                ux = wt * s.y; uy = wt * s.z; uz = wt * s.w;
            }
        }
        ux = wave_sum(ux); uy = wave_sum(uy); uz = wave_sum(uz);
        xl.x += ux * g.dt; xl.y += uy * g.dt; xl.z += uz * g.dt;
        if (lane == 0) {
            m.U[nd] = make_float4(ux, uy, uz, 0.0f);
            m.X[nd] = xl;
        }
    }
    const float P[3] = {o.x + xl.x, o.y + xl.y, o.z + xl.z};
    float xr[3];
    int base[3];
    node_base(g, B, P, xr, base);
    float wt = 0.0f;
    if (act) {
        const int idx[3] = {base[0] + off[0], base[1] + off[1], base[2] + off[2]};
        int w[3];
        if (wrap_index(B, 0, idx[0], w[0]) && wrap_index(B, 1, idx[1], w[1]) && wrap_index(B, 2, idx[2], w[2])) {
            if (g.kernel == LBFS_IBM_LAGRANGE4)
                wt = (float)(lagrange4(g, B, 0, base[0], off[0], (double)P[0]) *
                             lagrange4(g, B, 1, base[1], off[1], (double)P[1]) *
                             lagrange4(g, B, 2, base[2], off[2], (double)P[2]));
            else
                wt = cell_weight(g, B, idx, w, cell_internal(g, B, w), P, xr);
        }
    }
    const float ws = wave_sum(wt);
    const float inv = g.kernel == LBFS_IBM_COS4 ? (ws > 0.0f ? 1.0f / ws : 0.0f) : 1.0f;
    g.wnode[64 * (size_t)nd + lane] = wt * inv;
    if (lane == 0) {
        const int key = B.bin0 + bin_coord(B, 0, base[0]) +
                        B.tn[0] * (bin_coord(B, 1, base[1]) + B.tn[1] * bin_coord(B, 2, base[2]));
        g.nbase[nd] = make_int4(base[0], base[1], base[2], key);
        if (INTERP && g.binned) {
            g.bin_key[nd] = key;
            g.bin_rank[nd] = atomicAdd(&g.bin_count[key], 1);
        }
    }
}

// exclusive scan of the bin counts in chunks of 1024 (counts are cleared for
// the next step), then the chunk offsets
__global__ __launch_bounds__(1024) void k_ibm_scan1(IbmArgs g)
{
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * 1024 + threadIdx.x;
    const int v = b < g.n_bins ? g.bin_count[b] : 0;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        int t = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            const int y = __shfl_up(t, off, 64);
            if (lane >= off) t += y;
        }
        if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    const int excl = x - v + (wave > 0 ? wsum[wave - 1] : 0);
    if (b < g.n_bins) {
        g.bin_start[b] = excl;
        g.bin_count[b] = 0;
    }
    if (threadIdx.x == 0) g.scan_part[blockIdx.x] = wsum[15];
}

__global__ __launch_bounds__(1024) void k_ibm_scan2(IbmArgs g)
{
    __shared__ int offset;
    if (threadIdx.x == 0) {
        int s = 0;
        for (int k = 0; k < (int)blockIdx.x; k++) s += g.scan_part[k];
        offset = s;
        if (blockIdx.x == gridDim.x - 1) g.bin_start[g.n_bins] = s + g.scan_part[blockIdx.x];
    }
    __syncthreads();
    const int b = blockIdx.x * 1024 + threadIdx.x;
    if (b < g.n_bins) g.bin_start[b] += offset;
}

__global__ __launch_bounds__(256) void k_ibm_place(MemArgs m, IbmArgs g)
{
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= m.n_nodes) return;
    g.sorted[g.bin_start[g.bin_key[nd]] + g.bin_rank[nd]] = nd;
}

// ---------------------------------------------------------------------------
// IBM: spreading.  One workgroup per node bin (the 4x4x4 block of cells that
// holds the nodes' stencil base): its nodes' stencils stay inside a 7x7x7
// window, where the weighted forces are accumulated in LDS (one wavefront per
// node, one lane per stencil slot), then the window is added to the per-cell
// accumulators in global memory.  All sums are 64-bit fixed point integer
// additions, so the result does not depend on the order of the nodes, lanes
// or workgroups.  k_ibm_finalize then turns the accumulators of the touched
// tiles into force densities and clears them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ibm_spread(MemArgs m, IbmArgs g)
{
    const int b = blockIdx.x;
    const int lo = g.bin_start[b], hi = g.bin_start[b + 1];
    if (lo == hi) return;  // uniform over the workgroup
    const IbmBlock B = g.blk[g.bin_block[b]];
    const int lb = b - B.bin0;
    const int bo[3] = {4 * (lb % B.tn[0]), 4 * ((lb / B.tn[0]) % B.tn[1]), 4 * (lb / (B.tn[0] * B.tn[1]))};
    __shared__ unsigned long long racc[3][343];
    __shared__ int cnd[64];
    __shared__ int4 cbase[64];
    __shared__ float4 cF[64];
    __shared__ float cw[64][64];   // the chunk's cached stencil weights, one row per node
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int q = tid; q < 3 * 343; q += 256) racc[q / 343][q % 343] = 0ull;
    const int S = stencil_size(g);
    int off[3];
    lane_offset(S, lane, off);
    const bool act = lane < S * S * S;
    for (int c0 = lo; c0 < hi; c0 += 64) {
        const int cnt = min(64, hi - c0);
        __syncthreads();
        if (tid < cnt) {
            const int p = g.sorted[c0 + tid];
            cnd[tid] = p;
            cbase[tid] = g.nbase[p];
            cF[tid] = m.F[p];
        }
        __syncthreads();
        // all weight rows of the chunk in flight at once (one coalesced row per node)
        for (int q = tid; q < 64 * cnt; q += 256) cw[q >> 6][q & 63] = g.wnode[64 * (size_t)cnd[q >> 6] + (q & 63)];
        __syncthreads();
        for (int jn = wave; jn < cnt; jn += 4) {
            const float wt = cw[jn][lane];
            if (!act || wt == 0.0f) continue;
            const int4 nb = cbase[jn];
            const int bs[3] = {nb.x, nb.y, nb.z};
            int loc = 0, mul = 1;
            bool in = true;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                int c = bs[a] + off[a];
                int l;
                if (B.per[a]) {
                    // window coordinate of the wrapped base, then the slot offset
                    // (nodes advect freely across periodic boundaries, so
                    // the base may lie several periods away)
                    const int wb = ((bs[a] % B.n[a]) + B.n[a]) % B.n[a];
                    l = wb - bo[a] + off[a];
                } else {
                    in = in && c >= 0 && c < B.n[a];
                    l = c - bo[a];
                }
                in = in && l >= 0 && l < 7;
                loc += l * mul;
                mul *= 7;
            }
            if (!in) continue;
            const float4 F = cF[jn];
            atomicAdd(&racc[0][loc], (unsigned long long)llrint((double)(F.x * wt) * FIX_SCALE));
            atomicAdd(&racc[1][loc], (unsigned long long)llrint((double)(F.y * wt) * FIX_SCALE));
            atomicAdd(&racc[2][loc], (unsigned long long)llrint((double)(F.z * wt) * FIX_SCALE));
        }
    }
    __syncthreads();
    for (int q = tid; q < 343; q += 256) {
        const unsigned long long v0 = racc[0][q], v1 = racc[1][q], v2 = racc[2][q];
        if ((v0 | v1 | v2) == 0ull) continue;
        const int l[3] = {q % 7, (q / 7) % 7, q / 49};
        int w[3];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            int c = bo[a] + l[a];
            if (c >= B.n[a]) {
                if (!B.per[a]) { ok = false; break; }
                c -= B.n[a];
                if (c >= B.n[a]) c %= B.n[a];
            }
            w[a] = c;
        }
        if (!ok) continue;
        const int ic = cell_internal(g, B, w);
        unsigned long long *acc = g.acc + 3 * (size_t)ic;
        if (v0) atomicAdd(acc, v0);
        if (v1) atomicAdd(acc + 1, v1);
        if (v2) atomicAdd(acc + 2, v2);
        g.tile_touched[B.bin0 + (w[0] >> 2) + B.tn[0] * ((w[1] >> 2) + B.tn[1] * (w[2] >> 2))] = 1;
    }
}

// Spreading for small node counts: one wavefront per node, one lane per
// stencil slot, adding its weighted force straight into the cell's 64-bit
// fixed-point accumulators (integer additions: order independent).  No node
// binning is needed.
__global__ __launch_bounds__(256) void k_ibm_spread_nodes(MemArgs m, IbmArgs g)
{
    const int lane = threadIdx.x & 63;
    const int nd = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (nd >= m.n_nodes) return;
    const int S = stencil_size(g);
    if (lane >= S * S * S) return;
    const float wt = g.wnode[64 * (size_t)nd + lane];
    if (wt == 0.0f) return;
    const IbmBlock B = g.blk[g.node_block[nd]];
    const int4 nb = g.nbase[nd];
    int off[3];
    lane_offset(S, lane, off);
    const int idx[3] = {nb.x + off[0], nb.y + off[1], nb.z + off[2]};
    int w[3];
    if (!(wrap_index(B, 0, idx[0], w[0]) && wrap_index(B, 1, idx[1], w[1]) && wrap_index(B, 2, idx[2], w[2]))) return;
    const float4 F = m.F[nd];
    const int ic = cell_internal(g, B, w);
    unsigned long long *acc = g.acc + 3 * (size_t)ic;
    atomicAdd(acc, (unsigned long long)llrint((double)(F.x * wt) * FIX_SCALE));
    atomicAdd(acc + 1, (unsigned long long)llrint((double)(F.y * wt) * FIX_SCALE));
    atomicAdd(acc + 2, (unsigned long long)llrint((double)(F.z * wt) * FIX_SCALE));
    g.tile_touched[B.bin0 + (w[0] >> 2) + B.tn[0] * ((w[1] >> 2) + B.tn[1] * (w[2] >> 2))] = 1;
}

// per 4x4x4 tile: force density of the touched cells from their accumulators
// (which are cleared), static force back where IBM force was held last step
__global__ __launch_bounds__(64) void k_ibm_finalize(IbmArgs g)
{
    const int t = blockIdx.x;
    const int touched = g.tile_touched[t], flag = g.tile_flag[t];
    if (!touched && !flag) return;
    const IbmBlock B = g.blk[g.bin_block[t]];
    const int lt = t - B.bin0;
    const int tc[3] = {lt % B.tn[0], (lt / B.tn[0]) % B.tn[1], lt / (B.tn[0] * B.tn[1])};
    const int lane = threadIdx.x;
    const int w[3] = {4 * tc[0] + (lane & 3), 4 * tc[1] + ((lane >> 2) & 3), 4 * tc[2] + (lane >> 4)};
    if (w[0] < B.n[0] && w[1] < B.n[1] && w[2] < B.n[2]) {
        const int ic = cell_internal(g, B, w);
        if (touched) {
            const float sc = g.kernel == LBFS_IBM_LINEAR2 ? B.inv_h * B.inv_h * B.inv_h : 1.0f / g.cen[ic].w;
            unsigned long long *acc = g.acc + 3 * (size_t)ic;
            for (int a = 0; a < 3; a++) {
                const double v = (double)(long long)acc[a] / FIX_SCALE;
                g.cell_force[3 * (size_t)ic + a] = g.static_force[3 * (size_t)ic + a] + (float)v * sc;
                acc[a] = 0ull;
            }
        } else {
            for (int a = 0; a < 3; a++) g.cell_force[3 * (size_t)ic + a] = g.static_force[3 * (size_t)ic + a];
        }
    }
    if (lane == 0) {
        g.tile_flag[t] = touched;
        g.tile_touched[t] = 0;
    }
}

// ---------------------------------------------------------------------------
// membrane
// ---------------------------------------------------------------------------
// volume, area and MAD sums in two levels: partial sums over chunks of about
// 256 triangles (and the matching share of the MAD edges), then one
// wavefront per membrane adds its chunks' partials; both in a fixed order
__global__ __launch_bounds__(256) void k_mem_partials(MemArgs m)
{
    const int *q = m.chunk + 5 * (size_t)blockIdx.x;
    const int t0 = q[1], t1 = q[2], e0 = q[3], e1 = q[4];
    __shared__ double sh[3][256];
    double V = 0.0, A = 0.0, M = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += 256) {
        const int4 qq = m.tri[t];
        const f3 a = ld3(m.X[qq.x]), b = ld3(m.X[qq.y]), c = ld3(m.X[qq.z]);
        const f3 N = cross(b - a, c - a);
        A += 0.5 * (double)sqrtf(dot(N, N));
        V += (double)dot(N, (a + b + c) * (1.0f / 3.0f)) / 6.0;
    }
    for (int e = e0 + threadIdx.x; e < e1; e += 256) {
        const float4 s = m.edge_state[e];
        M += (double)s.x * ((double)s.y + (double)s.z) / 3.0;
    }
    sh[0][threadIdx.x] = V;
    sh[1][threadIdx.x] = A;
    sh[2][threadIdx.x] = M;
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + stp];
        __syncthreads();
    }
    if (threadIdx.x < 3) m.part[3 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_mem_globals(MemArgs m)
{
    const int mem = blockIdx.x, lane = threadIdx.x;
    double V = 0.0, A = 0.0, M = 0.0;
    for (int c = m.mem_chunk_start[mem] + lane; c < m.mem_chunk_start[mem + 1]; c += 64) {
        V += m.part[3 * (size_t)c];
        A += m.part[3 * (size_t)c + 1];
        M += m.part[3 * (size_t)c + 2];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        V += __shfl_xor(V, off, 64);
        A += __shfl_xor(A, off, 64);
        M += __shfl_xor(M, off, 64);
    }
    if (lane == 0) {
        double *gl = m.glob + 8 * (size_t)mem;
        gl[2] = M / gl[6];  // MAD of the previous call, over the area it was computed with
        gl[0] = V;
        gl[1] = A;
        gl[6] = A;
    }
}

namespace {
// derivatives of N = (q - p) x (r - p) and of N . centroid for the nine
// coordinates p.xyz q.xyz r.xyz: dN = e_k x (next - prev) (spring_network.cu:1348-1412)
__device__ __forceinline__ void tri_derivs(f3 p, f3 q, f3 r, f3 N, f3 dN[9], float dV[9])
{
    const f3 v[3] = {p, q, r};
    const f3 cen = (p + q + r) * (1.0f / 3.0f);
#pragma unroll
    for (int vt = 0; vt < 3; vt++) {
        const f3 dd = v[(vt + 1) % 3] - v[(vt + 2) % 3];
        const f3 dk[3] = {mk(0.0f, -dd.z, dd.y), mk(dd.z, 0.0f, -dd.x), mk(-dd.y, dd.x, 0.0f)};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            dN[3 * vt + k] = dk[k];
            dV[3 * vt + k] = cmp(N, k) * (1.0f / 3.0f) + dot(dk[k], cen);
        }
    }
}
}  // namespace

__global__ __launch_bounds__(128) void k_mem_edges(MemArgs m)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.n_edges) return;
    const int mem = m.edge_mem[e];
    const MemParamDev p = m.params[mem];
    const double *gl = m.glob + 8 * (size_t)mem;
    // global coefficients in FP64 (volume and area constraints, MAD)
    const float vol_coef = (float)(-(double)p.volume_modulus * (gl[0] / gl[3] - 1.0) / 6.0);
    const float area_coef = (float)(-(double)p.global_area_modulus * (gl[1] / gl[4] - 1.0));
    const float mad_coef = (float)(-2.0 * (double)p.global_bending * (gl[2] - gl[5]) / gl[4]);
    const int4 q = m.edges[e];
    const f3 A = ld3(m.X[q.x]), B = ld3(m.X[q.y]), C = ld3(m.X[q.z]), D = ld3(m.X[q.w]);
    const f3 l = C - A;
    const float length = sqrtf(dot(l, l));
    const f3 N1 = cross(B - A, C - A), N2 = cross(C - A, D - A);
    const float n1 = sqrtf(dot(N1, N1)), n2 = sqrtf(dot(N2, N2));
    f3 dN1[9], dN2[9];
    float dV1[9], dV2[9];
    tri_derivs(A, B, C, N1, dN1, dV1);
    tri_derivs(C, D, A, N2, dN2, dV2);
    const float in1 = 1.0f / n1, in2 = 1.0f / n2;
    float dn1[9], dn2[9];
#pragma unroll
    for (int j = 0; j < 9; j++) {
        dn1[j] = dot(dN1[j], N1) * in1;
        dn2[j] = dot(dN2[j], N2) * in2;
    }
    // slots A(0-2) B(3-5) C(6-8) D(9-11); triangle 2 is (C, D, A)
    float d_area[12], d_vol[12], d_ang[12];
#pragma unroll
    for (int j = 0; j < 9; j++) { d_area[j] = dn1[j]; d_vol[j] = dV1[j]; }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        d_area[9 + j] = dn2[3 + j];
        d_vol[9 + j] = dV2[3 + j];
        d_area[6 + j] += dn2[j];
        d_vol[6 + j] += dV2[j];
        d_area[j] += dn2[6 + j];
        d_vol[j] += dV2[6 + j];
    }
#pragma unroll
    for (int j = 0; j < 12; j++) { d_area[j] *= (1.0f / 6.0f); d_vol[j] *= (1.0f / 18.0f); }
    float cs = dot(N1, N2) * in1 * in2;
    cs = fminf(fmaxf(cs, -1.0f), 1.0f);
    const f3 c1 = (A + B + C) * (1.0f / 3.0f), c2 = (A + C + D) * (1.0f / 3.0f);
    const float dir = dot(N1, c2 - c1);
    const float angle = acosf(cs) * (float)((dir < 0.0f) - (dir > 0.0f));
    const float il = 1.0f / length;
    const f3 dL = l * il;
    {
        const f3 h1 = cross(N1 * in1, dL), h2 = cross(dL, N2 * in2);
        float a2[9];
#pragma unroll
        for (int j = 0; j < 9; j++) {
            d_ang[j] = -in1 * dot(h1, dN1[j]);
            a2[j] = -in2 * dot(h2, dN2[j]);
        }
        d_ang[9] = a2[3]; d_ang[10] = a2[4]; d_ang[11] = a2[5];
        d_ang[6] += a2[0]; d_ang[7] += a2[1]; d_ang[8] += a2[2];
        d_ang[0] += a2[6]; d_ang[1] += a2[7]; d_ang[2] += a2[8];
    }
    float d_curve[12];
#pragma unroll
    for (int j = 0; j < 12; j++) d_curve[j] = d_ang[j] * length;
    d_curve[0] -= dL.x * angle; d_curve[1] -= dL.y * angle; d_curve[2] -= dL.z * angle;
    d_curve[6] += dL.x * angle; d_curve[7] += dL.y * angle; d_curve[8] += dL.z * angle;
    f3 Fo[4] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
    // local and global area: one opposite node per triangle
    const float2 a0 = m.edge_area0[e];
    const float ca = -p.local_area_modulus * (n1 / a0.x - 1.0f) + area_coef;
    const float cb = -p.local_area_modulus * (n2 / a0.y - 1.0f) + area_coef;
    Fo[1] = Fo[1] + mk(dn1[3], dn1[4], dn1[5]) * ca;
    Fo[3] = Fo[3] + mk(dn2[3], dn2[4], dn2[5]) * cb;
    // volume
#pragma unroll
    for (int j = 0; j < 4; j++) Fo[j] = Fo[j] + mk(d_vol[3 * j], d_vol[3 * j + 1], d_vol[3 * j + 2]) * vol_coef;
    // membrane viscosity
    {
        const float4 va = m.U[q.x], vc = m.U[q.z];
        const f3 vca = mk(vc.x - va.x, vc.y - va.y, vc.z - va.z);
        const f3 vf = dL * (-dot(dL, vca) * 4.0f / 1.7320508075688772f * p.viscosity);
        Fo[2] = Fo[2] + vf;
        Fo[0] = Fo[0] - vf;
    }
    // bending
    const float4 ec = m.edge_const[e];
    const float kappa = angle * length / ((n1 + n2) * (1.0f / 6.0f));
    const float dk0 = ec.z - kappa;
    const float lbc = -0.5f * p.local_bending * dk0 * dk0;
    const float ccf = -p.global_bending * (kappa - ec.z) + mad_coef;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        Fo[j] = Fo[j] + mk(d_area[3 * j], d_area[3 * j + 1], d_area[3 * j + 2]) * lbc;
        Fo[j] = Fo[j] + mk(d_curve[3 * j], d_curve[3 * j + 1], d_curve[3 * j + 2]) * ccf;
    }
    // worm-like chain spring
    {
        const float x = length / ec.x;
        const float om = 1.0f - x;
        const float fc = -ec.y * (0.25f / (om * om) - 0.25f + x - p.spring_c);
        Fo[2] = Fo[2] + dL * fc;
        Fo[0] = Fo[0] - dL * fc;
    }
    m.edge_state[e] = make_float4(kappa, n1, n2, 0.0f);
#pragma unroll
    for (int j = 0; j < 4; j++) m.edge_force[4 * (size_t)e + j] = make_float4(Fo[j].x, Fo[j].y, Fo[j].z, 0.0f);
}

// eight lanes per node, each summing every eighth incidence; the eight
// partial sums are combined by a butterfly, so the order is fixed
__global__ __launch_bounds__(256) void k_mem_nodes(MemArgs m)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = gid >> 3, sub = gid & 7;
    const bool live = nd < m.n_nodes;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (live)
        for (int j = m.inc_start[nd] + sub; j < m.inc_start[nd + 1]; j += 8) {
            const float4 f = m.edge_force[m.inc[j]];
            fx += f.x; fy += f.y; fz += f.z;
        }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
    }
    if (live && sub == 0) m.F[nd] = make_float4(fx, fy, fz, 0.0f);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_ibm_interp(const MemArgs &m, const IbmArgs &g, const float4 *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_ibm_interp<true>, dim3((m.n_nodes + 3) / 4), dim3(256), 0, s, m, g, st);
    return hipGetLastError();
}
hipError_t launch_ibm_weights(const MemArgs &m, const IbmArgs &g, hipStream_t s)
{
    hipLaunchKernelGGL(k_ibm_interp<false>, dim3((m.n_nodes + 3) / 4), dim3(256), 0, s, m, g, nullptr);
    return hipGetLastError();
}
hipError_t launch_ibm_scan(const IbmArgs &g, hipStream_t s)
{
    const int blocks = (g.n_bins + 1023) / 1024;
    hipLaunchKernelGGL(k_ibm_scan1, dim3(blocks), dim3(1024), 0, s, g);
    hipLaunchKernelGGL(k_ibm_scan2, dim3(blocks), dim3(1024), 0, s, g);
    return hipGetLastError();
}
hipError_t launch_ibm_place(const MemArgs &m, const IbmArgs &g, hipStream_t s)
{
    hipLaunchKernelGGL(k_ibm_place, dim3((m.n_nodes + 255) / 256), dim3(256), 0, s, m, g);
    return hipGetLastError();
}
hipError_t launch_mem_globals(const MemArgs &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_partials, dim3(m.n_chunks), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_mem_globals, dim3(m.n_mem), dim3(64), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_mem_edges(const MemArgs &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_edges, dim3((m.n_edges + 127) / 128), dim3(128), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_mem_nodes(const MemArgs &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_nodes, dim3((8 * m.n_nodes + 255) / 256), dim3(256), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_ibm_spread(const MemArgs &m, const IbmArgs &g, hipStream_t s)
{
    if (g.binned)
        hipLaunchKernelGGL(k_ibm_spread, dim3(g.n_bins), dim3(256), 0, s, m, g);
    else
        hipLaunchKernelGGL(k_ibm_spread_nodes, dim3((m.n_nodes + 3) / 4), dim3(256), 0, s, m, g);
    hipLaunchKernelGGL(k_ibm_finalize, dim3(g.n_bins), dim3(64), 0, s, g);
    return hipGetLastError();
}

}  // namespace lbfs

namespace lbfs {

// ---------------------------------------------------------------------------
// internal viscosity (immersed_boundary_method.cu:389-527, Solver_gpu.cu:2153-2208)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mem_bbox(MemArgs m)
{
    const int *mem_node_start = m.mem_node_start;
    const int mem = blockIdx.x;
    __shared__ float lo[3][256], hi[3][256];
    float l[3] = {1e30f, 1e30f, 1e30f}, h[3] = {-1e30f, -1e30f, -1e30f};
    const float4 o = m.origin[mem];
    for (int n = mem_node_start[mem] + threadIdx.x; n < mem_node_start[mem + 1]; n += blockDim.x) {
        const float4 x = m.X[n];
        const float p[3] = {o.x + x.x, o.y + x.y, o.z + x.z};
        for (int a = 0; a < 3; a++) { l[a] = fminf(l[a], p[a]); h[a] = fmaxf(h[a], p[a]); }
    }
    for (int a = 0; a < 3; a++) { lo[a][threadIdx.x] = l[a]; hi[a][threadIdx.x] = h[a]; }
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp)
            for (int a = 0; a < 3; a++) {
                lo[a][threadIdx.x] = fminf(lo[a][threadIdx.x], lo[a][threadIdx.x + stp]);
                hi[a][threadIdx.x] = fmaxf(hi[a][threadIdx.x], hi[a][threadIdx.x + stp]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        m.bbox[2 * mem] = make_float4(lo[0][0], lo[1][0], lo[2][0], 0.0f);
        m.bbox[2 * mem + 1] = make_float4(hi[0][0], hi[1][0], hi[2][0], 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_internal_visc(MemArgs m, IbmArgs g, float delta_h, float *visc_factor)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n_cells) return;
    const float4 c = g.cen[i];
    const int cb = g.cell_block ? g.cell_block[i] : 0;
    float out = 1.0f;
    bool any = false;
    for (int mem = 0; mem < m.n_mem; mem++) {
        const float r = m.params[mem].visc_ratio;
        if (r == 1.0f || m.mem_block[mem] != cb) continue;
        const float4 lo = m.bbox[2 * mem], hi = m.bbox[2 * mem + 1];
        const float xl = lo.x - 2.0f * delta_h, xh = hi.x + 2.0f * delta_h;
        if (c.x < xl || c.x > xh || c.y < lo.y - 2.0f * delta_h || c.y > hi.y + 2.0f * delta_h ||
            c.z < lo.z - 2.0f * delta_h || c.z > hi.z + 2.0f * delta_h)
            continue;
        const float4 o = m.origin[mem];
        const float pc0 = c.x * c.y - xh * c.y, pc1 = c.x * c.z - xh * c.z, pc2 = c.x - xh;
        int count = 0;
        float dmin = 1e9f;
        for (int t = m.mem_tri_start[mem]; t < m.mem_tri_start[mem + 1]; t++) {
            const int4 q = m.tri[t];
            const int ids[3] = {q.x, q.y, q.z};
            float p[3][3];
            for (int k = 0; k < 3; k++) {
                const float4 x = m.X[ids[k]];
                p[k][0] = o.x + x.x; p[k][1] = o.y + x.y; p[k][2] = o.z + x.z;
            }
            if (!(p[0][0] > c.x && p[1][0] > c.x && p[2][0] > c.x)) continue;
            float side[3];
            for (int k = 0; k < 3; k++) {
                const float *u = p[k], *v = p[(k + 1) % 3];
                const float e0 = u[1] * v[2] - v[1] * u[2], e1 = u[2] - v[2], e2 = v[1] - u[1];
                side[k] = pc0 * e1 + pc1 * e2 + pc2 * e0;
            }
            if ((side[0] < 0 && side[1] < 0 && side[2] < 0) || (side[0] > 0 && side[1] > 0 && side[2] > 0)) count++;
            for (int k = 0; k < 3; k++) {
                const float dx = c.x - p[k][0], dy = c.y - p[k][1], dz = c.z - p[k][2];
                dmin = fminf(dmin, sqrtf(dx * dx + dy * dy + dz * dz));
            }
        }
        const float factor = 2.0f * fminf(0.5f, dmin * 0.5f / delta_h + 1.0f / PI_F * sinf(PI_F * dmin / delta_h) * 0.5f);
        const float v = 0.5f * ((count & 1) ? (r + 1.0f) + (r - 1.0f) * factor : (r + 1.0f) + (1.0f - r) * factor);
        // several membranes: keep the value furthest from 1
        if (!any || fabsf(v - 1.0f) > fabsf(out - 1.0f)) out = v;
        any = true;
    }
    visc_factor[i] = out;
}

// Output-step energies of each membrane (one workgroup per membrane, FP64
// sums over its edges): WLC springs (get_spring_energy, spring_network.cu:1721-1768,
// halved as there), local area (get_area_energy :1687-1698, scaled by the
// global area modulus as there), local bending (get_bending_energy :1703-1715,
// edge e divided by entry e of the flat |N| array as the launch passes it),
// then global area, volume and ADE from the globals (Solver_gpu.cu:1060-1139).
// out: spring, local area, local bending, global area, volume, ADE, total.
__global__ __launch_bounds__(256) void k_mem_energy(MemArgs m, double *out)
{
    const int mem = blockIdx.x;
    const MemParamDev p = m.params[mem];
    const double x0 = p.wlc_x0;
    const double c1 = 0.25 / (1.0 - x0) / (1.0 - x0) - 0.25 + x0;
    const double c2 = x0 * x0 * (3.0 - 2.0 * x0) / (1.0 - x0) - 4.0 * c1 * x0;
    __shared__ double sh[3][256];
    double es = 0.0, ela = 0.0, elb = 0.0;
    const int e0 = m.mem_edge_start[mem], e1 = m.mem_edge_start[mem + 1];
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int4 q = m.edges[e];
        const f3 l = ld3(m.X[q.z]) - ld3(m.X[q.x]);
        const float4 kc = m.edge_const[e];
        const float4 st = m.edge_state[e];
        const float2 a0 = m.edge_area0[e];
        const double Lm = kc.x, x = (double)sqrtf(dot(l, l)) / Lm;
        es += 0.5 * (double)kc.y * fabs(0.25 * Lm * (x * x * (3.0 - 2.0 * x) / (1.0 - x) - 4.0 * x * c1 - c2));
        const double d1 = (double)st.y - a0.x, d2 = (double)st.z - a0.y;
        ela += (double)p.global_area_modulus * 0.5 * (d1 * d1 / a0.x + d2 * d2 / a0.y);
        const int flat = e - e0;                       // entry of the membrane's 2E |N| array
        const float4 sf = m.edge_state[e0 + (flat >> 1)];
        const double nflat = (flat & 1) ? sf.z : sf.y;
        const double dk = (double)st.x - kc.z;
        elb += 0.5 * (double)p.local_bending * dk * dk / nflat;
    }
    sh[0][threadIdx.x] = es;
    sh[1][threadIdx.x] = ela;
    sh[2][threadIdx.x] = elb;
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + stp];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double *gl = m.glob + 8 * (size_t)mem;
        const double dv = gl[0] - gl[3], da = gl[1] - gl[4], dm = gl[2] - gl[5];
        double *o = out + 7 * (size_t)mem;
        o[0] = sh[0][0];
        o[1] = sh[1][0];
        o[2] = sh[2][0];
        o[3] = 0.5 * (double)p.global_area_modulus * da * da / gl[4];
        o[4] = 0.5 * (double)p.volume_modulus * dv * dv / gl[3];
        o[5] = 0.5 * (double)p.global_bending * dm * dm / gl[4];
        o[6] = o[0] + o[1] + o[2] + o[3] + o[4] + o[5];
    }
}

hipError_t launch_mem_energy(const MemArgs &m, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_energy, dim3(m.n_mem), dim3(256), 0, s, m, out);
    return hipGetLastError();
}

hipError_t launch_internal_visc(const MemArgs &m, const IbmArgs &g, float delta_h, float *visc_factor, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_bbox, dim3(m.n_mem), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_internal_visc, dim3((g.n_cells + 255) / 256), dim3(256), 0, s, m, g, delta_h, visc_factor);
    return hipGetLastError();
}

}  // namespace lbfs
