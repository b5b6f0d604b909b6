// Immersed-boundary coupling and spring-network membrane kernels for gfx950.
//
//   k_ibm_interp   per node (one per lane): fluid velocity at the node
//                  (immersed_boundary_method.cu:14-97 or the 4-point kernels),
//                  explicit advection X += U dt (immersed_boundary_method.cu:234-264),
//                  the advected node listed in its fluid tile
//   k_mem_volume   per triangle: volume and area sums (lagrangian_object.cpp:601-650)
//                  in 64-bit fixed point, wavefront-reduced
//   k_mem_edges    per spring: all membrane force terms (spring_network.cu:613-971),
//                  written per edge and slot (no atomics); the MAD sum of the next
//                  force call (spring_network.cu:1186-1238)
//   k_mem_nodes    per node: sum of its edge slots in a fixed order, combined
//                  with the membrane's global coefficients
//   k_ibm_spread_tiles
//                  per fluid tile: the forces of the nodes listed around the tile
//                  spread onto its cells (immersed_boundary_method.cu:947-1024) in
//                  LDS, summed in 64-bit fixed point so the value does not depend
//                  on the order the nodes are visited in
//   k_ibm_apply    per fluid tile: the spread force density added to the updated state
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "coupling.hpp"
#include "ibm_device.hpp"
#include "membrane_force.hpp"
#include "xcd.hpp"

namespace lbfs {

// Add one value per lane to the membrane sums (64-bit fixed point, so the
// result does not depend on the order): one atomic per wavefront when its
// lanes belong to one membrane (mem < 0: no value), lane by lane otherwise.
__device__ __forceinline__ void mem_sum_add(unsigned long long *acc, int mem, const long long v[2], int nv)
{
    int m0 = mem;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m0 = max(m0, __shfl_xor(m0, off, 64));
    if (m0 < 0) return;
    if (__all(mem < 0 || mem == m0)) {
        long long s[2] = {v[0], v[1]};
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int k = 0; k < 2; k++) s[k] += __shfl_xor(s[k], off, 64);
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < nv; k++) atomicAdd(acc + 4 * (size_t)m0 + k, (unsigned long long)s[k]);
    } else if (mem >= 0) {
        for (int k = 0; k < nv; k++) atomicAdd(acc + 4 * (size_t)mem + k, (unsigned long long)v[k]);
    }
}

__device__ __forceinline__ long long mem_fx(double x, int *err)
{
    return (long long)fix_term(x, MEM_SUM_SCALE, err, ERR_FIX_MEM);
}

// volume and area of triangle t (thread t) added to its membrane's sums
__device__ __forceinline__ void triangle_sums(const MemArgs &m, int t)
{
    int tm = -1;
    long long v[2] = {0, 0};
    if (t < m.n_tri) {
        const int4 qq = m.tri[t];
        const f3 a = ld3(m.X[qq.x]), b = ld3(m.X[qq.y]), c = ld3(m.X[qq.z]);
        const f3 N = cross(b - a, c - a);
        v[0] = mem_fx((double)dot(N, (a + b + c) * (1.0f / 3.0f)) / 6.0, m.err);
        v[1] = mem_fx(0.5 * (double)sqrtf(dot(N, N)), m.err);
        tm = qq.w;
    }
    mem_sum_add(m.mem_acc, tm, v, 2);
}

// the membrane's global coefficients in FP64 from the running volume and area
// sums and the MAD of this force call, rounded once to the working precision
__device__ __forceinline__ void mem_coefs(const MemArgs &m, int mem, float &vol, float &area, float &mad)
{
    const double *gl = m.glob + 8 * (size_t)mem;
    const unsigned long long *acc = m.mem_acc + 4 * (size_t)mem;
    const double V = (double)(long long)acc[0] * (1.0 / MEM_SUM_SCALE), A = (double)(long long)acc[1] * (1.0 / MEM_SUM_SCALE);
    const MemParamDev &p = m.params[mem];
    vol = (float)(-(double)p.volume_modulus * (V / gl[3] - 1.0) / 6.0);
    area = (float)(-(double)p.global_area_modulus * (A / gl[4] - 1.0));
    mad = (float)(-2.0 * (double)p.global_bending * (gl[7] - gl[5]) / gl[4]);
}


// ---------------------------------------------------------------------------
// IBM stencils.  The stencil of a node at P is the S x S x S block of cells
// from its base index (S = 2 for the 2-point kernel, 4 for the others);
// indices wrap on periodic axes, cells beyond a wall drop out (weight 0).
// Each node's normalised weights are computed once per step, at its advected
// position, and kept (64 floats, lane order i + S j + S^2 k): the spreading
// of that step and the interpolation of the next read them.
// ---------------------------------------------------------------------------
template <int K>
struct Stencil {
    static constexpr int S = K == LBFS_IBM_LINEAR2 ? 2 : 4;
};

// unnormalised weight of stencil cell (i, j, k) of a node at P (grid
// coordinate xr, base index base); w: the wrapped cell index, c its centroid
template <int K>
__device__ __forceinline__ float stencil_weight(const IbmArgs &g, const IbmBlock &B, const float P[3], const float xr[3],
                                                const int base[3], int i, int j, int k, const int w[3], const float4 c)
{
    if constexpr (K == LBFS_IBM_LINEAR2) {
        float wt = 1.0f;
        wt *= linear2_factor(base[0] + i, xr[0]);
        wt *= linear2_factor(base[1] + j, xr[1]);
        wt *= linear2_factor(base[2] + k, xr[2]);
        return wt;
    } else if constexpr (K == LBFS_IBM_LAGRANGE4) {
        return (float)(lagrange4(g, B, 0, base[0], i, (double)P[0]) * lagrange4(g, B, 1, base[1], j, (double)P[1]) *
                       lagrange4(g, B, 2, base[2], k, (double)P[2]));
    } else {
        const int d[3] = {base[0] + i - w[0], base[1] + j - w[1], base[2] + k - w[2]};
        return cos4_cell(B, c, d, P);
    }
}

__device__ __forceinline__ int node_tile(const IbmBlock &B, const int base[3])
{
    return B.tile0 + tile_coord(B, 0, base[0]) + B.nt[0] * (tile_coord(B, 1, base[1]) + B.nt[1] * tile_coord(B, 2, base[2]));
}

// sum over the L lanes of a node's lane group (aligned groups of L lanes,
// butterfly order)
template <int L>
__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Per node, S x S lanes, lane (j, k) taking the S cells of stencil row (j, k).
// INTERP: the fluid velocity at the node from its kept weights
// (immersed_boundary_method.cu:14-97, or the 4-point kernels) and the
// advection X += U dt (immersed_boundary_method.cu:234-264).  Always: the
// weights of the (new) position, normalised for the cosine kernel, kept with
// the stencil base and its fluid tile; INTERP also lists the node in that
// tile for the spreading.
template <int K, bool INTERP>
__global__ __launch_bounds__(256) void k_ibm_interp(MemArgs m, IbmArgs g, const float4 *__restrict__ st)
{
    constexpr int S = Stencil<K>::S, L = S * S;
    // consecutive nodes (neighbours on a membrane) on one XCD share its L2
    const int gid = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int r = gid % L, j = r % S, k = r / S;
    const bool live = gid / L < m.n_nodes;
    const int nd = live ? gid / L : m.n_nodes - 1;  // dead lanes shadow a live node for the group sums
    const IbmBlock B = node_block(g, nd);
    const float4 o = m.node_origin[nd];
    float4 xl = m.X[nd];
    float *row = g.wnode + 64 * (size_t)nd + S * r;  // this lane's S weights
    if constexpr (INTERP) {
        const int4 nb = g.nbase[nd];
        float ux = 0.0f, uy = 0.0f, uz = 0.0f;
        int w[3] = {0, 0, 0};
        const bool rowok = wrap_index(B, 1, nb.y + j, w[1]) && wrap_index(B, 2, nb.z + k, w[2]);
        // the row's cell indices first, then its states and kept weights (a
        // cell off a wall loads cell 0, unused), so the loads are in flight
        // together rather than one cell at a time
        bool ok[S];
        int ci[S];
#pragma unroll
        for (int i = 0; i < S; i++) {
            ok[i] = rowok && wrap_index(B, 0, nb.x + i, w[0]);
            ci[i] = ok[i] ? cell_internal(g, B, w) : 0;
        }
        float4 sv[S];
        float rw[S];
#pragma unroll
        for (int i = 0; i < S; i++) {
            sv[i] = st[ci[i]];
            rw[i] = row[i];
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            if (!ok[i]) continue;
            ux = fmaf(rw[i], sv[i].y, ux);
            uy = fmaf(rw[i], sv[i].z, uy);
            uz = fmaf(rw[i], sv[i].w, uz);
        }
        ux = group_sum<L>(ux);
        uy = group_sum<L>(uy);
        uz = group_sum<L>(uz);
        xl.x += ux * g.dt;
        xl.y += uy * g.dt;
        xl.z += uz * g.dt;
        if (live && r == 0) {
            m.U[nd] = make_float4(ux, uy, uz, 0.0f);
            m.X[nd] = xl;
        }
    }
    const float P[3] = {o.x + xl.x, o.y + xl.y, o.z + xl.z};
    float xr[3];
    int base[3];
    node_base(g, B, P, xr, base);
    float wt[S];
    float ws = 0.0f;
    {
        int w[3] = {0, 0, 0};
        const bool rowok = wrap_index(B, 1, base[1] + j, w[1]) && wrap_index(B, 2, base[2] + k, w[2]);
        // the row's wrapped indices, then (cosine kernel) every centroid at
        // once, then the weights
        bool ok[S];
        int wx[S];
#pragma unroll
        for (int i = 0; i < S; i++) {
            wx[i] = 0;
            ok[i] = rowok && wrap_index(B, 0, base[0] + i, wx[i]);
        }
        float4 c[S];
#pragma unroll
        for (int i = 0; i < S; i++) {
            c[i] = float4{};
            if constexpr (K == LBFS_IBM_COS4) {
                const int wi[3] = {wx[i], w[1], w[2]};
                c[i] = g.cen[ok[i] ? cell_internal(g, B, wi) : 0];
            }
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            wt[i] = 0.0f;
            if (!ok[i]) continue;
            const int wi[3] = {wx[i], w[1], w[2]};
            wt[i] = stencil_weight<K>(g, B, P, xr, base, i, j, k, wi, c[i]);
            ws += wt[i];
        }
    }
    if constexpr (K == LBFS_IBM_COS4) {
        ws = group_sum<L>(ws);
        const float inv = ws > 0.0f ? 1.0f / ws : 0.0f;
#pragma unroll
        for (int i = 0; i < S; i++) wt[i] *= inv;
    }
    if (live) {
        if constexpr (S == 4) {
            *reinterpret_cast<float4 *>(row) = make_float4(wt[0], wt[1], wt[2], wt[3]);
        } else {
            *reinterpret_cast<float2 *>(row) = make_float2(wt[0], wt[1]);
        }
    }
    const int key = node_tile(B, base);
    if (live && r == 0) g.nbase[nd] = make_int4(base[0], base[1], base[2], key);
    if (INTERP && g.binned != IBM_SPREAD_NODES) {
        // listed in its tile, one atomic per tile among the wavefront's nodes
        // (the bin scattering also lists the tile once); a full tile sends the
        // node to the overflow list.  Every lane takes part in the shuffles.
        const int lane = threadIdx.x & 63;
        const int mykey = live && r == 0 ? key : -1;
        int rank = 0, cnt = 0, first = lane;
#pragma unroll
        for (int p = 0; p < 64; p += L) {
            const int kp = __shfl(mykey, p, 64);
            if (kp == mykey) {
                cnt++;
                if (p < lane) rank++;
                first = min(first, p);
            }
        }
        int base_q = 0;
        if (mykey >= 0 && first == lane) base_q = atomicAdd(&g.tile_count[key], cnt);
        const int q = __shfl(base_q, first, 64) + rank;
        if (mykey < 0) return;
        if (q == 0 && g.binned == IBM_SPREAD_BINS) {
            const int li = atomicAdd(g.n_bin_list, 1);
            if (li < g.n_tiles) g.bin_list[li] = key;
        }
        if (q < g.tile_cap) {
            g.tile_slots[(size_t)key * g.tile_cap + q] = nd;
        } else {
            const int e = atomicAdd(g.n_overflow, 1);
            if (e < m.n_nodes) g.overflow[e] = nd;
        }
    }
}

// ---------------------------------------------------------------------------
// IBM: spreading (immersed_boundary_method.cu:947-1024).  Tile spreading: one
// workgroup per fluid tile gathers the nodes whose stencils reach it (those
// listed in the tiles holding stencil bases from S - 1 cells before the tile
// to its end, and the overflow list), keeps those that reach it, and spreads
// them one wavefront each (one lane per stencil cell, the node's kept
// weights) into per-cell sums of the tile in LDS; it then writes the tile's
// force densities (no global atomics).  Per-node spreading: one wavefront
// per node straight into per-cell accumulators in global memory.  All sums
// are 64-bit fixed-point integer sums, so their value depends neither on the
// order of the nodes nor on the path: both give the same bits.
// ---------------------------------------------------------------------------
// sum of a node's edge slots by eight lanes (sub = lane & 7; 8 or more: no
// slots), fixed order
__device__ __forceinline__ float4 node_force8(const MemArgs &m, int nd, int sub)
{
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (sub < 8) {
        // two incidences per pass, their loads in flight together (the sum
        // keeps the order j, j + 8, j + 16, ...)
        const int j1 = m.inc_start[nd + 1];
        for (int j = m.inc_start[nd] + sub; j < j1; j += 16) {
            const bool two = j + 8 < j1;
            const int e0 = m.inc[j], e1 = m.inc[two ? j + 8 : j];
            const float4 f0 = m.edge_force[e0], f1 = m.edge_force[e1];
            fx += f0.x; fy += f0.y; fz += f0.z;
            if (two) { fx += f1.x; fy += f1.y; fz += f1.z; }
        }
    }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
    }
    return make_float4(fx, fy, fz, 0.0f);
}

// one node by one wavefront (lane = stencil cell) into the global accumulators
template <int K>
__device__ __forceinline__ void spread_node_global(const MemArgs &m, const IbmArgs &g, int nd, int lane, float4 F)
{
    constexpr int S = Stencil<K>::S;
    if (lane >= S * S * S) return;
    const float wt = g.wnode[64 * (size_t)nd + lane];
    if (wt == 0.0f) return;
    const IbmBlock B = node_block(g, nd);
    const int4 nb = g.nbase[nd];
    int off[3], w[3];
    lane_offset(S, lane, off);
    if (!(wrap_index(B, 0, nb.x + off[0], w[0]) && wrap_index(B, 1, nb.y + off[1], w[1]) &&
          wrap_index(B, 2, nb.z + off[2], w[2])))
        return;
    const int ic = cell_internal(g, B, w);
    unsigned long long *acc = g.acc + 3 * (size_t)ic;
    atomicAdd(acc, fix_term_f(F.x * wt, g.err, ERR_FIX_IBM));
    atomicAdd(acc + 1, fix_term_f(F.y * wt, g.err, ERR_FIX_IBM));
    atomicAdd(acc + 2, fix_term_f(F.z * wt, g.err, ERR_FIX_IBM));
    g.tile_acc[g.cell_tile[ic]] = 1;
}

// offsets along one axis whose (wrapped) cell lies in [lo, lo + dim), and
// their local index
template <int S>
__device__ __forceinline__ unsigned axis_in(const IbmBlock &B, int a, int base, int lo, int dim, int loc[S])
{
    unsigned in = 0u;
#pragma unroll
    for (int o = 0; o < S; o++) {
        int w = 0;
        const bool ok = wrap_index(B, a, base + o, w);
        loc[o] = w - lo;
        if (ok && loc[o] >= 0 && loc[o] < dim) in |= 1u << o;
    }
    return in;
}

template <int K>
__global__ __launch_bounds__(SPREAD_NT) void k_ibm_spread_tiles(MemArgs m, IbmArgs g)
{
    constexpr int S = Stencil<K>::S, C = S * S * S, NPW = 64 / C, NW = SPREAD_NT / 64, NS = S * S * S;
    extern __shared__ int4 tile_lds[];  // candidates (bases, forces, ids), then 3 x max_tile_cells sums
    __shared__ int s_tile[NS], s_pre[NS + 1], s_total, s_nov, s_nk, s_tc[3][S], s_ntc[3];
    const int T = xcd_remap(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int4 tl = g.tile_lo[T];
    const IbmBlock B = box(g, tl.w);
    const int lo[3] = {tl.x, tl.y, tl.z};
    const int dim[3] = {min(B.tw[0], B.n[0] - lo[0]), min(B.tw[1], B.n[1] - lo[1]), min(B.tw[2], B.n[2] - lo[2])};
    if (tid == 0) {
        // tiles holding the bases of nodes that can reach this tile: per axis
        // those of the cells lo - (S - 1) ... lo (wrapped, or clamped at a wall
        // as the node tiles are); constant axis indices keep the box in registers
#pragma unroll
        for (int a = 0; a < 3; a++) {
            int n = 0;
#pragma unroll
            for (int o = S - 1; o >= 0; o--) {
                int w = lo[a] - o;
                if (w < 0) w = B.per[a] ? ((w % B.n[a]) + B.n[a]) % B.n[a] : 0;
                const int t = w / B.tw[a];
                bool seen = false;
                for (int q = 0; q < n; q++) seen |= s_tc[a][q] == t;
                if (!seen) s_tc[a][n++] = t;
            }
            s_ntc[a] = n;
        }
    }
    __syncthreads();
    if (wave == 0) {
        // one lane per neighbour tile: its node count, then a scan for the
        // candidates' offsets
        const int nx = s_ntc[0], ny = s_ntc[1], ncomb = nx * ny * s_ntc[2];
        int c = 0, id = 0;
        if (lane < ncomb) {
            const int x = lane % nx, y = (lane / nx) % ny, z = lane / (nx * ny);
            id = B.tile0 + s_tc[0][x] + B.nt[0] * (s_tc[1][y] + B.nt[1] * s_tc[2][z]);
            c = min(g.tile_count[id], g.tile_cap);
        }
        int incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane < ncomb) {
            s_tile[lane] = id;
            s_pre[lane] = incl - c;
        }
        if (lane == ncomb - 1) {
            s_pre[ncomb] = incl;
            s_total = incl;
        }
        if (lane == 0) s_nov = min(*g.n_overflow, m.n_nodes);
    }
    __syncthreads();
    const int tot = s_total, nov = s_nov;
    if (tot + nov == 0) {  // uniform over the workgroup
        if (tid == 0) g.tile_now[T] = 0;
        return;
    }
    const int cap = g.max_tile_cells;
    const int nct = dim[0] * dim[1] * dim[2];
    int4 *cbase = tile_lds;                                              // [SPREAD_NT] stencil bases
    float4 *cF = reinterpret_cast<float4 *>(cbase + SPREAD_NT);          // [SPREAD_NT] forces
    int *cand = reinterpret_cast<int *>(cF + SPREAD_NT);                 // [SPREAD_NT] node ids
    unsigned long long *tacc = reinterpret_cast<unsigned long long *>(cand + SPREAD_NT);  // 3 x cap sums
    for (int i = tid; i < nct; i += SPREAD_NT) tacc[i] = tacc[cap + i] = tacc[2 * cap + i] = 0ull;
    for (int q0 = 0; q0 < tot + nov; q0 += SPREAD_NT) {
        if (tid == 0) s_nk = 0;
        __syncthreads();  // the previous chunk is spread
        const int q = q0 + tid;
        if (q < tot + nov) {
            int nd;
            if (q < tot) {
                int k = 0;
                while (q >= s_pre[k + 1]) k++;
                nd = g.tile_slots[(size_t)s_tile[k] * g.tile_cap + (q - s_pre[k])];
            } else {
                nd = g.overflow[q - tot];
            }
            if (q < tot || g.node_block[nd] == tl.w) {
                const int4 nb = g.nbase[nd];
                int loc[S];
                const unsigned ix = axis_in<S>(B, 0, nb.x, lo[0], dim[0], loc);
                const unsigned iy = axis_in<S>(B, 1, nb.y, lo[1], dim[1], loc);
                const unsigned iz = axis_in<S>(B, 2, nb.z, lo[2], dim[2], loc);
                if (ix && iy && iz) {
                    // keep it: the spreading pass below reads it from LDS
                    const int e = atomicAdd(&s_nk, 1);
                    cand[e] = nd;
                    cbase[e] = nb;
                    cF[e] = m.F[nd];
                }
            }
        }
        __syncthreads();
        const int nk = s_nk;
        int off[3];
        lane_offset(S, lane % C, off);
        // UN nodes per wavefront and pass: their weight loads in flight together
        constexpr int UN = 4;
        for (int e0 = wave * NPW * UN; e0 < nk; e0 += NW * NPW * UN) {
            float wt[UN];
#pragma unroll
            for (int u = 0; u < UN; u++) {
                const int e = e0 + u * NPW + lane / C;
                wt[u] = e < nk ? g.wnode[64 * (size_t)cand[e] + lane % C] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < UN; u++) {
                if (wt[u] == 0.0f) continue;
                const int e = e0 + u * NPW + lane / C;
                const int4 nb = cbase[e];
                int w[3];
                if (!(wrap_index(B, 0, nb.x + off[0], w[0]) && wrap_index(B, 1, nb.y + off[1], w[1]) &&
                      wrap_index(B, 2, nb.z + off[2], w[2])))
                    continue;
                const int l0 = w[0] - lo[0], l1 = w[1] - lo[1], l2 = w[2] - lo[2];
                if (l0 < 0 || l0 >= dim[0] || l1 < 0 || l1 >= dim[1] || l2 < 0 || l2 >= dim[2]) continue;
                const int li = l0 + dim[0] * (l1 + dim[1] * l2);
                const float4 F = cF[e];
                atomicAdd(tacc + li, fix_term_f(F.x * wt[u], g.err, ERR_FIX_IBM));
                atomicAdd(tacc + cap + li, fix_term_f(F.y * wt[u], g.err, ERR_FIX_IBM));
                atomicAdd(tacc + 2 * cap + li, fix_term_f(F.z * wt[u], g.err, ERR_FIX_IBM));
            }
        }
    }
    __syncthreads();
    const int c0 = g.tile_cell_start[T];
    int any = 0;
    for (int i = tid; i < nct; i += SPREAD_NT) {
        const unsigned long long v[3] = {tacc[i], tacc[cap + i], tacc[2 * cap + i]};
        const float sc = g.spread_sc[c0 + i];
#pragma unroll
        for (int c = 0; c < 3; c++)
            g.spread_f[3 * (size_t)(c0 + i) + c] = (float)((double)(long long)v[c] * (1.0 / FIX_SCALE)) * sc;
        any |= (v[0] | v[1] | v[2]) != 0ull;
    }
    any = __syncthreads_or(any);
    if (tid == 0) g.tile_now[T] = any;
}

// Bin scattering: the workgroups loop over work items of up to 32 nodes of
// the listed bins (IBM_BIN_CAP / 32 items per bin, empty ones skipped), so a
// full bin is shared by several workgroups.  Each item gathers its nodes'
// forces from the edge records (eight lanes per node, the fixed order of
// the node-force pass, which this path replaces: every node is in one bin
// slot or on the overflow list) and writes them out, spreads its nodes (one
// wavefront each, one lane per stencil cell, the kept weights) into an LDS
// window of the bin and the S - 1 cells beyond its upper faces, then adds the
// window's nonzero sums to the per-cell accumulators.
constexpr int BIN_NT = 512, BIN_CH = 32;

template <int K>
__global__ __launch_bounds__(BIN_NT) void k_ibm_spread_bins(MemArgs m, IbmArgs g)
{
    constexpr int S = Stencil<K>::S, C = S * S * S, NPW = 64 / C, NW = BIN_NT / 64;
    constexpr int NCH = IBM_BIN_CAP / BIN_CH, WD = 4 + S - 1, NWIN = WD * WD * WD;
    constexpr int CWN = (C * BIN_CH + BIN_NT - 1) / BIN_NT;  // weight loads per thread and item
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ unsigned long long racc[3][NWIN];
    __shared__ int cnd[BIN_CH];
    __shared__ int4 cbase[BIN_CH];
    __shared__ float4 cF[BIN_CH];
    __shared__ float cw[BIN_CH][C];  // the item's kept weights, one row per node
    int off[3];
    lane_offset(S, lane % C, off);
    const int n_items = min(*g.n_bin_list, g.n_tiles) * NCH;  // uniform over the workgroup
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int b = g.bin_list[it / NCH];
        const int cnt_all = min(g.tile_count[b], IBM_BIN_CAP);
        const int c_begin = (it % NCH) * BIN_CH, cnt = min(cnt_all, c_begin + BIN_CH) - c_begin;
        if (cnt <= 0) continue;  // uniform over the workgroup
        const int4 tl = g.tile_lo[b];
        const IbmBlock B = box(g, tl.w);
        const int bo[3] = {tl.x, tl.y, tl.z};
        __syncthreads();  // the previous item's window is flushed
        for (int q = tid; q < 3 * NWIN; q += BIN_NT) racc[q / NWIN][q % NWIN] = 0ull;
        if (tid < cnt) {
            const int p = g.tile_slots[(size_t)b * IBM_BIN_CAP + c_begin + tid];
            cnd[tid] = p;
            cbase[tid] = g.nbase[p];
        }
        __syncthreads();
        // all weight rows of the item in flight at once (one coalesced row
        // per node), held in registers while the node forces are gathered
        float wv[CWN];
#pragma unroll
        for (int u = 0; u < CWN; u++) {
            const int q = tid + u * BIN_NT;
            wv[u] = q < C * cnt ? g.wnode[64 * (size_t)cnd[q / C] + q % C] : 0.0f;
        }
        if (tid < 8 * BIN_CH) {  // whole wavefronts: the eight-lane sums shuffle
            const int jn = tid >> 3;
            const bool ok = jn < cnt;
            const float4 F = node_force8(m, ok ? cnd[jn] : 0, ok ? (tid & 7) : 8);
            if (ok && (tid & 7) == 0) {
                cF[jn] = F;
                m.F[cnd[jn]] = F;
            }
        }
#pragma unroll
        for (int u = 0; u < CWN; u++) {
            const int q = tid + u * BIN_NT;
            if (q < C * cnt) cw[q / C][q % C] = wv[u];
        }
        __syncthreads();
        for (int jn = wave * NPW + lane / C; jn < cnt; jn += NW * NPW) {
            const float wt = cw[jn][lane % C];
            if (wt == 0.0f) continue;
            const int4 nb = cbase[jn];
            const int bs[3] = {nb.x, nb.y, nb.z};
            int loc = 0, mul = 1;
            bool in = true;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                int l;
                if (B.per[a]) {
                    // window coordinate of the wrapped base (nodes advect freely
                    // across periodic boundaries: the base may lie periods away)
                    l = ((bs[a] % B.n[a]) + B.n[a]) % B.n[a] - bo[a] + off[a];
                } else {
                    const int c = bs[a] + off[a];
                    in = in && c >= 0 && c < B.n[a];
                    l = c - bo[a];
                }
                in = in && l >= 0 && l < WD;
                loc += l * mul;
                mul *= WD;
            }
            if (!in) continue;
            const float4 F = cF[jn];
            atomicAdd(&racc[0][loc], fix_term_f(F.x * wt, g.err, ERR_FIX_IBM));
            atomicAdd(&racc[1][loc], fix_term_f(F.y * wt, g.err, ERR_FIX_IBM));
            atomicAdd(&racc[2][loc], fix_term_f(F.z * wt, g.err, ERR_FIX_IBM));
        }
        __syncthreads();
        for (int q = tid; q < NWIN; q += BIN_NT) {
            const unsigned long long v0 = racc[0][q], v1 = racc[1][q], v2 = racc[2][q];
            if ((v0 | v1 | v2) == 0ull) continue;
            const int l[3] = {q % WD, (q / WD) % WD, q / (WD * WD)};
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
            g.tile_acc[g.cell_tile[ic]] = 1;
        }
    }
    // nodes of full bins: one wavefront each (its force from the edge records
    // by lanes 0-7), straight into the accumulators
    const int nov = min(*g.n_overflow, m.n_nodes);
    for (int e = blockIdx.x * NW + wave; e < nov; e += gridDim.x * NW) {
        const int nd = g.overflow[e];
        float4 F = node_force8(m, nd, lane < 8 ? lane : 8);
        F = make_float4(__shfl(F.x, 0, 64), __shfl(F.y, 0, 64), __shfl(F.z, 0, 64), 0.0f);
        if (lane == 0) m.F[nd] = F;
        spread_node_global<K>(m, g, nd, lane, F);
    }
}

// per-node spreading of every node: one wavefront each
template <int K>
__global__ __launch_bounds__(256) void k_ibm_spread_nodes(MemArgs m, IbmArgs g)
{
    const int nd = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (nd >= m.n_nodes) return;
    spread_node_global<K>(m, g, nd, threadIdx.x & 63, m.F[nd]);
}

// One workgroup per fluid tile: the tile's spread force densities (written
// by the tile spreading, or from the per-node accumulators, which are
// cleared) added to the state the face pass updated with the static force
// (rho u += dt f, the update's own form: u_new = u* + dt f / rho_new, rho_new
// unchanged by the force) and written to the cell force output.  Clears the
// tile's node count for the next step.
__global__ __launch_bounds__(256) void k_ibm_apply(IbmArgs g, ApplyArgs p)
{
    const int t = xcd_remap(blockIdx.x, gridDim.x), tid = threadIdx.x;
    const bool gather = g.binned == IBM_SPREAD_GATHER;
    if (t == 0) {
        // the listed bins' counts (the next interpolation lists the nodes anew)
        if (g.binned == IBM_SPREAD_BINS) {
            const int nl = min(*g.n_bin_list, g.n_tiles);
            for (int e = tid; e < nl; e += blockDim.x) g.tile_count[g.bin_list[e]] = 0;
        }
        __syncthreads();
        if (tid == 0 && g.binned != IBM_SPREAD_NODES) {
            *g.n_overflow = 0;
            *g.n_bin_list = 0;
        }
        // the step's membrane sums: V, A, the MAD the step used, A as the next
        // A_prev and the MAD of the next force call (numerator / A)
        for (int k = tid; k < p.n_mem; k += blockDim.x) {
            unsigned long long *s = p.mem_acc + 4 * (size_t)k;
            double *gl = p.glob + 8 * (size_t)k;
            const double V = (double)(long long)s[0] * (1.0 / MEM_SUM_SCALE), A = (double)(long long)s[1] * (1.0 / MEM_SUM_SCALE);
            const double M = (double)(long long)s[2] * (1.0 / MEM_SUM_SCALE);
            gl[0] = V;
            gl[1] = A;
            gl[2] = gl[7];
            gl[6] = A;
            gl[7] = M / A;
            s[0] = s[1] = s[2] = 0ull;
        }
    }
    if (tid == 0 && gather) g.tile_count[t] = 0;
    const int now = gather ? g.tile_now[t] : g.tile_acc[t], prev = g.tile_prev[t];
    if (!(now | prev | p.all)) return;  // uniform over the workgroup
    int any = 0;
    for (int i = p.tile_cell_start[t] + tid; i < p.tile_cell_start[t + 1]; i += blockDim.x) {
        float f[3] = {0.0f, 0.0f, 0.0f};
        // the cell's state, static force and (per-cell sums) scale loaded
        // with its spread force rather than after it
        const float4 s0 = p.st[i];
        float sfc[3];
#pragma unroll
        for (int c = 0; c < 3; c++) sfc[c] = g.static_force[3 * (size_t)i + c];
        if (now) {
            if (gather) {
#pragma unroll
                for (int c = 0; c < 3; c++) f[c] = g.spread_f[3 * (size_t)i + c];
            } else {
                unsigned long long *a = g.acc + 3 * (size_t)i;
                const unsigned long long v[3] = {a[0], a[1], a[2]};
                const float sc = g.spread_sc[i];
                if ((v[0] | v[1] | v[2]) != 0ull) {
#pragma unroll
                    for (int c = 0; c < 3; c++) f[c] = (float)((double)(long long)v[c] * (1.0 / FIX_SCALE)) * sc;
                    a[0] = a[1] = a[2] = 0ull;
                }
            }
            if (f[0] != 0.0f || f[1] != 0.0f || f[2] != 0.0f) {
                float4 s = s0;
                const float q = p.dt / (p.rho0 + s.x);
                s.y += q * f[0];
                s.z += q * f[1];
                s.w += q * f[2];
                p.st[i] = s;
                any = 1;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) p.cell_force_out[3 * (size_t)i + c] = sfc[c] + f[c];
    }
    any = __syncthreads_or(any);
    if (tid == 0) {
        if (!gather) g.tile_acc[t] = 0;
        if (!p.all) g.tile_prev[t] = any;
    }
}

// ---------------------------------------------------------------------------
// membrane
// ---------------------------------------------------------------------------

// constants of one edge
__device__ __forceinline__ memf::EdgeConst edge_const(const MemArgs &m, const MemParamDev &p, int e)
{
    const float4 ec = m.edge_const[e];
    const float2 a0 = m.edge_area0[e];
    return {ec.x, ec.z, a0.x, a0.y, p.local_area_modulus, p.viscosity, p.spring_c, ec.w,
            p.spring_law == LBFS_SPRING_WLC_POW ? 1 : 0};
}


// per triangle: volume and area sums
__global__ __launch_bounds__(256) void k_mem_volume(MemArgs m)
{
    triangle_sums(m, blockIdx.x * blockDim.x + threadIdx.x);
}

// Per spring: the force on its four nodes with the membrane's global
// coefficients (from the volume and area sums of k_mem_volume and the MAD of
// this call), one record per edge slot; its MAD contribution (the MAD of the
// next force call) added to the membrane's sum.
__global__ __launch_bounds__(128) void k_mem_edges(MemArgs m)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int em = -1;
    long long mv[2] = {0, 0};
    if (e < m.n_edges) {
        const int mem = m.edge_mem[e];
        const MemParamDev p = m.params[mem];
        memf::EdgeIn<float> in;
        mem_coefs(m, mem, in.vol_coef, in.area_coef, in.mad_coef);
        const int4 q = m.edges[e];
        const float4 xa = m.X[q.x], xb = m.X[q.y], xc = m.X[q.z], xd = m.X[q.w], va = m.U[q.x], vc = m.U[q.z];
        in.A = {xa.x, xa.y, xa.z};
        in.B = {xb.x, xb.y, xb.z};
        in.C = {xc.x, xc.y, xc.z};
        in.D = {xd.x, xd.y, xd.z};
        in.Ua = {va.x, va.y, va.z};
        in.Uc = {vc.x, vc.y, vc.z};
        in.ks = m.edge_const[e].y;
        in.lb = p.local_bending;
        in.gb = p.global_bending;
        memf::EdgeOut<float> out;
        memf::edge_force(in, edge_const(m, p, e), out);
        m.edge_state[e] = make_float4(out.kappa, out.n1, out.n2, 0.0f);
#pragma unroll
        for (int j = 0; j < 4; j++)
            m.edge_force[4 * (size_t)e + j] = make_float4(out.F[j].x, out.F[j].y, out.F[j].z, 0.0f);
        if (e - m.mem_edge_start[mem] < p.n_mad) {
            em = mem;
            mv[0] = mem_fx((double)out.kappa * ((double)out.n1 + (double)out.n2) / 3.0, m.err);
        }
    }
    // the MAD sum goes to the third slot
    mem_sum_add(m.mem_acc + 2, em, mv, 1);
}

// eight lanes per node, each summing every eighth incidence
__global__ __launch_bounds__(256) void k_mem_nodes(MemArgs m)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = gid >> 3, sub = gid & 7;
    const bool live = nd < m.n_nodes;
    const float4 F = node_force8(m, live ? nd : 0, live ? sub : 8);
    if (live && sub == 0) m.F[nd] = F;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <bool INTERP>
static hipError_t launch_interp_kernel(const MemArgs &m, const IbmArgs &g, const float4 *st, hipStream_t s)
{
    // S x S lanes per node
    const int lanes = g.kernel == LBFS_IBM_LINEAR2 ? 4 : 16;
    const dim3 grid((unsigned)(((size_t)m.n_nodes * lanes + 255) / 256)), blk(256);
    if (g.kernel == LBFS_IBM_LINEAR2)
        hipLaunchKernelGGL((k_ibm_interp<LBFS_IBM_LINEAR2, INTERP>), grid, blk, 0, s, m, g, st);
    else if (g.kernel == LBFS_IBM_LAGRANGE4)
        hipLaunchKernelGGL((k_ibm_interp<LBFS_IBM_LAGRANGE4, INTERP>), grid, blk, 0, s, m, g, st);
    else
        hipLaunchKernelGGL((k_ibm_interp<LBFS_IBM_COS4, INTERP>), grid, blk, 0, s, m, g, st);
    return hipGetLastError();
}
hipError_t launch_ibm_interp(const MemArgs &m, const IbmArgs &g, const float4 *st, hipStream_t s)
{
    return launch_interp_kernel<true>(m, g, st, s);
}
hipError_t launch_ibm_weights(const MemArgs &m, const IbmArgs &g, hipStream_t s)
{
    return launch_interp_kernel<false>(m, g, nullptr, s);
}
hipError_t launch_mem_edges(const MemArgs &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_edges, dim3((m.n_edges + 127) / 128), dim3(128), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_ibm_apply(const IbmArgs &g, const ApplyArgs &p, hipStream_t s)
{
    hipLaunchKernelGGL(k_ibm_apply, dim3(p.n_tiles), dim3(256), 0, s, g, p);
    return hipGetLastError();
}
hipError_t launch_mem_volume(const MemArgs &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_volume, dim3((m.n_tri + 255) / 256), dim3(256), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_mem_nodes(const MemArgs &m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_nodes, dim3((8 * m.n_nodes + 255) / 256), dim3(256), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_ibm_spread(const MemArgs &m, const IbmArgs &g, hipStream_t s)
{
    if (g.binned == IBM_SPREAD_BINS) {
        const dim3 grid(std::min(g.n_tiles * (IBM_BIN_CAP / BIN_CH), 1024)), blk(BIN_NT);
        if (g.kernel == LBFS_IBM_LINEAR2)
            hipLaunchKernelGGL(k_ibm_spread_bins<LBFS_IBM_LINEAR2>, grid, blk, 0, s, m, g);
        else if (g.kernel == LBFS_IBM_LAGRANGE4)
            hipLaunchKernelGGL(k_ibm_spread_bins<LBFS_IBM_LAGRANGE4>, grid, blk, 0, s, m, g);
        else
            hipLaunchKernelGGL(k_ibm_spread_bins<LBFS_IBM_COS4>, grid, blk, 0, s, m, g);
    } else if (g.binned == IBM_SPREAD_GATHER) {
        const size_t lds = spread_tile_lds(g.max_tile_cells);
        const dim3 grid(g.n_tiles), blk(SPREAD_NT);
        if (g.kernel == LBFS_IBM_LINEAR2)
            hipLaunchKernelGGL(k_ibm_spread_tiles<LBFS_IBM_LINEAR2>, grid, blk, lds, s, m, g);
        else if (g.kernel == LBFS_IBM_LAGRANGE4)
            hipLaunchKernelGGL(k_ibm_spread_tiles<LBFS_IBM_LAGRANGE4>, grid, blk, lds, s, m, g);
        else
            hipLaunchKernelGGL(k_ibm_spread_tiles<LBFS_IBM_COS4>, grid, blk, lds, s, m, g);
    } else {
        const dim3 grid((m.n_nodes + 3) / 4), blk(256);
        if (g.kernel == LBFS_IBM_LINEAR2) hipLaunchKernelGGL(k_ibm_spread_nodes<LBFS_IBM_LINEAR2>, grid, blk, 0, s, m, g);
        else if (g.kernel == LBFS_IBM_LAGRANGE4)
            hipLaunchKernelGGL(k_ibm_spread_nodes<LBFS_IBM_LAGRANGE4>, grid, blk, 0, s, m, g);
        else hipLaunchKernelGGL(k_ibm_spread_nodes<LBFS_IBM_COS4>, grid, blk, 0, s, m, g);
    }
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

// Per triangle and cell, shared by both searches so they give the same bits
// (no product fused into an add): the Plücker sidedness of the ray from the
// cell along +x against the triangle's edges (all of one strict sign: the
// ray crosses it), and the distance from the cell to a vertex.
__device__ __forceinline__ bool visc_ray_crosses(const float p[3][3], float pc0, float pc1, float pc2)
{
#pragma clang fp contract(off)
    float side[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float *u = p[k], *v = p[(k + 1) % 3];
        const float e0 = u[1] * v[2] - v[1] * u[2], e1 = u[2] - v[2], e2 = v[1] - u[1];
        side[k] = pc0 * e1 + pc1 * e2 + pc2 * e0;
    }
    return (side[0] < 0 && side[1] < 0 && side[2] < 0) || (side[0] > 0 && side[1] > 0 && side[2] > 0);
}

__device__ __forceinline__ float visc_dist(const float4 c, const float p[3])
{
#pragma clang fp contract(off)
    const float dx = c.x - p[0], dy = c.y - p[1], dz = c.z - p[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

// a triangle's vertices relative to the membrane origin (as stored): the
// sidedness products then stay small and exact to a few ulps of the cell size
__device__ __forceinline__ void visc_tri(const MemArgs &m, int t, float p[3][3])
{
    const int4 q = m.tri[t];
    const int ids[3] = {q.x, q.y, q.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float4 x = m.X[ids[k]];
        p[k][0] = x.x; p[k][1] = x.y; p[k][2] = x.z;
    }
}

// order-preserving keys of floats (NaN aside) for an integer maximum
__device__ __forceinline__ unsigned fkey(float f)
{
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_value(unsigned k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ int visc_bin(float v, float v0, float inv, int n)
{
    const int b = (int)floorf((v - v0) * inv);
    return b < 0 ? 0 : (b >= n ? n - 1 : b);
}

// One workgroup per membrane with a viscosity ratio other than 1: its
// triangles into the y-z bins their (slightly widened) projections overlap,
// the largest smallest-x over each node's triangles (a node counts for the
// nearest-vertex distance of a cell exactly when one of its triangles lies
// wholly at x beyond the cell), and its nodes into 3-D bins.  Bin counts
// and cursors in LDS; entries within a bin in any order (the field does not
// depend on it).
__global__ __launch_bounds__(256) void k_visc_build(MemArgs m, ViscGrid v, float delta_h)
{
    constexpr int NB2 = VISC_G2 * VISC_G2, NB3 = VISC_G3 * VISC_G3 * VISC_G3, PER = NB3 / 256;
    const int mem = blockIdx.x, tid = threadIdx.x;
    if (m.params[mem].visc_ratio == 1.0f) return;  // uniform over the workgroup
    __shared__ int c2[NB2], c3[NB3], part[256], s_tot;
    const float4 lo = m.bbox[2 * mem], hi = m.bbox[2 * mem + 1], o = m.origin[mem];
    const float pad = 2.0f * delta_h;
    // grids in membrane coordinates (relative to its origin), over its box
    // widened by 2 delta_h
    const float y0 = lo.y - o.y - pad, z0 = lo.z - o.z - pad, x0 = lo.x - o.x - pad;
    const float ey = hi.y - o.y + pad - y0, ez = hi.z - o.z + pad - z0, ex = hi.x - o.x + pad - x0;
    const float iy = ey > 0.0f ? VISC_G2 / ey : 0.0f, iz = ez > 0.0f ? VISC_G2 / ez : 0.0f;
    const float jx = ex > 0.0f ? VISC_G3 / ex : 0.0f, jy = ey > 0.0f ? VISC_G3 / ey : 0.0f,
                jz = ez > 0.0f ? VISC_G3 / ez : 0.0f;
    // a triangle joins every bin its projection, widened by a small margin
    // (far beyond rounding), overlaps
    const float margin = 1e-4f * (ey + ez);
    if (tid == 0) {
        v.geo[3 * (size_t)mem] = make_float4(y0, z0, iy, iz);
        v.geo[3 * (size_t)mem + 1] = make_float4(x0, y0, z0, 0.0f);
        v.geo[3 * (size_t)mem + 2] = make_float4(jx, jy, jz, 0.0f);
    }
    const int t0 = m.mem_tri_start[mem], t1 = m.mem_tri_start[mem + 1];
    const int n0 = m.mem_node_start[mem], n1 = m.mem_node_start[mem + 1];
    for (int b = tid; b < NB2; b += 256) c2[b] = 0;
    for (int b = tid; b < NB3; b += 256) c3[b] = 0;
    for (int n = n0 + tid; n < n1; n += 256) v.qx[n] = fkey(-INFINITY);
    __syncthreads();
    auto tri_bins = [&](const float p[3][3], int &by0, int &by1, int &bz0, int &bz1) {
        const float ylo = fminf(fminf(p[0][1], p[1][1]), p[2][1]), yhi = fmaxf(fmaxf(p[0][1], p[1][1]), p[2][1]);
        const float zlo = fminf(fminf(p[0][2], p[1][2]), p[2][2]), zhi = fmaxf(fmaxf(p[0][2], p[1][2]), p[2][2]);
        by0 = visc_bin(ylo - margin, y0, iy, VISC_G2); by1 = visc_bin(yhi + margin, y0, iy, VISC_G2);
        bz0 = visc_bin(zlo - margin, z0, iz, VISC_G2); bz1 = visc_bin(zhi + margin, z0, iz, VISC_G2);
    };
    for (int t = t0 + tid; t < t1; t += 256) {
        float p[3][3];
        visc_tri(m, t, p);
        int by0, by1, bz0, bz1;
        tri_bins(p, by0, by1, bz0, bz1);
        for (int bz = bz0; bz <= bz1; bz++)
            for (int by = by0; by <= by1; by++) atomicAdd(&c2[by + VISC_G2 * bz], 1);
        const int4 q = m.tri[t];
        const unsigned kx = fkey(fminf(fminf(p[0][0], p[1][0]), p[2][0]));
        atomicMax(v.qx + q.x, kx);
        atomicMax(v.qx + q.y, kx);
        atomicMax(v.qx + q.z, kx);
    }
    __syncthreads();
    // exclusive offsets of the y-z bins (one per thread)
    {
        const int c = tid < NB2 ? c2[tid] : 0;
        part[tid] = c;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int add = tid >= off ? part[tid - off] : 0;
            __syncthreads();
            part[tid] += add;
            __syncthreads();
        }
        if (tid == 255) s_tot = part[255];
        const int excl = part[tid] - c;
        int *off = v.tri_off + (size_t)mem * (NB2 + 1);
        if (tid < NB2) {
            off[tid] = excl;
            c2[tid] = excl;  // cursor
        }
        __syncthreads();
        if (tid == 0) off[NB2] = s_tot;
    }
    const bool fits = s_tot <= VISC_TRI_CAP * (t1 - t0);
    if (tid == 0) v.ovf[mem] = fits ? 0 : 1;
    if (fits)
        for (int t = t0 + tid; t < t1; t += 256) {
            float p[3][3];
            visc_tri(m, t, p);
            int by0, by1, bz0, bz1;
            tri_bins(p, by0, by1, bz0, bz1);
            for (int bz = bz0; bz <= bz1; bz++)
                for (int by = by0; by <= by1; by++) {
                    const int slot = atomicAdd(&c2[by + VISC_G2 * bz], 1);
                    v.tri_ent[(size_t)VISC_TRI_CAP * t0 + slot] = t;
                }
        }
    // nodes into 3-D bins
    auto node_bin = [&](int n) {
        const float4 x = m.X[n];
        return visc_bin(x.x, x0, jx, VISC_G3) + VISC_G3 * (visc_bin(x.y, y0, jy, VISC_G3) + VISC_G3 * visc_bin(x.z, z0, jz, VISC_G3));
    };
    for (int n = n0 + tid; n < n1; n += 256) atomicAdd(&c3[node_bin(n)], 1);
    __syncthreads();
    {
        int sum = 0;
        for (int k = 0; k < PER; k++) sum += c3[tid * PER + k];
        part[tid] = sum;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int add = tid >= off ? part[tid - off] : 0;
            __syncthreads();
            part[tid] += add;
            __syncthreads();
        }
        int run = part[tid] - sum;
        int *off = v.node_off + (size_t)mem * (NB3 + 1);
        for (int k = 0; k < PER; k++) {
            const int b = tid * PER + k, c = c3[b];
            off[b] = run;
            c3[b] = run;
            run += c;
        }
        if (tid == 255) off[NB3] = part[255];
        __syncthreads();
    }
    for (int n = n0 + tid; n < n1; n += 256) {
        const int slot = atomicAdd(&c3[node_bin(n)], 1);
        v.node_ent[n0 + slot] = n;
    }
}

// Per cell: for every membrane with a viscosity ratio other than 1 whose
// bounding box (widened by 2 delta_h) holds the cell, the parity of the
// crossings of the ray from the cell along +x with the triangles lying
// wholly at larger x (inside or outside) and the distance to the nearest
// vertex of those triangles give the smoothed factor; several membranes:
// the value furthest from 1.  With the search structure the crossings come
// from the triangles of the cell's y-z bin and the distance from the nodes
// within 2 delta_h (beyond 1.5 delta_h the smoothed factor is exactly 1 in
// float, as it is with no vertex at all): the same bits as every triangle.
__global__ __launch_bounds__(256) void k_internal_visc(MemArgs m, IbmArgs g, ViscGrid v, float delta_h, float *visc_factor)
{
    constexpr int NB2 = VISC_G2 * VISC_G2, NB3 = VISC_G3 * VISC_G3 * VISC_G3;
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
        // membrane coordinates: the cell relative to the membrane origin
        const float4 o = m.origin[mem];
        const float4 cr = make_float4(c.x - o.x, c.y - o.y, c.z - o.z, 0.0f);
        const float xr = xh - o.x;
        const float pc0 = cr.x * cr.y - xr * cr.y, pc1 = cr.x * cr.z - xr * cr.z, pc2 = cr.x - xr;
        int count = 0;
        float dmin = 1e9f;
        if (!v.on || v.ovf[mem]) {
            for (int t = m.mem_tri_start[mem]; t < m.mem_tri_start[mem + 1]; t++) {
                float p[3][3];
                visc_tri(m, t, p);
                if (!(p[0][0] > cr.x && p[1][0] > cr.x && p[2][0] > cr.x)) continue;
                if (visc_ray_crosses(p, pc0, pc1, pc2)) count++;
                for (int k = 0; k < 3; k++) dmin = fminf(dmin, visc_dist(cr, p[k]));
            }
        } else {
            const float4 g0 = v.geo[3 * (size_t)mem], g1 = v.geo[3 * (size_t)mem + 1], g2 = v.geo[3 * (size_t)mem + 2];
            const int b2 = visc_bin(cr.y, g0.x, g0.z, VISC_G2) + VISC_G2 * visc_bin(cr.z, g0.y, g0.w, VISC_G2);
            const int *toff = v.tri_off + (size_t)mem * (NB2 + 1);
            const int *tent = v.tri_ent + (size_t)VISC_TRI_CAP * m.mem_tri_start[mem];
            for (int e = toff[b2]; e < toff[b2 + 1]; e++) {
                float p[3][3];
                visc_tri(m, tent[e], p);
                if (!(p[0][0] > cr.x && p[1][0] > cr.x && p[2][0] > cr.x)) continue;
                if (visc_ray_crosses(p, pc0, pc1, pc2)) count++;
            }
            const float R = 2.02f * delta_h;
            const int xa = visc_bin(cr.x - R, g1.x, g2.x, VISC_G3), xb = visc_bin(cr.x + R, g1.x, g2.x, VISC_G3);
            const int ya = visc_bin(cr.y - R, g1.y, g2.y, VISC_G3), yb = visc_bin(cr.y + R, g1.y, g2.y, VISC_G3);
            const int za = visc_bin(cr.z - R, g1.z, g2.z, VISC_G3), zb = visc_bin(cr.z + R, g1.z, g2.z, VISC_G3);
            const int *noff = v.node_off + (size_t)mem * (NB3 + 1);
            for (int bz = za; bz <= zb; bz++)
                for (int by = ya; by <= yb; by++)
                    for (int bx = xa; bx <= xb; bx++) {
                        const int b = bx + VISC_G3 * (by + VISC_G3 * bz);
                        for (int e = noff[b]; e < noff[b + 1]; e++) {
                            const int n = v.node_ent[m.mem_node_start[mem] + e];
                            if (!(fkey_value(v.qx[n]) > cr.x)) continue;
                            const float4 x = m.X[n];
                            const float p[3] = {x.x, x.y, x.z};
                            dmin = fminf(dmin, visc_dist(cr, p));
                        }
                    }
        }
        const float factor = 2.0f * fminf(0.5f, dmin * 0.5f / delta_h + 1.0f / PI_F * sinf(PI_F * dmin / delta_h) * 0.5f);
        const float val = 0.5f * ((count & 1) ? (r + 1.0f) + (r - 1.0f) * factor : (r + 1.0f) + (1.0f - r) * factor);
        // several membranes: keep the value furthest from 1
        if (!any || fabsf(val - 1.0f) > fabsf(out - 1.0f)) out = val;
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
        if (p.spring_law == LBFS_SPRING_WLC_POW) {
            // no energy routine in the reference: the integral of the WLC-POW
            // force from the rest length, k_s Lm [g(x) - g(x0)] - k_p (1/L - 1/L0),
            // g(x) = 1/(1-x) - x + 2x^2
            const double g = 1.0 / (1.0 - x) - x + 2.0 * x * x, g0 = 1.0 / (1.0 - x0) - x0 + 2.0 * x0 * x0;
            es += (double)kc.y * Lm * (g - g0) - (double)kc.y * (double)kc.w * (1.0 / (x * Lm) - 1.0 / (x0 * Lm));
        } else {
            es += 0.5 * (double)kc.y * fabs(0.25 * Lm * (x * x * (3.0 - 2.0 * x) / (1.0 - x) - 4.0 * x * c1 - c2));
        }
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

// Staging of lbfs_get_membrane_energies_noalloc: 7 FP64 energies per
// membrane, at most NOALLOC_MAX_MEMBRANES membranes (448 KiB per device, held
// until the process ends).  Host code never passes `scratch` itself (its host
// name is not a device address): energy_scratch_setup takes each device's
// address once, with that device current, and the kernels get that pointer.
static __device__ __align__(256) double scratch[7 * NOALLOC_MAX_MEMBRANES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

hipError_t energy_scratch_setup(int device, double **out)
{
    *out = nullptr;
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_internal_visc(const MemArgs &m, const IbmArgs &g, const ViscGrid &v, float delta_h,
                                float *visc_factor, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_bbox, dim3(m.n_mem), dim3(256), 0, s, m);
    if (v.on) hipLaunchKernelGGL(k_visc_build, dim3(m.n_mem), dim3(256), 0, s, m, v, delta_h);
    hipLaunchKernelGGL(k_internal_visc, dim3((g.n_cells + 255) / 256), dim3(256), 0, s, m, g, v, delta_h, visc_factor);
    return hipGetLastError();
}

}  // namespace lbfs
