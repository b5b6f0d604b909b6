// Immersed-boundary coupling and spring-network membrane kernels for gfx950.
//
//   k_ibm_interp   per node: fluid velocity at the node (immersed_boundary_method.cu:14-97
//                  or the 4-point cosine kernel), explicit advection X += U dt
//                  (immersed_boundary_method.cu:234-264), the advected node put in
//                  its bin
//   k_mem_volume   per triangle: volume and area sums (lagrangian_object.cpp:601-650)
//                  in 64-bit fixed point, wavefront-reduced
//   k_mem_edges    per spring: all membrane force terms (spring_network.cu:613-971),
//                  written per edge and slot (no atomics); the MAD sum of the next
//                  force call (spring_network.cu:1186-1238)
//   k_mem_nodes    per node: sum of its edge slots in a fixed order, combined
//                  with the membrane's global coefficients
//   k_ibm_spread   per 4x4x4 bin tile: each cell gathers the forces of the nodes
//                  binned around it (immersed_boundary_method.cu:947-1024); the sum
//                  is accumulated in 64-bit fixed point so its value does not depend
//                  on the order the nodes are visited in
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "coupling.hpp"
#include "ibm_device.hpp"
#include "membrane_force.hpp"

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
    const double V = (double)(long long)acc[0] / MEM_SUM_SCALE, A = (double)(long long)acc[1] / MEM_SUM_SCALE;
    const MemParamDev &p = m.params[mem];
    vol = (float)(-(double)p.volume_modulus * (V / gl[3] - 1.0) / 6.0);
    area = (float)(-(double)p.global_area_modulus * (A / gl[4] - 1.0));
    mad = (float)(-2.0 * (double)p.global_bending * (gl[7] - gl[5]) / gl[4]);
}


// ---------------------------------------------------------------------------
// IBM: interpolation, advection, weights, binning.  One wavefront per node,
// one lane per stencil cell.
// ---------------------------------------------------------------------------
// INTERP: interpolate the cell velocity with the node's cached weights (those
// of its current position), advect the node and count it into its bin.
// Always: weights of the (new) position, normalised for the cosine kernel,
// cached for spreading and the next interpolation.
// (held to 64 registers: 8 wavefronts per SIMD keep more nodes' loads in flight)
template <bool INTERP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_ibm_interp(MemArgs m, IbmArgs g,
                                                                                         const float4 *__restrict__ st)
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
            // straight into the bin's slots; a full bin sends the node to the
            // overflow list (spread node by node)
            const int r = atomicAdd(&g.bin_count[key], 1);
            if (r == 0 && g.use_list) {  // once per bin and step
                const int li = atomicAdd(g.n_bin_list, 1);
                if (li < g.n_bins) g.bin_list[li] = key;
            }
            if (r < IBM_BIN_CAP) {
                g.bin_slots[(size_t)key * IBM_BIN_CAP + r] = nd;
            } else {
                const int o = atomicAdd(g.n_overflow, 1);
                if (o < m.n_nodes) g.overflow[o] = nd;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// IBM: spreading.  Per node bin (the 4x4x4 block of cells that holds the
// nodes' stencil base; its nodes were put in the bin's slots, and the bin in
// the list of occupied bins, by the interpolation), or per chunk of a bin's
// nodes: the nodes' stencils stay inside a 7x7x7 window, where the weighted
// forces are accumulated in LDS (one wavefront per node, one lane per
// stencil slot), then the window is added to the per-cell accumulators in
// global memory.  The node forces are gathered from their edge records here
// (eight lanes per node, the order of k_mem_nodes) and stored.  Nodes of full
// bins are spread one wavefront each after the bins.  All sums are 64-bit
// fixed point integer additions, so the result does not depend on the order
// of the nodes, lanes or workgroups.  Each write flags the cell's fluid tile;
// k_ibm_apply turns the accumulators of the flagged tiles into force
// densities and clears them.
// ---------------------------------------------------------------------------
// sum of a node's edge slots by eight lanes (sub = lane & 7; 8 or more: no
// slots), fixed order
__device__ __forceinline__ float4 node_force8(const MemArgs &m, int nd, int sub)
{
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (sub < 8)
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
    return make_float4(fx, fy, fz, 0.0f);
}

// one node by one wavefront straight into the global accumulators
__device__ __forceinline__ void spread_node_global(const MemArgs &m, const IbmArgs &g, int nd, float4 F, int lane)
{
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
    const int ic = cell_internal(g, B, w);
    unsigned long long *acc = g.acc + 3 * (size_t)ic;
    atomicAdd(acc, fix_term((double)(F.x * wt), FIX_SCALE, g.err, ERR_FIX_IBM));
    atomicAdd(acc + 1, fix_term((double)(F.y * wt), FIX_SCALE, g.err, ERR_FIX_IBM));
    atomicAdd(acc + 2, fix_term((double)(F.z * wt), FIX_SCALE, g.err, ERR_FIX_IBM));
    g.tile_now[g.cell_tile[ic]] = 1;
}

// LIST (few nodes): the workgroups loop over work items of CH nodes of the
// listed bins (IBM_BIN_CAP / CH items per bin, empty ones skipped), so a full
// bin is shared by several workgroups; the apply pass clears the bin counts.
// Otherwise one workgroup per bin of the grid, its nodes in chunks of CH
// (small workgroups measured fastest there: 128 threads, 16-node chunks).
template <int NT, int CH, bool LIST>
__global__ __launch_bounds__(NT) void k_ibm_spread(MemArgs m, IbmArgs g)
{
    constexpr int NW = NT / 64;
    constexpr int NCH = LIST ? IBM_BIN_CAP / CH : 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ unsigned long long racc[3][343];
    __shared__ int cnd[CH];
    __shared__ int4 cbase[CH];
    __shared__ float4 cF[CH];
    __shared__ float cw[CH][64];   // the chunk's cached stencil weights, one row per node
    const int S = stencil_size(g);
    int off[3];
    lane_offset(S, lane, off);
    const bool act = lane < S * S * S;
    // uniform over the workgroup
    const int n_items = LIST ? min(*g.n_bin_list, g.n_bins) * NCH : g.n_bins;
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int b = LIST ? g.bin_list[it / NCH] : it;
        const int cnt_all = min(g.bin_count[b], IBM_BIN_CAP);
        const int c_begin = LIST ? (it % NCH) * CH : 0;
        const int c_end = LIST ? min(cnt_all, c_begin + CH) : cnt_all;
        if (c_begin >= c_end) continue;  // uniform over the workgroup
        const IbmBlock B = g.blk[g.bin_block[b]];
        const int lb = b - B.bin0;
        const int bo[3] = {4 * (lb % B.tn[0]), 4 * ((lb / B.tn[0]) % B.tn[1]), 4 * (lb / (B.tn[0] * B.tn[1]))};
        __syncthreads();  // the previous item's window is flushed
        for (int q = tid; q < 3 * 343; q += NT) racc[q / 343][q % 343] = 0ull;
        const int *slots = g.bin_slots + (size_t)b * IBM_BIN_CAP;
        for (int c0 = c_begin; c0 < c_end; c0 += CH) {
            const int cnt = min(CH, c_end - c0);
            __syncthreads();
            if (tid < cnt) {
                const int p = slots[c0 + tid];
                cnd[tid] = p;
                cbase[tid] = g.nbase[p];
            }
            __syncthreads();
            // node forces from the edge records: eight lanes per node
            for (int jn = tid >> 3; jn < cnt; jn += NT / 8) {
                const int p = cnd[jn];
                const float4 F = node_force8(m, p, tid & 7);
                if ((tid & 7) == 0) {
                    cF[jn] = F;
                    m.F[p] = F;
                }
            }
            // all weight rows of the chunk in flight at once (one coalesced row per node)
            for (int q = tid; q < 64 * cnt; q += NT) cw[q >> 6][q & 63] = g.wnode[64 * (size_t)cnd[q >> 6] + (q & 63)];
            __syncthreads();
            for (int jn = wave; jn < cnt; jn += NW) {
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
                atomicAdd(&racc[0][loc], fix_term((double)(F.x * wt), FIX_SCALE, g.err, ERR_FIX_IBM));
                atomicAdd(&racc[1][loc], fix_term((double)(F.y * wt), FIX_SCALE, g.err, ERR_FIX_IBM));
                atomicAdd(&racc[2][loc], fix_term((double)(F.z * wt), FIX_SCALE, g.err, ERR_FIX_IBM));
            }
        }
        __syncthreads();
        for (int q = tid; q < 343; q += NT) {
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
            g.tile_now[g.cell_tile[ic]] = 1;
        }
        // (the grid form) read by this workgroup only; cleared for the next step
        if (!LIST && tid == 0) g.bin_count[b] = 0;
    }
    // nodes of full bins, one wavefront each
    const int nov = min(*g.n_overflow, m.n_nodes);
    for (int k = blockIdx.x * NW + wave; k < nov; k += gridDim.x * NW) {
        const int nd = g.overflow[k];
        float4 F = node_force8(m, nd, lane & 7);
        F.x = __shfl(F.x, 0, 64);
        F.y = __shfl(F.y, 0, 64);
        F.z = __shfl(F.z, 0, 64);
        if (lane == 0) m.F[nd] = F;
        spread_node_global(m, g, nd, F, lane);
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
    spread_node_global(m, g, nd, m.F[nd], lane);
}

// One workgroup per fluid tile: the spread sums of the tile's cells turned
// into force densities (static force + sum * scale), added to the state the
// face pass updated with the static force (rho u += dt f, the update's own
// form: u_new = u* + dt f / rho_new, rho_new unchanged by the force), written
// to the cell force output and cleared.
__global__ __launch_bounds__(256) void k_ibm_apply(IbmArgs g, ApplyArgs p)
{
    const int t = blockIdx.x, tid = threadIdx.x;
    if (t == 0) {
        // the listed bins' counts (the spreading shares a bin between workgroups)
        if (g.binned && g.use_list) {
            const int nl = min(*g.n_bin_list, g.n_bins);
            for (int e = tid; e < nl; e += blockDim.x) g.bin_count[g.bin_list[e]] = 0;
        }
        __syncthreads();
        if (tid == 0 && g.binned) {
            *g.n_overflow = 0;
            *g.n_bin_list = 0;
        }
        // the step's membrane sums: V, A, the MAD the step used, A as the next
        // A_prev and the MAD of the next force call (numerator / A)
        for (int k = tid; k < p.n_mem; k += blockDim.x) {
            unsigned long long *s = p.mem_acc + 4 * (size_t)k;
            double *gl = p.glob + 8 * (size_t)k;
            const double V = (double)(long long)s[0] / MEM_SUM_SCALE, A = (double)(long long)s[1] / MEM_SUM_SCALE;
            const double M = (double)(long long)s[2] / MEM_SUM_SCALE;
            gl[0] = V;
            gl[1] = A;
            gl[2] = gl[7];
            gl[6] = A;
            gl[7] = M / A;
            s[0] = s[1] = s[2] = 0ull;
        }
    }
    const int now = g.tile_now[t], prev = g.tile_prev[t];
    if (!(now | prev | p.all)) return;  // uniform over the workgroup
    for (int i = p.tile_cell_start[t] + tid; i < p.tile_cell_start[t + 1]; i += blockDim.x) {
        float f[3] = {0.0f, 0.0f, 0.0f};
        if (now) {
            unsigned long long *a = g.acc + 3 * (size_t)i;
            const float sc = g.spread_sc[i];
            bool any = false;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const long long v = (long long)a[c];
                f[c] = (float)((double)v / FIX_SCALE) * sc;
                any |= v != 0;
            }
            if (any) {
                a[0] = a[1] = a[2] = 0ull;
                float4 s = p.st[i];
                const float q = p.dt / (p.rho0 + s.x);
                s.y += q * f[0];
                s.z += q * f[1];
                s.w += q * f[2];
                p.st[i] = s;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) p.cell_force_out[3 * (size_t)i + c] = g.static_force[3 * (size_t)i + c] + f[c];
    }
    __syncthreads();
    if (tid == 0) {
        if (now) g.tile_now[t] = 0;
        if (!p.all) g.tile_prev[t] = now;
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
    // the workgroups loop over the listed bins (or take one bin of the grid
    // each), then over the nodes of full bins
    if (g.binned && g.use_list)
        hipLaunchKernelGGL((k_ibm_spread<512, 32, true>), dim3(std::min(g.n_bins * (IBM_BIN_CAP / 32), 1024)), dim3(512),
                           0, s, m, g);
    else if (g.binned)
        hipLaunchKernelGGL((k_ibm_spread<128, 16, false>), dim3(g.n_bins), dim3(128), 0, s, m, g);
    else
        hipLaunchKernelGGL(k_ibm_spread_nodes, dim3((m.n_nodes + 3) / 4), dim3(256), 0, s, m, g);
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

hipError_t launch_internal_visc(const MemArgs &m, const IbmArgs &g, float delta_h, float *visc_factor, hipStream_t s)
{
    hipLaunchKernelGGL(k_mem_bbox, dim3(m.n_mem), dim3(256), 0, s, m);
    hipLaunchKernelGGL(k_internal_visc, dim3((g.n_cells + 255) / 256), dim3(256), 0, s, m, g, delta_h, visc_factor);
    return hipGetLastError();
}

}  // namespace lbfs
