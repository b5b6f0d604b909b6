// Adjoint of the IBM coupling and of the spring-network membrane for gfx950
// (4-point cosine kernel).  In reverse order of the forward step:
//
//   k_adj_spread      per node (one wavefront, one lane per stencil cell) at
//                     X^n+1: adjoint of the node force through spreading
//                     (immersed_boundary_method.cu:947-1024), and of X^n+1
//                     through the normalised weights
//   k_adj_mem_edges   per edge: the edge force (spring_network.cu:613-971) in
//                     forward-mode, one input direction at a time, contracted
//                     with the adjoint of the four node forces
//   k_adj_mem_partials per chunk of edges: partial sums of the coefficient
//                     and modulus adjoints
//   k_adj_mem_coef    per membrane (fixed order, in every workgroup that needs
//                     it): adjoints of the global volume/area/MAD coefficients
//                     and the modulus gradients; then per edge the volume and
//                     area constraint terms at X^n+1 and the MAD term, a
//                     function of X^n (one step lag)
//   k_adj_mem_nodes   per node: fixed-order sum of its edge slots; adds the
//                     advection X^n+1 = X^n + dt U^n in reverse
//   k_adj_interp      per node at X^n: adjoint of X^n through the weights of
//                     the interpolation (immersed_boundary_method.cu:14-97),
//                     and the node velocity adjoint spread onto the cells in
//                     64-bit fixed point (order independent)
#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "coupling.hpp"
#include "ibm_device.hpp"
#include "membrane_force.hpp"

namespace lbfs {

namespace {

// stencil cell of lane `lane` for a node at P: weight, its gradient and the
// internal cell id (-1 outside a wall)
__device__ __forceinline__ int stencil_cell(const IbmArgs &g, const IbmBlock &B, const float P[3], int lane, float &wt,
                                            float dw[3])
{
    float xr[3];
    int base[3], off[3];
    node_base(g, B, P, xr, base);
    lane_offset(4, lane, off);
    const int idx[3] = {base[0] + off[0], base[1] + off[1], base[2] + off[2]};
    int w[3];
    wt = 0.0f;
    dw[0] = dw[1] = dw[2] = 0.0f;
    if (!(wrap_index(B, 0, idx[0], w[0]) && wrap_index(B, 1, idx[1], w[1]) && wrap_index(B, 2, idx[2], w[2])))
        return -1;
    const int ic = cell_internal(g, B, w);
    wt = cos4_weight_grad(B, g.cen[ic], idx, w, P, dw);
    return ic;
}

__device__ __forceinline__ float3 f3of(float4 v) { return make_float3(v.x, v.y, v.z); }

// Gradient of sum_c w_c b_c for normalised weights w_c = wh_c / sum wh:
// (sum_c dwh_c b_c - (sum_c dwh_c) (sum_c w_c b_c)) / sum wh
__device__ __forceinline__ void weight_term(float wh, const float dw[3], float inv, float b, float out[3])
{
    const float sb = wave_sum(wh * inv * b);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float s1 = wave_sum(dw[a] * b), s2 = wave_sum(dw[a]);
        out[a] = (s1 - s2 * sb) * inv;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// spreading: f_c = static_c + (1 / V_c) sum_n w_nc F_n
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_spread(MemArgs m, IbmArgs g, AdjMemArgs j)
{
    const int lane = threadIdx.x & 63;
    const int nd = blockIdx.x * 4 + (threadIdx.x >> 6);
    // the node-velocity maximum of this reverse step starts from zero (the
    // previous step's readers ran before this kernel on the same stream)
    if (blockIdx.x == 0 && threadIdx.x == 0) *j.umax = 0u;
    if (nd >= m.n_nodes) return;
    const IbmBlock B = node_block(g, nd);
    const float4 o = m.origin[m.node_mem[nd]], xl = m.X[nd];
    const float P[3] = {o.x + xl.x, o.y + xl.y, o.z + xl.z};
    float wh, dw[3];
    const int ic = stencil_cell(g, B, P, lane, wh, dw);
    float3 a = make_float3(0.0f, 0.0f, 0.0f);
    if (ic >= 0) {
        const float4 l = j.lcf[ic];
        const float iv = 1.0f / g.cen[ic].w;
        a = make_float3(l.x * iv, l.y * iv, l.z * iv);
    }
    const float ws = wave_sum(wh);
    const float inv = ws > 0.0f ? 1.0f / ws : 0.0f;
    const float w = wh * inv;
    const float lfx = wave_sum(w * a.x), lfy = wave_sum(w * a.y), lfz = wave_sum(w * a.z);
    const float4 F = m.F[nd];
    const float b = F.x * a.x + F.y * a.y + F.z * a.z;
    float lx[3];
    weight_term(wh, dw, inv, b, lx);
    if (lane == 0) {
        j.lam_F[nd] = make_float4(lfx, lfy, lfz, 0.0f);
        const float4 l = j.lam_X[nd];
        j.lam_X[nd] = make_float4(l.x + lx[0], l.y + lx[1], l.z + lx[2], 0.0f);
    }
}

// ---------------------------------------------------------------------------
// membrane
// ---------------------------------------------------------------------------
namespace {

// the 24 differentiated inputs of an edge, in the order of the adjoint record:
// A B C D (12), Ua Uc (6), vol/area/MAD coefficients (3), ks, local and
// global bending moduli (3)
__device__ __forceinline__ memf::EdgeIn<memf::Dual> seeded(const float x[24], int d)
{
    using memf::Dual;
    Dual y[24];
#pragma unroll
    for (int k = 0; k < 24; k++) y[k] = Dual(x[k], k == d ? 1.0f : 0.0f);
    memf::EdgeIn<Dual> in;
    in.A = {y[0], y[1], y[2]};
    in.B = {y[3], y[4], y[5]};
    in.C = {y[6], y[7], y[8]};
    in.D = {y[9], y[10], y[11]};
    in.Ua = {y[12], y[13], y[14]};
    in.Uc = {y[15], y[16], y[17]};
    in.vol_coef = y[18];
    in.area_coef = y[19];
    in.mad_coef = y[20];
    in.ks = y[21];
    in.lb = y[22];
    in.gb = y[23];
    return in;
}

__device__ __forceinline__ void edge_inputs(const MemArgs &m, const float4 *X, int e, const MemParamDev &p,
                                            float x[24])
{
    const int4 q = m.edges[e];
    const float4 xa = X[q.x], xb = X[q.y], xc = X[q.z], xd = X[q.w], va = m.U[q.x], vc = m.U[q.z];
    const float v[18] = {xa.x, xa.y, xa.z, xb.x, xb.y, xb.z, xc.x, xc.y, xc.z,
                         xd.x, xd.y, xd.z, va.x, va.y, va.z, vc.x, vc.y, vc.z};
#pragma unroll
    for (int k = 0; k < 18; k++) x[k] = v[k];
    const double *gl = m.glob + 8 * (size_t)m.edge_mem[e];
    x[18] = (float)(-(double)p.volume_modulus * (gl[0] / gl[3] - 1.0) / 6.0);
    x[19] = (float)(-(double)p.global_area_modulus * (gl[1] / gl[4] - 1.0));
    x[20] = (float)(-2.0 * (double)p.global_bending * (gl[2] - gl[5]) / gl[4]);
    x[21] = m.edge_const[e].y;
    x[22] = p.local_bending;
    x[23] = p.global_bending;
}

__device__ __forceinline__ memf::EdgeConst edge_consts(const MemArgs &m, const MemParamDev &p, int e)
{
    const float4 ec = m.edge_const[e];
    const float2 a0 = m.edge_area0[e];
    return {ec.x, ec.z, a0.x, a0.y, p.local_area_modulus, p.viscosity, p.spring_c, ec.w,
            p.spring_law == LBFS_SPRING_WLC_POW ? 1 : 0};
}

}  // namespace

// one thread per (edge, input direction)
__global__ __launch_bounds__(256) void k_adj_mem_edges(MemArgs m, AdjMemArgs j)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = gid / 24, d = gid - 24 * e;
    if (e >= m.n_edges) return;
    const MemParamDev p = m.params[m.edge_mem[e]];
    float x[24];
    edge_inputs(m, m.X, e, p, x);
    const int4 q = m.edges[e];
    const float4 l0 = j.lam_F[q.x], l1 = j.lam_F[q.y], l2 = j.lam_F[q.z], l3 = j.lam_F[q.w];
    memf::EdgeOut<memf::Dual> out;
    memf::edge_force(seeded(x, d), edge_consts(m, p, e), out);
    j.erec[24 * (size_t)e + d] = l0.x * out.F[0].x.d + l0.y * out.F[0].y.d + l0.z * out.F[0].z.d +
                                 l1.x * out.F[1].x.d + l1.y * out.F[1].y.d + l1.z * out.F[1].z.d +
                                 l2.x * out.F[2].x.d + l2.y * out.F[2].y.d + l2.z * out.F[2].z.d +
                                 l3.x * out.F[3].x.d + l3.y * out.F[3].y.d + l3.z * out.F[3].z.d;
}

// Per membrane: sums of the coefficient and modulus adjoints over its edges
// (FP64, fixed order), chained through the FP64 coefficient formulas:
//   vol  = -k_v (V / V0 - 1) / 6     -> lambda_V   = l_vol (-k_v / (6 V0))
//   area = -k_A (A / A0 - 1)          -> lambda_A   = l_area (-k_A / A0)
//   mad  = -2 k_gb (MAD - MAD0) / A0  -> lambda_MAD = l_mad (-2 k_gb / A0),
//                                        d/dk_gb   = l_mad (-2 (MAD - MAD0) / A0)
// first level: per chunk of at most 256 edges of one membrane
__global__ __launch_bounds__(256) void k_adj_mem_partials(MemArgs m, AdjMemArgs j)
{
    __shared__ double sh[6][256];
    const int *q = j.echunk + 3 * (size_t)blockIdx.x;
    const int tid = threadIdx.x;
    const int e = q[1] + tid;
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (e < q[2]) {
        const float *r = j.erec + 24 * (size_t)e + 18;
        s[0] = r[0];
        s[1] = r[1];
        s[2] = r[2];
        s[3] = (double)r[3] * (double)m.edge_const[e].y;
        s[4] = r[4];
        s[5] = r[5];
    }
    for (int k = 0; k < 6; k++) sh[k][tid] = s[k];
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if (tid < stp)
            for (int k = 0; k < 6; k++) sh[k][tid] += sh[k][tid + stp];
        __syncthreads();
    }
    if (tid < 6) j.epart[6 * (size_t)blockIdx.x + tid] = sh[tid][0];
}

// Second level, fused with the per-edge terms that use its result.  Every
// workgroup first sums, one wavefront per membrane of its edges, the chunk
// partials of those membranes in a fixed order (so all workgroups get the same
// bits) and chains them through the coefficient formulas; the workgroup
// holding a membrane's first edge also adds the step's modulus gradients.
// Then, per edge, lane 12 of its ADJ_COEF_LANES lanes writes
// lambda_V dV/dX + lambda_A dA/dX at X^n+1 per edge slot, and lanes 0-11 the
// gradient at X^n of c1 [MAD edge] kappa (|N1| + |N2|) / 3 - c2 (|N1| + |N2|) / 6,
// c1 = lambda_MAD / A(X^n), c2 = lambda_MAD MAD / A(X^n) (the sum of the second
// term over the edges is c2 A).  The first step's MAD comes from the set-up (a
// constant of the run, like the reference areas and volume).
constexpr int ADJ_COEF_LANES = 13;
constexpr int ADJ_COEF_NT = 256;
constexpr int ADJ_COEF_MAX_MEM = ADJ_COEF_NT / ADJ_COEF_LANES + 2;  // membranes a workgroup's edges can span

__global__ __launch_bounds__(ADJ_COEF_NT) void k_adj_mem_coef(MemArgs m, AdjMemArgs j)
{
    __shared__ double s_mg[ADJ_COEF_MAX_MEM][5];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gid = blockIdx.x * ADJ_COEF_NT + tid;
    const int e_lo = blockIdx.x * ADJ_COEF_NT / ADJ_COEF_LANES;
    const int e_hi = min((blockIdx.x * ADJ_COEF_NT + ADJ_COEF_NT - 1) / ADJ_COEF_LANES, m.n_edges - 1);
    const int mem_lo = m.edge_mem[e_lo], n_mem_blk = m.edge_mem[e_hi] - mem_lo + 1;
    for (int k = wave; k < n_mem_blk; k += ADJ_COEF_NT / 64) {
        const int mem = mem_lo + k;
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int c = j.echunk_start[mem] + lane; c < j.echunk_start[mem + 1]; c += 64)
            for (int t = 0; t < 6; t++) s[t] += j.epart[6 * (size_t)c + t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int t = 0; t < 6; t++) s[t] += __shfl_xor(s[t], off, 64);
        if (lane == 0) {
            const MemParamDev p = m.params[mem];
            const double *gl = m.glob + 8 * (size_t)mem;
            const double *gp = j.glob_prev + 8 * (size_t)mem;
            const double lmad = s[2] * (-2.0 * (double)p.global_bending / gl[4]);
            double *mg = s_mg[k];
            mg[0] = s[0] * (-(double)p.volume_modulus / (6.0 * gl[3]));
            mg[1] = s[1] * (-(double)p.global_area_modulus / gl[4]);
            mg[2] = lmad;
            // MAD = M / A(X^n), M = sum over the MAD edges of kappa (|N1| + |N2|) / 3
            mg[3] = lmad / gp[6];
            mg[4] = lmad * gl[2] / gp[6];
            const int ef = m.mem_edge_start[mem];
            if (ef >= e_lo && ef <= e_hi) {
                double *gm = j.gmem + 4 * (size_t)mem;
                gm[0] += s[3];
                gm[1] += s[4];
                gm[2] += s[5] + s[2] * (-2.0 * (gl[2] - gl[5]) / gl[4]);
            }
        }
    }
    __syncthreads();
    const int e = gid / ADJ_COEF_LANES, d = gid - ADJ_COEF_LANES * e;
    if (e >= m.n_edges) return;
    const int mem = m.edge_mem[e];
    const double *mg = s_mg[mem - mem_lo];
    if (d == ADJ_COEF_LANES - 1) {
        const float lv = (float)mg[0], la = (float)mg[1];
        const int4 q = m.edges[e];
        const float4 xa = m.X[q.x], xb = m.X[q.y], xc = m.X[q.z], xd = m.X[q.w];
        memf::EdgeGeometry<float> g;
        memf::edge_geometry<float>({xa.x, xa.y, xa.z}, {xb.x, xb.y, xb.z}, {xc.x, xc.y, xc.z}, {xd.x, xd.y, xd.z},
                                   g);
#pragma unroll
        for (int sl = 0; sl < 4; sl++)
            j.grec[4 * (size_t)e + sl] = make_float4(lv * g.d_vol[3 * sl] + la * g.d_area[3 * sl],
                                                     lv * g.d_vol[3 * sl + 1] + la * g.d_area[3 * sl + 1],
                                                     lv * g.d_vol[3 * sl + 2] + la * g.d_area[3 * sl + 2], 0.0f);
        return;
    }
    float *rec = j.mrec + 12 * (size_t)e + d;
    if (!j.mad_lag || mg[2] == 0.0) {
        *rec = 0.0f;
        return;
    }
    const MemParamDev p = m.params[mem];
    const bool in_mad = e - m.mem_edge_start[mem] < p.n_mad;
    const float c1 = in_mad ? (float)mg[3] : 0.0f, c2 = (float)mg[4];
    float x[24];
    edge_inputs(m, j.Xn, e, p, x);
    memf::EdgeOut<memf::Dual> out;
    memf::edge_force(seeded(x, d), edge_consts(m, p, e), out);
    const memf::Dual nn = out.n1 + out.n2;
    *rec = c1 * (out.kappa.d * nn.v + out.kappa.v * nn.d) * (1.0f / 3.0f) - c2 * nn.d * (1.0f / 6.0f);
}

// Per node (eight lanes, fixed order): lambda_X^n+1 is complete once the
// edge, constraint and spreading shares are in; then the advection
// X^n+1 = X^n + dt U^n in reverse: lambda_U^n += dt lambda_X^n+1,
// lambda_X^n = lambda_X^n+1 + the MAD share (the interpolation weights add theirs next)
__global__ __launch_bounds__(256) void k_adj_mem_nodes(MemArgs m, IbmArgs g, AdjMemArgs j)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = gid >> 3, sub = gid & 7;
    const bool live = nd < m.n_nodes;
    float x[3] = {0, 0, 0}, u[3] = {0, 0, 0}, xm[3] = {0, 0, 0};
    if (live)
        for (int k = m.inc_start[nd] + sub; k < m.inc_start[nd + 1]; k += 8) {
            const int r = m.inc[k];
            const int e = r >> 2, s = r & 3;
            const float *er = j.erec + 24 * (size_t)e;
            const float4 gr = j.grec[r];
            const float *mr = j.mrec + 12 * (size_t)e + 3 * s;
            x[0] += er[3 * s] + gr.x;
            x[1] += er[3 * s + 1] + gr.y;
            x[2] += er[3 * s + 2] + gr.z;
            xm[0] += mr[0]; xm[1] += mr[1]; xm[2] += mr[2];
            if (s == 0 || s == 2) {
                const int b = s == 0 ? 12 : 15;
                u[0] += er[b]; u[1] += er[b + 1]; u[2] += er[b + 2];
            }
        }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            x[c] += __shfl_xor(x[c], off, 64);
            u[c] += __shfl_xor(u[c], off, 64);
            xm[c] += __shfl_xor(xm[c], off, 64);
        }
    if (live && sub == 0) {
        const float4 l = j.lam_X[nd];
        const float lx[3] = {l.x + x[0], l.y + x[1], l.z + x[2]};
        const float lu[3] = {u[0] + g.dt * lx[0], u[1] + g.dt * lx[1], u[2] + g.dt * lx[2]};
        j.lam_U[nd] = make_float4(lu[0], lu[1], lu[2], 0.0f);
        j.lam_X[nd] = make_float4(lx[0] + xm[0], lx[1] + xm[1], lx[2] + xm[2], 0.0f);
        const float mx = fmaxf(fabsf(lu[0]), fmaxf(fabsf(lu[1]), fabsf(lu[2])));
        atomicMax(j.umax, __float_as_uint(mx));
    }
}

// ---------------------------------------------------------------------------
// interpolation: U_n = sum_c w_nc(X^n) u_c
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_interp(MemArgs m, IbmArgs g, AdjMemArgs j)
{
    const int lane = threadIdx.x & 63;
    const int nd = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (nd >= m.n_nodes) return;
    const IbmBlock B = node_block(g, nd);
    const float4 o = m.origin[m.node_mem[nd]], xl = j.Xn[nd];
    const float P[3] = {o.x + xl.x, o.y + xl.y, o.z + xl.z};
    float wh, dw[3];
    const int ic = stencil_cell(g, B, P, lane, wh, dw);
    const float4 lu = j.lam_U[nd];
    float b = 0.0f;
    if (ic >= 0) {
        const float4 s = j.st_n[ic];
        b = lu.x * s.y + lu.y * s.z + lu.z * s.w;
    }
    const float ws = wave_sum(wh);
    const float inv = ws > 0.0f ? 1.0f / ws : 0.0f;
    float lx[3];
    weight_term(wh, dw, inv, b, lx);
    if (lane == 0) {
        const float4 l = j.lam_X[nd];
        j.lam_X[nd] = make_float4(l.x + lx[0], l.y + lx[1], l.z + lx[2], 0.0f);
    }
    const double scale = adj_fix_scale(*j.umax);
    const float w = wh * inv;
    if (ic >= 0 && w != 0.0f && scale > 0.0) {
        unsigned long long *acc = j.acc + 3 * (size_t)ic;
        atomicAdd(acc, fix_term((double)(w * lu.x), scale, g.err, ERR_FIX_ADJ));
        atomicAdd(acc + 1, fix_term((double)(w * lu.y), scale, g.err, ERR_FIX_ADJ));
        atomicAdd(acc + 2, fix_term((double)(w * lu.z), scale, g.err, ERR_FIX_ADJ));
    }
}

hipError_t launch_adj_coupling(const MemArgs &m, const IbmArgs &g, const AdjMemArgs &j, hipStream_t s)
{
    const int nb4 = (m.n_nodes + 3) / 4;
    hipLaunchKernelGGL(k_adj_spread, dim3(nb4), dim3(256), 0, s, m, g, j);
    hipLaunchKernelGGL(k_adj_mem_edges, dim3((24 * m.n_edges + 255) / 256), dim3(256), 0, s, m, j);
    hipLaunchKernelGGL(k_adj_mem_partials, dim3(j.n_echunks), dim3(256), 0, s, m, j);
    hipLaunchKernelGGL(k_adj_mem_coef, dim3((ADJ_COEF_LANES * m.n_edges + ADJ_COEF_NT - 1) / ADJ_COEF_NT),
                       dim3(ADJ_COEF_NT), 0, s, m, j);
    hipLaunchKernelGGL(k_adj_mem_nodes, dim3((8 * m.n_nodes + 255) / 256), dim3(256), 0, s, m, g, j);
    hipLaunchKernelGGL(k_adj_interp, dim3(nb4), dim3(256), 0, s, m, g, j);
    return hipGetLastError();
}

}  // namespace lbfs
