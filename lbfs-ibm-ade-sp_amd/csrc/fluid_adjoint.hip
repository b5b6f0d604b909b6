// Adjoint (reverse-mode derivative) of the fluid step, hand-written for gfx950.
//
//   k_adj_update  per cell: adjoint of the forward-Euler update
//                 (LBFS.cu:2089-2172) with respect to the state, the cell
//                 residual and the cell force
//   k_adj_flux    one workgroup per tile of cells, as the forward flux pass:
//                 the transpose of each face's reconstruction + equilibrium +
//                 flux (LBFS.cu:2574-2810) is evaluated in chunks of faces
//                 staged in LDS, then every cell gathers its own side of its
//                 six faces (state and gradient adjoints) in a fixed order
//   k_adj_grad    per cell: transpose of the least-squares gradients
//                 (LBFS.cu:2177-2240): its own stencil rows plus, per
//                 neighbour, the neighbour's row that points back to it
//                 (weights gathered once on the host), plus the
//                 interpolation adjoint held in fixed point
//
// Strict upwinding (ties to the owner) is piecewise smooth; its derivative is
// taken on the side the forward step took.  No floating-point atomics:
// results are bitwise repeatable.
#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "fluid_device.hpp"

namespace lbfs {

// ---------------------------------------------------------------------------
// update
// ---------------------------------------------------------------------------
// forward: drho' = drho + dt r0, rho' = rho0 + drho',
//          u'_t  = u_t + dt (r_t + f_t - u_t r0) / rho'
// Cell i of the update's adjoint from the step's state adjoint L, its start
// state s0, residual r and cell force cf; p: the cell's body-force terms.
__device__ __forceinline__ void adj_update_cell(const FluidArgs &a, const AdjFluidArgs &j, const float *cf_all, int i,
                                                float4 L, float4 s0, float4 r, double p[3])
{
    float fx = a.body_force[0], fy = a.body_force[1], fz = a.body_force[2];
    if (cf_all) {
        const float *cf = cf_all + 3 * (size_t)i;
        fx += cf[0]; fy += cf[1]; fz += cf[2];
    }
    const float r0 = r.x, r1 = r.y + fx, r2 = r.z + fy, r3 = r.w + fz;
    const float dt = a.dt;
    const float rho = a.rho0 + (s0.x + dt * r0);
    const float inv = 1.0f / rho;
    const float a1 = dt * (r1 - s0.y * r0), a2 = dt * (r2 - s0.z * r0), a3 = dt * (r3 - s0.w * r0);
    const float lrho = -inv * inv * (L.y * a1 + L.z * a2 + L.w * a3);
    const float ld = L.x + lrho;
    const float lr0 = dt * ld - dt * inv * (L.y * s0.y + L.z * s0.z + L.w * s0.w);
    const float c = dt * inv;
    const float lr1 = L.y * c, lr2 = L.z * c, lr3 = L.w * c;
    const float keep = 1.0f - dt * r0 * inv;
    const float iv = a.inv_vol[i];
    j.mu[i] = make_float4(lr0 * iv, lr1 * iv, lr2 * iv, lr3 * iv);
    j.lcf[i] = make_float4(lr1, lr2, lr3, 0.0f);
    j.ls[i] = make_float4(ld, L.y * keep, L.z * keep, L.w * keep);
    p[0] = lr1; p[1] = lr2; p[2] = lr3;
}

// the workgroup's body-force terms (256 threads, one cell each) added to its
// own partials: the running sums need no atomics
__device__ __forceinline__ void adj_force_partials(const AdjFluidArgs &j, const double p[3])
{
    __shared__ double sh[3][256];
    for (int t = 0; t < 3; t++) sh[t][threadIdx.x] = p[t];
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp)
            for (int t = 0; t < 3; t++) sh[t][threadIdx.x] += sh[t][threadIdx.x + stp];
        __syncthreads();
    }
    if (threadIdx.x < 3) j.fpart[3 * (size_t)blockIdx.x + threadIdx.x] += sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_adj_update(FluidArgs a, AdjFluidArgs j, const float4 *__restrict__ st,
                                                    const float4 *__restrict__ res)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double p[3] = {0.0, 0.0, 0.0};
    if (i < a.n_cells) adj_update_cell(a, j, a.cell_force, i, j.lam_in[i], st[i], res[i], p);
    adj_force_partials(j, p);
}

// fixed-order sum of the partials into the body-force gradient
__global__ __launch_bounds__(256) void k_adj_force_sum(AdjFluidArgs j)
{
    __shared__ double sh[3][256];
    double p[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < j.n_fpart; b += 256)
        for (int t = 0; t < 3; t++) p[t] += j.fpart[3 * (size_t)b + t];
    for (int t = 0; t < 3; t++) sh[t][threadIdx.x] = p[t];
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp)
            for (int t = 0; t < 3; t++) sh[t][threadIdx.x] += sh[t][threadIdx.x + stp];
        __syncthreads();
    }
    if (threadIdx.x < 3) j.gforce[threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// face flux
// ---------------------------------------------------------------------------
// adjoint of g = w (q0 + (rho0 + q0) P), P = eu (3 + 4.5 ig eu) - 1.5 ig |q_u|^2
__device__ __forceinline__ void eq_point_vjp(const float q[4], float eu, float lg, float w, float rho0, float ig,
                                             float lq[4], float &leu)
{
    const float uu = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    const float P = eu * (3.0f + 4.5f * ig * eu) - 1.5f * ig * uu;
    const float R = lg * w * (rho0 + q[0]);
    lq[0] = lg * w * (1.0f + P);
    leu = R * (3.0f + 9.0f * ig * eu);
    const float c = -3.0f * ig * R;
    lq[1] = c * q[1];
    lq[2] = c * q[2];
    lq[3] = c * q[3];
}

// adjoint moments of the face: S0 += gs, S1 += e gd, T += e (e.n) gs
struct FaceAdj {
    float lS0, lS1x, lS1y, lS1z, lTx, lTy, lTz;
};

// One pair of opposite directions in reverse: the forward reconstruction of
// the two lattice points (owner side along s e, neighbour side along -s e,
// s = sign(e.n); a tie feeds both from the owner) and the transpose of it.
template <class L, int K>
__device__ __forceinline__ void pair_vjp(const FaceAdj &fa, const float po[4], const float pn[4], const float Go[12],
                                         const float Gn[12], float nx, float ny, float nz, float rho0, float ig,
                                         float lpo[4], float lpn[4], float lGo[12], float lGn[12])
{
    constexpr int ex = L::ex[K], ey = L::ey[K], ez = L::ez[K];
    constexpr float w = L::w(K);
    const float en = edot<ex, ey, ez>(nx, ny, nz);
    const bool tie = en == 0.0f;
    const float sg = en > 0.0f ? 1.0f : -1.0f;
    float ego[4], egn[4], qo[4], qn[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        ego[v] = edot<ex, ey, ez>(Go[3 * v], Go[3 * v + 1], Go[3 * v + 2]);
        egn[v] = edot<ex, ey, ez>(Gn[3 * v], Gn[3 * v + 1], Gn[3 * v + 2]);
        qo[v] = po[v] - sg * ego[v];
        qn[v] = tie ? po[v] - ego[v] : pn[v] + sg * egn[v];
    }
    const float eu_o = sg * edot<ex, ey, ez>(qo[1], qo[2], qo[3]);
    const float eu_n = -sg * edot<ex, ey, ez>(qn[1], qn[2], qn[3]);
    const float lgs = fa.lS0 + en * edot<ex, ey, ez>(fa.lTx, fa.lTy, fa.lTz);
    const float lgd = edot<ex, ey, ez>(fa.lS1x, fa.lS1y, fa.lS1z);
    const float lgo = lgs + sg * lgd, lgn = lgs - sg * lgd;
    float lqo[4], lqn[4], leo, len;
    eq_point_vjp(qo, eu_o, lgo, w, rho0, ig, lqo, leo);
    eq_point_vjp(qn, eu_n, lgn, w, rho0, ig, lqn, len);
    leo *= sg;
    len *= -sg;
    eacc<ex>(lqo[1], leo); eacc<ey>(lqo[2], leo); eacc<ez>(lqo[3], leo);
    eacc<ex>(lqn[1], len); eacc<ey>(lqn[2], len); eacc<ez>(lqn[3], len);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        // qo = po - s ego; qn = pn + s egn, or po - ego on a tie
        float lego = -sg * lqo[v], legn = 0.0f;
        lpo[v] += lqo[v];
        if (tie) {
            lpo[v] += lqn[v];
            lego -= lqn[v];
        } else {
            lpn[v] += lqn[v];
            legn = sg * lqn[v];
        }
        eacc<ex>(lGo[3 * v], lego); eacc<ey>(lGo[3 * v + 1], lego); eacc<ez>(lGo[3 * v + 2], lego);
        eacc<ex>(lGn[3 * v], legn); eacc<ey>(lGn[3 * v + 1], legn); eacc<ez>(lGn[3 * v + 2], legn);
    }
}

template <class L, int... P>
__device__ __forceinline__ void all_pair_vjps(std::integer_sequence<int, P...>, const FaceAdj &fa, const float po[4],
                                              const float pn[4], const float Go[12], const float Gn[12], float nx,
                                              float ny, float nz, float rho0, float ig, float lpo[4], float lpn[4],
                                              float lGo[12], float lGn[12])
{
    (pair_vjp<L, 2 * P + 1>(fa, po, pn, Go, Gn, nx, ny, nz, rho0, ig, lpo, lpn, lGo, lGn), ...);
}

// Adjoint of one tile-face's flux (the forward of face_flux in
// fluid_kernels.hip, strict upwinding), from the face's moments S0, S1 as the
// forward pass computed them.  out[0] state and out[1..3] gradient
// adjoint of the owner, out[4..7] the same for the neighbour (zero for a
// boundary face, whose ghost is a function of the owner).
template <class L, bool GHOST>
__device__ __forceinline__ void face_vjp(const FluidArgs &a, const float4 *__restrict__ st,
                                         const float4 *__restrict__ grad, const float4 *__restrict__ mu, int tf,
                                         float4 out[8])
{
    const FaceRec fr = load_face(a, tf);
    const float4 fb = fr.d1, fcv = fr.d2;
    const int o = fr.o, nb = fr.nb;
    constexpr bool ghost = GHOST;  // boundary faces follow the interior faces in every tile list
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 8; k++) out[k] = zero;
    const float nx = fr.nx, ny = fr.ny, nz = fr.nz, A = fr.A, dt = fr.dt;
    const float4 so = st[o];
    const float4 mo = mu[o];
    float go[12], gn[12];
    load_grad(grad, o, go);
    const float sov[4] = {so.x, so.y, so.z, so.w};
    float po[4], pn[4], p0[4];
    float4 lam;
    int ty[4] = {0, 0, 0, 0};
    float inv_dmag = 0.0f, dir[3] = {0.0f, 0.0f, 0.0f}, fac[3] = {1.0f, 1.0f, 1.0f};
    if (!ghost) {
        const float4 sn = st[nb], mn4 = mu[nb];
        lam = make_float4(mn4.x - mo.x, mn4.y - mo.y, mn4.z - mo.z, mn4.w - mo.w);
        load_grad(grad, nb, gn);
        const float snv[4] = {sn.x, sn.y, sn.z, sn.w};
#pragma unroll
        for (int v = 0; v < 4; v++) {
            po[v] = sov[v] - (fb.x * go[3 * v] + fb.y * go[3 * v + 1] + fb.z * go[3 * v + 2]);
            pn[v] = snv[v] - (fcv.x * gn[3 * v] + fcv.y * gn[3 * v + 1] + fcv.z * gn[3 * v + 2]);
            p0[v] = 0.5f * (po[v] + pn[v]);
        }
    } else {
        const int bc = nb - a.n_cells;
        const int2 t2 = a.bc_type[bc];
        if (t2.y == LBFS_BC_EMPTY) return;
        lam = make_float4(-mo.x, -mo.y, -mo.z, -mo.w);
        const float4 bv = a.bc_val[bc];
        const float bvv[4] = {bv.x, bv.y, bv.z, bv.w};
        const float d1n = sqrtf(fb.x * fb.x + fb.y * fb.y + fb.z * fb.z);
        inv_dmag = 2.0f / d1n;
        if (a.flags & LBFS_FLAG_GHOST_DIR_GPU) { dir[0] = -fb.x / d1n; dir[1] = -fb.y / d1n; dir[2] = -fb.z / d1n; }
        else { dir[0] = nx; dir[1] = ny; dir[2] = nz; }
        const float nv[3] = {nx, ny, nz};
#pragma unroll
        for (int c = 0; c < 3; c++) fac[c] = (a.flags & LBFS_FLAG_NEUMANN_SIGNED) ? 1.0f - nv[c] : 1.0f - fabsf(nv[c]);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            ty[v] = v == 0 ? t2.x : t2.y;
            if (ty[v] == 1 || ty[v] == 3) {
                const float dphi = 2.0f * (bvv[v] - sov[v]) * inv_dmag;
#pragma unroll
                for (int c = 0; c < 3; c++) gn[3 * v + c] = dphi * dir[c];
                p0[v] = ty[v] == 1 ? bvv[v] : sov[v] - (fb.x * gn[3 * v] + fb.y * gn[3 * v + 1] + fb.z * gn[3 * v + 2]);
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) gn[3 * v + c] = go[3 * v + c] * fac[c];
                p0[v] = sov[v] - (fb.x * gn[3 * v] + fb.y * gn[3 * v + 1] + fb.z * gn[3 * v + 2]);
            }
            po[v] = p0[v];
            pn[v] = p0[v];
        }
    }
    // LBFS_FLAG_GHOST_FACE_GRAD: the owner-side points of a boundary face take
    // the ghost gradient (LBFS.cu:391-394); the reverse below routes their
    // gradient adjoint into the ghost gradient's
    const bool ghost_grad_all = ghost && (a.flags & LBFS_FLAG_GHOST_FACE_GRAD);
    if (ghost_grad_all) {
#pragma unroll
        for (int t = 0; t < 12; t++) go[t] = gn[t];
    }
    float Go[12], Gn[12];
#pragma unroll
    for (int t = 0; t < 12; t++) { Go[t] = go[t] * dt; Gn[t] = gn[t] * dt; }
    const float rho0 = a.rho0, ig = a.inv_gamma;
    // forward moments S0, S1 of this face, kept by the forward pass
    const float4 fm = a.face_moments[tf];
    FaceSums f;
    f.S0 = fm.x;
    f.S1x = fm.y;
    f.S1y = fm.z;
    f.S1z = fm.w;
    const float P0 = -1.5f * (p0[1] * p0[1] + p0[2] * p0[2] + p0[3] * p0[3]) * ig;
    const float rho_f = rho0 + f.S0;
    const float mnrm = f.S1x * nx + f.S1y * ny + f.S1z * nz;
    const float vf = a.visc_factor ? a.visc_factor[o] : 1.0f;
    const float tau = a.visc3_over_gamma * vf / dt;
    const float c1 = 1.0f - tau;
    const float irf = 1.0f / rho_f;
    const float q = mnrm * irf * ig;
    // reverse of the flux from the moments
    const float Lx = lam.y, Ly = lam.z, Lz = L::dims == 3 ? lam.w : 0.0f;
    const float lq = A * c1 * (Lx * f.S1x + Ly * f.S1y + Lz * f.S1z);
    const float lmn = lam.x * A + lq * ig * irf;
    const float lrf = -lq * mnrm * ig * irf * irf;
    FaceAdj fj;
    fj.lS0 = A * c1 * (Lx * nx + Ly * ny + Lz * nz) * (1.0f / 3.0f) + lrf;
    fj.lS1x = Lx * A * c1 * q + lmn * nx;
    fj.lS1y = Ly * A * c1 * q + lmn * ny;
    fj.lS1z = Lz * A * c1 * q + lmn * nz;
    fj.lTx = Lx * A * tau;
    fj.lTy = Ly * A * tau;
    fj.lTz = Lz * A * tau;
    // rest point
    float lp0[4];
    {
        const float w0 = L::w(0);
        lp0[0] = fj.lS0 * w0 * (1.0f + P0);
        const float c = fj.lS0 * w0 * (rho0 + p0[0]) * (-3.0f * ig);
        lp0[1] = c * p0[1];
        lp0[2] = c * p0[2];
        lp0[3] = c * p0[3];
    }
    float lpo[4] = {0.0f, 0.0f, 0.0f, 0.0f}, lpn[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float lGo[12], lGn[12];
#pragma unroll
    for (int t = 0; t < 12; t++) { lGo[t] = 0.0f; lGn[t] = 0.0f; }
    all_pair_vjps<L>(std::make_integer_sequence<int, (L::Q - 1) / 2>{}, fj, po, pn, Go, Gn, nx, ny, nz, rho0, ig, lpo,
                     lpn, lGo, lGn);
    float lso[4], lgo_[12];
    if (!ghost) {
        float lsn[4], lgn_[12];
        const float d1[3] = {fb.x, fb.y, fb.z}, d2[3] = {fcv.x, fcv.y, fcv.z};
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float a_o = lpo[v] + 0.5f * lp0[v], a_n = lpn[v] + 0.5f * lp0[v];
            lso[v] = a_o;
            lsn[v] = a_n;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                lgo_[3 * v + c] = dt * lGo[3 * v + c] - d1[c] * a_o;
                lgn_[3 * v + c] = dt * lGn[3 * v + c] - d2[c] * a_n;
            }
        }
        out[4] = make_float4(lsn[0], lsn[1], lsn[2], lsn[3]);
        out[5] = make_float4(lgn_[0], lgn_[1], lgn_[2], lgn_[3]);
        out[6] = make_float4(lgn_[4], lgn_[5], lgn_[6], lgn_[7]);
        out[7] = make_float4(lgn_[8], lgn_[9], lgn_[10], lgn_[11]);
    } else {
        const float d1[3] = {fb.x, fb.y, fb.z};
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float lp = lpo[v] + lpn[v] + lp0[v];  // every lattice point starts from the face value
            float lgv[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (ghost_grad_all) {
                    lgo_[3 * v + c] = 0.0f;
                    lgv[c] = dt * (lGn[3 * v + c] + lGo[3 * v + c]);
                } else {
                    lgo_[3 * v + c] = dt * lGo[3 * v + c];
                    lgv[c] = dt * lGn[3 * v + c];
                }
            }
            lso[v] = 0.0f;
            if (ty[v] == 1 || ty[v] == 3) {
                if (ty[v] == 3) {
                    lso[v] += lp;
#pragma unroll
                    for (int c = 0; c < 3; c++) lgv[c] -= d1[c] * lp;
                }
                const float ldphi = dir[0] * lgv[0] + dir[1] * lgv[1] + dir[2] * lgv[2];
                lso[v] -= 2.0f * inv_dmag * ldphi;
            } else {
                lso[v] += lp;
#pragma unroll
                for (int c = 0; c < 3; c++) lgo_[3 * v + c] += fac[c] * (lgv[c] - d1[c] * lp);
            }
        }
    }
    out[0] = make_float4(lso[0], lso[1], lso[2], lso[3]);
    out[1] = make_float4(lgo_[0], lgo_[1], lgo_[2], lgo_[3]);
    out[2] = make_float4(lgo_[4], lgo_[5], lgo_[6], lgo_[7]);
    out[3] = make_float4(lgo_[8], lgo_[9], lgo_[10], lgo_[11]);
}

constexpr int ADJ_THREADS = 512;
constexpr int ADJ_CHUNK = 512;       // faces per LDS chunk (8 float4 each: 64 KiB), held as 8 planes
                                     // of ADJ_CHUNK float4 so that a wavefront's stores are contiguous

__device__ __forceinline__ void add4(float *acc, float4 v)
{
    acc[0] += v.x; acc[1] += v.y; acc[2] += v.z; acc[3] += v.w;
}

// ADJ_CELLS: cells per thread (1: tiles of up to 512 cells, 2: up to 1024)
template <class L, int ADJ_CELLS>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_flux(FluidArgs a, AdjFluidArgs j, const float4 *__restrict__ st,
                                                          const float4 *__restrict__ grad)
{
    extern __shared__ float4 lds_adj[];
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int f0 = a.tile_face_start[t], f1 = a.tile_face_start[t + 1];
    const int c0 = a.tile_cell_start[t], c1 = a.tile_cell_start[t + 1];
    const int fb = a.tile_bface_start[t];
    float acc[ADJ_CELLS][16];
    uint32_t slot[ADJ_CELLS][6];
#pragma unroll
    for (int u = 0; u < ADJ_CELLS; u++) {
#pragma unroll
        for (int k = 0; k < 16; k++) acc[u][k] = 0.0f;
        const int i = c0 + (int)threadIdx.x + u * ADJ_THREADS;
        if (i < c1) {
            const uint32_t *sl = reinterpret_cast<const uint32_t *>(a.cell_slots) + 3 * (size_t)i;
            const uint32_t s01 = sl[0], s23 = sl[1], s45 = sl[2];
            slot[u][0] = s01 & 0xffffu; slot[u][1] = s01 >> 16;
            slot[u][2] = s23 & 0xffffu; slot[u][3] = s23 >> 16;
            slot[u][4] = s45 & 0xffffu; slot[u][5] = s45 >> 16;
        } else {
#pragma unroll
            for (int k = 0; k < 6; k++) slot[u][k] = 0xffffffffu;
        }
    }
    for (int cs = f0; cs < f1; cs += ADJ_CHUNK) {
        const int ce = cs + ADJ_CHUNK < f1 ? cs + ADJ_CHUNK : f1;
        for (int f = cs + threadIdx.x; f < ce; f += ADJ_THREADS) {
            float4 o[8];
            if (f < fb) face_vjp<L, false>(a, st, grad, j.mu, f, o);
            else face_vjp<L, true>(a, st, grad, j.mu, f, o);
#pragma unroll
            for (int k = 0; k < 8; k++) lds_adj[k * ADJ_CHUNK + (f - cs)] = o[k];
        }
        __syncthreads();
        const uint32_t base = (uint32_t)(cs - f0), n = (uint32_t)(ce - cs);
#pragma unroll
        for (int u = 0; u < ADJ_CELLS; u++)
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const uint32_t s = slot[u][k];
                if (s == 0xffffffffu) continue;
                const uint32_t r = (s & 0x7fffu) - base;
                if (r >= n) continue;
                const int side = (s & 0x8000u) ? 0 : 4;  // bit 15 set: the cell owns the face
                const float4 *src = lds_adj + side * ADJ_CHUNK + r;
#pragma unroll
                for (int q = 0; q < 4; q++) add4(&acc[u][4 * q], src[q * ADJ_CHUNK]);
            }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < ADJ_CELLS; u++) {
        const int i = c0 + (int)threadIdx.x + u * ADJ_THREADS;
        if (i >= c1) continue;
        const float4 l = j.ls[i];
        j.ls_acc[i] = make_float4(l.x + acc[u][0], l.y + acc[u][1], l.z + acc[u][2], l.w + acc[u][3]);
        j.lam_g[3 * (size_t)i + 0] = make_float4(acc[u][4], acc[u][5], acc[u][6], acc[u][7]);
        j.lam_g[3 * (size_t)i + 1] = make_float4(acc[u][8], acc[u][9], acc[u][10], acc[u][11]);
        j.lam_g[3 * (size_t)i + 2] = make_float4(acc[u][12], acc[u][13], acc[u][14], acc[u][15]);
    }
}

// ---------------------------------------------------------------------------
// gradients
// ---------------------------------------------------------------------------
// forward: grad_i[v] = sum_k W_k(i) d_k[v], d_k = st_j - st_i (interior),
// 2 (bc - st_i) for a mirrored ghost value, 0 for a copied one
__device__ __forceinline__ void lam_grad(const float4 *__restrict__ lg, int i, float g[12])
{
    const float4 a = lg[3 * (size_t)i], b = lg[3 * (size_t)i + 1], c = lg[3 * (size_t)i + 2];
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w;
    g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
    g[8] = c.x; g[9] = c.y; g[10] = c.z; g[11] = c.w;
}

// The transpose of the gradients for cell i: its state adjoint (written to
// lam_out and returned).
__device__ __forceinline__ float4 adj_grad_cell(const FluidArgs &a, const AdjFluidArgs &j, int i)
{
    const int n = a.n_cells;
    const float4 l0 = j.ls_acc[i];
    float acc[4] = {l0.x, l0.y, l0.z, l0.w};
    float g[12];
    lam_grad(j.lam_g, i, g);
    // every stencil index and weight first, then the neighbours' rows and
    // adjoint gradients three entries at a time (a boundary entry loads the
    // cell's own, unused), so the loads are in flight together rather than
    // one index and one neighbour at a time
    int jk[6];
#pragma unroll
    for (int k = 0; k < 6; k++) jk[k] = a.stencil[(size_t)k * n + i];
    float wk[18];
#pragma unroll
    for (int k = 0; k < 18; k++) wk[k] = a.gw[(size_t)k * n + i];
#pragma unroll
    for (int k0 = 0; k0 < 6; k0 += 3) {
        float4 rk[3];
        float gk[3][12];
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int jn = jk[k0 + u];
            rk[u] = j.rev_w[(size_t)(k0 + u) * n + i];
            lam_grad(j.lam_g, jn >= 0 ? jn : i, gk[u]);
        }
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int k = k0 + u, jn = jk[k];
            const float wx = wk[3 * k], wy = wk[3 * k + 1], wz = wk[3 * k + 2];
            float c[4] = {1.0f, 1.0f, 1.0f, 1.0f};
            if (jn < 0) {
                const int2 ty = a.bc_type[-1 - jn];
                const bool mv = ty.y == 1 || ty.y == 3;
                c[0] = ty.x == 1 ? 2.0f : 0.0f;
                c[1] = c[2] = c[3] = mv ? 2.0f : 0.0f;
            }
#pragma unroll
            for (int v = 0; v < 4; v++) acc[v] -= c[v] * (wx * g[3 * v] + wy * g[3 * v + 1] + wz * g[3 * v + 2]);
            if (jn >= 0) {
                // the neighbour's stencil row that points back here
                const float4 r = rk[u];
                const float *gc = gk[u];
#pragma unroll
                for (int v = 0; v < 4; v++) acc[v] += r.x * gc[3 * v] + r.y * gc[3 * v + 1] + r.z * gc[3 * v + 2];
            }
        }
    }
    if (j.acc_rw) {
        // the interpolation adjoint is held as v 2^(40 - e), e the binary
        // exponent of the largest adjoint node velocity component
        const double inv_scale = adj_fix_inv_scale(*j.umax);
        unsigned long long *p = j.acc_rw + 3 * (size_t)i;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const long long v = (long long)p[t];
            if (v != 0) {
                acc[1 + t] += (float)((double)v * inv_scale);
                p[t] = 0ull;
            }
        }
    }
    const float4 L = make_float4(acc[0], acc[1], acc[2], acc[3]);
    j.lam_out[i] = L;
    return L;
}

// UPD: the update adjoint of the step before follows in the same pass (its
// state adjoint is the one just computed; every reader of the update
// adjoint's outputs of this step has finished): one launch fewer per step
template <bool UPD>
__global__ __launch_bounds__(256) void k_adj_grad(FluidArgs a, AdjFluidArgs j, AdjPrevStep prev)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double p[3] = {0.0, 0.0, 0.0};
    if (i < a.n_cells) {
        const float4 L = adj_grad_cell(a, j, i);
        if (UPD) adj_update_cell(a, j, prev.cf, i, L, prev.st[i], prev.res[i], p);
    }
    if (UPD) adj_force_partials(j, p);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_adj_update(const FluidArgs &a, const AdjFluidArgs &j, const float4 *st, const float4 *res,
                             hipStream_t s)
{
    const int blocks = (a.n_cells + 255) / 256;
    hipLaunchKernelGGL(k_adj_update, dim3(blocks), dim3(256), 0, s, a, j, st, res);
    return hipGetLastError();
}

hipError_t launch_adj_force_sum(const AdjFluidArgs &j, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_force_sum, dim3(1), dim3(256), 0, s, j);
    return hipGetLastError();
}

template <class L>
static void launch_adj_flux_l(const FluidArgs &a, const AdjFluidArgs &j, const float4 *st, const float4 *grad,
                              int cells_per_thread, size_t lds, hipStream_t s)
{
    if (cells_per_thread == 1)
        hipLaunchKernelGGL((k_adj_flux<L, 1>), dim3(a.n_tiles), dim3(ADJ_THREADS), lds, s, a, j, st, grad);
    else
        hipLaunchKernelGGL((k_adj_flux<L, 2>), dim3(a.n_tiles), dim3(ADJ_THREADS), lds, s, a, j, st, grad);
}

hipError_t launch_adj_flux(const FluidArgs &a, const AdjFluidArgs &j, const float4 *st, const float4 *grad,
                           int max_tile_cells, hipStream_t s)
{
    const size_t lds = sizeof(float4) * 8 * (size_t)ADJ_CHUNK;
    const int cpt = max_tile_cells <= ADJ_THREADS ? 1 : 2;
    switch (a.lattice) {
    case LBFS_D3Q15: launch_adj_flux_l<D3Q15>(a, j, st, grad, cpt, lds, s); break;
    case LBFS_D3Q19: launch_adj_flux_l<D3Q19>(a, j, st, grad, cpt, lds, s); break;
    case LBFS_D2Q9: launch_adj_flux_l<D2Q9>(a, j, st, grad, cpt, lds, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adj_grad(const FluidArgs &a, const AdjFluidArgs &j, const AdjPrevStep *prev, hipStream_t s)
{
    const dim3 grid((a.n_cells + 255) / 256), blk(256);
    if (prev) hipLaunchKernelGGL(k_adj_grad<true>, grid, blk, 0, s, a, j, *prev);
    else hipLaunchKernelGGL(k_adj_grad<false>, grid, blk, 0, s, a, j, AdjPrevStep{});
    return hipGetLastError();
}

}  // namespace lbfs
