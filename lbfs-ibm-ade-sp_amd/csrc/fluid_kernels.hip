// LBFS fluid kernels for gfx950 (FP32 working precision, density stored as
// rho - rho0).
//
//   k_gradients      per-cell least-squares gradients of (rho, u, v, w)
//                    (LBFS.cu:2177-2240, ghost rows of LBFS.cu:1933-2048 folded in)
//   k_flux_update    one workgroup per tile of cells: every face touching the
//                    tile is reconstructed once into LDS (LBFS.cu:2574-2810),
//                    then each cell gathers its six face fluxes in a fixed
//                    order and advances (LBFS.cu:2242-2281, 2089-2172).  No
//                    atomics: results are bitwise repeatable.
//
// Face reconstruction works on deviations from the rest state: with
// g_k = feq_k - w_k rho0, the interface density deviation and momentum are
// sum g_k and sum e_k g_k, and the flux of the interface equilibrium is taken
// from its closed-form moments (rho cs^2 I + rho u u / gamma), which the
// discrete sums of the reference reproduce exactly on these lattices.
#include <hip/hip_runtime.h>

#include <utility>

#include "fluid_device.hpp"
#include "fluid_kernels.hpp"
#include "lattice.hpp"

namespace lbfs {

// ---------------------------------------------------------------------------
// gradients
// ---------------------------------------------------------------------------
// least-squares gradient of cell i (12 components: rho, u, v, w along x, y, z):
// the loads of its stencil (every index, then every neighbour state, a ghost
// entry loading the cell's own, and every weight, so they are in flight
// together rather than one index and one state at a time), then the sums
struct CellStencil {
    float4 si, sk[6];
    int jk[6];
    float wk[18];
};
__device__ __forceinline__ void load_stencil(const FluidArgs &a, const float4 *__restrict__ st, int i, CellStencil &c)
{
    const int n = a.n_cells;
    c.si = st[i];
#pragma unroll
    for (int k = 0; k < 6; k++) c.jk[k] = a.stencil[(size_t)k * n + i];
#pragma unroll
    for (int k = 0; k < 6; k++) c.sk[k] = st[c.jk[k] >= 0 ? c.jk[k] : i];
#pragma unroll
    for (int k = 0; k < 18; k++) c.wk[k] = a.gw[(size_t)k * n + i];
}
__device__ __forceinline__ void stencil_gradient(const FluidArgs &a, const CellStencil &c, float R[12])
{
    const float4 si = c.si;
#pragma unroll
    for (int t = 0; t < 12; t++) R[t] = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int j = c.jk[k];
        float dr, du, dv, dw;
        if (j >= 0) {
            const float4 sj = c.sk[k];
            dr = sj.x - si.x; du = sj.y - si.y; dv = sj.z - si.z; dw = sj.w - si.w;
        } else {
            // ghost: Dirichlet mirror 2 bc - owner (velocity codes 1 and 3),
            // Neumann and empty copy (Solution.cpp:759-833, LBFS.cu:1843-1928)
            const int bc = -1 - j;
            const int2 ty = a.bc_type[bc];
            const float4 bv = a.bc_val[bc];
            const bool mv = ty.y == 1 || ty.y == 3;
            dr = ty.x == 1 ? 2.0f * (bv.x - si.x) : 0.0f;
            du = mv ? 2.0f * (bv.y - si.y) : 0.0f;
            dv = mv ? 2.0f * (bv.z - si.z) : 0.0f;
            dw = mv ? 2.0f * (bv.w - si.w) : 0.0f;
        }
        const float wx = c.wk[3 * k], wy = c.wk[3 * k + 1], wz = c.wk[3 * k + 2];
        R[0] += wx * dr; R[1] += wy * dr; R[2] += wz * dr;
        R[3] += wx * du; R[4] += wy * du; R[5] += wz * du;
        R[6] += wx * dv; R[7] += wy * dv; R[8] += wz * dv;
        R[9] += wx * dw; R[10] += wy * dw; R[11] += wz * dw;
    }
}
__device__ __forceinline__ void cell_gradient(const FluidArgs &a, const float4 *__restrict__ st, int i, float R[12])
{
    CellStencil c;
    load_stencil(a, st, i, c);
    stencil_gradient(a, c, R);
}

__global__ __launch_bounds__(256) void k_gradients(FluidArgs a, const float4 *__restrict__ st,
                                                   float4 *__restrict__ grad)
{
    const int nblk = gridDim.x;
    const int b = xcd_remap(blockIdx.x, nblk);
    const int i = b * 256 + threadIdx.x;
    if (i >= a.n_cells) return;
    float R[12];
    cell_gradient(a, st, i, R);
    grad[3 * (size_t)i + 0] = make_float4(R[0], R[1], R[2], R[3]);
    grad[3 * (size_t)i + 1] = make_float4(R[4], R[5], R[6], R[7]);
    grad[3 * (size_t)i + 2] = make_float4(R[8], R[9], R[10], R[11]);
}

// Sources of the cell gradients a face reads: the gradient array of the
// two-pass step, or the fused step's tile (the gradients of the tile's cells
// and of the cells across its border in LDS, at the face's slots)
struct GlobalGrad {
    const float4 *__restrict__ grad;
    __device__ __forceinline__ void load_owner(int, int i, float g[12]) const { load_grad(grad, i, g); }
    __device__ __forceinline__ void load_nbr(int, int i, float g[12]) const { load_grad(grad, i, g); }
};
struct TileGrad {
    const float4 *lg;              // LDS: 3 float4 per gradient slot
    const uint32_t *__restrict__ slots;  // per tile-face: owner slot | neighbour slot << 16
    __device__ __forceinline__ void at(int q, float g[12]) const
    {
        const float4 x = lg[3 * q], y = lg[3 * q + 1], z = lg[3 * q + 2];
        g[0] = x.x; g[1] = x.y; g[2] = x.z; g[3] = x.w;
        g[4] = y.x; g[5] = y.y; g[6] = y.z; g[7] = y.w;
        g[8] = z.x; g[9] = z.y; g[10] = z.z; g[11] = z.w;
    }
    __device__ __forceinline__ void load_owner(int tf, int, float g[12]) const { at(slots[tf] & 0xffffu, g); }
    __device__ __forceinline__ void load_nbr(int tf, int, float g[12]) const { at(slots[tf] >> 16, g); }
};

// ---------------------------------------------------------------------------
// face reconstruction and flux
// ---------------------------------------------------------------------------
template <class L, bool BAND, class G>
__device__ __forceinline__ float4 face_flux(const FluidArgs &a, const float4 *__restrict__ st, const G &gsrc, int tf)
{
    static_assert(opposite_pairs<L>(), "lattice directions must come in opposite pairs after the rest point");
    const FaceRec fr = load_face(a, tf);
    const float4 fb = fr.d1, fcv = fr.d2;
    const int o = fr.o, nb = fr.nb;
    const bool ghost = nb >= a.n_cells;
    const float nx = fr.nx, ny = fr.ny, nz = fr.nz, A = fr.A, dt = fr.dt;
    const float4 so = st[o];
    float go[12], gn[12];
    gsrc.load_owner(tf, o, go);
    const float sov[4] = {so.x, so.y, so.z, so.w};
    float po[4], pn[4], p0[4];
    if (!ghost) {
        const float4 sn = st[nb];
        gsrc.load_nbr(tf, nb, gn);
        const float snv[4] = {sn.x, sn.y, sn.z, sn.w};
#pragma unroll
        for (int v = 0; v < 4; v++) {
            po[v] = sov[v] - (fb.x * go[3 * v] + fb.y * go[3 * v + 1] + fb.z * go[3 * v + 2]);
            pn[v] = snv[v] - (fcv.x * gn[3 * v] + fcv.y * gn[3 * v + 1] + fcv.z * gn[3 * v + 2]);
            p0[v] = 0.5f * (po[v] + pn[v]);
        }
    } else {
        // ghost gradient (gradients.cpp:393-505 / LBFS.cu:1933-2048) and face
        // value from the boundary condition (Solver.cpp:855-900)
        const int bc = nb - a.n_cells;
        const int2 ty = a.bc_type[bc];
        if (ty.y == LBFS_BC_EMPTY) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // Solver.cpp:308-310
        const float4 bv = a.bc_val[bc];
        const float bvv[4] = {bv.x, bv.y, bv.z, bv.w};
        const float d1n = sqrtf(fb.x * fb.x + fb.y * fb.y + fb.z * fb.z);
        const float inv_dmag = 2.0f / d1n;  // ghost centroid sits halfway to the face
        float dir[3];
        if (a.flags & LBFS_FLAG_GHOST_DIR_GPU) { dir[0] = -fb.x / d1n; dir[1] = -fb.y / d1n; dir[2] = -fb.z / d1n; }
        else { dir[0] = nx; dir[1] = ny; dir[2] = nz; }
        const float nv[3] = {nx, ny, nz};
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int type = v == 0 ? ty.x : ty.y;
            if (type == 1 || type == 3) {
                const float dphi = 2.0f * (bvv[v] - sov[v]) * inv_dmag;
#pragma unroll
                for (int c = 0; c < 3; c++) gn[3 * v + c] = dphi * dir[c];
                // code 3 (parabolic/shear targets) extrapolates the face value
                // with the ghost gradient (LBFS.cu:358-378)
                p0[v] = type == 1 ? bvv[v] : sov[v] - (fb.x * gn[3 * v] + fb.y * gn[3 * v + 1] + fb.z * gn[3 * v + 2]);
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float fac = (a.flags & LBFS_FLAG_NEUMANN_SIGNED) ? 1.0f - nv[c] : 1.0f - fabsf(nv[c]);
                    gn[3 * v + c] = go[3 * v + c] * fac;
                }
                p0[v] = sov[v] - (fb.x * gn[3 * v] + fb.y * gn[3 * v + 1] + fb.z * gn[3 * v + 2]);
            }
            po[v] = p0[v];
            pn[v] = p0[v];
        }
        // the GPU reconstruction hands the ghost gradient to the owner-side
        // points too (grad_*_1 = grad_*_2, LBFS.cu:391-394)
        if (a.flags & LBFS_FLAG_GHOST_FACE_GRAD) {
#pragma unroll
            for (int t = 0; t < 12; t++) go[t] = gn[t];
        }
    }
#pragma unroll
    for (int t = 0; t < 12; t++) { go[t] *= dt; gn[t] *= dt; }
    const float rho0 = a.rho0, ig = a.inv_gamma;
    const float bw = 0.001f * dt;
    FaceSums f;
    f.S1x = f.S1y = f.S1z = f.Tx = f.Ty = f.Tz = 0.0f;
    {
        const float uu = p0[1] * p0[1] + p0[2] * p0[2] + p0[3] * p0[3];
        const float P = -1.5f * uu * ig;
        f.S0 = L::w(0) * (p0[0] + (rho0 + p0[0]) * P);
    }
    all_pairs<L, BAND>(std::make_integer_sequence<int, (L::Q - 1) / 2>{}, f, po, pn, p0, go, gn, nx, ny, nz, bw, ghost,
                       rho0, ig);
    const float S0 = f.S0, S1x = f.S1x, S1y = f.S1y, S1z = f.S1z, Tx = f.Tx, Ty = f.Ty, Tz = f.Tz;
    if (a.face_moments) a.face_moments[tf] = make_float4(S0, S1x, S1y, S1z);
    const float rho_f = rho0 + S0;
    const float mn = S1x * nx + S1y * ny + S1z * nz;
    const float vf = a.visc_factor ? a.visc_factor[o] : 1.0f;
    const float tau = a.visc3_over_gamma * vf * __builtin_amdgcn_rcpf(dt);
    const float c1 = 1.0f - tau;
    const float third = S0 * (1.0f / 3.0f);
    const float q = mn * __builtin_amdgcn_rcpf(rho_f) * ig;
    float4 r;
    r.x = mn * A;
    r.y = A * (c1 * (third * nx + S1x * q) + tau * Tx);
    r.z = A * (c1 * (third * ny + S1y * q) + tau * Ty);
    // a planar lattice carries no z momentum flux (its sum e_z e_z w is zero)
    r.w = L::dims == 3 ? A * (c1 * (third * nz + S1z * q) + tau * Tz) : 0.0f;
    return r;
}

// the cell update of a tile from its face fluxes in LDS: each cell gathers
// its six signed face slots in a fixed order (LBFS.cu:2242-2281), adds the
// force density and advances (LBFS.cu:2089-2172)
template <int NT>
__device__ __forceinline__ void tile_update(const FluidArgs &a, const float4 *lds_flux, int c0, int c1,
                                            const float4 *__restrict__ st, float4 *__restrict__ st_out,
                                            float4 *__restrict__ res_out, int do_update)
{
    for (int i = c0 + threadIdx.x; i < c1; i += NT) {
        const uint32_t *sl = reinterpret_cast<const uint32_t *>(a.cell_slots) + 3 * (size_t)i;
        const uint32_t s01 = sl[0], s23 = sl[1], s45 = sl[2];
        // the cell's state and force loaded with its slots, not after the sums
        const float4 s0 = do_update ? st[i] : float4{};
        float cfv[3] = {0.0f, 0.0f, 0.0f};
        if (do_update && a.cell_force) {
            const float *cf = a.cell_force + 3 * (size_t)i;
            cfv[0] = cf[0]; cfv[1] = cf[1]; cfv[2] = cf[2];
        }
        const uint32_t s[6] = {s01 & 0xffffu, s01 >> 16, s23 & 0xffffu, s23 >> 16, s45 & 0xffffu, s45 >> 16};
        float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const float4 fl = lds_flux[s[k] & 0x7fffu];
            const float sg = (s[k] & 0x8000u) ? -1.0f : 1.0f;
            r0 += sg * fl.x; r1 += sg * fl.y; r2 += sg * fl.z; r3 += sg * fl.w;
        }
        const float iv = a.inv_vol[i];
        r0 *= iv; r1 *= iv; r2 *= iv; r3 *= iv;
        if (res_out) res_out[i] = make_float4(r0, r1, r2, r3);
        if (!do_update) continue;
        float fx = a.body_force[0], fy = a.body_force[1], fz = a.body_force[2];
        if (a.cell_force) {
            fx += cfv[0]; fy += cfv[1]; fz += cfv[2];
        }
        r1 += fx; r2 += fy; r3 += fz;
        float dt = a.dt;
        if (a.ts_mode == LBFS_TS_LOCAL) {
            const float2 c = a.cfl_coef[i];
            dt = a.dt / (c.x + c.y / (a.rho0 + s0.x));
        } else if (a.ts_mode == LBFS_TS_MIN) {
            dt = __uint_as_float(*a.dt_min);
        }
        const float drho = s0.x + dt * r0;
        const float rho = a.rho0 + drho;
        const float inv_rho = 1.0f / rho;
        float4 sn;
        sn.x = drho;
        sn.y = s0.y + dt * (r1 - s0.y * r0) * inv_rho;
        sn.z = s0.z + dt * (r2 - s0.z * r0) * inv_rho;
        sn.w = s0.w + dt * (r3 - s0.w * r0) * inv_rho;
        st_out[i] = sn;
    }
}

// two-pass step, second pass: the tile's faces from the gradient array
template <class L, bool BAND, int NT>
__global__ __launch_bounds__(NT) void k_flux_update(FluidArgs a, const float4 *__restrict__ st,
                                                     const float4 *__restrict__ grad, float4 *__restrict__ st_out,
                                                     float4 *__restrict__ res_out, int do_update)
{
    extern __shared__ float4 lds_flux[];
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int f0 = a.tile_face_start[t], f1 = a.tile_face_start[t + 1];
    const GlobalGrad gsrc{grad};
    for (int f = f0 + threadIdx.x; f < f1; f += blockDim.x)
        lds_flux[f - f0] = face_flux<L, BAND>(a, st, gsrc, f);
    __syncthreads();
    tile_update<NT>(a, lds_flux, a.tile_cell_start[t], a.tile_cell_start[t + 1], st, st_out, res_out, do_update);
}

// Fused step, one workgroup per tile: the gradients of the tile's cells and
// of the cells across its border that its faces reach into LDS (the same
// arithmetic as the gradient pass), the tile's faces from them, then the cell
// update.  Results equal the two-pass step bit for bit; the gradient array is
// not read, and written (the tile's own cells) only when a.grad_out is set.
template <class L, bool BAND, int NT>
__global__ __launch_bounds__(NT) void k_fluid_fused(FluidArgs a, const float4 *__restrict__ st,
                                                     float4 *__restrict__ st_out, float4 *__restrict__ res_out,
                                                     int do_update)
{
    extern __shared__ float4 lds[];
    float4 *lg = lds;                                   // 3 per gradient slot
    float4 *lds_flux = lds + 3 * (size_t)a.max_grad_slots;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int c0 = a.tile_cell_start[t], c1 = a.tile_cell_start[t + 1], nc = c1 - c0;
    const int h0 = a.tile_halo_start[t], nq = nc + a.tile_halo_start[t + 1] - h0;
    for (int q = threadIdx.x; q < nq; q += NT) {
        const int i = q < nc ? c0 + q : a.halo_ids[h0 + q - nc];
        float R[12];
        cell_gradient(a, st, i, R);
        float4 *o = lg + 3 * q;
        o[0] = make_float4(R[0], R[1], R[2], R[3]);
        o[1] = make_float4(R[4], R[5], R[6], R[7]);
        o[2] = make_float4(R[8], R[9], R[10], R[11]);
        if (a.grad_out && q < nc) {
            float4 *g = a.grad_out + 3 * (size_t)i;
            g[0] = o[0];
            g[1] = o[1];
            g[2] = o[2];
        }
    }
    __syncthreads();
    const int f0 = a.tile_face_start[t], f1 = a.tile_face_start[t + 1];
    const TileGrad gsrc{lg, a.face_gslot};
    for (int f = f0 + threadIdx.x; f < f1; f += NT) lds_flux[f - f0] = face_flux<L, BAND>(a, st, gsrc, f);
    __syncthreads();
    tile_update<NT>(a, lds_flux, c0, c1, st, st_out, res_out, do_update);
}

template <class L, bool BAND>
__global__ __launch_bounds__(256) void k_face_flux_all(FluidArgs a, const float4 *__restrict__ st,
                                                       const float4 *__restrict__ grad, float4 *__restrict__ out,
                                                       int n_tile_faces)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n_tile_faces)
        out[f] = face_flux<L, BAND>(a, st, GlobalGrad{grad}, f);
}

// ---------------------------------------------------------------------------
// global minimum of the local time steps (min<128> + min_block,
// Solver_gpu.cu:773-776): positive floats order like their bit patterns, so an
// unsigned atomic minimum is exact and independent of the visiting order
// ---------------------------------------------------------------------------
__global__ void k_dt_reset(unsigned int *dt_min) { *dt_min = 0x7f800000u; }

__global__ __launch_bounds__(256) void k_dt_min(FluidArgs a, const float4 *__restrict__ st)
{
    __shared__ float sh[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float v = INFINITY;
    if (i < a.n_cells) {
        const float2 c = a.cfl_coef[i];
        v = a.dt / (c.x + c.y / (a.rho0 + st[i].x));
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp) sh[threadIdx.x] = fminf(sh[threadIdx.x], sh[threadIdx.x + stp]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(a.dt_min, __float_as_uint(sh[0]));
}

// ---------------------------------------------------------------------------
// FP64 diagnostics: per-block partial sums (fixed order), summed on the host
// ---------------------------------------------------------------------------
// (force: the momentum force density added to the residual, 3 per cell, as
// the reference adds force_x/y/z before squaring, Solver_gpu.cu:1008-1046)
__global__ __launch_bounds__(256) void k_diag_partials(int n, float rho0, const float4 *__restrict__ st,
                                                       const float4 *__restrict__ res, const float *__restrict__ force,
                                                       const float *__restrict__ inv_vol, double *__restrict__ part)
{
    __shared__ double sh[6][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const float4 s = st[i];
        const float4 r = res[i];
        double r1 = r.y, r2 = r.z, r3 = r.w;
        if (force) {
            r1 += (double)force[3 * (size_t)i];
            r2 += (double)force[3 * (size_t)i + 1];
            r3 += (double)force[3 * (size_t)i + 2];
        }
        v[0] = (double)r.x * r.x;
        v[1] = r1 * r1;
        v[2] = r2 * r2;
        v[3] = r3 * r3;
        v[4] = ((double)rho0 + (double)s.x) / (double)inv_vol[i];
        const double sp = sqrt((double)s.y * s.y + (double)s.z * s.z + (double)s.w * s.w);
        v[5] = (isfinite(s.x) && isfinite(s.y) && isfinite(s.z) && isfinite(s.w)) ? sp : INFINITY;
    }
    for (int t = 0; t < 6; t++) sh[t][threadIdx.x] = v[t];
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if ((int)threadIdx.x < stp) {
            for (int t = 0; t < 5; t++) sh[t][threadIdx.x] += sh[t][threadIdx.x + stp];
            sh[5][threadIdx.x] = fmax(sh[5][threadIdx.x], sh[5][threadIdx.x + stp]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_gradients(const FluidArgs &a, const float4 *st, float4 *grad, hipStream_t s)
{
    const int blocks = (a.n_cells + 255) / 256;
    hipLaunchKernelGGL(k_gradients, dim3(blocks), dim3(256), 0, s, a, st, grad);
    return hipGetLastError();
}

// instantiate a kernel template for the context's lattice and tie-band mode
template <template <class, bool> class K, class... Args>
static hipError_t dispatch(const FluidArgs &a, dim3 grid, size_t lds, hipStream_t s, Args... args)
{
    const bool band = (a.flags & LBFS_FLAG_UPWIND_BAND) != 0;
    switch (a.lattice * 2 + (band ? 1 : 0)) {
    case 0: K<D3Q15, false>::launch(grid, lds, s, a, args...); break;
    case 1: K<D3Q15, true>::launch(grid, lds, s, a, args...); break;
    case 2: K<D3Q19, false>::launch(grid, lds, s, a, args...); break;
    case 3: K<D3Q19, true>::launch(grid, lds, s, a, args...); break;
    case 4: K<D2Q9, false>::launch(grid, lds, s, a, args...); break;
    case 5: K<D2Q9, true>::launch(grid, lds, s, a, args...); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class L, bool BAND>
struct FluxUpdate {
    static void launch(dim3 g, size_t lds, hipStream_t s, const FluidArgs &a, const float4 *st, const float4 *grad,
                       float4 *st_out, float4 *res_out, int do_update)
    {
        if (a.block_threads == 576)
            hipLaunchKernelGGL((k_flux_update<L, BAND, 576>), g, dim3(576), lds, s, a, st, grad, st_out, res_out, do_update);
        else if (a.block_threads == 512)
            hipLaunchKernelGGL((k_flux_update<L, BAND, 512>), g, dim3(512), lds, s, a, st, grad, st_out, res_out, do_update);
        else
            hipLaunchKernelGGL((k_flux_update<L, BAND, 256>), g, dim3(256), lds, s, a, st, grad, st_out, res_out, do_update);
    }
};

template <class L, bool BAND>
struct FaceFluxAll {
    static void launch(dim3 g, size_t lds, hipStream_t s, const FluidArgs &a, const float4 *st, const float4 *grad,
                       float4 *out, int n)
    {
        hipLaunchKernelGGL((k_face_flux_all<L, BAND>), g, dim3(256), lds, s, a, st, grad, out, n);
    }
};

hipError_t launch_flux_update(const FluidArgs &a, const float4 *st, const float4 *grad, float4 *st_out,
                              float4 *res_out, int do_update, hipStream_t s)
{
    const size_t lds = sizeof(float4) * (size_t)a.max_tile_faces;
    return dispatch<FluxUpdate>(a, dim3(a.n_tiles), lds, s, st, grad, st_out, res_out, do_update);
}

template <class L, bool BAND>
struct FluidFused {
    static void launch(dim3 g, size_t lds, hipStream_t s, const FluidArgs &a, const float4 *st, float4 *st_out,
                       float4 *res_out, int do_update)
    {
        if (a.block_threads == 576)
            hipLaunchKernelGGL((k_fluid_fused<L, BAND, 576>), g, dim3(576), lds, s, a, st, st_out, res_out, do_update);
        else if (a.block_threads == 512)
            hipLaunchKernelGGL((k_fluid_fused<L, BAND, 512>), g, dim3(512), lds, s, a, st, st_out, res_out, do_update);
        else
            hipLaunchKernelGGL((k_fluid_fused<L, BAND, 256>), g, dim3(256), lds, s, a, st, st_out, res_out, do_update);
    }
};

size_t fluid_fused_lds(const FluidArgs &a)
{
    return sizeof(float4) * (3 * (size_t)a.max_grad_slots + (size_t)a.max_tile_faces);
}

hipError_t launch_fluid_fused(const FluidArgs &a, const float4 *st, float4 *st_out, float4 *res_out, int do_update,
                              hipStream_t s)
{
    return dispatch<FluidFused>(a, dim3(a.n_tiles), fluid_fused_lds(a), s, st, st_out, res_out, do_update);
}

hipError_t launch_face_flux_all(const FluidArgs &a, const float4 *st, const float4 *grad, float4 *out, int n_tile_faces,
                                hipStream_t s)
{
    const int blocks = (n_tile_faces + 255) / 256;
    return dispatch<FaceFluxAll>(a, dim3(blocks), 0, s, st, grad, out, n_tile_faces);
}

hipError_t launch_dt_min(const FluidArgs &a, const float4 *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_reset, dim3(1), dim3(1), 0, s, a.dt_min);
    hipLaunchKernelGGL(k_dt_min, dim3((a.n_cells + 255) / 256), dim3(256), 0, s, a, st);
    return hipGetLastError();
}

hipError_t launch_diag_partials(int n, float rho0, const float4 *st, const float4 *res, const float *force,
                                const float *inv_vol, double *part,
                                hipStream_t s)
{
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_diag_partials, dim3(blocks), dim3(256), 0, s, n, rho0, st, res, force, inv_vol, part);
    return hipGetLastError();
}

}  // namespace lbfs
