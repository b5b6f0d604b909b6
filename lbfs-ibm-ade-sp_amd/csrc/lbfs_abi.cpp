// C ABI implementation: context set-up, state transfer and the time step.
#include "lbfs_abi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "fluid_kernels.hpp"
#include "geometry_gpu.hpp"
#include "mesh_host.hpp"

using namespace lbfs;

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return ctx->fail(LBFS_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int lbfs_ctx::fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

template <class T>
static int dev_upload(lbfs_ctx *ctx, T **dst, const T *src, size_t n)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, n * sizeof(T) + 16));
    ctx->allocs.push_back(p);
    if (src) HIPCHK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = (T *)p;
    return LBFS_OK;
}

static thread_local std::string g_create_error;

extern "C" const char *lbfs_last_error(const lbfs_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

static int fluid_setup(lbfs_ctx *ctx, const lbfs_mesh_desc &m, const lbfs_params &p)
{
    std::string err;
    const bool closed = (p.flags & LBFS_FLAG_CLOSED_FACES) != 0, det_compat = (p.flags & LBFS_FLAG_LS_DET_COMPAT) != 0;
    std::vector<double> lhs, rhs;
    int rc;
    if (p.flags & LBFS_FLAG_HOST_GEOMETRY) {
        rc = compute_geometry(m, p.visc, p.gamma, p.fneq_min, closed, ctx->geo, err);
        if (rc) return ctx->fail(rc, "geometry: %s", err.c_str());
        rc = compute_ls(ctx->geo, p.gradient_type, det_compat, lhs, rhs, err);
        if (rc) return ctx->fail(rc, "least squares: %s", err.c_str());
    } else {
        // the face and cell passes on the GPU, the same bits as the host's
        rc = gpu_geometry(m, p.visc, p.gamma, p.fneq_min, closed, p.gradient_type, det_compat, ctx->stream, ctx->geo,
                          lhs, rhs, err);
        if (rc) return ctx->fail(rc, "geometry: %s", err.c_str());
    }
    int tile[3] = {8, 8, 4};
    if (p.tile[0] || p.tile[1] || p.tile[2]) {
        if (p.tile[0] < 1 || p.tile[1] < 1 || p.tile[2] < 1 || p.tile[0] * p.tile[1] * p.tile[2] > 2048)
            return ctx->fail(LBFS_E_INVALID_ARG, "tile %dx%dx%d: each side >= 1, at most 2048 cells", p.tile[0],
                             p.tile[1], p.tile[2]);
        for (int t = 0; t < 3; t++) tile[t] = p.tile[t];
    }
    if (p.block_threads != 0 && p.block_threads != 256 && p.block_threads != 512 && p.block_threads != 576)
        return ctx->fail(LBFS_E_INVALID_ARG, "block_threads must be 256, 512 or 576");
    rc = mesh_blocks(m, p.body_force, ctx->blocks, err);
    if (rc) return ctx->fail(rc, "mesh blocks: %s", err.c_str());
    rc = build_layout(m, ctx->geo, lhs, rhs, p.rho0, p.body_force, p.lattice == LBFS_D2Q9 ? 2 : 3, tile, ctx->blocks,
                      ctx->lay, err);
    if (rc) return ctx->fail(rc, "layout: %s", err.c_str());
    const DeviceLayout &d = ctx->lay;
    const int n = d.n_cells, nbc = d.n_bc;
    FluidArgs &a = ctx->args;
    memset(&a, 0, sizeof(a));
    a.n_cells = n;
    a.n_bc = nbc;
    a.n_tiles = d.n_tiles;
    a.max_tile_faces = d.max_tile_faces;
    a.block_threads = p.block_threads == 512 || p.block_threads == 576 ? p.block_threads : 256;
    if ((size_t)d.max_tile_faces * 16 > 96 * 1024)
        return ctx->fail(LBFS_E_INVALID_ARG, "tile too large: %d faces need more than 96 KiB of LDS", d.max_tile_faces);
    for (int t = 0; t < d.n_tiles; t++)
        a.max_tile_cells = std::max(a.max_tile_cells, d.tile_cell_start[t + 1] - d.tile_cell_start[t]);
    a.max_grad_slots = d.max_grad_slots;
    {
        int *ths, *hid;
        uint32_t *fgs;
        if ((rc = dev_upload(ctx, &ths, d.tile_halo_start.data(), d.tile_halo_start.size()))) return rc;
        if ((rc = dev_upload(ctx, &hid, d.halo_ids.data(), d.halo_ids.size()))) return rc;
        if ((rc = dev_upload(ctx, &fgs, d.face_gslot.data(), d.face_gslot.size()))) return rc;
        a.tile_halo_start = ths;
        a.halo_ids = hid;
        a.face_gslot = fgs;
    }
    // the fused fluid pass keeps the tile's gradients in LDS beside its face
    // fluxes; larger tiles run the two passes
    ctx->fused_ok = fluid_fused_lds(a) <= 96 * 1024;
    ctx->fused = ctx->fused_ok;
    a.lattice = p.lattice;
    a.flags = p.flags;
    a.rho0 = (float)p.rho0;
    a.inv_gamma = (float)(1.0 / p.gamma);
    a.visc3_over_gamma = (float)(p.visc * 3.0 / p.gamma);
    a.dt = (float)p.dt;
    for (int t = 0; t < 3; t++) a.body_force[t] = (float)p.body_force[t];
    int *stencil; float *gw, *inv_vol; int16_t *slots; int *tcs, *tfs;
    float4 *fA, *fB, *fC; int2 *bct; float4 *bcv;
    if ((rc = dev_upload(ctx, &stencil, d.stencil.data(), d.stencil.size()))) return rc;
    if ((rc = dev_upload(ctx, &gw, d.gw.data(), d.gw.size()))) return rc;
    if ((rc = dev_upload(ctx, &inv_vol, d.inv_vol.data(), d.inv_vol.size()))) return rc;
    if ((rc = dev_upload(ctx, &slots, d.cell_slots.data(), d.cell_slots.size()))) return rc;
    if ((rc = dev_upload(ctx, &tcs, d.tile_cell_start.data(), d.tile_cell_start.size()))) return rc;
    if ((rc = dev_upload(ctx, &tfs, d.tile_face_start.data(), d.tile_face_start.size()))) return rc;
    int *tbs;
    if ((rc = dev_upload(ctx, &tbs, d.tile_bface_start.data(), d.tile_bface_start.size()))) return rc;
    a.tile_bface_start = tbs;
    const size_t ntf = d.face_ref.size();
    if ((rc = dev_upload(ctx, &fA, (const float4 *)d.fA.data(), ntf))) return rc;
    if ((rc = dev_upload(ctx, &fB, (const float4 *)d.fB.data(), ntf))) return rc;
    if ((rc = dev_upload(ctx, &fC, (const float4 *)d.fC.data(), ntf))) return rc;
    std::vector<int2> bt(nbc > 0 ? nbc : 1);
    ctx->bc_vel_type.assign(nbc, LBFS_BC_NEUMANN);
    std::vector<float4> bvals(nbc > 0 ? nbc : 1);
    for (int b = 0; b < nbc; b++) {
        const int rt = m.bc_rho_type ? m.bc_rho_type[b] : LBFS_BC_NEUMANN;
        const int vt = m.bc_vel_type ? m.bc_vel_type[b] : LBFS_BC_NEUMANN;
        const bool rt_ok = rt == LBFS_BC_DIRICHLET || rt == LBFS_BC_NEUMANN || rt == LBFS_BC_EMPTY;
        const bool vt_ok = vt == LBFS_BC_DIRICHLET || vt == LBFS_BC_NEUMANN || vt == LBFS_BC_PARABOLIC ||
                           vt == LBFS_BC_SHEAR || vt == LBFS_BC_EMPTY;
        if (!rt_ok || !vt_ok)
            return ctx->fail(LBFS_E_UNSUPPORTED, "boundary face %d: unsupported BC types (rho %d, vel %d)", b, rt, vt);
        if ((rt == LBFS_BC_EMPTY) != (vt == LBFS_BC_EMPTY))
            return ctx->fail(LBFS_E_INVALID_ARG, "boundary face %d: empty must be set for rho and velocity", b);
        if ((vt == LBFS_BC_PARABOLIC) && !(p.channel_diameter > 0.0))
            return ctx->fail(LBFS_E_INVALID_ARG, "parabolic velocity BC needs channel_diameter > 0");
        const double *bv = m.bc_values ? m.bc_values + 4 * (size_t)b : nullptr;
        double u = bv ? bv[1] : 0.0;
        // device codes: rho 1 mirror / 2 copy; velocity 1 mirror with the BC
        // value at the face, 3 mirror about a target evaluated once at the
        // ghost centroid with the face value extrapolated (LBFS.cu:1892-1918,
        // 358-378), 2 copy, LBFS_BC_EMPTY no flux
        int dvt = vt;
        if (vt == LBFS_BC_PARABOLIC || vt == LBFS_BC_SHEAR) {
            const double *c = &ctx->geo.cc[3 * ((size_t)n + b)];
            if (vt == LBFS_BC_PARABOLIC) {
                const double H = p.channel_diameter, H2 = 0.5 * H, H4 = H * H * H * H;
                u = 16.0 * u * (H2 + c[1]) * (H2 - c[1]) * (H2 + c[2]) * (H2 - c[2]) / H4;
            } else {
                u = c[1] * u;
            }
            dvt = 3;
        }
        bt[b] = make_int2(rt == LBFS_BC_DIRICHLET ? 1 : 2, dvt);
        ctx->bc_vel_type[b] = vt;
        bvals[b] = make_float4(bv ? (float)(bv[0] - p.rho0) : 0.0f, (float)u, bv ? (float)bv[2] : 0.0f,
                               bv ? (float)bv[3] : 0.0f);
    }
    if ((rc = dev_upload(ctx, &bct, bt.data(), bt.size()))) return rc;
    if ((rc = dev_upload(ctx, &bcv, bvals.data(), bvals.size()))) return rc;
    a.stencil = stencil; a.gw = gw; a.inv_vol = inv_vol; a.cell_slots = slots;
    a.tile_cell_start = tcs; a.tile_face_start = tfs;
    a.fA = fA; a.fB = fB; a.fC = fC;
    a.bc_type = bct; a.bc_val = bcv;
    a.visc_factor = nullptr;
    // cell force density: body force + closure term, IBM forces are added on top
    float *sf, *cf;
    if ((rc = dev_upload(ctx, &sf, d.static_force.data(), d.static_force.size()))) return rc;
    if ((rc = dev_upload(ctx, &cf, d.static_force.data(), d.static_force.size()))) return rc;
    ctx->static_force = sf;
    ctx->cell_force = cf;  // output: static + spread force of the last coupled step
    a.cell_force = sf;
    // local / minimum time steps: per-cell coefficients of dt V / lambda
    if (p.time_stepping == LBFS_TS_LOCAL || p.time_stepping == LBFS_TS_MIN) {
        if (!(p.max_velocity > 0.0)) return ctx->fail(LBFS_E_INVALID_ARG, "local time stepping needs max_velocity > 0");
        const Geometry &g = ctx->geo;
        std::vector<double> Ac(3 * (size_t)n, 0.0);
        for (int f = 0; f < g.n_faces; f++)
            for (int t = 0; t < 3; t++) {
                const double v = 0.5 * std::fabs(g.normal[3 * (size_t)f + t] * g.area[f]);
                Ac[3 * (size_t)g.owner[f] + t] += v;
                if (g.nbr[f] < n) Ac[3 * (size_t)g.nbr[f] + t] += v;
            }
        const double U = p.max_velocity, gm = p.gamma;
        const double c = U * std::sqrt(1.0 - gm + std::pow(gm / std::sqrt(3.0) / U, 2.0));
        std::vector<float2> coef(n);
        for (int i = 0; i < n; i++) {
            const int cc = d.perm[i];
            const double *A = &Ac[3 * (size_t)cc];
            const double V = g.vol[cc];
            coef[i] = make_float2((float)((U + c) * (A[0] + A[1] + A[2]) / (gm * V)),
                                  (float)(8.0 * p.visc * (A[0] * A[0] + A[1] * A[1] + A[2] * A[2]) / (gm * V * V)));
        }
        float2 *dcoef;
        unsigned int *dmin;
        if ((rc = dev_upload(ctx, &dcoef, coef.data(), coef.size()))) return rc;
        if ((rc = dev_upload<unsigned int>(ctx, &dmin, nullptr, 1))) return rc;
        a.cfl_coef = dcoef;
        a.dt_min = dmin;
        a.ts_mode = p.time_stepping;
    } else if (p.time_stepping != 0 && p.time_stepping != LBFS_TS_CONSTANT) {
        return ctx->fail(LBFS_E_INVALID_ARG, "unknown time stepping %d", p.time_stepping);
    }
    for (int t = 0; t < 3; t++) a.body_force[t] = 0.0f;
    // state buffers
    if ((rc = dev_upload<float4>(ctx, &ctx->st[0], nullptr, n))) return rc;
    if ((rc = dev_upload<float4>(ctx, &ctx->st[1], nullptr, n))) return rc;
    if ((rc = dev_upload<float4>(ctx, &ctx->grad, nullptr, 3 * (size_t)n))) return rc;
    if ((rc = dev_upload<float4>(ctx, &ctx->res, nullptr, n))) return rc;
    if ((rc = dev_upload<float4>(ctx, &ctx->face_scratch, nullptr, ntf))) return rc;
    if ((rc = dev_upload<double>(ctx, &ctx->diag_part, nullptr, 6 * (size_t)((n + 255) / 256)))) return rc;
    if ((rc = dev_upload<int>(ctx, &ctx->dev_err, nullptr, 1))) return rc;
    HIPCHK(hipMemsetAsync(ctx->dev_err, 0, sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->st[0], 0, sizeof(float4) * n, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->st[1], 0, sizeof(float4) * n, ctx->stream));
    ctx->cur = 0;
    return LBFS_OK;
}

extern "C" int lbfs_create(const lbfs_mesh_desc *mesh, const lbfs_params *params, lbfs_ctx **out)
{
    g_create_error.clear();
    if (!out) return LBFS_E_INVALID_ARG;
    *out = nullptr;
    if (!mesh || !params) { g_create_error = "null mesh or params"; return LBFS_E_INVALID_ARG; }
    if (params->lattice < 0 || params->lattice > 2) { g_create_error = "unknown lattice"; return LBFS_E_INVALID_ARG; }
    if (!(params->gamma > 0.0) || !(params->visc >= 0.0) || !(params->dt > 0.0)) {
        g_create_error = "gamma and dt must be positive, visc non-negative";
        return LBFS_E_INVALID_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_error = "no HIP device available";
        return LBFS_E_NO_DEVICE;
    }
    lbfs_ctx *ctx = new lbfs_ctx();
    ctx->params = *params;
    if (hipGetDevice(&ctx->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "could not create a HIP stream";
        delete ctx;
        return LBFS_E_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->side2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join2, hipEventDisableTiming) != hipSuccess) {
        g_create_error = "could not create a HIP stream or event";
        lbfs_destroy(ctx);
        return LBFS_E_HIP;
    }
    int rc = fluid_setup(ctx, *mesh, *params);
    if (rc) {
        g_create_error = ctx->err;
        lbfs_destroy(ctx);
        return rc;
    }
    // the static staging of lbfs_get_membrane_energies_noalloc on this device
    const hipError_t se = energy_scratch_setup(ctx->device, &ctx->energy_scratch);
    if (se != hipSuccess) {
        g_create_error = std::string("hipGetSymbolAddress of the energy staging: ") + hipGetErrorString(se);
        lbfs_destroy(ctx);
        return LBFS_E_HIP;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        g_create_error = "stream synchronisation failed during set-up";
        lbfs_destroy(ctx);
        return LBFS_E_HIP;
    }
    *out = ctx;
    return LBFS_OK;
}

extern "C" int lbfs_destroy(lbfs_ctx *ctx)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->drop_graph();
    ctx->release_extensions();
    for (void *p : ctx->allocs) hipFree(p);
    ctx->allocs.clear();
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->side2) (void)hipStreamSynchronize(ctx->side2);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_fork2) (void)hipEventDestroy(ctx->ev_fork2);
    if (ctx->ev_join2) (void)hipEventDestroy(ctx->ev_join2);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->side2) (void)hipStreamDestroy(ctx->side2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LBFS_OK;
}

int lbfs_ctx::check_device_error()
{
    lbfs_ctx *ctx = this;
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, dev_err, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (h & ERR_FIX_IBM)
        return fail(LBFS_E_NONFINITE, "IBM spreading: a node force term is not finite or out of the fixed-point "
                                      "range (|F w| >= 4096)");
    if (h & ERR_FIX_MEM)
        return fail(LBFS_E_NONFINITE, "membrane sums: a volume, area or curvature term is not finite or out of "
                                      "the fixed-point range");
    if (h & ERR_FIX_ADJ) return fail(LBFS_E_NONFINITE, "adjoint interpolation: a term is not finite");
    return LBFS_OK;
}

extern "C" int lbfs_sync(lbfs_ctx *ctx)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ctx->check_device_error();
}

extern "C" int lbfs_set_state(lbfs_ctx *ctx, const double *rho, const double *u, const double *v, const double *w)
{
    if (!ctx || !rho || !u || !v || !w) return LBFS_E_INVALID_ARG;
    const int n = ctx->lay.n_cells;
    std::vector<float4> h(n);
    for (int i = 0; i < n; i++) {
        const int c = ctx->lay.perm[i];
        if (!std::isfinite(rho[c]) || !std::isfinite(u[c]) || !std::isfinite(v[c]) || !std::isfinite(w[c]))
            return ctx->fail(LBFS_E_NONFINITE, "non-finite state at cell %d", c);
        h[i] = make_float4((float)(rho[c] - ctx->params.rho0), (float)u[c], (float)v[c], (float)w[c]);
    }
    HIPCHK(hipMemcpyAsync(ctx->st[ctx->cur], h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}

extern "C" int lbfs_get_state(lbfs_ctx *ctx, double *rho, double *u, double *v, double *w)
{
    if (!ctx || !rho || !u || !v || !w) return LBFS_E_INVALID_ARG;
    const int n = ctx->lay.n_cells;
    std::vector<float4> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->st[ctx->cur], sizeof(float4) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; i++) {
        const int c = ctx->lay.perm[i];
        rho[c] = ctx->params.rho0 + (double)h[i].x;
        u[c] = h[i].y; v[c] = h[i].z; w[c] = h[i].w;
    }
    return LBFS_OK;
}

void lbfs_ctx::drop_graph()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
}

// capture two consecutive steps (state ping-pong back to buffer 0) into a graph
static int capture_graph(lbfs_ctx *ctx)
{
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = ctx->run_stages(false);
    if (!rc) rc = ctx->run_stages(false);
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return ctx->fail(LBFS_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        ctx->graph_exec = nullptr;
        return ctx->fail(LBFS_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return LBFS_OK;
}

int lbfs_ctx::enqueue_steps(int nsteps)
{
    lbfs_ctx *ctx = this;
    int s = 0;
    if (use_graph && nsteps >= 2) {
        if (cur != 0) {
            int rc = enqueue_step();
            if (rc) return rc;
            s++;
        }
        if (!graph_exec) {
            const int saved = cur;
            int rc = capture_graph(this);
            cur = saved;  // capture only recorded the launches
            if (rc) return rc;
        }
        // graph pairs, falling back to single eager steps where the internal
        // viscosity field is due
        while (s + 2 <= nsteps) {
            if (visc_due(t_step) || visc_due(t_step + 1)) {
                int rc = enqueue_step();
                if (rc) return rc;
                s++;
                if (cur != 0) {
                    rc = enqueue_step();
                    if (rc) return rc;
                    s++;
                }
                continue;
            }
            HIPCHK(hipGraphLaunch(graph_exec, stream));
            t_step += 2;
            s += 2;
        }
    }
    for (; s < nsteps; s++) {
        int rc = enqueue_step();
        if (rc) return rc;
    }
    return LBFS_OK;
}

extern "C" int lbfs_step(lbfs_ctx *ctx, int nsteps)
{
    if (!ctx || nsteps < 0) return LBFS_E_INVALID_ARG;
    return ctx->enqueue_steps(nsteps);
}

extern "C" int lbfs_set_option(lbfs_ctx *ctx, int option, int value)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    if (option == LBFS_OPT_GRAPH) {
        ctx->use_graph = value != 0;
        ctx->drop_graph();
        return LBFS_OK;
    }
    if (option == LBFS_OPT_IBM_BINNED) {
        if (value < -1 || value > 2) return ctx->fail(LBFS_E_INVALID_ARG, "LBFS_OPT_IBM_BINNED takes -1, 0, 1 or 2");
        ctx->ibm_binned_option = value;  // applies to the next lbfs_add_membranes
        return LBFS_OK;
    }
    if (option == LBFS_OPT_FUSED_FLUID) {
        if (value < 0 || value > 1) return ctx->fail(LBFS_E_INVALID_ARG, "LBFS_OPT_FUSED_FLUID takes 0 or 1");
        if (value && !ctx->fused_ok)
            return ctx->fail(LBFS_E_UNSUPPORTED, "the tiles are too large for the fused fluid pass");
        ctx->fused = value != 0;
        ctx->drop_graph();
        return LBFS_OK;
    }
    if (option == LBFS_OPT_VISC_SEARCH) {
        if (value < 0 || value > 1) return ctx->fail(LBFS_E_INVALID_ARG, "LBFS_OPT_VISC_SEARCH takes 0 or 1");
        ctx->visc_search = value;
        if (ctx->visc_factor) ctx->visc_grid.on = value;
        return LBFS_OK;
    }
    if (option == LBFS_OPT_STEP_SCHEDULE) {
        if (value < 0 || value > 2) return ctx->fail(LBFS_E_INVALID_ARG, "LBFS_OPT_STEP_SCHEDULE takes 0, 1 or 2");
        ctx->schedule_option = value;
        ctx->drop_graph();
        return LBFS_OK;
    }
    return ctx->fail(LBFS_E_INVALID_ARG, "unknown option %d", option);
}

extern "C" int lbfs_stage_gradients(lbfs_ctx *ctx, double *grad_out)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    HIPCHK(launch_gradients(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->stream));
    if (!grad_out) return LBFS_OK;
    const int n = ctx->lay.n_cells;
    std::vector<float> h(12 * (size_t)n);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->grad, sizeof(float) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; i++) {
        const int c = ctx->lay.perm[i];
        for (int t = 0; t < 12; t++) grad_out[12 * (size_t)c + t] = h[12 * (size_t)i + t];
    }
    return LBFS_OK;
}

extern "C" int lbfs_stage_face_flux(lbfs_ctx *ctx, double *flux_out)
{
    if (!ctx || !flux_out) return LBFS_E_INVALID_ARG;
    const int ntf = (int)ctx->lay.face_ref.size();
    HIPCHK(launch_gradients(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->stream));
    HIPCHK(launch_face_flux_all(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->face_scratch, ntf, ctx->stream));
    std::vector<float4> h(ntf);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->face_scratch, sizeof(float4) * ntf, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the device flux leaves out the rest pressure rho0/3 n A (it cancels per
    // cell); it is added back here so the result is the reference face flux
    const double p0 = ctx->params.rho0 / 3.0;
    const double p0z = ctx->params.lattice == LBFS_D2Q9 ? 0.0 : p0;
    for (int f = 0; f < ctx->geo.n_faces; f++) {
        const float4 r = h[ctx->lay.face_first[f]];
        const bool empty = f >= ctx->geo.n_int && ctx->bc_vel_type[f - ctx->geo.n_int] == LBFS_BC_EMPTY;
        const double A = empty ? 0.0 : ctx->geo.area[f];
        const double *n = &ctx->geo.normal[3 * (size_t)f];
        flux_out[4 * (size_t)f] = r.x;
        flux_out[4 * (size_t)f + 1] = r.y + p0 * n[0] * A;
        flux_out[4 * (size_t)f + 2] = r.z + p0 * n[1] * A;
        flux_out[4 * (size_t)f + 3] = r.w + p0z * n[2] * A;
    }
    return LBFS_OK;
}

extern "C" int lbfs_stage_residual(lbfs_ctx *ctx, double *res_out)
{
    if (!ctx || !res_out) return LBFS_E_INVALID_ARG;
    const int n = ctx->lay.n_cells;
    HIPCHK(launch_gradients(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->stream));
    HIPCHK(launch_flux_update(ctx->args, ctx->st[ctx->cur], ctx->grad, nullptr, ctx->res, 0, ctx->stream));
    std::vector<float4> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->res, sizeof(float4) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // add back the rest-pressure closure term -rho0/3 sum_f n A / V (zero on closed cells)
    const Geometry &g = ctx->geo;
    const std::vector<double> &S = ctx->lay.closure;
    const double p0 = ctx->params.rho0 / 3.0;
    const double p0z = ctx->params.lattice == LBFS_D2Q9 ? 0.0 : p0;
    for (int i = 0; i < n; i++) {
        const int c = ctx->lay.perm[i];
        res_out[4 * (size_t)c] = h[i].x;
        res_out[4 * (size_t)c + 1] = h[i].y - p0 * S[3 * (size_t)c] / g.vol[c];
        res_out[4 * (size_t)c + 2] = h[i].z - p0 * S[3 * (size_t)c + 1] / g.vol[c];
        res_out[4 * (size_t)c + 3] = h[i].w - p0z * S[3 * (size_t)c + 2] / g.vol[c];
    }
    return LBFS_OK;
}

extern "C" int lbfs_get_cell_force(lbfs_ctx *ctx, double *f_out)
{
    if (!ctx || !f_out) return LBFS_E_INVALID_ARG;
    const int n = ctx->lay.n_cells;
    std::vector<float> cf(3 * (size_t)n), sf(3 * (size_t)n);
    HIPCHK(hipMemcpyAsync(cf.data(), ctx->cell_force, sizeof(float) * cf.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(sf.data(), ctx->static_force, sizeof(float) * sf.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // body force plus the spread IBM force (the closure term is not a force of the reference)
    for (int i = 0; i < n; i++) {
        const int c = ctx->lay.perm[i];
        const double *bf = cell_body_force(ctx->blocks, c, ctx->params.body_force);
        for (int t = 0; t < 3; t++)
            f_out[3 * (size_t)c + t] = bf[t] + ((double)cf[3 * (size_t)i + t] - (double)sf[3 * (size_t)i + t]);
    }
    return LBFS_OK;
}

// FP64 partial sums of k_diag_partials over the state `st`, the flux residual
// in ctx->res and the momentum force density `force`, closed out on the host
// in block order (residuals::l2_norm_rms_moukallad, residuals.cpp:114-132)
static int finish_diagnostics(lbfs_ctx *ctx, const float4 *st, const float *force, double ref_rho, double ref_u,
                              lbfs_diag *out)
{
    const int n = ctx->lay.n_cells, nb = (n + 255) / 256;
    HIPCHK(launch_diag_partials(n, ctx->args.rho0, st, ctx->res, force, ctx->args.inv_vol, ctx->diag_part,
                                ctx->stream));
    std::vector<double> h(6 * (size_t)nb);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->diag_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s[5] = {0, 0, 0, 0, 0}, mx = 0.0;
    for (int b = 0; b < nb; b++) {
        for (int t = 0; t < 5; t++) s[t] += h[6 * (size_t)b + t];
        mx = std::fmax(mx, h[6 * (size_t)b + 5]);
    }
    out->rho_rms = std::sqrt(s[0] / n) / ref_rho;
    out->u_rms = std::sqrt(s[1] / n) / (ref_rho * ref_u);
    out->v_rms = std::sqrt(s[2] / n) / (ref_rho * ref_u);
    out->w_rms = std::sqrt(s[3] / n) / (ref_rho * ref_u);
    out->total_mass = s[4];
    out->max_speed = mx;
    out->finite = std::isfinite(mx) && std::isfinite(s[0]) && std::isfinite(s[1]) ? 1 : 0;
    int rc = ctx->check_device_error();
    if (rc) {
        out->finite = 0;
        return rc;
    }
    return out->finite ? LBFS_OK : ctx->fail(LBFS_E_NONFINITE, "non-finite fluid state");
}

extern "C" int lbfs_reduce_diagnostics(lbfs_ctx *ctx, double ref_rho, double ref_u, lbfs_diag *out)
{
    if (!ctx || !out || !(ref_rho > 0.0) || !(ref_u > 0.0)) return LBFS_E_INVALID_ARG;
    // residual of the step that starts from the current state, with the
    // static force (body force) added to the momentum rows
    HIPCHK(launch_gradients(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->stream));
    HIPCHK(launch_flux_update(ctx->args, ctx->st[ctx->cur], ctx->grad, nullptr, ctx->res, 0, ctx->stream));
    return finish_diagnostics(ctx, ctx->st[ctx->cur], ctx->static_force, ref_rho, ref_u, out);
}

extern "C" int lbfs_step_residual(lbfs_ctx *ctx, double ref_rho, double ref_u, lbfs_diag *out)
{
    if (!ctx || !out || !(ref_rho > 0.0) || !(ref_u > 0.0)) return LBFS_E_INVALID_ARG;
    // one eager step whose face pass also writes the flux residual; the
    // step's force density is the static force plus, with membranes, the
    // force the step spread (the cell force output of the apply pass)
    ctx->res_track = true;
    int rc = ctx->enqueue_step();
    ctx->res_track = false;
    if (rc) return rc;
    const float *force = ctx->have_membranes ? ctx->cell_force : ctx->static_force;
    return finish_diagnostics(ctx, ctx->st[ctx->cur], force, ref_rho, ref_u, out);
}

extern "C" int lbfs_run_to_tolerance(lbfs_ctx *ctx, int max_steps, double tolerance, int check_every, double ref_rho,
                                     double ref_u, int *steps_done, int *converged, lbfs_diag *out)
{
    if (!ctx || max_steps < 0 || check_every < 1 || !(ref_rho > 0.0) || !(ref_u > 0.0) || !steps_done ||
        !converged || !out)
        return LBFS_E_INVALID_ARG;
    *steps_done = 0;
    *converged = 0;
    memset(out, 0, sizeof(*out));
    int done = 0;
    while (done < max_steps) {
        const int chunk = std::min(check_every, max_steps - done);
        int rc = chunk > 1 ? ctx->enqueue_steps(chunk - 1) : LBFS_OK;
        if (!rc) rc = lbfs_step_residual(ctx, ref_rho, ref_u, out);
        if (rc) {
            *steps_done = done;
            return rc;
        }
        done += chunk;
        // residuals::max_error (residuals.cpp:158-166): the momentum rows only
        if (std::fmax(std::fmax(out->w_rms, out->u_rms), out->v_rms) < tolerance) {
            *converged = 1;
            break;
        }
    }
    *steps_done = done;
    return LBFS_OK;
}

// ---------------------------------------------------------------------------
// host-side mesh helpers
// ---------------------------------------------------------------------------
extern "C" int lbfs_mesh_hex_counts(int nx, int ny, int nz, const int *periodic, int *n_points, int *n_faces,
                                    int *n_internal)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !periodic || !n_points || !n_faces || !n_internal) return LBFS_E_INVALID_ARG;
    const HexCounts c = hex_counts(nx, ny, nz, periodic);
    *n_points = c.n_points; *n_faces = c.n_faces; *n_internal = c.n_internal;
    return LBFS_OK;
}

extern "C" int lbfs_mesh_hex_generate(int nx, int ny, int nz, const double *lengths, const double *origin,
                                      const int *periodic, double jitter, uint64_t seed, double *points,
                                      int *face_vertices, int *owner, int *neighbour, double *face_shift, int *patch)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !lengths || !origin || !periodic || !points || !face_vertices || !owner ||
        !neighbour || !face_shift || !patch || jitter < 0.0 || jitter >= 0.5)
        return LBFS_E_INVALID_ARG;
    hex_generate(nx, ny, nz, lengths, origin, periodic, jitter, seed, points, face_vertices, owner, neighbour, face_shift,
                 patch);
    return LBFS_OK;
}

extern "C" int lbfs_mesh_geometry(const lbfs_mesh_desc *mesh, double visc, double gamma, double fneq_min, int flags,
                                  double *area, double *normal, double *face_centroid, double *face_dt,
                                  double *centroid, double *volume)
{
    if (!mesh || !area || !normal || !face_centroid || !face_dt || !centroid || !volume) return LBFS_E_INVALID_ARG;
    Geometry g;
    std::string err;
    int rc = compute_geometry(*mesh, visc, gamma, fneq_min, (flags & LBFS_FLAG_CLOSED_FACES) != 0, g, err);
    if (rc) { g_create_error = err; return rc; }
    std::copy(g.area.begin(), g.area.end(), area);
    std::copy(g.normal.begin(), g.normal.end(), normal);
    std::copy(g.fc.begin(), g.fc.end(), face_centroid);
    std::copy(g.fdt.begin(), g.fdt.end(), face_dt);
    std::copy(g.cc.begin(), g.cc.end(), centroid);
    std::copy(g.vol.begin(), g.vol.end(), volume);
    return LBFS_OK;
}
