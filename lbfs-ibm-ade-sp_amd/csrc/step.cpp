// Time-step orchestration behind lbfs_step: the loop body of
// Solver_gpu.cu:749-1246 as an ordered list of kernel stages, plus the
// event-based measurement entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "context.hpp"

using namespace lbfs;

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return ctx->fail(LBFS_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

std::vector<lbfs_ctx::Stage> lbfs_ctx::step_stages(bool with_visc, bool fused_fluid)
{
    std::vector<Stage> s;
    lbfs_ctx *ctx = this;
    const double n = lay.n_cells, nf = geo.n_faces;
    if (have_membranes) {
        // IBM interpolation from the state at the start of the step and node
        // advection, then membrane forces at the advected nodes and spreading
        // (Solver_gpu.cu:808-915).  Interpolation per node: X read and written
        // 32, U 16, node indices 8, stencil base read and written 32, the kept
        // weight row read and the new one written 512, list slot 8; per cell
        // the stencils reach (about 8 per node, each cell reached by about 8
        // nodes) state 16 and centroid 16
        const double nn = mem.n_nodes, ne = mem.n_edges, nt = mem.n_tri;
        const double touched = std::min(n, 8.0 * nn);
        s.push_back({"ibm_interp", [ctx]() -> int {
                         HIPCHK(launch_ibm_interp(ctx->mem, ctx->ibm, ctx->st[ctx->cur], ctx->stream));
                         return LBFS_OK;
                     },
                     608.0 * nn + 32.0 * touched});
        // volume and area sums: triangle 16 + nodes 48 per triangle; edge
        // forces: nodes 96 + indices 24 + constants 24 + state 16 + forces
        // written 64 per edge
        s.push_back({"mem_volume", [ctx]() -> int {
                         HIPCHK(launch_mem_volume(ctx->mem, ctx->stream));
                         return LBFS_OK;
                     },
                     64.0 * nt});
        s.push_back({"mem_edges", [ctx]() -> int {
                         HIPCHK(launch_mem_edges(ctx->mem, ctx->stream));
                         return LBFS_OK;
                     },
                     224.0 * ne});
        // node forces: per edge its four force slots 64 and their incidence
        // ids 16; per node the incidence range 4 and the force written 16 (in
        // the spreading when it gathers them itself)
        const double node_force_bytes = 80.0 * ne + 20.0 * nn;
        const bool gather_forces = spread_gathers_forces(ibm);
        if (!gather_forces)
            s.push_back({"mem_nodes", [ctx]() -> int {
                             HIPCHK(launch_mem_nodes(ctx->mem, ctx->stream));
                             return LBFS_OK;
                         },
                         node_force_bytes});
        if (with_visc)
            s.push_back({"internal_visc", [ctx]() -> int {
                             HIPCHK(launch_internal_visc(ctx->mem, ctx->ibm, ctx->visc_grid, ctx->delta_h, ctx->visc_factor,
                                                          ctx->stream));
                             return LBFS_OK;
                         },
                         8.0 * n});
        // spreading per node: weight row 256, force 16, stencil base 16; per
        // touched cell the force density written 12 (tile gathering) or the
        // fixed-point sums read and written 48 (per-node and bin scattering)
        const double cell_bytes = ibm.binned == IBM_SPREAD_GATHER ? 12.0 : 48.0;
        s.push_back({"ibm_spread", [ctx]() -> int {
                         HIPCHK(launch_ibm_spread(ctx->mem, ctx->ibm, ctx->stream));
                         return LBFS_OK;
                     },
                     288.0 * nn + cell_bytes * touched + (gather_forces ? node_force_bytes : 0.0)});
    }
    const bool fused = fused_fluid && fused_ok && args.ts_mode != LBFS_TS_MIN;
    if (args.ts_mode == LBFS_TS_MIN)
        s.push_back({"dt_min", [ctx]() -> int {
                         HIPCHK(launch_dt_min(ctx->args, ctx->st[ctx->cur], ctx->stream));
                         return LBFS_OK;
                     },
                     24.0 * n});
    // pass G: state 16 + stencil 24 + RHS 72 + LS matrix 36 + gradients written 48 (B per cell)
    if (!fused)
        s.push_back({"gradients", [ctx]() -> int {
                         HIPCHK(launch_gradients(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->stream));
                         return LBFS_OK;
                     },
                     196.0 * n});
    if (have_membranes) {
        // The face pass updates the cell with the static force (it does not
        // need the spread force, so it runs beside the coupling chain); the
        // apply pass then adds the spread force density in the tiles the
        // spreading wrote into.  Face pass as below; apply: force density 12
        // + static force 12 + state read and written 32 + force written 12
        // per touched cell (about 8 per node)
        const double touched = std::min(n, 8.0 * (double)mem.n_nodes);
        if (fused)
            // the fused pass: both passes' bytes less the gradients written
            // and read back (they stay in LDS) and the second read of the state
            s.push_back({"fluid", [ctx]() -> int {
                             HIPCHK(launch_fluid_fused(ctx->args, ctx->st[ctx->cur], ctx->st[ctx->cur ^ 1],
                                                       ctx->res_out(), 1, ctx->stream));
                             return LBFS_OK;
                         },
                         192.0 * n + 40.0 * nf});
        else
            s.push_back({"flux_update", [ctx]() -> int {
                             HIPCHK(launch_flux_update(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->st[ctx->cur ^ 1],
                                                       ctx->res_out(), 1, ctx->stream));
                             return LBFS_OK;
                         },
                         108.0 * n + 40.0 * nf});
        s.push_back({"ibm_apply", [ctx]() -> int {
                         HIPCHK(launch_ibm_apply(ctx->ibm, ctx->apply_args(ctx->st[ctx->cur ^ 1], ctx->cell_force, 0),
                                                 ctx->stream));
                         ctx->cur ^= 1;
                         return LBFS_OK;
                     },
                     68.0 * touched});
        return s;
    }
    // pass F: state 16 + gradients 48 + volume/centroid 16 + force 12 + state written 16 per cell,
    // plus owner/nb 8, normal 12, face centroid 12, area 4, dt 4 per face; the
    // fused pass moves G + F less the gradients written and read back (they
    // stay in LDS) and the second read of the state: 192 per cell
    if (fused)
        s.push_back({"fluid", [ctx]() -> int {
                         HIPCHK(launch_fluid_fused(ctx->args, ctx->st[ctx->cur], ctx->st[ctx->cur ^ 1], ctx->res_out(),
                                                   1, ctx->stream));
                         ctx->cur ^= 1;
                         return LBFS_OK;
                     },
                     192.0 * n + 40.0 * nf});
    else
        s.push_back({"flux_update", [ctx]() -> int {
                         HIPCHK(launch_flux_update(ctx->args, ctx->st[ctx->cur], ctx->grad, ctx->st[ctx->cur ^ 1],
                                                   ctx->res_out(), 1, ctx->stream));
                         ctx->cur ^= 1;
                         return LBFS_OK;
                     },
                     108.0 * n + 40.0 * nf});
    return s;
}

int lbfs_ctx::run_stages(bool with_visc)
{
    lbfs_ctx *ctx = this;
    const bool fz = fused && fused_ok && args.ts_mode != LBFS_TS_MIN;
    if (!have_membranes || !side || with_visc) {
        // (the internal viscosity field feeds the face pass: in the steps
        // that rebuild it, everything runs in order on one stream)
        for (auto &st : step_stages(with_visc, fz)) {
            int rc = st.run();
            if (rc) return rc;
        }
        return LBFS_OK;
    }
    // With membranes the gradients and the face pass (which read only the
    // state at the start of the step) run beside the coupling chain (a
    // latency-bound sequence of small kernels); the apply pass waits for the
    // face pass and for spreading.  Same kernels, same inputs: the results are
    // those of the sequential order.
    // Schedule 0 (the default): the fluid passes on the main stream from the
    // start of the step, the whole coupling chain on a second.  Schedule 1:
    // the interpolation first and alone on the main stream (beside the fluid
    // passes it runs several times slower), then the fluid passes on the main
    // stream and the membrane forces and spreading on a second.  Schedule 2:
    // the fluid passes on a second stream from the start, the membrane forces
    // on a third after the interpolation.  (The fluid passes are enqueued
    // first: a captured graph launches its branches in that order.)  In a
    // graph each wait across streams costs 10-20 us; schedules 0 and 2 measure
    // the same, schedule 1 some 2 us more per step.
    if (schedule_option == 0 || schedule_option == 1) {
        const bool interp_first = schedule_option == 1;
        if (interp_first) HIPCHK(launch_ibm_interp(mem, ibm, st[cur], stream));
        HIPCHK(hipEventRecord(ev_fork, stream));
        if (fz) {
            HIPCHK(launch_fluid_fused(args, st[cur], st[cur ^ 1], res_out(), 1, stream));
        } else {
            HIPCHK(launch_gradients(args, st[cur], grad, stream));
            HIPCHK(launch_flux_update(args, st[cur], grad, st[cur ^ 1], res_out(), 1, stream));
        }
        HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
        if (!interp_first) HIPCHK(launch_ibm_interp(mem, ibm, st[cur], side));
        HIPCHK(launch_mem_volume(mem, side));
        HIPCHK(launch_mem_edges(mem, side));
        if (!spread_gathers_forces(ibm)) HIPCHK(launch_mem_nodes(mem, side));
        HIPCHK(launch_ibm_spread(mem, ibm, side));
        HIPCHK(hipEventRecord(ev_join, side));
        HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
        HIPCHK(launch_ibm_apply(ibm, apply_args(st[cur ^ 1], cell_force, 0), stream));
        cur ^= 1;
        return LBFS_OK;
    }
    HIPCHK(hipEventRecord(ev_fork, stream));
    HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
    if (fz) {
        HIPCHK(launch_fluid_fused(args, st[cur], st[cur ^ 1], res_out(), 1, side));
    } else {
        HIPCHK(launch_gradients(args, st[cur], grad, side));
        HIPCHK(launch_flux_update(args, st[cur], grad, st[cur ^ 1], res_out(), 1, side));
    }
    HIPCHK(hipEventRecord(ev_join, side));
    HIPCHK(launch_ibm_interp(mem, ibm, st[cur], stream));
    HIPCHK(hipEventRecord(ev_fork2, stream));
    HIPCHK(hipStreamWaitEvent(side2, ev_fork2, 0));
    HIPCHK(launch_mem_volume(mem, side2));
    HIPCHK(launch_mem_edges(mem, side2));
    if (!spread_gathers_forces(ibm)) HIPCHK(launch_mem_nodes(mem, side2));
    HIPCHK(hipEventRecord(ev_join2, side2));
    HIPCHK(hipStreamWaitEvent(stream, ev_join2, 0));
    HIPCHK(launch_ibm_spread(mem, ibm, stream));
    HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
    HIPCHK(launch_ibm_apply(ibm, apply_args(st[cur ^ 1], cell_force, 0), stream));
    cur ^= 1;
    return LBFS_OK;
}

int lbfs_ctx::enqueue_step()
{
    int rc = run_stages(visc_due(t_step));
    if (!rc) t_step++;
    return rc;
}

// HIP events destroyed on every exit path
struct EventSet {
    std::vector<hipEvent_t> ev;
    ~EventSet()
    {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    hipError_t create(size_t n)
    {
        ev.assign(n, nullptr);
        for (hipEvent_t &e : ev) {
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
};

extern "C" int lbfs_time_steps(lbfs_ctx *ctx, int nsteps, double *ms_total)
{
    if (!ctx || nsteps <= 0 || !ms_total) return LBFS_E_INVALID_ARG;
    EventSet E;
    HIPCHK(E.create(2));
    HIPCHK(hipEventRecord(E.ev[0], ctx->stream));
    int rc = ctx->enqueue_steps(nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E.ev[1], ctx->stream));
    HIPCHK(hipEventSynchronize(E.ev[1]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, E.ev[0], E.ev[1]));
    *ms_total = ms;
    return LBFS_OK;
}

extern "C" int lbfs_time_kernels(lbfs_ctx *ctx, int nsteps, int max_count, const char **names, double *ms, int *count)
{
    if (!ctx || nsteps <= 0 || !ms || !count || max_count <= 0) return LBFS_E_INVALID_ARG;
    // whole steps, one stage at a time with an event between stages (the
    // stage list of a step that does not rebuild the internal viscosity
    // field; such steps are refused so the schedule stays that of lbfs_step)
    for (int s = 0; s < nsteps; s++)
        if (ctx->visc_due(ctx->t_step + s))
            return ctx->fail(LBFS_E_STATE, "a timed step would rebuild the internal viscosity field");
    std::vector<lbfs_ctx::Stage> stages = ctx->step_stages(false, ctx->fused);
    const int ns = (int)stages.size();
    EventSet E;
    HIPCHK(E.create((size_t)nsteps * (ns + 1)));
    for (int s = 0; s < nsteps; s++) {
        for (int k = 0; k < ns; k++) {
            HIPCHK(hipEventRecord(E.ev[(size_t)s * (ns + 1) + k], ctx->stream));
            int rc = stages[k].run();
            if (rc) return rc;
        }
        HIPCHK(hipEventRecord(E.ev[(size_t)s * (ns + 1) + ns], ctx->stream));
        ctx->t_step++;  // a whole step was taken
    }
    HIPCHK(hipEventSynchronize(E.ev.back()));
    std::vector<double> acc(ns, 0.0);
    for (int s = 0; s < nsteps; s++)
        for (int k = 0; k < ns; k++) {
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, E.ev[(size_t)s * (ns + 1) + k], E.ev[(size_t)s * (ns + 1) + k + 1]));
            acc[k] += t;
        }
    *count = ns < max_count ? ns : max_count;
    for (int k = 0; k < *count; k++) {
        ms[k] = acc[k] / nsteps;
        if (names) names[k] = stages[k].name;
    }
    return LBFS_OK;
}

extern "C" int lbfs_time_stage(lbfs_ctx *ctx, const char *name, int n, double *ms_avg)
{
    if (!ctx || !name || n <= 0 || !ms_avg) return LBFS_E_INVALID_ARG;
    // the fluid passes read the state and write scratch or the other state
    // buffer, so back-to-back launches leave the state as it was
    if (std::strcmp(name, "gradients") != 0 && std::strcmp(name, "flux_update") != 0 && std::strcmp(name, "fluid") != 0)
        return ctx->fail(LBFS_E_INVALID_ARG,
                         "lbfs_time_stage times the fluid passes only (gradients, flux_update, fluid)");
    std::vector<lbfs_ctx::Stage> stages = ctx->step_stages(false, std::strcmp(name, "fluid") == 0);
    const lbfs_ctx::Stage *st = nullptr;
    for (auto &s : stages)
        if (std::strcmp(s.name, name) == 0) st = &s;
    if (!st) return ctx->fail(LBFS_E_INVALID_ARG, "no stage %s", name);
    EventSet E;
    HIPCHK(E.create(2));
    const int cur = ctx->cur;
    HIPCHK(hipEventRecord(E.ev[0], ctx->stream));
    for (int k = 0; k < n; k++) {
        int rc = st->run();
        ctx->cur = cur;
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(E.ev[1], ctx->stream));
    HIPCHK(hipEventSynchronize(E.ev[1]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, E.ev[0], E.ev[1]));
    *ms_avg = ms / n;
    return LBFS_OK;
}

extern "C" int lbfs_model_bytes(lbfs_ctx *ctx, int max_count, double *bytes, int *count)
{
    if (!ctx || !bytes || !count || max_count <= 0) return LBFS_E_INVALID_ARG;
    std::vector<lbfs_ctx::Stage> stages = ctx->step_stages(false, ctx->fused);
    *count = (int)stages.size() < max_count ? (int)stages.size() : max_count;
    for (int k = 0; k < *count; k++) bytes[k] = stages[k].model_bytes;
    return LBFS_OK;
}
