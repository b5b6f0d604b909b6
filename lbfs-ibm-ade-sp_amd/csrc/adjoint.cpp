// Discrete adjoint of the coupled time step behind lbfs_adjoint_*: a forward
// run that keeps the state every K steps, then a reverse sweep over segments
// of K steps.  Each segment is replayed from its checkpoint, keeping what the
// reverse step needs of every step (start state, gradients, residual, cell
// force, nodes before and after advection, node velocities and forces,
// membrane globals), written straight into per-step records; then the
// segment's steps are reversed by the adjoint kernels, last first.  The
// replays run the very kernels of lbfs_step on the same inputs, so the
// replayed trajectory is the recorded one bit for bit.
//
// Also the two setters the parameter gradients refer to: the uniform body
// force and the membrane moduli.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.hpp"
#include "context.hpp"

using namespace lbfs;

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return ctx->fail(LBFS_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace lbfs {

// the device state a step starts from
struct Snapshot {
    float4 *st = nullptr, *X = nullptr, *edge_state = nullptr;
    float *wnode = nullptr;
    int4 *nbase = nullptr;
    double *glob = nullptr;
    long long t = 0;
};

// what the reverse of one step reads: the forward quantities of that step
struct StepRec {
    float4 *st = nullptr;       // state at the start (aliases the checkpoint for a segment's first step)
    float4 *grad = nullptr;     // gradients of st
    float4 *res = nullptr;      // flux residual (before forces)
    float4 *mom = nullptr;      // face moments S0, S1 per tile face
    float *cf = nullptr;        // cell force of the step [3 n]
    float4 *X = nullptr, *Xn1 = nullptr, *U = nullptr, *F = nullptr;  // X^n, X^n+1, U^n, F^n
    double *glob_prev = nullptr, *glob = nullptr;  // globals at the start, after the step's sums
    // X and glob_prev alias the previous record's Xn1 and glob (the checkpoint's for the first)
};

struct AdjointState {
    std::vector<void *> allocs;
    int n_steps = 0, every = 0;
    int max_tile_cells = 0;
    bool recorded = false;
    bool last_in_rec = false;     // the records hold the last segment (written by the forward pass)
    std::vector<Snapshot> ckpt;   // start states of steps 0, K, 2K, ...
    std::vector<StepRec> rec;     // the steps of one segment
    std::vector<StepRec> rec2;    // a second segment's steps: the backward pass replays the segment
                                  // before into one set while it reverses the other
    Snapshot start, end;          // before and after the recorded run
    hipStream_t rs = nullptr;     // stream of the overlapped replays: one, so that the process's
                                  // streams stay within its hardware queues (four by default) and
                                  // no replay kernel queues behind a reverse kernel
    hipEvent_t rev[7] = {};       // replay fork and joins, records ready (2), records reversed (2)
    // fluid adjoint
    float4 *lam_st[2] = {nullptr, nullptr};
    float4 *mu = nullptr, *lcf = nullptr, *ls = nullptr, *ls_acc = nullptr, *lam_g = nullptr;
    double *fpart = nullptr, *gforce = nullptr;
    int n_fpart = 0;
    float4 *rev_w = nullptr;
    // coupling adjoint
    float4 *lam_X = nullptr, *lam_U = nullptr, *lam_F = nullptr, *grec = nullptr;
    float *erec = nullptr, *mrec = nullptr;
    double *gmem = nullptr, *epart = nullptr;
    int *echunk = nullptr, *echunk_start = nullptr;
    int n_echunks = 0;
    unsigned int *umax = nullptr;
    unsigned long long *acc = nullptr;  // fixed-point cell sums of the coupling adjoint
};

void adjoint_release(lbfs_ctx *ctx)
{
    if (!ctx->adj) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    AdjointState &A = *ctx->adj;
    if (A.rs) {
        (void)hipStreamSynchronize(A.rs);
        (void)hipStreamDestroy(A.rs);
    }
    for (hipEvent_t &e : A.rev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : A.allocs) (void)hipFree(p);
    delete ctx->adj;
    ctx->adj = nullptr;
}

}  // namespace lbfs

namespace {

template <class T>
int adj_alloc(lbfs_ctx *ctx, T **dst, size_t n)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, (n > 0 ? n : 1) * sizeof(T)));
    ctx->adj->allocs.push_back(p);
    HIPCHK(hipMemsetAsync(p, 0, (n > 0 ? n : 1) * sizeof(T), ctx->stream));  // ordered before the kernels
    *dst = (T *)p;
    return LBFS_OK;
}

int snap_alloc(lbfs_ctx *ctx, Snapshot &s)
{
    int rc;
    if ((rc = adj_alloc(ctx, &s.st, (size_t)ctx->lay.n_cells))) return rc;
    if (!ctx->have_membranes) return LBFS_OK;
    const MemArgs &m = ctx->mem;
    if ((rc = adj_alloc(ctx, &s.X, (size_t)m.n_nodes))) return rc;
    if ((rc = adj_alloc(ctx, &s.edge_state, (size_t)m.n_edges))) return rc;
    if ((rc = adj_alloc(ctx, &s.wnode, 64 * (size_t)m.n_nodes))) return rc;
    if ((rc = adj_alloc(ctx, &s.nbase, (size_t)m.n_nodes))) return rc;
    if ((rc = adj_alloc(ctx, &s.glob, 8 * (size_t)m.n_mem))) return rc;
    return LBFS_OK;
}

// save: live state -> s; otherwise s -> live state (state into buffer 0);
// on stream `q` (the context's stream if null)
int snap_copy(lbfs_ctx *ctx, Snapshot &s, bool save, hipStream_t q = nullptr)
{
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    hipStream_t const stream = q ? q : ctx->stream;
    const size_t n = (size_t)ctx->lay.n_cells;
    if (save) {
        HIPCHK(hipMemcpyAsync(s.st, ctx->st[ctx->cur], sizeof(float4) * n, k, stream));
        s.t = ctx->t_step;
    } else {
        HIPCHK(hipMemcpyAsync(ctx->st[0], s.st, sizeof(float4) * n, k, stream));
        ctx->cur = 0;
        ctx->t_step = s.t;
    }
    if (!ctx->have_membranes) return LBFS_OK;
    MemArgs &m = ctx->mem;
    IbmArgs &g = ctx->ibm;
    const size_t nn = (size_t)m.n_nodes, ne = (size_t)m.n_edges, nm = (size_t)m.n_mem;
    if (save) {
        HIPCHK(hipMemcpyAsync(s.X, m.X, sizeof(float4) * nn, k, stream));
        HIPCHK(hipMemcpyAsync(s.edge_state, m.edge_state, sizeof(float4) * ne, k, stream));
        HIPCHK(hipMemcpyAsync(s.wnode, g.wnode, sizeof(float) * 64 * nn, k, stream));
        HIPCHK(hipMemcpyAsync(s.nbase, g.nbase, sizeof(int4) * nn, k, stream));
        HIPCHK(hipMemcpyAsync(s.glob, m.glob, sizeof(double) * 8 * nm, k, stream));
    } else {
        HIPCHK(hipMemcpyAsync(m.X, s.X, sizeof(float4) * nn, k, stream));
        HIPCHK(hipMemcpyAsync(m.edge_state, s.edge_state, sizeof(float4) * ne, k, stream));
        HIPCHK(hipMemcpyAsync(g.wnode, s.wnode, sizeof(float) * 64 * nn, k, stream));
        HIPCHK(hipMemcpyAsync(g.nbase, s.nbase, sizeof(int4) * nn, k, stream));
        HIPCHK(hipMemcpyAsync(m.glob, s.glob, sizeof(double) * 8 * nm, k, stream));
    }
    return LBFS_OK;
}

int save(lbfs_ctx *ctx, Snapshot &s) { return snap_copy(ctx, s, true); }
int restore(lbfs_ctx *ctx, Snapshot &s, hipStream_t q = nullptr) { return snap_copy(ctx, s, false, q); }

// the conditions under which the adjoint is the derivative of the step
int adjoint_supported(lbfs_ctx *ctx)
{
    if (ctx->args.ts_mode != 0) return ctx->fail(LBFS_E_UNSUPPORTED, "the adjoint needs constant time steps");
    if (ctx->args.flags & LBFS_FLAG_UPWIND_BAND)
        return ctx->fail(LBFS_E_UNSUPPORTED, "the adjoint follows strict upwinding (no tie band)");
    for (int t = 0; t < ctx->lay.n_tiles; t++)
        if (ctx->lay.tile_cell_start[t + 1] - ctx->lay.tile_cell_start[t] > 1024)
            return ctx->fail(LBFS_E_UNSUPPORTED, "the adjoint takes tiles of at most 1024 cells");
    if (ctx->have_membranes) {
        if (ctx->ibm.kernel != LBFS_IBM_COS4)
            return ctx->fail(LBFS_E_UNSUPPORTED, "the adjoint supports the 4-point cosine IBM kernel");
        if (ctx->visc_factor)
            return ctx->fail(LBFS_E_UNSUPPORTED, "the adjoint needs internal viscosity ratios of 1");
    }
    return LBFS_OK;
}

// the step records of one segment
int alloc_records(lbfs_ctx *ctx, std::vector<StepRec> &recs, int every)
{
    const int n = ctx->lay.n_cells;
    int rc;
    recs.resize(every);
    for (int k = 0; k < every; k++) {
        StepRec &r = recs[k];
        if (k > 0 && (rc = adj_alloc(ctx, &r.st, (size_t)n))) return rc;
        if ((rc = adj_alloc(ctx, &r.grad, 3 * (size_t)n))) return rc;
        if ((rc = adj_alloc(ctx, &r.res, (size_t)n))) return rc;
        if ((rc = adj_alloc(ctx, &r.mom, ctx->lay.fA.size() / 4))) return rc;
        if (ctx->have_membranes && (rc = adj_alloc(ctx, &r.cf, 3 * (size_t)n))) return rc;
        if (ctx->have_membranes) {
            const size_t nn = ctx->mem.n_nodes, nm = ctx->mem.n_mem;
            if ((rc = adj_alloc(ctx, &r.Xn1, nn))) return rc;
            if ((rc = adj_alloc(ctx, &r.U, nn))) return rc;
            if ((rc = adj_alloc(ctx, &r.F, nn))) return rc;
            if ((rc = adj_alloc(ctx, &r.glob, 8 * nm))) return rc;
        }
    }
    return LBFS_OK;
}

int adjoint_setup(lbfs_ctx *ctx, int n_steps, int every)
{
    if (ctx->adj && ctx->adj->n_steps == n_steps && ctx->adj->every == every) {
        ctx->adj->recorded = false;  // same shape: the buffers are reused
        return LBFS_OK;
    }
    adjoint_release(ctx);
    ctx->adj = new AdjointState();
    AdjointState &A = *ctx->adj;
    A.n_steps = n_steps;
    A.every = every;
    for (int t = 0; t < ctx->lay.n_tiles; t++)
        A.max_tile_cells = std::max(A.max_tile_cells, ctx->lay.tile_cell_start[t + 1] - ctx->lay.tile_cell_start[t]);
    const int n = ctx->lay.n_cells;
    int rc;
    const int n_ckpt = (n_steps + every - 1) / every;
    A.ckpt.resize(n_ckpt);
    for (auto &s : A.ckpt)
        if ((rc = snap_alloc(ctx, s))) return rc;
    if ((rc = alloc_records(ctx, A.rec, every))) return rc;
    if (n_ckpt > 1) {
        // the overlapped replays: a second record set, its own stream
        // (default priority) and events
        if ((rc = alloc_records(ctx, A.rec2, every))) return rc;
        HIPCHK(hipStreamCreateWithFlags(&A.rs, hipStreamNonBlocking));
        for (hipEvent_t &e : A.rev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if ((rc = snap_alloc(ctx, A.end))) return rc;
    for (auto &b : A.lam_st)
        if ((rc = adj_alloc(ctx, &b, (size_t)n))) return rc;
    if ((rc = adj_alloc(ctx, &A.mu, (size_t)n))) return rc;
    if ((rc = adj_alloc(ctx, &A.lcf, (size_t)n))) return rc;
    if ((rc = adj_alloc(ctx, &A.ls, (size_t)n))) return rc;
    if ((rc = adj_alloc(ctx, &A.ls_acc, (size_t)n))) return rc;
    if ((rc = adj_alloc(ctx, &A.lam_g, 3 * (size_t)n))) return rc;
    A.n_fpart = (n + 255) / 256;
    if ((rc = adj_alloc(ctx, &A.fpart, 3 * (size_t)A.n_fpart))) return rc;
    if ((rc = adj_alloc(ctx, &A.gforce, 3))) return rc;
    // reverse stencil rows: for slot k of cell i with neighbour c, the sum of
    // the weight vectors of c's slots that point back to i; every such slot
    // of c is claimed by exactly one slot of i (the stencil is symmetric on
    // face-neighbour meshes; anything else is refused)
    {
        const std::vector<int> &sten = ctx->lay.stencil;
        const std::vector<float> &gw = ctx->lay.gw;
        std::vector<float4> rw(6 * (size_t)n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        std::vector<unsigned char> used(6 * (size_t)n, 0);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 6; k++) {
                const int c = sten[(size_t)k * n + i];
                if (c < 0) continue;
                for (int kb = 0; kb < 6; kb++) {
                    if (used[(size_t)kb * n + c] || sten[(size_t)kb * n + c] != i) continue;
                    used[(size_t)kb * n + c] = 1;
                    float4 &w = rw[(size_t)k * n + i];
                    w.x += gw[(size_t)(3 * kb) * n + c];
                    w.y += gw[(size_t)(3 * kb + 1) * n + c];
                    w.z += gw[(size_t)(3 * kb + 2) * n + c];
                    break;
                }
            }
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 6; k++)
                if (sten[(size_t)k * n + i] >= 0 && !used[(size_t)k * n + i])
                    return ctx->fail(LBFS_E_UNSUPPORTED, "the adjoint needs a symmetric gradient stencil");
        if ((rc = adj_alloc(ctx, &A.rev_w, rw.size()))) return rc;
        HIPCHK(hipMemcpyAsync(A.rev_w, rw.data(), sizeof(float4) * rw.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));  // (rw is freed at the end of the block)
    }
    if (ctx->have_membranes) {
        const MemArgs &m = ctx->mem;
        const size_t nn = m.n_nodes, ne = m.n_edges, nm = m.n_mem;
        if ((rc = adj_alloc(ctx, &A.lam_X, nn))) return rc;
        if ((rc = adj_alloc(ctx, &A.lam_U, nn))) return rc;
        if ((rc = adj_alloc(ctx, &A.lam_F, nn))) return rc;
        if ((rc = adj_alloc(ctx, &A.grec, 4 * ne))) return rc;
        if ((rc = adj_alloc(ctx, &A.erec, 24 * ne))) return rc;
        if ((rc = adj_alloc(ctx, &A.mrec, 12 * ne))) return rc;
        if ((rc = adj_alloc(ctx, &A.gmem, 4 * nm))) return rc;
        if ((rc = adj_alloc(ctx, &A.umax, 1))) return rc;
        if ((rc = adj_alloc(ctx, &A.acc, 3 * (size_t)n))) return rc;
        // chunks of at most 256 edges per membrane for the two-level sums
        std::vector<int> ch, cs(1, 0);
        for (int i = 0; i < (int)nm; i++) {
            for (int e = ctx->mem_edge_start[i]; e < ctx->mem_edge_start[i + 1]; e += 256) {
                ch.push_back(i);
                ch.push_back(e);
                ch.push_back(std::min(e + 256, ctx->mem_edge_start[i + 1]));
            }
            cs.push_back((int)ch.size() / 3);
        }
        A.n_echunks = cs.back();
        if ((rc = adj_alloc(ctx, &A.echunk, ch.size()))) return rc;
        if ((rc = adj_alloc(ctx, &A.echunk_start, cs.size()))) return rc;
        if ((rc = adj_alloc(ctx, &A.epart, 6 * (size_t)A.n_echunks))) return rc;
        HIPCHK(hipMemcpyAsync(A.echunk, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(A.echunk_start, cs.data(), sizeof(int) * cs.size(), hipMemcpyHostToDevice,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    // the buffers were cleared on the context's stream
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}

// One step of a segment replay, from the live state with its fluid state in
// `st` (the record's start state), keeping what the reverse step reads; the
// new state goes to `st_next`.  The kernels and their order are those of
// lbfs_step.
// streams and events a replay runs on: the context's own, or those of the
// overlapped replays
struct Lanes {
    hipStream_t main, side, side2;
    hipEvent_t fork, join, join2;
};
Lanes ctx_lanes(lbfs_ctx *ctx) { return {ctx->stream, ctx->side, ctx->side2, ctx->ev_fork, ctx->ev_join, ctx->ev_join2}; }
Lanes replay_lanes(AdjointState &A) { return {A.rs, A.rs, A.rs, A.rev[0], A.rev[1], A.rev[2]}; }

// the fluid pass of a replayed step, keeping its gradients and face moments:
// the fused pass where the step uses it (it writes the gradients as well),
// else the gradient pass and the flux pass
hipError_t replay_fluid(const lbfs_ctx *ctx, FluidArgs a, StepRec &r, const float4 *st, float4 *st_next,
                        hipStream_t s)
{
    a.face_moments = r.mom;
    if (ctx->fused && ctx->fused_ok && a.ts_mode != LBFS_TS_MIN) {
        a.grad_out = r.grad;
        return launch_fluid_fused(a, st, st_next, r.res, 1, s);
    }
    const hipError_t e = launch_gradients(a, st, r.grad, s);
    return e != hipSuccess ? e : launch_flux_update(a, st, r.grad, st_next, r.res, 1, s);
}

int replay_step(lbfs_ctx *ctx, StepRec &r, const float4 *st, float4 *st_next, const Lanes &q)
{
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    const size_t n = (size_t)ctx->lay.n_cells;
    if (!ctx->have_membranes) {
        HIPCHK(replay_fluid(ctx, ctx->args, r, st, st_next, q.main));
        return LBFS_OK;
    }
    // as in the time step: after the interpolation, the gradients and the
    // residual pass beside the coupling chain, the membrane forces beside the
    // node binning; the node kernels write into the record (X^n is advected
    // in r.Xn1; U^n, F^n and the step's globals land in r.U, r.F, r.glob)
    MemArgs m = ctx->mem;
    const size_t nn = m.n_nodes, nm = m.n_mem;
    HIPCHK(hipMemcpyAsync(r.Xn1, r.X, sizeof(float4) * nn, k, q.main));
    HIPCHK(hipMemcpyAsync(r.glob, r.glob_prev, sizeof(double) * 8 * nm, k, q.main));
    m.X = r.Xn1;
    m.U = r.U;
    m.F = r.F;
    m.glob = r.glob;
    HIPCHK(launch_ibm_interp(m, ctx->ibm, st, q.main));
    HIPCHK(hipEventRecord(q.fork, q.main));
    HIPCHK(hipStreamWaitEvent(q.side, q.fork, 0));
    HIPCHK(replay_fluid(ctx, ctx->args, r, st, st_next, q.side));
    HIPCHK(hipEventRecord(q.join, q.side));
    HIPCHK(hipStreamWaitEvent(q.side2, q.fork, 0));
    HIPCHK(launch_mem_volume(m, q.side2));
    HIPCHK(launch_mem_edges(m, q.side2));
    if (!spread_gathers_forces(ctx->ibm)) HIPCHK(launch_mem_nodes(m, q.side2));
    HIPCHK(hipEventRecord(q.join2, q.side2));
    HIPCHK(hipStreamWaitEvent(q.main, q.join2, 0));
    HIPCHK(launch_ibm_spread(m, ctx->ibm, q.main));
    HIPCHK(hipStreamWaitEvent(q.main, q.join, 0));
    // the step's cell force goes straight into the record (every cell)
    lbfs::ApplyArgs p = ctx->apply_args(st_next, r.cf, 1);
    p.glob = m.glob;
    HIPCHK(launch_ibm_apply(ctx->ibm, p, q.main));
    return LBFS_OK;
}

// One reverse step.  updated: the update adjoint of this step already ran
// (fused into the gradient transpose of the step after); prev: the record of
// the step before, whose update adjoint then runs fused into this step's
// gradient transpose (null: it runs at the start of that step).
int reverse_step(lbfs_ctx *ctx, const StepRec &r, bool first, int &cur_l, bool updated, const StepRec *prev)
{
    AdjointState &A = *ctx->adj;
    FluidArgs a = ctx->args;
    if (ctx->have_membranes) a.cell_force = r.cf;
    a.face_moments = r.mom;
    AdjFluidArgs j{};
    j.lam_in = A.lam_st[cur_l];
    j.lam_out = A.lam_st[cur_l ^ 1];
    j.mu = A.mu; j.lcf = A.lcf; j.ls = A.ls; j.ls_acc = A.ls_acc; j.lam_g = A.lam_g;
    j.fpart = A.fpart; j.gforce = A.gforce; j.n_fpart = A.n_fpart;
    j.rev_w = A.rev_w;
    j.acc = ctx->have_membranes ? A.acc : nullptr;
    j.acc_rw = ctx->have_membranes ? A.acc : nullptr;
    j.umax = A.umax;
    AdjPrevStep pv{};
    if (prev) pv = {prev->st, prev->res, ctx->have_membranes ? prev->cf : ctx->args.cell_force};
    if (!updated) HIPCHK(launch_adj_update(a, j, r.st, r.res, ctx->stream));
    if (!ctx->have_membranes) {
        HIPCHK(launch_adj_flux(a, j, r.st, r.grad, A.max_tile_cells, ctx->stream));
        HIPCHK(launch_adj_grad(a, j, prev ? &pv : nullptr, ctx->stream));
        cur_l ^= 1;
        return LBFS_OK;
    }
    // the face adjoint (side stream) beside the coupling adjoint chain; both
    // read only the update's adjoint, and the gradient transpose waits for both
    HIPCHK(hipEventRecord(ctx->ev_fork, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
    HIPCHK(launch_adj_flux(a, j, r.st, r.grad, A.max_tile_cells, ctx->side));
    HIPCHK(hipEventRecord(ctx->ev_join, ctx->side));
    MemArgs m = ctx->mem;
    m.X = r.Xn1; m.U = r.U; m.F = r.F; m.glob = r.glob;
    AdjMemArgs jm{};
    jm.Xn = r.X; jm.glob_prev = r.glob_prev;
    jm.lam_X = A.lam_X; jm.lam_U = A.lam_U; jm.lam_F = A.lam_F;
    jm.erec = A.erec; jm.grec = A.grec; jm.mrec = A.mrec; jm.gmem = A.gmem;
    jm.umax = A.umax; jm.lcf = A.lcf; jm.st_n = r.st; jm.acc = A.acc;
    jm.mad_lag = first ? 0 : 1;
    jm.echunk = A.echunk; jm.echunk_start = A.echunk_start; jm.epart = A.epart; jm.n_echunks = A.n_echunks;
    HIPCHK(launch_adj_coupling(m, ctx->ibm, jm, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    HIPCHK(launch_adj_grad(a, j, prev ? &pv : nullptr, ctx->stream));
    cur_l ^= 1;
    return LBFS_OK;
}

// Replays segment sg from its checkpoint into the step records; the state
// after the segment is left in ctx->st[1].
// a segment replay enqueued step by step (so that it can be interleaved with
// the reverse of another segment on other streams)
struct SegReplay {
    int sg = -1, s0 = 0, s1 = 0, t = 0;
    std::vector<StepRec> *recs = nullptr;
    Lanes q{};
};

int replay_begin(lbfs_ctx *ctx, int sg, std::vector<StepRec> &recs, const Lanes &q, SegReplay &p)
{
    AdjointState &A = *ctx->adj;
    p.sg = sg;
    p.s0 = sg * A.every;
    p.s1 = std::min(p.s0 + A.every, A.n_steps);
    p.t = p.s0;
    p.recs = &recs;
    p.q = q;
    return restore(ctx, A.ckpt[sg], q.main);
}

bool replay_pending(const SegReplay &p) { return p.recs && p.t < p.s1; }

int replay_next(lbfs_ctx *ctx, SegReplay &p)
{
    AdjointState &A = *ctx->adj;
    std::vector<StepRec> &recs = *p.recs;
    const int t = p.t, s0 = p.s0, s1 = p.s1;
    StepRec &r = recs[t - s0];
    const StepRec *pr = t > s0 ? &recs[t - 1 - s0] : nullptr;
    if (!pr) r.st = A.ckpt[p.sg].st;
    r.X = pr ? pr->Xn1 : A.ckpt[p.sg].X;
    r.glob_prev = pr ? pr->glob : A.ckpt[p.sg].glob;
    float4 *next = t + 1 < s1 ? recs[t + 1 - s0].st : ctx->st[1];
    p.t++;
    return replay_step(ctx, r, t == s0 ? ctx->st[0] : r.st, next, p.q);
}

// replays segment sg from its checkpoint into the step records `recs`
int record_segment(lbfs_ctx *ctx, int sg, std::vector<StepRec> &recs, const Lanes &q)
{
    SegReplay p;
    int rc;
    if ((rc = replay_begin(ctx, sg, recs, q, p))) return rc;
    while (replay_pending(p))
        if ((rc = replay_next(ctx, p))) return rc;
    return LBFS_OK;
}

}  // namespace

extern "C" int lbfs_adjoint_forward(lbfs_ctx *ctx, int n_steps, int checkpoint_every)
{
    if (!ctx || n_steps <= 0 || checkpoint_every <= 0) return LBFS_E_INVALID_ARG;
    int rc;
    if ((rc = adjoint_supported(ctx))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int every = std::min(checkpoint_every, n_steps);
    if ((rc = adjoint_setup(ctx, n_steps, every))) return rc;
    AdjointState &A = *ctx->adj;
    A.start = A.ckpt[0];
    if ((rc = save(ctx, A.ckpt[0]))) return rc;
    // the steps before the last segment, checkpointed every K steps
    const int n_seg = (n_steps + every - 1) / every, s_last = (n_seg - 1) * every;
    for (int t = 0; t < s_last; t++) {
        if ((rc = ctx->enqueue_step())) return rc;
        if ((t + 1) % every == 0)
            if ((rc = save(ctx, A.ckpt[(t + 1) / every]))) return rc;
    }
    // the last segment straight into the step records (the same kernels on
    // the same inputs), so the backward pass starts without replaying it;
    // then the solver takes the end state from the records
    if ((rc = record_segment(ctx, n_seg - 1, A.rec, ctx_lanes(ctx)))) return rc;
    {
        const StepRec &r = A.rec[n_steps - 1 - s_last];
        const hipMemcpyKind k = hipMemcpyDeviceToDevice;
        ctx->cur = 1;
        ctx->t_step += n_steps - s_last;
        if (ctx->have_membranes) {
            MemArgs &m = ctx->mem;
            const size_t nn = m.n_nodes, nm = m.n_mem;
            HIPCHK(hipMemcpyAsync(m.X, r.Xn1, sizeof(float4) * nn, k, ctx->stream));
            HIPCHK(hipMemcpyAsync(m.U, r.U, sizeof(float4) * nn, k, ctx->stream));
            HIPCHK(hipMemcpyAsync(m.F, r.F, sizeof(float4) * nn, k, ctx->stream));
            HIPCHK(hipMemcpyAsync(m.glob, r.glob, sizeof(double) * 8 * nm, k, ctx->stream));
            HIPCHK(hipMemcpyAsync(ctx->cell_force, r.cf, sizeof(float) * 3 * (size_t)ctx->lay.n_cells, k,
                                  ctx->stream));
            // the cell force output now holds spread force anywhere: the
            // next apply pass resets every tile it does not write
            HIPCHK(hipMemsetD32Async((hipDeviceptr_t)ctx->ibm.tile_prev, 1, ctx->lay.n_tiles, ctx->stream));
        }
    }
    A.last_in_rec = true;
    if ((rc = save(ctx, A.end))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    A.recorded = true;
    return LBFS_OK;
}

extern "C" int lbfs_adjoint_rewind(lbfs_ctx *ctx)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    if (!ctx->adj || !ctx->adj->recorded) return ctx->fail(LBFS_E_INVALID_ARG, "no recorded forward run");
    int rc;
    if ((rc = restore(ctx, ctx->adj->start))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}

extern "C" int lbfs_adjoint_backward(lbfs_ctx *ctx, const double *lam_nodes, const double *lam_state,
                                     double *grad_body_force, double *grad_moduli, double *lam_state0,
                                     double *lam_nodes0)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    if (!ctx->adj || !ctx->adj->recorded) return ctx->fail(LBFS_E_INVALID_ARG, "no recorded forward run");
    if ((lam_nodes || lam_nodes0) && !ctx->have_membranes)
        return ctx->fail(LBFS_E_INVALID_ARG, "node adjoints given without membranes");
    AdjointState &A = *ctx->adj;
    const int n = ctx->lay.n_cells;
    int rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // seeds: adjoints of the final state and node positions, cleared sums;
    // all on the context's stream (it does not wait for the null stream)
    {
        std::vector<float4> h(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        if (lam_state)
            for (int i = 0; i < n; i++) {
                const int c = ctx->lay.perm[i];
                h[i] = make_float4((float)lam_state[4 * (size_t)c], (float)lam_state[4 * (size_t)c + 1],
                                   (float)lam_state[4 * (size_t)c + 2], (float)lam_state[4 * (size_t)c + 3]);
            }
        HIPCHK(hipMemcpyAsync(A.lam_st[0], h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemsetAsync(A.gforce, 0, 3 * sizeof(double), ctx->stream));
        HIPCHK(hipMemsetAsync(A.fpart, 0, 3 * sizeof(double) * A.n_fpart, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (ctx->have_membranes) {
        const int nn = ctx->mem.n_nodes;
        std::vector<float4> h(nn, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        if (lam_nodes)
            for (int i = 0; i < nn; i++)
                h[i] = make_float4((float)lam_nodes[3 * (size_t)i], (float)lam_nodes[3 * (size_t)i + 1],
                                   (float)lam_nodes[3 * (size_t)i + 2], 0.0f);
        HIPCHK(hipMemcpyAsync(A.lam_X, h.data(), sizeof(float4) * nn, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemsetAsync(A.gmem, 0, 4 * sizeof(double) * ctx->mem.n_mem, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    int cur_l = 0;
    const int K = A.every, N = A.n_steps;
    const int n_seg = (N + K - 1) / K;
    // Segment sg's records: set (n_seg - 1 - sg) % 2 (the forward pass left
    // the last segment in set 0).  With several segments the replay of the
    // segment before runs on its own stream while the reverse of this one
    // runs on the context's: it waits until the reverse of the segment after
    // has released its set, and the reverse of a segment waits for its replay.
    std::vector<StepRec> *sets[2] = {&A.rec, &A.rec2};
    hipEvent_t *ready = A.rev + 3, *reversed = A.rev + 5;
    if (!A.last_in_rec) {
        if (n_seg > 1) {
            HIPCHK(hipEventRecord(ready[1], ctx->stream));  // the seeds above come first
            HIPCHK(hipStreamWaitEvent(A.rs, ready[1], 0));
            if ((rc = record_segment(ctx, n_seg - 1, A.rec, replay_lanes(A)))) return rc;
            HIPCHK(hipEventRecord(ready[0], A.rs));
        } else if ((rc = record_segment(ctx, n_seg - 1, A.rec, ctx_lanes(ctx)))) {
            return rc;
        }
    } else if (n_seg > 1) {
        HIPCHK(hipEventRecord(ready[0], ctx->stream));
    }
    A.last_in_rec = false;
    for (int sg = n_seg - 1; sg >= 0; sg--) {
        const int s0 = sg * K, s1 = std::min(s0 + K, N), b = (n_seg - 1 - sg) & 1;
        SegReplay p;
        if (sg > 0) {
            // replay the segment before into the other set
            if (sg + 1 < n_seg) HIPCHK(hipStreamWaitEvent(A.rs, reversed[b ^ 1], 0));
            else HIPCHK(hipStreamWaitEvent(A.rs, ready[b], 0));  // the live state is free
            if ((rc = replay_begin(ctx, sg - 1, *sets[b ^ 1], replay_lanes(A), p))) return rc;
        }
        if (n_seg > 1) HIPCHK(hipStreamWaitEvent(ctx->stream, ready[b], 0));
        // the two streams' steps enqueued in turn, so both have work at once;
        // within the segment each step's update adjoint runs fused into the
        // gradient transpose of the step after (the segment's first step
        // takes the record of the segment before, which is still being
        // replayed: its update adjoint runs on its own)
        for (int t = s1 - 1; t >= s0; t--) {
            if (replay_pending(p) && (rc = replay_next(ctx, p))) return rc;
            const StepRec *prev = t > s0 ? &(*sets[b])[t - 1 - s0] : nullptr;
            if ((rc = reverse_step(ctx, (*sets[b])[t - s0], t == 0, cur_l, t < s1 - 1, prev))) return rc;
        }
        while (replay_pending(p))
            if ((rc = replay_next(ctx, p))) return rc;
        if (sg > 0) HIPCHK(hipEventRecord(ready[b ^ 1], A.rs));
        if (n_seg > 1) HIPCHK(hipEventRecord(reversed[b], ctx->stream));
    }
    if (n_seg == 1) A.last_in_rec = true;  // the records still hold the only segment
    {
        AdjFluidArgs j{};
        j.fpart = A.fpart; j.gforce = A.gforce; j.n_fpart = A.n_fpart;
        HIPCHK(launch_adj_force_sum(j, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (grad_body_force) HIPCHK(hipMemcpy(grad_body_force, A.gforce, 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (grad_moduli && ctx->have_membranes) {
        const int nm = ctx->mem.n_mem;
        std::vector<double> g(4 * (size_t)nm);
        HIPCHK(hipMemcpy(g.data(), A.gmem, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < nm; i++) {
            // the spring constants are proportional to the shear modulus
            grad_moduli[3 * (size_t)i] = ctx->mem_shear[i] != 0.0 ? g[4 * (size_t)i] / ctx->mem_shear[i] : 0.0;
            grad_moduli[3 * (size_t)i + 1] = g[4 * (size_t)i + 1];
            grad_moduli[3 * (size_t)i + 2] = g[4 * (size_t)i + 2];
        }
    }
    if (lam_state0) {
        std::vector<float4> h(n);
        HIPCHK(hipMemcpy(h.data(), A.lam_st[cur_l], sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            const int c = ctx->lay.perm[i];
            lam_state0[4 * (size_t)c] = h[i].x;
            lam_state0[4 * (size_t)c + 1] = h[i].y;
            lam_state0[4 * (size_t)c + 2] = h[i].z;
            lam_state0[4 * (size_t)c + 3] = h[i].w;
        }
    }
    if (lam_nodes0) {
        const int nn = ctx->mem.n_nodes;
        std::vector<float4> h(nn);
        HIPCHK(hipMemcpy(h.data(), A.lam_X, sizeof(float4) * nn, hipMemcpyDeviceToHost));
        for (int i = 0; i < nn; i++) {
            lam_nodes0[3 * (size_t)i] = h[i].x;
            lam_nodes0[3 * (size_t)i + 1] = h[i].y;
            lam_nodes0[3 * (size_t)i + 2] = h[i].z;
        }
    }
    // leave the context as the forward run left it
    if ((rc = restore(ctx, A.end))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}

extern "C" int lbfs_adjoint_run(lbfs_ctx *ctx, int n_steps, int checkpoint_every, const double *target_nodes,
                                double *loss, double *grad_body_force, double *grad_moduli)
{
    if (!ctx || !target_nodes || !loss) return LBFS_E_INVALID_ARG;
    if (!ctx->have_membranes) return ctx->fail(LBFS_E_INVALID_ARG, "the node loss needs membranes");
    int rc;
    if ((rc = lbfs_adjoint_forward(ctx, n_steps, checkpoint_every))) return rc;
    const int nn = ctx->mem.n_nodes;
    std::vector<double> X(3 * (size_t)nn), lam(3 * (size_t)nn);
    if ((rc = lbfs_get_membrane_nodes(ctx, X.data(), nullptr, nullptr))) return rc;
    double L = 0.0;
    for (size_t k = 0; k < X.size(); k++) {
        const double d = X[k] - target_nodes[k];
        L += d * d;
        lam[k] = 2.0 * d;
    }
    *loss = L;
    return lbfs_adjoint_backward(ctx, lam.data(), nullptr, grad_body_force, grad_moduli, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
extern "C" int lbfs_set_body_force(lbfs_ctx *ctx, const double *f)
{
    if (!ctx || !f) return LBFS_E_INVALID_ARG;
    for (int t = 0; t < 3; t++)
        if (!std::isfinite(f[t])) return ctx->fail(LBFS_E_INVALID_ARG, "non-finite body force");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int n = ctx->lay.n_cells;
    const int dims = ctx->params.lattice == LBFS_D2Q9 ? 2 : 3;
    const double rho0 = ctx->params.rho0;
    std::vector<float> sf(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const int c = ctx->lay.perm[i];
        for (int t = 0; t < 3; t++)
            sf[3 * (size_t)i + t] =
                (float)(f[t] - (t < dims ? rho0 / 3.0 * ctx->lay.closure[3 * (size_t)c + t] / ctx->geo.vol[c] : 0.0));
    }
    // (copies on the context's stream: it does not wait for the null stream)
    HIPCHK(hipMemcpyAsync(ctx->static_force, sf.data(), sizeof(float) * sf.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    // every cell holds the static force until the next spreading overwrites it
    HIPCHK(hipMemcpyAsync(ctx->cell_force, sf.data(), sizeof(float) * sf.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int t = 0; t < 3; t++) ctx->params.body_force[t] = f[t];
    for (lbfs::HostBlock &b : ctx->blocks)
        for (int t = 0; t < 3; t++) b.force[t] = f[t];
    return LBFS_OK;
}

extern "C" int lbfs_set_membrane_moduli(lbfs_ctx *ctx, int index, double shear_modulus, double local_bending_modulus,
                                        double global_bending_modulus)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    if (!ctx->have_membranes || index < 0 || index >= ctx->mem.n_mem)
        return ctx->fail(LBFS_E_INVALID_ARG, "no membrane %d", index);
    if (!std::isfinite(shear_modulus) || !std::isfinite(local_bending_modulus) || !std::isfinite(global_bending_modulus))
        return ctx->fail(LBFS_E_INVALID_ARG, "non-finite modulus");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    MemParamDev &p = ctx->mem_params[index];
    p.local_bending = (float)local_bending_modulus;
    p.global_bending = (float)global_bending_modulus;
    ctx->mem_shear[index] = shear_modulus;
    HIPCHK(hipMemcpyAsync((void *)(ctx->mem.params + index), &p, sizeof(MemParamDev), hipMemcpyHostToDevice,
                          ctx->stream));
    const int e0 = ctx->mem_edge_start[index], e1 = ctx->mem_edge_start[index + 1];
    for (int e = e0; e < e1; e++) ctx->mem_edge_const[e].y = (float)(ctx->mem_ks_unit[e] * shear_modulus);
    HIPCHK(hipMemcpyAsync((void *)(ctx->mem.edge_const + e0), ctx->mem_edge_const.data() + e0,
                          sizeof(float4) * (e1 - e0), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}
