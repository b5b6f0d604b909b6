// Host side of the membrane / IBM coupling: set-up of the device arrays of all
// membranes (Solver_gpu.cu:473-611, 659-732 for one membrane; here any number,
// each with its own parameters and connectivity offsets) and state readback.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "coupling.hpp"
#include "membrane_host.hpp"

using namespace lbfs;

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return ctx->fail(LBFS_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <class T>
static int cp_alloc(lbfs_ctx *ctx, T **dst, const T *src, size_t n)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, (n > 0 ? n : 1) * sizeof(T) + 16));
    ctx->coupling_allocs.push_back(p);
    if (src && n) HIPCHK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    else HIPCHK(hipMemsetAsync(p, 0, (n > 0 ? n : 1) * sizeof(T), ctx->stream));  // ordered before the kernels
    *dst = (T *)p;
    return LBFS_OK;
}

void lbfs_ctx::release_extensions()
{
    adjoint_release(this);
    for (void *p : coupling_allocs) (void)hipFree(p);
    coupling_allocs.clear();
    have_membranes = false;
    mem_energy = nullptr;
    visc_factor = nullptr;
    args.visc_factor = nullptr;
    visc_grid = ViscGrid{};
}

extern "C" int lbfs_add_membranes(lbfs_ctx *ctx, const lbfs_membrane_desc *mem, int count, int ibm_kernel)
{
    if (!ctx || !mem || count <= 0) return LBFS_E_INVALID_ARG;
    if (ibm_kernel != LBFS_IBM_LINEAR2 && ibm_kernel != LBFS_IBM_COS4 && ibm_kernel != LBFS_IBM_LAGRANGE4)
        return ctx->fail(LBFS_E_INVALID_ARG, "unknown IBM kernel %d", ibm_kernel);
    if (ctx->args.ts_mode != 0) return ctx->fail(LBFS_E_UNSUPPORTED, "membranes need constant time steps");
    if (ctx->blocks.empty() || !(ctx->blocks[0].spacing > 0.0))
        return ctx->fail(LBFS_E_UNSUPPORTED, "membranes need a mesh with a logical i,j,k structure and spacing");
    for (int i = 0; i < count; i++)
        if (mem[i].block < 0 || mem[i].block >= (int)ctx->blocks.size())
            return ctx->fail(LBFS_E_INVALID_ARG, "membrane %d: block %d does not exist", i, mem[i].block);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->drop_graph();
    ctx->release_extensions();
    // restore the force field without IBM contributions
    HIPCHK(hipMemcpyAsync(ctx->cell_force, ctx->static_force, sizeof(float) * 3 * (size_t)ctx->lay.n_cells,
                          hipMemcpyDeviceToDevice, ctx->stream));
    std::vector<MembraneSetup> ms(count);
    std::string err;
    for (int i = 0; i < count; i++) {
        int rc = membrane_setup(mem[i], ms[i], err);
        if (rc) return ctx->fail(rc, "membrane %d: %s", i, err.c_str());
    }
    // concatenated host arrays
    std::vector<float4> X, origin(count);
    std::vector<int> node_mem, mem_tri_start(1, 0), mem_edge_start(1, 0), mem_node_start(1, 0), edge_mem;
    std::vector<int4> tri, edges;
    std::vector<float4> edge_const, edge_state;
    std::vector<float2> edge_area0;
    std::vector<MemParamDev> params(count);
    std::vector<double> glob(8 * (size_t)count, 0.0);
    for (int i = 0; i < count; i++) {
        const MembraneSetup &s = ms[i];
        const int nb = mem_node_start.back(), eb = mem_edge_start.back();
        double c[3] = {0, 0, 0};
        for (int n = 0; n < s.n_nodes; n++)
            for (int a = 0; a < 3; a++) c[a] += s.X[3 * (size_t)n + a] / s.n_nodes;
        origin[i] = make_float4((float)c[0], (float)c[1], (float)c[2], 0.0f);
        const double oc[3] = {(double)origin[i].x, (double)origin[i].y, (double)origin[i].z};
        for (int n = 0; n < s.n_nodes; n++) {
            X.push_back(make_float4((float)(s.X[3 * (size_t)n] - oc[0]), (float)(s.X[3 * (size_t)n + 1] - oc[1]),
                                    (float)(s.X[3 * (size_t)n + 2] - oc[2]), 0.0f));
            node_mem.push_back(i);
        }
        for (int t = 0; t < s.n_tri; t++)
            tri.push_back(make_int4(nb + s.tri[3 * t], nb + s.tri[3 * t + 1], nb + s.tri[3 * t + 2], i));
        for (int e = 0; e < s.n_edges; e++) {
            const int *q = &s.springs[4 * (size_t)e];
            edges.push_back(make_int4(nb + q[0], nb + q[1], nb + q[2], nb + q[3]));
            edge_mem.push_back(i);
            edge_const.push_back(make_float4((float)s.contour_length[e], (float)s.spring_constant[e],
                                             (float)s.curvature_0[e], (float)s.pow_ratio[e]));
            edge_area0.push_back(make_float2((float)s.spring_area_0[2 * (size_t)e], (float)s.spring_area_0[2 * (size_t)e + 1]));
            edge_state.push_back(make_float4((float)s.curvature[e], (float)s.spring_area[2 * (size_t)e],
                                             (float)s.spring_area[2 * (size_t)e + 1], 0.0f));
        }
        mem_node_start.push_back(nb + s.n_nodes);
        mem_tri_start.push_back(mem_tri_start.back() + s.n_tri);
        mem_edge_start.push_back(eb + s.n_edges);
        const double r = mem[i].wlc_ratio;
        MemParamDev &p = params[i];
        p.spring_c = (float)(0.25 * r * r / ((r - 1.0) * (r - 1.0)) - 0.25 + 1.0 / r);
        p.wlc_x0 = (float)(1.0 / r);
        p.spring_law = mem[i].spring_law;
        p.volume_modulus = (float)mem[i].volume_modulus;
        p.local_area_modulus = (float)mem[i].local_area_modulus;
        p.global_area_modulus = (float)mem[i].global_area_modulus;
        p.global_bending = (float)mem[i].global_bending_modulus;
        p.local_bending = (float)mem[i].local_bending_modulus;
        p.viscosity = (float)mem[i].viscosity;
        p.n_mad = mem[i].mad_compat ? std::min(s.n_nodes, s.n_edges) : s.n_edges;
        p.visc_ratio = mem[i].internal_viscosity_ratio > 0.0 ? (float)mem[i].internal_viscosity_ratio : 1.0f;
        double *gl = &glob[8 * (size_t)i];
        gl[0] = s.V; gl[1] = s.A; gl[2] = s.MAD; gl[3] = s.V0; gl[4] = s.A0; gl[5] = s.MAD0; gl[6] = s.A;
        gl[7] = s.MAD;  // the first force call uses the MAD of the initial shape
    }
    const int nn = (int)X.size(), ne = (int)edges.size();
    // node -> edge-slot incidences in (edge, slot) order
    std::vector<int> inc_start(nn + 1, 0), inc(4 * (size_t)ne);
    for (int e = 0; e < ne; e++) {
        const int q[4] = {edges[e].x, edges[e].y, edges[e].z, edges[e].w};
        for (int j = 0; j < 4; j++) inc_start[q[j] + 1]++;
    }
    for (int n = 0; n < nn; n++) inc_start[n + 1] += inc_start[n];
    {
        std::vector<int> fill(inc_start.begin(), inc_start.end() - 1);
        for (int e = 0; e < ne; e++) {
            const int q[4] = {edges[e].x, edges[e].y, edges[e].z, edges[e].w};
            for (int j = 0; j < 4; j++) inc[fill[q[j]]++] = 4 * e + j;
        }
    }
    MemArgs &m = ctx->mem;
    memset(&m, 0, sizeof(m));
    m.n_mem = count; m.n_nodes = nn; m.n_tri = (int)tri.size(); m.n_edges = ne;
    int rc;
    float4 *dX, *dU, *dF, *dorig, *dec, *des, *def;
    int *dnm, *dts, *des0, *dem, *dis, *dinc;
    int4 *dtri, *dedg;
    float2 *dea;
    MemParamDev *dpar;
    double *dgl;
    if ((rc = cp_alloc(ctx, &dX, X.data(), nn))) return rc;
    if ((rc = cp_alloc<float4>(ctx, &dU, nullptr, nn))) return rc;
    if ((rc = cp_alloc<float4>(ctx, &dF, nullptr, nn))) return rc;
    if ((rc = cp_alloc(ctx, &dorig, origin.data(), count))) return rc;
    if ((rc = cp_alloc(ctx, &dnm, node_mem.data(), nn))) return rc;
    std::vector<float4> node_origin(nn);
    for (size_t i = 0; i < nn; i++) node_origin[i] = origin[node_mem[i]];
    float4 *dnorig;
    if ((rc = cp_alloc(ctx, &dnorig, node_origin.data(), nn))) return rc;
    if ((rc = cp_alloc(ctx, &dtri, tri.data(), tri.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dts, mem_tri_start.data(), mem_tri_start.size()))) return rc;
    if ((rc = cp_alloc(ctx, &des0, mem_edge_start.data(), mem_edge_start.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dedg, edges.data(), ne))) return rc;
    if ((rc = cp_alloc(ctx, &dem, edge_mem.data(), ne))) return rc;
    if ((rc = cp_alloc(ctx, &dec, edge_const.data(), ne))) return rc;
    if ((rc = cp_alloc(ctx, &dea, edge_area0.data(), ne))) return rc;
    if ((rc = cp_alloc(ctx, &des, edge_state.data(), ne))) return rc;
    if ((rc = cp_alloc<float4>(ctx, &def, nullptr, 4 * (size_t)ne))) return rc;
    if ((rc = cp_alloc(ctx, &dis, inc_start.data(), inc_start.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dinc, inc.data(), inc.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dpar, params.data(), count))) return rc;
    if ((rc = cp_alloc(ctx, &dgl, glob.data(), glob.size()))) return rc;
    unsigned long long *dmacc;
    if ((rc = cp_alloc<unsigned long long>(ctx, &dmacc, nullptr, 4 * (size_t)count))) return rc;
    int *dns;
    float4 *dbb;
    if ((rc = cp_alloc(ctx, &dns, mem_node_start.data(), mem_node_start.size()))) return rc;
    if ((rc = cp_alloc<float4>(ctx, &dbb, nullptr, 2 * (size_t)count))) return rc;
    if ((rc = cp_alloc<double>(ctx, &ctx->mem_energy, nullptr, 7 * (size_t)count))) return rc;
    m.mem_node_start = dns;
    m.bbox = dbb;
    m.X = dX; m.U = dU; m.F = dF; m.node_mem = dnm; m.origin = dorig; m.node_origin = dnorig; m.tri = dtri;
    m.mem_tri_start = dts; m.mem_edge_start = des0; m.edges = dedg; m.edge_mem = dem;
    m.edge_const = dec; m.edge_area0 = dea; m.edge_state = des; m.edge_force = def;
    m.inc_start = dis; m.inc = dinc; m.params = dpar; m.glob = dgl; m.mem_acc = dmacc;
    ctx->mem_node_start = mem_node_start;
    ctx->mem_params = params;
    ctx->mem_edge_const = edge_const;
    ctx->mem_edge_start = mem_edge_start;
    ctx->mem_shear.assign(count, 0.0);
    ctx->mem_ks_unit.assign(ne, 0.0);
    for (int i = 0; i < count; i++) {
        ctx->mem_shear[i] = mem[i].shear_modulus;
        for (int e = 0; e < ms[i].n_edges; e++)
            ctx->mem_ks_unit[mem_edge_start[i] + e] =
                mem[i].shear_modulus != 0.0 ? ms[i].spring_constant[e] / mem[i].shear_modulus : 0.0;
    }
    ctx->mem_origin.assign(3 * (size_t)count, 0.0);
    for (int i = 0; i < count; i++) {
        ctx->mem_origin[3 * (size_t)i] = origin[i].x;
        ctx->mem_origin[3 * (size_t)i + 1] = origin[i].y;
        ctx->mem_origin[3 * (size_t)i + 2] = origin[i].z;
    }
    // IBM grids: one per logical box, bins (= spreading tiles) numbered box after box
    IbmArgs &g = ctx->ibm;
    memset(&g, 0, sizeof(g));
    // the device error word (fixed-point range of the spread and membrane sums)
    g.err = ctx->dev_err;
    m.err = ctx->dev_err;
    HIPCHK(hipMemsetAsync(ctx->dev_err, 0, sizeof(int), ctx->stream));
    g.kernel = ibm_kernel;
    // spreading path: dense suspensions (a node per 100 cells or more, where
    // most fluid tiles hold membrane) gather per fluid tile without atomics;
    // sparser ones scatter per occupied 4x4x4 bin, whose short work items
    // spread over the whole chip (measured: one 2562-node cell in a
    // 128x64x64 channel 13 us against 70 us; 380k nodes in 256x128x128
    // 0.28 ms against 0.85 ms)
    const int nc0 = ctx->lay.n_cells;
    if (ctx->ibm_binned_option >= 0) g.binned = ctx->ibm_binned_option;
    else g.binned = (double)nn >= 0.01 * nc0 ? IBM_SPREAD_GATHER : IBM_SPREAD_BINS;
    if (g.binned == IBM_SPREAD_GATHER && spread_tile_lds(ctx->args.max_tile_cells) > 150 * 1024)
        g.binned = IBM_SPREAD_BINS;  // tiles beyond the LDS budget
    const int nblk = (int)ctx->blocks.size();
    std::vector<IbmBlock> blk(nblk);
    std::vector<int4> tile_lo;  // first cell of each node-list tile in its box, and the box
    std::vector<int> cell_blk_caller(ctx->lay.n_cells, 0);
    int ftile0 = 0;
    for (int b = 0; b < nblk; b++) {
        const HostBlock &hb = ctx->blocks[b];
        IbmBlock &B = blk[b];
        for (int a = 0; a < 3; a++) {
            B.n[a] = hb.n[a];
            B.origin[a] = (float)hb.origin[a];
            B.L[a] = (float)(hb.n[a] * hb.spacing);
            B.per[a] = 0;
            B.ftw[a] = ctx->lay.tile[a];
            B.fnt[a] = (hb.n[a] + B.ftw[a] - 1) / B.ftw[a];
            // node lists per fluid tile (gathering) or per 4x4x4 bin
            B.tw[a] = g.binned == IBM_SPREAD_GATHER ? B.ftw[a] : 4;
            B.nt[a] = (hb.n[a] + B.tw[a] - 1) / B.tw[a];
        }
        B.h = (float)hb.spacing;
        B.inv_h = (float)(1.0 / hb.spacing);
        B.cell0 = hb.cell0;
        // the fluid tiles of the box come in the order of the device layout
        auto log2_or = [](int v) { int s = 0; while ((1 << s) < v) s++; return (1 << s) == v ? s : -1; };
        B.regular = 1;
        for (int a = 0; a < 3; a++) {
            B.fsh[a] = log2_or(B.ftw[a]);
            B.tsh[a] = log2_or(B.tw[a]);
            B.regular &= B.fsh[a] >= 0 && hb.n[a] % B.ftw[a] == 0;
        }
        B.icell0 = ctx->lay.tile_cell_start[ftile0];
        ftile0 += B.fnt[0] * B.fnt[1] * B.fnt[2];
        B.tile0 = (int)tile_lo.size();
        for (int tz = 0; tz < B.nt[2]; tz++)
            for (int ty = 0; ty < B.nt[1]; ty++)
                for (int tx = 0; tx < B.nt[0]; tx++)
                    tile_lo.push_back(make_int4(tx * B.tw[0], ty * B.tw[1], tz * B.tw[2], b));
        const int nb_cells = hb.n[0] * hb.n[1] * hb.n[2];
        for (int c = 0; c < nb_cells; c++) cell_blk_caller[hb.cell0 + c] = b;
    }
    if (ftile0 != ctx->lay.n_tiles)
        return ctx->fail(LBFS_E_UNSUPPORTED, "the boxes' tiles (%d) do not match the layout's (%d)", ftile0,
                         ctx->lay.n_tiles);
    // periodic axes of each box from the translations of its wrap faces
    for (int f = 0; f < ctx->geo.n_faces; f++)
        for (int a = 0; a < 3; a++)
            if (ctx->geo.shift[3 * (size_t)f + a] != 0.0) {
                IbmBlock &B = blk[cell_blk_caller[ctx->geo.owner[f]]];
                B.per[a] = 1;
                B.L[a] = (float)std::fabs(ctx->geo.shift[3 * (size_t)f + a]);
            }
    for (IbmBlock &B : blk)
        for (int a = 0; a < 3; a++) B.period_step[a] = B.L[a] / (float)B.n[a];
    // 4-point Lagrange kernel: per box and axis the vertex planes along the
    // first grid line (face centroids of its cells, the reference's plateau
    // arrays) and the cell-line centroids
    std::vector<double> axis;
    for (int b = 0; b < nblk; b++) {
        const HostBlock &hb = ctx->blocks[b];
        const Geometry &gm = ctx->geo;
        const int stride[3] = {1, hb.n[0], hb.n[0] * hb.n[1]};
        for (int a = 0; a < 3; a++) {
            blk[b].ax_off[a] = (int)axis.size();
            const int n = hb.n[a];
            std::vector<double> plane(n + 1), line(n);
            for (int i = 0; i < n; i++) {
                const int c = hb.cell0 + i * stride[a];
                line[i] = gm.cc[3 * (size_t)c + a];
                double lo_n = 2.0, hi_n = -2.0;
                for (int k = 0; k < 6; k++) {
                    const int f = gm.gface[6 * (size_t)c + k];
                    const bool nb = gm.grole[6 * (size_t)c + k] == 1;
                    const double na = (nb ? -1.0 : 1.0) * gm.normal[3 * (size_t)f + a];
                    const double x = gm.fc[3 * (size_t)f + a] - (nb ? gm.shift[3 * (size_t)f + a] : 0.0);
                    if (na < lo_n) { lo_n = na; plane[i] = x; }
                    if (i == n - 1 && na > hi_n) { hi_n = na; plane[n] = x; }
                }
            }
            axis.insert(axis.end(), plane.begin(), plane.end());
            axis.insert(axis.end(), line.begin(), line.end());
        }
    }
    double *dax;
    if ((rc = cp_alloc(ctx, &dax, axis.data(), axis.size()))) return rc;
    g.axis = dax;
    g.n_blocks = nblk;
    g.n_tiles = (int)tile_lo.size();
    g.dt = (float)ctx->params.dt;
    std::vector<int> node_block(nn), mem_block(count);
    for (int i = 0; i < count; i++) {
        mem_block[i] = mem[i].block;
        for (int k = mem_node_start[i]; k < mem_node_start[i + 1]; k++) node_block[k] = mem[i].block;
    }
    IbmBlock *dblk;
    int4 *dtl;
    int *dnb, *dmb;
    if ((rc = cp_alloc(ctx, &dblk, blk.data(), blk.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dtl, tile_lo.data(), tile_lo.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dnb, node_block.data(), node_block.size()))) return rc;
    if ((rc = cp_alloc(ctx, &dmb, mem_block.data(), mem_block.size()))) return rc;
    g.blk = dblk;
    g.blk0 = blk[0];
    g.tile_lo = dtl;
    g.tile_cell_start = ctx->args.tile_cell_start;
    g.max_tile_cells = ctx->args.max_tile_cells;
    g.node_block = dnb;
    ctx->mem.mem_block = dmb;
    if (nblk > 1) {
        std::vector<int> cb(ctx->lay.n_cells);
        for (int i = 0; i < ctx->lay.n_cells; i++) cb[i] = cell_blk_caller[ctx->lay.perm[i]];
        int *dcb;
        if ((rc = cp_alloc(ctx, &dcb, cb.data(), cb.size()))) return rc;
        g.cell_block = dcb;
    }
    const int nc = ctx->lay.n_cells;
    std::vector<float4> cen(nc);
    for (int i = 0; i < nc; i++) {
        const int c = ctx->lay.perm[i];
        cen[i] = make_float4((float)ctx->geo.cc[3 * (size_t)c], (float)ctx->geo.cc[3 * (size_t)c + 1],
                             (float)ctx->geo.cc[3 * (size_t)c + 2], (float)ctx->geo.vol[c]);
    }
    // force density scale of a cell's spread sum: 1 / V (the physical-distance
    // kernels), 1 / h^3 of its box (the 2-point kernel); float as the kernels
    std::vector<float> ssc(nc);
    for (int i = 0; i < nc; i++) {
        if (ibm_kernel == LBFS_IBM_LINEAR2) {
            const IbmBlock &B = blk[nblk > 1 ? cell_blk_caller[ctx->lay.perm[i]] : 0];
            ssc[i] = B.inv_h * B.inv_h * B.inv_h;
        } else {
            ssc[i] = 1.0f / cen[i].w;
        }
    }
    float *dssc;
    if ((rc = cp_alloc(ctx, &dssc, ssc.data(), ssc.size()))) return rc;
    g.spread_sc = dssc;
    int *dinv, *tcount, *slots, *ov, *nov;
    float4 *dcen;
    float *wn, *sf;
    int4 *nbs;
    // node slots per fluid tile: twice its cells (a membrane sheet crossing a
    // tile twice at the finest resolution stays below that); per bin 128; the
    // rest overflow
    g.tile_cap = g.binned == IBM_SPREAD_GATHER ? std::max(64, 2 * ctx->args.max_tile_cells) : IBM_BIN_CAP;
    int *blist, *nblist;
    if ((rc = cp_alloc<int>(ctx, &blist, nullptr, g.n_tiles))) return rc;
    if ((rc = cp_alloc<int>(ctx, &nblist, nullptr, 1))) return rc;
    g.bin_list = blist;
    g.n_bin_list = nblist;
    if ((rc = cp_alloc(ctx, &dinv, ctx->lay.inv.data(), nc))) return rc;
    if ((rc = cp_alloc(ctx, &dcen, cen.data(), nc))) return rc;
    if ((rc = cp_alloc<int>(ctx, &tcount, nullptr, g.n_tiles))) return rc;
    if ((rc = cp_alloc<int>(ctx, &slots, nullptr, (size_t)g.tile_cap * g.n_tiles))) return rc;
    if ((rc = cp_alloc<int>(ctx, &ov, nullptr, nn))) return rc;
    if ((rc = cp_alloc<int>(ctx, &nov, nullptr, 1))) return rc;
    unsigned long long *dacc;
    if ((rc = cp_alloc<unsigned long long>(ctx, &dacc, nullptr, 3 * (size_t)nc))) return rc;
    g.acc = dacc;
    if ((rc = cp_alloc<float>(ctx, &sf, nullptr, 3 * (size_t)nc))) return rc;
    g.spread_f = sf;
    if ((rc = cp_alloc<float>(ctx, &wn, nullptr, 64 * (size_t)nn))) return rc;
    if ((rc = cp_alloc<int4>(ctx, &nbs, nullptr, nn))) return rc;
    g.inv = dinv; g.cen = dcen; g.tile_count = tcount; g.tile_slots = slots; g.overflow = ov; g.n_overflow = nov;
    g.wnode = wn; g.nbase = nbs;
    g.static_force = ctx->static_force;
    g.cell_force = ctx->cell_force;
    g.n_cells = nc;
    // fluid tile of each cell, and the per-tile flags of the apply pass
    {
        const int nt = ctx->lay.n_tiles;
        std::vector<int> ct(nc);
        for (int t = 0; t < nt; t++)
            for (int i = ctx->lay.tile_cell_start[t]; i < ctx->lay.tile_cell_start[t + 1]; i++) ct[i] = t;
        int *dct, *dnow, *dacc2, *dprev;
        if ((rc = cp_alloc(ctx, &dct, ct.data(), ct.size()))) return rc;
        if ((rc = cp_alloc<int>(ctx, &dnow, nullptr, nt))) return rc;
        if ((rc = cp_alloc<int>(ctx, &dacc2, nullptr, nt))) return rc;
        if ((rc = cp_alloc<int>(ctx, &dprev, nullptr, nt))) return rc;
        g.cell_tile = dct;
        g.tile_now = dnow;
        g.tile_acc = dacc2;
        g.tile_prev = dprev;
    }
    // internal viscosity field, only when some membrane has a ratio other than 1
    bool any_ratio = false;
    for (int i = 0; i < count; i++) any_ratio |= params[i].visc_ratio != 1.0f;
    if (any_ratio) {
        std::vector<float> ones(nc, 1.0f);
        float *vf;
        if ((rc = cp_alloc(ctx, &vf, ones.data(), nc))) return rc;
        ctx->visc_factor = vf;
        ctx->args.visc_factor = vf;
        ViscGrid &v = ctx->visc_grid;
        v = ViscGrid{};
        if ((rc = cp_alloc<float4>(ctx, &v.geo, nullptr, 3 * (size_t)count))) return rc;
        if ((rc = cp_alloc<int>(ctx, &v.tri_off, nullptr, (size_t)count * (VISC_G2 * VISC_G2 + 1)))) return rc;
        if ((rc = cp_alloc<int>(ctx, &v.tri_ent, nullptr, (size_t)VISC_TRI_CAP * tri.size()))) return rc;
        if ((rc = cp_alloc<int>(ctx, &v.ovf, nullptr, (size_t)count))) return rc;
        if ((rc = cp_alloc<int>(ctx, &v.node_off, nullptr, (size_t)count * (VISC_G3 * VISC_G3 * VISC_G3 + 1))))
            return rc;
        if ((rc = cp_alloc<int>(ctx, &v.node_ent, nullptr, nn))) return rc;
        if ((rc = cp_alloc<unsigned>(ctx, &v.qx, nullptr, nn))) return rc;
        v.on = ctx->visc_search;
    }
    double hmin = ctx->blocks[0].spacing;
    for (const HostBlock &hb : ctx->blocks) hmin = std::min(hmin, hb.spacing);
    ctx->delta_h = (float)std::min(0.5 * hmin, 0.1);
    ctx->t_step = 0;
    // stencil weights of the initial positions (the first interpolation uses them)
    HIPCHK(launch_ibm_weights(m, g, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipDeviceSynchronize());
    ctx->have_membranes = true;
    return LBFS_OK;
}

extern "C" int lbfs_get_membrane_nodes(lbfs_ctx *ctx, double *X, double *U, double *F)
{
    if (!ctx) return LBFS_E_INVALID_ARG;
    if (!ctx->have_membranes) return ctx->fail(LBFS_E_STATE, "no membranes attached");
    int rc = ctx->check_device_error();
    if (rc) return rc;
    const int nn = ctx->mem.n_nodes;
    std::vector<float4> hx(nn), hu(nn), hf(nn);
    std::vector<int> nm(nn);
    HIPCHK(hipMemcpyAsync(hx.data(), ctx->mem.X, sizeof(float4) * nn, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hu.data(), ctx->mem.U, sizeof(float4) * nn, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hf.data(), ctx->mem.F, sizeof(float4) * nn, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int mi = 0;
    for (int n = 0; n < nn; n++) {
        while (n >= ctx->mem_node_start[mi + 1]) mi++;
        const double *o = &ctx->mem_origin[3 * (size_t)mi];
        if (X) { X[3 * (size_t)n] = o[0] + hx[n].x; X[3 * (size_t)n + 1] = o[1] + hx[n].y; X[3 * (size_t)n + 2] = o[2] + hx[n].z; }
        if (U) { U[3 * (size_t)n] = hu[n].x; U[3 * (size_t)n + 1] = hu[n].y; U[3 * (size_t)n + 2] = hu[n].z; }
        if (F) { F[3 * (size_t)n] = hf[n].x; F[3 * (size_t)n + 1] = hf[n].y; F[3 * (size_t)n + 2] = hf[n].z; }
    }
    return LBFS_OK;
}

extern "C" int lbfs_get_membrane_globals(lbfs_ctx *ctx, double *out6)
{
    if (!ctx || !out6) return LBFS_E_INVALID_ARG;
    if (!ctx->have_membranes) return ctx->fail(LBFS_E_STATE, "no membranes attached");
    const int nm = ctx->mem.n_mem;
    std::vector<double> g(8 * (size_t)nm);
    HIPCHK(hipMemcpyAsync(g.data(), ctx->mem.glob, sizeof(double) * g.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < nm; i++)
        for (int q = 0; q < 6; q++) out6[6 * (size_t)i + q] = g[8 * (size_t)i + q];
    return LBFS_OK;
}

extern "C" int lbfs_get_membrane_energies(lbfs_ctx *ctx, double *out7)
{
    if (!ctx || !out7) return LBFS_E_INVALID_ARG;
    if (!ctx->have_membranes) return ctx->fail(LBFS_E_STATE, "no membranes attached");
    const int nm = ctx->mem.n_mem;
    HIPCHK(launch_mem_energy(ctx->mem, ctx->mem_energy, ctx->stream));
    HIPCHK(hipMemcpyAsync(out7, ctx->mem_energy, sizeof(double) * 7 * (size_t)nm, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}

static_assert(NOALLOC_MAX_MEMBRANES == LBFS_NOALLOC_MAX_MEMBRANES, "staging size of the ABI header");

extern "C" int lbfs_get_membrane_energies_noalloc(lbfs_ctx *ctx, double *out7)
{
    if (!ctx || !out7) return LBFS_E_INVALID_ARG;
    if (!ctx->have_membranes) return ctx->fail(LBFS_E_STATE, "no membranes attached");
    const int nm = ctx->mem.n_mem;
    if (nm > NOALLOC_MAX_MEMBRANES)
        return ctx->fail(LBFS_E_INVALID_ARG, "lbfs_get_membrane_energies_noalloc: %d membranes, at most %d", nm,
                         NOALLOC_MAX_MEMBRANES);
    if (!ctx->energy_scratch) return ctx->fail(LBFS_E_STATE, "no energy staging address on this device");
    // every entry read back is written by this launch first
    HIPCHK(launch_mem_energy(ctx->mem, ctx->energy_scratch, ctx->stream));
    HIPCHK(hipMemcpyAsync(out7, ctx->energy_scratch, sizeof(double) * 7 * (size_t)nm, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return LBFS_OK;
}

extern "C" int lbfs_mesh_icosphere(int frequency, double *nodes, int *tri)
{
    if (frequency < 1 || frequency > 64 || !nodes || !tri) return LBFS_E_INVALID_ARG;
    std::vector<double> X;
    std::vector<int> T;
    icosphere(frequency, X, T);
    std::copy(X.begin(), X.end(), nodes);
    std::copy(T.begin(), T.end(), tri);
    return LBFS_OK;
}

extern "C" int lbfs_mesh_evans_fung(int n_nodes, const double *unit_nodes, double d0, double *nodes)
{
    if (n_nodes <= 0 || !unit_nodes || !nodes || !(d0 > 0.0)) return LBFS_E_INVALID_ARG;
    std::vector<double> u(unit_nodes, unit_nodes + 3 * (size_t)n_nodes), out;
    evans_fung(u, d0, out);
    std::copy(out.begin(), out.end(), nodes);
    return LBFS_OK;
}

extern "C" int lbfs_mesh_springs(int n_nodes, int n_tri, const int *tri, int *springs, int *n_springs)
{
    if (n_nodes <= 0 || n_tri <= 0 || !tri || !springs || !n_springs) return LBFS_E_INVALID_ARG;
    std::vector<int> T(tri, tri + 3 * (size_t)n_tri), S;
    std::string err;
    int rc = build_springs(n_nodes, T, S, err);
    if (rc) return rc;
    std::copy(S.begin(), S.end(), springs);
    *n_springs = (int)S.size() / 4;
    return LBFS_OK;
}
