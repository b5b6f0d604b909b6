// Device-side argument blocks and launchers of the membrane and IBM kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "lbfs_abi.h"

namespace lbfs {

// scale of the membranes' 64-bit fixed-point sums (volume, area, MAD numerator)
constexpr double MEM_SUM_SCALE = 1073741824.0;  // 2^30

// spreading paths; all give the same bits
enum {
    IBM_SPREAD_NODES = 0,   // one wavefront per node into per-cell accumulators (global atomics)
    IBM_SPREAD_GATHER = 1,  // nodes listed per fluid tile; each tile gathers the nodes whose
                            // stencils reach it and sums their forces in LDS (no global atomics)
    IBM_SPREAD_BINS = 2,    // nodes listed per 4x4x4 bin; each occupied bin spreads its nodes into an
                            // LDS window (the bin and S - 1 cells beyond) added to the accumulators
};
constexpr int IBM_BIN_CAP = 128;  // node slots per bin of the bin scattering

struct MemParamDev {
    float spring_c;           // WLC constant making F(L0) = 0
    float volume_modulus, local_area_modulus, global_area_modulus;
    float global_bending, local_bending, viscosity;
    int n_mad;                // edges summed for MAD
    float visc_ratio;         // internal viscosity ratio (1 = off)
    float wlc_x0;             // 1 / wlc_ratio: rest length over contour length
    int spring_law;           // LBFS_SPRING_WLC or LBFS_SPRING_WLC_POW
};

struct MemArgs {
    int n_mem, n_nodes, n_tri, n_edges;
    float4 *X;                // node position relative to the membrane origin
    float4 *U;                // node velocity (interpolated)
    float4 *F;                // node force
    const int *node_mem;
    const float4 *origin;     // per membrane
    const float4 *node_origin;  // per node: its membrane's origin (one load where origin[node_mem] takes two)
    const int4 *tri;          // a, b, c, membrane
    const int *mem_tri_start, *mem_edge_start, *mem_node_start;
    const int4 *edges;        // a, b, c, d (edge a-c)
    const int *edge_mem;
    const float4 *edge_const; // contour length, spring constant, curvature_0, k_p / k_s (WLC-POW; 0 with WLC)
    const float2 *edge_area0; // |N1|, |N2| of the reference state
    float4 *edge_state;       // curvature, |N1|, |N2|, 0 (feeds the MAD of the next step)
    float4 *edge_force;       // 4 per edge: forces on a, b, c, d
    const int *inc_start, *inc;  // node -> (edge * 4 + slot) incidences
    const MemParamDev *params;
    double *glob;             // 8 per membrane: V, A, MAD (of the last force call), V0, A0, MAD0,
                              // A_prev, MAD of the next force call
    unsigned long long *mem_acc;  // 4 per membrane: running sums of V, A and the MAD numerator
                              // (fixed point, MEM_SUM_SCALE), closed out by the cell update
    float4 *bbox;             // 2 per membrane: lower and upper corner of the nodes
    const int *mem_block;     // box of each membrane
    int *err;                 // the context's device error word (ERR_FIX_* bits, fixed_point.hpp)
};

// one logical box of cells (an ensemble member, or the whole mesh)
struct IbmBlock {
    int n[3], per[3];
    int tw[3];                // node-list tile size (cells per tile along each axis; the last tile may be smaller)
    int nt[3];                // node-list tiles per axis
    int tile0;                // first node-list tile id of the box (tiles x fastest, then y, then z)
    int cell0;                // first caller cell id of the box (cells i + j nx + k nx ny after it)
    int ftw[3], fnt[3];       // fluid tile size and tiles per axis (the device cell order)
    int fsh[3];               // log2 of the fluid tile size
    int tsh[3];               // log2 of the node-list tile size, -1 if not a power of two
    int regular;              // 1: every fluid tile of the box is whole and a power of two along each
                              // axis, so internal ids follow from i, j, k by shifts
    int icell0;               // internal id of the box's first cell (its first fluid tile's first cell)
    int ax_off[3];            // 4-point Lagrange kernel: offset of the axis data in IbmArgs::axis
    float origin[3], L[3], h, inv_h;
    float period_step[3];     // L / n: shift of a cell centroid per period of a wrapped stencil index
};

struct IbmArgs {
    int kernel;               // LBFS_IBM_*
    int n_cells;
    int n_blocks, n_tiles;
    int binned;               // IBM_SPREAD_*: how the forces are spread
    const IbmBlock *blk;      // [n_blocks]
    IbmBlock blk0;            // the box of a single-box mesh, by value (no per-node lookup)
    const int4 *tile_lo;      // [n_tiles] first cell of each node-list tile in its box (x, y, z) and the box
    const int *tile_cell_start;  // [n_tiles + 1] internal cell ids of each tile (tile gathering: the
                              // node-list tiles are the fluid tiles)
    int max_tile_cells;
    int *bin_list, *n_bin_list;  // bin scattering: the tiles holding nodes this step (the count is
                              // cleared by the apply pass)
    const int *node_block;    // [n_nodes] box of each node
    const int *cell_block;    // [n_cells] box of each cell (internal order), null for one box
    const double *axis;       // per box and axis: n + 1 vertex plane coordinates, then n cell-line centroids
    float dt;
    const int *inv;           // caller cell id -> internal id
    const float4 *cen;        // internal order: centroid xyz, volume
    int *tile_count;          // [n_tiles] nodes listed in each tile this step (cleared by the apply pass)
    int *tile_slots;          // tile_cap node ids per tile
    int tile_cap;
    int *overflow, *n_overflow;  // nodes of full tiles, gathered by every tile (the count is cleared by the
                              // apply pass)
    float *wnode;             // 64 per node: normalised stencil weights of its current position
                              // (lane order i + S j + S^2 k)
    int4 *nbase;              // per node: stencil base cell (logical) of the current position and its tile
    const float *spread_sc;   // [n_cells] force density per unit spread sum: 1/V, or 1/h^3 (2-point)
    float *spread_f;          // [3 n_cells] force density of the tiles the tile spreading wrote
    unsigned long long *acc;  // 3 per cell (internal order): fixed-point spread force of the per-node
                              // spreading (and scratch of the adjoint)
    const float *static_force;
    float *cell_force;
    const int *cell_tile;     // [n_cells] fluid tile of each cell
    int *tile_now;            // [n_tiles] 1: the tile spreading wrote a nonzero force density this step
    int *tile_acc;            // [n_tiles] 1: the per-node spreading wrote into the tile's accumulators
    int *tile_prev;           // [n_tiles] 1: the cell force output holds spread force there
    int *err;                 // the context's device error word (ERR_FIX_* bits, fixed_point.hpp)
};

// The end of the coupled step: adds the spread force density to the updated
// state (the face pass updated it with the static force), writes the cell
// force output and clears the per-node accumulators, tile by tile: only the
// tiles the spreading wrote into (all = 0; the tiles of the previous step get
// their output reset to the static force), or every tile (all = 1, for
// per-step force records).  Also closes out the membrane sums of the step and
// clears the tile node counts and the overflow count for the next step.
// the bin scattering gathers the node forces from the edge records itself
// (every node is in one bin slot or on the overflow list), so the node-force
// pass does not run before it
inline bool spread_gathers_forces(const IbmArgs &g) { return g.binned == IBM_SPREAD_BINS; }

struct ApplyArgs {
    const int *tile_cell_start;
    int n_tiles;
    float dt, rho0;
    float4 *st;               // the updated state
    float *cell_force_out;    // [3 n]
    int all;
    unsigned long long *mem_acc;  // [4 n_mem] membrane sums
    double *glob;             // [8 n_mem]
    int n_mem;
};

// threads of the tile spreading, also the candidate nodes it tests at a time
constexpr int SPREAD_NT = 512;
// LDS of the tile spreading: the staged candidate nodes (base 16 B, force
// 16 B, id 4 B), then three 64-bit sums per tile cell
inline size_t spread_tile_lds(int max_tile_cells) { return (size_t)SPREAD_NT * 36 + (size_t)max_tile_cells * 24; }

hipError_t launch_ibm_interp(const MemArgs &m, const IbmArgs &g, const float4 *st, hipStream_t s);
// weights and tiles of the current node positions (after attaching membranes
// or restoring them)
hipError_t launch_ibm_weights(const MemArgs &m, const IbmArgs &g, hipStream_t s);
hipError_t launch_mem_volume(const MemArgs &m, hipStream_t s);
hipError_t launch_mem_edges(const MemArgs &m, hipStream_t s);
hipError_t launch_mem_nodes(const MemArgs &m, hipStream_t s);
// spreading: per fluid tile into spread_f, or per node into the fixed-point
// cell accumulators; launch_ibm_apply adds the force densities to the state
hipError_t launch_ibm_spread(const MemArgs &m, const IbmArgs &g, hipStream_t s);
hipError_t launch_ibm_apply(const IbmArgs &g, const ApplyArgs &p, hipStream_t s);
// Binned search structure of the internal-viscosity field, rebuilt with it:
// per membrane its triangles by the y-z bins their projections overlap (a ray
// along +x from a cell can only cross those of the cell's bin) and its nodes
// by 3-D bins (the nearest-vertex distance only matters within 2 delta_h).
constexpr int VISC_G2 = 16;       // y-z bins per axis of a membrane's triangle grid
constexpr int VISC_G3 = 16;       // bins per axis of its node grid
constexpr int VISC_TRI_CAP = 8;   // triangle-bin entries per triangle; a membrane needing more is searched whole
struct ViscGrid {
    int on;                       // 0: every triangle of each membrane (the definition; the same bits)
    float4 *geo;                  // 3 per membrane: (y0, z0, bins/len y, bins/len z), (x0, y0, z0, -), bins/len xyz
    int *tri_off;                 // [n_mem (VISC_G2^2 + 1)] entry offsets of each y-z bin
    int *tri_ent;                 // [VISC_TRI_CAP n_tri] triangle ids (a membrane's from VISC_TRI_CAP tri_start)
    int *ovf;                     // [n_mem] 1: the membrane's entries did not fit (searched whole)
    int *node_off;                // [n_mem (VISC_G3^3 + 1)]
    int *node_ent;                // [n_nodes] node ids (a membrane's from its node_start)
    unsigned *qx;                 // [n_nodes] order key of the largest smallest-x of the node's triangles
};
// internal viscosity field (immersed_boundary_method.cu:389-527)
hipError_t launch_internal_visc(const MemArgs &m, const IbmArgs &g, const ViscGrid &v, float delta_h,
                                float *visc_factor, hipStream_t s);
// output-step energies, 7 doubles per membrane (Solver_gpu.cu:1060-1139)
hipError_t launch_mem_energy(const MemArgs &m, double *out, hipStream_t s);
// device address of the static energy staging of the current device (taken
// once per device with that device current); NOALLOC_MAX_MEMBRANES membranes
constexpr int NOALLOC_MAX_MEMBRANES = 8192;
hipError_t energy_scratch_setup(int device, double **out);

}  // namespace lbfs
