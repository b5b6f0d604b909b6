// Device-side argument block and launchers of the fluid kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "lbfs_abi.h"

namespace lbfs {

struct FluidArgs {
    int n_cells, n_bc, n_tiles, max_tile_faces;
    int max_tile_cells;            // largest tile
    int max_grad_slots;            // largest tile plus its halo (LDS of the fused step)
    const int *tile_halo_start, *halo_ids;  // per tile: cells across its border its faces reach
    const uint32_t *face_gslot;    // per tile-face: LDS gradient slots of owner | neighbour << 16
    int block_threads;             // workgroup size of the flux/update kernel (256 or 512)
    int lattice, flags;
    float rho0, inv_gamma, visc3_over_gamma, dt;
    float body_force[3];
    // per cell (internal tile order)
    const int *stencil;            // [6 n] SoA
    const float *gw;               // [18 n] SoA
    const float *inv_vol;          // [n]
    const int16_t *cell_slots;     // [6 n]
    const float *visc_factor;      // [n] or null
    const float *cell_force;       // [3 n] or null (IBM force density)
    float4 *face_moments;          // [tile faces] or null: S0, S1 of each face (kept for the adjoint)
    float4 *grad_out;              // [3 n] or null: the fused pass also writes the gradients of its
                                   // tile's cells (kept for the adjoint, the layout of the gradient pass)
    // tiles
    const int *tile_cell_start, *tile_face_start;
    const int *tile_bface_start;   // first boundary face of each tile's list
    // tile faces
    const float4 *fA, *fB, *fC;    // normal * area, dt | owner offset, owner | neighbour offset, neighbour
    // boundary conditions
    const int2 *bc_type;           // (rho type, vel type)
    const float4 *bc_val;          // (rho - rho0, u, v, w)
    // time stepping: 0 constant dt; LBFS_TS_LOCAL: dt_i = dt / (c.x + c.y / rho_i);
    // LBFS_TS_MIN: the minimum of those, reduced into *dt_min before the update
    int ts_mode;
    const float2 *cfl_coef;        // [n] or null
    unsigned int *dt_min;          // float bits of the global step (LBFS_TS_MIN)
};

hipError_t launch_gradients(const FluidArgs &a, const float4 *st, float4 *grad, hipStream_t s);
hipError_t launch_flux_update(const FluidArgs &a, const float4 *st, const float4 *grad, float4 *st_out,
                              float4 *res_out, int do_update, hipStream_t s);
// the whole fluid step in one pass per tile (gradients in LDS); equals
// launch_gradients + launch_flux_update bit for bit, gradients included when
// a.grad_out is set
hipError_t launch_fluid_fused(const FluidArgs &a, const float4 *st, float4 *st_out, float4 *res_out, int do_update,
                              hipStream_t s);
size_t fluid_fused_lds(const FluidArgs &a);
hipError_t launch_face_flux_all(const FluidArgs &a, const float4 *st, const float4 *grad, float4 *out,
                                int n_tile_faces, hipStream_t s);
// global minimum of the local steps into *a.dt_min (LBFS_TS_MIN)
hipError_t launch_dt_min(const FluidArgs &a, const float4 *st, hipStream_t s);
hipError_t launch_diag_partials(int n, float rho0, const float4 *st, const float4 *res, const float *force,
                                const float *inv_vol, double *part, hipStream_t s);

}  // namespace lbfs
