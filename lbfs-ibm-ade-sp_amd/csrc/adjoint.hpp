// Discrete adjoint of one coupled time step: argument blocks and launchers.
//
// The forward step n maps (st^n, X^n) to (st^n+1, X^n+1):
//   U^n    = I(X^n) u^n                  interpolation (cached weights of X^n)
//   X^n+1  = X^n + dt U^n                advection
//   F^n    = M(X^n+1, U^n; X^n)          membrane forces (the MAD term lags one step)
//   f^n    = static + S(X^n+1) F^n / V   spreading
//   st^n+1 = Phi(st^n, f^n)              gradients, face fluxes, update
// The reverse step takes the adjoints of (st^n+1, X^n+1) to those of
// (st^n, X^n) and adds the step's share to the parameter gradients.
#pragma once

#include <hip/hip_runtime.h>

#include "coupling.hpp"
#include "fluid_kernels.hpp"

namespace lbfs {

struct AdjFluidArgs {
    const float4 *lam_in;      // [n] adjoint of st^n+1 (drho, u, v, w)
    float4 *lam_out;           // [n] adjoint of st^n
    float4 *mu;                // [n] adjoint of the cell residual divided by the volume
    float4 *lcf;               // [n] adjoint of the cell force (x, y, z, 0)
    float4 *ls;                // [n] adjoint of st^n through the update
    float4 *ls_acc;            // [n] ls plus the face contributions
    float4 *lam_g;             // [3 n] adjoint of the gradients, in the gradient layout
    double *fpart;             // [3 blocks] running per-block sums of the cell-force adjoint
    double *gforce;            // [3] gradient of the uniform body force (sum of fpart)
    const float4 *rev_w;       // [6 n] SoA: for stencil slot k of cell i (neighbour c), the weight
                               // vector of c's stencil rows that point back to i (x, y, z)
    const unsigned long long *acc;  // [3 n] fixed-point adjoint of the interpolated velocity, or null
    unsigned long long *acc_rw;
    const unsigned int *umax;  // float bits of the largest |adjoint node velocity| (fixed-point scale)
    int n_fpart;
};

struct AdjMemArgs {
    const float4 *Xn;          // [nodes] positions at the start of the step (X^n)
    const double *glob_prev;   // [8 mem] globals at the start of the step (A(X^n) in slot 6)
    float4 *lam_X;             // [nodes] adjoint of X^n+1 in, of X^n out
    float4 *lam_U;             // [nodes] adjoint of U^n
    float4 *lam_F;             // [nodes] adjoint of F^n
    float *erec;               // [24 edges] adjoint of each edge's differentiated inputs
    float4 *grec;              // [4 edges] volume/area constraint share per edge slot (X^n+1)
    float *mrec;               // [12 edges] MAD share per edge coordinate (X^n)
    double *gmem;              // [4 mem] running sums: lambda_ks ks, local bending, global bending
    unsigned int *umax;
    const float4 *lcf;         // [n cells] adjoint of the cell force
    const float4 *st_n;        // [n cells] state at the start of the step
    unsigned long long *acc;   // [3 n cells] fixed-point accumulators
    int mad_lag;               // 0 on the first step: its MAD is a set-up constant
    const int *echunk;         // 3 per chunk: membrane, first edge, end edge (at most 256 edges)
    const int *echunk_start;   // [mem + 1] first chunk of each membrane
    double *epart;             // [6 chunks] partial sums
    int n_echunks;
};

// Fixed-point scale of the interpolation adjoint: values are held as
// v 2^(40 - e), e the binary exponent of the largest |component| (bits m of a
// non-negative float), so every node's share stays below 2^40 in magnitude.
__device__ __forceinline__ double adj_fix_scale(unsigned int m)
{
    const float x = __uint_as_float(m);
    if (!(x > 0.0f)) return 0.0;
    int e;
    frexpf(x, &e);
    return ldexp(1.0, 40 - e);
}
__device__ __forceinline__ double adj_fix_inv_scale(unsigned int m)
{
    const float x = __uint_as_float(m);
    if (!(x > 0.0f)) return 0.0;
    int e;
    frexpf(x, &e);
    return ldexp(1.0, e - 40);
}

hipError_t launch_adj_update(const FluidArgs &a, const AdjFluidArgs &j, const float4 *st, const float4 *res,
                             hipStream_t s);
// a.face_moments holds the forward moments of the step; tiles of at most 1024 cells
hipError_t launch_adj_flux(const FluidArgs &a, const AdjFluidArgs &j, const float4 *st, const float4 *grad,
                           int max_tile_cells, hipStream_t s);
// the record of the step before, for the update adjoint fused into the
// gradient transpose: its start state, residual and cell force
struct AdjPrevStep {
    const float4 *st, *res;
    const float *cf;
};
// prev: run the update adjoint of the step before in the same pass (null: not)
hipError_t launch_adj_grad(const FluidArgs &a, const AdjFluidArgs &j, const AdjPrevStep *prev, hipStream_t s);
// gforce = sum of the per-block partials, in a fixed order
hipError_t launch_adj_force_sum(const AdjFluidArgs &j, hipStream_t s);

hipError_t launch_adj_coupling(const MemArgs &m, const IbmArgs &g, const AdjMemArgs &j, hipStream_t s);

}  // namespace lbfs
