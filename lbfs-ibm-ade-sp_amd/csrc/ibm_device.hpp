// Device helpers shared by the IBM/membrane kernels and their adjoints:
// small 3-vectors, the IBM stencil (base cell, periodic wrap, weights of the
// 2-point, 4-point cosine and 4-point Lagrange kernels) and wavefront sums.
#pragma once

#include <hip/hip_runtime.h>

#include "coupling.hpp"
#include "fixed_point.hpp"

namespace lbfs {
namespace {

constexpr double FIX_SCALE = 1099511627776.0;  // 2^40
constexpr float PI_F = 3.14159265358979f;

// 64-bit fixed-point term of v at `scale`, rounded to nearest; a term that is
// not finite or beyond the range fix_in_range allows is replaced by 0 and
// flagged in the error word, so an unstable run reports an error instead of
// wrapping the integer sums
__device__ __forceinline__ unsigned long long fix_term(double v, double scale, int *err, int bit)
{
    const double s = v * scale;
    if (fix_in_range(s)) return (unsigned long long)__double2ll_rn(s);
    atomicOr(err, bit);
    return 0ull;
}

// the same term for a float value: v * 2^40 is exact in float, so the range
// test and the rounding give the bits fix_term gives for (double)v
__device__ __forceinline__ unsigned long long fix_term_f(float v, int *err, int bit)
{
    const float s = v * 1099511627776.0f;  // 2^40
    if (fabsf(s) < 4503599627370496.0f)    // 2^52 (false for NaN)
        return (unsigned long long)(long long)rintf(s);
    atomicOr(err, bit);
    return 0ull;
}

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ f3 ld3(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float cmp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
// IBM stencil helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int stencil_size(const IbmArgs &g) { return g.kernel == LBFS_IBM_LINEAR2 ? 2 : 4; }

// nearest vertex plane, ties to the larger index (find_index_closest,
// immersed_boundary_method.cu:588-636), by bisection
__device__ __forceinline__ int closest_vertex(const double *v, int n, double t)
{
    if (t <= v[0]) return 0;
    if (t >= v[n - 1]) return n - 1;
    int lo = 0, hi = n - 1;  // v[lo] < t <= v[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < t) lo = mid; else hi = mid;
    }
    return (t - v[lo] >= v[hi] - t) ? hi : lo;
}

// cell-line coordinate at unwrapped index c (periodic image or wall continuation)
__device__ __forceinline__ double line_coord(const IbmArgs &g, const IbmBlock &B, int a, int c)
{
    const int n = B.n[a];
    const double *l = g.axis + B.ax_off[a] + n + 1;
    if (c >= 0 && c < n) return l[c];
    if (B.per[a]) {
        const int k = c >= 0 ? c / n : -((-c + n - 1) / n);
        return l[c - k * n] + (double)k * (double)B.L[a];
    }
    const double d0 = n > 1 ? l[1] - l[0] : (double)B.h, d1 = n > 1 ? l[n - 1] - l[n - 2] : (double)B.h;
    return c < 0 ? l[0] + c * d0 : l[n - 1] + (c - (n - 1)) * d1;
}

// cubic Lagrange basis i of the 4 line points from `base` at x (the first
// column of the inverse Vandermonde matrix, vandemonde_inverse :919-941)
__device__ __forceinline__ double lagrange4(const IbmArgs &g, const IbmBlock &B, int a, int base, int i, double x)
{
    double p[4];
#pragma unroll
    for (int k = 0; k < 4; k++) p[k] = line_coord(g, B, a, base + k);
    double v = 1.0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (k != i) v *= (x - p[k]) / (p[i] - p[k]);
    return v;
}

__device__ __forceinline__ void node_base(const IbmArgs &g, const IbmBlock &B, const float P[3], float xr[3], int base[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        xr[a] = (P[a] - B.origin[a]) * B.inv_h;
        if (g.kernel == LBFS_IBM_LAGRANGE4)
            base[a] = closest_vertex(g.axis + B.ax_off[a], B.n[a] + 1, (double)P[a]) - 2;
        else
            base[a] = g.kernel == LBFS_IBM_LINEAR2 ? (int)roundf(xr[a]) - 1 : (int)floorf(xr[a] - 1.5f);
    }
}

// wrap a logical index into the grid; false if it lies outside a wall
__device__ __forceinline__ bool wrap_index(const IbmBlock &B, int a, int i, int &w)
{
    const int n = B.n[a];
    if (i >= 0 && i < n) { w = i; return true; }
    if (!B.per[a]) return false;
    // one period away in almost every case; the remainder only beyond that
    int v = i < 0 ? i + n : i - n;
    if (__builtin_expect(v < 0 || v >= n, 0)) v = ((i % n) + n) % n;
    w = v;
    return true;
}

// unnormalised weight of the cell with unwrapped index idx / wrapped index w
// phi(r) = (1 + cos(pi r / 2)) / 4 for |r| < 2, written as (1 - sin^2(pi r / 4)) / 2
// with sin from its Taylor series to x^11 (|x| <= pi/2: truncation below 6e-8)
__device__ __forceinline__ float cos4_phi(float r)
{
    const float x = r * (PI_F * 0.25f), x2 = x * x;
    float p = -2.5052108385e-8f;                 // -1/11!
    p = fmaf(p, x2, 2.7557319224e-6f);           //  1/9!
    p = fmaf(p, x2, -1.9841269841e-4f);          // -1/7!
    p = fmaf(p, x2, 8.3333333333e-3f);           //  1/5!
    p = fmaf(p, x2, -1.6666666667e-1f);          // -1/3!
    const float sn = fmaf(p * x2, x, x);
    return 0.5f * fmaf(-sn, sn, 1.0f);
}

// d phi / d r of the same polynomial: phi = (1 - sn^2) / 2, sn(x) as above,
// so phi' = -sn sn'(x) pi / 4 with sn' the term-by-term derivative
__device__ __forceinline__ float cos4_dphi(float r)
{
    const float x = r * (PI_F * 0.25f), x2 = x * x;
    float p = -2.5052108385e-8f;
    p = fmaf(p, x2, 2.7557319224e-6f);
    p = fmaf(p, x2, -1.9841269841e-4f);
    p = fmaf(p, x2, 8.3333333333e-3f);
    p = fmaf(p, x2, -1.6666666667e-1f);
    const float sn = fmaf(p * x2, x, x);
    float d = 11.0f * -2.5052108385e-8f;
    d = fmaf(d, x2, 9.0f * 2.7557319224e-6f);
    d = fmaf(d, x2, 7.0f * -1.9841269841e-4f);
    d = fmaf(d, x2, 5.0f * 8.3333333333e-3f);
    d = fmaf(d, x2, 3.0f * -1.6666666667e-1f);
    const float dsn = fmaf(d, x2, 1.0f);
    return -sn * dsn * (PI_F * 0.25f);
}

// 2-point kernel factor of stencil index idx along one axis (xr: the node's
// grid coordinate)
__device__ __forceinline__ float linear2_factor(int idx, float xr) { return 1.0f - fabsf((float)idx + 0.5f - xr); }

// 4-point cosine weight of a cell with centroid c for a node at P; d: the
// periods between the stencil index and the wrapped cell along each axis
__device__ __forceinline__ float cos4_cell(const IbmBlock &B, const float4 c, const int d[3], const float P[3])
{
    const float cc[3] = {c.x, c.y, c.z};
    float wt = 1.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float xc = fmaf((float)d[a], B.period_step[a], cc[a]);
        const float r = (xc - P[a]) * B.inv_h;
        wt *= fabsf(r) < 2.0f ? cos4_phi(r) : 0.0f;
    }
    return wt;
}

// unnormalised weight of the stencil cell with unwrapped index idx, wrapped
// index w and internal id `internal`, for a node at P (grid coordinate xr)
__device__ __forceinline__ float cell_weight(const IbmArgs &g, const IbmBlock &B, const int idx[3], const int w[3],
                                             int internal, const float P[3], const float xr[3])
{
    float wt = 1.0f;
    if (g.kernel == LBFS_IBM_LINEAR2) {
#pragma unroll
        for (int a = 0; a < 3; a++) wt *= linear2_factor(idx[a], xr[a]);
    } else {
        const int d[3] = {idx[0] - w[0], idx[1] - w[1], idx[2] - w[2]};
        wt = cos4_cell(B, g.cen[internal], d, P);
    }
    return wt;
}

// unnormalised 4-point cosine weight of a cell and its gradient with respect
// to the node position P (r = (x_c - P) / h, so dr/dP = -1/h)
__device__ __forceinline__ float cos4_weight_grad(const IbmBlock &B, const float4 c, const int idx[3], const int w[3],
                                                  const float P[3], float dw[3])
{
    const float cc[3] = {c.x, c.y, c.z};
    float ph[3], dph[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float xc = fmaf((float)(idx[a] - w[a]), B.period_step[a], cc[a]);
        const float r = (xc - P[a]) * B.inv_h;
        const bool in = fabsf(r) < 2.0f;
        ph[a] = in ? cos4_phi(r) : 0.0f;
        dph[a] = in ? -cos4_dphi(r) * B.inv_h : 0.0f;
    }
    dw[0] = dph[0] * ph[1] * ph[2];
    dw[1] = ph[0] * dph[1] * ph[2];
    dw[2] = ph[0] * ph[1] * dph[2];
    float wt = 1.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) wt *= ph[a];
    return wt;
}

// internal id of the cell with (wrapped) index w: from the tile layout in a
// box of whole tiles (tiles x fastest, cells x fastest within a tile), else
// through the id map
// the box of node nd, and box b (a single box comes by value with the arguments)
__device__ __forceinline__ IbmBlock node_block(const IbmArgs &g, int nd)
{
    return g.n_blocks == 1 ? g.blk0 : g.blk[g.node_block[nd]];
}
__device__ __forceinline__ IbmBlock box(const IbmArgs &g, int b) { return g.n_blocks == 1 ? g.blk0 : g.blk[b]; }

__device__ __forceinline__ int cell_internal(const IbmArgs &g, const IbmBlock &B, const int w[3])
{
    if (B.regular) {
        const int t0 = w[0] >> B.fsh[0], t1 = w[1] >> B.fsh[1], t2 = w[2] >> B.fsh[2];
        const int l = (w[0] & (B.ftw[0] - 1)) + ((w[1] & (B.ftw[1] - 1)) << B.fsh[0]) +
                      ((w[2] & (B.ftw[2] - 1)) << (B.fsh[0] + B.fsh[1]));
        return B.icell0 + ((t0 + B.fnt[0] * (t1 + B.fnt[1] * t2)) << (B.fsh[0] + B.fsh[1] + B.fsh[2])) + l;
    }
    return g.inv[B.cell0 + w[0] + B.n[0] * (w[1] + B.n[1] * w[2])];
}

// fluid tile coordinate along axis a of a node whose stencil base index is
// `base` (wrapped on periodic axes, clamped into the box on the others)
__device__ __forceinline__ int tile_coord(const IbmBlock &B, int a, int base)
{
    const int n = B.n[a];
    int b = base;
    if (B.per[a]) b = ((b % n) + n) % n;
    else b = b < 0 ? 0 : (b >= n ? n - 1 : b);
    return B.tsh[a] >= 0 ? b >> B.tsh[a] : b / B.tw[a];
}

// ---------------------------------------------------------------------------
// wavefront helpers: one wavefront per node, one lane per stencil cell (64
// for the 4-point kernels, 8 for the 2-point one); sums are butterfly
// reductions, so their order is fixed
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void lane_offset(int S, int lane, int off[3])
{
    if (S == 4) { off[0] = lane & 3; off[1] = (lane >> 2) & 3; off[2] = lane >> 4; }
    else { off[0] = lane & 1; off[1] = (lane >> 1) & 1; off[2] = lane >> 2; }
}

}  // namespace
}  // namespace lbfs
