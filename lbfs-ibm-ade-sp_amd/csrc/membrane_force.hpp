// The spring-network force of one edge (spring_network.cu:613-971), written
// once over a scalar type T: `float` for the force kernel, `Dual` (a value
// with one directional derivative) for the adjoint, which differentiates the
// very arithmetic of the forward kernel direction by direction.
#pragma once

#include <hip/hip_runtime.h>

namespace lbfs {
namespace memf {

// value and one directional derivative (forward-mode differentiation)
struct Dual {
    float v, d;
    __device__ __forceinline__ Dual() : v(0.0f), d(0.0f) {}
    __device__ __forceinline__ Dual(float x) : v(x), d(0.0f) {}
    __device__ __forceinline__ Dual(float x, float dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const float q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, float b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(float a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, float b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(float a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, float b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(float a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, float b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(float a, Dual b)
{
    const float q = a / b.v;
    return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
__device__ __forceinline__ Dual &operator*=(Dual &a, float b) { return a = a * b; }

__device__ __forceinline__ float val(float x) { return x; }
__device__ __forceinline__ float val(Dual x) { return x.v; }
__device__ __forceinline__ float tsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ Dual tsqrt(Dual x)
{
    const float s = sqrtf(x.v);
    return {s, 0.5f * x.d / s};
}
// acos of a cosine clamped to [-1, 1]; at the clamp the derivative is zero
__device__ __forceinline__ float tacos_clamped(float c) { return acosf(fminf(fmaxf(c, -1.0f), 1.0f)); }
__device__ __forceinline__ Dual tacos_clamped(Dual c)
{
    if (c.v >= 1.0f) return {acosf(1.0f), 0.0f};
    if (c.v <= -1.0f) return {acosf(-1.0f), 0.0f};
    return {acosf(c.v), -c.d / sqrtf(fmaxf(1.0f - c.v * c.v, 1e-12f))};
}

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
__device__ __forceinline__ V3<T> v3(T x, T y, T z) { return {x, y, z}; }
template <class T>
__device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
__device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T, class S>
__device__ __forceinline__ V3<T> operator*(V3<T> a, S s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
__device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
__device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
__device__ __forceinline__ T cmp(V3<T> a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// derivatives of N = (q - p) x (r - p) and of N . centroid for the nine
// coordinates p.xyz q.xyz r.xyz: dN = e_k x (next - prev) (spring_network.cu:1348-1412)
template <class T>
__device__ __forceinline__ void tri_derivs(V3<T> p, V3<T> q, V3<T> r, V3<T> N, V3<T> dN[9], T dV[9])
{
    const V3<T> v[3] = {p, q, r};
    const V3<T> cen = (p + q + r) * (1.0f / 3.0f);
#pragma unroll
    for (int vt = 0; vt < 3; vt++) {
        const V3<T> dd = v[(vt + 1) % 3] - v[(vt + 2) % 3];
        const T z(0.0f);
        const V3<T> dk[3] = {v3<T>(z, -dd.z, dd.y), v3<T>(dd.z, z, -dd.x), v3<T>(-dd.y, dd.x, z)};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            dN[3 * vt + k] = dk[k];
            dV[3 * vt + k] = cmp(N, k) * (1.0f / 3.0f) + dot(dk[k], cen);
        }
    }
}

// Per-edge shares of the area and volume gradients: slots A(0-2) B(3-5)
// C(6-8) D(9-11), triangle 1 (A, B, C), triangle 2 (C, D, A).  Every
// triangle belongs to three edges, so the shares of all edges add up to
// dA/dX and dV/dX of the closed surface.
template <class T>
struct EdgeGeometry {
    V3<T> l, N1, N2, dN1[9], dN2[9];
    T length, n1, n2, in1, in2;
    T dn1[9], dn2[9], d_area[12], d_vol[12];
};

template <class T>
__device__ __forceinline__ void edge_geometry(V3<T> A, V3<T> B, V3<T> C, V3<T> D, EdgeGeometry<T> &g)
{
    g.l = C - A;
    g.length = tsqrt(dot(g.l, g.l));
    g.N1 = cross(B - A, C - A);
    g.N2 = cross(C - A, D - A);
    g.n1 = tsqrt(dot(g.N1, g.N1));
    g.n2 = tsqrt(dot(g.N2, g.N2));
    T dV1[9], dV2[9];
    tri_derivs(A, B, C, g.N1, g.dN1, dV1);
    tri_derivs(C, D, A, g.N2, g.dN2, dV2);
    g.in1 = 1.0f / g.n1;
    g.in2 = 1.0f / g.n2;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        g.dn1[j] = dot(g.dN1[j], g.N1) * g.in1;
        g.dn2[j] = dot(g.dN2[j], g.N2) * g.in2;
    }
#pragma unroll
    for (int j = 0; j < 9; j++) { g.d_area[j] = g.dn1[j]; g.d_vol[j] = dV1[j]; }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        g.d_area[9 + j] = g.dn2[3 + j];
        g.d_vol[9 + j] = dV2[3 + j];
        g.d_area[6 + j] += g.dn2[j];
        g.d_vol[6 + j] += dV2[j];
        g.d_area[j] += g.dn2[6 + j];
        g.d_vol[j] += dV2[6 + j];
    }
#pragma unroll
    for (int j = 0; j < 12; j++) { g.d_area[j] *= (1.0f / 6.0f); g.d_vol[j] *= (1.0f / 18.0f); }
}

// inputs of one edge force evaluation that the adjoint differentiates
template <class T>
struct EdgeIn {
    V3<T> A, B, C, D;                 // node positions relative to the membrane origin
    V3<T> Ua, Uc;                     // velocities of the edge's end nodes (membrane viscosity)
    T vol_coef, area_coef, mad_coef;  // global volume / area / MAD coefficients
    T ks, lb, gb;                     // spring constant, local and global bending moduli
};

// per-edge constants
struct EdgeConst {
    float lmax, kappa0;               // contour length, reference curvature
    float a01, a02;                   // reference |N1|, |N2|
    float la_mod, visc, spring_c;     // local area modulus, membrane viscosity, WLC constant
    float kp_ks;                      // WLC-POW: k_p / k_s (the POW term scales with the spring constant)
    int pow_law;                      // 1: WLC-POW spring, 0: WLC
};

template <class T>
struct EdgeOut {
    V3<T> F[4];                       // forces on A, B, C, D
    T kappa, n1, n2;                  // curvature and |N1|, |N2| (the MAD of the next step)
};

template <class T>
__device__ __forceinline__ void edge_force(const EdgeIn<T> &in, const EdgeConst &k, EdgeOut<T> &out)
{
    const V3<T> A = in.A, B = in.B, C = in.C, D = in.D;
    EdgeGeometry<T> g;
    edge_geometry(A, B, C, D, g);
    const T in1 = g.in1, in2 = g.in2, n1 = g.n1, n2 = g.n2, length = g.length;
    T cs = dot(g.N1, g.N2) * in1 * in2;
    const V3<T> c1 = (A + B + C) * (1.0f / 3.0f), c2 = (A + C + D) * (1.0f / 3.0f);
    const T dir = dot(g.N1, c2 - c1);
    const T angle = tacos_clamped(cs) * (float)((val(dir) < 0.0f) - (val(dir) > 0.0f));
    const T il = 1.0f / length;
    const V3<T> dL = g.l * il;
    T d_ang[12];
    {
        const V3<T> h1 = cross(g.N1 * in1, dL), h2 = cross(dL, g.N2 * in2);
        T a2[9];
#pragma unroll
        for (int j = 0; j < 9; j++) {
            d_ang[j] = -in1 * dot(h1, g.dN1[j]);
            a2[j] = -in2 * dot(h2, g.dN2[j]);
        }
        d_ang[9] = a2[3]; d_ang[10] = a2[4]; d_ang[11] = a2[5];
        d_ang[6] += a2[0]; d_ang[7] += a2[1]; d_ang[8] += a2[2];
        d_ang[0] += a2[6]; d_ang[1] += a2[7]; d_ang[2] += a2[8];
    }
    T d_curve[12];
#pragma unroll
    for (int j = 0; j < 12; j++) d_curve[j] = d_ang[j] * length;
    d_curve[0] -= dL.x * angle; d_curve[1] -= dL.y * angle; d_curve[2] -= dL.z * angle;
    d_curve[6] += dL.x * angle; d_curve[7] += dL.y * angle; d_curve[8] += dL.z * angle;
    const T z(0.0f);
    V3<T> Fo[4] = {v3(z, z, z), v3(z, z, z), v3(z, z, z), v3(z, z, z)};
    // local and global area: one opposite node per triangle
    const T ca = -k.la_mod * (n1 / k.a01 - 1.0f) + in.area_coef;
    const T cb = -k.la_mod * (n2 / k.a02 - 1.0f) + in.area_coef;
    Fo[1] = Fo[1] + v3(g.dn1[3], g.dn1[4], g.dn1[5]) * ca;
    Fo[3] = Fo[3] + v3(g.dn2[3], g.dn2[4], g.dn2[5]) * cb;
    // volume
#pragma unroll
    for (int j = 0; j < 4; j++) Fo[j] = Fo[j] + v3(g.d_vol[3 * j], g.d_vol[3 * j + 1], g.d_vol[3 * j + 2]) * in.vol_coef;
    // membrane viscosity
    {
        const V3<T> vca = in.Uc - in.Ua;
        const V3<T> vf = dL * (-dot(dL, vca) * 4.0f / 1.7320508075688772f * k.visc);
        Fo[2] = Fo[2] + vf;
        Fo[0] = Fo[0] - vf;
    }
    // bending
    const T kappa = angle * length / ((n1 + n2) * (1.0f / 6.0f));
    const T dk0 = k.kappa0 - kappa;
    const T lbc = -0.5f * in.lb * dk0 * dk0;
    const T ccf = -in.gb * (kappa - k.kappa0) + in.mad_coef;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        Fo[j] = Fo[j] + v3(g.d_area[3 * j], g.d_area[3 * j + 1], g.d_area[3 * j + 2]) * lbc;
        Fo[j] = Fo[j] + v3(d_curve[3 * j], d_curve[3 * j + 1], d_curve[3 * j + 2]) * ccf;
    }
    // spring
    {
        const T x = length / k.lmax;
        const T om = 1.0f - x;
        // worm-like chain (spring_network.cu:932-936), or WLC-POW
        // -(k_s (1/(1-x)^2 - 1 + 4x) + k_p / L^2) (spring_network.cu:221-223)
        const T fc = k.pow_law ? -in.ks * (1.0f / (om * om) - 1.0f + 4.0f * x + k.kp_ks * (il * il))
                               : -in.ks * (0.25f / (om * om) - 0.25f + x - k.spring_c);
        Fo[2] = Fo[2] + dL * fc;
        Fo[0] = Fo[0] - dL * fc;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) out.F[j] = Fo[j];
    out.kappa = kappa;
    out.n1 = n1;
    out.n2 = n2;
}

}  // namespace memf
}  // namespace lbfs
