// Device helpers shared by the fluid kernels and their adjoints: gradient
// loads, compile-time lattice direction sums and the equilibrium deviation of
// one lattice point.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "fluid_kernels.hpp"
#include "lattice.hpp"
#include "xcd.hpp"

namespace lbfs {

__device__ __forceinline__ void load_grad(const float4 *__restrict__ grad, int i, float g[12])
{
    const float4 a = grad[3 * (size_t)i], b = grad[3 * (size_t)i + 1], c = grad[3 * (size_t)i + 2];
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w;
    g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
    g[8] = c.x; g[9] = c.y; g[10] = c.z; g[11] = c.w;
}

// one tile-face record: unit normal and area from the area-weighted normal
struct FaceRec {
    float nx, ny, nz, A, dt;
    float4 d1, d2;  // owner and neighbour image centroid - face centroid (w: the ids' bits)
    int o, nb;
};
__device__ __forceinline__ FaceRec load_face(const FluidArgs &a, int tf)
{
    const float4 fa = a.fA[tf];
    FaceRec r;
    r.d1 = a.fB[tf];
    r.d2 = a.fC[tf];
    r.A = sqrtf(fa.x * fa.x + fa.y * fa.y + fa.z * fa.z);
    const float iA = r.A > 0.0f ? 1.0f / r.A : 0.0f;
    r.nx = fa.x * iA;
    r.ny = fa.y * iA;
    r.nz = fa.z * iA;
    r.dt = fa.w;
    r.o = __float_as_int(r.d1.w);
    r.nb = __float_as_int(r.d2.w);
    return r;
}

template <class L>
constexpr bool opposite_pairs()
{
    for (int k = 1; k + 1 < L::Q; k += 2)
        if (L::ex[k] != -L::ex[k + 1] || L::ey[k] != -L::ey[k + 1] || L::ez[k] != -L::ez[k + 1]) return false;
    return (L::Q % 2) == 1;
}

// Sums over the components of a lattice direction with the zero components
// dropped at compile time: 0 * x cannot be folded away at run time (x may be
// infinite or NaN), so these are written out per direction.  -0.0f is the
// exact identity of addition, so the first term folds to itself.
template <int EX, int EY, int EZ>
__device__ __forceinline__ float edot(float a, float b, float c)
{
    float s = -0.0f;
    if constexpr (EX == 1) s += a; else if constexpr (EX == -1) s -= a;
    if constexpr (EY == 1) s += b; else if constexpr (EY == -1) s -= b;
    if constexpr (EZ == 1) s += c; else if constexpr (EZ == -1) s -= c;
    return s;
}

template <int E>
__device__ __forceinline__ void eacc(float &s, float v)
{
    if constexpr (E == 1) s += v; else if constexpr (E == -1) s -= v;
}

// equilibrium deviation g = w (drho + rho P) of one lattice point with state
// ph = (drho, u, v, w): P = 3 e.u + 4.5 (e.u)^2 / gamma - 1.5 u.u / gamma
__device__ __forceinline__ float eq_point_eu(const float ph[4], float eu, float w, float rho0, float ig)
{
    const float uu = ph[1] * ph[1] + ph[2] * ph[2] + ph[3] * ph[3];
    const float P = eu * (3.0f + 4.5f * ig * eu) - 1.5f * ig * uu;
    return w * (ph[0] + (rho0 + ph[0]) * P);
}

template <int EX, int EY, int EZ>
__device__ __forceinline__ float eq_point(const float ph[4], float w, float rho0, float ig)
{
    const float eu = edot<EX, EY, EZ>(ph[1], ph[2], ph[3]);
    const float uu = ph[1] * ph[1] + ph[2] * ph[2] + ph[3] * ph[3];
    const float P = eu * (3.0f + 4.5f * ig * eu) - 1.5f * ig * uu;
    return w * (ph[0] + (rho0 + ph[0]) * P);
}

struct FaceSums {
    float S0, S1x, S1y, S1z, Tx, Ty, Tz;
};

// one pair of opposite directions (e, -e): one e.n and one e.grad per side
template <class L, bool BAND, int K>
__device__ __forceinline__ void pair_step(FaceSums &f, const float po[4], const float pn[4], const float p0[4],
                                          const float go[12], const float gn[12], float nx, float ny, float nz,
                                          float bw, bool ghost, float rho0, float ig)
{
    constexpr int ex = L::ex[K], ey = L::ey[K], ez = L::ez[K];
    static_assert(L::w(K) == L::w(K + 1), "opposite directions carry equal weights");
    const float en = edot<ex, ey, ez>(nx, ny, nz);
    float pk[4], pb[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const float ego = edot<ex, ey, ez>(go[3 * v], go[3 * v + 1], go[3 * v + 2]);
        const float egn = edot<ex, ey, ez>(gn[3 * v], gn[3 * v + 1], gn[3 * v + 2]);
        if (!BAND) {
            // point x_f - e dt sits on the neighbour side when e.n < 0 (ties to the owner)
            pk[v] = en < 0.0f ? pn[v] - egn : po[v] - ego;
            pb[v] = en > 0.0f ? pn[v] + egn : po[v] + ego;
        } else {
            const float avg = ghost ? egn : 0.5f * (ego + egn);
            pk[v] = en > bw ? po[v] - ego : (en < -bw ? pn[v] - egn : p0[v] - avg);
            pb[v] = en < -bw ? po[v] + ego : (en > bw ? pn[v] + egn : p0[v] + avg);
        }
    }
    const float gk = eq_point<ex, ey, ez>(pk, L::w(K), rho0, ig);
    const float gb = eq_point<-ex, -ey, -ez>(pb, L::w(K + 1), rho0, ig);
    const float gs = gk + gb, gd = gk - gb;
    f.S0 += gs;
    eacc<ex>(f.S1x, gd); eacc<ey>(f.S1y, gd); eacc<ez>(f.S1z, gd);
    const float eng = en * gs;
    eacc<ex>(f.Tx, eng); eacc<ey>(f.Ty, eng); eacc<ez>(f.Tz, eng);
}

template <class L, bool BAND, int... P>
__device__ __forceinline__ void all_pairs(std::integer_sequence<int, P...>, FaceSums &f, const float po[4],
                                          const float pn[4], const float p0[4], const float go[12],
                                          const float gn[12], float nx, float ny, float nz, float bw, bool ghost,
                                          float rho0, float ig)
{
    (pair_step<L, BAND, 2 * P + 1>(f, po, pn, p0, go, gn, nx, ny, nz, bw, ghost, rho0, ig), ...);
}

}  // namespace lbfs
