/*
 * FP64 CPU oracle: spring-network membrane (WLC springs, local/global area,
 * volume, curvature and area-difference bending, membrane viscosity).
 *
 * TEST INFRASTRUCTURE ONLY (see lbfs_oracle.h).
 *
 * Springs are edges a-c shared by the triangles (a,b,c) and (a,c,d), built
 * as in lagrangian_object.cpp:322-490 (ordered one-ring of every node, one
 * spring per consecutive neighbour triple, duplicates removed keeping the
 * first).  Forces restate get_spring_forces_atomics
 * (spring_network.cu:613-971) with get_d_normals / get_d_norms
 * (spring_network.cu:1348-1428); the per-step orchestration follows
 * run_spring_network_functions_atomics (Solver_gpu.cu:2311-2361) and
 * lagrangian_object::add_volume / add_area / add_MAD
 * (lagrangian_object.cpp:601-650); set-up follows get_contour_lengths
 * (spring_network.cu:516-604, get_spring_constant 1600-1615),
 * get_reference_curvatures (spring_network.cu:389-513) and the initial call
 * of Solver_gpu.cu:726-732.
 */
#include "lbfs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double x, y, z;
} d3;
static d3 mk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 scl(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 crs(d3 a, d3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static d3 ld(const double *X, int i) { return mk(X[3 * i], X[3 * i + 1], X[3 * i + 2]); }
static double cmp(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
static d3 unit(int k) { return mk(k == 0, k == 1, k == 2); }

/* ------------------------------------------------------------------ */
/* spring connectivity                                                 */
/* ------------------------------------------------------------------ */
int or_build_springs(int n_nodes, int n_tri, const int *tri, int *spring_out)
{
    int *deg = (int *)calloc((size_t)n_nodes, sizeof(int));
    int *start = (int *)calloc((size_t)n_nodes + 1, sizeof(int));
    int *pairs, *fill, *ring, *rlen, i, t, n_spring = 0, maxdeg = 0;
    long long *keys;
    for (t = 0; t < 3 * n_tri; t++) deg[tri[t]]++;
    for (i = 0; i < n_nodes; i++) { start[i + 1] = start[i] + deg[i]; if (deg[i] > maxdeg) maxdeg = deg[i]; }
    pairs = (int *)malloc(sizeof(int) * 2 * (size_t)start[n_nodes]);
    fill = (int *)calloc((size_t)n_nodes, sizeof(int));
    /* per node, the (next, prev) pair of every triangle in triangle order
     * (lagrangian_object.cpp:300-320) */
    for (t = 0; t < n_tri; t++)
        for (i = 0; i < 3; i++) {
            int c = tri[3 * t + i], p1 = tri[3 * t + (i + 1) % 3], p2 = tri[3 * t + (i + 2) % 3];
            int s = start[c] + fill[c]++;
            pairs[2 * s] = p1;
            pairs[2 * s + 1] = p2;
        }
    /* ordered ring: first pair, then repeatedly the pair starting with the last node */
    ring = (int *)malloc(sizeof(int) * (size_t)start[n_nodes]);
    rlen = (int *)calloc((size_t)n_nodes, sizeof(int));
    for (i = 0; i < n_nodes; i++) {
        int *R = ring + start[i], d = deg[i], k;
        if (d == 0) continue;
        R[0] = pairs[2 * start[i]];
        R[1] = pairs[2 * start[i] + 1];
        rlen[i] = 2;
        while (rlen[i] < d) {
            int last = R[rlen[i] - 1], found = -1;
            for (k = 0; k < d; k++)
                if (pairs[2 * (start[i] + k)] == last) { found = pairs[2 * (start[i] + k) + 1]; break; }
            if (found < 0 || found == R[0]) break;
            R[rlen[i]++] = found;
        }
    }
    /* springs (i, r_j, r_j+1, r_j+2), deduplicated on the sorted node set */
    keys = (long long *)malloc(sizeof(long long) * 3 * (size_t)n_tri);
    for (i = 0; i < n_nodes; i++) {
        int *R = ring + start[i], d = rlen[i], j;
        for (j = 0; j < d; j++) {
            int q[4] = {i, R[j], R[(j + 1) % d], R[(j + 2) % d]}, a, b, dup = 0, s;
            long long key = 0;
            for (a = 0; a < 4; a++)
                for (b = a + 1; b < 4; b++)
                    if (q[b] < q[a]) { int tmp = q[a]; q[a] = q[b]; q[b] = tmp; }
            for (a = 0; a < 4; a++) key = key * (long long)(n_nodes + 1) + q[a];
            for (s = 0; s < n_spring; s++)
                if (keys[s] == key) { dup = 1; break; }
            if (dup) continue;
            keys[n_spring] = key;
            spring_out[4 * n_spring] = i;
            spring_out[4 * n_spring + 1] = R[j];
            spring_out[4 * n_spring + 2] = R[(j + 1) % d];
            spring_out[4 * n_spring + 3] = R[(j + 2) % d];
            n_spring++;
        }
    }
    free(deg); free(start); free(pairs); free(fill); free(ring); free(rlen); free(keys);
    (void)maxdeg;
    return n_spring;
}

/* ------------------------------------------------------------------ */
/* triangle derivatives                                                */
/* ------------------------------------------------------------------ */
/* N = (q - p) x (r - p) of triangle (p, q, r).  dN[j] = dN/d(coord j) for the
 * nine coordinates p.xyz q.xyz r.xyz: e_k x (next - prev) (get_d_normals,
 * spring_network.cu:1348-1389).  dV[j] = d(N . centroid)/d(coord j)
 * (spring_network.cu:1394-1412). */
static void tri_derivs(d3 p, d3 q, d3 r, d3 N, d3 *dN, double *dV)
{
    d3 v[3] = {p, q, r};
    d3 cen = scl(add(add(p, q), r), 1.0 / 3.0);
    int vtx, k;
    for (vtx = 0; vtx < 3; vtx++) {
        d3 nxt = v[(vtx + 1) % 3], prv = v[(vtx + 2) % 3];
        for (k = 0; k < 3; k++) {
            d3 d = crs(unit(k), sub(nxt, prv));
            dN[3 * vtx + k] = d;
            dV[3 * vtx + k] = cmp(N, k) / 3.0 + dot(d, cen);
        }
    }
}

/* ------------------------------------------------------------------ */
/* globals: volume and area (update_spring_network_tet_parameters,        */
/* spring_network.cu:1186-1238 + total<128> sums)                         */
/* ------------------------------------------------------------------ */
void or_membrane_globals(or_membrane *mb, const double *X, double *vol, double *area)
{
    double V = 0.0, A = 0.0;
    int t;
    for (t = 0; t < mb->n_tri; t++) {
        d3 a = ld(X, mb->tri[3 * t]), b = ld(X, mb->tri[3 * t + 1]), c = ld(X, mb->tri[3 * t + 2]);
        d3 N = crs(sub(b, a), sub(c, a));
        d3 cen = scl(add(add(a, b), c), 1.0 / 3.0);
        A += sqrt(dot(N, N)) * 0.5;
        V += dot(N, cen) / 6.0;
    }
    *vol = V;
    *area = A;
}

/* ------------------------------------------------------------------ */
/* forces                                                              */
/* ------------------------------------------------------------------ */
static void edge_forces(or_membrane *mb, const or_mem_params *mp, const double *X, const double *Vn, int e,
                        int initials, double vol_coef, double area_coef, double mad_coef, double spring_c, d3 *Fo)
{
    const int *s = mb->spring + 4 * e;
    d3 A = ld(X, s[0]), B = ld(X, s[1]), C = ld(X, s[2]), D = ld(X, s[3]);
    d3 l = sub(C, A);
    double length = sqrt(dot(l, l));
    d3 N1 = crs(sub(B, A), sub(C, A)), N2 = crs(sub(C, A), sub(D, A));
    double n1 = sqrt(dot(N1, N1)), n2 = sqrt(dot(N2, N2));
    d3 dN1[9], dN2[9];
    double dV1[9], dV2[9], dn1[9], dn2[9], d_area[12], d_vol[12], d_ang[12], d_curve[12];
    double cosang, dir, angle;
    int j;
    tri_derivs(A, B, C, N1, dN1, dV1);
    tri_derivs(C, D, A, N2, dN2, dV2);
    for (j = 0; j < 9; j++) {
        dn1[j] = dot(dN1[j], N1) / n1;
        dn2[j] = dot(dN2[j], N2) / n2;
    }
    /* slot order A(0-2) B(3-5) C(6-8) D(9-11); triangle 2 is (C, D, A) */
    for (j = 0; j < 12; j++) d_area[j] = d_vol[j] = 0.0;
    for (j = 0; j < 9; j++) {
        d_area[j] += dn1[j];
        d_vol[j] += dV1[j];
    }
    for (j = 0; j < 3; j++) {
        d_area[6 + j] += dn2[j];
        d_area[9 + j] += dn2[3 + j];
        d_area[j] += dn2[6 + j];
        d_vol[6 + j] += dV2[j];
        d_vol[9 + j] += dV2[3 + j];
        d_vol[j] += dV2[6 + j];
    }
    for (j = 0; j < 12; j++) {
        d_area[j] *= 1.0 / 6.0;
        d_vol[j] *= 1.0 / 18.0;
    }
    cosang = dot(N1, N2) / (n1 * n2);
    cosang = fmax(fmin(cosang, 1.0), -1.0);
    {
        d3 c1 = scl(add(add(A, B), C), 1.0 / 3.0), c2 = scl(add(add(A, C), D), 1.0 / 3.0);
        dir = dot(N1, sub(c2, c1));
    }
    angle = acos(cosang) * (double)((dir < 0) - (dir > 0));
    if (initials) {
        mb->spring_angle_0[e] = angle;
        mb->spring_area_0[2 * e] = n1;
        mb->spring_area_0[2 * e + 1] = n2;
    }
    {
        d3 h1 = crs(scl(N1, 1.0 / n1), scl(l, 1.0 / length));
        d3 h2 = crs(scl(l, 1.0 / length), scl(N2, 1.0 / n2));
        double a1[9], a2[9];
        for (j = 0; j < 9; j++) {
            a1[j] = -1.0 / n1 * dot(h1, dN1[j]);
            a2[j] = -1.0 / n2 * dot(h2, dN2[j]);
        }
        for (j = 0; j < 12; j++) d_ang[j] = 0.0;
        for (j = 0; j < 9; j++) d_ang[j] += a1[j];
        for (j = 0; j < 3; j++) {
            d_ang[6 + j] += a2[j];
            d_ang[9 + j] += a2[3 + j];
            d_ang[j] += a2[6 + j];
        }
    }
    {
        d3 dL = scl(l, 1.0 / length);
        double sat, kappa, lbc, cc, x, fcoef;
        d3 vca, vf;
        for (j = 0; j < 12; j++) d_curve[j] = d_ang[j] * length;
        d_curve[0] -= dL.x * angle; d_curve[1] -= dL.y * angle; d_curve[2] -= dL.z * angle;
        d_curve[6] += dL.x * angle; d_curve[7] += dL.y * angle; d_curve[8] += dL.z * angle;
        for (j = 0; j < 4; j++) Fo[j] = mk(0, 0, 0);
        /* local + global area (one opposite node per triangle) */
        {
            double c1 = -mp->local_area_modulus * (n1 / mb->spring_area_0[2 * e] - 1.0) + area_coef;
            double c2 = -mp->local_area_modulus * (n2 / mb->spring_area_0[2 * e + 1] - 1.0) + area_coef;
            Fo[1] = add(Fo[1], scl(mk(dn1[3], dn1[4], dn1[5]), c1));
            Fo[3] = add(Fo[3], scl(mk(dn2[3], dn2[4], dn2[5]), c2));
        }
        /* volume */
        for (j = 0; j < 4; j++) Fo[j] = add(Fo[j], scl(mk(d_vol[3 * j], d_vol[3 * j + 1], d_vol[3 * j + 2]), vol_coef));
        /* membrane viscosity (Osay) */
        vca = sub(ld(Vn, s[2]), ld(Vn, s[0]));
        vf = scl(dL, -dot(dL, vca) * 4.0 / sqrt(3.0) * mp->viscosity);
        Fo[2] = add(Fo[2], vf);
        Fo[0] = sub(Fo[0], vf);
        /* bending */
        sat = (n1 + n2) / 6.0;
        mb->spring_area[2 * e] = n1;
        mb->spring_area[2 * e + 1] = n2;
        kappa = angle * length / sat;
        mb->curvature[e] = kappa;
        lbc = -0.5 * mp->local_bending_modulus * (mb->curvature_0[e] - kappa) * (mb->curvature_0[e] - kappa);
        for (j = 0; j < 4; j++) Fo[j] = add(Fo[j], scl(mk(d_area[3 * j], d_area[3 * j + 1], d_area[3 * j + 2]), lbc));
        cc = -mp->global_bending_modulus * (kappa - mb->curvature_0[e]) + mad_coef;
        for (j = 0; j < 4; j++) Fo[j] = add(Fo[j], scl(mk(d_curve[3 * j], d_curve[3 * j + 1], d_curve[3 * j + 2]), cc));
        x = length / mb->contour_length[e];
        if (mp->spring_law == 1) {
            /* WLC-POW spring (spring_network.cu:221-223): -(k_s (1/(1-x)^2 - 1 + 4x) + k_p / L^2) */
            fcoef = -(mb->spring_constant[e] * (1.0 / ((1.0 - x) * (1.0 - x)) - 1.0 + 4.0 * x) +
                      mb->pow_constant[e] / (length * length));
        } else {
            /* worm-like-chain spring */
            fcoef = -mb->spring_constant[e] * (0.25 / ((1.0 - x) * (1.0 - x)) - 0.25 + x - spring_c);
        }
        Fo[2] = add(Fo[2], scl(dL, fcoef));
        Fo[0] = sub(Fo[0], scl(dL, fcoef));
    }
}

void or_membrane_forces(or_membrane *mb, const or_mem_params *mp, const double *X, const double *V, int initials,
                        double *F)
{
    double vol, area, r = mp->wlc_ratio;
    double spring_c = 0.25 * r * r / ((r - 1.0) * (r - 1.0)) - 0.25 + 1.0 / r;
    double vol_coef, area_coef, mad_coef, mad;
    int e, j, n_mad;
    or_membrane_globals(mb, X, &vol, &area);
    mb->total_volume = vol;
    mb->total_area = area;
    if (initials) {
        mb->total_volume_0 = vol * mp->volume_ratio;
        mb->total_area_0 = area;
    }
    vol_coef = -mp->volume_modulus * (mb->total_volume / mb->total_volume_0 - 1.0) / 6.0;
    area_coef = -mp->global_area_modulus * (mb->total_area / mb->total_area_0 - 1.0);
    mad_coef = -2.0 * mp->global_bending_modulus * (mb->mad - mb->mad_0) / mb->total_area_0;
    memset(F, 0, sizeof(double) * 3 * (size_t)mb->n_nodes);
    for (e = 0; e < mb->n_edges; e++) {
        d3 Fo[4];
        edge_forces(mb, mp, X, V, e, initials, vol_coef, area_coef, mad_coef, spring_c, Fo);
        for (j = 0; j < 4; j++) {
            int nd = mb->spring[4 * e + j];
            F[3 * nd] += Fo[j].x;
            F[3 * nd + 1] += Fo[j].y;
            F[3 * nd + 2] += Fo[j].z;
        }
    }
    /* MAD = sum kappa (|N1| + |N2|) / 3 / A (sum_product_curvature<128>,
     * lagrangian_object.cpp:636-650); mad_compat sums only the first n_nodes
     * edges as the launch of Solver_gpu.cu:2355 does */
    n_mad = mp->mad_compat ? (mb->n_nodes < mb->n_edges ? mb->n_nodes : mb->n_edges) : mb->n_edges;
    mad = 0.0;
    for (e = 0; e < n_mad; e++) mad += mb->curvature[e] * (mb->spring_area[2 * e] + mb->spring_area[2 * e + 1]) / 3.0;
    mb->mad = mad / mb->total_area;
    if (initials) mb->mad_0 = mb->mad;
}

/* set-up: contour lengths and spring constants on the stress-free shape
 * (get_contour_lengths, pivkin = false), reference curvatures on the
 * curvature shape, then the initial force call on the initial shape. */
void or_membrane_setup(or_membrane *mb, const or_mem_params *mp, const double *X_stress_free,
                       const double *X_curvature, const double *X_initial, double *F_initial)
{
    const double r = mp->wlc_ratio;
    const double cons = 0.5 * pow(1.0 - 1.0 / r, -3.0) + 1.0;
    int e;
    double *V0 = (double *)calloc(3 * (size_t)mb->n_nodes, sizeof(double));
    for (e = 0; e < mb->n_edges; e++) {
        const int *s = mb->spring + 4 * e;
        d3 l = sub(ld(X_stress_free, s[2]), ld(X_stress_free, s[0]));
        double L = sqrt(dot(l, l));
        mb->spring_constant[e] = 4.0 * r * L * mp->shear_modulus / (sqrt(3.0) * cons);
        mb->contour_length[e] = L * r;
        if (mp->spring_law == 1) {
            /* get_spring_pow_constants (spring_network.cu:1655-1682) on the rest
             * length L: the linear shear modulus of the network equals the
             * membrane's, and the force vanishes at L */
            double x = 1.0 / r;
            double alpha = sqrt(3.0) / L;
            double beta = x * 0.5 / pow(1.0 - x, 3.0) - 0.25 / pow(1.0 - x, 2.0) + 0.25;
            double gamma = sqrt(3.0) * 3.0 * 0.25 / pow(L, 3.0);
            double beta2 = 1.0 / pow(1.0 - x, 2.0) - 1.0 + 4.0 * x;
            double ks = mp->shear_modulus / (alpha * beta + gamma * L * L * beta2);
            mb->spring_constant[e] = ks;
            mb->pow_constant[e] = -L * L * ks * beta2;
        }
    }
    for (e = 0; e < mb->n_edges; e++) {
        const int *s = mb->spring + 4 * e;
        d3 A = ld(X_curvature, s[0]), B = ld(X_curvature, s[1]), C = ld(X_curvature, s[2]), D = ld(X_curvature, s[3]);
        d3 N1 = crs(sub(B, A), sub(C, A)), N2 = crs(sub(C, A), sub(D, A));
        double n1 = sqrt(dot(N1, N1)), n2 = sqrt(dot(N2, N2)), cosang, dir, angle, L;
        d3 c1 = scl(add(add(A, B), C), 1.0 / 3.0), c2 = scl(add(add(A, C), D), 1.0 / 3.0);
        cosang = fmax(fmin(dot(N1, N2) / (n1 * n2), 1.0), -1.0);
        dir = dot(sub(N2, N1), sub(c2, c1));
        angle = acos(cosang) * (double)((dir > 0) - (dir < 0));
        L = sqrt(dot(sub(C, A), sub(C, A)));
        mb->curvature[e] = angle * L / ((n1 + n2) / 6.0);
        mb->spring_area[2 * e] = n1;
        mb->spring_area[2 * e + 1] = n2;
        mb->curvature_0[e] = mp->sphere_reference ? mp->curvature_ratio / 3.265483301 : 0.0;
        mb->spring_angle_0[e] = angle;
        mb->spring_area_0[2 * e] = n1;
        mb->spring_area_0[2 * e + 1] = n2;
    }
    mb->mad = 0.0;
    mb->mad_0 = 0.0;
    mb->total_volume_0 = 1.0;
    mb->total_area_0 = 1.0;
    or_membrane_forces(mb, mp, X_initial, V0, 1, F_initial);
    free(V0);
}

/*
 * Output-step energies (Solver_gpu.cu:1060-1139): volume, global area and
 * ADE from the globals; local area (get_area_energy, spring_network.cu:1687-1698,
 * which the reference scales with the global area modulus), local bending
 * (get_bending_energy :1703-1715, launched with the flat |N| array as the
 * nodal area, so edge e divides by entry e of the 2E array) and WLC springs
 * (get_spring_energy :1721-1768, launched with 1 / wlc_ratio).  The reference
 * adds only the first n_tet_Blocks block partials of the edge sums; here the
 * sums run over every edge.
 * out: spring, local area, local bending, global area, volume, ADE, total.
 */
void or_membrane_energies(const or_membrane *mb, const or_mem_params *mp, const double *X, double *out)
{
    const double x0 = 1.0 / mp->wlc_ratio;
    const double e1 = 0.25 / (1.0 - x0) / (1.0 - x0) - 0.25 + x0;
    const double e2 = x0 * x0 * (3.0 - 2.0 * x0) / (1.0 - x0) - 4.0 * e1 * x0;
    double dv = mb->total_volume - mb->total_volume_0, da = mb->total_area - mb->total_area_0, dm = mb->mad - mb->mad_0;
    double e_spring = 0.0, e_la = 0.0, e_lb = 0.0;
    int e;
    for (e = 0; e < mb->n_edges; e++) {
        const int *s = mb->spring + 4 * e;
        d3 l = sub(ld(X, s[2]), ld(X, s[0]));
        double Lm = mb->contour_length[e], x = sqrt(dot(l, l)) / Lm;
        double a1 = mb->spring_area[2 * e] - mb->spring_area_0[2 * e];
        double a2 = mb->spring_area[2 * e + 1] - mb->spring_area_0[2 * e + 1];
        double dk = mb->curvature[e] - mb->curvature_0[e];
        e_la += mp->global_area_modulus * 0.5 *
                (a1 * a1 / mb->spring_area_0[2 * e] + a2 * a2 / mb->spring_area_0[2 * e + 1]);
        e_lb += 0.5 * mp->local_bending_modulus * dk * dk / mb->spring_area[e];
        if (mp->spring_law == 1) {
            /* WLC-POW (no energy routine in the reference): the integral of
             * the spring force from the rest length L0 = x0 Lm,
             * k_s Lm [g(x) - g(x0)] - k_p (1/L - 1/L0), g(x) = 1/(1-x) - x + 2x^2 */
            double L = x * Lm, L0 = x0 * Lm;
            double g = 1.0 / (1.0 - x) - x + 2.0 * x * x, g0 = 1.0 / (1.0 - x0) - x0 + 2.0 * x0 * x0;
            e_spring += mb->spring_constant[e] * Lm * (g - g0) - mb->pow_constant[e] * (1.0 / L - 1.0 / L0);
        } else {
            e_spring += 0.5 * mb->spring_constant[e] *
                        fabs(0.25 * Lm * (x * x * (3.0 - 2.0 * x) / (1.0 - x) - 4.0 * x * e1 - e2));
        }
    }
    out[0] = e_spring;
    out[1] = e_la;
    out[2] = e_lb;
    out[3] = 0.5 * mp->global_area_modulus * da * da / mb->total_area_0;
    out[4] = 0.5 * mp->volume_modulus * dv * dv / mb->total_volume_0;
    out[5] = 0.5 * mp->global_bending_modulus * dm * dm / mb->total_area_0;
    out[6] = out[0] + out[1] + out[2] + out[3] + out[4] + out[5];
}
