/*
 * FP64 CPU oracle: mesh geometry, least-squares setup and the LBFS fluid step.
 *
 * TEST INFRASTRUCTURE ONLY (see lbfs_oracle.h).  Restates the reference CPU
 * solver path; citations are file:line in the reference repository.
 */
#include "lbfs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* small vector helpers                                                */
/* ------------------------------------------------------------------ */
static void v_sub(const double *a, const double *b, double *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void v_cross(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double v_norm(const double *a) { return sqrt(v_dot(a, a)); }

/* ------------------------------------------------------------------ */
/* lattice tables                                                      */
/* ------------------------------------------------------------------ */
/* D3Q15 in the reference order and weights: Solver.cpp:1398-1445,
 * Solver_gpu.cu:1407-1445 (2/9 rest, 1/9 axis, 1/72 corners). */
static const int E15[15][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
/* D3Q19 and D2Q9 have no reference counterpart; same construction, own order. */
static const int E19[19][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
static const int E9[9][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};

int or_lattice(int lattice, int *e, double *w)
{
    int q, k;
    if (lattice == OR_D3Q15) {
        q = 15;
        for (k = 0; k < q; k++) {
            memcpy(e + 3 * k, E15[k], sizeof(int) * 3);
            w[k] = k == 0 ? 2.0 / 9.0 : (k < 7 ? 1.0 / 9.0 : 1.0 / 72.0);
        }
    } else if (lattice == OR_D3Q19) {
        q = 19;
        for (k = 0; k < q; k++) {
            memcpy(e + 3 * k, E19[k], sizeof(int) * 3);
            w[k] = k == 0 ? 1.0 / 3.0 : (k < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
        }
    } else if (lattice == OR_D2Q9) {
        q = 9;
        for (k = 0; k < q; k++) {
            memcpy(e + 3 * k, E9[k], sizeof(int) * 3);
            w[k] = k == 0 ? 4.0 / 9.0 : (k < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
        }
    } else {
        return -1;
    }
    return q;
}

/* feq_k = w_k rho (1 + 3 e.u + (4.5 (e.u)^2 - 1.5 u.u) / gamma): the
 * polynomial of Solver.cpp:1278-1374 (uu2, uv ... divided by gamma, linear
 * term not) written with the lattice table. */
void or_feq(int lattice, double rho, const double *u, double gamma, double *feq)
{
    int e[81];
    double w[27];
    int q = or_lattice(lattice, e, w), k;
    double uu = v_dot(u, u);
    for (k = 0; k < q; k++) {
        double eu = e[3 * k] * u[0] + e[3 * k + 1] * u[1] + e[3 * k + 2] * u[2];
        feq[k] = w[k] * rho * (1.0 + 3.0 * eu + (4.5 * eu * eu - 1.5 * uu) / gamma);
    }
}

/* Hand-expanded D3Q15 equilibrium of lattice point k, one case per reference
 * branch: Solver.cpp:1278-1374 (populate_feq) and Solver.cpp:681-759
 * (interface feq). */
static double feq15_literal(int k, double rho, double u, double v, double w, double gamma)
{
    const double W0 = 2.0 / 9.0, W1 = 1.0 / 9.0, W7 = 1.0 / 72.0;
    double uu2 = u * u / gamma, vv2 = v * v / gamma, ww2 = w * w / gamma;
    double s = (uu2 + vv2 + ww2) * 1.5;
    double uv = u * v * 9.0 / gamma, uw = u * w * 9.0 / gamma, vw = v * w * 9.0 / gamma;
    double q2 = 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - s;
    switch (k) {
    case 0: return W0 * rho * (1.0 - s);
    case 1: return W1 * rho * (1.0 + 3.0 * u + 4.5 * uu2 - s);
    case 2: return W1 * rho * (1.0 - 3.0 * u + 4.5 * uu2 - s);
    case 3: return W1 * rho * (1.0 + 3.0 * v + 4.5 * vv2 - s);
    case 4: return W1 * rho * (1.0 - 3.0 * v + 4.5 * vv2 - s);
    case 5: return W1 * rho * (1.0 + 3.0 * w + 4.5 * ww2 - s);
    case 6: return W1 * rho * (1.0 - 3.0 * w + 4.5 * ww2 - s);
    case 7: return W7 * rho * (1.0 + 3.0 * u + 3.0 * v + 3.0 * w + uv + uw + vw + q2);
    case 8: return W7 * rho * (1.0 - 3.0 * u - 3.0 * v - 3.0 * w + uv + uw + vw + q2);
    case 9: return W7 * rho * (1.0 + 3.0 * u + 3.0 * v - 3.0 * w + uv - uw - vw + q2);
    case 10: return W7 * rho * (1.0 - 3.0 * u - 3.0 * v + 3.0 * w + uv - uw - vw + q2);
    case 11: return W7 * rho * (1.0 + 3.0 * u - 3.0 * v + 3.0 * w - uv + uw - vw + q2);
    case 12: return W7 * rho * (1.0 - 3.0 * u + 3.0 * v - 3.0 * w - uv + uw - vw + q2);
    case 13: return W7 * rho * (1.0 - 3.0 * u + 3.0 * v + 3.0 * w - uv - uw + vw + q2);
    default: return W7 * rho * (1.0 + 3.0 * u - 3.0 * v - 3.0 * w - uv - uw + vw + q2);
    }
}

/* table-driven equilibrium of lattice point k (same polynomial as or_feq) */
static double feq_one(const int *ek, double wk, double rho, const double *u, double gamma)
{
    double eu = ek[0] * u[0] + ek[1] * u[1] + ek[2] * u[2];
    double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    return wk * rho * (1.0 + 3.0 * eu + (4.5 * eu * eu - 1.5 * uu) / gamma);
}

/* ------------------------------------------------------------------ */
/* mesh geometry                                                       */
/* ------------------------------------------------------------------ */
/*
 * Face area / centroid / normal by a fan of triangles from vertex 0
 * (unstructured_mesh.cpp:1170-1204, cell_tet_area; 1609-1640), cell volume by
 * origin tetrahedra (unstructured_mesh.cpp:1015-1045, tet_volume), cell
 * centroid as the area-weighted mean of face centroids
 * (unstructured_mesh.cpp:1627-1677), face streaming dt
 * min(0.5 sqrt(A), 3 nu / (gamma fneq_min)) (unstructured_mesh.cpp:1435-1443),
 * ghost centroid halfway between owner and face (unstructured_mesh.cpp:1963-1996),
 * gradient stencil in face order (unstructured_mesh.cpp:1647-1671).
 * Extensions: interior faces may carry a periodic translation `shift` (the
 * neighbour sees the face displaced by -shift); `closed` replaces the face
 * area by the magnitude of its area vector.
 */
int or_geometry(int n_cells, int n_faces, int n_int, const double *pts, const int *fverts,
                const int *owner, const int *nbr_in, const double *shift,
                double visc, double gamma, double fneq_min, int closed,
                double *area, double *normal, double *fc, double *fdt, double *cc, double *vol,
                int *gcell, int *gface, int *grole, int *nbr_out)
{
    int n_bc = n_faces - n_int, f, i, t;
    double *tot_area = (double *)calloc((size_t)n_cells, sizeof(double));
    int *cnt = (int *)calloc((size_t)n_cells, sizeof(int));
    if (!tot_area || !cnt) { free(tot_area); free(cnt); return -1; }
    memset(vol, 0, sizeof(double) * (size_t)n_cells);
    memset(cc, 0, sizeof(double) * 3 * (size_t)(n_cells + n_bc));
    for (f = 0; f < n_faces; f++) {
        const double *P[4];
        double csum[3] = {0, 0, 0}, nsum[3] = {0, 0, 0}, asum = 0.0;
        const double *s = shift ? shift + 3 * f : NULL;
        int o = owner[f], nb = f < n_int ? nbr_in[f] : n_cells + f - n_int;
        for (i = 0; i < 4; i++) P[i] = pts + 3 * fverts[4 * f + i];
        for (t = 0; t < 2; t++) {
            const double *a = P[0], *b = P[t + 1], *c = P[t + 2];
            double ab[3], ac[3], N[3], cen[3], tv, ta;
            v_sub(b, a, ab);
            v_sub(c, a, ac);
            v_cross(ab, ac, N);
            tv = v_dot(a, N) / 6.0;
            vol[o] += tv;
            if (f < n_int) vol[nb] -= tv - (s ? v_dot(s, N) / 6.0 : 0.0);
            ta = 0.5 * v_norm(N);
            for (i = 0; i < 3; i++) {
                cen[i] = (a[i] + b[i] + c[i]) / 3.0;
                csum[i] += cen[i] * ta;
                nsum[i] += N[i];
            }
            asum += ta;
        }
        /* closed: area = |sum N| / 2 so that n A is the exact area vector and
         * every cell closes (extension); otherwise the sum of triangle areas
         * as in the reference */
        area[f] = closed ? 0.5 * v_norm(nsum) : asum;
        {
            double nn = v_norm(nsum);
            for (i = 0; i < 3; i++) {
                fc[3 * f + i] = csum[i] / asum;
                normal[3 * f + i] = nsum[i] / nn;
            }
        }
        {
            double fd = 0.5 * sqrt(asum);
            if (fneq_min > 0.0) {
                double ref = visc * 3.0 / gamma / fneq_min;
                if (ref < fd) fd = ref;
            }
            fdt[f] = fd;
        }
        tot_area[o] += asum;
        for (i = 0; i < 3; i++) cc[3 * o + i] += csum[i];
        if (cnt[o] >= 6) { free(tot_area); free(cnt); return -2; }
        gcell[6 * o + cnt[o]] = nb;
        gface[6 * o + cnt[o]] = f;
        grole[6 * o + cnt[o]] = 0;
        cnt[o]++;
        if (f < n_int) {
            tot_area[nb] += asum;
            for (i = 0; i < 3; i++) cc[3 * nb + i] += csum[i] - (s ? s[i] * asum : 0.0);
            if (cnt[nb] >= 6) { free(tot_area); free(cnt); return -2; }
            gcell[6 * nb + cnt[nb]] = o;
            gface[6 * nb + cnt[nb]] = f;
            grole[6 * nb + cnt[nb]] = 1;
            cnt[nb]++;
        }
        nbr_out[f] = nb;
    }
    for (i = 0; i < n_cells; i++) {
        if (cnt[i] != 6) { free(tot_area); free(cnt); return -3; }
        for (t = 0; t < 3; t++) cc[3 * i + t] /= tot_area[i];
    }
    for (f = n_int; f < n_faces; f++) {
        int g = n_cells + f - n_int, o = owner[f];
        for (t = 0; t < 3; t++) cc[3 * g + t] = cc[3 * o + t] + 0.5 * (fc[3 * f + t] - cc[3 * o + t]);
    }
    free(tot_area);
    free(cnt);
    return 0;
}

/* image of stencil entry k of cell i relative to i: displacement d = x_nb - x_i
 * and the face centroid offset xf - x_i, both in the frame of cell i */
static void stencil_vectors(const or_mesh *m, int i, int k, double *d, double *dxf)
{
    int i1 = m->gcell[6 * i + k], f = m->gface[6 * i + k], r = m->grole[6 * i + k], t;
    const double *s = m->shift ? m->shift + 3 * f : NULL;
    for (t = 0; t < 3; t++) {
        double sh = s ? s[t] : 0.0;
        if (r == 0) {
            d[t] = m->cc[3 * i1 + t] + sh - m->cc[3 * i + t];
            dxf[t] = m->fc[3 * f + t] - m->cc[3 * i + t];
        } else {
            d[t] = m->cc[3 * i1 + t] - sh - m->cc[3 * i + t];
            dxf[t] = m->fc[3 * f + t] - sh - m->cc[3 * i + t];
        }
    }
}

/*
 * Hybrid least-squares / Green-Gauss weights: gradients.cpp:214-313
 * (pre_fill_LHS_and_RHS_matrix) and gradients.cpp:149-210
 * (add_non_macrovariable_contributions).  grad_type 0 hybrid, 1 LS, 2 GG.
 * det_compat reproduces the determinant of gradients.cpp:282-284.
 */
int or_ls_setup(const or_mesh *m, int grad_type, int det_compat, double *lhs, double *rhs)
{
    int i, k;
    for (i = 0; i < m->n_cells; i++) {
        double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double max_area = 0.0, max_disp = 0.0, aspect, beta, det;
        double a, b, c, d_, e, f_, g, h, kk;
        for (k = 0; k < 6; k++) {
            double d[3], dxf[3];
            int f = m->gface[6 * i + k];
            stencil_vectors(m, i, k, d, dxf);
            if (m->area[f] > max_area) max_area = m->area[f];
            if (v_norm(d) > max_disp) max_disp = v_norm(d);
        }
        aspect = 2.0 * fabs(max_disp) * max_area / m->vol[i];
        if (grad_type == 0) beta = fmin(1.0, 2.0 / aspect);
        else if (grad_type == 1) beta = 1.0;
        else beta = 0.0;
        L[0] += 2.0 * (1.0 - beta) * m->vol[i];
        L[4] += 2.0 * (1.0 - beta) * m->vol[i];
        L[8] += 2.0 * (1.0 - beta) * m->vol[i];
        for (k = 0; k < 6; k++) {
            double d[3], dxf[3], Lc, Ln, ln, alpha, w;
            int f = m->gface[6 * i + k], r = m->grole[6 * i + k], p, q;
            const double *n = m->normal + 3 * f;
            double sgn = r == 0 ? 1.0 : -1.0;
            stencil_vectors(m, i, k, d, dxf);
            ln = v_dot(dxf, n);
            Lc = v_norm(d);
            Ln = v_dot(d, n);
            alpha = pow(2.0 * ln / Ln, 2.0);
            w = alpha * m->area[f] / Lc;
            for (p = 0; p < 3; p++)
                for (q = 0; q < 3; q++) L[3 * p + q] += w * d[p] * d[q] * beta;
            for (p = 0; p < 3; p++)
                rhs[(6 * i + k) * 3 + p] = beta * w * d[p] + 2.0 * (1.0 - beta) * 0.5 * n[p] * sgn * m->area[f];
        }
        a = L[0]; b = L[1]; c = L[2]; d_ = L[3]; e = L[4]; f_ = L[5]; g = L[6]; h = L[7]; kk = L[8];
        if (det_compat)
            det = a * (e * kk - f_ * h) + b * (d_ * kk - f_ - g) + c * (d_ * h - e * g);
        else
            det = a * (e * kk - f_ * h) - b * (d_ * kk - f_ * g) + c * (d_ * h - e * g);
        lhs[9 * i + 0] = (e * kk - f_ * h) / det;
        lhs[9 * i + 1] = (c * h - b * kk) / det;
        lhs[9 * i + 2] = (b * f_ - c * e) / det;
        lhs[9 * i + 3] = (f_ * g - d_ * kk) / det;
        lhs[9 * i + 4] = (a * kk - c * g) / det;
        lhs[9 * i + 5] = (c * d_ - a * f_) / det;
        lhs[9 * i + 6] = (d_ * h - e * g) / det;
        lhs[9 * i + 7] = (b * g - a * h) / det;
        lhs[9 * i + 8] = (a * e - b * d_) / det;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* fluid stages                                                        */
/* ------------------------------------------------------------------ */

/* Dirichlet value of variable v (0 rho, 1-3 velocity) at boundary b: the BC
 * value, or for the GPU-only velocity types the target evaluated at the
 * ghost centroid, LBFS.cu:1892-1918: parabolic duct inlet
 * u = 16 u_max (H/2 + y)(H/2 - y)(H/2 + z)(H/2 - z) / H^4, shear u = y u_s. */
static double bc_target(const or_mesh *m, const or_params *p, int b, int v)
{
    const double *bc = m->bc_val + 4 * b;
    const double *c = m->cc + 3 * (m->n_cells + b);
    int type = v == 0 ? m->bc_rho_type[b] : m->bc_vel_type[b];
    if (v == 1 && type == OR_BC_PARABOLIC) {
        double H = p->channel_diameter, H2 = 0.5 * H, H4 = H * H * H * H;
        return 16.0 * bc[1] * (H2 + c[1]) * (H2 - c[1]) * (H2 + c[2]) * (H2 - c[2]) / H4;
    }
    if (v == 1 && type == OR_BC_SHEAR) return c[1] * bc[1];
    return bc[v];
}

/* the ghost mirrors a Dirichlet value for these types */
static int bc_mirrored(const or_mesh *m, int b, int v)
{
    int type = v == 0 ? m->bc_rho_type[b] : m->bc_vel_type[b];
    if (v == 0) return type == OR_BC_DIRICHLET;
    return type == OR_BC_DIRICHLET || type == OR_BC_PARABOLIC || type == OR_BC_SHEAR;
}

/* Ghost states: Solution.cpp:759-833 and LBFS.cu:1843-1928 (types 1, 9, 10
 * mirror 2 bc - owner, Neumann and empty copy the owner with dx = 0). */
void or_bc_update(const or_mesh *m, const or_params *p, double *st)
{
    int b, v;
    for (b = 0; b < m->n_bc; b++) {
        int j = m->n_cells + b, o = m->owner[m->n_int + b];
        for (v = 0; v < 4; v++) {
            if (bc_mirrored(m, b, v)) {
                double t = bc_target(m, p, b, v);
                st[4 * j + v] = t - (st[4 * o + v] - t);
            } else {
                st[4 * j + v] = st[4 * o + v];
            }
        }
    }
}

/*
 * Interior LS gradients grad = LHS^-1 sum_k RHS_k (phi_k - phi_i)
 * (gradients.cpp:319-380, add_LS_contributions gradients.cpp:119-147) and the
 * ghost gradients of gradients.cpp:393-505 / get_ghost_grads 592-648:
 * Dirichlet 2 (bc - owner) / |x_ghost - x_owner| along the face normal (or
 * along owner->ghost with OR_FLAG_GHOST_DIR_GPU, LBFS.cu:1965-1985);
 * Neumann owner gradient times (1 - |n|) (LBFS.cu:1997-2001) or (1 - n) with
 * OR_FLAG_NEUMANN_SIGNED (gradients.cpp:412-414).
 */
void or_gradients(const or_mesh *m, const or_params *p, const double *st, double *g)
{
    int i, k, v, t, b;
    for (i = 0; i < m->n_cells; i++) {
        double R[4][3] = {{0}};
        for (k = 0; k < 6; k++) {
            int i1 = m->gcell[6 * i + k];
            const double *w = m->rhs + (6 * i + k) * 3;
            for (v = 0; v < 4; v++) {
                double dphi = st[4 * i1 + v] - st[4 * i + v];
                for (t = 0; t < 3; t++) R[v][t] += w[t] * dphi;
            }
        }
        for (v = 0; v < 4; v++)
            for (t = 0; t < 3; t++) {
                const double *L = m->lhs + 9 * i + 3 * t;
                g[12 * i + 3 * v + t] = L[0] * R[v][0] + L[1] * R[v][1] + L[2] * R[v][2];
            }
    }
    for (b = 0; b < m->n_bc; b++) {
        int f = m->n_int + b, j = m->n_cells + b, o = m->owner[f];
        const double *n = m->normal + 3 * f;
        double d[3], dmag, dir[3];
        v_sub(m->cc + 3 * j, m->cc + 3 * o, d);
        dmag = v_norm(d);
        for (t = 0; t < 3; t++) dir[t] = (p->flags & OR_FLAG_GHOST_DIR_GPU) ? d[t] / dmag : n[t];
        for (v = 0; v < 4; v++) {
            if (bc_mirrored(m, b, v)) {
                double dphi = 2.0 * (bc_target(m, p, b, v) - st[4 * o + v]);
                for (t = 0; t < 3; t++) g[12 * j + 3 * v + t] = dphi / dmag * dir[t];
            } else {
                for (t = 0; t < 3; t++) {
                    double fac = (p->flags & OR_FLAG_NEUMANN_SIGNED) ? (1.0 - n[t]) : (1.0 - fabs(n[t]));
                    g[12 * j + 3 * v + t] = g[12 * o + 3 * v + t] * fac;
                }
            }
        }
    }
}

/*
 * LBFS face flux, Solver.cpp:286-381 (face loop) with
 * populate_lattice_macros Solver.cpp:840-1274, populate_feq 1278-1374,
 * interface moments 314-339 and calculate_flux_at_interface 655-837.
 * Lattice point k sits at x_f - e_k dt_f; it takes the neighbour-side
 * extrapolation when -e_k.n > 0, the owner side otherwise (ties to the owner),
 * the rest point the average.  Boundary faces (ghost neighbour) use the face
 * value from the BC and the ghost gradient on the outer side
 * (Solver.cpp:855-900).  OR_FLAG_UPWIND_BAND switches to the GPU tie band
 * |e.n| <= 0.001 dt with the averaged gradient (LBFS.cu:349-352, 456-932),
 * OR_FLAG_GHOST_FACE_GRAD to the GPU's ghost gradient at every lattice point
 * of a boundary face (LBFS.cu:391-394).
 * The relaxation factor carries the owner's viscosity ratio as on the GPU
 * (LBFS.cu:2676, 1488).
 */
void or_face_flux(const or_mesh *m, const or_params *p, const double *st, const double *g, double *flux)
{
    int e[81];
    double w[27];
    int q = or_lattice(p->lattice, e, w);
    int literal = (p->flags & OR_FLAG_LITERAL_D3Q15) && p->lattice == OR_D3Q15;
    int f, k, v, t;
    for (f = 0; f < m->n_faces; f++) {
        int i = m->owner[f], nb = m->nbr[f];
        int ghost = nb >= m->n_cells;
        const double *n = m->normal + 3 * f;
        const double *s = m->shift ? m->shift + 3 * f : NULL;
        double dt = m->fdt[f], A = m->area[f];
        double t1[3], t2[3], phi_i[4], phi_nb[4], phi0[4];
        const double *gi = g + 12 * i, *gn = g + 12 * nb;
        /* GPU form: the owner-side points of a boundary face use the ghost
         * gradient too (LBFS.cu:391-394); Solver.cpp:918-1000 keeps the owner's */
        if (ghost && (p->flags & OR_FLAG_GHOST_FACE_GRAD)) gi = gn;
        double lat[27][4], feq[27], rho_f, u_f[3], tau, fneq_fac;
        for (t = 0; t < 3; t++) {
            t1[t] = m->cc[3 * i + t] - m->fc[3 * f + t];
            t2[t] = m->cc[3 * nb + t] + (s ? s[t] : 0.0) - m->fc[3 * f + t];
        }
        if (ghost && m->bc_vel_type[nb - m->n_cells] == OR_BC_EMPTY) {
            /* empty patch: no flux (Solver.cpp:308-310) */
            for (v = 0; v < 4; v++) flux[4 * f + v] = 0.0;
            continue;
        }
        if (ghost) {
            int b = nb - m->n_cells;
            const double *bc = m->bc_val + 4 * b;
            for (v = 0; v < 4; v++) {
                /* only plain Dirichlet substitutes the BC value at the face;
                 * the parabolic/shear types extrapolate with the ghost
                 * gradient (LBFS.cu:358-378) */
                int type = v == 0 ? m->bc_rho_type[b] : m->bc_vel_type[b];
                if (type == OR_BC_DIRICHLET) phi0[v] = bc[v];
                else phi0[v] = st[4 * i + v] - v_dot(t1, gn + 3 * v);
                phi_i[v] = phi_nb[v] = phi0[v];
            }
        } else {
            for (v = 0; v < 4; v++) {
                phi_i[v] = st[4 * i + v] - v_dot(t1, gi + 3 * v);
                phi_nb[v] = st[4 * nb + v] - v_dot(t2, gn + 3 * v);
                phi0[v] = (phi_i[v] + phi_nb[v]) * 0.5;
            }
        }
        for (v = 0; v < 4; v++) lat[0][v] = phi0[v];
        for (k = 1; k < q; k++) {
            const int *ek = e + 3 * k;
            double en = ek[0] * n[0] + ek[1] * n[1] + ek[2] * n[2];
            for (v = 0; v < 4; v++) {
                double dgi = ek[0] * gi[3 * v] + ek[1] * gi[3 * v + 1] + ek[2] * gi[3 * v + 2];
                double dgn = ek[0] * gn[3 * v] + ek[1] * gn[3 * v + 1] + ek[2] * gn[3 * v + 2];
                if (p->flags & OR_FLAG_UPWIND_BAND) {
                    double band = 0.001 * dt;
                    if (en > band) lat[k][v] = phi_i[v] - dgi * dt;
                    else if (en < -band) lat[k][v] = phi_nb[v] - dgn * dt;
                    else lat[k][v] = phi0[v] - (ghost ? dgn : 0.5 * (dgi + dgn)) * dt;
                } else {
                    if (-en > 0.0) lat[k][v] = phi_nb[v] - dgn * dt;
                    else lat[k][v] = phi_i[v] - dgi * dt;
                }
            }
        }
        /* equilibria at the lattice points and the interface moments */
        for (k = 0; k < q; k++) {
            if (literal) feq[k] = feq15_literal(k, lat[k][0], lat[k][1], lat[k][2], lat[k][3], p->gamma);
            else feq[k] = feq_one(e + 3 * k, w[k], lat[k][0], &lat[k][1], p->gamma);
        }
        rho_f = 0.0;
        u_f[0] = u_f[1] = u_f[2] = 0.0;
        for (k = 0; k < q; k++) {
            rho_f += feq[k];
            for (t = 0; t < 3; t++) u_f[t] += e[3 * k + t] * feq[k];
        }
        for (t = 0; t < 3; t++) u_f[t] = 1.0 / rho_f * u_f[t];
        fneq_fac = m->visc_factor ? m->visc_factor[i] : 1.0;
        tau = p->visc * fneq_fac * 3.0 / dt / p->gamma;
        {
            double Px = 0, Py = 0, Pz = 0, M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (k = 1; k < q; k++) {
                double fi = literal ? feq15_literal(k, rho_f, u_f[0], u_f[1], u_f[2], p->gamma)
                                    : feq_one(e + 3 * k, w[k], rho_f, u_f, p->gamma);
                double fk = fi - tau * (fi - feq[k]);
                const int *ek = e + 3 * k;
                Px += ek[0] * fk;
                Py += ek[1] * fk;
                Pz += ek[2] * fk;
                for (v = 0; v < 3; v++)
                    for (t = 0; t < 3; t++) M[3 * v + t] += ek[v] * ek[t] * fk;
            }
            flux[4 * f + 0] = (Px * n[0] + Py * n[1] + Pz * n[2]) * A;
            for (v = 0; v < 3; v++)
                flux[4 * f + 1 + v] = (M[v] * n[0] + M[3 + v] * n[1] + M[6 + v] * n[2]) * A;
        }
        (void)s;
    }
}

/* Scatter of face fluxes to cells: Solver.cpp:353-366 (owner minus,
 * neighbour plus, each divided by its volume); ghost cells receive nothing. */
void or_residual(const or_mesh *m, const double *flux, double *res)
{
    int f, v;
    memset(res, 0, sizeof(double) * 4 * (size_t)m->n_cells);
    for (f = 0; f < m->n_faces; f++) {
        int i = m->owner[f], nb = m->nbr[f];
        for (v = 0; v < 4; v++) {
            res[4 * i + v] -= flux[4 * f + v] / m->vol[i];
            if (nb < m->n_cells) res[4 * nb + v] += flux[4 * f + v] / m->vol[nb];
        }
    }
}

/* Forward-Euler conservative update (one Runge-Kutta stage, RungeKutta.h:12-14,
 * Solver.cpp:405-438; force added to the momentum residual as on the GPU,
 * Solver_gpu.cu:1008-1010 / LBFS.cu:2089-2172).  cons holds (rho, rho u). */
/*
 * Local time steps (get_cfl_device, LBFS.cu:2871-2936; areas of
 * populate_cfl_areas, Solver_gpu.cu:1657-1690): with A_t = sum over the
 * cell's faces of |n_t| A / 2 and c = U sqrt(1 - gamma + (gamma / sqrt 3 / U)^2),
 * lambda = (U + c)(A_x + A_y + A_z) / gamma + 4 (2 nu / (gamma rho V)) |A|^2 and
 * dt_i = CFL V / lambda.  OR_TS_MIN replaces every dt_i by their minimum
 * (min<128> + min_block, Solver_gpu.cu:773-783); OR_TS_CONSTANT uses dt.
 */
void or_local_dt(const or_mesh *m, const or_params *p, const double *st, double *dt_cell)
{
    int i, f, t;
    double U = p->max_velocity, g = p->gamma, mn = 1e300;
    double c = U * sqrt(1.0 - g + pow(g / sqrt(3.0) / U, 2.0));
    double *Ac = (double *)calloc(3 * (size_t)m->n_cells, sizeof(double));
    for (f = 0; f < m->n_faces; f++) {
        int o = m->owner[f], nb = m->nbr[f];
        for (t = 0; t < 3; t++) {
            double a = fabs(m->normal[3 * f + t] * m->area[f]);
            Ac[3 * o + t] += 0.5 * a;
            if (nb < m->n_cells) Ac[3 * nb + t] += 0.5 * a;
        }
    }
    for (i = 0; i < m->n_cells; i++) {
        const double *A = Ac + 3 * i;
        double lam = (U + c) * (A[0] + A[1] + A[2]) / g;
        double ve = 2.0 * p->visc / g / st[4 * i] / m->vol[i] * (A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        dt_cell[i] = p->dt * m->vol[i] / (lam + 4.0 * ve);
        if (dt_cell[i] < mn) mn = dt_cell[i];
    }
    if (p->time_stepping == OR_TS_MIN)
        for (i = 0; i < m->n_cells; i++) dt_cell[i] = mn;
    free(Ac);
}

void or_update(const or_mesh *m, const or_params *p, const double *res, const double *force,
               double *cons, double *st)
{
    int i, v;
    double *dtc = NULL;
    if (p->time_stepping == OR_TS_LOCAL || p->time_stepping == OR_TS_MIN) {
        dtc = (double *)malloc(sizeof(double) * (size_t)m->n_cells);
        or_local_dt(m, p, st, dtc);
    }
    for (i = 0; i < m->n_cells; i++) {
        double dt = dtc ? dtc[i] : p->dt;
        cons[4 * i] += dt * res[4 * i];
        for (v = 1; v < 4; v++) cons[4 * i + v] += dt * (res[4 * i + v] + (force ? force[3 * i + v - 1] : 0.0));
        st[4 * i] = cons[4 * i];
        for (v = 1; v < 4; v++) st[4 * i + v] = cons[4 * i + v] / cons[4 * i];
    }
    free(dtc);
}

/* One full fluid time step (Solver.cpp:266-439 with RK stages = 1). */
int or_fluid_step(const or_mesh *m, const or_params *p, double *st, double *cons, const double *force,
                  double *g, double *flux, double *res)
{
    int i;
    or_bc_update(m, p, st);
    or_gradients(m, p, st, g);
    or_face_flux(m, p, st, g, flux);
    or_residual(m, flux, res);
    or_update(m, p, res, force, cons, st);
    if (force) {
        for (i = 0; i < m->n_cells; i++) {
            res[4 * i + 1] += force[3 * i];
            res[4 * i + 2] += force[3 * i + 1];
            res[4 * i + 3] += force[3 * i + 2];
        }
    }
    for (i = 0; i < 4 * m->n_cells; i++)
        if (!isfinite(st[i])) return -1;
    return 0;
}

/* Residual RMS: residuals.cpp:114-132 (sqrt(sum r^2 / N) / rho_ref and
 * / (rho_ref U)). */
void or_residual_rms(int n_cells, const double *res, double ref_rho, double ref_u, double *out4)
{
    double s[4] = {0, 0, 0, 0};
    int i, v;
    for (i = 0; i < n_cells; i++)
        for (v = 0; v < 4; v++) s[v] += res[4 * i + v] * res[4 * i + v];
    out4[0] = sqrt(s[0] / n_cells) / ref_rho;
    for (v = 1; v < 4; v++) out4[v] = sqrt(s[v] / n_cells) / (ref_rho * ref_u);
}
