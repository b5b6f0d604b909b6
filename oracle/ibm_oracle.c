/*
 * FP64 CPU oracle: immersed-boundary interpolation and spreading.
 *
 * TEST INFRASTRUCTURE ONLY (see lbfs_oracle.h).
 *
 * OR_IBM_LINEAR2 restates the active reference kernels
 * interpolate_velocities_on_nodes (immersed_boundary_method.cu:14-97) and
 * spread_forces_on_structured_grid (immersed_boundary_method.cu:947-1024):
 * x_ref = (X - origin)/h, cells round(x_ref)-1 .. round(x_ref), weight
 * 1 - |i + 0.5 - x_ref| per axis on the logical grid, spreading F w / h^3.
 * Extension: periodic axes wrap the cell index instead of dropping it.
 *
 * OR_IBM_COS4 is the 4-point cosine kernel phi(r) = (1 + cos(pi r / 2)) / 4
 * of the inactive reference kernels (immersed_boundary_method.cu:100-165,
 * 1028-1091) made 3D, evaluated on physical centroid distances r = (x_c - X)/h
 * over the 4x4x4 logical stencil and renormalised to sum 1; spreading uses
 * the same weights, F w / V_c, so sum_c f_c V_c = F and interpolation and
 * spreading are adjoint.
 *
 * OR_IBM_LAGRANGE4 restates the non-uniform-grid kernels
 * interpolate_velocities_on_nodes_non_uniform_grid and
 * spread_forces_on_non_uniform_grid (immersed_boundary_method.cu:640-872):
 * per axis the vertex nearest the node (find_index_closest :588-636, ties to
 * the larger index) less 2 gives the first of 4 cells; the weights are the
 * first column of the inverse Vandermonde matrix of the 4 cell-line centroid
 * offsets (vandemonde_inverse :919-941), i.e. the cubic Lagrange basis at the
 * node; spreading F w / V_c.  Extensions: periodic axes wrap the cell index
 * and shift its line coordinate by the period; on walls the line coordinates
 * continue with the end spacing and cells outside the grid are dropped (the
 * reference reads outside its arrays there).
 */
#include "lbfs_oracle.h"

#include <math.h>

#define OR_PI 3.14159265358979323846

static int wrap(int i, int n, int periodic, int *ok)
{
    if (i >= 0 && i < n) return i;
    if (!periodic) { *ok = 0; return 0; }
    i %= n;
    if (i < 0) i += n;
    return i;
}

/* nearest vertex, ties to the larger index (find_index_closest) */
static int closest_vertex(const double *v, int n, double t)
{
    int u = 0;
    if (t <= v[0]) return 0;
    if (t >= v[n - 1]) return n - 1;
    while (u < n && v[u] < t) u++;
    return (t - v[u - 1] >= v[u] - t) ? u : u - 1;
}

/* coordinate of the cell line at (unwrapped) index c along axis a */
static double line_coord(const or_grid *g, int a, int c, int n, int periodic)
{
    const double *l = g->line[a];
    if (c >= 0 && c < n) return l[c];
    if (periodic) {
        int k = c >= 0 ? c / n : -((-c + n - 1) / n);
        return l[c - k * n] + k * g->L[a];
    }
    if (c < 0) return l[0] + c * (n > 1 ? l[1] - l[0] : g->h);
    return l[n - 1] + (c - (n - 1)) * (n > 1 ? l[n - 1] - l[n - 2] : g->h);
}

/* Lagrange weights of 4 line points p at x (first column of the inverse
 * Vandermonde matrix of the offsets p - x) */
static void lagrange4(const double p[4], double x, double w[4])
{
    int i, j;
    for (i = 0; i < 4; i++) {
        double v = 1.0;
        for (j = 0; j < 4; j++)
            if (j != i) v *= (x - p[j]) / (p[i] - p[j]);
        w[i] = v;
    }
}

/* stencil of node X: up to 64 (cell id, weight) pairs; returns count */
static int stencil(int kernel, const or_grid *g, const double *X, int *cells, double *wts)
{
    if (kernel == OR_IBM_LAGRANGE4) {
        const int n3[3] = {g->nx, g->ny, g->nz};
        int base[3], a, i, j, k, cnt = 0;
        double wa[3][4];
        for (a = 0; a < 3; a++) {
            double p[4];
            base[a] = closest_vertex(g->plateau[a], n3[a] + 1, X[a]) - 2;
            for (i = 0; i < 4; i++) p[i] = line_coord(g, a, base[a] + i, n3[a], g->periodic[a]);
            lagrange4(p, X[a], wa[a]);
        }
        for (i = 0; i < 4; i++)
            for (j = 0; j < 4; j++)
                for (k = 0; k < 4; k++) {
                    int idx[3] = {base[0] + i, base[1] + j, base[2] + k}, w3[3], ok = 1;
                    for (a = 0; a < 3; a++) w3[a] = wrap(idx[a], n3[a], g->periodic[a], &ok);
                    if (!ok) continue;
                    cells[cnt] = w3[0] + w3[1] * g->nx + w3[2] * g->nx * g->ny;
                    wts[cnt] = wa[0][i] * wa[1][j] * wa[2][k];
                    cnt++;
                }
        return cnt;
    }
    {
    const int n3[3] = {g->nx, g->ny, g->nz};
    int base[3], S = kernel == OR_IBM_LINEAR2 ? 2 : 4, a, i, j, k, cnt = 0;
    double xr[3], wsum = 0.0;
    for (a = 0; a < 3; a++) {
        xr[a] = (X[a] - g->origin[a]) / g->h;
        base[a] = kernel == OR_IBM_LINEAR2 ? (int)round(xr[a]) - 1 : (int)floor(xr[a] - 1.5);
    }
    for (i = 0; i < S; i++)
        for (j = 0; j < S; j++)
            for (k = 0; k < S; k++) {
                int idx[3] = {base[0] + i, base[1] + j, base[2] + k}, w3[3], ok = 1, c;
                double w = 1.0;
                for (a = 0; a < 3; a++) w3[a] = wrap(idx[a], n3[a], g->periodic[a], &ok);
                if (!ok) continue;
                c = w3[0] + w3[1] * g->nx + w3[2] * g->nx * g->ny;
                for (a = 0; a < 3; a++) {
                    double r;
                    if (kernel == OR_IBM_LINEAR2) {
                        r = idx[a] + 0.5 - xr[a];
                        w *= 1.0 - fabs(r);
                    } else {
                        double xc = g->cc[3 * c + a];
                        /* periodic image of the centroid nearest to the node */
                        if (g->periodic[a]) xc += (idx[a] - w3[a]) * (g->L[a] / n3[a]);
                        r = (xc - X[a]) / g->h;
                        w *= fabs(r) < 2.0 ? 0.25 * (1.0 + cos(OR_PI * r / 2.0)) : 0.0;
                    }
                }
                cells[cnt] = c;
                wts[cnt] = w;
                wsum += w;
                cnt++;
            }
    if (kernel == OR_IBM_COS4 && wsum > 0.0)
        for (i = 0; i < cnt; i++) wts[i] /= wsum;
    return cnt;
    }
}

int or_ibm_interpolate(int kernel, const or_grid *g, int n_nodes, const double *X, const double *st, double *U)
{
    int n, t, a;
    int cells[64];
    double w[64];
    for (n = 0; n < n_nodes; n++) {
        int cnt = stencil(kernel, g, X + 3 * n, cells, w);
        for (a = 0; a < 3; a++) U[3 * n + a] = 0.0;
        for (t = 0; t < cnt; t++)
            for (a = 0; a < 3; a++) U[3 * n + a] += st[4 * cells[t] + 1 + a] * w[t];
    }
    return 0;
}

int or_ibm_spread(int kernel, const or_grid *g, int n_nodes, const double *X, const double *F, double *f_cell)
{
    int n, t, a;
    int cells[64];
    double w[64];
    const double ih3 = 1.0 / (g->h * g->h * g->h);
    for (n = 0; n < n_nodes; n++) {
        int cnt = stencil(kernel, g, X + 3 * n, cells, w);
        for (t = 0; t < cnt; t++) {
            double s = kernel == OR_IBM_LINEAR2 ? ih3 : 1.0 / g->vol[cells[t]];
            for (a = 0; a < 3; a++) f_cell[3 * cells[t] + a] += F[3 * n + a] * w[t] * s;
        }
    }
    return 0;
}

/*
 * Internal viscosity field: update_interior_viscosities_plucker
 * (immersed_boundary_method.cu:389-527) driven by gpu_solver::update_viscosities
 * (Solver_gpu.cu:2153-2208).  Cells whose centroid lies in the node bounding box
 * widened by 2 delta_h cast a ray in +x; every triangle with all three nodes at
 * larger x is tested with Plucker side operators (immersed_boundary_method.cu:
 * 357-384); odd crossings = inside.  factor = 2 min(0.5, d/(2 delta_h) +
 * sin(pi d / delta_h) / (2 pi)) with d the distance to the nearest tested node;
 * visc = ((r + 1) +- (r - 1) factor) / 2.  Cells outside the box keep 1.
 * Deviation: cell 0 is not skipped (the reference tests i > 0).
 */
void or_internal_viscosity(int n_cells, const double *cc, int n_nodes, const double *X, int n_tri, const int *tri,
                           double ratio, double delta_h, double *visc_factor)
{
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    int i, t, a;
    for (i = 0; i < n_nodes; i++)
        for (a = 0; a < 3; a++) {
            if (X[3 * i + a] < lo[a]) lo[a] = X[3 * i + a];
            if (X[3 * i + a] > hi[a]) hi[a] = X[3 * i + a];
        }
    for (a = 0; a < 3; a++) { lo[a] -= 2.0 * delta_h; hi[a] += 2.0 * delta_h; }
    for (i = 0; i < n_cells; i++) {
        const double *c = cc + 3 * i;
        double pc[3], dmin = 1e9, factor, visc;
        int count = 0;
        if (c[0] < lo[0] || c[0] > hi[0] || c[1] < lo[1] || c[1] > hi[1] || c[2] < lo[2] || c[2] > hi[2]) {
            visc_factor[i] = 1.0;
            continue;
        }
        /* Plucker coordinates of the ray from the centroid to x = hi[0] */
        pc[0] = c[0] * c[1] - hi[0] * c[1];
        pc[1] = c[0] * c[2] - hi[0] * c[2];
        pc[2] = c[0] - hi[0];
        for (t = 0; t < n_tri; t++) {
            const double *p[3] = {X + 3 * tri[3 * t], X + 3 * tri[3 * t + 1], X + 3 * tri[3 * t + 2]};
            double side[3];
            int k;
            if (!(p[0][0] > c[0] && p[1][0] > c[0] && p[2][0] > c[0])) continue;
            for (k = 0; k < 3; k++) {
                const double *u = p[k], *v = p[(k + 1) % 3];
                double e0 = u[1] * v[2] - v[1] * u[2], e1 = u[2] - v[2], e2 = v[1] - u[1];
                side[k] = pc[0] * e1 + pc[1] * e2 + pc[2] * e0;
            }
            if ((side[0] < 0 && side[1] < 0 && side[2] < 0) || (side[0] > 0 && side[1] > 0 && side[2] > 0)) count++;
            for (k = 0; k < 3; k++) {
                double d = sqrt((c[0] - p[k][0]) * (c[0] - p[k][0]) + (c[1] - p[k][1]) * (c[1] - p[k][1]) +
                                (c[2] - p[k][2]) * (c[2] - p[k][2]));
                if (d < dmin) dmin = d;
            }
        }
        factor = 2.0 * fmin(0.5, dmin * 0.5 / delta_h + 1.0 / OR_PI * sin(OR_PI * dmin / delta_h) * 0.5);
        if (count % 2) visc = (ratio + 1.0) + (ratio - 1.0) * factor;
        else visc = (ratio + 1.0) + (1.0 - ratio) * factor;
        visc_factor[i] = visc * 0.5;
    }
}
