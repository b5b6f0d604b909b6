/*
 * FP64 CPU oracle for the LBFS + IBM + spring-network membrane time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * Every function restates the reference algorithm in plain C99 double
 * precision and cites the reference file:line it follows (paths relative to
 * the reference repository root).  Deviations from the reference (periodic
 * translation vectors on wrap faces, D3Q19/D2Q9 lattices, the quirk flags) are
 * named where they occur.
 */
#ifndef LBFS_ORACLE_H
#define LBFS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* lattice ids */
#define OR_D3Q15 0
#define OR_D3Q19 1
#define OR_D2Q9 2

/* quirk / semantics flags (bitmask in or_params.flags) */
#define OR_FLAG_UPWIND_BAND 1   /* GPU tie band |e.n| <= 0.001 dt (LBFS.cu:349-352) instead of strict sign */
#define OR_FLAG_NEUMANN_SIGNED 2 /* CPU Neumann ghost gradient (1 - n) (gradients.cpp:412-414) instead of (1 - |n|) */
#define OR_FLAG_GHOST_DIR_GPU 4  /* Dirichlet ghost gradient along owner->ghost (LBFS.cu:1965-1985) instead of face normal */
#define OR_FLAG_LITERAL_D3Q15 8  /* use the hand-expanded D3Q15 formulas instead of the table-driven lattice */
#define OR_FLAG_GHOST_FACE_GRAD 64 /* boundary faces: every lattice point takes the ghost gradient, as the GPU
                                      reconstruction sets grad_*_1 = grad_*_2 there (LBFS.cu:391-394) */

typedef struct {
    int n_cells, n_faces, n_int, n_bc;
    const int *owner;        /* [n_faces] */
    const int *nbr;          /* [n_faces]; boundary face f -> ghost n_cells + f - n_int */
    const double *shift;     /* [3 n_faces] translation of the neighbour image (periodic wraps), owner frame */
    const double *area;      /* [n_faces] */
    const double *normal;    /* [3 n_faces] unit normal owner -> neighbour */
    const double *fc;        /* [3 n_faces] face centroid, owner frame */
    const double *fdt;       /* [n_faces] streaming dt of the face */
    const double *cc;        /* [3 (n_cells + n_bc)] cell and ghost centroids */
    const double *vol;       /* [n_cells] */
    const int *gcell;        /* [6 n_cells] gradient stencil */
    const int *gface;        /* [6 n_cells] face of each stencil entry */
    const int *grole;        /* [6 n_cells] 0: the cell owns the face, 1: it is the neighbour */
    const double *lhs;       /* [9 n_cells] stored inverse of the LS matrix */
    const double *rhs;       /* [18 n_cells] RHS weights ((i*6 + k)*3 + xyz) */
    const int *bc_rho_type;  /* [n_bc] 1 Dirichlet, 2 Neumann */
    const int *bc_vel_type;  /* [n_bc] 1 Dirichlet, 2 Neumann */
    const double *bc_val;    /* [4 n_bc] rho, u, v, w */
    const double *visc_factor; /* [n_cells] or NULL (1) */
} or_mesh;

typedef struct {
    int lattice;
    int flags;
    double visc;     /* kinematic viscosity nu */
    double gamma;    /* preconditioning gamma (1 = none) */
    double dt;       /* time-marching step */
    double channel_diameter; /* H of the parabolic (9) and shear (10) velocity BCs, domain.Y in the reference */
    int time_stepping;       /* OR_TS_*: gpu_time_stepping of Solver_gpu.cu:767-790 */
    double max_velocity;     /* U of the CFL estimate (global_variables) */
} or_params;

/* time stepping: 1 local, 2 constant dt, 3 global minimum of the local steps;
 * with 1 and 3, dt is the CFL factor (time_marching_step) */
#define OR_TS_LOCAL 1
#define OR_TS_CONSTANT 2
#define OR_TS_MIN 3

/* boundary condition types (preprocessor.cpp:469-491); OR_BC_EMPTY marks the
 * faces of an "empty" patch, which carry no flux (Solver.cpp:308-310) */
#define OR_BC_DIRICHLET 1
#define OR_BC_NEUMANN 2
#define OR_BC_PARABOLIC 9
#define OR_BC_SHEAR 10
#define OR_BC_EMPTY 11

/* ---- geometry and setup (unstructured_mesh.cpp, gradients.cpp) ---- */
int or_geometry(int n_cells, int n_faces, int n_int, const double *pts, const int *fverts,
                const int *owner, const int *nbr_in, const double *shift,
                double visc, double gamma, double fneq_min, int closed,
                double *area, double *normal, double *fc, double *fdt, double *cc, double *vol,
                int *gcell, int *gface, int *grole, int *nbr_out);

int or_ls_setup(const or_mesh *m, int grad_type, int det_compat, double *lhs, double *rhs);

/* ---- lattice ---- */
int or_lattice(int lattice, int *e /* [27*3] */, double *w /* [27] */);
void or_feq(int lattice, double rho, const double *u, double gamma, double *feq);

/* ---- fluid stages (Solver.cpp CPU path; state AoS st[4 i + {rho,u,v,w}]) ---- */
void or_bc_update(const or_mesh *m, const or_params *p, double *st);
void or_gradients(const or_mesh *m, const or_params *p, const double *st, double *g /* [12 n_total] */);
void or_face_flux(const or_mesh *m, const or_params *p, const double *st, const double *g,
                  double *flux /* [4 n_faces]: mass, mom x, y, z */);
void or_residual(const or_mesh *m, const double *flux, double *res /* [4 n_cells] */);
void or_local_dt(const or_mesh *m, const or_params *p, const double *st, double *dt_cell);
void or_update(const or_mesh *m, const or_params *p, const double *res, const double *force,
               double *cons /* [4 n_cells] rho, rho u, rho v, rho w */, double *st);
int or_fluid_step(const or_mesh *m, const or_params *p, double *st, double *cons, const double *force,
                  double *g_scratch, double *flux_scratch, double *res_out);
void or_residual_rms(int n_cells, const double *res, double ref_rho, double ref_u, double *out4);

/* ---- immersed boundary (immersed_boundary_method.cu) ---- */
#define OR_IBM_LINEAR2 0  /* 2-point trilinear on the logical grid (immersed_boundary_method.cu:14-97, 947-1024) */
#define OR_IBM_COS4 1     /* 4-point cosine on physical centroid distances, renormalised */
#define OR_IBM_LAGRANGE4 2 /* 4-point Lagrange per axis on tensor-product grids (immersed_boundary_method.cu:588-872) */
/* logical grid: cell id i + j nx + k nx ny, nominal spacing h from `origin`;
 * periodic axes wrap (lengths L); cc / st / vol are in cell-id order */
typedef struct {
    int nx, ny, nz;
    int periodic[3];
    double origin[3], L[3], h;
    const double *cc;   /* [3 n] centroids */
    const double *vol;  /* [n] */
    /* OR_IBM_LAGRANGE4 only: per axis the vertex coordinates (n_a + 1, the
     * reference's plateau arrays) and the centroid coordinates of the cells
     * along the first grid line of that axis (n_a) */
    const double *plateau[3];
    const double *line[3];
} or_grid;
int or_ibm_interpolate(int kernel, const or_grid *g, int n_nodes, const double *X, const double *st,
                       double *U);
int or_ibm_spread(int kernel, const or_grid *g, int n_nodes, const double *X, const double *F,
                  double *f_cell /* [3 n], accumulated */);

/* ---- spring-network membrane (spring_network.cu) ---- */
typedef struct {
    double shear_modulus, wlc_ratio, volume_modulus, local_area_modulus, global_area_modulus;
    double global_bending_modulus, local_bending_modulus, viscosity, curvature_ratio, volume_ratio;
    int sphere_reference;
    int mad_compat;   /* 1: MAD summed over the first n_nodes edges (Solver_gpu.cu:2355) */
    int spring_law;   /* 0: WLC (get_spring_forces_atomics); 1: WLC-POW (spring_network.cu:221-223, constants
                         of get_spring_pow_constants :1655-1682) */
} or_mem_params;

typedef struct {
    int n_nodes, n_tri, n_edges;
    const int *tri;        /* [3 n_tri] */
    const int *spring;     /* [4 n_edges] a, b, c, d; edge a-c */
    /* per-edge reference data */
    double *contour_length, *spring_constant, *curvature_0, *spring_angle_0, *spring_area_0 /* [2E] */;
    double *curvature, *spring_area /* [2E] */;
    /* globals */
    double total_volume, total_volume_0, total_area, total_area_0, mad, mad_0;
    double *pow_constant;  /* [n_edges] POW constant k_p of the WLC-POW law (unused with WLC) */
} or_membrane;

int or_build_springs(int n_nodes, int n_tri, const int *tri, int *spring_out /* [4 * 3 n_tri / 2] */);
void or_membrane_setup(or_membrane *mb, const or_mem_params *mp, const double *X_stress_free,
                       const double *X_curvature, const double *X_initial, double *F_initial);
void or_membrane_forces(or_membrane *mb, const or_mem_params *mp, const double *X, const double *V,
                        int initials, double *F);
void or_membrane_globals(or_membrane *mb, const double *X, double *vol, double *area);
void or_membrane_energies(const or_membrane *mb, const or_mem_params *mp, const double *X,
                          double *out /* [7]: spring, local area, local bending, global area, volume, ADE, total */);

#ifdef __cplusplus
}
#endif
#endif
