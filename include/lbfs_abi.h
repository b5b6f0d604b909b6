/*
 * C ABI of the MI355X LBFS + IBM + spring-network membrane time step.
 *
 * Plain pointers and sizes only.  All device memory is owned by the context;
 * the caller owns every host buffer it passes in.  Every call returns an int
 * status (LBFS_OK or a negative LBFS_E_* code) and never exits the process;
 * lbfs_last_error() gives the message of the last failure on a context.
 * One context = one HIP stream; call a context from one host thread at a time.
 *
 * Cell ids at this boundary are the caller's (the mesh description's) ids;
 * the library keeps its own tile-blocked order internally.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   gpu_solver::General_Purpose_Solver_mk_i   Solver_gpu.h:28-31, Solver_gpu.cu:178-1275
 *   device array set-up / marshalling          Solver_gpu.cu:394-742, 1821-2151
 *   per-stage kernels                          LBFS.hpp:44-96, immersed_boundary_method.hpp:11-49,
 *                                              spring_network.hpp:18-106
 */
#ifndef LBFS_ABI_H
#define LBFS_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define LBFS_OK 0
#define LBFS_E_INVALID_ARG -1
#define LBFS_E_HIP -2
#define LBFS_E_NONFINITE -3
#define LBFS_E_UNSUPPORTED -4
#define LBFS_E_NO_DEVICE -5
#define LBFS_E_STATE -6

/* ---- lattices ---- */
#define LBFS_D3Q15 0 /* the reference lattice (Solver_gpu.cu:1407-1445) */
#define LBFS_D3Q19 1
#define LBFS_D2Q9 2

/* ---- gradient reconstruction (gradients.cpp:255-265) ---- */
#define LBFS_GRAD_HYBRID 0
#define LBFS_GRAD_LS 1
#define LBFS_GRAD_GG 2

/* ---- boundary condition types (preprocessor.cpp:471-493) ---- */
#define LBFS_BC_DIRICHLET 1
#define LBFS_BC_NEUMANN 2
#define LBFS_BC_PARABOLIC 9  /* velocity only: duct inlet 16 u_max (H/2+y)(H/2-y)(H/2+z)(H/2-z)/H^4 at the
                                ghost centroid, u_max = bc u value, H = channel_diameter (LBFS.cu:1892-1906) */
#define LBFS_BC_SHEAR 10     /* velocity only: u = y * bc u value at the ghost centroid (LBFS.cu:1907-1918) */
#define LBFS_BC_EMPTY 11     /* rho and velocity: face of an "empty" patch, no flux (Solver.cpp:308-310);
                                the reference marks these by patch name, this ABI by type */

/* ---- semantics flags for reference quirks (lbfs_params.flags) ---- */
#define LBFS_FLAG_UPWIND_BAND 1      /* tie band |e.n| <= 0.001 dt_f, LBFS.cu:349-352 */
#define LBFS_FLAG_NEUMANN_SIGNED 2   /* Neumann ghost gradient (1 - n), gradients.cpp:412-414 */
#define LBFS_FLAG_GHOST_DIR_GPU 4    /* Dirichlet ghost gradient along owner->ghost, LBFS.cu:1965-1985 */
#define LBFS_FLAG_LS_DET_COMPAT 16   /* LS determinant as written in gradients.cpp:282-284 */
#define LBFS_FLAG_CLOSED_FACES 32    /* face area |sum of triangle normals| / 2 instead of the sum of
                                        triangle areas (unstructured_mesh.cpp:1193-1196): cells close exactly */
#define LBFS_FLAG_GHOST_FACE_GRAD 64 /* on a boundary face every lattice point takes the ghost gradient, as the
                                        GPU reconstruction sets grad_*_1 = grad_*_2 there (LBFS.cu:391-394);
                                        without it the owner-side points take the owner's own gradient
                                        (Solver.cpp:918-1000) */
#define LBFS_FLAG_HOST_GEOMETRY 128 /* set-up only (no counterpart in the reference): the FP64 mesh geometry and
                                        least-squares weights on the host instead of the GPU; both give the
                                        same bits */
/* the reference GPU solver's semantics (calc_face_flux / calc_face_flux_bc, LBFS.cu:328-936, 1933-2048) */
#define LBFS_FLAGS_REFERENCE_GPU (LBFS_FLAG_UPWIND_BAND | LBFS_FLAG_GHOST_DIR_GPU | LBFS_FLAG_GHOST_FACE_GRAD)

/* ---- IBM kernels ---- */
#define LBFS_IBM_LINEAR2 0 /* 2-point trilinear on the logical grid, immersed_boundary_method.cu:14-97, 947-1024 */
#define LBFS_IBM_COS4 1    /* 4-point cosine on physical distances, renormalised */
#define LBFS_IBM_LAGRANGE4 2 /* 4-point Lagrange per axis for tensor-product (stretched) grids,
                                immersed_boundary_method.cu:588-872, 919-941 */

/*
 * Hexahedral mesh in OpenFOAM polyMesh form (unstructured_mesh.cpp:670-941):
 * internal faces first, then boundary faces.  Quad faces, vertices ordered so
 * that the normal points from owner to neighbour.  Periodic wrap faces are
 * internal faces whose neighbour image is translated by face_shift.
 */
typedef struct {
    int n_points;
    const double *points;        /* [3 n_points] */
    int n_cells;
    int n_faces;
    int n_internal;
    const int *face_vertices;    /* [4 n_faces] */
    const int *owner;            /* [n_faces] */
    const int *neighbour;        /* [n_internal] */
    const double *face_shift;    /* [3 n_faces] or NULL */
    const int *bc_rho_type;      /* [n_faces - n_internal] */
    const int *bc_vel_type;      /* [n_faces - n_internal] */
    const double *bc_values;     /* [4 (n_faces - n_internal)] rho, u, v, w */
    /* logical structure (cell id = i + j nx + k nx ny) used for tiling and IBM; 0 if unknown */
    int nx, ny, nz;
    double origin[3];
    double spacing;              /* nominal cell size (IBM stencil unit) */
    /* ensemble of independent boxes stepped together (no counterpart in the
     * reference): cells numbered box after box, each box logical on its own
     * (i + j nx + k nx ny after the box's first cell); nx, ny, nz, origin and
     * spacing above are then unused.  0 or 1 = the single box above. */
    int n_blocks;
    const int *block_dims;       /* [3 n_blocks] */
    const double *block_origin;  /* [3 n_blocks] */
    const double *block_spacing; /* [n_blocks] */
    /* body force of each box (the same per-cell force term as
     * lbfs_params.body_force, box by box), or NULL: lbfs_params.body_force in
     * every box */
    const double *block_body_force; /* [3 n_blocks] or NULL */
} lbfs_mesh_desc;

typedef struct {
    int lattice;
    int gradient_type;
    int flags;
    double visc;        /* kinematic viscosity (global_variables.cpp:24) */
    double gamma;       /* preconditioning gamma, 1 = none */
    double fneq_min;    /* face dt cap 3 nu / (gamma fneq_min), <= 0 disables */
    double dt;          /* time-marching step */
    double rho0;        /* reference density; the device stores rho - rho0 */
    double body_force[3];
    double channel_diameter; /* H of LBFS_BC_PARABOLIC (domain.Y, Solver_gpu.cu:935); 0 if unused */
    int time_stepping;  /* LBFS_TS_*; 0 = constant dt */
    double max_velocity;/* U of the CFL estimate (LBFS_TS_LOCAL / LBFS_TS_MIN) */
    int tile[3];        /* cells per tile along x, y, z of the logical grid (no counterpart in the
                           reference); 0 0 0 = 8 8 4.  Results do not depend on it. */
    int block_threads;  /* workgroup size of the fluid kernels: 0 (= 256), 256, 512 or 576 (three faces per
                           thread on 8x8x8 tiles: 1728 faces) */
} lbfs_params;

/* time stepping (gpu_time_stepping, Solver_gpu.cu:767-790; get_cfl_device
 * LBFS.cu:2871-2936): with LOCAL and MIN, dt above is the CFL factor and each
 * cell's step is dt V / lambda with lambda = (U + c)(A_x + A_y + A_z) / gamma
 * + 8 nu |A|^2 / (gamma rho V), A_t = sum over the cell's faces of |n_t| A / 2,
 * c = U sqrt(1 - gamma + (gamma / sqrt 3 / U)^2); MIN uses their minimum.
 * Membranes need constant stepping. */
#define LBFS_TS_LOCAL 1
#define LBFS_TS_CONSTANT 2
#define LBFS_TS_MIN 3

typedef struct lbfs_ctx lbfs_ctx;

/* context lifetime: Solver_gpu.cu:394-742 (set-up), no counterpart for destroy */
int lbfs_create(const lbfs_mesh_desc *mesh, const lbfs_params *params, lbfs_ctx **out);
int lbfs_destroy(lbfs_ctx *ctx);
const char *lbfs_last_error(const lbfs_ctx *ctx);
int lbfs_sync(lbfs_ctx *ctx);

/* fluid state in caller cell order, FP64 host arrays of n_cells
 * (Solution::clone(double4*) Solution.cpp:308-319, Solver_gpu.cu:2138-2151) */
int lbfs_set_state(lbfs_ctx *ctx, const double *rho, const double *u, const double *v, const double *w);
int lbfs_get_state(lbfs_ctx *ctx, double *rho, double *u, double *v, double *w);

/* whole time steps: the time loop body of Solver_gpu.cu:749-1246 */
int lbfs_step(lbfs_ctx *ctx, int nsteps);

/* options (no counterpart in the reference) */
#define LBFS_OPT_GRAPH 1 /* 1 (default): replay the step from a captured hipGraph, 0: launch kernels one by one */
#define LBFS_OPT_IBM_BINNED 2 /* spreading for the next lbfs_add_membranes: -1 (default) by node density
                                 (1 from one node per 100 cells, else 2); 0 one wavefront per node into
                                 integer cell accumulators in global memory; 1 the nodes listed per fluid
                                 tile, each tile gathering the nodes around it (sums in LDS, no global
                                 atomics); 2 the nodes listed per 4x4x4 bin, each occupied bin spreading
                                 into an LDS window added to the accumulators; the result is the same
                                 bit for bit */
#define LBFS_OPT_STEP_SCHEDULE 3 /* coupled step schedule (the result is the same bit for bit):
                                    0 (default) the fluid passes on the main stream from the start
                                    of the step, the whole coupling chain on a second; 1 the
                                    interpolation alone, then the fluid passes on the main stream
                                    and the membrane forces and spreading on a second; 2 the fluid
                                    passes on a second stream from the start, the membrane forces
                                    on a third after the interpolation */
#define LBFS_OPT_FUSED_FLUID 4 /* 1 (default where the tiles fit): whole steps run the fluid passes as one
                                 launch per tile (the tile's gradients in LDS); 0: gradients, then the
                                 face pass; the result is the same bit for bit */
#define LBFS_OPT_VISC_SEARCH 5 /* 1 (default): the internal-viscosity field searches binned triangles and
                                 nodes; 0: every triangle of each membrane (immersed_boundary_method.cu:
                                 389-527); the field is the same bit for bit */
int lbfs_set_option(lbfs_ctx *ctx, int option, int value);

/* per-stage entry points (parity tests) */
int lbfs_stage_gradients(lbfs_ctx *ctx, double *grad /* [12 n_cells] or NULL */);   /* LBFS.cu:2177-2240 */
int lbfs_stage_face_flux(lbfs_ctx *ctx, double *flux /* [4 n_faces] */);           /* LBFS.cu:2574-2810 */
int lbfs_stage_residual(lbfs_ctx *ctx, double *res /* [4 n_cells] */);             /* LBFS.cu:2242-2281 */

/* named stages of one step, in order (no counterpart in the reference as an
 * API; the groups are the reference's kernel groups of Solver_gpu.cu:749-1246):
 *   ibm_interp    interpolate_velocities_on_nodes + update_node_positions
 *                 (immersed_boundary_method.cu:14-97, 234-264)
 *   mem_volume, mem_edges, mem_nodes  run_spring_network_functions_atomics
 *                 (Solver_gpu.cu:2311-2361) with the volume, area and MAD sums
 *                 (lagrangian_object.cpp:601-650); mem_nodes only without
 *                 binned spreading
 *   internal_visc update_interior_viscosities_plucker (:389-527), when due
 *   ibm_spread    spread_forces_on_structured_grid (immersed_boundary_method.cu:947-1024)
 *   gradients     get_interior_gradients + get_bc_gradients (LBFS.cu:1933-2240)
 *   dt_min        get_cfl_device + min (LBFS.cu:2871-2936), time_stepping 3 only
 *   flux_update   calc_face_flux(_bc) + add_face_flux_to_cell + add + time_integration
 *                 (LBFS.cu:2574-2810, 2242-2281, 2089-2172) with the static force
 *   ibm_apply     with membranes: the spread force density added to the update
 *                 (common_kernels.cu:220-233) in the tiles the spreading reached
 * Running them all in order, then lbfs_stage_end_step, equals lbfs_step(ctx, 1). */
int lbfs_stage_names(lbfs_ctx *ctx, int max_count, const char **names, int *count);
int lbfs_stage_run(lbfs_ctx *ctx, const char *name);
int lbfs_stage_end_step(lbfs_ctx *ctx);

/* cell force density added to the momentum residual (Solver_gpu.cu:834-846, 1008-1010) */
int lbfs_get_cell_force(lbfs_ctx *ctx, double *f /* [3 n_cells] */);

/* diagnostics: residual RMS (square + total<128> + residuals::l2_norm_rms_moukallad,
 * Solver_gpu.cu:1032-1046, residuals.cpp:114-132: sqrt(sum r^2 / n_cells) / ref_rho for
 * rho, / (ref_rho ref_u) for the momentum rows, whose residual includes the force density
 * as after add(), Solver_gpu.cu:1008-1010), FP64 mass and the largest speed; finite = 0
 * and LBFS_E_NONFINITE when the state or a fixed-point sum of the coupling is not finite */
typedef struct {
    double rho_rms, u_rms, v_rms, w_rms;
    double total_mass;
    double max_speed;
    int finite;
} lbfs_diag;
/* residual of the step that starts from the current state with the static force
 * (body force); the state is unchanged (equals the next step's RMS without membranes) */
int lbfs_reduce_diagnostics(lbfs_ctx *ctx, double ref_rho, double ref_u, lbfs_diag *out);
/* one step (eager launches) and the residual RMS of that step as the reference
 * computes it: flux residual of the state at its start plus the force density the
 * step applied (static force and, with membranes, the force it spread) */
int lbfs_step_residual(lbfs_ctx *ctx, double ref_rho, double ref_u, lbfs_diag *out);
/* the time loop's exit test (Solver_gpu.cu:1206: max_error() < tolerance or time up;
 * residuals::max_error residuals.cpp:158-166 = max of the u, v, w RMS): steps until
 * a checked step's RMS is below tolerance or max_steps are done; every check_every-th
 * step is a lbfs_step_residual step (1 = every step, the reference's test), the others
 * run as lbfs_step.  *steps_done counts the steps taken, *converged is 1 when the
 * tolerance was met, *out holds the last checked step's diagnostics. */
int lbfs_run_to_tolerance(lbfs_ctx *ctx, int max_steps, double tolerance, int check_every, double ref_rho,
                          double ref_u, int *steps_done, int *converged, lbfs_diag *out);

/* ---- spring-network membranes and immersed-boundary coupling ----
 * One descriptor per membrane (lagrangian_object, lagrangian_object.h:17-141;
 * parameters parsed in preprocessor.cpp:120-401).  Node positions are
 * absolute.  The three shapes are those of Solver_gpu.cu:659-732: contour
 * lengths from the stress-free shape, reference angles from the curvature
 * shape, positions from the initial shape (NULL = the initial shape). */
typedef struct {
    int n_nodes;
    int n_tri;
    const double *nodes;              /* [3 n_nodes] initial shape */
    const double *nodes_stress_free;  /* or NULL */
    const double *nodes_curvature;    /* or NULL */
    const int *tri;                   /* [3 n_tri], outward winding */
    double shear_modulus, wlc_ratio;
    double volume_modulus, local_area_modulus, global_area_modulus;
    double global_bending_modulus, local_bending_modulus;
    double viscosity;                 /* membrane viscosity (Osay dissipative term) */
    double curvature_ratio, volume_ratio;
    int sphere_reference;             /* 1: curvature_0 = curvature_ratio / 3.265483301 */
    int mad_compat;                   /* 1: MAD over the first n_nodes edges (Solver_gpu.cu:2355) */
    double internal_viscosity_ratio;  /* interior / exterior viscosity; 1 leaves the field at 1
                                         (Solver_gpu.cu:884-889, lagrangian_object.h:85) */
    int n_springs;                    /* 0: springs built from the triangles (lagrangian_object.cpp:322-490) */
    const int *springs;               /* [4 n_springs] a, b, c, d as read from an MZ file (mesh_read) */
    int block;                        /* box of the ensemble the membrane lives in (0 for one box) */
    int spring_law;                   /* LBFS_SPRING_WLC (0) or LBFS_SPRING_WLC_POW */
} lbfs_membrane_desc;
/* spring laws: the WLC of get_spring_forces_atomics (spring_network.cu:631-633, 932-936), or
 * WLC-POW, -(k_s (1/(1-x)^2 - 1 + 4x) + k_p / L^2) (spring_network.cu:221-223) with k_s, k_p from
 * the shear modulus and the rest length (get_spring_pow_constants, spring_network.cu:1655-1682) */
#define LBFS_SPRING_WLC 0
#define LBFS_SPRING_WLC_POW 1

/* attach membranes (replaces any attached before); requires the mesh's logical
 * structure (nx, ny, nz, origin, spacing); ibm_kernel = LBFS_IBM_*.
 * Set-up of Solver_gpu.cu:659-732; per step run_spring_network_functions_atomics
 * (Solver_gpu.cu:2311-2361), interpolate/advect (Solver_gpu.cu:808-828) and
 * spreading (Solver_gpu.cu:899-915) join the fluid step. */
int lbfs_add_membranes(lbfs_ctx *ctx, const lbfs_membrane_desc *mem, int count, int ibm_kernel);
/* concatenated node state of all membranes (positions absolute) */
int lbfs_get_membrane_nodes(lbfs_ctx *ctx, double *X, double *U, double *F);
/* per membrane: V, A, MAD, V0, A0, MAD0 (FP64, lagrangian_object.cpp:601-650) */
int lbfs_get_membrane_globals(lbfs_ctx *ctx, double *out6);

/* output-step membrane energies, 7 per membrane: WLC spring, local area,
 * local bending, global area, volume, ADE, total (Solver_gpu.cu:1060-1139,
 * spring_network.cu:1687-1768); evaluated from the last force call's state */
int lbfs_get_membrane_energies(lbfs_ctx *ctx, double *out /* [7 n_membranes] */);
/* The same energies, bit for bit, staged in a fixed device array instead of
 * memory the context allocates: this call allocates nothing (no hipMalloc, no
 * hipFree, no allocating library routine).  No counterpart in the reference;
 * it is for callers that cannot allocate (serving loops, code that must not
 * call hipMalloc).  At most LBFS_NOALLOC_MAX_MEMBRANES membranes (else
 * LBFS_E_INVALID_ARG); the staging is 7 * 8192 doubles (448 KiB) per device,
 * shared by every context on that device and held until the process ends, so
 * calls on one device from different threads or overlapping streams race. */
#define LBFS_NOALLOC_MAX_MEMBRANES 8192
int lbfs_get_membrane_energies_noalloc(lbfs_ctx *ctx, double *out /* [7 n_membranes] */);

/* ---- parameters the adjoint differentiates ----
 * Uniform body force (the reference's pressure gradient along its direction,
 * Solver_gpu.cu:834-839, preprocessor.cpp:722-736); replaces lbfs_params.body_force. */
int lbfs_set_body_force(lbfs_ctx *ctx, const double *f /* [3] */);  /* every box of an ensemble */
/* shear modulus (spring constants k_s are proportional to it, spring_network.cu:1600-1615),
 * local and global bending moduli of membrane `index` (lagrangian_object.h) */
int lbfs_set_membrane_moduli(lbfs_ctx *ctx, int index, double shear_modulus, double local_bending_modulus,
                             double global_bending_modulus);

/* ---- discrete adjoint of the coupled step (no counterpart in the reference:
 * gradients of a loss on the state after n steps, for callers that fit
 * membrane moduli or forcing) ----
 * lbfs_adjoint_forward runs n_steps steps from the current state, keeping the
 * state every checkpoint_every steps.  lbfs_adjoint_backward takes the
 * loss's adjoints of the final node positions (3 per node, absolute) and of
 * the final state (4 per cell: rho, u, v, w; caller order), either may be
 * NULL for zero, and returns dL/d(body force) [3], per membrane dL/d(shear
 * modulus, local bending, global bending) [3 n_membranes], and the adjoints
 * of the initial state [4 n_cells] and node positions [3 n_nodes] (with the
 * membranes' reference areas, volume and MAD, set up from the initial shape,
 * held fixed); any output may be NULL.  Each segment of checkpoint_every steps is replayed from its
 * checkpoint; the context is left in the state after the forward run.
 * Needs constant time steps, strict upwinding, the 4-point cosine IBM kernel
 * and internal viscosity ratios of 1.  lbfs_adjoint_rewind restores the
 * state the last forward run started from.  lbfs_adjoint_run: forward +
 * backward for L = sum over nodes |X - target|^2 (target 3 per node). */
int lbfs_adjoint_forward(lbfs_ctx *ctx, int n_steps, int checkpoint_every);
int lbfs_adjoint_backward(lbfs_ctx *ctx, const double *lam_nodes, const double *lam_state, double *grad_body_force,
                          double *grad_moduli, double *lam_state0, double *lam_nodes0);
int lbfs_adjoint_rewind(lbfs_ctx *ctx);
int lbfs_adjoint_run(lbfs_ctx *ctx, int n_steps, int checkpoint_every, const double *target_nodes, double *loss,
                     double *grad_body_force, double *grad_moduli);

/* host-side membrane meshes: frequency-n geodesic sphere (10 n^2 + 2 nodes) */
int lbfs_mesh_icosphere(int frequency, double *nodes, int *tri);
/* Evans-Fung biconcave disc of diameter d0 from a unit sphere (symmetry axis z) */
int lbfs_mesh_evans_fung(int n_nodes, const double *unit_nodes, double d0, double *nodes);
/* springs (a,b,c,d), edge a-c, of a closed triangulation (lagrangian_object.cpp:322-490) */
int lbfs_mesh_springs(int n_nodes, int n_tri, const int *tri, int *springs, int *n_springs);

/* membrane mesh files (lagrangian_object.cpp): Osay "N T" + xyz + 1-based
 * triangles (osay_mesh_read 238-302); MZ "N S T" + xyz + triangles + 1-based
 * springs a b c d + per-node neighbour lists (mesh_read 495-587).  Read
 * positions are shifted by centre (the object's centre_x/y/z) as there. */
#define LBFS_MESH_OSAY 0
#define LBFS_MESH_MZ 1
int lbfs_membrane_mesh_counts(const char *path, int format, int *n_nodes, int *n_tri, int *n_springs);
int lbfs_membrane_mesh_read(const char *path, int format, const double centre[3], double *nodes, int *tri,
                            int *springs /* [4 n_springs] or NULL */);
int lbfs_membrane_mesh_write(const char *path, int format, int n_nodes, const double *nodes, int n_tri,
                             const int *tri, int n_springs, const int *springs);

/* ---- case files in the reference's XML schema (preprocessor.cpp:87-857) ----
 * The parts the time step reads.  Derived values follow global_variables.cpp:20-68
 * (visc = mach_no * ref_length / reynolds_no; preconditioning "true" with
 * gamma 1 gives gamma = 2 sqrt(3) mach_no) and preprocessor.cpp:419-530 (BC
 * velocities scaled by mach_no / scale). */
#define LBFS_CASE_MAX_PATCHES 32
#define LBFS_CASE_MAX_OBJECTS 64
typedef struct {
    char name[64];
    int rho_type, vel_type;    /* preprocessor.cpp numbering; "empty" -> LBFS_BC_EMPTY; unknown -> 0 */
    double values[4];          /* rho, u, v, w */
} lbfs_case_patch;
typedef struct {
    char name[64];
    int type;                  /* 1 rigid_cylinder, 2 spring_network, 0 other */
    int num_nodes;
    double radius, centre[3];
    char mesh_file[256], stress_free_mesh[256], curvature_mesh[256];
    double shear_modulus, wlc_ratio, volume_modulus, local_area_modulus, global_area_modulus;
    double global_bending_modulus, local_bending_modulus, viscosity, curvature_ratio, volume_ratio;
    double internal_viscosity_ratio;
    int sphere_reference, periodic, mz_importer;
} lbfs_case_object;
typedef struct {
    char import_file[512], import_format[32];
    int mesh_type, gpu_solver, gpu_device, output_step, time_stepping, gradient_type;
    double max_velocity, reynolds_number, visc, gamma, fneq_min, time_marching_step, simulation_length, scale;
    double domain[3], spacing[3], origin[3], streaming_dt, ref_length;
    double average_rho, initial_velocity[3], body_force[3];
    int n_patches;
    lbfs_case_patch patches[LBFS_CASE_MAX_PATCHES];
    int n_objects;
    lbfs_case_object objects[LBFS_CASE_MAX_OBJECTS];
    char error[256];
} lbfs_case;
int lbfs_case_read(const char *path, lbfs_case *out);

/* ---- measurement (no counterpart in the reference; used by bench.py) ----
 * HIP events recorded on the context's stream.  lbfs_time_steps: wall time of
 * nsteps whole steps.  lbfs_time_kernels: average duration per launch of each
 * kernel of the step, names[i] / ms[i] for i < *count (at most max_count). */
int lbfs_time_steps(lbfs_ctx *ctx, int nsteps, double *ms_total);
int lbfs_time_kernels(lbfs_ctx *ctx, int nsteps, int max_count, const char **names, double *ms, int *count);
/* average duration of one fluid pass ("gradients" or "flux_update") over n
 * back-to-back launches on the current state, which they leave unchanged */
int lbfs_time_stage(lbfs_ctx *ctx, const char *name, int n, double *ms_avg);
/* device bytes touched per launch by the model of each kernel (same order as lbfs_time_kernels) */
int lbfs_model_bytes(lbfs_ctx *ctx, int max_count, double *bytes, int *count);

/* ---- host-side mesh helpers (no device work) ---- */
/* counts of a generated nx*ny*nz hex box (unstructured_mesh.cpp:239-666) */
int lbfs_mesh_hex_counts(int nx, int ny, int nz, const int *periodic /* [3] */,
                         int *n_points, int *n_faces, int *n_internal);
/* seeded jittered hex box; patch ids of boundary faces: 0 left 1 right 2 bottom 3 top 4 back 5 front */
int lbfs_mesh_hex_generate(int nx, int ny, int nz, const double *lengths /* [3] */, const double *origin /* [3] */,
                           const int *periodic /* [3] */, double jitter, uint64_t seed,
                           double *points, int *face_vertices, int *owner, int *neighbour,
                           double *face_shift, int *patch);
/* OpenFOAM polyMesh ASCII files in `dir` (points, faces, owner, neighbour,
 * boundary): import_openfoam_mesh (unstructured_mesh.cpp:670-941) and
 * import_openfoam_BC_tags (unstructured_mesh.cpp:2533-2608).  Quad faces only.
 * face_patch: patch index of each face (-1 for internal faces). */
int lbfs_polymesh_counts(const char *dir, int *n_points, int *n_cells, int *n_faces, int *n_internal, int *n_patches);
int lbfs_polymesh_read(const char *dir, double *points, int *face_vertices, int *owner, int *neighbour, int *face_patch,
                       double *face_shift /* [3 n_faces] or NULL; periodic translations from dir/lbfsFaceShift */);
int lbfs_polymesh_patch(const char *dir, int index, char *name, int name_len, char *type, int type_len);
int lbfs_polymesh_write(const char *dir, int n_points, const double *points, int n_cells, int n_faces, int n_internal,
                        const int *face_vertices, const int *owner, const int *neighbour, int n_patches,
                        const char *const *patch_names, const char *const *patch_types, const int *patch_start,
                        const int *patch_size, const double *face_shift /* or NULL */);
/* ---- restart and output files (host only, no TECIO) ----
 * State files <dir>/rho.txt, u.txt, v.txt, w.txt, one line "i ,value" per cell
 * in caller order: Solution::output (Solution.cpp:221-251) and
 * Solution::import (Solution.cpp:987-1064).  precision <= 0: the reference's
 * default 6 significant digits; 17 round-trips a double. */
int lbfs_state_txt_write(const char *dir, int n_cells, const double *rho, const double *u, const double *v,
                         const double *w, int precision);
int lbfs_state_txt_read(const char *dir, int n_cells, double *rho, double *u, double *v, double *w);
/* membrane zone file of tecplot_output::spring_network_output_gpu
 * (tecplot_output.cpp:844-915): header, per node x y z, force, velocity,
 * nodal area (a third of the incident triangles, spring_network.cu:1486),
 * index, curvature (0.5 theta L over the node's springs, convex positive,
 * spring_network.cu:1532), then the 1-based triangle vertex ids one per line.
 * nodal_area / curvature [n_nodes] receive the written columns (or NULL). */
int lbfs_membrane_dat_write(const char *path, int n_nodes, const double *X, const double *F, const double *U, int n_tri,
                            const int *tri, int time, int precision, double *nodal_area, double *curvature);

/* FP64 geometry of a mesh (unstructured_mesh.cpp:1592-1692) */
int lbfs_mesh_geometry(const lbfs_mesh_desc *mesh, double visc, double gamma, double fneq_min, int flags,
                       double *area, double *normal, double *face_centroid, double *face_dt,
                       double *centroid /* [3 (n_cells + n_bc)] */, double *volume);

#ifdef __cplusplus
}
#endif
#endif
