/*
 * plbfs_hip.h -- C-ABI of the MI355X PLBFS solver (unstructured finite-volume
 * lattice Boltzmann flux solver + immersed boundary method).
 *
 * This is the boundary that replaces the reference's GPU solver call
 *   gpu_solver::General_Purpose_Solver_mk_i(unstructured_mesh&, Solution&,
 *       Boundary_Conditions&, external_forces&, global_variables&, domain_geometry&,
 *       initial_conditions&, unstructured_bcs&, int mg, Solution& residual, int fmg,
 *       post_processing&, std::vector<lagrangian_object>&)
 *   (src/Solver_gpu.h:28-31, called from src/program.cpp:107-110).
 * The reference passes C++ objects; here the same information crosses as plain
 * pointers and sizes.  plbfs_program_run (include/plbfs_case.h) is the same
 * hand-off made from this project's own case reader, in the order of
 * src/program.cpp:40-137; a reference-side caller fills these structs from its
 * unstructured_mesh, Boundary_Conditions, Solution and lagrangian_object.
 *
 * Conventions
 *   - All arrays are in the caller's (OpenFOAM file) numbering.  Cells are
 *     0..n_cells-1, ghost cells are n_cells + (face - n_internal_faces) for
 *     boundary faces (src/unstructured_mesh.cpp:1346-1375).
 *   - Caller-owned host buffers are copied in; the context owns all device
 *     memory.  One context = one device + one HIP stream; a context is not
 *     thread safe.
 *   - Every call returns PLBFS_OK (0) or a negative PLBFS_ERR_* code; the last
 *     error text is available from plbfs_last_error().  Nothing calls exit()
 *     (the reference exits on any CUDA error, src/Solver_gpu.cu:1419-1431).
 */
#ifndef PLBFS_HIP_H
#define PLBFS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLBFS_OK 0
#define PLBFS_ERR_INPUT -1      /* bad sizes, indices or parameters */
#define PLBFS_ERR_NAN -2        /* NaN in rho or u (src/Solver.cpp:448-470) */
#define PLBFS_ERR_RHO -3        /* rho / average_rho > 1000 (src/Solver.cpp:471-493) */
#define PLBFS_ERR_BCTYPE -4     /* boundary-condition type not supported */
#define PLBFS_ERR_HIP -5        /* HIP runtime error */
#define PLBFS_ERR_NODEVICE -6   /* no GPU visible */
#define PLBFS_ERR_IBM -7        /* Lagrangian node left its precomputed support */

/* gradient_calc (src/gradients.cpp:259-269) */
#define PLBFS_GRAD_LS 0
#define PLBFS_GRAD_HYBRID 1
#define PLBFS_GRAD_GG 2
/* time_stepping (src/Solver.cpp:1752-1773) */
#define PLBFS_DT_LOCAL 0
#define PLBFS_DT_CONSTANT 1
#define PLBFS_DT_MIN 2
#define PLBFS_DT_TALTS 3
/* storage precision of fields, gradients and fluxes on the device */
#define PLBFS_FP64 0
#define PLBFS_FP32 1
/* immersed-boundary delta kernels */
#define PLBFS_DELTA_HAT2 0      /* 2-point hat, src/immersed_boundary_method.cu:44-84 */
#define PLBFS_DELTA_PESKIN4 1   /* 4-point Peskin (2002) kernel */

/* Mesh after geometry precompute: the content of the reference's
 * unstructured_mesh object that the solver reads (src/unstructured_mesh.h). */
typedef struct {
    int32_t n_cells;            /* interior cells N (get_n_cells) */
    int32_t n_faces;            /* all faces F (get_n_faces) */
    int32_t n_internal_faces;   /* NI (get_n_neighbours) */
    const int32_t *owner;       /* [F] */
    const int32_t *neighbour;   /* [F] interior neighbour, or ghost index N+(f-NI) */
    const double *cell_centroid;/* [3*(N+F-NI)] xyz interleaved, ghosts included */
    const double *cell_volume;  /* [N] */
    const double *face_centroid;/* [3*F] */
    const double *face_normal;  /* [3*F] unit, owner -> neighbour */
    const double *face_area;    /* [F] */
    const double *face_dt;      /* [F] streaming dt (src/unstructured_mesh.cpp:979-983) */
} plbfs_mesh_view;

/* One boundary patch (src/Boundary_Conditions.cpp:388-452): types use the
 * reference codes of src/preprocessor.cpp:239-258, 1 dirichlet, 2 neumann,
 * 3 periodic, 4 parabolic-N, 5 parabolic-W, 6 inlet, 7 wall, 8 symmetry-x,
 * 9 parabolic, for density and velocity alike.  The ghost update
 * (src/Solution.cpp:759-833) writes u and v of types 4 and 5 from a profile
 * at the centroid of cell b, b the boundary index (domain_x, domain_y of
 * plbfs_params), and writes nothing for density 4, 5, 9, velocity 7, 9 and w
 * of 4, 5: those keep the ghost's initial state (plbfs_set_ghost_state).  The
 * ghost gradient of all of them is the gradient of cell 0, the reference's
 * periodic node of a patch that is not periodic (src/gradients.cpp:434-440,
 * 488-502, src/Boundary_Conditions.cpp:416).  Velocities are already
 * multiplied by mach_no/scale.
 * A periodic face's ghost takes the state and gradient of its partner: the
 * interior cell nearest to the owner's centroid + per_translate
 * (src/Boundary_Conditions.cpp:425-445, whose search starts from an
 * uninitialised cell index; the owner is used here). */
typedef struct {
    int32_t rho_type;
    int32_t vel_type;
    double rho, u, v, w;
    int32_t is_empty;           /* OpenFOAM patch type "empty" (src/Solver.cpp:308) */
    double per_translate[3];    /* periodic_translation (src/preprocessor.cpp:299-301) */
} plbfs_patch;

/* Run parameters (the parts of global_variables / initial_conditions the
 * solver reads; src/global_variables.cpp:20-56, src/preprocessor.cpp:54-84). */
typedef struct {
    double visc;                /* nu = mach_no * ref_length / Re */
    double gamma;               /* pre_conditioned_gamma */
    double fneq_min;
    double cfl;                 /* time_marching_step */
    double max_velocity;        /* mach_no */
    double ref_rho;             /* average_rho */
    double force;               /* uniform x body force = -pressure_gradient */
    double tolerance;           /* convergence threshold on max(u,v,w) rms residual */
    int32_t gradient_type;      /* PLBFS_GRAD_* */
    int32_t time_stepping;      /* PLBFS_DT_* */
    int32_t precision;          /* PLBFS_FP64 or PLBFS_FP32 */
    int32_t check_interval;     /* steps between host convergence checks (>=1) */
    int32_t lsq_exact_det;      /* 0: the reference's LSQ determinant expression
                                   (src/gradients.cpp:282-284, the parity mode);
                                   1: the true 3x3 determinant.  The two agree when
                                   LHS_xy = 0 (axis-aligned hexes); on jittered
                                   tetrahedra the reference expression can flip the
                                   gradient's sign and the run diverges. */
    double domain_x, domain_y;  /* domain.X, domain.Y: geometry x, y times scale
                                   (src/preprocessor.cpp:48,467-468), the lengths
                                   of the parabolic profiles of types 4 and 5 */
} plbfs_params;

/* Lagrangian nodes of the immersed objects (reference: lagrangian_object,
 * src/lagrangian_object.h; GPU update src/immersed_boundary_method.cu:161-190,244-295). */
typedef struct {
    int32_t n_nodes;
    const double *x_ref;        /* [3*n] anchor positions */
    const double *x;            /* [3*n] initial positions */
    double stiffness;           /* spring constant k */
    double node_area;           /* surface area per node dA: F = -k dA (x - x_ref) */
    double dt;                  /* node time step (reference: time_marching_step) */
    double h;                   /* delta-kernel width (mesh spacing) */
    int32_t kernel;             /* PLBFS_DELTA_* */
} plbfs_ibm_nodes;

typedef struct {
    int32_t status;             /* PLBFS_OK, PLBFS_ERR_NAN, PLBFS_ERR_RHO, PLBFS_ERR_IBM */
    int32_t steps;              /* steps taken by this call */
    int32_t converged;          /* 1 if max(u,v,w) residual < tolerance */
    int32_t total_steps;        /* steps since the state was last set */
    double residual[4];         /* rho, u, v, w rms residual of the last step */
    double solve_ms;            /* device time of the time loop (hipEvent) */
} plbfs_result;

typedef struct plbfs_ctx plbfs_ctx;

const char *plbfs_last_error(void);

/* Number of visible HIP devices (0 when none). */
int plbfs_device_count(void);

/* Builds the solver: LSQ precompute with the reference's quirks
 * (src/gradients.cpp:214-313), CFL areas (src/Solver.cpp:1653-1686),
 * space-filling-curve cell order, block face lists, device upload.
 * bface_patch[F-NI] gives the patch of each boundary face. */
int plbfs_create(const plbfs_mesh_view *mesh, const plbfs_params *params, const plbfs_patch *patches,
                 int32_t n_patches, const int32_t *bface_patch, plbfs_ctx **out);
int plbfs_destroy(plbfs_ctx *ctx);

/* Initial primitive state of interior cells (src/Solution.cpp:94-211); the
 * conservative accumulator starts as the same values (src/Solver.cpp:195). */
int plbfs_set_state(plbfs_ctx *ctx, const double *rho, const double *u, const double *v, const double *w);
/* Initial state of the ghost cells, [F - NI] each in boundary-face order (the
 * reference's soln entries n_cells.., src/Solution.cpp:94-211 over all cells).
 * The boundary types that leave a ghost component unwritten (density 4, 5, 9;
 * velocity 7, 9; w of velocity 4, 5) keep these values; other ghosts ignore
 * them.  A context with such a patch refuses to run until this was called. */
int plbfs_set_ghost_state(plbfs_ctx *ctx, const double *rho, const double *u, const double *v, const double *w);
/* Final primitive state (soln) of interior cells. */
int plbfs_get_state(plbfs_ctx *ctx, double *rho, double *u, double *v, double *w);

/* Immersed boundary (optional).  Replaces any previous node set. */
int plbfs_set_ibm(plbfs_ctx *ctx, const plbfs_ibm_nodes *nodes);
/* Node positions, interpolated velocities and forces, each [3*n]. */
int plbfs_get_ibm(plbfs_ctx *ctx, double *x, double *vel, double *force);
/* Eulerian IBM body force per cell [3*N] from the last spreading. */
int plbfs_get_body_force(plbfs_ctx *ctx, double *f);

/* Advances up to max_steps time steps (4 RK stages each), stopping after the
 * first step whose residual is below params.tolerance (src/Solver.cpp:522). */
int plbfs_run(plbfs_ctx *ctx, int32_t max_steps, plbfs_result *out);
/* plbfs_run without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It runs the same
 * kernels with the same results as plbfs_run, but launches them directly (no
 * graph, no events; solve_ms is host wall time) and keeps the control block,
 * per-block residual partials, dt minima and the residual history in a
 * fixed array of 8 MiB per device, held until the process ends.  A call needs
 * about 256 B + 40 B per flux block + 8 B per 256 cells + 32 B per step of it
 * and fails with PLBFS_ERR_INPUT beyond that.  hist_out [4*max_steps] (or
 * NULL) receives the residual history; plbfs_get_residual_history keeps to
 * plbfs_run.  Calls that overlap on one device share the array and race. */
int plbfs_run_noalloc(plbfs_ctx *ctx, int32_t max_steps, plbfs_result *out, double *hist_out);
/* Per-step residual history of the last plbfs_run call, [4*steps]. */
int plbfs_get_residual_history(plbfs_ctx *ctx, double *out, int32_t max_steps);

/* Diagnostics: one residual evaluation (ghost update, gradients, fluxes) of the
 * current state without advancing it.  grads [12*N] (rho,u,v,w x,y,z), R [4*N]. */
int plbfs_eval_residual(plbfs_ctx *ctx, double *grads, double *R);
/* Local time step of every interior cell for the current state [N]. */
int plbfs_get_dt(plbfs_ctx *ctx, double *dt);

/* Per-kernel device timing over the last plbfs_run (hipEvents recorded around
 * every launch when enabled).  name_out receives a NUL-separated list. */
int plbfs_enable_timing(plbfs_ctx *ctx, int32_t on);
int plbfs_kernel_stats(plbfs_ctx *ctx, int32_t which, double *total_ms, int32_t *launches);

/* Layout facts for roofline accounting, written to info[0..9]: cells, faces,
 * face work items per stage (faces on block borders are evaluated by both
 * blocks), stencil slots, boundary faces, flux blocks, most items in one
 * block, gather entries, cells per flux block, LDS bytes per flux block. */
int plbfs_layout_info(plbfs_ctx *ctx, int64_t *info10);

/* ---- Post-processing (src/post_processing.cpp:50-173) ---------------------
 * Surface force on the boundary faces of one patch from the current state,
 * with the reference's sums over "wall" faces: p = rho_owner / 3 gamma,
 * nu' = nu gamma, ghost-cell gradients of u and v,
 *   drag = sum ((p - nu' du/dx) n_x - nu' du/dy n_y) A,
 *   lift = sum (nu' dv/dx n_x - (p - nu' dv/dy) n_y) A.
 * mesh and bface_patch are the arrays given to plbfs_create. */
int plbfs_wall_forces(plbfs_ctx *ctx, const plbfs_mesh_view *mesh, const int32_t *bface_patch, int32_t patch,
                      double *drag_lift2);
/* Force of the fluid on the immersed nodes: minus the sum of their spring
 * forces (the sphere's drag and lift for a body held by the springs). */
int plbfs_ibm_body_force(plbfs_ctx *ctx, double *force3);

/* ---- Batched cases ----------------------------------------------------------
 * B independent cases on one mesh topology, advanced together: the mesh,
 * layout and LSQ weights are shared; each case has its own viscosity (its
 * Reynolds number), state, local time steps and immersed nodes.  A case that
 * converges (or fails) is frozen while the others continue; plbfs_batch_run
 * returns once every case is frozen or max_steps have passed.  The reference
 * has no batched solver: each reference run is one call of
 * gpu_solver::General_Purpose_Solver_mk_i (src/Solver_gpu.h:28-31), and a batch
 * replaces B such calls on one mesh.  Time stepping local or constant.  With
 * fneq_min > 0 the face streaming dt is min(0.5 sqrt(A), 3 nu_b / (gamma
 * fneq_min)) per case b, the mesh builder's formula for that case's viscosity
 * (the mesh view's face_dt is then not used); with fneq_min = 0 it is the
 * view's face_dt. */
typedef struct plbfs_batch plbfs_batch;

int plbfs_batch_create(const plbfs_mesh_view *mesh, const plbfs_params *params, const plbfs_patch *patches,
                       int32_t n_patches, const int32_t *bface_patch, int32_t n_cases, const double *case_visc,
                       plbfs_batch **out);
int plbfs_batch_destroy(plbfs_batch *batch);
/* State of case case_id (file cell order, [N] each); case_id < 0 sets every
 * case.  Resets every case's step count and convergence flag. */
int plbfs_batch_set_state(plbfs_batch *batch, int32_t case_id, const double *rho, const double *u, const double *v,
                          const double *w);
int plbfs_batch_get_state(plbfs_batch *batch, int32_t case_id, double *rho, double *u, double *v, double *w);
/* Immersed nodes of one case; kernel, h and dt must agree across cases. */
int plbfs_batch_set_ibm(plbfs_batch *batch, int32_t case_id, const plbfs_ibm_nodes *nodes);
/* results[n_cases]: per-case status, steps, convergence and residual; the
 * return value is the first failing case's status, or PLBFS_OK. */
int plbfs_batch_run(plbfs_batch *batch, int32_t max_steps, plbfs_result *results);
/* Tiling: cells per flux block (multiple of 64, 64..1024), cases per
 * workgroup (a divisor of 256), and whether the tile stages its own (cell,
 * case) records in LDS. */
int plbfs_batch_set_tiling(plbfs_batch *batch, int32_t block_cells, int32_t cases_per_block, int32_t stage);
/* Times trial_steps steps of every tiling that fits the LDS, keeps the
 * fastest, and restores the state and node positions set before the call. */
int plbfs_batch_autotune(plbfs_batch *batch, int32_t trial_steps, int32_t *block_cells, int32_t *cases_per_block,
                         int32_t *stage, double *ms_per_step);
/* info[0..4]: cases, cells per flux block, cases per workgroup, LDS bytes, staging. */
int plbfs_batch_layout_info(plbfs_batch *batch, int64_t *info5);

#define PLBFS_KSTAT_GRAD 0
#define PLBFS_KSTAT_FLUX 1
#define PLBFS_KSTAT_IBM 2
#define PLBFS_KSTAT_REDUCE 3

#ifdef __cplusplus
}
#endif

#endif /* PLBFS_HIP_H */
