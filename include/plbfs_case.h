/*
 * plbfs_case.h -- C-ABI of the host-side case layer that sits above the solver
 * boundary: OpenFOAM polyMesh reader + FP64 geometry (the reference's
 * unstructured_mesh, src/unstructured_mesh.cpp:26-94,236-507,1133-1233,1912-1987),
 * XML case reading and preprocessing (src/preprocessor.cpp:27-84,187-302,438-515,
 * src/global_variables.cpp:20-56), boundary-condition assignment
 * (src/Boundary_Conditions.cpp:388-452), initial conditions
 * (src/Solution.cpp:94-211), output (src/Solution.cpp:221-251,
 * src/tecplot_output.cpp:273-466) and the whole program run (src/program.cpp:40-137).
 */
#ifndef PLBFS_CASE_H
#define PLBFS_CASE_H

#include <stdint.h>

#include "plbfs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plbfs_mesh plbfs_mesh;

/* Reads <prefix>owner, points, neighbour, faces, boundary (the reference's
 * import_file prefix, which must end in '/') and computes the FP64 geometry from
 * the single-precision points.  visc, gamma and fneq_min set the face streaming
 * dt cap (src/unstructured_mesh.cpp:979-983). */
int plbfs_mesh_read_openfoam(const char *prefix, double visc, double gamma, double fneq_min, plbfs_mesh **out);

/* Same mesh from arrays already in memory (points float32; faces as CSR;
 * boundary faces grouped into n_patches contiguous patches of the given type
 * tokens and sizes). */
int plbfs_mesh_from_arrays(int32_t n_points, const float *px, const float *py, const float *pz, int32_t n_cells,
                           int32_t n_faces, int32_t n_internal, const int32_t *face_off, const int32_t *face_verts,
                           const int32_t *owner, const int32_t *neighbour, int32_t n_patches,
                           const char *const *patch_types, const int32_t *patch_sizes, double visc, double gamma,
                           double fneq_min, plbfs_mesh **out);

void plbfs_mesh_free(plbfs_mesh *m);

/* info: n_cells, n_faces, n_internal_faces, n_points, n_patches, n_ghosts */
int plbfs_mesh_info(const plbfs_mesh *m, int64_t *info6);
/* Pointers into the mesh's own arrays (valid until plbfs_mesh_free). */
int plbfs_mesh_view_get(const plbfs_mesh *m, plbfs_mesh_view *view);
/* Boundary-face patch indices [F-NI]. */
const int32_t *plbfs_mesh_bface_patch(const plbfs_mesh *m);
/* Patch type token (the BC key, src/unstructured_mesh.cpp:1951-1964). */
int plbfs_mesh_patch_type(const plbfs_mesh *m, int32_t patch, char *buf, int32_t len);
/* Raw topology and points as read. */
int plbfs_mesh_raw(const plbfs_mesh *m, float *px, float *py, float *pz, int32_t *owner, int32_t *neighbour_internal,
                   int32_t *face_off, int32_t *face_verts);
int64_t plbfs_mesh_n_face_verts(const plbfs_mesh *m);

/* ---- XML case (schema of src/preprocessor.cpp) ----------------------------- */
typedef struct plbfs_case plbfs_case;

int plbfs_case_load(const char *xml_path, plbfs_case **out);
void plbfs_case_free(plbfs_case *c);
/* Derived solver parameters and text settings. */
int plbfs_case_params(const plbfs_case *c, plbfs_params *p);
/* Import prefix, output directory, time-stepping / gradient tokens. */
int plbfs_case_string(const plbfs_case *c, const char *key, char *buf, int32_t len);
double plbfs_case_number(const plbfs_case *c, const char *key);
/* Patch table matched by type token to the mesh's patches (reference lookup
 * src/Boundary_Conditions.cpp:404-405).  Fails with PLBFS_ERR_BCTYPE on an
 * unknown token (the reference reads past the end of its vector there). */
int plbfs_case_patches(const plbfs_case *c, const plbfs_mesh *m, plbfs_patch *out, int32_t n_patches);
/* Initial conditions over all cells incl. ghosts ([N+G] each). */
int plbfs_case_initial_state(const plbfs_case *c, const plbfs_mesh *m, double *rho, double *u, double *v, double *w);
/* Lagrangian objects: number of objects / nodes, and the node set of object k
 * (reference rigid cylinder, src/lagrangian_object.cpp:79-113, or the sphere
 * with Fibonacci nodes, which has no reference counterpart). */
int plbfs_case_n_objects(const plbfs_case *c);
int plbfs_case_object_nodes(const plbfs_case *c, int32_t k, double *x_ref3, int32_t max_nodes, int32_t *n_out,
                            double *stiffness, double *node_area, int32_t *kernel);

/* ---- Output ------------------------------------------------------------- */
/* Tecplot ASCII in place of the reference's TecIO .plt files.  Grid file
 * (FILETYPE=GRID): node coordinates and polyhedral faces, left element =
 * owner + 1, right = neighbour + 1 or 0 (src/unstructured_mesh.cpp:2073-2200). */
int plbfs_write_tecplot_grid(const plbfs_mesh *m, const char *path);
/* Solution file (FILETYPE=SOLUTION) with the reference's cell-centred
 * variables p u v w x y z vort stf res_rho res_u res_v res_w dt fneq
 * (src/tecplot_output.cpp:273-466): p = rho, velocities over max_velocity,
 * centroids, zero vorticity and stream function (their computation is
 * commented out in the reference), the residual res4 [4*N] (rho, u, v, w per
 * cell, as plbfs_eval_residual gives it; NULL: zeros), local time steps
 * dt_local [N] (NULL: zeros) and tau' = 3 visc / (dt_face gamma) of the last
 * face each cell owns (src/Solver.cpp:680). */
int plbfs_write_tecplot(const plbfs_mesh *m, const char *path, const double *rho, const double *u, const double *v,
                        const double *w, const double *res4, const double *dt_local, double max_velocity,
                        double visc, double gamma, double time);
/* One Lagrangian object (src/tecplot_output.cpp:784-1062): nodx nody nodz
 * (positions over the domain lengths X, Y, Z) vel_x.. force_z per node; a
 * rigid cylinder of radial_nodes x depth_nodes nodes as a FEPOLYGON zone with
 * the reference's edge and element tables, any other node set (0, 0) as
 * points.  vel3, force3 may be NULL (zeros). */
int plbfs_write_tecplot_nodes(const char *path, const char *name, int32_t n, const double *x3, const double *vel3,
                              const double *force3, double X, double Y, double Z, int32_t radial_nodes,
                              int32_t depth_nodes, double time);
/* Full-precision text fields rho/u/v/w.txt, the reference's "i ,value" lines. */
int plbfs_write_fields(const char *dir, int32_t n, const double *rho, const double *u, const double *v,
                       const double *w);
int plbfs_read_fields(const char *dir, int32_t n, double *rho, double *u, double *v, double *w);

/* ---- Whole program (src/main.cpp:18-28, src/program.cpp:40-137) ---------- */
/* Reads the XML, builds mesh/BCs/ICs, runs the MI355X solver, writes outputs.
 * result may be NULL. */
int plbfs_program_run(const char *xml_path, plbfs_result *result);

#ifdef __cplusplus
}
#endif

#endif /* PLBFS_CASE_H */
