/*
 * Declarations of the reference members that the reference-side binding in
 * INTEGRATION.md touches, with the reference's names and types (each line
 * cites the header that declares it).  Only for checking that the binding
 * compiles against those names: nothing here is defined or built.
 */
#ifndef REF_DECLS_SOLVER_GPU_H
#define REF_DECLS_SOLVER_GPU_H
#include <string>
#include <vector>

class Mesh {
public:
    double get_centroid_x(int i);   /* Mesh.h:20 */
    double get_centroid_y(int i);   /* Mesh.h:21 */
    double get_centroid_z(int i);   /* Mesh.h:22 */
    double get_cell_volume(int i);  /* Mesh.h:74 */
};

class unstructured_mesh : public Mesh {
public:
    int get_n_cells();                  /* unstructured_mesh.h:17 */
    int get_n_faces();                  /* unstructured_mesh.h:18 */
    int get_n_neighbours();             /* unstructured_mesh.h:19 */
    std::vector<std::string> bc_types;  /* unstructured_mesh.h:29 */
    int get_mesh_neighbour(int i);      /* unstructured_mesh.h:30 */
    int get_mesh_owner(int i);          /* unstructured_mesh.h:31 */
    double get_face_i(int i);           /* unstructured_mesh.h:38 */
    double get_face_j(int i);           /* unstructured_mesh.h:39 */
    double get_face_k(int i);           /* unstructured_mesh.h:40 */
    double get_face_x(int i);           /* unstructured_mesh.h:41 */
    double get_face_y(int i);           /* unstructured_mesh.h:42 */
    double get_face_z(int i);           /* unstructured_mesh.h:43 */
    double get_face_area(int i);        /* unstructured_mesh.h:44 */
    double get_delta_t_face(int i);     /* unstructured_mesh.h:45 */
};

class Solution {
public:
    double get_rho(int i);              /* Solution.h:22 */
    double get_u(int i);                /* Solution.h:23 */
    double get_v(int i);                /* Solution.h:24 */
    double get_w(int i);                /* Solution.h:25 */
    void set_rho(int i, double arg);    /* Solution.h:30 */
    void set_u(int i, double arg);      /* Solution.h:31 */
    void set_v(int i, double arg);      /* Solution.h:32 */
    void set_w(int i, double arg);      /* Solution.h:33 */
};

class unstructured_bcs {
public:
    std::vector<std::string> name;                              /* unstructured_bcs.h:12 */
    std::vector<int> rho_type, vel_type;                        /* unstructured_bcs.h:13 */
    std::vector<double> rho, u, v, w;                           /* unstructured_bcs.h:14 */
    double per_translate_x, per_translate_y, per_translate_z;   /* unstructured_bcs.h:16 */
};

class global_variables {
public:
    double tolerance;                   /* global_variables.h:34 */
    double pre_conditioned_gamma = 1;   /* global_variables.h:36 */
    double simulation_length = 2000;    /* global_variables.h:37 */
    double time_marching_step;          /* global_variables.h:38 */
    double max_velocity;                /* global_variables.h:40 */
    double ref_rho;                     /* global_variables.h:46 */
    double visc;                        /* global_variables.h:49 */
    double fneq_min = 0.0;              /* global_variables.h:57 */
    std::string gradient_calc_type;     /* global_variables.h:66 */
    std::string time_stepping;          /* global_variables.h:69 */
};

class domain_geometry {
public:
    double dt;                          /* domain_geometry.h:12 */
};

class initial_conditions {
public:
    double pressure_gradient = 0.0;     /* initial_conditions.h:15 */
};

class lagrangian_object {
public:
    int num_nodes;                                          /* lagrangian_object.h:13 */
    double radius, depth;                                   /* lagrangian_object.h:15 */
    double stiffness;                                       /* lagrangian_object.h:17 */
    double *node_x, *node_y, *node_z;                       /* lagrangian_object.h:18 */
    double *node_x_ref, *node_y_ref, *node_z_ref;           /* lagrangian_object.h:19 */
    double *node_vel_x, *node_vel_y, *node_vel_z;           /* lagrangian_object.h:20 */
    double *node_force_x, *node_force_y, *node_force_z;     /* lagrangian_object.h:21 */
};

/* passed through, never touched */
class Boundary_Conditions;
class external_forces;
class post_processing;

class gpu_solver {                      /* Solver_gpu.h:22-31 */
public:
    void General_Purpose_Solver_mk_i(unstructured_mesh &Mesh, Solution &soln, Boundary_Conditions &bc,
                                     external_forces &source, global_variables &globals, domain_geometry &domain,
                                     initial_conditions &init_conds, unstructured_bcs &quad_bcs_orig, int mg,
                                     Solution &residual, int fmg, post_processing &pp,
                                     std::vector<lagrangian_object> &object_vec);
};
#endif
