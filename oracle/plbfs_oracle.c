/*
 * plbfs_oracle.c -- CPU FP64 restatement of the reference solver, used ONLY as
 * the reference result in tests.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the reference result (or as
 * the timed CPU baseline).  The product path (3d-plfbs-gpu_amd/) never links it.
 *
 * It restates, in plain single-threaded C, the reference's CPU solver
 * (src/Solver.cpp) together with the mesh geometry (src/unstructured_mesh.cpp),
 * the least-squares gradient precompute and evaluation (src/gradients.cpp), the
 * ghost-cell update (src/Solution.cpp), the residual (src/residuals.cpp) and the
 * RK tables (src/RungeKutta.h).  Every function cites the reference lines it
 * follows.  Quirks are reproduced on purpose:
 *   - points are single precision (unstructured_mesh.h:50, .cpp:357);
 *   - the LSQ determinant is the reference's buggy expression (gradients.cpp:282-284);
 *   - CFL areas are all accumulated into the x slot (Solver.cpp:1679-1681);
 *   - a neighbour index is "boundary" only if strictly greater than n_cells
 *     (Solver.cpp:861), so the first boundary face takes the interior branch;
 *   - the "empty" patch skip tests face > n_internal (Solver.cpp:308);
 *   - the conservative RK accumulator starts as a copy of the primitive state
 *     (Solver.cpp:195);
 *   - the stage coefficients are alpha={1,1/2,1/2,1}, beta={1/6,1/3,1/3,1/6}
 *     applied as U_k = U0 + alpha_k dt R (Solver.cpp:405-438, RungeKutta.h:16-18).
 *
 * The immersed-boundary part restates the reference GPU IBM
 * (src/immersed_boundary_method.cu:13-90,161-190,244-375;
 *  src/Solver_gpu.cu:687-718,766-768), plus the unstructured pair-list
 * formulation with the 4-point Peskin kernel that the product adds (no
 * reference counterpart).
 *
 * Pins (tests/test_oracle_pins.py, tests/test_ibm_oracle.py, tests/test_host.py;
 * the reference ships no tests, fixtures or sample meshes and cannot be built
 * here, so closed forms and invariants do the pinning):
 *   build_geometry     cube closed forms, ghost mirror, tet closure, total
 *                      volume on jittered meshes, single-precision points;
 *   build_lsq          inv*LHS = I (true determinant) or s*I (reference
 *                      expression), exact gradients of linear fields on
 *                      axis-aligned hexes (where the reference expression is exact);
 *   build_cfl_areas, get_cfl   closed-form dt on uniform cubes (x-slot areas);
 *   face_flux          Euler flux of a uniform state, zero residual at rest,
 *                      global conservation against the boundary fluxes;
 *   om_run (RK)        amplification polynomial 1+Z+7/12Z^2+5/24Z^3+1/24Z^4,
 *                      body force adds dt*f per step;
 *   IBM                hat weights equal the reference's structured formula,
 *                      linear fields reproduced, force conservation, adjointness,
 *                      node motion X += U dt;
 *   update_bcs         ghost states of types 1, 2, 6 and 7: the uniform free
 *                      stream is steady in the channel whichever of them the
 *                      inlet uses, and a density jump against the inlet patch
 *                      leaves a residual there for 6 and 7 (patch value) but
 *                      not for 2 (owner value);
 *   ghost gradients    type 2 keeps the sign of n in (1 - n): a field mirror-
 *                      symmetric about a cube's mid-plane gets a residual that
 *                      is symmetric except in the cells on the two Neumann walls;
 *   type 8 (symmetry-x) the ghost is the owner's mirror image (rho, u, w copied,
 *                      v negated: exact gradients of even/odd linear fields on
 *                      the plane), and its velocity gradient is the Dirichlet
 *                      one toward the patch value (a uniform stream is steady
 *                      when the patch holds it, slowed or sped up in proportion
 *                      to the difference otherwise);
 *   type 7 density     the ghost gradient is the owner's: with the patch value
 *                      on a linear field, the boundary flux equals the flux of
 *                      the same face inside a longer box (types 1, 2, 6 do not);
 *   empty patches      "face > n_internal": on a one-cell-thick box with empty x
 *                      sides every residual equals the periodic-x one except
 *                      the owner of boundary face 0, which keeps that face's flux;
 *   "neighbour > n_cells" with y, z periodic and a quadratic density of x, every
 *                      x = 0 face gives the same flux except boundary face 0,
 *                      which takes the two-sided branch;
 *   LSQ determinant    on hexes sheared in x by y (LHS_xz = LHS_yz = 0) the
 *                      reference inverse is the true one times
 *                      (xx yy - xy^2)/(xx yy + xy^2): the sign of its xy term;
 *                      sheared in x by y and z (LHS_yz = 0), the reference
 *                      expression is zz (xx yy + xy^2) - xy xz - yy xz^2: the
 *                      "- yz - zx" part (tests/test_oracle_flux_pins.py);
 *   face_flux (non-equilibrium part, tests/test_oracle_flux_pins.py)
 *                      a linear shear on uniform hexes gives the closed-form
 *                      moments (tau' = 3 nu/(dt gamma), offsets -e.grad dt,
 *                      1/gamma in feq) at gamma 1 and 0.3; on jittered hexes
 *                      and tets every interior face equals the flux written
 *                      from its definition (lattice point x_f - e dt from the
 *                      side of the face plane it lies on, rest point the mean
 *                      of the two face values), and on unjittered hexes points
 *                      on the face plane take the owner side;
 *   ghost gradients    inlet (6) faces: zero; symmetry-x (8) faces: zero
 *                      density slot, Dirichlet velocity formula (the boundary
 *                      flux equals the definition with these gradients);
 *   parabolic ghosts   types 4 and 5 at the centroid of cell b (b the boundary
 *                      index) on an unpermuted box, u of type 5 with (X - y),
 *                      w and type 9 at the initial state.
 * Every pin in tests/test_oracle_flux_pins.py also rejects a copy of this file
 * built with a planted mistake (OM_PLANT below).  No function is left
 * unpinned.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OM_OK 0
#define OM_ERR_INPUT -1
#define OM_ERR_NAN -2
#define OM_ERR_RHO -3
#define OM_ERR_BCTYPE -4

/* Planted mistakes, for the tests that show each pin fails when the
 * restatement is wrong in a plausible way.  The library the tests use as the
 * oracle is built with OM_PLANT 0 (none); oracle/Makefile's `plants` target
 * builds one library per mistake:
 *   1 tau' = 3 nu / dt without the 1/gamma      2 tau' doubled
 *   3 lattice offsets +e.grad dt (sign flipped)  4 lattice side flipped (neighbour iff e.n > 0)
 *   5 rest point from the owner side only        6 LSQ determinant "- yz * zx" for "- yz - zx"
 *   7 type-6 ghost gradients copy the owner's    8 type-8 density ghost gradient copies the owner's
 *   9 parabolic ghosts at the owner's centroid instead of cell i's
 *  10 type-5 ghost u with (X - x) for (X - y)
 *  11 lattice points on the face plane (e.n = 0) from the neighbour side */
#ifndef OM_PLANT
#define OM_PLANT 0
#endif

/* ------------------------------------------------------------------------- */
/* Inputs                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n_points, n_cells, n_faces, n_internal;
    const float *px, *py, *pz;     /* points, single precision */
    const int *face_off;           /* [n_faces+1] offsets into face_verts */
    const int *face_verts;
    const int *owner;              /* [n_faces] */
    const int *neighbour;          /* [n_internal] */
    const int *bface_patch;        /* [n_faces-n_internal] patch index per boundary face */
} om_mesh_in;

typedef struct {
    int rho_type, vel_type;        /* reference codes 1..9 (preprocessor.cpp:243-290) */
    double rho, u, v, w;           /* BC values, velocity already scaled by mach factor */
    int is_empty;                  /* OpenFOAM patch type "empty" */
    double per_translate[3];       /* periodic (type 3): partner = nearest cell to owner + this */
} om_patch;

typedef struct {
    double visc;                   /* nu = U L / Re (global_variables.cpp:24) */
    double gamma;                  /* pre_conditioned_gamma */
    double fneq_min;
    double cfl;                    /* time_marching_step */
    double max_velocity;           /* U = mach_no */
    double ref_rho;                /* average_rho */
    double force;                  /* -pressure_gradient (Solver.cpp:213) */
    double tolerance;
    int gradient_type;             /* 0 LS, 1 hybrid, 2 Green-Gauss */
    int time_stepping;             /* 0 local, 1 constant, 2 min, 3 talts */
    int lsq_exact_det;             /* 0: reference determinant expression; 1: true determinant */
    double domain_x, domain_y;     /* domain.X, domain.Y (geometry x, y times scale, preprocessor.cpp:48,467-468) */
} om_params;

/* ------------------------------------------------------------------------- */
/* Context                                                                   */
/* ------------------------------------------------------------------------- */

typedef struct {
    om_mesh_in m;
    om_params p;
    int n_patches;
    om_patch *patch;
    int N, G, T, F, NI;            /* cells, ghosts, total, faces, internal faces */
    int *nbr;                      /* [F] neighbour incl. ghost index for boundary faces */
    int *partner;                  /* [G] periodic partner cell (Boundary_Conditions.cpp:428-445), else 0 */
    /* geometry */
    double *cx, *cy, *cz, *vol, *tot_area;   /* [T] */
    double *fx, *fy, *fz, *fa, *ni, *nj, *nk, *fdt; /* [F] */
    int *st_off, *st_cell, *st_face;         /* stencil CSR over interior cells */
    /* LSQ */
    double *Lxx, *Lxy, *Lxz, *Lyx, *Lyy, *Lyz, *Lzx, *Lzy, *Lzz; /* [N] inverse */
    double *Rx, *Ry, *Rz;                    /* [N*6] */
    double *cfl_area;                        /* [N] x slot only */
    /* state */
    double *s_rho, *s_u, *s_v, *s_w;         /* soln [T] */
    double *t_rho, *t_u, *t_v, *t_w;         /* temp [T] */
    double *a_rho, *a_u, *a_v, *a_w;         /* soln_t0 [T] */
    double *b_rho, *b_u, *b_v, *b_w;         /* soln_t1 [T] */
    double *r_rho, *r_u, *r_v, *r_w;         /* residual [T] */
    double *g;                               /* gradients [T][12]: rho xyz,u xyz,v xyz,w xyz */
    double *dtl;                             /* [N] */
    double *force_x, *force_y, *force_z;     /* IBM body force [N], added to R (GPU convention) */
    int use_body_force;
    /* immersed boundary (Lagrangian nodes) */
    int ibm_n, ibm_kernel, ibm_on;
    double ibm_h, ibm_dt, ibm_k, ibm_dA;
    double *ibm_x, *ibm_x0, *ibm_xref, *ibm_u, *ibm_f; /* [3*n] */
    double res[4];                           /* last residual rho,u,v,w */
    int steps_done;
} om_ctx;

static const int EX[15] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
static const int EY[15] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1};
static const int EZ[15] = {0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1};

/* lattice weights (Solver.cpp:1409-1445) */
static double lattice_weight(int k) {
    if (k > 0 && k < 7) return 1.0 / 9.0;
    if (k > 6) return 1.0 / 72.0;
    return 2.0 / 9.0;
}

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

void om_destroy(om_ctx *c);

/* ------------------------------------------------------------------------- */
/* Geometry: unstructured_mesh.cpp:577-607 (tet_volume), 732-766          */
/* (cell_tet_area), 979-983 (face dt), 1133-1233 (generate_internal_mesh), */
/* 1346-1375 (ghost centroid); vector_var.cpp (factor divides, average /3) */
/* ------------------------------------------------------------------------- */

static double mag3(double x, double y, double z) { return sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2)); }

static void build_geometry(om_ctx *c) {
    const om_mesh_in *m = &c->m;
    int N = c->N, F = c->F, NI = c->NI;
    /* stencil counts */
    int *cnt = xcalloc(N, sizeof(int));
    for (int f = 0; f < F; f++) {
        cnt[m->owner[f]]++;
        if (f < NI) cnt[m->neighbour[f]]++;
    }
    c->st_off = xcalloc(N + 1, sizeof(int));
    for (int i = 0; i < N; i++) c->st_off[i + 1] = c->st_off[i] + cnt[i];
    c->st_cell = xcalloc(c->st_off[N], sizeof(int));
    c->st_face = xcalloc(c->st_off[N], sizeof(int));
    memset(cnt, 0, sizeof(int) * N);

    for (int f = 0; f < F; f++) {
        int nv = m->face_off[f + 1] - m->face_off[f];
        const int *fv = m->face_verts + m->face_off[f];
        int o = m->owner[f];
        double ax = m->px[fv[0]], ay = m->py[fv[0]], az = m->pz[fv[0]];
        /* tet_volume over the fan (unstructured_mesh.cpp:1149-1151, 577-607) */
        for (int t = 0; t < nv - 2; t++) {
            double bx = m->px[fv[t + 1]] - ax, by = m->py[fv[t + 1]] - ay, bz = m->pz[fv[t + 1]] - az;
            double qx = m->px[fv[t + 2]] - ax, qy = m->py[fv[t + 2]] - ay, qz = m->pz[fv[t + 2]] - az;
            double xx = by * qz - bz * qy, yy = bz * qx - bx * qz, zz = bx * qy - by * qx;
            double v = (ax * xx + ay * yy + az * zz) / 6;
            c->vol[o] = c->vol[o] + v;
            if (f < NI) c->vol[m->neighbour[f]] = c->vol[m->neighbour[f]] - v;
        }
        /* cell_tet_area over the fan (unstructured_mesh.cpp:1158-1160, 732-766) */
        double fa = 0.0, fcx = 0.0, fcy = 0.0, fcz = 0.0, fnx = 0.0, fny = 0.0, fnz = 0.0;
        for (int t = 0; t < nv - 2; t++) {
            double bx0 = m->px[fv[t + 1]], by0 = m->py[fv[t + 1]], bz0 = m->pz[fv[t + 1]];
            double qx0 = m->px[fv[t + 2]], qy0 = m->py[fv[t + 2]], qz0 = m->pz[fv[t + 2]];
            double dx = (ax + bx0 + qx0) / 3, dy = (ay + by0 + qy0) / 3, dz = (az + bz0 + qz0) / 3;
            double bx = bx0 - ax, by = by0 - ay, bz = bz0 - az;
            double qx = qx0 - ax, qy = qy0 - ay, qz = qz0 - az;
            double xx = by * qz - bz * qy, yy = bz * qx - bx * qz, zz = bx * qy - by * qx;
            double area = 0.5 * mag3(xx, yy, zz);
            fa = fa + area;
            double inv = 1 / area;
            dx = dx / inv; dy = dy / inv; dz = dz / inv;
            fcx = fcx + dx; fcy = fcy + dy; fcz = fcz + dz;
            fnx = fnx + xx; fny = fny + yy; fnz = fnz + zz;
        }
        c->fa[f] = fa;
        c->tot_area[o] = c->tot_area[o] + fa;
        c->cx[o] = c->cx[o] + fcx; c->cy[o] = c->cy[o] + fcy; c->cz[o] = c->cz[o] + fcz;
        c->fx[f] = fcx / fa; c->fy[f] = fcy / fa; c->fz[f] = fcz / fa;
        double nm = mag3(fnx, fny, fnz);
        c->ni[f] = fnx / nm; c->nj[f] = fny / nm; c->nk[f] = fnz / nm;
        /* calc_face_delta_t (unstructured_mesh.cpp:979-983) */
        double dtf = 0.5 * sqrt(fa);
        double refdt = c->p.visc * 3 / c->p.gamma / c->p.fneq_min;
        c->fdt[f] = (refdt < dtf) ? refdt : dtf;
        if (f > NI - 1) {
            int gh = f - NI + N;
            c->st_cell[c->st_off[o] + cnt[o]] = gh;
            c->st_face[c->st_off[o] + cnt[o]] = f;
            cnt[o]++;
            c->nbr[f] = gh;
        } else {
            int nb = m->neighbour[f];
            c->tot_area[nb] = c->tot_area[nb] + fa;
            c->cx[nb] = c->cx[nb] + fcx; c->cy[nb] = c->cy[nb] + fcy; c->cz[nb] = c->cz[nb] + fcz;
            c->st_cell[c->st_off[o] + cnt[o]] = nb;
            c->st_face[c->st_off[o] + cnt[o]] = f;
            cnt[o]++;
            c->st_cell[c->st_off[nb] + cnt[nb]] = o;
            c->st_face[c->st_off[nb] + cnt[nb]] = f;
            cnt[nb]++;
            c->nbr[f] = nb;
        }
    }
    for (int i = 0; i < N; i++) {
        c->cx[i] = c->cx[i] / c->tot_area[i];
        c->cy[i] = c->cy[i] / c->tot_area[i];
        c->cz[i] = c->cz[i] / c->tot_area[i];
    }
    /* ghost centroids (unstructured_mesh.cpp:1346-1375) */
    for (int f = NI; f < F; f++) {
        int gh = N + f - NI, j = m->owner[f];
        double rx = c->fx[f] - c->cx[j], ry = c->fy[f] - c->cy[j], rz = c->fz[f] - c->cz[j];
        rx = rx / 0.5; ry = ry / 0.5; rz = rz / 0.5;
        c->cx[gh] = c->cx[j] + rx * fabs(c->ni[f]);
        c->cy[gh] = c->cy[j] + ry * fabs(c->nj[f]);
        c->cz[gh] = c->cz[j] + rz * fabs(c->nk[f]);
    }
    free(cnt);
}

/* ------------------------------------------------------------------------- */
/* LSQ precompute: gradients.cpp:214-313 with add_non_macrovariable_         */
/* contributions gradients.cpp:149-210 (buggy determinant kept)              */
/* ------------------------------------------------------------------------- */

static void build_lsq(om_ctx *c) {
    int N = c->N;
    for (int i = 0; i < N; i++) {
        double xx = 0, xy = 0, xz = 0, yx = 0, yy = 0, yz = 0, zx = 0, zy = 0, zz = 0;
        double max_area = 0.0, max_disp = 0.0;
        for (int s = c->st_off[i]; s < c->st_off[i + 1]; s++) {
            int i1 = c->st_cell[s], f = c->st_face[s];
            max_area = fmax(max_area, c->fa[f]);
            double dx = c->cx[i1] - c->cx[i], dy = c->cy[i1] - c->cy[i], dz = c->cz[i1] - c->cz[i];
            double disp = sqrt(dx * dx + dy * dy + dz * dz);
            max_disp = (max_disp < disp) ? disp : max_disp; /* std::max(disp,max_disp) */
        }
        double ar = 2 * fabs(max_disp) * max_area / c->vol[i];
        double beta;
        if (c->p.gradient_type == 1) {
            double q = 2 / ar;
            beta = (q < 1.0) ? q : 1.0; /* std::min(1.0, 2/ar) */
        } else if (c->p.gradient_type == 0) {
            beta = 1.0;
        } else {
            beta = 0.0;
        }
        xx = xx + 2 * (1 - beta) * c->vol[i];
        yy = yy + 2 * (1 - beta) * c->vol[i];
        zz = zz + 2 * (1 - beta) * c->vol[i];
        int slot = 0;
        for (int s = c->st_off[i]; s < c->st_off[i + 1]; s++, slot++) {
            int i1 = c->st_cell[s], f = c->st_face[s];
            double dx = c->fx[f] - c->cx[i], dy = c->fy[f] - c->cy[i], dz = c->fz[f] - c->cz[i];
            double l_norm = dx * c->ni[f] + dy * c->nj[f] + dz * c->nk[f];
            dx = c->cx[i1] - c->cx[i]; dy = c->cy[i1] - c->cy[i]; dz = c->cz[i1] - c->cz[i];
            double L = sqrt(pow(dx, 2.0) + pow(dy, 2.0) + pow(dz, 2.0));
            double sa = c->fa[f];
            double L_norm = dx * c->ni[f] + dy * c->nj[f] + dz * c->nk[f];
            double alpha_w = pow(2 * l_norm / L_norm, 2.0);
            double w = alpha_w * sa / L;
            xx = xx + w * dx * dx * beta; xy = xy + w * dx * dy * beta; xz = xz + w * dx * dz * beta;
            yx = yx + w * dy * dx * beta; yy = yy + w * dy * dy * beta; yz = yz + w * dy * dz * beta;
            zx = zx + w * dz * dx * beta; zy = zy + w * dz * dy * beta; zz = zz + w * dz * dz * beta;
            if (slot < 6) {
                if (c->m.owner[f] == i) {
                    c->Rx[i * 6 + slot] = (beta * w * dx + 2 * (1 - beta) * 0.5 * c->ni[f] * sa);
                    c->Ry[i * 6 + slot] = (beta * w * dy + 2 * (1 - beta) * 0.5 * c->nj[f] * sa);
                    c->Rz[i * 6 + slot] = (beta * w * dz + 2 * (1 - beta) * 0.5 * c->nk[f] * sa);
                } else {
                    c->Rx[i * 6 + slot] = (beta * w * dx + 2 * (1 - beta) * 0.5 * c->ni[f] * -1 * sa);
                    c->Ry[i * 6 + slot] = (beta * w * dy + 2 * (1 - beta) * 0.5 * c->nj[f] * -1 * sa);
                    c->Rz[i * 6 + slot] = (beta * w * dz + 2 * (1 - beta) * 0.5 * c->nk[f] * -1 * sa);
                }
            }
        }
        double det = xx * (yy * zz - yz * zy) + xy * (yx * zz - yz - zx) + xz * (yx * zy - yy * zx);
        if (OM_PLANT == 6) det = xx * (yy * zz - yz * zy) + xy * (yx * zz - yz * zx) + xz * (yx * zy - yy * zx);
        if (c->p.lsq_exact_det) det = xx * (yy * zz - yz * zy) - xy * (yx * zz - yz * zx) + xz * (yx * zy - yy * zx);
        double a = xx, b = xy, cc = xz, d = yx, e = yy, ff = yz, g = zx, h = zy, k = zz;
        c->Lxx[i] = 1 / det * (e * k - ff * h);
        c->Lxy[i] = 1 / det * (cc * h - b * k);
        c->Lxz[i] = 1 / det * (b * ff - cc * e);
        c->Lyx[i] = 1 / det * (ff * g - d * k);
        c->Lyy[i] = 1 / det * (a * k - cc * g);
        c->Lyz[i] = 1 / det * (cc * d - a * ff);
        c->Lzx[i] = 1 / det * (d * h - e * g);
        c->Lzy[i] = 1 / det * (b * g - a * h);
        c->Lzz[i] = 1 / det * (a * e - b * d);
    }
}

/* populate_cfl_areas (Solver.cpp:1653-1686): all three sums into the x slot */
static void build_cfl_areas(om_ctx *c) {
    for (int i = 0; i < c->N; i++) {
        double ax = 0, ay = 0, az = 0;
        for (int s = c->st_off[i]; s < c->st_off[i + 1]; s++) {
            int f = c->st_face[s];
            ax = ax + fabs(c->ni[f] * c->fa[f]);
            ay = ay + fabs(c->nj[f] * c->fa[f]);
            az = az + fabs(c->nk[f] * c->fa[f]);
        }
        double v = 0.0;
        v = v + ax / 2;
        v = v + ay / 2;
        v = v + az / 2;
        c->cfl_area[i] = v;
    }
}

/* ------------------------------------------------------------------------- */
/* Periodic partners: Boundary_Conditions.cpp:425-445.  The reference takes   */
/* the centroid of an uninitialised index there; the owner of the boundary    */
/* face is used here.  Partner = the interior cell nearest to owner centroid  */
/* + per_translate (first minimum in cell order); 0 for other patches, the    */
/* reference's default.                                                       */
/* ------------------------------------------------------------------------- */

static void build_partners(om_ctx *c) {
    c->partner = xcalloc(c->G ? c->G : 1, sizeof(int));
    for (int b = 0; b < c->G; b++) {
        const om_patch *P = &c->patch[c->m.bface_patch[b]];
        if (P->rho_type != 3 && P->vel_type != 3) continue;
        int o = c->m.owner[c->NI + b];
        double x = c->cx[o] + P->per_translate[0], y = c->cy[o] + P->per_translate[1],
               z = c->cz[o] + P->per_translate[2];
        double best = 10000000;
        for (int i = 0; i < c->N; i++) {
            double dx = c->cx[i] - x, dy = c->cy[i] - y, dz = c->cz[i] - z;
            double d = sqrt(pow(dx, 2) + pow(dy, 2) + pow(dz, 2));
            if (d < best) {
                best = d;
                c->partner[b] = i;
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Creation                                                                  */
/* ------------------------------------------------------------------------- */

/* Type codes 1..9 of preprocessor.cpp:243-262.  Codes the ghost update does
 * not write (density 4, 5, 9; velocity 7, 9, and w of 4 and 5) leave the
 * ghost at its value in soln, which no step writes: the initial state. */
static int check_patch_types(const om_patch *p, int n) {
    for (int i = 0; i < n; i++) {
        int r = p[i].rho_type, v = p[i].vel_type;
        if (r < 1 || r > 9 || v < 1 || v > 9) return OM_ERR_BCTYPE;
    }
    return OM_OK;
}

om_ctx *om_create(const om_mesh_in *m, const om_params *p, const om_patch *patches, int n_patches, int *err) {
    *err = OM_OK;
    if (!m || !p || m->n_cells <= 0 || m->n_faces <= 0 || m->n_internal < 0 || m->n_internal > m->n_faces) {
        *err = OM_ERR_INPUT;
        return NULL;
    }
    if (check_patch_types(patches, n_patches) != OM_OK) {
        *err = OM_ERR_BCTYPE;
        return NULL;
    }
    om_ctx *c = xcalloc(1, sizeof(om_ctx));
    c->m = *m;
    c->p = *p;
    c->n_patches = n_patches;
    c->patch = xcalloc(n_patches, sizeof(om_patch));
    memcpy(c->patch, patches, sizeof(om_patch) * n_patches);
    c->N = m->n_cells; c->F = m->n_faces; c->NI = m->n_internal;
    c->G = c->F - c->NI; c->T = c->N + c->G;
    for (int b = 0; b < c->G; b++) {
        if (m->bface_patch[b] < 0 || m->bface_patch[b] >= n_patches) {
            *err = OM_ERR_INPUT;
            om_destroy(c);
            return NULL;
        }
    }
    int T = c->T, F = c->F, N = c->N;
    c->nbr = xcalloc(F, sizeof(int));
    double **cells[] = {&c->cx, &c->cy, &c->cz, &c->vol, &c->tot_area,
                        &c->s_rho, &c->s_u, &c->s_v, &c->s_w, &c->t_rho, &c->t_u, &c->t_v, &c->t_w,
                        &c->a_rho, &c->a_u, &c->a_v, &c->a_w, &c->b_rho, &c->b_u, &c->b_v, &c->b_w,
                        &c->r_rho, &c->r_u, &c->r_v, &c->r_w};
    for (size_t q = 0; q < sizeof(cells) / sizeof(cells[0]); q++) *cells[q] = xcalloc(T, sizeof(double));
    double **faces[] = {&c->fx, &c->fy, &c->fz, &c->fa, &c->ni, &c->nj, &c->nk, &c->fdt};
    for (size_t q = 0; q < sizeof(faces) / sizeof(faces[0]); q++) *faces[q] = xcalloc(F, sizeof(double));
    double **inv[] = {&c->Lxx, &c->Lxy, &c->Lxz, &c->Lyx, &c->Lyy, &c->Lyz, &c->Lzx, &c->Lzy, &c->Lzz,
                      &c->cfl_area, &c->dtl, &c->force_x, &c->force_y, &c->force_z};
    for (size_t q = 0; q < sizeof(inv) / sizeof(inv[0]); q++) *inv[q] = xcalloc(N, sizeof(double));
    c->Rx = xcalloc((size_t)N * 6, sizeof(double));
    c->Ry = xcalloc((size_t)N * 6, sizeof(double));
    c->Rz = xcalloc((size_t)N * 6, sizeof(double));
    c->g = xcalloc((size_t)T * 12, sizeof(double));
    build_geometry(c);
    build_lsq(c);
    build_cfl_areas(c);
    build_partners(c);
    return c;
}

void om_destroy(om_ctx *c) {
    if (!c) return;
    void *ptrs[] = {c->patch, c->nbr, c->partner, c->cx, c->cy, c->cz, c->vol, c->tot_area, c->fx, c->fy, c->fz, c->fa,
                    c->ni, c->nj, c->nk, c->fdt, c->st_off, c->st_cell, c->st_face, c->Lxx, c->Lxy, c->Lxz,
                    c->Lyx, c->Lyy, c->Lyz, c->Lzx, c->Lzy, c->Lzz, c->Rx, c->Ry, c->Rz, c->cfl_area,
                    c->s_rho, c->s_u, c->s_v, c->s_w, c->t_rho, c->t_u, c->t_v, c->t_w, c->a_rho, c->a_u,
                    c->a_v, c->a_w, c->b_rho, c->b_u, c->b_v, c->b_w, c->r_rho, c->r_u, c->r_v, c->r_w,
                    c->g, c->dtl, c->force_x, c->force_y, c->force_z, c->ibm_x, c->ibm_x0, c->ibm_xref,
                    c->ibm_u, c->ibm_f};
    for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); q++) free(ptrs[q]);
    free(c);
}

/* ------------------------------------------------------------------------- */
/* Accessors for tests                                                       */
/* ------------------------------------------------------------------------- */

int om_n_total(const om_ctx *c) { return c->T; }
int om_stencil_size(const om_ctx *c) { return c->st_off[c->N]; }

/* cell geometry over all T cells: centroid xyz, volume */
void om_get_cells(const om_ctx *c, double *cx, double *cy, double *cz, double *vol) {
    memcpy(cx, c->cx, sizeof(double) * c->T);
    memcpy(cy, c->cy, sizeof(double) * c->T);
    memcpy(cz, c->cz, sizeof(double) * c->T);
    memcpy(vol, c->vol, sizeof(double) * c->T);
}

/* face geometry: centroid xyz, unit normal, area, streaming dt, neighbour incl. ghosts */
void om_get_faces(const om_ctx *c, double *fx, double *fy, double *fz, double *ni, double *nj, double *nk,
                  double *fa, double *fdt, int *nbr) {
    size_t F = (size_t)c->F;
    memcpy(fx, c->fx, sizeof(double) * F); memcpy(fy, c->fy, sizeof(double) * F);
    memcpy(fz, c->fz, sizeof(double) * F); memcpy(ni, c->ni, sizeof(double) * F);
    memcpy(nj, c->nj, sizeof(double) * F); memcpy(nk, c->nk, sizeof(double) * F);
    memcpy(fa, c->fa, sizeof(double) * F); memcpy(fdt, c->fdt, sizeof(double) * F);
    memcpy(nbr, c->nbr, sizeof(int) * F);
}

/* LSQ inverse [N*9] row-major, RHS [N*6*3] (slot-major xyz), cfl area [N] */
void om_get_lsq(const om_ctx *c, double *inv9, double *rhs, double *cfl_area) {
    for (int i = 0; i < c->N; i++) {
        double *q = inv9 + (size_t)i * 9;
        q[0] = c->Lxx[i]; q[1] = c->Lxy[i]; q[2] = c->Lxz[i];
        q[3] = c->Lyx[i]; q[4] = c->Lyy[i]; q[5] = c->Lyz[i];
        q[6] = c->Lzx[i]; q[7] = c->Lzy[i]; q[8] = c->Lzz[i];
        for (int s = 0; s < 6; s++) {
            rhs[((size_t)i * 6 + s) * 3 + 0] = c->Rx[i * 6 + s];
            rhs[((size_t)i * 6 + s) * 3 + 1] = c->Ry[i * 6 + s];
            rhs[((size_t)i * 6 + s) * 3 + 2] = c->Rz[i * 6 + s];
        }
        cfl_area[i] = c->cfl_area[i];
    }
}

void om_get_partner(const om_ctx *c, int *out) { memcpy(out, c->partner, sizeof(int) * (size_t)c->G); }

void om_get_stencil(const om_ctx *c, int *off, int *cell, int *face) {
    memcpy(off, c->st_off, sizeof(int) * (c->N + 1));
    memcpy(cell, c->st_cell, sizeof(int) * c->st_off[c->N]);
    memcpy(face, c->st_face, sizeof(int) * c->st_off[c->N]);
}

/* Initial primitive state over T cells (soln), and soln_t1 = clone(soln) (Solver.cpp:195). */
void om_set_state(om_ctx *c, const double *rho, const double *u, const double *v, const double *w) {
    size_t T = (size_t)c->T;
    memcpy(c->s_rho, rho, sizeof(double) * T); memcpy(c->s_u, u, sizeof(double) * T);
    memcpy(c->s_v, v, sizeof(double) * T); memcpy(c->s_w, w, sizeof(double) * T);
    memcpy(c->b_rho, rho, sizeof(double) * T); memcpy(c->b_u, u, sizeof(double) * T);
    memcpy(c->b_v, v, sizeof(double) * T); memcpy(c->b_w, w, sizeof(double) * T);
    c->steps_done = 0;
}

void om_get_state(const om_ctx *c, double *rho, double *u, double *v, double *w) {
    size_t N = (size_t)c->N;
    memcpy(rho, c->s_rho, sizeof(double) * N); memcpy(u, c->s_u, sizeof(double) * N);
    memcpy(v, c->s_v, sizeof(double) * N); memcpy(w, c->s_w, sizeof(double) * N);
}

void om_get_residual(const om_ctx *c, double *res4) { memcpy(res4, c->res, sizeof(double) * 4); }

/* Body force per interior cell added to R every stage (reference GPU IBM convention,
 * Solver_gpu.cu:766-768).  Pass NULL to switch it off. */
void om_set_body_force(om_ctx *c, const double *fxv, const double *fyv, const double *fzv) {
    if (!fxv) {
        c->use_body_force = 0;
        return;
    }
    memcpy(c->force_x, fxv, sizeof(double) * c->N);
    memcpy(c->force_y, fyv, sizeof(double) * c->N);
    memcpy(c->force_z, fzv, sizeof(double) * c->N);
    c->use_body_force = 1;
}

/* ------------------------------------------------------------------------- */
/* Ghost-cell update: Solution.cpp:759-833 (handled types only)             */
/* ------------------------------------------------------------------------- */

static void update_bcs(om_ctx *c, double *rho, double *u, double *v, double *w) {
    for (int b = 0; b < c->G; b++) {
        int j = b + c->N, face = b + c->NI, nb = c->m.owner[face], pn = c->partner[b];
        const om_patch *P = &c->patch[c->m.bface_patch[b]];
        double dx = 0;
        if (P->rho_type == 1) rho[j] = P->rho - (rho[nb] - P->rho);
        else if (P->rho_type == 2) rho[j] = rho[nb] + dx * P->rho;
        else if (P->rho_type == 3) rho[j] = rho[pn];
        else if (P->rho_type == 6) rho[j] = P->rho;
        else if (P->rho_type == 7) rho[j] = P->rho;
        else if (P->rho_type == 8) rho[j] = rho[nb];
        if (P->vel_type == 1) {
            u[j] = P->u - (u[nb] - P->u);
            v[j] = P->v - (v[nb] - P->v);
            w[j] = P->w - (w[nb] - P->w);
        } else if (P->vel_type == 2) {
            u[j] = u[nb] + dx * P->u;
            v[j] = v[nb] + dx * P->v;
            w[j] = w[nb] + dx * P->w;
        } else if (P->vel_type == 3) {
            u[j] = u[pn];
            v[j] = v[pn];
            w[j] = w[pn];
        } else if (P->vel_type == 4) {
            /* parabolic-N (Solution.cpp:804-808): the profile is evaluated at
             * the centroid of cell b (the boundary index), w is not written */
            int ci = (OM_PLANT == 9) ? nb : b;
            double Y = c->p.domain_y;
            u[j] = 4 * P->u / pow(Y, 2) * c->cy[ci] * (Y - c->cy[ci]);
            v[j] = 4 * P->v / pow(Y, 2) * c->cy[ci] * (Y - c->cy[ci]);
        } else if (P->vel_type == 5) {
            /* parabolic-W (Solution.cpp:809-813): u takes x (X - y) of cell b */
            int ci = (OM_PLANT == 9) ? nb : b;
            double X = c->p.domain_x;
            u[j] = 4 * P->u / pow(X, 2) * c->cx[ci] * (X - ((OM_PLANT == 10) ? c->cx[ci] : c->cy[ci]));
            v[j] = 4 * P->v / pow(X, 2) * c->cx[ci] * (X - c->cx[ci]);
        } else if (P->vel_type == 6) {
            double taper = 1.0;
            u[j] = P->u * taper;
            v[j] = P->v * taper;
            w[j] = P->w * taper;
        } else if (P->vel_type == 8) {
            u[j] = u[nb];
            v[j] = -v[nb];
            w[j] = w[nb];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Gradients: gradients.cpp:319-506 (interior) and 592-648 (ghost, Dirichlet) */
/* ------------------------------------------------------------------------- */

#define GI(i, q) c->g[(size_t)(i) * 12 + (q)]

static void ghost_grads_dirichlet(om_ctx *c, const double *rho, const double *u, const double *v,
                                  const double *w, int b, int o, int rho_type) {
    const om_patch *P = &c->patch[c->m.bface_patch[b]];
    int i1 = b + c->N, f = b + c->NI;
    double dx = c->cx[i1] - c->cx[o], dy = c->cy[i1] - c->cy[o], dz = c->cz[i1] - c->cz[o];
    double d_mag = sqrt(pow(dx, 2.0) + pow(dy, 2.0) + pow(dz, 2.0));
    double d_u = 2 * (P->u - u[o]), d_v = 2 * (P->v - v[o]), d_w = 2 * (P->w - w[o]);
    double d_rho = 2 * (P->rho - rho[o]);
    double n_i = c->ni[f], n_j = c->nj[f], n_k = c->nk[f];
    if (rho_type) {
        GI(i1, 0) = d_rho / d_mag * n_i;
        GI(i1, 1) = d_rho / d_mag * n_j;
        GI(i1, 2) = d_rho / d_mag * n_k;
    } else {
        GI(i1, 3) = d_u / d_mag * n_i; GI(i1, 6) = d_v / d_mag * n_i; GI(i1, 9) = d_w / d_mag * n_i;
        GI(i1, 4) = d_u / d_mag * n_j; GI(i1, 7) = d_v / d_mag * n_j; GI(i1, 10) = d_w / d_mag * n_j;
        GI(i1, 5) = d_u / d_mag * n_k; GI(i1, 8) = d_v / d_mag * n_k; GI(i1, 11) = d_w / d_mag * n_k;
    }
}

static void compute_gradients(om_ctx *c, const double *rho, const double *u, const double *v, const double *w) {
    int N = c->N;
    for (int i = 0; i < N; i++) {
        double Rr[3] = {0, 0, 0}, Ru[3] = {0, 0, 0}, Rv[3] = {0, 0, 0}, Rw[3] = {0, 0, 0};
        int slot = 0;
        for (int s = c->st_off[i]; s < c->st_off[i + 1]; s++, slot++) {
            int i1 = c->st_cell[s];
            double du = (u[i1] - u[i]), dv = (v[i1] - v[i]), dw = (w[i1] - w[i]), dr = (rho[i1] - rho[i]);
            double wx = c->Rx[i * 6 + slot], wy = c->Ry[i * 6 + slot], wz = c->Rz[i * 6 + slot];
            Rr[0] = Rr[0] + wx * dr; Rr[1] = Rr[1] + wy * dr; Rr[2] = Rr[2] + wz * dr;
            Ru[0] = Ru[0] + wx * du; Ru[1] = Ru[1] + wy * du; Ru[2] = Ru[2] + wz * du;
            Rv[0] = Rv[0] + wx * dv; Rv[1] = Rv[1] + wy * dv; Rv[2] = Rv[2] + wz * dv;
            Rw[0] = Rw[0] + wx * dw; Rw[1] = Rw[1] + wy * dw; Rw[2] = Rw[2] + wz * dw;
        }
        double *R[4] = {Rr, Ru, Rv, Rw};
        for (int q = 0; q < 4; q++) {
            GI(i, 3 * q + 0) = c->Lxx[i] * R[q][0] + c->Lxy[i] * R[q][1] + c->Lxz[i] * R[q][2];
            GI(i, 3 * q + 1) = c->Lyx[i] * R[q][0] + c->Lyy[i] * R[q][1] + c->Lyz[i] * R[q][2];
            GI(i, 3 * q + 2) = c->Lzx[i] * R[q][0] + c->Lzy[i] * R[q][1] + c->Lzz[i] * R[q][2];
        }
    }
    for (int b = 0; b < c->G; b++) {
        int f = c->NI + b, o = c->m.owner[f], j = b + N;
        const om_patch *P = &c->patch[c->m.bface_patch[b]];
        if (P->rho_type == 1) {
            ghost_grads_dirichlet(c, rho, u, v, w, b, o, 1);
        } else if (P->rho_type == 2) {
            GI(j, 0) = GI(o, 0) * (1 - c->ni[f]);
            GI(j, 1) = GI(o, 1) * (1 - c->nj[f]);
            GI(j, 2) = GI(o, 2) * (1 - c->nk[f]);
        } else if (P->rho_type == 6) {
            GI(j, 0) = 0.0; GI(j, 1) = 0.0; GI(j, 2) = 0.0;
            if (OM_PLANT == 7) { GI(j, 0) = GI(o, 0); GI(j, 1) = GI(o, 1); GI(j, 2) = GI(o, 2); }
        } else if (P->rho_type == 7) {
            GI(j, 0) = GI(o, 0); GI(j, 1) = GI(o, 1); GI(j, 2) = GI(o, 2);
        } else if (P->rho_type == 8) {
            /* the velocity formula; the density slots are never written
             * (gradients.cpp:429-431): zero, as in a freshly allocated array */
            ghost_grads_dirichlet(c, rho, u, v, w, b, o, 0);
            if (OM_PLANT == 8) { GI(j, 0) = GI(o, 0); GI(j, 1) = GI(o, 1); GI(j, 2) = GI(o, 2); }
        } else {
            /* periodic: the partner's gradient (gradients.cpp:434-440) */
            const int pn = c->partner[b];
            GI(j, 0) = GI(pn, 0); GI(j, 1) = GI(pn, 1); GI(j, 2) = GI(pn, 2);
        }
        if (P->vel_type == 1) {
            ghost_grads_dirichlet(c, rho, u, v, w, b, o, 0);
        } else if (P->vel_type == 2) {
            for (int q = 1; q < 4; q++) {
                GI(j, 3 * q + 0) = GI(o, 3 * q + 0) * (1 - c->ni[f]);
                GI(j, 3 * q + 1) = GI(o, 3 * q + 1) * (1 - c->nj[f]);
                GI(j, 3 * q + 2) = GI(o, 3 * q + 2) * (1 - c->nk[f]);
            }
        } else if (P->vel_type == 6) {
            for (int q = 3; q < 12; q++) GI(j, q) = (OM_PLANT == 7) ? GI(o, q) : 0.0;
        } else if (P->vel_type == 8) {
            ghost_grads_dirichlet(c, rho, u, v, w, b, o, 0);
        } else {
            /* periodic (gradients.cpp:488-502) */
            const int pn = c->partner[b];
            for (int q = 3; q < 12; q++) GI(j, q) = GI(pn, q);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Face flux: Solver.cpp:286-381 (driver), 840-1274 (lattice macros),        */
/* 1278-1374 (feq), 655-837 (relaxed populations and flux)                   */
/* ------------------------------------------------------------------------- */

static void face_flux(om_ctx *c, int face, const double *rho, const double *u, const double *v, const double *w,
                      double out[4], double *rho_face_out) {
    int i = c->m.owner[face], nb = c->nbr[face], N = c->N;
    double dt = c->fdt[face];
    double n[3] = {c->ni[face], c->nj[face], c->nk[face]};
    double A = c->fa[face];
    double t1x = c->cx[i] - c->fx[face], t1y = c->cy[i] - c->fy[face], t1z = c->cz[i] - c->fz[face];
    double t2x = c->cx[nb] - c->fx[face], t2y = c->cy[nb] - c->fy[face], t2z = c->cz[nb] - c->fz[face];
    double lr[15], lu[15], lv[15], lw[15];
    double rho_i, rho_nb, u_i, u_nb, v_i, v_nb, w_i, w_nb;
    const double *gi = &c->g[(size_t)i * 12], *gn = &c->g[(size_t)nb * 12];
    if (nb > N) {
        const om_patch *P = &c->patch[c->m.bface_patch[nb - N]];
        if (P->rho_type == 1) lr[0] = P->rho;
        else if (P->rho_type == 8) lr[0] = rho[i];
        else lr[0] = rho[i] - (t1x * gn[0] + t1y * gn[1] + t1z * gn[2]);
        if (P->vel_type == 1) {
            lu[0] = P->u; lv[0] = P->v; lw[0] = P->w;
        } else if (P->vel_type == 8) {
            lu[0] = u[i]; lv[0] = 0; lw[0] = w[i];
        } else {
            lu[0] = u[i] - (t1x * gn[3] + t1y * gn[4] + t1z * gn[5]);
            lv[0] = v[i] - (t1x * gn[6] + t1y * gn[7] + t1z * gn[8]);
            lw[0] = w[i] - (t1x * gn[9] + t1y * gn[10] + t1z * gn[11]);
        }
        rho_i = rho_nb = lr[0];
        u_i = u_nb = lu[0];
        v_i = v_nb = lv[0];
        w_i = w_nb = lw[0];
    } else {
        rho_i = rho[i] - (t1x * gi[0] + t1y * gi[1] + t1z * gi[2]);
        rho_nb = rho[nb] - (t2x * gn[0] + t2y * gn[1] + t2z * gn[2]);
        lr[0] = (rho_i + rho_nb) * 0.5;
        if (OM_PLANT == 5) lr[0] = rho_i;
        u_i = u[i] - (t1x * gi[3] + t1y * gi[4] + t1z * gi[5]);
        u_nb = u[nb] - (t2x * gn[3] + t2y * gn[4] + t2z * gn[5]);
        lu[0] = (u_i + u_nb) * 0.5;
        if (OM_PLANT == 5) lu[0] = u_i;
        v_i = v[i] - (t1x * gi[6] + t1y * gi[7] + t1z * gi[8]);
        v_nb = v[nb] - (t2x * gn[6] + t2y * gn[7] + t2z * gn[8]);
        lv[0] = (v_i + v_nb) * 0.5;
        w_i = w[i] - (t1x * gi[9] + t1y * gi[10] + t1z * gi[11]);
        w_nb = w[nb] - (t2x * gn[9] + t2y * gn[10] + t2z * gn[11]);
        lw[0] = (w_i + w_nb) * 0.5;
    }
    /* lattice points k=1..14: the side is the neighbour iff e_k . n < 0, written
     * as the reference's sum of signed normal components > 0 (Solver.cpp:958-1273) */
    for (int k = 1; k < 15; k++) {
        double cond = (-EX[k]) * n[0] + (-EY[k]) * n[1] + (-EZ[k]) * n[2];
        int use_nb = (OM_PLANT == 4) ? cond < 0 : (OM_PLANT == 11) ? cond >= 0 : cond > 0;
        const double *gg = use_nb ? gn : gi;
        double br = use_nb ? rho_nb : rho_i, bu = use_nb ? u_nb : u_i;
        double bv = use_nb ? v_nb : v_i, bw = use_nb ? w_nb : w_i;
        /* rho_k = base - e_x g_x dt - e_y g_y dt - e_z g_z dt, left to right */
        double vals[4] = {br, bu, bv, bw};
        for (int q = 0; q < 4; q++) {
            double r = vals[q];
            if (OM_PLANT == 3) r = r + 2 * (EX[k] * gg[3 * q + 0] + EY[k] * gg[3 * q + 1] + EZ[k] * gg[3 * q + 2]) * dt;
            if (EX[k] > 0) r = r - gg[3 * q + 0] * dt; else if (EX[k] < 0) r = r + gg[3 * q + 0] * dt;
            if (EY[k] > 0) r = r - gg[3 * q + 1] * dt; else if (EY[k] < 0) r = r + gg[3 * q + 1] * dt;
            if (EZ[k] > 0) r = r - gg[3 * q + 2] * dt; else if (EZ[k] < 0) r = r + gg[3 * q + 2] * dt;
            vals[q] = r;
        }
        lr[k] = vals[0]; lu[k] = vals[1]; lv[k] = vals[2]; lw[k] = vals[3];
    }
    double gam = c->p.gamma;
    double feq[15];
    for (int k = 0; k < 15; k++) {
        double uu2 = lu[k] * lu[k] / gam, vv2 = lv[k] * lv[k] / gam, ww2 = lw[k] * lw[k] / gam;
        double u2v2w2 = (uu2 + vv2 + ww2) * 1.5;
        double uv = lu[k] * lv[k] * 9.0 / gam, uw = lu[k] * lw[k] * 9.0 / gam, vw = lv[k] * lw[k] * 9.0 / gam;
        double uu = lu[k], vv = lv[k], ww = lw[k];
        double wk = lattice_weight(k), rk = lr[k], t;
        switch (k) {
        case 0: t = (1.0 - u2v2w2); break;
        case 1: t = (1.0 + 3.0 * uu + 4.5 * uu2 - u2v2w2); break;
        case 2: t = (1.0 - 3.0 * uu + 4.5 * uu2 - u2v2w2); break;
        case 3: t = (1.0 + 3.0 * vv + 4.5 * vv2 - u2v2w2); break;
        case 4: t = (1.0 - 3.0 * vv + 4.5 * vv2 - u2v2w2); break;
        case 5: t = (1.0 + 3.0 * ww + 4.5 * ww2 - u2v2w2); break;
        case 6: t = (1.0 - 3.0 * ww + 4.5 * ww2 - u2v2w2); break;
        case 7: t = (1.0 + 3.0 * uu + 3.0 * vv + 3.0 * ww + uv + uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 8: t = (1.0 - 3.0 * uu - 3.0 * vv - 3.0 * ww + uv + uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 9: t = (1.0 + 3.0 * uu + 3.0 * vv - 3.0 * ww + uv - uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 10: t = (1.0 - 3.0 * uu - 3.0 * vv + 3.0 * ww + uv - uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 11: t = (1.0 + 3.0 * uu - 3.0 * vv + 3.0 * ww - uv + uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 12: t = (1.0 - 3.0 * uu + 3.0 * vv - 3.0 * ww - uv + uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 13: t = (1.0 - 3.0 * uu + 3.0 * vv + 3.0 * ww - uv - uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        default: t = (1.0 + 3.0 * uu - 3.0 * vv - 3.0 * ww - uv - uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        }
        feq[k] = wk * rk * t;
    }
    /* interface moments (Solver.cpp:329-347) */
    double rI = feq[0] + feq[1] + feq[2] + feq[3] + feq[4] + feq[5] + feq[6] + feq[7] + feq[8] + feq[9] +
                feq[10] + feq[11] + feq[12] + feq[13] + feq[14];
    double uIx = 1 / rI * (feq[1] - feq[2] + feq[7] - feq[8] + feq[9] - feq[10] + feq[11] - feq[12] - feq[13] + feq[14]);
    double uIy = 1 / rI * (feq[3] - feq[4] + feq[7] - feq[8] + feq[9] - feq[10] - feq[11] + feq[12] + feq[13] - feq[14]);
    double uIz = 1 / rI * (feq[5] - feq[6] + feq[7] - feq[8] - feq[9] + feq[10] + feq[11] - feq[12] + feq[13] - feq[14]);
    if (rho_face_out) *rho_face_out = rI;
    /* calculate_flux_at_interface (Solver.cpp:655-837) */
    double uu2 = uIx * uIx / gam, vv2 = uIy * uIy / gam, ww2 = uIz * uIz / gam;
    double u2v2w2 = (uu2 + vv2 + ww2) * 1.5;
    double uu = uIx, vv = uIy, ww = uIz;
    double uv = uu * vv * 9.0 / gam, uw = uu * ww * 9.0 / gam, vw = vv * ww * 9.0 / gam;
    double tau = (c->p.visc * 3 / dt / gam);
    if (OM_PLANT == 1) tau = c->p.visc * 3 / dt;
    if (OM_PLANT == 2) tau = 2 * tau;
    double fi[15];
    for (int k = 1; k < 15; k++) {
        double t;
        switch (k) {
        case 1: t = (1.0 + 3.0 * uu + 4.5 * uu2 - u2v2w2); break;
        case 2: t = (1.0 - 3.0 * uu + 4.5 * uu2 - u2v2w2); break;
        case 3: t = (1.0 + 3.0 * vv + 4.5 * vv2 - u2v2w2); break;
        case 4: t = (1.0 - 3.0 * vv + 4.5 * vv2 - u2v2w2); break;
        case 5: t = (1.0 + 3.0 * ww + 4.5 * ww2 - u2v2w2); break;
        case 6: t = (1.0 - 3.0 * ww + 4.5 * ww2 - u2v2w2); break;
        case 7: t = (1.0 + 3.0 * uu + 3.0 * vv + 3.0 * ww + uv + uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 8: t = (1.0 - 3.0 * uu - 3.0 * vv - 3.0 * ww + uv + uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 9: t = (1.0 + 3.0 * uu + 3.0 * vv - 3.0 * ww + uv - uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 10: t = (1.0 - 3.0 * uu - 3.0 * vv + 3.0 * ww + uv - uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 11: t = (1.0 + 3.0 * uu - 3.0 * vv + 3.0 * ww - uv + uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 12: t = (1.0 - 3.0 * uu + 3.0 * vv - 3.0 * ww - uv + uw - vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        case 13: t = (1.0 - 3.0 * uu + 3.0 * vv + 3.0 * ww - uv - uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        default: t = (1.0 + 3.0 * uu - 3.0 * vv - 3.0 * ww - uv - uw + vw + 4.5 * uu2 + 4.5 * vv2 + 4.5 * ww2 - u2v2w2); break;
        }
        double fe = lattice_weight(k) * rI * t;
        fi[k] = fe - tau * (fe - feq[k]);
    }
    double xP = (fi[1] - fi[2] + fi[7] - fi[8] + fi[9] - fi[10] + fi[11] - fi[12] - fi[13] + fi[14]);
    double xmx = (fi[1] + fi[2] + fi[7] + fi[8] + fi[9] + fi[10] + fi[11] + fi[12] + fi[13] + fi[14]);
    double xmy = fi[7] + fi[8] + fi[9] + fi[10] - fi[11] - fi[12] - fi[13] - fi[14];
    double xmz = fi[7] + fi[8] - fi[9] - fi[10] + fi[11] + fi[12] - fi[13] - fi[14];
    double yP = (fi[3] - fi[4] + fi[7] - fi[8] + fi[9] - fi[10] - fi[11] + fi[12] + fi[13] - fi[14]);
    double ymx = xmy;
    double ymy = (fi[3] + fi[4] + fi[7] + fi[8] + fi[9] + fi[10] + fi[11] + fi[12] + fi[13] + fi[14]);
    double ymz = fi[7] + fi[8] - fi[9] - fi[10] - fi[11] - fi[12] + fi[13] + fi[14];
    double zP = (fi[5] - fi[6] + fi[7] - fi[8] - fi[9] + fi[10] + fi[11] - fi[12] + fi[13] - fi[14]);
    double zmx = xmz, zmy = ymz;
    double zmz = (fi[5] + fi[6] + fi[7] + fi[8] + fi[9] + fi[10] + fi[11] + fi[12] + fi[13] + fi[14]);
    out[0] = (xP * n[0] + yP * n[1] + zP * n[2]) * A;
    out[1] = (xmx * n[0] + ymx * n[1] + zmx * n[2]) * A;
    out[2] = (xmy * n[0] + ymy * n[1] + zmy * n[2]) * A;
    out[3] = (xmz * n[0] + ymz * n[1] + zmz * n[2]) * A;
}

/* ------------------------------------------------------------------------- */
/* CFL: Solver.cpp:1693-1776                                                  */
/* ------------------------------------------------------------------------- */

static void get_cfl(om_ctx *c, const double *rho, const double *u, const double *v, const double *w) {
    double factor = c->p.cfl, U = c->p.max_velocity, gam = c->p.gamma;
    double ceff = U * sqrt(1 - gam + pow(gam / sqrt(3) / U, 2));
    double min_dt = 100000000000;
    for (int i = 0; i < c->N; i++) {
        double ax = c->cfl_area[i], ay = 0.0, az = 0.0;
        double eig = (fabs(u[i]) + ceff) * ax + (fabs(v[i]) + ceff) * ay + (fabs(w[i]) + ceff) * az;
        eig = eig / gam;
        double ve = 2 * c->p.visc / gam / rho[i] / c->vol[i];
        ve = ve * (ax * ax + ay * ay + az * az);
        eig = eig + 4 * ve;
        double temp = factor * c->vol[i] / eig;
        if (temp < 0) min_dt = temp;
        if (temp < min_dt) min_dt = temp;
        if (c->p.time_stepping == 0 || c->p.time_stepping == 3) c->dtl[i] = temp;
        else c->dtl[i] = factor;
    }
    if (c->p.time_stepping == 2) {
        for (int i = 0; i < c->N; i++) c->dtl[i] = min_dt;
    }
    if (c->p.time_stepping == 3) {
        for (int i = 0; i < c->N; i++) {
            double fr = pow(2, floor(log2(c->dtl[i] / min_dt)));
            c->dtl[i] = min_dt * fr;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* One stage's residual: Solver.cpp:272-381                                  */
/* ------------------------------------------------------------------------- */

static void stage_residual(om_ctx *c) {
    size_t T = (size_t)c->T;
    update_bcs(c, c->t_rho, c->t_u, c->t_v, c->t_w);
    memset(c->r_rho, 0, sizeof(double) * T); memset(c->r_u, 0, sizeof(double) * T);
    memset(c->r_v, 0, sizeof(double) * T); memset(c->r_w, 0, sizeof(double) * T);
    compute_gradients(c, c->t_rho, c->t_u, c->t_v, c->t_w);
    for (int f = 0; f < c->F; f++) {
        if (f > c->NI && c->patch[c->m.bface_patch[f - c->NI]].is_empty) continue;
        double Fl[4];
        face_flux(c, f, c->t_rho, c->t_u, c->t_v, c->t_w, Fl, NULL);
        int i = c->m.owner[f], nb = c->nbr[f];
        c->r_rho[i] = c->r_rho[i] + (-Fl[0] / c->vol[i]);
        c->r_u[i] = c->r_u[i] + (-Fl[1] / c->vol[i]);
        c->r_v[i] = c->r_v[i] + (-Fl[2] / c->vol[i]);
        c->r_w[i] = c->r_w[i] + (-Fl[3] / c->vol[i]);
        if (nb < c->N) { /* ghost slots receive F/0 in the reference and are never read */
            c->r_rho[nb] = c->r_rho[nb] + (+Fl[0] / c->vol[nb]);
            c->r_u[nb] = c->r_u[nb] + (+Fl[1] / c->vol[nb]);
            c->r_v[nb] = c->r_v[nb] + (Fl[2] / c->vol[nb]);
            c->r_w[nb] = c->r_w[nb] + (Fl[3] / c->vol[nb]);
        }
    }
    if (c->use_body_force) {
        for (int i = 0; i < c->N; i++) {
            c->r_u[i] = c->r_u[i] + c->force_x[i];
            c->r_v[i] = c->r_v[i] + c->force_y[i];
            c->r_w[i] = c->r_w[i] + c->force_z[i];
        }
    }
}

/* Exposed for unit tests: one stage's residual from a given temp state over T cells
 * (ghost values are recomputed from the BCs first).  Outputs R [N], gradients [T*12]. */
void om_eval_residual(om_ctx *c, const double *rho, const double *u, const double *v, const double *w,
                      double *R4, double *grads) {
    size_t T = (size_t)c->T;
    memcpy(c->t_rho, rho, sizeof(double) * T); memcpy(c->t_u, u, sizeof(double) * T);
    memcpy(c->t_v, v, sizeof(double) * T); memcpy(c->t_w, w, sizeof(double) * T);
    stage_residual(c);
    for (int i = 0; i < c->N; i++) {
        R4[(size_t)i * 4 + 0] = c->r_rho[i]; R4[(size_t)i * 4 + 1] = c->r_u[i];
        R4[(size_t)i * 4 + 2] = c->r_v[i]; R4[(size_t)i * 4 + 3] = c->r_w[i];
    }
    if (grads) memcpy(grads, c->g, sizeof(double) * T * 12);
}

/* Exposed for unit tests: flux of one face from the current temp state and gradients
 * (call om_eval_residual first). */
void om_face_flux(om_ctx *c, int face, double *out4) {
    face_flux(c, face, c->t_rho, c->t_u, c->t_v, c->t_w, out4, NULL);
}

/* Exposed for unit tests: the temp state over T cells after om_eval_residual
 * (interior cells as given, ghosts from the boundary update). */
void om_get_temp(const om_ctx *c, double *rho, double *u, double *v, double *w) {
    size_t T = (size_t)c->T;
    memcpy(rho, c->t_rho, sizeof(double) * T); memcpy(u, c->t_u, sizeof(double) * T);
    memcpy(v, c->t_v, sizeof(double) * T); memcpy(w, c->t_w, sizeof(double) * T);
}

void om_get_dt(om_ctx *c, double *dt) {
    get_cfl(c, c->s_rho, c->s_u, c->s_v, c->s_w);
    memcpy(dt, c->dtl, sizeof(double) * c->N);
}

/* ------------------------------------------------------------------------- */
/* Immersed boundary: the reference GPU sequence (Solver_gpu.cu:687-718):     */
/* interpolate node velocities from the step-start state                    */
/* (immersed_boundary_method.cu:13-90), move nodes X = X0 + U dt             */
/* (:244-295), spring force F = -k dA (X - X_ref) (:161-190), spread to the   */
/* cells (:299-375); the cell force is added to R every stage (:766-768).    */
/* On unstructured meshes the delta weights are normalised per node:        */
/*   W_cn = delta(x_c - X_n) V_c / sum_c' delta(x_c' - X_n) V_c'             */
/*   U_n = sum_c W_cn u_c,   f_c = sum_n F_n W_cn / V_c                      */
/* which is the reference's formula on its uniform unit grid (sum w = 1).   */
/* Kernels: hat  phi(r) = max(0, 1-|r|)   (the reference's 2-point stencil);  */
/*          Peskin (2002) 4-point phi on |r| <= 2.                            */
/* ------------------------------------------------------------------------- */

static double delta_phi(int kernel, double r) {
    r = fabs(r);
    if (kernel == 0) return r < 1.0 ? 1.0 - r : 0.0;
    if (r <= 1.0) return (3.0 - 2.0 * r + sqrt(1.0 + 4.0 * r - 4.0 * r * r)) / 8.0;
    if (r <= 2.0) return (5.0 - 2.0 * r - sqrt(-7.0 + 12.0 * r - 4.0 * r * r)) / 8.0;
    return 0.0;
}

static double delta3(const om_ctx *c, int cell, const double *X) {
    double ih = 1.0 / c->ibm_h;
    return delta_phi(c->ibm_kernel, (c->cx[cell] - X[0]) * ih) * delta_phi(c->ibm_kernel, (c->cy[cell] - X[1]) * ih) *
           delta_phi(c->ibm_kernel, (c->cz[cell] - X[2]) * ih) * (ih * ih * ih);
}

void om_set_ibm(om_ctx *c, int n, const double *xref, const double *x, double stiffness, double dA, double dt,
                double h, int kernel) {
    free(c->ibm_x); free(c->ibm_x0); free(c->ibm_xref); free(c->ibm_u); free(c->ibm_f);
    c->ibm_n = n;
    c->ibm_on = n > 0;
    c->ibm_kernel = kernel;
    c->ibm_h = h; c->ibm_dt = dt; c->ibm_k = stiffness; c->ibm_dA = dA;
    c->ibm_x = xcalloc((size_t)3 * n, sizeof(double));
    c->ibm_x0 = xcalloc((size_t)3 * n, sizeof(double));
    c->ibm_xref = xcalloc((size_t)3 * n, sizeof(double));
    c->ibm_u = xcalloc((size_t)3 * n, sizeof(double));
    c->ibm_f = xcalloc((size_t)3 * n, sizeof(double));
    memcpy(c->ibm_x, x, sizeof(double) * 3 * n);
    memcpy(c->ibm_x0, x, sizeof(double) * 3 * n);
    memcpy(c->ibm_xref, xref, sizeof(double) * 3 * n);
    c->use_body_force = c->ibm_on;
    memset(c->force_x, 0, sizeof(double) * c->N);
    memset(c->force_y, 0, sizeof(double) * c->N);
    memset(c->force_z, 0, sizeof(double) * c->N);
}

/* Interpolated velocity at X from the primitive state (brute force over cells). */
static void ibm_interp(const om_ctx *c, const double *X, const double *u, const double *v, const double *w,
                       double U[3]) {
    double S = 0, su = 0, sv = 0, sw = 0;
    for (int i = 0; i < c->N; i++) {
        double d = delta3(c, i, X);
        if (d == 0.0) continue;
        double wv = d * c->vol[i];
        S += wv; su += wv * u[i]; sv += wv * v[i]; sw += wv * w[i];
    }
    U[0] = su / S; U[1] = sv / S; U[2] = sw / S;
}

static double ibm_norm(const om_ctx *c, const double *X) {
    double S = 0;
    for (int i = 0; i < c->N; i++) {
        double d = delta3(c, i, X);
        if (d != 0.0) S += d * c->vol[i];
    }
    return S;
}

/* One IBM update from the given primitive state; fills the cell force arrays. */
void om_ibm_step(om_ctx *c, const double *u, const double *v, const double *w) {
    int n = c->ibm_n;
    memset(c->force_x, 0, sizeof(double) * c->N);
    memset(c->force_y, 0, sizeof(double) * c->N);
    memset(c->force_z, 0, sizeof(double) * c->N);
    for (int k = 0; k < n; k++) {
        double *X = c->ibm_x + 3 * k, *X0 = c->ibm_x0 + 3 * k, *Xr = c->ibm_xref + 3 * k;
        double *U = c->ibm_u + 3 * k, *Fn = c->ibm_f + 3 * k;
        ibm_interp(c, X, u, v, w, U);
        for (int a = 0; a < 3; a++) {
            X[a] = X0[a] + U[a] * c->ibm_dt;
            X0[a] = X[a];
            Fn[a] = (-c->ibm_k * c->ibm_dA * (X[a] - Xr[a]));
        }
    }
    for (int k = 0; k < n; k++) {
        double *X = c->ibm_x + 3 * k, *Fn = c->ibm_f + 3 * k;
        double S = ibm_norm(c, X);
        for (int i = 0; i < c->N; i++) {
            double d = delta3(c, i, X);
            if (d == 0.0) continue;
            double q = d / S;
            c->force_x[i] += Fn[0] * q;
            c->force_y[i] += Fn[1] * q;
            c->force_z[i] += Fn[2] * q;
        }
    }
}

void om_get_ibm(const om_ctx *c, double *x, double *u, double *f) {
    size_t n = (size_t)3 * c->ibm_n;
    if (x) memcpy(x, c->ibm_x, sizeof(double) * n);
    if (u) memcpy(u, c->ibm_u, sizeof(double) * n);
    if (f) memcpy(f, c->ibm_f, sizeof(double) * n);
}

void om_get_body_force(const om_ctx *c, double *f3) {
    for (int i = 0; i < c->N; i++) {
        f3[(size_t)3 * i] = c->force_x[i];
        f3[(size_t)3 * i + 1] = c->force_y[i];
        f3[(size_t)3 * i + 2] = c->force_z[i];
    }
}

/* ------------------------------------------------------------------------- */
/* Time loop: Solver.cpp:245-560                                             */
/* ------------------------------------------------------------------------- */

static const double RK_ALPHA[4] = {1.0, 0.5, 0.5, 1.0};
static const double RK_BETA[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

/* Runs up to nsteps time steps; writes per-step residuals (rho,u,v,w) into hist
 * (may be NULL).  Returns OM_OK, OM_ERR_NAN or OM_ERR_RHO; *converged set when
 * max(u,v,w) residual < tolerance (the solve stops after that step). */
int om_run(om_ctx *c, int nsteps, double *hist, int *steps_done, int *converged) {
    int N = c->N;
    size_t T = (size_t)c->T;
    double force = c->p.force;
    *converged = 0;
    *steps_done = 0;
    for (int t = 0; t < nsteps; t++) {
        memcpy(c->a_rho, c->b_rho, sizeof(double) * T); memcpy(c->a_u, c->b_u, sizeof(double) * T);
        memcpy(c->a_v, c->b_v, sizeof(double) * T); memcpy(c->a_w, c->b_w, sizeof(double) * T);
        memcpy(c->t_rho, c->s_rho, sizeof(double) * T); memcpy(c->t_u, c->s_u, sizeof(double) * T);
        memcpy(c->t_v, c->s_v, sizeof(double) * T); memcpy(c->t_w, c->s_w, sizeof(double) * T);
        get_cfl(c, c->t_rho, c->t_u, c->t_v, c->t_w);
        if (c->ibm_on) om_ibm_step(c, c->t_u, c->t_v, c->t_w);
        for (int rk = 0; rk < 4; rk++) {
            stage_residual(c);
            for (int i = 0; i < N; i++) {
                double dti = c->dtl[i];
                double f1 = c->a_rho[i] + c->r_rho[i] * dti * RK_ALPHA[rk];
                double f2 = c->a_u[i] + (c->r_u[i] + force) * dti * RK_ALPHA[rk];
                double f3 = c->a_v[i] + c->r_v[i] * dti * RK_ALPHA[rk];
                double f4 = c->a_w[i] + c->r_w[i] * dti * RK_ALPHA[rk];
                f2 = f2 / f1; f3 = f3 / f1; f4 = f4 / f1;
                c->t_rho[i] = f1; c->t_u[i] = f2; c->t_v[i] = f3; c->t_w[i] = f4;
                c->b_rho[i] = c->b_rho[i] + dti * RK_BETA[rk] * c->r_rho[i];
                c->b_u[i] = c->b_u[i] + dti * RK_BETA[rk] * (c->r_u[i] + force);
                c->b_v[i] = c->b_v[i] + dti * RK_BETA[rk] * c->r_v[i];
                c->b_w[i] = c->b_w[i] + dti * RK_BETA[rk] * c->r_w[i];
                c->s_rho[i] = c->b_rho[i];
                c->s_u[i] = c->b_u[i] / c->b_rho[i];
                c->s_v[i] = c->b_v[i] / c->b_rho[i];
                c->s_w[i] = c->b_w[i] / c->b_rho[i];
            }
        }
        /* residual + guards (Solver.cpp:444-498; residuals.cpp:35-44,79-91,139-147) */
        double nr = 0, nu = 0, nv = 0, nw = 0, den = 0;
        int err = OM_OK;
        for (int i = 0; i < N; i++) {
            nr = nr + pow(c->r_rho[i], 2.0); nu = nu + pow(c->r_u[i], 2.0);
            nv = nv + pow(c->r_v[i], 2.0); nw = nw + pow(c->r_w[i], 2.0);
            den = den + 1;
            if (isnan(c->t_rho[i]) || isnan(c->t_u[i])) { err = OM_ERR_NAN; break; }
            if (c->t_rho[i] / c->p.ref_rho > 1000.0) { err = OM_ERR_RHO; break; }
        }
        *steps_done = t + 1;
        c->steps_done++;
        if (err != OM_OK) return err;
        double ref_mom = c->p.ref_rho * c->p.max_velocity;
        c->res[0] = sqrt(nr / den) / c->p.ref_rho;
        c->res[1] = sqrt(nu / den) / ref_mom;
        c->res[2] = sqrt(nv / den) / ref_mom;
        c->res[3] = sqrt(nw / den) / ref_mom;
        if (hist) memcpy(hist + (size_t)t * 4, c->res, sizeof(double) * 4);
        double me = fmax(fmax(c->res[3], c->res[1]), c->res[2]);
        if (me < c->p.tolerance) {
            *converged = 1;
            return OM_OK;
        }
    }
    return OM_OK;
}
