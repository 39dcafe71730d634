// Post-processing of a solver context: surface forces on wall patches
// (src/post_processing.cpp:50-173) and the force on the immersed objects.
// Host code over the context's state and gradients.
#include <cmath>
#include <vector>

#include "ctx.hpp"

extern "C" int plbfs_wall_forces(plbfs_ctx *c, const plbfs_mesh_view *mesh, const int32_t *bface_patch,
                                 int32_t patch, double *out2) {
    if (!c || !mesh || !out2 || (mesh->n_faces > mesh->n_internal_faces && !bface_patch))
        return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    const int N = mesh->n_cells, F = mesh->n_faces, NI = mesh->n_internal_faces, G = F - NI;
    if (N != c->L.N || F != c->L.F) return plbfs_fail(PLBFS_ERR_INPUT, "mesh does not match the context");
    if (patch < 0 || patch >= (int)c->patches.size()) return plbfs_fail(PLBFS_ERR_INPUT, "patch %d", patch);
    std::vector<double> rho(N), u(N), v(N), w(N), g((size_t)12 * N);
    int rc;
    if ((rc = plbfs_get_state(c, rho.data(), u.data(), v.data(), w.data()))) return rc;
    if ((rc = plbfs_eval_residual(c, g.data(), nullptr))) return rc;
    const plbfs_patch &P = c->patches[patch];
    const double gam = c->params.gamma, nuv = c->params.visc * gam;
    const double *cc = mesh->cell_centroid;
    double drag = 0, lift = 0;
    for (int b = 0; b < G; b++) {
        if (bface_patch[b] != patch) continue;
        const int f = NI + b, o = mesh->owner[f], gh = N + b;
        const double n[3] = {mesh->face_normal[3 * f], mesh->face_normal[3 * f + 1], mesh->face_normal[3 * f + 2]};
        const double A = mesh->face_area[f];
        // ghost-cell gradients of u and v (src/gradients.cpp:446-502, 592-648)
        double gu[3], gv[3];
        if (P.vel_type == 1 || P.vel_type == 8) {
            const double dx = cc[3 * gh] - cc[3 * o], dy = cc[3 * gh + 1] - cc[3 * o + 1],
                         dz = cc[3 * gh + 2] - cc[3 * o + 2];
            const double dm = std::sqrt(std::pow(dx, 2.0) + std::pow(dy, 2.0) + std::pow(dz, 2.0));
            const double du = 2 * (P.u - u[o]), dv = 2 * (P.v - v[o]);
            for (int a = 0; a < 3; a++) {
                gu[a] = du / dm * n[a];
                gv[a] = dv / dm * n[a];
            }
        } else if (P.vel_type == 2) {
            for (int a = 0; a < 3; a++) {
                gu[a] = g[(size_t)12 * o + 3 + a] * (1 - n[a]);
                gv[a] = g[(size_t)12 * o + 6 + a] * (1 - n[a]);
            }
        } else if (P.vel_type == 3 || plbfs::vel_keeps_ghost(P.vel_type)) {
            // the periodic partner's gradient; cell 0's where the ghost is not written
            const int p = c->L.perm[c->L.b_partner[b]];
            for (int a = 0; a < 3; a++) {
                gu[a] = g[(size_t)12 * p + 3 + a];
                gv[a] = g[(size_t)12 * p + 6 + a];
            }
        } else {
            for (int a = 0; a < 3; a++) gu[a] = gv[a] = 0.0;
        }
        const double pr = rho[o] / 3.0 * gam;
        drag = drag + ((pr - nuv * gu[0]) * n[0] - nuv * gu[1] * n[1]) * A;
        lift = lift + (nuv * gv[0] * n[0] - (pr - nuv * gv[1]) * n[1]) * A;
    }
    out2[0] = drag;
    out2[1] = lift;
    return PLBFS_OK;
}

extern "C" int plbfs_ibm_body_force(plbfs_ctx *c, double *out3) {
    if (!c || !out3) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    if (!c->ibm.on) return plbfs_fail(PLBFS_ERR_INPUT, "no immersed boundary set");
    const int n = c->ibm.n;
    std::vector<double> f((size_t)3 * n);
    int rc = plbfs_get_ibm(c, nullptr, nullptr, f.data());
    if (rc) return rc;
    // the spring force acts on the fluid; the fluid pushes the body the other way
    double s[3] = {0, 0, 0};
    for (int k = 0; k < n; k++)
        for (int a = 0; a < 3; a++) s[a] -= f[(size_t)3 * k + a];
    for (int a = 0; a < 3; a++) out3[a] = s[a];
    return PLBFS_OK;
}
