// Host precompute of the solver's device layout (FP64, built with
// -ffp-contract=off so the LSQ arithmetic follows the reference's order).
//
// LSQ matrices: src/gradients.cpp:214-313 (pre_fill_LHS_and_RHS_matrix) and
// :149-210 (add_non_macrovariable_contributions), including the reference's
// determinant expression at :282-284, which is NOT the determinant of the
// assembled matrix when LHS_xy != 0.  The gradient the reference evaluates,
// inv * sum_j RHS_j (phi_j - phi_i) (:341-389), is folded here into one
// 3-vector weight per stencil slot, W_j = inv * RHS_j.
// CFL areas: src/Solver.cpp:1653-1686 (all three sums land in the x slot).
#include "layout.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "host_common.hpp"

namespace plbfs {

// Hilbert index via the transpose form of Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707 (2004): axes -> transpose, then bit interleave.
uint64_t hilbert_key3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; i++) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; i++) X[i] ^= t;
    uint64_t key = 0;
    for (int q = bits - 1; q >= 0; q--)
        for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> q) & 1u);
    return key;
}

// the type codes of src/preprocessor.cpp:239-258
static bool type_ok(int t) { return t >= 1 && t <= 9; }

// Interior cell nearest to x (file index; ties to the smaller index), found
// through a uniform bin grid of the interior centroids.
struct NearestCell {
    double lo[3], hb;
    int nb[3];
    std::vector<int32_t> off, idx;
    const double *cc;
    int N;
    NearestCell(const double *cc_, int N_) : cc(cc_), N(N_) {
        double hi[3];
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = cc[a];
        for (int i = 0; i < N; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], cc[3 * i + a]);
                hi[a] = std::max(hi[a], cc[3 * i + a]);
            }
        double vol = 1;
        for (int a = 0; a < 3; a++) vol *= std::max(hi[a] - lo[a], 1e-12);
        hb = std::max(std::cbrt(vol / N) * 2.0, 1e-12);
        for (int a = 0; a < 3; a++) nb[a] = std::max(1, (int)std::ceil((hi[a] - lo[a]) / hb) + 1);
        const size_t nbin = (size_t)nb[0] * nb[1] * nb[2];
        off.assign(nbin + 1, 0);
        std::vector<size_t> bin(N);
        for (int i = 0; i < N; i++) {
            bin[i] = key(cc + 3 * i);
            off[bin[i] + 1]++;
        }
        for (size_t k = 0; k < nbin; k++) off[k + 1] += off[k];
        idx.resize(N);
        std::vector<int32_t> fill(off.begin(), off.end() - 1);
        for (int i = 0; i < N; i++) idx[fill[bin[i]]++] = i;
    }
    int coord(double x, int a) const { return std::min(nb[a] - 1, std::max(0, (int)std::floor((x - lo[a]) / hb))); }
    size_t key(const double *x) const {
        return ((size_t)coord(x[0], 0) * nb[1] + coord(x[1], 1)) * nb[2] + coord(x[2], 2);
    }
    int find(const double x[3]) const {
        const int c[3] = {coord(x[0], 0), coord(x[1], 1), coord(x[2], 2)};
        double best = 1e300;
        int arg = -1;
        for (int r = 0;; r++) {
            // shell of bins at Chebyshev distance r
            for (int i = c[0] - r; i <= c[0] + r; i++)
                for (int j = c[1] - r; j <= c[1] + r; j++)
                    for (int k = c[2] - r; k <= c[2] + r; k++) {
                        if (std::max(std::abs(i - c[0]), std::max(std::abs(j - c[1]), std::abs(k - c[2]))) != r)
                            continue;
                        if (i < 0 || j < 0 || k < 0 || i >= nb[0] || j >= nb[1] || k >= nb[2]) continue;
                        const size_t b = ((size_t)i * nb[1] + j) * nb[2] + k;
                        for (int q = off[b]; q < off[b + 1]; q++) {
                            const int n = idx[q];
                            const double dx = cc[3 * n] - x[0], dy = cc[3 * n + 1] - x[1], dz = cc[3 * n + 2] - x[2];
                            const double d = std::sqrt(std::pow(dx, 2) + std::pow(dy, 2) + std::pow(dz, 2));
                            if (d < best || (d == best && n < arg)) {
                                best = d;
                                arg = n;
                            }
                        }
                    }
            // every unvisited bin is at least r * hb away from x's bin
            if (arg >= 0 && best < r * hb) return arg;
            if (r > nb[0] + nb[1] + nb[2]) return arg;
        }
    }
};

int build_layout(const plbfs_mesh_view &m, const plbfs_params &p, const plbfs_patch *patches, int n_patches,
                 const int32_t *bface_patch, int block_cells, bool split_border, Layout &L) {
    const int N = m.n_cells, F = m.n_faces, NI = m.n_internal_faces, G = F - NI;
    if (N <= 0 || F <= 0 || NI < 0 || NI > F || !m.owner || !m.neighbour || !m.cell_centroid || !m.cell_volume ||
        !m.face_centroid || !m.face_normal || !m.face_area || !m.face_dt)
        return plbfs_fail(PLBFS_ERR_INPUT, "incomplete mesh view");
    if (n_patches <= 0 || !patches || (G > 0 && !bface_patch)) return plbfs_fail(PLBFS_ERR_INPUT, "missing patches");
    for (int q = 0; q < n_patches; q++) {
        if (!type_ok(patches[q].rho_type) || !type_ok(patches[q].vel_type))
            return plbfs_fail(PLBFS_ERR_BCTYPE, "patch %d: unsupported bc types rho=%d vel=%d", q,
                              patches[q].rho_type, patches[q].vel_type);
        if ((patches[q].rho_type == 3) != (patches[q].vel_type == 3))
            return plbfs_fail(PLBFS_ERR_BCTYPE, "patch %d: periodic must apply to both rho and velocity", q);
    }
    for (int b = 0; b < G; b++)
        if (bface_patch[b] < 0 || bface_patch[b] >= n_patches)
            return plbfs_fail(PLBFS_ERR_INPUT, "boundary face %d: patch %d out of range", b, bface_patch[b]);
    for (int f = 0; f < F; f++) {
        if (m.owner[f] < 0 || m.owner[f] >= N) return plbfs_fail(PLBFS_ERR_INPUT, "face %d: owner", f);
        if (f < NI && (m.neighbour[f] < 0 || m.neighbour[f] >= N))
            return plbfs_fail(PLBFS_ERR_INPUT, "face %d: neighbour", f);
        if (f >= NI && m.neighbour[f] != N + f - NI)
            return plbfs_fail(PLBFS_ERR_INPUT, "face %d: boundary neighbour must be the ghost index", f);
    }
    if (block_cells < 64 || block_cells > 1024 || block_cells % 64)
        return plbfs_fail(PLBFS_ERR_INPUT, "block_cells %d", block_cells);
    L = Layout();
    L.N = N; L.F = F; L.NI = NI; L.G = G;
    L.block_cells = block_cells;
    const double *cc = m.cell_centroid;
    const double *fc = m.face_centroid;
    const double *fn = m.face_normal;

    // stencil in file order (src/unstructured_mesh.cpp:1197-1222)
    std::vector<int32_t> s_off(N + 1, 0), s_cell, s_face;
    for (int f = 0; f < F; f++) {
        s_off[m.owner[f] + 1]++;
        if (f < NI) s_off[m.neighbour[f] + 1]++;
    }
    for (int i = 0; i < N; i++) {
        if (s_off[i + 1] > kMaxSlots)
            return plbfs_fail(PLBFS_ERR_INPUT, "cell %d has %d faces; at most %d are supported", i, s_off[i + 1],
                              kMaxSlots);
        s_off[i + 1] += s_off[i];
    }
    s_cell.resize(s_off[N]);
    s_face.resize(s_off[N]);
    {
        std::vector<int32_t> fill(s_off.begin(), s_off.end() - 1);
        for (int f = 0; f < F; f++) {
            const int o = m.owner[f];
            s_cell[fill[o]] = m.neighbour[f];
            s_face[fill[o]++] = f;
            if (f < NI) {
                const int nb = m.neighbour[f];
                s_cell[fill[nb]] = o;
                s_face[fill[nb]++] = f;
            }
        }
    }

    // LSQ precompute (gradients.cpp:214-313)
    L.lsq_inv9.assign((size_t)N * 9, 0.0);
    L.lsq_rhs.assign((size_t)N * 18, 0.0);
    L.cfl_area.assign(N, 0.0);
    std::vector<double> cfl_file(N, 0.0);
    for (int i = 0; i < N; i++) {
        double xx = 0, xy = 0, xz = 0, yx = 0, yy = 0, yz = 0, zx = 0, zy = 0, zz = 0;
        double max_area = 0.0, max_disp = 0.0;
        const double vi = m.cell_volume[i];
        for (int s = s_off[i]; s < s_off[i + 1]; s++) {
            const int i1 = s_cell[s], f = s_face[s];
            max_area = std::max(max_area, m.face_area[f]);
            const double dx = cc[3 * i1] - cc[3 * i], dy = cc[3 * i1 + 1] - cc[3 * i + 1],
                         dz = cc[3 * i1 + 2] - cc[3 * i + 2];
            const double disp = std::sqrt(dx * dx + dy * dy + dz * dz);
            max_disp = std::max(disp, max_disp);
        }
        const double ar = 2 * std::fabs(max_disp) * max_area / vi;
        double beta;
        if (p.gradient_type == PLBFS_GRAD_HYBRID) beta = std::min(1.0, (2 / ar));
        else if (p.gradient_type == PLBFS_GRAD_LS) beta = 1.0;
        else beta = 0.0;
        xx = xx + 2 * (1 - beta) * vi;
        yy = yy + 2 * (1 - beta) * vi;
        zz = zz + 2 * (1 - beta) * vi;
        int slot = 0;
        for (int s = s_off[i]; s < s_off[i + 1]; s++, slot++) {
            const int i1 = s_cell[s], f = s_face[s];
            const double ni = fn[3 * f], nj = fn[3 * f + 1], nk = fn[3 * f + 2];
            double dx = fc[3 * f] - cc[3 * i], dy = fc[3 * f + 1] - cc[3 * i + 1], dz = fc[3 * f + 2] - cc[3 * i + 2];
            const double l_norm = dx * ni + dy * nj + dz * nk;
            dx = cc[3 * i1] - cc[3 * i];
            dy = cc[3 * i1 + 1] - cc[3 * i + 1];
            dz = cc[3 * i1 + 2] - cc[3 * i + 2];
            const double Ld = std::sqrt(std::pow(dx, 2.0) + std::pow(dy, 2.0) + std::pow(dz, 2.0));
            const double sa = m.face_area[f];
            const double L_norm = dx * ni + dy * nj + dz * nk;
            const double alpha_w = std::pow(2 * l_norm / L_norm, 2.0);
            const double w = alpha_w * sa / Ld;
            xx = xx + w * dx * dx * beta; xy = xy + w * dx * dy * beta; xz = xz + w * dx * dz * beta;
            yx = yx + w * dy * dx * beta; yy = yy + w * dy * dy * beta; yz = yz + w * dy * dz * beta;
            zx = zx + w * dz * dx * beta; zy = zy + w * dz * dy * beta; zz = zz + w * dz * dz * beta;
            double *r = &L.lsq_rhs[((size_t)i * 6 + slot) * 3];
            if (m.owner[f] == i) {
                r[0] = (beta * w * dx + 2 * (1 - beta) * 0.5 * ni * sa);
                r[1] = (beta * w * dy + 2 * (1 - beta) * 0.5 * nj * sa);
                r[2] = (beta * w * dz + 2 * (1 - beta) * 0.5 * nk * sa);
            } else {
                r[0] = (beta * w * dx + 2 * (1 - beta) * 0.5 * ni * -1 * sa);
                r[1] = (beta * w * dy + 2 * (1 - beta) * 0.5 * nj * -1 * sa);
                r[2] = (beta * w * dz + 2 * (1 - beta) * 0.5 * nk * -1 * sa);
            }
        }
        // the reference's determinant expression (gradients.cpp:282-284)
        double det = xx * (yy * zz - yz * zy) + xy * (yx * zz - yz - zx) + xz * (yx * zy - yy * zx);
        if (p.lsq_exact_det) det = xx * (yy * zz - yz * zy) - xy * (yx * zz - yz * zx) + xz * (yx * zy - yy * zx);
        const double a = xx, b = xy, c = xz, d = yx, e = yy, f = yz, g = zx, h = zy, k = zz;
        double *q = &L.lsq_inv9[(size_t)i * 9];
        q[0] = 1 / det * (e * k - f * h);
        q[1] = 1 / det * (c * h - b * k);
        q[2] = 1 / det * (b * f - c * e);
        q[3] = 1 / det * (f * g - d * k);
        q[4] = 1 / det * (a * k - c * g);
        q[5] = 1 / det * (c * d - a * f);
        q[6] = 1 / det * (d * h - e * g);
        q[7] = 1 / det * (b * g - a * h);
        q[8] = 1 / det * (a * e - b * d);
        // CFL areas (Solver.cpp:1653-1686)
        double ax = 0, ay = 0, az = 0;
        for (int s = s_off[i]; s < s_off[i + 1]; s++) {
            const int f2 = s_face[s];
            ax = ax + std::fabs(fn[3 * f2] * m.face_area[f2]);
            ay = ay + std::fabs(fn[3 * f2 + 1] * m.face_area[f2]);
            az = az + std::fabs(fn[3 * f2 + 2] * m.face_area[f2]);
        }
        double v = 0.0;
        v = v + ax / 2;
        v = v + ay / 2;
        v = v + az / 2;
        cfl_file[i] = v;
    }

    // Hilbert order of interior centroids
    L.perm.resize(N);
    L.inv.resize(N);
    {
        double lo[3] = {cc[0], cc[1], cc[2]}, hi[3] = {cc[0], cc[1], cc[2]};
        for (int i = 0; i < N; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], cc[3 * i + a]);
                hi[a] = std::max(hi[a], cc[3 * i + a]);
            }
        double span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        if (!(span > 0)) span = 1.0;
        const int bits = 16;
        const double scale = ((1u << bits) - 1) / span;
        std::vector<std::pair<uint64_t, int32_t>> keys(N);
        for (int i = 0; i < N; i++) {
            uint32_t q3[3];
            for (int a = 0; a < 3; a++) {
                double t = (cc[3 * i + a] - lo[a]) * scale;
                t = std::min(std::max(t, 0.0), (double)((1u << bits) - 1));
                q3[a] = (uint32_t)t;
            }
            keys[i] = {hilbert_key3(q3[0], q3[1], q3[2], bits), i};
        }
        std::sort(keys.begin(), keys.end());
        for (int n = 0; n < N; n++) {
            L.perm[n] = keys[n].second;
            L.inv[keys[n].second] = n;
        }
    }

    // per-cell arrays in the new order, folded stencil weights
    L.st_cnt.assign(N, 0);
    L.st_nbr.assign((size_t)kMaxSlots * N, 0);
    L.st_w.assign((size_t)kMaxSlots * 3 * N, 0.0);
    L.vol.resize(N);
    L.inv_vol.resize(N);
    L.centroid.resize((size_t)3 * N);
    for (int n = 0; n < N; n++) {
        const int i = L.perm[n];
        for (int a = 0; a < 3; a++) L.centroid[(size_t)a * N + n] = cc[3 * i + a];
        const int cnt = s_off[i + 1] - s_off[i];
        L.st_cnt[n] = (uint8_t)cnt;
        const double *q = &L.lsq_inv9[(size_t)i * 9];
        for (int s = 0; s < cnt; s++) {
            const int i1 = s_cell[s_off[i] + s];
            L.st_nbr[(size_t)s * N + n] = (i1 < N) ? L.inv[i1] : -1 - (i1 - N);
            const double *r = &L.lsq_rhs[((size_t)i * 6 + s) * 3];
            for (int c = 0; c < 3; c++)
                L.st_w[((size_t)s * 3 + c) * N + n] = q[3 * c] * r[0] + q[3 * c + 1] * r[1] + q[3 * c + 2] * r[2];
        }
        for (int s = cnt; s < kMaxSlots; s++) L.st_nbr[(size_t)s * N + n] = n;
        L.vol[n] = m.cell_volume[i];
        L.inv_vol[n] = 1.0 / m.cell_volume[i];
        L.cfl_area[n] = cfl_file[i];
    }

    // residual of the uniform rest state rho0 (see layout.hpp)
    {
        std::vector<double> cf((size_t)3 * N, 0.0);
        const double third = p.ref_rho / 3.0;
        for (int f = 0; f < F; f++) {
            if (f > NI && patches[bface_patch[f - NI]].is_empty) continue;
            const int o = m.owner[f];
            for (int a = 0; a < 3; a++) {
                const double v = third * fn[3 * f + a] * m.face_area[f];
                cf[3 * o + a] -= v / m.cell_volume[o];
                if (f < NI) cf[3 * m.neighbour[f] + a] += v / m.cell_volume[m.neighbour[f]];
            }
        }
        L.corr.resize((size_t)3 * N);
        for (int n = 0; n < N; n++)
            for (int a = 0; a < 3; a++) L.corr[(size_t)a * N + n] = cf[3 * L.perm[n] + a];
    }

    // periodic partners: the interior cell nearest to the owner's centroid
    // + per_translate (src/Boundary_Conditions.cpp:425-445, owner in place of
    // the reference's uninitialised index); -1 on other patches
    L.b_partner.assign(G, -1);
    {
        bool any = false;
        for (int q = 0; q < n_patches; q++) any = any || patches[q].rho_type == 3 || patches[q].vel_type == 3;
        if (any) {
            NearestCell nc(cc, N);
            for (int b = 0; b < G; b++) {
                const plbfs_patch &P = patches[bface_patch[b]];
                if (P.rho_type != 3 && P.vel_type != 3) continue;
                const int o = m.owner[NI + b];
                const double x[3] = {cc[3 * o] + P.per_translate[0], cc[3 * o + 1] + P.per_translate[1],
                                     cc[3 * o + 2] + P.per_translate[2]};
                L.b_partner[b] = L.inv[nc.find(x)];
            }
            // a periodic ghost is its partner (state and gradient), so the
            // gradient stencil reads the partner directly
            for (size_t k = 0; k < L.st_nbr.size(); k++) {
                const int j = L.st_nbr[k];
                if (j < 0 && L.b_partner[-1 - j] >= 0) L.st_nbr[k] = L.b_partner[-1 - j];
            }
        }
    }

    // ghosts the reference leaves (partly) unwritten: fixed values, and the
    // gradient of cell 0, its periodic node on a patch that is not periodic
    // (src/Boundary_Conditions.cpp:416); parabolic u, v at the centroid of
    // cell b, b the boundary index (src/Solution.cpp:804-813)
    L.b_fixed.assign((size_t)4 * G, 0.0);
    L.needs_ghost_state = false;
    for (int b = 0; b < G; b++) {
        const plbfs_patch &P = patches[bface_patch[b]];
        const bool kr = rho_keeps_ghost(P.rho_type), kv = vel_keeps_ghost(P.vel_type);
        if (!kr && !kv) continue;
        L.needs_ghost_state = true;
        L.b_partner[b] = L.inv[0];
        const double xb = cc[3 * b], yb = cc[3 * b + 1];
        if (P.vel_type == 4) {
            const double Y = p.domain_y;
            L.b_fixed[(size_t)4 * b + 1] = 4 * P.u / std::pow(Y, 2) * yb * (Y - yb);
            L.b_fixed[(size_t)4 * b + 2] = 4 * P.v / std::pow(Y, 2) * yb * (Y - yb);
        } else if (P.vel_type == 5) {
            const double X = p.domain_x;
            L.b_fixed[(size_t)4 * b + 1] = 4 * P.u / std::pow(X, 2) * xb * (X - yb);
            L.b_fixed[(size_t)4 * b + 2] = 4 * P.v / std::pow(X, 2) * xb * (X - xb);
        }
    }

    // boundary faces
    L.b_patch.resize(G);
    L.b_dmag.resize(G);
    L.b_owner.resize(G);
    for (int b = 0; b < G; b++) {
        const int f = NI + b, o = m.owner[f], g = N + b;
        L.b_patch[b] = bface_patch[b];
        const double dx = cc[3 * g] - cc[3 * o], dy = cc[3 * g + 1] - cc[3 * o + 1], dz = cc[3 * g + 2] - cc[3 * o + 2];
        L.b_dmag[b] = std::sqrt(std::pow(dx, 2.0) + std::pow(dy, 2.0) + std::pow(dz, 2.0));
        L.b_owner[b] = L.inv[o];
    }

    // block face items and per-cell gather lists
    const int C = block_cells;
    L.n_blocks = (N + C - 1) / C;
    L.blk_item_off.assign(L.n_blocks + 1, 0);
    L.g_off.assign(N + 1, 0);
    L.g_ent.clear();
    L.g_ent.reserve(s_off[N]);
    std::vector<int32_t> mark(F, -1), slot_of(F, 0), border_of(F, -1);
    std::vector<int32_t> item_face, border_face;
    item_face.reserve((size_t)F + F / 4);
    int max_items = 0;
    for (int blk = 0; blk < L.n_blocks; blk++) {
        const int c0 = blk * C, c1 = std::min(N, c0 + C);
        const int first_item = (int)item_face.size();
        for (int n = c0; n < c1; n++) {
            const int i = L.perm[n];
            for (int s = s_off[i]; s < s_off[i + 1]; s++) {
                const int f = s_face[s];
                if (f > NI && patches[bface_patch[f - NI]].is_empty) continue;  // Solver.cpp:308
                const int own = m.owner[f] == i ? 1 : 0;
                if (split_border && f < NI && L.inv[m.owner[f]] / C != L.inv[m.neighbour[f]] / C) {
                    if (border_of[f] < 0) {
                        border_of[f] = (int32_t)border_face.size();
                        border_face.push_back(f);
                    }
                    L.g_ent.push_back((border_of[f] << 2) | 2 | own);
                    continue;
                }
                if (mark[f] != blk) {
                    mark[f] = blk;
                    slot_of[f] = (int)item_face.size() - first_item;
                    item_face.push_back(f);
                }
                L.g_ent.push_back((slot_of[f] << 2) | own);
            }
            L.g_off[n + 1] = (int32_t)L.g_ent.size();
        }
        L.blk_item_off[blk + 1] = (int32_t)item_face.size();
        max_items = std::max(max_items, (int)item_face.size() - first_item);
    }
    L.max_items_per_block = max_items;
    L.border0 = (int64_t)item_face.size();
    L.n_border = (int64_t)border_face.size();
    item_face.insert(item_face.end(), border_face.begin(), border_face.end());
    const size_t E = item_face.size();
    L.n_items = (int64_t)E;
    L.it_own.resize(E);
    L.it_nbr.resize(E);
    L.it_side.resize(E);
    L.it_geo.resize((size_t)kGeoPerItem * E);
    L.it_fc.resize((size_t)3 * E);
    if (G > 0)
        for (int a = 0; a < 3; a++) L.ghost0[a] = cc[3 * N + a];
    static const int kPairs[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
    for (size_t e = 0; e < E; e++) {
        const int f = item_face[e];
        const int o = m.owner[f], nb = m.neighbour[f];
        L.it_own[e] = L.inv[o];
        L.it_nbr[e] = (f < NI) ? L.inv[nb] : -1 - (f - NI);
        uint32_t side = 0;
        for (int pr = 0; pr < 7; pr++) {
            double c = 0.0;
            for (int a = 0; a < 3; a++)
                if (kPairs[pr][a]) c = c + kPairs[pr][a] * fn[3 * f + a];
            if (c < 0.0) side |= 1u << (2 * pr);
            if (c > 0.0) side |= 1u << (2 * pr + 1);
        }
        L.it_side[e] = side;
        double g[kGeoPerItem];
        g[0] = fn[3 * f];
        g[1] = fn[3 * f + 1];
        g[2] = fn[3 * f + 2];
        g[3] = m.face_area[f];
        g[4] = m.face_dt[f];
        g[5] = cc[3 * o] - fc[3 * f];
        g[6] = cc[3 * o + 1] - fc[3 * f + 1];
        g[7] = cc[3 * o + 2] - fc[3 * f + 2];
        g[8] = cc[3 * nb] - fc[3 * f];
        g[9] = cc[3 * nb + 1] - fc[3 * f + 1];
        g[10] = cc[3 * nb + 2] - fc[3 * f + 2];
        for (int q = 0; q < kGeoPerItem; q++) L.it_geo[(size_t)q * E + e] = g[q];
        for (int a = 0; a < 3; a++) L.it_fc[(size_t)a * E + e] = fc[3 * f + a];
    }
    return PLBFS_OK;
}

}  // namespace plbfs
