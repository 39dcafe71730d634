// Host mesh: OpenFOAM ASCII polyMesh reader and FP64 cell/face geometry.
//
// Follows the reference's unstructured_mesh (src/unstructured_mesh.cpp):
//   import_openfoam_mesh        :236-507   (counts from the note on line 12 of owner,
//                                           list bodies between '(' and ')' lines,
//                                           points read as float)
//   import_openfoam_BC_tags     :1912-1987 (patch type token per boundary face)
//   generate_internal_mesh      :1133-1233 (fan triangulation from vertex 0)
//   tet_volume / cell_tet_area  :577-607, 732-766
//   calc_face_delta_t           :979-983
//   populate_ghost_cell_from_foam_face :1346-1375
// Geometry is evaluated in IEEE double in the reference's operation order
// (this file is built with -ffp-contract=off).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "plbfs_case.h"
#include "host_common.hpp"

struct plbfs_mesh {
    int32_t n_points = 0, n_cells = 0, n_faces = 0, n_internal = 0;
    std::vector<float> px, py, pz;
    std::vector<int32_t> face_off, face_verts, owner, neighbour;  // neighbour [F] incl. ghosts
    std::vector<std::string> patch_types;
    std::vector<int32_t> patch_sizes, bface_patch;
    // geometry
    std::vector<double> cell_centroid;  // 3*(N+G)
    std::vector<double> cell_volume;    // N+G (ghosts 0)
    std::vector<double> face_centroid, face_normal, face_area, face_dt;
};

namespace {

bool read_file(const std::string &path, std::string &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    size_t got = n > 0 ? std::fread(&out[0], 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == out.size();
}

// Splits into lines and calls fn(line_index, begin, end) for list-body lines:
// a line whose first character is '(' opens the list (the line itself is not
// data), a line whose first character is ')' closes it.  As in the reference,
// the "first character" of an empty line is the previous line's.
template <class Fn>
void for_list_lines(const std::string &text, Fn fn) {
    bool start = false;
    char comp = 0;
    size_t pos = 0, n = text.size();
    long t = 0;
    while (pos < n) {
        size_t e = text.find('\n', pos);
        if (e == std::string::npos) e = n;
        size_t le = e;
        if (le > pos && text[le - 1] == '\r') le--;
        if (le > pos) comp = text[pos];
        if (comp == ')') start = false;
        if (start) fn(t, text.data() + pos, text.data() + le);
        if (comp == '(') start = true;
        pos = e + 1;
        t++;
    }
}

bool parse_note(const std::string &line, const char *key, long &val) {
    size_t p = line.find(key);
    if (p == std::string::npos) return false;
    const char *s = line.c_str() + p + std::strlen(key);
    char *end = nullptr;
    errno = 0;
    val = std::strtol(s, &end, 10);
    return end != s && errno == 0;
}

}  // namespace

// Geometry from points/faces (src/unstructured_mesh.cpp:1133-1233).
static int build_geometry(plbfs_mesh &m, double visc, double gamma, double fneq_min) {
    const int N = m.n_cells, F = m.n_faces, NI = m.n_internal, G = F - NI, T = N + G;
    m.cell_centroid.assign((size_t)3 * T, 0.0);
    m.cell_volume.assign(T, 0.0);
    m.face_centroid.assign((size_t)3 * F, 0.0);
    m.face_normal.assign((size_t)3 * F, 0.0);
    m.face_area.assign(F, 0.0);
    m.face_dt.assign(F, 1.0);
    std::vector<double> tot_area(T, 0.0);
    double *cc = m.cell_centroid.data();
    const double refdt = visc * 3 / gamma / fneq_min;
    for (int f = 0; f < F; f++) {
        const int32_t *fv = m.face_verts.data() + m.face_off[f];
        const int nv = m.face_off[f + 1] - m.face_off[f];
        const int o = m.owner[f];
        const double ax = m.px[fv[0]], ay = m.py[fv[0]], az = m.pz[fv[0]];
        for (int t = 0; t < nv - 2; t++) {
            const double bx = m.px[fv[t + 1]] - ax, by = m.py[fv[t + 1]] - ay, bz = m.pz[fv[t + 1]] - az;
            const double qx = m.px[fv[t + 2]] - ax, qy = m.py[fv[t + 2]] - ay, qz = m.pz[fv[t + 2]] - az;
            const double cx = by * qz - bz * qy, cy = bz * qx - bx * qz, cz = bx * qy - by * qx;
            const double vol = (ax * cx + ay * cy + az * cz) / 6;
            m.cell_volume[o] = m.cell_volume[o] + vol;
            if (f < NI) m.cell_volume[m.neighbour[f]] = m.cell_volume[m.neighbour[f]] - vol;
        }
        double area = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
        for (int t = 0; t < nv - 2; t++) {
            const double bx0 = m.px[fv[t + 1]], by0 = m.py[fv[t + 1]], bz0 = m.pz[fv[t + 1]];
            const double qx0 = m.px[fv[t + 2]], qy0 = m.py[fv[t + 2]], qz0 = m.pz[fv[t + 2]];
            double gx = (ax + bx0 + qx0) / 3, gy = (ay + by0 + qy0) / 3, gz = (az + bz0 + qz0) / 3;
            const double bx = bx0 - ax, by = by0 - ay, bz = bz0 - az;
            const double qx = qx0 - ax, qy = qy0 - ay, qz = qz0 - az;
            const double cx = by * qz - bz * qy, cy = bz * qx - bx * qz, cz = bx * qy - by * qx;
            const double a = 0.5 * std::sqrt(std::pow(cx, 2) + std::pow(cy, 2) + std::pow(cz, 2));
            area = area + a;
            const double inv = 1 / a;  // vector_var::factor divides (src/vector_var.cpp:102-107)
            gx = gx / inv;
            gy = gy / inv;
            gz = gz / inv;
            sx = sx + gx; sy = sy + gy; sz = sz + gz;
            nx = nx + cx; ny = ny + cy; nz = nz + cz;
        }
        m.face_area[f] = area;
        tot_area[o] = tot_area[o] + area;
        cc[3 * o] = cc[3 * o] + sx;
        cc[3 * o + 1] = cc[3 * o + 1] + sy;
        cc[3 * o + 2] = cc[3 * o + 2] + sz;
        m.face_centroid[3 * f] = sx / area;
        m.face_centroid[3 * f + 1] = sy / area;
        m.face_centroid[3 * f + 2] = sz / area;
        const double nm = std::sqrt(std::pow(nx, 2) + std::pow(ny, 2) + std::pow(nz, 2));
        m.face_normal[3 * f] = nx / nm;
        m.face_normal[3 * f + 1] = ny / nm;
        m.face_normal[3 * f + 2] = nz / nm;
        const double d = 0.5 * std::sqrt(area);
        m.face_dt[f] = (refdt < d) ? refdt : d;  // std::min(delta_t_face, reference_dt)
        if (f < NI) {
            const int nb = m.neighbour[f];
            tot_area[nb] = tot_area[nb] + area;
            cc[3 * nb] = cc[3 * nb] + sx;
            cc[3 * nb + 1] = cc[3 * nb + 1] + sy;
            cc[3 * nb + 2] = cc[3 * nb + 2] + sz;
        }
    }
    for (int i = 0; i < N; i++) {
        if (!(tot_area[i] > 0.0)) return plbfs_fail(PLBFS_ERR_INPUT, "cell %d has no faces", i);
        cc[3 * i] = cc[3 * i] / tot_area[i];
        cc[3 * i + 1] = cc[3 * i + 1] / tot_area[i];
        cc[3 * i + 2] = cc[3 * i + 2] / tot_area[i];
    }
    for (int f = NI; f < F; f++) {
        const int g = N + f - NI, j = m.owner[f];
        m.neighbour[f] = g;
        double rx = m.face_centroid[3 * f] - cc[3 * j];
        double ry = m.face_centroid[3 * f + 1] - cc[3 * j + 1];
        double rz = m.face_centroid[3 * f + 2] - cc[3 * j + 2];
        rx = rx / 0.5;
        ry = ry / 0.5;
        rz = rz / 0.5;
        cc[3 * g] = cc[3 * j] + rx * std::fabs(m.face_normal[3 * f]);
        cc[3 * g + 1] = cc[3 * j + 1] + ry * std::fabs(m.face_normal[3 * f + 1]);
        cc[3 * g + 2] = cc[3 * j + 2] + rz * std::fabs(m.face_normal[3 * f + 2]);
    }
    return PLBFS_OK;
}

static int validate_topology(const plbfs_mesh &m) {
    if (m.n_cells <= 0 || m.n_faces <= 0 || m.n_internal < 0 || m.n_internal > m.n_faces)
        return plbfs_fail(PLBFS_ERR_INPUT, "bad mesh sizes N=%d F=%d NI=%d", m.n_cells, m.n_faces, m.n_internal);
    if ((int)m.face_off.size() != m.n_faces + 1) return plbfs_fail(PLBFS_ERR_INPUT, "face list size mismatch");
    for (int f = 0; f < m.n_faces; f++) {
        int nv = m.face_off[f + 1] - m.face_off[f];
        if (nv < 3) return plbfs_fail(PLBFS_ERR_INPUT, "face %d has %d vertices", f, nv);
        for (int q = m.face_off[f]; q < m.face_off[f + 1]; q++)
            if (m.face_verts[q] < 0 || m.face_verts[q] >= m.n_points)
                return plbfs_fail(PLBFS_ERR_INPUT, "face %d vertex out of range", f);
        if (m.owner[f] < 0 || m.owner[f] >= m.n_cells) return plbfs_fail(PLBFS_ERR_INPUT, "face %d owner", f);
        if (f < m.n_internal && (m.neighbour[f] < 0 || m.neighbour[f] >= m.n_cells))
            return plbfs_fail(PLBFS_ERR_INPUT, "face %d neighbour", f);
    }
    long total = 0;
    for (int32_t s : m.patch_sizes) {
        if (s < 0) return plbfs_fail(PLBFS_ERR_INPUT, "negative patch size %d", s);
        total += s;
    }
    if (total != m.n_faces - m.n_internal)
        return plbfs_fail(PLBFS_ERR_INPUT, "patch sizes sum %ld != boundary faces %d", total,
                          m.n_faces - m.n_internal);
    return PLBFS_OK;
}

static void fill_bface_patch(plbfs_mesh &m) {
    m.bface_patch.clear();
    for (size_t p = 0; p < m.patch_sizes.size(); p++)
        for (int q = 0; q < m.patch_sizes[p]; q++) m.bface_patch.push_back((int32_t)p);
}

static int read_openfoam(const char *prefix, double visc, double gamma, double fneq_min, plbfs_mesh **out) {
    std::string pre(prefix), text;
    std::unique_ptr<plbfs_mesh> m(new plbfs_mesh());
    auto bail = [](int rc) { return rc; };
    // owner: counts from the note on 0-based line 12
    if (!read_file(pre + "owner", text)) return bail(plbfs_fail(PLBFS_ERR_INPUT, "cannot read %sowner", prefix));
    {
        size_t pos = 0;
        std::string line;
        for (int t = 0; t <= 12 && pos <= text.size(); t++) {
            size_t e = text.find('\n', pos);
            if (e == std::string::npos) e = text.size();
            line = text.substr(pos, e - pos);
            pos = e + 1;
        }
        long nc = 0, nf = 0, ni = 0, np = 0;
        if (!parse_note(line, "nCells:", nc) || !parse_note(line, "nFaces:", nf) ||
            !parse_note(line, "nInternalFaces:", ni) || !parse_note(line, "nPoints:", np))
            return bail(plbfs_fail(PLBFS_ERR_INPUT, "owner line 12 lacks the nPoints/nCells/nFaces note"));
        // every owner entry takes at least two bytes of the file
        if (nc <= 0 || nf <= 0 || np <= 0 || ni < 0 || ni > nf || nc > INT32_MAX || np > INT32_MAX ||
            nf > (long)(text.size() / 2))
            return bail(plbfs_fail(PLBFS_ERR_INPUT, "owner note counts out of range: nPoints %ld nCells %ld nFaces %ld "
                                   "nInternalFaces %ld", np, nc, nf, ni));
        m->n_cells = (int32_t)nc;
        m->n_faces = (int32_t)nf;
        m->n_internal = (int32_t)ni;
        m->n_points = (int32_t)np;
    }
    m->owner.assign(m->n_faces, 0);
    {
        int i = 0;
        for_list_lines(text, [&](long, const char *b, const char *) {
            if (i < m->n_faces) m->owner[i++] = (int32_t)std::strtol(b, nullptr, 10);
        });
        if (i != m->n_faces) return bail(plbfs_fail(PLBFS_ERR_INPUT, "owner has %d entries, expected %d", i, m->n_faces));
    }
    if (!read_file(pre + "points", text)) return bail(plbfs_fail(PLBFS_ERR_INPUT, "cannot read %spoints", prefix));
    // a point line "(x y z)" takes at least eight bytes
    if ((size_t)m->n_points > text.size() / 8)
        return bail(plbfs_fail(PLBFS_ERR_INPUT, "points file too short for %d points", m->n_points));
    m->px.assign(m->n_points, 0.f);
    m->py.assign(m->n_points, 0.f);
    m->pz.assign(m->n_points, 0.f);
    {
        int i = 0;
        bool bad = false;
        for_list_lines(text, [&](long, const char *b, const char *) {
            if (i >= m->n_points) {
                bad = true;
                return;
            }
            const char *s = b + 1;  // skip '('
            char *e1, *e2, *e3;
            m->px[i] = std::strtof(s, &e1);
            m->py[i] = std::strtof(e1, &e2);
            m->pz[i] = std::strtof(e2, &e3);
            if (e1 == s || e2 == e1 || e3 == e2) bad = true;
            i++;
        });
        if (bad || i != m->n_points) return bail(plbfs_fail(PLBFS_ERR_INPUT, "points list malformed"));
    }
    if (!read_file(pre + "neighbour", text)) return bail(plbfs_fail(PLBFS_ERR_INPUT, "cannot read %sneighbour", prefix));
    m->neighbour.assign(m->n_faces, 0);  // boundary faces 0 until ghosts are assigned (:413-416)
    {
        int i = 0;
        for_list_lines(text, [&](long, const char *b, const char *) {
            if (i < m->n_internal) m->neighbour[i] = (int32_t)std::strtol(b, nullptr, 10);
            i++;
        });
        if (i < m->n_internal) return bail(plbfs_fail(PLBFS_ERR_INPUT, "neighbour list too short"));
    }
    if (!read_file(pre + "faces", text)) return bail(plbfs_fail(PLBFS_ERR_INPUT, "cannot read %sfaces", prefix));
    m->face_off.assign(1, 0);
    m->face_off.reserve(m->n_faces + 1);
    m->face_verts.reserve((size_t)m->n_faces * 4);
    {
        bool bad = false;
        for_list_lines(text, [&](long, const char *b, const char *) {
            char *e;
            long nv = std::strtol(b, &e, 10);
            if (e == b || *e != '(' || nv < 3) {
                bad = true;
                return;
            }
            const char *s = e + 1;
            for (long q = 0; q < nv; q++) {
                char *e2;
                long v = std::strtol(s, &e2, 10);
                if (e2 == s) {
                    bad = true;
                    return;
                }
                m->face_verts.push_back((int32_t)v);
                s = e2;
            }
            m->face_off.push_back((int32_t)m->face_verts.size());
        });
        if (bad || (int)m->face_off.size() != m->n_faces + 1)
            return bail(plbfs_fail(PLBFS_ERR_INPUT, "faces list malformed (%zu faces, expected %d)",
                                   m->face_off.size() - 1, m->n_faces));
    }
    if (!read_file(pre + "boundary", text)) return bail(plbfs_fail(PLBFS_ERR_INPUT, "cannot read %sboundary", prefix));
    {
        std::string bc_type;
        bool bad = false;
        for_list_lines(text, [&](long, const char *b, const char *e) {
            std::istringstream iss(std::string(b, e));
            std::string t1, t2;
            iss >> t1 >> t2;
            if (t1 == "type" && !t2.empty()) {
                t2.pop_back();
                bc_type = t2;
            }
            if (t1 == "nFaces" && !t2.empty()) {
                t2.pop_back();
                char *end = nullptr;
                errno = 0;
                long nfp = std::strtol(t2.c_str(), &end, 10);
                if (end == t2.c_str() || *end != '\0' || errno != 0 || nfp < 0 || nfp > INT32_MAX) {
                    bad = true;
                    return;
                }
                m->patch_types.push_back(bc_type);
                m->patch_sizes.push_back((int32_t)nfp);
            }
        });
        if (bad) return bail(plbfs_fail(PLBFS_ERR_INPUT, "boundary: nFaces is not a face count"));
    }
    int rc = validate_topology(*m);
    if (rc != PLBFS_OK) return bail(rc);
    fill_bface_patch(*m);
    rc = build_geometry(*m, visc, gamma, fneq_min);
    if (rc != PLBFS_OK) return bail(rc);
    *out = m.release();
    return PLBFS_OK;
}

extern "C" int plbfs_mesh_read_openfoam(const char *prefix, double visc, double gamma, double fneq_min,
                                        plbfs_mesh **out) {
    if (!prefix || !out) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    *out = nullptr;
    return plbfs_guarded("plbfs_mesh_read_openfoam", [&] { return read_openfoam(prefix, visc, gamma, fneq_min, out); });
}

static int from_arrays(int32_t n_points, const float *px, const float *py, const float *pz, int32_t n_cells,
                       int32_t n_faces, int32_t n_internal, const int32_t *face_off, const int32_t *face_verts,
                       const int32_t *owner, const int32_t *neighbour, int32_t n_patches,
                       const char *const *patch_types, const int32_t *patch_sizes, double visc, double gamma,
                       double fneq_min, plbfs_mesh **out) {
    std::unique_ptr<plbfs_mesh> m(new plbfs_mesh());
    m->n_points = n_points;
    m->n_cells = n_cells;
    m->n_faces = n_faces;
    m->n_internal = n_internal;
    m->px.assign(px, px + n_points);
    m->py.assign(py, py + n_points);
    m->pz.assign(pz, pz + n_points);
    m->face_off.assign(face_off, face_off + n_faces + 1);
    for (int f = 0; f < n_faces; f++)
        if (face_off[f + 1] < face_off[f]) return plbfs_fail(PLBFS_ERR_INPUT, "face_off decreases at face %d", f);
    if (m->face_off[0] != 0) return plbfs_fail(PLBFS_ERR_INPUT, "face_off[0] must be 0");
    m->face_verts.assign(face_verts, face_verts + face_off[n_faces]);
    m->owner.assign(owner, owner + n_faces);
    m->neighbour.assign(n_faces, 0);
    for (int f = 0; f < n_internal; f++) m->neighbour[f] = neighbour[f];
    for (int p = 0; p < n_patches; p++) {
        m->patch_types.push_back(patch_types && patch_types[p] ? patch_types[p] : "");
        m->patch_sizes.push_back(patch_sizes[p]);
    }
    int rc = validate_topology(*m);
    if (rc != PLBFS_OK) return rc;
    fill_bface_patch(*m);
    rc = build_geometry(*m, visc, gamma, fneq_min);
    if (rc != PLBFS_OK) return rc;
    *out = m.release();
    return PLBFS_OK;
}

extern "C" int plbfs_mesh_from_arrays(int32_t n_points, const float *px, const float *py, const float *pz,
                                      int32_t n_cells, int32_t n_faces, int32_t n_internal, const int32_t *face_off,
                                      const int32_t *face_verts, const int32_t *owner, const int32_t *neighbour,
                                      int32_t n_patches, const char *const *patch_types,
                                      const int32_t *patch_sizes, double visc, double gamma, double fneq_min,
                                      plbfs_mesh **out) {
    if (!out || !px || !py || !pz || !face_off || !face_verts || !owner || (n_internal > 0 && !neighbour) ||
        (n_patches > 0 && !patch_sizes) || n_points <= 0 || n_cells <= 0 || n_faces <= 0 || n_internal < 0 ||
        n_internal > n_faces || n_patches < 0)
        return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments to plbfs_mesh_from_arrays");
    *out = nullptr;
    return plbfs_guarded("plbfs_mesh_from_arrays", [&] {
        return from_arrays(n_points, px, py, pz, n_cells, n_faces, n_internal, face_off, face_verts, owner, neighbour,
                           n_patches, patch_types, patch_sizes, visc, gamma, fneq_min, out);
    });
}

extern "C" void plbfs_mesh_free(plbfs_mesh *m) { delete m; }

extern "C" int plbfs_mesh_info(const plbfs_mesh *m, int64_t *info) {
    if (!m || !info) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    info[0] = m->n_cells;
    info[1] = m->n_faces;
    info[2] = m->n_internal;
    info[3] = m->n_points;
    info[4] = (int64_t)m->patch_types.size();
    info[5] = m->n_faces - m->n_internal;
    return PLBFS_OK;
}

extern "C" int plbfs_mesh_view_get(const plbfs_mesh *m, plbfs_mesh_view *v) {
    if (!m || !v) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    v->n_cells = m->n_cells;
    v->n_faces = m->n_faces;
    v->n_internal_faces = m->n_internal;
    v->owner = m->owner.data();
    v->neighbour = m->neighbour.data();
    v->cell_centroid = m->cell_centroid.data();
    v->cell_volume = m->cell_volume.data();
    v->face_centroid = m->face_centroid.data();
    v->face_normal = m->face_normal.data();
    v->face_area = m->face_area.data();
    v->face_dt = m->face_dt.data();
    return PLBFS_OK;
}

extern "C" const int32_t *plbfs_mesh_bface_patch(const plbfs_mesh *m) { return m ? m->bface_patch.data() : nullptr; }

extern "C" int plbfs_mesh_patch_type(const plbfs_mesh *m, int32_t p, char *buf, int32_t len) {
    if (!m || !buf || len <= 0 || p < 0 || p >= (int32_t)m->patch_types.size())
        return plbfs_fail(PLBFS_ERR_INPUT, "bad patch index");
    std::snprintf(buf, (size_t)len, "%s", m->patch_types[p].c_str());
    return PLBFS_OK;
}

extern "C" int64_t plbfs_mesh_n_face_verts(const plbfs_mesh *m) { return m ? (int64_t)m->face_verts.size() : -1; }

extern "C" int plbfs_mesh_raw(const plbfs_mesh *m, float *px, float *py, float *pz, int32_t *owner,
                              int32_t *neighbour_internal, int32_t *face_off, int32_t *face_verts) {
    if (!m) return plbfs_fail(PLBFS_ERR_INPUT, "null mesh");
    if (px) std::memcpy(px, m->px.data(), sizeof(float) * m->n_points);
    if (py) std::memcpy(py, m->py.data(), sizeof(float) * m->n_points);
    if (pz) std::memcpy(pz, m->pz.data(), sizeof(float) * m->n_points);
    if (owner) std::memcpy(owner, m->owner.data(), sizeof(int32_t) * m->n_faces);
    if (neighbour_internal) std::memcpy(neighbour_internal, m->neighbour.data(), sizeof(int32_t) * m->n_internal);
    if (face_off) std::memcpy(face_off, m->face_off.data(), sizeof(int32_t) * (m->n_faces + 1));
    if (face_verts) std::memcpy(face_verts, m->face_verts.data(), sizeof(int32_t) * m->face_verts.size());
    return PLBFS_OK;
}
