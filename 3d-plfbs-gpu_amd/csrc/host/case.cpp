// XML case layer and whole-program run (include/plbfs_case.h).
//
// Parameter derivation follows src/preprocessor.cpp:27-84 (velocities scaled by
// mach_no/scale), :187-302 (boundary types by token), :438-515 (globals and
// initial conditions), src/global_variables.cpp:20-56 (nu, tau, preconditioned
// gamma), src/Boundary_Conditions.cpp:388-452 (patch lookup by type token),
// src/Solution.cpp:94-189 (initial conditions) and src/program.cpp:40-137 (the
// run sequence).  Lagrangian objects: src/lagrangian_object.cpp:79-113 (rigid
// cylinder); the rigid sphere with a Fibonacci node lattice has no reference
// counterpart.
#include <sys/stat.h>

#include <chrono>
#include <climits>
#include <fstream>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <algorithm>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "plbfs_case.h"
#include "xml.hpp"

struct plbfs_mesh;  // defined in mesh.cpp

namespace {

// Upper bound on the nodes of one Lagrangian object read from a case.
constexpr int kMaxObjectNodes = 100000000;

struct BcSpec {
    std::string name;
    double rho = 0, u = 0, v = 0, w = 0;  // u,v,w scaled by the mach factor
    int vel_type = 0, rho_type = 0;
};

struct ObjSpec {
    std::string name, type, mesh_file;
    int num_nodes = 0;
    double radius = 0, depth = 0, stiffness = 0, cx = 0, cy = 0, cz = 0;
    int kernel = PLBFS_DELTA_HAT2;
};

int bc_code(const std::string &t) {
    static const std::map<std::string, int> m = {{"dirichlet", 1}, {"neumann", 2},    {"periodic", 3},
                                                 {"parabolic-N", 4}, {"parabolic-W", 5}, {"inlet", 6},
                                                 {"wall", 7},      {"symmetry-x", 8}, {"parabolic", 9}};
    auto it = m.find(t);
    return it == m.end() ? 0 : it->second;
}

bool read_text(const std::string &path, std::string &out) {
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

void mkdirs(const std::string &p) {
    std::string cur;
    for (size_t i = 0; i < p.size(); i++) {
        cur += p[i];
        if (p[i] == '/' || i + 1 == p.size()) ::mkdir(cur.c_str(), 0755);
    }
}

}  // namespace

struct plbfs_case {
    std::map<std::string, double> num;      // "section/key" -> number
    std::map<std::string, std::string> str;  // "section/key" -> text
    std::vector<BcSpec> bcs;
    std::vector<ObjSpec> objects;
    plbfs_params params{};
    std::string xml_path;

    double n(const std::string &k, double def = 0.0) const {
        auto it = num.find(k);
        return it == num.end() ? def : it->second;
    }
    std::string s(const std::string &k, const std::string &def = "") const {
        auto it = str.find(k);
        return it == str.end() ? def : it->second;
    }
};

static void flatten(const plbfs::XmlNode &node, const std::string &prefix, plbfs_case &c) {
    for (auto &ch : node.children) {
        const std::string key = prefix.empty() ? ch->name : prefix + "/" + ch->name;
        if (ch->children.empty()) {
            c.str[key] = ch->text;
            char *end = nullptr;
            double v = std::strtod(ch->text.c_str(), &end);
            if (end != ch->text.c_str() && end && *end == '\0') c.num[key] = v;
        } else if (ch->name != "bc" && ch->name != "object") {
            flatten(*ch, key, c);
        }
    }
}

static int case_load(const char *xml_path, plbfs_case **out) {
    std::string text, err;
    if (!read_text(xml_path, text)) return plbfs_fail(PLBFS_ERR_INPUT, "cannot read %s", xml_path);
    auto root = plbfs::xml_parse(text, err);
    if (!root) return plbfs_fail(PLBFS_ERR_INPUT, "%s: %s", xml_path, err.c_str());
    std::unique_ptr<plbfs_case> c(new plbfs_case());
    c->xml_path = xml_path;
    flatten(*root, "", *c);
    // required elements (the reference dereferences each without checks)
    const char *required[] = {"global_variables/tolerance", "global_variables/simulation_length",
                              "global_variables/time_marching_step", "global_variables/reynolds_no",
                              "global_variables/mach_no", "global_variables/import_file",
                              "geometry/ref_length", "initial_conditions/average_rho"};
    for (const char *k : required) {
        if (!c->num.count(k) && !c->str.count(k))
            return plbfs_fail(PLBFS_ERR_INPUT, "%s: missing element %s", xml_path, k);
        if (std::strcmp(k, "global_variables/import_file") != 0 && !c->num.count(k))
            return plbfs_fail(PLBFS_ERR_INPUT, "%s: element %s is not a number", xml_path, k);
    }
    const double U = c->n("global_variables/mach_no");
    const double scale = c->n("global_variables/scale", 1.0);
    const double factor = U / scale;  // preprocessor.cpp:58,190
    if (const plbfs::XmlNode *ub = root->child("uns_boundary_conditions")) {
        for (auto &bc : ub->children) {
            if (bc->name != "bc") continue;
            BcSpec b;
            auto t = [&](const char *k) {
                const plbfs::XmlNode *e = bc->child(k);
                return e ? e->text : std::string();
            };
            b.name = t("name");
            b.rho = std::atof(t("rho").c_str());
            b.u = std::atof(t("u").c_str()) * factor;
            b.v = std::atof(t("v").c_str()) * factor;
            b.w = std::atof(t("w").c_str()) * factor;
            b.vel_type = bc_code(t("vel_type"));
            b.rho_type = bc_code(t("rho_type"));
            c->bcs.push_back(b);
        }
    }
    if (const plbfs::XmlNode *lo = root->child("lagrangian_objects")) {
        for (auto &ob : lo->children) {
            if (ob->name != "object") continue;
            ObjSpec o;
            auto t = [&](const char *k) {
                const plbfs::XmlNode *e = ob->child(k);
                return e ? e->text : std::string();
            };
            o.name = t("name");
            o.type = t("type");
            const double nn = std::atof(t("num_nodes").c_str());
            o.num_nodes = nn >= 1 && nn <= kMaxObjectNodes ? (int)nn : -1;
            o.radius = std::atof(t("radius").c_str());
            o.depth = std::atof(t("depth").c_str());
            o.stiffness = std::atof(t("stiffness").c_str());
            o.cx = std::atof(t("centre_x").c_str());
            o.cy = std::atof(t("centre_y").c_str());
            o.cz = std::atof(t("centre_z").c_str());
            o.mesh_file = t("mesh_file");
            const std::string k = t("delta");
            o.kernel = (k == "peskin4" || (k.empty() && o.type == "rigid_sphere")) ? PLBFS_DELTA_PESKIN4
                                                                                    : PLBFS_DELTA_HAT2;
            c->objects.push_back(o);
        }
    }
    // global_variables::initialise (global_variables.cpp:20-56)
    plbfs_params &p = c->params;
    p.visc = U * c->n("geometry/ref_length") / c->n("global_variables/reynolds_no");
    p.gamma = c->n("global_variables/pre_condition_gamma", 1.0);
    if (c->s("global_variables/preconditioning") == "true" && std::fabs(p.gamma - 1) < 0.00001)
        p.gamma = U * 2.0 * std::sqrt(3);
    p.fneq_min = c->n("global_variables/fneq_min", 0.0);
    p.cfl = c->n("global_variables/time_marching_step");
    p.max_velocity = U;
    p.ref_rho = c->n("initial_conditions/average_rho");
    p.force = -c->n("initial_conditions/pressure_gradient", 0.0);
    p.tolerance = c->n("global_variables/tolerance");
    const std::string gc = c->s("global_variables/gradient_calc", "LS");
    p.gradient_type = gc == "LS" ? PLBFS_GRAD_LS : (gc == "hybrid" ? PLBFS_GRAD_HYBRID : PLBFS_GRAD_GG);
    const std::string ts = c->s("global_variables/time_stepping", "local");
    p.time_stepping = ts == "local" ? PLBFS_DT_LOCAL
                      : ts == "min" ? PLBFS_DT_MIN
                      : ts == "talts" ? PLBFS_DT_TALTS
                                      : PLBFS_DT_CONSTANT;
    p.precision = c->s("global_variables/precision") == "fp32" ? PLBFS_FP32 : PLBFS_FP64;
    p.check_interval = (int32_t)c->n("global_variables/check_interval", 10);
    p.lsq_exact_det = c->s("global_variables/lsq_determinant") == "exact" ? 1 : 0;
    // domain.X, domain.Y after domain_geometry::scale_geometries (domain_geometry.cpp:21-24)
    p.domain_x = c->n("geometry/x", 1.0) * c->n("global_variables/scale", 1.0);
    p.domain_y = c->n("geometry/y", 1.0) * c->n("global_variables/scale", 1.0);
    if (!std::isfinite(p.visc))
        return plbfs_fail(PLBFS_ERR_INPUT, "%s: mach_no * ref_length / reynolds_no is not finite", xml_path);
    *out = c.release();
    return PLBFS_OK;
}

extern "C" int plbfs_case_load(const char *xml_path, plbfs_case **out) {
    if (!xml_path || !out) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    *out = nullptr;
    return plbfs_guarded("plbfs_case_load", [&] { return case_load(xml_path, out); });
}

extern "C" void plbfs_case_free(plbfs_case *c) { delete c; }

extern "C" int plbfs_case_params(const plbfs_case *c, plbfs_params *p) {
    if (!c || !p) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    *p = c->params;
    return PLBFS_OK;
}

extern "C" int plbfs_case_string(const plbfs_case *c, const char *key, char *buf, int32_t len) {
    if (!c || !key || !buf || len <= 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    auto it = c->str.find(key);
    if (it == c->str.end()) return plbfs_fail(PLBFS_ERR_INPUT, "no element %s", key);
    std::snprintf(buf, (size_t)len, "%s", it->second.c_str());
    return PLBFS_OK;
}

extern "C" double plbfs_case_number(const plbfs_case *c, const char *key) {
    if (!c || !key) return NAN;
    auto it = c->num.find(key);
    return it == c->num.end() ? NAN : it->second;
}

extern "C" int plbfs_case_patches(const plbfs_case *c, const plbfs_mesh *m, plbfs_patch *out, int32_t n_patches) {
    if (!c || !m || !out) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    int64_t info[6];
    plbfs_mesh_info(m, info);
    if (n_patches != info[4]) return plbfs_fail(PLBFS_ERR_INPUT, "mesh has %lld patches", (long long)info[4]);
    char buf[256];
    for (int q = 0; q < n_patches; q++) {
        plbfs_mesh_patch_type(m, q, buf, sizeof buf);
        const BcSpec *b = nullptr;
        for (auto &x : c->bcs)
            if (x.name == buf) {
                b = &x;
                break;
            }
        if (!b) return plbfs_fail(PLBFS_ERR_BCTYPE, "no <bc> with name '%s' for mesh patch %d", buf, q);
        out[q].rho_type = b->rho_type;
        out[q].vel_type = b->vel_type;
        out[q].rho = b->rho;
        out[q].u = b->u;
        out[q].v = b->v;
        out[q].w = b->w;
        out[q].is_empty = std::strcmp(buf, "empty") == 0;
        // one translation for every periodic patch (src/preprocessor.cpp:299-301)
        out[q].per_translate[0] = c->n("uns_boundary_conditions/periodic_translation/x", 0.0);
        out[q].per_translate[1] = c->n("uns_boundary_conditions/periodic_translation/y", 0.0);
        out[q].per_translate[2] = c->n("uns_boundary_conditions/periodic_translation/z", 0.0);
    }
    return PLBFS_OK;
}

extern "C" int plbfs_case_initial_state(const plbfs_case *c, const plbfs_mesh *m, double *rho, double *u, double *v,
                                        double *w) {
    if (!c || !m || !rho || !u || !v || !w) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    int64_t info[6];
    plbfs_mesh_info(m, info);
    plbfs_mesh_view view;
    plbfs_mesh_view_get(m, &view);
    const int64_t T = info[0] + info[5];
    const double factor = c->n("global_variables/mach_no") / c->n("global_variables/scale", 1.0);
    const double ox = c->n("initial_conditions/origin_location/x"), oy = c->n("initial_conditions/origin_location/y"),
                 oz = c->n("initial_conditions/origin_location/z");
    const double rmx = c->n("initial_conditions/rho/origin_magnitude/x"),
                 rmy = c->n("initial_conditions/rho/origin_magnitude/y"),
                 rmz = c->n("initial_conditions/rho/origin_magnitude/z");
    const double rgx = c->n("initial_conditions/rho/gradient/x"), rgy = c->n("initial_conditions/rho/gradient/y"),
                 rgz = c->n("initial_conditions/rho/gradient/z");
    const double vmx = c->n("initial_conditions/vel/origin_magnitude/x") * factor,
                 vmy = c->n("initial_conditions/vel/origin_magnitude/y") * factor,
                 vmz = c->n("initial_conditions/vel/origin_magnitude/z") * factor;
    const double vgx = c->n("initial_conditions/vel/gradient/x") * factor,
                 vgy = c->n("initial_conditions/vel/gradient/y") * factor,
                 vgz = c->n("initial_conditions/vel/gradient/z") * factor;
    for (int64_t t = 0; t < T; t++) {
        const double dx = view.cell_centroid[3 * t] - ox, dy = view.cell_centroid[3 * t + 1] - oy,
                     dz = view.cell_centroid[3 * t + 2] - oz;
        const double tx = rmx + rgx * dx, ty = rmy + rgy * dy, tz = rmz + rgz * dz;
        rho[t] = std::sqrt(std::pow(tx, 2) + std::pow(ty, 2) + std::pow(tz, 2));
        const double ux = vmx + vgx * dx, uy = vmy + vgy * dy, uz = vmz + vgz * dz;
        u[t] = ux + uy + uz;
        v[t] = 0.0;
        w[t] = 0.0;
    }
    return PLBFS_OK;
}

extern "C" int plbfs_case_n_objects(const plbfs_case *c) { return c ? (int)c->objects.size() : 0; }

static int object_nodes(const plbfs_case *c, int32_t k, double *x, int32_t max_nodes, int32_t *n_out,
                        double *stiffness, double *node_area, int32_t *kernel) {
    const ObjSpec &o = c->objects[k];
    if ((o.type == "rigid_cylinder" || o.type == "rigid_sphere") && o.num_nodes < 1)
        return plbfs_fail(PLBFS_ERR_INPUT, "object '%s': num_nodes must be 1..%d", o.name.c_str(), kMaxObjectNodes);
    const double PI = 3.14159265358979323846;
    std::vector<double> pts;
    double area = 0;
    if (o.type == "rigid_cylinder") {
        // lagrangian_object.cpp:79-113; dA as in update_node_forces (immersed_boundary_method.cu:168)
        const int depth_nodes = (int)std::sqrt((double)o.num_nodes), radial = depth_nodes;
        const double delta_z = o.depth / depth_nodes;
        for (int j = 0; j < depth_nodes; j++)
            for (int i = 0; i < radial; i++) {
                pts.push_back(o.cx + o.radius * std::sin(2 * PI * i / radial));
                pts.push_back(o.cy + o.radius * std::cos(2 * PI * i / radial));
                pts.push_back(o.cz + o.depth / 2 - j * o.depth / depth_nodes - delta_z / 2);
            }
        area = 2 * PI * o.radius / o.num_nodes * o.depth;
    } else if (o.type == "rigid_sphere") {
        const int n = o.num_nodes;
        const double ga = PI * (3.0 - std::sqrt(5.0));
        for (int i = 0; i < n; i++) {
            const double z = 1.0 - (2.0 * i + 1.0) / n;
            const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
            const double ph = ga * i;
            pts.push_back(o.cx + o.radius * r * std::cos(ph));
            pts.push_back(o.cy + o.radius * r * std::sin(ph));
            pts.push_back(o.cz + o.radius * z);
        }
        area = 4 * PI * o.radius * o.radius / n;
    } else if (o.type == "spring_network") {
        // lagrangian_object::mesh_read (lagrangian_object.cpp:116-213): node,
        // spring and triangle counts, node coordinates (used as given), then
        // triangle and spring connectivity and per-node neighbour lists, all
        // 1-based.  The solver moves every node on its own anchor spring like
        // the rigid objects; the area per node is the reference's cylinder
        // expression with the file's node count (immersed_boundary_method.cu:168).
        std::ifstream in(o.mesh_file);
        if (!in) return plbfs_fail(PLBFS_ERR_INPUT, "object '%s': cannot open mesh_file '%s'", o.name.c_str(),
                                   o.mesh_file.c_str());
        long nn = -1, ns = -1, nt = -1;
        in >> nn >> ns >> nt;
        if (!in || nn <= 0 || ns < 0 || nt < 0 || nn > kMaxObjectNodes || ns > 16L * kMaxObjectNodes ||
            nt > 16L * kMaxObjectNodes)
            return plbfs_fail(PLBFS_ERR_INPUT, "object '%s': bad counts in '%s'", o.name.c_str(), o.mesh_file.c_str());
        pts.resize((size_t)3 * nn);
        for (long i = 0; i < 3 * nn; i++) in >> pts[i];
        long idx = 0;
        for (long i = 0; i < 3 * nt + 4 * ns && in; i++) {
            in >> idx;
            if (idx < 1 || idx > nn)
                return plbfs_fail(PLBFS_ERR_INPUT, "object '%s': connectivity index %ld out of 1..%ld",
                                  o.name.c_str(), idx, nn);
        }
        for (long i = 0; i < nn && in; i++) {
            long deg = 0;
            in >> deg;
            for (long j = 0; j < deg && in; j++) in >> idx;
        }
        if (!in) return plbfs_fail(PLBFS_ERR_INPUT, "object '%s': '%s' ends early", o.name.c_str(), o.mesh_file.c_str());
        area = 2 * PI * o.radius / nn * o.depth;
    } else {
        return plbfs_fail(PLBFS_ERR_INPUT, "object '%s': unsupported type '%s'", o.name.c_str(), o.type.c_str());
    }
    const int n = (int)pts.size() / 3;
    if (n_out) *n_out = n;
    if (x) {
        if (n > max_nodes) return plbfs_fail(PLBFS_ERR_INPUT, "object has %d nodes > %d", n, max_nodes);
        std::memcpy(x, pts.data(), pts.size() * sizeof(double));
    }
    if (stiffness) *stiffness = o.stiffness;
    if (node_area) *node_area = area;
    if (kernel) *kernel = o.kernel;
    return PLBFS_OK;
}

extern "C" int plbfs_case_object_nodes(const plbfs_case *c, int32_t k, double *x, int32_t max_nodes, int32_t *n_out,
                                       double *stiffness, double *node_area, int32_t *kernel) {
    if (!c || k < 0 || k >= (int)c->objects.size()) return plbfs_fail(PLBFS_ERR_INPUT, "bad object index");
    return plbfs_guarded("plbfs_case_object_nodes",
                         [&] { return object_nodes(c, k, x, max_nodes, n_out, stiffness, node_area, kernel); });
}

// ---- output ----------------------------------------------------------------

extern "C" int plbfs_write_fields(const char *dir, int32_t n, const double *rho, const double *u, const double *v,
                                  const double *w) {
    if (!dir || n < 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    mkdirs(dir);
    const char *names[4] = {"rho", "u", "v", "w"};
    const double *src[4] = {rho, u, v, w};
    for (int q = 0; q < 4; q++) {
        std::string path = std::string(dir) + "/" + names[q] + ".txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return plbfs_fail(PLBFS_ERR_INPUT, "cannot write %s", path.c_str());
        // the reference's "i ,value" lines (Solution.cpp:238-244), at full precision
        for (int i = 0; i < n; i++) std::fprintf(f, "%d ,%.17g\n", i, src[q][i]);
        std::fclose(f);
    }
    return PLBFS_OK;
}

extern "C" int plbfs_read_fields(const char *dir, int32_t n, double *rho, double *u, double *v, double *w) {
    if (!dir || n < 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    const char *names[4] = {"rho", "u", "v", "w"};
    double *dst[4] = {rho, u, v, w};
    for (int q = 0; q < 4; q++) {
        std::string path = std::string(dir) + "/" + names[q] + ".txt", text;
        if (!read_text(path, text)) return plbfs_fail(PLBFS_ERR_INPUT, "cannot read %s", path.c_str());
        int i = 0;
        size_t pos = 0;
        while (pos < text.size() && i < n) {
            size_t e = text.find('\n', pos);
            if (e == std::string::npos) e = text.size();
            size_t comma = text.find(',', pos);
            if (comma != std::string::npos && comma < e) dst[q][i++] = std::strtod(text.c_str() + comma + 1, nullptr);
            pos = e + 1;
        }
        if (i != n) return plbfs_fail(PLBFS_ERR_INPUT, "%s has %d values, expected %d", path.c_str(), i, n);
    }
    return PLBFS_OK;
}

// Tecplot ASCII output without TecIO.  Like the reference, the grid and the
// solution are separate files: a grid file with the node coordinates and the
// polyhedral face connectivity (unstructured_mesh.cpp:2073-2200; left element
// = owner + 1, right = neighbour + 1 or 0), and per output time a solution
// file with the cell-centred variables of tecplot_output.cpp:273-466.

static void write_block(FILE *f, const std::vector<double> &a) {
    const size_t n = a.size();
    for (size_t i = 0; i < n; i++) std::fprintf(f, "%.17g%c", a[i], (i % 6 == 5 || i == n - 1) ? '\n' : ' ');
}

static void polyhedral_zone_header(FILE *f, int P, int N, int F, size_t nfv, double time) {
    std::fprintf(f,
                 "ZONE T=\"%d\" ZONETYPE=FEPOLYHEDRON NODES=%d ELEMENTS=%d FACES=%d TOTALNUMFACENODES=%zu "
                 "NUMCONNECTEDBOUNDARYFACES=0 TOTALNUMBOUNDARYCONNECTIONS=0 DATAPACKING=BLOCK",
                 N, P, N, F, nfv);
    std::fprintf(f, " SOLUTIONTIME=%.17g", time);
}

extern "C" int plbfs_write_tecplot_grid(const plbfs_mesh *m, const char *path) {
    if (!m || !path) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    int64_t info[6];
    plbfs_mesh_info(m, info);
    const int N = (int)info[0], F = (int)info[1], NI = (int)info[2], P = (int)info[3];
    std::vector<float> px(P), py(P), pz(P);
    std::vector<int32_t> owner(F), nb(NI > 0 ? NI : 1), fo(F + 1), fv((size_t)plbfs_mesh_n_face_verts(m));
    plbfs_mesh_raw(m, px.data(), py.data(), pz.data(), owner.data(), nb.data(), fo.data(), fv.data());
    FILE *f = std::fopen(path, "w");
    if (!f) return plbfs_fail(PLBFS_ERR_INPUT, "cannot write %s", path);
    std::fprintf(f, "TITLE = \"LBFS grid\"\nFILETYPE = GRID\nVARIABLES = \"X\" \"Y\" \"Z\"\n");
    polyhedral_zone_header(f, P, N, F, fv.size(), 0.0);
    std::fprintf(f, "\n");
    write_block(f, std::vector<double>(px.begin(), px.end()));
    write_block(f, std::vector<double>(py.begin(), py.end()));
    write_block(f, std::vector<double>(pz.begin(), pz.end()));
    for (int k = 0; k < F; k++) std::fprintf(f, "%d%c", fo[k + 1] - fo[k], (k % 16 == 15 || k == F - 1) ? '\n' : ' ');
    for (int k = 0; k < F; k++) {
        for (int q = fo[k]; q < fo[k + 1]; q++) std::fprintf(f, "%d ", fv[q] + 1);
        std::fputc('\n', f);
    }
    for (int k = 0; k < F; k++) std::fprintf(f, "%d%c", owner[k] + 1, (k % 16 == 15 || k == F - 1) ? '\n' : ' ');
    for (int k = 0; k < F; k++)
        std::fprintf(f, "%d%c", k < NI ? nb[k] + 1 : 0, (k % 16 == 15 || k == F - 1) ? '\n' : ' ');
    std::fclose(f);
    return PLBFS_OK;
}

extern "C" int plbfs_write_tecplot(const plbfs_mesh *m, const char *path, const double *rho, const double *u,
                                   const double *v, const double *w, const double *res4, const double *dt_local,
                                   double max_velocity, double visc, double gamma, double time) {
    if (!m || !path || !rho || !u || !v || !w) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    if (!(max_velocity != 0) || !(gamma != 0)) return plbfs_fail(PLBFS_ERR_INPUT, "max_velocity and gamma");
    int64_t info[6];
    plbfs_mesh_info(m, info);
    const int N = (int)info[0], F = (int)info[1], NI = (int)info[2], P = (int)info[3];
    const size_t nfv = (size_t)plbfs_mesh_n_face_verts(m);
    plbfs_mesh_view mv;
    plbfs_mesh_view_get(m, &mv);
    const int32_t *bp = plbfs_mesh_bface_patch(m);
    // fneq: tau' = 3 nu / (dt_face gamma) of the last face a cell owns in face
    // order, skipping the "empty" faces past the first boundary face
    // (Solver.cpp:308,680: local_fneq[owner] is overwritten face by face); 0
    // where a cell owns no such face (the reference leaves it unset)
    std::vector<double> fneq(N, 0.0);
    {
        char buf[256];
        std::vector<int> empty((size_t)info[4], 0);
        for (int q = 0; q < (int)info[4]; q++) {
            plbfs_mesh_patch_type(m, q, buf, sizeof buf);
            empty[q] = std::strcmp(buf, "empty") == 0;
        }
        for (int fc = 0; fc < F; fc++) {
            if (fc > NI && empty[bp[fc - NI]]) continue;
            fneq[mv.owner[fc]] = visc * 3 / mv.face_dt[fc] / gamma;
        }
    }
    FILE *f = std::fopen(path, "w");
    if (!f) return plbfs_fail(PLBFS_ERR_INPUT, "cannot write %s", path);
    std::fprintf(f,
                 "TITLE = \"LBFS\"\nFILETYPE = SOLUTION\nVARIABLES = \"p\" \"u\" \"v\" \"w\" \"x\" \"y\" \"z\" "
                 "\"vort\" \"stf\" \"res_rho\" \"res_u\" \"res_v\" \"res_w\" \"dt\" \"fneq\"\n");
    polyhedral_zone_header(f, P, N, F, nfv, time);
    std::fprintf(f, " VARLOCATION=([1-15]=CELLCENTERED)\n");
    std::vector<double> col(N);
    auto put = [&](auto get) {
        for (int i = 0; i < N; i++) col[i] = get(i);
        write_block(f, col);
    };
    // p is the density, velocities over max_velocity, x y z the centroids
    put([&](int i) { return rho[i]; });
    put([&](int i) { return u[i] / max_velocity; });
    put([&](int i) { return v[i] / max_velocity; });
    put([&](int i) { return w[i] / max_velocity; });
    for (int a = 0; a < 3; a++) put([&](int i) { return mv.cell_centroid[3 * (size_t)i + a]; });
    // vorticity and stream function stay at their initial zeros: the
    // reference's calc_vorticity call is commented out (Solver.cpp:537,562)
    put([&](int) { return 0.0; });
    put([&](int) { return 0.0; });
    for (int q = 0; q < 4; q++) put([&](int i) { return res4 ? res4[(size_t)4 * i + q] : 0.0; });
    put([&](int i) { return dt_local ? dt_local[i] : 0.0; });
    put([&](int i) { return fneq[i]; });
    std::fclose(f);
    return PLBFS_OK;
}

extern "C" int plbfs_write_tecplot_nodes(const char *path, const char *name, int32_t n, const double *x3,
                                         const double *vel3, const double *force3, double X, double Y, double Z,
                                         int32_t radial_nodes, int32_t depth_nodes, double time) {
    if (!path || !name || n <= 0 || !x3 || !(X != 0 && Y != 0 && Z != 0))
        return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    FILE *f = std::fopen(path, "w");
    if (!f) return plbfs_fail(PLBFS_ERR_INPUT, "cannot write %s", path);
    std::fprintf(f, "TITLE = \"%s\"\nVARIABLES = \"nodx\" \"nody\" \"nodz\" \"vel_x\" \"vel_y\" \"vel_z\" "
                 "\"force_x\" \"force_y\" \"force_z\"\n", name);
    const bool poly = radial_nodes >= 2 && depth_nodes >= 2 && (int64_t)radial_nodes * depth_nodes == n;
    const int R = radial_nodes, D = depth_nodes;
    if (poly) {
        // the cylinder surface as quads between consecutive rings, edges as
        // faces (tecplot_output.cpp:784-1062)
        std::fprintf(f,
                     "ZONE T=\"%d\" ZONETYPE=FEPOLYGON NODES=%d ELEMENTS=%d FACES=%d NUMCONNECTEDBOUNDARYFACES=0 "
                     "TOTALNUMBOUNDARYCONNECTIONS=0 DATAPACKING=BLOCK SOLUTIONTIME=%.17g\n",
                     R * (D - 1), n, R * (D - 1), R * D + R * (D - 1), time);
    } else {
        std::fprintf(f, "ZONE T=\"%s\" I=%d DATAPACKING=BLOCK SOLUTIONTIME=%.17g\n", name, n, time);
    }
    // positions over the domain lengths, as the reference writes them
    const double L3[3] = {X, Y, Z};
    std::vector<double> col(n);
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < n; k++) col[k] = x3[3 * (size_t)k + a] / L3[a];
        write_block(f, col);
    }
    for (const double *src : {vel3, force3})
        for (int a = 0; a < 3; a++) {
            for (int k = 0; k < n; k++) col[k] = src ? src[3 * (size_t)k + a] : 0.0;
            write_block(f, col);
        }
    if (poly) {
        // face nodes (1-based): ring edges, closing edge of each ring, then
        // the axial edges between rings
        for (int j = 0; j < D; j++) {
            for (int k = 1; k < R; k++) std::fprintf(f, "%d %d\n", R * j + k, R * j + k + 1);
            std::fprintf(f, "%d %d\n", R * (j + 1), 1 + R * j);
        }
        for (int j = 0; j < D - 1; j++)
            for (int k = 0; k < R; k++) std::fprintf(f, "%d %d\n", R * j + k + 1, R * (j + 1) + k + 1);
        // left and right elements of the faces in the same order (0: none)
        std::vector<int> left, right;
        for (int ii = 0; ii < R; ii++) left.push_back(0), right.push_back(ii + 1);
        for (int j = 1; j < D - 1; j++)
            for (int ii = 0; ii < R; ii++) left.push_back(R * (j - 1) + ii + 1), right.push_back(R * j + ii + 1);
        for (int ii = 0; ii < R; ii++) left.push_back(R * (D - 2) + ii + 1), right.push_back(0);
        for (int j = 0; j < D - 1; j++) {
            left.push_back(R * j + 1), right.push_back(R * (j + 1));
            for (int ii = 1; ii < R; ii++) left.push_back(R * j + ii + 1), right.push_back(R * j + ii);
        }
        for (int e : left) std::fprintf(f, "%d\n", e);
        for (int e : right) std::fprintf(f, "%d\n", e);
    }
    std::fclose(f);
    return PLBFS_OK;
}

// ---- program ---------------------------------------------------------------

static int program_run(const char *xml_path, plbfs_result *result) {
    const auto t_start = std::chrono::steady_clock::now();
    plbfs_case *c = nullptr;
    int rc = plbfs_case_load(xml_path, &c);
    if (rc) return rc;
    std::unique_ptr<plbfs_case, void (*)(plbfs_case *)> cg(c, plbfs_case_free);
    const plbfs_params &p = c->params;
    if (c->s("global_variables/import_format", "openfoam") != "openfoam")
        return plbfs_fail(PLBFS_ERR_INPUT, "only import_format openfoam is supported");
    plbfs_mesh *m = nullptr;
    rc = plbfs_mesh_read_openfoam(c->s("global_variables/import_file").c_str(), p.visc, p.gamma, p.fneq_min, &m);
    if (rc) return rc;
    std::unique_ptr<plbfs_mesh, void (*)(plbfs_mesh *)> mg(m, plbfs_mesh_free);
    int64_t info[6];
    plbfs_mesh_info(m, info);
    const int N = (int)info[0], T = (int)(info[0] + info[5]), NP = (int)info[4];
    std::vector<plbfs_patch> patches(NP);
    if ((rc = plbfs_case_patches(c, m, patches.data(), NP))) return rc;
    std::vector<double> rho(T), u(T), v(T), w(T);
    std::string outdir = c->s("global_variables/output_directory", ".");
    mkdirs(outdir);
    // initial conditions over all cells; a restart replaces the interior ones
    if ((rc = plbfs_case_initial_state(c, m, rho.data(), u.data(), v.data(), w.data()))) return rc;
    if (c->s("global_variables/restart_analysis") == "true") {
        if ((rc = plbfs_read_fields(outdir.c_str(), N, rho.data(), u.data(), v.data(), w.data()))) return rc;
    }
    plbfs_mesh_view view;
    plbfs_mesh_view_get(m, &view);
    plbfs_ctx *ctx = nullptr;
    if ((rc = plbfs_create(&view, &p, patches.data(), NP, plbfs_mesh_bface_patch(m), &ctx))) return rc;
    std::unique_ptr<plbfs_ctx, int (*)(plbfs_ctx *)> xg(ctx, plbfs_destroy);
    if ((rc = plbfs_set_state(ctx, rho.data(), u.data(), v.data(), w.data()))) return rc;
    // the initial state over the ghosts (soln's entries past n_cells), which
    // some boundary types keep (src/Solution.cpp:759-833)
    if ((rc = plbfs_set_ghost_state(ctx, rho.data() + N, u.data() + N, v.data() + N, w.data() + N))) return rc;
    std::vector<int32_t> obj_nodes;  // node count of each object, in node order
    if (!c->objects.empty()) {
        std::vector<double> all;
        double stiff = 0, area = 0;
        int kernel = 0;
        for (int k = 0; k < (int)c->objects.size(); k++) {
            int32_t n = 0, ker = 0;
            double s = 0, a = 0;
            if ((rc = plbfs_case_object_nodes(c, k, nullptr, 0, &n, &s, &a, &ker))) return rc;
            std::vector<double> x((size_t)3 * n);
            if ((rc = plbfs_case_object_nodes(c, k, x.data(), n, &n, &s, &a, &ker))) return rc;
            all.insert(all.end(), x.begin(), x.end());
            obj_nodes.push_back(n);
            stiff = s;  // the reference applies object 0's parameters to all nodes (Solver_gpu.cu:708-709)
            if (k == 0) {
                area = a;
                kernel = ker;
            }
        }
        plbfs_ibm_nodes nodes{};
        nodes.n_nodes = (int32_t)(all.size() / 3);
        nodes.x_ref = all.data();
        nodes.x = all.data();
        nodes.stiffness = stiff;
        nodes.node_area = area;
        nodes.dt = p.cfl;
        nodes.h = c->n("global_variables/delta_width", 2.0 * c->n("geometry/streaming_dt", 0.5));
        nodes.kernel = kernel;
        if ((rc = plbfs_set_ibm(ctx, &nodes))) return rc;
    }
    const double length = std::ceil(c->n("global_variables/simulation_length") * c->n("global_variables/scale", 1.0));
    if (!(length >= 0 && length <= INT32_MAX))
        return plbfs_fail(PLBFS_ERR_INPUT, "simulation_length * scale = %g steps is out of range", length);
    const int steps = (int)length;
    plbfs_result r{};
    rc = plbfs_run(ctx, steps, &r);
    if (rc != PLBFS_OK && rc != PLBFS_ERR_NAN && rc != PLBFS_ERR_RHO && rc != PLBFS_ERR_IBM) return rc;
    const int run_rc = rc;
    // residual history (error.txt columns of Solver.cpp:506-509)
    {
        std::vector<double> h((size_t)4 * (r.steps > 0 ? r.steps : 1));
        plbfs_get_residual_history(ctx, h.data(), r.steps);
        FILE *f = std::fopen((outdir + "/error.txt").c_str(), "w");
        if (f) {
            for (int t = 0; t < r.steps; t++) {
                const double *q = &h[(size_t)4 * t];
                std::fprintf(f, "%d, %.17g, %.17g, %.17g, %.17g, %.17g\n", t,
                             std::fmax(std::fmax(q[3], q[1]), q[2]), q[0], q[1], q[2], q[3]);
            }
            std::fclose(f);
        }
    }
    std::vector<double> sr(N), su(N), sv(N), sw(N);
    if ((rc = plbfs_get_state(ctx, sr.data(), su.data(), sv.data(), sw.data()))) return rc;
    if ((rc = plbfs_write_fields(outdir.c_str(), N, sr.data(), su.data(), sv.data(), sw.data()))) return rc;
    // Tecplot: the grid once, the solution with the residual of the final
    // state and its local time steps, and each immersed object's nodes
    // (tecplot_output.cpp:273-466,784-1062; Solver_gpu.cu:911-912)
    mkdirs(outdir + "/plt");
    if ((rc = plbfs_write_tecplot_grid(m, (outdir + "/plt/grid.dat").c_str()))) return rc;
    {
        std::vector<double> res((size_t)4 * N), dtl(N);
        if ((rc = plbfs_eval_residual(ctx, nullptr, res.data())) || (rc = plbfs_get_dt(ctx, dtl.data()))) return rc;
        const std::string plt = outdir + "/plt/" + std::to_string(r.total_steps) + ".dat";
        if ((rc = plbfs_write_tecplot(m, plt.c_str(), sr.data(), su.data(), sv.data(), sw.data(), res.data(),
                                      dtl.data(), p.max_velocity, p.visc, p.gamma, r.total_steps)))
            return rc;
    }
    if (!obj_nodes.empty()) {
        int32_t ntot = 0;
        for (int32_t n : obj_nodes) ntot += n;
        std::vector<double> x((size_t)3 * ntot), vel((size_t)3 * ntot), force((size_t)3 * ntot);
        if ((rc = plbfs_get_ibm(ctx, x.data(), vel.data(), force.data()))) return rc;
        mkdirs(outdir + "/plt/object");
        const double X = c->n("geometry/x", 1.0) * c->n("global_variables/scale", 1.0);
        const double Y = c->n("geometry/y", 1.0) * c->n("global_variables/scale", 1.0);
        const double Z = c->n("geometry/z", 1.0);
        size_t k0 = 0;
        for (size_t k = 0; k < obj_nodes.size(); k++) {
            const ObjSpec &o = c->objects[k];
            const int32_t n = obj_nodes[k];
            // rings of a rigid cylinder (lagrangian_object.cpp:84-85); other objects are points
            const int32_t rings = o.type == "rigid_cylinder" ? (int32_t)std::sqrt((double)o.num_nodes) : 0;
            const std::string path = outdir + "/plt/object/" + o.name + std::to_string(r.total_steps) + ".dat";
            if ((rc = plbfs_write_tecplot_nodes(path.c_str(), o.name.c_str(), n, x.data() + 3 * k0,
                                                vel.data() + 3 * k0, force.data() + 3 * k0, X, Y, Z, rings, rings,
                                                r.total_steps)))
                return rc;
            k0 += (size_t)n;
        }
    }
    // forces: every "wall" patch with the reference's coefficients
    // (src/post_processing.cpp:170-171), and the immersed objects
    if (FILE *f = std::fopen((outdir + "/drag.txt").c_str(), "w")) {
        const double rho0 = p.ref_rho, U = p.max_velocity;
        const double X = c->n("geometry/x", 1.0), Y = c->n("geometry/y", 1.0), Z = c->n("geometry/z", 1.0);
        std::fprintf(f, "patch , drag , lift , drag_coef , lift_coef\n");
        char buf[256];
        for (int q = 0; q < NP; q++) {
            plbfs_mesh_patch_type(m, q, buf, sizeof buf);
            if (std::strcmp(buf, "wall") != 0) continue;
            double dl[2];
            if ((rc = plbfs_wall_forces(ctx, &view, plbfs_mesh_bface_patch(m), q, dl))) {
                std::fclose(f);
                return rc;
            }
            std::fprintf(f, "%d , %.17g , %.17g , %.17g , %.17g\n", q, dl[0], dl[1],
                         2 * dl[0] / rho0 / std::pow(U, 2) / Y / Z, 2 * dl[1] / rho0 / std::pow(U, 2) / X / Z);
        }
        if (!c->objects.empty()) {
            double fb[3];
            if ((rc = plbfs_ibm_body_force(ctx, fb)) == PLBFS_OK)
                std::fprintf(f, "immersed , %.17g , %.17g , %.17g\n", fb[0], fb[1], fb[2]);
        }
        std::fclose(f);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (FILE *f = std::fopen((outdir + "/globals.txt").c_str(), "w")) {
        // program.cpp:192-234 fields that the solver determines
        std::fprintf(f, "Runtime:%gs\nSolve:%gms\nSteps:%d\nConverged:%d\n", secs, r.solve_ms, r.total_steps,
                     r.converged);
        std::fprintf(f, "Tau:%.17g\nMach:%.17g\nReynolds: %.17g\n",
                     3 * p.visc / c->n("geometry/streaming_dt", 0.5) + 0.5, p.max_velocity * std::sqrt(3),
                     c->n("global_variables/reynolds_no"));
        std::fclose(f);
    }
    if (result) *result = r;
    return run_rc;
}

extern "C" int plbfs_program_run(const char *xml_path, plbfs_result *result) {
    return plbfs_guarded("plbfs_program_run", [&] { return program_run(xml_path, result); });
}
