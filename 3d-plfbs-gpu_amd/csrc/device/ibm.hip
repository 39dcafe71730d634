// Immersed boundary method on unstructured meshes (gfx950).
//
// Follows the reference GPU IBM's per-step sequence (src/Solver_gpu.cu:687-718):
// interpolate node velocities from the step-start state
// (src/immersed_boundary_method.cu:13-90), move the nodes X = X0 + U dt
// (:244-295), spring force F = -k dA (X - X_ref) (:161-190) and spread the
// force onto the cells (:299-375); the cell force enters R every stage
// (src/Solver_gpu.cu:766-768).  The reference indexes a structured grid and
// uses a 2-point hat; here every node has a precomputed list of candidate
// cells (centroids within the kernel support plus a drift margin of the
// anchor), sorted by cell, so that
//   interpolation  = one 64-wide wavefront per node, lanes stride over the
//                    node's pairs, shuffle butterfly reduction (fixed order);
//   spreading      = one thread per affected cell summing over its pairs in
//                    node order (pairs regrouped by cell on the host).
// Weights are normalised per node, W_cn = delta V_c / sum delta V, which is
// the reference's formula on its uniform unit grid.  No atomics.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"
#include "ibm_kernels.hpp"

using namespace plbfs;

namespace {

IbmArgs make_args(plbfs_ctx *c, bool rebuild = false) {
    const IbmState &s = c->ibm;
    IbmArgs a;
    a.xc = s.xc;
    a.rebuild = rebuild ? 1 : 0;
    a.n = s.n;
    a.N = c->L.N;
    a.B = 1;
    a.node_kdA = nullptr;
    a.node_case = nullptr;
    a.cp_case = nullptr;
    a.case_done = nullptr;
    a.case_status = nullptr;
    a.kernel = s.kernel;
    a.ih = 1.0 / s.h;
    a.dt = s.dt;
    a.k_dA = s.stiffness * s.node_area;
    a.margin = s.margin;
    a.x = s.x;
    a.x0 = s.x0;
    a.xref = s.xref;
    a.vel = s.vel;
    a.force = s.force;
    a.norm = s.norm;
    a.np_off = s.np_off;
    a.np_cell = s.np_cell;
    a.cp_cell = s.cp_cell;
    a.cp_off = s.cp_off;
    a.cp_node = s.cp_node;
    a.cp_pair = s.cp_pair;
    a.pw = s.pw;
    a.n_affected = s.n_affected;
    a.cc = s.cell_c;
    a.vol = c->dm.vol;
    return a;
}

}  // namespace

int ibm_launch(plbfs_ctx *c, Ctrl *ctrl, bool timed, bool rebuild) {
    if (!c->ibm.on) return PLBFS_OK;
    const IbmArgs a = make_args(c, rebuild);
    TimedLaunch t(c, PLBFS_KSTAT_IBM, timed);
    const dim3 grid((a.n + kNodesPerBlock - 1) / kNodesPerBlock);
    if (c->f32)
        hipLaunchKernelGGL(k_ibm_nodes<float4>, grid, dim3(kIbmThreads), 0, c->stream, a, ctrl, (const float4 *)c->stA);
    else
        hipLaunchKernelGGL(k_ibm_nodes<double4>, grid, dim3(kIbmThreads), 0, c->stream, a, ctrl,
                           (const double4 *)c->stA);
    hipLaunchKernelGGL(k_ibm_spread, dim3((a.n_affected + kSpreadPerBlock - 1) / kSpreadPerBlock), dim3(kIbmThreads), 0, c->stream, a,
                       ctrl, c->fbody);
    HIP_TRY(hipGetLastError());
    return PLBFS_OK;
}

int build_ibm_pairs(const plbfs::Layout &L, const plbfs_ibm_nodes *nodes, IbmPairs &P) {
    const int n = nodes->n_nodes;
    if (n <= 0 || !nodes->x_ref || !nodes->x || !(nodes->h > 0) || nodes->node_area < 0 ||
        (nodes->kernel != PLBFS_DELTA_HAT2 && nodes->kernel != PLBFS_DELTA_PESKIN4))
        return plbfs_fail(PLBFS_ERR_INPUT, "bad immersed-boundary node set");
    const int N = L.N;
    const double h = nodes->h;
    const double support = (nodes->kernel == PLBFS_DELTA_HAT2 ? 1.0 : 2.0) * h;
    // drift margin around the position a node's list is built at: a node that
    // moves half of it gets a new list (plbfs_run); one that moves all of it
    // may miss cells and stops the run
    const double margin = 1.0 * h;
    const double reach = support + margin;
    P.margin = margin;
    const double *cc = L.centroid.data();
    const double *xs = nodes->x;  // lists around the nodes' current positions
    // bin the cells near the node cloud
    double lo[3], hi[3];
    for (int q = 0; q < 3; q++) {
        lo[q] = hi[q] = xs[q];
        for (int k = 0; k < n; k++) {
            lo[q] = std::min(lo[q], xs[3 * k + q]);
            hi[q] = std::max(hi[q], xs[3 * k + q]);
        }
        lo[q] -= reach;
        hi[q] += reach;
    }
    auto key3 = [&](int64_t bx, int64_t by, int64_t bz) { return (bx * 1048576 + by) * 1048576 + bz; };
    std::unordered_map<int64_t, std::vector<int32_t>> bins;
    for (int i = 0; i < N; i++) {
        const double x = cc[i], y = cc[(size_t)N + i], z = cc[2 * (size_t)N + i];
        if (x < lo[0] || x > hi[0] || y < lo[1] || y > hi[1] || z < lo[2] || z > hi[2]) continue;
        bins[key3((int64_t)std::floor((x - lo[0]) / reach), (int64_t)std::floor((y - lo[1]) / reach),
                  (int64_t)std::floor((z - lo[2]) / reach))]
            .push_back(i);
    }
    std::vector<int32_t> &np_off = P.np_off, &np_cell = P.np_cell;
    np_off.assign(n + 1, 0);
    np_cell.clear();
    std::vector<std::array<int32_t, 3>> cell_node;  // (cell, node, node-major pair)
    for (int k = 0; k < n; k++) {
        const double *X = xs + 3 * k;
        const int64_t bx = (int64_t)std::floor((X[0] - lo[0]) / reach), by = (int64_t)std::floor((X[1] - lo[1]) / reach),
                      bz = (int64_t)std::floor((X[2] - lo[2]) / reach);
        std::vector<int32_t> cand;
        for (int64_t dx = -1; dx <= 1; dx++)
            for (int64_t dy = -1; dy <= 1; dy++)
                for (int64_t dz = -1; dz <= 1; dz++) {
                    auto it = bins.find(key3(bx + dx, by + dy, bz + dz));
                    if (it == bins.end()) continue;
                    for (int32_t i : it->second) {
                        if (std::fabs(cc[i] - X[0]) < reach && std::fabs(cc[(size_t)N + i] - X[1]) < reach &&
                            std::fabs(cc[2 * (size_t)N + i] - X[2]) < reach)
                            cand.push_back(i);
                    }
                }
        std::sort(cand.begin(), cand.end());
        if (cand.empty()) return plbfs_fail(PLBFS_ERR_IBM, "node %d has no cell within its kernel support", k);
        for (int32_t i : cand) {
            cell_node.push_back({i, k, (int32_t)np_cell.size()});
            np_cell.push_back(i);
        }
        np_off[k + 1] = (int32_t)np_cell.size();
    }
    std::sort(cell_node.begin(), cell_node.end());
    std::vector<int32_t> &cp_cell = P.cp_cell, &cp_off = P.cp_off, &cp_node = P.cp_node, &cp_pair = P.cp_pair;
    cp_cell.clear();
    cp_node.clear();
    cp_pair.clear();
    cp_off.assign(1, 0);
    for (size_t p = 0; p < cell_node.size(); p++) {
        if (p == 0 || cell_node[p][0] != cell_node[p - 1][0]) {
            if (p > 0) cp_off.push_back((int32_t)p);
            cp_cell.push_back(cell_node[p][0]);
        }
        cp_node.push_back(cell_node[p][1]);
        cp_pair.push_back(cell_node[p][2]);
    }
    cp_off.push_back((int32_t)cell_node.size());
    return PLBFS_OK;
}

static int set_ibm(plbfs_ctx *c, const plbfs_ibm_nodes *nodes, int rebuilds) {
    IbmPairs P;
    int rc = build_ibm_pairs(c->L, nodes, P);
    if (rc) return rc;
    const int n = nodes->n_nodes;
    const int N = c->L.N;
    const double h = nodes->h;
    const double margin = P.margin;
    const std::vector<int32_t> &np_off = P.np_off, &np_cell = P.np_cell, &cp_cell = P.cp_cell, &cp_off = P.cp_off,
                               &cp_node = P.cp_node;
    // device arrays; the previous node set's arrays are freed once no queued
    // work can read them (the captured step graph that references them is
    // dropped below)
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->graph_ok = false;
    for (void *p : c->ibm_allocs) (void)hipFree(p);
    c->ibm_allocs.clear();
    c->dm.blk_force = nullptr;
    c->dm32.blk_force = nullptr;
    c->prm.use_body_force = 0;  // until the new node set is complete
    std::vector<void *> &M = c->ibm_allocs;
    IbmState &s = c->ibm;
    s = IbmState();
    s.n = n;
    s.kernel = nodes->kernel;
    s.h = h;
    s.dt = nodes->dt;
    s.stiffness = nodes->stiffness;
    s.node_area = nodes->node_area;
    s.margin = margin;
    s.n_affected = (int)cp_cell.size();
    s.n_pairs = (int64_t)np_cell.size();
    s.xref_host.assign(nodes->x_ref, nodes->x_ref + 3 * (size_t)n);
    s.rebuilds = rebuilds;
    std::vector<double> xr((size_t)3 * n), xx((size_t)3 * n), z((size_t)3 * n, 0.0), zn(n, 0.0);
    for (int k = 0; k < n; k++)
        for (int q = 0; q < 3; q++) {
            xr[(size_t)q * n + k] = nodes->x_ref[3 * k + q];
            xx[(size_t)q * n + k] = nodes->x[3 * k + q];
        }
    if ((rc = upload(M, &s.xref, xr)) || (rc = upload(M, &s.x, xx)) || (rc = upload(M, &s.x0, xx)) ||
        (rc = upload(M, &s.xc, xx)) ||
        (rc = upload(M, &s.vel, z)) || (rc = upload(M, &s.force, z)) || (rc = upload(M, &s.norm, zn)) ||
        (rc = upload(M, &s.np_off, np_off)) || (rc = upload(M, &s.np_cell, np_cell)) ||
        (rc = upload(M, &s.cp_cell, cp_cell)) || (rc = upload(M, &s.cp_off, cp_off)) ||
        (rc = upload(M, &s.cp_node, cp_node)) || (rc = upload(M, &s.cp_pair, P.cp_pair)) ||
        (rc = upload(M, &s.cell_c, c->L.centroid)) || (rc = dalloc(M, (void **)&s.pw, np_cell.size() * sizeof(double))))
        return rc;
    {
        // flux blocks that hold an affected cell: only these read the body force
        std::vector<int32_t> bf((size_t)c->L.n_blocks, 0);
        for (int32_t cell : cp_cell) bf[(size_t)cell / c->L.block_cells] = 1;
        if ((rc = upload(M, &s.blk_force, bf))) return rc;
        c->dm.blk_force = s.blk_force;
        c->dm32.blk_force = s.blk_force;
    }
    {
        const IbmArgs a = make_args(c);
        hipLaunchKernelGGL(k_ibm_weights, dim3((n + kNodesPerBlock - 1) / kNodesPerBlock), dim3(kIbmThreads), 0,
                           c->stream, a);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    HIP_TRY(hipMemset(c->fbody, 0, (size_t)3 * N * sizeof(double)));
    s.on = true;
    c->prm.use_body_force = 1;
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->graph_ok = false;
    return PLBFS_OK;
}

extern "C" int plbfs_set_ibm(plbfs_ctx *c, const plbfs_ibm_nodes *nodes) {
    if (!c || !nodes) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    return set_ibm(c, nodes, 0);
}

int ibm_rebuild(plbfs_ctx *c) {
    if (!c->ibm.on) return PLBFS_OK;
    const IbmState &s = c->ibm;
    const int n = s.n;
    std::vector<double> x((size_t)3 * n);
    int rc = plbfs_get_ibm(c, x.data(), nullptr, nullptr);
    if (rc) return rc;
    const std::vector<double> xr = s.xref_host;
    const plbfs_ibm_nodes nodes{n, xr.data(), x.data(), s.stiffness, s.node_area, s.dt, s.h, s.kernel};
    return set_ibm(c, &nodes, s.rebuilds + 1);
}

extern "C" int plbfs_get_ibm(plbfs_ctx *c, double *x, double *vel, double *force) {
    if (!c) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    if (!c->ibm.on) return plbfs_fail(PLBFS_ERR_INPUT, "no immersed boundary set");
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int n = c->ibm.n;
    double *src[3] = {c->ibm.x, c->ibm.vel, c->ibm.force};
    double *dst[3] = {x, vel, force};
    std::vector<double> h((size_t)3 * n);
    for (int a = 0; a < 3; a++) {
        if (!dst[a]) continue;
        HIP_TRY(hipMemcpy(h.data(), src[a], h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < n; k++)
            for (int q = 0; q < 3; q++) dst[a][3 * k + q] = h[(size_t)q * n + k];
    }
    return PLBFS_OK;
}
