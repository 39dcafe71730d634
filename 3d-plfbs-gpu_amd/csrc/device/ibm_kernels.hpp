// Immersed-boundary kernels shared by the single-case solver (ibm.hip) and
// the batched solver (batch.hip).
//
// Per step, before the RK stages (src/Solver_gpu.cu:687-718):
//   k_ibm_nodes   one 64-wide wavefront per node: interpolate the node velocity
//                 from the step-start state (src/immersed_boundary_method.cu:13-90),
//                 move X = X0 + U dt (:244-295), spring force F = -k dA (X - X_ref)
//                 (:161-190), and the spreading normaliser at the new position;
//   k_ibm_spread  16 lanes per affected (cell, case) record: sum of the
//                 node forces over its pairs (:299-375).
// Weights are normalised per node, W_cn = delta V_c / sum delta V, which is the
// reference's formula on its uniform unit grid.  No atomics.
//
// For a batch, node k belongs to case node_case[k] and reads the state record
// cell * B + case; affected records carry their case in cp_case.  Nodes and
// records of a case whose case_done flag is set are skipped.
#pragma once
#include "lbfs_device.hpp"
#include "plbfs_hip.h"

namespace plbfs {
namespace {

constexpr int kIbmThreads = 256;
constexpr int kNodesPerBlock = kIbmThreads / 64;
constexpr int kIbmUnroll = 4;  // pairs per lane whose loads are issued together

__device__ __forceinline__ double ibm_phi(int kernel, double r) {
    r = fabs(r);
    if (kernel == PLBFS_DELTA_HAT2) return r < 1.0 ? 1.0 - r : 0.0;
    // Peskin, "The immersed boundary method", Acta Numerica 11 (2002), 4-point function
    if (r <= 1.0) return (3.0 - 2.0 * r + sqrt(1.0 + 4.0 * r - 4.0 * r * r)) / 8.0;
    if (r <= 2.0) return (5.0 - 2.0 * r - sqrt(-7.0 + 12.0 * r - 4.0 * r * r)) / 8.0;
    return 0.0;
}

__device__ __forceinline__ double ibm_delta(int kernel, double ih, const double *__restrict__ cc, int N, int c,
                                            const double X[3]) {
    return ibm_phi(kernel, (cc[c] - X[0]) * ih) * ibm_phi(kernel, (cc[(size_t)N + c] - X[1]) * ih) *
           ibm_phi(kernel, (cc[2 * (size_t)N + c] - X[2]) * ih) * (ih * ih * ih);
}

__device__ __forceinline__ double wave_allsum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

struct IbmArgs {
    int n, N, kernel;
    int B;  // cases per cell record (1 for a single case)
    double ih, dt, k_dA, margin;
    double *x, *x0, *xref, *vel, *force, *norm;
    const double *xc;        // [3][n] positions the candidate lists were built around
    int rebuild;             // 1: ask the host for new lists at half the margin (single case)
    const double *node_kdA;  // per-node k dA, or null for the scalar k_dA
    const int32_t *node_case;  // per-node case, or null (case 0)
    const int32_t *np_off, *np_cell;
    const int32_t *cp_cell, *cp_off, *cp_node;
    const int32_t *cp_pair;  // node-major pair index of every cell-major entry
    const int32_t *cp_case;  // per affected record case, or null (case 0)
    double *pw;              // [n_pairs] kernel weight delta(x_c - X_k) at the current node position
    const int32_t *case_done;  // per-case freeze flags, or null
    int *case_status;          // per-case status, or null
    int n_affected;
    const double *cc, *vol;
};

template <typename SV>
__global__ __launch_bounds__(kIbmThreads) void k_ibm_nodes(IbmArgs a, Ctrl *__restrict__ ctrl,
                                                           const SV *__restrict__ st) {
    if (ctrl->done) return;
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * kNodesPerBlock + (threadIdx.x >> 6);
    if (k >= a.n) return;
    const int b = a.node_case ? a.node_case[k] : 0;
    if (a.case_done && a.case_done[b]) return;
    const int n = a.n;
    const double X[3] = {a.x[k], a.x[(size_t)n + k], a.x[2 * (size_t)n + k]};
    const int p0 = a.np_off[k], p1 = a.np_off[k + 1];
    double S = 0, su = 0, sv = 0, sw = 0;
    // weights at X: computed at set-up or by the previous step's second pass.
    // A lane takes kIbmUnroll of its pairs at a time and issues all their
    // loads before the first use, so their latencies overlap; the sums still
    // run over the lane's pairs in list order.
    for (int pb = p0 + lane; pb < p1; pb += 64 * kIbmUnroll) {
        int c[kIbmUnroll];
        double d[kIbmUnroll], vol[kIbmUnroll];
        SV sc[kIbmUnroll];
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            const int p = pb + 64 * u;
            c[u] = p < p1 ? a.np_cell[p] : 0;
            d[u] = p < p1 ? a.pw[p] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            if (d[u] != 0.0) {
                vol[u] = a.vol[c[u]];
                sc[u] = st[(size_t)c[u] * a.B + b];
            }
        }
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            if (d[u] != 0.0) {
                const double wv = d[u] * vol[u];
                S += wv;
                su += wv * sc[u].y;
                sv += wv * sc[u].z;
                sw += wv * sc[u].w;
            }
        }
    }
    S = wave_allsum(S);
    su = wave_allsum(su);
    sv = wave_allsum(sv);
    sw = wave_allsum(sw);
    // a node with no cell in its support (S = 0) is an error below; it stays
    // where it is, so that only finite positions and forces are written
    const bool supported = S > 0.0;
    const double U[3] = {supported ? su / S : 0.0, supported ? sv / S : 0.0, supported ? sw / S : 0.0};
    const double kdA = a.node_kdA ? a.node_kdA[k] : a.k_dA;
    double Xn[3], F[3], drift = 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double x0 = a.x0[(size_t)q * n + k], xr = a.xref[(size_t)q * n + k];
        Xn[q] = x0 + U[q] * a.dt;
        F[q] = (-kdA * (Xn[q] - xr));
        drift = fmax(drift, fabs(Xn[q] - a.xc[(size_t)q * n + k]));
    }
    double S2 = 0;
    for (int pb = p0 + lane; pb < p1; pb += 64 * kIbmUnroll) {
        int c[kIbmUnroll];
        double d[kIbmUnroll], vol[kIbmUnroll];
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) c[u] = pb + 64 * u < p1 ? a.np_cell[pb + 64 * u] : 0;
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            const int p = pb + 64 * u;
            d[u] = 0.0;
            if (p < p1) {
                d[u] = ibm_delta(a.kernel, a.ih, a.cc, a.N, c[u], Xn);
                a.pw[p] = d[u];
            }
        }
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++)
            if (d[u] != 0.0) vol[u] = a.vol[c[u]];
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++)
            if (d[u] != 0.0) S2 += d[u] * vol[u];
    }
    S2 = wave_allsum(S2);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 3; q++) {
            a.x[(size_t)q * n + k] = Xn[q];
            a.x0[(size_t)q * n + k] = Xn[q];
            a.vel[(size_t)q * n + k] = U[q];
            a.force[(size_t)q * n + k] = F[q];
        }
        a.norm[k] = S2;
        // past half the margin the lists are rebuilt around the new positions
        // after this step; past the margin they may miss cells: an error
        if (a.rebuild && drift > 0.5 * a.margin) ctrl->rebuild = 1;
        if (!supported || !(S2 > 0.0) || drift > a.margin) {
            if (a.case_status) {
                // batch: only this case stops
                a.case_status[b] = PLBFS_ERR_IBM;
            } else {
                ctrl->status = PLBFS_ERR_IBM;
                ctrl->done = 1;
            }
        }
    }
}

// Kernel weights of every pair at the nodes' current positions (set-up).
__global__ __launch_bounds__(kIbmThreads) void k_ibm_weights(IbmArgs a) {
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * kNodesPerBlock + (threadIdx.x >> 6);
    if (k >= a.n) return;
    const int n = a.n;
    const double X[3] = {a.x[k], a.x[(size_t)n + k], a.x[2 * (size_t)n + k]};
    for (int p = a.np_off[k] + lane; p < a.np_off[k + 1]; p += 64)
        a.pw[p] = ibm_delta(a.kernel, a.ih, a.cc, a.N, a.np_cell[p], X);
}

// kSpreadLanes lanes per affected record (kSpreadPerBlock records per
// workgroup): the lanes stride over the record's pairs in node order and a fixed
// butterfly inside the lane group sums them, so every record's force has one
// summation order wherever it is spread from.  A record has a few dozen pairs
// (a node's Peskin-4 support spans 4^3 cells, its candidate list with the drift
// margin 6^3), so a whole wavefront per record would leave most lanes idle.
// fbody is [3][B*N] component-major over records.
constexpr int kSpreadLanes = 16;
constexpr int kSpreadPerBlock = kIbmThreads / kSpreadLanes;

__device__ __forceinline__ double group_allsum(double x) {
#pragma unroll
    for (int off = kSpreadLanes / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__global__ __launch_bounds__(kIbmThreads) void k_ibm_spread(IbmArgs a, const Ctrl *__restrict__ ctrl,
                                                            double *__restrict__ fbody) {
    if (ctrl->done) return;
    const int lane = threadIdx.x & (kSpreadLanes - 1);
    const int j = blockIdx.x * kSpreadPerBlock + (int)(threadIdx.x / kSpreadLanes);
    // every lane of a group has the same j and case, so a group leaves as a
    // whole and the shuffles below never read a lane that has left
    if (j >= a.n_affected) return;
    const int c = a.cp_cell[j];
    const int b = a.cp_case ? a.cp_case[j] : 0;
    if (a.case_done && a.case_done[b]) return;
    const int n = a.n;
    double f0 = 0, f1 = 0, f2 = 0;
    const int q0 = a.cp_off[j], q1 = a.cp_off[j + 1];
    // kIbmUnroll pairs per lane at a time, loads first (as in k_ibm_nodes)
    for (int pb = q0 + lane; pb < q1; pb += kSpreadLanes * kIbmUnroll) {
        int k[kIbmUnroll];
        double d[kIbmUnroll], nk[kIbmUnroll], fx[kIbmUnroll], fy[kIbmUnroll], fz[kIbmUnroll];
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            const int p = pb + kSpreadLanes * u;
            k[u] = p < q1 ? a.cp_node[p] : 0;
            d[u] = p < q1 ? a.pw[a.cp_pair[p]] : 0.0;  // at the node's new position
        }
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            if (d[u] != 0.0) {
                nk[u] = a.norm[k[u]];
                fx[u] = a.force[k[u]];
                fy[u] = a.force[(size_t)n + k[u]];
                fz[u] = a.force[2 * (size_t)n + k[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kIbmUnroll; u++) {
            if (d[u] != 0.0) {
                const double q = d[u] / nk[u];
                f0 += fx[u] * q;
                f1 += fy[u] * q;
                f2 += fz[u] * q;
            }
        }
    }
    f0 = group_allsum(f0);
    f1 = group_allsum(f1);
    f2 = group_allsum(f2);
    if (lane == 0) {
        const size_t NR = (size_t)a.N * a.B, r = (size_t)c * a.B + b;
        fbody[r] = f0;
        fbody[NR + r] = f1;
        fbody[2 * NR + r] = f2;
    }
}

}  // namespace
}  // namespace plbfs
