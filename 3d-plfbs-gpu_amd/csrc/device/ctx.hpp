// Solver context shared by the HIP translation units (solver.hip, ibm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../host/host_common.hpp"
#include "../host/layout.hpp"
#include "lbfs_device.hpp"
#include "plbfs_hip.h"

#define HIP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return plbfs_fail(PLBFS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                                                  \
    } while (0)

namespace plbfs {
constexpr int kBlock = 256;
// default cells per flux block (face slots of one block share its LDS)
constexpr int kFluxBlockCells = 256;
// Dynamic LDS per flux workgroup.  gfx950 gives a workgroup 160 KiB
// (hipDeviceProp_t::sharedMemPerBlock), static and dynamic LDS together; the
// flux kernels' static arrays take at most 160 B, and a launch whose total
// exceeds 160 KiB is not refused by the runtime but faults, so 1 KiB stays free.
constexpr int kMaxLds = 159 * 1024;
}

struct KStat {
    double ms = 0;
    int launches = 0;
};

// Immersed-boundary state of a context (node arrays component-major [3][n]).
struct IbmState {
    bool on = false;
    int n = 0;
    int kernel = 0;
    double h = 1, dt = 0, stiffness = 0, node_area = 0, margin = 0;
    double *x = nullptr, *x0 = nullptr, *xref = nullptr, *vel = nullptr, *force = nullptr, *norm = nullptr;
    double *xc = nullptr;            // [3][n] positions the candidate lists were built around
    std::vector<double> xref_host;   // [3n] interleaved anchors, kept for list rebuilds
    int rebuilds = 0;                // list rebuilds since plbfs_set_ibm
    int32_t *np_off = nullptr, *np_cell = nullptr;           // node -> candidate cells
    int32_t *cp_cell = nullptr, *cp_off = nullptr, *cp_node = nullptr;  // affected cell -> nodes
    int32_t *cp_pair = nullptr;  // node-major pair index of each cell-major entry
    int32_t *blk_force = nullptr;  // [n_blocks] 1 where a flux block holds an affected cell
    double *pw = nullptr;        // [n_pairs] kernel weights at the current node positions
    int n_affected = 0;
    int64_t n_pairs = 0;
    double *cell_c = nullptr;  // [3][N] centroids, Hilbert order
};

struct plbfs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    plbfs::Layout L;
    plbfs_params params{};
    std::vector<plbfs_patch> patches;
    plbfs::DevParams prm{};
    std::vector<void *> allocs;
    std::vector<void *> ibm_allocs;  // the immersed-boundary arrays, freed when plbfs_set_ibm replaces them
    bool f32 = false;                 // FP32 storage of fields, gradients and face data
    plbfs::DevMesh<double> dm{};      // FP64 storage (also holds the shared int/FP64 arrays)
    plbfs::DevMesh<float> dm32{};     // FP32 storage
    double *stA = nullptr, *stB = nullptr;  // cell records (rho,u,v,w), ping-pong
    double *t0 = nullptr, *t1 = nullptr;    // conservative RK accumulators (rho, rho u, rho v, rho w)
    double *grad = nullptr;                 // 12 per cell
    double *dtl = nullptr;                  // [N]
    double *fbody = nullptr;                // [3][N] immersed-boundary force density
    void *b_fix = nullptr;                  // [G] fixed ghost states (storage precision)
    std::vector<double> ghost_fixed;        // [4*G] the same in FP64 on the host
    bool ghost_state_set = true;            // false until plbfs_set_ghost_state where needed
    void *bflux = nullptr;                  // [4][n_border] border-face fluxes (storage precision)
    double *partials = nullptr;             // [n_blocks*5]
    double *bmin = nullptr;                 // [n_blocks]
    double *hist = nullptr;
    int hist_cap = 0;
    plbfs::Ctrl *ctrl = nullptr;
    size_t lds_bytes = 0;
    int cell_grid = 0;  // blocks of kBlock cells for the per-cell kernels
    int total_steps = 0;
    std::vector<double> last_hist;
    IbmState ibm;
    // timing
    bool timing = false;
    KStat kstat[4];
    struct EvPair {
        int which;
        hipEvent_t a, b;
    };
    std::vector<EvPair> pending;
    // graph of one step
    hipGraphExec_t graph_exec = nullptr;
    bool graph_ok = false;
    bool use_graph = true;
    // this device's copy of the file-scope scratch array of solver.hip
    // (plbfs_run_noalloc), resolved once per device at creation
    char *scratch = nullptr;
};

inline int dalloc(std::vector<void *> &list, void **p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    HIP_TRY(hipMalloc(p, bytes));
    list.push_back(*p);
    return PLBFS_OK;
}

inline int dalloc(plbfs_ctx *c, void **p, size_t bytes) { return dalloc(c->allocs, p, bytes); }

template <typename T>
inline int upload(std::vector<void *> &list, T **dst, const std::vector<T> &src) {
    int rc = dalloc(list, (void **)dst, src.size() * sizeof(T));
    if (rc) return rc;
    if (!src.empty()) HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return PLBFS_OK;
}

template <typename T>
inline int upload(plbfs_ctx *c, T **dst, const std::vector<T> &src) {
    return upload(c->allocs, dst, src);
}

// Brackets one launch with events on the context's stream (timing mode only);
// the pairs are resolved after the run synchronises, so nothing blocks here.
struct TimedLaunch {
    plbfs_ctx *c;
    int which;
    bool on;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(plbfs_ctx *cc, int w, bool timed) : c(cc), which(w), on(timed) {
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~TimedLaunch() {
        if (on) {
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back({which, a, b});
        }
    }
};

// plbfs_create with an explicit flux block size (64..1024 cells, a multiple of
// 64), LDS sized for lds_cases cases per block (the batched solver sizes its
// own), staging of the flux block's own records, and border faces either
// evaluated once by a separate pass (split_border) or by both blocks.
int create_ctx(const plbfs_mesh_view *mesh, const plbfs_params *params, const plbfs_patch *patches,
               int32_t n_patches, const int32_t *bface_patch, int block_cells, int lds_cases, bool flux_stage,
               bool split_border, plbfs_ctx **out);

// Launches the per-step immersed-boundary kernels (ibm.hip) against the control
// block ctrl; no-op when off.  timed: bracket them with timing events;
// rebuild: flag drifted nodes for a list rebuild (ctrl->rebuild).
int ibm_launch(plbfs_ctx *c, plbfs::Ctrl *ctrl, bool timed, bool rebuild);
// Rebuilds the candidate lists around the nodes' current positions, as
// plbfs_set_ibm with those positions would (ibm.hip).
int ibm_rebuild(plbfs_ctx *c);

// Node-major candidate cells of every node (centroid within the kernel support
// plus a drift margin of the node's position) and the same pairs regrouped by cell.
struct IbmPairs {
    double margin = 0;
    std::vector<int32_t> np_off, np_cell;          // node -> candidate cells (sorted)
    std::vector<int32_t> cp_cell, cp_off, cp_node;  // affected cell -> nodes (node order)
    std::vector<int32_t> cp_pair;                   // ... and their node-major pair index
};
int build_ibm_pairs(const plbfs::Layout &L, const plbfs_ibm_nodes *nodes, IbmPairs &P);
