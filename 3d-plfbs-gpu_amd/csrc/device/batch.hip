// Batched PLBFS: B independent cases on one mesh topology (gfx950).
//
// The cases share the mesh, its layout (Hilbert order, face items, gather
// lists) and the LSQ weights; each case has its own viscosity (Reynolds
// number), state, local time steps and immersed nodes.  Cell records are
// stored case-fastest, record i * B + b, so the cases of one cell are adjacent:
// a gather of cell j for C consecutive cases is one contiguous run, and the
// geometry of a face item is loaded once for C cases.
//
// One RK stage of all cases is two launches:
//   kb_grad   one thread per (cell, case) record;
//   kb_flux   one workgroup per (flux block, case group): face items x C cases
//             into LDS, then the records of the block's cells x C cases gather
//             their slots in the reference face order and update.
// After the fourth stage kb_finalize (one workgroup per case) reduces the
// per-block partials in a fixed order into the case's residual and freezes a
// converged or failed case; kb_alldone stops the replay once every case is
// frozen.  One step is captured once as a hipGraph and replayed.
//
// Tiling (cells per flux block, cases per workgroup) is picked by
// plbfs_batch_autotune from timed trial steps of the candidates that fit LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "ibm_kernels.hpp"
#include "stage.hpp"

using namespace plbfs;

namespace {

struct BatchArgs {
    int B, C;
    int stage;              // 1: stage the tile's own (cell, case) records in LDS
    const double *visc;     // [B] nu
    const double *v3g;      // [B] 3 nu / gamma
    const double *dtcap;    // [B] streaming dt cap 3 nu / (gamma fneq_min), or infinity
    const int32_t *done;    // [B] frozen cases
    const uint8_t *force_rec;  // [N*B] 1 where a (cell, case) record receives immersed-boundary force
};

// Per-case control written by kb_finalize.
struct CaseCtl {
    int32_t *done, *status, *steps, *converged;
    double *res;  // [B][4]
};

template <typename S>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void kb_grad(
    DevMesh<S> m, BatchArgs a, const Ctrl *__restrict__ ctrl, const V4<S> *__restrict__ st,
    V4<S> *__restrict__ grad) {
    if (ctrl->done) return;
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int i = (int)(idx / a.B);
    const int b = (int)(idx - (int64_t)i * a.B);
    if (i >= m.N || a.done[b]) return;
    S g[12];
    lsq_grad<S>(m, i, GlobalStates<S, RecB>{st, RecB{a.B, b}}, g);
    // the batch keeps the state apart: three gradient V4 per (cell, case) record
    V4<S> *gi = grad + 3 * ((size_t)i * a.B + b);
    gi[0] = V4<S>{g[0], g[1], g[2], g[3]};
    gi[1] = V4<S>{g[4], g[5], g[6], g[7]};
    gi[2] = V4<S>{g[8], g[9], g[10], g[11]};
}

// Border pass (split border): faces between two flux blocks, one thread per
// (face, case) record with the cases of a face adjacent, evaluated once into
// bflux [4][n_border][B] that both blocks' gathers read.
template <typename S>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(std::is_same<S, float>::value ? 6 : 1)))
void kb_flux_border(DevMesh<S> m, DevParams prm, BatchArgs a, const Ctrl *__restrict__ ctrl,
                    const V4<S> *__restrict__ st_in, const V4<S> *__restrict__ grad) {
    if (ctrl->done) return;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int j = (int)(k / a.B), b = (int)(k - (int64_t)j * a.B);
    if (j >= m.n_border || a.done[b]) return;
    const SideLoader<S, RecB, false> ld{st_in, grad, m.cgeo, RecB{a.B, b}, nullptr, nullptr, 0, 0, 1, 0};
    S F[4];
    item_flux<S>(m, prm, m.border0 + j, ld, a.v3g[b], a.dtcap[b], F);
    const size_t nbd = (size_t)m.n_border;
#pragma unroll
    for (int q = 0; q < 4; q++) m.bflux[(q * nbd + j) * a.B + b] = F[q];
}

template <typename S>
__global__ __launch_bounds__(kBlock) void kb_flux(DevMesh<S> m, DevParams prm, BatchArgs a, StageArgs sa,
                                                  const Ctrl *__restrict__ ctrl, const V4<S> *__restrict__ st_in,
                                                  const V4<S> *__restrict__ grad, V4<S> *__restrict__ st_out,
                                                  double4 *__restrict__ t0, Acc<S> *__restrict__ t1,
                                                  double *__restrict__ dtl, const double *__restrict__ fbody,
                                                  double *__restrict__ partials) {
    extern __shared__ double lds[];
    if (ctrl->done) return;
    const int N = m.N, B = a.B, C = a.C;
    const int blk = blockIdx.x, b0 = blockIdx.y * C;
    const int e0 = m.blk_item_off[blk], e1 = m.blk_item_off[blk + 1];
    const int stride = m.max_items * C;
    S *sF = reinterpret_cast<S *>(lds);
    const int c0 = blk * m.block_cells;
    const int ncell = min(N, c0 + m.block_cells) - c0;
    V4<S> *sRec = nullptr;
    double4 *sCg = nullptr;
    int nst = 0;
    if (a.stage) {
        char *base = reinterpret_cast<char *>(lds) + staged_offset<S>(m.max_items, C);
        sRec = reinterpret_cast<V4<S> *>(base);
        if (std::is_same<S, double>::value)
            sCg = reinterpret_cast<double4 *>(base + (size_t)m.block_cells * C * 4 * sizeof(V4<S>));
        stage_records<S, false>(st_in, grad, std::is_same<S, double>::value ? m.cgeo : nullptr, B, b0, C, c0, ncell, sRec,
                         sCg);
        nst = ncell;
        __syncthreads();
    }
    // kBlock is a multiple of C: thread t serves case lane c = t % C, items
    // t / C, t / C + kBlock / C, ..., so its case's flags and viscosity are
    // read once
    const int c = threadIdx.x % C, b = b0 + c;
    const bool live = b < B && !a.done[b];
    if (live) {
        const double v3g = a.v3g[b], dtcap = a.dtcap[b];
        const SideLoader<S, RecB, false> ld{st_in, grad, m.cgeo, RecB{B, b}, sRec, sCg, c0, nst, C, c};
        for (int el = threadIdx.x / C; el < e1 - e0; el += kBlock / C) {
            S F[4];
            item_flux<S>(m, prm, e0 + el, ld, v3g, dtcap, F);
#pragma unroll
            for (int q = 0; q < 4; q++) sF[q * stride + el * C + c] = F[q];
        }
    }
    __syncthreads();
    double sq[4] = {0, 0, 0, 0};
    int err = 0;
    const size_t NR = (size_t)N * B;
    const double visc = live ? a.visc[b] : 0.0;
    for (int il = threadIdx.x / C; live && il < ncell; il += kBlock / C) {
        const int i = c0 + il;
        double R[4];
        gather_residual<S>(m, i, sF, stride, C, c, R, B, b);
        const size_t r = (size_t)i * B + b;
        if (prm.use_body_force && a.force_rec[r]) {
            R[1] += fbody[r];
            R[2] += fbody[NR + r];
            R[3] += fbody[2 * NR + r];
        }
        const int e = rk_update<S>(m, prm, visc, sa, i, r, R, st_out, t0, t1, dtl, sq);
        if (e == 1) err = 1;
        else if (e == 2 && err == 0) err = 2;
    }
    if (sa.stage == 3) {
        __syncthreads();  // the face slots are no longer read; reuse the LDS
        double *red = lds;
#pragma unroll
        for (int q = 0; q < 4; q++) red[threadIdx.x * 5 + q] = sq[q];
        red[threadIdx.x * 5 + 4] = (double)err;
        __syncthreads();
        if ((int)threadIdx.x < C) {
            const int b = b0 + threadIdx.x;
            if (b < B && !a.done[b]) {
                double s[5] = {0, 0, 0, 0, 0};
                for (int t = threadIdx.x; t < kBlock; t += C) {
                    for (int q = 0; q < 4; q++) s[q] += red[t * 5 + q];
                    s[4] = fmax(s[4], red[t * 5 + 4]);
                }
                for (int q = 0; q < 5; q++) partials[((size_t)blk * B + b) * 5 + q] = s[q];
            }
        }
    }
}

constexpr int kFinThreads = 256;

// One workgroup per case: fixed-order sum of the block partials
// (src/residuals.cpp:35-44,139-147) and the case's stop test.
__global__ __launch_bounds__(kFinThreads) void kb_finalize(int nblk, int B, double n_cells, double ref_rho,
                                                           double ref_mom, double tol, const Ctrl *__restrict__ ctrl,
                                                           CaseCtl cc, const double *__restrict__ partials) {
    if (ctrl->done) return;
    const int b = blockIdx.x;
    if (cc.done[b]) return;
    __shared__ double red[kFinThreads][5];
    double s[5] = {0, 0, 0, 0, 0};
    // a thread's blocks k = tid, tid + T, ... in that order; four at a time
    // with all their loads issued before the sums
    constexpr int kU = 4;
    for (int k0 = threadIdx.x; k0 < nblk; k0 += kFinThreads * kU) {
        double v[kU][5];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int k = k0 + u * kFinThreads;
            if (k < nblk) {
                const double *p = partials + ((size_t)k * B + b) * 5;
                for (int c = 0; c < 5; c++) v[u][c] = p[c];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            if (k0 + u * kFinThreads < nblk) {
                for (int c = 0; c < 4; c++) s[c] += v[u][c];
                s[4] = fmax(s[4], v[u][4]);
            }
        }
    }
    for (int c = 0; c < 5; c++) red[threadIdx.x][c] = s[c];
    __syncthreads();
    for (int off = kFinThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int c = 0; c < 4; c++) red[threadIdx.x][c] += red[threadIdx.x + off][c];
            red[threadIdx.x][4] = fmax(red[threadIdx.x][4], red[threadIdx.x + off][4]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double r[4];
        r[0] = sqrt(red[0][0] / n_cells) / ref_rho;
        for (int c = 1; c < 4; c++) r[c] = sqrt(red[0][c] / n_cells) / ref_mom;
        for (int c = 0; c < 4; c++) cc.res[(size_t)b * 4 + c] = r[c];
        cc.steps[b] += 1;
        const int e = (int)red[0][4];
        // a status set earlier in the step (the immersed nodes) is the cause
        // and is kept; NaN or rho blow-up is reported only on its own
        if (cc.status[b] == PLBFS_OK) {
            if (e == 1) cc.status[b] = PLBFS_ERR_NAN;
            else if (e == 2) cc.status[b] = PLBFS_ERR_RHO;
        }
        if (cc.status[b] != PLBFS_OK) {
            cc.done[b] = 1;
        } else if (fmax(fmax(r[3], r[1]), r[2]) < tol) {
            cc.converged[b] = 1;
            cc.done[b] = 1;
        }
    }
}

__global__ __launch_bounds__(kFinThreads) void kb_alldone(int B, Ctrl *__restrict__ ctrl,
                                                          const int32_t *__restrict__ done) {
    if (ctrl->done) return;
    __shared__ int live;
    if (threadIdx.x == 0) live = 0;
    __syncthreads();
    int mine = 0;
    for (int b = threadIdx.x; b < B; b += kFinThreads) mine |= done[b] ? 0 : 1;
    if (mine) live = 1;  // benign same-value race
    __syncthreads();
    if (threadIdx.x == 0) {
        ctrl->step = ctrl->step + 1;
        if (!live) ctrl->done = 1;
    }
}

template <typename S>
__global__ __launch_bounds__(kBlock) void kb_init_dt(DevMesh<S> m, DevParams prm, BatchArgs a,
                                                     const V4<S> *__restrict__ st, double *__restrict__ dtl) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int i = (int)(idx / a.B);
    const int b = (int)(idx - (int64_t)i * a.B);
    if (i >= m.N || a.done[b]) return;
    const double r0 = std::is_same<S, float>::value ? prm.rho0 : 0.0;
    const V4<S> q = st[idx];
    if (prm.time_stepping == PLBFS_DT_CONSTANT) dtl[idx] = prm.cfl;
    else dtl[idx] = local_dt(prm, a.visc[b], r0 + (double)q.x, (double)q.y, m.cfl_area[i], m.vol[i]);
}

// Case b's records <- one cell-ordered state (Hilbert order); b < 0: all cases.
template <typename S>
__global__ __launch_bounds__(kBlock) void kb_scatter(int N, int B, int b, double rho0, const double4 *__restrict__ src,
                                                     double4 *__restrict__ t0, V4<S> *__restrict__ st) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int i = (int)(idx / B), bb = (int)(idx - (int64_t)i * B);
    if (i >= N || (b >= 0 && bb != b)) return;
    const double4 q = src[i];
    t0[idx] = q;
    st[idx] = V4<S>{(S)(q.x - rho0), (S)q.y, (S)q.z, (S)q.w};
}

__global__ __launch_bounds__(kBlock) void kb_gather(int N, int B, int b, const double4 *__restrict__ src,
                                                    double4 *__restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    dst[i] = src[(size_t)i * B + b];
}

}  // namespace

// Host-side node set of one case.
struct BatchIbmCase {
    bool on = false;
    std::vector<double> xref, x;  // [3n] interleaved
    double kdA = 0;
    IbmPairs pairs;
};

struct plbfs_batch {
    // inputs, kept for re-tiling
    int N = 0, F = 0, NI = 0, T = 0;
    std::vector<int32_t> owner, nbr, bpatch;
    std::vector<double> cc, vol, fc, fn, fa, fdt;
    plbfs_params params{};
    std::vector<plbfs_patch> patches;
    int B = 0;
    std::vector<double> visc;
    bool f32 = false;
    // shared mesh, layout and stream
    plbfs_ctx *base = nullptr;
    int block_cells = 256, C = 1, stage = 0;
    int split = 0;             // 1: faces between flux blocks evaluated once by kb_flux_border
    void *bflux = nullptr;     // [4][n_border][B] border fluxes (storage precision)
    size_t bflux_cap = 0;
    std::vector<void *> allocs;
    // device arrays over N * B records
    void *stA = nullptr, *stB = nullptr, *grad = nullptr, *t1 = nullptr;
    double *t0 = nullptr, *dtl = nullptr, *fbody = nullptr, *partials = nullptr, *scratch = nullptr;
    size_t partials_cap = 0;
    double *d_visc = nullptr, *d_v3g = nullptr, *d_dtcap = nullptr;
    int32_t *done = nullptr, *status = nullptr, *steps = nullptr, *conv = nullptr;
    double *res = nullptr;
    Ctrl *ctrl = nullptr;
    // immersed boundary, all cases concatenated
    std::vector<BatchIbmCase> ibm_case;
    bool ibm_dirty = false, ibm_on = false;
    int ibm_kernel = 0;
    double ibm_h = 1, ibm_dt = 0, ibm_margin = 0;
    IbmArgs ia{};
    std::vector<void *> ibm_allocs;
    uint8_t *force_rec = nullptr;  // in ibm_allocs
    // graph of one step
    hipGraphExec_t graph = nullptr;
    bool graph_ok = false;
};

namespace {

int balloc(std::vector<void *> &list, void **p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    HIP_TRY(hipMalloc(p, bytes));
    list.push_back(*p);
    return PLBFS_OK;
}
int balloc(plbfs_batch *b, void **p, size_t bytes) { return balloc(b->allocs, p, bytes); }

template <typename T>
int bupload(std::vector<void *> &list, T **dst, const std::vector<T> &src) {
    int rc = balloc(list, (void **)dst, src.size() * sizeof(T));
    if (rc) return rc;
    if (!src.empty()) HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return PLBFS_OK;
}
template <typename T>
int bupload(plbfs_batch *b, T **dst, const std::vector<T> &src) {
    return bupload(b->allocs, dst, src);
}

void free_list(std::vector<void *> &list) {
    for (void *p : list) (void)hipFree(p);
    list.clear();
}

void drop_graph(plbfs_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_ok = false;
}

plbfs_mesh_view view_of(plbfs_batch *b) {
    return plbfs_mesh_view{b->N, b->F, b->NI, b->owner.data(), b->nbr.data(), b->cc.data(), b->vol.data(),
                           b->fc.data(), b->fn.data(), b->fa.data(), b->fdt.data()};
}

size_t lds_bytes_for(const plbfs::Layout &L, int C, bool f32, bool stage) {
    const size_t b = f32 ? flux_lds_bytes<float>(L.max_items_per_block, L.block_cells, stage, C)
                        : flux_lds_bytes<double>(L.max_items_per_block, L.block_cells, stage, C);
    return std::max(b, (size_t)kBlock * 5 * sizeof(double));
}

// (Re)builds the shared layout for a flux block of block_cells cells and C
// cases, staging the tile's records in LDS when stage is set.  The new layout
// is built beside the current one and replaces it only once it is complete:
// a tiling that is refused (LDS, allocation) leaves the batch as it was.
int retile(plbfs_batch *b, int block_cells, int C, int stage) {
    if (block_cells < 64 || block_cells > 1024 || block_cells % 64)
        return plbfs_fail(PLBFS_ERR_INPUT, "block_cells %d: need a multiple of 64 in [64, 1024]", block_cells);
    if (C < 1 || C > kBlock || kBlock % C)
        return plbfs_fail(PLBFS_ERR_INPUT, "cases_per_block %d: need a divisor of %d", C, kBlock);
    plbfs_mesh_view v = view_of(b);
    plbfs_ctx *nb = nullptr;
    int rc = create_ctx(&v, &b->params, b->patches.data(), (int32_t)b->patches.size(),
                        b->bpatch.empty() ? nullptr : b->bpatch.data(), block_cells, C, false, b->split != 0, &nb);
    if (rc) return rc;
    const size_t lds = lds_bytes_for(nb->L, C, b->f32, stage != 0);
    if (lds > (size_t)kMaxLds) {
        plbfs_destroy(nb);
        return plbfs_fail(PLBFS_ERR_INPUT, "tile of %d cells x %d cases needs %zu B of LDS (limit %d)", block_cells,
                          C, lds, kMaxLds);
    }
    nb->lds_bytes = lds;
    const size_t need = (size_t)nb->L.n_blocks * b->B * 5;
    if (need > b->partials_cap) {
        // the old array stays in b->allocs (freed with the batch): queued work may still read it
        double *np = nullptr;
        if ((rc = balloc(b, (void **)&np, need * sizeof(double)))) {
            plbfs_destroy(nb);
            return rc;
        }
        b->partials = np;
        b->partials_cap = need;
    }
    const size_t nbf = (size_t)4 * nb->L.n_border * b->B * (b->f32 ? 4 : 8);
    if (nbf > b->bflux_cap) {
        // as the partials: an array that queued work may read stays allocated
        void *np = nullptr;
        if ((rc = balloc(b, &np, nbf))) {
            plbfs_destroy(nb);
            return rc;
        }
        b->bflux = np;
        b->bflux_cap = nbf;
    }
    drop_graph(b);
    if (b->base) plbfs_destroy(b->base);
    b->base = nb;
    b->block_cells = block_cells;
    b->C = C;
    b->stage = stage;
    return PLBFS_OK;
}

BatchArgs args_of(plbfs_batch *b) {
    return BatchArgs{b->B, b->C, b->stage, b->d_visc, b->d_v3g, b->d_dtcap, b->done, b->force_rec};
}

int ibm_upload(plbfs_batch *b) {
    b->ibm_dirty = false;
    b->ibm_on = false;
    drop_graph(b);
    HIP_TRY(hipStreamSynchronize(b->base->stream));
    free_list(b->ibm_allocs);
    b->force_rec = nullptr;
    std::vector<void *> &M = b->ibm_allocs;
    std::vector<double> xr, xx, kda;
    std::vector<int32_t> ncase, np_off(1, 0), np_cell, cp_cell, cp_case, cp_off(1, 0), cp_node, cp_pair;
    int n_total = 0;
    for (int c = 0; c < b->B; c++) {
        const BatchIbmCase &ic = b->ibm_case[c];
        if (!ic.on) continue;
        const int n = (int)ic.xref.size() / 3;
        const int32_t pair0 = (int32_t)np_cell.size();
        for (int k = 0; k < n; k++) {
            ncase.push_back(c);
            kda.push_back(ic.kdA);
            for (int p = ic.pairs.np_off[k]; p < ic.pairs.np_off[k + 1]; p++) np_cell.push_back(ic.pairs.np_cell[p]);
            np_off.push_back((int32_t)np_cell.size());
        }
        for (size_t j = 0; j < ic.pairs.cp_cell.size(); j++) {
            cp_cell.push_back(ic.pairs.cp_cell[j]);
            cp_case.push_back(c);
            for (int p = ic.pairs.cp_off[j]; p < ic.pairs.cp_off[j + 1]; p++) {
                cp_node.push_back(n_total + ic.pairs.cp_node[p]);
                cp_pair.push_back(pair0 + ic.pairs.cp_pair[p]);
            }
            cp_off.push_back((int32_t)cp_node.size());
        }
        n_total += n;
    }
    if (n_total == 0) return PLBFS_OK;
    // node arrays component-major [3][n_total]
    xr.assign((size_t)3 * n_total, 0.0);
    xx.assign((size_t)3 * n_total, 0.0);
    int k0 = 0;
    for (int c = 0; c < b->B; c++) {
        const BatchIbmCase &ic = b->ibm_case[c];
        if (!ic.on) continue;
        const int n = (int)ic.xref.size() / 3;
        for (int k = 0; k < n; k++)
            for (int q = 0; q < 3; q++) {
                xr[(size_t)q * n_total + k0 + k] = ic.xref[3 * k + q];
                xx[(size_t)q * n_total + k0 + k] = ic.x[3 * k + q];
            }
        k0 += n;
    }
    IbmArgs &a = b->ia;
    a = IbmArgs{};
    a.n = n_total;
    a.N = b->N;
    a.B = b->B;
    a.kernel = b->ibm_kernel;
    a.ih = 1.0 / b->ibm_h;
    a.dt = b->ibm_dt;
    a.k_dA = 0;
    a.margin = b->ibm_margin;
    a.n_affected = (int)cp_cell.size();
    std::vector<double> z((size_t)3 * n_total, 0.0), zn(n_total, 0.0);
    int rc;
    int32_t *nc = nullptr, *npo = nullptr, *npc = nullptr, *cpc = nullptr, *cpk = nullptr, *cpo = nullptr,
            *cpn = nullptr, *cpp = nullptr;
    double *kd = nullptr, *ccd = nullptr, *pw = nullptr;
    double *xcd = nullptr;  // the lists were built around the positions given to set_ibm
    if ((rc = bupload(M, &a.xref, xr)) || (rc = bupload(M, &a.x, xx)) || (rc = bupload(M, &a.x0, xx)) ||
        (rc = bupload(M, &xcd, xx)) ||
        (rc = bupload(M, &a.vel, z)) || (rc = bupload(M, &a.force, z)) || (rc = bupload(M, &a.norm, zn)) ||
        (rc = bupload(M, &kd, kda)) || (rc = bupload(M, &nc, ncase)) || (rc = bupload(M, &npo, np_off)) ||
        (rc = bupload(M, &npc, np_cell)) || (rc = bupload(M, &cpc, cp_cell)) || (rc = bupload(M, &cpk, cp_case)) ||
        (rc = bupload(M, &cpo, cp_off)) || (rc = bupload(M, &cpn, cp_node)) || (rc = bupload(M, &cpp, cp_pair)) ||
        (rc = bupload(M, &ccd, b->base->L.centroid)) ||
        (rc = balloc(M, (void **)&pw, np_cell.size() * sizeof(double))))
        return rc;
    a.xc = xcd;
    a.rebuild = 0;  // a batch case stops at the margin (PLBFS_ERR_IBM)
    a.node_kdA = kd;
    a.node_case = nc;
    a.np_off = npo;
    a.np_cell = npc;
    a.cp_cell = cpc;
    a.cp_case = cpk;
    a.cp_off = cpo;
    a.cp_node = cpn;
    a.cp_pair = cpp;
    a.pw = pw;
    a.cc = ccd;
    a.vol = b->base->dm.vol;
    a.case_done = b->done;
    a.case_status = b->status;
    {
        // records the spreading writes: only these read the body force
        std::vector<uint8_t> fr((size_t)b->N * b->B, 0);
        for (size_t j = 0; j < cp_cell.size(); j++) fr[(size_t)cp_cell[j] * b->B + cp_case[j]] = 1;
        if ((rc = bupload(M, &b->force_rec, fr))) return rc;
    }
    HIP_TRY(hipMemset(b->fbody, 0, (size_t)3 * b->N * b->B * sizeof(double)));
    hipLaunchKernelGGL(k_ibm_weights, dim3((n_total + kNodesPerBlock - 1) / kNodesPerBlock), dim3(kIbmThreads), 0,
                       b->base->stream, a);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->base->stream));
    b->ibm_on = true;
    return PLBFS_OK;
}

template <typename S>
void launch_batch_stages(plbfs_batch *b, const DevMesh<S> &dm_base) {
    plbfs_ctx *c = b->base;
    // the base context's border array holds one case; the batch's holds B
    DevMesh<S> dm = dm_base;
    dm.bflux = (S *)b->bflux;
    const int64_t nbr = (int64_t)dm.n_border * b->B;
    const BatchArgs a = args_of(b);
    const int64_t NR = (int64_t)b->N * b->B;
    const dim3 rgrid((unsigned)((NR + kBlock - 1) / kBlock));
    const dim3 fgrid(c->L.n_blocks, (b->B + b->C - 1) / b->C);
    DevParams prm = c->prm;
    prm.use_body_force = b->ibm_on ? 1 : 0;
    static const double kAl[4] = {1.0, 0.5, 0.5, 1.0};
    static const double kBe[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    void *in = b->stA, *out = b->stB;
    for (int k = 0; k < 4; k++) {
        hipLaunchKernelGGL(kb_grad<S>, rgrid, dim3(kBlock), 0, c->stream, dm, a, b->ctrl, (const V4<S> *)in,
                           (V4<S> *)b->grad);
        StageArgs sa{kAl[k], kBe[k], k};
        if (nbr > 0)
            hipLaunchKernelGGL(kb_flux_border<S>, dim3((unsigned)((nbr + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               c->stream, dm, prm, a, b->ctrl, (const V4<S> *)in, (const V4<S> *)b->grad);
        hipLaunchKernelGGL(kb_flux<S>, fgrid, dim3(kBlock), c->lds_bytes, c->stream, dm, prm, a, sa, b->ctrl,
                           (const V4<S> *)in, (const V4<S> *)b->grad, (V4<S> *)out, (double4 *)b->t0,
                           (Acc<S> *)b->t1, b->dtl, b->fbody, b->partials);
        std::swap(in, out);
    }
}

int launch_batch_step(plbfs_batch *b) {
    plbfs_ctx *c = b->base;
    if (b->ibm_on) {
        const dim3 grid((b->ia.n + kNodesPerBlock - 1) / kNodesPerBlock);
        if (b->f32)
            hipLaunchKernelGGL(k_ibm_nodes<float4>, grid, dim3(kIbmThreads), 0, c->stream, b->ia, b->ctrl,
                               (const float4 *)b->stA);
        else
            hipLaunchKernelGGL(k_ibm_nodes<double4>, grid, dim3(kIbmThreads), 0, c->stream, b->ia, b->ctrl,
                               (const double4 *)b->stA);
        hipLaunchKernelGGL(k_ibm_spread, dim3((b->ia.n_affected + kSpreadPerBlock - 1) / kSpreadPerBlock),
                           dim3(kIbmThreads), 0, c->stream, b->ia, b->ctrl, b->fbody);
    }
    if (b->f32) launch_batch_stages<float>(b, c->dm32);
    else launch_batch_stages<double>(b, c->dm);
    CaseCtl cc{b->done, b->status, b->steps, b->conv, b->res};
    hipLaunchKernelGGL(kb_finalize, dim3(b->B), dim3(kFinThreads), 0, c->stream, c->L.n_blocks, b->B,
                       (double)b->N, c->prm.ref_rho, c->prm.ref_mom, c->prm.tol, b->ctrl, cc, b->partials);
    hipLaunchKernelGGL(kb_alldone, dim3(1), dim3(kFinThreads), 0, c->stream, b->B, b->ctrl, b->done);
    HIP_TRY(hipGetLastError());
    return PLBFS_OK;
}

int build_batch_graph(plbfs_batch *b) {
    if (b->graph_ok) return PLBFS_OK;
    hipStream_t s = b->base->stream;
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = launch_batch_step(b);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    if (e != hipSuccess) return plbfs_fail(PLBFS_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return plbfs_fail(PLBFS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    b->graph_ok = true;
    return PLBFS_OK;
}

// Clears step count, convergence, status and residual of case case_id (every
// case when case_id < 0).  The other cases keep theirs: get_state reads a
// case's step count to know whether its t0 holds primitives or conservatives.
int reset_case_ctl(plbfs_batch *b, int case_id = -1) {
    const size_t k0 = case_id < 0 ? 0 : (size_t)case_id;
    const size_t n = case_id < 0 ? (size_t)b->B : 1;
    HIP_TRY(hipMemset(b->done + k0, 0, n * sizeof(int32_t)));
    HIP_TRY(hipMemset(b->status + k0, 0, n * sizeof(int32_t)));
    HIP_TRY(hipMemset(b->steps + k0, 0, n * sizeof(int32_t)));
    HIP_TRY(hipMemset(b->conv + k0, 0, n * sizeof(int32_t)));
    HIP_TRY(hipMemset(b->res + 4 * k0, 0, n * 4 * sizeof(double)));
    return PLBFS_OK;
}

// Runs up to max_steps steps of every live case; ms_out gets the device time.
int run_steps(plbfs_batch *b, int max_steps, float *ms_out) {
    plbfs_ctx *c = b->base;
    int rc;
    if (b->ibm_dirty && (rc = ibm_upload(b))) return rc;
    Ctrl h{};
    HIP_TRY(hipMemcpyAsync(b->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
    const int64_t NR = (int64_t)b->N * b->B;
    const dim3 rgrid((unsigned)((NR + kBlock - 1) / kBlock));
    if (b->f32)
        hipLaunchKernelGGL(kb_init_dt<float>, rgrid, dim3(kBlock), 0, c->stream, c->dm32, c->prm, args_of(b),
                           (const float4 *)b->stA, b->dtl);
    else
        hipLaunchKernelGGL(kb_init_dt<double>, rgrid, dim3(kBlock), 0, c->stream, c->dm, c->prm, args_of(b),
                           (const double4 *)b->stA, b->dtl);
    HIP_TRY(hipGetLastError());
    if ((rc = build_batch_graph(b))) return rc;
    const int interval = std::max(1, b->params.check_interval);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int s = 0; s < max_steps; s++) {
        HIP_TRY(hipGraphLaunch(b->graph, c->stream));
        if ((s + 1) % interval == 0 || s == max_steps - 1) {
            HIP_TRY(hipMemcpyAsync(&h, b->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            if (h.done) break;
        }
    }
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_out) *ms_out = ms;
    return PLBFS_OK;
}

}  // namespace

extern "C" int plbfs_batch_create(const plbfs_mesh_view *mesh, const plbfs_params *params,
                                  const plbfs_patch *patches, int32_t n_patches, const int32_t *bface_patch,
                                  int32_t n_cases, const double *case_visc, plbfs_batch **out) {
    if (!mesh || !params || !out || !case_visc || n_cases <= 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    *out = nullptr;
    if (params->time_stepping != PLBFS_DT_LOCAL && params->time_stepping != PLBFS_DT_CONSTANT)
        return plbfs_fail(PLBFS_ERR_INPUT, "batched cases support local and constant time stepping");
    for (int c = 0; c < n_cases; c++)
        if (!(case_visc[c] > 0)) return plbfs_fail(PLBFS_ERR_INPUT, "case %d: viscosity must be positive", c);
    for (int q = 0; q < n_patches; q++)
        if (rho_keeps_ghost(patches[q].rho_type) || vel_keeps_ghost(patches[q].vel_type))
            return plbfs_fail(PLBFS_ERR_BCTYPE,
                              "patch %d: types that keep the ghost's initial state (density 4, 5, 9, velocity "
                              "4, 5, 7, 9) need a per-case ghost state, which batches do not hold", q);
    auto *b = new plbfs_batch();
    const int N = mesh->n_cells, F = mesh->n_faces, NI = mesh->n_internal_faces, T = N + F - NI;
    b->N = N;
    b->F = F;
    b->NI = NI;
    b->T = T;
    b->owner.assign(mesh->owner, mesh->owner + F);
    b->nbr.assign(mesh->neighbour, mesh->neighbour + F);
    b->cc.assign(mesh->cell_centroid, mesh->cell_centroid + 3 * (size_t)T);
    b->vol.assign(mesh->cell_volume, mesh->cell_volume + N);
    b->fc.assign(mesh->face_centroid, mesh->face_centroid + 3 * (size_t)F);
    b->fn.assign(mesh->face_normal, mesh->face_normal + 3 * (size_t)F);
    b->fa.assign(mesh->face_area, mesh->face_area + F);
    b->fdt.assign(mesh->face_dt, mesh->face_dt + F);
    if (params->fneq_min != 0.0) {
        // the cap 3 nu / (gamma fneq_min) differs per case: keep the geometric
        // streaming dt here (the host mesh formula) and cap it per case on the device
        for (int f = 0; f < F; f++) b->fdt[f] = 0.5 * std::sqrt(b->fa[f]);
    }
    if (bface_patch) b->bpatch.assign(bface_patch, bface_patch + (F - NI));
    b->params = *params;
    b->patches.assign(patches, patches + n_patches);
    b->B = n_cases;
    b->visc.assign(case_visc, case_visc + n_cases);
    b->f32 = params->precision == PLBFS_FP32;
    b->ibm_case.resize(n_cases);
    if (const char *sp = std::getenv("PLBFS_BATCH_SPLIT_BORDER")) b->split = std::atoi(sp) != 0;
    auto fail = [&](int rc) {
        plbfs_batch_destroy(b);
        return rc;
    };
    int rc = retile(b, 256, 1, 0);
    if (rc) return fail(rc);
    const size_t NR = (size_t)N * n_cases;
    const size_t sb = b->f32 ? 4 : 8;
    std::vector<double> v3g(n_cases);
    std::vector<double> dtcap(n_cases, INFINITY);
    for (int c = 0; c < n_cases; c++) {
        v3g[c] = case_visc[c] * 3 / params->gamma;
        if (params->fneq_min != 0.0) dtcap[c] = case_visc[c] * 3 / params->gamma / params->fneq_min;
    }
    if ((rc = balloc(b, &b->stA, NR * 4 * sb)) || (rc = balloc(b, &b->stB, NR * 4 * sb)) ||
        (rc = balloc(b, &b->grad, NR * 12 * sb)) ||
        (rc = balloc(b, &b->t1, NR * 4 * sb)) ||
        (rc = balloc(b, (void **)&b->t0, NR * 4 * sizeof(double))) ||
        (rc = balloc(b, (void **)&b->dtl, NR * sizeof(double))) ||
        (rc = balloc(b, (void **)&b->fbody, NR * 3 * sizeof(double))) ||
        (rc = balloc(b, (void **)&b->scratch, (size_t)N * 4 * sizeof(double))) ||
        (rc = bupload(b, &b->d_visc, b->visc)) || (rc = bupload(b, &b->d_v3g, v3g)) ||
        (rc = bupload(b, &b->d_dtcap, dtcap)) ||
        (rc = balloc(b, (void **)&b->done, n_cases * sizeof(int32_t))) ||
        (rc = balloc(b, (void **)&b->status, n_cases * sizeof(int32_t))) ||
        (rc = balloc(b, (void **)&b->steps, n_cases * sizeof(int32_t))) ||
        (rc = balloc(b, (void **)&b->conv, n_cases * sizeof(int32_t))) ||
        (rc = balloc(b, (void **)&b->res, n_cases * 4 * sizeof(double))) ||
        (rc = balloc(b, (void **)&b->ctrl, sizeof(Ctrl))))
        return fail(rc);
    if (hipMemset(b->fbody, 0, NR * 3 * sizeof(double)) != hipSuccess || reset_case_ctl(b))
        return fail(plbfs_fail(PLBFS_ERR_HIP, "batch initialisation"));
    *out = b;
    return PLBFS_OK;
}

extern "C" int plbfs_batch_destroy(plbfs_batch *b) {
    if (!b) return PLBFS_OK;
    drop_graph(b);
    if (b->base) {
        hipStreamSynchronize(b->base->stream);
        plbfs_destroy(b->base);
    }
    free_list(b->ibm_allocs);
    free_list(b->allocs);
    delete b;
    return PLBFS_OK;
}

extern "C" int plbfs_batch_set_state(plbfs_batch *b, int32_t case_id, const double *rho, const double *u,
                                     const double *v, const double *w) {
    if (!b || !rho || !u || !v || !w || case_id >= b->B) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    const int N = b->N;
    const std::vector<int32_t> &perm = b->base->L.perm;
    std::vector<double> h((size_t)4 * N);
    for (int n = 0; n < N; n++) {
        const int i = perm[n];
        h[(size_t)4 * n] = rho[i];
        h[(size_t)4 * n + 1] = u[i];
        h[(size_t)4 * n + 2] = v[i];
        h[(size_t)4 * n + 3] = w[i];
    }
    hipStream_t s = b->base->stream;
    HIP_TRY(hipMemcpyAsync(b->scratch, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    const int64_t NR = (int64_t)N * b->B;
    const dim3 grid((unsigned)((NR + kBlock - 1) / kBlock));
    const double r0 = b->f32 ? b->params.ref_rho : 0.0;
    if (b->f32)
        hipLaunchKernelGGL(kb_scatter<float>, grid, dim3(kBlock), 0, s, N, b->B, (int)case_id, r0,
                           (const double4 *)b->scratch, (double4 *)b->t0, (float4 *)b->stA);
    else
        hipLaunchKernelGGL(kb_scatter<double>, grid, dim3(kBlock), 0, s, N, b->B, (int)case_id, r0,
                           (const double4 *)b->scratch, (double4 *)b->t0, (double4 *)b->stA);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return reset_case_ctl(b, case_id);
}

extern "C" int plbfs_batch_get_state(plbfs_batch *b, int32_t case_id, double *rho, double *u, double *v,
                                     double *w) {
    if (!b || case_id < 0 || case_id >= b->B) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    const int N = b->N;
    hipStream_t s = b->base->stream;
    // the conservative state t0 holds the solution in both precisions
    hipLaunchKernelGGL(kb_gather, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, N, b->B, (int)case_id,
                       (const double4 *)b->t0, (double4 *)b->scratch);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)4 * N);
    HIP_TRY(hipMemcpyAsync(h.data(), b->scratch, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int32_t steps = 0;
    HIP_TRY(hipMemcpy(&steps, b->steps + case_id, sizeof(int32_t), hipMemcpyDeviceToHost));
    const std::vector<int32_t> &perm = b->base->L.perm;
    for (int n = 0; n < N; n++) {
        const double *q = &h[(size_t)4 * n];
        const int i = perm[n];
        // before any step t0 holds the initial primitives, afterwards conservatives
        const double r = q[0];
        if (rho) rho[i] = r;
        if (u) u[i] = steps > 0 ? q[1] / r : q[1];
        if (v) v[i] = steps > 0 ? q[2] / r : q[2];
        if (w) w[i] = steps > 0 ? q[3] / r : q[3];
    }
    return PLBFS_OK;
}

extern "C" int plbfs_batch_set_ibm(plbfs_batch *b, int32_t case_id, const plbfs_ibm_nodes *nodes) {
    if (!b || !nodes || case_id < 0 || case_id >= b->B) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    bool any = false;
    for (const auto &ic : b->ibm_case) any = any || ic.on;
    if (any && (nodes->kernel != b->ibm_kernel || nodes->h != b->ibm_h || nodes->dt != b->ibm_dt))
        return plbfs_fail(PLBFS_ERR_INPUT, "all cases of a batch share the delta kernel, its width and dt");
    BatchIbmCase ic;
    int rc = build_ibm_pairs(b->base->L, nodes, ic.pairs);
    if (rc) return rc;
    const int n = nodes->n_nodes;
    ic.on = true;
    ic.xref.assign(nodes->x_ref, nodes->x_ref + 3 * (size_t)n);
    ic.x.assign(nodes->x, nodes->x + 3 * (size_t)n);
    ic.kdA = nodes->stiffness * nodes->node_area;
    b->ibm_case[case_id] = std::move(ic);
    b->ibm_kernel = nodes->kernel;
    b->ibm_h = nodes->h;
    b->ibm_dt = nodes->dt;
    b->ibm_margin = b->ibm_case[case_id].pairs.margin;
    b->ibm_dirty = true;
    drop_graph(b);
    return PLBFS_OK;
}

extern "C" int plbfs_batch_run(plbfs_batch *b, int32_t max_steps, plbfs_result *results) {
    if (!b || max_steps < 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    float ms = 0;
    int rc = run_steps(b, max_steps, &ms);
    if (rc) return rc;
    const int B = b->B;
    std::vector<int32_t> st(B), steps(B), conv(B);
    std::vector<double> res((size_t)4 * B);
    HIP_TRY(hipMemcpy(st.data(), b->status, B * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(steps.data(), b->steps, B * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(conv.data(), b->conv, B * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(res.data(), b->res, res.size() * sizeof(double), hipMemcpyDeviceToHost));
    int worst = PLBFS_OK;
    for (int c = 0; c < B; c++) {
        if (results) {
            plbfs_result &r = results[c];
            r.status = st[c];
            r.steps = steps[c];
            r.converged = conv[c];
            r.total_steps = steps[c];
            for (int q = 0; q < 4; q++) r.residual[q] = res[(size_t)4 * c + q];
            r.solve_ms = ms;
        }
        if (st[c] != PLBFS_OK && worst == PLBFS_OK) worst = st[c];
    }
    return worst;
}

extern "C" int plbfs_batch_set_tiling(plbfs_batch *b, int32_t block_cells, int32_t cases_per_block, int32_t stage) {
    if (!b) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    const bool ibm = std::any_of(b->ibm_case.begin(), b->ibm_case.end(), [](const BatchIbmCase &c) { return c.on; });
    int rc = retile(b, block_cells, cases_per_block, stage ? 1 : 0);
    if (rc) return rc;
    if (ibm) b->ibm_dirty = true;  // the centroid and volume arrays belong to the new layout
    return PLBFS_OK;
}

extern "C" int plbfs_batch_autotune(plbfs_batch *b, int32_t trial_steps, int32_t *block_cells,
                                    int32_t *cases_per_block, int32_t *stage, double *ms_per_step) {
    if (!b || trial_steps <= 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    // snapshot what a trial changes: state, RK accumulator start, node positions
    const size_t NR = (size_t)b->N * b->B, sb = b->f32 ? 4 : 8;
    int rc;
    if (b->ibm_dirty && (rc = ibm_upload(b))) return rc;
    std::vector<char> s_st(NR * 4 * sb), s_t0(NR * 4 * sizeof(double));
    HIP_TRY(hipMemcpy(s_st.data(), b->stA, s_st.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(s_t0.data(), b->t0, s_t0.size(), hipMemcpyDeviceToHost));
    std::vector<std::vector<BatchIbmCase>> keep_ibm(1, b->ibm_case);
    const int cand_cells[] = {64, 128, 256, 512};
    const int cand_cases[] = {1, 2, 4, 8, 16, 32};
    double best = 1e300;
    int bc_best = b->block_cells, c_best = b->C, s_best = b->stage;
    const bool ibm = std::any_of(b->ibm_case.begin(), b->ibm_case.end(), [](const BatchIbmCase &c) { return c.on; });
    for (int bc : cand_cells) {
        for (int C : cand_cases) {
            for (int stage = 0; stage < 2; stage++) {
                if (C > b->B || bc * C < kBlock) continue;  // at least one (cell, case) per thread
                if ((rc = retile(b, bc, C, stage))) {
                    if (rc == PLBFS_ERR_INPUT) continue;  // does not fit LDS
                    return rc;
                }
                if (ibm && (rc = ibm_upload(b))) return rc;
                float ms = 0;
                if ((rc = run_steps(b, 1, nullptr)) || (rc = run_steps(b, trial_steps, &ms))) return rc;
                const double per = ms / trial_steps;
                if (per < best) {
                    best = per;
                    bc_best = bc;
                    c_best = C;
                    s_best = stage;
                }
                // restore
                HIP_TRY(hipMemcpy(b->stA, s_st.data(), s_st.size(), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(b->t0, s_t0.data(), s_t0.size(), hipMemcpyHostToDevice));
                if ((rc = reset_case_ctl(b))) return rc;
            }
        }
    }
    if ((rc = retile(b, bc_best, c_best, s_best))) return rc;
    b->ibm_case = keep_ibm[0];
    b->ibm_dirty = ibm;
    if (block_cells) *block_cells = bc_best;
    if (cases_per_block) *cases_per_block = c_best;
    if (stage) *stage = s_best;
    if (ms_per_step) *ms_per_step = best;
    return PLBFS_OK;
}

extern "C" int plbfs_batch_layout_info(plbfs_batch *b, int64_t *info5) {
    if (!b || !info5) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    if (!b->base) return plbfs_fail(PLBFS_ERR_INPUT, "batch has no layout");
    info5[0] = b->B;
    info5[1] = b->block_cells;
    info5[2] = b->C;
    info5[3] = (int64_t)b->base->lds_bytes;
    info5[4] = b->stage;
    return PLBFS_OK;
}
