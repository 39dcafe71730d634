// PLBFS solver on one MI355X (gfx950): kernels, context and the C-ABI of
// include/plbfs_hip.h.
//
// Per RK stage two kernels run back to back on the context's stream:
//   k_grad         one thread per cell: least-squares gradients of rho,u,v,w
//                  from the folded stencil weights (src/gradients.cpp:319-389);
//                  ghost neighbours are evaluated from the owner and its patch.
//   k_flux_update  one block per 256 Hilbert-consecutive cells: every face of the
//                  block is evaluated once into an LDS slot (border faces by both
//                  blocks), then each cell gathers its faces' slots in the
//                  reference face order and applies the RK stage update
//                  (src/Solver.cpp:286-438).  The last stage also writes the new
//                  solution, the next local time step (src/Solver.cpp:1693-1776)
//                  and per-block residual partial sums.
// After the fourth stage k_finalize sums the partials in a fixed order
// (src/residuals.cpp:35-44,139-147) and sets the stop flag on convergence,
// NaN or rho blow-up; every kernel returns immediately once the flag is set.
// No float atomics are used anywhere, so runs are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "stage.hpp"

using namespace plbfs;

namespace {

constexpr int kFinalizeThreads = 256;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Byte offset of the staged records behind the face slots in a flux block's LDS.

// Face-item record loader of flux block blk; with staging, first copies the
// block's own cell records into LDS behind the face slots (then barriers).
template <typename S>
__device__ __forceinline__ SideLoader<S, Rec1> block_loader(const DevMesh<S> &m, const V4<S> *st, const V4<S> *grad,
                                                             double *lds, int blk) {
    SideLoader<S, Rec1> ld;
    ld.st = st;
    ld.grad = grad;
    ld.cgeo = m.cgeo;
    ld.c0 = blk * m.block_cells;
    ld.nc = 0;
    ld.C = 1;
    ld.cl = 0;
    ld.sRec = nullptr;
    ld.sCg = nullptr;
    if (m.stage) {
        char *base = reinterpret_cast<char *>(lds) + staged_offset<S>(m.max_items);
        ld.nc = min(m.N, ld.c0 + m.block_cells) - ld.c0;
        V4<S> *sRec = reinterpret_cast<V4<S> *>(base);
        double4 *sCg = std::is_same<S, double>::value
                           ? reinterpret_cast<double4 *>(base + (size_t)m.block_cells * 4 * sizeof(V4<S>))
                           : nullptr;
        stage_records<S>(st, grad, std::is_same<S, double>::value ? m.cgeo : nullptr, 1, 0, 1, ld.c0, ld.nc, sRec,
                         sCg);
        ld.sRec = sRec;
        ld.sCg = sCg;
        __syncthreads();
    }
    return ld;
}

// Item index records (owner, neighbour, side bits, 0).
static std::vector<int4> item_idx(const Layout &L) {
    std::vector<int4> v((size_t)L.n_items);
    for (size_t e = 0; e < v.size(); e++) v[e] = int4{L.it_own[e], L.it_nbr[e], (int)L.it_side[e], 0};
    return v;
}

// Occupancy 8 waves/SIMD (<= 64 VGPRs) keeps enough gathers in flight; the
// compiler otherwise settles at 92 VGPRs (5 waves) for the FP64 instance.
// There the limit costs a 28 B per lane spill; 7 waves (69 VGPRs, no spill)
// measured the same time.
template <typename S>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_grad(DevMesh<S> m, const Ctrl *__restrict__ ctrl,
                                                 const V4<S> *__restrict__ st, V4<S> *__restrict__ grad) {
    if (ctrl->done) return;
    const int c0 = xcd_block(blockIdx.x, gridDim.x, m.xcd_remap) * kBlock;
    const int i = c0 + threadIdx.x;
    S g[12];
    V4<S> si;
    const int nc = min(m.N - c0, kBlock);
    const bool live = i < m.N;
    if (m.grad_stage) {
        // the block's own states in LDS: most stencil neighbours are in the block
        __shared__ V4<S> sSt[kBlock];
        if ((int)threadIdx.x < nc) sSt[threadIdx.x] = st[i];
        __syncthreads();
        if (live) {
            lsq_grad<S>(m, i, StagedStates<S>{st, sSt, c0, nc}, g);
            si = sSt[threadIdx.x];
        }
    } else if (live) {
        lsq_grad<S>(m, i, GlobalStates<S, Rec1>{st, Rec1{}}, g);
        si = st[i];
    }
    if constexpr (kGradOff<S> == 1) {
        // 64 B records (state + gradient): assembled in LDS, then stored by the
        // block as one contiguous run, lane by lane
        __shared__ V4<S> sOut[kBlock * kGradRec<S>];
        if (live) {
            V4<S> *o = sOut + kGradRec<S> * threadIdx.x;
            o[0] = si;
            o[1] = V4<S>{g[0], g[1], g[2], g[3]};
            o[2] = V4<S>{g[4], g[5], g[6], g[7]};
            o[3] = V4<S>{g[8], g[9], g[10], g[11]};
        }
        __syncthreads();
        V4<S> *out = grad + (size_t)kGradRec<S> * c0;
        for (int k = threadIdx.x; k < kGradRec<S> * nc; k += kBlock) out[k] = sOut[k];
    } else {
        if (!live) return;
        V4<S> *gi = grad + (size_t)kGradRec<S> * i;
        gi[0] = V4<S>{g[0], g[1], g[2], g[3]};
        gi[1] = V4<S>{g[4], g[5], g[6], g[7]};
        gi[2] = V4<S>{g[8], g[9], g[10], g[11]};
    }
}

// Faces between two flux blocks, one thread per face: flux into the border
// array that both blocks' cells read in their gather.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_flux_border(DevMesh<S> m, DevParams prm, const Ctrl *__restrict__ ctrl,
                                                        const V4<S> *__restrict__ st_in,
                                                        const V4<S> *__restrict__ grad) {
    if (ctrl->done) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= m.n_border) return;
    const SideLoader<S, Rec1> ld{st_in, grad, m.cgeo, Rec1{}, nullptr, nullptr, 0, 0, 1, 0};
    S F[4];
    item_flux<S>(m, prm, m.border0 + t, ld, prm.visc3_invgam, INFINITY, F);
    const size_t nbd = (size_t)m.n_border;
#pragma unroll
    for (int q = 0; q < 4; q++) m.bflux[q * nbd + t] = F[q];
}

// FP32: 6 waves per SIMD (80 VGPRs, no spill) instead of the compiler's 5
// (87 VGPRs) took the 4.1M-tet flux from 0.310 to 0.301 ms.  FP64 is left to
// the compiler (143 VGPRs, 3 waves): forcing 4 waves spills 76 B per lane and
// measured 0.684 ms against 0.543.
template <typename S>
constexpr int kFluxMinWaves = std::is_same<S, float>::value ? 6 : 1;

template <typename S>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kFluxMinWaves<S>))) void k_flux_update(
    DevMesh<S> m, DevParams prm, StageArgs sa, const Ctrl *__restrict__ ctrl, const V4<S> *__restrict__ st_in,
    const V4<S> *__restrict__ grad, V4<S> *__restrict__ st_out, double4 *__restrict__ t0, Acc<S> *__restrict__ t1,
    double *__restrict__ dtl, const double *__restrict__ fbody, double *__restrict__ partials) {
    extern __shared__ double lds[];
    if (ctrl->done) return;
    const int N = m.N;
    const int blk = xcd_block(blockIdx.x, gridDim.x, m.xcd_remap);
    const int e0 = m.blk_item_off[blk], e1 = m.blk_item_off[blk + 1];
    const int nslot = m.max_items;
    S *sF = reinterpret_cast<S *>(lds);
    const SideLoader<S, Rec1> ld = block_loader<S>(m, st_in, grad, lds, blk);
    for (int e = e0 + threadIdx.x; e < e1; e += kBlock) {
        S F[4];
        item_flux<S>(m, prm, e, ld, prm.visc3_invgam, INFINITY, F);
        const int sl = e - e0;
#pragma unroll
        for (int q = 0; q < 4; q++) sF[q * nslot + sl] = F[q];
    }
    __syncthreads();
    double sq[4] = {0, 0, 0, 0};
    int err = 0;
    const int c_end = min(N, (blk + 1) * m.block_cells);
    for (int i = blk * m.block_cells + threadIdx.x; i < c_end; i += kBlock) {
        double R[4];
        gather_residual<S>(m, i, sF, nslot, 1, 0, R);
        if (prm.use_body_force && (!m.blk_force || m.blk_force[blk])) {
            R[1] += fbody[i];
            R[2] += fbody[(size_t)N + i];
            R[3] += fbody[2 * (size_t)N + i];
        }
        const int e = rk_update<S>(m, prm, prm.visc, sa, i, (size_t)i, R, st_out, t0, t1, dtl, sq);
        if (e == 1) err = 1;
        else if (e == 2 && err == 0) err = 2;
    }
    if (sa.stage == 3) {
        __shared__ double red[kBlock / 64][5];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        double e = (double)err;
        double r0 = wave_sum(sq[0]), r1 = wave_sum(sq[1]), r2 = wave_sum(sq[2]), r3 = wave_sum(sq[3]);
        // error code: max over the wave
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_xor(e, off, 64));
        if (lane == 0) {
            red[wv][0] = r0; red[wv][1] = r1; red[wv][2] = r2; red[wv][3] = r3; red[wv][4] = e;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double s[5] = {0, 0, 0, 0, 0};
            for (int q = 0; q < kBlock / 64; q++) {
                for (int c = 0; c < 4; c++) s[c] += red[q][c];
                s[4] = fmax(s[4], red[q][4]);
            }
            for (int c = 0; c < 5; c++) partials[(size_t)blk * 5 + c] = s[c];
        }
    }
}

// Global minimum of the local time steps for the "min" and "talts" modes
// (Solver.cpp:1746-1773): every block reduces all cells (cheap; N is read once
// per block from L2) -- done in two kernels to keep it deterministic.
__global__ __launch_bounds__(kBlock) void k_dt_blockmin(int N, const Ctrl *__restrict__ ctrl,
                                                        const double *__restrict__ dtl, double *__restrict__ bmin) {
    if (ctrl->done) return;
    __shared__ double red[kBlock / 64];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v = (i < N) ? dtl[i] : 1e300;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0];
        for (int q = 1; q < kBlock / 64; q++) s = fmin(s, red[q]);
        bmin[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_dt_apply(int N, int nblk, int mode, const Ctrl *__restrict__ ctrl,
                                                     const double *__restrict__ bmin, double *__restrict__ dtl) {
    if (ctrl->done) return;
    __shared__ double smin;
    if (threadIdx.x == 0) {
        double s = 100000000000.0;
        for (int q = 0; q < nblk; q++) s = (bmin[q] < s) ? bmin[q] : s;
        smin = s;
    }
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const double mn = smin;
    if (mode == PLBFS_DT_MIN) dtl[i] = mn;
    else dtl[i] = mn * pow(2.0, floor(log2(dtl[i] / mn)));
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_init_dt(int N, DevParams prm, const Ctrl *__restrict__ ctrl,
                                                    const double *__restrict__ cfl_area,
                                                    const double *__restrict__ vol, const V4<S> *__restrict__ st,
                                                    double *__restrict__ dtl) {
    if (ctrl->done) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const double r0 = std::is_same<S, float>::value ? prm.rho0 : 0.0;
    if (prm.time_stepping == PLBFS_DT_CONSTANT) dtl[i] = prm.cfl;
    else dtl[i] = local_dt(prm, prm.visc, r0 + (double)st[i].x, (double)st[i].y, cfl_area[i], vol[i]);
}

__global__ __launch_bounds__(kFinalizeThreads) void k_finalize(int nblk, double n_cells, double ref_rho,
                                                               double ref_mom, double tol, int stop_on_rebuild,
                                                               Ctrl *__restrict__ ctrl,
                                                               const double *__restrict__ partials,
                                                               double *__restrict__ hist, int hist_cap) {
    if (ctrl->done) return;
    __shared__ double red[kFinalizeThreads][5];
    double s[5] = {0, 0, 0, 0, 0};
    // a thread's partials b = tid, tid + T, ... in that order; four at a time
    // with all their loads issued before the sums
    constexpr int kU = 4;
    for (int b0 = threadIdx.x; b0 < nblk; b0 += kFinalizeThreads * kU) {
        double v[kU][5];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int b = b0 + u * kFinalizeThreads;
            if (b < nblk)
                for (int c = 0; c < 5; c++) v[u][c] = partials[(size_t)b * 5 + c];
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            if (b0 + u * kFinalizeThreads < nblk) {
                for (int c = 0; c < 4; c++) s[c] += v[u][c];
                s[4] = fmax(s[4], v[u][4]);
            }
        }
    }
    for (int c = 0; c < 5; c++) red[threadIdx.x][c] = s[c];
    __syncthreads();
    for (int off = kFinalizeThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            for (int c = 0; c < 4; c++) red[threadIdx.x][c] += red[threadIdx.x + off][c];
            red[threadIdx.x][4] = fmax(red[threadIdx.x][4], red[threadIdx.x + off][4]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int step = ctrl->step;
        double r[4];
        r[0] = sqrt(red[0][0] / n_cells) / ref_rho;
        for (int c = 1; c < 4; c++) r[c] = sqrt(red[0][c] / n_cells) / ref_mom;
        for (int c = 0; c < 4; c++) ctrl->res[c] = r[c];
        if (step < hist_cap)
            for (int c = 0; c < 4; c++) hist[(size_t)step * 4 + c] = r[c];
        ctrl->step = step + 1;
        const int e = (int)red[0][4];
        if (e == 1) {
            ctrl->status = PLBFS_ERR_NAN;
            ctrl->done = 1;
        } else if (e == 2) {
            ctrl->status = PLBFS_ERR_RHO;
            ctrl->done = 1;
        } else {
            const double me = fmax(fmax(r[3], r[1]), r[2]);
            if (me < tol) {
                ctrl->converged = 1;
                ctrl->done = 1;
            }
        }
        // immersed nodes asked for new candidate lists: the host builds them
        // and continues with the next step
        if (stop_on_rebuild && ctrl->rebuild) ctrl->done = 1;
    }
}

// One residual evaluation for diagnostics: flux gather only (stage-independent).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_residual_only(DevMesh<S> m, DevParams prm,
                                                          const V4<S> *__restrict__ st_in,
                                                          const V4<S> *__restrict__ grad,
                                                          const double *__restrict__ fbody,
                                                          double *__restrict__ Rout) {
    extern __shared__ double lds[];
    const int N = m.N;
    const int blk = xcd_block(blockIdx.x, gridDim.x, m.xcd_remap);
    const int e0 = m.blk_item_off[blk], e1 = m.blk_item_off[blk + 1];
    const int nslot = m.max_items;
    S *sF = reinterpret_cast<S *>(lds);
    const SideLoader<S, Rec1> ld = block_loader<S>(m, st_in, grad, lds, blk);
    for (int e = e0 + threadIdx.x; e < e1; e += kBlock) {
        S F[4];
        item_flux<S>(m, prm, e, ld, prm.visc3_invgam, INFINITY, F);
#pragma unroll
        for (int q = 0; q < 4; q++) sF[q * nslot + (e - e0)] = F[q];
    }
    __syncthreads();
    const int c_end = min(N, (blk + 1) * m.block_cells);
    for (int i = blk * m.block_cells + threadIdx.x; i < c_end; i += kBlock) {
        double R[4];
        gather_residual<S>(m, i, sF, nslot, 1, 0, R);
        if (prm.use_body_force) {
            R[1] += fbody[i];
            R[2] += fbody[(size_t)N + i];
            R[3] += fbody[2 * (size_t)N + i];
        }
        for (int q = 0; q < 4; q++) Rout[(size_t)q * N + i] = R[q];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

// ---- the time loop without allocation --------------------------------------
// plbfs_run_noalloc works in this array instead of the context's control block,
// partials, dt minima and history; it allocates nothing, builds no graph and
// creates no events.  Each device holds its own copy until the process ends;
// kScratchBytes bounds what one call may use.  Calls that overlap on one
// device share the array and race.
constexpr size_t kScratchBytes = (size_t)8 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxScratchDevices];

// This device's address of the scratch array; called by create_ctx with the
// context's device current, once per device.
static int resolve_scratch(int device, char **out) {
    if (device < 0 || device >= kMaxScratchDevices)
        return plbfs_fail(PLBFS_ERR_HIP, "device %d: scratch table holds %d devices", device, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return plbfs_fail(PLBFS_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[device];
    return PLBFS_OK;
}

static int release(plbfs_ctx *c) {
    if (c->graph_exec) {
        hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    for (void *p : c->allocs) hipFree(p);
    c->allocs.clear();
    for (void *p : c->ibm_allocs) hipFree(p);
    c->ibm_allocs.clear();
    if (c->stream) {
        hipStreamDestroy(c->stream);
        c->stream = nullptr;
    }
    return PLBFS_OK;
}

extern "C" int plbfs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static bool env_flag(const char *name, bool def) {
    const char *v = std::getenv(name);
    if (!v || !v[0]) return def;
    return v[0] != '0';
}

extern "C" int plbfs_create(const plbfs_mesh_view *mesh, const plbfs_params *params, const plbfs_patch *patches,
                            int32_t n_patches, const int32_t *bface_patch, plbfs_ctx **out) {
    // cells per flux block (a multiple of 64): larger blocks duplicate fewer
    // border faces but need more LDS
    int block_cells = kFluxBlockCells;
    if (const char *bc = std::getenv("PLBFS_BLOCK_CELLS")) block_cells = std::atoi(bc);
    if (block_cells < 64 || block_cells % 64 || block_cells > 4 * kBlock)
        return plbfs_fail(PLBFS_ERR_INPUT, "PLBFS_BLOCK_CELLS must be a multiple of 64 up to %d", 4 * kBlock);
    // measured on MI355X (4.09M tets): staging the flux block's own records in
    // LDS costs more (staging phase + barrier) than the L2 traffic it saves,
    // while staging the gradient kernel's 256 states saves 12 %; flux staging
    // is therefore opt-in (PLBFS_FLUX_STAGE=1) and gradient staging opt-out
    const bool flux_stage = env_flag("PLBFS_FLUX_STAGE", false);
    // faces between two blocks: both blocks evaluate them, or, with
    // PLBFS_SPLIT_BORDER=1, a separate border pass evaluates each once.
    // Measured on MI355X (4.09M tets): the border pass makes the flux step
    // slower (FP64 0.555 -> 0.617 ms, FP32 0.331 -> 0.364 ms; its faces
    // gather both sides from global memory), so duplication is the default
    const bool split = env_flag("PLBFS_SPLIT_BORDER", false);
    return create_ctx(mesh, params, patches, n_patches, bface_patch, block_cells, 1, flux_stage, split, out);
}

// Fixed ghost states to the device: (rho, u, v, w) per boundary face in the
// storage precision (FP32: rho as the deviation from rho0).
static int upload_ghost_fixed(plbfs_ctx *c) {
    const size_t G = (size_t)c->L.G;
    if (G == 0) return PLBFS_OK;
    if (!c->f32) {
        HIP_TRY(hipMemcpy(c->b_fix, c->ghost_fixed.data(), G * 4 * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> h(G * 4);
        for (size_t k = 0; k < G * 4; k++)
            h[k] = (float)(c->ghost_fixed[k] - ((k & 3) == 0 ? c->prm.rho0 : 0.0));
        HIP_TRY(hipMemcpy(c->b_fix, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return PLBFS_OK;
}

int create_ctx(const plbfs_mesh_view *mesh, const plbfs_params *params, const plbfs_patch *patches,
               int32_t n_patches, const int32_t *bface_patch, int block_cells, int lds_cases, bool flux_stage,
               bool split_border, plbfs_ctx **out) {
    if (!mesh || !params || !out) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    *out = nullptr;
    const plbfs_params &p = *params;
    if (!(p.visc > 0) || !(p.gamma > 0) || !(p.cfl > 0) || !(p.max_velocity > 0) || !(p.ref_rho > 0))
        return plbfs_fail(PLBFS_ERR_INPUT, "visc, gamma, cfl, max_velocity and ref_rho must be positive");
    if (p.time_stepping < 0 || p.time_stepping > 3 || p.gradient_type < 0 || p.gradient_type > 2)
        return plbfs_fail(PLBFS_ERR_INPUT, "bad time_stepping or gradient_type");
    if (p.precision != PLBFS_FP64 && p.precision != PLBFS_FP32)
        return plbfs_fail(PLBFS_ERR_INPUT, "bad precision %d", p.precision);
    if (plbfs_device_count() <= 0) return plbfs_fail(PLBFS_ERR_NODEVICE, "no HIP device visible");
    auto *c = new plbfs_ctx();
    c->params = p;
    c->f32 = p.precision == PLBFS_FP32;
    c->patches.assign(patches, patches + n_patches);
    int rc = build_layout(*mesh, p, patches, n_patches, bface_patch, block_cells, split_border, c->L);
    if (rc) {
        delete c;
        return rc;
    }
    const Layout &L = c->L;
    const size_t sbytes = c->f32 ? sizeof(float) : sizeof(double);
    const bool stage = flux_stage;
    c->lds_bytes = lds_cases == 1
                       ? (c->f32 ? flux_lds_bytes<float>(L.max_items_per_block, block_cells, stage)
                                 : flux_lds_bytes<double>(L.max_items_per_block, block_cells, stage))
                       : (size_t)L.max_items_per_block * 4 * sbytes * lds_cases;
    if (c->lds_bytes > kMaxLds) {
        // L refers into c: read what the message needs before c is freed
        const size_t need = c->lds_bytes;
        const int slots = L.max_items_per_block;
        delete c;
        return plbfs_fail(PLBFS_ERR_INPUT, "a block needs %zu B of LDS for %d face slots x %d cases (limit %d B)",
                          need, slots, lds_cases, kMaxLds);
    }
    auto fail = [&](int r) {
        release(c);
        delete c;
        return r;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return fail(plbfs_fail(PLBFS_ERR_HIP, "hipGetDevice"));
    if ((rc = resolve_scratch(c->device, &c->scratch))) return fail(rc);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(plbfs_fail(PLBFS_ERR_HIP, "hipStreamCreate"));
    const int N = L.N;
    DevMesh<double> &d = c->dm;
    d.N = N;
    d.G = L.G;
    d.n_blocks = L.n_blocks;
    d.max_items = L.max_items_per_block;
    d.E = L.n_items;
    d.border0 = L.border0;
    d.n_border = (int)L.n_border;
    d.bflux = nullptr;
    d.blk_force = nullptr;
    d.block_cells = L.block_cells;
    d.stage = stage ? 1 : 0;
    d.grad_stage = env_flag("PLBFS_GRAD_STAGE", true) ? 1 : 0;
    d.g0x = L.ghost0[0];
    d.g0y = L.ghost0[1];
    d.g0z = L.ghost0[2];
    c->cell_grid = (N + kBlock - 1) / kBlock;
    // measured on MI355X (4.09M tets): the XCD-contiguous remap is ~5% slower
    // than plain dispatch order for the flux kernel, so it is opt-in
    const char *nx = std::getenv("PLBFS_XCD_REMAP");
    d.xcd_remap = (nx && nx[0] == '1') ? 1 : 0;
    std::vector<DevPatch> dp(n_patches);
    for (int q = 0; q < n_patches; q++) {
        dp[q].rho_type = patches[q].rho_type;
        dp[q].vel_type = patches[q].vel_type;
        dp[q].is_empty = patches[q].is_empty;
        const bool keeps = rho_keeps_ghost(patches[q].rho_type) || vel_keeps_ghost(patches[q].vel_type);
        dp[q].flags = (patches[q].rho_type == 3 || patches[q].vel_type == 3 || keeps ? kGhostGradFromCell : 0) |
                      (keeps ? kGhostFixed : 0);
        dp[q].rho = patches[q].rho;
        dp[q].u = patches[q].u;
        dp[q].v = patches[q].v;
        dp[q].w = patches[q].w;
    }
    std::vector<int32_t> bpatch = L.b_patch;
    if (bpatch.empty()) bpatch.push_back(0);
    std::vector<int32_t> bpartner = L.b_partner;
    if (bpartner.empty()) bpartner.push_back(-1);
    std::vector<double> bdmag = L.b_dmag;
    if (bdmag.empty()) bdmag.push_back(1.0);
    c->ghost_fixed = L.b_fixed;
    c->ghost_state_set = !L.needs_ghost_state;
    if ((rc = dalloc(c, (void **)&c->b_fix, (size_t)std::max(L.G, 1) * 4 * sbytes))) return fail(rc);
    d.b_fix = c->f32 ? nullptr : (const double4 *)c->b_fix;
    d.corr = nullptr;
    if ((rc = upload(c, (uint8_t **)&d.st_cnt, L.st_cnt)) || (rc = upload(c, (int32_t **)&d.st_nbr, L.st_nbr)) ||
        (rc = upload(c, (double **)&d.vol, L.vol)) || (rc = upload(c, (double **)&d.inv_vol, L.inv_vol)) ||
        (rc = upload(c, (double **)&d.cfl_area, L.cfl_area)) ||
        (rc = upload(c, (int32_t **)&d.blk_item_off, L.blk_item_off)) ||
        (rc = upload(c, (int4 **)&d.it_x, item_idx(L))) ||
        (rc = upload(c, (int32_t **)&d.g_off, L.g_off)) || (rc = upload(c, (int32_t **)&d.g_ent, L.g_ent)) ||
        (rc = upload(c, (int32_t **)&d.b_patch, bpatch)) || (rc = upload(c, (int32_t **)&d.b_partner, bpartner)) ||
        (rc = upload(c, (DevPatch **)&d.patches, dp)))
        return fail(rc);
    if (!c->f32) {
        // n, A, dt only; offsets come from the face and cell centroids
        const size_t E = (size_t)L.n_items;
        std::vector<double4> geo(2 * E);
        for (size_t e = 0; e < E; e++) {
            const double *g = L.it_geo.data();
            geo[2 * e] = double4{g[e], g[E + e], g[2 * E + e], g[3 * E + e]};
            geo[2 * e + 1] = double4{g[4 * E + e], L.it_fc[e], L.it_fc[E + e], L.it_fc[2 * E + e]};
        }
        std::vector<double> cg((size_t)4 * N);
        for (int n = 0; n < N; n++) {
            for (int a = 0; a < 3; a++) cg[(size_t)4 * n + a] = L.centroid[(size_t)a * N + n];
            cg[(size_t)4 * n + 3] = L.inv_vol[n];
        }
        if ((rc = upload(c, (double **)&d.st_w, L.st_w)) || (rc = upload(c, (double4 **)&d.it_d, geo)) ||
            (rc = upload(c, (double **)&d.cgeo, cg)) ||
            (rc = upload(c, (double **)&d.b_dmag, bdmag)))
            return fail(rc);
    } else {
        // FP32 storage: the same arrays rounded once from the FP64 precompute;
        // patch densities become deviations from rho0
        DevMesh<float> &e = c->dm32;
        e.N = d.N; e.G = d.G; e.n_blocks = d.n_blocks; e.max_items = d.max_items; e.E = d.E;
        e.border0 = d.border0; e.n_border = d.n_border; e.bflux = nullptr; e.blk_force = nullptr;
        e.xcd_remap = d.xcd_remap;
        e.block_cells = d.block_cells;
        e.stage = d.stage;
        e.grad_stage = d.grad_stage;
        e.it_d = nullptr;
        e.cgeo = nullptr;
        e.g0x = d.g0x; e.g0y = d.g0y; e.g0z = d.g0z;
        e.st_cnt = d.st_cnt; e.st_nbr = d.st_nbr; e.vol = d.vol; e.inv_vol = d.inv_vol; e.cfl_area = d.cfl_area;
        e.blk_item_off = d.blk_item_off; e.it_x = nullptr; e.g_off = d.g_off;
        e.g_ent = d.g_ent; e.b_patch = d.b_patch; e.b_partner = d.b_partner;
        e.b_fix = (const float4 *)c->b_fix;
        auto to32 = [](const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); };
        std::vector<DevPatch> dp32 = dp;
        for (auto &q : dp32) q.rho -= p.ref_rho;
        const size_t E = (size_t)L.n_items;
        std::vector<float4> geo(3 * E);
        for (size_t it = 0; it < E; it++) {
            float g[12] = {0};
            for (int q = 0; q < kGeoPerItem; q++) g[q] = (float)L.it_geo[(size_t)q * E + it];
            geo[3 * it] = float4{g[0], g[1], g[2], g[3]};
            geo[3 * it + 1] = float4{g[4], g[5], g[6], g[7]};
            float sb;
            const uint32_t side = L.it_side[it];
            std::memcpy(&sb, &side, sizeof sb);
            geo[3 * it + 2] = float4{g[8], g[9], g[10], sb};
        }
        std::vector<int2> on(E);
        for (size_t it = 0; it < E; it++) on[it] = int2{L.it_own[it], L.it_nbr[it]};
        if ((rc = upload(c, (float **)&e.st_w, to32(L.st_w))) || (rc = upload(c, (float4 **)&e.it_f, geo)) ||
            (rc = upload(c, (int2 **)&e.it_on, on)) ||
            (rc = upload(c, (float **)&e.b_dmag, to32(bdmag))) || (rc = upload(c, (float **)&e.corr, to32(L.corr))) ||
            (rc = upload(c, (DevPatch **)&e.patches, dp32)))
            return fail(rc);
    }
    const size_t n4 = (size_t)4 * N * sbytes;
    const size_t nacc = (size_t)4 * N * (c->f32 ? sizeof(float) : sizeof(double));
    if ((rc = dalloc(c, (void **)&c->stA, n4)) || (rc = dalloc(c, (void **)&c->stB, (size_t)4 * N * sizeof(double))) ||
        (rc = dalloc(c, (void **)&c->t0, (size_t)4 * N * sizeof(double))) || (rc = dalloc(c, (void **)&c->t1, nacc)) ||
        (rc = dalloc(c, (void **)&c->grad, (size_t)4 * (c->f32 ? kGradRec<float> : kGradRec<double>) * N * sbytes)) ||
        (rc = dalloc(c, (void **)&c->dtl, (size_t)N * sizeof(double))) ||
        (rc = dalloc(c, (void **)&c->fbody, (size_t)3 * N * sizeof(double))) ||
        (rc = dalloc(c, (void **)&c->partials, (size_t)L.n_blocks * 5 * sizeof(double))) ||
        (rc = dalloc(c, (void **)&c->bmin, (size_t)c->cell_grid * sizeof(double))) ||
        (rc = dalloc(c, (void **)&c->ctrl, sizeof(Ctrl))) ||
        (rc = dalloc(c, &c->bflux, (size_t)4 * std::max<int64_t>(L.n_border, 1) * sbytes)))
        return fail(rc);
    c->dm.bflux = (double *)c->bflux;
    c->dm32.bflux = (float *)c->bflux;
    {
        std::vector<double> z((size_t)3 * N, 0.0);
        if (hipMemcpy(c->fbody, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return fail(plbfs_fail(PLBFS_ERR_HIP, "upload body force"));
        Ctrl h{};
        if (hipMemcpy(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess)
            return fail(plbfs_fail(PLBFS_ERR_HIP, "upload ctrl"));
    }
    DevParams &prm = c->prm;
    prm.visc = p.visc;
    prm.gamma = p.gamma;
    prm.invgam = 1.0 / p.gamma;
    prm.visc3_invgam = p.visc * 3 / p.gamma;
    prm.force = p.force;
    prm.cfl = p.cfl;
    prm.ceff = p.max_velocity * std::sqrt(1 - p.gamma + std::pow(p.gamma / std::sqrt(3) / p.max_velocity, 2));
    prm.ref_rho = p.ref_rho;
    prm.rho0 = p.ref_rho;
    prm.ref_mom = p.ref_rho * p.max_velocity;
    prm.tol = p.tolerance;
    prm.time_stepping = p.time_stepping;
    prm.use_body_force = 0;
    const char *g = std::getenv("PLBFS_NO_GRAPH");
    c->use_graph = !(g && g[0] == '1');
    if ((rc = upload_ghost_fixed(c))) return fail(rc);
    *out = c;
    return PLBFS_OK;
}

extern "C" int plbfs_destroy(plbfs_ctx *c) {
    if (!c) return PLBFS_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    release(c);
    delete c;
    return PLBFS_OK;
}

static int to_device_state(plbfs_ctx *c, const double *rho, const double *u, const double *v, const double *w,
                           double *dst) {
    const int N = c->L.N;
    std::vector<double> h((size_t)4 * N);  // cell records (rho,u,v,w), Hilbert order
    const double *src[4] = {rho, u, v, w};
    for (int n = 0; n < N; n++)
        for (int q = 0; q < 4; q++) h[(size_t)4 * n + q] = src[q][c->L.perm[n]];
    HIP_TRY(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return PLBFS_OK;
}

// Copies nf interleaved values per cell back to file order (one array per value).
static int from_device_fields(plbfs_ctx *c, const double *src, int nf, double *const *out) {
    const int N = c->L.N;
    std::vector<double> h((size_t)nf * N);
    HIP_TRY(hipMemcpy(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int q = 0; q < nf; q++)
        if (out[q])
            for (int n = 0; n < N; n++) out[q][c->L.perm[n]] = h[(size_t)nf * n + q];
    return PLBFS_OK;
}

extern "C" int plbfs_set_state(plbfs_ctx *c, const double *rho, const double *u, const double *v, const double *w) {
    if (!c || !rho || !u || !v || !w) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    // the conservative accumulator starts as a copy of the primitives (Solver.cpp:195)
    int rc = to_device_state(c, rho, u, v, w, c->t0);
    if (rc) return rc;
    const int N = c->L.N;
    if (!c->f32) {
        HIP_TRY(hipMemcpy(c->stA, c->t0, (size_t)4 * N * sizeof(double), hipMemcpyDeviceToDevice));
    } else {
        std::vector<float> h((size_t)4 * N);
        const double *src[4] = {rho, u, v, w};
        for (int n = 0; n < N; n++)
            for (int q = 0; q < 4; q++)
                h[(size_t)4 * n + q] = (float)(src[q][c->L.perm[n]] - (q == 0 ? c->prm.rho0 : 0.0));
        HIP_TRY(hipMemcpy(c->stA, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    c->total_steps = 0;
    return PLBFS_OK;
}

extern "C" int plbfs_set_ghost_state(plbfs_ctx *c, const double *rho, const double *u, const double *v,
                                     const double *w) {
    if (!c || !rho || !u || !v || !w) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    // only the components the reference's ghost update leaves unwritten take
    // these values (src/Solution.cpp:759-833); the parabolic u, v stay
    for (int b = 0; b < c->L.G; b++) {
        const plbfs_patch &P = c->patches[c->L.b_patch[b]];
        double *q = &c->ghost_fixed[(size_t)4 * b];
        if (rho_keeps_ghost(P.rho_type)) q[0] = rho[b];
        if (P.vel_type == 7 || P.vel_type == 9) {
            q[1] = u[b];
            q[2] = v[b];
        }
        if (vel_keeps_ghost(P.vel_type)) q[3] = w[b];
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    int rc = upload_ghost_fixed(c);
    if (rc) return rc;
    c->ghost_state_set = true;
    return PLBFS_OK;
}

// Refuses to advance a context whose patches keep initial ghost values that
// were never given.
static int check_ghost_state(const plbfs_ctx *c) {
    if (c->ghost_state_set) return PLBFS_OK;
    return plbfs_fail(PLBFS_ERR_INPUT,
                      "a patch keeps the ghost's initial state (density 4, 5, 9 or velocity 4, 5, 7, 9): "
                      "call plbfs_set_ghost_state first");
}

extern "C" int plbfs_get_state(plbfs_ctx *c, double *rho, double *u, double *v, double *w) {
    if (!c) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    double *outs[4] = {rho, u, v, w};
    if (!c->f32) return from_device_fields(c, c->stA, 4, outs);
    // FP32 storage: the FP64 conservative state t0 holds the solution (before
    // any step it still holds the initial primitives)
    const int N = c->L.N;
    std::vector<double> h((size_t)4 * N);
    HIP_TRY(hipMemcpy(h.data(), c->t0, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int n = 0; n < N; n++) {
        const double *q = &h[(size_t)4 * n];
        const int i = c->L.perm[n];
        const double r = q[0];
        const bool cons = c->total_steps > 0;
        if (rho) rho[i] = r;
        if (u) u[i] = cons ? q[1] / r : q[1];
        if (v) v[i] = cons ? q[2] / r : q[2];
        if (w) w[i] = cons ? q[3] / r : q[3];
    }
    return PLBFS_OK;
}

// ---- launches -------------------------------------------------------------

static void resolve_timing(plbfs_ctx *c) {
    for (auto &p : c->pending) {
        hipEventSynchronize(p.b);
        float ms = 0;
        hipEventElapsedTime(&ms, p.a, p.b);
        c->kstat[p.which].ms += ms;
        c->kstat[p.which].launches++;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    c->pending.clear();
}

static const double kAlpha[4] = {1.0, 0.5, 0.5, 1.0};
static const double kBeta[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// Device memory a time loop works in besides the context's fields: the control
// block, per-block residual partials, per-block dt minima and the residual
// history.  plbfs_run uses the context's own arrays, plbfs_run_noalloc carves
// them out of the file-scope scratch array.
struct RunScratch {
    Ctrl *ctrl;
    double *partials;  // [n_blocks * 5]
    double *bmin;      // [cell_grid]
    double *hist;      // [hist_cap * 4]
    int hist_cap;
};

static RunScratch ctx_scratch(plbfs_ctx *c) { return RunScratch{c->ctrl, c->partials, c->bmin, c->hist, c->hist_cap}; }

static void launch_dt_modes(plbfs_ctx *c, const RunScratch &w) {
    const int N = c->L.N, nb = c->cell_grid;
    if (c->params.time_stepping == PLBFS_DT_MIN || c->params.time_stepping == PLBFS_DT_TALTS) {
        hipLaunchKernelGGL(k_dt_blockmin, dim3(nb), dim3(kBlock), 0, c->stream, N, w.ctrl, c->dtl, w.bmin);
        hipLaunchKernelGGL(k_dt_apply, dim3(nb), dim3(kBlock), 0, c->stream, N, nb, c->params.time_stepping,
                           w.ctrl, w.bmin, c->dtl);
    }
}

template <typename S>
static void launch_stages(plbfs_ctx *c, const DevMesh<S> &dm, const RunScratch &w, bool timed) {
    const int nb = c->L.n_blocks;
    void *in = c->stA, *outb = c->stB;
    for (int k = 0; k < 4; k++) {
        {
            TimedLaunch t(c, PLBFS_KSTAT_GRAD, timed);
            hipLaunchKernelGGL(k_grad<S>, dim3(c->cell_grid), dim3(kBlock), 0, c->stream, dm, w.ctrl, (const V4<S> *)in,
                               (V4<S> *)c->grad);
        }
        StageArgs sa{kAlpha[k], kBeta[k], k};
        {
            TimedLaunch t(c, PLBFS_KSTAT_FLUX, timed);
            if (dm.n_border > 0)
                hipLaunchKernelGGL(k_flux_border<S>, dim3((dm.n_border + kBlock - 1) / kBlock), dim3(kBlock), 0,
                                   c->stream, dm, c->prm, w.ctrl, (const V4<S> *)in, (const V4<S> *)c->grad);
            hipLaunchKernelGGL(k_flux_update<S>, dim3(nb), dim3(kBlock), c->lds_bytes, c->stream, dm, c->prm, sa,
                               w.ctrl, (const V4<S> *)in, (const V4<S> *)c->grad, (V4<S> *)outb,
                               (double4 *)c->t0, (Acc<S> *)c->t1, c->dtl, c->fbody, w.partials);
        }
        std::swap(in, outb);
    }
}

template <typename S>
static void launch_init_dt(plbfs_ctx *c, const DevMesh<S> &dm, const Ctrl *ctrl) {
    hipLaunchKernelGGL(k_init_dt<S>, dim3(c->cell_grid), dim3(kBlock), 0, c->stream, c->L.N, c->prm, ctrl,
                       dm.cfl_area, dm.vol, (const V4<S> *)c->stA, c->dtl);
}

// The loop's first launches: the local time steps of the current state.
static int launch_run_start(plbfs_ctx *c, const RunScratch &w) {
    if (c->f32) launch_init_dt<float>(c, c->dm32, w.ctrl);
    else launch_init_dt<double>(c, c->dm, w.ctrl);
    launch_dt_modes(c, w);
    HIP_TRY(hipGetLastError());
    return PLBFS_OK;
}

// rebuild: immersed nodes that drift half their list margin end the step
// replay so that the host rebuilds the lists (plbfs_run); without it they run
// on to the margin (plbfs_run_noalloc, which cannot allocate new lists).
static int launch_step(plbfs_ctx *c, const RunScratch &w, bool timed, bool rebuild) {
    const int nb = c->L.n_blocks;
    int rc = ibm_launch(c, w.ctrl, timed, rebuild);  // once per step, before the stages (Solver_gpu.cu:687-718)
    if (rc) return rc;
    if (c->f32) launch_stages<float>(c, c->dm32, w, timed);
    else launch_stages<double>(c, c->dm, w, timed);
    launch_dt_modes(c, w);
    {
        TimedLaunch t(c, PLBFS_KSTAT_REDUCE, timed);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, c->stream, nb, (double)c->L.N,
                           c->prm.ref_rho, c->prm.ref_mom, c->prm.tol, rebuild ? 1 : 0, w.ctrl, w.partials,
                           w.hist, w.hist_cap);
    }
    HIP_TRY(hipGetLastError());
    return PLBFS_OK;
}

static int ensure_hist(plbfs_ctx *c, int steps) {
    if (steps <= c->hist_cap) return PLBFS_OK;
    if (c->graph_exec) {
        hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
        c->graph_ok = false;
    }
    int cap = std::max(steps, 1024);
    int rc = dalloc(c, (void **)&c->hist, (size_t)cap * 4 * sizeof(double));
    if (rc) return rc;
    c->hist_cap = cap;
    return PLBFS_OK;
}

static int build_graph(plbfs_ctx *c) {
    if (c->graph_ok || !c->use_graph || c->timing) return PLBFS_OK;
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(c, ctx_scratch(c), false, true);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return plbfs_fail(PLBFS_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return plbfs_fail(PLBFS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    c->graph_ok = true;
    return PLBFS_OK;
}

static void fill_result(plbfs_ctx *c, const Ctrl &h, float ms, plbfs_result *out) {
    c->total_steps += h.step;
    if (!out) return;
    out->status = h.status;
    out->steps = h.step;
    out->converged = h.converged;
    out->total_steps = c->total_steps;
    for (int q = 0; q < 4; q++) out->residual[q] = h.res[q];
    out->solve_ms = ms;
}

extern "C" int plbfs_run(plbfs_ctx *c, int32_t max_steps, plbfs_result *out) {
    if (!c || max_steps < 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    int rc = check_ghost_state(c);
    if (rc) return rc;
    if ((rc = ensure_hist(c, max_steps))) return rc;
    if (rc) return rc;
    Ctrl h{};
    HIP_TRY(hipMemcpyAsync(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
    const RunScratch w = ctx_scratch(c);
    if ((rc = launch_run_start(c, w))) return rc;
    if ((rc = build_graph(c))) return rc;
    const int interval = std::max(1, c->params.check_interval);
    hipEvent_t ev0, ev1;
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    HIP_TRY(hipEventRecord(ev0, c->stream));
    for (int s = 0; s < max_steps; s++) {
        if (c->graph_ok && !c->timing) {
            HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
        } else if ((rc = launch_step(c, w, c->timing, true))) {
            return rc;
        }
        if ((s + 1) % interval == 0 || s == max_steps - 1) {
            HIP_TRY(hipMemcpyAsync(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            if (h.done && h.rebuild && h.status == PLBFS_OK && !h.converged) {
                // new candidate lists around the nodes' positions after step
                // h.step, then on from there: the launches after the stop were
                // no-ops, so the loop resumes at the step count
                if ((rc = ibm_rebuild(c)) || (rc = ensure_hist(c, max_steps)) || (rc = build_graph(c))) return rc;
                h.done = 0;
                h.rebuild = 0;
                HIP_TRY(hipMemcpyAsync(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
                s = h.step - 1;
                continue;
            }
            if (h.done) break;
        }
    }
    HIP_TRY(hipEventRecord(ev1, c->stream));
    HIP_TRY(hipEventSynchronize(ev1));
    resolve_timing(c);
    float ms = 0;
    hipEventElapsedTime(&ms, ev0, ev1);
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
    HIP_TRY(hipMemcpy(&h, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    c->last_hist.assign((size_t)h.step * 4, 0.0);
    if (h.step > 0)
        HIP_TRY(hipMemcpy(c->last_hist.data(), c->hist, (size_t)std::min(h.step, c->hist_cap) * 4 * sizeof(double),
                          hipMemcpyDeviceToHost));
    fill_result(c, h, ms, out);
    return h.status;
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of scratch a run of max_steps steps needs on context c.
static size_t scratch_need(const plbfs_ctx *c, int64_t max_steps) {
    return round256(sizeof(Ctrl)) + round256((size_t)c->L.n_blocks * 5 * sizeof(double)) +
           round256((size_t)c->cell_grid * sizeof(double)) + (size_t)max_steps * 4 * sizeof(double);
}

__global__ void k_clear_ctrl(Ctrl *ctrl) {
    if (threadIdx.x == 0) *ctrl = Ctrl{};
}

extern "C" int plbfs_run_noalloc(plbfs_ctx *c, int32_t max_steps, plbfs_result *out, double *hist_out) {
    if (!c || max_steps < 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    if (!c->scratch) return plbfs_fail(PLBFS_ERR_HIP, "no scratch address for this context's device");
    if (int e = check_ghost_state(c)) return e;
    const size_t need = scratch_need(c, max_steps);
    if (need > kScratchBytes)
        return plbfs_fail(PLBFS_ERR_INPUT,
                          "plbfs_run_noalloc: %d cells in %d blocks and %d steps need %zu B of scratch, limit %zu B",
                          c->L.N, c->L.n_blocks, max_steps, need, kScratchBytes);
    char *p = c->scratch;
    RunScratch w;
    w.ctrl = reinterpret_cast<Ctrl *>(p);
    p += round256(sizeof(Ctrl));
    w.partials = reinterpret_cast<double *>(p);
    p += round256((size_t)c->L.n_blocks * 5 * sizeof(double));
    w.bmin = reinterpret_cast<double *>(p);
    p += round256((size_t)c->cell_grid * sizeof(double));
    w.hist = reinterpret_cast<double *>(p);
    w.hist_cap = max_steps;
    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(k_clear_ctrl, dim3(1), dim3(64), 0, c->stream, w.ctrl);
    int rc = launch_run_start(c, w);
    if (rc) return rc;
    const int interval = std::max(1, c->params.check_interval);
    Ctrl h{};
    for (int s = 0; s < max_steps; s++) {
        if ((rc = launch_step(c, w, false, false))) return rc;
        if ((s + 1) % interval == 0 || s == max_steps - 1) {
            HIP_TRY(hipMemcpyAsync(&h, w.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            if (h.done) break;
        }
    }
    HIP_TRY(hipMemcpyAsync(&h, w.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (hist_out && h.step > 0)
        HIP_TRY(hipMemcpy(hist_out, w.hist, (size_t)h.step * 4 * sizeof(double), hipMemcpyDeviceToHost));
    c->last_hist.clear();  // plbfs_get_residual_history covers plbfs_run only
    fill_result(c, h, (float)ms, out);
    return h.status;
}

extern "C" int plbfs_get_residual_history(plbfs_ctx *c, double *out, int32_t max_steps) {
    if (!c || !out) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    const size_t n = std::min((size_t)max_steps * 4, c->last_hist.size());
    std::memcpy(out, c->last_hist.data(), n * sizeof(double));
    return PLBFS_OK;
}

extern "C" int plbfs_eval_residual(plbfs_ctx *c, double *grads, double *R) {
    if (!c) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    if (int e = check_ghost_state(c)) return e;
    const int N = c->L.N, nb = c->L.n_blocks;
    Ctrl h{};
    HIP_TRY(hipMemcpyAsync(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
    double *Rd = c->stB;  // scratch: stB is rewritten before it is read in a step
    if (!c->f32) {
        hipLaunchKernelGGL(k_grad<double>, dim3(c->cell_grid), dim3(kBlock), 0, c->stream, c->dm, c->ctrl,
                           (const double4 *)c->stA, (double4 *)c->grad);
        if (c->dm.n_border > 0)
            hipLaunchKernelGGL(k_flux_border<double>, dim3((c->dm.n_border + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               c->stream, c->dm, c->prm, c->ctrl, (const double4 *)c->stA,
                               (const double4 *)c->grad);
        hipLaunchKernelGGL(k_residual_only<double>, dim3(nb), dim3(kBlock), c->lds_bytes, c->stream, c->dm, c->prm,
                           (const double4 *)c->stA, (const double4 *)c->grad, c->fbody, Rd);
    } else {
        hipLaunchKernelGGL(k_grad<float>, dim3(c->cell_grid), dim3(kBlock), 0, c->stream, c->dm32, c->ctrl,
                           (const float4 *)c->stA, (float4 *)c->grad);
        if (c->dm32.n_border > 0)
            hipLaunchKernelGGL(k_flux_border<float>, dim3((c->dm32.n_border + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               c->stream, c->dm32, c->prm, c->ctrl, (const float4 *)c->stA,
                               (const float4 *)c->grad);
        hipLaunchKernelGGL(k_residual_only<float>, dim3(nb), dim3(kBlock), c->lds_bytes, c->stream, c->dm32,
                           c->prm, (const float4 *)c->stA, (const float4 *)c->grad, c->fbody, Rd);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (grads) {
        // records of kGradRec V4 per cell, the gradient from V4 kGradOff on
        const int rs = 4 * (c->f32 ? kGradRec<float> : kGradRec<double>);
        const int ro = 4 * (c->f32 ? kGradOff<float> : kGradOff<double>);
        std::vector<double> h2((size_t)rs * N);
        if (!c->f32) {
            HIP_TRY(hipMemcpy(h2.data(), c->grad, h2.size() * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> h1((size_t)rs * N);
            HIP_TRY(hipMemcpy(h1.data(), c->grad, h1.size() * sizeof(float), hipMemcpyDeviceToHost));
            h2.assign(h1.begin(), h1.end());
        }
        for (int n = 0; n < N; n++)
            for (int q = 0; q < 12; q++) grads[(size_t)c->L.perm[n] * 12 + q] = h2[(size_t)rs * n + ro + q];
    }
    if (R) {
        std::vector<double> h4((size_t)4 * N);
        HIP_TRY(hipMemcpy(h4.data(), Rd, h4.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int n = 0; n < N; n++)
            for (int q = 0; q < 4; q++) R[(size_t)c->L.perm[n] * 4 + q] = h4[(size_t)q * N + n];
    }
    return PLBFS_OK;
}

extern "C" int plbfs_get_dt(plbfs_ctx *c, double *dt) {
    if (!c || !dt) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    Ctrl h{};
    HIP_TRY(hipMemcpyAsync(c->ctrl, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream));
    int rc = launch_run_start(c, ctx_scratch(c));
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    double *outs[1] = {dt};
    return from_device_fields(c, c->dtl, 1, outs);
}

extern "C" int plbfs_enable_timing(plbfs_ctx *c, int32_t on) {
    if (!c) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    c->timing = on != 0;
    for (auto &k : c->kstat) k = KStat();
    return PLBFS_OK;
}

extern "C" int plbfs_kernel_stats(plbfs_ctx *c, int32_t which, double *total_ms, int32_t *launches) {
    if (!c || which < 0 || which > 3) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    if (total_ms) *total_ms = c->kstat[which].ms;
    if (launches) *launches = c->kstat[which].launches;
    return PLBFS_OK;
}

extern "C" int plbfs_layout_info(plbfs_ctx *c, int64_t *info) {
    if (!c || !info) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    const Layout &L = c->L;
    int64_t slots = 0;
    for (int n = 0; n < L.N; n++) slots += L.st_cnt[n];
    info[0] = L.N;
    info[1] = L.F;
    info[2] = L.n_items;
    info[3] = slots;
    info[4] = L.G;
    info[5] = L.n_blocks;
    info[6] = L.max_items_per_block;
    info[7] = (int64_t)L.g_ent.size();
    info[8] = L.block_cells;
    info[9] = (int64_t)c->lds_bytes;
    return PLBFS_OK;
}

extern "C" int plbfs_get_body_force(plbfs_ctx *c, double *f) {
    if (!c || !f) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    const int N = c->L.N;
    std::vector<double> h((size_t)3 * N);
    HIP_TRY(hipMemcpy(h.data(), c->fbody, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int n = 0; n < N; n++)
        for (int q = 0; q < 3; q++) f[(size_t)c->L.perm[n] * 3 + q] = h[(size_t)q * N + n];
    return PLBFS_OK;
}
