// Warm-start message-passing network on the cell graph (gfx950).
//
// 6 layers, hidden width 64, every MLP Linear -> SiLU -> Linear:
//   h  = enc(x)                                  x: [N,16] node inputs
//   repeat 6: m_e = msg([h[dst_e], h[src_e], f_e])    f_e: [E,16] edge inputs
//             a_t = sum_{dst_e = t} m_e / max(deg_t, 1)
//             h   = h + upd([h, a])
//   y  = dec(h)                                  y: [N,16]
// Directed edges are sorted by (dst, src); aggregation and every backward
// scatter walk fixed per-node edge lists, so no float atomics are used and
// results are bitwise reproducible.
//
// Every product runs on the matrix cores: k_gemm tiles C = op(A) op(B) into
// 64x64 blocks of 4 wavefronts, stages 64x32 / 32x64 operand tiles in LDS as
// bfloat16 (round to nearest even from the FP32 activations and master
// weights), and issues v_mfma_f32_16x16x16_bf16 with FP32 accumulation.  Weight
// gradients (reductions over all rows) are split over row ranges into FP32
// partials that k_sum_splits adds in a fixed order.  Adam keeps FP32 master
// weights and moments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ctx.hpp"
#include "plbfs_mpnn.h"

namespace {

constexpr int kH = 64;           // hidden width
constexpr int kDin = 16;         // node input width (padded)
constexpr int kDedge = 16;       // edge input width (padded)
constexpr int kDout = 16;        // output width (padded)
constexpr int kLayers = 6;
constexpr int kKmsg = 2 * kH + kDedge;  // 144
constexpr int kKupd = 2 * kH;           // 128
constexpr int kMlps = 2 + 2 * kLayers;  // enc, (msg, upd) x 6, dec
constexpr int kT = 256;                 // threads per block

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f) {
    // round to nearest even (inputs are finite)
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float silu(float a) { return a / (1.0f + expf(-a)); }
__device__ __forceinline__ float dsilu(float a) {
    const float s = 1.0f / (1.0f + expf(-a));
    return s * (1.0f + a * (1.0f - s));
}

// C[M,N] (+)= op(A)[M,K] * op(B)[K,N] over k in [k_begin, k_end) of this
// z-slice.  op(A)(m,k) = TA ? A[k*lda+m] : A[m*lda+k]; op(B)(k,n) = TB ?
// B[n*ldb+k] : B[k*ldb+n].  Epilogue: + bias[n] if bias; C2 = silu(value) if C2.
// Slice z writes to C + z * split_stride.
constexpr int kTM = 64, kTN = 64, kTK = 32;
constexpr int kLdsA = kTK + 8;  // row pitch (bf16) of the A tile [m][k]
constexpr int kLdsB = kTK + 8;  // row pitch of the B tile stored transposed [n][k]

template <bool TA, bool TB>
__global__ __launch_bounds__(kT) void k_gemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                             const float *__restrict__ B, int ldb, float *__restrict__ C, int ldc,
                                             int k_chunk, size_t split_stride, const float *__restrict__ bias,
                                             float *__restrict__ C2) {
    __shared__ unsigned short sA[kTM * kLdsA];
    __shared__ unsigned short sB[kTN * kLdsB];
    const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
    const int kb = blockIdx.z * k_chunk, ke = min(K, kb + k_chunk);
    C += blockIdx.z * split_stride;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;  // this wave's 32x32 sub-tile
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        // stage A (64 x 32) and B (32 x 64, stored [n][k]); 8 elements each per thread
#pragma unroll
        for (int r = 0; r < (kTM * kTK) / kT; r++) {
            const int idx = tid + r * kT;
            int m, k;
            if (TA) { m = idx % kTM; k = idx / kTM; } else { k = idx % kTK; m = idx / kTK; }
            const int gm = m0 + m, gk = k0 + k;
            float v = 0.f;
            if (gm < M && gk < ke) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            sA[m * kLdsA + k] = to_bf16(v);
        }
#pragma unroll
        for (int r = 0; r < (kTN * kTK) / kT; r++) {
            const int idx = tid + r * kT;
            int n, k;
            if (TB) { k = idx % kTK; n = idx / kTK; } else { n = idx % kTN; k = idx / kTN; }
            const int gn = n0 + n, gk = k0 + k;
            float v = 0.f;
            if (gn < N && gk < ke) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            sB[n * kLdsB + k] = to_bf16(v);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            // A fragment: row = lane % 16, k = kk + 4 * (lane / 16) + j
            // B fragment: col = lane % 16, k = kk + 4 * (lane / 16) + j
            bf16x4 af[2], bfr[2];
            const int kq = kk + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const unsigned short *pa = &sA[(wm + i * 16 + (lane & 15)) * kLdsA + kq];
                af[i] = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                const unsigned short *pb = &sB[(wn + i * 16 + (lane & 15)) * kLdsB + kq];
                bfr[i] = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D fragment: col = lane % 16, row = 4 * (lane / 16) + j
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int gn = n0 + wn + j * 16 + (lane & 15);
            if (gn >= N) continue;
            const float bv = bias ? bias[gn] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int gm = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
                if (gm >= M) continue;
                const float v = acc[i][j][q] + bv;
                C[(size_t)gm * ldc + gn] = v;
                if (C2) C2[(size_t)gm * ldc + gn] = silu(v);
            }
        }
}

// out[i] = sum_z parts[z * stride + i] in z order (deterministic split-K sum)
__global__ __launch_bounds__(kT) void k_sum_splits(int n, int splits, size_t stride, const float *__restrict__ parts,
                                                   float *__restrict__ out) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < splits; z++) s += parts[z * stride + i];
    out[i] = s;
}

// db[c] = sum over rows of d[r * ld + c] (one block per column, fixed tree order)
__global__ __launch_bounds__(kT) void k_colsum(int rows, int ld, const float *__restrict__ d, float *__restrict__ db) {
    __shared__ float red[kT];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += kT) s += d[(size_t)r * ld + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// z[e] = [h[dst_e], h[src_e], f_e]
__global__ __launch_bounds__(kT) void k_gather_msg(int E, const int *__restrict__ src, const int *__restrict__ dst,
                                                   const float *__restrict__ h, const float *__restrict__ f,
                                                   float *__restrict__ z) {
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (int64_t)E * kKmsg) return;
    const int e = (int)(t / kKmsg), c = (int)(t - (int64_t)e * kKmsg);
    float v;
    if (c < kH) v = h[(size_t)dst[e] * kH + c];
    else if (c < 2 * kH) v = h[(size_t)src[e] * kH + c - kH];
    else v = f[(size_t)e * kDedge + c - 2 * kH];
    z[t] = v;
}

// u[t] = [h[t], inv_deg[t] * sum_{e in in(t)} m[e]]
__global__ __launch_bounds__(kT) void k_aggregate(int N, const int *__restrict__ in_off, const float *__restrict__ inv_deg,
                                                  const float *__restrict__ h, const float *__restrict__ m,
                                                  float *__restrict__ u) {
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (int64_t)N * kH) return;
    const int i = (int)(t / kH), c = (int)(t - (int64_t)i * kH);
    float s = 0.f;
    for (int e = in_off[i]; e < in_off[i + 1]; e++) s += m[(size_t)e * kH + c];
    u[(size_t)i * kKupd + c] = h[t];
    u[(size_t)i * kKupd + kH + c] = s * inv_deg[i];
}

// y[i] += x[i] (residual stream)
__global__ __launch_bounds__(kT) void k_axpy(int64_t n, const float *__restrict__ x, float *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// d[i] *= silu'(a[i])
__global__ __launch_bounds__(kT) void k_dsilu(int64_t n, const float *__restrict__ a, float *__restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) d[i] *= dsilu(a[i]);
}

// Backward of aggregate + gather for one layer:
//   dm[e]  = du[dst_e, 64:128] * inv_deg[dst_e]
__global__ __launch_bounds__(kT) void k_dmsg(int E, const int *__restrict__ dst, const float *__restrict__ inv_deg,
                                             const float *__restrict__ du, float *__restrict__ dm) {
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (int64_t)E * kH) return;
    const int e = (int)(t / kH), c = (int)(t - (int64_t)e * kH);
    const int d = dst[e];
    dm[t] = du[(size_t)d * kKupd + kH + c] * inv_deg[d];
}

//   dh[t] += du[t, 0:64] + sum_{e in in(t)} dz[e, 0:64] + sum_{e in out(t)} dz[e, 64:128]
__global__ __launch_bounds__(kT) void k_dh_layer(int N, const int *__restrict__ in_off, const int *__restrict__ out_off,
                                                 const int *__restrict__ out_edge, const float *__restrict__ du,
                                                 const float *__restrict__ dz, float *__restrict__ dh) {
    const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (int64_t)N * kH) return;
    const int i = (int)(t / kH), c = (int)(t - (int64_t)i * kH);
    float s = dh[t] + du[(size_t)i * kKupd + c];
    for (int e = in_off[i]; e < in_off[i + 1]; e++) s += dz[(size_t)e * kKmsg + c];
    for (int q = out_off[i]; q < out_off[i + 1]; q++) s += dz[(size_t)out_edge[q] * kKmsg + kH + c];
    dh[t] = s;
}

// dy = 2 (y - t) mask / n_used; per-block partial sums of the squared error
__global__ __launch_bounds__(kT) void k_mse(int64_t n, float inv_n, const float *__restrict__ y,
                                            const float *__restrict__ tgt, const float *__restrict__ mask,
                                            float *__restrict__ dy, double *__restrict__ part) {
    __shared__ double red[kT];
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    double s = 0;
    if (i < n) {
        const float d = (y[i] - tgt[i]) * mask[i];
        dy[i] = 2.f * d * inv_n;
        s = (double)d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kT) void k_adam(int n, float lr, float b1, float b2, float eps, float c1, float c2,
                                             const float *__restrict__ g, float *__restrict__ w,
                                             float *__restrict__ mo, float *__restrict__ ve) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.f - b1) * gi;
    const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    w[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

struct MlpShape {
    int k, o;
};

MlpShape mlp_shape(int q) {
    if (q == 0) return {kDin, kH};
    if (q == kMlps - 1) return {kH, kDout};
    return (q % 2 == 1) ? MlpShape{kKmsg, kH} : MlpShape{kKupd, kH};
}

// offsets of W1, b1, W2, b2 of MLP q in the flat parameter vector
struct MlpOff {
    size_t W1, b1, W2, b2, end;
};

MlpOff mlp_off(int q) {
    size_t p = 0;
    for (int r = 0; r < q; r++) {
        const MlpShape s = mlp_shape(r);
        p += (size_t)s.k * kH + kH + (size_t)kH * s.o + s.o;
    }
    const MlpShape s = mlp_shape(q);
    MlpOff o;
    o.W1 = p;
    o.b1 = o.W1 + (size_t)s.k * kH;
    o.W2 = o.b1 + kH;
    o.b2 = o.W2 + (size_t)kH * s.o;
    o.end = o.b2 + s.o;
    return o;
}

size_t n_params_total() { return mlp_off(kMlps - 1).end; }

}  // namespace

// Activation caches of one MLP application: input z [rows, k], a1 and s1 [rows, 64].
struct MlpCache {
    float *z = nullptr, *a1 = nullptr, *s1 = nullptr;
    int rows = 0;
};

struct plbfs_mpnn {
    int N = 0, E = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    int *src = nullptr, *dst = nullptr, *in_off = nullptr, *out_off = nullptr, *out_edge = nullptr;
    float *inv_deg = nullptr, *ef = nullptr;
    std::vector<int32_t> h_order;  // sorted edge -> caller edge
    size_t P = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    int adam_t = 0;
    // activations
    float *x = nullptr, *y = nullptr, *tgt = nullptr, *mask = nullptr, *dy = nullptr;
    float *h[kLayers + 1] = {};
    MlpCache cache[kMlps];
    float *out_tmp = nullptr;  // [max(E,N) * 64] MLP output scratch (messages, update deltas)
    float *dbuf0 = nullptr, *dbuf1 = nullptr, *dbuf2 = nullptr;  // backward scratch
    float *dh = nullptr;
    float *split = nullptr;
    size_t split_cap = 0;
    double *mse_part = nullptr;
    int mse_blocks = 0;
    float mse_scale = 1.f;
    bool train = false;
};

namespace {

int malloc_t(plbfs_mpnn *n, void **p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    HIP_TRY(hipMalloc(p, bytes));
    n->allocs.push_back(*p);
    return PLBFS_OK;
}

template <typename T>
int upload_t(plbfs_mpnn *n, T **dst, const std::vector<T> &src) {
    int rc = malloc_t(n, (void **)dst, src.size() * sizeof(T));
    if (rc) return rc;
    if (!src.empty()) HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return PLBFS_OK;
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kT - 1) / kT)); }

// C[M,N] = op(A) op(B) (+ bias, C2 = silu) on the matrix cores.
template <bool TA, bool TB>
void gemm(plbfs_mpnn *n, int M, int Nn, int K, const float *A, int lda, const float *B, int ldb, float *C, int ldc,
          const float *bias, float *C2) {
    const dim3 grid((M + kTM - 1) / kTM, (Nn + kTN - 1) / kTN, 1);
    hipLaunchKernelGGL((k_gemm<TA, TB>), grid, dim3(kT), 0, n->stream, M, Nn, K, A, lda, B, ldb, C, ldc, K,
                       (size_t)0, bias, C2);
}

// Weight gradient dW[K, Nn] = A^T[K, M] B[M, Nn] with the reduction over M
// split into row ranges (deterministic fixed-order sum of the partials).
int gemm_tn_split(plbfs_mpnn *n, int M, int Nn, int K, const float *A, int lda, const float *B, int ldb, float *C) {
    const int chunk = 4096;
    const int splits = std::max(1, (M + chunk - 1) / chunk);
    const size_t stride = (size_t)K * Nn;
    if ((size_t)splits * stride > n->split_cap) {
        int rc = malloc_t(n, (void **)&n->split, (size_t)splits * stride * sizeof(float));
        if (rc) return rc;
        n->split_cap = (size_t)splits * stride;
    }
    // output rows = K (columns of A), reduction dimension = M (rows of A and B)
    const dim3 grid((K + kTM - 1) / kTM, (Nn + kTN - 1) / kTN, splits);
    hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(kT), 0, n->stream, K, Nn, M, A, lda, B, ldb, n->split, Nn,
                       chunk, stride, (const float *)nullptr, (float *)nullptr);
    hipLaunchKernelGGL(k_sum_splits, grid1((int64_t)stride), dim3(kT), 0, n->stream, (int)stride, splits, stride,
                       n->split, C);
    return PLBFS_OK;
}

// y = W2 silu(W1 z + b1) + b2 over rows rows; caches z, a1, s1.
void mlp_fwd(plbfs_mpnn *n, int q, int rows, float *z, float *out) {
    const MlpShape s = mlp_shape(q);
    const MlpOff o = mlp_off(q);
    MlpCache &c = n->cache[q];
    c.z = z;
    c.rows = rows;
    gemm<false, false>(n, rows, kH, s.k, z, s.k, n->w + o.W1, kH, c.a1, kH, n->w + o.b1, c.s1);
    gemm<false, false>(n, rows, s.o, kH, c.s1, kH, n->w + o.W2, s.o, out, s.o, n->w + o.b2, nullptr);
}

// Backward of MLP q given dy [rows, o]: writes its parameter gradients and
// dz [rows, k] (if dz).  Scratch: dbuf0 [rows, 64].
int mlp_bwd(plbfs_mpnn *n, int q, const float *dy, float *dz) {
    const MlpShape s = mlp_shape(q);
    const MlpOff o = mlp_off(q);
    const MlpCache &c = n->cache[q];
    const int rows = c.rows;
    int rc;
    if ((rc = gemm_tn_split(n, rows, s.o, kH, c.s1, kH, dy, s.o, n->g + o.W2))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(s.o), dim3(kT), 0, n->stream, rows, s.o, dy, n->g + o.b2);
    float *da1 = n->dbuf0;
    gemm<false, true>(n, rows, kH, s.o, dy, s.o, n->w + o.W2, s.o, da1, kH, nullptr, nullptr);
    hipLaunchKernelGGL(k_dsilu, grid1((int64_t)rows * kH), dim3(kT), 0, n->stream, (int64_t)rows * kH, c.a1, da1);
    if ((rc = gemm_tn_split(n, rows, kH, s.k, c.z, s.k, da1, kH, n->g + o.W1))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(kH), dim3(kT), 0, n->stream, rows, kH, da1, n->g + o.b1);
    if (dz) gemm<false, true>(n, rows, s.k, kH, da1, kH, n->w + o.W1, kH, dz, s.k, nullptr, nullptr);
    return PLBFS_OK;
}

void forward(plbfs_mpnn *n) {
    const int N = n->N, E = n->E;
    mlp_fwd(n, 0, N, n->x, n->h[0]);
    for (int l = 0; l < kLayers; l++) {
        const int qm = 1 + 2 * l, qu = 2 + 2 * l;
        hipLaunchKernelGGL(k_gather_msg, grid1((int64_t)E * kKmsg), dim3(kT), 0, n->stream, E, n->src, n->dst,
                           n->h[l], n->ef, n->cache[qm].z);
        mlp_fwd(n, qm, E, n->cache[qm].z, n->out_tmp);
        hipLaunchKernelGGL(k_aggregate, grid1((int64_t)N * kH), dim3(kT), 0, n->stream, N, n->in_off, n->inv_deg,
                           n->h[l], n->out_tmp, n->cache[qu].z);
        mlp_fwd(n, qu, N, n->cache[qu].z, n->h[l + 1]);
        hipLaunchKernelGGL(k_axpy, grid1((int64_t)N * kH), dim3(kT), 0, n->stream, (int64_t)N * kH, n->h[l],
                           n->h[l + 1]);
    }
    mlp_fwd(n, kMlps - 1, N, n->h[kLayers], n->y);
}

int backward(plbfs_mpnn *n) {
    const int N = n->N, E = n->E;
    int rc;
    // dh = d dec / d h
    if ((rc = mlp_bwd(n, kMlps - 1, n->dy, n->dh))) return rc;
    for (int l = kLayers - 1; l >= 0; l--) {
        const int qm = 1 + 2 * l, qu = 2 + 2 * l;
        float *du = n->dbuf1;  // [N, 128]
        if ((rc = mlp_bwd(n, qu, n->dh, du))) return rc;
        float *dm = n->out_tmp;  // [E, 64]
        hipLaunchKernelGGL(k_dmsg, grid1((int64_t)E * kH), dim3(kT), 0, n->stream, E, n->dst, n->inv_deg, du, dm);
        float *dz = n->dbuf2;  // [E, 144]
        if ((rc = mlp_bwd(n, qm, dm, dz))) return rc;
        hipLaunchKernelGGL(k_dh_layer, grid1((int64_t)N * kH), dim3(kT), 0, n->stream, N, n->in_off, n->out_off,
                           n->out_edge, du, dz, n->dh);
    }
    return mlp_bwd(n, 0, n->dh, nullptr);
}

}  // namespace

extern "C" int plbfs_mpnn_n_params(void) { return (int)n_params_total(); }

extern "C" int plbfs_mpnn_create(int32_t n_nodes, int32_t n_edges, const int32_t *src, const int32_t *dst,
                                 const float *edge_feat, int32_t train, plbfs_mpnn **out) {
    if (!out || n_nodes <= 0 || n_edges < 0 || (n_edges > 0 && (!src || !dst || !edge_feat)))
        return plbfs_fail(PLBFS_ERR_INPUT, "bad graph");
    *out = nullptr;
    for (int e = 0; e < n_edges; e++)
        if (src[e] < 0 || src[e] >= n_nodes || dst[e] < 0 || dst[e] >= n_nodes)
            return plbfs_fail(PLBFS_ERR_INPUT, "edge %d: node index out of range", e);
    if (plbfs_device_count() <= 0) return plbfs_fail(PLBFS_ERR_NODEVICE, "no HIP device visible");
    auto *n = new plbfs_mpnn();
    auto fail = [&](int rc) {
        plbfs_mpnn_destroy(n);
        return rc;
    };
    const int N = n_nodes, E = n_edges;
    n->N = N;
    n->E = E;
    n->train = train != 0;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(plbfs_fail(PLBFS_ERR_HIP, "hipStreamCreate"));
    // sort edges by (dst, src), stable
    std::vector<int32_t> ord(E);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return dst[a] != dst[b] ? dst[a] < dst[b] : src[a] < src[b];
    });
    n->h_order = ord;
    std::vector<int32_t> s(E), d(E), in_off(N + 1, 0), out_off(N + 1, 0), out_edge(E);
    std::vector<float> f((size_t)E * kDedge), inv_deg(N);
    for (int e = 0; e < E; e++) {
        s[e] = src[ord[e]];
        d[e] = dst[ord[e]];
        std::memcpy(&f[(size_t)e * kDedge], edge_feat + (size_t)ord[e] * kDedge, kDedge * sizeof(float));
        in_off[d[e] + 1]++;
        out_off[s[e] + 1]++;
    }
    for (int i = 0; i < N; i++) {
        inv_deg[i] = 1.0f / (float)std::max(in_off[i + 1], 1);
        in_off[i + 1] += in_off[i];
        out_off[i + 1] += out_off[i];
    }
    {
        std::vector<int32_t> fill(out_off.begin(), out_off.end() - 1);
        for (int e = 0; e < E; e++) out_edge[fill[s[e]]++] = e;  // increasing edge order per source
    }
    int rc;
    if ((rc = upload_t(n, &n->src, s)) || (rc = upload_t(n, &n->dst, d)) || (rc = upload_t(n, &n->in_off, in_off)) ||
        (rc = upload_t(n, &n->out_off, out_off)) || (rc = upload_t(n, &n->out_edge, out_edge)) ||
        (rc = upload_t(n, &n->inv_deg, inv_deg)) || (rc = upload_t(n, &n->ef, f)))
        return fail(rc);
    n->P = n_params_total();
    const size_t P = n->P;
    std::vector<float> zero(P, 0.f);
    if ((rc = upload_t(n, &n->w, zero)) || (rc = upload_t(n, &n->g, zero)) || (rc = upload_t(n, &n->m1, zero)) ||
        (rc = upload_t(n, &n->m2, zero)))
        return fail(rc);
    const size_t NE = (size_t)std::max(N, E);
    auto fa = [&](float **p, size_t count) { return malloc_t(n, (void **)p, count * sizeof(float)); };
    if ((rc = fa(&n->x, (size_t)N * kDin)) || (rc = fa(&n->y, (size_t)N * kDout)) ||
        (rc = fa(&n->tgt, (size_t)N * kDout)) || (rc = fa(&n->mask, (size_t)N * kDout)) ||
        (rc = fa(&n->dy, (size_t)N * kDout)) || (rc = fa(&n->out_tmp, NE * kH)) ||
        (rc = fa(&n->dbuf0, NE * kH)) || (rc = fa(&n->dbuf1, (size_t)N * kKupd)) ||
        (rc = fa(&n->dh, (size_t)N * kH)))
        return fail(rc);
    if (n->train && (rc = fa(&n->dbuf2, (size_t)E * kKmsg + 1))) return fail(rc);
    for (int l = 0; l <= kLayers; l++)
        if ((rc = fa(&n->h[l], (size_t)N * kH))) return fail(rc);
    // activations kept for the backward pass; without training every message
    // layer shares one buffer set and every update layer another
    for (int q = 0; q < kMlps; q++) {
        const MlpShape sh = mlp_shape(q);
        const size_t rows = (q % 2 == 1 && q < kMlps - 1) ? (size_t)E : (size_t)N;
        MlpCache &c = n->cache[q];
        if (!n->train && q >= 3 && q < kMlps - 1) {
            c = n->cache[q % 2 == 1 ? 1 : 2];
            continue;
        }
        if ((q > 0 && q < kMlps - 1 && (rc = fa(&c.z, rows * sh.k + 1))) || (rc = fa(&c.a1, rows * kH + 1)) ||
            (rc = fa(&c.s1, rows * kH + 1)))
            return fail(rc);
    }
    n->mse_blocks = (int)(((int64_t)N * kDout + kT - 1) / kT);
    if ((rc = malloc_t(n, (void **)&n->mse_part, (size_t)n->mse_blocks * sizeof(double)))) return fail(rc);
    *out = n;
    return PLBFS_OK;
}

extern "C" int plbfs_mpnn_destroy(plbfs_mpnn *n) {
    if (!n) return PLBFS_OK;
    if (n->stream) hipStreamSynchronize(n->stream);
    for (void *p : n->allocs) hipFree(p);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
    return PLBFS_OK;
}

extern "C" int plbfs_mpnn_set_params(plbfs_mpnn *n, const float *params) {
    if (!n || !params) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    HIP_TRY(hipMemcpy(n->w, params, n->P * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(n->m1, 0, n->P * sizeof(float)));
    HIP_TRY(hipMemset(n->m2, 0, n->P * sizeof(float)));
    n->adam_t = 0;
    return PLBFS_OK;
}

extern "C" int plbfs_mpnn_get_params(plbfs_mpnn *n, float *params) {
    if (!n || !params) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    HIP_TRY(hipStreamSynchronize(n->stream));
    HIP_TRY(hipMemcpy(params, n->w, n->P * sizeof(float), hipMemcpyDeviceToHost));
    return PLBFS_OK;
}

extern "C" int plbfs_mpnn_forward(plbfs_mpnn *n, const float *x, float *y) {
    if (!n || !x || !y) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    HIP_TRY(hipMemcpyAsync(n->x, x, (size_t)n->N * kDin * sizeof(float), hipMemcpyHostToDevice, n->stream));
    n->cache[0].z = n->x;
    forward(n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, n->y, (size_t)n->N * kDout * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    return PLBFS_OK;
}

namespace {

int loss_grad_device(plbfs_mpnn *n, double *loss) {
    const int64_t ny = (int64_t)n->N * kDout;
    n->cache[0].z = n->x;
    forward(n);
    hipLaunchKernelGGL(k_mse, grid1(ny), dim3(kT), 0, n->stream, ny, n->mse_scale, n->y, n->tgt, n->mask, n->dy,
                       n->mse_part);
    int rc = backward(n);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    if (loss) {
        std::vector<double> part(n->mse_blocks);
        HIP_TRY(hipMemcpyAsync(part.data(), n->mse_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                               n->stream));
        HIP_TRY(hipStreamSynchronize(n->stream));
        double s = 0;
        for (double v : part) s += v;
        *loss = s * n->mse_scale;
    }
    return PLBFS_OK;
}

}  // namespace

extern "C" int plbfs_mpnn_loss_grad(plbfs_mpnn *n, const float *x, const float *target, const float *mask,
                                    double *loss, float *grads) {
    if (!n || !x || !target || !mask) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    if (!n->train) return plbfs_fail(PLBFS_ERR_INPUT, "network created without training buffers");
    const size_t ny = (size_t)n->N * kDout;
    double used = 0;
    for (size_t i = 0; i < ny; i++) used += mask[i] != 0.f ? 1.0 : 0.0;
    n->mse_scale = (float)(1.0 / std::max(used, 1.0));
    HIP_TRY(hipMemcpyAsync(n->x, x, (size_t)n->N * kDin * sizeof(float), hipMemcpyHostToDevice, n->stream));
    HIP_TRY(hipMemcpyAsync(n->tgt, target, ny * sizeof(float), hipMemcpyHostToDevice, n->stream));
    HIP_TRY(hipMemcpyAsync(n->mask, mask, ny * sizeof(float), hipMemcpyHostToDevice, n->stream));
    int rc = loss_grad_device(n, loss);
    if (rc) return rc;
    if (grads) HIP_TRY(hipMemcpy(grads, n->g, n->P * sizeof(float), hipMemcpyDeviceToHost));
    return PLBFS_OK;
}

extern "C" int plbfs_mpnn_adam_step(plbfs_mpnn *n, float lr, float beta1, float beta2, float eps) {
    if (!n) return plbfs_fail(PLBFS_ERR_INPUT, "null argument");
    n->adam_t++;
    const float c1 = 1.f / (1.f - std::pow(beta1, (float)n->adam_t));
    const float c2 = 1.f / (1.f - std::pow(beta2, (float)n->adam_t));
    hipLaunchKernelGGL(k_adam, grid1((int64_t)n->P), dim3(kT), 0, n->stream, (int)n->P, lr, beta1, beta2, eps, c1, c2,
                       n->g, n->w, n->m1, n->m2);
    HIP_TRY(hipGetLastError());
    return PLBFS_OK;
}

extern "C" int plbfs_mpnn_train(plbfs_mpnn *n, int32_t n_samples, const float *X, const float *T, const float *mask,
                                int32_t epochs, float lr, uint64_t seed, double *loss_hist) {
    if (!n || n_samples <= 0 || !X || !T || !mask || epochs <= 0) return plbfs_fail(PLBFS_ERR_INPUT, "bad arguments");
    if (!n->train) return plbfs_fail(PLBFS_ERR_INPUT, "network created without training buffers");
    const size_t nx = (size_t)n->N * kDin, ny = (size_t)n->N * kDout;
    float *dX = nullptr, *dT = nullptr;
    int rc;
    if ((rc = malloc_t(n, (void **)&dX, nx * n_samples * sizeof(float))) ||
        (rc = malloc_t(n, (void **)&dT, ny * n_samples * sizeof(float))))
        return rc;
    HIP_TRY(hipMemcpy(dX, X, nx * n_samples * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dT, T, ny * n_samples * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(n->mask, mask, ny * sizeof(float), hipMemcpyHostToDevice));
    double used = 0;
    for (size_t i = 0; i < ny; i++) used += mask[i] != 0.f ? 1.0 : 0.0;
    n->mse_scale = (float)(1.0 / std::max(used, 1.0));
    std::vector<int> order(n_samples);
    std::iota(order.begin(), order.end(), 0);
    uint64_t st = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    auto release = [&]() {
        for (float *p : {dX, dT}) {
            n->allocs.erase(std::find(n->allocs.begin(), n->allocs.end(), (void *)p));
            hipFree(p);
        }
    };
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = n_samples - 1; i > 0; i--) std::swap(order[i], order[rnd() % (uint64_t)(i + 1)]);
        const double lr_min = lr / 20.0, ph = epochs > 1 ? (double)ep / (epochs - 1) : 0.0;
        const float lr_ep = (float)(lr_min + 0.5 * (lr - lr_min) * (1.0 + std::cos(M_PI * ph)));
        double sum = 0;
        for (int k = 0; k < n_samples; k++) {
            const int sidx = order[k];
            HIP_TRY(hipMemcpyAsync(n->x, dX + nx * sidx, nx * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
            HIP_TRY(hipMemcpyAsync(n->tgt, dT + ny * sidx, ny * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
            double l = 0;
            if ((rc = loss_grad_device(n, &l)) || (rc = plbfs_mpnn_adam_step(n, lr_ep, 0.9f, 0.999f, 1e-8f))) {
                hipStreamSynchronize(n->stream);
                release();
                return rc;
            }
            sum += l;
        }
        if (loss_hist) loss_hist[ep] = sum / n_samples;
    }
    HIP_TRY(hipStreamSynchronize(n->stream));
    release();
    return PLBFS_OK;
}
