// Per-cell and per-face pieces of one RK stage, shared by the single-case
// kernels (solver.hip) and the batched kernels (batch.hip).
//
// A "record index" names where cell c's state and gradient live: c for one
// case, c * B + b for case b of a batch stored case-fastest (so the B cases of
// one cell are adjacent and a gather of one cell for consecutive cases is one
// contiguous run).
#pragma once
#include <type_traits>

#include "lbfs_device.hpp"
#include "plbfs_hip.h"

namespace plbfs {

struct Rec1 {
    __device__ __forceinline__ size_t operator()(int c) const { return (size_t)c; }
};
struct RecB {
    int B, b;
    __device__ __forceinline__ size_t operator()(int c) const { return (size_t)c * B + b; }
};

// Per-cell RK accumulator between stages: the running sum t1 itself in FP64
// storage; in FP32 storage only the increment sum_k beta_k dt R_k (the
// conservative state stays in t0 in FP64, so small increments keep their bits).
template <typename S>
struct AccT {
    using type = double4;
};
template <>
struct AccT<float> {
    using type = float4;
};
template <typename S>
using Acc = typename AccT<S>::type;

struct StageArgs {
    double alpha, beta;
    int stage;  // 0..3
};

__device__ __forceinline__ double local_dt(const DevParams &prm, double visc, double rho, double u, double ax,
                                           double vol) {
    // Solver.cpp:1715-1744 with the x-slot-only CFL areas (A_y = A_z = 0)
    double eig = (fabs(u) + prm.ceff) * ax;
    eig = eig / prm.gamma;
    double ve = 2 * visc / prm.gamma / rho / vol;
    ve = ve * (ax * ax);
    eig = eig + 4 * ve;
    return prm.cfl * vol / eig;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Least-squares gradient of cell i (src/gradients.cpp:319-389): folded stencil
// weights times neighbour differences; ghost neighbours from the owner's patch.
// Passes a value through an empty asm statement (no instruction is emitted).
// Without it the optimiser merges "load from LDS" and "load from global" on
// the two sides of a branch into one flat load through a selected pointer,
// which is slower than either and waits on both memory counters.
template <typename V>
__device__ __forceinline__ V opaque4(V v) {
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
    return v;
}

template <typename S, typename R>
struct GlobalStates {
    const V4<S> *__restrict__ st;
    R rec;
    __device__ __forceinline__ V4<S> operator()(int j) const { return st[rec(j)]; }
};

// States of cells [c0, c0 + nc) from an LDS copy, the rest from global memory.
template <typename S>
struct StagedStates {
    const V4<S> *__restrict__ st;
    const V4<S> *sSt;
    int c0, nc;
    __device__ __forceinline__ V4<S> operator()(int j) const {
        const unsigned rel = (unsigned)(j - c0);
        if (rel < (unsigned)nc) return opaque4(sSt[rel]);
        return st[j];
    }
};

template <typename S, typename ST>
__device__ __forceinline__ void lsq_grad(const DevMesh<S> &m, int i, const ST &state, S g[12]) {
    const int N = m.N;
    const V4<S> q0 = state(i);
    const S p0[4] = {q0.x, q0.y, q0.z, q0.w};
#pragma unroll
    for (int q = 0; q < 12; q++) g[q] = (S)0;
    const int cnt = m.st_cnt[i];
#pragma unroll
    for (int s = 0; s < kMaxSlots; s++) {
        if (s < cnt) {
            const int j = m.st_nbr[(size_t)s * N + i];
            S pj[4];
            if (j >= 0) {
                const V4<S> qj = state(j);
                pj[0] = qj.x;
                pj[1] = qj.y;
                pj[2] = qj.z;
                pj[3] = qj.w;
            } else {
                // (periodic ghosts were replaced by their partner cell in st_nbr)
                const int b = -1 - j;
                const DevPatch P = m.patches[m.b_patch[b]];
                S fx[4] = {0, 0, 0, 0};
                if (P.flags & kGhostFixed) {
                    const V4<S> f = m.b_fix[b];
                    fx[0] = f.x; fx[1] = f.y; fx[2] = f.z; fx[3] = f.w;
                }
                ghost_state<S>(P, p0, p0, fx, pj);
            }
            const S wx = m.st_w[((size_t)s * 3 + 0) * N + i];
            const S wy = m.st_w[((size_t)s * 3 + 1) * N + i];
            const S wz = m.st_w[((size_t)s * 3 + 2) * N + i];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const S d = pj[q] - p0[q];
                g[3 * q] += wx * d;
                g[3 * q + 1] += wy * d;
                g[3 * q + 2] += wz * d;
            }
        }
    }
}

template <typename S>
__device__ __forceinline__ void unpack_side(V4<S> q, V4<S> a, V4<S> b, V4<S> d, S p[4], S g[12]) {
    p[0] = q.x;
    p[1] = q.y;
    p[2] = q.z;
    p[3] = q.w;
    g[0] = a.x; g[1] = a.y; g[2] = a.z; g[3] = a.w;
    g[4] = b.x; g[5] = b.y; g[6] = b.z; g[7] = b.w;
    g[8] = d.x; g[9] = d.y; g[10] = d.z; g[11] = d.w;
}

// Where a face item reads a cell's state + gradient record (and, in FP64
// storage, its centroid): cells of the block's own range [c0, c0 + nc) from
// the copy the block staged in LDS (sRec: state, grad x3 per cell; sCg:
// centroid + 1/V), every other cell from global memory.  With nc = 0 every
// read goes to global memory.
// COMB: gradient records carry the state (kGradRec<S> V4 per record, the
// single-case FP32 layout); otherwise the state comes from `st` and the record
// is the three gradient V4 (FP64, and the batched solver in either precision).
template <typename S, typename R, bool COMB = (kGradOff<S> == 1)>
struct SideLoader {
    const V4<S> *__restrict__ st;
    const V4<S> *__restrict__ grad;
    const double4 *__restrict__ cgeo;
    R rec;
    const V4<S> *sRec;  // staged records, cell-major then case lane: 4 * (rel * C + cl)
    const double4 *sCg;
    int c0, nc;
    int C, cl;  // cases staged per cell and this item's case lane (1, 0 for one case)
    __device__ __forceinline__ void side(int c, S p[4], S g[12]) const {
        const unsigned rel = (unsigned)(c - c0);
        if (rel < (unsigned)nc) {
            const V4<S> *q = sRec + 4 * (rel * C + cl);
            unpack_side<S>(opaque4(q[0]), opaque4(q[1]), opaque4(q[2]), opaque4(q[3]), p, g);
        } else {
            const size_t r = rec(c);
            if constexpr (COMB) {
                const V4<S> *gc = grad + kGradRec<S> * r;
                unpack_side<S>(gc[0], gc[1], gc[2], gc[3], p, g);
            } else {
                const V4<S> *gc = grad + 3 * r;
                unpack_side<S>(st[r], gc[0], gc[1], gc[2], p, g);
            }
        }
    }
    __device__ __forceinline__ double4 centroid(int c) const {
        const unsigned rel = (unsigned)(c - c0);
        if (rel < (unsigned)nc) return opaque4(sCg[rel]);
        return cgeo[c];
    }
};

// Copies the records of cells [c0, c0 + nc) into LDS, 4 V4 per (cell, case
// lane) for the C cases b0 .. b0 + C - 1 of a batch of B (C = B = 1 for one
// case; lanes past the last case are left unset), and, when cgeo is given,
// the cells' centroid records.
template <typename S, bool COMB = (kGradOff<S> == 1)>
__device__ __forceinline__ void stage_records(const V4<S> *__restrict__ st, const V4<S> *__restrict__ grad,
                                              const double4 *__restrict__ cgeo, int B, int b0, int C, int c0,
                                              int nc, V4<S> *sRec, double4 *sCg) {
    for (int k = threadIdx.x; k < nc * C * 4; k += blockDim.x) {
        const int r = k >> 2, q = k & 3;
        const int il = r / C, cl = r - il * C;
        if (b0 + cl >= B) continue;
        const size_t g = (size_t)(c0 + il) * B + b0 + cl;
        if constexpr (COMB) sRec[k] = grad[kGradRec<S> * g + q];
        else sRec[k] = q == 0 ? st[g] : grad[3 * g + q - 1];
    }
    if (cgeo)
        for (int k = threadIdx.x; k < nc; k += blockDim.x) sCg[k] = cgeo[c0 + k];
}

// Byte offset of the staged records behind the face slots (C cases per slot)
// in a flux block's LDS.
template <typename S>
__host__ __device__ inline size_t staged_offset(int max_items, int C = 1) {
    return ((size_t)4 * max_items * C * sizeof(S) + 31) / 32 * 32;
}

// Dynamic LDS of one flux block: face slots, and with staging the block's own
// cell records for C cases (+ centroid records in FP64 storage).
template <typename S>
inline size_t flux_lds_bytes(int max_items, int block_cells, bool stage, int C = 1) {
    size_t b = staged_offset<S>(max_items, C);
    if (stage) {
        b += (size_t)block_cells * C * 4 * sizeof(V4<S>);
        if (std::is_same<S, double>::value) b += (size_t)block_cells * sizeof(double4);
    }
    return b;
}

// Flux of face work item e (src/Solver.cpp:286-347 with populate_lattice_macros
// :840-1274); v3g = 3 nu / gamma of the case.  dtcap caps the item's streaming
// dt: a batch stores the geometric 0.5 sqrt(A) and caps it per case with
// 3 nu / (gamma fneq_min), the min the host applies for one case
// (src/unstructured_mesh.cpp:979-983); a single case passes infinity.
template <typename S, typename LD>
__device__ __forceinline__ void item_flux(const DevMesh<S> &m, const DevParams &prm, int64_t e, const LD &ld,
                                          double v3g, double dtcap, S F[4]) {
    int o, nb;
    unsigned side;
    S geo[kGeoPerItem];  // n(3), A, dt, r_owner(3), r_neighbour(3)
    if constexpr (std::is_same<S, double>::value) {
        const int4 ix = m.it_x[e];
        o = ix.x;
        nb = ix.y;
        side = (unsigned)ix.z;
        const double4 g0 = m.it_d[2 * e], g1 = m.it_d[2 * e + 1];
        geo[0] = g0.x; geo[1] = g0.y; geo[2] = g0.z; geo[3] = g0.w;
        geo[4] = g1.x;
        const double fx = g1.y, fy = g1.z, fz = g1.w;
        const double4 co = ld.centroid(o);
        geo[5] = co.x - fx;
        geo[6] = co.y - fy;
        geo[7] = co.z - fz;
        if (nb >= 0) {
            const double4 cn = ld.centroid(nb);
            geo[8] = cn.x - fx;
            geo[9] = cn.y - fy;
            geo[10] = cn.z - fz;
        } else {
            geo[8] = m.g0x - fx;
            geo[9] = m.g0y - fy;
            geo[10] = m.g0z - fz;
        }
    } else {
        const int2 ix = m.it_on[e];
        o = ix.x;
        nb = ix.y;
        const float4 g0 = m.it_f[3 * e], g1 = m.it_f[3 * e + 1], g2 = m.it_f[3 * e + 2];
        geo[0] = g0.x; geo[1] = g0.y; geo[2] = g0.z; geo[3] = g0.w;
        geo[4] = g1.x; geo[5] = g1.y; geo[6] = g1.z; geo[7] = g1.w;
        geo[8] = g2.x; geo[9] = g2.y; geo[10] = g2.z;
        side = __float_as_uint(g2.w);
    }
    const S n[3] = {geo[0], geo[1], geo[2]};
    const S A = geo[3], dt = fmin(geo[4], (S)dtcap);
    S po[4], go[12];
    ld.side(o, po, go);
    S pI[4], pN[4], p0[4], gN[12];
    if (nb >= 0 || nb == -1) {
        // interior branch: a real neighbour, or the ghost whose index equals
        // n_cells (the reference's strict "neighbour > n_cells" test, Solver.cpp:861)
        S pn[4];
        if (nb >= 0) {
            ld.side(nb, pn, gN);
        } else {
            const DevPatch P = m.patches[m.b_patch[0]];
            S pp[4], gp[12], fx[4] = {0, 0, 0, 0};
            if (P.flags & kGhostGradFromCell) ld.side(m.b_partner[0], pp, gp);
            if (P.flags & kGhostFixed) {
                const V4<S> f = m.b_fix[0];
                fx[0] = f.x; fx[1] = f.y; fx[2] = f.z; fx[3] = f.w;
            }
            ghost_state<S>(P, po, pp, fx, pn);
            ghost_grad<S>(P, n, m.b_dmag[0], po, go, gp, gN);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            pI[q] = po[q] - (geo[5] * go[3 * q] + geo[6] * go[3 * q + 1] + geo[7] * go[3 * q + 2]);
            pN[q] = pn[q] - (geo[8] * gN[3 * q] + geo[9] * gN[3 * q + 1] + geo[10] * gN[3 * q + 2]);
            p0[q] = (pI[q] + pN[q]) * (S)0.5;
        }
    } else {
        const int b = -1 - nb;
        const DevPatch P = m.patches[m.b_patch[b]];
        S gp[12];
        if (P.flags & kGhostGradFromCell) {
            S pp[4];
            ld.side(m.b_partner[b], pp, gp);
        }
        ghost_grad<S>(P, n, m.b_dmag[b], po, go, gp, gN);
        S fv[4];
        if (P.rho_type == 1) fv[0] = (S)P.rho;
        else if (P.rho_type == 8) fv[0] = po[0];
        else fv[0] = po[0] - (geo[5] * gN[0] + geo[6] * gN[1] + geo[7] * gN[2]);
        if (P.vel_type == 1) {
            fv[1] = (S)P.u;
            fv[2] = (S)P.v;
            fv[3] = (S)P.w;
        } else if (P.vel_type == 8) {
            fv[1] = po[1];
            fv[2] = (S)0;
            fv[3] = po[3];
        } else {
#pragma unroll
            for (int q = 1; q < 4; q++) fv[q] = po[q] - (geo[5] * gN[3 * q] + geo[6] * gN[3 * q + 1] + geo[7] * gN[3 * q + 2]);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) pI[q] = pN[q] = p0[q] = fv[q];
    }
    lattice_flux<std::is_same<S, float>::value, S>(n, A, dt, (S)v3g, side, p0, pI, go, pN, gN, (S)prm.invgam,
                                                   (S)prm.rho0, F);
}

// Residual of cell i from its faces' LDS slots, summed in the reference's face
// order (src/Solver.cpp:353-366).  Slot sl of lane c lives at sF[q*stride + sl*C + c].
template <typename S>
__device__ __forceinline__ void gather_residual(const DevMesh<S> &m, int i, const S *sF, int stride, int C, int c,
                                                double R[4], int B = 1, int b = 0) {
    const double iv = m.inv_vol[i];
    R[0] = R[1] = R[2] = R[3] = 0;
    const int k0 = m.g_off[i], k1 = m.g_off[i + 1];
    for (int k = k0; k < k1; k++) {
        const int ent = m.g_ent[k];
        S v[4];
        if (ent & 2) {
            // a face between two blocks: its flux was evaluated once by the border
            // pass, [4][n_border][B] with the case fastest (B = 1 for one case)
            const size_t j = (size_t)(ent >> 2), nbd = (size_t)m.n_border;
#pragma unroll
            for (int q = 0; q < 4; q++) v[q] = m.bflux[(q * nbd + j) * B + b];
        } else {
            const int sl = (ent >> 2) * C + c;
#pragma unroll
            for (int q = 0; q < 4; q++) v[q] = sF[q * stride + sl];
        }
        if (ent & 1) {
#pragma unroll
            for (int q = 0; q < 4; q++) R[q] = R[q] + (-(double)v[q] * iv);
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) R[q] = R[q] + ((double)v[q] * iv);
        }
    }
    if constexpr (std::is_same<S, float>::value) {
        // rho0/3 sum(+-n A)/V of the cell, omitted from the single-precision face fluxes
        const int N = m.N;
        R[1] += (double)m.corr[i];
        R[2] += (double)m.corr[(size_t)N + i];
        R[3] += (double)m.corr[2 * (size_t)N + i];
    }
}

// RK stage update of one cell record r (cell i) from its residual R
// (src/Solver.cpp:405-438, src/RungeKutta.h:16-18).  The last stage writes the
// new primitive solution, t0 and the next local time step, adds R^2 to sq and
// returns 1 for NaN or 2 for rho blow-up (src/Solver.cpp:448-493).
template <typename S>
__device__ __forceinline__ int rk_update(const DevMesh<S> &m, const DevParams &prm, double visc, StageArgs sa,
                                         int i, size_t r, const double R[4], V4<S> *__restrict__ st_out,
                                         double4 *__restrict__ t0, Acc<S> *__restrict__ t1,
                                         double *__restrict__ dtl, double sq[4]) {
    constexpr bool F32 = std::is_same<S, float>::value;
    const double dti = dtl[r];
    const double4 ta = t0[r];
    const double a0 = ta.x, a1 = ta.y, a2 = ta.z, a3 = ta.w;
    // Solver.cpp:412-416
    double f1 = a0 + R[0] * dti * sa.alpha;
    double f2 = a1 + (R[1] + prm.force) * dti * sa.alpha;
    double f3 = a2 + R[2] * dti * sa.alpha;
    double f4 = a3 + R[3] * dti * sa.alpha;
    f2 = f2 / f1;
    f3 = f3 / f1;
    f4 = f4 / f1;
    // Solver.cpp:424-428
    const double db = dti * sa.beta;
    double b0, b1, b2, b3;
    if (!F32) {
        if (sa.stage == 0) {
            b0 = a0; b1 = a1; b2 = a2; b3 = a3;
        } else {
            const Acc<S> tb = t1[r];
            b0 = tb.x; b1 = tb.y; b2 = tb.z; b3 = tb.w;
        }
        b0 = b0 + db * R[0];
        b1 = b1 + db * (R[1] + prm.force);
        b2 = b2 + db * R[2];
        b3 = b3 + db * R[3];
    } else {
        // increments only; the conservative state is t0 + increments
        double d0 = 0, d1 = 0, d2 = 0, d3 = 0;
        if (sa.stage != 0) {
            const Acc<S> tb = t1[r];
            d0 = tb.x; d1 = tb.y; d2 = tb.z; d3 = tb.w;
        }
        d0 += db * R[0];
        d1 += db * (R[1] + prm.force);
        d2 += db * R[2];
        d3 += db * R[3];
        if (sa.stage < 3) {
            b0 = d0; b1 = d1; b2 = d2; b3 = d3;
        } else {
            b0 = a0 + d0; b1 = a1 + d1; b2 = a2 + d2; b3 = a3 + d3;
        }
    }
    const double r0 = F32 ? prm.rho0 : 0.0;  // single-precision density is stored as rho - rho0
    if (sa.stage < 3) {
        st_out[r] = V4<S>{(S)(f1 - r0), (S)f2, (S)f3, (S)f4};
        using E = typename std::conditional<F32, float, double>::type;
        t1[r] = Acc<S>{(E)b0, (E)b1, (E)b2, (E)b3};
        return 0;
    }
    // soln = t1 as primitives (Solver.cpp:431-436); it is the next step's temp
    const double s0 = b0, s1 = b1 / b0, s2 = b2 / b0, s3 = b3 / b0;
    st_out[r] = V4<S>{(S)(s0 - r0), (S)s1, (S)s2, (S)s3};
    t0[r] = double4{b0, b1, b2, b3};
    if (prm.time_stepping == PLBFS_DT_CONSTANT) dtl[r] = prm.cfl;
    else dtl[r] = local_dt(prm, visc, s0, s1, m.cfl_area[i], m.vol[i]);
    sq[0] += R[0] * R[0];
    sq[1] += R[1] * R[1];
    sq[2] += R[2] * R[2];
    sq[3] += R[3] * R[3];
    if (isnan(f1) || isnan(f2)) return 1;
    if (f1 / prm.ref_rho > 1000.0) return 2;
    return 0;
}

}  // namespace plbfs
