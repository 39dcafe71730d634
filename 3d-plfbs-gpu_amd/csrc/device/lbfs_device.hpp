// Device-side data structures and numerics of the PLBFS solver (gfx950).
//
// Numerics follow the reference's CPU solver (src/Solver.cpp); the
// per-face equilibrium-flux evaluation is reorganised for registers:
// instead of materialising the 15 lattice populations (src/Solver.cpp:286-347,
// 655-837), each lattice point is reconstructed, its equilibrium evaluated and
// folded into running moments S0 = sum feq, S1 = sum e feq, Pi = sum e e^T feq.
// With f_k = feq_k(I) - tau (feq_k(I) - feq_k) and the D3Q15 isotropy
// sum w e_a e_b = I/3, sum w e_a e_b e_c e_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc)/9,
// the reference's flux sums are exactly
//   sum_k e_k f_k       = (1 - tau) rho_I u_I + tau S1      (= S1, since rho_I u_I = S1)
//   sum_k e_k e_k^T f_k = (1 - tau) rho_I (I/3 + u_I u_I^T / gamma) + tau Pi
// so only S0, S1 and Pi are needed.  They are accumulated as unweighted
// pair sums per lattice weight class (13 values in registers) and weighted
// once per face.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plbfs {

// flags: the ghost gradient copies cell b_partner[b]'s (periodic partner, or
// cell 0 where the reference's update leaves the ghost unwritten), and the
// ghost state holds per-ghost fixed values b_fix[b] (those patches)
constexpr int32_t kGhostGradFromCell = 1;
constexpr int32_t kGhostFixed = 2;

struct DevPatch {
    int32_t rho_type, vel_type, is_empty, flags;
    double rho, u, v, w;
};

// Step control block, written by the reduction kernel, read by every kernel.
struct Ctrl {
    int32_t done;    // nonzero: skip all further work
    int32_t status;  // 0 ok, PLBFS_ERR_NAN, PLBFS_ERR_RHO, PLBFS_ERR_IBM
    int32_t step;    // steps completed in this run
    int32_t converged;
    int32_t rebuild;  // nonzero: an immersed node drifted half its list margin
    double res[4];
};

// Per-cell records gathered by neighbours are stored array-of-structures so a
// gather touches one 32 B (state) or 96 B (gradient) run instead of 4 or 12
// separate cache lines.
template <typename S>
struct V4T;
template <>
struct V4T<double> {
    using type = double4;
};
template <>
struct V4T<float> {
    using type = float4;
};
template <typename S>
using V4 = typename V4T<S>::type;

// Gradient records gathered by the face items.  FP64: three V4 (d/dx, d/dy,
// d/dz of rho, u, v, w), 96 B.  FP32: the cell's state followed by its three
// gradient V4, 64 B, so that a face side is one aligned 64 B run; the
// gradient kernel writes all four (the state it read), so its stores stay
// contiguous.
template <typename S>
inline constexpr int kGradRec = std::is_same<S, float>::value ? 4 : 3;  // V4 per record
template <typename S>
inline constexpr int kGradOff = std::is_same<S, float>::value ? 1 : 0;  // first gradient V4

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx.x mod 8), and
// each XCD has its own L2.  Remapping gives XCD x a contiguous range of logical
// blocks, so Hilbert-adjacent cells (and their shared halo) meet in one L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int bid, int nblk, int remap) {
    if (!remap) return bid;
    const int q = nblk / kXcds, r = nblk % kXcds;
    const int x = bid % kXcds, k = bid / kXcds;
    return x * q + (x < r ? x : r) + k;
}

template <typename S>
struct DevMesh {
    int N, G, n_blocks, max_items;
    // border items [border0, border0 + n_border) and their fluxes [4][n_border]
    // (one evaluation per face between two blocks; 0 when items are duplicated)
    int64_t border0;
    int n_border;
    S *bflux;
    int xcd_remap;
    int block_cells;  // cells per flux block (a multiple of 64)
    int stage;        // 1: flux blocks stage their own cells' records in LDS
    int grad_stage;   // 1: gradient blocks stage their own cells' states in LDS
    int64_t E;
    const uint8_t *st_cnt;
    const int32_t *st_nbr;
    const S *st_w;
    const double *vol, *inv_vol, *cfl_area;
    const int32_t *blk_item_off;
    // Face work items, one record each so that an item is a few wide loads
    // from one address.  FP64 storage: it_x[e] = (owner, neighbour or -1-b,
    // lattice side bits (see lattice_pair), 0) and it_d[2e..2e+1] = (n, A),
    // (dt, x_face); the centroid-to-face offsets are formed on the fly from cgeo
    // (same bits as the host's cc - fc).  FP32 storage (56 B per item):
    // it_on[e] = (owner, neighbour or -1-b) and it_f[3e..3e+2] = (n, A),
    // (dt, r_owner), (r_neighbour, side bits stored in the float's bits).
    const int4 *it_x;
    const double4 *it_d;
    const int2 *it_on;
    const float4 *it_f;
    const double4 *cgeo;  // FP64 storage: per cell (cx, cy, cz, 1/V)
    double g0x, g0y, g0z;  // centroid of the ghost of boundary face 0
    const int32_t *g_off, *g_ent;
    // [n_blocks] nonzero where a flux block holds a cell the immersed boundary
    // spreads force to; null: every block reads the body force
    const int32_t *blk_force;
    const int32_t *b_patch;
    const int32_t *b_partner;  // cell whose gradient the ghost takes (kGhostGradFromCell), -1 elsewhere
    const V4<S> *b_fix;        // [G] fixed ghost states (kGhostFixed); FP32: rho - rho0
    const S *b_dmag;
    const DevPatch *patches;  // FP32 storage: rho values shifted by -rho0
    const S *corr;            // FP32 storage: [3][N] rho0/3 sum(+-n A)/V per cell
};

struct DevParams {
    double visc, gamma, invgam, visc3_invgam, force, cfl, ceff, ref_rho, ref_mom, tol;
    double rho0;  // FP32 storage: density offset (rho = rho0 + stored value)
    int time_stepping;
    int use_body_force;
};

// ---------------------------------------------------------------------------
// Ghost cells (src/Solution.cpp:759-833, src/gradients.cpp:395-504,592-648)
// ---------------------------------------------------------------------------

// pp: state of the periodic partner (read only by type 3); fx: the ghost's
// fixed values (read only by the types the reference's update leaves at the
// ghost's initial state, and the parabolic u, v of types 4 and 5).
template <typename T>
__device__ __forceinline__ void ghost_state(const DevPatch &P, const T po[4], const T pp[4], const T fx[4],
                                            T pg[4]) {
    const int rt = P.rho_type, vt = P.vel_type;
    if (rt == 1) pg[0] = (T)P.rho - (po[0] - (T)P.rho);
    else if (rt == 3) pg[0] = pp[0];
    else if (rt == 6 || rt == 7) pg[0] = (T)P.rho;
    else if (rt == 4 || rt == 5 || rt == 9) pg[0] = fx[0];
    else pg[0] = po[0];  // 2 (neumann, dx = 0) and 8 (symmetry)
    if (vt == 4 || vt == 5 || vt == 7 || vt == 9) {
        pg[1] = fx[1];
        pg[2] = fx[2];
        pg[3] = fx[3];
    } else if (vt == 1) {
        pg[1] = (T)P.u - (po[1] - (T)P.u);
        pg[2] = (T)P.v - (po[2] - (T)P.v);
        pg[3] = (T)P.w - (po[3] - (T)P.w);
    } else if (vt == 6) {
        pg[1] = (T)P.u;
        pg[2] = (T)P.v;
        pg[3] = (T)P.w;
    } else if (vt == 8) {
        pg[1] = po[1];
        pg[2] = -po[2];
        pg[3] = po[3];
    } else if (vt == 3) {
        pg[1] = pp[1];
        pg[2] = pp[2];
        pg[3] = pp[3];
    } else {
        pg[1] = po[1];
        pg[2] = po[2];
        pg[3] = po[3];
    }
}

// Ghost gradient from the owner's state and gradient (gp: the gradient of
// cell b_partner[b], read by type 3 and by the types whose ghost the
// reference's update leaves unwritten).  rho_type 8 leaves the reference's
// zero-initialised rho slot untouched, so it is zero here too.
template <typename T>
__device__ __forceinline__ void ghost_grad(const DevPatch &P, const T n[3], T dmag, const T po[4], const T go[12],
                                           const T gp[12], T gg[12]) {
    const int rt = P.rho_type, vt = P.vel_type;
    if (rt == 1) {
        const T q = (T)2 * ((T)P.rho - po[0]) / dmag;
        gg[0] = q * n[0];
        gg[1] = q * n[1];
        gg[2] = q * n[2];
    } else if (rt == 2) {
        gg[0] = go[0] * ((T)1 - n[0]);
        gg[1] = go[1] * ((T)1 - n[1]);
        gg[2] = go[2] * ((T)1 - n[2]);
    } else if (rt == 7) {
        gg[0] = go[0];
        gg[1] = go[1];
        gg[2] = go[2];
    } else if (rt == 3 || rt == 4 || rt == 5 || rt == 9) {
        gg[0] = gp[0];
        gg[1] = gp[1];
        gg[2] = gp[2];
    } else {
        gg[0] = gg[1] = gg[2] = (T)0;
    }
    if (vt == 1 || vt == 8) {
        const T qu = (T)2 * ((T)P.u - po[1]) / dmag;
        const T qv = (T)2 * ((T)P.v - po[2]) / dmag;
        const T qw = (T)2 * ((T)P.w - po[3]) / dmag;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            gg[3 + c] = qu * n[c];
            gg[6 + c] = qv * n[c];
            gg[9 + c] = qw * n[c];
        }
    } else if (vt == 2) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            gg[3 + c] = go[3 + c] * ((T)1 - n[c]);
            gg[6 + c] = go[6 + c] * ((T)1 - n[c]);
            gg[9 + c] = go[9 + c] * ((T)1 - n[c]);
        }
    } else if (vt == 3 || vt == 4 || vt == 5 || vt == 7 || vt == 9) {
#pragma unroll
        for (int q = 3; q < 12; q++) gg[q] = gp[q];
    } else {
#pragma unroll
        for (int q = 3; q < 12; q++) gg[q] = (T)0;
    }
}

// ---------------------------------------------------------------------------
// Lattice evaluation (src/Solver.cpp:840-1374 restructured as moments)
// ---------------------------------------------------------------------------

// Unweighted pair sums h = feq / w.  Each axis pair (weight 1/9) feeds
// exactly one component of each moment, so its sum S and difference D are kept
// as they are; the four body diagonals (weight 1/72) all feed S0 and every
// diagonal of Pi with the same sum, which is therefore accumulated once.  The
// weights are applied once per face.
template <typename T>
struct PairSums {
    T ax_s[3], ax_d[3];           // axis pairs x, y, z: S and D
    T dg_s, dg_x, dg_y, dg_z;     // diagonals: sum S, sum e_a D
    T dg_xy, dg_xz, dg_yz;        // diagonals: sum e_a e_b S
};

// Equilibrium of one lattice point e at state (r, u, v, w), divided by its
// weight.  With DEV (FP32 storage) the density is a deviation r = rho - rho0
// and the value returned is h = (feq - w rho0) / w = r (1 + X) + rho0 X,
// X = 3 e.u + (4.5 (e.u)^2 - 1.5 u^2)/gamma, so the O(1) part rho0 never enters
// single precision.
template <int EX, int EY, int EZ, bool DEV, typename T>
__device__ __forceinline__ T pop_eq(T r, T u, T v, T w, T c45, T c15, T rho0) {
    const T eu = (T)EX * u + (T)EY * v + (T)EZ * w;
    const T usq = u * u + v * v + w * w;
    // X = 3 e.u + (4.5 (e.u)^2 - 1.5 u^2) / gamma, as e.u (3 + c45 e.u) - c15 u^2
    // with c45 = 4.5 / gamma and c15 = 1.5 / gamma
    const T X = eu * ((T)3 + c45 * eu) - c15 * usq;
    if constexpr (DEV) return r + (r + rho0) * X;
    else return r + r * X;
}

// e.g of one lattice direction from a gradient triple g (x, y, z) and its sum
// s = gx + gy + gz: the four body diagonals share s.
template <int EX, int EY, int EZ, typename T>
__device__ __forceinline__ T dir_deriv(const T *g, T s) {
    if constexpr (EX != 0 && EY != 0 && EZ != 0) {
        // EX, EY, EZ in {+-1}: at most one of them is -1 in kLatticePairs' order
        if constexpr (EX < 0) return s - (T)2 * g[0];
        else if constexpr (EY < 0) return s - (T)2 * g[1];
        else if constexpr (EZ < 0) return s - (T)2 * g[2];
        else return s;
    } else {
        return (T)EX * g[0] + (T)EY * g[1] + (T)EZ * g[2];
    }
}

// Lattice pair +e / -e (pair index P, order of kLatticePairs).  Point +e takes
// the neighbour side iff e.n < 0 and point -e iff e.n > 0 (the reference's "sum
// of signed normal components > 0" tests, src/Solver.cpp:958-1273).  The two
// decisions per pair are precomputed on the host in FP64 (bits 2P and 2P+1 of
// `side`): for faces containing a lattice direction, e.n is zero analytically
// and its sign is rounding noise, so evaluating it in another precision would
// pick the other side.  The pair enters the moments through its sum
// S = h(+e) + h(-e) (even moments) and difference D = h(+e) - h(-e) (odd ones).
template <int P, int EX, int EY, int EZ, bool DEV, typename T>
__device__ __forceinline__ void lattice_pair(PairSums<T> &M, unsigned side, T dt, const T pI[4], const T gO[12],
                                             const T sO[4], const T pN[4], const T gN[12], const T sN[4], T c45,
                                             T c15, T rho0) {
    const bool up = (side >> (2 * P)) & 1u;      // +e point from the neighbour side
    const bool dn = (side >> (2 * P + 1)) & 1u;  // -e point from the neighbour side
    T a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const T po = dir_deriv<EX, EY, EZ, T>(gO + 3 * q, sO[q]);
        const T pn = dir_deriv<EX, EY, EZ, T>(gN + 3 * q, sN[q]);
        a[q] = up ? (pN[q] - pn * dt) : (pI[q] - po * dt);
        b[q] = dn ? (pN[q] + pn * dt) : (pI[q] + po * dt);
    }
    const T fp = pop_eq<EX, EY, EZ, DEV, T>(a[0], a[1], a[2], a[3], c45, c15, rho0);
    const T fm = pop_eq<-EX, -EY, -EZ, DEV, T>(b[0], b[1], b[2], b[3], c45, c15, rho0);
    const T S = fp + fm, D = fp - fm;
    if constexpr (EX == 0 || EY == 0 || EZ == 0) {
        constexpr int ax = EX ? 0 : (EY ? 1 : 2);
        M.ax_s[ax] = S;
        M.ax_d[ax] = D;
    } else if constexpr (P == 3) {
        // the first diagonal (+1, +1, +1) starts the sums
        static_assert(EX == 1 && EY == 1 && EZ == 1, "pair 3 is the (+1, +1, +1) diagonal");
        M.dg_s = S;
        M.dg_x = D;
        M.dg_y = D;
        M.dg_z = D;
        M.dg_xy = S;
        M.dg_xz = S;
        M.dg_yz = S;
    } else {
        M.dg_s += S;
        M.dg_x += (T)EX * D;
        M.dg_y += (T)EY * D;
        M.dg_z += (T)EZ * D;
        M.dg_xy += (T)(EX * EY) * S;
        M.dg_xz += (T)(EX * EZ) * S;
        M.dg_yz += (T)(EY * EZ) * S;
    }
}

// Reciprocal: correctly rounded in FP64; in FP32 the hardware estimate
// (1 ulp), whose error is far below the FP32 variant's tolerance.
template <typename T>
__device__ __forceinline__ T lattice_rcp(T x) {
    if constexpr (std::is_same<T, float>::value) return __builtin_amdgcn_rcpf(x);
    else return (T)1 / x;
}

// Flux through a face of area A and unit normal n (owner -> neighbour); tau
// is passed as its parts v3g / dt.  With DEV the momentum flux omits the
// constant rho0/3 n A, whose sum over a cell's faces is a per-cell geometric
// constant added separately.
template <bool DEV, typename T>
__device__ __forceinline__ void lattice_flux(const T n[3], T A, T dt, T v3g, unsigned side, const T p0[4],
                                             const T pI[4], const T gO[12], const T pN[4], const T gN[12], T invgam,
                                             T rho0, T F[4]) {
    PairSums<T> M;
    const T w0 = (T)(2.0 / 9.0), w1 = (T)(1.0 / 9.0), w7 = (T)(1.0 / 72.0);
    const T c45 = (T)4.5 * invgam, c15 = (T)1.5 * invgam;
    T sO[4], sN[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        sO[q] = gO[3 * q] + gO[3 * q + 1] + gO[3 * q + 2];
        sN[q] = gN[3 * q] + gN[3 * q + 1] + gN[3 * q + 2];
    }
    const T h0 = pop_eq<0, 0, 0, DEV, T>(p0[0], p0[1], p0[2], p0[3], c45, c15, rho0);
    lattice_pair<0, 1, 0, 0, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    lattice_pair<1, 0, 1, 0, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    lattice_pair<2, 0, 0, 1, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    lattice_pair<3, 1, 1, 1, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    lattice_pair<4, 1, 1, -1, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    lattice_pair<5, 1, -1, 1, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    lattice_pair<6, -1, 1, 1, DEV, T>(M, side, dt, pI, gO, sO, pN, gN, sN, c45, c15, rho0);
    // weighted moments: S0, S1 and the diagonal of Pi
    const T dS = w7 * M.dg_s;
    const T s0 = w0 * h0 + w1 * (M.ax_s[0] + M.ax_s[1] + M.ax_s[2]) + dS;
    const T sx = w1 * M.ax_d[0] + w7 * M.dg_x;
    const T sy = w1 * M.ax_d[1] + w7 * M.dg_y;
    const T sz = w1 * M.ax_d[2] + w7 * M.dg_z;
    const T pxx = w1 * M.ax_s[0] + dS, pyy = w1 * M.ax_s[1] + dS, pzz = w1 * M.ax_s[2] + dS;
    const T pxy = w7 * M.dg_xy, pxz = w7 * M.dg_xz, pyz = w7 * M.dg_yz;
    const T tau = v3g * lattice_rcp(dt);
    const T rI = DEV ? rho0 + s0 : s0;
    const T ir = lattice_rcp(rI);
    const T ux = sx * ir, uy = sy * ir, uz = sz * ir;
    const T om = (T)1 - tau;
    // sum e e^T feq(I) = rho_I (I/3 + u u^T/gamma); in DEV form rho0 I/3 is dropped
    const T iso = DEV ? s0 * (T)(1.0 / 3.0) : rI * (T)(1.0 / 3.0);
    const T k = om * rI;
    const T mxx = om * iso + k * (ux * ux * invgam) + tau * pxx;
    const T myy = om * iso + k * (uy * uy * invgam) + tau * pyy;
    const T mzz = om * iso + k * (uz * uz * invgam) + tau * pzz;
    const T mxy = k * (ux * uy * invgam) + tau * pxy;
    const T mxz = k * (ux * uz * invgam) + tau * pxz;
    const T myz = k * (uy * uz * invgam) + tau * pyz;
    F[0] = (sx * n[0] + sy * n[1] + sz * n[2]) * A;
    F[1] = (mxx * n[0] + mxy * n[1] + mxz * n[2]) * A;
    F[2] = (mxy * n[0] + myy * n[1] + myz * n[2]) * A;
    F[3] = (mxz * n[0] + myz * n[1] + mzz * n[2]) * A;
}

}  // namespace plbfs
