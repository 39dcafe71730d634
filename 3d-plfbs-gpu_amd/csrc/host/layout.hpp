// Device data layout of the PLBFS solver, built on the host in FP64.
//
// Cells are renumbered along a 3-D Hilbert curve of their centroids and every
// per-cell array is stored structure-of-arrays in that order, so that a block
// of consecutive threads touches a compact region of the mesh.  Ghost cells
// are never stored: a ghost's state and gradient are functions of its owner
// and its patch, evaluated where needed.
//
// Face work is organised per block of `block_cells` consecutive cells: every
// face with at least one cell in the block becomes a work item of that block,
// its flux goes to a block-local LDS slot, and each cell then sums its own
// faces' slots in the reference's face order (the order of
// src/Solver.cpp:286-366).  A face on the border of two blocks is either
// evaluated by both (duplicated items), or, with split_border, once by a
// separate border pass into a global array that both blocks read.  No float
// atomics are involved.
#pragma once
#include <cstdint>
#include <vector>

#include "plbfs_hip.h"

namespace plbfs {

constexpr int kMaxSlots = 6;  // LSQ slots per cell (the reference's RHS arrays hold 6, gradients.cpp)
constexpr int kGeoPerItem = 11;  // n(3), area, dt, r_owner(3), r_neighbour(3)

struct Layout {
    int N = 0, F = 0, NI = 0, G = 0;
    int block_cells = 256;
    int n_blocks = 0;
    int max_items_per_block = 0;
    int64_t n_items = 0;
    // border items (split_border): items [border0, border0 + n_border) are the
    // faces between two blocks, each evaluated once into a global flux array
    int64_t border0 = 0, n_border = 0;
    std::vector<int32_t> perm;  // new cell -> file cell
    std::vector<int32_t> inv;   // file cell -> new cell
    // LSQ stencil, slot-major: element (s, i) at s*N + i
    std::vector<uint8_t> st_cnt;   // [N]
    std::vector<int32_t> st_nbr;   // [kMaxSlots*N]: new index, or -1-b for the ghost of boundary face b
    std::vector<double> st_w;      // [kMaxSlots*3*N]: folded weights inv(LHS) * RHS_slot, (s*3+c)*N + i
    // per cell
    std::vector<double> vol, inv_vol, cfl_area;  // [N]
    std::vector<double> centroid;                // [3][N] component-major
    // rho0/3 * sum over the cell's faces of -+ n A / V (owner -, neighbour +):
    // the residual of the constant density rho0 at rest, which the FP32
    // deviation-form fluxes leave out (nonzero only where faces are warped)
    std::vector<double> corr;  // [3][N]
    // face work items
    std::vector<int32_t> blk_item_off;  // [n_blocks+1]
    std::vector<int32_t> it_own, it_nbr;  // [E] new owner index; neighbour new index or -1-b
    // [E] upwind side of the 14 moving lattice points: for pair p (directions
    // (1,0,0),(0,1,0),(0,0,1),(1,1,1),(1,1,-1),(1,-1,1),(-1,1,1)) bit 2p is
    // set when +e takes the neighbour side (e.n < 0), bit 2p+1 when -e does
    // (e.n > 0); e.n summed in x,y,z order in FP64 as src/Solver.cpp:958-1273
    std::vector<uint32_t> it_side;
    std::vector<double> it_geo;  // [kGeoPerItem*E], component-major
    std::vector<double> it_fc;   // [3*E] face centroid, component-major
    double ghost0[3] = {0, 0, 0};  // centroid of the ghost of boundary face 0
    // per-cell gather list in file face order, entries (index << 2) | border << 1 | is_owner:
    // index is the block-local slot, or with the border bit the border item's index
    std::vector<int32_t> g_off;  // [N+1]
    std::vector<int32_t> g_ent;
    // boundary faces (b = face - NI)
    std::vector<int32_t> b_patch;  // [G]
    std::vector<double> b_dmag;    // [G] |x_ghost - x_owner|
    std::vector<int32_t> b_owner;  // [G] new owner index
    // [G] cell whose gradient the ghost takes (new index): the periodic partner
    // on periodic patches, cell 0 (file order) on patches whose ghost the
    // reference leaves unwritten (density 4, 5, 9; velocity 4, 5, 7, 9);
    // -1 elsewhere
    std::vector<int32_t> b_partner;
    // [4*G] (rho, u, v, w) of ghosts that keep fixed values: u, v of the
    // parabolic profiles (velocity 4, 5) here; the components the reference
    // leaves at the ghost's initial state come from plbfs_set_ghost_state
    std::vector<double> b_fixed;
    bool needs_ghost_state = false;  // some patch keeps initial ghost values
    // LSQ inverse and RHS in the reference layout, file order (for tests)
    std::vector<double> lsq_inv9;  // [N*9]
    std::vector<double> lsq_rhs;   // [N*6*3]
};

// Boundary types whose ghost the reference's update leaves (partly) at its
// initial value, and whose ghost gradient is then the periodic node's: cell 0
// (src/Solution.cpp:759-833, src/gradients.cpp:434-440,488-502).
inline bool rho_keeps_ghost(int t) { return t == 4 || t == 5 || t == 9; }
inline bool vel_keeps_ghost(int t) { return t == 4 || t == 5 || t == 7 || t == 9; }

// Builds the layout.  Returns PLBFS_OK or an error code (message via plbfs_fail).
int build_layout(const plbfs_mesh_view &m, const plbfs_params &p, const plbfs_patch *patches, int n_patches,
                 const int32_t *bface_patch, int block_cells, bool split_border, Layout &L);

// 3-D Hilbert key of integer coordinates with `bits` bits per axis.
uint64_t hilbert_key3(uint32_t x, uint32_t y, uint32_t z, int bits);

}  // namespace plbfs
