// Internal layout of the statics_hip library.
//
// Device layout (one set of arrays for all structures of a context; a batch
// concatenates them, structure after structure):
//   bricks   8^3 cells each, numbered in Morton order within a structure; only
//            bricks holding a free cell exist. Per brick: the range of its free
//            cells [cbase[b], cbase[b+1]) and of its nodes [nbase[b], nbase[b+1]).
//   cells    the free cells, brick-major, slot order within a brick (slot
//            ((z&7)*8 + (y&7))*8 + (x&7), lattice/src/lib.rs:310-314). Per cell:
//            e, g (FP64 and FP32), the live-beam mask, its six face nodes
//            (VOXEL_FACES order; bit 31 set where the node is not free), and its
//            coordinates.
//   nodes    every face of a free cell, brick-major. A face lies between cells
//            c- and c+ = c- + e_o; it is stored with c+ ("lower" slot of c+) when
//            c+ is free, else with c- ("upper" slot of c-). So every node is kept
//            by a free cell next to it: that cell is its key in the octree.
//   vectors  6 values per node (u, w), zero on held and gone nodes.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/statics_hip.h"
#include "st_beam.hpp"

namespace st {

constexpr uint32_t kNotFree = 0x80000000u;  // in a cell's face list: held or gone node
constexpr uint8_t kHeld = 1, kGone = 2;

struct StError {
  int code;
  std::string what;
};

#define ST_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw ::st::StError{e_ == hipErrorOutOfMemory ? ST_OOM : ST_DEVICE,                      \
                          std::string(#call) + ": " + hipGetErrorString(e_)};                 \
  } while (0)

// Device bytes the library's buffers hold in this process, and the most they held since
// the last reset (StStats.mem_peak).
struct MemTally {
  std::atomic<int64_t> held{0}, peak{0};
  void add(int64_t b) {
    const int64_t now = held.fetch_add(b) + b;
    int64_t p = peak.load();
    while (now > p && !peak.compare_exchange_weak(p, now)) {
    }
  }
  void reset_peak() { peak.store(held.load()); }
};
inline MemTally& mem_tally() {
  static MemTally t;
  return t;
}

// A device buffer that grows on demand and is never shrunk (freed at context end).
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t n) {
    if (n <= bytes) return;
    release();
    const size_t want = n < 256 ? 256 : n;
    ST_HIP(hipMalloc(&p, want));
    bytes = want;
    mem_tally().add((int64_t)want);
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      mem_tally().add(-(int64_t)bytes);
    }
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Host-side description of one uploaded structure (its share of the device arrays).
struct StructInfo {
  int64_t brick0, nbricks;
  int64_t cell0, ncells;
  int64_t node0, nnodes;
  int32_t origin[3];  // lattice coordinates of local (0,0,0)
  int32_t ext[3];     // extent in cells (local coordinates lie in [0, ext))
  int32_t levels_cap; // ceil(log2(max extent))
  double spacing;
};

// What the host builds from cells (st_upload); copied to the device as is.
struct HostGeo {
  std::vector<uint32_t> cbase, nbase;         // per brick + 1
  std::vector<uint32_t> bstruct;              // per brick: structure
  std::vector<double> ce, cg;                 // per free cell
  std::vector<double> csec;                   // 15 per free cell (beam sections over SECTION s^2),
                                              // or empty: every beam SECTION s^2
  std::vector<uint16_t> cmask;                // per free cell, alive beams
  std::vector<uint32_t> cfaces;               // 6 per free cell
  std::vector<uint32_t> ccoord;               // per free cell: x | y << 10 | z << 20 (local)
  std::vector<uint32_t> cstruct;              // per free cell: structure
  std::vector<int64_t> cell_of_input;         // per input cell: compact id or -1 (held)
  std::vector<uint8_t> nflags;                // per node: kHeld
  std::vector<int32_t> ncell;                 // 2 per node: c-, c+ (compact ids, -1 absent)
  std::vector<uint8_t> nkind;                 // per node: 2*o + side (side 1: stored with c-)
  std::vector<uint32_t> nkey;                 // per node: its key cell (the cell keeping it), packed as ccoord
  std::vector<int32_t> nface;                 // 4 per node: o, x, y, z (+ side cell, lattice units)
  std::vector<StructInfo> structs;
  void clear() { *this = HostGeo(); }
};

// Device copies of the geometry.
struct DevGeo {
  int64_t nbricks = 0, ncells = 0, nnodes = 0;
  int32_t nstruct = 0;
  DBuf cbase, nbase, bstruct;
  DBuf blist;                 // bricks by node-count class (the cell-by-cell operator's launches)
  int64_t nclass[2] = {0, 0};  // bricks per class
  uint32_t class_nodes[2] = {0, 0};  // largest node count in each class
  DBuf ce64, cg64, ce32, cg32, cmask_user, cmask, cfaces, ccoord, cstruct;
  DBuf csec64, csec32;  // 15 per free cell when the cells gave beam sections
  bool has_sec = false;
  DBuf nflags, ncell, nkind, nkey;
  DBuf bgeo;  // BeamGeo per structure
};

// Schedules of the fine operator: the kind (0: node by node with the brick's
// 768 nodes / 256 cells staged in LDS, 1: node by node from memory, 2: cell by cell
// with LDS rows per node, 3: node by node from memory with both cells' loads issued
// together), whether bricks are dealt to XCDs in contiguous runs, and the
// workgroup's threads (fewer threads: more workgroups, each a brick, in flight per CU).
struct ApplySchedule {
  int tile, xcd, threads;
};
// (64- and 128-thread workgroups were measured at 256^3 and never won: 0.58-1.38 ms against
// 0.58 (FP32) and 0.93 ms (FP64) with 256 threads.)
constexpr ApplySchedule kSchedules[] = {{0, 1, 256}, {1, 1, 256}, {2, 1, 256}, {0, 0, 256},
                                        {1, 0, 256}, {2, 0, 256}, {3, 1, 256}, {3, 0, 256}};
constexpr int kNumSchedules = (int)(sizeof(kSchedules) / sizeof(kSchedules[0]));

// A coarse block in FP16 for the mixed V-cycle: the 36 entries (row-major) over the
// block's scale (its largest entry maps to 2^14), the scale, and the block's column, in
// 80 bytes (five aligned 16-byte loads).
struct alignas(16) HalfBlock {
  _Float16 a[36];
  float scale;
  int col;
};
static_assert(sizeof(HalfBlock) == 80, "HalfBlock is five 16-byte words");

// A block of a prolongator in FP32 with the index of the other end of its row or column
// (160 bytes: ten aligned 16-byte loads).
struct alignas(16) PBlock32 {
  float a[36];
  int idx;
  int pad[3];
};
static_assert(sizeof(PBlock32) == 160, "PBlock32 is ten 16-byte words");

// One coarse level (aggregates of the level below, 6 DOF each), block-sparse.
struct Level {
  int64_t n = 0;        // aggregates
  int64_t nblocks = 0;  // stored blocks incl. diagonal
  int32_t cube = 0;     // octree level of the cubes its aggregates lie in
  DBuf rowptr;          // n+1 (int32): blocks of row I at [rowptr[I], rowptr[I+1]); diagonal first
  DBuf col;             // nblocks (int32)
  DBuf a64, a32;        // 36 per block, row-major
  DBuf hblk;            // the FP16 copy of the blocks (HalfBlock records, when half)
  bool half = false;    // the mixed V-cycle reads the FP16 copy
  DBuf dinv64, dinv32;  // 36 per aggregate
  DBuf centre;          // 3 per aggregate (double, metres, structure-local)
  DBuf parent;          // per aggregate of this level: its aggregate in the next level (-1 none)
  DBuf rep;             // per aggregate: one fine node of it (int32)
  DBuf key;             // per aggregate: cube coordinates packed (uint64: struct, x, y, z)
  DBuf kid_ptr, kids;   // children (aggregates of the level below) grouped by this level's aggregate
  // smoothed prolongator from the next level to this one (rows: this level's aggregates),
  // block-sparse by rows, and the same blocks grouped by column (the next level's aggregates)
  int64_t pblocks = 0;
  DBuf prow, pcol, pval64, pval32, pt_ptr, pt_blk, pblk_row;
  // (mixed, large levels) the FP32 blocks as PBlock32 records in row order (column index
  // inside) and in column order (row index inside), for the V-cycle's transfers; pval32 goes
  DBuf prr, prc;
  bool prh = false;  // prr and prc hold HalfBlock records (FP16 over a per-block scale)
  DBuf agg_struct;      // per aggregate: structure
  int layout32 = 0, layout64 = 0;  // the residual kernel's tuned layout per precision (see k_c_resid)
  std::vector<double> step;  // smoother step per structure
  DBuf step_d;               // per structure (double)
  // V-cycle vectors at this level
  DBuf r64, x64, t64, u64, v64, r32, x32, t32, u32, v32;
  std::vector<int64_t> seg;  // per structure: its aggregates [seg[s], seg[s+1])
  // dense top: per structure the inverse of its (6 n_s)^2 operator, FP64, symmetric
  bool dense = false;
  int32_t dense_mmax = 0;
  DBuf dense_a, dense_off, dense_seg;  // matrices; per structure offset (int64, ns+1); seg (int64)
};

// A lattice as the reference holds it (nodes, then bonds; materials six values each)
struct GatherOut {
  uint32_t level = 0;
  std::vector<double> x, m, node_mat, p, n, a, d, bond_mat;
  std::vector<uint8_t> held;
  std::vector<int64_t> around, ij;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  HostGeo host;
  DevGeo geo;
  bool uploaded = false;
  std::vector<uint8_t> gone_host;  // per node, last set
  // fine-level smoother
  DBuf dinv64, dinv32;              // 21 per node (upper triangle), zero for non-free nodes
  std::vector<double> step0;        // per structure
  DBuf step0_d;
  // fine-node aggregation (level 1) and the union-find scratch
  DBuf label1, uf, uf2, flag, scan_tmp, sort_tmp;
  DBuf nodes_by_agg, agg_ptr;
  // per brick, its level-1 aggregates [brick_agg1[b], brick_agg1[b + 1]) (int32, nbricks + 1);
  // restrict_fused: every level-1 aggregate lies in one brick, so the operator's epilogue
  // restricts the residual brick by brick
  DBuf brick_agg1;
  bool restrict_fused = false;
  // smoothed prolongator from level 1 to the fine nodes (rows: nodes), used to build level 1
  DBuf p0row, p0col, p0val, p0t_ptr, p0t_blk, p0blk_row;
  int64_t p0blocks = 0;
  bool smoothed = true;  // smoothed aggregation (false: rigid aggregates only)
  int sa_levels = 2;     // transfers smoothed, from the finest up
  int nu = 1;            // fine smoothing steps before and after the coarse correction
  int nu1 = 1;           // the same at level 1
  int nu_c = 1;          // the same above level 1
  int gamma = 1;         // coarse corrections per visit above level 1 (1: V-cycle, 2: W-cycle)
  int gamma_from = 1;    // levels above this index (0: level 1) are solved by gamma cycles
  int gamma_to = 1000;   // ... up to this index
  std::vector<int> gammas;  // if set: cycles per visit of each level index (1 beyond the list);
                            // if empty, gamma on the indices (gamma_from, gamma_to]
  int f64_from = 4;      // first level (0: level 1) kept in FP64 by the mixed V-cycle
  bool fold_switch = true;  // the switch to FP64 folded into the restriction and prolongation
  bool p_half = true;       // the V-cycle's prolongator records in FP16
  bool half_coarse = true;  // the mixed V-cycle's largest coarse operators stored in FP16
  bool graphs = false;   // replay each CG iteration's tail as a captured graph
  bool dense_top = true; // invert a small coarse level densely and drop the levels above
  int apply_sched[2] = {0, 0};  // operator schedule per precision (FP32, FP64), see kSchedules
  std::map<const void*, size_t> lds_attr;  // per kernel: the dynamic LDS limit this context set
  bool tuned = false;           // apply_sched chosen for the uploaded structures
  std::vector<Level> levels;        // coarse levels 1..
  // the coarse residual layout chosen per (level index, FP32, aggregates, blocks), kept
  // for the hierarchies rebuilt on this upload (cleared by st_upload)
  std::vector<std::pair<std::array<int64_t, 4>, int>> coarse_layouts;
  // the Galerkin product's task layout (1: by pairs, 0: grouped by column) per (level
  // index, row width class, 0, 0), kept for the context's life (a choice of speed only)
  std::vector<std::pair<std::array<int64_t, 4>, int>> gal_modes;
  // The hierarchy (fine smoother and coarse levels) is kept between solves of the same
  // operator: anything that changes the operator clears hier_valid; hier_key holds the
  // settings it was built with.
  struct HierKey {
    bool mixed, smoothed, dense_top;
    int sa_levels, f64_from;
    bool half;
    bool operator==(const HierKey& o) const {
      return mixed == o.mixed && smoothed == o.smoothed && dense_top == o.dense_top && sa_levels == o.sa_levels &&
             f64_from == o.f64_from && half == o.half;
    }
  };
  bool hier_valid = false;
  HierKey hier_key{};
  // solver vectors
  DBuf f64, x64, r64, p64, ap64, z64, t64, u64, w64;
  DBuf x32, r32, p32, ap32, z32, t32, w32;
  DBuf partial, scal;               // reductions: per-block partials, per-structure scalars
  DBuf done;                        // per structure: 1 once converged
  DBuf fres, xsol;                  // resident loads and the last solution (FP64, 6 per node)
  bool loads_set = false;
  bool have_solution = false;
  StSolveOpts last_opts{};
  std::vector<int64_t> dof_per_struct;  // free nodes per structure, from the last setup
  int64_t last_ndof = 0;
  // per-launch timing of one kernel (st_kernel_timing_enable)
  int timing_which = -1;
  std::vector<hipEvent_t> tev;      // pairs
  size_t tev_used = 0;
  double timing_bytes = 0.0;
  double timing_ms_done = 0.0;
  int64_t timing_launches = 0;
  // the last gathered lattice, kept until read, and each material's full-cell measure
  // (rho, e, g, ft, fc, fs, then section and strength ratio of both beam classes)
  GatherOut gathered;
  std::vector<std::array<double, 10>> full_known;
};

// host upload (st_upload.cpp)
void build_host_geo(HostGeo& hg, const StCells* const* cells, int32_t n);
void copy_geo_to_device(Ctx& c);

}  // namespace st
