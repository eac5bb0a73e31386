// Host side of st_upload: cells -> 8^3 bricks, node numbering, face tables.
// Plays the part of lattice::build (statics/src/lattice.rs:124-163): the same
// nodes (every face of a free cell; held iff the cell across is held) and the
// same beams (every face pair of a free cell), but stored per brick and cell
// instead of as node and bond lists.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <numeric>
#include <thread>

#include "st_internal.hpp"

namespace st {
namespace {

inline int popc(uint64_t v) { return __builtin_popcountll(v); }

uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
  };
  return spread(x) | spread(y) << 1 | spread(z) << 2;
}

struct Occ {  // one occupied brick of a structure, before compaction
  uint64_t freem[8] = {}, heldm[8] = {};
  int32_t input[512];
  Occ() { std::fill(std::begin(input), std::end(input), -1); }
};

struct Builder {
  const StCells& C;
  int32_t lo[3], ext[3], nbk[3];
  std::vector<int32_t> btab;  // dense brick grid -> index into occ (or -1)
  std::vector<Occ> occ;

  explicit Builder(const StCells& c) : C(c) {}

  int32_t brick_at(int32_t bx, int32_t by, int32_t bz) const {
    if (bx < 0 || by < 0 || bz < 0 || bx >= nbk[0] || by >= nbk[1] || bz >= nbk[2]) return -1;
    return btab[((size_t)bz * nbk[1] + by) * nbk[0] + bx];
  }
  static int slot_of(int x, int y, int z) { return ((z & 7) * 8 + (y & 7)) * 8 + (x & 7); }
  bool test(int32_t x, int32_t y, int32_t z, bool held) const {
    if (x < 0 || y < 0 || z < 0) return false;
    int32_t b = brick_at(x >> 3, y >> 3, z >> 3);
    if (b < 0) return false;
    int s = slot_of(x, y, z);
    return ((held ? occ[b].heldm : occ[b].freem)[s >> 6] >> (s & 63)) & 1;
  }
  bool is_free(int32_t x, int32_t y, int32_t z) const { return test(x, y, z, false); }
  bool is_held(int32_t x, int32_t y, int32_t z) const { return test(x, y, z, true); }
  int32_t input_at(int32_t x, int32_t y, int32_t z) const {
    int32_t b = brick_at(x >> 3, y >> 3, z >> 3);
    return b < 0 ? -1 : occ[b].input[slot_of(x, y, z)];
  }
};

}  // namespace

namespace {

struct Act {  // an active brick (one with a free cell)
  uint64_t key;
  int32_t occ;
  int32_t bx, by, bz;
};

// One structure's bricks, cell numbering and node masks: everything but the arrays.
struct Prep {
  const StCells* C = nullptr;
  Builder B;
  std::vector<Act> act;
  std::vector<int32_t> act_of_occ;          // occ index -> position in act
  std::vector<uint32_t> cell_base_local;    // per active brick + 1
  std::vector<uint64_t> nmask;              // 48 per active brick: node masks per kind and word
  std::vector<uint32_t> npre;               // 48 per active brick: nodes before each mask word
  std::vector<uint32_t> node_base_local;    // per active brick + 1
  StructInfo info{};
  explicit Prep(const StCells& c) : C(&c), B(c) {}
};

void prepare(Prep& P) {
  const StCells& C = *P.C;
  Builder& B = P.B;
  if (C.n_cells < 0 || (C.n_cells > 0 && (!C.sites || !C.held || !C.e || !C.g)) || !(C.spacing > 0))
    throw StError{ST_BAD_INPUT, "cells: missing arrays, negative count or non-positive spacing"};
  int64_t nc = C.n_cells;
  int32_t mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  for (int64_t i = 0; i < nc; i++)
    for (int k = 0; k < 3; k++) {
      int32_t v = C.sites[3 * i + k];
      if (i == 0 || v < mn[k]) mn[k] = v;
      if (i == 0 || v > mx[k]) mx[k] = v;
    }
  for (int k = 0; k < 3; k++) {
    B.lo[k] = (int32_t)(std::floor(mn[k] / 8.0) * 8);
    B.ext[k] = nc ? mx[k] - B.lo[k] + 1 : 1;
    if (B.ext[k] > 1016) throw StError{ST_BAD_INPUT, "cells: a structure may span at most 1016 cells per axis"};
    B.nbk[k] = (B.ext[k] + 7) / 8;
  }
  B.btab.assign((size_t)B.nbk[0] * B.nbk[1] * B.nbk[2], -1);
  for (int64_t i = 0; i < nc; i++) {
    int32_t x = C.sites[3 * i] - B.lo[0], y = C.sites[3 * i + 1] - B.lo[1], z = C.sites[3 * i + 2] - B.lo[2];
    size_t bi = ((size_t)(z >> 3) * B.nbk[1] + (y >> 3)) * B.nbk[0] + (x >> 3);
    if (B.btab[bi] < 0) {
      B.btab[bi] = (int32_t)B.occ.size();
      B.occ.emplace_back();
    }
    Occ& o = B.occ[B.btab[bi]];
    int s = Builder::slot_of(x, y, z);
    if (o.input[s] >= 0) throw StError{ST_BAD_INPUT, "cells: a site appears twice"};
    o.input[s] = (int32_t)i;
    if (C.held[i])
      o.heldm[s >> 6] |= 1ull << (s & 63);
    else {
      if (!(C.e[i] > 0) || !(C.g[i] > 0) || !std::isfinite(C.e[i]) || !std::isfinite(C.g[i]))
        throw StError{ST_BAD_INPUT, "cells: a free cell has a non-positive or non-finite stiffness"};
      if (C.section)
        for (int q = 0; q < 15; q++)
          if (!(C.section[15 * i + q] > 0) || !std::isfinite(C.section[15 * i + q]))
            throw StError{ST_BAD_INPUT, "cells: a beam section is non-positive or non-finite"};
      o.freem[s >> 6] |= 1ull << (s & 63);
    }
  }
  // Active bricks: those with a free cell, in Morton order.
  auto& act = P.act;
  for (int32_t bz = 0; bz < B.nbk[2]; bz++)
    for (int32_t by = 0; by < B.nbk[1]; by++)
      for (int32_t bx = 0; bx < B.nbk[0]; bx++) {
        int32_t o = B.brick_at(bx, by, bz);
        if (o < 0) continue;
        bool any = false;
        for (int w = 0; w < 8; w++) any |= B.occ[o].freem[w] != 0;
        if (any) act.push_back({morton3(bx, by, bz), o, bx, by, bz});
      }
  std::sort(act.begin(), act.end(), [](const Act& a, const Act& b) { return a.key < b.key; });
  StructInfo& info = P.info;
  info.nbricks = (int64_t)act.size();
  for (int k = 0; k < 3; k++) {
    info.origin[k] = B.lo[k];
    info.ext[k] = B.ext[k];
  }
  int32_t emax = std::max(B.ext[0], std::max(B.ext[1], B.ext[2]));
  info.levels_cap = 0;
  while ((1 << info.levels_cap) < emax) info.levels_cap++;
  info.spacing = C.spacing;
  // Free-cell compact ids: brick order, slot order.
  P.act_of_occ.assign(B.occ.size(), -1);
  for (size_t a = 0; a < act.size(); a++) P.act_of_occ[act[a].occ] = (int32_t)a;
  P.cell_base_local.assign(act.size() + 1, 0);
  for (size_t a = 0; a < act.size(); a++) {
    int cnt = 0;
    for (int w = 0; w < 8; w++) cnt += popc(B.occ[act[a].occ].freem[w]);
    P.cell_base_local[a + 1] = P.cell_base_local[a] + cnt;
  }
  // Node masks per active brick: kind k = 2 o + side.
  P.nmask.assign(act.size() * 48, 0);
  P.npre.assign(act.size() * 48, 0);
  P.node_base_local.assign(act.size() + 1, 0);
  for (size_t a = 0; a < act.size(); a++) {
    const Occ& o = B.occ[act[a].occ];
    uint32_t cnt = 0;
    for (int k = 0; k < 6; k++) {
      int ori = k >> 1, side = k & 1;
      for (int w = 0; w < 8; w++) {
        uint64_t m = 0;
        uint64_t fm = o.freem[w];
        while (fm) {
          int bit = __builtin_ctzll(fm);
          fm &= fm - 1;
          int s = w * 64 + bit;
          int x = act[a].bx * 8 + (s & 7), y = act[a].by * 8 + ((s >> 3) & 7), z = act[a].bz * 8 + (s >> 6);
          if (side == 0) {
            m |= 1ull << bit;
          } else {
            int nx = x + (ori == 0), ny = y + (ori == 1), nz = z + (ori == 2);
            if (!B.is_free(nx, ny, nz)) m |= 1ull << bit;
          }
        }
        P.nmask[a * 48 + k * 8 + w] = m;
        P.npre[a * 48 + k * 8 + w] = cnt;
        cnt += popc(m);
      }
    }
    P.node_base_local[a + 1] = P.node_base_local[a] + cnt;
  }
  info.ncells = P.cell_base_local.back();
  info.nnodes = P.node_base_local.back();
}

// Fills one structure's share of the arrays (sized and offset by the caller).
void fill(const Prep& P, int32_t si, HostGeo& hg, int64_t in_beg) {
  const StCells& C = *P.C;
  const Builder& B = P.B;
  const auto& act = P.act;
  const StructInfo& info = P.info;
  const int64_t c0 = info.cell0, n0 = info.node0;
  auto cell_id = [&](int32_t x, int32_t y, int32_t z) -> int64_t {  // compact id of a free cell
    int32_t o = B.brick_at(x >> 3, y >> 3, z >> 3);
    const int32_t ai = P.act_of_occ[o];
    int s = Builder::slot_of(x, y, z);
    const uint64_t* fm = B.occ[o].freem;
    int r = 0;
    for (int w = 0; w < (s >> 6); w++) r += popc(fm[w]);
    r += popc(fm[s >> 6] & ((1ull << (s & 63)) - 1));
    return c0 + P.cell_base_local[ai] + r;
  };
  auto node_id = [&](int32_t x, int32_t y, int32_t z, int k) -> int64_t {
    int32_t o = B.brick_at(x >> 3, y >> 3, z >> 3);
    int32_t a = P.act_of_occ[o];
    int s = Builder::slot_of(x, y, z);
    uint64_t m = P.nmask[a * 48 + k * 8 + (s >> 6)];
    return n0 + P.node_base_local[a] + P.npre[a * 48 + k * 8 + (s >> 6)] + popc(m & ((1ull << (s & 63)) - 1));
  };
  // The node holding face f of free cell (x, y, z).
  auto face_node = [&](int32_t x, int32_t y, int32_t z, int f) -> int64_t {
    int ori = f >> 1;
    if ((f & 1) == 0) return node_id(x, y, z, 2 * ori);
    int nx = x + (ori == 0), ny = y + (ori == 1), nz = z + (ori == 2);
    if (B.is_free(nx, ny, nz)) return node_id(nx, ny, nz, 2 * ori);
    return node_id(x, y, z, 2 * ori + 1);
  };
  for (size_t a = 0; a < act.size(); a++) {
    const Occ& o = B.occ[act[a].occ];
    const int64_t bidx = info.brick0 + (int64_t)a;
    hg.bstruct[bidx] = (uint32_t)si;
    hg.cbase[bidx + 1] = (uint32_t)(c0 + P.cell_base_local[a + 1]);
    hg.nbase[bidx + 1] = (uint32_t)(n0 + P.node_base_local[a + 1]);
    for (int w = 0; w < 8; w++) {
      uint64_t fm = o.freem[w];
      while (fm) {
        int bit = __builtin_ctzll(fm);
        fm &= fm - 1;
        int s = w * 64 + bit;
        int x = act[a].bx * 8 + (s & 7), y = act[a].by * 8 + ((s >> 3) & 7), z = act[a].bz * 8 + (s >> 6);
        int64_t c = cell_id(x, y, z);
        int32_t in = o.input[s];
        hg.cell_of_input[in_beg + in] = c;
        hg.ce[c] = C.e[in];
        hg.cg[c] = C.g[in];
        if (C.section)
          for (int q = 0; q < 15; q++) hg.csec[15 * c + q] = C.section[15 * in + q];
        hg.cmask[c] = C.beams ? (uint16_t)(C.beams[in] & 0x7fff) : (uint16_t)0x7fff;
        hg.ccoord[c] = (uint32_t)x | (uint32_t)y << 10 | (uint32_t)z << 20;
        hg.cstruct[c] = (uint32_t)si;
        for (int f = 0; f < 6; f++) {
          int64_t v = face_node(x, y, z, f);
          hg.cfaces[c * 6 + f] = (uint32_t)v;
          int ori = f >> 1;
          // the node's cells: [c-, c+]
          if (f & 1)
            hg.ncell[v * 2 + 0] = (int32_t)c;
          else
            hg.ncell[v * 2 + 1] = (int32_t)c;
          // held: the cell across is held
          int ax = x + (ori == 0 ? ((f & 1) ? 1 : -1) : 0), ay = y + (ori == 1 ? ((f & 1) ? 1 : -1) : 0),
              az = z + (ori == 2 ? ((f & 1) ? 1 : -1) : 0);
          if (B.is_held(ax, ay, az)) hg.nflags[v] |= kHeld;
          // face name: orientation and the + side cell, lattice units
          int px = (f & 1) ? ax : x, py = (f & 1) ? ay : y, pz = (f & 1) ? az : z;
          hg.nface[v * 4 + 0] = ori;
          hg.nface[v * 4 + 1] = px + B.lo[0];
          hg.nface[v * 4 + 2] = py + B.lo[1];
          hg.nface[v * 4 + 3] = pz + B.lo[2];
          bool upper_slot = (f & 1) && !B.is_free(ax, ay, az);
          hg.nkind[v] = (uint8_t)(2 * ori + (upper_slot ? 1 : 0));
          bool kept_here = !(f & 1) || upper_slot;
          hg.nkey[v] = kept_here ? hg.ccoord[c] : ((uint32_t)px | (uint32_t)py << 10 | (uint32_t)pz << 20);
        }
      }
    }
  }
}

// Runs f(i) for i in [0, n) on up to `threads` host threads; rethrows the first error.
template <class F>
void parallel_for(int32_t n, F f) {
  const int threads = std::max(1, std::min<int>(n, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
  if (threads <= 1) {
    for (int32_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int32_t> next{0};
  std::vector<std::exception_ptr> err(threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      try {
        for (int32_t i; (i = next.fetch_add(1)) < n;) f(i);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

}  // namespace

// Structures are prepared in parallel (bricks, numbering, node masks), laid out one
// after another, and filled in parallel into their own ranges of the arrays.
void build_host_geo(HostGeo& hg, const StCells* const* cells, int32_t n) {
  hg.clear();
  std::vector<std::unique_ptr<Prep>> prep(n);
  for (int32_t si = 0; si < n; si++) prep[si].reset(new Prep(*cells[si]));
  parallel_for(n, [&](int32_t si) { prepare(*prep[si]); });
  int64_t bricks = 0, ncells = 0, nnodes = 0, ninput = 0;
  std::vector<int64_t> in_beg(n);
  for (int32_t si = 0; si < n; si++) {
    StructInfo& info = prep[si]->info;
    info.brick0 = bricks;
    info.cell0 = ncells;
    info.node0 = nnodes;
    bricks += info.nbricks;
    ncells += info.ncells;
    nnodes += info.nnodes;
    in_beg[si] = ninput;
    ninput += cells[si]->n_cells;
  }
  if (nnodes > 0x7fffffffLL) throw StError{ST_BAD_INPUT, "cells: more than 2^31 nodes"};
  hg.cbase.assign(bricks + 1, 0);
  hg.nbase.assign(bricks + 1, 0);
  hg.bstruct.assign(bricks, 0);
  hg.ce.resize(ncells);
  bool any_sec = false;
  for (int32_t si = 0; si < n; si++) any_sec |= cells[si]->section != nullptr;
  if (any_sec) hg.csec.assign(ncells * 15, 1.0);
  hg.cg.resize(ncells);
  hg.cmask.resize(ncells);
  hg.cfaces.resize(ncells * 6);
  hg.ccoord.resize(ncells);
  hg.cstruct.resize(ncells);
  hg.nflags.assign(nnodes, 0);
  hg.ncell.assign(nnodes * 2, -1);
  hg.nkind.assign(nnodes, 0);
  hg.nkey.assign(nnodes, 0);
  hg.nface.assign(nnodes * 4, 0);
  hg.cell_of_input.assign(ninput, -1);
  parallel_for(n, [&](int32_t si) { fill(*prep[si], si, hg, in_beg[si]); });
  for (int32_t si = 0; si < n; si++) hg.structs.push_back(prep[si]->info);
}

}  // namespace st
