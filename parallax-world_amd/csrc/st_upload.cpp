// Host side of st_upload: cells -> 8^3 bricks, node numbering, face tables.
// Plays the part of lattice::build (statics/src/lattice.rs:124-163): the same
// nodes (every face of a free cell; held iff the cell across is held) and the
// same beams (every face pair of a free cell), but stored per brick and cell
// instead of as node and bond lists.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

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

void build_host_geo(HostGeo& hg, const StCells* const* cells, int32_t n) {
  hg.clear();
  hg.cbase.push_back(0);
  hg.nbase.push_back(0);
  for (int32_t si = 0; si < n; si++) {
    const StCells& C = *cells[si];
    if (C.n_cells < 0 || (C.n_cells > 0 && (!C.sites || !C.held || !C.e || !C.g)) || !(C.spacing > 0))
      throw StError{ST_BAD_INPUT, "cells: missing arrays, negative count or non-positive spacing"};
    Builder B(C);
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
        o.freem[s >> 6] |= 1ull << (s & 63);
      }
    }
    // Active bricks: those with a free cell, in Morton order.
    struct Act {
      uint64_t key;
      int32_t occ;
      int32_t bx, by, bz;
    };
    std::vector<Act> act;
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
    StructInfo info{};
    info.brick0 = (int64_t)hg.cbase.size() - 1;
    info.nbricks = (int64_t)act.size();
    info.cell0 = hg.cbase.back();
    info.node0 = hg.nbase.back();
    for (int k = 0; k < 3; k++) {
      info.origin[k] = B.lo[k];
      info.ext[k] = B.ext[k];
    }
    int32_t emax = std::max(B.ext[0], std::max(B.ext[1], B.ext[2]));
    info.levels_cap = 0;
    while ((1 << info.levels_cap) < emax) info.levels_cap++;
    info.spacing = C.spacing;
    // Free-cell compact ids: brick order, slot order.
    std::unordered_map<int32_t, int32_t> act_of_occ;  // occ index -> position in act
    for (size_t a = 0; a < act.size(); a++) act_of_occ[act[a].occ] = (int32_t)a;
    std::vector<uint32_t> cell_base_local(act.size() + 1, 0);
    for (size_t a = 0; a < act.size(); a++) {
      int cnt = 0;
      for (int w = 0; w < 8; w++) cnt += popc(B.occ[act[a].occ].freem[w]);
      cell_base_local[a + 1] = cell_base_local[a] + cnt;
    }
    const int64_t c0 = info.cell0;
    auto cell_id = [&](int32_t x, int32_t y, int32_t z) -> int64_t {  // compact id of a free cell
      int32_t o = B.brick_at(x >> 3, y >> 3, z >> 3);
      auto it = act_of_occ.find(o);
      int s = Builder::slot_of(x, y, z);
      const uint64_t* fm = B.occ[o].freem;
      int r = 0;
      for (int w = 0; w < (s >> 6); w++) r += popc(fm[w]);
      r += popc(fm[s >> 6] & ((1ull << (s & 63)) - 1));
      return c0 + cell_base_local[it->second] + r;
    };
    // Node masks per active brick: kind k = 2 o + side.
    std::vector<uint64_t> nmask(act.size() * 48, 0);
    std::vector<uint32_t> npre(act.size() * 48, 0);
    std::vector<uint32_t> node_base_local(act.size() + 1, 0);
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
          nmask[a * 48 + k * 8 + w] = m;
          npre[a * 48 + k * 8 + w] = cnt;
          cnt += popc(m);
        }
      }
      node_base_local[a + 1] = node_base_local[a] + cnt;
    }
    const int64_t n0 = info.node0;
    auto node_id = [&](int32_t x, int32_t y, int32_t z, int k) -> int64_t {
      int32_t o = B.brick_at(x >> 3, y >> 3, z >> 3);
      int32_t a = act_of_occ.at(o);
      int s = Builder::slot_of(x, y, z);
      uint64_t m = nmask[a * 48 + k * 8 + (s >> 6)];
      return n0 + node_base_local[a] + npre[a * 48 + k * 8 + (s >> 6)] + popc(m & ((1ull << (s & 63)) - 1));
    };
    // The node holding face f of free cell (x, y, z).
    auto face_node = [&](int32_t x, int32_t y, int32_t z, int f) -> int64_t {
      int ori = f >> 1;
      if ((f & 1) == 0) return node_id(x, y, z, 2 * ori);
      int nx = x + (ori == 0), ny = y + (ori == 1), nz = z + (ori == 2);
      if (B.is_free(nx, ny, nz)) return node_id(nx, ny, nz, 2 * ori);
      return node_id(x, y, z, 2 * ori + 1);
    };
    const int64_t ncells_s = cell_base_local.back(), nnodes_s = node_base_local.back();
    info.ncells = ncells_s;
    info.nnodes = nnodes_s;
    size_t cbeg = hg.ce.size(), nbeg = hg.nflags.size();
    hg.ce.resize(cbeg + ncells_s);
    hg.cg.resize(cbeg + ncells_s);
    hg.cmask.resize(cbeg + ncells_s);
    hg.cfaces.resize((cbeg + ncells_s) * 6);
    hg.ccoord.resize(cbeg + ncells_s);
    hg.cstruct.resize(cbeg + ncells_s, (uint32_t)si);
    hg.nflags.resize(nbeg + nnodes_s, 0);
    hg.ncell.resize((nbeg + nnodes_s) * 2, -1);
    hg.nkind.resize(nbeg + nnodes_s, 0);
    hg.nkey.resize(nbeg + nnodes_s, 0);
    hg.nface.resize((nbeg + nnodes_s) * 4, 0);
    size_t in_beg = hg.cell_of_input.size();
    hg.cell_of_input.resize(in_beg + nc, -1);
    for (size_t a = 0; a < act.size(); a++) {
      const Occ& o = B.occ[act[a].occ];
      hg.bstruct.push_back((uint32_t)si);
      // cells
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
          hg.cmask[c] = C.beams ? (uint16_t)(C.beams[in] & 0x7fff) : (uint16_t)0x7fff;
          hg.ccoord[c] = (uint32_t)x | (uint32_t)y << 10 | (uint32_t)z << 20;
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
      hg.cbase.push_back((uint32_t)(c0 + cell_base_local[a + 1]));
      hg.nbase.push_back((uint32_t)(n0 + node_base_local[a + 1]));
    }
    hg.structs.push_back(info);
    if (hg.nbase.back() > 0x7fffffffu) throw StError{ST_BAD_INPUT, "cells: more than 2^31 nodes"};
  }
}

}  // namespace st
