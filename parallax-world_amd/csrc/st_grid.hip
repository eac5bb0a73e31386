// Cells of a world tile held as the reference's brick grid (lattice::Grid,
// lattice/src/lib.rs:182-205, or worldgen's VoxelGrid, worldgen/src/grid.rs:69-81, which
// shares its brick layout): which sites are cells, in the order statics' `solids` takes
// them, and which of them hold_bulk holds (statics/src/lattice.rs:107-119), both on the
// device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "st_solver.hpp"

namespace st {
namespace {

constexpr int kGridBrick = 8;  // lattice/src/lib.rs:101 (BRICK)

struct GridView {
  int n0, n1, n2;   // the box's size in sites
  int nb0, nb1;     // bricks along x and y
  const int32_t* bricks;
  const uint32_t* solid;  // a bit per site (lattice::Grid), or
  const uint8_t* bytes;   // a byte per site (worldgen's VoxelGrid), a cell from bmin up
  uint8_t bmin;
  int64_t nsites;
};

// Grid::solid (or VoxelGrid::solid >= bmin) at a site relative to the box's least one
// (place, index: lattice/src/lib.rs:293-313, worldgen/src/grid.rs:530-537)
__device__ __forceinline__ bool grid_solid(const GridView& g, int x, int y, int z, int64_t* index = nullptr) {
  if (x < 0 || x >= g.n0 || y < 0 || y >= g.n1 || z < 0 || z >= g.n2) return false;
  const int64_t place = ((int64_t)(z >> 3) * g.nb1 + (y >> 3)) * g.nb0 + (x >> 3);
  const int32_t at = g.bricks[place];
  if (at < 0) return false;
  const int64_t i = (int64_t)at * 512 + (((z & 7) * kGridBrick + (y & 7)) * kGridBrick + (x & 7));
  if (i >= g.nsites) return false;
  if (index) *index = i;
  if (g.bytes) return g.bytes[i] >= g.bmin;
  return (g.solid[i >> 5] >> (i & 31)) & 1u;
}

// the box's sites in solids' order (statics/src/lattice.rs:87-102: x outermost, z innermost)
__device__ __forceinline__ void box_site(const GridView& g, int64_t t, int* s) {
  const int64_t yz = (int64_t)g.n1 * g.n2;
  s[0] = (int)(t / yz);
  s[1] = (int)((t / g.n2) % g.n1);
  s[2] = (int)(t % g.n2);
}

__global__ void k_grid_flags(GridView g, int64_t total, int* flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int s[3];
  box_site(g, t, s);
  flag[t] = grid_solid(g, s[0], s[1], s[2]) ? 1 : 0;
}

// 64-bit exclusive scan in two levels: per block of 1024 sites, then the block totals
__global__ void k_block_count(int64_t total, const int* flag, int64_t* bsum) {
  __shared__ int part[256];
  const int64_t b0 = (int64_t)blockIdx.x * 1024;
  int m = 0;
  for (int k = threadIdx.x; k < 1024; k += 256)
    if (b0 + k < total) m += flag[b0 + k];
  part[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = part[0];
}

// One thread per block of 1024 sites writes its cells in order (a block's offset from the
// block totals' prefix sums on the host).
__global__ void k_grid_emit(GridView g, int64_t total, int64_t nblk, const int* flag, const int64_t* boff,
                            int32_t* sites, int64_t* index) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  int64_t p = boff[b];
  const int64_t t1 = std::min<int64_t>(total, (b + 1) * 1024);
  for (int64_t t = b * 1024; t < t1; t++) {
    if (!flag[t]) continue;
    int s[3];
    box_site(g, t, s);
    int64_t i = -1;
    grid_solid(g, s[0], s[1], s[2], &i);
    sites[3 * p] = s[0];
    sites[3 * p + 1] = s[1];
    sites[3 * p + 2] = s[2];
    index[p] = i;
    p++;
  }
}

// hold_bulk: a cell is held if every site within the radius (the offsets `near`) is a cell
__global__ void k_hold_bulk(GridView g, int64_t m, const int32_t* sites, int nnear, const int32_t* near,
                            uint8_t* held) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int x = sites[3 * k], y = sites[3 * k + 1], z = sites[3 * k + 2];
  bool all = true;
  for (int j = 0; j < nnear && all; j++)
    all = grid_solid(g, x + near[3 * j], y + near[3 * j + 1], z + near[3 * j + 2]);
  held[k] = all ? 1 : 0;
}

template <class T>
void upload_vec(DBuf& d, const T* h, size_t n, hipStream_t s) {
  d.ensure(std::max<size_t>(n, 1) * sizeof(T));
  if (n) ST_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

}  // namespace
}  // namespace st

using namespace st;

int st_grid_cells(StCtx* c, const StGrid* grid, int32_t level, double hold_radius, int64_t* count_out,
                  int32_t* sites_out, uint8_t* held_out, int64_t* index_out) {
  return guard(c, [&] {
    if (!grid || !count_out) throw StError{ST_BAD_INPUT, "grid cells: grid or count_out is NULL"};
    if (level < 0 || level > 20 || !(hold_radius >= 0) || !std::isfinite(hold_radius))
      throw StError{ST_BAD_INPUT, "grid cells: level or hold radius"};
    const int n0 = grid->n[0], n1 = grid->n[1], n2 = grid->n[2];
    if (n0 < 0 || n1 < 0 || n2 < 0) throw StError{ST_BAD_INPUT, "grid cells: negative box size"};
    const int nb0 = (n0 + kGridBrick - 1) / kGridBrick, nb1 = (n1 + kGridBrick - 1) / kGridBrick,
              nb2 = (n2 + kGridBrick - 1) / kGridBrick;
    if (grid->nbricks != (int64_t)nb0 * nb1 * nb2 || (grid->nbricks && !grid->bricks))
      throw StError{ST_BAD_INPUT, "grid cells: bricks does not cover the box (one per 8^3 place)"};
    const bool bytes = grid->solid_bytes != nullptr;
    if (bytes && grid->solid_min == 0) throw StError{ST_BAD_INPUT, "grid cells: solid_min must be at least 1"};
    if (grid->nsites < 0 || (!bytes && (grid->nsites % 512 || (grid->nsites && !grid->solid))))
      throw StError{ST_BAD_INPUT, "grid cells: nsites is not a multiple of 512 or solid is NULL"};
    for (int64_t b = 0; b < grid->nbricks; b++)
      if (grid->bricks[b] >= 0 && ((int64_t)grid->bricks[b] + 1) * 512 > grid->nsites)
        throw StError{ST_BAD_INPUT, "grid cells: a brick's sites run past the grid's"};
    ST_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    DBuf bricks, solid, flag, bsum, boff, sites, index, nearb, held;
    upload_vec(bricks, grid->bricks, (size_t)grid->nbricks, s);
    DBuf sbytes;
    if (bytes)
      upload_vec(sbytes, grid->solid_bytes, (size_t)grid->nsites, s);
    else
      upload_vec(solid, grid->solid, (size_t)(grid->nsites / 32), s);
    GridView g{n0, n1, n2, nb0, nb1, bricks.as<int32_t>(), solid.as<uint32_t>(), sbytes.as<uint8_t>(),
               grid->solid_min, grid->nsites};
    const int64_t total = (int64_t)n0 * n1 * n2;
    const int64_t nblk = (total + 1023) / 1024;
    flag.ensure(std::max<int64_t>(total, 1) * sizeof(int));
    bsum.ensure(std::max<int64_t>(nblk, 1) * sizeof(int64_t));
    if (total) {
      k_grid_flags<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(g, total, flag.as<int>());
      k_block_count<<<(unsigned)nblk, 256, 0, s>>>(total, flag.as<int>(), bsum.as<int64_t>());
    }
    ST_HIP(hipGetLastError());
    std::vector<int64_t> hb((size_t)nblk), ho((size_t)nblk + 1, 0);
    if (nblk) ST_HIP(hipMemcpyAsync(hb.data(), bsum.p, nblk * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    ST_HIP(hipStreamSynchronize(s));
    for (int64_t b = 0; b < nblk; b++) ho[b + 1] = ho[b] + hb[b];
    const int64_t m = ho[nblk];
    *count_out = m;
    if (!sites_out && !held_out && !index_out) return;
    upload_vec(boff, ho.data(), (size_t)nblk, s);
    sites.ensure(std::max<int64_t>(m, 1) * 3 * sizeof(int32_t));
    index.ensure(std::max<int64_t>(m, 1) * sizeof(int64_t));
    if (nblk)
      k_grid_emit<<<(unsigned)((nblk + 255) / 256), 256, 0, s>>>(g, total, nblk, flag.as<int>(), boff.as<int64_t>(),
                                                                sites.as<int32_t>(), index.as<int64_t>());
    // hold_bulk's offsets: within the radius at the level's spacing, in FP64 as the reference
    // decides them (statics/src/lattice.rs:108-113; sites here are in the grid's own units)
    const double sp = 0.125 * std::ldexp(1.0, level);  // UNIT * 2^level (lattice.rs:16, 34-36)
    const int64_t r = (int64_t)std::ceil(hold_radius / sp);
    std::vector<int32_t> near;
    for (int64_t a = -r; a <= r; a++)
      for (int64_t b2 = -r; b2 <= r; b2++)
        for (int64_t c2 = -r; c2 <= r; c2++)
          if (std::sqrt((double)(a * a + b2 * b2 + c2 * c2)) * sp <= hold_radius) {
            near.push_back((int32_t)a);
            near.push_back((int32_t)b2);
            near.push_back((int32_t)c2);
          }
    upload_vec(nearb, near.data(), near.size(), s);
    held.ensure(std::max<int64_t>(m, 1));
    if (m)
      k_hold_bulk<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(g, m, sites.as<int32_t>(), (int)(near.size() / 3),
                                                             nearb.as<int32_t>(), held.as<uint8_t>());
    ST_HIP(hipGetLastError());
    if (m) {
      if (sites_out) {
        ST_HIP(hipMemcpyAsync(sites_out, sites.p, m * 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      }
      if (held_out) ST_HIP(hipMemcpyAsync(held_out, held.p, m, hipMemcpyDeviceToHost, s));
      if (index_out) ST_HIP(hipMemcpyAsync(index_out, index.p, m * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    }
    ST_HIP(hipStreamSynchronize(s));
    if (sites_out)  // the box's least site back in
      for (int64_t k = 0; k < m; k++)
        for (int d = 0; d < 3; d++) sites_out[3 * k + d] += grid->min[d];
  });
}
