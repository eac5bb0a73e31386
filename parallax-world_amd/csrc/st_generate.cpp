// Seeded synthetic structures: a held ground layer, square pillars at
// Poisson-disk spacing standing on it, and one-voxel floor slabs at random
// heights with 2x2 holes; per-cell log-stiffness N(0, sigma^2).
//
// Random numbers: SplitMix64 seeds a xoshiro256** stream (both published by
// Vigna); uniform doubles take the top 53 bits; normals by Box-Muller. The
// same seed gives the same structure on every machine with this libm.
// Pillar centres: Bridson's Poisson-disk sampling (k = 30 candidates per
// active point) over [0, nx) x [0, ny).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/statics_hip.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto& w : s) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      w = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  int64_t range(int64_t lo, int64_t hi) {  // inclusive
    return lo + (int64_t)(uniform() * (double)(hi - lo + 1));
  }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// Bridson: points in [0,w) x [0,h), no two closer than r.
std::vector<std::pair<double, double>> poisson_disk(Rng& rng, double w, double h, double r) {
  std::vector<std::pair<double, double>> pts;
  if (r <= 0 || w <= 0 || h <= 0) return pts;
  double cell = r / std::sqrt(2.0);
  int gw = (int)std::ceil(w / cell), gh = (int)std::ceil(h / cell);
  std::vector<int> grid((size_t)gw * gh, -1);
  std::vector<int> active;
  auto put = [&](double x, double y) {
    pts.emplace_back(x, y);
    int id = (int)pts.size() - 1;
    grid[(size_t)(int)(y / cell) * gw + (int)(x / cell)] = id;
    active.push_back(id);
  };
  put(rng.uniform() * w, rng.uniform() * h);
  while (!active.empty()) {
    size_t pick = (size_t)(rng.uniform() * (double)active.size());
    if (pick >= active.size()) pick = active.size() - 1;
    auto [px, py] = pts[active[pick]];
    bool found = false;
    for (int k = 0; k < 30; k++) {
      double a = 2.0 * M_PI * rng.uniform();
      double d = r * (1.0 + rng.uniform());
      double x = px + d * std::cos(a), y = py + d * std::sin(a);
      if (x < 0 || y < 0 || x >= w || y >= h) continue;
      int gx = (int)(x / cell), gy = (int)(y / cell);
      bool ok = true;
      for (int yy = std::max(0, gy - 2); yy <= std::min(gh - 1, gy + 2) && ok; yy++)
        for (int xx = std::max(0, gx - 2); xx <= std::min(gw - 1, gx + 2); xx++) {
          int q = grid[(size_t)yy * gw + xx];
          if (q >= 0) {
            double dx = pts[q].first - x, dy = pts[q].second - y;
            if (dx * dx + dy * dy < r * r) {
              ok = false;
              break;
            }
          }
        }
      if (ok) {
        put(x, y);
        found = true;
        break;
      }
    }
    if (!found) {
      active[pick] = active.back();
      active.pop_back();
    }
  }
  return pts;
}

}  // namespace

extern "C" int st_generate(const StGenParams* p, uint8_t* occ, float* theta) {
  if (!p || !occ || p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return ST_BAD_INPUT;
  if (p->pillar_min < 1 || p->pillar_max < p->pillar_min || p->slab_min < 1 || p->slab_max < p->slab_min)
    return ST_BAD_INPUT;
  const int64_t nx = p->nx, ny = p->ny, nz = p->nz;
  auto at = [&](int64_t x, int64_t y, int64_t z) -> uint8_t& { return occ[(x * ny + y) * nz + z]; };
  std::memset(occ, 0, (size_t)(nx * ny * nz));
  Rng rng(p->seed);
  // Ground: the z = 0 layer, held.
  for (int64_t x = 0; x < nx; x++)
    for (int64_t y = 0; y < ny; y++) at(x, y, 0) = 2;
  // Pillars from z = 1 to the top.
  auto centres = poisson_disk(rng, (double)nx, (double)ny, p->pillar_spacing);
  for (auto [cx, cy] : centres) {
    int64_t w = rng.range(p->pillar_min, p->pillar_max);
    int64_t x0 = (int64_t)std::floor(cx - 0.5 * (double)w), y0 = (int64_t)std::floor(cy - 0.5 * (double)w);
    for (int64_t x = std::max<int64_t>(0, x0); x < std::min(nx, x0 + w); x++)
      for (int64_t y = std::max<int64_t>(0, y0); y < std::min(ny, y0 + w); y++)
        for (int64_t z = 1; z < nz; z++) at(x, y, z) = 1;
  }
  // Slabs: one voxel thick, every slab_min..slab_max, 2x2 tiles removed with hole_p.
  int64_t z = 0;
  for (;;) {
    z += rng.range(p->slab_min, p->slab_max);
    if (z >= nz) break;
    for (int64_t tx = 0; tx < nx; tx += 2)
      for (int64_t ty = 0; ty < ny; ty += 2) {
        bool hole = rng.uniform() < p->hole_p;
        if (hole) continue;
        for (int64_t x = tx; x < std::min(nx, tx + 2); x++)
          for (int64_t y = ty; y < std::min(ny, ty + 2); y++) at(x, y, z) = 1;
      }
  }
  if (theta) {
    // One normal per site in index order, so theta does not depend on the fill.
    Rng trng(p->seed ^ 0x5DEECE66Dull);
    const int64_t n = nx * ny * nz;
    for (int64_t i = 0; i < n; i++) theta[i] = p->sigma > 0 ? (float)(p->sigma * trng.normal()) : 0.0f;
  }
  return ST_OK;
}
