// The oracle: a CPU FP64 restatement of the reference's statics crate, for
// tests and the CPU baseline only. It is test infrastructure: nothing in the
// product links, loads or calls it.
//
// Followed, function by function (all paths relative to the reference repo):
//   statics/src/model.rs    Mat, Node, Bond::{force, used, stiffness}, Mat::used
//   statics/src/lattice.rs  spacing, centre, coarse_site, class, beam, hold_bulk, build
//   statics/src/bodies.rs   shift, each, times*, Level::{new, of, most_asked, apply,
//                           jacobi, ask, coarser}, merge, Cholesky, conjugate_gradients,
//                           cycle, forces, moved, motions, bodies_of, settle, used,
//                           used_standing, joined_to_held, invert, find, as_generated
//   statics/src/gather.rs   Cell::{handle, pieces}, responses, beams, solve, full,
//                           measured, gathered (and Bond::energy, model.rs:132-136)
// Rust semantics mirrored: no contraction of a * b + c (built with -ffp-contract=off);
// sums are sequential left folds; `max_by` keeps the
// LAST maximum (invert's pivot); `>>` on i64 floors; f64::round is half away
// from zero (std::round); HashMaps are only looked up, never iterated.
//
// Pinned by tests/test_oracle.py: closed forms (SECTION identity, a single
// bond, principal stresses), invariants (symmetry, rigid motions, equilibrium,
// f.x = x.K.x), brute-force dense solves on small lattices, independent BFS
// and hold_bulk restatements, numpy inverses/Cholesky, and the level-0
// "used" numbers the reference documents in docs/statics.md:236-245.
// The exact iteration path at the reference's hardcoded 1e-6 (iteration counts,
// Work totals, most_asked's 12-step value) cannot be pinned without running
// the Rust code, which has no toolchain here; parity is defined on the
// converged solution.
//
// Additions that the reference does not have (oracle for new functionality):
//   orc_scale_cell scales one cell's stiffness (its bonds' E and G), so tests take
//                  central differences of compliance and soft-max stress w.r.t.
//                  per-cell log-stiffness from oracle solves; the device adjoint
//                  is checked against those and the scaling invariants.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace {

using V3 = std::array<double, 3>;
using Six = std::array<double, 6>;
using M6 = std::array<std::array<double, 6>, 6>;
using M3 = std::array<std::array<double, 3>, 3>;
using Site = std::array<int64_t, 3>;

// model.rs:16-36
V3 add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V3 sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 mul(V3 a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross(V3 a, V3 b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double norm(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a) { return mul(a, 1.0 / norm(a)); }

constexpr double GRAVITY = 9.81;  // model.rs:38

// model.rs:43-51 (name dropped: it is only printed by the examples)
struct Mat {
  double rho, e, g, ft, fc, fs;
  // model.rs:66-68
  double used(double s1, double s3) const {
    return std::max(std::max(s1 / ft, -s3 / fc), (s1 - s3) / (2.0 * fs));
  }
};

// model.rs:92-98
struct Node {
  V3 x;
  double m;
  Mat mat;
  bool held;
};

// model.rs:101-111
struct Bond {
  size_t i, j;
  V3 p, n;
  double a, d;
  Mat mat;
  // model.rs:115-119
  V3 force(V3 delta) const {
    double along = dot(delta, n);
    double k = a / d;
    return add(mul(n, mat.e * k * along), mul(sub(delta, mul(n, along)), mat.g * k));
  }
  // model.rs:123-129
  double used(V3 f) const {
    V3 t = mul(f, 1.0 / a);
    double sigma = dot(t, n);
    double tau2 = std::max(dot(t, t) - sigma * sigma, 0.0);
    double r = std::sqrt(sigma * sigma / 4.0 + tau2);
    return mat.used(sigma / 2.0 + r, sigma / 2.0 - r);
  }
  // model.rs:132-136
  double energy(V3 f) const {
    double k = a / d;
    double along = dot(f, n);
    return 0.5 * (along * along / (mat.e * k) + (dot(f, f) - along * along) / (mat.g * k));
  }
  // model.rs:139-149
  M3 stiffness() const {
    double k = a / d;
    M3 s{};
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        double nn = n[r] * n[c];
        s[r][c] = mat.e * k * nn + mat.g * k * ((r == c ? 1.0 : 0.0) - nn);
      }
    return s;
  }
};

// lattice.rs:16-25
constexpr double UNIT = 0.125;
constexpr size_t PER_BEAM = 4;
constexpr double SECTION[2] = {0.336, 0.336};
// lattice/src/lib.rs:64
constexpr int VOXEL_FACES[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

// lattice.rs:29-32
size_t cls(V3 a, V3 b, V3 c) {
  V3 u = sub(a, c), v = sub(b, c);
  return (size_t)std::round(1.0 + dot(u, v) / (norm(u) * norm(v)));
}
// lattice.rs:34-36
double spacing(uint32_t level) { return UNIT * (double)(uint64_t(1) << level); }
// lattice.rs:39-41
V3 centre(Site site, uint32_t level) {
  V3 c;
  for (int k = 0; k < 3; k++) c[k] = ((double)site[k] + 0.5 * (double)(uint64_t(1) << level)) * UNIT;
  return c;
}
// lattice.rs:45-47: arithmetic shift floors, as Rust's >> on i64
Site coarse_site(Site p, uint32_t l) {
  Site q;
  for (int k = 0; k < 3; k++) q[k] = (p[k] >> l) << l;
  return q;
}

struct Solid {
  Site site;
  Mat mat;
  bool held;
};

using Around = std::array<Site, 2>;
struct AroundHash {
  size_t operator()(const Around& a) const {
    uint64_t h = 1469598103934665603ull;
    for (auto& s : a)
      for (auto v : s) h = (h ^ (uint64_t)v) * 1099511628211ull;
    return (size_t)h;
  }
};
struct SiteHash {
  size_t operator()(const Site& s) const {
    uint64_t h = 1469598103934665603ull;
    for (auto v : s) h = (h ^ (uint64_t)v) * 1099511628211ull;
    return (size_t)h;
  }
};

// lattice.rs:61-71 (+ which cell emitted each beam, for the gradient oracle)
struct Lattice {
  uint32_t level;
  std::vector<Node> nodes;
  std::vector<Bond> bonds;
  std::vector<Around> around;
  std::vector<size_t> beam_cell;  // per beam (PER_BEAM bonds), the index into the solids
};

// lattice.rs:76-84
void beam(size_t i, size_t j, V3 a, V3 b, double area, Mat mat, std::vector<Bond>& out) {
  V3 n = unit(sub(b, a));
  V3 e1 = unit(cross(n, std::fabs(n[0]) < 0.9 ? V3{1.0, 0.0, 0.0} : V3{0.0, 1.0, 0.0}));
  V3 e2 = cross(n, e1);
  double r = std::sqrt(area / (2.0 * M_PI));
  V3 middle = mul(add(a, b), 0.5);
  double d = norm(sub(b, a));
  for (V3 o : {mul(e1, r), mul(e1, -r), mul(e2, r), mul(e2, -r)})
    out.push_back(Bond{i, j, add(middle, o), n, area / 4.0, d, mat});
}

// lattice.rs:107-119
void hold_bulk(uint32_t level, std::vector<Solid>& solids, double radius) {
  int64_t r = (int64_t)std::ceil(radius / spacing(level));
  std::vector<Site> near;
  for (int64_t a = -r; a <= r; a++)
    for (int64_t b = -r; b <= r; b++)
      for (int64_t c = -r; c <= r; c++)
        if (std::sqrt((double)(a * a + b * b + c * c)) * spacing(level) <= radius)
          near.push_back({a << level, b << level, c << level});
  std::unordered_map<Site, char, SiteHash> filled;
  for (auto& c : solids) filled[c.site] = 1;
  std::vector<char> held(solids.size());
  for (size_t i = 0; i < solids.size(); i++) {
    bool all = true;
    for (auto& o : near) {
      Site s{solids[i].site[0] + o[0], solids[i].site[1] + o[1], solids[i].site[2] + o[2]};
      if (!filled.count(s)) {
        all = false;
        break;
      }
    }
    held[i] = all;
  }
  for (size_t i = 0; i < solids.size(); i++) solids[i].held = held[i];
}

// lattice.rs:124-163
Lattice build(uint32_t level, const std::vector<Solid>& solids) {
  double s = spacing(level);
  std::unordered_map<Site, char, SiteHash> held;
  for (auto& c : solids)
    if (c.held) held[c.site] = 1;
  std::unordered_map<Around, size_t, AroundHash> index;
  Lattice lat;
  lat.level = level;
  for (size_t ci = 0; ci < solids.size(); ci++) {
    const Solid& cell = solids[ci];
    if (cell.held) continue;
    double m = cell.mat.rho * s * s * s / 6.0;
    size_t ends[6];
    for (int f = 0; f < 6; f++) {
      Site other;
      for (int k = 0; k < 3; k++) other[k] = cell.site[k] + ((int64_t)VOXEL_FACES[f][k] << level);
      Around key = cell.site < other ? Around{cell.site, other} : Around{other, cell.site};
      auto it = index.find(key);
      size_t k;
      if (it == index.end()) {
        V3 x = mul(add(centre(cell.site, level), centre(other, level)), 0.5);
        lat.nodes.push_back(Node{x, 0.0, cell.mat, held.count(other) > 0});
        lat.around.push_back(key);
        k = lat.nodes.size() - 1;
        index.emplace(key, k);
      } else {
        k = it->second;
      }
      Node& node = lat.nodes[k];
      node.m += m;
      if (cell.mat.rho > node.mat.rho) node.mat = cell.mat;
      ends[f] = k;
    }
    V3 middle = centre(cell.site, level);
    for (int a = 0; a < 6; a++)
      for (int b = a + 1; b < 6; b++)
        if (!(lat.nodes[ends[a]].held && lat.nodes[ends[b]].held)) {
          V3 x = lat.nodes[ends[a]].x, y = lat.nodes[ends[b]].x;
          double area = SECTION[cls(x, y, middle)] * s * s;
          beam(ends[a], ends[b], x, y, area, cell.mat, lat.bonds);
          lat.beam_cell.push_back(ci);
        }
  }
  return lat;
}

// ---------------------------------------------------------------- bodies.rs

// bodies.rs:44-47
Six shift(Six x, V3 r) {
  V3 u = add({x[0], x[1], x[2]}, cross({x[3], x[4], x[5]}, r));
  return {u[0], u[1], u[2], x[3], x[4], x[5]};
}

// bodies.rs:51-69: contiguous chunks over the host's hardware threads; serial below 256.
int g_threads = 0;  // 0: std::thread::hardware_concurrency()
// progress lines on stderr for long runs (an addition of the oracle): every level built,
// and every g_progress CG iterations; 0: none
int g_progress = 0;
int threads_now() {
  int t = g_threads > 0 ? g_threads : (int)std::thread::hardware_concurrency();
  return t < 1 ? 1 : t;
}
template <class T, class F>
std::vector<T> each(size_t n, F f) {
  int threads = threads_now();
  std::vector<T> out(n);
  if (n < 256 || threads == 1) {
    for (size_t i = 0; i < n; i++) out[i] = f(i);
    return out;
  }
  size_t size = (n + threads - 1) / threads;
  std::vector<std::thread> pool;
  for (size_t c = 0; c * size < n; c++) {
    pool.emplace_back([&, c]() {
      size_t hi = std::min(n, (c + 1) * size);
      for (size_t k = c * size; k < hi; k++) out[k] = f(k);
    });
  }
  for (auto& t : pool) t.join();
  return out;
}

// bodies.rs:71-77
Six times(const M6& a, const Six& x) {
  Six y;
  for (int s = 0; s < 6; s++) {
    double acc = 0.0;
    for (int t = 0; t < 6; t++) acc += a[s][t] * x[t];
    y[s] = acc;
  }
  return y;
}
Six times_transposed(const M6& a, const Six& x) {
  Six y;
  for (int s = 0; s < 6; s++) {
    double acc = 0.0;
    for (int t = 0; t < 6; t++) acc += a[t][s] * x[t];
    y[s] = acc;
  }
  return y;
}
Six add6v(const Six& a, const Six& b) {
  Six y;
  for (int s = 0; s < 6; s++) y[s] = a[s] + b[s];
  return y;
}
// bodies.rs:83-89
M6 product(const M6& a, const M6& b) {
  M6 c;
  for (int r = 0; r < 6; r++)
    for (int q = 0; q < 6; q++) {
      double acc = 0.0;
      for (int k = 0; k < 6; k++) acc += a[r][k] * b[k][q];
      c[r][q] = acc;
    }
  return c;
}
M6 transpose(const M6& a) {
  M6 t;
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) t[r][c] = a[c][r];
  return t;
}
// bodies.rs:91-97
void add6(M6& to, const M6& a, double by) {
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) to[r][c] += by * a[r][c];
}
// bodies.rs:100-109
template <class F>
M6 matrix(F f) {
  M6 a{};
  for (int c = 0; c < 6; c++) {
    Six e{};
    e[c] = 1.0;
    Six col = f(e);
    for (int r = 0; r < 6; r++) a[r][c] = col[r];
  }
  return a;
}
using Row = std::vector<std::pair<size_t, M6>>;
// bodies.rs:112-121
void merge(Row& row, size_t c, const M6& m, double by) {
  for (auto& e : row)
    if (e.first == c) {
      add6(e.second, m, by);
      return;
    }
  M6 to{};
  add6(to, m, by);
  row.emplace_back(c, to);
}

// bodies.rs:576-602: Gauss-Jordan, the pivot the LAST greatest |a[x][col]| (Rust max_by).
M6 invert(const M6& m) {
  M6 a = m, inv{};
  for (int k = 0; k < 6; k++) inv[k][k] = 1.0;
  for (int col = 0; col < 6; col++) {
    int piv = col;
    for (int x = col + 1; x < 6; x++)
      if (!(std::fabs(a[x][col]) < std::fabs(a[piv][col]))) piv = x;  // ties move to the later row
    std::swap(a[col], a[piv]);
    std::swap(inv[col], inv[piv]);
    double d = a[col][col];
    for (int k = 0; k < 6; k++) {
      a[col][k] /= d;
      inv[col][k] /= d;
    }
    for (int r = 0; r < 6; r++) {
      double f = a[r][col];
      if (r != col && f != 0.0)
        for (int k = 0; k < 6; k++) {
          a[r][k] -= f * a[col][k];
          inv[r][k] -= f * inv[col][k];
        }
    }
  }
  return inv;
}

double dot6(const std::vector<Six>& a, const std::vector<Six>& b) {
  double acc = 0.0;
  for (size_t k = 0; k < a.size(); k++) {
    double s6 = 0.0;
    for (int s = 0; s < 6; s++) s6 += a[k][s] * b[k][s];
    acc += s6;
  }
  return acc;
}

struct Cholesky {
  size_t n = 0;
  std::vector<double> l;
  std::vector<Six> solve(const std::vector<Six>& r) const;
};

struct Pair {
  size_t i, j;
  M6 m;
};

struct Level {
  std::vector<V3> centre;
  std::vector<M6> own;
  std::vector<Pair> pairs;
  std::vector<std::vector<std::tuple<size_t, size_t, bool>>> rows;
  std::vector<M6> inverse;
  double step = 1.0;
  std::vector<Row> down;
  std::vector<std::vector<std::pair<size_t, size_t>>> from;
  bool has_direct = false;
  Cholesky direct;

  // bodies.rs:193-206
  std::vector<Six> apply(const std::vector<Six>& x, double& work) const {
    work += (double)(x.size() + pairs.size());
    return each<Six>(x.size(), [&](size_t i) {
      Six y = times(own[i], x[i]);
      for (auto& [o, p, tr] : rows[i]) {
        const M6& m = pairs[p].m;
        Six a = tr ? times_transposed(m, x[o]) : times(m, x[o]);
        for (int s = 0; s < 6; s++) y[s] += a[s];
      }
      return y;
    });
  }
  // bodies.rs:208-210
  std::vector<Six> jacobi(const std::vector<Six>& r) const {
    return each<Six>(r.size(), [&](size_t i) { return times(inverse[i], r[i]); });
  }
  // bodies.rs:180-190
  double most_asked() const {
    std::vector<Six> v(own.size());
    for (size_t k = 0; k < own.size(); k++)
      for (int s = 0; s < 6; s++) v[k][s] = 1.0 + (double)((7 * k + 3 * s) % 11) / 11.0;
    double most = 1.0;
    for (int it = 0; it < 12; it++) {
      double w0 = 0.0;
      std::vector<Six> w = jacobi(apply(v, w0));
      most = std::sqrt(dot6(w, w)) / std::sqrt(dot6(v, v));
      double scale = 1.0 / std::sqrt(dot6(w, w));
      for (auto& x : w)
        for (auto& a : x) a *= scale;
      v = std::move(w);
    }
    return most;
  }
  // bodies.rs:213-225
  template <class F>
  Row ask(size_t k, F m) const {
    Row row;
    for (auto& [c, b] : m(k)) merge(row, c, product(own[k], b), 1.0);
    for (auto& [o, p, tr] : rows[k]) {
      M6 a = tr ? transpose(pairs[p].m) : pairs[p].m;
      for (auto& [c, b] : m(o)) merge(row, c, product(a, b), 1.0);
    }
    return row;
  }
  Level coarser(const std::vector<size_t>& up, size_t count);
};

// bodies.rs:166-176
Level level_of(std::vector<V3> centre, std::vector<M6> own, std::vector<Pair> pairs) {
  Level lv;
  lv.rows.assign(own.size(), {});
  for (size_t p = 0; p < pairs.size(); p++) {
    lv.rows[pairs[p].i].emplace_back(pairs[p].j, p, false);
    lv.rows[pairs[p].j].emplace_back(pairs[p].i, p, true);
  }
  lv.inverse.resize(own.size());
  for (size_t k = 0; k < own.size(); k++) lv.inverse[k] = invert(own[k]);
  lv.centre = std::move(centre);
  lv.own = std::move(own);
  lv.pairs = std::move(pairs);
  lv.step = 4.0 / 3.0 / lv.most_asked();
  return lv;
}

struct PairHash {
  size_t operator()(const std::pair<size_t, size_t>& p) const { return p.first * 1000003u ^ p.second; }
};

// bodies.rs:126-164
Level level_new(const Lattice& lat, const std::vector<int64_t>& part, size_t parts, const std::vector<size_t>& live) {
  std::vector<V3> centre(parts, V3{0, 0, 0});
  for (size_t v = 0; v < part.size(); v++)
    if (part[v] >= 0) centre[part[v]] = lat.nodes[v].x;
  std::vector<M6> own(parts, M6{});
  std::vector<Pair> pairs;
  std::unordered_map<std::pair<size_t, size_t>, size_t, PairHash> pair;
  using T3 = std::array<std::array<double, 6>, 3>;
  for (size_t b : live) {
    const Bond& bond = lat.bonds[b];
    M3 s = bond.stiffness();
    auto seen = [&](size_t v) -> T3 {
      V3 r = sub(bond.p, lat.nodes[v].x);
      return T3{{{1.0, 0.0, 0.0, 0.0, r[2], -r[1]}, {0.0, 1.0, 0.0, -r[2], 0.0, r[0]}, {0.0, 0.0, 1.0, r[1], -r[0], 0.0}}};
    };
    auto through = [&](const T3& a, const T3& bb) -> M6 {
      T3 sb;
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 6; c++) {
          double acc = 0.0;
          for (int k = 0; k < 3; k++) acc += s[r][k] * bb[k][c];
          sb[r][c] = acc;
        }
      M6 out;
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) {
          double acc = 0.0;
          for (int k = 0; k < 3; k++) acc += a[k][r] * sb[k][c];
          out[r][c] = acc;
        }
      return out;
    };
    int64_t pi = part[bond.i], pj = part[bond.j];
    T3 ti, tj;
    if (pi >= 0) ti = seen(bond.i);
    if (pj >= 0) tj = seen(bond.j);
    if (pi >= 0) add6(own[pi], through(ti, ti), 1.0);
    if (pj >= 0) add6(own[pj], through(tj, tj), 1.0);
    if (pi >= 0 && pj >= 0) {
      size_t i = pi, j = pj;
      const T3* a = &ti;
      const T3* bb = &tj;
      if (!(i < j)) {
        std::swap(i, j);
        std::swap(a, bb);
      }
      auto it = pair.find({i, j});
      size_t k;
      if (it == pair.end()) {
        pairs.push_back(Pair{i, j, M6{}});
        k = pairs.size() - 1;
        pair.emplace(std::make_pair(i, j), k);
      } else {
        k = it->second;
      }
      add6(pairs[k].m, through(*a, *bb), -1.0);
    }
  }
  return level_of(std::move(centre), std::move(own), std::move(pairs));
}

// bodies.rs:228-282
Level Level::coarser(const std::vector<size_t>& up, size_t count) {
  std::vector<V3> c(count, V3{0, 0, 0});
  std::vector<double> n(count, 0.0);
  for (size_t k = 0; k < up.size(); k++) {
    c[up[k]] = add(c[up[k]], centre[k]);
    n[up[k]] += 1.0;
  }
  for (size_t q = 0; q < count; q++) c[q] = mul(c[q], 1.0 / n[q]);
  std::vector<M6> rigid(up.size());
  for (size_t k = 0; k < up.size(); k++) {
    V3 r = sub(centre[k], c[up[k]]);
    rigid[k] = matrix([&](Six y) { return shift(y, r); });
  }
  auto rigid_row = [&](size_t k) { return Row{{up[k], rigid[k]}}; };
  down = each<Row>(up.size(), [&](size_t k) {
    Row row = rigid_row(k);
    for (auto& [cc, m] : ask(k, rigid_row)) merge(row, cc, product(inverse[k], m), -step);
    return row;
  });
  from.assign(count, {});
  for (size_t k = 0; k < down.size(); k++)
    for (size_t q = 0; q < down[k].size(); q++) from[down[k][q].first].emplace_back(k, q);
  const Level& fine = *this;
  // kp[k] = (K P)_k is needed by the coarse rows whose from lists hold k. The reference
  // keeps it for every fine row at once; here it is formed per run of coarse rows, for the
  // fine rows that run needs, which bounds its memory (a kp row is ~15 6x6 blocks) and gives
  // the same values: kp[k] is a function of k alone, and each coarse row sums its from list
  // in the same order.
  std::vector<Row> rows_c(count);
  const size_t run = std::max<size_t>(4096, count / 16);
  for (size_t a = 0; a < count; a += run) {
    const size_t b = std::min(count, a + run);
    std::vector<size_t> need;
    for (size_t cc = a; cc < b; cc++)
      for (auto& kq : fine.from[cc]) need.push_back(kq.first);
    std::sort(need.begin(), need.end());
    need.erase(std::unique(need.begin(), need.end()), need.end());
    std::vector<Row> kp =
        each<Row>(need.size(), [&](size_t i) { return fine.ask(need[i], [&](size_t o) { return fine.down[o]; }); });
    std::vector<Row> part = each<Row>(b - a, [&](size_t i) {
      const size_t cc = a + i;
      Row row;
      for (auto& [k, q] : fine.from[cc]) {
        M6 pt = transpose(fine.down[k][q].second);
        const Row& kr = kp[std::lower_bound(need.begin(), need.end(), k) - need.begin()];
        for (auto& [d, m] : kr)
          if (d >= cc) merge(row, d, product(pt, m), 1.0);
      }
      return row;
    });
    for (size_t i = 0; i < b - a; i++) rows_c[a + i] = std::move(part[i]);
  }
  std::vector<M6> own_c(count, M6{});
  std::vector<Pair> pairs_c;
  for (size_t cc = 0; cc < count; cc++)
    for (auto& [d, m] : rows_c[cc]) {
      if (d == cc)
        own_c[cc] = m;
      else
        pairs_c.push_back(Pair{cc, d, m});
    }
  return level_of(std::move(c), std::move(own_c), std::move(pairs_c));
}

// bodies.rs:293-315
Cholesky cholesky_of(const Level& level) {
  size_t parts = level.centre.size();
  size_t n = 6 * parts;
  std::vector<double> a(n * n, 0.0);
  double w = 0.0;
  for (size_t c = 0; c < n; c++) {
    std::vector<Six> e(parts, Six{});
    e[c / 6][c % 6] = 1.0;
    std::vector<Six> v = level.apply(e, w);
    for (size_t k = 0; k < parts; k++)
      for (int s = 0; s < 6; s++) a[(6 * k + s) * n + c] = v[k][s];
  }
  for (size_t j = 0; j < n; j++) {
    double acc = 0.0;
    for (size_t k = 0; k < j; k++) acc += a[j * n + k] * a[j * n + k];
    double d = std::sqrt(a[j * n + j] - acc);
    a[j * n + j] = d;
    for (size_t i = j + 1; i < n; i++) {
      double acc2 = 0.0;
      for (size_t k = 0; k < j; k++) acc2 += a[i * n + k] * a[j * n + k];
      a[i * n + j] = (a[i * n + j] - acc2) / d;
    }
  }
  Cholesky ch;
  ch.n = n;
  ch.l = std::move(a);
  return ch;
}

// bodies.rs:317-327
std::vector<Six> Cholesky::solve(const std::vector<Six>& r) const {
  std::vector<double> y;
  y.reserve(n);
  for (auto& v : r)
    for (double a : v) y.push_back(a);
  for (size_t i = 0; i < n; i++) {
    double acc = 0.0;
    for (size_t k = 0; k < i; k++) acc += l[i * n + k] * y[k];
    y[i] = (y[i] - acc) / l[i * n + i];
  }
  for (size_t i = n; i-- > 0;) {
    double acc = 0.0;
    for (size_t k = i + 1; k < n; k++) acc += l[k * n + i] * y[k];
    y[i] = (y[i] - acc) / l[i * n + i];
  }
  std::vector<Six> out(n / 6);
  for (size_t k = 0; k < n / 6; k++)
    for (int s = 0; s < 6; s++) out[k][s] = y[6 * k + s];
  return out;
}

size_t find(std::vector<size_t>& p, size_t a) {  // bodies.rs:628-634
  while (p[a] != a) {
    p[a] = p[p[a]];
    a = p[a];
  }
  return a;
}

// bodies.rs:472-496
size_t bodies_of(const Level& level, const std::vector<int64_t>& part, const std::vector<Site>& key,
                 std::vector<size_t>& up, std::vector<int64_t>& next) {
  size_t parts = level.centre.size();
  std::vector<Site> their(parts, Site{0, 0, 0});
  for (size_t v = 0; v < part.size(); v++)
    if (part[v] >= 0) their[part[v]] = key[v];
  std::vector<size_t> parent(parts);
  for (size_t k = 0; k < parts; k++) parent[k] = k;
  for (auto& pr : level.pairs)
    if (their[pr.i] == their[pr.j]) {
      size_t a = find(parent, pr.i), c = find(parent, pr.j);
      parent[a] = c;
    }
  std::unordered_map<size_t, size_t> id;
  up.assign(parts, 0);
  for (size_t k = 0; k < parts; k++) {
    size_t n = id.size();
    auto it = id.emplace(find(parent, k), n).first;
    up[k] = it->second;
  }
  next.assign(part.size(), -1);
  for (size_t v = 0; v < part.size(); v++)
    if (part[v] >= 0) next[v] = (int64_t)up[part[v]];
  return id.size();
}

// bodies.rs:377-403
std::vector<Six> cycle(const std::vector<Level>& levels, size_t l, const std::vector<Six>& r, double& work) {
  const Level& level = levels[l];
  if (l + 1 == levels.size()) return level.direct.solve(r);
  auto rest = [&](const std::vector<Six>& x) {
    std::vector<Six> ax = level.apply(x, work);
    std::vector<Six> out(r.size());
    for (size_t k = 0; k < r.size(); k++)
      for (int s = 0; s < 6; s++) out[k][s] = r[k][s] - ax[k][s];
    return out;
  };
  std::vector<Six> x = level.jacobi(r);
  for (auto& v : x)
    for (auto& a : v) a = a * level.step;
  std::vector<Six> left = rest(x);
  std::vector<Six> coarse = each<Six>(level.from.size(), [&](size_t c) {
    Six a{};
    for (auto& [k, q] : level.from[c]) a = add6v(a, times_transposed(level.down[k][q].second, left[k]));
    return a;
  });
  std::vector<Six> y = cycle(levels, l + 1, coarse, work);
  std::vector<Six> moved = each<Six>(x.size(), [&](size_t k) {
    Six a{};
    for (auto& [c, p] : level.down[k]) a = add6v(a, times(p, y[c]));
    return a;
  });
  for (size_t k = 0; k < x.size(); k++) x[k] = add6v(x[k], moved[k]);
  left = rest(x);
  std::vector<Six> d = level.jacobi(left);
  for (size_t k = 0; k < x.size(); k++)
    for (int s = 0; s < 6; s++) x[k][s] += level.step * d[k][s];
  return x;
}

// bodies.rs:338-373 (tol and the iteration cap are parameters here; the
// reference hardcodes 1e-6 and 100,000)
std::vector<Six> conjugate_gradients(const Level& level, const std::vector<Six>& loads,
                                     const std::function<std::vector<Six>(const std::vector<Six>&)>& pre,
                                     double tol, int max_iter, double& work, int& iters) {
  std::vector<Six> x(loads.size(), Six{});
  std::vector<Six> r = loads;
  std::vector<Six> z = pre(r);
  std::vector<Six> p = z;
  double rz = dot6(r, z);
  double goal = tol * std::sqrt(dot6(loads, loads));
  iters = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < max_iter; it++) {
    if (g_progress > 0 && it % g_progress == 0) {
      std::fprintf(stderr, "oracle: iteration %d, |r|/|f| %.3e, %.1f s\n", it,
                   std::sqrt(dot6(r, r)) / std::sqrt(dot6(loads, loads)),
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      std::fflush(stderr);
    }
    if (std::sqrt(dot6(r, r)) <= goal) break;
    std::vector<Six> ap = level.apply(p, work);
    double alpha = rz / dot6(p, ap);
    for (size_t k = 0; k < x.size(); k++)
      for (int s = 0; s < 6; s++) {
        x[k][s] += alpha * p[k][s];
        r[k][s] -= alpha * ap[k][s];
      }
    z = pre(r);
    double next = dot6(r, z);
    for (size_t k = 0; k < x.size(); k++)
      for (int s = 0; s < 6; s++) p[k][s] = z[k][s] + next / rz * p[k][s];
    rz = next;
    iters = it + 1;
  }
  return x;
}

struct MotionStats {
  int iters = 0;
  double work = 0.0;
  double setup_s = 0.0, solve_s = 0.0;
  int levels = 0;
  size_t top_parts = 0;
};

// bodies.rs:429-468
std::vector<Six> motions(const Lattice& lat, const std::vector<char>& alive, const std::vector<char>& gone,
                         const std::function<Six(size_t)>& loads, double tol, int max_iter, MotionStats& st) {
  auto t0 = std::chrono::steady_clock::now();
  constexpr size_t FEW = 64;
  const auto& nodes = lat.nodes;
  const auto& bonds = lat.bonds;
  std::vector<size_t> live;
  for (size_t b = 0; b < bonds.size(); b++)
    if (alive[b] && !gone[bonds[b].i] && !gone[bonds[b].j]) live.push_back(b);
  std::vector<int64_t> part(nodes.size(), -1);
  size_t count = 0;
  for (size_t v = 0; v < nodes.size(); v++)
    if (!nodes[v].held && !gone[v]) part[v] = (int64_t)count++;
  std::vector<int64_t> node = part;
  std::vector<Level> levels;
  levels.push_back(level_new(lat, part, count, live));
  if (g_progress > 0) {
    std::fprintf(stderr, "oracle: level 0: %zu parts, %zu pairs, %.1f s\n", count, levels.back().pairs.size(),
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    std::fflush(stderr);
  }
  std::vector<Site> key(nodes.size());
  for (size_t v = 0; v < nodes.size(); v++) key[v] = lat.around[v][0];
  uint32_t l = lat.level;
  while (count > FEW) {
    l += 1;
    for (auto& k : key) k = coarse_site(k, l);
    std::vector<size_t> up;
    std::vector<int64_t> next;
    size_t bodies = bodies_of(levels.back(), part, key, up, next);
    if (bodies == count) break;
    Level up_level = levels.back().coarser(up, bodies);
    levels.push_back(std::move(up_level));
    part = std::move(next);
    count = bodies;
    if (g_progress > 0) {
      std::fprintf(stderr, "oracle: level %zu: %zu parts, %zu pairs, %.1f s\n", levels.size() - 1, count,
                   levels.back().pairs.size(),
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      std::fflush(stderr);
    }
  }
  Level& top = levels.back();
  top.direct = cholesky_of(top);
  top.has_direct = true;
  std::vector<Six> ld;
  for (size_t v = 0; v < nodes.size(); v++)
    if (node[v] >= 0) ld.push_back(loads(v));
  auto t1 = std::chrono::steady_clock::now();
  double inner = 0.0;
  auto vcyc = [&](const std::vector<Six>& r) { return cycle(levels, 0, r, inner); };
  std::vector<Six> x = conjugate_gradients(levels[0], ld, vcyc, tol, max_iter, st.work, st.iters);
  st.work += inner;
  auto t2 = std::chrono::steady_clock::now();
  st.setup_s = std::chrono::duration<double>(t1 - t0).count();
  st.solve_s = std::chrono::duration<double>(t2 - t1).count();
  st.levels = (int)levels.size();
  st.top_parts = levels.back().centre.size();
  std::vector<Six> out(nodes.size(), Six{});
  for (size_t v = 0; v < nodes.size(); v++)
    if (node[v] >= 0) out[v] = x[node[v]];
  return out;
}

// bodies.rs:420-423
std::vector<V3> moved(const Lattice& lat, const std::vector<Six>& x) {
  std::vector<V3> f(lat.bonds.size());
  for (size_t b = 0; b < lat.bonds.size(); b++) {
    const Bond& bd = lat.bonds[b];
    Six aj = shift(x[bd.j], sub(bd.p, lat.nodes[bd.j].x));
    Six ai = shift(x[bd.i], sub(bd.p, lat.nodes[bd.i].x));
    f[b] = bd.force(sub({aj[0], aj[1], aj[2]}, {ai[0], ai[1], ai[2]}));
  }
  return f;
}

// bodies.rs:407-417
std::vector<V3> forces(const Lattice& lat, const std::vector<char>& alive, const std::vector<char>& gone, double tol,
                       MotionStats& st) {
  auto weight = [&](size_t v) { return Six{0.0, -lat.nodes[v].m * GRAVITY, 0.0, 0.0, 0.0, 0.0}; };
  std::vector<Six> x = motions(lat, alive, gone, weight, tol, 100000, st);
  std::vector<V3> f = moved(lat, x);
  for (size_t b = 0; b < lat.bonds.size(); b++)
    if (!alive[b] || gone[lat.bonds[b].i] || gone[lat.bonds[b].j]) f[b] = V3{0, 0, 0};
  return f;
}

// bodies.rs:544-546
std::vector<double> used(const Lattice& lat, const std::vector<V3>& f) {
  std::vector<double> u(lat.bonds.size());
  for (size_t b = 0; b < lat.bonds.size(); b++) u[b] = lat.bonds[b].used(f[b]);
  return u;
}

// bodies.rs:557-574
std::vector<char> joined_to_held(const std::vector<Node>& nodes, const std::vector<Bond>& bonds,
                                 const std::vector<char>& alive) {
  std::vector<std::vector<size_t>> next(nodes.size());
  for (size_t b = 0; b < bonds.size(); b++)
    if (alive[b]) {
      next[bonds[b].i].push_back(bonds[b].j);
      next[bonds[b].j].push_back(bonds[b].i);
    }
  std::vector<char> held(nodes.size());
  std::vector<size_t> stack;
  for (size_t v = 0; v < nodes.size(); v++) {
    held[v] = nodes[v].held;
    if (held[v]) stack.push_back(v);
  }
  while (!stack.empty()) {
    size_t v = stack.back();
    stack.pop_back();
    for (size_t w : next[v])
      if (!held[w]) {
        held[w] = 1;
        stack.push_back(w);
      }
  }
  return held;
}

struct Outcome {
  double fallen;
  std::vector<char> gone;
  int rounds;
  double work;
};

// bodies.rs:511-541
Outcome settle(const Lattice& lat, double tol) {
  const auto& nodes = lat.nodes;
  const auto& bonds = lat.bonds;
  std::vector<char> alive(bonds.size(), 1), gone(nodes.size(), 0);
  double work = 0.0;
  int rounds = 0;
  for (;;) {
    rounds += 1;
    std::vector<char> held = joined_to_held(nodes, bonds, alive);
    for (size_t v = 0; v < nodes.size(); v++) gone[v] |= !held[v];
    MotionStats st;
    std::vector<V3> f = forces(lat, alive, gone, tol, st);
    work += st.work;
    std::vector<double> u = used(lat, f);
    bool broke = false;
    for (size_t beam0 = 0; beam0 < bonds.size(); beam0 += PER_BEAM) {
      bool any = false;
      for (size_t b = beam0; b < beam0 + PER_BEAM && b < bonds.size(); b++)
        any |= alive[b] && !gone[bonds[b].i] && !gone[bonds[b].j] && u[b] > 1.0;
      if (any) {
        for (size_t b = beam0; b < beam0 + PER_BEAM && b < bonds.size(); b++) alive[b] = 0;
        broke = true;
      }
    }
    if (!broke) break;
  }
  double freem = 0.0, fell = 0.0;
  for (auto& v : nodes)
    if (!v.held) freem += v.m;
  for (size_t v = 0; v < nodes.size(); v++)
    if (gone[v]) fell += nodes[v].m;
  return Outcome{freem > 0.0 ? fell / freem : 0.0, gone, rounds, work};
}

// bodies.rs:550-554
std::vector<double> used_standing(const Lattice& lat, double tol) {
  std::vector<char> alive(lat.bonds.size(), 1);
  std::vector<char> held = joined_to_held(lat.nodes, lat.bonds, alive);
  std::vector<char> gone(held.size());
  for (size_t v = 0; v < held.size(); v++) gone[v] = !held[v];
  MotionStats st;
  return used(lat, forces(lat, alive, gone, tol, st));
}

// bodies.rs:604-625: strength no less than what `generated` needs under its own weight,
// MARGIN times, beam by beam (a beam named by its end nodes' sites, lesser first).
constexpr double MARGIN = 2.0;
void as_generated(Lattice& lattice, const Lattice& generated, double tol) {
  using Key = std::array<Around, 2>;
  auto beam = [](const Lattice& l, const Bond& b) {
    const Around& i = l.around[b.i];
    const Around& j = l.around[b.j];
    return i < j ? Key{i, j} : Key{j, i};
  };
  std::map<Key, double> need;
  std::vector<double> u = used_standing(generated, tol);
  for (size_t b = 0; b < generated.bonds.size(); b++) {
    auto it = need.emplace(beam(generated, generated.bonds[b]), 0.0).first;
    it->second = std::max(it->second, u[b]);
  }
  std::vector<double> by(lattice.bonds.size());
  for (size_t b = 0; b < lattice.bonds.size(); b++) {
    auto it = need.find(beam(lattice, lattice.bonds[b]));
    by[b] = it == need.end() ? 1.0 : std::max(it->second * MARGIN, 1.0);
  }
  // Mat::stronger (model.rs:71-73)
  for (size_t b = 0; b < lattice.bonds.size(); b++) {
    Mat& m = lattice.bonds[b].mat;
    m.ft *= by[b];
    m.fc *= by[b];
    m.fs *= by[b];
  }
}

// ---------------------------------------------------------------- gather.rs
// A coarser level measured from a finer one. Materials are told apart by their
// six values here; the reference tells them apart by name, which is the same for
// its materials (each name has its own values; only bonds' copies are made
// stronger, and gathering reads nodes' materials).

bool same_mat(const Mat& a, const Mat& b) {
  return a.rho == b.rho && a.e == b.e && a.g == b.g && a.ft == b.ft && a.fc == b.fc && a.fs == b.fs;
}
constexpr Mat STONE_MAT{2500.0, 30e9, 12e9, 5e6, 100e6, 10e6};  // model.rs:53

// bodies.rs:628-634
size_t find_root(std::vector<size_t>& p, size_t a) {
  while (p[a] != a) {
    p[a] = p[p[a]];
    a = p[a];
  }
  return a;
}

// gather.rs:25-51: free parts (finer nodes inside), faces (named by the cell across),
// the finer bonds between its places (free parts first, then a handle per face)
struct GCell {
  std::vector<size_t> free, across;
  std::vector<std::array<size_t, 3>> bonds;
  size_t handle(size_t other) {
    size_t h = 0;
    while (h < across.size() && across[h] != other) h++;
    if (h == across.size()) across.push_back(other);
    return free.size() + h;
  }
  std::vector<size_t> pieces() const {
    size_t nf = free.size();
    std::vector<size_t> parent(nf + across.size());
    for (size_t k = 0; k < parent.size(); k++) parent[k] = k;
    for (auto& b : bonds) {
      size_t a = find_root(parent, b[0]), c = find_root(parent, b[1]);
      parent[a] = c;
    }
    std::vector<size_t> out(across.size());
    for (size_t h = 0; h < across.size(); h++) out[h] = find_root(parent, nf + h);
    return out;
  }
};

// gather.rs:165-186: elimination with partial pivoting (the LAST greatest |a[x][c]|, as
// max_by), a direction it cannot tell solved as zero. a: n rows of n + rhs.
std::vector<std::vector<double>> gsolve(std::vector<std::vector<double>> a, size_t rhs) {
  size_t n = a.size();
  double scale = 0.0;
  for (size_t r = 0; r < n; r++) scale = std::max(scale, std::fabs(a[r][r]));
  auto small = [&](double v) { return std::fabs(v) <= 1e-12 * scale; };
  for (size_t c = 0; c < n; c++) {
    size_t p = c;
    for (size_t x = c + 1; x < n; x++)
      if (!(std::fabs(a[x][c]) < std::fabs(a[p][c]))) p = x;
    std::swap(a[c], a[p]);
    if (small(a[c][c])) continue;
    std::vector<double> pivot = a[c];
    for (size_t r = 0; r < n; r++) {
      double f = a[r][c] / pivot[c];
      if (r != c && f != 0.0)
        for (size_t k = c; k < n + rhs; k++) a[r][k] -= f * pivot[k];
    }
  }
  std::vector<std::vector<double>> x(n, std::vector<double>(rhs));
  for (size_t r = 0; r < n; r++)
    for (size_t k = 0; k < rhs; k++) x[r][k] = small(a[r][r]) ? 0.0 : a[r][n + k] / a[r][r];
  return x;
}

// gather.rs:53-114: the cell's matter as a lattice of its own (handles held at `at`), and
// the finer bonds' forces for each handle moved alone along each axis, the free parts
// at rest where that leaves them.
std::pair<Lattice, std::vector<std::vector<V3>>> responses(const Lattice& fine, const GCell& cell,
                                                           const std::vector<V3>& at) {
  size_t nf = cell.free.size(), nh = cell.across.size();
  Lattice lat;
  lat.level = fine.level;
  for (size_t v : cell.free) {
    Node nd = fine.nodes[v];
    nd.held = false;
    lat.nodes.push_back(nd);
  }
  for (const V3& x : at) lat.nodes.push_back(Node{x, 0.0, STONE_MAT, true});
  for (auto& t : cell.bonds) {
    Bond b = fine.bonds[t[2]];
    b.i = t[0];
    b.j = t[1];
    lat.bonds.push_back(b);
  }
  size_t n = 6 * nf, m = 3 * nh;
  auto column = [&](size_t v, size_t d, V3 p) -> V3 {
    V3 e{0.0, 0.0, 0.0};
    e[d % 3] = 1.0;
    return d < 3 ? e : cross(e, sub(p, lat.nodes[v].x));
  };
  auto dof = [&](size_t v, size_t d) -> int64_t {
    if (v < nf) return (int64_t)(6 * v + d);
    if (d < 3) return (int64_t)(n + 3 * (v - nf) + d);
    return -1;
  };
  std::vector<std::vector<double>> a(n, std::vector<double>(n + m, 0.0));
  for (const Bond& b : lat.bonds) {
    M3 st = b.stiffness();
    const std::pair<size_t, double> ends[2] = {{b.i, -1.0}, {b.j, 1.0}};
    for (auto [v, sv] : ends)
      for (size_t d = 0; d < 6; d++) {
        int64_t r = dof(v, d);
        if (r < 0 || (size_t)r >= n) continue;
        V3 cv = column(v, d, b.p);
        V3 scv{dot(st[0], cv), dot(st[1], cv), dot(st[2], cv)};
        for (auto [w, sw] : ends)
          for (size_t e = 0; e < 6; e++) {
            int64_t c = dof(w, e);
            if (c < 0) continue;
            double sign = (size_t)c < n ? 1.0 : -1.0;
            a[r][c] += sign * sv * sw * dot(scv, column(w, e, b.p));
          }
      }
  }
  auto x = gsolve(std::move(a), m);
  std::vector<std::vector<V3>> cases;
  for (size_t c = 0; c < m; c++) {
    std::vector<Six> motion(nf + nh, Six{});
    for (size_t v = 0; v < nf; v++)
      for (size_t d = 0; d < 6; d++) motion[v][d] = x[6 * v + d][c];
    motion[nf + c / 3][c % 3] = 1.0;
    cases.push_back(moved(lat, motion));
  }
  return {std::move(lat), std::move(cases)};
}

struct Found {
  size_t a, b;
  double k, stronger;
};

// gather.rs:116-161: the beams (a < b, stiffness along each) nearest a stiffness between
// handles, none negative (least squares, dropping the most negative till none is; the
// FIRST most negative, as min_by), none between handles the matter does not join.
std::vector<std::tuple<size_t, size_t, double>> beams(const std::vector<std::vector<double>>& k,
                                                      const std::vector<V3>& at,
                                                      const std::function<bool(size_t, size_t)>& joined) {
  size_t n = 3 * at.size();
  std::vector<std::pair<size_t, size_t>> pairs;
  for (size_t a = 0; a < at.size(); a++)
    for (size_t b = a + 1; b < at.size(); b++)
      if (joined(a, b)) pairs.push_back({a, b});
  std::vector<std::vector<double>> units;
  for (auto [a, b] : pairs) {
    V3 e = unit(sub(at[b], at[a]));
    std::vector<double> d(n * n, 0.0);
    for (size_t r = 0; r < 3; r++)
      for (size_t c = 0; c < 3; c++) {
        double v = e[r] * e[c];
        d[(3 * a + r) * n + 3 * a + c] += v;
        d[(3 * b + r) * n + 3 * b + c] += v;
        d[(3 * a + r) * n + 3 * b + c] -= v;
        d[(3 * b + r) * n + 3 * a + c] -= v;
      }
    units.push_back(std::move(d));
  }
  std::vector<double> target(n * n);
  for (size_t i = 0; i < n * n; i++) target[i] = 0.5 * (k[i / n][i % n] + k[i % n][i / n]);
  auto vdot = [](const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
  };
  std::vector<char> on(pairs.size(), 1);
  for (;;) {
    std::vector<size_t> idx;
    for (size_t p = 0; p < pairs.size(); p++)
      if (on[p]) idx.push_back(p);
    size_t m = idx.size();
    std::vector<std::vector<double>> g(m, std::vector<double>(m + 1, 0.0));
    for (size_t r = 0; r < m; r++) {
      for (size_t c = 0; c < m; c++) g[r][c] = vdot(units[idx[r]], units[idx[c]]);
      g[r][m] = vdot(units[idx[r]], target);
    }
    auto x = gsolve(std::move(g), 1);
    int64_t worst = -1;
    for (size_t r = 0; r < m; r++)
      if (x[r][0] < 0.0 && (worst < 0 || x[r][0] < x[(size_t)worst][0])) worst = (int64_t)r;
    if (worst >= 0) {
      on[idx[(size_t)worst]] = 0;
      continue;
    }
    std::vector<std::tuple<size_t, size_t, double>> out;
    for (size_t r = 0; r < m; r++)
      if (x[r][0] > 0.0) out.push_back({pairs[idx[r]].first, pairs[idx[r]].second, x[r][0]});
    return out;
  }
}

using Full = std::array<std::pair<double, double>, 2>;
Lattice measured(const Lattice& fine, bool against_full);
// When set, measured() keeps each measured cell's stiffness between its handles (a test hook).
std::vector<std::vector<double>>* g_keep_k = nullptr;

// lattice.rs:87-102 (solids), for a box all of one material
std::vector<Solid> solids_full(uint32_t level, V3 lo, V3 hi, Mat mat) {
  double s = spacing(level);
  std::vector<Solid> found;
  for (int64_t a = (int64_t)std::floor(lo[0] / s); a < (int64_t)std::ceil(hi[0] / s); a++)
    for (int64_t b = (int64_t)std::floor(lo[1] / s); b < (int64_t)std::ceil(hi[1] / s); b++)
      for (int64_t c = (int64_t)std::floor(lo[2] / s); c < (int64_t)std::ceil(hi[2] / s); c++)
        found.push_back(Solid{{a << level, b << level, c << level}, mat, false});
  return found;
}

// gather.rs:200-223: a full cell's beams by class (section over spacing squared, how much
// stronger than the material), measured once per material
Full full(Mat mat) {
  static std::mutex lock;
  static std::vector<std::pair<Mat, Full>> known;
  {
    std::lock_guard<std::mutex> g(lock);
    for (auto& k : known)
      if (same_mat(k.first, mat)) return k.second;
  }
  double s = spacing(1);
  Lattice coarse = measured(build(0, solids_full(0, {-1.5 * s, -1.5 * s, -1.5 * s}, {1.5 * s, 1.5 * s, 1.5 * s}, mat)),
                            false);
  double sum[2][3] = {{0, 0, 0}, {0, 0, 0}};
  const Site zero{0, 0, 0};
  for (size_t q = 0; q < coarse.bonds.size(); q += PER_BEAM) {
    const Bond& b = coarse.bonds[q];
    auto has = [&](size_t v) { return coarse.around[v][0] == zero || coarse.around[v][1] == zero; };
    if (has(b.i) && has(b.j)) {
      size_t k = cls(coarse.nodes[b.i].x, coarse.nodes[b.j].x, centre(zero, 1));
      sum[k][0] += 4.0 * b.a / (s * s);
      sum[k][1] += b.mat.ft / mat.ft;
      sum[k][2] += 1.0;
    }
  }
  Full f;
  for (int k = 0; k < 2; k++)
    f[k] = sum[k][2] > 0.0 ? std::make_pair(sum[k][0] / sum[k][2], sum[k][1] / sum[k][2]) : std::make_pair(1.0, 1.0);
  std::lock_guard<std::mutex> g(lock);
  known.push_back({mat, f});
  return f;
}

// gather.rs:225-395
Lattice measured(const Lattice& fine, bool against_full) {
  uint32_t level = fine.level + 1;
  std::unordered_map<Site, size_t, SiteHash> index;
  std::vector<Site> sites;
  auto cell_of = [&](Site site) {
    Site key = coarse_site(site, level);
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    sites.push_back(key);
    index.emplace(key, sites.size() - 1);
    return sites.size() - 1;
  };
  std::vector<std::array<size_t, 2>> between;
  for (auto& a : fine.around) {
    size_t x = cell_of(a[0]);
    size_t y = cell_of(a[1]);
    between.push_back({x, y});
  }
  std::vector<size_t> lies;
  for (auto& b : fine.bonds) {
    const Around& ai = fine.around[b.i];
    const Around& aj = fine.around[b.j];
    const Site& s0 = (ai[0] == aj[0] || ai[0] == aj[1]) ? ai[0] : ai[1];
    if (!(s0 == aj[0] || s0 == aj[1])) throw std::runtime_error("a beam is inside a cell");
    lies.push_back(cell_of(s0));
  }
  size_t n = sites.size();
  auto face = [&](size_t a, size_t b) { return sites[a] < sites[b] ? std::make_pair(a, b) : std::make_pair(b, a); };
  std::vector<GCell> cells(n);
  std::vector<std::vector<std::pair<Mat, double>>> mats(n);
  std::map<std::pair<size_t, size_t>, std::pair<double, bool>> faces;
  for (size_t v = 0; v < between.size(); v++) {
    size_t x = between[v][0], y = between[v][1];
    const Node& part = fine.nodes[v];
    for (size_t c : (x == y ? std::vector<size_t>{x} : std::vector<size_t>{x, y})) {
      bool hit = false;
      for (auto& mw : mats[c])
        if (same_mat(mw.first, part.mat)) {
          mw.second += part.m;
          hit = true;
          break;
        }
      if (!hit) mats[c].push_back({part.mat, part.m});
    }
    if (x == y) {
      cells[x].free.push_back(v);
    } else {
      auto& f = faces[face(x, y)];
      f.first += part.m;
      f.second = f.second || part.held;
    }
  }
  std::vector<size_t> place(fine.nodes.size(), SIZE_MAX);
  for (auto& cell : cells)
    for (size_t k = 0; k < cell.free.size(); k++) place[cell.free[k]] = k;
  for (size_t b = 0; b < fine.bonds.size(); b++) {
    size_t c = lies[b];
    auto at = [&](size_t v) {
      size_t x = between[v][0], y = between[v][1];
      if (x == c && y == c) return place[v];
      return cells[c].handle(x == c ? y : x);
    };
    size_t i = at(fine.bonds[b].i);
    size_t j = at(fine.bonds[b].j);
    cells[c].bonds.push_back({i, j, b});
  }
  std::vector<char> held(n);
  for (size_t c = 0; c < n; c++) {
    bool all = !cells[c].free.empty();
    for (size_t v : cells[c].free) all = all && fine.nodes[v].held;
    held[c] = all;
  }
  for (size_t c = 0; c < n; c++)
    for (size_t o : cells[c].across) faces[face(c, o)];
  auto mat_of = [&](std::vector<size_t> cs) {
    std::vector<std::pair<Mat, double>> all;
    for (size_t c : cs)
      for (auto& mw : mats[c]) {
        bool hit = false;
        for (auto& x : all)
          if (same_mat(x.first, mw.first)) {
            x.second += mw.second;
            hit = true;
            break;
          }
        if (!hit) all.push_back(mw);
      }
    if (all.empty()) return STONE_MAT;
    size_t best = 0;
    for (size_t k = 1; k < all.size(); k++)
      if (!(all[k].second < all[best].second)) best = k;  // max_by keeps the last maximum
    return all[best].first;
  };
  auto ccentre = [&](size_t c) { return centre(sites[c], level); };
  std::map<std::pair<size_t, size_t>, size_t> node;
  Lattice out;
  out.level = level;
  for (auto& [f, mh] : faces) {
    node[f] = out.nodes.size();
    out.nodes.push_back(Node{mul(add(ccentre(f.first), ccentre(f.second)), 0.5), mh.first,
                             mat_of({f.first, f.second}), mh.second || held[f.first] || held[f.second]});
    out.around.push_back({sites[f.first], sites[f.second]});
  }
  double s = spacing(level);
  std::vector<std::vector<size_t>> ends(n);
  std::vector<std::vector<V3>> at(n);
  std::vector<Mat> mat(n);
  std::vector<Full> fulls(n);
  for (size_t c = 0; c < n; c++) {
    for (size_t o : cells[c].across) ends[c].push_back(node[face(c, o)]);
    for (size_t k : ends[c]) at[c].push_back(out.nodes[k].x);
    mat[c] = mat_of({c});
  }
  for (size_t c = 0; c < n; c++) fulls[c] = against_full ? full(mat[c]) : Full{{{1.0, 1.0}, {1.0, 1.0}}};
  if (g_keep_k) g_keep_k->assign(n, {});
  auto found = each<std::vector<Found>>(n, [&](size_t c) {
    const GCell& cell = cells[c];
    const std::vector<V3>& pts = at[c];
    const Mat& mt = mat[c];
    const Full& fl = fulls[c];
    std::vector<Found> fd;
    if (held[c] || pts.empty()) return fd;
    auto [lat, cases] = responses(fine, cell, pts);
    size_t nf = cell.free.size(), m = cases.size();
    std::vector<std::vector<double>> k(m, std::vector<double>(m, 0.0));
    for (size_t cs = 0; cs < m; cs++)
      for (size_t b = 0; b < lat.bonds.size(); b++) {
        const Bond& bd = lat.bonds[b];
        const V3& f = cases[cs][b];
        const std::pair<size_t, double> sides[2] = {{bd.j, 1.0}, {bd.i, -1.0}};
        for (auto [v, sign] : sides)
          if (v >= nf)
            for (size_t q = 0; q < 3; q++) k[3 * (v - nf) + q][cs] += sign * f[q];
      }
    if (g_keep_k) {
      std::vector<double> flat;
      for (auto& row : k) flat.insert(flat.end(), row.begin(), row.end());
      static std::mutex keep;
      std::lock_guard<std::mutex> g(keep);
      (*g_keep_k)[c] = std::move(flat);
    }
    std::vector<size_t> piece = cell.pieces();
    for (auto [a, b, stiff] : beams(k, pts, [&](size_t x, size_t y) { return piece[x] == piece[y]; })) {
      V3 e = unit(sub(pts[b], pts[a]));
      std::vector<V3> pulled(lat.bonds.size());
      for (size_t q = 0; q < lat.bonds.size(); q++) {
        V3 f{0.0, 0.0, 0.0};
        for (size_t t = 0; t < 3; t++)
          f = add(f, add(mul(cases[3 * b + t][q], e[t] / 2.0), mul(cases[3 * a + t][q], -e[t] / 2.0)));
        pulled[q] = f;
      }
      double hd = 0.0, work = 0.0;
      std::vector<double> u = used(lat, pulled);
      for (size_t q = 0; q < lat.bonds.size(); q++) {
        double w = lat.bonds[q].energy(pulled[q]);
        hd += w * u[q];
        work += w;
      }
      double fine_used = work > 0.0 ? hd / work : 0.0;
      double length = norm(sub(pts[b], pts[a]));
      double own = mt.used(mt.e / length, 0.0);
      double stronger = fine_used > 0.0 ? own / fine_used : 1.0;
      size_t kc = cls(pts[a], pts[b], ccentre(c));
      double section = stiff * length / mt.e / (s * s);
      if (against_full) section = section * SECTION[kc] / fl[kc].first;
      fd.push_back(Found{a, b, section * s * s * mt.e / length, stronger / fl[kc].second});
    }
    return fd;
  });
  for (size_t c = 0; c < n; c++) {
    const GCell& cell = cells[c];
    double mass = 0.0;
    for (size_t v : cell.free) mass += fine.nodes[v].m;
    std::vector<double> joined(at[c].size(), 0.0);
    for (auto& f : found[c]) {
      joined[f.a] += f.k;
      joined[f.b] += f.k;
    }
    double total = 0.0;
    for (double j : joined) total += j;
    for (size_t h = 0; h < ends[c].size(); h++)
      out.nodes[ends[c][h]].m += mass * (total > 0.0 ? joined[h] / total : 1.0 / (double)at[c].size());
    for (auto& f : found[c])
      if (!(out.nodes[ends[c][f.a]].held && out.nodes[ends[c][f.b]].held)) {
        double length = norm(sub(at[c][f.b], at[c][f.a]));
        Mat st = mat[c];
        st.ft *= f.stronger;
        st.fc *= f.stronger;
        st.fs *= f.stronger;
        beam(ends[c][f.a], ends[c][f.b], at[c][f.a], at[c][f.b], f.k * length / mat[c].e, st, out.bonds);
        out.beam_cell.push_back(c);
      }
  }
  return out;
}

}  // namespace

// ======================================================================= C ABI
// Used from Python (ctypes) by tests and by bench.py's CPU-baseline leg.

extern "C" {

struct OrcLattice;

void orc_set_threads(int t) { g_threads = t; }
void orc_set_progress(int every) { g_progress = every; }
int orc_threads() { return threads_now(); }

// solids: n sites (3 int64 each, level-0 units), mats n x 6 (rho, e, g, ft, fc, fs), held n bytes.
OrcLattice* orc_build(uint32_t level, size_t n, const int64_t* sites, const double* mats, const uint8_t* held) {
  std::vector<Solid> s(n);
  for (size_t i = 0; i < n; i++) {
    s[i].site = {sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]};
    const double* m = mats + 6 * i;
    s[i].mat = Mat{m[0], m[1], m[2], m[3], m[4], m[5]};
    s[i].held = held[i] != 0;
  }
  return reinterpret_cast<OrcLattice*>(new Lattice(build(level, s)));
}

void orc_free(OrcLattice* h) { delete reinterpret_cast<Lattice*>(h); }

void orc_hold_bulk(uint32_t level, size_t n, const int64_t* sites, double radius, uint8_t* held_out) {
  std::vector<Solid> s(n);
  for (size_t i = 0; i < n; i++) s[i].site = {sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]};
  hold_bulk(level, s, radius);
  for (size_t i = 0; i < n; i++) held_out[i] = s[i].held;
}

void orc_sizes(const OrcLattice* h, size_t* nodes, size_t* bonds) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  *nodes = L->nodes.size();
  *bonds = L->bonds.size();
}

// x: 3 per node, m: 1, held: 1 byte, around: 6 int64 (lesser site, then greater)
void orc_nodes(const OrcLattice* h, double* x, double* m, uint8_t* held, int64_t* around) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  for (size_t v = 0; v < L->nodes.size(); v++) {
    for (int k = 0; k < 3; k++) x[3 * v + k] = L->nodes[v].x[k];
    m[v] = L->nodes[v].m;
    held[v] = L->nodes[v].held;
    for (int a = 0; a < 2; a++)
      for (int k = 0; k < 3; k++) around[6 * v + 3 * a + k] = L->around[v][a][k];
  }
}

// ij: 2 per bond, p and n: 3, a, d: 1, e, g: 1, beam_cell: one per beam
void orc_bonds(const OrcLattice* h, int64_t* ij, double* p, double* n, double* a, double* d, double* e, double* g,
               int64_t* beam_cell) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  for (size_t b = 0; b < L->bonds.size(); b++) {
    const Bond& bd = L->bonds[b];
    ij[2 * b] = (int64_t)bd.i;
    ij[2 * b + 1] = (int64_t)bd.j;
    for (int k = 0; k < 3; k++) {
      p[3 * b + k] = bd.p[k];
      n[3 * b + k] = bd.n[k];
    }
    a[b] = bd.a;
    d[b] = bd.d;
    e[b] = bd.mat.e;
    g[b] = bd.mat.g;
  }
  for (size_t q = 0; q < L->beam_cell.size(); q++) beam_cell[q] = (int64_t)L->beam_cell[q];
}

void orc_joined_to_held(const OrcLattice* h, const uint8_t* alive, uint8_t* out) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  std::vector<char> al(alive, alive + L->bonds.size());
  auto r = joined_to_held(L->nodes, L->bonds, al);
  for (size_t v = 0; v < r.size(); v++) out[v] = r[v];
}

// loads: 6 per node (read for parts only). Returns 0, or 1 if the iteration cap was met.
int orc_motions(const OrcLattice* h, const uint8_t* alive, const uint8_t* gone, const double* loads, double tol,
                int max_iter, double* x_out, int* iters, double* work, double* setup_s, double* solve_s,
                int* levels) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  std::vector<char> al(alive, alive + L->bonds.size()), gn(gone, gone + L->nodes.size());
  MotionStats st;
  auto ld = [&](size_t v) {
    Six s;
    for (int k = 0; k < 6; k++) s[k] = loads[6 * v + k];
    return s;
  };
  auto x = motions(*L, al, gn, ld, tol, max_iter, st);
  for (size_t v = 0; v < x.size(); v++)
    for (int k = 0; k < 6; k++) x_out[6 * v + k] = x[v][k];
  if (iters) *iters = st.iters;
  if (work) *work = st.work;
  if (setup_s) *setup_s = st.setup_s;
  if (solve_s) *solve_s = st.solve_s;
  if (levels) *levels = st.levels;
  return st.iters >= max_iter ? 1 : 0;
}

// Level 0's stiffness applied to x (6 per node, zero rows for held and gone nodes), as Level::apply.
void orc_apply(const OrcLattice* h, const uint8_t* alive, const uint8_t* gone, const double* x, double* y) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  std::vector<size_t> live;
  for (size_t b = 0; b < L->bonds.size(); b++)
    if (alive[b] && !gone[L->bonds[b].i] && !gone[L->bonds[b].j]) live.push_back(b);
  std::vector<int64_t> part(L->nodes.size(), -1);
  size_t count = 0;
  for (size_t v = 0; v < L->nodes.size(); v++)
    if (!L->nodes[v].held && !gone[v]) part[v] = (int64_t)count++;
  Level lv = level_new(*L, part, count, live);
  std::vector<Six> xs(count);
  for (size_t v = 0; v < L->nodes.size(); v++)
    if (part[v] >= 0)
      for (int k = 0; k < 6; k++) xs[part[v]][k] = x[6 * v + k];
  double w = 0.0;
  auto ys = lv.apply(xs, w);
  for (size_t v = 0; v < L->nodes.size(); v++)
    for (int k = 0; k < 6; k++) y[6 * v + k] = part[v] >= 0 ? ys[part[v]][k] : 0.0;
}

void orc_moved(const OrcLattice* h, const double* x, double* f_out) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  std::vector<Six> xs(L->nodes.size());
  for (size_t v = 0; v < xs.size(); v++)
    for (int k = 0; k < 6; k++) xs[v][k] = x[6 * v + k];
  auto f = moved(*L, xs);
  for (size_t b = 0; b < f.size(); b++)
    for (int k = 0; k < 3; k++) f_out[3 * b + k] = f[b][k];
}

void orc_forces(const OrcLattice* h, const uint8_t* alive, const uint8_t* gone, double tol, double* f_out) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  std::vector<char> al(alive, alive + L->bonds.size()), gn(gone, gone + L->nodes.size());
  MotionStats st;
  auto f = forces(*L, al, gn, tol, st);
  for (size_t b = 0; b < f.size(); b++)
    for (int k = 0; k < 3; k++) f_out[3 * b + k] = f[b][k];
}

// Bond::used with the bond's own material; mats_used lets a test give the strengths (ft, fc, fs) per bond.
void orc_used(const OrcLattice* h, const double* f, const double* strength, double* u_out) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  for (size_t b = 0; b < L->bonds.size(); b++) {
    Bond bd = L->bonds[b];
    if (strength) {
      bd.mat.ft = strength[3 * b];
      bd.mat.fc = strength[3 * b + 1];
      bd.mat.fs = strength[3 * b + 2];
    }
    u_out[b] = bd.used({f[3 * b], f[3 * b + 1], f[3 * b + 2]});
  }
}

void orc_used_standing(const OrcLattice* h, double tol, double* u_out) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  auto u = used_standing(*L, tol);
  for (size_t b = 0; b < u.size(); b++) u_out[b] = u[b];
}

void orc_settle(const OrcLattice* h, double tol, uint8_t* gone_out, double* fallen, int* rounds, double* work) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  Outcome o = settle(*L, tol);
  for (size_t v = 0; v < o.gone.size(); v++) gone_out[v] = o.gone[v];
  *fallen = o.fallen;
  *rounds = o.rounds;
  *work = o.work;
}

// Scales the stiffness (e and g) of every bond of cell `cell` by `by` (a log-stiffness step of log(by)).
void orc_as_generated(OrcLattice* h, const OrcLattice* generated, double tol) {
  as_generated(*reinterpret_cast<Lattice*>(h), *reinterpret_cast<const Lattice*>(generated), tol);
}

void orc_scale_cell(OrcLattice* h, int64_t cell, double by) {
  auto* L = reinterpret_cast<Lattice*>(h);
  for (size_t q = 0; q < L->beam_cell.size(); q++)
    if ((int64_t)L->beam_cell[q] == cell)
      for (size_t b = q * PER_BEAM; b < (q + 1) * PER_BEAM; b++) {
        L->bonds[b].mat.e *= by;
        L->bonds[b].mat.g *= by;
      }
}

// Test hooks for the small dense routines.
void orc_invert6(const double* m, double* out) {
  M6 a;
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) a[r][c] = m[6 * r + c];
  M6 inv = invert(a);
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) out[6 * r + c] = inv[r][c];
}

void orc_coarse_site(const int64_t* p, uint32_t l, int64_t* out) {
  Site s = coarse_site({p[0], p[1], p[2]}, l);
  for (int k = 0; k < 3; k++) out[k] = s[k];
}

double orc_spacing(uint32_t level) { return spacing(level); }

// gather::gathered (against_full = 1) or measured(fine, false): a new lattice, freed by orc_free.
OrcLattice* orc_gathered(const OrcLattice* h, int against_full) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  return reinterpret_cast<OrcLattice*>(new Lattice(measured(*L, against_full != 0)));
}

uint32_t orc_level(const OrcLattice* h) { return reinterpret_cast<const Lattice*>(h)->level; }

// materials: 6 per node (rho, e, g, ft, fc, fs), 6 per bond
void orc_mats(const OrcLattice* h, double* node_mat, double* bond_mat) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  auto put = [](const Mat& m, double* o) {
    o[0] = m.rho, o[1] = m.e, o[2] = m.g, o[3] = m.ft, o[4] = m.fc, o[5] = m.fs;
  };
  for (size_t v = 0; v < L->nodes.size(); v++) put(L->nodes[v].mat, node_mat + 6 * v);
  for (size_t b = 0; b < L->bonds.size(); b++) put(L->bonds[b].mat, bond_mat + 6 * b);
}

// measured(fine, false)'s stiffness between handles of each measured cell (gather.rs:291-303):
// k_out gets, cell after cell, nh_out[i] then its (3 nh)^2 entries row by row; returns how many
// cells (at most cap) and the doubles written in *len.
size_t orc_cell_stiffness(const OrcLattice* h, size_t cap, double* k_out, size_t* len) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  std::vector<std::vector<double>> ks;
  g_keep_k = &ks;
  Lattice out = measured(*L, false);
  g_keep_k = nullptr;
  size_t cells = 0, at = 0;
  for (auto& k : ks) {
    if (k.empty() || cells == cap) continue;
    size_t m = (size_t)std::llround(std::sqrt((double)k.size()));
    k_out[at++] = (double)(m / 3);
    for (double v : k) k_out[at++] = v;
    cells++;
  }
  *len = at;
  return cells;
}

// Bond::energy (model.rs:132-136) of each bond bearing f (3 per bond)
void orc_energy(const OrcLattice* h, const double* f, double* out) {
  auto* L = reinterpret_cast<const Lattice*>(h);
  for (size_t b = 0; b < L->bonds.size(); b++) out[b] = L->bonds[b].energy({f[3 * b], f[3 * b + 1], f[3 * b + 2]});
}

// gather.rs:165-186 on an n x (n + rhs) row-major matrix; x: n x rhs
void orc_gather_solve(size_t n, size_t rhs, const double* a, double* x) {
  std::vector<std::vector<double>> m(n, std::vector<double>(n + rhs));
  for (size_t r = 0; r < n; r++)
    for (size_t c = 0; c < n + rhs; c++) m[r][c] = a[r * (n + rhs) + c];
  auto s = gsolve(std::move(m), rhs);
  for (size_t r = 0; r < n; r++)
    for (size_t k = 0; k < rhs; k++) x[r * rhs + k] = s[r][k];
}

// gather.rs:116-161 for handles at `at` (nh x 3), k (3nh x 3nh), joined (nh x nh bytes).
// out: up to nh(nh-1)/2 triples (a, b, stiffness); returns how many.
size_t orc_beams(size_t nh, const double* at, const double* k, const uint8_t* joined, double* out) {
  std::vector<V3> pts(nh);
  for (size_t h = 0; h < nh; h++) pts[h] = {at[3 * h], at[3 * h + 1], at[3 * h + 2]};
  size_t n = 3 * nh;
  std::vector<std::vector<double>> km(n, std::vector<double>(n));
  for (size_t r = 0; r < n; r++)
    for (size_t c = 0; c < n; c++) km[r][c] = k[r * n + c];
  auto found = beams(km, pts, [&](size_t a, size_t b) { return joined[a * nh + b] != 0; });
  for (size_t q = 0; q < found.size(); q++) {
    out[3 * q] = (double)std::get<0>(found[q]);
    out[3 * q + 1] = (double)std::get<1>(found[q]);
    out[3 * q + 2] = std::get<2>(found[q]);
  }
  return found.size();
}
}
