// The oracle of the mechanics solver: a CPU FP64 restatement of the reference's
// mechanics::solve (a symmetric system over cells, three unknowns to a cell,
// preconditioned CG with V-cycles over rigid-body parts of the octree), for tests and the
// CPU baseline only. It is test infrastructure: nothing in the product links, loads or
// calls it.
//
// Followed, function by function (paths relative to the reference repo):
//   mechanics/src/solve.rs    System::times 61-72, lift 76-98, embed 100-106,
//                             transpose6 108-110, plus6 112-118, times6/times6t 120-126,
//                             inverse 129-139, cholesky 141-166, solve_cholesky 168-187,
//                             Coarse 190-207, aggregate 210-213, find/union/sets 216-250,
//                             coarsen 253-301, step 309-327, Coarse::apply 330-340,
//                             Coarse::factor 343-372, Solver::new 387-420,
//                             Solver::cycle 423-469, Solver::precondition 472-497,
//                             Solver::solve 508-537
//   mechanics/src/element.rs  the unit cube's stiffness (strain 26-32, stress_of 35-40,
//                             unit 49-76) and its turned blocks (stiffness 145-150), so that
//                             tests build systems as mechanics/src/lib.rs:638-676 does
// Rust semantics mirrored: no contraction (built with -ffp-contract=off); sums are
// sequential left folds in the reference's order; union at the lesser root; parts
// numbered as first met. The one order the reference leaves to a hash map (coarsen's
// pairs, solve.rs:280-296, a FastMap iterated) is taken here in (p, q) order: it changes
// only the order of a level's product's sums, so rounding.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace {

using V3 = std::array<double, 3>;
using M3 = std::array<V3, 3>;
using V6 = std::array<double, 6>;
using M6 = std::array<V6, 6>;
using Key = std::array<int64_t, 3>;

constexpr int kRounds = 300;  // solve.rs:21
constexpr size_t kDirect = 24;  // solve.rs:23

V3 add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V3 sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 scale(V3 a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross(V3 a, V3 b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
V3 times(const M3& m, V3 v) { return {dot(m[0], v), dot(m[1], v), dot(m[2], v)}; }
V3 times_t(const M3& m, V3 v) {
  V3 o;
  for (int c = 0; c < 3; c++) {
    double s = 0.0;
    for (int r = 0; r < 3; r++) s += m[r][c] * v[r];
    o[c] = s;
  }
  return o;
}

struct System {
  std::vector<Key> sites;
  std::vector<V3> at;
  std::vector<M3> diag;
  std::vector<std::pair<std::array<uint32_t, 2>, M3>> pairs;
  // solve.rs:61-72
  void apply(const std::vector<V3>& x, std::vector<V3>& y) const {
    for (size_t i = 0; i < diag.size(); i++) y[i] = times(diag[i], x[i]);
    for (const auto& p : pairs) {
      const size_t i = p.first[0], j = p.first[1];
      y[i] = add(y[i], times(p.second, x[j]));
      y[j] = add(y[j], times_t(p.second, x[i]));
    }
  }
};

// solve.rs:76-98: T(r)^T M T(s)
M6 lift(const M6& m, V3 r, V3 s) {
  auto x = [](V3 r) -> M3 { return M3{V3{0.0, r[2], -r[1]}, V3{-r[2], 0.0, r[0]}, V3{r[1], -r[0], 0.0}}; };
  const M3 xr = x(r), xs = x(s);
  M6 t = m;
  for (int row = 0; row < 6; row++)
    for (int c = 0; c < 3; c++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc += m[row][k] * xs[k][c];
      t[row][3 + c] += acc;
    }
  M6 out = t;
  for (int r3 = 0; r3 < 3; r3++)
    for (int col = 0; col < 6; col++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc += xr[k][r3] * t[k][col];
      out[3 + r3][col] += acc;
    }
  return out;
}

M6 embed(const M3& b) {  // solve.rs:100-106
  M6 m{};
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) m[r][c] = b[r][c];
  return m;
}
M6 transpose6(const M6& m) {
  M6 t;
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) t[r][c] = m[c][r];
  return t;
}
void plus6(M6& a, const M6& b) {
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) a[r][c] += b[r][c];
}
V6 times6(const M6& m, const V6& v) {
  V6 o;
  for (int r = 0; r < 6; r++) {
    double s = 0.0;
    for (int c = 0; c < 6; c++) s += m[r][c] * v[c];
    o[r] = s;
  }
  return o;
}
V6 times6t(const M6& m, const V6& v) {
  V6 o;
  for (int r = 0; r < 6; r++) {
    double s = 0.0;
    for (int c = 0; c < 6; c++) s += m[c][r] * v[c];
    o[r] = s;
  }
  return o;
}

// solve.rs:141-166: L L^T in place; pivots too small to keep dropped
std::vector<bool> cholesky(std::vector<double>& a, size_t n) {
  double largest = 0.0;
  for (size_t i = 0; i < n; i++) largest = std::fmax(largest, a[i * n + i]);
  std::vector<bool> dropped(n, false);
  for (size_t j = 0; j < n; j++) {
    double d = a[j * n + j];
    for (size_t k = 0; k < j; k++) d -= a[j * n + k] * a[j * n + k];
    if (d <= largest * 1e-12) {
      dropped[j] = true;
      for (size_t i = j; i < n; i++) a[i * n + j] = 0.0;
      continue;
    }
    d = std::sqrt(d);
    a[j * n + j] = d;
    for (size_t i = j + 1; i < n; i++) {
      double s = a[i * n + j];
      for (size_t k = 0; k < j; k++) s -= a[i * n + k] * a[j * n + k];
      a[i * n + j] = s / d;
    }
  }
  return dropped;
}

// solve.rs:168-187
void solve_cholesky(const std::vector<double>& a, const std::vector<bool>& dropped, size_t n, double* b) {
  for (size_t i = 0; i < n; i++) {
    if (dropped[i]) {
      b[i] = 0.0;
      continue;
    }
    double s = 0.0;
    for (size_t k = 0; k < i; k++) s += a[i * n + k] * b[k];
    b[i] = (b[i] - s) / a[i * n + i];
  }
  for (size_t i = n; i-- > 0;) {
    if (dropped[i]) {
      b[i] = 0.0;
      continue;
    }
    double s = 0.0;
    for (size_t k = i + 1; k < n; k++) s += a[k * n + i] * b[k];
    b[i] = (b[i] - s) / a[i * n + i];
  }
}

// solve.rs:129-139
template <size_t N>
std::array<std::array<double, N>, N> inverse(const std::array<std::array<double, N>, N>& m) {
  std::vector<double> a(N * N);
  for (size_t r = 0; r < N; r++)
    for (size_t c = 0; c < N; c++) a[r * N + c] = m[r][c];
  const std::vector<bool> l = cholesky(a, N);
  std::array<std::array<double, N>, N> out{};
  for (size_t c = 0; c < N; c++) {
    std::vector<double> e(N, 0.0);
    e[c] = 1.0;
    solve_cholesky(a, l, N, e.data());
    for (size_t r = 0; r < N; r++) out[c][r] = e[r];  // from_fn(|c| from_fn(|r| e[r])): row c holds column c
  }
  return out;
}

// solve.rs:216-250
size_t find(std::vector<size_t>& up, size_t i) {
  while (up[i] != i) {
    up[i] = up[up[i]];
    i = up[i];
  }
  return i;
}
void unite(std::vector<size_t>& up, size_t i, size_t j) {
  const size_t r = find(up, i), s = find(up, j);
  up[std::max(r, s)] = std::min(r, s);
}

int64_t div_euclid2(int64_t v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }
Key parent(const Key& k) { return {div_euclid2(k[0]), div_euclid2(k[1]), div_euclid2(k[2])}; }

// sets + aggregate (solve.rs:210-213, 237-250): the pairs within one octree cell a level up
std::vector<uint32_t> aggregate(const std::vector<Key>& keys, const std::vector<std::array<uint32_t, 2>>& pairs,
                                size_t& count) {
  const size_t n = keys.size();
  std::vector<size_t> up(n);
  for (size_t i = 0; i < n; i++) up[i] = i;
  for (const auto& p : pairs)
    if (parent(keys[p[0]]) == parent(keys[p[1]])) unite(up, p[0], p[1]);
  std::vector<uint32_t> index(n, UINT32_MAX), of(n);
  count = 0;
  for (size_t i = 0; i < n; i++) {
    const size_t root = find(up, i);
    if (index[root] == UINT32_MAX) index[root] = (uint32_t)count++;
    of[i] = index[root];
  }
  return of;
}

struct Coarse {
  std::vector<uint32_t> of;
  std::vector<V3> r;
  std::vector<Key> keys;
  std::vector<V3> centre;
  std::vector<M6> diag;
  std::vector<std::pair<std::array<uint32_t, 2>, M6>> pairs;
  std::vector<M6> inv;
  bool direct = false;
  std::vector<double> fact;
  std::vector<bool> dropped;
  double size = 0.0, step = 1.0;

  void apply(const std::vector<V6>& x, std::vector<V6>& y) const {  // solve.rs:330-340
    for (size_t p = 0; p < diag.size(); p++) y[p] = times6(diag[p], x[p]);
    for (const auto& pr : pairs) {
      const size_t p = pr.first[0], q = pr.first[1];
      const V6 a = times6(pr.second, x[q]), c = times6t(pr.second, x[p]);
      for (int k = 0; k < 6; k++) {
        y[p][k] += a[k];
        y[q][k] += c[k];
      }
    }
  }
  void factor() {  // solve.rs:343-372
    const size_t n = centre.size(), m = 6 * n;
    std::vector<double> a(36 * n * n, 0.0);
    for (size_t p = 0; p < n; p++) {
      const M6& d = diag[p];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) a[(6 * p + r) * m + 6 * p + c] += d[r][c];
      const double t = (d[0][0] + d[1][1] + d[2][2]) / 3.0;
      for (int k = 3; k < 6; k++) a[(6 * p + k) * m + 6 * p + k] += 1e-6 * t * size * size;
    }
    for (const auto& pr : pairs) {
      const size_t p = pr.first[0], q = pr.first[1];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) {
          a[(6 * p + r) * m + 6 * q + c] += pr.second[r][c];
          a[(6 * q + c) * m + 6 * p + r] += pr.second[r][c];
        }
    }
    dropped = cholesky(a, m);
    fact = std::move(a);
    direct = true;
  }
};

// solve.rs:309-327: 4/3 over the greatest eigenvalue of D^-1 A by ten powers
template <size_t N, class Apply, class Over>
double step_of(size_t n, Apply apply, Over over) {
  using VN = std::array<double, N>;
  auto norm = [](const std::vector<VN>& v) {
    double s = 0.0;
    for (const auto& e : v)
      for (double x : e) s += x * x;
    return std::sqrt(s);
  };
  std::vector<VN> v(n), a(n), w(n);
  for (size_t k = 0; k < n; k++)
    for (size_t s = 0; s < N; s++) v[k][s] = 1.0 + (double)((7 * k + 3 * s) % 11) / 11.0;
  double most = 1.0;
  for (int it = 0; it < 10; it++) {
    apply(v, a);
    over(a, w);
    const double nw = norm(w), nv = norm(v);
    if (nw == 0.0) break;
    most = nw / nv;
    for (size_t k = 0; k < n; k++)
      for (size_t s = 0; s < N; s++) v[k][s] = w[k][s] / nw;
  }
  return 4.0 / 3.0 / most;
}

// solve.rs:253-301
std::unique_ptr<Coarse> coarsen(const std::vector<Key>& keys, const std::vector<V3>& at, const std::vector<M6>& diag,
                                const std::vector<std::pair<std::array<uint32_t, 2>, M6>>& pairs, double size) {
  std::vector<std::array<uint32_t, 2>> ij(pairs.size());
  for (size_t k = 0; k < pairs.size(); k++) ij[k] = pairs[k].first;
  size_t n = 0;
  auto c = std::make_unique<Coarse>();
  c->of = aggregate(keys, ij, n);
  c->centre.assign(n, V3{0, 0, 0});
  std::vector<double> count(n, 0.0);
  c->keys.assign(n, Key{0, 0, 0});
  for (size_t i = 0; i < c->of.size(); i++) {
    const uint32_t a = c->of[i];
    c->centre[a] = add(c->centre[a], at[i]);
    count[a] += 1.0;
    c->keys[a] = parent(keys[i]);
  }
  for (size_t a = 0; a < n; a++) c->centre[a] = scale(c->centre[a], 1.0 / count[a]);
  c->r.resize(c->of.size());
  for (size_t i = 0; i < c->of.size(); i++) c->r[i] = sub(at[i], c->centre[c->of[i]]);
  c->diag.assign(n, M6{});
  for (size_t i = 0; i < diag.size(); i++) plus6(c->diag[c->of[i]], lift(diag[i], c->r[i], c->r[i]));
  std::map<std::pair<uint32_t, uint32_t>, M6> map;
  for (const auto& pr : pairs) {
    const size_t i = pr.first[0], j = pr.first[1];
    const uint32_t a = c->of[i], cc = c->of[j];
    const M6 m = lift(pr.second, c->r[i], c->r[j]);
    if (a == cc) {
      plus6(c->diag[a], m);
      plus6(c->diag[a], transpose6(m));
    } else if (a < cc) {
      auto it = map.emplace(std::make_pair(a, cc), M6{}).first;
      plus6(it->second, m);
    } else {
      auto it = map.emplace(std::make_pair(cc, a), M6{}).first;
      plus6(it->second, transpose6(m));
    }
  }
  for (auto& kv : map) c->pairs.push_back({{kv.first.first, kv.first.second}, kv.second});
  c->inv.resize(n);
  for (size_t p = 0; p < n; p++) {
    M6 d = c->diag[p];
    const double t = (d[0][0] + d[1][1] + d[2][2]) / 3.0;
    for (int k = 3; k < 6; k++) d[k][k] += 1e-6 * t * size * size;
    c->inv[p] = inverse<6>(d);
  }
  c->size = size;
  c->step = 1.0;
  Coarse* cp = c.get();
  c->step = step_of<6>(
      n, [cp](const std::vector<V6>& x, std::vector<V6>& y) { cp->apply(x, y); },
      [cp](const std::vector<V6>& x, std::vector<V6>& y) {
        for (size_t p = 0; p < x.size(); p++) y[p] = times6(cp->inv[p], x[p]);
      });
  if (n <= kDirect) c->factor();
  return c;
}

struct Solver {
  std::vector<M3> inv;
  double step = 1.0;
  std::vector<std::unique_ptr<Coarse>> levels;

  Solver(const System& s, double size) {  // solve.rs:387-420
    inv.resize(s.diag.size());
    for (size_t i = 0; i < s.diag.size(); i++) inv[i] = inverse<3>(s.diag[i]);
    step = step_of<3>(
        s.sites.size(), [&s](const std::vector<V3>& x, std::vector<V3>& y) { s.apply(x, y); },
        [this](const std::vector<V3>& x, std::vector<V3>& y) {
          for (size_t i = 0; i < x.size(); i++) y[i] = times(inv[i], x[i]);
        });
    std::vector<M6> diag(s.diag.size());
    for (size_t i = 0; i < s.diag.size(); i++) diag[i] = embed(s.diag[i]);
    std::vector<std::pair<std::array<uint32_t, 2>, M6>> pairs(s.pairs.size());
    for (size_t k = 0; k < s.pairs.size(); k++) pairs[k] = {s.pairs[k].first, embed(s.pairs[k].second)};
    double sz = size;
    auto next = coarsen(s.sites, s.at, diag, pairs, sz * 2.0);
    for (;;) {
      sz *= 2.0;
      const size_t n = next->centre.size(), before = levels.empty() ? s.sites.size() : levels.back()->centre.size();
      if (next->direct) {
        levels.push_back(std::move(next));
        break;
      }
      if (n == before) {
        next->factor();
        levels.push_back(std::move(next));
        break;
      }
      levels.push_back(std::move(next));
      const Coarse& l = *levels.back();
      next = coarsen(l.keys, l.centre, l.diag, l.pairs, sz * 2.0);
    }
  }

  std::vector<V6> cycle(size_t l, const std::vector<V6>& b) const {  // solve.rs:423-469
    const Coarse& level = *levels[l];
    const size_t n = level.centre.size();
    if (level.direct) {
      std::vector<double> x(6 * n);
      for (size_t p = 0; p < n; p++)
        for (int k = 0; k < 6; k++) x[6 * p + k] = b[p][k];
      solve_cholesky(level.fact, level.dropped, 6 * n, x.data());
      std::vector<V6> out(n);
      for (size_t p = 0; p < n; p++)
        for (int k = 0; k < 6; k++) out[p][k] = x[6 * p + k];
      return out;
    }
    auto smooth = [&](std::vector<V6>& x, const std::vector<V6>& r) {
      for (size_t p = 0; p < n; p++) {
        const V6 d = times6(level.inv[p], r[p]);
        for (int k = 0; k < 6; k++) x[p][k] = x[p][k] + level.step * d[k];
      }
    };
    auto residual = [&](const std::vector<V6>& x) {
      std::vector<V6> y(n);
      level.apply(x, y);
      for (size_t p = 0; p < n; p++)
        for (int k = 0; k < 6; k++) y[p][k] = b[p][k] - y[p][k];
      return y;
    };
    std::vector<V6> x(n, V6{});
    smooth(x, b);
    std::vector<V6> r = residual(x);
    const Coarse& up = *levels[l + 1];
    std::vector<V6> rb(up.centre.size(), V6{});
    for (size_t p = 0; p < n; p++) {
      const size_t a = up.of[p];
      const V3 t = cross(up.r[p], V3{r[p][0], r[p][1], r[p][2]});
      for (int k = 0; k < 3; k++) {
        rb[a][k] += r[p][k];
        rb[a][3 + k] += r[p][3 + k] + t[k];
      }
    }
    const std::vector<V6> e = cycle(l + 1, rb);
    for (size_t p = 0; p < n; p++) {
      const size_t a = up.of[p];
      const V3 u = cross(V3{e[a][3], e[a][4], e[a][5]}, up.r[p]);
      for (int k = 0; k < 3; k++) {
        x[p][k] += e[a][k] + u[k];
        x[p][3 + k] += e[a][3 + k];
      }
    }
    r = residual(x);
    smooth(x, r);
    return x;
  }

  // solve.rs:472-497
  void precondition(const System& s, const std::vector<V3>& b, std::vector<V3>& out, std::vector<V3>& scratch) const {
    const size_t n = b.size();
    for (size_t i = 0; i < n; i++) out[i] = scale(times(inv[i], b[i]), step);
    s.apply(out, scratch);
    const Coarse& up = *levels[0];
    std::vector<V6> rb(up.centre.size(), V6{});
    for (size_t i = 0; i < n; i++) {
      const V3 r = sub(b[i], scratch[i]);
      const size_t a = up.of[i];
      const V3 t = cross(up.r[i], r);
      for (int k = 0; k < 3; k++) {
        rb[a][k] += r[k];
        rb[a][3 + k] += t[k];
      }
    }
    const std::vector<V6> e = cycle(0, rb);
    for (size_t i = 0; i < n; i++) {
      const size_t a = up.of[i];
      out[i] = add(out[i], add(V3{e[a][0], e[a][1], e[a][2]}, cross(V3{e[a][3], e[a][4], e[a][5]}, up.r[i])));
    }
    s.apply(out, scratch);
    for (size_t i = 0; i < n; i++) out[i] = add(out[i], scale(times(inv[i], sub(b[i], scratch[i])), step));
  }

  // solve.rs:508-537; -1 for None
  int64_t solve(const System& s, const std::vector<V3>& b, std::vector<V3>& x, double tolerance) const {
    const size_t n = b.size();
    if (n == 0) return 0;
    auto measure = [](const std::vector<V3>& r, const std::vector<V3>& z) {
      double acc = 0.0;
      for (size_t i = 0; i < r.size(); i++) acc += dot(r[i], z[i]);
      return acc;
    };
    std::vector<V3> z(n, V3{0, 0, 0}), scratch(n, V3{0, 0, 0});
    precondition(s, b, z, scratch);
    const double target = tolerance * tolerance * measure(b, z);
    if (std::isnan(target) || target < 0.0) return -1;
    std::vector<V3> ax(n);
    s.apply(x, ax);
    std::vector<V3> r(n);
    for (size_t i = 0; i < n; i++) r[i] = sub(b[i], ax[i]);
    precondition(s, r, z, scratch);
    std::vector<V3> p = z;
    double rz = measure(r, z);
    std::vector<V3> ap(n);
    for (int round = 0; round < kRounds; round++) {
      if (rz < 0.0) return -1;
      if (rz <= target) return round;
      s.apply(p, ap);
      double pap = 0.0;
      for (size_t i = 0; i < n; i++) pap += dot(p[i], ap[i]);
      if (pap <= 0.0) return round;
      const double alpha = rz / pap;
      for (size_t i = 0; i < n; i++) {
        x[i] = add(x[i], scale(p[i], alpha));
        r[i] = sub(r[i], scale(ap[i], alpha));
      }
      precondition(s, r, z, scratch);
      const double next = measure(r, z);
      const double beta = next / rz;
      rz = next;
      for (size_t i = 0; i < n; i++) p[i] = add(z[i], scale(p[i], beta));
    }
    return kRounds;
  }
};

System system_of(int64_t n, const int64_t* sites, const double* at, const double* diag, int64_t np, const uint32_t* ij,
                 const double* blocks) {
  System s;
  s.sites.resize(n);
  s.at.resize(n);
  s.diag.resize(n);
  for (int64_t i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      s.sites[i][k] = sites ? sites[3 * i + k] : 0;
      s.at[i][k] = at ? at[3 * i + k] : 0.0;
    }
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) s.diag[i][r][c] = diag[9 * i + 3 * r + c];
  }
  s.pairs.resize(np);
  for (int64_t k = 0; k < np; k++) {
    s.pairs[k].first = {ij[2 * k], ij[2 * k + 1]};
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) s.pairs[k].second[r][c] = blocks[9 * k + 3 * r + c];
  }
  return s;
}

// element.rs:14-76: Poisson's ratio, the corners, a node's strain at a point, stress of
// strain, the unit cube's stiffness by 2x2x2 Gauss points
constexpr double kNu = 0.3;
V3 sign_of(int c) { return {(double)(2 * (c & 1) - 1), (double)(2 * ((c >> 1) & 1) - 1), (double)(2 * ((c >> 2) & 1) - 1)}; }
std::array<V3, 6> strain(int a, V3 at) {
  const V3 s = sign_of(a);
  V3 g;
  for (int k = 0; k < 3; k++) {
    double prod = 1.0;
    for (int j = 0; j < 3; j++)
      if (j != k) prod *= 1.0 + s[j] * at[j];
    g[k] = s[k] / 4.0 * prod;
  }
  return {V3{g[0], 0.0, 0.0}, V3{0.0, g[1], 0.0}, V3{0.0, 0.0, g[2]},
          V3{0.0, g[2], g[1]}, V3{g[2], 0.0, g[0]}, V3{g[1], g[0], 0.0}};
}
std::array<double, 6> stress_of(const std::array<double, 6>& e) {
  const double l = kNu / ((1.0 + kNu) * (1.0 - 2.0 * kNu)), m = 1.0 / (2.0 * (1.0 + kNu));
  const double tr = e[0] + e[1] + e[2];
  return {l * tr + 2.0 * m * e[0], l * tr + 2.0 * m * e[1], l * tr + 2.0 * m * e[2], m * e[3], m * e[4], m * e[5]};
}
void unit_k(double* k) {  // k[((a * 8 + c) * 3 + i) * 3 + j]
  const double g = 1.0 / std::sqrt(3.0);
  std::memset(k, 0, sizeof(double) * 8 * 8 * 9);
  for (int p = 0; p < 8; p++) {
    const V3 sp = sign_of(p);
    const V3 at = {sp[0] * g, sp[1] * g, sp[2] * g};
    std::array<std::array<V3, 6>, 8> b;
    for (int a = 0; a < 8; a++) b[a] = strain(a, at);
    for (int a = 0; a < 8; a++)
      for (int c = 0; c < 8; c++)
        for (int j = 0; j < 3; j++) {
          std::array<double, 6> e;
          for (int r = 0; r < 6; r++) e[r] = b[c][r][j];
          const std::array<double, 6> s = stress_of(e);
          for (int i = 0; i < 3; i++) {
            double acc = 0.0;
            for (int r = 0; r < 6; r++) acc += b[a][r][i] * s[r];
            k[((a * 8 + c) * 3 + i) * 3 + j] += acc / 8.0;
          }
        }
  }
}

}  // namespace

extern "C" {

// Solver::new (solve.rs:387-420) on a system given as flat arrays: sites (3 i64 per
// unknown), at (3 f64), diag (9 f64 row-major), pairs (2 u32 + 9 f64 each).
void* mor_new(int64_t n, const int64_t* sites, const double* at, const double* diag, int64_t np, const uint32_t* ij,
              const double* blocks, double size) {
  const System s = system_of(n, sites, at, diag, np, ij, blocks);
  return new Solver(s, size);
}
void mor_free(void* h) { delete static_cast<Solver*>(h); }

// Solver::solve (solve.rs:508-537): x holds the start and receives the solution; the
// rounds taken, or -1 for None.
int64_t mor_solve(void* h, int64_t n, const double* diag, int64_t np, const uint32_t* ij, const double* blocks,
                  const double* b, double* x, double tolerance) {
  const System s = system_of(n, nullptr, nullptr, diag, np, ij, blocks);
  std::vector<V3> bv(n), xv(n);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      bv[i][k] = b[3 * i + k];
      xv[i][k] = x[3 * i + k];
    }
  const int64_t rounds = static_cast<Solver*>(h)->solve(s, bv, xv, tolerance);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) x[3 * i + k] = xv[i][k];
  return rounds;
}

// The preconditioner's output for one residual (solve.rs:472-497), for the tests' checks.
void mor_precondition(void* h, int64_t n, const double* diag, int64_t np, const uint32_t* ij, const double* blocks,
                      const double* b, double* out) {
  const System s = system_of(n, nullptr, nullptr, diag, np, ij, blocks);
  std::vector<V3> bv(n), z(n), scratch(n);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) bv[i][k] = b[3 * i + k];
  static_cast<Solver*>(h)->precondition(s, bv, z, scratch);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) out[3 * i + k] = z[i][k];
}

// System::times (solve.rs:61-72).
void mor_times(int64_t n, const double* diag, int64_t np, const uint32_t* ij, const double* blocks, const double* x,
               double* y) {
  const System s = system_of(n, nullptr, nullptr, diag, np, ij, blocks);
  std::vector<V3> xv(n), yv(n);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) xv[i][k] = x[3 * i + k];
  s.apply(xv, yv);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) y[3 * i + k] = yv[i][k];
}

// The levels: count; per level (up to cap) its parts, pairs, step, whether it is solved
// directly; step0 = the finest level's step.
int mor_levels(void* h, int cap, int64_t* parts, int64_t* npairs, double* steps, int32_t* direct, double* step0) {
  const Solver* s = static_cast<Solver*>(h);
  if (step0) *step0 = s->step;
  for (int l = 0; l < (int)s->levels.size() && l < cap; l++) {
    parts[l] = (int64_t)s->levels[l]->centre.size();
    npairs[l] = (int64_t)s->levels[l]->pairs.size();
    steps[l] = s->levels[l]->step;
    direct[l] = s->levels[l]->direct ? 1 : 0;
  }
  return (int)s->levels.size();
}

// Level l's part of each finer part (solve.rs:190-207 `of`), and its offset r from the middle.
void mor_level_map(void* h, int l, uint32_t* of, double* r) {
  const Coarse& c = *static_cast<Solver*>(h)->levels[l];
  for (size_t i = 0; i < c.of.size(); i++) {
    of[i] = c.of[i];
    for (int k = 0; k < 3; k++) r[3 * i + k] = c.r[i][k];
  }
}

// The 6x6 lift T(r)^T M T(s) (solve.rs:76-98) and the inverses (solve.rs:129-139), for tests.
void mor_lift(const double* m, const double* r, const double* s, double* out) {
  M6 a;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) a[i][j] = m[6 * i + j];
  const M6 o = lift(a, V3{r[0], r[1], r[2]}, V3{s[0], s[1], s[2]});
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) out[6 * i + j] = o[i][j];
}
void mor_inverse6(const double* m, double* out) {
  M6 a;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) a[i][j] = m[6 * i + j];
  const M6 o = inverse<6>(a);
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) out[6 * i + j] = o[i][j];
}

// The unit cube's stiffness (element.rs:49-76): k[a][c] 3x3, 8 x 8 blocks.
void mor_unit_k(double* k) { unit_k(k); }

}  // extern "C"
