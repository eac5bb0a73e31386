// The next level up, measured from a finer one (statics/src/gather.rs:193-395): each
// coarse cell's matter condensed onto its faces, and the beams nearest that stiffness.
//
// The host does the bookkeeping the reference does in `measured` (which coarse cell each
// finer node and bond lies in, the faces, the handles, the materials, the weights); the
// device measures the cells, one wave per cell: the elimination of the free parts
// (responses, :59-114), the stiffness between handles, the non-negative beam fit (beams,
// :119-161) and each beam's share of the finer strength. Every choice in that measure (a
// pivot, a beam dropped for being negative, a beam kept for being positive) follows from
// FP64 values compared exactly, and full cells are symmetric, so ties are the rule, not
// the exception: each value is computed with the reference's operations in the reference's
// order, without contraction (the pragma below), so the choices, and the lattice, are the
// reference's. A wave measures a cell; its lanes split only work whose values do not depend
// on how it is split.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

#include "st_solver.hpp"

#pragma clang fp contract(off)

namespace st {
namespace {

constexpr int kMaxFree = 12;   // finer nodes inside a coarse cell: its 12 inner faces
constexpr int kMaxHandles = 6;  // a coarse cell's faces
constexpr int kMaxBonds = 480;  // 8 finer cells x 15 beams x PER_BEAM
constexpr int kMaxPairs = kMaxHandles * (kMaxHandles - 1) / 2;
constexpr int kRow = 6 * kMaxFree + 3 * kMaxHandles;  // the elimination's row pitch (90)
constexpr int kN2 = 3 * kMaxHandles;                   // handle motions (18)
constexpr double kSection[2] = {0.336, 0.336};         // lattice.rs:25
constexpr double kUnit = 0.125;                        // lattice.rs:16
constexpr int kPerBeam = 4;                            // lattice.rs:19

struct G3 {
  double c[3];
};
__host__ __device__ inline G3 gadd(G3 a, G3 b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
__host__ __device__ inline G3 gsub(G3 a, G3 b) { return {{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}}; }
__host__ __device__ inline G3 gmul(G3 a, double s) { return {{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
__host__ __device__ inline double gdot(G3 a, G3 b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; }
__host__ __device__ inline G3 gcross(G3 a, G3 b) {
  return {{a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2], a.c[0] * b.c[1] - a.c[1] * b.c[0]}};
}
__host__ __device__ inline double gnorm(G3 a) { return sqrt(gdot(a, a)); }
__host__ __device__ inline G3 gunit(G3 a) { return gmul(a, 1.0 / gnorm(a)); }
// std::max's choice (the first unless the second is greater), as the reference's f64::max
// gives for the values it meets here
__host__ __device__ inline double gmax(double a, double b) { return a < b ? b : a; }
// lattice.rs:29-32
__host__ __device__ inline int gclass(G3 a, G3 b, G3 c) {
  G3 u = gsub(a, c), v = gsub(b, c);
  return (int)round(1.0 + gdot(u, v) / (gnorm(u) * gnorm(v)));
}
// Mat::used (model.rs:66-68) for strengths ft, fc, fs
__host__ __device__ inline double mat_used(double ft, double fc, double fs, double s1, double s3) {
  return gmax(gmax(s1 / ft, -s3 / fc), (s1 - s3) / (2.0 * fs));
}

// The cells to measure, structure of arrays; a cell's bonds, free parts and handles are
// contiguous ranges.
struct GJob {
  int32_t n;
  const int32_t *nf, *nh, *nb;
  const int64_t *boff, *foff, *hoff;
  const double* fx;       // free parts' positions, 3 each
  const double* hx;       // handles' positions, 3 each
  const int32_t* piece;   // per handle, the piece of matter it is joined to
  const int32_t* bij;     // per bond of a cell, its two places (free parts first, then handles)
  const int64_t* bidx;    // per bond of a cell, the finer lattice's bond
  const double *bp, *bn, *ba, *bd, *bm;  // the finer lattice's bonds: p, n (3 each), a, d, material (6)
  const double* cmat;     // per cell: e, ft, fc, fs of its material
  const double* cfull;    // per cell: section and strength ratio of both classes
  const double* ccentre;  // per cell: its centre
  double s;               // the coarse spacing
  int32_t against_full;   // sections against a full cell's (gathered) or as measured
  int32_t* nfound;        // per cell: how many beams
  int32_t* fab;           // per cell, kMaxPairs x 2: the beams' handles
  double* fval;           // per cell, kMaxPairs x 2: stiffness along, how much stronger
};

// the solution's row r, right-hand side k
__device__ inline double sol_x(const double* A, int pitch, int n, int r, int k, double tiny) {
  const double d = A[r * pitch + r];
  return fabs(d) <= tiny ? 0.0 : A[r * pitch + n + k] / d;
}

// gather.rs:165-186 by one wave (every lane calls it): each step's pivot found over the
// lanes (the last of the greatest, as max_by), the other rows updated one lane each, eight
// entries at a time, each entry as the reference updates it. Returns the scale's tiny.
__device__ double wave_solve(double* M, int pitch, int n, int rhs, int lane) {
  double scale = 0.0;
  for (int r = 0; r < n; r++) scale = gmax(scale, fabs(M[r * pitch + r]));
  const double tiny = 1e-12 * scale;
  const int W = n + rhs;
  for (int cc = 0; cc < n; cc++) {
    double bv = 0.0;
    int bi = -1;
    for (int x = cc + lane; x < n; x += 64) {
      const double v = fabs(M[x * pitch + cc]);
      if (bi < 0 || !(v < bv)) bv = v, bi = x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o);
      const int oi = __shfl_xor(bi, o);
      if (oi >= 0 && (bi < 0 || bv < ov || (!(ov < bv) && oi > bi))) bv = ov, bi = oi;
    }
    const int p = __shfl(bi, 0);  // every lane holds the same; taken from one so the wave agrees
    if (p != cc) {
      for (int k = lane; k < W; k += 64) {
        const double t = M[cc * pitch + k];
        M[cc * pitch + k] = M[p * pitch + k];
        M[p * pitch + k] = t;
      }
      __syncthreads();
    }
    const double pv = M[cc * pitch + cc];
    if (fabs(pv) <= tiny) continue;
    for (int r = lane; r < n; r += 64) {
      if (r == cc) continue;
      const double f = M[r * pitch + cc] / pv;
      if (f == 0.0) continue;
      int k = cc;
      for (; k + 8 <= W; k += 8) {
        double u[8], v[8];
#pragma unroll
        for (int t = 0; t < 8; t++) u[t] = M[cc * pitch + k + t], v[t] = M[r * pitch + k + t];
#pragma unroll
        for (int t = 0; t < 8; t++) M[r * pitch + k + t] = v[t] - f * u[t];
      }
      for (; k < W; k++) M[r * pitch + k] -= f * M[cc * pitch + k];
    }
    __syncthreads();
  }
  return tiny;
}

constexpr int kRowL = kRow + 1;  // LDS row pitch of the elimination (odd: a column spreads over the banks)
constexpr int64_t kCaseDoubles = (int64_t)kN2 * kMaxBonds * 3;  // global workspace per cell: the cases

// One wave measures one coarse cell. Lanes share the work only where each value's operations
// and their order stay those of the reference: the rows of the assembly and of an
// elimination step, the handle motions, the bonds of a pulled beam; the sums over bonds run
// in bond order, each on one lane.
__global__ __launch_bounds__(64) void k_measure_cells(GJob j, double* ws) {
  const int c = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ double A[6 * kMaxFree * kRowL];
  __shared__ double X[6 * kMaxFree * kN2];
  __shared__ double Ks[kN2 * kN2];
  __shared__ double Gs[kMaxPairs * (kMaxPairs + 1)];
  __shared__ double fx_s[kMaxPairs];
  __shared__ int fa_s[kMaxPairs], fb_s[kMaxPairs];
  __shared__ int nfound_s;
  double* cases = ws + (int64_t)c * kCaseDoubles;
  const int nf = j.nf[c], nh = j.nh[c], nb = j.nb[c];
  const int n = 6 * nf, m = 3 * nh, W = n + m;
  const int32_t* piece = j.piece + j.hoff[c];
  const double* fx = j.fx + 3 * j.foff[c];
  const double* hx = j.hx + 3 * j.hoff[c];
  const int32_t* bij = j.bij + 2 * j.boff[c];
  const int64_t* bidx = j.bidx + j.boff[c];
  auto BP = [&](int b) -> G3 {
    const double* q = j.bp + 3 * bidx[b];
    return {{q[0], q[1], q[2]}};
  };
  auto BN = [&](int b) -> G3 {
    const double* q = j.bn + 3 * bidx[b];
    return {{q[0], q[1], q[2]}};
  };
  auto BA = [&](int b) { return j.ba[bidx[b]]; };
  auto BD = [&](int b) { return j.bd[bidx[b]]; };
  auto BM = [&](int b, int k) { return j.bm[6 * bidx[b] + k]; };  // rho, e, g, ft, fc, fs
  auto pos = [&](int v) -> G3 {
    const double* q = v < nf ? fx + 3 * v : hx + 3 * (v - nf);
    return {{q[0], q[1], q[2]}};
  };
  auto dof = [&](int v, int d) -> int { return v < nf ? 6 * v + d : (d < 3 ? n + 3 * (v - nf) + d : -1); };
  auto column = [&](int v, int d, G3 p) -> G3 {
    G3 e{{0.0, 0.0, 0.0}};
    e.c[d % 3] = 1.0;
    return d < 3 ? e : gcross(e, gsub(p, pos(v)));
  };
  auto CASE = [&](int cs, int b, int q) -> double& { return cases[((int64_t)cs * kMaxBonds + b) * 3 + q]; };

  // responses (gather.rs:59-103): the free parts' equations, the handles' columns on the
  // right. A lane owns whole rows and walks the bonds in order, adding each bond's twelve
  // entries of its row: every entry takes its bonds in the reference's order, no barrier.
  for (int t = lane; t < n * kRowL; t += 64) A[t] = 0.0;
  __syncthreads();  // rows are zeroed by other lanes than their owners
  for (int r = lane; r < n; r += 64) {
    const int v = r / 6, d = r % 6;
    for (int b = 0; b < nb; b++) {
      const int vs[2] = {bij[2 * b], bij[2 * b + 1]};
      const int x = vs[0] == v ? 0 : (vs[1] == v ? 1 : -1);
      if (x < 0) continue;
      const G3 p = BP(b), nv = BN(b);
      const double kk = BA(b) / BD(b);
      const double e = BM(b, 1), g = BM(b, 2);
      const double sg[2] = {-1.0, 1.0};
      double st[3][3];  // Bond::stiffness (model.rs:139-149)
      for (int a2 = 0; a2 < 3; a2++)
        for (int q = 0; q < 3; q++) {
          const double nn = nv.c[a2] * nv.c[q];
          st[a2][q] = e * kk * nn + g * kk * ((a2 == q ? 1.0 : 0.0) - nn);
        }
      const G3 cv = column(v, d, p);
      const G3 scv{{gdot(G3{{st[0][0], st[0][1], st[0][2]}}, cv), gdot(G3{{st[1][0], st[1][1], st[1][2]}}, cv),
                    gdot(G3{{st[2][0], st[2][1], st[2][2]}}, cv)}};
      for (int y = 0; y < 2; y++)
        for (int d2 = 0; d2 < 6; d2++) {
          const int col = dof(vs[y], d2);
          if (col < 0) continue;
          const double sign = col < n ? 1.0 : -1.0;
          A[r * kRowL + col] += sign * sg[x] * sg[y] * gdot(scv, column(vs[y], d2, p));
        }
    }
  }
  __syncthreads();

  // gather.rs:165-186, then the solutions x (the free parts' motions for each handle motion)
  const double tiny_a = wave_solve(A, kRowL, n, m, lane);
  for (int t = lane; t < n * m; t += 64) X[t] = sol_x(A, kRowL, n, t / m, t % m, tiny_a);
  __syncthreads();

  // each handle moved alone along each axis (one lane each): the bonds' forces (moved,
  // bodies.rs:420-423), and what holds each handle (gather.rs:291-303), bond after bond
  for (int t = lane; t < kN2 * kN2; t += 64) Ks[t] = 0.0;
  __syncthreads();
  if (lane < m) {
    const int cs = lane;
    for (int b = 0; b < nb; b++) {
      const G3 p = BP(b), nv = BN(b);
      G3 tr[2];
      for (int x = 0; x < 2; x++) {
        const int v = bij[2 * b + x];
        double mo[6];
        if (v < nf) {
          for (int d = 0; d < 6; d++) mo[d] = X[(6 * v + d) * m + cs];
        } else {
          for (int d = 0; d < 6; d++) mo[d] = 0.0;
          if (v - nf == cs / 3) mo[cs % 3] = 1.0;
        }
        // shift (bodies.rs:44-47): the translation at the bond's point
        tr[x] = gadd(G3{{mo[0], mo[1], mo[2]}}, gcross(G3{{mo[3], mo[4], mo[5]}}, gsub(p, pos(v))));
      }
      const G3 delta = gsub(tr[1], tr[0]);
      // Bond::force (model.rs:115-119)
      const double along = gdot(delta, nv);
      const double kk = BA(b) / BD(b);
      const G3 f = gadd(gmul(nv, BM(b, 1) * kk * along), gmul(gsub(delta, gmul(nv, along)), BM(b, 2) * kk));
      for (int q = 0; q < 3; q++) CASE(cs, b, q) = f.c[q];
      const int vj = bij[2 * b + 1], vi = bij[2 * b];
      if (vj >= nf)
        for (int q = 0; q < 3; q++) Ks[(3 * (vj - nf) + q) * kN2 + cs] += 1.0 * f.c[q];
      if (vi >= nf)
        for (int q = 0; q < 3; q++) Ks[(3 * (vi - nf) + q) * kN2 + cs] += -1.0 * f.c[q];
    }
  }
  __syncthreads();

  // beams (gather.rs:119-161): the normal equations' entries one lane each, their
  // elimination over the wave; every lane keeps the same active set
  __shared__ int pa[kMaxPairs], pb[kMaxPairs];
  __shared__ double ue[3 * kMaxPairs];
  int np = 0;
  for (int a = 0; a < nh; a++)
    for (int b = a + 1; b < nh; b++)
      if (piece[a] == piece[b]) {
        if (lane == 0) pa[np] = a, pb[np] = b;
        np++;
      }
  __syncthreads();
  for (int p = lane; p < np; p += 64) {
    const G3 u = gunit(gsub(pos(nf + pb[p]), pos(nf + pa[p])));
    for (int k = 0; k < 3; k++) ue[3 * p + k] = u.c[k];
  }
  __syncthreads();
  // A unit (a beam of unit stiffness along its direction, as a matrix) is nonzero only in the
  // four 3x3 blocks of its two handles: +e e^T on the diagonal blocks, -e e^T off them. Its
  // products with another matrix are summed over those 36 entries in increasing index; the
  // reference sums over all (3 nh)^2, but every other term is a signed zero, which leaves a
  // sum begun at +0 unchanged, so the sums are the same to the bit.
  auto U = [&](int p, int row, int col) -> double {
    const int hr = row / 3, hc = col / 3;
    const bool ra = hr == pa[p], rb = hr == pb[p], ca = hc == pa[p], cb = hc == pb[p];
    if (!((ra || rb) && (ca || cb))) return 0.0;
    const double v = ue[3 * p + row % 3] * ue[3 * p + col % 3];
    return (ra == ca) ? 0.0 + v : 0.0 - v;
  };
  auto T = [&](int row, int col) { return 0.5 * (Ks[row * kN2 + col] + Ks[col * kN2 + row]); };
  auto udot = [&](int p, int q) {
    const int hs[2] = {pa[p], pb[p]};
    double sum = 0.0;
    for (int x = 0; x < 2; x++)
      for (int r = 0; r < 3; r++) {
        const int row = 3 * hs[x] + r;
        for (int y = 0; y < 2; y++)
          for (int c2 = 0; c2 < 3; c2++) {
            const int col = 3 * hs[y] + c2;
            sum += U(p, row, col) * (q >= 0 ? U(q, row, col) : T(row, col));
          }
      }
    return sum;
  };
  const int gp = kMaxPairs + 1;
  uint32_t on = np ? ((1u << np) - 1u) : 0u;
  int idx[kMaxPairs], mm = 0;
  double xs[kMaxPairs];
  for (;;) {
    mm = 0;
    for (int p = 0; p < np; p++)
      if (on >> p & 1u) idx[mm++] = p;
    for (int e = lane; e < mm * (mm + 1); e += 64) {
      const int r = e / (mm + 1), q = e % (mm + 1);
      Gs[r * gp + q] = q < mm ? udot(idx[r], idx[q]) : udot(idx[r], -1);
    }
    __syncthreads();
    const double tiny_g = wave_solve(Gs, gp, mm, 1, lane);
    int worst = -1;
    for (int r = 0; r < mm; r++) {
      xs[r] = sol_x(Gs, gp, mm, r, 0, tiny_g);
      if (xs[r] < 0.0 && (worst < 0 || xs[r] < xs[worst])) worst = r;  // the first most negative
    }
    __syncthreads();  // every lane has read the solutions before the next round writes
    if (worst < 0) break;
    on &= ~(1u << idx[worst]);
  }
  if (lane == 0) {
    int found = 0;
    for (int r = 0; r < mm; r++)
      if (xs[r] > 0.0) {
        fa_s[found] = pa[idx[r]];
        fb_s[found] = pb[idx[r]];
        fx_s[found] = xs[r];
        found++;
      }
    nfound_s = found;
  }
  __syncthreads();

  // each beam kept: pulled alone, a unit apart (lanes on bonds), how much of the finer
  // strength it uses (summed in bond order on one lane)
  const double e_c = j.cmat[4 * c], ft_c = j.cmat[4 * c + 1], fc_c = j.cmat[4 * c + 2], fs_c = j.cmat[4 * c + 3];
  const G3 centre{{j.ccentre[3 * c], j.ccentre[3 * c + 1], j.ccentre[3 * c + 2]}};
  const double s = j.s;
  double* wu = A;  // the elimination is done with: the bonds' work and its used share
  const int nfd = nfound_s;
  for (int k = 0; k < nfd; k++) {
    const int a = fa_s[k], b = fb_s[k];
    const G3 at_a = pos(nf + a), at_b = pos(nf + b);
    const G3 e = gunit(gsub(at_b, at_a));
    for (int q = lane; q < nb; q += 64) {
      G3 f{{0.0, 0.0, 0.0}};
      for (int t = 0; t < 3; t++) {
        const G3 cb{{CASE(3 * b + t, q, 0), CASE(3 * b + t, q, 1), CASE(3 * b + t, q, 2)}};
        const G3 ca{{CASE(3 * a + t, q, 0), CASE(3 * a + t, q, 1), CASE(3 * a + t, q, 2)}};
        f = gadd(f, gadd(gmul(cb, e.c[t] / 2.0), gmul(ca, -e.c[t] / 2.0)));
      }
      const G3 nv = BN(q);
      const double ba = BA(q), bd = BD(q);
      // Bond::used (model.rs:123-129)
      const G3 tq = gmul(f, 1.0 / ba);
      const double sigma = gdot(tq, nv);
      const double tau2 = gmax(gdot(tq, tq) - sigma * sigma, 0.0);
      const double rr = sqrt(sigma * sigma / 4.0 + tau2);
      const double u = mat_used(BM(q, 3), BM(q, 4), BM(q, 5), sigma / 2.0 + rr, sigma / 2.0 - rr);
      // Bond::energy (model.rs:132-136)
      const double kk = ba / bd;
      const double along = gdot(f, nv);
      const double wq = 0.5 * (along * along / (BM(q, 1) * kk) + (gdot(f, f) - along * along) / (BM(q, 2) * kk));
      wu[2 * q] = wq * u;
      wu[2 * q + 1] = wq;
    }
    __syncthreads();
    if (lane == 0) {
      double held = 0.0, work = 0.0;
      for (int q = 0; q < nb; q++) {
        held += wu[2 * q];
        work += wu[2 * q + 1];
      }
      const double fine_used = work > 0.0 ? held / work : 0.0;
      const double length = gnorm(gsub(at_b, at_a));
      const double own = mat_used(ft_c, fc_c, fs_c, e_c / length, 0.0);
      const double stronger = fine_used > 0.0 ? own / fine_used : 1.0;
      const int kc = gclass(at_a, at_b, centre);
      double section = fx_s[k] * length / e_c / (s * s);
      if (j.against_full) section = section * kSection[kc] / j.cfull[4 * c + 2 * kc];
      j.fab[(int64_t)c * 2 * kMaxPairs + 2 * k] = a;
      j.fab[(int64_t)c * 2 * kMaxPairs + 2 * k + 1] = b;
      j.fval[(int64_t)c * 2 * kMaxPairs + 2 * k] = section * s * s * e_c / length;
      j.fval[(int64_t)c * 2 * kMaxPairs + 2 * k + 1] = stronger / j.cfull[4 * c + 2 * kc + 1];
    }
    __syncthreads();
  }
  if (lane == 0) j.nfound[c] = nfd;
}

// ------------------------------------------------------------------ host side

using Mat6 = std::array<double, 6>;  // rho, e, g, ft, fc, fs
using Site3 = std::array<int64_t, 3>;
struct Site3Hash {
  size_t operator()(const Site3& s) const {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (int64_t v : s) h = (h ^ (uint64_t)v) * 0x100000001b3ull + (h >> 29);
    return (size_t)h;
  }
};

// A lattice in host memory (lattice.rs:61-71)
struct HLat {
  uint32_t level = 0;
  std::vector<G3> x;
  std::vector<double> m;
  std::vector<Mat6> nmat;
  std::vector<uint8_t> held;
  std::vector<std::array<Site3, 2>> around;
  std::vector<std::array<int64_t, 2>> ij;
  std::vector<G3> p, n;
  std::vector<double> a, d;
  std::vector<Mat6> bmat;
};

double hspacing(uint32_t level) { return kUnit * (double)(uint64_t(1) << level); }  // lattice.rs:34-36
G3 hcentre(const Site3& site, uint32_t level) {                                      // lattice.rs:39-41
  G3 c;
  for (int k = 0; k < 3; k++) c.c[k] = ((double)site[k] + 0.5 * (double)(uint64_t(1) << level)) * kUnit;
  return c;
}

// lattice.rs:76-84: PER_BEAM bonds about the beam's middle
void hbeam(HLat& L, int64_t i, int64_t j, G3 a, G3 b, double area, const Mat6& mat) {
  const G3 n = gunit(gsub(b, a));
  const G3 e1 = gunit(gcross(n, fabs(n.c[0]) < 0.9 ? G3{{1.0, 0.0, 0.0}} : G3{{0.0, 1.0, 0.0}}));
  const G3 e2 = gcross(n, e1);
  const double r = sqrt(area / (2.0 * M_PI));
  const G3 middle = gmul(gadd(a, b), 0.5);
  const double d = gnorm(gsub(b, a));
  const G3 off[kPerBeam] = {gmul(e1, r), gmul(e1, -r), gmul(e2, r), gmul(e2, -r)};
  for (const G3& o : off) {
    L.ij.push_back({i, j});
    L.p.push_back(gadd(middle, o));
    L.n.push_back(n);
    L.a.push_back(area / 4.0);
    L.d.push_back(d);
    L.bmat.push_back(mat);
  }
}

// A finer lattice as flat arrays (StLattice's layout), read in place
struct LatIn {
  uint32_t level;
  int64_t nn, nb;
  const double *x, *m, *nmat;
  const uint8_t* held;
  const int64_t *around, *ij;
  const double *p, *n, *a, *d, *bmat;
};
LatIn view_of(const HLat& L) {
  return {L.level, (int64_t)L.m.size(), (int64_t)L.a.size(), reinterpret_cast<const double*>(L.x.data()),
          L.m.data(), reinterpret_cast<const double*>(L.nmat.data()), L.held.data(),
          reinterpret_cast<const int64_t*>(L.around.data()), reinterpret_cast<const int64_t*>(L.ij.data()),
          reinterpret_cast<const double*>(L.p.data()), reinterpret_cast<const double*>(L.n.data()), L.a.data(),
          L.d.data(), reinterpret_cast<const double*>(L.bmat.data())};
}

void measure(Ctx* c, const LatIn& fine, bool against_full, HLat& out);

// f(begin, end) over [0, n) in contiguous ranges on up to 16 host threads (pure arithmetic only)
template <class F>
void par_for(int64_t n, F f) {
  const int64_t t = std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 65536 || t == 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve((size_t)t);
  const int64_t chunk = (n + t - 1) / t;
  int64_t a = 0;
  try {
    for (; a < n; a += chunk) pool.emplace_back(f, a, std::min(n, a + chunk));
  } catch (const std::exception&) {
    // a thread that could not start: its range, and the rest, run here
  }
  for (; a < n; a += chunk) f(a, std::min(n, a + chunk));
  for (auto& th : pool) th.join();
}

// lattice.rs:76-84 into bonds [q, q + PER_BEAM) of L
void hbeam_at(HLat& L, int64_t q, int64_t i, int64_t j, G3 a, G3 b, double area, const Mat6& mat) {
  const G3 n = gunit(gsub(b, a));
  const G3 e1 = gunit(gcross(n, fabs(n.c[0]) < 0.9 ? G3{{1.0, 0.0, 0.0}} : G3{{0.0, 1.0, 0.0}}));
  const G3 e2 = gcross(n, e1);
  const double r = sqrt(area / (2.0 * M_PI));
  const G3 middle = gmul(gadd(a, b), 0.5);
  const double d = gnorm(gsub(b, a));
  const G3 off[kPerBeam] = {gmul(e1, r), gmul(e1, -r), gmul(e2, r), gmul(e2, -r)};
  for (int k = 0; k < kPerBeam; k++) {
    L.ij[q + k] = {i, j};
    L.p[q + k] = gadd(middle, off[k]);
    L.n[q + k] = n;
    L.a[q + k] = area / 4.0;
    L.d[q + k] = d;
    L.bmat[q + k] = mat;
  }
}

// gather.rs:200-223: a full cell's beams by class, measured once per material. The
// reference measures the cell at the origin of a 6^3 block of level-0 cells; the cell's
// own problem (its 8 finer cells, their nodes and bonds, in the same order, none held) is
// the same in the 2^3 block of those 8 cells alone, and so are its beams, which are all
// the block's beams with both ends on the origin cell's faces.
std::array<double, 4> full_of(Ctx* c, const Mat6& mat) {
  for (auto& k : c->full_known)
    if (std::equal(mat.begin(), mat.end(), k.begin())) return {k[6], k[7], k[8], k[9]};
  HLat block;  // build(0, the 8 cells) (lattice.rs:124-163; weights left out: unused)
  const int faces[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  std::map<std::array<Site3, 2>, int64_t> index;
  for (int64_t ax = 0; ax < 2; ax++)
    for (int64_t ay = 0; ay < 2; ay++)
      for (int64_t az = 0; az < 2; az++) {
        const Site3 site{ax, ay, az};
        int64_t ends[6];
        for (int f = 0; f < 6; f++) {
          const Site3 other{site[0] + faces[f][0], site[1] + faces[f][1], site[2] + faces[f][2]};
          const std::array<Site3, 2> key = site < other ? std::array<Site3, 2>{site, other}
                                                        : std::array<Site3, 2>{other, site};
          auto it = index.find(key);
          if (it == index.end()) {
            it = index.emplace(key, (int64_t)block.x.size()).first;
            block.x.push_back(gmul(gadd(hcentre(site, 0), hcentre(other, 0)), 0.5));
            block.m.push_back(0.0);
            block.nmat.push_back(mat);
            block.held.push_back(0);
            block.around.push_back(key);
          }
          ends[f] = it->second;
        }
        const G3 middle = hcentre(site, 0);
        const double s0 = hspacing(0);
        for (int a = 0; a < 6; a++)
          for (int b = a + 1; b < 6; b++) {
            const G3 x = block.x[ends[a]], y = block.x[ends[b]];
            hbeam(block, ends[a], ends[b], x, y, kSection[gclass(x, y, middle)] * s0 * s0, mat);
          }
      }
  HLat coarse;
  measure(c, view_of(block), false, coarse);
  const double s = hspacing(1);
  const Site3 zero{0, 0, 0};
  double sum[2][3] = {{0, 0, 0}, {0, 0, 0}};
  for (size_t q = 0; q < coarse.a.size(); q += kPerBeam) {
    const int64_t i = coarse.ij[q][0], j = coarse.ij[q][1];
    auto on = [&](int64_t v) { return coarse.around[v][0] == zero || coarse.around[v][1] == zero; };
    if (on(i) && on(j)) {
      const int k = gclass(coarse.x[i], coarse.x[j], hcentre(zero, 1));
      sum[k][0] += 4.0 * coarse.a[q] / (s * s);
      sum[k][1] += coarse.bmat[q][3] / mat[3];
      sum[k][2] += 1.0;
    }
  }
  std::array<double, 4> f;
  for (int k = 0; k < 2; k++) {
    f[2 * k] = sum[k][2] > 0.0 ? sum[k][0] / sum[k][2] : 1.0;
    f[2 * k + 1] = sum[k][2] > 0.0 ? sum[k][1] / sum[k][2] : 1.0;
  }
  std::array<double, 10> kept;
  std::copy(mat.begin(), mat.end(), kept.begin());
  std::copy(f.begin(), f.end(), kept.begin() + 6);
  c->full_known.push_back(kept);
  return f;
}

template <class T>
void put(DBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.ensure(std::max<size_t>(h.size(), 1) * sizeof(T));
  if (!h.empty()) ST_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// gather.rs:228-395
void measure(Ctx* ctx, const LatIn& fine, bool against_full, HLat& out) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  const uint32_t level = fine.level + 1;
  const int64_t nn = fine.nn, nbond = fine.nb;
  // The caller's bond arrays go to the device on a host thread while the bookkeeping runs
  // (only the kernel reads them). The guard joins it however this call ends.
  DBuf dbp, dbn, dba, dbd, dbm;
  hipError_t up_err = hipSuccess;
  std::thread uploader;
  struct Join {
    std::thread& t;
    ~Join() {
      if (t.joinable()) t.join();
    }
  } join{uploader};
  if (nbond > 0) {
    ST_HIP(hipSetDevice(ctx->device));
    const size_t nb8 = (size_t)nbond * sizeof(double);
    dbp.ensure(3 * nb8), dbn.ensure(3 * nb8), dba.ensure(nb8), dbd.ensure(nb8), dbm.ensure(6 * nb8);
    auto upload = [&, nb8] {
      hipError_t e = hipSetDevice(ctx->device);
      const std::pair<DBuf*, std::pair<const double*, size_t>> parts[5] = {
          {&dbp, {fine.p, 3 * nb8}}, {&dbn, {fine.n, 3 * nb8}}, {&dba, {fine.a, nb8}}, {&dbd, {fine.d, nb8}},
          {&dbm, {fine.bmat, 6 * nb8}}};
      for (auto& pt : parts)
        if (e == hipSuccess) e = hipMemcpy(pt.first->p, pt.second.first, pt.second.second, hipMemcpyHostToDevice);
      // a pageable host-to-device hipMemcpy may return before its DMA lands; the kernel
      // reading these arrays runs on another stream, so wait for the device here
      if (e == hipSuccess) e = hipDeviceSynchronize();
      up_err = e;
    };
    try {
      uploader = std::thread(upload);
    } catch (const std::exception&) {
      upload();  // no thread to be had: upload here
    }
  }
  auto site_of = [&](int64_t v, int k) {
    const int64_t* q = fine.around + 6 * v + 3 * k;
    return Site3{q[0], q[1], q[2]};
  };
  // the coarse cells, numbered as first met, each finer node's two in turn
  // (an open-addressing table of cell numbers: slots at least twice the nodes, linear probing)
  size_t cap = 64;
  while (cap < 2 * (size_t)nn + 16) cap <<= 1;
  std::vector<int64_t> table(cap, -1);
  const Site3Hash hasher;
  std::vector<Site3> sites;
  std::vector<std::array<int64_t, 2>> between((size_t)nn);
  for (int64_t v = 0; v < nn; v++)
    for (int k = 0; k < 2; k++) {
      Site3 key = site_of(v, k);
      for (int t = 0; t < 3; t++) key[t] = (key[t] >> level) << level;
      size_t h = hasher(key) & (cap - 1);
      while (table[h] >= 0 && sites[table[h]] != key) h = (h + 1) & (cap - 1);
      if (table[h] < 0) {
        table[h] = (int64_t)sites.size();
        sites.push_back(key);
      }
      between[v][k] = table[h];
    }
  table = std::vector<int64_t>();
  // each bond's cell: that of the site its two nodes share (the first of node i's two that
  // node j also lies by), already numbered through node i
  const int64_t ncell = (int64_t)sites.size();
  std::vector<int64_t> lies((size_t)nbond), bcount((size_t)ncell + 1, 0);
  std::atomic<bool> apart{false};
  par_for(nbond, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; b++) {
      const int64_t i = fine.ij[2 * b], j = fine.ij[2 * b + 1];
      const Site3 i0 = site_of(i, 0), i1 = site_of(i, 1), j0 = site_of(j, 0), j1 = site_of(j, 1);
      const int k = (i0 == j0 || i0 == j1) ? 0 : 1;
      if (k == 1 && !(i1 == j0 || i1 == j1)) apart = true;
      lies[b] = between[i][k];
    }
  });
  if (apart) throw StError{ST_BAD_INPUT, "gather: a bond's nodes share no cell"};
  for (int64_t b = 0; b < nbond; b++) bcount[lies[b] + 1]++;
  for (int64_t c = 0; c < ncell; c++) bcount[c + 1] += bcount[c];
  auto face_key = [&](int64_t a, int64_t b) {
    return sites[a] < sites[b] ? std::array<int64_t, 2>{a, b} : std::array<int64_t, 2>{b, a};
  };
  struct Cell {
    std::vector<int64_t> free, across;
  };
  std::vector<Cell> cells((size_t)ncell);
  std::vector<std::vector<std::pair<Mat6, double>>> mats((size_t)ncell);
  auto add_mat = [](std::vector<std::pair<Mat6, double>>& to, const Mat6& m, double w) {
    for (auto& x : to)
      if (x.first == m) {
        x.second += w;
        return;
      }
    to.push_back({m, w});
  };
  auto nmat = [&](int64_t v) {
    Mat6 m;
    for (int k = 0; k < 6; k++) m[k] = fine.nmat[6 * v + k];
    return m;
  };
  // the faces: (key, finer node on it, or -1 for a face only named by a handle), to be
  // sorted by key as the reference's ordered map keeps them
  struct FaceRec {
    std::array<int64_t, 2> key;
    int64_t v;
  };
  std::vector<FaceRec> recs;
  for (int64_t v = 0; v < nn; v++) {
    const int64_t x = between[v][0], y = between[v][1];
    const Mat6 mv = nmat(v);
    add_mat(mats[x], mv, fine.m[v]);
    if (x != y) add_mat(mats[y], mv, fine.m[v]);
    if (x == y)
      cells[x].free.push_back(v);
    else
      recs.push_back({face_key(x, y), v});
  }
  std::vector<int64_t> place((size_t)nn, -1);
  for (auto& cell : cells)
    for (size_t k = 0; k < cell.free.size(); k++) place[cell.free[k]] = (int64_t)k;
  // each cell's bonds in order, as places (free parts first, then a handle per face, in
  // the order the cell's bonds first name them): cells over threads
  std::vector<int32_t> bij(2 * (size_t)nbond);
  std::vector<int64_t> bidx((size_t)nbond), bfill(bcount.begin(), bcount.end() - 1);
  for (int64_t b = 0; b < nbond; b++) bidx[bfill[lies[b]]++] = b;
  for (auto& cell : cells) cell.across.reserve(kMaxHandles);  // the threads below never allocate
  std::atomic<bool> too_many{false};
  par_for(ncell, [&](int64_t c0, int64_t c1) {
    for (int64_t c = c0; c < c1; c++) {
      Cell& cell = cells[c];
      auto at = [&](int64_t v) -> int64_t {
        const int64_t x = between[v][0], y = between[v][1];
        if (x == c && y == c) return place[v];
        const int64_t other = x == c ? y : x;
        size_t h = 0;
        while (h < cell.across.size() && cell.across[h] != other) h++;
        if (h == cell.across.size()) {
          if (h == (size_t)kMaxHandles) {  // a cell has at most six faces
            too_many = true;
            return 0;
          }
          cell.across.push_back(other);
        }
        return (int64_t)(cell.free.size() + h);
      };
      for (int64_t q = bcount[c]; q < bcount[c + 1]; q++) {
        const int64_t b = bidx[q];
        bij[2 * q] = (int32_t)at(fine.ij[2 * b]);
        bij[2 * q + 1] = (int32_t)at(fine.ij[2 * b + 1]);
      }
    }
  });
  if (too_many) throw StError{ST_BAD_INPUT, "gather: a coarse cell's bonds name more than six faces"};
  std::vector<uint8_t> held((size_t)ncell);
  for (int64_t c = 0; c < ncell; c++) {
    bool all = !cells[c].free.empty();
    for (int64_t v : cells[c].free) all = all && fine.held[v];
    held[c] = all;
    for (int64_t o : cells[c].across) recs.push_back({face_key(c, o), -1});
  }
  {  // stable sort by key: counting sort on the first cell, then each bucket by the second
    std::vector<int64_t> start((size_t)ncell + 1, 0);
    for (const FaceRec& r : recs) start[r.key[0] + 1]++;
    for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    std::vector<FaceRec> sorted(recs.size());
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (const FaceRec& r : recs) sorted[fill[r.key[0]]++] = r;
    for (int64_t c = 0; c < ncell; c++)  // a few each: insertion sort, which keeps equal keys in order
      for (int64_t i = start[c] + 1; i < start[c + 1]; i++) {
        const FaceRec r = sorted[i];
        int64_t j = i;
        for (; j > start[c] && sorted[j - 1].key[1] > r.key[1]; j--) sorted[j] = sorted[j - 1];
        sorted[j] = r;
      }
    recs.swap(sorted);
  }
  auto mat_of = [&](std::initializer_list<int64_t> cs) {
    std::vector<std::pair<Mat6, double>> all;
    for (int64_t c : cs)
      for (auto& mw : mats[c]) add_mat(all, mw.first, mw.second);
    if (all.empty()) return Mat6{2500.0, 30e9, 12e9, 5e6, 100e6, 10e6};  // STONE (model.rs:53)
    size_t best = 0;
    for (size_t k = 1; k < all.size(); k++)
      if (!(all[k].second < all[best].second)) best = k;  // the last of the heaviest, as max_by
    return all[best].first;
  };
  out = HLat();
  out.level = level;
  std::vector<std::array<int64_t, 2>> keys;
  for (size_t r = 0; r < recs.size();) {
    const auto key = recs[r].key;
    double m = 0.0;
    bool on_held = false;
    for (; r < recs.size() && recs[r].key == key; r++)
      if (recs[r].v >= 0) {
        m += fine.m[recs[r].v];
        on_held = on_held || fine.held[recs[r].v];
      }
    keys.push_back(key);
    out.x.push_back(gmul(gadd(hcentre(sites[key[0]], level), hcentre(sites[key[1]], level)), 0.5));
    out.m.push_back(m);
    out.nmat.push_back(mat_of({key[0], key[1]}));
    out.held.push_back(on_held || held[key[0]] || held[key[1]]);
    out.around.push_back({sites[key[0]], sites[key[1]]});
  }
  recs = std::vector<FaceRec>();
  const double s = hspacing(level);
  std::vector<std::vector<int64_t>> ends((size_t)ncell);
  std::vector<Mat6> mat((size_t)ncell);
  for (int64_t c = 0; c < ncell; c++) {
    for (int64_t o : cells[c].across) {
      const auto key = face_key(c, o);
      ends[c].push_back(std::lower_bound(keys.begin(), keys.end(), key) - keys.begin());
    }
    mat[c] = mat_of({c});
  }
  const auto t1 = clock::now();

  // the cells to measure
  std::vector<int64_t> job;
  for (int64_t c = 0; c < ncell; c++)
    if (!held[c] && !cells[c].across.empty()) job.push_back(c);
  const int64_t nj = (int64_t)job.size();
  std::vector<int32_t> nf(nj), nh(nj), nb(nj);
  std::vector<int64_t> boff(nj), foff(nj + 1, 0), hoff(nj + 1, 0);
  std::vector<double> cmat(4 * nj), cfull(4 * nj), ccentre(3 * nj);
  // sizes and offsets in order; the full-cell measures (each material's once, which may
  // measure a block of its own) in order too
  for (int64_t q = 0; q < nj; q++) {
    const int64_t c = job[q];
    const Cell& cell = cells[c];
    const int64_t cb = bcount[c + 1] - bcount[c];
    if (cell.free.size() > (size_t)kMaxFree || cell.across.size() > (size_t)kMaxHandles || cb > kMaxBonds)
      throw StError{ST_BAD_INPUT, "gather: a coarse cell holds more than 12 free nodes, 6 faces or 480 bonds"};
    nf[q] = (int32_t)cell.free.size();
    nh[q] = (int32_t)cell.across.size();
    nb[q] = (int32_t)cb;
    foff[q + 1] = foff[q] + nf[q];
    hoff[q + 1] = hoff[q] + nh[q];
    boff[q] = bcount[c];
    const std::array<double, 4> f = against_full ? full_of(ctx, mat[c]) : std::array<double, 4>{1.0, 1.0, 1.0, 1.0};
    for (int k = 0; k < 4; k++) cfull[4 * q + k] = f[k];
  }
  std::vector<double> fx(3 * (size_t)foff[nj]), hx(3 * (size_t)hoff[nj]);
  std::vector<int32_t> piece((size_t)hoff[nj]);
  par_for(nj, [&](int64_t q0, int64_t q1) {
    int64_t parent[kMaxFree + kMaxHandles];  // (the threads allocate nothing)
    for (int64_t q = q0; q < q1; q++) {
      const int64_t c = job[q];
      const Cell& cell = cells[c];
      for (size_t k = 0; k < cell.free.size(); k++)
        for (int t = 0; t < 3; t++) fx[3 * (foff[q] + (int64_t)k) + t] = fine.x[3 * cell.free[k] + t];
      for (size_t k = 0; k < ends[c].size(); k++)
        for (int t = 0; t < 3; t++) hx[3 * (hoff[q] + (int64_t)k) + t] = out.x[ends[c][k]].c[t];
      // which of its handles its matter joins (gather.rs:42-50): a piece of each
      for (size_t k = 0; k < cell.free.size() + cell.across.size(); k++) parent[k] = (int64_t)k;
      auto root = [&](int64_t a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
      };
      for (int64_t t = bcount[c]; t < bcount[c + 1]; t++) {
        const int64_t ra = root(bij[2 * t]);
        parent[ra] = root(bij[2 * t + 1]);
      }
      for (size_t h = 0; h < cell.across.size(); h++)
        piece[hoff[q] + (int64_t)h] = (int32_t)root((int64_t)(cell.free.size() + h));
      cmat[4 * q] = mat[c][1];
      for (int k = 0; k < 3; k++) cmat[4 * q + 1 + k] = mat[c][3 + k];
      const G3 ce = hcentre(sites[c], level);
      for (int k = 0; k < 3; k++) ccentre[3 * q + k] = ce.c[k];
    }
  });
  foff.pop_back();
  hoff.pop_back();
  const auto t2 = clock::now();
  std::vector<int32_t> nfound(nj), fab((size_t)nj * 2 * kMaxPairs);
  std::vector<double> fval((size_t)nj * 2 * kMaxPairs);
  if (nj) {
    ST_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DBuf dnf, dnh, dnb, dboff, dfoff, dhoff, dfx, dhx, dpiece, dbij, dbidx, dcmat, dcfull, dcc, dfound, dfab, dfval,
        dws;
    put(dnf, nf, st), put(dnh, nh, st), put(dnb, nb, st), put(dboff, boff, st), put(dfoff, foff, st);
    put(dhoff, hoff, st), put(dfx, fx, st), put(dhx, hx, st), put(dpiece, piece, st), put(dbij, bij, st);
    put(dbidx, bidx, st), put(dcmat, cmat, st), put(dcfull, cfull, st), put(dcc, ccentre, st);
    if (uploader.joinable()) uploader.join();
    if (up_err != hipSuccess)
      throw StError{up_err == hipErrorOutOfMemory ? ST_OOM : ST_DEVICE,
                    std::string("gather: bond upload: ") + hipGetErrorString(up_err)};
    dfound.ensure(nj * sizeof(int32_t));
    dfab.ensure(fab.size() * sizeof(int32_t));
    dfval.ensure(fval.size() * sizeof(double));
    // cells per launch (a wave each; the cases are 207 KB a cell in global memory)
    const int64_t chunk = std::min<int64_t>(nj, 16384);
    dws.ensure((size_t)chunk * kCaseDoubles * sizeof(double));
    for (int64_t c0 = 0; c0 < nj; c0 += chunk) {
      const int64_t cn = std::min(chunk, nj - c0);
      GJob j;
      j.n = (int32_t)cn;
      j.nf = dnf.as<int32_t>() + c0, j.nh = dnh.as<int32_t>() + c0, j.nb = dnb.as<int32_t>() + c0;
      j.boff = dboff.as<int64_t>() + c0, j.foff = dfoff.as<int64_t>() + c0, j.hoff = dhoff.as<int64_t>() + c0;
      j.fx = dfx.as<double>(), j.hx = dhx.as<double>(), j.piece = dpiece.as<int32_t>();
      j.bij = dbij.as<int32_t>(), j.bidx = dbidx.as<int64_t>();
      j.bp = dbp.as<double>(), j.bn = dbn.as<double>(), j.ba = dba.as<double>(), j.bd = dbd.as<double>();
      j.bm = dbm.as<double>();
      j.cmat = dcmat.as<double>() + 4 * c0, j.cfull = dcfull.as<double>() + 4 * c0, j.ccentre = dcc.as<double>() + 3 * c0;
      j.s = s;
      j.against_full = against_full ? 1 : 0;
      j.nfound = dfound.as<int32_t>() + c0;
      j.fab = dfab.as<int32_t>() + c0 * 2 * kMaxPairs;
      j.fval = dfval.as<double>() + c0 * 2 * kMaxPairs;
      k_measure_cells<<<(unsigned)cn, 64, 0, st>>>(j, dws.as<double>());
      ST_HIP(hipGetLastError());
    }
    ST_HIP(hipMemcpyAsync(nfound.data(), dfound.p, nj * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ST_HIP(hipMemcpyAsync(fab.data(), dfab.p, fab.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ST_HIP(hipMemcpyAsync(fval.data(), dfval.p, fval.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
  }
  const auto t3 = clock::now();

  // each cell's weight on its faces, as much as each is joined (in cell order: a face takes
  // from both its cells), and where its beams go (gather.rs:369-392)
  std::vector<int64_t> slot((size_t)ncell, -1);
  for (int64_t q = 0; q < nj; q++) slot[job[q]] = q;
  std::vector<int64_t> first((size_t)ncell + 1, 0);
  for (int64_t c = 0; c < ncell; c++) {
    const Cell& cell = cells[c];
    const int64_t q = slot[c];
    const int nfd = q >= 0 ? nfound[q] : 0;
    const int32_t* ab = q >= 0 ? fab.data() + q * 2 * kMaxPairs : nullptr;
    const double* val = q >= 0 ? fval.data() + q * 2 * kMaxPairs : nullptr;
    double mass = 0.0;
    for (int64_t v : cell.free) mass += fine.m[v];
    std::vector<double> joined(ends[c].size(), 0.0);
    for (int k = 0; k < nfd; k++) {
      joined[ab[2 * k]] += val[2 * k];
      joined[ab[2 * k + 1]] += val[2 * k];
    }
    double total = 0.0;
    for (double jv : joined) total += jv;
    for (size_t h = 0; h < ends[c].size(); h++)
      out.m[ends[c][h]] += mass * (total > 0.0 ? joined[h] / total : 1.0 / (double)ends[c].size());
    int64_t kept = 0;
    for (int k = 0; k < nfd; k++) kept += !(out.held[ends[c][ab[2 * k]]] && out.held[ends[c][ab[2 * k + 1]]]);
    first[c + 1] = first[c] + kPerBeam * kept;
  }
  const int64_t nbo = first[ncell];
  out.ij.resize(nbo), out.p.resize(nbo), out.n.resize(nbo), out.a.resize(nbo), out.d.resize(nbo);
  out.bmat.resize(nbo);
  par_for(ncell, [&](int64_t c0, int64_t c1) {
    for (int64_t c = c0; c < c1; c++) {
      const int64_t q = slot[c];
      if (q < 0) continue;
      const int32_t* ab = fab.data() + q * 2 * kMaxPairs;
      const double* val = fval.data() + q * 2 * kMaxPairs;
      int64_t at = first[c];
      for (int k = 0; k < nfound[q]; k++) {
        const int64_t ea = ends[c][ab[2 * k]], eb = ends[c][ab[2 * k + 1]];
        if (out.held[ea] && out.held[eb]) continue;
        const G3 xa = out.x[ea], xb = out.x[eb];
        const double length = gnorm(gsub(xb, xa));
        Mat6 st = mat[c];
        for (int t = 3; t < 6; t++) st[t] *= val[2 * k + 1];  // Mat::stronger (model.rs:71-73)
        hbeam_at(out, at, ea, eb, xa, xb, val[2 * k] * length / mat[c][1], st);
        at += kPerBeam;
      }
    }
  });
  if (std::getenv("STATICS_HIP_VERBOSE")) {
    auto ms = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::fprintf(stderr, "[statics-hip] gather: %lld cells (%lld measured): bookkeeping %.1f ms, packing %.1f ms, "
                 "device %.1f ms, assembly %.1f ms\n", (long long)ncell, (long long)nj, ms(t0, t1), ms(t1, t2),
                 ms(t2, t3), ms(t3, clock::now()));
  }
}

}  // namespace
}  // namespace st

using namespace st;

int st_gather(StCtx* c, const StLattice* fine, int against_full, int64_t* nnodes_out, int64_t* nbonds_out) {
  return guard(c, [&] {
    if (!fine || fine->nnodes < 0 || fine->nbonds < 0 || fine->level > 40)
      throw StError{ST_BAD_INPUT, "gather: no lattice, or negative sizes"};
    if (fine->nnodes && (!fine->x || !fine->m || !fine->node_mat || !fine->held || !fine->around))
      throw StError{ST_BAD_INPUT, "gather: a node array is NULL"};
    if (fine->nbonds && (!fine->ij || !fine->p || !fine->n || !fine->a || !fine->d || !fine->bond_mat))
      throw StError{ST_BAD_INPUT, "gather: a bond array is NULL"};
    for (int64_t b = 0; b < 2 * fine->nbonds; b++)
      if (fine->ij[b] < 0 || fine->ij[b] >= fine->nnodes)
        throw StError{ST_BAD_INPUT, "gather: a bond's node is out of range"};
    const LatIn L{fine->level, fine->nnodes, fine->nbonds, fine->x, fine->m, fine->node_mat, fine->held,
                  fine->around, fine->ij, fine->p, fine->n, fine->a, fine->d, fine->bond_mat};
    HLat out;
    measure(c, L, against_full != 0, out);
    GatherOut& g = c->gathered;
    g = GatherOut();
    g.level = out.level;
    const size_t on = out.x.size(), ob = out.a.size();
    g.x.resize(3 * on), g.m = out.m, g.node_mat.resize(6 * on), g.held = out.held, g.around.resize(6 * on);
    for (size_t v = 0; v < on; v++) {
      for (int k = 0; k < 3; k++) g.x[3 * v + k] = out.x[v].c[k];
      for (int k = 0; k < 6; k++) g.node_mat[6 * v + k] = out.nmat[v][k];
      for (int a = 0; a < 2; a++)
        for (int k = 0; k < 3; k++) g.around[6 * v + 3 * a + k] = out.around[v][a][k];
    }
    g.ij.resize(2 * ob), g.p.resize(3 * ob), g.n.resize(3 * ob), g.a = out.a, g.d = out.d, g.bond_mat.resize(6 * ob);
    par_for((int64_t)ob, [&](int64_t b0, int64_t b1) {
      for (int64_t b = b0; b < b1; b++) {
        g.ij[2 * b] = out.ij[b][0], g.ij[2 * b + 1] = out.ij[b][1];
        for (int k = 0; k < 3; k++) g.p[3 * b + k] = out.p[b].c[k], g.n[3 * b + k] = out.n[b].c[k];
        for (int k = 0; k < 6; k++) g.bond_mat[6 * b + k] = out.bmat[b][k];
      }
    });
    if (nnodes_out) *nnodes_out = (int64_t)on;
    if (nbonds_out) *nbonds_out = (int64_t)ob;
  });
}

int st_gather_read(StCtx* c, uint32_t* level, double* x, double* m, double* node_mat, uint8_t* held, int64_t* around,
                   int64_t* ij, double* p, double* n, double* a, double* d, double* bond_mat) {
  return guard(c, [&] {
    const GatherOut& g = c->gathered;
    auto copy = [](auto* to, const auto& from) {
      if (to && !from.empty()) std::memcpy(to, from.data(), from.size() * sizeof(from[0]));
    };
    if (level) *level = g.level;
    copy(x, g.x), copy(m, g.m), copy(node_mat, g.node_mat), copy(held, g.held), copy(around, g.around);
    copy(ij, g.ij), copy(p, g.p), copy(n, g.n), copy(a, g.a), copy(d, g.d), copy(bond_mat, g.bond_mat);
  });
}
