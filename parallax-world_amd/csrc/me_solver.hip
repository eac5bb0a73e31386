// mechanics::solve::Solver on the device (mechanics/src/solve.rs:372-537).
//
// The system's unknowns are three to a node (level 0); a level up, six to a part (its
// translation at its middle and its turn). Each level keeps its blocks as a block CSR with
// both halves (a pair (i, j, B) gives B in row i and B^T in row j), its diagonal blocks
// apart, and their inverses. The host does the bookkeeping of a level, as the reference's
// coarsen does it (solve.rs:253-301): the parts (union-find over the pairs within one octree
// cell a level up, numbered as first met), their middles and keys, which finer blocks add
// into which part's and which pair's; the device forms the blocks (the Galerkin lifts
// T(r)^T M T(s), solve.rs:76-98), their inverses, the smoothing steps (solve.rs:309-327)
// and runs the iterations. The top level is factored on the host as the reference does
// (solve.rs:343-372) and solved on the device by one wave.
//
// Every sum over a part's members and over a coarse pair's finer pairs is taken in the
// reference's order, with contraction off as in the reference, so the first coarse level
// is the reference's to the bit; the products' and dots' sums are taken in a fixed order
// of this build's own (rounding differs from the reference's sequential sums).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mechanics_hip.h"
#include "st_internal.hpp"

#pragma clang fp contract(off)

namespace me {
namespace {

using st::DBuf;
using st::StError;

constexpr int kRounds = 300;         // solve.rs:21
constexpr int64_t kDirect = 24;      // solve.rs:23
constexpr int64_t kMaxTop = 1024;    // parts a top level solved directly may hold here
constexpr int kBlock = 256;
constexpr int kDotBlocks = 1024;     // partial sums of a dot product

enum Scal { RZ = 0, PAP = 1, NEXT = 2, NW = 3, NV = 4, TARGET = 5, kScal = 8 };

unsigned nblk(int64_t n, int b = kBlock) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

// ------------------------------------------------------------ device math

// Cholesky with dropped pivots (solve.rs:141-166), in place on a (n x n, row-major).
template <int N>
__device__ void cholesky(double* a, bool* dropped) {
  double largest = 0.0;
  for (int i = 0; i < N; i++) largest = fmax(largest, a[i * N + i]);
  for (int j = 0; j < N; j++) {
    dropped[j] = false;
    double d = a[j * N + j];
    for (int k = 0; k < j; k++) d -= a[j * N + k] * a[j * N + k];
    if (d <= largest * 1e-12) {
      dropped[j] = true;
      for (int i = j; i < N; i++) a[i * N + j] = 0.0;
      continue;
    }
    d = sqrt(d);
    a[j * N + j] = d;
    for (int i = j + 1; i < N; i++) {
      double s = a[i * N + j];
      for (int k = 0; k < j; k++) s -= a[i * N + k] * a[j * N + k];
      a[i * N + j] = s / d;
    }
  }
}

// solve.rs:168-187
template <int N>
__device__ void solve_cholesky(const double* a, const bool* dropped, double* b) {
  for (int i = 0; i < N; i++) {
    if (dropped[i]) {
      b[i] = 0.0;
      continue;
    }
    double s = 0.0;
    for (int k = 0; k < i; k++) s += a[i * N + k] * b[k];
    b[i] = (b[i] - s) / a[i * N + i];
  }
  for (int i = N - 1; i >= 0; i--) {
    if (dropped[i]) {
      b[i] = 0.0;
      continue;
    }
    double s = 0.0;
    for (int k = i + 1; k < N; k++) s += a[k * N + i] * b[k];
    b[i] = (b[i] - s) / a[i * N + i];
  }
}

// inverse (solve.rs:129-139): row c of the result is the solution for the unit vector c.
template <int N>
__device__ void inverse(double* a, double* out) {
  bool dropped[N];
  cholesky<N>(a, dropped);
  for (int c = 0; c < N; c++) {
    double e[N];
    for (int r = 0; r < N; r++) e[r] = r == c ? 1.0 : 0.0;
    solve_cholesky<N>(a, dropped, e);
    for (int r = 0; r < N; r++) out[c * N + r] = e[r];
  }
}

// T(r)^T m T(s) in place (solve.rs:76-98): the right half of each row gains its left half
// times X(s), then the lower half gains X(r)^T times the upper half; X(r) w = w x r.
__device__ void lift(double* m, const double* r, const double* s) {
  const double xs[9] = {0.0, s[2], -s[1], -s[2], 0.0, s[0], s[1], -s[0], 0.0};
  const double xr[9] = {0.0, r[2], -r[1], -r[2], 0.0, r[0], r[1], -r[0], 0.0};
  for (int row = 0; row < 6; row++) {
    double add[3];
    for (int c = 0; c < 3; c++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc += m[row * 6 + k] * xs[k * 3 + c];
      add[c] = acc;
    }
    for (int c = 0; c < 3; c++) m[row * 6 + 3 + c] += add[c];
  }
  for (int r3 = 0; r3 < 3; r3++)
    for (int col = 0; col < 6; col++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc += xr[k * 3 + r3] * m[k * 6 + col];
      m[(3 + r3) * 6 + col] += acc;
    }
}

// A finer block as 6x6 (embed, solve.rs:100-106, for the three-unknown level).
template <int B>
__device__ void load6(const double* src, double* m) {
  if constexpr (B == 6) {
    for (int i = 0; i < 36; i++) m[i] = src[i];
  } else {
    for (int i = 0; i < 36; i++) m[i] = 0.0;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) m[r * 6 + c] = src[r * 3 + c];
  }
}

// ------------------------------------------------------------ kernels

// Block CSR values from the pairs' blocks: slot s takes pair src[s] >> 1, transposed if
// its low bit is set.
template <int B>
__global__ void k_scatter(int64_t nslots, const int64_t* src, const double* blocks, double* val) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int64_t k = src[s] >> 1;
  const bool t = src[s] & 1;
  const double* b = blocks + (size_t)B * B * k;
  double* o = val + (size_t)B * B * s;
  for (int r = 0; r < B; r++)
    for (int c = 0; c < B; c++) o[r * B + c] = t ? b[c * B + r] : b[r * B + c];
}

// The inverses of the diagonal blocks; for parts (B = 6) each turn first held by a little
// of what its translation is, at the level's size (solve.rs:286-297).
template <int B>
__global__ __launch_bounds__(64) void k_inverse(int64_t n, const double* diag, double size, double* inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[B * B], o[B * B];
  for (int k = 0; k < B * B; k++) a[k] = diag[(size_t)B * B * i + k];
  if constexpr (B == 6) {
    const double t = (a[0] + a[7] + a[14]) / 3.0;
    for (int k = 3; k < 6; k++) a[k * 6 + k] += 1e-6 * t * size * size;
  }
  inverse<B>(a, o);
  for (int k = 0; k < B * B; k++) inv[(size_t)B * B * i + k] = o[k];
}

// y = D x + sum of the row's blocks times x (System::times / Coarse::apply, solve.rs:61-72,
// 330-340), then by mode: 0 out = y; 1 out = b - y; 2 out = x + step inv (b - y).
// kRowLanes lanes to a row: lane g takes the row's blocks g, g + kRowLanes, ... (the lanes
// of a row read neighbouring blocks), lane 0 the diagonal block; their sums meet by
// shuffles in a fixed order.
constexpr int kRowLanes = 8;
template <int B, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n, const double* __restrict__ diag,
                                                 const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                 const double* __restrict__ val, const double* __restrict__ x,
                                                 const double* __restrict__ b, const double* __restrict__ inv,
                                                 double step, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / kRowLanes;
  const int g = (int)(t % kRowLanes);
  const bool live = i < n;
  double y[B], xi[B];
  for (int k = 0; k < B; k++) y[k] = 0.0;
  if (live) {
    if (g == 0) {
      for (int k = 0; k < B; k++) xi[k] = x[(size_t)B * i + k];
      const double* d = diag + (size_t)B * B * i;
      for (int r = 0; r < B; r++) {
        double s = 0.0;
        for (int c = 0; c < B; c++) s += d[r * B + c] * xi[c];
        y[r] = s;
      }
    }
    for (int64_t e = rowptr[i] + g; e < rowptr[i + 1]; e += kRowLanes) {
      const int32_t j = col[e];
      const double* m = val + (size_t)B * B * e;
      double xj[B];
      for (int k = 0; k < B; k++) xj[k] = x[(size_t)B * j + k];
      for (int r = 0; r < B; r++) {
        double s = 0.0;
        for (int c = 0; c < B; c++) s += m[r * B + c] * xj[c];
        y[r] += s;
      }
    }
  }
  for (int o = 1; o < kRowLanes; o <<= 1)
    for (int k = 0; k < B; k++) y[k] += __shfl_xor(y[k], o);
  if (!live || g != 0) return;
  if constexpr (MODE == 0) {
    for (int k = 0; k < B; k++) out[(size_t)B * i + k] = y[k];
  } else {
    double res[B];
    for (int k = 0; k < B; k++) res[k] = b[(size_t)B * i + k] - y[k];
    if constexpr (MODE == 1) {
      for (int k = 0; k < B; k++) out[(size_t)B * i + k] = res[k];
    } else {
      const double* v = inv + (size_t)B * B * i;
      for (int r = 0; r < B; r++) {
        double s = 0.0;
        for (int c = 0; c < B; c++) s += v[r * B + c] * res[c];
        out[(size_t)B * i + r] = xi[r] + s * step;
      }
    }
  }
}

// x = step inv b (the first smoothing step from nothing); step 1: inv b (step's `over`).
template <int B>
__global__ void k_jacobi_set(int64_t n, const double* inv, double step, const double* b, double* x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* v = inv + (size_t)B * B * i;
  double bi[B];
  for (int k = 0; k < B; k++) bi[k] = b[(size_t)B * i + k];
  for (int r = 0; r < B; r++) {
    double s = 0.0;
    for (int c = 0; c < B; c++) s += v[r * B + c] * bi[c];
    x[(size_t)B * i + r] = B == 3 ? s * step : 0.0 + step * s;
  }
}

// What a level leaves unbalanced goes up as forces and torques: each part sums its
// members' (in order) residuals and their moments about its middle (solve.rs:444-452,
// 478-486).
template <int B>
__global__ void k_restrict(int64_t n, const int64_t* mptr, const int32_t* mlist, const double* roff,
                           const double* res, double* rb) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t e = mptr[a]; e < mptr[a + 1]; e++) {
    const int64_t p = mlist[e];
    const double* r = roff + 3 * (size_t)p;
    const double* v = res + (size_t)B * p;
    const double t[3] = {r[1] * v[2] - r[2] * v[1], r[2] * v[0] - r[0] * v[2], r[0] * v[1] - r[1] * v[0]};
    for (int k = 0; k < 3; k++) {
      acc[k] += v[k];
      if constexpr (B == 6)
        acc[3 + k] += v[3 + k] + t[k];
      else
        acc[3 + k] += t[k];
    }
  }
  for (int k = 0; k < 6; k++) rb[6 * (size_t)a + k] = acc[k];
}

// ... and comes back as motions: each finer part moves with its part, translation plus the
// turn about the middle (solve.rs:454-461, 488-492).
template <int B>
__global__ void k_prolong(int64_t n, const uint32_t* of, const double* roff, const double* e, double* x) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double* ea = e + 6 * (size_t)of[p];
  const double* r = roff + 3 * (size_t)p;
  const double u[3] = {ea[4] * r[2] - ea[5] * r[1], ea[5] * r[0] - ea[3] * r[2], ea[3] * r[1] - ea[4] * r[0]};
  double* xp = x + (size_t)B * p;
  if constexpr (B == 3) {
    for (int k = 0; k < 3; k++) xp[k] = xp[k] + (ea[k] + u[k]);
  } else {
    for (int k = 0; k < 3; k++) {
      xp[k] += ea[k] + u[k];
      xp[3 + k] += ea[3 + k];
    }
  }
}

// A part's block (solve.rs:276-286): its members' blocks lifted to its middle, in order,
// then the pairs within it, in order, each with its transpose.
template <int FB>
__global__ __launch_bounds__(64) void k_lift_diag(int64_t n, const int64_t* mptr, const int32_t* mlist, const int64_t* iptr,
                            const int64_t* ilist, const uint32_t* fij, const double* roff, const double* fdiag,
                            const double* fblocks, double* out) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double acc[36];
  for (int k = 0; k < 36; k++) acc[k] = 0.0;
  double m[36];
  for (int64_t e = mptr[a]; e < mptr[a + 1]; e++) {
    const int64_t i = mlist[e];
    load6<FB>(fdiag + (size_t)FB * FB * i, m);
    lift(m, roff + 3 * (size_t)i, roff + 3 * (size_t)i);
    for (int k = 0; k < 36; k++) acc[k] += m[k];
  }
  for (int64_t e = iptr[a]; e < iptr[a + 1]; e++) {
    const int64_t k = ilist[e];
    const uint32_t i = fij[2 * k], j = fij[2 * k + 1];
    load6<FB>(fblocks + (size_t)FB * FB * k, m);
    lift(m, roff + 3 * (size_t)i, roff + 3 * (size_t)j);
    for (int t = 0; t < 36; t++) acc[t] += m[t];
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) acc[r * 6 + c] += m[c * 6 + r];
  }
  for (int k = 0; k < 36; k++) out[36 * (size_t)a + k] = acc[k];
}

// A pair of parts' block (solve.rs:287-296): the lifts of the finer pairs between them, in
// order, transposed where the finer pair runs from the greater part (low bit of the entry).
template <int FB>
__global__ __launch_bounds__(64) void k_lift_pairs(int64_t npairs, const int64_t* sptr, const int64_t* slist, const uint32_t* fij,
                             const double* roff, const double* fblocks, double* out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npairs) return;
  double acc[36];
  for (int k = 0; k < 36; k++) acc[k] = 0.0;
  double m[36];
  for (int64_t e = sptr[s]; e < sptr[s + 1]; e++) {
    const int64_t k = slist[e] >> 1;
    const bool t = slist[e] & 1;
    const uint32_t i = fij[2 * k], j = fij[2 * k + 1];
    load6<FB>(fblocks + (size_t)FB * FB * k, m);
    lift(m, roff + 3 * (size_t)i, roff + 3 * (size_t)j);
    if (t) {
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) acc[r * 6 + c] += m[c * 6 + r];
    } else {
      for (int k2 = 0; k2 < 36; k2++) acc[k2] += m[k2];
    }
  }
  for (int k = 0; k < 36; k++) out[36 * (size_t)s + k] = acc[k];
}

// Dot products (and squared norms, b == a) in a fixed order: per block, then the blocks'
// partial sums in order by one block.
__global__ __launch_bounds__(kBlock) void k_dot_partial(int64_t n, const double* a, const double* b, double* part) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(kBlock) void k_dot_final(int nparts, const double* part, double* scal, int idx) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[idx] = sh[0];
}

// The power iteration's start (solve.rs:312): v[k][s] = 1 + ((7k + 3s) mod 11) / 11.
__global__ void k_power_start(int64_t n, int B, double* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * B) return;
  const int64_t k = i / B, s = i % B;
  v[i] = 1.0 + (double)((7 * k + 3 * s) % 11) / 11.0;
}
__global__ void k_scale_by(int64_t n, const double* w, const double* scal, int idx, double* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = w[i] / sqrt(scal[idx]);
}

// CG updates (solve.rs:522-535), the step lengths from the device's dots.
__global__ void k_cg_update(int64_t n, const double* p, const double* ap, double* x, double* r, const double* scal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(scal[PAP] > 0.0)) return;
  const double alpha = scal[RZ] / scal[PAP];
  x[i] = x[i] + p[i] * alpha;
  r[i] = r[i] - ap[i] * alpha;
}
__global__ void k_cg_p(int64_t n, const double* z, double* p, const double* scal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = scal[NEXT] / scal[RZ];
  p[i] = z[i] + p[i] * beta;
}
__global__ void k_cg_shift(double* scal) { scal[RZ] = scal[NEXT]; }
__global__ void k_sub(int64_t n, const double* a, const double* b, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}

// The top level solved with its factor (solve.rs:426-430) by one wave: each step's sum over
// the lanes in a fixed order.
__global__ __launch_bounds__(64) void k_dense_solve(int m, const double* a, const uint8_t* dropped, double* b) {
  extern __shared__ double sb[];
  for (int i = threadIdx.x; i < m; i += 64) sb[i] = b[i];
  __syncthreads();
  auto wave_sum = [](double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
  };
  for (int i = 0; i < m; i++) {
    double s = 0.0;
    for (int k = threadIdx.x; k < i; k += 64) s += a[(size_t)i * m + k] * sb[k];
    s = wave_sum(s);
    __syncthreads();
    if (threadIdx.x == 0) sb[i] = dropped[i] ? 0.0 : (sb[i] - s) / a[(size_t)i * m + i];
    __syncthreads();
  }
  for (int i = m - 1; i >= 0; i--) {
    double s = 0.0;
    for (int k = i + 1 + threadIdx.x; k < m; k += 64) s += a[(size_t)k * m + i] * sb[k];
    s = wave_sum(s);
    __syncthreads();
    if (threadIdx.x == 0) sb[i] = dropped[i] ? 0.0 : (sb[i] - s) / a[(size_t)i * m + i];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < m; i += 64) b[i] = sb[i];
}

// ------------------------------------------------------------ host side

using Key = std::array<int64_t, 3>;
using P3 = std::array<double, 3>;

int64_t div_euclid2(int64_t v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

// f(t, lo, hi) over [0, n) in contiguous chunks, one host thread each (chunk t before
// chunk t + 1, so per-chunk results merged in chunk order keep the sequential order).
int host_threads() {
  static const int t = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  return t;
}
template <class F>
void chunks(int64_t n, int nt, F f) {
  if (nt <= 1 || n < 65536) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) th.emplace_back([&, t] { f(t, n * t / nt, n * (t + 1) / nt); });
  for (auto& x : th) x.join();
}
Key parent(const Key& k) { return {div_euclid2(k[0]), div_euclid2(k[1]), div_euclid2(k[2])}; }

template <class T>
void up(DBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.ensure(std::max<size_t>(h.size() * sizeof(T), 1));
  if (!h.empty()) ST_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// A block CSR with both halves of each pair: row i gets pair k as (j, 2k), row j as (i, 2k+1).
struct Csr {
  int64_t nslots = 0;
  DBuf rowptr, col, src, val;
};
// Entry e of the both-halves listing: pair e >> 1, in the row of its first unknown (e even)
// or its second (e odd, transposed).
__global__ void k_csr_entries(int64_t ne, const uint32_t* ij, uint32_t* key, int64_t* val) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  key[e] = ij[e];  // ij[2k + h]: the row of half h of pair k
  val[e] = e;
}
// After the stable sort by row: each row's first entry (binary search), and each entry's
// column (the pair's other unknown).
__global__ void k_csr_rows(int64_t n, int64_t ne, const uint32_t* key, int64_t* rowptr) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  int64_t lo = 0, hi = ne;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if ((int64_t)key[mid] < r)
      lo = mid + 1;
    else
      hi = mid;
  }
  rowptr[r] = lo;
}
__global__ void k_csr_cols(int64_t ne, const int64_t* src, const uint32_t* ij, int32_t* col) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  col[e] = (int32_t)ij[src[e] ^ 1];
}

// The block CSR of a level's pairs (device ij, np pairs) with both halves of each pair: row i
// gets pair k as (j, 2k), row j as (i, 2k+1), each row's entries in pair order (a stable
// sort of the listing by row).
void csr_build(Csr& c, int64_t n, const uint32_t* dij, int64_t np, hipStream_t s) {
  const int64_t ne = 2 * np;
  c.nslots = ne;
  c.rowptr.ensure((size_t)(n + 1) * sizeof(int64_t));
  c.col.ensure(std::max<size_t>(ne * sizeof(int32_t), 1));
  c.src.ensure(std::max<size_t>(ne * sizeof(int64_t), 1));
  if (ne == 0) {
    ST_HIP(hipMemsetAsync(c.rowptr.p, 0, (size_t)(n + 1) * sizeof(int64_t), s));
    return;
  }
  if (ne >= (int64_t(1) << 31)) throw StError{ST_BAD_INPUT, "system: more than 2^30 pairs"};
  DBuf kin, kout, vin, tmp;
  kin.ensure(ne * sizeof(uint32_t));
  kout.ensure(ne * sizeof(uint32_t));
  vin.ensure(ne * sizeof(int64_t));
  k_csr_entries<<<nblk(ne), kBlock, 0, s>>>(ne, dij, kin.as<uint32_t>(), vin.as<int64_t>());
  ST_HIP(hipGetLastError());
  int bits = 1;
  while (bits < 32 && (int64_t(1) << bits) < n) bits++;
  size_t bytes = 0;
  ST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin.as<uint32_t>(), kout.as<uint32_t>(),
                                            vin.as<int64_t>(), c.src.as<int64_t>(), (int)ne, 0, bits, s));
  tmp.ensure(bytes + 16);
  ST_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin.as<uint32_t>(), kout.as<uint32_t>(),
                                            vin.as<int64_t>(), c.src.as<int64_t>(), (int)ne, 0, bits, s));
  k_csr_rows<<<nblk(n + 1), kBlock, 0, s>>>(n, ne, kout.as<uint32_t>(), c.rowptr.as<int64_t>());
  k_csr_cols<<<nblk(ne), kBlock, 0, s>>>(ne, c.src.as<int64_t>(), dij, c.col.as<int32_t>());
  ST_HIP(hipGetLastError());
  ST_HIP(hipStreamSynchronize(s));
  for (DBuf* b : {&kin, &kout, &vin, &tmp}) b->release();
}

struct Level {
  int64_t n = 0, nfiner = 0, npairs = 0;
  double step = 1.0, size = 0.0;
  bool direct = false;
  std::vector<Key> keys;
  std::vector<P3> centre;
  std::vector<std::array<uint32_t, 2>> ij;  // its pairs (a < c), in (a, c) order
  DBuf of, roff;                            // per finer part: its part, its offset from the middle
  DBuf mptr, mlist;                         // each part's members, in order
  DBuf pij;                                 // the pairs, on the device
  DBuf diag, inv, pairval;                  // 36 per part, per part, per pair
  Csr csr;
  DBuf fact, dropped;                       // the top: its factor and dropped pivots
  DBuf b, x, r, x2;                         // the cycle's vectors
};

}  // namespace

struct Solver {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int64_t n = 0;
  double step0 = 1.0;
  DBuf inv0;                                  // the setup system's diagonal inverses
  // the current system (level 0): its pattern (to see whether the next one differs), blocks
  std::vector<std::array<uint32_t, 2>> pattern;
  int64_t npairs0 = 0;
  DBuf diag0, blocks0, ij0;
  Csr csr0;
  std::vector<Level> levels;
  DBuf x, r, z, p, ap, bvec, t1, t2, part, scal;
  double* hscal = nullptr;
  double setup_ms = 0.0, solve_ms = 0.0;
  // device time of the CG's K p (the level-0 product), when asked for: event pairs around
  // each launch, folded into the totals at the end of each solve
  bool timing = false;
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  int64_t t_launches = 0;
  double t_ms = 0.0;

  ~Solver();
  void timing_collect();
  void upload_system(const MeSystem& s, bool setup);
  double power_step(int B, int64_t nn, const double* diag, const Csr& c, const double* inv, DBuf& v, DBuf& a, DBuf& w);
  void coarsen(const std::vector<Key>& keys, const std::vector<P3>& at, const std::vector<std::array<uint32_t, 2>>& fij,
               const double* fdiag, const double* fblocks, const uint32_t* dfij, int FB, double size, Level& L);
  void factor(Level& L);
  double* cycle(size_t l, double* b);
  void precondition(const double* b, double* out);
  double dot(const double* a, const double* b, int64_t nn, int idx, bool read);
  void build(const MeSystem& s, double size);
  int32_t solve(const MeSystem& s, const double* b, double* x, double tol);
};

namespace {

void check_system(const MeSystem& s, int64_t n_expect) {
  if (s.n < 0 || s.n >= (int64_t(1) << 31)) throw StError{ST_BAD_INPUT, "system: n out of range"};
  if (n_expect >= 0 && s.n != n_expect)
    throw StError{ST_BAD_INPUT, "system: not the unknowns the solver was made for"};
  if (s.n && !s.diag) throw StError{ST_BAD_INPUT, "system: diag is NULL"};
  if (s.npairs < 0 || (s.npairs && (!s.pairs || !s.blocks))) throw StError{ST_BAD_INPUT, "system: pairs missing"};
  for (int64_t k = 0; k < s.npairs; k++) {
    const uint32_t i = s.pairs[2 * k], j = s.pairs[2 * k + 1];
    if ((int64_t)i >= s.n || (int64_t)j >= s.n || i == j)
      throw StError{ST_BAD_INPUT, "system: pair " + std::to_string(k) + " names a missing unknown or one twice"};
  }
}

}  // namespace

void release_level(Level& L) {
  for (DBuf* b : {&L.of, &L.roff, &L.mptr, &L.mlist, &L.pij, &L.diag, &L.inv, &L.pairval, &L.csr.rowptr, &L.csr.col,
                  &L.csr.src, &L.csr.val, &L.fact, &L.dropped, &L.b, &L.x, &L.r, &L.x2})
    b->release();
}

void Solver::timing_collect() {
  for (size_t i = 0; i < tev_used; i++) {
    float ms = 0.0f;
    ST_HIP(hipEventElapsedTime(&ms, tev[2 * i], tev[2 * i + 1]));
    t_ms += ms;
    t_launches++;
  }
  tev_used = 0;
}

Solver::~Solver() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (hipEvent_t e : tev) (void)hipEventDestroy(e);
  for (Level& L : levels) release_level(L);
  for (DBuf* b : {&inv0, &diag0, &blocks0, &ij0, &csr0.rowptr, &csr0.col, &csr0.src, &csr0.val, &x, &r, &z, &p, &ap,
                  &bvec, &t1, &t2, &part, &scal})
    b->release();
  if (hscal) (void)hipHostFree(hscal);
  if (stream) (void)hipStreamDestroy(stream);
}

// The current system's blocks on the device; its CSR again only when its pairs differ.
void Solver::upload_system(const MeSystem& s, bool setup) {
  std::vector<std::array<uint32_t, 2>> ij(s.npairs);
  for (int64_t k = 0; k < s.npairs; k++) ij[k] = {s.pairs[2 * k], s.pairs[2 * k + 1]};
  diag0.ensure(std::max<size_t>(9 * sizeof(double) * s.n, 1));
  if (s.n) ST_HIP(hipMemcpyAsync(diag0.p, s.diag, 9 * sizeof(double) * s.n, hipMemcpyHostToDevice, stream));
  blocks0.ensure(std::max<size_t>(9 * sizeof(double) * s.npairs, 1));
  if (s.npairs) ST_HIP(hipMemcpyAsync(blocks0.p, s.blocks, 9 * sizeof(double) * s.npairs, hipMemcpyHostToDevice, stream));
  if (setup || ij != pattern) {
    pattern = ij;
    npairs0 = s.npairs;
    up(ij0, ij, stream);
    csr_build(csr0, s.n, ij0.as<uint32_t>(), s.npairs, stream);
    csr0.val.ensure(std::max<size_t>(9 * sizeof(double) * csr0.nslots, 1));
  }
  if (csr0.nslots)
    k_scatter<3><<<nblk(csr0.nslots), kBlock, 0, stream>>>(csr0.nslots, csr0.src.as<int64_t>(), blocks0.as<double>(),
                                                            csr0.val.as<double>());
  ST_HIP(hipGetLastError());
}

double Solver::dot(const double* a, const double* b, int64_t nn, int idx, bool read) {
  k_dot_partial<<<kDotBlocks, kBlock, 0, stream>>>(nn, a, b, part.as<double>());
  k_dot_final<<<1, kBlock, 0, stream>>>(kDotBlocks, part.as<double>(), scal.as<double>(), idx);
  ST_HIP(hipGetLastError());
  if (!read) return 0.0;
  ST_HIP(hipMemcpyAsync(hscal, scal.p, kScal * sizeof(double), hipMemcpyDeviceToHost, stream));
  ST_HIP(hipStreamSynchronize(stream));
  return hscal[idx];
}

// step (solve.rs:309-327): ten powers of inv A from the fixed start.
double Solver::power_step(int B, int64_t nn, const double* diag, const Csr& c, const double* inv, DBuf& v, DBuf& a,
                          DBuf& w) {
  const int64_t len = nn * B;
  if (nn == 0) return 4.0 / 3.0;
  k_power_start<<<nblk(len), kBlock, 0, stream>>>(nn, B, v.as<double>());
  double most = 1.0;
  for (int it = 0; it < 10; it++) {
    if (B == 3) {
      k_spmv<3, 0><<<nblk(nn * kRowLanes), kBlock, 0, stream>>>(nn, diag, c.rowptr.as<int64_t>(), c.col.as<int32_t>(),
                                                     c.val.as<double>(), v.as<double>(), nullptr, nullptr, 0.0,
                                                     a.as<double>());
      k_jacobi_set<3><<<nblk(nn), kBlock, 0, stream>>>(nn, inv, 1.0, a.as<double>(), w.as<double>());
    } else {
      k_spmv<6, 0><<<nblk(nn * kRowLanes), kBlock, 0, stream>>>(nn, diag, c.rowptr.as<int64_t>(), c.col.as<int32_t>(),
                                                     c.val.as<double>(), v.as<double>(), nullptr, nullptr, 0.0,
                                                     a.as<double>());
      k_jacobi_set<6><<<nblk(nn), kBlock, 0, stream>>>(nn, inv, 1.0, a.as<double>(), w.as<double>());
    }
    ST_HIP(hipGetLastError());
    dot(w.as<double>(), w.as<double>(), len, NW, false);
    const double nv2 = dot(v.as<double>(), v.as<double>(), len, NV, true);
    const double nw = std::sqrt(hscal[NW]), nv = std::sqrt(nv2);
    if (nw == 0.0) break;
    most = nw / nv;
    k_scale_by<<<nblk(len), kBlock, 0, stream>>>(len, w.as<double>(), scal.as<double>(), NW, v.as<double>());
    ST_HIP(hipGetLastError());
  }
  return 4.0 / 3.0 / most;
}

// The level above a finer one (solve.rs:253-301): host bookkeeping, device blocks.
void Solver::coarsen(const std::vector<Key>& keys, const std::vector<P3>& at,
                     const std::vector<std::array<uint32_t, 2>>& fij, const double* fdiag, const double* fblocks,
                     const uint32_t* dfij, int FB, double size, Level& L) {
  const int64_t nf = (int64_t)keys.size();
  const bool verbose = std::getenv("STATICS_HIP_VERBOSE") != nullptr && fij.size() > 1000000;
  const auto t0 = std::chrono::steady_clock::now();
  auto mark = [&](const char* what) {
    if (verbose)
      std::fprintf(stderr, "mechanics_hip:   coarsen %s at %.1f ms\n", what,
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  };
  // the parts: union at the lesser root over the pairs within one octree cell a level up,
  // numbered as first met (solve.rs:210-250)
  std::vector<int64_t> upv(nf);
  for (int64_t i = 0; i < nf; i++) upv[i] = i;
  auto find = [&](int64_t i) {
    while (upv[i] != i) {
      upv[i] = upv[upv[i]];
      i = upv[i];
    }
    return i;
  };
  const int nt = host_threads();
  std::vector<Key> par(nf);
  chunks(nf, nt, [&](int, int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; i++) par[i] = parent(keys[i]);
  });
  const int64_t nfp = (int64_t)fij.size();
  std::vector<uint8_t> within(nfp);
  chunks(nfp, nt, [&](int, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; k++) within[k] = par[fij[k][0]] == par[fij[k][1]];
  });
  for (int64_t k = 0; k < nfp; k++)
    if (within[k]) {
      const int64_t r = find(fij[k][0]), s2 = find(fij[k][1]);
      upv[std::max(r, s2)] = std::min(r, s2);
    }
  mark("union-find");
  std::vector<uint32_t> index(nf, UINT32_MAX), of(nf);
  int64_t n = 0;
  for (int64_t i = 0; i < nf; i++) {
    const int64_t root = find(i);
    if (index[root] == UINT32_MAX) index[root] = (uint32_t)n++;
    of[i] = index[root];
  }
  L.n = n;
  L.nfiner = nf;
  L.size = size;
  // middles, keys, offsets (solve.rs:255-266)
  L.centre.assign(n, P3{0, 0, 0});
  std::vector<double> count(n, 0.0);
  L.keys.assign(n, Key{0, 0, 0});
  for (int64_t i = 0; i < nf; i++) {
    const uint32_t a = of[i];
    for (int k = 0; k < 3; k++) L.centre[a][k] = L.centre[a][k] + at[i][k];
    count[a] += 1.0;
    L.keys[a] = par[i];
  }
  for (int64_t a = 0; a < n; a++)
    for (int k = 0; k < 3; k++) L.centre[a][k] = L.centre[a][k] * (1.0 / count[a]);
  std::vector<double> roff(3 * nf);
  for (int64_t i = 0; i < nf; i++)
    for (int k = 0; k < 3; k++) roff[3 * i + k] = at[i][k] - L.centre[of[i]][k];
  // members in order; the pairs within each part, in order; the pairs between parts
  std::vector<int64_t> mptr(n + 1, 0);
  for (int64_t i = 0; i < nf; i++) mptr[of[i] + 1]++;
  for (int64_t a = 0; a < n; a++) mptr[a + 1] += mptr[a];
  std::vector<int32_t> mlist(nf);
  {
    std::vector<int64_t> f(mptr.begin(), mptr.end() - 1);
    for (int64_t i = 0; i < nf; i++) mlist[f[of[i]]++] = (int32_t)i;
  }
  mark("parts, middles, members");
  // the pairs within each part (in order), and the pairs between parts bucketed by their
  // lesser part (stable, so in order within a bucket), then each bucket sorted by the
  // greater part: runs of one (a, c) are a coarse pair's finer pairs in order, the coarse
  // pairs in (a, c) order
  // per chunk of pairs its counts per part, then each chunk's place in each bucket (chunks in
  // order), then the fills; at most 64 MiB of counts
  const int ntb = (int)std::max<int64_t>(1, std::min<int64_t>(nt, (int64_t(1) << 26) / std::max<int64_t>(16 * n, 1)));
  std::vector<std::vector<int64_t>> ci(ntb, std::vector<int64_t>(n, 0)), cb(ntb, std::vector<int64_t>(n, 0));
  chunks(nfp, ntb, [&](int t, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; k++) {
      const uint32_t a = of[fij[k][0]], c = of[fij[k][1]];
      if (a == c)
        ci[t][a]++;
      else
        cb[t][std::min(a, c)]++;
    }
  });
  std::vector<int64_t> iptr(n + 1, 0), bptr(n + 1, 0);
  for (int64_t a = 0; a < n; a++) {
    int64_t si = iptr[a], sb = bptr[a];
    for (int t = 0; t < ntb; t++) {
      const int64_t xi = ci[t][a], xb = cb[t][a];
      ci[t][a] = si;  // chunk t's first place in bucket a
      cb[t][a] = sb;
      si += xi;
      sb += xb;
    }
    iptr[a + 1] = si;
    bptr[a + 1] = sb;
  }
  std::vector<int64_t> ilist(iptr[n]);
  std::vector<std::pair<uint32_t, int64_t>> bl(bptr[n]);  // (greater part, finer pair k << 1 | transposed)
  chunks(nfp, ntb, [&](int t, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; k++) {
      const uint32_t a = of[fij[k][0]], c = of[fij[k][1]];
      if (a == c)
        ilist[ci[t][a]++] = k;
      else
        bl[cb[t][std::min(a, c)]++] = {std::max(a, c), 2 * k + (a > c ? 1 : 0)};
    }
  });
  ci.clear();
  cb.clear();
  mark("pairs bucketed");
  chunks(n, nt, [&](int, int64_t lo, int64_t hi) {
    for (int64_t a = lo; a < hi; a++) std::sort(bl.begin() + bptr[a], bl.begin() + bptr[a + 1]);
  });
  mark("buckets sorted");
  L.ij.clear();
  std::vector<int64_t> sptr(1, 0), slist(bl.size());
  for (int64_t a = 0; a < n; a++)
    for (int64_t e = bptr[a]; e < bptr[a + 1]; e++) {
      if (e == bptr[a] || bl[e].first != bl[e - 1].first) {
        L.ij.push_back({(uint32_t)a, bl[e].first});
        sptr.push_back(sptr.back());
      }
      slist[e] = bl[e].second;
      sptr.back()++;
    }
  const int64_t np = (int64_t)L.ij.size();
  L.npairs = np;
  mark("coarse pairs listed");
  // to the device: the maps, then the blocks
  up(L.of, of, stream);
  up(L.roff, roff, stream);
  up(L.mptr, mptr, stream);
  up(L.mlist, mlist, stream);
  up(L.pij, L.ij, stream);
  DBuf dip, dil, dsp, dsl;
  up(dip, iptr, stream);
  up(dil, ilist, stream);
  up(dsp, sptr, stream);
  up(dsl, slist, stream);
  L.diag.ensure(std::max<size_t>(36 * sizeof(double) * n, 1));
  L.inv.ensure(std::max<size_t>(36 * sizeof(double) * n, 1));
  L.pairval.ensure(std::max<size_t>(36 * sizeof(double) * np, 1));
  if (n) {
    if (FB == 3)
      k_lift_diag<3><<<nblk(n, 64), 64, 0, stream>>>(n, L.mptr.as<int64_t>(), L.mlist.as<int32_t>(),
                                                     dip.as<int64_t>(), dil.as<int64_t>(), dfij, L.roff.as<double>(),
                                                     fdiag, fblocks, L.diag.as<double>());
    else
      k_lift_diag<6><<<nblk(n, 64), 64, 0, stream>>>(n, L.mptr.as<int64_t>(), L.mlist.as<int32_t>(),
                                                     dip.as<int64_t>(), dil.as<int64_t>(), dfij, L.roff.as<double>(),
                                                     fdiag, fblocks, L.diag.as<double>());
    k_inverse<6><<<nblk(n, 64), 64, 0, stream>>>(n, L.diag.as<double>(), size, L.inv.as<double>());
  }
  if (np) {
    if (FB == 3)
      k_lift_pairs<3><<<nblk(np, 64), 64, 0, stream>>>(np, dsp.as<int64_t>(), dsl.as<int64_t>(), dfij,
                                                       L.roff.as<double>(), fblocks, L.pairval.as<double>());
    else
      k_lift_pairs<6><<<nblk(np, 64), 64, 0, stream>>>(np, dsp.as<int64_t>(), dsl.as<int64_t>(), dfij,
                                                       L.roff.as<double>(), fblocks, L.pairval.as<double>());
  }
  ST_HIP(hipGetLastError());
  mark("maps uploaded, lifts launched");
  csr_build(L.csr, n, L.pij.as<uint32_t>(), np, stream);
  mark("csr built");
  L.csr.val.ensure(std::max<size_t>(36 * sizeof(double) * L.csr.nslots, 1));
  if (L.csr.nslots)
    k_scatter<6><<<nblk(L.csr.nslots), kBlock, 0, stream>>>(L.csr.nslots, L.csr.src.as<int64_t>(),
                                                             L.pairval.as<double>(), L.csr.val.as<double>());
  ST_HIP(hipGetLastError());
  for (DBuf* b : {&L.b, &L.x, &L.r, &L.x2}) b->ensure(std::max<size_t>(6 * sizeof(double) * n, 1));
  L.step = power_step(6, n, L.diag.as<double>(), L.csr, L.inv.as<double>(), L.x, L.r, L.x2);
  ST_HIP(hipStreamSynchronize(stream));
  mark("step");
  for (DBuf* t : {&dip, &dil, &dsp, &dsl}) t->release();
  if (n <= kDirect) factor(L);
}

// Coarse::factor (solve.rs:343-372), on the host with the reference's arithmetic.
void Solver::factor(Level& L) {
  const int64_t n = L.n, m = 6 * n;
  if (n > kMaxTop)
    throw StError{ST_BAD_INPUT, "solver: " + std::to_string(n) + " parts left apart at the top (at most " +
                                    std::to_string(kMaxTop) + " are solved directly here)"};
  std::vector<double> diag(36 * n), pv(36 * L.npairs);
  ST_HIP(hipStreamSynchronize(stream));
  if (n) ST_HIP(hipMemcpy(diag.data(), L.diag.p, diag.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (L.npairs) ST_HIP(hipMemcpy(pv.data(), L.pairval.p, pv.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> a((size_t)m * m, 0.0);
  for (int64_t p = 0; p < n; p++) {
    const double* d = diag.data() + 36 * p;
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) a[(6 * p + r) * m + 6 * p + c] += d[r * 6 + c];
    const double t = (d[0] + d[7] + d[14]) / 3.0;
    for (int k = 3; k < 6; k++) a[(6 * p + k) * m + 6 * p + k] += 1e-6 * t * L.size * L.size;
  }
  for (int64_t k = 0; k < L.npairs; k++) {
    const int64_t p = L.ij[k][0], q = L.ij[k][1];
    const double* b = pv.data() + 36 * k;
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) {
        a[(6 * p + r) * m + 6 * q + c] += b[r * 6 + c];
        a[(6 * q + c) * m + 6 * p + r] += b[r * 6 + c];
      }
  }
  double largest = 0.0;
  for (int64_t i = 0; i < m; i++) largest = std::fmax(largest, a[i * m + i]);
  std::vector<uint8_t> dropped(std::max<int64_t>(m, 1), 0);
  for (int64_t j = 0; j < m; j++) {
    double d = a[j * m + j];
    for (int64_t k = 0; k < j; k++) d -= a[j * m + k] * a[j * m + k];
    if (d <= largest * 1e-12) {
      dropped[j] = 1;
      for (int64_t i = j; i < m; i++) a[i * m + j] = 0.0;
      continue;
    }
    d = std::sqrt(d);
    a[j * m + j] = d;
    for (int64_t i = j + 1; i < m; i++) {
      double s = a[i * m + j];
      for (int64_t k = 0; k < j; k++) s -= a[i * m + k] * a[j * m + k];
      a[i * m + j] = s / d;
    }
  }
  up(L.fact, a, stream);
  up(L.dropped, dropped, stream);
  ST_HIP(hipStreamSynchronize(stream));
  L.direct = true;
}

// Solver::new (solve.rs:387-420).
void Solver::build(const MeSystem& s, double size) {
  const auto t0 = std::chrono::steady_clock::now();
  const bool verbose = std::getenv("STATICS_HIP_VERBOSE") != nullptr;
  auto mark = [&](const char* what) {
    if (!verbose) return;
    ST_HIP(hipStreamSynchronize(stream));
    std::fprintf(stderr, "mechanics_hip: setup %s at %.1f ms\n", what,
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  };
  n = s.n;
  upload_system(s, true);
  mark("system uploaded");
  inv0.ensure(std::max<size_t>(9 * sizeof(double) * n, 1));
  for (DBuf* b : {&x, &r, &z, &p, &ap, &bvec, &t1, &t2}) b->ensure(std::max<size_t>(3 * sizeof(double) * n, 1));
  part.ensure(kDotBlocks * sizeof(double));
  scal.ensure(kScal * sizeof(double));
  if (!hscal) ST_HIP(hipHostMalloc((void**)&hscal, kScal * sizeof(double)));
  if (n) {
    k_inverse<3><<<nblk(n, 64), 64, 0, stream>>>(n, diag0.as<double>(), 0.0, inv0.as<double>());
    ST_HIP(hipGetLastError());
  }
  step0 = power_step(3, n, diag0.as<double>(), csr0, inv0.as<double>(), t1, t2, z);
  mark("level 0 inverses and step");
  for (Level& L : levels) release_level(L);
  levels.clear();
  std::vector<Key> keys(n);
  std::vector<P3> at(n);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      keys[i][k] = s.sites ? s.sites[3 * i + k] : 0;
      at[i][k] = s.at ? s.at[3 * i + k] : 0.0;
    }
  double sz = size;
  levels.emplace_back();
  coarsen(keys, at, pattern, diag0.as<double>(), blocks0.as<double>(), ij0.as<uint32_t>(), 3, sz * 2.0,
          levels.back());
  mark("level 1");
  for (;;) {
    sz *= 2.0;
    const int64_t nl = levels.back().n, before = levels.size() == 1 ? n : levels[levels.size() - 2].n;
    if (levels.back().direct) break;
    if (nl == before) {
      factor(levels.back());
      break;
    }
    Level& l = levels.back();
    std::vector<P3> centre = l.centre;
    std::vector<Key> lkeys = l.keys;
    std::vector<std::array<uint32_t, 2>> lij = l.ij;
    levels.emplace_back();
    Level& prev = levels[levels.size() - 2];
    coarsen(lkeys, centre, lij, prev.diag.as<double>(), prev.pairval.as<double>(), prev.pij.as<uint32_t>(), 6,
            sz * 2.0, levels.back());
    mark("a level");
  }
  ST_HIP(hipStreamSynchronize(stream));
  setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Solver::cycle (solve.rs:423-469): the correction from level l for residual b (device).
double* Solver::cycle(size_t l, double* b) {
  Level& L = levels[l];
  const int64_t nl = L.n;
  if (L.direct) {
    if (b != L.x.as<double>())
      ST_HIP(hipMemcpyAsync(L.x.p, b, 6 * sizeof(double) * nl, hipMemcpyDeviceToDevice, stream));
    if (nl) {
      k_dense_solve<<<1, 64, 6 * nl * sizeof(double), stream>>>((int)(6 * nl), L.fact.as<double>(),
                                                                 L.dropped.as<uint8_t>(), L.x.as<double>());
      ST_HIP(hipGetLastError());
    }
    return L.x.as<double>();
  }
  Level& U = levels[l + 1];
  double* x = L.x.as<double>();
  k_jacobi_set<6><<<nblk(nl), kBlock, 0, stream>>>(nl, L.inv.as<double>(), L.step, b, x);
  k_spmv<6, 1><<<nblk(nl * kRowLanes), kBlock, 0, stream>>>(nl, L.diag.as<double>(), L.csr.rowptr.as<int64_t>(),
                                                 L.csr.col.as<int32_t>(), L.csr.val.as<double>(), x, b, nullptr, 0.0,
                                                 L.r.as<double>());
  k_restrict<6><<<nblk(U.n), kBlock, 0, stream>>>(U.n, U.mptr.as<int64_t>(), U.mlist.as<int32_t>(),
                                                   U.roff.as<double>(), L.r.as<double>(), U.b.as<double>());
  ST_HIP(hipGetLastError());
  const double* e = cycle(l + 1, U.b.as<double>());
  k_prolong<6><<<nblk(nl), kBlock, 0, stream>>>(nl, U.of.as<uint32_t>(), U.roff.as<double>(), e, x);
  k_spmv<6, 2><<<nblk(nl * kRowLanes), kBlock, 0, stream>>>(nl, L.diag.as<double>(), L.csr.rowptr.as<int64_t>(),
                                                 L.csr.col.as<int32_t>(), L.csr.val.as<double>(), x, b,
                                                 L.inv.as<double>(), L.step, L.x2.as<double>());
  ST_HIP(hipGetLastError());
  return L.x2.as<double>();
}

// Solver::precondition (solve.rs:472-497): out = P b.
void Solver::precondition(const double* b, double* out) {
  Level& U = levels[0];
  k_jacobi_set<3><<<nblk(n), kBlock, 0, stream>>>(n, inv0.as<double>(), step0, b, t1.as<double>());
  k_spmv<3, 1><<<nblk(n * kRowLanes), kBlock, 0, stream>>>(n, diag0.as<double>(), csr0.rowptr.as<int64_t>(),
                                                csr0.col.as<int32_t>(), csr0.val.as<double>(), t1.as<double>(), b,
                                                nullptr, 0.0, t2.as<double>());
  k_restrict<3><<<nblk(U.n), kBlock, 0, stream>>>(U.n, U.mptr.as<int64_t>(), U.mlist.as<int32_t>(),
                                                   U.roff.as<double>(), t2.as<double>(), U.b.as<double>());
  ST_HIP(hipGetLastError());
  const double* e = cycle(0, U.b.as<double>());
  k_prolong<3><<<nblk(n), kBlock, 0, stream>>>(n, U.of.as<uint32_t>(), U.roff.as<double>(), e, t1.as<double>());
  k_spmv<3, 2><<<nblk(n * kRowLanes), kBlock, 0, stream>>>(n, diag0.as<double>(), csr0.rowptr.as<int64_t>(),
                                                csr0.col.as<int32_t>(), csr0.val.as<double>(), t1.as<double>(), b,
                                                inv0.as<double>(), step0, out);
  ST_HIP(hipGetLastError());
}

// Solver::solve (solve.rs:508-537); -1 for None.
int32_t Solver::solve(const MeSystem& s, const double* bh, double* xh, double tol) {
  const auto t0 = std::chrono::steady_clock::now();
  if (s.n == 0) return 0;
  upload_system(s, false);
  const int64_t n3 = 3 * n;
  ST_HIP(hipMemcpyAsync(bvec.p, bh, n3 * sizeof(double), hipMemcpyHostToDevice, stream));
  ST_HIP(hipMemcpyAsync(x.p, xh, n3 * sizeof(double), hipMemcpyHostToDevice, stream));
  double* b = bvec.as<double>();
  auto finish = [&](int32_t rounds) {
    ST_HIP(hipMemcpyAsync(xh, x.p, n3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    ST_HIP(hipStreamSynchronize(stream));
    timing_collect();
    solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rounds;
  };
  precondition(b, z.as<double>());
  const double target = tol * tol * dot(b, z.as<double>(), n3, TARGET, true);
  if (std::isnan(target) || target < 0.0) return finish(-1);
  // r = b - K x
  k_spmv<3, 1><<<nblk(n * kRowLanes), kBlock, 0, stream>>>(n, diag0.as<double>(), csr0.rowptr.as<int64_t>(),
                                                csr0.col.as<int32_t>(), csr0.val.as<double>(), x.as<double>(), b,
                                                nullptr, 0.0, r.as<double>());
  ST_HIP(hipGetLastError());
  precondition(r.as<double>(), z.as<double>());
  ST_HIP(hipMemcpyAsync(p.p, z.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, stream));
  double rz = dot(r.as<double>(), z.as<double>(), n3, RZ, true);
  for (int round = 0; round < kRounds; round++) {
    if (rz < 0.0) return finish(-1);
    if (rz <= target) return finish(round);
    if (timing) {
      while (tev.size() < 2 * tev_used + 2) {
        hipEvent_t e;
        ST_HIP(hipEventCreate(&e));
        tev.push_back(e);
      }
      ST_HIP(hipEventRecord(tev[2 * tev_used], stream));
    }
    k_spmv<3, 0><<<nblk(n * kRowLanes), kBlock, 0, stream>>>(n, diag0.as<double>(), csr0.rowptr.as<int64_t>(),
                                                  csr0.col.as<int32_t>(), csr0.val.as<double>(), p.as<double>(),
                                                  nullptr, nullptr, 0.0, ap.as<double>());
    ST_HIP(hipGetLastError());
    if (timing) ST_HIP(hipEventRecord(tev[2 * tev_used++ + 1], stream));
    dot(p.as<double>(), ap.as<double>(), n3, PAP, false);
    k_cg_update<<<nblk(n3), kBlock, 0, stream>>>(n3, p.as<double>(), ap.as<double>(), x.as<double>(),
                                                  r.as<double>(), scal.as<double>());
    ST_HIP(hipGetLastError());
    precondition(r.as<double>(), z.as<double>());
    dot(r.as<double>(), z.as<double>(), n3, NEXT, false);
    k_cg_p<<<nblk(n3), kBlock, 0, stream>>>(n3, z.as<double>(), p.as<double>(), scal.as<double>());
    k_cg_shift<<<1, 1, 0, stream>>>(scal.as<double>());
    ST_HIP(hipGetLastError());
    ST_HIP(hipMemcpyAsync(hscal, scal.p, kScal * sizeof(double), hipMemcpyDeviceToHost, stream));
    ST_HIP(hipStreamSynchronize(stream));
    if (!(hscal[PAP] > 0.0)) return finish(round);  // x and r were left as they were
    rz = hscal[RZ];
  }
  return finish(kRounds);
}

}  // namespace me

struct MeSolver {
  me::Solver s;
};

namespace {
template <class F>
int guard(MeSolver* h, F f) {
  if (!h) return ST_BAD_INPUT;
  try {
    ST_HIP(hipSetDevice(h->s.device));
    f();
    h->s.err.clear();
    return ST_OK;
  } catch (const st::StError& e) {
    h->s.err = e.what;
    return e.code;
  } catch (const std::bad_alloc&) {
    h->s.err = "host allocation failed";
    return ST_OOM;
  }
}
}  // namespace

extern "C" int me_solver_new(MeSolver** out, const MeSystem* system, double size, int device) {
  if (!out || !system) return ST_BAD_INPUT;
  *out = nullptr;
  std::unique_ptr<MeSolver> h(new MeSolver());
  h->s.device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return ST_DEVICE;
  const int rc = guard(h.get(), [&] {
    if (!(size > 0) || !std::isfinite(size)) throw st::StError{ST_BAD_INPUT, "solver: size must be positive"};
    me::check_system(*system, -1);
    ST_HIP(hipStreamCreateWithFlags(&h->s.stream, hipStreamNonBlocking));
    h->s.build(*system, size);
  });
  if (rc != ST_OK) {
    std::fprintf(stderr, "me_solver_new: %s\n", h->s.err.c_str());
    return rc;
  }
  *out = h.release();
  return ST_OK;
}

extern "C" int me_solve(MeSolver* s, const MeSystem* system, const double* b, double* x, double tolerance,
                        int32_t* rounds) {
  return guard(s, [&] {
    if (!system || (system->n && (!b || !x))) throw st::StError{ST_BAD_INPUT, "solve: system, b or x is NULL"};
    me::check_system(*system, s->s.n);
    const int32_t r = s->s.solve(*system, b, x, tolerance);
    if (rounds) *rounds = r;
  });
}

extern "C" int me_solver_free(MeSolver* s) {
  if (!s) return ST_OK;
  (void)hipSetDevice(s->s.device);
  delete s;
  return ST_OK;
}

extern "C" const char* me_last_error(const MeSolver* s) { return s ? s->s.err.c_str() : "no solver"; }

extern "C" int me_solver_info(const MeSolver* s, int32_t cap, int32_t* count, int64_t* parts, int64_t* pairs,
                              double* steps, int32_t* direct, double* step0, double* setup_ms, double* solve_ms) {
  if (!s) return ST_BAD_INPUT;
  const auto& L = s->s.levels;
  if (count) *count = (int32_t)L.size();
  for (int32_t l = 0; l < cap && l < (int32_t)L.size(); l++) {
    if (parts) parts[l] = L[l].n;
    if (pairs) pairs[l] = L[l].npairs;
    if (steps) steps[l] = L[l].step;
    if (direct) direct[l] = L[l].direct ? 1 : 0;
  }
  if (step0) *step0 = s->s.step0;
  if (setup_ms) *setup_ms = s->s.setup_ms;
  if (solve_ms) *solve_ms = s->s.solve_ms;
  return ST_OK;
}

extern "C" int me_kernel_timing(MeSolver* s, int enable, int64_t* launches, double* total_ms,
                                double* bytes_per_launch) {
  return guard(s, [&] {
    me::Solver& v = s->s;
    if (enable > 0) {
      v.timing = true;
      v.t_launches = 0;
      v.t_ms = 0.0;
    } else if (enable == 0) {
      v.timing = false;
    }
    if (launches) *launches = v.t_launches;
    if (total_ms) *total_ms = v.t_ms;
    // read: the diagonal blocks, row pointers, column indices and blocks once, x once, y
    // once; written: y
    if (bytes_per_launch)
      *bytes_per_launch = (double)v.n * (72.0 + 8.0 + 24.0 + 24.0) + 8.0 + (double)v.csr0.nslots * (4.0 + 72.0);
  });
}
