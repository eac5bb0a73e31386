// The solve: conjugate gradients preconditioned by an octree multigrid
// V-cycle, all on the device (statics/src/bodies.rs:338-403, 429-468).
//
// Hierarchy. Level 1's aggregates are the connected pieces of the free nodes
// keyed in each 2^3 cube of cells, level l's those of level l-1's aggregates
// in each 2^l cube: "bodies within bodies" as in bodies_of
// (statics/src/bodies.rs:472-496), here with every node keyed by the free
// cell that keeps it. Pieces are found by a lock-free union-find over the
// fine nodes whose edges are the live beams inside each cube. Coarsening goes
// on until a cube holds the whole structure, where the aggregates are the
// structure's connected pieces and the top operator is block diagonal, so the
// top solve is exact by 6x6 inverses.
//
// Transfer. Each aggregate moves rigidly about its centre (the reference's
// P0 = shift, statics/src/bodies.rs:44-47, 239); the coarse operator is the
// Galerkin P0^T K P0. Level 1's is assembled directly from the beams that
// cross aggregates (a beam inside an aggregate moves rigidly and adds nothing),
// coarser ones from the level below. Unlike the reference the prolongator is
// not smoothed: the reference's smoothing step widens every coarse stencil
// (about 20 blocks a row at level 1, 40 at level 2) and would need a sparse
// triple product per level on the device.
//
// Smoother: damped block Jacobi with the reference's step 4/3 over the
// largest eigenvalue of D^-1 A (statics/src/bodies.rs:172-190), estimated by
// power iteration per level and structure.
//
// Precision: ST_FP64 runs everything in FP64. ST_MIXED keeps the CG (vectors,
// K p, inner products, step lengths) in FP64 and runs the V-cycle in FP32 on the
// fine level and the first four coarse levels, FP64 above and in the dense top. In both, when the
// recursive residual meets the tolerance the true FP64 residual f - K x replaces
// it and the iteration goes on from it if it falls short (unless the caller asks
// for the reference's rule, where the recursive residual decides).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "st_solver.hpp"

namespace st {

namespace {

constexpr int kScal = 16;  // doubles of CG state per structure
enum { S_RZ = 0, S_RR, S_PAP, S_ALPHA, S_BETA, S_GOAL2, S_FF, S_RZNEW, S_RRTRUE, S_RRLAST, S_LAM_A, S_LAM_B, S_BEST, S_STALL };
constexpr int kMaxJ = 96;  // distinct neighbours of one aggregate

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

// --------------------------------------------------------------- geometry
// Rigid motion y about a centre seen at offset r: [u + w x r; w] (statics/src/bodies.rs:44-47)
template <class T>
__device__ __forceinline__ void shift_add(const T* y, const double* r, T* out) {
  const T r0 = (T)r[0], r1 = (T)r[1], r2 = (T)r[2];
  out[0] += y[0] + (y[4] * r2 - y[5] * r1);
  out[1] += y[1] + (y[5] * r0 - y[3] * r2);
  out[2] += y[2] + (y[3] * r1 - y[4] * r0);
  out[3] += y[3];
  out[4] += y[4];
  out[5] += y[5];
}
// Inverse of a symmetric positive definite 6x6 block (FP64, row-major) by its Cholesky
// factor, A^-1 = L^-T L^-1, every loop unrolled so the block stays in registers. Returns
// false when a pivot is not positive (the block is not SPD).
__device__ bool invert6(const double* a_in, double* inv) {
  double L[6][6];
#pragma unroll
  for (int j = 0; j < 6; j++) {
    double d = a_in[6 * j + j];
#pragma unroll
    for (int k = 0; k < j; k++) d -= L[j][k] * L[j][k];
    if (!(d > 0.0)) return false;
    L[j][j] = sqrt(d);
    const double r = 1.0 / L[j][j];
#pragma unroll
    for (int i = j + 1; i < 6; i++) {
      double t = 0.5 * (a_in[6 * i + j] + a_in[6 * j + i]);
#pragma unroll
      for (int k = 0; k < j; k++) t -= L[i][k] * L[j][k];
      L[i][j] = t * r;
    }
  }
  double M[6][6];  // L^-1, lower triangular
#pragma unroll
  for (int j = 0; j < 6; j++) {
    M[j][j] = 1.0 / L[j][j];
#pragma unroll
    for (int i = j + 1; i < 6; i++) {
      double t = 0.0;
#pragma unroll
      for (int k = j; k < i; k++) t -= L[i][k] * M[k][j];
      M[i][j] = t / L[i][i];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = i; j < 6; j++) {
      double t = 0.0;
#pragma unroll
      for (int k = j; k < 6; k++) t += M[k][i] * M[k][j];
      inv[6 * i + j] = t;
      inv[6 * j + i] = t;
    }
  return true;
}

// The beam geometry of pair q at spacing geo: n (unit), class.
__device__ __forceinline__ void pair_n(int q, const BeamGeo& G, double* n, int& cls) {
  cls = pair_opposite(q) ? 0 : 1;
  const double ns = G.nscale[cls];
  for (int k = 0; k < 3; k++) n[k] = ns * (double)pair_dir(q, k);
}

// Beam q of cell c: its section over SECTION s^2 (1 when the cells gave none).
__device__ __forceinline__ double beam_sec(const double* csec, int64_t c, int q) {
  return csec ? csec[15 * c + q] : 1.0;
}

// Own block of one beam end (6x6, row-major) added into D: sgn -1 for the beam's i end (face a), +1 for j.
// rho: the beam's section over SECTION s^2.
__device__ void beam_own_add(const BeamGeo& G, int q, int sgn, double e, double g, double rho, double* D) {
  double n[3];
  int cls;
  pair_n(q, G, n, cls);
  const double K4 = G.K4[cls] * rho, R2 = G.R2[cls] * (rho * rho), hd = G.hd[cls];
  const double kg = K4 * g, ke = K4 * (e - g);
  double nx[9] = {0, -n[2], n[1], n[2], 0, -n[0], -n[1], n[0], 0};  // [n]x
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      const double nn = n[r] * n[c], id = r == c ? 1.0 : 0.0;
      D[6 * r + c] += kg * id + ke * nn;                          // uu
      const double uw = sgn * hd * kg * nx[3 * r + c];            // uw = sgn hd kg [n]x
      D[6 * r + 3 + c] += uw;
      D[6 * (3 + c) + r] += uw;                                   // wu = uw^T
      D[6 * (3 + r) + 3 + c] += hd * hd * kg * (id - nn) + R2 * (e * id + (g + g - e) * nn);
    }
}

// G_v = B_v Sh(r_v) = [sgn I, hd[n]x - sgn [r]x]   (3x6)
__device__ __forceinline__ void gmat(int sgn, const double* n, double hd, const double* r, double* Gm) {
  const double nx[9] = {0, -n[2], n[1], n[2], 0, -n[0], -n[1], n[0], 0};
  const double rx[9] = {0, -r[2], r[1], r[2], 0, -r[0], -r[1], r[0], 0};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      Gm[6 * a + b] = a == b ? (double)sgn : 0.0;
      Gm[6 * a + 3 + b] = hd * nx[3 * a + b] - sgn * rx[3 * a + b];
    }
}

// out += Gv^T Sf Gw + sv*sw*diag(0, Rm), Sf = K4(gI + (e-g)nn^T), Rm = R2(eI + (2g-e)nn^T)
__device__ void beam_coarse_add(const double* Gv, const double* Gw, int sv, int sw, const double* n, double K4,
                                double R2, double e, double g, double* out) {
  double Sf[9], Rm[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      const double nn = n[r] * n[c], id = r == c ? 1.0 : 0.0;
      Sf[3 * r + c] = K4 * (g * id + (e - g) * nn);
      Rm[3 * r + c] = R2 * (e * id + (g + g - e) * nn);
    }
  double SG[18];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 6; c++) {
      double s = 0;
      for (int k = 0; k < 3; k++) s += Sf[3 * r + k] * Gw[6 * k + c];
      SG[6 * r + c] = s;
    }
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) {
      double s = 0;
      for (int k = 0; k < 3; k++) s += Gv[6 * k + r] * SG[6 * k + c];
      out[6 * r + c] += s;
    }
  const double ss = (double)(sv * sw);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) out[6 * (3 + r) + 3 + c] += ss * Rm[3 * r + c];
}

// Row vector pc^T K_b for the beam block K_b = Gv^T Sf Gw + sv sw diag(0, Rm) of
// beam_coarse_add at zero offsets (Gv = [sv I, hd [n]x], Gw = [sw I, hd [n]x]):
// a = Gv pc, b = Sf a, then pc^T K_b = [sw b, -hd n x b + sv sw Rm pc_w].
__device__ __forceinline__ void beam_row(const double* pc, int sv, int sw, const double* n, double hd, double K4,
                                         double R2, double e, double g, double* qr) {
  double a[3];
  a[0] = sv * pc[0] + hd * (n[1] * pc[5] - n[2] * pc[4]);
  a[1] = sv * pc[1] + hd * (n[2] * pc[3] - n[0] * pc[5]);
  a[2] = sv * pc[2] + hd * (n[0] * pc[4] - n[1] * pc[3]);
  const double na = n[0] * a[0] + n[1] * a[1] + n[2] * a[2];
  double b[3];
  for (int k = 0; k < 3; k++) b[k] = K4 * (g * a[k] + (e - g) * na * n[k]);
  const double nw = n[0] * pc[3] + n[1] * pc[4] + n[2] * pc[5];
  const double ss = (double)(sv * sw);
  for (int k = 0; k < 3; k++) qr[k] = sw * b[k];
  qr[3] = -hd * (n[1] * b[2] - n[2] * b[1]) + ss * R2 * (e * pc[3] + (g + g - e) * nw * n[0]);
  qr[4] = -hd * (n[2] * b[0] - n[0] * b[2]) + ss * R2 * (e * pc[4] + (g + g - e) * nw * n[1]);
  qr[5] = -hd * (n[0] * b[1] - n[1] * b[0]) + ss * R2 * (e * pc[5] + (g + g - e) * nw * n[2]);
}

// ------------------------------------------------------------- union-find
__device__ __forceinline__ int uf_find(int* p, int x) {
  volatile int* vp = p;
  for (;;) {
    const int px = vp[x];
    if (px == x) return x;
    const int ppx = vp[px];
    if (ppx != px) vp[x] = ppx;  // path halving: only ever points further up
    x = ppx;
  }
}
__device__ __forceinline__ void uf_unite(int* p, int a, int b) {
  for (;;) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&p[a], a, b) == a) return;
  }
}

__global__ void k_iota(int64_t n, int* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

__device__ __forceinline__ bool same_cube(uint32_t a, uint32_t b, int L) {
  return (((a & 1023u) >> L) == ((b & 1023u) >> L)) && ((((a >> 10) & 1023u) >> L) == (((b >> 10) & 1023u) >> L)) &&
         ((((a >> 20) & 1023u) >> L) == (((b >> 20) & 1023u) >> L));
}

// Unites the free ends of every live beam whose ends are keyed in one 2^L cube
// but not in one 2^(L-1) cube (those were united before).
__global__ void k_unite_level(int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask, const uint32_t* nkey, int L,
                              int* uf) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  uint32_t fi[6], key[6];
  for (int f = 0; f < 6; f++) {
    fi[f] = cfaces[6 * c + f];
    key[f] = (fi[f] & kNotFree) ? 0u : nkey[fi[f]];
  }
  const uint32_t m = cmask[c];
  for (int q = 0; q < 15; q++) {
    if (!(m & (1u << q))) continue;
    const int a = kPairA[q], b = kPairB[q];
    if ((fi[a] | fi[b]) & kNotFree) continue;
    if (!same_cube(key[a], key[b], L)) continue;
    if (L > 1 && same_cube(key[a], key[b], L - 1)) continue;
    uf_unite(uf, (int)fi[a], (int)fi[b]);
  }
}

// Every alive beam (any ends): for reachability from held nodes.
__global__ void k_unite_all(int64_t ncells, const uint32_t* cfaces, const uint16_t* mask, int* uf) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const uint32_t m = mask[c];
  for (int q = 0; q < 15; q++)
    if (m & (1u << q))
      uf_unite(uf, (int)(cfaces[6 * c + kPairA[q]] & ~kNotFree), (int)(cfaces[6 * c + kPairB[q]] & ~kNotFree));
}

__global__ void k_compress(int64_t n, int* uf) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) uf[v] = uf_find(uf, (int)v);
}

// flag[v] = 1 for free nodes that are roots
__global__ void k_root_flags(int64_t n, const int* uf, const uint8_t* nflags, int* flag) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  flag[v] = (!(nflags[v] & (kHeld | kGone)) && uf[v] == (int)v) ? 1 : 0;
}

// label[v] = id of v's root (ids = exclusive scan of root flags); rep[id] = root node
__global__ void k_label(int64_t n, const int* uf, const uint8_t* nflags, const int* ids, int* label, int* rep) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (nflags[v] & (kHeld | kGone)) {
    label[v] = -1;
    return;
  }
  const int r = uf[v];
  label[v] = ids[r];
  if (r == (int)v) rep[ids[r]] = (int)v;
}

__global__ void k_held_roots(int64_t n, const int* uf, const uint8_t* nflags, int* flag) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && (nflags[v] & kHeld)) flag[uf[v]] = 1;
}
__global__ void k_joined(int64_t n, const int* uf, const int* flag, uint8_t* joined) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) joined[v] = flag[uf[v]] ? 1 : 0;
}

__global__ void k_fill_int(int64_t n, int* p, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// parent[I] = label[rep[I]]
__global__ void k_parent(int64_t n, const int* rep, const int* label, int* parent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = label[rep[i]];
}

// segment starts of sorted keys: ptr[key[i]] = i for the first of each key; ptr[nseg] = count
__global__ void k_seg_ptr(int64_t count, const int* keys, int* ptr, int nseg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    if (i == 0 || keys[i] != keys[i - 1]) ptr[keys[i]] = (int)i;
  }
  if (i == 0) ptr[nseg] = (int)count;
}

// sort key: label for free nodes, INT_MAX otherwise
__global__ void k_sort_keys(int64_t n, const int* label, int* keys, int* vals) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  keys[v] = label[v] >= 0 ? label[v] : 0x7fffffff;
  vals[v] = (int)v;
}

// ------------------------------------------------------- fine smoother D^-1
// D_v = sum of the own blocks of the live beams at node v (statics/src/bodies.rs:151-153).
__global__ __launch_bounds__(256) void k_dinv(int64_t nnodes, const uint8_t* nflags, const int32_t* ncell, const uint8_t* nkind,
                       const double* ce, const double* cg, const double* csec, const uint16_t* cmask,
                       const uint32_t* cstruct, const BeamGeo* bgeo, double* dinv64, SymHalf* dinv32, int* bad) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nnodes) return;
  double D[36], inv[36];
  for (int i = 0; i < 36; i++) D[i] = 0.0;
  const bool isfree = !(nflags[v] & (kHeld | kGone));
  if (isfree) {
    const int o = nkind[v] >> 1;
    for (int side = 0; side < 2; side++) {
      const int c = ncell[2 * v + side];
      if (c < 0) continue;
      const int f = side == 0 ? 2 * o + 1 : 2 * o;
      const BeamGeo G = bgeo[cstruct[c]];
      const double e = ce[c], g = cg[c];
      const uint32_t m = (uint32_t)cmask[c] & face_beams(f);
      for (int q = 0; q < 15; q++)
        if (m & (1u << q)) beam_own_add(G, q, kPairA[q] == f ? -1 : 1, e, g, beam_sec(csec, c, q), D);
    }
    if (!invert6(D, inv)) {
      atomicExch(bad, 1);
      for (int i = 0; i < 36; i++) inv[i] = 0.0;
    }
  } else {
    for (int i = 0; i < 36; i++) inv[i] = 0.0;
  }
  double m = 0.0;
  for (int i = 0; i < 6; i++)
    for (int j = i; j < 6; j++) {
      // the inverse of a symmetric block, symmetrised against rounding
      const double a = 0.5 * (inv[6 * i + j] + inv[6 * j + i]);
      dinv64[kSym * v + sym_idx(i, j)] = a;
      m = fmax(m, fabs(a));
    }
  // the FP32 V-cycle's copy: SymHalf (FP16 entries over the node's scale)
  const double sc = m > 0.0 ? m / 16384.0 : 1.0;
  const float scf = (float)sc;
  SymHalf h;
  for (int k = 0; k < 22; k++) h.e[k] = (_Float16)0.0f;
  for (int k = 0; k < kSym; k++) h.e[k] = (_Float16)(float)(dinv64[kSym * v + k] / (double)scf);
  h.scale = scf;
  dinv32[v] = h;
}

// ----------------------------------------------------------- brick vector ops
// All fine-level vector kernels run one workgroup per brick so a structure's
// share is a contiguous run of workgroups and per-structure sums are exact
// brick by brick in a fixed order.
template <class T>
__global__ __launch_bounds__(kBlock) void k_jacobi_fine(const uint32_t* nbase, const uint32_t* bstruct, const uint8_t* done,
                                                        const T* dinv, const double* step, const T* r, T* x) {
  const uint32_t b = blockIdx.x, s = bstruct[b];
  if (done && done[s]) return;
  const T w = step ? (T)step[s] : T(1);
  for (uint32_t n = nbase[b] + threadIdx.x; n < nbase[b + 1]; n += blockDim.x) {
    T rv[6], d[6];
    load6<T>(r + 6 * (size_t)n, rv);
    sym6v<T>(dinv, n, rv, d);
    for (int k = 0; k < 6; k++) d[k] *= w;
    store6<T>(x + 6 * (size_t)n, d);
  }
}

// x += a p ; r -= a ap ; partial r.r      (TX may be wider than T)
template <class T>
__global__ __launch_bounds__(kBlock) void k_update_xr(const uint32_t* nbase, const uint32_t* bstruct, const uint8_t* done,
                                                      const double* scal, T* x, T* r, const T* p, const T* ap,
                                                      double* partial, float* r32 = nullptr) {
  __shared__ double red[kBlock / 64];
  const uint32_t b = blockIdx.x, s = bstruct[b];
  if (done[s]) {
    if (threadIdx.x == 0) partial[b] = 0.0;
    return;
  }
  const double a = scal[kScal * s + S_ALPHA];
  double acc = 0.0;
  for (uint32_t i = 6 * nbase[b] + threadIdx.x; i < 6 * nbase[b + 1]; i += blockDim.x) {
    x[i] = (T)((double)x[i] + a * (double)p[i]);
    const T rn = (T)((double)r[i] - a * (double)ap[i]);
    r[i] = rn;
    if (r32) r32[i] = (float)rn;
    acc += (double)rn * (double)rn;
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) partial[b] = t;
}

// p = z + beta p
template <class T, class TZ>
__global__ __launch_bounds__(kBlock) void k_update_p(const uint32_t* nbase, const uint32_t* bstruct, const uint8_t* done,
                                                     const double* scal, const TZ* z, T* p) {
  const uint32_t b = blockIdx.x, s = bstruct[b];
  if (done[s]) return;
  const double be = scal[kScal * s + S_BETA];
  for (uint32_t i = 6 * nbase[b] + threadIdx.x; i < 6 * nbase[b + 1]; i += blockDim.x)
    p[i] = (T)((double)z[i] + be * (double)p[i]);
}


// partial a.b per brick
template <class T>
__global__ __launch_bounds__(kBlock) void k_dot_fine(const uint32_t* nbase, const T* a, const T* bb, double* partial) {
  __shared__ double red[kBlock / 64];
  const uint32_t b = blockIdx.x;
  double acc = 0.0;
  for (uint32_t i = 6 * nbase[b] + threadIdx.x; i < 6 * nbase[b + 1]; i += blockDim.x)
    acc += (double)a[i] * (double)bb[i];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) partial[b] = t;
}

// dst = src (converting), brick-blocked, skipping done structures
template <class TD, class TS>
__global__ __launch_bounds__(kBlock) void k_copy_fine(const uint32_t* nbase, const uint32_t* bstruct, const uint8_t* done,
                                                      const TS* src, TD* dst) {
  const uint32_t b = blockIdx.x, s = bstruct[b];
  if (done && done[s]) return;
  for (uint32_t i = 6 * nbase[b] + threadIdx.x; i < 6 * nbase[b + 1]; i += blockDim.x) dst[i] = (TD)src[i];
}


__global__ __launch_bounds__(kBlock) void k_zero_fine(const uint32_t* nbase, float* v) {
  const uint32_t b = blockIdx.x;
  for (uint32_t i = 6 * nbase[b] + threadIdx.x; i < 6 * nbase[b + 1]; i += blockDim.x) v[i] = 0.0f;
}

// x += (double) dx and dx = 0, for every structure (the FP32 CG's steps into the FP64 solution)
__global__ __launch_bounds__(kBlock) void k_flush_fine(const uint32_t* nbase, float* dx, double* x) {
  const uint32_t b = blockIdx.x;
  for (uint32_t i = 6 * nbase[b] + threadIdx.x; i < 6 * nbase[b + 1]; i += blockDim.x) {
    x[i] += (double)dx[i];
    dx[i] = 0.0f;
  }
}

// zero rows of held and gone nodes (a load on them is not read)
__global__ void k_mask_rows(int64_t n, const uint8_t* nflags, double* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (nflags[i] & (kHeld | kGone))
    for (int k = 0; k < 6; k++) v[6 * i + k] = 0.0;
}

// out[s * stride + slot] = sum of partial over [seg[s], seg[s+1]) (fixed order)
// (kReduceThreads threads a structure, each adding every kReduceThreads-th entry in order,
// four loads in flight: a 256^3 structure has 32,000 bricks, which 256 threads took 54 us
// to read one after another)
constexpr int kReduceThreads = 1024;
__global__ __launch_bounds__(kReduceThreads) void k_reduce_seg(const double* partial, const int64_t* seg, double* out,
                                                               int stride, int slot) {
  __shared__ double red[kReduceThreads / 64];
  const int s = blockIdx.x;
  constexpr int K = kReduceThreads;
  const int64_t e = seg[s + 1];
  double acc = 0.0;
  int64_t i = seg[s] + threadIdx.x;
  for (; i + 3 * K < e; i += 4 * K) {
    const double v0 = partial[i], v1 = partial[i + K], v2 = partial[i + 2 * K], v3 = partial[i + 3 * K];
    acc += v0;
    acc += v1;
    acc += v2;
    acc += v3;
  }
  for (; i < e; i += K) acc += partial[i];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) out[(int64_t)s * stride + slot] = t;
}

// ------------------------------------------------------------- CG scalars
__global__ void k_cg_start(int ns, double* scal, uint8_t* done, int* status, int* iters, double tol, int mixed,
                           uint8_t* want) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  double* S = scal + kScal * s;
  S[S_GOAL2] = tol * tol * S[S_FF];
  S[S_RRLAST] = S[S_RR];
  S[S_BEST] = S[S_RR];
  S[S_STALL] = 0.0;
  iters[s] = 0;
  status[s] = 0;
  want[s] = 0;
  done[s] = (S[S_RR] <= S[S_GOAL2]) ? 1 : 0;
  if (!isfinite(S[S_RR])) {
    done[s] = 1;
    status[s] = ST_NAN;
  }
  (void)mixed;
}
__global__ void k_cg_alpha(int ns, double* scal, uint8_t* done, int* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns || done[s]) return;
  double* S = scal + kScal * s;
  const double pap = S[S_PAP];
  if (!isfinite(pap)) {
    status[s] = ST_NAN;
    done[s] = 1;
    S[S_ALPHA] = 0.0;
  } else if (!(pap > 0.0)) {
    status[s] = ST_NOT_SPD;
    done[s] = 1;
    S[S_ALPHA] = 0.0;
  } else {
    S[S_ALPHA] = S[S_RZ] / pap;
  }
}
// after x, r update. mixed: ask for a true residual instead of stopping
__global__ void k_cg_check(int ns, double* scal, uint8_t* done, int* status, int* iters, int mixed, uint8_t* want,
                           double drop2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns || done[s]) return;
  double* S = scal + kScal * s;
  iters[s] += 1;
  const double rr = S[S_RR];
  if (!isfinite(rr)) {
    status[s] = ST_NAN;
    done[s] = 1;
    return;
  }
  if (mixed) {
    if (rr <= S[S_GOAL2] || rr <= drop2 * S[S_RRLAST]) want[s] = 1;
  } else if (rr <= S[S_GOAL2]) {
    done[s] = 1;
  }
}
// r.z <= 0 while r is not yet small means the preconditioner is not positive on r (it can
// lose definiteness on nearly singular structures): the structure stops as not SPD.
__global__ void k_cg_beta(int ns, double* scal, uint8_t* done, int* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns || done[s]) return;
  double* S = scal + kScal * s;
  const double rzn = S[S_RZNEW];
  if (!(rzn > 0.0) && S[S_RR] > S[S_GOAL2]) {
    status[s] = ST_NOT_SPD;
    done[s] = 1;
    return;
  }
  S[S_BETA] = S[S_RZ] != 0.0 ? rzn / S[S_RZ] : 0.0;
  S[S_RZ] = rzn;
}
// After a true residual: it is the residual from now on. A structure whose true
// residual has not fallen by a tenth over kStall replacements stops (the
// precision's attainable accuracy is reached) and is reported as not converged.
// (stall_limit and improve, the squared fall that counts, are kStall and 0.81 for the FP64
// CG; the FP32 CG gives up sooner, after 3 replacements gaining less than half each.)
constexpr int kStall = 8;
__global__ void k_cg_true(int ns, double* scal, uint8_t* done, int* status, uint8_t* want, int* refreshes,
                          int stall_limit, double improve) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns || done[s]) return;
  double* S = scal + kScal * s;
  refreshes[s] += 1;
  const double rt = S[S_RRTRUE];
  S[S_RR] = rt;
  S[S_RRLAST] = rt;
  want[s] = 0;
  if (!isfinite(rt)) {
    status[s] = ST_NAN;
    done[s] = 1;
  } else if (rt <= S[S_GOAL2]) {
    done[s] = 1;
  } else {
    if (rt < improve * S[S_BEST]) {
      S[S_BEST] = rt;
      S[S_STALL] = 0.0;
    } else {
      S[S_STALL] += 1.0;
      if (S[S_STALL] >= stall_limit) {
        status[s] = ST_NO_CONVERGENCE;
        done[s] = 1;
      }
    }
  }
}

// ST_MIXED32 under the reference's rule: a structure whose recursive residual met the
// tolerance is done; the others take the true residual as their residual and go on.
__global__ void k_cg_true_rec(int ns, double* scal, uint8_t* done, int* status, uint8_t* want, int* refreshes) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns || done[s]) return;
  double* S = scal + kScal * s;
  refreshes[s] += 1;
  const bool met = S[S_RR] <= S[S_GOAL2];
  const double rt = S[S_RRTRUE];
  want[s] = 0;
  if (!isfinite(rt)) {
    status[s] = ST_NAN;
    done[s] = 1;
  } else if (met) {
    done[s] = 1;  // S_RR keeps the recursive residual the rule judged
  } else {
    S[S_RR] = rt;
    S[S_RRLAST] = rt;
  }
}

// ------------------------------------------------------------ level 1 setup
// centre of each level-1 aggregate: mean of its nodes' positions (fixed order)
__global__ void k_centre1(int64_t n1, const int* agg_ptr, const int* nodes, const uint32_t* nkey, const uint8_t* nkind,
                          const uint32_t* nstruct_of_agg, const BeamGeo* bgeo, double* centre) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n1) return;
  const double s = 2.0 * bgeo[nstruct_of_agg[I]].hd[0];
  double c[3] = {0, 0, 0};
  const int b = agg_ptr[I], e = agg_ptr[I + 1];
  for (int i = b; i < e; i++) {
    double p[3];
    node_pos((uint32_t)nodes[i], nkey, nkind, s, p);
    c[0] += p[0];
    c[1] += p[1];
    c[2] += p[2];
  }
  const double inv = 1.0 / (double)(e - b);
  centre[3 * I] = c[0] * inv;
  centre[3 * I + 1] = c[1] * inv;
  centre[3 * I + 2] = c[2] * inv;
}

// structure of each aggregate (from its representative node's first cell)
__global__ void k_agg_struct(int64_t n, const int* rep, const int32_t* ncell, const uint32_t* cstruct, uint32_t* out) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const int v = rep[I];
  int c = ncell[2 * v];
  if (c < 0) c = ncell[2 * v + 1];
  out[I] = cstruct[c];
}

// Level-1 Galerkin rows from the beams crossing aggregates. pass 0: count
// distinct neighbours; pass 1: fill blocks (row's diagonal first).
__global__ void k_a1(int pass, int64_t n1, const int* agg_ptr, const int* nodes, const int* label, const double* centre,
                     const uint32_t* agg_struct, const uint32_t* nkey, const uint8_t* nkind, const int32_t* ncell,
                     const uint32_t* cfaces, const double* ce, const double* cg, const double* csec,
                     const uint16_t* cmask, const BeamGeo* bgeo, int* rowcnt, const int* rowptr, int* col, double* A,
                     int* overflow) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n1) return;
  const BeamGeo G = bgeo[agg_struct[I]];
  const double s = 2.0 * G.hd[0];
  const double cI[3] = {centre[3 * I], centre[3 * I + 1], centre[3 * I + 2]};
  int js[kMaxJ];
  int nj = 0;
  int base = 0, len = 0;
  if (pass == 1) {
    base = rowptr[I];
    len = rowptr[I + 1] - base;
    col[base] = (int)I;
    for (int i = 0; i < 36 * len; i++) A[36 * (int64_t)base + i] = 0.0;
    nj = 1;  // filled by the slots already in col
  }
  for (int idx = agg_ptr[I]; idx < agg_ptr[I + 1]; idx++) {
    const uint32_t v = (uint32_t)nodes[idx];
    const int o = nkind[v] >> 1;
    double pv[3], rv[3];
    node_pos(v, nkey, nkind, s, pv);
    for (int k = 0; k < 3; k++) rv[k] = pv[k] - cI[k];
    for (int side = 0; side < 2; side++) {
      const int c = ncell[2 * v + side];
      if (c < 0) continue;
      const int f = side == 0 ? 2 * o + 1 : 2 * o;
      const uint32_t m = (uint32_t)cmask[c] & face_beams(f);
      if (!m) continue;
      const double e = ce[c], g = cg[c];
      for (int q = 0; q < 15; q++) {
        if (!(m & (1u << q))) continue;
        const int f2 = kPairA[q] == f ? kPairB[q] : kPairA[q];
        const uint32_t wf = cfaces[6 * c + f2];
        const int J = (wf & kNotFree) ? -1 : label[wf];
        if (J == (int)I) continue;
        if (pass == 0) {
          if (J >= 0) {
            bool seen = false;
            for (int t = 0; t < nj; t++) seen |= js[t] == J;
            if (!seen) {
              if (nj < kMaxJ)
                js[nj++] = J;
              else
                atomicExch(overflow, 1);
            }
          }
          continue;
        }
        double n[3];
        int cls;
        pair_n(q, G, n, cls);
        const int sv = kPairA[q] == f ? -1 : 1, sw = -sv;
        const double rho = beam_sec(csec, c, q), K4 = G.K4[cls] * rho, R2 = G.R2[cls] * (rho * rho);
        double Gv[18];
        gmat(sv, n, G.hd[cls], rv, Gv);
        beam_coarse_add(Gv, Gv, sv, sv, n, K4, R2, e, g, A + 36 * (int64_t)base);
        if (J >= 0) {
          double pw[3], rw[3];
          node_pos(wf, nkey, nkind, s, pw);
          for (int k = 0; k < 3; k++) rw[k] = pw[k] - centre[3 * (int64_t)J + k];
          double Gw[18];
          gmat(sw, n, G.hd[cls], rw, Gw);
          int slot = -1;
          for (int t = 1; t < len; t++)
            if (col[base + t] == J) slot = t;
          if (slot < 0) {
            // first time: claim the next free slot (filled in order)
            for (int t = 1; t < len; t++)
              if (col[base + t] == -1) {
                col[base + t] = J;
                slot = t;
                break;
              }
          }
          if (slot > 0) beam_coarse_add(Gv, Gw, sv, sw, n, K4, R2, e, g, A + 36 * ((int64_t)base + slot));
        }
      }
    }
  }
  if (pass == 0) rowcnt[I] = 1 + nj;
}

// Galerkin from the level below: rows of parent J = sum over children i, blocks (i, j):
// Sh(r_i)^T A_ij Sh(r_j) into (J, parent(j)).
__device__ void shT_A_sh(const double* A, const double* ri, const double* rj, double* out) {
  // M = A Sh(rj): columns 3..5 get  -A[:,0:3][rj]x
  double M[36];
  const double rx[9] = {0, -rj[2], rj[1], rj[2], 0, -rj[0], -rj[1], rj[0], 0};
  for (int r = 0; r < 6; r++) {
    for (int c = 0; c < 3; c++) M[6 * r + c] = A[6 * r + c];
    for (int c = 0; c < 3; c++) {
      double s = A[6 * r + 3 + c];
      for (int k = 0; k < 3; k++) s -= A[6 * r + k] * rx[3 * k + c];
      M[6 * r + 3 + c] = s;
    }
  }
  // Sh(ri)^T M: rows 3..5 += [ri]x rows 0..2
  const double qx[9] = {0, -ri[2], ri[1], ri[2], 0, -ri[0], -ri[1], ri[0], 0};
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 6; c++) out[6 * r + c] += M[6 * r + c];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 6; c++) {
      double s = M[6 * (3 + r) + c];
      for (int k = 0; k < 3; k++) s += qx[3 * r + k] * M[6 * k + c];
      out[6 * (3 + r) + c] += s;
    }
}

__global__ void k_galerkin(int pass, int64_t n, const int* kid_ptr, const int* kids, const double* centre,
                           const double* centre_c, const int* parent_c, const int* rowptr_c, const int* col_c,
                           const double* A_c, int* rowcnt, const int* rowptr, int* col, double* A, int* overflow) {
  const int64_t J = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= n) return;
  const double cJ[3] = {centre[3 * J], centre[3 * J + 1], centre[3 * J + 2]};
  int js[kMaxJ];
  int nj = 0;
  int base = 0, len = 0;
  if (pass == 1) {
    base = rowptr[J];
    len = rowptr[J + 1] - base;
    col[base] = (int)J;
    for (int i = 0; i < 36 * len; i++) A[36 * (int64_t)base + i] = 0.0;
  }
  for (int t = kid_ptr[J]; t < kid_ptr[J + 1]; t++) {
    const int i = kids[t];
    double ri[3];
    for (int k = 0; k < 3; k++) ri[k] = centre_c[3 * (int64_t)i + k] - cJ[k];
    for (int bq = rowptr_c[i]; bq < rowptr_c[i + 1]; bq++) {
      const int j = col_c[bq];
      const int P = parent_c[j];
      if (pass == 0) {
        if (P != (int)J) {
          bool seen = false;
          for (int u = 0; u < nj; u++) seen |= js[u] == P;
          if (!seen) {
            if (nj < kMaxJ)
              js[nj++] = P;
            else
              atomicExch(overflow, 1);
          }
        }
        continue;
      }
      double rj[3];
      for (int k = 0; k < 3; k++) rj[k] = centre_c[3 * (int64_t)j + k] - centre[3 * (int64_t)P + k];
      int slot = 0;
      if (P != (int)J) {
        slot = -1;
        for (int u = 1; u < len; u++)
          if (col[base + u] == P) slot = u;
        if (slot < 0)
          for (int u = 1; u < len; u++)
            if (col[base + u] == -1) {
              col[base + u] = P;
              slot = u;
              break;
            }
      }
      if (slot >= 0) shT_A_sh(A_c + 36 * (int64_t)bq, ri, rj, A + 36 * ((int64_t)base + slot));
    }
  }
  if (pass == 0) rowcnt[J] = 1 + nj;
}

// centre of a coarse aggregate: mean of its children's centres (statics/src/bodies.rs:229-237)
__global__ void k_centre_c(int64_t n, const int* kid_ptr, const int* kids, const double* centre_c, double* centre) {
  const int64_t J = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= n) return;
  double c[3] = {0, 0, 0};
  const int b = kid_ptr[J], e = kid_ptr[J + 1];
  for (int t = b; t < e; t++)
    for (int k = 0; k < 3; k++) c[k] += centre_c[3 * (int64_t)kids[t] + k];
  for (int k = 0; k < 3; k++) centre[3 * J + k] = c[k] / (double)(e - b);
}

__global__ void k_coarse_dinv(int64_t n, const int* rowptr, const double* A, double* d64, float* d32, int* bad) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  double inv[36];
  if (!invert6(A + 36 * (int64_t)rowptr[I], inv)) {
    atomicExch(bad, 1);
    for (int i = 0; i < 36; i++) inv[i] = 0.0;
  }
  for (int i = 0; i < 36; i++) {
    d64[36 * I + i] = inv[i];
    d32[36 * I + i] = (float)inv[i];
  }
}

// FP16 copy of a level's blocks, each scaled so that its largest entry is 2^14.
// FP16 copy of a level's blocks (see HalfBlock): each block over its own scale, its
// largest entry mapped to 2^14
__global__ void k_to_half_blocks(int64_t nb, const double* a, const int* col, HalfBlock* h) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double m = 0.0;
  for (int k = 0; k < 36; k++) m = fmax(m, fabs(a[36 * b + k]));
  const double sc = m > 0.0 ? m / 16384.0 : 1.0;
  const float scf = (float)sc;
  HalfBlock out;
  for (int k = 0; k < 36; k++) out.a[k] = (_Float16)(float)(a[36 * b + k] / (double)scf);
  out.scale = scf;
  out.col = col[b];
  h[b] = out;
}

__global__ void k_to_f32(int64_t n, const double* a, float* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_fill_neg(int64_t n, int* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = -1;
}

// --------------------------------------------------------- coarse vector ops
// Every coarse kernel: one thread per aggregate, skipping structures that are done.
template <class T>
__global__ void k_c_jacobi(int64_t n, const uint32_t* ast, const uint8_t* done, const T* dinv, const double* step,
                           const T* r, T* x) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const uint32_t s = ast[I];
  if (done && done[s]) return;
  T rv[6], d[6];
  load6<T>(r + 6 * I, rv);
  mat6v<T>(dinv + 36 * I, rv, d);
  const T w = step ? (T)step[s] : T(1);
  for (int k = 0; k < 6; k++) d[k] *= w;
  store6<T>(x + 6 * I, d);
}

// t = r - A x  (mode 0)   or   z = x + step D^-1 (r - A x)  (mode 1)
template <class A, class B>
__global__ void k_convert(int64_t n, const A* a, B* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}

// y = a y + x  (a = 0: copy)
template <class T>
__global__ void k_axpy(int64_t n, T a, const T* x, T* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a == T(0) ? x[i] : a * y[i] + x[i];
}

// ---- dense top: a coarse level's operator per structure, inverted by Gauss-Jordan
// (SPD: no pivoting), symmetrised, and applied as a dense product.
// Scatter the level's blocks into each structure's dense matrix (zeroed before).
__global__ void k_dense_fill(int64_t n, const uint32_t* ast, const int64_t* seg, const int64_t* off,
                             const int* rowptr, const int* col, const double* A, double* D) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const uint32_t s = ast[I];
  const int64_t a0 = seg[s], m = 6 * (seg[s + 1] - a0);
  double* Ds = D + off[s];
  for (int b = rowptr[I]; b < rowptr[I + 1]; b++) {
    const int64_t J = col[b];
    for (int r = 0; r < 6; r++)
      for (int cc = 0; cc < 6; cc++) Ds[(6 * (I - a0) + r) * m + 6 * (J - a0) + cc] = A[36 * (int64_t)b + 6 * r + cc];
  }
}
// The Gauss-Jordan kernels run over the structures still active at step k (those with
// more than k rows): ord lists the structures by size, largest first; pre1/pre2 are the
// running sums of m and m^2 in that order, so thread t finds its structure by bisection.
__device__ __forceinline__ int seg_of(const int64_t* pre, int n, int64_t t) {
  int lo = 0, hi = n - 1;  // the last i with pre[i] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}
// Step k, first half: save column k and row k (row k with a 1 at k), zero column k.
__global__ void k_gj_save(int k, int nact, const int* ord, const int64_t* pre1, const int64_t* seg, const int64_t* off,
                          double* D, double* colk, double* rowk, int mmax) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pre1[nact]) return;
  const int i = seg_of(pre1, nact, t);
  const int s = ord[i];
  const int64_t m = 6 * (seg[s + 1] - seg[s]), r = t - pre1[i];
  double* Ds = D + off[s];
  colk[(int64_t)s * mmax + r] = Ds[r * m + k];
  rowk[(int64_t)s * mmax + r] = r == k ? 1.0 : Ds[(int64_t)k * m + r];
  if (r != k) Ds[r * m + k] = 0.0;
}
// Step k, second half: row k /= pivot; every other row -= its column-k entry times row k.
__global__ void k_gj_update(int k, int nact, const int* ord, const int64_t* pre2, const int64_t* seg,
                            const int64_t* off, double* D, const double* colk, const double* rowk, int mmax) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pre2[nact]) return;
  const int i0 = seg_of(pre2, nact, t);
  const int s = ord[i0];
  const int64_t m = 6 * (seg[s + 1] - seg[s]), e = t - pre2[i0];
  const int64_t i = e / m, j = e % m;
  const double p = 1.0 / colk[(int64_t)s * mmax + k];
  double* Ds = D + off[s];
  const double rk = rowk[(int64_t)s * mmax + j] * p;
  if (i == k)
    Ds[e] = rk;
  else
    Ds[e] -= colk[(int64_t)s * mmax + i] * rk;
}
// (D + D^T) / 2 in place, every structure
__global__ void k_dense_sym(int nact, const int* ord, const int64_t* pre2, const int64_t* seg, const int64_t* off,
                            double* D) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pre2[nact]) return;
  const int i0 = seg_of(pre2, nact, t);
  const int s = ord[i0];
  const int64_t m = 6 * (seg[s + 1] - seg[s]), e = t - pre2[i0];
  const int64_t i = e / m, j = e % m;
  if (j <= i) return;
  double* Ds = D + off[s];
  const double a = 0.5 * (Ds[i * m + j] + Ds[j * m + i]);
  Ds[i * m + j] = a;
  Ds[j * m + i] = a;
}
// x = D r per structure (D symmetric: row i read as column i, contiguous). One wave per
// row: its 64 lanes take every 64th j, then meet by a butterfly (a fixed order), so the
// rows of a 1536-row top fill 384 workgroups (a 1024-thread block per 32 rows filled 48).
template <class T>
__global__ __launch_bounds__(256) void k_dense_apply(const int64_t* seg, const int64_t* off, const double* D,
                                                     const uint8_t* done, const T* r, T* x) {
  const int s = blockIdx.y;
  if (done && done[s]) return;
  const int64_t a0 = seg[s], m = 6 * (seg[s + 1] - a0);
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= m) return;  // the whole wave
  const int lane = threadIdx.x & 63;
  const double* Di = D + off[s] + i * m;
  const T* rs = r + 6 * a0;
  double acc = 0.0;
#pragma unroll 4
  for (int64_t j = lane; j < m; j += 64) acc += Di[j] * (double)rs[j];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) x[6 * a0 + i] = (T)acc;
}

// Coarse block rows are worked by groups of 8 lanes: lane i < 6 of a group computes
// component i of its row, reading row i of every 6x6 block (the group reads each
// block as one contiguous 144-byte run); lanes 6 and 7 idle.
template <class T>
__device__ __forceinline__ T row6_dot(const T* __restrict__ m, const T* v) {
  T s = 0;
#pragma unroll
  for (int j = 0; j < 6; j++) s += m[j] * v[j];
  return s;
}
// The six components of the group's vector, gathered from its lanes 0..5.
template <class T>
__device__ __forceinline__ void group_gather6(T a, T* all) {
  const int base = (threadIdx.x & 63) & ~7;
#pragma unroll
  for (int j = 0; j < 6; j++) all[j] = __shfl(a, base + j, 64);
}

// Row i of block row I times x, summed over the row's blocks b0..b1 in order; four
// blocks' loads are issued before their products, to keep more loads in flight.
// six entries of a TX vector into T (widened or narrowed as a separate conversion would)
template <class T, class TX>
__device__ __forceinline__ void load6_as(const TX* p, T* v) {
  if constexpr (std::is_same<T, TX>::value) {
    load6<T>(p, v);
  } else {
    TX w[6];
    load6<TX>(p, w);
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = (T)w[k];
  }
}
template <class T, class TX = T>
__device__ __forceinline__ T block_row_dot(int b0, int b1, const int* col, const T* A, const TX* x, int i) {
  T acc = 0;
  int b = b0;
  for (; b + 4 <= b1; b += 4) {
    int cb[4];
    T xv[4][6], av[4][6];
#pragma unroll
    for (int k = 0; k < 4; k++) cb[k] = col[b + k];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      load6_as<T, TX>(x + 6 * (int64_t)cb[k], xv[k]);
      load6<T>(A + 36 * (int64_t)(b + k) + 6 * i, av[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) acc += row6_dot<T>(av[k], xv[k]);
  }
  for (; b < b1; b++) {
    T xv[6];
    load6_as<T, TX>(x + 6 * (int64_t)col[b], xv);
    acc += row6_dot<T>(A + 36 * (int64_t)b + 6 * i, xv);
  }
  return acc;
}

// Coarse residual and smoothing, lanes mapped onto each block's 36 entries: a row is
// worked by L = 36 / E lanes (9 in FP32, 18 in FP64), lane i loading entries
// E i .. E i + E - 1 of every block of the row as one 16-byte load, so a group reads
// each block as one contiguous 144 or 288-byte run. Each lane keeps its entries' sums
// over the row's blocks; at the row's end the per-entry sums meet in LDS and are
// added in a fixed order. 64 / L rows per wave, four waves per block.
// mode 0: out = r - A x; mode 1: out = x + w D^-1 (r - A x)
template <class T>
struct CoarseLanes {
  static constexpr int E = 16 / (int)sizeof(T);  // entries per lane
  static constexpr int L = 36 / E;               // lanes per row
  static constexpr int R = 64 / L;               // rows per wave
  static constexpr int kRowsPerBlock = 4 * R;
};
template <class T>
__device__ __forceinline__ void load_entries(const T* p, T* v);
template <>
__device__ __forceinline__ void load_entries<float>(const float* p, float* v) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
template <>
__device__ __forceinline__ void load_entries<double>(const double* p, double* v) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  v[0] = a.x; v[1] = a.y;
}
// The same loads marked non-temporal (the blocks stream through once per product; the
// gathered x entries are what the caches should keep).
typedef float st_f4 __attribute__((ext_vector_type(4)));
typedef double st_d2 __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ void load_entries_nt(const T* p, T* v);
template <>
__device__ __forceinline__ void load_entries_nt<float>(const float* p, float* v) {
  const st_f4 a = __builtin_nontemporal_load(reinterpret_cast<const st_f4*>(p));
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
template <>
__device__ __forceinline__ void load_entries_nt<double>(const double* p, double* v) {
  const st_d2 a = __builtin_nontemporal_load(reinterpret_cast<const st_d2*>(p));
  v[0] = a.x; v[1] = a.y;
}
// Entries E i .. E i + E - 1 of block b.
template <class T, int kNT>
__device__ __forceinline__ void block_entries(const T* A, int64_t b, int i, T* v) {
  constexpr int E = CoarseLanes<T>::E;
  if (kNT) {
    load_entries_nt<T>(A + 36 * b + E * i, v);
  } else {
    load_entries<T>(A + 36 * b + E * i, v);
  }
}
// the x components lane entries E i .. E i + E - 1 multiply: columns c0, c0 + 1, ... mod 6
template <class T>
__device__ __forceinline__ void load_columns(const T* xr, int c0, T* v);
template <>
__device__ __forceinline__ void load_columns<float>(const float* xr, int c0, float* v) {
  const float2 a = *reinterpret_cast<const float2*>(xr + c0);
  const float2 b = *reinterpret_cast<const float2*>(xr + (c0 + 2) % 6);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <>
__device__ __forceinline__ void load_columns<double>(const double* xr, int c0, double* v) {
  const double2 a = *reinterpret_cast<const double2*>(xr + c0);
  v[0] = a.x; v[1] = a.y;
}

// kWide = 0 ("narrow"): R rows per wave, one group of L lanes per row walking the row's
// blocks in order. kWide = 1: one row per wave, its blocks dealt round-robin to the
// wave's 64 / L groups (G), whose per-entry sums meet in LDS and are added in group
// order: a short chain of loads per lane where a level has too few rows to fill the
// GPU with narrow rows (its rows are then latency-bound chains of dependent loads).
// The sums' order is fixed either way (per lane in block order, then lanes/groups in
// order), so each variant is deterministic; they differ from each other by rounding.
//
// mode 0: out = r - A x, and (out2 not NULL) out2 = w D^-1 (r - A x), the damped Jacobi
//         start of a cycle whose right-hand side is this residual;
// mode 1: out = x + w D^-1 (r - A x) (+ add, when add is not NULL).
// kU blocks' loads are issued together per lane (4 or 8); kNT: the blocks' loads are
// non-temporal. Neither changes the sums or their order.
template <class T, int kWide, int kU, int kNT>
__global__ __launch_bounds__(256) void k_c_resid(int mode, int64_t n, const uint32_t* ast, const uint8_t* done,
                                                 const int* rowptr, const int* col, const T* A, const T* dinv,
                                                 const double* step, const T* r, const T* x, T* out,
                                                 const T* add, T* out2) {
  using CL = CoarseLanes<T>;
  constexpr int E = CL::E, L = CL::L;
  constexpr int R = kWide ? 1 : CL::R;  // rows per wave
  constexpr int G = kWide ? 64 / L : 1;  // lane groups per row
  constexpr int NG = 64 / L;             // lane groups per wave
  __shared__ T part[4][NG][36];
  __shared__ T vres[4][R][6];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int grp = lane / L, i = lane - grp * L;
  const int g = kWide ? 0 : grp;   // the row of the wave this group works
  const int gi = kWide ? grp : 0;  // the group's place among the row's groups
  const int64_t I = ((int64_t)blockIdx.x * 4 + wv) * R + g;
  const bool live = grp < NG && g < R && I < n && !(done && done[ast[I]]);
  if (live) {
    T acc[E];
#pragma unroll
    for (int k = 0; k < E; k++) acc[k] = T(0);
    const int c0 = (E * i) % 6;
    int b = rowptr[I] + gi;
    const int b1 = rowptr[I + 1];
    for (; b + (kU - 1) * G < b1; b += kU * G) {
      int cb[kU];
      T av[kU][E], xv[kU][E];
#pragma unroll
      for (int q = 0; q < kU; q++) cb[q] = col[b + q * G];
#pragma unroll
      for (int q = 0; q < kU; q++) {
        block_entries<T, kNT>(A, (int64_t)(b + q * G), i, av[q]);
        load_columns<T>(x + 6 * (int64_t)cb[q], c0, xv[q]);
      }
#pragma unroll
      for (int q = 0; q < kU; q++)
#pragma unroll
        for (int k = 0; k < E; k++) acc[k] += av[q][k] * xv[q][k];
    }
    for (; b < b1; b += G) {
      T av[E], xv[E];
      block_entries<T, kNT>(A, (int64_t)b, i, av);
      load_columns<T>(x + 6 * (int64_t)col[b], c0, xv);
#pragma unroll
      for (int k = 0; k < E; k++) acc[k] += av[k] * xv[k];
    }
#pragma unroll
    for (int k = 0; k < E; k++) part[wv][grp][E * i + k] = acc[k];
  }
  __syncthreads();
  T v = T(0);
  if (live && gi == 0 && i < 6) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 6; c++) {
      T e = part[wv][g][6 * i + c];
#pragma unroll
      for (int h = 1; h < G; h++) e += part[wv][h][6 * i + c];
      s += e;
    }
    v = r[6 * I + i] - s;
    if (mode == 0) out[6 * I + i] = v;
    vres[wv][g][i] = v;
  }
  if (mode == 0 && !out2) return;
  __syncthreads();
  if (live && gi == 0 && i < 6) {
    T d = T(0);
#pragma unroll
    for (int j = 0; j < 6; j++) d += dinv[36 * I + 6 * i + j] * vres[wv][g][j];
    d *= (T)step[ast[I]];
    if (mode == 0)
      out2[6 * I + i] = d;
    else
      out[6 * I + i] = x[6 * I + i] + d + (add ? add[6 * I + i] : T(0));
  }
}
template <class T>
inline unsigned coarse_blocks(int64_t n, bool wide) {
  const int64_t rpb = wide ? 4 : CoarseLanes<T>::kRowsPerBlock;
  return (unsigned)std::max<int64_t>(1, (n + rpb - 1) / rpb);
}

// The same residual kernel on FP16 blocks (HalfBlock records, five aligned 16-byte loads
// each, the column and scale inside): G lanes per row, each lane whole blocks in turn
// (blocks gl, gl + G, ... of the row, two in flight), y = sum of scale (A_h x[col]) in FP32;
// the lanes' sums meet by a butterfly (the same bits in every lane of the group), and
// lanes 0-5 finish components 0-5 as k_c_resid does. A lane load covers a whole block
// (one load instruction per 16 bytes, against one per 8 bytes and two more for x in the
// lane layout above), which the texture units address at a quarter of the cost. kU
// blocks' loads are issued together per lane. (Non-temporal block loads ran at half
// speed; dealing each XCD one contiguous run of rows, so that neighbouring aggregates'
// x meet in its L2, changed nothing.)
typedef unsigned st_u4 __attribute__((ext_vector_type(4)));
typedef _Float16 st_h2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load_half_block(const HalfBlock* p, st_u4* q) {
  const st_u4* s = reinterpret_cast<const st_u4*>(p);
#pragma unroll
  for (int k = 0; k < 5; k++) q[k] = s[k];
}
__device__ __forceinline__ void load_x6(const float* p, float* v) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float2 a = *reinterpret_cast<const float2*>(p + 2 * k);
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
}
__device__ __forceinline__ void half_block_madd(const st_u4* q, const float* xv, float* y) {
  unsigned w[20];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    w[4 * k] = q[k].x;
    w[4 * k + 1] = q[k].y;
    w[4 * k + 2] = q[k].z;
    w[4 * k + 3] = q[k].w;
  }
  const float sc = __builtin_bit_cast(float, w[18]);
#pragma unroll
  for (int r = 0; r < 6; r++) {
    float t = 0.0f;
#pragma unroll
    for (int j = 0; j < 6; j++) {
      const int e = 6 * r + j;
      const st_h2 h = __builtin_bit_cast(st_h2, w[e >> 1]);
      t += (float)((e & 1) ? h.y : h.x) * xv[j];
    }
    y[r] += sc * t;
  }
}
template <int G, int kU>
__global__ __launch_bounds__(256) void k_c_resid_h(int mode, int64_t n, const uint32_t* ast, const uint8_t* done,
                                                   const int* rowptr, const HalfBlock* A, const float* dinv,
                                                   const double* step, const float* r, const float* x, float* out,
                                                   const float* add, float* out2) {
  static_assert(G >= 8 && G <= 64 && (G & (G - 1)) == 0, "G: 8 to 64 lanes, a power of two");
  const int gl = threadIdx.x & (G - 1);
  const int64_t I = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const bool live = I < n && !(done && done[ast[I]]);
  float y[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (live) {
    int b = rowptr[I] + gl;
    const int b1 = rowptr[I + 1];
    for (; b + (kU - 1) * G < b1; b += kU * G) {
      st_u4 q[kU][5];
      float xv[kU][6];
#pragma unroll
      for (int u = 0; u < kU; u++) load_half_block(A + b + u * G, q[u]);
#pragma unroll
      for (int u = 0; u < kU; u++) load_x6(x + 6 * (int64_t)(int)q[u][4].w, xv[u]);
#pragma unroll
      for (int u = 0; u < kU; u++) half_block_madd(q[u], xv[u], y);
    }
    for (; b < b1; b += G) {
      st_u4 q0[5];
      load_half_block(A + b, q0);
      float x0[6];
      load_x6(x + 6 * (int64_t)(int)q0[4].w, x0);
      half_block_madd(q0, x0, y);
    }
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1)
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] += __shfl_xor(y[k], o, 64);
  if (!live || gl >= 6) return;
  const int i = gl;
  float v[6];
  load_x6(r + 6 * I, v);
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] -= y[k];
  if (mode == 0) out[6 * I + i] = v[i];
  if (mode == 0 && !out2) return;
  float d = 0.0f;
#pragma unroll
  for (int j = 0; j < 6; j++) d += dinv[36 * I + 6 * i + j] * v[j];
  d *= (float)step[ast[I]];
  if (mode == 0)
    out2[6 * I + i] = d;
  else
    out[6 * I + i] = x[6 * I + i] + d + (add ? add[6 * I + i] : 0.0f);
}
template <int G>
inline unsigned half_blocks_grid(int64_t n) {
  return (unsigned)std::max<int64_t>(1, (n * G + 255) / 256);
}

// y = A x (for power iterations), then partial sums of x.x per aggregate are not needed: w = D^-1 A x
template <class T>
__global__ void k_c_dinvA(int64_t n, const int* rowptr, const int* col, const T* A, const T* dinv, const T* x, T* w) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t I = t >> 3;
  const int i = (int)(t & 7);
  T acc = 0;
  if (I < n && i < 6) acc = block_row_dot<T>(rowptr[I], rowptr[I + 1], col, A, x, i);
  T all[6];
  group_gather6<T>(acc, all);
  if (I < n && i < 6) w[6 * I + i] = row6_dot<T>(dinv + 36 * I + 6 * i, all);
}

template <class T>
__global__ void k_c_sq(int64_t n, const T* x, double* out) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  double s = 0;
  for (int k = 0; k < 6; k++) s += (double)x[6 * I + k] * (double)x[6 * I + k];
  out[I] = s;
}

__global__ void k_c_init(int64_t n, double* x) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  for (int k = 0; k < 6; k++) x[6 * I + k] = 1.0 + (double)((7 * I + 3 * k) % 11) / 11.0;
}
__global__ void k_c_init32(int64_t n, float* x) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  for (int k = 0; k < 6; k++) x[6 * I + k] = 1.0 + (float)((7 * I + 3 * k) % 11) / 11.0;
}

// restrict fine -> level 1: c[I] = sum over nodes v of I of Sh(r_v)^T t[v]
// brick_agg[b] = the first level-1 aggregate whose first node lies in brick b or later
// (aggregates are numbered in the order of their first nodes); brick_agg[nb] = n1
__global__ void k_brick_agg1(int64_t nb, const uint32_t* nbase, int64_t n1, const int* agg_ptr, const int* nodes,
                             int* brick_agg) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  if (b == nb) {
    brick_agg[b] = (int)n1;
    return;
  }
  const uint32_t start = nbase[b];
  int64_t lo = 0, hi = n1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((uint32_t)nodes[agg_ptr[mid]] < start)
      lo = mid + 1;
    else
      hi = mid;
  }
  brick_agg[b] = (int)lo;
}
// Flags an aggregate whose nodes (sorted) do not all lie in the brick of its first node.
__global__ void k_agg_straddle(int64_t n1, const int* agg_ptr, const int* nodes, int64_t nb, const uint32_t* nbase,
                               int* bad) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n1) return;
  const uint32_t first = (uint32_t)nodes[agg_ptr[I]], last = (uint32_t)nodes[agg_ptr[I + 1] - 1];
  int64_t lo = 0, hi = nb;  // the brick of `first`: the last b with nbase[b] <= first
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (nbase[mid] <= first)
      lo = mid;
    else
      hi = mid;
  }
  if (last >= nbase[lo + 1] || first < nbase[lo]) atomicOr(bad, 1);
}

template <class T>
__global__ void k_restrict1(int64_t n1, const uint32_t* ast, const uint8_t* done, const int* agg_ptr, const int* nodes,
                            const double* centre, const uint32_t* nkey, const uint8_t* nkind, const BeamGeo* bgeo,
                            const T* t, T* c) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n1) return;
  const uint32_t s = ast[I];
  if (done && done[s]) return;
  const double sp = 2.0 * bgeo[s].hd[0];
  T acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = agg_ptr[I]; i < agg_ptr[I + 1]; i++) {
    const uint32_t v = (uint32_t)nodes[i];
    double p[3];
    node_pos(v, nkey, nkind, sp, p);
    for (int k = 0; k < 3; k++) p[k] -= centre[3 * I + k];
    T tv[6];
    load6<T>(t + 6 * (size_t)v, tv);
    shiftT_add<T>(tv, p, acc);
  }
  store6<T>(c + 6 * I, acc);
}

// prolong level 1 -> fine (brick-blocked): x[v] += Sh(r_v) y[label[v]]
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong1(const uint32_t* nbase, const uint32_t* bstruct, const uint8_t* done,
                                                     const int* label, const double* centre, const uint32_t* nkey,
                                                     const uint8_t* nkind, const BeamGeo* bgeo, const T* y, T* x) {
  const uint32_t b = blockIdx.x, s = bstruct[b];
  if (done && done[s]) return;
  const double sp = 2.0 * bgeo[s].hd[0];
  for (uint32_t v = nbase[b] + threadIdx.x; v < nbase[b + 1]; v += blockDim.x) {
    const int I = label[v];
    if (I < 0) continue;
    double p[3];
    node_pos(v, nkey, nkind, sp, p);
    for (int k = 0; k < 3; k++) p[k] -= centre[3 * (int64_t)I + k];
    T xv[6], yv[6];
    load6<T>(x + 6 * (size_t)v, xv);
    load6<T>(y + 6 * (int64_t)I, yv);
    shift_add<T>(yv, p, xv);
    store6<T>(x + 6 * (size_t)v, xv);
  }
}

// restrict level l -> l+1 over children lists
// With x0 not NULL, also the damped Jacobi start of the coarse level's first cycle,
// x0 = w D^-1 c (dinv, step: the coarse level's).
template <class T>
__global__ void k_restrict_c(int64_t n, const uint32_t* ast, const uint8_t* done, const int* kid_ptr, const int* kids,
                             const double* centre, const double* centre_c, const T* t, T* c, const T* dinv,
                             const double* step, T* x0) {
  const int64_t J = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= n) return;
  if (done && done[ast[J]]) return;
  T acc[6] = {0, 0, 0, 0, 0, 0};
  for (int u = kid_ptr[J]; u < kid_ptr[J + 1]; u++) {
    const int i = kids[u];
    double r[3];
    for (int k = 0; k < 3; k++) r[k] = centre_c[3 * (int64_t)i + k] - centre[3 * J + k];
    T tv[6];
    load6<T>(t + 6 * (int64_t)i, tv);
    shiftT_add<T>(tv, r, acc);
  }
  store6<T>(c + 6 * J, acc);
  if (x0) {
    T d[6];
    mat6v<T>(dinv + 36 * J, acc, d);
    const T w = (T)step[ast[J]];
#pragma unroll
    for (int k = 0; k < 6; k++) d[k] *= w;
    store6<T>(x0 + 6 * J, d);
  }
}

// prolong level l+1 -> l: x[i] += Sh(r_i) y[parent[i]]
template <class T>
__global__ void k_prolong_c(int64_t n, const uint32_t* ast, const uint8_t* done, const int* parent, const double* centre,
                            const double* centre_p, const T* y, T* x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (done && done[ast[i]]) return;
  const int P = parent[i];
  double r[3];
  for (int k = 0; k < 3; k++) r[k] = centre[3 * i + k] - centre_p[3 * (int64_t)P + k];
  T xv[6], yv[6];
  load6<T>(x + 6 * i, xv);
  load6<T>(y + 6 * (int64_t)P, yv);
  shift_add<T>(yv, r, xv);
  store6<T>(x + 6 * i, xv);
}

// structure ranges of an aggregate numbering: out[s] = ids[first node of structure s]
__global__ void k_seg_from_ids(int ns, const int64_t* node0, const int* ids, const int* flag, int64_t nnodes,
                               int64_t* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > ns) return;
  if (s < ns)
    out[s] = node0[s] < nnodes ? ids[node0[s]] : (nnodes ? ids[nnodes - 1] + flag[nnodes - 1] : 0);
  else
    out[s] = nnodes ? ids[nnodes - 1] + flag[nnodes - 1] : 0;
}

// ------------------------------------------------------- smoothed aggregation
// The prolongator P = (I - w D^-1 A) P0 (statics/src/bodies.rs:238-247), built
// explicitly row by row, and the Galerkin product P^T A P (:255-269) gathered
// row by row of the coarse operator over the columns of P: every output row is
// owned by one thread, so there are no atomics and the sums run in a fixed order.
constexpr int kMaxP = 16;   // blocks in one row of the fine P0 (a node has at most 10 beams)
constexpr int kMaxPc = 96;  // blocks in one row of a coarse P

__device__ __forceinline__ void mat6_mul_add(const double* a, const double* b, double* out) {  // out += a b
  for (int r = 0; r < 6; r++)
    for (int c2 = 0; c2 < 6; c2++) {
      double s = 0;
      for (int k = 0; k < 6; k++) s += a[6 * r + k] * b[6 * k + c2];
      out[6 * r + c2] += s;
    }
}
__device__ __forceinline__ void mat6_tmul(const double* a, const double* b, double* out) {  // out = a^T b
  for (int r = 0; r < 6; r++)
    for (int c2 = 0; c2 < 6; c2++) {
      double s = 0;
      for (int k = 0; k < 6; k++) s += a[6 * k + r] * b[6 * k + c2];
      out[6 * r + c2] = s;
    }
}
// Sh(r) as a 6x6 matrix: [[I, -[r]x], [0, I]]
__device__ __forceinline__ void shift_mat(const double* r, double* S) {
  for (int i = 0; i < 36; i++) S[i] = (i % 7 == 0) ? 1.0 : 0.0;
  S[0 * 6 + 4] = r[2];
  S[0 * 6 + 5] = -r[1];
  S[1 * 6 + 3] = -r[2];
  S[1 * 6 + 5] = r[0];
  S[2 * 6 + 3] = r[1];
  S[2 * 6 + 4] = -r[0];
}
// out = M Sh(r) for a 6x6 M
__device__ __forceinline__ void mat6_shift(const double* M, const double* r, double* out) {
  double S[36];
  shift_mat(r, S);
  for (int i = 0; i < 36; i++) out[i] = 0.0;
  mat6_mul_add(M, S, out);
}
__device__ __forceinline__ int slot_of(const int* cols, int len, int J) {
  for (int t = 0; t < len; t++)
    if (cols[t] == J) return t;
  return -1;
}

// One live beam at fine node v, seen from v: the beam's pair q, v's sign, the other end.
struct BeamAt {
  int q, sv, c;
  uint32_t wf;  // other end's face entry (kNotFree if held)
};
// Enumerates the live beams at v (over its two cells) into list; returns the count (<= 10).
__device__ int beams_at(uint32_t v, const uint8_t* nkind, const int32_t* ncell, const uint32_t* cfaces,
                        const uint16_t* cmask, BeamAt* list) {
  const int o = nkind[v] >> 1;
  int nb = 0;
  for (int side = 0; side < 2; side++) {
    const int c = ncell[2 * v + side];
    if (c < 0) continue;
    const int f = side == 0 ? 2 * o + 1 : 2 * o;
    const uint32_t m = (uint32_t)cmask[c] & face_beams(f);
    for (int q = 0; q < 15; q++) {
      if (!(m & (1u << q))) continue;
      const int f2 = kPairA[q] == f ? kPairB[q] : kPairA[q];
      list[nb].q = q;
      list[nb].sv = kPairA[q] == f ? -1 : 1;
      list[nb].c = c;
      list[nb].wf = cfaces[6 * c + f2];
      nb++;
    }
  }
  return nb;
}

// The ten beam places at v (five at its face of each of its two cells, in beams_at's order)
// with dead or absent beams marked q = -1: a list indexed by constants, so it stays in
// registers.
__device__ __forceinline__ void beams_at10(uint32_t v, const uint8_t* nkind, const int32_t* ncell,
                                           const uint32_t* cfaces, const uint16_t* cmask, BeamAt* list) {
  const int o = nkind[v] >> 1;
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int c = ncell[2 * v + side];
    const int f = side == 0 ? 2 * o + 1 : 2 * o;
    const uint32_t m = c >= 0 ? ((uint32_t)cmask[c] & face_beams(f)) : 0u;
#pragma unroll
    for (int jj = 0; jj < 5; jj++) {
      int q = 0;
      for (int qq = 0, k = 0; qq < 15; qq++)
        if (kPairA[qq] == f || kPairB[qq] == f) {
          if (k == jj) q = qq;
          k++;
        }
      BeamAt& b = list[5 * side + jj];
      const bool live = (m >> q) & 1u;
      b.q = live ? q : -1;
      b.sv = kPairA[q] == f ? -1 : 1;
      b.c = c;
      b.wf = live ? cfaces[6 * c + (kPairA[q] == f ? kPairB[q] : kPairA[q])] : kNotFree;
    }
  }
}

// Rows of P0 = (I - w D^-1 K) P_rigid at the fine nodes. pass 0: count blocks; pass 1: fill.
// (K chi_J)(v), the load at v when aggregate J moves rigidly: for v in J only the beams
// leaving J count (a beam inside J moves rigidly), K_b,vv Sh(r_v); for v not in J the beams
// reaching into J, K_b,vw Sh(r_w).
__global__ __launch_bounds__(256) void k_p0(int pass, int64_t nn, const uint8_t* nflags, const int* label, const double* centre,
                     const uint32_t* nkey, const uint8_t* nkind, const int32_t* ncell, const uint32_t* cfaces,
                     const double* ce, const double* cg, const double* csec, const uint16_t* cmask,
                     const uint32_t* cstruct, const BeamGeo* bgeo, const double* dinv, const double* step, int* cnt,
                     const int* prow, int* pcol, double* pval, int* overflow) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  if (nflags[v] & (kHeld | kGone)) {
    if (pass == 0) cnt[v] = 0;
    return;
  }
  BeamAt bl[10];
  beams_at10((uint32_t)v, nkind, ncell, cfaces, cmask, bl);
  const int I0 = label[v];
  int js[kMaxP];
  int nj = 1;
  js[0] = I0;
#pragma unroll
  for (int t = 0; t < 10; t++) {
    if (bl[t].q < 0 || (bl[t].wf & kNotFree)) continue;
    const int J = label[bl[t].wf];
    if (J != I0 && slot_of(js, nj, J) < 0) {
      if (nj < kMaxP)
        js[nj++] = J;
      else
        atomicExch(overflow, 1);
    }
  }
  if (pass == 0) {
    cnt[v] = nj;
    return;
  }
  const int base = prow[v];
  double* out = pval + 36 * (int64_t)base;
  for (int t = 0; t < nj; t++) pcol[base + t] = js[t];
  const uint32_t sidx = cstruct[ncell[2 * v] >= 0 ? ncell[2 * v] : ncell[2 * v + 1]];
  const BeamGeo G = bgeo[sidx];
  const double sp = 2.0 * G.hd[0];
  double pv[3], rv[3];
  node_pos((uint32_t)v, nkey, nkind, sp, pv);
  for (int k = 0; k < 3; k++) rv[k] = pv[k] - centre[3 * (int64_t)I0 + k];
  const double zero[3] = {0, 0, 0};
  const double w = step[sidx];
  // P[v, J] = [J = I0] Sh(r_v) - w D^-1 (K chi_J)(v), one block at a time in registers: the
  // beams leaving I0 add K_b,vv Sh(r_v) into I0's block, a beam reaching into J != I0 adds
  // K_b,vw Sh(r_w) into J's (a beam inside I0 moves rigidly and adds nothing)
  for (int t = 0; t < nj; t++) {
    const int J = js[t];
    double acc[36];
#pragma unroll
    for (int i = 0; i < 36; i++) acc[i] = 0.0;
#pragma unroll 1
    for (int bt = 0; bt < 10; bt++) {
      const BeamAt& b = bl[bt];
      if (b.q < 0) continue;
      const bool wfree = !(b.wf & kNotFree);
      const int Jw = wfree ? label[b.wf] : -1;
      if (Jw == I0) continue;
      const bool vv = t == 0, vw = Jw == J;
      if (!vv && !vw) continue;
      double n[3];
      int cls;
      pair_n(b.q, G, n, cls);
      const double e = ce[b.c], g = cg[b.c];
      const double rho = beam_sec(csec, b.c, b.q), K4 = G.K4[cls] * rho, R2 = G.R2[cls] * (rho * rho);
      double Bv[18];
      gmat(b.sv, n, G.hd[cls], zero, Bv);
      if (vv) {
        double Gv[18];
        gmat(b.sv, n, G.hd[cls], rv, Gv);
        beam_coarse_add(Bv, Gv, b.sv, b.sv, n, K4, R2, e, g, acc);  // K_b,vv Sh(r_v)
      }
      if (vw) {
        double pw[3], rw[3], Gw[18];
        node_pos(b.wf, nkey, nkind, sp, pw);
        for (int k = 0; k < 3; k++) rw[k] = pw[k] - centre[3 * (int64_t)Jw + k];
        gmat(-b.sv, n, G.hd[cls], rw, Gw);
        beam_coarse_add(Bv, Gw, b.sv, -b.sv, n, K4, R2, e, g, acc);  // K_b,vw Sh(r_w)
      }
    }
    double D[36];
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) D[6 * i + j] = dinv[kSym * v + sym_idx(i, j)];
    double tmp[36];
#pragma unroll
    for (int i = 0; i < 36; i++) tmp[i] = 0.0;
    mat6_mul_add(D, acc, tmp);
#pragma unroll
    for (int i = 0; i < 36; i++) acc[i] = -w * tmp[i];
    if (t == 0) {
      double S[36];
      shift_mat(rv, S);
#pragma unroll
      for (int i = 0; i < 36; i++) acc[i] += S[i];
    }
    double* blk = out + 36 * t;
#pragma unroll
    for (int i = 0; i < 36; i++) blk[i] = acc[i];
  }
}


// Rows of P = (I - w D^-1 A) P_rigid at an explicit level (rows i, columns the next level's J).
__global__ __launch_bounds__(256) void k_pc(int pass, int smooth, int64_t n, const int* rowptr, const int* colA, const double* A,
                     const double* dinv, const double* step, const uint32_t* ast, const int* parent,
                     const double* centre, const double* centre_up, int* cnt, const int* prow, int* pcol, double* pval,
                     int* overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int J0 = parent[i];
  int js[kMaxPc];
  int nj = 1;
  js[0] = J0;
  for (int b = smooth ? rowptr[i] : rowptr[i + 1]; b < rowptr[i + 1]; b++) {
    const int J = parent[colA[b]];
    if (slot_of(js, nj, J) < 0) {
      if (nj < kMaxPc)
        js[nj++] = J;
      else
        atomicExch(overflow, 1);
    }
  }
  if (pass == 0) {
    cnt[i] = nj;
    return;
  }
  const int base = prow[i];
  double* out = pval + 36 * (int64_t)base;
  for (int t = 0; t < nj; t++) pcol[base + t] = js[t];
  for (int k = 0; k < 36 * nj; k++) out[k] = 0.0;
  for (int b = smooth ? rowptr[i] : rowptr[i + 1]; b < rowptr[i + 1]; b++) {
    const int j = colA[b];
    const int J = parent[j];
    double r[3], M[36];
    for (int k = 0; k < 3; k++) r[k] = centre[3 * (int64_t)j + k] - centre_up[3 * (int64_t)J + k];
    mat6_shift(A + 36 * (int64_t)b, r, M);
    double* blk = out + 36 * slot_of(js, nj, J);
    for (int k = 0; k < 36; k++) blk[k] += M[k];
  }
  const double w = step[ast[i]];
  const double* D = dinv + 36 * i;
  for (int t = 0; t < nj; t++) {
    double* blk = out + 36 * t;
    double tmp[36];
    for (int k = 0; k < 36; k++) tmp[k] = 0.0;
    mat6_mul_add(D, blk, tmp);
    for (int k = 0; k < 36; k++) blk[k] = -w * tmp[k];
  }
  double r0[3], S[36];
  for (int k = 0; k < 3; k++) r0[k] = centre[3 * i + k] - centre_up[3 * (int64_t)J0 + k];
  shift_mat(r0, S);
  for (int k = 0; k < 36; k++) out[k] += S[k];
}


// ---- Galerkin rows, one wavefront per output row. The row's columns go into an LDS
// hash set (open addressing), are sorted (diagonal first, then ascending) and numbered,
// and the 6x6 products accumulate into LDS with native ds_add_f64. pass 0 counts the
// row's columns; pass 1 fills col and A. kSlots bounds the columns of one row.
constexpr int kHash = 2048;

template <int kH>
__device__ __forceinline__ unsigned hslot(int J) {
  return ((unsigned)J * 2654435761u) & (kH - 1);
}
template <int kH>
__device__ __forceinline__ void hinsert(int* keys, int J) {
  unsigned h = hslot<kH>(J);
  for (;;) {
    const int old = atomicCAS(&keys[h], -1, J);
    if (old == -1 || old == J) return;
    h = (h + 1) & (kH - 1);
  }
}
template <int kH>
__device__ __forceinline__ int hfind(const int* keys, const int* slot, int J) {
  unsigned h = hslot<kH>(J);
  for (;;) {
    const int k = keys[h];
    if (k == J) return slot[h];
    if (k == -1) return -1;
    h = (h + 1) & (kH - 1);
  }
}

// Shared part of both row kernels once the hash holds the row's columns: count them
// (pass 0) or sort and number them (pass 1, when they fit in cap). Returns the count,
// the diagonal included.
template <int kH>
__device__ int row_columns(int pass, int I, int* keys, int* slot, int* sorted, int cap, int* ncount) {
  __syncthreads();
  if (threadIdx.x == 0) *ncount = 0;
  __syncthreads();
  for (int h = threadIdx.x; h < kH; h += blockDim.x)
    if (keys[h] != -1 && keys[h] != I) {
      const int p = atomicAdd(ncount, 1);
      if (pass == 1 && p < cap) sorted[p] = keys[h];
    }
  __syncthreads();
  const int m = *ncount;
  if (pass == 1 && m + 1 <= cap) {
    // rank sort (m is small): an entry's place is the number of smaller keys
    for (int a = threadIdx.x; a < m; a += blockDim.x) {
      const int key = sorted[a];
      int rank = 0;
      for (int b = 0; b < m; b++) rank += sorted[b] < key;
      unsigned h = hslot<kH>(key);
      while (keys[h] != key) h = (h + 1) & (kH - 1);
      slot[h] = 1 + rank;
    }
    if (threadIdx.x == 0) {
      unsigned h = hslot<kH>(I);
      while (keys[h] != I) h = (h + 1) & (kH - 1);
      slot[h] = 0;
    }
  }
  __syncthreads();
  return m + 1;
}

// Inclusive sum over a 256-thread block (wsum: 4 ints of LDS).
__device__ __forceinline__ int block_scan256(int v, int* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  int add = 0;
  for (int k = 0; k < w; k++) add += wsum[k];
  __syncthreads();
  return v + add;
}

// Next level = P^T A P by rows, one workgroup per row I (rows[blockIdx.x], or
// blockIdx.x): row I = sum over the nodes i of column I of P, 256 at a time, of
// P[i,I]^T A[i,j] P[j,:]. Pass 1 groups the chunk's terms by j: one task per column j met
// in the chunk's rows of A forms T = sum over i of P[i,I]^T A[i,j] (A[i,j] = A[j,i]^T, row
// j of A read against the chunk's nodes, looked up in an LDS hash), then adds T P[j,J] for
// the J of P's row j: one pass over P's row j per j instead of one per (i, j). A chunk
// meeting kWL columns or more, or every chunk when pairs is set, takes one task per (i, j).
template <int pass, int kSlots, int kH, int kRG = 1>
__global__ __launch_bounds__(256) void k_sa_gal_w(const int* rows, int64_t nI, const int* pt_ptr,
                                                  const int* pt_blk, const int* pblk_row, const int* prow,
                                                  const int* pcol, const double* pval, const int* rowptrA,
                                                  const int* colA, const double* Aa, int* rowcnt, const int* rowptr,
                                                  int* col, double* A, int* overflow, int pairs) {
  // hashes of the chunk's nodes i (-> their block of P) and of its columns j, and the
  // columns j one chunk groups by (kWH - kWL > 256); smaller beside the widest rows' LDS
  constexpr int kIH = pass ? 512 : 1;
  constexpr int kWH = pass ? (kSlots >= 448 ? 1024 : 2048) : 1;
  constexpr int kWL = kSlots >= 448 ? 512 : 1024;
  __shared__ int keys[kH];
  __shared__ int slot[kH];
  __shared__ int sorted[kSlots];
  __shared__ int ncount;
  __shared__ int wsum[4];
  __shared__ int pre[257];
  __shared__ int rowi[256], blki[256];
  __shared__ double acc[36 * kSlots];
  __shared__ int ikey[kIH], ival[kIH], wkey[kWH], wl[pass ? kWL : 1];
  __shared__ int nw;
  if ((int64_t)blockIdx.x >= nI) return;
  const int I = rows ? rows[blockIdx.x] : (int)blockIdx.x;
  for (int h = threadIdx.x; h < kH; h += blockDim.x) keys[h] = -1;
  __syncthreads();
  if (threadIdx.x == 0) hinsert<kH>(keys, I);
  const int k0 = pt_ptr[I], k1 = pt_ptr[I + 1];
  constexpr int NRG = 6 / kRG;
  int m = 0;
  for (int run = 0; run < 2; run++) {  // run 0: columns; run 1 (pass 1): products
    if (run == 1) {
      m = row_columns<kH>(pass, I, keys, slot, sorted, kSlots, &ncount);
      if (pass == 0) {
        if (threadIdx.x == 0) rowcnt[I] = m;
        return;
      }
      if (m > kSlots) {
        if (threadIdx.x == 0) atomicMax(overflow, m);
        return;
      }
      for (int t = threadIdx.x; t < 36 * m; t += blockDim.x) acc[t] = 0.0;
      __syncthreads();
    }
    for (int c0 = k0; c0 < k1; c0 += 256) {
      const int nc = min(256, k1 - c0);
      if (run == 1) {
        for (int h = threadIdx.x; h < kIH; h += blockDim.x) ikey[h] = -1;
        for (int h = threadIdx.x; h < kWH; h += blockDim.x) wkey[h] = -1;
        if (threadIdx.x == 0) nw = 0;
        __syncthreads();
      }
      int len = 0;
      if ((int)threadIdx.x < nc) {
        const int blk = pt_blk[c0 + threadIdx.x];
        const int i = pblk_row[blk];
        blki[threadIdx.x] = blk;
        rowi[threadIdx.x] = i;
        len = rowptrA[i + 1] - rowptrA[i];
        if (run == 1) {  // a node appears once in a column of P
          unsigned h = hslot<kIH>(i);
          while (atomicCAS(&ikey[h], -1, i) != -1) h = (h + 1) & (kIH - 1);
          ival[h] = blk;
        }
      }
      const int incl = block_scan256(len, wsum);
      pre[threadIdx.x + 1] = incl;
      if (threadIdx.x == 0) pre[0] = 0;
      __syncthreads();
      const int total = pre[nc];
      // run 0: the columns J of P's rows j; run 1: the distinct j (at most kWL + 255 enter
      // the hash, which therefore never fills)
      for (int it = threadIdx.x; it < total; it += blockDim.x) {
        int lo = 0, hi = nc - 1;  // the last t with pre[t] <= it
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (pre[mid] <= it)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int j = colA[rowptrA[rowi[lo]] + (it - pre[lo])];
        if (run == 0) {
          for (int k2 = prow[j]; k2 < prow[j + 1]; k2++) hinsert<kH>(keys, pcol[k2]);
          continue;
        }
        if (*(volatile int*)&nw >= kWL) continue;
        unsigned h = hslot<kWH>(j);
        for (;;) {
          const int old = atomicCAS(&wkey[h], -1, j);
          if (old == -1) {
            const int p = atomicAdd(&nw, 1);
            if (p < kWL) wl[p] = j;
            break;
          }
          if (old == j) break;
          h = (h + 1) & (kWH - 1);
        }
      }
      __syncthreads();
      if (run == 0) continue;
      const int nwc = nw;
      if (!pairs && nwc < kWL) {
        for (int it2 = threadIdx.x; it2 < nwc * NRG; it2 += blockDim.x) {
          const int rg = it2 % NRG, j = wl[it2 / NRG];
          // rows rg kRG .. of T = sum over the chunk's i of P[i,I]^T A[i,j]
          double qr[kRG][6];
          for (int rr = 0; rr < kRG; rr++)
            for (int k = 0; k < 6; k++) qr[rr][k] = 0.0;
          for (int b = rowptrA[j]; b < rowptrA[j + 1]; b++) {
            const int i = colA[b];
            unsigned h = hslot<kIH>(i);
            int blk = -1;
            for (;;) {
              const int k = ikey[h];
              if (k == i) {
                blk = ival[h];
                break;
              }
              if (k == -1) break;
              h = (h + 1) & (kIH - 1);
            }
            if (blk < 0) continue;
            const double* PiI = pval + 36 * (int64_t)blk;
            const double* Ab = Aa + 36 * (int64_t)b;  // A[j,i]: A[i,j][i2][k] = Ab[6 k + i2]
            for (int i2 = 0; i2 < 6; i2++)
              for (int k = 0; k < 6; k++) {
                const double ab = Ab[6 * k + i2];
                for (int rr = 0; rr < kRG; rr++) qr[rr][k] += PiI[6 * i2 + rg * kRG + rr] * ab;
              }
          }
          const int k2a = prow[j], len2 = prow[j + 1] - k2a;
          for (int kk = 0; kk < len2; kk++) {
            const int k2 = k2a + (kk + it2) % len2;
            const int J = pcol[k2];
            if (J < I) continue;  // the lower blocks are the transposes of other rows' upper ones
            const int sl = hfind<kH>(keys, slot, J);
            const double* Pj = pval + 36 * (int64_t)k2;
            double o6[kRG][6];
            for (int rr = 0; rr < kRG; rr++)
              for (int cc = 0; cc < 6; cc++) o6[rr][cc] = 0.0;
            for (int k = 0; k < 6; k++)
              for (int cc = 0; cc < 6; cc++) {
                const double pk = Pj[6 * k + cc];
                for (int rr = 0; rr < kRG; rr++) o6[rr][cc] += qr[rr][k] * pk;
              }
            double* out = acc + 36 * sl + 6 * rg * kRG;
            for (int rr = 0; rr < kRG; rr++)
              for (int cc = 0; cc < 6; cc++) atomicAdd(&out[6 * rr + cc], o6[rr][cc]);
          }
        }
      } else {
        // one task per (i, j): rows of P[i,I]^T A[i,j], then of its products with P[j,:]
        for (int it2 = threadIdx.x; it2 < total * NRG; it2 += blockDim.x) {
          const int it = it2 / NRG, rg = it2 % NRG;
          int lo = 0, hi = nc - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= it)
              lo = mid;
            else
              hi = mid - 1;
          }
          const int b = rowptrA[rowi[lo]] + (it - pre[lo]);
          const int j = colA[b];
          const double* PiI = pval + 36 * (int64_t)blki[lo];
          const double* Ab = Aa + 36 * (int64_t)b;
          double qr[kRG][6];
          for (int rr = 0; rr < kRG; rr++)
            for (int k = 0; k < 6; k++) qr[rr][k] = 0.0;
          for (int i2 = 0; i2 < 6; i2++)
            for (int k = 0; k < 6; k++) {
              const double ab = Ab[6 * i2 + k];
              for (int rr = 0; rr < kRG; rr++) qr[rr][k] += PiI[6 * i2 + rg * kRG + rr] * ab;
            }
          const int k2a = prow[j], len2 = prow[j + 1] - k2a;
          for (int kk = 0; kk < len2; kk++) {
            const int k2 = k2a + (kk + it) % len2;
            const int J = pcol[k2];
            if (J < I) continue;
            const int sl = hfind<kH>(keys, slot, J);
            const double* Pj = pval + 36 * (int64_t)k2;
            double o6[kRG][6];
            for (int rr = 0; rr < kRG; rr++)
              for (int cc = 0; cc < 6; cc++) o6[rr][cc] = 0.0;
            for (int k = 0; k < 6; k++)
              for (int cc = 0; cc < 6; cc++) {
                const double pk = Pj[6 * k + cc];
                for (int rr = 0; rr < kRG; rr++) o6[rr][cc] += qr[rr][k] * pk;
              }
            double* out = acc + 36 * sl + 6 * rg * kRG;
            for (int rr = 0; rr < kRG; rr++)
              for (int cc = 0; cc < 6; cc++) atomicAdd(&out[6 * rr + cc], o6[rr][cc]);
          }
        }
      }
      __syncthreads();
    }
  }
  const int base = rowptr[I];
  if (threadIdx.x == 0) col[base] = I;
  for (int h = threadIdx.x; h < kH; h += blockDim.x)
    if (keys[h] != -1 && keys[h] != I) col[base + slot[h]] = keys[h];
  for (int t = threadIdx.x; t < 36 * m; t += blockDim.x) A[36 * (int64_t)base + t] = acc[t];
}

// Width classes of Galerkin rows (columns held in LDS by each kernel instance).
constexpr int kRowClass[4] = {64, 128, 256, 448};

// Rows by width class: lists[k * n + i] for class k (order within a class is free:
// every row is computed on its own).
__global__ void k_row_classes(int64_t n, const int* cnt, int* lists, int* ctr) {
  __shared__ int local[4], base[4];
  if (threadIdx.x < 4) local[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int k = -1, pos = 0;
  if (i < n) {
    const int w = cnt[i];
    k = w <= kRowClass[0] ? 0 : (w <= kRowClass[1] ? 1 : (w <= kRowClass[2] ? 2 : 3));
    pos = atomicAdd(&local[k], 1);
  }
  __syncthreads();
  if (threadIdx.x < 4) base[threadIdx.x] = local[threadIdx.x] ? atomicAdd(&ctr[threadIdx.x], local[threadIdx.x]) : 0;
  __syncthreads();
  if (k >= 0) lists[k * n + base[k] + pos] = (int)i;
}

// The blocks below the diagonal of a symmetric operator from the transposes of those
// above it: (I, J) with J < I is (J, I)^T, found in row J by bisection (its columns
// are ascending after the diagonal).
__global__ void k_fill_lower(int64_t n, const int* rowptr, const int* col, double* A) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  for (int p = rowptr[I] + 1; p < rowptr[I + 1]; p++) {
    const int J = col[p];
    if (J >= I) break;
    int lo = rowptr[J] + 1, hi = rowptr[J + 1] - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col[mid] < (int)I)
        lo = mid + 1;
      else
        hi = mid;
    }
    const double* src = A + 36 * (int64_t)lo;
    double* dst = A + 36 * (int64_t)p;
    for (int r = 0; r < 6; r++)
      for (int c2 = 0; c2 < 6; c2++) dst[6 * r + c2] = src[6 * c2 + r];
  }
}

// The five beams touching face f (pair indices, ascending).
__device__ __forceinline__ int face_beam_q(int f, int t) {
  int k = 0;
  for (int q = 0; q < 15; q++)
    if (kPairA[q] == f || kPairB[q] == f) {
      if (k == t) return q;
      k++;
    }
  return 0;
}

// Level 1 = P0^T K P0 by rows, one workgroup per row I: work items are (node v in
// column I of P0, one of the ten beams ending on v), adjacent threads on different
// nodes. kH: hash slots (4x kSlots; pass 0 sizes with the largest instance).
template <int pass, int kSlots, int kH, int kRG = 1>
__global__ __launch_bounds__(256) void k_sa_a1_w(const int* rows, int64_t n1, const int* pt_ptr,
                                                 const int* pt_blk,
                                                 const int* pblk_row, const int* prow, const int* pcol,
                                                 const double* pval, const uint8_t* nkind, const int32_t* ncell,
                                                 const uint32_t* cfaces, const double* ce, const double* cg,
                                                 const double* csec, const uint16_t* cmask, const uint32_t* cstruct,
                                                 const BeamGeo* bgeo,
                                                 int* rowcnt, const int* rowptr, int* col, double* A, int* overflow) {
  __shared__ int keys[kH];
  __shared__ int slot[kH];
  __shared__ int sorted[kSlots];
  __shared__ int ncount;
  __shared__ double acc[pass ? 36 * kSlots : 1];
  // the column's nodes, a chunk at a time: each node's P0 block, its ten beams' other
  // ends (kDeadBeam where the beam is absent or dead), its orientation, cells' e and g
  constexpr int kV = 128;
  constexpr uint32_t kDeadBeam = 0xfffffffeu;
  __shared__ int s_blk[kV], s_v[kV], s_sidx[kV];
  __shared__ uint8_t s_o[kV];
  __shared__ uint32_t s_wf[10][kV];
  __shared__ double s_e[2][kV], s_g[2][kV];
  __shared__ int s_c[2][kV];
  if ((int64_t)blockIdx.x >= n1) return;
  const int I = rows ? rows[blockIdx.x] : (int)blockIdx.x;
  for (int h = threadIdx.x; h < kH; h += blockDim.x) keys[h] = -1;
  __syncthreads();
  if (threadIdx.x == 0) hinsert<kH>(keys, I);
  const int k0 = pt_ptr[I], nv = pt_ptr[I + 1] - k0;
  auto load_chunk = [&](int base, int nvc) {
    __syncthreads();
    for (int t = threadIdx.x; t < nvc; t += blockDim.x) {
      const int blk = pt_blk[k0 + base + t];
      const int v = pblk_row[blk];
      const int o = nkind[v] >> 1;
      s_blk[t] = blk;
      s_v[t] = v;
      s_o[t] = (uint8_t)o;
      int sidx = 0;
      for (int side = 0; side < 2; side++) {
        const int c = ncell[2 * v + side];
        const int f = side == 0 ? 2 * o + 1 : 2 * o;
        const uint32_t mask = c >= 0 ? (uint32_t)cmask[c] : 0u;
        for (int jj = 0; jj < 5; jj++) {
          const int q = face_beam_q(f, jj);
          s_wf[5 * side + jj][t] =
              (mask & (1u << q)) ? cfaces[6 * c + (kPairA[q] == f ? kPairB[q] : kPairA[q])] : kDeadBeam;
        }
        s_c[side][t] = c;
        if (c >= 0) {
          s_e[side][t] = ce[c];
          s_g[side][t] = cg[c];
          sidx = (int)cstruct[c];
        }
      }
      s_sidx[t] = sidx;
    }
    __syncthreads();
  };
  // columns of the row: the P0 rows of the column's nodes and of their beams' other ends
  for (int base = 0; base < nv; base += kV) {
    const int nvc = min(kV, nv - base);
    load_chunk(base, nvc);
    for (int it = threadIdx.x; it < nvc * 10; it += blockDim.x) {
      const int vi = it % nvc, j = it / nvc;
      if (j == 0) {
        const int v = s_v[vi];
        for (int k2 = prow[v]; k2 < prow[v + 1]; k2++) hinsert<kH>(keys, pcol[k2]);
      }
      const uint32_t wf = s_wf[j][vi];
      if (wf == kDeadBeam || (wf & kNotFree)) continue;
      for (int k2 = prow[wf]; k2 < prow[wf + 1]; k2++) hinsert<kH>(keys, pcol[k2]);
    }
  }
  const int m = row_columns<kH>(pass, I, keys, slot, sorted, kSlots, &ncount);
  if (pass == 0) {
    if (threadIdx.x == 0) rowcnt[I] = m;
    return;
  }
  if (m > kSlots) {
    if (threadIdx.x == 0) atomicMax(overflow, m);
    return;
  }
  for (int x = threadIdx.x; x < 36 * m; x += blockDim.x) acc[x] = 0.0;
  // products split by groups of kRG output rows: thread task (row group, node, j), j < 10
  // the block K[v, w] of beam j's far end w, j = 10 the node's own block K[v, v], the sum
  // over its beams (one pass over P0's row of v for all ten); each task reads a column
  // block P0[., J] once for its kRG rows
  constexpr int NRG = 6 / kRG;
  for (int base = 0; base < nv; base += kV) {
    const int nvc = min(kV, nv - base);
    load_chunk(base, nvc);
    for (int it = threadIdx.x; it < nvc * 11 * NRG; it += blockDim.x) {
      const int rg = it % NRG, rest = it / NRG;
      const int vi = rest % nvc, j = rest / nvc;
      const int o = s_o[vi];
      const double* PvI = pval + 36 * (int64_t)s_blk[vi];
      double pc[kRG][6];  // columns rg kRG .. of P0[v, I]
      for (int rr = 0; rr < kRG; rr++)
        for (int i = 0; i < 6; i++) pc[rr][i] = PvI[6 * i + rg * kRG + rr];
      const BeamGeo G = bgeo[s_sidx[vi]];
      double qr[kRG][6];  // rows rg kRG .. of P0[v,I]^T K[v, row]
      int row;
      if (j < 10) {
        const uint32_t wf = s_wf[j][vi];
        if (wf == kDeadBeam || (wf & kNotFree)) continue;
        const int side = j / 5, f = side == 0 ? 2 * o + 1 : 2 * o;
        const int q = face_beam_q(f, j % 5);
        const int sv = kPairA[q] == f ? -1 : 1;
        double n[3];
        int cls;
        pair_n(q, G, n, cls);
        const double rho = beam_sec(csec, s_c[side][vi], q);
        for (int rr = 0; rr < kRG; rr++)
          beam_row(pc[rr], sv, -sv, n, G.hd[cls], G.K4[cls] * rho, G.R2[cls] * (rho * rho), s_e[side][vi],
                   s_g[side][vi], qr[rr]);
        row = (int)wf;
      } else {
        bool any = false;
        for (int rr = 0; rr < kRG; rr++)
          for (int k = 0; k < 6; k++) qr[rr][k] = 0.0;
        for (int jb = 0; jb < 10; jb++) {
          if (s_wf[jb][vi] == kDeadBeam) continue;
          any = true;
          const int side = jb / 5, f = side == 0 ? 2 * o + 1 : 2 * o;
          const int q = face_beam_q(f, jb % 5);
          const int sv = kPairA[q] == f ? -1 : 1;
          double n[3];
          int cls;
          pair_n(q, G, n, cls);
          const double rho = beam_sec(csec, s_c[side][vi], q);
          for (int rr = 0; rr < kRG; rr++) {
            double t[6];
            beam_row(pc[rr], sv, sv, n, G.hd[cls], G.K4[cls] * rho, G.R2[cls] * (rho * rho), s_e[side][vi],
                     s_g[side][vi], t);
            for (int k = 0; k < 6; k++) qr[rr][k] += t[k];
          }
        }
        if (!any) continue;
        row = s_v[vi];
      }
      // start each task's column loop at a different column, so that the tasks of a
      // wave (mostly nodes of one aggregate) do not add into the same block at once
      const int k2a = prow[row], len = prow[row + 1] - k2a;
      for (int kk = 0; kk < len; kk++) {
        const int k2 = k2a + (kk + rest) % len;
        const int J = pcol[k2];
        if (J < I) continue;  // the lower blocks are the transposes of other rows' upper ones
        const int sl = hfind<kH>(keys, slot, J);
        const double* Pj = pval + 36 * (int64_t)k2;
        double o6[kRG][6];
        for (int rr = 0; rr < kRG; rr++)
          for (int cc = 0; cc < 6; cc++) o6[rr][cc] = 0.0;
        for (int k = 0; k < 6; k++)
          for (int cc = 0; cc < 6; cc++) {
            const double pk = Pj[6 * k + cc];
            for (int rr = 0; rr < kRG; rr++) o6[rr][cc] += qr[rr][k] * pk;
          }
        double* out = acc + 36 * sl + 6 * rg * kRG;
        for (int rr = 0; rr < kRG; rr++)
          for (int cc = 0; cc < 6; cc++) atomicAdd(&out[6 * rr + cc], o6[rr][cc]);
      }
    }
  }
  __syncthreads();
  const int base = rowptr[I];
  if (threadIdx.x == 0) col[base] = I;
  for (int h = threadIdx.x; h < kH; h += blockDim.x)
    if (keys[h] != -1 && keys[h] != I) col[base + slot[h]] = keys[h];
  for (int x = threadIdx.x; x < 36 * m; x += blockDim.x) A[36 * (int64_t)base + x] = acc[x];
}

__global__ void k_blk_rows(int64_t n, const int* prow, int* blk_row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = prow[i]; k < prow[i + 1]; k++) blk_row[k] = (int)i;
}

// V-cycle transfers with an explicit P. restrict: c[J] = sum over column J of P^T t.
// With x0 not NULL, also x0 = w D^-1 c for the coarse level (as k_restrict_c).
template <class T>
__global__ void k_restrict_p(int64_t nJ, const uint32_t* ast, const uint8_t* done, const int* pt_ptr, const int* pt_blk,
                             const int* pblk_row, const T* pval, const T* t, T* c, const T* dinv, const double* step,
                             T* x0) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t J = g >> 3;
  const int a = (int)(g & 7);
  // the group of 8 lanes of one J stays whole (lanes 6 and 7 idle) for the shuffles below
  if (J >= nJ) return;
  if (done && done[ast[J]]) return;
  // column a of each block of column J of P, four blocks' loads issued together
  T acc = 0;
  int kk = a < 6 ? pt_ptr[J] : 0;
  const int k1 = a < 6 ? pt_ptr[J + 1] : 0;
  for (; kk + 4 <= k1; kk += 4) {
    int blk[4], row[4];
    T tv[4][6], mv[4][6];
#pragma unroll
    for (int q = 0; q < 4; q++) blk[q] = pt_blk[kk + q];
#pragma unroll
    for (int q = 0; q < 4; q++) row[q] = pblk_row[blk[q]];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      load6<T>(t + 6 * (int64_t)row[q], tv[q]);
#pragma unroll
      for (int b = 0; b < 6; b++) mv[q][b] = pval[36 * (int64_t)blk[q] + 6 * b + a];
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int b = 0; b < 6; b++) acc += mv[q][b] * tv[q][b];
  }
  for (; kk < k1; kk++) {
    const int blk = pt_blk[kk];
    const T* M = pval + 36 * (int64_t)blk + a;
    T tv[6];
    load6<T>(t + 6 * (int64_t)pblk_row[blk], tv);
#pragma unroll
    for (int b = 0; b < 6; b++) acc += M[6 * b] * tv[b];
  }
  if (a < 6) c[6 * J + a] = acc;
  if (x0) {
    T all[6];
    group_gather6<T>(acc, all);
    if (a < 6) x0[6 * J + a] = (T)step[ast[J]] * row6_dot<T>(dinv + 36 * J + 6 * a, all);
  }
}
// The same restriction for a level of few aggregates J (each with a long column of P):
// one wave per J, its lanes in ten groups of six, group g taking blocks g, g + 10, ... of
// the column; the groups' sums meet in LDS and are added in group order. U: the
// coarse level's precision (wider than T where the level above switches to FP64: the
// sums are widened as a separate conversion would, and x0 formed as k_c_jacobi does).
template <class T, class U = T>
__global__ __launch_bounds__(256) void k_restrict_pw(int64_t nJ, const uint32_t* ast, const uint8_t* done,
                                                     const int* pt_ptr, const int* pt_blk, const int* pblk_row,
                                                     const T* pval, const T* t, U* c, const U* dinv,
                                                     const double* step, U* x0) {
  __shared__ T part[4][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t J = (int64_t)blockIdx.x * 4 + wv;
  const bool live = J < nJ && !(done && done[ast[J]]);
  const int grp = lane / 6, a = lane - 6 * grp;
  T acc = 0;
  if (live && grp < 10) {
    const int k1 = pt_ptr[J + 1];
    for (int kk = pt_ptr[J] + grp; kk < k1; kk += 10) {
      const int blk = pt_blk[kk];
      const T* M = pval + 36 * (int64_t)blk + a;
      T tv[6];
      load6<T>(t + 6 * (int64_t)pblk_row[blk], tv);
#pragma unroll
      for (int b = 0; b < 6; b++) acc += M[6 * b] * tv[b];
    }
  }
  part[wv][lane] = acc;
  __syncthreads();
  if (!live || lane >= 6) return;
  U all[6];
#pragma unroll
  for (int j = 0; j < 6; j++) {
    T sum = 0;
    for (int g2 = 0; g2 < 10; g2++) sum += part[wv][6 * g2 + j];
    all[j] = (U)sum;
  }
  c[6 * J + lane] = all[lane];
  if (x0) x0[6 * J + lane] = row6_dot<U>(dinv + 36 * J + 6 * lane, all) * (U)step[ast[J]];
}
// The transfers of a large FP32 level through PBlock32 records, G lanes per aggregate,
// each lane whole blocks in turn (ten 16-byte loads and the other end's six entries),
// the lanes' sums meeting by a butterfly; lanes 0-5 finish as k_restrict_p and
// k_prolong_p do. Restriction: c[J] = sum over column J of P^T t (x0 = w D^-1 c when set).
typedef float st_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load_pblock(const PBlock32* p, float* a, int& idx) {
  const st_f4v* q = reinterpret_cast<const st_f4v*>(p);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const st_f4v v = q[k];
    a[4 * k] = v.x;
    a[4 * k + 1] = v.y;
    a[4 * k + 2] = v.z;
    a[4 * k + 3] = v.w;
  }
  idx = __builtin_bit_cast(int, q[9].x);
}
// (the records in FP16: HalfBlock, the other end's index in its column field)
__device__ __forceinline__ void load_pblock(const HalfBlock* p, float* a, int& idx) {
  st_u4 q[5];
  load_half_block(p, q);
  unsigned w[20];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    w[4 * k] = q[k].x;
    w[4 * k + 1] = q[k].y;
    w[4 * k + 2] = q[k].z;
    w[4 * k + 3] = q[k].w;
  }
  const float sc = __builtin_bit_cast(float, w[18]);
#pragma unroll
  for (int e = 0; e < 36; e++) {
    const st_h2 h = __builtin_bit_cast(st_h2, w[e >> 1]);
    a[e] = (float)((e & 1) ? h.y : h.x) * sc;
  }
  idx = (int)w[19];
}
template <int G, class PB>
__global__ __launch_bounds__(256) void k_restrict_rec(int64_t nJ, const uint32_t* ast, const uint8_t* done,
                                                      const int* pt_ptr, const PB* pc, const float* t, float* c,
                                                      const float* dinv, const double* step, float* x0) {
  const int gl = threadIdx.x & (G - 1);
  const int64_t J = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const bool live = J < nJ && !(done && done[ast[J]]);
  float y[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (live) {
    const int k1 = pt_ptr[J + 1];
    for (int kk = pt_ptr[J] + gl; kk < k1; kk += G) {
      float a[36], tv[6];
      int i;
      load_pblock(pc + kk, a, i);
      load6<float>(t + 6 * (int64_t)i, tv);
#pragma unroll
      for (int b = 0; b < 6; b++)
#pragma unroll
        for (int j = 0; j < 6; j++) y[j] += a[6 * b + j] * tv[b];
    }
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1)
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] += __shfl_xor(y[k], o, 64);
  if (!live || gl >= 6) return;
  c[6 * J + gl] = y[gl];
  if (x0) x0[6 * J + gl] = (float)step[ast[J]] * row6_dot<float>(dinv + 36 * J + 6 * gl, y);
}
// Prolongation: x[i] += sum over row i of P y.
template <int G, class PB>
__global__ __launch_bounds__(256) void k_prolong_rec(int64_t n, const uint32_t* ast, const uint8_t* done,
                                                     const int* prow, const PB* pr, const float* y, float* x) {
  const int gl = threadIdx.x & (G - 1);
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const bool live = i < n && !(done && done[ast[i]]);
  float acc[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (live) {
    const int k1 = prow[i + 1];
    for (int k = prow[i] + gl; k < k1; k += G) {
      float a[36], yv[6];
      int J;
      load_pblock(pr + k, a, J);
      load6<float>(y + 6 * (int64_t)J, yv);
#pragma unroll
      for (int r = 0; r < 6; r++)
#pragma unroll
        for (int j = 0; j < 6; j++) acc[r] += a[6 * r + j] * yv[j];
    }
  }
#pragma unroll
  for (int o = 1; o < G; o <<= 1)
#pragma unroll
    for (int k = 0; k < 6; k++) acc[k] += __shfl_xor(acc[k], o, 64);
  if (!live || gl >= 6) return;
  x[6 * i + gl] += acc[gl];
}
// The records: rows (block k in storage order, its column) or columns (position kk of
// the column grouping, block pt_blk[kk], its row).
__device__ __forceinline__ void fill_pblock(const float* a, int idx, PBlock32& r) {
  for (int e = 0; e < 36; e++) r.a[e] = a[e];
  r.idx = idx;
  r.pad[0] = r.pad[1] = r.pad[2] = 0;
}
__device__ __forceinline__ void fill_pblock(const float* a, int idx, HalfBlock& r) {
  float m = 0.0f;
  for (int e = 0; e < 36; e++) m = fmaxf(m, fabsf(a[e]));
  const float sc = m > 0.0f ? m / 16384.0f : 1.0f;
  for (int e = 0; e < 36; e++) r.a[e] = (_Float16)(a[e] / sc);
  r.scale = sc;
  r.col = idx;
}
template <class PB>
__global__ void k_pblocks_rows(int64_t nb, const int* pcol, const float* pval, PB* out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  PB r;
  fill_pblock(pval + 36 * k, pcol[k], r);
  out[k] = r;
}
template <class PB>
__global__ void k_pblocks_cols(int64_t nb, const int* pt_blk, const int* pblk_row, const float* pval, PB* out) {
  const int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= nb) return;
  const int blk = pt_blk[kk];
  PB r;
  fill_pblock(pval + 36 * (int64_t)blk, pblk_row[blk], r);
  out[kk] = r;
}
// prolong: x[i] += sum over row i of P y
template <class T, class TY = T>
__global__ void k_prolong_p(int64_t n, const uint32_t* ast, const uint8_t* done, const int* prow, const int* pcol,
                            const T* pval, const TY* y, T* x) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = g >> 3;
  const int a = (int)(g & 7);
  if (i >= n || a >= 6) return;
  if (done && done[ast[i]]) return;
  x[6 * i + a] += block_row_dot<T, TY>(prow[i], prow[i + 1], pcol, pval, y, a);
}

// rigid prolongation level 1 -> fine, overwriting: x[v] = Sh(r_v) y[label[v]] (0 off aggregates)
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong1_set(const uint32_t* nbase, const uint32_t* bstruct,
                                                         const uint8_t* done, const int* label, const double* centre,
                                                         const uint32_t* nkey, const uint8_t* nkind, const BeamGeo* bgeo,
                                                         const T* y, T* x) {
  const uint32_t b = blockIdx.x, s = bstruct[b];
  if (done && done[s]) return;
  const double sp = 2.0 * bgeo[s].hd[0];
  for (uint32_t v = nbase[b] + threadIdx.x; v < nbase[b + 1]; v += blockDim.x) {
    const int I = label[v];
    T xv[6] = {0, 0, 0, 0, 0, 0};
    if (I >= 0) {
      double p[3];
      node_pos(v, nkey, nkind, sp, p);
      for (int k = 0; k < 3; k++) p[k] -= centre[3 * (int64_t)I + k];
      T yv[6];
      load6<T>(y + 6 * (int64_t)I, yv);
      shift_add<T>(yv, p, xv);
    }
    store6<T>(x + 6 * (size_t)v, xv);
  }
}

template <class T>
T* ptr(DBuf& b) {
  return b.as<T>();
}

}  // namespace

// ============================================================== host side

// A captured launch sequence, replayed: run(stream, f) captures f's launches on the
// first call (f must only enqueue work on the stream) and launches the graph.
struct GraphExec {
  hipGraphExec_t exec = nullptr;
  template <class F>
  void run(hipStream_t st, F f) {
    if (!exec) {
      ST_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        f();
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
      }
      hipGraph_t gr = nullptr;
      ST_HIP(hipStreamEndCapture(st, &gr));
      const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      ST_HIP(e);
    }
    ST_HIP(hipGraphLaunch(exec, st));
  }
  ~GraphExec() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

struct Solver {
  Ctx& c;
  hipStream_t st;
  int ns;
  DevGeo& g;
  std::vector<int64_t> sbrick;  // per structure: brick offsets (ns+1)
  DBuf sbrick_d, node0_d;
  DBuf sstatus, siters, srefresh, swant, sbad;
  std::vector<std::vector<int64_t>> lseg;  // per level: aggregate offsets per structure
  std::vector<DBuf*> lseg_d;
  DBuf seg_tmp;
  int64_t ndof = 0;
  bool mixed = false;  // the solve's precision: the FP32 levels need no FP64 copies afterwards
  explicit Solver(Ctx& cc) : c(cc), st(cc.stream), ns(cc.geo.nstruct), g(cc.geo) {}
  ~Solver() {
    sbrick_d.release();
    node0_d.release();
    sstatus.release();
    siters.release();
    srefresh.release();
    swant.release();
    sbad.release();
    seg_tmp.release();
  }

  void reduce_bricks(DBuf& partial, int slot) {
    k_reduce_seg<<<ns, kReduceThreads, 0, st>>>(partial.as<double>(), sbrick_d.as<int64_t>(), c.scal.as<double>(), kScal, slot);
  }

  void scan(int* in, int* out, int64_t n) {
    size_t bytes = 0;
    ST_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st));
    c.scan_tmp.ensure(bytes + 16);
    ST_HIP(hipcub::DeviceScan::ExclusiveSum(c.scan_tmp.p, bytes, in, out, (int)n, st));
  }
  void sort_pairs(const int* kin, int* kout, const int* vin, int* vout, int64_t n) {
    size_t bytes = 0;
    ST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, 32, st));
    c.sort_tmp.ensure(bytes + 16);
    ST_HIP(hipcub::DeviceRadixSort::SortPairs(c.sort_tmp.p, bytes, kin, kout, vin, vout, (int)n, 0, 32, st));
  }
  int read_int(const int* d) {
    int h = 0;
    ST_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    return h;
  }

  void init_struct_arrays() {
    sbrick.assign(ns + 1, 0);
    std::vector<int64_t> node0(ns + 1, 0);
    for (int s = 0; s < ns; s++) {
      sbrick[s] = c.host.structs[s].brick0;
      node0[s] = c.host.structs[s].node0;
    }
    sbrick[ns] = g.nbricks;
    node0[ns] = g.nnodes;
    sbrick_d.ensure((ns + 1) * sizeof(int64_t));
    node0_d.ensure((ns + 1) * sizeof(int64_t));
    ST_HIP(hipMemcpyAsync(sbrick_d.p, sbrick.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    ST_HIP(hipMemcpyAsync(node0_d.p, node0.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    sstatus.ensure(ns * sizeof(int));
    siters.ensure(ns * sizeof(int));
    srefresh.ensure(ns * sizeof(int));
    swant.ensure(ns);
    sbad.ensure(sizeof(int));
    c.scal.ensure((size_t)ns * kScal * sizeof(double));
    c.done.ensure(ns);
    c.partial.ensure(std::max<int64_t>(g.nbricks, 1) * sizeof(double));
    ST_HIP(hipMemsetAsync(c.scal.p, 0, (size_t)ns * kScal * sizeof(double), st));
    ST_HIP(hipMemsetAsync(srefresh.p, 0, ns * sizeof(int), st));
  }

  // ---------------------------------------------------------------- setup
  void setup_fine() {
    const int64_t nn = g.nnodes;
    c.dinv64.ensure(std::max<int64_t>(nn, 1) * kSym * sizeof(double));
    c.dinv32.ensure(std::max<int64_t>(nn, 1) * sizeof(SymHalf));
    ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
    if (nn)
      k_dinv<<<nblk(nn), 256, 0, st>>>(nn, g.nflags.as<uint8_t>(), g.ncell.as<int32_t>(), g.nkind.as<uint8_t>(),
                                       g.ce64.as<double>(), g.cg64.as<double>(), g.csec64.as<double>(), g.cmask.as<uint16_t>(),
                                       g.cstruct.as<uint32_t>(), g.bgeo.as<BeamGeo>(), c.dinv64.as<double>(),
                                       c.dinv32.as<SymHalf>(), sbad.as<int>());
    ST_HIP(hipGetLastError());
    if (read_int(sbad.as<int>())) throw StError{ST_NOT_SPD, "a free node has a singular stiffness block"};
    // smoother step: 4/3 over the greatest eigenvalue of D^-1 K (statics/src/bodies.rs:174, 180-190)
    const size_t vb = (size_t)std::max<int64_t>(nn, 1) * 6 * sizeof(double);
    c.t64.ensure(vb);
    c.u64.ensure(vb);
    c.z64.ensure(vb);
    double* v = c.t64.as<double>();
    double* w = c.u64.as<double>();
    double* tmp = c.z64.as<double>();
    if (nn) {
      k_c_init<<<nblk(nn), 256, 0, st>>>(nn, v);
      k_mask_rows<<<nblk(nn), 256, 0, st>>>(nn, g.nflags.as<uint8_t>(), v);
    }
    const int iters = 10;
    std::vector<double> nrm_prev(ns, 0.0), nrm(ns, 0.0);
    for (int it = 0; it < iters; it++) {
      EpiStore<double> e;
      e.y = tmp;
      launch_apply<double>(c, v, e);
      if (g.nbricks)
        k_jacobi_fine<double><<<(unsigned)g.nbricks, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), g.bstruct.as<uint32_t>(),
                                                                      nullptr, c.dinv64.as<double>(), nullptr, tmp, w);
      if (it >= iters - 2) {
        if (g.nbricks)
          k_dot_fine<double><<<(unsigned)g.nbricks, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), w, w, c.partial.as<double>());
        reduce_bricks(c.partial, it == iters - 2 ? S_LAM_A : S_LAM_B);
      }
      std::swap(v, w);
    }
    std::vector<double> sc((size_t)ns * kScal);
    ST_HIP(hipMemcpyAsync(sc.data(), c.scal.p, sc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    c.step0.assign(ns, 1.0);
    for (int s = 0; s < ns; s++) {
      const double a = sc[(size_t)s * kScal + S_LAM_A], b = sc[(size_t)s * kScal + S_LAM_B];
      const double lam = (a > 0 && b > 0) ? std::sqrt(b / a) : 1.0;
      c.step0[s] = 4.0 / 3.0 / lam;
    }
    c.step0_d.ensure(ns * sizeof(double));
    ST_HIP(hipMemcpyAsync(c.step0_d.p, c.step0.data(), ns * sizeof(double), hipMemcpyHostToDevice, st));
  }

  // Labels at cube level L over the fine union-find (already holding level L-1's unions).
  // Returns the aggregate count; ids/rep/label written.
  int64_t label_level(int L, int* label, DBuf& rep, std::vector<int64_t>& seg) {
    const int64_t nn = g.nnodes;
    int* uf = c.uf.as<int>();
    if (g.ncells)
      k_unite_level<<<nblk(g.ncells), 256, 0, st>>>(g.ncells, g.cfaces.as<uint32_t>(), g.cmask.as<uint16_t>(),
                                                    g.nkey.as<uint32_t>(), L, uf);
    k_compress<<<nblk(nn), 256, 0, st>>>(nn, uf);
    int* flag = c.flag.as<int>();
    int* ids = c.uf2.as<int>();
    k_root_flags<<<nblk(nn), 256, 0, st>>>(nn, uf, g.nflags.as<uint8_t>(), flag);
    scan(flag, ids, nn);
    seg.assign(ns + 1, 0);
    seg_tmp.ensure((ns + 1) * sizeof(int64_t));
    k_seg_from_ids<<<nblk(ns + 1), 256, 0, st>>>(ns, node0_d.as<int64_t>(), ids, flag, nn, seg_tmp.as<int64_t>());
    ST_HIP(hipMemcpyAsync(seg.data(), seg_tmp.p, (ns + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    const int64_t count = seg[ns];
    rep.ensure(std::max<int64_t>(count, 1) * sizeof(int));
    k_label<<<nblk(nn), 256, 0, st>>>(nn, uf, g.nflags.as<uint8_t>(), ids, label, rep.as<int>());
    ST_HIP(hipGetLastError());
    return count;
  }

  void build_rows_a1(Level& lv) {
    const int64_t n = lv.n;
    DBuf cnt;
    cnt.ensure((n + 1) * sizeof(int));
    lv.rowptr.ensure((n + 1) * sizeof(int));
    ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
    ST_HIP(hipMemsetAsync(cnt.p, 0, (n + 1) * sizeof(int), st));
    auto launch = [&](int pass) {
      k_a1<<<nblk(n, 64), 64, 0, st>>>(pass, n, c.agg_ptr.as<int>(), c.nodes_by_agg.as<int>(), c.label1.as<int>(),
                                      lv.centre.as<double>(), lv.agg_struct.as<uint32_t>(), g.nkey.as<uint32_t>(),
                                      g.nkind.as<uint8_t>(), g.ncell.as<int32_t>(), g.cfaces.as<uint32_t>(),
                                      g.ce64.as<double>(), g.cg64.as<double>(), g.csec64.as<double>(), g.cmask.as<uint16_t>(),
                                      g.bgeo.as<BeamGeo>(), cnt.as<int>(), lv.rowptr.as<int>(), lv.col.as<int>(),
                                      lv.a64.as<double>(), sbad.as<int>());
      ST_HIP(hipGetLastError());
    };
    launch(0);
    finish_rows(lv, cnt, launch);
    cnt.release();
  }

  template <class F>
  void finish_rows(Level& lv, DBuf& cnt, F launch) {
    const int64_t n = lv.n;
    if (read_int(sbad.as<int>())) throw StError{ST_DEVICE, "an aggregate has more neighbours than the row capacity"};
    scan(cnt.as<int>(), lv.rowptr.as<int>(), n + 1);
    int nb = 0;
    ST_HIP(hipMemcpyAsync(&nb, lv.rowptr.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    lv.nblocks = nb;
    lv.col.ensure(std::max(nb, 1) * sizeof(int));
    lv.a64.ensure(std::max(nb, 1) * 36 * sizeof(double));
    lv.a32.ensure(std::max(nb, 1) * 36 * sizeof(float));
    k_fill_neg<<<nblk(nb), 256, 0, st>>>(nb, lv.col.as<int>());
    launch(1);
    k_to_f32<<<nblk((int64_t)nb * 36), 256, 0, st>>>((int64_t)nb * 36, lv.a64.as<double>(), lv.a32.as<float>());
    lv.dinv64.ensure(std::max<int64_t>(n, 1) * 36 * sizeof(double));
    lv.dinv32.ensure(std::max<int64_t>(n, 1) * 36 * sizeof(float));
    ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
    k_coarse_dinv<<<nblk(n), 256, 0, st>>>(n, lv.rowptr.as<int>(), lv.a64.as<double>(), lv.dinv64.as<double>(),
                                           lv.dinv32.as<float>(), sbad.as<int>());
    ST_HIP(hipGetLastError());
    if (read_int(sbad.as<int>())) throw StError{ST_NOT_SPD, "a coarse diagonal block is singular"};
  }

  // Two-pass construction of a block-sparse prolongator (rows nrows, columns ncols) and of
  // its blocks grouped by column (stable sort by column keeps rows ascending within a column).
  template <class F>
  int64_t build_p(int64_t nrows, int64_t ncols, DBuf& prow, DBuf& pcol, DBuf& pval, DBuf& pt_ptr, DBuf& pt_blk,
                  DBuf& pblk_row, F launch) {
    DBuf cnt;
    cnt.ensure((nrows + 1) * sizeof(int));
    prow.ensure((nrows + 1) * sizeof(int));
    ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
    ST_HIP(hipMemsetAsync(cnt.p, 0, (nrows + 1) * sizeof(int), st));
    launch(0, cnt.as<int>(), prow.as<int>(), (int*)nullptr, (double*)nullptr);
    if (read_int(sbad.as<int>())) throw StError{ST_DEVICE, "a prolongator row has more blocks than its capacity"};
    scan(cnt.as<int>(), prow.as<int>(), nrows + 1);
    int nb = 0;
    ST_HIP(hipMemcpyAsync(&nb, prow.as<int>() + nrows, sizeof(int), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    pcol.ensure(std::max(nb, 1) * sizeof(int));
    pval.ensure((size_t)std::max(nb, 1) * 36 * sizeof(double));
    launch(1, cnt.as<int>(), prow.as<int>(), pcol.as<int>(), pval.as<double>());
    ST_HIP(hipGetLastError());
    cnt.release();
    // blocks by column
    pblk_row.ensure(std::max(nb, 1) * sizeof(int));
    k_blk_rows<<<nblk(nrows), 256, 0, st>>>(nrows, prow.as<int>(), pblk_row.as<int>());
    DBuf vals, skeys;
    vals.ensure(std::max(nb, 1) * sizeof(int));
    skeys.ensure(std::max(nb, 1) * sizeof(int));
    pt_blk.ensure(std::max(nb, 1) * sizeof(int));
    pt_ptr.ensure((ncols + 1) * sizeof(int));
    k_iota<<<nblk(nb), 256, 0, st>>>(nb, vals.as<int>());
    sort_pairs(pcol.as<int>(), skeys.as<int>(), vals.as<int>(), pt_blk.as<int>(), nb);
    k_fill_int<<<nblk(ncols + 1), 256, 0, st>>>(ncols + 1, pt_ptr.as<int>(), 0);
    k_seg_ptr<<<nblk(nb), 256, 0, st>>>(nb, skeys.as<int>(), pt_ptr.as<int>(), (int)ncols);
    ST_HIP(hipGetLastError());
    vals.release();
    skeys.release();
    return nb;
  }

  int max_int(const int* d, int64_t n) {
    DBuf out;
    out.ensure(sizeof(int));
    size_t bytes = 0;
    ST_HIP(hipcub::DeviceReduce::Max(nullptr, bytes, d, out.as<int>(), (int)n, st));
    c.scan_tmp.ensure(bytes + 16);
    ST_HIP(hipcub::DeviceReduce::Max(c.scan_tmp.p, bytes, d, out.as<int>(), (int)n, st));
    const int h = read_int(out.as<int>());
    out.release();
    return h;
  }

  // Rows of a Galerkin product, one workgroup per row: launch(pass, slots, cnt, rows,
  // nrows). Pass 0 counts every row's columns; pass 1 runs, for each width class, the
  // kernel instance holding that many columns in LDS over the class's rows.
  template <class F>
  void wave_rows(Level& lv, F launch) {
    const int64_t n = lv.n;
    DBuf cnt, lists, ctr;
    cnt.ensure((n + 1) * sizeof(int));
    lv.rowptr.ensure((n + 1) * sizeof(int));
    ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
    ST_HIP(hipMemsetAsync(cnt.p, 0, (n + 1) * sizeof(int), st));
    launch(0, 0, cnt.as<int>(), (const int*)nullptr, n);
    const int widest = max_int(cnt.as<int>(), n);
    if (widest > kRowClass[3]) throw StError{ST_DEVICE, "a coarse operator row has more than 448 blocks"};
    lists.ensure(4 * n * sizeof(int));
    ctr.ensure(4 * sizeof(int));
    ST_HIP(hipMemsetAsync(ctr.p, 0, 4 * sizeof(int), st));
    k_row_classes<<<nblk(n), 256, 0, st>>>(n, cnt.as<int>(), lists.as<int>(), ctr.as<int>());
    int count[4];
    ST_HIP(hipMemcpyAsync(count, ctr.p, sizeof(count), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    finish_rows(lv, cnt, [&](int pass) {
      if (pass == 1) {
        for (int k = 0; k < 4; k++)
          if (count[k]) launch(1, kRowClass[k], cnt.as<int>(), lists.as<int>() + k * n, (int64_t)count[k]);
        k_fill_lower<<<nblk(n), 256, 0, st>>>(n, lv.rowptr.as<int>(), lv.col.as<int>(), lv.a64.as<double>());
      }
    });
    cnt.release();
    lists.release();
    ctr.release();
  }

  // Level 1 of smoothed aggregation: P0 at the fine nodes, then A1 = P0^T K P0.
  void build_sa_level1(Level& lv) {
    const int64_t nn = g.nnodes, n = lv.n;
    c.p0blocks = build_p(nn, n, c.p0row, c.p0col, c.p0val, c.p0t_ptr, c.p0t_blk, c.p0blk_row,
                         [&](int pass, int* cnt, int* prow, int* pcol, double* pval) {
                           k_p0<<<nblk(nn), 256, 0, st>>>(pass, nn, g.nflags.as<uint8_t>(), c.label1.as<int>(),
                                                          lv.centre.as<double>(), g.nkey.as<uint32_t>(),
                                                          g.nkind.as<uint8_t>(), g.ncell.as<int32_t>(),
                                                          g.cfaces.as<uint32_t>(), g.ce64.as<double>(),
                                                          g.cg64.as<double>(), g.csec64.as<double>(), g.cmask.as<uint16_t>(),
                                                          g.cstruct.as<uint32_t>(), g.bgeo.as<BeamGeo>(),
                                                          c.dinv64.as<double>(), c.step0_d.as<double>(), cnt, prow,
                                                          pcol, pval, sbad.as<int>());
                         });
    // output rows per task of the A1 product (STATICS_HIP_A1_ROWS: 1, 2, 3 or 6): three rows
    // read each P0 block once for three; at 256^3 two and three rows within 1% of each other,
    // six rows (213 VGPRs, two waves per SIMD) 60 ms slower
    int a1_rows = 3;
    if (const char* e = std::getenv("STATICS_HIP_A1_ROWS")) a1_rows = std::atoi(e);
    wave_rows(lv, [&](int pass, int slots, int* cnt, const int* rows, int64_t nr) {
#define ST_A1(P, S, H, RG)                                                                                               \
  k_sa_a1_w<P, S, H, RG><<<(unsigned)nr, 256, 0, st>>>(rows, nr, c.p0t_ptr.as<int>(), c.p0t_blk.as<int>(),          \
                                               c.p0blk_row.as<int>(), c.p0row.as<int>(), c.p0col.as<int>(),         \
                                               c.p0val.as<double>(), g.nkind.as<uint8_t>(), g.ncell.as<int32_t>(),  \
                                               g.cfaces.as<uint32_t>(), g.ce64.as<double>(), g.cg64.as<double>(),   \
                                               g.csec64.as<double>(), g.cmask.as<uint16_t>(), g.cstruct.as<uint32_t>(), \
                                               g.bgeo.as<BeamGeo>(), cnt, lv.rowptr.as<int>(), lv.col.as<int>(),    \
                                               lv.a64.as<double>(), sbad.as<int>())
      if (pass == 0)
        ST_A1(0, 1, 2048, 1);
      else if (slots <= 64) {
        if (a1_rows == 6)
          ST_A1(1, 64, 256, 6);
        else if (a1_rows == 3)
          ST_A1(1, 64, 256, 3);
        else if (a1_rows == 2)
          ST_A1(1, 64, 256, 2);
        else
          ST_A1(1, 64, 256, 1);
      } else if (slots <= 128)
        ST_A1(1, 128, 512, 1);
      else if (slots <= 256)
        ST_A1(1, 256, 1024, 1);
      else
        ST_A1(1, 448, 2048, 1);
#undef ST_A1
      ST_HIP(hipGetLastError());
    });
    // the explicit fine prolongator only serves this product (the V-cycle applies it matrix-free)
    for (DBuf* b : {&c.p0row, &c.p0col, &c.p0val, &c.p0t_ptr, &c.p0t_blk, &c.p0blk_row}) b->release();
  }

  // Above level 1: P of the level below (smoothed, or rigid when smooth is false), then this level = P^T A P.
  void build_sa_coarse(Level& below, Level& lv, bool smooth) {
    const int64_t nb_rows = below.n, n = lv.n;
    below.pblocks = build_p(nb_rows, n, below.prow, below.pcol, below.pval64, below.pt_ptr, below.pt_blk,
                            below.pblk_row, [&](int pass, int* cnt, int* prow, int* pcol, double* pval) {
                              k_pc<<<nblk(nb_rows), 256, 0, st>>>(pass, smooth ? 1 : 0, nb_rows,
                                                                  below.rowptr.as<int>(),
                                                                  below.col.as<int>(), below.a64.as<double>(),
                                                                  below.dinv64.as<double>(), below.step_d.as<double>(),
                                                                  below.agg_struct.as<uint32_t>(),
                                                                  below.parent.as<int>(), below.centre.as<double>(),
                                                                  lv.centre.as<double>(), cnt, prow, pcol, pval,
                                                                  sbad.as<int>());
                            });
    // output rows per task of the product (STATICS_HIP_GAL_ROWS: 1, 2 or 3): three rows read
    // each block of A and P once for three; at 256^3 the setup's Galerkin products take
    // 58 ms less than with one
    int gal_rows = 3;
    if (const char* e = std::getenv("STATICS_HIP_GAL_ROWS")) gal_rows = std::atoi(e);
    // Tasks by (i, j) pairs or grouped by j (the same terms added in another order): the
    // grouping saves the passes over P's rows where the chunk's nodes share many j, and
    // costs a scan of row j of A per j, the dearer where A's rows are long (at 256^3 level 2
    // by j 37 ms against 64, level 3 by pairs 2 ms against 10). Chosen per level and row
    // width class by timing both the first time this context builds it (kept across
    // uploads: a batch's chunks time it once); STATICS_HIP_GAL_PAIRS=0 or 1 forces.
    const char* gp = std::getenv("STATICS_HIP_GAL_PAIRS");
    const int gal_force = gp ? (std::atoi(gp) ? 1 : 0) : -1;
    below.pval32.ensure((size_t)std::max<int64_t>(below.pblocks, 1) * 36 * sizeof(float));
    k_to_f32<<<nblk(below.pblocks * 36), 256, 0, st>>>(below.pblocks * 36, below.pval64.as<double>(),
                                                       below.pval32.as<float>());
    wave_rows(lv, [&](int pass, int slots, int* cnt, const int* rows, int64_t nr) {
#define ST_GAL(P, S, H, RG)                                                                                              \
  k_sa_gal_w<P, S, H, RG><<<(unsigned)nr, 256, 0, st>>>(rows, nr, below.pt_ptr.as<int>(), below.pt_blk.as<int>(),    \
                                                below.pblk_row.as<int>(), below.prow.as<int>(),                     \
                                                below.pcol.as<int>(), below.pval64.as<double>(),                    \
                                                below.rowptr.as<int>(), below.col.as<int>(), below.a64.as<double>(),\
                                                cnt, lv.rowptr.as<int>(), lv.col.as<int>(), lv.a64.as<double>(),    \
                                                sbad.as<int>(), pairs)
      auto launch = [&](int pairs) {
        if (pass == 0)
          ST_GAL(0, 1, 2048, 1);
        else if (slots <= 64) {
          if (gal_rows == 3)
            ST_GAL(1, 64, 256, 3);
          else if (gal_rows == 2)
            ST_GAL(1, 64, 256, 2);
          else
            ST_GAL(1, 64, 256, 1);
        } else if (slots <= 128) {
          if (gal_rows == 3)
            ST_GAL(1, 128, 512, 3);
          else if (gal_rows == 2)
            ST_GAL(1, 128, 512, 2);
          else
            ST_GAL(1, 128, 512, 1);
        } else if (slots <= 256)
          ST_GAL(1, 256, 1024, 1);
        else
          ST_GAL(1, 448, 2048, 1);
        ST_HIP(hipGetLastError());
      };
      if (pass == 0 || gal_force >= 0) {
        launch(gal_force > 0 ? 1 : 0);
        return;
      }
      const std::array<int64_t, 4> shape{(int64_t)c.levels.size(), (int64_t)slots, 0, 0};
      for (auto& kv : c.gal_modes)
        if (kv.first == shape) {
          launch(kv.second);
          return;
        }
      hipEvent_t e[3];
      for (auto& x : e) ST_HIP(hipEventCreate(&x));
      ST_HIP(hipEventRecord(e[0], st));
      launch(1);
      ST_HIP(hipEventRecord(e[1], st));
      launch(0);
      ST_HIP(hipEventRecord(e[2], st));
      ST_HIP(hipEventSynchronize(e[2]));
      float ms1 = 0, ms0 = 0;
      ST_HIP(hipEventElapsedTime(&ms1, e[0], e[1]));
      ST_HIP(hipEventElapsedTime(&ms0, e[1], e[2]));
      for (auto& x : e) (void)hipEventDestroy(x);
      const int mode = ms1 < ms0 ? 1 : 0;
      if (mode == 1) launch(1);  // the rows as every later build of this shape makes them
      c.gal_modes.emplace_back(shape, mode);
      if (std::getenv("STATICS_HIP_VERBOSE"))
        std::fprintf(stderr, "statics_hip: Galerkin rows of width <= %d (%lld rows): by pairs %.2f ms, by j %.2f ms\n",
                     slots, (long long)nr, ms1, ms0);
#undef ST_GAL
    });
  }

  // The level's FP16 operator for the mixed V-cycle (each block scaled to its own largest
  // entry): made right after the level is built when the level will run in FP16, so that
  // its smoother's step is estimated on the same operator.
  void make_half(Level& lv) {
    lv.hblk.ensure((size_t)lv.nblocks * sizeof(HalfBlock));
    k_to_half_blocks<<<nblk(lv.nblocks), 256, 0, st>>>(lv.nblocks, lv.a64.as<double>(), lv.col.as<int>(),
                                                       lv.hblk.as<HalfBlock>());
    ST_HIP(hipGetLastError());
    lv.a32.release();
    lv.half = true;
  }

  // The smoother's step 4/(3 lambda) per structure, lambda the largest eigenvalue of D^-1 A
  // by 12 power iterations (statics/src/bodies.rs:180-190). A level that runs in FP16 is
  // iterated on its FP16 blocks in FP32 (a quarter of the bytes of the FP64 blocks; the
  // estimate only sets a damping factor).
  void coarse_step(Level& lv, const std::vector<int64_t>& seg) {
    const int64_t n = lv.n;
    DBuf v, w, sq, zr, neg;
    const size_t vb = lv.half ? sizeof(float) : sizeof(double);
    v.ensure(n * 6 * vb);
    w.ensure(n * 6 * vb);
    sq.ensure(n * sizeof(double));
    DBuf segd;
    segd.ensure((ns + 1) * sizeof(int64_t));
    ST_HIP(hipMemcpyAsync(segd.p, seg.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    const int iters = 12;
    if (lv.half) {
      // w = D^-1 A v as the coarse residual kernel's out2 = step D^-1 (r - A v) with r = 0 and
      // step = -1 for every structure; no structure is skipped (done = NULL)
      zr.ensure(n * 6 * sizeof(float));
      ST_HIP(hipMemsetAsync(zr.p, 0, n * 6 * sizeof(float), st));
      std::vector<double> m1(ns, -1.0);
      neg.ensure(ns * sizeof(double));
      ST_HIP(hipMemcpyAsync(neg.p, m1.data(), ns * sizeof(double), hipMemcpyHostToDevice, st));
      DBuf scratch;
      scratch.ensure(n * 6 * sizeof(float));
      float* a = v.as<float>();
      float* b = w.as<float>();
      k_c_init32<<<nblk(n), 256, 0, st>>>(n, a);
      for (int it = 0; it < iters; it++) {
        k_c_resid_h<16, 4><<<half_blocks_grid<16>(n), 256, 0, st>>>(
            0, n, lv.agg_struct.as<uint32_t>(), nullptr, lv.rowptr.as<int>(), lv.hblk.as<HalfBlock>(),
            lv.dinv32.as<float>(), neg.as<double>(), zr.as<float>(), a, scratch.as<float>(), nullptr, b);
        if (it >= iters - 2) {
          k_c_sq<float><<<nblk(n), 256, 0, st>>>(n, b, sq.as<double>());
          k_reduce_seg<<<ns, kReduceThreads, 0, st>>>(sq.as<double>(), segd.as<int64_t>(), c.scal.as<double>(), kScal,
                                              it == iters - 2 ? S_LAM_A : S_LAM_B);
        }
        std::swap(a, b);
      }
      ST_HIP(hipGetLastError());
      ST_HIP(hipStreamSynchronize(st));  // before the scratch vectors go
      scratch.release();
    } else {
      double* a = v.as<double>();
      double* b = w.as<double>();
      k_c_init<<<nblk(n), 256, 0, st>>>(n, a);
      for (int it = 0; it < iters; it++) {
        k_c_dinvA<double><<<nblk(8 * n), 256, 0, st>>>(n, lv.rowptr.as<int>(), lv.col.as<int>(), lv.a64.as<double>(),
                                                   lv.dinv64.as<double>(), a, b);
        if (it >= iters - 2) {
          k_c_sq<double><<<nblk(n), 256, 0, st>>>(n, b, sq.as<double>());
          k_reduce_seg<<<ns, kReduceThreads, 0, st>>>(sq.as<double>(), segd.as<int64_t>(), c.scal.as<double>(), kScal,
                                              it == iters - 2 ? S_LAM_A : S_LAM_B);
        }
        std::swap(a, b);
      }
    }
    std::vector<double> sc((size_t)ns * kScal);
    ST_HIP(hipMemcpyAsync(sc.data(), c.scal.p, sc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
    lv.step.assign(ns, 1.0);
    for (int s = 0; s < ns; s++) {
      const double x = sc[(size_t)s * kScal + S_LAM_A], y = sc[(size_t)s * kScal + S_LAM_B];
      const double lam = (x > 0 && y > 0) ? std::sqrt(y / x) : 1.0;
      lv.step[s] = 4.0 / 3.0 / lam;
    }
    lv.step_d.ensure(ns * sizeof(double));
    ST_HIP(hipMemcpyAsync(lv.step_d.p, lv.step.data(), ns * sizeof(double), hipMemcpyHostToDevice, st));
    ST_HIP(hipStreamSynchronize(st));
    v.release();
    w.release();
    sq.release();
    segd.release();
    zr.release();
    neg.release();
  }

  void alloc_level_vectors(Level& lv) {
    const size_t n = std::max<int64_t>(lv.n, 1);
    lv.r64.ensure(n * 6 * 8);
    lv.x64.ensure(n * 6 * 8);
    lv.t64.ensure(n * 6 * 8);
    lv.u64.ensure(n * 6 * 8);
    lv.v64.ensure(n * 6 * 8);
    lv.r32.ensure(n * 6 * 4);
    lv.x32.ensure(n * 6 * 4);
    lv.t32.ensure(n * 6 * 4);
    lv.u32.ensure(n * 6 * 4);
    lv.v32.ensure(n * 6 * 4);
  }

  void setup_hierarchy() {
    release_levels(c);
    const int64_t nn = g.nnodes;
    c.dof_per_struct.assign(ns, 0);
    c.last_ndof = 0;
    if (!nn) return;
    c.uf.ensure(nn * sizeof(int));
    c.uf2.ensure((nn + 1) * sizeof(int));
    c.flag.ensure((nn + 1) * sizeof(int));
    c.label1.ensure(nn * sizeof(int));
    k_iota<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>());
    int cap = 0;
    for (auto& si : c.host.structs) cap = std::max(cap, si.levels_cap);
    // number of free nodes
    k_root_flags<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>(), g.nflags.as<uint8_t>(), c.flag.as<int>());
    scan(c.flag.as<int>(), c.uf2.as<int>(), nn);
    {
      std::vector<int64_t> seg(ns + 1, 0);
      seg_tmp.ensure((ns + 1) * sizeof(int64_t));
      k_seg_from_ids<<<nblk(ns + 1), 256, 0, st>>>(ns, node0_d.as<int64_t>(), c.uf2.as<int>(), c.flag.as<int>(), nn,
                                                   seg_tmp.as<int64_t>());
      ST_HIP(hipMemcpyAsync(seg.data(), seg_tmp.p, (ns + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      ST_HIP(hipStreamSynchronize(st));
      ndof = seg[ns];
      c.dof_per_struct.assign(ns, 0);
      for (int s2 = 0; s2 < ns; s2++) c.dof_per_struct[s2] = seg[s2 + 1] - seg[s2];
      c.last_ndof = ndof;
    }
    if (!ndof) return;
    // the mixed V-cycle's three largest coarse operators (levels 1-3, nearly all of the
    // coarse blocks) in FP16, each block with its own scale: at 256^3 the same iterations and
    // 3.14 s a step against 3.20 s with level 3 in FP32; the 512-structure batch stays at 96
    // iterations at most. FP16 on level 4 as well left some of the batch's structures at 248.
    // STATICS_HIP_HALF_LEVELS sets how many levels, STATICS_HIP_HALF_BLOCKS keeps the
    // operators of fewer blocks than it in FP32, STATICS_HIP_HALF=0 keeps them all.
    int64_t kHalfBlocks = 0;
    if (const char* e = std::getenv("STATICS_HIP_HALF_BLOCKS")) kHalfBlocks = std::max<int64_t>(0, std::atoll(e));
    int kHalfLevels = 3;
    if (const char* e = std::getenv("STATICS_HIP_HALF_LEVELS")) kHalfLevels = std::max(0, std::atoi(e));
    auto maybe_half = [&](Level& lv) {
      const int li = (int)c.levels.size();
      if (mixed && c.half_coarse && li < c.f64_from && li < kHalfLevels && lv.nblocks >= kHalfBlocks) make_half(lv);
    };
    int64_t prev = ndof;
    DBuf label_tmp;
    label_tmp.ensure(nn * sizeof(int));
    for (int L = 1; L <= std::max(cap, 1); L++) {
      Level lv;
      std::vector<int64_t> seg;
      int* lab = c.levels.empty() ? c.label1.as<int>() : label_tmp.as<int>();
      const int64_t cnt = label_level(L, lab, lv.rep, seg);
      lv.cube = L;
      if (cnt == prev && L < std::max(cap, 1)) {  // nothing merged: try the next cube
        lv.rep.release();
        continue;
      }
      if (cnt == prev && !c.levels.empty()) {
        lv.rep.release();
        break;
      }
      lv.n = cnt;
      lv.seg = seg;
      lv.agg_struct.ensure(cnt * sizeof(uint32_t));
      k_agg_struct<<<nblk(cnt), 256, 0, st>>>(cnt, lv.rep.as<int>(), g.ncell.as<int32_t>(), g.cstruct.as<uint32_t>(),
                                              lv.agg_struct.as<uint32_t>());
      lv.centre.ensure(cnt * 3 * sizeof(double));
      lv.parent.ensure(cnt * sizeof(int));
      k_fill_int<<<nblk(cnt), 256, 0, st>>>(cnt, lv.parent.as<int>(), -1);
      if (c.levels.empty()) {
        // level 1: nodes grouped by aggregate
        DBuf keys, vals, skeys;
        keys.ensure(nn * sizeof(int));
        vals.ensure(nn * sizeof(int));
        skeys.ensure(nn * sizeof(int));
        c.nodes_by_agg.ensure(nn * sizeof(int));
        c.agg_ptr.ensure((cnt + 1) * sizeof(int));
        k_sort_keys<<<nblk(nn), 256, 0, st>>>(nn, c.label1.as<int>(), keys.as<int>(), vals.as<int>());
        sort_pairs(keys.as<int>(), skeys.as<int>(), vals.as<int>(), c.nodes_by_agg.as<int>(), nn);
        k_seg_ptr<<<nblk(ndof), 256, 0, st>>>(ndof, skeys.as<int>(), c.agg_ptr.as<int>(), (int)cnt);
        k_centre1<<<nblk(cnt), 256, 0, st>>>(cnt, c.agg_ptr.as<int>(), c.nodes_by_agg.as<int>(), g.nkey.as<uint32_t>(),
                                             g.nkind.as<uint8_t>(), lv.agg_struct.as<uint32_t>(), g.bgeo.as<BeamGeo>(),
                                             lv.centre.as<double>());
        keys.release();
        vals.release();
        skeys.release();
        // the brick ranges of the level-1 aggregates, for the fused restriction
        c.brick_agg1.ensure((g.nbricks + 1) * sizeof(int));
        k_brick_agg1<<<nblk(g.nbricks + 1), 256, 0, st>>>(g.nbricks, g.nbase.as<uint32_t>(), cnt, c.agg_ptr.as<int>(),
                                                         c.nodes_by_agg.as<int>(), c.brick_agg1.as<int>());
        ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
        k_agg_straddle<<<nblk(cnt), 256, 0, st>>>(cnt, c.agg_ptr.as<int>(), c.nodes_by_agg.as<int>(), g.nbricks,
                                                  g.nbase.as<uint32_t>(), sbad.as<int>());
        c.restrict_fused = read_int(sbad.as<int>()) == 0 && std::getenv("STATICS_HIP_NOFUSE") == nullptr;
        if (c.smoothed)
          build_sa_level1(lv);
        else
          build_rows_a1(lv);
      } else {
        Level& below = c.levels.back();
        // parent of each aggregate below
        k_parent<<<nblk(below.n), 256, 0, st>>>(below.n, below.rep.as<int>(), lab, below.parent.as<int>());
        DBuf vals, skeys;
        vals.ensure(below.n * sizeof(int));
        skeys.ensure(below.n * sizeof(int));
        lv.kids.ensure(below.n * sizeof(int));
        lv.kid_ptr.ensure((cnt + 1) * sizeof(int));
        k_iota<<<nblk(below.n), 256, 0, st>>>(below.n, vals.as<int>());
        sort_pairs(below.parent.as<int>(), skeys.as<int>(), vals.as<int>(), lv.kids.as<int>(), below.n);
        k_seg_ptr<<<nblk(below.n), 256, 0, st>>>(below.n, skeys.as<int>(), lv.kid_ptr.as<int>(), (int)cnt);
        k_centre_c<<<nblk(cnt), 256, 0, st>>>(cnt, lv.kid_ptr.as<int>(), lv.kids.as<int>(), below.centre.as<double>(),
                                              lv.centre.as<double>());
        vals.release();
        skeys.release();
        if (c.smoothed) {
          build_sa_coarse(below, lv, (int)c.levels.size() < c.sa_levels);
          maybe_half(lv);
          coarse_step(lv, seg);
          alloc_level_vectors(lv);
          c.levels.push_back(std::move(lv));
          prev = cnt;
          if (c.levels.back().nblocks == c.levels.back().n) break;  // block diagonal: exact top
          continue;
        }
        DBuf cntb;
        cntb.ensure((cnt + 1) * sizeof(int));
        lv.rowptr.ensure((cnt + 1) * sizeof(int));
        ST_HIP(hipMemsetAsync(sbad.p, 0, sizeof(int), st));
        ST_HIP(hipMemsetAsync(cntb.p, 0, (cnt + 1) * sizeof(int), st));
        auto launch = [&](int pass) {
          k_galerkin<<<nblk(cnt, 64), 64, 0, st>>>(pass, cnt, lv.kid_ptr.as<int>(), lv.kids.as<int>(),
                                                   lv.centre.as<double>(), below.centre.as<double>(),
                                                   below.parent.as<int>(), below.rowptr.as<int>(), below.col.as<int>(),
                                                   below.a64.as<double>(), cntb.as<int>(), lv.rowptr.as<int>(),
                                                   lv.col.as<int>(), lv.a64.as<double>(), sbad.as<int>());
          ST_HIP(hipGetLastError());
        };
        launch(0);
        finish_rows(lv, cntb, launch);
        cntb.release();
      }
      maybe_half(lv);
      coarse_step(lv, seg);
      alloc_level_vectors(lv);
      c.levels.push_back(std::move(lv));
      prev = cnt;
      if (c.levels.back().nblocks == c.levels.back().n) break;  // block diagonal: exact top
    }
    label_tmp.release();
    if (c.dense_top) make_dense_top();
    // (a level made the dense top keeps its FP64 operator: it is inverted, not iterated)
    for (Level& lv : c.levels)
      if (lv.dense && lv.half) {
        lv.hblk.release();
        lv.half = false;
      }
    if (mixed)  // the FP32 levels' FP64 copies served the Galerkin products and the step estimates only
      for (size_t li = 0; li < c.levels.size() && (int)li < c.f64_from; li++)
        for (DBuf* b : {&c.levels[li].a64, &c.levels[li].dinv64, &c.levels[li].pval64}) b->release();
    // the FP32 transfers below levels of more than kWideRestrict aggregates through records
    // (STATICS_HIP_PREC=0: through the block arrays)
    const char* prec_env = std::getenv("STATICS_HIP_PREC");
    if (mixed && !(prec_env && std::atoi(prec_env) == 0))
      for (size_t li = 0; li + 1 < c.levels.size() && (int)li < c.f64_from; li++) {
        Level& lv = c.levels[li];
        if (!lv.pblocks || c.levels[li + 1].n <= kWideRestrict) continue;
        lv.prh = c.p_half && lv.half;
        const size_t rb = lv.prh ? sizeof(HalfBlock) : sizeof(PBlock32);
        lv.prr.ensure((size_t)lv.pblocks * rb);
        lv.prc.ensure((size_t)lv.pblocks * rb);
        if (lv.prh) {
          k_pblocks_rows<<<nblk(lv.pblocks), 256, 0, st>>>(lv.pblocks, lv.pcol.as<int>(), lv.pval32.as<float>(),
                                                           lv.prr.as<HalfBlock>());
          k_pblocks_cols<<<nblk(lv.pblocks), 256, 0, st>>>(lv.pblocks, lv.pt_blk.as<int>(), lv.pblk_row.as<int>(),
                                                           lv.pval32.as<float>(), lv.prc.as<HalfBlock>());
        } else {
          k_pblocks_rows<<<nblk(lv.pblocks), 256, 0, st>>>(lv.pblocks, lv.pcol.as<int>(), lv.pval32.as<float>(),
                                                           lv.prr.as<PBlock32>());
          k_pblocks_cols<<<nblk(lv.pblocks), 256, 0, st>>>(lv.pblocks, lv.pt_blk.as<int>(), lv.pblk_row.as<int>(),
                                                           lv.pval32.as<float>(), lv.prc.as<PBlock32>());
        }
        ST_HIP(hipGetLastError());
        lv.pval32.release();
      }
    tune_coarse_layout();
  }

  // Per coarse level, the fastest layout of its residual kernel in the precision the solve
  // runs it in (FP64 and FP32: narrow or wide rows, four or eight blocks in flight, plain or
  // non-temporal block loads, see k_c_resid; FP16: lanes per row and blocks in flight, see
  // k_c_resid_h), timed on the level itself after a warm launch. STATICS_HIP_WIDE=k (0-7)
  // forces layout k.
  void tune_coarse_layout() {
    const char* force = std::getenv("STATICS_HIP_WIDE");
    hipEvent_t e0, e1;
    ST_HIP(hipEventCreate(&e0));
    ST_HIP(hipEventCreate(&e1));
    for (size_t li = 0; li < c.levels.size(); li++) {
      Level& lv = c.levels[li];
      if (lv.dense || exact_level(li)) continue;
      const bool f32 = mixed && (int)li < c.f64_from;
      if (force) {
        (f32 ? lv.layout32 : lv.layout64) = std::max(0, std::min(7, std::atoi(force)));
        continue;
      }
      // a rebuilt hierarchy of the same upload has levels of the same shapes: the choice
      // made for a level of this index, precision and shape stands
      const std::array<int64_t, 4> shape{(int64_t)li, f32 ? (lv.half ? 2 : 1) : 0, lv.n, lv.nblocks};
      bool known = false;
      for (auto& kv : c.coarse_layouts)
        if (kv.first == shape) {
          (f32 ? lv.layout32 : lv.layout64) = kv.second;
          known = true;
        }
      if (known) continue;
      float ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const int nlay = 8;  // the FP16 kernel's layouts, or the lane layouts
      for (int w = 0; w < nlay; w++) {
        (f32 ? lv.layout32 : lv.layout64) = w;
        auto run = [&] {
          if (f32) {
            ST_HIP(hipMemsetAsync(lv.x32.p, 0, lv.n * 6 * sizeof(float), st));
            c_resid<float>(lv, 0, lv.r32.as<float>(), lv.x32.as<float>(), lv.t32.as<float>());
          } else {
            ST_HIP(hipMemsetAsync(lv.x64.p, 0, lv.n * 6 * sizeof(double), st));
            c_resid<double>(lv, 0, lv.r64.as<double>(), lv.x64.as<double>(), lv.t64.as<double>());
          }
        };
        // one warm launch, timed: a level whose launch takes under 50 us is timed over 30
        // launches (its few microseconds vary from launch to launch), others over 3
        ST_HIP(hipEventRecord(e0, st));
        run();
        ST_HIP(hipEventRecord(e1, st));
        ST_HIP(hipEventSynchronize(e1));
        float warm = 0;
        ST_HIP(hipEventElapsedTime(&warm, e0, e1));
        const int reps = warm < 0.05f ? 30 : 3;
        ST_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < reps; k++) run();
        ST_HIP(hipEventRecord(e1, st));
        ST_HIP(hipEventSynchronize(e1));
        ST_HIP(hipEventElapsedTime(&ms[w], e0, e1));
        ms[w] /= (float)reps;
      }
      int best = 0;
      for (int w = 1; w < nlay; w++)
        if (ms[w] < ms[best]) best = w;
      (f32 ? lv.layout32 : lv.layout64) = best;
      c.coarse_layouts.emplace_back(shape, best);
      if (std::getenv("STATICS_HIP_VERBOSE"))
        std::fprintf(stderr,
                     "statics_hip: level %zu residual layout %d (%s): %.1f %.1f %.1f %.1f %.1f %.1f %.1f %.1f us\n",
                     li + 1, best, f32 ? (lv.half ? "fp16" : "fp32") : "fp64", 1e3 * ms[0], 1e3 * ms[1],
                     1e3 * ms[2], 1e3 * ms[3], 1e3 * ms[4], 1e3 * ms[5], 1e3 * ms[6], 1e3 * ms[7]);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }

  // The first level from index 2 up where every structure has at most kDenseMax
  // aggregates becomes the top: its operator is inverted densely per structure, and
  // the levels above it are dropped (visiting them cost hundreds of tiny launches an
  // iteration in the gamma-cycles).
  void make_dense_top() {
    // (a top of up to 1,024 aggregates, at 256^3 level 5 with 820, saved 2% of an
    // iteration and took 0.5 s more to invert: not kept)
    constexpr int64_t kDenseMax = 256;
    int ld = -1;
    for (size_t li = 2; li + 1 < c.levels.size(); li++) {
      const auto& sg = c.levels[li].seg;
      int64_t mx = 0;
      for (int s2 = 0; s2 < ns; s2++) mx = std::max(mx, sg[s2 + 1] - sg[s2]);
      if (mx <= kDenseMax) {
        ld = (int)li;
        break;
      }
    }
    if (ld < 0) return;
    for (size_t li = ld + 1; li < c.levels.size(); li++) {
      Level& l2 = c.levels[li];
      for (DBuf* b : {&l2.rowptr, &l2.col, &l2.a64, &l2.a32, &l2.hblk, &l2.dinv64, &l2.dinv32,
                      &l2.centre, &l2.parent, &l2.rep, &l2.key, &l2.kid_ptr, &l2.kids, &l2.agg_struct, &l2.step_d,
                      &l2.r64, &l2.x64, &l2.t64, &l2.u64, &l2.v64, &l2.r32, &l2.x32, &l2.t32, &l2.u32, &l2.v32,
                      &l2.prow, &l2.pcol, &l2.pval64, &l2.pval32, &l2.prr, &l2.prc, &l2.pt_ptr, &l2.pt_blk, &l2.pblk_row})
        b->release();
    }
    c.levels.resize(ld + 1);
    Level& lv = c.levels[ld];
    std::vector<int64_t> off(ns + 1, 0);
    int64_t mmax = 0;
    for (int s2 = 0; s2 < ns; s2++) {
      const int64_t m = 6 * (lv.seg[s2 + 1] - lv.seg[s2]);
      off[s2 + 1] = off[s2] + m * m;
      mmax = std::max(mmax, m);
    }
    lv.dense = true;
    lv.dense_mmax = (int32_t)mmax;
    lv.dense_a.ensure(std::max<int64_t>(off[ns], 1) * sizeof(double));
    lv.dense_off.ensure((ns + 1) * sizeof(int64_t));
    lv.dense_seg.ensure((ns + 1) * sizeof(int64_t));
    ST_HIP(hipMemcpyAsync(lv.dense_off.p, off.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    ST_HIP(hipMemcpyAsync(lv.dense_seg.p, lv.seg.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    ST_HIP(hipMemsetAsync(lv.dense_a.p, 0, std::max<int64_t>(off[ns], 1) * sizeof(double), st));
    if (!mmax) return;
    const int64_t* sgd = lv.dense_seg.as<int64_t>();
    const int64_t* ofd = lv.dense_off.as<int64_t>();
    double* D = lv.dense_a.as<double>();
    k_dense_fill<<<nblk(lv.n), 256, 0, st>>>(lv.n, lv.agg_struct.as<uint32_t>(), sgd, ofd, lv.rowptr.as<int>(),
                                             lv.col.as<int>(), lv.a64.as<double>(), D);
    // structures by size, largest first, and running sums of m and m^2 in that order
    std::vector<int> ord(ns);
    for (int s2 = 0; s2 < ns; s2++) ord[s2] = s2;
    auto msz = [&](int s2) { return 6 * (lv.seg[s2 + 1] - lv.seg[s2]); };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return msz(a) > msz(b); });
    std::vector<int64_t> pre1(ns + 1, 0), pre2(ns + 1, 0);
    for (int i = 0; i < ns; i++) {
      pre1[i + 1] = pre1[i] + msz(ord[i]);
      pre2[i + 1] = pre2[i] + msz(ord[i]) * msz(ord[i]);
    }
    DBuf colk, rowk, ordd, pre1d, pre2d;
    colk.ensure((size_t)ns * mmax * sizeof(double));
    rowk.ensure((size_t)ns * mmax * sizeof(double));
    ordd.ensure(ns * sizeof(int));
    pre1d.ensure((ns + 1) * sizeof(int64_t));
    pre2d.ensure((ns + 1) * sizeof(int64_t));
    ST_HIP(hipMemcpyAsync(ordd.p, ord.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
    ST_HIP(hipMemcpyAsync(pre1d.p, pre1.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    ST_HIP(hipMemcpyAsync(pre2d.p, pre2.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    int nact = ns;
    for (int k = 0; k < mmax; k++) {
      while (nact > 0 && msz(ord[nact - 1]) <= k) nact--;  // structures with more than k rows
      k_gj_save<<<nblk(pre1[nact]), 256, 0, st>>>(k, nact, ordd.as<int>(), pre1d.as<int64_t>(), sgd, ofd, D,
                                                  colk.as<double>(), rowk.as<double>(), (int)mmax);
      k_gj_update<<<nblk(pre2[nact]), 256, 0, st>>>(k, nact, ordd.as<int>(), pre2d.as<int64_t>(), sgd, ofd, D,
                                                    colk.as<double>(), rowk.as<double>(), (int)mmax);
    }
    k_dense_sym<<<nblk(pre2[ns]), 256, 0, st>>>(ns, ordd.as<int>(), pre2d.as<int64_t>(), sgd, ofd, D);
    ST_HIP(hipGetLastError());
    ST_HIP(hipStreamSynchronize(st));
    for (DBuf* b2 : {&colk, &rowk, &ordd, &pre1d, &pre2d}) b2->release();
  }

  // --------------------------------------------------------------- V-cycle
  template <class T>
  T* lvbuf(Level& lv, int which) {
    DBuf* b64[5] = {&lv.r64, &lv.x64, &lv.t64, &lv.u64, &lv.v64};
    DBuf* b32[5] = {&lv.r32, &lv.x32, &lv.t32, &lv.u32, &lv.v32};
    return std::is_same<T, double>::value ? (T*)b64[which]->p : (T*)b32[which]->p;
  }
  template <class T>
  const T* lvA(Level& lv) {
    return std::is_same<T, double>::value ? (const T*)lv.a64.p : (const T*)lv.a32.p;
  }
  template <class T>
  const T* lvD(Level& lv) {
    return std::is_same<T, double>::value ? (const T*)lv.dinv64.p : (const T*)lv.dinv32.p;
  }

  // Level li is solved exactly by one launch: its dense top, or a block-diagonal top.
  bool exact_level(size_t li) const {
    const Level& lv = c.levels[li];
    return lv.dense || (li + 1 == c.levels.size() && lv.nblocks == lv.n);
  }

  // The coarse residual kernel of level lv in its tuned layout (narrow or wide rows).
  // mode 0: out = r - A x (and out2 = w D^-1 (r - A x) when out2 is set);
  // mode 1: out = x + w D^-1 (r - A x) (+ add).
  template <class T, int W, int U, int NT>
  void c_resid_v(Level& lv, int mode, const T* r, const T* x, T* out, const T* add, T* out2, const T* dinv,
                 const double* step) {
    k_c_resid<T, W, U, NT><<<coarse_blocks<T>(lv.n, W != 0), 256, 0, st>>>(
          mode, lv.n, lv.agg_struct.as<uint32_t>(), c.done.as<uint8_t>(), lv.rowptr.as<int>(), lv.col.as<int>(),
          lvA<T>(lv), dinv, step, r, x, out, add, out2);
  }
  template <int G, int U>
  void c_resid_h(Level& lv, int mode, const float* r, const float* x, float* out, const float* add, float* out2,
                 const float* dinv, const double* step) {
    k_c_resid_h<G, U><<<half_blocks_grid<G>(lv.n), 256, 0, st>>>(
        mode, lv.n, lv.agg_struct.as<uint32_t>(), c.done.as<uint8_t>(), lv.rowptr.as<int>(), lv.hblk.as<HalfBlock>(),
        dinv, step, r, x, out, add, out2);
  }
  template <class T>
  void c_resid(Level& lv, int mode, const T* r, const T* x, T* out, const T* add = nullptr, T* out2 = nullptr) {
    const int layout = std::is_same<T, double>::value ? lv.layout64 : lv.layout32;
    const T* dinv = (mode == 1 || out2) ? lvD<T>(lv) : nullptr;
    const double* step = (mode == 1 || out2) ? lv.step_d.as<double>() : nullptr;
    if constexpr (std::is_same<T, float>::value) {
      if (lv.half) {
        switch (layout) {  // (lanes per row, blocks in flight a lane)
          case 0: c_resid_h<8, 2>(lv, mode, r, x, out, add, out2, dinv, step); break;
          case 1: c_resid_h<16, 2>(lv, mode, r, x, out, add, out2, dinv, step); break;
          case 2: c_resid_h<32, 2>(lv, mode, r, x, out, add, out2, dinv, step); break;
          case 3: c_resid_h<8, 4>(lv, mode, r, x, out, add, out2, dinv, step); break;
          case 4: c_resid_h<16, 4>(lv, mode, r, x, out, add, out2, dinv, step); break;
          case 5: c_resid_h<32, 4>(lv, mode, r, x, out, add, out2, dinv, step); break;
          case 6: c_resid_h<64, 4>(lv, mode, r, x, out, add, out2, dinv, step); break;
          default: c_resid_h<16, 8>(lv, mode, r, x, out, add, out2, dinv, step); break;
        }
        return;
      }
    }
    switch (layout) {  // bit 0: wide rows; bit 1: eight blocks in flight; bit 2: non-temporal blocks
      case 0: c_resid_v<T, 0, 4, 0>(lv, mode, r, x, out, add, out2, dinv, step); break;
      case 1: c_resid_v<T, 1, 4, 0>(lv, mode, r, x, out, add, out2, dinv, step); break;
      case 2: c_resid_v<T, 0, 8, 0>(lv, mode, r, x, out, add, out2, dinv, step); break;
      case 3: c_resid_v<T, 1, 8, 0>(lv, mode, r, x, out, add, out2, dinv, step); break;
      case 4: c_resid_v<T, 0, 4, 1>(lv, mode, r, x, out, add, out2, dinv, step); break;
      case 5: c_resid_v<T, 1, 4, 1>(lv, mode, r, x, out, add, out2, dinv, step); break;
      case 6: c_resid_v<T, 0, 8, 1>(lv, mode, r, x, out, add, out2, dinv, step); break;
      default: c_resid_v<T, 1, 8, 1>(lv, mode, r, x, out, add, out2, dinv, step); break;
    }
  }

  // The coarse solve of level li for a V-cycle in precision T, on the right-hand side in
  // the level's buffer 0; returns the buffer holding the result. x0_ready: whoever wrote
  // the right-hand side also wrote the first cycle's damped Jacobi start w D^-1 r into
  // buffer 1. Levels from c.f64_from up run in FP64 whatever T is (their operators are
  // the most ill-conditioned and the smallest), the right-hand side and the result
  // converted at the switch.
  template <class T>
  T* descend(size_t li, bool x0_ready) {
    if (std::is_same<T, float>::value && (int)li >= c.f64_from) {
      Level& lv = c.levels[li];
      const int64_t m = lv.n * 6;
      k_convert<float, double><<<nblk(m), 256, 0, st>>>(m, lv.r32.as<float>(), lv.r64.as<double>());
      const double* y = coarse_cycle<double>(li, false);
      k_convert<double, float><<<nblk(m), 256, 0, st>>>(m, y, lv.x32.as<float>());
      return (T*)lv.x32.p;
    }
    return coarse_cycle<T>(li, x0_ready);
  }

  // Solves A_li x = r approximately: one cycle, or (levels with a gamma count above one)
  // gamma of them, each later one on the remaining residual (x += cycle(r - A x)), which
  // keeps the preconditioner symmetric. Buffers: 0 r, 1 each cycle's start, 3 each later
  // cycle's right-hand side, 2 and 4 the results in turn.
  template <class T>
  T* coarse_cycle(size_t li, bool x0_ready) {
    Level& lv = c.levels[li];
    const uint8_t* done = c.done.as<uint8_t>();
    T* r = lvbuf<T>(lv, 0);
    T* x = lvbuf<T>(lv, 1);
    const int64_t n = lv.n;
    const uint32_t* ast = lv.agg_struct.as<uint32_t>();
    if (lv.dense) {  // the dense top: exact
      const dim3 gd(nblk(lv.dense_mmax, 4), (unsigned)ns);
      k_dense_apply<T><<<gd, 256, 0, st>>>(lv.dense_seg.as<int64_t>(), lv.dense_off.as<int64_t>(),
                                          lv.dense_a.as<double>(), done, r, x);
      return x;
    }
    if (exact_level(li)) {  // block diagonal: exact
      k_c_jacobi<T><<<nblk(n), 256, 0, st>>>(n, ast, done, lvD<T>(lv), nullptr, r, x);
      return x;
    }
    if (!x0_ready) k_c_jacobi<T><<<nblk(n), 256, 0, st>>>(n, ast, done, lvD<T>(lv), lv.step_d.as<double>(), r, x);
    T* res = one_cycle<T>(li, r, x, lvbuf<T>(lv, 2), nullptr);
    const int reps = !c.gammas.empty() ? (li < c.gammas.size() ? c.gammas[li] : 1)
                                       : (((int)li > c.gamma_from && (int)li <= c.gamma_to) ? c.gamma : 1);
    T* rk = lvbuf<T>(lv, 3);
    for (int k = 1; k < reps; k++) {
      c_resid<T>(lv, 0, r, res, rk, nullptr, x);  // rk = r - A res; x = w D^-1 rk
      T* out = res == lvbuf<T>(lv, 2) ? lvbuf<T>(lv, 4) : lvbuf<T>(lv, 2);
      res = one_cycle<T>(li, rk, x, out, res);  // res + cycle(rk)
    }
    return res;
  }

  // Restrictions onto a level of at most this many aggregates take a wave per aggregate
  // (k_restrict_pw): such levels are deep, with long columns of P, and few enough that the
  // eight-lane groups of k_restrict_p leave most of the GPU idle.
  static constexpr int64_t kWideRestrict = 8192;

  // One cycle at level li on rhs from the damped Jacobi start x (updated in place): the
  // residual, restricted; the coarser solve, prolonged onto x; the post-smoothing
  // x + w D^-1 (rhs - A x) (+ add) into out, which also serves as the residual's buffer
  // (out differs from rhs, x and add). Returns out.
  template <class T>
  T* one_cycle(size_t li, const T* rhs, T* x, T* out, const T* add) {
    Level& lv = c.levels[li];
    const uint8_t* done = c.done.as<uint8_t>();
    const int64_t n = lv.n;
    const uint32_t* ast = lv.agg_struct.as<uint32_t>();
    if (li + 1 == c.levels.size()) {  // not block diagonal (should not happen): smoothing sweeps
      for (int k = 0; k < 8; k++) {
        c_resid<T>(lv, 1, rhs, x, out);
        k_axpy<T><<<nblk(6 * n), 256, 0, st>>>(6 * n, T(0), out, x);
      }
      c_resid<T>(lv, 1, rhs, x, out, add);
      return out;
    }
    Level& up = c.levels[li + 1];
    const int nu = li == 0 ? c.nu1 : c.nu_c;
    for (int i = 1; i < nu; i++) {  // further smoothing steps x <- x + w D^-1 (rhs - A x)
      c_resid<T>(lv, 1, rhs, x, out);
      k_axpy<T><<<nblk(6 * n), 256, 0, st>>>(6 * n, T(0), out, x);
    }
    const T* pv = std::is_same<T, double>::value ? (const T*)lv.pval64.p : (const T*)lv.pval32.p;
    T* u = out;
    c_resid<T>(lv, 0, rhs, x, u);
    if constexpr (std::is_same<T, float>::value) {
      // the level above runs in FP64 and is small: the restriction writes its FP64
      // right-hand side and damped Jacobi start, and the prolongation reads its FP64
      // result (the conversions and the start's launch folded in, the same values)
      if (c.fold_switch && (int)(li + 1) >= c.f64_from && lv.pblocks && lv.prc.p == nullptr &&
          up.n <= kWideRestrict) {
        const bool fx = !exact_level(li + 1);
        k_restrict_pw<float, double><<<nblk(up.n, 4), 256, 0, st>>>(
            up.n, up.agg_struct.as<uint32_t>(), done, lv.pt_ptr.as<int>(), lv.pt_blk.as<int>(), lv.pblk_row.as<int>(),
            pv, u, lvbuf<double>(up, 0), fx ? lvD<double>(up) : nullptr, fx ? up.step_d.as<double>() : nullptr,
            fx ? lvbuf<double>(up, 1) : nullptr);
        const double* y64 = coarse_cycle<double>(li + 1, fx);
        k_prolong_p<float, double><<<nblk(8 * n), 256, 0, st>>>(n, ast, done, lv.prow.as<int>(), lv.pcol.as<int>(),
                                                                pv, y64, x);
        c_resid<T>(lv, 1, rhs, x, out, add);
        return out;
      }
    }
    // the coarser level's first cycle starts from w D^-1 of its right-hand side: written
    // with it, unless that level is solved exactly or switches precision
    const bool fuse = !exact_level(li + 1) && !(std::is_same<T, float>::value && (int)(li + 1) >= c.f64_from);
    T* up_r = lvbuf<T>(up, 0);
    T* up_x0 = fuse ? lvbuf<T>(up, 1) : nullptr;
    const T* up_d = fuse ? lvD<T>(up) : nullptr;
    const double* up_step = fuse ? up.step_d.as<double>() : nullptr;
    bool recs = false;
    if constexpr (std::is_same<T, float>::value) recs = lv.prc.p != nullptr;
    if (recs) {
      if constexpr (std::is_same<T, float>::value) {
        const unsigned gr = nblk(up.n * 16);
        if (lv.prh)
          k_restrict_rec<16><<<gr, 256, 0, st>>>(up.n, up.agg_struct.as<uint32_t>(), done, lv.pt_ptr.as<int>(),
                                                 lv.prc.as<HalfBlock>(), u, up_r, up_d, up_step, up_x0);
        else
          k_restrict_rec<16><<<gr, 256, 0, st>>>(up.n, up.agg_struct.as<uint32_t>(), done, lv.pt_ptr.as<int>(),
                                                 lv.prc.as<PBlock32>(), u, up_r, up_d, up_step, up_x0);
      }
    } else if (lv.pblocks && up.n <= kWideRestrict)
      k_restrict_pw<T><<<nblk(up.n, 4), 256, 0, st>>>(
          up.n, up.agg_struct.as<uint32_t>(), done, lv.pt_ptr.as<int>(), lv.pt_blk.as<int>(), lv.pblk_row.as<int>(), pv,
          u, up_r, up_d, up_step, up_x0);
    else if (lv.pblocks)
      k_restrict_p<T><<<nblk(8 * up.n), 256, 0, st>>>(up.n, up.agg_struct.as<uint32_t>(), done, lv.pt_ptr.as<int>(),
                                                     lv.pt_blk.as<int>(), lv.pblk_row.as<int>(), pv, u, up_r, up_d,
                                                     up_step, up_x0);
    else
      k_restrict_c<T><<<nblk(up.n), 256, 0, st>>>(up.n, up.agg_struct.as<uint32_t>(), done, up.kid_ptr.as<int>(),
                                                  up.kids.as<int>(), up.centre.as<double>(), lv.centre.as<double>(), u,
                                                  up_r, up_d, up_step, up_x0);
    const T* y = descend<T>(li + 1, fuse);
    if (recs) {
      if constexpr (std::is_same<T, float>::value) {
        const unsigned gp = nblk(n * 8);
        if (lv.prh)
          k_prolong_rec<8><<<gp, 256, 0, st>>>(n, ast, done, lv.prow.as<int>(), lv.prr.as<HalfBlock>(), y, x);
        else
          k_prolong_rec<8><<<gp, 256, 0, st>>>(n, ast, done, lv.prow.as<int>(), lv.prr.as<PBlock32>(), y, x);
      }
    } else if (lv.pblocks)
      k_prolong_p<T><<<nblk(8 * n), 256, 0, st>>>(n, ast, done, lv.prow.as<int>(), lv.pcol.as<int>(), pv, y, x);
    else
      k_prolong_c<T><<<nblk(n), 256, 0, st>>>(n, ast, done, lv.parent.as<int>(), lv.centre.as<double>(),
                                              up.centre.as<double>(), y, x);
    c_resid<T>(lv, 1, rhs, x, out, add);
    return out;
  }

  // u = b - K x into u, and its rigid restriction P_rigid^T u into level 1's right-hand side:
  // fused into the operator's epilogue when every level-1 aggregate lies in one brick
  template <class T>
  void residual_restrict(const T* x, const T* b, T* u, Level& l1) {
    const uint8_t* done = c.done.as<uint8_t>();
    if (c.restrict_fused) {
      EpiResidRestrict<T> er;
      er.done = done;
      er.b = b;
      er.r = u;
      er.brick_agg = c.brick_agg1.as<int>();
      er.agg_ptr = c.agg_ptr.as<int>();
      er.nodes = c.nodes_by_agg.as<int>();
      er.centre = l1.centre.as<double>();
      er.nkey = g.nkey.as<uint32_t>();
      er.nkind = g.nkind.as<uint8_t>();
      er.bgeo = g.bgeo.as<BeamGeo>();
      er.c = lvbuf<T>(l1, 0);
      launch_apply<T>(c, x, er);
      return;
    }
    EpiResid<T, T> er;
    er.done = done;
    er.partial = c.partial.as<double>();
    er.b = b;
    er.r = u;
    launch_apply<T>(c, x, er);
    k_restrict1<T><<<nblk(l1.n), 256, 0, st>>>(l1.n, l1.agg_struct.as<uint32_t>(), done, c.agg_ptr.as<int>(),
                                               c.nodes_by_agg.as<int>(), l1.centre.as<double>(), g.nkey.as<uint32_t>(),
                                               g.nkind.as<uint8_t>(), g.bgeo.as<BeamGeo>(), u, lvbuf<T>(l1, 0));
  }

  // z = M^-1 r at the fine level; partial r.z per brick into c.partial
  // rdot: if set, the partial is rdot . z in FP64 instead of r . z
  template <class T>
  void vcycle(const T* r, T* z, T* x1, T* r1, T* s1, const double* rdot = nullptr) {
    const uint8_t* done = c.done.as<uint8_t>();
    const unsigned nb = (unsigned)g.nbricks;
    const T* dinv = std::is_same<T, double>::value ? (const T*)c.dinv64.p : (const T*)c.dinv32.p;
    k_jacobi_fine<T><<<nb, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), g.bstruct.as<uint32_t>(), done, dinv,
                                            c.step0_d.as<double>(), r, x1);
    // further damped Jacobi steps x1 += w D^-1 (r - K x1), ping-ponging through s1
    auto smooth_steps = [&](int k) {
      for (int i = 0; i < k; i++) {
        EpiSmooth<T> es;
        es.done = done;
        es.partial = c.partial.as<double>();
        es.b = r;
        es.dinv = dinv;
        es.step = c.step0_d.as<double>();
        es.z = s1;
        launch_apply<T>(c, x1, es);
        k_copy_fine<T, T><<<nb, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), g.bstruct.as<uint32_t>(), done, s1, x1);
      }
    };
    smooth_steps(c.nu - 1);
    if (!c.levels.empty() && c.smoothed) {
      // The smoothed restriction's argument is u = r1 - K w D^-1 r1 with r1 = r - K x1, that
      // is u = r - K s1 for s1 = x1 + w D^-1 (r - K x1): one operator pass writes s1 (and
      // no r1), the next forms u from r. Then c = P_rigid^T u; coarse; s1 = P_rigid y;
      // x1 += s1 - w D^-1 K s1; z = x1 + w D^-1 (r - K x1)
      EpiResidJac<T> ej;
      ej.done = done;
      ej.b = r;
      ej.dinv = dinv;
      ej.step = c.step0_d.as<double>();
      ej.sout = s1;
      launch_apply<T>(c, x1, ej);
      Level& l1 = c.levels[0];
      residual_restrict<T>(s1, r, r1, l1);
      const T* y1 = descend<T>(0, false);
      k_prolong1_set<T><<<nb, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), g.bstruct.as<uint32_t>(), done,
                                               c.label1.as<int>(), l1.centre.as<double>(), g.nkey.as<uint32_t>(),
                                               g.nkind.as<uint8_t>(), g.bgeo.as<BeamGeo>(), y1, s1);
      EpiProlong<T> ep;
      ep.done = done;
      ep.dinv = dinv;
      ep.step = c.step0_d.as<double>();
      ep.x = x1;
      launch_apply<T>(c, s1, ep);
    } else if (!c.levels.empty()) {
      Level& l1 = c.levels[0];
      residual_restrict<T>(x1, r, r1, l1);
      const T* y1 = descend<T>(0, false);
      k_prolong1<T><<<nb, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), g.bstruct.as<uint32_t>(), done, c.label1.as<int>(),
                                           l1.centre.as<double>(), g.nkey.as<uint32_t>(), g.nkind.as<uint8_t>(),
                                           g.bgeo.as<BeamGeo>(), y1, x1);
    }
    smooth_steps(c.nu - 1);
    EpiSmooth<T> es;
    es.done = done;
    es.partial = c.partial.as<double>();
    es.b = r;
    es.dinv = dinv;
    es.step = c.step0_d.as<double>();
    es.bdot = rdot;
    es.z = z;
    launch_apply<T>(c, x1, es);
    ST_HIP(hipGetLastError());
  }

  // ------------------------------------------------------------------ PCG
  void sync_flags(std::vector<uint8_t>& done, std::vector<uint8_t>& want) {
    done.resize(ns);
    want.resize(ns);
    ST_HIP(hipMemcpyAsync(done.data(), c.done.p, ns, hipMemcpyDeviceToHost, st));
    ST_HIP(hipMemcpyAsync(want.data(), swant.p, ns, hipMemcpyDeviceToHost, st));
    ST_HIP(hipStreamSynchronize(st));
  }

  // f64: loads (device, zero on non-free rows). x64: initial guess (zeros or warm) and result.
  void pcg(int precision, double tol, int max_iter, int refresh) {
    const unsigned nb = (unsigned)g.nbricks;
    const int64_t nn = g.nnodes;
    const size_t v64 = (size_t)std::max<int64_t>(nn, 1) * 6 * 8, v32 = (size_t)std::max<int64_t>(nn, 1) * 6 * 4;
    const uint32_t* nbase = g.nbase.as<uint32_t>();
    const uint32_t* bst = g.bstruct.as<uint32_t>();
    uint8_t* done = c.done.as<uint8_t>();
    double* scal = c.scal.as<double>();
    double* partial = c.partial.as<double>();
    const bool mixed = precision == ST_MIXED;
    c.r64.ensure(v64);
    // ||f||^2 and the initial residual r = f - K x
    k_dot_fine<double><<<nb, kBlock, 0, st>>>(nbase, c.f64.as<double>(), c.f64.as<double>(), partial);
    reduce_bricks(c.partial, S_FF);
    ST_HIP(hipMemsetAsync(done, 0, ns, st));
    {
      EpiResid<double, double> er;
      er.partial = partial;
      er.b = c.f64.as<double>();
      er.r = c.r64.as<double>();
      launch_apply<double>(c, c.x64.as<double>(), er);
      reduce_bricks(c.partial, S_RR);
    }
    k_cg_start<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>(), siters.as<int>(), tol, mixed ? 1 : 0,
                                         swant.as<uint8_t>());
    if (precision == ST_MIXED32) {
      pcg32(max_iter, refresh);
      return;
    }
    {
      c.p64.ensure(v64);
      c.ap64.ensure(v64);
      c.z64.ensure(v64);
      c.t64.ensure(v64);
      c.u64.ensure(v64);
      c.w64.ensure(v64);
      double *x = c.x64.as<double>(), *r = c.r64.as<double>(), *p = c.p64.as<double>(), *ap = c.ap64.as<double>(),
             *z = c.z64.as<double>();
      // z = M^-1 r and the partial r.z: the V-cycle in FP64, or (mixed) in FP32 on the
      // rounded residual with z widened back to FP64 and r.z summed in FP64
      if (mixed) {
        c.r32.ensure(v32);
        c.z32.ensure(v32);
        c.t32.ensure(v32);
        c.p32.ensure(v32);
        c.w32.ensure(v32);
      }
      // (mixed: r32 is kept equal to r by the residual update; z stays FP32 in z32)
      float* r32 = mixed ? c.r32.as<float>() : nullptr;
      float* z32 = mixed ? c.z32.as<float>() : nullptr;
      // r.z of the mixed CG from the rounded r32 the V-cycle reads anyway (STATICS_HIP_RZ64=1:
      // from the FP64 r, 48 more bytes a node in the last smoothing pass)
      const char* rz = std::getenv("STATICS_HIP_RZ64");
      const double* rdot = rz && std::atoi(rz) ? r : nullptr;
      auto precond = [&](int slot) {
        if (!mixed)
          vcycle<double>(r, z, c.t64.as<double>(), c.u64.as<double>(), c.w64.as<double>());
        else
          vcycle<float>(r32, z32, c.t32.as<float>(), c.p32.as<float>(), c.w32.as<float>(), rdot);
        reduce_bricks(c.partial, slot);
      };
      if (mixed) k_copy_fine<float, double><<<nb, kBlock, 0, st>>>(nbase, bst, done, r, r32);
      precond(S_RZ);
      if (mixed)
        k_copy_fine<double, float><<<nb, kBlock, 0, st>>>(nbase, bst, done, z32, p);
      else
        k_copy_fine<double, double><<<nb, kBlock, 0, st>>>(nbase, bst, done, z, p);
      std::vector<uint8_t> dh, wh;
      int it = 0;
      // Every launch of an iteration (K p, the step length, the updates, the preconditioner
      // with its hundreds of coarse-level launches, the next direction) has fixed arguments:
      // with c.graphs, `chunk` whole iterations are captured once into one graph and each
      // chunk is one graph launch; the device-side done flags make a converged structure's
      // work a no-op inside it, and the host looks at the flags once per chunk.
      const int chunk = c.graphs ? 8 : 4;
      auto tail = [&]() {
        reduce_bricks(c.partial, S_PAP);
        k_cg_alpha<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>());
        k_update_xr<double><<<nb, kBlock, 0, st>>>(nbase, bst, done, scal, x, r, p, ap, partial, r32);
        reduce_bricks(c.partial, S_RR);
        // refresh < 0: the reference's rule, the recursive residual decides (bodies.rs:350-352);
        // otherwise it only calls for the true residual, which must meet the tolerance itself
        k_cg_check<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>(), siters.as<int>(), refresh < 0 ? 0 : 1,
                                             swant.as<uint8_t>(), 0.0);
        precond(S_RZNEW);
        k_cg_beta<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>());
        if (mixed)
          k_update_p<double, float><<<nb, kBlock, 0, st>>>(nbase, bst, done, scal, z32, p);
        else
          k_update_p<double, double><<<nb, kBlock, 0, st>>>(nbase, bst, done, scal, z, p);
      };
      auto iteration = [&]() {
        EpiDot<double> ed;
        ed.done = done;
        ed.partial = partial;
        ed.y = ap;
        launch_apply<double>(c, p, ed);
        tail();
      };
      GraphExec chunk_graph;
      while (it < max_iter) {
        const int k_end = std::min(max_iter, it + chunk);
        if (c.graphs && k_end - it == chunk) {
          chunk_graph.run(st, [&] {
            for (int k = 0; k < chunk; k++) iteration();
          });
          it = k_end;
        } else {
          for (; it < k_end; it++) iteration();
        }
        ST_HIP(hipGetLastError());
        sync_flags(dh, wh);
        bool any_want = false;
        for (int s = 0; s < ns; s++) any_want |= !dh[s] && wh[s];
        if (any_want || it >= max_iter) {
          // the recursive residual says done: replace it by the true one, f - K x, and go on if short
          EpiResid<double, double> er;
          er.done = done;
          er.partial = partial;
          er.b = c.f64.as<double>();
          er.r = r;
          launch_apply<double>(c, x, er);
          reduce_bricks(c.partial, S_RRTRUE);
          k_cg_true<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>(), swant.as<uint8_t>(),
                                              srefresh.as<int>(), kStall, 0.81);
          sync_flags(dh, wh);
        }
        if (std::all_of(dh.begin(), dh.end(), [](uint8_t d) { return d != 0; })) break;
      }
    }
    ST_HIP(hipGetLastError());
  }

  // The CG in FP32 (ST_MIXED32) after pcg's start (||f||^2, r64 = f - K x64, the scalars):
  // r, p, K p, z and the steps dx since the last replacement are FP32; inner products,
  // step lengths and x stay FP64. When a structure's FP32 residual has fallen a hundredfold
  // since the last replacement, or meets the tolerance, the steps are added into x in FP64
  // and the FP64 true residual f - K x replaces r (the direction p is kept).
  void pcg32(int max_iter, int refresh) {
    const unsigned nb = (unsigned)g.nbricks;
    const int64_t nn = g.nnodes;
    const size_t v32 = (size_t)std::max<int64_t>(nn, 1) * 6 * 4;
    const uint32_t* nbase = g.nbase.as<uint32_t>();
    const uint32_t* bst = g.bstruct.as<uint32_t>();
    uint8_t* done = c.done.as<uint8_t>();
    double* scal = c.scal.as<double>();
    double* partial = c.partial.as<double>();
    for (DBuf* b : {&c.r32, &c.z32, &c.t32, &c.p32, &c.w32, &c.ap32, &c.x32}) b->ensure(v32);
    float *r = c.r32.as<float>(), *z = c.z32.as<float>(), *p = c.p32.as<float>(), *q = c.ap32.as<float>(),
          *dx = c.x32.as<float>();
    double* x = c.x64.as<double>();
    // a fall of the recursive residual's square by this much since the last replacement asks for one
    const double drop2 = 1e-4;
    k_copy_fine<float, double><<<nb, kBlock, 0, st>>>(nbase, bst, nullptr, c.r64.as<double>(), r);
    k_zero_fine<<<nb, kBlock, 0, st>>>(nbase, dx);
    // (the V-cycle's scratch: t32, q (free until the next K p) and w32)
    auto precond = [&](int slot) {
      vcycle<float>(r, z, c.t32.as<float>(), q, c.w32.as<float>());
      reduce_bricks(c.partial, slot);
    };
    precond(S_RZ);
    k_copy_fine<float, float><<<nb, kBlock, 0, st>>>(nbase, bst, done, z, p);
    std::vector<uint8_t> dh, wh;
    int it = 0, last_replace = 0;
    const int chunk = 4;
    // a replacement at least every 32 iterations: an FP32 recursion that stagnates (stiff
    // materials) is then seen by the stall count, and the FP64 CG takes over
    const int every = 32;
    while (it < max_iter) {
      const int k_end = std::min(max_iter, it + chunk);
      for (; it < k_end; it++) {
        EpiDot<float> ed;
        ed.done = done;
        ed.partial = partial;
        ed.y = q;
        launch_apply<float>(c, p, ed);
        reduce_bricks(c.partial, S_PAP);
        k_cg_alpha<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>());
        k_update_xr<float><<<nb, kBlock, 0, st>>>(nbase, bst, done, scal, dx, r, p, q, partial, nullptr);
        reduce_bricks(c.partial, S_RR);
        // refresh < 0: the reference's rule, the recursive residual decides (bodies.rs:350-352);
        // it still asks for the replacements, which keep the FP32 recursion near the true residual
        k_cg_check<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>(), siters.as<int>(), 1,
                                             swant.as<uint8_t>(), drop2);
        precond(S_RZNEW);
        k_cg_beta<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>());
        k_update_p<float, float><<<nb, kBlock, 0, st>>>(nbase, bst, done, scal, z, p);
      }
      ST_HIP(hipGetLastError());
      sync_flags(dh, wh);
      bool any_want = false;
      for (int s2 = 0; s2 < ns; s2++) any_want |= !dh[s2] && wh[s2];
      if (any_want || it >= max_iter || it - last_replace >= every) {
        last_replace = it;
        k_flush_fine<<<nb, kBlock, 0, st>>>(nbase, dx, x);
        EpiResid<double, double> er;
        er.done = done;
        er.partial = partial;
        er.b = c.f64.as<double>();
        er.r = c.r64.as<double>();
        launch_apply<double>(c, x, er);
        reduce_bricks(c.partial, S_RRTRUE);
        if (refresh < 0) {
          // the recursive residual decides: the structures whose FP32 residual met the
          // tolerance stop here; the others go on from the true residual
          k_cg_true_rec<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>(), swant.as<uint8_t>(),
                                                  srefresh.as<int>());
        } else {
          k_cg_true<<<nblk(ns), 256, 0, st>>>(ns, scal, done, sstatus.as<int>(), swant.as<uint8_t>(),
                                              srefresh.as<int>(), 3, 0.25);
        }
        k_copy_fine<float, double><<<nb, kBlock, 0, st>>>(nbase, bst, done, c.r64.as<double>(), r);
        sync_flags(dh, wh);
      }
      if (std::all_of(dh.begin(), dh.end(), [](uint8_t d) { return d != 0; })) break;
    }
    k_flush_fine<<<nb, kBlock, 0, st>>>(nbase, dx, x);
    ST_HIP(hipGetLastError());
  }

  // final true residual per structure (FP64), into scal S_RRTRUE; ||f||^2 in S_FF
  void true_residual() {
    const unsigned nb = (unsigned)g.nbricks;
    uint8_t* done = c.done.as<uint8_t>();
    ST_HIP(hipMemsetAsync(done, 0, ns, st));
    EpiResid<double, double> er;
    er.partial = c.partial.as<double>();
    er.b = c.f64.as<double>();
    er.r = c.r64.as<double>();
    launch_apply<double>(c, c.x64.as<double>(), er);
    reduce_bricks(c.partial, S_RRTRUE);
    k_dot_fine<double><<<nb, kBlock, 0, st>>>(g.nbase.as<uint32_t>(), c.f64.as<double>(), c.f64.as<double>(),
                                              c.partial.as<double>());
    reduce_bricks(c.partial, S_FF);
  }
};

// ============================================================ entry helpers

void release_levels(Ctx& c) {
  c.hier_valid = false;
  c.restrict_fused = false;
  for (auto& lv : c.levels)
    for (DBuf* b : {&lv.rowptr, &lv.col, &lv.a64, &lv.a32, &lv.hblk, &lv.dinv64, &lv.dinv32, &lv.centre,
                    &lv.parent, &lv.rep,
                    &lv.key, &lv.kid_ptr, &lv.kids, &lv.agg_struct, &lv.step_d, &lv.r64, &lv.x64, &lv.t64, &lv.u64,
                    &lv.v64, &lv.r32, &lv.x32, &lv.t32, &lv.u32, &lv.v32, &lv.dense_a, &lv.dense_off,
                    &lv.dense_seg, &lv.prow, &lv.pcol, &lv.pval64, &lv.pval32, &lv.prr, &lv.prc, &lv.pt_ptr, &lv.pt_blk,
                    &lv.pblk_row})
      b->release();
  c.levels.clear();
}

// The hierarchy's device memory: the coarse levels, the fine smoother and the level-1
// aggregation with its prolongator. The next solve builds them again.
void release_hierarchy(Ctx& c) {
  release_levels(c);
  for (DBuf* b : {&c.dinv64, &c.dinv32, &c.label1, &c.nodes_by_agg, &c.agg_ptr, &c.brick_agg1, &c.p0row, &c.p0col,
                  &c.p0val, &c.p0t_ptr, &c.p0t_blk, &c.p0blk_row})
    b->release();
  c.p0blocks = 0;
}

void release_all(Ctx& c) {
  DevGeo& d = c.geo;
  for (DBuf* b : {&d.cbase, &d.nbase, &d.bstruct, &d.blist, &d.ce64, &d.cg64, &d.ce32, &d.cg32, &d.csec64,
                  &d.csec32, &d.cmask_user, &d.cmask, &d.cfaces, &d.ccoord, &d.cstruct, &d.nflags, &d.ncell, &d.nkind, &d.nkey,
                  &d.bgeo})
    b->release();
  for (DBuf* b : {&c.fres, &c.xsol, &c.dinv64, &c.dinv32, &c.step0_d, &c.label1, &c.uf, &c.uf2, &c.flag, &c.scan_tmp, &c.sort_tmp,
                  &c.nodes_by_agg, &c.agg_ptr, &c.f64, &c.x64, &c.r64, &c.p64, &c.ap64, &c.z64, &c.t64, &c.u64, &c.x32,
                  &c.r32, &c.p32, &c.ap32, &c.z32, &c.t32, &c.partial, &c.scal, &c.done, &c.w64, &c.w32, &c.brick_agg1, &c.p0row,
                  &c.p0col, &c.p0val, &c.p0t_ptr, &c.p0t_blk, &c.p0blk_row})
    b->release();
  release_levels(c);
}

void set_gone_flags(Ctx& c, const uint8_t* gone) {
  const int64_t nn = c.geo.nnodes;
  for (int64_t v = 0; v < nn; v++) {
    uint8_t& f = c.host.nflags[v];
    f = (uint8_t)((f & ~kGone) | ((gone && gone[v]) ? kGone : 0));
  }
  c.gone_host.assign(nn, 0);
  if (gone) std::memcpy(c.gone_host.data(), gone, nn);
  if (nn) ST_HIP(hipMemcpyAsync(c.geo.nflags.p, c.host.nflags.data(), nn, hipMemcpyHostToDevice, c.stream));
  refresh_cell_flags(c);
  ST_HIP(hipStreamSynchronize(c.stream));
  c.have_solution = false;
}

void set_beam_masks(Ctx& c, const uint16_t* beams) {
  HostGeo& h = c.host;
  const int64_t nin = (int64_t)h.cell_of_input.size();
  for (int64_t i = 0; i < nin; i++) {
    const int64_t cc = h.cell_of_input[i];
    if (cc >= 0) h.cmask[cc] = beams ? (uint16_t)(beams[i] & 0x7fff) : (uint16_t)0x7fff;
  }
  if (!h.cmask.empty())
    ST_HIP(hipMemcpyAsync(c.geo.cmask_user.p, h.cmask.data(), h.cmask.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                          c.stream));
  refresh_cell_flags(c);
  ST_HIP(hipStreamSynchronize(c.stream));
  c.have_solution = false;
}

// Runs setup + PCG for all uploaded structures. loads: device FP64 (c.f64), x0 in c.x64.
void solve_all(Ctx& c, const StSolveOpts& o, std::vector<StStats>& stats) {
  mem_tally().reset_peak();
  Solver S(c);
  const int ns = S.ns;
  hipEvent_t e0, e1, e2;
  ST_HIP(hipEventCreate(&e0));
  ST_HIP(hipEventCreate(&e1));
  ST_HIP(hipEventCreate(&e2));
  ST_HIP(hipEventRecord(e0, c.stream));
  // smoothed aggregation unless STATICS_HIP_RIGID is set (rigid aggregates only, for comparison)
  c.smoothed = std::getenv("STATICS_HIP_RIGID") == nullptr;
  // how many transfers are smoothed, from the finest up (the others are rigid): smoothing
  // every level widens the coarse stencils about 1.5x a level (STATICS_HIP_SA_LEVELS)
  c.sa_levels = 2;
  if (const char* e = std::getenv("STATICS_HIP_SA_LEVELS")) c.sa_levels = std::max(1, std::atoi(e));
  c.nu = 1;
  if (const char* e = std::getenv("STATICS_HIP_NU")) c.nu = std::max(1, std::atoi(e));
  // graph replay of the iteration tail: opt-in (STATICS_HIP_GRAPHS=1); at 256^3 the launch
  // queue stays full without it, and rocprofv3's kernel tracing does not follow captures
  // graph replay of whole iterations: on for batches (where the launches of many small
  // structures' tails would otherwise cost more than their work), off while K p is timed
  // (its events are host-side records) and for one structure, where the launch queue
  // stays full (STATICS_HIP_GRAPHS=0 or 1 forces it)
  c.graphs = c.geo.nstruct > 1 && c.timing_which < 0;
  if (const char* e = std::getenv("STATICS_HIP_GRAPHS")) c.graphs = std::atoi(e) != 0;
  if (c.timing_which >= 0) c.graphs = false;
  c.dense_top = std::getenv("STATICS_HIP_NODENSE") == nullptr;
  c.nu1 = 1;
  if (const char* e = std::getenv("STATICS_HIP_NU1")) c.nu1 = std::max(1, std::atoi(e));
  c.nu_c = 1;
  if (const char* e = std::getenv("STATICS_HIP_NUC")) c.nu_c = std::max(1, std::atoi(e));
  // gamma-cycles: each visit of coarse level index i runs gammas[i] cycles (the later
  // ones on the remaining residual). Measured at 256^3 (mixed, 1e-8): one cycle
  // everywhere 360 iterations; two on levels 3-6, 144; (1, 2, 4, 2, 2) on levels 1-5,
  // 84 iterations and the least time (level 6 is its dense top). On the 512^3 world
  // (10 levels) two cycles on level 6 as well take 192 iterations to 168.
  // STATICS_HIP_GAMMAS="a,b,..." sets the list; STATICS_HIP_GAMMA / GFROM / GTO set
  // one count for a range of level indices instead.
  c.gammas = {1, 2, 4, 2, 2, 2};
  c.gamma = 2;
  c.gamma_from = 1;
  c.gamma_to = 5;
  const char* eg = std::getenv("STATICS_HIP_GAMMA");
  const char* ef = std::getenv("STATICS_HIP_GFROM");
  const char* et = std::getenv("STATICS_HIP_GTO");
  if (eg || ef || et) {
    if (eg) c.gamma = std::max(1, std::atoi(eg));
    if (ef) c.gamma_from = std::max(0, std::atoi(ef));
    if (et) c.gamma_to = std::max(0, std::atoi(et));
    c.gammas.clear();
  }
  if (const char* e = std::getenv("STATICS_HIP_GAMMAS")) {
    c.gammas.clear();
    for (const char* p = e; *p;) {
      c.gammas.push_back(std::max(1, std::atoi(p)));
      while (*p && *p != ',') p++;
      if (*p == ',') p++;
    }
  }
  // the mixed V-cycle runs levels 1-4 in FP32 and levels 5 up (and the dense top) in FP64:
  // at 256^3 the iteration count is the same with every level in FP64 from 3, 5 or 7
  // up, and the FP32 levels cut 7% of the solve
  c.f64_from = 4;
  if (const char* e = std::getenv("STATICS_HIP_F64FROM")) c.f64_from = std::max(0, std::atoi(e));
  // the prolongator records in FP16 (block-scaled as the coarse operators) where the level
  // below runs its operator in FP16; STATICS_HIP_PHALF=0 keeps them FP32
  c.p_half = true;
  if (const char* e = std::getenv("STATICS_HIP_PHALF")) c.p_half = std::atoi(e) != 0;
  c.fold_switch = true;  // STATICS_HIP_FOLD=0: the switch to FP64 through separate conversions
  if (const char* e = std::getenv("STATICS_HIP_FOLD")) c.fold_switch = std::atoi(e) != 0;
  c.half_coarse = true;
  if (const char* e = std::getenv("STATICS_HIP_HALF")) c.half_coarse = std::atoi(e) != 0;
  autotune_apply(c);
  S.mixed = o.precision != ST_FP64;  // the FP32 V-cycle
  S.init_struct_arrays();
  const Ctx::HierKey key{S.mixed, c.smoothed, c.dense_top, c.sa_levels, c.f64_from, c.half_coarse};
  const bool kept = c.hier_valid && c.hier_key == key;
  if (kept) {
    S.ndof = c.last_ndof;
  } else {
    S.setup_fine();
    S.setup_hierarchy();
    c.hier_valid = true;
    c.hier_key = key;
  }
  ST_HIP(hipEventRecord(e1, c.stream));
  if (std::getenv("STATICS_HIP_VERBOSE")) {  // the hierarchy's shape, for tuning
    size_t mfree = 0, mtotal = 0;
    (void)hipMemGetInfo(&mfree, &mtotal);
    std::fprintf(stderr, "statics_hip: fine %lld nodes (%lld free), device memory in use %.1f GB\n",
                 (long long)c.geo.nnodes, (long long)S.ndof, (double)(mtotal - mfree) / 1e9);
    for (size_t l = 0; l < c.levels.size(); l++)
      std::fprintf(stderr, "statics_hip: level %zu cube %d: %lld aggregates, %lld blocks (%.1f/row), P below %lld\n",
                   l + 1, c.levels[l].cube, (long long)c.levels[l].n, (long long)c.levels[l].nblocks,
                   c.levels[l].n ? (double)c.levels[l].nblocks / c.levels[l].n : 0.0,
                   (long long)(l ? c.levels[l - 1].pblocks : c.p0blocks));
  }
  // ST_MIXED32: a structure whose FP32 CG stops short (stiff materials: FP32 K p loses the
  // cancellation a 1e-8 residual needs) goes on with the FP64 CG from where it stopped, on
  // the same FP32 V-cycle and hierarchy (the others are done at once: their true residual
  // already meets the tolerance).
  int prec_used = o.precision;
  std::vector<int> it_fp32(ns, 0), st_fp32(ns, ST_OK);
  bool fell_back = false;
  if (S.ndof > 0) {
    S.pcg(o.precision, o.tol, o.max_iter, o.refresh);
    if (o.precision == ST_MIXED32) {
      std::vector<int> sts(ns);
      ST_HIP(hipMemcpy(sts.data(), S.sstatus.p, ns * sizeof(int), hipMemcpyDeviceToHost));
      if (std::any_of(sts.begin(), sts.end(), [](int v) { return v == ST_NOT_SPD || v == ST_NO_CONVERGENCE; })) {
        fell_back = true;
        st_fp32 = sts;
        ST_HIP(hipMemcpy(it_fp32.data(), S.siters.p, ns * sizeof(int), hipMemcpyDeviceToHost));
        if (std::getenv("STATICS_HIP_VERBOSE"))
          std::fprintf(stderr, "statics_hip: the FP32 CG stopped short; on with the FP64 CG\n");
        prec_used = ST_MIXED;
        S.pcg(prec_used, o.tol, o.max_iter, o.refresh);
      }
    }
  }
  // The fast cycle settings (gamma counts up to 4, FP32 on coarse levels 3-4) can lose
  // definiteness on nearly singular structures, such as a settle round where the structure
  // comes apart. A solve they fail is repeated once, from zero, with the robust settings:
  // two cycles on levels 3-6 and FP64 from level 3 (a rebuilt hierarchy).
  std::vector<int> it_first(ns, 0), st_first(ns, ST_OK);
  bool retried = false;
  if (S.ndof > 0 && !std::getenv("STATICS_HIP_NORETRY")) {
    std::vector<int> sts(ns);
    ST_HIP(hipMemcpy(sts.data(), S.sstatus.p, ns * sizeof(int), hipMemcpyDeviceToHost));
    const bool bad = std::any_of(sts.begin(), sts.end(), [](int v) { return v == ST_NOT_SPD || v == ST_NO_CONVERGENCE; });
    const std::vector<int> robust = {1, 1, 2, 2, 2, 2};
    if (bad && (c.gammas != robust || c.f64_from > 2 || c.half_coarse)) {
      retried = true;
      st_first = sts;
      ST_HIP(hipMemcpy(it_first.data(), S.siters.p, ns * sizeof(int), hipMemcpyDeviceToHost));
      if (std::getenv("STATICS_HIP_VERBOSE"))
        std::fprintf(stderr, "statics_hip: solve failed with the fast cycle settings; again with the robust ones\n");
      c.gammas = robust;
      c.f64_from = std::min(c.f64_from, 2);
      c.half_coarse = false;
      S.setup_fine();
      S.setup_hierarchy();
      c.hier_valid = true;
      c.hier_key = Ctx::HierKey{S.mixed, c.smoothed, c.dense_top, c.sa_levels, c.f64_from, c.half_coarse};
      ST_HIP(hipMemsetAsync(c.x64.p, 0, (size_t)c.geo.nnodes * 6 * sizeof(double), c.stream));
      S.pcg(prec_used, o.tol, o.max_iter, o.refresh);
    }
  }
  S.true_residual();
  ST_HIP(hipEventRecord(e2, c.stream));
  ST_HIP(hipEventSynchronize(e2));
  float ms_setup = 0, ms_solve = 0;
  ST_HIP(hipEventElapsedTime(&ms_setup, e0, e1));
  ST_HIP(hipEventElapsedTime(&ms_solve, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  std::vector<double> sc((size_t)ns * kScal);
  std::vector<int> it(ns), sts(ns), rf(ns);
  ST_HIP(hipMemcpy(sc.data(), c.scal.p, sc.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (S.ndof > 0) {
    ST_HIP(hipMemcpy(it.data(), S.siters.p, ns * sizeof(int), hipMemcpyDeviceToHost));
    ST_HIP(hipMemcpy(sts.data(), S.sstatus.p, ns * sizeof(int), hipMemcpyDeviceToHost));
    ST_HIP(hipMemcpy(rf.data(), S.srefresh.p, ns * sizeof(int), hipMemcpyDeviceToHost));
  }
  stats.assign(ns, StStats{});
  for (int s = 0; s < ns; s++) {
    StStats& t = stats[s];
    const StructInfo& si = c.host.structs[s];
    t.iters = it[s] + it_first[s] + it_fp32[s];
    t.levels = 1 + (int)c.levels.size();
    t.refreshes = rf[s];
    const double ff = sc[(size_t)s * kScal + S_FF], rr = sc[(size_t)s * kScal + S_RRTRUE];
    t.rel_residual = ff > 0 ? std::sqrt(rr / ff) : std::sqrt(rr);
    t.status = sts[s];
    if (o.refresh < 0) {  // the recursive residual decides; rel_residual still reports the true one
      const double rr_rec = sc[(size_t)s * kScal + S_RR];
      if (t.status == ST_OK && !(rr_rec <= sc[(size_t)s * kScal + S_GOAL2] * (1.0 + 1e-9)) && ff > 0)
        t.status = ST_NO_CONVERGENCE;
    } else if (t.status == ST_OK && !(t.rel_residual <= o.tol * (1.0 + 1e-9)) && ff > 0) {
      t.status = ST_NO_CONVERGENCE;
    }
    if (t.status == ST_OK && !std::isfinite(t.rel_residual)) t.status = ST_NAN;
    t.setup_ms = ms_setup;
    t.solve_ms = ms_solve;
    t.nodes = si.nnodes;
    t.cells = si.ncells;
    t.bricks = si.nbricks;
    t.dof_nodes = s < (int)c.dof_per_struct.size() ? c.dof_per_struct[s] : 0;
    t.retries = (fell_back ? 1 : 0) + (retried ? 1 : 0);
    t.first_status = fell_back ? st_fp32[s] : retried ? st_first[s] : ST_OK;
  }
  if (o.release_hierarchy) release_hierarchy(c);
  const int64_t peak = mem_tally().peak.load();
  for (auto& t : stats) t.mem_peak = peak;
}

__global__ void k_gone_from_joined(int64_t n, const uint8_t* joined, uint8_t* nflags, int* count) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  nflags[v] = (uint8_t)((nflags[v] & ~kGone) | (joined[v] ? 0 : kGone));
  if (joined[v]) atomicAdd(count, 1);
}

// Reachability from held nodes over alive beams (statics/src/bodies.rs:557-574), then
// gone = !joined on the device.
void classify_device(Ctx& c, int64_t* joined_count, uint8_t* joined_host) {
  DevGeo& g = c.geo;
  const int64_t nn = g.nnodes;
  hipStream_t st = c.stream;
  if (!nn) {
    if (joined_count) *joined_count = 0;
    return;
  }
  c.uf.ensure(nn * sizeof(int));
  c.flag.ensure((nn + 1) * sizeof(int));
  c.uf2.ensure((nn + 1) * sizeof(int));
  uint8_t* out = reinterpret_cast<uint8_t*>(c.uf2.p);
  k_iota<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>());
  if (g.ncells)
    k_unite_all<<<nblk(g.ncells), 256, 0, st>>>(g.ncells, g.cfaces.as<uint32_t>(), g.cmask_user.as<uint16_t>(),
                                                c.uf.as<int>());
  k_compress<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>());
  ST_HIP(hipMemsetAsync(c.flag.p, 0, (nn + 1) * sizeof(int), st));
  k_held_roots<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>(), g.nflags.as<uint8_t>(), c.flag.as<int>());
  k_joined<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>(), c.flag.as<int>(), out);
  int* cnt = c.flag.as<int>() + nn;  // the last slot of flag, past the nodes
  ST_HIP(hipMemsetAsync(cnt, 0, sizeof(int), st));
  k_gone_from_joined<<<nblk(nn), 256, 0, st>>>(nn, out, g.nflags.as<uint8_t>(), cnt);
  ST_HIP(hipGetLastError());
  refresh_cell_flags(c);
  if (joined_host) ST_HIP(hipMemcpyAsync(joined_host, out, nn, hipMemcpyDeviceToHost, st));
  int h = 0;
  ST_HIP(hipMemcpyAsync(&h, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
  ST_HIP(hipStreamSynchronize(st));
  if (joined_count) *joined_count = h;
  c.have_solution = false;
}

// ---- joined_to_held without allocating: the scratch lives in a fixed device array
//
// The union-find parents and the held-root flags of st_joined_to_held_static, for up to
// kStaticNodes nodes, and one counter: 2 * 4 Mi + 64 int32 (32 MiB + 256 B), which each
// GPU the library runs on holds until the process ends. Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device address): it
// passes the address hipGetSymbolAddress gave for the device at st_create.
namespace {
constexpr int64_t kStaticNodes = int64_t(1) << 22;
__device__ __align__(256) int scratch[2 * kStaticNodes + 64];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
int* scratch_dev[kMaxDevices] = {};

// Sums joined[0..n) into *count, one atomic add per block.
__global__ __launch_bounds__(256) void k_count_joined(int64_t n, const uint8_t* joined, int* count) {
  __shared__ int part[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  part[threadIdx.x] = (i < n && joined[i]) ? 1 : 0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(count, part[0]);
}

}  // namespace

// The device address of the scratch array on `device`, which must be current. Called by
// st_create (once per device; later calls find it resolved).
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw StError{ST_BAD_INPUT, "device index beyond 63"};
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return;
  void* ptr = nullptr;
  ST_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw StError{ST_DEVICE, "hipGetSymbolAddress gave no address for the scratch array"};
  scratch_dev[device] = static_cast<int*>(ptr);
}

// joined_to_held over the context's beam masks into joined (device memory, one byte per
// node) and, if count is not NULL, the number joined into *count (device); the context's
// gone set is left as it is. Launches only, on stream s: no allocation, no copy, no wait.
void joined_static(Ctx& c, hipStream_t s, uint8_t* joined, int32_t* count) {
  DevGeo& g = c.geo;
  const int64_t nn = g.nnodes;
  if (nn > kStaticNodes) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "st_joined_to_held_static: %lld nodes, more than its scratch holds (%lld)",
                  (long long)nn, (long long)kStaticNodes);
    throw StError{ST_BAD_INPUT, msg};
  }
  int* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    base = (c.device >= 0 && c.device < kMaxDevices) ? scratch_dev[c.device] : nullptr;
  }
  if (!base) throw StError{ST_DEVICE, "st_joined_to_held_static: the scratch array was not set up on this device"};
  int* uf = base;
  int* flag = base + kStaticNodes;
  int* cnt = count ? count : base + 2 * kStaticNodes;
  k_fill_int<<<1, 64, 0, s>>>(1, cnt, 0);
  if (nn) {
    k_iota<<<nblk(nn), 256, 0, s>>>(nn, uf);
    if (g.ncells)
      k_unite_all<<<nblk(g.ncells), 256, 0, s>>>(g.ncells, g.cfaces.as<uint32_t>(), g.cmask_user.as<uint16_t>(), uf);
    k_compress<<<nblk(nn), 256, 0, s>>>(nn, uf);
    k_fill_int<<<nblk(nn), 256, 0, s>>>(nn, flag, 0);
    k_held_roots<<<nblk(nn), 256, 0, s>>>(nn, uf, g.nflags.as<uint8_t>(), flag);
    k_joined<<<nblk(nn), 256, 0, s>>>(nn, uf, flag, joined);
    k_count_joined<<<nblk(nn), 256, 0, s>>>(nn, joined, cnt);
  }
  ST_HIP(hipGetLastError());
}

double device_dot_f_x(Ctx& c, const double* a, const double* b) {
  DevGeo& g = c.geo;
  if (!g.nbricks) return 0.0;
  c.partial.ensure(g.nbricks * sizeof(double));
  k_dot_fine<double><<<(unsigned)g.nbricks, kBlock, 0, c.stream>>>(g.nbase.as<uint32_t>(), a, b, c.partial.as<double>());
  std::vector<double> h(g.nbricks);
  ST_HIP(hipMemcpyAsync(h.data(), c.partial.p, g.nbricks * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  ST_HIP(hipStreamSynchronize(c.stream));
  double sum = 0;
  for (double v : h) sum += v;
  return sum;
}

// Reachability from held nodes over alive beams (statics/src/bodies.rs:557-574).
void joined_to_held(Ctx& c, std::vector<uint8_t>& joined) {
  DevGeo& g = c.geo;
  const int64_t nn = g.nnodes;
  hipStream_t st = c.stream;
  joined.assign(nn, 0);
  if (!nn) return;
  c.uf.ensure(nn * sizeof(int));
  c.flag.ensure((nn + 1) * sizeof(int));
  DBuf out;
  out.ensure(nn);
  k_iota<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>());
  if (g.ncells)
    k_unite_all<<<nblk(g.ncells), 256, 0, st>>>(g.ncells, g.cfaces.as<uint32_t>(), g.cmask_user.as<uint16_t>(),
                                                c.uf.as<int>());
  k_compress<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>());
  ST_HIP(hipMemsetAsync(c.flag.p, 0, nn * sizeof(int), st));
  k_held_roots<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>(), g.nflags.as<uint8_t>(), c.flag.as<int>());
  k_joined<<<nblk(nn), 256, 0, st>>>(nn, c.uf.as<int>(), c.flag.as<int>(), out.as<uint8_t>());
  ST_HIP(hipGetLastError());
  ST_HIP(hipMemcpyAsync(joined.data(), out.p, nn, hipMemcpyDeviceToHost, st));
  ST_HIP(hipStreamSynchronize(st));
  out.release();
}

// Unites the ends of every alive edge of a general graph (a lattice's bonds).
__global__ void k_unite_edges(int64_t ne, const int64_t* ij, const uint8_t* alive, int* uf) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ne || (alive && !alive[b])) return;
  uf_unite(uf, (int)ij[2 * b], (int)ij[2 * b + 1]);
}

// bodies::joined_to_held (statics/src/bodies.rs:557-574) over a lattice's nodes and bonds as
// given: reachability from held nodes over the alive bonds, by union-find on the device.
void joined_edges(Ctx& c, int64_t nn, const uint8_t* held, int64_t ne, const int64_t* ij, const uint8_t* alive,
                  uint8_t* joined) {
  if (nn <= 0) return;
  if (nn >= (int64_t(1) << 31)) throw StError{ST_BAD_INPUT, "joined_edges: more nodes than int32 indexes"};
  for (int64_t b = 0; b < ne; b++)
    if (ij[2 * b] < 0 || ij[2 * b] >= nn || ij[2 * b + 1] < 0 || ij[2 * b + 1] >= nn)
      throw StError{ST_BAD_INPUT, "joined_edges: a bond names a missing node"};
  hipStream_t st = c.stream;
  DBuf uf, flag, hd, ijd, ad, out;
  uf.ensure(nn * sizeof(int));
  flag.ensure(nn * sizeof(int));
  hd.ensure(nn);
  out.ensure(nn);
  ijd.ensure(std::max<int64_t>(ne, 1) * 2 * sizeof(int64_t));
  if (alive) ad.ensure(std::max<int64_t>(ne, 1));
  for (int64_t v = 0; v < nn; v++)
    if (held[v] > 1) throw StError{ST_BAD_INPUT, "joined_edges: held is 0 or 1 per node"};
  ST_HIP(hipMemcpyAsync(hd.p, held, nn, hipMemcpyHostToDevice, st));
  if (ne) ST_HIP(hipMemcpyAsync(ijd.p, ij, ne * 2 * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (alive && ne) ST_HIP(hipMemcpyAsync(ad.p, alive, ne, hipMemcpyHostToDevice, st));
  k_iota<<<nblk(nn), 256, 0, st>>>(nn, uf.as<int>());
  if (ne)
    k_unite_edges<<<nblk(ne), 256, 0, st>>>(ne, ijd.as<int64_t>(), alive ? ad.as<uint8_t>() : nullptr, uf.as<int>());
  k_compress<<<nblk(nn), 256, 0, st>>>(nn, uf.as<int>());
  k_fill_int<<<nblk(nn), 256, 0, st>>>(nn, flag.as<int>(), 0);
  k_held_roots<<<nblk(nn), 256, 0, st>>>(nn, uf.as<int>(), hd.as<uint8_t>(), flag.as<int>());
  k_joined<<<nblk(nn), 256, 0, st>>>(nn, uf.as<int>(), flag.as<int>(), out.as<uint8_t>());
  ST_HIP(hipGetLastError());
  ST_HIP(hipMemcpyAsync(joined, out.p, nn, hipMemcpyDeviceToHost, st));
  ST_HIP(hipStreamSynchronize(st));
}

}  // namespace st

using namespace st;

extern "C" int st_joined_edges(StCtx* c, int64_t nnodes, const uint8_t* held, int64_t nbonds, const int64_t* ij,
                               const uint8_t* alive, uint8_t* joined_out) {
  return guard(c, [&] {
    if (nnodes < 0 || nbonds < 0 || (nnodes && (!held || !joined_out)) || (nbonds && !ij))
      throw StError{ST_BAD_INPUT, "st_joined_edges: bad arguments"};
    ST_HIP(hipSetDevice(c->device));
    joined_edges(*c, nnodes, held, nbonds, ij, alive, joined_out);
  });
}

extern "C" int st_joined_to_held_static(StCtx* c, void* stream, uint8_t* joined_out, int32_t* count_out) {
  return guard(c, [&] {
    need_upload(c);
    if (!joined_out) throw StError{ST_BAD_INPUT, "st_joined_to_held_static: joined_out is NULL"};
    ST_HIP(hipSetDevice(c->device));
    joined_static(*c, stream ? static_cast<hipStream_t>(stream) : c->stream, joined_out, count_out);
  });
}
