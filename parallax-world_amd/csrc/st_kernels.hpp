// Device kernels of the fine level: the matrix-free stiffness K x over 8^3
// bricks, its fused epilogues, and brick-blocked vector operations.
//
// K x (the reference's Level::apply at level 0, statics/src/bodies.rs:193-206,
// with the level's blocks assembled per bond in Level::new, :126-164) is
// computed cell by cell from the closed-form beams of st_beam.hpp. One
// workgroup owns one brick: it applies the brick's free cells and, for the
// nodes on the brick's lower faces, the free cells across them (the "halo"),
// accumulating into LDS with native ds_add; then every node of the brick has
// its complete row and an epilogue consumes it (store, residual, smoothing,
// dot product). No global atomics: every node is written by exactly one
// workgroup, and the sum into each node runs in a fixed set of terms.
#pragma once
#include "st_internal.hpp"

namespace st {

constexpr int kBlock = 256;

template <class T>
struct GeoD {
  const uint32_t *cbase, *nbase, *hbase, *hcell, *bstruct, *cfaces;
  const uint8_t* ho;
  const T *ce, *cg;
  const uint16_t* cmask;
  const BeamGeo* bgeo;
};

template <class T>
__device__ __forceinline__ void load6(const T* __restrict__ p, T* v) {
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = p[k];
}
template <>
__device__ __forceinline__ void load6<double>(const double* __restrict__ p, double* v) {
  const double2* q = reinterpret_cast<const double2*>(p);
  double2 a = q[0], b = q[1], c = q[2];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
template <>
__device__ __forceinline__ void load6<float>(const float* __restrict__ p, float* v) {
  const float2* q = reinterpret_cast<const float2*>(p);
  float2 a = q[0], b = q[1], c = q[2];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
template <class T>
__device__ __forceinline__ void store6(T* __restrict__ p, const T* v) {
#pragma unroll
  for (int k = 0; k < 6; k++) p[k] = v[k];
}
template <>
__device__ __forceinline__ void store6<double>(double* __restrict__ p, const double* v) {
  double2* q = reinterpret_cast<double2*>(p);
  q[0] = make_double2(v[0], v[1]);
  q[1] = make_double2(v[2], v[3]);
  q[2] = make_double2(v[4], v[5]);
}
template <>
__device__ __forceinline__ void store6<float>(float* __restrict__ p, const float* v) {
  float2* q = reinterpret_cast<float2*>(p);
  q[0] = make_float2(v[0], v[1]);
  q[1] = make_float2(v[2], v[3]);
  q[2] = make_float2(v[4], v[5]);
}

// One free cell's beams applied to x; the loads of its faces kept in this brick
// (node index in [n0, n0 + nn), free) are added into acc. keep: beams to apply;
// only: if >= 0, only that face's loads are kept (halo cells).
template <class T>
__device__ __forceinline__ void cell_work(const GeoD<T>& g, const BeamGeo& G, const T* __restrict__ x, uint32_t c,
                                          uint32_t keep, int only, T* acc, uint32_t n0, uint32_t nn) {
  T xv[6][6], yv[6][6];
  uint32_t fi[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    fi[f] = g.cfaces[6 * c + f];
    if (fi[f] & kNotFree) {
#pragma unroll
      for (int k = 0; k < 6; k++) xv[f][k] = T(0);
    } else {
      load6<T>(x + 6 * (size_t)fi[f], xv[f]);
    }
#pragma unroll
    for (int k = 0; k < 6; k++) yv[f][k] = T(0);
  }
  const T e = g.ce[c], gg = g.cg[c];
  const uint32_t mask = (uint32_t)g.cmask[c] & keep;
  const CellCoef<T> c0 = cell_coef<T>(G, 0, e, gg), c1 = cell_coef<T>(G, 1, e, gg);
  cell_apply<T>(c0, c1, mask, xv, yv);
#pragma unroll
  for (int f = 0; f < 6; f++) {
    if (only >= 0 && f != only) continue;
    if (fi[f] & kNotFree) continue;
    const uint32_t local = fi[f] - n0;
    if (local >= nn) continue;
#pragma unroll
    for (int k = 0; k < 6; k++) atomicAdd(&acc[6 * local + k], yv[f][k]);
  }
}

// Block-wide sum of one double per thread; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
  return s;
}

// y = K x over bricks, then epi.node(n, row, s) for each node of the brick.
// Epi::kReduce: the per-node values returned are summed per brick into epi.partial[b].
template <class T, class Epi>
__global__ __launch_bounds__(kBlock) void k_apply(GeoD<T> g, const T* __restrict__ x, Epi epi) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* acc = reinterpret_cast<T*>(smem_raw);
  __shared__ double red[kBlock / 64];
  const uint32_t b = blockIdx.x;
  const uint32_t s = g.bstruct[b];
  if (epi.skip(s)) {
    if (Epi::kReduce && threadIdx.x == 0) epi.partial[b] = 0.0;
    return;
  }
  const uint32_t n0 = g.nbase[b], nn = g.nbase[b + 1] - n0;
  for (uint32_t i = threadIdx.x; i < nn * 6; i += blockDim.x) acc[i] = T(0);
  const BeamGeo G = g.bgeo[s];
  __syncthreads();
  const uint32_t c_end = g.cbase[b + 1];
  for (uint32_t c = g.cbase[b] + threadIdx.x; c < c_end; c += blockDim.x) cell_work<T>(g, G, x, c, 0x7fffu, -1, acc, n0, nn);
  const uint32_t h_end = g.hbase[b + 1];
  for (uint32_t h = g.hbase[b] + threadIdx.x; h < h_end; h += blockDim.x) {
    const int f = 2 * (int)g.ho[h] + 1;
    cell_work<T>(g, G, x, g.hcell[h], face_beams(f), f, acc, n0, nn);
  }
  __syncthreads();
  double part = 0.0;
  for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x) part += epi.node(n0 + i, acc + 6 * i, s);
  if (Epi::kReduce) {
    double t = block_sum(part, red);
    if (threadIdx.x == 0) epi.partial[b] = t;
  }
}

// ---------------------------------------------------------------- epilogues
struct EpiBase {
  const uint8_t* done = nullptr;  // per structure: skip when set
  double* partial = nullptr;
  __device__ bool skip(uint32_t s) const { return done && done[s]; }
};

template <class T>
struct EpiStore : EpiBase {
  static constexpr bool kReduce = false;
  T* y;
  __device__ double node(uint32_t n, const T* a, uint32_t) const {
    store6<T>(y + 6 * (size_t)n, a);
    return 0.0;
  }
};

// y = K p and the partial p . y
template <class T>
struct EpiDot : EpiBase {
  static constexpr bool kReduce = true;
  T* y;
  const T* p;
  __device__ double node(uint32_t n, const T* a, uint32_t) const {
    T pv[6];
    load6<T>(p + 6 * (size_t)n, pv);
    store6<T>(y + 6 * (size_t)n, a);
    double d = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) d += (double)pv[k] * (double)a[k];
    return d;
  }
};

// r = b - K x (optionally the partial r . r)
template <class T, class TB>
struct EpiResid : EpiBase {
  static constexpr bool kReduce = true;
  const TB* b;
  T* r;
  __device__ double node(uint32_t n, const T* a, uint32_t) const {
    TB bv[6];
    load6<TB>(b + 6 * (size_t)n, bv);
    T rv[6];
    double d = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      rv[k] = (T)((T)bv[k] - a[k]);
      d += (double)rv[k] * (double)rv[k];
    }
    store6<T>(r + 6 * (size_t)n, rv);
    return d;
  }
};

// 6x6 row-major times vector
template <class T>
__device__ __forceinline__ void mat6v(const T* __restrict__ m, const T* v, T* out) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    T s = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) s += m[6 * i + j] * v[j];
    out[i] = s;
  }
}

// z = xin + step D^-1 (b - K xin); partial b . z
template <class T>
struct EpiSmooth : EpiBase {
  static constexpr bool kReduce = true;
  const T* b;
  const T* xin;
  const T* dinv;
  const double* step;  // per structure
  T* z;
  __device__ double node(uint32_t n, const T* a, uint32_t s) const {
    T bv[6], xv[6], t[6], d[6];
    load6<T>(b + 6 * (size_t)n, bv);
    load6<T>(xin + 6 * (size_t)n, xv);
#pragma unroll
    for (int k = 0; k < 6; k++) t[k] = bv[k] - a[k];
    mat6v<T>(dinv + 36 * (size_t)n, t, d);
    const T w = (T)step[s];
    double acc = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      xv[k] += w * d[k];
      acc += (double)bv[k] * (double)xv[k];
    }
    store6<T>(z + 6 * (size_t)n, xv);
    return acc;
  }
};

// r1 = b - K x and s = step D^-1 r1 (the restriction's smoothing step, fused)
template <class T>
struct EpiResidJac : EpiBase {
  static constexpr bool kReduce = false;
  const T* b;
  const T* dinv;
  const double* step;
  T* r1;
  T* sout;
  __device__ double node(uint32_t n, const T* a, uint32_t s) const {
    T bv[6], t[6], d[6];
    load6<T>(b + 6 * (size_t)n, bv);
#pragma unroll
    for (int k = 0; k < 6; k++) t[k] = bv[k] - a[k];
    store6<T>(r1 + 6 * (size_t)n, t);
    mat6v<T>(dinv + 36 * (size_t)n, t, d);
    const T w = (T)step[s];
#pragma unroll
    for (int k = 0; k < 6; k++) d[k] *= w;
    store6<T>(sout + 6 * (size_t)n, d);
    return 0.0;
  }
};

// x += u - step D^-1 K u, u the rigid prolongation (the smoothed prolongator, matrix-free)
template <class T>
struct EpiProlong : EpiBase {
  static constexpr bool kReduce = false;
  const T* u;
  const T* dinv;
  const double* step;
  T* x;
  __device__ double node(uint32_t n, const T* a, uint32_t s) const {
    T uv[6], xv[6], d[6];
    load6<T>(u + 6 * (size_t)n, uv);
    load6<T>(x + 6 * (size_t)n, xv);
    mat6v<T>(dinv + 36 * (size_t)n, a, d);
    const T w = (T)step[s];
#pragma unroll
    for (int k = 0; k < 6; k++) xv[k] += uv[k] - w * d[k];
    store6<T>(x + 6 * (size_t)n, xv);
    return 0.0;
  }
};

}  // namespace st
