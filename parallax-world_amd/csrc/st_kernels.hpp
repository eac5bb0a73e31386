// Device kernels of the fine level: the matrix-free stiffness K x over 8^3
// bricks, its fused epilogues, and brick-blocked vector operations.
//
// K x (the reference's Level::apply at level 0, statics/src/bodies.rs:193-206,
// with the level's blocks assembled per bond in Level::new, :126-164) is
// computed node by node from the closed-form beams of st_beam.hpp. One
// workgroup owns one brick: it stages the brick's node motions and cell table in
// LDS with coalesced reads, then one thread per node gathers the motions of the
// faces of the node's (at most two) cells and sums the loads of the (at most
// ten) beams ending on the node, and an epilogue consumes the row in registers
// (store, residual, smoothing, dot product). No atomics: every row is summed by
// one thread in a fixed order.
#pragma once
#include "st_internal.hpp"

namespace st {

constexpr int kBlock = 256;

template <class T>
struct GeoD {
  const uint32_t *cbase, *nbase, *bstruct, *cfaces;
  const uint8_t *nflags, *nkind;
  const int32_t* ncell;
  int xcd;  // deal bricks to XCDs in contiguous runs
  const T *ce, *cg;
  const uint16_t* cmask;
  const BeamGeo* bgeo;
  const T* carea;  // 15 section factors per cell, or NULL (every beam SECTION s^2)
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

// A brick's share of the operator's inputs staged in LDS: the motions of its first
// ns nodes and the face tables, e, g and beam masks of its first ncs cells (both
// read coalesced). Faces and cells outside the staged ranges (the neighbours across
// the brick's lower faces, and the rest of very full bricks) are read from memory.
// How much to stage is a schedule choice (LDS per workgroup against waves per CU):
// the kernel is instantiated for several (nodes, cells) capacities and the fastest
// is picked by timing on the uploaded structure.
// Pointers into LDS, so that loads through them are ds_read (not flat) loads.
template <class T>
using lds_cptr = const T __attribute__((address_space(3)))*;

template <class T>
struct Tile {
  lds_cptr<T> xs;
  lds_cptr<uint32_t> cf;
  lds_cptr<T> ce;
  lds_cptr<T> cg;
  lds_cptr<uint16_t> cm;
  uint32_t n0, ns, c0, ncs;
};

template <class T>
__device__ __forceinline__ void lds_load6(lds_cptr<T> p, T* v) {
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = p[k];
}

// The load at face F of cell c from the cell's live beams that end there (the five beams
// (F, f2)), written into y. F is known at compile time, so each beam's direction, class and
// end are constants. kSec: the cells carry per-beam sections.
template <class T, int F, bool kStaged, bool kSec>
__device__ __forceinline__ void cell_face_row(const GeoD<T>& g, const Tile<T>& t, const BeamGeo& G,
                                              const T* __restrict__ x, uint32_t c, T* y) {
  const uint32_t lc = c - t.c0;
  const bool staged = kStaged && lc < t.ncs;
  const uint32_t mask = (uint32_t)(staged ? t.cm[lc] : g.cmask[c]) & face_beams(F);
  if (!mask) {
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] = T(0);
    return;
  }
  uint32_t fi[6];
  if (staged) {
#pragma unroll
    for (int f = 0; f < 6; f++) fi[f] = t.cf[6 * lc + f];
  } else {
    const uint2* fp = reinterpret_cast<const uint2*>(g.cfaces + 6 * (size_t)c);
    const uint2 f01 = fp[0], f23 = fp[1], f45 = fp[2];
    fi[0] = f01.x; fi[1] = f01.y; fi[2] = f23.x; fi[3] = f23.y; fi[4] = f45.x; fi[5] = f45.y;
  }
  // Every face is a node (held and gone ones too), so its row is loaded unconditionally
  // and zeroed after for a node that is not free: the loads issue back to back.
  T xv[6][6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const uint32_t v = fi[f] & ~kNotFree;
    const uint32_t ln = v - t.n0;
    if (kStaged && ln < t.ns)
      lds_load6<T>(t.xs + 6 * ln, xv[f]);
    else
      load6<T>(x + 6 * (size_t)v, xv[f]);
  }
#pragma unroll
  for (int f = 0; f < 6; f++)
#pragma unroll
    for (int k = 0; k < 6; k++) xv[f][k] = (fi[f] & kNotFree) ? T(0) : xv[f][k];
  const T e = staged ? t.ce[lc] : g.ce[c], gg = staged ? t.cg[lc] : g.cg[c];
  const CellCoef<T> c0 = cell_coef<T>(G, 0, e, gg), c1 = cell_coef<T>(G, 1, e, gg);
  const T* rho = kSec ? g.carea + 15 * (size_t)c : nullptr;
  if (mask == face_beams(F)) {
    face_row_full<T, F, kSec>(c0, c1, rho, xv, y);
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] = T(0);
    face_row_masked<T, F, kSec>(c0, c1, mask, rho, xv, y);
  }
}

// Row of K x at a free node of orientation O: face 2O+1 of the cell below (c-) and
// face 2O of the cell above (c+), where those cells are free.
// (y is written: the sum of the two cells' loads in that order, or the one cell's.)
template <class T, int O, bool kStaged, bool kSec>
__device__ __forceinline__ void node_row(const GeoD<T>& g, const Tile<T>& t, const BeamGeo& G,
                                         const T* __restrict__ x, int2 cells, T* y) {
  T ya[6], yb[6];
  const bool ha = cells.x >= 0, hb = cells.y >= 0;
  if (ha) cell_face_row<T, 2 * O + 1, kStaged, kSec>(g, t, G, x, (uint32_t)cells.x, ya);
  if (hb) cell_face_row<T, 2 * O, kStaged, kSec>(g, t, G, x, (uint32_t)cells.y, yb);
#pragma unroll
  for (int k = 0; k < 6; k++) y[k] = ha ? (hb ? ya[k] + yb[k] : ya[k]) : (hb ? yb[k] : T(0));
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

// y = K x, one workgroup per brick and one thread per node: every node gathers the
// beams of its two cells that end on it, so each row is summed by one thread in a
// fixed order, with no atomics, and epi.node(n, row, s, xn) consumes it in
// registers (xn: the node's own entries of x, from the tile where staged).
// Epi::kReduce: the per-node values returned are summed per brick into epi.partial[b].
template <class T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// Workgroups are dealt to the 8 XCDs in turn (workgroup i runs on XCD i mod 8); this
// maps each XCD's workgroups to one contiguous run of bricks, so bricks that are
// neighbours in Morton order, and share halo nodes and cells, meet in the same L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_brick(uint32_t w, uint32_t nb) {
  const uint32_t q = nb / kXcds, r = nb % kXcds;  // XCD k takes q bricks, plus one if k < r
  const uint32_t k = w % kXcds, i = w / kXcds;
  return k * q + min(k, r) + i;
}

// Waves per SIMD the operator's register allocation aims at (it is latency-bound:
// more waves in flight hide the gathers).
#ifndef ST_APPLY_WAVES
#define ST_APPLY_WAVES 1
#endif

template <class T, class Epi, int kNodes, int kCells, bool kSec, int kThreads = kBlock>
__global__ __launch_bounds__(kThreads, ST_APPLY_WAVES) void k_apply(GeoD<T> g, const T* __restrict__ x, Epi epi) {
  using V2 = typename Vec2<T>::type;
  __shared__ V2 xs2[kNodes > 0 ? kNodes * 3 : 1];
  __shared__ uint2 cf2[kCells > 0 ? kCells * 3 : 1];
  __shared__ T ces[kCells > 0 ? kCells : 1], cgs[kCells > 0 ? kCells : 1];
  __shared__ uint16_t cms[kCells > 0 ? kCells : 1];
  __shared__ double red[kBlock / 64];
  const uint32_t b = g.xcd ? xcd_brick(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t s = g.bstruct[b];
  if (epi.skip(s)) {
    if (Epi::kReduce && threadIdx.x == 0) epi.partial[b] = 0.0;
    return;
  }
  Tile<T> t;
  t.n0 = g.nbase[b];
  const uint32_t n_end = g.nbase[b + 1];
  t.ns = min(n_end - t.n0, (uint32_t)kNodes);
  t.c0 = g.cbase[b];
  t.ncs = min(g.cbase[b + 1] - t.c0, (uint32_t)kCells);
  {
    const V2* xg = reinterpret_cast<const V2*>(x + 6 * (size_t)t.n0);
    for (uint32_t i = threadIdx.x; i < 3 * t.ns; i += blockDim.x) xs2[i] = xg[i];
    const uint2* cg2 = reinterpret_cast<const uint2*>(g.cfaces + 6 * (size_t)t.c0);
    for (uint32_t i = threadIdx.x; i < 3 * t.ncs; i += blockDim.x) cf2[i] = cg2[i];
    for (uint32_t i = threadIdx.x; i < t.ncs; i += blockDim.x) {
      ces[i] = g.ce[t.c0 + i];
      cgs[i] = g.cg[t.c0 + i];
      cms[i] = g.cmask[t.c0 + i];
    }
  }
  t.xs = (lds_cptr<T>)(reinterpret_cast<const T*>(xs2));
  t.cf = (lds_cptr<uint32_t>)(reinterpret_cast<const uint32_t*>(cf2));
  t.ce = (lds_cptr<T>)ces;
  t.cg = (lds_cptr<T>)cgs;
  t.cm = (lds_cptr<uint16_t>)cms;
  const BeamGeo G = g.bgeo[s];
  __syncthreads();
  double part = 0.0;
  for (uint32_t v = t.n0 + threadIdx.x; v < n_end; v += blockDim.x) {
    T y[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (!(g.nflags[v] & (kHeld | kGone))) {
      const int o = g.nkind[v] >> 1;
      const int2 cells = reinterpret_cast<const int2*>(g.ncell)[v];
      if (o == 0)
        node_row<T, 0, (kNodes > 0), kSec>(g, t, G, x, cells, y);
      else if (o == 1)
        node_row<T, 1, (kNodes > 0), kSec>(g, t, G, x, cells, y);
      else
        node_row<T, 2, (kNodes > 0), kSec>(g, t, G, x, cells, y);
    }
    const uint32_t lv = v - t.n0;
    T xo[6];  // the node's own entries of x, for epilogues that use them
    if (lv < t.ns)
      lds_load6<T>(t.xs + 6 * lv, xo);
    else
      load6<T>(x + 6 * (size_t)v, xo);
    part += epi.node(v, y, s, xo);
  }
  if (Epi::kReduce) {
    const double tot = block_sum(part, red);
    if (threadIdx.x == 0) epi.partial[b] = tot;
  }
  if constexpr (Epi::kTail) {
    __syncthreads();
    epi.tail(b, s);
  }
}

// Face F's load in one cell from the cell's mask (all beams), its face motions xv (zero on
// faces that are not free), e and g: written into y.
template <class T, int F, bool kSec>
__device__ __forceinline__ void face_row_of(const BeamGeo& G, T e, T gg, uint32_t mask, const T* rho,
                                            const T (*xv)[6], T* y) {
  mask &= face_beams(F);
  const CellCoef<T> c0 = cell_coef<T>(G, 0, e, gg), c1 = cell_coef<T>(G, 1, e, gg);
  if (mask == face_beams(F)) {
    face_row_full<T, F, kSec>(c0, c1, rho, xv, y);
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] = T(0);
    if (mask) face_row_masked<T, F, kSec>(c0, c1, mask, rho, xv, y);
  }
}

__device__ __forceinline__ void load_faces(const uint32_t* cfaces, uint32_t c, uint32_t* fi) {
  const uint2* fp = reinterpret_cast<const uint2*>(cfaces + 6 * (size_t)c);
  const uint2 f01 = fp[0], f23 = fp[1], f45 = fp[2];
  fi[0] = f01.x; fi[1] = f01.y; fi[2] = f23.x; fi[3] = f23.y; fi[4] = f45.x; fi[5] = f45.y;
}

// The operator node by node with all of a node's loads issued before its arithmetic: its
// two cells' face tables, masks, e and g together, then their twelve face rows together,
// so that a node waits on three trips to memory (node, cells, faces) where node_row's
// cell-after-cell order waits on five. One workgroup per brick, no LDS tile.
template <class T, class Epi, bool kSec>
__global__ __launch_bounds__(kBlock) void k_apply_nodes2(GeoD<T> g, const T* __restrict__ x, Epi epi) {
  __shared__ double red[kBlock / 64];
  const uint32_t b = g.xcd ? xcd_brick(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t s = g.bstruct[b];
  if (epi.skip(s)) {
    if (Epi::kReduce && threadIdx.x == 0) epi.partial[b] = 0.0;
    return;
  }
  const uint32_t n0 = g.nbase[b], n_end = g.nbase[b + 1];
  const BeamGeo G = g.bgeo[s];
  double part = 0.0;
  for (uint32_t v = n0 + threadIdx.x; v < n_end; v += blockDim.x) {
    typename Epi::Pre pr;
    if constexpr (Epi::kPre) pr = epi.pre(v);  // in flight during the gathers below
    const uint8_t fl = g.nflags[v], kind = g.nkind[v];
    const int2 cells = reinterpret_cast<const int2*>(g.ncell)[v];
    const bool live = !(fl & (kHeld | kGone));
    const bool ha = live && cells.x >= 0, hb = live && cells.y >= 0;
    // cell 0 stands in for an absent cell: its loads are in bounds and their values unused
    const uint32_t ca = ha ? (uint32_t)cells.x : 0u, cb = hb ? (uint32_t)cells.y : 0u;
    uint32_t fa[6], fb[6];
    load_faces(g.cfaces, ca, fa);
    load_faces(g.cfaces, cb, fb);
    const uint32_t ma = g.cmask[ca], mb = g.cmask[cb];
    const T ea = g.ce[ca], ga = g.cg[ca], eb = g.ce[cb], gb = g.cg[cb];
    T xa[6][6], xb[6][6];
#pragma unroll
    for (int f = 0; f < 6; f++) {
      load6<T>(x + 6 * (size_t)(fa[f] & ~kNotFree), xa[f]);
      load6<T>(x + 6 * (size_t)(fb[f] & ~kNotFree), xb[f]);
    }
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
      for (int k = 0; k < 6; k++) {
        xa[f][k] = (fa[f] & kNotFree) ? T(0) : xa[f][k];
        xb[f][k] = (fb[f] & kNotFree) ? T(0) : xb[f][k];
      }
    const T* rhoa = kSec ? g.carea + 15 * (size_t)ca : nullptr;
    const T* rhob = kSec ? g.carea + 15 * (size_t)cb : nullptr;
    T ya[6], yb[6], xo[6];
    const int o = kind >> 1;
    // the node is face 2o+1 of the cell below it (a) and face 2o of the cell above (b)
    if (o == 0) {
      face_row_of<T, 1, kSec>(G, ea, ga, ma, rhoa, xa, ya);
      face_row_of<T, 0, kSec>(G, eb, gb, mb, rhob, xb, yb);
#pragma unroll
      for (int k = 0; k < 6; k++) xo[k] = ha ? xa[1][k] : xb[0][k];
    } else if (o == 1) {
      face_row_of<T, 3, kSec>(G, ea, ga, ma, rhoa, xa, ya);
      face_row_of<T, 2, kSec>(G, eb, gb, mb, rhob, xb, yb);
#pragma unroll
      for (int k = 0; k < 6; k++) xo[k] = ha ? xa[3][k] : xb[2][k];
    } else {
      face_row_of<T, 5, kSec>(G, ea, ga, ma, rhoa, xa, ya);
      face_row_of<T, 4, kSec>(G, eb, gb, mb, rhob, xb, yb);
#pragma unroll
      for (int k = 0; k < 6; k++) xo[k] = ha ? xa[5][k] : xb[4][k];
    }
    T y[6];
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] = ha ? (hb ? ya[k] + yb[k] : ya[k]) : (hb ? yb[k] : T(0));
    if (!ha && !hb) load6<T>(x + 6 * (size_t)v, xo);  // a held or gone node (its row of x)
    if constexpr (Epi::kPre)
      part += epi.node_pre(v, y, s, xo, pr);
    else
      part += epi.node(v, y, s, xo);
  }
  if (Epi::kReduce) {
    const double tot = block_sum(part, red);
    if (threadIdx.x == 0) epi.partial[b] = tot;
  }
  if constexpr (Epi::kTail) {
    __syncthreads();
    epi.tail(b, s);
  }
}

// The operator cell by cell (the other schedule): one workgroup per brick, one
// thread per free cell applies the cell's 15 beams once and adds the loads of each
// face into the LDS row of that face's node when the node is the brick's. A node
// gets at most two adds, one from each of its cells, and 0 + a + b is the same in
// either order, so the rows are exact and deterministic. A node on the brick's lower
// faces whose other cell lies in another brick takes that cell's five beams ending on
// it (the halo). Then the epilogue runs per node. blist: the bricks of this launch
// (bricks are launched in classes by node count, each with LDS for its largest).
template <class T, class Epi, bool kSec, int kThreads = kBlock>
__global__ __launch_bounds__(kThreads) void k_apply_cells(GeoD<T> g, const uint32_t* __restrict__ blist,
                                                        const T* __restrict__ x, Epi epi) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* yl = reinterpret_cast<T*>(smem_raw);
  __shared__ double red[kBlock / 64];
  const uint32_t b = blist[g.xcd ? xcd_brick(blockIdx.x, gridDim.x) : blockIdx.x];
  const uint32_t s = g.bstruct[b];
  if (epi.skip(s)) {
    if (Epi::kReduce && threadIdx.x == 0) epi.partial[b] = 0.0;
    return;
  }
  const uint32_t n0 = g.nbase[b], nn = g.nbase[b + 1] - n0;
  const uint32_t c0 = g.cbase[b], c1 = g.cbase[b + 1];
  for (uint32_t i = threadIdx.x; i < 6 * nn; i += blockDim.x) yl[i] = T(0);
  const BeamGeo G = g.bgeo[s];
  __syncthreads();
  double part = 0.0;
  for (uint32_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
    const uint2* fp = reinterpret_cast<const uint2*>(g.cfaces + 6 * (size_t)c);
    const uint2 f01 = fp[0], f23 = fp[1], f45 = fp[2];
    const uint32_t fi[6] = {f01.x, f01.y, f23.x, f23.y, f45.x, f45.y};
    T xv[6][6], yv[6][6];
#pragma unroll
    for (int f = 0; f < 6; f++) load6<T>(x + 6 * (size_t)(fi[f] & ~kNotFree), xv[f]);  // back to back
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
      for (int k = 0; k < 6; k++) xv[f][k] = (fi[f] & kNotFree) ? T(0) : xv[f][k];
    const T e = g.ce[c], gg = g.cg[c];
    cell_apply<T, kSec>(cell_coef<T>(G, 0, e, gg), cell_coef<T>(G, 1, e, gg), (uint32_t)g.cmask[c], xv, yv,
                        kSec ? g.carea + 15 * (size_t)c : nullptr);
    if constexpr (Epi::kCellDot) {  // every (cell, face) counted once, in the cell's own brick
#pragma unroll
      for (int f = 0; f < 6; f++)
#pragma unroll
        for (int k = 0; k < 6; k++) part += (double)xv[f][k] * (double)yv[f][k];
    }
#pragma unroll
    for (int f = 0; f < 6; f++) {
      if (fi[f] & kNotFree) continue;
      const uint32_t l = fi[f] - n0;
      if (l >= nn) continue;  // kept by the next brick, which takes this cell as its halo
#pragma unroll
      for (int k = 0; k < 6; k++) atomicAdd(&yl[6 * l + k], yv[f][k]);
    }
  }
  Tile<T> none{};  // the halo reads memory
  for (uint32_t l = threadIdx.x; l < nn; l += blockDim.x) {
    const uint32_t v = n0 + l;
    const uint8_t kind = g.nkind[v];
    if ((kind & 1) || (g.nflags[v] & (kHeld | kGone))) continue;
    const int32_t cm = g.ncell[2 * (size_t)v];
    if (cm < 0 || ((uint32_t)cm >= c0 && (uint32_t)cm < c1)) continue;
    T y6[6];
    const int o = kind >> 1;
    if (o == 0)
      cell_face_row<T, 1, false, kSec>(g, none, G, x, (uint32_t)cm, y6);
    else if (o == 1)
      cell_face_row<T, 3, false, kSec>(g, none, G, x, (uint32_t)cm, y6);
    else
      cell_face_row<T, 5, false, kSec>(g, none, G, x, (uint32_t)cm, y6);
#pragma unroll
    for (int k = 0; k < 6; k++) atomicAdd(&yl[6 * l + k], y6[k]);
  }
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < nn; l += blockDim.x) {
    const uint32_t v = n0 + l;
    T y[6];
#pragma unroll
    for (int k = 0; k < 6; k++) y[k] = yl[6 * l + k];
    if constexpr (Epi::kCellDot) {
      epi.store(v, y);
    } else {
      T xo[6];
      load6<T>(x + 6 * (size_t)v, xo);
      part += epi.node(v, y, s, xo);
    }
  }
  if (Epi::kReduce) {
    const double tot = block_sum(part, red);
    if (threadIdx.x == 0) epi.partial[b] = tot;
  }
  if constexpr (Epi::kTail) {
    __syncthreads();
    epi.tail(b, s);
  }
}

// ---------------------------------------------------------------- node geometry
// Position of node v (its face's centre, metres, in its structure's frame) at spacing s.
__device__ __forceinline__ void node_pos(uint32_t v, const uint32_t* nkey, const uint8_t* nkind, double s, double* p) {
  const uint32_t k = nkey[v];
  const int kind = nkind[v], o = kind >> 1;
  p[0] = (double)(k & 1023u) + 0.5;
  p[1] = (double)((k >> 10) & 1023u) + 0.5;
  p[2] = (double)((k >> 20) & 1023u) + 0.5;
  p[o] += (kind & 1) ? 0.5 : -0.5;
  p[0] *= s;
  p[1] *= s;
  p[2] *= s;
}

// A load [f; m] at offset r taken to the centre, [f; m + r x f]: the transpose of the
// rigid shift [u + w x r; w] (statics/src/bodies.rs:44-47).
template <class T>
__device__ __forceinline__ void shiftT_add(const T* f, const double* r, T* out) {
  const T r0 = (T)r[0], r1 = (T)r[1], r2 = (T)r[2];
  out[0] += f[0];
  out[1] += f[1];
  out[2] += f[2];
  out[3] += f[3] + (r1 * f[2] - r2 * f[1]);
  out[4] += f[4] + (r2 * f[0] - r0 * f[2]);
  out[5] += f[5] + (r0 * f[1] - r1 * f[0]);
}

// ---------------------------------------------------------------- epilogues
struct EpiBase {
  const uint8_t* done = nullptr;  // per structure: skip when set
  double* partial = nullptr;
  __device__ bool skip(uint32_t s) const { return done && done[s]; }
  // kTail: after every node of the brick is done (and the block synchronised), the
  // workgroup runs tail(b, s) once more over the brick
  static constexpr bool kTail = false;
  __device__ void tail(uint32_t, uint32_t) const {}
  // kCellDot: the reduced value is x.(K x), which the cell-by-cell schedule sums cell by
  // cell from the motions and loads it already holds (x^T K x = sum over cells of
  // x_c . (K_c x_c)), so that its epilogue only stores the row (store()).
  static constexpr bool kCellDot = false;
  // kPre: the epilogue's own reads of node n (Pre, from pre(n)) can be issued before the
  // operator's gathers and consumed by node_pre(); k_apply_nodes2 does so, overlapping
  // their latency with the gathers'
  static constexpr bool kPre = false;
  struct Pre {};
};

template <class T>
struct EpiStore : EpiBase {
  static constexpr bool kReduce = false;
  T* y;
  __device__ double node(uint32_t n, const T* a, uint32_t, const T*) const {
    store6<T>(y + 6 * (size_t)n, a);
    return 0.0;
  }
};

// y = K p and the partial p . y (p is the operator's argument)
template <class T>
struct EpiDot : EpiBase {
  static constexpr bool kReduce = true;
  static constexpr bool kCellDot = true;
  T* y;
  __device__ void store(uint32_t n, const T* a) const { store6<T>(y + 6 * (size_t)n, a); }
  __device__ double node(uint32_t n, const T* a, uint32_t, const T* pn) const {
    T pv[6];
    load6<T>(pn, pv);
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
  __device__ double node(uint32_t n, const T* a, uint32_t, const T*) const {
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

// The residual u = b - K x (into r, in place allowed) followed, per brick, by the rigid
// restriction to level 1, c_I = sum over the aggregate's nodes v of Sh(p_v - centre_I)^T u_v
// (statics/src/bodies.rs:44-47 transposed). Eight lanes take an aggregate, lane k its nodes
// k, k + 8, ... in node order, and a butterfly over the eight lanes adds their sums (a fixed
// order). Level-1 aggregates lie inside one brick (checked at setup), so the workgroup that
// wrote a brick's residuals restricts them; brick_agg[b] .. brick_agg[b + 1] are the
// brick's aggregates.
template <class T>
struct EpiResidRestrict : EpiBase {
  static constexpr bool kReduce = false;
  static constexpr bool kTail = true;
  const T* b;
  T* r;
  const int *brick_agg, *agg_ptr, *nodes;
  const double* centre;
  const uint32_t* nkey;
  const uint8_t* nkind;
  const BeamGeo* bgeo;
  T* c;
  static constexpr bool kPre = true;
  struct Pre {
    T b[6];
  };
  __device__ Pre pre(uint32_t n) const {
    Pre p;
    load6<T>(b + 6 * (size_t)n, p.b);
    return p;
  }
  __device__ double node(uint32_t n, const T* a, uint32_t s, const T* xn) const { return node_pre(n, a, s, xn, pre(n)); }
  __device__ double node_pre(uint32_t n, const T* a, uint32_t, const T*, const Pre& p) const {
    T rv[6];
#pragma unroll
    for (int k = 0; k < 6; k++) rv[k] = p.b[k] - a[k];
    store6<T>(r + 6 * (size_t)n, rv);
    return 0.0;
  }
  __device__ void tail(uint32_t bk, uint32_t s) const {
    const double sp = 2.0 * bgeo[s].hd[0];
    const int lane = (int)(threadIdx.x & 7), group = (int)(threadIdx.x >> 3), groups = (int)(blockDim.x >> 3);
    const int a0 = brick_agg[bk], a1 = brick_agg[bk + 1];
    // every lane of a wave runs the same number of rounds, so the shuffles see whole groups
    const int rounds = (a1 - a0 + groups - 1) / groups;
    for (int q = 0; q < rounds; q++) {
      const int I = a0 + q * groups + group;
      T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
      if (I < a1) {
        double cI[3];
        for (int k = 0; k < 3; k++) cI[k] = centre[3 * (size_t)I + k];
        for (int i = agg_ptr[I] + lane; i < agg_ptr[I + 1]; i += 8) {
          const uint32_t v = (uint32_t)nodes[i];
          double p[3];
          node_pos(v, nkey, nkind, sp, p);
          for (int k = 0; k < 3; k++) p[k] -= cI[k];
          T tv[6];
          load6<T>(r + 6 * (size_t)v, tv);
          shiftT_add<T>(tv, p, acc);
        }
      }
#pragma unroll
      for (int m = 1; m < 8; m <<= 1)
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] += __shfl_xor(acc[k], m, 64);
      T o = acc[0];
#pragma unroll
      for (int k = 1; k < 6; k++) o = lane == k ? acc[k] : o;
      if (I < a1 && lane < 6) c[6 * (size_t)I + lane] = o;
    }
  }
};

// Symmetric 6x6 blocks (the inverse diagonal blocks D^-1) are stored as their upper
// triangle row by row: 21 values, (i, j) with i <= j at 6 i - i (i - 1) / 2 + j - i.
__host__ __device__ constexpr int sym_idx(int i, int j) {
  return i <= j ? 6 * i - i * (i - 1) / 2 + (j - i) : 6 * j - j * (j - 1) / 2 + (i - j);
}
constexpr int kSym = 21;
// A node's packed D^-1 (the 21 entries of the upper triangle): FP64 as 21 doubles; in
// the FP32 V-cycle as a 48-byte record of the entries in FP16 over the node's own scale
// (its largest entry maps to 2^14), read as three 16-byte loads instead of 21 4-byte
// ones (84 bytes).
struct alignas(16) SymHalf {
  _Float16 e[22];  // e[21] unused
  float scale;
};
static_assert(sizeof(SymHalf) == 48, "SymHalf is three 16-byte words");
template <class T>
__device__ __forceinline__ void load_sym(const T* __restrict__ dinv, size_t n, T* m) {
#pragma unroll
  for (int k = 0; k < kSym; k++) m[k] = dinv[kSym * n + k];
}
template <>
__device__ __forceinline__ void load_sym<float>(const float* __restrict__ dinv, size_t n, float* m) {
  typedef unsigned st_sym_u4 __attribute__((ext_vector_type(4)));
  typedef _Float16 st_sym_h2 __attribute__((ext_vector_type(2)));
  const st_sym_u4* q = reinterpret_cast<const st_sym_u4*>(reinterpret_cast<const SymHalf*>(dinv) + n);
  const st_sym_u4 a = q[0], b = q[1], c = q[2];
  const unsigned w[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
  const float sc = __builtin_bit_cast(float, w[11]);
#pragma unroll
  for (int k = 0; k < kSym; k++) {
    const st_sym_h2 h = __builtin_bit_cast(st_sym_h2, w[k >> 1]);
    m[k] = (float)((k & 1) ? h.y : h.x) * sc;
  }
}

// out = M v, M symmetric 6x6 given by its upper triangle a (registers)
template <class T>
__device__ __forceinline__ void sym6v_reg(const T* a, const T* v, T* out) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    T s = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) s += a[sym_idx(i, j)] * v[j];
    out[i] = s;
  }
}
// D^-1 of node n (packed, see SymHalf) times v
template <class T>
__device__ __forceinline__ void sym6v(const T* __restrict__ dinv, size_t n, const T* v, T* out) {
  T a[kSym];
  load_sym<T>(dinv, n, a);
  sym6v_reg<T>(a, v, out);
}

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

// z = x + step D^-1 (b - K x) (x the operator's argument); partial b . z, or with
// bdot set, the FP64 bdot . z
template <class T>
struct EpiSmooth : EpiBase {
  static constexpr bool kReduce = true;
  const T* b;
  const T* dinv;
  const double* step;  // per structure
  const double* bdot = nullptr;
  T* z;
  static constexpr bool kPre = true;
  struct Pre {
    T b[6], m[kSym];
  };
  __device__ Pre pre(uint32_t n) const {
    Pre p;
    load6<T>(b + 6 * (size_t)n, p.b);
    load_sym<T>(dinv, n, p.m);
    return p;
  }
  __device__ double node(uint32_t n, const T* a, uint32_t s, const T* xn) const { return node_pre(n, a, s, xn, pre(n)); }
  __device__ double node_pre(uint32_t n, const T* a, uint32_t s, const T* xn, const Pre& p) const {
    T xv[6], t[6], d[6];
    const T* bv = p.b;
    load6<T>(xn, xv);
#pragma unroll
    for (int k = 0; k < 6; k++) t[k] = bv[k] - a[k];
    sym6v_reg<T>(p.m, t, d);
    const T w = (T)step[s];
#pragma unroll
    for (int k = 0; k < 6; k++) xv[k] += w * d[k];
    double acc = 0;
    if (bdot) {
      double bd[6];
      load6<double>(bdot + 6 * (size_t)n, bd);
#pragma unroll
      for (int k = 0; k < 6; k++) acc += bd[k] * (double)xv[k];
    } else {
#pragma unroll
      for (int k = 0; k < 6; k++) acc += (double)bv[k] * (double)xv[k];
    }
    store6<T>(z + 6 * (size_t)n, xv);
    return acc;
  }
};

// sout = x + step D^-1 (b - K x) without the dot (the V-cycle's smoothed restriction
// forms its argument b - K sout from it)
template <class T>
struct EpiResidJac : EpiBase {
  static constexpr bool kReduce = false;
  const T* b;
  const T* dinv;
  const double* step;
  T* sout;
  static constexpr bool kPre = true;
  struct Pre {
    T b[6], m[kSym];
  };
  __device__ Pre pre(uint32_t n) const {
    Pre p;
    load6<T>(b + 6 * (size_t)n, p.b);
    load_sym<T>(dinv, n, p.m);
    return p;
  }
  __device__ double node(uint32_t n, const T* a, uint32_t s, const T* xn) const { return node_pre(n, a, s, xn, pre(n)); }
  __device__ double node_pre(uint32_t n, const T* a, uint32_t s, const T* xn, const Pre& p) const {
    T t[6], d[6], xv[6];
    load6<T>(xn, xv);
#pragma unroll
    for (int k = 0; k < 6; k++) t[k] = p.b[k] - a[k];
    sym6v_reg<T>(p.m, t, d);
    const T w = (T)step[s];
#pragma unroll
    for (int k = 0; k < 6; k++) xv[k] += w * d[k];
    store6<T>(sout + 6 * (size_t)n, xv);
    return 0.0;
  }
};

// x += u - step D^-1 K u, u the operator's argument (the rigid prolongation of the
// coarse correction): the smoothed prolongator, matrix-free
template <class T>
struct EpiProlong : EpiBase {
  static constexpr bool kReduce = false;
  const T* dinv;
  const double* step;
  T* x;
  static constexpr bool kPre = true;
  struct Pre {
    T x[6], m[kSym];
  };
  __device__ Pre pre(uint32_t n) const {
    Pre p;
    load6<T>(x + 6 * (size_t)n, p.x);
    load_sym<T>(dinv, n, p.m);
    return p;
  }
  __device__ double node(uint32_t n, const T* a, uint32_t s, const T* un) const { return node_pre(n, a, s, un, pre(n)); }
  __device__ double node_pre(uint32_t n, const T* a, uint32_t s, const T* un, const Pre& p) const {
    T uv[6], xv[6], d[6];
    load6<T>(un, uv);
#pragma unroll
    for (int k = 0; k < 6; k++) xv[k] = p.x[k];
    sym6v_reg<T>(p.m, a, d);
    const T w = (T)step[s];
#pragma unroll
    for (int k = 0; k < 6; k++) xv[k] += uv[k] - w * d[k];
    store6<T>(x + 6 * (size_t)n, xv);
    return 0.0;
  }
};

}  // namespace st
