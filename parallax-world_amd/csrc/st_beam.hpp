// One beam of a cell, in closed form: the reference's four bonds
// (statics/src/lattice.rs:76-84) at middle +- r e1, +- r e2 with the same
// spring S = k [g I + (e - g) n n^T] (statics/src/model.rs:139-149).
//
// Because the four offsets o_k sum to zero and sum o_k o_k^T = 2 r^2 (I - n n^T),
// the four bonds act on the beam's end nodes i, j exactly as
//   delta = (u_j - u_i) + (d/2) n x (w_i + w_j)            (relative motion of the middle)
//   F     = 4 S delta = K4 [g delta + (e - g)(n.delta) n]    (force on j; -F on i)
//   M     = R2 [e dw + (2g - e)(n.dw) n],  dw = w_j - w_i    (couple of the offsets)
//   (K x)_j = [F;  -(d/2) n x F + M],   (K x)_i = [-F; -(d/2) n x F - M]
// with K4 = 4k = A/d and R2 = 2 r^2 k = A^2 / (4 pi d), A the beam's section.
// This equals the sum over the four bonds of T^T S T (statics/src/bodies.rs:137-161)
// up to rounding, and does not depend on the choice of e1, e2.
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#define ST_HD __host__ __device__ __forceinline__
#else
#define ST_HD inline
#endif

namespace st {

constexpr double kSection = 0.336;  // statics/src/lattice.rs:25, both classes
constexpr double kPi = 3.14159265358979323846;
constexpr double kRs2 = 0.70710678118654752440;  // 1/sqrt(2)

// Face pairs (a < b) in VOXEL_FACES order -x,+x,-y,+y,-z,+z (lattice/src/lib.rs:64):
// the order the reference emits a cell's beams in (statics/src/lattice.rs:152-159).
constexpr int kPairA[15] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4};
constexpr int kPairB[15] = {1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};
// Unit outward normal of face f, component c.
ST_HD constexpr int face_normal(int f, int c) { return (f >> 1) == c ? ((f & 1) ? 1 : -1) : 0; }
// The beam's direction n = unit(N_b - N_a), times sqrt(2) for perpendicular faces
// (so components are -1, 0, 1), and whether the faces are opposite (class 0).
ST_HD constexpr bool pair_opposite(int q) { return kPairB[q] == (kPairA[q] ^ 1); }
ST_HD constexpr int pair_dir(int q, int c) { return face_normal(kPairB[q], c) - face_normal(kPairA[q], c) == 0 ? 0 : ((face_normal(kPairB[q], c) - face_normal(kPairA[q], c)) > 0 ? 1 : -1); }

// Per-class geometric constants at spacing s (all in SI units).
struct BeamGeo {
  double K4[2];  // A / d
  double R2[2];  // A^2 / (4 pi d)
  double hd[2];  // d / 2
  double nscale[2];  // 1 (class 0) or 1/sqrt(2) (class 1): n = nscale * pair_dir
};
inline BeamGeo beam_geo(double s) {
  BeamGeo g{};
  const double A = kSection * s * s;
  const double d[2] = {s, s * kRs2};
  for (int c = 0; c < 2; c++) {
    g.K4[c] = A / d[c];
    g.R2[c] = A * A / (4.0 * kPi * d[c]);
    g.hd[c] = 0.5 * d[c];
    g.nscale[c] = c == 0 ? 1.0 : kRs2;
  }
  return g;
}

// Per-cell coefficients for one class, for a beam direction n = ns N with N in {-1, 0, 1}^3
// (ns = 1 for opposite faces, 1/sqrt(2) for perpendicular ones):
//   delta = du + hdn (N x w),  F = kg delta + ken2 (N.delta) N,
//   M = re dw + rxn2 (N.dw) N,  lever -(d/2) n x F = hkn (N x delta),
// with hdn = hd ns, ken2 = K4 (e - g) ns^2, rxn2 = R2 (2g - e) ns^2, hkn = -hd ns K4 g.
template <class T>
struct CellCoef {
  T kg, ken2, re, rxn2, hdn, hkn;
};

template <class T>
ST_HD CellCoef<T> cell_coef(const BeamGeo& G, int cls, T e, T g) {
  const T ns2 = cls == 0 ? T(1) : T(0.5);
  CellCoef<T> c;
  c.kg = (T)G.K4[cls] * g;
  c.ken2 = (T)G.K4[cls] * (e - g) * ns2;
  c.re = (T)G.R2[cls] * e;
  c.rxn2 = (T)G.R2[cls] * (g + g - e) * ns2;
  c.hdn = (T)(G.hd[cls] * G.nscale[cls]);
  c.hkn = -c.hdn * c.kg;
  return c;
}

// A beam of its own section rho * SECTION s^2 (a gathered lattice's, statics/src/gather.rs:
// 386-390): the force terms scale with the section, the couple of the offsets with its
// square (K4 = A / d, R2 = A^2 / (4 pi d)).
template <class T>
ST_HD CellCoef<T> scaled(const CellCoef<T>& c, T rho) {
  CellCoef<T> o = c;
  o.kg = c.kg * rho;
  o.ken2 = c.ken2 * rho;
  o.hkn = c.hkn * rho;
  o.re = c.re * (rho * rho);
  o.rxn2 = c.rxn2 * (rho * rho);
  return o;
}

// A a + B b for compile-time A, B in {-1, 0, 1}, not both 0, with no operation on a zero
// coefficient (0 * x cannot be folded away in IEEE arithmetic, so it is never written).
template <int A, int B, class T>
ST_HD T lin2(T a, T b) {
  static_assert(A != 0 || B != 0, "lin2 of two zero coefficients");
  if constexpr (A == 0)
    return B > 0 ? b : -b;
  else if constexpr (B == 0)
    return A > 0 ? a : -a;
  else if constexpr (A > 0)
    return B > 0 ? a + b : a - b;
  else
    return B > 0 ? b - a : -(a + b);
}
template <int A, int B, int C, class T>
ST_HD T lin3(T a, T b, T c) {
  if constexpr (A == 0)
    return lin2<B, C>(b, c);
  else if constexpr (B == 0 && C == 0)
    return A > 0 ? a : -a;
  else {
    const T r = lin2<B, C == 0 ? 0 : C>(b, c);  // B or C non-zero here
    return A > 0 ? a + r : r - a;
  }
}
// (N x v)_K for N = (NX, NY, NZ): N_{K+1} v_{K+2} - N_{K+2} v_{K+1}; structurally zero when
// both of those components of N are.
template <int K, int NX, int NY, int NZ>
ST_HD constexpr bool cross_zero() {
  constexpr int N[3] = {NX, NY, NZ};
  return N[(K + 1) % 3] == 0 && N[(K + 2) % 3] == 0;
}
template <int K, int NX, int NY, int NZ, class T>
ST_HD T cross_n(const T* v) {
  constexpr int N[3] = {NX, NY, NZ};
  return lin2<N[(K + 1) % 3], -N[(K + 2) % 3]>(v[(K + 2) % 3], v[(K + 1) % 3]);
}
template <int K, int NX, int NY, int NZ>
ST_HD constexpr int ncomp() {
  constexpr int N[3] = {NX, NY, NZ};
  return N[K];
}

// The beam's end loads from its end motions xi, xj (6 each: u, w): the force F on end j
// (-F on end i) and the rotational rows Lj = L + M (end j) and Li = L - M (end i).
template <class T, int NX, int NY, int NZ>
struct BeamLoads {
  T F[3], Lj[3], Li[3];
};
template <int K, class T, int NX, int NY, int NZ>
ST_HD void beam_delta(const CellCoef<T>& c, const T* xi, const T* xj, const T* w, T* d) {
  const T du = xj[K] - xi[K];
  if constexpr (cross_zero<K, NX, NY, NZ>())
    d[K] = du;
  else
    d[K] = du + c.hdn * cross_n<K, NX, NY, NZ>(w);
}
template <int K, class T, int NX, int NY, int NZ>
ST_HD void beam_rows(const CellCoef<T>& c, const T* d, const T* q, T t, T tq, BeamLoads<T, NX, NY, NZ>& o) {
  constexpr int N = ncomp<K, NX, NY, NZ>();
  T f = c.kg * d[K], m = c.re * q[K];
  if constexpr (N > 0) {
    f = f + t;
    m = m + tq;
  } else if constexpr (N < 0) {
    f = f - t;
    m = m - tq;
  }
  o.F[K] = f;
  if constexpr (cross_zero<K, NX, NY, NZ>()) {
    o.Lj[K] = m;
    o.Li[K] = -m;
  } else {
    const T l = c.hkn * cross_n<K, NX, NY, NZ>(d);
    o.Lj[K] = l + m;
    o.Li[K] = l - m;
  }
}
template <class T, int NX, int NY, int NZ>
ST_HD BeamLoads<T, NX, NY, NZ> beam_loads(const CellCoef<T>& c, const T* xi, const T* xj) {
  T w[3], d[3], q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    w[k] = xi[3 + k] + xj[3 + k];  // (components the direction does not use are dropped)
    q[k] = xj[3 + k] - xi[3 + k];
  }
  beam_delta<0, T, NX, NY, NZ>(c, xi, xj, w, d);
  beam_delta<1, T, NX, NY, NZ>(c, xi, xj, w, d);
  beam_delta<2, T, NX, NY, NZ>(c, xi, xj, w, d);
  const T t = c.ken2 * lin3<NX, NY, NZ>(d[0], d[1], d[2]);
  const T tq = c.rxn2 * lin3<NX, NY, NZ>(q[0], q[1], q[2]);
  BeamLoads<T, NX, NY, NZ> o;
  beam_rows<0, T, NX, NY, NZ>(c, d, q, t, tq, o);
  beam_rows<1, T, NX, NY, NZ>(c, d, q, t, tq, o);
  beam_rows<2, T, NX, NY, NZ>(c, d, q, t, tq, o);
  return o;
}

// Applies beam q of a cell to end motions xi = x[a], xj = x[b] (6 each: u, w) and adds
// the beam's loads into yi, yj. N = (NX, NY, NZ) in {-1,0,1}^3, the direction over ns.
template <class T, int NX, int NY, int NZ>
ST_HD void beam_apply(const CellCoef<T>& c, const T* xi, const T* xj, T* yi, T* yj) {
  const BeamLoads<T, NX, NY, NZ> o = beam_loads<T, NX, NY, NZ>(c, xi, xj);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    yj[k] += o.F[k];
    yi[k] -= o.F[k];
    yj[3 + k] += o.Lj[k];
    yi[3 + k] += o.Li[k];
  }
}

// The same beam's load at one end only: end j (J) or end i, added into y.
template <class T, int NX, int NY, int NZ, bool J>
ST_HD void beam_end(const CellCoef<T>& c, const T* xi, const T* xj, T* y) {
  const BeamLoads<T, NX, NY, NZ> o = beam_loads<T, NX, NY, NZ>(c, xi, xj);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (J) {
      y[k] += o.F[k];
      y[3 + k] += o.Lj[k];
    } else {
      y[k] -= o.F[k];
      y[3 + k] += o.Li[k];
    }
  }
}

// The coefficients of beam Q of a cell: its class's, scaled by the beam's own section when
// the cells carry sections (kSec: rho holds the cell's 15 factors).
template <class T, bool kSec, int Q>
ST_HD CellCoef<T> beam_coef(const CellCoef<T>& c0, const CellCoef<T>& c1, const T* rho) {
  const CellCoef<T>& cc = pair_opposite(Q) ? c0 : c1;
  if constexpr (kSec)
    return scaled<T>(cc, rho[Q]);
  else
    return cc;
}

// Beam Q of a cell applied to its faces' motions x[f][6], its loads added into y[f][6]
// (FA, FB: the first load of face a or b in this cell, assigned rather than added onto a zero).
template <class T, int Q, bool FA, bool FB>
ST_HD void cell_beam(const CellCoef<T>& c, const T (*x)[6], T (*y)[6]) {
  constexpr int a = kPairA[Q], b = kPairB[Q];
  const BeamLoads<T, pair_dir(Q, 0), pair_dir(Q, 1), pair_dir(Q, 2)> o =
      beam_loads<T, pair_dir(Q, 0), pair_dir(Q, 1), pair_dir(Q, 2)>(c, x[a], x[b]);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if constexpr (FB) {
      y[b][k] = o.F[k];
      y[b][3 + k] = o.Lj[k];
    } else {
      y[b][k] += o.F[k];
      y[b][3 + k] += o.Lj[k];
    }
    if constexpr (FA) {
      y[a][k] = -o.F[k];
      y[a][3 + k] = o.Li[k];
    } else {
      y[a][k] -= o.F[k];
      y[a][3 + k] += o.Li[k];
    }
  }
}
// Every beam of a cell whose 15 beams are all alive: straight-line code, each face's first
// load assigned (face 0 and 1 first meet beam 0, face f >= 2 beam f - 1).
template <class T, bool kSec, int Q = 0>
ST_HD void cell_apply_full(const CellCoef<T>& c0, const CellCoef<T>& c1, const T* rho, const T (*x)[6], T (*y)[6]) {
  cell_beam<T, Q, Q == 0, Q <= 4>(beam_coef<T, kSec, Q>(c0, c1, rho), x, y);
  if constexpr (Q < 14) cell_apply_full<T, kSec, Q + 1>(c0, c1, rho, x, y);
}
// The live beams of a cell (mask bit q = beam q alive), onto y zeroed before.
template <class T, bool kSec, int Q = 0>
ST_HD void cell_apply_masked(const CellCoef<T>& c0, const CellCoef<T>& c1, uint32_t mask, const T* rho,
                             const T (*x)[6], T (*y)[6]) {
  if (mask & (1u << Q)) cell_beam<T, Q, false, false>(beam_coef<T, kSec, Q>(c0, c1, rho), x, y);
  if constexpr (Q < 14) cell_apply_masked<T, kSec, Q + 1>(c0, c1, mask, rho, x, y);
}

// All live beams of one cell: x[f][6] the motions of its faces (f in VOXEL_FACES
// order), y[f][6] receives each face's load (assigned). mask bit q = beam q alive.
// kSec: rho holds the cell's 15 section factors (else every beam is SECTION s^2).
template <class T, bool kSec>
ST_HD void cell_apply(const CellCoef<T>& c0, const CellCoef<T>& c1, uint32_t mask, const T (*x)[6], T (*y)[6],
                      const T* rho) {
  if (mask == 0x7fffu) {
    cell_apply_full<T, kSec>(c0, c1, rho, x, y);
  } else {
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
      for (int k = 0; k < 6; k++) y[f][k] = T(0);
    cell_apply_masked<T, kSec>(c0, c1, mask, rho, x, y);
  }
}

// The load at face F of one cell from its beams (F, f2), onto y: assigned by the first beam
// (First) or added.
template <class T, int F, int Q, bool First>
ST_HD void face_beam_end(const CellCoef<T>& c, const T (*x)[6], T* y) {
  constexpr bool J = kPairB[Q] == F;
  const BeamLoads<T, pair_dir(Q, 0), pair_dir(Q, 1), pair_dir(Q, 2)> o =
      beam_loads<T, pair_dir(Q, 0), pair_dir(Q, 1), pair_dir(Q, 2)>(c, x[kPairA[Q]], x[kPairB[Q]]);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const T f = J ? o.F[k] : -o.F[k];
    const T l = J ? o.Lj[k] : o.Li[k];
    if constexpr (First) {
      y[k] = f;
      y[3 + k] = l;
    } else {
      y[k] += f;
      y[3 + k] += l;
    }
  }
}
// Face F's five beams, all alive.
template <class T, int F, bool kSec, int Q = 0, bool First = true>
ST_HD void face_row_full(const CellCoef<T>& c0, const CellCoef<T>& c1, const T* rho, const T (*x)[6], T* y) {
  if constexpr (kPairA[Q] == F || kPairB[Q] == F) {
    face_beam_end<T, F, Q, First>(beam_coef<T, kSec, Q>(c0, c1, rho), x, y);
    if constexpr (Q < 14) face_row_full<T, F, kSec, Q + 1, false>(c0, c1, rho, x, y);
  } else if constexpr (Q < 14) {
    face_row_full<T, F, kSec, Q + 1, First>(c0, c1, rho, x, y);
  }
}
// Face F's live beams (mask holds only beams of face F), onto y zeroed before.
template <class T, int F, bool kSec, int Q = 0>
ST_HD void face_row_masked(const CellCoef<T>& c0, const CellCoef<T>& c1, uint32_t mask, const T* rho,
                           const T (*x)[6], T* y) {
  if constexpr (kPairA[Q] == F || kPairB[Q] == F) {
    if (mask & (1u << Q)) face_beam_end<T, F, Q, false>(beam_coef<T, kSec, Q>(c0, c1, rho), x, y);
  }
  if constexpr (Q < 14) face_row_masked<T, F, kSec, Q + 1>(c0, c1, mask, rho, x, y);
}

// Beam masks of the beams touching face f (5 of 15).
ST_HD constexpr uint32_t face_beams(int f) {
  uint32_t m = 0;
  for (int q = 0; q < 15; q++)
    if (kPairA[q] == f || kPairB[q] == f) m |= 1u << q;
  return m;
}

}  // namespace st
