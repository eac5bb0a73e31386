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

// Per-cell coefficients for one class: F = kg*delta + ke*(n.delta) n; M = re*dw + rx*(n.dw) n;
// the lever term -(d/2) n x F = -hk * (n x delta) with hk = hd*kg.
template <class T>
struct CellCoef {
  T kg, ke, re, rx, hk, ns, hd;
};

template <class T>
ST_HD CellCoef<T> cell_coef(const BeamGeo& G, int cls, T e, T g) {
  CellCoef<T> c;
  c.kg = (T)G.K4[cls] * g;
  c.ke = (T)G.K4[cls] * (e - g);
  c.re = (T)G.R2[cls] * e;
  c.rx = (T)G.R2[cls] * (g + g - e);
  c.hk = (T)G.hd[cls] * c.kg;
  c.ns = (T)G.nscale[cls];
  c.hd = (T)G.hd[cls];
  return c;
}

// A beam of its own section rho * SECTION s^2 (a gathered lattice's, statics/src/gather.rs:
// 386-390): the force terms scale with the section, the couple of the offsets with its
// square (K4 = A / d, R2 = A^2 / (4 pi d)).
template <class T>
ST_HD CellCoef<T> scaled(const CellCoef<T>& c, T rho) {
  CellCoef<T> o = c;
  o.kg = c.kg * rho;
  o.ke = c.ke * rho;
  o.hk = c.hk * rho;
  o.re = c.re * (rho * rho);
  o.rx = c.rx * (rho * rho);
  return o;
}

// Applies beam q of a cell to end motions xi = x[a], xj = x[b] (6 each: u, w) and adds
// the beam's loads into yi, yj. n components nx, ny, nz in {-1,0,1} times c.ns.
template <class T, int NX, int NY, int NZ>
ST_HD void beam_apply(const CellCoef<T>& c, const T* xi, const T* xj, T* yi, T* yj) {
  const T ns = c.ns;
  const T n0 = (T)NX * ns, n1 = (T)NY * ns, n2 = (T)NZ * ns;
  // w = w_i + w_j; delta = du + hd * n x w
  const T w0 = xi[3] + xj[3], w1 = xi[4] + xj[4], w2 = xi[5] + xj[5];
  const T cx0 = n1 * w2 - n2 * w1, cx1 = n2 * w0 - n0 * w2, cx2 = n0 * w1 - n1 * w0;
  const T d0 = (xj[0] - xi[0]) + c.hd * cx0;
  const T d1 = (xj[1] - xi[1]) + c.hd * cx1;
  const T d2 = (xj[2] - xi[2]) + c.hd * cx2;
  const T nd = n0 * d0 + n1 * d1 + n2 * d2;
  const T F0 = c.kg * d0 + c.ke * nd * n0;
  const T F1 = c.kg * d1 + c.ke * nd * n1;
  const T F2 = c.kg * d2 + c.ke * nd * n2;
  const T q0 = xj[3] - xi[3], q1 = xj[4] - xi[4], q2 = xj[5] - xi[5];
  const T nq = n0 * q0 + n1 * q1 + n2 * q2;
  const T M0 = c.re * q0 + c.rx * nq * n0;
  const T M1 = c.re * q1 + c.rx * nq * n1;
  const T M2 = c.re * q2 + c.rx * nq * n2;
  // lever: -(d/2) n x F = -hk (n x delta)
  const T L0 = -c.hk * (n1 * d2 - n2 * d1);
  const T L1 = -c.hk * (n2 * d0 - n0 * d2);
  const T L2 = -c.hk * (n0 * d1 - n1 * d0);
  yj[0] += F0; yj[1] += F1; yj[2] += F2;
  yi[0] -= F0; yi[1] -= F1; yi[2] -= F2;
  yj[3] += L0 + M0; yj[4] += L1 + M1; yj[5] += L2 + M2;
  yi[3] += L0 - M0; yi[4] += L1 - M1; yi[5] += L2 - M2;
}

// The same beam's load at one end only: end j (J) or end i, added into y.
template <class T, int NX, int NY, int NZ, bool J>
ST_HD void beam_end(const CellCoef<T>& c, const T* xi, const T* xj, T* y) {
  const T ns = c.ns;
  const T n0 = (T)NX * ns, n1 = (T)NY * ns, n2 = (T)NZ * ns;
  const T w0 = xi[3] + xj[3], w1 = xi[4] + xj[4], w2 = xi[5] + xj[5];
  const T cx0 = n1 * w2 - n2 * w1, cx1 = n2 * w0 - n0 * w2, cx2 = n0 * w1 - n1 * w0;
  const T d0 = (xj[0] - xi[0]) + c.hd * cx0;
  const T d1 = (xj[1] - xi[1]) + c.hd * cx1;
  const T d2 = (xj[2] - xi[2]) + c.hd * cx2;
  const T nd = n0 * d0 + n1 * d1 + n2 * d2;
  const T F0 = c.kg * d0 + c.ke * nd * n0;
  const T F1 = c.kg * d1 + c.ke * nd * n1;
  const T F2 = c.kg * d2 + c.ke * nd * n2;
  const T q0 = xj[3] - xi[3], q1 = xj[4] - xi[4], q2 = xj[5] - xi[5];
  const T nq = n0 * q0 + n1 * q1 + n2 * q2;
  const T M0 = c.re * q0 + c.rx * nq * n0;
  const T M1 = c.re * q1 + c.rx * nq * n1;
  const T M2 = c.re * q2 + c.rx * nq * n2;
  const T L0 = -c.hk * (n1 * d2 - n2 * d1);
  const T L1 = -c.hk * (n2 * d0 - n0 * d2);
  const T L2 = -c.hk * (n0 * d1 - n1 * d0);
  if (J) {
    y[0] += F0; y[1] += F1; y[2] += F2;
    y[3] += L0 + M0; y[4] += L1 + M1; y[5] += L2 + M2;
  } else {
    y[0] -= F0; y[1] -= F1; y[2] -= F2;
    y[3] += L0 - M0; y[4] += L1 - M1; y[5] += L2 - M2;
  }
}

// All live beams of one cell: x[f][6] the motions of its faces (f in VOXEL_FACES
// order), y[f][6] accumulates. mask bit q = beam q alive. only_face >= 0 restricts
// to the beams touching that face (the others' loads are still added, but the
// caller keeps only y[only_face]).
// rho: the cell's 15 section factors, or NULL (every beam SECTION s^2).
template <class T>
ST_HD void cell_apply(const CellCoef<T>& c0, const CellCoef<T>& c1, uint32_t mask, const T (*x)[6], T (*y)[6],
                      const T* rho = nullptr) {
#define ST_BEAM(q, a, b, cc, NX, NY, NZ) \
  if (mask & (1u << q))                  \
    beam_apply<T, NX, NY, NZ>(rho ? scaled<T>(cc, rho[q]) : cc, x[a], x[b], y[a], y[b]);
  ST_BEAM(0, 0, 1, c0, 1, 0, 0)
  ST_BEAM(1, 0, 2, c1, 1, -1, 0)
  ST_BEAM(2, 0, 3, c1, 1, 1, 0)
  ST_BEAM(3, 0, 4, c1, 1, 0, -1)
  ST_BEAM(4, 0, 5, c1, 1, 0, 1)
  ST_BEAM(5, 1, 2, c1, -1, -1, 0)
  ST_BEAM(6, 1, 3, c1, -1, 1, 0)
  ST_BEAM(7, 1, 4, c1, -1, 0, -1)
  ST_BEAM(8, 1, 5, c1, -1, 0, 1)
  ST_BEAM(9, 2, 3, c0, 0, 1, 0)
  ST_BEAM(10, 2, 4, c1, 0, 1, -1)
  ST_BEAM(11, 2, 5, c1, 0, 1, 1)
  ST_BEAM(12, 3, 4, c1, 0, -1, -1)
  ST_BEAM(13, 3, 5, c1, 0, -1, 1)
  ST_BEAM(14, 4, 5, c0, 0, 0, 1)
#undef ST_BEAM
}

// Beam masks of the beams touching face f (5 of 15).
ST_HD constexpr uint32_t face_beams(int f) {
  uint32_t m = 0;
  for (int q = 0; q < 15; q++)
    if (kPairA[q] == f || kPairB[q] == f) m |= 1u << q;
  return m;
}

}  // namespace st
