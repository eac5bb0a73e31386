// Entry points of the C ABI for the solve, its batches, classification, bond
// forces and the gradients.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>

#include "st_solver.hpp"

// The layouts include/statics_hip.h states for bindings in other languages.
static_assert(sizeof(StSolveOpts) == 32 && offsetof(StSolveOpts, tol) == 8 && offsetof(StSolveOpts, max_iter) == 16 &&
                  offsetof(StSolveOpts, refresh) == 20 && offsetof(StSolveOpts, warm_start) == 24 &&
                  offsetof(StSolveOpts, release_hierarchy) == 28,
              "StSolveOpts layout");
static_assert(sizeof(StStats) == 88 && offsetof(StStats, rel_residual) == 16 && offsetof(StStats, setup_ms) == 24 &&
                  offsetof(StStats, solve_ms) == 32 && offsetof(StStats, nodes) == 40 &&
                  offsetof(StStats, dof_nodes) == 48 && offsetof(StStats, cells) == 56 &&
                  offsetof(StStats, bricks) == 64 && offsetof(StStats, retries) == 72 &&
                  offsetof(StStats, first_status) == 76 && offsetof(StStats, mem_peak) == 80,
              "StStats layout");

namespace st {
namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

// e1, e2 of pair q as in lattice.rs:78-80, from the unit direction n.
__device__ __forceinline__ void pair_frame(const double* n, double* e1, double* e2) {
  double ref[3] = {0, 0, 0};
  if (fabs(n[0]) < 0.9)
    ref[0] = 1.0;
  else
    ref[1] = 1.0;
  double c[3] = {n[1] * ref[2] - n[2] * ref[1], n[2] * ref[0] - n[0] * ref[2], n[0] * ref[1] - n[1] * ref[0]};
  const double inv = 1.0 / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int k = 0; k < 3; k++) e1[k] = c[k] * inv;
  e2[0] = n[1] * e1[2] - n[2] * e1[1];
  e2[1] = n[2] * e1[0] - n[0] * e1[2];
  e2[2] = n[0] * e1[1] - n[1] * e1[0];
}

struct BeamState {
  double n[3], hd, K4, R2, k;  // k = a/d per bond
  double r;                    // offset radius sqrt(A / 2 pi)
};
// Beam q of a cell whose section is rho SECTION s^2 (rho 1: lattice::build's).
__device__ __forceinline__ void beam_state(const BeamGeo& G, int q, double rho, BeamState& b) {
  const int cls = pair_opposite(q) ? 0 : 1;
  for (int k = 0; k < 3; k++) b.n[k] = G.nscale[cls] * (double)pair_dir(q, k);
  b.hd = G.hd[cls];
  b.K4 = G.K4[cls] * rho;
  b.R2 = G.R2[cls] * (rho * rho);
  b.k = 0.25 * b.K4;
  // A = K4 * d ; r = sqrt(A / (2 pi))
  b.r = sqrt(b.K4 * 2.0 * G.hd[cls] / (2.0 * kPi));
}
// Beam q of cell c: its section over SECTION s^2 (1 when the cells gave none).
__device__ __forceinline__ double beam_sec(const double* csec, int64_t c, int q) {
  return csec ? csec[15 * c + q] : 1.0;
}

// middle relative motion delta and dw of a beam between end motions xi, xj
__device__ __forceinline__ void beam_delta(const BeamState& b, const double* xi, const double* xj, double* d,
                                           double* dw) {
  const double w[3] = {xi[3] + xj[3], xi[4] + xj[4], xi[5] + xj[5]};
  d[0] = xj[0] - xi[0] + b.hd * (b.n[1] * w[2] - b.n[2] * w[1]);
  d[1] = xj[1] - xi[1] + b.hd * (b.n[2] * w[0] - b.n[0] * w[2]);
  d[2] = xj[2] - xi[2] + b.hd * (b.n[0] * w[1] - b.n[1] * w[0]);
  for (int k = 0; k < 3; k++) dw[k] = xj[3 + k] - xi[3 + k];
}

__device__ __forceinline__ void load_faces(const uint32_t* cfaces, int64_t c, const double* x, double (*xv)[6]) {
  for (int f = 0; f < 6; f++) {
    const uint32_t v = cfaces[6 * c + f];
    for (int k = 0; k < 6; k++) xv[f][k] = (v & kNotFree) ? 0.0 : x[6 * (size_t)v + k];
  }
}

// The forces on the j side of beam q's four bonds (statics/src/bodies.rs:420-423 with
// model.rs:115-119), f[k] in the reference's bond order: offsets +r e1, -r e1, +r e2, -r e2
// (lattice.rs:76-84).
__device__ __forceinline__ void beam_bond_forces(const BeamState& b, double e, double g, const double* xi,
                                                 const double* xj, double (*f)[3]) {
  double d[3], dw[3], e1[3], e2[3];
  beam_delta(b, xi, xj, d, dw);
  pair_frame(b.n, e1, e2);
  for (int k = 0; k < 4; k++) {
    const double* ek = k < 2 ? e1 : e2;
    const double sg = (k & 1) ? -b.r : b.r;
    const double off[3] = {sg * ek[0], sg * ek[1], sg * ek[2]};
    const double dl[3] = {d[0] + dw[1] * off[2] - dw[2] * off[1], d[1] + dw[2] * off[0] - dw[0] * off[2],
                          d[2] + dw[0] * off[1] - dw[1] * off[0]};
    const double along = dl[0] * b.n[0] + dl[1] * b.n[1] + dl[2] * b.n[2];
    for (int t = 0; t < 3; t++) f[k][t] = b.n[t] * (e * b.k * along) + (dl[t] - b.n[t] * along) * (g * b.k);
  }
}

// Bond::used (statics/src/model.rs:123-129) with Mat::used (model.rs:66-68): the share of
// the material's strengths (ft, fc, fs) that force f uses in a bond of area a along n.
__device__ __forceinline__ double bond_used(const double* f, const double* n, double a, double ft, double fc,
                                            double fs) {
  const double ia = 1.0 / a;
  const double t[3] = {f[0] * ia, f[1] * ia, f[2] * ia};
  const double sigma = t[0] * n[0] + t[1] * n[1] + t[2] * n[2];
  const double tau2 = fmax(t[0] * t[0] + t[1] * t[1] + t[2] * t[2] - sigma * sigma, 0.0);
  const double r = sqrt(sigma * sigma / 4.0 + tau2);
  const double s1 = sigma / 2.0 + r, s3 = sigma / 2.0 - r;
  return fmax(fmax(s1 / ft, -s3 / fc), (s1 - s3) / (2.0 * fs));
}

// A bond's area: a quarter of its beam's section rho SECTION s^2 (lattice.rs:21-25, 81-83).
__device__ __forceinline__ double bond_area(const BeamGeo& G, double rho) {
  const double s = 2.0 * G.hd[0];
  return kSection * s * s / 4.0 * rho;
}

// Per bond force on the j side: out[c][q][k][3], zero for dead beams.
__global__ void k_bond_forces(int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask, const double* ce,
                              const double* cg, const uint32_t* cstruct, const BeamGeo* bgeo, const double* csec, const double* x,
                              double* out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double xv[6][6];
  load_faces(cfaces, c, x, xv);
  const BeamGeo G = bgeo[cstruct[c]];
  const double e = ce[c], g = cg[c];
  const uint32_t m = cmask[c];
  for (int q = 0; q < 15; q++) {
    double* o = out + ((size_t)c * 15 + q) * 12;
    if (!(m & (1u << q))) {
      for (int i = 0; i < 12; i++) o[i] = 0.0;
      continue;
    }
    BeamState b;
    beam_state(G, q, beam_sec(csec, c, q), b);
    double f[4][3];
    beam_bond_forces(b, e, g, xv[kPairA[q]], xv[kPairB[q]], f);
    for (int k = 0; k < 4; k++)
      for (int t = 0; t < 3; t++) o[3 * k + t] = f[k][t];
  }
}

// Per bond used share (bodies.rs:544-546): out[c][q][k], zero for dead beams.
// strength: ft, fc, fs per beam, 45 per cell (beam q at 3 q).
__global__ void k_bond_used(int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask, const double* ce,
                            const double* cg, const uint32_t* cstruct, const BeamGeo* bgeo, const double* csec, const double* strength,
                            const double* x, double* out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double xv[6][6];
  load_faces(cfaces, c, x, xv);
  const BeamGeo G = bgeo[cstruct[c]];
  const double e = ce[c], g = cg[c];
  const uint32_t m = cmask[c];
  for (int q = 0; q < 15; q++) {
    const double* sq = strength + 45 * c + 3 * q;
    const double ft = sq[0], fc = sq[1], fs = sq[2];
    double* o = out + ((size_t)c * 15 + q) * 4;
    if (!(m & (1u << q))) {
      for (int k = 0; k < 4; k++) o[k] = 0.0;
      continue;
    }
    BeamState b;
    beam_state(G, q, beam_sec(csec, c, q), b);
    double f[4][3];
    beam_bond_forces(b, e, g, xv[kPairA[q]], xv[kPairB[q]], f);
    for (int k = 0; k < 4; k++) o[k] = bond_used(f[k], b.n, bond_area(G, beam_sec(csec, c, q)), ft, fc, fs);
  }
}

// One round's breaking (statics/src/bodies.rs:523-533): a live beam (alive, neither end
// gone) one of whose four bonds uses more than its strength (used > 1) dies whole; its
// bit is cleared in the caller's mask. Counts the beams broken.
__global__ void k_break(int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask, uint16_t* cmask_user,
                        const double* ce, const double* cg, const uint32_t* cstruct, const BeamGeo* bgeo, const double* csec,
                        const double* strength, const double* x, int* broken) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const uint32_t m = cmask[c];
  if (!m) return;
  double xv[6][6];
  load_faces(cfaces, c, x, xv);
  const BeamGeo G = bgeo[cstruct[c]];
  const double e = ce[c], g = cg[c];
  uint32_t dead = 0;
  for (int q = 0; q < 15; q++) {
    if (!(m & (1u << q))) continue;
    const double* sq = strength + 45 * c + 3 * q;
    const double ft = sq[0], fc = sq[1], fs = sq[2];
    BeamState b;
    beam_state(G, q, beam_sec(csec, c, q), b);
    double f[4][3];
    beam_bond_forces(b, e, g, xv[kPairA[q]], xv[kPairB[q]], f);
    bool any = false;
    const double a = bond_area(G, beam_sec(csec, c, q));
    for (int k = 0; k < 4; k++) any |= bond_used(f[k], b.n, a, ft, fc, fs) > 1.0;
    if (any) dead |= 1u << q;
  }
  if (dead) {
    cmask_user[c] = (uint16_t)(cmask_user[c] & ~dead);
    atomicAdd(broken, __popc(dead));
  }
}

// sigma of bond k of beam q: (e/d)(n.delta + dw.(o_k x n))
__device__ __forceinline__ double bond_sigma(const BeamState& b, double e, const double* d, const double* dw,
                                             const double* ox) {
  return (e / (2.0 * b.hd)) * (b.n[0] * d[0] + b.n[1] * d[1] + b.n[2] * d[2] + dw[0] * ox[0] + dw[1] * ox[1] +
                               dw[2] * ox[2]);
}
__device__ __forceinline__ void bond_oxn(const BeamState& b, int k, double* ox) {
  double e1[3], e2[3];
  pair_frame(b.n, e1, e2);
  const double* ek = k < 2 ? e1 : e2;
  const double sg = (k & 1) ? -b.r : b.r;
  const double o[3] = {sg * ek[0], sg * ek[1], sg * ek[2]};
  ox[0] = o[1] * b.n[2] - o[2] * b.n[1];
  ox[1] = o[2] * b.n[0] - o[0] * b.n[2];
  ox[2] = o[0] * b.n[1] - o[1] * b.n[0];
}

// pass 0: per-cell max sigma; pass 1: per-cell sum exp(beta (sigma - m))
__global__ void k_softmax_stats(int pass, int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask,
                                const double* ce, const uint32_t* cstruct, const BeamGeo* bgeo, const double* csec, const double* x,
                                double beta, const double* mref, double* out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double xv[6][6];
  load_faces(cfaces, c, x, xv);
  const BeamGeo G = bgeo[cstruct[c]];
  const double e = ce[c];
  const uint32_t m = cmask[c];
  double acc = pass == 0 ? -INFINITY : 0.0;
  for (int q = 0; q < 15; q++) {
    if (!(m & (1u << q))) continue;
    if (cfaces[6 * c + kPairA[q]] & cfaces[6 * c + kPairB[q]] & kNotFree) continue;  // held-held: no such beam
    BeamState b;
    beam_state(G, q, beam_sec(csec, c, q), b);
    double d[3], dw[3];
    beam_delta(b, xv[kPairA[q]], xv[kPairB[q]], d, dw);
    for (int k = 0; k < 4; k++) {
      double ox[3];
      bond_oxn(b, k, ox);
      const double s = bond_sigma(b, e, d, dw, ox);
      if (pass == 0)
        acc = fmax(acc, s);
      else
        acc += exp(beta * (s - mref[0]));
    }
  }
  out[c] = acc;
}

// d J / d x for the soft-max: per bond weight w = exp(beta(sigma - m)) / Z, scattered to the
// bond's end nodes (global FP64 atomics; run once per backward).
__global__ void k_softmax_rhs(int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask, const double* ce,
                              const uint32_t* cstruct, const BeamGeo* bgeo, const double* csec, const double* x, double beta,
                              const double* mz, double* g) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double xv[6][6];
  load_faces(cfaces, c, x, xv);
  const BeamGeo G = bgeo[cstruct[c]];
  const double e = ce[c];
  const uint32_t m = cmask[c];
  const double mm = mz[0], Z = mz[1];
  // the cell's contributions summed per face in registers, then added into g once per
  // face and component (36 atomics a cell instead of 12 a beam)
  double acc[6][6];
#pragma unroll
  for (int f = 0; f < 6; f++)
#pragma unroll
    for (int t = 0; t < 6; t++) acc[f][t] = 0.0;
#pragma unroll
  for (int q = 0; q < 15; q++) {
    if (!(m & (1u << q))) continue;
    if (cfaces[6 * c + kPairA[q]] & cfaces[6 * c + kPairB[q]] & kNotFree) continue;  // held-held: no such beam
    BeamState b;
    beam_state(G, q, beam_sec(csec, c, q), b);
    double d[3], dw[3];
    beam_delta(b, xv[kPairA[q]], xv[kPairB[q]], d, dw);
    double gu[3] = {0, 0, 0}, gw[3] = {0, 0, 0};
    for (int k = 0; k < 4; k++) {
      double ox[3];
      bond_oxn(b, k, ox);
      const double w = exp(beta * (bond_sigma(b, e, d, dw, ox) - mm)) / Z;
      const double sc = w * e / (2.0 * b.hd);
      for (int t = 0; t < 3; t++) {
        gu[t] += sc * b.n[t];
        gw[t] += sc * ox[t];
      }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
      acc[kPairB[q]][t] += gu[t];
      acc[kPairB[q]][3 + t] += gw[t];
      acc[kPairA[q]][t] -= gu[t];
      acc[kPairA[q]][3 + t] -= gw[t];
    }
  }
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const uint32_t v = cfaces[6 * c + f];
    if (v & kNotFree) continue;
#pragma unroll
    for (int t = 0; t < 6; t++) atomicAdd(&g[6 * (size_t)v + t], acc[f][t]);
  }
}

// Per-cell sensitivity: -lambda_c^T K_c x_c (+ the soft-max's explicit sum w sigma when beta > 0).
__global__ void k_sensitivity(int64_t ncells, const uint32_t* cfaces, const uint16_t* cmask, const double* ce,
                              const double* cg, const uint32_t* cstruct, const BeamGeo* bgeo, const double* csec, const double* x,
                              const double* lam, int softmax, double beta, const double* mz, double* out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double xv[6][6], lv[6][6];
  load_faces(cfaces, c, x, xv);
  load_faces(cfaces, c, lam, lv);
  const BeamGeo G = bgeo[cstruct[c]];
  const double e = ce[c], g = cg[c];
  const uint32_t m = cmask[c];
  double acc = 0.0;
  for (int q = 0; q < 15; q++) {
    if (!(m & (1u << q))) continue;
    if (cfaces[6 * c + kPairA[q]] & cfaces[6 * c + kPairB[q]] & kNotFree) continue;  // held-held: no such beam
    BeamState b;
    beam_state(G, q, beam_sec(csec, c, q), b);
    double d[3], dw[3], dl[3], dwl[3];
    beam_delta(b, xv[kPairA[q]], xv[kPairB[q]], d, dw);
    beam_delta(b, lv[kPairA[q]], lv[kPairB[q]], dl, dwl);
    const double nd = b.n[0] * d[0] + b.n[1] * d[1] + b.n[2] * d[2];
    const double nq = b.n[0] * dw[0] + b.n[1] * dw[1] + b.n[2] * dw[2];
    double F[3], M[3];
    for (int t = 0; t < 3; t++) {
      F[t] = b.K4 * (g * d[t] + (e - g) * nd * b.n[t]);
      M[t] = b.R2 * (e * dw[t] + (g + g - e) * nq * b.n[t]);
    }
    acc -= dl[0] * F[0] + dl[1] * F[1] + dl[2] * F[2] + dwl[0] * M[0] + dwl[1] * M[1] + dwl[2] * M[2];
    if (softmax)
      for (int k = 0; k < 4; k++) {
        double ox[3];
        bond_oxn(b, k, ox);
        const double s = bond_sigma(b, e, d, dw, ox);
        acc += exp(beta * (s - mz[0])) / mz[1] * s;
      }
  }
  out[c] = acc;
}

__global__ void k_gravity(int64_t nnodes, const int32_t* ncell, const uint8_t* nflags, double* f) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nnodes) return;
  const int cnt = (ncell[2 * v] >= 0) + (ncell[2 * v + 1] >= 0);
  for (int k = 0; k < 6; k++) f[6 * v + k] = 0.0;
  if (!(nflags[v] & (kHeld | kGone))) f[6 * v + 2] = -(double)cnt / 6.0;
}

__global__ void k_mask_rows2(int64_t n, const uint8_t* nflags, double* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (nflags[i] & (kHeld | kGone))
    for (int k = 0; k < 6; k++) v[6 * i + k] = 0.0;
}

void check_opts(const StSolveOpts* o) {
  if (!o) throw StError{ST_BAD_INPUT, "opts is NULL"};
  if (o->precision != ST_FP64 && o->precision != ST_MIXED && o->precision != ST_MIXED32)
    throw StError{ST_BAD_INPUT, "opts.precision"};
  if (!(o->tol >= 0) || o->max_iter < 0) throw StError{ST_BAD_INPUT, "opts: tol or max_iter"};
}

int worst(const std::vector<StStats>& s) {
  for (auto& t : s)
    if (t.status != ST_OK) return t.status;
  return ST_OK;
}

// Host reduction of per-cell partials (fixed order).
double host_sum(const std::vector<double>& v) {
  double s = 0;
  for (double a : v) s += a;
  return s;
}

void set_loads(StCtx* c, const double* loads) {
  const int64_t nn = c->geo.nnodes;
  const size_t bytes = (size_t)std::max<int64_t>(nn, 1) * 6 * sizeof(double);
  c->fres.ensure(bytes);
  // kept as given: each solve masks the rows of the nodes held or gone at that time
  if (nn) ST_HIP(hipMemcpyAsync(c->fres.p, loads, nn * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  c->loads_set = true;
}

// Solves with the resident loads; the solution stays in c->xsol.
std::vector<StStats> solve_resident(StCtx* c, const StSolveOpts& o) {
  if (!c->loads_set) throw StError{ST_BAD_INPUT, "solve: no loads set"};
  const int64_t nn = c->geo.nnodes;
  const size_t bytes = (size_t)std::max<int64_t>(nn, 1) * 6 * sizeof(double);
  c->xsol.ensure(bytes);
  c->x64.ensure(bytes);
  c->f64.ensure(bytes);
  if (nn) {
    // the loads of the nodes free now (the held and gone ones may have changed since they were set)
    ST_HIP(hipMemcpyAsync(c->f64.p, c->fres.p, nn * 6 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    k_mask_rows2<<<nblk(nn), 256, 0, c->stream>>>(nn, c->geo.nflags.as<uint8_t>(), c->f64.as<double>());
    if (o.warm_start && c->have_solution) {
      ST_HIP(hipMemcpyAsync(c->x64.p, c->xsol.p, nn * 6 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      k_mask_rows2<<<nblk(nn), 256, 0, c->stream>>>(nn, c->geo.nflags.as<uint8_t>(), c->x64.as<double>());
    } else {
      ST_HIP(hipMemsetAsync(c->x64.p, 0, nn * 6 * sizeof(double), c->stream));
    }
  }
  std::vector<StStats> out;
  solve_all(*c, o, out);
  std::swap(c->x64, c->xsol);
  c->have_solution = true;
  c->last_opts = o;
  return out;
}

// ft, fc, fs of every beam of every device cell (45 per cell) from the caller's: 3 per
// input cell (per_beam = 0, the cell's material) or 45 (per_beam != 0: beam q's at 3 q,
// as st_lattice_cells gives them for lattices whose beams are of their own strengths).
// Held cells have no beams and are skipped; so are beams a cell does not hold.
void upload_strength(StCtx* c, const double* strength, int per_beam, DBuf& out) {
  const auto& coi = c->host.cell_of_input;
  const int64_t ncl = c->geo.ncells;
  const int stride = per_beam ? 45 : 3;
  std::vector<double> h((size_t)std::max<int64_t>(ncl, 1) * 45, 1.0);
  for (size_t i = 0; i < coi.size(); i++) {
    const int64_t k = coi[i];
    if (k < 0) continue;
    const uint32_t held_beams = c->host.cmask[k];
    for (int q = 0; q < 15; q++)
      for (int t = 0; t < 3; t++) {
        const double v = strength[stride * i + (per_beam ? 3 * q : 0) + t];
        if (!(v > 0) || !std::isfinite(v)) {
          if (per_beam && !(held_beams & (1u << q))) continue;  // a beam the cell does not hold
          throw StError{ST_BAD_INPUT, "strength: a free cell has a non-positive or non-finite ft, fc or fs"};
        }
        h[45 * (size_t)k + 3 * q + t] = v;
      }
  }
  out.ensure(h.size() * sizeof(double));
  ST_HIP(hipMemcpyAsync(out.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  ST_HIP(hipStreamSynchronize(c->stream));
}
}  // namespace
}  // namespace st

using namespace st;

extern "C" {


int st_set_loads(StCtx* c, const double* loads) {
  return guard(c, [&] {
    need_upload(c);
    if (!loads) throw StError{ST_BAD_INPUT, "loads is NULL"};
    ST_HIP(hipSetDevice(c->device));
    set_loads(c, loads);
    ST_HIP(hipStreamSynchronize(c->stream));
  });
}

int st_solve_resident(StCtx* c, const StSolveOpts* opts, StStats* stats) {
  std::vector<StStats> st_out;
  int rc = guard(c, [&] {
    need_upload(c);
    check_opts(opts);
    ST_HIP(hipSetDevice(c->device));
    st_out = solve_resident(c, *opts);
    if (stats) *stats = st_out.empty() ? StStats{} : st_out[0];
  });
  if (rc == ST_OK && !st_out.empty()) rc = st_out[0].status;
  return rc;
}

int st_get_solution(StCtx* c, double* x_out) {
  return guard(c, [&] {
    need_upload(c);
    if (!x_out) throw StError{ST_BAD_INPUT, "x_out is NULL"};
    if (!c->have_solution) throw StError{ST_BAD_INPUT, "no solution"};
    const int64_t nn = c->geo.nnodes;
    if (nn) ST_HIP(hipMemcpyAsync(x_out, c->xsol.p, nn * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
  });
}

int st_classify(StCtx* c, int64_t* joined_count) {
  return guard(c, [&] {
    need_upload(c);
    ST_HIP(hipSetDevice(c->device));
    classify_device(*c, joined_count, nullptr);
  });
}

int st_solve(StCtx* c, const double* loads, const StSolveOpts* opts, double* x_out, StStats* stats) {
  std::vector<StStats> st_out;
  int rc = guard(c, [&] {
    need_upload(c);
    check_opts(opts);
    if (!loads || !x_out) throw StError{ST_BAD_INPUT, "solve: loads or x_out is NULL"};
    ST_HIP(hipSetDevice(c->device));
    const int64_t nn = c->geo.nnodes;
    set_loads(c, loads);
    if (opts->warm_start) {
      c->xsol.ensure((size_t)std::max<int64_t>(nn, 1) * 6 * sizeof(double));
      if (nn) ST_HIP(hipMemcpyAsync(c->xsol.p, x_out, nn * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
      c->have_solution = true;
    }
    st_out = solve_resident(c, *opts);
    if (nn) ST_HIP(hipMemcpyAsync(x_out, c->xsol.p, nn * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
    if (stats) *stats = st_out.empty() ? StStats{} : st_out[0];
  });
  if (rc == ST_OK && !st_out.empty()) rc = st_out[0].status;
  return rc;
}

int st_solve_batch(StCtx* c, int32_t n, const StCells* const* cells, const double* const* loads,
                   const StSolveOpts* opts, double* const* x_out, StStats* stats) {
  std::vector<StStats> st_out;
  int rc = guard(c, [&] {
    check_opts(opts);
    if (n <= 0 || !cells) throw StError{ST_BAD_INPUT, "batch: n <= 0 or cells is NULL"};
    for (int i = 0; i < n; i++)
      if (!cells[i]) throw StError{ST_BAD_INPUT, "batch: a cells entry is NULL"};
    ST_HIP(hipSetDevice(c->device));
    build_host_geo(c->host, cells, n);
    copy_geo_to_device(*c);
    c->gone_host.assign(c->geo.nnodes, 0);
    refresh_cell_flags(*c);
    c->uploaded = true;
    // as forces does (statics/src/bodies.rs:407-409 under settle's 519-522): what is not
    // joined to a held node is gone before the solve
    classify_device(*c, nullptr, nullptr);
    const int64_t nn = c->geo.nnodes;
    const size_t bytes = (size_t)std::max<int64_t>(nn, 1) * 6 * sizeof(double);
    c->f64.ensure(bytes);
    c->x64.ensure(bytes);
    if (nn) {
      k_gravity<<<nblk(nn), 256, 0, c->stream>>>(nn, c->geo.ncell.as<int32_t>(), c->geo.nflags.as<uint8_t>(),
                                                 c->f64.as<double>());
      for (int i = 0; i < n; i++)
        if (loads && loads[i]) {
          const StructInfo& si = c->host.structs[i];
          ST_HIP(hipMemcpyAsync(c->f64.as<double>() + 6 * si.node0, loads[i], si.nnodes * 6 * sizeof(double),
                                hipMemcpyHostToDevice, c->stream));
        }
      k_mask_rows2<<<nblk(nn), 256, 0, c->stream>>>(nn, c->geo.nflags.as<uint8_t>(), c->f64.as<double>());
      ST_HIP(hipMemsetAsync(c->x64.p, 0, nn * 6 * sizeof(double), c->stream));
    }
    solve_all(*c, *opts, st_out);
    for (int i = 0; i < n; i++)
      if (x_out && x_out[i]) {
        const StructInfo& si = c->host.structs[i];
        ST_HIP(hipMemcpyAsync(x_out[i], c->x64.as<double>() + 6 * si.node0, si.nnodes * 6 * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
      }
    ST_HIP(hipStreamSynchronize(c->stream));
    c->have_solution = false;
    if (stats)
      for (int i = 0; i < n; i++) stats[i] = st_out[i];
  });
  if (rc == ST_OK) rc = worst(st_out);
  return rc;
}

int st_joined_to_held(StCtx* c, const uint16_t* beams, uint8_t* joined_out, int keep_gone) {
  return guard(c, [&] {
    need_upload(c);
    ST_HIP(hipSetDevice(c->device));
    if (beams) set_beam_masks(*c, beams);
    if (keep_gone) {
      std::vector<uint8_t> j;
      joined_to_held(*c, j);
      if (joined_out) std::memcpy(joined_out, j.data(), j.size());
    } else {
      classify_device(*c, nullptr, joined_out);
    }
  });
}

int st_bond_forces(StCtx* c, const double* x, double* out) {
  return guard(c, [&] {
    need_upload(c);
    if (!x || !out) throw StError{ST_BAD_INPUT, "bond forces: x or out is NULL"};
    ST_HIP(hipSetDevice(c->device));
    const int64_t nn = c->geo.nnodes, ncl = c->geo.ncells;
    DBuf xd, od;
    xd.ensure(std::max<int64_t>(nn, 1) * 6 * sizeof(double));
    od.ensure(std::max<int64_t>(ncl, 1) * 180 * sizeof(double));
    if (nn) ST_HIP(hipMemcpyAsync(xd.p, x, nn * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (ncl)
      k_bond_forces<<<nblk(ncl, 128), 128, 0, c->stream>>>(ncl, c->geo.cfaces.as<uint32_t>(), c->geo.cmask.as<uint16_t>(),
                                                      c->geo.ce64.as<double>(), c->geo.cg64.as<double>(),
                                                      c->geo.cstruct.as<uint32_t>(), c->geo.bgeo.as<BeamGeo>(), c->geo.csec64.as<double>(),
                                                      xd.as<double>(), od.as<double>());
    ST_HIP(hipGetLastError());
    std::vector<double> h((size_t)ncl * 180);
    if (ncl) ST_HIP(hipMemcpyAsync(h.data(), od.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
    xd.release();
    od.release();
    // upload order: one 180-block per input cell; held cells get zeros
    const auto& coi = c->host.cell_of_input;
    for (size_t i = 0; i < coi.size(); i++) {
      if (coi[i] < 0)
        std::memset(out + 180 * i, 0, 180 * sizeof(double));
      else
        std::memcpy(out + 180 * i, h.data() + 180 * (size_t)coi[i], 180 * sizeof(double));
    }
  });
}


int st_bond_used(StCtx* c, const double* strength, int per_beam, const double* x, double* out) {
  return guard(c, [&] {
    need_upload(c);
    if (!strength || !x || !out) throw StError{ST_BAD_INPUT, "bond used: strength, x or out is NULL"};
    ST_HIP(hipSetDevice(c->device));
    const int64_t nn = c->geo.nnodes, ncl = c->geo.ncells;
    DBuf sd, xd, od;
    upload_strength(c, strength, per_beam, sd);
    xd.ensure(std::max<int64_t>(nn, 1) * 6 * sizeof(double));
    od.ensure(std::max<int64_t>(ncl, 1) * 60 * sizeof(double));
    if (nn) ST_HIP(hipMemcpyAsync(xd.p, x, nn * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (ncl)
      k_bond_used<<<nblk(ncl, 128), 128, 0, c->stream>>>(ncl, c->geo.cfaces.as<uint32_t>(), c->geo.cmask.as<uint16_t>(),
                                                    c->geo.ce64.as<double>(), c->geo.cg64.as<double>(),
                                                    c->geo.cstruct.as<uint32_t>(), c->geo.bgeo.as<BeamGeo>(), c->geo.csec64.as<double>(),
                                                    sd.as<double>(), xd.as<double>(), od.as<double>());
    ST_HIP(hipGetLastError());
    std::vector<double> h((size_t)ncl * 60);
    if (ncl) ST_HIP(hipMemcpyAsync(h.data(), od.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
    const auto& coi = c->host.cell_of_input;
    for (size_t i = 0; i < coi.size(); i++) {
      if (coi[i] < 0)
        std::memset(out + 60 * i, 0, 60 * sizeof(double));
      else
        std::memcpy(out + 60 * i, h.data() + 60 * (size_t)coi[i], 60 * sizeof(double));
    }
  });
}

int st_settle(StCtx* c, const double* strength, int per_beam, const double* loads, const StSolveOpts* opts, uint8_t* gone_out,
              uint16_t* beams_out, int32_t* rounds_out, StStats* stats) {
  std::vector<StStats> st_out;
  int rc = guard(c, [&] {
    need_upload(c);
    check_opts(opts);
    if (!strength || !loads) throw StError{ST_BAD_INPUT, "settle: strength or loads is NULL"};
    ST_HIP(hipSetDevice(c->device));
    DevGeo& g = c->geo;
    const int64_t nn = g.nnodes, ncl = g.ncells;
    DBuf sd, broken;
    upload_strength(c, strength, per_beam, sd);
    set_loads(c, loads);
    broken.ensure(sizeof(int));
    StSolveOpts o = *opts;
    StStats total{};
    int rounds = 0;
    for (;;) {
      rounds++;
      // bodies.rs:518-519: what no chain of alive beams joins to a held node is gone
      classify_device(*c, nullptr, nullptr);
      // bodies.rs:520: forces under the loads (later rounds may start from the last round's x)
      o.warm_start = rounds > 1 ? opts->warm_start : 0;
      st_out = solve_resident(c, o);
      const int bad = worst(st_out);
      if (bad != ST_OK) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "settle: the solve of round %d failed after %d iterations (residual %.3g)",
                      rounds, st_out[0].iters, st_out[0].rel_residual);
        throw StError{bad, msg};
      }
      total.iters += st_out[0].iters;
      total.setup_ms += st_out[0].setup_ms;
      total.solve_ms += st_out[0].solve_ms;
      total.retries += st_out[0].retries;
      total.mem_peak = std::max(total.mem_peak, st_out[0].mem_peak);
      // under the reference's rule the recursive residual decides: a round whose true
      // residual is above tol is flagged (first_status) with the worst such residual
      if (o.refresh < 0 && !(st_out[0].rel_residual <= o.tol * (1.0 + 1e-9))) {
        total.first_status = ST_NO_CONVERGENCE;
        total.rel_residual = std::max(total.rel_residual, st_out[0].rel_residual);
      }
      const bool verbose = std::getenv("STATICS_HIP_VERBOSE") != nullptr;
      // bodies.rs:523-533: beams with an overused bond break
      ST_HIP(hipMemsetAsync(broken.p, 0, sizeof(int), c->stream));
      if (ncl)
        k_break<<<nblk(ncl, 128), 128, 0, c->stream>>>(ncl, g.cfaces.as<uint32_t>(), g.cmask.as<uint16_t>(),
                                                      g.cmask_user.as<uint16_t>(), g.ce64.as<double>(),
                                                      g.cg64.as<double>(), g.cstruct.as<uint32_t>(),
                                                      g.bgeo.as<BeamGeo>(), g.csec64.as<double>(), sd.as<double>(), c->xsol.as<double>(),
                                                      broken.as<int>());
      ST_HIP(hipGetLastError());
      int nb = 0;
      ST_HIP(hipMemcpyAsync(&nb, broken.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      ST_HIP(hipStreamSynchronize(c->stream));
      if (verbose)
        std::fprintf(stderr, "statics_hip: settle round %d: %d iterations, residual %.3g, %d beams broken\n", rounds,
                     st_out[0].iters, st_out[0].rel_residual, nb);
      if (!nb) break;
      refresh_cell_flags(*c);  // the masks changed: the kept hierarchy no longer fits
    }
    // the host copy of the masks follows the device's
    HostGeo& h = c->host;
    if (!h.cmask.empty())
      ST_HIP(hipMemcpyAsync(h.cmask.data(), g.cmask_user.p, h.cmask.size() * sizeof(uint16_t), hipMemcpyDeviceToHost,
                            c->stream));
    std::vector<uint8_t> fl(nn);
    if (nn) ST_HIP(hipMemcpyAsync(fl.data(), g.nflags.p, nn, hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
    if (gone_out)
      for (int64_t v = 0; v < nn; v++) gone_out[v] = (fl[v] & kGone) ? 1 : 0;
    if (beams_out) {
      const auto& coi = h.cell_of_input;
      for (size_t i = 0; i < coi.size(); i++) beams_out[i] = coi[i] < 0 ? 0 : h.cmask[coi[i]];
    }
    if (rounds_out) *rounds_out = rounds;
    if (stats) {
      *stats = st_out[0];
      stats->iters = total.iters;
      stats->setup_ms = total.setup_ms;
      stats->solve_ms = total.solve_ms;
      stats->retries = total.retries;
      stats->first_status = total.first_status;
      stats->mem_peak = total.mem_peak;
      if (total.first_status != ST_OK) stats->rel_residual = total.rel_residual;
    }
  });
  return rc;
}

int st_backward(StCtx* c, int objective, double beta, const StSolveOpts* opts, double* J_out, double* dJ,
                StStats* stats) {
  std::vector<StStats> st_out;
  const bool value_only = (objective & ST_VALUE_ONLY) != 0;
  objective &= ~ST_VALUE_ONLY;
  int rc = guard(c, [&] {
    need_upload(c);
    check_opts(opts);
    if (!c->have_solution) throw StError{ST_BAD_INPUT, "backward: no solve to differentiate (call st_solve first)"};
    if (objective != ST_COMPLIANCE && objective != ST_SOFTMAX_STRESS) throw StError{ST_BAD_INPUT, "objective"};
    if (objective == ST_SOFTMAX_STRESS && !(beta > 0)) throw StError{ST_BAD_INPUT, "beta must be > 0"};
    ST_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    DevGeo& g = c->geo;
    const int64_t nn = g.nnodes, ncl = g.ncells;
    const size_t bytes = (size_t)std::max<int64_t>(nn, 1) * 6 * sizeof(double);
    DBuf xd, lam, cellv, mz;
    xd.ensure(bytes);
    lam.ensure(bytes);
    cellv.ensure(std::max<int64_t>(ncl, 1) * sizeof(double));
    mz.ensure(2 * sizeof(double));
    if (nn) ST_HIP(hipMemcpyAsync(xd.p, c->xsol.p, nn * 6 * sizeof(double), hipMemcpyDeviceToDevice, s));
    double J = 0.0, mzh[2] = {0.0, 1.0};
    std::vector<double> hc(ncl);
    const bool soft = objective == ST_SOFTMAX_STRESS;
    if (!soft) {
      // compliance f.x; self-adjoint: lambda = x
      J = device_dot_f_x(*c, c->fres.as<double>(), xd.as<double>());
      if (nn) ST_HIP(hipMemcpyAsync(lam.p, xd.p, nn * 6 * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
      // J = m + log(sum exp(beta (sigma - m))) / beta, m the greatest sigma
      if (ncl) {
        k_softmax_stats<<<nblk(ncl, 128), 128, 0, s>>>(0, ncl, g.cfaces.as<uint32_t>(), g.cmask.as<uint16_t>(),
                                                  g.ce64.as<double>(), g.cstruct.as<uint32_t>(), g.bgeo.as<BeamGeo>(), g.csec64.as<double>(),
                                                  xd.as<double>(), beta, nullptr, cellv.as<double>());
        ST_HIP(hipMemcpyAsync(hc.data(), cellv.p, ncl * sizeof(double), hipMemcpyDeviceToHost, s));
        ST_HIP(hipStreamSynchronize(s));
      }
      double m = -INFINITY;
      for (double v : hc) m = std::max(m, v);
      if (!std::isfinite(m)) throw StError{ST_BAD_INPUT, "backward: no live bond"};
      mzh[0] = m;
      ST_HIP(hipMemcpyAsync(mz.p, mzh, 2 * sizeof(double), hipMemcpyHostToDevice, s));
      k_softmax_stats<<<nblk(ncl, 128), 128, 0, s>>>(1, ncl, g.cfaces.as<uint32_t>(), g.cmask.as<uint16_t>(),
                                                g.ce64.as<double>(), g.cstruct.as<uint32_t>(), g.bgeo.as<BeamGeo>(), g.csec64.as<double>(),
                                                xd.as<double>(), beta, mz.as<double>(), cellv.as<double>());
      ST_HIP(hipMemcpyAsync(hc.data(), cellv.p, ncl * sizeof(double), hipMemcpyDeviceToHost, s));
      ST_HIP(hipStreamSynchronize(s));
      mzh[1] = host_sum(hc);
      J = m + std::log(mzh[1]) / beta;
      if (value_only) {  // the objective only: no adjoint solve
        xd.release();
        lam.release();
        cellv.release();
        mz.release();
        if (J_out) *J_out = J;
        return;
      }
      ST_HIP(hipMemcpyAsync(mz.p, mzh, 2 * sizeof(double), hipMemcpyHostToDevice, s));
      // adjoint: K lambda = dJ/dx, solved with the same operator (and the forward solve's
      // hierarchy, when its precision mode is the same)
      c->f64.ensure(bytes);
      c->x64.ensure(bytes);
      ST_HIP(hipMemsetAsync(c->f64.p, 0, bytes, s));
      k_softmax_rhs<<<nblk(ncl, 128), 128, 0, s>>>(ncl, g.cfaces.as<uint32_t>(), g.cmask.as<uint16_t>(), g.ce64.as<double>(),
                                              g.cstruct.as<uint32_t>(), g.bgeo.as<BeamGeo>(), g.csec64.as<double>(), xd.as<double>(), beta,
                                              mz.as<double>(), c->f64.as<double>());
      k_mask_rows2<<<nblk(nn), 256, 0, s>>>(nn, g.nflags.as<uint8_t>(), c->f64.as<double>());
      ST_HIP(hipMemsetAsync(c->x64.p, 0, bytes, s));
      solve_all(*c, *opts, st_out);
      ST_HIP(hipMemcpyAsync(lam.p, c->x64.p, bytes, hipMemcpyDeviceToDevice, s));
    }
    if (ncl && !value_only)
      k_sensitivity<<<nblk(ncl, 128), 128, 0, s>>>(ncl, g.cfaces.as<uint32_t>(), g.cmask.as<uint16_t>(), g.ce64.as<double>(),
                                              g.cg64.as<double>(), g.cstruct.as<uint32_t>(), g.bgeo.as<BeamGeo>(), g.csec64.as<double>(),
                                              xd.as<double>(), lam.as<double>(), soft ? 1 : 0, beta, mz.as<double>(),
                                              cellv.as<double>());
    ST_HIP(hipGetLastError());
    if (ncl && dJ && !value_only)
      ST_HIP(hipMemcpyAsync(hc.data(), cellv.p, ncl * sizeof(double), hipMemcpyDeviceToHost, s));
    ST_HIP(hipStreamSynchronize(s));
    xd.release();
    lam.release();
    cellv.release();
    mz.release();
    if (J_out) *J_out = J;
    if (dJ && !value_only) {
      const auto& coi = c->host.cell_of_input;
      for (size_t i = 0; i < coi.size(); i++) dJ[i] = coi[i] < 0 ? 0.0 : hc[coi[i]];
    }
    if (stats) *stats = st_out.empty() ? StStats{} : st_out[0];
  });
  if (rc == ST_OK && !st_out.empty()) rc = st_out[0].status;
  return rc;
}

}  // extern "C"
