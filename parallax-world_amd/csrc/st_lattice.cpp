// st_lattice_cells: a Lattice as the reference holds it (statics/src/lattice.rs:61-71,
// model.rs:92-111) back to the cells lattice::build made it from (lattice.rs:124-163):
// each beam's cell is the site its two end nodes share (the technique of
// statics/src/gather.rs:245), its pair q the two faces of that cell the nodes are, lesser
// first; a cell whose face node is held has the cell across that face held (lattice.rs:138).
//
// The device's operator takes each cell's beams between face centres with one e, g per
// cell, and per beam a section (as a factor on SECTION s^2) and strengths. So a lattice is
// taken when it is one build could have made at its level, up to per-beam strengths (what
// bodies::as_generated changes, statics/src/bodies.rs:611-625), or one gather::measured
// made (statics/src/gather.rs:386-390: beams of their own sections and strengths, one
// material to a cell): every beam four bonds of one area a in the order +r e1, -r e1,
// +r e2, -r e2 with r = sqrt(4 a / 2 pi) (lattice::beam, lattice.rs:76-84), the bonds of
// a cell of one e and g, the bonds of a beam of one strength.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "st_internal.hpp"

namespace st {
namespace {

constexpr double kUnit = 0.125;  // statics/src/lattice.rs:16

struct SiteKey {
  int64_t x, y, z;
  bool operator==(const SiteKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct SiteKeyHash {
  size_t operator()(const SiteKey& k) const {
    uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

bool close(double a, double b, double scale) { return std::fabs(a - b) <= 1e-9 * scale; }

[[noreturn]] void refuse(const std::string& why) { throw StError{ST_BAD_INPUT, "st_lattice_cells: " + why}; }

}  // namespace

void lattice_cells(const StLattice& L, StLatticeCells& out) {
  const int64_t nn = L.nnodes, nb = L.nbonds;
  if (nn < 0 || nb < 0 || (nb % 4) != 0) refuse("node or bond count (bonds come four to a beam)");
  if (nn && (!L.x || !L.held || !L.around)) refuse("node arrays missing");
  if (nb && (!L.ij || !L.p || !L.n || !L.a || !L.d || !L.bond_mat)) refuse("bond arrays missing");
  if (L.level > 20) refuse("level");
  const int lv = (int)L.level;
  const double s = kUnit * (double)(int64_t(1) << lv);
  // the nodes: orientation and the two sites (level units), lesser first
  std::vector<SiteKey> lo(nn), hi(nn);
  std::vector<int8_t> ori(nn);
  for (int64_t v = 0; v < nn; v++) {
    const int64_t* a = L.around + 6 * v;
    SiteKey A{a[0] >> lv, a[1] >> lv, a[2] >> lv}, B{a[3] >> lv, a[4] >> lv, a[5] >> lv};
    const int64_t dx = B.x - A.x, dy = B.y - A.y, dz = B.z - A.z;
    int o = -1;
    if (dx == 1 && dy == 0 && dz == 0) o = 0;
    if (dx == 0 && dy == 1 && dz == 0) o = 1;
    if (dx == 0 && dy == 0 && dz == 1) o = 2;
    if (o < 0) refuse("node " + std::to_string(v) + " does not lie between two face neighbours, lesser first");
    lo[v] = A;
    hi[v] = B;
    ori[v] = (int8_t)o;
  }
  // the cells that own beams, numbered as first met through the bonds
  std::unordered_map<SiteKey, int64_t, SiteKeyHash> cell_of;
  std::vector<SiteKey> sites;
  std::vector<uint16_t> beams;
  std::vector<double> e, g, rho, strength, section;  // strength: 45 per cell, section: 15
  std::vector<uint8_t> have_mat;
  std::vector<int64_t> bond_index((size_t)nb);
  const double full = 0.336 * s * s;  // SECTION s^2, statics/src/lattice.rs:21-25, both classes
  int64_t first_own = -1;             // the first bond of a beam of a section of its own
  for (int64_t b = 0; b < nb; b += 4) {
    const int64_t i = L.ij[2 * b], j = L.ij[2 * b + 1];
    if (i < 0 || j < 0 || i >= nn || j >= nn || i == j) refuse("bond " + std::to_string(b) + " names a missing node");
    for (int k = 1; k < 4; k++)
      if (L.ij[2 * (b + k)] != i || L.ij[2 * (b + k) + 1] != j)
        refuse("bonds " + std::to_string(b) + ".." + std::to_string(b + 3) + " are not one beam's four");
    // the site both ends lie on
    const SiteKey* ci[2] = {&lo[i], &hi[i]};
    const SiteKey* cj[2] = {&lo[j], &hi[j]};
    int si = -1, sj = -1;
    for (int u = 0; u < 2; u++)
      for (int w = 0; w < 2; w++)
        if (*ci[u] == *cj[w]) {
          if (si >= 0) refuse("beam at bond " + std::to_string(b) + ": its ends share two sites");
          si = u;
          sj = w;
        }
    if (si < 0) refuse("beam at bond " + std::to_string(b) + " joins nodes of no common cell (not a built lattice)");
    const SiteKey c = *ci[si];
    // the face of the cell each end is: the lesser site's + face, or the greater site's - face
    const int fi = 2 * ori[i] + (si == 0 ? 1 : 0), fj = 2 * ori[j] + (sj == 0 ? 1 : 0);
    if (fi == fj) refuse("beam at bond " + std::to_string(b) + ": its ends are one face");
    // the device's beams run from the lesser face a to the greater b (build's, lattice.rs:
    // 152-153); gather::measured's may run the other way (its handles' order, gather.rs:386-389)
    const bool flip = fi > fj;
    const int fa = flip ? fj : fi, fb = flip ? fi : fj;
    int q = -1;
    for (int t = 0; t < 15; t++)
      if (kPairA[t] == fa && kPairB[t] == fb) q = t;
    // the geometry of lattice::beam between the two face centres, from a to b
    const double* xi = L.x + 3 * (flip ? j : i);
    const double* xj = L.x + 3 * (flip ? i : j);
    double cc[3] = {((double)c.x + 0.5) * s, ((double)c.y + 0.5) * s, ((double)c.z + 0.5) * s};
    for (int t = 0; t < 3; t++) {
      const double want_i = cc[t] + 0.5 * s * face_normal(fa, t), want_j = cc[t] + 0.5 * s * face_normal(fb, t);
      if (!close(xi[t], want_i, s) || !close(xj[t], want_j, s))
        refuse("beam at bond " + std::to_string(b) + ": an end is not its face's centre");
    }
    double nv[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    const double d = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (double& t : nv) t /= d;
    double ref[3] = {0, 0, 0};
    ref[std::fabs(nv[0]) < 0.9 ? 0 : 1] = 1.0;
    double e1[3] = {nv[1] * ref[2] - nv[2] * ref[1], nv[2] * ref[0] - nv[0] * ref[2], nv[0] * ref[1] - nv[1] * ref[0]};
    const double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (double& t : e1) t /= l1;
    const double e2[3] = {nv[1] * e1[2] - nv[2] * e1[1], nv[2] * e1[0] - nv[0] * e1[2], nv[0] * e1[1] - nv[1] * e1[0]};
    // the beam's section: four times its bonds' area
    const double area = 4.0 * L.a[b];
    if (!(area > 0) || !std::isfinite(area)) refuse("beam at bond " + std::to_string(b) + ": a bond area is not positive");
    if (first_own < 0 && !close(area, full, full)) first_own = b;
    const double r = std::sqrt(area / (2.0 * 3.14159265358979323846));
    // run from b to a, the beam's n and e1 are negated, e2 kept: its first two bonds are the
    // device's in the other order
    const double sn = flip ? -1.0 : 1.0;
    for (int kl = 0; kl < 4; kl++) {
      const int64_t bk = b + kl;
      const int k = flip && kl < 2 ? kl ^ 1 : kl;  // the device's place of the bond
      if (L.a[bk] != L.a[b]) refuse("beam at bond " + std::to_string(b) + ": its bonds differ in area");
      if (!close(L.d[bk], d, s)) refuse("beam at bond " + std::to_string(b) + ": d is not its length");
      const double* ek = k < 2 ? e1 : e2;
      const double sg = (k & 1) ? -r : r;
      for (int t = 0; t < 3; t++) {
        if (!close(L.n[3 * bk + t], sn * nv[t], 1.0))
          refuse("beam at bond " + std::to_string(b) + ": n is not its direction");
        if (!close(L.p[3 * bk + t], 0.5 * (xi[t] + xj[t]) + sg * ek[t], s))
          refuse("beam at bond " + std::to_string(b) + ": the bonds are not at middle +- r e1, +- r e2 in order");
      }
      const double* m = L.bond_mat + 6 * bk;
      const double* m0 = L.bond_mat + 6 * b;
      if (m[3] != m0[3] || m[4] != m0[4] || m[5] != m0[5] || m[1] != m0[1] || m[2] != m0[2] || m[0] != m0[0])
        refuse("beam at bond " + std::to_string(b) + ": its bonds differ in material");
    }
    auto it = cell_of.find(c);
    int64_t ci_idx;
    if (it == cell_of.end()) {
      ci_idx = (int64_t)sites.size();
      cell_of.emplace(c, ci_idx);
      sites.push_back(c);
      beams.push_back(0);
      e.push_back(0.0);
      g.push_back(0.0);
      rho.push_back(0.0);
      strength.resize(strength.size() + 45, 0.0);
      section.resize(section.size() + 15, 1.0);
      have_mat.push_back(0);
    } else {
      ci_idx = it->second;
    }
    const double* m = L.bond_mat + 6 * b;
    if (!have_mat[ci_idx]) {
      e[ci_idx] = m[1];
      g[ci_idx] = m[2];
      rho[ci_idx] = m[0];
      have_mat[ci_idx] = 1;
    } else if (e[ci_idx] != m[1] || g[ci_idx] != m[2]) {
      refuse("the beams of one cell differ in stiffness (cell of bond " + std::to_string(b) + ")");
    }
    if (beams[ci_idx] & (1u << q)) refuse("beam at bond " + std::to_string(b) + " appears twice");
    beams[ci_idx] |= (uint16_t)(1u << q);
    for (int t = 0; t < 3; t++) strength[45 * ci_idx + 3 * q + t] = m[3 + t];
    section[15 * ci_idx + q] = area / full;
    for (int kl = 0; kl < 4; kl++) {
      const int64_t at = 60 * ci_idx + 4 * q + (flip && kl < 2 ? kl ^ 1 : kl);
      bond_index[b + kl] = flip ? -1 - at : at;
    }
  }
  const int64_t nfree = (int64_t)sites.size();
  // held cells: across each held node from a cell with beams (a held node lies between a
  // free cell and a held one, lattice.rs:138)
  std::vector<uint8_t> held(nfree, 0);
  for (int64_t v = 0; v < nn; v++) {
    const bool lo_free = cell_of.count(lo[v]) != 0, hi_free = cell_of.count(hi[v]) != 0;
    if (L.held[v]) {
      if (lo_free && hi_free) refuse("node " + std::to_string(v) + " is held but both its cells have beams");
      if (!lo_free && !hi_free) continue;  // a face no beam reaches: nothing to solve there
      const SiteKey hc = lo_free ? hi[v] : lo[v];
      if (!cell_of.count(hc)) {
        cell_of.emplace(hc, (int64_t)sites.size());
        sites.push_back(hc);
        held.push_back(1);
      }
    }
  }
  // every free node lies on a cell with beams and, across, on a free cell or nothing
  for (int64_t v = 0; v < nn; v++) {
    if (L.held[v]) continue;
    const auto a = cell_of.find(lo[v]), b = cell_of.find(hi[v]);
    const bool af = a != cell_of.end() && !held[a->second], bf = b != cell_of.end() && !held[b->second];
    if (!af && !bf) refuse("free node " + std::to_string(v) + " is on no cell with beams");
    if ((a != cell_of.end() && held[a->second]) || (b != cell_of.end() && held[b->second]))
      refuse("free node " + std::to_string(v) + " lies on a held cell");
  }
  const int64_t ncell = (int64_t)sites.size();
  out.n_cells = ncell;
  out.spacing = s;
  if (!out.sites) return;  // the sizing call
  if (first_own >= 0 && !out.section)
    refuse("beam at bond " + std::to_string(first_own) +
           " has a section other than SECTION s^2 (a gathered lattice: give out->section for the sections)");
  for (int64_t k = 0; k < ncell; k++) {
    const SiteKey& c = sites[k];
    if (c.x < INT32_MIN || c.x > INT32_MAX || c.y < INT32_MIN || c.y > INT32_MAX || c.z < INT32_MIN || c.z > INT32_MAX)
      refuse("a site beyond int32");
    out.sites[3 * k] = (int32_t)c.x;
    out.sites[3 * k + 1] = (int32_t)c.y;
    out.sites[3 * k + 2] = (int32_t)c.z;
    const bool f = k < nfree;
    if (out.held) out.held[k] = held[k];
    if (out.e) out.e[k] = f ? e[k] : 0.0;
    if (out.g) out.g[k] = f ? g[k] : 0.0;
    if (out.rho) out.rho[k] = f ? rho[k] : 0.0;
    if (out.beams) out.beams[k] = f ? beams[k] : 0;
    if (out.strength)
      for (int t = 0; t < 45; t++) out.strength[45 * k + t] = f ? strength[45 * k + t] : 0.0;
    if (out.section)
      for (int t = 0; t < 15; t++) out.section[15 * k + t] = f ? section[15 * k + t] : 1.0;
  }
  if (out.bond_index) std::memcpy(out.bond_index, bond_index.data(), (size_t)nb * sizeof(int64_t));
}

}  // namespace st

extern "C" int st_lattice_cells(const StLattice* lat, StLatticeCells* out, char* message, int64_t message_len) {
  try {
    if (!lat || !out) throw st::StError{ST_BAD_INPUT, "st_lattice_cells: lat or out is NULL"};
    st::lattice_cells(*lat, *out);
    if (message && message_len > 0) message[0] = 0;
    return ST_OK;
  } catch (const st::StError& e) {
    if (message && message_len > 0) std::snprintf(message, (size_t)message_len, "%s", e.what.c_str());
    return e.code;
  } catch (const std::bad_alloc&) {
    if (message && message_len > 0) std::snprintf(message, (size_t)message_len, "host allocation failed");
    return ST_OOM;
  }
}
