// Context, upload and the operator entry points of the C ABI.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

#include "st_kernels.hpp"
#include "st_solver.hpp"

namespace st {

namespace {

template <class V>
void up(DBuf& d, const std::vector<V>& h, hipStream_t s) {
  d.ensure(std::max<size_t>(h.size(), 1) * sizeof(V));
  if (!h.empty()) ST_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, s));
}

// Marks faces that are not free (held or gone) in each cell's face list, and
// drops beams touching a gone node (statics/src/bodies.rs:434).
__global__ void k_cell_flags(int64_t ncells, uint32_t* cfaces, const uint8_t* nflags, const uint16_t* mask_user,
                             uint16_t* mask) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  uint32_t m = mask_user[c];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    uint32_t v = cfaces[6 * c + f] & ~kNotFree;
    uint8_t fl = nflags[v];
    cfaces[6 * c + f] = v | ((fl & (kHeld | kGone)) ? kNotFree : 0u);
    if (fl & kGone) m &= ~face_beams(f);
  }
  mask[c] = (uint16_t)m;
}

__global__ void k_to_float(int64_t n, const double* a, float* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

}  // namespace

void copy_geo_to_device(Ctx& c) {
  HostGeo& h = c.host;
  DevGeo& d = c.geo;
  hipStream_t s = c.stream;
  d.nbricks = (int64_t)h.cbase.size() - 1;
  d.ncells = (int64_t)h.ce.size();
  d.nnodes = (int64_t)h.nflags.size();
  d.nstruct = (int32_t)h.structs.size();
  c.tuned = false;
  c.hier_valid = false;
  c.coarse_layouts.clear();
  {
    // bricks in two classes by node count (Morton order kept inside each)
    constexpr uint32_t kSmall = 640;
    std::vector<uint32_t> small, big;
    d.class_nodes[0] = d.class_nodes[1] = 0;
    for (size_t b = 0; b + 1 < h.nbase.size(); b++) {
      const uint32_t nn = h.nbase[b + 1] - h.nbase[b];
      const int k = nn <= kSmall ? 0 : 1;
      (k ? big : small).push_back((uint32_t)b);
      d.class_nodes[k] = std::max(d.class_nodes[k], nn);
    }
    d.nclass[0] = (int64_t)small.size();
    d.nclass[1] = (int64_t)big.size();
    small.insert(small.end(), big.begin(), big.end());
    up(d.blist, small, s);
  }
  up(d.cbase, h.cbase, s);
  up(d.nbase, h.nbase, s);
  up(d.bstruct, h.bstruct, s);
  up(d.ce64, h.ce, s);
  up(d.cg64, h.cg, s);
  up(d.cmask_user, h.cmask, s);
  up(d.cmask, h.cmask, s);
  up(d.cfaces, h.cfaces, s);
  up(d.ccoord, h.ccoord, s);
  up(d.cstruct, h.cstruct, s);
  up(d.nflags, h.nflags, s);
  up(d.ncell, h.ncell, s);
  up(d.nkind, h.nkind, s);
  up(d.nkey, h.nkey, s);
  std::vector<BeamGeo> bg;
  for (auto& si : h.structs) bg.push_back(beam_geo(si.spacing));
  up(d.bgeo, bg, s);
  d.ce32.ensure(std::max<int64_t>(d.ncells, 1) * sizeof(float));
  d.cg32.ensure(std::max<int64_t>(d.ncells, 1) * sizeof(float));
  if (d.ncells) {
    const int nb = (int)((d.ncells + 255) / 256);
    k_to_float<<<nb, 256, 0, s>>>(d.ncells, d.ce64.as<double>(), d.ce32.as<float>());
    k_to_float<<<nb, 256, 0, s>>>(d.ncells, d.cg64.as<double>(), d.cg32.as<float>());
    ST_HIP(hipGetLastError());
  }
  d.has_sec = !h.csec.empty();
  if (d.has_sec) {
    up(d.csec64, h.csec, s);
    d.csec32.ensure(h.csec.size() * sizeof(float));
    k_to_float<<<(unsigned)((h.csec.size() + 255) / 256), 256, 0, s>>>((int64_t)h.csec.size(),
                                                                       d.csec64.as<double>(), d.csec32.as<float>());
    ST_HIP(hipGetLastError());
  } else {
    d.csec64.release();
    d.csec32.release();
  }
}

// Applies the node flags (held, gone) to the cells' face lists and live masks.
void refresh_cell_flags(Ctx& c) {
  c.hier_valid = false;  // the operator's live beams may have changed
  DevGeo& d = c.geo;
  if (!d.ncells) return;
  k_cell_flags<<<(unsigned)((d.ncells + 255) / 256), 256, 0, c.stream>>>(d.ncells, d.cfaces.as<uint32_t>(), d.nflags.as<uint8_t>(),
                                                                   d.cmask_user.as<uint16_t>(), d.cmask.as<uint16_t>());
  ST_HIP(hipGetLastError());
}

template <class T>
GeoD<T> geo_view(Ctx& c) {
  DevGeo& d = c.geo;
  GeoD<T> g;
  g.cbase = d.cbase.as<uint32_t>();
  g.nbase = d.nbase.as<uint32_t>();
  g.bstruct = d.bstruct.as<uint32_t>();
  g.cfaces = d.cfaces.as<uint32_t>();
  g.nflags = d.nflags.as<uint8_t>();
  g.nkind = d.nkind.as<uint8_t>();
  g.ncell = d.ncell.as<int32_t>();
  g.ce = std::is_same<T, double>::value ? (const T*)d.ce64.p : (const T*)d.ce32.p;
  g.cg = std::is_same<T, double>::value ? (const T*)d.cg64.p : (const T*)d.cg32.p;
  g.cmask = d.cmask.as<uint16_t>();
  g.bgeo = d.bgeo.as<BeamGeo>();
  g.carea = !d.has_sec ? nullptr
            : std::is_same<T, double>::value ? (const T*)d.csec64.p : (const T*)d.csec32.p;
  return g;
}
template GeoD<double> geo_view<double>(Ctx&);
template GeoD<float> geo_view<float>(Ctx&);

// Folds the recorded event pairs into the running total (keeps the pool small).
void timing_collect(Ctx& c) {
  if (!c.tev_used) return;
  ST_HIP(hipEventSynchronize(c.tev[2 * c.tev_used - 1]));
  for (size_t i = 0; i < c.tev_used; i++) {
    float ms = 0;
    ST_HIP(hipEventElapsedTime(&ms, c.tev[2 * i], c.tev[2 * i + 1]));
    c.timing_ms_done += ms;
  }
  c.tev_used = 0;
}

// One application by the cell-centric kernel (kind 2) with kThreads a workgroup.
template <class T, class Epi, bool kSec, int kThreads>
void launch_cells(Ctx& c, const GeoD<T>& g, const T* x, const Epi& epi) {
  // the dynamic LDS limit is raised once per context and kernel instance (a context
  // owns its device; each keeps what it set, so contexts on other threads or devices
  // neither share nor race on it)
  const void* fn = reinterpret_cast<const void*>(&k_apply_cells<T, Epi, kSec, kThreads>);
  const DevGeo& d = c.geo;
  const size_t need = (size_t)std::max(d.class_nodes[0], d.class_nodes[1]) * 6 * sizeof(T);
  size_t& have = c.lds_attr[fn];
  if (need > have) {
    ST_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    have = need;
  }
  const uint32_t* bl = d.blist.as<uint32_t>();
  for (int k = 0; k < 2; k++) {
    if (d.nclass[k])
      k_apply_cells<T, Epi, kSec, kThreads>
          <<<(unsigned)d.nclass[k], kThreads, (size_t)d.class_nodes[k] * 6 * sizeof(T), c.stream>>>(g, bl, x, epi);
    bl += d.nclass[k];
  }
}

// One application in schedule s; kSec: the cells carry per-beam sections.
template <class T, class Epi, bool kSec>
void launch_schedule(Ctx& c, const ApplySchedule& s, const GeoD<T>& g, const T* x, const Epi& epi) {
  const unsigned nb = (unsigned)c.geo.nbricks;
  if (s.tile == 0)
    k_apply<T, Epi, 768, 256, kSec><<<nb, kBlock, 0, c.stream>>>(g, x, epi);
  else if (s.tile == 1)
    k_apply<T, Epi, 0, 0, kSec><<<nb, kBlock, 0, c.stream>>>(g, x, epi);
  else if (s.tile == 3)
    k_apply_nodes2<T, Epi, kSec><<<nb, kBlock, 0, c.stream>>>(g, x, epi);
  else
    launch_cells<T, Epi, kSec, kBlock>(c, g, x, epi);
}

template <class T, class Epi>
void launch_apply(Ctx& c, const T* x, const Epi& epi) {
  if (!c.geo.nbricks) return;
  // the CG's K p: FP64, or FP32 under ST_MIXED32
  const bool timed = std::is_same<Epi, EpiDot<T>>::value && c.timing_which == 0;
  if (timed) {
    if (c.tev_used * 2 + 2 > c.tev.size()) {
      if (c.tev.size() >= 4096) timing_collect(c);
      while (c.tev_used * 2 + 2 > c.tev.size()) {
        hipEvent_t e;
        ST_HIP(hipEventCreate(&e));
        c.tev.push_back(e);
      }
    }
    ST_HIP(hipEventRecord(c.tev[2 * c.tev_used], c.stream));
  }
  GeoD<T> g = geo_view<T>(c);
  const int sched = c.apply_sched[sizeof(T) == 8 ? 1 : 0];
  g.xcd = kSchedules[sched].xcd;
  if (g.carea)
    launch_schedule<T, Epi, true>(c, kSchedules[sched], g, x, epi);
  else
    launch_schedule<T, Epi, false>(c, kSchedules[sched], g, x, epi);
  ST_HIP(hipGetLastError());
  if (timed) {
    ST_HIP(hipEventRecord(c.tev[2 * c.tev_used + 1], c.stream));
    c.tev_used++;
    c.timing_launches++;
    // algorithmic bytes: x read and y written at every free node; e, g, the beam mask and
    // the six face indices of every free cell
    c.timing_bytes = (double)c.last_ndof * 6 * 2 * sizeof(T) + (double)c.geo.ncells * (2 * sizeof(T) + 2 + 6 * 4);
  }
}

// Explicit instantiations used by the solver.
#define ST_INST_APPLY(T, E) template void launch_apply<T, E>(Ctx&, const T*, const E&);
ST_INST_APPLY(double, EpiStore<double>)
ST_INST_APPLY(float, EpiStore<float>)
ST_INST_APPLY(double, EpiDot<double>)
ST_INST_APPLY(float, EpiDot<float>)
ST_INST_APPLY(double, EpiResidRestrict<double>)
ST_INST_APPLY(float, EpiResidRestrict<float>)
ST_INST_APPLY(double, SINGLE_ARG(EpiResid<double, double>))
ST_INST_APPLY(float, SINGLE_ARG(EpiResid<float, float>))
ST_INST_APPLY(float, SINGLE_ARG(EpiResid<float, double>))
ST_INST_APPLY(double, EpiSmooth<double>)
ST_INST_APPLY(float, EpiSmooth<float>)
ST_INST_APPLY(double, EpiResidJac<double>)
ST_INST_APPLY(float, EpiResidJac<float>)
ST_INST_APPLY(double, EpiProlong<double>)
ST_INST_APPLY(float, EpiProlong<float>)

namespace {
template <class T>
int tune_one(Ctx& c, T* x, T* y) {
  hipEvent_t e0, e1;
  ST_HIP(hipEventCreate(&e0));
  ST_HIP(hipEventCreate(&e1));
  EpiStore<T> es;
  es.y = y;
  int best = 0;
  float best_ms = 1e30f;
  const int saved = c.apply_sched[sizeof(T) == 8 ? 1 : 0];
  for (int k = 0; k < kNumSchedules; k++) {
    c.apply_sched[sizeof(T) == 8 ? 1 : 0] = k;
    launch_apply<T>(c, x, es);  // warm
    ST_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < 3; r++) launch_apply<T>(c, x, es);
    ST_HIP(hipEventRecord(e1, c.stream));
    ST_HIP(hipEventSynchronize(e1));
    float ms = 0;
    ST_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (std::getenv("STATICS_HIP_VERBOSE"))
      std::fprintf(stderr, "statics_hip: schedule %d (kind %d, xcd %d, %d threads) %s: %.3f ms\n", k,
                   kSchedules[k].tile, kSchedules[k].xcd, kSchedules[k].threads, sizeof(T) == 8 ? "fp64" : "fp32",
                   ms / 3);
    if (ms < best_ms) {
      best_ms = ms;
      best = k;
    }
  }
  c.apply_sched[sizeof(T) == 8 ? 1 : 0] = saved;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}
}  // namespace

// Picks the operator's schedule (LDS tile capacity, brick-to-XCD mapping) for each
// precision by timing every one on the uploaded structures; STATICS_HIP_SCHEDULE
// forces one. Runs once per upload.
void autotune_apply(Ctx& c) {
  if (c.tuned || !c.geo.nbricks) return;
  if (const char* e = std::getenv("STATICS_HIP_SCHEDULE")) {
    const int k = std::max(0, std::min(kNumSchedules - 1, std::atoi(e)));
    c.apply_sched[0] = c.apply_sched[1] = k;
    c.tuned = true;
    return;
  }
  const size_t n = (size_t)c.geo.nnodes * 6;
  DBuf a, b;
  a.ensure(n * sizeof(double));
  b.ensure(n * sizeof(double));
  ST_HIP(hipMemsetAsync(a.p, 0, n * sizeof(double), c.stream));
  c.apply_sched[0] = tune_one<float>(c, a.as<float>(), b.as<float>());
  c.apply_sched[1] = tune_one<double>(c, a.as<double>(), b.as<double>());
  ST_HIP(hipStreamSynchronize(c.stream));
  a.release();
  b.release();
  c.tuned = true;
}

}  // namespace st

using namespace st;

extern "C" {

int st_create(StCtx** out, int device) {
  if (!out) return ST_BAD_INPUT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ST_DEVICE;
  if (device < 0 || device >= n) return ST_BAD_INPUT;
  auto* c = new StCtx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ST_DEVICE;
  }
  // the fixed scratch of st_joined_to_held_static: its address on this device
  try {
    scratch_setup(device);
  } catch (const StError& e) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return e.code;
  }
  *out = c;
  return ST_OK;
}

int st_destroy(StCtx* c) {
  if (!c) return ST_BAD_INPUT;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  release_all(*c);
  for (auto e : c->tev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return ST_OK;
}

const char* st_last_error(const StCtx* c) { return c ? c->err.c_str() : "null context"; }

int st_upload(StCtx* c, const StCells* cells) {
  return guard(c, [&] {
    if (!cells) throw StError{ST_BAD_INPUT, "cells is NULL"};
    ST_HIP(hipSetDevice(c->device));
    const StCells* list[1] = {cells};
    build_host_geo(c->host, list, 1);
    copy_geo_to_device(*c);
    c->gone_host.assign(c->geo.nnodes, 0);
    refresh_cell_flags(*c);
    ST_HIP(hipStreamSynchronize(c->stream));
    c->uploaded = true;
    c->have_solution = false;
  });
}

int st_sizes(StCtx* c, int64_t* nodes, int64_t* cells, int64_t* bricks) {
  return guard(c, [&] {
    need_upload(c);
    if (nodes) *nodes = c->geo.nnodes;
    if (cells) *cells = c->geo.ncells;
    if (bricks) *bricks = c->geo.nbricks;
  });
}

int st_node_faces(StCtx* c, int32_t* faces) {
  return guard(c, [&] {
    need_upload(c);
    if (!faces) throw StError{ST_BAD_INPUT, "faces is NULL"};
    std::memcpy(faces, c->host.nface.data(), c->host.nface.size() * sizeof(int32_t));
  });
}

int st_lookup_faces(StCtx* c, int64_t n, const int32_t* faces, int64_t* index) {
  return guard(c, [&] {
    need_upload(c);
    if (n < 0 || (n && (!faces || !index))) throw StError{ST_BAD_INPUT, "lookup: bad arguments"};
    struct H {
      size_t operator()(const std::array<int32_t, 4>& a) const {
        uint64_t h = 1469598103934665603ull;
        for (auto v : a) h = (h ^ (uint32_t)v) * 1099511628211ull;
        return (size_t)h;
      }
    };
    std::unordered_map<std::array<int32_t, 4>, int64_t, H> m;
    m.reserve(c->geo.nnodes * 2);
    const auto& nf = c->host.nface;
    for (int64_t v = 0; v < c->geo.nnodes; v++) m[{nf[4 * v], nf[4 * v + 1], nf[4 * v + 2], nf[4 * v + 3]}] = v;
    for (int64_t i = 0; i < n; i++) {
      auto it = m.find({faces[4 * i], faces[4 * i + 1], faces[4 * i + 2], faces[4 * i + 3]});
      index[i] = it == m.end() ? -1 : it->second;
    }
  });
}

int st_node_held(StCtx* c, uint8_t* held) {
  return guard(c, [&] {
    need_upload(c);
    if (!held) throw StError{ST_BAD_INPUT, "held is NULL"};
    for (int64_t v = 0; v < c->geo.nnodes; v++) held[v] = (c->host.nflags[v] & kHeld) ? 1 : 0;
  });
}

int st_set_gone(StCtx* c, const uint8_t* gone) {
  return guard(c, [&] {
    need_upload(c);
    set_gone_flags(*c, gone);
  });
}

int st_set_beams(StCtx* c, const uint16_t* beams) {
  return guard(c, [&] {
    need_upload(c);
    set_beam_masks(*c, beams);
  });
}

int st_set_stiffness(StCtx* c, const double* e, const double* g) {
  return guard(c, [&] {
    need_upload(c);
    if (!e || !g) throw StError{ST_BAD_INPUT, "set_stiffness: e and g are required"};
    ST_HIP(hipSetDevice(c->device));
    HostGeo& h = c->host;
    const auto& coi = h.cell_of_input;
    for (size_t i = 0; i < coi.size(); i++) {
      const int64_t k = coi[i];
      if (k < 0) continue;  // held
      if (!(e[i] > 0) || !(g[i] > 0) || !std::isfinite(e[i]) || !std::isfinite(g[i]))
        throw StError{ST_BAD_INPUT, "set_stiffness: a free cell has a non-positive or non-finite stiffness"};
      h.ce[k] = e[i];
      h.cg[k] = g[i];
    }
    DevGeo& d = c->geo;
    up(d.ce64, h.ce, c->stream);
    up(d.cg64, h.cg, c->stream);
    if (d.ncells) {
      const int nb = (int)((d.ncells + 255) / 256);
      k_to_float<<<nb, 256, 0, c->stream>>>(d.ncells, d.ce64.as<double>(), d.ce32.as<float>());
      k_to_float<<<nb, 256, 0, c->stream>>>(d.ncells, d.cg64.as<double>(), d.cg32.as<float>());
      ST_HIP(hipGetLastError());
    }
    ST_HIP(hipStreamSynchronize(c->stream));
    c->have_solution = false;
    c->hier_valid = false;
  });
}

int st_kernel_timing_enable(StCtx* c, int which, int enable) {
  return guard(c, [&] {
    if (which != 0 && which != 1) throw StError{ST_BAD_INPUT, "kernel timing: which is 0 or 1"};
    timing_collect(*c);
    if (enable) {
      c->timing_which = which;
      c->timing_ms_done = 0.0;
      c->timing_launches = 0;
    } else {
      c->timing_which = -1;
    }
  });
}

int st_kernel_timing(StCtx* c, int64_t* launches, double* total_ms, double* bytes) {
  return guard(c, [&] {
    timing_collect(*c);
    if (launches) *launches = c->timing_launches;
    if (total_ms) *total_ms = c->timing_ms_done;
    if (bytes) *bytes = c->timing_bytes;
  });
}

int st_apply_schedule(StCtx* c, int fp64, int* index, const char** kernel, const char** description,
                      int* dispatches) {
  return guard(c, [&] {
    need_upload(c);
    autotune_apply(*c);
    const int sched = c->apply_sched[fp64 ? 1 : 0];
    const ApplySchedule& k = kSchedules[sched];
    if (index) *index = sched;
    static const char* const names[4] = {"k_apply", "k_apply", "k_apply_cells", "k_apply_nodes2"};
    static const char* const what[4] = {"node-centric, LDS tile of 768 nodes / 256 cells",
                                        "node-centric, no LDS tile", "cell-centric, brick rows in LDS",
                                        "node-centric, no LDS tile, both cells' loads together"};
    static std::string desc[kNumSchedules];
    static std::once_flag once;
    std::call_once(once, [] {
      for (int i = 0; i < kNumSchedules; i++)
        desc[i] = std::string(what[kSchedules[i].tile]) +
                  (kSchedules[i].xcd ? ", XCD brick mapping" : ", bricks in launch order") + ", " +
                  std::to_string(kSchedules[i].threads) + " threads a brick";
    });
    if (kernel) *kernel = names[k.tile];
    if (description) *description = desc[sched].c_str();
    if (dispatches) *dispatches = k.tile == 2 ? (c->geo.nclass[0] > 0) + (c->geo.nclass[1] > 0) : 1;
  });
}

int st_apply(StCtx* c, const double* x, double* y) {
  return guard(c, [&] {
    need_upload(c);
    if (!x || !y) throw StError{ST_BAD_INPUT, "apply: x or y is NULL"};
    ST_HIP(hipSetDevice(c->device));
    const size_t bytes = (size_t)c->geo.nnodes * 6 * sizeof(double);
    autotune_apply(*c);
    c->t64.ensure(bytes);
    c->u64.ensure(bytes);
    ST_HIP(hipMemcpyAsync(c->t64.p, x, bytes, hipMemcpyHostToDevice, c->stream));
    EpiStore<double> e;
    e.y = c->u64.as<double>();
    launch_apply<double>(*c, c->t64.as<double>(), e);
    ST_HIP(hipMemcpyAsync(y, c->u64.p, bytes, hipMemcpyDeviceToHost, c->stream));
    ST_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
