// Declarations shared by the library's translation units.
#pragma once
#include <exception>
#include "st_kernels.hpp"

#define SINGLE_ARG(...) __VA_ARGS__

namespace st {

template <class T>
GeoD<T> geo_view(Ctx& c);
template <class T, class Epi>
void launch_apply(Ctx& c, const T* x, const Epi& epi);
void refresh_cell_flags(Ctx& c);
void release_all(Ctx& c);
void release_levels(Ctx& c);
void release_hierarchy(Ctx& c);
void solve_all(Ctx& c, const StSolveOpts& o, std::vector<StStats>& stats);
void joined_to_held(Ctx& c, std::vector<uint8_t>& joined);
void copy_geo_to_device(Ctx& c);
void timing_collect(Ctx& c);
void classify_device(Ctx& c, int64_t* joined_count, uint8_t* joined_host);
double device_dot_f_x(Ctx& c, const double* a, const double* b);
void set_gone_flags(Ctx& c, const uint8_t* gone);
void set_beam_masks(Ctx& c, const uint16_t* beams);
void autotune_apply(Ctx& c);
void scratch_setup(int device);
void joined_static(Ctx& c, hipStream_t s, uint8_t* joined, int32_t* count);

}  // namespace st

struct StCtx : st::Ctx {};

namespace st {
template <class F>
int guard(StCtx* c, F f) {
  if (!c) return ST_BAD_INPUT;
  try {
    c->err.clear();
    f();
    return ST_OK;
  } catch (const StError& e) {
    c->err = e.what;
    return e.code;
  } catch (const std::bad_alloc&) {
    c->err = "host allocation failed";
    return ST_OOM;
  } catch (const std::exception& e) {  // e.g. a host thread that could not start
    c->err = std::string("host error: ") + e.what();
    return ST_DEVICE;
  }
}
inline void need_upload(StCtx* c) {
  if (!c->uploaded) throw StError{ST_NO_STRUCTURE, "no structure uploaded"};
}
}  // namespace st
