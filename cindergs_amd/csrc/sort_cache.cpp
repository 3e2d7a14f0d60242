// Temporal hysteresis for Stage 3: SortCache (src/render/pipeline.rs:40-63,
// 66-193).  The cache remembers the view matrix, the flattened cage vertices,
// the frame size and the frame parameters of the last committed binning; a
// later frame may take the binning back when its view matrix moved by at most
// 1e-5 (largest absolute element difference), no cage vertex coordinate moved
// by more than 1e-5, the size is the same and, for the *_for_frame forms, the
// splat count and the bits of focal_x, focal_y, the split threshold and the
// mip variance are equal.  The binning itself stays with the caller (the HIP
// runtime keeps it in the frame arena); the cache tracks whether one is held.
#include <cmath>
#include <cstring>
#include <limits>
#include <new>

#include "cgs_internal.h"
#include "cgs_sort_cache.h"

namespace {

uint32_t bits_of(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

// f32::max: a NaN operand yields the other operand
float fmax_rs(float a, float b) { return std::fmax(a, b); }

float max_abs_difference(const float* a, const float* b) {
  float m = 0.0f;
  for (int i = 0; i < 16; ++i) m = fmax_rs(m, std::fabs(a[i] - b[i]));
  return m;
}

float max_cage_difference(const std::vector<float>& a, const float* b, uint64_t nv) {
  if (a.size() != nv * 3) return std::numeric_limits<float>::infinity();
  float m = 0.0f;
  for (uint64_t i = 0; i < nv * 3; ++i) m = fmax_rs(m, std::fabs(a[i] - b[i]));
  return m;
}

}  // namespace

bool cgs_sort_cache::can_reuse(const float* v, const float* cage, uint64_t nv, uint32_t w,
                               uint32_t h) const {
  if (!has_view || !has_binning || width != w || height != h) return false;
  return max_abs_difference(view, v) <= CGS_CAMERA_MOTION_EPSILON &&
         max_cage_difference(cage_state, cage, nv) <= CGS_CAGE_MOTION_EPSILON;
}

bool cgs_sort_cache::take(bool params_match, const float* v, const float* cage, uint64_t nv,
                          uint32_t w, uint32_t h) {
  if (params_match && can_reuse(v, cage, nv, w, h)) {
    ++reuse_count;
    has_binning = false;
    return true;
  }
  has_binning = false;
  return false;
}

bool cgs_sort_cache::params_match(const cgs_sort_frame& f) const {
  return splat_count == f.splat_count && focal_x_bits == bits_of(f.focal_x) &&
         focal_y_bits == bits_of(f.focal_y) && split_bits == bits_of(f.split_threshold) &&
         mip_bits == bits_of(f.mip_var);
}

void cgs_sort_cache::commit(const float* v, const float* cage, uint64_t nv, uint32_t w,
                            uint32_t h) {
  memcpy(view, v, sizeof(view));
  has_view = true;
  cage_state.assign(cage, cage + nv * 3);
  width = w;
  height = h;
  has_binning = true;
}

void cgs_sort_cache::commit_frame(const cgs_sort_frame& f, const float* v, const float* cage,
                                  uint64_t nv, uint32_t w, uint32_t h) {
  splat_count = f.splat_count;
  focal_x_bits = bits_of(f.focal_x);
  focal_y_bits = bits_of(f.focal_y);
  split_bits = bits_of(f.split_threshold);
  mip_bits = bits_of(f.mip_var);
  commit(v, cage, nv, w, h);
}

void cgs_sort_cache::invalidate() {
  has_binning = false;
  has_view = false;
  cage_state.clear();
  splat_count = 0;
}

// ---- C ABI ----

namespace {
int bad(const char* what) { return cgs_fail(CGS_ERR_INVALID_ARG, what); }
bool cage_ok(const float* cage, uint64_t nv) { return nv == 0 || cage != nullptr; }
}  // namespace

extern "C" {

int cgs_sort_cache_create(cgs_sort_cache** out) {
  if (!out) return bad("null output");
  *out = new (std::nothrow) cgs_sort_cache();
  return *out ? CGS_OK : bad("allocation failed");
}

void cgs_sort_cache_destroy(cgs_sort_cache* s) { delete s; }

int cgs_sort_cache_can_reuse(const cgs_sort_cache* s, const float* view, const float* cage,
                             uint64_t num_cage_vertices, uint32_t width, uint32_t height,
                             uint32_t* reusable) {
  if (!s || !view || !reusable || !cage_ok(cage, num_cage_vertices)) return bad("null argument");
  *reusable = s->can_reuse(view, cage, num_cage_vertices, width, height) ? 1u : 0u;
  return CGS_OK;
}

int cgs_sort_cache_take_if_reusable(cgs_sort_cache* s, const float* view, const float* cage,
                                    uint64_t num_cage_vertices, uint32_t width, uint32_t height,
                                    uint32_t* reused) {
  if (!s || !view || !reused || !cage_ok(cage, num_cage_vertices)) return bad("null argument");
  *reused = s->take(true, view, cage, num_cage_vertices, width, height) ? 1u : 0u;
  return CGS_OK;
}

int cgs_sort_cache_take_if_reusable_for_frame(cgs_sort_cache* s, const float* view,
                                              const float* cage, uint64_t num_cage_vertices,
                                              uint32_t width, uint32_t height,
                                              const cgs_sort_frame* frame, uint32_t* reused) {
  if (!s || !view || !frame || !reused || !cage_ok(cage, num_cage_vertices))
    return bad("null argument");
  *reused = s->take(s->params_match(*frame), view, cage, num_cage_vertices, width, height) ? 1u
                                                                                           : 0u;
  return CGS_OK;
}

int cgs_sort_cache_commit(cgs_sort_cache* s, const float* view, const float* cage,
                          uint64_t num_cage_vertices, uint32_t width, uint32_t height) {
  if (!s || !view || !cage_ok(cage, num_cage_vertices)) return bad("null argument");
  s->commit(view, cage, num_cage_vertices, width, height);
  return CGS_OK;
}

int cgs_sort_cache_commit_for_frame(cgs_sort_cache* s, const float* view, const float* cage,
                                    uint64_t num_cage_vertices, uint32_t width, uint32_t height,
                                    const cgs_sort_frame* frame) {
  if (!s || !view || !frame || !cage_ok(cage, num_cage_vertices)) return bad("null argument");
  s->commit_frame(*frame, view, cage, num_cage_vertices, width, height);
  return CGS_OK;
}

int cgs_sort_cache_invalidate(cgs_sort_cache* s) {
  if (!s) return bad("null argument");
  s->invalidate();
  return CGS_OK;
}

int cgs_sort_cache_reuse_count(const cgs_sort_cache* s, uint64_t* count) {
  if (!s || !count) return bad("null argument");
  *count = s->reuse_count;
  return CGS_OK;
}

}  // extern "C"
