// cgs_sort_cache.h — SortCache state (src/render/pipeline.rs:66-193) shared
// by sort_cache.cpp (its C ABI) and runtime.cpp (the hysteresis frame mode).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/cindergs.h"

// The binning the state stands for is held by its owner.
#define CGS_CAMERA_MOTION_EPSILON 1e-5f
#define CGS_CAGE_MOTION_EPSILON 1e-5f
struct cgs_sort_cache {
  bool has_view = false;
  float view[16] = {};
  std::vector<float> cage_state;  // 3 floats per cage vertex
  uint32_t width = 0, height = 0;
  uint64_t splat_count = 0;
  uint32_t focal_x_bits = 0, focal_y_bits = 0, split_bits = 0, mip_bits = 0;
  bool has_binning = false;
  uint64_t reuse_count = 0;

  bool can_reuse(const float* v, const float* cage, uint64_t nv, uint32_t w, uint32_t h) const;
  bool params_match(const cgs_sort_frame& f) const;
  bool take(bool params_match, const float* v, const float* cage, uint64_t nv, uint32_t w,
            uint32_t h);
  void commit(const float* v, const float* cage, uint64_t nv, uint32_t w, uint32_t h);
  void commit_frame(const cgs_sort_frame& f, const float* v, const float* cage, uint64_t nv,
                    uint32_t w, uint32_t h);
  void invalidate();
};
