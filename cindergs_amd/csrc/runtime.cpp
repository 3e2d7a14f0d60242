// runtime.cpp — host runtime behind include/cindergs.h.
//
// Re-creates the state machine of the reference's GpuSceneState /
// GpuFrameArena / GpuPipelineState (src/render/gpu_state.rs) on HIP:
//   * scene residency with the same fingerprint rule (count, data pointers,
//     64-sample FNV-1a; gpu_state.rs:65-79) and cage topology / vertex diffing
//     (gpu_state.rs:215-343) that drives the Stage 1 dirty flag;
//   * an arena of device buffers grown by 1.5x with a 256-word floor
//     (gpu_state.rs:412-429, gpu_pool.rs:43-49);
//   * the Stage 3 cache keyed on deform generation + camera + size + culling
//     parameters (gpu_state.rs:454-499, 956-981);
//   * execute_frame (gpu_state.rs:613-1243).
// Differences by design: counts (split children, survivors, tile entries)
// stay on the device for the whole frame and are read once at the end; the
// tile-entry arena is sized from the previous frames and a frame that
// overflows it is re-run from Stage 3 with a larger arena.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cindergs.h"
#include "cgs_internal.h"
#include "cgs_scene_math.h"
#include "cgs_sort_cache.h"

namespace {

thread_local std::string g_err;
// set while a view batch is being captured into a HIP graph: buffers are sized
// before the capture, so a growth inside it is an internal error
thread_local bool g_no_alloc = false;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(CGS_ERR_EXECUTION, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap_words = 0;  // 4-byte words
};

// 1.5x growth with a 256-word floor (calculate_exponential_growth, gpu_pool.rs:43-49)
int ensure_words(DevBuf& b, size_t words) {
  size_t req = words < 1 ? 1 : words;
  if (b.p && b.cap_words >= req) return CGS_OK;
  if (g_no_alloc) return fail(CGS_ERR_EXECUTION, "internal: buffer growth during graph capture");
  size_t cap;
  if (!b.p) {
    cap = req < 256 ? 256 : req;
  } else {
    cap = b.cap_words + b.cap_words / 2;
    if (cap < req) cap = req;
  }
  if (b.p) {
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap_words = 0;
  }
  HIPCHK(hipMalloc(&b.p, cap * 4));
  b.cap_words = cap;
  return CGS_OK;
}

void release(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap_words = 0;
}

template <class T>
T* P(DevBuf& b) {
  return static_cast<T*>(b.p);
}
template <typename T>
T* P(const DevBuf& b) {
  return static_cast<T*>(b.p);
}

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
  const uint8_t* p = static_cast<const uint8_t*>(data);
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001B3ull;
  }
  return h;
}
uint64_t fnv_u64(uint64_t h, uint64_t v) { return fnv1a(h, &v, 8); }

// hash_f32_sample / hash_u32_sample (gpu_state.rs:37-63)
uint64_t hash_sample(uint64_t h, const uint32_t* data, size_t len) {
  h = fnv_u64(h, (uint64_t)len);
  size_t step = len / 64;
  if (step < 1) step = 1;
  for (size_t i = 0; i < len; i += step) h = fnv1a(h, &data[i], 4);
  if (len) h = fnv1a(h, &data[len - 1], 4);
  return h;
}

const uint64_t FNV_OFFSET = 0xcbf29ce484222325ull;

// Scratch of cgs_exclusive_scan_u32_noalloc: [0] total, [1] partition ticket,
// [4 ..) one u64 look-back status word per 4096-element partition.  A copy
// lives on every GPU; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (kept in g_scratch).
constexpr uint32_t NOALLOC_SCAN_WORDS = 4u + 2u * (CGS_SCAN_NOALLOC_MAX_N / 4096u);
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCAN_WORDS];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[MAX_DEVICES] = {};

// the scratch address on `device` (which must be current), resolved once
int scratch_for_device(int device, uint32_t** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(CGS_ERR_UNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CGS_ERR_RUNTIME, std::string("scratch symbol lookup failed: ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch[device];
  return CGS_OK;
}

struct Stage3Key {
  uint64_t deform_generation;
  uint32_t view_bits[16];
  uint32_t fx, fy, cx, cy, width, height, near_bits, mip_bits, min_op_bits, min_rad_bits;
  uint32_t tile_w, tile_h;
  bool operator==(const Stage3Key& o) const { return memcmp(this, &o, sizeof(*this)) == 0; }
};

uint32_t fbits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

}  // namespace

// Per-view frame buffers (GpuFrameArena, gpu_state.rs:384-429): the projected
// splats, the tile entries and their sort buffers, the tile ranges, the frame
// control words and the framebuffer.  A context owns one; a view batch owns one
// per concurrent lane.
struct FrameArena {
  DevBuf geom, color, extents;
  DevBuf e_keys, e_ids, e_depths, e_alt_k, e_alt_v, e_perm;
  uint32_t entry_cap = 0;
  DevBuf ranges, out_rgba, cnt_dbg, off_dbg, comp_order;
  DevBuf hitm;   // hit masks from k_count to k_emit
  DevBuf ctrl3;  // frame control, see frame_layout()
  // presort (tile_sort = 2): the splats' depth keys, their sorted order, and
  // the control words of that sort; `order` is the order of the last Stage 3
  // (null: it was not presorted)
  DevBuf sp_keys, sp_vals, sp_alt_k, sp_alt_v, sp_ctrl;
  const uint32_t* order = nullptr;
  uint32_t tiles_x = 0, tiles_y = 0, padded_row = 0;
  bool mode64 = false;
  uint32_t depth_bits = 0;
  uint32_t* sorted_keys = nullptr;
  uint32_t* sorted_ids = nullptr;
  bool capture = false;
  bool captured = false;
  cgs_tuning tune = {16384u, 8u, 64u, 1u, 0u, 16u, 16u};
  // bucket mode (k_tilesort.hip): histograms, bases, buckets, the sort's
  // tables and control words, and the launch plan of the segmented sort
  DevBuf bk_hist, bk_base, bk_bucket, bk_alt, bk_tab, bk_ctl;
  bool bucket = false;             // the last Stage 3 ran in bucket mode
  uint32_t bk_chunk_blocks = 0;    // chunk blocks / merge rounds it launched
  uint32_t bk_rounds = 0;
  bool bk_planned = false;         // the plan of the next launch: chunks / rounds
  uint32_t bk_need_chunks = 0, bk_need_rounds = 0;
  // what the last frame of each geometry (tiles, splat slots) needed; a view
  // batch sets the plan of each view itself (plan_external)
  std::vector<std::pair<uint64_t, std::pair<uint32_t, uint32_t>>> bk_plans;
  uint64_t bk_key = 0;
  bool plan_external = false;
  // projected records read by Stages 3-4 when they live outside the arena
  // (the view batch's multi-view projection); null: geom / color / extents
  const float* geom_in = nullptr;
  const float* color_in = nullptr;
  const uint32_t* extents_in = nullptr;

  void release_all() {
    DevBuf* bufs[] = {&geom,    &color,  &extents, &e_keys,   &e_ids,     &e_depths,
                      &e_alt_k, &e_alt_v, &e_perm,  &ranges,   &out_rgba,  &cnt_dbg,
                      &off_dbg, &comp_order, &hitm, &ctrl3,    &bk_hist,   &bk_base,
                      &bk_bucket, &bk_alt, &bk_tab, &bk_ctl, &sp_keys, &sp_vals,
                      &sp_alt_k, &sp_alt_v, &sp_ctrl};
    for (DevBuf* b : bufs) release(*b);
    order = nullptr;
    entry_cap = 0;
    bk_planned = false;
    bk_plans.clear();
  }
};

struct NearCache;
static void near_cache_free(NearCache* n);

struct cgs_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {};
  char name[256] = {0};

  // ---- scene residency ----
  bool scene_known = false;
  uint64_t scene_hash = 0;
  uint32_t n_rest = 0;
  bool f16 = false;
  uint32_t f16_degree = 3;
  DevBuf rest_geo, rest_cov_h, rest_sh, rest_sh_h, staging;  // wave-blocked

  // ---- cages ----
  bool has_cages = false;
  uint64_t topo_hash = 0;
  uint32_t num_tets = 0, num_verts = 0;
  DevBuf elems, inv_dm, bind_tet, bind_w, verts;
  std::vector<float> last_verts;
  uint32_t last_params[2] = {0, 0};
  bool deformed_valid = false;
  uint64_t deform_generation = 0;

  // ---- arena ----
  DevBuf jac, rot, stretch, tet_pack, def_geo, split_meta;
  uint32_t def_cap = 0;  // splat capacity of def_geo
  DevBuf ctrl1;  // [split_count, ticket, pad, pad][u64 status...]
  FrameArena fa;
  uint32_t* h_counters = nullptr;  // pinned [4]: survivors, split_count, total, pad
  uint32_t* scan_scratch = nullptr;  // this device's copy of `scratch` (never the symbol itself)
  NearCache* near_cache = nullptr;  // the last near/far binding's cage buffers and grid
  DevBuf prim_ctrl, prim_alt_k, prim_alt_v, prim_alt_d;
  // MVC scene deformation: per-call working set
  DevBuf mvc_work;
  // scene ingest: the vertex payload of a PLY file
  DevBuf ingest;
  // export of a deformed scene: per-tet SH band matrices
  DevBuf tet_shm;

  // ---- last frame ----
  uint32_t last_w = 0, last_h = 0;
  uint32_t max_g = 0;          // splat slots of the projected arrays
  uint32_t active = 0;         // active splats of the last frame
  uint32_t n_entries = 0;
  bool s3_valid = false;
  Stage3Key s3_key{};
  uint32_t s3_survivors = 0;
  uint32_t s3_max_g = 0;       // splat slots when the held tile order was sorted
  // SortCache hysteresis mode (cgs_set_sort_hysteresis)
  bool hysteresis = false;
  cgs_sort_cache sort_cache;
  CgsProf prof;
  int prof_count = 0;
  double prof_ms[CgsProf::MAX];
  const char* prof_names[CgsProf::MAX];
};

namespace {

// frame control words (zeroed once per frame before Stage 2)
// Stage 3 control words: [0, zero_words) is cleared before every Stage 3 run;
// the per-partition histogram rows after it are fully written by the binning.
struct FrameLayout {
  size_t total;                   // [entries, survivors]
  size_t rA_tickets, rB_tickets;  // 8 each
  size_t histA, histB;            // 1024 each, contiguous (64-bit path reduces both at once)
  size_t block_sums, block_surv;  // [bin_parts] each, fully written by the binning
  size_t rA_status, rB_status;
  size_t zero_words;
  size_t bin_hist;                // [bin_parts][hist_words]
  size_t words;
  uint32_t bin_parts, r_parts;
};

FrameLayout frame_layout(uint32_t max_g, uint32_t cap) {
  FrameLayout L;
  L.total = 2;
  L.rA_tickets = 4;   // 8 words: up to 8 passes of 4-bit digits
  L.rB_tickets = 12;
  L.histA = 20;
  L.histB = 20 + 1024;
  size_t w = 20 + 2048;
  if (w & 1) w++;
  L.bin_parts = cgs_bin_num_parts(max_g);
  L.block_sums = w;
  w += (size_t)(L.bin_parts ? L.bin_parts : 1);
  L.block_surv = w;
  w += (size_t)(L.bin_parts ? L.bin_parts : 1);
  L.r_parts = cgs_radix_num_parts(cap);
  size_t rs = 4 * (size_t)(L.r_parts ? L.r_parts : 1) * 256;
  L.rA_status = w;
  w += rs;
  L.rB_status = w;
  w += rs;
  L.zero_words = w;
  L.bin_hist = w;
  w += (size_t)(L.bin_parts ? L.bin_parts : 1) * 2048;
  L.words = w;
  return L;
}

int ensure_ctx(cgs_context* c) {
  if (!c) return fail(CGS_ERR_INVALID_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  return CGS_OK;
}

#define TRY(expr)              \
  do {                         \
    int r_ = (expr);           \
    if (r_ != CGS_OK) return r_; \
  } while (0)

// host records src[i*k_in + c] -> channels chan0.. of a k_out-channel blocked array
int upload_blocked_u32(cgs_context* c, DevBuf& dst, const void* src, uint32_t n, uint32_t k_in,
                       uint32_t k_out, uint32_t chan0) {
  TRY(ensure_words(dst, cgs_blk_words(n, k_out)));
  if ((size_t)n * k_in == 0) return CGS_OK;
  TRY(ensure_words(c->staging, (size_t)n * k_in));
  HIPCHK(hipMemcpyAsync(c->staging.p, src, (size_t)n * k_in * 4, hipMemcpyHostToDevice, c->stream));
  cgs_launch_to_blocked_u32(P<uint32_t>(c->staging), P<uint32_t>(dst), n, k_in, k_out, chan0,
                            c->stream);
  HIPCHK(hipGetLastError());
  return CGS_OK;
}

uint64_t scene_fingerprint(const float* pos, const float* cov, const float* opa, const float* sh,
                           uint64_t n, size_t cov_words, size_t sh_words) {
  uint64_t h = FNV_OFFSET;
  h = fnv_u64(h, n);
  h = fnv_u64(h, (uint64_t)(uintptr_t)pos);
  h = fnv_u64(h, (uint64_t)(uintptr_t)cov);
  h = fnv_u64(h, (uint64_t)(uintptr_t)opa);
  h = fnv_u64(h, (uint64_t)(uintptr_t)sh);
  h = hash_sample(h, reinterpret_cast<const uint32_t*>(pos), (size_t)n * 3);
  h = hash_sample(h, reinterpret_cast<const uint32_t*>(cov), (size_t)n * cov_words);
  h = hash_sample(h, reinterpret_cast<const uint32_t*>(opa), (size_t)n);
  h = hash_sample(h, reinterpret_cast<const uint32_t*>(sh), (size_t)n * sh_words);
  return h;
}

void scene_changed(cgs_context* c) {
  c->deformed_valid = false;
  c->topo_hash = 0;  // force topology re-upload (bindings are per splat)
  c->has_cages = false;
  c->deform_generation++;
}


struct FrameGeom {
  uint32_t W, H, tiles_x, tiles_y, T, padded_row;
  uint32_t tile_w, tile_h;  // tile size in pixels
};

// the frame's tile grid (TileGrid::new, tiles/mod.rs:39-49, with the tuned
// tile size in place of TILE_SIZE)
FrameGeom frame_geom(const cgs_compositor_uniforms* comp, const cgs_tuning& tune) {
  FrameGeom g;
  g.W = comp->screen_width < 1 ? 1 : comp->screen_width;
  g.H = comp->screen_height < 1 ? 1 : comp->screen_height;
  g.tile_w = tune.tile_w;
  g.tile_h = tune.tile_h;
  g.tiles_x = (g.W + g.tile_w - 1u) / g.tile_w;
  g.tiles_y = (g.H + g.tile_h - 1u) / g.tile_h;
  g.T = g.tiles_x * g.tiles_y;
  g.padded_row = (g.W + 63) / 64 * 64;  // padded_rgba8_row_u32 (gpu_state.rs:25-27)
  return g;
}

void init_entry_cap(FrameArena& a, uint32_t max_g) {
  if (a.entry_cap != 0) return;
  uint64_t guess = (uint64_t)max_g * 4u;
  if (guess < (1u << 20)) guess = 1u << 20;
  if (guess > (1u << 29)) guess = 1u << 29;
  a.entry_cap = (uint32_t)guess;
}

uint32_t ceil_log2_u32(uint32_t v) { return v <= 1u ? 0u : 32u - (uint32_t)__builtin_clz(v - 1u); }

// control words of the bucket mode: [0, 4) stats (max tile entries, merge
// rounds, entries, chunks), [4] ticket, then one u64 look-back status per
// 128-tile group
size_t bucket_ctl_words(uint32_t tiles_pad) { return 8u + 2u * (size_t)(tiles_pad / 128u); }

// buffer sizes of the bucket mode for max_g splat slots, T tiles and an entry
// capacity of cap
struct BucketSizes {
  uint32_t tiles_pad, parts, subs, grid, windows, win_slots, chunk_slots;
  size_t tab_words, ctl_words;
};
BucketSizes bucket_sizes(uint32_t max_g, uint32_t T, uint32_t cap) {
  BucketSizes z;
  z.tiles_pad = (T + 127u) / 128u * 128u;
  z.parts = cgs_bin_num_parts(max_g);
  // about 512 blocks; one block covers at most 127 x 512 < 65536 splats (the
  // u16 histogram halves)
  z.subs = (z.parts + 511u) / 512u;
  if (z.subs < 1u) z.subs = 1u;
  if (z.subs > 127u) z.subs = 127u;
  z.grid = (z.parts + z.subs - 1u) / z.subs;
  z.windows = cgs_seg_sort_windows(cap);
  z.win_slots = z.windows + 1u;
  z.chunk_slots = cap / CGS_SEG_CHUNK + cap / CGS_SEG_WINDOW + 2u;
  z.tab_words = 2u * ((size_t)T + 1u) + z.win_slots + z.chunk_slots;
  z.ctl_words = bucket_ctl_words(z.tiles_pad);
  return z;
}

bool bucket_mode(const FrameArena& a, uint32_t T, uint32_t max_g) {
  return a.tune.tile_sort == 1u && T <= CGS_BUCKET_MAX_TILES && cgs_bin_num_parts(max_g) != 0;
}

// the presort's buffers for max_g splat slots
int size_presort(FrameArena& a, uint32_t max_g) {
  const size_t n = max_g ? max_g : 1u;
  TRY(ensure_words(a.sp_keys, n));
  TRY(ensure_words(a.sp_vals, n));
  TRY(ensure_words(a.sp_alt_k, n));
  TRY(ensure_words(a.sp_alt_v, n));
  TRY(ensure_words(a.sp_ctrl, cgs_radix_control_words((uint32_t)n)));
  return CGS_OK;
}

int size_bucket(FrameArena& a, uint32_t max_g, uint32_t T) {
  const uint32_t cap = a.entry_cap;
  const BucketSizes z = bucket_sizes(max_g, T, cap);
  TRY(ensure_words(a.bk_bucket, (size_t)cap * 2u));
  TRY(ensure_words(a.bk_alt, (size_t)cap * 2u));
  TRY(ensure_words(a.bk_hist, (size_t)z.grid * (z.tiles_pad / 2u)));
  TRY(ensure_words(a.bk_base, (size_t)z.grid * z.tiles_pad));
  TRY(ensure_words(a.bk_tab, z.tab_words));
  TRY(ensure_words(a.bk_ctl, z.ctl_words));
  return CGS_OK;
}

// Stage 3 in bucket mode (k_count with tile histograms, k_bucket_scan,
// k_emit into tile buckets, k_seg_sort and its merge rounds): the same keys,
// ids and ranges as the LSD sort.  The segmented sort's grid and merge rounds
// follow the previous frame's needs; the host re-runs a frame that needed
// more (bucket_rerun).
int launch_stage3_bucket(FrameArena& a, hipStream_t s, CgsProf* prof, const FrameGeom& g,
                         float near_plane, uint32_t num_rest, const uint32_t* split_count,
                         uint32_t max_g) {
  const uint32_t cap = a.entry_cap;
  const uint32_t T = g.T;
  const BucketSizes z = bucket_sizes(max_g, T, cap);
  const uint32_t tiles_pad = z.tiles_pad, parts = z.parts, subs = z.subs, grid = z.grid;
  const uint32_t windows = z.windows, win_slots = z.win_slots, chunk_slots = z.chunk_slots;
  const size_t ctl_words = z.ctl_words;
  TRY(ensure_words(a.e_keys, cap));
  TRY(ensure_words(a.e_ids, cap));
  TRY(ensure_words(a.ranges, (size_t)T * 2));
  TRY(size_bucket(a, max_g, T));
  const FrameLayout L = frame_layout(max_g, cap);
  TRY(ensure_words(a.ctrl3, L.words));
  uint32_t* ctrl = P<uint32_t>(a.ctrl3);
  TRY(ensure_words(a.cnt_dbg, max_g));
  TRY(ensure_words(a.hitm, cgs_bin_mask_words(max_g)));
  a.captured = false;
  if (a.capture) {
    TRY(ensure_words(a.off_dbg, max_g));
    a.captured = true;
  }
  uint32_t tile_bits = T <= 1 ? 1u : 32u - (uint32_t)__builtin_clz(T - 1u);
  uint32_t depth_bits = 32u - tile_bits;
  if (depth_bits > 28u) depth_bits = 28u;
  a.depth_bits = depth_bits;

  uint32_t* tab = P<uint32_t>(a.bk_tab);
  uint32_t* tile_start = tab;
  uint32_t* chunk_start = tab + (T + 1u);
  uint32_t* win_first = tab + 2u * (T + 1u);
  uint32_t* chunk_tile = win_first + win_slots;
  uint32_t* bctl = P<uint32_t>(a.bk_ctl);

  CgsBinArgs B;
  memset(&B, 0, sizeof(B));
  // cleared by k_count: the frame's scalars, and the bucket control words
  B.zero_at[0] = ctrl;
  B.zero_len[0] = (uint32_t)L.block_sums;
  B.zero_at[1] = bctl;
  B.zero_len[1] = (uint32_t)ctl_words;
  B.zero_at[2] = bctl;
  B.zero_len[2] = 0u;
  B.geom = a.geom_in ? a.geom_in : P<float>(a.geom);
  B.extents = a.extents_in ? a.extents_in : P<uint32_t>(a.extents);
  B.split_count = split_count;
  B.num_rest = num_rest;
  B.max_gaussians = max_g;
  B.width = g.W;
  B.height = g.H;
  B.tiles_x = g.tiles_x;
  B.tiles_y = g.tiles_y;
  B.tile_w = g.tile_w;
  B.tile_h = g.tile_h;
  B.mode64 = 0u;
  float np = near_plane < 1e-6f ? 1e-6f : near_plane;
  B.base_depth_bits = fbits(np) & 0x7F800000u;
  B.depth_bits = depth_bits;
  B.cap = cap;
  B.ids = P<uint32_t>(a.e_ids);  // not written in bucket mode
  B.counts = P<uint32_t>(a.cnt_dbg);
  B.offsets_out = a.capture ? P<uint32_t>(a.off_dbg) : nullptr;
  cgs_bin_masks(P<uint32_t>(a.hitm), max_g, &B.small_masks, &B.chunk_masks);
  B.total = ctrl + L.total;
  B.block_sums = ctrl + L.block_sums;
  B.block_surv = ctrl + L.block_surv;
  B.hist_partial = nullptr;
  B.sort_bits = tile_bits + depth_bits > 32u ? 32u : tile_bits + depth_bits;
  B.tile_bits = tile_bits;
  B.digit_bits = 8u;
  B.tile_hist = P<uint32_t>(a.bk_hist);
  B.tile_base = P<uint32_t>(a.bk_base);
  B.bucket = P<uint64_t>(a.bk_bucket);
  B.tiles_pad = tiles_pad;
  B.subs = subs;
  B.parts = parts;

  CgsBucketArgs K;
  K.hist = P<uint32_t>(a.bk_hist);
  K.rows = grid;
  K.tiles = T;
  K.tiles_pad = tiles_pad;
  K.cap = cap;
  K.big = CGS_SEG_WINDOW;
  K.window = CGS_SEG_WINDOW;
  K.base = P<uint32_t>(a.bk_base);
  K.tile_start = tile_start;
  K.chunk_start = chunk_start;
  K.ranges = P<uint32_t>(a.ranges);
  K.win_first = win_first;
  K.win_slots = win_slots;
  K.chunk_tile = chunk_tile;
  K.chunk_slots = chunk_slots;
  K.stats = bctl;
  K.ticket = bctl + 4;
  K.status = reinterpret_cast<uint64_t*>(bctl + 8);
  cgs_launch_bin(B, s, prof, &K);

  CgsSegSortArgs S;
  S.bucket = P<uint64_t>(a.bk_bucket);
  S.alt = P<uint64_t>(a.bk_alt);
  S.keys = P<uint32_t>(a.e_keys);
  S.ids = P<uint32_t>(a.e_ids);
  S.tile_start = tile_start;
  S.chunk_start = chunk_start;
  S.win_first = win_first;
  S.chunk_tile = chunk_tile;
  S.stats = bctl;
  S.tiles = T;
  S.cap = cap;
  S.window = CGS_SEG_WINDOW;
  S.windows = windows;
  S.big = CGS_SEG_WINDOW;
  S.chunk_slots = chunk_slots;
  // the previous frame's chunks with a quarter more room, and its merge rounds
  uint32_t cb = chunk_slots, rounds = 3u;
  if (a.bk_planned) {
    const uint64_t want = (uint64_t)a.bk_need_chunks + a.bk_need_chunks / 4u + 64u;
    cb = want < chunk_slots ? (uint32_t)want : chunk_slots;
    rounds = a.bk_need_rounds;
  }
  a.bk_chunk_blocks = cb;
  a.bk_rounds = rounds;
  cgs_launch_seg_sort(S, cb, rounds, s, prof);
  a.sorted_keys = P<uint32_t>(a.e_keys);
  a.sorted_ids = P<uint32_t>(a.e_ids);
  HIPCHK(hipGetLastError());
  return CGS_OK;
}

// Stage 3 on the projected splats in the arena (geom/extents, max_g slots,
// active = num_rest + *split_count).  Launch-only; the caller checks the
// total against the capacity after synchronising.
int launch_stage3(FrameArena& a, hipStream_t s, CgsProf* prof, const FrameGeom& g,
                  float near_plane, uint32_t num_rest, const uint32_t* split_count,
                  uint32_t max_g) {
  a.tiles_x = g.tiles_x;
  a.tiles_y = g.tiles_y;
  a.mode64 = g.T > 65535u;
  a.order = nullptr;  // set again below when this frame presorts
  init_entry_cap(a, max_g);
  const uint32_t cap = a.entry_cap;
  const bool bucket = bucket_mode(a, g.T, max_g);
  a.bucket = bucket;
  if (bucket && !a.plan_external) {
    // the plan this geometry's last frame left
    a.bk_key = ((uint64_t)(g.tile_w * 64u + g.tile_h) << 48) | ((uint64_t)g.T << 32) | max_g;
    a.bk_planned = false;
    for (const auto& p : a.bk_plans)
      if (p.first == a.bk_key) {
        a.bk_planned = true;
        a.bk_need_chunks = p.second.first;
        a.bk_need_rounds = p.second.second;
      }
  }
  if (bucket) return launch_stage3_bucket(a, s, prof, g, near_plane, num_rest, split_count, max_g);
  TRY(ensure_words(a.e_keys, cap));
  TRY(ensure_words(a.e_ids, cap));
  TRY(ensure_words(a.e_alt_k, cap));
  TRY(ensure_words(a.e_alt_v, cap));
  const bool presort = a.tune.tile_sort == 2u;
  if (a.mode64 && !presort) {
    TRY(ensure_words(a.e_depths, cap));
    TRY(ensure_words(a.e_perm, cap));
  }
  TRY(ensure_words(a.ranges, (size_t)g.T * 2));
  if (presort) TRY(size_presort(a, max_g));
  const FrameLayout L = frame_layout(max_g, cap);
  TRY(ensure_words(a.ctrl3, L.words));
  uint32_t* ctrl = P<uint32_t>(a.ctrl3);
  // to clear: the frame's scalars, sort tickets and digit histograms (up to
  // the per-block sums, which k_count writes whole), the sort's look-back
  // status (the second set only for 64-bit keys) and the tile ranges.  k_count
  // clears them; without splats there is no k_count and they are filled here.
  uint32_t* zero_at[3] = {ctrl, ctrl + L.rA_status, P<uint32_t>(a.ranges)};
  const uint32_t zero_len[3] = {(uint32_t)L.block_sums,
                                (uint32_t)((a.mode64 ? 2u : 1u) * (L.rB_status - L.rA_status)),
                                g.T * 2u};
  if (cgs_bin_num_parts(max_g) == 0)
    for (int z = 0; z < 3; ++z) cgs_launch_fill_u32(zero_at[z], 0u, zero_len[z], s);
  TRY(ensure_words(a.cnt_dbg, max_g));  // per-splat tile counts (always needed)
  TRY(ensure_words(a.hitm, cgs_bin_mask_words(max_g)));
  a.captured = false;
  if (a.capture) {
    TRY(ensure_words(a.off_dbg, max_g));
    a.captured = true;
  }

  uint32_t tile_bits, depth_bits;
  tile_bits = g.T <= 1 ? 1u : 32u - (uint32_t)__builtin_clz(g.T - 1u);
  depth_bits = 32u - tile_bits;
  if (depth_bits > 28u) depth_bits = 28u;
  a.depth_bits = depth_bits;

  CgsBinArgs B;
  memset(&B, 0, sizeof(B));
  for (int z = 0; z < 3; ++z) {
    B.zero_at[z] = zero_at[z];
    B.zero_len[z] = zero_len[z];
  }
  B.geom = a.geom_in ? a.geom_in : P<float>(a.geom);
  B.extents = a.extents_in ? a.extents_in : P<uint32_t>(a.extents);
  B.split_count = split_count;
  B.num_rest = num_rest;
  B.max_gaussians = max_g;
  B.width = g.W;
  B.height = g.H;
  B.tiles_x = g.tiles_x;
  B.tiles_y = g.tiles_y;
  B.tile_w = g.tile_w;
  B.tile_h = g.tile_h;
  B.mode64 = a.mode64 ? 1u : 0u;
  float np = near_plane < 1e-6f ? 1e-6f : near_plane;
  B.base_depth_bits = fbits(np) & 0x7F800000u;
  B.depth_bits = depth_bits;
  B.cap = cap;
  B.keys = P<uint32_t>(a.e_keys);
  B.depths = a.mode64 && !presort ? P<uint32_t>(a.e_depths) : nullptr;
  B.ids = P<uint32_t>(a.e_ids);
  B.counts = P<uint32_t>(a.cnt_dbg);
  B.offsets_out = a.capture ? P<uint32_t>(a.off_dbg) : nullptr;
  cgs_bin_masks(P<uint32_t>(a.hitm), max_g, &B.small_masks, &B.chunk_masks);
  B.total = ctrl + L.total;
  B.block_sums = ctrl + L.block_sums;
  B.block_surv = ctrl + L.block_surv;
  B.hist_partial = ctrl + L.bin_hist;
  const uint32_t tile_passes = (tile_bits + 7u) / 8u;
  B.hist_words = presort ? 256u * tile_passes : a.mode64 ? 2048u : 1024u;
  uint32_t sort_bits = tile_bits + depth_bits;
  if (sort_bits > 32) sort_bits = 32;
  B.sort_bits = sort_bits;
  B.tile_bits = tile_bits;
  B.subs = 1u;
  B.parts = cgs_bin_num_parts(max_g);
  const uint32_t db = a.mode64 || presort ? 8u : a.tune.radix_digit_bits;
  B.digit_bits = db;
  a.order = nullptr;
  if (presort && max_g != 0) {
    // the splats sorted by their keys' depth bits (stable: ties by index)
    CgsRadixBuffers SR;
    SR.keys = P<uint32_t>(a.sp_keys);
    SR.vals = P<uint32_t>(a.sp_vals);
    SR.keys_alt = P<uint32_t>(a.sp_alt_k);
    SR.vals_alt = P<uint32_t>(a.sp_alt_v);
    uint32_t* sc = P<uint32_t>(a.sp_ctrl);
    const size_t parts = cgs_radix_num_parts(max_g) ? cgs_radix_num_parts(max_g) : 1u;
    const size_t status_words = 4 * parts * 256;
    // layout: hist[1024] | status | tickets[4] | partial histograms
    SR.hist = sc;
    SR.status = sc + 1024;
    SR.tickets = sc + 1024 + status_words;
    SR.hist_partial = sc + 1024 + status_words + 4;
    SR.cap = max_g;
    const uint32_t key_bits = a.mode64 ? 32u : depth_bits;
    prof->mark("k_splat_depth_keys", s);
    const uint32_t rows = cgs_launch_splat_depth_keys(
        B, SR.keys, SR.hist, (uint32_t)(1024 + status_words + 4), key_bits, SR.hist_partial,
        cgs_radix_hist_blocks(), s);
    // its launches are timed together (the entries' sort keeps its own names)
    prof->mark("k_presort", s);
    cgs_launch_hist_reduce(SR.hist_partial, rows, 1024u, 256u * ((key_bits + 7u) / 8u), SR.hist, s,
                           true);
    // partitions of 4096 splats up to 4M slots: enough blocks to fill the
    // chip, where the entries' partitions would leave most of it idle
    const uint32_t pk = max_g <= (1u << 22) ? 4096u : a.tune.radix_partition_keys;
    const int alt = cgs_launch_radix_sort(SR, nullptr, max_g, key_bits, true, true, s, nullptr,
                                          pk, 8u);
    B.order = alt ? SR.vals_alt : SR.vals;
    a.order = B.order;
  }
  cgs_launch_bin(B, s, prof);
  // digit histograms of every pass, summed over the binning partitions
  const uint32_t hist_bins = presort     ? 256u * tile_passes
                             : a.mode64 ? 1024u + 256u * tile_passes
                                        : (1u << db) * ((sort_bits + db - 1u) / db);
  prof->mark("k_hist_reduce", s);
  cgs_launch_hist_reduce(ctrl + L.bin_hist, L.bin_parts, B.hist_words, hist_bins, ctrl + L.histA,
                         s, true);  // cleared with the frame's control words

  const uint32_t* d_n = ctrl + L.total;
  if (!a.mode64) {
    CgsRadixBuffers R;
    R.keys = P<uint32_t>(a.e_keys);
    R.vals = P<uint32_t>(a.e_ids);
    R.keys_alt = P<uint32_t>(a.e_alt_k);
    R.vals_alt = P<uint32_t>(a.e_alt_v);
    R.hist = ctrl + L.histA;
    R.hist_partial = nullptr;
    R.status = ctrl + L.rA_status;
    R.tickets = ctrl + L.rA_tickets;
    R.cap = cap;
    // presort: the entries are in (depth, splat) order already; the tile bits
    // above depth_bits are left
    int alt = cgs_launch_radix_sort(R, d_n, 0, sort_bits, false, true, s, prof,
                                    a.tune.radix_partition_keys, db, presort ? depth_bits : 0u);
    a.sorted_keys = alt ? R.keys_alt : R.keys;
    a.sorted_ids = alt ? R.vals_alt : R.vals;
    prof->mark("k_tile_ranges", s);
    cgs_launch_tile_ranges(a.sorted_keys, d_n, cap, depth_bits, g.T, P<uint32_t>(a.ranges),
                           s);
  } else if (presort) {
    // 64-bit keys of presorted splats: the entries are in (depth, splat) order
    // already, so one stable sort of (tile id, splat id) by the tile bits
    CgsRadixBuffers Bt;
    Bt.keys = P<uint32_t>(a.e_keys);
    Bt.vals = P<uint32_t>(a.e_ids);
    Bt.keys_alt = P<uint32_t>(a.e_alt_k);
    Bt.vals_alt = P<uint32_t>(a.e_alt_v);
    Bt.hist = ctrl + L.histA;
    Bt.hist_partial = nullptr;
    Bt.status = ctrl + L.rA_status;
    Bt.tickets = ctrl + L.rA_tickets;
    Bt.cap = cap;
    int altB = cgs_launch_radix_sort(Bt, d_n, 0, tile_bits, false, true, s, prof,
                                     a.tune.radix_partition_keys, 8u);
    a.sorted_keys = altB ? Bt.keys_alt : Bt.keys;
    a.sorted_ids = altB ? Bt.vals_alt : Bt.vals;
    prof->mark("k_tile_ranges", s);
    cgs_launch_tile_ranges(a.sorted_keys, d_n, cap, 0, g.T, P<uint32_t>(a.ranges), s);
  } else {
    // 64-bit fallback (gpu_radix_sort_u64_pairs, radix_sort.rs:377-477): stable
    // sort by depth with the entry index as payload, gather, stable sort by tile
    CgsRadixBuffers A;
    A.keys = P<uint32_t>(a.e_depths);
    A.vals = P<uint32_t>(a.e_perm);
    A.keys_alt = P<uint32_t>(a.e_alt_k);
    A.vals_alt = P<uint32_t>(a.e_alt_v);
    A.hist = ctrl + L.histA;
    A.hist_partial = nullptr;
    A.status = ctrl + L.rA_status;
    A.tickets = ctrl + L.rA_tickets;
    A.cap = cap;
    int altA = cgs_launch_radix_sort(A, d_n, 0, 32, true, true, s, prof,
                                     a.tune.radix_partition_keys, 8u);
    const uint32_t* perm = altA ? A.vals_alt : A.vals;
    uint32_t* gk = altA ? P<uint32_t>(a.e_depths) : P<uint32_t>(a.e_alt_k);
    uint32_t* gv = altA ? P<uint32_t>(a.e_perm) : P<uint32_t>(a.e_alt_v);
    // with 4 passes altA == 0: perm in e_perm, gather into the alt pair
    prof->mark("k_gather2", s);
    if (!cgs_launch_gather2(perm, P<uint32_t>(a.e_keys), P<uint32_t>(a.e_ids), gk, gv, d_n, cap,
                            s))
      return fail(CGS_ERR_INVALID_ARG, "internal: gather buffers not 16-byte aligned");
    CgsRadixBuffers Bt;
    Bt.keys = gk;
    Bt.vals = gv;
    Bt.keys_alt = P<uint32_t>(a.e_keys);
    Bt.vals_alt = P<uint32_t>(a.e_ids);
    Bt.hist = ctrl + L.histB;
    Bt.hist_partial = nullptr;
    Bt.status = ctrl + L.rB_status;
    Bt.tickets = ctrl + L.rB_tickets;
    Bt.cap = cap;
    int altB = cgs_launch_radix_sort(Bt, d_n, 0, tile_bits, false, true, s, prof,
                                     a.tune.radix_partition_keys, 8u);
    a.sorted_keys = altB ? Bt.keys_alt : Bt.keys;
    a.sorted_ids = altB ? Bt.vals_alt : Bt.vals;
    prof->mark("k_tile_ranges", s);
    cgs_launch_tile_ranges(a.sorted_keys, d_n, cap, 0, g.T, P<uint32_t>(a.ranges), s);
  }
  HIPCHK(hipGetLastError());
  return CGS_OK;
}

// Stage 4 into out (padded rows of g.padded_row words), or into the arena's
// framebuffer when out is null
int launch_stage4(FrameArena& a, hipStream_t s, CgsProf* prof, const FrameGeom& g,
                  const cgs_compositor_uniforms* comp, uint32_t* out = nullptr) {
  if (!out) {
    TRY(ensure_words(a.out_rgba, (size_t)g.padded_row * g.H));
    out = P<uint32_t>(a.out_rgba);
  }
  a.padded_row = g.padded_row;
  CgsCompositeArgs A;
  A.geom = a.geom_in ? a.geom_in : P<float>(a.geom);
  A.color = a.color_in ? a.color_in : P<float>(a.color);
  A.sorted_ids = a.sorted_ids ? a.sorted_ids : P<uint32_t>(a.e_ids);
  A.ranges = P<uint32_t>(a.ranges);
  A.width = g.W;
  A.height = g.H;
  A.tiles_x = g.tiles_x;
  A.tiles_y = g.tiles_y;
  A.tile_w = g.tile_w;
  A.tile_h = g.tile_h;
  A.padded_row = g.padded_row;
  for (int i = 0; i < 4; ++i) A.bg[i] = comp->background_color[i];
  A.out = out;
  TRY(ensure_words(a.comp_order, g.T));
  A.order = a.tune.tile_order ? P<uint32_t>(a.comp_order) : nullptr;
  cgs_launch_composite(A, s, prof, a.tune.composite_batch);
  HIPCHK(hipGetLastError());
  return CGS_OK;
}

// After a Stage 3 in bucket mode: records the chunks and merge rounds the
// frame needed (stats = the bucket control words read back); true when it
// needed more than were launched, so that its Stage 3 must run again.
bool bucket_rerun(FrameArena& a, const uint32_t* stats) {
  if (!a.bucket) return false;
  const bool more = stats[3] > a.bk_chunk_blocks || stats[1] > a.bk_rounds;
  a.bk_need_chunks = stats[3];
  a.bk_need_rounds = stats[1];
  a.bk_planned = true;
  if (!a.plan_external) {
    bool found = false;
    for (auto& p : a.bk_plans)
      if (p.first == a.bk_key) {
        p.second = {stats[3], stats[1]};
        found = true;
      }
    if (!found) {
      if (a.bk_plans.size() >= 64) a.bk_plans.erase(a.bk_plans.begin());
      a.bk_plans.push_back({a.bk_key, {stats[3], stats[1]}});
    }
  }
  return more;
}

// read the entry total, the survivors and the split count (one small copy + wait):
// h_counters[2] = entries, [3] = survivors, [4] = split children, [12..16) the
// bucket mode's stats
int read_counters(cgs_context* c, bool with_split) {
  const uint32_t* ctrl = P<uint32_t>(c->fa.ctrl3);
  HIPCHK(hipMemcpyAsync(c->h_counters, ctrl, 4 * 4, hipMemcpyDeviceToHost, c->stream));
  if (c->fa.bucket)
    HIPCHK(hipMemcpyAsync(c->h_counters + 12, c->fa.bk_ctl.p, 4 * 4, hipMemcpyDeviceToHost,
                          c->stream));
  if (with_split && c->ctrl1.p)
    HIPCHK(hipMemcpyAsync(c->h_counters + 4, c->ctrl1.p, 4, hipMemcpyDeviceToHost, c->stream));
  else
    c->h_counters[4] = 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
  return ms;
}


// ---------------------------------------------------------------------------
// frame pieces shared by the single-frame path and the view batch
// ---------------------------------------------------------------------------

// the scene side of one frame after Stage 1: what Stages 2-3 read
struct SceneFrame {
  CgsSceneDev S;
  const uint32_t* split_count = nullptr;  // device word: split children, or null
  uint32_t N = 0, max_g = 0;
  bool split_possible = false;
};

// Stage 1 (gpu_state.rs:706-713, 852-894): the scene is re-deformed when the
// cage vertices or (tau, mip) changed since the last deformation; fills the
// scene pointers Stage 2 reads.  deformed = Stage 1 ran.
int stage1(cgs_context* c, const cgs_frame_params* prm, hipStream_t s, CgsProf* prof,
           SceneFrame& F, bool& deformed) {
  deformed = false;
  const uint32_t N = c->n_rest;
  F.N = N;
  F.split_possible = c->has_cages && prm->strain_split_threshold > 0.0f;
  F.max_g = F.split_possible ? 2 * N : N;
  memset(&F.S, 0, sizeof(F.S));
  F.split_count = nullptr;
  if (c->has_cages) {
    if (c->f16) return fail(CGS_ERR_UNSUPPORTED, "cage deformation of the f16 storage variant");
    const uint32_t p0 = fbits(prm->strain_split_threshold), p1 = fbits(prm->mip_filter_3d_var);
    if (p0 != c->last_params[0] || p1 != c->last_params[1]) {
      c->last_params[0] = p0;
      c->last_params[1] = p1;
      c->deformed_valid = false;
    }
    TRY(ensure_words(c->jac, (size_t)c->num_tets * 9));
    TRY(ensure_words(c->rot, (size_t)c->num_tets * 9));
    TRY(ensure_words(c->stretch, c->num_tets));
    TRY(ensure_words(c->tet_pack, (size_t)c->num_tets * CGS_TET_PACK));
    if (c->def_cap < F.max_g) {
      c->def_cap = F.max_g;
      c->deformed_valid = false;
    }
    TRY(ensure_words(c->def_geo, cgs_blk_words(c->def_cap, CGS_GEO_K)));
    TRY(ensure_words(c->split_meta, (size_t)2 * N));
    const uint32_t dparts = cgs_deform_num_parts(N);
    TRY(ensure_words(c->ctrl1, 4 + 2 * (size_t)(dparts ? dparts : 1)));
    if (!c->deformed_valid) {
      cgs_launch_fill_u32(P<uint32_t>(c->ctrl1), 0u, 4 + 2 * (size_t)(dparts ? dparts : 1), s);
      CgsTetArgs TA;
      TA.verts = P<float>(c->verts);
      TA.elems = P<uint32_t>(c->elems);
      TA.inv_dm = P<float>(c->inv_dm);
      TA.jac = P<float>(c->jac);
      TA.rot = P<float>(c->rot);
      TA.stretch = P<float>(c->stretch);
      TA.pack = P<float>(c->tet_pack);
      TA.num_tets = c->num_tets;
      prof->mark("k_tet_transforms", s);
      cgs_launch_tet_transforms(TA, s);
      CgsDeformArgs D;
      D.verts = P<float>(c->verts);
      D.elems = P<uint32_t>(c->elems);
      D.tet_pack = P<float>(c->tet_pack);
      D.bind_tet = P<uint32_t>(c->bind_tet);
      D.bind_w = P<float>(c->bind_w);
      D.rest_geo = P<float>(c->rest_geo);
      D.out_geo = P<float>(c->def_geo);
      D.split_meta = P<uint32_t>(c->split_meta);
      D.split_count = P<uint32_t>(c->ctrl1);
      D.status = reinterpret_cast<uint64_t*>(P<uint32_t>(c->ctrl1) + 4);
      D.ticket = P<uint32_t>(c->ctrl1) + 1;
      D.split_threshold = F.split_possible ? prm->strain_split_threshold : 0.0f;
      D.mip = prm->mip_filter_3d_var;
      D.num_gaussians = N;
      D.num_tets = c->num_tets;
      prof->mark("k_deform", s);
      if (!cgs_launch_deform(D, s))
        return fail(CGS_ERR_INVALID_ARG, "internal: tet records or weights not 16-byte aligned");
      HIPCHK(hipGetLastError());
      c->deformed_valid = true;
      c->deform_generation++;
      deformed = true;
    }
    F.S.geo = P<float>(c->def_geo);
    F.split_count = P<uint32_t>(c->ctrl1);
  } else {
    F.S.geo = P<float>(c->rest_geo);
    if (c->f16) F.S.cov_h = P<uint16_t>(c->rest_cov_h);
  }
  if (c->f16)
    F.S.sh_h = P<uint16_t>(c->rest_sh_h);
  else
    F.S.sh = P<float>(c->rest_sh);
  return CGS_OK;
}

// the Stage 3 cache key fields (gpu_state.rs:956-981)
Stage3Key stage3_key(const cgs_context* c, const cgs_camera_uniforms* cam, const FrameGeom& g,
                     const cgs_frame_params* prm) {
  Stage3Key key;
  memset(&key, 0, sizeof(key));
  key.deform_generation = c->deform_generation;
  for (int i = 0; i < 16; ++i) key.view_bits[i] = fbits(cam->view_matrix[i]);
  key.fx = fbits(cam->focal_x);
  key.fy = fbits(cam->focal_y);
  key.cx = fbits(cam->principal_x);
  key.cy = fbits(cam->principal_y);
  key.width = g.W;
  key.height = g.H;
  key.tile_w = g.tile_w;
  key.tile_h = g.tile_h;
  key.near_bits = fbits(cam->near_plane);
  key.mip_bits = fbits(prm->mip_filter_3d_var);
  key.min_op_bits = fbits(prm->min_opacity);
  key.min_rad_bits = fbits(prm->min_radius_px);
  return key;
}

// The Stage 2 arguments of the scene side of a frame (project_v2 dispatch,
// gpu_state.rs:882-948); camera and outputs are filled by the caller
int project_args(const cgs_context* c, const SceneFrame& F, const cgs_compositor_uniforms* comp,
                 const cgs_frame_params* prm, CgsProjectArgs& PA) {
  memset(&PA, 0, sizeof(PA));
  PA.scene = F.S;
  PA.bind_tet = c->has_cages ? P<uint32_t>(c->bind_tet) : nullptr;
  PA.tet_rot = c->has_cages ? P<float>(c->rot) : nullptr;
  PA.split_meta = c->has_cages ? P<uint32_t>(c->split_meta) : nullptr;
  PA.split_count = F.split_count;
  PA.mip = c->has_cages ? 0.0f : prm->mip_filter_3d_var;  // gpu_state.rs:894
  PA.min_opacity = prm->min_opacity;
  PA.min_radius = prm->min_radius_px;
  PA.sh_degree = comp->sh_degree > 3 ? 3 : comp->sh_degree;
  if (c->f16 && PA.sh_degree != c->f16_degree)
    return fail(CGS_ERR_INVALID_ARG, "sh_degree must equal the quantized scene's degree");
  PA.has_deformation = c->has_cages ? 1u : 0u;
  PA.num_tets = c->num_tets;
  PA.num_rest = F.N;
  PA.max_gaussians = F.max_g;
  return CGS_OK;
}

// Stage 2 of one view into the arena a
int stage2(const cgs_context* c, const SceneFrame& F, FrameArena& a, hipStream_t s,
           CgsProf* prof, const cgs_camera_uniforms* cam, const cgs_compositor_uniforms* comp,
           const cgs_frame_params* prm, bool zero_culled = false) {
  const uint32_t max_g = F.max_g;
  TRY(ensure_words(a.geom, (size_t)8 * max_g));
  TRY(ensure_words(a.color, (size_t)4 * max_g));
  TRY(ensure_words(a.extents, max_g));
  init_entry_cap(a, max_g);
  const FrameLayout L = frame_layout(max_g, a.entry_cap);
  TRY(ensure_words(a.ctrl3, L.words));
  CgsProjectArgs PA;
  TRY(project_args(c, F, comp, prm, PA));
  memcpy(PA.camera, cam, 96);
  for (int i = 0; i < 3; ++i) PA.cam_pos[i] = comp->camera_position[i];
  PA.geom = P<float>(a.geom);
  PA.color = P<float>(a.color);
  PA.extents = P<uint32_t>(a.extents);
  PA.zero_culled = zero_culled ? 1u : 0u;
  prof->mark("k_project", s);
  cgs_launch_project(PA, s);
  HIPCHK(hipGetLastError());
  return CGS_OK;
}

int check_tuning(const cgs_tuning* t) {
  if (t->radix_partition_keys != 16384u && t->radix_partition_keys != 12288u &&
      t->radix_partition_keys != 10240u && t->radix_partition_keys != 8192u &&
      t->radix_partition_keys != 4096u)
    return fail(CGS_ERR_INVALID_ARG,
                "radix_partition_keys must be 16384, 12288, 10240, 8192 or 4096");
  if (t->radix_digit_bits != 8u && t->radix_digit_bits != 4u)
    return fail(CGS_ERR_INVALID_ARG, "radix_digit_bits must be 8 or 4");
  if (t->composite_batch != 64u && t->composite_batch != 128u && t->composite_batch != 256u &&
      t->composite_batch != 512u)
    return fail(CGS_ERR_INVALID_ARG, "composite_batch must be 64, 128, 256 or 512");
  if (t->composite_batch == 512u && !(t->tile_w == 16u && t->tile_h == 16u))
    return fail(CGS_ERR_INVALID_ARG, "composite_batch 512 needs 16x16 tiles");
  if (t->tile_order > 1u) return fail(CGS_ERR_INVALID_ARG, "tile_order must be 0 or 1");
  if (t->tile_sort > 2u) return fail(CGS_ERR_INVALID_ARG, "tile_sort must be 0, 1 or 2");
  const uint32_t tw = t->tile_w, th = t->tile_h;
  if (!((tw == 16u && th == 16u) || (tw == 8u && th == 8u) || (tw == 16u && th == 32u) ||
        (tw == 32u && th == 16u)))
    return fail(CGS_ERR_INVALID_ARG, "tile size must be 16x16, 8x8, 16x32 or 32x16");
  // an 8 x 8 tile is one wave, which stages 64 splats per step
  if (tw == 8u && t->composite_batch != 64u)
    return fail(CGS_ERR_INVALID_ARG, "8x8 tiles take composite_batch 64");
  return CGS_OK;
}

// the arena of an empty scene's frame: no tile has entries
int background_arena(FrameArena& a, const FrameGeom& g, hipStream_t s) {
  a.order = nullptr;
  TRY(ensure_words(a.ranges, (size_t)g.T * 2));
  cgs_launch_fill_u32(P<uint32_t>(a.ranges), 0u, (uint64_t)g.T * 2, s);
  TRY(ensure_words(a.geom, 8));
  TRY(ensure_words(a.color, 4));
  TRY(ensure_words(a.e_ids, 1));
  a.sorted_ids = P<uint32_t>(a.e_ids);
  return CGS_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char* cgs_last_error(void) { return g_err.c_str(); }

int cgs_create(int device, cgs_context** out) {
  if (!out) return fail(CGS_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(CGS_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(CGS_ERR_INVALID_ARG, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (prop.warpSize != 64)
    return fail(CGS_ERR_UNSUPPORTED, std::string("wave64 device required, got ") + prop.gcnArchName);
  cgs_context* c = new cgs_context();
  c->device = device;
  snprintf(c->name, sizeof(c->name), "%s (%s)", prop.name, prop.gcnArchName);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(CGS_ERR_RUNTIME, "stream creation failed");
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  for (auto& e : c->prof.ev) (void)hipEventCreate(&e);
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_counters), 16 * 4, hipHostMallocDefault) != hipSuccess) {
    delete c;
    return fail(CGS_ERR_RUNTIME, "pinned allocation failed");
  }
  memset(c->h_counters, 0, 64);
  const int rs = scratch_for_device(device, &c->scan_scratch);
  if (rs != CGS_OK) {
    (void)hipHostFree(c->h_counters);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rs;
  }
  *out = c;
  return CGS_OK;
}

void cgs_destroy(cgs_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->rest_geo, &c->rest_cov_h, &c->rest_sh, &c->rest_sh_h, &c->staging,
                    &c->elems, &c->inv_dm, &c->bind_tet, &c->bind_w, &c->verts, &c->jac, &c->rot,
                    &c->stretch, &c->tet_pack, &c->def_geo, &c->split_meta, &c->ctrl1, &c->prim_ctrl,
                    &c->prim_alt_k, &c->prim_alt_v, &c->prim_alt_d, &c->mvc_work,
                    &c->ingest, &c->tet_shm};
  for (DevBuf* b : bufs) release(*b);
  near_cache_free(c->near_cache);
  c->near_cache = nullptr;
  c->fa.release_all();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->prof.ev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_counters) (void)hipHostFree(c->h_counters);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int cgs_device_name(cgs_context* c, char* buf, size_t len) {
  if (!c || !buf || len == 0) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  snprintf(buf, len, "%s", c->name);
  return CGS_OK;
}

int cgs_synchronize(cgs_context* c) {
  TRY(ensure_ctx(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_invalidate(cgs_context* c) {
  if (!c) return fail(CGS_ERR_INVALID_ARG, "null context");
  c->scene_known = false;
  c->topo_hash = 0;
  c->last_verts.clear();
  c->deformed_valid = false;
  c->deform_generation++;
  c->s3_valid = false;
  c->sort_cache.invalidate();
  return CGS_OK;
}

int cgs_invalidate_stage3(cgs_context* c) {
  if (!c) return fail(CGS_ERR_INVALID_ARG, "null context");
  c->s3_valid = false;
  c->sort_cache.invalidate();
  return CGS_OK;
}

int cgs_set_sort_hysteresis(cgs_context* c, uint32_t enable) {
  if (!c) return fail(CGS_ERR_INVALID_ARG, "null context");
  c->hysteresis = enable != 0;
  c->sort_cache = cgs_sort_cache();
  c->s3_valid = false;
  return CGS_OK;
}

int cgs_sort_hysteresis_reuse_count(cgs_context* c, uint64_t* count) {
  if (!c || !count) return fail(CGS_ERR_INVALID_ARG, "null argument");
  *count = c->sort_cache.reuse_count;
  return CGS_OK;
}

int cgs_sync_scene(cgs_context* c, const float* pos, const float* cov, const float* opa,
                   const float* sh, uint64_t count) {
  TRY(ensure_ctx(c));
  if (count > 0 && (!pos || !cov || !opa || !sh))
    return fail(CGS_ERR_INVALID_ARG, "null scene array");
  if (count >= (1ull << 28)) return fail(CGS_ERR_INVALID_ARG, "scene too large (>= 2^28 splats)");
  const uint64_t h = scene_fingerprint(pos, cov, opa, sh, count, 6, 48);
  if (c->scene_known && !c->f16 && c->scene_hash == h && c->n_rest == count) return CGS_OK;
  const uint32_t n = (uint32_t)count;
  TRY(upload_blocked_u32(c, c->rest_geo, pos, n, 3, CGS_GEO_K, CGS_CH_X));
  TRY(upload_blocked_u32(c, c->rest_geo, cov, n, 6, CGS_GEO_K, CGS_CH_COV));
  TRY(upload_blocked_u32(c, c->rest_geo, opa, n, 1, CGS_GEO_K, CGS_CH_OPA));
  TRY(upload_blocked_u32(c, c->rest_sh, sh, n, 48, CGS_SH_K, 0));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->n_rest = n;
  c->f16 = false;
  c->scene_hash = h;
  c->scene_known = true;
  scene_changed(c);
  return CGS_OK;
}

int cgs_sync_scene_f16(cgs_context* c, const float* pos, const uint32_t* cov_packed,
                       const float* opa, const uint32_t* sh_packed, uint32_t sh_degree,
                       uint64_t count) {
  TRY(ensure_ctx(c));
  if (count > 0 && (!pos || !cov_packed || !opa || !sh_packed))
    return fail(CGS_ERR_INVALID_ARG, "null scene array");
  if (count >= (1ull << 28)) return fail(CGS_ERR_INVALID_ARG, "scene too large (>= 2^28 splats)");
  if (sh_degree > 3) sh_degree = 3;
  const uint32_t bases = sh_degree == 0 ? 1 : sh_degree == 1 ? 4 : sh_degree == 2 ? 9 : 16;
  const uint32_t sh_words = (bases * 3 + 1) / 2;
  const uint64_t h = scene_fingerprint(pos, reinterpret_cast<const float*>(cov_packed), opa,
                                       reinterpret_cast<const float*>(sh_packed), count, 3,
                                       sh_words) ^ (0x9e3779b97f4a7c15ull * (sh_degree + 1));
  if (c->scene_known && c->f16 && c->scene_hash == h && c->n_rest == count) return CGS_OK;
  const uint32_t n = (uint32_t)count;
  TRY(upload_blocked_u32(c, c->rest_geo, pos, n, 3, CGS_GEO_K, CGS_CH_X));
  TRY(upload_blocked_u32(c, c->rest_geo, opa, n, 1, CGS_GEO_K, CGS_CH_OPA));
  // covariances: 6 halves per splat -> 6 blocked half channels
  TRY(ensure_words(c->rest_cov_h, (cgs_blk_words(n, 6) + 1) / 2));
  TRY(ensure_words(c->rest_sh_h, (cgs_blk_words(n, CGS_SH_K) + 1) / 2));
  if (n) {
    TRY(ensure_words(c->staging, (size_t)n * (sh_words > 3 ? sh_words : 3)));
    HIPCHK(hipMemcpyAsync(c->staging.p, cov_packed, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    cgs_launch_to_blocked_u16(P<uint16_t>(c->staging), P<uint16_t>(c->rest_cov_h), n, 6, 6, 0,
                              c->stream);
    HIPCHK(hipMemcpyAsync(c->staging.p, sh_packed, (size_t)n * sh_words * 4, hipMemcpyHostToDevice,
                          c->stream));
    cgs_launch_sh_f16_blocked(P<uint16_t>(c->staging), P<uint16_t>(c->rest_sh_h), n, sh_words * 2,
                              bases - 1, c->stream);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->n_rest = n;
  c->f16 = true;
  c->f16_degree = sh_degree;
  c->scene_hash = h;
  c->scene_known = true;
  scene_changed(c);
  return CGS_OK;
}

int cgs_sync_cages(cgs_context* c, const uint32_t* elems, const float* inv_dm, uint32_t num_tets,
                   const float* verts, uint32_t num_verts, const uint32_t* bind_tet,
                   const float* bind_w, uint64_t count) {
  TRY(ensure_ctx(c));
  if (num_tets == 0) {
    // no active cages (gpu_state.rs:215-229)
    const bool was = c->has_cages;
    c->has_cages = false;
    c->num_tets = 0;
    c->num_verts = 0;
    c->topo_hash = 0;
    c->last_verts.clear();
    c->deformed_valid = !was;
    if (was) c->deform_generation++;
    return CGS_OK;
  }
  if (!elems || !inv_dm || !verts || (count && (!bind_tet || !bind_w)))
    return fail(CGS_ERR_INVALID_ARG, "null cage array");
  if (count != c->n_rest) return fail(CGS_ERR_INVALID_ARG, "binding count != scene count");
  for (uint64_t i = 0; i < (uint64_t)num_tets * 4; ++i)
    if (elems[i] >= num_verts) return fail(CGS_ERR_INVALID_ARG, "tet element index out of range");
  uint64_t th = FNV_OFFSET;
  th = fnv_u64(th, num_tets);
  th = fnv_u64(th, num_verts);
  th = fnv_u64(th, (uint64_t)(uintptr_t)elems);
  th = fnv_u64(th, (uint64_t)(uintptr_t)inv_dm);
  th = fnv_u64(th, (uint64_t)(uintptr_t)bind_tet);
  th = fnv_u64(th, count);
  th = hash_sample(th, elems, (size_t)num_tets * 4);
  th = hash_sample(th, reinterpret_cast<const uint32_t*>(inv_dm), (size_t)num_tets * 9);
  th = hash_sample(th, bind_tet, count);
  th = hash_sample(th, reinterpret_cast<const uint32_t*>(bind_w), count * 4);
  if (th == 0) th = 1;
  if (!c->has_cages || c->topo_hash != th) {
    TRY(ensure_words(c->elems, (size_t)num_tets * 4));
    TRY(ensure_words(c->inv_dm, (size_t)num_tets * 9));
    TRY(ensure_words(c->bind_tet, count));
    TRY(ensure_words(c->bind_w, count * 4));
    HIPCHK(hipMemcpyAsync(c->elems.p, elems, (size_t)num_tets * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->inv_dm.p, inv_dm, (size_t)num_tets * 36, hipMemcpyHostToDevice, c->stream));
    if (count) {
      HIPCHK(hipMemcpyAsync(c->bind_tet.p, bind_tet, count * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->bind_w.p, bind_w, count * 16, hipMemcpyHostToDevice, c->stream));
    }
    c->num_tets = num_tets;
    c->has_cages = true;
    c->topo_hash = th;
    c->last_verts.clear();
    c->deformed_valid = false;
  }
  const size_t nv = (size_t)num_verts * 3;
  const bool changed = c->num_verts != num_verts || c->last_verts.size() != nv ||
                       memcmp(c->last_verts.data(), verts, nv * 4) != 0;
  if (changed) {
    TRY(ensure_words(c->verts, nv));
    HIPCHK(hipMemcpyAsync(c->verts.p, verts, nv * 4, hipMemcpyHostToDevice, c->stream));
    c->last_verts.assign(verts, verts + nv);
    c->num_verts = num_verts;
    c->deformed_valid = false;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_set_capture(cgs_context* c, uint32_t enable) {
  if (!c) return fail(CGS_ERR_INVALID_ARG, "null context");
  c->fa.capture = enable != 0;
  return CGS_OK;
}

int cgs_execute_frame(cgs_context* c, const cgs_camera_uniforms* cam,
                      const cgs_compositor_uniforms* comp, const cgs_frame_params* prm,
                      cgs_stage_timings* tim) {
  TRY(ensure_ctx(c));
  if (!cam || !comp || !prm) return fail(CGS_ERR_INVALID_ARG, "null frame argument");
  cgs_stage_timings T;
  memset(&T, 0, sizeof(T));
  const FrameGeom g = frame_geom(comp, c->fa.tune);
  if (g.W != c->last_w || g.H != c->last_h) {
    c->s3_valid = false;
    c->last_w = g.W;
    c->last_h = g.H;
  }
  const bool prof = prm->profile_stages != 0;
  hipStream_t s = c->stream;
  c->prof.on = prm->profile_stages >= 2;
  c->prof.n = 0;
  c->prof_count = 0;

  if (c->n_rest == 0) {
    // background frame (gpu_state.rs:658-704)
    TRY(background_arena(c->fa, g, s));
    if (prof) HIPCHK(hipEventRecord(c->ev[3], s));
    TRY(launch_stage4(c->fa, s, &c->prof, g, comp));
    if (prof) HIPCHK(hipEventRecord(c->ev[4], s));
    HIPCHK(hipStreamSynchronize(s));
    if (prof) T.stage4_composite_ms = elapsed(c->ev[3], c->ev[4]);
    T.stage1_skipped = 1;
    c->max_g = 0;
    c->active = 0;
    c->n_entries = 0;
    c->s3_valid = false;
    if (tim) *tim = T;
    return CGS_OK;
  }

  if (prof) HIPCHK(hipEventRecord(c->ev[0], s));
  // ---- Stage 1 ----
  SceneFrame F;
  bool deformed = false;
  TRY(stage1(c, prm, s, &c->prof, F, deformed));
  T.stage1_skipped = deformed ? 0u : 1u;
  const uint32_t N = F.N, max_g = F.max_g;
  c->max_g = max_g;
  if (prof) HIPCHK(hipEventRecord(c->ev[1], s));

  // ---- Stage 3 cache key (gpu_state.rs:956-981) ----
  const Stage3Key key = stage3_key(c, cam, g, prm);
  bool use_cache = !prm->force_stage3 && c->s3_valid && c->s3_key == key;
  // or the SortCache hysteresis of the CPU fallback frame (pipeline.rs:552-600):
  // cage state = the vertices of this frame's cages, flattened
  const cgs_sort_frame sf = {c->n_rest, cam->focal_x, cam->focal_y,
                             prm->strain_split_threshold, prm->mip_filter_3d_var};
  const float* cage_xyz = c->has_cages ? c->last_verts.data() : nullptr;
  const uint64_t cage_nv = c->has_cages ? c->last_verts.size() / 3 : 0;
  if (c->hysteresis) {
    const bool hit = c->sort_cache.take(c->sort_cache.params_match(sf), cam->view_matrix,
                                        cage_xyz, cage_nv, g.W, g.H);
    use_cache = hit && !prm->force_stage3 && c->s3_valid && c->s3_max_g == max_g;
  }

  // ---- Stage 2 ----
  // (a splat culled now but present in a reused tile order gets a record of
  // zero opacity, so that it contributes nothing)
  TRY(stage2(c, F, c->fa, s, &c->prof, cam, comp, prm, c->hysteresis && use_cache));
  if (prof) HIPCHK(hipEventRecord(c->ev[2], s));

  // ---- Stage 3 ----
  if (!use_cache)
    TRY(launch_stage3(c->fa, s, &c->prof, g, cam->near_plane, N, F.split_count, max_g));
  if (prof) HIPCHK(hipEventRecord(c->ev[3], s));
  // ---- Stage 4 ----
  TRY(launch_stage4(c->fa, s, &c->prof, g, comp));
  if (prof) HIPCHK(hipEventRecord(c->ev[4], s));

  TRY(read_counters(c, c->has_cages));
  const uint32_t splits = c->h_counters[4];
  uint32_t survivors = use_cache ? c->s3_survivors : c->h_counters[3];
  if (!use_cache) {
    uint32_t total = c->h_counters[2];
    // the tile-entry arena overflowed (grow it), or the segmented sort needed
    // more chunk blocks or merge rounds than it launched: re-run Stages 3-4
    int reruns = 0;
    bool more = bucket_rerun(c->fa, c->h_counters + 12);
    while ((total > c->fa.entry_cap || more) && reruns < 4) {
      if (total > c->fa.entry_cap) {
        uint64_t want = (uint64_t)total + total / 2;
        if (want >= (1ull << 30)) return fail(CGS_ERR_EXECUTION, "tile entries exceed 2^30");
        c->fa.entry_cap = (uint32_t)want;
      }
      TRY(launch_stage3(c->fa, s, &c->prof, g, cam->near_plane, N, F.split_count, max_g));
      if (prof) HIPCHK(hipEventRecord(c->ev[3], s));
      TRY(launch_stage4(c->fa, s, &c->prof, g, comp));
      if (prof) HIPCHK(hipEventRecord(c->ev[4], s));
      TRY(read_counters(c, c->has_cages));
      total = c->h_counters[2];
      survivors = c->h_counters[3];
      more = bucket_rerun(c->fa, c->h_counters + 12);
      reruns++;
    }
    if (total > c->fa.entry_cap || more)
      return fail(CGS_ERR_EXECUTION, "Stage 3 did not settle after 4 re-runs (internal error)");
    c->n_entries = total;
    c->s3_key = key;
    c->s3_valid = true;
    c->s3_survivors = survivors;
    c->s3_max_g = max_g;
    T.stage3_reruns = (uint32_t)reruns;
  } else {
    T.stage3_cached = 1;
  }
  if (c->hysteresis) c->sort_cache.commit_frame(sf, cam->view_matrix, cage_xyz, cage_nv, g.W, g.H);
  c->active = N + (F.split_possible ? splits : 0u);
  if (c->prof.on) {
    c->prof.mark("end", s);
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i + 1 < c->prof.n; ++i) {
      c->prof_names[i] = c->prof.names[i];
      c->prof_ms[i] = elapsed(c->prof.ev[i], c->prof.ev[i + 1]);
    }
    c->prof_count = c->prof.n > 0 ? c->prof.n - 1 : 0;
    c->prof.on = false;
  }
  if (prof) {
    T.stage1_deform_ms = elapsed(c->ev[0], c->ev[1]);
    T.stage2_project_ms = elapsed(c->ev[1], c->ev[2]);
    T.stage3_bin_sort_ms = use_cache ? 0.0 : elapsed(c->ev[2], c->ev[3]);
    T.stage4_composite_ms = elapsed(c->ev[3], c->ev[4]);
    T.used_hardware_timestamps = 1;
  }
  T.active_gaussians = c->active;
  T.surviving_gaussians = survivors;
  T.total_tile_entries = c->n_entries;
  if (tim) *tim = T;
  return CGS_OK;
}

int cgs_kernel_times(cgs_context* c, char* names, double* ms, uint32_t cap, uint32_t* count) {
  if (!c || !count) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  uint32_t n = (uint32_t)c->prof_count;
  if (n > cap) n = cap;
  for (uint32_t i = 0; i < n; ++i) {
    if (names) snprintf(names + 32 * i, 32, "%s", c->prof_names[i]);
    if (ms) ms[i] = c->prof_ms[i];
  }
  *count = n;
  return CGS_OK;
}

// ---- launch tuning ----

// 16 x 16 tiles: the 60 choices of the LSD radix sort (partition x digit
// width x compositor batch x tile order), then the 6 of the bucket mode (batch
// x order); then the other tile sizes with fewer choices each: 8 x 8 (batch
// 64) radix partitions 16384 / 8192 x order, then bucket x order; 16 x 32 and
// 32 x 16 radix partitions 16384 / 8192 x batch 256 / 128 x order, then
// bucket x batch x order.  Then the presort (tile_sort 2): 16 x 16 with
// partitions 16384 / 10240 / 8192 x batch 256 / 128 x order, 8 x 8 with
// partitions 16384 / 8192 x order.  Last, 16 x 16 with compositor batches of
// 512 (two splats staged per thread): LSD partitions 16384 / 12288 / 10240 /
// 8192 x order.
uint32_t cgs_tuning_candidates(void) { return 120u; }

int cgs_tuning_candidate(uint32_t i, cgs_tuning* out) {
  if (!out || i >= 120u) return fail(CGS_ERR_INVALID_ARG, "candidate index out of range");
  static const uint32_t parts[5] = {16384u, 12288u, 10240u, 8192u, 4096u},
                        digits[2] = {8u, 4u}, batches[3] = {256u, 128u, 64u},
                        orders[2] = {1u, 0u}, parts2[2] = {16384u, 8192u},
                        parts3[3] = {16384u, 10240u, 8192u};
  out->radix_partition_keys = 16384u;
  out->radix_digit_bits = 8u;
  out->tile_w = 16u;
  out->tile_h = 16u;
  if (i < 60u) {
    out->radix_partition_keys = parts[i % 5u];
    out->radix_digit_bits = digits[(i / 5u) % 2u];
    out->composite_batch = batches[(i / 10u) % 3u];
    out->tile_order = orders[i / 30u];
    out->tile_sort = 0u;
  } else if (i < 66u) {
    i -= 60u;
    out->composite_batch = batches[i % 3u];
    out->tile_order = orders[i / 3u];
    out->tile_sort = 1u;
  } else if (i < 72u) {
    i -= 66u;
    out->tile_w = out->tile_h = 8u;
    out->composite_batch = 64u;
    if (i < 4u) {
      out->radix_partition_keys = parts2[i % 2u];
      out->tile_order = orders[i / 2u];
      out->tile_sort = 0u;
    } else {
      out->tile_order = orders[i - 4u];
      out->tile_sort = 1u;
    }
  } else if (i >= 112u) {
    i -= 112u;
    static const uint32_t parts4[4] = {16384u, 12288u, 10240u, 8192u};
    out->radix_partition_keys = parts4[i % 4u];
    out->composite_batch = 512u;
    out->tile_order = orders[i / 4u];
    out->tile_sort = 0u;
  } else if (i >= 96u) {
    i -= 96u;
    out->tile_sort = 2u;
    if (i < 12u) {
      out->radix_partition_keys = parts3[i % 3u];
      out->composite_batch = batches[(i / 3u) % 2u];
      out->tile_order = orders[i / 6u];
    } else {
      i -= 12u;
      out->tile_w = out->tile_h = 8u;
      out->radix_partition_keys = parts2[i % 2u];
      out->composite_batch = 64u;
      out->tile_order = orders[i / 2u];
    }
  } else {
    i -= 72u;
    const bool tall = i < 12u;
    if (!tall) i -= 12u;
    out->tile_w = tall ? 16u : 32u;
    out->tile_h = tall ? 32u : 16u;
    if (i < 8u) {
      out->radix_partition_keys = parts2[i % 2u];
      out->composite_batch = batches[(i / 2u) % 2u];
      out->tile_order = orders[i / 4u];
      out->tile_sort = 0u;
    } else {
      i -= 8u;
      out->composite_batch = batches[i % 2u];
      out->tile_order = orders[i / 2u];
      out->tile_sort = 1u;
    }
  }
  return CGS_OK;
}

int cgs_set_tuning(cgs_context* c, const cgs_tuning* t) {
  if (!c || !t) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  TRY(check_tuning(t));
  if (t->tile_w != c->fa.tune.tile_w || t->tile_h != c->fa.tune.tile_h) {
    // the kept Stage 3 (and the hysteresis cache's) is for the old tile grid
    c->s3_valid = false;
    c->sort_cache.invalidate();
  }
  c->fa.tune = *t;
  return CGS_OK;
}

int cgs_get_tuning(cgs_context* c, cgs_tuning* t) {
  if (!c || !t) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  *t = c->fa.tune;
  return CGS_OK;
}

int cgs_autotune(cgs_context* c, const cgs_camera_uniforms* cam,
                 const cgs_compositor_uniforms* comp, const cgs_frame_params* prm, uint32_t reps,
                 uint32_t flags, cgs_tuning* best, double* ms, uint32_t n_ms) {
  TRY(ensure_ctx(c));
  if (!cam || !comp || !prm) return fail(CGS_ERR_INVALID_ARG, "null frame argument");
  if (flags & ~CGS_TUNE_TILE_SHAPES) return fail(CGS_ERR_INVALID_ARG, "unknown autotune flags");
  if (reps == 0) reps = 1;
  cgs_frame_params p = *prm;
  p.force_stage3 = 1;
  p.profile_stages = 0;
  const cgs_tuning start = c->fa.tune;
  // per tile size, the frame every candidate of that size must equal: the
  // current setting's for its size, else the size's first candidate's
  struct SizeRef {
    uint32_t tw, th;
    uint64_t entries;
    std::vector<uint8_t> rgba;
    std::vector<uint32_t> ranges;
  };
  std::vector<SizeRef> refs;
  uint32_t ran_w = start.tile_w, ran_h = start.tile_h;  // tile size of the last frame run
  auto restore = [&](int rc) {
    c->fa.tune = start;
    if (ran_w != start.tile_w || ran_h != start.tile_h) {
      c->s3_valid = false;
      c->sort_cache.invalidate();
    }
    return rc;
  };
  // one frame of the current setting; its framebuffer, ranges and entry count
  auto frame = [&](SizeRef& r) -> int {
    const FrameGeom g = frame_geom(comp, c->fa.tune);
    cgs_stage_timings T;
    ran_w = c->fa.tune.tile_w;
    ran_h = c->fa.tune.tile_h;
    TRY(cgs_execute_frame(c, cam, comp, &p, &T));
    r.tw = c->fa.tune.tile_w;
    r.th = c->fa.tune.tile_h;
    r.entries = T.total_tile_entries;
    r.rgba.resize((size_t)g.W * g.H * 4);
    TRY(cgs_readback_rgba8(c, r.rgba.data(), g.W, g.H));
    r.ranges.assign(c->n_rest ? (size_t)g.T * 2 : 0, 0u);
    if (c->n_rest)
      HIPCHK(hipMemcpy(r.ranges.data(), c->fa.ranges.p, r.ranges.size() * 4,
                       hipMemcpyDeviceToHost));
    return CGS_OK;
  };
  refs.emplace_back();
  if (int rc = frame(refs.back())) return restore(rc);
  const uint32_t n = cgs_tuning_candidates();
  double best_ms = 1e300;
  cgs_tuning best_t = start;
  SizeRef got;
  for (uint32_t i = 0; i < n; ++i) {
    cgs_tuning cand;
    TRY(cgs_tuning_candidate(i, &cand));
    if (ms && i < n_ms) ms[i] = -1.0;
    if (!(flags & CGS_TUNE_TILE_SHAPES) &&
        (cand.tile_w != start.tile_w || cand.tile_h != start.tile_h))
      continue;
    c->fa.tune = cand;
    if (int rc = frame(got)) return restore(rc);
    const SizeRef* ref = nullptr;
    for (const SizeRef& r : refs)
      if (r.tw == cand.tile_w && r.th == cand.tile_h) ref = &r;
    if (!ref) {
      refs.push_back(got);
    } else if (got.entries != ref->entries || got.rgba != ref->rgba || got.ranges != ref->ranges) {
      return restore(fail(CGS_ERR_EXECUTION, "a tuning candidate changed the frame (internal error)"));
    }
    cgs_stage_timings T;
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    for (uint32_t r = 0; r < reps; ++r)
      if (int rc = cgs_execute_frame(c, cam, comp, &p, &T)) return restore(rc);
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[1]));
    const double t_ms = elapsed(c->ev[0], c->ev[1]) / reps;
    if (ms && i < n_ms) ms[i] = t_ms;
    if (t_ms < best_ms) {
      best_ms = t_ms;
      best_t = cand;
    }
  }
  c->fa.tune = best_t;
  if (ran_w != best_t.tile_w || ran_h != best_t.tile_h) {
    c->s3_valid = false;
    c->sort_cache.invalidate();
  }
  if (best) *best = best_t;
  return CGS_OK;
}

int cgs_readback_rgba8(cgs_context* c, uint8_t* dst, uint32_t width, uint32_t height) {
  TRY(ensure_ctx(c));
  if (!dst) return fail(CGS_ERR_INVALID_ARG, "null destination");
  if (!c->fa.out_rgba.p || width != c->last_w || height != c->last_h)
    return fail(CGS_ERR_SHORT_READBACK, "framebuffer size differs from the last frame");
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)width * 4, c->fa.out_rgba.p, (size_t)c->fa.padded_row * 4,
                          (size_t)width * 4, height, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_framebuffer_device_ptr(cgs_context* c, void** ptr, uint32_t* padded_row_u32) {
  if (!c || !ptr || !padded_row_u32) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  *ptr = c->fa.out_rgba.p;
  *padded_row_u32 = c->fa.padded_row;
  return CGS_OK;
}

int cgs_read_projected(cgs_context* c, float* attrs10, uint32_t* extents, uint64_t n) {
  TRY(ensure_ctx(c));
  if (n > c->max_g) return fail(CGS_ERR_SHORT_READBACK, "more splats requested than projected");
  std::vector<float> geom((size_t)n * 8), color((size_t)n * 4);
  std::vector<uint32_t> ext(n);
  if (n) {
    HIPCHK(hipMemcpyAsync(geom.data(), c->fa.geom.p, n * 32, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(color.data(), c->fa.color.p, n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ext.data(), c->fa.extents.p, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint64_t i = 0; i < n; ++i) {
    if (extents) extents[i] = ext[i];
    if (!attrs10) continue;
    float* a = attrs10 + i * 10;
    if (ext[i] == 0) {
      for (int k = 0; k < 10; ++k) a[k] = 0.0f;
      continue;
    }
    const float* gg = &geom[i * 8];
    const float* cc = &color[i * 4];
    a[0] = gg[0]; a[1] = gg[1]; a[2] = gg[2]; a[3] = gg[3]; a[4] = gg[4]; a[5] = gg[5];
    a[6] = cc[0]; a[7] = cc[1]; a[8] = cc[2]; a[9] = gg[6];
  }
  return CGS_OK;
}

int cgs_read_stage3(cgs_context* c, uint32_t* counts, uint32_t* offsets, uint32_t* sorted_keys,
                    uint32_t* sorted_ids, uint32_t* ranges, uint64_t* n_entries) {
  TRY(ensure_ctx(c));
  if ((counts || offsets) && !c->fa.captured)
    return fail(CGS_ERR_INVALID_ARG, "counts/offsets need cgs_set_capture(ctx, 1) before the frame");
  const uint32_t E = c->n_entries;
  const size_t T = (size_t)c->fa.tiles_x * c->fa.tiles_y;
  // after a presort the binning's arrays are per position: order[j] is the
  // splat at position j
  std::vector<uint32_t> order, pos_counts, pos_offsets;
  const bool permuted = c->fa.order && c->max_g && (counts || offsets);
  if (permuted) {
    order.resize(c->max_g);
    HIPCHK(hipMemcpyAsync(order.data(), c->fa.order, (size_t)c->max_g * 4, hipMemcpyDeviceToHost,
                          c->stream));
    if (counts) pos_counts.resize(c->max_g);
    if (offsets) pos_offsets.resize(c->max_g);
  }
  uint32_t* cdst = permuted && counts ? pos_counts.data() : counts;
  uint32_t* odst = permuted && offsets ? pos_offsets.data() : offsets;
  if (counts && c->max_g)
    HIPCHK(hipMemcpyAsync(cdst, c->fa.cnt_dbg.p, (size_t)c->max_g * 4, hipMemcpyDeviceToHost, c->stream));
  if (offsets && c->max_g)
    HIPCHK(hipMemcpyAsync(odst, c->fa.off_dbg.p, (size_t)c->max_g * 4, hipMemcpyDeviceToHost, c->stream));
  if (sorted_keys && E)
    HIPCHK(hipMemcpyAsync(sorted_keys, c->fa.sorted_keys, (size_t)E * 4, hipMemcpyDeviceToHost, c->stream));
  if (sorted_ids && E)
    HIPCHK(hipMemcpyAsync(sorted_ids, c->fa.sorted_ids, (size_t)E * 4, hipMemcpyDeviceToHost, c->stream));
  if (ranges && T)
    HIPCHK(hipMemcpyAsync(ranges, c->fa.ranges.p, T * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (permuted) {
    for (uint32_t j = 0; j < c->max_g; ++j) {
      const uint32_t sid = order[j];
      if (sid >= c->max_g) return fail(CGS_ERR_EXECUTION, "internal: presort order out of range");
      if (counts) counts[sid] = pos_counts[j];
      if (offsets) offsets[sid] = pos_offsets[j];
    }
  }
  if (n_entries) *n_entries = E;
  return CGS_OK;
}

int cgs_read_deformed(cgs_context* c, float* pos, float* cov, float* opa, uint32_t* split_meta,
                      uint32_t* split_count) {
  TRY(ensure_ctx(c));
  if (!c->has_cages || !c->def_geo.p) return fail(CGS_ERR_INVALID_ARG, "no deformation has run");
  uint32_t splits = 0;
  HIPCHK(hipMemcpyAsync(&splits, c->ctrl1.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const uint32_t m = c->n_rest + splits;
  if (m > c->def_cap) return fail(CGS_ERR_EXECUTION, "split count exceeds the deformed arena");
  std::vector<float> g(cgs_blk_words(m, CGS_GEO_K));
  if (!g.empty())
    HIPCHK(hipMemcpyAsync(g.data(), c->def_geo.p, g.size() * 4, hipMemcpyDeviceToHost, c->stream));
  if (split_meta && splits)
    HIPCHK(hipMemcpyAsync(split_meta, c->split_meta.p, (size_t)splits * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < m; ++i) {
    if (pos)
      for (uint32_t k = 0; k < 3; ++k) pos[(size_t)i * 3 + k] = g[cgs_blk(i, k, CGS_GEO_K)];
    if (cov)
      for (uint32_t k = 0; k < 6; ++k) cov[(size_t)i * 6 + k] = g[cgs_blk(i, 3 + k, CGS_GEO_K)];
    if (opa) opa[i] = g[cgs_blk(i, CGS_CH_OPA, CGS_GEO_K)];
  }
  if (split_count) *split_count = splits;
  return CGS_OK;
}

int cgs_export_deformed_scene(cgs_context* c, uint32_t sh_degree, float* d_pos, float* d_cov,
                              float* d_opa, float* d_sh, uint64_t capacity, uint64_t* count) {
  TRY(ensure_ctx(c));
  if (!count) return fail(CGS_ERR_INVALID_ARG, "null count");
  *count = 0;
  if (!c->has_cages || !c->def_geo.p || !c->rot.p)
    return fail(CGS_ERR_INVALID_ARG, "no deformation has run");
  if (c->f16) return fail(CGS_ERR_UNSUPPORTED, "cage deformation of the f16 storage variant");
  if (c->n_rest == 0) return CGS_OK;
  uint32_t splits = 0;
  HIPCHK(hipMemcpyAsync(&splits, c->ctrl1.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const uint64_t m = (uint64_t)c->n_rest + splits;
  if (m > c->def_cap) return fail(CGS_ERR_EXECUTION, "split count exceeds the deformed arena");
  if (!d_pos && !d_cov && !d_opa && !d_sh) {  // a query of the record count
    *count = m;
    return CGS_OK;
  }
  if (!d_pos || !d_cov || !d_opa || !d_sh) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  if (reinterpret_cast<uintptr_t>(d_sh) % 16u != 0u)
    return fail(CGS_ERR_INVALID_ARG, "d_sh must be 16-byte aligned");
  if (m > capacity) return fail(CGS_ERR_INVALID_ARG, "capacity below the deformed splat count");
  const uint32_t deg = sh_degree > 3u ? 3u : sh_degree;
  TRY(ensure_words(c->tet_shm, (size_t)c->num_tets * CGS_SH_MATS));
  cgs_context_prof_begin(c);  // per-launch device times, for cgs_kernel_times
  c->prof.mark("k_tet_sh_matrices", c->stream);
  cgs_launch_tet_sh_matrices(P<float>(c->rot), c->num_tets, P<float>(c->tet_shm), c->stream);
  CgsExportDeformedArgs A;
  A.geo = P<float>(c->def_geo);
  A.sh = P<float>(c->rest_sh);
  A.bind_tet = P<uint32_t>(c->bind_tet);
  A.split_meta = P<uint32_t>(c->split_meta);
  A.mats = P<float>(c->tet_shm);
  A.num_rest = c->n_rest;
  A.count = (uint32_t)m;
  A.num_tets = c->num_tets;
  A.degree = deg;
  A.pos = d_pos;
  A.cov = d_cov;
  A.opa = d_opa;
  A.out_sh = d_sh;
  c->prof.mark("k_export_deformed", c->stream);
  cgs_launch_export_deformed(A, c->stream);
  const hipError_t e = hipGetLastError();
  cgs_context_prof_end(c, e == hipSuccess);
  if (e != hipSuccess)
    return fail(CGS_ERR_EXECUTION, std::string("export deformed scene: ") + hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(c->stream));
  *count = m;
  return CGS_OK;
}

int cgs_read_tet_transforms(cgs_context* c, float* jac, float* rot, float* stretch) {
  TRY(ensure_ctx(c));
  if (!c->has_cages || !c->jac.p) return fail(CGS_ERR_INVALID_ARG, "no deformation has run");
  const size_t t = c->num_tets;
  if (jac) HIPCHK(hipMemcpyAsync(jac, c->jac.p, t * 36, hipMemcpyDeviceToHost, c->stream));
  if (rot) HIPCHK(hipMemcpyAsync(rot, c->rot.p, t * 36, hipMemcpyDeviceToHost, c->stream));
  if (stretch) HIPCHK(hipMemcpyAsync(stretch, c->stretch.p, t * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_load_projected(cgs_context* c, const float* attrs10, const uint32_t* extents, uint64_t n) {
  TRY(ensure_ctx(c));
  if (n && (!attrs10 || !extents)) return fail(CGS_ERR_INVALID_ARG, "null arrays");
  if (n >= (1ull << 28)) return fail(CGS_ERR_INVALID_ARG, "too many splats");
  std::vector<float> geom((size_t)n * 8), color((size_t)n * 4);
  for (uint64_t i = 0; i < n; ++i) {
    const float* a = attrs10 + i * 10;
    float* gg = &geom[i * 8];
    gg[0] = a[0]; gg[1] = a[1]; gg[2] = a[2]; gg[3] = a[3]; gg[4] = a[4]; gg[5] = a[5];
    gg[6] = a[9];
    gg[7] = extents[i] ? cgs_fclamp(2.0f * cgs_logf(255.0f * a[5]), 0.0f, 8.0f) : 0.0f;
    color[i * 4] = a[6]; color[i * 4 + 1] = a[7]; color[i * 4 + 2] = a[8]; color[i * 4 + 3] = 0.0f;
  }
  TRY(ensure_words(c->fa.geom, (size_t)8 * n));
  TRY(ensure_words(c->fa.color, (size_t)4 * n));
  TRY(ensure_words(c->fa.extents, n));
  if (n) {
    HIPCHK(hipMemcpyAsync(c->fa.geom.p, geom.data(), n * 32, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->fa.color.p, color.data(), n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->fa.extents.p, extents, n * 4, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->max_g = (uint32_t)n;
  c->active = (uint32_t)n;
  c->s3_valid = false;
  return CGS_OK;
}

int cgs_run_stage3(cgs_context* c, const cgs_camera_uniforms* cam,
                   const cgs_compositor_uniforms* comp, uint64_t* n_entries) {
  TRY(ensure_ctx(c));
  if (!cam || !comp) return fail(CGS_ERR_INVALID_ARG, "null argument");
  const FrameGeom g = frame_geom(comp, c->fa.tune);
  c->last_w = g.W;
  c->last_h = g.H;
  for (int attempt = 0; attempt < 5; ++attempt) {
    TRY(launch_stage3(c->fa, c->stream, &c->prof, g, cam->near_plane, c->active, nullptr, c->max_g));
    TRY(read_counters(c, false));
    const uint32_t total = c->h_counters[2];
    c->n_entries = total;
    const bool more = bucket_rerun(c->fa, c->h_counters + 12);
    if (total <= c->fa.entry_cap && !more) break;
    if (total <= c->fa.entry_cap) continue;
    uint64_t want = (uint64_t)total + total / 2;
    if (want >= (1ull << 30)) return fail(CGS_ERR_EXECUTION, "tile entries exceed 2^30");
    c->fa.entry_cap = (uint32_t)want;
  }
  if (n_entries) *n_entries = c->n_entries;
  return CGS_OK;
}

int cgs_run_stage4(cgs_context* c, const cgs_compositor_uniforms* comp) {
  TRY(ensure_ctx(c));
  if (!comp) return fail(CGS_ERR_INVALID_ARG, "null argument");
  const FrameGeom g = frame_geom(comp, c->fa.tune);
  if (g.tiles_x != c->fa.tiles_x || g.tiles_y != c->fa.tiles_y)
    return fail(CGS_ERR_INVALID_ARG, "compositor size differs from Stage 3");
  TRY(launch_stage4(c->fa, c->stream, &c->prof, g, comp));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_malloc(cgs_context* c, uint64_t bytes, void** dptr) {
  TRY(ensure_ctx(c));
  if (!dptr) return fail(CGS_ERR_INVALID_ARG, "null out pointer");
  HIPCHK(hipMalloc(dptr, bytes ? bytes : 4));
  return CGS_OK;
}

int cgs_free(cgs_context* c, void* dptr) {
  TRY(ensure_ctx(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dptr) HIPCHK(hipFree(dptr));
  return CGS_OK;
}

int cgs_memcpy_h2d(cgs_context* c, void* dst, const void* src, uint64_t bytes) {
  TRY(ensure_ctx(c));
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_memcpy_d2h(cgs_context* c, void* dst, const void* src, uint64_t bytes) {
  TRY(ensure_ctx(c));
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_memcpy_d2d(cgs_context* c, void* dst, const void* src, uint64_t bytes) {
  TRY(ensure_ctx(c));
  if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_exclusive_scan_u32(cgs_context* c, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                           uint32_t* total) {
  TRY(ensure_ctx(c));
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "n too large");
  if (n && (!d_in || !d_out)) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  const uint32_t parts = cgs_scan_num_parts((uint32_t)n);
  const size_t words = 4 + 2 * (size_t)(parts ? parts : 1);
  TRY(ensure_words(c->prim_ctrl, words));
  uint32_t* ctrl = P<uint32_t>(c->prim_ctrl);
  cgs_launch_fill_u32(ctrl, 0u, words, c->stream);
  cgs_launch_scan(d_in, d_out, (uint32_t)n, ctrl, reinterpret_cast<uint64_t*>(ctrl + 4), ctrl + 1,
                  c->stream);
  HIPCHK(hipGetLastError());
  uint32_t tot = 0;
  HIPCHK(hipMemcpyAsync(c->h_counters + 8, ctrl, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  tot = c->h_counters[8];
  if (total) *total = tot;
  return CGS_OK;
}

int cgs_exclusive_scan_u32_noalloc(cgs_context* c, const uint32_t* d_in, uint32_t* d_out,
                                   uint64_t n, uint32_t* total) {
  TRY(ensure_ctx(c));
  if (n > CGS_SCAN_NOALLOC_MAX_N)
    return fail(CGS_ERR_INVALID_ARG, "n exceeds CGS_SCAN_NOALLOC_MAX_N (2^28)");
  if (n && (!d_in || !d_out)) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  uint32_t* ctrl = c->scan_scratch;
  if (!ctrl) return fail(CGS_ERR_RUNTIME, "scan scratch not resolved");
  const uint32_t parts = cgs_scan_num_parts((uint32_t)n);
  const size_t words = 4 + 2 * (size_t)(parts ? parts : 1);
  // the previous call's total, ticket and status words: cleared by a kernel
  // (no memset node, so the call can be captured into a graph)
  cgs_launch_fill_u32(ctrl, 0u, words, c->stream);
  cgs_launch_scan(d_in, d_out, (uint32_t)n, ctrl, reinterpret_cast<uint64_t*>(ctrl + 4), ctrl + 1,
                  c->stream);
  HIPCHK(hipGetLastError());
  if (total) {
    HIPCHK(hipMemcpyAsync(c->h_counters + 9, ctrl, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *total = c->h_counters[9];
  }
  return CGS_OK;
}

static int radix_prim(cgs_context* c, uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t bits,
                      bool iota, int* in_alt) {
  TRY(ensure_words(c->prim_alt_k, n));
  TRY(ensure_words(c->prim_alt_v, n));
  const size_t words = cgs_radix_control_words(n);
  TRY(ensure_words(c->prim_ctrl, words));
  uint32_t* ctrl = P<uint32_t>(c->prim_ctrl);
  const size_t parts = cgs_radix_num_parts(n) ? cgs_radix_num_parts(n) : 1u;
  const size_t status_words = 4 * parts * 256;
  // layout: hist[1024] | status | tickets[4] | partial histograms
  cgs_launch_fill_u32(ctrl + 1024, 0u, status_words + 4, c->stream);
  CgsRadixBuffers R;
  R.keys = keys;
  R.vals = vals;
  R.keys_alt = P<uint32_t>(c->prim_alt_k);
  R.vals_alt = P<uint32_t>(c->prim_alt_v);
  R.hist = ctrl;
  R.status = ctrl + 1024;
  R.tickets = ctrl + 1024 + status_words;
  R.hist_partial = ctrl + 1024 + status_words + 4;
  R.cap = n;
  *in_alt = cgs_launch_radix_sort(R, nullptr, n, bits, iota, false, c->stream);
  HIPCHK(hipGetLastError());
  return CGS_OK;
}

int cgs_radix_sort_u32_pairs(cgs_context* c, uint32_t* keys, uint32_t* vals, uint64_t n,
                             uint32_t sort_bits) {
  TRY(ensure_ctx(c));
  if (n >= (1ull << 30)) return fail(CGS_ERR_INVALID_ARG, "n too large (>= 2^30)");
  if (n == 0 || sort_bits == 0) return CGS_OK;
  if (!keys || !vals) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  int alt = 0;
  TRY(radix_prim(c, keys, vals, (uint32_t)n, sort_bits, false, &alt));
  if (alt) {
    HIPCHK(hipMemcpyAsync(keys, c->prim_alt_k.p, n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(vals, c->prim_alt_v.p, n * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_radix_sort_u64_pairs(cgs_context* c, uint32_t* tiles, uint32_t* depths, uint32_t* ids,
                             uint64_t n, uint32_t tile_bits) {
  TRY(ensure_ctx(c));
  if (n >= (1ull << 30)) return fail(CGS_ERR_INVALID_ARG, "n too large (>= 2^30)");
  if (n == 0) return CGS_OK;
  if (!tiles || !depths || !ids) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  const uint32_t m = (uint32_t)n;
  // 1) stable sort by depth with the entry index as payload (4 passes: result in place)
  // three arrays of m words, each starting 16-byte aligned (the gather and
  // the sort use 16-byte accesses)
  const size_t mp = ((size_t)m + 3u) & ~(size_t)3u;
  TRY(ensure_words(c->prim_alt_d, mp * 3));
  uint32_t* perm = P<uint32_t>(c->prim_alt_d);
  uint32_t* t2 = perm + mp;
  uint32_t* i2 = perm + 2 * mp;
  int alt = 0;
  TRY(radix_prim(c, depths, perm, m, 32, true, &alt));
  if (alt) {
    HIPCHK(hipMemcpyAsync(depths, c->prim_alt_k.p, n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(perm, c->prim_alt_v.p, n * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  // 2) gather tiles and ids through the permutation
  TRY(ensure_words(c->staging, 1));
  uint32_t* d_n = P<uint32_t>(c->staging);
  cgs_launch_fill_u32(d_n, m, 1, c->stream);
  if (!cgs_launch_gather2(perm, tiles, ids, t2, i2, d_n, m, c->stream))
    return fail(CGS_ERR_INVALID_ARG, "internal: gather buffers not 16-byte aligned");
  // 3) stable sort by tile id
  TRY(radix_prim(c, t2, i2, m, tile_bits ? tile_bits : 1, false, &alt));
  const uint32_t* st = alt ? P<uint32_t>(c->prim_alt_k) : t2;
  const uint32_t* si = alt ? P<uint32_t>(c->prim_alt_v) : i2;
  // depths follow the final order: depth_sorted[perm2]; recompute through ids
  HIPCHK(hipMemcpyAsync(tiles, st, n * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(ids, si, n * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

// TetMesh::bind_gaussians / bind_gaussians_with_bounds (src/deformation/
// mod.rs:263-323) on the GPU.  Host arrays in, host arrays out.
int cgs_bind_tets(cgs_context* c, const float* rest_vertices, uint32_t num_vertices,
                  const uint32_t* elements, const float* inv_dm, uint32_t num_tets,
                  const float* positions, uint64_t n, const float* bounds, uint32_t* out_tet,
                  float* out_weights) {
  TRY(ensure_ctx(c));
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  if (!rest_vertices || !elements || !inv_dm || !positions || !out_tet || !out_weights)
    return fail(CGS_ERR_INVALID_ARG, "null buffer");
  if (num_tets == 0) return fail(CGS_ERR_INVALID_ARG, "a cage needs at least one tetrahedron");
  for (uint32_t t = 0; t < num_tets; ++t)
    for (int k = 0; k < 4; ++k)
      if (elements[(size_t)t * 4 + k] >= num_vertices)
        return fail(CGS_ERR_INVALID_ARG, "tet vertex index out of range");
  // per-tet record (X0, inv_dm) and the box that holds every point the tet can
  // accept: p = X0 + A w with A = inv_dm^-1, and a penalty <= 1e-4 keeps p
  // within 1e-4 of a diameter of the corners X0 + A {0, e1, e2, e3}
  std::vector<float> rec((size_t)num_tets * 12);
  std::vector<double> box((size_t)num_tets * 6);
  bool grid_ok = true;
  double glo[3] = {1e300, 1e300, 1e300}, ghi[3] = {-1e300, -1e300, -1e300};
  for (uint32_t t = 0; t < num_tets; ++t) {
    const float* x0 = rest_vertices + (size_t)elements[(size_t)t * 4] * 3;
    const float* m = inv_dm + (size_t)t * 9;  // column-major: m[col*3 + row]
    float* r = rec.data() + (size_t)t * 12;
    for (int k = 0; k < 3; ++k) r[k] = x0[k];
    for (int k = 0; k < 9; ++k) r[3 + k] = m[k];
    // A = inverse of M (row-major M[row][col] = m[col*3 + row]), in double
    double M[3][3];
    for (int row = 0; row < 3; ++row)
      for (int col = 0; col < 3; ++col) M[row][col] = m[col * 3 + row];
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    if (!(std::fabs(det) > 1e-30) || !std::isfinite(det)) {
      grid_ok = false;  // a tet that accepts an unbounded region: scan everything
      continue;
    }
    double Ai[3][3];
    Ai[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    Ai[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    Ai[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    Ai[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    Ai[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    Ai[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    Ai[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    Ai[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    Ai[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = x0[k];
    for (int e = 0; e < 3; ++e)
      for (int k = 0; k < 3; ++k) {
        const double v = x0[k] + Ai[k][e];
        lo[k] = std::min(lo[k], v);
        hi[k] = std::max(hi[k], v);
      }
    double diam = 0.0, mag = 0.0;
    for (int k = 0; k < 3; ++k) {
      diam = std::max(diam, hi[k] - lo[k]);
      mag = std::max(mag, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
    }
    const double pad = 1e-3 * diam + 1e-5 * mag + 1e-7;
    for (int k = 0; k < 3; ++k) {
      box[(size_t)t * 6 + k] = lo[k] - pad;
      box[(size_t)t * 6 + 3 + k] = hi[k] + pad;
      if (!std::isfinite(lo[k] - pad) || !std::isfinite(hi[k] + pad)) grid_ok = false;
      glo[k] = std::min(glo[k], lo[k] - pad);
      ghi[k] = std::max(ghi[k], hi[k] + pad);
    }
  }
  // grid of about 2 cells per tet edge length, at most 128 per axis
  uint32_t dims[3] = {1, 1, 1};
  std::vector<uint32_t> cell_start, cell_tets;
  double cell[3] = {1, 1, 1};
  if (grid_ok) {
    const double cells_target = std::max(1.0, 8.0 * num_tets);
    const double vol = std::max((ghi[0] - glo[0]) * (ghi[1] - glo[1]) * (ghi[2] - glo[2]), 1e-30);
    const double h = std::cbrt(vol / cells_target);
    for (int k = 0; k < 3; ++k) {
      dims[k] = (uint32_t)std::min(128.0, std::max(1.0, std::ceil((ghi[k] - glo[k]) / h)));
      cell[k] = (ghi[k] - glo[k]) / dims[k];
      if (!(cell[k] > 0.0)) cell[k] = 1.0;
    }
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<uint32_t> cnt(ncell + 1, 0u);
    auto range = [&](uint32_t t, int k, uint32_t& a, uint32_t& b) {
      const double fa = std::floor((box[(size_t)t * 6 + k] - glo[k]) / cell[k]);
      const double fb = std::floor((box[(size_t)t * 6 + 3 + k] - glo[k]) / cell[k]);
      a = (uint32_t)std::min<double>(dims[k] - 1, std::max(0.0, fa));
      b = (uint32_t)std::min<double>(dims[k] - 1, std::max(0.0, fb));
    };
    for (int pass = 0; pass < 2 && grid_ok; ++pass) {
      if (pass == 1) {
        size_t total = 0;
        for (size_t i = 0; i < ncell; ++i) total += cnt[i];
        if (total > (64u << 20)) {  // boxes too large to be worth a grid
          grid_ok = false;
          break;
        }
        cell_start.assign(ncell + 1, 0u);
        for (size_t i = 0; i < ncell; ++i) cell_start[i + 1] = cell_start[i] + cnt[i];
        cell_tets.assign(cell_start[ncell], 0u);
        std::fill(cnt.begin(), cnt.end(), 0u);
      }
      for (uint32_t t = 0; t < num_tets; ++t) {  // index order keeps lists ascending
        uint32_t a[3], b[3];
        for (int k = 0; k < 3; ++k) range(t, k, a[k], b[k]);
        for (uint32_t z = a[2]; z <= b[2]; ++z)
          for (uint32_t y = a[1]; y <= b[1]; ++y)
            for (uint32_t x = a[0]; x <= b[0]; ++x) {
              const size_t cidx = ((size_t)z * dims[1] + y) * dims[0] + x;
              if (pass == 1) cell_tets[cell_start[cidx] + cnt[cidx]] = t;
              cnt[cidx]++;
            }
      }
    }
  }
  // device buffers (freed before returning: binding is a one-off per cage)
  void *d_rec = nullptr, *d_pos = nullptr, *d_tet = nullptr, *d_w = nullptr;
  void *d_start = nullptr, *d_cells = nullptr;
  auto cleanup = [&]() {
    for (void* p : {d_rec, d_pos, d_tet, d_w, d_start, d_cells})
      if (p) (void)hipFree(p);
  };
  int rc = CGS_OK;
  do {
    if (hipMalloc(&d_rec, rec.size() * 4) != hipSuccess ||
        hipMalloc(&d_pos, n * 12) != hipSuccess || hipMalloc(&d_tet, n * 4) != hipSuccess ||
        hipMalloc(&d_w, n * 16) != hipSuccess) {
      rc = fail(CGS_ERR_RUNTIME, "hipMalloc failed (binding buffers)");
      break;
    }
    if (grid_ok && (hipMalloc(&d_start, cell_start.size() * 4) != hipSuccess ||
                    hipMalloc(&d_cells, std::max<size_t>(cell_tets.size(), 1) * 4) != hipSuccess)) {
      rc = fail(CGS_ERR_RUNTIME, "hipMalloc failed (binding grid)");
      break;
    }
    hipError_t e = hipMemcpyAsync(d_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice,
                                  c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_pos, positions, n * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && grid_ok) {
      e = hipMemcpyAsync(d_start, cell_start.data(), cell_start.size() * 4, hipMemcpyHostToDevice,
                         c->stream);
      if (e == hipSuccess && !cell_tets.empty())
        e = hipMemcpyAsync(d_cells, cell_tets.data(), cell_tets.size() * 4,
                           hipMemcpyHostToDevice, c->stream);
    }
    if (e != hipSuccess) {
      rc = fail(CGS_ERR_RUNTIME, std::string("upload: ") + hipGetErrorString(e));
      break;
    }
    CgsBindArgs A;
    memset(&A, 0, sizeof(A));
    A.tet_rec = (const float*)d_rec;
    A.num_tets = num_tets;
    A.pos = (const float*)d_pos;
    A.n = (uint32_t)n;
    A.has_bounds = bounds ? 1u : 0u;
    for (int k = 0; k < 3; ++k) {
      A.bmin[k] = bounds ? bounds[k] : 0.0f;
      A.bmax[k] = bounds ? bounds[3 + k] : 0.0f;
      A.lo[k] = (float)glo[k];
      A.inv_cell[k] = (float)(1.0 / cell[k]);
      A.dims[k] = dims[k];
    }
    A.cell_start = grid_ok ? (const uint32_t*)d_start : nullptr;
    A.cell_tets = grid_ok ? (const uint32_t*)d_cells : nullptr;
    A.out_tet = (uint32_t*)d_tet;
    A.out_w = (float*)d_w;
    cgs_launch_bind_tets(A, c->stream);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out_tet, d_tet, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out_weights, d_w, n * 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(CGS_ERR_EXECUTION, std::string("bind: ") + hipGetErrorString(e));
  } while (false);
  (void)hipStreamSynchronize(c->stream);
  cleanup();
  return rc;
}

// ---- mean value coordinates (src/deformation/mvc.rs) ----------------------

// Uploads a triangle cage and its vertex->triangle incidence lists (entries
// tri*3 + corner in ascending triangle order) into scratch device memory.
struct MvcCage {
  void* d_verts = nullptr;
  void* d_tris = nullptr;
  void* d_start = nullptr;
  void* d_list = nullptr;
  float h = 0.0f;
  ~MvcCage() {
    for (void* p : {d_verts, d_tris, d_start, d_list})
      if (p) (void)hipFree(p);
  }
};

static int mvc_upload(cgs_context* c, const float* verts, uint32_t nv, const uint32_t* tris,
                      uint32_t nt, MvcCage& M) {
  if (!verts || !tris) return fail(CGS_ERR_INVALID_ARG, "null cage");
  if (nv < 4 || nt < 4) return fail(CGS_ERR_INVALID_ARG, "a cage needs >= 4 vertices and triangles");
  if (cgs_mvc_lds_bytes(nv, nt) > 64 * 1024)
    return fail(CGS_ERR_INVALID_ARG, "cage too large for one wave's LDS (nv*24 + nt*12 > 64 KiB)");
  std::vector<uint32_t> start(nv + 1, 0u), list((size_t)nt * 3);
  for (uint32_t t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) {
      const uint32_t v = tris[(size_t)t * 3 + k];
      if (v >= nv) return fail(CGS_ERR_INVALID_ARG, "triangle vertex index out of range");
      start[v + 1]++;
    }
  for (uint32_t v = 0; v < nv; ++v) start[v + 1] += start[v];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint32_t t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) list[fill[tris[(size_t)t * 3 + k]]++] = t * 3 + (uint32_t)k;
  // jacobian_step_size (mvc.rs:485-495), f32
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k) {
      lo[k] = cgs_fmin(lo[k], verts[(size_t)v * 3 + k]);
      hi[k] = cgs_fmax(hi[k], verts[(size_t)v * 3 + k]);
    }
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  const float diag = cgs_fmax(sqrtf((dx * dx + dy * dy) + dz * dz), 1e-3f);
  M.h = cgs_fclamp(diag * 1e-3f, 1e-5f, 1e-2f);
  HIPCHK(hipMalloc(&M.d_verts, (size_t)nv * 12));
  HIPCHK(hipMalloc(&M.d_tris, (size_t)nt * 12));
  HIPCHK(hipMalloc(&M.d_start, (size_t)(nv + 1) * 4));
  HIPCHK(hipMalloc(&M.d_list, (size_t)nt * 12));
  HIPCHK(hipMemcpyAsync(M.d_verts, verts, (size_t)nv * 12, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(M.d_tris, tris, (size_t)nt * 12, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(M.d_start, start.data(), (size_t)(nv + 1) * 4, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(M.d_list, list.data(), (size_t)nt * 12, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  return CGS_OK;
}

// PolyCageMesh::bind_gaussians (mvc.rs:498-541).  Cage on the host; points
// and outputs in device memory: d_points [3n], d_weights [n*nv],
// d_gradients [n*nv*3] or NULL, d_inside [n] u32 or NULL.
int cgs_mvc_bind_exact(cgs_context* c, const float* cage_vertices, uint32_t num_vertices,
                       const uint32_t* triangles, uint32_t num_triangles, const float* d_points,
                       uint64_t n, float* d_weights, float* d_gradients, uint32_t* d_inside) {
  TRY(ensure_ctx(c));
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  if (!d_points || !d_weights) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  MvcCage M;
  TRY(mvc_upload(c, cage_vertices, num_vertices, triangles, num_triangles, M));
  CgsMvcArgs A;
  A.verts = (const float*)M.d_verts;
  A.nv = num_vertices;
  A.tris = (const uint32_t*)M.d_tris;
  A.nt = num_triangles;
  A.vt_start = (const uint32_t*)M.d_start;
  A.vt_list = (const uint32_t*)M.d_list;
  A.pts = d_points;
  A.n = (uint32_t)n;
  A.h = M.h;
  A.out_w = d_weights;
  A.out_g = d_gradients;
  A.mask = d_inside;
  A.list = nullptr;
  cgs_launch_mvc_bind(A, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

// PolyCageMesh::winding_number for n device points into d_out [n] f32.
int cgs_mvc_winding(cgs_context* c, const float* cage_vertices, uint32_t num_vertices,
                    const uint32_t* triangles, uint32_t num_triangles, const float* d_points,
                    uint64_t n, float* d_out) {
  TRY(ensure_ctx(c));
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  if (!d_points || !d_out) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  MvcCage M;
  TRY(mvc_upload(c, cage_vertices, num_vertices, triangles, num_triangles, M));
  CgsMvcArgs A;
  memset(&A, 0, sizeof(A));
  A.verts = (const float*)M.d_verts;
  A.nv = num_vertices;
  A.tris = (const uint32_t*)M.d_tris;
  A.nt = num_triangles;
  A.pts = d_points;
  A.n = (uint32_t)n;
  cgs_launch_mvc_winding(A, d_out, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

// ---- near/far split of the MVC sums (the learned binding) ----

// (A A^T)^-1 of A = [1 ... 1; v_1 ... v_nv] (4 x nv), in double, row-major
// into minv; false when the vertices are coplanar.
static bool linear_precision_minv(const float* verts, uint32_t nv, float minv[16]) {
  double G[4][4] = {{0}};
  for (uint32_t v = 0; v < nv; ++v) {
    const double a[4] = {1.0, verts[(size_t)v * 3], verts[(size_t)v * 3 + 1], verts[(size_t)v * 3 + 2]};
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) G[i][k] += a[i] * a[k];
  }
  double aug[4][8];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 8; ++k) aug[i][k] = k < 4 ? G[i][k] : (k - 4 == i ? 1.0 : 0.0);
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r2 = col + 1; r2 < 4; ++r2)
      if (std::fabs(aug[r2][col]) > std::fabs(aug[piv][col])) piv = r2;
    if (std::fabs(aug[piv][col]) < 1e-12) return false;
    for (int k = 0; k < 8; ++k) std::swap(aug[col][k], aug[piv][k]);
    const double d = aug[col][col];
    for (int k = 0; k < 8; ++k) aug[col][k] /= d;
    for (int r2 = 0; r2 < 4; ++r2)
      if (r2 != col) {
        const double f = aug[r2][col];
        for (int k = 0; k < 8; ++k) aug[r2][k] -= f * aug[col][k];
      }
  }
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) minv[i * 4 + k] = (float)aug[i][4 + k];
  return true;
}

static float cage_diagonal(const float* verts, uint32_t nv) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k) {
      lo[k] = cgs_fmin(lo[k], verts[(size_t)v * 3 + k]);
      hi[k] = cgs_fmax(hi[k], verts[(size_t)v * 3 + k]);
    }
  return sqrtf(((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1])) +
               (hi[2] - lo[2]) * (hi[2] - lo[2]));
}

static std::vector<float> triangle_centroids(const float* verts, const uint32_t* tris, uint32_t nt) {
  std::vector<float> cent((size_t)nt * 3);
  for (uint32_t t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k)
      cent[(size_t)t * 3 + k] = ((verts[(size_t)tris[t * 3] * 3 + k] + verts[(size_t)tris[t * 3 + 1] * 3 + k]) +
                                 verts[(size_t)tris[t * 3 + 2] * 3 + k]) / 3.0f;
  return cent;
}

// A uniform grid over the cage's bounding box (32 cells along the longest
// side) listing, per cell, every triangle whose centroid lies within
// r1 + h + (cell half-diagonal) of the cell centre: a superset of the
// triangles with phi > 0 at any point of the cell or h away from it.
struct NearGrid {
  void* d_cent = nullptr;
  void* d_start = nullptr;
  void* d_list = nullptr;
  float lo[3] = {0.f, 0.f, 0.f}, inv_cell = 0.f;
  uint32_t dim[3] = {1u, 1u, 1u};
  ~NearGrid() {
    for (void* p : {d_cent, d_start, d_list})
      if (p) (void)hipFree(p);
  }
};

static int near_grid_build(cgs_context* c, const float* verts, uint32_t nv, const uint32_t* tris,
                           uint32_t nt, float r1, float h, NearGrid& G) {
  const std::vector<float> cent = triangle_centroids(verts, tris, nt);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; ++v)
    for (int k = 0; k < 3; ++k) {
      lo[k] = cgs_fmin(lo[k], verts[(size_t)v * 3 + k]);
      hi[k] = cgs_fmax(hi[k], verts[(size_t)v * 3 + k]);
    }
  const float ext = cgs_fmax(cgs_fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  if (!(ext > 0.0f)) return fail(CGS_ERR_INVALID_ARG, "degenerate cage bounding box");
  const float cell = ext / 32.0f;
  for (int k = 0; k < 3; ++k) {
    G.lo[k] = lo[k];
    G.dim[k] = (uint32_t)((hi[k] - lo[k]) / cell) + 1u;
  }
  G.inv_cell = 1.0f / cell;
  const size_t cells = (size_t)G.dim[0] * G.dim[1] * G.dim[2];
  const double R = (double)r1 + (double)h + 0.5 * (double)cell * std::sqrt(3.0);
  std::vector<uint32_t> start(cells + 1, 0u), list;
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<uint32_t> fill;
    if (pass == 1) {
      for (size_t i = 0; i < cells; ++i) start[i + 1] += start[i];
      if (start[cells] > (1u << 28)) return fail(CGS_ERR_INVALID_ARG, "near-field lists too large");
      list.resize(start[cells] ? start[cells] : 1);
      fill.assign(start.begin(), start.end() - 1);
    }
    for (uint32_t t = 0; t < nt; ++t) {
      int a0[3], a1[3];
      for (int k = 0; k < 3; ++k) {
        const double ck = cent[(size_t)t * 3 + k];
        a0[k] = std::max(0, (int)std::floor((ck - R - lo[k]) / cell));
        a1[k] = std::min((int)G.dim[k] - 1, (int)std::floor((ck + R - lo[k]) / cell));
      }
      for (int z = a0[2]; z <= a1[2]; ++z)
        for (int y = a0[1]; y <= a1[1]; ++y)
          for (int x = a0[0]; x <= a1[0]; ++x) {
            const double dx = lo[0] + (x + 0.5) * cell - cent[(size_t)t * 3];
            const double dy = lo[1] + (y + 0.5) * cell - cent[(size_t)t * 3 + 1];
            const double dz = lo[2] + (z + 0.5) * cell - cent[(size_t)t * 3 + 2];
            if (dx * dx + dy * dy + dz * dz > R * R) continue;
            const size_t ci = ((size_t)z * G.dim[1] + y) * G.dim[0] + x;
            if (pass == 0)
              start[ci + 1]++;
            else
              list[fill[ci]++] = t;
          }
    }
  }
  HIPCHK(hipMalloc(&G.d_cent, (size_t)nt * 12));
  HIPCHK(hipMalloc(&G.d_start, (cells + 1) * 4));
  HIPCHK(hipMalloc(&G.d_list, list.size() * 4));
  HIPCHK(hipMemcpyAsync(G.d_cent, cent.data(), (size_t)nt * 12, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(G.d_start, start.data(), (cells + 1) * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(G.d_list, list.data(), list.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  return CGS_OK;
}

static int check_near_radii(float r0, float r1) {
  if (!(r0 >= 0.0f && r1 > r0 && r1 < 1e6f))
    return fail(CGS_ERR_INVALID_ARG, "near field radii need 0 <= r0 < r1");
  return CGS_OK;
}

// The far sums' target scale: num_vertices x diagonal / (4 pi), so that a
// far sum of the size of the smallest interior weight sum (4 pi / diagonal)
// spread over the vertices is 1.
static float far_sums_scale(const float* verts, uint32_t nv) {
  return (float)nv * cgs_fmax(cage_diagonal(verts, nv), 1e-6f) * 0.0795774715f;
}

// Training targets of the near/far network: the far share q of the exact
// weights (u - u_near) / sum u, times num_vertices, per point
// (CGS_FAR_SHARE), or the far sums u - u_near times far_sums_scale
// (CGS_FAR_SUMS).
int cgs_mvc_far_targets(cgs_context* c, const float* cage_vertices, uint32_t num_vertices,
                        const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                        float near_r1, const float* d_points, uint64_t n, uint32_t kind,
                        float* d_out) {
  TRY(ensure_ctx(c));
  TRY(check_near_radii(near_r0, near_r1));
  if (kind != CGS_FAR_SHARE && kind != CGS_FAR_SUMS)
    return fail(CGS_ERR_INVALID_ARG, "far target kind must be CGS_FAR_SHARE or CGS_FAR_SUMS");
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  if (!d_points || !d_out) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  MvcCage M;
  TRY(mvc_upload(c, cage_vertices, num_vertices, triangles, num_triangles, M));
  const std::vector<float> cent = triangle_centroids(cage_vertices, triangles, num_triangles);
  void* d_cent = nullptr;
  HIPCHK(hipMalloc(&d_cent, cent.size() * 4));
  hipError_t e = hipMemcpyAsync(d_cent, cent.data(), cent.size() * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    CgsMvcArgs A;
    memset(&A, 0, sizeof(A));
    A.verts = (const float*)M.d_verts;
    A.nv = num_vertices;
    A.tris = (const uint32_t*)M.d_tris;
    A.nt = num_triangles;
    A.vt_start = (const uint32_t*)M.d_start;
    A.vt_list = (const uint32_t*)M.d_list;
    A.pts = d_points;
    A.n = (uint32_t)n;
    const bool sums = kind == CGS_FAR_SUMS;
    cgs_launch_mvc_far_share(A, (const float*)d_cent, near_r0, near_r1,
                             sums ? far_sums_scale(cage_vertices, num_vertices) : (float)num_vertices,
                             sums ? 1 : 0, d_out, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_cent);
  if (e != hipSuccess) return fail(CGS_ERR_EXECUTION, std::string("far share: ") + hipGetErrorString(e));
  return CGS_OK;
}

int cgs_mvc_far_share(cgs_context* c, const float* cage_vertices, uint32_t num_vertices,
                      const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                      float near_r1, const float* d_points, uint64_t n, float* d_out) {
  return cgs_mvc_far_targets(c, cage_vertices, num_vertices, triangles, num_triangles, near_r0,
                             near_r1, d_points, n, CGS_FAR_SHARE, d_out);
}

// The cage buffers and the near-field grid of the last near/far binding,
// kept by the context and reused while the cage (vertices, triangles) and the
// radii stay the same: the grid costs more host time than a small binding.
struct NearCache {
  uint64_t key = 0;
  MvcCage M;
  NearGrid G;
  float minv[16];
  bool parity = false;          // every triangle's circumradius <= r1 + h
  void* d_cell_inside = nullptr;  // [cells]: 1 inside, 0 outside, 2 undecided
  ~NearCache() {
    if (d_cell_inside) (void)hipFree(d_cell_inside);
  }
};

// The side of every grid cell's centre (winding numbers, one point per lane)
// for the crossing-parity inside test, which needs every triangle that meets
// a cell in the cell's list: true when no triangle reaches farther than
// r1 + h from its centroid.
static int near_cells_classify(cgs_context* c, const float* verts, const uint32_t* tris, uint32_t nt,
                               float r1, NearCache& C) {
  const std::vector<float> cent = triangle_centroids(verts, tris, nt);
  float reach = 0.0f;
  for (uint32_t t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) {
      const float* v = verts + (size_t)tris[t * 3 + k] * 3;
      const float dx = v[0] - cent[t * 3], dy = v[1] - cent[t * 3 + 1], dz = v[2] - cent[t * 3 + 2];
      reach = cgs_fmax(reach, sqrtf((dx * dx + dy * dy) + dz * dz));
    }
  C.parity = reach <= r1 + C.M.h;
  if (!C.parity) return CGS_OK;
  const NearGrid& G = C.G;
  const size_t cells = (size_t)G.dim[0] * G.dim[1] * G.dim[2];
  const float cell = 1.0f / G.inv_cell;
  std::vector<float> centres(cells * 3);
  for (uint32_t z = 0; z < G.dim[2]; ++z)
    for (uint32_t y = 0; y < G.dim[1]; ++y)
      for (uint32_t x = 0; x < G.dim[0]; ++x) {
        const size_t ci = ((size_t)z * G.dim[1] + y) * G.dim[0] + x;
        centres[ci * 3] = G.lo[0] + ((float)x + 0.5f) * cell;
        centres[ci * 3 + 1] = G.lo[1] + ((float)y + 0.5f) * cell;
        centres[ci * 3 + 2] = G.lo[2] + ((float)z + 0.5f) * cell;
      }
  void *d_c = nullptr, *d_w = nullptr;
  std::vector<float> w(cells);
  hipError_t e = hipMalloc(&d_c, cells * 12);
  if (e == hipSuccess) e = hipMalloc(&d_w, cells * 4);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_c, centres.data(), cells * 12, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    cgs_launch_mvc_winding_lane((const float*)C.M.d_verts, (const uint32_t*)C.M.d_tris, nt,
                                (const float*)d_c, (uint32_t)cells, (float*)d_w, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(w.data(), d_w, cells * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_c) (void)hipFree(d_c);
  if (d_w) (void)hipFree(d_w);
  if (e != hipSuccess) return fail(CGS_ERR_EXECUTION, std::string("cell classification: ") + hipGetErrorString(e));
  std::vector<uint32_t> side(cells);
  for (size_t i = 0; i < cells; ++i) {
    const float a = fabsf(w[i]);
    side[i] = a >= 0.75f ? 1u : (a <= 0.25f ? 0u : 2u);
  }
  HIPCHK(hipMalloc(&C.d_cell_inside, cells * 4));
  HIPCHK(hipMemcpyAsync(C.d_cell_inside, side.data(), cells * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

static void near_cache_free(NearCache* n) { delete n; }

static uint64_t near_key(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float r0,
                         float r1) {
  uint64_t h = 0xcbf29ce484222325ull;
  auto mix = [&](const void* p, size_t bytes) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t k = 0; k < bytes; ++k) {
      h ^= b[k];
      h *= 0x100000001b3ull;
    }
  };
  mix(&nv, 4);
  mix(&nt, 4);
  mix(&r0, 4);
  mix(&r1, 4);
  mix(verts, (size_t)nv * 12);
  mix(tris, (size_t)nt * 12);
  return h;
}

// Shared set-up of a near/far binding: the cached cage buffers and grid,
// the winding numbers of every point (one point per lane) and the kernel
// arguments.
struct NearSetup {
  NearCache* C = nullptr;  // owned by the context
  void* d_wind = nullptr;
  CgsNearArgs a;
  ~NearSetup() {
    if (d_wind) (void)hipFree(d_wind);
  }
};

static int near_setup(cgs_context* c, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                      float r0, float r1, const float* d_points, uint64_t n, NearSetup& S) {
  TRY(check_near_radii(r0, r1));
  if (nv > CGS_NEARFAR_MAX_VERTICES)
    return fail(CGS_ERR_INVALID_ARG, "the near/far binding takes cages of at most 2048 vertices");
  if (!verts || !tris) return fail(CGS_ERR_INVALID_ARG, "null cage");
  const uint64_t key = near_key(verts, nv, tris, nt, r0, r1);
  if (!c->near_cache || c->near_cache->key != key) {
    near_cache_free(c->near_cache);
    c->near_cache = nullptr;
    NearCache* C = new NearCache();
    int rc = mvc_upload(c, verts, nv, tris, nt, C->M);
    if (rc == CGS_OK) rc = near_grid_build(c, verts, nv, tris, nt, r1, C->M.h, C->G);
    if (rc == CGS_OK) rc = near_cells_classify(c, verts, tris, nt, r1, *C);
    if (rc == CGS_OK && !linear_precision_minv(verts, nv, C->minv))
      rc = fail(CGS_ERR_INVALID_ARG, "cage vertices are coplanar");
    if (rc != CGS_OK) {
      delete C;
      return rc;
    }
    C->key = key;
    c->near_cache = C;
  }
  NearCache* C = c->near_cache;
  S.C = C;
  memset(&S.a, 0, sizeof(S.a));
  memcpy(S.a.minv, C->minv, sizeof(S.a.minv));
  if (!C->parity) {
    HIPCHK(hipMalloc(&S.d_wind, n * 4));
    cgs_launch_mvc_winding_lane((const float*)C->M.d_verts, (const uint32_t*)C->M.d_tris, nt, d_points,
                                (uint32_t)n, (float*)S.d_wind, c->stream);
  }
  CgsNearArgs& a = S.a;
  a.cell_inside = (const uint32_t*)C->d_cell_inside;
  a.verts = (const float*)C->M.d_verts;
  a.nv = nv;
  a.tris = (const uint32_t*)C->M.d_tris;
  a.nt = nt;
  a.cent = (const float*)C->G.d_cent;
  a.cell_start = (const uint32_t*)C->G.d_start;
  a.cell_list = (const uint32_t*)C->G.d_list;
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = C->G.lo[k];
    a.dim[k] = C->G.dim[k];
  }
  a.inv_cell = C->G.inv_cell;
  a.r0 = r0;
  a.r1 = r1;
  a.h = C->M.h;
  // every interior point has sum u >= 4 pi / diagonal; tau is 1e-4 of that,
  // so the near term is short by at most 5e-5 of the weights' sum where it
  // fades out (the projection restores the sum)
  a.tau = 1.2566371e-3f / cgs_fmax(cage_diagonal(verts, nv), 1e-6f);
  a.sums_scale = far_sums_scale(verts, nv);
  a.pts = d_points;
  a.n = (uint32_t)n;
  a.winding = (const float*)S.d_wind;
  return CGS_OK;
}

// The near/far combination with caller-supplied far targets: d_q [n x nv]
// (as cgs_mvc_far_targets writes them for `kind`) and d_dq [3n x nv] (row
// 3p + axis, the derivatives of those rows) or NULL.  Outputs as
// cgs_mvc_bind_exact plus d_residual [n] (|sum w v - x| before the
// projection; +inf where the near sums meet a vertex or a face).
int cgs_mvc_nearfar_bind_kind(cgs_context* c, const float* cage_vertices, uint32_t num_vertices,
                              const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                              float near_r1, const float* d_points, uint64_t n, const float* d_q,
                              const float* d_dq, uint32_t kind, float* d_weights,
                              float* d_gradients, uint32_t* d_inside, float* d_residual) {
  TRY(ensure_ctx(c));
  if (kind != CGS_FAR_SHARE && kind != CGS_FAR_SUMS)
    return fail(CGS_ERR_INVALID_ARG, "far target kind must be CGS_FAR_SHARE or CGS_FAR_SUMS");
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  if (!d_points || !d_q || !d_weights || !d_residual) return fail(CGS_ERR_INVALID_ARG, "null buffer");
  NearSetup S;
  TRY(near_setup(c, cage_vertices, num_vertices, triangles, num_triangles, near_r0, near_r1, d_points,
                 n, S));
  CgsNearArgs a = S.a;
  a.Q = d_q;
  a.DQ = d_dq;
  a.ldq = num_vertices;
  a.far_sums = kind == CGS_FAR_SUMS ? 1u : 0u;
  a.q_scale = a.far_sums ? a.sums_scale : (float)num_vertices;
  a.out_w = d_weights;
  a.out_g = d_gradients;
  a.residual = d_residual;
  a.inside = d_inside;
  cgs_launch_mvc_nearfar(a, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return CGS_OK;
}

int cgs_mvc_nearfar_bind(cgs_context* c, const float* cage_vertices, uint32_t num_vertices,
                         const uint32_t* triangles, uint32_t num_triangles, float near_r0,
                         float near_r1, const float* d_points, uint64_t n, const float* d_q,
                         const float* d_dq, float* d_weights, float* d_gradients,
                         uint32_t* d_inside, float* d_residual) {
  return cgs_mvc_nearfar_bind_kind(c, cage_vertices, num_vertices, triangles, num_triangles,
                                   near_r0, near_r1, d_points, n, d_q, d_dq, CGS_FAR_SHARE,
                                   d_weights, d_gradients, d_inside, d_residual);
}

// The surrogate binding: the network's weights and gradients for every
// point, exact MVC (the reference routine) for every point whose
// reconstruction residual |sum_j w_j v_j - x| exceeds residual_tol times the
// cage's bounding-box diagonal, zero rows outside the cage (same inside test
// as the exact binding).  Same outputs as cgs_mvc_bind_exact.  A network
// with a near field (cgs_mlp_set_near_field) gives the far share only: the
// near sums are evaluated exactly per point (k_mvc_nearfar) and the inside
// test runs one point per lane.
int cgs_mlp_bind(cgs_context* c, struct cgs_mlp* net, const float* cage_vertices,
                 uint32_t num_vertices, const uint32_t* triangles, uint32_t num_triangles,
                 const float* d_points, uint64_t n, float residual_tol, float* d_weights,
                 float* d_gradients, uint32_t* d_inside, uint64_t* fallback_count) {
  TRY(ensure_ctx(c));
  if (!net || !d_points || !d_weights || !d_inside)
    return fail(CGS_ERR_INVALID_ARG, "null argument");
  if (cgs_mlp_outputs(net) != num_vertices)
    return fail(CGS_ERR_INVALID_ARG, "the network's outputs do not match the cage's vertices");
  if (n == 0) {
    if (fallback_count) *fallback_count = 0;
    return CGS_OK;
  }
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  float r0 = 0.0f, r1 = 0.0f;
  const bool near = cgs_mlp_near(net, &r0, &r1);
  const float diag = cage_diagonal(cage_vertices, num_vertices);
  NearSetup S;
  MvcCage M0;
  MvcCage* M = &M0;
  void *d_wind = nullptr, *d_res = nullptr, *d_list = nullptr, *d_count = nullptr;
  auto cleanup = [&]() {
    for (void* p : {d_wind, d_res, d_list, d_count})
      if (p) (void)hipFree(p);
  };
  int rc = CGS_OK;
  do {
    if (hipMalloc(&d_res, n * 4) != hipSuccess || hipMalloc(&d_list, n * 4) != hipSuccess ||
        hipMalloc(&d_count, 4) != hipSuccess) {
      rc = fail(CGS_ERR_RUNTIME, "hipMalloc failed (surrogate binding)");
      break;
    }
    const float* wind = nullptr;  // null: the inside flags come from the near/far kernel
    if (near) {
      if ((rc = near_setup(c, cage_vertices, num_vertices, triangles, num_triangles, r0, r1, d_points,
                           n, S)) != CGS_OK)
        break;
      M = &S.C->M;
      CgsNearArgs a = S.a;
      a.residual = (float*)d_res;
      a.inside = d_inside;
      if ((rc = cgs_mlp_infer_nearfar(net, d_points, n, d_weights, d_gradients, a)) != CGS_OK) break;
    } else {
      if ((rc = mvc_upload(c, cage_vertices, num_vertices, triangles, num_triangles, M0)) != CGS_OK) break;
      if (hipMalloc(&d_wind, n * 4) != hipSuccess) {
        rc = fail(CGS_ERR_RUNTIME, "hipMalloc failed (surrogate binding)");
        break;
      }
      CgsMvcArgs W;
      memset(&W, 0, sizeof(W));
      W.verts = (const float*)M0.d_verts;
      W.nv = num_vertices;
      W.tris = (const uint32_t*)M0.d_tris;
      W.nt = num_triangles;
      W.pts = d_points;
      W.n = (uint32_t)n;
      cgs_launch_mvc_winding(W, (float*)d_wind, c->stream);
      wind = (const float*)d_wind;
      float minv[16];
      if (!linear_precision_minv(cage_vertices, num_vertices, minv)) {
        rc = fail(CGS_ERR_INVALID_ARG, "cage vertices are coplanar");
        break;
      }
      rc = cgs_mlp_infer_residual(net, d_points, n, d_weights, d_gradients, (const float*)M0.d_verts,
                                  minv, (float*)d_res);
      if (rc != CGS_OK) break;
    }
    // points whose residual fails the test, compacted into a list, take the
    // exact routine
    hipError_t e = hipMemsetAsync(d_count, 0, 4, c->stream);
    cgs_launch_mvc_select(wind, (const float*)d_res, residual_tol * diag, (uint32_t)n, num_vertices,
                          d_inside, d_weights, d_gradients, (uint32_t*)d_list, (uint32_t*)d_count,
                          c->stream);
    uint32_t count = 0;
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&count, d_count, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && count) {
      CgsMvcArgs A;
      memset(&A, 0, sizeof(A));
      A.verts = (const float*)M->d_verts;
      A.nv = num_vertices;
      A.tris = (const uint32_t*)M->d_tris;
      A.nt = num_triangles;
      A.vt_start = (const uint32_t*)M->d_start;
      A.vt_list = (const uint32_t*)M->d_list;
      A.pts = d_points;
      A.n = count;
      A.h = M->h;
      A.out_w = d_weights;
      A.out_g = d_gradients;
      A.mask = d_inside;
      A.list = (const uint32_t*)d_list;
      cgs_launch_mvc_bind(A, c->stream);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
      rc = fail(CGS_ERR_EXECUTION, std::string("surrogate binding: ") + hipGetErrorString(e));
      break;
    }
    if (fallback_count) *fallback_count = count;
  } while (false);
  (void)hipStreamSynchronize(c->stream);
  cleanup();
  return rc;
}

// evaluate_position + MvcBindings::local_jacobian (mvc.rs:66-78, 473-481) for
// n bound points: d_out_positions [3n] (rest position where not inside),
// d_out_jacobians [9n] column-major or NULL (needs d_gradients).  The deformed
// cage vertices come from the host (num_vertices x 3).
int cgs_mvc_deform(cgs_context* c, const float* d_weights, const float* d_gradients,
                   const uint32_t* d_inside, const float* d_rest_positions,
                   const float* deformed_vertices, uint32_t num_vertices, uint64_t n,
                   float* d_out_positions, float* d_out_jacobians) {
  TRY(ensure_ctx(c));
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 31)) return fail(CGS_ERR_INVALID_ARG, "too many points");
  if (!d_weights || !d_rest_positions || !deformed_vertices || !d_out_positions)
    return fail(CGS_ERR_INVALID_ARG, "null buffer");
  if (d_out_jacobians && !d_gradients)
    return fail(CGS_ERR_INVALID_ARG, "jacobians need the weight gradients");
  void* dv = nullptr;
  HIPCHK(hipMalloc(&dv, (size_t)num_vertices * 12));
  hipError_t e = hipMemcpyAsync(dv, deformed_vertices, (size_t)num_vertices * 12,
                                hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    cgs_launch_mvc_deform(d_weights, d_gradients, d_inside, d_rest_positions, (const float*)dv,
                          num_vertices, (uint32_t)n, d_out_positions, d_out_jacobians, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dv);
  if (e != hipSuccess) return fail(CGS_ERR_EXECUTION, std::string("mvc deform: ") + hipGetErrorString(e));
  return CGS_OK;
}



int cgs_mvc_deform_scene(cgs_context* c, const float* d_weights, const float* d_gradients,
                         const uint32_t* d_inside, const float* deformed_vertices,
                         uint32_t num_vertices, uint64_t n, const float* d_rest_pos,
                         const float* d_rest_cov, const float* d_rest_opa, const float* d_rest_sh,
                         uint32_t sh_degree, float split_threshold, float mip_filter_3d_var,
                         float* d_out_pos, float* d_out_cov, float* d_out_opa, float* d_out_sh,
                         uint64_t* out_count) {
  TRY(ensure_ctx(c));
  if (!out_count) return fail(CGS_ERR_INVALID_ARG, "null out_count");
  *out_count = 0;
  if (n == 0) return CGS_OK;
  if (n >= (1ull << 30)) return fail(CGS_ERR_INVALID_ARG, "n too large");
  if (!d_weights || !d_gradients || !deformed_vertices || num_vertices == 0 || !d_rest_pos ||
      !d_rest_cov || !d_rest_opa || !d_rest_sh || !d_out_pos || !d_out_cov || !d_out_opa ||
      !d_out_sh)
    return fail(CGS_ERR_INVALID_ARG, "null argument (the weight gradients are required)");
  if (reinterpret_cast<uintptr_t>(d_out_sh) % 16u != 0u)
    return fail(CGS_ERR_INVALID_ARG, "d_out_sh must be 16-byte aligned");
  const size_t nn = (size_t)n;
  // working set (4-byte words, grown 1.5x and kept by the context): deformed
  // vertices [3 nv], x' [3n], F [9n], split flags [n], ranks [n], second
  // children [10n]
  const size_t nv_words = ((size_t)num_vertices * 3 + 63) & ~(size_t)63;
  TRY(ensure_words(c->mvc_work, nv_words + nn * (3 + 9 + 1 + 1 + 10)));
  float* dv = P<float>(c->mvc_work);
  float* def_pos = dv + nv_words;
  float* def_F = def_pos + nn * 3;
  uint32_t* flags = reinterpret_cast<uint32_t*>(def_F + nn * 9);
  uint32_t* ranks = flags + nn;
  float* child_b = reinterpret_cast<float*>(ranks + nn);
  int rc = CGS_OK;
  hipError_t e = hipMemcpyAsync(dv, deformed_vertices, (size_t)num_vertices * 12,
                                hipMemcpyHostToDevice, c->stream);
  // per-launch device times of this call, for cgs_kernel_times
  c->prof.on = true;
  c->prof.n = 0;
  c->prof_count = 0;
  if (e == hipSuccess) {
    // x' = sum w v', F = sum v' grad(w)^T (evaluate_position, local_jacobian)
    c->prof.mark("k_mvc_deform", c->stream);
    cgs_launch_mvc_deform(d_weights, d_gradients, d_inside, d_rest_pos, dv, num_vertices,
                          (uint32_t)n, def_pos, def_F, c->stream);
    CgsMvcSceneArgs A;
    A.mask = d_inside;
    A.def_pos = def_pos;
    A.def_F = def_F;
    A.rest_pos = d_rest_pos;
    A.rest_cov = d_rest_cov;
    A.rest_opa = d_rest_opa;
    A.rest_sh = d_rest_sh;
    A.n = (uint32_t)n;
    A.sh_degree = sh_degree > 3 ? 3 : sh_degree;
    A.tau = split_threshold;
    A.mip = mip_filter_3d_var;
    A.out_pos = d_out_pos;
    A.out_cov = d_out_cov;
    A.out_opa = d_out_opa;
    A.out_sh = d_out_sh;
    A.flags = flags;
    A.child_b = child_b;
    c->prof.mark("k_mvc_scene", c->stream);
    cgs_launch_mvc_scene(A, c->stream);
    e = hipGetLastError();
    uint32_t splits = 0;
    c->prof.mark("k_scan", c->stream);
    if (e == hipSuccess) rc = cgs_exclusive_scan_u32(c, flags, ranks, n, &splits);
    if (e == hipSuccess && rc == CGS_OK && splits) {
      c->prof.mark("k_mvc_scene_children", c->stream);
      cgs_launch_mvc_scene_children(A, ranks, c->stream);
      e = hipGetLastError();
    }
    c->prof.mark("end", c->stream);
    if (e == hipSuccess && rc == CGS_OK) {
      e = hipStreamSynchronize(c->stream);
      if (e == hipSuccess) *out_count = n + splits;
    }
    if (e == hipSuccess && rc == CGS_OK) {
      for (int i = 0; i + 1 < c->prof.n; ++i) {
        c->prof_names[i] = c->prof.names[i];
        c->prof_ms[i] = elapsed(c->prof.ev[i], c->prof.ev[i + 1]);
      }
      c->prof_count = c->prof.n > 0 ? c->prof.n - 1 : 0;
    }
  }
  c->prof.on = false;
  if (rc != CGS_OK) return rc;
  if (e != hipSuccess)
    return fail(CGS_ERR_EXECUTION, std::string("mvc deform scene: ") + hipGetErrorString(e));
  return CGS_OK;
}

}  // extern "C"

int cgs_fail(int code, const char* msg) { return fail(code, msg ? msg : "error"); }
int cgs_context_enter(cgs_context* c) { return ensure_ctx(c); }
void cgs_context_prof_begin(cgs_context* c) {
  c->prof.on = true;
  c->prof.n = 0;
  c->prof_count = 0;
}
void cgs_context_prof_mark(cgs_context* c, const char* name) { c->prof.mark(name, c->stream); }
void cgs_context_prof_end(cgs_context* c, bool ok) {
  if (ok && c->prof.on) {
    c->prof.mark("end", c->stream);
    if (hipStreamSynchronize(c->stream) == hipSuccess) {
      for (int i = 0; i + 1 < c->prof.n; ++i) {
        c->prof_names[i] = c->prof.names[i];
        c->prof_ms[i] = elapsed(c->prof.ev[i], c->prof.ev[i + 1]);
      }
      c->prof_count = c->prof.n > 0 ? c->prof.n - 1 : 0;
    }
  }
  c->prof.on = false;
}
int cgs_context_staging(cgs_context* c, uint64_t bytes, void** p) {
  TRY(ensure_ctx(c));
  TRY(ensure_words(c->ingest, (size_t)((bytes + 3u) / 4u)));
  *p = c->ingest.p;
  return CGS_OK;
}
hipStream_t cgs_context_stream(cgs_context* c) { return c ? c->stream : nullptr; }


// ===========================================================================
// View batch: many views of several resident scenes from one HIP graph
// ===========================================================================
//
// The reference renders one view per execute_frame_gpu call (gpu_state.rs:
// 613-1243), waiting for its counts between stages.  A batch renders n views
// of up to n scenes with no host round trip between them:
//   * Stage 1 of every scene runs first, on the scene's own stream (it is
//     shared by all views of the scene);
//   * Stages 2-4 of every view are captured once into a HIP graph and
//     replayed while the views, the scenes' device buffers and the batch's own
//     buffers are unchanged (a fingerprint of all of them keys the graph);
//   * the views are dealt to `lanes` arenas, longest first (an estimate of
//     splats + pixels/2), and each lane is a chain of views on its own branch
//     of the graph, so the latency-bound steps of one view (look-back sort
//     passes, the tile-order block, launch gaps) overlap the others;
//   * each view composites into its own framebuffer and leaves its tile-entry
//     total and survivor count in a device slot; these are read once after
//     the replay, and a view whose entries overflowed its lane's arena has
//     the arena grown (1.5x) and the batch replayed.
// A view's pixels are those cgs_execute_frame gives for the same camera: the
// kernels and their inputs are the same.

struct cgs_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  std::vector<hipStream_t> lane_stream;
  std::vector<hipEvent_t> lane_done;
  std::vector<FrameArena> lanes;
  DevBuf fb;        // every view's framebuffer, 256-byte aligned slices
  DevBuf counters;  // [view][8]: control words 0-3 of the view's Stage 3, then
                    // its bucket-mode stats (zero in the LSD mode)
  uint32_t* h_counters = nullptr;
  size_t h_counter_words = 0;
  std::vector<size_t> fb_off;  // words
  std::vector<uint32_t> fb_w, fb_h, fb_row, view_lane;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t key = 0;
  CgsProf noprof;  // never on
  // multi-view projection: every view's projected records (geom 8, color 4,
  // extents 1 words per splat slot) and the per-view descriptors, grouped by
  // (scene, SH degree) in runs of at most 64 views
  bool shared_projection = true;
  DevBuf proj, vproj;
  std::vector<size_t> proj_off;  // words, per view
  struct Group {
    uint32_t scene, first, count, view0;  // descriptors [first, first + count)
  };
  std::vector<Group> groups;
  std::vector<uint32_t> vproj_index;  // view -> descriptor
  // tile entries of each view in the last replay of the same views, used to
  // deal the views to lanes by their measured weight
  uint64_t views_hash = 0;
  std::vector<uint32_t> last_entries;
  // per view, the segmented sort's needs in the last replay (bucket mode)
  // and what the captured graph launches
  std::vector<uint8_t> plan_known, view_bucket;
  // views sorted with the depth presort (see cgs_execute_views_batched)
  std::vector<uint8_t> view_presort;
  std::vector<uint32_t> plan_chunks, plan_rounds, launched_chunks, launched_rounds;
};

namespace {

void batch_drop_graph(cgs_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->key = 0;
}

// every buffer the views of a lane touch in Stages 2-4: up to max_g splat
// slots and T tiles, and in bucket mode up to bucket_g slots and bucket_T
// tiles (the views of at most CGS_BUCKET_MAX_TILES tiles; 0: none)
int size_arena(FrameArena& a, uint32_t max_g, uint32_t T, bool mode64, bool own_records,
               uint32_t bucket_g, uint32_t bucket_T) {
  init_entry_cap(a, max_g);
  const uint32_t cap = a.entry_cap;
  TRY(ensure_words(a.geom, own_records ? (size_t)8 * max_g : 8));
  TRY(ensure_words(a.color, own_records ? (size_t)4 * max_g : 4));
  TRY(ensure_words(a.extents, own_records ? max_g : 1));
  TRY(ensure_words(a.cnt_dbg, max_g));
  TRY(ensure_words(a.hitm, cgs_bin_mask_words(max_g)));
  TRY(ensure_words(a.e_keys, cap));
  TRY(ensure_words(a.e_ids, cap));
  TRY(ensure_words(a.e_alt_k, cap));
  TRY(ensure_words(a.e_alt_v, cap));
  if (mode64) {
    TRY(ensure_words(a.e_depths, cap));
    TRY(ensure_words(a.e_perm, cap));
  }
  TRY(ensure_words(a.ranges, (size_t)T * 2));
  TRY(ensure_words(a.comp_order, T));
  TRY(ensure_words(a.ctrl3, frame_layout(max_g, cap).words));
  if (bucket_T != 0 && bucket_mode(a, bucket_T, bucket_g)) TRY(size_bucket(a, bucket_g, bucket_T));
  if (a.tune.tile_sort == 2u) TRY(size_presort(a, max_g));
  return CGS_OK;
}

uint64_t fnv_bytes(uint64_t h, const void* p, size_t n) { return fnv1a(h, p, n); }

uint64_t fnv_buf(uint64_t h, const DevBuf& d) {
  h = fnv_u64(h, (uint64_t)(uintptr_t)d.p);
  return fnv_u64(h, d.cap_words);
}

// Stages 2-4 of view v on lane stream s (no host synchronisation)
int batch_enqueue_view(cgs_batch* b, cgs_context* c, const SceneFrame& F, FrameArena& a,
                       hipStream_t s, const cgs_camera_uniforms* cam,
                       const cgs_compositor_uniforms* comp, const cgs_frame_params* prm,
                       uint32_t v) {
  const FrameGeom g = frame_geom(comp, a.tune);
  uint32_t* out = P<uint32_t>(b->fb) + b->fb_off[v];
  uint32_t* slot = P<uint32_t>(b->counters) + 8 * (size_t)v;
  b->view_bucket[v] = 0;
  if (F.N == 0) {
    // background frame (gpu_state.rs:658-704)
    TRY(background_arena(a, g, s));
    cgs_launch_fill_u32(slot, 0u, 8, s);
    return launch_stage4(a, s, &b->noprof, g, comp, out);
  }
  if (b->shared_projection) {
    // the view's records were written by the multi-view projection
    uint32_t* rec = P<uint32_t>(b->proj) + b->proj_off[v];
    a.geom_in = reinterpret_cast<const float*>(rec);
    a.color_in = reinterpret_cast<const float*>(rec + (size_t)8 * F.max_g);
    a.extents_in = rec + (size_t)12 * F.max_g;
  } else {
    TRY(stage2(c, F, a, s, &b->noprof, cam, comp, prm));
  }
  // the segmented sort's plan is this view's, from its last replay
  a.plan_external = true;
  a.bk_planned = b->plan_known[v] != 0;
  a.bk_need_chunks = b->plan_chunks[v];
  a.bk_need_rounds = b->plan_rounds[v];
  int rc = launch_stage3(a, s, &b->noprof, g, cam->near_plane, F.N, F.split_count, F.max_g);
  if (rc == CGS_OK) {
    cgs_launch_copy_words(P<uint32_t>(a.ctrl3), slot, 4, s);
    if (a.bucket) {
      cgs_launch_copy_words(P<uint32_t>(a.bk_ctl), slot + 4, 4, s);
      b->view_bucket[v] = 1;
      b->launched_chunks[v] = a.bk_chunk_blocks;
      b->launched_rounds[v] = a.bk_rounds;
    } else {
      cgs_launch_fill_u32(slot + 4, 0u, 4, s);
    }
    rc = launch_stage4(a, s, &b->noprof, g, comp, out);
  }
  a.geom_in = a.color_in = nullptr;
  a.extents_in = nullptr;
  return rc;
}

int batch_capture(cgs_batch* b, cgs_context* const* scenes, const std::vector<SceneFrame>& SF,
                  const uint32_t* view_scene, const cgs_camera_uniforms* cams,
                  const cgs_compositor_uniforms* comps, uint32_t n_views,
                  const cgs_frame_params* prm) {
  batch_drop_graph(b);
  const uint32_t L = (uint32_t)b->lanes.size();
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  g_no_alloc = true;
  int rc = CGS_OK;
  hipError_t e = hipSuccess;
  if (b->shared_projection) {
    for (const cgs_batch::Group& G : b->groups) {
      if (rc != CGS_OK) break;
      CgsProjectArgs PA;
      rc = project_args(scenes[G.scene], SF[G.scene], &comps[G.view0], prm, PA);
      if (rc == CGS_OK)
        cgs_launch_project_views(PA, reinterpret_cast<const CgsViewProjection*>(b->vproj.p) + G.first,
                                 G.count, b->stream);
    }
  }
  if (rc == CGS_OK) e = hipEventRecord(b->fork, b->stream);
  for (uint32_t k = 0; k < L && e == hipSuccess && rc == CGS_OK; ++k)
    e = hipStreamWaitEvent(b->lane_stream[k], b->fork, 0);
  if (e != hipSuccess && rc == CGS_OK)
    rc = fail(CGS_ERR_RUNTIME, std::string("graph fork: ") + hipGetErrorString(e));
  for (uint32_t v = 0; v < n_views && rc == CGS_OK; ++v) {
    const uint32_t k = b->view_lane[v], si = view_scene[v];
    FrameArena& a = b->lanes[k];
    const uint32_t lane_sort = a.tune.tile_sort;
    if (b->view_presort[v]) a.tune.tile_sort = 2u;
    rc = batch_enqueue_view(b, scenes[si], SF[si], a, b->lane_stream[k], &cams[v], &comps[v],
                            prm, v);
    a.tune.tile_sort = lane_sort;
  }
  for (uint32_t k = 0; k < L && rc == CGS_OK; ++k) {
    e = hipEventRecord(b->lane_done[k], b->lane_stream[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->lane_done[k], 0);
    if (e != hipSuccess) rc = fail(CGS_ERR_RUNTIME, std::string("graph join: ") + hipGetErrorString(e));
  }
  g_no_alloc = false;
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->stream, &graph);
  if (rc != CGS_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ee != hipSuccess || !graph)
    return fail(CGS_ERR_RUNTIME, std::string("graph capture: ") + hipGetErrorString(ee));
  b->graph = graph;
  const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    batch_drop_graph(b);
    return fail(CGS_ERR_RUNTIME, std::string("graph instantiate: ") + hipGetErrorString(ei));
  }
  return CGS_OK;
}

}  // namespace

extern "C" {

int cgs_batch_create(int device, uint32_t lanes, cgs_batch** out) {
  if (!out) return fail(CGS_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  if (lanes < 1 || lanes > 16) return fail(CGS_ERR_INVALID_ARG, "lanes must be in [1, 16]");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(CGS_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(CGS_ERR_INVALID_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  cgs_batch* b = new cgs_batch();
  b->device = device;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreate(&b->t0) == hipSuccess && hipEventCreate(&b->t1) == hipSuccess;
  b->lanes.resize(lanes);
  b->lane_stream.assign(lanes, nullptr);
  b->lane_done.assign(lanes, nullptr);
  for (uint32_t k = 0; k < lanes && ok; ++k)
    ok = hipStreamCreateWithFlags(&b->lane_stream[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->lane_done[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    cgs_batch_destroy(b);
    return fail(CGS_ERR_RUNTIME, "stream or event creation failed");
  }
  *out = b;
  return CGS_OK;
}

void cgs_batch_destroy(cgs_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (FrameArena& a : b->lanes) a.release_all();
  release(b->fb);
  release(b->counters);
  release(b->proj);
  release(b->vproj);
  if (b->h_counters) (void)hipHostFree(b->h_counters);
  for (hipStream_t s : b->lane_stream)
    if (s) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->lane_done)
    if (e) (void)hipEventDestroy(e);
  hipEvent_t evs[] = {b->fork, b->t0, b->t1};
  for (hipEvent_t e : evs)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int cgs_execute_views_batched(cgs_batch* b, cgs_context* const* scenes, uint32_t n_scenes,
                              const uint32_t* view_scene, const cgs_camera_uniforms* cams,
                              const cgs_compositor_uniforms* comps, uint32_t n_views,
                              const cgs_frame_params* prm, cgs_batch_stats* stats) {
  if (!b || !scenes || !prm || (n_views && (!view_scene || !cams || !comps)))
    return fail(CGS_ERR_INVALID_ARG, "null batch argument");
  if (n_scenes == 0 && n_views) return fail(CGS_ERR_INVALID_ARG, "views without scenes");
  HIPCHK(hipSetDevice(b->device));
  cgs_batch_stats St;
  memset(&St, 0, sizeof(St));
  St.views = n_views;
  St.lanes = (uint32_t)b->lanes.size();
  for (uint32_t i = 0; i < n_scenes; ++i) {
    if (!scenes[i]) return fail(CGS_ERR_INVALID_ARG, "null scene context");
    if (scenes[i]->device != b->device) return fail(CGS_ERR_INVALID_ARG, "scene on another device");
  }
  for (uint32_t v = 0; v < n_views; ++v)
    if (view_scene[v] >= n_scenes) return fail(CGS_ERR_INVALID_ARG, "view scene index out of range");

  // ---- Stage 1 of every scene, on its own stream ----
  std::vector<SceneFrame> SF(n_scenes);
  for (uint32_t i = 0; i < n_scenes; ++i) {
    cgs_context* c = scenes[i];
    if (c->n_rest == 0) continue;
    bool deformed = false;
    TRY(stage1(c, prm, c->stream, &b->noprof, SF[i], deformed));
    HIPCHK(hipStreamSynchronize(c->stream));
  }

  // ---- views -> lanes (longest first onto the least loaded lane) ----
  // weight of a view in tile-entry units: its entries (measured in the last
  // replay of the same views, else estimated as 3 per splat), plus pixels / 2
  // and splats / 5 for the per-pixel and per-splat work
  uint64_t vh = FNV_OFFSET;
  vh = fnv_u64(vh, n_views);
  vh = fnv_bytes(vh, view_scene, 4 * (size_t)n_views);
  vh = fnv_bytes(vh, cams, sizeof(cgs_camera_uniforms) * n_views);
  vh = fnv_bytes(vh, comps, sizeof(cgs_compositor_uniforms) * n_views);
  for (uint32_t i = 0; i < n_scenes; ++i) vh = fnv_u64(vh, SF[i].max_g);
  const bool measured = vh == b->views_hash && b->last_entries.size() == n_views;
  const uint32_t L = (uint32_t)b->lanes.size();
  std::vector<FrameGeom> G(n_views);
  std::vector<double> cost(n_views);
  std::vector<uint32_t> order(n_views);
  for (uint32_t v = 0; v < n_views; ++v) {
    G[v] = frame_geom(&comps[v], b->lanes[0].tune);
    const double mg = (double)SF[view_scene[v]].max_g;
    const double e = measured ? (double)b->last_entries[v] : 3.0 * mg;
    cost[v] = e + 0.5 * (double)G[v].W * G[v].H + 0.2 * mg;
    order[v] = v;
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  std::vector<double> load(L, 0.0);
  b->view_lane.assign(n_views, 0);
  for (uint32_t v : order) {
    uint32_t best = 0;
    for (uint32_t k = 1; k < L; ++k)
      if (load[k] < load[best]) best = k;
    b->view_lane[v] = best;
    load[best] += cost[v];
  }
  // with the LSD sort, a view whose entries outnumber its splat slots more
  // than 4 to 1 in the last replay is sorted with the splats presorted by
  // depth: 3-4 passes over the splats and 2-3 over the entries' tile bits
  // instead of 4-7 over the entries (measured: 4K views of the 4M-splat
  // scene, 16 entries per splat, and the 8K frame, 9 per splat, gain; 1080p
  // views, 2.3 per splat, lose)
  b->view_presort.assign(n_views, 0);
  for (uint32_t v = 0; v < n_views; ++v) {
    const uint32_t mg = SF[view_scene[v]].max_g;
    if (measured && b->lanes[b->view_lane[v]].tune.tile_sort == 0u && SF[view_scene[v]].N != 0 &&
        (double)b->last_entries[v] > 4.0 * (double)mg) {
      b->view_presort[v] = 1;
      St.presort_views++;
    }
  }
  // ---- framebuffers and counter slots ----
  b->fb_off.assign(n_views, 0);
  b->fb_w.assign(n_views, 0);
  b->fb_h.assign(n_views, 0);
  b->fb_row.assign(n_views, 0);
  size_t words = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    b->fb_off[v] = words;
    b->fb_w[v] = G[v].W;
    b->fb_h[v] = G[v].H;
    b->fb_row[v] = G[v].padded_row;
    words += ((size_t)G[v].padded_row * G[v].H + 63) / 64 * 64;
  }
  TRY(ensure_words(b->fb, words));
  TRY(ensure_words(b->counters, 8 * (size_t)(n_views ? n_views : 1)));
  if (b->h_counter_words < 8 * (size_t)n_views) {
    if (b->h_counters) (void)hipHostFree(b->h_counters);
    b->h_counters = nullptr;
    b->h_counter_words = 0;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->h_counters), 32 * (size_t)n_views,
                         hipHostMallocDefault));
    b->h_counter_words = 8 * (size_t)n_views;
  }
  if (b->plan_known.size() != n_views || b->views_hash != vh) {
    b->plan_known.assign(n_views, 0);
    b->plan_chunks.assign(n_views, 0);
    b->plan_rounds.assign(n_views, 0);
  }
  b->view_bucket.assign(n_views, 0);
  b->launched_chunks.assign(n_views, 0);
  b->launched_rounds.assign(n_views, 0);
  if (n_views == 0) {
    if (stats) *stats = St;
    return CGS_OK;
  }

  // ---- multi-view projection: records per view, descriptors per group ----
  std::vector<CgsViewProjection> desc;
  if (b->shared_projection) {
    b->proj_off.assign(n_views, 0);
    size_t pw = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
      b->proj_off[v] = pw;
      const uint32_t mg = SF[view_scene[v]].max_g;
      pw += ((size_t)13 * mg + 63) / 64 * 64;
    }
    TRY(ensure_words(b->proj, pw));
    b->groups.clear();
    b->vproj_index.assign(n_views, 0);
    // views of one scene and SH degree, in view order, runs of <= 64
    std::vector<uint32_t> byg(n_views);
    for (uint32_t v = 0; v < n_views; ++v) byg[v] = v;
    std::stable_sort(byg.begin(), byg.end(), [&](uint32_t x, uint32_t y) {
      const uint32_t dx = std::min(comps[x].sh_degree, 3u), dy = std::min(comps[y].sh_degree, 3u);
      return view_scene[x] != view_scene[y] ? view_scene[x] < view_scene[y] : dx < dy;
    });
    for (uint32_t v : byg) {
      const uint32_t si = view_scene[v];
      if (SF[si].N == 0) continue;
      const uint32_t deg = std::min(comps[v].sh_degree, 3u);
      cgs_batch::Group* G = b->groups.empty() ? nullptr : &b->groups.back();
      if (!G || G->scene != si || std::min(comps[G->view0].sh_degree, 3u) != deg ||
          G->count == 64u) {
        b->groups.push_back({si, (uint32_t)desc.size(), 0u, v});
        G = &b->groups.back();
      }
      CgsViewProjection d;
      memset(&d, 0, sizeof(d));
      memcpy(d.camera, &cams[v], 96);
      for (int i = 0; i < 3; ++i) d.cam_pos[i] = comps[v].camera_position[i];
      uint32_t* rec = P<uint32_t>(b->proj) + b->proj_off[v];
      const uint32_t mg = SF[si].max_g;
      d.geom = reinterpret_cast<float*>(rec);
      d.color = reinterpret_cast<float*>(rec + (size_t)8 * mg);
      d.extents = rec + (size_t)12 * mg;
      b->vproj_index[v] = (uint32_t)desc.size();
      desc.push_back(d);
      G->count++;
    }
    TRY(ensure_words(b->vproj, desc.size() * sizeof(CgsViewProjection) / 4 + 1));
  }

  for (int attempt = 0;; ++attempt) {
    // ---- size every lane for its largest view ----
    for (uint32_t k = 0; k < L; ++k) {
      uint32_t mg = 1, T = 1, bg = 0, bT = 0;
      bool m64 = false;
      for (uint32_t v = 0; v < n_views; ++v) {
        if (b->view_lane[v] != k) continue;
        mg = std::max(mg, SF[view_scene[v]].max_g);
        T = std::max(T, G[v].T);
        m64 = m64 || G[v].T > 65535u;
        if (G[v].T <= CGS_BUCKET_MAX_TILES && SF[view_scene[v]].N != 0) {
          bg = std::max(bg, SF[view_scene[v]].max_g);
          bT = std::max(bT, G[v].T);
        }
      }
      TRY(size_arena(b->lanes[k], mg, T, m64, !b->shared_projection, bg, bT));
      bool presort = false;
      for (uint32_t v = 0; v < n_views; ++v)
        presort = presort || (b->view_lane[v] == k && b->view_presort[v]);
      if (presort) TRY(size_presort(b->lanes[k], mg));
    }
    // ---- the graph's fingerprint: every pointer and value it bakes in ----
    uint64_t h = FNV_OFFSET;
    h = fnv_u64(h, n_views);
    h = fnv_u64(h, L);
    h = fnv_bytes(h, view_scene, 4 * (size_t)n_views);
    h = fnv_bytes(h, b->view_lane.data(), 4 * (size_t)n_views);
    h = fnv_bytes(h, cams, sizeof(cgs_camera_uniforms) * n_views);
    h = fnv_bytes(h, comps, sizeof(cgs_compositor_uniforms) * n_views);
    const float pv[4] = {prm->strain_split_threshold, prm->mip_filter_3d_var, prm->min_opacity,
                         prm->min_radius_px};
    h = fnv_bytes(h, pv, sizeof(pv));
    for (uint32_t i = 0; i < n_scenes; ++i) {
      const cgs_context* c = scenes[i];
      h = fnv_u64(h, (uint64_t)(uintptr_t)c);
      h = fnv_bytes(h, &SF[i].S, sizeof(SF[i].S));
      h = fnv_u64(h, (uint64_t)(uintptr_t)SF[i].split_count);
      h = fnv_u64(h, ((uint64_t)SF[i].N << 32) | SF[i].max_g);
      h = fnv_u64(h, ((uint64_t)c->num_tets << 8) | (c->has_cages ? 1u : 0u) |
                         (c->f16 ? 2u : 0u) | ((uint64_t)c->f16_degree << 2));
      h = fnv_buf(h, c->bind_tet);
      h = fnv_buf(h, c->rot);
      h = fnv_buf(h, c->split_meta);
    }
    for (const FrameArena& a : b->lanes) {
      const DevBuf* bufs[] = {&a.geom,      &a.color,     &a.extents, &a.cnt_dbg, &a.e_keys,
                              &a.e_ids,     &a.e_alt_k,   &a.e_alt_v, &a.e_depths, &a.e_perm,
                              &a.ranges,    &a.comp_order, &a.hitm,   &a.ctrl3,   &a.bk_hist,
                              &a.bk_base,   &a.bk_bucket, &a.bk_alt,  &a.bk_tab,  &a.bk_ctl,
                              &a.sp_keys,   &a.sp_vals,   &a.sp_alt_k, &a.sp_alt_v, &a.sp_ctrl};
      for (const DevBuf* d : bufs) h = fnv_buf(h, *d);
      h = fnv_u64(h, a.entry_cap);
    }
    h = fnv_buf(h, b->fb);
    h = fnv_buf(h, b->counters);
    h = fnv_bytes(h, b->plan_known.data(), n_views);
    h = fnv_bytes(h, b->plan_chunks.data(), 4 * (size_t)n_views);
    h = fnv_bytes(h, b->plan_rounds.data(), 4 * (size_t)n_views);
    h = fnv_bytes(h, b->view_presort.data(), n_views);
    h = fnv_u64(h, b->shared_projection ? 1u : 0u);
    h = fnv_buf(h, b->proj);
    h = fnv_buf(h, b->vproj);
    if (h == 0) h = 1;
    if (!b->exec || b->key != h) {
      if (!desc.empty())
        HIPCHK(hipMemcpy(b->vproj.p, desc.data(), desc.size() * sizeof(CgsViewProjection),
                         hipMemcpyHostToDevice));
      TRY(batch_capture(b, scenes, SF, view_scene, cams, comps, n_views, prm));
      b->key = h;
      St.graph_built = 1;
    }
    // ---- replay ----
    HIPCHK(hipEventRecord(b->t0, b->stream));
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    HIPCHK(hipEventRecord(b->t1, b->stream));
    HIPCHK(hipMemcpyAsync(b->h_counters, b->counters.p, 32 * (size_t)n_views,
                          hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    St.gpu_ms += elapsed(b->t0, b->t1);
    // ---- arena overflow: grow the lane and replay ----
    std::vector<uint32_t> need(L, 0);
    bool overflow = false;
    uint64_t entries = 0;
    bool replan = false;
    for (uint32_t v = 0; v < n_views; ++v) {
      const uint32_t total = b->h_counters[8 * v + 2];
      entries += total;
      FrameArena& a = b->lanes[b->view_lane[v]];
      need[b->view_lane[v]] = std::max(need[b->view_lane[v]], total);
      if (total > a.entry_cap) overflow = true;
      if (b->view_bucket[v]) {
        // the segmented sort's needs: a view that needed more chunk blocks or
        // merge rounds than its graph launched is planned anew and replayed
        const uint32_t* st = b->h_counters + 8 * v + 4;
        if (st[3] > b->launched_chunks[v] || st[1] > b->launched_rounds[v]) replan = true;
        b->plan_known[v] = 1;
        b->plan_chunks[v] = st[3];
        b->plan_rounds[v] = st[1];
      }
    }
    St.total_tile_entries = entries;
    b->views_hash = vh;
    b->last_entries.resize(n_views);
    for (uint32_t v = 0; v < n_views; ++v) b->last_entries[v] = b->h_counters[8 * v + 2];
    if (!overflow && !replan) break;
    if (attempt >= 4) return fail(CGS_ERR_EXECUTION, "tile-entry arena still overflowing");
    for (uint32_t k = 0; k < L; ++k) {
      FrameArena& a = b->lanes[k];
      if (need[k] <= a.entry_cap) continue;
      const uint64_t want = (uint64_t)need[k] + need[k] / 2;
      if (want >= (1ull << 30)) return fail(CGS_ERR_EXECUTION, "tile entries exceed 2^30");
      a.entry_cap = (uint32_t)want;
    }
    St.reruns++;
  }
  if (stats) *stats = St;
  return CGS_OK;
}

int cgs_batch_set_tuning(cgs_batch* b, const cgs_tuning* t) {
  if (!b || !t) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  TRY(check_tuning(t));
  for (FrameArena& a : b->lanes) a.tune = *t;
  batch_drop_graph(b);  // the launches change
  return CGS_OK;
}

int cgs_batch_set_shared_projection(cgs_batch* b, uint32_t enable) {
  if (!b) return fail(CGS_ERR_INVALID_ARG, "null batch");
  b->shared_projection = enable != 0;
  batch_drop_graph(b);
  return CGS_OK;
}

int cgs_batch_framebuffer(cgs_batch* b, uint32_t view, void** ptr, uint32_t* padded_row_u32) {
  if (!b || !ptr || !padded_row_u32) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  if (view >= b->fb_off.size()) return fail(CGS_ERR_INVALID_ARG, "view index out of range");
  *ptr = P<uint32_t>(b->fb) + b->fb_off[view];
  *padded_row_u32 = b->fb_row[view];
  return CGS_OK;
}

int cgs_batch_readback_rgba8(cgs_batch* b, uint32_t view, uint8_t* dst, uint32_t width,
                             uint32_t height) {
  if (!b || !dst) return fail(CGS_ERR_INVALID_ARG, "bad arguments");
  if (view >= b->fb_off.size()) return fail(CGS_ERR_INVALID_ARG, "view index out of range");
  if (width != b->fb_w[view] || height != b->fb_h[view])
    return fail(CGS_ERR_SHORT_READBACK, "size differs from the view's framebuffer");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)width * 4, P<uint32_t>(b->fb) + b->fb_off[view],
                          (size_t)b->fb_row[view] * 4, (size_t)width * 4, height,
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return CGS_OK;
}

int cgs_batch_view_counts(cgs_batch* b, uint32_t view, uint32_t* tile_entries,
                          uint32_t* survivors) {
  if (!b) return fail(CGS_ERR_INVALID_ARG, "null batch");
  if (view >= b->fb_off.size() || 8 * (size_t)view >= b->h_counter_words)
    return fail(CGS_ERR_INVALID_ARG, "view index out of range");
  if (tile_entries) *tile_entries = b->h_counters[8 * view + 2];
  if (survivors) *survivors = b->h_counters[8 * view + 3];
  return CGS_OK;
}

}  // extern "C"
