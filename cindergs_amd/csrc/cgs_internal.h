// cgs_internal.h — device-side building blocks shared by the gfx950 kernels
// and the launcher declarations used by the host runtime.
//
// Device data layout (all buffers owned by the runtime, see runtime.cpp):
//   scene (rest pose)   wave-blocked SoA: splats in blocks of 64 (one wave),
//                       channel-major inside a block, element (i, k) of a
//                       K-channel array at ((i/64)*K + k)*64 + i%64.
//                       geo K=10 {x, y, z, cov xx,xy,xz,yy,yz,zz, opacity},
//                       sh K=48 (f32, or f16 for the storage variant),
//                       cov_h K=6 f16 (storage variant).  A wave's loads are
//                       256 contiguous bytes each and stay inside one 2.5 KB
//                       (geo) + 12 KB (sh) region: few pages, few TLB entries.
//   deformed scene      geo layout with capacity 2N when splitting is on
//   projected splat     geom[8] f32 = {mu_x, mu_y, a, b, c, alpha, tz, q_cut}
//                       color[4] f32 = {r, g, b, 0}; extents u32 (ext_x | ext_y<<16)
//   stage 3             keys/ids u32[E] (+ ping-pong copies), ranges u32[2T]
// Wave size is 64 (CDNA); every block size below is a multiple of 64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cgs_detmath.h"

#define CGS_WAVE 64

// Flags of the decoupled look-back status words.
#define CGS_LB_AGG 1u
#define CGS_LB_INC 2u

#define CGS_GEO_K 10
#define CGS_CH_X 0
#define CGS_CH_COV 3
#define CGS_CH_OPA 9
#define CGS_SH_K 48

// index of element (i, k) of a K-channel wave-blocked array
__host__ __device__ __forceinline__ size_t cgs_blk(uint32_t i, uint32_t k, uint32_t K) {
  return ((size_t)(i >> 6) * K + k) * 64u + (i & 63u);
}
// words of a K-channel wave-blocked array holding n elements
__host__ __device__ __forceinline__ size_t cgs_blk_words(uint32_t n, uint32_t K) {
  return (size_t)((n + 63u) / 64u) * K * 64u;
}

namespace cgs {

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// number of set bits of `mask` below this lane
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// inclusive scan across the 64 lanes of a wave
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += t;
  }
  return v;
}

// Exclusive scan of one value per thread over a block of NT threads.
// `s_tmp` must hold NT/64 words.  Returns the exclusive prefix; `total` gets the sum.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp,
                                                         uint32_t& total) {
  constexpr int NW = NT / 64;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / 64;
  uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) s_tmp[wave] = inc;
  __syncthreads();
  uint32_t wave_prefix = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    uint32_t s = s_tmp[w];
    if (w < (int)wave) wave_prefix += s;
    tot += s;
  }
  total = tot;
  __syncthreads();
  return wave_prefix + inc - v;
}

__device__ __forceinline__ uint32_t ld_acquire_u32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release_u32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_acquire_u64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back, windowed.  Partition `part` publishes its aggregate,
// then reads the statuses of up to WIN predecessors at once and consumes
// them nearest first: an aggregate is added and the walk goes on, an
// inclusive prefix is added and ends it, a status not yet published stops the
// window and is polled again.  On MI355X each status read is a cross-XCD
// coherent load (~125 ns), so reading a window per round trip lets the prefix
// travel WIN partitions per memory latency instead of one.  Predecessors
// hold smaller partition ids taken from an atomic ticket (or launched earlier),
// so they are resident or finished and every wait ends.  Called by one thread
// per status column.
// With PUBLISHED the caller has already stored this partition's aggregate
// (lookback_publish), as early as it was known, so that the successors'
// walks find it sooner.
template <class U, int VBITS>
__device__ __forceinline__ void lookback_publish(U* status, uint32_t part, uint32_t stride,
                                                 U aggregate) {
  const U VMASK = ((U)1 << VBITS) - (U)1;
  if (part != 0)
    __hip_atomic_store(&status[(size_t)part * stride],
                       ((U)CGS_LB_AGG << VBITS) | (aggregate & VMASK), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

template <class U, int VBITS, int WIN, bool PUBLISHED = false>
__device__ __forceinline__ U lookback_windowed(U* status, uint32_t part, uint32_t stride,
                                               U aggregate) {
  const U VMASK = ((U)1 << VBITS) - (U)1;
  if (part == 0) {
    __hip_atomic_store(&status[0], ((U)CGS_LB_INC << VBITS) | (aggregate & VMASK),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (!PUBLISHED)
    lookback_publish<U, VBITS>(status, part, stride, aggregate);
  U excl = 0;
  int64_t q = (int64_t)part - 1;  // nearest predecessor not yet accounted for
  while (true) {
    U s[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
      const int64_t qq = q - k;
      s[k] = qq >= 0 ? __hip_atomic_load(&status[(size_t)qq * stride], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT)
                     : ((U)CGS_LB_INC << VBITS);  // never consumed: partition 0 is inclusive
    }
    int used = 0;
    bool done = false;
#pragma unroll
    for (int k = 0; k < WIN; ++k) {
      if (done || used != k) continue;
      const uint32_t flag = (uint32_t)(s[k] >> VBITS);
      if (flag == 0) continue;  // not published yet: poll again from here
      excl += s[k] & VMASK;
      used = k + 1;
      if (flag == CGS_LB_INC) done = true;
    }
    if (done) break;
    q -= used;
    if (used < WIN) __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_store(&status[(size_t)part * stride],
                     ((U)CGS_LB_INC << VBITS) | ((excl + aggregate) & VMASK), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// 64-bit status words {flag:2 | value:62}; window 4 (measured: the split
// compaction of k_deform -6 % against 16)
__device__ __forceinline__ uint64_t lookback_u64(uint64_t* status, uint32_t part,
                                                 uint64_t aggregate) {
  return lookback_windowed<uint64_t, 62, 4>(status, part, 1u, aggregate);
}

// 32-bit status words {flag:2 | value:30} (radix digit counts).  A window
// of 4 measured fastest for the sort (1, 2, 4, 8, 16, 32 and 64 tried): each
// of its 256 digit threads issues the window's loads together, and wider
// windows cost more in loads than they save in round trips.
__device__ __forceinline__ uint32_t lookback_u32(uint32_t* status, uint32_t part,
                                                 uint32_t stride, uint32_t aggregate) {
  return lookback_windowed<uint32_t, 30, 4>(status, part, stride, aggregate);
}
// the same walk for a partition that published its aggregate already
__device__ __forceinline__ uint32_t lookback_walk_u32(uint32_t* status, uint32_t part,
                                                      uint32_t stride, uint32_t aggregate) {
  return lookback_windowed<uint32_t, 30, 4, true>(status, part, stride, aggregate);
}

}  // namespace cgs

// ---------------------------------------------------------------------------
// Launchers (defined in the k_*.hip files, called by runtime.cpp)
// ---------------------------------------------------------------------------

struct CgsSceneDev {
  const float* geo;       // blocked K=10: x,y,z, cov[6], opacity
  const uint16_t* cov_h;  // blocked K=6 f16 covariances (storage variant) or null
  const float* sh;        // blocked K=48 f32, or null when sh_h is used
  const uint16_t* sh_h;   // blocked K=48 f16 (storage variant)
};

struct CgsProjectArgs {
  CgsSceneDev scene;
  const uint32_t* bind_tet;     // [n_rest] or null
  const float* tet_rot;         // [9*num_tets] col-major or null
  const uint32_t* split_meta;   // [2*n_rest] (parent, tet) or null
  const uint32_t* split_count;  // device scalar or null (== 0)
  float camera[24];
  float cam_pos[3];
  float mip;
  float min_opacity;
  float min_radius;
  uint32_t sh_degree;
  uint32_t has_deformation;
  uint32_t num_tets;
  uint32_t num_rest;
  uint32_t max_gaussians;
  uint32_t zero_culled;  // write zero records for culled / inactive splats
  float* geom;       // [8*max]
  float* color;      // [4*max]
  uint32_t* extents; // [max]
};

// One view of a multi-view projection: camera (CameraUniforms words),
// camera position, and the view's output records
struct CgsViewProjection {
  float camera[24];
  float cam_pos[3];
  float pad;
  float* geom;        // [8*max]
  float* color;       // [4*max]
  uint32_t* extents;  // [max]
};

struct CgsBinArgs {
  const float* geom;
  const uint32_t* extents;
  const uint32_t* split_count;  // device scalar or null
  uint32_t num_rest;
  uint32_t max_gaussians;
  uint32_t width, height, tiles_x, tiles_y;
  uint32_t tile_w, tile_h;   // tile size in pixels (16 x 16: the reference's)
  uint32_t mode64;           // 0: packed 32-bit keys, 1: (tile, depth bits) pairs
  uint32_t base_depth_bits;  // 32-bit mode
  uint32_t depth_bits;       // 32-bit mode
  uint32_t cap;              // capacity of the entry arrays
  uint32_t* keys;            // 32-bit mode: packed keys; 64-bit mode: tile ids
  uint32_t* depths;          // 64-bit mode only (null: not written)
  uint32_t* ids;
  uint32_t* counts;          // [max] per-splat tile counts
  uint32_t* offsets_out;     // optional [max] (exclusive scan)
  uint32_t* total;           // device scalars: [0] total entries (may exceed cap),
                             // [1] survivors (splats with nonzero extents)
  uint32_t* block_sums;      // [num_parts] entries per block
  uint32_t* block_surv;      // [num_parts] survivors per block
  uint32_t* hist_partial;    // optional [num_parts][hist_words]: digit histograms of the
                             // emitted keys (32-bit mode: 4 passes of the key; 64-bit
                             // mode: 4 passes of the depth bits, then the tile passes)
  uint32_t hist_words;       // 1024 (32-bit) or 2048 (64-bit)
  uint32_t sort_bits;        // 32-bit mode: bits of the packed key
  uint32_t tile_bits;        // 64-bit mode
  uint32_t digit_bits;       // 32-bit mode: digit width of the sort (8 or 4); 64-bit: 8
  // hit results k_count leaves for k_emit (both null: k_emit re-tests):
  uint8_t* small_masks;      // [max] per splat with a box of <= 8 tiles: bit k = tile k hit
  uint64_t* chunk_masks;     // [num_parts * 8 * 32] per wave: the hit masks of its first
                             // 32 wave-walk chunks
  // words k_count clears for the kernels after it (sort tickets and look-back
  // status, tile ranges), so that no fill launch precedes it: three spans
  uint32_t* zero_at[3];
  uint32_t zero_len[3];
  // bucket mode (tile_hist null: the LSD mode, one sub-block per block)
  uint32_t* tile_hist;        // [grid][tiles_pad / 2] u16-pair tile counts per block (k_count)
  const uint32_t* tile_base;  // [grid][tiles_pad] bucket base of every (block, tile) (k_emit)
  uint64_t* bucket;           // [cap] (key << 32 | splat id), grouped by tile (k_emit)
  uint32_t tiles_pad;         // multiple of 128
  uint32_t subs;              // sub-blocks of 512 splats per block
  uint32_t parts;             // sub-blocks in all
  // depth-ordered emission (presort): position j of the binning takes splat
  // order[j] (a stable sort of the splats by their key's depth bits), so the
  // entries come out ordered by (depth, splat id) and only the tile bits are
  // left to sort; counts, offsets and hit masks are per position, the
  // entries' ids are splat ids.  The histogram covers the tile digits only
  // (8-bit digits of the tile id at hist_partial[p * 256]); 64-bit mode
  // writes no depths.  Null: positions are splat indices.
  const uint32_t* order;
};

// presort keys: keys[s] = the depth bits of splat s's sort key (32-bit mode:
// the packed key's depth digit; 64-bit mode: the raw f32 bits of max(tz, 0)),
// 0 for splats that emit nothing (s >= active or extents 0), s < max_g; also
// clears zero[0 .. zero_len) (the presort's histogram, look-back status and
// tickets) and writes per-block digit histograms of the sort's passes (8-bit
// digits over key_bits bits) as rows of 1024 words of `partial`.  Returns the
// rows written (<= max_rows).
uint32_t cgs_launch_splat_depth_keys(const CgsBinArgs& a, uint32_t* keys, uint32_t* zero,
                                     uint32_t zero_len, uint32_t key_bits, uint32_t* partial,
                                     uint32_t max_rows, hipStream_t s);

// ---- Stage 3 in bucket mode (k_tilesort.hip) ----
#define CGS_SEG_WINDOW 2048u  // entries per sort window; larger tiles are sorted in chunks
#define CGS_SEG_CHUNK 4096u   // entries per chunk of a large tile
struct CgsBucketArgs {
  const uint32_t* hist;   // [rows][tiles_pad / 2] u16-pair tile counts of the binning blocks
  uint32_t rows, tiles, tiles_pad, cap;
  uint32_t big;           // tiles with more entries are sorted in chunks
  uint32_t window;        // CGS_SEG_WINDOW
  uint32_t* base;         // [rows][tiles_pad] bucket base of every (block, tile)
  uint32_t* tile_start;   // [tiles + 1] monotone tile starts
  uint32_t* chunk_start;  // [tiles + 1] first chunk of every tile (large tiles)
  uint32_t* ranges;       // [2 tiles] (start, end), (0, 0) when empty
  uint32_t* win_first;    // [win_slots] first tile starting at or after k x window
  uint32_t win_slots;
  uint32_t* chunk_tile;   // [chunk_slots] tile of every chunk
  uint32_t chunk_slots;
  uint64_t* status;       // [tiles_pad / 128] look-back, zeroed
  uint32_t* ticket;       // zeroed
  uint32_t* stats;        // [4] zeroed: max tile entries, merge rounds, entries, chunks
};
struct CgsSegSortArgs {
  uint64_t* bucket;       // [cap] entries grouped by tile (from k_emit)
  uint64_t* alt;          // [cap] chunk runs of the large tiles
  uint32_t* keys;         // [cap] sorted keys (output)
  uint32_t* ids;          // [cap] sorted splat ids (output)
  const uint32_t* tile_start;
  const uint32_t* chunk_start;
  const uint32_t* win_first;
  const uint32_t* chunk_tile;
  const uint32_t* stats;
  uint32_t tiles, cap, window, windows, big, chunk_slots;
};
uint32_t cgs_seg_sort_windows(uint32_t cap);
struct CgsProf;
void cgs_launch_bucket_scan(const CgsBucketArgs& a, hipStream_t s);
void cgs_launch_seg_sort(const CgsSegSortArgs& a, uint32_t chunk_blocks, uint32_t rounds,
                         hipStream_t s, CgsProf* prof);

// Optional per-launch timing: mark(name) records an event before a launch;
// the time of launch i is the interval between marks i and i+1.
struct CgsProf {
  static constexpr int MAX = 64;
  hipEvent_t ev[MAX];
  const char* names[MAX];
  int n = 0;
  bool on = false;
  void mark(const char* name, hipStream_t s) {
    if (!on || n >= MAX) return;
    names[n] = name;
    (void)hipEventRecord(ev[n], s);
    ++n;
  }
};

extern "C++" {
void cgs_launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
// dst[0, n) = src[0, n) with one small block (a frame's control words)
void cgs_launch_copy_words(const uint32_t* src, uint32_t* dst, uint32_t n, hipStream_t s);
// per-splat records in[i*k_in + c] -> blocked out[(i, chan0 + c) of k_out channels]
void cgs_launch_to_blocked_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t k_in,
                               uint32_t k_out, uint32_t chan0, hipStream_t s);
void cgs_launch_to_blocked_u16(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t k_in,
                               uint32_t k_out, uint32_t chan0, hipStream_t s);
// trimmed channel-major f16 SH records (stride halves) -> blocked K=48 halves
void cgs_launch_sh_f16_blocked(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t stride,
                               uint32_t rest, hipStream_t s);
void cgs_launch_project(const CgsProjectArgs& a, hipStream_t s);
// Stage 2 of n_views <= 64 views of the scene in `a` (its camera, cam_pos and
// outputs unused) into each view's records; `views` in device memory
void cgs_launch_project_views(const CgsProjectArgs& a, const CgsViewProjection* views,
                              uint32_t n_views, hipStream_t s);
uint32_t cgs_bin_num_parts(uint32_t max_gaussians);
// words of the hit-mask buffer k_count fills for k_emit (small masks, then
// chunk masks 8-byte aligned); cgs_bin_masks splits it
size_t cgs_bin_mask_words(uint32_t max_gaussians);
void cgs_bin_masks(uint32_t* buf, uint32_t max_gaussians, uint8_t** small_masks,
                   uint64_t** chunk_masks);
// tiles the bucket mode's LDS histograms hold (larger frames use the LSD mode)
#define CGS_BUCKET_MAX_TILES 32768u
uint32_t cgs_bin_grid(const CgsBinArgs& a);
void cgs_launch_bin(const CgsBinArgs& a, hipStream_t s, CgsProf* prof = nullptr,
                    const CgsBucketArgs* between = nullptr);
void cgs_launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total,
                     uint64_t* status, uint32_t* ticket, hipStream_t s);
uint32_t cgs_scan_num_parts(uint32_t n);

// radix sort of (key, value) u32 pairs on `bits` low key bits, n read from
// the device scalar `d_n` (or n_static when d_n is null).  `cap` bounds n.
struct CgsRadixBuffers {
  uint32_t* keys;
  uint32_t* vals;
  uint32_t* keys_alt;
  uint32_t* vals_alt;
  uint32_t* hist;          // [4][256] digit histograms
  uint32_t* hist_partial;  // [cgs_radix_hist_blocks()][4*256] (when computed here)
  uint32_t* status;        // [4][max_parts][256] zeroed
  uint32_t* tickets;       // [4] zeroed
  uint32_t cap;
};
uint32_t cgs_radix_num_parts(uint32_t cap);
uint32_t cgs_radix_hist_blocks();
size_t cgs_radix_control_words(uint32_t cap);
// hist[b] = sum_r partial[r*row_words + b] for b < bins
// hist[0..bins) = column sums of partial[rows][row_words]; zeroed: hist is
// already zero (else it is cleared first)
void cgs_launch_hist_reduce(const uint32_t* partial, uint32_t rows, uint32_t row_words,
                            uint32_t bins, uint32_t* hist, hipStream_t s, bool zeroed);
// Returns 1 if the sorted result is in the *_alt buffers, else 0.  With
// hist_given the caller has already filled b.hist for the passes (pass p's
// 2^digit_bits bins at b.hist + p * 2^digit_bits).  part_keys: 16384 (one
// 1024-thread block per CU), 10240 (640 threads: ~256 partitions for 2.5M
// keys), 8192 (two 512-thread blocks per CU) or 4096 (256 threads, four
// blocks per CU);
// digit_bits: 8 (4 passes over 32 bits) or 4 (8 passes).  Every choice gives
// the same stable order.  begin_bit: the key bits [begin_bit, bits) are
// sorted (the entries arrive ordered by the bits below it); the histogram of
// pass p is that of the digit at begin_bit + p * digit_bits.
int cgs_launch_radix_sort(const CgsRadixBuffers& b, const uint32_t* d_n, uint32_t n_static,
                          uint32_t bits, bool iota_vals, bool hist_given, hipStream_t s,
                          CgsProf* prof = nullptr, uint32_t part_keys = 16384,
                          uint32_t digit_bits = 8, uint32_t begin_bit = 0);
// a_out[i] = a_in[perm[i]], b_out likewise; perm, a_out, b_out 16-byte
// aligned (false, nothing launched, otherwise)
bool cgs_launch_gather2(const uint32_t* perm, const uint32_t* a_in, const uint32_t* b_in,
                        uint32_t* a_out, uint32_t* b_out, const uint32_t* d_n, uint32_t cap,
                        hipStream_t s);
void cgs_launch_tile_ranges(const uint32_t* sorted, const uint32_t* d_n, uint32_t cap,
                            uint32_t shift, uint32_t total_tiles, uint32_t* ranges,
                            hipStream_t s);

struct CgsCompositeArgs {
  const float* geom;
  const float* color;
  const uint32_t* sorted_ids;
  const uint32_t* ranges;
  uint32_t width, height, tiles_x, tiles_y, padded_row;
  uint32_t tile_w, tile_h;  // 8 x 8, 16 x 16, 16 x 32 or 32 x 16 pixels
  float bg[4];
  uint32_t* out;
  uint32_t* order;  // [tiles] dispatch order scratch (null: tile order)
};
// batch: splats staged per compositor step (64, 128 or 256; the same pixels
// for every choice; 8 x 8 tiles take 64)
void cgs_launch_composite(const CgsCompositeArgs& a, hipStream_t s, CgsProf* prof = nullptr,
                          uint32_t batch = 256);

struct CgsBindArgs {
  const float* tet_rec;  // [T][12]: X0 xyz, inv_dm column-major
  uint32_t num_tets;
  const float* pos;      // [n][3]
  uint32_t n;
  uint32_t has_bounds;
  float bmin[3], bmax[3];
  // uniform grid: cell (x, y, z) = ((z * dims[1] + y) * dims[0] + x); null = none
  const uint32_t* cell_start;  // [cells + 1]
  const uint32_t* cell_tets;   // tet ids per cell, ascending
  float lo[3], inv_cell[3];
  uint32_t dims[3];
  uint32_t* out_tet;  // [n], 0xFFFFFFFF = unbound
  float* out_w;       // [n][4]
};
void cgs_launch_bind_tets(const CgsBindArgs& a, hipStream_t s);

struct CgsMvcArgs {
  const float* verts;       // [nv][3] rest cage vertices
  uint32_t nv;
  const uint32_t* tris;     // [nt][3] outward triangles
  uint32_t nt;
  const uint32_t* vt_start; // [nv + 1] incidence lists: entries tri*3 + corner,
  const uint32_t* vt_list;  //   ascending triangle order per vertex
  const float* pts;         // [n][3]
  uint32_t n;
  float h;                  // central-difference step (jacobian_step_size)
  float* out_w;             // [n][nv]
  float* out_g;             // [n][nv][3] or null
  uint32_t* mask;           // [n] inside flags or null
  const uint32_t* list;     // [n] indices of the points to bind (others untouched),
                            //   or null: points 0..n-1
};
size_t cgs_mvc_lds_bytes(uint32_t nv, uint32_t nt);
void cgs_launch_mvc_bind(const CgsMvcArgs& a, hipStream_t s);
void cgs_launch_mvc_winding(const CgsMvcArgs& a, float* out, hipStream_t s);
// inside = |winding| >= 0.5 (winding given) or the flags already in
// `inside` (winding null); points inside whose residual exceeds tol are
// appended to list (count in *count, zeroed by the caller; the order is
// arbitrary); rows of points outside zeroed (w [n x nv], g [n x nv x 3] or null)
void cgs_launch_mvc_select(const float* winding, const float* residual, float tol, uint32_t n,
                           uint32_t nv, uint32_t* inside, float* w, float* g, uint32_t* list,
                           uint32_t* count, hipStream_t s);
// Near/far split of the MVC sums (k_mvc.hip): triangle t's terms at point x
// carry the weight phi_t(x) = smoothstep((r1 - |x - c_t|) / (r1 - r0)) (c_t
// the triangle's centroid): the near sums u_near = sum_t phi_t mu_t are
// evaluated exactly, the far share q = (u - u_near) / sum(u) is learned, or
// the far sums u_far = u - u_near themselves (sums = 1).
// Training targets: q * q_scale (or u_far * q_scale) per point, [n x nv],
// degenerate points zero.
void cgs_launch_mvc_far_share(const CgsMvcArgs& a, const float* cent, float r0, float r1,
                              float q_scale, int sums, float* out, hipStream_t s);
// winding numbers, one point per lane (the same sum as winding_number,
// mvc.rs:326-347, accumulated in triangle order)
void cgs_launch_mvc_winding_lane(const float* verts, const uint32_t* tris, uint32_t nt,
                                 const float* pts, uint32_t n, float* out, hipStream_t s);
struct CgsNearArgs {
  const float* verts;          // [nv][3]
  uint32_t nv;
  const uint32_t* tris;        // [nt][3]
  uint32_t nt;
  const float* cent;           // [nt][3] triangle centroids
  const uint32_t* cell_start;  // [cells + 1] candidate lists: every triangle whose
  const uint32_t* cell_list;   //   centroid is within r1 + h + cell half-diagonal
  float lo[3], inv_cell;       //   of the cell centre, ascending
  uint32_t dim[3];
  float r0, r1, h, tau;        // tau: scale below which the near term fades out
  const float* pts;            // [n][3]
  uint32_t n;
  const float* winding;        // [n] or null: the inside test is the crossing
                               //   parity below
  const uint32_t* cell_inside; // [cells] 1 = cell centre inside the cage (parity
                               //   reference), 2 = undecided (points there are
                               //   flagged for the exact path)
  const float* Q;              // [n x ldq] far share * q_scale (far_sums: u_far * q_scale)
  const float* DQ;             // [3n x ldq] its derivatives (row 3p + axis) or null
  const uint16_t* Qh;          // the same in bf16 (used when set)
  const uint16_t* DQh;
  uint32_t ldq;
  float q_scale;
  uint32_t far_sums;           // 1: Q holds the far sums, w = (u_near + u_far) / sum
  float sums_scale;            // the far sums' target scale for this cage (near setup)
  float minv[16];              // (A A^T)^-1, A = [1; v] (4 x nv), row-major
  float* out_w;                // [n][nv]
  float* out_g;                // [n][nv][3] or null (then DQ unused)
  float* residual;             // [n]: |sum w v - x| before the projection; +inf
                               //   where the near sums hit a vertex or a face
  uint32_t* inside;            // [n] or null
};
#define CGS_NEARFAR_MAX_VERTICES 2048u
void cgs_launch_mvc_nearfar(const CgsNearArgs& a, hipStream_t s);
// deform_scene_cpu (mvc.rs:564-749) after the per-point position and Jacobian
struct CgsMvcSceneArgs {
  const uint32_t* mask;      // [n] bound flags, or null (all bound)
  const float* def_pos;      // [3n] x'
  const float* def_F;        // [9n] F (column-major)
  const float* rest_pos;     // [3n]
  const float* rest_cov;     // [6n]
  const float* rest_opa;     // [n]
  const float* rest_sh;      // [48n]
  uint32_t n, sh_degree;
  float tau, mip;
  float* out_pos;            // [3 * 2n]
  float* out_cov;            // [6 * 2n]
  float* out_opa;            // [2n]
  float* out_sh;             // [48 * 2n]
  uint32_t* flags;           // [n] split flags
  float* child_b;            // [10n] second child of split splats (pos, cov, opacity)
};
void cgs_launch_mvc_scene(const CgsMvcSceneArgs& a, hipStream_t s);
void cgs_launch_mvc_scene_children(const CgsMvcSceneArgs& a, const uint32_t* ranks,
                                   hipStream_t s);
void cgs_launch_mvc_deform(const float* w, const float* g, const uint32_t* mask,
                           const float* rest_pos, const float* dverts, uint32_t nv, uint32_t n,
                           float* out_pos, float* out_F, hipStream_t s);

// ---- surrogate network (k_mlp.hip, mlp.cpp) ----
#define CGS_MLP_LAYERS 5
struct CgsGemmArgs {       // C = A B^T, A [M x K], B [N x K] bf16 (K % 32 == 0, ld % 8 == 0)
  const uint16_t* A;
  uint32_t lda;
  const uint16_t* B;
  uint32_t ldb;
  uint32_t M, N, K;
  const float* bias;       // bias[col * bias_stride] or null
  uint32_t bias_stride;
  uint32_t relu;
  const uint16_t* mask;    // C *= (mask[(row / mask_div) * ldm + col] > 0) or null
  uint32_t ldm, mask_div;
  float* Cf;               // f32 [M x ldcf] or null
  uint32_t ldcf;
  uint16_t* Cb;            // bf16 [M x ldcb] or null
  uint32_t ldcb;
  uint16_t* Ct;            // bf16 transposed [N x ldct] or null (M % 4 == 0)
  uint32_t ldct;
  // split-K (ksplit > 1): block z of the grid sums the K range z of ksplit
  // equal 32-aligned ranges into Cf + z * split_stride (f32 partials only: no
  // bias, ReLU, mask, bf16 or transposed output); the partials are summed by
  // their reader in a fixed order
  uint32_t ksplit;
  uint64_t split_stride;
};
struct CgsAdamArgs {
  float* w;                // flat master parameters (layer l: rows x (cols + 32), bias in col `cols`)
  float* m;
  float* v;
  const float* grad;
  uint64_t total;
  uint64_t off[CGS_MLP_LAYERS];
  uint32_t rows[CGS_MLP_LAYERS], cols[CGS_MLP_LAYERS];
  uint16_t* wb[CGS_MLP_LAYERS];  // bf16 [rows x cols]
  uint16_t* wt[CGS_MLP_LAYERS];  // bf16 [cols x rows]
  float b1, b2, lr_t, eps_t;
  uint32_t splits;         // the gradient is the sum of `splits` partials,
  uint64_t split_stride;   // grad[e + s * split_stride], s = 0 .. splits-1
};
void cgs_launch_gemm_nt(const CgsGemmArgs& g, hipStream_t s);
// grad[e] = sum over s < splits of grad[e + s * stride] (in place), e < total
void cgs_launch_sum_splits(float* grad, uint64_t total, uint32_t splits, uint64_t stride,
                           hipStream_t s);
void cgs_launch_fourier(const float* X, const uint32_t* idx, uint32_t P, const float* Bm,
                        uint16_t* A0, uint16_t* A0t, uint32_t ldt, uint16_t* T0, hipStream_t s);
void cgs_launch_mlp_loss(const float* Z, uint32_t ldz, const float* targets, const uint32_t* idx,
                         uint32_t P, uint32_t V, uint16_t* dZ, uint16_t* dZt, uint32_t ldt,
                         float* row_loss, bool affine, hipStream_t s);
// Whole-network inference in one kernel (k_mlp.hip): Fourier features, four
// hidden layers and the output layer with the activations resident in LDS.
// Z [P x ldz] = outputs (no normalisation), TZ [3P x ldz] = their derivatives
// along x, y, z (row 3p + axis) or null.
struct CgsMlpInferArgs {
  const float* X;                       // [P x 3]
  uint32_t P;
  const float* fourier;                 // [32 x 3]
  const uint16_t* wb[CGS_MLP_LAYERS];   // bf16 [rows x cols]
  const float* w;                       // master weights (bias column)
  uint64_t off[CGS_MLP_LAYERS];
  uint32_t n_out_p;                     // output rows, a multiple of 32
  float* Z;
  float* TZ;
  uint16_t* Zh;                         // bf16 outputs instead (when set)
  uint16_t* TZh;
  uint32_t ldz;
};
void cgs_launch_mlp_infer(const CgsMlpInferArgs& a, hipStream_t s);
void cgs_launch_sum(const float* v, uint32_t n, float div, float* out, hipStream_t s);
void cgs_launch_adam(const CgsAdamArgs& a, bool update, hipStream_t s);
void cgs_launch_sample(uint32_t* idx, uint32_t P, uint64_t n, uint64_t seed, uint32_t step,
                       hipStream_t s);
struct CgsMlpPostArgs {
  const float* Z;       // [P x ldz]
  const float* TZ;      // [3P x ldz] tangents or null
  uint32_t ldz, P, V;
  const float* X;       // [P x 3]
  const float* cage;    // [V x 3] rest cage or null (no residual, no projection)
  float minv[16];       // (A A^T)^-1, A = [1; v] (4 x V), row-major
  float* out_w;         // [P x V]
  float* out_g;         // [P x V x 3] or null
  float* residual;      // [P] or null
  int raw;              // 1: outputs and tangents copied as they are
};
void cgs_launch_mlp_post(const CgsMlpPostArgs& a, hipStream_t s);

struct cgs_mlp;
int cgs_mlp_infer_residual(struct cgs_mlp* M, const float* d_x, uint64_t n, float* d_w,
                           float* d_grad, const float* d_cage, const float* minv16,
                           float* d_residual);
int cgs_mlp_infer_nearfar(struct cgs_mlp* M, const float* d_x, uint64_t n, float* d_w,
                          float* d_grad, const CgsNearArgs& near);
bool cgs_mlp_near(struct cgs_mlp* M, float* r0, float* r1);
hipStream_t cgs_mlp_stream(struct cgs_mlp* M);
uint32_t cgs_mlp_outputs(struct cgs_mlp* M);

// ---- scene ingest (k_ingest.hip, ply.cpp) ----
// decode + activation of n raw vertex records (a PLY payload) into the SoA
struct CgsActivateArgs {
  const uint8_t* records;    // n * stride bytes, 4-byte aligned
  uint64_t n;
  uint32_t stride;           // bytes per record
  uint32_t le;               // little-endian scalars
  uint32_t per_block;        // records per block (<= 128, per_block * stride <= 65528)
  uint32_t word_path;        // every present slot is a 4-byte aligned LE float32 and stride % 4 == 0
  uint32_t slot_offset[62];
  uint8_t slot_type[62];
  float* pos;                // [3n]
  float* cov;                // [6n]
  float* opa;                // [n]
  float* sh;                 // [48n]
};
void cgs_launch_activate_raw(const CgsActivateArgs& a, hipStream_t s);
void cgs_launch_quantize(const float* cov, const float* sh, uint64_t n, uint32_t degree,
                         uint32_t* cov_f16, uint32_t* rot, uint32_t* sh_f16, hipStream_t s);
void cgs_launch_export_raw(const float* pos, const float* cov, const float* opa,
                           const float* sh, uint64_t n, float* records, hipStream_t s);
// the deformed scene of the last Stage 1 run as a SoA, SH rotated per tet
// the band 1-3 matrices of one rotation (9, 25, 49 floats) at float offsets
// 0, 12 and 40, each padded to whole float4 (16-byte loads)
#define CGS_SH_MATS 92u
#define CGS_SH_M2 12u
#define CGS_SH_M3 40u
struct CgsExportDeformedArgs {
  const float* geo;          // Stage 1 output, blocked K=10
  const float* sh;           // rest SH, blocked K=48 f32
  const uint32_t* bind_tet;  // [num_rest]
  const uint32_t* split_meta;  // [2 * splits] (parent, tet)
  const float* mats;         // [num_tets * CGS_SH_MATS]
  uint32_t num_rest, count, num_tets, degree;
  float* pos;                // [3 count]
  float* cov;                // [6 count]
  float* opa;                // [count]
  float* out_sh;             // [48 count], 16-byte aligned
};
void cgs_launch_tet_sh_matrices(const float* rot, uint32_t num_tets, float* mats,
                                hipStream_t s);
void cgs_launch_export_deformed(const CgsExportDeformedArgs& a, hipStream_t s);

// error reporting and context access shared by the host files
int cgs_fail(int code, const char* msg);

hipStream_t cgs_context_stream(struct cgs_context* c);
// the context's device made current (for ABI files outside runtime.cpp)
int cgs_context_enter(struct cgs_context* c);
// a device buffer of at least `bytes`, owned by the context (grown 1.5x, kept)
int cgs_context_staging(struct cgs_context* c, uint64_t bytes, void** p);
// per-launch HIP events on the context stream, read by cgs_kernel_times
void cgs_context_prof_begin(struct cgs_context* c);
void cgs_context_prof_mark(struct cgs_context* c, const char* name);
void cgs_context_prof_end(struct cgs_context* c, bool ok);

struct CgsTetArgs {
  const float* verts;     // [3*V] interleaved
  const uint32_t* elems;  // [4*T]
  const float* inv_dm;    // [9*T] col-major
  float* jac;             // [9*T]
  float* rot;             // [9*T]
  float* stretch;         // [T]
  float* pack;            // [CGS_TET_PACK*T] the per-tet record k_deform reads
  uint32_t num_tets;
};
// per tet: corners v0..v3 (xyz), J (9, column-major), stretch, 2 pad floats
#define CGS_TET_PACK 24u
void cgs_launch_tet_transforms(const CgsTetArgs& a, hipStream_t s);

struct CgsDeformArgs {
  const float* verts;
  const uint32_t* elems;
  const float* tet_pack;     // [CGS_TET_PACK*T] (see CgsTetArgs), 16-byte aligned
  const uint32_t* bind_tet;  // [n]
  const float* bind_w;       // [4n], 16-byte aligned
  const float* rest_geo;     // blocked K=10
  float* out_geo;            // blocked K=10, capacity 2n when splitting
  uint32_t* split_meta;      // [2n] or null
  uint32_t* split_count;     // device scalar (written)
  uint64_t* status;          // [parts] zeroed
  uint32_t* ticket;          // zeroed
  float split_threshold;
  float mip;
  uint32_t num_gaussians;
  uint32_t num_tets;
};
uint32_t cgs_deform_num_parts(uint32_t n);
// false (nothing launched) when tet_pack or bind_w is not 16-byte aligned
bool cgs_launch_deform(const CgsDeformArgs& a, hipStream_t s);
}
