// k_binning.hip — Stage 3 binning: per-splat tile counts, their exclusive
// scan and key emission in ONE single-pass kernel, plus tile-range detection.
//
// Semantics follow count_tiles_kernel / emit_tile_keys_kernel /
// emit_tile_keys_u64_kernel (src/render/kernels/binning.rs:90-328) with the
// exact quadratic ellipse-vs-tile test ellipse_intersects_tile_gpu (35-83)
// and gpu_exclusive_scan_total (scan.rs:198-333); ranges follow
// identify_tile_ranges(_u32_tile)_kernel (binning.rs:344-417).
//
// MI355X design: the reference runs count -> 3-level scan -> host readback of
// the total -> emit.  Here two kernels: k_count writes per-splat counts and
// per-block sums; k_emit derives its block prefix by reducing the sums of the
// blocks before it (a few loads per thread: no look-back chain, no ticket, no
// same-address atomics), scans locally and emits; the total stays on the
// device and emission is bounded by the arena capacity (an overflowing frame
// is re-run by the host with a larger arena).  The emitting blocks also build
// the radix digit histograms of the entries they store.  The per-splat cutoff
// q_cut = clamp(2 ln(255 alpha), 0, 8) was computed once by Stage 2 with the
// same expression the reference evaluates per tile, so the test is bit-equal.
#include "cgs_internal.h"

namespace {

constexpr int BT = 256;  // threads per block
constexpr int BS = 4;    // consecutive splats per thread
constexpr int BP = BT * BS;

struct TileBox {
  uint32_t tx0, tx1, ty0, ty1;
};

__device__ __forceinline__ TileBox tile_box(float mu_x, float mu_y, uint32_t packed,
                                            uint32_t tiles_x, uint32_t tiles_y) {
  const float ext_x = (float)(packed & 0xFFFFu);
  const float ext_y = (float)((packed >> 16) & 0xFFFFu);
  const int32_t mtx = (int32_t)tiles_x - 1, mty = (int32_t)tiles_y - 1;
  TileBox b;
  b.tx0 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_x - ext_x) / 16.0f)), 0, mtx);
  b.tx1 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_x + ext_x) / 16.0f)), 0, mtx);
  b.ty0 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_y - ext_y) / 16.0f)), 0, mty);
  b.ty1 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_y + ext_y) / 16.0f)), 0, mty);
  return b;
}

struct Splat2D {
  float mu_x, mu_y, a, b, c, alpha, tz, qcut;
};

// ellipse_intersects_tile_gpu (binning.rs:35-83)
__device__ __forceinline__ bool hits(const Splat2D& s, uint32_t tx, uint32_t ty, uint32_t W,
                                     uint32_t H) {
  if (s.alpha < (1.0f / 255.0f) || W == 0u || H == 0u) return false;
  const uint32_t xs = tx * 16u, ys = ty * 16u;
  if (xs >= W || ys >= H) return false;
  const uint32_t xe = min(xs + 15u, W - 1u), ye = min(ys + 15u, H - 1u);
  const float dx0 = ((float)xs + 0.5f) - s.mu_x;
  const float dx1 = ((float)xe + 0.5f) - s.mu_x;
  const float dy0 = ((float)ys + 0.5f) - s.mu_y;
  const float dy1 = ((float)ye + 0.5f) - s.mu_y;
  const float x_near = cgs_fclamp(0.0f, dx0, dx1);
  const float y1 = cgs_fclamp(-s.b * x_near / cgs_fmax(s.c, 1e-12f), dy0, dy1);
  const float q1 = s.a * x_near * x_near + 2.0f * s.b * x_near * y1 + s.c * y1 * y1;
  const float y_near = cgs_fclamp(0.0f, dy0, dy1);
  const float x2 = cgs_fclamp(-s.b * y_near / cgs_fmax(s.a, 1e-12f), dx0, dx1);
  const float q2 = s.a * x2 * x2 + 2.0f * s.b * x2 * y_near + s.c * y_near * y_near;
  return cgs_fmin(q1, q2) <= s.qcut + 1e-4f;
}

__device__ __forceinline__ Splat2D load_splat(const float* geom, uint32_t i) {
  const float4* g = reinterpret_cast<const float4*>(geom + (size_t)i * 8);
  const float4 g0 = g[0], g1 = g[1];
  Splat2D s;
  s.mu_x = g0.x;
  s.mu_y = g0.y;
  s.a = g0.z;
  s.b = g0.w;
  s.c = g1.x;
  s.alpha = g1.y;
  s.tz = g1.z;
  s.qcut = g1.w;
  return s;
}

// Splats are handled in blocks of BT*BS consecutive splats, thread t owning
// splats base + t*BS .. + BS-1 (so a thread's entries are contiguous).
__device__ __forceinline__ uint32_t count_splat(const CgsBinArgs& A, uint32_t s, uint32_t active,
                                                bool& survived) {
  survived = false;
  if (s >= active) return 0u;
  const uint32_t packed = A.extents[s];
  if (packed == 0u) return 0u;
  survived = true;
  if (A.tiles_x == 0u || A.tiles_y == 0u) return 0u;
  const Splat2D sp = load_splat(A.geom, s);
  const TileBox bx = tile_box(sp.mu_x, sp.mu_y, packed, A.tiles_x, A.tiles_y);
  uint32_t cnt = 0;
  for (uint32_t ty = bx.ty0; ty <= bx.ty1; ++ty)
    for (uint32_t tx = bx.tx0; tx <= bx.tx1; ++tx) cnt += hits(sp, tx, ty, A.width, A.height) ? 1u : 0u;
  return cnt;
}

// count_tiles_kernel (binning.rs:90-164): per-splat tile counts + per-block sums
__global__ __launch_bounds__(BT) void k_count(CgsBinArgs A) {
  __shared__ uint32_t s_tmp[BT / 64];
  const uint32_t t = threadIdx.x;
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  const uint32_t s0 = blockIdx.x * (uint32_t)BP + t * (uint32_t)BS;
  uint32_t sum = 0, surv = 0;
#pragma unroll
  for (int j = 0; j < BS; ++j) {
    const uint32_t s = s0 + (uint32_t)j;
    bool sv;
    const uint32_t c = count_splat(A, s, active, sv);
    if (s < A.max_gaussians) A.counts[s] = c;
    sum += c;
    surv += sv ? 1u : 0u;
  }
  uint32_t tot, stot;
  (void)cgs::block_exclusive_scan<BT>(sum, s_tmp, tot);
  (void)cgs::block_exclusive_scan<BT>(surv, s_tmp, stot);
  if (t == 0) {
    A.block_sums[blockIdx.x] = tot;
    A.block_surv[blockIdx.x] = stot;
  }
}

// digit histograms of one emitted entry for every radix pass that will sort it
__device__ __forceinline__ uint32_t dmask(uint32_t bits, uint32_t shift) {
  const uint32_t w = bits - shift >= 8u ? 8u : bits - shift;
  return (1u << w) - 1u;
}
__device__ __forceinline__ void hist_key(uint32_t* s_hist, uint32_t key, uint32_t bits) {
  for (uint32_t p = 0; p * 8u < bits; ++p)
    atomicAdd(&s_hist[p * 256u + ((key >> (p * 8u)) & dmask(bits, p * 8u))], 1u);
}

// gpu_exclusive_scan_total + emit_tile_keys(_u64)_kernel (scan.rs:198-333,
// binning.rs:168-328).  The prefix of block b is the sum of the block sums
// before it, reduced here by the block itself (no look-back chain, no ticket).
__global__ __launch_bounds__(BT) void k_emit(CgsBinArgs A) {
  __shared__ uint32_t s_tmp[BT / 64];
  __shared__ uint32_t s_hist[2048];
  const uint32_t t = threadIdx.x;
  const uint32_t b = blockIdx.x;
  if (A.hist_partial)
    for (uint32_t i = t; i < A.hist_words; i += BT) s_hist[i] = 0u;
  // prefix of this block and, in the last block, the totals
  uint32_t pre = 0, all_surv = 0;
  const bool last = b == gridDim.x - 1;
  for (uint32_t j = t; j < b; j += BT) pre += A.block_sums[j];
  if (last)
    for (uint32_t j = t; j < gridDim.x; j += BT) all_surv += A.block_surv[j];
  uint32_t pre_tot, surv_tot;
  (void)cgs::block_exclusive_scan<BT>(pre, s_tmp, pre_tot);
  (void)cgs::block_exclusive_scan<BT>(all_surv, s_tmp, surv_tot);
  // thread-blocked offsets inside the block
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  const uint32_t s0 = b * (uint32_t)BP + t * (uint32_t)BS;
  uint32_t cnt[BS];
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < BS; ++j) {
    const uint32_t s = s0 + (uint32_t)j;
    cnt[j] = s < A.max_gaussians ? A.counts[s] : 0u;
    sum += cnt[j];
  }
  uint32_t btot;
  uint32_t off = pre_tot + cgs::block_exclusive_scan<BT>(sum, s_tmp, btot);
  if (last && t == 0) {
    A.total[0] = pre_tot + btot;
    A.total[1] = surv_tot;
  }
  const bool tiles_ok = A.tiles_x != 0u && A.tiles_y != 0u;
#pragma unroll 1
  for (int j = 0; j < BS; ++j) {
    const uint32_t s = s0 + (uint32_t)j;
    if (s >= A.max_gaussians) break;
    if (A.offsets_out) A.offsets_out[s] = off;
    const uint32_t c = cnt[j];
    if (c != 0u && s < active && tiles_ok) {
      const uint32_t packed = A.extents[s];
      const Splat2D sp = load_splat(A.geom, s);
      const TileBox bx = tile_box(sp.mu_x, sp.mu_y, packed, A.tiles_x, A.tiles_y);
      uint32_t cur = off;
      if (!A.mode64) {
        const uint32_t raw = cgs_f2u(cgs_fmax(sp.tz, 1e-6f));
        const uint32_t rel =
            raw > A.base_depth_bits ? min(raw - A.base_depth_bits, 0x0FFFFFFFu) : 0u;
        const uint32_t dq = rel >> (28u - A.depth_bits);
        for (uint32_t ty = bx.ty0; ty <= bx.ty1; ++ty)
          for (uint32_t tx = bx.tx0; tx <= bx.tx1; ++tx)
            if (hits(sp, tx, ty, A.width, A.height)) {
              const uint32_t key = ((ty * A.tiles_x + tx) << A.depth_bits) | dq;
              // only stored entries are histogrammed: the sort reads
              // min(total, cap) entries and its digit offsets must add up to that
              if (cur < A.cap) {
                A.keys[cur] = key;
                A.ids[cur] = s;
                if (A.hist_partial) hist_key(s_hist, key, A.sort_bits);
              }
              ++cur;
            }
      } else {
        const uint32_t dbits = cgs_f2u(cgs_fmax(sp.tz, 0.0f));
        for (uint32_t ty = bx.ty0; ty <= bx.ty1; ++ty)
          for (uint32_t tx = bx.tx0; tx <= bx.tx1; ++tx)
            if (hits(sp, tx, ty, A.width, A.height)) {
              const uint32_t tile = ty * A.tiles_x + tx;
              if (cur < A.cap) {
                A.keys[cur] = tile;
                A.depths[cur] = dbits;
                A.ids[cur] = s;
                if (A.hist_partial) {
                  hist_key(s_hist, dbits, 32u);
                  hist_key(s_hist + 1024, tile, A.tile_bits);
                }
              }
              ++cur;
            }
      }
    }
    off += c;
  }
  if (A.hist_partial) {
    __syncthreads();
    for (uint32_t i = t; i < A.hist_words; i += BT)
      A.hist_partial[(size_t)b * A.hist_words + i] = s_hist[i];
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256u)
    p[i] = v;
}

// identify_tile_ranges_kernel (binning.rs:344-379); shift = depth_bits for
// packed keys, 0 for sorted tile ids (383-417).  Ranges must be zeroed.
__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* sorted, const uint32_t* d_n,
                                                     uint32_t cap, uint32_t shift,
                                                     uint32_t total_tiles, uint32_t* ranges) {
  const uint32_t n = min(*d_n, cap);
  const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
  if (idx >= n) return;
  const uint32_t tile = sorted[idx] >> shift;
  if (tile >= total_tiles) return;
  if (idx == 0u) {
    ranges[tile * 2u] = 0u;
  } else {
    const uint32_t prev = sorted[idx - 1u] >> shift;
    if (tile != prev) {
      if (prev < total_tiles) ranges[prev * 2u + 1u] = idx;
      ranges[tile * 2u] = idx;
    }
  }
  if (idx + 1u == n) ranges[tile * 2u + 1u] = n;
}

__global__ __launch_bounds__(256) void k_gather2(const uint32_t* perm, const uint32_t* a_in,
                                                 const uint32_t* b_in, uint32_t* a_out,
                                                 uint32_t* b_out, const uint32_t* d_n,
                                                 uint32_t cap) {
  const uint32_t n = min(*d_n, cap);
  const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
  if (idx >= n) return;
  const uint32_t p = perm[idx];
  a_out[idx] = a_in[p];
  b_out[idx] = b_in[p];
}

}  // namespace

void cgs_launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  uint64_t blocks = (n + 255u) / 256u;
  if (blocks > 4096u) blocks = 4096u;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n);
}

uint32_t cgs_bin_num_parts(uint32_t max_gaussians) {
  return (max_gaussians + (uint32_t)BP - 1u) / (uint32_t)BP;
}

void cgs_launch_bin(const CgsBinArgs& a, hipStream_t s, CgsProf* prof) {
  const uint32_t parts = cgs_bin_num_parts(a.max_gaussians);
  if (parts == 0) return;
  if (prof) prof->mark("k_count", s);
  hipLaunchKernelGGL(k_count, dim3(parts), dim3(BT), 0, s, a);
  if (prof) prof->mark("k_emit", s);
  hipLaunchKernelGGL(k_emit, dim3(parts), dim3(BT), 0, s, a);
}

void cgs_launch_tile_ranges(const uint32_t* sorted, const uint32_t* d_n, uint32_t cap,
                            uint32_t shift, uint32_t total_tiles, uint32_t* ranges,
                            hipStream_t s) {
  if (cap == 0) return;
  hipLaunchKernelGGL(k_tile_ranges, dim3((cap + 255u) / 256u), dim3(256), 0, s, sorted, d_n, cap,
                     shift, total_tiles, ranges);
}

void cgs_launch_gather2(const uint32_t* perm, const uint32_t* a_in, const uint32_t* b_in,
                        uint32_t* a_out, uint32_t* b_out, const uint32_t* d_n, uint32_t cap,
                        hipStream_t s) {
  if (cap == 0) return;
  hipLaunchKernelGGL(k_gather2, dim3((cap + 255u) / 256u), dim3(256), 0, s, perm, a_in, b_in,
                     a_out, b_out, d_n, cap);
}
