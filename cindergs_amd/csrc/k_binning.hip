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

// threads per block (8 wave64s), one splat per thread: 512 measured faster
// than 1024 (k_emit -27 % at 1080p, -31 % on 65M-entry 4K views) and than 256
constexpr int BT = 512;
constexpr int BP = BT;        // splats per block
constexpr uint32_t SMALL_BOX = 8;  // boxes of at most this many tiles: walked by their thread
constexpr uint32_t KCHUNK = 32;    // hit masks kept per wave for the wave walk
// words of the bucket mode's LDS tile histogram: 8192 tiles (1080p: 8160) or
// 32768 tiles (2160p: 32400)
constexpr int BIN_HW_SMALL = 4096;
constexpr int BIN_HW_LARGE = 16384;

struct TileBox {
  uint32_t tx0, tx1, ty0, ty1;
};

// compute_gaussian_tile_bounds_xy (tiles/mod.rs:97-130, the GPU form in
// binning.rs:52-63) with the tile size as a parameter.  The tile sizes are
// powers of two (8, 16, 32), so x / w and x * (1 / w) are the same number
// (1 / w is exact and the product is the exact quotient, rounded once): the
// box is the reference's division bit for bit, without the division
// sequence.  inv_tw, inv_th = 1 / tile width, 1 / tile height.
__device__ __forceinline__ TileBox tile_box(float mu_x, float mu_y, uint32_t packed,
                                            uint32_t tiles_x, uint32_t tiles_y, float inv_tw,
                                            float inv_th) {
  const float ext_x = (float)(packed & 0xFFFFu);
  const float ext_y = (float)((packed >> 16) & 0xFFFFu);
  const int32_t mtx = (int32_t)tiles_x - 1, mty = (int32_t)tiles_y - 1;
  TileBox b;
  b.tx0 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_x - ext_x) * inv_tw)), 0, mtx);
  b.tx1 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_x + ext_x) * inv_tw)), 0, mtx);
  b.ty0 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_y - ext_y) * inv_th)), 0, mty);
  b.ty1 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_y + ext_y) * inv_th)), 0, mty);
  return b;
}

// ellipse_intersects_tile_gpu (binning.rs:35-83) on the tile's pixel-centre
// rectangle; d0 = (mu_x, mu_y, a, b), d1 = (c, alpha, q_cut, -)
__device__ __forceinline__ bool hits(const float4& d0, const float4& d1, uint32_t tx,
                                     uint32_t ty, uint32_t W, uint32_t H, uint32_t tw,
                                     uint32_t th) {
  const float mu_x = d0.x, mu_y = d0.y, a = d0.z, b = d0.w, c = d1.x;
  if (d1.y < (1.0f / 255.0f) || W == 0u || H == 0u) return false;
  const uint32_t xs = tx * tw, ys = ty * th;
  if (xs >= W || ys >= H) return false;
  const uint32_t xe = min(xs + tw - 1u, W - 1u), ye = min(ys + th - 1u, H - 1u);
  const float dx0 = ((float)xs + 0.5f) - mu_x;
  const float dx1 = ((float)xe + 0.5f) - mu_x;
  const float dy0 = ((float)ys + 0.5f) - mu_y;
  const float dy1 = ((float)ye + 0.5f) - mu_y;
  const float x_near = cgs_fclamp(0.0f, dx0, dx1);
  const float y1 = cgs_fclamp(-b * x_near / cgs_fmax(c, 1e-12f), dy0, dy1);
  const float q1 = a * x_near * x_near + 2.0f * b * x_near * y1 + c * y1 * y1;
  const float y_near = cgs_fclamp(0.0f, dy0, dy1);
  const float x2 = cgs_fclamp(-b * y_near / cgs_fmax(a, 1e-12f), dx0, dx1);
  const float q2 = a * x2 * x2 + 2.0f * b * x2 * y_near + c * y_near * y_near;
  return cgs_fmin(q1, q2) <= d1.z + 1e-4f;
}

// 32-bit mode key bits of a splat: the depth digit of the packed key
// (emit_tile_keys_kernel, binning.rs:168-247); 64-bit mode: raw depth bits
__device__ __forceinline__ uint32_t splat_key_bits(const CgsBinArgs& A, float tz) {
  if (!A.mode64) {
    const uint32_t raw = cgs_f2u(cgs_fmax(tz, 1e-6f));
    const uint32_t rel = raw > A.base_depth_bits ? min(raw - A.base_depth_bits, 0x0FFFFFFFu) : 0u;
    return rel >> (28u - A.depth_bits);
  }
  return cgs_f2u(cgs_fmax(tz, 0.0f));
}

// the splat at binning position s (identity unless the splats were presorted)
__device__ __forceinline__ uint32_t splat_at(const CgsBinArgs& A, uint32_t s) {
  return (A.order && s < A.max_gaussians) ? A.order[s] : s;
}

// One splat as the binning sees it (thread-owned)
struct BinSplat {
  float4 d0, d1;  // (mu_x, mu_y, a, b), (c, alpha, q_cut, key bits)
  TileBox box;
  uint32_t area;  // tiles of the box (0: hits nothing)
  bool surv;      // extents != 0 and index < active
};

__device__ __forceinline__ BinSplat load_bin_splat(const CgsBinArgs& A, uint32_t s,
                                                   uint32_t active) {
  BinSplat b;
  b.area = 0;
  b.surv = false;
  if (s >= active) return b;
  const uint32_t packed = A.extents[s];
  if (packed == 0u) return b;
  b.surv = true;
  if (A.tiles_x == 0u || A.tiles_y == 0u) return b;
  const float4* g = reinterpret_cast<const float4*>(A.geom + (size_t)s * 8);
  const float4 g0 = g[0], g1 = g[1];
  b.d0 = g0;
  b.d1 = make_float4(g1.x, g1.y, g1.w, cgs_u2f(splat_key_bits(A, g1.z)));
  // tile sizes are powers of two (checked by cgs_set_tuning): exact reciprocals
  b.box = tile_box(g0.x, g0.y, packed, A.tiles_x, A.tiles_y, 1.0f / (float)A.tile_w,
                   1.0f / (float)A.tile_h);
  // a splat below the opacity cut hits no tile (hits() rejects it first)
  if (g1.y >= (1.0f / 255.0f))
    b.area = (b.box.tx1 - b.box.tx0 + 1u) * (b.box.ty1 - b.box.ty0 + 1u);
  return b;
}

// A wave's splats with more than SMALL_BOX tiles ("big"), in splat order, and
// the start of each one's tiles in the wave's flattened (splat, tile) item
// list (tiles row-major inside the box: the reference's emission order).  The
// wave walks that list 64 items at a time, so a few large splats do not hold
// the other lanes of their wave.
struct WaveBig {
  uint32_t start[65];
  float4 d0[64];
  float4 d1[64];
  uint4 box[64];     // tx0, ty0, box width, splat index
  uint32_t acc[64];  // count: hits so far; emit: next output slot
  uint32_t mark[64];
};

// stages this lane's splat if it is big (in slot `slot`); returns the wave's
// item total
__device__ __forceinline__ uint32_t stage_big(WaveBig& W, const BinSplat& b, uint32_t s,
                                              uint32_t acc0, uint32_t lane, uint32_t& nv,
                                              uint32_t& slot) {
  const bool big = b.area > SMALL_BOX;
  const uint64_t mb = __ballot(big);
  nv = (uint32_t)__popcll(mb);
  const uint32_t a = big ? b.area : 0u;
  const uint32_t inc = cgs::wave_inclusive_scan(a);
  const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
  slot = cgs::mask_rank(mb);
  if (big) {
    const uint32_t p = slot;
    W.start[p] = inc - a;
    W.d0[p] = b.d0;
    W.d1[p] = b.d1;
    W.box[p] = make_uint4(b.box.tx0, b.box.ty0, b.box.tx1 - b.box.tx0 + 1u, s);
    W.acc[p] = acc0;
  }
  if (lane == 0) W.start[nv] = total;
  W.mark[lane] = 0u;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return total;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 64-item chunk of the flattened walk starting at item `base`: finds
// this lane's big splat (slot `sp`) and tile, and whether the tile is hit.
// `cp` is the first big splat whose items start at or after `base`.  Returns
// the lanes that start a splat run in the chunk, lane 0 always included.
// With `test` false the hit is taken from `known` (the mask k_count recorded
// for this chunk) instead of the ellipse test.
__device__ __forceinline__ uint64_t walk_chunk(const CgsBinArgs& A, WaveBig& W, uint32_t base,
                                               uint32_t total, uint32_t nv, uint32_t lane,
                                               uint32_t& cp, uint32_t& sp, uint32_t& tile_x,
                                               uint32_t& tile_y, bool& hit, bool test = true,
                                               uint64_t known = 0ull) {
  // big splats whose first item falls in the chunk mark their first lane
  const uint32_t j = cp + lane;
  const uint32_t st = j < nv ? W.start[j] : 0xFFFFFFFFu;
  const bool head = st < base + 64u;
  if (head) W.mark[st - base] = 1u;
  wave_sync_lds();
  const uint64_t M = __ballot(W.mark[lane] != 0u);
  W.mark[lane] = 0u;
  const uint32_t heads = (uint32_t)__popcll(__ballot(head));
  const uint32_t upto = cgs::mask_rank(M) + (uint32_t)((M >> lane) & 1ull);
  sp = cp + upto - 1u;  // cp - 1 = the splat continued from the previous chunk
  cp += heads;
  const uint32_t item = base + lane;
  hit = false;
  tile_x = tile_y = 0u;
  if (item < total) {
    const uint32_t k = item - W.start[sp];
    const uint4 bx = W.box[sp];
    // k / width through the reciprocal, corrected to the exact quotient
    uint32_t q = (uint32_t)((float)k * __builtin_amdgcn_rcpf((float)bx.z));
    int32_t r = (int32_t)k - (int32_t)(q * bx.z);
    if (r < 0) {
      q--;
      r += (int32_t)bx.z;
    } else if (r >= (int32_t)bx.z) {
      q++;
      r -= (int32_t)bx.z;
    }
    tile_x = bx.x + (uint32_t)r;
    tile_y = bx.y + q;
    hit = test ? hits(W.d0[sp], W.d1[sp], tile_x, tile_y, A.width, A.height, A.tile_w, A.tile_h)
               : ((known >> lane) & 1ull) != 0ull;
  }
  return M | 1ull;
}

// lanes [first lane of this lane's run, this lane)
__device__ __forceinline__ uint64_t run_before(uint64_t R, uint32_t lane) {
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint64_t le = lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull);
  const uint32_t h = 63u - (uint32_t)__builtin_clzll(R & le);
  return lt & ~((1ull << h) - 1ull);
}

// the whole run of a run-head lane
__device__ __forceinline__ uint64_t run_of_head(uint64_t R, uint32_t lane) {
  const uint64_t above = lane == 63u ? 0ull : R & (~0ull << (lane + 1u));
  const uint32_t nxt = above ? (uint32_t)__builtin_ctzll(above) : 64u;
  const uint64_t lo = nxt == 64u ? ~0ull : ((1ull << nxt) - 1ull);
  return lo & (~0ull << lane);
}

// one entry of a tile histogram of u16 pairs (tile t in the low / high half
// of word t / 2; a block covers at most 65535 splats, so no half overflows)
__device__ __forceinline__ void tile_hist_add(uint32_t* h, uint32_t tile) {
  atomicAdd(&h[tile >> 1], 1u << ((tile & 1u) << 4));
}

// count_tiles_kernel (binning.rs:90-164): per-splat tile counts and per
// sub-block sums; in bucket mode (HW > 0) also the block's tile histogram,
// written as row blockIdx.x of tile_hist.  A block takes `subs` consecutive
// sub-blocks of BP splats.
template <int HW>
__global__ __launch_bounds__(BT) void k_count(CgsBinArgs A) {
  __shared__ WaveBig s_w[BT / 64];
  __shared__ uint32_t s_tmp[BT / 64];
  __shared__ uint32_t s_th[HW > 0 ? HW : 1];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  WaveBig& W = s_w[wave];
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  // the spans the later kernels of the frame need cleared (none of them is
  // read or written by this kernel)
  for (int z = 0; z < 3; ++z)
    for (uint32_t i = blockIdx.x * (uint32_t)BT + t; i < A.zero_len[z]; i += gridDim.x * (uint32_t)BT)
      A.zero_at[z][i] = 0u;
  if (HW > 0) {
    for (uint32_t i = t; i < (uint32_t)HW; i += BT) s_th[i] = 0u;
    __syncthreads();
  }
  const uint32_t sb0 = blockIdx.x * A.subs, sb1 = min(sb0 + A.subs, A.parts);
  for (uint32_t sb = sb0; sb < sb1; ++sb) {
    const uint32_t s = sb * (uint32_t)BP + t;  // position (counts, masks)
    const uint32_t sid = splat_at(A, s);       // splat
    const BinSplat b = load_bin_splat(A, sid, active);
    // small boxes: the thread walks them; bit k of the splat's mask = box tile
    // k (row-major) is hit
    uint32_t cnt = 0;
    if (b.area != 0u && b.area <= SMALL_BOX) {
      uint32_t mask = 0u, k = 0u;
      for (uint32_t ty = b.box.ty0; ty <= b.box.ty1; ++ty)
        for (uint32_t tx = b.box.tx0; tx <= b.box.tx1; ++tx, ++k)
          if (hits(b.d0, b.d1, tx, ty, A.width, A.height, A.tile_w, A.tile_h)) {
            mask |= 1u << k;
            if (HW > 0) tile_hist_add(s_th, ty * A.tiles_x + tx);
          }
      cnt = (uint32_t)__popc(mask);
      if (A.small_masks) A.small_masks[s] = (uint8_t)mask;
    }
    // big boxes: the wave walks them together; the hit masks of its first
    // KCHUNK chunks are kept for k_emit
    uint32_t nv, slot;
    const uint32_t total = stage_big(W, b, sid, 0u, lane, nv, slot);
    uint64_t* masks =
        A.chunk_masks ? A.chunk_masks + ((size_t)sb * (BT / 64) + wave) * KCHUNK : nullptr;
    uint32_t cp = 0;
    for (uint32_t base = 0; base < total; base += 64u) {
      uint32_t sp, tx, ty;
      bool hit;
      const uint64_t R = walk_chunk(A, W, base, total, nv, lane, cp, sp, tx, ty, hit);
      const uint64_t Hm = __ballot(hit);
      if (HW > 0 && hit) tile_hist_add(s_th, ty * A.tiles_x + tx);
      if (masks && base / 64u < KCHUNK && lane == 0u) masks[base / 64u] = Hm;
      if (((R >> lane) & 1ull) && base + lane < total)
        W.acc[sp] += (uint32_t)__popcll(Hm & run_of_head(R, lane));
      wave_sync_lds();
    }
    if (b.area > SMALL_BOX) cnt = W.acc[slot];
    if (s < A.max_gaussians) A.counts[s] = cnt;
    uint32_t tot, stot;
    (void)cgs::block_exclusive_scan<BT>(cnt, s_tmp, tot);
    (void)cgs::block_exclusive_scan<BT>(b.surv ? 1u : 0u, s_tmp, stot);
    if (t == 0) {
      A.block_sums[sb] = tot;
      A.block_surv[sb] = stot;
    }
  }
  if (HW > 0) {
    __syncthreads();
    const uint32_t hw = A.tiles_pad / 2u;
    uint32_t* row = A.tile_hist + (size_t)blockIdx.x * hw;
    for (uint32_t i = t; i < hw; i += BT) row[i] = s_th[i];
  }
}

// digit histograms of the emitted entries; entries of the thread walks: one
// LDS atomic per digit (the lanes of a wave hold different splats)
// pass p of a sort with db-bit digits: bins [p << db, (p + 1) << db)
__device__ __forceinline__ uint32_t dmask(uint32_t bits, uint32_t shift, uint32_t db) {
  const uint32_t w = bits - shift >= db ? db : bits - shift;
  return (1u << w) - 1u;
}
__device__ __forceinline__ void hist_key(uint32_t* s_hist, uint32_t key, uint32_t bits,
                                         uint32_t db) {
  for (uint32_t p = 0; p * db < bits; ++p)
    atomicAdd(&s_hist[(p << db) + ((key >> (p * db)) & dmask(bits, p * db, db))], 1u);
}
// The digits of the passes p >= first_p only (one entry), and of the passes
// below first_p, made only of a splat's own bits, once with its entry count.
__device__ __forceinline__ void hist_key_from(uint32_t* s_hist, uint32_t key, uint32_t bits,
                                              uint32_t db, uint32_t first_p) {
  for (uint32_t p = first_p; p * db < bits; ++p)
    atomicAdd(&s_hist[(p << db) + ((key >> (p * db)) & dmask(bits, p * db, db))], 1u);
}
__device__ __forceinline__ void hist_const(uint32_t* s_hist, uint32_t key, uint32_t bits,
                                           uint32_t db, uint32_t passes, uint32_t count) {
  for (uint32_t p = 0; p < passes; ++p)
    atomicAdd(&s_hist[(p << db) + ((key >> (p * db)) & dmask(bits, p * db, db))], count);
}
// Histograms of the wave walk's stored entries.  The lanes of a chunk hold
// tiles of one or a few splats, so the digits made only of a splat's depth
// bits (the low `const_bits` of the key) are shared by a whole run of lanes:
// the run's first lane adds the run's stored count once.  Digits with tile
// bits get one atomic per entry, the top one summed over the wave first.
__device__ __forceinline__ void hist_walk(uint32_t* s_hist, uint32_t key, uint32_t bits,
                                          uint32_t const_bits, uint32_t db, bool store,
                                          uint64_t Hs, bool head, uint64_t run, uint32_t lane) {
  const uint32_t passes = (bits + db - 1u) / db;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t d = (key >> (p * db)) & dmask(bits, p * db, db);
    if ((p + 1u) * db <= const_bits) {
      const uint32_t c = head ? (uint32_t)__popcll(Hs & run) : 0u;
      if (c != 0u) atomicAdd(&s_hist[(p << db) + d], c);
    } else if (p + 1u < passes) {
      if (store) atomicAdd(&s_hist[(p << db) + d], 1u);
    } else {
      uint64_t m = Hs;
      while (m) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(m);
        const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)d, (int)leader);
        const uint64_t same = __ballot(store && d == d0);
        if (lane == leader) atomicAdd(&s_hist[(p << db) + d0], (uint32_t)__popcll(same));
        m &= ~same;
      }
    }
  }
}

// one entry: 32-bit packed key (tile << depth_bits | depth digit), or the
// (tile, raw depth) pair of the 64-bit mode
__device__ __forceinline__ void store_entry(const CgsBinArgs& A, uint32_t pos, uint32_t tile,
                                            uint32_t kb, uint32_t id) {
  if (!A.mode64) {
    A.keys[pos] = (tile << A.depth_bits) | kb;
  } else {
    A.keys[pos] = tile;
    if (A.depths) A.depths[pos] = kb;
  }
  A.ids[pos] = id;
}

// bucket mode: the entry goes to its tile's bucket at the (block, tile) base
// plus its rank among the block's entries of that tile (LDS cursor)
__device__ __forceinline__ void bucket_put(const CgsBinArgs& A, uint32_t* cur, uint32_t tile,
                                           uint32_t kb, uint32_t id) {
  const uint32_t sh = (tile & 1u) << 4;
  const uint32_t r = (atomicAdd(&cur[tile >> 1], 1u << sh) >> sh) & 0xFFFFu;
  const uint32_t slot = A.tile_base[(size_t)blockIdx.x * A.tiles_pad + tile] + r;
  if (slot < A.cap) A.bucket[slot] = ((uint64_t)((tile << A.depth_bits) | kb) << 32) | id;
}

// gpu_exclusive_scan_total + emit_tile_keys(_u64)_kernel (scan.rs:198-333,
// binning.rs:168-328).  The prefix of sub-block sb is the sum of the
// sub-block sums before it, reduced here by the block itself (no look-back
// chain, no ticket).  LSD mode: entries in emission order (splat-major) and
// the digit histograms of the sort; only stored entries are histogrammed (the
// sort reads min(total, cap) entries and its digit offsets must add up to
// that).  Bucket mode (HW > 0): entries into their tile buckets.
template <int HW>
__global__ __launch_bounds__(BT) void k_emit(CgsBinArgs A) {
  __shared__ WaveBig s_w[BT / 64];
  __shared__ uint32_t s_tmp[BT / 64];
  __shared__ uint32_t s_hist[HW > 0 ? HW : 2048];  // digit histograms, or tile cursors
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t blk = blockIdx.x;
  WaveBig& W = s_w[wave];
  const bool hist = HW == 0 && A.hist_partial != nullptr;
  const uint32_t hclear = HW > 0 ? (uint32_t)HW : (hist ? A.hist_words : 0u);
  for (uint32_t i = t; i < hclear; i += BT) s_hist[i] = 0u;
  const uint32_t hbits = A.mode64 ? 32u : A.sort_bits;
  const uint32_t db = A.digit_bits == 4u ? 4u : 8u;
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  const uint32_t sb0 = blk * A.subs, sb1 = min(sb0 + A.subs, A.parts);
  for (uint32_t sb = sb0; sb < sb1; ++sb) {
    // prefix of this sub-block and, in the last one, the totals
    uint32_t pre = 0, all_surv = 0;
    const bool last = sb == A.parts - 1u;
    for (uint32_t j = t; j < sb; j += BT) pre += A.block_sums[j];
    if (last)
      for (uint32_t j = t; j < A.parts; j += BT) all_surv += A.block_surv[j];
    uint32_t pre_tot, surv_tot;
    (void)cgs::block_exclusive_scan<BT>(pre, s_tmp, pre_tot);
    (void)cgs::block_exclusive_scan<BT>(all_surv, s_tmp, surv_tot);
    const uint32_t s = sb * (uint32_t)BP + t;
    const uint32_t c = s < A.max_gaussians ? A.counts[s] : 0u;
    uint32_t btot;
    const uint32_t off = pre_tot + cgs::block_exclusive_scan<BT>(c, s_tmp, btot);
    if (last && t == 0) {
      A.total[0] = pre_tot + btot;
      A.total[1] = surv_tot;
    }
    if (A.offsets_out && s < A.max_gaussians) A.offsets_out[s] = off;
    // staged exactly as k_count staged it (its walk's chunks, whose hit masks
    // are reused here, depend on every big box, hit or not)
    const uint32_t sid = splat_at(A, s);
    const BinSplat b = load_bin_splat(A, sid, active);
    // presort: the histogram of the tile digits only
    const bool tile_hist = hist && (A.mode64 || A.order);
    uint32_t* th = A.order ? s_hist : s_hist + 1024;
    // small boxes: the thread walks them (the tiles k_count found hit, when it
    // kept their mask)
    if (b.area != 0u && b.area <= SMALL_BOX) {
      uint32_t cur = off;
      const uint32_t kb = cgs_f2u(b.d1.w);
      const uint32_t mask = A.small_masks ? (uint32_t)A.small_masks[s] : 0u;
      // passes whose digits hold depth bits only (the splat's own): counted once
      // below with the splat's stored entries; the others per entry
      const uint32_t cpass = A.mode64 ? 4u : min(A.depth_bits, hbits) / db;
      uint32_t k = 0u;
      for (uint32_t ty = b.box.ty0; ty <= b.box.ty1; ++ty)
        for (uint32_t tx = b.box.tx0; tx <= b.box.tx1; ++tx, ++k)
          if (A.small_masks ? ((mask >> k) & 1u) != 0u
                            : hits(b.d0, b.d1, tx, ty, A.width, A.height, A.tile_w, A.tile_h)) {
            const uint32_t tile = ty * A.tiles_x + tx;
            if (HW > 0) {
              bucket_put(A, s_hist, tile, kb, sid);
            } else if (cur < A.cap) {
              store_entry(A, cur, tile, kb, sid);
              if (tile_hist)
                hist_key(th, tile, A.tile_bits, 8u);
              else if (hist)
                hist_key_from(s_hist, (tile << A.depth_bits) | kb, hbits, db, cpass);
            }
            ++cur;
          }
      const uint32_t stored = min(cur, A.cap) - min(off, A.cap);
      if (hist && stored != 0u && !A.order) {
        if (!A.mode64)
          hist_const(s_hist, kb, hbits, db, cpass, stored);
        else
          hist_const(s_hist, kb, 32u, 8u, 4u, stored);
      }
    }
    // big boxes: the wave walks them together; a hit goes to its splat's next
    // slot plus the hits of the same splat on the lanes before it
    uint32_t nv, slot;
    const uint32_t total = stage_big(W, b, sid, off, lane, nv, slot);
    const uint64_t* masks =
        A.chunk_masks ? A.chunk_masks + ((size_t)sb * (BT / 64) + wave) * KCHUNK : nullptr;
    uint32_t cp = 0;
    for (uint32_t base = 0; base < total; base += 64u) {
      uint32_t sp, tx, ty;
      bool hit;
      const bool known = masks && base / 64u < KCHUNK;
      const uint64_t R = walk_chunk(A, W, base, total, nv, lane, cp, sp, tx, ty, hit, !known,
                                    known ? masks[base / 64u] : 0ull);
      const uint64_t Hm = __ballot(hit);
      const bool act = base + lane < total;
      uint32_t pos = 0, id = 0, kb = 0;
      if (act) kb = cgs_f2u(W.d1[sp].w);  // the splat's depth bits, valid on every active lane
      if (hit) {
        pos = W.acc[sp] + (uint32_t)__popcll(Hm & run_before(R, lane));
        id = W.box[sp].w;
      }
      const bool head = ((R >> lane) & 1ull) && act;
      const uint64_t run = head ? run_of_head(R, lane) : 0ull;
      wave_sync_lds();
      if (head) W.acc[sp] += (uint32_t)__popcll(Hm & run);
      const uint32_t tile = ty * A.tiles_x + tx;
      if (HW > 0) {
        if (hit) bucket_put(A, s_hist, tile, kb, id);
      } else {
        const bool store = hit && pos < A.cap;
        const uint64_t Hs = __ballot(store);
        if (store) store_entry(A, pos, tile, kb, id);
        if (hist) {
          if (!A.mode64 && !A.order) {
            hist_walk(s_hist, (tile << A.depth_bits) | kb, hbits, A.depth_bits, db, store, Hs,
                      head, run, lane);
          } else {
            if (!A.order) hist_walk(s_hist, kb, 32u, 32u, 8u, store, Hs, head, run, lane);
            hist_walk(th, tile, A.tile_bits, 0u, 8u, store, Hs, head, run, lane);
          }
        }
      }
      wave_sync_lds();
    }
  }
  if (hist) {
    __syncthreads();
    for (uint32_t i = t; i < A.hist_words; i += BT)
      A.hist_partial[(size_t)blk * A.hist_words + i] = s_hist[i];
  }
}

// the depth bits of every splat's sort key, for the presort (0: emits
// nothing), and the block's digit histograms of the presort's passes (8-bit
// digits of key_bits bits) as row blockIdx.x of `partial`
__global__ __launch_bounds__(256) void k_splat_depth_keys(CgsBinArgs A, uint32_t* keys,
                                                          uint32_t* zero, uint32_t zero_len,
                                                          uint32_t key_bits, uint32_t* partial) {
  __shared__ uint32_t s_h[1024];
  for (uint32_t i = threadIdx.x; i < 1024u; i += 256u) s_h[i] = 0u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < zero_len; i += gridDim.x * 256u)
    zero[i] = 0u;
  __syncthreads();
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  const uint32_t passes = (key_bits + 7u) / 8u;
  for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < A.max_gaussians;
       s += gridDim.x * 256u) {
    uint32_t k = 0u;
    if (s < active && A.extents[s] != 0u) k = splat_key_bits(A, A.geom[(size_t)s * 8 + 6]);
    keys[s] = k;
    for (uint32_t p = 0; p < passes; ++p) atomicAdd(&s_h[(p << 8) + ((k >> (8u * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 1024u; i += 256u)
    partial[(size_t)blockIdx.x * 1024u + i] = s_h[i];
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256u)
    p[i] = v;
}

// identify_tile_ranges_kernel (binning.rs:344-379); shift = depth_bits for
// packed keys, 0 for sorted tile ids (383-417).  Ranges must be zeroed.  Each
// thread takes four consecutive entries (one 16-byte load; `sorted` is
// 16-byte aligned) and the entry before them.
__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* sorted, const uint32_t* d_n,
                                                     uint32_t cap, uint32_t shift,
                                                     uint32_t total_tiles, uint32_t* ranges) {
  const uint32_t n = min(*d_n, cap);
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  const uint32_t i0 = q * 4u;
  if (i0 >= n) return;
  uint32_t k[4];
  if (i0 + 4u <= n) {
    const uint4 v = reinterpret_cast<const uint4*>(sorted)[q];
    k[0] = v.x;
    k[1] = v.y;
    k[2] = v.z;
    k[3] = v.w;
  } else {
    for (uint32_t j = 0; j < 4u; ++j) k[j] = i0 + j < n ? sorted[i0 + j] : 0u;
  }
  uint32_t prev = i0 != 0u ? sorted[i0 - 1u] >> shift : 0u;
  for (uint32_t j = 0; j < 4u; ++j) {
    const uint32_t idx = i0 + j;
    if (idx >= n) break;
    const uint32_t tile = k[j] >> shift;
    if (tile < total_tiles) {
      if (idx == 0u) {
        ranges[tile * 2u] = 0u;
      } else if (tile != prev) {
        if (prev < total_tiles) ranges[prev * 2u + 1u] = idx;
        ranges[tile * 2u] = idx;
      }
      if (idx + 1u == n) ranges[tile * 2u + 1u] = n;
    }
    prev = tile;
  }
}

__global__ __launch_bounds__(256) void k_gather2(const uint32_t* perm, const uint32_t* a_in,
                                                 const uint32_t* b_in, uint32_t* a_out,
                                                 uint32_t* b_out, const uint32_t* d_n,
                                                 uint32_t cap) {
  // four entries per thread: one 16-byte load of the permutation, eight
  // independent gathers in flight, two 16-byte stores (perm, a_out and b_out
  // 16-byte aligned)
  const uint32_t n = min(*d_n, cap);
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  const uint32_t i0 = q * 4u;
  if (i0 >= n) return;
  if (i0 + 4u <= n) {
    const uint4 p = reinterpret_cast<const uint4*>(perm)[q];
    const uint4 a = make_uint4(a_in[p.x], a_in[p.y], a_in[p.z], a_in[p.w]);
    const uint4 b = make_uint4(b_in[p.x], b_in[p.y], b_in[p.z], b_in[p.w]);
    reinterpret_cast<uint4*>(a_out)[q] = a;
    reinterpret_cast<uint4*>(b_out)[q] = b;
  } else {
    for (uint32_t idx = i0; idx < n; ++idx) {
      const uint32_t p = perm[idx];
      a_out[idx] = a_in[p];
      b_out[idx] = b_in[p];
    }
  }
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

size_t cgs_bin_mask_words(uint32_t max_gaussians) {
  const size_t small = ((size_t)max_gaussians + 7u) / 8u * 2u;  // bytes -> words, 8-byte aligned
  const size_t chunk = (size_t)cgs_bin_num_parts(max_gaussians) * (BT / 64) * KCHUNK * 2u;
  return small + chunk;
}

void cgs_bin_masks(uint32_t* buf, uint32_t max_gaussians, uint8_t** small_masks,
                   uint64_t** chunk_masks) {
  const size_t small = ((size_t)max_gaussians + 7u) / 8u * 2u;
  *small_masks = reinterpret_cast<uint8_t*>(buf);
  *chunk_masks = reinterpret_cast<uint64_t*>(buf + small);
}

uint32_t cgs_bin_grid(const CgsBinArgs& a) {
  return a.subs ? (a.parts + a.subs - 1u) / a.subs : 0u;
}

// LSD mode: k_count, k_emit.  Bucket mode: k_count, `between` (the column
// scan), k_emit.
void cgs_launch_bin(const CgsBinArgs& a, hipStream_t s, CgsProf* prof,
                    const CgsBucketArgs* between) {
  const uint32_t grid = cgs_bin_grid(a);
  if (grid == 0) return;
  const uint32_t hw = a.tiles_pad / 2u;
  const int mode = !a.tile_hist ? 0 : hw <= (uint32_t)BIN_HW_SMALL ? 1 : 2;
  if (prof) prof->mark("k_count", s);
  if (mode == 0)
    hipLaunchKernelGGL(k_count<0>, dim3(grid), dim3(BT), 0, s, a);
  else if (mode == 1)
    hipLaunchKernelGGL(k_count<BIN_HW_SMALL>, dim3(grid), dim3(BT), 0, s, a);
  else
    hipLaunchKernelGGL(k_count<BIN_HW_LARGE>, dim3(grid), dim3(BT), 0, s, a);
  if (between) {
    if (prof) prof->mark("k_bucket_scan", s);
    cgs_launch_bucket_scan(*between, s);
  }
  if (prof) prof->mark("k_emit", s);
  if (mode == 0)
    hipLaunchKernelGGL(k_emit<0>, dim3(grid), dim3(BT), 0, s, a);
  else if (mode == 1)
    hipLaunchKernelGGL(k_emit<BIN_HW_SMALL>, dim3(grid), dim3(BT), 0, s, a);
  else
    hipLaunchKernelGGL(k_emit<BIN_HW_LARGE>, dim3(grid), dim3(BT), 0, s, a);
}

uint32_t cgs_launch_splat_depth_keys(const CgsBinArgs& a, uint32_t* keys, uint32_t* zero,
                                     uint32_t zero_len, uint32_t key_bits, uint32_t* partial,
                                     uint32_t max_rows, hipStream_t s) {
  const uint32_t n = max(a.max_gaussians, zero_len);
  if (n == 0 || max_rows == 0) return 0;
  // about 4 splats per thread
  const uint32_t grid = max(1u, min((n + 1023u) / 1024u, max_rows));
  hipLaunchKernelGGL(k_splat_depth_keys, dim3(grid), dim3(256), 0, s, a, keys, zero, zero_len,
                     key_bits, partial);
  return grid;
}

void cgs_launch_tile_ranges(const uint32_t* sorted, const uint32_t* d_n, uint32_t cap,
                            uint32_t shift, uint32_t total_tiles, uint32_t* ranges,
                            hipStream_t s) {
  if (cap == 0) return;
  const uint32_t quads = (cap + 3u) / 4u;
  hipLaunchKernelGGL(k_tile_ranges, dim3((quads + 255u) / 256u), dim3(256), 0, s, sorted, d_n,
                     cap, shift, total_tiles, ranges);
}

bool cgs_launch_gather2(const uint32_t* perm, const uint32_t* a_in, const uint32_t* b_in,
                        uint32_t* a_out, uint32_t* b_out, const uint32_t* d_n, uint32_t cap,
                        hipStream_t s) {
  if (((uintptr_t)perm | (uintptr_t)a_out | (uintptr_t)b_out) & 15u) return false;
  if (cap == 0) return true;
  const uint32_t quads = (cap + 3u) / 4u;
  hipLaunchKernelGGL(k_gather2, dim3((quads + 255u) / 256u), dim3(256), 0, s, perm, a_in, b_in,
                     a_out, b_out, d_n, cap);
  return true;
}
