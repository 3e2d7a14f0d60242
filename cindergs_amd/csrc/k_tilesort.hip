// k_tilesort.hip — Stage 3 as tile buckets and a segmented sort.
//
// The reference orders its tile entries with a stable LSD radix sort of the
// 32-bit keys (tile << depth_bits | depth) over the emission order, splat-major
// (gpu_radix_sort_u32_pairs, src/render/kernels/radix_sort.rs:292-368, called
// from gpu_state.rs:1047-1100).  Keys equal to each other keep their splat
// order, so the result is the order of the pairs (key, splat id), a total
// order: sorting every tile's entries by (key, id) gives the same array.
//
// MI355X design (the binning kernels of k_binning.hip do the first half):
//  * k_count keeps, per block, a tile histogram in LDS (u16 pairs: a block
//    covers at most 65535 splats and a splat hits a tile once) and writes it
//    as one row; no global atomics (contended ones cost ~0.4 ms a frame here);
//  * k_bucket_scan sums the rows per tile (a column scan, decoupled look-back
//    across tile groups): the tile starts, which are the tile ranges, the
//    bucket base of every (block, tile), and the plan of the sort below;
//  * k_emit writes each entry straight into its tile's bucket, ranked inside
//    the block by LDS atomics (so the order inside a bucket is arbitrary);
//  * k_seg_sort sorts (key << 32 | id) inside each tile: windows of 2048
//    entries take the tiles that start in them (at most 4096 entries) and a
//    256-thread block merge-sorts them in registers and LDS; a tile of more
//    than 2048 entries is cut into 4096-entry chunks, sorted the same way,
//    and k_seg_merge merges a big tile's chunks pairwise, one round per
//    launch, writing the final arrays in the tile's last round.
#include "cgs_internal.h"

namespace {

constexpr int SB_T = 128;        // tiles per k_bucket_scan block
constexpr int SB_W = 16;         // its waves: each sums a slice of the rows
constexpr int SB_U = 8;          // rows loaded per round trip
constexpr int SS_T = 256;        // threads of a sort block
constexpr int SS_I = 16;         // values per thread
constexpr uint32_t SS_N = SS_T * SS_I;  // 4096 values per sort block
constexpr uint64_t KMAX = ~0ull;

// LDS index of value i of a sort block: one pad word pair after every 16
// values, so that the blocked reads (16 consecutive values per thread) spread
// over the banks
__device__ __forceinline__ uint32_t pad(uint32_t i) { return i + (i >> 4); }
constexpr uint32_t SS_LDS = SS_N + SS_N / 16;

__device__ __forceinline__ uint32_t ceil_log2(uint32_t v) {
  return v <= 1u ? 0u : 32u - (uint32_t)__builtin_clz(v - 1u);
}

// ---------------------------------------------------------------------------
// column scan of the per-block tile histograms
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SB_W * 64) void k_bucket_scan(CgsBucketArgs A) {
  __shared__ uint32_t s_part[SB_W][SB_T];  // per-wave partial column sums
  __shared__ uint32_t s_start[SB_T];
  __shared__ uint32_t s_tmp[SB_W];
  __shared__ uint32_t s_grp;
  __shared__ uint64_t s_base;
  __shared__ uint32_t s_max[2];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  if (t == 0) {
    s_grp = atomicAdd(A.ticket, 1u);
    s_max[0] = 0u;
    s_max[1] = 0u;
  }
  __syncthreads();
  const uint32_t g = s_grp;
  const uint32_t tile0 = g * (uint32_t)SB_T;
  const uint32_t hw = A.tiles_pad / 2u;
  const uint32_t word = tile0 / 2u + lane;  // this lane's tiles: tile0 + 2 lane, + 1
  const uint32_t per = (A.rows + SB_W - 1u) / SB_W;
  const uint32_t r0 = min(wave * per, A.rows), r1 = min(r0 + per, A.rows);
  // SB_U independent loads per round trip (the rows are in HBM or L2)
  uint32_t lo = 0, hi = 0;
  for (uint32_t r = r0; r < r1; r += SB_U) {
    uint32_t w[SB_U];
#pragma unroll
    for (int u = 0; u < SB_U; ++u) w[u] = r + u < r1 ? A.hist[(size_t)(r + u) * hw + word] : 0u;
#pragma unroll
    for (int u = 0; u < SB_U; ++u) {
      lo += w[u] & 0xFFFFu;
      hi += w[u] >> 16;
    }
  }
  s_part[wave][2u * lane] = lo;
  s_part[wave][2u * lane + 1u] = hi;
  __syncthreads();
  const uint32_t tile = tile0 + t;
  uint32_t n = 0, chunks = 0;
  if (t < (uint32_t)SB_T && tile < A.tiles) {
#pragma unroll
    for (int w = 0; w < SB_W; ++w) n += s_part[w][t];
    chunks = n > A.big ? (n + SS_N - 1u) / SS_N : 0u;
  }
  uint32_t tot_n, tot_c;
  const uint32_t ex_n = cgs::block_exclusive_scan<SB_W * 64>(n, s_tmp, tot_n);
  const uint32_t ex_c = cgs::block_exclusive_scan<SB_W * 64>(chunks, s_tmp, tot_c);
  if (t == 0) s_base = cgs::lookback_u64(A.status, g, ((uint64_t)tot_c << 32) | tot_n);
  __syncthreads();
  const uint32_t base_n = (uint32_t)s_base, base_c = (uint32_t)(s_base >> 32);
  if (t < (uint32_t)SB_T) {
    const uint32_t start = base_n + ex_n;
    s_start[t] = start;
    if (tile < A.tiles) {
      A.tile_start[tile] = start;
      const uint32_t c0 = base_c + ex_c;
      A.chunk_start[tile] = c0;
      // ranges end at the arena's capacity: in a frame that overflowed it (and
      // will run again with a larger arena) Stage 4 reads only stored entries
      const uint32_t rs = min(start, A.cap), re = min(start + n, A.cap);
      A.ranges[2u * tile] = re > rs ? rs : 0u;
      A.ranges[2u * tile + 1u] = re > rs ? re : 0u;
      // window table: win_first[k] = the first tile whose start is >= k x
      // window; the window boundaries in (start, start + n] are followed by
      // the next tile
      if (tile == 0u) A.win_first[0] = 0u;
      for (uint32_t k = start / A.window + 1u; k <= (start + n) / A.window && k < A.win_slots; ++k)
        A.win_first[k] = tile + 1u;
      for (uint32_t c = 0; c < chunks; ++c)
        if (c0 + c < A.chunk_slots) A.chunk_tile[c0 + c] = tile;
      if (n) {
        atomicMax(&s_max[0], n);
        atomicMax(&s_max[1], ceil_log2(chunks));
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    if (s_max[0]) atomicMax(&A.stats[0], s_max[0]);
    if (s_max[1]) atomicMax(&A.stats[1], s_max[1]);
    if (tile0 + (uint32_t)SB_T >= A.tiles) {  // the last group: the totals
      A.tile_start[A.tiles] = base_n + tot_n;
      A.chunk_start[A.tiles] = base_c + tot_c;
      A.stats[2] = base_n + tot_n;
      A.stats[3] = base_c + tot_c;
    }
  }
  // bucket bases: each wave walks its rows again with the running prefix
  const uint32_t ta = 2u * lane, tb = ta + 1u;
  uint32_t run_a = s_start[ta], run_b = s_start[tb];
  for (uint32_t w = 0; w < wave; ++w) {
    run_a += s_part[w][ta];
    run_b += s_part[w][tb];
  }
  for (uint32_t r = r0; r < r1; r += SB_U) {
    uint32_t w[SB_U];
#pragma unroll
    for (int u = 0; u < SB_U; ++u) w[u] = r + u < r1 ? A.hist[(size_t)(r + u) * hw + word] : 0u;
#pragma unroll
    for (int u = 0; u < SB_U; ++u) {
      if (r + u < r1)
        reinterpret_cast<uint2*>(A.base + (size_t)(r + u) * A.tiles_pad + tile0)[lane] =
            make_uint2(run_a, run_b);
      run_a += w[u] & 0xFFFFu;
      run_b += w[u] >> 16;
    }
  }
}

// ---------------------------------------------------------------------------
// block merge sort of up to SS_N 64-bit values
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cas(uint64_t& a, uint64_t& b) {
  const bool sw = b < a;
  const uint64_t lo = sw ? b : a, hi = sw ? a : b;
  a = lo;
  b = hi;
}

// ascending bitonic network over 16 registers
__device__ __forceinline__ void sort16(uint64_t (&v)[SS_I]) {
#pragma unroll
  for (int k = 2; k <= SS_I; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < SS_I; ++i) {
        const int l = i ^ j;
        if (l > i) {
          if ((i & k) == 0)
            cas(v[i], v[l]);
          else
            cas(v[l], v[i]);
        }
      }
}

// Values s[pad(0 .. n_pad)) (n_pad a power of two, 16 <= n_pad <= SS_N) in
// LDS, unsorted: on return v holds positions [16 t, 16 t + 16) of their sorted
// order (threads with 16 t >= n_pad hold nothing useful).  Runs of 16 are
// sorted in registers, then merged pairwise through LDS: each thread finds
// where its 16 outputs start on the merge path of the two runs (binary
// search) and merges them serially.  Ends with a barrier.
__device__ __forceinline__ void block_sort(uint64_t (&v)[SS_I], uint64_t* s, uint32_t n_pad) {
  const uint32_t my0 = threadIdx.x * (uint32_t)SS_I;
  const bool act = my0 < n_pad;
#pragma unroll
  for (int i = 0; i < SS_I; ++i) v[i] = act ? s[pad(my0 + i)] : KMAX;
  sort16(v);
  for (uint32_t w = SS_I; w < n_pad; w <<= 1) {
    __syncthreads();
    if (act)
#pragma unroll
      for (int i = 0; i < SS_I; ++i) s[pad(my0 + i)] = v[i];
    __syncthreads();
    if (act) {
      const uint32_t a0 = my0 & ~(2u * w - 1u), b0 = a0 + w;
      const uint32_t diag = my0 - a0;
      uint32_t lo = diag > w ? diag - w : 0u, hi = min(diag, w);
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[pad(a0 + mid)] <= s[pad(b0 + diag - 1u - mid)])
          lo = mid + 1u;
        else
          hi = mid;
      }
      uint32_t ai = a0 + lo, bi = b0 + diag - lo;
      const uint32_t ae = a0 + w, be = b0 + w;
      uint64_t va = ai < ae ? s[pad(ai)] : KMAX, vb = bi < be ? s[pad(bi)] : KMAX;
#pragma unroll
      for (int i = 0; i < SS_I; ++i) {
        const bool ta = bi >= be || (ai < ae && va <= vb);
        v[i] = ta ? va : vb;
        if (ta) {
          ++ai;
          va = ai < ae ? s[pad(ai)] : KMAX;
        } else {
          ++bi;
          vb = bi < be ? s[pad(bi)] : KMAX;
        }
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n) {
  uint32_t p = (uint32_t)SS_I;
  while (p < n) p <<= 1;
  return p;
}

// sorts src[lo, hi) (hi - lo <= SS_N) into final keys / ids, or into dst64
__device__ __forceinline__ void sort_range(const CgsSegSortArgs& A, const uint64_t* src,
                                           uint32_t lo, uint32_t hi, uint64_t* dst64,
                                           uint64_t* s) {
  const uint32_t n = hi - lo, t = threadIdx.x;
  const uint32_t n_pad = pow2_at_least(n);
  for (uint32_t i = t; i < n_pad; i += SS_T) s[pad(i)] = i < n ? src[lo + i] : KMAX;
  __syncthreads();
  uint64_t v[SS_I];
  block_sort(v, s, n_pad);
  const uint32_t my0 = t * (uint32_t)SS_I;
  if (my0 < n_pad)
#pragma unroll
    for (int i = 0; i < SS_I; ++i) s[pad(my0 + i)] = v[i];
  __syncthreads();
  for (uint32_t i = t; i < n; i += SS_T) {
    const uint64_t x = s[pad(i)];
    if (dst64) {
      dst64[lo + i] = x;
    } else {
      A.keys[lo + i] = (uint32_t)(x >> 32);
      A.ids[lo + i] = (uint32_t)x;
    }
  }
}

// One block per 2048-entry window of the bucket array (the tiles that start
// in it, at most 4096 entries) or per 4096-entry chunk of a tile with more
// than 2048 entries.
__global__ __launch_bounds__(SS_T) void k_seg_sort(CgsSegSortArgs A) {
  __shared__ uint64_t s[SS_LDS];
  const uint32_t total = min(A.stats[2], A.cap);
  const uint32_t b = blockIdx.x;
  if (b < A.windows) {
    const uint32_t w0 = b * A.window;
    if (w0 >= total) return;
    const uint32_t t0 = A.win_first[b];
    uint32_t t1 = (b + 1u) * A.window <= total ? A.win_first[b + 1u] : A.tiles;
    if (t1 <= t0) return;
    // a tile of more than `big` entries that starts here is the last one
    // (it reaches past the window) and is sorted by its chunks
    if (A.tile_start[t1] - A.tile_start[t1 - 1u] > A.big) --t1;
    if (t1 <= t0) return;
    const uint32_t lo = min(A.tile_start[t0], total), hi = min(A.tile_start[t1], total);
    if (hi <= lo) return;
    sort_range(A, A.bucket, lo, min(hi, lo + SS_N), nullptr, s);
    return;
  }
  const uint32_t j = b - A.windows;
  if (j >= min(A.stats[3], A.chunk_slots)) return;
  const uint32_t tile = A.chunk_tile[j];
  const uint32_t ts = A.tile_start[tile], te = A.tile_start[tile + 1u];
  const uint32_t c = j - A.chunk_start[tile];
  const uint32_t lo = min(ts + c * SS_N, total), hi = min(min(lo + SS_N, te), total);
  if (hi <= lo) return;
  const uint32_t nch = (te - ts + SS_N - 1u) / SS_N;
  // a single chunk is final; otherwise the merge rounds take it from alt
  sort_range(A, A.bucket, lo, hi, nch > 1u ? A.alt : nullptr, s);
}

// ---------------------------------------------------------------------------
// merge rounds of the big tiles
// ---------------------------------------------------------------------------

// The number of A values among the first `diag` values of merge(A, B) (ties
// to A), found by the whole block: each round tests 256 evenly spaced
// candidates at once and keeps the interval between the last true and the
// first false test.
__device__ uint32_t coop_merge_path(const uint64_t* a, uint32_t la, const uint64_t* b,
                                    uint32_t lb, uint32_t diag, uint32_t* s_cnt) {
  uint32_t lo = diag > lb ? diag - lb : 0u, hi = min(diag, la);
  const uint32_t t = threadIdx.x;
  while (lo < hi) {
    const uint32_t range = hi - lo;
    const uint32_t step = (range + SS_T - 1u) / SS_T;
    const uint32_t k_n = (range + step - 1u) / step;  // candidates, <= 256
    bool pred = false;
    if (t < k_n) {
      const uint32_t i = lo + t * step;  // < hi <= min(diag, la)
      pred = a[i] <= b[diag - 1u - i];
    }
    const uint64_t m = __ballot(pred);
    if ((t & 63u) == 0u) s_cnt[t >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    const uint32_t cnt = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    __syncthreads();
    // candidates 0 .. cnt-1 are true, cnt .. k_n-1 false
    const uint32_t nlo = cnt ? lo + (cnt - 1u) * step + 1u : lo;
    const uint32_t nhi = cnt < k_n ? lo + cnt * step : hi;
    lo = nlo;
    hi = nhi;
    if (step == 1u) break;
  }
  return lo;
}

// Round r (1-based) of the big tiles: runs of 4096 << (r - 1) values are
// merged pairwise; block j writes output chunk j (the chunk numbering of
// k_seg_sort).  A tile whose last round this is writes the final arrays.
__global__ __launch_bounds__(SS_T) void k_seg_merge(CgsSegSortArgs A, uint32_t round,
                                                    const uint64_t* in, uint64_t* out) {
  __shared__ uint64_t s[SS_LDS];
  __shared__ uint32_t s_cnt[4];
  const uint32_t total = min(A.stats[2], A.cap);
  const uint32_t j = blockIdx.x;
  if (j >= min(A.stats[3], A.chunk_slots)) return;
  const uint32_t tile = A.chunk_tile[j];
  const uint32_t ts = A.tile_start[tile], te = min(A.tile_start[tile + 1u], total);
  if (te <= ts) return;
  const uint32_t n = te - ts;
  const uint32_t nch = (n + SS_N - 1u) / SS_N;
  const uint32_t rounds = ceil_log2(nch);
  if (round > rounds) return;
  const uint32_t c = j - A.chunk_start[tile];
  const uint32_t o0 = c * SS_N;
  if (o0 >= n) return;
  const uint32_t L = SS_N << (round - 1u);
  const uint32_t p0 = o0 & ~(2u * L - 1u);
  const uint32_t la = min(L, n - p0);
  const uint32_t lb = p0 + L < n ? min(L, n - p0 - L) : 0u;
  const uint64_t* a = in + ts + p0;
  const uint64_t* bb = in + ts + p0 + la;
  const uint32_t d0 = o0 - p0, d1 = min(d0 + SS_N, la + lb);
  const uint32_t m = d1 - d0;
  const uint32_t i0 = coop_merge_path(a, la, bb, lb, d0, s_cnt);
  const uint32_t i1 = coop_merge_path(a, la, bb, lb, d1, s_cnt);
  const uint32_t na = i1 - i0, nb = m - na;
  const uint32_t j0 = d0 - i0;
  const uint32_t t = threadIdx.x;
  // the two slices side by side in LDS: A part at [0, na), B part at [na, m)
  for (uint32_t i = t; i < m; i += SS_T)
    s[pad(i)] = i < na ? a[i0 + i] : bb[j0 + (i - na)];
  __syncthreads();
  uint64_t v[SS_I];
  const uint32_t my0 = t * (uint32_t)SS_I;
  if (my0 < m) {
    uint32_t lo = my0 > nb ? my0 - nb : 0u, hi = min(my0, na);
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s[pad(mid)] <= s[pad(na + my0 - 1u - mid)])
        lo = mid + 1u;
      else
        hi = mid;
    }
    uint32_t ai = lo, bi = na + my0 - lo;
    uint64_t va = ai < na ? s[pad(ai)] : KMAX, vb = bi < m ? s[pad(bi)] : KMAX;
#pragma unroll
    for (int i = 0; i < SS_I; ++i) {
      const bool ta = bi >= m || (ai < na && va <= vb);
      v[i] = ta ? va : vb;
      if (ta) {
        ++ai;
        va = ai < na ? s[pad(ai)] : KMAX;
      } else {
        ++bi;
        vb = bi < m ? s[pad(bi)] : KMAX;
      }
    }
  }
  __syncthreads();
  if (my0 < m)
#pragma unroll
    for (int i = 0; i < SS_I; ++i)
      if (my0 + (uint32_t)i < m) s[pad(my0 + i)] = v[i];
  __syncthreads();
  const bool last = round == rounds;
  for (uint32_t i = t; i < m; i += SS_T) {
    const uint64_t x = s[pad(i)];
    const uint32_t pos = ts + o0 + i;
    if (last) {
      A.keys[pos] = (uint32_t)(x >> 32);
      A.ids[pos] = (uint32_t)x;
    } else {
      out[pos] = x;
    }
  }
}

}  // namespace

uint32_t cgs_seg_sort_windows(uint32_t cap) { return (cap + CGS_SEG_WINDOW - 1u) / CGS_SEG_WINDOW; }

void cgs_launch_bucket_scan(const CgsBucketArgs& a, hipStream_t s) {
  const uint32_t groups = a.tiles_pad / (uint32_t)SB_T;
  if (groups == 0) return;
  hipLaunchKernelGGL(k_bucket_scan, dim3(groups), dim3(SB_W * 64), 0, s, a);
}

// k_seg_sort, then `rounds` merge rounds (the k-th reads the buffer the
// previous one wrote: alt, bucket, alt, ...)
void cgs_launch_seg_sort(const CgsSegSortArgs& a, uint32_t chunk_blocks, uint32_t rounds,
                         hipStream_t s, CgsProf* prof) {
  const uint32_t grid = a.windows + chunk_blocks;
  if (grid == 0) return;
  if (prof) prof->mark("k_seg_sort", s);
  hipLaunchKernelGGL(k_seg_sort, dim3(grid), dim3(SS_T), 0, s, a);
  for (uint32_t r = 1; r <= rounds && chunk_blocks; ++r) {
    if (prof) prof->mark("k_seg_merge", s);
    const uint64_t* in = (r & 1u) ? a.alt : a.bucket;
    uint64_t* out = (r & 1u) ? a.bucket : a.alt;
    hipLaunchKernelGGL(k_seg_merge, dim3(chunk_blocks), dim3(SS_T), 0, s, a, r, in, out);
  }
}
