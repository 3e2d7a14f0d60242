// k_radix.hip — stable LSD radix sort of (u32 key, u32 value) pairs.
//
// Replaces gpu_radix_sort_u32_pairs (src/render/kernels/radix_sort.rs:292-368):
// the reference runs 4-bit digits, each pass as histogram + 3-level scan +
// scatter with 16 per-digit block scans.  Here a pass is ONE kernel (the
// "onesweep" scheme): 8-bit (or 4-bit) digits; a partition of 4096 to 16384
// keys (a tuning choice) is ranked with wave64 ballots (8 ballots give the
// mask of lanes sharing a digit) and per-wave LDS digit counters, partitions
// chain their per-digit counts with a decoupled look-back, keys are staged in
// LDS in digit order and written out as contiguous runs.  All digit
// histograms of all passes come from one upfront read of the keys (or from
// the binning pass that emitted them) as per-block partials summed by a small
// reduction: no global atomics besides one partition ticket per live
// partition.  Ranking keeps the input order of equal digits, so each pass and
// the whole sort are stable, which is the reference's order.
#include "cgs_internal.h"

namespace {

constexpr int RI = 16;           // keys per thread (12 for the 12288-key partitions)
constexpr int RP_MIN = 256 * RI; // smallest partition (status arrays are sized for it)
constexpr int RADIX = 256;       // digits of the widest digit (8 bits)
constexpr int MAX_PASSES = 4;    // passes of 8-bit digits over 32 bits
constexpr int HIST_BLOCKS = 512; // partial histograms of the standalone sort

__device__ __forceinline__ uint32_t digit_mask(uint32_t bits, uint32_t shift, uint32_t db) {
  const uint32_t w = bits - shift >= db ? db : bits - shift;
  return (1u << w) - 1u;
}

// Partial digit histograms of all passes, one row of MAX_PASSES*256 words per
// block (no global atomics: same-address device atomics serialise).  Each
// thread reads four keys per step (one 16-byte load when the four are in
// range) and adds every digit of every pass to the block's LDS histogram.
// The passes sort key bits [begin, bits): the keys are read shifted right by
// begin, over bits - begin bits.
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, const uint32_t* d_n,
                                                    uint32_t n_static, uint32_t cap,
                                                    uint32_t bits, uint32_t db, uint32_t begin,
                                                    uint32_t* partial) {
  __shared__ uint32_t s_h[MAX_PASSES * RADIX];
  bits -= begin;
  const uint32_t passes = (bits + db - 1u) / db;
  const uint32_t radix = 1u << db;
  for (uint32_t i = threadIdx.x; i < MAX_PASSES * RADIX; i += 256u) s_h[i] = 0u;
  __syncthreads();
  const uint32_t n = min(d_n ? *d_n : n_static, cap);
  const uint32_t stride = gridDim.x * 1024u;
  for (uint32_t i0 = blockIdx.x * 1024u + threadIdx.x * 4u; i0 < n; i0 += stride) {
    uint32_t k[4];
    if (i0 + 3u < n && (reinterpret_cast<uintptr_t>(keys + i0) & 15u) == 0u) {
      const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
      k[0] = v.x >> begin; k[1] = v.y >> begin; k[2] = v.z >> begin; k[3] = v.w >> begin;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) k[j] = i0 + (uint32_t)j < n ? keys[i0 + j] >> begin : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (i0 + (uint32_t)j >= n) break;
      for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t sh = p * db;
        atomicAdd(&s_h[p * radix + ((k[j] >> sh) & digit_mask(bits, sh, db))], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < MAX_PASSES * RADIX; i += 256u)
    partial[(size_t)blockIdx.x * (MAX_PASSES * RADIX) + i] = s_h[i];
}

// hist[b] = sum over rows of partial[row][b], b < bins: each block owns 16 bins,
// its 16 row-chunks are summed in parallel and combined in LDS.
// hist[b] += sum over the rows of a slice of partial[row][b]: block (x, y)
// sums bins [64x, 64x + 64) over the y-th slice of the rows, four rows at a
// time (one wave per row, lanes over bins: 256-byte row reads), and adds its
// totals to hist, which must start at zero.
__global__ __launch_bounds__(256) void k_hist_reduce(const uint32_t* partial, uint32_t rows,
                                                     uint32_t row_words, uint32_t bins,
                                                     uint32_t* hist) {
  __shared__ uint32_t s_sum[4][64];
  const uint32_t t = threadIdx.x, lane = t & 63u, grp = t >> 6;
  const uint32_t bin = blockIdx.x * 64u + lane;
  const uint32_t per = (rows + gridDim.y - 1u) / gridDim.y;
  const uint32_t r0 = blockIdx.y * per, r1 = min(rows, r0 + per);
  uint32_t acc = 0;
  if (bin < bins) {
#pragma unroll 4
    for (uint32_t r = r0 + grp; r < r1; r += 4u) acc += partial[(size_t)r * row_words + bin];
  }
  s_sum[grp][lane] = acc;
  __syncthreads();
  if (t < 64u && bin < bins) {
    const uint32_t v = s_sum[0][t] + s_sum[1][t] + s_sum[2][t] + s_sum[3][t];
    if (v != 0u) atomicAdd(&hist[bin], v);
  }
}

struct PassArgs {
  const uint32_t* keys_in;
  const uint32_t* vals_in;  // null: value = input index (iota)
  uint32_t* keys_out;
  uint32_t* vals_out;
  const uint32_t* d_n;
  uint32_t n_static;
  uint32_t cap;
  uint32_t bits;
  uint32_t shift;
  const uint32_t* hist;  // this pass's 256 bins
  uint32_t* status;      // this pass's [parts][256]
  uint32_t* ticket;
};

// One pass over partitions of RT*KI keys with DB-bit digits (KI keys per thread).
template <int RT, int DB, int KI = RI>
__global__ __launch_bounds__(RT) void k_radix_onesweep(PassArgs A) {
  constexpr int RP = RT * KI;
  constexpr int RW = RT / 64;
  constexpr int NB = 1 << DB;  // digits
  __shared__ uint2 s_kv[RP];  // (key, value) staged in digit order
  __shared__ uint32_t s_whist[RW][NB];
  __shared__ uint32_t s_lstart[NB];
  __shared__ uint32_t s_gadj[NB];
  __shared__ uint32_t s_tmp[RW];
  __shared__ uint32_t s_part;

  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & 63u, wave = t >> 6;
  const uint32_t n = min(A.d_n ? *A.d_n : A.n_static, A.cap);
  // blocks beyond the live entries leave before taking a ticket (each ticket
  // is a same-address device atomic); the tickets taken are exactly 0..P-1
  if (blockIdx.x * (uint32_t)RP >= n) return;
  if (t == 0) s_part = atomicAdd(A.ticket, 1u);
  for (uint32_t i = t; i < (uint32_t)(RW * NB); i += RT) (&s_whist[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t part = s_part;
  const uint32_t base = part * (uint32_t)RP;
  const uint32_t cnt_part = min((uint32_t)RP, n - base);
  const uint32_t dmask = digit_mask(A.bits, A.shift, (uint32_t)DB);

  // rank: wave w owns keys [base + w*1024, base + (w+1)*1024), item i of lane l
  // at offset i*64 + l, so items are visited in input order
  // offs: each key's rank among the wave's keys of its digit (< 1024, or
  // 0xFFFF for no key), two per register, so that keys, values and ranks of
  // the 1024-thread shape fit its 128 VGPRs without spilling
  uint32_t keys[KI], vals[KI], offs[KI / 2];
  const uint32_t wbase = base + wave * (uint32_t)(KI * 64);
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const uint32_t idx = wbase + (uint32_t)i * 64u + lane;
    keys[i] = idx < n ? __builtin_nontemporal_load(&A.keys_in[idx]) : 0u;
  }
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const uint32_t idx = wbase + (uint32_t)i * 64u + lane;
    const bool valid = idx < n;
    const uint32_t d = (keys[i] >> A.shift) & dmask;
    // lanes with the same digit: for each bit, keep the lanes whose bit equals
    // this lane's (sx = all ones when the bit is set, so ~(bb ^ sx) keeps them)
    uint32_t mlo = (uint32_t)__ballot(valid), mhi = (uint32_t)(__ballot(valid) >> 32);
#pragma unroll
    for (int b = 0; b < DB; ++b) {
      const uint32_t sx = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
      const uint64_t bb = __ballot(sx != 0u);
      mlo &= ~((uint32_t)bb ^ sx);
      mhi &= ~((uint32_t)(bb >> 32) ^ sx);
    }
    const uint64_t m = ((uint64_t)mhi << 32) | mlo;
    const uint32_t prior = cgs::mask_rank(m);
    uint32_t wcount = 0;
    if (valid) wcount = s_whist[wave][d];
    // every lane has read its counter before the group leader bumps it
    __builtin_amdgcn_wave_barrier();
    if (valid && prior == 0u) s_whist[wave][d] = wcount + (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
    const uint32_t rank = valid ? wcount + prior : 0xFFFFu;
    if (i & 1)
      offs[i >> 1] |= rank << 16;
    else
      offs[i >> 1] = rank;
  }
  // the values are loaded once the keys are ranked: their latency overlaps the
  // digit scans and the look-back, and they are not live during the ranking
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const uint32_t idx = wbase + (uint32_t)i * 64u + lane;
    vals[i] = A.vals_in ? (idx < n ? __builtin_nontemporal_load(&A.vals_in[idx]) : 0u) : idx;
  }
  __syncthreads();

  // per digit (thread t == digit t < 256): wave prefixes, partition count;
  // the count is published for the successors' look-back at once, before
  // the block scans
  uint32_t dtotal = 0;
  if (t < (uint32_t)NB) {
#pragma unroll
    for (int w = 0; w < RW; ++w) {
      const uint32_t c = s_whist[w][t];
      s_whist[w][t] = dtotal;
      dtotal += c;
    }
    cgs::lookback_publish<uint32_t, 30>(A.status + t, part, NB, dtotal);
  }
  uint32_t htot;
  const uint32_t lstart = cgs::block_exclusive_scan<RT>(dtotal, s_tmp, htot);
  const uint32_t hv = t < (uint32_t)NB ? A.hist[t] : 0u;
  uint32_t hsum;
  const uint32_t dbase = cgs::block_exclusive_scan<RT>(hv, s_tmp, hsum);
  if (t < (uint32_t)NB) {
    s_lstart[t] = lstart;
    // look-back for this digit over the partitions before this one
    const uint32_t excl = cgs::lookback_walk_u32(A.status + t, part, NB, dtotal);
    s_gadj[t] = dbase + excl - lstart;
  }
  __syncthreads();
  // each wave's digit offsets become partition offsets (one LDS read per key
  // below instead of two)
  for (uint32_t i = t; i < (uint32_t)(RW * NB); i += RT) (&s_whist[0][0])[i] += s_lstart[i % NB];
  __syncthreads();

  // stage (key, value) pairs in LDS in digit order
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const uint32_t rank = (offs[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
    if (rank != 0xFFFFu) {
      const uint32_t d = (keys[i] >> A.shift) & dmask;
      s_kv[s_whist[wave][d] + rank] = make_uint2(keys[i], vals[i]);
    }
  }
  __syncthreads();

  // write out: equal digits are consecutive in LDS and in the output.  A
  // pass reads its input and writes its output once each: both go
  // non-temporal (measured: Stage 3 174 -> 170 us at 1080p; the loads alone
  // changed nothing)
  for (uint32_t j = t; j < cnt_part; j += RT) {
    const uint2 kv = s_kv[j];
    const uint32_t d = (kv.x >> A.shift) & dmask;
    const uint32_t dst = s_gadj[d] + j;
    if (dst < n) {  // always true when the histogram matches the n keys read
      __builtin_nontemporal_store(kv.x, &A.keys_out[dst]);
      __builtin_nontemporal_store(kv.y, &A.vals_out[dst]);
    }
  }
}

}  // namespace

uint32_t cgs_radix_num_parts(uint32_t cap) { return (cap + (uint32_t)RP_MIN - 1u) / (uint32_t)RP_MIN; }
uint32_t cgs_radix_hist_blocks() { return HIST_BLOCKS; }

size_t cgs_radix_control_words(uint32_t cap) {
  const size_t parts = cgs_radix_num_parts(cap) ? cgs_radix_num_parts(cap) : 1u;
  return (size_t)MAX_PASSES * RADIX + (size_t)MAX_PASSES * parts * RADIX + MAX_PASSES +
         (size_t)HIST_BLOCKS * MAX_PASSES * RADIX;
}

void cgs_launch_hist_reduce(const uint32_t* partial, uint32_t rows, uint32_t row_words,
                            uint32_t bins, uint32_t* hist, hipStream_t s, bool zeroed) {
  if (bins == 0) return;
  if (!zeroed) cgs_launch_fill_u32(hist, 0u, bins, s);
  if (rows == 0) return;
  // slices of >= 32 rows, at most 64 of them
  uint32_t slices = rows / 32u;
  slices = slices < 1u ? 1u : slices > 64u ? 64u : slices;
  hipLaunchKernelGGL(k_hist_reduce, dim3((bins + 63u) / 64u, slices), dim3(256), 0, s, partial,
                     rows, row_words, bins, hist);
}

int cgs_launch_radix_sort(const CgsRadixBuffers& b, const uint32_t* d_n, uint32_t n_static,
                          uint32_t bits, bool iota_vals, bool hist_given, hipStream_t s,
                          CgsProf* prof, uint32_t part_keys, uint32_t digit_bits,
                          uint32_t begin_bit) {
  if (bits > 32) bits = 32;
  if (bits <= begin_bit || b.cap == 0) return 0;
  const uint32_t db = digit_bits == 4u ? 4u : 8u;
  const uint32_t radix = 1u << db;
  const uint32_t passes = (bits - begin_bit + db - 1u) / db;
  const uint32_t parts_alloc = cgs_radix_num_parts(b.cap);  // status rows per pass
  const uint32_t pk = part_keys == 4096u    ? 4096u
                      : part_keys == 8192u  ? 8192u
                      : part_keys == 10240u ? 10240u
                      : part_keys == 12288u ? 12288u
                                            : 16384u;
  const uint32_t grid = (b.cap + pk - 1u) / pk;
  if (!hist_given) {
    if (prof) prof->mark("k_radix_hist", s);
    hipLaunchKernelGGL(k_radix_hist, dim3(HIST_BLOCKS), dim3(256), 0, s, b.keys, d_n, n_static,
                       b.cap, bits, db, begin_bit, b.hist_partial);
    if (prof) prof->mark("k_hist_reduce", s);
    cgs_launch_hist_reduce(b.hist_partial, HIST_BLOCKS, MAX_PASSES * RADIX, passes * radix,
                           b.hist, s, false);
  }
  const uint32_t* kin = b.keys;
  const uint32_t* vin = iota_vals ? nullptr : b.vals;
  int in_alt = 0;
  for (uint32_t p = 0; p < passes; ++p) {
    PassArgs A;
    A.keys_in = kin;
    A.vals_in = vin;
    A.keys_out = in_alt ? b.keys : b.keys_alt;
    A.vals_out = in_alt ? b.vals : b.vals_alt;
    A.d_n = d_n;
    A.n_static = n_static;
    A.cap = b.cap;
    A.bits = bits;
    A.shift = begin_bit + p * db;
    A.hist = b.hist + p * radix;
    A.status = b.status + (size_t)p * parts_alloc * radix;
    A.ticket = b.tickets + p;
    if (prof) prof->mark("k_radix_onesweep", s);
    if (db == 8u) {
      if (pk == 16384u)
        hipLaunchKernelGGL((k_radix_onesweep<1024, 8>), dim3(grid), dim3(1024), 0, s, A);
      else if (pk == 12288u)
        hipLaunchKernelGGL((k_radix_onesweep<1024, 8, 12>), dim3(grid), dim3(1024), 0, s, A);
      else if (pk == 10240u)
        hipLaunchKernelGGL((k_radix_onesweep<640, 8>), dim3(grid), dim3(640), 0, s, A);
      else if (pk == 8192u)
        hipLaunchKernelGGL((k_radix_onesweep<512, 8>), dim3(grid), dim3(512), 0, s, A);
      else
        hipLaunchKernelGGL((k_radix_onesweep<256, 8>), dim3(grid), dim3(256), 0, s, A);
    } else {
      if (pk == 16384u)
        hipLaunchKernelGGL((k_radix_onesweep<1024, 4>), dim3(grid), dim3(1024), 0, s, A);
      else if (pk == 12288u)
        hipLaunchKernelGGL((k_radix_onesweep<1024, 4, 12>), dim3(grid), dim3(1024), 0, s, A);
      else if (pk == 10240u)
        hipLaunchKernelGGL((k_radix_onesweep<640, 4>), dim3(grid), dim3(640), 0, s, A);
      else if (pk == 8192u)
        hipLaunchKernelGGL((k_radix_onesweep<512, 4>), dim3(grid), dim3(512), 0, s, A);
      else
        hipLaunchKernelGGL((k_radix_onesweep<256, 4>), dim3(grid), dim3(256), 0, s, A);
    }
    kin = A.keys_out;
    vin = A.vals_out;
    in_alt ^= 1;
  }
  return in_alt;
}
