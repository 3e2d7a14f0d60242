// k_scan.hip — single-pass exclusive prefix scan of u32 with total.
//
// Replaces gpu_exclusive_scan_total (src/render/kernels/scan.rs:198-333), a
// 3-level Blelloch scan (512 elements per 256-thread block, 3 launches and a
// host readback of the total).  Here: one launch; each 256-thread partition
// scans 4096 elements (16 per thread, loaded striped for coalescing and
// re-read blocked from LDS), then chains partition totals with a decoupled
// look-back; the total is written to a device scalar.  u32 wrap-around
// matches the reference's u32 accumulation.
#include "cgs_internal.h"

namespace {

constexpr int ST = 256;
constexpr int SI = 16;
constexpr int SP = ST * SI;

__global__ __launch_bounds__(ST) void k_scan(const uint32_t* in, uint32_t* out, uint32_t n,
                                             uint32_t* total_out, uint64_t* status,
                                             uint32_t* ticket) {
  __shared__ uint32_t s_v[SP];
  __shared__ uint32_t s_tmp[ST / 64];
  __shared__ uint32_t s_part;
  __shared__ uint32_t s_base;
  const uint32_t t = threadIdx.x;
  if (t == 0) s_part = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t part = s_part;
  const uint32_t base = part * (uint32_t)SP;
#pragma unroll
  for (int i = 0; i < SI; ++i) {
    const uint32_t g = base + (uint32_t)i * ST + t;
    s_v[i * ST + t] = g < n ? in[g] : 0u;
  }
  __syncthreads();
  uint32_t loc[SI];
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < SI; ++j) {
    loc[j] = s_v[t * SI + j];
    sum += loc[j];
  }
  uint32_t tot;
  const uint32_t ex = cgs::block_exclusive_scan<ST>(sum, s_tmp, tot);
  if (t == 0) {
    const uint64_t e = cgs::lookback_u64(status, part, (uint64_t)tot);
    s_base = (uint32_t)e;
    if (part == gridDim.x - 1) *total_out = (uint32_t)(e + tot);
  }
  __syncthreads();
  uint32_t run = s_base + ex;
#pragma unroll
  for (int j = 0; j < SI; ++j) {
    s_v[t * SI + j] = run;
    run += loc[j];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SI; ++i) {
    const uint32_t g = base + (uint32_t)i * ST + t;
    if (g < n) out[g] = s_v[i * ST + t];
  }
}

}  // namespace

uint32_t cgs_scan_num_parts(uint32_t n) { return (n + (uint32_t)SP - 1u) / (uint32_t)SP; }

void cgs_launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total,
                     uint64_t* status, uint32_t* ticket, hipStream_t s) {
  const uint32_t parts = cgs_scan_num_parts(n);
  if (parts == 0) {
    cgs_launch_fill_u32(total, 0u, 1, s);
    return;
  }
  hipLaunchKernelGGL(k_scan, dim3(parts), dim3(ST), 0, s, in, out, n, total, status, ticket);
}
