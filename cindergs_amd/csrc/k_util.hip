// k_util.hip — layout conversion on upload: the caller's per-splat records
// (GaussianSceneSoa / QuantizedGaussianSceneSoa, src/scene/soa.rs:30-45,
// 429-500) become wave-blocked device arrays (64 splats per block,
// channel-major inside the block) so that every wave of the per-splat kernels
// reads contiguous 256-byte segments from one small region.
#include "cgs_internal.h"

namespace {

template <class T>
__global__ __launch_bounds__(256) void k_to_blocked(const T* in, T* out, uint32_t n,
                                                    uint32_t k_in, uint32_t k_out,
                                                    uint32_t chan0) {
  const uint64_t total = (uint64_t)n * k_in;
  for (uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x; j < total;
       j += (uint64_t)gridDim.x * 256u) {
    const uint32_t i = (uint32_t)(j / k_in), c = (uint32_t)(j % k_in);
    out[cgs_blk(i, chan0 + c, k_out)] = in[j];
  }
}

// f16 SH record (channel-major, degree-trimmed, `stride` halves per splat) ->
// 48 blocked channels in the f32 record order [dc r,g,b, rest_r 15, rest_g 15,
// rest_b 15]; coefficients beyond the degree are zero.
__global__ __launch_bounds__(256) void k_sh_f16_blocked(const uint16_t* in, uint16_t* out,
                                                        uint32_t n, uint32_t stride,
                                                        uint32_t rest) {
  const uint64_t total = (uint64_t)n * 48u;
  for (uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x; j < total;
       j += (uint64_t)gridDim.x * 256u) {
    const uint32_t i = (uint32_t)(j / 48u), k = (uint32_t)(j % 48u);
    int src = -1;
    if (k < 3u) {
      src = (int)k;
    } else {
      const uint32_t ch = (k - 3u) / 15u, r = (k - 3u) % 15u;
      if (r < rest) src = (int)(3u + ch * rest + r);
    }
    out[cgs_blk(i, k, 48u)] = src >= 0 ? in[(size_t)i * stride + (uint32_t)src] : (uint16_t)0;
  }
}

// dst[i] = src[i], i < n (n small: a frame's control words)
__global__ __launch_bounds__(64) void k_copy_words(const uint32_t* src, uint32_t* dst, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += 64u) dst[i] = src[i];
}

uint32_t grid_for(uint64_t total) {
  uint64_t b = (total + 255u) / 256u;
  if (b > 8192u) b = 8192u;
  if (b == 0) b = 1;
  return (uint32_t)b;
}

}  // namespace

void cgs_launch_to_blocked_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t k_in,
                               uint32_t k_out, uint32_t chan0, hipStream_t s) {
  if ((uint64_t)n * k_in == 0) return;
  hipLaunchKernelGGL(k_to_blocked<uint32_t>, dim3(grid_for((uint64_t)n * k_in)), dim3(256), 0, s,
                     in, out, n, k_in, k_out, chan0);
}

void cgs_launch_to_blocked_u16(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t k_in,
                               uint32_t k_out, uint32_t chan0, hipStream_t s) {
  if ((uint64_t)n * k_in == 0) return;
  hipLaunchKernelGGL(k_to_blocked<uint16_t>, dim3(grid_for((uint64_t)n * k_in)), dim3(256), 0, s,
                     in, out, n, k_in, k_out, chan0);
}

void cgs_launch_sh_f16_blocked(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t stride,
                               uint32_t rest, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_sh_f16_blocked, dim3(grid_for((uint64_t)n * 48u)), dim3(256), 0, s, in,
                     out, n, stride, rest);
}

void cgs_launch_copy_words(const uint32_t* src, uint32_t* dst, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy_words, dim3(1), dim3(64), 0, s, src, dst, n);
}
