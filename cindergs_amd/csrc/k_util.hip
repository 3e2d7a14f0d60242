// k_util.hip — layout conversion on upload: the caller's per-splat records
// (GaussianSceneSoa / QuantizedGaussianSceneSoa, src/scene/soa.rs:30-45,
// 429-500) become planar device arrays (one plane per component) so that the
// per-splat kernels read one contiguous 256-byte span per wave instruction.
#include "cgs_internal.h"

namespace {

// out[c * out_stride + i] = in[i * k + c]
__global__ __launch_bounds__(256) void k_deinterleave_u32(const uint32_t* in, uint32_t* out,
                                                          uint32_t n, uint32_t k,
                                                          uint32_t out_stride) {
  const uint64_t total = (uint64_t)n * k;
  for (uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x; j < total;
       j += (uint64_t)gridDim.x * 256u) {
    const uint32_t i = (uint32_t)(j / k), c = (uint32_t)(j % k);
    out[(size_t)c * out_stride + i] = in[j];
  }
}

__global__ __launch_bounds__(256) void k_deinterleave_u16(const uint16_t* in, uint16_t* out,
                                                          uint32_t n, uint32_t k,
                                                          uint32_t out_stride) {
  const uint64_t total = (uint64_t)n * k;
  for (uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x; j < total;
       j += (uint64_t)gridDim.x * 256u) {
    const uint32_t i = (uint32_t)(j / k), c = (uint32_t)(j % k);
    out[(size_t)c * out_stride + i] = in[j];
  }
}

// f16 SH record (channel-major, degree-trimmed, `stride` halves per splat) ->
// 48 planes in the f32 record order [dc r,g,b, rest_r 15, rest_g 15, rest_b 15].
__global__ __launch_bounds__(256) void k_sh_f16_planes(const uint16_t* in, uint16_t* out,
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
    out[(size_t)k * n + i] = src >= 0 ? in[(size_t)i * stride + (uint32_t)src] : (uint16_t)0;
  }
}

uint32_t grid_for(uint64_t total) {
  uint64_t b = (total + 255u) / 256u;
  if (b > 8192u) b = 8192u;
  if (b == 0) b = 1;
  return (uint32_t)b;
}

}  // namespace

void cgs_launch_deinterleave_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t k,
                                 uint32_t out_stride, hipStream_t s) {
  if ((uint64_t)n * k == 0) return;
  hipLaunchKernelGGL(k_deinterleave_u32, dim3(grid_for((uint64_t)n * k)), dim3(256), 0, s, in,
                     out, n, k, out_stride);
}

void cgs_launch_deinterleave_u16(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t k,
                                 uint32_t out_stride, hipStream_t s) {
  if ((uint64_t)n * k == 0) return;
  hipLaunchKernelGGL(k_deinterleave_u16, dim3(grid_for((uint64_t)n * k)), dim3(256), 0, s, in,
                     out, n, k, out_stride);
}

void cgs_launch_sh_f16_planes(const uint16_t* in, uint16_t* out, uint32_t n, uint32_t stride,
                              uint32_t rest, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_sh_f16_planes, dim3(grid_for((uint64_t)n * 48u)), dim3(256), 0, s, in,
                     out, n, stride, rest);
}
