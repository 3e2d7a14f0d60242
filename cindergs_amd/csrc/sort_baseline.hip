// sort_baseline.hip — the vendor-library baseline for the Stage 3 sort.
//
// Not part of the C ABI of include/cindergs.h and not loaded by the render
// path: a separate library (libcgs_sort_baseline.so) that bench.py loads to
// time hipCUB's DeviceRadixSort::SortPairs (rocPRIM's onesweep LSD radix
// sort) on as many (u32 key, u32 value) pairs, over as many key bits, as a
// frame's Stage 3 sorts, beside the library's own sort of that frame.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

namespace {

// splitmix64 of (seed, i): the pair's key (low `bits` bits) and value i
__global__ void k_fill_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t bits,
                             uint64_t seed) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1u);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const uint32_t mask = bits >= 32u ? 0xFFFFFFFFu : ((1u << bits) - 1u);
  keys[i] = (uint32_t)z & mask;
  vals[i] = i;
}

}  // namespace

// Mean device time (HIP events over `reps` sorts after 3 untimed ones) of
// hipcub::DeviceRadixSort::SortPairs on n pairs, bits key bits; 0 on success,
// -1 on a HIP error.  Allocates and frees its own buffers.
extern "C" __attribute__((visibility("default"))) int cgs_sortbase_rocprim(
    uint32_t n, uint32_t bits, uint32_t reps, uint64_t seed, double* ms_per_sort) {
  if (n == 0 || bits == 0 || bits > 32 || reps == 0 || !ms_per_sort) return -1;
  uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
  void* tmp = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  size_t tb = 0;
  float ms = 0.0f;
  bool ok = hipMalloc(&k0, (size_t)n * 4) == hipSuccess &&
            hipMalloc(&v0, (size_t)n * 4) == hipSuccess &&
            hipMalloc(&k1, (size_t)n * 4) == hipSuccess &&
            hipMalloc(&v1, (size_t)n * 4) == hipSuccess &&
            hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_fill_pairs, dim3((n + 255u) / 256u), dim3(256), 0, s, k0, v0, n, bits,
                       seed);
    ok = hipGetLastError() == hipSuccess &&
         hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, (int)n, 0, (int)bits,
                                            s) == hipSuccess &&
         hipMalloc(&tmp, tb) == hipSuccess;
  }
  for (int r = 0; ok && r < 3; ++r)
    ok = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, (int)n, 0, (int)bits, s) ==
         hipSuccess;
  ok = ok && hipEventRecord(a, s) == hipSuccess;
  for (uint32_t r = 0; ok && r < reps; ++r)
    ok = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, (int)n, 0, (int)bits, s) ==
         hipSuccess;
  ok = ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
       hipEventElapsedTime(&ms, a, b) == hipSuccess;
  if (ok) *ms_per_sort = (double)ms / reps;
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (s) (void)hipStreamDestroy(s);
  for (void* p : {(void*)k0, (void*)v0, (void*)k1, (void*)v1, tmp})
    if (p) (void)hipFree(p);
  return ok ? 0 : -1;
}
