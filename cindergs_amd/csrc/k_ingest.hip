// k_ingest.hip — scene ingest and quantised storage on the GPU.
//
//  * k_activate_raw: PlyLoader::load_from_bytes + GaussianSceneSoa::
//    from_raw_vertices_par (src/scene/ply_loader.rs:622-654, soa.rs:137-189)
//    for any binary vertex layout.  A block stages the bytes of up to 256
//    consecutive records in LDS with coalesced 4-byte loads (the payload is
//    read from HBM once, whatever its layout), then each thread decodes its
//    record through the slot table (offset + scalar type of each of the 62
//    RawPlyVertex floats) and activates it (cgs_activate_record).  When every
//    present slot is an aligned little-endian float32 (the canonical 248-byte
//    3DGS record, or a reordered all-float one) the decode is one LDS word
//    read per slot; otherwise it assembles the bytes.
//  * k_quantize: QuantizedGaussianSceneSoa::from_soa_with_degree (soa.rs:
//    454-500), one thread per splat: f16 pairs, octahedral rotation,
//    degree-trimmed f16 SH.
//  * k_export_raw: GaussianSceneSoa::to_raw_vertices (soa.rs:306-328), one
//    thread per splat, 248-byte records written as 31 float2 stores.
#include "cgs_internal.h"
#include "cgs_ingest_math.h"

namespace {

template <bool WORDS>
__global__ __launch_bounds__(256) void k_activate_raw(CgsActivateArgs A) {
  extern __shared__ uint32_t s_words[];
  const uint64_t r0 = (uint64_t)blockIdx.x * A.per_block;
  if (r0 >= A.n) return;
  const uint32_t cnt = (uint32_t)min((uint64_t)A.per_block, A.n - r0);
  const uint64_t total = A.n * A.stride;
  const uint64_t b0 = r0 * A.stride, b1 = b0 + (uint64_t)cnt * A.stride;
  const uint64_t w0 = b0 / 4u, w1 = (b1 + 3u) / 4u;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(A.records);
  for (uint64_t w = w0 + threadIdx.x; w < w1; w += 256u) {
    uint32_t v;
    if (4u * w + 4u <= total) {
      v = __builtin_nontemporal_load(src + w);
    } else {  // the payload's last partial word: bytes inside it only
      v = 0u;
      for (uint32_t k = 0; k < 4u; ++k)
        if (4u * w + k < total) v |= (uint32_t)A.records[4u * w + k] << (8u * k);
    }
    s_words[w - w0] = v;
  }
  __syncthreads();
  const uint32_t t = threadIdx.x;
  if (t >= cnt) return;
  const uint32_t rec_byte = (uint32_t)(b0 - 4u * w0) + t * A.stride;
  float raw[CGS_RAW_FLOATS];
  if (WORDS) {
    const uint32_t rw = rec_byte / 4u;
#pragma unroll
    for (uint32_t k = 0; k < CGS_RAW_FLOATS; ++k)
      raw[k] = A.slot_type[k] ? cgs_u2f(s_words[rw + A.slot_offset[k] / 4u]) : 0.0f;
  } else {
    const uint8_t* rec = reinterpret_cast<const uint8_t*>(s_words) + rec_byte;
    for (uint32_t k = 0; k < CGS_RAW_FLOATS; ++k)
      raw[k] = A.slot_type[k] ? cgs_pt_read(rec + A.slot_offset[k], A.slot_type[k], A.le != 0u)
                              : 0.0f;
  }
  float pos[3], cov[6], opa, sh[48];
  cgs_activate_record(raw, pos, cov, &opa, sh);
  const uint64_t i = r0 + t;
  for (int k = 0; k < 3; ++k) A.pos[i * 3 + k] = pos[k];
  float2* c2 = reinterpret_cast<float2*>(A.cov + i * 6);
  for (int k = 0; k < 3; ++k) c2[k] = make_float2(cov[2 * k], cov[2 * k + 1]);
  A.opa[i] = opa;
  float4* s4 = reinterpret_cast<float4*>(A.sh + i * 48);
  for (int k = 0; k < 12; ++k)
    s4[k] = make_float4(sh[4 * k], sh[4 * k + 1], sh[4 * k + 2], sh[4 * k + 3]);
}

__global__ __launch_bounds__(256) void k_quantize(const float* cov, const float* sh, uint64_t n,
                                                  uint32_t degree, uint32_t* cov_f16,
                                                  uint32_t* rot, uint32_t* sh_f16) {
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  float c[6], s[48];
  for (int k = 0; k < 6; ++k) c[k] = cov[i * 6 + k];
  const float4* s4 = reinterpret_cast<const float4*>(sh + i * 48);
  for (int k = 0; k < 12; ++k) {
    const float4 v = s4[k];
    s[4 * k] = v.x; s[4 * k + 1] = v.y; s[4 * k + 2] = v.z; s[4 * k + 3] = v.w;
  }
  uint32_t cw[3], r, sw[24];
  cgs_quantize_record(c, s, degree, cw, &r, sw);
  for (int k = 0; k < 3; ++k) cov_f16[i * 3 + k] = cw[k];
  rot[i] = r;
  const uint32_t words = cgs_sh_words(degree);
  for (uint32_t k = 0; k < words; ++k) sh_f16[i * words + k] = sw[k];
}

__global__ __launch_bounds__(256) void k_export_raw(const float* pos, const float* cov,
                                                    const float* opa, const float* sh, uint64_t n,
                                                    float* records) {
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  float p[3], c[6], s[48], raw[CGS_RAW_FLOATS];
  for (int k = 0; k < 3; ++k) p[k] = pos[i * 3 + k];
  for (int k = 0; k < 6; ++k) c[k] = cov[i * 6 + k];
  const float4* s4 = reinterpret_cast<const float4*>(sh + i * 48);
  for (int k = 0; k < 12; ++k) {
    const float4 v = s4[k];
    s[4 * k] = v.x; s[4 * k + 1] = v.y; s[4 * k + 2] = v.z; s[4 * k + 3] = v.w;
  }
  cgs_export_record(p, c, opa[i], s, raw);
  float2* out = reinterpret_cast<float2*>(records + i * CGS_RAW_FLOATS);
  for (int k = 0; k < 31; ++k) out[k] = make_float2(raw[2 * k], raw[2 * k + 1]);
}

}  // namespace

void cgs_launch_activate_raw(const CgsActivateArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  const uint64_t blocks = (a.n + a.per_block - 1u) / a.per_block;
  const size_t lds = (((size_t)a.per_block * a.stride + 3u) / 4u + 1u) * 4u;
  if (a.word_path)
    hipLaunchKernelGGL(k_activate_raw<true>, dim3((uint32_t)blocks), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL(k_activate_raw<false>, dim3((uint32_t)blocks), dim3(256), lds, s, a);
}

void cgs_launch_quantize(const float* cov, const float* sh, uint64_t n, uint32_t degree,
                         uint32_t* cov_f16, uint32_t* rot, uint32_t* sh_f16, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_quantize, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, cov, sh,
                     n, degree, cov_f16, rot, sh_f16);
}

void cgs_launch_export_raw(const float* pos, const float* cov, const float* opa,
                           const float* sh, uint64_t n, float* records, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_export_raw, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, pos,
                     cov, opa, sh, n, records);
}
