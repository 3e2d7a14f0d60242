// k_ingest.hip — scene ingest and quantised storage on the GPU.
//
//  * k_activate_raw: PlyLoader::load_from_bytes + GaussianSceneSoa::
//    from_raw_vertices_par (src/scene/ply_loader.rs:622-654, soa.rs:137-189)
//    for any binary vertex layout.  A 128-thread block stages the bytes of up
//    to 128 consecutive records in LDS with coalesced 16-byte loads (the
//    payload is read from HBM once, whatever its layout), each thread decodes
//    its record through the slot table (offset + scalar type of each of the
//    62 RawPlyVertex floats) and activates it (cgs_activate_record), and the
//    block's outputs go back through the same LDS so that every global store
//    is a coalesced 16-byte vector.  When every present slot is an aligned
//    little-endian float32 (the canonical 248-byte 3DGS record, or a
//    reordered all-float one) the decode is one LDS word read per slot;
//    otherwise it assembles the bytes.
//  * k_quantize: QuantizedGaussianSceneSoa::from_soa_with_degree (soa.rs:
//    454-500), one thread per splat: f16 pairs, octahedral rotation,
//    degree-trimmed f16 SH.
//  * k_export_raw: GaussianSceneSoa::to_raw_vertices (soa.rs:306-328), one
//    thread per splat, 248-byte records.
//  Both are bound by the per-splat eigen solve in double, not by HBM.
#include "cgs_internal.h"
#include "cgs_ingest_math.h"

namespace {

constexpr uint32_t AT = 128;  // threads (and at most records) per activation block
constexpr uint32_t OUT_WORDS = 60;  // pos 3 + cov 7 + opacity 1 + SH 49 per record in LDS

// copies cnt records of W words, held in LDS at an odd stride S (no bank
// conflicts when thread t writes record t), to W*cnt consecutive words in
// global memory: 16-byte stores when W % 4 == 0, 8-byte when W % 2 == 0 (and
// dst is aligned for them), else words
template <uint32_t W, uint32_t S, typename T>
__device__ __forceinline__ void lds_to_global(const uint32_t* lds, T* dst, uint32_t cnt,
                                              uint32_t t) {
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  const uintptr_t a = reinterpret_cast<uintptr_t>(d);
  if (W % 4u == 0u && (a & 15u) == 0u) {
    const uint32_t q = cnt * (W / 4u);
    for (uint32_t k = t; k < q; k += AT) {
      const uint32_t r = k / (W / 4u), c = 4u * (k % (W / 4u));
      const uint32_t* p = lds + r * S + c;
      reinterpret_cast<uint4*>(d)[k] = make_uint4(p[0], p[1], p[2], p[3]);
    }
  } else if (W % 2u == 0u && (a & 7u) == 0u) {
    const uint32_t q = cnt * (W / 2u);
    for (uint32_t k = t; k < q; k += AT) {
      const uint32_t r = k / (W / 2u), c = 2u * (k % (W / 2u));
      const uint32_t* p = lds + r * S + c;
      reinterpret_cast<uint2*>(d)[k] = make_uint2(p[0], p[1]);
    }
  } else {
    for (uint32_t k = t; k < W * cnt; k += AT) d[k] = lds[(k / W) * S + (k % W)];
  }
}

template <bool WORDS>
__global__ __launch_bounds__(AT) void k_activate_raw(CgsActivateArgs A) {
  extern __shared__ uint32_t s_words[];
  const uint64_t r0 = (uint64_t)blockIdx.x * A.per_block;
  if (r0 >= A.n) return;
  const uint32_t cnt = (uint32_t)min((uint64_t)A.per_block, A.n - r0);
  const uint64_t total = A.n * A.stride;
  const uint64_t b0 = r0 * A.stride, b1 = b0 + (uint64_t)cnt * A.stride;
  const uint64_t w0 = b0 / 4u, w1 = (b1 + 3u) / 4u;
  const uint32_t t = threadIdx.x;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(A.records);
  // stage the records: 16-byte loads over the aligned middle, words at the ends
  {
    const uint64_t q0 = (w0 + 3u) / 4u, q1 = w1 / 4u;  // whole uint4 of the range
    const bool vec = q1 > q0 && 16u * q1 <= total &&
                     (reinterpret_cast<uintptr_t>(A.records) & 15u) == 0u;
    const uint64_t head_end = vec ? 4u * q0 : w1, tail_start = vec ? 4u * q1 : w1;
    for (uint64_t w = w0 + t; w < head_end; w += AT) {
      uint32_t v;
      if (4u * w + 4u <= total) {
        v = __builtin_nontemporal_load(src + w);
      } else {
        v = 0u;
        for (uint32_t k = 0; k < 4u; ++k)
          if (4u * w + k < total) v |= (uint32_t)A.records[4u * w + k] << (8u * k);
      }
      s_words[w - w0] = v;
    }
    if (vec) {
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      const u32x4* src4 = reinterpret_cast<const u32x4*>(A.records);
      for (uint64_t q = q0 + t; q < q1; q += AT) {
        const u32x4 v = __builtin_nontemporal_load(src4 + q);
        const uint64_t o = 4u * q - w0;
        s_words[o] = v.x; s_words[o + 1u] = v.y; s_words[o + 2u] = v.z; s_words[o + 3u] = v.w;
      }
      for (uint64_t w = tail_start + t; w < w1; w += AT) {
        uint32_t v;
        if (4u * w + 4u <= total) {
          v = __builtin_nontemporal_load(src + w);
        } else {
          v = 0u;
          for (uint32_t k = 0; k < 4u; ++k)
            if (4u * w + k < total) v |= (uint32_t)A.records[4u * w + k] << (8u * k);
        }
        s_words[w - w0] = v;
      }
    }
  }
  __syncthreads();
  float out[58];
  if (t < cnt) {
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
        raw[k] = A.slot_type[k] ? cgs_pt_read(rec + A.slot_offset[k], A.slot_type[k],
                                              A.le != 0u)
                                : 0.0f;
    }
    cgs_activate_record(raw, out, out + 3, out + 9, out + 10);
  }
  __syncthreads();  // the records are consumed: the LDS now holds the outputs
  // planes of the block at odd strides: pos [cnt*3] | cov [cnt*7] | opacity
  // [cnt] | SH [cnt*49]
  uint32_t* o_pos = s_words;
  uint32_t* o_cov = o_pos + 3u * cnt;
  uint32_t* o_opa = o_cov + 7u * cnt;
  uint32_t* o_sh = o_opa + cnt;
  if (t < cnt) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o_pos[3u * t + k] = cgs_f2u(out[k]);
#pragma unroll
    for (int k = 0; k < 6; ++k) o_cov[7u * t + k] = cgs_f2u(out[3 + k]);
    o_opa[t] = cgs_f2u(out[9]);
#pragma unroll
    for (int k = 0; k < 48; ++k) o_sh[49u * t + k] = cgs_f2u(out[10 + k]);
  }
  __syncthreads();
  lds_to_global<3u, 3u>(o_pos, A.pos + r0 * 3u, cnt, t);
  lds_to_global<6u, 7u>(o_cov, A.cov + r0 * 6u, cnt, t);
  lds_to_global<1u, 1u>(o_opa, A.opa + r0, cnt, t);
  lds_to_global<48u, 49u>(o_sh, A.sh + r0 * 48u, cnt, t);
}

// quantize and export: one thread per splat, direct loads and stores.  Their
// cost is the per-splat eigen solve in double (decompose_covariance_3d), so
// occupancy matters more than coalescing: staging them through LDS like the
// activation measured 2.2x (quantize) and 1.4x (export) slower.
template <uint32_t DEG>
__global__ __launch_bounds__(256) void k_quantize(const float* cov, const float* sh, uint64_t n,
                                                  uint32_t* cov_f16, uint32_t* rot,
                                                  uint32_t* sh_f16) {
  constexpr uint32_t words = DEG == 0u ? 2u : DEG == 1u ? 6u : DEG == 2u ? 14u : 24u;
  const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  float c[6], s[48];
  const float2* c2 = reinterpret_cast<const float2*>(cov + i * 6);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float2 v = c2[k];
    c[2 * k] = v.x; c[2 * k + 1] = v.y;
  }
  const float4* s4 = reinterpret_cast<const float4*>(sh + i * 48);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const float4 v = s4[k];
    s[4 * k] = v.x; s[4 * k + 1] = v.y; s[4 * k + 2] = v.z; s[4 * k + 3] = v.w;
  }
  uint32_t cw[3], r, sw[words];
  cgs_quantize_record_t<DEG>(c, s, cw, &r, sw);
#pragma unroll
  for (int k = 0; k < 3; ++k) cov_f16[i * 3 + k] = cw[k];
  rot[i] = r;
#pragma unroll
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

// ShRotator::new for every tet (sh_rotation.rs:404-425): the f32 rotation
// matrices of bands 1-3 of the tet's polar rotation, 9 + 25 + 49 floats per
// tet, built once and shared by every splat bound to it
// ShRotator::new per tet (sh_rotation.rs:412-430): thread (t, band l =
// blockIdx.y + 1) writes band l's matrix of tet t, with the generated
// straight-line form of cgs_sh_band_matrix (same operations, same values)
__global__ __launch_bounds__(64) void k_tet_sh_matrices(const float* rot, uint32_t num_tets,
                                                        float* mats) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  if (t >= num_tets) return;
  float R[9];
  for (int k = 0; k < 9; ++k) R[k] = rot[(size_t)t * 9 + k];
  float* m = mats + (size_t)t * CGS_SH_MATS;
  if (blockIdx.y == 0) {
    cgs_sh_band1_matrix(R, m);
    for (uint32_t k = 9; k < CGS_SH_M2; ++k) m[k] = 0.0f;  // the pads
  } else if (blockIdx.y == 1) {
    cgs_sh_band2_matrix(R, m + CGS_SH_M2);
    for (uint32_t k = CGS_SH_M2 + 25u; k < CGS_SH_M3; ++k) m[k] = 0.0f;
  } else {
    cgs_sh_band3_matrix(R, m + CGS_SH_M3);
    for (uint32_t k = CGS_SH_M3 + 49u; k < CGS_SH_MATS; ++k) m[k] = 0.0f;
  }
}

// one record of the deformed scene for export: the Stage 1 geometry, and the
// SH of its source splat (a split child's parent) rotated by ShRotator::apply
// (sh_rotation.rs:427-446) of its tet; unbound records keep their SH.  The
// outputs go through LDS at odd strides so that the block writes its rows of
// every output array with coalesced stores.
__global__ __launch_bounds__(AT) void k_export_deformed(CgsExportDeformedArgs A) {
  __shared__ uint32_t o_sh[AT * 49u];
  __shared__ uint32_t o_pos[AT * 3u], o_cov[AT * 7u], o_opa[AT];
  const uint32_t t = threadIdx.x, r0 = blockIdx.x * AT, i = r0 + t;
  const uint32_t cnt = min(AT, A.count - r0);
  if (i < A.count) {
    for (uint32_t k = 0; k < 3u; ++k) o_pos[3u * t + k] = cgs_f2u(A.geo[cgs_blk(i, k, CGS_GEO_K)]);
    for (uint32_t k = 0; k < 6u; ++k)
      o_cov[7u * t + k] = cgs_f2u(A.geo[cgs_blk(i, 3u + k, CGS_GEO_K)]);
    o_opa[t] = cgs_f2u(A.geo[cgs_blk(i, CGS_CH_OPA, CGS_GEO_K)]);
    uint32_t src = i, tet = A.num_tets;
    if (i < A.num_rest) {
      tet = A.bind_tet[i];
    } else {
      src = A.split_meta[2u * (i - A.num_rest)];
      tet = A.split_meta[2u * (i - A.num_rest) + 1u];
    }
    float sh[48];
    for (uint32_t k = 0; k < 48u; ++k) sh[k] = A.sh[cgs_blk(src, k, CGS_SH_K)];
    if (A.degree > 0u && tet < A.num_tets) {
      // band by band (the bands are disjoint blocks, so the per-channel order
      // of ShRotator::apply gives the same values), each matrix fetched with
      // 16-byte loads: the records' tets differ from lane to lane
      const float4* m4 = reinterpret_cast<const float4*>(A.mats + (size_t)tet * CGS_SH_MATS);
      float M[52];
      for (uint32_t q = 0; q < 3u; ++q) reinterpret_cast<float4*>(M)[q] = m4[q];
      for (uint32_t ch = 0; ch < 3u; ++ch) cgs_sh_apply_block(sh, 3u + ch * 15u, 3, M);
      if (A.degree >= 2u) {
        for (uint32_t q = 0; q < 7u; ++q) reinterpret_cast<float4*>(M)[q] = m4[CGS_SH_M2 / 4u + q];
        for (uint32_t ch = 0; ch < 3u; ++ch) cgs_sh_apply_block(sh, 6u + ch * 15u, 5, M);
      }
      if (A.degree >= 3u) {
        for (uint32_t q = 0; q < 13u; ++q) reinterpret_cast<float4*>(M)[q] = m4[CGS_SH_M3 / 4u + q];
        for (uint32_t ch = 0; ch < 3u; ++ch) cgs_sh_apply_block(sh, 11u + ch * 15u, 7, M);
      }
    }
    for (uint32_t k = 0; k < 48u; ++k) o_sh[49u * t + k] = cgs_f2u(sh[k]);
  }
  __syncthreads();
  lds_to_global<3u, 3u>(o_pos, A.pos + (size_t)r0 * 3u, cnt, t);
  lds_to_global<6u, 7u>(o_cov, A.cov + (size_t)r0 * 6u, cnt, t);
  lds_to_global<1u, 1u>(o_opa, A.opa + r0, cnt, t);
  lds_to_global<48u, 49u>(o_sh, A.out_sh + (size_t)r0 * 48u, cnt, t);
}

}  // namespace

void cgs_launch_tet_sh_matrices(const float* rot, uint32_t num_tets, float* mats,
                                hipStream_t s) {
  if (num_tets == 0) return;
  hipLaunchKernelGGL(k_tet_sh_matrices, dim3((num_tets + 63u) / 64u, 3u), dim3(64), 0, s, rot,
                     num_tets, mats);
}

void cgs_launch_export_deformed(const CgsExportDeformedArgs& a, hipStream_t s) {
  if (a.count == 0) return;
  hipLaunchKernelGGL(k_export_deformed, dim3((a.count + AT - 1u) / AT), dim3(AT), 0, s, a);
}

void cgs_launch_activate_raw(const CgsActivateArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  const uint64_t blocks = (a.n + a.per_block - 1u) / a.per_block;
  // LDS: the staged records (plus the partial words at both ends), reused for
  // the outputs of the block
  const size_t in_words = ((size_t)a.per_block * a.stride + 3u) / 4u + 1u;
  const size_t out_words = (size_t)a.per_block * OUT_WORDS;
  const size_t lds = (in_words > out_words ? in_words : out_words) * 4u;
  if (a.word_path)
    hipLaunchKernelGGL(k_activate_raw<true>, dim3((uint32_t)blocks), dim3(AT), lds, s, a);
  else
    hipLaunchKernelGGL(k_activate_raw<false>, dim3((uint32_t)blocks), dim3(AT), lds, s, a);
}

void cgs_launch_quantize(const float* cov, const float* sh, uint64_t n, uint32_t degree,
                         uint32_t* cov_f16, uint32_t* rot, uint32_t* sh_f16, hipStream_t s) {
  if (n == 0) return;
  const dim3 grid((uint32_t)((n + 255u) / 256u));
  if (degree == 0u)
    hipLaunchKernelGGL(k_quantize<0u>, grid, dim3(256), 0, s, cov, sh, n, cov_f16, rot, sh_f16);
  else if (degree == 1u)
    hipLaunchKernelGGL(k_quantize<1u>, grid, dim3(256), 0, s, cov, sh, n, cov_f16, rot, sh_f16);
  else if (degree == 2u)
    hipLaunchKernelGGL(k_quantize<2u>, grid, dim3(256), 0, s, cov, sh, n, cov_f16, rot, sh_f16);
  else
    hipLaunchKernelGGL(k_quantize<3u>, grid, dim3(256), 0, s, cov, sh, n, cov_f16, rot, sh_f16);
}

void cgs_launch_export_raw(const float* pos, const float* cov, const float* opa,
                           const float* sh, uint64_t n, float* records, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_export_raw, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, pos,
                     cov, opa, sh, n, records);
}
