// cgs_ingest_math.h — per-splat math of scene ingest and quantised storage.
//
// Shared by the HIP kernels (k_ingest.hip) and the CPU oracle, so that both
// produce the same bits (-ffp-contract=off on both sides):
//   * PLY scalars to f32 (PlyScalarType::read_f32, src/scene/ply_loader.rs:
//     80-111): floats and doubles in either byte order, integers widened;
//   * the RawPlyVertex record (src/scene/raw_vertex.rs:16-37) as 62 floats:
//     position 0-2, normal 3-5, f_dc 6-8, f_rest 9-53 (15 per channel),
//     opacity 54, log-scale 55-57, quaternion w x y z 58-61;
//   * its activation (GaussianSceneSoa::push_raw, soa.rs:113-126):
//     activate_scale / activate_rotation / activate_opacity and
//     compute_covariance_3d (math/covariance.rs:13-94) in glam's scalar order;
//   * decompose_covariance_3d (covariance.rs:265-281) and its quaternion,
//     encode_octahedral_rotation (soa.rs:363-402), and the inverse
//     activations of GaussianSceneSoa::to_raw_vertices (soa.rs:306-328).
// glam's Quat is a 4-wide SSE vector on x86-64: its length is
// sqrt((x^2 + z^2) + (y^2 + w^2)), the order used here.
#pragma once

#include <stdint.h>

#include "cgs_detmath.h"
#include "cgs_scene_math.h"

// PLY scalar types (PlyScalarType, ply_loader.rs:34-45); 0 = property absent
#define CGS_PT_NONE 0u
#define CGS_PT_I8 1u
#define CGS_PT_U8 2u
#define CGS_PT_I16 3u
#define CGS_PT_U16 4u
#define CGS_PT_I32 5u
#define CGS_PT_U32 6u
#define CGS_PT_F32 7u
#define CGS_PT_F64 8u

#define CGS_RAW_FLOATS 62u

CGS_HD uint32_t cgs_pt_size(uint32_t t) {
  return t == CGS_PT_I8 || t == CGS_PT_U8 ? 1u
       : t == CGS_PT_I16 || t == CGS_PT_U16 ? 2u
       : t == CGS_PT_F64 ? 8u
       : t == CGS_PT_NONE ? 0u : 4u;
}

// one scalar at p (any alignment), widened to f32 as `as f32` does
CGS_HD float cgs_pt_read(const uint8_t* p, uint32_t t, bool le) {
  const uint32_t n = cgs_pt_size(t);
  uint64_t v = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t b = p[i];
    v |= (uint64_t)b << (8u * (le ? i : n - 1u - i));
  }
  switch (t) {
    case CGS_PT_I8: return (float)(int8_t)(uint8_t)v;
    case CGS_PT_U8: return (float)(uint8_t)v;
    case CGS_PT_I16: return (float)(int16_t)(uint16_t)v;
    case CGS_PT_U16: return (float)(uint16_t)v;
    case CGS_PT_I32: return (float)(int32_t)(uint32_t)v;
    case CGS_PT_U32: return (float)(uint32_t)v;
    case CGS_PT_F32: return cgs_u2f((uint32_t)v);
    case CGS_PT_F64: return (float)__builtin_bit_cast(double, v);
    default: return 0.0f;
  }
}

// activate_rotation (covariance.rs:36-49): quaternion (x, y, z, w) from the
// PLY's (w, x, y, z), normalised; identity below norm 1e-6
CGS_HD void cgs_activate_rotation(const float* wxyz, float* q) {
  const float w = wxyz[0], x = wxyz[1], y = wxyz[2], z = wxyz[3];
  const float norm = cgs_sqrtf(x * x + y * y + z * z + w * w);
  if (norm < 1e-6f) {
    q[0] = 0.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 1.0f;
    return;
  }
  const float inv = 1.0f / norm;
  q[0] = x * inv; q[1] = y * inv; q[2] = z * inv; q[3] = w * inv;
}

// Mat3::from_quat (column-major); q = (x, y, z, w)
CGS_HD void cgs_m3_from_quat(const float* q, float* m) {
  const float x2 = q[0] + q[0], y2 = q[1] + q[1], z2 = q[2] + q[2];
  const float xx = q[0] * x2, xy = q[0] * y2, xz = q[0] * z2;
  const float yy = q[1] * y2, yz = q[1] * z2, zz = q[2] * z2;
  const float wx = q[3] * x2, wy = q[3] * y2, wz = q[3] * z2;
  m[0] = 1.0f - (yy + zz); m[1] = xy + wz;          m[2] = xz - wy;
  m[3] = xy - wz;          m[4] = 1.0f - (xx + zz); m[5] = yz + wx;
  m[6] = xz + wy;          m[7] = yz - wx;          m[8] = 1.0f - (xx + yy);
}

// compute_covariance_3d (covariance.rs:58-94): M = R(q) diag(s), Sigma = M M^T
CGS_HD void cgs_covariance_from_scale_rot(const float* s, const float* q, float* cov) {
  float R[9], D[9] = {s[0], 0.0f, 0.0f, 0.0f, s[1], 0.0f, 0.0f, 0.0f, s[2]}, M[9], Mt[9], S[9];
  cgs_m3_from_quat(q, R);
  cgs_m3_mul(R, D, M);
  cgs_m3_transpose(M, Mt);
  cgs_m3_mul(M, Mt, S);
  cgs_m3_to_cov(S, cov);
}

// GaussianSceneSoa::push_raw of one decoded record: position [3], cov [6],
// opacity, SH [48] (f_dc then f_rest)
CGS_HD void cgs_activate_record(const float* raw, float* pos, float* cov, float* opa, float* sh) {
  pos[0] = raw[0]; pos[1] = raw[1]; pos[2] = raw[2];
  float s[3], q[4];
  for (int i = 0; i < 3; ++i) s[i] = cgs_fmax(cgs_expf(raw[55 + i]), 1e-4f);
  cgs_activate_rotation(raw + 58, q);
  cgs_covariance_from_scale_rot(s, q, cov);
  *opa = 1.0f / (1.0f + cgs_expf(-raw[54]));
  for (int i = 0; i < 48; ++i) sh[i] = raw[6 + i];
}

// ---- quaternions (glam Quat, SSE dot order) ----

CGS_HD float cgs_quat_dot(const float* a, const float* b) {
  return (a[0] * b[0] + a[2] * b[2]) + (a[1] * b[1] + a[3] * b[3]);
}

// Quat::normalize: q * (1 / |q|)
CGS_HD void cgs_quat_normalize(const float* q, float* o) {
  const float r = 1.0f / cgs_sqrtf(cgs_quat_dot(q, q));
  for (int k = 0; k < 4; ++k) o[k] = q[k] * r;
}

// Quat::from_mat3 (glam's from_rotation_axes, after DirectXMath's
// XMQuaternionRotationMatrix): the branch on the largest of x^2, y^2, z^2, w^2
CGS_HD void cgs_quat_from_m3(const float* m, float* q) {
  const float m00 = m[0], m01 = m[1], m02 = m[2];
  const float m10 = m[3], m11 = m[4], m12 = m[5];
  const float m20 = m[6], m21 = m[7], m22 = m[8];
  if (m22 <= 0.0f) {
    const float dif10 = m11 - m00, omm22 = 1.0f - m22;
    if (dif10 <= 0.0f) {
      const float f = omm22 - dif10, inv = 0.5f / cgs_sqrtf(f);
      q[0] = f * inv; q[1] = (m01 + m10) * inv; q[2] = (m02 + m20) * inv; q[3] = (m12 - m21) * inv;
    } else {
      const float f = omm22 + dif10, inv = 0.5f / cgs_sqrtf(f);
      q[0] = (m01 + m10) * inv; q[1] = f * inv; q[2] = (m12 + m21) * inv; q[3] = (m20 - m02) * inv;
    }
  } else {
    const float sum10 = m11 + m00, opm22 = 1.0f + m22;
    if (sum10 <= 0.0f) {
      const float f = opm22 - sum10, inv = 0.5f / cgs_sqrtf(f);
      q[0] = (m02 + m20) * inv; q[1] = (m12 + m21) * inv; q[2] = f * inv; q[3] = (m01 - m10) * inv;
    } else {
      const float f = opm22 + sum10, inv = 0.5f / cgs_sqrtf(f);
      q[0] = (m12 - m21) * inv; q[1] = (m20 - m02) * inv; q[2] = (m01 - m10) * inv; q[3] = f * inv;
    }
  }
}

// decompose_covariance_3d (covariance.rs:265-281): principal scales and the
// rotation of the right-handed eigenbasis
CGS_HD void cgs_decompose_covariance(const float* cov, float* scale, float* q) {
  float ev[3], basis[9], qr[4];
  cgs_sym_eigen(cov, ev, basis);
  const float min_var = 1e-4f * 1e-4f;
  for (int i = 0; i < 3; ++i) scale[i] = cgs_sqrtf(cgs_fmax(ev[i], min_var));
  cgs_quat_from_m3(basis, qr);
  cgs_quat_normalize(qr, q);
}

CGS_HD float cgs_sign_not_zero(float v) { return v >= 0.0f ? 1.0f : -1.0f; }

// encode_octahedral_rotation (soa.rs:369-402): 11-bit x, 11-bit y of the
// octahedral axis, 10-bit angle in [0, pi].  Quat::to_axis_angle is
// angle = 2 atan2(|v|, w), axis = v / |v| (identity below |v| = 1e-8).
CGS_HD uint32_t cgs_encode_octahedral(const float* rot) {
  float q[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  if (cgs_quat_dot(rot, rot) > 1e-12f) cgs_quat_normalize(rot, q);
  if (q[3] < 0.0f)
    for (int k = 0; k < 4; ++k) q[k] = -q[k];
  float axis[3] = {1.0f, 0.0f, 0.0f}, angle = 0.0f;
  const float len = cgs_sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  if (len >= 1e-8f) {
    angle = 2.0f * cgs_atan2f(len, q[3]);
    for (int k = 0; k < 3; ++k) axis[k] = q[k] / len;
  }
  float d[3] = {0.0f, 0.0f, 1.0f};
  const float al2 = axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2];
  if (!(angle <= 1e-6f || al2 <= 1e-12f)) {
    const float r = 1.0f / cgs_sqrtf(al2);
    for (int k = 0; k < 3; ++k) d[k] = axis[k] * r;
  }
  const float l1 = cgs_fmax(fabsf(d[0]) + fabsf(d[1]) + fabsf(d[2]), 1e-8f);
  float ox = d[0] / l1, oy = d[1] / l1;
  if (d[2] < 0.0f) {
    const float fx = (1.0f - fabsf(oy)) * cgs_sign_not_zero(ox);
    const float fy = (1.0f - fabsf(ox)) * cgs_sign_not_zero(oy);
    ox = fx;
    oy = fy;
  }
  const uint32_t ix = (uint32_t)roundf(cgs_fclamp(ox * 0.5f + 0.5f, 0.0f, 1.0f) * 2047.0f);
  const uint32_t iy = (uint32_t)roundf(cgs_fclamp(oy * 0.5f + 0.5f, 0.0f, 1.0f) * 2047.0f);
  const uint32_t ia = (uint32_t)roundf(cgs_fclamp(angle / CGS_PI_F, 0.0f, 1.0f) * 1023.0f);
  return (ix & 0x7FFu) | ((iy & 0x7FFu) << 11) | ((ia & 0x3FFu) << 22);
}

// ---- QuantizedGaussianSceneSoa::from_soa_with_degree (soa.rs:454-500) ----

CGS_HD uint32_t cgs_sh_bases(uint32_t degree) {
  return degree == 0u ? 1u : degree == 1u ? 4u : degree == 2u ? 9u : 16u;
}
CGS_HD uint32_t cgs_sh_words(uint32_t degree) { return (cgs_sh_bases(degree) * 3u + 1u) / 2u; }

CGS_HD uint32_t cgs_pack_f16_pair(float a, float b) {
  return (uint32_t)cgs_f32_to_f16(a) | ((uint32_t)cgs_f32_to_f16(b) << 16);
}

// one splat at SH degree DEG: cov_f16 [3], the octahedral rotation,
// sh_f16 [cgs_sh_words(DEG)] (indexes resolved at compile time)
template <uint32_t DEG>
CGS_HD void cgs_quantize_record_t(const float* cov, const float* sh, uint32_t* cov_f16,
                                  uint32_t* rot, uint32_t* sh_f16) {
  cov_f16[0] = cgs_pack_f16_pair(cov[0], cov[1]);
  cov_f16[1] = cgs_pack_f16_pair(cov[2], cov[3]);
  cov_f16[2] = cgs_pack_f16_pair(cov[4], cov[5]);
  float scale[3], q[4];
  cgs_decompose_covariance(cov, scale, q);
  *rot = cgs_encode_octahedral(q);
  // DC, then the first (bases - 1) rest coefficients of each channel, zero padded
  constexpr uint32_t bases = DEG == 0u ? 1u : DEG == 1u ? 4u : DEG == 2u ? 9u : 16u;
  constexpr uint32_t rest = bases - 1u, words = (bases * 3u + 1u) / 2u;
#pragma unroll
  for (uint32_t w = 0; w < words; ++w) {
    float pair[2];
#pragma unroll
    for (uint32_t h = 0; h < 2u; ++h) {
      const uint32_t i = 2u * w + h;
      float v = 0.0f;
      if (i < 3u) {
        v = sh[i];
      } else if (rest > 0u && i < 3u + 3u * rest) {
        const uint32_t ch = (i - 3u) / (rest > 0u ? rest : 1u);
        const uint32_t k = (i - 3u) % (rest > 0u ? rest : 1u);
        v = sh[3u + ch * 15u + k];
      }
      pair[h] = v;
    }
    sh_f16[w] = cgs_pack_f16_pair(pair[0], pair[1]);
  }
}

CGS_HD void cgs_quantize_record(const float* cov, const float* sh, uint32_t degree,
                                uint32_t* cov_f16, uint32_t* rot, uint32_t* sh_f16) {
  switch (degree) {
    case 0u: cgs_quantize_record_t<0u>(cov, sh, cov_f16, rot, sh_f16); break;
    case 1u: cgs_quantize_record_t<1u>(cov, sh, cov_f16, rot, sh_f16); break;
    case 2u: cgs_quantize_record_t<2u>(cov, sh, cov_f16, rot, sh_f16); break;
    default: cgs_quantize_record_t<3u>(cov, sh, cov_f16, rot, sh_f16); break;
  }
}

// ---- GaussianSceneSoa::to_raw_vertices (soa.rs:306-328): one record ----

CGS_HD void cgs_export_record(const float* pos, const float* cov, float opa, const float* sh,
                              float* raw) {
  for (uint32_t k = 0; k < CGS_RAW_FLOATS; ++k) raw[k] = 0.0f;
  raw[0] = pos[0]; raw[1] = pos[1]; raw[2] = pos[2];
  // inverse_activate_opacity (covariance.rs:161-165)
  const float c = cgs_fclamp(opa, 1e-6f, 1.0f - 1e-6f);
  raw[54] = cgs_logf(c / (1.0f - c));
  float scale[3], q[4];
  cgs_decompose_covariance(cov, scale, q);
  // inverse_activate_scale (168-175)
  for (int i = 0; i < 3; ++i) raw[55 + i] = cgs_logf(cgs_fmax(scale[i], 1e-4f));
  raw[58] = q[3]; raw[59] = q[0]; raw[60] = q[1]; raw[61] = q[2];
  for (int i = 0; i < 48; ++i) raw[6 + i] = sh[i];
}
