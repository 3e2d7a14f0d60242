// k_deform.hip — Stage 1: per-tet deformation gradient, Newton-Schulz polar
// rotation and largest principal stretch; per-splat covariance transport,
// opacity density law, 3D mip filter and strain splitting.
//
// Follows compute_tet_transforms_kernel (src/render/kernels/deformation.rs:
// 682-865) and deform_and_split_compact_kernel (872-1211) operation for
// operation (-ffp-contract=off; ln/exp/pow/acos/cos from cgs_detmath.h), with
// one deliberate change: the reference hands out child slots with an atomic
// counter (deformation.rs:1191), so child order varies from run to run.  Here
// each 1024-splat partition scans its split flags and chains the partition
// totals with a decoupled look-back, so child k of the frame is always the
// k-th splitting parent in index order — the order of the CPU path
// (src/deformation/cpu_reference.rs:316-377) — and keys stay reproducible.
// Outputs use the wave-blocked geo layout read by Stage 2.
#include "cgs_internal.h"

namespace {

__global__ __launch_bounds__(64) void k_tet_transforms(CgsTetArgs A) {
  const uint32_t idx = blockIdx.x * 64u + threadIdx.x;
  if (idx >= A.num_tets) return;
  const uint32_t* e = A.elems + (size_t)idx * 4;
  const uint32_t v0 = e[0], v1 = e[1], v2 = e[2], v3 = e[3];
  const float* V = A.verts;
  const float x0x = V[v0 * 3], x0y = V[v0 * 3 + 1], x0z = V[v0 * 3 + 2];
  const float x1x = V[v1 * 3], x1y = V[v1 * 3 + 1], x1z = V[v1 * 3 + 2];
  const float x2x = V[v2 * 3], x2y = V[v2 * 3 + 1], x2z = V[v2 * 3 + 2];
  const float x3x = V[v3 * 3], x3y = V[v3 * 3 + 1], x3z = V[v3 * 3 + 2];
  const float ds00 = x1x - x0x, ds10 = x1y - x0y, ds20 = x1z - x0z;
  const float ds01 = x2x - x0x, ds11 = x2y - x0y, ds21 = x2z - x0z;
  const float ds02 = x3x - x0x, ds12 = x3y - x0y, ds22 = x3z - x0z;
  const float* M = A.inv_dm + (size_t)idx * 9;
  const float a00 = M[0], a10 = M[1], a20 = M[2], a01 = M[3], a11 = M[4], a21 = M[5];
  const float a02 = M[6], a12 = M[7], a22 = M[8];
  const float j00 = ds00 * a00 + ds01 * a10 + ds02 * a20;
  const float j10 = ds10 * a00 + ds11 * a10 + ds12 * a20;
  const float j20 = ds20 * a00 + ds21 * a10 + ds22 * a20;
  const float j01 = ds00 * a01 + ds01 * a11 + ds02 * a21;
  const float j11 = ds10 * a01 + ds11 * a11 + ds12 * a21;
  const float j21 = ds20 * a01 + ds21 * a11 + ds22 * a21;
  const float j02 = ds00 * a02 + ds01 * a12 + ds02 * a22;
  const float j12 = ds10 * a02 + ds11 * a12 + ds12 * a22;
  const float j22 = ds20 * a02 + ds21 * a12 + ds22 * a22;
  float* J = A.jac + (size_t)idx * 9;
  J[0] = j00; J[1] = j10; J[2] = j20; J[3] = j01; J[4] = j11; J[5] = j21;
  J[6] = j02; J[7] = j12; J[8] = j22;

  // largest stretch from C = J^T J (trigonometric eigenvalue formula)
  float stretch_v = 0.0f;
  const float c00 = j00 * j00 + j10 * j10 + j20 * j20;
  const float c01 = j00 * j01 + j10 * j11 + j20 * j21;
  const float c02 = j00 * j02 + j10 * j12 + j20 * j22;
  const float c11 = j01 * j01 + j11 * j11 + j21 * j21;
  const float c12 = j01 * j02 + j11 * j12 + j21 * j22;
  const float c22 = j02 * j02 + j12 * j12 + j22 * j22;
  const float p1 = c01 * c01 + c02 * c02 + c12 * c12;
  if (p1 <= 1e-10f) {
    const float dmax = cgs_fmax(c00, cgs_fmax(c11, c22));
    stretch_v = cgs_sqrtf(cgs_fmax(dmax, 0.0f));
  } else {
    const float q = (c00 + c11 + c22) / 3.0f;
    const float d0 = c00 - q, d1 = c11 - q, d2 = c22 - q;
    const float p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0f * p1;
    const float p = cgs_sqrtf(cgs_fmax(p2 / 6.0f, 1e-12f));
    const float ip = 1.0f / p;
    const float b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip;
    const float b01 = c01 * ip, b02 = c02 * ip, b12 = c12 * ip;
    const float det_b = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                        b02 * (b01 * b12 - b11 * b02);
    const float r = cgs_fclamp(0.5f * det_b, -1.0f, 1.0f);
    const float phi = cgs_acosf(r) / 3.0f;
    const float emax = q + 2.0f * p * cgs_cosf(phi);
    stretch_v = cgs_sqrtf(cgs_fmax(emax, 0.0f));
  }

  // Newton-Schulz (Higham) polar rotation, 5 steps, chirality fix
  float* R = A.rot + (size_t)idx * 9;
  const float det_j = j00 * (j11 * j22 - j12 * j21) - j01 * (j10 * j22 - j12 * j20) +
                      j02 * (j10 * j21 - j11 * j20);
  const float adet = fabsf(det_j);
  if (adet < 1e-8f) {
    R[0] = 1.0f; R[1] = 0.0f; R[2] = 0.0f; R[3] = 0.0f; R[4] = 1.0f;
    R[5] = 0.0f; R[6] = 0.0f; R[7] = 0.0f; R[8] = 1.0f;
    return;
  }
  const float gamma = cgs_expf(-cgs_logf(adet) / 3.0f);
  float r00 = j00 * gamma, r10 = j10 * gamma, r20 = j20 * gamma;
  float r01 = j01 * gamma, r11 = j11 * gamma, r21 = j21 * gamma;
  float r02 = j02 * gamma, r12 = j12 * gamma, r22 = j22 * gamma;
#pragma unroll 1
  for (int it = 0; it < 5; ++it) {
    const float f00 = r11 * r22 - r21 * r12;
    const float f10 = r21 * r02 - r01 * r22;
    const float f20 = r01 * r12 - r11 * r02;
    const float f01 = r12 * r20 - r22 * r10;
    const float f11 = r22 * r00 - r02 * r20;
    const float f21 = r02 * r10 - r12 * r00;
    const float f02 = r10 * r21 - r20 * r11;
    const float f12 = r20 * r01 - r00 * r21;
    const float f22 = r00 * r11 - r10 * r01;
    const float det_r = r00 * f00 + r10 * f10 + r20 * f20;
    if (fabsf(det_r) > 1e-12f) {
      const float inv = 1.0f / det_r;
      r00 = 0.5f * (r00 + f00 * inv);
      r10 = 0.5f * (r10 + f10 * inv);
      r20 = 0.5f * (r20 + f20 * inv);
      r01 = 0.5f * (r01 + f01 * inv);
      r11 = 0.5f * (r11 + f11 * inv);
      r21 = 0.5f * (r21 + f21 * inv);
      r02 = 0.5f * (r02 + f02 * inv);
      r12 = 0.5f * (r12 + f12 * inv);
      r22 = 0.5f * (r22 + f22 * inv);
    }
  }
  const float det_f = r00 * (r11 * r22 - r21 * r12) + r10 * (r21 * r02 - r01 * r22) +
                      r20 * (r01 * r12 - r11 * r02);
  if (det_f < 0.0f) {
    r02 = -r02;
    r12 = -r12;
    r22 = -r22;
  }
  R[0] = r00; R[1] = r10; R[2] = r20; R[3] = r01; R[4] = r11; R[5] = r21;
  R[6] = r02; R[7] = r12; R[8] = r22;
  A.stretch[idx] = stretch_v;
  // what k_deform reads of the tet, in six 16-byte words: the four deformed
  // corners, J (column-major) and the stretch
  float4* P4 = reinterpret_cast<float4*>(A.pack + (size_t)idx * CGS_TET_PACK);
  P4[0] = make_float4(x0x, x0y, x0z, x1x);
  P4[1] = make_float4(x1y, x1z, x2x, x2y);
  P4[2] = make_float4(x2z, x3x, x3y, x3z);
  P4[3] = make_float4(j00, j10, j20, j01);
  P4[4] = make_float4(j11, j21, j02, j12);
  P4[5] = make_float4(j22, stretch_v, 0.0f, 0.0f);
}

__device__ __forceinline__ float det3s(float xx, float xy, float xz, float yy, float yz,
                                       float zz) {
  return xx * (yy * zz - yz * yz) - xy * (xy * zz - xz * yz) + xz * (xy * yz - xz * yy);
}

__device__ __forceinline__ void store_splat(const CgsDeformArgs& A, uint32_t i, float x, float y,
                                            float z, float c0, float c1, float c2, float c3,
                                            float c4, float c5, float o) {
  float* g = A.out_geo;
  g[cgs_blk(i, 0, CGS_GEO_K)] = x;
  g[cgs_blk(i, 1, CGS_GEO_K)] = y;
  g[cgs_blk(i, 2, CGS_GEO_K)] = z;
  g[cgs_blk(i, 3, CGS_GEO_K)] = c0;
  g[cgs_blk(i, 4, CGS_GEO_K)] = c1;
  g[cgs_blk(i, 5, CGS_GEO_K)] = c2;
  g[cgs_blk(i, 6, CGS_GEO_K)] = c3;
  g[cgs_blk(i, 7, CGS_GEO_K)] = c4;
  g[cgs_blk(i, 8, CGS_GEO_K)] = c5;
  g[cgs_blk(i, 9, CGS_GEO_K)] = o;
}

struct Child {
  float x, y, z, c00, c01, c02, c11, c12, c22, alpha;
  uint32_t tet;
};

// One splat of deform_and_split_compact_kernel: writes the splat (or its first
// child) at idx and returns true with the second child in `ch` when it splits.
__device__ __forceinline__ bool deform_splat(const CgsDeformArgs& A, uint32_t idx, Child& ch) {
  const bool valid = true;
  const float* RG = A.rest_geo;
  bool split = false;
  float px = 0, py = 0, pz = 0, dx = 0, dy = 0, dz = 0;
  float c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0, child_alpha = 0;
  uint32_t tet_raw = 0xFFFFFFFFu;
  if (valid) {
    tet_raw = A.bind_tet[idx];
    const float s_xx = RG[cgs_blk(idx, 3, CGS_GEO_K)], s_xy = RG[cgs_blk(idx, 4, CGS_GEO_K)];
    const float s_xz = RG[cgs_blk(idx, 5, CGS_GEO_K)], s_yy = RG[cgs_blk(idx, 6, CGS_GEO_K)];
    const float s_yz = RG[cgs_blk(idx, 7, CGS_GEO_K)], s_zz = RG[cgs_blk(idx, 8, CGS_GEO_K)];
    const float alpha = RG[cgs_blk(idx, 9, CGS_GEO_K)];
    const float mip = A.mip;
    if (tet_raw >= A.num_tets) {
      const float x = RG[cgs_blk(idx, 0, CGS_GEO_K)], y = RG[cgs_blk(idx, 1, CGS_GEO_K)];
      const float z = RG[cgs_blk(idx, 2, CGS_GEO_K)];
      if (mip > 0.0f) {
        const float det_orig = cgs_fmax(det3s(s_xx, s_xy, s_xz, s_yy, s_yz, s_zz), 1e-24f);
        const float r_xx = s_xx + mip, r_yy = s_yy + mip, r_zz = s_zz + mip;
        const float det_reg = cgs_fmax(det3s(r_xx, s_xy, s_xz, r_yy, s_yz, r_zz), det_orig);
        const float comp = cgs_sqrtf(cgs_fclamp(det_orig / det_reg, 0.0f, 1.0f));
        store_splat(A, idx, x, y, z, r_xx, s_xy, s_xz, r_yy, s_yz, r_zz,
                    cgs_fclamp(alpha * comp, 0.0f, 1.0f));
      } else {
        store_splat(A, idx, x, y, z, s_xx, s_xy, s_xz, s_yy, s_yz, s_zz, alpha);
      }
    } else {
      // the splat's weights and its tet's packed record (corners, J, stretch):
      // seven 16-byte loads
      const float4 W = *reinterpret_cast<const float4*>(A.bind_w + (size_t)idx * 4);
      const float w0 = W.x, w1 = W.y, w2 = W.z, w3 = W.w;
      const float4* T4 = reinterpret_cast<const float4*>(A.tet_pack + (size_t)tet_raw * CGS_TET_PACK);
      const float4 q0 = T4[0], q1 = T4[1], q2 = T4[2], q3 = T4[3], q4 = T4[4], q5 = T4[5];
      px = w0 * q0.x + w1 * q0.w + w2 * q1.z + w3 * q2.y;
      py = w0 * q0.y + w1 * q1.x + w2 * q1.w + w3 * q2.z;
      pz = w0 * q0.z + w1 * q1.y + w2 * q2.x + w3 * q2.w;
      const float j00 = q3.x, j10 = q3.y, j20 = q3.z, j01 = q3.w, j11 = q4.x, j21 = q4.y;
      const float j02 = q4.z, j12 = q4.w, j22 = q5.x;
      const float tet_stretch = q5.y;
      const float k00 = j00 * s_xx + j01 * s_xy + j02 * s_xz;
      const float k01 = j00 * s_xy + j01 * s_yy + j02 * s_yz;
      const float k02 = j00 * s_xz + j01 * s_yz + j02 * s_zz;
      const float k10 = j10 * s_xx + j11 * s_xy + j12 * s_xz;
      const float k11 = j10 * s_xy + j11 * s_yy + j12 * s_yz;
      const float k12 = j10 * s_xz + j11 * s_yz + j12 * s_zz;
      const float k20 = j20 * s_xx + j21 * s_xy + j22 * s_xz;
      const float k21 = j20 * s_xy + j21 * s_yy + j22 * s_yz;
      const float k22 = j20 * s_xz + j21 * s_yz + j22 * s_zz;
      const float m00 = cgs_fmax(1e-4f, k00 * j00 + k01 * j01 + k02 * j02);
      const float m01 =
          0.5f * ((k00 * j10 + k01 * j11 + k02 * j12) + (k10 * j00 + k11 * j01 + k12 * j02));
      const float m02 =
          0.5f * ((k00 * j20 + k01 * j21 + k02 * j22) + (k20 * j00 + k21 * j01 + k22 * j02));
      const float m11 = cgs_fmax(1e-4f, k10 * j10 + k11 * j11 + k12 * j12);
      const float m12 =
          0.5f * ((k10 * j20 + k11 * j21 + k12 * j22) + (k20 * j10 + k21 * j11 + k22 * j12));
      const float m22 = cgs_fmax(1e-4f, k20 * j20 + k21 * j21 + k22 * j22);
      const float det_j = j00 * (j11 * j22 - j12 * j21) - j01 * (j10 * j22 - j12 * j20) +
                          j02 * (j10 * j21 - j11 * j20);
      const float s_det = cgs_fclamp(det_j, 0.6f, 1.8f);
      const float safe_base = cgs_fclamp(1.0f - alpha, 1e-4f, 0.999f);
      const float alpha_p =
          cgs_fclamp(1.0f - cgs_powf_pos(safe_base, 1.0f / s_det), 0.05f, 0.99f);
      split = A.split_threshold > 0.0f && tet_stretch > A.split_threshold;
      if (!split) {
        if (mip > 0.0f) {
          const float det_orig = cgs_fmax(det3s(m00, m01, m02, m11, m12, m22), 1e-24f);
          const float r00 = m00 + mip, r11 = m11 + mip, r22 = m22 + mip;
          const float det_reg = cgs_fmax(det3s(r00, m01, m02, r11, m12, r22), det_orig);
          const float comp = cgs_sqrtf(cgs_fclamp(det_orig / det_reg, 0.0f, 1.0f));
          store_splat(A, idx, px, py, pz, r00, m01, m02, r11, m12, r22,
                      cgs_fclamp(alpha_p * comp, 0.0f, 1.0f));
        } else {
          store_splat(A, idx, px, py, pz, m00, m01, m02, m11, m12, m22, alpha_p);
        }
      } else {
        const float p1 = m01 * m01 + m02 * m02 + m12 * m12;
        float lambda0 = cgs_fmax(m00, cgs_fmax(m11, m22));
        float ux = 1.0f, uy = 0.0f, uz = 0.0f;
        if (p1 <= 1e-10f) {
          if (m11 >= m00 && m11 >= m22) {
            ux = 0.0f; uy = 1.0f; uz = 0.0f;
          } else if (m22 >= m00 && m22 >= m11) {
            ux = 0.0f; uy = 0.0f; uz = 1.0f;
          }
        } else {
          const float q = (m00 + m11 + m22) / 3.0f;
          const float d0 = m00 - q, d1 = m11 - q, d2 = m22 - q;
          const float p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0f * p1;
          const float p = cgs_sqrtf(cgs_fmax(p2 / 6.0f, 1e-12f));
          const float ip = 1.0f / p;
          const float b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip;
          const float b01 = m01 * ip, b02 = m02 * ip, b12 = m12 * ip;
          const float det_b = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
                              b02 * (b01 * b12 - b11 * b02);
          const float r = cgs_fclamp(0.5f * det_b, -1.0f, 1.0f);
          const float phi = cgs_acosf(r) / 3.0f;
          lambda0 = q + 2.0f * p * cgs_cosf(phi);
          const float r0x = m00 - lambda0, r0y = m01, r0z = m02;
          const float r1x = m01, r1y = m11 - lambda0, r1z = m12;
          const float r2x = m02, r2y = m12, r2z = m22 - lambda0;
          const float c01x = r0y * r1z - r0z * r1y, c01y = r0z * r1x - r0x * r1z,
                      c01z = r0x * r1y - r0y * r1x;
          const float n01 = c01x * c01x + c01y * c01y + c01z * c01z;
          const float c12x = r1y * r2z - r1z * r2y, c12y = r1z * r2x - r1x * r2z,
                      c12z = r1x * r2y - r1y * r2x;
          const float n12 = c12x * c12x + c12y * c12y + c12z * c12z;
          const float c02x = r0y * r2z - r0z * r2y, c02y = r0z * r2x - r0x * r2z,
                      c02z = r0x * r2y - r0y * r2x;
          const float n02 = c02x * c02x + c02y * c02y + c02z * c02z;
          if (n01 >= n12 && n01 >= n02 && n01 > 1e-12f) {
            const float inv = 1.0f / cgs_sqrtf(n01);
            ux = c01x * inv; uy = c01y * inv; uz = c01z * inv;
          } else if (n12 >= n02 && n12 > 1e-12f) {
            const float inv = 1.0f / cgs_sqrtf(n12);
            ux = c12x * inv; uy = c12y * inv; uz = c12z * inv;
          } else if (n02 > 1e-12f) {
            const float inv = 1.0f / cgs_sqrtf(n02);
            ux = c02x * inv; uy = c02y * inv; uz = c02z * inv;
          }
        }
        const float safe_l0 = cgs_fmax(lambda0, 1e-8f);
        const float mag = 0.5f * cgs_sqrtf(safe_l0);
        dx = ux * mag;
        dy = uy * mag;
        dz = uz * mag;
        const float shrink = 0.75f * safe_l0;
        c00 = cgs_fmax(1e-8f, m00 - shrink * ux * ux);
        c01 = m01 - shrink * ux * uy;
        c02 = m02 - shrink * ux * uz;
        c11 = cgs_fmax(1e-8f, m11 - shrink * uy * uy);
        c12 = m12 - shrink * uy * uz;
        c22 = cgs_fmax(1e-8f, m22 - shrink * uz * uz);
        child_alpha = alpha_p;
        if (mip > 0.0f) {
          const float det_orig = cgs_fmax(det3s(c00, c01, c02, c11, c12, c22), 1e-24f);
          c00 += mip;
          c11 += mip;
          c22 += mip;
          const float det_reg = cgs_fmax(det3s(c00, c01, c02, c11, c12, c22), det_orig);
          const float comp = cgs_sqrtf(cgs_fclamp(det_orig / det_reg, 0.0f, 1.0f));
          child_alpha = cgs_fclamp(child_alpha * comp, 0.0f, 1.0f);
        }
        store_splat(A, idx, px - dx, py - dy, pz - dz, c00, c01, c02, c11, c12, c22,
                    child_alpha);
      }
    }
  }

  ch.x = px + dx;
  ch.y = py + dy;
  ch.z = pz + dz;
  ch.c00 = c00;
  ch.c01 = c01;
  ch.c02 = c02;
  ch.c11 = c11;
  ch.c12 = c12;
  ch.c22 = c22;
  ch.alpha = child_alpha;
  ch.tet = tet_raw;
  return split;
}

constexpr int DI = 4;  // splats per thread: a partition is 1024 splats

__global__ __launch_bounds__(256) void k_deform(CgsDeformArgs A) {
  __shared__ uint32_t s_part;
  __shared__ uint32_t s_tmp[4];
  __shared__ uint64_t s_base;
  const uint32_t t = threadIdx.x;
  if (t == 0) s_part = atomicAdd(A.ticket, 1u);
  __syncthreads();
  const uint32_t part = s_part;
  Child ch[DI];
  bool sp[DI];
#pragma unroll
  for (int i = 0; i < DI; ++i) {
    const uint32_t idx = part * (256u * DI) + (uint32_t)i * 256u + t;
    sp[i] = idx < A.num_gaussians ? deform_splat(A, idx, ch[i]) : false;
  }
  // deterministic child slots: split flags scanned in splat-index order
  // (row i of the partition before row i+1), partitions chained by look-back
  uint32_t loc[DI];
  uint32_t run = 0;
#pragma unroll
  for (int i = 0; i < DI; ++i) {
    uint32_t tot;
    loc[i] = run + cgs::block_exclusive_scan<256>(sp[i] ? 1u : 0u, s_tmp, tot);
    run += tot;
  }
  if (t == 0) {
    const uint64_t excl = cgs::lookback_u64(A.status, part, (uint64_t)run);
    s_base = excl;
    if (part == gridDim.x - 1) *A.split_count = (uint32_t)(excl + run);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < DI; ++i) {
    if (!sp[i]) continue;
    const uint32_t idx = part * (256u * DI) + (uint32_t)i * 256u + t;
    const uint32_t off = (uint32_t)s_base + loc[i];
    store_splat(A, A.num_gaussians + off, ch[i].x, ch[i].y, ch[i].z, ch[i].c00, ch[i].c01,
                ch[i].c02, ch[i].c11, ch[i].c12, ch[i].c22, ch[i].alpha);
    A.split_meta[2 * off] = idx;
    A.split_meta[2 * off + 1] = ch[i].tet;
  }
}

}  // namespace

void cgs_launch_tet_transforms(const CgsTetArgs& a, hipStream_t s) {
  if (a.num_tets == 0) return;
  hipLaunchKernelGGL(k_tet_transforms, dim3((a.num_tets + 63u) / 64u), dim3(64), 0, s, a);
}

uint32_t cgs_deform_num_parts(uint32_t n) { return (n + 256u * DI - 1u) / (256u * DI); }

bool cgs_launch_deform(const CgsDeformArgs& a, hipStream_t s) {
  // the tet record and the weights are read with 16-byte loads
  if (((uintptr_t)a.tet_pack | (uintptr_t)a.bind_w) & 15u) return false;
  const uint32_t parts = cgs_deform_num_parts(a.num_gaussians);
  if (parts == 0) return true;
  hipLaunchKernelGGL(k_deform, dim3(parts), dim3(256), 0, s, a);
  return true;
}
