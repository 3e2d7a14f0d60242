// k_project.hip — Stage 2: EWA projection, opacity-aware tight extents and
// SH degree 0-3 colour along R^T d, one thread per splat.
//
// Follows project_gaussians_v2_kernel (src/render/kernels/project.rs:1018-1340)
// and eval_sh_f32 (797-895), and its f16-storage twin (1344-1679), operation
// for operation (compiled with -ffp-contract=off; ln via cgs_logf) so that the
// outputs are bit-identical to the CPU restatement.
//
// MI355X layout: the scene is wave-blocked (64 splats per block, channel-major
// inside), so every load instruction of a wave reads 256 contiguous bytes of
// one small region; the outputs are a
// 32-byte geometry record and a 16-byte colour record per splat, both aligned
// so the compositor's gathers never straddle a cache line.  The reference's
// surviving_ids compaction is never read downstream (gpu_state.rs:900-938);
// only its count is kept, and the binning pass derives it from the extents
// (same-address device atomics serialise at ~12 ns each on MI355X, so the
// projection itself does none).
#include "cgs_internal.h"

namespace {

template <bool F16>
__device__ __forceinline__ float ld_cov(const CgsSceneDev& s, uint32_t k, uint32_t i) {
  if (F16) return (float)__builtin_bit_cast(_Float16, s.cov_h[cgs_blk(i, k, 6u)]);
  return s.geo[cgs_blk(i, CGS_CH_COV + k, CGS_GEO_K)];
}

// SH coefficient k (0..47 in the f32 record order) of splat i.  The f32 SH
// stream is read once per frame and is most of Stage 2's bytes: non-temporal
// loads keep it from evicting the projected records Stage 3 reads next
// (measured at 1.07M splats, 1080p: Stage 2 61 -> 52.5 us, Stage 3 -3 %).
// Measured slower and left plain: the same for the position and covariance
// loads, and for the f16 storage's 2-byte SH loads (34.7 -> 37.3 us).
template <bool F16>
__device__ __forceinline__ float ld_sh(const CgsSceneDev& s, uint32_t k, uint32_t i) {
  if (F16) return (float)__builtin_bit_cast(_Float16, s.sh_h[cgs_blk(i, k, CGS_SH_K)]);
  return __builtin_nontemporal_load(&s.sh[cgs_blk(i, k, CGS_SH_K)]);
}

// eval_sh_f32 (project.rs:797-895); SH(k) gives coefficient k of the record
template <class SHL>
__device__ __forceinline__ void eval_sh(const SHL& SH, float dx, float dy, float dz, uint32_t deg,
                                        float& r, float& g, float& b) {
  float cr = 0.28209479177387814f * SH(0);
  float cg = 0.28209479177387814f * SH(1);
  float cb = 0.28209479177387814f * SH(2);
  if (deg >= 1) {
    float y1m1 = -0.4886025119029199f * dy;
    float y10 = 0.4886025119029199f * dz;
    float y1p1 = -0.4886025119029199f * dx;
    cr += y1m1 * SH(3) + y10 * SH(4) + y1p1 * SH(5);
    cg += y1m1 * SH(18) + y10 * SH(19) + y1p1 * SH(20);
    cb += y1m1 * SH(33) + y10 * SH(34) + y1p1 * SH(35);
    if (deg >= 2) {
      float xx = dx * dx, yy = dy * dy, zz = dz * dz;
      float xy = dx * dy, yz = dy * dz, xz = dx * dz;
      float y2m2 = 1.0925484305920792f * xy;
      float y2m1 = -1.0925484305920792f * yz;
      float y20 = 0.31539156525252005f * (2.0f * zz - xx - yy);
      float y2p1 = -1.0925484305920792f * xz;
      float y2p2 = 0.5462742152960396f * (xx - yy);
      cr += y2m2 * SH(6) + y2m1 * SH(7) + y20 * SH(8) +
            y2p1 * SH(9) + y2p2 * SH(10);
      cg += y2m2 * SH(21) + y2m1 * SH(22) +
            y20 * SH(23) + y2p1 * SH(24) +
            y2p2 * SH(25);
      cb += y2m2 * SH(36) + y2m1 * SH(37) +
            y20 * SH(38) + y2p1 * SH(39) +
            y2p2 * SH(40);
      if (deg >= 3) {
        float y3m3 = -0.5900435899266435f * dy * (3.0f * xx - yy);
        float y3m2 = 2.890611442640554f * xy * dz;
        float y3m1 = -0.4570457994644658f * dy * (4.0f * zz - xx - yy);
        float y30 = 0.3731763325901154f * dz * (2.0f * zz - 3.0f * xx - 3.0f * yy);
        float y3p1 = -0.4570457994644658f * dx * (4.0f * zz - xx - yy);
        float y3p2 = 1.445305721320277f * dz * (xx - yy);
        float y3p3 = -0.5900435899266435f * dx * (xx - 3.0f * yy);
        cr += y3m3 * SH(11) + y3m2 * SH(12) +
              y3m1 * SH(13) + y30 * SH(14) +
              y3p1 * SH(15) + y3p2 * SH(16) +
              y3p3 * SH(17);
        cg += y3m3 * SH(26) + y3m2 * SH(27) +
              y3m1 * SH(28) + y30 * SH(29) +
              y3p1 * SH(30) + y3p2 * SH(31) +
              y3p3 * SH(32);
        cb += y3m3 * SH(41) + y3m2 * SH(42) +
              y3m1 * SH(43) + y30 * SH(44) +
              y3p1 * SH(45) + y3p2 * SH(46) +
              y3p3 * SH(47);
      }
    }
  }
  r = cgs_fclamp(cr + 0.5f, 0.0f, 1.0f);
  g = cgs_fclamp(cg + 0.5f, 0.0f, 1.0f);
  b = cgs_fclamp(cb + 0.5f, 0.0f, 1.0f);
}

// The camera-dependent geometry of one splat (project_gaussians_v2_kernel,
// project.rs:1018-1284 up to the colour): returns the packed extents (0 =
// culled) and the 32-byte record in g0, g1.  cov = the 3D covariance (xx, xy,
// xz, yy, yz, zz) as stored, op = the activated opacity.
__device__ __forceinline__ uint32_t project_geometry(const float* cam, float px, float py,
                                                     float pz, float op, const float* cov,
                                                     float mip, float min_opacity,
                                                     float min_radius, float4& g0, float4& g1) {
  const float w00 = cam[0], w10 = cam[1], w20 = cam[2];
  const float w01 = cam[4], w11 = cam[5], w21 = cam[6];
  const float w02 = cam[8], w12 = cam[9], w22 = cam[10];
  const float w03 = cam[12], w13 = cam[13], w23 = cam[14];
  const float focal_x = cam[16], focal_y = cam[17];
  const float principal_x = cam[18], principal_y = cam[19];
  const float vw = cam[20], vh = cam[21], near_plane = cam[22];
  const float tx = w00 * px + w01 * py + w02 * pz + w03;
  const float ty = w10 * px + w11 * py + w12 * pz + w13;
  const float tz = w20 * px + w21 * py + w22 * pz + w23;
  if (tz <= near_plane) return 0u;
  if (op < min_opacity) return 0u;
  float v_xx = cov[0], v_xy = cov[1], v_xz = cov[2], v_yy = cov[3], v_yz = cov[4], v_zz = cov[5];
  float alpha_3d = op;
  if (mip > 0.0f) {
    float det_orig = cgs_fmax(v_xx * (v_yy * v_zz - v_yz * v_yz) -
                                  v_xy * (v_xy * v_zz - v_xz * v_yz) +
                                  v_xz * (v_xy * v_yz - v_xz * v_yy),
                              1e-24f);
    v_xx += mip;
    v_yy += mip;
    v_zz += mip;
    float det_reg = cgs_fmax(v_xx * (v_yy * v_zz - v_yz * v_yz) -
                                 v_xy * (v_xy * v_zz - v_xz * v_yz) +
                                 v_xz * (v_xy * v_yz - v_xz * v_yy),
                             det_orig);
    float comp = cgs_sqrtf(cgs_fclamp(det_orig / det_reg, 0.0f, 1.0f));
    alpha_3d = cgs_fclamp(op * comp, 0.0f, 1.0f);
    if (alpha_3d < min_opacity) return 0u;
  }
  const float k00 = w00 * v_xx + w01 * v_xy + w02 * v_xz;
  const float k01 = w00 * v_xy + w01 * v_yy + w02 * v_yz;
  const float k02 = w00 * v_xz + w01 * v_yz + w02 * v_zz;
  const float k10 = w10 * v_xx + w11 * v_xy + w12 * v_xz;
  const float k11 = w10 * v_xy + w11 * v_yy + w12 * v_yz;
  const float k12 = w10 * v_xz + w11 * v_yz + w12 * v_zz;
  const float k20 = w20 * v_xx + w21 * v_xy + w22 * v_xz;
  const float cam_xx = k00 * w00 + k01 * w01 + k02 * w02;
  const float cam_xy = k00 * w10 + k01 * w11 + k02 * w12;
  const float cam_xz = k00 * w20 + k01 * w21 + k02 * w22;
  const float cam_yy = k10 * w10 + k11 * w11 + k12 * w12;
  const float cam_yz = k10 * w20 + k11 * w21 + k12 * w22;
  const float k21 = w20 * v_xy + w21 * v_yy + w22 * v_yz;
  const float k22 = w20 * v_xz + w21 * v_yz + w22 * v_zz;
  const float cam_zz = k20 * w20 + k21 * w21 + k22 * w22;

  const float inv_tz = 1.0f / tz;
  const float inv_tz2 = inv_tz * inv_tz;
  const float j00 = focal_x * inv_tz;
  const float j02 = -focal_x * tx * inv_tz2;
  const float j11 = focal_y * inv_tz;
  const float j12 = -focal_y * ty * inv_tz2;
  const float m00 = j00 * cam_xx + j02 * cam_xz;
  const float m01 = j00 * cam_xy + j02 * cam_yz;
  const float m02 = j00 * cam_xz + j02 * cam_zz;
  const float m11 = j11 * cam_yy + j12 * cam_yz;
  const float m12 = j11 * cam_yz + j12 * cam_zz;
  const float s_xx = m00 * j00 + m02 * j02;
  const float s_xy = m01 * j11 + m02 * j12;
  const float s_yy = m11 * j11 + m12 * j12;

  const float raw_mid = 0.5f * (s_xx + s_yy);
  const float raw_det = cgs_fmax(s_xx * s_yy - s_xy * s_xy, 0.0f);
  const float raw_discr = cgs_sqrtf(cgs_fmax(raw_mid * raw_mid - raw_det, 0.0f));
  const float raw_lmax = cgs_fmax(raw_mid + raw_discr, 0.0f);
  if (3.0f * cgs_sqrtf(raw_lmax) < min_radius) return 0u;
  const float sp_xx = s_xx + 0.3f;
  const float sp_xy = s_xy;
  const float sp_yy = s_yy + 0.3f;
  const float det = sp_xx * sp_yy - sp_xy * sp_xy;
  if (det <= 1e-6f) return 0u;
  const float comp2 = cgs_sqrtf(cgs_fclamp(raw_det / det, 0.0f, 1.0f));
  const float alpha_2d = cgs_fclamp(alpha_3d * comp2, 0.0f, 1.0f);
  if (alpha_2d < min_opacity || alpha_2d <= (1.0f / 255.0f)) return 0u;
  const float q_max = cgs_fclamp(2.0f * cgs_logf(255.0f * alpha_2d), 0.0f, 8.0f);
  const float sigma_cut = cgs_sqrtf(q_max);
  const float ext_x =
      cgs_fclamp(ceilf(sigma_cut * cgs_sqrtf(cgs_fmax(sp_xx, 0.0f))), 1.0f, 1024.0f);
  const float ext_y =
      cgs_fclamp(ceilf(sigma_cut * cgs_sqrtf(cgs_fmax(sp_yy, 0.0f))), 1.0f, 1024.0f);
  const float mu_x = focal_x * tx * inv_tz + principal_x;
  const float mu_y = focal_y * ty * inv_tz + principal_y;
  if (mu_x + ext_x < 0.0f || mu_x - ext_x >= vw || mu_y + ext_y < 0.0f || mu_y - ext_y >= vh)
    return 0u;
  const float inv_det = 1.0f / det;
  const float ca = sp_yy * inv_det;
  const float cb = -sp_xy * inv_det;
  const float cc = sp_xx * inv_det;
  g0 = make_float4(mu_x, mu_y, ca, cb);
  g1 = make_float4(cc, alpha_2d, tz, q_max);
  return ((uint32_t)ext_x & 0xFFFFu) | (((uint32_t)ext_y & 0xFFFFu) << 16);
}

// The SH record index of splat idx (split children read their parent's) and
// its tet (num_tets when it has none)
__device__ __forceinline__ void sh_source(const CgsProjectArgs& A, uint32_t idx, uint32_t& sh_idx,
                                          uint32_t& tet_raw) {
  sh_idx = idx;
  tet_raw = A.num_tets;
  if (idx < A.num_rest) {
    if (A.has_deformation) tet_raw = A.bind_tet[idx];
  } else {
    const uint32_t off = idx - A.num_rest;
    sh_idx = A.split_meta[2 * off];
    tet_raw = A.split_meta[2 * off + 1];
  }
}

// The SH direction: the normalised camera-to-splat vector, taken into the
// bound tet's rest frame by R^T when the splat is deformed (project.rs:1286-1318)
__device__ __forceinline__ void sh_direction(const CgsProjectArgs& A, float px, float py, float pz,
                                             const float* cam_pos, uint32_t tet_raw, float& dx,
                                             float& dy, float& dz) {
  const float vx = px - cam_pos[0];
  const float vy = py - cam_pos[1];
  const float vz = pz - cam_pos[2];
  const float inv_len = 1.0f / cgs_fmax(0.0001f, cgs_sqrtf(vx * vx + vy * vy + vz * vz));
  dx = vx * inv_len;
  dy = vy * inv_len;
  dz = vz * inv_len;
  if (A.has_deformation && tet_raw < A.num_tets) {
    const float* R = A.tet_rot + (size_t)tet_raw * 9;
    const float r00 = R[0], r10 = R[1], r20 = R[2];
    const float r01 = R[3], r11 = R[4], r21 = R[5];
    const float r02 = R[6], r12 = R[7], r22 = R[8];
    const float ldx = r00 * dx + r10 * dy + r20 * dz;
    const float ldy = r01 * dx + r11 * dy + r21 * dz;
    const float ldz = r02 * dx + r12 * dy + r22 * dz;
    dx = ldx;
    dy = ldy;
    dz = ldz;
  }
}

// The SH coefficients a splat of degree DEG uses, loaded together once the
// splat survives (the degree is a template argument, so the bands above it
// are neither loaded nor evaluated).
template <int DEG>
struct ShRec {
  static constexpr int NB = (DEG + 1) * (DEG + 1);  // bases per channel
  float c[3 * NB];
};
// record index of basis j of channel ch: dc r g b, then 15 rest per channel
__device__ __forceinline__ constexpr uint32_t sh_k(int ch, int j) {
  return j == 0 ? (uint32_t)ch : (uint32_t)(3 + ch * 15 + (j - 1));
}
template <bool F16, int DEG>
__device__ __forceinline__ void load_rec(const CgsSceneDev& S, uint32_t i, ShRec<DEG>& R) {
#pragma unroll
  for (int ch = 0; ch < 3; ++ch)
#pragma unroll
    for (int j = 0; j < ShRec<DEG>::NB; ++j)
      R.c[ch * ShRec<DEG>::NB + j] = ld_sh<F16>(S, sh_k(ch, j), i);
}
template <int DEG>
struct ShView {
  const ShRec<DEG>& R;
  // coefficient k of the record (k constant after inlining)
  __device__ __forceinline__ float operator()(uint32_t k) const {
    const int ch = k < 3u ? (int)k : (int)((k - 3u) / 15u);
    const int j = k < 3u ? 0 : (int)((k - 3u) % 15u) + 1;
    return R.c[ch * ShRec<DEG>::NB + j];
  }
};

template <bool F16, int DEG>
__global__ __launch_bounds__(256) void k_project(CgsProjectArgs A) {
  const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
  if (idx >= A.max_gaussians) return;
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  if (idx >= active) {
    A.extents[idx] = 0u;
    if (A.zero_culled) {
      float4* g = reinterpret_cast<float4*>(A.geom + (size_t)idx * 8);
      g[0] = g[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return;
  }
  const CgsSceneDev& S = A.scene;
  const float px = S.geo[cgs_blk(idx, 0, CGS_GEO_K)];
  const float py = S.geo[cgs_blk(idx, 1, CGS_GEO_K)];
  const float pz = S.geo[cgs_blk(idx, 2, CGS_GEO_K)];
  const float op = S.geo[cgs_blk(idx, CGS_CH_OPA, CGS_GEO_K)];
  float cov[6];
#pragma unroll
  for (uint32_t k = 0; k < 6; ++k) cov[k] = ld_cov<F16>(S, k, idx);
  uint32_t sh_idx, tet_raw;
  sh_source(A, idx, sh_idx, tet_raw);
  ShRec<DEG> R;
  float4 g0, g1;
  const uint32_t ext =
      project_geometry(A.camera, px, py, pz, op, cov, A.mip, A.min_opacity, A.min_radius, g0, g1);
  if (ext != 0u) {
    load_rec<F16, DEG>(S, sh_idx, R);
    float dx, dy, dz;
    sh_direction(A, px, py, pz, A.cam_pos, tet_raw, dx, dy, dz);
    float cr, cg, cbl;
    eval_sh(ShView<DEG>{R}, dx, dy, dz, (uint32_t)DEG, cr, cg, cbl);
    float4* g = reinterpret_cast<float4*>(A.geom + (size_t)idx * 8);
    g[0] = g0;
    g[1] = g1;
    reinterpret_cast<float4*>(A.color)[idx] = make_float4(cr, cg, cbl, 0.0f);
  } else if (A.zero_culled) {
    float4* g = reinterpret_cast<float4*>(A.geom + (size_t)idx * 8);
    g[0] = g[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  A.extents[idx] = ext;
}

// Stage 2 of up to 64 views of one scene in one pass: the splat's position,
// opacity and covariance are read once, its SH record once if it survives in
// any view (then evaluated per view from registers), and each view gets the
// records k_project would write for it alone.
template <bool F16>
__global__ __launch_bounds__(256) void k_project_views(CgsProjectArgs A,
                                                       const CgsViewProjection* views,
                                                       uint32_t n_views) {
  const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
  if (idx >= A.max_gaussians) return;
  const uint32_t active = A.num_rest + (A.split_count ? A.split_count[0] : 0u);
  if (idx >= active) {
    for (uint32_t v = 0; v < n_views; ++v) views[v].extents[idx] = 0u;
    return;
  }
  const CgsSceneDev& S = A.scene;
  const float px = S.geo[cgs_blk(idx, 0, CGS_GEO_K)];
  const float py = S.geo[cgs_blk(idx, 1, CGS_GEO_K)];
  const float pz = S.geo[cgs_blk(idx, 2, CGS_GEO_K)];
  const float op = S.geo[cgs_blk(idx, CGS_CH_OPA, CGS_GEO_K)];
  float cov[6];
#pragma unroll
  for (uint32_t k = 0; k < 6; ++k) cov[k] = ld_cov<F16>(S, k, idx);
  uint64_t alive = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    const CgsViewProjection& V = views[v];
    float4 g0, g1;
    const uint32_t ext =
        project_geometry(V.camera, px, py, pz, op, cov, A.mip, A.min_opacity, A.min_radius, g0, g1);
    if (ext != 0u) {
      float4* g = reinterpret_cast<float4*>(V.geom + (size_t)idx * 8);
      g[0] = g0;
      g[1] = g1;
      alive |= 1ull << v;
    }
    V.extents[idx] = ext;
  }
  if (alive == 0ull) return;
  uint32_t sh_idx, tet_raw;
  sh_source(A, idx, sh_idx, tet_raw);
  const uint32_t deg = A.sh_degree;
  const uint32_t rest = (deg + 1u) * (deg + 1u) - 1u;  // rest coefficients per channel
  float sh[48];
#pragma unroll
  for (uint32_t k = 0; k < 48; ++k) {
    // record order [dc r g b, rest_r 15, rest_g 15, rest_b 15]: load what the
    // degree evaluates
    const bool used = k < 3u || (k - 3u) % 15u < rest;
    sh[k] = used ? ld_sh<F16>(S, k, sh_idx) : 0.0f;
  }
  for (uint32_t v = 0; v < n_views; ++v) {
    if (!((alive >> v) & 1ull)) continue;
    const CgsViewProjection& V = views[v];
    float dx, dy, dz;
    sh_direction(A, px, py, pz, V.cam_pos, tet_raw, dx, dy, dz);
    float cr, cg, cbl;
    eval_sh([&](uint32_t k) { return sh[k]; }, dx, dy, dz, deg, cr, cg, cbl);
    reinterpret_cast<float4*>(V.color)[idx] = make_float4(cr, cg, cbl, 0.0f);
  }
}

}  // namespace

void cgs_launch_project_views(const CgsProjectArgs& a, const CgsViewProjection* views,
                              uint32_t n_views, hipStream_t s) {
  if (a.max_gaussians == 0 || n_views == 0) return;
  const uint32_t blocks = (a.max_gaussians + 255u) / 256u;
  if (a.scene.sh_h)
    hipLaunchKernelGGL(k_project_views<true>, dim3(blocks), dim3(256), 0, s, a, views, n_views);
  else
    hipLaunchKernelGGL(k_project_views<false>, dim3(blocks), dim3(256), 0, s, a, views, n_views);
}

void cgs_launch_project(const CgsProjectArgs& a, hipStream_t s) {
  if (a.max_gaussians == 0) return;
  const dim3 g((a.max_gaussians + 255u) / 256u), b(256);
  const uint32_t d = a.sh_degree > 3u ? 3u : a.sh_degree;
#define KP(F, D) hipLaunchKernelGGL((k_project<F, D>), g, b, 0, s, a)
  if (a.scene.sh_h) {
    if (d == 0) KP(true, 0);
    else if (d == 1) KP(true, 1);
    else if (d == 2) KP(true, 2);
    else KP(true, 3);
  } else {
    if (d == 0) KP(false, 0);
    else if (d == 1) KP(false, 1);
    else if (d == 2) KP(false, 2);
    else KP(false, 3);
  }
#undef KP
}
