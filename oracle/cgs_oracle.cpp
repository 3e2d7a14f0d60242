// cgs_oracle.cpp — CPU restatement of CinderGS's render and cage-deformation
// path, used ONLY as the checker by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  Nothing in the product links or calls it.
//
// It restates, statement by statement and in the same floating-point operation
// order, the reference's v2 GPU pipeline (the bit-exact target for keys, scan
// offsets, tile counts and ranges) so that it can run on the host:
//   Stage 1  src/render/kernels/deformation.rs:682-865  (tet transforms)
//            src/render/kernels/deformation.rs:872-1211 (deform + split)
//            with children appended in parent order, as the CPU path does
//            (src/deformation/cpu_reference.rs:316-377), instead of the
//            reference GPU's atomic (non-deterministic) order.
//   Stage 2  src/render/kernels/project.rs:1018-1340 (+ eval_sh_f32 797-895,
//            f16 storage variant 1344-1679)
//   Stage 3  src/render/kernels/binning.rs:12-417, scan.rs (exclusive scan),
//            radix_sort.rs (stable LSD sort == stable sort by key)
//   Stage 4  src/render/kernels/compositor.rs:721-883
//   MVC      src/deformation/mvc.rs:340-541 (inside test, exact 3D MVC,
//            central-difference gradients)
//   MVC scene src/deformation/mvc.rs:564-749 (deform_scene_cpu: polar
//            rotation, F Sigma F^T, SH rotation, strain split, 3D mip filter)
//   Ingest   src/scene/ply_loader.rs:540-598 (parse_vertex_payload),
//            soa.rs:113-126 (push_raw), 454-500 (quantize), 306-328 (export)
//
// Every routine is a template over the scalar type: the float instantiation is
// the FP32 oracle (elementary functions from cgs_detmath.h so that the result
// is defined bit for bit), the double instantiation is the FP64 twin used to
// state each FP32 error against an FP64 run.
//
// Parity pinning status (what checks each function besides the GPU):
//   count/emit/ellipse test : brute-force per-pixel coverage on random ellipses
//                              (tiles/mod.rs:302-401 procedure), closed forms
//   scan / sort / ranges    : reference vectors (gpu_parity_tests.rs:512-662),
//                              std::stable_sort, structural invariants
//   projection              : closed forms (projection.rs:556-740 analogues),
//                              SH constants/special directions (sh.rs:144-182),
//                              FP64 twin
//   compositor              : hand-computed single/double splat fixtures,
//                              early-termination bound (T < 1e-3), FP64 twin
//   tet transforms / deform : identity, uniform scale 2, pure rotation,
//                              reflection closed forms (integration_tests.rs
//                              1146-1319), polar identities, FP64 twin
//   MVC                     : partition of unity, linear reproduction,
//                              FP64 twin, finite-difference consistency
//   ingest (cgs_ingest_math.h, shared with the kernels): decode against the
//                              tests' own struct unpacking, activation against
//                              a float64 closed form (R(q) diag(s^2) R^T,
//                              sigmoid, exp), f16 against numpy's float16,
//                              atan2 against libm, the octahedral code by its
//                              decode round trip, export by reload
//   MVC scene (per-splat math of cgs_scene_math.h, shared with the kernel):
//                              SH rotation orthogonal / matches the rotated
//                              direction under this file's eval_sh / identity
//                              no-op / degree 0 untouched (sh_rotation.rs
//                              484-566), polar identities (polar.rs:116-205),
//                              eigen roundtrip, mip attenuation, split quarters
//                              the major variance (covariance.rs:388-417), and
//                              the whole record against a float64 restatement
//                              built on SVD / eigh / least squares
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../cindergs_amd/csrc/cgs_detmath.h"
#include "../cindergs_amd/csrc/cgs_scene_math.h"
#include "../cindergs_amd/csrc/cgs_ingest_math.h"

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

template <class T> struct Mth;
template <> struct Mth<float> {
  static float ln(float x) { return cgs_logf(x); }
  static float ex(float x) { return cgs_expf(x); }
  static float sq(float x) { return cgs_sqrtf(x); }
  static float pw(float x, float y) { return cgs_powf_pos(x, y); }
  static float ac(float x) { return cgs_acosf(x); }
  static float co(float x) { return cgs_cosf(x); }
  static float mx(float a, float b) { return cgs_fmax(a, b); }
  static float mn(float a, float b) { return cgs_fmin(a, b); }
  static float cl(float x, float lo, float hi) { return cgs_fclamp(x, lo, hi); }
  static float fl(float x) { return floorf(x); }
  static float ce(float x) { return ceilf(x); }
  static float rd(float x) { return roundf(x); }
  static float ab(float x) { return fabsf(x); }
  // compositor exponent: libm (the GPU uses the hardware exp2 within 1/255)
  static float cex(float x) { return expf(x); }
};
template <> struct Mth<double> {
  static double ln(double x) { return log(x); }
  static double ex(double x) { return exp(x); }
  static double sq(double x) { return sqrt(x); }
  static double pw(double x, double y) { return pow(x, y); }
  static double ac(double x) { return acos(x); }
  static double co(double x) { return cos(x); }
  static double mx(double a, double b) { return fmax(a, b); }
  static double mn(double a, double b) { return fmin(a, b); }
  static double cl(double x, double lo, double hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
  }
  static double fl(double x) { return floor(x); }
  static double ce(double x) { return ceil(x); }
  static double rd(double x) { return round(x); }
  static double ab(double x) { return fabs(x); }
  static double cex(double x) { return exp(x); }
};

static int g_threads = 1;

// ---------------------------------------------------------------------------
// Stage 2: project_gaussians_v2_kernel (project.rs:1018-1340)
// ---------------------------------------------------------------------------

// eval_sh_f32 (project.rs:797-895); `c` is the 48-coefficient record
// [dc r,g,b, rest_r 15, rest_g 15, rest_b 15] (sh.rs:56-61).
template <class T>
static void eval_sh(const T* c, T dx, T dy, T dz, uint32_t deg, T out[3]) {
  typedef Mth<T> m;
  // K(x): the reference's f32 literal; the FP64 twin uses the same (f32) values.
#define K(x) ((T)(x##f))
  T cr = K(0.28209479177387814) * c[0];
  T cg = K(0.28209479177387814) * c[1];
  T cb = K(0.28209479177387814) * c[2];
  if (deg >= 1) {
    T y1m1 = K(-0.4886025119029199) * dy;
    T y10 = K(0.4886025119029199) * dz;
    T y1p1 = K(-0.4886025119029199) * dx;
    cr += y1m1 * c[3] + y10 * c[4] + y1p1 * c[5];
    cg += y1m1 * c[18] + y10 * c[19] + y1p1 * c[20];
    cb += y1m1 * c[33] + y10 * c[34] + y1p1 * c[35];
    if (deg >= 2) {
      T xx = dx * dx, yy = dy * dy, zz = dz * dz;
      T xy = dx * dy, yz = dy * dz, xz = dx * dz;
      T y2m2 = K(1.0925484305920792) * xy;
      T y2m1 = K(-1.0925484305920792) * yz;
      T y20 = K(0.31539156525252005) * ((T)2 * zz - xx - yy);
      T y2p1 = K(-1.0925484305920792) * xz;
      T y2p2 = K(0.5462742152960396) * (xx - yy);
      cr += y2m2 * c[6] + y2m1 * c[7] + y20 * c[8] + y2p1 * c[9] + y2p2 * c[10];
      cg += y2m2 * c[21] + y2m1 * c[22] + y20 * c[23] + y2p1 * c[24] + y2p2 * c[25];
      cb += y2m2 * c[36] + y2m1 * c[37] + y20 * c[38] + y2p1 * c[39] + y2p2 * c[40];
      if (deg >= 3) {
        T y3m3 = K(-0.5900435899266435) * dy * ((T)3 * xx - yy);
        T y3m2 = K(2.890611442640554) * xy * dz;
        T y3m1 = K(-0.4570457994644658) * dy * ((T)4 * zz - xx - yy);
        T y30 = K(0.3731763325901154) * dz * ((T)2 * zz - (T)3 * xx - (T)3 * yy);
        T y3p1 = K(-0.4570457994644658) * dx * ((T)4 * zz - xx - yy);
        T y3p2 = K(1.445305721320277) * dz * (xx - yy);
        T y3p3 = K(-0.5900435899266435) * dx * (xx - (T)3 * yy);
        cr += y3m3 * c[11] + y3m2 * c[12] + y3m1 * c[13] + y30 * c[14] + y3p1 * c[15] +
              y3p2 * c[16] + y3p3 * c[17];
        cg += y3m3 * c[26] + y3m2 * c[27] + y3m1 * c[28] + y30 * c[29] + y3p1 * c[30] +
              y3p2 * c[31] + y3p3 * c[32];
        cb += y3m3 * c[41] + y3m2 * c[42] + y3m1 * c[43] + y30 * c[44] + y3p1 * c[45] +
              y3p2 * c[46] + y3p3 * c[47];
      }
    }
  }
#undef K
  out[0] = m::cl(cr + (T)0.5, (T)0, (T)1);
  out[1] = m::cl(cg + (T)0.5, (T)0, (T)1);
  out[2] = m::cl(cb + (T)0.5, (T)0, (T)1);
}

struct ProjParams {
  const float* camera;  // 24 floats (CameraUniforms, camera.rs:9-36)
  float cam_pos[3];
  float mip;
  float min_opacity;
  float min_radius;
  uint32_t sh_degree;
  uint32_t has_deformation;
  uint32_t num_tets;
  uint32_t num_rest;
};

// One splat of project_gaussians_v2_kernel. Returns true when the splat
// survives; `attrs` = [mu_x, mu_y, a, b, c, alpha_2d, r, g, b, tz].
template <class T>
static bool project_one(const ProjParams& P, const T pos[3], const T cov[6], T op,
                        const T* sh48, const float* rot9, T attrs[10], uint32_t* ext_packed) {
  typedef Mth<T> m;
  const float* cam = P.camera;
  T w00 = cam[0], w10 = cam[1], w20 = cam[2];
  T w01 = cam[4], w11 = cam[5], w21 = cam[6];
  T w02 = cam[8], w12 = cam[9], w22 = cam[10];
  T w03 = cam[12], w13 = cam[13], w23 = cam[14];
  T focal_x = cam[16], focal_y = cam[17], principal_x = cam[18], principal_y = cam[19];
  T vw = cam[20], vh = cam[21], near_plane = cam[22];

  T px = pos[0], py = pos[1], pz = pos[2];
  T tx = w00 * px + w01 * py + w02 * pz + w03;
  T ty = w10 * px + w11 * py + w12 * pz + w13;
  T tz = w20 * px + w21 * py + w22 * pz + w23;
  if (tz <= near_plane) return false;

  if (op < (T)P.min_opacity) return false;
  T v_xx = cov[0], v_xy = cov[1], v_xz = cov[2], v_yy = cov[3], v_yz = cov[4], v_zz = cov[5];
  T alpha_3d = op;
  T mip = (T)P.mip;
  if (mip > (T)0) {
    T det_orig = m::mx(v_xx * (v_yy * v_zz - v_yz * v_yz) - v_xy * (v_xy * v_zz - v_xz * v_yz) +
                           v_xz * (v_xy * v_yz - v_xz * v_yy),
                       (T)1e-24f);
    v_xx += mip;
    v_yy += mip;
    v_zz += mip;
    T det_reg = m::mx(v_xx * (v_yy * v_zz - v_yz * v_yz) - v_xy * (v_xy * v_zz - v_xz * v_yz) +
                          v_xz * (v_xy * v_yz - v_xz * v_yy),
                      det_orig);
    T comp = m::sq(m::cl(det_orig / det_reg, (T)0, (T)1));
    alpha_3d = m::cl(op * comp, (T)0, (T)1);
    if (alpha_3d < (T)P.min_opacity) return false;
  }

  T k00 = w00 * v_xx + w01 * v_xy + w02 * v_xz;
  T k01 = w00 * v_xy + w01 * v_yy + w02 * v_yz;
  T k02 = w00 * v_xz + w01 * v_yz + w02 * v_zz;
  T k10 = w10 * v_xx + w11 * v_xy + w12 * v_xz;
  T k11 = w10 * v_xy + w11 * v_yy + w12 * v_yz;
  T k12 = w10 * v_xz + w11 * v_yz + w12 * v_zz;
  T k20 = w20 * v_xx + w21 * v_xy + w22 * v_xz;
  T k21 = w20 * v_xy + w21 * v_yy + w22 * v_yz;
  T k22 = w20 * v_xz + w21 * v_yz + w22 * v_zz;

  T cam_xx = k00 * w00 + k01 * w01 + k02 * w02;
  T cam_xy = k00 * w10 + k01 * w11 + k02 * w12;
  T cam_xz = k00 * w20 + k01 * w21 + k02 * w22;
  T cam_yy = k10 * w10 + k11 * w11 + k12 * w12;
  T cam_yz = k10 * w20 + k11 * w21 + k12 * w22;
  T cam_zz = k20 * w20 + k21 * w21 + k22 * w22;
  (void)k21;
  (void)k22;

  T inv_tz = (T)1 / tz;
  T inv_tz2 = inv_tz * inv_tz;
  T j00 = focal_x * inv_tz;
  T j02 = -focal_x * tx * inv_tz2;
  T j11 = focal_y * inv_tz;
  T j12 = -focal_y * ty * inv_tz2;

  T m00 = j00 * cam_xx + j02 * cam_xz;
  T m01 = j00 * cam_xy + j02 * cam_yz;
  T m02 = j00 * cam_xz + j02 * cam_zz;
  T m11 = j11 * cam_yy + j12 * cam_yz;
  T m12 = j11 * cam_yz + j12 * cam_zz;

  T s_xx = m00 * j00 + m02 * j02;
  T s_xy = m01 * j11 + m02 * j12;
  T s_yy = m11 * j11 + m12 * j12;

  T raw_mid = (T)0.5 * (s_xx + s_yy);
  T raw_det = m::mx(s_xx * s_yy - s_xy * s_xy, (T)0);
  T raw_discr = m::sq(m::mx(raw_mid * raw_mid - raw_det, (T)0));
  T raw_lmax = m::mx(raw_mid + raw_discr, (T)0);
  if ((T)3 * m::sq(raw_lmax) < (T)P.min_radius) return false;

  T sp_xx = s_xx + (T)0.3f;
  T sp_xy = s_xy;
  T sp_yy = s_yy + (T)0.3f;
  T det = sp_xx * sp_yy - sp_xy * sp_xy;
  if (det <= (T)1e-6f) return false;
  T comp2 = m::sq(m::cl(raw_det / det, (T)0, (T)1));
  T alpha_2d = m::cl(alpha_3d * comp2, (T)0, (T)1);
  if (alpha_2d < (T)P.min_opacity || alpha_2d <= (T)(1.0f / 255.0f)) return false;
  T q_max = m::cl((T)2 * m::ln((T)255 * alpha_2d), (T)0, (T)8);
  T sigma_cut = m::sq(q_max);
  T ext_x = m::cl(m::ce(sigma_cut * m::sq(m::mx(sp_xx, (T)0))), (T)1, (T)1024);
  T ext_y = m::cl(m::ce(sigma_cut * m::sq(m::mx(sp_yy, (T)0))), (T)1, (T)1024);
  T mu_x = focal_x * tx * inv_tz + principal_x;
  T mu_y = focal_y * ty * inv_tz + principal_y;
  if (mu_x + ext_x < (T)0 || mu_x - ext_x >= vw || mu_y + ext_y < (T)0 || mu_y - ext_y >= vh)
    return false;
  T inv_det = (T)1 / det;
  T a = sp_yy * inv_det;
  T b = -sp_xy * inv_det;
  T c = sp_xx * inv_det;

  T vx = px - (T)P.cam_pos[0];
  T vy = py - (T)P.cam_pos[1];
  T vz = pz - (T)P.cam_pos[2];
  T inv_len = (T)1 / m::mx((T)0.0001f, m::sq(vx * vx + vy * vy + vz * vz));
  T dx = vx * inv_len, dy = vy * inv_len, dz = vz * inv_len;
  if (rot9) {
    T r00 = rot9[0], r10 = rot9[1], r20 = rot9[2];
    T r01 = rot9[3], r11 = rot9[4], r21 = rot9[5];
    T r02 = rot9[6], r12 = rot9[7], r22 = rot9[8];
    T ldx = r00 * dx + r10 * dy + r20 * dz;
    T ldy = r01 * dx + r11 * dy + r21 * dz;
    T ldz = r02 * dx + r12 * dy + r22 * dz;
    dx = ldx;
    dy = ldy;
    dz = ldz;
  }
  T rgb[3];
  eval_sh<T>(sh48, dx, dy, dz, P.sh_degree, rgb);
  attrs[0] = mu_x;
  attrs[1] = mu_y;
  attrs[2] = a;
  attrs[3] = b;
  attrs[4] = c;
  attrs[5] = alpha_2d;
  attrs[6] = rgb[0];
  attrs[7] = rgb[1];
  attrs[8] = rgb[2];
  attrs[9] = tz;
  uint32_t ex = (uint32_t)ext_x, ey = (uint32_t)ext_y;
  *ext_packed = (ex & 0xFFFFu) | ((ey & 0xFFFFu) << 16);
  return true;
}

// IEEE binary16 -> binary32, exact (the reference's native f16 path casts
// with f32::cast_from, project.rs:899-1009).
static float half_to_float(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t mant = h & 0x3ffu;
  if (e == 0) {
    if (mant == 0) return cgs_u2f(sign);
    float f = (float)mant * 5.9604644775390625e-8f;  // mant * 2^-24, exact
    return cgs_u2f(cgs_f2u(f) | sign);
  }
  if (e == 31) return cgs_u2f(sign | 0x7f800000u | (mant << 13));
  return cgs_u2f(sign | ((e + 112u) << 23) | (mant << 13));
}

template <class T, class OutT>
static uint32_t project_all(const ProjParams& P, const float* pos, const float* cov,
                            const uint16_t* cov_h, const float* opa, const float* sh,
                            const uint16_t* sh_h, const uint32_t* bind_tet,
                            const float* tet_rot, const uint32_t* split_meta,
                            uint32_t split_count, OutT* attrs10, uint32_t* extents,
                            uint32_t max_gaussians) {
  uint32_t active = P.num_rest + split_count;
  uint32_t survivors = 0;
  uint32_t bases = P.sh_degree == 0 ? 1 : P.sh_degree == 1 ? 4 : P.sh_degree == 2 ? 9 : 16;
  uint32_t h_stride = ((bases * 3 + 1) / 2) * 2;
#pragma omp parallel for schedule(static, 2048) reduction(+ : survivors) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)max_gaussians; ++ii) {
    uint32_t idx = (uint32_t)ii;
    if (idx >= active) {
      extents[idx] = 0;
      continue;
    }
    T p[3] = {(T)pos[idx * 3 + 0], (T)pos[idx * 3 + 1], (T)pos[idx * 3 + 2]};
    T cv[6];
    for (int k = 0; k < 6; ++k)
      cv[k] = cov_h ? (T)half_to_float(cov_h[(size_t)idx * 6 + k]) : (T)cov[(size_t)idx * 6 + k];
    uint32_t sh_idx = idx;
    uint32_t tet_raw = P.num_tets;
    if (idx < P.num_rest) {
      if (P.has_deformation) tet_raw = bind_tet[idx];
    } else {
      uint32_t off = idx - P.num_rest;
      sh_idx = split_meta[off * 2];
      tet_raw = split_meta[off * 2 + 1];
    }
    T shv[48];
    if (sh_h) {
      // channel-major trimmed f16 record -> 48-coefficient record
      for (int k = 0; k < 48; ++k) shv[k] = 0;
      const uint16_t* s = sh_h + (size_t)sh_idx * h_stride;
      shv[0] = half_to_float(s[0]);
      shv[1] = half_to_float(s[1]);
      shv[2] = half_to_float(s[2]);
      uint32_t rest = bases - 1;
      for (uint32_t k = 0; k < rest; ++k) {
        shv[3 + k] = half_to_float(s[3 + k]);
        shv[18 + k] = half_to_float(s[3 + rest + k]);
        shv[33 + k] = half_to_float(s[3 + 2 * rest + k]);
      }
    } else {
      for (int k = 0; k < 48; ++k) shv[k] = (T)sh[(size_t)sh_idx * 48 + k];
    }
    const float* rot =
        (P.has_deformation && tet_raw < P.num_tets) ? tet_rot + (size_t)tet_raw * 9 : nullptr;
    T at[10];
    uint32_t ep = 0;
    bool alive = project_one<T>(P, p, cv, (T)opa[idx], shv, rot, at, &ep);
    extents[idx] = alive ? ep : 0u;
    if (alive) {
      for (int k = 0; k < 10; ++k) attrs10[(size_t)idx * 10 + k] = (OutT)at[k];
      survivors += 1;
    }
  }
  return survivors;
}

// ---------------------------------------------------------------------------
// Stage 3: binning.rs
// ---------------------------------------------------------------------------

// ellipse_intersects_tile_gpu (binning.rs:35-83) with the per-splat cutoff
// q_cutoff = clamp(2 ln(255 alpha), 0, 8) hoisted (a pure function of alpha).
// Tiles of tw x th pixels (the reference's TILE_SIZE = 16, tiles/mod.rs:10,
// as a parameter): the tile's pixel-centre rectangle.
template <class T>
static bool ellipse_tile(T mu_x, T mu_y, T a, T b, T c, T alpha, T q_cut, uint32_t tx, uint32_t ty,
                         uint32_t W, uint32_t H, uint32_t tw, uint32_t th) {
  typedef Mth<T> m;
  if (alpha < (T)(1.0f / 255.0f) || W == 0 || H == 0) return false;
  uint32_t xs = tx * tw, ys = ty * th;
  if (xs >= W || ys >= H) return false;
  uint32_t xe = std::min(xs + tw - 1u, W - 1u), ye = std::min(ys + th - 1u, H - 1u);
  T tminx = (T)(float)xs + (T)0.5, tmaxx = (T)(float)xe + (T)0.5;
  T tminy = (T)(float)ys + (T)0.5, tmaxy = (T)(float)ye + (T)0.5;
  T dx0 = tminx - mu_x, dx1 = tmaxx - mu_x, dy0 = tminy - mu_y, dy1 = tmaxy - mu_y;
  T x_near = m::cl((T)0, dx0, dx1);
  T y1 = m::cl(-b * x_near / m::mx(c, (T)1e-12f), dy0, dy1);
  T q1 = a * x_near * x_near + (T)2 * b * x_near * y1 + c * y1 * y1;
  T y_near = m::cl((T)0, dy0, dy1);
  T x2 = m::cl(-b * y_near / m::mx(a, (T)1e-12f), dx0, dx1);
  T q2 = a * x2 * x2 + (T)2 * b * x2 * y_near + c * y_near * y_near;
  T min_quad = m::mn(q1, q2);
  return min_quad <= q_cut + (T)1e-4f;
}

struct Bounds {
  uint32_t tx0, tx1, ty0, ty1;
};

// compute_gaussian_tile_bounds_xy (tiles/mod.rs:97-130) with TILE_SIZE_F32
// replaced by the tile width / height
static Bounds tile_bounds(float mu_x, float mu_y, uint32_t packed, uint32_t tiles_x,
                          uint32_t tiles_y, uint32_t tw, uint32_t th) {
  float ext_x = (float)(packed & 0xFFFFu);
  float ext_y = (float)((packed >> 16) & 0xFFFFu);
  int32_t mtx = (int32_t)tiles_x - 1, mty = (int32_t)tiles_y - 1;
  const float fw = (float)tw, fh = (float)th;
  Bounds r;
  r.tx0 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_x - ext_x) / fw)), 0, mtx);
  r.tx1 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_x + ext_x) / fw)), 0, mtx);
  r.ty0 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_y - ext_y) / fh)), 0, mty);
  r.ty1 = (uint32_t)cgs_iclamp(cgs_f2i_sat(floorf((mu_y + ext_y) / fh)), 0, mty);
  return r;
}

static inline float qcut_of(float alpha) {
  return cgs_fclamp(2.0f * cgs_logf(255.0f * alpha), 0.0f, 8.0f);
}

}  // namespace

// ===========================================================================
// Exported C interface (ctypes).  All arrays are in the reference's layouts.
// ===========================================================================

ORC_API void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
ORC_API int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

ORC_API float orc_logf(float x) { return cgs_logf(x); }
ORC_API float orc_expf(float x) { return cgs_expf(x); }
ORC_API float orc_powf(float x, float y) { return cgs_powf_pos(x, y); }
ORC_API float orc_acosf(float x) { return cgs_acosf(x); }
ORC_API float orc_cosf(float x) { return cgs_cosf(x); }
ORC_API float orc_half_to_float(uint16_t h) { return half_to_float(h); }

// Stage 2, FP32 or FP64 (attrs written as float or double).  cov_h / sh_h select
// the f16 storage variant (project.rs:1344-1679) when non-null.
ORC_API uint32_t orc_project_v2(const float* pos, const float* cov, const uint16_t* cov_h,
                                const float* opa, const float* sh, const uint16_t* sh_h,
                                const uint32_t* bind_tet, const float* tet_rot,
                                const uint32_t* split_meta, uint32_t split_count,
                                const float* camera24, const float* cam_pos3, float mip,
                                float min_opacity, float min_radius, uint32_t sh_degree,
                                uint32_t has_deformation, uint32_t num_tets, uint32_t num_rest,
                                uint32_t max_gaussians, int fp64, void* attrs10,
                                uint32_t* extents) {
  ProjParams P;
  P.camera = camera24;
  P.cam_pos[0] = cam_pos3[0];
  P.cam_pos[1] = cam_pos3[1];
  P.cam_pos[2] = cam_pos3[2];
  P.mip = mip;
  P.min_opacity = min_opacity;
  P.min_radius = min_radius;
  P.sh_degree = sh_degree > 3 ? 3 : sh_degree;
  P.has_deformation = has_deformation;
  P.num_tets = num_tets;
  P.num_rest = num_rest;
  if (fp64)
    return project_all<double, double>(P, pos, cov, cov_h, opa, sh, sh_h, bind_tet, tet_rot,
                                       split_meta, split_count, (double*)attrs10, extents,
                                       max_gaussians);
  return project_all<float, float>(P, pos, cov, cov_h, opa, sh, sh_h, bind_tet, tet_rot,
                                   split_meta, split_count, (float*)attrs10, extents,
                                   max_gaussians);
}

// compute_packed_key_bits (binning.rs:12-20)
ORC_API void orc_packed_key_bits(uint32_t total_tiles, uint32_t* tile_bits, uint32_t* depth_bits) {
  uint32_t tb = total_tiles <= 1 ? 1u : 32u - (uint32_t)__builtin_clz(total_tiles - 1u);
  uint32_t db = 32u - tb;
  if (db > 28u) db = 28u;
  *tile_bits = tb;
  *depth_bits = db;
}

// count_tiles_kernel (binning.rs:90-164)
ORC_API void orc_count_tiles(const float* attrs10, const uint32_t* extents, uint32_t active,
                             uint32_t max_gaussians, uint32_t W, uint32_t H, uint32_t tiles_x,
                             uint32_t tiles_y, uint32_t tile_w, uint32_t tile_h,
                             uint32_t* counts) {
#pragma omp parallel for schedule(static, 2048) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)max_gaussians; ++ii) {
    uint32_t idx = (uint32_t)ii;
    if (idx >= active || tiles_x == 0 || tiles_y == 0 || extents[idx] == 0) {
      counts[idx] = 0;
      continue;
    }
    const float* a = attrs10 + (size_t)idx * 10;
    Bounds bd = tile_bounds(a[0], a[1], extents[idx], tiles_x, tiles_y, tile_w, tile_h);
    float qc = qcut_of(a[5]);
    uint32_t cnt = 0;
    for (uint32_t ty = bd.ty0; ty <= bd.ty1; ++ty)
      for (uint32_t tx = bd.tx0; tx <= bd.tx1; ++tx)
        if (ellipse_tile<float>(a[0], a[1], a[2], a[3], a[4], a[5], qc, tx, ty, W, H, tile_w,
                                tile_h)) cnt++;
    counts[idx] = cnt;
  }
}

// gpu_exclusive_scan_total (scan.rs:198-333) == serial exclusive prefix sum
// (tiles/cpu_reference.rs:100-115); u32 wrap-around as on the GPU.
ORC_API uint32_t orc_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n) {
  uint32_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t v = in[i];
    out[i] = acc;
    acc += v;
  }
  return acc;
}

// emit_tile_keys_kernel (binning.rs:168-250)
ORC_API void orc_emit_keys32(const float* attrs10, const uint32_t* extents,
                             const uint32_t* offsets, uint32_t active, uint32_t W, uint32_t H,
                             uint32_t tiles_x, uint32_t tiles_y, uint32_t tile_w,
                             uint32_t tile_h, uint32_t base_depth_bits, uint32_t depth_bits,
                             uint32_t* keys, uint32_t* ids) {
#pragma omp parallel for schedule(static, 2048) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)active; ++ii) {
    uint32_t idx = (uint32_t)ii;
    if (tiles_x == 0 || tiles_y == 0 || extents[idx] == 0) continue;
    const float* a = attrs10 + (size_t)idx * 10;
    uint32_t raw = cgs_f2u(cgs_fmax(a[9], 1e-6f));
    uint32_t rel = raw > base_depth_bits ? std::min(raw - base_depth_bits, 0x0FFFFFFFu) : 0u;
    uint32_t dq = rel >> (28u - depth_bits);
    Bounds bd = tile_bounds(a[0], a[1], extents[idx], tiles_x, tiles_y, tile_w, tile_h);
    float qc = qcut_of(a[5]);
    uint32_t cur = offsets[idx];
    for (uint32_t ty = bd.ty0; ty <= bd.ty1; ++ty)
      for (uint32_t tx = bd.tx0; tx <= bd.tx1; ++tx)
        if (ellipse_tile<float>(a[0], a[1], a[2], a[3], a[4], a[5], qc, tx, ty, W, H, tile_w,
                                tile_h)) {
          uint32_t tile = ty * tiles_x + tx;
          keys[cur] = (tile << depth_bits) | dq;
          ids[cur] = idx;
          cur++;
        }
  }
}

// emit_tile_keys_u64_kernel (binning.rs:254-328)
ORC_API void orc_emit_keys64(const float* attrs10, const uint32_t* extents,
                             const uint32_t* offsets, uint32_t active, uint32_t W, uint32_t H,
                             uint32_t tiles_x, uint32_t tiles_y, uint32_t tile_w,
                             uint32_t tile_h, uint32_t* tile_ids, uint32_t* depths,
                             uint32_t* ids) {
#pragma omp parallel for schedule(static, 2048) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)active; ++ii) {
    uint32_t idx = (uint32_t)ii;
    if (tiles_x == 0 || tiles_y == 0 || extents[idx] == 0) continue;
    const float* a = attrs10 + (size_t)idx * 10;
    uint32_t dbits = cgs_f2u(cgs_fmax(a[9], 0.0f));
    Bounds bd = tile_bounds(a[0], a[1], extents[idx], tiles_x, tiles_y, tile_w, tile_h);
    float qc = qcut_of(a[5]);
    uint32_t cur = offsets[idx];
    for (uint32_t ty = bd.ty0; ty <= bd.ty1; ++ty)
      for (uint32_t tx = bd.tx0; tx <= bd.tx1; ++tx)
        if (ellipse_tile<float>(a[0], a[1], a[2], a[3], a[4], a[5], qc, tx, ty, W, H, tile_w,
                                tile_h)) {
          tile_ids[cur] = ty * tiles_x + tx;
          depths[cur] = dbits;
          ids[cur] = idx;
          cur++;
        }
  }
}

// gpu_radix_sort_u32_pairs (radix_sort.rs:292-368): a stable LSD radix sort
// over `bits` key bits == a stable sort by (key & mask).
ORC_API void orc_sort_pairs32(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t bits) {
  uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
  std::vector<uint64_t> tmp(n);
  for (uint64_t i = 0; i < n; ++i) tmp[i] = ((uint64_t)(keys[i] & mask) << 32) | i;
  std::sort(tmp.begin(), tmp.end());  // (key, original index) total order == stable
  std::vector<uint32_t> k2(n), v2(n);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t src = tmp[i] & 0xFFFFFFFFull;
    k2[i] = keys[src];
    v2[i] = vals[src];
  }
  if (n) {
    memcpy(keys, k2.data(), n * 4);
    memcpy(vals, v2.data(), n * 4);
  }
}

// gpu_radix_sort_u64_pairs (radix_sort.rs:377-477): stable by depth, then
// stable by tile == stable sort by (tile & tile_mask, depth).
ORC_API void orc_sort64(uint32_t* tile_ids, uint32_t* depths, uint32_t* ids, uint64_t n,
                        uint32_t tile_bits) {
  uint32_t tmask = tile_bits >= 32 ? 0xFFFFFFFFu : ((1u << tile_bits) - 1u);
  std::vector<uint64_t> order(n);
  std::iota(order.begin(), order.end(), 0ull);
  std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) {
    uint32_t tx = tile_ids[x] & tmask, ty = tile_ids[y] & tmask;
    if (tx != ty) return tx < ty;
    return depths[x] < depths[y];
  });
  std::vector<uint32_t> t2(n), d2(n), i2(n);
  for (uint64_t i = 0; i < n; ++i) {
    t2[i] = tile_ids[order[i]];
    d2[i] = depths[order[i]];
    i2[i] = ids[order[i]];
  }
  if (n) {
    memcpy(tile_ids, t2.data(), n * 4);
    memcpy(depths, d2.data(), n * 4);
    memcpy(ids, i2.data(), n * 4);
  }
}

// identify_tile_ranges_kernel (binning.rs:344-379); tile of entry i =
// keys[i] >> shift (shift = depth_bits, or 0 for sorted tile ids).
ORC_API void orc_tile_ranges(const uint32_t* sorted, uint64_t n, uint32_t shift,
                             uint32_t total_tiles, uint32_t* ranges) {
  memset(ranges, 0, (size_t)total_tiles * 8);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t t = sorted[i] >> shift;
    if (t >= total_tiles) continue;
    if (i == 0) {
      ranges[t * 2] = 0;
    } else {
      uint32_t pt = sorted[i - 1] >> shift;
      if (t != pt) {
        if (pt < total_tiles) ranges[pt * 2 + 1] = (uint32_t)i;
        ranges[t * 2] = (uint32_t)i;
      }
    }
    if (i + 1 == n) ranges[t * 2 + 1] = (uint32_t)n;
  }
}

// ---------------------------------------------------------------------------
// Stage 4: tile_compositor_v2_kernel (compositor.rs:721-883).  The per-pixel
// result is the sequential front-to-back blend over the tile's range; the
// batching and tile-wide early exit of the kernel do not change it.
// ---------------------------------------------------------------------------
template <class T, class AT>
static void composite_impl(const AT* attrs10, const uint32_t* sorted_ids, const uint32_t* ranges,
                           uint32_t W, uint32_t H, uint32_t tiles_x, uint32_t tile_w,
                           uint32_t tile_h, uint32_t padded_row, const float* bg, uint32_t* out,
                           float* out_rgba_f) {
  typedef Mth<T> m;
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
    uint32_t py = (uint32_t)yy;
    for (uint32_t px = 0; px < W; ++px) {
      uint32_t tile = (py / tile_h) * tiles_x + (px / tile_w);
      uint32_t rs = ranges[tile * 2], re = ranges[tile * 2 + 1];
      bool empty = rs >= re;
      T cr = 0, cg = 0, cb = 0, t = 1;
      T pcx = (T)(float)px + (T)0.5, pcy = (T)(float)py + (T)0.5;
      if (!empty) {
        for (uint32_t e = rs; e < re; ++e) {
          const AT* a = attrs10 + (size_t)sorted_ids[e] * 10;
          T op = (T)a[5];
          if (!(op > (T)0.003921569f)) continue;
          T dx = pcx - (T)a[0], dy = pcy - (T)a[1];
          T q = (T)-0.5 * ((T)a[2] * dx * dx + (T)2 * (T)a[3] * dx * dy + (T)a[4] * dy * dy);
          if (q > (T)0) q = 0;
          if (q >= (T)-4) {
            T alpha = m::mn((T)0.99f, op * m::cex(q));
            if (alpha >= (T)0.003921569f) {
              T w = alpha * t;
              cr += w * (T)a[6];
              cg += w * (T)a[7];
              cb += w * (T)a[8];
              t *= (T)1 - alpha;
              if (t < (T)0.001f) {
                t = 0;
                break;
              }
            }
          }
        }
      }
      T fr = m::cl(cr + t * (T)bg[0], (T)0, (T)1);
      T fg = m::cl(cg + t * (T)bg[1], (T)0, (T)1);
      T fb = m::cl(cb + t * (T)bg[2], (T)0, (T)1);
      T fa = m::cl((T)1 - t, (T)0, (T)1);
      if (empty) fa = (T)bg[3];
      if (out_rgba_f) {
        float* o = out_rgba_f + ((size_t)py * W + px) * 4;
        o[0] = (float)fr;
        o[1] = (float)fg;
        o[2] = (float)fb;
        o[3] = (float)fa;
      }
      if (out) {
        uint32_t r8 = (uint32_t)m::cl(m::rd(fr * (T)255), (T)0, (T)255);
        uint32_t g8 = (uint32_t)m::cl(m::rd(fg * (T)255), (T)0, (T)255);
        uint32_t b8 = (uint32_t)m::cl(m::rd(fb * (T)255), (T)0, (T)255);
        uint32_t a8 = (uint32_t)m::cl(m::rd(fa * (T)255), (T)0, (T)255);
        out[(size_t)py * padded_row + px] = r8 | (g8 << 8) | (b8 << 16) | (a8 << 24);
      }
    }
  }
}

// out: RGBA8 words with row stride padded_row (may be null); out_f: unrounded
// RGBA floats, W*H*4 (may be null).  fp64 selects the FP64 twin (attrs must
// then be double).
ORC_API void orc_composite_v2(const void* attrs10, const uint32_t* sorted_ids,
                              const uint32_t* ranges, uint32_t W, uint32_t H, uint32_t tiles_x,
                              uint32_t tile_w, uint32_t tile_h, uint32_t padded_row,
                              const float* bg4, int fp64, uint32_t* out, float* out_f) {
  if (fp64)
    composite_impl<double, double>((const double*)attrs10, sorted_ids, ranges, W, H, tiles_x,
                                   tile_w, tile_h, padded_row, bg4, out, out_f);
  else
    composite_impl<float, float>((const float*)attrs10, sorted_ids, ranges, W, H, tiles_x,
                                 tile_w, tile_h, padded_row, bg4, out, out_f);
}

// ---------------------------------------------------------------------------
// Stage 1: compute_tet_transforms_kernel (deformation.rs:682-865) and
// deform_and_split_compact_kernel (deformation.rs:872-1211), children appended
// in parent order (cpu_reference.rs:316-377 ordering).
// ---------------------------------------------------------------------------
namespace {

template <class T>
static T det3s(T xx, T xy, T xz, T yy, T yz, T zz) {
  return xx * (yy * zz - yz * yz) - xy * (xy * zz - xz * yz) + xz * (xy * yz - xz * yy);
}

template <class T>
static void tet_one(const float* V, const uint32_t* e, const float* M, T J[9], T R[9],
                    T* stretch) {
  typedef Mth<T> m;
  const uint32_t v0 = e[0], v1 = e[1], v2 = e[2], v3 = e[3];
  T x0x = V[v0 * 3], x0y = V[v0 * 3 + 1], x0z = V[v0 * 3 + 2];
  T ds00 = (T)V[v1 * 3] - x0x, ds10 = (T)V[v1 * 3 + 1] - x0y, ds20 = (T)V[v1 * 3 + 2] - x0z;
  T ds01 = (T)V[v2 * 3] - x0x, ds11 = (T)V[v2 * 3 + 1] - x0y, ds21 = (T)V[v2 * 3 + 2] - x0z;
  T ds02 = (T)V[v3 * 3] - x0x, ds12 = (T)V[v3 * 3 + 1] - x0y, ds22 = (T)V[v3 * 3 + 2] - x0z;
  T a00 = M[0], a10 = M[1], a20 = M[2], a01 = M[3], a11 = M[4], a21 = M[5];
  T a02 = M[6], a12 = M[7], a22 = M[8];
  T j00 = ds00 * a00 + ds01 * a10 + ds02 * a20;
  T j10 = ds10 * a00 + ds11 * a10 + ds12 * a20;
  T j20 = ds20 * a00 + ds21 * a10 + ds22 * a20;
  T j01 = ds00 * a01 + ds01 * a11 + ds02 * a21;
  T j11 = ds10 * a01 + ds11 * a11 + ds12 * a21;
  T j21 = ds20 * a01 + ds21 * a11 + ds22 * a21;
  T j02 = ds00 * a02 + ds01 * a12 + ds02 * a22;
  T j12 = ds10 * a02 + ds11 * a12 + ds12 * a22;
  T j22 = ds20 * a02 + ds21 * a12 + ds22 * a22;
  J[0] = j00; J[1] = j10; J[2] = j20; J[3] = j01; J[4] = j11; J[5] = j21;
  J[6] = j02; J[7] = j12; J[8] = j22;
  T c00 = j00 * j00 + j10 * j10 + j20 * j20;
  T c01 = j00 * j01 + j10 * j11 + j20 * j21;
  T c02 = j00 * j02 + j10 * j12 + j20 * j22;
  T c11 = j01 * j01 + j11 * j11 + j21 * j21;
  T c12 = j01 * j02 + j11 * j12 + j21 * j22;
  T c22 = j02 * j02 + j12 * j12 + j22 * j22;
  T p1 = c01 * c01 + c02 * c02 + c12 * c12;
  if (p1 <= (T)1e-10f) {
    *stretch = m::sq(m::mx(m::mx(c00, m::mx(c11, c22)), (T)0));
  } else {
    T q = (c00 + c11 + c22) / (T)3;
    T d0 = c00 - q, d1 = c11 - q, d2 = c22 - q;
    T p2 = d0 * d0 + d1 * d1 + d2 * d2 + (T)2 * p1;
    T p = m::sq(m::mx(p2 / (T)6, (T)1e-12f));
    T ip = (T)1 / p;
    T b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip, b01 = c01 * ip, b02 = c02 * ip, b12 = c12 * ip;
    T det_b = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
              b02 * (b01 * b12 - b11 * b02);
    T r = m::cl((T)0.5 * det_b, (T)-1, (T)1);
    T phi = m::ac(r) / (T)3;
    T emax = q + (T)2 * p * m::co(phi);
    *stretch = m::sq(m::mx(emax, (T)0));
  }
  T det_j = j00 * (j11 * j22 - j12 * j21) - j01 * (j10 * j22 - j12 * j20) +
            j02 * (j10 * j21 - j11 * j20);
  T adet = m::ab(det_j);
  if (adet < (T)1e-8f) {
    for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? (T)1 : (T)0;
    return;
  }
  T gamma = m::ex(-m::ln(adet) / (T)3);
  T r00 = j00 * gamma, r10 = j10 * gamma, r20 = j20 * gamma;
  T r01 = j01 * gamma, r11 = j11 * gamma, r21 = j21 * gamma;
  T r02 = j02 * gamma, r12 = j12 * gamma, r22 = j22 * gamma;
  for (int it = 0; it < 5; ++it) {
    T f00 = r11 * r22 - r21 * r12, f10 = r21 * r02 - r01 * r22, f20 = r01 * r12 - r11 * r02;
    T f01 = r12 * r20 - r22 * r10, f11 = r22 * r00 - r02 * r20, f21 = r02 * r10 - r12 * r00;
    T f02 = r10 * r21 - r20 * r11, f12 = r20 * r01 - r00 * r21, f22 = r00 * r11 - r10 * r01;
    T det_r = r00 * f00 + r10 * f10 + r20 * f20;
    if (m::ab(det_r) > (T)1e-12f) {
      T inv = (T)1 / det_r;
      r00 = (T)0.5 * (r00 + f00 * inv);
      r10 = (T)0.5 * (r10 + f10 * inv);
      r20 = (T)0.5 * (r20 + f20 * inv);
      r01 = (T)0.5 * (r01 + f01 * inv);
      r11 = (T)0.5 * (r11 + f11 * inv);
      r21 = (T)0.5 * (r21 + f21 * inv);
      r02 = (T)0.5 * (r02 + f02 * inv);
      r12 = (T)0.5 * (r12 + f12 * inv);
      r22 = (T)0.5 * (r22 + f22 * inv);
    }
  }
  T det_f = r00 * (r11 * r22 - r21 * r12) + r10 * (r21 * r02 - r01 * r22) +
            r20 * (r01 * r12 - r11 * r02);
  if (det_f < (T)0) {
    r02 = -r02;
    r12 = -r12;
    r22 = -r22;
  }
  R[0] = r00; R[1] = r10; R[2] = r20; R[3] = r01; R[4] = r11; R[5] = r21;
  R[6] = r02; R[7] = r12; R[8] = r22;
}

template <class T>
struct Splat10 {
  T x, y, z, c[6], o;
};

// One splat of deform_and_split_compact_kernel.  Returns true when it splits;
// `a` receives the splat (or first child), `b` the second child.
template <class T>
static bool deform_one(const float* V, const uint32_t* elems, const T* jac, const T* stretch,
                       uint32_t tet_raw, const float* w4, const float* pos, const float* cov,
                       float opa, float tau, float mip, uint32_t num_tets, Splat10<T>& a,
                       Splat10<T>& b) {
  typedef Mth<T> m;
  T s_xx = cov[0], s_xy = cov[1], s_xz = cov[2], s_yy = cov[3], s_yz = cov[4], s_zz = cov[5];
  T alpha = opa;
  T mp = (T)mip;
  if (tet_raw >= num_tets) {
    a.x = pos[0];
    a.y = pos[1];
    a.z = pos[2];
    if (mp > (T)0) {
      T det_orig = m::mx(det3s(s_xx, s_xy, s_xz, s_yy, s_yz, s_zz), (T)1e-24f);
      T r_xx = s_xx + mp, r_yy = s_yy + mp, r_zz = s_zz + mp;
      T det_reg = m::mx(det3s(r_xx, s_xy, s_xz, r_yy, s_yz, r_zz), det_orig);
      T comp = m::sq(m::cl(det_orig / det_reg, (T)0, (T)1));
      T cc[6] = {r_xx, s_xy, s_xz, r_yy, s_yz, r_zz};
      for (int k = 0; k < 6; ++k) a.c[k] = cc[k];
      a.o = m::cl(alpha * comp, (T)0, (T)1);
    } else {
      T cc[6] = {s_xx, s_xy, s_xz, s_yy, s_yz, s_zz};
      for (int k = 0; k < 6; ++k) a.c[k] = cc[k];
      a.o = alpha;
    }
    return false;
  }
  T w0 = w4[0], w1 = w4[1], w2 = w4[2], w3 = w4[3];
  const uint32_t* E = elems + (size_t)tet_raw * 4;
  uint32_t v0 = E[0], v1 = E[1], v2 = E[2], v3 = E[3];
  T px = w0 * (T)V[v0 * 3] + w1 * (T)V[v1 * 3] + w2 * (T)V[v2 * 3] + w3 * (T)V[v3 * 3];
  T py = w0 * (T)V[v0 * 3 + 1] + w1 * (T)V[v1 * 3 + 1] + w2 * (T)V[v2 * 3 + 1] + w3 * (T)V[v3 * 3 + 1];
  T pz = w0 * (T)V[v0 * 3 + 2] + w1 * (T)V[v1 * 3 + 2] + w2 * (T)V[v2 * 3 + 2] + w3 * (T)V[v3 * 3 + 2];
  const T* J = jac + (size_t)tet_raw * 9;
  T j00 = J[0], j10 = J[1], j20 = J[2], j01 = J[3], j11 = J[4], j21 = J[5];
  T j02 = J[6], j12 = J[7], j22 = J[8];
  T k00 = j00 * s_xx + j01 * s_xy + j02 * s_xz;
  T k01 = j00 * s_xy + j01 * s_yy + j02 * s_yz;
  T k02 = j00 * s_xz + j01 * s_yz + j02 * s_zz;
  T k10 = j10 * s_xx + j11 * s_xy + j12 * s_xz;
  T k11 = j10 * s_xy + j11 * s_yy + j12 * s_yz;
  T k12 = j10 * s_xz + j11 * s_yz + j12 * s_zz;
  T k20 = j20 * s_xx + j21 * s_xy + j22 * s_xz;
  T k21 = j20 * s_xy + j21 * s_yy + j22 * s_yz;
  T k22 = j20 * s_xz + j21 * s_yz + j22 * s_zz;
  T m00 = m::mx((T)1e-4f, k00 * j00 + k01 * j01 + k02 * j02);
  T m01 = (T)0.5 * ((k00 * j10 + k01 * j11 + k02 * j12) + (k10 * j00 + k11 * j01 + k12 * j02));
  T m02 = (T)0.5 * ((k00 * j20 + k01 * j21 + k02 * j22) + (k20 * j00 + k21 * j01 + k22 * j02));
  T m11 = m::mx((T)1e-4f, k10 * j10 + k11 * j11 + k12 * j12);
  T m12 = (T)0.5 * ((k10 * j20 + k11 * j21 + k12 * j22) + (k20 * j10 + k21 * j11 + k22 * j12));
  T m22 = m::mx((T)1e-4f, k20 * j20 + k21 * j21 + k22 * j22);
  T det_j = j00 * (j11 * j22 - j12 * j21) - j01 * (j10 * j22 - j12 * j20) +
            j02 * (j10 * j21 - j11 * j20);
  T s_det = m::cl(det_j, (T)0.6f, (T)1.8f);
  T safe_base = m::cl((T)1 - alpha, (T)1e-4f, (T)0.999f);
  T alpha_p = m::cl((T)1 - m::pw(safe_base, (T)1 / s_det), (T)0.05f, (T)0.99f);
  bool split = tau > 0.0f && stretch[tet_raw] > (T)tau;
  if (!split) {
    a.x = px;
    a.y = py;
    a.z = pz;
    if (mp > (T)0) {
      T det_orig = m::mx(det3s(m00, m01, m02, m11, m12, m22), (T)1e-24f);
      T r00 = m00 + mp, r11 = m11 + mp, r22 = m22 + mp;
      T det_reg = m::mx(det3s(r00, m01, m02, r11, m12, r22), det_orig);
      T comp = m::sq(m::cl(det_orig / det_reg, (T)0, (T)1));
      T cc[6] = {r00, m01, m02, r11, m12, r22};
      for (int k = 0; k < 6; ++k) a.c[k] = cc[k];
      a.o = m::cl(alpha_p * comp, (T)0, (T)1);
    } else {
      T cc[6] = {m00, m01, m02, m11, m12, m22};
      for (int k = 0; k < 6; ++k) a.c[k] = cc[k];
      a.o = alpha_p;
    }
    return false;
  }
  T p1 = m01 * m01 + m02 * m02 + m12 * m12;
  T lambda0 = m::mx(m00, m::mx(m11, m22));
  T ux = 1, uy = 0, uz = 0;
  if (p1 <= (T)1e-10f) {
    if (m11 >= m00 && m11 >= m22) {
      ux = 0; uy = 1; uz = 0;
    } else if (m22 >= m00 && m22 >= m11) {
      ux = 0; uy = 0; uz = 1;
    }
  } else {
    T q = (m00 + m11 + m22) / (T)3;
    T d0 = m00 - q, d1 = m11 - q, d2 = m22 - q;
    T p2 = d0 * d0 + d1 * d1 + d2 * d2 + (T)2 * p1;
    T p = m::sq(m::mx(p2 / (T)6, (T)1e-12f));
    T ip = (T)1 / p;
    T b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip, b01 = m01 * ip, b02 = m02 * ip, b12 = m12 * ip;
    T det_b = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) +
              b02 * (b01 * b12 - b11 * b02);
    T r = m::cl((T)0.5 * det_b, (T)-1, (T)1);
    T phi = m::ac(r) / (T)3;
    lambda0 = q + (T)2 * p * m::co(phi);
    T r0x = m00 - lambda0, r0y = m01, r0z = m02;
    T r1x = m01, r1y = m11 - lambda0, r1z = m12;
    T r2x = m02, r2y = m12, r2z = m22 - lambda0;
    T c01x = r0y * r1z - r0z * r1y, c01y = r0z * r1x - r0x * r1z, c01z = r0x * r1y - r0y * r1x;
    T n01 = c01x * c01x + c01y * c01y + c01z * c01z;
    T c12x = r1y * r2z - r1z * r2y, c12y = r1z * r2x - r1x * r2z, c12z = r1x * r2y - r1y * r2x;
    T n12 = c12x * c12x + c12y * c12y + c12z * c12z;
    T c02x = r0y * r2z - r0z * r2y, c02y = r0z * r2x - r0x * r2z, c02z = r0x * r2y - r0y * r2x;
    T n02 = c02x * c02x + c02y * c02y + c02z * c02z;
    if (n01 >= n12 && n01 >= n02 && n01 > (T)1e-12f) {
      T inv = (T)1 / m::sq(n01);
      ux = c01x * inv; uy = c01y * inv; uz = c01z * inv;
    } else if (n12 >= n02 && n12 > (T)1e-12f) {
      T inv = (T)1 / m::sq(n12);
      ux = c12x * inv; uy = c12y * inv; uz = c12z * inv;
    } else if (n02 > (T)1e-12f) {
      T inv = (T)1 / m::sq(n02);
      ux = c02x * inv; uy = c02y * inv; uz = c02z * inv;
    }
  }
  T safe_l0 = m::mx(lambda0, (T)1e-8f);
  T mag = (T)0.5 * m::sq(safe_l0);
  T dx = ux * mag, dy = uy * mag, dz = uz * mag;
  T shrink = (T)0.75 * safe_l0;
  T c00 = m::mx((T)1e-8f, m00 - shrink * ux * ux);
  T c01 = m01 - shrink * ux * uy;
  T c02 = m02 - shrink * ux * uz;
  T c11 = m::mx((T)1e-8f, m11 - shrink * uy * uy);
  T c12 = m12 - shrink * uy * uz;
  T c22 = m::mx((T)1e-8f, m22 - shrink * uz * uz);
  T child_alpha = alpha_p;
  if (mp > (T)0) {
    T det_orig = m::mx(det3s(c00, c01, c02, c11, c12, c22), (T)1e-24f);
    c00 += mp;
    c11 += mp;
    c22 += mp;
    T det_reg = m::mx(det3s(c00, c01, c02, c11, c12, c22), det_orig);
    T comp = m::sq(m::cl(det_orig / det_reg, (T)0, (T)1));
    child_alpha = m::cl(child_alpha * comp, (T)0, (T)1);
  }
  T cc[6] = {c00, c01, c02, c11, c12, c22};
  a.x = px - dx; a.y = py - dy; a.z = pz - dz;
  b.x = px + dx; b.y = py + dy; b.z = pz + dz;
  for (int k = 0; k < 6; ++k) a.c[k] = b.c[k] = cc[k];
  a.o = b.o = child_alpha;
  return true;
}

template <class T>
static uint32_t deform_all(const float* V, const uint32_t* elems, const float* inv_dm,
                           uint32_t num_tets, const uint32_t* bind_tet, const float* bind_w,
                           const float* pos, const float* cov, const float* opa, uint32_t n,
                           float tau, float mip, T* J, T* R, T* stretch, T* out_pos, T* out_cov,
                           T* out_opa, uint32_t* split_meta) {
#pragma omp parallel for schedule(static, 64) num_threads(g_threads)
  for (int64_t t = 0; t < (int64_t)num_tets; ++t)
    tet_one<T>(V, elems + t * 4, inv_dm + t * 9, J + t * 9, R + t * 9, stretch + t);
  std::vector<uint8_t> flag(n);
  std::vector<Splat10<T>> child(n);
#pragma omp parallel for schedule(static, 1024) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    Splat10<T> a, b;
    bool sp = deform_one<T>(V, elems, J, stretch, bind_tet[ii], bind_w + ii * 4, pos + ii * 3,
                            cov + ii * 6, opa[ii], tau, mip, num_tets, a, b);
    out_pos[ii * 3] = a.x;
    out_pos[ii * 3 + 1] = a.y;
    out_pos[ii * 3 + 2] = a.z;
    for (int k = 0; k < 6; ++k) out_cov[ii * 6 + k] = a.c[k];
    out_opa[ii] = a.o;
    flag[ii] = sp ? 1 : 0;
    if (sp) child[ii] = b;
  }
  uint32_t splits = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!flag[i]) continue;
    const size_t c = (size_t)n + splits;
    out_pos[c * 3] = child[i].x;
    out_pos[c * 3 + 1] = child[i].y;
    out_pos[c * 3 + 2] = child[i].z;
    for (int k = 0; k < 6; ++k) out_cov[c * 6 + k] = child[i].c[k];
    out_opa[c] = child[i].o;
    split_meta[splits * 2] = i;
    split_meta[splits * 2 + 1] = bind_tet[i];
    splits++;
  }
  return splits;
}

}  // namespace

// Stage 1.  Outputs in the reference layouts, sized for 2n splats (positions
// [3*2n], covariances [6*2n], opacities [2n], split_meta [2n]); J/R/stretch
// per tet.  fp64 selects the FP64 twin (all float outputs are then double).
// Returns the number of split children.
ORC_API uint32_t orc_deform(const float* verts, const uint32_t* elems, const float* inv_dm,
                            uint32_t num_tets, const uint32_t* bind_tet, const float* bind_w,
                            const float* pos, const float* cov, const float* opa, uint32_t n,
                            float tau, float mip, int fp64, void* J, void* R, void* stretch,
                            void* out_pos, void* out_cov, void* out_opa, uint32_t* split_meta) {
  if (fp64)
    return deform_all<double>(verts, elems, inv_dm, num_tets, bind_tet, bind_w, pos, cov, opa, n,
                              tau, mip, (double*)J, (double*)R, (double*)stretch,
                              (double*)out_pos, (double*)out_cov, (double*)out_opa, split_meta);
  return deform_all<float>(verts, elems, inv_dm, num_tets, bind_tet, bind_w, pos, cov, opa, n, tau,
                           mip, (float*)J, (float*)R, (float*)stretch, (float*)out_pos,
                           (float*)out_cov, (float*)out_opa, split_meta);
}

// ---------------------------------------------------------------------------
// Cage construction and splat binding (src/deformation/mod.rs)
// ---------------------------------------------------------------------------

namespace {

// glam 0.28 Vec3::cross / dot (scalar f32, operation order of the crate)
struct V3 {
  float x, y, z;
};
static inline V3 v3_sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 v3_cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
static inline float v3_dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// One point against one tet: bind_point's barycentric weights
// [w1 w2 w3] = inv_dm (p - X0) (glam Mat3 * Vec3: x_axis*v.x + y_axis*v.y +
// z_axis*v.z, left to right), w0 = 1 - (w1 + w2 + w3), and the penalty
// sum of max(-w_k, 0) (mod.rs:280-293).
static inline float tet_weights(const float* X0, const float* m, const float p[3], float w[4]) {
  const float dx = p[0] - X0[0], dy = p[1] - X0[1], dz = p[2] - X0[2];
  w[1] = (m[0] * dx + m[3] * dy) + m[6] * dz;
  w[2] = (m[1] * dx + m[4] * dy) + m[7] * dz;
  w[3] = (m[2] * dx + m[5] * dy) + m[8] * dz;
  w[0] = 1.0f - ((w[1] + w[2]) + w[3]);
  return ((cgs_fmax(-w[0], 0.0f) + cgs_fmax(-w[1], 0.0f)) + cgs_fmax(-w[2], 0.0f)) +
         cgs_fmax(-w[3], 0.0f);
}

}  // namespace

// TetMesh::new (mod.rs:202-227): D_m = [X1-X0, X2-X0, X3-X0] (columns),
// inverse by glam's Mat3::inverse (cofactors from column cross products,
// det = z . (x cross y), scaled by 1/det, transposed) when |det| > 1e-8, else
// identity; written column-major (to_cols_array).
ORC_API void orc_tet_inv_dm(const float* verts, const uint32_t* elems, uint32_t num_tets,
                            float* inv_dm) {
  for (uint32_t t = 0; t < num_tets; ++t) {
    const float* a = verts + (size_t)elems[t * 4 + 0] * 3;
    const float* b = verts + (size_t)elems[t * 4 + 1] * 3;
    const float* c = verts + (size_t)elems[t * 4 + 2] * 3;
    const float* d = verts + (size_t)elems[t * 4 + 3] * 3;
    const V3 x0 = {a[0], a[1], a[2]};
    const V3 cx = v3_sub({b[0], b[1], b[2]}, x0);
    const V3 cy = v3_sub({c[0], c[1], c[2]}, x0);
    const V3 cz = v3_sub({d[0], d[1], d[2]}, x0);
    float* o = inv_dm + (size_t)t * 9;
    const float det = v3_dot(cz, v3_cross(cx, cy));
    if (!(fabsf(det) > 1e-8f)) {
      for (int k = 0; k < 9; ++k) o[k] = (k % 4 == 0) ? 1.0f : 0.0f;
      continue;
    }
    const V3 t0 = v3_cross(cy, cz), t1 = v3_cross(cz, cx), t2 = v3_cross(cx, cy);
    const float det2 = v3_dot(cz, t2);
    const float inv = 1.0f / det2;
    // from_cols(t0*inv, t1*inv, t2*inv).transpose(): column j of the result
    // holds component j of t0, t1, t2
    const V3 r0 = {t0.x * inv, t0.y * inv, t0.z * inv};
    const V3 r1 = {t1.x * inv, t1.y * inv, t1.z * inv};
    const V3 r2 = {t2.x * inv, t2.y * inv, t2.z * inv};
    o[0] = r0.x; o[1] = r1.x; o[2] = r2.x;
    o[3] = r0.y; o[4] = r1.y; o[5] = r2.y;
    o[6] = r0.z; o[7] = r1.z; o[8] = r2.z;
  }
}

// TetMesh::bind_point (mod.rs:267-310) for every point, brute force over the
// tets in index order: the first tet with penalty <= 1e-4 wins, else the tet
// of smallest penalty (first on ties); with bounds (min xyz, max xyz) a point
// outside them, or whose best penalty exceeds 0.15, is unbound (tet = ~0,
// weights 0) (bind_gaussians_with_bounds, 315-323).
ORC_API void orc_bind_tets(const float* verts, const uint32_t* elems, const float* inv_dm,
                           uint32_t num_tets, const float* pos, uint64_t n, const float* bounds,
                           uint32_t* out_tet, float* out_w) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const float p[3] = {pos[ii * 3 + 0], pos[ii * 3 + 1], pos[ii * 3 + 2]};
    float* ow = out_w + ii * 4;
    if (bounds && (p[0] < bounds[0] || p[0] > bounds[3] || p[1] < bounds[1] ||
                   p[1] > bounds[4] || p[2] < bounds[2] || p[2] > bounds[5])) {
      out_tet[ii] = 0xFFFFFFFFu;
      ow[0] = ow[1] = ow[2] = ow[3] = 0.0f;
      continue;
    }
    uint32_t best_t = 0;
    float best_pen = INFINITY;
    float best_w[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    bool found = false;
    for (uint32_t t = 0; t < num_tets; ++t) {
      float w[4];
      const float pen = tet_weights(verts + (size_t)elems[t * 4] * 3, inv_dm + (size_t)t * 9, p, w);
      if (pen <= 1e-4f) {
        out_tet[ii] = t;
        for (int k = 0; k < 4; ++k) ow[k] = w[k];
        found = true;
        break;
      }
      if (pen < best_pen) {
        best_pen = pen;
        best_t = t;
        for (int k = 0; k < 4; ++k) best_w[k] = w[k];
      }
    }
    if (found) continue;
    if (bounds && best_pen > 0.15f) {
      out_tet[ii] = 0xFFFFFFFFu;
      ow[0] = ow[1] = ow[2] = ow[3] = 0.0f;
    } else {
      out_tet[ii] = best_t;
      for (int k = 0; k < 4; ++k) ow[k] = best_w[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Mean value coordinates on a closed triangle cage (src/deformation/mvc.rs)
// ---------------------------------------------------------------------------

namespace {

template <class T> struct Tr;
template <> struct Tr<float> {
  static float asin_(float x) { return asinf(x); }
  static float sin_(float x) { return sinf(x); }
  static float atan2_(float y, float x) { return atan2f(y, x); }
  static float sqrt_(float x) { return sqrtf(x); }
};
template <> struct Tr<double> {
  static double asin_(double x) { return asin(x); }
  static double sin_(double x) { return sin(x); }
  static double atan2_(double y, double x) { return atan2(y, x); }
  static double sqrt_(double x) { return sqrt(x); }
};

template <class T> struct W3 {
  T x, y, z;
};
template <class T> static inline W3<T> w_sub(W3<T> a, W3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> static inline T w_dot(W3<T> a, W3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> static inline W3<T> w_cross(W3<T> a, W3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> static inline T w_len(W3<T> a) { return Tr<T>::sqrt_(w_dot(a, a)); }
template <class T> static inline W3<T> w_at(const float* v, uint32_t i) {
  return {(T)v[i * 3], (T)v[i * 3 + 1], (T)v[i * 3 + 2]};
}
template <class T> static inline T w_max(T a, T b) { return a > b ? a : (b != b ? a : b); }
template <class T> static inline T w_clamp(T x, T lo, T hi) { return x < lo ? lo : (x > hi ? hi : x); }

// PolyCageMesh::winding_number (mvc.rs:326-347)
template <class T>
static T winding_number(const float* V, const uint32_t* F, uint32_t nt, W3<T> p) {
  T sum = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const W3<T> a = w_sub(w_at<T>(V, F[t * 3]), p), b = w_sub(w_at<T>(V, F[t * 3 + 1]), p),
                c = w_sub(w_at<T>(V, F[t * 3 + 2]), p);
    const T la = w_len(a), lb = w_len(b), lc = w_len(c);
    if (la < (T)1e-7f || lb < (T)1e-7f || lc < (T)1e-7f) return 1;
    const T det = w_dot(a, w_cross(b, c));
    const T denom = ((la * lb * lc + w_dot(a, b) * lc) + w_dot(b, c) * la) + w_dot(c, a) * lb;
    sum += (T)2 * Tr<T>::atan2_(det, denom);
  }
  return sum / ((T)4 * (T)3.14159265358979323846f);
}

// PolyCageMesh::compute_mvc_weights (mvc.rs:366-471), operation order of the
// reference (left to right, glam's dot / cross / length / determinant)
template <class T>
static void mvc_weights(const float* V, uint32_t nv, const uint32_t* F, uint32_t nt,
                        const T x[3], T* w, T* dist, W3<T>* dir) {
  typedef Tr<T> m;
  const T PI_ = (T)3.14159265358979323846f;
  const W3<T> X = {x[0], x[1], x[2]};
  for (uint32_t i = 0; i < nv; ++i) {
    const W3<T> diff = w_sub(w_at<T>(V, i), X);
    const T d = w_len(diff);
    if (d < (T)1e-6f) {
      for (uint32_t k = 0; k < nv; ++k) w[k] = 0;
      w[i] = 1;
      return;
    }
    dist[i] = d;
    dir[i] = {diff.x / d, diff.y / d, diff.z / d};
  }
  for (uint32_t k = 0; k < nv; ++k) w[k] = 0;
  T total = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t i0 = F[t * 3], i1 = F[t * 3 + 1], i2 = F[t * 3 + 2];
    const W3<T> u0 = dir[i0], u1 = dir[i1], u2 = dir[i2];
    const T l0 = w_len(w_sub(u1, u2)), l1 = w_len(w_sub(u2, u0)), l2 = w_len(w_sub(u0, u1));
    const T th0 = (T)2 * m::asin_(w_clamp(l0 * (T)0.5, (T)0, (T)1));
    const T th1 = (T)2 * m::asin_(w_clamp(l1 * (T)0.5, (T)0, (T)1));
    const T th2 = (T)2 * m::asin_(w_clamp(l2 * (T)0.5, (T)0, (T)1));
    const T h = (T)0.5 * ((th0 + th1) + th2);
    if (fabs((double)(PI_ - h)) < (double)(T)1e-5f) {
      const T w0 = m::sin_(th0) * dist[i1] * dist[i2];
      const T w1 = m::sin_(th1) * dist[i2] * dist[i0];
      const T w2 = m::sin_(th2) * dist[i0] * dist[i1];
      const T sum = (w0 + w1) + w2;
      for (uint32_t k = 0; k < nv; ++k) w[k] = 0;
      if (fabs((double)sum) > (double)(T)1e-12f) {
        w[i0] = w0 / sum;
        w[i1] = w1 / sum;
        w[i2] = w2 / sum;
      }
      return;
    }
    const T s_t0 = m::sin_(th0), s_t1 = m::sin_(th1), s_t2 = m::sin_(th2);
    if (fabs((double)s_t0) < (double)(T)1e-6f || fabs((double)s_t1) < (double)(T)1e-6f ||
        fabs((double)s_t2) < (double)(T)1e-6f)
      continue;
    const T sin_h = m::sin_(h);
    const T c0 = w_clamp(((T)2 * sin_h * m::sin_(h - th0)) / (s_t1 * s_t2) - (T)1, (T)-1, (T)1);
    const T c1 = w_clamp(((T)2 * sin_h * m::sin_(h - th1)) / (s_t2 * s_t0) - (T)1, (T)-1, (T)1);
    const T c2 = w_clamp(((T)2 * sin_h * m::sin_(h - th2)) / (s_t0 * s_t1) - (T)1, (T)-1, (T)1);
    const T det = w_dot(u2, w_cross(u0, u1));
    const T sgn = det >= (T)0 ? (T)1 : (T)-1;
    const T s0 = sgn * m::sqrt_(w_max((T)1 - c0 * c0, (T)0));
    const T s1 = sgn * m::sqrt_(w_max((T)1 - c1 * c1, (T)0));
    const T s2 = sgn * m::sqrt_(w_max((T)1 - c2 * c2, (T)0));
    if (fabs((double)s0) <= (double)(T)1e-6f || fabs((double)s1) <= (double)(T)1e-6f ||
        fabs((double)s2) <= (double)(T)1e-6f)
      continue;
    const T mu0 = (th0 - c1 * th2 - c2 * th1) / ((T)2 * dist[i0] * s_t1 * s2);
    const T mu1 = (th1 - c2 * th0 - c0 * th2) / ((T)2 * dist[i1] * s_t2 * s0);
    const T mu2 = (th2 - c0 * th1 - c1 * th0) / ((T)2 * dist[i2] * s_t0 * s1);
    w[i0] += mu0;
    w[i1] += mu1;
    w[i2] += mu2;
    total += (mu0 + mu1) + mu2;
  }
  if (fabs((double)total) > (double)(T)1e-12f) {
    const T inv = (T)1 / total;
    for (uint32_t k = 0; k < nv; ++k) w[k] *= inv;
  } else {
    uint32_t best = 0;
    for (uint32_t k = 1; k < nv; ++k)
      if (dist[k] < dist[best]) best = k;
    w[best] = 1;
  }
}

// Far share of the MVC sums for the near/far network: triangle t's terms
// scaled by 1 - phi_t(x), phi_t = smoothstep((r1 - |x - c_t|) / (r1 - r0)),
// c_t = (v0 + v1 + v2) / 3 in f32; normalised by the full sum, times nv
// (sums_scale > 0: the far sums themselves, not normalised, times
// sums_scale).  Points on a vertex or a face, or with a zero sum, give a
// zero row.
template <class T>
static void mvc_far_share(const float* V, uint32_t nv, const uint32_t* F, uint32_t nt, const T x[3],
                          T r0, T r1, T* q, T* dist, W3<T>* dir, T sums_scale = 0) {
  typedef Tr<T> m;
  const T PI_ = (T)3.14159265358979323846f;
  const W3<T> X = {x[0], x[1], x[2]};
  for (uint32_t k = 0; k < nv; ++k) q[k] = 0;
  for (uint32_t i = 0; i < nv; ++i) {
    const W3<T> diff = w_sub(w_at<T>(V, i), X);
    const T d = w_len(diff);
    if (d < (T)1e-6f) return;
    dist[i] = d;
    dir[i] = {diff.x / d, diff.y / d, diff.z / d};
  }
  T total = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t i0 = F[t * 3], i1 = F[t * 3 + 1], i2 = F[t * 3 + 2];
    const W3<T> u0 = dir[i0], u1 = dir[i1], u2 = dir[i2];
    const T th0 = (T)2 * m::asin_(w_clamp(w_len(w_sub(u1, u2)) * (T)0.5, (T)0, (T)1));
    const T th1 = (T)2 * m::asin_(w_clamp(w_len(w_sub(u2, u0)) * (T)0.5, (T)0, (T)1));
    const T th2 = (T)2 * m::asin_(w_clamp(w_len(w_sub(u0, u1)) * (T)0.5, (T)0, (T)1));
    const T h = (T)0.5 * ((th0 + th1) + th2);
    if (fabs((double)(PI_ - h)) < (double)(T)1e-5f) {
      for (uint32_t k = 0; k < nv; ++k) q[k] = 0;
      return;
    }
    const T s_t0 = m::sin_(th0), s_t1 = m::sin_(th1), s_t2 = m::sin_(th2);
    if (fabs((double)s_t0) < (double)(T)1e-6f || fabs((double)s_t1) < (double)(T)1e-6f ||
        fabs((double)s_t2) < (double)(T)1e-6f)
      continue;
    const T sin_h = m::sin_(h);
    const T c0 = w_clamp(((T)2 * sin_h * m::sin_(h - th0)) / (s_t1 * s_t2) - (T)1, (T)-1, (T)1);
    const T c1 = w_clamp(((T)2 * sin_h * m::sin_(h - th1)) / (s_t2 * s_t0) - (T)1, (T)-1, (T)1);
    const T c2 = w_clamp(((T)2 * sin_h * m::sin_(h - th2)) / (s_t0 * s_t1) - (T)1, (T)-1, (T)1);
    const T sgn = w_dot(u2, w_cross(u0, u1)) >= (T)0 ? (T)1 : (T)-1;
    const T s0 = sgn * m::sqrt_(w_max((T)1 - c0 * c0, (T)0));
    const T s1 = sgn * m::sqrt_(w_max((T)1 - c1 * c1, (T)0));
    const T s2 = sgn * m::sqrt_(w_max((T)1 - c2 * c2, (T)0));
    if (fabs((double)s0) <= (double)(T)1e-6f || fabs((double)s1) <= (double)(T)1e-6f ||
        fabs((double)s2) <= (double)(T)1e-6f)
      continue;
    const T mu0 = (th0 - c1 * th2 - c2 * th1) / ((T)2 * dist[i0] * s_t1 * s2);
    const T mu1 = (th1 - c2 * th0 - c0 * th2) / ((T)2 * dist[i1] * s_t2 * s0);
    const T mu2 = (th2 - c0 * th1 - c1 * th0) / ((T)2 * dist[i2] * s_t0 * s1);
    total += (mu0 + mu1) + mu2;
    float cf[3];
    for (int k = 0; k < 3; ++k)
      cf[k] = ((V[(size_t)i0 * 3 + k] + V[(size_t)i1 * 3 + k]) + V[(size_t)i2 * 3 + k]) / 3.0f;
    const W3<T> dc = {X.x - (T)cf[0], X.y - (T)cf[1], X.z - (T)cf[2]};
    const T tt = w_clamp((r1 - w_len(dc)) / (r1 - r0), (T)0, (T)1);
    const T far = (T)1 - tt * tt * ((T)3 - (T)2 * tt);
    q[i0] += far * mu0;
    q[i1] += far * mu1;
    q[i2] += far * mu2;
  }
  if (!(fabs((double)total) > (double)(T)1e-12f)) {
    for (uint32_t k = 0; k < nv; ++k) q[k] = 0;
    return;
  }
  const T scale = sums_scale > (T)0 ? sums_scale : (T)nv / total;
  for (uint32_t k = 0; k < nv; ++k) q[k] *= scale;
}

// jacobian_step_size (mvc.rs:485-495)
template <class T> static T mvc_step(const float* V, uint32_t nv) {
  T lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], (T)V[i * 3 + k]);
      hi[k] = std::max(hi[k], (T)V[i * 3 + k]);
    }
  const W3<T> d = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const T diag = w_max(w_len(d), (T)1e-3f);
  return w_clamp(diag * (T)1e-3f, (T)1e-5f, (T)1e-2f);
}

template <class T>
static void mvc_bind_all(const float* V, uint32_t nv, const uint32_t* F, uint32_t nt,
                         const float* pts, uint64_t n, T* out_w, T* out_g, uint8_t* mask) {
  const T h = mvc_step<T>(V, nv);
  const T inv_2h = (T)0.5 / h;
#pragma omp parallel num_threads(g_threads)
  {
    std::vector<T> dist(nv), wp(nv), wm(nv);
    std::vector<W3<T>> dir(nv);
#pragma omp for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      const T p[3] = {(T)pts[ii * 3], (T)pts[ii * 3 + 1], (T)pts[ii * 3 + 2]};
      T* w = out_w + (size_t)ii * nv;
      T* g = out_g ? out_g + (size_t)ii * nv * 3 : nullptr;
      const T wn = winding_number<T>(V, F, nt, {p[0], p[1], p[2]});
      const bool inside = fabs((double)wn) >= 0.5;
      if (mask) mask[ii] = inside ? 1 : 0;
      if (!inside) {
        for (uint32_t k = 0; k < nv; ++k) w[k] = 0;
        if (g)
          for (uint32_t k = 0; k < nv * 3; ++k) g[k] = 0;
        continue;
      }
      mvc_weights<T>(V, nv, F, nt, p, w, dist.data(), dir.data());
      if (!g) continue;
      for (int axis = 0; axis < 3; ++axis) {
        T pp[3] = {p[0], p[1], p[2]}, pm[3] = {p[0], p[1], p[2]};
        pp[axis] += h;
        pm[axis] -= h;
        mvc_weights<T>(V, nv, F, nt, pp, wp.data(), dist.data(), dir.data());
        mvc_weights<T>(V, nv, F, nt, pm, wm.data(), dist.data(), dir.data());
        for (uint32_t j = 0; j < nv; ++j) g[(size_t)j * 3 + axis] = (wp[j] - wm[j]) * inv_2h;
      }
    }
  }
}

}  // namespace

// PolyCageMesh::bind_gaussians (mvc.rs:498-541): per point the inside mask
// (|winding number| >= 0.5), the MVC weights [n*nv] and their central
// differences [n*nv*3] (step = clamp(1e-3 diag, 1e-5, 1e-2)); zeros outside.
// out_grad may be null (weights only).  fp64 selects the FP64 twin (double
// outputs).
ORC_API void orc_mvc_bind(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                          const float* pts, uint64_t n, int fp64, void* out_w, void* out_grad,
                          uint8_t* mask) {
  if (fp64)
    mvc_bind_all<double>(verts, nv, tris, nt, pts, n, (double*)out_w, (double*)out_grad, mask);
  else
    mvc_bind_all<float>(verts, nv, tris, nt, pts, n, (float*)out_w, (float*)out_grad, mask);
}

// winding number only (contains_point = |w| >= 0.5)
ORC_API void orc_mvc_winding(const float* verts, const uint32_t* tris, uint32_t nt,
                             const float* pts, uint64_t n, int fp64, double* out) {
#pragma omp parallel for schedule(static, 64) num_threads(g_threads)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    if (fp64)
      out[ii] = winding_number<double>(verts, tris, nt,
                                       {(double)pts[ii * 3], (double)pts[ii * 3 + 1],
                                        (double)pts[ii * 3 + 2]});
    else
      out[ii] = winding_number<float>(verts, tris, nt,
                                      {pts[ii * 3], pts[ii * 3 + 1], pts[ii * 3 + 2]});
  }
}

// compute_mvc_weights alone (no inside test), per point: [n*nv]
ORC_API void orc_mvc_weights(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                             const float* pts, uint64_t n, int fp64, void* out_w) {
  for (uint64_t i = 0; i < n; ++i) {
    if (fp64) {
      std::vector<double> dist(nv);
      std::vector<W3<double>> dir(nv);
      const double p[3] = {pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
      mvc_weights<double>(verts, nv, tris, nt, p, (double*)out_w + i * nv, dist.data(), dir.data());
    } else {
      std::vector<float> dist(nv);
      std::vector<W3<float>> dir(nv);
      const float p[3] = {pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
      mvc_weights<float>(verts, nv, tris, nt, p, (float*)out_w + i * nv, dist.data(), dir.data());
    }
  }
}

// ---------------------------------------------------------------------------
// far shares of n points: [n*nv], times nv (fp64: double output)
ORC_API void orc_mvc_far_share(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                               const float* pts, uint64_t n, float r0, float r1, int fp64, void* out) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    if (fp64) {
      std::vector<double> dist(nv);
      std::vector<W3<double>> dir(nv);
      const double p[3] = {pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
      mvc_far_share<double>(verts, nv, tris, nt, p, r0, r1, (double*)out + i * nv, dist.data(), dir.data());
    } else {
      std::vector<float> dist(nv);
      std::vector<W3<float>> dir(nv);
      const float p[3] = {pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
      mvc_far_share<float>(verts, nv, tris, nt, p, r0, r1, (float*)out + i * nv, dist.data(), dir.data());
    }
  }
}

// far sums of n points: [n*nv], the far share's sums before the
// normalisation, times scale (fp64: double output)
ORC_API void orc_mvc_far_sums(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                              const float* pts, uint64_t n, float r0, float r1, float scale,
                              int fp64, void* out) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    if (fp64) {
      std::vector<double> dist(nv);
      std::vector<W3<double>> dir(nv);
      const double p[3] = {pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
      mvc_far_share<double>(verts, nv, tris, nt, p, r0, r1, (double*)out + i * nv, dist.data(),
                            dir.data(), (double)scale);
    } else {
      std::vector<float> dist(nv);
      std::vector<W3<float>> dir(nv);
      const float p[3] = {pts[i * 3], pts[i * 3 + 1], pts[i * 3 + 2]};
      mvc_far_share<float>(verts, nv, tris, nt, p, r0, r1, (float*)out + i * nv, dist.data(),
                           dir.data(), scale);
    }
  }
}

// MVC scene deformation: PolyCageMesh::deform_scene_cpu (mvc.rs:564-749)
// ---------------------------------------------------------------------------
//
// Given each point's deformed position x' and Jacobian F (evaluate_position /
// local_jacobian), restated in the reference's three passes: (1) per point the
// deformed record (mvc.rs:600-667), (2) the split flags (669-697), (3) one
// sequential walk that mip-filters unsplit records and replaces each split
// one by its first child while appending the second (699-740).

ORC_API uint64_t orc_mvc_scene(const float* def_pos, const float* def_F, const uint32_t* mask,
                               const float* rest_pos, const float* rest_cov,
                               const float* rest_opa, const float* rest_sh, uint64_t n,
                               uint32_t sh_degree, float tau, float mip, float* out_pos,
                               float* out_cov, float* out_opa, float* out_sh) {
  if (n == 0) return 0;
  static double tables[CGS_SH_TABLE_DOUBLES];
  cgs_sh_basis_tables(tables);
  const int degree = sh_degree > 3 ? 3 : (int)sh_degree;
  std::vector<float> positions(n * 3), covs(n * 6), opas(n), shs(n * 48);
  for (uint64_t i = 0; i < n; ++i) {
    const bool bound = mask ? mask[i] != 0u : true;
    for (int k = 0; k < 48; ++k) shs[i * 48 + k] = rest_sh[i * 48 + k];
    opas[i] = rest_opa[i];
    if (!bound) {
      for (int k = 0; k < 3; ++k) positions[i * 3 + k] = rest_pos[i * 3 + k];
      for (int k = 0; k < 6; ++k) covs[i * 6 + k] = rest_cov[i * 6 + k];
      continue;
    }
    const float* F = def_F + i * 9;
    float R[9], S[9], FS[9], Ft[9], C[9];
    cgs_polar_rotation(F, R);
    cgs_cov_to_m3(rest_cov + i * 6, S);
    cgs_m3_mul(F, S, FS);
    cgs_m3_transpose(F, Ft);
    cgs_m3_mul(FS, Ft, C);
    for (int k = 0; k < 3; ++k) positions[i * 3 + k] = def_pos[i * 3 + k];
    cgs_m3_to_cov(C, &covs[i * 6]);
    if (degree > 0) cgs_sh_rotate(&shs[i * 48], R, degree, tables);
  }
  std::vector<uint8_t> needs_split;
  if (tau > 0.0f) {
    needs_split.assign(n, 0);
    for (uint64_t i = 0; i < n; ++i) {
      if (mask && mask[i] == 0u) continue;
      needs_split[i] = cgs_max_stretch(def_F + i * 9) > tau ? 1 : 0;
    }
  }
  uint64_t total = n;
  for (uint64_t i = 0; i < n; ++i) {
    const bool split = tau > 0.0f && needs_split[i] != 0;
    float cov[6];
    for (int k = 0; k < 6; ++k) cov[k] = covs[i * 6 + k];
    float alpha = opas[i];
    if (!split) {
      if (mip > 0.0f) {
        cgs_mip_filter_3d(cov, &alpha, mip);
        for (int k = 0; k < 6; ++k) covs[i * 6 + k] = cov[k];
        opas[i] = alpha;
      }
      continue;
    }
    float child[6], off[3];
    cgs_split_major(cov, child, off);
    float cov_a[6], cov_b[6], alpha_a = alpha, alpha_b = alpha;
    for (int k = 0; k < 6; ++k) cov_a[k] = cov_b[k] = child[k];
    if (mip > 0.0f) {
      cgs_mip_filter_3d(cov_a, &alpha_a, mip);
      cgs_mip_filter_3d(cov_b, &alpha_b, mip);
    }
    float center[3] = {positions[i * 3], positions[i * 3 + 1], positions[i * 3 + 2]};
    for (int k = 0; k < 3; ++k) positions[i * 3 + k] = center[k] + (-off[k]);
    for (int k = 0; k < 6; ++k) covs[i * 6 + k] = cov_a[k];
    opas[i] = alpha_a;
    for (int k = 0; k < 3; ++k) positions.push_back(center[k] + off[k]);
    for (int k = 0; k < 6; ++k) covs.push_back(cov_b[k]);
    opas.push_back(alpha_b);
    for (int k = 0; k < 48; ++k) shs.push_back(shs[i * 48 + k]);
    ++total;
  }
  memcpy(out_pos, positions.data(), total * 12);
  memcpy(out_cov, covs.data(), total * 24);
  memcpy(out_opa, opas.data(), total * 4);
  memcpy(out_sh, shs.data(), total * 192);
  return total;
}

// The SH of deform_scene_cpu's output (cpu_reference.rs:141-147, 166-182,
// 368-369): ShRotator::new once per tet (sh_rotation.rs:412-430), applied to
// each bound record's SH (sh_rotation.rs:433-446); unbound records and tets
// out of range keep the rest SH; split child s (split_meta[2s] = parent,
// [2s+1] = its tet) copies the rotated SH of its parent.  rot: [T][9]
// column-major.
ORC_API void orc_deformed_sh(const float* rest_sh, uint32_t n, const float* rot,
                             uint32_t num_tets, const uint32_t* bind_tet,
                             const uint32_t* split_meta, uint32_t splits, int degree,
                             float* out_sh) {
  double t[CGS_SH_TABLE_DOUBLES];
  cgs_sh_basis_tables(t);
  if (degree > 3) degree = 3;
  std::vector<float> mats;
  if (degree > 0) {
    mats.resize((size_t)num_tets * 83);
    for (uint32_t k = 0; k < num_tets; ++k) {
      float* m = &mats[(size_t)k * 83];
      cgs_sh_band_matrix(rot + (size_t)k * 9, 1, t, m);
      if (degree >= 2) cgs_sh_band_matrix(rot + (size_t)k * 9, 2, t, m + 9);
      if (degree >= 3) cgs_sh_band_matrix(rot + (size_t)k * 9, 3, t, m + 34);
    }
  }
  for (uint32_t i = 0; i < n; ++i) {
    float* o = out_sh + (size_t)i * 48;
    memcpy(o, rest_sh + (size_t)i * 48, 192);
    const uint32_t tet = bind_tet[i];
    if (degree == 0 || tet >= num_tets) continue;
    const float* m = &mats[(size_t)tet * 83];
    for (int ch = 0; ch < 3; ++ch) {
      const int rest = 3 + ch * 15;
      cgs_sh_apply_block(o, rest, 3, m);
      if (degree >= 2) cgs_sh_apply_block(o, rest + 3, 5, m + 9);
      if (degree >= 3) cgs_sh_apply_block(o, rest + 8, 7, m + 34);
    }
  }
  for (uint32_t s = 0; s < splits; ++s)
    memcpy(out_sh + (size_t)(n + s) * 48, out_sh + (size_t)split_meta[2 * s] * 48, 192);
}

// the per-splat pieces, for their own checks
ORC_API void orc_sh_tables(double* out) { cgs_sh_basis_tables(out); }
ORC_API void orc_sh_band_matrix(const float* R, int l, float* M) {
  double t[CGS_SH_TABLE_DOUBLES];
  cgs_sh_basis_tables(t);
  cgs_sh_band_matrix(R, l, t, M);
}
ORC_API void orc_sh_rotate(float* sh, const float* R, int degree) {
  double t[CGS_SH_TABLE_DOUBLES];
  cgs_sh_basis_tables(t);
  cgs_sh_rotate(sh, R, degree, t);
}
// the kernel's straight-line form of the same rotation (bit-identical)
ORC_API void orc_sh_rotate_fast(float* sh, const float* R, int degree) {
  cgs_sh_rotate_fast(sh, R, degree);
}
// and of the band matrices
ORC_API void orc_sh_band_matrix_fast(const float* R, int l, float* M) {
  if (l == 1) cgs_sh_band1_matrix(R, M);
  else if (l == 2) cgs_sh_band2_matrix(R, M);
  else cgs_sh_band3_matrix(R, M);
}
ORC_API void orc_polar_rotation(const float* J, float* R) { cgs_polar_rotation(J, R); }
ORC_API void orc_sym_eigen(const float* cov, float* evals, float* basis) {
  cgs_sym_eigen(cov, evals, basis);
}
ORC_API float orc_max_stretch(const float* J) { return cgs_max_stretch(J); }
ORC_API void orc_split_major(const float* cov, float* child_cov, float* offset) {
  cgs_split_major(cov, child_cov, offset);
}
ORC_API void orc_mip_filter_3d(float* cov, float* alpha, float var) {
  cgs_mip_filter_3d(cov, alpha, var);
}
// Stage 2's SH colour (eval_sh, FP64 twin): clamp(0.5 + sum c Y(d), 0, 1)
ORC_API void orc_eval_sh64(const double* coeffs, const double* dir, uint32_t degree,
                           double* out) {
  eval_sh<double>(coeffs, dir[0], dir[1], dir[2], degree, out);
}

// ---------------------------------------------------------------------------
// Scene ingest: parse_vertex_payload + push_raw (ply_loader.rs:540-598,
// soa.rs:113-126), sequentially, record by record
// ---------------------------------------------------------------------------

ORC_API void orc_decode_activate(const uint8_t* payload, uint64_t n, uint32_t stride, int le,
                                 const uint32_t* slot_offset, const uint8_t* slot_type,
                                 float* raw_out, float* pos, float* cov, float* opa, float* sh) {
  for (uint64_t i = 0; i < n; ++i) {
    const uint8_t* rec = payload + i * stride;
    float raw[CGS_RAW_FLOATS];
    for (uint32_t k = 0; k < CGS_RAW_FLOATS; ++k)
      raw[k] = slot_type[k] ? cgs_pt_read(rec + slot_offset[k], slot_type[k], le != 0) : 0.0f;
    if (raw_out) memcpy(raw_out + i * CGS_RAW_FLOATS, raw, sizeof(raw));
    cgs_activate_record(raw, pos + i * 3, cov + i * 6, opa + i, sh + i * 48);
  }
}

ORC_API void orc_quantize(const float* cov, const float* sh, uint64_t n, uint32_t degree,
                          uint32_t* cov_f16, uint32_t* rot, uint32_t* sh_f16) {
  const uint32_t deg = degree > 3u ? 3u : degree, words = cgs_sh_words(deg);
  for (uint64_t i = 0; i < n; ++i)
    cgs_quantize_record(cov + i * 6, sh + i * 48, deg, cov_f16 + i * 3, rot + i,
                        sh_f16 + i * words);
}

ORC_API void orc_export_raw(const float* pos, const float* cov, const float* opa,
                            const float* sh, uint64_t n, float* raw) {
  for (uint64_t i = 0; i < n; ++i)
    cgs_export_record(pos + i * 3, cov + i * 6, opa[i], sh + i * 48, raw + i * CGS_RAW_FLOATS);
}

ORC_API void orc_f32_to_f16(const float* in, uint64_t n, uint16_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = cgs_f32_to_f16(in[i]);
}
ORC_API float orc_atan2f(float y, float x) { return cgs_atan2f(y, x); }
ORC_API void orc_quat_from_m3(const float* m, float* q) { cgs_quat_from_m3(m, q); }
ORC_API uint32_t orc_encode_octahedral(const float* q) { return cgs_encode_octahedral(q); }
ORC_API void orc_decompose_covariance(const float* cov, float* scale, float* q) {
  cgs_decompose_covariance(cov, scale, q);
}
