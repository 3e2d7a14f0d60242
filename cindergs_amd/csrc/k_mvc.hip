// k_mvc.hip — exact 3D mean value coordinates on a closed triangle cage.
//
// Follows PolyCageMesh (src/deformation/mvc.rs): winding_number / contains_point
// (326-363), compute_mvc_weights (366-471, Ju-Schaefer-Warren with the vertex,
// on-face and degenerate-triangle branches), bind_gaussians (498-541: weights
// plus central-difference gradients with the step jacobian_step_size,
// 485-495), evaluate_position (473-481) and MvcBindings::local_jacobian (66-78).
//
// MI355X design: one wave64 per point.  The point's unit directions and
// distances to every cage vertex live in LDS (float4 per vertex); lanes take
// triangles t = lane + 64k and write each triangle's three mu terms to LDS;
// each vertex then sums its incident triangles' terms in triangle order (a
// CSR incidence list built on the host), which is the order in which the
// reference adds them, and the normaliser is a fixed-order wave reduction, so
// the weights are deterministic.  The first vertex closer than 1e-6 and the
// first triangle on which the point lies (|pi - h| < 1e-5) are found with wave
// minima, matching the reference's early returns.  The seven evaluations of a
// bind (the point and p +- h e_axis) run back to back in the same wave.
// Transcendentals are the device library's asinf / sinf / atan2f, so the
// weights agree with the reference to rounding, not bit for bit.
#include "cgs_internal.h"
#include "cgs_scene_math.h"

namespace {

constexpr float MVC_PI = 3.14159265358979323846f;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Sum over the wave with DPP row operations (six adds, then lane 63 read
// to a scalar): quad swaps, half-row and row mirrors, then the row-15 and
// row-31 broadcasts carry the row sums to the last row.  Used where the
// summation order is free (the learned binding); wave_sum keeps the
// butterfly order of the exact kernels.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v += dpp_f32<0xB1>(v);        // quad_perm [1, 0, 3, 2]
  v += dpp_f32<0x4E>(v);        // quad_perm [2, 3, 0, 1]
  v += dpp_f32<0x141>(v);       // row_half_mirror
  v += dpp_f32<0x140>(v);       // row_mirror: every lane holds its row's sum
  v += dpp_f32<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f32<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

struct V3f {
  float x, y, z;
};
__device__ __forceinline__ V3f vsub(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float vdot(V3f a, V3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3f vcross(V3f a, V3f b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float vlen(V3f a) { return sqrtf(vdot(a, a)); }
__device__ __forceinline__ V3f vert(const float* V, uint32_t i) {
  return {V[i * 3], V[i * 3 + 1], V[i * 3 + 2]};
}

// winding_number (mvc.rs:326-347); triangles split over the lanes
__device__ float winding(const CgsMvcArgs& A, V3f p) {
  const uint32_t lane = threadIdx.x & 63u;
  float part = 0.0f;
  uint32_t on_vertex = 0u;
  for (uint32_t t = lane; t < A.nt; t += 64u) {
    const V3f a = vsub(vert(A.verts, A.tris[t * 3]), p);
    const V3f b = vsub(vert(A.verts, A.tris[t * 3 + 1]), p);
    const V3f c = vsub(vert(A.verts, A.tris[t * 3 + 2]), p);
    const float la = vlen(a), lb = vlen(b), lc = vlen(c);
    if (la < 1e-7f || lb < 1e-7f || lc < 1e-7f) {
      on_vertex = 1u;
      continue;
    }
    const float det = vdot(a, vcross(b, c));
    const float den = ((la * lb * lc + vdot(a, b) * lc) + vdot(b, c) * la) + vdot(c, a) * lb;
    part += 2.0f * atan2f(det, den);
  }
  if (__ballot(on_vertex != 0u) != 0ull) return 1.0f;
  return wave_sum(part) / (4.0f * MVC_PI);
}

// The weights of one point, left in s_w[v] (v < nv).  s_dir holds the
// directions and distances, s_mu the per-triangle terms (3 per triangle).
__device__ void mvc_eval(const CgsMvcArgs& A, V3f x, float4* s_dir, float* s_mu, float* s_w) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nv = A.nv, nt = A.nt;
  // vertices: directions, distances, first coincident vertex
  uint32_t hit_v = 0xFFFFFFFFu;
  for (uint32_t v = lane; v < nv; v += 64u) {
    const V3f diff = vsub(vert(A.verts, v), x);
    const float d = vlen(diff);
    if (d < 1e-6f) hit_v = min(hit_v, v);
    s_dir[v] = make_float4(diff.x / d, diff.y / d, diff.z / d, d);
  }
  hit_v = wave_min_u32(hit_v);
  if (hit_v != 0xFFFFFFFFu) {
    for (uint32_t v = lane; v < nv; v += 64u) s_w[v] = v == hit_v ? 1.0f : 0.0f;
    return;
  }
  __builtin_amdgcn_wave_barrier();
  // triangles: mu terms, first triangle the point lies on
  float part = 0.0f;
  uint32_t hit_t = 0xFFFFFFFFu;
  for (uint32_t t = lane; t < nt; t += 64u) {
    const uint32_t i0 = A.tris[t * 3], i1 = A.tris[t * 3 + 1], i2 = A.tris[t * 3 + 2];
    const float4 d0 = s_dir[i0], d1 = s_dir[i1], d2 = s_dir[i2];
    const V3f u0 = {d0.x, d0.y, d0.z}, u1 = {d1.x, d1.y, d1.z}, u2 = {d2.x, d2.y, d2.z};
    const float l0 = vlen(vsub(u1, u2)), l1 = vlen(vsub(u2, u0)), l2 = vlen(vsub(u0, u1));
    const float th0 = 2.0f * asinf(cgs_fclamp(l0 * 0.5f, 0.0f, 1.0f));
    const float th1 = 2.0f * asinf(cgs_fclamp(l1 * 0.5f, 0.0f, 1.0f));
    const float th2 = 2.0f * asinf(cgs_fclamp(l2 * 0.5f, 0.0f, 1.0f));
    const float h = 0.5f * ((th0 + th1) + th2);
    float mu0 = 0.0f, mu1 = 0.0f, mu2 = 0.0f;
    if (fabsf(MVC_PI - h) < 1e-5f) {
      hit_t = min(hit_t, t);
    } else {
      const float st0 = sinf(th0), st1 = sinf(th1), st2 = sinf(th2);
      if (!(fabsf(st0) < 1e-6f || fabsf(st1) < 1e-6f || fabsf(st2) < 1e-6f)) {
        const float sh = sinf(h);
        const float c0 = cgs_fclamp((2.0f * sh * sinf(h - th0)) / (st1 * st2) - 1.0f, -1.0f, 1.0f);
        const float c1 = cgs_fclamp((2.0f * sh * sinf(h - th1)) / (st2 * st0) - 1.0f, -1.0f, 1.0f);
        const float c2 = cgs_fclamp((2.0f * sh * sinf(h - th2)) / (st0 * st1) - 1.0f, -1.0f, 1.0f);
        const float det = vdot(u2, vcross(u0, u1));
        const float sg = det >= 0.0f ? 1.0f : -1.0f;
        const float s0 = sg * sqrtf(cgs_fmax(1.0f - c0 * c0, 0.0f));
        const float s1 = sg * sqrtf(cgs_fmax(1.0f - c1 * c1, 0.0f));
        const float s2 = sg * sqrtf(cgs_fmax(1.0f - c2 * c2, 0.0f));
        if (!(fabsf(s0) <= 1e-6f || fabsf(s1) <= 1e-6f || fabsf(s2) <= 1e-6f)) {
          mu0 = (th0 - c1 * th2 - c2 * th1) / (2.0f * d0.w * st1 * s2);
          mu1 = (th1 - c2 * th0 - c0 * th2) / (2.0f * d1.w * st2 * s0);
          mu2 = (th2 - c0 * th1 - c1 * th0) / (2.0f * d2.w * st0 * s1);
        }
      }
    }
    s_mu[t * 3] = mu0;
    s_mu[t * 3 + 1] = mu1;
    s_mu[t * 3 + 2] = mu2;
    part += (mu0 + mu1) + mu2;
  }
  hit_t = wave_min_u32(hit_t);
  __builtin_amdgcn_wave_barrier();
  if (hit_t != 0xFFFFFFFFu) {
    // the point is on triangle hit_t: weights from the opposite-angle sines
    const uint32_t i0 = A.tris[hit_t * 3], i1 = A.tris[hit_t * 3 + 1], i2 = A.tris[hit_t * 3 + 2];
    const float4 d0 = s_dir[i0], d1 = s_dir[i1], d2 = s_dir[i2];
    const V3f u0 = {d0.x, d0.y, d0.z}, u1 = {d1.x, d1.y, d1.z}, u2 = {d2.x, d2.y, d2.z};
    const float th0 = 2.0f * asinf(cgs_fclamp(vlen(vsub(u1, u2)) * 0.5f, 0.0f, 1.0f));
    const float th1 = 2.0f * asinf(cgs_fclamp(vlen(vsub(u2, u0)) * 0.5f, 0.0f, 1.0f));
    const float th2 = 2.0f * asinf(cgs_fclamp(vlen(vsub(u0, u1)) * 0.5f, 0.0f, 1.0f));
    const float w0 = sinf(th0) * d1.w * d2.w;
    const float w1 = sinf(th1) * d2.w * d0.w;
    const float w2 = sinf(th2) * d0.w * d1.w;
    const float sum = (w0 + w1) + w2;
    const bool ok = fabsf(sum) > 1e-12f;
    for (uint32_t v = lane; v < nv; v += 64u)
      s_w[v] = !ok ? 0.0f : v == i0 ? w0 / sum : v == i1 ? w1 / sum : v == i2 ? w2 / sum : 0.0f;
    __builtin_amdgcn_wave_barrier();
    return;
  }
  const float total = wave_sum(part);
  if (fabsf(total) > 1e-12f) {
    const float inv = 1.0f / total;
    for (uint32_t v = lane; v < nv; v += 64u) {
      float acc = 0.0f;  // incident triangles in triangle order
      for (uint32_t j = A.vt_start[v]; j < A.vt_start[v + 1]; ++j) acc += s_mu[A.vt_list[j]];
      s_w[v] = acc * inv;
    }
  } else {
    // nearest vertex (first on ties)
    float best = __builtin_inff();
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t v = lane; v < nv; v += 64u)
      if (s_dir[v].w < best) {
        best = s_dir[v].w;
        bi = v;
      }
    // wave arg-min: smallest distance, then smallest index
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const float ob = __shfl_xor(best, o, 64);
      const uint32_t oi = (uint32_t)__shfl_xor((int)bi, o, 64);
      if (ob < best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    for (uint32_t v = lane; v < nv; v += 64u) s_w[v] = v == bi ? 1.0f : 0.0f;
  }
  __builtin_amdgcn_wave_barrier();
}

// bind_gaussians (mvc.rs:498-541): one wave per point.
__global__ __launch_bounds__(64) void k_mvc_bind(CgsMvcArgs A) {
  extern __shared__ float4 s_dyn[];
  float4* s_dir = s_dyn;
  float* s_mu = reinterpret_cast<float*>(s_dir + A.nv);
  float* s_w = s_mu + 3 * A.nt;
  float* s_w2 = s_w + A.nv;
  const uint32_t lane = threadIdx.x;
  if (blockIdx.x >= A.n) return;
  const uint32_t i = A.list ? A.list[blockIdx.x] : blockIdx.x;
  const V3f p = {A.pts[(size_t)i * 3], A.pts[(size_t)i * 3 + 1], A.pts[(size_t)i * 3 + 2]};
  float* row = A.out_w + (size_t)i * A.nv;
  float* grow = A.out_g ? A.out_g + (size_t)i * A.nv * 3 : nullptr;
  const bool inside = fabsf(winding(A, p)) >= 0.5f;
  if (A.mask && lane == 0) A.mask[i] = inside ? 1u : 0u;
  if (!inside) {
    for (uint32_t v = lane; v < A.nv; v += 64u) row[v] = 0.0f;
    if (grow)
      for (uint32_t v = lane; v < 3 * A.nv; v += 64u) grow[v] = 0.0f;
    return;
  }
  mvc_eval(A, p, s_dir, s_mu, s_w);
  for (uint32_t v = lane; v < A.nv; v += 64u) row[v] = s_w[v];
  if (!grow) return;
  const float inv_2h = 0.5f / A.h;
  for (int axis = 0; axis < 3; ++axis) {
    V3f pp = p, pm = p;
    if (axis == 0) {
      pp.x += A.h;
      pm.x -= A.h;
    } else if (axis == 1) {
      pp.y += A.h;
      pm.y -= A.h;
    } else {
      pp.z += A.h;
      pm.z -= A.h;
    }
    mvc_eval(A, pp, s_dir, s_mu, s_w2);
    mvc_eval(A, pm, s_dir, s_mu, s_w);
    for (uint32_t v = lane; v < A.nv; v += 64u) grow[(size_t)v * 3 + axis] = (s_w2[v] - s_w[v]) * inv_2h;
    __builtin_amdgcn_wave_barrier();
  }
}

// winding numbers only (contains_point = |w| >= 0.5)
__global__ __launch_bounds__(64) void k_mvc_winding(CgsMvcArgs A, float* out) {
  const uint32_t i = blockIdx.x;
  if (i >= A.n) return;
  const V3f p = {A.pts[(size_t)i * 3], A.pts[(size_t)i * 3 + 1], A.pts[(size_t)i * 3 + 2]};
  const float w = winding(A, p);
  if (threadIdx.x == 0) out[i] = w;
}

// evaluate_position (mvc.rs:473-481) and MvcBindings::local_jacobian (66-78)
// for every bound point: x' = sum_j w_j v'_j, F = sum_j v'_j g_j^T.  One wave
// per point (four per block), lanes over vertices with 12-byte loads of the
// gradient and vertex triples, then a reduce-scatter butterfly that leaves
// the 12 sums (x' and the 9 entries of F) on 12 lanes after 14 cross-lane
// exchanges.  The reduction order is fixed, so the result is reproducible.
struct F3 {
  float x, y, z;
};

// lanes whose `bit` is clear keep the first half of a[0..2h), the others the
// second half, each adding its partner's copy of the half it keeps
template <int H>
__device__ __forceinline__ void keep_half(float* a, uint32_t lane, int bit) {
  const bool hi = (lane & (uint32_t)bit) != 0u;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float send = hi ? a[k] : a[H + k];
    const float keep = hi ? a[H + k] : a[k];
    a[k] = keep + __shfl_xor(send, bit, 64);
  }
}

__global__ __launch_bounds__(256) void k_mvc_deform(const float* w, const float* g,
                                                    const uint32_t* mask, const float* rest_pos,
                                                    const float* dverts, uint32_t nv, uint32_t n,
                                                    float* out_pos, float* out_F) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (i >= n) return;
  if (mask && mask[i] == 0u) {
    if (lane < 3) out_pos[(size_t)i * 3 + lane] = rest_pos[(size_t)i * 3 + lane];
    if (out_F && lane < 9) out_F[(size_t)i * 9 + lane] = (lane % 4 == 0) ? 1.0f : 0.0f;
    return;
  }
  const float* wr = w + (size_t)i * nv;
  const F3* gr = (g && out_F) ? reinterpret_cast<const F3*>(g + (size_t)i * nv * 3) : nullptr;
  const F3* vv = reinterpret_cast<const F3*>(dverts);
  // a[0..2] = x', a[3 + 3c + r] = F (column-major), a[12..15] padding
  float a[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) a[k] = 0.0f;
  if (gr) {
    for (uint32_t v = lane; v < nv; v += 64u) {
      const F3 p = vv[v];
      const float wv = __builtin_nontemporal_load(wr + v);
      const F3 gv = gr[v];
      a[0] += p.x * wv; a[1] += p.y * wv; a[2] += p.z * wv;
      a[3] += p.x * gv.x; a[4] += p.y * gv.x; a[5] += p.z * gv.x;
      a[6] += p.x * gv.y; a[7] += p.y * gv.y; a[8] += p.z * gv.y;
      a[9] += p.x * gv.z; a[10] += p.y * gv.z; a[11] += p.z * gv.z;
    }
  } else {
    for (uint32_t v = lane; v < nv; v += 64u) {
      const F3 p = vv[v];
      const float wv = __builtin_nontemporal_load(wr + v);
      a[0] += p.x * wv; a[1] += p.y * wv; a[2] += p.z * wv;
    }
  }
  // 16 values -> 8 (bit 32) -> 4 (bit 16) -> 2 (bit 8) -> 1 (bit 4), then the
  // last two butterflies over bits 2 and 1: lane l ends with the sum of
  // value index ((l >> 5) & 1) * 8 + ((l >> 4) & 1) * 4 + ((l >> 3) & 1) * 2 + ((l >> 2) & 1)
  keep_half<8>(a, lane, 32);
  keep_half<4>(a, lane, 16);
  keep_half<2>(a, lane, 8);
  keep_half<1>(a, lane, 4);
  a[0] += __shfl_xor(a[0], 2, 64);
  a[0] += __shfl_xor(a[0], 1, 64);
  if ((lane & 3u) == 0u) {
    const uint32_t k = ((lane >> 5) & 1u) * 8u + ((lane >> 4) & 1u) * 4u + ((lane >> 3) & 1u) * 2u +
                       ((lane >> 2) & 1u);
    if (k < 3u)
      out_pos[(size_t)i * 3 + k] = a[0];
    else if (k < 12u && out_F && gr)
      out_F[(size_t)i * 9 + (k - 3u)] = a[0];
  }
}

__global__ __launch_bounds__(256) void k_mvc_select(const float* winding, const float* residual,
                                                    float tol, uint32_t n, uint32_t nv,
                                                    uint32_t* inside, float* w, float* g,
                                                    uint32_t* list, uint32_t* count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const bool in = winding ? fabsf(winding[i]) >= 0.5f : inside[i] != 0u;
  if (winding) inside[i] = in ? 1u : 0u;
  if (in && residual[i] > tol) list[atomicAdd(count, 1u)] = i;
  if (!in) {
    for (uint32_t v = 0; v < nv; ++v) w[(size_t)i * nv + v] = 0.0f;
    if (g)
      for (uint32_t v = 0; v < 3 * nv; ++v) g[(size_t)i * nv * 3 + v] = 0.0f;
  }
}


// deform_scene_cpu (mvc.rs:564-749) per splat, from the point's deformed
// position x' and Jacobian F (k_mvc_deform): bound splats get Sigma' =
// F Sigma F^T and their SH rotated by the polar rotation of F (rest opacity
// kept); unbound ones keep the rest record.  A bound splat whose largest
// stretch of F exceeds `tau` is split along the major axis of Sigma' (child A
// replaces it, child B goes to the scratch record of i, flags[i] = 1); the mip
// filter applies to every record that is not split and to both children.
__global__ __launch_bounds__(64) void k_mvc_scene(CgsMvcSceneArgs A) {
  const uint32_t i = blockIdx.x * 64u + threadIdx.x;
  if (i >= A.n) return;
  const bool bound = A.mask ? A.mask[i] != 0u : true;
  float cov[6], sh[48];
  for (int k = 0; k < 6; ++k) cov[k] = A.rest_cov[(size_t)i * 6 + k];
  for (int k = 0; k < 48; ++k) sh[k] = A.rest_sh[(size_t)i * 48 + k];
  float alpha = A.rest_opa[i];
  float pos[3];
  bool split = false;
  if (bound) {
    float F[9];
    for (int k = 0; k < 3; ++k) pos[k] = A.def_pos[(size_t)i * 3 + k];
    for (int k = 0; k < 9; ++k) F[k] = A.def_F[(size_t)i * 9 + k];
    float R[9], S[9], FS[9], Ft[9], FSF[9];
    cgs_polar_rotation(F, R);
    cgs_cov_to_m3(cov, S);
    cgs_m3_mul(F, S, FS);
    cgs_m3_transpose(F, Ft);
    cgs_m3_mul(FS, Ft, FSF);
    cgs_m3_to_cov(FSF, cov);
    cgs_sh_rotate_fast(sh, R, (int)A.sh_degree);
    split = A.tau > 0.0f && cgs_max_stretch(F) > A.tau;
  } else {
    for (int k = 0; k < 3; ++k) pos[k] = A.rest_pos[(size_t)i * 3 + k];
  }
  if (!split) {
    cgs_mip_filter_3d(cov, &alpha, A.mip);
  } else {
    float child[6], off[3], cb[6];
    cgs_split_major(cov, child, off);
    float aa = alpha, ab = alpha;
    for (int k = 0; k < 6; ++k) cb[k] = child[k];
    cgs_mip_filter_3d(child, &aa, A.mip);
    cgs_mip_filter_3d(cb, &ab, A.mip);
    float* b = A.child_b + (size_t)i * 10;
    for (int k = 0; k < 3; ++k) b[k] = pos[k] + off[k];
    for (int k = 0; k < 6; ++k) b[3 + k] = cb[k];
    b[9] = ab;
    for (int k = 0; k < 3; ++k) pos[k] = pos[k] + (-off[k]);
    for (int k = 0; k < 6; ++k) cov[k] = child[k];
    alpha = aa;
  }
  A.flags[i] = split ? 1u : 0u;
  for (int k = 0; k < 3; ++k) A.out_pos[(size_t)i * 3 + k] = pos[k];
  for (int k = 0; k < 6; ++k) A.out_cov[(size_t)i * 6 + k] = cov[k];
  A.out_opa[i] = alpha;
  for (int k = 0; k < 48; ++k) A.out_sh[(size_t)i * 48 + k] = sh[k];
}

// child B of split splat i goes to record n + ranks[i] (ranks = exclusive scan
// of the flags: children in parent order), with the parent's rotated SH
__global__ __launch_bounds__(256) void k_mvc_scene_children(CgsMvcSceneArgs A,
                                                            const uint32_t* ranks) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= A.n || A.flags[i] == 0u) return;
  const size_t j = (size_t)A.n + ranks[i];
  const float* b = A.child_b + (size_t)i * 10;
  for (int k = 0; k < 3; ++k) A.out_pos[j * 3 + k] = b[k];
  for (int k = 0; k < 6; ++k) A.out_cov[j * 6 + k] = b[3 + k];
  A.out_opa[j] = b[9];
  for (int k = 0; k < 48; ++k) A.out_sh[j * 48 + k] = A.out_sh[(size_t)i * 48 + k];
}

// ---------------------------------------------------------------------------
// Near/far split for the learned binding.  Triangle t's terms at x get the
// weight phi_t(x) = smoothstep((r1 - |x - c_t|) / (r1 - r0)), c_t its
// centroid: 1 within r0, 0 beyond r1.  The near sums u_near = sum_t phi_t
// mu_t (the singular, fast-varying part next to the cage) are evaluated
// exactly; the far sums are smooth and are what the network learns, either
// as the share q = u_far / (u_near + u_far) of the normalised weights or as
// the far sums u_far themselves.
__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float near_phi(float r, float r0, float r1) {
  const float t = cgs_fclamp((r1 - r) / (r1 - r0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// Far-share training targets, one wave per point: the loop of mvc_eval with
// each triangle's terms scaled by 1 - phi_t before the per-vertex sums, the
// normaliser over all terms (sums = 1: the far sums, not normalised).  Points
// on a vertex or a face, or with a zero normaliser, get a zero row (they do
// not occur among interior samples).
__global__ __launch_bounds__(64) void k_mvc_far_share(CgsMvcArgs A, const float* cent, float r0,
                                                      float r1, float q_scale, int sums,
                                                      float* out) {
  extern __shared__ float4 s_dyn[];
  float4* s_dir = s_dyn;
  float* s_mu = reinterpret_cast<float*>(s_dir + A.nv);
  const uint32_t lane = threadIdx.x, i = blockIdx.x;
  if (i >= A.n) return;
  const V3f x = {A.pts[(size_t)i * 3], A.pts[(size_t)i * 3 + 1], A.pts[(size_t)i * 3 + 2]};
  float* row = out + (size_t)i * A.nv;
  uint32_t bad = 0u;
  for (uint32_t v = lane; v < A.nv; v += 64u) {
    const V3f diff = vsub(vert(A.verts, v), x);
    const float d = vlen(diff);
    if (d < 1e-6f) bad = 1u;
    s_dir[v] = make_float4(diff.x / d, diff.y / d, diff.z / d, d);
  }
  if (__ballot(bad != 0u) != 0ull) {
    for (uint32_t v = lane; v < A.nv; v += 64u) row[v] = 0.0f;
    return;
  }
  __builtin_amdgcn_wave_barrier();
  float part = 0.0f;
  for (uint32_t t = lane; t < A.nt; t += 64u) {
    const uint32_t i0 = A.tris[t * 3], i1 = A.tris[t * 3 + 1], i2 = A.tris[t * 3 + 2];
    const float4 d0 = s_dir[i0], d1 = s_dir[i1], d2 = s_dir[i2];
    const V3f u0 = {d0.x, d0.y, d0.z}, u1 = {d1.x, d1.y, d1.z}, u2 = {d2.x, d2.y, d2.z};
    const float th0 = 2.0f * asinf(cgs_fclamp(vlen(vsub(u1, u2)) * 0.5f, 0.0f, 1.0f));
    const float th1 = 2.0f * asinf(cgs_fclamp(vlen(vsub(u2, u0)) * 0.5f, 0.0f, 1.0f));
    const float th2 = 2.0f * asinf(cgs_fclamp(vlen(vsub(u0, u1)) * 0.5f, 0.0f, 1.0f));
    const float h = 0.5f * ((th0 + th1) + th2);
    float mu0 = 0.0f, mu1 = 0.0f, mu2 = 0.0f;
    if (fabsf(MVC_PI - h) < 1e-5f) {
      bad = 1u;
    } else {
      const float st0 = sinf(th0), st1 = sinf(th1), st2 = sinf(th2);
      if (!(fabsf(st0) < 1e-6f || fabsf(st1) < 1e-6f || fabsf(st2) < 1e-6f)) {
        const float sh = sinf(h);
        const float c0 = cgs_fclamp((2.0f * sh * sinf(h - th0)) / (st1 * st2) - 1.0f, -1.0f, 1.0f);
        const float c1 = cgs_fclamp((2.0f * sh * sinf(h - th1)) / (st2 * st0) - 1.0f, -1.0f, 1.0f);
        const float c2 = cgs_fclamp((2.0f * sh * sinf(h - th2)) / (st0 * st1) - 1.0f, -1.0f, 1.0f);
        const float sg = vdot(u2, vcross(u0, u1)) >= 0.0f ? 1.0f : -1.0f;
        const float s0 = sg * sqrtf(cgs_fmax(1.0f - c0 * c0, 0.0f));
        const float s1 = sg * sqrtf(cgs_fmax(1.0f - c1 * c1, 0.0f));
        const float s2 = sg * sqrtf(cgs_fmax(1.0f - c2 * c2, 0.0f));
        if (!(fabsf(s0) <= 1e-6f || fabsf(s1) <= 1e-6f || fabsf(s2) <= 1e-6f)) {
          mu0 = (th0 - c1 * th2 - c2 * th1) / (2.0f * d0.w * st1 * s2);
          mu1 = (th1 - c2 * th0 - c0 * th2) / (2.0f * d1.w * st2 * s0);
          mu2 = (th2 - c0 * th1 - c1 * th0) / (2.0f * d2.w * st0 * s1);
        }
      }
    }
    part += (mu0 + mu1) + mu2;
    const V3f c = vert(cent, t);
    const float far = 1.0f - near_phi(vlen(vsub(x, c)), r0, r1);
    s_mu[t * 3] = far * mu0;
    s_mu[t * 3 + 1] = far * mu1;
    s_mu[t * 3 + 2] = far * mu2;
  }
  const float total = wave_sum(part);
  const bool skip = __ballot(bad != 0u) != 0ull || !(fabsf(total) > 1e-12f);
  __builtin_amdgcn_wave_barrier();
  const float scale = skip ? 0.0f : sums ? q_scale : q_scale / total;
  for (uint32_t v = lane; v < A.nv; v += 64u) {
    float acc = 0.0f;
    for (uint32_t j = A.vt_start[v]; j < A.vt_start[v + 1]; ++j) acc += s_mu[A.vt_list[j]];
    row[v] = acc * scale;
  }
}

// winding_number (mvc.rs:326-347) with one point per lane: the triangle loop
// index is the same on every lane, so its vertices come through scalar loads
// and the whole wave works on 64 points at once.
__global__ __launch_bounds__(256) void k_mvc_winding_lane(const float* verts, const uint32_t* tris,
                                                          uint32_t nt, const float* pts, uint32_t n,
                                                          float* out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t ic = i < n ? i : n - 1u;
  const V3f p = {pts[(size_t)ic * 3], pts[(size_t)ic * 3 + 1], pts[(size_t)ic * 3 + 2]};
  float sum = 0.0f;
  bool on_vertex = false;
  for (uint32_t t = 0; t < nt; ++t) {
    const V3f a = vsub(vert(verts, tris[t * 3]), p);
    const V3f b = vsub(vert(verts, tris[t * 3 + 1]), p);
    const V3f c = vsub(vert(verts, tris[t * 3 + 2]), p);
    const float la = vlen(a), lb = vlen(b), lc = vlen(c);
    if (la < 1e-7f || lb < 1e-7f || lc < 1e-7f) {
      on_vertex = true;
      continue;
    }
    const float det = vdot(a, vcross(b, c));
    const float den = ((la * lb * lc + vdot(a, b) * lc) + vdot(b, c) * la) + vdot(c, a) * lb;
    sum += 2.0f * atan2f(det, den);
  }
  if (i < n) out[i] = on_vertex ? 1.0f : sum / (4.0f * MVC_PI);
}

// JSW terms of one triangle at x: the loop body of mvc_eval with the
// directions computed for this triangle alone and every sine and cosine taken
// from the chord lengths l_i = |u_j - u_k| instead of sinf: with
// a_i = theta_i / 2 = asin(l_i / 2), sin a_i = l_i / 2, cos a_i =
// sqrt((1 - l_i/2)(1 + l_i/2)), sin theta_i = 2 sin a_i cos a_i,
// cos theta_i = 1 - 2 sin^2 a_i, and h = a_0 + a_1 + a_2 by the addition
// formulas (the same quantities, three asinf per triangle instead of three
// asinf and seven sinf).  Returns 0 when the triangle adds nothing, 1 with mu
// set, 2 when x is on a vertex or on the triangle.
__device__ __forceinline__ int jsw_terms(V3f x, V3f v0, V3f v1, V3f v2, float* mu) {
  const V3f e0 = vsub(v0, x), e1 = vsub(v1, x), e2 = vsub(v2, x);
  const float d0 = vlen(e0), d1 = vlen(e1), d2 = vlen(e2);
  if (d0 < 1e-6f || d1 < 1e-6f || d2 < 1e-6f) return 2;
  const V3f u0 = {e0.x / d0, e0.y / d0, e0.z / d0};
  const V3f u1 = {e1.x / d1, e1.y / d1, e1.z / d1};
  const V3f u2 = {e2.x / d2, e2.y / d2, e2.z / d2};
  const float s0h = cgs_fclamp(vlen(vsub(u1, u2)) * 0.5f, 0.0f, 1.0f);
  const float s1h = cgs_fclamp(vlen(vsub(u2, u0)) * 0.5f, 0.0f, 1.0f);
  const float s2h = cgs_fclamp(vlen(vsub(u0, u1)) * 0.5f, 0.0f, 1.0f);
  const float c0h = sqrtf((1.0f - s0h) * (1.0f + s0h));
  const float c1h = sqrtf((1.0f - s1h) * (1.0f + s1h));
  const float c2h = sqrtf((1.0f - s2h) * (1.0f + s2h));
  const float th0 = 2.0f * asinf(s0h), th1 = 2.0f * asinf(s1h), th2 = 2.0f * asinf(s2h);
  const float h = 0.5f * ((th0 + th1) + th2);
  if (fabsf(MVC_PI - h) < 1e-5f) return 2;
  const float st0 = 2.0f * s0h * c0h, st1 = 2.0f * s1h * c1h, st2 = 2.0f * s2h * c2h;
  if (fabsf(st0) < 1e-6f || fabsf(st1) < 1e-6f || fabsf(st2) < 1e-6f) return 0;
  const float ct0 = 1.0f - 2.0f * s0h * s0h, ct1 = 1.0f - 2.0f * s1h * s1h, ct2 = 1.0f - 2.0f * s2h * s2h;
  const float s01 = s0h * c1h + c0h * s1h, c01 = c0h * c1h - s0h * s1h;
  const float sh = s01 * c2h + c01 * s2h, ch = c01 * c2h - s01 * s2h;
  const float c0 = cgs_fclamp((2.0f * sh * (sh * ct0 - ch * st0)) / (st1 * st2) - 1.0f, -1.0f, 1.0f);
  const float c1 = cgs_fclamp((2.0f * sh * (sh * ct1 - ch * st1)) / (st2 * st0) - 1.0f, -1.0f, 1.0f);
  const float c2 = cgs_fclamp((2.0f * sh * (sh * ct2 - ch * st2)) / (st0 * st1) - 1.0f, -1.0f, 1.0f);
  const float sg = vdot(u2, vcross(u0, u1)) >= 0.0f ? 1.0f : -1.0f;
  const float s0 = sg * sqrtf(cgs_fmax(1.0f - c0 * c0, 0.0f));
  const float s1 = sg * sqrtf(cgs_fmax(1.0f - c1 * c1, 0.0f));
  const float s2 = sg * sqrtf(cgs_fmax(1.0f - c2 * c2, 0.0f));
  if (fabsf(s0) <= 1e-6f || fabsf(s1) <= 1e-6f || fabsf(s2) <= 1e-6f) return 0;
  mu[0] = (th0 - c1 * th2 - c2 * th1) / (2.0f * d0 * st1 * s2);
  mu[1] = (th1 - c2 * th0 - c0 * th2) / (2.0f * d1 * st2 * s0);
  mu[2] = (th2 - c0 * th1 - c1 * th0) / (2.0f * d2 * st0 * s1);
  return 1;
}

// jsw_terms with the divisions of the near field as hardware reciprocals
// (v_rcp_f32 / v_rsq_f32, about 1 ulp) and contracted multiply-adds: the same
// quantities to a few ulp, which is far inside the learned binding's error
// budget (its weights are not held to the exact kernel's bits).  The unit
// directions come from one rsqrt per corner; 1 / (st_j st_k) = st_i / (st0
// st1 st2) shares one reciprocal between the three cosines.
__device__ __forceinline__ int jsw_terms_fast(V3f x, V3f v0, V3f v1, V3f v2, float* mu) {
#pragma clang fp contract(fast)
  const V3f e0 = vsub(v0, x), e1 = vsub(v1, x), e2 = vsub(v2, x);
  const float q0 = vdot(e0, e0), q1 = vdot(e1, e1), q2 = vdot(e2, e2);
  if (q0 < 1e-12f || q1 < 1e-12f || q2 < 1e-12f) return 2;
  const float r0 = __builtin_amdgcn_rsqf(q0), r1 = __builtin_amdgcn_rsqf(q1),
              r2 = __builtin_amdgcn_rsqf(q2);
  const float d0 = q0 * r0, d1 = q1 * r1, d2 = q2 * r2;
  const V3f u0 = {e0.x * r0, e0.y * r0, e0.z * r0};
  const V3f u1 = {e1.x * r1, e1.y * r1, e1.z * r1};
  const V3f u2 = {e2.x * r2, e2.y * r2, e2.z * r2};
  const V3f a0 = vsub(u1, u2), a1 = vsub(u2, u0), a2 = vsub(u0, u1);
  const float s0h = cgs_fclamp(__builtin_amdgcn_sqrtf(vdot(a0, a0)) * 0.5f, 0.0f, 1.0f);
  const float s1h = cgs_fclamp(__builtin_amdgcn_sqrtf(vdot(a1, a1)) * 0.5f, 0.0f, 1.0f);
  const float s2h = cgs_fclamp(__builtin_amdgcn_sqrtf(vdot(a2, a2)) * 0.5f, 0.0f, 1.0f);
  const float c0h = __builtin_amdgcn_sqrtf((1.0f - s0h) * (1.0f + s0h));
  const float c1h = __builtin_amdgcn_sqrtf((1.0f - s1h) * (1.0f + s1h));
  const float c2h = __builtin_amdgcn_sqrtf((1.0f - s2h) * (1.0f + s2h));
  const float th0 = 2.0f * asinf(s0h), th1 = 2.0f * asinf(s1h), th2 = 2.0f * asinf(s2h);
  const float h = 0.5f * ((th0 + th1) + th2);
  if (fabsf(MVC_PI - h) < 1e-5f) return 2;
  const float st0 = 2.0f * s0h * c0h, st1 = 2.0f * s1h * c1h, st2 = 2.0f * s2h * c2h;
  if (fabsf(st0) < 1e-6f || fabsf(st1) < 1e-6f || fabsf(st2) < 1e-6f) return 0;
  const float ct0 = 1.0f - 2.0f * s0h * s0h, ct1 = 1.0f - 2.0f * s1h * s1h, ct2 = 1.0f - 2.0f * s2h * s2h;
  const float s01 = s0h * c1h + c0h * s1h, c01 = c0h * c1h - s0h * s1h;
  const float sh = s01 * c2h + c01 * s2h, ch = c01 * c2h - s01 * s2h;
  const float rst = __builtin_amdgcn_rcpf(st0 * st1 * st2);
  const float c0 = cgs_fclamp((2.0f * sh * (sh * ct0 - ch * st0)) * (st0 * rst) - 1.0f, -1.0f, 1.0f);
  const float c1 = cgs_fclamp((2.0f * sh * (sh * ct1 - ch * st1)) * (st1 * rst) - 1.0f, -1.0f, 1.0f);
  const float c2 = cgs_fclamp((2.0f * sh * (sh * ct2 - ch * st2)) * (st2 * rst) - 1.0f, -1.0f, 1.0f);
  const float sg = vdot(u2, vcross(u0, u1)) >= 0.0f ? 1.0f : -1.0f;
  const float s0 = sg * __builtin_amdgcn_sqrtf(cgs_fmax(1.0f - c0 * c0, 0.0f));
  const float s1 = sg * __builtin_amdgcn_sqrtf(cgs_fmax(1.0f - c1 * c1, 0.0f));
  const float s2 = sg * __builtin_amdgcn_sqrtf(cgs_fmax(1.0f - c2 * c2, 0.0f));
  if (fabsf(s0) <= 1e-6f || fabsf(s1) <= 1e-6f || fabsf(s2) <= 1e-6f) return 0;
  mu[0] = (th0 - c1 * th2 - c2 * th1) * __builtin_amdgcn_rcpf(2.0f * d0 * st1 * s2);
  mu[1] = (th1 - c2 * th0 - c0 * th2) * __builtin_amdgcn_rcpf(2.0f * d1 * st2 * s0);
  mu[2] = (th2 - c0 * th1 - c1 * th0) * __builtin_amdgcn_rcpf(2.0f * d2 * st0 * s1);
  return 1;
}

// The learned binding's weights, one wave per point:
//   w = (1 - sum q) u_near U_near / (U_near^2 + tau^2) + q,
// u_near = sum over the point's near triangles of phi_t mu_t, U_near its sum,
// q = the network's far share.  The factor U / (U^2 + tau^2) is 1 / U_near
// wherever |U_near| >> tau and fades the near term out continuously as
// U_near -> 0 at the near shell's outer edge; tau is far below any interior
// point's weight sum (>= 4 pi / diagonal), so the near term is short by at
// most tau / (2 sum u) of the weights' sum there.  Gradients: central
// differences (step h) of the weights with the network's share linearised
// from its tangents, q(p +- h e_a) = q +- h dq_a:
//   g_a = (1 - Q) (un+ - un-) / 2h - dQ_a (un+ + un-) / 2 + dq_a.
// With far sums (far_sums = 1) the network gives f = u_far instead and
//   w = (u_near + f) / (U_near + F),  F = sum f,
// which needs no fade: the denominator is the whole MVC sum, bounded away
// from 0 inside the cage, and next to the cage, where u_near is singular, the
// network's error is divided by it.  Gradients: the same central
// differences, f(p +- h e_a) = f +- h df_a.  A sum at p below tau, or one at
// p +- h e_a within tau of 0, flags the point for the exact path.
// The residual |sum w v - x| is taken before the linear-precision projection
// w' = w + A^T Minv ([1; x] - A w) (and g'_a likewise with e_a), which the
// rows receive.
//
// Layout: the triangles of the point's grid cell within reach are compacted
// into LDS, then the (triangle, evaluation point) pairs are spread over the
// lanes so that every lane does about the same number of JSW evaluations;
// the per-vertex sums of the seven evaluations live in a small LDS hash table
// keyed by vertex (NEAR_SLOTS slots, open addressing; a point whose near
// triangles touch more vertices than fit is flagged for the exact path: the
// ~60 near triangles of the 642-vertex cage touch ~40 vertices), so a wave
// needs 6 KiB of LDS whatever the cage size.  The network's rows are loaded
// into registers (VPL vertices per lane) after the near loop, and the
// weights stay there until they are written.  Five waves per SIMD (96 VGPRs,
// a few spilled) measured 3 % faster than four; the reductions are DPP sums.
constexpr uint32_t NEAR_SLOTS = 128u, NEAR_TRIS = 256u, NEAR_EMPTY = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t near_hash(uint32_t v) { return (v * 0x9E3779B1u) >> 25; }

// slot of vertex v (inserted when absent); NEAR_EMPTY when the table is full
__device__ __forceinline__ uint32_t near_slot_insert(uint32_t* keys, uint32_t v) {
  uint32_t s = near_hash(v);
  for (uint32_t k = 0; k < NEAR_SLOTS; ++k) {
    const uint32_t old = atomicCAS(&keys[s], NEAR_EMPTY, v);
    if (old == NEAR_EMPTY || old == v) return s;
    s = (s + 1u) & (NEAR_SLOTS - 1u);
  }
  return NEAR_EMPTY;
}

__device__ __forceinline__ uint32_t near_slot_find(const uint32_t* keys, uint32_t v) {
  uint32_t s = near_hash(v);
  for (uint32_t k = 0; k < NEAR_SLOTS; ++k) {
    const uint32_t key = keys[s];
    if (key == v) return s;
    if (key == NEAR_EMPTY) return NEAR_EMPTY;
    s = (s + 1u) & (NEAR_SLOTS - 1u);
  }
  return NEAR_EMPTY;
}

// Does the segment [p, c] cross triangle (v0, v1, v2)?  Moller-Trumbore:
// 0 = no, 1 = yes, 2 = too close to call (the segment passes within 1e-5,
// in barycentric or segment-length terms, of an edge, a vertex or its own
// end points, or runs parallel to the triangle's plane inside it).
__device__ __forceinline__ int segment_crosses(V3f p, V3f c, V3f v0, V3f v1, V3f v2) {
  const float EPS = 1e-5f;
  const V3f d = vsub(c, p), e1 = vsub(v1, v0), e2 = vsub(v2, v0);
  const V3f pv = vcross(d, e2);
  const float det = vdot(e1, pv);
  const float scale = vlen(e1) * vlen(e2) * vlen(d);
  const V3f tv = vsub(p, v0);
  if (fabsf(det) <= 1e-7f * scale) {
    // parallel: undecided only if the segment lies in the plane
    const V3f nrm = vcross(e1, e2);
    const float dist = fabsf(vdot(tv, nrm)) / fmaxf(vlen(nrm), 1e-30f);
    return dist <= EPS * vlen(d) ? 2 : 0;
  }
  const float inv = 1.0f / det;
  const float u = vdot(tv, pv) * inv;
  const V3f qv = vcross(tv, e1);
  const float v = vdot(d, qv) * inv;
  const float t = vdot(e2, qv) * inv;
  const float w = 1.0f - u - v;
  const bool near_in = u > -EPS && v > -EPS && w > -EPS && t > -EPS && t < 1.0f + EPS;
  if (!near_in) return 0;
  const bool in = u > EPS && v > EPS && w > EPS && t > EPS && t < 1.0f - EPS;
  return in ? 1 : 2;
}

template <int VPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_mvc_nearfar(CgsNearArgs A) {
  __shared__ uint32_t s_key[NEAR_SLOTS];
  __shared__ float s_val[7][NEAR_SLOTS];
  __shared__ uint32_t s_tri[NEAR_TRIS];
  __shared__ uint32_t s_tslot[NEAR_TRIS];
  const uint32_t lane = threadIdx.x, i = blockIdx.x;
  if (i >= A.n) return;
  const uint32_t nv = A.nv;
  const V3f p = {A.pts[(size_t)i * 3], A.pts[(size_t)i * 3 + 1], A.pts[(size_t)i * 3 + 2]};
  float* row = A.out_w + (size_t)i * nv;
  float* grow = A.out_g ? A.out_g + (size_t)i * nv * 3 : nullptr;
  const float fx = (p.x - A.lo[0]) * A.inv_cell, fy = (p.y - A.lo[1]) * A.inv_cell,
              fz = (p.z - A.lo[2]) * A.inv_cell;
  const bool in_grid = fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)A.dim[0] &&
                       fy < (float)A.dim[1] && fz < (float)A.dim[2];
  uint32_t c0 = 0u, c1 = 0u, cell = 0u;
  V3f cc = p;
  if (in_grid) {
    const uint32_t cx = min((uint32_t)fx, A.dim[0] - 1u), cy = min((uint32_t)fy, A.dim[1] - 1u),
                   cz = min((uint32_t)fz, A.dim[2] - 1u);
    cell = (cz * A.dim[1] + cy) * A.dim[0] + cx;
    c0 = A.cell_start[cell];
    c1 = A.cell_start[cell + 1];
    const float cs = 1.0f / A.inv_cell;
    cc = {A.lo[0] + ((float)cx + 0.5f) * cs, A.lo[1] + ((float)cy + 0.5f) * cs,
          A.lo[2] + ((float)cz + 0.5f) * cs};
  }
  // inside test: the winding numbers when given, else the parity of the
  // segment from p to its cell's centre (whose side is known) against the
  // cell's triangles (every triangle meeting the cell is in its list);
  // undecided points go to the exact path with the inside flag set
  bool in;
  uint32_t degenerate = 0u;
  if (A.winding) {
    in = fabsf(A.winding[i]) >= 0.5f;
  } else if (!in_grid) {
    in = false;  // outside the bounding box
  } else {
    const uint32_t ref = A.cell_inside[cell];
    uint32_t crossings = 0u, unsure = ref > 1u ? 1u : 0u;
    for (uint32_t j = c0 + lane; j < c1; j += 64u) {
      const uint32_t t = A.cell_list[j];
      const int k = segment_crosses(p, cc, vert(A.verts, A.tris[t * 3]), vert(A.verts, A.tris[t * 3 + 1]),
                                    vert(A.verts, A.tris[t * 3 + 2]));
      crossings += k == 1 ? 1u : 0u;
      unsure |= k == 2 ? 1u : 0u;
    }
    const bool odd = ((uint32_t)wave_sum_dpp((float)crossings) & 1u) != 0u;
    degenerate = __ballot(unsure != 0u) != 0ull ? 1u : 0u;
    in = degenerate ? true : ((ref == 1u) != odd);
  }
  if (A.inside && lane == 0) A.inside[i] = in ? 1u : 0u;
  if (!in) {
    for (uint32_t v = lane; v < nv; v += 64u) row[v] = 0.0f;
    if (grow)
      for (uint32_t v = lane; v < 3 * nv; v += 64u) grow[v] = 0.0f;
    if (lane == 0) A.residual[i] = 0.0f;
    return;
  }
  const float zs = 1.0f / A.q_scale;
  const bool tan = grow && (A.DQ || A.DQh);
  const size_t qrow = (size_t)i * A.ldq;
  const uint32_t evals = grow ? 7u : 1u;
  for (uint32_t k = lane; k < NEAR_SLOTS; k += 64u) {
    s_key[k] = NEAR_EMPTY;
#pragma unroll
    for (int e = 0; e < 7; ++e) s_val[e][k] = 0.0f;
  }
  float U[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float reach = A.r1 + A.h;
  // chunks of up to NEAR_TRIS candidates: compact those within reach, then
  // the (triangle, evaluation) pairs over the lanes
  for (uint32_t base = c0; base < c1; base += NEAR_TRIS) {
    const uint32_t end = min(c1, base + NEAR_TRIS);
    uint32_t m = 0u;
    __syncthreads();
    for (uint32_t j0 = base; j0 < end; j0 += 64u) {
      const uint32_t j = j0 + lane;
      uint32_t t = 0u;
      bool keep = false;
      if (j < end) {
        t = A.cell_list[j];
        keep = vlen(vsub(p, vert(A.cent, t))) < reach;
      }
      const uint64_t bal = __ballot(keep);
      if (keep) s_tri[m + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = t;
      m += (uint32_t)__popcll(bal);
    }
    __syncthreads();
    // the near triangles' vertices take their hash slots once per triangle,
    // not once per evaluation: three 8-bit slots per triangle, bit 24 set
    // when the table is full (the point then goes to the exact path)
    for (uint32_t j = lane; j < m; j += 64u) {
      const uint32_t t = s_tri[j];
      uint32_t packed = 0u;
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        const uint32_t sl = near_slot_insert(s_key, A.tris[t * 3 + c3]);
        packed |= sl == NEAR_EMPTY ? (1u << 24) : sl << (8 * c3);
      }
      if (packed >> 24) degenerate = 1u;
      s_tslot[j] = packed;
    }
    __syncthreads();
    for (uint32_t item = lane; item < m * evals; item += 64u) {
      const uint32_t j = evals == 7u ? item / 7u : item;
      const int e = (int)(item - j * evals);
      const uint32_t t = s_tri[j];
      const V3f c = vert(A.cent, t);
      V3f x = p;
      if (e > 0) {
        const float step = (e & 1) ? A.h : -A.h;
        const int axis = (e - 1) >> 1;
        if (axis == 0) x.x += step;
        else if (axis == 1) x.y += step;
        else x.z += step;
      }
      const float phi = near_phi(vlen(vsub(x, c)), A.r0, A.r1);
      if (!(phi > 0.0f)) continue;
      const uint32_t iv[3] = {A.tris[t * 3], A.tris[t * 3 + 1], A.tris[t * 3 + 2]};
      float mu[3];
      const int k =
          jsw_terms_fast(x, vert(A.verts, iv[0]), vert(A.verts, iv[1]), vert(A.verts, iv[2]), mu);
      if (k == 2) degenerate = 1u;
      if (k != 1) continue;
      const uint32_t packed = s_tslot[j];
      if (packed >> 24) continue;
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3)
        atomicAdd(&s_val[e][(packed >> (8 * c3)) & 0xFFu], phi * mu[c3]);
#pragma unroll
      for (int ee = 0; ee < 7; ++ee)
        if (ee == e) U[ee] += phi * ((mu[0] + mu[1]) + mu[2]);
    }
  }
  bool degen = __ballot(degenerate != 0u) != 0ull;
  const bool sums = A.far_sums != 0u;
  float inv[7], Usum[7];
  const float tau2 = A.tau * A.tau;
#pragma unroll
  for (int e = 0; e < 7; ++e) {
    const float Ue = (uint32_t)e < evals ? wave_sum_dpp(U[e]) : 0.0f;
    Usum[e] = Ue;
    inv[e] = Ue / (Ue * Ue + tau2);
  }
  float wv[VPL], gv[3][VPL];
#pragma unroll
  for (int k = 0; k < VPL; ++k) {
    const uint32_t v = lane + 64u * k;
    // the network's rows are read once: non-temporal loads (7 % faster
    // over 4M points)
    wv[k] = v >= nv ? 0.0f
                    : (A.Qh ? bf16_f32(__builtin_nontemporal_load(&A.Qh[qrow + v])) : A.Q[qrow + v]) * zs;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const size_t t = ((size_t)i * 3 + a) * A.ldq + v;
      gv[a][k] = (tan && v < nv)
                     ? (A.DQh ? bf16_f32(__builtin_nontemporal_load(&A.DQh[t])) : A.DQ[t]) * zs
                     : 0.0f;
    }
  }
  float sq = 0.0f, sdq[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < VPL; ++k) {
    sq += wv[k];
#pragma unroll
    for (int a = 0; a < 3; ++a) sdq[a] += gv[a][k];
  }
  const float Fsum = wave_sum_dpp(sq), keep = 1.0f - Fsum, inv_2h = 0.5f / A.h;
  float dQ[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) dQ[a] = tan ? wave_sum_dpp(sdq[a]) : 0.0f;
  if (sums) {
    // reciprocals of the whole sums at p and at p +- h e_a (far part
    // linearised); inv[] is reused for them
    float den[7];
    den[0] = Usum[0] + Fsum;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      den[1 + 2 * a] = Usum[1 + 2 * a] + (Fsum + A.h * dQ[a]);
      den[2 + 2 * a] = Usum[2 + 2 * a] + (Fsum - A.h * dQ[a]);
    }
    // the whole sum at an interior point is >= 4 pi / diagonal (>> tau); at
    // p +- h e_a, which may lie outside the cage, it may take either sign, as
    // in the exact binding's differences, but not vanish
    bool bad = !(den[0] > A.tau);
#pragma unroll
    for (int e = 1; e < 7; ++e) bad = bad || (grow && !(fabsf(den[e]) > A.tau));
    degen = degen || bad;
#pragma unroll
    for (int e = 0; e < 7; ++e) inv[e] = bad ? 0.0f : 1.0f / den[e];
  }
  __syncthreads();
  // the combined rows and the sums of the projection (sum w, sum w v; the
  // same for each g_a)
  float m4[4] = {0.f, 0.f, 0.f, 0.f}, mg[3][4] = {{0.f}};
#pragma unroll
  for (int k = 0; k < VPL; ++k) {
    const uint32_t v = lane + 64u * k;
    if (v >= nv) continue;
    const uint32_t sl = near_slot_find(s_key, v);
    const V3f cv = vert(A.verts, v);
    const float wv0 = wv[k];  // the network's value (far share or far sum)
    float w;
    if (sums) {
      w = ((sl != NEAR_EMPTY ? s_val[0][sl] : 0.0f) + wv0) * inv[0];
    } else {
      const float un = sl != NEAR_EMPTY ? s_val[0][sl] * inv[0] : 0.0f;
      w = keep * un + wv0;
    }
    wv[k] = w;
    m4[0] += w;
    m4[1] += w * cv.x;
    m4[2] += w * cv.y;
    m4[3] += w * cv.z;
    if (grow) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float g;
        if (sums) {
          const float hd = A.h * gv[a][k];
          const float np = (sl != NEAR_EMPTY ? s_val[1 + 2 * a][sl] : 0.0f) + (wv0 + hd);
          const float nm = (sl != NEAR_EMPTY ? s_val[2 + 2 * a][sl] : 0.0f) + (wv0 - hd);
          g = (np * inv[1 + 2 * a] - nm * inv[2 + 2 * a]) * inv_2h;
        } else {
          const float up = sl != NEAR_EMPTY ? s_val[1 + 2 * a][sl] * inv[1 + 2 * a] : 0.0f;
          const float um = sl != NEAR_EMPTY ? s_val[2 + 2 * a][sl] * inv[2 + 2 * a] : 0.0f;
          g = (keep * ((up - um) * inv_2h) - dQ[a] * (0.5f * (up + um))) + gv[a][k];
        }
        gv[a][k] = g;
        mg[a][0] += g;
        mg[a][1] += g * cv.x;
        mg[a][2] += g * cv.y;
        mg[a][3] += g * cv.z;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) m4[k] = wave_sum_dpp(m4[k]);
  if (grow)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int k = 0; k < 4; ++k) mg[a][k] = wave_sum_dpp(mg[a][k]);
  if (lane == 0) {
    const float rx = m4[1] - p.x, ry = m4[2] - p.y, rz = m4[3] - p.z;
    A.residual[i] = degen ? __builtin_inff() : sqrtf((rx * rx + ry * ry) + rz * rz);
  }
  const float r[4] = {1.0f - m4[0], p.x - m4[1], p.y - m4[2], p.z - m4[3]};
  float lam[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    lam[k] = ((A.minv[k * 4] * r[0] + A.minv[k * 4 + 1] * r[1]) + A.minv[k * 4 + 2] * r[2]) +
             A.minv[k * 4 + 3] * r[3];
  float la[3][4] = {{0.f}};
  if (grow)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float e4[4] = {-mg[a][0], (a == 0 ? 1.0f : 0.0f) - mg[a][1],
                           (a == 1 ? 1.0f : 0.0f) - mg[a][2], (a == 2 ? 1.0f : 0.0f) - mg[a][3]};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        la[a][k] = ((A.minv[k * 4] * e4[0] + A.minv[k * 4 + 1] * e4[1]) + A.minv[k * 4 + 2] * e4[2]) +
                   A.minv[k * 4 + 3] * e4[3];
    }
#pragma unroll
  for (int k = 0; k < VPL; ++k) {
    const uint32_t v = lane + 64u * k;
    if (v >= nv) continue;
    const V3f cv = vert(A.verts, v);
    // the rows are the binding's output, not read again by this launch:
    // non-temporal stores (2-3 % faster over 4M points)
    __builtin_nontemporal_store(
        wv[k] + (((lam[0] + lam[1] * cv.x) + lam[2] * cv.y) + lam[3] * cv.z), &row[v]);
    if (grow) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        __builtin_nontemporal_store(
            gv[a][k] + (((la[a][0] + la[a][1] * cv.x) + la[a][2] * cv.y) + la[a][3] * cv.z),
            &grow[(size_t)v * 3 + a]);
    }
  }
}

}  // namespace

void cgs_launch_mvc_select(const float* winding, const float* residual, float tol, uint32_t n,
                           uint32_t nv, uint32_t* inside, float* w, float* g, uint32_t* list,
                           uint32_t* count, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mvc_select, dim3((n + 255u) / 256u), dim3(256), 0, s, winding, residual, tol,
                     n, nv, inside, w, g, list, count);
}

size_t cgs_mvc_lds_bytes(uint32_t nv, uint32_t nt) {
  return (size_t)nv * 16 + (size_t)nt * 12 + (size_t)nv * 8;
}

void cgs_launch_mvc_bind(const CgsMvcArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_mvc_bind, dim3(a.n), dim3(64), cgs_mvc_lds_bytes(a.nv, a.nt), s, a);
}

void cgs_launch_mvc_far_share(const CgsMvcArgs& a, const float* cent, float r0, float r1,
                              float q_scale, int sums, float* out, hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_mvc_far_share, dim3(a.n), dim3(64), (size_t)a.nv * 16 + (size_t)a.nt * 12, s,
                     a, cent, r0, r1, q_scale, sums, out);
}

void cgs_launch_mvc_winding_lane(const float* verts, const uint32_t* tris, uint32_t nt,
                                 const float* pts, uint32_t n, float* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mvc_winding_lane, dim3((n + 255u) / 256u), dim3(256), 0, s, verts, tris, nt,
                     pts, n, out);
}

void cgs_launch_mvc_nearfar(const CgsNearArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  const size_t lds = 0;
  if (a.nv <= 256)
    hipLaunchKernelGGL(k_mvc_nearfar<4>, dim3(a.n), dim3(64), lds, s, a);
  else if (a.nv <= 512)
    hipLaunchKernelGGL(k_mvc_nearfar<8>, dim3(a.n), dim3(64), lds, s, a);
  else if (a.nv <= 768)
    hipLaunchKernelGGL(k_mvc_nearfar<12>, dim3(a.n), dim3(64), lds, s, a);
  else if (a.nv <= 1024)
    hipLaunchKernelGGL(k_mvc_nearfar<16>, dim3(a.n), dim3(64), lds, s, a);
  else
    hipLaunchKernelGGL(k_mvc_nearfar<32>, dim3(a.n), dim3(64), lds, s, a);
}

void cgs_launch_mvc_winding(const CgsMvcArgs& a, float* out, hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_mvc_winding, dim3(a.n), dim3(64), 0, s, a, out);
}

void cgs_launch_mvc_deform(const float* w, const float* g, const uint32_t* mask,
                           const float* rest_pos, const float* dverts, uint32_t nv, uint32_t n,
                           float* out_pos, float* out_F, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mvc_deform, dim3((n + 3u) / 4u), dim3(256), 0, s, w, g, mask, rest_pos,
                     dverts, nv, n, out_pos, out_F);
}

void cgs_launch_mvc_scene(const CgsMvcSceneArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_mvc_scene, dim3((a.n + 63u) / 64u), dim3(64), 0, s, a);
}

void cgs_launch_mvc_scene_children(const CgsMvcSceneArgs& a, const uint32_t* ranks,
                                   hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_mvc_scene_children, dim3((a.n + 255u) / 256u), dim3(256), 0, s, a, ranks);
}
