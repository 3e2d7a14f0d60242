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
  const uint32_t i = blockIdx.x;
  if (i >= A.n) return;
  if (A.only && A.only[i] == 0u) return;
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
                                                    uint32_t* inside, uint32_t* flags, float* w,
                                                    float* g, uint32_t* block_counts) {
  __shared__ uint32_t s_c[4];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  uint32_t f = 0u;
  if (i < n) {
    const bool in = fabsf(winding[i]) >= 0.5f;
    f = (in && residual[i] > tol) ? 1u : 0u;
    inside[i] = in ? 1u : 0u;
    flags[i] = f;
    if (!in) {
      for (uint32_t v = 0; v < nv; ++v) w[(size_t)i * nv + v] = 0.0f;
      if (g)
        for (uint32_t v = 0; v < 3 * nv; ++v) g[(size_t)i * nv * 3 + v] = 0.0f;
    }
  }
  const uint32_t c = (uint32_t)__popcll(__ballot(f != 0u));
  if ((threadIdx.x & 63u) == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
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

}  // namespace

uint32_t cgs_mvc_select_blocks(uint32_t n) { return (n + 255u) / 256u; }

void cgs_launch_mvc_select(const float* winding, const float* residual, float tol, uint32_t n,
                           uint32_t nv, uint32_t* inside, uint32_t* flags, float* w, float* g,
                           uint32_t* block_counts, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mvc_select, dim3(cgs_mvc_select_blocks(n)), dim3(256), 0, s, winding,
                     residual, tol, n, nv, inside, flags, w, g, block_counts);
}

size_t cgs_mvc_lds_bytes(uint32_t nv, uint32_t nt) {
  return (size_t)nv * 16 + (size_t)nt * 12 + (size_t)nv * 8;
}

void cgs_launch_mvc_bind(const CgsMvcArgs& a, hipStream_t s) {
  if (a.n == 0) return;
  hipLaunchKernelGGL(k_mvc_bind, dim3(a.n), dim3(64), cgs_mvc_lds_bytes(a.nv, a.nt), s, a);
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
