// k_composite.hip — Stage 4: 16x16 tile compositor into RGBA8.
//
// Follows tile_compositor_v2_kernel (src/render/kernels/compositor.rs:721-883):
// front-to-back alpha blending of the tile's depth-sorted splats, q clamped to
// <= 0, skip q < -4, alpha = min(0.99, o e^q), skip alpha < 1/255, saturation
// at T < 1e-3, background blend, alpha = 1 - T (bg alpha for an empty tile),
// RGBA8 = round(255 x) packed r | g<<8 | b<<16 | a<<24 into rows padded to
// 64 pixels (gpu_state.rs:25-27).
//
// MI355X design:
//  * one 256-thread block (4 wave64s) per tile; wave w owns the 8x8 quadrant
//    (w&1, w>>1) of the tile, one pixel per lane;
//  * a batch of 256 splats is gathered once into LDS (two float4 + one float
//    per splat) and, by the loading thread, tested against the 4 quadrants
//    with the exact ellipse-vs-rectangle test the binning uses
//    (binning.rs:35-83, on the quadrant's pixel-centre rectangle); the four
//    results become 4 ballot masks per loading wave;
//  * each wave then walks only the set bits of its 256-bit mask, in order
//    (front to back), with a scalar loop: no divergent branches, and the
//    blend itself is branch-free (select instead of if/break);
//  * a wave stops when none of its pixels is alive, the tile stops after a
//    batch in which no pixel is alive (__syncthreads_or);
//  * the conic is stored pre-scaled by -1/2 (and b by -1), which is exact, so
//    q has the same bits as the reference expression; e^q uses the hardware
//    exp2 (the colour parity bound is 1/255);
//  * block b composites tile b: the hardware deals consecutive blocks to the 8
//    XCDs in turn, so XCD k gets the tile columns x = k (mod 8) (the row
//    length 120 at 1080p is a multiple of 8) and the dense middle of the
//    image is spread over all eight; vertical neighbours share an L2;
//  * the splat ids of the next batch are loaded while the current batch is
//    composited, so a batch waits for one memory round trip, not two.
#include "cgs_internal.h"

namespace {

constexpr int CT = 256;
constexpr float ALPHA_MIN = 0.003921569f;  // the reference's literal

// ellipse_intersects_tile_gpu (binning.rs:35-83) on an arbitrary pixel rectangle
__device__ __forceinline__ bool ellipse_hits_rect(float mu_x, float mu_y, float a, float b,
                                                  float c, float alpha, float qcut, uint32_t xs,
                                                  uint32_t ys, uint32_t w, uint32_t h,
                                                  uint32_t W, uint32_t H) {
  if (alpha < (1.0f / 255.0f) || xs >= W || ys >= H) return false;
  const uint32_t xe = min(xs + w - 1u, W - 1u), ye = min(ys + h - 1u, H - 1u);
  const float dx0 = ((float)xs + 0.5f) - mu_x;
  const float dx1 = ((float)xe + 0.5f) - mu_x;
  const float dy0 = ((float)ys + 0.5f) - mu_y;
  const float dy1 = ((float)ye + 0.5f) - mu_y;
  const float x_near = cgs_fclamp(0.0f, dx0, dx1);
  const float y1 = cgs_fclamp(-b * x_near / cgs_fmax(c, 1e-12f), dy0, dy1);
  const float q1 = a * x_near * x_near + 2.0f * b * x_near * y1 + c * y1 * y1;
  const float y_near = cgs_fclamp(0.0f, dy0, dy1);
  const float x2 = cgs_fclamp(-b * y_near / cgs_fmax(a, 1e-12f), dx0, dx1);
  const float q2 = a * x2 * x2 + 2.0f * b * x2 * y_near + c * y_near * y_near;
  return cgs_fmin(q1, q2) <= qcut + 1e-4f;
}

__global__ __launch_bounds__(CT) void k_composite(CgsCompositeArgs A) {
  __shared__ float4 s_g[CT];  // -a/2, -b, -c/2, opacity (0 when <= 1/255)
  __shared__ float4 s_m[CT];  // mu_x, mu_y, r, g
  __shared__ float s_b[CT];   // blue
  __shared__ uint64_t s_mask[4][CT / 64];  // [quadrant][loading wave]

  const uint32_t total_tiles = A.tiles_x * A.tiles_y;
  const uint32_t tile = blockIdx.x;
  if (tile >= total_tiles) return;
  const uint32_t tx = tile % A.tiles_x, ty = tile / A.tiles_x;
  const uint32_t t = threadIdx.x;
  const uint32_t wave = t >> 6, lane = t & 63u;
  const uint32_t qx0 = tx * 16u + (wave & 1u) * 8u, qy0 = ty * 16u + (wave >> 1) * 8u;
  const uint32_t px = qx0 + (lane & 7u), py = qy0 + (lane >> 3);
  const bool inside = px < A.width && py < A.height;

  const uint32_t rs = A.ranges[tile * 2u], re = A.ranges[tile * 2u + 1u];
  const bool empty = rs >= re;
  const uint32_t n = empty ? 0u : re - rs;
  const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;

  float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = 1.0f;
  bool alive = inside;
  uint32_t g_next = (rs + t < re) ? A.sorted_ids[rs + t] : 0u;
  for (uint32_t b0 = 0; b0 < n; b0 += CT) {
    // ---- gather the batch and test it against the 4 quadrants ----
    const uint32_t e = rs + b0 + t;
    const uint32_t g = g_next;
    if (e + (uint32_t)CT < re) g_next = A.sorted_ids[e + CT];  // next batch's id
    bool q0 = false, q1 = false, q2 = false, q3 = false;
    if (e < re) {
      const float4* gp = reinterpret_cast<const float4*>(A.geom + (size_t)g * 8);
      const float4 g0 = gp[0];  // mu_x, mu_y, a, b
      const float4 g1 = gp[1];  // c, alpha, tz, q_cut
      const float4 col = reinterpret_cast<const float4*>(A.color)[g];
      const float op = g1.y > ALPHA_MIN ? g1.y : 0.0f;
      s_g[t] = make_float4(-0.5f * g0.z, -g0.w, -0.5f * g1.x, op);
      s_m[t] = make_float4(g0.x, g0.y, col.x, col.y);
      s_b[t] = col.z;
      const uint32_t bx = tx * 16u, by = ty * 16u;
      q0 = ellipse_hits_rect(g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.w, bx, by, 8u, 8u,
                             A.width, A.height);
      q1 = ellipse_hits_rect(g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.w, bx + 8u, by, 8u, 8u,
                             A.width, A.height);
      q2 = ellipse_hits_rect(g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.w, bx, by + 8u, 8u, 8u,
                             A.width, A.height);
      q3 = ellipse_hits_rect(g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.w, bx + 8u, by + 8u, 8u,
                             8u, A.width, A.height);
    }
    const uint64_t m0 = __ballot(q0), m1 = __ballot(q1), m2 = __ballot(q2), m3 = __ballot(q3);
    if (lane == 0) {
      s_mask[0][wave] = m0;
      s_mask[1][wave] = m1;
      s_mask[2][wave] = m2;
      s_mask[3][wave] = m3;
    }
    __syncthreads();

    // ---- this wave's quadrant: only the splats that can touch it ----
    if (__ballot(alive) != 0ull) {
#pragma unroll 1
      for (uint32_t v = 0; v < (uint32_t)(CT / 64); ++v) {
        // the mask is wave-uniform; readfirstlane returns int, so widen through
        // uint32_t (a sign extension would set the upper 32 bits)
        const uint64_t mv = s_mask[wave][v];
        const uint32_t mlo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)mv);
        const uint32_t mhi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(mv >> 32));
        uint64_t m = ((uint64_t)mhi << 32) | (uint64_t)mlo;
        // two splats per iteration: their LDS reads and q / e^q chains are
        // independent, only the blend is sequential (front to back)
        while (m != 0ull) {
          const uint32_t k0 = v * 64u + (uint32_t)__builtin_ctzll(m);
          m &= m - 1ull;
          const bool two = m != 0ull;  // wave-uniform
          const uint32_t k1 = two ? v * 64u + (uint32_t)__builtin_ctzll(m) : k0;
          m = two ? (m & (m - 1ull)) : m;
          const float4 g0 = s_g[k0], m0 = s_m[k0];
          const float4 g1 = s_g[k1], m1 = s_m[k1];
          const float b0 = s_b[k0], b1 = s_b[k1];
          const float dx0 = pcx - m0.x, dy0 = pcy - m0.y;
          const float dx1 = pcx - m1.x, dy1 = pcy - m1.y;
          // == -0.5f * (a*dx*dx + 2*b*dx*dy + c*dy*dy) bit for bit
          float q0 = g0.x * dx0 * dx0 + g0.y * dx0 * dy0 + g0.z * dy0 * dy0;
          float q1 = g1.x * dx1 * dx1 + g1.y * dx1 * dy1 + g1.z * dy1 * dy1;
          q0 = q0 > 0.0f ? 0.0f : q0;
          q1 = q1 > 0.0f ? 0.0f : q1;
          const float a0 = fminf(0.99f, g0.w * __expf(q0));
          const float a1 = fminf(0.99f, g1.w * __expf(q1));
          {
            const bool ok = alive && q0 >= -4.0f && a0 >= ALPHA_MIN;
            const float w = ok ? a0 * T : 0.0f;
            cr = __builtin_fmaf(w, m0.z, cr);
            cg = __builtin_fmaf(w, m0.w, cg);
            cb = __builtin_fmaf(w, b0, cb);
            const float Tn = T * (1.0f - a0);
            const bool sat = ok && Tn < 0.001f;
            T = ok ? (sat ? 0.0f : Tn) : T;
            alive = alive && !sat;
          }
          {
            const bool ok = two && alive && q1 >= -4.0f && a1 >= ALPHA_MIN;
            const float w = ok ? a1 * T : 0.0f;
            cr = __builtin_fmaf(w, m1.z, cr);
            cg = __builtin_fmaf(w, m1.w, cg);
            cb = __builtin_fmaf(w, b1, cb);
            const float Tn = T * (1.0f - a1);
            const bool sat = ok && Tn < 0.001f;
            T = ok ? (sat ? 0.0f : Tn) : T;
            alive = alive && !sat;
          }
          if (__ballot(alive) == 0ull) break;
        }
        if (__ballot(alive) == 0ull) break;
      }
    }
    if (!__syncthreads_or(alive)) break;
  }

  if (inside) {
    const float fr = cgs_fclamp(cr + T * A.bg[0], 0.0f, 1.0f);
    const float fg = cgs_fclamp(cg + T * A.bg[1], 0.0f, 1.0f);
    const float fb = cgs_fclamp(cb + T * A.bg[2], 0.0f, 1.0f);
    float fa = cgs_fclamp(1.0f - T, 0.0f, 1.0f);
    if (empty) fa = A.bg[3];
    const uint32_t r8 = (uint32_t)cgs_fclamp(roundf(fr * 255.0f), 0.0f, 255.0f);
    const uint32_t g8 = (uint32_t)cgs_fclamp(roundf(fg * 255.0f), 0.0f, 255.0f);
    const uint32_t b8 = (uint32_t)cgs_fclamp(roundf(fb * 255.0f), 0.0f, 255.0f);
    const uint32_t a8 = (uint32_t)cgs_fclamp(roundf(fa * 255.0f), 0.0f, 255.0f);
    A.out[(size_t)py * A.padded_row + px] = r8 | (g8 << 8) | (b8 << 16) | (a8 << 24);
  }
}

}  // namespace

void cgs_launch_composite(const CgsCompositeArgs& a, hipStream_t s) {
  const uint32_t total = a.tiles_x * a.tiles_y;
  if (total == 0) return;
  hipLaunchKernelGGL(k_composite, dim3(total), dim3(CT), 0, s, a);
}
