// k_composite.hip — Stage 4: 16x16 tile compositor into RGBA8.
//
// Follows tile_compositor_v2_kernel (src/render/kernels/compositor.rs:721-883):
// front-to-back alpha blending of the tile's depth-sorted splats, q clamped to
// <= 0, skip q < -4, alpha = min(0.99, o e^q), skip alpha < 1/255, saturation
// at T < 1e-3, background blend, alpha = 1 - T (bg alpha for an empty tile),
// RGBA8 = round(255 x) packed r | g<<8 | b<<16 | a<<24 into rows padded to
// 64 pixels (gpu_state.rs:25-27).
//
// MI355X design: one 256-thread block (4 wave64s) per tile; each batch of 256
// splats is gathered once into LDS as two float4 records + one float and read
// back as wave-wide broadcasts (no bank conflicts); the tile stops as soon as
// no pixel is alive (__syncthreads_or instead of the reference's
// double-buffered shared atomic flag).  The conic is stored pre-scaled by
// -1/2 (and b by -1), which is exact, so the Mahalanobis term q has the same
// bits as the reference expression; e^q uses the hardware exp2 (the parity
// bound for colour is 1/255).  Tiles are dealt to blocks so that each of the
// 8 XCDs (round-robin block placement) gets one contiguous band of tiles,
// keeping neighbouring tiles' shared splats in the same L2.
#include "cgs_internal.h"

namespace {

constexpr int CT = 256;

__global__ __launch_bounds__(CT) void k_composite(CgsCompositeArgs A, uint32_t per_xcd) {
  __shared__ float4 s_g[CT];   // -a/2, -b, -c/2, opacity
  __shared__ float4 s_m[CT];   // mu_x, mu_y, r, g
  __shared__ float s_b[CT];    // blue

  const uint32_t total_tiles = A.tiles_x * A.tiles_y;
  const uint32_t tile = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
  if (tile >= total_tiles) return;
  const uint32_t tx = tile % A.tiles_x, ty = tile / A.tiles_x;
  const uint32_t t = threadIdx.x;
  const uint32_t px = tx * 16u + (t & 15u), py = ty * 16u + (t >> 4);
  const bool inside = px < A.width && py < A.height;

  const uint32_t rs = A.ranges[tile * 2u], re = A.ranges[tile * 2u + 1u];
  const bool empty = rs >= re;
  const uint32_t n = empty ? 0u : re - rs;
  const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;

  float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = 1.0f;
  bool alive = inside;
  for (uint32_t b0 = 0; b0 < n; b0 += CT) {
    const uint32_t e = rs + b0 + t;
    if (e < re) {
      const uint32_t g = A.sorted_ids[e];
      const float4* gp = reinterpret_cast<const float4*>(A.geom + (size_t)g * 8);
      const float4 g0 = gp[0];
      const float4 g1 = gp[1];
      const float4 col = reinterpret_cast<const float4*>(A.color)[g];
      s_g[t] = make_float4(-0.5f * g0.z, -g0.w, -0.5f * g1.x, g1.y);
      s_m[t] = make_float4(g0.x, g0.y, col.x, col.y);
      s_b[t] = col.z;
    }
    __syncthreads();
    const uint32_t cnt = min((uint32_t)CT, n - b0);
    if (alive) {
      for (uint32_t k = 0; k < cnt; ++k) {
        const float4 gk = s_g[k];
        const float op = gk.w;
        if (!(op > 0.003921569f)) continue;
        const float4 mk = s_m[k];
        const float dx = pcx - mk.x;
        const float dy = pcy - mk.y;
        // == -0.5f * (a*dx*dx + 2*b*dx*dy + c*dy*dy) bit for bit
        float q = gk.x * dx * dx + gk.y * dx * dy + gk.z * dy * dy;
        if (q > 0.0f) q = 0.0f;
        if (q >= -4.0f) {
          const float alpha = fminf(0.99f, op * __expf(q));
          if (alpha >= 0.003921569f) {
            const float w = alpha * T;
            cr = __builtin_fmaf(w, mk.z, cr);
            cg = __builtin_fmaf(w, mk.w, cg);
            cb = __builtin_fmaf(w, s_b[k], cb);
            T = T * (1.0f - alpha);
            if (T < 0.001f) {
              T = 0.0f;
              alive = false;
              break;
            }
          }
        }
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
  const uint32_t per_xcd = (total + 7u) / 8u;
  hipLaunchKernelGGL(k_composite, dim3(per_xcd * 8u), dim3(CT), 0, s, a, per_xcd);
}
