// k_composite.hip — Stage 4: tile compositor into RGBA8 (16x16 tiles as the
// reference; 8x8, 16x32 and 32x16 as tuning choices).
//
// Follows tile_compositor_v2_kernel (src/render/kernels/compositor.rs:721-883):
// front-to-back alpha blending of the tile's depth-sorted splats, q clamped to
// <= 0, skip q < -4, alpha = min(0.99, o e^q), skip alpha < 1/255, saturation
// at T < 1e-3, background blend, alpha = 1 - T (bg alpha for an empty tile),
// RGBA8 = round(255 x) packed r | g<<8 | b<<16 | a<<24 into rows padded to
// 64 pixels (gpu_state.rs:25-27).
//
// MI355X design:
//  * one block per tile of QX x QY quadrants of 8x8 pixels (16x16: 2 x 2,
//    256 threads); wave w owns quadrant (w % QX, w / QX), one pixel per lane;
//  * k_composite_q (the default, composite_batch 64): each quadrant's wave
//    walks the tile's list on its own, 64 splats per batch, with no block
//    barrier, and leaves as soon as its 64 pixels are saturated;
//  * k_composite (batches of 128-512 per block): a batch of CB splats is
//    gathered once into LDS (two float4 + one float
//    per splat) and, by the loading thread, tested against the QX x QY
//    quadrants with the exact ellipse-vs-rectangle test the binning uses
//    (binning.rs:35-83, on the quadrant's pixel-centre rectangle); the
//    results become one ballot mask per quadrant and loading wave;
//  * each wave compacts the set bits of its 256-bit mask into an LDS list of
//    splat slots (order kept) and walks the list front to back four splats
//    per step: the four LDS reads and q / e^q chains are independent, only
//    the blend is sequential; the blend is branch-free (select, not if/break);
//  * a wave stops when none of its pixels is alive, the tile stops after a
//    batch in which no pixel is alive (__syncthreads_or);
//  * q = -1/2 (a dx^2 + 2 b dx dy + c dy^2) is computed with the reference's
//    operations in the reference's order, on the conic stored as -a/2, -b,
//    -c/2 (power-of-two scalings: the same bits): the cut q >= -4, where the
//    blended alpha jumps by up to 0.99 e^-4, falls on the same side as in the
//    oracle; e^q is the hardware exp2 of q log2(e) and the blend folds
//    T (1 - alpha) into T - alpha T, which change only rounding (the colour
//    parity bound is 1/255);
//  * block b composites tile b: the hardware deals consecutive blocks to the 8
//    XCDs in turn, so XCD k gets the tile columns x = k (mod 8) (the row
//    length 120 at 1080p is a multiple of 8) and the dense middle of the
//    image is spread over all eight; vertical neighbours share an L2;
//  * the records of the next batch are loaded (and the ids of the one after)
//    while the current batch is composited, so the memory round trips of a
//    batch overlap the previous batch's blending.
#include "cgs_internal.h"

namespace {

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

// The same test with the two divisions as hardware reciprocals and the
// margin doubled (2e-4): it accepts every splat the exact test accepts (a
// reciprocal moves the line minimum y1 or x2 by about one ulp, which changes
// q there only at second order, ~1e-8 of q).  The compositor may take a
// superset of a quadrant's splats: each pixel still applies the exact q and
// alpha cuts, so the frame is the same bit for bit.
__device__ __forceinline__ bool ellipse_may_hit_rect(float mu_x, float mu_y, float a, float b,
                                                     float c, float alpha, float qcut,
                                                     uint32_t xs, uint32_t ys, uint32_t w,
                                                     uint32_t h, uint32_t W, uint32_t H) {
  if (alpha < (1.0f / 255.0f) || xs >= W || ys >= H) return false;
  const uint32_t xe = min(xs + w - 1u, W - 1u), ye = min(ys + h - 1u, H - 1u);
  const float dx0 = ((float)xs + 0.5f) - mu_x;
  const float dx1 = ((float)xe + 0.5f) - mu_x;
  const float dy0 = ((float)ys + 0.5f) - mu_y;
  const float dy1 = ((float)ye + 0.5f) - mu_y;
  const float x_near = cgs_fclamp(0.0f, dx0, dx1);
  const float y1 = cgs_fclamp(-b * x_near * __builtin_amdgcn_rcpf(cgs_fmax(c, 1e-12f)), dy0, dy1);
  const float q1 = a * x_near * x_near + 2.0f * b * x_near * y1 + c * y1 * y1;
  const float y_near = cgs_fclamp(0.0f, dy0, dy1);
  const float x2 = cgs_fclamp(-b * y_near * __builtin_amdgcn_rcpf(cgs_fmax(a, 1e-12f)), dx0, dx1);
  const float q2 = a * x2 * x2 + 2.0f * b * x2 * y_near + c * y_near * y_near;
  return cgs_fmin(q1, q2) <= qcut + 2e-4f;
}

// One splat's compositor inputs as staged in LDS.
struct Splat {
  float4 g;  // -a/2, -b, -c/2, opacity (0 when <= 1/255)
  float4 m;  // mu_x, mu_y, r, g
  float b;   // blue
};

constexpr float LOG2E = 1.4426950408889634f;

// min(q, 0) with q = -1/2 (a dx^2 + 2 b dx dy + c dy^2) in the order of
// compositor.rs (a dx dx + 2 b dx dy + c dy dy, left to right, then the
// factor): with the factor folded into the staged conic the products and sums
// round to the same bits (no fused operations: -ffp-contract=off)
__device__ __forceinline__ float splat_q(const Splat& s, float pcx, float pcy) {
  const float dx = pcx - s.m.x, dy = pcy - s.m.y;
  const float q = (s.g.x * dx * dx + s.g.y * dx * dy) + s.g.z * dy * dy;
  return fminf(q, 0.0f);
}

// Front-to-back blend of one splat into the pixel state (branch-free):
// w = alpha T, C += w c, T -= w (= T (1 - alpha)); T < 1e-3 saturates the
// pixel (T = 0) after the splat is added.  A pixel that is not alive has
// ok = false, so w = 0 and its T stays 0 (saturated) or 1 (off screen).
__device__ __forceinline__ void blend(const Splat& s, float q, float a, bool valid, bool& alive,
                                      float& T, float& cr, float& cg, float& cb) {
  const bool ok = valid && alive && q >= -4.0f && a >= ALPHA_MIN;
  const float w = (ok ? a : 0.0f) * T;
  cr = __builtin_fmaf(w, s.m.z, cr);
  cg = __builtin_fmaf(w, s.m.w, cg);
  cb = __builtin_fmaf(w, s.b, cb);
  const float Tn = T - w;
  const bool sat = Tn < 0.001f;
  T = sat ? 0.0f : Tn;
  alive = alive && !sat;
}

// blend for a pixel whose liveness is T > 0 (k_composite_q): the same
// arithmetic as blend() on a live pixel; a pixel at T = 0 adds w = 0 and
// stays at 0
__device__ __forceinline__ void blend_t(const Splat& s, float q, float a, float& T, float& cr,
                                        float& cg, float& cb) {
  const bool ok = q >= -4.0f && a >= ALPHA_MIN;
  const float w = (ok ? a : 0.0f) * T;
  cr = __builtin_fmaf(w, s.m.z, cr);
  cg = __builtin_fmaf(w, s.m.w, cg);
  cb = __builtin_fmaf(w, s.b, cb);
  const float Tn = T - w;
  T = Tn < 0.001f ? 0.0f : Tn;
}

template <int CB>
__device__ __forceinline__ Splat staged(const float4 (*st)[CB], uint32_t off) {
  const char* p = reinterpret_cast<const char*>(&st[0][0]) + off;
  Splat s;
  s.g = *reinterpret_cast<const float4*>(p);
  s.m = *reinterpret_cast<const float4*>(p + CB * 16);
  s.b = *reinterpret_cast<const float*>(p + 2 * CB * 16);
  return s;
}

// Dispatch order of the tiles: the heaviest first (entries > 2048, then
// > 512, then > 64, then the rest), so that the few tiles that take longest
// start at the beginning of the launch.  One 1024-thread block; thread t owns
// the contiguous tiles [t*per, (t+1)*per) (per <= 32), so one block scan of the
// per-thread class counts orders every class by tile index.  Counts of two
// classes share a word (tiles <= 32768 < 2^16).
__global__ __launch_bounds__(1024) void k_tile_order(const uint32_t* ranges, uint32_t tiles,
                                                     uint32_t* order) {
  __shared__ uint32_t s_tmp[16];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (tiles + 1023u) / 1024u;
  const uint32_t t0 = min(t * per, tiles), t1 = min(t0 + per, tiles);
  uint64_t cls = 0;  // 2 bits per owned tile
  uint32_t c01 = 0, c23 = 0;
  // unrolled so that a thread's range loads are in flight together
#pragma unroll 8
  for (uint32_t i = t0; i < t1; ++i) {
    const uint2 r = reinterpret_cast<const uint2*>(ranges)[i];
    const uint32_t n = r.y > r.x ? r.y - r.x : 0u;
    const uint32_t k = n > 2048u ? 0u : n > 512u ? 1u : n > 64u ? 2u : 3u;
    cls |= (uint64_t)k << (2u * (i - t0));
    c01 += k == 0u ? 1u : k == 1u ? 0x10000u : 0u;
    c23 += k == 2u ? 1u : k == 3u ? 0x10000u : 0u;
  }
  uint32_t a, c;
  const uint32_t o01 = cgs::block_exclusive_scan<1024>(c01, s_tmp, a);
  const uint32_t o23 = cgs::block_exclusive_scan<1024>(c23, s_tmp, c);
  uint32_t pos[4];
  pos[0] = o01 & 0xFFFFu;
  pos[1] = (a & 0xFFFFu) + (o01 >> 16);
  pos[2] = (a & 0xFFFFu) + (a >> 16) + (o23 & 0xFFFFu);
  pos[3] = (a & 0xFFFFu) + (a >> 16) + (c & 0xFFFFu) + (o23 >> 16);
  for (uint32_t i = t0; i < t1; ++i) {
    const uint32_t k = (uint32_t)(cls >> (2u * (i - t0))) & 3u;
    order[pos[k]++] = i;
  }
}

// Splats per step of the inner loop and the next batch's prefetch, measured
// on the 1.07M-splat 1080p orbit (k_composite, batch 256): 4 splats with
// prefetch 216 us (82 VGPRs, 5 waves per SIMD), without 219 us; 2 splats
// 246 us with or without (64-66 VGPRs, 7-8 waves per SIMD).  The wave is
// VALU bound (about 26 VALU instructions per splat evaluation, most of them
// the reference's q and the sequential blend), so the independent q / e^q
// chains of 4 splats are worth more than the occupancy of 2.
constexpr int COMPOSITE_STEP = 4;
constexpr bool COMPOSITE_PREFETCH = true;

// Tile of QX x QY quadrants of 8x8 pixels (64 x QX x QY threads); CB splats
// per batch (64, 128 or 256, at most the thread count), staged by threads
// t < CB; STEP splats walked per step of the inner loop; PF: the records of
// the next batch are loaded while this one is composited
template <int QX, int QY, int CB, int STEP, bool PF>
__global__ __launch_bounds__(64 * QX * QY) void k_composite(CgsCompositeArgs A) {
  constexpr int NQ = QX * QY;
  constexpr int NT = 64 * NQ;                    // threads
  constexpr int SU = CB > NT ? CB / NT : 1;      // splats staged per thread
  static_assert(CB <= NT || CB % NT == 0, "a batch is staged by SU splats per thread");
  // staged batch: plane 0 = Splat::g, 1 = Splat::m, 2 = (blue, -, -, -); the
  // quadrant lists hold byte offsets into a plane, so a splat's three reads
  // share one address register
  __shared__ float4 s_stage[3][CB];
  __shared__ uint64_t s_mask[NQ][CB / 64];     // [quadrant][loading wave]
  __shared__ __align__(16) uint32_t s_list[NQ][CB];  // [quadrant] splats that can touch it
  const bool loader = threadIdx.x < (uint32_t)CB;  // (every thread when SU > 1)

  const uint32_t total_tiles = A.tiles_x * A.tiles_y;
  if (blockIdx.x >= total_tiles) return;
  const uint32_t tile = A.order ? A.order[blockIdx.x] : blockIdx.x;
  const uint32_t tx = tile % A.tiles_x, ty = tile / A.tiles_x;
  const uint32_t t = threadIdx.x;
  const uint32_t wave = t >> 6, lane = t & 63u;
  const uint32_t bx = tx * (8u * QX), by = ty * (8u * QY);
  const uint32_t qx0 = bx + (wave % QX) * 8u, qy0 = by + (wave / QX) * 8u;
  const uint32_t px = qx0 + (lane & 7u), py = qy0 + (lane >> 3);
  const bool inside = px < A.width && py < A.height;

  const uint32_t rs = A.ranges[tile * 2u], re = A.ranges[tile * 2u + 1u];
  const bool empty = rs >= re;
  const uint32_t n = empty ? 0u : re - rs;
  const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
  const float4* geom4 = reinterpret_cast<const float4*>(A.geom);
  const float4* color4 = reinterpret_cast<const float4*>(A.color);

  // with PF, the records of the batch about to be staged are loaded one batch
  // ahead.  The loads are unconditional, at indexes clamped to the tile's last
  // entry, so that nothing waits for them before they are used a batch later.
  float4 r0[SU], r1[SU], rc[SU];
  uint32_t g_ahead[SU];
  const uint32_t last = empty ? rs : re - 1u;
#pragma unroll
  for (int u = 0; u < SU; ++u) {
    r0[u] = r1[u] = rc[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    g_ahead[u] = 0u;
    const uint32_t slot = t + (uint32_t)(u * NT);
    if (PF && !empty && loader) {
      const uint32_t g = A.sorted_ids[min(rs + slot, last)];
      r0[u] = geom4[(size_t)g * 2];
      r1[u] = geom4[(size_t)g * 2 + 1];
      rc[u] = color4[g];
      g_ahead[u] = A.sorted_ids[min(rs + slot + (uint32_t)CB, last)];
    }
  }

  float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = 1.0f;
  bool alive = inside;
  for (uint32_t b0 = 0; b0 < n; b0 += CB) {
    // ---- stage the batch and test it against the quadrants ----
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const uint32_t slot = t + (uint32_t)(u * NT);
      const uint32_t e = rs + b0 + slot;
      bool qh[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) qh[q] = false;
      if (!PF && loader && e < re) {
        const uint32_t g = A.sorted_ids[e];
        r0[u] = geom4[(size_t)g * 2];
        r1[u] = geom4[(size_t)g * 2 + 1];
        rc[u] = color4[g];
      }
      if (loader && e < re) {
        // r0 = mu_x, mu_y, a, b; r1 = c, alpha, tz, q_cut; rc = r, g, b
        const float4 a0 = r0[u], a1 = r1[u], ac = rc[u];
        const float op = a1.y > ALPHA_MIN ? a1.y : 0.0f;
        s_stage[0][slot] = make_float4(-0.5f * a0.z, -a0.w, -0.5f * a1.x, op);
        s_stage[1][slot] = make_float4(a0.x, a0.y, ac.x, ac.y);
        reinterpret_cast<float*>(&s_stage[2][slot])[0] = ac.z;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          qh[q] = ellipse_hits_rect(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.w, bx + (q % QX) * 8u,
                                    by + (q / QX) * 8u, 8u, 8u, A.width, A.height);
      }
      const uint32_t lw = wave + (uint32_t)(u * (NT / 64));  // the loading wave of this slot
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const uint64_t mq = __ballot(qh[q]);
        if (lane == 0 && lw < (uint32_t)(CB / 64)) s_mask[q][lw] = mq;
      }
      // the next batch's records are loaded while this batch is composited
      if (PF && loader) {
        const uint32_t en = e + CB;
        r0[u] = geom4[(size_t)g_ahead[u] * 2];
        r1[u] = geom4[(size_t)g_ahead[u] * 2 + 1];
        rc[u] = color4[g_ahead[u]];
        g_ahead[u] = A.sorted_ids[min(en + (uint32_t)CB, last)];
      }
    }
    __syncthreads();

    // ---- compact this wave's quadrant list (splat order kept) ----
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t v = 0; v < (uint32_t)(CB / 64); ++v) {
      const uint64_t mv = s_mask[wave][v];
      if ((mv >> lane) & 1ull) s_list[wave][cnt + cgs::mask_rank(mv)] = (v * 64u + lane) * 16u;
      cnt += (uint32_t)__popcll(mv);
    }
    cnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)cnt);
    __builtin_amdgcn_wave_barrier();

    // ---- walk it STEP splats at a time, front to back ----
    if (__ballot(alive) != 0ull) {
#pragma unroll 1
      for (uint32_t i = 0; i < cnt; i += (uint32_t)STEP) {
        uint32_t k[STEP];
        if (STEP == 4) {
          const uint4 id4 = *reinterpret_cast<const uint4*>(&s_list[wave][i]);
          k[0] = id4.x;
          k[1 % STEP] = id4.y;
          k[2 % STEP] = id4.z;
          k[3 % STEP] = id4.w;
        } else {
          const uint2 id2 = *reinterpret_cast<const uint2*>(&s_list[wave][i]);
          k[0] = id2.x;
          k[1 % STEP] = id2.y;
        }
        // wave-uniform: entries past cnt repeat the first (always valid) one
        bool v[STEP];
        Splat S[STEP];
        float q[STEP], al[STEP];
#pragma unroll
        for (int j = 0; j < STEP; ++j) {
          v[j] = i + (uint32_t)j < cnt;
          S[j] = staged<CB>(s_stage, v[j] ? k[j] : k[0]);
          q[j] = splat_q(S[j], pcx, pcy);
          al[j] = fminf(0.99f, S[j].g.w * __builtin_amdgcn_exp2f(q[j] * LOG2E));
        }
#pragma unroll
        for (int j = 0; j < STEP; ++j) blend(S[j], q[j], al[j], v[j], alive, T, cr, cg, cb);
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

// Tile of QX x QY quadrants of 8x8 pixels with one wave per quadrant and no
// block barrier: each wave walks the tile's whole list on its own, 64 splats
// per batch (one per lane: loaded, tested against the wave's own quadrant
// with the same exact ellipse test, staged in the wave's own LDS region),
// compacts the hits and composites them as the block-wide kernel does.  A
// wave whose pixels are all saturated leaves at once while the tile's other
// quadrants go on.  Each wave reads the tile's records itself; the waves of a
// block share a CU, so the repeats are L1 / L2 hits.
template <int QX, int QY, int STEP>
__global__ __launch_bounds__(64 * QX * QY) void k_composite_q(CgsCompositeArgs A) {
  constexpr int NQ = QX * QY;
  static_assert(STEP <= 4, "the list holds 4 entries past its 64 hits");
  // slot 64 of each wave's region is a splat of opacity 0 (it blends
  // nothing): the list entries past a batch's hit count point at it, so the
  // walk needs no test of the entry count
  __shared__ float4 s_stage[NQ][3][65];
  __shared__ __align__(16) uint32_t s_list[NQ][68];

  const uint32_t total_tiles = A.tiles_x * A.tiles_y;
  if (blockIdx.x >= total_tiles) return;
  const uint32_t tile = A.order ? A.order[blockIdx.x] : blockIdx.x;
  const uint32_t tx = tile % A.tiles_x, ty = tile / A.tiles_x;
  const uint32_t t = threadIdx.x;
  const uint32_t wave = t >> 6, lane = t & 63u;
  const uint32_t qx0 = tx * (8u * QX) + (wave % QX) * 8u, qy0 = ty * (8u * QY) + (wave / QX) * 8u;
  const uint32_t px = qx0 + (lane & 7u), py = qy0 + (lane >> 3);
  const bool inside = px < A.width && py < A.height;

  const uint32_t rs = A.ranges[tile * 2u], re = A.ranges[tile * 2u + 1u];
  const bool empty = rs >= re;
  const uint32_t n = empty ? 0u : re - rs;
  const float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
  const float4* geom4 = reinterpret_cast<const float4*>(A.geom);
  const float4* color4 = reinterpret_cast<const float4*>(A.color);
  float4(*st)[65] = s_stage[wave];
  const uint32_t wbase = wave * (uint32_t)sizeof(s_stage[0]);

  // a pixel is live while T > 0: an off-screen pixel starts at T = 0 and a
  // saturated one is set to 0, and a pixel at T = 0 blends nothing (w = a T)
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = inside ? 1.0f : 0.0f;
  // a quadrant with no pixel on screen has nothing to composite
  if (!empty && __ballot(inside) != 0ull) {
    // the records of the batch about to be tested are loaded one batch ahead,
    // at indexes clamped to the tile's last entry
    const uint32_t last = re - 1u;
    uint32_t g = A.sorted_ids[min(rs + lane, last)];
    float4 r0 = geom4[(size_t)g * 2], r1 = geom4[(size_t)g * 2 + 1], rc = color4[g];
    uint32_t g_ahead = A.sorted_ids[min(rs + lane + 64u, last)];
    if (lane == 0) {
      st[0][64] = make_float4(0.f, 0.f, 0.f, 0.f);
      st[1][64] = make_float4(0.f, 0.f, 0.f, 0.f);
      st[2][64] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    for (uint32_t b0 = 0; b0 < n; b0 += 64u) {
      const uint32_t e = rs + b0 + lane;
      bool hit = false;
      __builtin_amdgcn_wave_barrier();
      if (e < re) {
        // r0 = mu_x, mu_y, a, b; r1 = c, alpha, tz, q_cut; rc = r, g, b
        const float op = r1.y > ALPHA_MIN ? r1.y : 0.0f;
        st[0][lane] = make_float4(-0.5f * r0.z, -r0.w, -0.5f * r1.x, op);
        st[1][lane] = make_float4(r0.x, r0.y, rc.x, rc.y);
        reinterpret_cast<float*>(&st[2][lane])[0] = rc.z;
        hit = ellipse_may_hit_rect(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.w, qx0, qy0, 8u, 8u,
                                   A.width, A.height);
      }
      // the next batch's records load while this one is composited
      r0 = geom4[(size_t)g_ahead * 2];
      r1 = geom4[(size_t)g_ahead * 2 + 1];
      rc = color4[g_ahead];
      g_ahead = A.sorted_ids[min(e + 128u, last)];
      const uint64_t m = __ballot(hit);
      // list entries are byte offsets from s_stage[0] (this wave's region
      // included), so the walk's LDS addresses need no add
      const uint32_t cnt = (uint32_t)__popcll(m);
      if (hit) s_list[wave][cgs::mask_rank(m)] = wbase + lane * 16u;
      if (lane < (uint32_t)STEP) s_list[wave][cnt + lane] = wbase + 64u * 16u;
      __builtin_amdgcn_wave_barrier();
#pragma unroll 1
      for (uint32_t i = 0; i < cnt; i += (uint32_t)STEP) {
        uint32_t k[STEP];
#pragma unroll
        for (int j = 0; j < STEP; ++j) k[j] = s_list[wave][i + (uint32_t)j];
        Splat S[STEP];
        float q[STEP], al[STEP];
#pragma unroll
        for (int j = 0; j < STEP; ++j) {
          S[j] = staged<65>(s_stage[0], k[j]);
          q[j] = splat_q(S[j], pcx, pcy);
          al[j] = fminf(0.99f, S[j].g.w * __builtin_amdgcn_exp2f(q[j] * LOG2E));
        }
#pragma unroll
        for (int j = 0; j < STEP; ++j) blend_t(S[j], q[j], al[j], T, cr, cg, cb);
        if (__ballot(T > 0.0f) == 0ull) break;
      }
      if (__ballot(T > 0.0f) == 0ull) break;
    }
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

void cgs_launch_composite(const CgsCompositeArgs& a, hipStream_t s, CgsProf* prof,
                          uint32_t batch) {
  const uint32_t total = a.tiles_x * a.tiles_y;
  if (total == 0) return;
  CgsCompositeArgs A = a;
  if (A.order && total <= 32768u) {
    if (prof) prof->mark("k_tile_order", s);
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, s, A.ranges, total, A.order);
  } else {
    A.order = nullptr;  // tile order (8K frames: one block would take too long)
  }
  if (prof) prof->mark("k_composite", s);
  constexpr int ST = COMPOSITE_STEP;
  constexpr bool PF = COMPOSITE_PREFETCH;
  const dim3 grid(total);
  if (A.tile_w == 8u && A.tile_h == 8u) {
    hipLaunchKernelGGL((k_composite<1, 1, 64, ST, PF>), grid, dim3(64), 0, s, A);
  } else if (A.tile_w == 16u && A.tile_h == 32u) {
    if (batch == 64u)
      hipLaunchKernelGGL((k_composite_q<2, 4, ST>), grid, dim3(512), 0, s, A);
    else if (batch == 128u)
      hipLaunchKernelGGL((k_composite<2, 4, 128, ST, PF>), grid, dim3(512), 0, s, A);
    else
      hipLaunchKernelGGL((k_composite<2, 4, 256, ST, PF>), grid, dim3(512), 0, s, A);
  } else if (A.tile_w == 32u && A.tile_h == 16u) {
    if (batch == 64u)
      hipLaunchKernelGGL((k_composite_q<4, 2, ST>), grid, dim3(512), 0, s, A);
    else if (batch == 128u)
      hipLaunchKernelGGL((k_composite<4, 2, 128, ST, PF>), grid, dim3(512), 0, s, A);
    else
      hipLaunchKernelGGL((k_composite<4, 2, 256, ST, PF>), grid, dim3(512), 0, s, A);
  } else {
    if (batch == 64u)
      hipLaunchKernelGGL((k_composite_q<2, 2, ST>), grid, dim3(256), 0, s, A);
    else if (batch == 128u)
      hipLaunchKernelGGL((k_composite<2, 2, 128, ST, PF>), grid, dim3(256), 0, s, A);
    else if (batch == 512u)
      hipLaunchKernelGGL((k_composite<2, 2, 512, ST, PF>), grid, dim3(256), 0, s, A);
    else
      hipLaunchKernelGGL((k_composite<2, 2, 256, ST, PF>), grid, dim3(256), 0, s, A);
  }
}
