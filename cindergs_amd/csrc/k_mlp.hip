// k_mlp.hip — the MVC surrogate network: a coordinate MLP, trained and run
// with hand-written BF16 MFMA kernels (FP32 accumulation, FP32 master weights).
//
// Network (no counterpart in the reference, which always evaluates exact MVC,
// src/deformation/mvc.rs:366-541): x (3) -> Fourier features
// [sin 2pi Bx, cos 2pi Bx] (B: 32 x 3, seeded N(0, sigma^2)) -> 4 x (Linear 256
// + ReLU) -> Linear n_out -> affine normalisation w = z + (1 - sum z) / n_out, so
// the weights sum to 1 as mean value coordinates do.  Loss: mean squared error
// against exact MVC weights.
//
// gfx950 design:
//  * one GEMM kernel, C = A B^T with A [M x K] and B [N x K] both K-contiguous
//    bf16, v_mfma_f32_16x16x32_bf16 tiles (a 64 x 64 block tile, four waves of
//    32 x 32), fragments loaded straight from global memory as 16-byte rows;
//    epilogues fuse bias, ReLU, the ReLU mask of the backward pass, f32 or
//    bf16 stores and a transposed bf16 copy (the layout the next GEMM needs);
//  * every layer's weights are stored with 32 extra columns, column `cols`
//    holding the bias: the activations' transposed copies carry a row of ones
//    there, so the weight-gradient GEMM produces the bias gradient too;
//  * all parameters live in one flat f32 buffer, updated by one Adam kernel
//    that also refreshes the bf16 weights and their transposes;
//  * inference propagates the three input tangents through the same GEMMs
//    (forward-mode derivative, masked by the primal ReLU pattern), which gives
//    the weight gradients dw/dx that a binding needs.
#include "cgs_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// C = A B^T: a block computes a 64 x BN tile (BN = 64 or 128) with four
// waves of 32 x BN/2, K staged through LDS in chunks of 64, two buffers:
// while the MFMAs consume one chunk the block's global loads of the next are
// in flight.  Rows of a staged tile are 128 bytes (eight 16-byte chunks);
// chunk c of row r sits at c ^ (r & 7), so the 16 rows a fragment read
// touches fall in different banks.
constexpr uint32_t GBK = 64;

__device__ __forceinline__ uint32_t swz(uint32_t row, uint32_t chunk) {
  return row * GBK + ((chunk ^ (row & 7u)) << 3);
}

template <int BN>
__global__ __launch_bounds__(256) void k_gemm_nt(CgsGemmArgs G) {
  constexpr int NJ = BN / 32;                    // MFMA tiles per wave along N
  constexpr int AL = 64 * 8 / 256;               // A chunks per thread (2)
  constexpr int BL = BN * 8 / 256;               // B chunks per thread (2 or 4)
  __shared__ __align__(16) uint16_t sA[2][64 * GBK];
  __shared__ __align__(16) uint16_t sB[2][BN * GBK];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t bm = blockIdx.y * 64u, bn = blockIdx.x * (uint32_t)BN;
  const uint32_t wm = (wave >> 1) * 32u, wn = (wave & 1u) * (uint32_t)(BN / 2);
  uint32_t k_begin = 0u, k_end = G.K;
  float* Cf = G.Cf;
  if (G.ksplit > 1u) {
    const uint32_t kc = ((G.K / 32u + G.ksplit - 1u) / G.ksplit) * 32u;
    k_begin = min(G.K, blockIdx.z * kc);
    k_end = min(G.K, k_begin + kc);
    Cf += (size_t)blockIdx.z * G.split_stride;
  }
  // the epilogue's operands, loaded before the K loop (unconditionally, at
  // indexes clamped into the matrix; out-of-range elements are not stored)
  float bias_v[NJ];
  uint16_t mask_v[2][NJ][4];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t col = min(bn + wn + j * 16u + (lane & 15u), G.N - 1u);
    bias_v[j] = G.bias ? G.bias[(size_t)col * G.bias_stride] : 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t row = min(bm + wm + i * 16u + (lane >> 4) * 4u + (uint32_t)q, G.M - 1u);
        mask_v[i][j][q] = G.mask ? G.mask[(size_t)(row / G.mask_div) * G.ldm + col] : (uint16_t)0;
      }
  }
  // this thread's staging chunks: id = t + 256 u -> (row id >> 3, chunk id & 7)
  const bf16x8 zero = {};
  bf16x8 ra[AL], rb[BL];
  auto load = [&](uint32_t k0) {
#pragma unroll
    for (int u = 0; u < AL; ++u) {
      const uint32_t id = t + 256u * (uint32_t)u, row = id >> 3, k = k0 + (id & 7u) * 8u;
      const uint32_t gr = bm + row;
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(G.A + (size_t)min(gr, G.M - 1u) * G.lda +
                                                        min(k, k_end - 8u));
      ra[u] = gr < G.M && k < k_end ? v : zero;
    }
#pragma unroll
    for (int u = 0; u < BL; ++u) {
      const uint32_t id = t + 256u * (uint32_t)u, row = id >> 3, k = k0 + (id & 7u) * 8u;
      const uint32_t gc = bn + row;
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(G.B + (size_t)min(gc, G.N - 1u) * G.ldb +
                                                        min(k, k_end - 8u));
      rb[u] = gc < G.N && k < k_end ? v : zero;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int u = 0; u < AL; ++u) {
      const uint32_t id = t + 256u * (uint32_t)u;
      *reinterpret_cast<bf16x8*>(&sA[buf][swz(id >> 3, id & 7u)]) = ra[u];
    }
#pragma unroll
    for (int u = 0; u < BL; ++u) {
      const uint32_t id = t + 256u * (uint32_t)u;
      *reinterpret_cast<bf16x8*>(&sB[buf][swz(id >> 3, id & 7u)]) = rb[u];
    }
  };
  f32x4 acc[2][NJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint32_t fr = lane & 15u, fc = lane >> 4;  // fragment row, 16-byte chunk in a k step
  if (k_begin < k_end) {
    load(k_begin);
    stage(0);
    __syncthreads();
    int buf = 0;
    for (uint32_t k0 = k_begin; k0 < k_end; k0 += GBK) {
      const bool more = k0 + GBK < k_end;  // block-uniform
      if (more) load(k0 + GBK);
#pragma unroll
      for (uint32_t ks = 0; ks < GBK / 32u; ++ks) {
        bf16x8 a[2], b[NJ];
#pragma unroll
        for (int i = 0; i < 2; ++i)
          a[i] = *reinterpret_cast<const bf16x8*>(&sA[buf][swz(wm + i * 16u + fr, ks * 4u + fc)]);
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          b[j] = *reinterpret_cast<const bf16x8*>(&sB[buf][swz(wn + j * 16u + fr, ks * 4u + fc)]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      if (more) stage(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
  // lane holds C[row0 + q][col] for q = 0..3
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const uint32_t col = bn + wn + j * 16u + (lane & 15u);
      const uint32_t row0 = bm + wm + i * 16u + (lane >> 4) * 4u;
      if (col >= G.N) continue;
      float v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float x = acc[i][j][q] + bias_v[j];
        if (G.relu) x = fmaxf(x, 0.0f);
        if (G.mask) {
          const uint16_t mv = mask_v[i][j][q];
          x = ((mv & 0x8000u) == 0u && mv != 0u) ? x : 0.0f;
        }
        v[q] = x;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t row = row0 + q;
        if (row >= G.M) continue;
        if (Cf) Cf[(size_t)row * G.ldcf + col] = v[q];
        if (G.Cb) G.Cb[(size_t)row * G.ldcb + col] = f2bf(v[q]);
      }
      if (G.Ct && row0 + 3u < G.M) {
        const uint64_t packed = (uint64_t)f2bf(v[0]) | ((uint64_t)f2bf(v[1]) << 16) |
                                ((uint64_t)f2bf(v[2]) << 32) | ((uint64_t)f2bf(v[3]) << 48);
        *reinterpret_cast<uint64_t*>(G.Ct + (size_t)col * G.ldct + row0) = packed;
      }
    }
}

// Fourier features of P points (gathered through idx when given): A0 [P x 64]
// and its transpose [64(+32) x P]; tangents T0 [3P x 64] (row 3p + axis).
__global__ __launch_bounds__(256) void k_fourier(const float* X, const uint32_t* idx, uint32_t P,
                                                 const float* Bm, uint16_t* A0, uint16_t* A0t,
                                                 uint32_t ldt, uint16_t* T0) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t p = t >> 5, k = t & 31u;
  if (p >= P) return;
  const uint32_t src = idx ? idx[p] : p;
  const float x = X[(size_t)src * 3], y = X[(size_t)src * 3 + 1], z = X[(size_t)src * 3 + 2];
  const float b0 = Bm[k * 3], b1 = Bm[k * 3 + 1], b2 = Bm[k * 3 + 2];
  const float TWO_PI = 6.283185307179586f;
  const float s = TWO_PI * ((b0 * x + b1 * y) + b2 * z);
  float sn, cs;
  sincosf(s, &sn, &cs);
  A0[(size_t)p * 64 + k] = f2bf(sn);
  A0[(size_t)p * 64 + 32 + k] = f2bf(cs);
  if (A0t) {
    A0t[(size_t)k * ldt + p] = f2bf(sn);
    A0t[(size_t)(32 + k) * ldt + p] = f2bf(cs);
  }
  if (T0) {
    const float bb[3] = {b0, b1, b2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      T0[((size_t)p * 3 + a) * 64 + k] = f2bf(TWO_PI * bb[a] * cs);
      T0[((size_t)p * 3 + a) * 64 + 32 + k] = f2bf(-TWO_PI * bb[a] * sn);
    }
  }
}

__device__ __forceinline__ float wave_sum64(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Affine normalisation, squared error and its gradient:
// w = z + (1 - sum z)/V, e = w - target, loss_row = sum e^2,
// dL/dz_j = g_j - mean(g) with g = 2 e / (P V).  With affine = 0 the outputs
// are compared as they are (w = z, dL/dz = g): the far-field share of a
// near/far network has no sum constraint.  Writes dZ [P x ldz] and its
// transpose [ldz(+32) x P] in bf16 (columns >= V stay zero).
//
// One wave per point (lanes over the outputs): the row sums, then dZ.  The
// transpose dZ^T is written by k_transpose_bf16 (a column of dZ is one row of
// dZ^T; written here directly each lane would store 2 bytes to its own row).
__global__ __launch_bounds__(256) void k_mlp_loss(const float* Z, uint32_t ldz,
                                                  const float* targets, const uint32_t* idx,
                                                  uint32_t P, uint32_t V, uint16_t* dZ,
                                                  float* row_loss, int affine) {
  const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (p >= P) return;
  const float* z = Z + (size_t)p * ldz;
  const float* tg = targets + (size_t)(idx ? idx[p] : p) * V;
  float s = 0.0f;
  for (uint32_t j = lane; j < V; j += 64u) s += z[j];
  s = wave_sum64(s);
  const float shift = affine ? (1.0f - s) / (float)V : 0.0f;
  const float scale = 2.0f / ((float)P * (float)V);
  float gsum = 0.0f, l2 = 0.0f;
  for (uint32_t j = lane; j < V; j += 64u) {
    const float e = (z[j] + shift) - tg[j];
    l2 += e * e;
    gsum += scale * e;
  }
  gsum = wave_sum64(gsum);
  l2 = wave_sum64(l2);
  const float gmean = affine ? gsum / (float)V : 0.0f;
  for (uint32_t j = lane; j < V; j += 64u) {
    const float e = (z[j] + shift) - tg[j];
    dZ[(size_t)p * ldz + j] = f2bf(scale * e - gmean);
  }
  if (lane == 0 && row_loss) row_loss[p] = l2;
}

// out[c * ldo + r] = in[r * ldi + c], r < R, c < C (bf16), through a 64 x 64
// LDS tile: each wave reads and writes 128-byte row segments
__global__ __launch_bounds__(256) void k_transpose_bf16(const uint16_t* in, uint32_t ldi,
                                                        uint32_t R, uint32_t C, uint16_t* out,
                                                        uint32_t ldo) {
  __shared__ uint16_t t[64][66];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t r0 = blockIdx.y * 64u, c0 = blockIdx.x * 64u;
  for (uint32_t i = wave; i < 64u; i += 4u) {
    const uint32_t r = r0 + i, c = c0 + lane;
    t[i][lane] = r < R && c < C ? in[(size_t)r * ldi + c] : (uint16_t)0;
  }
  __syncthreads();
  for (uint32_t i = wave; i < 64u; i += 4u) {
    const uint32_t c = c0 + i, r = r0 + lane;
    if (c < C && r < R) out[(size_t)c * ldo + r] = t[lane][i];
  }
}

// one block: out = sum(v[0..n)) / div
__global__ __launch_bounds__(1024) void k_sum(const float* v, uint32_t n, float div, float* out) {
  __shared__ float s[16];
  float a = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 1024u) a += v[i];
  a = wave_sum64(a);
  if ((threadIdx.x & 63u) == 0) s[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < 16; ++w) t += s[w];
    *out = t / div;
  }
}

// Adam on the flat parameter buffer, then the bf16 weights [rows x cols] and
// transposed [cols x rows] of each layer (bias column excluded) refreshed.
__global__ __launch_bounds__(256) void k_adam(CgsAdamArgs A) {
  const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (e >= A.total) return;
  float g = A.grad[e];
  for (uint32_t sp = 1; sp < A.splits; ++sp) g += A.grad[e + sp * A.split_stride];
  const float m = A.b1 * A.m[e] + (1.0f - A.b1) * g;
  const float v = A.b2 * A.v[e] + (1.0f - A.b2) * g * g;
  A.m[e] = m;
  A.v[e] = v;
  const float w = A.w[e] - A.lr_t * m / (sqrtf(v) + A.eps_t);
  A.w[e] = w;
  int l = 0;
#pragma unroll
  for (int k = 1; k < CGS_MLP_LAYERS; ++k)
    if (e >= A.off[k]) l = k;
  const uint64_t r = e - A.off[l];
  const uint32_t ld = A.cols[l] + 32u;
  const uint32_t row = (uint32_t)(r / ld), col = (uint32_t)(r % ld);
  if (col < A.cols[l]) {
    const uint16_t b = f2bf(w);
    A.wb[l][(size_t)row * A.cols[l] + col] = b;
    A.wt[l][(size_t)col * A.rows[l] + row] = b;
  }
}

// the split-K partials of the gradient summed into the first (in place)
__global__ __launch_bounds__(256) void k_sum_splits(float* grad, uint64_t total, uint32_t splits,
                                                    uint64_t stride) {
  const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  float g = grad[e];
  for (uint32_t sp = 1; sp < splits; ++sp) g += grad[e + sp * stride];
  grad[e] = g;
}

// bf16 weights and transposes from the master weights (no update)
__global__ __launch_bounds__(256) void k_refresh(CgsAdamArgs A) {
  const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (e >= A.total) return;
  int l = 0;
#pragma unroll
  for (int k = 1; k < CGS_MLP_LAYERS; ++k)
    if (e >= A.off[k]) l = k;
  const uint64_t r = e - A.off[l];
  const uint32_t ld = A.cols[l] + 32u;
  const uint32_t row = (uint32_t)(r / ld), col = (uint32_t)(r % ld);
  if (col < A.cols[l]) {
    const uint16_t b = f2bf(A.w[e]);
    A.wb[l][(size_t)row * A.cols[l] + col] = b;
    A.wt[l][(size_t)col * A.rows[l] + row] = b;
  }
}

// batch indices: a 64-bit mix of (seed, step, i) modulo n
__global__ __launch_bounds__(256) void k_sample(uint32_t* idx, uint32_t P, uint64_t n,
                                                uint64_t seed, uint32_t step) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= P) return;
  uint64_t x = seed ^ ((uint64_t)step << 32) ^ (uint64_t)i * 0x9E3779B97F4A7C15ull;
  x ^= x >> 33;
  x *= 0xFF51AFD7ED558CCDull;
  x ^= x >> 33;
  x *= 0xC4CEB9FE1A85EC53ull;
  x ^= x >> 33;
  idx[i] = (uint32_t)(x % n);
}

// Whole-network inference, one block of eight waves per 128 activation rows.
// With tangents (TAN) a block takes 32 points and row 4p + a holds point p's
// activations (a = 0) or their derivative along axis a - 1 (a = 1..3);
// without, 128 points, one row each.  The rows live in LDS as bf16
// (128 x 256, 64 KiB, 16-byte chunks XOR-swizzled by row so that the sixteen
// rows of an MFMA fragment read fall on different banks); every layer is
// v_mfma_f32_16x16x32_bf16 over the resident rows, each wave producing 32
// output columns for all 128 rows (eight waves: the 256 hidden columns in
// one pass), its weight fragments read from global memory (the weights,
// 0.8 MiB in bf16, stay in L2) one K step ahead.  Measured on MI355X against
// four waves of 64 columns (0.73 ms per 65536 points with tangents) and a
// 2 x 2 split of rows and columns (0.88 ms): 0.47 ms, the weight-fragment
// loads per wave being what limits the wider tiles.  The MFMA output layout
// puts rows 4q..4q+3 of one column in one lane, i.e. a point's value and its
// three tangents, so the ReLU mask of the tangents is applied in registers.
// The activations never leave the chip; only the output layer is written, as
// f32.
__device__ __forceinline__ uint32_t act_idx(uint32_t row, uint32_t k) {
  return row * 256u + ((((k >> 3) ^ (row & 7u)) << 3) | (k & 7u));
}

template <bool TAN>
__global__ __launch_bounds__(512) void k_mlp_infer(CgsMlpInferArgs A) {
  __shared__ __attribute__((aligned(16))) uint16_t act[128 * 256];
  constexpr uint32_t PTS = TAN ? 32u : 128u;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t r = lane & 15u, kq = (lane >> 4) * 8u;
  const uint32_t p0 = blockIdx.x * PTS;
  const float TWO_PI = 6.283185307179586f;
  for (uint32_t t = threadIdx.x; t < PTS * 32u; t += 512u) {
    const uint32_t pl = t >> 5, k = t & 31u, p = p0 + pl;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (p < A.P) {
      x = A.X[(size_t)p * 3];
      y = A.X[(size_t)p * 3 + 1];
      z = A.X[(size_t)p * 3 + 2];
    }
    const float bb[3] = {A.fourier[k * 3], A.fourier[k * 3 + 1], A.fourier[k * 3 + 2]};
    const float sarg = TWO_PI * ((bb[0] * x + bb[1] * y) + bb[2] * z);
    float sn, cs;
    sincosf(sarg, &sn, &cs);
    const uint32_t row = TAN ? pl * 4u : pl;
    act[act_idx(row, k)] = f2bf(sn);
    act[act_idx(row, 32u + k)] = f2bf(cs);
    if (TAN) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        act[act_idx(row + 1u + a, k)] = f2bf(TWO_PI * bb[a] * cs);
        act[act_idx(row + 1u + a, 32u + k)] = f2bf(-TWO_PI * bb[a] * sn);
      }
    }
  }
  __syncthreads();
  const bf16x8 zero = {};
  f32x4 acc[8][2];
#pragma unroll 1
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    const uint32_t K = l == 0 ? 64u : 256u;
    const uint32_t N = l == CGS_MLP_LAYERS - 1 ? A.n_out_p : 256u;
    const uint16_t* W = A.wb[l];
    const float* bias = A.w + A.off[l] + K;
#pragma unroll 1
    for (uint32_t c0 = wave * 32u; c0 < N; c0 += 256u) {
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      // weight fragments one K step ahead
      bf16x8 b[2], bn[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t col = c0 + j * 16u + r;
        b[j] = col < N ? *reinterpret_cast<const bf16x8*>(W + (size_t)col * K + kq) : zero;
      }
      for (uint32_t k = 0; k < K; k += 32u) {
        const uint32_t kn = k + 32u < K ? k + 32u : k;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const uint32_t col = c0 + j * 16u + r;
          bn[j] = col < N ? *reinterpret_cast<const bf16x8*>(W + (size_t)col * K + kn + kq) : zero;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&act[act_idx(i * 16u + r, k + kq)]);
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = bn[j];
      }
      if (l == CGS_MLP_LAYERS - 1) {
        // output layer: f32 rows of Z (values) and TZ (tangents)
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const uint32_t col = c0 + j * 16u + (lane & 15u);
            if (col >= N) continue;
            const float bv = bias[(size_t)col * (K + 32u)];
            if (TAN) {
              const uint32_t p = p0 + i * 4u + (lane >> 4);
              if (p >= A.P) continue;
              if (A.Zh) {
                A.Zh[(size_t)p * A.ldz + col] = f2bf(acc[i][j][0] + bv);
                if (A.TZh) {
#pragma unroll
                  for (int a = 0; a < 3; ++a)
                    A.TZh[((size_t)p * 3 + a) * A.ldz + col] = f2bf(acc[i][j][1 + a]);
                }
                continue;
              }
              A.Z[(size_t)p * A.ldz + col] = acc[i][j][0] + bv;
              if (A.TZ) {
#pragma unroll
                for (int a = 0; a < 3; ++a) A.TZ[((size_t)p * 3 + a) * A.ldz + col] = acc[i][j][1 + a];
              }
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                const uint32_t p = p0 + i * 16u + (lane >> 4) * 4u + q;
                if (p >= A.P) continue;
                if (A.Zh)
                  A.Zh[(size_t)p * A.ldz + col] = f2bf(acc[i][j][q] + bv);
                else
                  A.Z[(size_t)p * A.ldz + col] = acc[i][j][q] + bv;
              }
            }
          }
      }
    }
    if (l == CGS_MLP_LAYERS - 1) break;
    __syncthreads();  // every wave has read this layer's input rows
    const uint32_t c0 = wave * 32u;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t col = c0 + j * 16u + (lane & 15u);
        const uint32_t row0 = i * 16u + (lane >> 4) * 4u;
        const float bv = bias[(size_t)col * (K + 32u)];
        if (TAN) {
          const float pre = acc[i][j][0] + bv;
          const bool on = pre > 0.0f;
          act[act_idx(row0, col)] = f2bf(on ? pre : 0.0f);
#pragma unroll
          for (int a = 1; a < 4; ++a) act[act_idx(row0 + a, col)] = f2bf(on ? acc[i][j][a] : 0.0f);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) act[act_idx(row0 + q, col)] = f2bf(fmaxf(acc[i][j][q] + bv, 0.0f));
        }
      }
    __syncthreads();
  }
}

// Surrogate outputs of P points (one wave per point): w = z + (1 - sum z)/V
// and, from the tangent rows, dw/dx_a = t_a - (sum t_a)/V.  With a cage: the
// reconstruction residual |sum_j w_j v_j - x| of these weights (the fallback
// test), then the weights projected onto linear precision,
// w' = w + A^T Minv ([1; x] - A w) with A = [1 ... 1; v_1 ... v_V] and
// Minv = (A A^T)^-1 (so sum w' = 1 and sum w' v = x), and their derivatives
// g'_a = g_a + A^T Minv (e_{a+1} - A g_a).
__global__ __launch_bounds__(256) void k_mlp_post(CgsMlpPostArgs A) {
  const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (p >= A.P) return;
  const uint32_t V = A.V;
  const float* z = A.Z + (size_t)p * A.ldz;
  float s = 0.0f;
  for (uint32_t j = lane; j < V; j += 64u) s += z[j];
  s = wave_sum64(s);
  const float shift = A.raw ? 0.0f : (1.0f - s) / (float)V;
  float* wrow = A.out_w + (size_t)p * V;
  float* grow = A.out_g ? A.out_g + (size_t)p * V * 3 : nullptr;
  if (A.TZ && grow) {
    for (int a = 0; a < 3; ++a) {
      const float* t = A.TZ + ((size_t)p * 3 + a) * A.ldz;
      float ts = 0.0f;
      for (uint32_t j = lane; j < V; j += 64u) ts += t[j];
      ts = A.raw ? 0.0f : wave_sum64(ts) / (float)V;
      for (uint32_t j = lane; j < V; j += 64u) grow[(size_t)j * 3 + a] = t[j] - ts;
    }
  }
  if (!A.cage) {
    for (uint32_t j = lane; j < V; j += 64u) wrow[j] = z[j] + shift;
    return;
  }
  const float* C = A.cage;
  float s0 = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (uint32_t j = lane; j < V; j += 64u) {
    const float w = z[j] + shift;
    s0 += w;
    sx += w * C[j * 3];
    sy += w * C[j * 3 + 1];
    sz += w * C[j * 3 + 2];
  }
  s0 = wave_sum64(s0);
  sx = wave_sum64(sx);
  sy = wave_sum64(sy);
  sz = wave_sum64(sz);
  const float x = A.X[(size_t)p * 3], y = A.X[(size_t)p * 3 + 1], zz = A.X[(size_t)p * 3 + 2];
  if (A.residual && lane == 0)
    A.residual[p] = sqrtf(((sx - x) * (sx - x) + (sy - y) * (sy - y)) + (sz - zz) * (sz - zz));
  const float r[4] = {1.0f - s0, x - sx, y - sy, zz - sz};
  float lam[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    lam[i] = ((A.minv[i * 4] * r[0] + A.minv[i * 4 + 1] * r[1]) + A.minv[i * 4 + 2] * r[2]) +
             A.minv[i * 4 + 3] * r[3];
  for (uint32_t j = lane; j < V; j += 64u)
    wrow[j] = (z[j] + shift) + ((lam[0] + lam[1] * C[j * 3]) + lam[2] * C[j * 3 + 1]) +
              lam[3] * C[j * 3 + 2];
  if (!(A.TZ && grow)) return;
  for (int a = 0; a < 3; ++a) {
    float c0 = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (uint32_t j = lane; j < V; j += 64u) {
      const float g = grow[(size_t)j * 3 + a];
      c0 += g;
      cx += g * C[j * 3];
      cy += g * C[j * 3 + 1];
      cz += g * C[j * 3 + 2];
    }
    c0 = wave_sum64(c0);
    cx = wave_sum64(cx);
    cy = wave_sum64(cy);
    cz = wave_sum64(cz);
    const float e[4] = {-c0, (a == 0 ? 1.0f : 0.0f) - cx, (a == 1 ? 1.0f : 0.0f) - cy,
                        (a == 2 ? 1.0f : 0.0f) - cz};
    float la[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      la[i] = ((A.minv[i * 4] * e[0] + A.minv[i * 4 + 1] * e[1]) + A.minv[i * 4 + 2] * e[2]) +
              A.minv[i * 4 + 3] * e[3];
    for (uint32_t j = lane; j < V; j += 64u)
      grow[(size_t)j * 3 + a] += ((la[0] + la[1] * C[j * 3]) + la[2] * C[j * 3 + 1]) +
                                 la[3] * C[j * 3 + 2];
  }
}

}  // namespace

void cgs_launch_gemm_nt(const CgsGemmArgs& g, hipStream_t s) {
  if (g.M == 0 || g.N == 0) return;
  const uint32_t z = g.ksplit > 1u ? g.ksplit : 1u;
  // 64 x 128 tiles when they still give at least 256 blocks, else 64 x 64
  const uint32_t m_blocks = (g.M + 63u) / 64u;
  if ((uint64_t)m_blocks * ((g.N + 127u) / 128u) * z >= 256u)
    hipLaunchKernelGGL(k_gemm_nt<128>, dim3((g.N + 127u) / 128u, m_blocks, z), dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL(k_gemm_nt<64>, dim3((g.N + 63u) / 64u, m_blocks, z), dim3(256), 0, s, g);
}

void cgs_launch_sum_splits(float* grad, uint64_t total, uint32_t splits, uint64_t stride,
                           hipStream_t s) {
  if (total == 0 || splits <= 1u) return;
  hipLaunchKernelGGL(k_sum_splits, dim3((uint32_t)((total + 255u) / 256u)), dim3(256), 0, s, grad,
                     total, splits, stride);
}

void cgs_launch_fourier(const float* X, const uint32_t* idx, uint32_t P, const float* Bm,
                        uint16_t* A0, uint16_t* A0t, uint32_t ldt, uint16_t* T0, hipStream_t s) {
  if (P == 0) return;
  hipLaunchKernelGGL(k_fourier, dim3((P * 32u + 255u) / 256u), dim3(256), 0, s, X, idx, P, Bm, A0,
                     A0t, ldt, T0);
}

void cgs_launch_mlp_infer(const CgsMlpInferArgs& a, hipStream_t s) {
  if (a.P == 0) return;
  if (a.TZ || a.TZh)
    hipLaunchKernelGGL(k_mlp_infer<true>, dim3((a.P + 31u) / 32u), dim3(512), 0, s, a);
  else
    hipLaunchKernelGGL(k_mlp_infer<false>, dim3((a.P + 127u) / 128u), dim3(512), 0, s, a);
}

void cgs_launch_mlp_loss(const float* Z, uint32_t ldz, const float* targets, const uint32_t* idx,
                         uint32_t P, uint32_t V, uint16_t* dZ, uint16_t* dZt, uint32_t ldt,
                         float* row_loss, bool affine, hipStream_t s) {
  if (P == 0) return;
  hipLaunchKernelGGL(k_mlp_loss, dim3((P + 3u) / 4u), dim3(256), 0, s, Z, ldz, targets, idx, P, V,
                     dZ, row_loss, affine ? 1 : 0);
  hipLaunchKernelGGL(k_transpose_bf16, dim3((V + 63u) / 64u, (P + 63u) / 64u), dim3(256), 0, s, dZ,
                     ldz, P, V, dZt, ldt);
}

void cgs_launch_sum(const float* v, uint32_t n, float div, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, v, n, div, out);
}

void cgs_launch_adam(const CgsAdamArgs& a, bool update, hipStream_t s) {
  const uint32_t blocks = (uint32_t)((a.total + 255u) / 256u);
  if (update)
    hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_refresh, dim3(blocks), dim3(256), 0, s, a);
}

void cgs_launch_sample(uint32_t* idx, uint32_t P, uint64_t n, uint64_t seed, uint32_t step,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_sample, dim3((P + 255u) / 256u), dim3(256), 0, s, idx, P, n, seed, step);
}

void cgs_launch_mlp_post(const CgsMlpPostArgs& a, hipStream_t s) {
  if (a.P == 0) return;
  hipLaunchKernelGGL(k_mlp_post, dim3((a.P + 3u) / 4u), dim3(256), 0, s, a);
}
