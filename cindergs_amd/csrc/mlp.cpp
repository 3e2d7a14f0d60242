// mlp.cpp — host side of the MVC surrogate network (kernels in k_mlp.hip):
// parameter layout and initialisation, the training loop (sampling, forward,
// loss, backward, Adam, all enqueued on the context's stream), inference with
// forward-mode input derivatives, and the surrogate binding with its residual
// test and exact-MVC fallback.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "../../include/cindergs.h"
#include "cgs_internal.h"

namespace {

constexpr uint32_t NF = 32;    // Fourier frequencies (64 features)
constexpr uint32_t NIN = 64;
constexpr uint32_t NH = 256;
constexpr uint32_t CHUNK = 65536;  // inference points per pass
// split-K partials of the weight gradients: their GEMMs have K = the batch
// (8192) but only rows/64 x (cols+32)/64 output tiles (20 for a hidden
// layer), too few blocks for the chip without the split
constexpr uint32_t GRAD_SPLITS = 16;

#define MCHK(expr)                                                                    \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return cgs_fail(CGS_ERR_EXECUTION,                                              \
                      (std::string(#expr " failed: ") + hipGetErrorString(e_)).c_str()); \
  } while (0)

template <class T> struct Dev {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct cgs_mlp {
  cgs_context* ctx = nullptr;
  hipStream_t s = nullptr;
  uint32_t n_out = 0, n_out_p = 0;
  uint32_t rows[CGS_MLP_LAYERS], cols[CGS_MLP_LAYERS];
  uint64_t off[CGS_MLP_LAYERS + 1];
  uint64_t seed = 0;
  uint32_t step = 0;
  Dev<float> w, m, v, grad, fourier;
  Dev<uint16_t> wb[CGS_MLP_LAYERS], wt[CGS_MLP_LAYERS];
  // training buffers (batch P)
  uint32_t P = 0;
  Dev<uint32_t> idx;
  Dev<uint16_t> a0, a0t, hb[4], ht[4], dz, dzt, gb[2], gt[2];
  Dev<float> z, row_loss, losses;
  // inference buffers (CHUNK points): outputs and their tangents
  bool inf_ready = false;
  Dev<float> iz, itz;
  Dev<uint16_t> izh, itzh;  // bf16 copies for the near/far binding
  // near/far split radii (0, 0: the network stands for the whole weights)
  // and what the network learns with it (CGS_FAR_SHARE or CGS_FAR_SUMS)
  float near_r0 = 0.0f, near_r1 = 0.0f;
  uint32_t far_kind = CGS_FAR_SHARE;
  std::vector<void*> owned;
};

namespace {

template <class T> int alloc(cgs_mlp* M, Dev<T>& d, size_t n, bool zero = true) {
  if (d.p && d.n >= n) return CGS_OK;
  if (d.p) {
    (void)hipFree(d.p);
    for (auto& q : M->owned)
      if (q == d.p) q = nullptr;
  }
  d.p = nullptr;
  MCHK(hipMalloc(&d.p, (n ? n : 1) * sizeof(T)));
  d.n = n;
  M->owned.push_back(d.p);
  if (zero) MCHK(hipMemsetAsync(d.p, 0, (n ? n : 1) * sizeof(T), M->s));
  return CGS_OK;
}

#define TRYM(expr)             \
  do {                         \
    int r_ = (expr);           \
    if (r_ != CGS_OK) return r_; \
  } while (0)

// a bf16 ones row at `row` of a [rows x ld] bf16 matrix
int ones_row(cgs_mlp* M, uint16_t* base, uint32_t row, uint32_t ld) {
  std::vector<uint16_t> ones(ld, (uint16_t)0x3F80u);
  MCHK(hipMemcpyAsync(base + (size_t)row * ld, ones.data(), (size_t)ld * 2,
                      hipMemcpyHostToDevice, M->s));
  MCHK(hipStreamSynchronize(M->s));
  return CGS_OK;
}

CgsAdamArgs adam_args(cgs_mlp* M) {
  CgsAdamArgs A;
  memset(&A, 0, sizeof(A));
  A.w = M->w.p;
  A.m = M->m.p;
  A.v = M->v.p;
  A.grad = M->grad.p;
  A.total = M->off[CGS_MLP_LAYERS];
  A.splits = GRAD_SPLITS;
  A.split_stride = M->off[CGS_MLP_LAYERS];
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    A.off[l] = M->off[l];
    A.rows[l] = M->rows[l];
    A.cols[l] = M->cols[l];
    A.wb[l] = M->wb[l].p;
    A.wt[l] = M->wt[l].p;
  }
  return A;
}

CgsGemmArgs gemm(const uint16_t* A, uint32_t lda, const uint16_t* B, uint32_t ldb, uint32_t Mr,
                 uint32_t N, uint32_t K) {
  CgsGemmArgs g;
  memset(&g, 0, sizeof(g));
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.M = Mr;
  g.N = N;
  g.K = K;
  g.mask_div = 1;
  g.bias_stride = 1;
  return g;
}

int ensure_train(cgs_mlp* M, uint32_t P) {
  if (M->P == P) return CGS_OK;
  M->P = P;
  TRYM(alloc(M, M->idx, P));
  TRYM(alloc(M, M->a0, (size_t)P * NIN));
  TRYM(alloc(M, M->a0t, (size_t)(NIN + 32) * P));
  TRYM(ones_row(M, M->a0t.p, NIN, P));
  for (int l = 0; l < 4; ++l) {
    TRYM(alloc(M, M->hb[l], (size_t)P * NH));
    TRYM(alloc(M, M->ht[l], (size_t)(NH + 32) * P));
    TRYM(ones_row(M, M->ht[l].p, NH, P));
  }
  TRYM(alloc(M, M->z, (size_t)P * M->n_out_p));
  TRYM(alloc(M, M->dz, (size_t)P * M->n_out_p));
  TRYM(alloc(M, M->dzt, (size_t)M->n_out_p * P));
  for (int k = 0; k < 2; ++k) {
    TRYM(alloc(M, M->gb[k], (size_t)P * NH));
    TRYM(alloc(M, M->gt[k], (size_t)NH * P));
  }
  TRYM(alloc(M, M->row_loss, P));
  return CGS_OK;
}

// forward + loss + backward of the batch whose inputs are X[idx] (idx null:
// the first P points); gradients into M->grad, the mean loss into *d_loss
int forward_backward(cgs_mlp* M, const float* X, const float* T, const uint32_t* idx, uint32_t P,
                     float* d_loss) {
  hipStream_t s = M->s;
  cgs_launch_fourier(X, idx, P, M->fourier.p, M->a0.p, M->a0t.p, P, nullptr, s);
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    const uint16_t* in = l == 0 ? M->a0.p : M->hb[l - 1].p;
    CgsGemmArgs g = gemm(in, M->cols[l], M->wb[l].p, M->cols[l], P, M->rows[l], M->cols[l]);
    g.bias = M->w.p + M->off[l] + M->cols[l];
    g.bias_stride = M->cols[l] + 32;
    if (l < 4) {
      g.relu = 1;
      g.Cb = M->hb[l].p;
      g.ldcb = NH;
      g.Ct = M->ht[l].p;
      g.ldct = P;
    } else {
      g.Cf = M->z.p;
      g.ldcf = M->n_out_p;
    }
    cgs_launch_gemm_nt(g, s);
  }
  cgs_launch_mlp_loss(M->z.p, M->n_out_p, T, idx, P, M->n_out, M->dz.p, M->dzt.p, P,
                      d_loss ? M->row_loss.p : nullptr, !(M->near_r1 > 0.0f), s);
  if (d_loss) cgs_launch_sum(M->row_loss.p, P, (float)P * (float)M->n_out, d_loss, s);
  // backward: G5 = dZ; dW_l = G_{l+1}^T [H_l | 1]; G_l = (G_{l+1} W_l) * (H_l > 0)
  const uint16_t* gB = M->dz.p;   // G_{l+1} [P x rows_l]
  const uint16_t* gT = M->dzt.p;  // its transpose [rows_l x P]
  for (int l = CGS_MLP_LAYERS - 1; l >= 0; --l) {
    const uint16_t* hT = l == 0 ? M->a0t.p : M->ht[l - 1].p;  // [H_l | 1]^T [(cols+32) x P]
    CgsGemmArgs gw = gemm(gT, P, hT, P, M->rows[l], M->cols[l] + 32, P);
    gw.Cf = M->grad.p + M->off[l];
    gw.ldcf = M->cols[l] + 32;
    gw.ksplit = GRAD_SPLITS;
    gw.split_stride = M->off[CGS_MLP_LAYERS];
    cgs_launch_gemm_nt(gw, s);
    if (l == 0) break;
    const int k = l & 1;
    CgsGemmArgs gd = gemm(gB, M->rows[l], M->wt[l].p, M->rows[l], P, M->cols[l], M->rows[l]);
    gd.mask = M->hb[l - 1].p;
    gd.ldm = NH;
    gd.Cb = M->gb[k].p;
    gd.ldcb = NH;
    gd.Ct = M->gt[k].p;
    gd.ldct = P;
    cgs_launch_gemm_nt(gd, s);
    gB = M->gb[k].p;
    gT = M->gt[k].p;
  }
  return CGS_OK;
}

int ensure_infer(cgs_mlp* M) {
  if (M->inf_ready) return CGS_OK;
  TRYM(alloc(M, M->iz, (size_t)CHUNK * M->n_out_p, false));
  TRYM(alloc(M, M->itz, (size_t)3 * CHUNK * M->n_out_p, false));
  TRYM(alloc(M, M->izh, (size_t)CHUNK * M->n_out_p, false));
  TRYM(alloc(M, M->itzh, (size_t)3 * CHUNK * M->n_out_p, false));
  M->inf_ready = true;
  return CGS_OK;
}

CgsMlpInferArgs infer_args(cgs_mlp* M) {
  CgsMlpInferArgs a;
  memset(&a, 0, sizeof(a));
  a.fourier = M->fourier.p;
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    a.wb[l] = M->wb[l].p;
    a.off[l] = M->off[l];
  }
  a.w = M->w.p;
  a.n_out_p = M->n_out_p;
  a.ldz = M->n_out_p;
  return a;
}

// Network outputs of points X[0..n) in chunks of CHUNK (k_mlp_infer), then
// per chunk either the affine normalisation (+ residual and projection with a
// cage, k_mlp_post) into out_w / out_g, or, with `near`, the near/far
// combination (k_mvc_nearfar) whose per-point pointers are offset here.
int infer(cgs_mlp* M, const float* X, uint64_t n, float* out_w, float* out_g, const float* cage,
          const float* minv16, float* residual, const CgsNearArgs* near) {
  TRYM(ensure_infer(M));
  hipStream_t s = M->s;
  CgsMlpInferArgs ia = infer_args(M);
  if (near) {
    // one stream: running the network on a second stream alongside the
    // near/far kernel was measured slower on MI355X, the two contending for
    // the same CUs.  The far shares pass between them in bf16 (half the
    // traffic of f32; the network computes in bf16 anyway).
    for (uint64_t b = 0; b < n; b += CHUNK) {
      const uint32_t Q = (uint32_t)((n - b) < CHUNK ? (n - b) : CHUNK);
      ia.X = X + b * 3;
      ia.P = Q;
      ia.Zh = M->izh.p;
      ia.TZh = out_g ? M->itzh.p : nullptr;
      cgs_launch_mlp_infer(ia, s);
      CgsNearArgs na = *near;
      na.pts = X + b * 3;
      na.n = Q;
      na.winding = near->winding ? near->winding + b : nullptr;
      na.Q = nullptr;
      na.DQ = nullptr;
      na.Qh = M->izh.p;
      na.DQh = out_g ? M->itzh.p : nullptr;
      na.ldq = M->n_out_p;
      na.far_sums = M->far_kind == CGS_FAR_SUMS ? 1u : 0u;
      na.q_scale = na.far_sums ? near->sums_scale : (float)M->n_out;
      na.out_w = out_w + b * M->n_out;
      na.out_g = out_g ? out_g + b * M->n_out * 3 : nullptr;
      na.residual = near->residual + b;
      na.inside = near->inside ? near->inside + b : nullptr;
      cgs_launch_mvc_nearfar(na, s);
    }
    MCHK(hipGetLastError());
    return CGS_OK;
  }
  for (uint64_t b = 0; b < n; b += CHUNK) {
    const uint32_t Q = (uint32_t)((n - b) < CHUNK ? (n - b) : CHUNK);
    const float* x = X + b * 3;
    ia.X = x;
    ia.P = Q;
    ia.Z = M->iz.p;
    ia.TZ = out_g ? M->itz.p : nullptr;
    cgs_launch_mlp_infer(ia, s);
    CgsMlpPostArgs pa;
    memset(&pa, 0, sizeof(pa));
    pa.Z = M->iz.p;
    pa.TZ = out_g ? M->itz.p : nullptr;
    pa.ldz = M->n_out_p;
    pa.P = Q;
    pa.V = M->n_out;
    pa.X = x;
    pa.cage = cage;
    if (minv16) memcpy(pa.minv, minv16, sizeof(pa.minv));
    pa.out_w = out_w + b * M->n_out;
    pa.out_g = out_g ? out_g + b * M->n_out * 3 : nullptr;
    pa.residual = residual ? residual + b : nullptr;
    pa.raw = (M->near_r1 > 0.0f && !cage) ? 1 : 0;
    cgs_launch_mlp_post(pa, s);
  }
  MCHK(hipGetLastError());
  return CGS_OK;
}

}  // namespace

extern "C" {

int cgs_mlp_create(cgs_context* ctx, uint32_t n_out, float fourier_sigma, uint64_t seed,
                   cgs_mlp** out) {
  if (!ctx || !out) return cgs_fail(CGS_ERR_INVALID_ARG, "null argument");
  if (n_out < 4 || n_out > 4096) return cgs_fail(CGS_ERR_INVALID_ARG, "n_out must be in [4, 4096]");
  cgs_mlp* M = new cgs_mlp();
  M->ctx = ctx;
  M->s = cgs_context_stream(ctx);
  M->n_out = n_out;
  M->n_out_p = (n_out + 31u) / 32u * 32u;
  M->seed = seed;
  const uint32_t rows[CGS_MLP_LAYERS] = {NH, NH, NH, NH, M->n_out_p};
  const uint32_t cols[CGS_MLP_LAYERS] = {NIN, NH, NH, NH, NH};
  M->off[0] = 0;
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    M->rows[l] = rows[l];
    M->cols[l] = cols[l];
    M->off[l + 1] = M->off[l] + (uint64_t)rows[l] * (cols[l] + 32);
  }
  const uint64_t total = M->off[CGS_MLP_LAYERS];
  // He-uniform weights, zero biases and pads; Fourier matrix N(0, sigma^2)
  std::mt19937_64 rng(seed);
  std::vector<float> h(total, 0.0f);
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    const float bound = sqrtf(6.0f / (float)cols[l]) * (l == CGS_MLP_LAYERS - 1 ? 0.1f : 1.0f);
    std::uniform_real_distribution<float> U(-bound, bound);
    const uint32_t valid_rows = l == CGS_MLP_LAYERS - 1 ? n_out : rows[l];
    for (uint32_t r = 0; r < valid_rows; ++r)
      for (uint32_t c = 0; c < cols[l]; ++c) h[M->off[l] + (uint64_t)r * (cols[l] + 32) + c] = U(rng);
  }
  std::normal_distribution<float> G(0.0f, fourier_sigma);
  std::vector<float> fb(NF * 3);
  for (auto& x : fb) x = G(rng);
  int rc = CGS_OK;
  do {
    if ((rc = alloc(M, M->w, total)) != CGS_OK) break;
    if ((rc = alloc(M, M->m, total)) != CGS_OK) break;
    if ((rc = alloc(M, M->v, total)) != CGS_OK) break;
    if ((rc = alloc(M, M->grad, total * GRAD_SPLITS)) != CGS_OK) break;
    if ((rc = alloc(M, M->fourier, NF * 3)) != CGS_OK) break;
    for (int l = 0; l < CGS_MLP_LAYERS && rc == CGS_OK; ++l) {
      rc = alloc(M, M->wb[l], (size_t)rows[l] * cols[l]);
      if (rc == CGS_OK) rc = alloc(M, M->wt[l], (size_t)rows[l] * cols[l]);
    }
    if (rc != CGS_OK) break;
    if (hipMemcpyAsync(M->w.p, h.data(), total * 4, hipMemcpyHostToDevice, M->s) != hipSuccess ||
        hipMemcpyAsync(M->fourier.p, fb.data(), NF * 12, hipMemcpyHostToDevice, M->s) !=
            hipSuccess) {
      rc = cgs_fail(CGS_ERR_RUNTIME, "upload of the initial parameters failed");
      break;
    }
    cgs_launch_adam(adam_args(M), false, M->s);
    if (hipStreamSynchronize(M->s) != hipSuccess)
      rc = cgs_fail(CGS_ERR_EXECUTION, "parameter initialisation failed");
  } while (false);
  if (rc != CGS_OK) {
    cgs_mlp_destroy(M);
    return rc;
  }
  *out = M;
  return CGS_OK;
}

void cgs_mlp_destroy(cgs_mlp* M) {
  if (!M) return;
  (void)hipStreamSynchronize(M->s);
  for (void* p : M->owned)
    if (p) (void)hipFree(p);
  delete M;
}

uint64_t cgs_mlp_param_count(const cgs_mlp* M) { return M ? M->off[CGS_MLP_LAYERS] : 0; }

int cgs_mlp_layout(const cgs_mlp* M, uint32_t* rows, uint32_t* cols, uint64_t* offsets) {
  if (!M) return cgs_fail(CGS_ERR_INVALID_ARG, "null network");
  for (int l = 0; l < CGS_MLP_LAYERS; ++l) {
    if (rows) rows[l] = M->rows[l];
    if (cols) cols[l] = M->cols[l];
    if (offsets) offsets[l] = M->off[l];
  }
  return CGS_OK;
}

int cgs_mlp_get_params(cgs_mlp* M, float* params, float* fourier) {
  if (!M) return cgs_fail(CGS_ERR_INVALID_ARG, "null network");
  if (params)
    MCHK(hipMemcpyAsync(params, M->w.p, M->off[CGS_MLP_LAYERS] * 4, hipMemcpyDeviceToHost, M->s));
  if (fourier) MCHK(hipMemcpyAsync(fourier, M->fourier.p, NF * 12, hipMemcpyDeviceToHost, M->s));
  MCHK(hipStreamSynchronize(M->s));
  return CGS_OK;
}

int cgs_mlp_set_params(cgs_mlp* M, const float* params, const float* fourier) {
  if (!M) return cgs_fail(CGS_ERR_INVALID_ARG, "null network");
  if (params)
    MCHK(hipMemcpyAsync(M->w.p, params, M->off[CGS_MLP_LAYERS] * 4, hipMemcpyHostToDevice, M->s));
  if (fourier) MCHK(hipMemcpyAsync(M->fourier.p, fourier, NF * 12, hipMemcpyHostToDevice, M->s));
  cgs_launch_adam(adam_args(M), false, M->s);
  MCHK(hipStreamSynchronize(M->s));
  return CGS_OK;
}

int cgs_mlp_train(cgs_mlp* M, const float* d_x, const float* d_targets, uint64_t n, uint32_t steps,
                  uint32_t batch, float lr, float* losses) {
  if (!M || !d_x || !d_targets) return cgs_fail(CGS_ERR_INVALID_ARG, "null argument");
  if (n == 0 || batch == 0 || batch % 64 != 0)
    return cgs_fail(CGS_ERR_INVALID_ARG, "batch must be a positive multiple of 64");
  TRYM(ensure_train(M, batch));
  if (losses) TRYM(alloc(M, M->losses, steps ? steps : 1));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t i = 0; i < steps; ++i) {
    M->step++;
    cgs_launch_sample(M->idx.p, batch, n, M->seed, M->step, M->s);
    TRYM(forward_backward(M, d_x, d_targets, M->idx.p, batch, losses ? M->losses.p + i : nullptr));
    CgsAdamArgs A = adam_args(M);
    const double t = (double)M->step;
    A.b1 = b1;
    A.b2 = b2;
    A.lr_t = (float)(lr * sqrt(1.0 - pow(b2, t)) / (1.0 - pow(b1, t)));
    A.eps_t = (float)(eps * sqrt(1.0 - pow(b2, t)));
    cgs_launch_adam(A, true, M->s);
  }
  MCHK(hipGetLastError());
  if (losses && steps)
    MCHK(hipMemcpyAsync(losses, M->losses.p, (size_t)steps * 4, hipMemcpyDeviceToHost, M->s));
  MCHK(hipStreamSynchronize(M->s));
  return CGS_OK;
}

int cgs_mlp_gradient(cgs_mlp* M, const float* d_x, const float* d_targets, uint32_t n,
                     float* grad, float* loss) {
  if (!M || !d_x || !d_targets || !grad) return cgs_fail(CGS_ERR_INVALID_ARG, "null argument");
  if (n == 0 || n % 64 != 0) return cgs_fail(CGS_ERR_INVALID_ARG, "n must be a multiple of 64");
  TRYM(ensure_train(M, n));
  TRYM(alloc(M, M->losses, 1));
  TRYM(forward_backward(M, d_x, d_targets, nullptr, n, M->losses.p));
  cgs_launch_sum_splits(M->grad.p, M->off[CGS_MLP_LAYERS], GRAD_SPLITS, M->off[CGS_MLP_LAYERS],
                        M->s);
  MCHK(hipMemcpyAsync(grad, M->grad.p, M->off[CGS_MLP_LAYERS] * 4, hipMemcpyDeviceToHost, M->s));
  if (loss) MCHK(hipMemcpyAsync(loss, M->losses.p, 4, hipMemcpyDeviceToHost, M->s));
  MCHK(hipStreamSynchronize(M->s));
  return CGS_OK;
}

int cgs_mlp_set_near_field(cgs_mlp* M, float r0, float r1) {
  if (!M) return cgs_fail(CGS_ERR_INVALID_ARG, "null network");
  const bool off = r0 == 0.0f && r1 == 0.0f;
  if (!off && !(r0 >= 0.0f && r1 > r0 && r1 < 1e6f))
    return cgs_fail(CGS_ERR_INVALID_ARG, "near field radii need 0 <= r0 < r1 (or both 0)");
  M->near_r0 = r0;
  M->near_r1 = r1;
  return CGS_OK;
}

int cgs_mlp_set_far_kind(cgs_mlp* M, uint32_t kind) {
  if (!M) return cgs_fail(CGS_ERR_INVALID_ARG, "null network");
  if (kind != CGS_FAR_SHARE && kind != CGS_FAR_SUMS)
    return cgs_fail(CGS_ERR_INVALID_ARG, "far target kind must be CGS_FAR_SHARE or CGS_FAR_SUMS");
  M->far_kind = kind;
  return CGS_OK;
}

int cgs_mlp_infer(cgs_mlp* M, const float* d_x, uint64_t n, float* d_w, float* d_grad) {
  if (!M || !d_x || !d_w) return cgs_fail(CGS_ERR_INVALID_ARG, "null argument");
  if (n == 0) return CGS_OK;
  TRYM(infer(M, d_x, n, d_w, d_grad, nullptr, nullptr, nullptr, nullptr));
  MCHK(hipStreamSynchronize(M->s));
  return CGS_OK;
}

}  // extern "C"

// used by the surrogate binding in runtime.cpp
int cgs_mlp_infer_residual(cgs_mlp* M, const float* d_x, uint64_t n, float* d_w, float* d_grad,
                           const float* d_cage, const float* minv16, float* d_residual) {
  return infer(M, d_x, n, d_w, d_grad, d_cage, minv16, d_residual, nullptr);
}
int cgs_mlp_infer_nearfar(cgs_mlp* M, const float* d_x, uint64_t n, float* d_w, float* d_grad,
                          const CgsNearArgs& near) {
  return infer(M, d_x, n, d_w, d_grad, nullptr, nullptr, nullptr, &near);
}
bool cgs_mlp_near(cgs_mlp* M, float* r0, float* r1) {
  *r0 = M->near_r0;
  *r1 = M->near_r1;
  return M->near_r1 > 0.0f;
}
hipStream_t cgs_mlp_stream(cgs_mlp* M) { return M->s; }
uint32_t cgs_mlp_outputs(cgs_mlp* M) { return M->n_out; }
