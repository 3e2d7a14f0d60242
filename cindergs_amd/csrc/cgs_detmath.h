// cgs_detmath.h — deterministic single-precision elementary functions.
//
// These functions define the arithmetic of the render path: the HIP kernels and
// the CPU oracle both call them, so a value computed on the GPU and on the host
// from the same inputs has the same bits.  Every function below uses only IEEE
// single-precision +, -, *, / and sqrt (each correctly rounded on gfx950 and on
// x86-64 SSE) in a fixed order, and must be compiled with -ffp-contract=off so
// that no multiply-add is fused differently on the two sides.
//
// Methods (textbook, coefficients derived from the Taylor series, not taken
// from any vendor library):
//   log : x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
//         s = (m-1)/(m+1), atanh series through s^11 (Cody & Waite style
//         argument reduction, ln 2 split into a 16-bit head and a tail).
//   exp : x = k ln2 + r, |r| <= ln2/2 (Cody & Waite two-constant reduction),
//         exp r by its Taylor polynomial through r^7, then scaled by 2^k.
//   pow : exp(y * log(x)) for x > 0.
//   acos: half-angle identities onto asin(z), |z| <= 1/2, asin Taylor series
//         through z^21.
//   cos : Taylor series through x^18, valid on [-pi, pi].
//   atan2: |y|/|x| or |x|/|y| in [0, 1]; above tan(pi/8), atan t =
//         pi/4 + atan((t-1)/(t+1)); atan u by its Taylor series through u^17
//         on |u| <= tan(pi/8); then the octant and quadrant identities.
//   f32 -> f16: IEEE round to nearest, ties to even, on the bit patterns
//         (integer operations only), subnormal halves included.
// Accuracy is pinned by tests against double-precision libm (a few ulp).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define CGS_HD __host__ __device__ __forceinline__
#else
#define CGS_HD static inline
#endif

CGS_HD uint32_t cgs_f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
CGS_HD float cgs_u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// Rust f32::max / f32::min semantics (IEEE maxNum/minNum: a NaN operand yields the other).
CGS_HD float cgs_fmax(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}
CGS_HD float cgs_fmin(float a, float b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}
// Rust f32::clamp(self, lo, hi): NaN stays NaN.
CGS_HD float cgs_fclamp(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// Rust `f as i32`: saturating, NaN -> 0.
CGS_HD int32_t cgs_f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483520.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}
CGS_HD int32_t cgs_iclamp(int32_t x, int32_t lo, int32_t hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

#define CGS_LN2_HI 0.693145751953125f        /* 45426 / 65536, exact in 16 bits */
#define CGS_LN2_LO 1.42860682030941723212e-6f /* ln 2 - CGS_LN2_HI */
#define CGS_INV_LN2 1.44269504088896340736f

CGS_HD float cgs_logf(float x) {
  if (x != x || x < 0.0f) return cgs_u2f(0x7fc00000u);
  if (x == 0.0f) return cgs_u2f(0xff800000u);
  uint32_t u = cgs_f2u(x);
  if (u >= 0x7f800000u) return x;  // +inf
  int32_t e_adj = 0;
  if (u < 0x00800000u) {  // subnormal: scale by 2^24
    x = x * 16777216.0f;
    u = cgs_f2u(x);
    e_adj = -24;
  }
  int32_t e = (int32_t)((u >> 23) & 0xffu) - 127 + e_adj;
  float m = cgs_u2f((u & 0x007fffffu) | 0x3f800000u);  // [1, 2)
  if (m > 1.41421356237309504880f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float f = m - 1.0f;  // exact (Sterbenz)
  float s = f / (2.0f + f);
  float z = s * s;
  float w = z * (0.666666666666666667f +
                 z * (0.4f + z * (0.285714285714285714f +
                                  z * (0.222222222222222222f + z * 0.181818181818181818f))));
  float lnm = (2.0f * s) + s * w;
  float ef = (float)e;
  return ef * CGS_LN2_HI + (ef * CGS_LN2_LO + lnm);
}

CGS_HD float cgs_expf(float x) {
  if (x != x) return x;
  if (x > 88.7228393554688f) return cgs_u2f(0x7f800000u);
  if (x < -103.972084045410f) return 0.0f;
  float kf = x * CGS_INV_LN2;
  // round half away from zero, deterministic on both sides
  kf = kf >= 0.0f ? (float)(int32_t)(kf + 0.5f) : -(float)(int32_t)(0.5f - kf);
  int32_t k = (int32_t)kf;
  float r = (x - kf * CGS_LN2_HI) - kf * CGS_LN2_LO;
  float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666666666666667f +
                                                r * (0.0416666666666666667f +
                                                     r * (0.00833333333333333333f +
                                                          r * (0.00138888888888888889f +
                                                               r * 0.000198412698412698413f))))));
  if (k > 127) {
    p = p * 2.0f;
    k = k - 1;
  }
  if (k >= -126) return p * cgs_u2f((uint32_t)(k + 127) << 23);
  // subnormal result: two exact power-of-two steps
  p = p * cgs_u2f((uint32_t)(k + 126 + 127) << 23);  // 2^(k+126), k+126 in [-24, -1]
  return p * cgs_u2f(1u << 23);                        // 2^-126
}

CGS_HD float cgs_powf_pos(float x, float y) {
  // x > 0 (callers clamp); y finite.
  return cgs_expf(y * cgs_logf(x));
}

CGS_HD float cgs_asinf_small(float x) {
  // |x| <= 0.5; Taylor series of asin through x^21.
  float z = x * x;
  float p = 0.166666666666666667f +
            z * (0.075f +
                 z * (0.0446428571428571429f +
                      z * (0.0303819444444444444f +
                           z * (0.0223721590909090909f +
                                z * (0.0173527644230769231f +
                                     z * (0.0139648437500000000f +
                                          z * (0.0115518008961397059f +
                                               z * (0.00976160952919407895f +
                                                    z * 0.00839033580961681548f))))))));
  return x + (x * z) * p;
}

#define CGS_PI_F 3.14159265358979323846f
#define CGS_HALF_PI_F 1.57079632679489661923f

CGS_HD float cgs_acosf(float x) {
  if (x != x) return x;
  if (x > 1.0f || x < -1.0f) return cgs_u2f(0x7fc00000u);
  if (x > 0.5f) {
    float t = __builtin_sqrtf((1.0f - x) * 0.5f);
    return 2.0f * cgs_asinf_small(t);
  }
  if (x < -0.5f) {
    float t = __builtin_sqrtf((1.0f + x) * 0.5f);
    return CGS_PI_F - 2.0f * cgs_asinf_small(t);
  }
  return CGS_HALF_PI_F - cgs_asinf_small(x);
}

CGS_HD float cgs_cosf(float x) {
  // valid on [-pi, pi]
  float z = x * x;
  float p = 1.0f + z * (-0.5f +
            z * (0.0416666666666666667f +
            z * (-0.00138888888888888889f +
            z * (2.48015873015873016e-5f +
            z * (-2.75573192239858907e-7f +
            z * (2.08767569878680990e-9f +
            z * (-1.14707455977297247e-11f +
            z * (4.77947733238738530e-14f +
            z * -1.56192069685862265e-16f))))))));
  return p;
}

CGS_HD float cgs_sqrtf(float x) { return __builtin_sqrtf(x); }

// atan on [0, 1] (see the method above)
CGS_HD float cgs_atan_unit(float t) {
  const bool red = t > 0.414213562373095049f;
  const float u = red ? (t - 1.0f) / (t + 1.0f) : t;
  const float z = u * u;
  const float p = 1.0f + z * (-0.333333333333333333f +
                  z * (0.2f +
                  z * (-0.142857142857142857f +
                  z * (0.111111111111111111f +
                  z * (-0.0909090909090909091f +
                  z * (0.0769230769230769231f +
                  z * (-0.0666666666666666667f +
                  z * 0.0588235294117647059f)))))));
  const float r = u * p;
  return red ? 0.785398163397448310f + r : r;
}

// atan2(y, x) with the quadrant conventions of C's atan2 for finite inputs
CGS_HD float cgs_atan2f(float y, float x) {
  if (x != x || y != y) return x + y;
  const float ax = x < 0.0f ? -x : x, ay = y < 0.0f ? -y : y;
  const bool yneg = (cgs_f2u(y) >> 31) != 0u, xneg = (cgs_f2u(x) >> 31) != 0u;
  float a;
  if (ax == 0.0f && ay == 0.0f) {
    a = xneg ? CGS_PI_F : 0.0f;
  } else if (ay > ax) {
    a = CGS_HALF_PI_F - cgs_atan_unit(ax / ay);
    if (xneg) a = CGS_PI_F - a;
  } else {
    a = cgs_atan_unit(ay / ax);
    if (xneg) a = CGS_PI_F - a;
  }
  return yneg ? -a : a;
}

// f32 -> IEEE binary16 bits, round to nearest even; NaN stays a quiet NaN
// with the top payload bits, overflow gives infinity
CGS_HD uint16_t cgs_f32_to_f16(float f) {
  const uint32_t x = cgs_f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) {
    if (ax == 0x7F800000u) return (uint16_t)(sign | 0x7C00u);
    return (uint16_t)(sign | 0x7E00u | ((ax >> 13) & 0x03FFu));
  }
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // >= 65520 rounds to inf
  if (ax >= 0x38800000u) {
    // normal half: rebias the exponent by 127 - 15 and round the 13 dropped bits
    const uint32_t r = ax - 0x38000000u + 0x0FFFu + ((ax >> 13) & 1u);
    return (uint16_t)(sign | (r >> 13));
  }
  // subnormal half: round |f| / 2^-24 to an integer
  const uint32_t e = ax >> 23;
  if (e < 102u) return (uint16_t)sign;  // below 2^-25: rounds to zero
  const uint32_t m = (ax & 0x007FFFFFu) | 0x00800000u;
  const uint32_t sh = 126u - e;  // 14..24
  uint32_t q = m >> sh;
  const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
  if (rem > half || (rem == half && (q & 1u))) ++q;
  return (uint16_t)(sign | q);
}

