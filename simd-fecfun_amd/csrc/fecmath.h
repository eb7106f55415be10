// Device math for the sum-product decoder and the noise generator.
//
// tanh/atanh: the reference evaluates them with vectorclass (src/LDPCDec/min_find.pl:39,45).
// The double decoder uses the forms written from the mathematics below -- exp by
// Cody-Waite reduction plus a Taylor polynomial, log from the series
// 2*atanh((m-1)/(m+1)), atanh from its Taylor series below 0.25 and
// 0.5*log((1+x)/(1-x)) above -- evaluated in a fixed order (the library is compiled
// with -ffp-contract=off), so its results are reproducible bit for bit on the CPU.
// The float decoder uses tanh_half_fast / two_atanh_fast, built on the hardware
// exp2 / log2 / reciprocal approximations.
//
// fastm2log / sincos_norm: the reference's own approximations of -2*ln(x) and of
// sin/cos(pi*x) used in its Box-Muller transform (src/RNGfunc/prng_functions.cpp:81-141).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fecmath {

__device__ __forceinline__ float expm1_poly(float r) {
  float p = 1.0f / 362880.0f;
  p = p * r + 1.0f / 40320.0f;
  p = p * r + 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  return p * r;
}

__device__ __forceinline__ float expm1_own(float y) {
  if (y < -88.0f) return -1.0f;
  if (y > 88.0f) return __int_as_float(0x7f800000);
  if (fabsf(y) < 0.34657359f) return expm1_poly(y);
  float k = __builtin_rintf(y * 1.44269504f);
  float r = (y - k * 0.693145751953125f) - k * 1.42860677e-06f;
  float p = expm1_poly(r);
  float s = __int_as_float(((int)k + 127) << 23);
  return s * p + (s - 1.0f);
}

__device__ __forceinline__ float log_own(float v) {
  uint32_t u = __float_as_uint(v);
  int e = (int)((u >> 23) & 0xff) - 127;
  float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = 1.0f / 13.0f;
  p = p * s2 + 1.0f / 11.0f;
  p = p * s2 + 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  p = p * s2 + 1.0f;
  float lm = 2.0f * s * p;
  float fe = (float)e;
  return fe * 0.693145751953125f + (fe * 1.42860677e-06f + lm);
}

__device__ __forceinline__ float tanh_own(float x) {
  float a = fabsf(x);
  float e = expm1_own(-2.0f * a);
  float t = (0.0f - e) / (2.0f + e);
  return x < 0.0f ? -t : t;
}

__device__ __forceinline__ float atanh_own(float x) {
  float a = fabsf(x), r;
  if (a < 0.25f) {
    float s2 = a * a;
    float p = 1.0f / 13.0f;
    p = p * s2 + 1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + 1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + 1.0f / 3.0f;
    r = a + a * s2 * p;
  } else {
    r = 0.5f * log_own((1.0f + a) / (1.0f - a));
  }
  return x < 0.0f ? -r : r;
}

// ---- float32 sum-product factors from the hardware exp2 / log2 / reciprocal ----
//
// The float32 decoder evaluates tanh(A/2) and 2 atanh(p) for every edge; these forms
// take 13-15 instructions each, branch-free (both halves computed, one selected), in
// place of the series with IEEE divisions above.  Accuracy: within 4 units in the
// last place of the correctly rounded values (the polynomial halves within 2; the
// hardware v_exp_f32 / v_log_f32 / v_rcp_f32 are 1-unit approximations).

// tanh(a / 2) for a >= 0.  Below a = 1.5: x + x^3 q(x^2), x = a/2, q a degree-4
// minimax fit of (tanh(x)/x - 1)/x^2 on [0, 0.75].  Above: (1 - e) / (1 + e) with
// e = exp(-a) = 2^(-a log2(e)), where the cancellation in 1 - e costs nothing.
__device__ __forceinline__ float tanh_half_fast(float a) {
  const float x = a * 0.5f, s = x * x;
  float q = -0.004716447088867426f;
  q = __builtin_fmaf(q, s, 0.01958320289850235f);
  q = __builtin_fmaf(q, s, -0.053344596177339554f);
  q = __builtin_fmaf(q, s, 0.13325409591197968f);
  q = __builtin_fmaf(q, s, -0.33332958817481995f);
  const float tp = __builtin_fmaf(x * s, q, x);
  const float e = __builtin_amdgcn_exp2f(a * -1.44269504f);
  const float te = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
  return a < 1.5f ? tp : te;
}

// 2 atanh(p) for 0 <= p < 1.  Below p = 0.5: y + y p^2 q(p^2), y = 2p, q a degree-6
// minimax fit of (atanh(p)/p - 1)/p^2 on [0, 0.5].  Above: ln((1 + p) / (1 - p)) =
// ln 2 * log2((1 + p) * (1 / (1 - p))), with 1 - p exact for p >= 0.5.
__device__ __forceinline__ float two_atanh_fast(float p) {
  const float s = p * p, y = p + p;
  float q = 0.177582249045372f;
  q = __builtin_fmaf(q, s, 0.022989485412836075f);
  q = __builtin_fmaf(q, s, 0.1033991202712059f);
  q = __builtin_fmaf(q, s, 0.1095632016658783f);
  q = __builtin_fmaf(q, s, 0.14295953512191772f);
  q = __builtin_fmaf(q, s, 0.1999967098236084f);
  q = __builtin_fmaf(q, s, 0.3333333730697632f);
  const float mp = __builtin_fmaf(y * s, q, y);
  const float ml = __builtin_amdgcn_logf((1.0f + p) * __builtin_amdgcn_rcpf(1.0f - p)) * 0.693147182f;
  return p < 0.5f ? mp : ml;
}

// ---- double versions (Vec8d_prod) ----
__device__ __forceinline__ double expm1_poly(double r) {
  double p = 1.0 / 6227020800.0;
  const double c[12] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,
                        1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,     1.0 / 120.0,
                        1.0 / 24.0,        1.0 / 6.0,        0.5,             1.0};
#pragma unroll
  for (int i = 0; i < 12; ++i) p = p * r + c[i];
  return p * r;
}

__device__ __forceinline__ double expm1_own(double y) {
  if (y < -708.0) return -1.0;
  if (y > 709.0) return __longlong_as_double(0x7ff0000000000000ll);
  if (fabs(y) < 0.34657359027997264) return expm1_poly(y);
  double k = __builtin_rint(y * 1.4426950408889634);
  double r = (y - k * 0.6931471803691238) - k * 1.9082149292705877e-10;
  double p = expm1_poly(r);
  double s = __longlong_as_double((long long)((int)k + 1023) << 52);
  return s * p + (s - 1.0);
}

__device__ __forceinline__ double log_own(double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  int e = (int)((u >> 52) & 0x7ff) - 1023;
  double m = __longlong_as_double((long long)((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
  if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
  double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double p = 1.0 / 21.0;
#pragma unroll
  for (int k = 19; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
  double fe = (double)e;
  return fe * 0.6931471803691238 + (fe * 1.9082149292705877e-10 + 2.0 * s * p);
}

__device__ __forceinline__ double tanh_own(double x) {
  double a = fabs(x);
  double e = expm1_own(-2.0 * a);
  double t = (0.0 - e) / (2.0 + e);
  return x < 0.0 ? -t : t;
}

__device__ __forceinline__ double atanh_own(double x) {
  double a = fabs(x), r;
  if (a < 0.25) {
    double s2 = a * a, p = 1.0 / 29.0;
#pragma unroll
    for (int k = 27; k >= 3; k -= 2) p = p * s2 + 1.0 / (double)k;
    r = a + a * s2 * p;
  } else {
    r = 0.5 * log_own((1.0 + a) / (1.0 - a));
  }
  return x < 0.0 ? -r : r;
}

// ---- Box-Muller pieces of the reference (prng_functions.cpp:81-141) ----
__device__ __forceinline__ float fastm2log(float x) {
  const float ca = -0.6296735f, cb = 1.466967f, m2log2 = -2.0f * 0.6931472f;
  int32_t xi = __float_as_int(x);
  int32_t ex = xi >> 23;
  int32_t more1p5 = (xi & 0x00400000) ^ 0x00400000;
  int32_t div2 = more1p5 >> 22;
  int32_t expbits = (more1p5 << 1) | 0x3f000000;
  float frac = __int_as_float((xi & 0x007fffff) | expbits) - 1.0f;
  ex = ex - (126 | div2);
  float r = __builtin_fmaf(__builtin_fmaf(ca, frac, cb), frac, (float)ex);
  return r * m2log2;
}

// x in [-1, 1) is the angle divided by pi
__device__ __forceinline__ void sincos_norm(float x, float* s, float* c) {
  const float ca = 0.99940307f, cb = -4.8911858f, cc = 3.5838442f;
  uint32_t xs = __float_as_uint(x) & 0x80000000u;
  float ax = fabsf(x);
  float axm = ax - 0.5f;
  uint32_t gt = 0x80000000u & ~__float_as_uint(axm);
  float xcos = gt ? ax - 1.0f : ax;
  xcos = xcos * xcos;
  float xsin = axm * axm;
  float lcos = __builtin_fmaf(__builtin_fmaf(xcos, cc, cb), xcos, ca);
  float lsin = __builtin_fmaf(__builtin_fmaf(xsin, cc, cb), xsin, ca);
  *c = __uint_as_float(__float_as_uint(lcos) | gt);
  *s = __uint_as_float(__float_as_uint(lsin) | xs);
}

// SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014)
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace fecmath
