// Scalar / vector arithmetic shared by the HIP kernels and the host code.
//
// Every routine here is written so that the host compiler and the gfx950
// compiler evaluate it with the same IEEE-754 single-precision operations in
// the same order (all code is built with -ffp-contract=off, fused
// multiply-adds appear only as explicit fmaf calls, and division / sqrtf are
// correctly rounded on both sides).  That is what makes hit records and
// per-path radiance reproducible bit for bit between host and device.
//
// Operation order follows the reference's tensor helpers:
//   VectorDot / VectorCross / VectorNormalize / VectorReflect
//   (LibTensorRay/Tensor/Tensor.cpp:684-713, 811-851).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define TR_HD __host__ __device__ inline

namespace trc {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kPiOver2 = 1.57079632679489661923f;
constexpr float kPiOver4 = 0.78539816339744830962f;

// Reference epsilons (LibTensorRay/Renderer/Config.h:33-35).
constexpr float kEpsilon = 1e-6f;
constexpr float kEdgeEpsilon = 1e-8f;
constexpr float kShadowEpsilon = 1e-2f;

struct F3 {
  float x, y, z;
};

TR_HD F3 mk3(float x, float y, float z) { return F3{x, y, z}; }
TR_HD F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
TR_HD F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
TR_HD F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
TR_HD F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
TR_HD F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
TR_HD F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
TR_HD F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
TR_HD F3 operator/(F3 a, F3 b) { return F3{a.x / b.x, a.y / b.y, a.z / b.z}; }
TR_HD F3& operator+=(F3& a, F3 b) { a = a + b; return a; }
TR_HD F3& operator-=(F3& a, F3 b) { a = a - b; return a; }

TR_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TR_HD F3 cross(F3 a, F3 b) {
  return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
TR_HD float len2(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
TR_HD float length(F3 a) { return sqrtf(len2(a)); }
TR_HD F3 normalize(F3 a) { return a / sqrtf(len2(a)); }
// reflect(in, n) = in + 2 * dot(-in, n) * n
TR_HD F3 reflect(F3 in, F3 n) { return in + (2.0f * dot(-in, n)) * n; }
TR_HD F3 f3_get(const float* p) { return F3{p[0], p[1], p[2]}; }
TR_HD bool finite3(F3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
TR_HD float maxf(float a, float b) { return a > b ? a : b; }
TR_HD float minf(float a, float b) { return a < b ? a : b; }
TR_HD float clampf(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }
TR_HD float safe_sqrt(float x) { return sqrtf(maxf(x, 0.0f)); }

// ---------------------------------------------------------------------------
// Elementary functions.  Own implementations (Cody-Waite range reduction with
// a two-part pi/2 and truncated Taylor series, whose coefficients are 1/n!),
// so host and device agree exactly.  Accuracy is a few ulp on the ranges the
// integrator uses (|x| <= 4*pi for sin/cos, [-16, 0] for exp2).
// ---------------------------------------------------------------------------
TR_HD void tr_sincos(float x, float* s_out, float* c_out) {
  // k = nearest integer to x / (pi/2)
  const float two_over_pi = 0.636619772367581343f;
  float kf = rintf(x * two_over_pi);
  // pi/2 = hi + lo with hi = fl(pi/2)
  const float pio2_hi = 1.57079637050628662109375f;
  const float pio2_lo = -4.37113900018624283e-8f;
  float r = fmaf(-kf, pio2_hi, x);
  r = fmaf(-kf, pio2_lo, r);
  float r2 = r * r;
  // sin(r) = r - r^3/3! + r^5/5! - r^7/7! + r^9/9! - r^11/11!
  float ps = -2.50521083854417188e-8f;
  ps = fmaf(ps, r2, 2.75573192239858907e-6f);
  ps = fmaf(ps, r2, -1.98412698412698413e-4f);
  ps = fmaf(ps, r2, 8.33333333333333333e-3f);
  ps = fmaf(ps, r2, -1.66666666666666667e-1f);
  float sr = fmaf(ps * r2, r, r);
  // cos(r) = 1 - r^2/2! + r^4/4! - ... + r^12/12!
  float pc = 2.08767569878680990e-9f;
  pc = fmaf(pc, r2, -2.75573192239858907e-7f);
  pc = fmaf(pc, r2, 2.48015873015873016e-5f);
  pc = fmaf(pc, r2, -1.38888888888888889e-3f);
  pc = fmaf(pc, r2, 4.16666666666666667e-2f);
  pc = fmaf(pc, r2, -0.5f);
  float cr = fmaf(pc, r2, 1.0f);
  int k = (int)kf & 3;
  float s, c;
  if (k == 0) { s = sr; c = cr; }
  else if (k == 1) { s = cr; c = -sr; }
  else if (k == 2) { s = -sr; c = -cr; }
  else { s = -cr; c = sr; }
  *s_out = s;
  *c_out = c;
}
TR_HD float tr_sin(float x) { float s, c; tr_sincos(x, &s, &c); return s; }
TR_HD float tr_cos(float x) { float s, c; tr_sincos(x, &s, &c); return c; }

// 2^x for x in a moderate range: 2^x = 2^k * e^(f ln 2), f in [0, 1).
TR_HD float tr_exp2(float x) {
  if (x < -126.0f) return 0.0f;
  if (x > 127.0f) return INFINITY;
  float kf = floorf(x);
  float f = x - kf;
  const float ln2 = 0.693147180559945309f;
  float y = f * ln2;
  // e^y, Taylor to degree 10 (y <= ln 2)
  float p = 2.75573192239858907e-7f;
  p = fmaf(p, y, 2.75573192239858907e-6f);
  p = fmaf(p, y, 2.48015873015873016e-5f);
  p = fmaf(p, y, 1.98412698412698413e-4f);
  p = fmaf(p, y, 1.38888888888888889e-3f);
  p = fmaf(p, y, 8.33333333333333333e-3f);
  p = fmaf(p, y, 4.16666666666666667e-2f);
  p = fmaf(p, y, 1.66666666666666667e-1f);
  p = fmaf(p, y, 0.5f);
  p = fmaf(p, y, 1.0f);
  p = fmaf(p, y, 1.0f);
  return ldexpf(p, (int)kf);
}

// asin on [0, 0.5] by its Maclaurin series: sum_n (2n)! / (4^n (n!)^2 (2n+1)) x^(2n+1)
TR_HD float tr_asin_small(float x) {
  float x2 = x * x;
  float p = 6.447210311889649e-3f;                 // n = 12
  p = fmaf(p, x2, 7.3125258735988454e-3f);         // n = 11
  p = fmaf(p, x2, 8.390335809616815e-3f);          // n = 10
  p = fmaf(p, x2, 9.761609529194078e-3f);          // n = 9
  p = fmaf(p, x2, 1.1551800896139705e-2f);         // n = 8
  p = fmaf(p, x2, 1.396484375e-2f);                // n = 7 (143/10240)
  p = fmaf(p, x2, 1.7352764423076924e-2f);         // n = 6 (231/13312)
  p = fmaf(p, x2, 2.2372159090909092e-2f);         // n = 5 (63/2816)
  p = fmaf(p, x2, 3.0381944444444444e-2f);         // n = 4 (35/1152)
  p = fmaf(p, x2, 4.4642857142857144e-2f);         // n = 3 (5/112)
  p = fmaf(p, x2, 7.5e-2f);                        // n = 2 (3/40)
  p = fmaf(p, x2, 1.6666666666666666e-1f);         // n = 1 (1/6)
  return fmaf(p * x2, x, x);
}

// acos via asin: acos(x) = pi/2 - asin(x) for |x| <= 0.5,
// 2 asin(sqrt((1-x)/2)) for x > 0.5, pi - 2 asin(sqrt((1+x)/2)) for x < -0.5.
TR_HD float tr_acos(float x) {
  x = clampf(x, -1.0f, 1.0f);
  if (x > 0.5f) return 2.0f * tr_asin_small(sqrtf((1.0f - x) * 0.5f));
  if (x < -0.5f) return kPi - 2.0f * tr_asin_small(sqrtf((1.0f + x) * 0.5f));
  if (x >= 0.0f) return kPiOver2 - tr_asin_small(x);
  return kPiOver2 + tr_asin_small(-x);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

TR_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

TR_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = mulhi32(M0, c.x), lo0 = M0 * c.x;
    uint32_t hi1 = mulhi32(M1, c.z), lo1 = M1 * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 24 high bits -> float in [0, 1)
TR_HD float u32_to_unit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// Random-number domains: one key space per estimator so that interior,
// boundary and guiding samples never share a stream.
enum RngDomain : uint32_t {
  kRngInterior = 0x11u,
  kRngPrimaryEdge = 0x22u,
  kRngPixelEdge = 0x33u,
  kRngDirectEdge = 0x44u,
  kRngIndirectEdge = 0x55u,
  kRngGuide = 0x66u,
};

// Four uniforms for (seed, domain, sample index, block): the block index
// separates the draws of one sample (camera jitter, per-bounce light and
// BSDF samples, ...).
struct Rnd4 {
  float a, b, c, d;
};
TR_HD Rnd4 rng4(uint32_t seed, uint32_t domain, uint64_t index, uint32_t block) {
  U4 v = philox4x32_10(U4{(uint32_t)index, (uint32_t)(index >> 32), block, 0u}, seed, domain);
  return Rnd4{u32_to_unit(v.x), u32_to_unit(v.y), u32_to_unit(v.z), u32_to_unit(v.w)};
}

}  // namespace trc
