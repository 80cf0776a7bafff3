/*
 * st_math.h - f32 helpers shared by host and device code of the FLIP substep.
 *
 * WGSL leaves the precision of exp() and the evaluation of mix()/smoothstep()
 * to the implementation.  This project fixes one definition of each so the HIP
 * kernels give the same bits on every run: the definitions below, compiled with
 * floating-point contraction off, are the ones the kernels use.
 */
#ifndef ST_MATH_H
#define ST_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define ST_HD __host__ __device__ inline
#else
#define ST_HD static inline
#endif

/* exp(x) in f32.  Cody-Waite range reduction x = k*ln2 + r with ln2 split into
 * 355/512 (exact for k*hi, |k| < 2^15) and the remainder, then the degree-7
 * Taylor polynomial of e^r on |r| <= ln2/2 in Horner form (truncation below
 * 6e-9 relative), then an exact scaling by 2^k.  Overflow gives +inf,
 * underflow 0. */
ST_HD float st_expf(float x) {
  if (x != x) return x;
  if (x > 88.8f) return INFINITY;
  if (x < -104.0f) return 0.0f;
  const float inv_ln2 = 1.44269504088896341f;
  const float ln2_hi = 0.693359375f;             /* 355/512 */
  const float ln2_lo = -2.12194440054690583e-4f; /* ln2 - 355/512 */
  float kf = rintf(x * inv_ln2);
  float r = x - kf * ln2_hi;
  r = r - kf * ln2_lo;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int k = (int)kf;
  /* two exact steps so 2^k never overflows on its own for results near the
   * smallest normals */
  int k1 = k / 2;
  int k2 = k - k1;
  return ldexpf(ldexpf(p, k1), k2);
}

/* sin and cos of x in f32 (used for |x| < 8).  k = rint(x * 2/pi); r = x - k*pi/2
 * with pi/2 in three parts whose products with small k are exact (Cody-Waite);
 * Taylor polynomials of sin (to r^11) and cos (to r^12) on |r| <= pi/4, whose
 * truncation is below 2e-9; then the signs and swap of quadrant k mod 4. */
ST_HD void st_sincosf(float x, float* sn, float* cs) {
  const float two_over_pi = 0.636619772367581343f;
  const float dp1 = 1.5703125f;
  const float dp2 = 4.837512969970703125e-4f;
  const float dp3 = 7.54978995489188216e-8f;
  float kf = rintf(x * two_over_pi);
  float r = x - kf * dp1;
  r = r - kf * dp2;
  r = r - kf * dp3;
  float z = r * r;
  float sp = -2.5052108385441718e-8f;
  sp = sp * z + 2.7557319223985893e-6f;
  sp = sp * z - 1.9841269841269841e-4f;
  sp = sp * z + 8.3333333333333332e-3f;
  sp = sp * z - 1.6666666666666666e-1f;
  float s = r + r * (z * sp);
  float cp = 2.0876756987868099e-9f;
  cp = cp * z - 2.7557319223985888e-7f;
  cp = cp * z + 2.4801587301587302e-5f;
  cp = cp * z - 1.3888888888888889e-3f;
  cp = cp * z + 4.1666666666666664e-2f;
  cp = cp * z - 0.5f;
  float c = 1.0f + z * cp;
  int q = ((int)kf) & 3;
  *sn = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
  *cs = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

/* sin and cos of x in f32 for |x| < 1.6e6 (ocean wave phases grow with time).
 * The reduction x = k*pi/2 + r is done in f64 with pi/2 split into
 * hi = 6746518852 / 2^32 (33 significant bits, so k * hi is exact for
 * |k| < 2^20) and lo = pi/2 - hi; r rounded to f32 goes through the same
 * Taylor polynomials as st_sincosf. */
ST_HD void st_sincosf_wide(float x, float* sn, float* cs) {
  const double two_over_pi = 0.63661977236758134308;
  const double pio2_hi = 1.5707963267341256142;  /* 6746518852 / 2^32 */
  const double pio2_lo = 6.077100506506192e-11;
  double xd = (double)x;
  double kd = rint(xd * two_over_pi);
  double rd = xd - kd * pio2_hi;
  rd = rd - kd * pio2_lo;
  float r = (float)rd;
  float z = r * r;
  float sp = -2.5052108385441718e-8f;
  sp = sp * z + 2.7557319223985893e-6f;
  sp = sp * z - 1.9841269841269841e-4f;
  sp = sp * z + 8.3333333333333332e-3f;
  sp = sp * z - 1.6666666666666666e-1f;
  float s = r + r * (z * sp);
  float cp = 2.0876756987868099e-9f;
  cp = cp * z - 2.7557319223985888e-7f;
  cp = cp * z + 2.4801587301587302e-5f;
  cp = cp * z - 1.3888888888888889e-3f;
  cp = cp * z + 4.1666666666666664e-2f;
  cp = cp * z - 0.5f;
  float c = 1.0f + z * cp;
  int q = (int)(((long long)kd) & 3);
  *sn = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
  *cs = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

ST_HD float st_mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
ST_HD float st_clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
/* WGSL smoothstep: t = clamp((x - lo) / (hi - lo), 0, 1); t * t * (3 - 2 t) */
ST_HD float st_smoothstepf(float lo, float hi, float x) {
  float t = st_clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
ST_HD float st_fractf(float x) { return x - floorf(x); }
ST_HD uint32_t st_f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
ST_HD float st_u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
/* WGSL i32(f32): truncation toward zero; out-of-range values clamp to
 * [-2^31, 2^31 - 128] (the ends of the i32 range that f32 reaches); NaN -> 0 */
ST_HD int32_t st_f2i_sat_portable(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483520.0f) return 2147483520;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}
/* WGSL u32(f32): truncation toward zero, clamped to [0, 2^32 - 256]; NaN -> 0 */
ST_HD uint32_t st_f2u_sat_portable(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967040.0f) return 4294967040u;
  return (uint32_t)x;
}
/* The same conversions.  On the GPU, v_cvt_i32_f32 / v_cvt_u32_f32 truncate,
 * saturate to the full integer range and turn NaN into 0, so one instruction
 * and an integer min give the result (checked against the portable form on the
 * device for all 2^32 inputs by the test suite). */
ST_HD int32_t st_f2i_sat(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  int32_t r;
  __asm__("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r < 2147483520 ? r : 2147483520;
#else
  return st_f2i_sat_portable(x);
#endif
}
/* st_f2i_sat for a value known to be below 2147483520 (or NaN): the conversion
 * alone, without the clamp that cannot act */
ST_HD int32_t st_f2i_nosat(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  int32_t r;
  __asm__("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
#else
  return st_f2i_sat_portable(x);
#endif
}
ST_HD uint32_t st_f2u_sat(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  __asm__("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r < 4294967040u ? r : 4294967040u;
#else
  return st_f2u_sat_portable(x);
#endif
}

/* a / n for n in 1..6, equal to the correctly rounded f32 division, given
 * y = 1.0f / n (itself correctly rounded).  q = a*y is within an ulp of a/n; the
 * residual r = a - q*n is exact in one fused multiply-add, and q + r*y rounds to
 * the quotient (Markstein's correction).  An exhaustive run over all 2^32 values
 * of a for each n finds this equal to the division for 2^-124 <= |a| < inf; it
 * differs only below 2^-125 for n = 6 (gradual underflow), at +-0 and at +-inf,
 * so |a| outside that range (and NaN) takes the division itself. */
ST_HD float st_div_small(float a, float n, float y) {
  uint32_t ua = st_f2u(a) & 0x7fffffffu;
  if (ua - 0x01800000u < 0x7e000000u) {
    float q = a * y;
    float r = fmaf(-q, n, a);
    return fmaf(r, y, q);
  }
  return a / n;
}

#endif /* ST_MATH_H */
