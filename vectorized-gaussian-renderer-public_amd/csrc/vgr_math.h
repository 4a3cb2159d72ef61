// Deterministic elementary functions for the Gaussian renderer.
//
// Every function here is built only from IEEE-754 binary32 operations that are
// correctly rounded on both the x86-64 host (SSE/AVX, default round-to-nearest)
// and gfx950 (v_add/v_mul/v_fma/v_rndne/v_ldexp, correctly rounded divide and
// square root): add, multiply, fused multiply-add, rint, ldexp and integer bit
// manipulation.  Compiled with -ffp-contract=off on both sides, the same input
// therefore produces the same bits on the CPU and on the GPU.  That is what lets
// integer outcomes that hinge on a float comparison (ceil of a splat radius,
// the transmittance cut-off, the strictly-largest-weight semantic winner) agree
// exactly between the renderer and its CPU check.
//
// The reference renderer calls CUDA's fast intrinsics (__expf, __logf,
// powf under --use_fast_math; renderer_cuda.cu:199, :462, :164).  Those are
// not reproducible off NVIDIA hardware, so this header supplies its own
// approximations with stated accuracy:
//   vgr_exp2f : max relative error 1.7e-7 on [-126, 127], exact at integers
//   vgr_expf  : relative error <= 1.7e-7 + 7.5e-8 |x| (rounding of x*log2(e));
//               1.6e-6 on the compositor's [-20, 0], 6e-7 where alpha >= 1/255
//   vgr_logf  : error <= 4e-7 * max(1, |ln x|) for normal x
// These bounds are checked against float64 numpy in the test suite.
//
// The header is plain C99 when included from C, and host+device inline
// functions when included from HIP.
#ifndef VGR_MATH_H
#define VGR_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VGR_HD __host__ __device__ __forceinline__
#define VGR_RINTF(x) rintf(x)
#define VGR_FMAF(a, b, c) fmaf((a), (b), (c))
#define VGR_LDEXPF(x, e) ldexpf((x), (e))
#define VGR_SQRTF(x) sqrtf(x)
#define VGR_FLOORF(x) floorf(x)
#define VGR_CEILF(x) ceilf(x)
#else
#include <math.h>
#include <string.h>
#define VGR_HD static inline
#define VGR_RINTF(x) rintf(x)
#define VGR_FMAF(a, b, c) fmaf((a), (b), (c))
#define VGR_LDEXPF(x, e) ldexpf((x), (e))
#define VGR_SQRTF(x) sqrtf(x)
#define VGR_FLOORF(x) floorf(x)
#define VGR_CEILF(x) ceilf(x)
#endif

#define VGR_ALPHA_THRESHOLD (1.0f / 255.0f)
#define VGR_MAX_ALPHA 0.99f
#define VGR_TRANSMITTANCE_THRESHOLD 1.0e-4f
#define VGR_MIN_ANTIALIAS_COMPENSATION 0.005f

VGR_HD uint32_t vgr_float_bits(float value) {
#if defined(__HIPCC__)
  return __builtin_bit_cast(uint32_t, value);
#else
  uint32_t bits;
  memcpy(&bits, &value, sizeof(bits));
  return bits;
#endif
}

VGR_HD float vgr_bits_float(uint32_t bits) {
#if defined(__HIPCC__)
  return __builtin_bit_cast(float, bits);
#else
  float value;
  memcpy(&value, &bits, sizeof(value));
  return value;
#endif
}

// 2^x.  x is clamped to [-126, 127] so the result stays a normal float (or
// the largest finite power), which keeps denormal handling out of the picture.
VGR_HD float vgr_exp2f(float x) {
  x = x < -126.0f ? -126.0f : x;
  x = x > 127.0f ? 127.0f : x;
  const float n = VGR_RINTF(x);
  const float f = x - n;  // exact: |f| <= 0.5 and shares n's ulp scale
  // 2^f = 1 + f q(f) on [-0.5, 0.5], q of degree 4 (Lawson-weighted Chebyshev
  // least squares on relative error, coefficients rounded to binary32).  The
  // unit constant term makes 2^0 exactly 1.
  float q = 0x1.5bba14p-10f;
  q = VGR_FMAF(q, f, 0x1.3cea88p-7f);
  q = VGR_FMAF(q, f, 0x1.c6b752p-5f);
  q = VGR_FMAF(q, f, 0x1.ebf9bcp-3f);
  q = VGR_FMAF(q, f, 0x1.62e42ap-1f);
  const float p = VGR_FMAF(q, f, 1.0f);
  return VGR_LDEXPF(p, (int)n);
}

// e^x evaluated as 2^(x * log2(e)).
VGR_HD float vgr_expf(float x) {
  return vgr_exp2f(x * 1.44269504088896340736f);
}

// vgr_expf for x <= 0 (or NaN), where the upper clamp of vgr_exp2f never
// acts: the same result bit for bit with one compare and select fewer.  The
// compositor's exponents are Gaussian powers, which pass its test only when
// they are not positive.
VGR_HD float vgr_expf_nonpositive(float x) {
  float y = x * 1.44269504088896340736f;
  y = y < -126.0f ? -126.0f : y;
  const float n = VGR_RINTF(y);
  const float f = y - n;
  float q = 0x1.5bba14p-10f;
  q = VGR_FMAF(q, f, 0x1.3cea88p-7f);
  q = VGR_FMAF(q, f, 0x1.c6b752p-5f);
  q = VGR_FMAF(q, f, 0x1.ebf9bcp-3f);
  q = VGR_FMAF(q, f, 0x1.62e42ap-1f);
  const float p = VGR_FMAF(q, f, 1.0f);
  return VGR_LDEXPF(p, (int)n);
}

// Natural logarithm of a positive, normal, finite x.  x = m * 2^e with
// m in [sqrt(1/2), sqrt(2)); ln(m) = 2 atanh(s), s = (m - 1) / (m + 1),
// evaluated as an odd series through s^11 (|s| <= 0.1716).
VGR_HD float vgr_logf(float x) {
  const uint32_t bits = vgr_float_bits(x);
  int exponent = (int)((bits >> 23) & 0xffu) - 127;
  float mantissa = vgr_bits_float((bits & 0x007fffffu) | 0x3f800000u);
  if (mantissa > 1.41421356237309504880f) {
    mantissa = mantissa * 0.5f;
    exponent += 1;
  }
  const float s = (mantissa - 1.0f) / (mantissa + 1.0f);
  const float s2 = s * s;
  float series = 0x1.745d18p-4f;               // 1/11
  series = VGR_FMAF(series, s2, 0x1.c71c72p-4f);  // 1/9
  series = VGR_FMAF(series, s2, 0x1.24924ap-3f);  // 1/7
  series = VGR_FMAF(series, s2, 0x1.99999ap-3f);  // 1/5
  series = VGR_FMAF(series, s2, 0x1.555556p-2f);  // 1/3
  series = series * s2;
  const float log_mantissa = VGR_FMAF(2.0f * s, series, 2.0f * s);
  return VGR_FMAF((float)exponent, 0x1.62e430p-1f, log_mantissa);
}

VGR_HD float vgr_log2f(float x) {
  return vgr_logf(x) * 1.44269504088896340736f;
}

// Linear RGB to sRGB transfer (IEC 61966-2-1), clamped to [0, 1] first, as the
// reference's linear_to_srgb (renderer_cuda.cu:161-166).
VGR_HD float vgr_linear_to_srgb(float value) {
  value = value < 0.0f ? 0.0f : value;
  value = value > 1.0f ? 1.0f : value;
  if (value <= 0.0031308f) {
    return 12.92f * value;
  }
  const float power = vgr_exp2f(vgr_log2f(value) * (1.0f / 2.4f));
  return 1.055f * power - 0.055f;
}

#endif  // VGR_MATH_H
