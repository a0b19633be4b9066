// pbf_refmath.hpp - correctly rounded FP32 division and square root for the
// REF pair terms, cheaper than the general IEEE sequences on the inputs
// those terms see.
//
// The REF passes must reproduce the reference kernels' IEEE-754 results bit
// for bit: x / y and sqrt(x) correctly rounded (round to nearest even).  The
// general sequences handle every input (zero, subnormal, huge and special
// operands) by scaling the operands before a Newton-Raphson refinement and
// fixing up special cases after it.  On the operands of a pair term (a
// distance |r| < h, its components, a density) no scaling is needed, and
// the three divisions of a pair share one divisor, so:
//
//   * division (Markstein's FMA method: P. Markstein, "Computation of
//     elementary functions on the IBM RISC System/6000 processor", IBM J.
//     Res. Develop. 34(1), 1990): with y0 = rcp(b) within 1 ulp of 1/b,
//       e  = fma(-b, y0, 1),   y  = fma(e, y0, y0)      (once per divisor)
//       q0 = a y,              r0 = fma(-b, q0, a),   q1 = fma(r0, y, q0)
//                              r1 = fma(-b, q1, a),   q  = fma(r1, y, q1)
//     gives a / b correctly rounded while every intermediate stays in the
//     normal range: b in [2^-90, 2^23] and a = 0 or 2^-103 <= |a| <= 2^30 b
//     (then |a / b| >= 2^-126 and the residuals r0, r1 are exact).  These
//     are the steps the general sequence runs when its operand scaling is
//     the identity; its final fix-up then only sets the sign (sign a xor
//     sign b), which `with_sign` does for b > 0.
//   * square root: s = sqrt_approx(x) within 1 ulp, then the neighbouring
//     values s- and s+ are tested with the residuals fma(-s-, s, x) and
//     fma(-s+, s, x), whose signs are exact (x <= s- s: sqrt(x) rounds
//     below s; x > s+ s: above), correct for x in [2^-96, FLT_MAX] (no
//     operand scaling, no special value).
//
// Every caller checks these ranges on each pair and takes the general '/'
// and sqrtf outside them, so its results equal theirs for every input.
// Plain FP32 with explicit fmaf; the callers compile without contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace pbf {

// An FP32 divisor with its refined reciprocal.
struct RcpRN {
  float b, y;
};

__device__ __forceinline__ RcpRN rcp_refined(float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  const float e = __builtin_fmaf(-b, y0, 1.0f);
  return RcpRN{b, __builtin_fmaf(e, y0, y0)};
}

// |q| with the sign of a: the quotient's sign for a divisor b > 0, also for
// a = -0 (whose refinement yields +0).
__device__ __forceinline__ float with_sign(float q, float a) {
  return __uint_as_float((__float_as_uint(q) & 0x7fffffffu) |
                         (__float_as_uint(a) & 0x80000000u));
}

// a / b correctly rounded, for b in [2^-90, 2^23] and a = 0 or
// 2^-103 <= |a| <= 2^30 b.
__device__ __forceinline__ float div_rn(float a, RcpRN d) {
  const float q0 = a * d.y;
  const float r0 = __builtin_fmaf(-d.b, q0, a);
  const float q1 = __builtin_fmaf(r0, d.y, q0);
  const float r1 = __builtin_fmaf(-d.b, q1, a);
  return with_sign(__builtin_fmaf(r1, d.y, q1), a);
}

// 1 / b correctly rounded, for b in [2^-90, 2^90] (a = 1: q0 = y).
__device__ __forceinline__ float recip_rn(float b) {
  const RcpRN d = rcp_refined(b);
  const float r0 = __builtin_fmaf(-b, d.y, 1.0f);
  const float q1 = __builtin_fmaf(r0, d.y, d.y);
  const float r1 = __builtin_fmaf(-b, q1, 1.0f);
  return __builtin_fmaf(r1, d.y, q1);
}

// sqrt(x) correctly rounded, for x in [2^-96, FLT_MAX].
__device__ __forceinline__ float sqrt_rn(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  const float rm = __builtin_fmaf(-sm, s, x);
  const float rp = __builtin_fmaf(-sp, s, x);
  s = rm <= 0.0f ? sm : s;
  s = rp > 0.0f ? sp : s;
  return s;
}

// Whether each of three values is 0 or at least 2^-103 in magnitude (its
// biased exponent above 23): 2|bits| - 1, unsigned (0 wraps to the
// largest), is below 2 (24 << 23) - 1 exactly for the tiny non-zero ones.
__device__ __forceinline__ bool no_tiny3(float a, float b, float c) {
  const unsigned ta = (__float_as_uint(a) << 1) - 1u;
  const unsigned tb = (__float_as_uint(b) << 1) - 1u;
  const unsigned tc = (__float_as_uint(c) << 1) - 1u;
  return min(min(ta, tb), tc) >= 0x17ffffffu;
}

// lo <= x <= hi for positive finite bounds, by the bit patterns (false for
// negative values and NaN).
__device__ __forceinline__ bool in_range(float x, unsigned lo, unsigned hi) {
  return __float_as_uint(x) - lo <= hi - lo;
}

// 2^-90 <= b <= 2^90: recip_rn applies.
__device__ __forceinline__ bool recip_ok(float b) {
  return in_range(b, 0x12800000u, 0x6c800000u);
}

// 2^-96 <= x <= FLT_MAX: sqrt_rn applies.
__device__ __forceinline__ bool sqrt_ok(float x) {
  return in_range(x, 0x0f800000u, 0x7f7fffffu);
}

// A pair (r, r^2) with r^2 < h^2: sqrt_rn(r^2) and div_rn of each
// component by it apply when h^2 <= 2^40 and 2^-96 <= r^2 (then
// 2^-48 <= |r| < 2^20) and no component is tiny and non-zero (each is at
// most |r| (1 + 2^-22)).  One branch-free test (h2_ok is the scene's).
__device__ __forceinline__ bool pair_rn_ok(float rx, float ry, float rz,
                                           float r2, bool h2_ok) {
  return h2_ok & (r2 >= 0x1p-96f) & no_tiny3(rx, ry, rz);
}

// |r| and r / |r| of a pair with r^2 < h^2 (r^2 formed by the caller),
// each correctly rounded: sqrt(r^2), then one division per component.
// h2_ok: the scene's h^2 <= 2^40.  RN: the shortened sequences where the
// pair's operands allow (fewer instructions per pair, a range test per
// pair); without RN the general ones only (the same bits; a shorter
// dependent chain for the scenes whose time a wave's latency sets).
template <bool RN>
__device__ __forceinline__ float ref_unit(float rx, float ry, float rz,
                                         float r2, bool h2_ok, float &ux,
                                         float &uy, float &uz) {
  float rl;
  if (RN && pair_rn_ok(rx, ry, rz, r2, h2_ok)) {
    rl = sqrt_rn(r2);
    const RcpRN d = rcp_refined(rl);
    ux = div_rn(rx, d);
    uy = div_rn(ry, d);
    uz = div_rn(rz, d);
  } else {
    rl = sqrtf(r2);
    ux = rx / rl;
    uy = ry / rl;
    uz = rz / rl;
  }
  return rl;
}

// sqrt(x) correctly rounded (RN as for ref_unit).
template <bool RN>
__device__ __forceinline__ float ref_sqrt(float x) {
  return RN && sqrt_ok(x) ? sqrt_rn(x) : sqrtf(x);
}

// 1 / b correctly rounded (RN as for ref_unit).
template <bool RN>
__device__ __forceinline__ float ref_recip(float b) {
  return RN && recip_ok(b) ? recip_rn(b) : 1.0f / b;
}

}  // namespace pbf
