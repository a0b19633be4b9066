/* Host restatement of the FMA sequences of pbf_refmath.hpp, started from a
 * chosen first approximation, for the CPU tests: the sequences must give
 * the correctly rounded quotient / root from every float within 1 ulp of
 * the exact value (the accuracy of the hardware's reciprocal and square
 * root estimates), on the operand ranges the device code admits. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

/* a / b from the first reciprocal estimate y0 (b > 0). */
static float div_from(float a, float b, float y0) {
  const float e = fmaf(-b, y0, 1.0f);
  const float y = fmaf(e, y0, y0);
  const float q0 = a * y;
  const float r0 = fmaf(-b, q0, a);
  const float q1 = fmaf(r0, y, q0);
  const float r1 = fmaf(-b, q1, a);
  const float q = fmaf(r1, y, q1);
  return u2f((f2u(q) & 0x7fffffffu) | (f2u(a) & 0x80000000u));
}

/* 1 / b from y0 (the a = 1 form: q0 = y). */
static float recip_from(float b, float y0) {
  const float e = fmaf(-b, y0, 1.0f);
  const float y = fmaf(e, y0, y0);
  const float r0 = fmaf(-b, y, 1.0f);
  const float q1 = fmaf(r0, y, y);
  const float r1 = fmaf(-b, q1, 1.0f);
  return fmaf(r1, y, q1);
}

/* sqrt(x) from the first estimate s. */
static float sqrt_from(float x, float s) {
  const float sm = u2f(f2u(s) - 1u), sp = u2f(f2u(s) + 1u);
  const float rm = fmaf(-sm, s, x), rp = fmaf(-sp, s, x);
  s = rm <= 0.0f ? sm : s;
  s = rp > 0.0f ? sp : s;
  return s;
}

/* Each function returns the number of (input, estimate) cases whose result
 * differs in any bit from the IEEE operation, the estimates being the
 * correctly rounded value and its two neighbours; *first gets the first
 * failing input index (or -1). */
long check_div(const float *a, const float *b, long n, long *first) {
  long bad = 0;
  *first = -1;
  for (long i = 0; i < n; ++i) {
    const float want = a[i] / b[i];
    const uint32_t y = f2u(1.0f / b[i]);
    for (int d = -1; d <= 1; ++d) {
      if (f2u(div_from(a[i], b[i], u2f(y + (uint32_t)d))) != f2u(want)) {
        if (*first < 0) *first = i;
        ++bad;
      }
    }
  }
  return bad;
}

long check_recip(const float *b, long n, long *first) {
  long bad = 0;
  *first = -1;
  for (long i = 0; i < n; ++i) {
    const float want = 1.0f / b[i];
    const uint32_t y = f2u(want);
    for (int d = -1; d <= 1; ++d) {
      if (f2u(recip_from(b[i], u2f(y + (uint32_t)d))) != f2u(want)) {
        if (*first < 0) *first = i;
        ++bad;
      }
    }
  }
  return bad;
}

long check_sqrt(const float *x, long n, long *first) {
  long bad = 0;
  *first = -1;
  for (long i = 0; i < n; ++i) {
    const float want = sqrtf(x[i]);
    const uint32_t s = f2u(want);
    for (int d = -1; d <= 1; ++d) {
      if (f2u(sqrt_from(x[i], u2f(s + (uint32_t)d))) != f2u(want)) {
        if (*first < 0) *first = i;
        ++bad;
      }
    }
  }
  return bad;
}
