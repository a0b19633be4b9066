// Device check of pbf_refmath.hpp against the compiler's general IEEE
// sequences ('/', sqrtf, 1/x) on gfx950, for the GPU tests.  Every kernel
// reads n inputs and counts the results that differ in any bit (NaN results
// must both be NaN); the lowest failing index is kept.
#include <hip/hip_runtime.h>

#include "../../position-based-fluids_amd/csrc/pbf_refmath.hpp"

#pragma clang fp contract(off)

namespace {

__device__ bool same(float a, float b) {
  return __float_as_uint(a) == __float_as_uint(b) ||
         (__builtin_isnan(a) && __builtin_isnan(b));
}

__device__ void note(bool ok, long i, unsigned long long *bad,
                     long long *first) {
  if (!ok) {
    atomicAdd(bad, 1ull);
    atomicMin(first, (long long)i);
  }
}

// 0: div_rn(a, rcp_refined(b)) against a / b; in = (a, b) pairs.
// 1: recip_rn(b) against 1 / b.
// 2: sqrt_rn(x) against sqrtf(x).
// 3: ref_unit(r) against sqrtf(r^2) and r_k / |r|; in = (x, y, z) triples,
//    r^2 formed left to right without contraction; h^2 <= 2^40 assumed.
// 4: ref_sqrt(x) against sqrtf(x) (any x).
// 5: ref_recip(b) against 1 / b (any b).
// 6, 7, 8: no_tiny3(x, x, x), sqrt_ok(x), recip_ok(x) against the label
//    given beside each value; in = (x, label) pairs, label 0 or 1.
__global__ void k_check(int which, const float *__restrict__ in, long n,
                        unsigned long long *bad, long long *first) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (which) {
    case 0: {
      const float a = in[2 * i], b = in[2 * i + 1];
      note(same(pbf::div_rn(a, pbf::rcp_refined(b)), a / b), i, bad, first);
      break;
    }
    case 1: {
      const float b = in[i];
      note(same(pbf::recip_rn(b), 1.0f / b), i, bad, first);
      break;
    }
    case 2: {
      const float x = in[i];
      note(same(pbf::sqrt_rn(x), sqrtf(x)), i, bad, first);
      break;
    }
    case 3: {
      const float rx = in[3 * i], ry = in[3 * i + 1], rz = in[3 * i + 2];
      const float r2 = rx * rx + ry * ry + rz * rz;
      float ux, uy, uz;
      const float rl = pbf::ref_unit<true>(rx, ry, rz, r2, true, ux, uy, uz);
      const float rl0 = sqrtf(r2);
      note(same(rl, rl0) && same(ux, rx / rl0) && same(uy, ry / rl0) &&
               same(uz, rz / rl0),
           i, bad, first);
      break;
    }
    case 4: {
      const float x = in[i];
      note(same(pbf::ref_sqrt<true>(x), sqrtf(x)), i, bad, first);
      break;
    }
    case 5: {
      const float b = in[i];
      note(same(pbf::ref_recip<true>(b), 1.0f / b), i, bad, first);
      break;
    }
    case 6:
    case 7:
    case 8: {
      const float x = in[2 * i];
      const bool want = in[2 * i + 1] != 0.0f;
      const bool got = which == 6   ? pbf::no_tiny3(x, x, x)
                       : which == 7 ? pbf::sqrt_ok(x)
                                    : pbf::recip_ok(x);
      note(got == want, i, bad, first);
      break;
    }
    default:
      break;
  }
}

}  // namespace

// Runs check `which` over n inputs of (which: 0, 6, 7, 8 -> 2; 3 -> 3;
// else 1) floats each from host memory.  Returns the mismatch count (and the first
// failing index in *first, -1 if none), or -1 on a HIP error or a bad
// argument.
extern "C" long long refmath_check(int which, const float *host_in, long n,
                                   long long *first) {
  if (which < 0 || which > 8 || n < 0 || !first || (n > 0 && !host_in))
    return -1;
  *first = -1;
  if (n == 0) return 0;
  const int per = which == 3 ? 3 : (which == 0 || which >= 6) ? 2 : 1;
  float *d_in = nullptr;
  unsigned long long *d_bad = nullptr;
  long long *d_first = nullptr;
  const size_t bytes = (size_t)n * per * sizeof(float);
  long long result = -1;
  unsigned long long bad = 0;
  long long fi = 0;
  const long long none = 0x7fffffffffffffffll;
  if (hipMalloc(&d_in, bytes) != hipSuccess) goto done;
  if (hipMalloc(&d_bad, sizeof(bad)) != hipSuccess) goto done;
  if (hipMalloc(&d_first, sizeof(fi)) != hipSuccess) goto done;
  if (hipMemcpy(d_in, host_in, bytes, hipMemcpyHostToDevice) != hipSuccess)
    goto done;
  if (hipMemcpy(d_bad, &bad, sizeof(bad), hipMemcpyHostToDevice) !=
      hipSuccess)
    goto done;
  if (hipMemcpy(d_first, &none, sizeof(none), hipMemcpyHostToDevice) !=
      hipSuccess)
    goto done;
  k_check<<<(unsigned)((n + 255) / 256), 256>>>(which, d_in, n, d_bad,
                                                 d_first);
  if (hipGetLastError() != hipSuccess) goto done;
  if (hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost) !=
      hipSuccess)
    goto done;
  if (hipMemcpy(&fi, d_first, sizeof(fi), hipMemcpyDeviceToHost) !=
      hipSuccess)
    goto done;
  *first = bad ? fi : -1;
  result = (long long)bad;
done:
  if (d_in) (void)hipFree(d_in);
  if (d_bad) (void)hipFree(d_bad);
  if (d_first) (void)hipFree(d_first);
  return result;
}
