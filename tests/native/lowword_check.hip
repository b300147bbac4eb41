// Host-side check of the compensated state's 16-bit low word
// (sparc_amd/csrc/lowword.h): for FP64 values over the FP32 range,
//   * hi + decode(encode(v)) is within one low-word step of v (ulp(hi)/2^16,
//     2^-39 relative), within half a step unless the word clamps;
//   * hi + lo is exact in FP64 and encoding it again gives the same word;
//   * the word is 0 for an FP32-representable v.
// Prints "OK <count>" or the first failure.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>

namespace sparc {
#include "../../sparc_amd/csrc/lowword.h"
}

int main() {
  using namespace sparc;
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> m(1.0, 2.0);
  std::uniform_int_distribution<int> ex(-120, 120);
  long n = 0, clamped = 0;
  for (int i = 0; i < 2000000; ++i) {
    const double v = (i & 1 ? -1.0 : 1.0) * std::ldexp(m(g), ex(g));
    const float hi = (float)v;
    const short q = lo_encode(hi, v);
    const double lo = lo_decode(hi, q);
    const double w = (double)hi + lo;
    const double step = std::ldexp(1.0, lo_exponent(hi) - 166);
    const double err = std::fabs(w - v);
    const bool clamp = q == 32767 || q == -32767;
    clamped += clamp;
    if (err > (clamp ? 1.0 : 0.5) * step || err > std::ldexp(std::fabs(v), -39)) {
      std::printf("FAIL bound v=%.17g hi=%.9g q=%d err=%.3g step=%.3g\n", v, hi, q, err, step);
      return 1;
    }
    if (w - (double)hi != lo || lo_encode(hi, w) != q) {
      std::printf("FAIL exact v=%.17g q=%d\n", v, q);
      return 1;
    }
    if (lo_encode(hi, (double)hi) != 0) {
      std::printf("FAIL zero hi=%.9g\n", hi);
      return 1;
    }
    ++n;
  }
  // subnormal and zero high words take the smallest step
  for (float hi : {0.0f, 1e-45f, 1e-40f, -3e-39f}) {
    const double v = (double)hi + 1e-50;
    const short q = lo_encode(hi, v);
    if (std::fabs((double)hi + lo_decode(hi, q) - v) > 0.5 * std::ldexp(1.0, 1 - 166)) {
      std::printf("FAIL subnormal hi=%.9g\n", hi);
      return 1;
    }
    ++n;
  }
  std::printf("OK %ld clamped %ld\n", n, clamped);
  return 0;
}
