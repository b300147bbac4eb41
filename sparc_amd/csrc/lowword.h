// The 16-bit low word of the compensated FP32 state (host and device).
// Included inside namespace sparc by kernels.h.
#pragma once

// Low word of the compensated FP32 state: v = hi + q * 2^(e - 166), with q a
// 16-bit integer and e = max(biased exponent of hi, 1), i.e. q counts steps of
// ulp(hi) / 2^16. hi = fl32(v) leaves |v - hi| <= ulp(hi) / 2, so |q| <= 2^15
// (clamped to 32767) and hi + lo carries 40 significant bits. The scale is a
// power of two built from its exponent field, so encoding and decoding are
// exact apart from the rounding of q.
__host__ __device__ inline int lo_exponent(float hi) {
  union { float f; unsigned u; } b{hi};
  const int e = (int)((b.u >> 23) & 255u);
  return e < 1 ? 1 : e;
}
__host__ __device__ inline double lo_pow2(int k) {  // 2^k for -1022 <= k <= 1023
  union { unsigned long long u; double d; } b{(unsigned long long)(k + 1023) << 52};
  return b.d;
}
__host__ __device__ inline double lo_decode(float hi, short q) {
  return (double)q * lo_pow2(lo_exponent(hi) - 166);
}
__host__ __device__ inline short lo_encode(float hi, double v) {
  double r = rint((v - (double)hi) * lo_pow2(166 - lo_exponent(hi)));
  r = r > 32767.0 ? 32767.0 : (r < -32767.0 ? -32767.0 : r);
  return (short)r;
}
