// LDS-resident batched complex FFT (Stockham autosort, radices 4, 2, 3) used
// by the x- and y-line kernels.  Replaces the FFTW r2c/c2r plans of
// initialize.f90:323-337 and damping.f90:46-58: real lines are packed two per
// complex line (the derivative and filter multipliers are Hermitian-preserving,
// so real and imaginary parts stay separate), which halves the work of a
// real-to-complex pair without a separate r2c post-pass.
#pragma once
#include <hip/hip_runtime.h>

namespace sparc {

template <typename T>
struct C2 {
  T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// n = product of rad[0..nrad-1]; each radix in {2,3,4}
struct FFTPlan {
  int n;
  int nrad;
  int rad[20];
};

// DIR = -1 forward (exp(-i...)), +1 unnormalized inverse
template <typename T, int DIR, int R>
__device__ __forceinline__ void fft_stage(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst,
                                          int n, int nlines, int Ns,
                                          const C2<T>* __restrict__ tw) {
  const int nb = n / R;
  const int tstride = n / (Ns * R);
  for (int w = threadIdx.x; w < nlines * nb; w += blockDim.x) {
    const int line = w / nb;
    const int j = w - line * nb;
    const C2<T>* in = src + line * n;
    C2<T>* out = dst + line * n;
    const int k = j % Ns;
    C2<T> v0 = in[j], v1 = in[j + nb], v2, v3;
    if (R >= 3) v2 = in[j + 2 * nb];
    if (R == 4) v3 = in[j + 3 * nb];
    if (Ns > 1) {
      C2<T> t1 = tw[k * tstride];
      if (DIR > 0) t1.y = -t1.y;
      v1 = cmul(v1, t1);
      if (R >= 3) {
        C2<T> t2 = tw[2 * k * tstride];
        if (DIR > 0) t2.y = -t2.y;
        v2 = cmul(v2, t2);
      }
      if (R == 4) {
        C2<T> t3 = tw[3 * k * tstride];
        if (DIR > 0) t3.y = -t3.y;
        v3 = cmul(v3, t3);
      }
    }
    const int idxD = (j / Ns) * Ns * R + k;
    if (R == 2) {
      out[idxD] = {v0.x + v1.x, v0.y + v1.y};
      out[idxD + Ns] = {v0.x - v1.x, v0.y - v1.y};
    } else if (R == 4) {
      C2<T> t0 = {v0.x + v2.x, v0.y + v2.y};
      C2<T> t1 = {v0.x - v2.x, v0.y - v2.y};
      C2<T> t2 = {v1.x + v3.x, v1.y + v3.y};
      C2<T> d = {v1.x - v3.x, v1.y - v3.y};
      // forward: t3 = -i*d ; inverse: t3 = +i*d
      C2<T> t3 = (DIR < 0) ? C2<T>{d.y, -d.x} : C2<T>{-d.y, d.x};
      out[idxD] = {t0.x + t2.x, t0.y + t2.y};
      out[idxD + Ns] = {t1.x + t3.x, t1.y + t3.y};
      out[idxD + 2 * Ns] = {t0.x - t2.x, t0.y - t2.y};
      out[idxD + 3 * Ns] = {t1.x - t3.x, t1.y - t3.y};
    } else {  // R == 3
      const T h = (T)0.5;
      const T s3 = (T)0.86602540378443864676;  // sqrt(3)/2
      C2<T> s = {v1.x + v2.x, v1.y + v2.y};
      C2<T> d = {v1.x - v2.x, v1.y - v2.y};
      C2<T> m = {v0.x - h * s.x, v0.y - h * s.y};
      // forward: y1 = m - i*s3*d ; inverse: y1 = m + i*s3*d
      C2<T> id = (DIR < 0) ? C2<T>{s3 * d.y, -s3 * d.x} : C2<T>{-s3 * d.y, s3 * d.x};
      out[idxD] = {v0.x + s.x, v0.y + s.y};
      out[idxD + Ns] = {m.x + id.x, m.y + id.y};
      out[idxD + 2 * Ns] = {m.x - id.x, m.y - id.y};
    }
  }
}

// Runs all stages; returns the buffer holding the result (b0 or b1).
template <typename T, int DIR>
__device__ C2<T>* lds_fft(C2<T>* b0, C2<T>* b1, int nlines, const FFTPlan& P,
                          const C2<T>* __restrict__ tw) {
  int Ns = 1;
  C2<T>* src = b0;
  C2<T>* dst = b1;
  for (int s = 0; s < P.nrad; ++s) {
    const int R = P.rad[s];
    if (R == 4)
      fft_stage<T, DIR, 4>(src, dst, P.n, nlines, Ns, tw);
    else if (R == 2)
      fft_stage<T, DIR, 2>(src, dst, P.n, nlines, Ns, tw);
    else
      fft_stage<T, DIR, 3>(src, dst, P.n, nlines, Ns, tw);
    __syncthreads();
    C2<T>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

}  // namespace sparc
