// Line transforms of length 256 as a four-step FFT held in registers:
// 256 = 16 x 16, each line owned by 16 lanes with 16 values each.
//
//   forward  X[k1 + 16 k2] = sum_n2 W16^(n2 k2) W256^(n2 k1) sum_n1 W16^(n1 k1) x[16 n1 + n2]
//   inverse  x[n1 + 16 n2] = sum_k1 W16^-(k1 n2) W256^-(k1 n1) sum_k2 W16^-(k2 n1) X[k1 + 16 k2]
//
// Lane u of a line loads x[16 n1 + u] (n1 = 0..15), runs a 16-point DFT in
// registers, applies the W256 twiddles and exchanges with the other 15 lanes
// through LDS once; after the second 16-point DFT it holds X[u + 16 k2].  The
// spectral multiplier is applied there, the inverse runs the same steps back,
// and lane u ends with x[u + 16 n2].  A derivative is two LDS exchanges and
// three barriers, against seven LDS round trips for the radix-16 Stockham
// passes of fft_reg.h, which leaves the kernels bound by their loads and
// stores.  Same contracts as k_xlines_n / k_ylines_n.
// Included inside namespace sparc by kernels.h.
#pragma once

// Exchange tile of one line: element (row, col) at row*17 + col (16 x 16, the
// odd row stride keeps column reads on distinct banks).
constexpr int R16P = 17;

// W256^(u*k), k = 0..15, for lane u, from the block's table
template <typename T>
__device__ __forceinline__ C2<T> tw16(const C2<T>* tw2, int k, int u) {
  return tw2[k * 16 + u];
}

template <typename T, int DIR>
__device__ __forceinline__ void twiddle16(C2<T> (&v)[16], const C2<T>* tw2, int u) {
#pragma unroll
  for (int k = 1; k < 16; ++k) {
    C2<T> t = tw16(tw2, k, u);
    if (DIR > 0) t.y = -t.y;
    v[k] = cmul(v[k], t);
  }
}

// lane u holds row u's 16 entries after the exchange: writes (k, u), reads (u, k)
template <typename T>
__device__ __forceinline__ void exchange16(C2<T> (&v)[16], C2<T>* Sl, int u) {
#pragma unroll
  for (int k = 0; k < 16; ++k) Sl[k * R16P + u] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = Sl[u * R16P + k];
}

// block table W256^(u*k) at [k*16 + u]
template <typename T>
__device__ __forceinline__ void load_tw2(C2<T>* tw2, const C2<T>* __restrict__ twg) {
  for (int e = threadIdx.x; e < 256; e += blockDim.x) {
    const int k = e >> 4, u = e & 15;
    tw2[e] = twg[(u * k) & 255];
  }
}

// x-lines of length 256: 16 complex lines (32 real rows, or 16 row pairs of
// a plane pair) per block of 256 threads, lane u = threadIdx % 16 of line
// threadIdx / 16.  blockIdx.y selects the field.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_x256(int nrows, int nlines, XArgs<T> X, T beta,
                                              const T* __restrict__ mult,
                                              const C2<T>* __restrict__ twg, int ny, int nz,
                                              long nxy) {
  constexpr int N = 256, LPB = 16, LSX = 16 * R16P;  // LSX = 16 (mod 32)
  __shared__ C2<T> S[LPB * LSX];
  __shared__ C2<T> tw2[256];
  __shared__ T ml[256];
  const int fld = blockIdx.y;
  const T* __restrict__ a = X.a[fld];
  const T* __restrict__ s = X.s[fld];
  T* out = X.out[fld];
  C2<T>* spec = X.spec[fld];
  load_tw2(tw2, twg);
  if (MODE == XM_DERIV || MODE == XM_FILTER)
    for (int e = threadIdx.x; e < N; e += blockDim.x) ml[e] = mult[e];
  const int l = threadIdx.x >> 4, u = threadIdx.x & 15;
  const long L = (long)blockIdx.x * LPB + l;
  const bool act = L < nlines;
  C2<T>* Sl = S + l * LSX;
  C2<T> v[16];
  T ov0[MODE == XM_DAMP_INV ? 16 : 1], ov1[MODE == XM_DAMP_INV ? 16 : 1];
  // plane-pair addressing of the damping modes
  long o0 = 0;
  bool w0 = false, w1 = false;
  if (MODE == XM_DAMP_FWD || MODE == XM_DAMP_INV) {
    const long q = L / ny, j = L - q * ny;
    const long k0 = 2 * q, k1 = k0 + 1;
    o0 = (long)N * j + nxy * k0;
    if (MODE == XM_DAMP_FWD) {
      w1 = k1 < nz;
    } else {
      w0 = X.edges || (k0 != 0 && k0 != nz - 1);
      w1 = k1 < nz - 1 || (X.edges && k1 < nz);
    }
  }
  const long r0 = 2 * L, r1 = r0 + 1;
  const bool has1 = r1 < nrows;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int x = 16 * n + u;
    C2<T> z = {0, 0};
    if (act) {
      if (MODE == XM_DERIV || MODE == XM_FILTER) {
        T re = a[r0 * N + x];
        if (s) re = re + beta * s[r0 * N + x];
        T im = 0;
        if (has1) {
          im = a[r1 * N + x];
          if (s) im = im + beta * s[r1 * N + x];
        }
        z = {re, im};
      } else if (MODE == XM_DAMP_FWD) {
        T re = a[o0 + x];
        if (s) re = re + beta * s[o0 + x];
        T im = 0;
        if (w1) {
          im = a[o0 + nxy + x];
          if (s) im = im + beta * s[o0 + nxy + x];
        }
        z = {re, im};
      } else {  // XM_DAMP_INV: lane u loads X[u + 16 k2] and the RHS it updates
        z = spec[L * N + u + 16 * n];
        ov0[n] = w0 ? out[o0 + x] : (T)0;
        ov1[n] = w1 ? out[o0 + nxy + x] : (T)0;
      }
    }
    v[n] = z;
  }
  __syncthreads();  // tables
  if (MODE != XM_DAMP_INV) {
    dft_reg<T, 16, -1>(v);
    twiddle16<T, -1>(v, tw2, u);
    exchange16(v, Sl, u);
    dft_reg<T, 16, -1>(v);  // v[k2] = X[u + 16 k2]
    if (MODE == XM_DAMP_FWD) {
      if (act) {
#pragma unroll
        for (int k = 0; k < 16; ++k) spec[L * N + u + 16 * k] = v[k];
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const T c = ml[u + 16 * k];
      v[k] = (MODE == XM_DERIV) ? C2<T>{-c * v[k].y, c * v[k].x} : C2<T>{c * v[k].x, c * v[k].y};
    }
    __syncthreads();  // exchange tile reused
  }
  dft_reg<T, 16, +1>(v);
  twiddle16<T, +1>(v, tw2, u);
  exchange16(v, Sl, u);
  dft_reg<T, 16, +1>(v);  // v[n2] = x[u + 16 n2]
  if (!act) return;
  if (MODE == XM_DERIV || MODE == XM_FILTER) {
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int x = u + 16 * n;
      if (MODE == XM_DERIV && X.rhs) {  // row r = j + ny*k
        out[r0 * N + x] = rhs_of(X, r0 / ny, r0 * N + x, nz, v[n].x);
        if (has1) out[r1 * N + x] = rhs_of(X, r1 / ny, r1 * N + x, nz, v[n].y);
      } else {
        out[r0 * N + x] = v[n].x;
        if (has1) out[r1 * N + x] = v[n].y;
      }
    }
  } else {
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int x = u + 16 * n;
      if (w0) out[o0 + x] = ov0[n] - v[n].x;
      if (w1) out[o0 + nxy + x] = ov1[n] - v[n].y;
    }
  }
}

// y-lines of length 256 on one plane: 16 complex lines per block of 256
// threads, lane w = threadIdx / 16 of line c = threadIdx % 16, so the 16
// lines' lanes are adjacent in x.  DERIV/FILTER: line c packs columns
// i0 + 2c, i0 + 2c + 1 (i0 = 32 * blockIdx.x).  DAMP: line c is x-bin
// i0 + c (i0 = 16 * blockIdx.x) of a plane pair's spectrum, times the damping
// rates held as rates_o[m + (nx/2+1) * y].  blockIdx.z selects the field.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_y256(int nx, int nz, XArgs<T> X, T beta,
                                              const T* __restrict__ mult,
                                              const C2<T>* __restrict__ twg,
                                              const T* __restrict__ rates_o) {
  constexpr int NY = 256, LSY = 18 * R16P - 16;  // 290 = 2 (mod 32)
  __shared__ C2<T> S[16 * LSY];
  __shared__ C2<T> tw2[256];
  __shared__ T ml[256];
  const int fld = blockIdx.z;
  const T* __restrict__ a = X.a[fld];
  const T* __restrict__ s = X.s[fld];
  T* out = X.out[fld];
  C2<T>* spec = X.spec[fld];
  load_tw2(tw2, twg);
  if (MODE != YM_DAMP)
    for (int e = threadIdx.x; e < NY; e += blockDim.x) ml[e] = mult[e];
  const int c = threadIdx.x & 15, w = threadIdx.x >> 4;
  const long nxy = (long)nx * NY;
  const long k = blockIdx.y;
  C2<T>* Sl = S + c * LSY;
  C2<T> v[16];
  if (MODE == YM_DAMP) {
    const int i0 = blockIdx.x * 16;
#pragma unroll
    for (int n = 0; n < 16; ++n) v[n] = spec[(k * NY + 16 * n + w) * nx + i0 + c];
  } else {
    const long x0 = (long)blockIdx.x * 32 + 2 * c;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const long o = x0 + (long)nx * (16 * n + w) + nxy * k;
      C2<T> z = *reinterpret_cast<const C2<T>*>(a + o);
      if (s) {
        const C2<T> t = *reinterpret_cast<const C2<T>*>(s + o);
        z = {z.x + beta * t.x, z.y + beta * t.y};
      }
      v[n] = z;
    }
  }
  __syncthreads();  // tables
  dft_reg<T, 16, -1>(v);
  twiddle16<T, -1>(v, tw2, w);
  exchange16(v, Sl, w);
  dft_reg<T, 16, -1>(v);  // v[k2] = Y[w + 16 k2]
  if (MODE == YM_DAMP) {
    const int m = blockIdx.x * 16 + c;
    const int mi = (m <= nx / 2) ? m : nx - m;
    const long nh = nx / 2 + 1;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const T r = rates_o[mi + nh * (w + 16 * q)];
      v[q] = {r * v[q].x, r * v[q].y};
    }
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const T cm = ml[w + 16 * q];
      v[q] = (MODE == YM_DERIV) ? C2<T>{-cm * v[q].y, cm * v[q].x}
                                : C2<T>{cm * v[q].x, cm * v[q].y};
    }
  }
  __syncthreads();  // exchange tile reused
  dft_reg<T, 16, +1>(v);
  twiddle16<T, +1>(v, tw2, w);
  exchange16(v, Sl, w);
  dft_reg<T, 16, +1>(v);  // v[n2] = y-line value at w + 16 n2
  if (MODE == YM_DAMP) {
    const int i0 = blockIdx.x * 16;
#pragma unroll
    for (int n = 0; n < 16; ++n) spec[(k * NY + w + 16 * n) * nx + i0 + c] = v[n];
  } else {
    const long x0 = (long)blockIdx.x * 32 + 2 * c;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const long o = x0 + (long)nx * (w + 16 * n) + nxy * k;
      C2<T> z = v[n];
      if (MODE == YM_DERIV && X.rhs) z = {rhs_of(X, k, o, nz, z.x), rhs_of(X, k, o + 1, nz, z.y)};
      *reinterpret_cast<C2<T>*>(out + o) = z;
    }
  }
}
