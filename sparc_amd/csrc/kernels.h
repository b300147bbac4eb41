// Device kernels of the SPARC stepping engine (gfx950, wave64).
//
// Layout: every 3-D field is Fortran column-major (nx,ny,nz), x fastest, one
// array per field (structure of arrays).  Background fields are either full
// 3-D arrays or compact z-profiles (bg3d flag), so a horizontally uniform
// background costs nz loads instead of nx*ny*nz.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sparc_gpu.h"
#include "fft_lds.h"

namespace sparc {

// EPI_RHS3N: EPI_RHS3D without RHS_xi = temp v on the interior planes, which
// the damping forward x pass (k_xw256, XArgs::cp) writes instead
enum {
  EPI_STORE = 0, EPI_P3D = 1, EPI_RHS3D = 2, EPI_P2D = 3, EPI_RHS2D = 4, EPI_C2 = 5, EPI_V2 = 6,
  EPI_RHS3N = 7
};
// XM_DERIV_DAMPINV: the momentum RHS from the x derivative with the damping
// inverse of the same component subtracted, in one pass (k_xw256 only)
enum { XM_DERIV = 0, XM_FILTER = 1, XM_DAMP_FWD = 2, XM_DAMP_INV = 3, XM_DERIV_DAMPINV = 4 };
enum { YM_DERIV = 0, YM_FILTER = 1, YM_DAMP = 2 };

template <typename T>
struct Fld {
  T* a[SPARC_MAXFIELDS];
  T* scr[SPARC_MAXFIELDS];
  T* pmlvars[6];
  T* scrpml[6];
  T* M[12];  // option 10 work arrays
  const T* bg[SPARC_BG_COUNT];
  int bg3d[SPARC_BG_COUNT];
  const T* stretch;
  const T* g;
  T* W[8];
  const T* forcing;
  int nx, ny, nz, nxy;
  int npmlbot, npmltop;
  T beta;
  int src_plane;  // 0-based plane receiving the source this step, -1 if none
  T delta_width;
  const T* srcdep;  // option 2: source_dep(k) over planes |k - src_plane| <= 4
  const int32_t* src_planes;  // MHD: per-pixel 0-based source planes (null: src_plane)
  const T* src_dw;            // MHD: delta_width_2d per pixel
};


// Coefficients of DBYD2's right-hand side (dbyd2.f90:88-196), cast to T.
template <typename T>
struct TriCoef {
  T ci1, ci2;              // interior 6th order: d*7/9, d/36
  T c1, c2, c3, c4, c5;    // 5th order rows 2 and N2-1
  T b1, b2, b3, b4;        // 4th order rows 1 and N2
  T alpha2, beta1, cst;    // boundary preconditioning
  // Rows kc0 <= k < kc1 where the LU factor tables hold one value in T
  // (the factors converge geometrically away from the boundary rows):
  // there lowr = lw, 1/betad = ib, gam = gm (bitwise the table values).
  int kc0, kc1;
  T lw, ib, gm;
};

template <typename T>
__device__ __forceinline__ T bgv(const Fld<T>& F, int id, int c, int k) {
  return F.bg3d[id] ? F.bg[id][c + (long)F.nxy * k] : F.bg[id][k];
}

// The kernel-mode source term of the MHD displacement options at plane k of
// column c (displacement.f90:813-836, physics2d.f90:902-926): at the uniform
// plane, or at the pixel's erad_2d/orad_2d with its delta_width_2d.
template <typename T>
__device__ __forceinline__ T mhd_source(const Fld<T>& F, int c, int k) {
  if (F.src_plane < 0) return (T)0;
  if (F.src_planes) {
    if (k != F.src_planes[c]) return (T)0;
    return F.forcing[c] * F.src_dw[c] / bgv(F, SPARC_BG_RHO0, c, k);
  }
  if (k != F.src_plane) return (T)0;
  return F.forcing[c] * F.delta_width / bgv(F, SPARC_BG_RHO0, c, k);
}

__device__ __forceinline__ int pml_index(int k, int nz, int nbot, int ntop) {
  if (k < nbot) return k;
  if (k >= nz - ntop) return nbot + (k - (nz - ntop));
  return -1;
}

#include "fft_reg.h"
#include "fft_r16.h"

// ---------------------------------------------------------------------------
// x-lines: batched FFT along the contiguous index.  Real rows are packed two
// per complex line.  DERIV: out = IFFT(i k * FFT(a + beta*s)) (ddx,
// derivatives.f90:28-58); FILTER: in-place Fermi low-pass (physics.f90:796-802);
// DAMP_FWD/INV: first/last pass of the 2-D damping transform of z-plane pairs
// (damping.f90:135-160).
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_xlines(int n, int nrows, int nlines, int LPB,
                                                const T* __restrict__ a, const T* __restrict__ s,
                                                T beta, T* out, const T* __restrict__ mult,
                                                FFTPlan P, const C2<T>* __restrict__ tw,
                                                C2<T>* spec, int ny, int nz, long nxy,
                                                int edges) {
  extern __shared__ unsigned char smem_raw[];
  C2<T>* b0 = reinterpret_cast<C2<T>*>(smem_raw);
  C2<T>* b1 = b0 + LPB * n;
  const int L0 = blockIdx.x * LPB;
  const int nl = min(LPB, nlines - L0);
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    const int l = e / n, x = e - l * n;
    const long L = L0 + l;
    C2<T> v;
    if (MODE == XM_DERIV || MODE == XM_FILTER) {
      const long r0 = 2 * L, r1 = r0 + 1;
      T re = a[r0 * n + x];
      if (s) re = re + beta * s[r0 * n + x];
      T im = 0;
      if (r1 < nrows) {
        im = a[r1 * n + x];
        if (s) im = im + beta * s[r1 * n + x];
      }
      v = {re, im};
    } else if (MODE == XM_DAMP_FWD) {
      const long q = L / ny, j = L - q * ny;
      const long k0 = 2 * q, k1 = k0 + 1;
      const long o0 = x + (long)n * j + nxy * k0;
      T re = a[o0];
      if (s) re = re + beta * s[o0];
      T im = 0;
      if (k1 < nz) {
        im = a[o0 + nxy];
        if (s) im = im + beta * s[o0 + nxy];
      }
      v = {re, im};
    } else {
      v = spec[L * n + x];
    }
    b0[e] = v;
  }
  __syncthreads();
  C2<T>* r;
  if (MODE == XM_DAMP_INV)
    r = lds_fft<T, +1>(b0, b1, nl, P, tw);
  else
    r = lds_fft<T, -1>(b0, b1, nl, P, tw);
  if (MODE == XM_DAMP_FWD) {
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) spec[(long)L0 * n + e] = r[e];
    return;
  }
  if (MODE == XM_DERIV || MODE == XM_FILTER) {
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      const int x = e % n;
      const T c = mult[x];
      C2<T> v = r[e];
      if (MODE == XM_DERIV)
        v = {-c * v.y, c * v.x};
      else
        v = {c * v.x, c * v.y};
      r[e] = v;
    }
    __syncthreads();
    C2<T>* other = (r == b0) ? b1 : b0;
    r = lds_fft<T, +1>(r, other, nl, P, tw);
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      const int l = e / n, x = e - l * n;
      const long L = L0 + l;
      const long r0 = 2 * L, r1 = r0 + 1;
      out[r0 * n + x] = r[e].x;
      if (r1 < nrows) out[r1 * n + x] = r[e].y;
    }
  } else {  // XM_DAMP_INV: RHSv -= term; planes 1 and nz only with edges (options
            // 2 and 8 zero them before the damping, pml.f90:56-74,193-206)
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      const int l = e / n, x = e - l * n;
      const long L = L0 + l;
      const long q = L / ny, j = L - q * ny;
      const long k0 = 2 * q, k1 = k0 + 1;
      const long o0 = x + (long)n * j + nxy * k0;
      if (edges || (k0 != 0 && k0 != nz - 1)) out[o0] = out[o0] - r[e].x;
      if (k1 < nz - 1 || (edges && k1 < nz)) out[o0 + nxy] = out[o0 + nxy] - r[e].y;
    }
  }
}

// ---------------------------------------------------------------------------
// y-lines: FFT along the stride-nx index on a tile of TX columns of one plane
// (ddy, derivatives.f90:62-170, without the MPI transpose).  YM_DAMP works on
// the x-spectrum of a plane pair: forward y-FFT, times damping_rates, inverse.
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_ylines(int nx, int ny, int nz, int TX,
                                                const T* __restrict__ a, const T* __restrict__ s,
                                                T beta, T* out, const T* __restrict__ mult,
                                                FFTPlan P, const C2<T>* __restrict__ tw,
                                                C2<T>* spec, const T* __restrict__ rates) {
  extern __shared__ unsigned char smem_raw[];
  C2<T>* b0 = reinterpret_cast<C2<T>*>(smem_raw);
  const long nxy = (long)nx * ny;
  const int i0 = blockIdx.x * TX;
  const long k = blockIdx.y;
  if (MODE == YM_DERIV || MODE == YM_FILTER) {
    const int nl = TX / 2;
    C2<T>* b1 = b0 + nl * ny;
    T* b0r = reinterpret_cast<T*>(b0);
    for (int e = threadIdx.x; e < TX * ny; e += blockDim.x) {
      const int y = e / TX, xx = e - y * TX;
      const long o = (i0 + xx) + (long)nx * y + nxy * k;
      T v = a[o];
      if (s) v = v + beta * s[o];
      b0r[2 * ((xx >> 1) * ny + y) + (xx & 1)] = v;
    }
    __syncthreads();
    C2<T>* r = lds_fft<T, -1>(b0, b1, nl, P, tw);
    for (int e = threadIdx.x; e < nl * ny; e += blockDim.x) {
      const int y = e % ny;
      const T c = mult[y];
      C2<T> v = r[e];
      if (MODE == YM_DERIV)
        v = {-c * v.y, c * v.x};
      else
        v = {c * v.x, c * v.y};
      r[e] = v;
    }
    __syncthreads();
    C2<T>* other = (r == b0) ? b1 : b0;
    r = lds_fft<T, +1>(r, other, nl, P, tw);
    const T* rr = reinterpret_cast<const T*>(r);
    for (int e = threadIdx.x; e < TX * ny; e += blockDim.x) {
      const int y = e / TX, xx = e - y * TX;
      const long o = (i0 + xx) + (long)nx * y + nxy * k;
      out[o] = rr[2 * ((xx >> 1) * ny + y) + (xx & 1)];
    }
  } else {  // YM_DAMP on spec[(q*ny + y)*nx + m], q = blockIdx.y
    const int nl = TX;
    C2<T>* b1 = b0 + nl * ny;
    const long q = k;
    for (int e = threadIdx.x; e < nl * ny; e += blockDim.x) {
      const int y = e / nl, mm = e - y * nl;
      b0[mm * ny + y] = spec[(q * ny + y) * nx + i0 + mm];
    }
    __syncthreads();
    C2<T>* r = lds_fft<T, -1>(b0, b1, nl, P, tw);
    for (int e = threadIdx.x; e < nl * ny; e += blockDim.x) {
      const int mm = e / ny, y = e - mm * ny;
      const int m = i0 + mm;
      const int mi = (m <= nx / 2) ? m : nx - m;
      const T c = rates[y + (long)ny * mi];
      r[e] = {c * r[e].x, c * r[e].y};
    }
    __syncthreads();
    C2<T>* other = (r == b0) ? b1 : b0;
    r = lds_fft<T, +1>(r, other, nl, P, tw);
    for (int e = threadIdx.x; e < nl * ny; e += blockDim.x) {
      const int y = e / nl, mm = e - y * nl;
      spec[(q * ny + y) * nx + i0 + mm] = r[mm * ny + y];
    }
  }
}

#include "zsolve.h"

// ---------------------------------------------------------------------------
// a += dt*scr fused with the 11-point z filter (step.f90:114-118,
// physics.f90:760-788, physics2d.f90:527-560), or the filter alone
// (FILTER_VAR_Z, kernels.f90:845-881) when scr == nullptr.  One thread per
// column with an 11-value register window; output plane k is written after
// input plane k+5 is read, so the update is in place.
// ---------------------------------------------------------------------------
// Fields of one update launch, selected by blockIdx.y (scr == nullptr: the
// filter alone).
template <typename T>
struct UpdArgs {
  T* a[SPARC_MAXFIELDS];
  short* lo[SPARC_MAXFIELDS];  // low words of the compensated FP32 state
  const T* scr[SPARC_MAXFIELDS];
};

template <typename T>
__global__ __launch_bounds__(64) void k_update_filter(UpdArgs<T> A, T dt, int nxy, int nz,
                                                      int kst) {
  constexpr int U = 8;  // planes loaded one chunk ahead of use
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nxy) return;
  T* a = A.a[blockIdx.y];
  const T* __restrict__ scr = A.scr[blockIdx.y];
  const T c0 = (T)0.75390625, h1 = (T)(0.5 * 0.41015625), h2 = (T)(0.5 * -0.234375),
          h3 = (T)(0.5 * 0.087890625), h4 = (T)(0.5 * -0.01953125),
          h5 = (T)(0.5 * 0.001953125);
  auto L = [&](int k) -> T {
    if (k >= nz) return (T)0;
    const long q = c + (long)nxy * k;
    T v = a[q];
    if (scr) v = v + dt * scr[q];
    return v;
  };
  T w[11];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = 0;
#pragma unroll
  for (int i = 5; i < 11; ++i) w[i] = L(i - 5);
  T na[U], ns[U];
  auto load_chunk = [&](int kb) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int kk = kb + j;
      if (kk < nz) {
        const long q = c + (long)nxy * kk;
        na[j] = a[q];
        if (scr) ns[j] = scr[q];
      }
    }
  };
  load_chunk(6);
  for (int k0 = 0; k0 < nz; k0 += U) {
    T ca[U], cs[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      ca[j] = na[j];
      cs[j] = ns[j];
    }
    if (k0 + U + 6 < nz) load_chunk(k0 + U + 6);
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int k = k0 + j;
      if (k < nz) {
        T o;
        if (k == 0 || k == nz - 1)
          o = w[5];
        else if (k <= 4)
          o = (kst == 2) ? (T)0.5 * (w[6] + w[4]) : w[5];
        else if (k >= nz - 5)
          o = (T)0.5 * (w[6] + w[4]);
        else
          o = c0 * w[5] + h1 * (w[4] + w[6]) + h2 * (w[3] + w[7]) + h3 * (w[2] + w[8]) +
              h4 * (w[1] + w[9]) + h5 * (w[0] + w[10]);
        a[c + (long)nxy * k] = o;
#pragma unroll
        for (int i = 0; i < 10; ++i) w[i] = w[i + 1];
        w[10] = (k + 6 < nz) ? (scr ? ca[j] + dt * cs[j] : ca[j]) : (T)0;
      }
    }
  }
}

#include "lowword.h"

// Compensated form for an FP32 state held as hi + lo: the update and the
// filter run in FP64 on hi + lo + dt*scr and the result is split back into
// hi = fl32(v) and the 16-bit low word of v - hi, so the state carries 40
// significant bits and the per-step rounding of the sum does not accumulate
// over a solve.
__global__ __launch_bounds__(64) void k_update_filter_c(UpdArgs<float> A, double dt, int nxy,
                                                        int nz, int kst) {
  constexpr int U = 8;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nxy) return;
  float* a = A.a[blockIdx.y];
  short* lo = A.lo[blockIdx.y];
  const float* __restrict__ scr = A.scr[blockIdx.y];
  const double c0 = 0.75390625, h1 = 0.5 * 0.41015625, h2 = 0.5 * -0.234375,
               h3 = 0.5 * 0.087890625, h4 = 0.5 * -0.01953125, h5 = 0.5 * 0.001953125;
  auto V = [&](float x, short l, float s) -> double {
    return (double)x + lo_decode(x, l) + dt * (double)s;
  };
  double w[11];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = 0.0;
#pragma unroll
  for (int i = 5; i < 11; ++i) {
    const int k = i - 5;
    const long q = c + (long)nxy * k;
    w[i] = (k < nz) ? V(a[q], lo[q], scr[q]) : 0.0;
  }
  float na[U], ns[U];
  short nl[U];
  auto load_chunk = [&](int kb) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int kk = kb + j;
      if (kk < nz) {
        const long q = c + (long)nxy * kk;
        na[j] = a[q];
        nl[j] = lo[q];
        ns[j] = scr[q];
      }
    }
  };
  load_chunk(6);
  for (int k0 = 0; k0 < nz; k0 += U) {
    float ca[U], cs[U];
    short cl[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      ca[j] = na[j];
      cl[j] = nl[j];
      cs[j] = ns[j];
    }
    if (k0 + U + 6 < nz) load_chunk(k0 + U + 6);
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int k = k0 + j;
      if (k < nz) {
        double o;
        if (k == 0 || k == nz - 1)
          o = w[5];
        else if (k <= 4)
          o = (kst == 2) ? 0.5 * (w[6] + w[4]) : w[5];
        else if (k >= nz - 5)
          o = 0.5 * (w[6] + w[4]);
        else
          o = c0 * w[5] + h1 * (w[4] + w[6]) + h2 * (w[3] + w[7]) + h3 * (w[2] + w[8]) +
              h4 * (w[1] + w[9]) + h5 * (w[0] + w[10]);
        const long q = c + (long)nxy * k;
        const float hi = (float)o;
        a[q] = hi;
        lo[q] = lo_encode(hi, o);
#pragma unroll
        for (int i = 0; i < 10; ++i) w[i] = w[i + 1];
        w[10] = (k + 6 < nz) ? V(ca[j], cl[j], cs[j]) : 0.0;
      }
    }
  }
}

template <typename T>
__global__ void k_axpy(T* y, const T* __restrict__ x, T alpha, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + alpha * x[i];
}

template <typename T>
__global__ void k_forcing_series(T* forcing, const T* __restrict__ mask, T sval, int nxy) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nxy) forcing[c] = mask[c] * sval;
}

struct LagArgs {
  int nq;
  int idx[7];   // 0-based sample planes of vr
  double w[7];  // Lagrange weight of each sample at this time
};

template <typename T>
__global__ void k_forcing_lagrange(T* forcing, const T* __restrict__ vr, int nxy, LagArgs la,
                                   T scale) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nxy) return;
  double f = 0.0;
  for (int q = 0; q < la.nq; ++q) f += la.w[q] * (double)vr[c + (long)nxy * la.idx[q]];
  forcing[c] = (T)(f * (double)scale);
}

// sum of up to three fields: out = x + y (+ z)
template <typename T>
__global__ void k_sum3(T* out, const T* __restrict__ x, const T* __restrict__ y,
                       const T* __restrict__ z, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  T v = x[i];
  if (y) v = v + y[i];
  if (z) v = v + z[i];
  out[i] = v;
}

// out(:,:,k) = in(:,:,k) * prof(k)
template <typename T>
__global__ void k_scale_z(T* out, const T* __restrict__ in, const double* __restrict__ prof,
                          long nxy, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = in[i] * (T)prof[i / nxy];
}

// One snapshot pair of PRODUCE_KERNELS (kernels.f90:276-313) in FP64.
template <typename T>
__global__ void k_kern_accum(double* __restrict__ hess, double* __restrict__ kc2,
                             double* __restrict__ krho, const T* __restrict__ fxx,
                             const T* __restrict__ fxy, const T* __restrict__ fxz,
                             const T* __restrict__ fax, const T* __restrict__ fay,
                             const T* __restrict__ faz, const T* __restrict__ avx,
                             const T* __restrict__ avy, const T* __restrict__ avz,
                             const T* __restrict__ divf, const T* __restrict__ diva,
                             const T* __restrict__ d31, const T* __restrict__ d32,
                             const T* __restrict__ d33, const T* __restrict__ c2, int c2_3d,
                             const double* __restrict__ grav, double sff, double sfa,
                             long nxy, long n) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long k = q / nxy;
  const double Fax = fax[q] * sff, Fay = fay[q] * sff, Faz = faz[q] * sff;
  const double Fxx = fxx[q] * sff, Fxy = fxy[q] * sff, Fxz = fxz[q] * sff;
  const double Avx = avx[q] * sfa, Avy = avy[q] * sfa, Avz = avz[q] * sfa;
  const double Divf = divf[q] * sff, Diva = diva[q] * sfa;
  const double C2v = c2_3d ? (double)c2[q] : (double)c2[k];
  hess[q] = hess[q] + Faz * Faz + Fax * Fax + Fay * Fay;
  const double t1 = Divf * Diva;
  kc2[q] = kc2[q] + t1;
  double kr = krho[q] + t1 * C2v + Faz * Avz + Fax * Avx + Fay * Avy;
  const double D31 = d31[q] * sfa, D32 = d32 ? d32[q] * sfa : 0.0, D33 = d33[q] * sfa;
  kr = kr - grav[k] * Fxz * Diva + (Fxx * D31 + Fxy * D32 + Fxz * D33);
  krho[q] = kr;
}

// ---------------------------------------------------------------------------
// Option 10 (MP_MHD_PML_3D_DISPL, displacement.f90:663-849): pointwise parts
// between the derivative passes.  Stage temp = a + beta*scr is formed on the
// fly; M[] are the context's work arrays.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T tmp_of(const Fld<T>& F, int f, long q) {
  T v = F.a[f][q];
  if (F.beta != (T)0) v = v + F.beta * F.scr[f][q];
  return v;
}

// div xi -> p, rho with the horizontal background gradients (:694-700):
// W0 + W1 + W2 = div, out p -> W3, rho -> W4; and, from the same stage temp
// xi, flux = xi x B0 -> M0..M2 (:673-675: flux1 = -(xi_z boy - xi_y boz),
// ...), which the reference forms before the divergence: nothing between
// the two reads or writes M0..M2, and xi is read once
template <typename T>
__global__ void k_mhd_pr(Fld<T> F, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  const T xx = tmp_of(F, SPARC_XI_X, q), xy = tmp_of(F, SPARC_XI_Y, q), xz = tmp_of(F, SPARC_XI_Z, q);
  const T bx = bgv(F, SPARC_BG_BX, c, k), by = bgv(F, SPARC_BG_BY, c, k), bz = bgv(F, SPARC_BG_BZ, c, k);
  F.M[0][q] = -(xz * by - xy * bz);
  F.M[1][q] = -(xx * bz - xz * bx);
  F.M[2][q] = -(xy * bx - xx * by);
  const T div = F.W[0][q] + F.W[1][q] + F.W[2][q];
  F.W[3][q] = -bgv(F, SPARC_BG_C2RHO0, c, k) * div - xz * bgv(F, SPARC_BG_GRADP0_Z, c, k) -
              xx * bgv(F, SPARC_BG_GRADP0_X, c, k) - xy * bgv(F, SPARC_BG_GRADP0_Y, c, k);
  F.W[4][q] = -bgv(F, SPARC_BG_RHO0, c, k) * div - xz * bgv(F, SPARC_BG_GRADRHO0_Z, c, k) -
              xx * bgv(F, SPARC_BG_GRADRHO0_X, c, k) - xy * bgv(F, SPARC_BG_GRADRHO0_Y, c, k);
}

// PML of option 10, bottom layers only (:684-692 and the like): psi field w
// on derivative d over planes 0..npmlbot-1
template <typename T>
__global__ void k_pml_bottom(Fld<T> F, T* __restrict__ d, int w, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  T psi = F.pmlvars[w][q];
  if (F.beta != (T)0) psi = psi + F.beta * F.scrpml[w][q];
  const T dv = d[q];
  F.scrpml[w][q] = bgv(F, SPARC_BG_AZ, c, k) * dv + bgv(F, SPARC_BG_BZPML, c, k) * psi;
  d[q] = dv + psi;
}

// b = curl(flux) from its pieces (:744-746): M0 = dy f3 - dz f2,
// M1 = -dx f3 + dz f1, M2 = dx f2 - dy f1 with the pieces in M6..M11
template <typename T>
__global__ void k_mhd_b(Fld<T> F, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  F.M[0][q] = F.M[7][q] - F.M[8][q];
  F.M[1][q] = -F.M[6][q] + F.M[11][q];
  F.M[2][q] = F.M[9][q] - F.M[10][q];
}

// J = curl(b), the Lorentz force J x B0 + J0 x b, grad p, sponge, gravity,
// RHSxi = v, source, zeroed planes 1 and nz (:776-846).  gradp in M3..M5,
// b in M0..M2, curl pieces dx by, dy bz, dz bx, dx bz, dy bx, dz by in M6..M11,
// rho in W4.  The damping of v is subtracted afterwards.  vin: the damping
// forward pass has already stored the stage temp v as RHS_xi (zero on
// planes 1 and nz), so v is read from there and RHS_xi is not written.
template <typename T>
__global__ void k_mhd_rhs(Fld<T> F, const T* __restrict__ sponge, long n, int vin) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int nz = F.nz;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  T vx, vy, vz;
  if (vin) {
    vx = F.scr[SPARC_XI_X][q];
    vy = F.scr[SPARC_XI_Y][q];
    vz = F.scr[SPARC_XI_Z][q];
  } else {
    vx = tmp_of(F, SPARC_V_X, q);
    vy = tmp_of(F, SPARC_V_Y, q);
    vz = tmp_of(F, SPARC_V_Z, q);
  }
  if (k == 0 || k == nz - 1) {
#pragma unroll
    for (int f = 0; f < SPARC_NFIELDS; ++f) F.scr[f][q] = 0;
    return;
  }
  const T jx = F.M[7][q] - F.M[11][q];
  const T jy = -F.M[9][q] + F.M[8][q];
  const T jz = F.M[6][q] - F.M[10][q];
  const T bx = F.M[0][q], by = F.M[1][q], bz = F.M[2][q];
  const T b0x = bgv(F, SPARC_BG_BX, c, k), b0y = bgv(F, SPARC_BG_BY, c, k),
          b0z = bgv(F, SPARC_BG_BZ, c, k);
  const T j0x = bgv(F, SPARC_BG_CURLBX, c, k), j0y = bgv(F, SPARC_BG_CURLBY, c, k),
          j0z = bgv(F, SPARC_BG_CURLBZ, c, k);
  const T ri = bgv(F, SPARC_BG_RHOINV, c, k);
  T rx = ri * (-F.M[3][q] + (jy * b0z - jz * b0y)) + ri * (j0y * bz - j0z * by);
  T ry = ri * (-F.M[4][q] + (b0x * jz - jx * b0z)) + ri * (bx * j0z - j0x * bz);
  T rz = ri * (-F.M[5][q] + (jx * b0y - jy * b0x)) + ri * (j0x * by - j0y * bx);
  if (k + 1 >= nz - 35) {
    const T sp = sponge[k];
    rx = rx - sp * vx;
    ry = ry - sp * vy;
    rz = rz - sp * vz;
  }
  rz = rz - ri * F.W[4][q] * F.g[k];
  rz = rz + mhd_source(F, c, k);
  if (!vin) {
    F.scr[SPARC_XI_X][q] = vx;
    F.scr[SPARC_XI_Y][q] = vy;
    F.scr[SPARC_XI_Z][q] = vz;
  }
  F.scr[SPARC_V_X][q] = rx;
  F.scr[SPARC_V_Y][q] = ry;
  F.scr[SPARC_V_Z][q] = rz;
}

// Option 8 (MP_MHD_PML_3D, pml.f90:80-208), slots rho 0, v 1..3, p 4, b 5..7.
// flux = v x B0 (pml.f90:89-91) -> M0..M2
template <typename T>
__global__ void k_mhd8_flux(Fld<T> F, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  const T vx = tmp_of(F, 1, q), vy = tmp_of(F, 2, q), vz = tmp_of(F, 3, q);
  const T bx = bgv(F, SPARC_BG_BX, c, k), by = bgv(F, SPARC_BG_BY, c, k), bz = bgv(F, SPARC_BG_BZ, c, k);
  F.M[0][q] = -(vz * by - vy * bz);
  F.M[1][q] = -(vx * bz - vz * bx);
  F.M[2][q] = -(vy * bx - vx * by);
}

// Continuity, momentum (J x B0 + J0 x b), pressure and induction RHS, source
// over 9 planes, planes 1 and nz zeroed (pml.f90:119-190).  Derivatives:
// div parts W0..W2, grad p (W4, W3, W5), d_x f3 M3, d_z f1 M4, d_y f3 M5,
// d_z f2 M6, d_x f2 M7, d_y bz M8, d_z bx M9, d_x by M10, d_y bx M11,
// d_x bz W6, d_y f1 M1, d_z by M2.  The damping of v is subtracted afterwards.
template <typename T>
__global__ void k_mhd8_rhs(Fld<T> F, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int nz = F.nz;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  const T rho = tmp_of(F, 0, q), vx = tmp_of(F, 1, q), vy = tmp_of(F, 2, q), vz = tmp_of(F, 3, q);
  const T bx = tmp_of(F, 5, q), by = tmp_of(F, 6, q), bz = tmp_of(F, 7, q);
  if (k == 0 || k == nz - 1) {
#pragma unroll
    for (int f = 0; f < 8; ++f) F.scr[f][q] = 0;
    return;
  }
  const T rbx = F.M[5][q] - F.M[6][q];
  const T rby = -F.M[3][q] + F.M[4][q];
  const T rbz = F.M[7][q] - F.M[1][q];
  const T div = F.W[0][q] + F.W[1][q] + F.W[2][q];
  const T jx = F.M[8][q] - F.M[2][q];
  const T jy = -F.W[6][q] + F.M[9][q];
  const T jz = F.M[10][q] - F.M[11][q];
  const T rc = -bgv(F, SPARC_BG_GRADRHO0_X, c, k) * vx - bgv(F, SPARC_BG_GRADRHO0_Y, c, k) * vy -
               bgv(F, SPARC_BG_GRADRHO0_Z, c, k) * vz - bgv(F, SPARC_BG_RHO0, c, k) * div;
  const T rp = -bgv(F, SPARC_BG_C2RHO0, c, k) * div - vx * bgv(F, SPARC_BG_GRADP0_X, c, k) -
               vy * bgv(F, SPARC_BG_GRADP0_Y, c, k) - vz * bgv(F, SPARC_BG_GRADP0_Z, c, k);
  const T b0x = bgv(F, SPARC_BG_BX, c, k), b0y = bgv(F, SPARC_BG_BY, c, k),
          b0z = bgv(F, SPARC_BG_BZ, c, k);
  const T j0x = bgv(F, SPARC_BG_CURLBX, c, k), j0y = bgv(F, SPARC_BG_CURLBY, c, k),
          j0z = bgv(F, SPARC_BG_CURLBZ, c, k);
  const T ri = bgv(F, SPARC_BG_RHOINV, c, k);
  T rx = -ri * (F.W[4][q] - (jy * b0z - jz * b0y));
  T ry = -ri * (F.W[3][q] - (b0x * jz - jx * b0z));
  T rz = -ri * (F.W[5][q] - (jx * b0y - jy * b0x));
  rx = rx + ri * (j0y * bz - j0z * by);
  ry = ry + ri * (bx * j0z - j0x * bz);
  rz = rz + ri * (j0x * by - j0y * bx);
  rz = rz - ri * rho * F.g[k];
  if (F.src_plane >= 0 && k - F.src_plane <= 4 && F.src_plane - k <= 4)
    rz = rz + F.forcing[c] * F.srcdep[k];
  F.scr[0][q] = rc;
  F.scr[1][q] = rx;
  F.scr[2][q] = ry;
  F.scr[3][q] = rz;
  F.scr[4][q] = rp;
  F.scr[5][q] = rbx;
  F.scr[6][q] = rby;
  F.scr[7][q] = rbz;
}

// Option 18 (MP_MHD_PML_2D_DISPL, physics2d.f90:724-943), ny = 1.
// PML on every PML plane (top and bottom): psi field w on derivative d.
template <typename T>
__global__ void k_pml_planes(Fld<T> F, T* __restrict__ d, int w, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  const int pi = pml_index(k, F.nz, F.npmlbot, F.npmltop);
  if (pi < 0) return;
  const long pq = c + (long)F.nxy * pi;
  T psi = F.pmlvars[w][pq];
  if (F.beta != (T)0) psi = psi + F.beta * F.scrpml[w][pq];
  const T dv = d[q];
  F.scrpml[w][pq] = bgv(F, SPARC_BG_AZ, c, k) * dv + bgv(F, SPARC_BG_BZPML, c, k) * psi;
  d[q] = dv + psi;
}

// div xi = W0 + W2 -> p (W3), rho (W4); flux2 = -(xi_x boz - xi_z box) -> M0
// (physics2d.f90:767-769,815)
template <typename T>
__global__ void k_mhd2_pr(Fld<T> F, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  const T xx = tmp_of(F, SPARC_XI_X, q), xz = tmp_of(F, SPARC_XI_Z, q);
  const T div = F.W[0][q] + F.W[2][q];
  F.W[3][q] = -bgv(F, SPARC_BG_C2RHO0, c, k) * div - xz * bgv(F, SPARC_BG_GRADP0_Z, c, k) -
              xx * bgv(F, SPARC_BG_GRADP0_X, c, k);
  F.W[4][q] = -bgv(F, SPARC_BG_RHO0, c, k) * div - xz * bgv(F, SPARC_BG_GRADRHO0_Z, c, k) -
              xx * bgv(F, SPARC_BG_GRADRHO0_X, c, k);
  F.M[0][q] = -(xx * bgv(F, SPARC_BG_BZ, c, k) - xz * bgv(F, SPARC_BG_BX, c, k));
}

template <typename T>
__global__ void k_negate(T* __restrict__ d, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) d[q] = -d[q];
}

// curlby = -ddx(bz) + dzbx; RHSv = rhoinv (-grad p + curlby x B0 + curlB0 x b),
// gravity, RHSxi = v, source; plane 1 zeroed (physics2d.f90:876-933).  gradp_x
// M6, gradp_z M7, bx M1, bz M2, dzbx M3, ddx(bz) M4, rho W4.
template <typename T>
__global__ void k_mhd2_rhs(Fld<T> F, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  const T vx = tmp_of(F, SPARC_V_X, q), vz = tmp_of(F, SPARC_V_Z, q);
  if (k == 0) {
#pragma unroll
    for (int f = 0; f < SPARC_NFIELDS; ++f) F.scr[f][q] = 0;
    return;
  }
  const T ri = bgv(F, SPARC_BG_RHOINV, c, k);
  const T cby = -F.M[4][q] + F.M[3][q];
  const T j0y = bgv(F, SPARC_BG_CURLBY, c, k);
  T rx = ri * (-F.M[6][q] + cby * bgv(F, SPARC_BG_BZ, c, k));
  T rz = ri * (-F.M[7][q] + (-cby * bgv(F, SPARC_BG_BX, c, k)));
  rx = rx + ri * (j0y * F.M[2][q]);
  rz = rz + ri * (-j0y * F.M[1][q]);
  rz = rz - ri * F.W[4][q] * F.g[k];
  rz = rz + mhd_source(F, c, k);
  F.scr[SPARC_XI_X][q] = vx;
  F.scr[SPARC_XI_Z][q] = vz;
  F.scr[SPARC_V_X][q] = rx;
  F.scr[SPARC_V_Z][q] = rz;
}

// out[q] -= term[q] for planes 2..nz (damping applied before plane 1 is zeroed)
template <typename T>
__global__ void k_sub_above_plane1(T* __restrict__ out, const T* __restrict__ term, long nxy,
                                   long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nxy && q < n) out[q] = out[q] - term[q];
}

// ---------------------------------------------------------------------------
// Magnetic kernels (kernels.f90:336-430): vector operands are three arrays,
// or a background triple (B0 or curl B0) read through bgv.
// ---------------------------------------------------------------------------
template <typename T>
struct Vec3 {
  const T* p[3];  // arrays, or
  int bg;         // first background id of a triple (SPARC_BG_BX, SPARC_BG_CURLBX), -1 if arrays
  const T* s;     // optional scalar field multiplying the vector (nullptr: none)
};

template <typename T>
__device__ __forceinline__ void vec_at(const Fld<T>& F, const Vec3<T>& v, long q, int c, int k,
                                       T* o) {
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = v.bg >= 0 ? bgv(F, v.bg + i, c, k) : v.p[i][q];
  if (v.s) {
    const T m = v.s[q];
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] = o[i] * m;
  }
}

// c = a x b (CROSS_KERN, kernels.f90:757-766) into T arrays, or with ACC
// added to the FP64 sums times `scale` (dst = the kernel sums).
template <typename T, bool ACC>
__global__ void k_cross3(Fld<T> F, Vec3<T> a, Vec3<T> b, T* cx, T* cy, T* cz, double* kx,
                         double* ky, double* kz, double scale, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  T u[3], v[3];
  vec_at(F, a, q, c, k, u);
  vec_at(F, b, q, c, k, v);
  const T x = u[1] * v[2] - u[2] * v[1], y = v[0] * u[2] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
  if (ACC) {
    kx[q] = kx[q] + scale * (double)x;
    ky[q] = ky[q] + scale * (double)y;
    kz[q] = kz[q] + scale * (double)z;
  } else {
    cx[q] = x;
    cy[q] = y;
    cz[q] = z;
  }
}

// k += scale * v (v optionally times a scalar field)
template <typename T>
__global__ void k_acc_vec(Fld<T> F, Vec3<T> v, double* kx, double* ky, double* kz, double scale,
                          long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int k = (int)(q / F.nxy), c = (int)(q - (long)k * F.nxy);
  T u[3];
  vec_at(F, v, q, c, k, u);
  kx[q] = kx[q] + scale * (double)u[0];
  ky[q] = ky[q] + scale * (double)u[1];
  kz[q] = kz[q] + scale * (double)u[2];
}

// curl from its six derivative pieces (CURL_KERN, kernels.f90:779-789):
// c = (dy fz - dz fy, dz fx - dx fz, dx fy - dy fx); dyfx/dyfz may be null (ny = 1)
template <typename T>
__global__ void k_curl_combine(T* cx, T* cy, T* cz, const T* dxfy, const T* dyfx, const T* dyfz,
                               const T* dzfy, const T* dzfx, const T* dxfz, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const T a = dyfz ? dyfz[q] : (T)0, b = dyfx ? dyfx[q] : (T)0;
  cx[q] = a - dzfy[q];
  cy[q] = dzfx[q] - dxfz[q];
  cz[q] = dxfy[q] - b;
}

// Flow-kernel sums for one forward velocity component j (kernels.f90:317-332):
// K_v(i) += d_i v_f,j * v_a,j with d_x, d_y, d_z v_f,j given (dy == nullptr: 0).
template <typename T>
__global__ void k_kern_flow(double* __restrict__ kx, double* __restrict__ ky,
                            double* __restrict__ kz, const T* __restrict__ dx,
                            const T* __restrict__ dy, const T* __restrict__ dz,
                            const T* __restrict__ av, double sff, double sfa, long n) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double a = (double)av[q] * sfa;
  kx[q] = kx[q] + (double)dx[q] * sff * a;
  if (dy) ky[q] = ky[q] + (double)dy[q] * sff * a;
  kz[q] = kz[q] + (double)dz[q] * sff * a;
}

// -rhoinv * d with planes 1 and nz zeroed: a derivative turned into a
// momentum right-hand side (the generic-size path of XArgs::rhs)
template <typename T>
__global__ void k_rhs_grad(T* __restrict__ out, const T* __restrict__ d,
                           const T* __restrict__ rz, const T* __restrict__ r3, long nxy,
                           int nz, long n) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long k = q / nxy;
  out[q] = (k == 0 || k == nz - 1) ? (T)0 : -(r3 ? r3[q] : rz[k]) * d[q];
}

template <typename T>
__global__ void k_from_double(T* __restrict__ out, const double* __restrict__ in, long n) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) out[q] = (T)in[q];
}

// y component of CURL_KERN(fx, 0, fz) = d_z fx - d_x fz, into FP64
template <typename T>
__global__ void k_curl_y(double* __restrict__ out, const T* __restrict__ dzfx,
                         const T* __restrict__ dxfz, long n) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) out[q] = (double)dzfx[q] - (double)dxfz[q];
}

}  // namespace sparc
