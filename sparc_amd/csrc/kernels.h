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

enum { EPI_STORE = 0, EPI_P3D = 1, EPI_RHS3D = 2, EPI_P2D = 3, EPI_RHS2D = 4 };
enum { XM_DERIV = 0, XM_FILTER = 1, XM_DAMP_FWD = 2, XM_DAMP_INV = 3 };
enum { YM_DERIV = 0, YM_FILTER = 1, YM_DAMP = 2 };

template <typename T>
struct Fld {
  T* a[SPARC_NFIELDS];
  T* scr[SPARC_NFIELDS];
  T* pmlvars[2];
  T* scrpml[2];
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
};

// Coefficients of DBYD2's right-hand side (dbyd2.f90:88-196), cast to T.
template <typename T>
struct TriCoef {
  T ci1, ci2;              // interior 6th order: d*7/9, d/36
  T c1, c2, c3, c4, c5;    // 5th order rows 2 and N2-1
  T b1, b2, b3, b4;        // 4th order rows 1 and N2
  T alpha2, beta1, cst;    // boundary preconditioning
};

template <typename T>
__device__ __forceinline__ T bgv(const Fld<T>& F, int id, int c, int k) {
  return F.bg3d[id] ? F.bg[id][c + (long)F.nxy * k] : F.bg[id][k];
}

__device__ __forceinline__ int pml_index(int k, int nz, int nbot, int ntop) {
  if (k < nbot) return k;
  if (k >= nz - ntop) return nbot + (k - (nz - ntop));
  return -1;
}

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
                                                C2<T>* spec, int ny, int nz, long nxy) {
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
  } else {  // XM_DAMP_INV: RHSv -= term, boundary planes excluded (they are zeroed)
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      const int l = e / n, x = e - l * n;
      const long L = L0 + l;
      const long q = L / ny, j = L - q * ny;
      const long k0 = 2 * q, k1 = k0 + 1;
      const long o0 = x + (long)n * j + nxy * k0;
      if (k0 != 0 && k0 != nz - 1) out[o0] = out[o0] - r[e].x;
      if (k1 < nz - 1) out[o0 + nxy] = out[o0 + nxy] - r[e].y;
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
    const int nh = nx / 2 + 1;
    for (int e = threadIdx.x; e < nl * ny; e += blockDim.x) {
      const int mm = e / ny, y = e - mm * ny;
      const int m = i0 + mm;
      const int mi = (m <= nx / 2) ? m : nx - m;
      const T c = rates[mi + (long)nh * y];
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

// ---------------------------------------------------------------------------
// z-columns: DBYD2 + MTRIDAG ispecial=0 (dbyd2.f90:88-199, mtridag.f90:94-111)
// with the LU factors of the constant matrix precomputed in FP64 on the host,
// one thread per (x,y) column, lanes coalesced along x.  The forward sweep
// forms the right-hand side from a 6-value register window of B = a + beta*s;
// the backward sweep hands d/dz = x*stretch(k) to a fused epilogue (the
// pointwise part of the RHS operator of one LDDRK stage).
// ---------------------------------------------------------------------------
template <typename T, int EPI>
__global__ __launch_bounds__(256) void k_zsolve(Fld<T> F, TriCoef<T> tc,
                                                const T* __restrict__ lowr,
                                                const T* __restrict__ invb,
                                                const T* __restrict__ gam, int ibc,
                                                const T* __restrict__ inA,
                                                const T* __restrict__ inS, T beta_in,
                                                T* __restrict__ ud,
                                                const T* __restrict__ prov, T provscale,
                                                T* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= F.nxy) return;
  const long nxy = F.nxy;
  const int n = F.nz;
  auto B = [&](int k) -> T {
    const long q = c + nxy * k;
    T v = inA[q];
    if (inS) v = v + beta_in * inS[q];
    return v;
  };
  // ---- forward sweep
  T wm1 = 0, w0 = B(0), w1 = B(1), w2 = B(2), w3 = B(3), w4 = B(4);
  T r0;
  if (ibc == 0)
    r0 = 0;
  else if (ibc == 1 || ibc == 3)
    r0 = tc.b1 * w0 + tc.b2 * w1 + tc.b3 * w2 + tc.b4 * w3;
  else
    r0 = prov ? prov[c] * provscale : (T)0;
  const T r1 = tc.c1 * w0 + tc.c2 * w1 + tc.c3 * w2 + tc.c4 * w3 + tc.c5 * w4;
  if (ibc & 1) r0 = (tc.alpha2 * r0 - tc.beta1 * r1) * tc.cst;
  T u = r0 * invb[0];
  ud[c] = u;
  u = (r1 - lowr[1] * u) * invb[1];
  ud[c + nxy] = u;
  for (int k = 2; k <= n - 3; ++k) {
    const T r = tc.ci1 * (w3 - w1) + tc.ci2 * (w4 - w0);
    u = (r - lowr[k] * u) * invb[k];
    ud[c + nxy * k] = u;
    wm1 = w0;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
    if (k + 3 <= n - 1) w4 = B(k + 3);
  }
  // window now holds B(n-5..n-1) in wm1..w3
  const T rn2 = -(tc.c1 * w3 + tc.c2 * w2 + tc.c3 * w1 + tc.c4 * w0 + tc.c5 * wm1);
  u = (rn2 - lowr[n - 2] * u) * invb[n - 2];
  ud[c + nxy * (n - 2)] = u;
  T rn1;
  if (ibc == 0)
    rn1 = 0;
  else if (ibc == 1 || ibc == 2) {
    rn1 = -(tc.b1 * w3 + tc.b2 * w2 + tc.b3 * w1 + tc.b4 * w0);
    rn1 = (tc.alpha2 * rn1 - tc.beta1 * rn2) * tc.cst;
  } else {
    rn1 = prov ? prov[c] * provscale : (T)0;
  }
  u = (rn1 - lowr[n - 1] * u) * invb[n - 1];
  // ---- backward sweep + epilogue
  T x = u;
  for (int k = n - 1; k >= 0; --k) {
    if (k < n - 1) x = ud[c + nxy * k] - gam[k + 1] * x;
    const T val = x * F.stretch[k];
    const long q = c + nxy * k;
    if (EPI == EPI_STORE) {
      out[q] = val;
    } else if (EPI == EPI_P3D || EPI == EPI_P2D) {
      T dz = val;
      const int pi = pml_index(k, n, F.npmlbot, F.npmltop);
      if (pi >= 0) {
        const long pq = c + nxy * pi;
        T psi = F.pmlvars[0][pq];
        if (F.beta != (T)0) psi = psi + F.beta * F.scrpml[0][pq];
        F.scrpml[0][pq] = bgv(F, SPARC_BG_AZ, c, k) * dz + bgv(F, SPARC_BG_BZPML, c, k) * psi;
        dz = dz + psi;
      }
      T div;
      if (EPI == EPI_P3D)
        div = F.W[0][q] + F.W[1][q] + dz;
      else
        div = F.W[0][q] + dz;
      T xz = F.a[SPARC_XI_Z][q];
      if (F.beta != (T)0) xz = xz + F.beta * F.scr[SPARC_XI_Z][q];
      F.W[3][q] = -bgv(F, SPARC_BG_C2RHO0, c, k) * div - xz * bgv(F, SPARC_BG_GRADP0_Z, c, k);
      F.W[4][q] = -bgv(F, SPARC_BG_RHO0, c, k) * div - xz * bgv(F, SPARC_BG_GRADRHO0_Z, c, k);
    } else {  // EPI_RHS3D / EPI_RHS2D
      T gz = val;
      const int pi = pml_index(k, n, F.npmlbot, F.npmltop);
      if (pi >= 0) {
        const long pq = c + nxy * pi;
        T psi = F.pmlvars[1][pq];
        if (F.beta != (T)0) psi = psi + F.beta * F.scrpml[1][pq];
        F.scrpml[1][pq] = bgv(F, SPARC_BG_AZ, c, k) * gz + bgv(F, SPARC_BG_BZPML, c, k) * psi;
        gz = gz + psi;
      }
      const T ri = bgv(F, SPARC_BG_RHOINV, c, k);
      const T rho = F.W[4][q];
      if (EPI == EPI_RHS3D) {
        T rvx = -ri * F.W[0][q];
        T rvy = -ri * F.W[1][q];
        T rvz = -ri * gz;
        rvz = rvz - rho * F.g[k] * ri;
        if (k == F.src_plane)
          rvz = rvz + F.forcing[c] / bgv(F, SPARC_BG_RHO0, c, k) * F.delta_width;
        if (k == 0 || k == n - 1) {
#pragma unroll
          for (int f = 0; f < SPARC_NFIELDS; ++f) F.scr[f][q] = 0;
        } else {
          T tv[3];
#pragma unroll
          for (int f = 0; f < 3; ++f) {
            tv[f] = F.a[SPARC_V_X + f][q];
            if (F.beta != (T)0) tv[f] = tv[f] + F.beta * F.scr[SPARC_V_X + f][q];
          }
          F.scr[SPARC_XI_X][q] = tv[0];
          F.scr[SPARC_XI_Y][q] = tv[1];
          F.scr[SPARC_XI_Z][q] = tv[2];
          F.scr[SPARC_V_X][q] = rvx;
          F.scr[SPARC_V_Y][q] = rvy;
          F.scr[SPARC_V_Z][q] = rvz;
        }
      } else {
        T rvx = -ri * F.W[0][q];
        T rvz = -ri * (gz + rho * F.g[k]);
        if (k == F.src_plane)
          rvz = rvz + F.forcing[c] * F.delta_width / bgv(F, SPARC_BG_RHO0, c, k);
        if (k == 0) {
          F.scr[SPARC_XI_X][q] = 0;
          F.scr[SPARC_XI_Z][q] = 0;
          F.scr[SPARC_V_X][q] = 0;
          F.scr[SPARC_V_Z][q] = 0;
        } else {
          T tvx = F.a[SPARC_V_X][q], tvz = F.a[SPARC_V_Z][q];
          if (F.beta != (T)0) {
            tvx = tvx + F.beta * F.scr[SPARC_V_X][q];
            tvz = tvz + F.beta * F.scr[SPARC_V_Z][q];
          }
          F.scr[SPARC_XI_X][q] = tvx;
          F.scr[SPARC_XI_Z][q] = tvz;
          F.scr[SPARC_V_X][q] = rvx;
          F.scr[SPARC_V_Z][q] = rvz;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// a += dt*scr fused with the 11-point z filter (step.f90:114-118,
// physics.f90:760-788, physics2d.f90:527-560), or the filter alone
// (FILTER_VAR_Z, kernels.f90:845-881) when scr == nullptr.  One thread per
// column with an 11-value register window; output plane k is written after
// input plane k+5 is read, so the update is in place.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_update_filter(T* a, const T* __restrict__ scr, T dt,
                                                       int nxy, int nz, int kst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nxy) return;
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
  for (int k = 0; k < nz; ++k) {
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
    w[10] = L(k + 6);
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

}  // namespace sparc
