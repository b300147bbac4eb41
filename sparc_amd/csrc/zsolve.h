// z-columns: DBYD2 + MTRIDAG ispecial=0 (dbyd2.f90:88-199, mtridag.f90:94-111)
// with the LU factors of the constant matrix precomputed in FP64 on the host,
// one thread per (x,y) column, lanes coalesced along x.  The forward sweep
// forms the right-hand side from a register window of B = a + beta*s; the
// backward sweep hands d/dz = x*stretch(k) to a fused epilogue (the pointwise
// part of the RHS operator of one LDDRK stage).
//
// A grid of nx*ny columns gives one wave per SIMD at 256x256, so the sweeps
// are latency-bound unless many loads are in flight per wave: both sweeps
// run in chunks of planes whose inputs (B in the forward sweep; the
// forward-sweep result ud and every epilogue input in the backward sweep) are
// loaded into registers one chunk ahead of their use.
// Included inside namespace sparc by kernels.h.
#pragma once

// Everything the backward sweep reads at one plane.
template <typename T>
struct EIn {
  T ud;
  T w0, w1, w4;
  T a0, a1, a2;  // a of the fields read by the epilogue
  T s0, s1, s2;  // scr of the same fields (stage temp = a + beta*scr)
  T g0, g1, g2, g3;  // 3-D backgrounds (unused when z-profiles)
  T pv, ps, paz, pbz;  // PML psi and coefficients on PML planes
};

template <typename T, int EPI>
__device__ __forceinline__ void load_epi(const Fld<T>& F, const T* __restrict__ ud, int c,
                                         int k, EIn<T>& e) {
  const int n = F.nz;
  if (k < 0) return;
  const long nxy = F.nxy;
  const long q = c + nxy * k;
  if (k < n - 1) e.ud = ud[q];
  const bool st = F.beta != (T)0;
  if (EPI == EPI_P3D || EPI == EPI_P2D) {
    e.w0 = F.W[0][q];
    if (EPI == EPI_P3D) e.w1 = F.W[1][q];
    e.a2 = F.a[SPARC_XI_Z][q];
    if (st) e.s2 = F.scr[SPARC_XI_Z][q];
    if (F.bg3d[SPARC_BG_C2RHO0]) e.g0 = F.bg[SPARC_BG_C2RHO0][q];
    if (F.bg3d[SPARC_BG_GRADP0_Z]) e.g1 = F.bg[SPARC_BG_GRADP0_Z][q];
    if (F.bg3d[SPARC_BG_RHO0]) e.g2 = F.bg[SPARC_BG_RHO0][q];
    if (F.bg3d[SPARC_BG_GRADRHO0_Z]) e.g3 = F.bg[SPARC_BG_GRADRHO0_Z][q];
  } else if (EPI == EPI_RHS3D || EPI == EPI_RHS2D) {
    e.w0 = F.W[0][q];
    if (EPI == EPI_RHS3D) e.w1 = F.W[1][q];
    e.w4 = F.W[4][q];
    e.a0 = F.a[SPARC_V_X][q];
    e.a2 = F.a[SPARC_V_Z][q];
    if (EPI == EPI_RHS3D) e.a1 = F.a[SPARC_V_Y][q];
    if (st) {
      e.s0 = F.scr[SPARC_V_X][q];
      e.s2 = F.scr[SPARC_V_Z][q];
      if (EPI == EPI_RHS3D) e.s1 = F.scr[SPARC_V_Y][q];
    }
    if (F.bg3d[SPARC_BG_RHOINV]) e.g0 = F.bg[SPARC_BG_RHOINV][q];
  }
  if (EPI != EPI_STORE) {
    const int pi = pml_index(k, n, F.npmlbot, F.npmltop);
    if (pi >= 0) {
      const int w = (EPI == EPI_P3D || EPI == EPI_P2D) ? 0 : 1;
      const long pq = c + nxy * pi;
      e.pv = F.pmlvars[w][pq];
      if (st) e.ps = F.scrpml[w][pq];
      e.paz = bgv(F, SPARC_BG_AZ, c, k);
      e.pbz = bgv(F, SPARC_BG_BZPML, c, k);
    }
  }
}

template <typename T>
__device__ __forceinline__ T bg_or(const Fld<T>& F, int id, T v3d, int k) {
  return F.bg3d[id] ? v3d : F.bg[id][k];
}

template <typename T, int EPI>
__device__ __forceinline__ void epilogue(const Fld<T>& F, int c, int k, T val, const EIn<T>& e,
                                         T* __restrict__ out) {
  const int n = F.nz;
  const long q = c + (long)F.nxy * k;
  const bool st = F.beta != (T)0;
  if (EPI == EPI_STORE) {
    out[q] = val;
    return;
  }
  const int pi = pml_index(k, n, F.npmlbot, F.npmltop);
  if (EPI == EPI_P3D || EPI == EPI_P2D) {
    T dz = val;
    if (pi >= 0) {
      T psi = e.pv;
      if (st) psi = psi + F.beta * e.ps;
      F.scrpml[0][c + (long)F.nxy * pi] = e.paz * dz + e.pbz * psi;
      dz = dz + psi;
    }
    T div;
    if (EPI == EPI_P3D)
      div = e.w0 + e.w1 + dz;
    else
      div = e.w0 + dz;
    T xz = e.a2;
    if (st) xz = xz + F.beta * e.s2;
    const T c2r = bg_or(F, SPARC_BG_C2RHO0, e.g0, k);
    const T gp = bg_or(F, SPARC_BG_GRADP0_Z, e.g1, k);
    const T r0 = bg_or(F, SPARC_BG_RHO0, e.g2, k);
    const T gr = bg_or(F, SPARC_BG_GRADRHO0_Z, e.g3, k);
    F.W[3][q] = -c2r * div - xz * gp;
    F.W[4][q] = -r0 * div - xz * gr;
  } else {
    T gz = val;
    if (pi >= 0) {
      T psi = e.pv;
      if (st) psi = psi + F.beta * e.ps;
      F.scrpml[1][c + (long)F.nxy * pi] = e.paz * gz + e.pbz * psi;
      gz = gz + psi;
    }
    const T ri = bg_or(F, SPARC_BG_RHOINV, e.g0, k);
    const T rho = e.w4;
    if (EPI == EPI_RHS3D) {
      T rvx = -ri * e.w0;
      T rvy = -ri * e.w1;
      T rvz = -ri * gz;
      rvz = rvz - rho * F.g[k] * ri;
      if (k == F.src_plane) rvz = rvz + F.forcing[c] / bgv(F, SPARC_BG_RHO0, c, k) * F.delta_width;
      if (k == 0 || k == n - 1) {
#pragma unroll
        for (int f = 0; f < SPARC_NFIELDS; ++f) F.scr[f][q] = 0;
      } else {
        T tvx = e.a0, tvy = e.a1, tvz = e.a2;
        if (st) {
          tvx = tvx + F.beta * e.s0;
          tvy = tvy + F.beta * e.s1;
          tvz = tvz + F.beta * e.s2;
        }
        F.scr[SPARC_XI_X][q] = tvx;
        F.scr[SPARC_XI_Y][q] = tvy;
        F.scr[SPARC_XI_Z][q] = tvz;
        F.scr[SPARC_V_X][q] = rvx;
        F.scr[SPARC_V_Y][q] = rvy;
        F.scr[SPARC_V_Z][q] = rvz;
      }
    } else {
      T rvx = -ri * e.w0;
      T rvz = -ri * (gz + rho * F.g[k]);
      if (k == F.src_plane) rvz = rvz + F.forcing[c] * F.delta_width / bgv(F, SPARC_BG_RHO0, c, k);
      if (k == 0) {
        F.scr[SPARC_XI_X][q] = 0;
        F.scr[SPARC_XI_Z][q] = 0;
        F.scr[SPARC_V_X][q] = 0;
        F.scr[SPARC_V_Z][q] = 0;
      } else {
        T tvx = e.a0, tvz = e.a2;
        if (st) {
          tvx = tvx + F.beta * e.s0;
          tvz = tvz + F.beta * e.s2;
        }
        F.scr[SPARC_XI_X][q] = tvx;
        F.scr[SPARC_XI_Z][q] = tvz;
        F.scr[SPARC_V_X][q] = rvx;
        F.scr[SPARC_V_Z][q] = rvz;
      }
    }
  }
}

template <typename T, int EPI>
__global__ __launch_bounds__(64) void k_zsolve(Fld<T> F, TriCoef<T> tc, const T* __restrict__ lowr,
                                               const T* __restrict__ invb,
                                               const T* __restrict__ gam, int ibc,
                                               const T* __restrict__ inA,
                                               const T* __restrict__ inS, T beta_in,
                                               T* __restrict__ ud, const T* __restrict__ prov,
                                               T provscale, T* __restrict__ out) {
  constexpr int UF = 16;  // forward chunk (planes)
  constexpr int UB = 4;   // backward chunk (planes)
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
  // ---- forward sweep: rows 0 and 1
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
  // ---- interior rows 2..n-3; row k consumes B(k+3) after its update
  T na[UF], ns[UF];
  auto load_chunk = [&](int kb) {
#pragma unroll
    for (int j = 0; j < UF; ++j) {
      const int kk = kb + j;
      if (kk < n) {
        const long q = c + nxy * kk;
        na[j] = inA[q];
        if (inS) ns[j] = inS[q];
      }
    }
  };
  load_chunk(5);
  for (int k0 = 2; k0 <= n - 3; k0 += UF) {
    T ca[UF], cs[UF];
#pragma unroll
    for (int j = 0; j < UF; ++j) {
      ca[j] = na[j];
      cs[j] = ns[j];
    }
    if (k0 + UF <= n - 3) load_chunk(k0 + UF + 3);
#pragma unroll
    for (int j = 0; j < UF; ++j) {
      const int k = k0 + j;
      if (k <= n - 3) {
        const T r = tc.ci1 * (w3 - w1) + tc.ci2 * (w4 - w0);
        u = (r - lowr[k] * u) * invb[k];
        ud[c + nxy * k] = u;
        wm1 = w0;
        w0 = w1;
        w1 = w2;
        w2 = w3;
        w3 = w4;
        if (k + 3 < n) w4 = inS ? ca[j] + beta_in * cs[j] : ca[j];
      }
    }
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
  // ---- backward sweep + epilogue, chunks of UB planes loaded one chunk ahead
  EIn<T> nx[UB];
#pragma unroll
  for (int j = 0; j < UB; ++j) load_epi<T, EPI>(F, ud, c, n - 1 - j, nx[j]);
  T x = u;
  for (int kh = n - 1; kh >= 0; kh -= UB) {
    EIn<T> cu[UB];
#pragma unroll
    for (int j = 0; j < UB; ++j) cu[j] = nx[j];
    if (kh - UB >= 0) {
#pragma unroll
      for (int j = 0; j < UB; ++j) load_epi<T, EPI>(F, ud, c, kh - UB - j, nx[j]);
    }
#pragma unroll
    for (int j = 0; j < UB; ++j) {
      const int k = kh - j;
      if (k >= 0) {
        if (k < n - 1) x = cu[j].ud - gam[k + 1] * x;
        epilogue<T, EPI>(F, c, k, x * F.stretch[k], cu[j], out);
      }
    }
  }
}

