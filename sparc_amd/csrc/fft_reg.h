// Compile-time-sized batched complex FFT in LDS with register butterflies of
// radix 16/8/4/3 (Stockham autosort, natural order in and out).  Line length
// N is a template parameter so every index computation is by constants; the
// twiddle table is staged in LDS once per block.  Used for the sizes of the
// benchmark grids (32..512); other sizes use the generic core of fft_lds.h.
// Included inside namespace sparc by kernels.h.
#pragma once

template <int N>
struct RPlan;
template <>
struct RPlan<32> {
  static constexpr int P = 2;
  static constexpr int R[3] = {8, 4, 1};
};
template <>
struct RPlan<64> {
  static constexpr int P = 2;
  static constexpr int R[3] = {8, 8, 1};
};
template <>
struct RPlan<96> {
  static constexpr int P = 3;
  static constexpr int R[3] = {8, 4, 3};
};
template <>
struct RPlan<128> {
  static constexpr int P = 2;
  static constexpr int R[3] = {16, 8, 1};
};
template <>
struct RPlan<192> {
  static constexpr int P = 3;
  static constexpr int R[3] = {8, 8, 3};
};
template <>
struct RPlan<256> {
  static constexpr int P = 2;
  static constexpr int R[3] = {16, 16, 1};
};
template <>
struct RPlan<384> {
  static constexpr int P = 3;
  static constexpr int R[3] = {16, 8, 3};
};
template <>
struct RPlan<512> {
  static constexpr int P = 3;
  static constexpr int R[3] = {8, 8, 8};
};

// cos/sin(2 pi m / 16), m = 0..15
__device__ constexpr double kC16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double kS16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// x * exp(DIR * 2 pi i M / R), M and R compile-time (R a power of two <= 16)
template <typename T, int R, int M, int DIR>
__device__ __forceinline__ C2<T> twc(C2<T> x) {
  constexpr int m16 = ((M % R) * (16 / R)) % 16;
  if constexpr (m16 == 0) {
    return x;
  } else if constexpr (m16 == 8) {
    return {-x.x, -x.y};
  } else if constexpr (m16 == 4) {  // exp(DIR i pi/2) = DIR*i
    return DIR > 0 ? C2<T>{-x.y, x.x} : C2<T>{x.y, -x.x};
  } else if constexpr (m16 == 12) {
    return DIR > 0 ? C2<T>{x.y, -x.x} : C2<T>{-x.y, x.x};
  } else {
    const T c = (T)kC16[m16];
    const T s = (T)(DIR > 0 ? kS16[m16] : -kS16[m16]);
    return {x.x * c - x.y * s, x.x * s + x.y * c};
  }
}

__host__ __device__ constexpr int bitrev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b)
    if (i & (1 << b)) r |= 1 << (bits - 1 - b);
  return r;
}
__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

template <typename T, int R, int DIR, int LEN, int I, int K>
__device__ __forceinline__ void bfly(C2<T> (&v)[R]) {
  const C2<T> a = v[I + K];
  const C2<T> b = twc<T, LEN, K, DIR>(v[I + K + LEN / 2]);
  v[I + K] = {a.x + b.x, a.y + b.y};
  v[I + K + LEN / 2] = {a.x - b.x, a.y - b.y};
}

template <typename T, int R, int DIR, int LEN, int I, int K>
__device__ __forceinline__ void bfly_k(C2<T> (&v)[R]) {
  if constexpr (K < LEN / 2) {
    bfly<T, R, DIR, LEN, I, K>(v);
    bfly_k<T, R, DIR, LEN, I, K + 1>(v);
  }
}
template <typename T, int R, int DIR, int LEN, int I>
__device__ __forceinline__ void bfly_i(C2<T> (&v)[R]) {
  if constexpr (I < R) {
    bfly_k<T, R, DIR, LEN, I, 0>(v);
    bfly_i<T, R, DIR, LEN, I + LEN>(v);
  }
}
template <typename T, int R, int DIR, int LEN>
__device__ __forceinline__ void bfly_len(C2<T> (&v)[R]) {
  if constexpr (LEN <= R) {
    bfly_i<T, R, DIR, LEN, 0>(v);
    bfly_len<T, R, DIR, LEN * 2>(v);
  }
}

// In-register DFT of R points (R = 2, 4, 8, 16: radix-2 DIT; R = 3: direct).
template <typename T, int R, int DIR>
__device__ __forceinline__ void dft_reg(C2<T> (&v)[R]) {
  if constexpr (R == 3) {
    const T h = (T)0.5, s3 = (T)0.86602540378443864676;
    const C2<T> s = {v[1].x + v[2].x, v[1].y + v[2].y};
    const C2<T> d = {v[1].x - v[2].x, v[1].y - v[2].y};
    const C2<T> m = {v[0].x - h * s.x, v[0].y - h * s.y};
    const C2<T> id = (DIR < 0) ? C2<T>{s3 * d.y, -s3 * d.x} : C2<T>{-s3 * d.y, s3 * d.x};
    v[0] = {v[0].x + s.x, v[0].y + s.y};
    v[1] = {m.x + id.x, m.y + id.y};
    v[2] = {m.x - id.x, m.y - id.y};
  } else {
    constexpr int bits = ilog2(R);
    C2<T> t[R];
#pragma unroll
    for (int i = 0; i < R; ++i) t[i] = v[bitrev(i, bits)];
    bfly_len<T, R, DIR, 2>(t);
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = t[i];
  }
}

// LDS layout of a batch of lines: element i of a line at pad(i) = i + i/16
// (breaks the stride-16 bank conflicts of radix-16 Stockham writes), lines
// LS = pad(N) + 8 elements apart (rotates banks between lines).
__host__ __device__ constexpr int lpad(int i) { return i + (i >> 4); }
template <int N>
struct LineStride {
  static constexpr int LS = lpad(N) + 8;
};
// y tiles are filled and drained with consecutive lanes on consecutive lines:
// a stride of 1 (mod 32) complex values puts 16 such lanes on distinct banks
template <int N>
struct YLineStride {
  static constexpr int LS = (lpad(N) + 31) / 32 * 32 + 1;
};

// One in-place Stockham pass of radix R over LPB lines of length N in LDS:
// every thread loads its butterflies into registers, the block synchronizes,
// then the results are written back.
template <typename T, int N, int R, int NS, int DIR, int LPB, int NT, int LS>
__device__ __forceinline__ void rpass(C2<T>* buf, const C2<T>* __restrict__ tw) {
  constexpr int NB = N / R;
  constexpr int TOT = LPB * NB;
  constexpr int IT = (TOT + NT - 1) / NT;
  C2<T> v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (TOT % NT == 0 || w < TOT) {
      const int line = w / NB;
      const int j = w - line * NB;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[line * LS + lpad(j + r * NB)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (TOT % NT == 0 || w < TOT) {
      const int line = w / NB;
      const int j = w - line * NB;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          C2<T> t = tw[r * k * (N / (NS * R))];
          if (DIR > 0) t.y = -t.y;
          v[it][r] = cmul(v[it][r], t);
        }
      }
      dft_reg<T, R, DIR>(v[it]);
      const int idxD = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[line * LS + lpad(idxD + r * NS)] = v[it][r];
    }
  }
  __syncthreads();
}

// Full in-place FFT of LPB lines in buf.
template <typename T, int N, int DIR, int LPB, int NT, int LS = LineStride<N>::LS>
__device__ __forceinline__ void rfft(C2<T>* buf, const C2<T>* __restrict__ tw) {
  using PL = RPlan<N>;
  constexpr int R0 = PL::R[0], R1 = PL::R[1], R2 = PL::R[2];
  rpass<T, N, R0, 1, DIR, LPB, NT, LS>(buf, tw);
  rpass<T, N, R1, R0, DIR, LPB, NT, LS>(buf, tw);
  if constexpr (PL::P == 3) rpass<T, N, R2, R0 * R1, DIR, LPB, NT, LS>(buf, tw);
}

// lines per block: 16 KiB of complex values per block and 256 threads, i.e.
// ~8 (FP32) / 4 (FP64) values per thread in the load and store phases
// (measured on MI355X: 8 lines x 256 threads moves 4.5-5.1 TB/s at N = 256,
// 16 lines x 256 threads 2.9-3.9 TB/s)
template <typename T, int N>
struct XCfg {
  static constexpr int LPB = (16384 / (int)sizeof(C2<T>)) / N > 0
                                 ? (16384 / (int)sizeof(C2<T>)) / N
                                 : 1;
  static constexpr int NT = 256;
  static constexpr int LS = LineStride<N>::LS;
};

// x-lines with compile-time N: same contract as k_xlines (kernels.h).
// blockIdx.y selects one of up to 3 fields (a/s/out/spec given per field).
template <typename T>
struct XArgs {
  const T* a[3];
  const T* s[3];
  T* out[3];
  C2<T>* spec[3];
  int edges;  // damping inverse also on planes 1 and nz (option 2 zeroes before damping)
  // derivative written as a momentum right-hand side, -rhoinv * d, zero on
  // planes 1 and nz (displacement.f90:589-597,639-642): rhoinv as a z profile
  // (rz) or a 3-D array (r3)
  int rhs;
  const T* rz;
  const T* r3;
  // derivative combined into the output (curl assembly): out = ald d + gout out
  int acc;
  T ald, gout;
  // damping forward pass (k_xw256, k_xlines_n): the stage temp a + beta s it
  // transforms is also stored here (RHS_xi = temp v; zero on planes 1 and
  // nz), when set
  T* cp[3];
  // damping inverse pass (k_xw256, k_xlines_n): the RHS is read from here
  // (the result still goes to out), when set
  const T* rin[3];
  // damping inverse passes that also do the next stage's damping forward
  // pass (k_xw256_rhs_dampinv<..., true>, k_xlines_n<..., XM_DAMP_INV, ...,
  // true>): with the RHS r each line leaves in out, the next stage temp
  // pa + pbeta r is stored to pcp (RHS_xi, zero on planes 1 and nz) and its
  // forward transform to spec, over the spectrum the line has just read
  const T* pa[3];
  T* pcp[3];
  T pbeta;
};

// plain derivative store: d, or combined with what out holds (X.acc)
template <typename T>
__device__ __forceinline__ void store_d(const XArgs<T>& X, T* out, long o, T d) {
  out[o] = X.acc ? X.ald * d + X.gout * out[o] : d;
}

template <typename T>
__device__ __forceinline__ T rhs_of(const XArgs<T>& X, long k, long o, int nz, T d) {
  if (k == 0 || k == nz - 1) return (T)0;
  const T ri = X.r3 ? X.r3[o] : X.rz[k];
  return -ri * d;
}

// PREP (XM_DAMP_INV only): each line then also does the next stage's
// XM_DAMP_FWD from the RHS rows it has just written (X.pa, X.pbeta, X.pcp;
// the spectrum goes over the one the line read).
template <typename T, int N, int MODE, int LPB = XCfg<T, N>::LPB, int NT = XCfg<T, N>::NT,
          bool PREP = false>
__global__ __launch_bounds__(NT) void k_xlines_n(int nrows, int nlines, XArgs<T> X, T beta,
                                                 const T* __restrict__ mult,
                                                 const C2<T>* __restrict__ twg, int ny, int nz,
                                                 long nxy) {
  constexpr int LS = LineStride<N>::LS;
  __shared__ C2<T> b0[LPB * LS];
  __shared__ C2<T> tw[N];
  const int fld = blockIdx.y;
  const T* __restrict__ a = X.a[fld];
  const T* __restrict__ s = X.s[fld];
  T* out = X.out[fld];
  C2<T>* spec = X.spec[fld];
  for (int m = threadIdx.x; m < N; m += NT) tw[m] = twg[m];
  const int L0 = blockIdx.x * LPB;
  const int nl = min(LPB, nlines - L0);
  // damping inverse: the RHS values it updates, loaded with the spectrum.
  // Every load is issued before the first use (a + beta s, the LDS store).
  constexpr int IT = (LPB * N + NT - 1) / NT;
  constexpr bool RL = MODE == XM_DERIV || MODE == XM_FILTER || MODE == XM_DAMP_FWD;
  static_assert(!PREP || MODE == XM_DAMP_INV, "the next stage is prepared by the inverse pass");
  T ov0[MODE == XM_DAMP_INV ? IT : 1], ov1[MODE == XM_DAMP_INV ? IT : 1];
  T pa0[PREP ? IT : 1], pa1[PREP ? IT : 1];
  T ra[RL ? IT : 1], rs[RL ? IT : 1], ia[RL ? IT : 1], is[RL ? IT : 1];
  C2<T> sv[RL ? 1 : IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * NT;
    const int l = e / N, x = e - l * N;
    if (e < LPB * N && l < nl) {
      const long L = L0 + l;
      if (RL) {
        long o0, st;
        bool h1;
        if (MODE == XM_DAMP_FWD) {
          const long q = L / ny, j = L - q * ny;
          o0 = x + (long)N * j + nxy * (2 * q);
          st = nxy;
          h1 = 2 * q + 1 < nz;
        } else {
          o0 = 2 * L * N + x;
          st = N;
          h1 = 2 * L + 1 < nrows;
        }
        ra[it] = a[o0];
        if (s) rs[it] = s[o0];
        ia[it] = h1 ? a[o0 + st] : (T)0;
        if (s) is[it] = h1 ? s[o0 + st] : (T)0;
      } else {
        sv[it] = spec[L * N + x];
        if constexpr (MODE == XM_DAMP_INV) {
          // the RHS is read from X.rin when set (the result still goes to out)
          const T* __restrict__ ri = X.rin[fld] ? X.rin[fld] : out;
          const long q = L / ny, j = L - q * ny;
          const long k0 = 2 * q, k1 = k0 + 1;
          const long o0 = x + (long)N * j + nxy * k0;
          // PREP: rows the pass leaves alone keep the RHS held in out
          ov0[it] = (X.edges || (k0 != 0 && k0 != nz - 1)) ? ri[o0] : (PREP ? out[o0] : (T)0);
          ov1[it] = (k1 < nz - 1 || (X.edges && k1 < nz)) ? ri[o0 + nxy]
                                                          : ((PREP && k1 < nz) ? out[o0 + nxy] : (T)0);
          if constexpr (PREP) {
            pa0[it] = X.pa[fld][o0];
            pa1[it] = k1 < nz ? X.pa[fld][o0 + nxy] : (T)0;
          }
        }
      }
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int e = threadIdx.x + it * NT;
    const int l = e / N, x = e - l * N;
    if (e < LPB * N) {
      C2<T> v = {0, 0};
      if (l < nl) {
        if (RL)
          v = s ? C2<T>{ra[it] + beta * rs[it], ia[it] + beta * is[it]} : C2<T>{ra[it], ia[it]};
        else
          v = sv[it];
        if constexpr (MODE == XM_DAMP_FWD) {
          // the stage temp is also stored to X.cp when set, zero on planes 1 and nz
          T* __restrict__ cp = X.cp[fld];
          if (cp) {
            const long L = L0 + l, q = L / ny, j = L - q * ny, k0 = 2 * q;
            const long o0 = x + (long)N * j + nxy * k0;
            cp[o0] = (k0 == 0 || k0 == nz - 1) ? (T)0 : v.x;
            if (k0 + 1 < nz) cp[o0 + nxy] = (k0 + 1 == nz - 1) ? (T)0 : v.y;
          }
        }
      }
      b0[l * LS + lpad(x)] = v;
    }
  }
  __syncthreads();
  if (MODE == XM_DAMP_INV)
    rfft<T, N, +1, LPB, NT>(b0, tw);
  else
    rfft<T, N, -1, LPB, NT>(b0, tw);
  C2<T>* r = b0;
  if (MODE == XM_DAMP_FWD) {
    for (int e = threadIdx.x; e < nl * N; e += NT) {
      const int l = e / N, x = e - l * N;
      spec[(long)L0 * N + e] = r[l * LS + lpad(x)];
    }
    return;
  }
  if (MODE == XM_DERIV || MODE == XM_FILTER) {
#pragma unroll
    for (int e = threadIdx.x; e < LPB * N; e += NT) {
      const int l = e / N, x = e - l * N;
      const T c = mult[x];
      C2<T> v = r[l * LS + lpad(x)];
      if (MODE == XM_DERIV)
        v = {-c * v.y, c * v.x};
      else
        v = {c * v.x, c * v.y};
      r[l * LS + lpad(x)] = v;
    }
    __syncthreads();
    rfft<T, N, +1, LPB, NT>(b0, tw);
#pragma unroll
    for (int e = threadIdx.x; e < nl * N; e += NT) {
      const int l = e / N, x = e - l * N;
      const long L = L0 + l;
      const long r0 = 2 * L, r1 = r0 + 1;
      const C2<T> v = r[l * LS + lpad(x)];
      if (MODE == XM_DERIV && X.rhs) {  // row r = j + ny*k
        out[r0 * N + x] = rhs_of(X, r0 / ny, r0 * N + x, nz, v.x);
        if (r1 < nrows) out[r1 * N + x] = rhs_of(X, r1 / ny, r1 * N + x, nz, v.y);
      } else {
        store_d(X, out, r0 * N + x, v.x);
        if (r1 < nrows) store_d(X, out, r1 * N + x, v.y);
      }
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + it * NT;
      if (e >= nl * N) break;
      const int l = e / N, x = e - l * N;
      const long L = L0 + l;
      const long q = L / ny, j = L - q * ny;
      const long k0 = 2 * q, k1 = k0 + 1;
      const long o0 = x + (long)N * j + nxy * k0;
      const C2<T> v = r[l * LS + lpad(x)];
      const bool w0 = X.edges || (k0 != 0 && k0 != nz - 1);
      const bool w1 = k1 < nz - 1 || (X.edges && k1 < nz);
      if (w0) out[o0] = ov0[it] - v.x;
      if (w1) out[o0 + nxy] = ov1[it] - v.y;
      if constexpr (PREP) {
        // the next stage temp pa + pbeta r, as XM_DAMP_FWD forms and stores it
        const T r0 = w0 ? ov0[it] - v.x : ov0[it], r1 = w1 ? ov1[it] - v.y : ov1[it];
        const C2<T> u = {pa0[it] + X.pbeta * r0, pa1[it] + X.pbeta * r1};
        T* __restrict__ cp = X.pcp[fld];
        cp[o0] = (k0 == 0 || k0 == nz - 1) ? (T)0 : u.x;
        if (k1 < nz) cp[o0 + nxy] = (k1 == nz - 1) ? (T)0 : u.y;
        r[l * LS + lpad(x)] = u;
      }
    }
    if constexpr (PREP) {
      __syncthreads();
      rfft<T, N, -1, LPB, NT>(b0, tw);
      for (int e = threadIdx.x; e < nl * N; e += NT) {
        const int l = e / N, x = e - l * N;
        spec[(long)L0 * N + e] = r[l * LS + lpad(x)];
      }
    }
  }
}

// y-lines with compile-time NY: tile of TX columns of one plane (DERIV,
// FILTER: TX real columns packed in TX/2 complex lines) or TX complex x-bins of
// a plane-pair spectrum (DAMP).  blockIdx.z selects the field.
__host__ __device__ constexpr int pow2_floor(int v) { return v >= 2 ? 2 * pow2_floor(v / 2) : 1; }

// y tile of 32 KiB, rows read in 128-byte segments; the damping y pass takes
// 512 threads (112 against 124 us with 256 at 256x256x300 FP32)
constexpr int kYTileBytes = 32768, kYSegBytes = 128, kYDampNt = 512;
template <typename T, int NY, int MODE>
struct YCfg {
  // 32 KiB of complex values per block; row segments of 128 bytes at most;
  // derivative/filter tiles of 3072 values and more use 512 threads (measured
  // 4.7 vs 3.7 TB/s with 256 on 4096; on 384-point lines 141 vs 190 us, the
  // 256-thread tile at 138 VGPRs)
  static constexpr int LINES = pow2_floor((kYTileBytes / (int)sizeof(C2<T>)) / NY);
  static constexpr int CAP = kYSegBytes / (int)sizeof(C2<T>);  // complex lines per row segment
  static constexpr int NL = LINES < CAP ? LINES : CAP;
  static constexpr int TX = (MODE == YM_DAMP) ? NL : 2 * NL;
  // the damping pass of 384-point lines at 256 threads: 295 against 316 us
  // at 512 and 593 at 1024 (384 x 384 x 300 FP32)
  static constexpr int NT = MODE == YM_DAMP ? (NY == 384 ? 256 : kYDampNt)
                                            : (NL * NY < 3072 ? 256 : 512);
};

template <typename T, int NY, int MODE, int NL = YCfg<T, NY, MODE>::NL,
          int NT = YCfg<T, NY, MODE>::NT>
__global__ __launch_bounds__(NT) void k_ylines_n(int nx, int nz, XArgs<T> X, T beta,
                                                 const T* __restrict__ mult,
                                                 const C2<T>* __restrict__ twg,
                                                 const T* __restrict__ rates) {
  constexpr int TX = (MODE == YM_DAMP) ? NL : 2 * NL;
  constexpr int LS = YLineStride<NY>::LS;
  __shared__ C2<T> b0[NL * LS];
  __shared__ C2<T> tw[NY];
  const int fld = blockIdx.z;
  const T* __restrict__ a = X.a[fld];
  const T* __restrict__ s = X.s[fld];
  T* out = X.out[fld];
  C2<T>* spec = X.spec[fld];
  for (int m = threadIdx.x; m < NY; m += NT) tw[m] = twg[m];
  const long nxy = (long)nx * NY;
  const int i0 = blockIdx.x * TX;
  const long k = blockIdx.y;
  if (MODE == YM_DERIV || MODE == YM_FILTER) {
    T* b0r = reinterpret_cast<T*>(b0);
    // every load issued before the first use (a + beta s, the LDS store)
    constexpr int IT = (TX * NY + NT - 1) / NT;
    T va[IT], vs[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + it * NT;
      if (e < TX * NY) {
        const int y = e / TX, xx = e - y * TX;
        const long o = (i0 + xx) + (long)nx * y + nxy * k;
        va[it] = a[o];
        if (s) vs[it] = s[o];
      }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + it * NT;
      if (e < TX * NY) {
        const int y = e / TX, xx = e - y * TX;
        const T v = s ? va[it] + beta * vs[it] : va[it];
        b0r[2 * ((xx >> 1) * LS + lpad(y)) + (xx & 1)] = v;
      }
    }
    __syncthreads();
    rfft<T, NY, -1, NL, NT, LS>(b0, tw);
    C2<T>* r = b0;
#pragma unroll
    for (int e = threadIdx.x; e < NL * NY; e += NT) {
      const int l = e / NY, y = e - l * NY;
      const T c = mult[y];
      C2<T> v = r[l * LS + lpad(y)];
      if (MODE == YM_DERIV)
        v = {-c * v.y, c * v.x};
      else
        v = {c * v.x, c * v.y};
      r[l * LS + lpad(y)] = v;
    }
    __syncthreads();
    rfft<T, NY, +1, NL, NT, LS>(b0, tw);
    const T* rr = reinterpret_cast<const T*>(b0);
#pragma unroll
    for (int e = threadIdx.x; e < TX * NY; e += NT) {
      const int y = e / TX, xx = e - y * TX;
      const long o = (i0 + xx) + (long)nx * y + nxy * k;
      const T d = rr[2 * ((xx >> 1) * LS + lpad(y)) + (xx & 1)];
      if (MODE == YM_DERIV && X.rhs)
        out[o] = rhs_of(X, k, o, nz, d);
      else
        store_d(X, out, o, d);
    }
  } else {
    const long q = k;
    constexpr int IT = (NL * NY + NT - 1) / NT;
    C2<T> sv[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + it * NT;
      const int y = e / NL, mm = e - y * NL;
      if (e < NL * NY) sv[it] = spec[(q * NY + y) * nx + i0 + mm];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + it * NT;
      const int y = e / NL, mm = e - y * NL;
      if (e < NL * NY) b0[mm * LS + lpad(y)] = sv[it];
    }
    __syncthreads();
    rfft<T, NY, -1, NL, NT, LS>(b0, tw);
    C2<T>* r = b0;
#pragma unroll 4
    for (int e = threadIdx.x; e < NL * NY; e += NT) {
      const int mm = e / NY, y = e - mm * NY;
      const int m = i0 + mm;
      const int mi = (m <= nx / 2) ? m : nx - m;
      const T c = rates[y + (long)NY * mi];
      const int o = mm * LS + lpad(y);
      r[o] = {c * r[o].x, c * r[o].y};
    }
    __syncthreads();
    rfft<T, NY, +1, NL, NT, LS>(b0, tw);
#pragma unroll 4
    for (int e = threadIdx.x; e < NL * NY; e += NT) {
      const int y = e / NL, mm = e - y * NL;
      spec[(q * NY + y) * nx + i0 + mm] = b0[mm * LS + lpad(y)];
    }
  }
}
