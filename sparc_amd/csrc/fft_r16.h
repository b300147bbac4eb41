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

// lane u holds row u's 16 entries after the exchange: writes (k, u), reads
// (u, k); real parts first, then imaginary parts through the same tile of T,
// which halves the block's LDS (four blocks of 40 KB filled a CU; this leaves
// room for six)
template <typename T>
__device__ __forceinline__ void exchange16(C2<T> (&v)[16], T* Sl, int u) {
#pragma unroll
  for (int k = 0; k < 16; ++k) Sl[k * R16P + u] = v[k].x;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k].x = Sl[u * R16P + k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) Sl[k * R16P + u] = v[k].y;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k].y = Sl[u * R16P + k];
}

// Buffer view of `bytes` bytes at p, and complex loads and stores through it
// at a per-lane offset plus a wave-uniform one: the row offsets of a y line
// stay in scalar registers instead of one 64-bit address per row.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ybuf(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           0x00020000);
}
typedef unsigned int ybu2 __attribute__((ext_vector_type(2)));
typedef unsigned int ybu4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ C2<T> ybl(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  if constexpr (sizeof(T) == 4) {
    const ybu2 u = __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0);
    return C2<T>{__uint_as_float(u.x), __uint_as_float(u.y)};
  } else {
    const ybu4 u = __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, 0);
    return C2<T>{__hiloint2double((int)u.y, (int)u.x), __hiloint2double((int)u.w, (int)u.z)};
  }
}
template <typename T>
__device__ __forceinline__ void ybs(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so,
                                    C2<T> v) {
  if constexpr (sizeof(T) == 4) {
    const ybu2 u = {__float_as_uint(v.x), __float_as_uint(v.y)};
    __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)vo, (int)so, 0);
  } else {
    const long long x = __double_as_longlong(v.x), y = __double_as_longlong(v.y);
    const ybu4 u = {(unsigned)x, (unsigned)(x >> 32), (unsigned)y, (unsigned)(y >> 32)};
    __builtin_amdgcn_raw_buffer_store_b128(u, r, (int)vo, (int)so, 0);
  }
}
template <typename T>
__device__ __forceinline__ T ybl1(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  if constexpr (sizeof(T) == 4) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
  } else {
    const ybu2 u = __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0);
    return __hiloint2double((int)u.y, (int)u.x);
  }
}

// block table W256^(u*k) at [k*16 + u]
template <typename T>
__device__ __forceinline__ void load_tw2(C2<T>* tw2, const C2<T>* __restrict__ twg) {
  for (int e = threadIdx.x; e < 256; e += blockDim.x) {
    const int k = e >> 4, u = e & 15;
    tw2[e] = twg[(u * k) & 255];
  }
}

// y-lines of length 256 on one plane: 16 complex lines per block of 256
// threads, lane w = threadIdx / 16 of line c = threadIdx % 16, so the 16
// lines' lanes are adjacent in x.  DERIV/FILTER: line c packs columns
// i0 + 2c, i0 + 2c + 1 (i0 = 32 * blockIdx.x).  DAMP: line c is x-bin
// i0 + c (i0 = 16 * blockIdx.x) of a plane pair's spectrum, times the damping
// rates held as rates_o[m + (nx/2+1) * y].  blockIdx.z selects the field.
// line stride of the exchange tile in T: 292 = 4 (mod 32), so both the
// column writes and the row reads of a wave spread over all 32 banks
constexpr int kY256Lsy = 292;

// One block of k_y256: x-chunk bx, plane (pair) k, field fld; S, tw2 and ml
// are the block's LDS tiles.
template <typename T, int MODE>
__device__ __forceinline__ void y256_body(int nx, int nz, const XArgs<T>& X, T beta,
                                          const T* __restrict__ mult,
                                          const C2<T>* __restrict__ twg,
                                          const T* __restrict__ rates_o, T* S, C2<T>* tw2,
                                          T* ml, int bx, long k, int fld) {
  constexpr int NY = 256, LSY = kY256Lsy;
  const T* __restrict__ a = X.a[fld];
  const T* __restrict__ s = X.s[fld];
  T* out = X.out[fld];
  C2<T>* spec = X.spec[fld];
  load_tw2(tw2, twg);
  if (MODE != YM_DAMP)
    for (int e = threadIdx.x; e < NY; e += blockDim.x) ml[e] = mult[e];
  const int c = threadIdx.x & 15, w = threadIdx.x >> 4;
  const long nxy = (long)nx * NY;
  T* Sl = S + c * LSY;
  C2<T> v[16];
  // one plane (pair) per block: buffers over it, lane offset vo, row n at
  // the scalar offset n * rs
  const int i0 = MODE == YM_DAMP ? bx * 16 : bx * 32;
  const unsigned vo = MODE == YM_DAMP ? (unsigned)((w * nx + i0 + c) * sizeof(C2<T>))
                                      : (unsigned)((w * nx + i0 + 2 * c) * sizeof(T));
  const unsigned rs = MODE == YM_DAMP ? (unsigned)(16 * nx * sizeof(C2<T>))
                                      : (unsigned)(16 * nx * sizeof(T));
  const __amdgpu_buffer_rsrc_t rsp =
      MODE == YM_DAMP ? ybuf(spec + k * NY * nx, (unsigned)(NY * nx * sizeof(C2<T>)))
                      : ybuf(out + nxy * k, (unsigned)(nxy * sizeof(T)));
  if (MODE == YM_DAMP) {
#pragma unroll
    for (int n = 0; n < 16; ++n) v[n] = ybl<T>(rsp, vo, n * rs);
  } else {
    // all 16 (or 32) loads are issued before the stage temp a + beta s is
    // formed; combining them inside the loop made every row wait for its
    // own loads
    const __amdgpu_buffer_rsrc_t ra = ybuf(a + nxy * k, (unsigned)(nxy * sizeof(T)));
    if (s) {
      const __amdgpu_buffer_rsrc_t rsc = ybuf(s + nxy * k, (unsigned)(nxy * sizeof(T)));
      C2<T> t[16];
#pragma unroll
      for (int n = 0; n < 16; ++n) {
        v[n] = ybl<T>(ra, vo, n * rs);
        t[n] = ybl<T>(rsc, vo, n * rs);
      }
#pragma unroll
      for (int n = 0; n < 16; ++n) v[n] = {v[n].x + beta * t[n].x, v[n].y + beta * t[n].y};
    } else {
#pragma unroll
      for (int n = 0; n < 16; ++n) v[n] = ybl<T>(ra, vo, n * rs);
    }
  }
  __syncthreads();  // tables
  dft_reg<T, 16, -1>(v);
  twiddle16<T, -1>(v, tw2, w);
  exchange16(v, Sl, w);
  dft_reg<T, 16, -1>(v);  // v[k2] = Y[w + 16 k2]
  if (MODE == YM_DAMP) {
    const int m = bx * 16 + c;
    const int mi = (m <= nx / 2) ? m : nx - m;
    const int nh = nx / 2 + 1;
    const __amdgpu_buffer_rsrc_t rr = ybuf(rates_o, (unsigned)(nh * NY * sizeof(T)));
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const T r = ybl1<T>(rr, (unsigned)((mi + nh * w) * sizeof(T)),
                          (unsigned)(16 * q * nh * sizeof(T)));
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
#pragma unroll
    for (int n = 0; n < 16; ++n) ybs<T>(rsp, vo, n * rs, v[n]);
  } else {
    const long x0 = (long)i0 + 2 * c;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      C2<T> z = v[n];
      if (MODE == YM_DERIV && X.rhs) {
        const long o = x0 + (long)nx * (w + 16 * n) + nxy * k;
        z = {rhs_of(X, k, o, nz, z.x), rhs_of(X, k, o + 1, nz, z.y)};
      } else if (X.acc) {
        const C2<T> u = ybl<T>(rsp, vo, n * rs);
        z = {X.ald * z.x + X.gout * u.x, X.ald * z.y + X.gout * u.y};
      }
      ybs<T>(rsp, vo, n * rs, z);
    }
  }
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_y256(int nx, int nz, XArgs<T> X, T beta,
                                              const T* __restrict__ mult,
                                              const C2<T>* __restrict__ twg,
                                              const T* __restrict__ rates_o) {
  __shared__ T S[16 * kY256Lsy];
  __shared__ C2<T> tw2[256];
  __shared__ T ml[256];
  y256_body<T, MODE>(nx, nz, X, beta, mult, twg, rates_o, S, tw2, ml, blockIdx.x, blockIdx.y,
                     blockIdx.z);
}

// The two horizontal parts of div xi in one launch (displacement.f90:491-
// 513): blocks [0, nbx) the x derivative of X.a[0] (x lines of 256, one
// wavefront per row pair), the rest the y derivative of Y.a[0] (register
// four-step, x-chunks of 32 columns, then planes), each of the stage temp
// a + beta scr; neither pass's tail idles the device.
template <typename T, int LPW>
__global__ __launch_bounds__(256) void k_xy256_div(int nrows, int nlines, XArgs<T> X, XArgs<T> Y,
                                                   T beta, const T* __restrict__ multx,
                                                   const T* __restrict__ multy,
                                                   const C2<T>* __restrict__ twx,
                                                   const C2<T>* __restrict__ twy, int nx, int ny,
                                                   int nz, long nxy, int nbx) {
  __shared__ T S[16 * kY256Lsy];
  __shared__ C2<T> tw2[256];
  __shared__ T ml[256];
  int b = blockIdx.x;
  if (b < nbx) {
    xw256_body<T, XM_DERIV, LPW>(nrows, nlines, X, beta, multx, twx, ny, nz, nxy, 0, b);
  } else {
    b -= nbx;
    const int nbc = nx / 32;
    y256_body<T, YM_DERIV>(nx, nz, Y, beta, multy, twy, (const T*)nullptr, S, tw2, ml, b % nbc,
                           b / nbc, 0);
  }
}

// The damping y pass of three fields (D: nbd blocks, x-chunks of 16 bins
// fastest, then plane pairs, then fields) and the y derivative of one field
// (G, the rest: x-chunks of 32 columns, then planes) in one launch, so the
// blocks of the lighter derivative pass fill the damping pass's tail.
template <typename T>
__global__ __launch_bounds__(256) void k_y256_damp_deriv(int nx, int nz, XArgs<T> D, XArgs<T> G,
                                                         const T* __restrict__ mult,
                                                         const C2<T>* __restrict__ twg,
                                                         const T* __restrict__ rates_o,
                                                         int nbd) {
  __shared__ T S[16 * kY256Lsy];
  __shared__ C2<T> tw2[256];
  __shared__ T ml[256];
  int b = blockIdx.x;
  if (b < nbd) {
    const int nbx = nx / 16, np = (nz + 1) / 2;
    const int bx = b % nbx, t = b / nbx;
    y256_body<T, YM_DAMP>(nx, nz, D, (T)0, mult, twg, rates_o, S, tw2, ml, bx, t % np, t / np);
  } else {
    b -= nbd;
    const int nbx = nx / 32;
    y256_body<T, YM_DERIV>(nx, nz, G, (T)0, mult, twg, rates_o, S, tw2, ml, b % nbx, b / nbx, 0);
  }
}

// ---------------------------------------------------------------------------
// x-derivative lines of length 256, one wavefront per complex line (a row
// pair), no LDS tile.  Lane l holds v[j] = x[l + 64 j], j = 0..3, so every
// load and store is a full 256-byte piece of a row.  With n = n0 + 4 n1 +
// 16 n2 + 64 n3 (lane l = n0 + 4 n1 + 16 n2, register j = n3), the forward
// transform is four radix-4 stages (decimation in frequency):
//   S1  DFT4 over the registers (n3 -> k0), times W256^(l k0)
//   T1  swap the register digit with lane digit n2 (lane bits 4-5)
//   S2  DFT4 (n2 -> k1), times W64^((l & 15) k1)
//   T2  swap with lane digit n1 (bits 2-3)
//   S3  DFT4 (n1 -> k2), times W16^((l & 3) k2)
//   T3  swap with lane digit n0 (bits 0-1)
//   S4  DFT4 (n0 -> k3)
// leaving X[k0 + 4 k1 + 16 k2 + 64 k3] in lane k2 + 4 k1 + 16 k0, register
// k3.  The spectral multiplier is applied there and the inverse runs the
// same steps backwards with conjugate twiddles, ending in natural order.  A
// digit swap is two rounds of pairwise exchanges (lane ^ 2b, then lane ^ b)
// in which each lane sends the two values whose register bit differs from
// its lane bit.
// selects between values (a conditional on array lvalues would select
// between their addresses and keep the array in memory)
template <typename T>
__device__ __forceinline__ C2<T> xw_sel(bool c, C2<T> a, C2<T> b) {
  return C2<T>{c ? a.x : b.x, c ? a.y : b.y};
}

template <typename T, int B>
__device__ __forceinline__ void xw_swap(C2<T> (&v)[4], int lane) {
#pragma unroll
  for (int h = 1; h >= 0; --h) {  // register bit h <-> lane bit log2(B) + h
    const int m = B << h;
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      // registers i0 (bit h clear) and i1 (bit h set), other bit = o
      const int i0 = h ? o : 2 * o, i1 = i0 + (1 << h);
      const C2<T> snd = xw_sel(hi, v[i0], v[i1]);
      const C2<T> rcv = {__shfl_xor(snd.x, m), __shfl_xor(snd.y, m)};
      v[i0] = xw_sel(hi, rcv, v[i0]);
      v[i1] = xw_sel(hi, v[i1], rcv);
    }
  }
}

// The same digit swaps for FP32 without the LDS permute: lane bits 5 and 4
// by v_permlane32_swap / v_permlane16_swap (which swap the upper half of one
// register's lane group with the lower half of the other's: exactly the
// exchange above), lane bits 0-3 by DPP row shifts folded into the select
// (a lane with the bit set reads lane - m, one with it clear lane + m).
template <int CTRL>
__device__ __forceinline__ float xw_dpp(float x) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}

template <int M>
__device__ __forceinline__ void xw_pair_pl(float& a, float& b, bool hi) {
  if constexpr (M == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a),
                                                    __builtin_bit_cast(unsigned, b), false, false);
    a = __builtin_bit_cast(float, (unsigned)r[0]);
    b = __builtin_bit_cast(float, (unsigned)r[1]);
  } else if constexpr (M == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a),
                                                    __builtin_bit_cast(unsigned, b), false, false);
    a = __builtin_bit_cast(float, (unsigned)r[0]);
    b = __builtin_bit_cast(float, (unsigned)r[1]);
  } else {
    static_assert(M == 1 || M == 2 || M == 4 || M == 8, "row shifts within 16 lanes");
    // both shifted reads by every lane, then the selects: a conditional
    // read would run with the partner lanes switched off
    const float sb = xw_dpp<0x110 + M>(b);  // row_shr:M, lane - M
    const float sa = xw_dpp<0x100 + M>(a);  // row_shl:M, lane + M
    const float na = hi ? sb : a;
    const float nb = hi ? b : sa;
    a = na;
    b = nb;
  }
}

template <int B>
__device__ __forceinline__ void xw_swap_pl(C2<float> (&v)[4], int lane) {
#pragma unroll
  for (int h = 1; h >= 0; --h) {
    const bool hi = (lane & (B << h)) != 0;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      const int i0 = h ? o : 2 * o, i1 = i0 + (1 << h);
      if (h) {
        xw_pair_pl<B * 2>(v[i0].x, v[i1].x, hi);
        xw_pair_pl<B * 2>(v[i0].y, v[i1].y, hi);
      } else {
        xw_pair_pl<B>(v[i0].x, v[i1].x, hi);
        xw_pair_pl<B>(v[i0].y, v[i1].y, hi);
      }
    }
  }
}

template <typename T, int B>
__device__ __forceinline__ void xw_swap_any(C2<T> (&v)[4], int lane) {
  // FP32: lane bits 4, 5 by permlane swaps, lane bits 0-3 by DPP; FP64 by
  // the LDS-free shuffles of xw_swap
  constexpr bool PL = sizeof(T) == 4;
  if constexpr (PL)
    xw_swap_pl<B>(reinterpret_cast<C2<float>(&)[4]>(v), lane);
  else
    xw_swap<T, B>(v, lane);
}

template <typename T>
__device__ __forceinline__ void xw_twiddle(C2<T> (&v)[4], const C2<T> (&w)[3], bool conj) {
#pragma unroll
  for (int j = 1; j < 4; ++j) {
    C2<T> t = w[j - 1];
    if (conj) t.y = -t.y;
    v[j] = cmul(v[j], t);
  }
}

// the four forward stages (natural order in, digit-reversed lanes out) and
// the inverse stages (back to natural order), unnormalised
template <typename T>
__device__ __forceinline__ void xw_forward(C2<T> (&v)[4], const C2<T> (&w1)[3],
                                           const C2<T> (&w2)[3], const C2<T> (&w3)[3], int lane) {
  dft_reg<T, 4, -1>(v);
  xw_twiddle(v, w1, false);
  xw_swap_any<T, 16>(v, lane);
  dft_reg<T, 4, -1>(v);
  xw_twiddle(v, w2, false);
  xw_swap_any<T, 4>(v, lane);
  dft_reg<T, 4, -1>(v);
  xw_twiddle(v, w3, false);
  xw_swap_any<T, 1>(v, lane);
  dft_reg<T, 4, -1>(v);
}

template <typename T>
__device__ __forceinline__ void xw_inverse(C2<T> (&v)[4], const C2<T> (&w1)[3],
                                           const C2<T> (&w2)[3], const C2<T> (&w3)[3], int lane) {
  dft_reg<T, 4, +1>(v);
  xw_swap_any<T, 1>(v, lane);
  xw_twiddle(v, w3, true);
  dft_reg<T, 4, +1>(v);
  xw_swap_any<T, 4>(v, lane);
  xw_twiddle(v, w2, true);
  dft_reg<T, 4, +1>(v);
  xw_swap_any<T, 16>(v, lane);
  xw_twiddle(v, w1, true);
  dft_reg<T, 4, +1>(v);
}

// Same contracts as k_xlines_n<T, 256, MODE> for MODE = XM_DERIV (line L =
// rows 2L, 2L+1, x fastest; out = i k mult(k) applied, or the momentum RHS
// when X.rhs), XM_DAMP_FWD (line L = row j of planes 2q, 2q+1, L = j +
// ny q; spec[L N + k] = the forward transform of the stage temp a + beta s,
// which is also stored to X.cp when set, zero on planes 1 and nz) and
// XM_DAMP_INV (the same rows of out, or of X.rin when set, minus the inverse
// transform of spec, into out; planes 1 and nz skipped unless X.edges).
// XM_DERIV_DAMPINV takes the damping lines (row j of planes 2q, 2q+1) of
// the input X.a (no stage temp) and writes -rhoinv d/dx minus the inverse
// transform of spec, which is what the momentum-RHS derivative followed by
// the damping inverse leave there.
// PREP (XM_DERIV_DAMPINV and XM_DAMP_INV): each line then also does the
// next stage's XM_DAMP_FWD from the RHS rows it has just written (X.pa,
// X.pbeta, X.pcp; the spectrum goes over the one the line read).
// LPW lines per wave, the next line's loads issued before the current
// line's transform.
// One 256-thread block of k_xw256: field fld; wave w takes lines
// LPW (4 bx + w) .. LPW (4 bx + w) + LPW - 1.
template <typename T, int MODE, int LPW, bool PREP = false>
__device__ __forceinline__ void xw256_body(int nrows, int nlines, const XArgs<T>& X, T beta,
                                           const T* __restrict__ mult,
                                           const C2<T>* __restrict__ twg, int ny, int nz,
                                           long nxy, int fld, int bx) {
  static_assert(MODE == XM_DERIV || MODE == XM_DAMP_FWD || MODE == XM_DAMP_INV ||
                    MODE == XM_DERIV_DAMPINV,
                "k_xw256 modes");
  static_assert(!PREP || MODE == XM_DAMP_INV || MODE == XM_DERIV_DAMPINV,
                "the next stage is prepared by the damping inverse passes");
  constexpr bool DD = MODE == XM_DERIV_DAMPINV;
  constexpr int N = 256;
  const T* __restrict__ a = X.a[fld];
  const T* __restrict__ s = X.s[fld];
  T* __restrict__ out = X.out[fld];
  C2<T>* __restrict__ spec = X.spec[fld];
  // offset of the line's first row, stride to its second, whether it exists
  auto rows = [&](long L, long& o, long& st, bool& h1) {
    if (MODE == XM_DERIV) {
      o = 2 * L * N;
      st = N;
      h1 = 2 * L + 1 < nrows;
    } else {
      // lines number (nz + 1) / 2 * ny < 2^31: 32-bit division
      const int q = (int)((unsigned)L / (unsigned)ny), j = (int)L - q * ny;
      o = (long)N * j + nxy * (2 * q);
      st = nxy;
      h1 = 2 * q + 1 < nz;
    }
  };
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long L0 = ((long)bx * 4 + wv) * LPW;
  // per-lane twiddles of the three twiddled stages and the multiplier of the
  // four spectral values this lane holds
  C2<T> w1[3], w2[3], w3[3];
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    w1[k - 1] = twg[(lane * k) & 255];
    w2[k - 1] = twg[(4 * (lane & 15) * k) & 255];
    w3[k - 1] = twg[(16 * (lane & 3) * k) & 255];
  }
  const int kl = (lane >> 4) + 4 * ((lane >> 2) & 3) + 16 * (lane & 3);
  T ml[4];
  if (MODE == XM_DERIV || DD) {
#pragma unroll
    for (int j = 0; j < 4; ++j) ml[j] = mult[kl + 64 * j];
  }
  C2<T> nv[4];
  T no0[MODE == XM_DAMP_INV ? 4 : 1], no1[MODE == XM_DAMP_INV ? 4 : 1];
  C2<T> npa[PREP ? 4 : 1];  // PREP: rows of the next stage's a of the line
  // damping inverse: lane l takes X[kl + 64 j], the layout the inverse
  // stages start from, and the RHS values it updates
  auto edge_w = [&](long L, bool& w0, bool& w1) {
    const int k0 = 2 * (int)((unsigned)L / (unsigned)ny), k1 = k0 + 1;
    w0 = X.edges || (k0 != 0 && k0 != nz - 1);
    w1 = k1 < nz - 1 || (X.edges && k1 < nz);
  };
#define XW_LOAD_INV(LL)                                                      \
  {                                                                          \
    long o_, st_;                                                            \
    bool h1_, w0_, w1_;                                                      \
    rows((LL), o_, st_, h1_);                                                \
    edge_w((LL), w0_, w1_);                                                  \
    const T* __restrict__ ri_ = X.rin[fld] ? X.rin[fld] : out;               \
    _Pragma("unroll") for (int j = 0; j < 4; ++j) {                          \
      const int x = lane + 64 * j;                                           \
      nv[j] = spec[(LL) * N + kl + 64 * j];                                  \
      /* PREP: rows the pass leaves alone keep the RHS held in out */        \
      no0[j] = w0_ ? ri_[o_ + x] : (PREP ? out[o_ + x] : (T)0);              \
      no1[j] = w1_ ? ri_[o_ + st_ + x] : ((PREP && h1_) ? out[o_ + st_ + x] : (T)0); \
      if constexpr (PREP) {                                                  \
        npa[j].x = X.pa[fld][o_ + x];                                        \
        npa[j].y = h1_ ? X.pa[fld][o_ + st_ + x] : (T)0;                     \
      }                                                                      \
    }                                                                        \
  }
  // raw a and s values of the next line: the stage temp a + beta s is formed
  // when the line is used, so the loads stay in flight meanwhile
  C2<T> ns[4];
  C2<T> nsp[DD ? 4 : 1];  // DD: the damping spectrum, digit-reversed lanes
#define XW_LOAD(LL)                                                  \
  do {                                                               \
    if constexpr (MODE == XM_DAMP_INV) {                             \
      XW_LOAD_INV(LL)                                                \
    } else {                                                         \
      long o_, st_;                                                \
      bool h1_;                                                    \
      rows((LL), o_, st_, h1_);                                    \
      const T* __restrict__ a0_ = a + o_;                          \
      const T* __restrict__ s0_ = s ? s + o_ : nullptr;            \
      _Pragma("unroll") for (int j = 0; j < 4; ++j) {              \
        const int x = lane + 64 * j;                               \
        nv[j].x = a0_[x];                                          \
        if (s0_) ns[j].x = s0_[x];                                 \
        nv[j].y = h1_ ? a0_[st_ + x] : (T)0;                       \
        if (s0_) ns[j].y = h1_ ? s0_[st_ + x] : (T)0;              \
        if constexpr (DD) nsp[j] = spec[(LL) * N + kl + 64 * j];   \
        if constexpr (PREP) {                                      \
          npa[j].x = X.pa[fld][o_ + x];                            \
          npa[j].y = h1_ ? X.pa[fld][o_ + st_ + x] : (T)0;         \
        }                                                          \
      }                                                            \
    }                                                              \
  } while (0)
  // PREP: the next stage temp pa + pbeta r of line L from the RHS rows r0,
  // r1 just written, as XM_DAMP_FWD forms, stores and transforms it
  auto prep_line = [&](long L, const C2<T> (&pa)[4], const T (&r0)[4], const T (&r1)[4]) {
    long o, st;
    bool h1;
    rows(L, o, st, h1);
    const int k0 = 2 * (int)((unsigned)L / (unsigned)ny);
    const bool e0 = k0 == 0 || k0 == nz - 1, e1 = k0 + 1 == nz - 1;
    T* __restrict__ cp = X.pcp[fld];
    C2<T> u[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      u[j] = C2<T>{pa[j].x + X.pbeta * r0[j], pa[j].y + X.pbeta * r1[j]};
      cp[o + lane + 64 * j] = e0 ? (T)0 : u[j].x;
      if (h1) cp[o + st + lane + 64 * j] = e1 ? (T)0 : u[j].y;
    }
    xw_forward(u, w1, w2, w3, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) spec[L * N + kl + 64 * j] = u[j];
  };
  if (L0 < nlines) XW_LOAD(L0);
  for (int p = 0; p < LPW; ++p) {
    const long L = L0 + p;
    if (L >= nlines) break;
    C2<T> v[4], sp[DD ? 4 : 1], pa[PREP ? 4 : 1];
    T ov0[MODE == XM_DAMP_INV ? 4 : 1], ov1[MODE == XM_DAMP_INV ? 4 : 1];
    T r0a[PREP ? 4 : 1], r1a[PREP ? 4 : 1];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = nv[j];
      if constexpr (PREP) pa[j] = npa[j];
      if constexpr (DD) sp[j] = nsp[j];
      if (MODE != XM_DAMP_INV && s) v[j] = C2<T>{v[j].x + beta * ns[j].x, v[j].y + beta * ns[j].y};
      if (MODE == XM_DAMP_INV) {
        ov0[j] = no0[j];
        ov1[j] = no1[j];
      }
    }
    if (p + 1 < LPW && L + 1 < nlines) XW_LOAD(L + 1);
    if constexpr (MODE == XM_DAMP_FWD) {
      T* __restrict__ cp = X.cp[fld];
      if (cp) {
        long o, st;
        bool h1;
        rows(L, o, st, h1);
        const int k0 = 2 * (int)((unsigned)L / (unsigned)ny);
        const bool e0 = k0 == 0 || k0 == nz - 1, e1 = k0 + 1 == nz - 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          cp[o + lane + 64 * j] = e0 ? (T)0 : v[j].x;
          if (h1) cp[o + st + lane + 64 * j] = e1 ? (T)0 : v[j].y;
        }
      }
    }
    // DD: rows k0, k1 of the pair, which of them take the RHS (not planes 1
    // and nz) and the damping, and rhoinv as a z profile.  When both rows
    // take the damping and rhoinv is a profile (constant along the line),
    // the rows are scaled by -rhoinv before the transform and the damping
    // spectrum is subtracted from the derivative's: one inverse transform.
    long dd_o = 0, dd_st = 0;
    bool dd_h1 = false, dd_w0 = false, dd_w1 = false, dd_e0 = false, dd_e1 = false, one = false;
    T dd_z0 = 0, dd_z1 = 0;
    if constexpr (DD) {
      rows(L, dd_o, dd_st, dd_h1);
      edge_w(L, dd_w0, dd_w1);
      const int k0 = 2 * (int)((unsigned)L / (unsigned)ny), k1 = k0 + 1;
      dd_e0 = k0 == 0 || k0 == nz - 1;
      dd_e1 = k1 == nz - 1;
      dd_z0 = X.r3 ? (T)0 : X.rz[k0];
      dd_z1 = (X.r3 || !dd_h1) ? (T)0 : X.rz[k1];
      one = !X.r3 && dd_w0 && dd_w1;
      if (one) {
        const T s0 = dd_e0 ? (T)0 : -dd_z0, s1 = dd_e1 ? (T)0 : -dd_z1;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = C2<T>{s0 * v[j].x, s1 * v[j].y};
      }
    }
    if (MODE != XM_DAMP_INV) {  // forward transform
      xw_forward(v, w1, w2, w3, lane);
      if (MODE == XM_DAMP_FWD) {
        // natural order: each store fills one 512-byte block, lanes permuted
#pragma unroll
        for (int j = 0; j < 4; ++j) spec[L * N + kl + 64 * j] = v[j];
        continue;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = C2<T>{-ml[j] * v[j].y, ml[j] * v[j].x};
      if constexpr (DD) {
        if (one) {
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] = C2<T>{v[j].x - sp[j].x, v[j].y - sp[j].y};
        }
      }
    }
    xw_inverse(v, w1, w2, w3, lane);
    if constexpr (DD) {
      if (one) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int x = lane + 64 * j;
          out[dd_o + x] = v[j].x;
          if (dd_h1) out[dd_o + dd_st + x] = v[j].y;
          if constexpr (PREP) {
            r0a[j] = v[j].x;
            r1a[j] = dd_h1 ? v[j].y : (T)0;
          }
        }
        if constexpr (PREP) prep_line(L, pa, r0a, r1a);
        continue;
      }
      xw_inverse(sp, w1, w2, w3, lane);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = lane + 64 * j;
        const T ri0 = X.r3 ? X.r3[dd_o + x] : dd_z0;
        T r0 = dd_e0 ? (T)0 : -ri0 * v[j].x;
        if (dd_w0) r0 = r0 - sp[j].x;
        out[dd_o + x] = r0;
        T r1 = (T)0;
        if (dd_h1) {
          const T ri1 = X.r3 ? X.r3[dd_o + dd_st + x] : dd_z1;
          r1 = dd_e1 ? (T)0 : -ri1 * v[j].y;
          if (dd_w1) r1 = r1 - sp[j].y;
          out[dd_o + dd_st + x] = r1;
        }
        if constexpr (PREP) {
          r0a[j] = r0;
          r1a[j] = r1;
        }
      }
      if constexpr (PREP) prep_line(L, pa, r0a, r1a);
      continue;
    }
    if (MODE == XM_DAMP_INV) {
      long o, st;
      bool h1, w0, w1;
      rows(L, o, st, h1);
      edge_w(L, w0, w1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = lane + 64 * j;
        if (w0) out[o + x] = ov0[j] - v[j].x;
        if (w1) out[o + st + x] = ov1[j] - v[j].y;
        if constexpr (PREP) {
          r0a[j] = w0 ? ov0[j] - v[j].x : ov0[j];
          r1a[j] = w1 ? ov1[j] - v[j].y : ov1[j];
        }
      }
      if constexpr (PREP) prep_line(L, pa, r0a, r1a);
      continue;
    }
    const long r0 = 2 * L;
    const bool has1 = r0 + 1 < nrows;
    T* __restrict__ o0 = out + r0 * N;
    if (X.rhs) {
      // -rhoinv * d, zero on planes 1 and nz (rhs_of)
      // rows number ny * nz < 2^31: 32-bit division
      const int k0 = (int)((unsigned)r0 / (unsigned)ny), k1 = (int)((unsigned)(r0 + 1) / (unsigned)ny);
      const bool e0 = k0 == 0 || k0 == nz - 1, e1 = k1 == 0 || k1 == nz - 1;
      const T* __restrict__ q3 = X.r3 ? X.r3 + r0 * N : nullptr;
      const T z0 = X.r3 ? (T)0 : X.rz[k0], z1 = X.r3 ? (T)0 : X.rz[k1];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = lane + 64 * j;
        const T ri0 = q3 ? q3[x] : z0;
        o0[x] = e0 ? (T)0 : -ri0 * v[j].x;
        if (has1) {
          const T ri1 = q3 ? q3[N + x] : z1;
          o0[N + x] = e1 ? (T)0 : -ri1 * v[j].y;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = lane + 64 * j;
        store_d(X, o0, x, v[j].x);
        if (has1) store_d(X, o0, N + x, v[j].y);
      }
    }
  }
#undef XW_LOAD
#undef XW_LOAD_INV
}

template <typename T, int MODE, int LPW>
__global__ __launch_bounds__(256) void k_xw256(int nrows, int nlines, XArgs<T> X, T beta,
                                               const T* __restrict__ mult,
                                               const C2<T>* __restrict__ twg, int ny, int nz,
                                               long nxy) {
  xw256_body<T, MODE, LPW>(nrows, nlines, X, beta, mult, twg, ny, nz, nxy, blockIdx.y, blockIdx.x);
}

// The last two x passes of the option-4 stage in one launch, so the blocks of
// the lighter pass fill the other's tail: blockIdx.y = 0 the x derivative of
// p as the momentum RHS with the damping inverse of v_x (XM_DERIV_DAMPINV on
// H), 1 and 2 the damping inverse of fields 0 and 1 of D (XM_DAMP_INV).
// PREP: each of them also does the next stage's damping forward pass of its
// velocity (H field 0, D fields 0 and 1) from the RHS it has just written.
template <typename T, int LPW, bool PREP>
__global__ __launch_bounds__(256) void k_xw256_rhs_dampinv(int nrows, int nlines, XArgs<T> H,
                                                           XArgs<T> D, const T* __restrict__ mult,
                                                           const C2<T>* __restrict__ twg, int ny,
                                                           int nz, long nxy) {
  if (blockIdx.y == 0)
    xw256_body<T, XM_DERIV_DAMPINV, LPW, PREP>(nrows, nlines, H, (T)0, mult, twg, ny, nz, nxy,
                                               0, blockIdx.x);
  else
    xw256_body<T, XM_DAMP_INV, LPW, PREP>(nrows, nlines, D, (T)0, mult, twg, ny, nz, nxy,
                                          blockIdx.y - 1, blockIdx.x);
}

// DAMP_VELOCITY (damping.f90:125-185) on one 256 x 256 plane per block:
// out -= IDFT2(rates DFT2(in)), unnormalised transforms, rates[m + 129 ky]
// with 1/(nx ny) folded and m = min(kx, 256 - kx) (even in kx and ky).
// 16 wavefronts; wave w keeps row pairs q = 8w .. 8w+7 in registers as
// complex rows z_q = in[2q] + i in[2q+1] (lane l, register j: x = l + 64j)
// and transforms them along x (spectrum in the digit-reversed lanes of
// xw_forward).  The y lines of the half spectrum kx = 0..128 run in two
// chunks of x-bins closed under kx -> 256 - kx ({0..64, 192..255}, then
// {65..191}), staged through an LDS tile [bin][q]: each line unpacks V_y[kx]
// = (z_q[kx] + conj z_q[-kx]) / 2 (y = 2q) or -i (z_q[kx] - conj z_q[-kx]) / 2
// (y = 2q+1), is transformed, scaled, transformed back and repacked into
// both bins (the outputs are real, so bin -kx holds the conjugates).  The
// inverse x transforms then leave D in[2q] + i D in[2q+1].  Planes 1 and nz
// are skipped unless X.edges.
// kx = KL(lane) + 64 j: chunk 0 is registers 0 and 3 and bin 64 (lane 0,
// register 1), chunk 1 registers 1 and 2 without it; with c and j known at
// compile time the other chunk's registers are dead while the tile holds them
__device__ __forceinline__ bool d2_member(int c, int j, int kx) {
  return c == 0 ? (j == 0 || j == 3 || kx == 64) : ((j == 1 || j == 2) && kx != 64);
}
__device__ __forceinline__ int d2_col(int c, int kx) {
  return c == 0 ? (kx <= 64 ? kx : kx - 127) : kx - 65;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_damp2d(XArgs<T> X, const C2<T>* __restrict__ twg,
                                                 const T* __restrict__ rates, int nz, long nxy) {
  constexpr int N = 256, LQ = 129;
  __shared__ C2<T> tile[129 * LQ];
  const int k = blockIdx.x, fld = blockIdx.y;
  if (!X.edges && (k == 0 || k == nz - 1)) return;  // whole block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const T* __restrict__ in = X.a[fld] + nxy * k;
  T* __restrict__ out = X.out[fld] + nxy * k;
  C2<T> w1[3], w2[3], w3[3];
#pragma unroll
  for (int kk = 1; kk < 4; ++kk) {
    w1[kk - 1] = twg[(lane * kk) & 255];
    w2[kk - 1] = twg[(4 * (lane & 15) * kk) & 255];
    w3[kk - 1] = twg[(16 * (lane & 3) * kk) & 255];
  }
  C2<T> z[8][4];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const long o = (long)(2 * (8 * wv + r)) * N + lane;
#pragma unroll
    for (int j = 0; j < 4; ++j) z[r][j] = C2<T>{in[o + 64 * j], in[o + N + 64 * j]};
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) xw_forward(z[r], w1, w2, w3, lane);
  const int kl = (lane >> 4) + 4 * ((lane >> 2) & 3) + 16 * (lane & 3);
  const bool odd = lane & 1;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kx = kl + 64 * j;
      if (d2_member(c, j, kx)) {
        C2<T>* t = tile + d2_col(c, kx) * LQ + 8 * wv;
#pragma unroll
        for (int r = 0; r < 8; ++r) t[r] = z[r][j];
      }
    }
    __syncthreads();
    const int ncol = c ? 64 : 65, kx0 = c ? 65 : 0;
    for (int i = wv; i < ncol; i += 16) {  // wave-uniform
      const int kx = kx0 + i, mk = (N - kx) & (N - 1);
      T rt[4];  // loaded ahead of the transform
#pragma unroll
      for (int j = 0; j < 4; ++j) rt[j] = rates[kx + 129 * (kl + 64 * j)];
      C2<T>* ta = tile + d2_col(c, kx) * LQ;
      C2<T>* tb = tile + d2_col(c, mk) * LQ;
      C2<T> v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = (lane + 64 * j) >> 1;
        const C2<T> A = ta[q], B = tb[q];
        v[j] = odd ? C2<T>{A.y + B.y, B.x - A.x} : C2<T>{A.x + B.x, A.y - B.y};
      }
      xw_forward(v, w1, w2, w3, lane);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const T r = (T)0.5 * rt[j];
        v[j] = C2<T>{r * v[j].x, r * v[j].y};
      }
      xw_inverse(v, w1, w2, w3, lane);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const C2<T> P = {__shfl_xor(v[j].x, 1), __shfl_xor(v[j].y, 1)};
        const C2<T> E = odd ? P : v[j], O = odd ? v[j] : P;
        const int q = (lane + 64 * j) >> 1;
        if (mk == kx) {
          if (!odd) ta[q] = C2<T>{E.x, O.x};
        } else if (!odd) {
          ta[q] = C2<T>{E.x - O.y, E.y + O.x};
        } else {
          tb[q] = C2<T>{E.x + O.y, O.x - E.y};
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kx = kl + 64 * j;
      if (d2_member(c, j, kx)) {
        const C2<T>* t = tile + d2_col(c, kx) * LQ + 8 * wv;
#pragma unroll
        for (int r = 0; r < 8; ++r) z[r][j] = t[r];
      }
    }
    __syncthreads();  // tile reused by the next chunk
  }
  // the RHS rows of the next row pair are loaded before this one's inverse
  const long ob = (long)(2 * 8 * wv) * N + lane;
  T o0[4], o1[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    o0[j] = out[ob + 64 * j];
    o1[j] = out[ob + N + 64 * j];
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const long o = ob + (long)(2 * r) * N;
    T n0[4], n1[4];
    if (r < 7) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        n0[j] = out[o + 2 * N + 64 * j];
        n1[j] = out[o + 3 * N + 64 * j];
      }
    }
    xw_inverse(z[r], w1, w2, w3, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      out[o + 64 * j] = o0[j] - z[r][j].x;
      out[o + N + 64 * j] = o1[j] - z[r][j].y;
    }
    if (r < 7) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        o0[j] = n0[j];
        o1[j] = n1[j];
      }
    }
  }
}
