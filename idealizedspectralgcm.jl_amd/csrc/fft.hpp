// LDS-resident Stockham FFT (radix-4 with one radix-2 stage when log2(L) is
// odd) used by the longitude transforms.  One FFT of length L is done by
// T = L/4 threads that hold four points each per stage; a 256-thread
// workgroup therefore runs 256/T FFTs side by side.
#pragma once

#include <hip/hip_runtime.h>

namespace jgcm {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// Complex product with the fused multiply-adds spelled out, so every kernel
// instantiation rounds it the same way (left to the compiler, the choice of
// the product that is fused can differ between instantiations, and paths that
// must agree bit for bit, such as a batched and a single transform, would not).
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(__fma_rn(a.x, b.x, -(a.y * b.y)), __fma_rn(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// twiddle w^k with w = exp(SIGN * 2 pi i / L); tw[] holds exp(-2 pi i k / L)
template <int SIGN>
__device__ __forceinline__ double2 twid(const double2* tw, int k) {
  double2 t = tw[k];
  return SIGN < 0 ? t : cconj(t);
}

// Barrier among the threads of one FFT group.  When the group fits in one
// wavefront (L <= 256) a wavefront-scope fence suffices: LDS operations of a
// wave are processed in order, the fence makes the compiler wait for them and
// stops it from moving LDS accesses across.
template <int L>
__device__ __forceinline__ void gsync() {
  if constexpr (L / 4 <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <int L, int SIGN>
__device__ __forceinline__ void stage_r4(double2* buf, const double2* tw, int j, int Ns) {
  constexpr int T = L / 4;
  double2 v0 = buf[j], v1 = buf[j + T], v2 = buf[j + 2 * T], v3 = buf[j + 3 * T];
  int k = j & (Ns - 1);
  if (Ns > 1) {
    int step = L / (Ns * 4);
    v1 = cmul(v1, twid<SIGN>(tw, k * step));
    v2 = cmul(v2, twid<SIGN>(tw, 2 * k * step));
    v3 = cmul(v3, twid<SIGN>(tw, 3 * k * step));
  }
  double2 a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = csub(v1, v3);
  double2 y0 = cadd(a, c), y2 = csub(a, c);
  double2 y1, y3;
  if (SIGN < 0) {  // y1 = b - i d, y3 = b + i d
    y1 = make_double2(b.x + d.y, b.y - d.x);
    y3 = make_double2(b.x - d.y, b.y + d.x);
  } else {         // y1 = b + i d, y3 = b - i d
    y1 = make_double2(b.x - d.y, b.y + d.x);
    y3 = make_double2(b.x + d.y, b.y - d.x);
  }
  gsync<L>();
  int o = (j - k) * 4 + k;
  buf[o] = y0;
  buf[o + Ns] = y1;
  buf[o + 2 * Ns] = y2;
  buf[o + 3 * Ns] = y3;
  gsync<L>();
}

template <int L, int SIGN>
__device__ __forceinline__ void stage_r2(double2* buf, const double2* tw, int j, int Ns) {
  constexpr int T = L / 4;
  double2 y[2][2];
  int o[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int q = j + h * T;
    double2 v0 = buf[q], v1 = buf[q + L / 2];
    int k = q & (Ns - 1);
    if (Ns > 1) v1 = cmul(v1, twid<SIGN>(tw, k * (L / (Ns * 2))));
    y[h][0] = cadd(v0, v1);
    y[h][1] = csub(v0, v1);
    o[h] = (q - k) * 2 + k;
  }
  gsync<L>();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    buf[o[h]] = y[h][0];
    buf[o[h] + Ns] = y[h][1];
  }
  gsync<L>();
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Workgroup-wide gather: every thread first issues up to MAXIT independent
// loads, then performs the matching stores, so MAXIT memory requests are in
// flight per thread instead of one load-wait-store at a time.
template <int MAXIT, typename LoadF, typename StoreF>
__device__ __forceinline__ void staged_copy(int total, LoadF load, StoreF store) {
  using T = decltype(load(0));
  const int nt = blockDim.x;
  for (int base = 0; base < total; base += MAXIT * nt) {
    T tmp[MAXIT];
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
      int idx = base + it * nt + (int)threadIdx.x;
      if (idx < total) tmp[it] = load(idx);
    }
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
      int idx = base + it * nt + (int)threadIdx.x;
      if (idx < total) store(idx, tmp[it]);
    }
  }
}

// Full FFT of buf[0..L) (natural order in and out), called by every thread
// of the FFT group (gsync inside); thread j in [0, L/4) of the group works.  Radix sequence: 4, [2], 4, 4, ...
template <int L, int SIGN>
__device__ __forceinline__ void fft_lds(double2* buf, const double2* tw, int j) {
  constexpr int LG = ilog2(L);
  constexpr bool R2 = (LG & 1) != 0;
  constexpr int NR4 = LG / 2;
  int Ns = 1;
  stage_r4<L, SIGN>(buf, tw, j, Ns);
  Ns *= 4;
  if (R2) {
    stage_r2<L, SIGN>(buf, tw, j, Ns);
    Ns *= 2;
  }
#pragma unroll
  for (int s = 1; s < NR4; ++s) {
    stage_r4<L, SIGN>(buf, tw, j, Ns);
    Ns *= 4;
  }
}

// Streaming (non-temporal) stores for outputs written once in whole 128-byte
// lines: they leave the L2 as the kernel runs instead of in the write-back at
// its end, which otherwise delays the next kernel of the stream.
__device__ __forceinline__ void store_stream(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_stream(double2* p, double2 v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}

// ------------------------------------------------ register DFTs (four-step)
// Forward DFTs of 4, 8 and 16 points held in registers, natural order in and
// out: X[k] = sum_n x[n] exp(-2 pi i n k / N).  Constant twiddles are the
// exact values cos(j pi / 8), sin(j pi / 8).
__device__ __forceinline__ double2 mul_negi(double2 a) { return make_double2(a.y, -a.x); }

__device__ __forceinline__ void dft4r(double2& x0, double2& x1, double2& x2, double2& x3) {
  double2 a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = mul_negi(csub(x1, x3));
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, d);
  x3 = csub(b, d);
}

constexpr double COS_PI8 = 0.92387953251128675613;  // cos(pi / 8)
constexpr double SIN_PI8 = 0.38268343236508977173;  // sin(pi / 8)
constexpr double RSQRT2 = 0.70710678118654752440;   // cos(pi / 4)

template <int N>
__device__ __forceinline__ void dft_reg(double2* x) {
  static_assert(N == 4 || N == 8 || N == 16, "4, 8 or 16 points");
  if constexpr (N == 4) {
    dft4r(x[0], x[1], x[2], x[3]);
  } else if constexpr (N == 8) {
    // radix 2 x 4: n = 2a + b; inner 4-point over a, twiddle W8^{bc}, outer 2-point
    dft4r(x[0], x[2], x[4], x[6]);
    dft4r(x[1], x[3], x[5], x[7]);
    x[3] = cmul(x[3], make_double2(RSQRT2, -RSQRT2));
    x[5] = mul_negi(x[5]);
    x[7] = cmul(x[7], make_double2(-RSQRT2, -RSQRT2));
    double2 y[8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      y[c] = cadd(x[2 * c], x[2 * c + 1]);
      y[c + 4] = csub(x[2 * c], x[2 * c + 1]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = y[i];
  } else {
    // radix 4 x 4: n = 4a + b; inner 4-point over a, twiddle W16^{bc}, outer over b
#pragma unroll
    for (int b = 0; b < 4; ++b) dft4r(x[b], x[4 + b], x[8 + b], x[12 + b]);
    const double2 w1 = make_double2(COS_PI8, -SIN_PI8), w2 = make_double2(RSQRT2, -RSQRT2);
    const double2 w3 = make_double2(SIN_PI8, -COS_PI8), w6 = make_double2(-RSQRT2, -RSQRT2);
    const double2 w9 = make_double2(-COS_PI8, SIN_PI8);
    x[5] = cmul(x[5], w1);
    x[9] = cmul(x[9], w2);
    x[13] = cmul(x[13], w3);
    x[6] = cmul(x[6], w2);
    x[10] = mul_negi(x[10]);
    x[14] = cmul(x[14], w6);
    x[7] = cmul(x[7], w3);
    x[11] = cmul(x[11], w6);
    x[15] = cmul(x[15], w9);
#pragma unroll
    for (int c = 0; c < 4; ++c) dft4r(x[4 * c], x[4 * c + 1], x[4 * c + 2], x[4 * c + 3]);
    double2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = x[i];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int d = 0; d < 4; ++d) x[c + 4 * d] = t[4 * c + d];
  }
}

// Four-step FFT of L = 16 P points (P = 4, 8, 16) by the P lanes of one
// group, n = P n1 + n2, k = k1 + 16 k2:
//   step 1: lane n2 holds x[P n1 + n2], n1 = 0..15, and takes the 16-point
//           DFT over n1 (Y[n2][k1]), times W_L^{n2 k1};
//   step 2: through LDS, lane l takes the P-point DFTs over n2 of the 16 / P
//           columns k1 = l + P r, giving X[k1 + 16 k2] in x[r P + k2].
// ex: the group's exchange block, 16 (P + 1) entries.  tw[k] = exp(-2 pi i k / L).
template <int P>
__device__ __forceinline__ void fft_four_step(double2 (&x)[16], double2* ex,
                                              const double2* __restrict__ tw, int l) {
  constexpr int L = 16 * P, EP = P + 1, R = 16 / P;
  double2 w[16];
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) w[k1] = tw[(l * k1) & (L - 1)];
  dft_reg<16>(x);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) x[k1] = cmul(x[k1], w[k1]);
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) ex[k1 * EP + l] = x[k1];
  gsync<64>();  // the group's lanes are in one wavefront
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int n2 = 0; n2 < P; ++n2) x[r * P + n2] = ex[(l + P * r) * EP + n2];
    dft_reg<P>(x + r * P);
  }
}

}  // namespace jgcm
