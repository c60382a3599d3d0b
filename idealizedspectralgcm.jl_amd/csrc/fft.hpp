// LDS-resident Stockham FFT (radix-4 with one radix-2 stage when log2(L) is
// odd) used by the longitude transforms.  One FFT of length L is done by
// T = L/4 threads that hold four points each per stage; a 256-thread
// workgroup therefore runs 256/T FFTs side by side.
#pragma once

#include <hip/hip_runtime.h>

namespace jgcm {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
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

}  // namespace jgcm
