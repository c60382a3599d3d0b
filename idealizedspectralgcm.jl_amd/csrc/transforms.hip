// Spherical-harmonic transforms on gfx950.
//
// Forward  (Trans_Grid_To_Spherical!, Spectral_Spherical_Mesh.jl:221-293):
//   1. fft_fold_kernel: each workgroup takes `ppw` southern latitudes j' of
//      one level-field and their mirrors nθ-1-j'.  Row j' and its mirror go
//      into ONE complex FFT (z = a + i c); the wavenumbers m = 0..N of the
//      equatorially even / odd combinations E = g_j' + g_mirror and
//      O = g_j' - g_mirror come straight out of that spectrum.  Output: the
//      E/O block [field][m][parity][j'] (complex, FP64).
//   2. legendre_fwd_mfma: per (m, parity) segment, rows n = m+p, m+p+2, ..
//      S[n] = sum_j' P~[n][j'] X[j'] with the quadrature weight and the 1/2
//      folded into P~, as 16 x 16 MFMA tiles (rows x 16 real columns).
// Inverse (Trans_Spherical_To_Grid!, Spectral_Spherical_Mesh.jl:143-217):
//   1. legendre_inv_mfma: per (m, parity) segment, E/O[j'] = sum_n P[n][j'] S[n]
//      over degrees m..N+1, as MFMA tiles (latitudes x columns).
//   2. ifft_unfold_kernel: rebuilds g_j' = E+O, g_mirror = E-O, packs both
//      rows into one complex spectrum and runs one inverse FFT per pair.
// vor/div (Vor_Div_From_Grid_UV!, :349-393) reuse both stages: the FFT stage
// multiplies by cos θ on load, the Legendre stage contracts two tables
// (coef_alp_a against i*X^p and coef_alp_b against X^(1-p), the parity of
// dP/dμ being opposite to that of P).
//
// Precision variants of the Legendre stage (the FFT stage is always FP64):
//   FP64    : FP64 tables, FP64 MFMA.
//   FP32    : FP32 tables and operands, FP32 MFMA with FP32 accumulation.
//   SPLIT32 : x = x_hi + x_lo, P = P_hi + P_lo (FP32 pieces); per block of 16
//             contraction steps the products P_hi x_hi and the cross terms
//             P_hi x_lo + P_lo x_hi are summed by FP32 MFMA, the block sums
//             are accumulated in FP64 (compensated summation).
#include <algorithm>
#include <mutex>
#include <queue>
#include <type_traits>
#include <vector>

#include "fft.hpp"
#include "jgcm_internal.hpp"

namespace jgcm {

enum { P_FP64 = 0, P_FP32 = 1, P_SPLIT = 2, P_BF16 = 3 };

// A launch may gather its level-fields from up to MAXSEG separate arrays
// (e.g. vor, div, u cos, v cos, T and ln ps back-transformed together):
// global field b lives in segment k with start[k] <= b < start[k+1], at
// ptr[k] + (b - start[k]) * per-field size.  flag[k] bit 0 selects the cos θ
// multiply (forward FFT) or divide (inverse FFT) for that segment; for the
// inverse Legendre stage bits 1..3 name a spectral operator applied while the
// coefficients are staged (SPEC_OP_*), ptr2[k] being its second operand.
constexpr int MAXSEG = 12;
struct SegMap {
  int nseg;
  int start[MAXSEG + 1];
  double* ptr[MAXSEG];
  const double* ptr2[MAXSEG];
  int flag[MAXSEG];
};

// Spectral operators fused into the inverse transform's staging (flag bits 1..3):
enum {
  SPEC_OP_NONE = 0,
  SPEC_OP_GRAD_LAMBDA = 1,  // cos θ ∂λ h / a      (Compute_Gradient_Cos!, SSM:677-680)
  SPEC_OP_GRAD_THETA = 2,   // cos θ ∂θ h / a      (SSM:682-687)
  SPEC_OP_UCOS = 3,         // u cos θ from (vor, div)  (Compute_Ucos_Vcos_From_Vor_Div!, SSM:568-576)
  SPEC_OP_VCOS = 4,         // v cos θ from (vor, div)
  SPEC_OP_NEG_I = 5,        // -i h  (transposes of the α operator, SSM:449-456)
  SPEC_OP_NEG = 6,          // -h
};

// coefficient tables [m][ns+1] of those operators
struct SpecOpTables {
  const double* dlam;
  const double* dthm;
  const double* dthp;
  const double* uvc;
  const double* uvm;
  const double* uvp;
};

__host__ __device__ __forceinline__ int seg_find(const SegMap& s, int b) {
  int k = 0;
#pragma unroll
  for (int q = 1; q < MAXSEG; ++q)
    if (q < s.nseg && b >= s.start[q]) k = q;
  return k;
}

__host__ __device__ __forceinline__ double* seg_field(const SegMap& s, int b, size_t elems) {
  int k = seg_find(s, b);
  return s.ptr[k] + (size_t)(b - s.start[k]) * elems;
}

// First packed table row of wavenumber m (rows n = m..ns per m, m = 0..N):
// sum_{k<m} (ns + 1 - k).  Computed rather than loaded, so the table
// addresses of a Legendre workgroup do not wait on a memory round trip.
__device__ __forceinline__ int packed_row_start(int m, int ns) {
  return m * (ns + 1) - (m * (m - 1)) / 2;
}

static inline SegMap one_seg(const void* p, int nlev, int flag) {
  SegMap s{};
  s.nseg = 1;
  s.start[0] = 0;
  s.start[1] = nlev;
  s.ptr[0] = const_cast<double*>(reinterpret_cast<const double*>(p));
  s.flag[0] = flag;
  return s;
}

// ---------------------------------------------------------------- FFT stage
// PLAIN (all kernels of this stage): one array with no cos θ flag, so the
// segment lookup and the per-element scaling branches compile away.
template <int L, bool PLAIN>
__global__ __launch_bounds__(256) void fft_fold_kernel(
    SegMap in,                         // grid fields [nlat][L] each
    double2* __restrict__ eo,          // [B][nm][2][half]
    const double2* __restrict__ twg,   // [L]
    const double* __restrict__ cosv,   // [nlat]: rows multiplied by cos θ where flag & 1
    int nlat, int nm, int ppw) {
  constexpr int T = L / 4;
  constexpr int FPW = 256 / T;
  __shared__ double2 buf[FPW][L];
  __shared__ double2 tw[L];
  extern __shared__ double2 stage[];  // [ppw][2 nm + 1]: (m, parity) per latitude, padded
  const int half = nlat / 2;
  const int SP = 2 * nm + 1;
  const int b = blockIdx.y;
  const int j0 = blockIdx.x * ppw;
  const int seg = PLAIN ? 0 : seg_find(in, b);
  const double* __restrict__ grid = in.ptr[seg] + (size_t)(b - in.start[seg]) * nlat * L;
  // flag bit 0: rows scaled by cos θ; with bit 1 divided by it instead
  const double* __restrict__ cosmul = (!PLAIN && (in.flag[seg] & 1)) ? cosv : nullptr;
  const bool divcos = !PLAIN && (in.flag[seg] & 2) != 0;
  const int g = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const double scale = 1.0 / L;
  // twiddles and all passes' rows are loaded up front (one round trip, up to
  // PASSES x 8 + TWI loads in flight per thread)
  constexpr int TWI = (L + 255) / 256;
  double2 twr[TWI];
#pragma unroll
  for (int q = 0; q < TWI; ++q)
    if (q * 256 + (int)threadIdx.x < L) twr[q] = twg[q * 256 + threadIdx.x];
  constexpr int PASSES = (L == 256) ? 2 : ((L == 512) ? 4 : ((L == 1024) ? 8 : 1));
  double2 pre[PASSES][4];
#pragma unroll
  for (int q = 0; q < PASSES; ++q) {
    int jj = q * FPW + g;
    int jp = j0 + jj;
    if (q * FPW < ppw && jj < ppw && jp < half) {
      const double* ra = grid + (size_t)jp * L;
      const double* rc = grid + (size_t)(nlat - 1 - jp) * L;
#pragma unroll
      for (int r = 0; r < 4; ++r) pre[q][r] = make_double2(ra[t + r * T], rc[t + r * T]);
    }
  }
#pragma unroll
  for (int q = 0; q < TWI; ++q)
    if (q * 256 + (int)threadIdx.x < L) tw[q * 256 + threadIdx.x] = twr[q];
  __syncthreads();
  for (int pass = 0; pass < ppw; pass += FPW) {
    int jj = pass + g;
    int jp = j0 + jj;
    bool valid = (jj < ppw) && (jp < half);
    if (valid) {
      double s = cosmul ? cosmul[jp] : 1.0;
      double2 v[4];
      if (pass / FPW < PASSES) {
#pragma unroll
        for (int q = 0; q < PASSES; ++q)
          if (q == pass / FPW) {
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = pre[q][r];
          }
      } else {
        const double* ra = grid + (size_t)jp * L;
        const double* rc = grid + (size_t)(nlat - 1 - jp) * L;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = make_double2(ra[t + r * T], rc[t + r * T]);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double a = v[r].x, c = v[r].y;
        if (cosmul) {
          if (divcos) {
            a = a / s;
            c = c / s;
          } else {
            a = a * s;
            c = c * s;
          }
        }
        buf[g][t + r * T] = make_double2(a, c);
      }
    }
    gsync<L>();
    fft_lds<L, -1>(buf[g], tw, t);
    if (valid) {
      for (int m = t; m < nm; m += T) {
        double2 zm = buf[g][m];
        double2 zc = cconj(buf[g][(L - m) & (L - 1)]);
        double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        double2 d = make_double2(0.5 * (zm.x - zc.x), 0.5 * (zm.y - zc.y));
        double2 C = make_double2(d.y, -d.x);  // (zm - zc) / (2i)
        stage[jj * SP + m * 2 + 0] = make_double2((A.x + C.x) * scale, (A.y + C.y) * scale);
        stage[jj * SP + m * 2 + 1] = make_double2((A.x - C.x) * scale, (A.y - C.y) * scale);
      }
    }
    gsync<L>();
  }
  __syncthreads();
  const int total = nm * 2 * ppw;
  const bool whole_lines = ppw >= 8;  // (m, parity) rows of >= 128 bytes
  for (int idx = threadIdx.x; idx < total; idx += 256) {
    int jj = idx % ppw;
    int mp = idx / ppw;
    int jp = j0 + jj;
    if (jp < half) {
      double2* o = eo + ((size_t)b * nm * 2 + mp) * half + jp;
      if (whole_lines) store_stream(o, stage[jj * SP + mp]);
      else *o = stage[jj * SP + mp];
    }
  }
}

// Inverse: E/O block -> two grid rows per FFT.  Segments with flag & 1 have
// their output rows divided by cos θ (Divide_By_Cos!, Spectral_Spherical_Mesh.jl:333-338);
// with flag & 2 the E/O values are first weighted 1/L (m = 0) and 1/(2L)
// (m > 0), the transposed Hermitian weights of the forward FFT; flag & 4
// (with & 1) multiplies the rows by cos θ instead of dividing.
template <int L, bool PLAIN>
__global__ __launch_bounds__(256) void ifft_unfold_kernel(
    const double2* __restrict__ eo,    // [B][nm][2][half]
    SegMap out,                        // grid fields [nlat][L] each
    const double2* __restrict__ twg, const double* __restrict__ cosv,
    int nlat, int nm, int ppw) {
  constexpr int T = L / 4;
  constexpr int FPW = 256 / T;
  __shared__ double2 buf[FPW][L];
  __shared__ double2 tw[L];
  extern __shared__ double2 stage[];  // [ppw][2 nm + 1]
  const int half = nlat / 2;
  const int SP = 2 * nm + 1;
  const int b = blockIdx.y;
  const int j0 = blockIdx.x * ppw;
  const int seg = PLAIN ? 0 : seg_find(out, b);
  double* __restrict__ grid = out.ptr[seg] + (size_t)(b - out.start[seg]) * nlat * L;
  const double* __restrict__ divcos = (!PLAIN && (out.flag[seg] & 1)) ? cosv : nullptr;
  const bool adjw = !PLAIN && (out.flag[seg] & 2) != 0;
  const bool mulcos = !PLAIN && (out.flag[seg] & 4) != 0;
  const double w0 = 1.0 / L, w1 = 1.0 / (2.0 * L);
  // twiddle loads in flight together with the E/O gather below
  constexpr int TWI = (L + 255) / 256;
  double2 twr[TWI];
#pragma unroll
  for (int q = 0; q < TWI; ++q)
    if (q * 256 + (int)threadIdx.x < L) twr[q] = twg[q * 256 + threadIdx.x];
  const int total = nm * 2 * ppw;
  staged_copy<8>(
      total,
      [&](int idx) {
        int jj = idx % ppw;
        int mp = idx / ppw;
        int jp = j0 + jj;
        return (jp < half) ? eo[((size_t)b * nm * 2 + mp) * half + jp] : make_double2(0.0, 0.0);
      },
      [&](int idx, double2 v) {
        if (adjw) {
          const double s = (idx / ppw) < 2 ? w0 : w1;  // m = (idx / ppw) / 2
          v = make_double2(v.x * s, v.y * s);
        }
        stage[(idx % ppw) * SP + idx / ppw] = v;
      });
#pragma unroll
  for (int q = 0; q < TWI; ++q)
    if (q * 256 + (int)threadIdx.x < L) tw[q * 256 + threadIdx.x] = twr[q];
  __syncthreads();
  const int g = threadIdx.x / T;
  const int t = threadIdx.x % T;
  for (int pass = 0; pass < ppw; pass += FPW) {
    int jj = pass + g;
    int jp = j0 + jj;
    bool valid = (jj < ppw) && (jp < half);
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[g][t + r * T] = make_double2(0.0, 0.0);
    gsync<L>();
    if (valid) {
      for (int m = t; m < nm; m += T) {
        double2 E = stage[jj * SP + m * 2 + 0];
        double2 O = stage[jj * SP + m * 2 + 1];
        double2 ga = cadd(E, O), gc = csub(E, O);
        if (m == 0) {
          // g_0 halved, then 2 Re(.) -> Re(g_0) for each row (SSM:211,214)
          buf[g][0] = make_double2(ga.x, gc.x);
        } else {
          buf[g][m] = make_double2(ga.x - gc.y, ga.y + gc.x);          // ga + i gc
          buf[g][L - m] = make_double2(ga.x + gc.y, gc.x - ga.y);      // conj(ga) + i conj(gc)
        }
      }
    }
    gsync<L>();
    fft_lds<L, +1>(buf[g], tw, t);
    if (valid) {
      int jm = nlat - 1 - jp;
      double* ra = grid + (size_t)jp * L;
      double* rc = grid + (size_t)jm * L;
      double c = divcos ? divcos[jp] : 1.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int i = t + r * T;
        double2 z = buf[g][i];
        if (divcos && mulcos) {
          store_stream(ra + i, z.x * c);
          store_stream(rc + i, z.y * c);
        } else if (divcos) {
          store_stream(ra + i, z.x / c);
          store_stream(rc + i, z.y / c);
        } else {
          store_stream(ra + i, z.x);
          store_stream(rc + i, z.y);
        }
      }
    }
    gsync<L>();
  }
}

// Full Fourier output (Trans_Grid_To_Fourier!, SSM:295-330): four[b][j][k] =
// fft(row)/nλ for all k, two rows per complex FFT.
template <int L>
__global__ __launch_bounds__(256) void fft_full_kernel(const double* __restrict__ grid,
                                                       double2* __restrict__ four,
                                                       const double2* __restrict__ twg,
                                                       int nlat) {
  constexpr int T = L / 4;
  constexpr int FPW = 256 / T;
  __shared__ double2 buf[FPW][L];
  __shared__ double2 tw[L];
  const int half = nlat / 2;
  const int b = blockIdx.y;
  for (int i = threadIdx.x; i < L; i += 256) tw[i] = twg[i];
  __syncthreads();
  const int g = threadIdx.x / T;
  const int t = threadIdx.x % T;
  int jp = blockIdx.x * FPW + g;
  bool valid = jp < half;
  int jm = nlat - 1 - jp;
  if (valid) {
    for (int r = 0; r < 4; ++r) {
      int i = t + r * T;
      buf[g][i] = make_double2(grid[((size_t)b * nlat + jp) * L + i], grid[((size_t)b * nlat + jm) * L + i]);
    }
  }
  __syncthreads();
  fft_lds<L, -1>(buf[g], tw, t);
  if (valid) {
    const double scale = 1.0 / L;
    for (int r = 0; r < 4; ++r) {
      int k = t + r * T;
      double2 zm = buf[g][k];
      double2 zc = cconj(buf[g][(L - k) & (L - 1)]);
      double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 d = make_double2(0.5 * (zm.x - zc.x), 0.5 * (zm.y - zc.y));
      double2 C = make_double2(d.y, -d.x);
      four[((size_t)b * nlat + jp) * L + k] = make_double2(A.x * scale, A.y * scale);
      four[((size_t)b * nlat + jm) * L + k] = make_double2(C.x * scale, C.y * scale);
    }
  }
}

// ------------------------------------------ four-step FFT kernels (L = 128, 256)
// The same fold / unfold as the radix-4 kernels above for nλ = 16 P, P = 4, 8
// or 16: each latitude pair is one group of P lanes holding 16 points each in
// registers (fft_four_step), so a transform costs two register DFTs and one
// LDS exchange instead of log4(L) LDS passes.  G latitude pairs per
// workgroup (P G threads, 64 <= P G <= 256).  Per group, LDS holds the
// exchange block (16 (P + 1) entries, reused for the spectrum) and the
// (m, parity) staging row of 2 nm + 1 entries.
template <int P, int G, bool PLAIN>
__global__ __launch_bounds__(P * G) void fold_four_step_kernel(
    SegMap in, double2* __restrict__ eo, const double2* __restrict__ twg,
    const double* __restrict__ cosv, int nlat, int nm) {
  constexpr int L = 16 * P, EX = 16 * (P + 1), R = 16 / P;
  __shared__ double2 ex[G][EX];
  extern __shared__ double2 stage[];  // [2 nm][G]: (m, parity) rows of the G pairs
  const int half = nlat / 2;
  const int b = blockIdx.y;
  const int g = threadIdx.x / P, l = threadIdx.x % P;
  const int j0 = blockIdx.x * G;
  const int jp = j0 + g;
  const bool valid = jp < half;
  const int seg = PLAIN ? 0 : seg_find(in, b);
  const double* __restrict__ grid = in.ptr[seg] + (size_t)(b - in.start[seg]) * nlat * L;
  const int jr = valid ? jp : 0;
  const double* ra = grid + (size_t)jr * L;
  const double* rc = grid + (size_t)(nlat - 1 - jr) * L;
  double2 x[16];
#pragma unroll
  for (int n1 = 0; n1 < 16; ++n1) x[n1] = make_double2(ra[P * n1 + l], rc[P * n1 + l]);
  if (!PLAIN && (in.flag[seg] & 1)) {  // rows times cos θ, or with bit 1 divided by it
    const double c = cosv[jr];
    const bool div = (in.flag[seg] & 2) != 0;
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1)
      x[n1] = div ? make_double2(x[n1].x / c, x[n1].y / c) : make_double2(x[n1].x * c, x[n1].y * c);
  }
  fft_four_step<P>(x, ex[g], twg, l);
  // x[r P + k2] = Z[k1 + 16 k2], k1 = l + P r.  Z_{L-m} for m = k1 + 16 k2:
  // k1 != 0: lane (P - l) % P, register (R - 1 - r) P + (P - 1 - k2) when
  // l != 0; for l = 0 the lane's own register (R - r) % R, (P - k2) % P when
  // r = 0, else R - r, P - 1 - k2.
  // z = a + i c: A = (Z_m + conj Z_{L-m}) / 2, C = (Z_m - conj Z_{L-m}) / (2i);
  // E = (A + C) / L, O = (A - C) / L
  const int src = (threadIdx.x & 63) - l + ((P - l) & (P - 1));
  const double scale = 1.0 / L;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k2 = 0; k2 < P; ++k2) {
      if (P * r + 16 * k2 >= nm) break;  // no lane has m < nm (wave-uniform)
      const int m = l + P * r + 16 * k2;
      const double2 q = x[(R - 1 - r) * P + (P - 1 - k2)];
      double2 zc = make_double2(__shfl(q.x, src, 64), __shfl(q.y, src, 64));
      if (l == 0) zc = r == 0 ? x[(P - k2) & (P - 1)] : x[(R - r) * P + (P - 1 - k2)];
      zc.y = -zc.y;
      const double2 zm = x[r * P + k2];
      const double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      const double2 d = make_double2(0.5 * (zm.x - zc.x), 0.5 * (zm.y - zc.y));
      const double2 C = make_double2(d.y, -d.x);
      if (m < nm) {
        stage[(2 * m + 0) * G + g] = make_double2((A.x + C.x) * scale, (A.y + C.y) * scale);
        stage[(2 * m + 1) * G + g] = make_double2((A.x - C.x) * scale, (A.y - C.y) * scale);
      }
    }
  __syncthreads();
  // G >= 8: each (m, parity) row of the workgroup is a whole 128-byte line
  const int total = 2 * nm * G;
  for (int idx = threadIdx.x; idx < total; idx += P * G) {
    const int jj = idx % G, mp = idx / G;
    if (j0 + jj < half) {
      double2* o = eo + ((size_t)b * nm * 2 + mp) * half + j0 + jj;
      if constexpr (G >= 8) store_stream(o, stage[idx]);
      else *o = stage[idx];
    }
  }
}

template <int P, int G, bool PLAIN>
__global__ __launch_bounds__(P * G) void unfold_four_step_kernel(
    const double2* __restrict__ eo, SegMap out, const double2* __restrict__ twg,
    const double* __restrict__ cosv, int nlat, int nm) {
  constexpr int L = 16 * P, EX = 16 * (P + 1);
  __shared__ double2 ex[G][EX];
  extern __shared__ double2 stage[];  // [G][2 nm + 1]
  const int half = nlat / 2;
  const int SP = 2 * nm + 1;
  const int b = blockIdx.y;
  const int g = threadIdx.x / P, l = threadIdx.x % P;
  const int j0 = blockIdx.x * G;
  const int jp = j0 + g;
  const int seg = PLAIN ? 0 : seg_find(out, b);
  double* __restrict__ grid = out.ptr[seg] + (size_t)(b - out.start[seg]) * nlat * L;
  const int flag = PLAIN ? 0 : out.flag[seg];
  const double w0 = 1.0 / L, w1 = 1.0 / (2.0 * L);
  const int total = 2 * nm * G;
  staged_copy<8>(
      total,
      [&](int idx) {
        const int jj = idx % G, mp = idx / G;
        return (j0 + jj < half) ? eo[((size_t)b * nm * 2 + mp) * half + j0 + jj]
                                : make_double2(0.0, 0.0);
      },
      [&](int idx, double2 v) {
        const int jj = idx % G, mp = idx / G;
        if (flag & 2) {  // transposed Hermitian weights of the forward FFT
          const double s = mp < 2 ? w0 : w1;
          v = make_double2(v.x * s, v.y * s);
        }
        stage[jj * SP + mp] = v;
      });
  __syncthreads();
  // spectrum Z_m = g_a + i g_c (m >= 1) with Z_{L-m} = conj g_a + i conj g_c
  // and Z_0 = (Re g_a, Re g_c), g_a = E + O, g_c = E - O; the inverse FFT is
  // the conjugate of the forward four-step FFT of conj Z
  const double2* st = stage + g * SP;
  double2 x[16];
#pragma unroll
  for (int n1 = 0; n1 < 16; ++n1) {
    const int m = P * n1 + l;
    double2 z = make_double2(0.0, 0.0);
    if (m == 0) {
      const double2 E = st[0], O = st[1];
      z = make_double2(E.x + O.x, E.x - O.x);
    } else if (m < nm) {
      const double2 ga = cadd(st[2 * m], st[2 * m + 1]), gc = csub(st[2 * m], st[2 * m + 1]);
      z = make_double2(ga.x - gc.y, ga.y + gc.x);
    } else if (L - m < nm) {
      const int q = L - m;
      const double2 ga = cadd(st[2 * q], st[2 * q + 1]), gc = csub(st[2 * q], st[2 * q + 1]);
      z = make_double2(ga.x + gc.y, gc.x - ga.y);
    }
    x[n1] = cconj(z);
  }
  fft_four_step<P>(x, ex[g], twg, l);
  if (jp < half) {
    double* ra = grid + (size_t)jp * L;
    double* rc = grid + (size_t)(nlat - 1 - jp) * L;
    const double c = (flag & 1) ? cosv[jp] : 1.0;
#pragma unroll
    for (int r = 0; r < 16 / P; ++r)
#pragma unroll
      for (int k2 = 0; k2 < P; ++k2) {
        const int i = l + P * r + 16 * k2;
        double va = x[r * P + k2].x, vc = -x[r * P + k2].y;
        if (flag & 1) {  // divided by cos θ, or with bit 2 multiplied
          if (flag & 4) {
            va *= c;
            vc *= c;
          } else {
            va /= c;
            vc /= c;
          }
        }
        store_stream(ra + i, va);
        store_stream(rc + i, vc);
      }
  }
}

// ------------------------------------------------- general-length longitude DFT
// For nλ that is not a power of two (the reference's FFTW takes any length):
// direct sums over the twiddle table tw[k] = exp(-2πi k/nλ), indexed by
// (m λ) mod nλ, one workgroup per latitude pair (fold, unfold) or row (full)
// and field.  Same inputs, outputs, scaling and cos θ options as the radix-4
// kernels above; O(nλ (N+1)) per row pair.
__global__ __launch_bounds__(256) void dft_fold_kernel(SegMap in, double2* __restrict__ eo,
                                                       const double2* __restrict__ twg,
                                                       const double* __restrict__ cosv, int nlat,
                                                       int L, int nm) {
  extern __shared__ double2 dsh[];  // tw[L] | row a, row c (L doubles each)
  double2* tw = dsh;
  double* ra = reinterpret_cast<double*>(dsh + L);
  double* rc = ra + L;
  const int half = nlat / 2;
  const int jp = blockIdx.x, b = blockIdx.y;
  const int seg = seg_find(in, b);
  const double* __restrict__ grid = in.ptr[seg] + (size_t)(b - in.start[seg]) * nlat * L;
  const double s = (in.flag[seg] & 1) ? cosv[jp] : 1.0;  // cos θ is symmetric about the equator
  const int cmode = (in.flag[seg] & 1) ? ((in.flag[seg] & 2) ? 2 : 1) : 0;  // x cos, / cos
  const double* a = grid + (size_t)jp * L;
  const double* c = grid + (size_t)(nlat - 1 - jp) * L;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    tw[i] = twg[i];
    ra[i] = cmode == 1 ? a[i] * s : (cmode == 2 ? a[i] / s : a[i]);
    rc[i] = cmode == 1 ? c[i] * s : (cmode == 2 ? c[i] / s : c[i]);
  }
  __syncthreads();
  const double scale = 1.0 / L;
  for (int m = threadIdx.x; m < nm; m += blockDim.x) {
    double2 xa = make_double2(0.0, 0.0), xc = make_double2(0.0, 0.0);
    int k = 0;
    for (int l = 0; l < L; ++l) {
      const double2 w = tw[k];
      xa = make_double2(xa.x + ra[l] * w.x, xa.y + ra[l] * w.y);
      xc = make_double2(xc.x + rc[l] * w.x, xc.y + rc[l] * w.y);
      k += m;
      if (k >= L) k -= L;
    }
    eo[(((size_t)b * nm + m) * 2 + 0) * half + jp] =
        make_double2((xa.x + xc.x) * scale, (xa.y + xc.y) * scale);
    eo[(((size_t)b * nm + m) * 2 + 1) * half + jp] =
        make_double2((xa.x - xc.x) * scale, (xa.y - xc.y) * scale);
  }
}

__global__ __launch_bounds__(256) void dft_unfold_kernel(const double2* __restrict__ eo,
                                                         SegMap out,
                                                         const double2* __restrict__ twg,
                                                         const double* __restrict__ cosv,
                                                         int nlat, int L, int nm) {
  extern __shared__ double2 dsh[];  // tw[L] | ga[nm] | gc[nm]
  double2* tw = dsh;
  double2* ga = dsh + L;
  double2* gc = ga + nm;
  const int half = nlat / 2;
  const int jp = blockIdx.x, b = blockIdx.y;
  const int seg = seg_find(out, b);
  double* __restrict__ grid = out.ptr[seg] + (size_t)(b - out.start[seg]) * nlat * L;
  const bool divcos = out.flag[seg] & 1;
  const bool adjw = (out.flag[seg] & 2) != 0;
  for (int i = threadIdx.x; i < L; i += blockDim.x) tw[i] = twg[i];
  for (int m = threadIdx.x; m < nm; m += blockDim.x) {
    double2 E = eo[(((size_t)b * nm + m) * 2 + 0) * half + jp];
    double2 O = eo[(((size_t)b * nm + m) * 2 + 1) * half + jp];
    if (adjw) {  // transposed Hermitian weights, as in ifft_unfold_kernel
      const double s = m == 0 ? 1.0 / L : 1.0 / (2.0 * L);
      E = make_double2(E.x * s, E.y * s);
      O = make_double2(O.x * s, O.y * s);
    }
    ga[m] = cadd(E, O);
    gc[m] = csub(E, O);
  }
  __syncthreads();
  const double c = divcos ? cosv[jp] : 1.0;
  for (int l = threadIdx.x; l < L; l += blockDim.x) {
    // Re(g_0) + 2 Σ_m Re(g_m e^{+2πi m l / L})  (the reference's g₀/2 and 2 Re, SSM:211-215)
    double xa = ga[0].x, xc = gc[0].x;
    int k = l;
    for (int m = 1; m < nm; ++m) {
      const double2 w = tw[k];  // e^{-iθ}: Re(g e^{iθ}) = g.x w.x + g.y w.y
      xa = xa + 2.0 * (ga[m].x * w.x + ga[m].y * w.y);
      xc = xc + 2.0 * (gc[m].x * w.x + gc[m].y * w.y);
      k += l;
      if (k >= L) k -= L;
    }
    if (divcos && (out.flag[seg] & 4)) {
      xa = xa * c;
      xc = xc * c;
    } else if (divcos) {
      xa = xa / c;
      xc = xc / c;
    }
    grid[(size_t)jp * L + l] = xa;
    grid[(size_t)(nlat - 1 - jp) * L + l] = xc;
  }
}

__global__ __launch_bounds__(256) void dft_full_kernel(const double* __restrict__ grid,
                                                       double2* __restrict__ four,
                                                       const double2* __restrict__ twg, int nlat,
                                                       int L) {
  extern __shared__ double2 dsh[];  // tw[L] | row (L doubles)
  double2* tw = dsh;
  double* row = reinterpret_cast<double*>(dsh + L);
  const int j = blockIdx.x, b = blockIdx.y;
  const double* r = grid + ((size_t)b * nlat + j) * L;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    tw[i] = twg[i];
    row[i] = r[i];
  }
  __syncthreads();
  const double scale = 1.0 / L;
  for (int kk = threadIdx.x; kk < L; kk += blockDim.x) {
    double2 x = make_double2(0.0, 0.0);
    int k = 0;
    for (int l = 0; l < L; ++l) {
      const double2 w = tw[k];
      x = make_double2(x.x + row[l] * w.x, x.y + row[l] * w.y);
      k += kk;
      if (k >= L) k -= L;
    }
    four[((size_t)b * nlat + j) * L + kk] = make_double2(x.x * scale, x.y * scale);
  }
}

static bool pow2_fft(int L) { return L >= 16 && L <= 1024 && (L & (L - 1)) == 0; }

// Full Fourier array [B][nlat][nlon] -> E/O block (for Compute_Alpha_Operator!).
__global__ void fold_full_kernel(const double2* __restrict__ four, double2* __restrict__ eo,
                                 int nlat, int nlon, int nm, int B) {
  int half = nlat / 2;
  size_t total = (size_t)B * nm * half;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    int jp = idx % half;
    int m = (idx / half) % nm;
    int b = idx / ((size_t)half * nm);
    double2 a = four[((size_t)b * nlat + jp) * nlon + m];
    double2 c = four[((size_t)b * nlat + (nlat - 1 - jp)) * nlon + m];
    eo[(((size_t)b * nm + m) * 2 + 0) * half + jp] = cadd(a, c);
    eo[(((size_t)b * nm + m) * 2 + 1) * half + jp] = csub(a, c);
  }
}

// ------------------------------------------------------------ Legendre stage
// Legendre tile shape: LCOLS real columns per workgroup (LCOLS / 2 complex
// fields, one wave per 16 columns) and TL result tiles of 16 rows (forward)
// or 16 latitudes (inverse) per workgroup.  TL is a run-time choice per mesh
// and direction (jgcm_mesh_set_legendre_shape: 1, 2 or 4; 0 = DEF_TILES),
// picked by the autotuner; the results do not depend on it.
#ifndef JGCM_LCOLS
#define JGCM_LCOLS 64
#endif
constexpr int LCOLS = JGCM_LCOLS;
constexpr int LTHREADS = LCOLS * 4;
constexpr int DEF_TILES = 2;

// Table rows are read at wave-uniform addresses; going through the constant
// address space lets the compiler use scalar loads (one load per wave,
// operand in SGPRs) instead of per-lane vector loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

template <int PREC> struct LdsT { using T = double; };
template <> struct LdsT<P_FP32> { using T = float; };
template <> struct LdsT<P_SPLIT> { using T = float2; };   // (hi, lo)
template <> struct LdsT<P_BF16> { using T = ushort4; };   // (b1, b2, b3, 0)

template <int PREC>
__device__ __forceinline__ typename LdsT<PREC>::T to_lds(double v) {
  if constexpr (PREC == P_FP64) return v;
  else if constexpr (PREC == P_FP32) return (float)v;
  else if constexpr (PREC == P_SPLIT) {
    float h = (float)v;
    return make_float2(h, (float)(v - (double)h));
  } else {
    return split_bf16x3(v);
  }
}

struct FwdTables {
  const double* d;      // FP64 rows
  const float* hi;      // FP32 / split hi
  const float* lo;      // split lo
  const ushort4* b3;    // BF16 three-piece split
};

struct FwdJob {
  // term A: coefficient table tA against mulA * X_A^(par)
  const double2* srcA;  // E/O block of this job's first field
  FwdTables tA;
  int mulA_i;           // 1: X_A multiplied by i
  // term B (optional): table tB against signB * X_B^(1-par)
  const double2* srcB;
  FwdTables tB;
  double signB;
  int nf;               // fields in this job
  SegMap out;           // spectral fields [ns+1][nm] (complex) each
  // outputs scaled by scale0 (m = 0) or scale1 (m > 0), then stored or, with
  // accumulate, added to the output array
  double scale0 = 1.0, scale1 = 1.0;
  int accumulate = 0;
};

constexpr int MAXJOB = 4;
struct FwdJobs {
  FwdJob job[MAXJOB];
  int njobs;
  int zstart[MAXJOB + 1];  // first blockIdx.z of each job (field chunks of 32)
};

// ------------------------------------------------ Legendre stage on MFMA
// Per (m, parity) segment the Legendre sums are small dense products:
//   forward  S[k][c] = sum_j P[n_k][j] X[j][c]      (rows k x latitudes j)
//   inverse  E[j][c] = sum_k P[n_k][j] S[k][c]      (latitudes x rows)
// with c = 2 * field + (re, im).  Each wave owns 16 columns (8 fields) and
// accumulates 16 x 16 output tiles with v_mfma_f64_16x16x4f64 (FP64) or
// v_mfma_f32_16x16x4f32 (FP32; split-FP32 as hi*hi + (hi*lo + lo*hi) in FP32
// over blocks of 16 contraction steps, block sums added in FP64).  Both
// operands go through LDS: per chunk of KC contraction steps the workgroup
// stages the right-hand block (fields) and the segment's table rows with all
// loads in flight at once, then runs the MFMA loop.
// Operand layouts (both instructions): A[i][k] in lane i + 16 k, B[k][j] in
// lane j + 16 k.  Result D[i][j] in lane j + 16 (i % 4), register i / 4 for
// the FP64 instruction and lane j + 16 (i / 4), register i % 4 for FP32.
typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

constexpr int KC = 64;  // contraction steps (latitudes or rows) staged per pass
static_assert(KC <= 64, "the staging maps one lane to each contraction step of a chunk");
static_assert(LTHREADS % 64 == 0, "tile shape");
constexpr int KS1 = KC + 1;

__device__ __forceinline__ dv4 mma64(double a, double b, dv4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ fv4 mma32(float a, float b, fv4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// table element type in LDS: FP64, FP32, or (hi, lo)
template <int PREC> struct TabT { using T = double; };
template <> struct TabT<P_FP32> { using T = float; };
template <> struct TabT<P_SPLIT> { using T = float2; };
template <> struct TabT<P_BF16> { using T = ushort4; };

// contraction steps per MFMA: 4 (f64 / f32 16x16x4), 16 (bf16 16x16x16)
template <int PREC> constexpr int kstep() { return PREC == P_BF16 ? 16 : 4; }

template <int PREC>
__device__ __forceinline__ typename TabT<PREC>::T tab_load(const FwdTables& t, size_t i, bool ok) {
  if constexpr (PREC == P_FP64) return ok ? t.d[i] : 0.0;
  else if constexpr (PREC == P_FP32) return ok ? t.hi[i] : 0.0f;
  else if constexpr (PREC == P_SPLIT) return ok ? make_float2(t.hi[i], t.lo[i]) : make_float2(0.0f, 0.0f);
  else return ok ? t.b3[i] : make_ushort4(0, 0, 0, 0);
}

// row of the 16 x 16 result tile held in register r of a lane with lane/16 = lk
template <int PREC>
__device__ __forceinline__ int tile_row(int lk, int r) {
  if constexpr (PREC == P_FP64) return lk + 4 * r;
  else return 4 * lk + r;
}

// Up to 4 result tiles per wave, FP64 accumulators in every precision.
struct TileAcc {
  double v[4][4];
};

typedef short sv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ fv4 mma16(sv4 a, sv4 b, fv4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// acc[t] += sum over MFMA steps ks < KS of A_t(ks) B(ks); bval(ks, e) and
// aval(t, ks, e) give this lane's operand values (from LDS): e = 0 for the
// 16x16x4 instructions, e = 0..3 (four consecutive contraction steps of the
// lane) for the bf16 16x16x16 instruction.  In steps: mfma_run_begin, then
// mfma_run_steps over consecutive ranges of ks covering 0..KS (the staging of
// a chunk may arrive in pieces), then mfma_run_end; the operations and their
// order are those of one pass over 0..KS whatever the pieces: FP64 and FP32
// accumulate the whole chunk in the MFMA accumulators, split-FP32 and BF16x3
// close a block sum into acc after every 4 MFMA steps from the chunk's start
// and after the last one.
template <int PREC> struct MfmaRun {
  fv4 c[4];   // FP32: the sums; split: hi*hi; BF16x3: the block sums
  fv4 c2[4];  // split: the cross terms
};
template <> struct MfmaRun<P_FP64> {
  dv4 c[4];
};

template <int PREC>
__device__ __forceinline__ void mfma_run_begin(MfmaRun<PREC>& run) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if constexpr (PREC == P_FP64) {
      run.c[t] = dv4{0.0, 0.0, 0.0, 0.0};
    } else {
      run.c[t] = fv4{0.0f, 0.0f, 0.0f, 0.0f};
      run.c2[t] = fv4{0.0f, 0.0f, 0.0f, 0.0f};
    }
  }
}

// block sum of the split / BF16x3 forms into acc, accumulators cleared
template <int PREC>
__device__ __forceinline__ void mfma_run_close_block(TileAcc& acc, MfmaRun<PREC>& run) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if constexpr (PREC == P_SPLIT) acc.v[t][r] += (double)run.c[t][r] + (double)run.c2[t][r];
      else acc.v[t][r] += (double)run.c[t][r];
    }
    run.c[t] = fv4{0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (PREC == P_SPLIT) run.c2[t] = fv4{0.0f, 0.0f, 0.0f, 0.0f};
  }
}

template <int PREC, typename BF, typename AF>
__device__ __forceinline__ void mfma_run_steps(TileAcc& acc, MfmaRun<PREC>& run, int ntile, int ks0,
                                               int ks1, BF bval, AF aval) {
  if constexpr (PREC == P_FP64) {
#pragma unroll 4
    for (int ks = ks0; ks < ks1; ++ks) {
      double b = bval(ks, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t < ntile) run.c[t] = mma64(aval(t, ks, 0), b, run.c[t]);
    }
  } else if constexpr (PREC == P_FP32) {
#pragma unroll 4
    for (int ks = ks0; ks < ks1; ++ks) {
      float b = bval(ks, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t < ntile) run.c[t] = mma32(aval(t, ks, 0), b, run.c[t]);
    }
  } else if constexpr (PREC == P_SPLIT) {
    for (int ks = ks0; ks < ks1; ++ks) {
      float2 b = bval(ks, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t < ntile) {
          float2 a = aval(t, ks, 0);
          run.c[t] = mma32(a.x, b.x, run.c[t]);
          run.c2[t] = mma32(a.x, b.y, run.c2[t]);
          run.c2[t] = mma32(a.y, b.x, run.c2[t]);
        }
      }
      if ((ks & 3) == 3) mfma_run_close_block<PREC>(acc, run);
    }
  } else {
    // BF16 x 3: x = x1 + x2 + x3, P = p1 + p2 + p3; the six products with
    // piece orders summing to <= 2 (the rest are below 2^-24 |P x|), smallest
    // first, in FP32 over blocks of 64 contraction steps; block sums in FP64.
    for (int ks = ks0; ks < ks1; ++ks) {
      sv4 b1, b2, b3;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        ushort4 x = bval(ks, e);
        b1[e] = (short)x.x;
        b2[e] = (short)x.y;
        b3[e] = (short)x.z;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t < ntile) {
          sv4 a1, a2, a3;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            ushort4 p = aval(t, ks, e);
            a1[e] = (short)p.x;
            a2[e] = (short)p.y;
            a3[e] = (short)p.z;
          }
          run.c[t] = mma16(a1, b3, run.c[t]);
          run.c[t] = mma16(a2, b2, run.c[t]);
          run.c[t] = mma16(a3, b1, run.c[t]);
          run.c[t] = mma16(a1, b2, run.c[t]);
          run.c[t] = mma16(a2, b1, run.c[t]);
          run.c[t] = mma16(a1, b1, run.c[t]);
        }
      }
      if ((ks & 3) == 3) mfma_run_close_block<PREC>(acc, run);
    }
  }
}

// the chunk of KS MFMA steps is complete
template <int PREC>
__device__ __forceinline__ void mfma_run_end(TileAcc& acc, MfmaRun<PREC>& run, int KS) {
  if constexpr (PREC == P_FP64 || PREC == P_FP32) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc.v[t][r] += (double)run.c[t][r];
  } else {
    if (KS & 3) mfma_run_close_block<PREC>(acc, run);
  }
}

template <int PREC, typename BF, typename AF>
__device__ __forceinline__ void mfma_accumulate(TileAcc& acc, int ntile, int KS, BF bval, AF aval) {
  MfmaRun<PREC> run;
  mfma_run_begin<PREC>(run);
  mfma_run_steps<PREC>(acc, run, ntile, 0, KS, bval, aval);
  mfma_run_end<PREC>(acc, run, KS);
}

// Two staged copies whose global loads are issued together (up to IT + IT
// per thread in flight), then written to LDS.
#ifndef JGCM_STAGE_IT
#define JGCM_STAGE_IT 8
#endif
template <typename L1, typename S1, typename L2, typename S2>
__device__ __forceinline__ void staged_copy2(int n1, L1 load1, S1 store1, int n2, L2 load2,
                                             S2 store2) {
  using T1 = decltype(load1(0));
  using T2 = decltype(load2(0));
  constexpr int IT = JGCM_STAGE_IT;
  const int nt = blockDim.x;
  const int n = max(n1, n2);
  for (int base = 0; base < n; base += IT * nt) {
    T1 a[IT];
    T2 b[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int idx = base + it * nt + (int)threadIdx.x;
      if (idx < n1) a[it] = load1(idx);
      if (idx < n2) b[it] = load2(idx);
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      int idx = base + it * nt + (int)threadIdx.x;
      if (idx < n1) store1(idx, a[it]);
      if (idx < n2) store2(idx, b[it]);
    }
  }
}

__host__ __device__ constexpr size_t align16(size_t b) { return (b + 15) / 16 * 16; }

// Workgroup coordinates (x fastest) after an XCD-aware renumbering.  The
// dispatcher hands workgroup L to XCD L % 8 in launch order.  Renumbering
// deals chunks of XCD_CHUNK consecutive logical workgroups round-robin over
// the XCDs: consecutive segments (m, parity) -- 8 values of m per chunk, one
// 128-byte line of every spectral row, m being the fastest index of the
// spectral arrays -- meet in one L2, every XCD still gets chunks from the
// whole range of m (balanced triangle), and the heavy low-m segments are still
// launched first.  Indices past the last full round stay in place.
#ifndef JGCM_XCD_REMAP
#define JGCM_XCD_REMAP 1
#endif
constexpr unsigned NXCD = 8;
constexpr unsigned XCD_CHUNK = 16;
__device__ __forceinline__ uint3 wg_coords() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
#if JGCM_XCD_REMAP
  const unsigned T = gx * gy * gridDim.z;
  const unsigned full = T / (NXCD * XCD_CHUNK) * (NXCD * XCD_CHUNK);
  if (L < full) {
    const unsigned x = L % NXCD, slot = L / NXCD;
    L = (x + NXCD * (slot / XCD_CHUNK)) * XCD_CHUNK + slot % XCD_CHUNK;
  }
#endif
  return make_uint3(L % gx, (L / gx) % gy, L / (gx * gy));
}

// Forward tile wg: x = segment (m, par); y = group of TL 16-row tiles; z =
// (job, chunk of 32 fields).  256 threads: wave w owns columns 16w .. 16w+15.
// LDS: X [64 cols][KC+1], P [16 TL rows][KC+1].  Shared by the single-mesh
// kernel (wg from the launch grid) and the multi-mesh work-list kernel (wg
// from a work item).
template <int PREC, int TL>
__device__ __forceinline__ void legendre_fwd_tile(const FwdJobs& jobs, int ns, int nm, int half,
                                                  int nmax, int zero_top, const uint3 wg,
                                                  unsigned char* smem_raw) {
  constexpr int TROWS = 16 * TL;
  static_assert(TL == 1 || TL == 2 || TL == 4, "1, 2 or 4 tiles of 16 rows");
  using LT = typename LdsT<PREC>::T;
  using TT = typename TabT<PREC>::T;
  LT* X = reinterpret_cast<LT*>(smem_raw);
  TT* Pt = reinterpret_cast<TT*>(smem_raw + align16(sizeof(LT) * LCOLS * KS1));
  const int m = wg.x >> 1;
  const int par = wg.x & 1;
  int jobi = 0;
#pragma unroll
  for (int q = 1; q < MAXJOB; ++q)
    if (q < jobs.njobs && (int)wg.z >= jobs.zstart[q]) jobi = q;
  const FwdJob& job = jobs.job[jobi];
  const int f0 = (wg.z - jobs.zstart[jobi]) * (LCOLS / 2);
  const int nf = min(LCOLS / 2, job.nf - f0);
  const size_t spec_d = (size_t)(ns + 1) * nm * 2;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int nfirst = m + par;
  const int nr = (nfirst <= nmax) ? (nmax - nfirst) / 2 + 1 : 0;
  if (zero_top && par == 0 && wg.y == 0 && threadIdx.x < 2 * nf) {
    int fl = threadIdx.x >> 1;
    double* o = seg_field(job.out, f0 + fl, spec_d) + ((size_t)ns * nm + m) * 2;
    o[threadIdx.x & 1] = 0.0;
  }
  const int kk0 = wg.y * TROWS;
  const int nrg = min(TROWS, nr - kk0);
  if (nrg <= 0) return;  // uniform over the workgroup
  const int ntile = (nrg + 15) / 16;
  const size_t rowbase = (size_t)(packed_row_start(m, ns) + nfirst + 2 * kk0 - m) * half;
  TileAcc acc;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc.v[t][r] = 0.0;
  constexpr int KSTEP = kstep<PREC>();
  const int lko = KSTEP == 16 ? 4 * lk : lk;  // this lane's first contraction step
  const LT* Xc = X + (16 * wave + li) * KS1 + lko;
  const TT* Pc = Pt + li * KS1 + lko;
  // Staging in sub-chunks of SUB contraction steps (FP64: the products of one
  // sub-chunk overlap the arrival of the next; the faster products of the
  // other precisions measured best with one sub-chunk per chunk).  NG lane
  // groups of SUB lanes each; a lane loads FL fields and RL table rows of a
  // sub-chunk.
  constexpr int SUB = PREC == P_FP64 ? 16 : KC;
  constexpr int NSUB = KC / SUB;
  constexpr int NG = 64 / SUB;
  constexpr int NWV = LTHREADS / 64;
  constexpr int FPT = (LCOLS / 2 + NWV - 1) / NWV;   // fields per wave
  constexpr int RPT = (TROWS + NWV - 1) / NWV;       // table rows per wave
  constexpr int FL = FPT / NG, RL = (RPT + NG - 1) / NG;
  static_assert(FPT == NG * FL && KC % SUB == 0, "fields split evenly over the lane groups");
  const int jl = lane & (SUB - 1), fq = lane / SUB;
  for (int term = 0; term < 2; ++term) {
    const double2* src = term == 0 ? job.srcA : job.srcB;
    if (!src) break;
    const FwdTables& tab = term == 0 ? job.tA : job.tB;
    const int pt = term == 0 ? par : par ^ 1;
    const int mul_i = term == 0 ? job.mulA_i : 0;
    const double sgn = term == 0 ? 1.0 : job.signB;
    const double2* srow = src + (((size_t)f0 * nm + m) * 2 + pt) * half;
    const size_t fstride = (size_t)nm * 2 * half;
    for (int lat0 = 0; lat0 < half; lat0 += KC) {
      const int kcl = min(KC, half - lat0);
      const int KSc = (kcl + KSTEP - 1) / KSTEP;
      const int kpad = KSTEP * KSc;
      const int nsub = (kcl + SUB - 1) / SUB;
      __syncthreads();  // the previous chunk's operands are consumed
      // Every load of the chunk is issued first, in sub-chunks of SUB
      // latitudes (lane = latitude jl of the sub-chunk and field / row group
      // fq; the wave picks the fields and table rows: no index division);
      // then each sub-chunk goes to LDS as soon as its own loads are in, and
      // its MFMA steps run while the later sub-chunks are still arriving.
      double2 xv[NSUB][FL];
      TT pv[NSUB][RL];
#pragma unroll
      for (int sc = 0; sc < NSUB; ++sc) {
        const int j = SUB * sc + jl;
        const bool jin = sc < nsub && j < kcl;
#pragma unroll
        for (int p = 0; p < FL; ++p) {
          const int fl = wave + NWV * (fq + NG * p);
          xv[sc][p] = (fl < nf && jin) ? srow[fl * fstride + lat0 + j] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int p = 0; p < RL; ++p) {
          const int i = fq + NG * p;
          const int k = wave + NWV * i;
          const bool ok = i < RPT && k < nrg && jin;
          pv[sc][p] = tab_load<PREC>(tab, ok ? rowbase + (size_t)(2 * k) * half + lat0 + j : 0, ok);
        }
      }
      MfmaRun<PREC> run;
      mfma_run_begin<PREC>(run);
#pragma unroll
      for (int sc = 0; sc < NSUB; ++sc) {
        if (sc >= nsub) break;  // uniform over the workgroup
        const int j = SUB * sc + jl;
        if (j < kpad) {
#pragma unroll
          for (int p = 0; p < FL; ++p) {
            const int fl = wave + NWV * (fq + NG * p);
            double re = xv[sc][p].x, im = xv[sc][p].y;
            if (mul_i) { double t = re; re = -im; im = t; }
            re *= sgn;
            im *= sgn;
            X[(2 * fl) * KS1 + j] = to_lds<PREC>(re);
            X[(2 * fl + 1) * KS1 + j] = to_lds<PREC>(im);
          }
#pragma unroll
          for (int p = 0; p < RL; ++p) {
            const int i = fq + NG * p;
            const int k = wave + NWV * i;
            if (i < RPT && k < TROWS) Pt[k * KS1 + j] = pv[sc][p];
          }
        }
        __syncthreads();
        const int ks0 = SUB * sc / KSTEP, ks1 = min(KSc, SUB * (sc + 1) / KSTEP);
        mfma_run_steps<PREC>(
            acc, run, ntile, ks0, ks1, [&](int ks, int e) { return Xc[KSTEP * ks + e]; },
            [&](int t, int ks, int e) { return Pc[16 * t * KS1 + KSTEP * ks + e]; });
      }
      mfma_run_end<PREC>(acc, run, KSc);
    }
  }
  const int col = 16 * wave + li;
  const int fl = col >> 1, comp = col & 1;
  if (fl < nf) {
    double* base = seg_field(job.out, f0 + fl, spec_d) + (size_t)m * 2 + comp;
    const double sc = m == 0 ? job.scale0 : job.scale1;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= ntile) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int kk = kk0 + 16 * t + tile_row<PREC>(lk, r);
        if (kk < nr) {
          double* o = base + (size_t)(nfirst + 2 * kk) * nm * 2;
          double v = (double)acc.v[t][r] * sc;
          if (job.accumulate) v = *o + v;
          *o = v;
        }
      }
    }
  }
}

template <int PREC, int TL>
__global__ __launch_bounds__(LTHREADS) void legendre_fwd_mfma(FwdJobs jobs, int ns, int nm, int half,
                                                         int nmax, int zero_top) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  legendre_fwd_tile<PREC, TL>(jobs, ns, nm, half, nmax, zero_top, wg_coords(), smem_raw);
}

// Operator table values of one spectral row (m, n), loaded once per
// staging lane and shared by every field of the chunk (they do not depend on
// the field).  Only the tables of the operators in `op_mask` (bit 1 << op)
// are read.
struct OpRow {
  double dlam, dthm, dthp, uvc, uvm, uvp;
};

__device__ __forceinline__ OpRow op_row(const SpecOpTables& t, int op_mask, int m, int n, int ns) {
  const int c = m * (ns + 1) + n;
  OpRow r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (op_mask & (1 << SPEC_OP_GRAD_LAMBDA)) r.dlam = t.dlam[c];
  if (op_mask & (1 << SPEC_OP_GRAD_THETA)) {
    r.dthm = t.dthm[c];
    r.dthp = t.dthp[c];
  }
  if (op_mask & ((1 << SPEC_OP_UCOS) | (1 << SPEC_OP_VCOS))) {
    r.uvc = t.uvc[c];
    r.uvm = t.uvm[c];
    r.uvp = t.uvp[c];
  }
  return r;
}

// Spectral coefficient (n, m) after the operator `op` (SPEC_OP_*); h (and h2,
// the second operand of the u cos / v cos operators) point at the field's
// coefficients, r holds the row's table values.  The arithmetic and its order
// are those of the stand-alone operator kernels.
__device__ __forceinline__ double2 staged_coefficient_row(const double2* h, const double2* h2, int op,
                                                          const OpRow& r, int n, int m, int nm,
                                                          int ns) {
  const size_t i = (size_t)n * nm + m;
  if (op == SPEC_OP_NONE) return h[i];
  if (op == SPEC_OP_NEG_I) {
    double2 v = h[i];
    return make_double2(v.y, -v.x);
  }
  if (op == SPEC_OP_NEG) {
    double2 v = h[i];
    return make_double2(-v.x, -v.y);
  }
  if (op == SPEC_OP_GRAD_LAMBDA) {
    double2 v = h[i];
    return make_double2(r.dlam * (-v.y), r.dlam * v.x);
  }
  if (op == SPEC_OP_GRAD_THETA) {
    double2 q = make_double2(0.0, 0.0);
    if (n <= ns - 1) {
      double2 hp = h[i + nm];
      q = make_double2(r.dthp * hp.x, r.dthp * hp.y);
    }
    if (n >= 1) {
      double2 hm = h[i - nm];
      q = make_double2(q.x + r.dthm * hm.x, q.y + r.dthm * hm.y);
    }
    return q;
  }
  // u cos θ / v cos θ: h = vor, h2 = div
  const double2* vor = h;
  const double2* div = h2;
  const double cc = r.uvc;
  if (op == SPEC_OP_UCOS) {
    double2 d = div[i];
    double2 u = make_double2(cc * (-d.y), cc * d.x);
    if (n >= 1) {
      double2 vm = vor[i - nm];
      u = make_double2(u.x - r.uvm * vm.x, u.y - r.uvm * vm.y);
    }
    if (n <= ns - 1) {
      double2 vp = vor[i + nm];
      u = make_double2(u.x - r.uvp * vp.x, u.y - r.uvp * vp.y);
    }
    return u;
  }
  double2 v = vor[i];
  double2 w = make_double2(cc * (-v.y), cc * v.x);
  if (n >= 1) {
    double2 dm = div[i - nm];
    w = make_double2(w.x + r.uvm * dm.x, w.y + r.uvm * dm.y);
  }
  if (n <= ns - 1) {
    double2 dp = div[i + nm];
    w = make_double2(w.x + r.uvp * dp.x, w.y + r.uvp * dp.y);
  }
  return w;
}

// The same with the field given by its global index b in a segment map.
__device__ __forceinline__ double2 staged_coefficient_seg(const SegMap& in, const OpRow& r, int b,
                                                          size_t spec_d, int n, int m, int nm,
                                                          int ns) {
  const int k = seg_find(in, b);
  const size_t off = (size_t)(b - in.start[k]) * spec_d;
  const double2* h = reinterpret_cast<const double2*>(in.ptr[k] + off);
  const double2* h2 = in.ptr2[k] ? reinterpret_cast<const double2*>(in.ptr2[k] + off) : nullptr;
  return staged_coefficient_row(h, h2, (in.flag[k] >> 1) & 7, r, n, m, nm, ns);
}

// Inverse tile wg: x = segment (m, par), y = chunk of 32 fields, z = group of
// 16 TL latitudes; 256 threads, wave w owns columns 16w .. 16w+15.
// LDS: S [64 cols][KC+1], P [KC rows][16 TL + 1].  Shared like the forward
// tile by the single-mesh and the work-list kernels.
// PLAIN: every field comes from in.ptr[0] without an operator (checked by the
// launcher): the staging skips the segment and operator logic.
template <int PREC, bool PLAIN, int TL>
__device__ __forceinline__ void legendre_inv_tile(const SegMap& in, double2* __restrict__ eo,
                                                  const FwdTables& tab, int B, int nmax, int ns,
                                                  int nm, int half, const SpecOpTables& ops,
                                                  int out_mode, const uint3 wg,
                                                  unsigned char* smem_raw) {
  constexpr int TROWS = 16 * TL;
  static_assert(TL == 1 || TL == 2 || TL == 4, "1, 2 or 4 tiles of 16 latitudes");
  using LT = typename LdsT<PREC>::T;
  using TT = typename TabT<PREC>::T;
  LT* S = reinterpret_cast<LT*>(smem_raw);
  TT* Pt = reinterpret_cast<TT*>(smem_raw + align16(sizeof(LT) * LCOLS * KS1));
  const int m = wg.x >> 1;
  const int par = wg.x & 1;
  const int f0 = wg.y * (LCOLS / 2);
  const int nf = min(LCOLS / 2, B - f0);
  const int lat0 = wg.z * TROWS;
  const int nlg = min(TROWS, half - lat0);
  const int nlt = (nlg + 15) / 16;
  const int nfirst = m + par;
  const int nr = (nfirst <= nmax) ? (nmax - nfirst) / 2 + 1 : 0;
  const size_t spec_d = (size_t)(ns + 1) * nm * 2;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const size_t rowbase = (size_t)(packed_row_start(m, ns) + nfirst - m) * half + lat0;
  TileAcc acc;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc.v[t][r] = 0.0;
  constexpr int KSTEP = kstep<PREC>();
  const int lko = KSTEP == 16 ? 4 * lk : lk;  // this lane's first contraction step
  const LT* Sc = S + (16 * wave + li) * KS1 + lko;
  constexpr int PSI = TROWS + 1;
  const TT* Pc = Pt + lko * PSI + li;
  // the segment (array) of the chunk's fields, looked up once when the chunk
  // lies in a single one (a lookup per field costs more than the loads)
  const int seg_a = seg_find(in, f0), seg_b = seg_find(in, f0 + max(nf, 1) - 1);
  const bool seg_one = seg_a == seg_b;
  const size_t seg_off = (size_t)(f0 - in.start[seg_a]) * spec_d;
  const double2* seg_h = reinterpret_cast<const double2*>(in.ptr[seg_a] + seg_off);
  const double2* seg_h2 =
      in.ptr2[seg_a] ? reinterpret_cast<const double2*>(in.ptr2[seg_a] + seg_off) : nullptr;
  const int seg_op = (in.flag[seg_a] >> 1) & 7;
  int chunk_ops = 0;  // operators of the segments the chunk touches (bit 1 << op)
  for (int q = seg_a; q <= seg_b; ++q) chunk_ops |= 1 << ((in.flag[q] >> 1) & 7);
  if constexpr (PLAIN && PREC == P_FP64) {
    // Plain FP64 transforms stage in sub-chunks of SUBI rows: every load of a
    // chunk is issued first (lane = row jl of the sub-chunk and field group
    // fq), each sub-chunk goes to LDS when its own loads are in, and its MFMA
    // steps run while the later sub-chunks are still arriving.
    constexpr int SUBI = 16, NSUBI = KC / SUBI, NGI = 64 / SUBI;
    constexpr int NWV = LTHREADS / 64;
    constexpr int FPT = (LCOLS / 2 + NWV - 1) / NWV;
    constexpr int FLI = FPT / NGI;
    constexpr int TPS = (SUBI * TROWS + LTHREADS - 1) / LTHREADS;  // table loads per sub-chunk
    static_assert(FPT == NGI * FLI, "fields split evenly over the lane groups");
    const int jl = lane & (SUBI - 1), fq = lane / SUBI;
    const double2* h0 = reinterpret_cast<const double2*>(in.ptr[0]);
    for (int kk0 = 0; kk0 < nr; kk0 += KC) {
      const int kcr = min(KC, nr - kk0);
      const int KSc = (kcr + KSTEP - 1) / KSTEP;
      const int kpad = KSTEP * KSc;
      const int nsub = (kcr + SUBI - 1) / SUBI;
      __syncthreads();
      double2 sv[NSUBI][FLI];
      TT pv[NSUBI][TPS];
#pragma unroll
      for (int sc = 0; sc < NSUBI; ++sc) {
        const int kr = SUBI * sc + jl;
        const bool kin = sc < nsub && kr < kcr;
        const int nrow = nfirst + 2 * (kk0 + kr);
#pragma unroll
        for (int p = 0; p < FLI; ++p) {
          const int fl = wave + NWV * (fq + NGI * p);
          sv[sc][p] = (fl < nf && kin)
                          ? h0[(size_t)(f0 + fl) * (spec_d / 2) + (size_t)nrow * nm + m]
                          : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < TPS; ++i) {
          const int idx = (int)threadIdx.x + LTHREADS * i;
          const int j = idx % TROWS, k = SUBI * sc + idx / TROWS;
          const bool ok = sc < nsub && idx < SUBI * TROWS && k < kcr && j < nlg;
          pv[sc][i] = tab_load<PREC>(tab, ok ? rowbase + (size_t)(2 * (kk0 + k)) * half + j : 0, ok);
        }
      }
      MfmaRun<PREC> run;
      mfma_run_begin<PREC>(run);
#pragma unroll
      for (int sc = 0; sc < NSUBI; ++sc) {
        if (sc >= nsub) break;  // uniform over the workgroup
        const int kr = SUBI * sc + jl;
        if (kr < kpad) {
#pragma unroll
          for (int p = 0; p < FLI; ++p) {
            const int fl = wave + NWV * (fq + NGI * p);
            S[(2 * fl) * KS1 + kr] = to_lds<PREC>(sv[sc][p].x);
            S[(2 * fl + 1) * KS1 + kr] = to_lds<PREC>(sv[sc][p].y);
          }
        }
#pragma unroll
        for (int i = 0; i < TPS; ++i) {
          const int idx = (int)threadIdx.x + LTHREADS * i;
          const int k = SUBI * sc + idx / TROWS;
          if (idx < SUBI * TROWS && k < kpad) Pt[k * PSI + idx % TROWS] = pv[sc][i];
        }
        __syncthreads();
        const int ks0 = SUBI * sc / KSTEP, ks1 = min(KSc, SUBI * (sc + 1) / KSTEP);
        mfma_run_steps<PREC>(
            acc, run, nlt, ks0, ks1, [&](int ks, int e) { return Sc[KSTEP * ks + e]; },
            [&](int t, int ks, int e) { return Pc[(KSTEP * ks + e) * PSI + 16 * t]; });
      }
      mfma_run_end<PREC>(acc, run, KSc);
    }
  } else {
    for (int kk0 = 0; kk0 < nr; kk0 += KC) {
      const int kcr = min(KC, nr - kk0);
      const int KSc = (kcr + KSTEP - 1) / KSTEP;
      const int kpad = KSTEP * KSc;
      __syncthreads();
      {
        // coefficients: lane = contraction row of the chunk (kpad <= KC = 64),
        // the wave picks the fields (uniform segment lookup, no index division)
        constexpr int NWV = LTHREADS / 64;
        constexpr int FPT = (LCOLS / 2 + NWV - 1) / NWV;
        const int kr = lane;
        const bool kin = kr < kcr;
        const int nrow = nfirst + 2 * (kk0 + kr);
        double2 sv[FPT];
        if constexpr (PLAIN) {
          // the plain transform of one array (the common case): direct loads
          const double2* h0 = reinterpret_cast<const double2*>(in.ptr[0]);
  #pragma unroll
          for (int i = 0; i < FPT; ++i) {
            const int fl = wave + NWV * i;
            sv[i] = (fl < nf && kin)
                        ? h0[(size_t)(f0 + fl) * (spec_d / 2) + (size_t)nrow * nm + m]
                        : make_double2(0.0, 0.0);
          }
        } else if (seg_one) {
          // the chunk's fields lie in one array: pointer arithmetic per field,
          // one straight loop per operator (the operator is the workgroup's)
          const OpRow orow = op_row(ops, 1 << seg_op, m, kin ? nrow : m, ns);
          auto stage_op = [&](auto opc) {
            constexpr int OP = decltype(opc)::value;
  #pragma unroll
            for (int i = 0; i < FPT; ++i) {
              const int fl = wave + NWV * i;
              const size_t o = (size_t)fl * (spec_d / 2);
              sv[i] = (fl < nf && kin)
                          ? staged_coefficient_row(seg_h + o, seg_h2 ? seg_h2 + o : nullptr, OP,
                                                   orow, nrow, m, nm, ns)
                          : make_double2(0.0, 0.0);
            }
          };
          switch (seg_op) {
            case SPEC_OP_NONE: stage_op(std::integral_constant<int, SPEC_OP_NONE>{}); break;
            case SPEC_OP_GRAD_LAMBDA: stage_op(std::integral_constant<int, SPEC_OP_GRAD_LAMBDA>{}); break;
            case SPEC_OP_GRAD_THETA: stage_op(std::integral_constant<int, SPEC_OP_GRAD_THETA>{}); break;
            case SPEC_OP_UCOS: stage_op(std::integral_constant<int, SPEC_OP_UCOS>{}); break;
            case SPEC_OP_VCOS: stage_op(std::integral_constant<int, SPEC_OP_VCOS>{}); break;
            case SPEC_OP_NEG_I: stage_op(std::integral_constant<int, SPEC_OP_NEG_I>{}); break;
            default: stage_op(std::integral_constant<int, SPEC_OP_NEG>{}); break;
          }
        } else {
          // fields from several arrays: the tables of every operator in the
          // chunk are read once per lane, the array is resolved per field
          const OpRow orow = op_row(ops, chunk_ops, m, kin ? nrow : m, ns);
  #pragma unroll
          for (int i = 0; i < FPT; ++i) {
            const int fl = wave + NWV * i;
            sv[i] = (fl < nf && kin) ? staged_coefficient_seg(in, orow, f0 + fl, spec_d, nrow, m, nm, ns)
                                     : make_double2(0.0, 0.0);
          }
        }
        // table rows: idx -> (latitude, row) by shifts (TROWS is a power of two)
        constexpr int TPT = (KC * TROWS + LTHREADS - 1) / LTHREADS;
        TT pv[TPT];
  #pragma unroll
        for (int i = 0; i < TPT; ++i) {
          const int idx = (int)threadIdx.x + LTHREADS * i;
          const int j = idx % TROWS, k = idx / TROWS;
          const bool ok = k < kcr && j < nlg;
          pv[i] = tab_load<PREC>(tab, ok ? rowbase + (size_t)(2 * (kk0 + k)) * half + j : 0, ok);
        }
        if (kr < kpad) {
  #pragma unroll
          for (int i = 0; i < FPT; ++i) {
            const int fl = wave + NWV * i;
            if (fl >= LCOLS / 2) break;
            S[(2 * fl) * KS1 + kr] = to_lds<PREC>(sv[i].x);
            S[(2 * fl + 1) * KS1 + kr] = to_lds<PREC>(sv[i].y);
          }
        }
  #pragma unroll
        for (int i = 0; i < TPT; ++i) {
          const int idx = (int)threadIdx.x + LTHREADS * i;
          if (idx < kpad * TROWS) Pt[(idx / TROWS) * PSI + idx % TROWS] = pv[i];
        }
      }
      __syncthreads();
      mfma_accumulate<PREC>(
          acc, nlt, KSc, [&](int ks, int e) { return Sc[KSTEP * ks + e]; },
          [&](int t, int ks, int e) { return Pc[(KSTEP * ks + e) * PSI + 16 * t]; });
    }
  }
  // The tile goes through LDS ([field][latitude] complex) so that each
  // field's latitudes leave as one contiguous run of whole lines.
  const int col = 16 * wave + li;
  const int fl = col >> 1, comp = col & 1;
  double* O = reinterpret_cast<double*>(smem_raw);
  __syncthreads();  // the last chunk's operands are consumed
  if (fl < nf) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= nlt) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = 16 * t + tile_row<PREC>(lk, r);
        O[(fl * TROWS + j) * 2 + comp] = acc.v[t][r];
      }
    }
  }
  __syncthreads();
  // out_mode bit 0: write the opposite parity's block (transposed two-table
  // contraction); bit 1: add to the block instead of overwriting it
  const int pout = (out_mode & 1) ? par ^ 1 : par;
  const double2* O2 = reinterpret_cast<const double2*>(O);
  for (int idx = threadIdx.x; idx < nf * TROWS; idx += LTHREADS) {
    const int j = idx % TROWS, f = idx / TROWS;
    if (j < nlg) {
      double2* o = eo + (((size_t)(f0 + f) * nm + m) * 2 + pout) * half + lat0 + j;
      double2 v = O2[idx];
      if (out_mode & 2) v = cadd(*o, v);
      store_stream(o, v);
    }
  }
}

template <int PREC, bool PLAIN, int TL>
__global__ __launch_bounds__(LTHREADS) void legendre_inv_mfma(SegMap in, double2* __restrict__ eo,
                                                         FwdTables tab, int B, int nmax, int ns,
                                                         int nm, int half,
                                                         SpecOpTables ops, int out_mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  legendre_inv_tile<PREC, PLAIN, TL>(in, eo, tab, B, nmax, ns, nm, half, ops, out_mode,
                                     wg_coords(), smem_raw);
}

// ------------------------------------------------------------------ drivers
// Latitude pairs per FFT workgroup.  A workgroup runs 256/(L/4) FFTs per
// pass; at least 8 pairs per workgroup, except at L = 256 where one pass of 4
// measured faster for the fold at every batch size and for the unfold below
// ~48 level-fields (T85: fold -1.2 us at 20 and 61 level-fields, unfold
// -1.3 us at 20, +0.9 us at 61).
// Largest latitude-pair count whose staging buffer, with the kernels' static
// LDS (FPW + 1 rows of L complex values), fits in 64 KB.
// nλ = 128 and 256 with 2 (nm - 1) < nλ run the four-step kernels (at 64
// the radix-4 kernels measured faster: 4 lanes per transform leave too
// little work per wavefront)
static bool use_four_step(const jgcm_mesh* M) {
  const int L = M->nlon;
  return (L == 128 || L == 256) && 2 * (M->nm - 1) < L;
}

// four-step kernels: latitude pairs per workgroup G = 64 / P, 128 / P or
// 256 / P (one to four wavefronts), the largest limited to what fits in 64 KB
static size_t four_step_lds_per_pair(const jgcm_mesh* M) {
  const int P = M->nlon / 16;
  return (size_t)(16 * (P + 1) + 2 * M->nm + 1) * sizeof(double2);
}

static int four_step_pairs(const jgcm_mesh* M, int ppw) {
  const int P = M->nlon / 16;
  const size_t per = four_step_lds_per_pair(M);
  int g = 64 / P;
  // built-in choice: at least 8 pairs, so that the E/O rows are whole lines
  if (ppw <= 0) ppw = 8;
  while (g < ppw && 2 * g * P <= 256 && (size_t)(2 * g) * per <= 64 * 1024) g *= 2;
  return g;
}

int fft_max_ppw(const jgcm_mesh* M) {
  if (use_four_step(M)) return four_step_pairs(M, 1 << 20);
  const int L = M->nlon;
  const int fpw = 256 / (L / 4);
  const long stat = (long)(fpw + 1) * L * 16;
  const long per = (long)(2 * M->nm + 1) * 16;
  long mx = (64L * 1024 - stat) / per;
  if (mx > M->half) mx = M->half;
  return mx < 1 ? 1 : (int)mx;
}

static int pick_ppw(int L, int B, bool fold) {
  int fpw = 256 / (L / 4);
  if (fpw == 4 && (fold || B < 48)) return 4;
  return fpw >= 8 ? fpw : 8;
}

template <int P, int G>
static void launch_fold_fs(const jgcm_mesh* M, const SegMap& in, double2* eo, int B) {
  dim3 grd((M->half + G - 1) / G, B);
  size_t lds = (size_t)(2 * M->nm) * G * sizeof(double2);
  const double2* tw = reinterpret_cast<const double2*>(M->d_twiddle);
  if (in.nseg == 1 && in.flag[0] == 0)
    hipLaunchKernelGGL((fold_four_step_kernel<P, G, true>), grd, dim3(P * G), lds, M->stream, in,
                       eo, tw, M->d_cos, M->nlat, M->nm);
  else
    hipLaunchKernelGGL((fold_four_step_kernel<P, G, false>), grd, dim3(P * G), lds, M->stream, in,
                       eo, tw, M->d_cos, M->nlat, M->nm);
}

template <int P, int G>
static void launch_unfold_fs(const jgcm_mesh* M, const double2* eo, const SegMap& out, int B) {
  dim3 grd((M->half + G - 1) / G, B);
  size_t lds = (size_t)(2 * M->nm + 1) * G * sizeof(double2);
  const double2* tw = reinterpret_cast<const double2*>(M->d_twiddle);
  if (out.nseg == 1 && out.flag[0] == 0)
    hipLaunchKernelGGL((unfold_four_step_kernel<P, G, true>), grd, dim3(P * G), lds, M->stream,
                       eo, out, tw, M->d_cos, M->nlat, M->nm);
  else
    hipLaunchKernelGGL((unfold_four_step_kernel<P, G, false>), grd, dim3(P * G), lds, M->stream,
                       eo, out, tw, M->d_cos, M->nlat, M->nm);
}

// wavefronts per workgroup W = P G / 64 from the requested latitude pairs
#define JGCM_FS_DISPATCH(P_, G_, CALL)                                         \
  switch (P_ * 100 + G_) {                                                     \
    case 1604: { constexpr int PP = 16, GG = 4; CALL; } break;                 \
    case 1608: { constexpr int PP = 16, GG = 8; CALL; } break;                 \
    case 1616: { constexpr int PP = 16, GG = 16; CALL; } break;                \
    case 808: { constexpr int PP = 8, GG = 8; CALL; } break;                   \
    case 816: { constexpr int PP = 8, GG = 16; CALL; } break;                  \
    case 832: { constexpr int PP = 8, GG = 32; CALL; } break;                  \
    default: break;                                                            \
  }

static void fold_four_step(const jgcm_mesh* M, const SegMap& in, double2* eo, int B) {
  const int P = M->nlon / 16, G = four_step_pairs(M, M->fold_ppw);
  JGCM_FS_DISPATCH(P, G, (launch_fold_fs<PP, GG>(M, in, eo, B)));
}

static void unfold_four_step(const jgcm_mesh* M, const double2* eo, const SegMap& out, int B) {
  const int P = M->nlon / 16, G = four_step_pairs(M, M->unfold_ppw);
  JGCM_FS_DISPATCH(P, G, (launch_unfold_fs<PP, GG>(M, eo, out, B)));
}

template <int L>
static void launch_fold(const jgcm_mesh* M, const SegMap& in, double2* eo, int B) {
  int ppw = M->fold_ppw > 0 ? M->fold_ppw : pick_ppw(L, B, true);
  dim3 grd((M->half + ppw - 1) / ppw, B);
  size_t lds = (size_t)(M->nm * 2 + 1) * ppw * sizeof(double2);
  const double2* tw = reinterpret_cast<const double2*>(M->d_twiddle);
  if (in.nseg == 1 && in.flag[0] == 0)
    hipLaunchKernelGGL((fft_fold_kernel<L, true>), grd, dim3(256), lds, M->stream, in, eo, tw,
                       M->d_cos, M->nlat, M->nm, ppw);
  else
    hipLaunchKernelGGL((fft_fold_kernel<L, false>), grd, dim3(256), lds, M->stream, in, eo, tw,
                       M->d_cos, M->nlat, M->nm, ppw);
}

template <int L>
static void launch_unfold(const jgcm_mesh* M, const double2* eo, const SegMap& out, int B) {
  int ppw = M->unfold_ppw > 0 ? M->unfold_ppw : pick_ppw(L, B, false);
  dim3 grd((M->half + ppw - 1) / ppw, B);
  size_t lds = (size_t)(M->nm * 2 + 1) * ppw * sizeof(double2);
  const double2* tw = reinterpret_cast<const double2*>(M->d_twiddle);
  if (out.nseg == 1 && out.flag[0] == 0)
    hipLaunchKernelGGL((ifft_unfold_kernel<L, true>), grd, dim3(256), lds, M->stream, eo, out, tw,
                       M->d_cos, M->nlat, M->nm, ppw);
  else
    hipLaunchKernelGGL((ifft_unfold_kernel<L, false>), grd, dim3(256), lds, M->stream, eo, out, tw,
                       M->d_cos, M->nlat, M->nm, ppw);
}

template <int L>
static void launch_full(const jgcm_mesh* M, const double* grid, double2* four, int B) {
  constexpr int FPW = 256 / (L / 4);
  dim3 grd((M->half + FPW - 1) / FPW, B);
  hipLaunchKernelGGL(fft_full_kernel<L>, grd, dim3(256), 0, M->stream, grid, four,
                     reinterpret_cast<const double2*>(M->d_twiddle), M->nlat);
}

#define JGCM_DISPATCH_L(L_, CALL)                                              \
  switch (L_) {                                                                \
    case 16: { constexpr int LL = 16; CALL; } break;                           \
    case 32: { constexpr int LL = 32; CALL; } break;                           \
    case 64: { constexpr int LL = 64; CALL; } break;                           \
    case 128: { constexpr int LL = 128; CALL; } break;                         \
    case 256: { constexpr int LL = 256; CALL; } break;                         \
    case 512: { constexpr int LL = 512; CALL; } break;                         \
    case 1024: { constexpr int LL = 1024; CALL; } break;                       \
    default: break;                                                            \
  }

static int fft_fold_map(const jgcm_mesh* M, const SegMap& in, double2* eo) {
  if (use_four_step(M)) {
    fold_four_step(M, in, eo, in.start[in.nseg]);
  } else if (pow2_fft(M->nlon)) {
    JGCM_DISPATCH_L(M->nlon, launch_fold<LL>(M, in, eo, in.start[in.nseg]));
  } else {
    const size_t lds = (size_t)M->nlon * (sizeof(double2) + 2 * sizeof(double));
    hipLaunchKernelGGL(dft_fold_kernel, dim3(M->half, in.start[in.nseg]), dim3(256), lds,
                       M->stream, in, eo, reinterpret_cast<const double2*>(M->d_twiddle), M->d_cos,
                       M->nlat, M->nlon, M->nm);
  }
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

static int ifft_unfold_map(const jgcm_mesh* M, const double2* eo, const SegMap& out) {
  if (use_four_step(M)) {
    unfold_four_step(M, eo, out, out.start[out.nseg]);
  } else if (pow2_fft(M->nlon)) {
    JGCM_DISPATCH_L(M->nlon, launch_unfold<LL>(M, eo, out, out.start[out.nseg]));
  } else {
    const size_t lds = ((size_t)M->nlon + 2 * (size_t)M->nm) * sizeof(double2);
    hipLaunchKernelGGL(dft_unfold_kernel, dim3(M->half, out.start[out.nseg]), dim3(256), lds,
                       M->stream, eo, out, reinterpret_cast<const double2*>(M->d_twiddle),
                       M->d_cos, M->nlat, M->nlon, M->nm);
  }
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int fft_fold(const jgcm_mesh* M, const double* grid, double2* eo, const double* cosmul, int B) {
  return fft_fold_map(M, one_seg(grid, B, cosmul ? 1 : 0), eo);
}

int ifft_unfold(const jgcm_mesh* M, const double2* eo, double* grid, const double* divcos, int B) {
  return ifft_unfold_map(M, eo, one_seg(grid, B, divcos ? 1 : 0));
}

// ifft_unfold with the segment flag given (bits of ifft_unfold_kernel)
int ifft_unfold_flag(const jgcm_mesh* M, const double2* eo, double* grid, int flag, int B) {
  return ifft_unfold_map(M, eo, one_seg(grid, B, flag));
}

static FwdTables tables_for(const jgcm_mesh* M, int which) {
  const RowTables& r = M->rt;
  switch (which) {
    case 0: return {r.pfwd, r.pfwd_hi, r.pfwd_lo, r.pfwd_b3};
    case 1: return {r.alpa, r.alpa_hi, r.alpa_lo, r.alpa_b3};
    case 2: return {r.alpb, r.alpb_hi, r.alpb_lo, r.alpb_b3};
    default: return {r.pinv, r.pinv_hi, r.pinv_lo, r.pinv_b3};
  }
}

// LDS of the Legendre kernels: operand block [64][KC+1] + table block [64][KC+1]
static size_t legendre_lds(int precision, int tl) {
  const size_t trows = 16 * (size_t)tl;
  size_t e = precision == JGCM_PREC_FP32 ? sizeof(float) : 8;  // double, float2
  size_t pf = trows * KS1, pi = (size_t)KC * (trows + 1);
  size_t operands = align16(e * LCOLS * KS1) + e * (pf > pi ? pf : pi);
  size_t out_tile = (size_t)(LCOLS / 2) * trows * sizeof(double2);  // inverse epilogue
  return operands > out_tile ? operands : out_tile;
}

static int tiles_of(int t) { return t > 0 ? t : DEF_TILES; }

static int launch_legendre_fwd(const jgcm_mesh* M, FwdJobs& jobs, int nmax, int zero_top) {
  int z = 0;
  for (int j = 0; j < jobs.njobs; ++j) {
    jobs.zstart[j] = z;
    z += (jobs.job[j].nf + LCOLS / 2 - 1) / (LCOLS / 2);
  }
  jobs.zstart[jobs.njobs] = z;
  if (z == 0) return JGCM_OK;
  const int tl = tiles_of(M->leg_tiles_fwd);
  int maxrows = nmax / 2 + 1;
  int rgroups = ((maxrows + 15) / 16 + tl - 1) / tl;
  dim3 grd(2 * M->nm, rgroups, z);
  size_t lds = legendre_lds(M->precision, tl);
#define JGCM_FWD_LAUNCH(P, T)                                                                   \
  hipLaunchKernelGGL((legendre_fwd_mfma<P, T>), grd, dim3(LTHREADS), lds, M->stream, jobs,     \
                     M->ns, M->nm, M->half, nmax, zero_top)
#define JGCM_FWD_TILES(P)                                                                       \
  do {                                                                                          \
    if (tl == 1) JGCM_FWD_LAUNCH(P, 1);                                                         \
    else if (tl == 4) JGCM_FWD_LAUNCH(P, 4);                                                    \
    else JGCM_FWD_LAUNCH(P, 2);                                                                 \
  } while (0)
  switch (M->precision) {
    case JGCM_PREC_FP64: JGCM_FWD_TILES(P_FP64); break;
    case JGCM_PREC_FP32: JGCM_FWD_TILES(P_FP32); break;
    case JGCM_PREC_BF16X3: JGCM_FWD_TILES(P_BF16); break;
    default: JGCM_FWD_TILES(P_SPLIT); break;
  }
#undef JGCM_FWD_TILES
#undef JGCM_FWD_LAUNCH
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

static int legendre_inv_map(const jgcm_mesh* M, const SegMap& snm, double2* eo, int table,
                            int nmax, int out_mode = 0) {
  int B = snm.start[snm.nseg];
  const int tl = tiles_of(M->leg_tiles_inv);
  const int trows = 16 * tl;
  int fchunks = (B + LCOLS / 2 - 1) / (LCOLS / 2);
  dim3 grd(2 * M->nm, fchunks, (M->half + trows - 1) / trows);
  FwdTables tab = tables_for(M, table);
  size_t lds = legendre_lds(M->precision, tl);
  SpecOpTables ops{M->d_dlam, M->d_dthm, M->d_dthp, M->d_uvc, M->d_uvm, M->d_uvp};
  const bool plain = snm.nseg == 1 && ((snm.flag[0] >> 1) & 7) == SPEC_OP_NONE;
#define JGCM_INV_LAUNCH1(P, PL, T)                                                              \
  hipLaunchKernelGGL((legendre_inv_mfma<P, PL, T>), grd, dim3(LTHREADS), lds, M->stream, snm,   \
                     eo, tab, B, nmax, M->ns, M->nm, M->half, ops, out_mode)
#define JGCM_INV_TILES(P, PL)                                                                   \
  do {                                                                                          \
    if (tl == 1) JGCM_INV_LAUNCH1(P, PL, 1);                                                    \
    else if (tl == 4) JGCM_INV_LAUNCH1(P, PL, 4);                                               \
    else JGCM_INV_LAUNCH1(P, PL, 2);                                                            \
  } while (0)
#define JGCM_INV_LAUNCH(P)                                                                      \
  do {                                                                                          \
    if (plain) JGCM_INV_TILES(P, true);                                                         \
    else JGCM_INV_TILES(P, false);                                                              \
  } while (0)
  switch (M->precision) {
    case JGCM_PREC_FP64: JGCM_INV_LAUNCH(P_FP64); break;
    case JGCM_PREC_FP32: JGCM_INV_LAUNCH(P_FP32); break;
    case JGCM_PREC_BF16X3: JGCM_INV_LAUNCH(P_BF16); break;
    default: JGCM_INV_LAUNCH(P_SPLIT); break;
  }
#undef JGCM_INV_LAUNCH
#undef JGCM_INV_TILES
#undef JGCM_INV_LAUNCH1
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int legendre_inv_t(const jgcm_mesh* M, const double2* snm, double2* eo, int B, int table,
                   int nmax) {
  return legendre_inv_map(M, one_seg(snm, B, 0), eo, table, nmax);
}

int legendre_inv(const jgcm_mesh* M, const double2* snm, double2* eo, int B) {
  return legendre_inv_t(M, snm, eo, B, 3, M->ns);
}

// Transposed α-operator contraction: eo (+)= table^T (op h) over rows n <= nmax,
// written to the parity block of the rows (or the opposite one).
int legendre_alpha_transpose(const jgcm_mesh* M, const double2* h, double2* eo, int B, int table,
                             int op, int swap, int accumulate) {
  SegMap in = one_seg(h, B, op << 1);
  return legendre_inv_map(M, in, eo, table, M->N, (swap ? 1 : 0) | (accumulate ? 2 : 0));
}

static FwdJob plain_job(const jgcm_mesh* M, const double2* eo, int nf, const SegMap& out) {
  return FwdJob{eo, tables_for(M, 0), 0, nullptr, FwdTables{}, 0.0, nf, out};
}

int legendre_fwd_plain(const jgcm_mesh* M, const double2* eo, double2* snm, int B) {
  FwdJobs jobs{};
  jobs.njobs = 1;
  jobs.job[0] = plain_job(M, eo, B, one_seg(snm, B, 0));
  return launch_legendre_fwd(M, jobs, M->N, 1);
}

// vor = α(V, U, -1), div = α(U, V, +1) with the top degree zeroed (SSM:387-392)
static void vordiv_jobs(const jgcm_mesh* M, const double2* eoU, const double2* eoV, int B,
                        const SegMap& vor, const SegMap& div, FwdJob* out2) {
  out2[0] = FwdJob{eoV, tables_for(M, 1), 1, eoU, tables_for(M, 2), -1.0, B, vor};
  out2[1] = FwdJob{eoU, tables_for(M, 1), 1, eoV, tables_for(M, 2), +1.0, B, div};
}

int legendre_vordiv(const jgcm_mesh* M, const double2* eoU, const double2* eoV, double2* vor,
                    double2* div, int B) {
  FwdJobs jobs{};
  jobs.njobs = 2;
  vordiv_jobs(M, eoU, eoV, B, one_seg(vor, B, 0), one_seg(div, B, 0), jobs.job);
  return launch_legendre_fwd(M, jobs, M->N, 1);
}

int legendre_alpha(const jgcm_mesh* M, const double2* eoA, const double2* eoB, double isign,
                   double2* alpha, int B) {
  FwdJobs jobs{};
  jobs.njobs = 1;
  jobs.job[0] = FwdJob{eoA, tables_for(M, 1), 1, eoB, tables_for(M, 2), isign, B, one_seg(alpha, B, 0)};
  return launch_legendre_fwd(M, jobs, M->ns, 0);
}

// transpose of the inverse Legendre stage: s[n] = sum_j' Pinv[n][j'] eo^(n-m)[j'], n = m..ns,
// times scale0 (m = 0) or scale1 (m > 0), stored or added to the output
int legendre_inv_adjoint(const jgcm_mesh* M, const double2* eo, const SegMap& out, int B,
                         double scale0, double scale1, int accumulate) {
  FwdJobs jobs{};
  jobs.njobs = 1;
  jobs.job[0] = FwdJob{eo, tables_for(M, 3), 0, nullptr, FwdTables{}, 0.0, B, out};
  jobs.job[0].scale0 = scale0;
  jobs.job[0].scale1 = scale1;
  jobs.job[0].accumulate = accumulate;
  return launch_legendre_fwd(M, jobs, M->ns, 0);
}

int fft_full(const jgcm_mesh* M, const double* grid, double2* four, int B) {
  if (pow2_fft(M->nlon)) {
    JGCM_DISPATCH_L(M->nlon, launch_full<LL>(M, grid, four, B));
  } else {
    const size_t lds = (size_t)M->nlon * (sizeof(double2) + sizeof(double));
    hipLaunchKernelGGL(dft_full_kernel, dim3(M->nlat, B), dim3(256), lds, M->stream, grid, four,
                       reinterpret_cast<const double2*>(M->d_twiddle), M->nlat, M->nlon);
  }
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int fold_full(const jgcm_mesh* M, const double2* four, double2* eo, int B) {
  size_t total = (size_t)B * M->nm * M->half;
  int blocks = (int)std::min<size_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(fold_full_kernel, dim3(blocks), dim3(256), 0, M->stream, four, eo, M->nlat,
                     M->nlon, M->nm, B);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// Segment maps of a batched back-transform: `in` the spectral arrays with
// their operators, `out` the grid arrays with their ÷cos θ flags.  Returns
// the number of level-fields.
static int build_inverse(int nseg, const double* const* snm, const double* const* snm_b,
                         double* const* grid, const int* nlev, const int* flags, SegMap& in,
                         SegMap& out) {
  in = SegMap{};
  out = SegMap{};
  in.nseg = out.nseg = nseg;
  int B = 0;
  for (int k = 0; k < nseg; ++k) {
    in.start[k] = out.start[k] = B;
    in.ptr[k] = const_cast<double*>(snm[k]);
    in.ptr2[k] = snm_b ? snm_b[k] : nullptr;
    out.ptr[k] = grid[k];
    in.flag[k] = flags ? (flags[k] & ~1) : 0;
    out.flag[k] = flags ? (flags[k] & 1) : 0;
    B += nlev[k];
  }
  in.start[nseg] = out.start[nseg] = B;
  return B;
}

// Back-transform of several spectral arrays in one Legendre launch and one FFT
// launch; flag & 1 divides the rows of that array by cos θ, flag >> 1 names a
// spectral operator applied on the way in (SPEC_OP_*, second operand snm_b[k]).
int inverse_batch(const jgcm_mesh* M, int nseg, const double* const* snm,
                  const double* const* snm_b, double* const* grid, const int* nlev,
                  const int* flags) {
  SegMap in, out;
  build_inverse(nseg, snm, snm_b, grid, nlev, flags, in, out);
  double2* eo = reinterpret_cast<double2*>(M->d_four_a);
  if (int rc = legendre_inv_map(M, in, eo, 3, M->ns)) return rc;
  return ifft_unfold_map(M, eo, out);
}

// Segment map of a batched forward transform's grid inputs and its Legendre
// jobs reading the E/O block `eo`.  flags[k] == JGCM_BATCH_VORDIV pairs array
// k (u) with array k+1 (v) and writes vor to snm[k], div to snm[k+1]
// (Vor_Div_From_Grid_UV!); other arrays get the plain transform, and
// consecutive plain arrays share one job.
static int build_forward(const jgcm_mesh* M, int nseg, const double* const* grid,
                         double* const* snm, const int* nlev, const int* flags, double2* eo,
                         SegMap& in, FwdJobs& jobs) {
  in = SegMap{};
  in.nseg = nseg;
  int B = 0;
  for (int k = 0; k < nseg; ++k) {
    in.start[k] = B;
    in.ptr[k] = const_cast<double*>(grid[k]);
    int f = flags ? flags[k] : 0;
    bool uv = (f == JGCM_BATCH_VORDIV) || (k > 0 && flags && flags[k - 1] == JGCM_BATCH_VORDIV);
    in.flag[k] = uv ? 1 : 0;
    B += nlev[k];
  }
  in.start[nseg] = B;
  const size_t eo_field = (size_t)M->nm * 2 * M->half;
  jobs = FwdJobs{};
  int nj = 0;
  for (int k = 0; k < nseg;) {
    if (nj >= MAXJOB) {
      set_error("batched forward transform needs more than %d Legendre jobs", MAXJOB);
      return JGCM_ERR_ARG;
    }
    if (flags && flags[k] == JGCM_BATCH_VORDIV) {
      if (nj + 2 > MAXJOB) {
        set_error("batched forward transform needs more than %d Legendre jobs", MAXJOB);
        return JGCM_ERR_ARG;
      }
      vordiv_jobs(M, eo + in.start[k] * eo_field, eo + in.start[k + 1] * eo_field, nlev[k],
                  one_seg(snm[k], nlev[k], 0), one_seg(snm[k + 1], nlev[k], 0), jobs.job + nj);
      nj += 2;
      k += 2;
      continue;
    }
    // a run of plain arrays: one job, outputs through a segment map
    SegMap out{};
    int k0 = k, nf = 0;
    while (k < nseg && !(flags && flags[k] == JGCM_BATCH_VORDIV)) {
      out.start[k - k0] = nf;
      out.ptr[k - k0] = snm[k];
      nf += nlev[k];
      ++k;
    }
    out.nseg = k - k0;
    out.start[out.nseg] = nf;
    jobs.job[nj++] = plain_job(M, eo + in.start[k0] * eo_field, nf, out);
  }
  jobs.njobs = nj;
  int z = 0;
  for (int j = 0; j < nj; ++j) {
    jobs.zstart[j] = z;
    z += (jobs.job[j].nf + LCOLS / 2 - 1) / (LCOLS / 2);
  }
  jobs.zstart[nj] = z;
  return JGCM_OK;
}

// Forward transform of several grid arrays in one FFT launch and one Legendre
// launch (build_forward gives the arrays' roles).
int forward_batch(const jgcm_mesh* M, int nseg, const double* const* grid, double* const* snm,
                  const int* nlev, const int* flags) {
  double2* eo = reinterpret_cast<double2*>(M->d_four_a);
  SegMap in;
  FwdJobs jobs;
  if (int rc = build_forward(M, nseg, grid, snm, nlev, flags, eo, in, jobs)) return rc;
  if (int rc = fft_fold_map(M, in, eo)) return rc;
  return launch_legendre_fwd(M, jobs, M->N, 1);
}

// ------------------------------------------------ plans: several meshes per launch
// A plan holds the batched transform of several meshes that share one stream
// (an ensemble's truncations) and runs the Legendre stage of all of them in
// ONE launch, driven by a work list.  A work item is one tile of a one-mesh
// Legendre launch: (mesh, segment (m, parity), y, z), with y, z the grid
// coordinates that tile has in the one-mesh launch, so every tile computes
// exactly what it computes there.  The items are sorted by an estimate of
// their cost and dealt out longest-first to the least-loaded workgroup
// (persistent workgroups, each running its own item list) or, with one
// workgroup per item, launched heaviest first.

constexpr int PLAN_MAXMESH = 8;

struct InvDesc {
  SegMap in;
  double2* eo;
  FwdTables tab;
  SpecOpTables ops;
  int B, nmax, ns, nm, half, tl;
};

struct FwdDesc {
  FwdJobs jobs;
  int ns, nm, half, nmax, tl;
};

template <int PREC>
__global__ __launch_bounds__(LTHREADS) void legendre_inv_worklist(const InvDesc* __restrict__ desc,
                                                             const int4* __restrict__ items,
                                                             const int* __restrict__ wg_start) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int end = wg_start[blockIdx.x + 1];
  for (int i = wg_start[blockIdx.x]; i < end; ++i) {
    const int4 it = items[i];
    const InvDesc& d = desc[it.x];
    const uint3 wg = make_uint3(it.y, it.z, it.w);
    // each tile starts with a barrier before its first LDS store, so the
    // previous tile's epilogue has read the LDS it reuses
    if (d.tl == 1)
      legendre_inv_tile<PREC, false, 1>(d.in, d.eo, d.tab, d.B, d.nmax, d.ns, d.nm, d.half,
                                        d.ops, 0, wg, smem_raw);
    else if (d.tl == 4)
      legendre_inv_tile<PREC, false, 4>(d.in, d.eo, d.tab, d.B, d.nmax, d.ns, d.nm, d.half,
                                        d.ops, 0, wg, smem_raw);
    else
      legendre_inv_tile<PREC, false, 2>(d.in, d.eo, d.tab, d.B, d.nmax, d.ns, d.nm, d.half,
                                        d.ops, 0, wg, smem_raw);
  }
}

template <int PREC>
__global__ __launch_bounds__(LTHREADS) void legendre_fwd_worklist(const FwdDesc* __restrict__ desc,
                                                             const int4* __restrict__ items,
                                                             const int* __restrict__ wg_start) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int end = wg_start[blockIdx.x + 1];
  for (int i = wg_start[blockIdx.x]; i < end; ++i) {
    const int4 it = items[i];
    const FwdDesc& d = desc[it.x];
    const uint3 wg = make_uint3(it.y, it.z, it.w);
    if (d.tl == 1)
      legendre_fwd_tile<PREC, 1>(d.jobs, d.ns, d.nm, d.half, d.nmax, 1, wg, smem_raw);
    else if (d.tl == 4)
      legendre_fwd_tile<PREC, 4>(d.jobs, d.ns, d.nm, d.half, d.nmax, 1, wg, smem_raw);
    else
      legendre_fwd_tile<PREC, 2>(d.jobs, d.ns, d.nm, d.half, d.nmax, 1, wg, smem_raw);
  }
}

// Longest processing time first: items in decreasing cost (ties in index
// order), each to the workgroup with the least work so far (ties: the lowest
// workgroup).  order = the item indices grouped by workgroup, each group in
// assignment order; wg_start[w] .. wg_start[w+1] is workgroup w's range.
// nwg <= 0 or >= n: one workgroup per item, heaviest first.
int lpt_schedule(const std::vector<double>& cost, int nwg, std::vector<int>& order,
                 std::vector<int>& wg_start) {
  const int n = (int)cost.size();
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  if (nwg <= 0 || nwg >= n) {
    order = idx;
    wg_start.resize(n + 1);
    for (int i = 0; i <= n; ++i) wg_start[i] = i;
    return n;
  }
  using Load = std::pair<double, int>;
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
  for (int w = 0; w < nwg; ++w) heap.push({0.0, w});
  std::vector<std::vector<int>> bins(nwg);
  for (int i : idx) {
    Load l = heap.top();
    heap.pop();
    bins[l.second].push_back(i);
    heap.push({l.first + cost[i], l.second});
  }
  order.clear();
  wg_start.assign(1, 0);
  for (int w = 0; w < nwg; ++w) {
    order.insert(order.end(), bins[w].begin(), bins[w].end());
    wg_start.push_back((int)order.size());
  }
  return nwg;
}

}  // namespace jgcm

struct jgcm_plan {
  int direction = 0;  // 0: spherical -> grid, 1: grid -> spherical
  int nmesh = 0;
  int precision = 0;
  jgcm_mesh* mesh[jgcm::PLAN_MAXMESH] = {};
  jgcm::SegMap fft[jgcm::PLAN_MAXMESH];    // the FFT stage's grid arrays per mesh
  double2* eo[jgcm::PLAN_MAXMESH] = {};     // each mesh's E/O block
  void* d_desc = nullptr;                   // InvDesc[] or FwdDesc[]
  int4* d_items = nullptr;
  int* d_wg_start = nullptr;
  int nitems = 0, nwg = 0;
  size_t lds = 0;
  std::vector<int4> items;                  // host copy in launch order
  std::vector<int> wg_start;
};

namespace jgcm {

static int plan_occupancy(int direction, int precision, size_t lds, int* per_cu) {
  const void* k = nullptr;
  switch (precision) {
#define JGCM_PLAN_K(P)                                                                       \
  k = direction == 0 ? reinterpret_cast<const void*>(legendre_inv_worklist<P>)               \
                     : reinterpret_cast<const void*>(legendre_fwd_worklist<P>)
    case JGCM_PREC_FP64: JGCM_PLAN_K(P_FP64); break;
    case JGCM_PREC_FP32: JGCM_PLAN_K(P_FP32); break;
    case JGCM_PREC_BF16X3: JGCM_PLAN_K(P_BF16); break;
    default: JGCM_PLAN_K(P_SPLIT); break;
#undef JGCM_PLAN_K
  }
  JGCM_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, LTHREADS, lds));
  if (*per_cu < 1) *per_cu = 1;
  return JGCM_OK;
}

static void free_plan(jgcm_plan* p) {
  if (!p) return;
  jgcm_mesh* m0 = p->mesh[0];
  void* bufs[] = {p->d_desc, p->d_items, p->d_wg_start};
  if (m0) {
    (void)hipStreamSynchronize(m0->stream);
    // graphs that replay this plan's launches keep its buffers' addresses:
    // while any graph of the first mesh lives, the buffers go with the mesh
    const bool graphs = m0->live_graphs.load() > m0->live_plans.load();
    for (void* b : bufs)
      if (b) {
        if (graphs) m0->retired.push_back(b);
        else (void)hipFree(b);
      }
  } else {
    for (void* b : bufs)
      if (b) (void)hipFree(b);
  }
  for (int g = 0; g < p->nmesh; ++g) {
    jgcm_mesh* m = p->mesh[g];
    if (!m) continue;
    m->live_plans.fetch_sub(1);
    m->live_graphs.fetch_sub(1);
    mesh_release(m);
  }
  delete p;
}

// Shared part of the two create calls: meshes checked (one stream, one
// precision, each once, no capture open) and retained; the plan's scratch
// pinned like a graph's (growth keeps the buffers the plan addresses).
static int plan_begin(int direction, int nmesh, jgcm_mesh_t* const* meshes, jgcm_plan** out) {
  JGCM_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  JGCM_REQUIRE(nmesh >= 1 && nmesh <= PLAN_MAXMESH, "nmesh must be in 1..%d, got %d",
               PLAN_MAXMESH, nmesh);
  JGCM_REQUIRE(meshes != nullptr, "meshes is NULL");
  if (capture_ctx() != nullptr) {
    set_error("a plan allocates and uploads its work list: create it before capturing");
    return JGCM_ERR_STATE;
  }
  for (int g = 0; g < nmesh; ++g) {
    if (int rc = check_mesh(meshes[g])) return rc;
    JGCM_REQUIRE(meshes[g]->stream == meshes[0]->stream,
                 "mesh %d does not share mesh 0's stream (jgcm_mesh_share_stream)", g);
    JGCM_REQUIRE(meshes[g]->precision == meshes[0]->precision,
                 "mesh %d has another Legendre precision than mesh 0", g);
    for (int q = 0; q < g; ++q)
      JGCM_REQUIRE(meshes[q] != meshes[g], "mesh %d appears twice (one batch per mesh)", g);
  }
  jgcm_plan* p = new (std::nothrow) jgcm_plan();
  if (!p) {
    set_error("out of host memory");
    return JGCM_ERR_NOMEM;
  }
  p->direction = direction;
  p->precision = meshes[0]->precision;
  *out = p;
  return JGCM_OK;
}

static void plan_hold(jgcm_plan* p, jgcm_mesh* m) {
  mesh_retain(m);
  m->live_graphs.fetch_add(1);
  m->live_plans.fetch_add(1);
  p->mesh[p->nmesh++] = m;
}

// Work list from the items' costs, uploaded with the descriptors.
static int plan_finish(jgcm_plan* p, const void* desc, size_t desc_bytes,
                       const std::vector<int4>& items, const std::vector<double>& cost,
                       int workgroups, int tlmax) {
  p->lds = legendre_lds(p->precision, tlmax);
  int nwg = workgroups;
  if (workgroups < 0) {
    int per_cu = 1, cus = 1;
    if (int rc = plan_occupancy(p->direction, p->precision, p->lds, &per_cu)) return rc;
    JGCM_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                         p->mesh[0]->device));
    nwg = per_cu * cus;
  }
  std::vector<int> order;
  p->nwg = lpt_schedule(cost, nwg, order, p->wg_start);
  p->nitems = (int)items.size();
  p->items.resize(items.size());
  for (size_t i = 0; i < order.size(); ++i) p->items[i] = items[order[i]];
  JGCM_CHECK_HIP(hipMalloc(&p->d_desc, desc_bytes));
  JGCM_CHECK_HIP(hipMalloc(&p->d_items, sizeof(int4) * std::max<size_t>(1, p->items.size())));
  JGCM_CHECK_HIP(hipMalloc(&p->d_wg_start, sizeof(int) * p->wg_start.size()));
  JGCM_CHECK_HIP(hipMemcpy(p->d_desc, desc, desc_bytes, hipMemcpyHostToDevice));
  JGCM_CHECK_HIP(hipMemcpy(p->d_items, p->items.data(), sizeof(int4) * p->items.size(),
                           hipMemcpyHostToDevice));
  JGCM_CHECK_HIP(hipMemcpy(p->d_wg_start, p->wg_start.data(), sizeof(int) * p->wg_start.size(),
                           hipMemcpyHostToDevice));
  return JGCM_OK;
}

// Cost estimates of one tile in staged elements: both operand blocks over the
// contraction, the MFMA steps, the stores and a fixed part per tile.
static double inv_item_cost(int nr, int nf, int nlg) {
  const int nlt = (nlg + 15) / 16;
  return 256.0 + (double)nr * (nf + nlg + 4 * nlt) + (double)nf * nlg;
}

static double fwd_item_cost(int half, int terms, int nf, int nrg) {
  const int ntile = (nrg + 15) / 16;
  return 256.0 + (double)terms * half * (nf + nrg + 4 * ntile) + (double)nf * nrg;
}

static int plan_launch_legendre(const jgcm_plan* p) {
  if (p->nitems == 0) return JGCM_OK;
  hipStream_t st = p->mesh[0]->stream;
  const int4* it = p->d_items;
  const int* ws = p->d_wg_start;
#define JGCM_PLAN_LAUNCH(P)                                                                   \
  do {                                                                                        \
    if (p->direction == 0)                                                                    \
      hipLaunchKernelGGL(legendre_inv_worklist<P>, dim3(p->nwg), dim3(LTHREADS), p->lds, st,  \
                         reinterpret_cast<const InvDesc*>(p->d_desc), it, ws);                \
    else                                                                                      \
      hipLaunchKernelGGL(legendre_fwd_worklist<P>, dim3(p->nwg), dim3(LTHREADS), p->lds, st,  \
                         reinterpret_cast<const FwdDesc*>(p->d_desc), it, ws);                \
  } while (0)
  switch (p->precision) {
    case JGCM_PREC_FP64: JGCM_PLAN_LAUNCH(P_FP64); break;
    case JGCM_PREC_FP32: JGCM_PLAN_LAUNCH(P_FP32); break;
    case JGCM_PREC_BF16X3: JGCM_PLAN_LAUNCH(P_BF16); break;
    default: JGCM_PLAN_LAUNCH(P_SPLIT); break;
  }
#undef JGCM_PLAN_LAUNCH
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

}  // namespace jgcm

namespace jgcm {

// Argument checks of the batched transforms (the one-mesh calls and plans);
// *total = the level-fields of the batch.
static int check_inverse_args(int nseg, const double* const* snm, const double* const* snm_b,
                              double* const* pfield, const int* nlev, const int* flags,
                              int* total) {
  JGCM_REQUIRE(nseg >= 1 && nseg <= MAXSEG, "nseg must be in 1..%d, got %d", MAXSEG, nseg);
  JGCM_REQUIRE(snm && pfield && nlev, "NULL argument");
  int t = 0;
  for (int k = 0; k < nseg; ++k) {
    JGCM_REQUIRE(snm[k] && pfield[k], "NULL field pointer (array %d)", k);
    JGCM_REQUIRE(nlev[k] >= 1, "nlev[%d] must be >= 1, got %d", k, nlev[k]);
    int f = flags ? flags[k] : 0;
    int op = f >> 1;
    JGCM_REQUIRE(f >= 0 && op <= SPEC_OP_VCOS, "flags[%d] = %d is not valid", k, f);
    JGCM_REQUIRE(op < SPEC_OP_UCOS || (snm_b && snm_b[k]),
                 "array %d: u cos / v cos needs the divergence in snm_b[%d]", k, k);
    t += nlev[k];
  }
  *total = t;
  return JGCM_OK;
}

static int check_forward_args(int nseg, const double* const* pfield, double* const* snm,
                              const int* nlev, const int* flags, int* total) {
  JGCM_REQUIRE(nseg >= 1 && nseg <= MAXSEG, "nseg must be in 1..%d, got %d", MAXSEG, nseg);
  JGCM_REQUIRE(snm && pfield && nlev, "NULL argument");
  int t = 0;
  for (int k = 0; k < nseg; ++k) {
    JGCM_REQUIRE(snm[k] && pfield[k], "NULL field pointer (array %d)", k);
    JGCM_REQUIRE(nlev[k] >= 1, "nlev[%d] must be >= 1, got %d", k, nlev[k]);
    int f = flags ? flags[k] : 0;
    JGCM_REQUIRE(f == 0 || f == JGCM_BATCH_VORDIV, "flags[%d] = %d is not valid", k, f);
    if (f == JGCM_BATCH_VORDIV) {
      JGCM_REQUIRE(k + 1 < nseg && flags[k + 1] == 0 && nlev[k + 1] == nlev[k],
                   "array %d (u) needs a following v array with flag 0 and the same nlev", k);
      t += nlev[k];
      ++k;
      JGCM_REQUIRE(snm[k] && pfield[k], "NULL field pointer (array %d)", k);
    }
    t += nlev[k];
  }
  *total = t;
  return JGCM_OK;
}

}  // namespace jgcm

using namespace jgcm;

#define JGCM_MESH_PROLOGUE(mesh, nlev)                                           \
  if (int _rc = check_mesh(mesh)) return _rc;                                    \
  JGCM_REQUIRE((nlev) >= 1, "nlev must be >= 1, got %d", (nlev));                \
  if (int _rc = ensure_scratch(mesh, (nlev))) return _rc;

extern "C" {

int jgcm_trans_grid_to_spherical(jgcm_mesh_t* mesh, const double* pfield, double* snm, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(pfield && snm, "NULL field pointer");
  double2* eo = reinterpret_cast<double2*>(mesh->d_four_a);
  if (int rc = fft_fold(mesh, pfield, eo, nullptr, nlev)) return rc;
  return legendre_fwd_plain(mesh, eo, reinterpret_cast<double2*>(snm), nlev);
}

int jgcm_trans_spherical_to_grid(jgcm_mesh_t* mesh, const double* snm, double* pfield, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(pfield && snm, "NULL field pointer");
  double2* eo = reinterpret_cast<double2*>(mesh->d_four_a);
  if (int rc = legendre_inv(mesh, reinterpret_cast<const double2*>(snm), eo, nlev)) return rc;
  return ifft_unfold(mesh, eo, pfield, nullptr, nlev);
}

}  // extern "C"

// ---- allocation-free forward transform ---------------------------------------
// The E/O Fourier block between the FFT and the Legendre stage of
// jgcm_trans_grid_to_spherical_noalloc lives in this file-scope device array
// (each GPU holds its own copy for the life of the process).  Host code never
// passes `scratch` itself to a kernel: its device address on the current GPU is
// looked up once per device with hipGetSymbolAddress and kept in
// g_noalloc_ptr.  Every call on one GPU shares the array: calls that overlap
// on one GPU (different streams, or several host threads) race.  Nothing in
// it needs clearing: the FFT stage writes every element the Legendre stage
// reads.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;   // 16 MiB
static __device__ __align__(256) double2 scratch[NOALLOC_SCRATCH_BYTES / sizeof(double2)];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static double2* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the mesh's GPU (which must be current).
static int noalloc_scratch(const jgcm_mesh* mesh, double2** out) {
  int dev = -1;
  JGCM_CHECK_HIP(hipGetDevice(&dev));
  if (dev != mesh->device) {
    set_error("the mesh lives on device %d but device %d is current", mesh->device, dev);
    return JGCM_ERR_STATE;
  }
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("device %d is beyond the %d devices the static scratch supports", dev,
              NOALLOC_MAX_DEVICES);
    return JGCM_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[dev] == nullptr) {
    void* p = nullptr;
    JGCM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[dev] = static_cast<double2*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return JGCM_OK;
}

extern "C" {

int jgcm_trans_grid_to_spherical_noalloc(jgcm_mesh_t* mesh, const double* pfield, double* snm,
                                         int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(pfield && snm, "NULL field pointer");
  const size_t per_field = mesh->four_elems() * sizeof(double2);
  const size_t max_fields = NOALLOC_SCRATCH_BYTES / per_field;
  JGCM_REQUIRE((size_t)nlev <= max_fields,
               "nlev = %d level-fields need %zu bytes of staging; the static scratch holds %zu "
               "bytes, at most %zu level-fields at T%d on %d latitudes",
               nlev, (size_t)nlev * per_field, NOALLOC_SCRATCH_BYTES, max_fields, mesh->N,
               mesh->nlat);
  double2* eo = nullptr;
  if (int rc = noalloc_scratch(mesh, &eo)) return rc;
  if (int rc = fft_fold(mesh, pfield, eo, nullptr, nlev)) return rc;
  return legendre_fwd_plain(mesh, eo, reinterpret_cast<double2*>(snm), nlev);
}

int jgcm_trans_grid_to_fourier(jgcm_mesh_t* mesh, const double* pfield, double* fourier_g, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(pfield && fourier_g, "NULL field pointer");
  return fft_full(mesh, pfield, reinterpret_cast<double2*>(fourier_g), nlev);
}

int jgcm_vor_div_from_grid_uv(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                              double* vor, double* div, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, 2 * nlev);
  JGCM_REQUIRE(grid_u && grid_v && vor && div, "NULL field pointer");
  const double* g[2] = {grid_u, grid_v};
  double* s[2] = {vor, div};
  int nl[2] = {nlev, nlev};
  int fl[2] = {JGCM_BATCH_VORDIV, 0};
  return forward_batch(mesh, 2, g, s, nl, fl);
}

int jgcm_trans_spherical_to_grid_batch(jgcm_mesh_t* mesh, int nseg, const double* const* snm,
                                       const double* const* snm_b, double* const* pfield,
                                       const int* nlev, const int* flags) {
  if (int rc = check_mesh(mesh)) return rc;
  int total = 0;
  if (int rc = check_inverse_args(nseg, snm, snm_b, pfield, nlev, flags, &total)) return rc;
  if (int rc = ensure_scratch(mesh, total)) return rc;
  return inverse_batch(mesh, nseg, snm, snm_b, pfield, nlev, flags);
}

int jgcm_trans_grid_to_spherical_batch(jgcm_mesh_t* mesh, int nseg, const double* const* pfield,
                                       double* const* snm, const int* nlev, const int* flags) {
  if (int rc = check_mesh(mesh)) return rc;
  int total = 0;
  if (int rc = check_forward_args(nseg, pfield, snm, nlev, flags, &total)) return rc;
  if (int rc = ensure_scratch(mesh, total)) return rc;
  return forward_batch(mesh, nseg, pfield, snm, nlev, flags);
}

// Adjoint (transpose w.r.t. sum_ij x y on the grid and sum Re(conj(a) b) over
// the spectral array) of Trans_Grid_To_Spherical!: weighted-transpose
// Legendre (qwg/2 rows, degrees m..N), per-m scaling, conjugate FFT.
int jgcm_trans_grid_to_spherical_adjoint(jgcm_mesh_t* mesh, const double* snm, double* pfield,
                                         int nlev) {
  return jgcm_trans_grid_to_spherical_adjoint_batch(mesh, 1, &snm, &pfield, &nlev);
}

// The same for up to MAXSEG arrays in one Legendre launch and one FFT launch
// (the weights applied while the FFT stages its input).
int jgcm_trans_grid_to_spherical_adjoint_batch(jgcm_mesh_t* mesh, int nseg,
                                               const double* const* snm, double* const* pfield,
                                               const int* nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nseg >= 1 && nseg <= MAXSEG, "nseg must be in 1..%d, got %d", MAXSEG, nseg);
  JGCM_REQUIRE(snm && pfield && nlev, "NULL argument");
  int total = 0;
  for (int k = 0; k < nseg; ++k) {
    JGCM_REQUIRE(snm[k] && pfield[k], "NULL field pointer (array %d)", k);
    JGCM_REQUIRE(nlev[k] >= 1, "nlev[%d] must be >= 1, got %d", k, nlev[k]);
    total += nlev[k];
  }
  if (int rc = ensure_scratch(mesh, total)) return rc;
  SegMap in{}, out{};
  in.nseg = out.nseg = nseg;
  int B = 0;
  for (int k = 0; k < nseg; ++k) {
    in.start[k] = out.start[k] = B;
    in.ptr[k] = const_cast<double*>(snm[k]);
    out.ptr[k] = pfield[k];
    out.flag[k] = 2;  // transposed Hermitian weights
    B += nlev[k];
  }
  in.start[nseg] = out.start[nseg] = B;
  double2* eo = reinterpret_cast<double2*>(mesh->d_four_a);
  if (int rc = legendre_inv_map(mesh, in, eo, 0, mesh->N)) return rc;
  return ifft_unfold_map(mesh, eo, out);
}

// Adjoint of Trans_Spherical_To_Grid!: FFT with the transposed Hermitian
// weights (L for m = 0, 2L for m > 0, applied to the contraction's output),
// then the transposed inverse Legendre over degrees m..N+1.  The lower
// triangle (n < m) of the output is set to zero.
int jgcm_trans_spherical_to_grid_adjoint(jgcm_mesh_t* mesh, const double* pfield, double* snm,
                                         int nlev) {
  return jgcm_trans_spherical_to_grid_adjoint_batch(mesh, 1, &pfield, &snm, &nlev, nullptr, 0);
}

// The same for up to MAXSEG arrays in one FFT launch and one Legendre launch.
// flags[k] & JGCM_BATCH_DIVCOS: array k is divided by cos θ on the way in
// (the transpose of the inverse batch's ÷cos output).  accumulate: the
// results are added to snm[k] (degrees m..N+1 only; the outputs must not
// overlap); otherwise snm[k] is overwritten and its other entries zeroed.
int jgcm_trans_spherical_to_grid_adjoint_batch(jgcm_mesh_t* mesh, int nseg,
                                               const double* const* pfield, double* const* snm,
                                               const int* nlev, const int* flags, int accumulate) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nseg >= 1 && nseg <= MAXSEG, "nseg must be in 1..%d, got %d", MAXSEG, nseg);
  JGCM_REQUIRE(snm && pfield && nlev, "NULL argument");
  const size_t spec_d = mesh->spec_elems() * 2;
  int total = 0;
  for (int k = 0; k < nseg; ++k) {
    JGCM_REQUIRE(snm[k] && pfield[k], "NULL field pointer (array %d)", k);
    JGCM_REQUIRE(nlev[k] >= 1, "nlev[%d] must be >= 1, got %d", k, nlev[k]);
    int f = flags ? flags[k] : 0;
    JGCM_REQUIRE(f == 0 || f == JGCM_BATCH_DIVCOS, "flags[%d] = %d is not valid", k, f);
    if (accumulate) {
      for (int q = 0; q < k; ++q) {
        const double *a0 = snm[q], *a1 = snm[q] + nlev[q] * spec_d;
        const double *b0 = snm[k], *b1 = snm[k] + nlev[k] * spec_d;
        JGCM_REQUIRE(a1 <= b0 || b1 <= a0, "accumulated outputs %d and %d overlap", q, k);
      }
    }
    total += nlev[k];
  }
  if (int rc = ensure_scratch(mesh, total)) return rc;
  SegMap in{}, out{};
  in.nseg = out.nseg = nseg;
  int B = 0;
  for (int k = 0; k < nseg; ++k) {
    in.start[k] = out.start[k] = B;
    in.ptr[k] = const_cast<double*>(pfield[k]);
    in.flag[k] = (flags && flags[k]) ? 3 : 0;  // cos θ, divided
    out.ptr[k] = snm[k];
    B += nlev[k];
  }
  in.start[nseg] = out.start[nseg] = B;
  double2* eo = reinterpret_cast<double2*>(mesh->d_four_a);
  if (int rc = fft_fold_map(mesh, in, eo)) return rc;
  if (!accumulate) {
    // zero the outputs, one fill per run of adjacent arrays
    for (int k = 0; k < nseg;) {
      size_t count = (size_t)nlev[k] * spec_d;
      int q = k + 1;
      while (q < nseg && snm[q] == snm[k] + count) count += (size_t)nlev[q++] * spec_d;
      if (int rc = jgcm_fill(mesh, snm[k], 0.0, count)) return rc;
      k = q;
    }
  }
  return legendre_inv_adjoint(mesh, eo, out, B, (double)mesh->nlon, 2.0 * mesh->nlon, accumulate);
}

int jgcm_compute_alpha_operator(jgcm_mesh_t* mesh, const double* fourier_a, const double* fourier_b,
                                double isign, double* alpha, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(fourier_a && fourier_b && alpha, "NULL field pointer");
  double2* ea = reinterpret_cast<double2*>(mesh->d_four_a);
  double2* eb = reinterpret_cast<double2*>(mesh->d_four_b);
  if (int rc = fold_full(mesh, reinterpret_cast<const double2*>(fourier_a), ea, nlev)) return rc;
  if (int rc = fold_full(mesh, reinterpret_cast<const double2*>(fourier_b), eb, nlev)) return rc;
  return legendre_alpha(mesh, ea, eb, isign, reinterpret_cast<double2*>(alpha), nlev);
}

// ------------------------------------------------------------------ plans
int jgcm_plan_spherical_to_grid_create(int nmesh, jgcm_mesh_t* const* meshes, const int* nseg,
                                       const double* const* const* snm,
                                       const double* const* const* snm_b,
                                       double* const* const* pfield, const int* const* nlev,
                                       const int* const* flags, int workgroups,
                                       jgcm_plan_t** out) {
  JGCM_REQUIRE(nseg && snm && pfield && nlev, "NULL argument");
  jgcm_plan* p = nullptr;
  if (int rc = plan_begin(0, nmesh, meshes, &p)) return rc;
  std::vector<InvDesc> desc(nmesh);
  std::vector<int4> items;
  std::vector<double> cost;
  int tlmax = 1, rc = JGCM_OK;
  for (int g = 0; g < nmesh && rc == JGCM_OK; ++g) {
    jgcm_mesh* M = meshes[g];
    int total = 0;
    rc = check_inverse_args(nseg[g], snm[g], snm_b ? snm_b[g] : nullptr, pfield[g], nlev[g],
                            flags ? flags[g] : nullptr, &total);
    if (rc == JGCM_OK) rc = ensure_scratch(M, total);
    if (rc != JGCM_OK) break;
    plan_hold(p, M);
    InvDesc& d = desc[g];
    build_inverse(nseg[g], snm[g], snm_b ? snm_b[g] : nullptr, pfield[g], nlev[g],
                  flags ? flags[g] : nullptr, d.in, p->fft[g]);
    p->eo[g] = reinterpret_cast<double2*>(M->d_four_a);
    d.eo = p->eo[g];
    d.tab = tables_for(M, 3);
    d.ops = SpecOpTables{M->d_dlam, M->d_dthm, M->d_dthp, M->d_uvc, M->d_uvm, M->d_uvp};
    d.B = total;
    d.nmax = d.ns = M->ns;
    d.nm = M->nm;
    d.half = M->half;
    d.tl = tiles_of(M->leg_tiles_inv);
    tlmax = std::max(tlmax, d.tl);
    const int trows = 16 * d.tl;
    const int fchunks = (total + LCOLS / 2 - 1) / (LCOLS / 2);
    const int lgroups = (M->half + trows - 1) / trows;
    for (int x = 0; x < 2 * M->nm; ++x) {
      const int nfirst = (x >> 1) + (x & 1);
      const int nr = nfirst <= d.nmax ? (d.nmax - nfirst) / 2 + 1 : 0;
      for (int y = 0; y < fchunks; ++y) {
        const int nf = std::min(LCOLS / 2, total - y * (LCOLS / 2));
        for (int z = 0; z < lgroups; ++z) {
          const int nlg = std::min(trows, M->half - z * trows);
          items.push_back(make_int4(g, x, y, z));
          cost.push_back(inv_item_cost(nr, nf, nlg));
        }
      }
    }
  }
  if (rc == JGCM_OK)
    rc = plan_finish(p, desc.data(), sizeof(InvDesc) * nmesh, items, cost, workgroups, tlmax);
  if (rc != JGCM_OK) {
    free_plan(p);
    return rc;
  }
  *out = p;
  return JGCM_OK;
}

int jgcm_plan_grid_to_spherical_create(int nmesh, jgcm_mesh_t* const* meshes, const int* nseg,
                                       const double* const* const* pfield,
                                       double* const* const* snm, const int* const* nlev,
                                       const int* const* flags, int workgroups,
                                       jgcm_plan_t** out) {
  JGCM_REQUIRE(nseg && snm && pfield && nlev, "NULL argument");
  jgcm_plan* p = nullptr;
  if (int rc = plan_begin(1, nmesh, meshes, &p)) return rc;
  std::vector<FwdDesc> desc(nmesh);
  std::vector<int4> items;
  std::vector<double> cost;
  int tlmax = 1, rc = JGCM_OK;
  for (int g = 0; g < nmesh && rc == JGCM_OK; ++g) {
    jgcm_mesh* M = meshes[g];
    int total = 0;
    rc = check_forward_args(nseg[g], pfield[g], snm[g], nlev[g], flags ? flags[g] : nullptr,
                            &total);
    if (rc == JGCM_OK) rc = ensure_scratch(M, total);
    if (rc != JGCM_OK) break;
    plan_hold(p, M);
    FwdDesc& d = desc[g];
    p->eo[g] = reinterpret_cast<double2*>(M->d_four_a);
    rc = build_forward(M, nseg[g], pfield[g], snm[g], nlev[g], flags ? flags[g] : nullptr,
                       p->eo[g], p->fft[g], d.jobs);
    if (rc != JGCM_OK) break;
    d.ns = M->ns;
    d.nm = M->nm;
    d.half = M->half;
    d.nmax = M->N;
    d.tl = tiles_of(M->leg_tiles_fwd);
    tlmax = std::max(tlmax, d.tl);
    const int trows = 16 * d.tl;
    const int maxrows = d.nmax / 2 + 1;
    const int rgroups = ((maxrows + 15) / 16 + d.tl - 1) / d.tl;
    for (int j = 0; j < d.jobs.njobs; ++j) {
      const FwdJob& job = d.jobs.job[j];
      const int terms = job.srcB ? 2 : 1;
      for (int z = d.jobs.zstart[j]; z < d.jobs.zstart[j + 1]; ++z) {
        const int nf = std::min(LCOLS / 2, job.nf - (z - d.jobs.zstart[j]) * (LCOLS / 2));
        for (int x = 0; x < 2 * M->nm; ++x) {
          const int nfirst = (x >> 1) + (x & 1);
          const int nr = nfirst <= d.nmax ? (d.nmax - nfirst) / 2 + 1 : 0;
          for (int y = 0; y < rgroups; ++y) {
            const int nrg = std::min(trows, nr - y * trows);
            if (nrg <= 0) continue;  // a tile without rows writes nothing
            items.push_back(make_int4(g, x, y, z));
            cost.push_back(fwd_item_cost(M->half, terms, nf, nrg));
          }
        }
      }
    }
  }
  if (rc == JGCM_OK)
    rc = plan_finish(p, desc.data(), sizeof(FwdDesc) * nmesh, items, cost, workgroups, tlmax);
  if (rc != JGCM_OK) {
    free_plan(p);
    return rc;
  }
  *out = p;
  return JGCM_OK;
}

int jgcm_plan_execute(jgcm_plan_t* plan) {
  JGCM_REQUIRE(plan != nullptr, "plan is NULL");
  for (int g = 0; g < plan->nmesh; ++g) {
    jgcm_mesh* M = plan->mesh[g];
    if (int rc = check_mesh(M)) return rc;  // an open capture records the mesh
    if (M->stream != plan->mesh[0]->stream) {
      set_error("mesh %d no longer shares mesh 0's stream", g);
      return JGCM_ERR_STATE;
    }
    if (M->precision != plan->precision) {
      set_error("mesh %d's Legendre precision changed after the plan was created", g);
      return JGCM_ERR_STATE;
    }
  }
  if (plan->direction == 1) {
    for (int g = 0; g < plan->nmesh; ++g)
      if (int rc = fft_fold_map(plan->mesh[g], plan->fft[g], plan->eo[g])) return rc;
    return plan_launch_legendre(plan);
  }
  if (int rc = plan_launch_legendre(plan)) return rc;
  for (int g = 0; g < plan->nmesh; ++g)
    if (int rc = ifft_unfold_map(plan->mesh[g], plan->eo[g], plan->fft[g])) return rc;
  return JGCM_OK;
}

int jgcm_plan_info(jgcm_plan_t* plan, int* nitems, int* nworkgroups) {
  JGCM_REQUIRE(plan != nullptr, "plan is NULL");
  if (nitems) *nitems = plan->nitems;
  if (nworkgroups) *nworkgroups = plan->nwg;
  return JGCM_OK;
}

int jgcm_plan_work_list(jgcm_plan_t* plan, int* items, int* wg_start) {
  JGCM_REQUIRE(plan != nullptr, "plan is NULL");
  if (items)
    for (int i = 0; i < plan->nitems; ++i) {
      const int4 v = plan->items[i];
      items[4 * i] = v.x;
      items[4 * i + 1] = v.y;
      items[4 * i + 2] = v.z;
      items[4 * i + 3] = v.w;
    }
  if (wg_start)
    for (int w = 0; w <= plan->nwg; ++w) wg_start[w] = plan->wg_start[w];
  return JGCM_OK;
}

int jgcm_plan_destroy(jgcm_plan_t* plan) {
  free_plan(plan);
  return JGCM_OK;
}

int jgcm_lpt_schedule(int n, const double* cost, int nwg, int* order, int* wg_start,
                      int* used) {
  JGCM_REQUIRE(n >= 0 && (n == 0 || cost) && order && wg_start, "NULL argument");
  std::vector<double> c(cost, cost + n);
  std::vector<int> o, ws;
  int u = lpt_schedule(c, nwg, o, ws);
  std::copy(o.begin(), o.end(), order);
  std::copy(ws.begin(), ws.end(), wg_start);
  if (used) *used = u;
  return JGCM_OK;
}

}  // extern "C"
