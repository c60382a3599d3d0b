// FP64 CPU restatement of the JGCM transform pair (TEST INFRASTRUCTURE ONLY:
// the CPU baseline that bench.py times beside the GPU path, and a second
// check of the numpy oracle in tests/test_oracle_cpu.py; no product code
// links it).
//
// Restates Spectral_Spherical_Mesh.jl:
//   sht_forward  <- Trans_Grid_To_Spherical!  SSM:221-293: FFT / nλ of every
//                   latitude row (:258-260), even/odd sums of mirror rows,
//                   snm[m,n] = 1/2 Σ_{j south} (g_j ± g_j*) qwg[m,n,j]
//                   (:277-288), degree num_spherical zeroed (:290);
//   sht_inverse  <- Trans_Spherical_To_Grid!  SSM:143-217: even/odd sums over
//                   degrees m..num_spherical (:195-203), unfolded to both
//                   hemispheres (:204-205), g_0 halved (:211), 2 nλ Re(ifft)
//                   (:213-215).
// The reference's FFTW call is replaced by a textbook iterative radix-2
// Cooley-Tukey FFT (nλ a power of two; a direct DFT otherwise), two real rows
// per complex transform.  OpenMP threads share the row pairs of the FFT
// stage and the zonal wavenumbers of the Legendre stage.
//
// Layouts are the oracle's (C order): grid [nlev][nlat][nlon]; spectral
// complex [nlev][ns+1][N+1] as interleaved (re, im) doubles; tables
// [N+1][ns+1][nlat] (qnm for the inverse, qnm * w for the forward).
#include "fft_cpu.hpp"

namespace {
using jgcm_oracle::cd;
using jgcm_oracle::Fft;
using jgcm_oracle::set_threads;
}  // namespace

extern "C" {

// Trans_Grid_To_Spherical!  SSM:221-293
int sht_forward(int N, int ns, int nlat, int nlon, int nlev, const double* qwg,
                const double* grid, double* snm, int threads) {
  if (nlat % 2 || N < 0 || ns < N || nlev < 1 || nlon < 2 * N + 1) return -1;
  set_threads(threads);
  const int h = nlat / 2, nm = N + 1;
  const Fft fft(nlon);
  // even / odd sums of every row pair, [m][parity][level][j], real and
  // imaginary parts apart (contiguous in j for the quadrature)
  const size_t per_m = (size_t)2 * nlev * h;
  std::vector<double> xr((size_t)nm * per_m), xi((size_t)nm * per_m);
#pragma omp parallel
  {
    std::vector<cd> z(nlon), tmp;
    // FFT / nλ (:258-260) of row j and its mirror, packed as one complex transform
#pragma omp for schedule(static)
    for (int kj = 0; kj < nlev * h; ++kj) {
      const int k = kj / h, j = kj % h;
      const double* g = grid + (size_t)k * nlat * nlon;
      const double* a = g + (size_t)j * nlon;
      const double* c = g + (size_t)(nlat - 1 - j) * nlon;
      for (int i = 0; i < nlon; ++i) z[i] = cd(a[i], c[i]);
      fft.run(z.data(), -1, tmp);
      for (int m = 0; m < nm; ++m) {
        cd zm = z[m], zc = std::conj(z[(nlon - m) % nlon]);
        cd fa = 0.5 * (zm + zc) / (double)nlon;                       // row j
        cd d = zm - zc;
        cd fc = cd(0.5 * d.imag(), -0.5 * d.real()) / (double)nlon;   // mirror: d / 2i
        cd e = fa + fc, o = fa - fc;
        const size_t oe = (size_t)m * per_m + (size_t)k * h + j;
        const size_t oo = oe + (size_t)nlev * h;
        xr[oe] = e.real();
        xi[oe] = e.imag();
        xr[oo] = o.real();
        xi[oo] = o.imag();
      }
    }
    // quadrature over the southern latitudes, rows n = m..ns-1 (:277-288)
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m < nm; ++m) {
      for (int n = m; n < ns; ++n) {
        const int par = (n - m) % 2;
        const double* q = qwg + ((size_t)m * (ns + 1) + n) * nlat;
        for (int k = 0; k < nlev; ++k) {
          const size_t base = (size_t)m * per_m + ((size_t)par * nlev + k) * h;
          const double* ar_ = xr.data() + base;
          const double* ai_ = xi.data() + base;
          double ar = 0.0, ai = 0.0;
#pragma omp simd reduction(+ : ar, ai)
          for (int j = 0; j < h; ++j) {
            ar += ar_[j] * q[j];
            ai += ai_[j] * q[j];
          }
          double* s = snm + ((size_t)k * (ns + 1) * nm + (size_t)n * nm + m) * 2;
          s[0] = ar / 2.0;
          s[1] = ai / 2.0;
        }
      }
      for (int k = 0; k < nlev; ++k) {           // :290
        double* s = snm + ((size_t)k * (ns + 1) * nm + (size_t)ns * nm + m) * 2;
        s[0] = 0.0;
        s[1] = 0.0;
      }
    }
  }
  return 0;
}

// Trans_Spherical_To_Grid!  SSM:143-217
int sht_inverse(int N, int ns, int nlat, int nlon, int nlev, const double* qnm,
                const double* snm, double* grid, int threads) {
  if (nlat % 2 || N < 0 || ns < N || nlev < 1 || nlon < 2 * N + 1) return -1;
  set_threads(threads);
  const int h = nlat / 2, nm = N + 1;
  const Fft fft(nlon);
  // Fourier coefficients of both hemispheres, [level][j][m]
  std::vector<cd> gs((size_t)nlev * h * nm), gn((size_t)nlev * h * nm);
#pragma omp parallel
  {
    std::vector<cd> z(nlon), tmp;
    // even / odd accumulators [parity][level][j], real and imaginary parts
    std::vector<double> fr((size_t)2 * nlev * h), fi((size_t)2 * nlev * h);
#pragma omp for schedule(dynamic, 1)
    for (int m = 0; m < nm; ++m) {
      std::fill(fr.begin(), fr.end(), 0.0);
      std::fill(fi.begin(), fi.end(), 0.0);
      for (int n = m; n <= ns; ++n) {          // degrees m..num_spherical (:195-203)
        const double* q = qnm + ((size_t)m * (ns + 1) + n) * nlat;
        const size_t pb = (size_t)((n - m) % 2) * nlev * h;
        for (int k = 0; k < nlev; ++k) {
          const double* s = snm + ((size_t)k * (ns + 1) * nm + (size_t)n * nm + m) * 2;
          const double cr = s[0], ci = s[1];
          double* ar = fr.data() + pb + (size_t)k * h;
          double* ai = fi.data() + pb + (size_t)k * h;
          for (int j = 0; j < h; ++j) {
            ar[j] += cr * q[j];
            ai[j] += ci * q[j];
          }
        }
      }
      const size_t ob = (size_t)nlev * h;
      for (int k = 0; k < nlev; ++k)            // unfold (:204-205)
        for (int j = 0; j < h; ++j) {
          const size_t o = (size_t)k * h + j;
          cd fe(fr[o], fi[o]), fo(fr[ob + o], fi[ob + o]);
          gs[o * nm + m] = fe + fo;
          gn[o * nm + m] = fe - fo;
        }
    }
    // 2 nλ Re(ifft(g)) with g_0 halved (:211-215), a row and its mirror per transform
#pragma omp for schedule(static)
    for (int kj = 0; kj < nlev * h; ++kj) {
      const int k = kj / h, j = kj % h;
      const cd* a_ = gs.data() + (size_t)kj * nm;
      const cd* c_ = gn.data() + (size_t)kj * nm;
      std::fill(z.begin(), z.end(), cd(0.0, 0.0));
      z[0] = cd(a_[0].real(), c_[0].real());
      for (int m = 1; m < nm; ++m) {
        cd a = a_[m], c = c_[m];
        z[m] += a + cd(0.0, 1.0) * c;
        z[nlon - m] += std::conj(a) + cd(0.0, 1.0) * std::conj(c);
      }
      fft.run(z.data(), +1, tmp);
      double* g = grid + (size_t)k * nlat * nlon;
      double* ra = g + (size_t)j * nlon;
      double* rc = g + (size_t)(nlat - 1 - j) * nlon;
      for (int i = 0; i < nlon; ++i) {
        ra[i] = z[i].real();
        rc[i] = z[i].imag();
      }
    }
  }
  return 0;
}

int sht_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
