// Textbook FFT shared by the C++ restatements (TEST INFRASTRUCTURE ONLY, like
// the rest of oracle/): an iterative radix-2 Cooley-Tukey transform for
// lengths that are powers of two, a direct DFT otherwise.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

#include <omp.h>

namespace jgcm_oracle {

using cd = std::complex<double>;

struct Fft {
  int L = 0, lg = -1;
  std::vector<cd> tw;      // exp(-2 pi i k / L)
  std::vector<int> rev;

  explicit Fft(int n) : L(n) {
    tw.resize(n);
    for (int k = 0; k < n; ++k) tw[k] = std::polar(1.0, -2.0 * M_PI * k / n);
    if ((n & (n - 1)) == 0) {
      lg = 0;
      while ((1 << lg) < n) ++lg;
      rev.resize(n);
      for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
        rev[i] = r;
      }
    }
  }

  // in place, sign -1: sum_j x_j e^{-2 pi i jk/L}; sign +1 the conjugate kernel
  void run(cd* x, int sign, std::vector<cd>& tmp) const {
    if (lg < 0) {  // direct DFT
      tmp.assign(x, x + L);
      for (int k = 0; k < L; ++k) {
        cd s = 0.0;
        for (int j = 0; j < L; ++j) {
          cd w = tw[(size_t)j * k % L];
          s += tmp[j] * (sign < 0 ? w : std::conj(w));
        }
        x[k] = s;
      }
      return;
    }
    for (int i = 0; i < L; ++i)
      if (i < rev[i]) std::swap(x[i], x[rev[i]]);
    for (int len = 2; len <= L; len <<= 1) {
      const int half = len / 2, step = L / len;
      for (int i = 0; i < L; i += len)
        for (int j = 0; j < half; ++j) {
          cd w = tw[(size_t)j * step];
          if (sign > 0) w = std::conj(w);
          cd u = x[i + j], v = x[i + j + half] * w;
          x[i + j] = u + v;
          x[i + j + half] = u - v;
        }
    }
  }
};

inline void set_threads(int threads) {
  if (threads > 0) omp_set_num_threads(threads);
}

}  // namespace jgcm_oracle
