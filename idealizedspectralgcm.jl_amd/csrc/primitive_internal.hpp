// Vertical coordinate and semi-implicit solver state shared by the
// primitive-equation step (primitive.hip) and its adjoint (hs_adjoint.hip).
#pragma once

#include <vector>

#include "jgcm_internal.hpp"

struct jgcm_vert_coord {
  jgcm_mesh* mesh = nullptr;
  int nd = 0;
  int advect_scheme = 0;  // 0: second_centered_wts, 1: second_centered
  int zero_top = 1;
  double p_ref = 101325.0;
  std::vector<double> ak, bk, dak, dbk;
  std::vector<double> cf;  // pure sigma: ln p_full_k - ln ps (host copy of the device table)
  // [ak(nd+1) | bk(nd+1) | dak(nd) | dbk(nd) | lnb(nd+1) | cf(nd) | sf(nd)]
  double* d_coef = nullptr;
  // pure sigma (every a_k = 0, the top half level usable): the pressure
  // variables are ps times per-level constants and ln ps plus per-level
  // constants (lnb, cf, sf below), so a column needs one logarithm
  int pure_sigma = 0;
};

struct jgcm_semi_implicit {
  jgcm_vert_coord* vc = nullptr;
  jgcm_leapfrog_t* lf = nullptr;
  jgcm_atmo_params atmo{};
  int nd = 0;
  double ps_ref = 0.0;
  std::vector<double> dp_ref, p_half_ref, lnp_half_ref, lnp_full_ref, t_ref, h, div_mat;
  std::vector<double> nu, tau, gamma;
  int num_wavenumbers = 0;
  std::vector<double> wave_matrix;  // [num_wavenumbers+1][nd][nd]
  // [dp_ref | lnp_half_ref | lnp_full_ref | t_ref | h | skap | 1 / dp_ref] on
  // device; skap_k = (p_full_k / ps)^κ = exp(κ cf_k) on pure sigma (1 otherwise)
  double* d_ref = nullptr;
  double* d_wave = nullptr;
  double* d_scal = nullptr;         // correction scalars
  double* d_colsum = nullptr;       // column sums [2][nlat][nlon]
};

namespace jgcm {

constexpr int MAXND = 64;

// device view of a vertical coordinate: ak, bk (nd+1), Δak, Δbk (nd); for a
// pure sigma coordinate also lnb = ln b_k (nd+1; 0 where b_k = 0),
// cf_k = ln p_full_k - ln ps and sf_k = p_full_k / ps (nd), the
// Simmons-Burridge full-level constants (Press_And_Geopot.jl:55-60)
struct VC {
  int nd, zero_top, scheme, pure;
  const double* ak;
  const double* bk;
  const double* dak;
  const double* dbk;
  const double* lnb;
  const double* cf;
  const double* sf;
};

inline VC vc_of(const jgcm_vert_coord* v) {
  const int n1 = v->nd + 1, nd = v->nd;
  const double* c = v->d_coef;
  return VC{nd, v->zero_top, v->advect_scheme, v->pure_sigma, c, c + n1, c + 2 * n1,
            c + 2 * n1 + nd, c + 2 * n1 + 2 * nd, c + 3 * n1 + 2 * nd, c + 3 * n1 + 3 * nd};
}

// Column tiles: a workgroup owns COLS consecutive columns (64 by default, one
// per lane of a wave) and nq = blockDim.x / COLS level groups; group q works on
// levels q*KPT .. q*KPT + KPT - 1.  With COLS < 64 one wave holds 64 / COLS
// level groups of the same columns.
constexpr int TILE = 64;

struct TileIdx {
  int lane, q, nq, k0;
  int cols;   // columns per tile (lane = column within the tile, 0 .. cols-1)
  size_t c;   // column (grid point or spectral coefficient)
  bool col;   // column exists
};

template <int KPT, int COLS = TILE>
__device__ __forceinline__ TileIdx tile_idx(size_t ncol) {
  static_assert(COLS > 0 && TILE % COLS == 0, "a wave holds whole column tiles");
  TileIdx t;
  t.lane = threadIdx.x % COLS;
  t.q = threadIdx.x / COLS;
  t.nq = blockDim.x / COLS;
  t.k0 = t.q * KPT;
  t.cols = COLS;
  t.c = (size_t)blockIdx.x * COLS + t.lane;
  t.col = t.c < ncol;
  return t;
}

// Sum over the 64 lanes of a wave in a fixed order (result in lane 0).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = TILE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, TILE);
  return v;
}

// N per-thread partials summed, per column, over the level groups above
// (smaller level index) and below this one, in group order.  `red` holds
// N * blockDim.x doubles;
// every thread of the workgroup must call it (two barriers, `red` is free
// again on return).
template <int N>
__device__ inline void level_sums_n(double* red, const TileIdx& t, const double (&part)[N],
                                    double (&above)[N], double (&below)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) red[(i * t.nq + t.q) * t.cols + t.lane] = part[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < t.nq; ++w) {
      const double x = red[(i * t.nq + w) * t.cols + t.lane];
      if (w < t.q) a += x;
      if (w > t.q) b += x;
    }
    above[i] = a;
    below[i] = b;
  }
  __syncthreads();
}

}  // namespace jgcm
