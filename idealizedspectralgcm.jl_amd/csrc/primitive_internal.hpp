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
  double* d_coef = nullptr;  // [ak(nd+1) | bk(nd+1) | dak(nd) | dbk(nd)]
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
  double* d_ref = nullptr;          // [dp_ref | lnp_half_ref | lnp_full_ref | t_ref | h] on device
  double* d_wave = nullptr;
  double* d_scal = nullptr;         // correction scalars
  double* d_colsum = nullptr;       // column sums [2][nlat][nlon]
};

namespace jgcm {

constexpr int MAXND = 64;

// device view of a vertical coordinate: ak, bk (nd+1), Δak, Δbk (nd)
struct VC {
  int nd, zero_top, scheme;
  const double* ak;
  const double* bk;
  const double* dak;
  const double* dbk;
};

inline VC vc_of(const jgcm_vert_coord* v) {
  return VC{v->nd, v->zero_top, v->advect_scheme, v->d_coef, v->d_coef + (v->nd + 1),
            v->d_coef + 2 * (v->nd + 1), v->d_coef + 2 * (v->nd + 1) + v->nd};
}

}  // namespace jgcm
