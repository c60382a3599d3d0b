// Primitive-equation (Held-Suarez) step on gfx950: vertical coordinate,
// pressure / geopotential columns, Held-Suarez forcing, Four_In_One,
// vertical advection, the semi-implicit solver and the mass / energy
// corrections.
//
// Reference files (line ranges per function below):
//   Vert_Coordinate.jl, Press_And_Geopot.jl, src/Atmos_Param/HS_Forcing.jl,
//   Semi_Implicit.jl, Spectral_Dynamics.jl.
// Grid-space work runs one thread per (latitude, longitude) column with a
// loop over the nd levels (coalesced across longitude); spectral-space
// semi-implicit work runs one thread per (m, n) coefficient column.  The
// operation order of every expression follows the reference broadcast so
// the FP64 rounding matches term by term (contraction off).
#pragma clang fp contract(off)

#include <cmath>
#include <new>
#include <vector>

#include "jgcm_internal.hpp"

namespace jgcm {
int area_mean_device(const jgcm_mesh* M, const double* grid, double* out_dev);
}

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

static inline int nblk(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

struct VC {
  int nd, zero_top, scheme;
  const double* ak;
  const double* bk;
  const double* dak;
  const double* dbk;
};

static VC vc_of(const jgcm_vert_coord* v) {
  return VC{v->nd, v->zero_top, v->advect_scheme, v->d_coef, v->d_coef + (v->nd + 1),
            v->d_coef + 2 * (v->nd + 1), v->d_coef + 2 * (v->nd + 1) + v->nd};
}

// ---------------------------------------------------------------- columns
// Pressure_Variables! (Press_And_Geopot.jl:36-67) for one column; P = plane
// size (nlat * nlon), c = column index.
__device__ inline void pressure_column(const VC& v, double ps, double* p_half, double* dp,
                                       double* lnp_half, double* p_full, double* lnp_full,
                                       size_t P, size_t c) {
  int nd = v.nd;
  for (int k = 0; k <= nd; ++k) p_half[k * P + c] = v.ak[k] + v.bk[k] * ps;
  for (int k = 0; k < nd; ++k) dp[k * P + c] = p_half[(k + 1) * P + c] - p_half[k * P + c];
  int kt = v.zero_top ? 1 : 0;
  for (int k = kt; k <= nd; ++k) lnp_half[k * P + c] = log(p_half[k * P + c]);
  for (int k = kt; k < nd; ++k)
    lnp_full[k * P + c] = lnp_half[(k + 1) * P + c] +
                          p_half[k * P + c] * (lnp_half[(k + 1) * P + c] - lnp_half[k * P + c]) /
                              dp[k * P + c] - 1.0;
  if (v.zero_top) {
    lnp_half[c] = 0.0;
    lnp_full[c] = lnp_half[P + c] - 1.0;
  }
  for (int k = 0; k < nd; ++k) p_full[k * P + c] = exp(lnp_full[k * P + c]);
}

// Compute_Geopotential! (Press_And_Geopot.jl:71-111); `scale` divides the
// result (1/g for Compute_Pressures_And_Heights!, :3-19).
__device__ inline void geopotential_column(const VC& v, double rdgas, const double* lnp_half,
                                           const double* lnp_full, const double* t, double geopots,
                                           double* g_full, double* g_half, size_t P, size_t c,
                                           double grav_div) {
  int nd = v.nd;
  g_half[nd * P + c] = geopots;
  int kt = 0;
  if (v.zero_top) {
    kt = 1;
    g_half[c] = 0.0;
  }
  for (int k = nd - 1; k >= kt; --k)
    g_half[k * P + c] = g_half[(k + 1) * P + c] +
                        rdgas * t[k * P + c] * (lnp_half[(k + 1) * P + c] - lnp_half[k * P + c]);
  for (int k = 0; k < nd; ++k)
    g_full[k * P + c] = g_half[(k + 1) * P + c] +
                        rdgas * t[k * P + c] * (lnp_half[(k + 1) * P + c] - lnp_full[k * P + c]);
  if (grav_div != 0.0) {
    for (int k = 0; k < nd; ++k) g_full[k * P + c] = g_full[k * P + c] / grav_div;
    for (int k = 0; k <= nd; ++k) g_half[k * P + c] = g_half[k * P + c] / grav_div;
  }
}

// Vert_Advection! (Vert_Coordinate.jl:171-210), adding the result to `acc`
// (the callers do grid_δQ .+= afterwards, Spectral_Dynamics.jl:285-290).
__device__ inline void vert_advection_add(const VC& v, const double* r, const double* dz,
                                          const double* w, double* acc, double* rdt_out, size_t P,
                                          size_t c) {
  int nd = v.nd;
  double flux_k = 0.0;  // flux at half level k (top: 0)
  for (int k = 0; k < nd; ++k) {
    double flux_k1;
    if (k + 1 < nd) {
      double rk = r[k * P + c], rk1 = r[(k + 1) * P + c];
      if (v.scheme == 0) {
        double dzk = dz[k * P + c], dzk1 = dz[(k + 1) * P + c];
        flux_k1 = w[(k + 1) * P + c] * (rk + (rk1 - rk) * dzk / (dzk + dzk1));
      } else {
        flux_k1 = w[(k + 1) * P + c] * (rk + rk1) / 2.0;
      }
    } else {
      flux_k1 = 0.0;
    }
    double rdt = (flux_k - flux_k1 + r[k * P + c] * (w[(k + 1) * P + c] - w[k * P + c])) / dz[k * P + c];
    if (rdt_out) rdt_out[k * P + c] = rdt;
    if (acc) acc[k * P + c] = acc[k * P + c] + rdt;
    flux_k = flux_k1;
  }
}

struct HSParams {
  double sigma_b, k_f, k_a, k_s, dT_y, dth_z;
};

// HS_Forcing! (HS_Forcing.jl:1-33) with Rayleigh_Damping! (:101-130) and
// Newtonian_Damping! (:37-97) for one column.  Returns 1 if a σ assertion
// (:85,122) fails.
__device__ inline int hs_forcing_column(int nd, double kappa, double cp_air, double dt, double sinj,
                                        const double* u, const double* v, const double* p_half,
                                        const double* p_full, const double* t, double* du,
                                        double* dv, double* t_eq, double* dtt, const HSParams& h,
                                        size_t P, size_t c) {
  const double day = 86400;
  double kf = h.k_f / day;
  double ps = p_half[nd * P + c];
  int bad = 0;
  for (int k = 0; k < nd; ++k) {
    double s = p_full[k * P + c] / ps;
    if ((s - h.sigma_b) / (1.0 - h.sigma_b) > 1.0) bad = 1;
    double kv = kf / (1.0 - h.sigma_b) * fmax(0.0, s - h.sigma_b);
    du[k * P + c] = -kv * u[k * P + c];
    dv[k * P + c] = -kv * v[k * P + c];
  }
  for (int k = 0; k < nd; ++k) {
    double uu = u[k * P + c], vv = v[k * P + c], a = du[k * P + c], b = dv[k * P + c];
    dtt[k * P + c] = -((uu + 0.5 * a * dt) * a + (vv + 0.5 * b * dt) * b) / cp_air;
  }
  const double t_zero = 315.0, t_strat = 200.0, p_ref = 1.0e5;
  double ka = h.k_a / day, ks = h.k_s / day;
  double s2 = sinj * sinj;
  double c2 = 1.0 - s2;
  double c4 = c2 * c2;
  for (int k = 0; k < nd; ++k) {
    double pn = p_full[k * P + c] / p_ref;
    double te = (t_zero - h.dT_y * s2 - h.dth_z * c2 * log(pn)) * pow(pn, kappa);
    t_eq[k * P + c] = fmax(t_strat, te);
  }
  for (int k = 0; k < nd; ++k) {
    double s = p_full[k * P + c] / ps;
    if ((s - h.sigma_b) / (1.0 - h.sigma_b) > 1.0) bad = 1;
    double kt = ka + (ks - ka) / (1.0 - h.sigma_b) * fmax(0.0, s - h.sigma_b) * c4;
    dtt[k * P + c] = dtt[k * P + c] - kt * (t[k * P + c] - t_eq[k * P + c]);
  }
  return bad;
}

// Four_In_One! (Spectral_Dynamics.jl:88-183) for one column.
__device__ inline void four_in_one_column(const VC& v, double rdgas, double cp_air,
                                          const double* div, const double* u, const double* vv,
                                          const double* dp, const double* lnp_half,
                                          const double* lnp_full, const double* p_full, double dlps,
                                          double dtps, const double* t, double* M_half,
                                          double* w_full, double* du, double* dv, double* dps,
                                          double* dt, size_t P, size_t c) {
  int nd = v.nd;
  double kappa = rdgas / cp_air;
  double dmean_tot = 0.0;
  for (int k = 0; k < nd; ++k) {
    size_t o = k * P + c;
    double Dp = dp[o];
    double lhp = lnp_half[(k + 1) * P + c], lh = lnp_half[o], lf = lnp_full[o];
    double dlnp_p = lhp - lf;
    double dlnp_m = lf - lh;
    double dlnp = lhp - lh;
    double x1 = (v.bk[k] * dlnp_m + v.bk[k + 1] * dlnp_p) / Dp;
    double dl = x1 * dlps;
    double dth = x1 * dtps;
    double tk = t[o];
    du[o] = du[o] - rdgas * tk * dl;
    dv[o] = dv[o] - rdgas * tk * dth;
    double uk = u[o], vk = vv[o];
    double dmean = div[o] * Dp + v.dbk[k] * (uk * dlps + vk * dtps);
    double x5 = -(dmean_tot * dlnp + dmean * dlnp_p) / Dp + uk * dl + vk * dth;
    dt[o] = dt[o] + kappa * tk * x5;
    w_full[o] = x5 * p_full[o];
    dmean_tot = dmean_tot + dmean;
    M_half[(k + 1) * P + c] = -dmean_tot;
  }
  dps[c] = dps[c] - dmean_tot;
  for (int k = 0; k < nd - 1; ++k)
    M_half[(k + 1) * P + c] = M_half[(k + 1) * P + c] + dmean_tot * v.bk[k + 1];
  M_half[c] = 0.0;
  M_half[nd * P + c] = 0.0;
}

// ------------------------------------------------------------------ kernels
// Step part 1 (Spectral_Dynamics_Physics!, SD:452-467; Compute_Corrections_Init,
// SD:5-28; Pressure_Variables!, SD:256): pressure variables of ps_c, HS forcing
// on the p-fields, δps = 0, and the column integrands of the corrections
// (ps_p and sum_k E_k Δp_k(ps_p)).
__global__ void hs_physics_kernel(VC v, int nlat, int nlon, const double* __restrict__ sint,
                                  double kappa, double cp_air, double dt, HSParams hp,
                                  const double* ps_c, const double* u_p, const double* v_p,
                                  const double* t_p, const double* ps_p, double* p_half, double* dp,
                                  double* lnp_half, double* p_full, double* lnp_full, double* du,
                                  double* dv, double* t_eq, double* dtt, double* dps,
                                  double* colsum, int forcing, int energy, int* bad) {
  size_t P = (size_t)nlat * nlon;
  GRID_STRIDE(c, P) {
    int j = (int)(c / nlon);
    pressure_column(v, ps_c[c], p_half, dp, lnp_half, p_full, lnp_full, P, c);
    dps[c] = 0.0;
    if (forcing) {
      if (hs_forcing_column(v.nd, kappa, cp_air, dt, sint[j], u_p, v_p, p_half, p_full, t_p, du, dv,
                            t_eq, dtt, hp, P, c))
        atomicAdd(bad, 1);
    }
    if (energy) {
      double ps = ps_p[c];
      double vi = 0.0;
      for (int k = 0; k < v.nd; ++k) {
        size_t o = k * P + c;
        double a = u_p[o] + dt * du[o], b = v_p[o] + dt * dv[o];
        double e = 0.5 * (a * a + b * b) + cp_air * (t_p[o] + dt * dtt[o]);
        vi = vi + e * (v.dak[k] + v.dbk[k] * ps);
      }
      colsum[c] = vi;
    }
  }
}

// Step part 2 (SD:260-307 minus the transforms): ∇ps = ∇lnps * ps, Four_In_One,
// geopotential, δlnps = δps / ps, vertical advection of u, v, T, absolute
// vorticity flux, energy = Φ + KE.
__global__ void hs_column_dynamics_kernel(VC v, int nlat, int nlon, double rdgas, double cp_air,
                                          double two_omega, const double* __restrict__ sint,
                                          const double* div, const double* u, const double* vv,
                                          const double* ps, const double* dp,
                                          const double* lnp_half, const double* lnp_full,
                                          const double* p_full, double* dlps, double* dtps,
                                          const double* t, const double* vor,
                                          const double* geopots, double* M_half, double* w_full,
                                          double* du, double* dv, double* dps, double* dtt,
                                          double* g_half, double* g_full, double* dlnps,
                                          double* absvor, double* energy) {
  size_t P = (size_t)nlat * nlon;
  GRID_STRIDE(c, P) {
    int j = (int)(c / nlon);
    double p = ps[c];
    double a = dlps[c] * p, b = dtps[c] * p;
    dlps[c] = a;
    dtps[c] = b;
    four_in_one_column(v, rdgas, cp_air, div, u, vv, dp, lnp_half, lnp_full, p_full, a, b, t, M_half,
                       w_full, du, dv, dps, dtt, P, c);
    geopotential_column(v, rdgas, lnp_half, lnp_full, t, geopots[c], g_full, g_half, P, c, 0.0);
    dlnps[c] = dps[c] / p;
    vert_advection_add(v, u, dp, M_half, du, nullptr, P, c);
    vert_advection_add(v, vv, dp, M_half, dv, nullptr, P, c);
    vert_advection_add(v, t, dp, M_half, dtt, nullptr, P, c);
    double f = two_omega * sint[j];
    for (int k = 0; k < v.nd; ++k) {
      size_t o = k * P + c;
      double av = vor[o] + f;
      absvor[o] = av;
      double uk = u[o], vk = vv[o];
      du[o] = du[o] + av * vk;
      dv[o] = dv[o] - av * uk;
      energy[o] = g_full[o] + 0.5 * (uk * uk + vk * vk);
    }
  }
}

// Step part 3 of the corrections (SD:336-343): ps_n = exp(lnps) (times the
// mass factor when known), and column integrands of the energy correction.
__global__ void hs_ps_energy_kernel(VC v, int nlat, int nlon, double cp_air, const double* lnps,
                                    double* ps_n, const double* scal, int factor_idx,
                                    const double* u, const double* vv, const double* t,
                                    double* colsum, int energy) {
  size_t P = (size_t)nlat * nlon;
  GRID_STRIDE(c, P) {
    double p;
    if (factor_idx >= 0) {
      p = ps_n[c] * scal[factor_idx];
    } else {
      p = exp(lnps[c]);
    }
    ps_n[c] = p;
    if (energy) {
      double vi = 0.0;
      for (int k = 0; k < v.nd; ++k) {
        size_t o = k * P + c;
        double e = 0.5 * (u[o] * u[o] + vv[o] * vv[o]) + cp_air * t[o];
        vi = vi + e * (v.dak[k] + v.dbk[k] * p);
      }
      colsum[c] = vi;
    }
  }
}

// scalar finishing steps (one thread); slots: 0 mean_ps_p, 1 mean_energy_p
// (area mean of the column integral, divided by g here), 2 mean_ps_n,
// 3 factor, 4 log(factor), 5 mean_energy_n, 6 temperature correction
__global__ void hs_scalar_kernel(double* scal, int stage, double grav, double cp_air) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (stage == 0) {
    scal[1] = scal[1] / grav;
  } else if (stage == 1) {
    double f = scal[0] / scal[2];
    scal[3] = f;
    scal[4] = log(f);
  } else if (stage == 2) {
    scal[5] = scal[5] / grav;
    scal[6] = grav * (scal[1] - scal[5]) / (cp_air * scal[0]);
  }
}

// apply the corrections (SD:41-57): T_n += corr, spe_t_n[k](0,0) += corr,
// spe_lnps_n(0,0) += log(factor)
__global__ void hs_apply_corrections_kernel(int nd, size_t P, size_t spec_per_level, double* t_n,
                                            double2* spe_t_n, double2* spe_lnps_n,
                                            const double* scal, int mass, int energy) {
  size_t n = (size_t)nd * P;
  double corr = energy ? scal[6] : 0.0;
  GRID_STRIDE(i, n) {
    if (energy) t_n[i] = t_n[i] + corr;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (mass) spe_lnps_n[0].x = spe_lnps_n[0].x + scal[4];
    if (energy)
      for (int k = 0; k < nd; ++k) spe_t_n[k * spec_per_level].x = spe_t_n[k * spec_per_level].x + corr;
  }
}

// pressure variables and heights (SD:355 and Compute_Pressures_And_Heights!, SD:485)
__global__ void hs_pressure_heights_kernel(VC v, int nlat, int nlon, double rdgas, double grav,
                                           const double* ps, const double* t, const double* geopots,
                                           double* p_half, double* dp, double* lnp_half,
                                           double* p_full, double* lnp_full, double* z_full,
                                           double* z_half) {
  size_t P = (size_t)nlat * nlon;
  GRID_STRIDE(c, P) {
    pressure_column(v, ps[c], p_half, dp, lnp_half, p_full, lnp_full, P, c);
    if (z_full) geopotential_column(v, rdgas, lnp_half, lnp_full, t, geopots[c], z_full, z_half, P, c, grav);
  }
}

// Semi-implicit correction (Semi_Implicit.jl:360-404 and :462-513) for one
// (m, n) coefficient column over the nd levels.
struct SIRef {
  const double* dp;
  const double* lnp_half;
  const double* lnp_full;
  const double* t_ref;
  const double* h;
  const double* bk;
};

__device__ inline void linear_ps_t(int nd, double kappa, const SIRef& r, const double2* ddiv,
                                   double2* dps, double2* dtv) {
  // Linear_Ps_T_δdiv! (Semi_Implicit.jl:271-341)
  double2 dmt = make_double2(0.0, 0.0);
  double2 M[MAXND + 1];
  for (int k = 0; k < nd; ++k) {
    double dlp = r.lnp_half[k + 1] - r.lnp_full[k];
    double dl = r.lnp_half[k + 1] - r.lnp_half[k];
    double2 dm = make_double2(ddiv[k].x * r.dp[k], ddiv[k].y * r.dp[k]);
    double s = -kappa * r.t_ref[k];
    dtv[k] = make_double2(s * (dmt.x * dl + dm.x * dlp) / r.dp[k], s * (dmt.y * dl + dm.y * dlp) / r.dp[k]);
    dmt = make_double2(dmt.x + dm.x, dmt.y + dm.y);
    M[k + 1] = make_double2(-dmt.x, -dmt.y);
  }
  *dps = make_double2(-dmt.x, -dmt.y);
  for (int k = 0; k < nd - 1; ++k)
    M[k + 1] = make_double2(M[k + 1].x + dmt.x * r.bk[k + 1], M[k + 1].y + dmt.y * r.bk[k + 1]);
  M[0] = make_double2(0.0, 0.0);
  M[nd] = make_double2(0.0, 0.0);
  double2 mdt_prev = make_double2(0.0, 0.0);  // Mdt_half[0]
  for (int k = 0; k < nd; ++k) {
    double2 mdt_next;
    if (k + 1 < nd) {
      double dtr = r.t_ref[k + 1] - r.t_ref[k];
      mdt_next = make_double2(M[k + 1].x * dtr, M[k + 1].y * dtr);
    } else {
      mdt_next = make_double2(0.0, 0.0);
    }
    dtv[k] = make_double2(dtv[k].x - 0.5 * (mdt_next.x + mdt_prev.x) / r.dp[k],
                          dtv[k].y - 0.5 * (mdt_next.y + mdt_prev.y) / r.dp[k]);
    mdt_prev = mdt_next;
  }
}

__device__ inline void linear_geopot(int nd, double rdgas, const SIRef& r, const double2* dtv,
                                     double2* g) {
  // Linear_Geopot_δt! (Semi_Implicit.jl:213-241)
  double2 gh_next = make_double2(0.0, 0.0);  // half level nd
  double2 gh[MAXND + 1];
  gh[nd] = gh_next;
  for (int k = nd - 1; k >= 1; --k) {
    double d = r.lnp_half[k + 1] - r.lnp_half[k];
    gh[k] = make_double2(gh[k + 1].x + rdgas * (dtv[k].x * d), gh[k + 1].y + rdgas * (dtv[k].y * d));
  }
  for (int k = 0; k < nd; ++k) {
    double d = r.lnp_half[k + 1] - r.lnp_full[k];
    g[k] = make_double2(gh[k + 1].x + rdgas * (dtv[k].x * d), gh[k + 1].y + rdgas * (dtv[k].y * d));
  }
}

__global__ __launch_bounds__(128) void hs_implicit_kernel(
    int nd, int nm, int ns, size_t spec_per_level, double kappa, double rdgas, double ps_ref,
    double xi, SIRef r, const double* __restrict__ eig, const double* __restrict__ wave,
    int num_wavenumbers, const double2* div_c, const double2* div_p, const double2* lnps_c,
    const double2* lnps_p, const double2* t_c, const double2* t_p, double2* ddiv, double2* dlnps,
    double2* dtt) {
  GRID_STRIDE(idx, spec_per_level) {
    int m = (int)(idx % nm);
    int n = (int)(idx / nm);
    double2 a[MAXND], b[MAXND], g[MAXND];
    double2 dps;
    // Adjust_δlnps_δt_δdiv!: δdiv_temp = div_p - div_c
    for (int k = 0; k < nd; ++k) {
      size_t o = k * spec_per_level + idx;
      a[k] = make_double2(div_p[o].x - div_c[o].x, div_p[o].y - div_c[o].y);
    }
    linear_ps_t(nd, kappa, r, a, &dps, b);          // b = δt_temp, dps = δps_temp
    double2 dl = dlnps[idx];
    dl = make_double2(dl.x + dps.x / ps_ref, dl.y + dps.y / ps_ref);
    for (int k = 0; k < nd; ++k) {
      size_t o = k * spec_per_level + idx;
      double2 d = dtt[o];
      d = make_double2(d.x + b[k].x, d.y + b[k].y);
      dtt[o] = d;
      // δt_temp = t_p - t_c + ξ δt
      a[k] = make_double2(t_p[o].x - t_c[o].x + xi * d.x, t_p[o].y - t_c[o].y + xi * d.y);
    }
    double2 lt = make_double2(lnps_p[idx].x - lnps_c[idx].x + xi * dl.x,
                              lnps_p[idx].y - lnps_c[idx].y + xi * dl.y);
    linear_geopot(nd, rdgas, r, a, g);
    double e = eig[m * (ns + 1) + n];
    for (int k = 0; k < nd; ++k) {
      size_t o = k * spec_per_level + idx;
      double hp = r.h[k] * ps_ref;
      double2 d = ddiv[o];
      a[k] = make_double2(d.x - e * (g[k].x + hp * lt.x), d.y - e * (g[k].y + hp * lt.y));
    }
    // wave-matrix solve for n <= num_wavenumbers (Semi_Implicit.jl:487-497)
    if (n >= m && n <= num_wavenumbers) {
      const double* W = wave + (size_t)n * nd * nd;
      for (int k = 0; k < nd; ++k) {
        double2 s = make_double2(0.0, 0.0);
        for (int q = 0; q < nd; ++q) {
          double w = W[k * nd + q];
          s = make_double2(s.x + w * a[q].x, s.y + w * a[q].y);
        }
        b[k] = s;
      }
      for (int k = 0; k < nd; ++k) a[k] = b[k];
    }
    for (int k = 0; k < nd; ++k) ddiv[k * spec_per_level + idx] = a[k];
    linear_ps_t(nd, kappa, r, a, &dps, b);
    for (int k = 0; k < nd; ++k) {
      size_t o = k * spec_per_level + idx;
      double2 d = dtt[o];
      dtt[o] = make_double2(d.x + xi * b[k].x, d.y + xi * b[k].y);
    }
    double f = xi / ps_ref;
    dlnps[idx] = make_double2(dl.x + f * dps.x, dl.y + f * dps.y);
  }
}

__global__ void vert_advection_kernel(VC v, size_t P, const double* r, const double* dz,
                                      const double* w, double* rdt) {
  GRID_STRIDE(c, P) vert_advection_add(v, r, dz, w, nullptr, rdt, P, c);
}

__global__ void four_in_one_kernel(VC v, size_t P, double rdgas, double cp_air, const double* div,
                                   const double* u, const double* vv, const double* dp,
                                   const double* lnp_half, const double* lnp_full,
                                   const double* p_full, const double* dlps, const double* dtps,
                                   const double* t, double* M_half, double* w_full, double* du,
                                   double* dv, double* dps, double* dt) {
  GRID_STRIDE(c, P)
  four_in_one_column(v, rdgas, cp_air, div, u, vv, dp, lnp_half, lnp_full, p_full, dlps[c], dtps[c], t,
                     M_half, w_full, du, dv, dps, dt, P, c);
}

__global__ void geopotential_kernel(VC v, size_t P, double rdgas, const double* lnp_half,
                                    const double* lnp_full, const double* t, const double* geopots,
                                    double* g_full, double* g_half) {
  GRID_STRIDE(c, P)
  geopotential_column(v, rdgas, lnp_half, lnp_full, t, geopots[c], g_full, g_half, P, c, 0.0);
}

// area-weighted mean of two planes at once into scal[s0], scal[s1]
__global__ __launch_bounds__(256) void area_mean2_kernel(const double* __restrict__ f0,
                                                         const double* __restrict__ f1,
                                                         const double* __restrict__ w, int nlat,
                                                         int nlon, double* scal, int s0, int s1) {
  __shared__ double red[2][256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nlon; i += 256) {
    double ca = 0.0, cb = 0.0;
    for (int j = 0; j < nlat; ++j) {
      if (f0) ca += f0[(size_t)j * nlon + i] * w[j];
      if (f1) cb += f1[(size_t)j * nlon + i] * w[j];
    }
    a += ca;
    b += cb;
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (f0) scal[s0] = red[0][0] / (2.0 * nlon);
    if (f1) scal[s1] = red[1][0] / (2.0 * nlon);
  }
}

// ----------------------------------------------------------- host helpers
// in-place Gauss-Jordan inverse with partial pivoting (FP64)
static bool invert(std::vector<double>& A, int n) {
  std::vector<double> I(n * n, 0.0);
  for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(A[r * n + col]) > std::fabs(A[piv * n + col])) piv = r;
    if (A[piv * n + col] == 0.0) return false;
    if (piv != col)
      for (int q = 0; q < n; ++q) {
        std::swap(A[piv * n + q], A[col * n + q]);
        std::swap(I[piv * n + q], I[col * n + q]);
      }
    double d = A[col * n + col];
    for (int q = 0; q < n; ++q) {
      A[col * n + q] /= d;
      I[col * n + q] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      double f = A[r * n + col];
      if (f == 0.0) continue;
      for (int q = 0; q < n; ++q) {
        A[r * n + q] -= f * A[col * n + q];
        I[r * n + q] -= f * I[col * n + q];
      }
    }
  }
  A = I;
  return true;
}

static double xi_of(const jgcm_leapfrog* lf) {
  return lf->init_step ? lf->dt * lf->implicit_coef : 2 * lf->dt * lf->implicit_coef;
}

// Build_Wave_Matrices (Semi_Implicit.jl:165-180) and upload
static int build_wave(jgcm_semi_implicit* si) {
  int nd = si->nd;
  const jgcm_mesh* M = si->vc->mesh;
  double xi = xi_of(si->lf);
  si->wave_matrix.assign((size_t)(si->num_wavenumbers + 1) * nd * nd, 0.0);
  for (int i = 0; i <= si->num_wavenumbers; ++i) {
    double factor = xi * xi * M->laplacian_eig[(size_t)0 * (M->ns + 1) + i];
    std::vector<double> A(nd * nd, 0.0);
    for (int k = 0; k < nd; ++k) A[k * nd + k] = 1.0;
    for (int q = 0; q < nd * nd; ++q) A[q] = A[q] - factor * si->div_mat[q];
    if (!invert(A, nd)) {
      set_error("wave matrix %d is singular", i);
      return JGCM_ERR_STATE;
    }
    std::copy(A.begin(), A.end(), si->wave_matrix.begin() + (size_t)i * nd * nd);
  }
  if (!si->d_wave) JGCM_CHECK_HIP(hipMalloc(&si->d_wave, si->wave_matrix.size() * sizeof(double)));
  JGCM_CHECK_HIP(hipMemcpyAsync(si->d_wave, si->wave_matrix.data(),
                                si->wave_matrix.size() * sizeof(double), hipMemcpyHostToDevice,
                                M->stream));
  JGCM_CHECK_HIP(hipStreamSynchronize(M->stream));
  return JGCM_OK;
}

// host restatement of the 1-column reference profile functions used by the
// semi-implicit set-up (Semi_Implicit.jl:40-160)
static void host_pressure(const jgcm_vert_coord* v, double ps, std::vector<double>& ph,
                          std::vector<double>& dp, std::vector<double>& lh, std::vector<double>& pf,
                          std::vector<double>& lf) {
  int nd = v->nd;
  ph.assign(nd + 1, 0.0);
  dp.assign(nd, 0.0);
  lh.assign(nd + 1, 0.0);
  pf.assign(nd, 0.0);
  lf.assign(nd, 0.0);
  for (int k = 0; k <= nd; ++k) ph[k] = v->ak[k] + v->bk[k] * ps;
  for (int k = 0; k < nd; ++k) dp[k] = ph[k + 1] - ph[k];
  int kt = v->zero_top ? 1 : 0;
  for (int k = kt; k <= nd; ++k) lh[k] = std::log(ph[k]);
  for (int k = kt; k < nd; ++k) lf[k] = lh[k + 1] + ph[k] * (lh[k + 1] - lh[k]) / dp[k] - 1.0;
  if (v->zero_top) {
    lh[0] = 0.0;
    lf[0] = lh[1] - 1.0;
  }
  for (int k = 0; k < nd; ++k) pf[k] = std::exp(lf[k]);
}

}  // namespace jgcm

using namespace jgcm;

extern "C" {

int jgcm_vert_coord_create(jgcm_mesh_t* mesh, int nd, int vert_coord_option,
                           int vert_difference_option, int vert_advect_scheme, double p_ref,
                           int zero_top, jgcm_vert_coord_t** out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(out != nullptr, "out is NULL");
  JGCM_REQUIRE(nd >= 1 && nd <= MAXND, "nd must be in 1..%d", MAXND);
  if (vert_coord_option != JGCM_VC_EVEN_SIGMA) {
    set_error("Compute_Vert_Coord option %d is not a valid value (only even_sigma is supported)",
              vert_coord_option);
    return JGCM_ERR_STATE;
  }
  if (vert_difference_option != JGCM_VD_SIMMONS_BURRIDGE) {
    set_error("vert_difference_option %d is not a valid value", vert_difference_option);
    return JGCM_ERR_STATE;
  }
  if (vert_advect_scheme != JGCM_VA_SECOND_CENTERED_WTS && vert_advect_scheme != JGCM_VA_SECOND_CENTERED) {
    set_error("vert_advect_scheme %d is not a valid value", vert_advect_scheme);
    return JGCM_ERR_STATE;
  }
  jgcm_vert_coord* v = new (std::nothrow) jgcm_vert_coord();
  if (!v) return JGCM_ERR_NOMEM;
  v->mesh = mesh;
  v->nd = nd;
  v->advect_scheme = vert_advect_scheme == JGCM_VA_SECOND_CENTERED_WTS ? 0 : 1;
  v->zero_top = zero_top ? 1 : 0;
  v->p_ref = p_ref;
  // Compute_Even_Sigma (Vert_Coordinate.jl:112-117): a = 0, b = LinRange(0, 1, nd+1)
  v->ak.assign(nd + 1, 0.0);
  v->bk.resize(nd + 1);
  for (int k = 0; k <= nd; ++k) v->bk[k] = (double)k / nd;
  v->bk[nd] = 1.0;
  v->dak.resize(nd);
  v->dbk.resize(nd);
  for (int k = 0; k < nd; ++k) {
    v->dak[k] = v->ak[k + 1] - v->ak[k];
    v->dbk[k] = v->bk[k + 1] - v->bk[k];
  }
  std::vector<double> all;
  all.insert(all.end(), v->ak.begin(), v->ak.end());
  all.insert(all.end(), v->bk.begin(), v->bk.end());
  all.insert(all.end(), v->dak.begin(), v->dak.end());
  all.insert(all.end(), v->dbk.begin(), v->dbk.end());
  hipError_t e = hipMalloc(&v->d_coef, all.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(v->d_coef, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("vert coord upload failed: %s", hipGetErrorString(e));
    delete v;
    return JGCM_ERR_HIP;
  }
  *out = v;
  return JGCM_OK;
}

int jgcm_vert_coord_destroy(jgcm_vert_coord_t* v) {
  if (!v) return JGCM_OK;
  if (v->d_coef) (void)hipFree(v->d_coef);
  delete v;
  return JGCM_OK;
}

int jgcm_vert_coord_get(jgcm_vert_coord_t* v, double* ak, double* bk) {
  JGCM_REQUIRE(v != nullptr, "vert coord is NULL");
  for (int k = 0; k <= v->nd; ++k) {
    if (ak) ak[k] = v->ak[k];
    if (bk) bk[k] = v->bk[k];
  }
  return JGCM_OK;
}

int jgcm_pressure_variables(jgcm_vert_coord_t* v, const double* grid_ps, double* p_half,
                            double* dp, double* lnp_half, double* p_full, double* lnp_full) {
  JGCM_REQUIRE(v && grid_ps && p_half && dp && lnp_half && p_full && lnp_full, "NULL argument");
  jgcm_mesh* M = v->mesh;
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(hs_pressure_heights_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(v),
                     M->nlat, M->nlon, 0.0, 0.0, grid_ps, nullptr, nullptr, p_half, dp, lnp_half,
                     p_full, lnp_full, nullptr, nullptr);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_compute_pressures_and_heights(jgcm_vert_coord_t* v, const jgcm_atmo_params* atmo,
                                       const double* grid_ps, const double* grid_geopots,
                                       const double* grid_t, double* p_half, double* dp,
                                       double* lnp_half, double* p_full, double* lnp_full,
                                       double* z_full, double* z_half) {
  JGCM_REQUIRE(v && atmo && grid_ps && grid_geopots && grid_t && z_full && z_half, "NULL argument");
  jgcm_mesh* M = v->mesh;
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(hs_pressure_heights_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(v),
                     M->nlat, M->nlon, atmo->rdgas, atmo->grav, grid_ps, grid_t, grid_geopots,
                     p_half, dp, lnp_half, p_full, lnp_full, z_full, z_half);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_semi_implicit_create(jgcm_vert_coord_t* v, const jgcm_atmo_params* atmo,
                              jgcm_leapfrog_t* lf, double ps_ref, const double* t_ref,
                              jgcm_semi_implicit_t** out) {
  JGCM_REQUIRE(v && atmo && lf && t_ref && out, "NULL argument");
  jgcm_semi_implicit* si = new (std::nothrow) jgcm_semi_implicit();
  if (!si) return JGCM_ERR_NOMEM;
  int nd = v->nd;
  si->vc = v;
  si->lf = lf;
  si->atmo = *atmo;
  si->nd = nd;
  si->ps_ref = ps_ref;
  si->t_ref.assign(t_ref, t_ref + nd);
  std::vector<double> ph, dp, lh, pf, lf_;
  host_pressure(v, ps_ref, ph, dp, lh, pf, lf_);
  si->dp_ref = dp;
  si->p_half_ref = ph;
  si->lnp_half_ref = lh;
  si->lnp_full_ref = lf_;
  const std::vector<double>& bk = v->bk;
  double R = atmo->rdgas, kappa = atmo->kappa;
  // δlnp_half_ref (index 0 = 0/0 is never read) and δlnp_full_ref (Semi_Implicit.jl:54-66)
  std::vector<double> dlh(nd + 1, 0.0), dlf(nd, 0.0);
  for (int k = 1; k <= nd; ++k) dlh[k] = bk[k] / ph[k];
  for (int k = 0; k < nd; ++k) {
    double p = lh[k + 1] - lf_[k], mm = lf_[k] - lh[k];
    dlf[k] = (bk[k + 1] * p + bk[k] * mm) / dp[k];
  }
  // Build_Implicit_Matrices (Semi_Implicit.jl:112-160) by unit vectors
  si->nu.assign(nd, 0.0);
  si->tau.assign(nd * nd, 0.0);
  si->gamma.assign(nd * nd, 0.0);
  for (int kk = 0; kk < nd; ++kk) {
    // Linear_Ps_T_δdiv! on e_kk (real arithmetic suffices: the input is real)
    double dmt = 0.0;
    std::vector<double> dt(nd), M(nd + 1, 0.0);
    for (int k = 0; k < nd; ++k) {
      double dlp = lh[k + 1] - lf_[k], dl = lh[k + 1] - lh[k];
      double dm = (k == kk ? 1.0 : 0.0) * dp[k];
      dt[k] = -kappa * si->t_ref[k] * (dmt * dl + dm * dlp) / dp[k];
      dmt = dmt + dm;
      M[k + 1] = -dmt;
    }
    double dps = -dmt;
    for (int k = 0; k < nd - 1; ++k) M[k + 1] = M[k + 1] + dmt * bk[k + 1];
    M[0] = 0.0;
    M[nd] = 0.0;
    std::vector<double> Mdt(nd + 1, 0.0);
    for (int k = 1; k < nd; ++k) Mdt[k] = M[k] * (si->t_ref[k] - si->t_ref[k - 1]);
    for (int k = 0; k < nd; ++k) dt[k] = dt[k] - 0.5 * (Mdt[k + 1] + Mdt[k]) / dp[k];
    si->nu[kk] = -dps;
    for (int k = 0; k < nd; ++k) si->tau[k * nd + kk] = -dt[k];
    // Linear_Geopot_δt! on e_kk
    std::vector<double> gh(nd + 1, 0.0), g(nd, 0.0);
    for (int k = nd - 1; k >= 1; --k) gh[k] = gh[k + 1] + R * ((k == kk ? 1.0 : 0.0) * (lh[k + 1] - lh[k]));
    for (int k = 0; k < nd; ++k) g[k] = gh[k + 1] + R * ((k == kk ? 1.0 : 0.0) * (lh[k + 1] - lf_[k]));
    for (int k = 0; k < nd; ++k) si->gamma[k * nd + kk] = g[k];
  }
  // h1 = Linear_Press_Gradient_δps (:190-210), h2 = Linear_Geopot_δps (:243-268)
  std::vector<double> h1(nd), h2(nd), gh(nd + 1, 0.0);
  for (int k = 0; k < nd; ++k) {
    double p = lh[k + 1] - lf_[k], mm = lf_[k] - lh[k];
    h1[k] = R * si->t_ref[k] * (bk[k + 1] * p + bk[k] * mm) / dp[k];
  }
  for (int k = nd - 1; k >= 1; --k) gh[k] = gh[k + 1] + R * (si->t_ref[k] * (dlh[k + 1] - dlh[k]));
  for (int k = 0; k < nd; ++k) h2[k] = gh[k + 1] + R * (si->t_ref[k] * (dlh[k + 1] - dlf[k]));
  si->h.resize(nd);
  for (int k = 0; k < nd; ++k) si->h[k] = h1[k] + h2[k];
  si->div_mat.assign(nd * nd, 0.0);
  for (int a = 0; a < nd; ++a)
    for (int b = 0; b < nd; ++b) {
      double s = 0.0;
      for (int q = 0; q < nd; ++q) s += si->gamma[a * nd + q] * si->tau[q * nd + b];
      si->div_mat[a * nd + b] = si->h[a] * si->nu[b] + s;
    }
  si->num_wavenumbers = v->mesh->ns - 1;
  // device reference profiles
  std::vector<double> dev;
  dev.insert(dev.end(), dp.begin(), dp.end());
  dev.insert(dev.end(), lh.begin(), lh.end());
  dev.insert(dev.end(), lf_.begin(), lf_.end());
  dev.insert(dev.end(), si->t_ref.begin(), si->t_ref.end());
  dev.insert(dev.end(), si->h.begin(), si->h.end());
  hipError_t e = hipMalloc(&si->d_ref, dev.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(si->d_ref, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&si->d_scal, 16 * sizeof(double));
  if (e == hipSuccess) e = hipMemset(si->d_scal, 0, 16 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&si->d_colsum, 2 * v->mesh->grid_elems() * sizeof(double));
  if (e != hipSuccess) {
    set_error("semi-implicit upload failed: %s", hipGetErrorString(e));
    return JGCM_ERR_HIP;
  }
  if (int rc = build_wave(si)) return rc;
  *out = si;
  return JGCM_OK;
}

int jgcm_semi_implicit_destroy(jgcm_semi_implicit_t* si) {
  if (!si) return JGCM_OK;
  void* p[] = {si->d_ref, si->d_wave, si->d_scal, si->d_colsum};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete si;
  return JGCM_OK;
}

// Update_Init_Step!(semi_implicit)  Semi_Implicit.jl:182-188
int jgcm_semi_implicit_update_init_step(jgcm_semi_implicit_t* si) {
  JGCM_REQUIRE(si != nullptr, "semi-implicit solver is NULL");
  if (int rc = jgcm_leapfrog_update_init_step(si->lf)) return rc;
  return build_wave(si);
}

int jgcm_semi_implicit_get(jgcm_semi_implicit_t* si, int which, double* host_out, size_t count) {
  JGCM_REQUIRE(si && host_out, "NULL argument");
  const std::vector<double>* t = nullptr;
  switch (which) {
    case 0: t = &si->h; break;
    case 1: t = &si->div_mat; break;
    case 2: t = &si->wave_matrix; break;
    case 3: t = &si->nu; break;
    case 4: t = &si->tau; break;
    case 5: t = &si->gamma; break;
    case 6: t = &si->dp_ref; break;
    case 7: t = &si->lnp_half_ref; break;
    case 8: t = &si->lnp_full_ref; break;
    default: break;
  }
  JGCM_REQUIRE(t != nullptr, "unknown semi-implicit table %d", which);
  JGCM_REQUIRE(count >= t->size(), "table %d needs %zu doubles", which, t->size());
  std::copy(t->begin(), t->end(), host_out);
  return JGCM_OK;
}

static SIRef siref(const jgcm_semi_implicit* si) {
  int nd = si->nd;
  const double* b = si->d_ref;
  return SIRef{b, b + nd, b + nd + (nd + 1), b + nd + (nd + 1) + nd, b + nd + (nd + 1) + 2 * nd,
               si->vc->d_coef + (nd + 1)};
}

// Implicit_Correction!(semi_implicit, vert_coord, atmo_data, ...)  Semi_Implicit.jl:462-513
int jgcm_implicit_correction(jgcm_semi_implicit_t* si, const double* div_c, const double* div_p,
                             const double* lnps_c, const double* lnps_p, const double* t_c,
                             const double* t_p, double* ddiv, double* dlnps, double* dt) {
  JGCM_REQUIRE(si && div_c && div_p && lnps_c && lnps_p && t_c && t_p && ddiv && dlnps && dt,
               "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  size_t S = M->spec_elems();
  double xi = xi_of(si->lf);
  hipLaunchKernelGGL(hs_implicit_kernel, dim3((unsigned)((S + 127) / 128)), dim3(128), 0, M->stream,
                     si->nd, M->nm, M->ns, S, si->atmo.kappa, si->atmo.rdgas, si->ps_ref, xi,
                     siref(si), M->d_eig, si->d_wave, si->num_wavenumbers, (const double2*)div_c,
                     (const double2*)div_p, (const double2*)lnps_c, (const double2*)lnps_p,
                     (const double2*)t_c, (const double2*)t_p, (double2*)ddiv, (double2*)dlnps,
                     (double2*)dt);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// HS_Forcing!(atmo_data, Δt, sinθ, u, v, p_half, p_full, t, δu, δv, t_eq, δt, params)
// HS_Forcing.jl:1-33 (pressures are recomputed from grid_ps; see jgcm.h)
int jgcm_hs_physics(jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, int64_t dt,
                    const double* grid_ps_c, const double* grid_u_p, const double* grid_v_p,
                    const double* grid_t_p, const double* grid_ps_p, double* p_half, double* dp,
                    double* lnp_half, double* p_full, double* lnp_full, double* grid_du,
                    double* grid_dv, double* grid_t_eq, double* grid_dt, double* grid_dps,
                    int do_forcing) {
  JGCM_REQUIRE(si && hp, "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  HSParams h{hp->sigma_b, hp->k_f, hp->k_a, hp->k_s, hp->delta_T_y, hp->delta_theta_z};
  int energy = si->atmo.do_energy_correction;
  int* bad = reinterpret_cast<int*>(si->d_scal + 15);
  hipLaunchKernelGGL(hs_physics_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(si->vc), M->nlat,
                     M->nlon, M->d_sin, si->atmo.kappa, si->atmo.cp_air, (double)dt, h, grid_ps_c,
                     grid_u_p, grid_v_p, grid_t_p, grid_ps_p, p_half, dp, lnp_half, p_full, lnp_full,
                     grid_du, grid_dv, grid_t_eq, grid_dt, grid_dps, si->d_colsum, do_forcing,
                     energy, bad);
  JGCM_LAUNCH_CHECK();
  // Compute_Corrections_Init (SD:5-28): area means into the scalar slots
  hipLaunchKernelGGL(area_mean2_kernel, dim3(1), dim3(256), 0, M->stream,
                     si->atmo.do_mass_correction ? grid_ps_p : nullptr, energy ? si->d_colsum : nullptr,
                     M->d_wts, M->nlat, M->nlon, si->d_scal, 0, 1);
  JGCM_LAUNCH_CHECK();
  hipLaunchKernelGGL(hs_scalar_kernel, dim3(1), dim3(64), 0, M->stream, si->d_scal, 0, si->atmo.grav,
                     si->atmo.cp_air);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_hs_column_dynamics(jgcm_semi_implicit_t* si, const double* grid_div, const double* grid_u,
                            const double* grid_v, const double* grid_ps, const double* dp,
                            const double* lnp_half, const double* lnp_full, const double* p_full,
                            double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t,
                            const double* grid_vor, const double* grid_geopots, double* grid_M_half,
                            double* grid_w_full, double* grid_du, double* grid_dv,
                            double* grid_dps, double* grid_dt, double* geopot_half,
                            double* geopot_full, double* grid_dlnps, double* grid_absvor,
                            double* grid_energy) {
  JGCM_REQUIRE(si, "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(hs_column_dynamics_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(si->vc),
                     M->nlat, M->nlon, si->atmo.rdgas, si->atmo.cp_air, 2 * si->atmo.omega, M->d_sin,
                     grid_div, grid_u, grid_v, grid_ps, dp, lnp_half, lnp_full, p_full,
                     grid_dlambda_ps, grid_dtheta_ps, grid_t, grid_vor, grid_geopots, grid_M_half,
                     grid_w_full, grid_du, grid_dv, grid_dps, grid_dt, geopot_half, geopot_full,
                     grid_dlnps, grid_absvor, grid_energy);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// Compute_Corrections! (SD:30-68) after the back-transforms, with ps_n = exp(lnps) (SD:336)
int jgcm_hs_corrections(jgcm_semi_implicit_t* si, const double* grid_lnps, double* grid_ps_n,
                        const double* grid_u_n, const double* grid_v_n, double* grid_t_n,
                        double* spe_lnps_n, double* spe_t_n) {
  JGCM_REQUIRE(si, "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  int mass = si->atmo.do_mass_correction, energy = si->atmo.do_energy_correction;
  VC v = vc_of(si->vc);
  // ps_n = exp(lnps)
  hipLaunchKernelGGL(hs_ps_energy_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, v, M->nlat, M->nlon,
                     si->atmo.cp_air, grid_lnps, grid_ps_n, si->d_scal, -1, grid_u_n, grid_v_n,
                     grid_t_n, si->d_colsum, mass ? 0 : energy);
  JGCM_LAUNCH_CHECK();
  if (mass) {
    hipLaunchKernelGGL(area_mean2_kernel, dim3(1), dim3(256), 0, M->stream, grid_ps_n, nullptr, M->d_wts,
                       M->nlat, M->nlon, si->d_scal, 2, 0);
    hipLaunchKernelGGL(hs_scalar_kernel, dim3(1), dim3(64), 0, M->stream, si->d_scal, 1,
                       si->atmo.grav, si->atmo.cp_air);
    // ps_n *= factor and the energy integrand with the corrected ps
    hipLaunchKernelGGL(hs_ps_energy_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, v, M->nlat, M->nlon,
                       si->atmo.cp_air, grid_lnps, grid_ps_n, si->d_scal, 3, grid_u_n, grid_v_n,
                       grid_t_n, si->d_colsum, energy);
    JGCM_LAUNCH_CHECK();
  }
  if (energy) {
    hipLaunchKernelGGL(area_mean2_kernel, dim3(1), dim3(256), 0, M->stream, nullptr, si->d_colsum,
                       M->d_wts, M->nlat, M->nlon, si->d_scal, 0, 5);
    hipLaunchKernelGGL(hs_scalar_kernel, dim3(1), dim3(64), 0, M->stream, si->d_scal, 2,
                       si->atmo.grav, si->atmo.cp_air);
    JGCM_LAUNCH_CHECK();
  }
  if (mass || energy) {
    size_t n = (size_t)si->nd * P;
    hipLaunchKernelGGL(hs_apply_corrections_kernel, dim3(nblk(n)), dim3(256), 0, M->stream, si->nd, P,
                       M->spec_elems(), grid_t_n, (double2*)spe_t_n, (double2*)spe_lnps_n,
                       si->d_scal, mass, energy);
    JGCM_LAUNCH_CHECK();
  }
  return JGCM_OK;
}

int jgcm_hs_scalars(jgcm_semi_implicit_t* si, double* host_out, int count) {
  JGCM_REQUIRE(si && host_out && count >= 1 && count <= 16, "bad argument");
  jgcm_mesh* M = si->vc->mesh;
  JGCM_CHECK_HIP(hipMemcpyAsync(host_out, si->d_scal, count * sizeof(double), hipMemcpyDeviceToHost,
                                M->stream));
  JGCM_CHECK_HIP(hipStreamSynchronize(M->stream));
  return JGCM_OK;
}

// Vert_Advection!(vert_coord, r, dz, w, Δt, scheme, rdt)  Vert_Coordinate.jl:171-210
int jgcm_vert_advection(jgcm_vert_coord_t* v, const double* r, const double* dz, const double* w,
                        double* rdt) {
  JGCM_REQUIRE(v && r && dz && w && rdt, "NULL argument");
  jgcm_mesh* M = v->mesh;
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(vert_advection_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(v), P, r, dz,
                     w, rdt);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// Four_In_One!(vert_coord, atmo_data, div, u, v, ps, Δp, lnp_half, lnp_full, p_full,
//              dλ_ps, dθ_ps, t, M_half, w_full, δu, δv, δps, δt)  Spectral_Dynamics.jl:88-183
int jgcm_four_in_one(jgcm_vert_coord_t* v, const jgcm_atmo_params* atmo, const double* grid_div,
                     const double* grid_u, const double* grid_v, const double* grid_ps,
                     const double* dp, const double* lnp_half, const double* lnp_full,
                     const double* p_full, const double* grid_dlambda_ps,
                     const double* grid_dtheta_ps, const double* grid_t, double* grid_M_half,
                     double* grid_w_full, double* grid_du, double* grid_dv, double* grid_dps,
                     double* grid_dt) {
  JGCM_REQUIRE(v && atmo, "NULL argument");
  jgcm_mesh* M = v->mesh;
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(four_in_one_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(v), P,
                     atmo->rdgas, atmo->cp_air, grid_div, grid_u, grid_v, dp, lnp_half, lnp_full,
                     p_full, grid_dlambda_ps, grid_dtheta_ps, grid_t, grid_M_half, grid_w_full,
                     grid_du, grid_dv, grid_dps, grid_dt);
  JGCM_LAUNCH_CHECK();
  (void)grid_ps;
  return JGCM_OK;
}

// Compute_Geopotential!(vert_coord, atmo_data, lnp_half, lnp_full, t, geopots,
//                       geopot_full, geopot_half)  Press_And_Geopot.jl:71-111
int jgcm_compute_geopotential(jgcm_vert_coord_t* v, const jgcm_atmo_params* atmo,
                              const double* lnp_half, const double* lnp_full, const double* grid_t,
                              const double* grid_geopots, double* geopot_full,
                              double* geopot_half) {
  JGCM_REQUIRE(v && atmo && lnp_half && lnp_full && grid_t && grid_geopots && geopot_full &&
                   geopot_half, "NULL argument");
  jgcm_mesh* M = v->mesh;
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(geopotential_kernel, dim3(nblk(P)), dim3(256), 0, M->stream, vc_of(v), P,
                     atmo->rdgas, lnp_half, lnp_full, grid_t, grid_geopots, geopot_full, geopot_half);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

}  // extern "C"
