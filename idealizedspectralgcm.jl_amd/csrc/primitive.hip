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

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "primitive_internal.hpp"

namespace jgcm {
int area_mean_device(const jgcm_mesh* M, const double* grid, double* out_dev);
}

namespace jgcm {

static inline int nblk(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// ---------------------------------------------------------------- columns
// Serial one-thread-per-column forms, used by the stand-alone entry points
// (Compute_Geopotential!, Vert_Advection!, Four_In_One!).  The fused step
// kernels further down use the level-parallel column tiles instead.

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

// HS_Forcing! parameters as the kernels use them (formed on the host once per
// call): kv0 = k_f / (1 - σ_b), kt0 = (k_s - k_a) / (1 - σ_b) and
// rsb = 1 / (1 - σ_b) with the rates per second, rcp = 1 / c_p.
struct HSParams {
  double sigma_b, kv0, ka, kt0, rsb, rcp, dT_y, dth_z;
};

static HSParams hs_params(const jgcm_hs_params* hp, double cp_air) {
  const double day = 86400;
  const double kf = hp->k_f / day, ka = hp->k_a / day, ks = hp->k_s / day;
  return HSParams{hp->sigma_b, kf / (1.0 - hp->sigma_b), ka, (ks - ka) / (1.0 - hp->sigma_b),
                  1.0 / (1.0 - hp->sigma_b), 1.0 / cp_air, hp->delta_T_y, hp->delta_theta_z};
}

// latitude row of grid column c (32-bit division: a grid has < 2^31 columns)
__device__ __forceinline__ int lat_of(size_t c, int nlon) {
  return (int)((unsigned)c / (unsigned)nlon);
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

// -------------------------------------------------------------- column tiles
// The fused step kernels give each workgroup 64 consecutive columns (one per
// lane) and nq = ceil(nd / KPT) waves; wave q keeps levels q*KPT ..
// q*KPT + KPT - 1 of its 64 columns in registers, so the levels of a column
// are worked on in parallel.  Vertical sums (mass flux divergence,
// hydrostatic integral, column energy) add per-wave partial sums through LDS.
// Sums over levels are therefore grouped by wave rather than strictly
// sequential; the FP64 results differ from a sequential sum only by rounding.
constexpr int MAXNQ = 8;
// Columns per workgroup of the grid-point column kernels (physics, column
// dynamics, pressures, correction sums): one per lane of a wave.  (Half-wave
// tiles with two levels per thread measured slower at T85 L20: 1024
// workgroups of 148-156 registers per thread against 512 of 249.)
#ifndef JGCM_GRID_COLS
#define JGCM_GRID_COLS 64
#endif
constexpr int GCOLS = JGCM_GRID_COLS;

// Sum over the GCOLS columns of a tile held by lanes 0 .. GCOLS-1 of a wave,
// in a fixed order (result in lane 0).
__device__ inline double tile_sum(double v) {
#pragma unroll
  for (int o = GCOLS / 2; o > 0; o >>= 1) v += __shfl_down(v, o, GCOLS);
  return v;
}

__device__ __forceinline__ double2 d2(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ double2 d2add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 d2sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 d2scale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
// a if keep, else 0, selected per component (see hs_implicit_kernel)
__device__ __forceinline__ double2 d2keep(bool keep, double2 a) {
  return make_double2(keep ? a.x : 0.0, keep ? a.y : 0.0);
}

// Sums of a per-thread partial over the level groups above (smaller level
// index) and below this one, per column.  `red` holds blockDim.x entries;
// every thread of the workgroup must call it (two barriers, `red` is free
// again on return).
__device__ inline void level_sums(double2* red, const TileIdx& t, double2 part, double2& above,
                                  double2& below) {
  red[t.q * t.cols + t.lane] = part;
  __syncthreads();
  double2 a = d2(0.0, 0.0), b = d2(0.0, 0.0);
  for (int i = 0; i < t.nq; ++i) {
    double2 x = red[i * t.cols + t.lane];
    if (i < t.q) a = d2add(a, x);
    if (i > t.q) b = d2add(b, x);
  }
  __syncthreads();
  above = a;
  below = b;
}


// Pressure_Variables! (Press_And_Geopot.jl:36-67) for the levels of one tile:
// ph/lnh at half levels k0 .. k0+KPT, dp/lnf/pf at full levels k0 .. k0+KPT-1.
template <int KPT>
struct PressTile {
  double ph[KPT + 1], lnh[KPT + 1], dp[KPT], lnf[KPT], pf[KPT];
};

template <int KPT>
__device__ __forceinline__ void pressure_tile(const VC& v, int k0, double ps, PressTile<KPT>& r) {
  if (v.pure) {
    // pure sigma: one logarithm per column, the rest per-level constants
    const double lps = log(ps);
#pragma unroll
    for (int i = 0; i <= KPT; ++i) {
      int k = min(k0 + i, v.nd);
      r.ph[i] = v.ak[k] + v.bk[k] * ps;
      r.lnh[i] = (v.zero_top && k == 0) ? 0.0 : v.lnb[k] + lps;
    }
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      int k = min(k0 + i, v.nd - 1);
      r.dp[i] = r.ph[i + 1] - r.ph[i];
      r.lnf[i] = v.cf[k] + lps;
      r.pf[i] = v.sf[k] * ps;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i <= KPT; ++i) {
    int k = min(k0 + i, v.nd);
    r.ph[i] = v.ak[k] + v.bk[k] * ps;
    r.lnh[i] = (v.zero_top && k == 0) ? 0.0 : log(r.ph[i]);
  }
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    r.dp[i] = r.ph[i + 1] - r.ph[i];
    if (v.zero_top && k0 + i == 0)
      r.lnf[i] = r.lnh[i + 1] - 1.0;
    else
      r.lnf[i] = r.lnh[i + 1] + r.ph[i] * (r.lnh[i + 1] - r.lnh[i]) / r.dp[i] - 1.0;
    r.pf[i] = exp(r.lnf[i]);
  }
}

template <int KPT>
__device__ __forceinline__ void store_pressure(const VC& v, const TileIdx& t, size_t P,
                                               const PressTile<KPT>& r, double* p_half, double* dp,
                                               double* lnp_half, double* p_full, double* lnp_full) {
  if (!t.col) return;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = t.k0 + i;
    if (k < v.nd) {
      size_t o = (size_t)k * P + t.c;
      p_half[o] = r.ph[i];
      lnp_half[o] = r.lnh[i];
      dp[o] = r.dp[i];
      lnp_full[o] = r.lnf[i];
      p_full[o] = r.pf[i];
    }
    if (k + 1 == v.nd) {
      size_t o = (size_t)v.nd * P + t.c;
      p_half[o] = r.ph[i + 1];
      lnp_half[o] = r.lnh[i + 1];
    }
  }
}

// Compute_Geopotential! (Press_And_Geopot.jl:71-111) for one tile, given the
// temperature and log-pressures of its levels: g_half[k] = Φs + sum_{k'>=k}
// Rd T (lnp_half[k'+1] - lnp_half[k']), g_full[k] = g_half[k+1] + Rd T
// (lnp_half[k+1] - lnp_full[k]); with zero_top, g_half[0] = 0.  Results are
// divided by `grav_div` when it is non-zero (Compute_Pressures_And_Heights!).
// The first half: the tile's hydrostatic increments Rd T (lnp_half[k+1] -
// lnp_half[k]) and their sum (the caller sums them over the tile's levels
// below, with level_sums).
template <int KPT>
__device__ __forceinline__ double geopotential_terms(const VC& v, const TileIdx& t, double rdgas,
                                                     const double (&T)[KPT],
                                                     const double (&lnh)[KPT + 1],
                                                     double (&term)[KPT]) {
  double part = 0.0;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    term[i] = (t.k0 + i < v.nd) ? rdgas * T[i] * (lnh[i + 1] - lnh[i]) : 0.0;
    part += term[i];
  }
  return part;
}

// The second half, given the sum of the increments of the levels below the tile.
template <int KPT>
__device__ __forceinline__ void geopotential_finish(const VC& v, const TileIdx& t, double rdgas,
                                                    const double (&T)[KPT],
                                                    const double (&lnh)[KPT + 1],
                                                    const double (&lnf)[KPT], double geopots,
                                                    const double (&term)[KPT], double below,
                                                    double (&gfull)[KPT],
                                                    double (&ghalf)[KPT + 1], double grav_div) {
  double acc = geopots + below;  // g_half at level k0 + KPT (or nd)
  ghalf[KPT] = acc;
#pragma unroll
  for (int i = KPT - 1; i >= 0; --i) {
    gfull[i] = acc + rdgas * T[i] * (lnh[i + 1] - lnf[i]);
    acc = acc + term[i];
    ghalf[i] = (v.zero_top && t.k0 + i == 0) ? 0.0 : acc;
  }
  if (grav_div != 0.0) {
#pragma unroll
    for (int i = 0; i < KPT; ++i) gfull[i] = gfull[i] / grav_div;
#pragma unroll
    for (int i = 0; i <= KPT; ++i) ghalf[i] = ghalf[i] / grav_div;
  }
}

template <int KPT>
__device__ __forceinline__ void geopotential_tile(const VC& v, const TileIdx& t, double2* red,
                                                  double rdgas, const double (&T)[KPT],
                                                  const double (&lnh)[KPT + 1],
                                                  const double (&lnf)[KPT], double geopots,
                                                  double (&gfull)[KPT], double (&ghalf)[KPT + 1],
                                                  double grav_div) {
  double term[KPT];
  const double part = geopotential_terms<KPT>(v, t, rdgas, T, lnh, term);
  double2 above, below;
  level_sums(red, t, d2(part, 0.0), above, below);
  geopotential_finish<KPT>(v, t, rdgas, T, lnh, lnf, geopots, term, below.x, gfull, ghalf,
                           grav_div);
}

// One workgroup's contribution to an area-weighted global mean: sum over its
// GCOLS columns of value * w[latitude], written by thread 0.
// wj: the column's Gaussian weight (read by the caller with its other inputs)
__device__ inline void block_area_partial(const TileIdx& t, double wj, double v0, double v1,
                                          double2* partial) {
  if (t.q != 0) return;
  double s0 = tile_sum(t.col ? v0 * wj : 0.0);
  double s1 = tile_sum(t.col ? v1 * wj : 0.0);
  if (t.lane == 0) partial[blockIdx.x] = d2(s0, s1);
}

// HS_Forcing! (HS_Forcing.jl:1-33) of one level of a column, shared by the
// physics pass and the fused column pass so that both round alike: Rayleigh
// damping of u, v (Rayleigh_Damping!, :101-130) with its heating (:25) and
// Newtonian cooling of T (Newtonian_Damping!, :37-97).  The column's
// constants are formed once (reciprocals of 1 - σ_b, p_half[nd] and c_p;
// on pure sigma (p_full/p_ref)^κ = skap_k (ps/p_ref)^κ, one exponential per
// column).
struct HSCol {
  double kv0, ka, kt0, sb, rsb, s2, c2, c4, rpsb, epk, rcp, dt, kappa, dT_y, dth_z;
  const double* skap;
};

__device__ __forceinline__ HSCol hs_column_constants(const VC& v, const HSParams& hp,
                                                     const double* skap, double kappa,
                                                     double dt, double ps, double sj) {
  const double ln_pref = 11.512925464970229;   // ln(1e5)
  HSCol h;
  h.kv0 = hp.kv0;
  h.ka = hp.ka;
  h.kt0 = hp.kt0;
  h.sb = hp.sigma_b;
  h.rsb = hp.rsb;
  h.s2 = sj * sj;
  h.c2 = 1.0 - h.s2;
  h.c4 = h.c2 * h.c2;
  h.rpsb = 1.0 / (v.ak[v.nd] + v.bk[v.nd] * ps);   // 1 / p_half[nd]
  h.epk = v.pure ? exp(kappa * (log(ps) - ln_pref)) : 0.0;   // (ps / p_ref)^κ
  h.rcp = hp.rcp;
  h.dt = dt;
  h.kappa = kappa;
  h.dT_y = hp.dT_y;
  h.dth_z = hp.dth_z;
  h.skap = skap;
  return h;
}

// level k (clamped into range by the caller) with full-level pressure pf and
// its logarithm lnf; p-field values U, V, T; tendencies and T_eq out
__device__ __forceinline__ void hs_forcing_level(const VC& v, const HSCol& h, int k, double pf,
                                                 double lnf, double U, double V, double T,
                                                 bool ok, double& DU, double& DV, double& DT,
                                                 double& teq, int& isbad) {
  const double t_zero = 315.0, t_strat = 200.0, p_ref = 1.0e5;
  const double ln_pref = 11.512925464970229;   // ln(1e5)
  const double s = pf * h.rpsb;
  if (ok && (s - h.sb) * h.rsb > 1.0) isbad = 1;
  const double kv = h.kv0 * fmax(0.0, s - h.sb);
  const double a = -kv * U, b = -kv * V;
  DU = a;
  DV = b;
  const double d = -((U + 0.5 * a * h.dt) * a + (V + 0.5 * b * h.dt) * b) * h.rcp;
  // ln(p/p_ref) and (p/p_ref)^κ; on pure sigma ln p_full is ln ps plus a
  // level constant and the power a level constant times a column factor
  double lpn, pk;
  if (v.pure) {
    lpn = lnf - ln_pref;
    pk = h.skap[k] * h.epk;
  } else {
    const double pn = pf / p_ref;
    lpn = log(pn);
    pk = pow(pn, h.kappa);
  }
  const double te = (t_zero - h.dT_y * h.s2 - h.dth_z * h.c2 * lpn) * pk;
  teq = fmax(t_strat, te);
  const double kt = h.ka + h.kt0 * fmax(0.0, s - h.sb) * h.c4;
  DT = d - kt * (T - teq);
}

// ------------------------------------------------------------------ kernels
// Step part 1 (Spectral_Dynamics_Physics!, SD:452-467; Compute_Corrections_Init,
// SD:5-28; Pressure_Variables!, SD:256): pressure variables of ps_c, HS_Forcing!
// (HS_Forcing.jl:1-33 with Rayleigh_Damping! :101-130 and Newtonian_Damping!
// :37-97) on the p-fields, δps = 0, and the per-workgroup partial sums of the
// corrections' area means (ps_p; sum_k E_k Δp_k(ps_p)).
template <int KPT>
__global__ __launch_bounds__(512) void hs_physics_kernel(
    VC v, int nlat, int nlon, const double* __restrict__ sint, const double* __restrict__ wts,
    double kappa, double cp_air, double dt, HSParams hp, const double* ps_c, const double* u_p,
    const double* v_p, const double* t_p, const double* ps_p, double* p_half, double* dp,
    double* lnp_half, double* p_full, double* lnp_full, double* du, double* dv, double* t_eq,
    double* dtt, double* dps, double2* partial, int forcing, int mass, int energy, int* bad,
    const double* __restrict__ skap) {
  __shared__ double2 red[MAXNQ * TILE];
  const size_t P = (size_t)nlat * nlon;
  const TileIdx t = tile_idx<KPT, GCOLS>(P);
  const int nd = v.nd;
  const int j = t.col ? lat_of(t.c, nlon) : 0;
  const double ps = t.col ? ps_c[t.c] : 1.0e5;
  PressTile<KPT> pr;
  pressure_tile<KPT>(v, t.k0, ps, pr);
  // the pressure variables are stored only when asked for (the step's
  // column pass derives them from ps itself)
  if (p_half) store_pressure<KPT>(v, t, P, pr, p_half, dp, lnp_half, p_full, lnp_full);
  if (t.col && t.q == 0) dps[t.c] = 0.0;
  const double psp = (t.col && (mass || energy)) ? ps_p[t.c] : 0.0;
  const double wj = (t.col && t.q == 0 && (mass || energy)) ? wts[j] : 0.0;

  double U[KPT], V[KPT], T[KPT], DU[KPT], DV[KPT], DT[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    bool ok = t.col && t.k0 + i < nd;
    size_t o = (size_t)(t.k0 + i) * P + t.c;
    U[i] = ok ? u_p[o] : 0.0;
    V[i] = ok ? v_p[o] : 0.0;
    T[i] = ok ? t_p[o] : 0.0;
    if (!forcing) {
      DU[i] = ok ? du[o] : 0.0;
      DV[i] = ok ? dv[o] : 0.0;
      DT[i] = ok ? dtt[o] : 0.0;
    }
  }
  if (forcing) {
    const HSCol h = hs_column_constants(v, hp, skap, kappa, dt, ps, sint[j]);
    int isbad = 0;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      bool ok = t.col && t.k0 + i < nd;
      double teq;
      hs_forcing_level(v, h, min(t.k0 + i, nd - 1), pr.pf[i], pr.lnf[i], U[i], V[i], T[i], ok,
                       DU[i], DV[i], DT[i], teq, isbad);
      if (ok) {
        size_t o = (size_t)(t.k0 + i) * P + t.c;
        du[o] = DU[i];
        dv[o] = DV[i];
        t_eq[o] = teq;
        dtt[o] = DT[i];
      }
    }
    if (isbad) atomicAdd(bad, 1);
  }
  if (mass || energy) {
    double part = 0.0;
    if (energy) {
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        int k = t.k0 + i;
        if (t.col && k < nd) {
          double a = U[i] + dt * DU[i], b = V[i] + dt * DV[i];
          double e = 0.5 * (a * a + b * b) + cp_air * (T[i] + dt * DT[i]);
          part = part + e * (v.dak[k] + v.dbk[k] * psp);
        }
      }
    }
    double2 above, below;
    level_sums(red, t, d2(part, 0.0), above, below);
    block_area_partial(t, wj, psp, part + below.x, partial);
  }
}

// Step part 2 (SD:260-307 minus the transforms): ∇ps = ∇lnps * ps,
// Four_In_One! (SD:88-183), Compute_Geopotential!, δlnps = δps / ps,
// Vert_Advection! (Vert_Coordinate.jl:171-210) of u, v, T added to the
// tendencies, absolute-vorticity flux (Compute_Abs_Vor!, SD:296-303) and
// energy = Φ + KE (SD:306).
// Arguments of the physics pass when it runs inside the column pass
// (hs_physics_column_dynamics_kernel): HS_Forcing! on the p-fields and the
// partial sums of Compute_Corrections_Init.
struct PhysArgs {
  const double* wts;
  double kappa, cp_air, dt;
  HSParams hp;
  const double* u_p;
  const double* v_p;
  const double* t_p;
  const double* ps_p;
  double* t_eq;
  double2* partial;
  int mass, energy;
  int* bad;
  const double* skap;
  // ∂λ ln ps, ∂θ ln ps from these arrays instead of dlps / dtps (which then
  // only receive ∇ps); NULL: read from dlps / dtps
  const double* dl_lnps;
  const double* dt_lnps;
  // the p = c half of Time_Advance! for the grid u, v, T, ps: u_p, v_p, t_p,
  // ps_p receive the current values once the forcing has read them (each
  // element by the thread that read it)
  int advance_p;
};

#define JGCM_COLUMN_DYNAMICS_ARGS                                                            \
    VC v, int nlat, int nlon, double rdgas, double cp_air, double two_omega,                  \
    const double* __restrict__ sint, const double* div, const double* u, const double* vv,    \
    const double* ps, const double* dp, const double* lnp_half, const double* lnp_full,       \
    const double* p_full, double* dlps, double* dtps, const double* t, const double* vor,     \
    const double* geopots, double* M_half, double* w_full, double* du, double* dv, double* dps, \
    double* dtt, double* g_half, double* g_full, double* dlnps, double* absvor, double* energy, \
    const double* dl_t, const double* dt_t
#define JGCM_COLUMN_DYNAMICS_PASS                                                            \
    v, nlat, nlon, rdgas, cp_air, two_omega, sint, div, u, vv, ps, dp, lnp_half, lnp_full,     \
    p_full, dlps, dtps, t, vor, geopots, M_half, w_full, du, dv, dps, dtt, g_half, g_full,     \
    dlnps, absvor, energy, dl_t, dt_t

// OWN_P: the pressure variables are evaluated from ps here instead of read.
// PHYS: the physics pass (HS forcing of the p-fields, δps = 0, the
// correction partials) runs first in the same tile, its tendencies kept in
// registers instead of going through du, dv, dtt and dps (requires OWN_P).
template <int KPT, bool OWN_P, bool PHYS>
__device__ __forceinline__ void hs_column_dynamics_body(double* red, const PhysArgs& pa,
                                                        JGCM_COLUMN_DYNAMICS_ARGS) {
  static_assert(OWN_P || !PHYS, "the fused physics pass evaluates the pressures itself");
  const size_t P = (size_t)nlat * nlon;
  const TileIdx tx = tile_idx<KPT, GCOLS>(P);
  const int nd = v.nd;
  const int k0 = tx.k0;
  const size_t c = tx.c;
  const double kappa = rdgas / cp_air;
  const double p = tx.col ? ps[c] : 1.0;
  const double* dl_in = dlps;
  const double* dt_in = dtps;
  if constexpr (PHYS) {
    if (pa.dl_lnps) {
      dl_in = pa.dl_lnps;
      dt_in = pa.dt_lnps;
    }
  }
  const double a = tx.col ? dl_in[c] * p : 0.0;
  const double b = tx.col ? dt_in[c] * p : 0.0;
  const double ps_tend = (!PHYS && tx.col) ? dps[c] : 0.0;   // the physics pass sets δps = 0
  const double gs = tx.col ? geopots[c] : 0.0;
  const int jlat = tx.col ? lat_of(c, nlon) : 0;
  const double f = two_omega * sint[jlat];
  // Pressure_Variables! of ps (SD:256): read from the given fields, or, when
  // they are NULL, evaluated here with the arithmetic of the physics pass
  // (bit-identical to the stored values)
  constexpr bool own_p = OWN_P;
  PressTile<KPT> pr;
  if (own_p) pressure_tile<KPT>(v, k0, p, pr);
  // levels k0-1 .. k0+KPT of u, v, T, Δp (index i+1 <-> level k0+i)
  double U[KPT + 2], V[KPT + 2], T[KPT + 2], D[KPT + 2];
#pragma unroll
  for (int i = 0; i < KPT + 2; ++i) {
    int k = k0 + i - 1;
    bool ok = tx.col && k >= 0 && k < nd;
    size_t o = (size_t)k * P + c;
    U[i] = ok ? u[o] : 0.0;
    V[i] = ok ? vv[o] : 0.0;
    T[i] = ok ? t[o] : 0.0;
    if (own_p) {
      int kk = ok ? k : 0;
      D[i] = ok ? (v.ak[kk + 1] + v.bk[kk + 1] * p) - (v.ak[kk] + v.bk[kk] * p) : 1.0;
    } else {
      D[i] = ok ? dp[o] : 1.0;
    }
  }
  double LH[KPT + 1], LF[KPT], PF[KPT], DIV[KPT], DU[KPT], DV[KPT], DT[KPT], VOR[KPT];
  // the T gradients of the horizontal advection are used last; they are
  // requested here with the other inputs so that no load waits at the end
  double GLT[KPT], GTT[KPT];
#pragma unroll
  for (int i = 0; i <= KPT; ++i) {
    int k = min(k0 + i, nd);
    LH[i] = tx.col ? (own_p ? pr.lnh[i] : lnp_half[(size_t)k * P + c]) : 0.0;
  }
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    bool ok = tx.col && k0 + i < nd;
    size_t o = (size_t)(k0 + i) * P + c;
    GLT[i] = (ok && dl_t) ? dl_t[o] : 0.0;
    GTT[i] = (ok && dl_t) ? dt_t[o] : 0.0;
    LF[i] = ok ? (own_p ? pr.lnf[i] : lnp_full[o]) : 0.0;
    PF[i] = ok ? (own_p ? pr.pf[i] : p_full[o]) : 0.0;
    DIV[i] = ok ? div[o] : 0.0;
    if constexpr (!PHYS) {
      DU[i] = ok ? du[o] : 0.0;
      DV[i] = ok ? dv[o] : 0.0;
      DT[i] = ok ? dtt[o] : 0.0;
    }
    VOR[i] = ok ? vor[o] : 0.0;
  }
  // the column energy of the physics pass (Compute_Corrections_Init), summed
  // over the levels with the other two level sums below
  double epart = 0.0;
  double psp = 0.0, wj = 0.0;
  if constexpr (PHYS) {
    // Spectral_Dynamics_Physics! (SD:452-467): HS_Forcing! (HS_Forcing.jl:1-130)
    // on the p-fields with this column's pressures, the same arithmetic as
    // hs_physics_kernel; then the Compute_Corrections_Init partials (SD:5-28)
    double Up[KPT], Vp[KPT], Tp[KPT];
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      bool ok = tx.col && k0 + i < nd;
      size_t o = (size_t)(k0 + i) * P + c;
      Up[i] = ok ? pa.u_p[o] : 0.0;
      Vp[i] = ok ? pa.v_p[o] : 0.0;
      Tp[i] = ok ? pa.t_p[o] : 0.0;
    }
    psp = (tx.col && (pa.mass || pa.energy)) ? pa.ps_p[c] : 0.0;
    wj = (tx.col && tx.q == 0 && (pa.mass || pa.energy)) ? pa.wts[jlat] : 0.0;
    const HSCol h = hs_column_constants(v, pa.hp, pa.skap, pa.kappa, pa.dt, p,
                                        sint[jlat]);
    const double dt = pa.dt;
    int isbad = 0;
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      bool ok = tx.col && k0 + i < nd;
      double teq;
      hs_forcing_level(v, h, min(k0 + i, nd - 1), pr.pf[i], pr.lnf[i], Up[i], Vp[i], Tp[i], ok,
                       DU[i], DV[i], DT[i], teq, isbad);
      if (ok) pa.t_eq[(size_t)(k0 + i) * P + c] = teq;
    }
    if (isbad) atomicAdd(pa.bad, 1);
    if (pa.energy) {
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        int k = k0 + i;
        if (tx.col && k < nd) {
          double ea = Up[i] + dt * DU[i], eb = Vp[i] + dt * DV[i];
          double e = 0.5 * (ea * ea + eb * eb) + pa.cp_air * (Tp[i] + dt * DT[i]);
          epart = epart + e * (v.dak[k] + v.dbk[k] * psp);
        }
      }
    }
  }

  // Four_In_One!: pressure-gradient force, mass-flux divergence (prefix sum
  // over levels), ω-α term and the σ-dot mass flux M_half.
  // 1 / Δp of the tile's levels (each level divides by its Δp five times)
  double rD[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) rD[i] = 1.0 / D[i + 1];
  double dl[KPT], dth[KPT], dmean[KPT];
  double part = 0.0;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    bool ok = tx.col && k < nd;
    int kk = ok ? k : 0;
    double Dp = D[i + 1];
    double dlnp_p = LH[i + 1] - LF[i];
    double dlnp_m = LF[i] - LH[i];
    double x1 = (v.bk[kk] * dlnp_m + v.bk[kk + 1] * dlnp_p) * rD[i];
    dl[i] = x1 * a;
    dth[i] = x1 * b;
    DU[i] = DU[i] - rdgas * T[i + 1] * dl[i];
    DV[i] = DV[i] - rdgas * T[i + 1] * dth[i];
    dmean[i] = ok ? DIV[i] * Dp + v.dbk[kk] * (U[i + 1] * a + V[i + 1] * b) : 0.0;
    part = part + dmean[i];
  }
  // the hydrostatic increments of the geopotential
  double Tk[KPT], LHk[KPT + 1], gterm[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) Tk[i] = T[i + 1];
#pragma unroll
  for (int i = 0; i <= KPT; ++i) LHk[i] = LH[i];
  const double gpart = geopotential_terms<KPT>(v, tx, rdgas, Tk, LHk, gterm);
  // the three sums over the levels (column energy, mass-flux divergence,
  // geopotential) in one exchange: two barriers instead of six
  const double parts[3] = {epart, part, gpart};
  double sa[3], sb[3];
  level_sums_n<3>(red, tx, parts, sa, sb);
  if constexpr (PHYS) {
    if (pa.mass || pa.energy) block_area_partial(tx, wj, psp, epart + sb[0], pa.partial);
  }
  const double total = sa[1] + part + sb[1];
  // M_half at half levels k0 .. k0+KPT
  double Mh[KPT + 1];
  Mh[0] = (k0 == 0 || k0 >= nd) ? 0.0 : -sa[1] + total * v.bk[k0];
  double run = sa[1];
  double W[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    double dlnp_p = LH[i + 1] - LF[i];
    double dlnp = LH[i + 1] - LH[i];
    double x5 = -(run * dlnp + dmean[i] * dlnp_p) * rD[i] + U[i + 1] * dl[i] + V[i + 1] * dth[i];
    DT[i] = DT[i] + kappa * T[i + 1] * x5;
    W[i] = x5 * PF[i];
    run = run + dmean[i];
    Mh[i + 1] = (k + 1 < nd) ? -run + total * v.bk[k + 1] : 0.0;
  }
  const double dps_new = ps_tend - total;
  if (tx.col && tx.q == 0) {
    dlps[c] = a;
    dtps[c] = b;
    dps[c] = dps_new;
    dlnps[c] = dps_new / p;
    M_half[c] = 0.0;
    if constexpr (PHYS) {
      if (pa.advance_p) const_cast<double*>(pa.ps_p)[c] = p;
    }
  }

  // geopotential
  double gfull[KPT], ghalf[KPT + 1];
  geopotential_finish<KPT>(v, tx, rdgas, Tk, LHk, LF, gs, gterm, sb[2], gfull, ghalf, 0.0);

  // vertical advection of u, v, T with w = M_half, dz = Δp: fluxes through
  // the half levels k0+i (i = 0 .. KPT), between R[i] and R[i+1], with the
  // centred weight Δp_k / (Δp_k + Δp_k+1) formed once for the three fields
  double fu[KPT + 1], fv[KPT + 1], ft[KPT + 1];
#pragma unroll
  for (int i = 0; i <= KPT; ++i) {
    const int h = k0 + i;
    if (h <= 0 || h >= nd) {
      fu[i] = fv[i] = ft[i] = 0.0;
    } else if (v.scheme == 0) {
      const double w = D[i] / (D[i] + D[i + 1]);
      fu[i] = Mh[i] * (U[i] + (U[i + 1] - U[i]) * w);
      fv[i] = Mh[i] * (V[i] + (V[i + 1] - V[i]) * w);
      ft[i] = Mh[i] * (T[i] + (T[i + 1] - T[i]) * w);
    } else {
      fu[i] = Mh[i] * (U[i] + U[i + 1]) / 2.0;
      fv[i] = Mh[i] * (V[i] + V[i + 1]) / 2.0;
      ft[i] = Mh[i] * (T[i] + T[i + 1]) / 2.0;
    }
  }
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    if (!(tx.col && k < nd)) continue;
    size_t o = (size_t)k * P + c;
    double wd = Mh[i + 1] - Mh[i];
    double ru = (fu[i] - fu[i + 1] + U[i + 1] * wd) * rD[i];   // dz = Δp
    double rv = (fv[i] - fv[i + 1] + V[i + 1] * wd) * rD[i];
    double rt = (ft[i] - ft[i + 1] + T[i + 1] * wd) * rD[i];
    double du_k = DU[i] + ru, dv_k = DV[i] + rv;
    double av = VOR[i] + f;
    double uk = U[i + 1], vk = V[i + 1];
    du[o] = du_k + av * vk;
    dv[o] = dv_k - av * uk;
    double dt_k = DT[i] + rt;
    if (dl_t) dt_k = dt_k - (uk * GLT[i] + vk * GTT[i]);  // Add_Horizontal_Advection! (SSM:716)
    dtt[o] = dt_k;
    absvor[o] = av;
    w_full[o] = W[i];
    M_half[(size_t)(k + 1) * P + c] = Mh[i + 1];
    g_full[o] = gfull[i];
    g_half[o] = ghalf[i];
    energy[o] = gfull[i] + 0.5 * (uk * uk + vk * vk);
    if (k + 1 == nd) g_half[(size_t)nd * P + c] = ghalf[i + 1];
    if constexpr (PHYS) {
      if (pa.advance_p) {
        const_cast<double*>(pa.u_p)[o] = uk;
        const_cast<double*>(pa.v_p)[o] = vk;
        const_cast<double*>(pa.t_p)[o] = T[i + 1];
      }
    }
  }
}

template <int KPT>
__global__ __launch_bounds__(512) void hs_column_dynamics_kernel(JGCM_COLUMN_DYNAMICS_ARGS) {
  __shared__ double red[3 * MAXNQ * TILE];  // three level sums
  hs_column_dynamics_body<KPT, false, false>(red, PhysArgs{}, JGCM_COLUMN_DYNAMICS_PASS);
}

template <int KPT>
__global__ __launch_bounds__(512) void hs_column_dynamics_own_p_kernel(JGCM_COLUMN_DYNAMICS_ARGS) {
  __shared__ double red[3 * MAXNQ * TILE];  // three level sums
  hs_column_dynamics_body<KPT, true, false>(red, PhysArgs{}, JGCM_COLUMN_DYNAMICS_PASS);
}

template <int KPT>
__global__ __launch_bounds__(512) void hs_physics_column_dynamics_kernel(
    PhysArgs pa, JGCM_COLUMN_DYNAMICS_ARGS) {
  __shared__ double red[3 * MAXNQ * TILE];  // three level sums
  hs_column_dynamics_body<KPT, true, true>(red, pa, JGCM_COLUMN_DYNAMICS_PASS);
}

// Compute_Corrections! sums (SD:30-68) in one pass: ps_n = exp(ln ps) (SD:336)
// and per workgroup Σ w ps_n, Σ w Σ_k E_k Δa_k and Σ w ps_n Σ_k E_k Δb_k with
// E = KE + c_p T_n, so that the mass-weighted energy of the corrected
// ps = f ps_n is (Σ_a + f Σ_b) / (2 nλ g) without a second pass.
template <int KPT>
__global__ __launch_bounds__(512) void hs_corrections_sums_kernel(
    VC v, int nlat, int nlon, double cp_air, const double* __restrict__ wts, const double* lnps,
    double* ps_n, const double* u, const double* vv, const double* t, double* partial, int sums,
    int energy) {
  __shared__ double2 red[MAXNQ * TILE];
  const size_t P = (size_t)nlat * nlon;
  const TileIdx tx = tile_idx<KPT, GCOLS>(P);
  const double p = tx.col ? exp(lnps[tx.c]) : 0.0;
  const double wj = (sums && tx.col && tx.q == 0) ? wts[lat_of(tx.c, nlon)] : 0.0;
  double pa = 0.0, pb = 0.0;
  if (energy) {
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      int k = tx.k0 + i;
      if (tx.col && k < v.nd) {
        size_t o = (size_t)k * P + tx.c;
        double e = 0.5 * (u[o] * u[o] + vv[o] * vv[o]) + cp_air * t[o];
        pa = pa + e * v.dak[k];
        pb = pb + e * v.dbk[k];
      }
    }
  }
  double2 above, below;
  level_sums(red, tx, d2(pa, pb), above, below);
  if (tx.col && tx.q == 0) ps_n[tx.c] = p;
  if (sums && tx.q == 0) {
    double s0 = tile_sum(tx.col ? p * wj : 0.0);
    double s1 = tile_sum(tx.col ? (pa + below.x) * wj : 0.0);
    double s2 = tile_sum(tx.col ? p * (pb + below.y) * wj : 0.0);
    if (tx.lane == 0) {
      partial[3 * (size_t)blockIdx.x] = s0;
      partial[3 * (size_t)blockIdx.x + 1] = s1;
      partial[3 * (size_t)blockIdx.x + 2] = s2;
    }
  }
}

constexpr int APPLY_THREADS = 256;

// The correction means from the per-workgroup partials of the physics pass
// (Σ w ps_p, Σ w Σ E_p Δp) and of the sums pass, reduced by every workgroup
// that needs them in the same fixed order (so all agree, bit for bit, and no
// separate reduction launch is needed): out = {mean ps_p, energy_p, mean
// ps_n, f, e_n, corr}.
__device__ void correction_scalars(const double* part_p, const double* part_n, int nparts,
                                   int nlon, double grav, double cp_air, int mass, int energy,
                                   double (*red)[APPLY_THREADS], double* out) {
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nparts; i += APPLY_THREADS) {
    acc[0] = acc[0] + part_p[2 * (size_t)i];
    acc[1] = acc[1] + part_p[2 * (size_t)i + 1];
    acc[2] = acc[2] + part_n[3 * (size_t)i];
    acc[3] = acc[3] + part_n[3 * (size_t)i + 1];
    acc[4] = acc[4] + part_n[3 * (size_t)i + 2];
  }
  // butterfly sums within each wavefront (every lane ends with the same
  // value), then the wavefronts' sums in wave order: one barrier
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < 5; ++q) acc[q] = acc[q] + __shfl_xor(acc[q], off, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 5; ++q) red[q][wave] = acc[q];
  __syncthreads();
  double tot[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    tot[q] = red[q][0];
    for (int w = 1; w < APPLY_THREADS / 64; ++w) tot[q] = tot[q] + red[q][w];
  }
  const double mps_p = tot[0] / (2.0 * nlon);
  const double e_p = tot[1] / (2.0 * nlon) / grav;
  const double mps_n = tot[2] / (2.0 * nlon);
  const double f = mass ? mps_p / mps_n : 1.0;
  const double e_n = (tot[3] + f * tot[4]) / (2.0 * nlon) / grav;
  out[0] = mps_p;
  out[1] = e_p;
  out[2] = mps_n;
  out[3] = f;
  out[4] = e_n;
  out[5] = energy ? grav * (e_p - e_n) / (cp_air * mps_p) : 0.0;
}

// Compute_Corrections! (SD:30-68) fused into Time_Advance! (Dyn_Data.jl:326-364):
// per field p = c, then c = n where n first receives its correction — kind 1:
// T_n += corr (grid), 2: ps_n *= f (grid), 3: the (0, 0) coefficient of every
// level += corr (spectral T, level stride `spec_level` doubles), 4: the (0, 0)
// coefficient += log f (spectral ln ps).  Only workgroups whose elements need
// the scalars reduce the partials (all of kinds 1 and 2, the ones holding a
// (0, 0) coefficient for kinds 3 and 4); workgroup 0 of the first kind-1 field
// stores the seven scalars as hs_apply_corrections_kernel does.  blockIdx.y = field.
constexpr int MAXCADV = 16;
#ifndef JGCM_CADV_PER_THREAD
#define JGCM_CADV_PER_THREAD 4  // elements per thread of hs_correct_advance_kernel
#endif
// Compute_Corrections! on the new state in place (SD:30-68): T_n += corr and
// ps_n *= f on the grid, the (0, 0) coefficient of every level of T_n += corr
// and that of ln ps_n += log f.  Workgroup b covers elements b W .. (b+1) W - 1
// of T_n followed by ps_n (W = JGCM_CADV_PER_THREAD x APPLY_THREADS); it
// requests them before reducing the partials (both round trips overlap) and
// then writes them; workgroup 0 also updates the spectral coefficients and
// stores the seven scalars.  Every workgroup reduces the partials in the same
// fixed order (correction_scalars), and the corrections are those of
// hs_correct_advance_kernel, so both give the same values bit for bit.
__global__ __launch_bounds__(APPLY_THREADS) void hs_apply_corrections_kernel(
    int nd, size_t P, size_t spec_per_level, int nlon, double* t_n, double* ps_n,
    double2* spe_t_n, double2* spe_lnps_n, const double* part_p, const double* part_n, int nparts,
    double* scal, double grav, double cp_air, int mass, int energy) {
  __shared__ double red[5][APPLY_THREADS];
  constexpr int IT = JGCM_CADV_PER_THREAD;
  const size_t nT = energy ? (size_t)nd * P : 0;
  const size_t total = nT + (mass ? P : 0);
  const size_t lo = (size_t)blockIdx.x * IT * APPLY_THREADS;
  double x[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t i = lo + threadIdx.x + (size_t)it * APPLY_THREADS;
    x[it] = i < nT ? t_n[i] : (i < total ? ps_n[i - nT] : 0.0);
  }
  double sc[6];
  correction_scalars(part_p, part_n, nparts, nlon, grav, cp_air, mass, energy, red, sc);
  const double f = sc[3], corr = sc[5];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t i = lo + threadIdx.x + (size_t)it * APPLY_THREADS;
    if (i < nT) t_n[i] = x[it] + corr;
    else if (i < total) ps_n[i - nT] = x[it] * f;
  }
  if (blockIdx.x != 0) return;
  // one level's (0, 0) coefficient per thread: a loop in one thread would wait
  // for each store before the next load (the compiler cannot rule out aliasing)
  if (energy)
    for (int k = threadIdx.x; k < nd; k += APPLY_THREADS)
      spe_t_n[k * spec_per_level].x = spe_t_n[k * spec_per_level].x + corr;
  if (threadIdx.x == 0) {
    scal[0] = sc[0];
    if (energy) scal[1] = sc[1];
    if (mass) {
      scal[2] = sc[2];
      scal[3] = f;
      scal[4] = log(f);
      spe_lnps_n[0].x = spe_lnps_n[0].x + scal[4];
    }
    if (energy) {
      scal[5] = sc[4];
      scal[6] = corr;
    }
  }
}

struct CorrAdvanceSet {
  double* p[MAXCADV];     // NULL: p already holds the filtered c; only c = n remains
  double* c[MAXCADV];
  double* n[MAXCADV];
  size_t count[MAXCADV];  // doubles
  int kind[MAXCADV];
};

__global__ __launch_bounds__(APPLY_THREADS) void hs_correct_advance_kernel(
    CorrAdvanceSet s, int per_block, size_t spec_level, int nlon, const double* part_p,
    const double* part_n, int nparts, double* scal, double grav, double cp_air, int mass,
    int energy, int scal_field) {
  __shared__ double red[5][APPLY_THREADS];
  const int fld = blockIdx.y;
  const size_t cnt = s.count[fld];
  const size_t lo = (size_t)blockIdx.x * per_block;
  if (lo >= cnt) return;  // uniform over the workgroup
  const size_t hi = min(lo + (size_t)per_block, cnt);
  const int kind = (mass || energy) ? s.kind[fld] : 0;
  bool need = kind == 1 || kind == 2;
  if (kind == 3) need = (lo + spec_level - 1) / spec_level * spec_level < hi;  // holds a multiple
  if (kind == 4) need = lo == 0;
  double* __restrict__ p = s.p[fld];
  double* __restrict__ c = s.c[fld];
  double* __restrict__ n = s.n[fld];
  // the block's c and n values are requested before the correction scalars
  // are reduced, so that both round trips overlap
  constexpr int IT = JGCM_CADV_PER_THREAD;
  double xc[IT], xn[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t i = lo + threadIdx.x + (size_t)it * APPLY_THREADS;
    xc[it] = (i < hi && p) ? c[i] : 0.0;
    xn[it] = i < hi ? n[i] : 0.0;
  }
  double sc[6] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
  if (need)
    correction_scalars(part_p, part_n, nparts, nlon, grav, cp_air, mass, energy, red, sc);
  const double f = sc[3], corr = sc[5];
  const double logf = (kind == 4 && mass) ? log(f) : 0.0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const size_t i = lo + threadIdx.x + (size_t)it * APPLY_THREADS;
    if (i >= hi) break;
    if (p) p[i] = xc[it];
    double v = xn[it];
    if (kind == 1 && energy) {
      v = v + corr;
      n[i] = v;
    } else if (kind == 2 && mass) {
      v = v * f;
      n[i] = v;
    } else if (kind == 3 && energy && i % spec_level == 0) {
      v = v + corr;
      n[i] = v;
    } else if (kind == 4 && mass && i == 0) {
      v = v + logf;
      n[i] = v;
    }
    c[i] = v;
  }
  if ((mass || energy) && fld == scal_field && blockIdx.x == 0 && threadIdx.x == 0) {
    scal[0] = sc[0];
    if (energy) scal[1] = sc[1];
    if (mass) {
      scal[2] = sc[2];
      scal[3] = f;
      scal[4] = log(f);
    }
    if (energy) {
      scal[5] = sc[4];
      scal[6] = corr;
    }
  }
}

// Pressure_Variables! (SD:355) and Compute_Pressures_And_Heights! (SD:485,
// Press_And_Geopot.jl:3-19: geopotential / g) when z_full is given.
template <int KPT>
__global__ __launch_bounds__(512) void hs_pressure_heights_kernel(
    VC v, int nlat, int nlon, double rdgas, double grav, const double* ps, const double* t,
    const double* geopots, double* p_half, double* dp, double* lnp_half, double* p_full,
    double* lnp_full, double* z_full, double* z_half) {
  __shared__ double2 red[MAXNQ * TILE];
  const size_t P = (size_t)nlat * nlon;
  const TileIdx tx = tile_idx<KPT, GCOLS>(P);
  // T is requested before the pressure stores (the stores may alias it for
  // all the compiler knows, so it would not move the loads above them)
  double T[KPT], gfull[KPT], ghalf[KPT + 1];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    bool ok = z_full && tx.col && tx.k0 + i < v.nd;
    T[i] = ok ? t[(size_t)(tx.k0 + i) * P + tx.c] : 0.0;
  }
  const double gs = (z_full && tx.col) ? geopots[tx.c] : 0.0;
  PressTile<KPT> pr;
  pressure_tile<KPT>(v, tx.k0, tx.col ? ps[tx.c] : 1.0e5, pr);
  store_pressure<KPT>(v, tx, P, pr, p_half, dp, lnp_half, p_full, lnp_full);
  if (!z_full) return;  // uniform over the grid
  geopotential_tile<KPT>(v, tx, red, rdgas, T, pr.lnh, pr.lnf, gs, gfull, ghalf, grav);
  if (!tx.col) return;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = tx.k0 + i;
    if (k < v.nd) {
      size_t o = (size_t)k * P + tx.c;
      z_full[o] = gfull[i];
      z_half[o] = ghalf[i];
      if (k + 1 == v.nd) z_half[(size_t)v.nd * P + tx.c] = ghalf[i + 1];
    }
  }
}

// Semi-implicit correction (Semi_Implicit.jl:360-404 and :462-513): one tile
// column per (m, n) coefficient; complex values per level.
struct SIRef {
  const double* dp;
  const double* lnp_half;
  const double* lnp_full;
  const double* t_ref;
  const double* h;
  const double* bk;
  const double* rdp;  // 1 / dp
};

// Linear_Ps_T_δdiv! (Semi_Implicit.jl:271-341) on one tile: δps and δT of
// the linearised system for the divergence `A` of this tile's levels.
template <int KPT>
__device__ __forceinline__ void linear_ps_t_tile(int nd, double kappa, const SIRef& r,
                                                 const TileIdx& t, double2* red,
                                                 const double2 (&A)[KPT], double2& dps,
                                                 double2 (&B)[KPT]) {
  const int k0 = t.k0;
  double2 dm[KPT];
  double2 part = d2(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    dm[i] = (k < nd) ? d2scale(r.dp[k], A[i]) : d2(0.0, 0.0);
    part = d2add(part, dm[i]);
  }
  double2 above, below;
  level_sums(red, t, part, above, below);
  const double2 total = d2add(d2add(above, part), below);
  dps = d2(-total.x, -total.y);
  double2 Mh[KPT + 1];
  Mh[0] = (k0 == 0 || k0 >= nd) ? d2(0.0, 0.0) : d2add(d2(-above.x, -above.y), d2scale(r.bk[k0], total));
  double2 run = above;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    if (k < nd) {
      double dlp = r.lnp_half[k + 1] - r.lnp_full[k];
      double dl = r.lnp_half[k + 1] - r.lnp_half[k];
      double s = -kappa * r.t_ref[k];
      B[i] = d2(s * (run.x * dl + dm[i].x * dlp) * r.rdp[k], s * (run.y * dl + dm[i].y * dlp) * r.rdp[k]);
      run = d2add(run, dm[i]);
      Mh[i + 1] = (k + 1 < nd) ? d2add(d2(-run.x, -run.y), d2scale(r.bk[k + 1], total)) : d2(0.0, 0.0);
    } else {
      B[i] = d2(0.0, 0.0);
      Mh[i + 1] = d2(0.0, 0.0);
    }
  }
  // M_half * ΔT_ref at half levels 1 .. nd-1 (zero at the top and bottom)
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    if (k >= nd) continue;
    double2 next = (k + 1 < nd) ? d2scale(r.t_ref[k + 1] - r.t_ref[k], Mh[i + 1]) : d2(0.0, 0.0);
    double2 prev = (k >= 1) ? d2scale(r.t_ref[k] - r.t_ref[k - 1], Mh[i]) : d2(0.0, 0.0);
    B[i] = d2(B[i].x - 0.5 * (next.x + prev.x) * r.rdp[k], B[i].y - 0.5 * (next.y + prev.y) * r.rdp[k]);
  }
}

// Linear_Geopot_δt! (Semi_Implicit.jl:213-241) on one tile.
template <int KPT>
__device__ __forceinline__ void linear_geopot_tile(int nd, double rdgas, const SIRef& r,
                                                   const TileIdx& t, double2* red,
                                                   const double2 (&B)[KPT], double2 (&G)[KPT]) {
  const int k0 = t.k0;
  double2 term[KPT];
  double2 part = d2(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    int k = k0 + i;
    if (k >= 1 && k < nd) {
      double d = r.lnp_half[k + 1] - r.lnp_half[k];
      term[i] = d2(rdgas * (B[i].x * d), rdgas * (B[i].y * d));
    } else {
      term[i] = d2(0.0, 0.0);
    }
    part = d2add(part, term[i]);
  }
  double2 above, below;
  level_sums(red, t, part, above, below);
  double2 acc = below;  // geopotential at half level k0 + KPT
#pragma unroll
  for (int i = KPT - 1; i >= 0; --i) {
    int k = k0 + i;
    if (k < nd) {
      double d = r.lnp_half[k + 1] - r.lnp_full[k];
      G[i] = d2(acc.x + rdgas * (B[i].x * d), acc.y + rdgas * (B[i].y * d));
    } else {
      G[i] = d2(0.0, 0.0);
    }
    acc = d2add(acc, term[i]);
  }
}

// Compute_Spectral_Damping! (Time_Integrator.jl:76-80, when `damp`) and
// Filtered_Leapfrog! (:93-100) of one element held in registers: q = δQ in,
// damped δQ out; c_out, n_out = the filtered current and the new value.  The
// arithmetic and its order are those of stepper.hip's damped_leapfrog_elem.
__device__ __forceinline__ void leapfrog_reg(double2& q, double2 p, double2 c, double d, bool damp,
                                             int init_step, double dt, double r, double2& c_out,
                                             double2& n_out) {
  if (damp) {
    if (init_step) {
      double den = 1.0 + dt * d;
      q = make_double2((q.x - c.x * d) / den, (q.y - c.y * d) / den);
    } else {
      double den = 1.0 + (2.0 * dt) * d;
      q = make_double2((q.x - p.x * d) / den, (q.y - p.y * d) / den);
    }
  }
  if (init_step) {
    double2 nx = make_double2(c.x + dt * q.x, c.y + dt * q.y);
    n_out = nx;
    c_out = make_double2(c.x + r * (-1.0 * c.x + nx.x), c.y + r * (-1.0 * c.y + nx.y));
  } else {
    double2 c1 = make_double2(c.x + r * (p.x - 2.0 * c.x), c.y + r * (p.y - 2.0 * c.y));
    double2 nx = make_double2(p.x + (2.0 * dt) * q.x, p.y + (2.0 * dt) * q.y);
    n_out = nx;
    c_out = make_double2(c1.x + r * nx.x, c1.y + r * nx.y);
  }
}

// The step's damping + leapfrog fused into the semi-implicit solve (vorticity,
// divergence, T damped; ln ps not): the c arrays are written (Robert filter),
// the n arrays receive the new values.  on = 0: not fused.
struct LFFuse {
  double2* dvor;
  const double2* vor_p;
  double2* vor_c;
  double2* vor_n;
  double2* div_c;
  double2* div_n;
  double2* t_c;
  double2* t_n;
  double2* lnps_c;
  double2* lnps_n;
  const double* damping;  // [m][n]
  int on, init_step;
  double dt, r;
  // advance: the p = c half of the spectral Time_Advance! as well — p
  // receives the filtered current value instead of c, which keeps the old
  // value until jgcm_hs_corrections_advance sets c = n (the p arrays are those
  // read as div_p, t_p, lnps_p and vor_p; each element is read and written by
  // the same thread)
  int advance;
  double2* vor_pw;
  double2* div_pw;
  double2* t_pw;
  double2* lnps_pw;
};

template <int KPT, int COLS>
__global__ __launch_bounds__(512) void hs_implicit_kernel(
    int nd, int nm, int ns, size_t spec_per_level, double kappa, double rdgas, double ps_ref,
    double xi, SIRef rg, const double* __restrict__ eig, const double* __restrict__ wave,
    int num_wavenumbers, const double2* div_c, const double2* div_p, const double2* lnps_c,
    const double2* lnps_p, const double2* t_c, const double2* t_p, double2* ddiv, double2* dlnps,
    double2* dtt, double2* energy, int stage_w, int tri, LFFuse lf) {
  // tri: only the coefficients n >= m are read and written (those with n < m
  // keep their values), and a tile holding none of them returns at once.
  // LDS: nd * COLS double2 (level sums, then the wave-matrix operand); the
  // reference profiles (6 x (nd+1) doubles); when stage_w, the wave matrices
  // W_n of the tile's degrees n_lo .. n_hi.  Everything a thread reads from
  // global memory is requested at entry, before the first barrier, so the
  // column work below waits for one round trip rather than one per stage.
  extern __shared__ double2 lds[];
  const TileIdx tx = tile_idx<KPT, COLS>(spec_per_level);
  // the tile's COLS consecutive columns (idx = n nm + m) span degrees n_lo .. n_hi
  const size_t first = (size_t)blockIdx.x * COLS;
  const size_t last = min(first + COLS, spec_per_level) - 1;
  const int n_lo = (int)(first / nm);
  if (tri && (int)(last / nm) == n_lo && (int)(first % nm) > n_lo) return;  // all m > n
  const size_t idx = tx.c;
  const int n = tx.col ? (int)((unsigned)idx / (unsigned)nm) : 0;
  const int m = tx.col ? (int)idx - n * nm : 0;
  const bool live = tx.col && (!tri || n >= m);
  const int k0 = tx.k0;
  const int n1 = nd + 1;
  double* refs = reinterpret_cast<double*>(lds + (size_t)nd * COLS);
  for (int k = threadIdx.x; k < n1; k += blockDim.x) {
    const bool in = k < nd;
    refs[k] = in ? rg.dp[k] : 0.0;
    refs[n1 + k] = rg.lnp_half[k];
    refs[2 * n1 + k] = in ? rg.lnp_full[k] : 0.0;
    refs[3 * n1 + k] = in ? rg.t_ref[k] : 0.0;
    refs[4 * n1 + k] = in ? rg.h[k] : 0.0;
    refs[5 * n1 + k] = rg.bk[k];
    refs[6 * n1 + k] = in ? rg.rdp[k] : 0.0;
  }
  const SIRef r{refs, refs + n1, refs + 2 * n1, refs + 3 * n1, refs + 4 * n1, refs + 5 * n1,
                refs + 6 * n1};
  const int nnd = nd * nd;
  double* Wl = refs + 7 * n1;
  if (stage_w) {
    const int nW = (int)(last / nm) - n_lo + 1;
    for (int q = threadIdx.x; q < nW * nnd; q += blockDim.x) {
      const int nn = n_lo + q / nnd;
      Wl[q] = nn <= num_wavenumbers ? wave[(size_t)nn * nnd + q % nnd] : 0.0;
    }
  }
  // per-thread inputs: every load reads a valid address (element 0 stands in
  // for the columns and levels this thread does not own) and the components
  // are selected afterwards.  A conditional on double2 objects
  // (ok ? x[o] : zero) compiles to a select between the objects' addresses —
  // generic (flat) loads that also wait for this thread's earlier stores, and
  // a private copy of the zero
  double2 Dp[KPT], Dc[KPT], Tp[KPT], Tc[KPT], DT[KPT], DV[KPT], EN[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const bool ok = live && k0 + i < nd;
    const size_t o = ok ? (size_t)(k0 + i) * spec_per_level + idx : 0;
    Dp[i] = d2keep(ok, div_p[o]);
    Dc[i] = d2keep(ok, div_c[o]);
    Tp[i] = d2keep(ok, t_p[o]);
    Tc[i] = d2keep(ok, t_c[o]);
    DT[i] = d2keep(ok, dtt[o]);
    DV[i] = d2keep(ok, ddiv[o]);
    EN[i] = energy ? d2keep(ok, energy[o]) : d2(0.0, 0.0);
  }
  double2 QV[KPT], VP[KPT], VCc[KPT];
  double dmp = 0.0;
  if (lf.on) {
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const bool ok = live && k0 + i < nd;
      const size_t o = ok ? (size_t)(k0 + i) * spec_per_level + idx : 0;
      QV[i] = d2keep(ok, lf.dvor[o]);
      VP[i] = d2keep(ok, lf.vor_p[o]);
      VCc[i] = d2keep(ok, lf.vor_c[o]);
    }
    dmp = lf.damping[m * (ns + 1) + n];
  }
  const size_t ic = live ? idx : 0;
  const double2 dl0 = d2keep(live, dlnps[ic]);
  const double2 lnc = d2keep(live, lnps_c[ic]);
  const double2 lnp = d2keep(live, lnps_p[ic]);
  const double e = eig[m * (ns + 1) + n];
  __syncthreads();  // refs and Wl
  double2 A[KPT], B[KPT], G[KPT], DD[KPT];
  // Adjust_δlnps_δt_δdiv!: δdiv_temp = div_p - div_c
#pragma unroll
  for (int i = 0; i < KPT; ++i) A[i] = d2sub(Dp[i], Dc[i]);
  double2 dps;
  linear_ps_t_tile<KPT>(nd, kappa, r, tx, lds, A, dps, B);  // B = δt_temp, dps = δps_temp
  const double2 dl = d2(dl0.x + dps.x / ps_ref, dl0.y + dps.y / ps_ref);
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const bool ok = live && k0 + i < nd;
    const double2 d = d2add(DT[i], B[i]);
    DD[i] = d;
    // δt_temp = t_p - t_c + ξ δt
    A[i] = ok ? d2(Tp[i].x - Tc[i].x + xi * d.x, Tp[i].y - Tc[i].y + xi * d.y) : d2(0.0, 0.0);
  }
  const double2 lt = d2(lnp.x - lnc.x + xi * dl.x, lnp.y - lnc.y + xi * dl.y);
  linear_geopot_tile<KPT>(nd, rdgas, r, tx, lds, A, G);
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int k = k0 + i;
    const bool ok = live && k < nd;
    const size_t o = (size_t)k * spec_per_level + idx;
    const double hp = ok ? r.h[k] * ps_ref : 0.0;
    double2 d = DV[i];
    if (energy && ok) {
      // Apply_Laplacian!(spe_energy); spe_δdiv .-= spe_energy  (SD:309-310)
      const double2 lap = d2(e * EN[i].x, e * EN[i].y);
      energy[o] = lap;
      d = d2sub(d, lap);
    }
    A[i] = d2(d.x - e * (G[i].x + hp * lt.x), d.y - e * (G[i].y + hp * lt.y));
  }
  // wave-matrix solve for n <= num_wavenumbers (Semi_Implicit.jl:487-497):
  // every level of the column is needed, so the operand goes through LDS
#pragma unroll
  for (int i = 0; i < KPT; ++i)
    if (k0 + i < nd) lds[(k0 + i) * COLS + tx.lane] = A[i];
  __syncthreads();
  if (live && n >= m && n <= num_wavenumbers) {
    // the level qq of the operand is read once for the KPT rows of this
    // thread (each row still sums over qq in increasing order); two loops, so
    // that each reads through one address space (an LDS/global select would
    // turn every read into a generic access)
    const int kmax = nd - 1;
    auto solve = [&](const double* Wn) {
      double2 s[KPT];
#pragma unroll
      for (int i = 0; i < KPT; ++i) s[i] = d2(0.0, 0.0);
#pragma unroll 4
      for (int qq = 0; qq < nd; ++qq) {
        const double2 x = lds[qq * COLS + tx.lane];
#pragma unroll
        for (int i = 0; i < KPT; ++i) {
          const double w = Wn[min(k0 + i, kmax) * nd + qq];
          s[i] = d2(s[i].x + w * x.x, s[i].y + w * x.y);
        }
      }
#pragma unroll
      for (int i = 0; i < KPT; ++i)
        if (k0 + i < nd) A[i] = s[i];
    };
    if (stage_w) solve(Wl + (size_t)(n - n_lo) * nnd);
    else solve(wave + (size_t)n * nnd);
  }
  __syncthreads();
  double2 Dfin[KPT];
#pragma unroll
  for (int i = 0; i < KPT; ++i) Dfin[i] = A[i];
  linear_ps_t_tile<KPT>(nd, kappa, r, tx, lds, A, dps, B);
  const double fx = xi / ps_ref;
  double2 lq = d2(dl.x + fx * dps.x, dl.y + fx * dps.y);
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    if (!(live && k0 + i < nd)) continue;
    const size_t o = (size_t)(k0 + i) * spec_per_level + idx;
    double2 qd = Dfin[i];
    double2 qt = d2(DD[i].x + xi * B[i].x, DD[i].y + xi * B[i].y);
    if (lf.on) {
      double2 cn, nn;
      leapfrog_reg(qd, Dp[i], Dc[i], dmp, true, lf.init_step, lf.dt, lf.r, cn, nn);
      if (lf.advance) lf.div_pw[o] = cn;
      else lf.div_c[o] = cn;
      lf.div_n[o] = nn;
      leapfrog_reg(qt, Tp[i], Tc[i], dmp, true, lf.init_step, lf.dt, lf.r, cn, nn);
      if (lf.advance) lf.t_pw[o] = cn;
      else lf.t_c[o] = cn;
      lf.t_n[o] = nn;
      double2 qv = QV[i];
      leapfrog_reg(qv, VP[i], VCc[i], dmp, true, lf.init_step, lf.dt, lf.r, cn, nn);
      lf.dvor[o] = qv;
      if (lf.advance) lf.vor_pw[o] = cn;
      else lf.vor_c[o] = cn;
      lf.vor_n[o] = nn;
    }
    ddiv[o] = qd;
    dtt[o] = qt;
  }
  if (live && tx.q == 0) {
    dlnps[idx] = lq;
    if (lf.on) {
      double2 cn, nn;
      leapfrog_reg(lq, lnp, lnc, 0.0, false, lf.init_step, lf.dt, lf.r, cn, nn);
      if (lf.advance) lf.lnps_pw[idx] = cn;
      else lf.lnps_c[idx] = cn;
      lf.lnps_n[idx] = nn;
    }
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

// Tile shape of the grid column kernels for nd levels: KPT levels per thread,
// ceil(nd / KPT) level groups of GCOLS threads (at most 512 threads).
static int tile_kpt(int nd) {
  int k = 2;
  while ((nd + k - 1) / k * GCOLS > MAXNQ * TILE) k *= 2;
  return k;
}
static dim3 tile_block(int nd) {
  int kpt = tile_kpt(nd);
  return dim3(GCOLS * ((nd + kpt - 1) / kpt));
}
static dim3 tile_grid(size_t ncol) { return dim3((unsigned)((ncol + GCOLS - 1) / GCOLS)); }

#define JGCM_TILE_DISPATCH(nd, KERNEL, ...)                                          \
  switch (tile_kpt(nd)) {                                                            \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                       \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                       \
    default: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                      \
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

// The reference's two fixed sigma tables (b at the nd+1 half levels, top to
// surface; a = 0): the old model's 14 levels (Compute_Old_Model_Sigma,
// Vert_Coordinate.jl:157-165) and the 18 levels of v197 (Compute_V197_Sigma,
// :143-154).
static const double kSigmaMcm[15] = {0.0,    0.03,   0.0707, 0.1311, 0.2102,
                                     0.3036, 0.4062, 0.5138, 0.6226, 0.7284,
                                     0.8255, 0.9066, 0.9640, 0.9933, 1.0};
static const double kSigmaV197[19] = {0.0,       0.0089163, 0.0342936, 0.0740741, 0.1262002,
                                      0.1886145, 0.2592592, 0.3360768, 0.4170096, 0.5000000,
                                      0.5829904, 0.6639231, 0.7407407, 0.8113854, 0.8737997,
                                      0.9259259, 0.9657064, 0.9910837, 1.0};

// Compute_Uneven_Sigma (Vert_Coordinate.jl:120-141) as its docstring states
// it: ζ_k = 1 - k/nd, b_k = exp(-(surf_res ζ + (1 - surf_res) ζ^exponent)
// scale_heights), b_nd = 1, b_0 = 0 when zero_top; a = 0.  (The code
// evaluates exp(+…), which gives b > 1; the docstring's form is the one that
// yields a coordinate.)
static void uneven_sigma(int nd, double scale_heights, double surf_res, double exponent,
                         bool zero_top, std::vector<double>& b) {
  b.resize(nd + 1);
  for (int k = 0; k <= nd; ++k) {
    const double zeta = (double)(nd - k) / nd;
    const double z = -(surf_res * zeta + (1.0 - surf_res) * std::pow(zeta, exponent));
    b[k] = std::exp(z * scale_heights);
  }
  b[nd] = 1.0;
  if (zero_top) b[0] = 0.0;
}

// Transition (Vert_Coordinate.jl:93-110): 0 at or above p_press, 1 at or
// below p_sigma, sin² in between, on the σ values p.
static double transition(double p, double p_sigma, double p_press) {
  if (p <= p_press) return 0.0;
  if (p >= p_sigma) return 1.0;
  const double x = p - p_press, xx = p_sigma - p_press;
  const double s = std::sin(0.5 * M_PI * x / xx);
  return s * s;
}

}  // namespace jgcm

using namespace jgcm;

extern "C" {

int jgcm_vert_coord_create(jgcm_mesh_t* mesh, int nd, int vert_coord_option,
                           int vert_difference_option, int vert_advect_scheme, double p_ref,
                           int zero_top, jgcm_vert_coord_t** out) {
  return jgcm_vert_coord_create_ex(mesh, nd, vert_coord_option, vert_difference_option,
                                   vert_advect_scheme, p_ref, zero_top, 4.0, 1.0, 0.1, 0.3, 2.5,
                                   out);
}

int jgcm_vert_coord_create_ex(jgcm_mesh_t* mesh, int nd, int vert_coord_option,
                              int vert_difference_option, int vert_advect_scheme, double p_ref,
                              int zero_top, double scale_heights, double surf_res,
                              double p_press, double p_sigma, double exponent,
                              jgcm_vert_coord_t** out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(out != nullptr, "out is NULL");
  JGCM_REQUIRE(nd >= 1 && nd <= MAXND, "nd must be in 1..%d", MAXND);
  // Compute_Vert_Coord (Vert_Coordinate.jl:61-87)
  const double* table = nullptr;
  int table_nd = 0;
  std::vector<double> ua, ub;  // uneven_sigma / hybrid: a and b at the half levels
  switch (vert_coord_option) {
    case JGCM_VC_EVEN_SIGMA:
      break;
    case JGCM_VC_MCM:
      table = kSigmaMcm;
      table_nd = 14;
      break;
    case JGCM_VC_V197:
      table = kSigmaV197;
      table_nd = 18;
      break;
    case JGCM_VC_UNEVEN_SIGMA:
      // :69-70 (the reference passes a, b before they exist; called here with nd)
      JGCM_REQUIRE(scale_heights > 0.0 && exponent > 0.0, "scale_heights and exponent must be > 0");
      uneven_sigma(nd, scale_heights, surf_res, exponent, zero_top != 0, ub);
      ua.assign(nd + 1, 0.0);
      break;
    case JGCM_VC_HYBRID: {
      // :71-77 with nd supplied: (a_σ, b_σ) = uneven σ without a zero top,
      // (b_p, a_p) = the same pair swapped (pressure levels), blended by
      // Transition(b_σ) — a = p_ref b_σ (1 - trans), b = b_σ trans
      JGCM_REQUIRE(scale_heights > 0.0 && exponent > 0.0, "scale_heights and exponent must be > 0");
      JGCM_REQUIRE(p_sigma > p_press, "p_sigma must exceed p_press");
      std::vector<double> bs;
      uneven_sigma(nd, scale_heights, surf_res, exponent, false, bs);
      ua.resize(nd + 1);
      ub.resize(nd + 1);
      for (int k = 0; k <= nd; ++k) {
        const double tr = transition(bs[k], p_sigma, p_press);
        ua[k] = p_ref * (bs[k] * (1.0 - tr));
        ub[k] = bs[k] * tr;
      }
      break;
    }
    default:
      set_error("Compute_Vert_Coord option %d is not a valid value for option", vert_coord_option);
      return JGCM_ERR_STATE;
  }
  JGCM_REQUIRE(table == nullptr || nd == table_nd,
               "the %s coordinate has %d levels, got nd = %d",
               vert_coord_option == JGCM_VC_MCM ? "mcm" : "v197", table_nd, nd);
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
  // Compute_Even_Sigma (Vert_Coordinate.jl:112-117): a = 0, b = LinRange(0, 1, nd+1);
  // the tables: a = 0, b as listed; uneven σ and hybrid as computed above
  v->ak.assign(nd + 1, 0.0);
  v->bk.resize(nd + 1);
  for (int k = 0; k <= nd; ++k) v->bk[k] = table ? table[k] : (double)k / nd;
  v->bk[nd] = 1.0;
  if (!ub.empty()) {
    v->ak = ua;
    v->bk = ub;
  }
  v->dak.resize(nd);
  v->dbk.resize(nd);
  for (int k = 0; k < nd; ++k) {
    v->dak[k] = v->ak[k + 1] - v->ak[k];
    v->dbk[k] = v->bk[k + 1] - v->bk[k];
  }
  // pure sigma constants: ln p_half_k = ln b_k + ln ps, and the
  // Simmons-Burridge ln p_full_k = ln p_half_{k+1} + p_half_k (ln p_half_{k+1}
  // - ln p_half_k) / Δp_k - 1 = ln ps + cf_k (p_half_k / Δp_k = b_k / Δb_k)
  bool pure = v->zero_top || v->bk[0] > 0.0;
  for (int k = 0; k <= nd; ++k) pure = pure && v->ak[k] == 0.0;
  v->pure_sigma = pure ? 1 : 0;
  std::vector<double> lnb(nd + 1, 0.0), cf(nd, 0.0), sf(nd, 0.0);
  if (pure) {
    for (int k = 0; k <= nd; ++k) lnb[k] = v->bk[k] > 0.0 ? std::log(v->bk[k]) : 0.0;
    for (int k = 0; k < nd; ++k) {
      if (v->zero_top && k == 0)
        cf[k] = lnb[1] - 1.0;   // :57-60
      else
        cf[k] = lnb[k + 1] + v->bk[k] * (lnb[k + 1] - lnb[k]) / v->dbk[k] - 1.0;
      sf[k] = std::exp(cf[k]);
    }
  }
  v->cf = cf;
  std::vector<double> all;
  all.insert(all.end(), v->ak.begin(), v->ak.end());
  all.insert(all.end(), v->bk.begin(), v->bk.end());
  all.insert(all.end(), v->dak.begin(), v->dak.end());
  all.insert(all.end(), v->dbk.begin(), v->dbk.end());
  all.insert(all.end(), lnb.begin(), lnb.end());
  all.insert(all.end(), cf.begin(), cf.end());
  all.insert(all.end(), sf.begin(), sf.end());
  hipError_t e = hipMalloc(&v->d_coef, all.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(v->d_coef, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("vert coord upload failed: %s", hipGetErrorString(e));
    delete v;
    return JGCM_ERR_HIP;
  }
  mesh_retain(mesh);   // released by jgcm_vert_coord_destroy
  *out = v;
  return JGCM_OK;
}

int jgcm_vert_coord_destroy(jgcm_vert_coord_t* v) {
  if (!v) return JGCM_OK;
  if (v->mesh) (void)hipStreamSynchronize(v->mesh->stream);
  if (v->d_coef) (void)hipFree(v->d_coef);
  jgcm_mesh* m = v->mesh;
  delete v;
  mesh_release(m);
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
  JGCM_TILE_DISPATCH(v->nd, hs_pressure_heights_kernel, tile_grid(P), tile_block(v->nd), 0, M->stream,
                     vc_of(v), M->nlat, M->nlon, 0.0, 0.0, grid_ps, nullptr, nullptr, p_half, dp,
                     lnp_half, p_full, lnp_full, nullptr, nullptr);
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
  JGCM_TILE_DISPATCH(v->nd, hs_pressure_heights_kernel, tile_grid(P), tile_block(v->nd), 0, M->stream,
                     vc_of(v), M->nlat, M->nlon, atmo->rdgas, atmo->grav, grid_ps, grid_t,
                     grid_geopots, p_half, dp, lnp_half, p_full, lnp_full, z_full, z_half);
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
  for (int k = 0; k < nd; ++k)
    dev.push_back(v->pure_sigma ? std::exp(si->atmo.kappa * v->cf[k]) : 1.0);
  for (int k = 0; k < nd; ++k) dev.push_back(1.0 / dp[k]);
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

// Column tiles of the semi-implicit solve: SI_COLS spectral coefficients per
// workgroup (half a wave, so that T85's 7482 coefficients per level make 234
// workgroups), levels split into groups of si_kpt(nd) per thread.
#ifndef JGCM_SI_COLS
#define JGCM_SI_COLS 32
#endif
constexpr int SI_COLS = JGCM_SI_COLS;
static int si_kpt(int nd) {
  int k = 1;
  while ((nd + k - 1) / k * SI_COLS > 512) k *= 2;
  return k;
}

static SIRef siref(const jgcm_semi_implicit* si) {
  int nd = si->nd;
  const double* b = si->d_ref;
  return SIRef{b, b + nd, b + nd + (nd + 1), b + nd + (nd + 1) + nd, b + nd + (nd + 1) + 2 * nd,
               si->vc->d_coef + (nd + 1), b + 6 * nd + 1};
}

// (p_full_k / ps)^κ per level (pure sigma), after the reference profiles and
// before their reciprocal Δp
static const double* skap_of(const jgcm_semi_implicit* si) { return si->d_ref + 5 * si->nd + 1; }

static int implicit_correction(jgcm_semi_implicit_t* si, const double* div_c, const double* div_p,
                               const double* lnps_c, const double* lnps_p, const double* t_c,
                               const double* t_p, double* ddiv, double* dlnps, double* dt,
                               double* energy, LFFuse lf = LFFuse{}, int tri = 0) {
  JGCM_REQUIRE(si && div_c && div_p && lnps_c && lnps_p && t_c && t_p && ddiv && dlnps && dt,
               "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  size_t S = M->spec_elems();
  double xi = xi_of(si->lf);
  const int nd = si->nd;
  size_t lds = (size_t)nd * SI_COLS * sizeof(double2) + (size_t)7 * (nd + 1) * sizeof(double);
  // wave matrices of at most SI_COLS / nm + 2 degrees per tile, staged when they fit
  const size_t nw = std::min<size_t>((size_t)M->ns + 1, SI_COLS / (size_t)M->nm + 2);
  const size_t wbytes = nw * nd * nd * sizeof(double);
  const int stage_w = lds + wbytes <= 48 * 1024 ? 1 : 0;
  if (stage_w) lds += wbytes;
  const int kpt = si_kpt(nd);
  const dim3 grid((unsigned)((S + SI_COLS - 1) / SI_COLS)), block(SI_COLS * ((nd + kpt - 1) / kpt));
#define JGCM_SI_LAUNCH(K)                                                                       \
  hipLaunchKernelGGL((hs_implicit_kernel<K, SI_COLS>), grid, block, lds, M->stream, nd, M->nm,  \
                     M->ns, S, si->atmo.kappa, si->atmo.rdgas, si->ps_ref, xi, siref(si),        \
                     M->d_eig, si->d_wave, si->num_wavenumbers, (const double2*)div_c,           \
                     (const double2*)div_p, (const double2*)lnps_c, (const double2*)lnps_p,      \
                     (const double2*)t_c, (const double2*)t_p, (double2*)ddiv, (double2*)dlnps,  \
                     (double2*)dt, (double2*)energy, stage_w, tri, lf)
  switch (kpt) {
    case 1: JGCM_SI_LAUNCH(1); break;
    case 2: JGCM_SI_LAUNCH(2); break;
    case 4: JGCM_SI_LAUNCH(4); break;
    default: JGCM_SI_LAUNCH(8); break;
  }
#undef JGCM_SI_LAUNCH
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// Implicit_Correction!(semi_implicit, vert_coord, atmo_data, ...)  Semi_Implicit.jl:462-513
int jgcm_implicit_correction(jgcm_semi_implicit_t* si, const double* div_c, const double* div_p,
                             const double* lnps_c, const double* lnps_p, const double* t_c,
                             const double* t_p, double* ddiv, double* dlnps, double* dt) {
  return implicit_correction(si, div_c, div_p, lnps_c, lnps_p, t_c, t_p, ddiv, dlnps, dt, nullptr);
}

// Apply_Laplacian!(spe_energy); spe_δdiv .-= spe_energy; Implicit_Correction!  (SD:309-316)
int jgcm_hs_energy_implicit_leapfrog(jgcm_semi_implicit_t* si, double* spe_energy, double* div_c,
                                     const double* div_p, double* lnps_c, const double* lnps_p,
                                     double* t_c, const double* t_p, double* ddiv, double* dlnps,
                                     double* dt, double* dvor, const double* vor_p, double* vor_c,
                                     double* vor_n, double* div_n, double* lnps_n, double* t_n) {
  JGCM_REQUIRE(spe_energy && dvor && vor_p && vor_c && vor_n && div_n && lnps_n && t_n,
               "NULL argument");
  jgcm_leapfrog_t* L = si->lf;
  LFFuse lf{(double2*)dvor, (const double2*)vor_p, (double2*)vor_c, (double2*)vor_n,
            (double2*)div_c, (double2*)div_n, (double2*)t_c, (double2*)t_n, (double2*)lnps_c,
            (double2*)lnps_n, L->d_damping, 1, L->init_step ? 1 : 0, (double)L->dt,
            L->robert_coef};
  return implicit_correction(si, div_c, div_p, lnps_c, lnps_p, t_c, t_p, ddiv, dlnps, dt,
                             spe_energy, lf, 1);
}

int jgcm_hs_implicit_leapfrog_advance(jgcm_semi_implicit_t* si, double* spe_energy, double* ddiv,
                                      double* dlnps, double* dt, double* dvor,
                                      double* const* vor, double* const* div,
                                      double* const* lnps, double* const* t) {
  JGCM_REQUIRE(si && spe_energy && ddiv && dlnps && dt && dvor && vor && div && lnps && t,
               "NULL argument");
  for (int k = 0; k < 3; ++k)
    JGCM_REQUIRE(vor[k] && div[k] && lnps[k] && t[k], "NULL field pointer (time level %d)", k);
  jgcm_leapfrog_t* L = si->lf;
  LFFuse lf{(double2*)dvor, (const double2*)vor[0], (double2*)vor[1], (double2*)vor[2],
            (double2*)div[1], (double2*)div[2], (double2*)t[1], (double2*)t[2], (double2*)lnps[1],
            (double2*)lnps[2], L->d_damping, 1, L->init_step ? 1 : 0, (double)L->dt,
            L->robert_coef, 1, (double2*)vor[0], (double2*)div[0], (double2*)t[0],
            (double2*)lnps[0]};
  return implicit_correction(si, div[1], div[0], lnps[1], lnps[0], t[1], t[0], ddiv, dlnps, dt,
                             spe_energy, lf, 1);
}

int jgcm_hs_energy_implicit_correction(jgcm_semi_implicit_t* si, double* spe_energy,
                                       const double* div_c, const double* div_p,
                                       const double* lnps_c, const double* lnps_p,
                                       const double* t_c, const double* t_p, double* ddiv,
                                       double* dlnps, double* dt) {
  JGCM_REQUIRE(spe_energy != nullptr, "NULL argument");
  return implicit_correction(si, div_c, div_p, lnps_c, lnps_p, t_c, t_p, ddiv, dlnps, dt,
                             spe_energy);
}

// per-workgroup partial sums of the area means live in the column-sum scratch
static double2* area_partials(const jgcm_semi_implicit* si) {
  return reinterpret_cast<double2*>(si->d_colsum);
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
  {
    const int given = (p_half != nullptr) + (dp != nullptr) + (lnp_half != nullptr) +
                      (p_full != nullptr) + (lnp_full != nullptr);
    JGCM_REQUIRE(given == 0 || given == 5,
                 "the five pressure variables must all be given or all be NULL");
  }
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  HSParams h = hs_params(hp, si->atmo.cp_air);
  int mass = si->atmo.do_mass_correction, energy = si->atmo.do_energy_correction;
  int* bad = reinterpret_cast<int*>(si->d_scal + 15);
  int nd = si->nd;
  JGCM_TILE_DISPATCH(nd, hs_physics_kernel, tile_grid(P), tile_block(nd), 0, M->stream,
                     vc_of(si->vc), M->nlat, M->nlon, M->d_sin, M->d_wts, si->atmo.kappa,
                     si->atmo.cp_air, (double)dt, h, grid_ps_c, grid_u_p, grid_v_p, grid_t_p,
                     grid_ps_p, p_half, dp, lnp_half, p_full, lnp_full, grid_du, grid_dv, grid_t_eq,
                     grid_dt, grid_dps, area_partials(si), do_forcing, mass, energy, bad,
                     skap_of(si));
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
                            double* grid_energy, const double* grid_dlambda_t,
                            const double* grid_dtheta_t) {
  JGCM_REQUIRE(si, "NULL argument");
  JGCM_REQUIRE((grid_dlambda_t == nullptr) == (grid_dtheta_t == nullptr),
               "grid_dlambda_t and grid_dtheta_t must both be given or both be NULL");
  {
    const int given = (dp != nullptr) + (lnp_half != nullptr) + (lnp_full != nullptr) +
                      (p_full != nullptr);
    JGCM_REQUIRE(given == 0 || given == 4,
                 "dp, lnp_half, lnp_full and p_full must all be given or all be NULL");
  }
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  int nd = si->nd;
#define JGCM_COLUMN_DYNAMICS_LAUNCH(KERNEL)                                                    \
  JGCM_TILE_DISPATCH(nd, KERNEL, tile_grid(P), tile_block(nd), 0, M->stream, vc_of(si->vc),     \
                     M->nlat, M->nlon, si->atmo.rdgas, si->atmo.cp_air, 2 * si->atmo.omega,     \
                     M->d_sin, grid_div, grid_u, grid_v, grid_ps, dp, lnp_half, lnp_full, p_full, \
                     grid_dlambda_ps, grid_dtheta_ps, grid_t, grid_vor, grid_geopots,           \
                     grid_M_half, grid_w_full, grid_du, grid_dv, grid_dps, grid_dt, geopot_half, \
                     geopot_full, grid_dlnps, grid_absvor, grid_energy, grid_dlambda_t,         \
                     grid_dtheta_t)
  if (dp == nullptr) {
    JGCM_COLUMN_DYNAMICS_LAUNCH(hs_column_dynamics_own_p_kernel);
  } else {
    JGCM_COLUMN_DYNAMICS_LAUNCH(hs_column_dynamics_kernel);
  }
#undef JGCM_COLUMN_DYNAMICS_LAUNCH
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// Spectral_Dynamics_Physics! and the column part of Spectral_Dynamics! in one
// pass (the physics tendencies stay in registers); see jgcm.h
static int physics_column_dynamics(
    jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, int64_t dt, const double* grid_u_p,
    const double* grid_v_p, const double* grid_t_p, const double* grid_ps_p, double* grid_t_eq,
    const double* grid_div, const double* grid_u, const double* grid_v, const double* grid_ps,
    double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t, const double* grid_vor,
    const double* grid_geopots, double* grid_M_half, double* grid_w_full, double* grid_du,
    double* grid_dv, double* grid_dps, double* grid_dt, double* geopot_half, double* geopot_full,
    double* grid_dlnps, double* grid_absvor, double* grid_energy, const double* grid_dlambda_t,
    const double* grid_dtheta_t, const double* dlambda_lnps, const double* dtheta_lnps,
    int advance_p) {
  JGCM_REQUIRE(si && hp, "NULL argument");
  JGCM_REQUIRE(grid_u_p && grid_v_p && grid_t_p && grid_ps_p && grid_t_eq && grid_div && grid_u &&
                   grid_v && grid_ps && grid_dlambda_ps && grid_dtheta_ps && grid_t && grid_vor &&
                   grid_geopots && grid_M_half && grid_w_full && grid_du && grid_dv && grid_dps &&
                   grid_dt && geopot_half && geopot_full && grid_dlnps && grid_absvor &&
                   grid_energy,
               "NULL field pointer");
  JGCM_REQUIRE((grid_dlambda_t == nullptr) == (grid_dtheta_t == nullptr),
               "grid_dlambda_t and grid_dtheta_t must both be given or both be NULL");
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  int nd = si->nd;
  PhysArgs pa{M->d_wts, si->atmo.kappa, si->atmo.cp_air, (double)dt,
              hs_params(hp, si->atmo.cp_air),
              grid_u_p, grid_v_p, grid_t_p, grid_ps_p, grid_t_eq, area_partials(si),
              si->atmo.do_mass_correction, si->atmo.do_energy_correction,
              reinterpret_cast<int*>(si->d_scal + 15), skap_of(si), dlambda_lnps, dtheta_lnps,
              advance_p ? 1 : 0};
  const double* none = nullptr;
  JGCM_TILE_DISPATCH(nd, hs_physics_column_dynamics_kernel, tile_grid(P), tile_block(nd), 0,
                     M->stream, pa, vc_of(si->vc), M->nlat, M->nlon, si->atmo.rdgas,
                     si->atmo.cp_air, 2 * si->atmo.omega, M->d_sin, grid_div, grid_u, grid_v,
                     grid_ps, none, none, none, none, grid_dlambda_ps, grid_dtheta_ps, grid_t,
                     grid_vor, grid_geopots, grid_M_half, grid_w_full, grid_du, grid_dv, grid_dps,
                     grid_dt, geopot_half, geopot_full, grid_dlnps, grid_absvor, grid_energy,
                     grid_dlambda_t, grid_dtheta_t);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_hs_physics_column_dynamics(
    jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, int64_t dt, const double* grid_u_p,
    const double* grid_v_p, const double* grid_t_p, const double* grid_ps_p, double* grid_t_eq,
    const double* grid_div, const double* grid_u, const double* grid_v, const double* grid_ps,
    double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t, const double* grid_vor,
    const double* grid_geopots, double* grid_M_half, double* grid_w_full, double* grid_du,
    double* grid_dv, double* grid_dps, double* grid_dt, double* geopot_half, double* geopot_full,
    double* grid_dlnps, double* grid_absvor, double* grid_energy, const double* grid_dlambda_t,
    const double* grid_dtheta_t) {
  return physics_column_dynamics(si, hp, dt, grid_u_p, grid_v_p, grid_t_p, grid_ps_p, grid_t_eq,
                                 grid_div, grid_u, grid_v, grid_ps, grid_dlambda_ps, grid_dtheta_ps,
                                 grid_t, grid_vor, grid_geopots, grid_M_half, grid_w_full, grid_du,
                                 grid_dv, grid_dps, grid_dt, geopot_half, geopot_full, grid_dlnps,
                                 grid_absvor, grid_energy, grid_dlambda_t, grid_dtheta_t, nullptr,
                                 nullptr, 0);
}

int jgcm_hs_physics_column_dynamics_ex(
    jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, int64_t dt, double* grid_u_p,
    double* grid_v_p, double* grid_t_p, double* grid_ps_p, double* grid_t_eq,
    const double* grid_div, const double* grid_u, const double* grid_v, const double* grid_ps,
    double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t, const double* grid_vor,
    const double* grid_geopots, double* grid_M_half, double* grid_w_full, double* grid_du,
    double* grid_dv, double* grid_dps, double* grid_dt, double* geopot_half, double* geopot_full,
    double* grid_dlnps, double* grid_absvor, double* grid_energy, const double* grid_dlambda_t,
    const double* grid_dtheta_t, const double* dlambda_lnps, const double* dtheta_lnps,
    int advance_p) {
  JGCM_REQUIRE((dlambda_lnps == nullptr) == (dtheta_lnps == nullptr),
               "dlambda_lnps and dtheta_lnps must both be given or both be NULL");
  return physics_column_dynamics(si, hp, dt, grid_u_p, grid_v_p, grid_t_p, grid_ps_p, grid_t_eq,
                                 grid_div, grid_u, grid_v, grid_ps, grid_dlambda_ps, grid_dtheta_ps,
                                 grid_t, grid_vor, grid_geopots, grid_M_half, grid_w_full, grid_du,
                                 grid_dv, grid_dps, grid_dt, geopot_half, geopot_full, grid_dlnps,
                                 grid_absvor, grid_energy, grid_dlambda_t, grid_dtheta_t,
                                 dlambda_lnps, dtheta_lnps, advance_p);
}

// Compute_Corrections! (SD:30-68) after the back-transforms, with ps_n = exp(lnps) (SD:336)
int jgcm_hs_corrections(jgcm_semi_implicit_t* si, const double* grid_lnps, double* grid_ps_n,
                        const double* grid_u_n, const double* grid_v_n, double* grid_t_n,
                        double* spe_lnps_n, double* spe_t_n) {
  JGCM_REQUIRE(si, "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  size_t P = M->grid_elems();
  int mass = si->atmo.do_mass_correction, energy = si->atmo.do_energy_correction;
  int nd = si->nd;
  VC v = vc_of(si->vc);
  // ps_n = exp(lnps) and the correction sums (one pass), then the corrections; the
  // physics pass left the partials of Compute_Corrections_Init at the start of the
  // column-sum scratch, these go after them
  const int nparts = (int)tile_grid(P).x;
  double* part_p = si->d_colsum;
  double* part_n = si->d_colsum + 2 * (size_t)nparts;
  JGCM_TILE_DISPATCH(nd, hs_corrections_sums_kernel, tile_grid(P), tile_block(nd), 0, M->stream, v,
                     M->nlat, M->nlon, si->atmo.cp_air, M->d_wts, grid_lnps, grid_ps_n, grid_u_n,
                     grid_v_n, grid_t_n, part_n, mass || energy, energy);
  JGCM_LAUNCH_CHECK();
  if (mass || energy) {
    const size_t total = (energy ? (size_t)si->nd * P : 0) + (mass ? P : 0);
    const size_t per_block = (size_t)JGCM_CADV_PER_THREAD * APPLY_THREADS;
    const int blocks = (int)((total + per_block - 1) / per_block);
    hipLaunchKernelGGL(hs_apply_corrections_kernel, dim3(blocks), dim3(APPLY_THREADS), 0, M->stream,
                       si->nd, P, M->spec_elems(), M->nlon, grid_t_n, grid_ps_n, (double2*)spe_t_n,
                       (double2*)spe_lnps_n, part_p, part_n, nparts, si->d_scal, si->atmo.grav,
                       si->atmo.cp_air, mass, energy);
    JGCM_LAUNCH_CHECK();
  }
  return JGCM_OK;
}

int jgcm_hs_corrections_advance(jgcm_semi_implicit_t* si, const double* grid_lnps,
                                double* const* p, double* const* c, double* const* n,
                                const size_t* count) {
  JGCM_REQUIRE(si && grid_lnps && p && c && n && count, "NULL argument");
  jgcm_mesh* M = si->vc->mesh;
  for (int f = 0; f < 8; ++f)
    JGCM_REQUIRE(c[f] && n[f], "NULL field pointer (field %d)", f);
  const size_t P = M->grid_elems(), S = M->spec_elems();
  const int nd = si->nd;
  JGCM_REQUIRE(count[0] == nd * S * 2 && count[1] == nd * S * 2 && count[2] == S * 2 &&
                   count[3] == nd * S * 2 && count[4] == nd * P && count[5] == nd * P &&
                   count[6] == P && count[7] == nd * P,
               "field sizes do not match the Time_Advance! order (vor, div, lnps, T spectral; "
               "u, v, ps, T grid)");
  int mass = si->atmo.do_mass_correction, energy = si->atmo.do_energy_correction;
  VC v = vc_of(si->vc);
  const int nparts = (int)tile_grid(P).x;
  double* part_p = si->d_colsum;
  double* part_n = si->d_colsum + 2 * (size_t)nparts;
  // ps_n = exp(ln ps) and the sums of the new state (n[4] u, n[5] v, n[7] T, n[6] ps)
  JGCM_TILE_DISPATCH(nd, hs_corrections_sums_kernel, tile_grid(P), tile_block(nd), 0, M->stream, v,
                     M->nlat, M->nlon, si->atmo.cp_air, M->d_wts, grid_lnps, n[6], n[4], n[5], n[7],
                     part_n, mass || energy, energy);
  JGCM_LAUNCH_CHECK();
  CorrAdvanceSet s{};
  static const int kinds[8] = {0, 0, 4, 3, 0, 0, 2, 1};
  size_t maxc = 0;
  for (int f = 0; f < 8; ++f) {
    s.p[f] = p[f];
    s.c[f] = c[f];
    s.n[f] = n[f];
    s.count[f] = count[f];
    s.kind[f] = kinds[f];
    maxc = std::max(maxc, count[f]);
  }
  const int per_block = JGCM_CADV_PER_THREAD * APPLY_THREADS;
  const dim3 grid((unsigned)((maxc + per_block - 1) / per_block), 8);
  hipLaunchKernelGGL(hs_correct_advance_kernel, grid, dim3(APPLY_THREADS), 0, M->stream, s,
                     per_block, 2 * S, M->nlon, part_p, part_n, nparts, si->d_scal, si->atmo.grav,
                     si->atmo.cp_air, mass, energy, 7);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_hs_scalars_device(jgcm_semi_implicit_t* si, double* dev_out, int count) {
  JGCM_REQUIRE(si && dev_out && count >= 1 && count <= 16, "bad argument");
  jgcm_mesh* M = si->vc->mesh;
  JGCM_CHECK_HIP(hipMemcpyAsync(dev_out, si->d_scal, count * sizeof(double), hipMemcpyDeviceToDevice,
                                M->stream));
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
