// FP64 CPU restatement of the Held-Suarez step (TEST INFRASTRUCTURE ONLY: the
// CPU baseline that bench.py times beside the GPU step, checked against the
// numpy oracle by tests/test_oracle_cpu.py; no product code links it).
//
// hs_step <- Atmosphere_Update!  Spectral_Dynamics.jl:470-490, i.e.
//   Spectral_Dynamics_Physics!  SD:452-467 with HS_Forcing! HS_Forcing.jl:1-130
//   Spectral_Dynamics!          SD:206-365:
//     Compute_Corrections_Init!   SD:5-28      Pressure_Variables!  Press_And_Geopot.jl:36-67
//     Compute_Gradients!          SSM:721-741  Four_In_One!         SD:88-183
//     Compute_Geopotential!       Press_And_Geopot.jl:71-111
//     Vert_Advection!             Vert_Coordinate.jl:171-210
//     Add_Horizontal_Advection!   SSM:694-718  Vor_Div_From_Grid_UV!  SSM:349-459
//     Apply_Laplacian!            SSM:583-621
//     Implicit_Correction!        Semi_Implicit.jl:360-404, 462-513 (linear operators :213-341)
//     Compute_Spectral_Damping! / Filtered_Leapfrog!  Time_Integrator.jl:58-101
//     UV_Grid_From_Vor_Div!       SSM:462-580  Compute_Corrections!  SD:30-68
//     Time_Advance!               Dyn_Data.jl:326-364
//   Compute_Pressures_And_Heights!  Press_And_Geopot.jl:3-19
// following oracle/primitive.py statement by statement (same operand order
// where it fixes the rounding), with the transforms of oracle/sht_cpu.cpp.
// The semi-implicit matrices W_n, h and the reference profiles are the numpy
// oracle's (SemiImplicitSolver, Semi_Implicit.jl:40-188): set up once per run,
// not part of a step.
//
// Threads: every grid loop runs over the columns of the grid, every spectral
// loop over the (n, m) coefficients, the transforms as in sht_cpu.cpp.
// Layouts are the oracle's (C order): grid [nlev][nlat][nlon]; spectral complex
// [nlev][ns+1][N+1]; coefficient tables [N+1][ns+1] (m, n); Legendre-type
// tables [N+1][ns+1][nlat].
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "fft_cpu.hpp"

extern "C" int sht_forward(int N, int ns, int nlat, int nlon, int nlev, const double* qwg,
                           const double* grid, double* snm, int threads);
extern "C" int sht_inverse(int N, int ns, int nlat, int nlon, int nlev, const double* qnm,
                           const double* snm, double* grid, int threads);

// Field order is mirrored by oracle/cpu.py (_HsConfig).
struct HsConfig {
  int N, ns, nlat, nlon, nd, num_wavenumbers;
  int zero_top, centered;          // vert_advect_scheme: 0 second_centered_wts, 1 second_centered
  int do_mass, do_energy, init_step;
  double radius, grav, rdgas, cp_air, kappa;
  double dt, robert, implicit_coef, ps_ref;
  double sigma_b, k_f, k_a, k_s, dT_y, dth_z;
  const double *ak, *bk;                                        // [nd+1]
  const double *qnm, *qwg, *alp_a, *alp_b;                      // [N+1][ns+1][nlat]
  const double *uvc, *uvm, *uvp, *dlam, *dthm, *dthp, *eig, *damping;   // [N+1][ns+1]
  const double *wts, *cos_theta, *sin_theta, *coriolis;         // [nlat]
  const double *dp_ref, *lnp_half_ref, *lnp_full_ref, *t_ref, *h;  // [nd], [nd+1], [nd], [nd], [nd]
  const double* wave_matrix;                                    // [num_wavenumbers+1][nd][nd]
};

// The Dyn_Data arrays a step touches (Dyn_Data.jl:158-324), named as in the
// oracle with δ -> d, Δ -> D, λ -> lam, θ -> th.  Field order is mirrored by
// oracle/cpu.py (HS_STATE).
struct HsState {
  // grid, nd levels
  double *grid_u_p, *grid_v_p, *grid_t_p, *grid_u_c, *grid_v_c, *grid_t_c;
  double *grid_u_n, *grid_v_n, *grid_t_n, *grid_du, *grid_dv, *grid_dt, *grid_t_eq;
  double *grid_p_full, *grid_lnp_full, *grid_Dp, *grid_w_full, *grid_geopot_full;
  double *grid_energy_full, *grid_d_full1, *grid_vor, *grid_div, *grid_absvor, *grid_z_full;
  // grid, nd + 1 half levels
  double *grid_p_half, *grid_lnp_half, *grid_M_half, *grid_geopot_half, *grid_z_half;
  // grid, one level
  double *grid_ps_p, *grid_ps_c, *grid_ps_n, *grid_dps, *grid_dlnps, *grid_lnps;
  double *grid_dlam_ps, *grid_dth_ps, *grid_geopots;
  // spectral (complex as interleaved doubles), nd levels
  double *spe_vor_p, *spe_vor_c, *spe_vor_n, *spe_div_p, *spe_div_c, *spe_div_n;
  double *spe_t_p, *spe_t_c, *spe_t_n, *spe_dvor, *spe_ddiv, *spe_dt, *spe_energy;
  // spectral, one level
  double *spe_lnps_p, *spe_lnps_c, *spe_lnps_n, *spe_dlnps;
};

namespace {

using jgcm_oracle::cd;
using jgcm_oracle::Fft;

constexpr int BL = 64;    // columns per block of the column loops (levels apart by a whole grid)

struct Step {
  const HsConfig& c;
  HsState& s;
  int threads;
  int N, ns, nm, nsp, nlat, nlon, nd, P, S;
  double dt, xi;            // get_dt / get_xi (TI:46-56)
  bool bad_sigma = false;

  Step(const HsConfig& c_, HsState& s_, int th)
      : c(c_), s(s_), threads(th), N(c_.N), ns(c_.ns), nm(c_.N + 1), nsp(c_.ns + 1),
        nlat(c_.nlat), nlon(c_.nlon), nd(c_.nd), P(c_.nlat * c_.nlon), S((c_.ns + 1) * (c_.N + 1)) {
    dt = c.init_step ? c.dt : 2.0 * c.dt;
    xi = c.init_step ? c.dt * c.implicit_coef : 2.0 * c.dt * c.implicit_coef;
  }

  static cd* sp(double* p) { return reinterpret_cast<cd*>(p); }

  // ---------------------------------------------------------- transforms
  void inverse(const double* snm, double* grid, int nlev) const {
    sht_inverse(N, ns, nlat, nlon, nlev, c.qnm, snm, grid, threads);
  }
  void forward(const double* grid, double* snm, int nlev) const {
    sht_forward(N, ns, nlat, nlon, nlev, c.qwg, grid, snm, threads);
  }
  void divide_by_cos(double* g, int nlev) const {                   // SSM:333-338
#pragma omp parallel for schedule(static)
    for (int kj = 0; kj < nlev * nlat; ++kj) {
      const double cj = c.cos_theta[kj % nlat];
      double* row = g + (size_t)kj * nlon;
      for (int i = 0; i < nlon; ++i) row[i] /= cj;
    }
  }

  // Compute_Gradient_Cos! SSM:655-692 then Compute_Gradients! SSM:721-741
  void compute_gradients(const double* hs_, double* gl, double* gt, int nlev) const {
    const cd* hs = reinterpret_cast<const cd*>(hs_);
    std::vector<cd> a((size_t)nlev * S), b((size_t)nlev * S);
#pragma omp parallel for schedule(static)
    for (int kn = 0; kn < nlev * nsp; ++kn) {
      const int k = kn / nsp, n = kn % nsp;
      for (int m = 0; m < nm; ++m) {
        const size_t o = ((size_t)k * nsp + n) * nm + m;
        const int t = m * nsp + n;
        a[o] = c.dlam[t] * (cd(0.0, 1.0) * hs[o]);
        cd v = n < ns ? c.dthp[t] * hs[o + nm] : cd(0.0, 0.0);
        if (n >= 1) v += c.dthm[t] * hs[o - nm];
        b[o] = v;
      }
    }
    inverse(reinterpret_cast<double*>(a.data()), gl, nlev);
    inverse(reinterpret_cast<double*>(b.data()), gt, nlev);
    divide_by_cos(gl, nlev);
    divide_by_cos(gt, nlev);
  }

  // Compute_Ucos_Vcos_From_Vor_Div! SSM:517-580, UV_Grid_From_Vor_Div! SSM:462-482
  void uv_grid_from_vor_div(const double* vor_, const double* div_, double* gu, double* gv) const {
    const cd* vor = reinterpret_cast<const cd*>(vor_);
    const cd* div = reinterpret_cast<const cd*>(div_);
    std::vector<cd> uc((size_t)nd * S), vc((size_t)nd * S);
#pragma omp parallel for schedule(static)
    for (int kn = 0; kn < nd * nsp; ++kn) {
      const int k = kn / nsp, n = kn % nsp;
      for (int m = 0; m < nm; ++m) {
        const size_t o = ((size_t)k * nsp + n) * nm + m;
        const int t = m * nsp + n;
        cd u = c.uvc[t] * (cd(0.0, 1.0) * div[o]);
        cd v = c.uvc[t] * (cd(0.0, 1.0) * vor[o]);
        if (n >= 1) {
          u -= c.uvm[t] * vor[o - nm];
          v += c.uvm[t] * div[o - nm];
        }
        if (n < ns) {
          u -= c.uvp[t] * vor[o + nm];
          v += c.uvp[t] * div[o + nm];
        }
        uc[o] = u;
        vc[o] = v;
      }
    }
    inverse(reinterpret_cast<double*>(uc.data()), gu, nd);
    inverse(reinterpret_cast<double*>(vc.data()), gv, nd);
    divide_by_cos(gu, nd);
    divide_by_cos(gv, nd);
  }

  // Vor_Div_From_Grid_UV! SSM:349-393 with Compute_Alpha_Operator! SSM:396-459
  void vor_div_from_grid_uv(const double* gu, const double* gv, double* vor_, double* div_) const {
    cd* vor = sp(vor_);
    cd* div = sp(div_);
    // fft(u cos θ) / nλ and fft(v cos θ) / nλ, [m][k][j], one complex transform per row pair
    const size_t nf = (size_t)nd * nlat * nm;
    std::vector<double> fur(nf), fui(nf), fvr(nf), fvi(nf);
    const Fft fft(nlon);
#pragma omp parallel
    {
      std::vector<cd> z(nlon), tmp;
#pragma omp for schedule(static)
      for (int kj = 0; kj < nd * nlat; ++kj) {
        const double cj = c.cos_theta[kj % nlat];
        const double* a = gu + (size_t)kj * nlon;
        const double* b = gv + (size_t)kj * nlon;
        for (int i = 0; i < nlon; ++i) z[i] = cd(a[i] * cj, b[i] * cj);
        fft.run(z.data(), -1, tmp);
        for (int m = 0; m < nm; ++m) {
          const cd zm = z[m], zc = std::conj(z[(nlon - m) % nlon]);
          const cd d = zm - zc;
          const cd U = 0.5 * (zm + zc) / (double)nlon;
          const cd V = cd(0.5 * d.imag(), -0.5 * d.real()) / (double)nlon;
          const size_t o = (size_t)m * nd * nlat + kj;
          fur[o] = U.real();
          fui[o] = U.imag();
          fvr[o] = V.real();
          fvi[o] = V.imag();
        }
      }
      const double f = 1.0 / (2.0 * c.radius);
#pragma omp for schedule(dynamic, 1)
      for (int m = 0; m < nm; ++m)
        for (int k = 0; k < nd; ++k)
          for (int n = m; n <= ns; ++n) {
            const double* qa = c.alp_a + ((size_t)m * nsp + n) * nlat;
            const double* qb = c.alp_b + ((size_t)m * nsp + n) * nlat;
            const size_t b = ((size_t)m * nd + k) * nlat;
            const double *Ur = fur.data() + b, *Ui = fui.data() + b;
            const double *Vr = fvr.data() + b, *Vi = fvi.data() + b;
            double uar = 0, uai = 0, ubr = 0, ubi = 0, var = 0, vai = 0, vbr = 0, vbi = 0;
#pragma omp simd reduction(+ : uar, uai, ubr, ubi, var, vai, vbr, vbi)
            for (int j = 0; j < nlat; ++j) {
              uar += Ur[j] * qa[j];
              uai += Ui[j] * qa[j];
              ubr += Ur[j] * qb[j];
              ubi += Ui[j] * qb[j];
              var += Vr[j] * qa[j];
              vai += Vi[j] * qa[j];
              vbr += Vr[j] * qb[j];
              vbi += Vi[j] * qb[j];
            }
            const cd ua(uar, uai), ub(ubr, ubi), va(var, vai), vb(vbr, vbi);
            const size_t o = ((size_t)k * nsp + n) * nm + m;
            vor[o] = f * (cd(0.0, 1.0) * va - ub);        // isign = -1 (SSM:388)
            div[o] = f * (cd(0.0, 1.0) * ua + vb);        // isign = +1 (SSM:389)
          }
    }
    for (int k = 0; k < nd; ++k)                          // SSM:391-392
      for (int m = 0; m < nm; ++m) {
        vor[((size_t)k * nsp + ns) * nm + m] = 0.0;
        div[((size_t)k * nsp + ns) * nm + m] = 0.0;
      }
  }

  // Area_Weighted_Global_Mean SSM:746-760 (level 1)
  double area_weighted_global_mean(const double* g) const {
    double acc = 0.0;
    for (int i = 0; i < nlon; ++i) {
      double col = 0.0;
      for (int j = 0; j < nlat; ++j) col += g[(size_t)j * nlon + i] * c.wts[j];
      acc += col;
    }
    return acc / (2.0 * nlon);
  }

  // Mass_Weighted_Global_Integral  Vert_Coordinate.jl:215-238
  double mass_weighted_global_integral(const double* e, const double* ps) const {
    std::vector<double> vi(P);
#pragma omp parallel for schedule(static)
    for (int p0 = 0; p0 < P; p0 += BL) {
      const int nb = std::min(BL, P - p0);
      for (int q = 0; q < nb; ++q) vi[p0 + q] = 0.0;
      for (int k = 0; k < nd; ++k) {
        const double dak = c.ak[k + 1] - c.ak[k], dbk = c.bk[k + 1] - c.bk[k];
        const double* ek = e + (size_t)k * P + p0;
        for (int q = 0; q < nb; ++q) vi[p0 + q] += ek[q] * (dak + dbk * ps[p0 + q]);
      }
    }
    return area_weighted_global_mean(vi.data()) / c.grav;
  }

  // ------------------------------------------------------------ columns
  // Pressure_Variables!  Press_And_Geopot.jl:36-67
  void pressure_variables() const {
    const int kt = c.zero_top ? 1 : 0;
#pragma omp parallel for schedule(static)
    for (int p0 = 0; p0 < P; p0 += BL) {
      const int nb = std::min(BL, P - p0);
      const double* ps = s.grid_ps_c + p0;
      for (int k = 0; k <= nd; ++k) {
        double* ph = s.grid_p_half + (size_t)k * P + p0;
        double* lh = s.grid_lnp_half + (size_t)k * P + p0;
        for (int q = 0; q < nb; ++q) ph[q] = c.ak[k] + c.bk[k] * ps[q];
        if (k >= kt)
          for (int q = 0; q < nb; ++q) lh[q] = std::log(ph[q]);
      }
      for (int k = 0; k < nd; ++k) {
        const size_t o = (size_t)k * P + p0;
        const double* ph = s.grid_p_half + o;
        const double* lh = s.grid_lnp_half + o;
        double* dp = s.grid_Dp + o;
        double* lf = s.grid_lnp_full + o;
        for (int q = 0; q < nb; ++q) dp[q] = ph[q + P] - ph[q];
        if (k >= kt)
          for (int q = 0; q < nb; ++q) lf[q] = lh[q + P] + ph[q] * (lh[q + P] - lh[q]) / dp[q] - 1.0;
      }
      if (c.zero_top)
        for (int q = 0; q < nb; ++q) {
          s.grid_lnp_half[p0 + q] = 0.0;
          s.grid_lnp_full[p0 + q] = s.grid_lnp_half[P + p0 + q] - 1.0;
        }
      for (int k = 0; k < nd; ++k) {
        const size_t o = (size_t)k * P + p0;
        for (int q = 0; q < nb; ++q) s.grid_p_full[o + q] = std::exp(s.grid_lnp_full[o + q]);
      }
    }
  }

  // Compute_Geopotential!  Press_And_Geopot.jl:71-111 (temperature t)
  void compute_geopotential(const double* t, double* gf, double* gh) const {
    const double R = c.rdgas;
    const int kt = c.zero_top ? 1 : 0;
#pragma omp parallel for schedule(static)
    for (int p0 = 0; p0 < P; p0 += BL) {
      const int nb = std::min(BL, P - p0);
      for (int q = 0; q < nb; ++q) gh[(size_t)nd * P + p0 + q] = s.grid_geopots[p0 + q];
      if (c.zero_top)
        for (int q = 0; q < nb; ++q) gh[p0 + q] = 0.0;
      for (int k = nd - 1; k >= kt; --k) {
        const size_t o = (size_t)k * P + p0;
        const double* lh = s.grid_lnp_half + o;
        for (int q = 0; q < nb; ++q) gh[o + q] = gh[o + P + q] + R * t[o + q] * (lh[q + P] - lh[q]);
      }
      for (int k = 0; k < nd; ++k) {
        const size_t o = (size_t)k * P + p0;
        const double* lh = s.grid_lnp_half + o;
        const double* lf = s.grid_lnp_full + o;
        for (int q = 0; q < nb; ++q) gf[o + q] = gh[o + P + q] + R * t[o + q] * (lh[q + P] - lf[q]);
      }
    }
  }

  // HS_Forcing!  HS_Forcing.jl:1-33, Rayleigh_Damping! :101-130, Newtonian_Damping! :37-97
  void hs_forcing() {
    const double sb = c.sigma_b, day = 86400.0;
    const double kf = c.k_f / day, ka = c.k_a / day, ks = c.k_s / day;
    const double t_zero = 315.0, t_strat = 200.0, p_ref = 1.0e5;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int kj = 0; kj < nd * nlat; ++kj) {
      const int j = kj % nlat;
      const double s2 = c.sin_theta[j] * c.sin_theta[j], c2 = 1.0 - s2, c4 = c2 * c2;
      for (int i = 0; i < nlon; ++i) {
        const size_t o = (size_t)kj * nlon + i;
        const double ps = s.grid_p_half[(size_t)nd * P + (size_t)j * nlon + i];
        const double sig = s.grid_p_full[o] / ps;
        if ((sig - sb) / (1.0 - sb) > 1.0) bad = 1;
        const double kv = kf / (1.0 - sb) * std::max(0.0, sig - sb);
        const double u = s.grid_u_p[o], v = s.grid_v_p[o];
        const double du = -kv * u, dv = -kv * v;
        s.grid_du[o] = du;
        s.grid_dv[o] = dv;
        double dT = -((u + 0.5 * du * dt) * du + (v + 0.5 * dv * dt) * dv) / c.cp_air;
        const double pn = s.grid_p_full[o] / p_ref;
        const double te = (t_zero - c.dT_y * s2 - c.dth_z * c2 * std::log(pn)) * std::pow(pn, c.kappa);
        const double teq = std::max(t_strat, te);
        s.grid_t_eq[o] = teq;
        const double kt = ka + (ks - ka) / (1.0 - sb) * std::max(0.0, sig - sb) * c4;
        dT -= kt * (s.grid_t_p[o] - teq);
        s.grid_dt[o] = dT;
      }
    }
    bad_sigma = bad_sigma || bad;
  }

  // Four_In_One!  SD:88-183
  void four_in_one() const {
    const double R = c.rdgas, kappa = c.rdgas / c.cp_air;
#pragma omp parallel for schedule(static)
    for (int p0 = 0; p0 < P; p0 += BL) {
      const int nb = std::min(BL, P - p0);
      const double* dlam_ps = s.grid_dlam_ps + p0;
      const double* dth_ps = s.grid_dth_ps + p0;
      double dmean_tot[BL];
      for (int q = 0; q < nb; ++q) dmean_tot[q] = 0.0;
      for (int k = 0; k < nd; ++k) {
        const size_t o0 = (size_t)k * P + p0;
        for (int q = 0; q < nb; ++q) {
          const size_t o = o0 + q;
          const double dp = s.grid_Dp[o];
          const double lh0 = s.grid_lnp_half[o], lh1 = s.grid_lnp_half[o + P], lf = s.grid_lnp_full[o];
          const double dlnp_p = lh1 - lf, dlnp_m = lf - lh0, dlnp = lh1 - lh0;
          const double x1 = (c.bk[k] * dlnp_m + c.bk[k + 1] * dlnp_p) / dp;
          const double dl = x1 * dlam_ps[q], dth = x1 * dth_ps[q];
          const double t = s.grid_t_c[o], u = s.grid_u_c[o], v = s.grid_v_c[o];
          s.grid_du[o] -= R * t * dl;
          s.grid_dv[o] -= R * t * dth;
          const double dmean = s.grid_div[o] * dp + (c.bk[k + 1] - c.bk[k]) * (u * dlam_ps[q] + v * dth_ps[q]);
          const double x5 = -(dmean_tot[q] * dlnp + dmean * dlnp_p) / dp + u * dl + v * dth;
          s.grid_dt[o] += kappa * t * x5;
          s.grid_w_full[o] = x5 * s.grid_p_full[o];
          dmean_tot[q] = dmean_tot[q] + dmean;
          s.grid_M_half[o + P] = -dmean_tot[q];
        }
      }
      for (int q = 0; q < nb; ++q) s.grid_dps[p0 + q] -= dmean_tot[q];
      for (int k = 0; k < nd - 1; ++k) {
        double* mh = s.grid_M_half + (size_t)(k + 1) * P + p0;
        for (int q = 0; q < nb; ++q) mh[q] += dmean_tot[q] * c.bk[k + 1];
      }
      for (int q = 0; q < nb; ++q) {
        s.grid_M_half[p0 + q] = 0.0;
        s.grid_M_half[(size_t)nd * P + p0 + q] = 0.0;
      }
    }
  }

  // Vert_Advection!  Vert_Coordinate.jl:171-210: acc += the tendency of r
  void vert_advection_add(const double* r, double* acc) const {
    const double* dz = s.grid_Dp;
    const double* w = s.grid_M_half;
#pragma omp parallel for schedule(static)
    for (int p0 = 0; p0 < P; p0 += BL) {
      const int nb = std::min(BL, P - p0);
      double flux_k[BL];                           // flux at half level k (0 at the top)
      for (int q = 0; q < nb; ++q) flux_k[q] = 0.0;
      for (int k = 0; k < nd; ++k) {
        const size_t o0 = (size_t)k * P + p0;
        for (int q = 0; q < nb; ++q) {
          const size_t o = o0 + q;
          double flux_k1 = 0.0;                    // half level k + 1 (0 at the surface)
          if (k + 1 < nd) {
            const double r0 = r[o], r1 = r[o + P], w1 = w[o + P];
            flux_k1 = c.centered ? w1 * (r0 + r1) / 2.0
                                 : w1 * (r0 + (r1 - r0) * dz[o] / (dz[o] + dz[o + P]));
          }
          acc[o] += (flux_k[q] - flux_k1 + r[o] * (w[o + P] - w[o])) / dz[o];
          flux_k[q] = flux_k1;
        }
      }
    }
  }

  // --------------------------------------------------------- spectral
  // Linear_Ps_T_Ddiv!  Semi_Implicit.jl:271-341 on one coefficient's column
  void linear_ps_t_ddiv(const cd* x, cd& dps, cd* dtv, cd* Mh) const {
    cd dmean_tot = 0.0;
    for (int k = 0; k < nd; ++k) {
      const double dlnp_p = c.lnp_half_ref[k + 1] - c.lnp_full_ref[k];
      const double dlnp = c.lnp_half_ref[k + 1] - c.lnp_half_ref[k];
      const cd dmean = x[k] * c.dp_ref[k];
      dtv[k] = -c.kappa * c.t_ref[k] * (dmean_tot * dlnp + dmean * dlnp_p) / c.dp_ref[k];
      dmean_tot = dmean_tot + dmean;
      Mh[k + 1] = -dmean_tot;
    }
    dps = -dmean_tot;
    for (int k = 0; k < nd - 1; ++k) Mh[k + 1] += dmean_tot * c.bk[k + 1];
    Mh[0] = 0.0;
    Mh[nd] = 0.0;
    cd mdt_k = 0.0;                                // Mdt_half[k], 0 at the top
    for (int k = 0; k < nd; ++k) {
      const cd mdt_k1 = k + 1 < nd ? Mh[k + 1] * (c.t_ref[k + 1] - c.t_ref[k]) : cd(0.0, 0.0);
      dtv[k] -= 0.5 * (mdt_k1 + mdt_k) / c.dp_ref[k];
      mdt_k = mdt_k1;
    }
  }

  // Linear_Geopot_Dt!  Semi_Implicit.jl:213-241 (geopot_half[nd] = 0)
  void linear_geopot_dt(const cd* x, cd* g) const {
    const double R = c.rdgas;
    cd gh1 = 0.0;                                  // geopot_half[k + 1]
    for (int k = nd - 1; k >= 0; --k) {
      g[k] = gh1 + R * (x[k] * (c.lnp_half_ref[k + 1] - c.lnp_full_ref[k]));
      if (k >= 1) gh1 = gh1 + R * (x[k] * (c.lnp_half_ref[k + 1] - c.lnp_half_ref[k]));
    }
  }

  // Implicit_Correction!  Semi_Implicit.jl:462-513 with Adjust_Dlnps_Dt_Ddiv! :360-404
  void implicit_correction() const {
    cd* div_c = sp(s.spe_div_c);
    cd* div_p = sp(s.spe_div_p);
    cd* t_c = sp(s.spe_t_c);
    cd* t_p = sp(s.spe_t_p);
    cd* ddiv = sp(s.spe_ddiv);
    cd* dT = sp(s.spe_dt);
    cd* dlnps = sp(s.spe_dlnps);
    const cd* lnps_c = sp(s.spe_lnps_c);
    const cd* lnps_p = sp(s.spe_lnps_p);
#pragma omp parallel
    {
      std::vector<cd> x(nd), y(nd), g(nd), Mh(nd + 1);
#pragma omp for schedule(static)
      for (int o = 0; o < S; ++o) {
        const int n = o / nm, m = o % nm;
        for (int k = 0; k < nd; ++k) x[k] = div_p[(size_t)k * S + o] - div_c[(size_t)k * S + o];
        cd dps;
        linear_ps_t_ddiv(x.data(), dps, y.data(), Mh.data());
        for (int k = 0; k < nd; ++k) dT[(size_t)k * S + o] += y[k];
        dlnps[o] += dps / c.ps_ref;
        for (int k = 0; k < nd; ++k)
          x[k] = t_p[(size_t)k * S + o] - t_c[(size_t)k * S + o] + xi * dT[(size_t)k * S + o];
        const cd dlnps_tmp = lnps_p[o] - lnps_c[o] + xi * dlnps[o];
        linear_geopot_dt(x.data(), g.data());
        const double eig = c.eig[m * nsp + n];
        for (int k = 0; k < nd; ++k)
          ddiv[(size_t)k * S + o] -= eig * (g[k] + c.h[k] * c.ps_ref * dlnps_tmp);
        if (n >= m && n <= c.num_wavenumbers) {      // ddiv[:, n, m] = W_n ddiv[:, n, m]
          const double* W = c.wave_matrix + (size_t)n * nd * nd;
          for (int k = 0; k < nd; ++k) x[k] = ddiv[(size_t)k * S + o];
          for (int r = 0; r < nd; ++r) {
            cd acc = 0.0;
            for (int k = 0; k < nd; ++k) acc += W[(size_t)r * nd + k] * x[k];
            ddiv[(size_t)r * S + o] = acc;
          }
        }
        for (int k = 0; k < nd; ++k) x[k] = ddiv[(size_t)k * S + o];
        linear_ps_t_ddiv(x.data(), dps, y.data(), Mh.data());
        for (int k = 0; k < nd; ++k) dT[(size_t)k * S + o] += xi * y[k];
        dlnps[o] += xi / c.ps_ref * dps;
      }
    }
  }

  // Compute_Spectral_Damping!  TI:58-81
  void spectral_damping(const double* Qc_, const double* Qp_, double* dQ_) const {
    const cd* Qc = reinterpret_cast<const cd*>(Qc_);
    const cd* Qp = reinterpret_cast<const cd*>(Qp_);
    cd* dQ = sp(dQ_);
#pragma omp parallel for schedule(static)
    for (int ko = 0; ko < nd * S; ++ko) {
      const int o = ko % S;
      const double d = c.damping[(o % nm) * nsp + o / nm];
      if (c.init_step)
        dQ[ko] = (dQ[ko] - Qc[ko] * d) / (1.0 + c.dt * d);
      else
        dQ[ko] = (dQ[ko] - Qp[ko] * d) / (1.0 + 2.0 * c.dt * d);
    }
  }

  // Filtered_Leapfrog!  TI:86-101
  void filtered_leapfrog(const double* dQ_, const double* Qp_, double* Qc_, double* Qn_, int nlev) const {
    const cd* dQ = reinterpret_cast<const cd*>(dQ_);
    const cd* Qp = reinterpret_cast<const cd*>(Qp_);
    cd* Qc = sp(Qc_);
    cd* Qn = sp(Qn_);
    const double r = c.robert;
#pragma omp parallel for schedule(static)
    for (int ko = 0; ko < nlev * S; ++ko) {
      if (c.init_step) {
        Qn[ko] = Qc[ko] + c.dt * dQ[ko];
        Qc[ko] += r * (-1.0 * Qc[ko] + Qn[ko]);
      } else {
        Qc[ko] += r * (Qp[ko] - 2.0 * Qc[ko]);
        Qn[ko] = Qp[ko] + 2.0 * c.dt * dQ[ko];
        Qc[ko] += r * Qn[ko];
      }
    }
  }

  // ------------------------------------------------------------ the step
  int run() {
    const size_t G = (size_t)nd * P, SP = (size_t)nd * S * 2;
    // Spectral_Dynamics_Physics!  SD:452-467
    std::fill(s.grid_dps, s.grid_dps + P, 0.0);
    hs_forcing();
    // Spectral_Dynamics!  SD:206-365
    // Compute_Corrections_Init!  SD:5-28
    double mean_ps_p = 0.0, mean_energy_p = 0.0;
    if (c.do_mass || c.do_energy) mean_ps_p = area_weighted_global_mean(s.grid_ps_p);
    if (c.do_energy) {
#pragma omp parallel for schedule(static)
      for (size_t o = 0; o < G; ++o) {
        const double u = s.grid_u_p[o] + dt * s.grid_du[o], v = s.grid_v_p[o] + dt * s.grid_dv[o];
        s.grid_energy_full[o] = 0.5 * (u * u + v * v) + c.cp_air * (s.grid_t_p[o] + dt * s.grid_dt[o]);
      }
      mean_energy_p = mass_weighted_global_integral(s.grid_energy_full, s.grid_ps_p);
    }
    pressure_variables();
    compute_gradients(s.spe_lnps_c, s.grid_dlam_ps, s.grid_dth_ps, 1);
    for (int p = 0; p < P; ++p) {
      s.grid_dlam_ps[p] *= s.grid_ps_c[p];
      s.grid_dth_ps[p] *= s.grid_ps_c[p];
    }
    four_in_one();
    compute_geopotential(s.grid_t_c, s.grid_geopot_full, s.grid_geopot_half);
    for (int p = 0; p < P; ++p) s.grid_dlnps[p] = s.grid_dps[p] / s.grid_ps_c[p];
    forward(s.grid_dlnps, s.spe_dlnps, 1);
    vert_advection_add(s.grid_u_c, s.grid_du);
    vert_advection_add(s.grid_v_c, s.grid_dv);
    vert_advection_add(s.grid_t_c, s.grid_dt);
    {  // Add_Horizontal_Advection!  SSM:694-718
      std::vector<double> gl(G), gt(G);
      compute_gradients(s.spe_t_c, gl.data(), gt.data(), nd);
#pragma omp parallel for schedule(static)
      for (size_t o = 0; o < G; ++o) s.grid_dt[o] -= (s.grid_u_c[o] * gl[o] + s.grid_v_c[o] * gt[o]);
    }
    forward(s.grid_dt, s.spe_dt, nd);
#pragma omp parallel for schedule(static)
    for (size_t o = 0; o < G; ++o) {               // Compute_Abs_Vor!  Atmo_Data.jl:59-65
      const double av = s.grid_vor[o] + c.coriolis[(o / nlon) % nlat];
      s.grid_absvor[o] = av;
      s.grid_du[o] += av * s.grid_v_c[o];
      s.grid_dv[o] -= av * s.grid_u_c[o];
    }
    vor_div_from_grid_uv(s.grid_du, s.grid_dv, s.spe_dvor, s.spe_ddiv);
#pragma omp parallel for schedule(static)
    for (size_t o = 0; o < G; ++o)
      s.grid_energy_full[o] = s.grid_geopot_full[o] + 0.5 * (s.grid_u_c[o] * s.grid_u_c[o] +
                                                             s.grid_v_c[o] * s.grid_v_c[o]);
    forward(s.grid_energy_full, s.spe_energy, nd);
    {
      cd* e = sp(s.spe_energy);
      cd* dd = sp(s.spe_ddiv);
#pragma omp parallel for schedule(static)
      for (size_t ko = 0; ko < (size_t)nd * S; ++ko) {
        const size_t o = ko % S;
        e[ko] *= c.eig[(o % nm) * nsp + o / nm];     // Apply_Laplacian!  SSM:614-621
        dd[ko] -= e[ko];
      }
    }
    implicit_correction();
    spectral_damping(s.spe_vor_c, s.spe_vor_p, s.spe_dvor);
    spectral_damping(s.spe_div_c, s.spe_div_p, s.spe_ddiv);
    spectral_damping(s.spe_t_c, s.spe_t_p, s.spe_dt);
    filtered_leapfrog(s.spe_dvor, s.spe_vor_p, s.spe_vor_c, s.spe_vor_n, nd);
    filtered_leapfrog(s.spe_ddiv, s.spe_div_p, s.spe_div_c, s.spe_div_n, nd);
    filtered_leapfrog(s.spe_dlnps, s.spe_lnps_p, s.spe_lnps_c, s.spe_lnps_n, 1);
    filtered_leapfrog(s.spe_dt, s.spe_t_p, s.spe_t_c, s.spe_t_n, nd);
    inverse(s.spe_vor_n, s.grid_vor, nd);
    inverse(s.spe_div_n, s.grid_div, nd);
    uv_grid_from_vor_div(s.spe_vor_n, s.spe_div_n, s.grid_u_n, s.grid_v_n);
    inverse(s.spe_t_n, s.grid_t_n, nd);
    inverse(s.spe_lnps_n, s.grid_lnps, 1);
    for (int p = 0; p < P; ++p) s.grid_ps_n[p] = std::exp(s.grid_lnps[p]);
    // Compute_Corrections!  SD:30-68
    if (c.do_mass) {
      const double factor = mean_ps_p / area_weighted_global_mean(s.grid_ps_n);
      for (int p = 0; p < P; ++p) s.grid_ps_n[p] *= factor;
      s.spe_lnps_n[0] += std::log(factor);
    }
    if (c.do_energy) {
#pragma omp parallel for schedule(static)
      for (size_t o = 0; o < G; ++o)
        s.grid_energy_full[o] = 0.5 * (s.grid_u_n[o] * s.grid_u_n[o] + s.grid_v_n[o] * s.grid_v_n[o]) +
                                c.cp_air * s.grid_t_n[o];
      const double mean_energy_n = mass_weighted_global_integral(s.grid_energy_full, s.grid_ps_n);
      const double corr = c.grav * (mean_energy_p - mean_energy_n) / (c.cp_air * mean_ps_p);
#pragma omp parallel for schedule(static)
      for (size_t o = 0; o < G; ++o) s.grid_t_n[o] += corr;
      for (int k = 0; k < nd; ++k) s.spe_t_n[(size_t)k * S * 2] += corr;
    }
    // Time_Advance!  Dyn_Data.jl:326-364
    auto advance = [](double* p, double* c_, const double* n_, size_t count) {
#pragma omp parallel for schedule(static)
      for (size_t o = 0; o < count; ++o) {
        p[o] = c_[o];
        c_[o] = n_[o];
      }
    };
    advance(s.spe_vor_p, s.spe_vor_c, s.spe_vor_n, SP);
    advance(s.spe_div_p, s.spe_div_c, s.spe_div_n, SP);
    advance(s.spe_lnps_p, s.spe_lnps_c, s.spe_lnps_n, (size_t)S * 2);
    advance(s.spe_t_p, s.spe_t_c, s.spe_t_n, SP);
    advance(s.grid_u_p, s.grid_u_c, s.grid_u_n, G);
    advance(s.grid_v_p, s.grid_v_c, s.grid_v_n, G);
    advance(s.grid_ps_p, s.grid_ps_c, s.grid_ps_n, P);
    advance(s.grid_t_p, s.grid_t_c, s.grid_t_n, G);
    pressure_variables();
    // Compute_Pressures_And_Heights!  Press_And_Geopot.jl:3-19
    pressure_variables();
    compute_geopotential(s.grid_t_c, s.grid_z_full, s.grid_z_half);
#pragma omp parallel for schedule(static)
    for (size_t o = 0; o < G; ++o) s.grid_z_full[o] /= c.grav;
    for (size_t o = 0; o < G + P; ++o) s.grid_z_half[o] /= c.grav;
    return bad_sigma ? -2 : 0;
  }
};

}  // namespace

extern "C" {

// One Atmosphere_Update! (SD:470-490) of the state s in place.  Returns 0, -1
// for inconsistent sizes, -2 when a full level lies below the surface
// (σ > 1, the check of HS_Forcing.jl:57 and :118).
int hs_step(const HsConfig* c, HsState* s, int threads) {
  if (c->nlat % 2 || c->N < 0 || c->ns < c->N || c->nd < 2 || c->nlon < 2 * c->N + 1 ||
      c->num_wavenumbers > c->ns)
    return -1;
  jgcm_oracle::set_threads(threads);
  Step st(*c, *s, threads);
  return st.run();
}

int hs_config_size(void) { return (int)sizeof(HsConfig); }
int hs_state_size(void) { return (int)sizeof(HsState); }

}  // extern "C"
