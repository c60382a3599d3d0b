// Reverse-mode pieces of the Held-Suarez step (Atmosphere_Update!,
// Spectral_Dynamics.jl:206-490) that have no counterpart among the transform
// and leapfrog adjoints:
//
//   hs_forcing_adjoint_kernel  transpose of the HS forcing on the p-fields
//                              (HS_Forcing.jl) and of the energy integral of
//                              Compute_Corrections_Init (SD:5-28)
//   hs_column_adjoint_kernel   transpose of the grid-column dynamics of the
//                              step linearised at its state: Four_In_One!
//                              (SD:88-183), the hydrostatic integral, δln ps,
//                              vertical advection (Vert_Coordinate.jl:171-210),
//                              horizontal advection of T, the vorticity flux
//                              and Φ + KE (SD:259-307)
//   hs_corrections_adjoint_kernel  transpose of the mass and energy
//                              corrections (SD:30-68) with respect to the
//                              new-state grid fields
//   implicit_adjoint_kernel    transpose of ∇²(energy), δdiv -= and
//                              Implicit_Correction! (Semi_Implicit.jl:462-513)
//                              per (m, n) from the operator's matrix, which the
//                              host builds by applying the forward to unit
//                              vectors (as Build_Implicit_Matrices does for
//                              its operators, Semi_Implicit.jl:112-160)
//   gradient_cos_adjoint_kernel transpose of Compute_Gradient_Cos! (SSM:677-687)
//
// The vertical coordinate must be pure sigma (ak = 0: even sigma, mcm, v197)
// with a zero top: every pressure is then b_k ps, the ln-pressure
// differences are constants, and ps cancels out of the column dynamics, which
// leaves ps in the forcing's T_eq and in the corrections.
#pragma clang fp contract(off)

#include <cmath>

#include "primitive_internal.hpp"

namespace jgcm {

static inline int nblk_h(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

#define HS_GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// Per-level constants of a pure-sigma, zero-top coordinate (Pressure_Variables!,
// Press_And_Geopot.jl:36-67, at ps = 1; differences of ln p do not depend on ps):
//   Dp = ln p_half[k+1] - ln p_full[k],  Dl = ln p_half[k+1] - ln p_half[k] (k >= 1),
//   C1 = (b_k (ln p_full - ln p_half[k]) + b_{k+1} Dp) / Δb_k  (Four_In_One!'s x1 times ps),
//   sig = p_full[k] / ps.
struct LevelConst {
  double Dp, Dl, C1, sig;
};

__device__ inline LevelConst level_const(const VC& v, int k) {
  LevelConst c;
  double lh1 = log(v.bk[k + 1]);
  double lf;
  if (k == 0) {
    lf = lh1 - 1.0;
    c.Dl = 0.0;
    c.Dp = lh1 - lf;
    c.C1 = (v.bk[1] * c.Dp) / v.dbk[0];
  } else {
    double lh0 = log(v.bk[k]);
    lf = lh1 + v.bk[k] * (lh1 - lh0) / v.dbk[k] - 1.0;
    c.Dl = lh1 - lh0;
    c.Dp = lh1 - lf;
    c.C1 = (v.bk[k] * (lf - lh0) + v.bk[k + 1] * c.Dp) / v.dbk[k];
  }
  c.sig = exp(lf);
  return c;
}

// weight of r_h in the flux at half level h (second_centered_wts: Δb_{h-1}/(Δb_{h-1}+Δb_h))
__device__ inline double vadv_alpha(const VC& v, int h) {
  return v.scheme == 0 ? v.dbk[h - 1] / (v.dbk[h - 1] + v.dbk[h]) : 0.5;
}

struct HsColAdj {
  // state: c-fields [nd][P], grid gradients of T [nd][P] and of ln ps [P], p-fields
  const double *div, *u, *v, *t, *vor, *tl, *tt, *gl, *gt;
  const double *u_p, *v_p, *t_p, *ps_p, *ps_c;
  // cotangents of the grid tendencies δu, δv, δT [nd][P], of Φ+KE [nd][P] and δln ps [P]
  const double *du_b, *dv_b, *dt_b, *e_b, *dlnps_b;
  // cotangents out
  double *div_b, *u_b, *v_b, *t_b, *vor_b, *tl_b, *tt_b, *gl_b, *gt_b;
  double *up_b, *vp_b, *tp_b, *psp_b, *psc_b;
  double* param_b;  // [4][P] or NULL: cotangents of k_f, k_a, k_s, ΔT_y, added per column
};

struct HsAdjScalars {
  double rdgas, cp_air, kappa, grav, two_omega, dt;
  const double* scal_bar;  // device: [1] energy_p cotangent, [2] mean ps_p cotangent
  jgcm_hs_params hp;
};

// Column tiles: 64 columns per workgroup (one per lane), wave q on levels
// q*KPT .. q*KPT+KPT-1.  Forward quantities (ps factored out, pure sigma):
//   Q_k = D_k + u_k gλ + v_k gθ,  S_k = Σ_{j<k} Δb_j Q_j,  S = Σ_j Δb_j Q_j
//   x5_k = -(Dl_k/Δb_k) S_k - Dp_k Q_k + C1_k (u_k gλ + v_k gθ);  δT_k += κ T_k x5_k
//   δu_k -= R T_k C1_k gλ,  δv_k -= R T_k C1_k gθ,  δln ps = -S
//   m_h = -S_h + b_h S (h = 1..nd-1; M_half = ps m),  vertical advection
//   V(r)_k = (f_k - f_{k+1} + r_k (m_{k+1} - m_k)) / Δb_k, f_h = m_h ((1-α_h) r_{h-1} + α_h r_h)
//   δT_k -= u_k Tλ_k + v_k Tθ_k,  δu_k += (ζ_k + f) v_k,  δv_k -= (ζ_k + f) u_k
//   E_k = R Σ_{j>k} Dl_j T_j + R Dp_k T_k + KE_k
// (the forcing on the p-fields is hs_forcing_adjoint_kernel).
// The reverse sweep needs three vertical scans — the prefix S_k, the prefix of
// the energy cotangent and the suffix of S_bar — and per-column sums; each is
// formed from per-wave partials in wave order (level_sums_n), so the result
// differs from a sequential sweep by rounding only.  Outputs must not alias
// inputs (neighbouring levels of the input cotangents are read by other waves).
template <int KPT>
__global__ __launch_bounds__(1024) void hs_column_adjoint_kernel(HsColAdj s, VC v, int nlat,
                                                                 int nlon, const double* sint,
                                                                 const double* wts, HsAdjScalars c) {
  extern __shared__ double red[];  // [4][nq][64]
  const size_t P = (size_t)nlat * nlon;
  const int nd = v.nd;
  const TileIdx t = tile_idx<KPT>(P);
  const bool ok = t.col;
  const size_t col = ok ? t.c : 0;
  const int k0 = t.k0;
  const int j = (int)(col / nlon);
  const double R = c.rdgas, kap = c.kappa;
  const double f = c.two_omega * sint[j];
  const double gl = ok ? s.gl[col] : 0.0, gt = ok ? s.gt[col] : 0.0;

  // slot i of U/V/T holds level k0-1+i (i = 0..KPT); slot i of BU/BV/BT holds
  // δ*_bar / Δb at level k0-1+i (i = 0..KPT+1); out-of-range levels read as 0
  double U[KPT + 1], V[KPT + 1], T[KPT + 1], BU[KPT + 2], BV[KPT + 2], BT[KPT + 2];
#pragma unroll
  for (int i = 0; i < KPT + 2; ++i) {
    const int k = k0 - 1 + i;
    const bool in = ok && k >= 0 && k < nd;
    const size_t o = (size_t)(in ? k : 0) * P + col;
    const double db = in ? v.dbk[k] : 1.0;
    BU[i] = in ? s.du_b[o] / db : 0.0;
    BV[i] = in ? s.dv_b[o] / db : 0.0;
    BT[i] = in ? s.dt_b[o] / db : 0.0;
    if (i <= KPT) {
      U[i] = in ? s.u[o] : 0.0;
      V[i] = in ? s.v[o] : 0.0;
      T[i] = in ? s.t[o] : 0.0;
    }
  }

  // scan 1: S_k at own levels and at k0+KPT, and the column total S
  double Q[KPT], Sl[KPT + 1];
  double part = 0.0;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int k = k0 + i;
    Sl[i] = part;
    Q[i] = 0.0;
    if (ok && k < nd) {
      const size_t o = (size_t)k * P + col;
      Q[i] = s.div[o] + U[i + 1] * gl + V[i + 1] * gt;
      part = part + v.dbk[k] * Q[i];
    }
  }
  Sl[KPT] = part;
  double Stot;
  {
    const double pa[1] = {part};
    double ab[1], be[1];
    level_sums_n<1>(red, t, pa, ab, be);
#pragma unroll
    for (int i = 0; i <= KPT; ++i) Sl[i] = Sl[i] + ab[0];
    Stot = ab[0] + part + be[0];
  }
  double mh[KPT + 1];  // m at half levels k0 .. k0+KPT
#pragma unroll
  for (int i = 0; i <= KPT; ++i) {
    const int h = k0 + i;
    mh[i] = (h >= 1 && h <= nd - 1) ? -Sl[i] + v.bk[h] * Stot : 0.0;
  }

  const double lb = ok ? -s.dlnps_b[col] : 0.0;  // S_bar from δln ps = -S

  // level-local terms
  double Qbp[KPT], Sbl[KPT], E[KPT], DL[KPT], ubp[KPT], vbp[KPT], tbp[KPT];
  double esum = 0.0, sbsum = 0.0, mbsum = 0.0;
  double glb = 0.0, gtb = 0.0;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int k = k0 + i;
    Qbp[i] = Sbl[i] = E[i] = DL[i] = ubp[i] = vbp[i] = tbp[i] = 0.0;
    if (!(ok && k < nd)) continue;
    const size_t o = (size_t)k * P + col;
    const LevelConst L = level_const(v, k);
    DL[i] = L.Dl;
    const double db = v.dbk[k];
    const double u = U[i + 1], vv = V[i + 1], tk = T[i + 1], vor = s.vor[o];
    const double a = s.du_b[o], b = s.dv_b[o], cc = s.dt_b[o], e = s.e_b[o];
    const double mk = mh[i], mk1 = mh[i + 1];
    // vertical advection of u, v, T: contributions to r_bar at level k
    const double bu = BU[i + 1], bv = BV[i + 1], bt = BT[i + 1];
    double vbu = bu * (mk1 - mk), vbv = bv * (mk1 - mk), vbt = bt * (mk1 - mk);
    double mb = 0.0;  // m_bar at half level k
    if (k >= 1) {
      const double bum = BU[i], bvm = BV[i], btm = BT[i];
      const double al = vadv_alpha(v, k);
      const double um = U[i], vm = V[i], tm = T[i];
      vbu = vbu + (bu - bum) * mk * al;
      vbv = vbv + (bv - bvm) * mk * al;
      vbt = vbt + (bt - btm) * mk * al;
      mb = mb + (bum * um - bu * u + (bu - bum) * ((1.0 - al) * um + al * u));
      mb = mb + (bvm * vm - bv * vv + (bv - bvm) * ((1.0 - al) * vm + al * vv));
      mb = mb + (btm * tm - bt * tk + (bt - btm) * ((1.0 - al) * tm + al * tk));
    }
    if (k + 1 <= nd - 1) {
      const double al1 = vadv_alpha(v, k + 1);
      vbu = vbu + (BU[i + 2] - bu) * mk1 * (1.0 - al1);
      vbv = vbv + (BV[i + 2] - bv) * mk1 * (1.0 - al1);
      vbt = vbt + (BT[i + 2] - bt) * mk1 * (1.0 - al1);
    }
    // energy conversion κ T x5
    const double uvg = u * gl + vv * gt;
    const double x5 = (k >= 1 ? -(L.Dl / db) * Sl[i] : 0.0) - L.Dp * Q[i] + L.C1 * uvg;
    const double xb = kap * tk * cc;
    if (k >= 1) {
      Sbl[i] = -mb - (L.Dl / db) * xb;
      mbsum = mbsum + v.bk[k] * mb;
    }
    Qbp[i] = -L.Dp * xb;
    // local cotangents
    s.vor_b[o] = vv * a - u * b;
    ubp[i] = -((vor + f) * b) - s.tl[o] * cc + u * e + L.C1 * gl * xb + vbu;
    vbp[i] = (vor + f) * a - s.tt[o] * cc + vv * e + L.C1 * gt * xb + vbv;
    tbp[i] = R * L.Dp * e - R * L.C1 * (gl * a + gt * b) + kap * x5 * cc + vbt;
    s.tl_b[o] = -(u * cc);
    s.tt_b[o] = -(vv * cc);
    glb = glb - R * L.C1 * tk * a + L.C1 * u * xb;
    gtb = gtb - R * L.C1 * tk * b + L.C1 * vv * xb;
    E[i] = e;
    esum = esum + e;
    sbsum = sbsum + Sbl[i];
  }

  // scans 2 and 3: prefix of the energy cotangent, suffix of S_bar, S_bar total
  double ab[3], be[3];
  {
    const double pa[3] = {esum, sbsum, mbsum};
    level_sums_n<3>(red, t, pa, ab, be);
  }
  const double Stot_b = lb + (ab[2] + mbsum + be[2]);
  double eprefix = ab[0];
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    if (k0 + i >= 1) tbp[i] = tbp[i] + R * DL[i] * eprefix;
    eprefix = eprefix + E[i];
  }
  // Q_bar = partial + Δb_k (Σ_{k'>k} S_bar + S_bar_tot), bottom up
  double suffix = be[1];
#pragma unroll
  for (int i = KPT - 1; i >= 0; --i) {
    const int k = k0 + i;
    if (!(ok && k < nd)) continue;
    const size_t o = (size_t)k * P + col;
    const double qb = Qbp[i] + v.dbk[k] * (suffix + Stot_b);
    const double u = U[i + 1], vv = V[i + 1];
    s.div_b[o] = qb;
    s.u_b[o] = ubp[i] + gl * qb;
    s.v_b[o] = vbp[i] + gt * qb;
    s.t_b[o] = tbp[i];
    glb = glb + u * qb;
    gtb = gtb + vv * qb;
    suffix = suffix + Sbl[i];
  }

  // per-column sums over the waves; wave 0 writes
  double ca[2], cb[2];
  {
    const double pa[2] = {glb, gtb};
    level_sums_n<2>(red, t, pa, ca, cb);
  }
  if (ok && t.q == 0) {
    s.gl_b[col] = glb + cb[0];
    s.gt_b[col] = gtb + cb[1];
  }
}

// Held-Suarez forcing on the p-fields (HS_Forcing.jl:1-130: Rayleigh damping
// of u_p, v_p below σ_b, Newtonian relaxation of T_p to T_eq(φ, p)) and the
// energy integral of Compute_Corrections_Init (SD:5-28) — transposed, with the
// cotangents of ps_c (through T_eq), ps_p (through the energy integral and
// the mean ps_p) and of k_f, k_a, k_s, ΔT_y summed per column.  Same tiles as
// hs_column_adjoint_kernel; independent of it (its own inputs and outputs).
template <int KPT>
__global__ __launch_bounds__(1024) void hs_forcing_adjoint_kernel(HsColAdj s, VC v, int nlat,
                                                                  int nlon, const double* sint,
                                                                  const double* wts,
                                                                  HsAdjScalars c) {
  extern __shared__ double red[];  // [4][nq][64]
  const size_t P = (size_t)nlat * nlon;
  const int nd = v.nd;
  const TileIdx t = tile_idx<KPT>(P);
  const bool ok = t.col;
  const size_t col = ok ? t.c : 0;
  const int j = (int)(col / nlon);
  const double kap = c.kappa, cp = c.cp_air, dt = c.dt;
  const double ep_bar = c.scal_bar[1], mpsp_bar = c.scal_bar[2];
  const double aj = wts[j] / (2.0 * nlon);
  const double day = 86400.0, t_zero = 315.0, t_strat = 200.0, p_ref = 1.0e5;
  const double kf = c.hp.k_f / day, ka = c.hp.k_a / day, ks = c.hp.k_s / day;
  const double sj = sint[j], s2 = sj * sj, c2 = 1.0 - s2, c4 = c2 * c2;
  const double psc = ok ? s.ps_c[col] : 1.0, psp = ok ? s.ps_p[col] : 0.0;
  double pscb = 0.0, pspb = 0.0, kfb = 0.0, kab = 0.0, ksb = 0.0, dtyb = 0.0;
#pragma unroll
  for (int i = 0; i < KPT; ++i) {
    const int k = t.k0 + i;
    if (!(ok && k < nd)) continue;
    const size_t o = (size_t)k * P + col;
    const LevelConst L = level_const(v, k);
    const double db = v.dbk[k];
    const double a = s.du_b[o], b = s.dv_b[o], cc = s.dt_b[o];
    const double sg = L.sig;
    const double kv = kf / (1.0 - c.hp.sigma_b) * fmax(0.0, sg - c.hp.sigma_b);
    const double kt = ka + (ks - ka) / (1.0 - c.hp.sigma_b) * fmax(0.0, sg - c.hp.sigma_b) * c4;
    const double Up = s.u_p[o], Vp = s.v_p[o], Tp = s.t_p[o];
    const double pn = sg * psc / p_ref;
    const double A0 = t_zero - c.hp.delta_T_y * s2 - c.hp.delta_theta_z * c2 * log(pn);
    const double pk = pow(pn, kap);
    const double te = A0 * pk;
    const double teq = fmax(t_strat, te);
    const double g2 = kv * (1.0 - 0.5 * kv * dt);
    const double DT = g2 * (Up * Up + Vp * Vp) / cp - kt * (Tp - teq);
    const double uu = Up + dt * (-kv * Up), vq = Vp + dt * (-kv * Vp);
    const double eb = ep_bar / c.grav * aj * psp * db;
    const double ap = a + eb * dt * uu, bp = b + eb * dt * vq, cpb = cc + cp * dt * eb;
    s.up_b[o] = eb * uu - kv * ap + 2.0 * g2 * Up * cpb / cp;
    s.vp_b[o] = eb * vq - kv * bp + 2.0 * g2 * Vp * cpb / cp;
    s.tp_b[o] = cp * eb - kt * cpb;
    if (te > t_strat) {
      // ∂T_eq/∂ps_c = pn^(κ-1) (κ A0 - Δθ_z cos²φ) σ / p_ref
      const double dte = (pk / pn) * (kap * A0 - c.hp.delta_theta_z * c2) * sg / p_ref;
      pscb = pscb + kt * cpb * dte;
      dtyb = dtyb - cpb * kt * s2 * pk;  // ∂T_eq/∂ΔT_y = -sin²φ pn^κ
    }
    // k_v = k_f/day max(0, σ-σ_b)/(1-σ_b);  k_T = k_a/day + (k_s-k_a)/day max(0, σ-σ_b)/(1-σ_b) cos⁴φ
    const double sp = fmax(0.0, sg - c.hp.sigma_b) / (1.0 - c.hp.sigma_b);
    const double kvb = -(ap * Up) - bp * Vp + cpb * (1.0 - kv * dt) * (Up * Up + Vp * Vp) / cp;
    const double ktb = -(cpb * (Tp - teq));
    kfb = kfb + kvb * sp / day;
    kab = kab + ktb * (1.0 - sp * c4) / day;
    ksb = ksb + ktb * (sp * c4) / day;
    const double ep = 0.5 * (uu * uu + vq * vq) + cp * (Tp + dt * DT);
    pspb = pspb + ep_bar / c.grav * aj * db * ep;
  }
  double c1a[2], c1b[2], c2a[4], c2b[4];
  {
    const double pa[2] = {pscb, pspb};
    level_sums_n<2>(red, t, pa, c1a, c1b);
  }
  {
    const double pa[4] = {kfb, kab, ksb, dtyb};
    level_sums_n<4>(red, t, pa, c2a, c2b);
  }
  if (ok && t.q == 0) {
    s.psc_b[col] = pscb + c1b[0];
    s.psp_b[col] = (pspb + c1b[1]) + mpsp_bar * aj;
    if (s.param_b) {
      s.param_b[col] = s.param_b[col] + (kfb + c2b[0]);
      s.param_b[P + col] = s.param_b[P + col] + (kab + c2b[1]);
      s.param_b[2 * P + col] = s.param_b[2 * P + col] + (ksb + c2b[2]);
      s.param_b[3 * P + col] = s.param_b[3 * P + col] + (dtyb + c2b[3]);
    }
  }
}

// Transpose of the corrections (SD:30-68) with respect to the new-state grid
// fields u_n, v_n, T_n (before its energy correction) and ps_n (before its mass
// correction), given the cotangents E_n_bar of the energy integral and
// mps_n_bar of the mean surface pressure.  Inputs are the corrected fields
// (T_n' = T_n + corr, ps_n' = f ps_n).  Output ln ps cotangent = ps_n ps_n_bar
// (ps_n = exp(ln ps grid)).
__global__ void hs_corrections_adjoint_kernel(VC v, int nlat, int nlon, const double* wts,
                                              const double* u_n, const double* v_n,
                                              const double* t_nc, const double* ps_nc,
                                              const double* scal, const double* scal_bar,
                                              double grav, double cp, double* un_b, double* vn_b,
                                              double* tn_b, double* lnps_b) {
  const size_t P = (size_t)nlat * nlon;
  const int nd = v.nd;
  const double corr = scal[6], f = scal[3], en_bar = scal_bar[0], mpsn_bar = scal_bar[3];
  HS_GRID_STRIDE(col, P) {
    const int j = (int)(col / nlon);
    const double aj = wts[j] / (2.0 * nlon);
    const double psc = ps_nc[col];
    const double psn = psc / f;
    double acc = 0.0;
    for (int k = 0; k < nd; ++k) {
      size_t o = (size_t)k * P + col;
      const double u = u_n[o], vv = v_n[o], t = t_nc[o] - corr;
      const double e = 0.5 * (u * u + vv * vv) + cp * t;
      acc = acc + v.dbk[k] * e;
      const double eb = en_bar / grav * aj * psc * v.dbk[k];
      un_b[o] = eb * u;
      vn_b[o] = eb * vv;
      tn_b[o] = eb * cp;
    }
    const double psb = en_bar * f * aj * acc / grav + mpsn_bar * aj;
    lnps_b[col] = psn * psb;
  }
}

// Scalar algebra of the corrections' transpose (SD:30-68), one thread:
//   corr = g (E_p - E_n) / (c_p mps_p),  f = mps_p / mps_n,
//   T_c'(k, 0, 0) = T_n(k, 0, 0) + corr,  ln ps_c'(0, 0, 0) = ln ps_n(0, 0, 0) + log f,
//   E_n depends on f through ps_n' = f ps_n  (∂E_n/∂f = E_n / f).
// scal: the forward's slots (0 mps_p, 1 E_p, 2 mps_n, 3 f, 5 E_n, 6 corr);
// out: 0 E_n_bar, 1 E_p_bar, 2 mps_p_bar, 3 mps_n_bar.
__global__ void hs_correction_scalars_adjoint_kernel(const double* scal, const double2* t_c_bar,
                                                     const double2* lnps_c_bar, int nd,
                                                     size_t spec_per_level, double grav, double cp,
                                                     double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double mps_p = scal[0], mps_n = scal[2], f = scal[3], e_n = scal[5], corr = scal[6];
  double corr_b = 0.0;
  for (int k = 0; k < nd; ++k) corr_b = corr_b + t_c_bar[(size_t)k * spec_per_level].x;
  double f_b = lnps_c_bar[0].x / f;
  const double en_b = -corr_b * grav / (cp * mps_p);
  const double ep_b = corr_b * grav / (cp * mps_p);
  double mpsp_b = -corr_b * corr / mps_p;
  f_b = f_b + en_b * e_n / f;
  mpsp_b = mpsp_b + f_b / mps_n;
  out[0] = en_b;
  out[1] = ep_b;
  out[2] = mpsp_b;
  out[3] = -f_b * f / mps_n;
}

// out[i][n][m] = Σ_o A_n[o][i] in[o][n][m] for n >= m (zero below the
// diagonal), A_n the real (2nd+1) x (7nd+3) matrix of the linear map
// (δdiv, δln ps, δT, energy, div_c, div_p, ln ps_c, ln ps_p, T_c, T_p) ->
// (δdiv, δln ps, δT) of ∇²(energy) + Implicit_Correction!.
// Workgroup (n, output group, m chunk) of four waves: lane = m, each wave
// owns IPW consecutive outputs i, so the matrix entries it reads are the same
// for every lane (scalar loads) and each input cotangent is loaded once per
// wave for IPW products.  Sums run over o in increasing order.
constexpr int IMPL_IPW = 8;

__global__ __launch_bounds__(256) void implicit_adjoint_kernel(const double* A, int nd, int nm,
                                                               const double2* ddiv_b,
                                                               const double2* dlnps_b,
                                                               const double2* dt_b, double2* out,
                                                               size_t spec_per_level) {
  const int nin = 7 * nd + 3, nout = 2 * nd + 1;
  const int n = blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int lane = threadIdx.x & 63;
  const int i0 = (blockIdx.y * 4 + wave) * IMPL_IPW;
  if (i0 >= nin) return;  // the whole wave; no barriers follow
  const int m = blockIdx.z * 64 + lane;
  const bool in_row = m < nm;
  const bool valid = in_row && m <= n;
  const size_t idx = (size_t)n * nm + (in_row ? m : 0);
  const double* An = A + (size_t)n * nout * nin;
  double2 acc[IMPL_IPW];
#pragma unroll
  for (int r = 0; r < IMPL_IPW; ++r) acc[r] = make_double2(0.0, 0.0);
#pragma unroll 4
  for (int o = 0; o < nout; ++o) {
    double2 x = make_double2(0.0, 0.0);
    if (valid)
      x = o < nd ? ddiv_b[(size_t)o * spec_per_level + idx]
                 : (o == nd ? dlnps_b[idx] : dt_b[(size_t)(o - nd - 1) * spec_per_level + idx]);
    const double* Ao = An + (size_t)o * nin + i0;
#pragma unroll
    for (int r = 0; r < IMPL_IPW; ++r) {
      if (i0 + r < nin) {
        const double w = Ao[r];
        acc[r] = make_double2(acc[r].x + w * x.x, acc[r].y + w * x.y);
      }
    }
  }
  if (!in_row) return;
#pragma unroll
  for (int r = 0; r < IMPL_IPW; ++r)
    if (i0 + r < nin) out[(size_t)(i0 + r) * spec_per_level + idx] = acc[r];
}

// Transpose of Compute_Gradient_Cos! (SSM:677-687):
//   gλ[n] = dlam[n] (i h[n]),  gθ[n] = dthp[n] h[n+1] (n <= ns-1) + dthm[n] h[n-1] (n >= 1)
// h_bar (+)= dlam (-i gλ_bar) + dthp[k-1] gθ_bar[k-1] + dthm[k+1] gθ_bar[k+1].
__global__ void gradient_cos_adjoint_kernel(const double2* glb, const double2* gtb, double2* hb,
                                            const double* dlam, const double* dthm,
                                            const double* dthp, int nm, int ns, size_t n_all,
                                            int accumulate) {
  HS_GRID_STRIDE(i, n_all) {
    const int m = (int)(i % nm);
    const int k = (int)((i / nm) % (ns + 1));
    const int c = m * (ns + 1) + k;
    const double dl = dlam[c];
    const double2 l0 = glb[i];
    double2 z = make_double2(dl * l0.y, -(dl * l0.x));
    if (k + 1 <= ns) {
      const double w = dthm[c + 1];
      const double2 g = gtb[i + nm];
      z = make_double2(z.x + w * g.x, z.y + w * g.y);
    }
    if (k >= 1) {
      const double w = dthp[c - 1];
      const double2 g = gtb[i - nm];
      z = make_double2(z.x + w * g.x, z.y + w * g.y);
    }
    if (accumulate) {
      const double2 h = hb[i];
      z = make_double2(h.x + z.x, h.y + z.y);
    }
    hb[i] = z;
  }
}

static int check_pure_sigma(const jgcm_vert_coord* v) {
  JGCM_REQUIRE(v->zero_top == 1, "the adjoint needs a zero-top coordinate");
  for (double a : v->ak) JGCM_REQUIRE(a == 0.0, "the adjoint needs a pure sigma coordinate (ak = 0)");
  JGCM_REQUIRE(v->nd >= 1 && v->nd <= MAXND, "nd must be in 1..%d", MAXND);
  return JGCM_OK;
}

}  // namespace jgcm

using namespace jgcm;

extern "C" {

int jgcm_hs_column_adjoint(jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, double dt,
                           const double* scal_bar, const double* const* state,
                           const double* const* cot_in, double* const* cot_out,
                           double* param_bar) {
  JGCM_REQUIRE(si && hp && scal_bar && state && cot_in && cot_out, "NULL argument");
  if (int rc = check_pure_sigma(si->vc)) return rc;
  jgcm_mesh* M = si->vc->mesh;
  for (int i = 0; i < 14; ++i) JGCM_REQUIRE(state[i], "NULL state array %d", i);
  for (int i = 0; i < 5; ++i) JGCM_REQUIRE(cot_in[i], "NULL input cotangent %d", i);
  for (int i = 0; i < 14; ++i) JGCM_REQUIRE(cot_out[i], "NULL output cotangent %d", i);
  HsColAdj s{state[0], state[1], state[2], state[3], state[4], state[5], state[6], state[7],
             state[8], state[9], state[10], state[11], state[12], state[13],
             cot_in[0], cot_in[1], cot_in[2], cot_in[3], cot_in[4],
             cot_out[0], cot_out[1], cot_out[2], cot_out[3], cot_out[4], cot_out[5], cot_out[6],
             cot_out[7], cot_out[8], cot_out[9], cot_out[10], cot_out[11], cot_out[12],
             cot_out[13], param_bar};
  HsAdjScalars c{si->atmo.rdgas, si->atmo.cp_air, si->atmo.kappa, si->atmo.grav,
                 2.0 * si->atmo.omega, dt, scal_bar, *hp};
  const int nd = si->vc->nd;
  const size_t P = M->grid_elems();
  const dim3 grid((unsigned)((P + TILE - 1) / TILE));
  // nq = ceil(nd / KPT) waves of 64 lanes per workgroup (nq <= 16)
  if (nd <= 32) {
    const int nq = (nd + 1) / 2;
    const size_t lds = (size_t)4 * nq * TILE * sizeof(double);
    hipLaunchKernelGGL(hs_forcing_adjoint_kernel<2>, grid, dim3(nq * TILE), lds, M->stream, s,
                       vc_of(si->vc), M->nlat, M->nlon, M->d_sin, M->d_wts, c);
    hipLaunchKernelGGL(hs_column_adjoint_kernel<2>, grid, dim3(nq * TILE), lds, M->stream, s,
                       vc_of(si->vc), M->nlat, M->nlon, M->d_sin, M->d_wts, c);
  } else {
    const int nq = (nd + 3) / 4;
    const size_t lds = (size_t)4 * nq * TILE * sizeof(double);
    hipLaunchKernelGGL(hs_forcing_adjoint_kernel<4>, grid, dim3(nq * TILE), lds, M->stream, s,
                       vc_of(si->vc), M->nlat, M->nlon, M->d_sin, M->d_wts, c);
    hipLaunchKernelGGL(hs_column_adjoint_kernel<4>, grid, dim3(nq * TILE), lds, M->stream, s,
                       vc_of(si->vc), M->nlat, M->nlon, M->d_sin, M->d_wts, c);
  }
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_hs_corrections_adjoint(jgcm_semi_implicit_t* si, const double* scal,
                                const double* spe_t_c_bar, const double* spe_lnps_c_bar,
                                const double* grid_u_n, const double* grid_v_n,
                                const double* grid_t_n, const double* grid_ps_n, double* scal_bar,
                                double* u_n_bar, double* v_n_bar, double* t_n_bar,
                                double* lnps_n_bar) {
  JGCM_REQUIRE(si && scal && spe_t_c_bar && spe_lnps_c_bar && grid_u_n && grid_v_n && grid_t_n &&
                   grid_ps_n && scal_bar && u_n_bar && v_n_bar && t_n_bar && lnps_n_bar,
               "NULL argument");
  if (int rc = check_pure_sigma(si->vc)) return rc;
  jgcm_mesh* M = si->vc->mesh;
  hipLaunchKernelGGL(hs_correction_scalars_adjoint_kernel, dim3(1), dim3(64), 0, M->stream, scal,
                     (const double2*)spe_t_c_bar, (const double2*)spe_lnps_c_bar, si->nd,
                     M->spec_elems(), si->atmo.grav, si->atmo.cp_air, scal_bar);
  size_t P = M->grid_elems();
  hipLaunchKernelGGL(hs_corrections_adjoint_kernel, dim3(nblk_h(P)), dim3(256), 0, M->stream,
                     vc_of(si->vc), M->nlat, M->nlon, M->d_wts, grid_u_n, grid_v_n, grid_t_n,
                     grid_ps_n, scal, scal_bar, si->atmo.grav, si->atmo.cp_air, u_n_bar, v_n_bar,
                     t_n_bar, lnps_n_bar);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_implicit_adjoint(jgcm_mesh_t* mesh, int nd, const double* matrix, const double* ddiv_bar,
                          const double* dlnps_bar, const double* dt_bar, double* out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nd >= 1 && nd <= MAXND, "nd must be in 1..%d, got %d", MAXND, nd);
  JGCM_REQUIRE(matrix && ddiv_bar && dlnps_bar && dt_bar && out, "NULL argument");
  size_t S = mesh->spec_elems();
  const int nin = 7 * nd + 3;
  const dim3 grid((unsigned)(mesh->ns + 1), (unsigned)((nin + 4 * IMPL_IPW - 1) / (4 * IMPL_IPW)),
                  (unsigned)((mesh->nm + 63) / 64));
  hipLaunchKernelGGL(implicit_adjoint_kernel, grid, dim3(256), 0, mesh->stream, matrix, nd,
                     mesh->nm, (const double2*)ddiv_bar, (const double2*)dlnps_bar,
                     (const double2*)dt_bar, (double2*)out, S);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_gradient_cos_adjoint(jgcm_mesh_t* mesh, const double* grad_lambda_bar,
                              const double* grad_theta_bar, double* h_bar, int accumulate, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(grad_lambda_bar && grad_theta_bar && h_bar, "NULL argument");
  size_t n = (size_t)nlev * mesh->spec_elems();
  hipLaunchKernelGGL(gradient_cos_adjoint_kernel, dim3(nblk_h(n)), dim3(256), 0, mesh->stream,
                     (const double2*)grad_lambda_bar, (const double2*)grad_theta_bar,
                     (double2*)h_bar, mesh->d_dlam, mesh->d_dthm, mesh->d_dthp, mesh->nm, mesh->ns,
                     n, accumulate ? 1 : 0);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

}  // extern "C"
