// Adjoints (transposes with respect to the real inner products
//   grid:     <x, y> = sum x y
//   spectral: <a, b> = sum Re(conj(a) b)  over the whole [nlev][ns+1][nm] array)
// of the operators of the barotropic vorticity step (Barotropic_Dynamics.jl:2-46):
// Vor_Div_From_Grid_UV!, UV_Grid_From_Vor_Div! (through the u cos θ / v cos θ
// stencil), the absolute-vorticity flux and the damped, filtered leapfrog with
// Time_Advance!.  Composed, they give the reverse-mode step
// jgcm_barotropic_step_adjoint.
//
// The elementwise kernels read every input element before writing the
// outputs of the same index, so their outputs may alias their inputs (the
// step adjoint runs them in place); no pointer here is __restrict__.
#pragma clang fp contract(off)

#include "jgcm_internal.hpp"

namespace jgcm {

int ifft_unfold(const jgcm_mesh* M, const double2* eo, double* grid, const double* divcos, int B);
int ifft_unfold_flag(const jgcm_mesh* M, const double2* eo, double* grid, int flag, int B);
int legendre_alpha_transpose(const jgcm_mesh* M, const double2* h, double2* eo, int B, int table,
                             int op, int swap, int accumulate);

static inline int nblk(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// row tables of the α transposes (coef_alp_a, coef_alp_b) and the staging
// operators applied to the cotangent on the way in
enum { TAB_ALPA = 1, TAB_ALPB = 2 };
enum { OP_NONE = 0, OP_NEG_I = 5, OP_NEG = 6 };

__global__ void scale_by_cos_kernel(const double* in, double* out, const double* cost, int nlat,
                                    int nlon, size_t n, int divide) {
  GRID_STRIDE(i, n) {
    double c = cost[(i / nlon) % nlat];
    out[i] = divide ? in[i] / c : in[i] * c;
  }
}

// Transpose of the u cos θ / v cos θ stencil (SSM:568-576, coefficients [m][n], real):
//   ucos[n] = uvc[n] (i div[n]) - uvm[n] vor[n-1] - uvp[n] vor[n+1]
//   vcos[n] = uvc[n] (i vor[n]) + uvm[n] div[n-1] + uvp[n] div[n+1]
// so for row k
//   vor_bar[k] = uvc[k] (-i vcos_bar[k]) - uvm[k+1] ucos_bar[k+1] - uvp[k-1] ucos_bar[k-1]
//   div_bar[k] = uvc[k] (-i ucos_bar[k]) + uvm[k+1] vcos_bar[k+1] + uvp[k-1] vcos_bar[k-1]
// with the forward's guards (the uvm term exists for n >= 1, the uvp term for
// n <= ns-1).  add_a / add_b (may be NULL, may alias vor_b) are added to vor_bar.
__global__ void ucos_vcos_adjoint_kernel(const double2* ub, const double2* vb, double2* vor_b,
                                         double2* div_b, const double2* add_a, const double2* add_b,
                                         const double* uvc, const double* uvm, const double* uvp,
                                         int nm, int ns, size_t n_all) {
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int k = (int)((i / nm) % (ns + 1));
    int c = m * (ns + 1) + k;
    double cc = uvc[c];
    double2 u0 = ub[i], v0 = vb[i];
    double2 zv = make_double2(cc * v0.y, -(cc * v0.x));  // uvc (-i vcos_bar)
    double2 zd = make_double2(cc * u0.y, -(cc * u0.x));  // uvc (-i ucos_bar)
    if (k + 1 <= ns) {
      double a = uvm[c + 1];
      double2 up = ub[i + nm], vp = vb[i + nm];
      zv = make_double2(zv.x - a * up.x, zv.y - a * up.y);
      zd = make_double2(zd.x + a * vp.x, zd.y + a * vp.y);
    }
    if (k >= 1) {
      double a = uvp[c - 1];
      double2 um = ub[i - nm], vm = vb[i - nm];
      zv = make_double2(zv.x - a * um.x, zv.y - a * um.y);
      zd = make_double2(zd.x + a * vm.x, zd.y + a * vm.y);
    }
    if (add_a) {
      double2 t = add_a[i];
      zv = make_double2(t.x + zv.x, t.y + zv.y);
    }
    if (add_b) {
      double2 t = add_b[i];
      zv = make_double2(zv.x + t.x, zv.y + t.y);
    }
    vor_b[i] = zv;
    if (div_b) div_b[i] = zd;
  }
}

// Transpose of grid_δu = (ζ + f) v, grid_δv = -(ζ + f) u (Barotropic_Dynamics.jl:26-29)
// at the state (ζ, u, v), f = 2Ω sin θ_j:
//   ζ_bar = v δu_bar - u δv_bar,  u_bar = -(ζ + f) δv_bar,  v_bar = (ζ + f) δu_bar.
__global__ void vorticity_flux_adjoint_kernel(const double* vor, const double* u, const double* v,
                                              const double* sint, double two_omega,
                                              const double* dub, const double* dvb, double* vor_b,
                                              double* u_b, double* v_b, int nlat, int nlon,
                                              size_t n) {
  GRID_STRIDE(i, n) {
    double av = vor[i] + two_omega * sint[(i / nlon) % nlat];
    double a = dub[i], b = dvb[i];
    double z = v[i] * a - u[i] * b;
    double ur = -(av * b), vr = av * a;
    vor_b[i] = z;
    u_b[i] = ur;
    v_b[i] = vr;
  }
}

// Transpose of Compute_Spectral_Damping! + Filtered_Leapfrog! + Time_Advance!
// (Time_Integrator.jl:58-101, Dyn_Data.jl:326-364) for one field, from the
// cotangents of the advanced state (p', c') to those of (p, c) and of the
// undamped tendency G.  With d = damping, r = Robert coefficient, the step is
//   later steps:  δ = (G - p d)/(1 + 2Δt d),  c' = p + 2Δt δ,
//                 p' = (1-2r) c + 2r p + 2rΔt δ
//   first step:   δ = (G - c d)/(1 + Δt d),   c' = c + Δt δ,   p' = c + rΔt δ
__device__ __forceinline__ void leapfrog_adjoint_elem(size_t i, const double2* pb_next,
                                                      const double2* cb_next, double2* pb,
                                                      double2* cb, double2* gb,
                                                      const double* damping, int nm, int ns,
                                                      int init_step, double dt, double r) {
  {
    int m = (int)(i % nm);
    int k = (int)((i / nm) % (ns + 1));
    double d = damping ? damping[m * (ns + 1) + k] : 0.0;
    double2 P = pb_next[i], C = cb_next[i];
    if (init_step) {
      double den = 1.0 + dt * d;
      double2 db = make_double2((r * dt) * P.x + dt * C.x, (r * dt) * P.y + dt * C.y);
      double s = d / den;
      gb[i] = make_double2(db.x / den, db.y / den);
      cb[i] = make_double2(P.x + C.x - s * db.x, P.y + C.y - s * db.y);
      pb[i] = make_double2(0.0, 0.0);
    } else {
      double den = 1.0 + (2.0 * dt) * d;
      double2 db = make_double2((2.0 * r * dt) * P.x + (2.0 * dt) * C.x,
                                (2.0 * r * dt) * P.y + (2.0 * dt) * C.y);
      double s = d / den;
      gb[i] = make_double2(db.x / den, db.y / den);
      double2 np = make_double2((2.0 * r) * P.x + C.x - s * db.x, (2.0 * r) * P.y + C.y - s * db.y);
      cb[i] = make_double2((1.0 - 2.0 * r) * P.x, (1.0 - 2.0 * r) * P.y);
      pb[i] = np;
    }
  }
}

__global__ void damped_leapfrog_adjoint_kernel(const double2* pb_next, const double2* cb_next,
                                               double2* pb, double2* cb, double2* gb,
                                               const double* damping, int nm, int ns, size_t n_all,
                                               int init_step, double dt, double r) {
  GRID_STRIDE(i, n_all) {
    leapfrog_adjoint_elem(i, pb_next, cb_next, pb, cb, gb, damping, nm, ns, init_step, dt, r);
  }
}

// the same for up to MAXLFA fields of one integrator (blockIdx.y = field)
constexpr int MAXLFA = 8;
struct LeapAdjSet {
  const double2* pbn[MAXLFA];
  const double2* cbn[MAXLFA];
  double2* pb[MAXLFA];
  double2* cb[MAXLFA];
  double2* gb[MAXLFA];
  const double* damping[MAXLFA];
  size_t n[MAXLFA];
};

__global__ void damped_leapfrog_adjoint_batch_kernel(LeapAdjSet s, int nm, int ns, int init_step,
                                                     double dt, double r) {
  const int f = blockIdx.y;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < s.n[f];
       i += (size_t)gridDim.x * blockDim.x)
    leapfrog_adjoint_elem(i, s.pbn[f], s.cbn[f], s.pb[f], s.cb[f], s.gb[f], s.damping[f], nm, ns,
                          init_step, dt, r);
}

// (vor_bar, div_bar) -> (u_bar, v_bar): transposed α contractions into the E/O
// blocks, transposed FFT stage, then × cos θ (the forward multiplies by cos θ
// on load).  Uses the mesh's E/O scratch.
static int vor_div_adjoint(jgcm_mesh* M, const double2* vb, const double2* db, double* u_bar,
                           double* v_bar, int nlev) {
  double2* eu = reinterpret_cast<double2*>(M->d_four_a);
  double2* ev = reinterpret_cast<double2*>(M->d_four_b);
  // vor = alp_a (i V^p) - alp_b U^(1-p)    div = alp_a (i U^p) + alp_b V^(1-p)   (SSM:387-392)
  if (int rc = legendre_alpha_transpose(M, vb, ev, nlev, TAB_ALPA, OP_NEG_I, 0, 0)) return rc;
  if (int rc = legendre_alpha_transpose(M, vb, eu, nlev, TAB_ALPB, OP_NEG, 1, 0)) return rc;
  if (db) {
    if (int rc = legendre_alpha_transpose(M, db, eu, nlev, TAB_ALPA, OP_NEG_I, 0, 1)) return rc;
    if (int rc = legendre_alpha_transpose(M, db, ev, nlev, TAB_ALPB, OP_NONE, 1, 1)) return rc;
  }
  // the FFT stage applies the transposed Hermitian weights (flag 2) and the
  // final × cos θ (flags 1 | 4)
  if (int rc = ifft_unfold_flag(M, eu, u_bar, 1 | 2 | 4, nlev)) return rc;
  return ifft_unfold_flag(M, ev, v_bar, 1 | 2 | 4, nlev);
}

static void ucos_vcos_adjoint(jgcm_mesh* M, const double* su, const double* sv, double* vor_bar,
                              double* div_bar, const double* add_a, const double* add_b,
                              int nlev) {
  size_t n = (size_t)nlev * M->spec_elems();
  hipLaunchKernelGGL(ucos_vcos_adjoint_kernel, dim3(nblk(n)), dim3(256), 0, M->stream,
                     (const double2*)su, (const double2*)sv, (double2*)vor_bar, (double2*)div_bar,
                     (const double2*)add_a, (const double2*)add_b, M->d_uvc, M->d_uvm, M->d_uvp,
                     M->nm, M->ns, n);
}

// ---- shallow-water step (Shallow_Water_Dynamics.jl:7-21, :88-179) ----------

// Transpose of the semi-implicit correction (Shallow_Water_Dynamics.jl:36-51)
//   D1 = Gd + e (h_c - h_p),  H1 = Gh + h0 (div_c - div_p),
//   Dn = (D1 - μ e H1) / (1 - μ² e h0),  Hn = H1 - μ h0 Dn
// and of the energy term Gd -= ∇²E, E = F(KE) + h_c (:137-143), e = -n(n+1)/a².
// In: (gd, gh) = cotangents of (Dn, Hn).  Out: (gd, gh) = cotangents of the
// explicit tendencies (Gd, Gh); ebar = cotangent of E; the state cotangents
// gain their linear terms.  (The h_c terms e D1 and -e D1 cancel: the scheme
// replaces the explicit -∇²h_c by the implicit average.)
__global__ void sw_implicit_adjoint_kernel(double2* gd, double2* gh, double2* ebar, double2* cd,
                                           double2* pd, double2* ch, double2* ph, const double* eig,
                                           double h0, double mu, int nm, int ns, size_t n_all) {
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int k = (int)((i / nm) % (ns + 1));
    double e = eig[m * (ns + 1) + k];
    double2 Dn = gd[i], Hn = gh[i];
    double mh = mu * h0;
    double2 Dt = make_double2(Dn.x - mh * Hn.x, Dn.y - mh * Hn.y);
    double den = 1.0 - (mu * mu * e) * h0;
    double2 D1 = make_double2(Dt.x / den, Dt.y / den);
    double me = mu * e;
    double2 H1 = make_double2(Hn.x - me * D1.x, Hn.y - me * D1.y);
    gd[i] = D1;
    gh[i] = H1;
    double2 eb = make_double2(-(e * D1.x), -(e * D1.y));
    ebar[i] = eb;
    double2 c = ch[i], p = ph[i];
    ch[i] = make_double2(c.x + e * D1.x + eb.x, c.y + e * D1.y + eb.y);
    ph[i] = make_double2(p.x - e * D1.x, p.y - e * D1.y);
    double2 a = cd[i], b = pd[i];
    cd[i] = make_double2(a.x + h0 * H1.x, a.y + h0 * H1.y);
    pd[i] = make_double2(b.x - h0 * H1.x, b.y - h0 * H1.y);
  }
}

// Transpose of the grid part of the SW tendencies at the state (ζ, D, H, u, v):
//   δu = -κm u + (ζ+f) v,  δv = -κm v - (ζ+f) u,  δH = -κt (H - h_eq)   (:16-18, :126-130)
//   KE = (u² + v²)/2                                                     (:139)
//   δH -= u gλ + v gθ   (gλ, gθ the grid gradients of h_c, SSM:716)     (:147)
//   δH -= H D                                                            (:148)
// In: cotangents of δu, δv, δH and KE.  Out (any slot may alias an input
// slot): ζ_bar, D_bar, H_bar, and u_bar, v_bar, gλ_bar, gθ_bar divided by
// cos θ (the inverse transforms behind them divide by cos θ).
struct SwGridAdj {
  const double *vor, *div, *h, *u, *v, *gl, *gt;       // state
  const double *dub, *dvb, *dhb, *kb;                  // cotangents in
  double *zb, *db, *hb, *ubc, *vbc, *glbc, *gtbc;      // cotangents out
};

__global__ void sw_grid_adjoint_kernel(SwGridAdj s, const double* sint, const double* cost,
                                       double two_omega, double km, double kt, int nlat, int nlon,
                                       size_t n) {
  GRID_STRIDE(i, n) {
    int j = (int)((i / nlon) % nlat);
    double av = s.vor[i] + two_omega * sint[j];
    double c = cost[j];
    double U = s.u[i], V = s.v[i], H = s.h[i], D = s.div[i], GL = s.gl[i], GT = s.gt[i];
    double a = s.dub[i], b = s.dvb[i], q = s.dhb[i], kk = s.kb[i];
    double ub = -(km * a) - av * b - GL * q + U * kk;
    double vb = -(km * b) + av * a - GT * q + V * kk;
    double zb = V * a - U * b;
    double db = -(H * q);
    double hb = -(D * q) - kt * q;
    double glb = -(U * q), gtb = -(V * q);
    s.zb[i] = zb;
    s.db[i] = db;
    s.hb[i] = hb;
    s.ubc[i] = ub / c;
    s.vbc[i] = vb / c;
    s.glbc[i] = glb / c;
    s.gtbc[i] = gtb / c;
  }
}

// State cotangents gain the spectral transposes:
//   vor_bar += tz + stencil_vor^T(su, sv),  div_bar += td + stencil_div^T(su, sv)
//   h_bar   += th + grad^T(sgl, sgt)
// with the u cos / v cos stencil of ucos_vcos_adjoint_kernel and the gradient
// stencil of Compute_Gradient_Cos! (SSM:677-687):
//   gλ[n] = dlam[n] (i h[n]),  gθ[n] = dthp[n] h[n+1] (n <= ns-1) + dthm[n] h[n-1] (n >= 1)
struct SwSpecAdj {
  const double2 *tz, *td, *th, *su, *sv, *sgl, *sgt;
  double2 *vor_b, *div_b, *h_b;
};

__global__ void sw_spectral_adjoint_kernel(SwSpecAdj s, const double* uvc, const double* uvm,
                                           const double* uvp, const double* dlam, const double* dthm,
                                           const double* dthp, int nm, int ns, size_t n_all) {
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int k = (int)((i / nm) % (ns + 1));
    int c = m * (ns + 1) + k;
    double cc = uvc[c], dl = dlam[c];
    double2 u0 = s.su[i], v0 = s.sv[i], l0 = s.sgl[i];
    double2 zv = make_double2(cc * v0.y, -(cc * v0.x));
    double2 zd = make_double2(cc * u0.y, -(cc * u0.x));
    double2 zh = make_double2(dl * l0.y, -(dl * l0.x));
    if (k + 1 <= ns) {
      double a = uvm[c + 1], t = dthm[c + 1];
      double2 up = s.su[i + nm], vp = s.sv[i + nm], gp = s.sgt[i + nm];
      zv = make_double2(zv.x - a * up.x, zv.y - a * up.y);
      zd = make_double2(zd.x + a * vp.x, zd.y + a * vp.y);
      zh = make_double2(zh.x + t * gp.x, zh.y + t * gp.y);
    }
    if (k >= 1) {
      double a = uvp[c - 1], t = dthp[c - 1];
      double2 um = s.su[i - nm], vm = s.sv[i - nm], gm = s.sgt[i - nm];
      zv = make_double2(zv.x - a * um.x, zv.y - a * um.y);
      zd = make_double2(zd.x + a * vm.x, zd.y + a * vm.y);
      zh = make_double2(zh.x + t * gm.x, zh.y + t * gm.y);
    }
    double2 t0 = s.tz[i], t1 = s.td[i], t2 = s.th[i];
    double2 v = s.vor_b[i], d = s.div_b[i], h = s.h_b[i];
    s.vor_b[i] = make_double2(v.x + (t0.x + zv.x), v.y + (t0.y + zv.y));
    s.div_b[i] = make_double2(d.x + (t1.x + zd.x), d.y + (t1.y + zd.y));
    s.h_b[i] = make_double2(h.x + (t2.x + zh.x), h.y + (t2.y + zh.y));
  }
}

int inverse_batch(const jgcm_mesh* M, int nseg, const double* const* snm,
                  const double* const* snm_b, double* const* grid, const int* nlev,
                  const int* flags);

}  // namespace jgcm

using namespace jgcm;

extern "C" {

int jgcm_trans_spherical_to_grid_adjoint(jgcm_mesh_t* mesh, const double* pfield, double* snm,
                                         int nlev);

int jgcm_vor_div_from_grid_uv_adjoint(jgcm_mesh_t* mesh, const double* vor_bar, const double* div_bar,
                                      double* u_bar, double* v_bar, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(vor_bar && u_bar && v_bar, "NULL field pointer");
  if (int rc = ensure_scratch(mesh, nlev)) return rc;
  return vor_div_adjoint(mesh, reinterpret_cast<const double2*>(vor_bar),
                         reinterpret_cast<const double2*>(div_bar), u_bar, v_bar, nlev);
}

int jgcm_ucos_vcos_adjoint(jgcm_mesh_t* mesh, const double* ucos_bar, const double* vcos_bar,
                           double* vor_bar, double* div_bar, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(ucos_bar && vcos_bar && vor_bar, "NULL field pointer");
  ucos_vcos_adjoint(mesh, ucos_bar, vcos_bar, vor_bar, div_bar, nullptr, nullptr, nlev);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_uv_grid_from_vor_div_adjoint(jgcm_mesh_t* mesh, const double* u_bar, const double* v_bar,
                                      double* vor_bar, double* div_bar, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(u_bar && v_bar && vor_bar, "NULL field pointer");
  if (int rc = ensure_scratch(mesh, 2 * nlev)) return rc;  // before any scratch pointer is taken
  // ÷ cos θ, then the transposed inverse transform of u and v in one batch
  // (adjacent outputs: one zero fill)
  double* su = mesh->d_spec_a;
  double* sv = su + (size_t)nlev * mesh->spec_elems() * 2;
  const double* g[2] = {u_bar, v_bar};
  double* s[2] = {su, sv};
  const int lev[2] = {nlev, nlev}, fl[2] = {JGCM_BATCH_DIVCOS, JGCM_BATCH_DIVCOS};
  if (int rc = jgcm_trans_spherical_to_grid_adjoint_batch(mesh, 2, g, s, lev, fl, 0)) return rc;
  ucos_vcos_adjoint(mesh, su, sv, vor_bar, div_bar, nullptr, nullptr, nlev);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_vorticity_flux_adjoint(jgcm_mesh_t* mesh, const double* grid_vor, const double* grid_u,
                                const double* grid_v, double omega, const double* du_bar,
                                const double* dv_bar, double* vor_bar, double* u_bar, double* v_bar,
                                int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(grid_vor && grid_u && grid_v && du_bar && dv_bar && vor_bar && u_bar && v_bar,
               "NULL field pointer");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(vorticity_flux_adjoint_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream,
                     grid_vor, grid_u, grid_v, mesh->d_sin, 2.0 * omega, du_bar, dv_bar, vor_bar,
                     u_bar, v_bar, mesh->nlat, mesh->nlon, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_damped_leapfrog_adjoint(jgcm_leapfrog_t* lf, int init_step, int damp,
                                 const double* p_bar_next, const double* c_bar_next, double* p_bar,
                                 double* c_bar, double* g_bar, int nlev) {
  JGCM_REQUIRE(lf != nullptr && lf->mesh != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(p_bar_next && c_bar_next && p_bar && c_bar && g_bar, "NULL field pointer");
  jgcm_mesh* M = lf->mesh;
  size_t n = (size_t)nlev * M->spec_elems();
  hipLaunchKernelGGL(damped_leapfrog_adjoint_kernel, dim3(nblk(n)), dim3(256), 0, M->stream,
                     (const double2*)p_bar_next, (const double2*)c_bar_next, (double2*)p_bar,
                     (double2*)c_bar, (double2*)g_bar, damp ? lf->d_damping : nullptr, M->nm, M->ns, n,
                     init_step ? 1 : 0, (double)lf->dt, lf->robert_coef);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_damped_leapfrog_adjoint_batch(jgcm_leapfrog_t* lf, int init_step, int nfields,
                                       const int* damp, const double* const* p_bar_next,
                                       const double* const* c_bar_next, double* const* p_bar,
                                       double* const* c_bar, double* const* g_bar,
                                       const int* nlev) {
  JGCM_REQUIRE(lf != nullptr && lf->mesh != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nfields >= 1 && nfields <= MAXLFA, "nfields must be in 1..%d, got %d", MAXLFA,
               nfields);
  JGCM_REQUIRE(damp && p_bar_next && c_bar_next && p_bar && c_bar && g_bar && nlev,
               "NULL argument");
  jgcm_mesh* M = lf->mesh;
  LeapAdjSet s{};
  size_t maxn = 0;
  for (int f = 0; f < nfields; ++f) {
    JGCM_REQUIRE(p_bar_next[f] && c_bar_next[f] && p_bar[f] && c_bar[f] && g_bar[f],
                 "NULL field pointer (field %d)", f);
    JGCM_REQUIRE(nlev[f] >= 1, "nlev[%d] must be >= 1, got %d", f, nlev[f]);
    s.pbn[f] = (const double2*)p_bar_next[f];
    s.cbn[f] = (const double2*)c_bar_next[f];
    s.pb[f] = (double2*)p_bar[f];
    s.cb[f] = (double2*)c_bar[f];
    s.gb[f] = (double2*)g_bar[f];
    s.damping[f] = damp[f] ? lf->d_damping : nullptr;
    s.n[f] = (size_t)nlev[f] * M->spec_elems();
    if (s.n[f] > maxn) maxn = s.n[f];
  }
  hipLaunchKernelGGL(damped_leapfrog_adjoint_batch_kernel, dim3(nblk(maxn), nfields), dim3(256), 0,
                     M->stream, s, M->nm, M->ns, init_step ? 1 : 0, (double)lf->dt,
                     lf->robert_coef);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// One reverse step of Barotropic_Dynamics! (the first step of a run when
// init_step != 0), in place on (p_bar, c_bar):
//   (p_bar, c_bar, G_bar) <- leapfrog^T (p_bar, c_bar)
//   (δu_bar, δv_bar)      <- VorDiv^T (G_bar, 0)
//   (ζ_bar, u_bar, v_bar) <- flux^T at (grid_vor, grid_u, grid_v)
//   c_bar += T^-T ζ_bar + stencil^T(T^-T (u_bar / cos θ), T^-T (v_bar / cos θ))
// grid_vor, grid_u, grid_v are the grid fields of the step's current state
// (T^-1 vor_c and UV_Grid_From_Vor_Div!(vor_c, 0)).
int jgcm_barotropic_step_adjoint(jgcm_leapfrog_t* lf, int init_step, const double* grid_vor,
                                 const double* grid_u, const double* grid_v, double omega,
                                 double* p_bar, double* c_bar, int nlev) {
  JGCM_REQUIRE(lf != nullptr && lf->mesh != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(grid_vor && grid_u && grid_v && p_bar && c_bar, "NULL field pointer");
  jgcm_mesh* M = lf->mesh;
  if (int rc = check_mesh(M)) return rc;
  if (int rc = ensure_scratch(M, 2 * nlev)) return rc;  // before any scratch pointer is taken
  const size_t S = (size_t)nlev * M->spec_elems() * 2, G = (size_t)nlev * M->grid_elems();
  double* gbar = M->d_spec_a;
  double* tz = M->d_spec_a + S;
  double* su = M->d_spec_b;
  double* sv = M->d_spec_b + S;
  double* ga0 = M->d_grid_a;
  double* ga1 = M->d_grid_a + G;
  double* gb0 = M->d_grid_b;
  if (int rc = jgcm_damped_leapfrog_adjoint(lf, init_step, 1, p_bar, c_bar, p_bar, c_bar, gbar, nlev))
    return rc;
  if (int rc = vor_div_adjoint(M, reinterpret_cast<const double2*>(gbar), nullptr, ga0, ga1, nlev))
    return rc;
  hipLaunchKernelGGL(vorticity_flux_adjoint_kernel, dim3(nblk(G)), dim3(256), 0, M->stream,
                     grid_vor, grid_u, grid_v, M->d_sin, 2.0 * omega, ga0, ga1, gb0, ga0, ga1,
                     M->nlat, M->nlon, G);
  hipLaunchKernelGGL(scale_by_cos_kernel, dim3(nblk(G)), dim3(256), 0, M->stream, ga0, ga0,
                     M->d_cos, M->nlat, M->nlon, G, 1);
  hipLaunchKernelGGL(scale_by_cos_kernel, dim3(nblk(G)), dim3(256), 0, M->stream, ga1, ga1,
                     M->d_cos, M->nlat, M->nlon, G, 1);
  JGCM_LAUNCH_CHECK();
  if (int rc = jgcm_trans_spherical_to_grid_adjoint(M, gb0, tz, nlev)) return rc;
  if (int rc = jgcm_trans_spherical_to_grid_adjoint(M, ga0, su, nlev)) return rc;
  if (int rc = jgcm_trans_spherical_to_grid_adjoint(M, ga1, sv, nlev)) return rc;
  ucos_vcos_adjoint(M, su, sv, c_bar, nullptr, c_bar, tz, nlev);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_trans_grid_to_spherical_adjoint(jgcm_mesh_t* mesh, const double* snm, double* pfield,
                                         int nlev);

// One reverse step of Shallow_Water_Physics! + Shallow_Water_Dynamics!
// (Shallow_Water_Dynamics.jl:7-21, :88-179), in place on the six state
// cotangents.  grid_vor, grid_div, grid_h, grid_u, grid_v are the grid fields
// of the step's starting state and spe_h its spectral height (the field the
// height advection differentiates).  Scratch: four level-field slots in each
// of the mesh's spectral and grid buffers (S_a0..3, S_b0..3, G_a0..3, G_b0..3).
int jgcm_sw_step_adjoint(jgcm_leapfrog_t* lf, int init_step, const double* grid_vor,
                         const double* grid_div, const double* grid_h, const double* grid_u,
                         const double* grid_v, const double* spe_h, double omega, double kappa_m,
                         double kappa_t, double h_0, double* vor_p_bar, double* vor_c_bar,
                         double* div_p_bar, double* div_c_bar, double* h_p_bar, double* h_c_bar,
                         int nlev) {
  JGCM_REQUIRE(lf != nullptr && lf->mesh != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(grid_vor && grid_div && grid_h && grid_u && grid_v && spe_h, "NULL state pointer");
  JGCM_REQUIRE(vor_p_bar && vor_c_bar && div_p_bar && div_c_bar && h_p_bar && h_c_bar,
               "NULL cotangent pointer");
  jgcm_mesh* M = lf->mesh;
  if (int rc = check_mesh(M)) return rc;
  if (int rc = ensure_scratch(M, 4 * nlev)) return rc;  // before any scratch pointer is taken
  const size_t S = (size_t)nlev * M->spec_elems() * 2, G = (size_t)nlev * M->grid_elems();
  double* Sa[4] = {M->d_spec_a, M->d_spec_a + S, M->d_spec_a + 2 * S, M->d_spec_a + 3 * S};
  double* Sb[4] = {M->d_spec_b, M->d_spec_b + S, M->d_spec_b + 2 * S, M->d_spec_b + 3 * S};
  double* Ga[4] = {M->d_grid_a, M->d_grid_a + G, M->d_grid_a + 2 * G, M->d_grid_a + 3 * G};
  double* Gb[4] = {M->d_grid_b, M->d_grid_b + G, M->d_grid_b + 2 * G, M->d_grid_b + 3 * G};
  const size_t ns_all = (size_t)nlev * M->spec_elems();
  // damping + leapfrog + Time_Advance: tendency cotangents Gv, Gd, Gh -> S_a0..2
  if (int rc = jgcm_damped_leapfrog_adjoint(lf, init_step, 1, vor_p_bar, vor_c_bar, vor_p_bar,
                                            vor_c_bar, Sa[0], nlev)) return rc;
  if (int rc = jgcm_damped_leapfrog_adjoint(lf, init_step, 1, div_p_bar, div_c_bar, div_p_bar,
                                            div_c_bar, Sa[1], nlev)) return rc;
  if (int rc = jgcm_damped_leapfrog_adjoint(lf, init_step, 1, h_p_bar, h_c_bar, h_p_bar, h_c_bar,
                                            Sa[2], nlev)) return rc;
  // semi-implicit correction and the energy term; E_bar -> S_a3
  const double mu = lf->implicit_coef * (init_step ? (double)lf->dt : 2.0 * (double)lf->dt);
  hipLaunchKernelGGL(sw_implicit_adjoint_kernel, dim3(nblk(ns_all)), dim3(256), 0, M->stream,
                     (double2*)Sa[1], (double2*)Sa[2], (double2*)Sa[3], (double2*)div_c_bar,
                     (double2*)div_p_bar, (double2*)h_c_bar, (double2*)h_p_bar, M->d_eig, h_0, mu,
                     M->nm, M->ns, ns_all);
  JGCM_LAUNCH_CHECK();
  // δH_bar = F^T Gh_bar -> G_a0, KE_bar = F^T E_bar -> G_a1 (one call over S_a2..3)
  if (int rc = jgcm_trans_grid_to_spherical_adjoint(M, Sa[2], Ga[0], 2 * nlev)) return rc;
  // (δu_bar, δv_bar) = VorDiv^T(Gv_bar, Gd_bar) -> G_a2, G_a3
  if (int rc = vor_div_adjoint(M, reinterpret_cast<const double2*>(Sa[0]),
                               reinterpret_cast<const double2*>(Sa[1]), Ga[2], Ga[3], nlev))
    return rc;
  // grid gradients of h (the forward's Add_Horizontal_Advection!) -> G_b0, G_b1
  {
    const double* s[2] = {spe_h, spe_h};
    const double* s2[2] = {nullptr, nullptr};
    double* g[2] = {Gb[0], Gb[1]};
    int nl[2] = {nlev, nlev};
    int fl[2] = {JGCM_BATCH_DIVCOS | JGCM_INV_GRAD_LAMBDA, JGCM_BATCH_DIVCOS | JGCM_INV_GRAD_THETA};
    if (int rc = inverse_batch(M, 2, s, s2, g, nl, fl)) return rc;
  }
  SwGridAdj ga{grid_vor, grid_div, grid_h, grid_u, grid_v, Gb[0], Gb[1],
               Ga[2], Ga[3], Ga[0], Ga[1],
               Ga[0], Ga[1], Ga[2], Ga[3], Gb[0], Gb[1], Gb[2]};
  hipLaunchKernelGGL(sw_grid_adjoint_kernel, dim3(nblk(G)), dim3(256), 0, M->stream, ga, M->d_sin,
                     M->d_cos, 2.0 * omega, kappa_m, kappa_t, M->nlat, M->nlon, G);
  JGCM_LAUNCH_CHECK();
  // T^-T of (ζ, D, H, u/cos) -> S_a0..3 and of (v/cos, gλ/cos, gθ/cos) -> S_b0..2
  if (int rc = jgcm_trans_spherical_to_grid_adjoint(M, Ga[0], Sa[0], 4 * nlev)) return rc;
  if (int rc = jgcm_trans_spherical_to_grid_adjoint(M, Gb[0], Sb[0], 3 * nlev)) return rc;
  SwSpecAdj sa{(const double2*)Sa[0], (const double2*)Sa[1], (const double2*)Sa[2],
               (const double2*)Sa[3], (const double2*)Sb[0], (const double2*)Sb[1],
               (const double2*)Sb[2], (double2*)vor_c_bar, (double2*)div_c_bar, (double2*)h_c_bar};
  hipLaunchKernelGGL(sw_spectral_adjoint_kernel, dim3(nblk(ns_all)), dim3(256), 0, M->stream, sa,
                     M->d_uvc, M->d_uvm, M->d_uvp, M->d_dlam, M->d_dthm, M->d_dthp, M->nm, M->ns,
                     ns_all);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

}  // extern "C"
