// Time integration and grid-tendency kernels used by the barotropic and
// shallow-water steppers.
//
// Reference:
//   Time_Integrator.jl:20-101        Filtered_Leapfrog, damping, leapfrog + RA filter
//   Shallow_Water_Dynamics.jl:7-54   Shallow_Water_Physics!, SW Implicit_Correction!
//   Atmo_Data.jl:45,59-65            Coriolis parameter, Compute_Abs_Vor!
// Operation order inside each kernel follows the reference's broadcast
// expressions so the rounding matches term by term (FP contraction off).
#pragma clang fp contract(off)

#include <cmath>
#include <new>

#include "jgcm_internal.hpp"

namespace jgcm {

static inline int nblk(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// Compute_Spectral_Damping! (TI:76-80) of one element: δQ = (δQ - Q d) / (1 + Δt d)
// with Q = Q_c on the first step, (δQ - Q_p d) / (1 + 2Δt d) afterwards.
__device__ __forceinline__ double2 damp_reg(double2 q, double2 p, double2 c, double d,
                                            int init_step, double dt) {
  if (init_step) {
    double den = 1.0 + dt * d;
    return make_double2((q.x - c.x * d) / den, (q.y - c.y * d) / den);
  }
  double den = 1.0 + (2.0 * dt) * d;
  return make_double2((q.x - p.x * d) / den, (q.y - p.y * d) / den);
}

// Filtered_Leapfrog! (TI:93-100) of one element: the filtered current value
// and the new value.
__device__ __forceinline__ void leapfrog_reg(double2 q, double2 p, double2 c, int init_step,
                                             double dt, double r, double2& c_out, double2& n_out) {
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

// Compute_Spectral_Damping! (TI:76-80) fused with Filtered_Leapfrog! (TI:93-100)
// for element i.  `mode` bit 0: apply damping, bit 1: apply leapfrog.
__device__ __forceinline__ void damped_leapfrog_elem(double2* __restrict__ dQ,
                                                     const double2* __restrict__ Qp,
                                                     double2* __restrict__ Qc,
                                                     double2* __restrict__ Qn,
                                                     const double* __restrict__ damping, int nm,
                                                     int ns, size_t i, int init_step, double dt,
                                                     double r, int mode) {
  int m = (int)(i % nm);
  int n = (int)((i / nm) % (ns + 1));
  double2 q = dQ[i];
  if (mode & 1) {
    double d = damping[m * (ns + 1) + n];
    q = damp_reg(q, init_step ? make_double2(0.0, 0.0) : Qp[i], init_step ? Qc[i] : make_double2(0.0, 0.0),
                 d, init_step, dt);
    dQ[i] = q;
  }
  if (mode & 2) {
    double2 c = Qc[i];
    double2 p = init_step ? make_double2(0.0, 0.0) : Qp[i];
    double2 cn, nx;
    leapfrog_reg(q, p, c, init_step, dt, r, cn, nx);
    Qn[i] = nx;
    Qc[i] = cn;
  }
}

__global__ void damped_leapfrog_kernel(double2* __restrict__ dQ, const double2* __restrict__ Qp,
                                       double2* __restrict__ Qc, double2* __restrict__ Qn,
                                       const double* __restrict__ damping, int nm, int ns,
                                       size_t n_all, int init_step, double dt, double r, int mode) {
  GRID_STRIDE(i, n_all)
  damped_leapfrog_elem(dQ, Qp, Qc, Qn, damping, nm, ns, i, init_step, dt, r, mode);
}

// several prognostic fields in one launch: blockIdx.y = field
constexpr int MAXLF = 8;
struct LeapfrogSet {
  double2* dQ[MAXLF];
  const double2* Qp[MAXLF];
  double2* Qc[MAXLF];
  double2* Qn[MAXLF];
  size_t n[MAXLF];
  int mode[MAXLF];
};

__global__ void damped_leapfrog_batch_kernel(LeapfrogSet s, const double* __restrict__ damping,
                                             int nm, int ns, int init_step, double dt, double r) {
  const int f = blockIdx.y;
  GRID_STRIDE(i, s.n[f])
  damped_leapfrog_elem(s.dQ[f], s.Qp[f], s.Qc[f], s.Qn[f], damping, nm, ns, i, init_step, dt, r,
                       s.mode[f]);
}

// SW Implicit_Correction! (Shallow_Water_Dynamics.jl:36-51) of one element
// (D = δdiv, H = δh in and out; e = the Laplacian eigenvalue).
__device__ __forceinline__ void sw_implicit_reg(double2& D, double2& H, double2 hc, double2 hp,
                                                double2 dc, double2 dp, double e, double h0,
                                                double mu, double mu2) {
  D = make_double2(D.x + e * (hc.x - hp.x), D.y + e * (hc.y - hp.y));
  H = make_double2(H.x + h0 * (dc.x - dp.x), H.y + h0 * (dc.y - dp.y));
  double me = mu * e;
  double den = 1.0 - (mu2 * e) * h0;
  D = make_double2((D.x - me * H.x) / den, (D.y - me * H.y) / den);
  double mh = mu * h0;
  H = make_double2(H.x - mh * D.x, H.y - mh * D.y);
}

__global__ void sw_implicit_kernel(const double2* __restrict__ div_c, const double2* __restrict__ div_p,
                                   const double2* __restrict__ h_c, const double2* __restrict__ h_p,
                                   double h0, double2* __restrict__ ddiv, double2* __restrict__ dh,
                                   const double* __restrict__ eig, int nm, int ns, size_t n_all,
                                   double mu) {
  double mu2 = mu * mu;
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int n = (int)((i / nm) % (ns + 1));
    double e = eig[m * (ns + 1) + n];
    double2 D = ddiv[i], H = dh[i];
    sw_implicit_reg(D, H, h_c[i], h_p[i], div_c[i], div_p[i], e, h0, mu, mu2);
    ddiv[i] = D;
    dh[i] = H;
  }
}

// The spectral part of a shallow-water step after the forward transforms
// (Shallow_Water_Dynamics.jl:141-143 δdiv -= ∇²(KE + h), :150-157
// Implicit_Correction!, :159-165 Compute_Spectral_Damping! + Filtered_Leapfrog!
// of ζ, D, h) and the spectral part of Time_Advance! (Dyn_Data.jl:328-340:
// p = the filtered c, c = n) for one element, with the arithmetic of
// axpby_kernel, laplacian_kernel, sw_implicit_kernel and
// damped_leapfrog_elem.  energy receives ∇²(KE + h_c) as spe_energy does.
struct SWSpectral {
  const double2* kin;
  double2* energy;
  double2* dvor;
  double2* ddiv;
  double2* dh;
  double2* vor[3];  // p, c, n
  double2* div[3];
  double2* h[3];
};

// advance = 1: the spectral half of Time_Advance! (p = the filtered c, c = n =
// the new value); advance = 0: Filtered_Leapfrog!'s own writes (the filtered
// c in place, n = the new value), Time_Advance! left to the caller.
__global__ void sw_spectral_kernel(SWSpectral s, const double* __restrict__ eig,
                                   const double* __restrict__ damping, int nm, int ns, size_t n_all,
                                   double h0, double mu, int init_step, double dt, double r,
                                   int advance) {
  const double mu2 = mu * mu;
  GRID_STRIDE(i, n_all) {
    const int m = (int)(i % nm);
    const int n = (int)((i / nm) % (ns + 1));
    const int t = m * (ns + 1) + n;
    const double e = eig[t], d = damping[t];
    const double2 k = s.kin[i], hc = s.h[1][i], hp = s.h[0][i];
    const double2 dc = s.div[1][i], dp = s.div[0][i];
    const double2 vc = s.vor[1][i], vp = s.vor[0][i];
    double2 D = s.ddiv[i], H = s.dh[i], Q = s.dvor[i];
    double2 en = make_double2(1.0 * k.x + 1.0 * hc.x, 1.0 * k.y + 1.0 * hc.y);
    en = make_double2(e * en.x, e * en.y);
    s.energy[i] = en;
    D = make_double2(1.0 * D.x + -1.0 * en.x, 1.0 * D.y + -1.0 * en.y);
    sw_implicit_reg(D, H, hc, hp, dc, dp, e, h0, mu, mu2);
    double2 cn, nx;
    Q = damp_reg(Q, vp, vc, d, init_step, dt);
    s.dvor[i] = Q;
    leapfrog_reg(Q, vp, vc, init_step, dt, r, cn, nx);
    if (advance) {
      s.vor[0][i] = cn;
      s.vor[1][i] = nx;
    } else {
      s.vor[1][i] = cn;
    }
    s.vor[2][i] = nx;
    D = damp_reg(D, dp, dc, d, init_step, dt);
    s.ddiv[i] = D;
    leapfrog_reg(D, dp, dc, init_step, dt, r, cn, nx);
    if (advance) {
      s.div[0][i] = cn;
      s.div[1][i] = nx;
    } else {
      s.div[1][i] = cn;
    }
    s.div[2][i] = nx;
    H = damp_reg(H, hp, hc, d, init_step, dt);
    s.dh[i] = H;
    leapfrog_reg(H, hp, hc, init_step, dt, r, cn, nx);
    if (advance) {
      s.h[0][i] = cn;
      s.h[1][i] = nx;
    } else {
      s.h[1][i] = cn;
    }
    s.h[2][i] = nx;
  }
}

// absvor = vor + f_j (Compute_Abs_Vor!, Atmo_Data.jl:59-65) and the vorticity
// flux  δu (op)= absvor v,  δv (op)= -absvor u  (Barotropic_Dynamics.jl:28-29,
// Shallow_Water_Dynamics.jl:129-130, Spectral_Dynamics.jl:300-301).
// accumulate = 0: assign, 1: add.
// The explicit vorticity-flux residual and the kinetic energy with their
// fused multiply-adds written out, shared by the stand-alone kernels and the
// fused shallow-water kernel so that both round alike.
__device__ __forceinline__ void vorticity_flux_acc(double a, double ui, double vi, double& du,
                                                   double& dv) {
  du = __fma_rn(a, vi, du);    // δu += absvor v
  dv = __fma_rn(-a, ui, dv);   // δv += -absvor u
}
__device__ __forceinline__ double kinetic(double ui, double vi) {
  return 0.5 * __fma_rn(ui, ui, vi * vi);
}

__global__ void vorticity_flux_kernel(const double* __restrict__ vor, const double* __restrict__ u,
                                      const double* __restrict__ v, const double* __restrict__ sint,
                                      double two_omega, double* __restrict__ absvor,
                                      double* __restrict__ du, double* __restrict__ dv, int nlat,
                                      int nlon, size_t n, int accumulate) {
  GRID_STRIDE(i, n) {
    int j = (int)((i / nlon) % nlat);
    double f = two_omega * sint[j];
    double a = vor[i] + f;
    if (absvor) absvor[i] = a;
    if (!du) continue;
    if (accumulate) {
      double x = du[i], y = dv[i];
      vorticity_flux_acc(a, u[i], v[i], x, y);
      du[i] = x;
      dv[i] = y;
    } else {
      du[i] = a * v[i];
      dv[i] = -a * u[i];
    }
  }
}

// out = a x + b y  (x or y may be NULL)
__global__ void axpby_kernel(double a, const double* __restrict__ x, double b,
                             const double* __restrict__ y, double* __restrict__ out, size_t n) {
  GRID_STRIDE(i, n) {
    double r;
    if (x && y) r = a * x[i] + b * y[i];
    else if (x) r = a * x[i];
    else if (y) r = b * y[i];
    else r = 0.0;
    out[i] = r;
  }
}

// out = (accumulate ? out : 0) + sign * x * y
__global__ void xmy_kernel(double sign, const double* __restrict__ x, const double* __restrict__ y,
                           double* __restrict__ out, size_t n, int accumulate) {
  GRID_STRIDE(i, n) {
    double p = x[i] * y[i];
    double r = sign < 0 ? -p : p;
    out[i] = accumulate ? out[i] + r : r;
  }
}

// 0.5 (u^2 + v^2) [+ base]   (Shallow_Water_Dynamics.jl:139, Spectral_Dynamics.jl:307)
__global__ void kinetic_energy_kernel(const double* __restrict__ u, const double* __restrict__ v,
                                      const double* __restrict__ base, double* __restrict__ out,
                                      size_t n) {
  GRID_STRIDE(i, n) {
    double k = kinetic(u[i], v[i]);
    out[i] = base ? base[i] + k : k;
  }
}

// Shallow_Water_Physics! (Shallow_Water_Dynamics.jl:7-21) and the first
// grid steps of Shallow_Water_Dynamics! (:126-130 absolute vorticity and the
// explicit residual, :139 kinetic energy) in one pass over the grid: the
// arithmetic of sw_physics_kernel, vorticity_flux_kernel (accumulating) and
// kinetic_energy_kernel.
__global__ void sw_explicit_kernel(const double* __restrict__ u, const double* __restrict__ v,
                                   const double* __restrict__ h, const double* __restrict__ h_eq,
                                   const double* __restrict__ vor, const double* __restrict__ sint,
                                   double km, double kt, double two_omega,
                                   double* __restrict__ du, double* __restrict__ dv,
                                   double* __restrict__ dh, double* __restrict__ absvor,
                                   double* __restrict__ kin, double* __restrict__ u_p,
                                   double* __restrict__ v_p, double* __restrict__ h_p, int nlat,
                                   int nlon, size_t n) {
  GRID_STRIDE(i, n) {
    const int j = (int)((i / nlon) % nlat);
    const double ui = u[i], vi = v[i], hi = h[i];
    double x = -km * ui, y = -km * vi;
    dh[i] = -kt * (hi - h_eq[i]);
    if (u_p) {  // the p = c half of Time_Advance! for u, v, h
      u_p[i] = ui;
      v_p[i] = vi;
      h_p[i] = hi;
    }
    const double a = vor[i] + two_omega * sint[j];
    absvor[i] = a;
    vorticity_flux_acc(a, ui, vi, x, y);
    du[i] = x;
    dv[i] = y;
    kin[i] = kinetic(ui, vi);
  }
}

// The shallow-water height tendency's grid terms (Shallow_Water_Dynamics.jl:147-148):
// Add_Horizontal_Advection!'s δh -= u ∂λh + v ∂θh (SSM:716, the arithmetic of
// advection_kernel) and δh -= h div (xmy_kernel with sign -1, accumulating).
__global__ void sw_height_tendency_kernel(double* __restrict__ dh, const double* __restrict__ u,
                                          const double* __restrict__ v, const double* __restrict__ gl,
                                          const double* __restrict__ gt, const double* __restrict__ h,
                                          const double* __restrict__ div, size_t n) {
  GRID_STRIDE(i, n) {
    double x = advect(dh[i], u[i], v[i], gl[i], gt[i]);
    double p = h[i] * div[i];
    dh[i] = x + -p;
  }
}

// out = f(x): 0 exp(x), 1 log(x), 2 a x + b
__global__ void map_kernel(int op, const double* __restrict__ x, double a, double b,
                           double* __restrict__ out, size_t n) {
  GRID_STRIDE(i, n) {
    double v = x[i];
    out[i] = op == 0 ? exp(v) : (op == 1 ? log(v) : a * v + b);
  }
}

// Shallow_Water_Physics! (Shallow_Water_Dynamics.jl:16-18)
__global__ void sw_physics_kernel(const double* __restrict__ u, const double* __restrict__ v,
                                  const double* __restrict__ h, const double* __restrict__ h_eq,
                                  double km, double kt, double* __restrict__ du,
                                  double* __restrict__ dv, double* __restrict__ dh, size_t n) {
  GRID_STRIDE(i, n) {
    du[i] = -km * u[i];
    dv[i] = -km * v[i];
    dh[i] = -kt * (h[i] - h_eq[i]);
  }
}

}  // namespace jgcm

using namespace jgcm;

extern "C" {

int jgcm_leapfrog_create(jgcm_mesh_t* mesh, double robert_coef, int damping_order,
                         double damping_coef, double implicit_coef, int64_t dt, int init_step,
                         int64_t start_time, int64_t end_time, jgcm_leapfrog_t** out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(out != nullptr, "out is NULL");
  JGCM_REQUIRE(damping_order % 2 == 0, "damping_order must be even (Time_Integrator.jl:24)");
  jgcm_leapfrog* L = new (std::nothrow) jgcm_leapfrog();
  if (!L) {
    set_error("out of host memory");
    return JGCM_ERR_NOMEM;
  }
  L->mesh = mesh;
  L->robert_coef = robert_coef;
  L->damping_order = damping_order;
  L->damping_coef = damping_coef;
  L->implicit_coef = implicit_coef;
  L->dt = dt;
  L->init_step = init_step != 0;
  L->time = start_time;
  L->start_time = start_time;
  L->end_time = end_time;
  int N = mesh->N, ns = mesh->ns;
  size_t ne = (size_t)(N + 1) * (ns + 1);
  L->damping.assign(ne, 0.0);
  double e_ref = mesh->laplacian_eig[(size_t)0 * (ns + 1) + (ns - 1)];  // eigen[1, num_spherical]
  for (size_t k = 0; k < ne; ++k)
    L->damping[k] = damping_coef * std::pow(mesh->laplacian_eig[k] / e_ref, (double)damping_order);
  hipError_t e = hipMalloc(&L->d_damping, ne * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpy(L->d_damping, L->damping.data(), ne * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("leapfrog table upload failed: %s", hipGetErrorString(e));
    if (L->d_damping) (void)hipFree(L->d_damping);
    delete L;
    return JGCM_ERR_HIP;
  }
  mesh_retain(mesh);   // released by jgcm_leapfrog_destroy
  *out = L;
  return JGCM_OK;
}

int jgcm_leapfrog_destroy(jgcm_leapfrog_t* lf) {
  if (!lf) return JGCM_OK;
  if (lf->mesh) (void)hipStreamSynchronize(lf->mesh->stream);
  if (lf->d_damping) (void)hipFree(lf->d_damping);
  jgcm_mesh* m = lf->mesh;
  delete lf;
  mesh_release(m);
  return JGCM_OK;
}

int jgcm_leapfrog_update_init_step(jgcm_leapfrog_t* lf) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(lf->init_step, "Update_Init_Step! called twice (Time_Integrator.jl:41)");
  lf->init_step = false;
  return JGCM_OK;
}

int jgcm_leapfrog_state(jgcm_leapfrog_t* lf, int* init_step, int64_t* dt, int64_t* time) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  if (init_step) *init_step = lf->init_step ? 1 : 0;
  if (dt) *dt = lf->dt;
  if (time) *time = lf->time;
  return JGCM_OK;
}

int jgcm_leapfrog_set_time(jgcm_leapfrog_t* lf, int64_t time) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  lf->time = time;
  return JGCM_OK;
}

int jgcm_leapfrog_damping(jgcm_leapfrog_t* lf, double* host_out, size_t count) {
  JGCM_REQUIRE(lf != nullptr && host_out != nullptr, "NULL argument");
  JGCM_REQUIRE(count >= lf->damping.size(), "damping table needs %zu doubles", lf->damping.size());
  for (size_t k = 0; k < lf->damping.size(); ++k) host_out[k] = lf->damping[k];
  return JGCM_OK;
}

static int launch_damped_leapfrog(jgcm_leapfrog_t* lf, double* dQ, const double* Qp, double* Qc,
                                  double* Qn, int nlev, int mode) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1");
  jgcm_mesh* M = lf->mesh;
  size_t n = (size_t)nlev * M->spec_elems();
  hipLaunchKernelGGL(damped_leapfrog_kernel, dim3(nblk(n)), dim3(256), 0, M->stream, (double2*)dQ,
                     (const double2*)Qp, (double2*)Qc, (double2*)Qn, lf->d_damping, M->nm, M->ns, n,
                     lf->init_step ? 1 : 0, (double)lf->dt, lf->robert_coef, mode);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_compute_spectral_damping(jgcm_leapfrog_t* lf, const double* Qc, const double* Qp,
                                  double* dQ, int nlev) {
  JGCM_REQUIRE(Qc && Qp && dQ, "NULL field pointer");
  return launch_damped_leapfrog(lf, dQ, Qp, const_cast<double*>(Qc), nullptr, nlev, 1);
}

int jgcm_filtered_leapfrog(jgcm_leapfrog_t* lf, const double* dQ, const double* Qp, double* Qc,
                           double* Qn, int nlev) {
  JGCM_REQUIRE(dQ && Qp && Qc && Qn, "NULL field pointer");
  return launch_damped_leapfrog(lf, const_cast<double*>(dQ), Qp, Qc, Qn, nlev, 2);
}

int jgcm_damped_leapfrog(jgcm_leapfrog_t* lf, double* dQ, const double* Qp, double* Qc, double* Qn,
                         int nlev) {
  JGCM_REQUIRE(dQ && Qp && Qc && Qn, "NULL field pointer");
  return launch_damped_leapfrog(lf, dQ, Qp, Qc, Qn, nlev, 3);
}

int jgcm_damped_leapfrog_batch(jgcm_leapfrog_t* lf, int nfields, double* const* dQ,
                               const double* const* Qp, double* const* Qc, double* const* Qn,
                               const int* nlev, const int* mode) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nfields >= 1 && nfields <= MAXLF, "nfields must be in 1..%d, got %d", MAXLF, nfields);
  JGCM_REQUIRE(dQ && Qp && Qc && Qn && nlev && mode, "NULL argument");
  jgcm_mesh* M = lf->mesh;
  LeapfrogSet s{};
  size_t maxn = 0;
  for (int f = 0; f < nfields; ++f) {
    JGCM_REQUIRE(mode[f] >= 1 && mode[f] <= 3, "mode[%d] must be 1, 2 or 3, got %d", f, mode[f]);
    JGCM_REQUIRE(nlev[f] >= 1, "nlev[%d] must be >= 1", f);
    JGCM_REQUIRE(dQ[f] && Qp[f] && Qc[f] && ((mode[f] & 2) == 0 || Qn[f]),
                 "NULL field pointer (field %d)", f);
    s.dQ[f] = (double2*)dQ[f];
    s.Qp[f] = (const double2*)Qp[f];
    s.Qc[f] = (double2*)Qc[f];
    s.Qn[f] = (double2*)Qn[f];
    s.n[f] = (size_t)nlev[f] * M->spec_elems();
    s.mode[f] = mode[f];
    maxn = s.n[f] > maxn ? s.n[f] : maxn;
  }
  hipLaunchKernelGGL(damped_leapfrog_batch_kernel, dim3(nblk(maxn), nfields), dim3(256), 0, M->stream,
                     s, lf->d_damping, M->nm, M->ns, lf->init_step ? 1 : 0, (double)lf->dt,
                     lf->robert_coef);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_sw_implicit_correction(jgcm_leapfrog_t* lf, const double* div_c, const double* div_p,
                                const double* h_c, const double* h_p, double h_0, double* ddiv,
                                double* dh, int nlev) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(div_c && div_p && h_c && h_p && ddiv && dh, "NULL field pointer");
  jgcm_mesh* M = lf->mesh;
  size_t n = (size_t)nlev * M->spec_elems();
  double dtv = lf->init_step ? (double)lf->dt : 2.0 * lf->dt;
  double mu = lf->implicit_coef * dtv;
  hipLaunchKernelGGL(sw_implicit_kernel, dim3(nblk(n)), dim3(256), 0, M->stream, (const double2*)div_c,
                     (const double2*)div_p, (const double2*)h_c, (const double2*)h_p, h_0,
                     (double2*)ddiv, (double2*)dh, M->d_eig, M->nm, M->ns, n, mu);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_vorticity_flux(jgcm_mesh_t* mesh, const double* grid_vor, const double* grid_u,
                        const double* grid_v, double omega, double* grid_absvor, double* grid_du,
                        double* grid_dv, int nlev, int accumulate) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(grid_vor, "NULL field pointer");
  JGCM_REQUIRE((grid_du && grid_dv && grid_u && grid_v) || (!grid_du && !grid_dv && grid_absvor),
               "need (u, v, δu, δv) or an absvor output only");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(vorticity_flux_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream, grid_vor, grid_u,
                     grid_v, mesh->d_sin, 2 * omega, grid_absvor, grid_du, grid_dv, mesh->nlat,
                     mesh->nlon, n, accumulate);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_axpby(jgcm_mesh_t* mesh, size_t count, double a, const double* x, double b,
               const double* y, double* out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(out, "NULL output");
  if (count == 0) return JGCM_OK;
  hipLaunchKernelGGL(axpby_kernel, dim3(nblk(count)), dim3(256), 0, mesh->stream, a, x, b, y, out, count);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_xmy(jgcm_mesh_t* mesh, size_t count, double sign, const double* x, const double* y,
             double* out, int accumulate) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(x && y && out, "NULL pointer");
  if (count == 0) return JGCM_OK;
  hipLaunchKernelGGL(xmy_kernel, dim3(nblk(count)), dim3(256), 0, mesh->stream, sign, x, y, out, count,
                     accumulate);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_map(jgcm_mesh_t* mesh, int op, size_t count, const double* x, double a, double b,
             double* out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(x && out, "NULL pointer");
  JGCM_REQUIRE(op >= 0 && op <= 2, "unknown map op %d", op);
  if (count == 0) return JGCM_OK;
  hipLaunchKernelGGL(map_kernel, dim3(nblk(count)), dim3(256), 0, mesh->stream, op, x, a, b, out, count);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_kinetic_energy(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                        const double* grid_base, double* grid_out, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(grid_u && grid_v && grid_out, "NULL pointer");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(kinetic_energy_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream, grid_u, grid_v,
                     grid_base, grid_out, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_sw_explicit(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                     const double* grid_h, const double* h_eq, double kappa_m, double kappa_t,
                     const double* grid_vor, double omega, double* grid_du, double* grid_dv,
                     double* grid_dh, double* grid_absvor, double* grid_kin, double* grid_u_p,
                     double* grid_v_p, double* grid_h_p, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(grid_u && grid_v && grid_h && h_eq && grid_vor && grid_du && grid_dv && grid_dh &&
                   grid_absvor && grid_kin,
               "NULL pointer");
  JGCM_REQUIRE((grid_u_p && grid_v_p && grid_h_p) || (!grid_u_p && !grid_v_p && !grid_h_p),
               "give all three p arrays or none");
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(sw_explicit_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream, grid_u, grid_v,
                     grid_h, h_eq, grid_vor, mesh->d_sin, kappa_m, kappa_t, 2 * omega, grid_du,
                     grid_dv, grid_dh, grid_absvor, grid_kin, grid_u_p, grid_v_p, grid_h_p,
                     mesh->nlat, mesh->nlon, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_sw_height_tendency(jgcm_mesh_t* mesh, const double* spe_h, const double* grid_u,
                            const double* grid_v, const double* grid_h, const double* grid_div,
                            double* grid_dh, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(spe_h && grid_u && grid_v && grid_h && grid_div && grid_dh, "NULL field pointer");
  // the grid scratch for both gradients exists before they are handed out
  if (int rc = ensure_scratch(mesh, 2 * nlev)) return rc;
  if (int rc = jgcm_compute_gradients(mesh, spe_h, mesh->d_grid_a, mesh->d_grid_b, nlev)) return rc;
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(sw_height_tendency_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream, grid_dh,
                     grid_u, grid_v, mesh->d_grid_a, mesh->d_grid_b, grid_h, grid_div, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_sw_height_tendency_carried(jgcm_mesh_t* mesh, const double* grid_dlambda_h,
                                    const double* grid_dtheta_h, const double* grid_u,
                                    const double* grid_v, const double* grid_h,
                                    const double* grid_div, double* grid_dh, int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(grid_dlambda_h && grid_dtheta_h && grid_u && grid_v && grid_h && grid_div && grid_dh,
               "NULL field pointer");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(sw_height_tendency_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream, grid_dh,
                     grid_u, grid_v, grid_dlambda_h, grid_dtheta_h, grid_h, grid_div, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_sw_spectral_step(jgcm_leapfrog_t* lf, double h_0, const double* spe_kin, double* spe_energy,
                          double* spe_dvor, double* spe_ddiv, double* spe_dh, double* const* vor,
                          double* const* div, double* const* h, int nlev) {
  return jgcm_sw_spectral_step_ex(lf, h_0, spe_kin, spe_energy, spe_dvor, spe_ddiv, spe_dh, vor, div,
                                  h, nlev, 1);
}

int jgcm_sw_spectral_step_ex(jgcm_leapfrog_t* lf, double h_0, const double* spe_kin,
                             double* spe_energy, double* spe_dvor, double* spe_ddiv, double* spe_dh,
                             double* const* vor, double* const* div, double* const* h, int nlev,
                             int advance) {
  JGCM_REQUIRE(lf != nullptr, "leapfrog is NULL");
  JGCM_REQUIRE(nlev >= 1, "nlev must be >= 1, got %d", nlev);
  JGCM_REQUIRE(spe_kin && spe_energy && spe_dvor && spe_ddiv && spe_dh && vor && div && h,
               "NULL argument");
  SWSpectral s{};
  s.kin = (const double2*)spe_kin;
  s.energy = (double2*)spe_energy;
  s.dvor = (double2*)spe_dvor;
  s.ddiv = (double2*)spe_ddiv;
  s.dh = (double2*)spe_dh;
  for (int k = 0; k < 3; ++k) {
    JGCM_REQUIRE(vor[k] && div[k] && h[k], "NULL field pointer (time level %d)", k);
    s.vor[k] = (double2*)vor[k];
    s.div[k] = (double2*)div[k];
    s.h[k] = (double2*)h[k];
  }
  jgcm_mesh* M = lf->mesh;
  size_t n = (size_t)nlev * M->spec_elems();
  double dtv = lf->init_step ? (double)lf->dt : 2.0 * lf->dt;
  double mu = lf->implicit_coef * dtv;
  hipLaunchKernelGGL(sw_spectral_kernel, dim3(nblk(n)), dim3(256), 0, M->stream, s, M->d_eig,
                     lf->d_damping, M->nm, M->ns, n, h_0, mu, lf->init_step ? 1 : 0,
                     (double)lf->dt, lf->robert_coef, advance ? 1 : 0);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_shallow_water_physics(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                               const double* grid_h, const double* h_eq, double kappa_m,
                               double kappa_t, double* grid_du, double* grid_dv, double* grid_dh,
                               int nlev) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(grid_u && grid_v && grid_h && h_eq && grid_du && grid_dv && grid_dh, "NULL pointer");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(sw_physics_kernel, dim3(nblk(n)), dim3(256), 0, mesh->stream, grid_u, grid_v, grid_h,
                     h_eq, kappa_m, kappa_t, grid_du, grid_dv, grid_dh, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

}  // extern "C"
