// Spectral-space and grid-space operators of the transform core.
// Reference: src/Atmos_Spectral/Spectral_Spherical_Mesh.jl (SSM), lines cited
// per kernel.  All kernels are elementwise (or an n +- 1 stencil in spectral
// space) and HBM-bound; they run one thread per element with grid-stride loops.
#include <new>

#include "jgcm_internal.hpp"

namespace jgcm {

int fft_fold(const jgcm_mesh* M, const double* grid, double2* eo, const double* cosmul, int B);
int ifft_unfold(const jgcm_mesh* M, const double2* eo, double* grid, const double* divcos, int B);
int legendre_inv(const jgcm_mesh* M, const double2* snm, double2* eo, int B);
int inverse_batch(const jgcm_mesh* M, int nseg, const double* const* snm,
                  const double* const* snm_b, double* const* grid, const int* nlev,
                  const int* flags);

static inline int nblocks(size_t n, int per = 256) {
  size_t b = (n + per - 1) / per;
  return (int)(b > 8192 ? 8192 : (b == 0 ? 1 : b));
}

#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

__global__ void copy_kernel(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  GRID_STRIDE(i, n) dst[i] = src[i];
}

__global__ void copy_bytes_kernel(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src,
                                  size_t n) {
  GRID_STRIDE(i, n) dst[i] = src[i];
}

__global__ void fill_kernel(double* __restrict__ dst, double v, size_t n) {
  GRID_STRIDE(i, n) dst[i] = v;
}

// Time_Advance! (Dyn_Data.jl:326-364): blockIdx.y = field; the same thread
// moves c -> p and then n -> c for an element, so one launch covers all fields.
constexpr int MAXCOPY = 32;
struct CopySet {
  double* dst[MAXCOPY];
  const double* src[MAXCOPY];
  size_t count[MAXCOPY];  // doubles
};

// blockIdx.y = array; a grid-stride loop over its elements
__global__ void copy_batch_kernel(CopySet s) {
  const int f = blockIdx.y;
  double* __restrict__ d = s.dst[f];
  const double* __restrict__ x = s.src[f];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < s.count[f];
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = x[i];
}

// blockIdx.y = array: dst += src
__global__ void add_batch_kernel(CopySet s) {
  const int f = blockIdx.y;
  double* __restrict__ d = s.dst[f];
  const double* __restrict__ x = s.src[f];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < s.count[f];
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = d[i] + x[i];
}

constexpr int MAXADV = 16;
struct AdvanceSet {
  double* p[MAXADV];
  double* c[MAXADV];
  const double* n[MAXADV];
  size_t count[MAXADV];
};

__global__ void time_advance_kernel(AdvanceSet s) {
  const int f = blockIdx.y;
  double* __restrict__ p = s.p[f];
  double* __restrict__ c = s.c[f];
  const double* __restrict__ n = s.n[f];
  GRID_STRIDE(i, s.count[f]) {
    double x = c[i];
    p[i] = x;
    c[i] = n[i];
  }
}

// grid [B][nlat][nlon] scaled by cos θ_j: divide (SSM:333-338) or multiply (SSM:341-346)
__global__ void cos_scale_kernel(const double* __restrict__ in, double* __restrict__ out,
                                 const double* __restrict__ cost, int nlat, int nlon, size_t n,
                                 int divide) {
  GRID_STRIDE(i, n) {
    int j = (int)((i / nlon) % nlat);
    out[i] = divide ? in[i] / cost[j] : in[i] * cost[j];
  }
}

// SSM:517-580.  Spectral layout [b][n][m]; coefficient tables [m][n].
__global__ void ucos_vcos_kernel(const double2* __restrict__ vor, const double2* __restrict__ div,
                                 double2* __restrict__ ucos, double2* __restrict__ vcos,
                                 const double* __restrict__ uvc, const double* __restrict__ uvm,
                                 const double* __restrict__ uvp, int nm, int ns, size_t n_all) {
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int n = (int)((i / nm) % (ns + 1));
    size_t base = i - (size_t)n * nm;  // element (b, 0, m)
    int c = m * (ns + 1) + n;
    double2 d = div[i], v = vor[i];
    double cc = uvc[c];
    double2 u = make_double2(cc * (-d.y), cc * d.x);   // uvc * (i div)
    double2 w = make_double2(cc * (-v.y), cc * v.x);   // uvc * (i vor)
    if (n >= 1) {
      double a = uvm[c];
      double2 vm = vor[base + (size_t)(n - 1) * nm], dm = div[base + (size_t)(n - 1) * nm];
      u = make_double2(u.x - a * vm.x, u.y - a * vm.y);
      w = make_double2(w.x + a * dm.x, w.y + a * dm.y);
    }
    if (n <= ns - 1) {
      double a = uvp[c];
      double2 vp = vor[base + (size_t)(n + 1) * nm], dp = div[base + (size_t)(n + 1) * nm];
      u = make_double2(u.x - a * vp.x, u.y - a * vp.y);
      w = make_double2(w.x + a * dp.x, w.y + a * dp.y);
    }
    ucos[i] = u;
    vcos[i] = w;
  }
}

// SSM:614-621
__global__ void laplacian_kernel(double2* __restrict__ s, const double* __restrict__ eig, int nm,
                                 int ns, size_t n_all) {
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int n = (int)((i / nm) % (ns + 1));
    double e = eig[m * (ns + 1) + n];
    double2 v = s[i];
    s[i] = make_double2(e * v.x, e * v.y);
  }
}

// SSM:655-692
__global__ void gradient_cos_kernel(const double2* __restrict__ h, double2* __restrict__ dl,
                                    double2* __restrict__ dt, const double* __restrict__ cdl,
                                    const double* __restrict__ cdm, const double* __restrict__ cdp,
                                    int nm, int ns, size_t n_all) {
  GRID_STRIDE(i, n_all) {
    int m = (int)(i % nm);
    int n = (int)((i / nm) % (ns + 1));
    size_t base = i - (size_t)n * nm;
    int c = m * (ns + 1) + n;
    double2 v = h[i];
    double a = cdl[c];
    dl[i] = make_double2(a * (-v.y), a * v.x);
    double2 t = make_double2(0.0, 0.0);
    if (n <= ns - 1) {
      double2 hp = h[base + (size_t)(n + 1) * nm];
      t = make_double2(cdp[c] * hp.x, cdp[c] * hp.y);
    }
    if (n >= 1) {
      double2 hm = h[base + (size_t)(n - 1) * nm];
      t = make_double2(t.x + cdm[c] * hm.x, t.y + cdm[c] * hm.y);
    }
    dt[i] = t;
  }
}

// SSM:716: δh -= u dλh + v dθh
__global__ void advection_kernel(double* __restrict__ dh, const double* __restrict__ u,
                                 const double* __restrict__ v, const double* __restrict__ gl,
                                 const double* __restrict__ gt, size_t n) {
  GRID_STRIDE(i, n) dh[i] = advect(dh[i], u[i], v[i], gl[i], gt[i]);
}

// SSM:746-760 (level 1): sum_i sum_j f[j][i] w_j / (2 nλ), one workgroup,
// fixed summation order (deterministic).
__global__ __launch_bounds__(256) void area_mean_kernel(const double* __restrict__ f,
                                                        const double* __restrict__ w, int nlat,
                                                        int nlon, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nlon; i += 256) {
    double col = 0.0;
    for (int j = 0; j < nlat; ++j) col += f[(size_t)j * nlon + i] * w[j];
    s += col;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / (2.0 * nlon);
}

int area_mean_device(const jgcm_mesh* M, const double* grid, double* out_dev) {
  hipLaunchKernelGGL(area_mean_kernel, dim3(1), dim3(256), 0, M->stream, grid, M->d_wts, M->nlat,
                     M->nlon, out_dev);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int ucos_vcos(const jgcm_mesh* M, const double2* vor, const double2* div, double2* ucos,
              double2* vcos, int B) {
  size_t n = (size_t)B * M->spec_elems();
  hipLaunchKernelGGL(ucos_vcos_kernel, dim3(nblocks(n)), dim3(256), 0, M->stream, vor, div, ucos,
                     vcos, M->d_uvc, M->d_uvm, M->d_uvp, M->nm, M->ns, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int gradient_cos(const jgcm_mesh* M, const double2* h, double2* dl, double2* dt, int B) {
  size_t n = (size_t)B * M->spec_elems();
  hipLaunchKernelGGL(gradient_cos_kernel, dim3(nblocks(n)), dim3(256), 0, M->stream, h, dl, dt,
                     M->d_dlam, M->d_dthm, M->d_dthp, M->nm, M->ns, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

// Two inverse transforms with the rows divided by cos θ, in one Legendre launch
// and one FFT launch; the spectral operators (gradient, u cos / v cos) are
// applied while the Legendre stage stages its coefficients.
static int inverse_pair_divcos(const jgcm_mesh* M, const double* sa, const double* sa2, int op_a,
                               const double* sb, const double* sb2, int op_b, double* ga,
                               double* gb, int B) {
  const double* s[2] = {sa, sb};
  const double* s2[2] = {sa2, sb2};
  double* g[2] = {ga, gb};
  int nl[2] = {B, B};
  int fl[2] = {1 | op_a, 1 | op_b};
  return inverse_batch(M, 2, s, s2, g, nl, fl);
}

}  // namespace jgcm

using namespace jgcm;

#define JGCM_MESH_PROLOGUE(mesh, nlev)                                           \
  if (int _rc = check_mesh(mesh)) return _rc;                                    \
  JGCM_REQUIRE((nlev) >= 1, "nlev must be >= 1, got %d", (nlev));                \
  if (int _rc = ensure_scratch(mesh, (nlev))) return _rc;

extern "C" {

int jgcm_copy_batch(jgcm_mesh_t* mesh, int n, double* const* dst, const double* const* src,
                    const size_t* bytes) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(n >= 1 && n <= MAXCOPY, "n must be in 1..%d, got %d", MAXCOPY, n);
  JGCM_REQUIRE(dst && src && bytes, "NULL argument");
  CopySet s{};
  size_t maxc = 0;
  for (int f = 0; f < n; ++f) {
    JGCM_REQUIRE(dst[f] && src[f], "NULL pointer (array %d)", f);
    JGCM_REQUIRE(bytes[f] % 8 == 0 && (uintptr_t)dst[f] % 8 == 0 && (uintptr_t)src[f] % 8 == 0,
                 "array %d: sizes and addresses must be multiples of 8 bytes", f);
    s.dst[f] = dst[f];
    s.src[f] = src[f];
    s.count[f] = bytes[f] / 8;
    if (s.count[f] > maxc) maxc = s.count[f];
  }
  if (maxc == 0) return JGCM_OK;
  hipLaunchKernelGGL(copy_batch_kernel, dim3(nblocks(maxc, 1024), n), dim3(256), 0, mesh->stream,
                     s);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_add_batch(jgcm_mesh_t* mesh, int n, double* const* dst, const double* const* src,
                   const size_t* count) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(n >= 1 && n <= MAXCOPY, "n must be in 1..%d, got %d", MAXCOPY, n);
  JGCM_REQUIRE(dst && src && count, "NULL argument");
  CopySet s{};
  size_t maxc = 0;
  for (int f = 0; f < n; ++f) {
    JGCM_REQUIRE(dst[f] && src[f], "NULL pointer (array %d)", f);
    for (int q = 0; q < f; ++q)
      JGCM_REQUIRE(dst[q] + count[q] <= dst[f] || dst[f] + count[f] <= dst[q],
                   "destinations %d and %d overlap", q, f);
    s.dst[f] = dst[f];
    s.src[f] = src[f];
    s.count[f] = count[f];
    if (count[f] > maxc) maxc = count[f];
  }
  if (maxc == 0) return JGCM_OK;
  hipLaunchKernelGGL(add_batch_kernel, dim3(nblocks(maxc, 1024), n), dim3(256), 0, mesh->stream,
                     s);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_copy(jgcm_mesh_t* mesh, void* dst, const void* src, size_t bytes) {
  if (int rc = check_mesh(mesh)) return rc;
  if (bytes == 0) return JGCM_OK;
  JGCM_REQUIRE(dst && src, "NULL pointer");
  if (bytes % 8 == 0 && ((uintptr_t)dst % 8 == 0) && ((uintptr_t)src % 8 == 0)) {
    size_t n = bytes / 8;
    hipLaunchKernelGGL(copy_kernel, dim3(nblocks(n)), dim3(256), 0, mesh->stream, (double*)dst,
                       (const double*)src, n);
  } else {
    hipLaunchKernelGGL(copy_bytes_kernel, dim3(nblocks(bytes)), dim3(256), 0, mesh->stream,
                       (unsigned char*)dst, (const unsigned char*)src, bytes);
  }
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_time_advance(jgcm_mesh_t* mesh, int nfields, double* const* p, double* const* c,
                      const double* const* n, const size_t* count) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nfields >= 1 && nfields <= MAXADV, "nfields must be in 1..%d, got %d", MAXADV, nfields);
  JGCM_REQUIRE(p && c && n && count, "NULL argument");
  AdvanceSet s{};
  size_t maxc = 0;
  for (int f = 0; f < nfields; ++f) {
    JGCM_REQUIRE(p[f] && c[f] && n[f], "NULL field pointer (field %d)", f);
    s.p[f] = p[f];
    s.c[f] = c[f];
    s.n[f] = n[f];
    s.count[f] = count[f];
    maxc = count[f] > maxc ? count[f] : maxc;
  }
  if (maxc == 0) return JGCM_OK;
  hipLaunchKernelGGL(time_advance_kernel, dim3(nblocks(maxc, 1024), nfields), dim3(256), 0,
                     mesh->stream, s);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_fill(jgcm_mesh_t* mesh, double* dst, double value, size_t count) {
  if (int rc = check_mesh(mesh)) return rc;
  if (count == 0) return JGCM_OK;
  JGCM_REQUIRE(dst, "NULL pointer");
  hipLaunchKernelGGL(fill_kernel, dim3(nblocks(count)), dim3(256), 0, mesh->stream, dst, value, count);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_divide_by_cos(jgcm_mesh_t* mesh, double* grid, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(grid, "NULL field pointer");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(cos_scale_kernel, dim3(nblocks(n)), dim3(256), 0, mesh->stream, grid, grid,
                     mesh->d_cos, mesh->nlat, mesh->nlon, n, 1);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_multiply_by_cos(jgcm_mesh_t* mesh, const double* grid, double* grid_cos, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(grid && grid_cos, "NULL field pointer");
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(cos_scale_kernel, dim3(nblocks(n)), dim3(256), 0, mesh->stream, grid, grid_cos,
                     mesh->d_cos, mesh->nlat, mesh->nlon, n, 0);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_compute_ucos_vcos_from_vor_div(jgcm_mesh_t* mesh, const double* vor, const double* div,
                                        double* ucos, double* vcos, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(vor && div && ucos && vcos, "NULL field pointer");
  return ucos_vcos(mesh, (const double2*)vor, (const double2*)div, (double2*)ucos, (double2*)vcos, nlev);
}

int jgcm_uv_grid_from_vor_div(jgcm_mesh_t* mesh, const double* vor, const double* div,
                              double* grid_u, double* grid_v, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(vor && div && grid_u && grid_v, "NULL field pointer");
  if (int rc = ensure_scratch(mesh, 2 * nlev)) return rc;
  return inverse_pair_divcos(mesh, vor, div, JGCM_INV_UCOS, vor, div, JGCM_INV_VCOS, grid_u, grid_v,
                             nlev);
}

int jgcm_apply_laplacian(jgcm_mesh_t* mesh, double* snm, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(snm, "NULL field pointer");
  size_t n = (size_t)nlev * mesh->spec_elems();
  hipLaunchKernelGGL(laplacian_kernel, dim3(nblocks(n)), dim3(256), 0, mesh->stream, (double2*)snm,
                     mesh->d_eig, mesh->nm, mesh->ns, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_compute_gradient_cos(jgcm_mesh_t* mesh, const double* spe_hs, double* spe_cos_dlambda,
                              double* spe_cos_dtheta, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(spe_hs && spe_cos_dlambda && spe_cos_dtheta, "NULL field pointer");
  return gradient_cos(mesh, (const double2*)spe_hs, (double2*)spe_cos_dlambda,
                      (double2*)spe_cos_dtheta, nlev);
}

int jgcm_compute_gradients(jgcm_mesh_t* mesh, const double* spe_hs, double* grid_dlambda,
                           double* grid_dtheta, int nlev) {
  JGCM_MESH_PROLOGUE(mesh, nlev);
  JGCM_REQUIRE(spe_hs && grid_dlambda && grid_dtheta, "NULL field pointer");
  if (int rc = ensure_scratch(mesh, 2 * nlev)) return rc;
  return inverse_pair_divcos(mesh, spe_hs, nullptr, JGCM_INV_GRAD_LAMBDA, spe_hs, nullptr,
                             JGCM_INV_GRAD_THETA, grid_dlambda, grid_dtheta, nlev);
}

int jgcm_add_horizontal_advection(jgcm_mesh_t* mesh, const double* spe_hs, const double* grid_u,
                                  const double* grid_v, double* grid_dhs, int nlev) {
  // scratch for 2 nlev fields first: the gradients below must not reallocate
  // the grid scratch this function hands them
  JGCM_MESH_PROLOGUE(mesh, 2 * nlev);
  JGCM_REQUIRE(spe_hs && grid_u && grid_v && grid_dhs, "NULL field pointer");
  if (int rc = jgcm_compute_gradients(mesh, spe_hs, mesh->d_grid_a, mesh->d_grid_b, nlev)) return rc;
  size_t n = (size_t)nlev * mesh->grid_elems();
  hipLaunchKernelGGL(advection_kernel, dim3(nblocks(n)), dim3(256), 0, mesh->stream, grid_dhs, grid_u,
                     grid_v, mesh->d_grid_a, mesh->d_grid_b, n);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

int jgcm_area_weighted_global_mean(jgcm_mesh_t* mesh, const double* grid, double* result) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(grid && result, "NULL pointer");
  if (int rc = area_mean_device(mesh, grid, mesh->d_scalar)) return rc;
  JGCM_CHECK_HIP(hipMemcpyAsync(result, mesh->d_scalar, sizeof(double), hipMemcpyDeviceToHost,
                                mesh->stream));
  JGCM_CHECK_HIP(hipStreamSynchronize(mesh->stream));
  return JGCM_OK;
}

}  // extern "C"

// ---- graph capture -----------------------------------------------------------
// A graph keeps a reference to every mesh whose calls it recorded (the mesh it
// was begun on, meshes forked into the capture, meshes sharing its stream):
// their streams, events and scratch buffers are kernel arguments or launch
// targets of the graph, so none of them may be freed, and none of their
// scratch may be released by growth, while the graph exists.
struct jgcm_graph {
  jgcm_mesh* mesh = nullptr;
  std::vector<jgcm_mesh*> used;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int jgcm_mesh_reserve(jgcm_mesh_t* mesh, int nfields) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(nfields >= 1, "nfields must be >= 1");
  return ensure_scratch(mesh, nfields);
}

int jgcm_mesh_wait(jgcm_mesh_t* waiter, jgcm_mesh_t* signaller) {
  if (int rc = check_mesh(waiter)) return rc;
  if (int rc = check_mesh(signaller)) return rc;
  if (waiter->stream == signaller->stream) return JGCM_OK;
  JGCM_CHECK_HIP(hipEventRecord(signaller->signal, signaller->stream));
  JGCM_CHECK_HIP(hipStreamWaitEvent(waiter->stream, signaller->signal, 0));
  return JGCM_OK;
}

int jgcm_graph_begin(jgcm_mesh_t* mesh) {
  if (int rc = check_mesh(mesh)) return rc;
  if (capture_ctx() != nullptr) {
    set_error("this thread is already capturing a graph");
    return JGCM_ERR_STATE;
  }
  JGCM_CHECK_HIP(hipStreamBeginCapture(mesh->stream, hipStreamCaptureModeThreadLocal));
  capture_open(mesh);
  return JGCM_OK;
}

int jgcm_graph_end(jgcm_mesh_t* mesh, jgcm_graph_t** out) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(out != nullptr, "out is NULL");
  CaptureCtx* ctx = capture_ctx();
  if (ctx == nullptr || ctx->origin != mesh) {
    set_error("jgcm_graph_end: no capture was begun on this mesh by this thread");
    return JGCM_ERR_STATE;
  }
  jgcm_graph* g = new (std::nothrow) jgcm_graph();
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(mesh->stream, &graph);
  std::vector<jgcm_mesh*> used;
  capture_close(&used);
  if (!g) {
    if (graph) (void)hipGraphDestroy(graph);
    set_error("out of host memory");
    return JGCM_ERR_NOMEM;
  }
  g->mesh = mesh;
  g->graph = graph;
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    set_error("graph capture failed: %s", hipGetErrorString(e));
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return JGCM_ERR_HIP;
  }
  g->used = std::move(used);
  for (jgcm_mesh* m : g->used) {
    mesh_retain(m);
    m->live_graphs.fetch_add(1);
  }
  *out = g;
  return JGCM_OK;
}

int jgcm_graph_launch(jgcm_graph_t* graph) {
  JGCM_REQUIRE(graph != nullptr, "graph is NULL");
  JGCM_CHECK_HIP(hipGraphLaunch(graph->exec, graph->mesh->stream));
  return JGCM_OK;
}

int jgcm_graph_destroy(jgcm_graph_t* graph) {
  if (!graph) return JGCM_OK;
  (void)hipStreamSynchronize(graph->mesh->stream);
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  std::vector<jgcm_mesh*> used = std::move(graph->used);
  delete graph;
  for (jgcm_mesh* m : used) {
    m->live_graphs.fetch_sub(1);
    mesh_release(m);
  }
  return JGCM_OK;
}

}  // extern "C"
