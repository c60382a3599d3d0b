// Mesh construction for libjgcm_hip: Gaussian grid, Legendre tables and the
// packed row tables the transform kernels read.
//
// Reference: src/Atmos_Spectral/Spectral_Spherical_Mesh.jl:69-139 (constructor)
// and src/Atmos_Spectral/Gauss_And_Legendre.jl (nodes, weights, recurrences).
// The recurrences and the Newton iteration run in FP64 on the GPU; FP
// contraction is disabled in this file so each operation rounds as written.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "jgcm_internal.hpp"

namespace jgcm {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

static thread_local CaptureCtx* g_capture = nullptr;

CaptureCtx* capture_ctx() { return g_capture; }

void capture_open(jgcm_mesh* origin) {
  g_capture = new CaptureCtx();
  g_capture->origin = origin;
  g_capture->meshes.push_back(origin);
}

void capture_close(std::vector<jgcm_mesh*>* used) {
  if (!g_capture) return;
  if (used) used->swap(g_capture->meshes);
  delete g_capture;
  g_capture = nullptr;
}

int check_mesh(const jgcm_mesh* m) {
  if (m == nullptr) {
    set_error("mesh handle is NULL");
    return JGCM_ERR_ARG;
  }
  if (g_capture) {
    jgcm_mesh* mm = const_cast<jgcm_mesh*>(m);
    auto& v = g_capture->meshes;
    if (std::find(v.begin(), v.end(), mm) == v.end()) v.push_back(mm);
  }
  return JGCM_OK;
}

void mesh_retain(jgcm_mesh* m) {
  if (m) m->refs.fetch_add(1);
}

// Gauss_And_Legendre.jl:81-115, one thread per node pair (i = 1..n/2).
__global__ void gauss_nodes_kernel(int n, double* __restrict__ sin_theta,
                                   double* __restrict__ wts, int* __restrict__ nonconv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (i > n / 2) return;
  const int itermax = 10000;
  const double tol = 1.0e-15;
  double dp = 0.0;
  double z = cos(M_PI * (i - 0.25) / (n + 0.5));
  for (int iter = 1; iter <= itermax; ++iter) {
    double p2 = 0.0, p1 = 1.0;
    for (int j = 1; j <= n; ++j) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    dp = n * (z * p1 - p2) / (z * z - 1.0);
    double z1 = z;
    z = z1 - p1 / dp;
    if (fabs(z - z1) <= tol) break;
    if (iter == itermax) atomicAdd(nonconv, 1);
  }
  sin_theta[i - 1] = -z;
  sin_theta[n - i] = z;
  double w = 2.0 / ((1.0 - z * z) * dp * dp);
  wts[i - 1] = w;
  wts[n - i] = w;
}

__device__ inline double eps_ml(int m, int l) {
  // ε[m,l] = sqrt((l^2 - m^2) / (4 l^2 - 1))   Gauss_And_Legendre.jl:41
  return sqrt((double)(l * l - m * m) / (4.0 * l * l - 1.0));
}

// Gauss_And_Legendre.jl:21-62: one thread per (m, latitude).  Writes
// qnm[m][n][j] for n = 0..ns+1 (the top degree is working space) and
// dqnm[m][n][j] for n = 0..ns.
__global__ void legendre_kernel(int N, int ns, int nlat, const double* __restrict__ sin_theta,
                                double* __restrict__ qnm, double* __restrict__ dqnm) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (j >= nlat || m > N) return;
  const int nq = ns + 2;
  double s = sin_theta[j];
  double c = sqrt(1.0 - s * s);
  double* q = qnm + ((size_t)m * nq) * nlat + j;
  double* dq = dqnm + ((size_t)m * (ns + 1)) * nlat + j;
  for (int n = 0; n < m; ++n) q[(size_t)n * nlat] = 0.0;
  for (int n = 0; n < m; ++n) dq[(size_t)n * nlat] = 0.0;
  // sectoral P_mm: qnm[k,k] = sqrt((2k+1)/(2k)) cos qnm[k-1,k-1]   (:30-32)
  double pmm = 1.0;
  for (int k = 1; k <= m; ++k) pmm = sqrt((2.0 * k + 1) / (2.0 * k)) * c * pmm;
  q[(size_t)m * nlat] = pmm;
  // P_{m+1,m} = sqrt(2m+3) sin P_mm                                    (:34-36)
  if (m + 1 < nq) q[(size_t)(m + 1) * nlat] = sqrt(2.0 * m + 3) * s * pmm;
  // three-term recurrence                                              (:45-51)
  double qm2 = pmm, qm1 = (m + 1 < nq) ? q[(size_t)(m + 1) * nlat] : 0.0;
  for (int l = m + 2; l <= ns + 1; ++l) {
    double ql = (s * qm1 - eps_ml(m, l - 1) * qm2) / eps_ml(m, l);
    q[(size_t)l * nlat] = ql;
    qm2 = qm1;
    qm1 = ql;
  }
  // dP/dmu                                                             (:53-62)
  double c2 = c * c;
  for (int l = m; l <= ns; ++l) {
    double t = (-l * eps_ml(m, l + 1)) * q[(size_t)(l + 1) * nlat];
    if (l != m) t = t + ((l + 1) * eps_ml(m, l)) * q[(size_t)(l - 1) * nlat];
    dq[(size_t)l * nlat] = t / c2;
  }
}

// Packed row tables on the southern half: row(m, n) = row_start[m] + n - m.
__global__ void pack_rows_kernel(int N, int ns, int nlat, double radius,
                                 const int* __restrict__ row_start,
                                 const double* __restrict__ qnm,   // [N+1][ns+2][nlat]
                                 const double* __restrict__ dqnm,  // [N+1][ns+1][nlat]
                                 const double* __restrict__ wts, const double* __restrict__ cos_t,
                                 RowTables rt) {
  int half = nlat / 2;
  int jp = blockIdx.x * blockDim.x + threadIdx.x;
  int n = blockIdx.y;
  int m = blockIdx.z;
  if (jp >= half || n < m || n > ns || m > N) return;
  size_t row = (size_t)row_start[m] + (n - m);
  size_t o = row * half + jp;
  double q = qnm[((size_t)m * (ns + 2) + n) * nlat + jp];
  double dq = dqnm[((size_t)m * (ns + 1) + n) * nlat + jp];
  double w = wts[jp];
  double c = cos_t[jp];
  double inv2r = 1.0 / (2.0 * radius);
  double vals[4];
  vals[0] = q;                              // qnm                 (SSM:174)
  vals[1] = (q * w) / 2.0;                  // qwg / 2             (SSM:96-99, :283)
  vals[2] = ((w * q) / (c * c) * m) * inv2r;  // coef_alp_a/(2r)   (SSM:416)
  vals[3] = (-w * dq) * inv2r;              // coef_alp_b/(2r)     (SSM:417)
  double* d64[4] = {rt.pinv, rt.pfwd, rt.alpa, rt.alpb};
  float* hi[4] = {rt.pinv_hi, rt.pfwd_hi, rt.alpa_hi, rt.alpb_hi};
  float* lo[4] = {rt.pinv_lo, rt.pfwd_lo, rt.alpa_lo, rt.alpb_lo};
  ushort4* b3[4] = {rt.pinv_b3, rt.pfwd_b3, rt.alpa_b3, rt.alpb_b3};
  for (int t = 0; t < 4; ++t) {
    d64[t][o] = vals[t];
    float h = (float)vals[t];
    hi[t][o] = h;
    lo[t][o] = (float)(vals[t] - (double)h);
    b3[t][o] = split_bf16x3(vals[t]);
  }
}

int ensure_scratch(jgcm_mesh* m, int nfields) {
  if (nfields <= m->cap_fields) return JGCM_OK;
  int cap = nfields;
  // Growing synchronises and allocates, which a capture cannot record: the
  // scratch must be reserved before capturing.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  JGCM_CHECK_HIP(hipStreamIsCapturing(m->stream, &cs));
  if (cs != hipStreamCaptureStatusNone || g_capture != nullptr) {
    set_error("mesh scratch must grow from %d to %d level-fields during a graph capture; "
              "reserve it before capturing (jgcm_mesh_reserve)", m->cap_fields, nfields);
    return JGCM_ERR_STATE;
  }
  double* ptrs[6] = {m->d_four_a, m->d_four_b, m->d_spec_a, m->d_spec_b, m->d_grid_a, m->d_grid_b};
  JGCM_CHECK_HIP(hipStreamSynchronize(m->stream));
  if (m->live_graphs.load() > 0) {
    // captured graphs keep these addresses as kernel arguments: keep the old
    // buffers alive until the mesh is freed
    for (double* p : ptrs)
      if (p) m->retired.push_back(p);
  } else {
    for (double* p : ptrs)
      if (p) JGCM_CHECK_HIP(hipFree(p));
  }
  size_t four = (size_t)cap * m->four_elems() * 2 * sizeof(double);
  size_t spec = (size_t)cap * m->spec_elems() * 2 * sizeof(double);
  size_t grid = (size_t)cap * m->grid_elems() * sizeof(double);
  JGCM_CHECK_HIP(hipMalloc(&m->d_four_a, four));
  JGCM_CHECK_HIP(hipMalloc(&m->d_four_b, four));
  JGCM_CHECK_HIP(hipMalloc(&m->d_spec_a, spec));
  JGCM_CHECK_HIP(hipMalloc(&m->d_spec_b, spec));
  JGCM_CHECK_HIP(hipMalloc(&m->d_grid_a, grid));
  JGCM_CHECK_HIP(hipMalloc(&m->d_grid_b, grid));
  m->cap_fields = cap;
  return JGCM_OK;
}

static int upload(double** d, const std::vector<double>& h) {
  JGCM_CHECK_HIP(hipMalloc(d, h.size() * sizeof(double)));
  JGCM_CHECK_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return JGCM_OK;
}

static void free_mesh(jgcm_mesh* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void* ptrs[] = {m->d_sin, m->d_cos, m->d_wts, m->d_qnm_full, m->d_dqnm_full, m->d_row_start,
                  m->rt.pinv, m->rt.pfwd, m->rt.alpa, m->rt.alpb,
                  m->rt.pinv_hi, m->rt.pinv_lo, m->rt.pfwd_hi, m->rt.pfwd_lo,
                  m->rt.alpa_hi, m->rt.alpa_lo, m->rt.alpb_hi, m->rt.alpb_lo,
                  m->rt.pinv_b3, m->rt.pfwd_b3, m->rt.alpa_b3, m->rt.alpb_b3,
                  m->d_twiddle, m->d_eig, m->d_uvm, m->d_uvc, m->d_uvp, m->d_dlam, m->d_dthm,
                  m->d_dthp, m->d_four_a, m->d_four_b, m->d_spec_a, m->d_spec_b, m->d_grid_a,
                  m->d_grid_b, m->d_scalar};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (void* p : m->retired) (void)hipFree(p);
  if (m->signal) (void)hipEventDestroy(m->signal);
  if (m->fork_ev) (void)hipEventDestroy(m->fork_ev);
  if (m->join_ev) (void)hipEventDestroy(m->join_ev);
  if (m->side) (void)hipStreamDestroy(m->side);
  if (m->on_side) m->stream = m->main_stream;
  if (m->stream && !m->stream_borrowed) (void)hipStreamDestroy(m->stream);
  jgcm_mesh* owner = m->stream_owner;
  delete m;
  mesh_release(owner);
}

void mesh_release(jgcm_mesh* m) {
  if (m && m->refs.fetch_sub(1) == 1) free_mesh(m);
}

static int create_impl(int N, int ns, int nlat, int nlon, int nd, double radius, jgcm_mesh* m) {
  m->N = N;
  m->ns = ns;
  m->nlat = nlat;
  m->nlon = nlon;
  m->nd = nd;
  m->radius = radius;
  m->half = nlat / 2;
  m->nm = N + 1;
  JGCM_CHECK_HIP(hipGetDevice(&m->device));
  JGCM_CHECK_HIP(hipStreamCreate(&m->stream));
  JGCM_CHECK_HIP(hipEventCreateWithFlags(&m->signal, hipEventDisableTiming));
  JGCM_CHECK_HIP(hipStreamCreate(&m->side));
  JGCM_CHECK_HIP(hipEventCreateWithFlags(&m->fork_ev, hipEventDisableTiming));
  JGCM_CHECK_HIP(hipEventCreateWithFlags(&m->join_ev, hipEventDisableTiming));
  hipStream_t st = m->stream;

  // ---- Gaussian nodes and weights (FP64, GPU) ----
  int* d_nonconv = nullptr;
  JGCM_CHECK_HIP(hipMalloc(&m->d_sin, nlat * sizeof(double)));
  JGCM_CHECK_HIP(hipMalloc(&m->d_wts, nlat * sizeof(double)));
  JGCM_CHECK_HIP(hipMalloc(&d_nonconv, sizeof(int)));
  JGCM_CHECK_HIP(hipMemsetAsync(d_nonconv, 0, sizeof(int), st));
  int half = nlat / 2;
  hipLaunchKernelGGL(gauss_nodes_kernel, dim3((half + 63) / 64), dim3(64), 0, st, nlat, m->d_sin,
                     m->d_wts, d_nonconv);
  JGCM_LAUNCH_CHECK();
  int nonconv = 0;
  JGCM_CHECK_HIP(hipMemcpyAsync(&nonconv, d_nonconv, sizeof(int), hipMemcpyDeviceToHost, st));
  m->sin_theta.resize(nlat);
  m->wts.resize(nlat);
  JGCM_CHECK_HIP(hipMemcpyAsync(m->sin_theta.data(), m->d_sin, nlat * sizeof(double),
                                hipMemcpyDeviceToHost, st));
  JGCM_CHECK_HIP(hipMemcpyAsync(m->wts.data(), m->d_wts, nlat * sizeof(double),
                                hipMemcpyDeviceToHost, st));
  JGCM_CHECK_HIP(hipStreamSynchronize(st));
  JGCM_CHECK_HIP(hipFree(d_nonconv));
  if (nonconv)  // the reference logs @error and continues (Gauss_And_Legendre.jl:108-110)
    fprintf(stderr, "jgcm: Compute_Gaussian! does not converge (%d nodes)\n", nonconv);

  // ---- host-side public fields ----
  m->cos_theta.resize(nlat);
  m->theta_c.resize(nlat);
  for (int j = 0; j < nlat; ++j) {
    m->cos_theta[j] = std::sqrt(1.0 - m->sin_theta[j] * m->sin_theta[j]);
    m->theta_c[j] = std::asin(m->sin_theta[j]);
  }
  m->lambda_c.resize(nlon);
  for (int i = 0; i < nlon; ++i) m->lambda_c[i] = ((double)i / nlon) * (2.0 * M_PI);
  double dlam = 2.0 * M_PI / nlon;
  m->lambda_e.resize(nlon + 1);
  for (int i = 0; i <= nlon; ++i) m->lambda_e[i] = (-0.5 + i) * dlam;
  m->theta_e.assign(nlat + 1, 0.0);
  m->theta_e[0] = -0.5 * M_PI;
  double acc = 0.0;
  for (int j = 0; j < nlat; ++j) {
    acc = acc + m->wts[j];
    m->theta_e[j + 1] = std::asin(acc - 1.0);
  }
  m->theta_e[nlat] = 0.5 * M_PI;
  if (upload(&m->d_cos, m->cos_theta)) return JGCM_ERR_HIP;

  // ---- Legendre recurrences (FP64, GPU) ----
  size_t nq_full = (size_t)(N + 1) * (ns + 2) * nlat;
  size_t ndq_full = (size_t)(N + 1) * (ns + 1) * nlat;
  JGCM_CHECK_HIP(hipMalloc(&m->d_qnm_full, nq_full * sizeof(double)));
  JGCM_CHECK_HIP(hipMalloc(&m->d_dqnm_full, ndq_full * sizeof(double)));
  hipLaunchKernelGGL(legendre_kernel, dim3((nlat + 63) / 64, N + 1), dim3(64), 0, st, N, ns, nlat,
                     m->d_sin, m->d_qnm_full, m->d_dqnm_full);
  JGCM_LAUNCH_CHECK();

  // ---- packed rows ----
  m->row_start_h.resize(N + 2);
  int rs = 0;
  for (int mm = 0; mm <= N; ++mm) {
    m->row_start_h[mm] = rs;
    rs += ns + 1 - mm;
  }
  m->row_start_h[N + 1] = rs;
  m->nrows = rs;
  JGCM_CHECK_HIP(hipMalloc(&m->d_row_start, (N + 2) * sizeof(int)));
  JGCM_CHECK_HIP(hipMemcpy(m->d_row_start, m->row_start_h.data(), (N + 2) * sizeof(int),
                           hipMemcpyHostToDevice));
  size_t tab = (size_t)rs * half;
  double** d64[4] = {&m->rt.pinv, &m->rt.pfwd, &m->rt.alpa, &m->rt.alpb};
  float** f32[8] = {&m->rt.pinv_hi, &m->rt.pinv_lo, &m->rt.pfwd_hi, &m->rt.pfwd_lo,
                    &m->rt.alpa_hi, &m->rt.alpa_lo, &m->rt.alpb_hi, &m->rt.alpb_lo};
  for (auto p : d64) JGCM_CHECK_HIP(hipMalloc(p, tab * sizeof(double)));
  for (auto p : f32) JGCM_CHECK_HIP(hipMalloc(p, tab * sizeof(float)));
  ushort4** b16[4] = {&m->rt.pinv_b3, &m->rt.pfwd_b3, &m->rt.alpa_b3, &m->rt.alpb_b3};
  for (auto p : b16) JGCM_CHECK_HIP(hipMalloc(p, tab * sizeof(ushort4)));
  hipLaunchKernelGGL(pack_rows_kernel, dim3((half + 63) / 64, ns + 1, N + 1), dim3(64), 0, st, N,
                     ns, nlat, radius, m->d_row_start, m->d_qnm_full, m->d_dqnm_full, m->d_wts,
                     m->d_cos, m->rt);
  JGCM_LAUNCH_CHECK();

  // host copies of qnm (degrees 0..ns) and dqnm
  std::vector<double> qfull(nq_full);
  m->dqnm.resize(ndq_full);
  JGCM_CHECK_HIP(hipMemcpyAsync(qfull.data(), m->d_qnm_full, nq_full * sizeof(double),
                                hipMemcpyDeviceToHost, st));
  JGCM_CHECK_HIP(hipMemcpyAsync(m->dqnm.data(), m->d_dqnm_full, ndq_full * sizeof(double),
                                hipMemcpyDeviceToHost, st));
  JGCM_CHECK_HIP(hipStreamSynchronize(st));
  m->qnm.resize(ndq_full);
  for (int mm = 0; mm <= N; ++mm)
    for (int n = 0; n <= ns; ++n)
      std::memcpy(&m->qnm[((size_t)mm * (ns + 1) + n) * nlat],
                  &qfull[((size_t)mm * (ns + 2) + n) * nlat], nlat * sizeof(double));
  JGCM_CHECK_HIP(hipFree(m->d_qnm_full));
  m->d_qnm_full = nullptr;
  JGCM_CHECK_HIP(hipFree(m->d_dqnm_full));
  m->d_dqnm_full = nullptr;

  // ---- [m][n] coefficient tables (SSM:101-114 and the *_Init functions) ----
  size_t ne = (size_t)(N + 1) * (ns + 1);
  m->epsilon.assign(ne, 0.0);
  m->laplacian_eig.assign(ne, 0.0);
  m->coef_uvm.assign(ne, 0.0);
  m->coef_uvc.assign(ne, 0.0);
  m->coef_uvp.assign(ne, 0.0);
  m->coef_dlam.assign(ne, 0.0);
  m->coef_dthm.assign(ne, 0.0);
  m->coef_dthp.assign(ne, 0.0);
  auto at = [ns](int mm, int n) { return (size_t)mm * (ns + 1) + n; };
  for (int mm = 0; mm <= N; ++mm)
    for (int n = mm; n <= ns; ++n)
      m->epsilon[at(mm, n)] = std::sqrt((double)(n * n - mm * mm) / (4.0 * n * n - 1.0));
  for (int mm = 0; mm <= N; ++mm) {
    for (int n = mm; n <= ns; ++n) {
      if (n > 0) {
        m->coef_uvc[at(mm, n)] = -radius * mm / (n * (n + 1));          // SSM:497
        m->coef_uvm[at(mm, n)] = radius * m->epsilon[at(mm, n)] / n;    // SSM:504
      }
      m->coef_dlam[at(mm, n)] = mm / radius;                            // SSM:640
      m->coef_dthm[at(mm, n)] = -(n - 1) * m->epsilon[at(mm, n)] / radius;  // SSM:641
      m->laplacian_eig[at(mm, n)] = -n * (n + 1) / (radius * radius);   // SSM:607
    }
    for (int n = mm; n <= ns - 1; ++n) {
      m->coef_uvp[at(mm, n)] = -radius * m->epsilon[at(mm, n + 1)] / (n + 1);  // SSM:510
      m->coef_dthp[at(mm, n)] = (n + 2) * m->epsilon[at(mm, n + 1)] / radius;  // SSM:647
    }
  }
  if (upload(&m->d_eig, m->laplacian_eig) || upload(&m->d_uvm, m->coef_uvm) ||
      upload(&m->d_uvc, m->coef_uvc) || upload(&m->d_uvp, m->coef_uvp) ||
      upload(&m->d_dlam, m->coef_dlam) || upload(&m->d_dthm, m->coef_dthm) ||
      upload(&m->d_dthp, m->coef_dthp))
    return JGCM_ERR_HIP;

  // ---- FFT twiddles exp(-2 pi i k / nλ), k = 0..nλ-1 ----
  std::vector<double> tw(2 * (size_t)nlon);
  for (int k = 0; k < nlon; ++k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * k / nlon;
    tw[2 * k] = (double)cosl(a);
    tw[2 * k + 1] = (double)sinl(a);
  }
  if (upload(&m->d_twiddle, tw)) return JGCM_ERR_HIP;
  JGCM_CHECK_HIP(hipMalloc(&m->d_scalar, 64 * sizeof(double)));
  JGCM_CHECK_HIP(hipStreamSynchronize(st));
  return ensure_scratch(m, nd > 0 ? nd : 1);
}

}  // namespace jgcm

using namespace jgcm;

extern "C" {

const char* jgcm_last_error(void) { return g_err; }

int jgcm_device_count(int* count) {
  JGCM_REQUIRE(count != nullptr, "count is NULL");
  JGCM_CHECK_HIP(hipGetDeviceCount(count));
  return JGCM_OK;
}

int jgcm_mesh_create(int num_fourier, int num_spherical, int nlat, int nlon, int nd,
                     double radius, jgcm_mesh_t** out) {
  JGCM_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  JGCM_REQUIRE(num_fourier >= 0 && num_spherical >= num_fourier, "need num_spherical >= num_fourier >= 0");
  JGCM_REQUIRE(nlat >= 2 && nlat % 2 == 0, "nθ must be even (Spectral_Spherical_Mesh.jl:193,275), got %d", nlat);
  JGCM_REQUIRE(nd >= 1, "nd must be >= 1");
  JGCM_REQUIRE(radius > 0.0, "radius must be positive");
  // powers of two 16..1024 take the radix-4 LDS FFT; other even lengths up to
  // 2048 the direct DFT kernels
  bool pow2 = nlon >= 16 && nlon <= 1024 && (nlon & (nlon - 1)) == 0;
  if (!pow2 && (nlon < 16 || nlon > 2048 || nlon % 2 != 0)) {
    set_error("nλ=%d: the longitude transform supports even lengths 16..2048", nlon);
    return JGCM_ERR_UNSUPPORTED;
  }
  if (2 * num_fourier + 1 > nlon) {
    set_error("need nλ >= 2 num_fourier + 1 (Spectral_Spherical_Mesh.jl:231), got N=%d nλ=%d",
              num_fourier, nlon);
    return JGCM_ERR_UNSUPPORTED;
  }
  if (nlat / 2 > 256) {
    set_error("nθ=%d exceeds the supported maximum of 512", nlat);
    return JGCM_ERR_UNSUPPORTED;
  }
  // LDS budget of the FFT stage: 16 KB work buffer + twiddles + the E/O
  // staging of 8 latitude pairs (or 256/(nλ/4) when larger) x (N+1) x 2.
  if (!pow2) {
    // direct DFT: twiddles + two rows (fold) or twiddles + 2 (N+1) coefficients (unfold)
    size_t lds = (size_t)nlon * 16 + std::max((size_t)nlon * 16, (size_t)(num_fourier + 1) * 32);
    if (lds > 65536) {
      set_error("T%d on nλ=%d exceeds the longitude DFT's 64 KB LDS budget", num_fourier, nlon);
      return JGCM_ERR_UNSUPPORTED;
    }
  } else {
    int fpw = 256 / (nlon / 4);
    int ppw = fpw >= 8 ? fpw : 8;
    size_t lds = 16384 + (size_t)nlon * 16 + (size_t)((num_fourier + 1) * 2 + 1) * ppw * 16;
    if (lds > 65536) {
      set_error("T%d on nλ=%d exceeds the FFT stage's 64 KB LDS budget", num_fourier, nlon);
      return JGCM_ERR_UNSUPPORTED;
    }
  }
  jgcm_mesh* m = new (std::nothrow) jgcm_mesh();
  if (!m) {
    set_error("out of host memory");
    return JGCM_ERR_NOMEM;
  }
  int rc = create_impl(num_fourier, num_spherical, nlat, nlon, nd, radius, m);
  if (rc != JGCM_OK) {
    free_mesh(m);
    return rc;
  }
  *out = m;
  return JGCM_OK;
}

int jgcm_mesh_destroy(jgcm_mesh_t* mesh) {
  mesh_release(mesh);   // the creator's reference
  return JGCM_OK;
}

int jgcm_mesh_lifetime(jgcm_mesh_t* mesh, int* refs, int* live_graphs) {
  if (int rc = check_mesh(mesh)) return rc;
  if (refs) *refs = mesh->refs.load();
  if (live_graphs) *live_graphs = mesh->live_graphs.load();
  return JGCM_OK;
}

int jgcm_mesh_synchronize(jgcm_mesh_t* mesh) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_CHECK_HIP(hipStreamSynchronize(mesh->stream));
  if (mesh->side) JGCM_CHECK_HIP(hipStreamSynchronize(mesh->side));
  return JGCM_OK;
}

int jgcm_mesh_branch(jgcm_mesh_t* mesh, int side) {
  if (int rc = check_mesh(mesh)) return rc;
  if (side) {
    JGCM_REQUIRE(!mesh->on_side, "the mesh already works on its side stream");
    JGCM_CHECK_HIP(hipEventRecord(mesh->fork_ev, mesh->stream));
    JGCM_CHECK_HIP(hipStreamWaitEvent(mesh->side, mesh->fork_ev, 0));
    mesh->main_stream = mesh->stream;
    mesh->stream = mesh->side;
    mesh->on_side = true;
    mesh->forked = true;
  } else if (mesh->on_side) {
    mesh->stream = mesh->main_stream;
    mesh->on_side = false;
  }
  return JGCM_OK;
}

int jgcm_mesh_join(jgcm_mesh_t* mesh) {
  if (int rc = check_mesh(mesh)) return rc;
  if (mesh->on_side) {
    mesh->stream = mesh->main_stream;
    mesh->on_side = false;
  }
  if (mesh->forked) {
    JGCM_CHECK_HIP(hipEventRecord(mesh->join_ev, mesh->side));
    JGCM_CHECK_HIP(hipStreamWaitEvent(mesh->stream, mesh->join_ev, 0));
    mesh->forked = false;
  }
  return JGCM_OK;
}

void* jgcm_mesh_stream(jgcm_mesh_t* mesh) { return mesh ? (void*)mesh->stream : nullptr; }

int jgcm_mesh_share_stream(jgcm_mesh_t* mesh, jgcm_mesh_t* owner) {
  if (int rc = check_mesh(mesh)) return rc;
  if (int rc = check_mesh(owner)) return rc;
  JGCM_REQUIRE(mesh != owner, "a mesh cannot share its own stream");
  JGCM_REQUIRE(!mesh->stream_borrowed, "mesh already uses another mesh's stream");
  JGCM_CHECK_HIP(hipStreamSynchronize(mesh->stream));
  JGCM_CHECK_HIP(hipStreamSynchronize(owner->stream));
  JGCM_CHECK_HIP(hipStreamDestroy(mesh->stream));
  mesh->stream = owner->stream;
  mesh->stream_borrowed = true;
  mesh->stream_owner = owner;   // the owner (and its stream) lives while this mesh does
  mesh_retain(owner);
  return JGCM_OK;
}

int jgcm_mesh_set_precision(jgcm_mesh_t* mesh, int precision) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(precision == JGCM_PREC_FP64 || precision == JGCM_PREC_FP32 ||
                   precision == JGCM_PREC_SPLIT32 || precision == JGCM_PREC_BF16X3,
               "unknown precision %d", precision);
  mesh->precision = precision;
  return JGCM_OK;
}

int jgcm_mesh_set_fft_shape(jgcm_mesh_t* mesh, int fold_ppw, int unfold_ppw) {
  if (int rc = check_mesh(mesh)) return rc;
  const int limit = fft_max_ppw(mesh);
  JGCM_REQUIRE(fold_ppw >= 0 && fold_ppw <= limit, "fold_ppw must be in 0..%d, got %d", limit,
               fold_ppw);
  JGCM_REQUIRE(unfold_ppw >= 0 && unfold_ppw <= limit, "unfold_ppw must be in 0..%d, got %d",
               limit, unfold_ppw);
  mesh->fold_ppw = fold_ppw;
  mesh->unfold_ppw = unfold_ppw;
  return JGCM_OK;
}

int jgcm_mesh_set_legendre_shape(jgcm_mesh_t* mesh, int fwd_tiles, int inv_tiles) {
  if (int rc = check_mesh(mesh)) return rc;
  auto ok = [](int t) { return t == 0 || t == 1 || t == 2 || t == 4; };
  JGCM_REQUIRE(ok(fwd_tiles) && ok(inv_tiles),
               "Legendre tiles per workgroup must be 0, 1, 2 or 4, got %d / %d", fwd_tiles,
               inv_tiles);
  mesh->leg_tiles_fwd = fwd_tiles;
  mesh->leg_tiles_inv = inv_tiles;
  return JGCM_OK;
}

int jgcm_mesh_get_legendre_shape(jgcm_mesh_t* mesh, int* fwd_tiles, int* inv_tiles) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(fwd_tiles && inv_tiles, "NULL argument");
  *fwd_tiles = mesh->leg_tiles_fwd;
  *inv_tiles = mesh->leg_tiles_inv;
  return JGCM_OK;
}

int jgcm_mesh_get_fft_shape(jgcm_mesh_t* mesh, int* fold_ppw, int* unfold_ppw, int* max_ppw) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(fold_ppw && unfold_ppw && max_ppw, "NULL argument");
  *fold_ppw = mesh->fold_ppw;
  *unfold_ppw = mesh->unfold_ppw;
  *max_ppw = fft_max_ppw(mesh);
  return JGCM_OK;
}

int jgcm_mesh_get_precision(jgcm_mesh_t* mesh, int* precision) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(precision != nullptr, "precision is NULL");
  *precision = mesh->precision;
  return JGCM_OK;
}

static const std::vector<double>* table_of(jgcm_mesh* m, int which, std::vector<double>& tmp) {
  switch (which) {
    case JGCM_TABLE_SIN_THETA: return &m->sin_theta;
    case JGCM_TABLE_COS_THETA: return &m->cos_theta;
    case JGCM_TABLE_WTS: return &m->wts;
    case JGCM_TABLE_LAMBDA_C: return &m->lambda_c;
    case JGCM_TABLE_THETA_C: return &m->theta_c;
    case JGCM_TABLE_LAMBDA_E: return &m->lambda_e;
    case JGCM_TABLE_THETA_E: return &m->theta_e;
    case JGCM_TABLE_QNM: return &m->qnm;
    case JGCM_TABLE_DQNM: return &m->dqnm;
    case JGCM_TABLE_EPSILON: return &m->epsilon;
    case JGCM_TABLE_LAPLACIAN_EIG: return &m->laplacian_eig;
    case JGCM_TABLE_COEF_UVM: return &m->coef_uvm;
    case JGCM_TABLE_COEF_UVC: return &m->coef_uvc;
    case JGCM_TABLE_COEF_UVP: return &m->coef_uvp;
    case JGCM_TABLE_COEF_DLAMBDA: return &m->coef_dlam;
    case JGCM_TABLE_COEF_DTHETAM: return &m->coef_dthm;
    case JGCM_TABLE_COEF_DTHETAP: return &m->coef_dthp;
    case JGCM_TABLE_QWG:
    case JGCM_TABLE_COEF_ALP_A:
    case JGCM_TABLE_COEF_ALP_B: {
      int N = m->N, ns = m->ns, nlat = m->nlat;
      tmp.assign((size_t)(N + 1) * (ns + 1) * nlat, 0.0);
      for (int mm = 0; mm <= N; ++mm)
        for (int n = mm; n <= ns; ++n)
          for (int j = 0; j < nlat; ++j) {
            size_t o = ((size_t)mm * (ns + 1) + n) * nlat + j;
            double w = m->wts[j], q = m->qnm[o], c = m->cos_theta[j];
            if (which == JGCM_TABLE_QWG) tmp[o] = q * w;                       // SSM:98
            else if (which == JGCM_TABLE_COEF_ALP_A) tmp[o] = w * q / (c * c) * mm;  // SSM:416
            else tmp[o] = -w * m->dqnm[o];                                     // SSM:417
          }
      return &tmp;
    }
    default: return nullptr;
  }
}

int jgcm_mesh_table_size(jgcm_mesh_t* mesh, int which, size_t* count) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(count != nullptr, "count is NULL");
  size_t full = (size_t)(mesh->N + 1) * (mesh->ns + 1);
  switch (which) {
    case JGCM_TABLE_SIN_THETA: case JGCM_TABLE_COS_THETA: case JGCM_TABLE_WTS:
    case JGCM_TABLE_THETA_C: *count = mesh->nlat; return JGCM_OK;
    case JGCM_TABLE_LAMBDA_C: *count = mesh->nlon; return JGCM_OK;
    case JGCM_TABLE_LAMBDA_E: *count = mesh->nlon + 1; return JGCM_OK;
    case JGCM_TABLE_THETA_E: *count = mesh->nlat + 1; return JGCM_OK;
    case JGCM_TABLE_QNM: case JGCM_TABLE_DQNM: case JGCM_TABLE_QWG:
    case JGCM_TABLE_COEF_ALP_A: case JGCM_TABLE_COEF_ALP_B:
      *count = full * mesh->nlat; return JGCM_OK;
    case JGCM_TABLE_EPSILON: case JGCM_TABLE_LAPLACIAN_EIG: case JGCM_TABLE_COEF_UVM:
    case JGCM_TABLE_COEF_UVC: case JGCM_TABLE_COEF_UVP: case JGCM_TABLE_COEF_DLAMBDA:
    case JGCM_TABLE_COEF_DTHETAM: case JGCM_TABLE_COEF_DTHETAP:
      *count = full; return JGCM_OK;
    default:
      set_error("unknown table id %d", which);
      return JGCM_ERR_ARG;
  }
}

int jgcm_mesh_get_table(jgcm_mesh_t* mesh, int which, double* host_out, size_t count) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(host_out != nullptr, "host_out is NULL");
  size_t need = 0;
  if (int rc = jgcm_mesh_table_size(mesh, which, &need)) return rc;
  JGCM_REQUIRE(count >= need, "table %d needs %zu doubles, got %zu", which, need, count);
  std::vector<double> tmp;
  const std::vector<double>* t = table_of(mesh, which, tmp);
  JGCM_REQUIRE(t != nullptr, "unknown table id %d", which);
  std::memcpy(host_out, t->data(), need * sizeof(double));
  return JGCM_OK;
}

int jgcm_malloc(void** ptr, size_t bytes) {
  JGCM_REQUIRE(ptr != nullptr, "ptr is NULL");
  JGCM_CHECK_HIP(hipMalloc(ptr, bytes));
  return JGCM_OK;
}

int jgcm_free(void* ptr) {
  if (ptr) JGCM_CHECK_HIP(hipFree(ptr));
  return JGCM_OK;
}

int jgcm_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  JGCM_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return JGCM_OK;
}

int jgcm_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  JGCM_CHECK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return JGCM_OK;
}

}  // extern "C"
