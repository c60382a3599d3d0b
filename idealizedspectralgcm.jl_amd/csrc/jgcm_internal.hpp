// Internal definitions shared by the libjgcm_hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/jgcm.h"

namespace jgcm {

void set_error(const char* fmt, ...);

#define JGCM_CHECK_HIP(expr)                                                     \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      ::jgcm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),  \
                        __FILE__, __LINE__);                                     \
      return JGCM_ERR_HIP;                                                       \
    }                                                                            \
  } while (0)

#define JGCM_REQUIRE(cond, ...)                                                  \
  do {                                                                           \
    if (!(cond)) {                                                               \
      ::jgcm::set_error(__VA_ARGS__);                                            \
      return JGCM_ERR_ARG;                                                       \
    }                                                                            \
  } while (0)

#define JGCM_LAUNCH_CHECK()                                                      \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess) {                                                      \
      ::jgcm::set_error("kernel launch failed: %s (%s:%d)",                     \
                        hipGetErrorString(_e), __FILE__, __LINE__);              \
      return JGCM_ERR_HIP;                                                       \
    }                                                                            \
  } while (0)

// Packed per-m row tables: for each zonal wavenumber m = 0..N the rows are
// the degrees n = m..ns (ns = num_spherical), each row holding `half` (=nθ/2)
// values on the southern latitudes j' = 0..half-1 (the northern values follow
// from the equatorial symmetry P(-mu) = (-1)^(n-m) P(mu)).
struct RowTables {
  // FP64
  double* pinv = nullptr;   // qnm                        (inverse transform)
  double* pfwd = nullptr;   // qnm * w / 2                (forward transform)
  double* alpa = nullptr;   // w qnm m / cos^2 / (2 r)    (alpha operator, term a)
  double* alpb = nullptr;   // -w dqnm / (2 r)            (alpha operator, term b)
  // FP32 hi/lo split copies: x = hi + lo, hi = fl32(x), lo = fl32(x - hi)
  float* pinv_hi = nullptr; float* pinv_lo = nullptr;
  float* pfwd_hi = nullptr; float* pfwd_lo = nullptr;
  float* alpa_hi = nullptr; float* alpa_lo = nullptr;
  float* alpb_hi = nullptr; float* alpb_lo = nullptr;
  // BF16 three-piece split copies: x = b1 + b2 + b3 (x, y, z of a ushort4; w = 0)
  ushort4* pinv_b3 = nullptr;
  ushort4* pfwd_b3 = nullptr;
  ushort4* alpa_b3 = nullptr;
  ushort4* alpb_b3 = nullptr;
};

// Add_Horizontal_Advection!'s update δh - (u ∂λh + v ∂θh) (SSM:716) with its
// multiply-add written out, so that every kernel using it rounds alike.
__device__ __forceinline__ double advect(double dh, double u, double v, double gl, double gt) {
  return dh - __fma_rn(u, gl, v * gt);
}

// BF16 (round to nearest even) from FP32 and back; finite values only.
__host__ __device__ inline unsigned short f32_to_bf16(float x) {
  unsigned u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(unsigned short b) {
  unsigned u = (unsigned)b << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}
// x = b1 + b2 + b3 + O(2^-24 |x|): each piece the BF16 rounding of the rest.
__host__ __device__ inline ushort4 split_bf16x3(double v) {
  unsigned short b1 = f32_to_bf16((float)v);
  double r = v - (double)bf16_to_f32(b1);
  unsigned short b2 = f32_to_bf16((float)r);
  r = r - (double)bf16_to_f32(b2);
  unsigned short b3 = f32_to_bf16((float)r);
  return make_ushort4(b1, b2, b3, 0);
}

}  // namespace jgcm

struct jgcm_mesh {
  int N = 0;        // num_fourier
  int ns = 0;       // num_spherical
  int nlat = 0, nlon = 0, nd = 0, half = 0;
  int nm = 0;       // N + 1 zonal wavenumbers
  int nrows = 0;    // packed rows: sum_m (ns + 1 - m)
  double radius = 0.0;
  int precision = JGCM_PREC_FP64;
  // latitude pairs per FFT workgroup (0: the built-in rule), set by an autotuner
  int fold_ppw = 0, unfold_ppw = 0;
  // 16-row (forward) / 16-latitude (inverse) tiles per Legendre workgroup:
  // 1, 2 or 4 (0: the built-in 2), set by an autotuner
  int leg_tiles_fwd = 0, leg_tiles_inv = 0;
  hipStream_t stream = nullptr;
  bool stream_borrowed = false;   // stream owned by another mesh (jgcm_mesh_share_stream)
  hipEvent_t signal = nullptr;    // recorded on this mesh's stream by jgcm_mesh_wait
  // side stream for independent work of one step (jgcm_mesh_branch / jgcm_mesh_join)
  hipStream_t side = nullptr, main_stream = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool on_side = false, forked = false;
  int device = 0;

  // Lifetime.  The creator holds one reference (dropped by jgcm_mesh_destroy);
  // every leapfrog, vertical coordinate and graph that uses the mesh and every
  // mesh that borrows its stream holds one more, so the mesh (its stream,
  // events and buffers) is freed only after the last of them, whatever order
  // a garbage collector finalizes them in.
  std::atomic<int> refs{1};
  std::atomic<int> live_graphs{0};     // graphs and plans whose kernels use this mesh's scratch
  std::atomic<int> live_plans{0};      // of which plans (jgcm_plan_*_create)
  jgcm_mesh* stream_owner = nullptr;   // the mesh whose stream this one borrows
  std::vector<void*> retired;          // scratch replaced while graphs still held it

  // host copies of public fields
  std::vector<double> sin_theta, cos_theta, wts, lambda_c, theta_c, lambda_e, theta_e;
  std::vector<double> qnm, dqnm;               // [N+1][ns+1][nlat]
  std::vector<double> epsilon, laplacian_eig;  // [N+1][ns+1]
  std::vector<double> coef_uvm, coef_uvc, coef_uvp, coef_dlam, coef_dthm, coef_dthp;
  std::vector<int> row_start_h;                // [N+2]

  // device tables
  double* d_sin = nullptr;
  double* d_cos = nullptr;
  double* d_wts = nullptr;
  double* d_qnm_full = nullptr;   // [N+1][ns+2][nlat]   (recurrence workspace)
  double* d_dqnm_full = nullptr;  // [N+1][ns+1][nlat]
  int* d_row_start = nullptr;     // [N+2]
  jgcm::RowTables rt;
  double* d_twiddle = nullptr;    // nlon complex: exp(-2 pi i k / nlon)
  // [N+1][ns+1] spectral coefficient tables (reference layout [m][n] read as
  // row-major here, i.e. index m*(ns+1)+n)
  double* d_eig = nullptr;
  double* d_uvm = nullptr;
  double* d_uvc = nullptr;
  double* d_uvp = nullptr;
  double* d_dlam = nullptr;
  double* d_dthm = nullptr;
  double* d_dthp = nullptr;

  // scratch (grown on demand, capacity in level-fields)
  int cap_fields = 0;
  double* d_four_a = nullptr;   // E/O Fourier block [B][nm][2][half] complex
  double* d_four_b = nullptr;
  double* d_spec_a = nullptr;   // spectral scratch [B][ns+1][N+1] complex
  double* d_spec_b = nullptr;
  double* d_grid_a = nullptr;   // grid scratch [B][nlat][nlon]
  double* d_grid_b = nullptr;
  double* d_scalar = nullptr;   // small device scalars

  size_t spec_elems() const { return (size_t)(ns + 1) * (N + 1); }
  size_t grid_elems() const { return (size_t)nlat * nlon; }
  size_t four_elems() const { return (size_t)nm * 2 * half; }  // complex elements
};

// Filtered_Leapfrog state (Time_Integrator.jl:2-18); the damping table lives
// on the device.
struct jgcm_leapfrog {
  jgcm_mesh* mesh = nullptr;
  double robert_coef = 0.0;
  int damping_order = 0;
  double damping_coef = 0.0;
  double implicit_coef = 0.0;
  int64_t dt = 0;
  bool init_step = true;
  int64_t time = 0, start_time = 0, end_time = 0;
  double* d_damping = nullptr;   // [N+1][ns+1]
  std::vector<double> damping;
};

namespace jgcm {
int ensure_scratch(jgcm_mesh* m, int nfields);
// Validates the handle; while this thread captures a graph it also records
// the mesh as one whose buffers the graph will use.
int check_mesh(const jgcm_mesh* m);
void mesh_retain(jgcm_mesh* m);
void mesh_release(jgcm_mesh* m);   // frees the mesh with its last reference
// Meshes used by the capture this thread has open (jgcm_graph_begin..end).
struct CaptureCtx {
  jgcm_mesh* origin = nullptr;
  std::vector<jgcm_mesh*> meshes;
};
CaptureCtx* capture_ctx();         // nullptr when this thread captures nothing
void capture_open(jgcm_mesh* origin);
void capture_close(std::vector<jgcm_mesh*>* used);
int fft_max_ppw(const jgcm_mesh* m);  // largest FFT latitude-pair count per workgroup
}  // namespace jgcm
