// Device accumulation of the Held-Suarez diagnostics (Output_Manager.jl:81-107):
// daily zonal means of T, T_eq, u, v and the daily mean of ps, summed on the
// GPU every step so the drivers need no per-step device-to-host copy.
#include "jgcm_internal.hpp"

namespace jgcm {

// One wavefront per (field, level, latitude) row: the row mean over the nλ
// longitudes (Zonal_Mean, Output_Manager.jl:276-280) is added to acc[row].
// Rows 0..4*nd*nlat-1 cover T, T_eq, u, v; the ps plane is added after.
__global__ __launch_bounds__(256) void zonal_accumulate_kernel(
    const double* __restrict__ t, const double* __restrict__ teq, const double* __restrict__ u,
    const double* __restrict__ v, const double* __restrict__ ps, int nd, int nlat, int nlon,
    double* __restrict__ acc_t, double* __restrict__ acc_teq, double* __restrict__ acc_u,
    double* __restrict__ acc_v, double* __restrict__ acc_ps) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int rows = nd * nlat;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int r = wave; r < 4 * rows; r += nwaves) {
    int f = r / rows;
    int row = r % rows;
    const double* src = f == 0 ? t : (f == 1 ? teq : (f == 2 ? u : v));
    double* dst = f == 0 ? acc_t : (f == 1 ? acc_teq : (f == 2 ? acc_u : acc_v));
    const double* p = src + (size_t)row * nlon;
    double s = 0.0;
    for (int i = lane; i < nlon; i += 64) s += p[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) dst[row] = dst[row] + s / nlon;
  }
  size_t P = (size_t)nlat * nlon;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < P; i += (size_t)gridDim.x * blockDim.x)
    acc_ps[i] = acc_ps[i] + ps[i];
}

}  // namespace jgcm

using namespace jgcm;

extern "C" {

// Update_Output!(output_manager, dyn_data, current_time)  Output_Manager.jl:81-107
// (the day index is chosen by the caller; acc_* point at that day's slices:
// acc_t/teq/u/v [nd][nθ], acc_ps [nθ][nλ])
int jgcm_zonal_mean_accumulate(jgcm_mesh_t* mesh, const double* grid_t, const double* grid_t_eq,
                               const double* grid_u, const double* grid_v, const double* grid_ps,
                               int nd, double* acc_t, double* acc_t_eq, double* acc_u,
                               double* acc_v, double* acc_ps) {
  if (int rc = check_mesh(mesh)) return rc;
  JGCM_REQUIRE(grid_t && grid_t_eq && grid_u && grid_v && grid_ps && acc_t && acc_t_eq && acc_u &&
                   acc_v && acc_ps && nd >= 1,
               "bad argument");
  int rows = 4 * nd * mesh->nlat;
  int blocks = (rows * 64 + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zonal_accumulate_kernel, dim3(blocks), dim3(256), 0, mesh->stream, grid_t,
                     grid_t_eq, grid_u, grid_v, grid_ps, nd, mesh->nlat, mesh->nlon, acc_t,
                     acc_t_eq, acc_u, acc_v, acc_ps);
  JGCM_LAUNCH_CHECK();
  return JGCM_OK;
}

}  // extern "C"
