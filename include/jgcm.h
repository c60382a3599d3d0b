/*
 * jgcm.h -- C ABI of libjgcm_hip.so, the MI355X (gfx950) spectral-transform
 * core of JGCM (CliMA/IdealizedSpectralGCM.jl, src/Atmos_Spectral).
 *
 * Every entry point mirrors one reference Julia function; the comment above
 * each declaration names it (file:line under the reference tree).  Argument
 * order and in-place semantics follow the reference: trailing pointer
 * arguments are outputs, inputs are never modified unless the reference
 * modifies them.
 *
 * Memory: every field pointer is a DEVICE pointer (hipMalloc'd, or a torch /
 * Julia GPU array), laid out exactly like the reference's Julia arrays:
 *   spectral  complex128 [nlev][num_spherical+1][num_fourier+1]  (m fastest,
 *             interleaved re/im, i.e. Julia ComplexF64[N+1, N+2, nlev])
 *   grid      float64    [nlev][nlat][nlon]  (longitude fastest, latitude
 *             index south -> north, Julia Float64[nλ, nθ, nlev])
 * `nlev` is the number of level-fields in the call.  The reference accepts
 * only nlev == nd or 1; this library accepts any nlev >= 1 (a batch), which is
 * how several prognostic fields go through one launch.
 *
 * Errors: every int-returning call returns JGCM_OK (0) or a negative code;
 * jgcm_last_error() gives a thread-local message.  Shape/argument failures
 * correspond to the reference's @assert sites, state failures to its error()
 * sites.
 *
 * Threading: each mesh owns one HIP stream and scratch buffers; calls on one
 * mesh are ordered on that stream and a mesh is not re-entrant (the reference
 * mesh shares scratch the same way, Spectral_Spherical_Mesh.jl:51-64).  The
 * stream is a blocking stream, so work on the legacy default stream is
 * ordered with it.
 */
#ifndef JGCM_H
#define JGCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JGCM_OK 0
#define JGCM_ERR_ARG (-1)     /* shape / argument assertion (reference @assert) */
#define JGCM_ERR_HIP (-2)     /* HIP runtime failure                            */
#define JGCM_ERR_STATE (-3)   /* invalid state (reference error())              */
#define JGCM_ERR_NOMEM (-4)
#define JGCM_ERR_UNSUPPORTED (-5)

/* Legendre-contraction precision (the FFT stage and all I/O stay FP64). */
#define JGCM_PREC_FP64 0      /* FP64 tables, FP64 accumulation                 */
#define JGCM_PREC_FP32 1      /* FP32 tables and operands, FP32 accumulation    */
#define JGCM_PREC_SPLIT32 2   /* hi/lo FP32 split, 3 products, compensated sums */
#define JGCM_PREC_BF16X3 3    /* 3-piece BF16 split, 6 products on BF16 MFMA,
                                 FP32 block sums, FP64 accumulation            */

typedef struct jgcm_mesh jgcm_mesh_t;
typedef struct jgcm_leapfrog jgcm_leapfrog_t;
typedef struct jgcm_graph jgcm_graph_t;
typedef struct jgcm_plan jgcm_plan_t;
typedef struct jgcm_vert_coord jgcm_vert_coord_t;
typedef struct jgcm_semi_implicit jgcm_semi_implicit_t;

/* Atmo_Data scalars (Atmo_Data.jl:4-31); cp_air = rdgas / kappa (:46). */
typedef struct jgcm_atmo_params {
  double radius, omega, grav, rdgas, kappa, cp_air;
  int do_mass_correction, do_energy_correction;
} jgcm_atmo_params;

/* Held-Suarez physics_params (exp/HSt42/Run_HS.jl:12, HS_Forcing.jl:8-13);
 * rates in day^-1, temperatures in K. */
typedef struct jgcm_hs_params {
  double sigma_b, k_f, k_a, k_s, delta_T_y, delta_theta_z;
} jgcm_hs_params;

/* Vert_Coordinate options (Vert_Coordinate.jl:61-87, Press_And_Geopot.jl:63,
 * Vert_Coordinate.jl:195-203).  UNEVEN_SIGMA and HYBRID are named so callers
 * can pass the reference's strings; creation rejects them (see below). */
#define JGCM_VC_EVEN_SIGMA 0
#define JGCM_VC_UNEVEN_SIGMA 1
#define JGCM_VC_HYBRID 2
#define JGCM_VC_MCM 3
#define JGCM_VC_V197 4
#define JGCM_VD_SIMMONS_BURRIDGE 0
#define JGCM_VA_SECOND_CENTERED_WTS 0
#define JGCM_VA_SECOND_CENTERED 1

const char* jgcm_last_error(void);
int jgcm_device_count(int* count);

/* ---- device memory helpers (for callers without their own GPU arrays) ---- */
int jgcm_malloc(void** ptr, size_t bytes);
int jgcm_free(void* ptr);
int jgcm_memcpy_h2d(void* dst, const void* src, size_t bytes);
int jgcm_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Ordered device-to-device copy on the mesh stream (a kernel, so it can be
 * graph-captured).  Used for Time_Advance! (Dyn_Data.jl:326-364).          */
int jgcm_copy(jgcm_mesh_t* mesh, void* dst, const void* src, size_t bytes);
/* n (<= 32) such copies in one launch (no reference counterpart; used to stage
 * a reverse step's inputs into the buffers its captured graph reads).
 * Sizes and addresses must be multiples of 8 bytes.                        */
int jgcm_copy_batch(jgcm_mesh_t* mesh, int n, double* const* dst, const double* const* src,
                    const size_t* bytes);
/* dst[k][i] += src[k][i], i < count[k] doubles, for n (<= 32) pairs in one
 * launch (no reference counterpart; accumulates a reverse step's
 * cotangents).  The destinations must not overlap.                         */
int jgcm_add_batch(jgcm_mesh_t* mesh, int n, double* const* dst, const double* const* src,
                   const size_t* count);
int jgcm_fill(jgcm_mesh_t* mesh, double* dst, double value, size_t count);
/* Time_Advance! (Dyn_Data.jl:326-364) for nfields (<= 16) prognostic fields
 * in one launch: per element, p = c, then c = n.  count[f] = doubles in
 * field f (complex fields count twice).                                    */
int jgcm_time_advance(jgcm_mesh_t* mesh, int nfields, double* const* p, double* const* c,
                      const double* const* n, const size_t* count);

/* ---- mesh: Spectral_Spherical_Mesh(num_fourier, num_spherical, nθ, nλ, nd,
 *      radius)  Spectral_Spherical_Mesh.jl:69-139.  Gaussian nodes/weights
 *      (Gauss_And_Legendre.jl:69-118) and the Legendre recurrences
 *      (Gauss_And_Legendre.jl:4-66) are computed in FP64 on the GPU.
 *      nlon: any even length 16..2048 with nlon >= 2 num_fourier + 1 (powers
 *      of two up to 1024 take the radix-4 FFT, others a direct DFT); nlat <=
 *      512.  Otherwise JGCM_ERR_UNSUPPORTED.                                  */
int jgcm_mesh_create(int num_fourier, int num_spherical, int nlat, int nlon,
                     int nd, double radius, jgcm_mesh_t** out);
/* Drops the caller's reference.  Leapfrogs, vertical coordinates, graphs and
 * meshes that borrow this mesh's stream hold references of their own, so the
 * mesh is freed after the last of them is destroyed, in any order.          */
int jgcm_mesh_destroy(jgcm_mesh_t* mesh);
/* Diagnostics (no reference counterpart): the mesh's reference count and the
 * number of live graphs that recorded calls on it.                          */
int jgcm_mesh_lifetime(jgcm_mesh_t* mesh, int* refs, int* live_graphs);
int jgcm_mesh_synchronize(jgcm_mesh_t* mesh);
/* hipStream_t of the mesh, returned as void*. */
void* jgcm_mesh_stream(jgcm_mesh_t* mesh);
/* Make `mesh` enqueue on `owner`'s stream (so several meshes, e.g. the
 * truncations of a mixed ensemble, can be captured into one graph).  `mesh`
 * keeps a reference to `owner`, whose stream it uses.  No reference
 * counterpart.                                                              */
int jgcm_mesh_share_stream(jgcm_mesh_t* mesh, jgcm_mesh_t* owner);
int jgcm_mesh_set_precision(jgcm_mesh_t* mesh, int precision);
int jgcm_mesh_get_precision(jgcm_mesh_t* mesh, int* precision);
/* FFT workgroup shape (no reference counterpart: FFTW plans its own): the
 * number of latitude pairs one workgroup of the longitude FFT handles, for the
 * grid -> Fourier fold (forward transforms) and the Fourier -> grid unfold
 * (inverse transforms); 0 restores the built-in rule.  Results do not depend
 * on the shape (each FFT's arithmetic is the same), only the launch does.
 * Values above *max_ppw (the LDS limit) are rejected.  Set by the autotuner
 * (autotune.py), which times the candidates and keeps the fastest.          */
int jgcm_mesh_set_fft_shape(jgcm_mesh_t* mesh, int fold_ppw, int unfold_ppw);
int jgcm_mesh_get_fft_shape(jgcm_mesh_t* mesh, int* fold_ppw, int* unfold_ppw, int* max_ppw);
/* Legendre-stage tile shape (no reference counterpart): 16-row (forward) or
 * 16-latitude (inverse) result tiles per workgroup, 1, 2 or 4 (0 = the
 * built-in 2), per direction; set by the autotuner.  Results do not depend
 * on it.                                                                    */
int jgcm_mesh_set_legendre_shape(jgcm_mesh_t* mesh, int fwd_tiles, int inv_tiles);
int jgcm_mesh_get_legendre_shape(jgcm_mesh_t* mesh, int* fwd_tiles, int* inv_tiles);

/* Host copies of the mesh's public fields (the Julia struct's fields). */
#define JGCM_TABLE_SIN_THETA 0     /* [nθ]                         */
#define JGCM_TABLE_COS_THETA 1     /* [nθ]                         */
#define JGCM_TABLE_WTS 2           /* [nθ]                         */
#define JGCM_TABLE_LAMBDA_C 3      /* [nλ]                         */
#define JGCM_TABLE_THETA_C 4       /* [nθ]                         */
#define JGCM_TABLE_LAMBDA_E 5      /* [nλ+1]                       */
#define JGCM_TABLE_THETA_E 6       /* [nθ+1]                       */
#define JGCM_TABLE_QNM 7           /* [N+1][ns+1][nθ]  (C order: m, n, j) */
#define JGCM_TABLE_DQNM 8          /* [N+1][ns+1][nθ]              */
#define JGCM_TABLE_QWG 9           /* [N+1][ns+1][nθ]              */
#define JGCM_TABLE_EPSILON 10      /* [N+1][ns+1]                  */
#define JGCM_TABLE_LAPLACIAN_EIG 11 /* [N+1][ns+1]                 */
#define JGCM_TABLE_COEF_UVM 12
#define JGCM_TABLE_COEF_UVC 13
#define JGCM_TABLE_COEF_UVP 14
#define JGCM_TABLE_COEF_DLAMBDA 15
#define JGCM_TABLE_COEF_DTHETAM 16
#define JGCM_TABLE_COEF_DTHETAP 17
#define JGCM_TABLE_COEF_ALP_A 18   /* [N+1][ns+1][nθ]              */
#define JGCM_TABLE_COEF_ALP_B 19   /* [N+1][ns+1][nθ]              */
/* Number of doubles in table `which` (so callers can size buffers). */
int jgcm_mesh_table_size(jgcm_mesh_t* mesh, int which, size_t* count);
int jgcm_mesh_get_table(jgcm_mesh_t* mesh, int which, double* host_out, size_t count);

/* ---- transforms ----------------------------------------------------------
 * Trans_Grid_To_Spherical!(mesh, pfield, snm)   Spectral_Spherical_Mesh.jl:221-293
 *   forward: FFT/nλ, even/odd-folded Gaussian quadrature for degrees m..N,
 *   degree num_spherical (N+1) row set to 0, lower triangle (n<m) untouched. */
int jgcm_trans_grid_to_spherical(jgcm_mesh_t* mesh, const double* pfield,
                                 double* snm, int nlev);
/* Trans_Grid_To_Spherical! without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results (bit for bit) as jgcm_trans_grid_to_spherical;
 * the FFT -> Legendre staging lives in a static 16 MiB device array (at most
 * 95 level-fields at T85 on 128 latitudes, 381 at T42; more fails with
 * JGCM_ERR_ARG).  All calls on one GPU share that array: calls that overlap
 * on one GPU race.  The mesh's device must be current.                     */
int jgcm_trans_grid_to_spherical_noalloc(jgcm_mesh_t* mesh, const double* pfield, double* snm,
                                         int nlev);
/* Trans_Spherical_To_Grid!(mesh, snm, pfield)   Spectral_Spherical_Mesh.jl:143-217
 *   inverse: sums degrees m..N+1, g_0 halved, pfield = 2nλ Re(ifft(g)).      */
int jgcm_trans_spherical_to_grid(jgcm_mesh_t* mesh, const double* snm,
                                 double* pfield, int nlev);
/* Trans_Grid_To_Fourier!(mesh, pfield, fourier_g)  Spectral_Spherical_Mesh.jl:295-330
 *   fourier_g complex128 [nlev][nθ][nλ] = fft(pfield)/nλ.                    */
int jgcm_trans_grid_to_fourier(jgcm_mesh_t* mesh, const double* pfield,
                               double* fourier_g, int nlev);

/* Batched forms of the two transforms (no reference counterpart: the
 * reference transforms one array per call).  nseg (1..12) arrays of nlev[k]
 * levels go through ONE Legendre launch and ONE FFT launch, which is how a
 * time step pushes all its level-fields through the transform at once.
 *   inverse: flags[k] & JGCM_BATCH_DIVCOS divides the output rows of array k
 *            by cos θ; flags[k] may add ONE spectral operator applied to the
 *            input on the fly (no spectral temporary is written):
 *              JGCM_INV_GRAD_LAMBDA  cos θ ∂λ h / a  (Compute_Gradient_Cos!,
 *              JGCM_INV_GRAD_THETA   cos θ ∂θ h / a   SSM:655-692)
 *              JGCM_INV_UCOS         u cos θ from vor = snm[k], div = snm_b[k]
 *              JGCM_INV_VCOS         v cos θ  (Compute_Ucos_Vcos_From_Vor_Div!,
 *                                             SSM:517-580)
 *            so DIVCOS|GRAD_* is Compute_Gradients! (SSM:721-741) and
 *            DIVCOS|UCOS, DIVCOS|VCOS is UV_Grid_From_Vor_Div! (SSM:462-482).
 *            snm_b may be NULL when no array uses UCOS / VCOS.
 *   forward: flags[k] == JGCM_BATCH_VORDIV pairs array k (grid u) with array
 *            k+1 (grid v, flags[k+1] = 0, same nlev) and writes vor to snm[k],
 *            div to snm[k+1] (Vor_Div_From_Grid_UV!); other arrays get the
 *            plain forward transform.  At most 4 Legendre jobs: each run of
 *            plain arrays is one job, each u/v pair two.
 * flags may be NULL (all plain).  Results equal the one-array calls.       */
#define JGCM_BATCH_DIVCOS 1
#define JGCM_BATCH_VORDIV 2
#define JGCM_INV_GRAD_LAMBDA 2
#define JGCM_INV_GRAD_THETA 4
#define JGCM_INV_UCOS 6
#define JGCM_INV_VCOS 8
int jgcm_trans_spherical_to_grid_batch(jgcm_mesh_t* mesh, int nseg, const double* const* snm,
                                       const double* const* snm_b, double* const* pfield,
                                       const int* nlev, const int* flags);
int jgcm_trans_grid_to_spherical_batch(jgcm_mesh_t* mesh, int nseg, const double* const* pfield,
                                       double* const* snm, const int* nlev, const int* flags);

/* Plans: the batched transforms of several meshes in one launch per stage
 * (no reference counterpart: the reference steps one model at a time; this is
 * the mixed-truncation ensemble's launch form).  A plan takes, per mesh g, the
 * arguments of jgcm_trans_spherical_to_grid_batch / _grid_to_spherical_batch
 * (nseg[g], snm[g], snm_b[g], pfield[g], nlev[g], flags[g]; snm_b and flags
 * may be NULL as a whole or per mesh).  The meshes must share one stream
 * (jgcm_mesh_share_stream) and one Legendre precision, and each appears once.
 * Execute runs the FFT stage per mesh (the work of every latitude is the same)
 * and the Legendre stage of ALL meshes in ONE launch: every tile of the
 * one-mesh launches -- (mesh, m, parity, row or latitude group, field chunk),
 * with each mesh's tile shape (jgcm_mesh_set_legendre_shape) -- is a work
 * item; items are sorted by estimated cost and dealt longest-first to the
 * least-loaded workgroup.  workgroups: > 0 that many persistent workgroups,
 * < 0 as many as fit on the device at once, 0 one workgroup per item
 * (heaviest launched first).  Results equal the one-mesh batched calls bit
 * for bit.  Creation reserves each mesh's scratch, allocates and uploads the
 * work list (so it must happen before a graph capture); the arrays named at
 * creation are the ones every execute reads and writes, and the tile shapes
 * and precision are those of creation time (execute refuses a changed
 * precision).                                                              */
int jgcm_plan_spherical_to_grid_create(int nmesh, jgcm_mesh_t* const* meshes, const int* nseg,
                                       const double* const* const* snm,
                                       const double* const* const* snm_b,
                                       double* const* const* pfield, const int* const* nlev,
                                       const int* const* flags, int workgroups,
                                       jgcm_plan_t** out);
int jgcm_plan_grid_to_spherical_create(int nmesh, jgcm_mesh_t* const* meshes, const int* nseg,
                                       const double* const* const* pfield,
                                       double* const* const* snm, const int* const* nlev,
                                       const int* const* flags, int workgroups,
                                       jgcm_plan_t** out);
int jgcm_plan_execute(jgcm_plan_t* plan);
/* Work items and workgroups of the Legendre launch; the work list itself as
 * 4 ints per item (mesh, segment 2m + parity, y, z) in launch order, and
 * wg_start[nworkgroups + 1] (workgroup w runs items wg_start[w] ..
 * wg_start[w+1] - 1).                                                      */
int jgcm_plan_info(jgcm_plan_t* plan, int* nitems, int* nworkgroups);
int jgcm_plan_work_list(jgcm_plan_t* plan, int* items, int* wg_start);
int jgcm_plan_destroy(jgcm_plan_t* plan);
/* The plans' scheduler on the host (no GPU call): n costs dealt in
 * decreasing order (ties in index order) to the least-loaded of nwg bins
 * (ties: lowest bin); order[n] = the item indices grouped by bin,
 * wg_start[used + 1] the bins' ranges.  nwg <= 0 or >= n: one bin per item,
 * heaviest first.                                                          */
int jgcm_lpt_schedule(int n, const double* cost, int nwg, int* order, int* wg_start, int* used);

/* Adjoints of the transform pair (no reference counterpart; the building
 * block of the reverse-mode derivative of a step).  Transposes with respect
 * to <x, y> = sum x y over grid values and <a, b> = sum Re(conj(a) b) over
 * the spectral array: weighted-transpose Legendre plus conjugate FFT.
 * grid_to_spherical_adjoint: spectral in -> grid out (degrees m..N read).
 * spherical_to_grid_adjoint: grid in -> spectral out (degrees m..N+1
 * written, lower triangle set to zero).                                     */
int jgcm_trans_grid_to_spherical_adjoint(jgcm_mesh_t* mesh, const double* snm, double* pfield,
                                         int nlev);
/* grid_to_spherical_adjoint for up to 12 arrays in one Legendre and one FFT
 * launch (no reference counterpart: the adjoint of the forward batch).     */
int jgcm_trans_grid_to_spherical_adjoint_batch(jgcm_mesh_t* mesh, int nseg,
                                               const double* const* snm, double* const* pfield,
                                               const int* nlev);
int jgcm_trans_spherical_to_grid_adjoint(jgcm_mesh_t* mesh, const double* pfield, double* snm,
                                         int nlev);
/* spherical_to_grid_adjoint for up to 12 arrays in one FFT and one Legendre
 * launch (no reference counterpart: the adjoint of the inverse batch).
 * flags[k] (or NULL): JGCM_BATCH_DIVCOS divides array k by cos θ on the way
 * in.  accumulate = 0: snm[k] overwritten (lower triangle zeroed);
 * accumulate = 1: the result is added to snm[k] (degrees m..N+1), and the
 * outputs must not overlap.                                                 */
int jgcm_trans_spherical_to_grid_adjoint_batch(jgcm_mesh_t* mesh, int nseg,
                                               const double* const* pfield, double* const* snm,
                                               const int* nlev, const int* flags, int accumulate);

/* Adjoints of the barotropic step's operators (no reference counterpart; the
 * same inner products).  Outputs may not alias inputs.
 * vor_div_from_grid_uv_adjoint: transpose of Vor_Div_From_Grid_UV!
 *   (Spectral_Spherical_Mesh.jl:349-393); div_bar may be NULL (zero).       */
int jgcm_vor_div_from_grid_uv_adjoint(jgcm_mesh_t* mesh, const double* vor_bar, const double* div_bar,
                                      double* u_bar, double* v_bar, int nlev);
/* transpose of the u cos θ / v cos θ stencil of Compute_Ucos_Vcos_From_Vor_Div!
 * (Spectral_Spherical_Mesh.jl:517-580); div_bar may be NULL (not written).  */
int jgcm_ucos_vcos_adjoint(jgcm_mesh_t* mesh, const double* ucos_bar, const double* vcos_bar,
                           double* vor_bar, double* div_bar, int nlev);
/* transpose of UV_Grid_From_Vor_Div! (Spectral_Spherical_Mesh.jl:462-482);
 * div_bar may be NULL.                                                      */
int jgcm_uv_grid_from_vor_div_adjoint(jgcm_mesh_t* mesh, const double* u_bar, const double* v_bar,
                                      double* vor_bar, double* div_bar, int nlev);
/* transpose of the absolute-vorticity flux δu = (ζ+f) v, δv = -(ζ+f) u
 * (Barotropic_Dynamics.jl:26-29) linearised at (grid_vor, grid_u, grid_v).  */
int jgcm_vorticity_flux_adjoint(jgcm_mesh_t* mesh, const double* grid_vor, const double* grid_u,
                                const double* grid_v, double omega, const double* du_bar,
                                const double* dv_bar, double* vor_bar, double* u_bar, double* v_bar,
                                int nlev);
/* transpose of Compute_Spectral_Damping! + Filtered_Leapfrog! + Time_Advance!
 * (Time_Integrator.jl:58-101, Dyn_Data.jl:326-364) for one field:
 * (p_bar', c_bar') -> (p_bar, c_bar, g_bar), g_bar the cotangent of the
 * undamped tendency.  init_step selects the first-step form; damp = 0 for a
 * field the step does not damp (ln ps).  In place (p_bar == p_bar_next,
 * c_bar == c_bar_next) is allowed.                                           */
int jgcm_damped_leapfrog_adjoint(jgcm_leapfrog_t* lf, int init_step, int damp,
                                 const double* p_bar_next, const double* c_bar_next, double* p_bar,
                                 double* c_bar, double* g_bar, int nlev);
/* the same for up to 8 fields of one integrator in one launch (no reference
 * counterpart: the transpose of the batched leapfrog); damp[f] and nlev[f]
 * per field; bit-identical to the single calls.                            */
int jgcm_damped_leapfrog_adjoint_batch(jgcm_leapfrog_t* lf, int init_step, int nfields,
                                       const int* damp, const double* const* p_bar_next,
                                       const double* const* c_bar_next, double* const* p_bar,
                                       double* const* c_bar, double* const* g_bar,
                                       const int* nlev);
/* one reverse step of Barotropic_Dynamics! (Barotropic_Dynamics.jl:2-46), in
 * place on (p_bar, c_bar); grid_vor, grid_u, grid_v are the grid fields of
 * the step's starting state.                                                */
int jgcm_barotropic_step_adjoint(jgcm_leapfrog_t* lf, int init_step, const double* grid_vor,
                                 const double* grid_u, const double* grid_v, double omega,
                                 double* p_bar, double* c_bar, int nlev);
/* one reverse step of Shallow_Water_Physics! + Shallow_Water_Dynamics!
 * (Shallow_Water_Dynamics.jl:7-21, :88-179), in place on the six state
 * cotangents; the grid fields and spe_h (the spectral height) are those of
 * the step's starting state.                                                */
int jgcm_sw_step_adjoint(jgcm_leapfrog_t* lf, int init_step, const double* grid_vor,
                         const double* grid_div, const double* grid_h, const double* grid_u,
                         const double* grid_v, const double* spe_h, double omega, double kappa_m,
                         double kappa_t, double h_0, double* vor_p_bar, double* vor_c_bar,
                         double* div_p_bar, double* div_c_bar, double* h_p_bar, double* h_c_bar,
                         int nlev);

/* Pieces of the reverse Held-Suarez step (no reference counterpart; even-sigma
 * coordinate).  jgcm_hs_column_adjoint: transpose of the grid-column part of
 * Atmosphere_Update! (Spectral_Dynamics.jl:206-307, 452-467; HS_Forcing.jl:1-130;
 * Spectral_Dynamics.jl:5-28 for the energy integral of the p-fields) at a state.
 *   state[14]  : grid div, u, v, T, vor [nd]; ∂λT, ∂θT [nd]; ∂λ ln ps, ∂θ ln ps [1];
 *                u_p, v_p, T_p [nd]; ps_p, ps_c [1]
 *   cot_in[5]  : cotangents of the grid δu, δv, δT, Φ+KE [nd] and δln ps [1]
 *   cot_out[14]: cotangents of div, u, v, T, vor, ∂λT, ∂θT, ∂λ ln ps, ∂θ ln ps,
 *                u_p, v_p, T_p, ps_p, ps_c (same shapes as the state)
 *   scal_bar   : device; [1], [2] = cotangents of the energy integral and mean ps
 *                of the p-fields (from jgcm_hs_corrections_adjoint)
 *   param_bar  : NULL, or [4][nθ][nλ] doubles to which each column adds its share
 *                of the cotangents of k_f, k_a, k_s and ΔT_y (HS_Forcing.jl:8-13)  */
int jgcm_hs_column_adjoint(jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, double dt,
                           const double* scal_bar, const double* const* state,
                           const double* const* cot_in, double* const* cot_out,
                           double* param_bar);
/* transpose of the mass and energy corrections (Spectral_Dynamics.jl:30-68).
 * scal: the step's seven correction scalars (device, jgcm_hs_scalars_device);
 * spe_t_c_bar, spe_lnps_c_bar: cotangents of the corrected new T and ln ps.
 * Writes scal_bar (device, 4 doubles: energy_n, energy_p, mean ps_p, mean ps_n
 * cotangents; the last three are read by jgcm_hs_column_adjoint as [1], [2])
 * and the cotangents of the new-state grid u, v, T and ln ps (T_n and ps_n
 * passed after their corrections).                                          */
int jgcm_hs_corrections_adjoint(jgcm_semi_implicit_t* si, const double* scal,
                                const double* spe_t_c_bar, const double* spe_lnps_c_bar,
                                const double* grid_u_n, const double* grid_v_n,
                                const double* grid_t_n, const double* grid_ps_n, double* scal_bar,
                                double* u_n_bar, double* v_n_bar, double* t_n_bar,
                                double* lnps_n_bar);
/* transpose of ∇²(energy); δdiv -= energy; Implicit_Correction! (Spectral_Dynamics.jl:309-316,
 * Semi_Implicit.jl:462-513) from its per-n matrix [ns+1][2nd+1][7nd+3] (outputs
 * δdiv, δln ps, δT; inputs δdiv, δln ps, δT, energy, div_c, div_p, ln ps_c,
 * ln ps_p, T_c, T_p).  out: 7nd+3 spectral level-fields in that order.       */
int jgcm_implicit_adjoint(jgcm_mesh_t* mesh, int nd, const double* matrix, const double* ddiv_bar,
                          const double* dlnps_bar, const double* dt_bar, double* out);
/* transpose of Compute_Gradient_Cos! (Spectral_Spherical_Mesh.jl:677-687):
 * (∂λ_bar, ∂θ_bar) -> h_bar (added when accumulate != 0).                   */
int jgcm_gradient_cos_adjoint(jgcm_mesh_t* mesh, const double* grad_lambda_bar,
                              const double* grad_theta_bar, double* h_bar, int accumulate, int nlev);

/* ---- operators ---------------------------------------------------------- */
/* Divide_By_Cos!(cosθ, grid)                    Spectral_Spherical_Mesh.jl:333-338 */
int jgcm_divide_by_cos(jgcm_mesh_t* mesh, double* grid, int nlev);
/* Multiply_By_Cos!(cosθ, grid, grid_cos)        Spectral_Spherical_Mesh.jl:341-346 */
int jgcm_multiply_by_cos(jgcm_mesh_t* mesh, const double* grid, double* grid_cos, int nlev);
/* Vor_Div_From_Grid_UV!(mesh, u, v, vor, div)  Spectral_Spherical_Mesh.jl:349-393 */
int jgcm_vor_div_from_grid_uv(jgcm_mesh_t* mesh, const double* grid_u,
                              const double* grid_v, double* vor, double* div, int nlev);
/* Compute_Alpha_Operator!(mesh, a, b, isign, α) Spectral_Spherical_Mesh.jl:424-459
 *   fourier_a/b are full Fourier arrays as produced by Trans_Grid_To_Fourier!. */
int jgcm_compute_alpha_operator(jgcm_mesh_t* mesh, const double* fourier_a,
                                const double* fourier_b, double isign,
                                double* alpha, int nlev);
/* UV_Grid_From_Vor_Div!(mesh, vor, div, u, v)  Spectral_Spherical_Mesh.jl:462-482 */
int jgcm_uv_grid_from_vor_div(jgcm_mesh_t* mesh, const double* vor,
                              const double* div, double* grid_u, double* grid_v, int nlev);
/* Compute_Ucos_Vcos_From_Vor_Div!               Spectral_Spherical_Mesh.jl:517-580 */
int jgcm_compute_ucos_vcos_from_vor_div(jgcm_mesh_t* mesh, const double* vor,
                                        const double* div, double* ucos,
                                        double* vcos, int nlev);
/* Apply_Laplacian!(mesh, s)                     Spectral_Spherical_Mesh.jl:614-621 */
int jgcm_apply_laplacian(jgcm_mesh_t* mesh, double* snm, int nlev);
/* Compute_Gradient_Cos!(mesh, hs, dλ, dθ)       Spectral_Spherical_Mesh.jl:655-692 */
int jgcm_compute_gradient_cos(jgcm_mesh_t* mesh, const double* spe_hs,
                              double* spe_cos_dlambda, double* spe_cos_dtheta, int nlev);
/* Compute_Gradients!(mesh, hs, grid_dλ, grid_dθ) Spectral_Spherical_Mesh.jl:721-741 */
int jgcm_compute_gradients(jgcm_mesh_t* mesh, const double* spe_hs,
                           double* grid_dlambda, double* grid_dtheta, int nlev);
/* Add_Horizontal_Advection!(mesh, hs, u, v, δh) Spectral_Spherical_Mesh.jl:694-718 */
int jgcm_add_horizontal_advection(jgcm_mesh_t* mesh, const double* spe_hs,
                                  const double* grid_u, const double* grid_v,
                                  double* grid_dhs, int nlev);
/* Area_Weighted_Global_Mean(mesh, grid) (level 1) Spectral_Spherical_Mesh.jl:746-760.
 * Synchronises the mesh stream and returns the value on the host.          */
int jgcm_area_weighted_global_mean(jgcm_mesh_t* mesh, const double* grid, double* result);

/* ---- time integration: Filtered_Leapfrog ------------------------------------
 * Filtered_Leapfrog(robert_coef, damping_order, damping_coef, eigen, implicit_coef,
 *                   Δt, init_step, start_time, end_time)   Time_Integrator.jl:20-38
 * `eigen` is the mesh's laplacian_eig; the ∇^(2 order) damping table is
 * damping_coef (eig / eig[m=0, n=num_spherical-1])^order (Time_Integrator.jl:31). */
int jgcm_leapfrog_create(jgcm_mesh_t* mesh, double robert_coef, int damping_order,
                         double damping_coef, double implicit_coef, int64_t dt, int init_step,
                         int64_t start_time, int64_t end_time, jgcm_leapfrog_t** out);
int jgcm_leapfrog_destroy(jgcm_leapfrog_t* lf);
/* Update_Init_Step!(integrator)   Time_Integrator.jl:40-43 (error when called twice) */
int jgcm_leapfrog_update_init_step(jgcm_leapfrog_t* lf);
int jgcm_leapfrog_state(jgcm_leapfrog_t* lf, int* init_step, int64_t* dt, int64_t* time);
int jgcm_leapfrog_set_time(jgcm_leapfrog_t* lf, int64_t time);
int jgcm_leapfrog_damping(jgcm_leapfrog_t* lf, double* host_out, size_t count);
/* Compute_Spectral_Damping!(integrator, Qc, Qp, δQ)   Time_Integrator.jl:58-81 */
int jgcm_compute_spectral_damping(jgcm_leapfrog_t* lf, const double* Qc, const double* Qp,
                                  double* dQ, int nlev);
/* Filtered_Leapfrog!(integrator, δQ, Qp, Qc, Qn)      Time_Integrator.jl:86-101 */
int jgcm_filtered_leapfrog(jgcm_leapfrog_t* lf, const double* dQ, const double* Qp, double* Qc,
                           double* Qn, int nlev);
/* Compute_Spectral_Damping! followed by Filtered_Leapfrog! in one pass (same results). */
int jgcm_damped_leapfrog(jgcm_leapfrog_t* lf, double* dQ, const double* Qp, double* Qc,
                         double* Qn, int nlev);
/* Compute_Spectral_Damping! / Filtered_Leapfrog! for nfields (<= 8) prognostic
 * fields in one launch (no reference counterpart; Spectral_Dynamics.jl:319-328
 * makes seven calls).  mode[f]: 1 damping of dQ only, 2 leapfrog only, 3 both
 * (damping, then the leapfrog with the damped tendency).  Qn[f] may be NULL
 * for mode 1.                                                              */
int jgcm_damped_leapfrog_batch(jgcm_leapfrog_t* lf, int nfields, double* const* dQ,
                               const double* const* Qp, double* const* Qc, double* const* Qn,
                               const int* nlev, const int* mode);
/* Implicit_Correction!(integrator, div_c, div_p, h_c, h_p, h_0, δdiv, δh)
 *                                         Shallow_Water_Dynamics.jl:25-54 */
int jgcm_sw_implicit_correction(jgcm_leapfrog_t* lf, const double* div_c, const double* div_p,
                                const double* h_c, const double* h_p, double h_0, double* ddiv,
                                double* dh, int nlev);

/* ---- grid-space tendency helpers of the steppers --------------------------- */
/* Compute_Abs_Vor!(vor, coriolis, absvor) (Atmo_Data.jl:59-65, f = 2Ω sinθ :45) fused with
 * δu (=|+=) absvor v, δv (=|+=) -absvor u  (Barotropic_Dynamics.jl:28-29,
 * Shallow_Water_Dynamics.jl:129-130, Spectral_Dynamics.jl:300-301).
 * grid_absvor may be NULL; with grid_du = grid_dv = NULL only absvor is written
 * (plain Compute_Abs_Vor!).  accumulate: 0 assigns, 1 adds.                 */
int jgcm_vorticity_flux(jgcm_mesh_t* mesh, const double* grid_vor, const double* grid_u,
                        const double* grid_v, double omega, double* grid_absvor, double* grid_du,
                        double* grid_dv, int nlev, int accumulate);
/* out = a x + b y over `count` doubles (x or y may be NULL).  Reference broadcast
 * updates such as spe_energy .= spe_kin + spe_h_c (Shallow_Water_Dynamics.jl:141). */
int jgcm_axpby(jgcm_mesh_t* mesh, size_t count, double a, const double* x, double b,
               const double* y, double* out);
/* out = (accumulate ? out : 0) + sign * x * y, e.g. grid_δh .-= grid_h .* grid_div
 * (Shallow_Water_Dynamics.jl:148).                                           */
int jgcm_xmy(jgcm_mesh_t* mesh, size_t count, double sign, const double* x, const double* y,
             double* out, int accumulate);
/* out = exp(x) (op 0, e.g. grid_ps_n .= exp.(grid_lnps), Spectral_Dynamics.jl:336),
 * log(x) (op 1) or a x + b (op 2), over `count` doubles.                      */
#define JGCM_MAP_EXP 0
#define JGCM_MAP_LOG 1
#define JGCM_MAP_AFFINE 2
int jgcm_map(jgcm_mesh_t* mesh, int op, size_t count, const double* x, double a, double b,
             double* out);
/* grid_out = [grid_base +] 0.5 (u^2 + v^2)  (Shallow_Water_Dynamics.jl:139,
 * Spectral_Dynamics.jl:307); grid_base may be NULL.                           */
int jgcm_kinetic_energy(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                        const double* grid_base, double* grid_out, int nlev);
/* Shallow_Water_Physics!(dyn_data, kappa_m, kappa_t, h_eq)  Shallow_Water_Dynamics.jl:7-21 */
int jgcm_shallow_water_physics(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                               const double* grid_h, const double* h_eq, double kappa_m,
                               double kappa_t, double* grid_du, double* grid_dv, double* grid_dh,
                               int nlev);
/* No reference counterpart (a fusion of reference steps): Shallow_Water_Physics!
 * (Shallow_Water_Dynamics.jl:7-21) followed by the first grid steps of
 * Shallow_Water_Dynamics! (:126-130 absolute vorticity and δu += absvor v,
 * δv -= absvor u; :139 KE = (u² + v²)/2) in one pass; results equal
 * jgcm_shallow_water_physics + jgcm_vorticity_flux(accumulate) +
 * jgcm_kinetic_energy bit for bit.  grid_u_p, grid_v_p, grid_h_p (all or
 * none): also the p = c half of Time_Advance! (Dyn_Data.jl:328-362) for the
 * grid u, v and h, which the step reads only as c.                          */
int jgcm_sw_explicit(jgcm_mesh_t* mesh, const double* grid_u, const double* grid_v,
                     const double* grid_h, const double* h_eq, double kappa_m, double kappa_t,
                     const double* grid_vor, double omega, double* grid_du, double* grid_dv,
                     double* grid_dh, double* grid_absvor, double* grid_kin, double* grid_u_p,
                     double* grid_v_p, double* grid_h_p, int nlev);
/* The shallow-water height tendency's grid terms (Shallow_Water_Dynamics.jl
 * :147-148): jgcm_add_horizontal_advection(spe_h, u, v, δh) followed by
 * δh -= h div (jgcm_xmy with sign -1, accumulating), the second in the same
 * pass as the advection product.  Equal to the two calls bit for bit.  No
 * reference counterpart (a fusion).                                         */
int jgcm_sw_height_tendency(jgcm_mesh_t* mesh, const double* spe_h, const double* grid_u,
                            const double* grid_v, const double* grid_h, const double* grid_div,
                            double* grid_dh, int nlev);
/* jgcm_sw_height_tendency with the gradient of h given on the grid
 * (cos-weighted ∂λh, ∂θh divided by cos θ, as Compute_Gradients! leaves them;
 * e.g. formed by the previous step's closing inverse batch).  No reference
 * counterpart (a fusion).                                                   */
int jgcm_sw_height_tendency_carried(jgcm_mesh_t* mesh, const double* grid_dlambda_h,
                                    const double* grid_dtheta_h, const double* grid_u,
                                    const double* grid_v, const double* grid_h,
                                    const double* grid_div, double* grid_dh, int nlev);
/* The spectral part of a shallow-water step after its forward transforms, in
 * one pass: energy = ∇²(KE + h_c) and δdiv -= energy (Shallow_Water_Dynamics.jl
 * :141-143), Implicit_Correction! (:150-157, Shallow_Water_Dynamics.jl:25-54),
 * Compute_Spectral_Damping! + Filtered_Leapfrog! of ζ, D and h (:159-165,
 * Time_Integrator.jl:58-101), and the spectral part of Time_Advance!
 * (Dyn_Data.jl:328-340): vor/div/h are {p, c, n} arrays; on return p holds
 * the filtered current value, c and n the new one.  Equal, bit for bit, to
 * jgcm_axpby, jgcm_apply_laplacian, jgcm_axpby, jgcm_sw_implicit_correction,
 * jgcm_damped_leapfrog_batch (damping on) and jgcm_time_advance of the three
 * fields.  No reference counterpart (a fusion).                             */
int jgcm_sw_spectral_step(jgcm_leapfrog_t* lf, double h_0, const double* spe_kin, double* spe_energy,
                          double* spe_dvor, double* spe_ddiv, double* spe_dh, double* const* vor,
                          double* const* div, double* const* h, int nlev);
/* The same with advance = 0: Filtered_Leapfrog!'s own writes (the filtered c
 * in place, n = the new value) instead of the spectral Time_Advance!; for
 * steps that advance the state by renaming p, c, n (no reference
 * counterpart; the shallow-water step's rotate option).                     */
int jgcm_sw_spectral_step_ex(jgcm_leapfrog_t* lf, double h_0, const double* spe_kin,
                             double* spe_energy, double* spe_dvor, double* spe_ddiv, double* spe_dh,
                             double* const* vor, double* const* div, double* const* h, int nlev,
                             int advance);

/* ---- primitive equations (Held-Suarez) ------------------------------------- */
/* Vert_Coordinate(nλ, nθ, nd, vert_coord_option, vert_difference_option,
 *                 vert_advect_scheme, p_ref, zero_top)   Vert_Coordinate.jl:42-56
 * with the reference's defaults scale_heights = 4, surf_res = 1, p_press =
 * 0.1, p_sigma = 0.3, exponent = 2.5 (jgcm_vert_coord_create_ex takes them).
 * Pure sigma coordinates (ak = 0): even sigma (:112-117, any nd <= 64), the
 * fixed tables mcm (Compute_Old_Model_Sigma, :157-165, nd = 14) and v197
 * (Compute_V197_Sigma, :143-154, nd = 18; for the tables nd must equal their
 * length — the reference indexes the table with nd and silently truncates a
 * smaller one) and uneven sigma (Compute_Uneven_Sigma, :120-141).  hybrid
 * (:71-77, Transition :93-110): pressure levels aloft, sigma levels below
 * p_sigma.  The reference cannot run uneven_sigma and hybrid (:70 passes a,
 * b before they exist, :72-73 omit nd, :74 calls Transition without nd, and
 * :131-133 evaluate exp(+...) against the docstring's exp(-...)); they are
 * built here as the docstring and the call structure state them:
 *   uneven: ζ_k = 1 - k/nd, b_k = exp(-(surf_res ζ + (1-surf_res) ζ^exponent)
 *           scale_heights), b_nd = 1, b_0 = 0 when zero_top, a = 0;
 *   hybrid: b_u = uneven without a zero top, t_k = Transition(b_u,k),
 *           a_k = p_ref b_u,k (1 - t_k), b_k = b_u,k t_k.                    */
int jgcm_vert_coord_create(jgcm_mesh_t* mesh, int nd, int vert_coord_option,
                           int vert_difference_option, int vert_advect_scheme, double p_ref,
                           int zero_top, jgcm_vert_coord_t** out);
int jgcm_vert_coord_create_ex(jgcm_mesh_t* mesh, int nd, int vert_coord_option,
                              int vert_difference_option, int vert_advect_scheme, double p_ref,
                              int zero_top, double scale_heights, double surf_res,
                              double p_press, double p_sigma, double exponent,
                              jgcm_vert_coord_t** out);
int jgcm_vert_coord_destroy(jgcm_vert_coord_t* vc);
/* host copies of ak, bk ([nd+1] each; either may be NULL) */
int jgcm_vert_coord_get(jgcm_vert_coord_t* vc, double* ak, double* bk);
/* Pressure_Variables!(vert_coord, ps, p_half, Δp, lnp_half, p_full, lnp_full)
 *                                                       Press_And_Geopot.jl:36-67 */
int jgcm_pressure_variables(jgcm_vert_coord_t* vc, const double* grid_ps, double* p_half,
                            double* dp, double* lnp_half, double* p_full, double* lnp_full);
/* Compute_Pressures_And_Heights!(atmo, vert_coord, ps, geopots, t, p_half, Δp,
 *   lnp_half, p_full, lnp_full, z_full, z_half)          Press_And_Geopot.jl:3-19 */
int jgcm_compute_pressures_and_heights(jgcm_vert_coord_t* vc, const jgcm_atmo_params* atmo,
                                       const double* grid_ps, const double* grid_geopots,
                                       const double* grid_t, double* p_half, double* dp,
                                       double* lnp_half, double* p_full, double* lnp_full,
                                       double* z_full, double* z_half);
/* Compute_Geopotential!                                  Press_And_Geopot.jl:71-111 */
int jgcm_compute_geopotential(jgcm_vert_coord_t* vc, const jgcm_atmo_params* atmo,
                              const double* lnp_half, const double* lnp_full, const double* grid_t,
                              const double* grid_geopots, double* geopot_full,
                              double* geopot_half);
/* Vert_Advection!(vert_coord, r, dz, w, Δt, scheme, rdt)  Vert_Coordinate.jl:171-210 */
int jgcm_vert_advection(jgcm_vert_coord_t* vc, const double* r, const double* dz,
                        const double* w, double* rdt);
/* Four_In_One!(...)                                      Spectral_Dynamics.jl:88-183 */
int jgcm_four_in_one(jgcm_vert_coord_t* vc, const jgcm_atmo_params* atmo, const double* grid_div,
                     const double* grid_u, const double* grid_v, const double* grid_ps,
                     const double* dp, const double* lnp_half, const double* lnp_full,
                     const double* p_full, const double* grid_dlambda_ps,
                     const double* grid_dtheta_ps, const double* grid_t, double* grid_M_half,
                     double* grid_w_full, double* grid_du, double* grid_dv, double* grid_dps,
                     double* grid_dt);
/* Semi_Implicit_Solver(vert_coord, atmo_data, integrator, ps_ref, t_ref, wave_numbers)
 *   Semi_Implicit.jl:40-160: reference profiles, ν, τ, γ, h, div_mat by unit
 *   vectors (FP64, host) and the inverted wave matrices W_n uploaded.        */
int jgcm_semi_implicit_create(jgcm_vert_coord_t* vc, const jgcm_atmo_params* atmo,
                              jgcm_leapfrog_t* lf, double ps_ref, const double* t_ref,
                              jgcm_semi_implicit_t** out);
int jgcm_semi_implicit_destroy(jgcm_semi_implicit_t* si);
/* Update_Init_Step!(semi_implicit)  Semi_Implicit.jl:182-188 */
int jgcm_semi_implicit_update_init_step(jgcm_semi_implicit_t* si);
/* host copies: 0 h[nd], 1 div_mat[nd][nd], 2 wave_matrix[N+1][nd][nd], 3 ν, 4 τ,
 * 5 γ, 6 Δp_ref, 7 lnp_half_ref[nd+1], 8 lnp_full_ref */
int jgcm_semi_implicit_get(jgcm_semi_implicit_t* si, int which, double* host_out, size_t count);
/* Implicit_Correction!(semi_implicit, vert_coord, atmo_data, div_c, div_p,
 *   lnps_c, lnps_p, t_c, t_p, δdiv, δlnps, δt)           Semi_Implicit.jl:462-513 */
int jgcm_implicit_correction(jgcm_semi_implicit_t* si, const double* div_c, const double* div_p,
                             const double* lnps_c, const double* lnps_p, const double* t_c,
                             const double* t_p, double* ddiv, double* dlnps, double* dt);
/* Apply_Laplacian!(spe_energy); spe_δdiv .-= spe_energy; Implicit_Correction!
 * (Spectral_Dynamics.jl:309-316) in one pass: spe_energy is replaced by its
 * Laplacian, as in the reference.                                          */
int jgcm_hs_energy_implicit_correction(jgcm_semi_implicit_t* si, double* spe_energy,
                                       const double* div_c, const double* div_p,
                                       const double* lnps_c, const double* lnps_p,
                                       const double* t_c, const double* t_p, double* ddiv,
                                       double* dlnps, double* dt);
/* The same, followed by the step's Compute_Spectral_Damping! and
 * Filtered_Leapfrog! (Spectral_Dynamics.jl:319-328; vorticity, divergence and
 * T damped, ln ps not) in the same pass (no single reference function): the
 * c arrays of vor, div, ln ps and T receive the Robert filter, the n arrays
 * the new values, the tendencies the damped values.  Only the coefficients
 * with n >= m are read and written: those with n < m, zero in every state a
 * step produces (no transform or step writes them), keep their values.  On
 * the others equal to jgcm_hs_energy_implicit_correction followed by
 * jgcm_damped_leapfrog_batch bit for bit.                                   */
int jgcm_hs_energy_implicit_leapfrog(jgcm_semi_implicit_t* si, double* spe_energy, double* div_c,
                                     const double* div_p, double* lnps_c, const double* lnps_p,
                                     double* t_c, const double* t_p, double* ddiv, double* dlnps,
                                     double* dt, double* dvor, const double* vor_p, double* vor_c,
                                     double* vor_n, double* div_n, double* lnps_n, double* t_n);
/* Step part 1, one column pass: Pressure_Variables!(ps_c) (Spectral_Dynamics.jl:256),
 * Spectral_Dynamics_Physics! / HS_Forcing! on the p-fields (:452-467,
 * HS_Forcing.jl:1-130; skipped when do_forcing = 0, δps = 0 always) and
 * Compute_Corrections_Init (:5-28, results kept on the device).  The
 * pressure variables are evaluated in registers and stored only when p_half
 * is not NULL (then all five must be given).                               */
int jgcm_hs_physics(jgcm_semi_implicit_t* si, const jgcm_hs_params* hs, int64_t dt,
                    const double* grid_ps_c, const double* grid_u_p, const double* grid_v_p,
                    const double* grid_t_p, const double* grid_ps_p, double* p_half, double* dp,
                    double* lnp_half, double* p_full, double* lnp_full, double* grid_du,
                    double* grid_dv, double* grid_t_eq, double* grid_dt, double* grid_dps,
                    int do_forcing);
/* Step part 2, one column pass (Spectral_Dynamics.jl:260-307 without the
 * transforms): ∇ps *= ps, Four_In_One!, Compute_Geopotential!, δlnps = δps/ps,
 * Vert_Advection! of u, v, T, absolute-vorticity flux, energy = Φ + KE.
 * When dp is NULL the pressure variables (dp, lnp_half, lnp_full, p_full)
 * are evaluated from grid_ps in registers, bit-identical to
 * jgcm_pressure_variables, instead of being read.
 * With grid_dlambda_t / grid_dtheta_t (the grid gradients of T, may be NULL)
 * it also applies the grid part of Add_Horizontal_Advection! for T
 * (Spectral_Spherical_Mesh.jl:716): δT -= u ∂λT + v ∂θT.                    */
int jgcm_hs_column_dynamics(jgcm_semi_implicit_t* si, const double* grid_div, const double* grid_u,
                            const double* grid_v, const double* grid_ps, const double* dp,
                            const double* lnp_half, const double* lnp_full, const double* p_full,
                            double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t,
                            const double* grid_vor, const double* grid_geopots, double* grid_M_half,
                            double* grid_w_full, double* grid_du, double* grid_dv,
                            double* grid_dps, double* grid_dt, double* geopot_half,
                            double* geopot_full, double* grid_dlnps, double* grid_absvor,
                            double* grid_energy, const double* grid_dlambda_t,
                            const double* grid_dtheta_t);
/* Spectral_Dynamics_Physics! (Spectral_Dynamics.jl:452-467) followed by the
 * column part of Spectral_Dynamics! (:250-307), one column pass: the HS
 * forcing of the p-fields, δps = 0 and the Compute_Corrections_Init partial
 * sums (as jgcm_hs_physics with do_forcing = 1), then everything
 * jgcm_hs_column_dynamics does with the pressure variables evaluated from
 * grid_ps; the physics tendencies stay in registers instead of going
 * through grid_du, grid_dv, grid_dt and grid_dps, which are only written.
 * No reference counterpart as one call; equal to the two calls to rounding. */
int jgcm_hs_physics_column_dynamics(
    jgcm_semi_implicit_t* si, const jgcm_hs_params* hs, int64_t dt, const double* grid_u_p,
    const double* grid_v_p, const double* grid_t_p, const double* grid_ps_p, double* grid_t_eq,
    const double* grid_div, const double* grid_u, const double* grid_v, const double* grid_ps,
    double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t, const double* grid_vor,
    const double* grid_geopots, double* grid_M_half, double* grid_w_full, double* grid_du,
    double* grid_dv, double* grid_dps, double* grid_dt, double* geopot_half, double* geopot_full,
    double* grid_dlnps, double* grid_absvor, double* grid_energy, const double* grid_dlambda_t,
    const double* grid_dtheta_t);
/* No reference counterpart: jgcm_hs_physics_column_dynamics with two options.
 * dlambda_lnps, dtheta_lnps (both or neither): ∂λ ln ps and ∂θ ln ps are read
 * from them (the gradients of the step's current ln ps, carried from the
 * previous step's closing inverse batch) instead of from grid_dlambda_ps /
 * grid_dtheta_ps, which then only receive ∇ps (SD:260-261).  advance_p: the
 * p = c half of Time_Advance! (Dyn_Data.jl:346-362) for the grid u, v, T, ps
 * in the same pass — grid_u_p, grid_v_p, grid_t_p, grid_ps_p receive
 * grid_u, grid_v, grid_t, grid_ps once the forcing has read them; pass NULL
 * as their p pointers to jgcm_hs_corrections_advance afterwards.  Equal to
 * jgcm_hs_physics_column_dynamics (and the full advance) bit for bit.      */
int jgcm_hs_physics_column_dynamics_ex(
    jgcm_semi_implicit_t* si, const jgcm_hs_params* hp, int64_t dt, double* grid_u_p,
    double* grid_v_p, double* grid_t_p, double* grid_ps_p, double* grid_t_eq,
    const double* grid_div, const double* grid_u, const double* grid_v, const double* grid_ps,
    double* grid_dlambda_ps, double* grid_dtheta_ps, const double* grid_t, const double* grid_vor,
    const double* grid_geopots, double* grid_M_half, double* grid_w_full, double* grid_du,
    double* grid_dv, double* grid_dps, double* grid_dt, double* geopot_half, double* geopot_full,
    double* grid_dlnps, double* grid_absvor, double* grid_energy, const double* grid_dlambda_t,
    const double* grid_dtheta_t, const double* dlambda_lnps, const double* dtheta_lnps,
    int advance_p);
/* ps_n = exp(lnps) then Compute_Corrections!  Spectral_Dynamics.jl:336-343, :30-68 */
int jgcm_hs_corrections(jgcm_semi_implicit_t* si, const double* grid_lnps, double* grid_ps_n,
                        const double* grid_u_n, const double* grid_v_n, double* grid_t_n,
                        double* spe_lnps_n, double* spe_t_n);
/* Compute_Corrections! followed by Time_Advance! (Spectral_Dynamics.jl:336-346,
 * Dyn_Data.jl:326-364; no single reference function): the sums of the new
 * state, then one launch that applies the corrections to the new fields while
 * it copies p = c, c = n.  Fields in Time_Advance! order — spectral vor, div,
 * ln ps, T, then grid u, v, ps, T — as p/c/n pointer arrays with their sizes
 * in doubles; grid_lnps as for jgcm_hs_corrections.  Equal to
 * jgcm_hs_corrections followed by jgcm_time_advance, bit for bit.          */
int jgcm_hs_corrections_advance(jgcm_semi_implicit_t* si, const double* grid_lnps,
                                double* const* p, double* const* c, double* const* n,
                                const size_t* count);
/* No reference counterpart (a fusion): jgcm_hs_energy_implicit_leapfrog with
 * the p = c half of the spectral Time_Advance! (Dyn_Data.jl:328-340) in the
 * same pass.  vor, div, lnps, t are {p, c, n} arrays: on return p holds the
 * filtered current value and n the new one; c still holds the old value.
 * Pass NULL as the p pointers of these four spectral fields to
 * jgcm_hs_corrections_advance afterwards: for them it then only sets c = n,
 * with the (0, 0) corrections of T and ln ps (so for any field whose p is
 * NULL: the grid fields after jgcm_hs_physics_column_dynamics_ex with
 * advance_p).  The pair equals
 * jgcm_hs_energy_implicit_leapfrog + jgcm_hs_corrections_advance bit for bit.
 * Nothing may read the spectral c arrays between the two calls.            */
int jgcm_hs_implicit_leapfrog_advance(jgcm_semi_implicit_t* si, double* spe_energy, double* ddiv,
                                      double* dlnps, double* dt, double* dvor,
                                      double* const* vor, double* const* div,
                                      double* const* lnps, double* const* t);
/* device correction scalars to the host: mean_ps_p, mean_energy_p, mean_ps_n,
 * factor, log(factor), mean_energy_n, temperature correction (synchronises) */
int jgcm_hs_scalars(jgcm_semi_implicit_t* si, double* host_out, int count);
/* The same scalars copied to device memory on the mesh stream (no wait).    */
int jgcm_hs_scalars_device(jgcm_semi_implicit_t* si, double* dev_out, int count);

/* ---- diagnostics -------------------------------------------------------------
 * Update_Output!(output_manager, dyn_data, current_time)  Output_Manager.jl:81-107:
 * adds Zonal_Mean (:276-280) of T, T_eq, u, v ([nd][nθ] each) and ps
 * ([nθ][nλ]) into the given day's accumulators, on the device.              */
int jgcm_zonal_mean_accumulate(jgcm_mesh_t* mesh, const double* grid_t, const double* grid_t_eq,
                               const double* grid_u, const double* grid_v, const double* grid_ps,
                               int nd, double* acc_t, double* acc_t_eq, double* acc_u,
                               double* acc_v, double* acc_ps);

/* ---- graph capture of a sequence of calls on one mesh's stream ---------------
 * No reference counterpart (the reference has no device).  Capture records
 * every kernel the library enqueues on the mesh stream between begin and end;
 * replays re-run them with the argument values of capture time.  Scratch must
 * already be large enough (jgcm_mesh_reserve): growth during a capture fails
 * with JGCM_ERR_STATE, and growth while graphs exist keeps their buffers
 * until the mesh is freed.  No synchronising call (e.g.
 * jgcm_area_weighted_global_mean) may be made while capturing.  A graph holds
 * a reference to every mesh whose calls it recorded.                        */
int jgcm_mesh_reserve(jgcm_mesh_t* mesh, int nfields);
int jgcm_graph_begin(jgcm_mesh_t* mesh);
int jgcm_graph_end(jgcm_mesh_t* mesh, jgcm_graph_t** out);
int jgcm_graph_launch(jgcm_graph_t* graph);
int jgcm_graph_destroy(jgcm_graph_t* graph);
/* Order two meshes' streams (no reference counterpart): work enqueued on
 * `waiter` after this call starts after everything enqueued on `signaller`
 * before it.  Inside a capture begun on `signaller` it forks `waiter` into the
 * capture (its later calls become parallel branches of the same graph); the
 * reverse call joins the branch back before jgcm_graph_end.  No-op when both
 * meshes use one stream.                                                      */
int jgcm_mesh_wait(jgcm_mesh_t* waiter, jgcm_mesh_t* signaller);
/* Independent work within one step (no reference counterpart).  After
 * jgcm_mesh_branch(mesh, 1) the mesh's calls go to its side stream, which
 * first waits for everything already enqueued; jgcm_mesh_branch(mesh, 0)
 * sends later calls to the main stream again while the side work runs;
 * jgcm_mesh_join(mesh) makes the main stream wait for the side work (and
 * returns to it).  Inside a capture the side work becomes a parallel branch
 * of the graph, joined before jgcm_graph_end.  The caller guarantees that the
 * two branches touch disjoint memory (the mesh scratch belongs to the main
 * branch); jgcm_mesh_synchronize waits for both streams.                     */
int jgcm_mesh_branch(jgcm_mesh_t* mesh, int side);
int jgcm_mesh_join(jgcm_mesh_t* mesh);

#ifdef __cplusplus
}
#endif
#endif /* JGCM_H */
