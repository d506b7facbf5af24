/*
 * libboilsim_hip.so — C ABI of the MI355X (gfx950) pressure projection.
 *
 * Plain pointers and sizes only.  Every pointer argument is a device address
 * in the calling process on the current HIP device; cell fields are
 * C-contiguous (nx, ny, nz) with k fastest, faces are ux (nx+1,ny,nz),
 * uy (nx,ny+1,nz), uz (nx,ny,nz+1) — the layout of the reference's Warp arrays
 * (python/boilingsim/geometry.py:254-282).  The library allocates no device
 * memory: the caller owns every buffer, as in the reference
 * (crates/cuda-kernels/src/pressure_solve_pcg.cu:82-88).  Return values are
 * hipError_t codes (0 = success), as the reference's launchers return
 * cudaError_t.  Entry points without a `stream` argument use the null stream,
 * like the reference's default-stream launchers.  The iterative solvers
 * (bs_pressure_solve_pcg_launch, bs_pcg_solve, bs_pcg_solve_noalloc,
 * bs_fcg_solve, bs_refine_solve, bs_fcg_refine_solve) share process-wide
 * polling and graph state and hold one lock for the whole call: calls from
 * several host threads run one after another.  A grid with a dimension <= 0
 * is a silent no-op, as in the reference's launchers
 * (crates/cuda-kernels/src/jacobi_pressure.cu:193, 238;
 * pressure_solve_pcg.cu:120-123): no buffer is touched except that the
 * statistics, iteration counts and device maxima (max|div|, max|u|) an entry
 * point reports read 0.
 */
#ifndef BOILSIM_HIP_H
#define BOILSIM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque stream handle; identical to hipStream_t. NULL = the null stream. */
typedef struct ihipStream_t* bs_stream_t;

/* ---- boundary helpers ---------------------------------------------------- */

/* ABI revision of this header. */
int bs_abi_version(void);
/* hipGetErrorString for a code returned by this library. */
const char* bs_error_string(int code);
/* hipSetDevice(device_id) and a one-line description (SimCore::new / device_info,
 * crates/sim-core/src/cuda.rs:25-40). */
int bs_device_open(int device_id, char* buf, int buflen);
int bs_device_synchronize(void);
/* In-place d_a[i] *= scale, synchronous.  Replaces scale_inplace_launch
 * (crates/cuda-kernels/src/scale.cu:22-34, lib.rs:23). */
int bs_scale_inplace_launch(float* d_a, int n, float scale);

/* ---- Jacobi pressure solve ------------------------------------------------- */

/* One masked 7-point Jacobi sweep, bit-identical to jacobi_pressure_step
 * (python/boilingsim/fluid.py:243-276).  Replaces jacobi_pressure_launch
 * (crates/cuda-kernels/src/jacobi_pressure.cu:184-208, lib.rs:32-45).
 * p_new and p_old must not overlap.  Asynchronous. */
int bs_jacobi_pressure_launch(float* p_new, const float* p_old, const float* div_u, const int* mat,
                              int nx, int ny, int nz, float dx, float dt, float rho, int mat_fluid,
                              int mat_air);
int bs_jacobi_pressure_launch_stream(float* p_new, const float* p_old, const float* div_u,
                                     const int* mat, int nx, int ny, int nz, float dx, float dt,
                                     float rho, int mat_fluid, int mat_air, bs_stream_t stream);

/* n_iter ping-pong sweeps; the result ends in p whatever the parity of n_iter,
 * and p_tmp ends as the reference leaves it.  Replaces pressure_solve_launch
 * (crates/cuda-kernels/src/jacobi_pressure.cu:228-270, lib.rs:52-66).
 * p and p_tmp must be disjoint.  Asynchronous. */
int bs_pressure_solve_launch(float* p, float* p_tmp, const float* div_u, const int* mat, int nx,
                             int ny, int nz, float dx, float dt, float rho, int n_iter,
                             int mat_fluid, int mat_air);
/* Same with a stream; zero_init != 0 promises p == 0 on entry (the
 * projection's cold start, fluid.py:627-629) so the first sweep skips reading
 * it. */
int bs_jacobi_solve_launch(float* p, float* p_tmp, const float* div_u, const int* mat, int nx,
                           int ny, int nz, float dx, float dt, float rho, int n_iter, int mat_fluid,
                           int mat_air, int zero_init, bs_stream_t stream);

/* Diagnostic (no counterpart in the reference): the Jacobi kernels' exact
 * division by 6, fma(x, RN(1/6), x * RN(1/6 - RN(1/6))) for finite
 * |x| >= 2^-96 and the IEEE divide otherwise, against x / 6.0f for all 2^32
 * float bit patterns on the device.  counts (three device uint64s) receives
 * the patterns where they differ, the patterns on the fma path, and the
 * patterns where the one-term x * RN(1/6) differs.  Asynchronous. */
int bs_div6_check(unsigned long long* counts, bs_stream_t stream);

/* ---- projection kernels (the Warp side of pressure_projection) ------------ */

/* compute_divergence (fluid.py:201-219). */
int bs_divergence_launch(float* div_u, const float* ux, const float* uy, const float* uz,
                         const int* mat, int nx, int ny, int nz, float dx, int mat_fluid,
                         bs_stream_t stream);
/* subtract_pressure_gradient_{x,y,z} (fluid.py:279-330), one launch. */
int bs_subtract_gradient_launch(float* ux, float* uy, float* uz, const float* p, const int* mat,
                                int nx, int ny, int nz, float dx, float dt, float rho, int mat_fluid,
                                bs_stream_t stream);
/* max|x| as float bits in *out_bits (device uint32); replaces the full-field
 * copy of fluid.py:733-734. */
int bs_max_abs_launch(const float* x, long long n, unsigned int* out_bits, bs_stream_t stream);
/* pressure_projection with the Jacobi solver (fluid.py:612-734): zero p/p_tmp,
 * divergence, n_iter sweeps, gradient, divergence, max|div| -> *max_div_bits. */
int bs_projection_jacobi_launch(float* p, float* p_tmp, float* div_u, float* ux, float* uy, float* uz,
                                const int* mat, int nx, int ny, int nz, float dx, float dt, float rho,
                                int n_iter, int mat_fluid, int mat_air, unsigned int* max_div_bits,
                                bs_stream_t stream);

/* ---- no-slip faces and the CFL velocity maximum --------------------------- */

/* enforce_no_slip (python/boilingsim/fluid.py:111-162): one launch for the
 * three enforce_no_slip_u{x,y,z} kernels.  Faces on the domain edge or next
 * to a non-fluid cell become +0.0; the others are not written. */
int bs_enforce_no_slip_launch(float* ux, float* uy, float* uz, const int* mat, int nx, int ny, int nz,
                              int mat_fluid, bs_stream_t stream);
/* enforce_no_slip then compute_divergence (pipeline.py:606-608 followed by
 * fluid.py:632-637) in one pass over the faces. */
int bs_divergence_no_slip_launch(float* div_u, float* ux, float* uy, float* uz, const int* mat, int nx,
                                 int ny, int nz, float dx, int mat_fluid, bs_stream_t stream);
/* compute_divergence plus, from the same pass, max|div| (fluid.py:727-734)
 * and max|u| over every face (compute_max_velocity, fluid.py:807-833), as
 * float bits in max_bits[0] and max_bits[1] (two device uint32s). */
int bs_divergence_max_launch(float* div_u, const float* ux, const float* uy, const float* uz,
                             const int* mat, int nx, int ny, int nz, float dx, int mat_fluid,
                             unsigned int* max_bits, bs_stream_t stream);
/* compute_max_velocity's device reduction (fluid.py:807-833: one zeroing
 * and three atomic-max launches there) in one launch; max|u| as float bits in
 * *out_bits (a device uint32). */
int bs_max_velocity_launch(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz,
                           unsigned int* out_bits, bs_stream_t stream);
/* One step's no-slip + projection + no-slip (pipeline.py:606-620) with the
 * Jacobi solver: the pre no-slip fused into the first divergence, the post
 * no-slip omitted (it would change no bit: the faces it zeroes were zeroed by
 * the pre pass and are never written by the solve or the gradient), and
 * max_bits[0] = max|div| after the correction (pressure_projection's return
 * value), max_bits[1] = max|u| of the final faces (what compute_max_velocity
 * returns after the step). */
int bs_projection_step_jacobi_launch(float* p, float* p_tmp, float* div_u, float* ux, float* uy,
                                     float* uz, const int* mat, int nx, int ny, int nz, float dx,
                                     float dt, float rho, int n_iter, int mat_fluid, int mat_air,
                                     unsigned int* max_bits, bs_stream_t stream);

/* ---- semi-Lagrangian advection -------------------------------------------- */

/* extend_temperature_into_solids (python/boilingsim/fluid.py:338-378): each
 * non-fluid cell takes T of its first fluid neighbour (order i-1, i+1, j-1,
 * j+1, k-1, k+1), fluid cells and isolated cells keep their own. */
int bs_extend_temperature_launch(float* t_ext, const float* t, const int* mat, int nx, int ny, int nz,
                                 int mat_fluid, bs_stream_t stream);
/* advect_ux, advect_uy, advect_uz and advect_temperature (fluid.py:381-492)
 * in one launch: face and cell values traced back along the trilinearly
 * sampled velocity, written to the *_new arrays (disjoint from the inputs);
 * t_ext from bs_extend_temperature_launch; origin (ox, oy, oz) and dx as
 * float32.  Grids whose face arrays exceed 2^31 elements are rejected. */
int bs_advect_launch(float* ux_new, float* uy_new, float* uz_new, float* t_new, const float* ux,
                     const float* uy, const float* uz, const float* t, const float* t_ext, const int* mat,
                     int nx, int ny, int nz, float ox, float oy, float oz, float dx, float dt, int mat_fluid,
                     bs_stream_t stream);
/* advect_all (fluid.py:737-775): extension, advection into the workspace
 * arrays (the reference's ws.ux_tmp, uy_tmp, uz_tmp, T_tmp, T_ext), and the
 * copies back into ux, uy, uz, t. */
int bs_advect_all_launch(float* ux, float* uy, float* uz, float* t, float* ux_tmp, float* uy_tmp, float* uz_tmp,
                         float* t_tmp, float* t_ext, const int* mat, int nx, int ny, int nz, float ox, float oy,
                         float oz, float dx, float dt, int mat_fluid, bs_stream_t stream);

/* ---- PCG ------------------------------------------------------------------ */

/* Jacobi-preconditioned CG with the reference's control flow and workspace
 * layout; FP32 vectors, FP64 dots/scalars.  Replaces pressure_solve_pcg_launch
 * (crates/cuda-kernels/src/pressure_solve_pcg.cu:89-118, lib.rs:131-153).
 * ws_* are n floats each, dot_workspace >= 1024 floats, the seven scalar slots
 * 1 float each (they receive float copies of the final scalars).  Synchronous:
 * returns after the solve with the iteration count in *iter_count_host. */
int bs_pressure_solve_pcg_launch(float* p_out, const float* div_u, const int* mat, int nx, int ny,
                                 int nz, float dx, float dt, float rho, int mat_fluid, int mat_air,
                                 float pressure_tol, int max_iter, float* ws_b, float* ws_r,
                                 float* ws_z, float* ws_p_search, float* ws_Ap, float* dot_workspace,
                                 float* dev_alpha, float* dev_beta, float* dev_rzold,
                                 float* dev_rznew, float* dev_bsq, float* dev_rsq, float* dev_pAp,
                                 int* iter_count_host);

/* Statistics of bs_pcg_solve. */
typedef struct {
    int iterations;
    int converged;
    int restarts;
    int reserved;
    double relres_recursive; /* sqrt(r.r / b.b) of the recursive residual */
    double relres_true;      /* ||b - A x||_2 / ||b||_2, FP64 accumulation */
    double bnorm;            /* ||b||_2 */
} bs_solve_stats;

/* Bytes of workspace bs_pcg_solve needs for an nx*ny*nz grid. */
size_t bs_pcg_workspace_bytes(int nx, int ny, int nz);
/* Jacobi-PCG to a TRUE relative residual tol (new; the reference tests only
 * its recursive FP32 residual, pressure_solve_pcg.cu:205-216): convergence is
 * tested every iteration on device, then ||b - A x|| is formed in FP64 and the
 * solve restarts from x (residual replacement) while it is above tol, at most
 * max_restarts times.  warm_start != 0 starts from x's contents. */
int bs_pcg_solve(float* x, const float* div_u, const int* mat, int nx, int ny, int nz, float dx,
                 float dt, float rho, int mat_fluid, int mat_air, double tol, int max_iter,
                 int warm_start, int max_restarts, void* workspace, size_t workspace_bytes,
                 bs_stream_t stream, bs_solve_stats* stats);
/* Bytes of the fixed device scratch array of bs_pcg_solve_noalloc (64 MiB). */
size_t bs_pcg_noalloc_scratch_bytes(void);
/* bs_pcg_solve without a workspace argument and without allocating anything:
 * no hipMalloc/hipFree, no pinned host memory, no events or graphs.  It has no
 * counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its workspace is one fixed array per device of
 * bs_pcg_noalloc_scratch_bytes() bytes, held from the library's load until
 * the process ends, so grids with bs_pcg_workspace_bytes(nx,ny,nz) above that
 * (about 3.9 M cells) are rejected with hipErrorInvalidValue before any work
 * is queued.  Every call on a device shares that device's array: calls that
 * overlap on one device (several host threads, or streams that run
 * concurrently) race.  Same arguments and results as bs_pcg_solve otherwise;
 * it blocks on a plain device-to-host copy of the done flag after each batch
 * of iterations. */
int bs_pcg_solve_noalloc(float* x, const float* div_u, const int* mat, int nx, int ny, int nz, float dx,
                         float dt, float rho, int mat_fluid, int mat_air, double tol, int max_iter,
                         int warm_start, int max_restarts, bs_stream_t stream, bs_solve_stats* stats);

/* ---- the reference PCG's building blocks, standalone ---------------------- */

/* Ap = 6p - S(p), non-fluid rows 0 (laplacian_spmv.cu:145-163). */
int bs_laplacian_spmv_launch(float* Ap, const float* p, const int* mat, int nx, int ny, int nz,
                             int mat_fluid, int mat_air);
/* z = r / (6 - #solid neighbours), z = r if that is <= 0, non-fluid 0
 * (diag_inverse_apply.cu:91-111). */
int bs_diag_inverse_apply_launch(float* z, const float* r, const int* mat, int nx, int ny, int nz,
                                 int mat_fluid, int mat_air);
/* *result = sum x*y accumulated in FP64 (dot_reduce.cu:93-115, which sums in
 * FP32); workspace >= 1024 floats. */
int bs_dot_launch(const float* x, const float* y, int n, float* workspace, float* result);
int bs_dot_launch_f64(const float* x, const float* y, int n, float* workspace, float* result,
                      double* result64);
/* axpy_device.cu:34-45, 66-77, 93-100, 126-138, 142-145. */
int bs_axpy_device_launch(float* y, const float* x, const float* alpha_ptr, int n, int negate_alpha);
int bs_scaled_axpy_launch(float* p_out, const float* z, const float* p_old, const float* beta_ptr,
                          int n);
int bs_divide_scalars_launch(const float* num_ptr, const float* den_ptr, float* result_ptr);
int bs_compute_b_launch(float* b, const float* div_u, const int* mat, int nx, int ny, int nz,
                        float scale, int mat_fluid);
int bs_zero_buffer_launch(float* buf, int n);


/* ---- learned preconditioner (no counterpart in the reference) -------------- */

/* A 3-level 3D conv U-Net (channels 8/16/32, 3x3x3, BF16 activations and
 * weights, FP32 accumulation on MFMA) mapping a residual r to a correction z;
 * the parameter layout is, per conv, W[27][cin][cout] then bias[cout] (FP32). */
size_t bs_unet_num_params(void);
/* Bytes of the packed BF16 weight fragments (forward and backward). */
size_t bs_unet_packed_bytes(void);
/* Workspace bytes for an nx*ny*nz grid; training != 0 adds the backward buffers. */
size_t bs_unet_workspace_bytes(int nx, int ny, int nz, int training);
/* Packs FP32 params into the MFMA fragment order (after every weight update). */
int bs_unet_pack(const float* params, void* packed, bs_stream_t stream);
/* One-byte class codes from the int32 mask: 0 air, 1 solid, 2 + #solid
 * neighbours for fluid (out of bounds = air). */
int bs_build_codes(const int* mat, int nx, int ny, int nz, int mat_fluid, int mat_air,
                   unsigned char* code, bs_stream_t stream);
/* z = rms(r) * U-Net(r / rms(r), classes) on fluid cells, 0 elsewhere;
 * n_fluid = number of fluid cells. */
int bs_unet_forward(const float* params, const void* packed, const float* r, const unsigned char* code,
                    int nx, int ny, int nz, double n_fluid, float* z, void* workspace,
                    size_t workspace_bytes, int training, bs_stream_t stream);
/* After bs_unet_forward(training = 1): accumulates into grads the gradient of
 * (z-e)^T A (z-e) / (e.r) for the target correction e (A e = r); loss_dev
 * (2 device doubles) receives the numerator and e.r. */
int bs_unet_backward(const float* params, const void* packed, const float* r, const float* e,
                     const unsigned char* code, int nx, int ny, int nz, float* grads,
                     double* loss_dev, void* workspace, size_t workspace_bytes, bs_stream_t stream);
/* Statistics of bs_fcg_solve. */
typedef struct {
    int iterations;     /* flexible-CG iterations */
    int converged;      /* true residual <= tol at exit */
    int fallback;       /* the flexible CG stalled and plain PCG finished the solve */
    int pcg_iterations; /* plain PCG iterations after a stall or residual drift */
    int restarts;
    int reserved;
    double relres_recursive;
    double relres_true; /* ||b - A x||_2 / ||b||_2, FP64 accumulation */
    double bnorm;
} bs_fcg_stats;

/* Workspace bytes of bs_fcg_solve for preconditioner kind `precond`. */
size_t bs_fcg_workspace_bytes(int nx, int ny, int nz, int precond);
/* Flexible CG (Polak-Ribiere beta) with a learned preconditioner, to a true
 * relative residual tol: precond 1 = the U-Net on this grid, 2 = a V-cycle on
 * this grid and its 2x coarsening (smoothers: Chebyshev polynomials in the
 * Jacobi splitting) with the U-Net solving the 4x-coarsened grid, its
 * correction polished by Chebyshev steps, 3 = the same V-cycle with one
 * Jacobi sweep and the same polish on the coarsest grid (no network), 4 = as
 * 2 with Jacobi sweeps before the network, which then corrects their
 * residual, 5 = as 3 with 60 Chebyshev steps (no network; more work than the
 * network and its polish), 6 = as 2 with two more geometric levels below the
 * coarsest grid and a one-workgroup Jacobi-PCG bottom solve in place of the
 * network (no network; an error when the lowest level exceeds 16384 cells).
 * unet_params and unet_packed may be null for kinds 3, 5 and 6.  A stall (no
 * 2 % residual drop in 20 iterations) hands over to plain Jacobi-PCG from the
 * current x. */
int bs_fcg_solve(float* x, const float* div_u, const int* mat, int nx, int ny, int nz, float dx, float dt,
                 float rho, int mat_fluid, int mat_air, double tol, int max_iter, int warm_start,
                 int precond, const float* unet_params, const void* unet_packed, void* workspace,
                 size_t workspace_bytes, bs_stream_t stream, bs_fcg_stats* stats);
/* Pass timing of the learned flexible CG (diagnostic; no counterpart in the
 * reference).  enable != 0 clears the records and turns timing on: later
 * bs_fcg_solve / bs_fcg_refine_solve calls launch their kernels directly (no
 * graphs) and record HIP events on their stream around each fine-level pass
 * (0: fcg_spmv_dir_kernel — direction, SpMV, p.Ap and the deferred x step;
 * 1: pre2_restrict_kernel<true> — the V-cycle's first fine pass with the
 * folded residual update; 2: prolong_post_kernel — its last fine pass).
 * enable == 0 waits for the records, turns timing off and writes per pass the
 * median and mean launch time in microseconds and the number of launches
 * timed (arrays of 3; null pointers are skipped).  Not thread-safe. */
int bs_fcg_pass_timing(int enable, double* us_median, double* us_mean, int* counts);
/* The V-cycle's first pass on one level (diagnostic; no counterpart in the
 * reference): z = two smoothing steps from zero on A z = r over class codes
 * `code` (0 air, 1 solid, 2 + solid neighbours for fluid) — damped Jacobi
 * (w = 0.8) for lo <= 0, the two-step Chebyshev polynomial over [lo, 2] of
 * D^-1 A otherwise (the V-cycle uses lo = 0.3) — and rc[I,J,K] = 0.5 x the sum
 * of r - A z over the fluid children of coarse cell (I,J,K), coarse dims
 * ((nx+1)/2, (ny+1)/2, (nz+1)/2).  fused = 1 runs the single-pass kernel the
 * V-cycle uses where the shape allows (nx, ny even, nz % 4 == 0, 16-byte
 * aligned vectors; otherwise an error) on the 256-thread tiles of its 2x
 * level, fused = 2 the same kernel on the taller tiles of its fine level,
 * fused = 0 the two-kernel path.
 * Asynchronous on `stream`. */
int bs_mg_smooth_restrict(const unsigned char* code, const float* r, float* z, float* rc, int nx, int ny, int nz,
                          int fused, double lo, bs_stream_t stream);
/* The V-cycle's last pass on the fine level (diagnostic; no counterpart in the
 * reference): zt = z_in + P ec on fluid cells (P the cell-centred trilinear
 * prolongation, weights 3/4 and 1/4 per axis, from the coarse grid of
 * ((nx+1)/2, (ny+1)/2, (nz+1)/2) cells; z_in elsewhere), then z_out = zt +
 * w D^-1 (r - A zt) on fluid cells (0 elsewhere); part[0..G) receives per
 * workgroup partials of r.z_out and part[G..2G) of r.z_old (z_old may be
 * null), *groups = G.  fused = 1 runs the single-pass kernel the V-cycle uses
 * (nz % 4 == 0, 16-byte aligned vectors; otherwise an error) on 256-thread
 * tiles, fused = 2 on the taller tiles of the V-cycle's fine level, fused = 0 the
 * prolongation and the sweep as two kernels through tmp (nx*ny*nz floats).
 * z_out must not alias z_in; an empty grid writes *groups = 0 and nothing
 * else.  Asynchronous on `stream`. */
int bs_mg_prolong_post(const unsigned char* code, const float* z_in, const float* ec, const float* r,
                       const float* z_old, float* z_out, float* tmp, double* part, size_t part_len, int* groups,
                       int nx, int ny, int nz, int fused, double w, bs_stream_t stream);
/* `steps` (even, > 0) Chebyshev steps over [lo, 2] of D^-1 A on A x = r from
 * x, the result in x (tmp: device scratch of nx*ny*nz floats): the V-cycle's
 * post-smoother (4 steps, lo = 0.02) and coarsest-level polish (20 steps,
 * lo = 0.02).  Diagnostic; no counterpart in the reference. */
int bs_mg_chebyshev(const unsigned char* code, const float* r, float* x, float* tmp, int nx, int ny, int nz,
                    int steps, double lo, bs_stream_t stream);
/* ---- Mixed-precision iterative refinement ------------------------------
 * No single counterpart in the reference: it wraps the Jacobi-PCG of
 * crates/cuda-kernels/src/pressure_solve_pcg.cu:89-243 (same A, M and b) in
 * an FP64 outer loop, x_{k+1} = x_k + PCG(b - A x_k), with the inner PCG's
 * vectors stored in FP32, FP16 or BF16. */
typedef struct {
    int outer;              /* refinement steps taken */
    int inner_iterations;   /* inner PCG iterations, all steps */
    int converged;          /* FP64 true relative residual <= tol */
    int storage;            /* 0 FP32, 1 FP16, 2 BF16 */
    double relres_true;     /* ||b - A x||_2 / ||b||_2 of the FP64 iterate */
    double bnorm;           /* ||b||_2 */
    double relres_inner_last; /* recursive relative residual of the last inner solve */
    int fallback;           /* bs_fcg_refine_solve: an inner flexible CG stalled and plain PCG took over */
    int pcg_iterations;     /* bs_fcg_refine_solve: inner plain-PCG iterations after that */
} bs_refine_stats;

size_t bs_refine_workspace_bytes(int nx, int ny, int nz);
/* x (FP64, nx*ny*nz) solves A x = b to a true relative residual tol, the
 * residual formed in FP64; each step runs the inner PCG to a recursive
 * relative residual inner_tol (at most max_inner iterations).  warm_start uses
 * x's contents as the initial guess; x32 (nullable) receives x rounded to
 * FP32. */
int bs_refine_solve(double* x, float* x32, const float* div_u, const int* mat, int nx, int ny, int nz,
                    float dx, float dt, float rho, int mat_fluid, int mat_air, double tol, int max_outer,
                    int storage, double inner_tol, int max_inner, int warm_start, void* workspace,
                    size_t workspace_bytes, bs_stream_t stream, bs_refine_stats* stats);

size_t bs_fcg_refine_workspace_bytes(int nx, int ny, int nz, int precond);
/* The same refinement with bs_fcg_solve's flexible CG (preconditioner
 * `precond`, FP32 vectors) as the inner solver, cold-started on each scaled
 * residual and run until its recursive residual drops by inner_tol; at most
 * max_inner inner iterations in all, as in bs_refine_solve.  When an inner
 * flexible CG stalls (bs_fcg_solve's test), the remaining steps run
 * bs_refine_solve's plain Jacobi-PCG inner solve (FP32 vectors) instead, as
 * bs_fcg_solve hands a stalled solve over to plain PCG. */
int bs_fcg_refine_solve(double* x, float* x32, const float* div_u, const int* mat, int nx, int ny, int nz,
                        float dx, float dt, float rho, int mat_fluid, int mat_air, double tol, int max_outer,
                        double inner_tol, int max_inner, int warm_start, int precond, const float* unet_params,
                        const void* unet_packed, void* workspace, size_t workspace_bytes, bs_stream_t stream,
                        bs_refine_stats* stats);

/* ---- Backward-Euler Jacobi heat conduction --------------------------------- */

/* One sweep of apply_implicit_conduction_update
 * (python/boilingsim/thermal.py:259-388): per non-air cell
 * T_dst = (T_old + sum g*T_nbr)/(1 + sum g) with the harmonic face
 * conductivity between materials and h_conv_air*dx at air faces; air cells
 * are not written.  k_arr, rho_arr, cp_arr: n_mat (<= 16) floats on the
 * device, indexed by material id.  Bit-identical to the reference arithmetic
 * with each float operation rounded separately.  Asynchronous. */
int bs_conduction_sweep_launch(float* t_dst, const float* t_iter, const float* t_old, const int* mat,
                               const float* k_arr, const float* rho_arr, const float* cp_arr, int n_mat,
                               int nx, int ny, int nz, float dx, float dt, float h_conv_air, float t_amb_k,
                               int mat_air, int mat_fluid, bs_stream_t stream);
/* The backward-Euler branch of conduct_one_step (thermal.py:648-665):
 * t_old = t, n_iter sweeps ping-ponging between t and t_alt, the result in t
 * (an odd count copies t_alt to t, as the reference).  Asynchronous. */
int bs_conduction_solve_launch(float* t, float* t_alt, float* t_old, const int* mat, const float* k_arr,
                               const float* rho_arr, const float* cp_arr, int n_mat, int nx, int ny, int nz,
                               float dx, float dt, float h_conv_air, float t_amb_k, int n_iter, int mat_air,
                               int mat_fluid, bs_stream_t stream);

/* Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 params; gscale
 * multiplies the gradient (1/batch). */
int bs_unet_adam(float* params, const float* grads, float* m, float* v, int step, float lr,
                 float gscale, bs_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* BOILSIM_HIP_H */
