// Internal interfaces between the Krylov driver (pcg.hip), the U-Net
// (unet.hip) and the multigrid hierarchy (mg.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bs {

constexpr unsigned kUNetPartWg = 1024;  // workgroups of the U-Net's reduction partials

// Workspace of one U-Net application on an nx*ny*nz grid (inference only).
size_t unet_ws_bytes(int nx, int ny, int nz);

// z = N(r) on fluid cells; n_fluid points to the grid's fluid-cell count in
// device memory.  rr_part (G partials of r.r) may be null.  Every kernel
// returns at once when *done != 0.
int unet_apply(hipStream_t st, void* ws, int nx, int ny, int nz, const float* params, const void* packed,
               const float* r, const uint8_t* code, const unsigned* n_fluid, const double* rr_part, unsigned G,
               float* z, const int* done);

// Multigrid-augmented preconditioner: a V-cycle on the fine grid and its 2x
// coarsening, with the 4x-coarsened level solved by the U-Net (coarse_kind 1,
// 2) or without it: one Jacobi sweep + 20 Chebyshev steps (0), + 60 (3), or
// two more geometric levels and a one-workgroup PCG bottom solve (4).
size_t mg_ws_bytes(int nx, int ny, int nz);
// Builds the coarse classes and codes from the fine mask (or from its fine
// class codes fcode when not null) and the coarsest level's fluid count, all
// on the device (asynchronous on st).
// deep: also the two levels below the 4x one (coarse kind 4; an error when
// the lowest has more than 16384 cells).
int mg_setup(hipStream_t st, void* ws, int nx, int ny, int nz, const int* mat, int mf, int ma, const uint8_t* fcode,
             bool deep = false);
// The flexible CG's residual update folded into the V-cycle's first fine
// pass: the pass reads r_old (mg_apply's r) and Ap, works on and writes
// r_out = r_old - (*alpha) Ap (r_out must not alias r_old) and one r.r
// partial per workgroup into rr_part (mg_fold_groups of them); `after` is
// enqueued right after that pass (the convergence check), and the rest of
// the cycle uses r_out.
struct MgFold {
    const float* ap;
    const double* alpha;
    float* r_out;
    double* rr_part;
    void (*after)(hipStream_t, void*);
    void* after_arg;
};
// Workgroups (r.r partials) of the folded first pass; 0 when the fine level
// cannot take it (nx, ny even and nz % 4 == 0 are needed).
unsigned mg_fold_groups(int nx, int ny, int nz);
// z = V-cycle(r); also writes partials of r.z and r.z_old (z_old may be null)
// into dots_part, mg_dots_groups(nx,ny,nz) workgroups of 2 values.
int mg_apply(hipStream_t st, void* ws, int nx, int ny, int nz, const uint8_t* code0, const float* r, float* z,
             const float* z_old, double* dots_part, int coarse_kind, const float* params, const void* packed,
             const int* done, const MgFold* fold = nullptr);
unsigned mg_dots_groups(int nx, int ny, int nz);

// e *= (r.e)/(e.A e) (0 when not positive): the A-norm-optimal multiple of a
// correction e for A x = r.  part: step_ws_bytes() of device scratch.
size_t step_ws_bytes();
void scale_to_optimal(hipStream_t st, int nx, int ny, int nz, const uint8_t* code, const float* r, float* e,
                      double* part, const int* done);

// Timing marks of the flexible CG's fine-level passes (diagnostic): while
// pass timing is on (bs_fcg_pass_timing), pass_mark(st, pass, 0) records a
// HIP event on st right before pass `pass` is enqueued and pass_mark(st,
// pass, 1) right after it (0: direction + SpMV, 1: the V-cycle's first fine
// pass with the folded residual update, 2: its last fine pass); otherwise it
// does nothing.
void pass_mark(hipStream_t st, int pass, int edge);

// Resolves (once per device, with it current) the current device's address
// of the fixed scratch array of bs_pcg_solve_noalloc and returns it through
// *out when out is not null.
int scratch_setup(void** out);

}  // namespace bs
