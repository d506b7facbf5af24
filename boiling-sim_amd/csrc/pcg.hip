// Jacobi-preconditioned conjugate gradient for the masked pressure Poisson
// system on gfx950, FP32 vectors with FP64 scalars.
//
// System (crates/cuda-kernels/src/pressure_solve_pcg.cu:89-243 and the
// kernels it drives): A p = b with A = 6I - S (S sums the six neighbours
// under the _pressure_neighbour rules of python/boilingsim/fluid.py:222-240,
// non-fluid rows are zero; laplacian_spmv.cu:60-143), M = diag(6 - #solid
// neighbours) with z = r where that is <= 0 (diag_inverse_apply.cu:63-89),
// b = (-rho*dx*dx/dt) * div on fluid cells (pressure_solve_pcg.cu:137,
// axpy_device.cu:104-123).  Out-of-bounds neighbours are air.
//
// Per iteration, three kernels and no host round trip:
//   spmv_dot : Ap = A p and p.Ap; the last workgroup forms alpha   (9+1 B/cell)
//   update   : x += alpha p, r -= alpha Ap, r.r and r.z with z = r/diag; the
//              last workgroup tests convergence and forms beta     (24+1 B/cell)
//   direction: p = r/diag + beta p                                 (12+1 B/cell)
// The cell class and the Jacobi diagonal live in one byte per cell (the
// `code` array, built once per solve from the int32 mask), so no kernel reads
// the 4-byte mask inside the loop.  Dots are FP64 single-pass grid reductions
// (reduce.hpp).  Every kernel first reads the device `done` flag and returns
// at once when the solve has finished, so the host can queue iterations in
// batches and poll the flag asynchronously instead of syncing every few
// iterations (the reference does a blocking copy every 5 iterations,
// pressure_solve_pcg.cu:205-216).
#include "boilsim_hip.h"
#include "reduce.hpp"

#include <math.h>
#include <stddef.h>
#include <string.h>

namespace bs {

// Device control block of one solve (lives in caller-provided memory).
struct PcgCtrl {
    double bsq;    // b.b
    double stop;   // tol^2 * bsq
    double rz;     // r.z of the current iterate
    double alpha;
    double beta;
    double pAp;
    double rr;     // r.r after the last update
    double rz_new;
    double rr_true;  // ||b - A x||^2 from the last residual pass
    unsigned int tk[4];  // reduction tickets: init, spmv, update, aux
    int done;
    int iter;
    int max_iter;
    int mode;      // 0: reference cadence (iter >= 8 && iter % 5 == 0), 1: every iteration
    int converged;
    int pad[3];
};
static_assert(sizeof(PcgCtrl) <= 256, "control block must fit 256 bytes");

// ---------------------------------------------------------------------------
// Tile mapping.  A tile is 256 threads x VEC consecutive cells of one (j,k)
// plane chunk at one i.  Workgroups of a grid-stride kernel take contiguous
// runs of tiles ordered i-fastest, so a workgroup walks down i for a fixed
// chunk and its i-1 neighbour plane was loaded by itself one tile earlier.
// ---------------------------------------------------------------------------
struct Tiles {
    int nx, ny, nz;
    FastDiv nzv;        // nz / VEC
    uint32_t chunks;    // 256-thread chunks per plane
    uint32_t ntiles;    // chunks * nx
};

__device__ __forceinline__ void tile_span(uint32_t ntiles, uint32_t& t0, uint32_t& t1) {
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x);
    t0 = (uint32_t)(((uint64_t)w * ntiles) / gridDim.x);
    t1 = (uint32_t)(((uint64_t)(w + 1) * ntiles) / gridDim.x);
}

// Tile t -> (i, j, k0) for this thread; false when the thread is past the plane.
template <int VEC>
__device__ __forceinline__ bool tile_cell(const Tiles& T, uint32_t t, int& i, int& j, int& k0) {
    const uint32_t chunk = t / (uint32_t)T.nx;
    i = (int)(t - chunk * (uint32_t)T.nx);
    const uint32_t q = chunk * 256u + threadIdx.x;
    if (q >= (uint32_t)T.ny * T.nzv.d) return false;
    j = (int)fdiv(q, T.nzv);
    k0 = (int)(q - (uint32_t)j * T.nzv.d) * VEC;
    return true;
}

template <int VEC>
__device__ __forceinline__ void ldf(const float* __restrict__ p, int64_t off, float (&v)[VEC]) {
    if constexpr (VEC == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p + off);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else if constexpr (VEC == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p + off);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = p[off];
    }
}

template <int VEC>
__device__ __forceinline__ void stf(float* __restrict__ p, int64_t off, const float (&v)[VEC]) {
    if constexpr (VEC == 4) {
        *reinterpret_cast<float4*>(p + off) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (VEC == 2) {
        *reinterpret_cast<float2*>(p + off) = make_float2(v[0], v[1]);
    } else {
        p[off] = v[0];
    }
}

template <int VEC>
__device__ __forceinline__ void ldc(const uint8_t* __restrict__ p, int64_t off, uint8_t (&v)[VEC]) {
    if constexpr (VEC == 4) {
        const uint32_t t = *reinterpret_cast<const uint32_t*>(p + off);
        v[0] = t & 0xFF; v[1] = (t >> 8) & 0xFF; v[2] = (t >> 16) & 0xFF; v[3] = t >> 24;
    } else if constexpr (VEC == 2) {
        const uint16_t t = *reinterpret_cast<const uint16_t*>(p + off);
        v[0] = t & 0xFF; v[1] = t >> 8;
    } else {
        v[0] = p[off];
    }
}

template <int VEC>
__device__ __forceinline__ void stc(uint8_t* __restrict__ p, int64_t off, const uint8_t (&v)[VEC]) {
    if constexpr (VEC == 4) {
        *reinterpret_cast<uint32_t*>(p + off) =
            (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
    } else if constexpr (VEC == 2) {
        *reinterpret_cast<uint16_t*>(p + off) = (uint16_t)(v[0] | (v[1] << 8));
    } else {
        p[off] = v[0];
    }
}

// z = M^-1 r for one cell from its code (diag_inverse_apply.cu:63-89).
__device__ __forceinline__ float precond(float r, uint8_t code) {
    if (code < kCodeFluid0) return 0.0f;
    const int dg = 8 - (int)code;  // 6 - n_solid
    return dg > 0 ? r / (float)dg : r;
}

// Neighbour term of the SpMV from the neighbour's code (fluid.py:222-240).
__device__ __forceinline__ float nbr_term(uint8_t c, float v, float ps) {
    return c >= kCodeFluid0 ? v : (c == kCodeSolid ? ps : 0.0f);
}

__device__ __forceinline__ int is_solid_m(int m, int mf, int ma) { return (m != mf && m != ma) ? 1 : 0; }

// ---------------------------------------------------------------------------
// Set-up: class codes, b, and (cold start) x = 0, r = b, p = z = M^-1 r, with
// b.b and r.z reduced.  Reads the int32 mask once (+ neighbours via cache).
// ---------------------------------------------------------------------------
template <int VEC>
__global__ __launch_bounds__(256) void pcg_setup_kernel(
    Tiles T, const int* __restrict__ mat, const float* __restrict__ div, float bscale, int mf, int ma,
    int cold, uint8_t* __restrict__ code, float* __restrict__ b, float* __restrict__ x,
    float* __restrict__ r, float* __restrict__ p, double* __restrict__ partials, PcgCtrl* ctrl) {
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double bb = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        int m[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) m[e] = mat[c0 + e];
        float dv[VEC], bv[VEC], zv[VEC], zero[VEC];
        uint8_t cv[VEC];
        ldf<VEC>(div, c0, dv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            zero[e] = 0.0f;
            const int k = k0 + e;
            if (m[e] != mf) {
                cv[e] = (m[e] == ma) ? kCodeAir : kCodeSolid;
                bv[e] = 0.0f;
                zv[e] = 0.0f;
                continue;
            }
            int ns = 0;
            if (i > 0) ns += is_solid_m(mat[c0 + e - plane], mf, ma);
            if (i < T.nx - 1) ns += is_solid_m(mat[c0 + e + plane], mf, ma);
            if (j > 0) ns += is_solid_m(mat[c0 + e - T.nz], mf, ma);
            if (j < T.ny - 1) ns += is_solid_m(mat[c0 + e + T.nz], mf, ma);
            if (k > 0) ns += is_solid_m(mat[c0 + e - 1], mf, ma);
            if (k < T.nz - 1) ns += is_solid_m(mat[c0 + e + 1], mf, ma);
            cv[e] = (uint8_t)(kCodeFluid0 + ns);
            bv[e] = bscale * dv[e];
            zv[e] = precond(bv[e], cv[e]);
            bb += (double)bv[e] * (double)bv[e];
            rz += (double)bv[e] * (double)zv[e];
        }
        stc<VEC>(code, c0, cv);
        stf<VEC>(b, c0, bv);
        if (cold) {
            stf<VEC>(x, c0, zero);
            stf<VEC>(r, c0, bv);
            stf<VEC>(p, c0, zv);
        }
    }
    double v[2] = {bb, rz};
    if (grid_reduce<2>(v, partials, &ctrl->tk[0]) && threadIdx.x == 0) {
        ctrl->bsq = v[0];
        ctrl->stop = ctrl->stop * v[0];  // host stored tol^2 in stop
        if (cold) {
            ctrl->rz = v[1];
            if (v[0] == 0.0 || ctrl->max_iter <= 0) ctrl->done = 1;
        }
    }
}

// r = b - A x formed in FP64 (stored rounded to FP32), r.r and r.z reduced;
// with set_direction also p = z = M^-1 r (warm start / residual replacement).
template <int VEC>
__global__ __launch_bounds__(256) void pcg_residual_kernel(
    Tiles T, const uint8_t* __restrict__ code, const float* __restrict__ b, const float* __restrict__ x,
    float* __restrict__ r, float* __restrict__ p, double* __restrict__ partials, PcgCtrl* ctrl,
    int set_direction) {
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC], cl[VEC], cr[VEC], cd[VEC], cu[VEC];
        float xc[VEC], xl[VEC], xr[VEC], xd[VEC], xu[VEC], bv[VEC], rv[VEC], zv[VEC];
        ldc<VEC>(code, c0, cc);
        ldf<VEC>(x, c0, xc);
        ldf<VEC>(b, c0, bv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) { cl[e] = cr[e] = cd[e] = cu[e] = kCodeAir; xl[e] = xr[e] = xd[e] = xu[e] = 0.0f; }
        if (i > 0) { ldc<VEC>(code, c0 - plane, cl); ldf<VEC>(x, c0 - plane, xl); }
        if (i < T.nx - 1) { ldc<VEC>(code, c0 + plane, cr); ldf<VEC>(x, c0 + plane, xr); }
        if (j > 0) { ldc<VEC>(code, c0 - T.nz, cd); ldf<VEC>(x, c0 - T.nz, xd); }
        if (j < T.ny - 1) { ldc<VEC>(code, c0 + T.nz, cu); ldf<VEC>(x, c0 + T.nz, xu); }
        uint8_t cb = kCodeAir, cf = kCodeAir;
        float xb = 0.0f, xf = 0.0f;
        if (k0 > 0) { cb = code[c0 - 1]; xb = x[c0 - 1]; }
        if (k0 + VEC < T.nz) { cf = code[c0 + VEC]; xf = x[c0 + VEC]; }
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (cc[e] < kCodeFluid0) { rv[e] = 0.0f; zv[e] = 0.0f; continue; }
            const float ps = xc[e];
            const uint8_t c_b = e > 0 ? cc[e - 1] : cb;
            const float v_b = e > 0 ? xc[e - 1] : xb;
            const uint8_t c_f = e < VEC - 1 ? cc[e + 1] : cf;
            const float v_f = e < VEC - 1 ? xc[e + 1] : xf;
            // FP64: the true residual must not inherit FP32 cancellation error.
            double s = (double)nbr_term(cl[e], xl[e], ps) + (double)nbr_term(cr[e], xr[e], ps);
            s += (double)nbr_term(cd[e], xd[e], ps);
            s += (double)nbr_term(cu[e], xu[e], ps);
            s += (double)nbr_term(c_b, v_b, ps);
            s += (double)nbr_term(c_f, v_f, ps);
            const double r64 = (double)bv[e] - (6.0 * (double)ps - s);
            rv[e] = (float)r64;
            zv[e] = precond(rv[e], cc[e]);
            rr += r64 * r64;
            rz += (double)rv[e] * (double)zv[e];
        }
        stf<VEC>(r, c0, rv);
        if (set_direction) stf<VEC>(p, c0, zv);
    }
    double v[2] = {rr, rz};
    if (grid_reduce<2>(v, partials, &ctrl->tk[3]) && threadIdx.x == 0) {
        ctrl->rr_true = v[0];
        if (set_direction) {
            ctrl->rr = v[0];
            ctrl->rz = v[1];
            if (v[0] <= ctrl->stop || ctrl->max_iter <= 0) ctrl->done = 1;
        }
    }
}

// Ap = A p, p.Ap -> alpha = rz / pAp (0 if pAp == 0, axpy_device.cu:82-91).
template <int VEC>
__global__ __launch_bounds__(256) void pcg_spmv_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                       const float* __restrict__ p, float* __restrict__ ap,
                                                       double* __restrict__ partials, PcgCtrl* ctrl) {
    if (ctrl->done) return;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double pap = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC], cl[VEC], cr[VEC], cd[VEC], cu[VEC];
        float pc[VEC], pl[VEC], pr[VEC], pd[VEC], pu[VEC], out[VEC];
        ldc<VEC>(code, c0, cc);
        ldf<VEC>(p, c0, pc);
#pragma unroll
        for (int e = 0; e < VEC; ++e) { cl[e] = cr[e] = cd[e] = cu[e] = kCodeAir; pl[e] = pr[e] = pd[e] = pu[e] = 0.0f; }
        if (i > 0) { ldc<VEC>(code, c0 - plane, cl); ldf<VEC>(p, c0 - plane, pl); }
        if (i < T.nx - 1) { ldc<VEC>(code, c0 + plane, cr); ldf<VEC>(p, c0 + plane, pr); }
        if (j > 0) { ldc<VEC>(code, c0 - T.nz, cd); ldf<VEC>(p, c0 - T.nz, pd); }
        if (j < T.ny - 1) { ldc<VEC>(code, c0 + T.nz, cu); ldf<VEC>(p, c0 + T.nz, pu); }
        uint8_t cb = kCodeAir, cf = kCodeAir;
        float vb = 0.0f, vf = 0.0f;
        if (k0 > 0) { cb = code[c0 - 1]; vb = p[c0 - 1]; }
        if (k0 + VEC < T.nz) { cf = code[c0 + VEC]; vf = p[c0 + VEC]; }
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (cc[e] < kCodeFluid0) { out[e] = 0.0f; continue; }
            const float ps = pc[e];
            const uint8_t c_b = e > 0 ? cc[e - 1] : cb;
            const float v_b = e > 0 ? pc[e - 1] : vb;
            const uint8_t c_f = e < VEC - 1 ? cc[e + 1] : cf;
            const float v_f = e < VEC - 1 ? pc[e + 1] : vf;
            float s = nbr_term(cl[e], pl[e], ps) + nbr_term(cr[e], pr[e], ps);
            s = s + nbr_term(cd[e], pd[e], ps);
            s = s + nbr_term(cu[e], pu[e], ps);
            s = s + nbr_term(c_b, v_b, ps);
            s = s + nbr_term(c_f, v_f, ps);
            out[e] = 6.0f * ps - s;
            pap += (double)ps * (double)out[e];
        }
        stf<VEC>(ap, c0, out);
    }
    double v[1] = {pap};
    if (grid_reduce<1>(v, partials, &ctrl->tk[1]) && threadIdx.x == 0) {
        ctrl->pAp = v[0];
        ctrl->alpha = v[0] != 0.0 ? ctrl->rz / v[0] : 0.0;
    }
}

// x += alpha p, r -= alpha Ap; r.r and r.z (z = M^-1 r); convergence test and
// beta = rz_new / rz (the reference's iteration body, pressure_solve_pcg.cu:
// 192-238).  In `mode` 0 the test runs only when iter >= 8 && iter % 5 == 0
// and a converged solve reports iter + 1 (pressure_solve_pcg.cu:205-216).
template <int VEC>
__global__ __launch_bounds__(256) void pcg_update_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                         const float* __restrict__ p,
                                                         const float* __restrict__ ap,
                                                         float* __restrict__ x, float* __restrict__ r,
                                                         double* __restrict__ partials, PcgCtrl* ctrl) {
    if (ctrl->done) return;
    const float af = (float)ctrl->alpha;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC];
        float pv[VEC], av[VEC], xv[VEC], rv[VEC];
        ldc<VEC>(code, c0, cc);
        ldf<VEC>(p, c0, pv);
        ldf<VEC>(ap, c0, av);
        ldf<VEC>(x, c0, xv);
        ldf<VEC>(r, c0, rv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            xv[e] = xv[e] + af * pv[e];
            rv[e] = rv[e] - af * av[e];
            const float z = precond(rv[e], cc[e]);
            rr += (double)rv[e] * (double)rv[e];
            rz += (double)rv[e] * (double)z;
        }
        stf<VEC>(x, c0, xv);
        stf<VEC>(r, c0, rv);
    }
    double v[2] = {rr, rz};
    if (grid_reduce<2>(v, partials, &ctrl->tk[2]) && threadIdx.x == 0) {
        const int it = ctrl->iter;
        ctrl->rr = v[0];
        const bool check = ctrl->mode != 0 || (it >= 8 && (it % 5) == 0);
        if (check && v[0] < ctrl->stop) {
            ctrl->iter = it + 1;
            ctrl->converged = 1;
            ctrl->done = 1;
        } else {
            ctrl->rz_new = v[1];
            ctrl->beta = ctrl->rz != 0.0 ? v[1] / ctrl->rz : 0.0;
            ctrl->rz = v[1];
            ctrl->iter = it + 1;
            if (it + 1 >= ctrl->max_iter) ctrl->done = 1;
        }
    }
}

// p = M^-1 r + beta p.  Runs after update; when update just finished the
// solve, `done` is set and this is skipped.
template <int VEC>
__global__ __launch_bounds__(256) void pcg_direction_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                            const float* __restrict__ r,
                                                            float* __restrict__ p, PcgCtrl* ctrl) {
    if (ctrl->done) return;
    const float bf = (float)ctrl->beta;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC];
        float rv[VEC], pv[VEC];
        ldc<VEC>(code, c0, cc);
        ldf<VEC>(r, c0, rv);
        ldf<VEC>(p, c0, pv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) pv[e] = precond(rv[e], cc[e]) + bf * pv[e];
        stf<VEC>(p, c0, pv);
    }
}

__global__ void pcg_ctrl_init_kernel(PcgCtrl* ctrl, double tol2, int max_iter, int mode) {
    if (threadIdx.x != 0) return;
    PcgCtrl c;
    memset(&c, 0, sizeof(c));
    c.stop = tol2;
    c.max_iter = max_iter;
    c.mode = mode;
    *ctrl = c;
}

__global__ void pcg_clear_done_kernel(PcgCtrl* ctrl) {
    if (threadIdx.x == 0) ctrl->done = 0;
}

// Float copies of the final scalars into the reference's 1-float slots.
__global__ void pcg_export_kernel(const PcgCtrl* ctrl, float* alpha, float* beta, float* rzold,
                                  float* rznew, float* bsq, float* rsq, float* pAp) {
    if (threadIdx.x != 0) return;
    if (alpha) *alpha = (float)ctrl->alpha;
    if (beta) *beta = (float)ctrl->beta;
    if (rzold) *rzold = (float)ctrl->rz;
    if (rznew) *rznew = (float)ctrl->rz_new;
    if (bsq) *bsq = (float)ctrl->bsq;
    if (rsq) *rsq = (float)ctrl->rr;
    if (pAp) *pAp = (float)ctrl->pAp;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct PcgBuffers {
    uint8_t* code;
    float *b, *x, *r, *p, *ap;
    double* partials;
    PcgCtrl* ctrl;
    unsigned nwg;
};

static int pcg_vec(int nz, std::initializer_list<const void*> ptrs) {
    int vec = (nz % 4 == 0) ? 4 : (nz % 2 == 0) ? 2 : 1;
    for (const void* q : ptrs) {
        const uintptr_t a = (uintptr_t)q;
        while (vec > 1 && (a % (uintptr_t)(4 * vec)) != 0) vec >>= 1;
    }
    return vec;
}

static Tiles make_tiles(int nx, int ny, int nz, int vec) {
    Tiles T;
    T.nx = nx; T.ny = ny; T.nz = nz;
    const uint32_t nzv = (uint32_t)(nz / vec);
    T.nzv = make_fastdiv(nzv);
    T.chunks = (uint32_t)(((uint64_t)ny * nzv + 255) / 256);
    T.ntiles = T.chunks * (uint32_t)nx;
    return T;
}

constexpr unsigned kMaxWg = 2048;

// Pinned host mirror of (done, iter) for asynchronous polling, one per process.
struct HostPoll {
    int* slots = nullptr;     // [2][2]: (done, iter) of the last two batches
    PcgCtrl* final = nullptr; // full control block after the solve
};
static HostPoll g_poll;

static int ensure_poll() {
    if (g_poll.slots) return hipSuccess;
    void* h = nullptr;
    BS_RETURN_IF_ERR(hipHostMalloc(&h, 512, hipHostMallocDefault));
    g_poll.slots = (int*)h;
    g_poll.final = (PcgCtrl*)((char*)h + 256);
    return hipSuccess;
}

#define BS_PCG_LAUNCH(KERNEL, VEC, GRID, ST, ...)                                   \
    do {                                                                           \
        if ((VEC) == 4) KERNEL<4><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);          \
        else if ((VEC) == 2) KERNEL<2><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);     \
        else KERNEL<1><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);                     \
    } while (0)

// Queues PCG iterations in batches until the device reports done; returns the
// final (iter, converged) through the pinned poll buffer.
static int pcg_iterate(const Tiles& T, int vec, const PcgBuffers& B, int max_iter, hipStream_t st,
                       int* iters_out, int* converged_out) {
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    hipEvent_t ev[2];
    BS_RETURN_IF_ERR(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    BS_RETURN_IF_ERR(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const size_t done_off = offsetof(PcgCtrl, done);
    int launched = 0, batch = 8, slot = 0;
    bool pending[2] = {false, false};
    int rc = hipSuccess;
    g_poll.slots[0] = g_poll.slots[2] = 0;
    while (launched < max_iter) {
        const int nb = (max_iter - launched) < batch ? (max_iter - launched) : batch;
        for (int it = 0; it < nb; ++it) {
            BS_PCG_LAUNCH(pcg_spmv_kernel, vec, B.nwg, st, T, B.code, B.p, B.ap, B.partials, B.ctrl);
            BS_PCG_LAUNCH(pcg_update_kernel, vec, B.nwg, st, T, B.code, B.p, B.ap, B.x, B.r, B.partials, B.ctrl);
            BS_PCG_LAUNCH(pcg_direction_kernel, vec, B.nwg, st, T, B.code, B.r, B.p, B.ctrl);
        }
        launched += nb;
        rc = hipGetLastError();
        if (rc != hipSuccess) break;
        rc = hipMemcpyAsync(g_poll.slots + 2 * slot, (const char*)B.ctrl + done_off, 2 * sizeof(int),
                            hipMemcpyDeviceToHost, st);
        if (rc != hipSuccess) break;
        rc = hipEventRecord(ev[slot], st);
        if (rc != hipSuccess) break;
        pending[slot] = true;
        // Wait for the previous batch (one batch stays queued ahead of the host).
        const int prev = slot ^ 1;
        if (pending[prev]) {
            rc = hipEventSynchronize(ev[prev]);
            if (rc != hipSuccess) break;
            pending[prev] = false;
            if (g_poll.slots[2 * prev] != 0) break;
        }
        slot ^= 1;
        if (batch < 64) batch *= 2;
    }
    if (rc == hipSuccess) {
        rc = hipMemcpyAsync(g_poll.final, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st);
        if (rc == hipSuccess) rc = hipStreamSynchronize(st);
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc != hipSuccess) return rc;
    if (iters_out) *iters_out = g_poll.final->iter;
    if (converged_out) *converged_out = g_poll.final->converged;
    return hipSuccess;
}

static unsigned pick_nwg(const Tiles& T, unsigned cap) {
    unsigned nwg = T.ntiles < kMaxWg ? T.ntiles : kMaxWg;
    if (nwg > cap) nwg = cap;
    return nwg < 1 ? 1 : nwg;
}

}  // namespace bs

using namespace bs;

// Drop-in for pressure_solve_pcg_launch (crates/cuda-kernels/src/
// pressure_solve_pcg.cu:89-118, crates/cuda-kernels/src/lib.rs:131-153).
// Workspace use: ws_b = b, ws_r = r, ws_p_search = p, ws_Ap = Ap; ws_z holds
// the one-byte class codes followed by the FP64 reduction partials (the z
// vector itself is never stored); dot_workspace (>= 1024 floats) holds the
// FP64 control block and, for very small grids, the partials.  The seven
// 1-float scalar slots receive float copies of the final alpha, beta, r.z
// (old and new), b.b, r.r and p.Ap.  Cold start, as the reference: p_out's
// initial contents are ignored.  Allocates no device memory.
extern "C" int bs_pressure_solve_pcg_launch(
    float* p_out, const float* div_u, const int* mat, int nx, int ny, int nz, float dx, float dt,
    float rho, int mat_fluid, int mat_air, float pressure_tol, int max_iter, float* ws_b, float* ws_r,
    float* ws_z, float* ws_p_search, float* ws_Ap, float* dot_workspace, float* dev_alpha,
    float* dev_beta, float* dev_rzold, float* dev_rznew, float* dev_bsq, float* dev_rsq,
    float* dev_pAp, int* iter_count_host) {
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        if (iter_count_host) *iter_count_host = 0;
        return hipSuccess;
    }
    hipStream_t st = nullptr;
    const int64_t n = (int64_t)nx * ny * nz;
    const int vec = pcg_vec(nz, {p_out, div_u, mat, ws_b, ws_r, ws_z, ws_p_search, ws_Ap});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    PcgBuffers B;
    B.code = (uint8_t*)ws_z;
    B.b = ws_b; B.x = p_out; B.r = ws_r; B.p = ws_p_search; B.ap = ws_Ap;
    B.ctrl = (PcgCtrl*)dot_workspace;
    // Partials after the codes in ws_z when there is room, else in dot_workspace.
    const uint64_t code_bytes = ((uint64_t)n + 255) & ~(uint64_t)255;
    const uint64_t zbytes = (uint64_t)n * 4;
    unsigned cap = zbytes > code_bytes ? (unsigned)((zbytes - code_bytes) / (2 * sizeof(double))) : 0;
    if (cap >= 64) {
        B.partials = (double*)((char*)ws_z + code_bytes);
    } else {
        B.partials = (double*)((char*)dot_workspace + 256);
        cap = (1024 * 4 - 256) / (2 * sizeof(double));
    }
    if (cap > kMaxWg) cap = kMaxWg;
    B.nwg = pick_nwg(T, cap);
    const double tol2 = (double)pressure_tol * (double)pressure_tol;
    pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, tol2, max_iter, 0);
    const float bscale = -rho * dx * dx / dt;  // pressure_solve_pcg.cu:137
    BS_PCG_LAUNCH(pcg_setup_kernel, vec, B.nwg, st, T, mat, div_u, bscale, mat_fluid, mat_air, 1, B.code,
                  B.b, B.x, B.r, B.p, B.partials, B.ctrl);
    BS_RETURN_IF_ERR(hipGetLastError());
    int iters = 0, conv = 0;
    if (max_iter > 0) {
        int rc = pcg_iterate(T, vec, B, max_iter, st, &iters, &conv);
        if (rc != hipSuccess) return rc;
    } else {
        BS_RETURN_IF_ERR(hipStreamSynchronize(st));
    }
    pcg_export_kernel<<<1, 64, 0, st>>>(B.ctrl, dev_alpha, dev_beta, dev_rzold, dev_rznew, dev_bsq,
                                        dev_rsq, dev_pAp);
    BS_RETURN_IF_ERR(hipGetLastError());
    BS_RETURN_IF_ERR(hipStreamSynchronize(st));
    if (iter_count_host) *iter_count_host = iters;
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Solver with its own workspace: warm start, true-residual target, every-
// iteration convergence test, asynchronous stream.
// ---------------------------------------------------------------------------

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

extern "C" size_t bs_pcg_workspace_bytes(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    return align256(n) + 4 * align256(n * 4) + align256(2 * kMaxWg * sizeof(double)) + 256;
}

static void carve(void* ws, size_t n, PcgBuffers& B) {
    char* c = (char*)ws;
    B.code = (uint8_t*)c; c += align256(n);
    B.b = (float*)c; c += align256(n * 4);
    B.r = (float*)c; c += align256(n * 4);
    B.p = (float*)c; c += align256(n * 4);
    B.ap = (float*)c; c += align256(n * 4);
    B.partials = (double*)c; c += align256(2 * kMaxWg * sizeof(double));
    B.ctrl = (PcgCtrl*)c;
}

// Solves A x = b (b from div_u as above) to ||b - A x||_2 <= tol ||b||_2 on
// the true residual: CG runs until the recursive residual meets tol, then the
// true residual is formed in one stencil pass (FP64 accumulation); if it is
// still above tol the solve restarts from x with r = b - A x (residual
// replacement), at most `max_restarts` times.  warm_start != 0 uses x's
// contents as the initial guess, otherwise x starts at 0.
extern "C" int bs_pcg_solve(float* x, const float* div_u, const int* mat, int nx, int ny, int nz,
                            float dx, float dt, float rho, int mat_fluid, int mat_air, double tol,
                            int max_iter, int warm_start, int max_restarts, void* workspace,
                            size_t workspace_bytes, hipStream_t st, bs_solve_stats* stats) {
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (workspace_bytes < bs_pcg_workspace_bytes(nx, ny, nz)) return (int)hipErrorInvalidValue;
    const size_t n = (size_t)nx * ny * nz;
    PcgBuffers B;
    carve(workspace, n, B);
    B.x = x;
    const int vec = pcg_vec(nz, {x, div_u, mat, B.b, B.r, B.p, B.ap, B.code});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    B.nwg = pick_nwg(T, kMaxWg);
    const double tol2 = tol * tol;
    const float bscale = -rho * dx * dx / dt;
    pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, tol2, max_iter, 1);
    BS_PCG_LAUNCH(pcg_setup_kernel, vec, B.nwg, st, T, mat, div_u, bscale, mat_fluid, mat_air,
                  warm_start ? 0 : 1, B.code, B.b, B.x, B.r, B.p, B.partials, B.ctrl);
    if (warm_start)
        BS_PCG_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partials, B.ctrl, 1);
    BS_RETURN_IF_ERR(hipGetLastError());
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    int total = 0, restarts = 0, conv = 0;
    PcgCtrl h;
    for (;;) {
        int iters = 0;
        int remaining = max_iter - total;
        if (remaining > 0) {
            int rc = pcg_iterate(T, vec, B, remaining, st, &iters, &conv);
            if (rc != hipSuccess) return rc;
        }
        // pcg_iterate's iteration counter continues across restarts.
        total = iters;
        // True residual of the current x (does not touch p).
        BS_PCG_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partials, B.ctrl, 0);
        BS_RETURN_IF_ERR(hipGetLastError());
        BS_RETURN_IF_ERR(hipMemcpyAsync(g_poll.final, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st));
        BS_RETURN_IF_ERR(hipStreamSynchronize(st));
        h = *g_poll.final;
        const double rel_true = h.bsq > 0.0 ? sqrt(h.rr_true / h.bsq) : 0.0;
        if (stats) {
            stats->iterations = h.iter;
            stats->restarts = restarts;
            stats->relres_true = rel_true;
            stats->bnorm = sqrt(h.bsq);
        }
        if (h.bsq == 0.0 || rel_true <= tol) { conv = 1; break; }
        if (restarts >= max_restarts || h.iter >= max_iter) { conv = 0; break; }
        // Residual replacement: r = b - A x, p = z, continue.
        ++restarts;
        pcg_clear_done_kernel<<<1, 64, 0, st>>>(B.ctrl);
        BS_PCG_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partials, B.ctrl, 1);
        BS_RETURN_IF_ERR(hipGetLastError());
        (void)total;
    }
    if (stats) {
        stats->converged = conv;
        stats->relres_recursive = h.bsq > 0.0 ? sqrt(h.rr / h.bsq) : 0.0;
    }
    return hipSuccess;
}
