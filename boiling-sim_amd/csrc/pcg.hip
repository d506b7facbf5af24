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
//   spmv     : Ap = A p, partials of p.Ap                          (8+1 B/cell)
//   update   : alpha = r.z / p.Ap; x += alpha p, r -= alpha Ap;
//              partials of r.r and r.z with z = r/diag             (24+1 B/cell)
//   direction: convergence test on r.r, beta = r.z_new / r.z;
//              p = r/diag + beta p                                 (12+1 B/cell)
// The cell class and the Jacobi diagonal live in one byte per cell (the
// `code` array, built once per solve from the int32 mask), so no kernel reads
// the 4-byte mask inside the loop.  Dots are FP64: each kernel writes one
// partial per workgroup and the next kernel sums them in a fixed order in
// every workgroup (grid.hpp), so scalars never leave the device, results are
// deterministic and no inter-workgroup fence is needed.  Every kernel first
// reads the device `done` flag and returns
// at once when the solve has finished, so the host can queue iterations in
// batches and poll the flag asynchronously instead of syncing every few
// iterations (the reference does a blocking copy every 5 iterations,
// pressure_solve_pcg.cu:205-216).
#include "boilsim_hip.h"
#include "grid.hpp"
#include "precond_api.hpp"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>

namespace bs {

// Device control block of one solve (lives in caller-provided memory).
struct PcgCtrl {
    double bsq;      // b.b
    double stop;     // tol^2 * bsq
    double rz[2];    // r.z of the current iterate, slot = iteration parity
    double alpha;
    double beta;
    double pAp;
    double rr;       // recursive r.r after the last update
    double rz_new;
    double rr_true;  // ||b - A x||^2 from the last residual pass
    int done;
    int iter;
    int converged;
    int max_iter;
    int mode;        // 0: reference cadence (iter >= 8 && iter % 5 == 0), 1: every iteration
    int best_it;     // flexible CG: iteration of the best residual so far
    int fallback;    // flexible CG: stalled or lost positivity
    int n_fluid;     // flexible CG with the fine-grid network: fluid cells (it normalises r)
    double best_rr;
    double dots[2];  // flexible CG: r.z_new, r.z_old (reduced)
    double apz;      // flexible CG, fused first pass: Ap.z of the iteration (z its preconditioned residual)
    int fresh;       // flexible CG: the next direction is p = z (first iteration of a solve)
    int polls;       // flexible CG: iterations posted to the host ring since the first direction
};
static_assert(sizeof(PcgCtrl) <= 256, "control block must fit 256 bytes");

// What the host polls after each flexible-CG iteration: the last kernel of
// the iteration stores it into a ring in mapped pinned host memory (slot
// (seq - 1) mod kPollRing), in place of a copy of the control block per
// launch unit.
struct PollRec {
    double rr;
    double stop;
    int iter;
    int done;
    int seq;  // 1 + the iteration's index since the first direction
    int pad;
};
constexpr int kPollRing = 8;  // > the iterations in flight (two units of at most two)

__device__ __forceinline__ void post_poll(PcgCtrl* ctrl, PollRec* ring) {
    const int q = ctrl->polls;
    ctrl->polls = q + 1;
    PollRec rec;
    rec.rr = ctrl->rr;
    rec.stop = ctrl->stop;
    rec.iter = ctrl->iter;
    rec.done = ctrl->done;
    rec.seq = q + 1;
    rec.pad = 0;
    ring[q & (kPollRing - 1)] = rec;
}

struct PcgBuffers {
    uint8_t* code;
    float *b, *x, *r, *p, *ap;
    double* partA;  // [G]     p.Ap
    double* partB;  // [2G]    r.r, r.z after the update
    double* partS;  // [2G]    set-up / residual pass
    double* partT;  // [ntiles] p.Ap per tile (one-tile SpMV), or null
    PcgCtrl* ctrl;
    unsigned nwg;
};

__device__ __forceinline__ int is_solid_m(int m, int mf, int ma) { return (m != mf && m != ma) ? 1 : 0; }

// ---------------------------------------------------------------------------
// Set-up: class codes, b, and (cold start) x = 0, r = b, p = z = M^-1 r (p
// may be null); partials of b.b and r.z.  Reads the int32 mask once
// (neighbours via cache).
// ---------------------------------------------------------------------------
template <int VEC>
__global__ __launch_bounds__(256) void pcg_setup_kernel(
    Tiles T, const int* __restrict__ mat, const float* __restrict__ div, float bscale, int mf, int ma,
    int cold, uint8_t* __restrict__ code, float* __restrict__ b, float* __restrict__ x,
    float* __restrict__ r, float* __restrict__ p, double* __restrict__ partS) {
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double bb = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        float dv[VEC], bv[VEC], zv[VEC], zero[VEC];
        uint8_t cv[VEC];
        ldf<VEC>(div, c0, dv);
        // the cell's row of VEC classes and its four neighbour rows as vector
        // loads, the k-neighbours across the row's ends as scalars; out of
        // bounds counts as air (not solid)
        int mv[VEC], ml[VEC], mr[VEC], md[VEC], mu[VEC];
        ldi<VEC>(mat, c0, mv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) ml[e] = mr[e] = md[e] = mu[e] = ma;
        if (i > 0) ldi<VEC>(mat, c0 - plane, ml);
        if (i < T.nx - 1) ldi<VEC>(mat, c0 + plane, mr);
        if (j > 0) ldi<VEC>(mat, c0 - T.nz, md);
        if (j < T.ny - 1) ldi<VEC>(mat, c0 + T.nz, mu);
        const int mkb = k0 > 0 ? mat[c0 - 1] : ma;
        const int mkf = k0 + VEC < T.nz ? mat[c0 + VEC] : ma;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            zero[e] = 0.0f;
            const int m = mv[e];
            if (m != mf) {
                cv[e] = (m == ma) ? kCodeAir : kCodeSolid;
                bv[e] = 0.0f;
                zv[e] = 0.0f;
                continue;
            }
            const int ns = is_solid_m(ml[e], mf, ma) + is_solid_m(mr[e], mf, ma) + is_solid_m(md[e], mf, ma) +
                           is_solid_m(mu[e], mf, ma) + is_solid_m(e > 0 ? mv[e - 1] : mkb, mf, ma) +
                           is_solid_m(e < VEC - 1 ? mv[e + 1] : mkf, mf, ma);
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
            if (p) stf<VEC>(p, c0, zv);  // null: the caller forms its own first direction
        }
    }
    double v[2] = {bb, rz};
    write_partials<2>(v, partS);
}

// r = b - A x formed in FP64 (stored rounded to FP32 unless r is null, for
// the test of the true residual alone); partials of r.r and r.z; with
// set_direction also p = z = M^-1 r (warm start / residual replacement).
template <int VEC>
__global__ __launch_bounds__(256) void pcg_residual_kernel(
    Tiles T, const uint8_t* __restrict__ code, const float* __restrict__ b, const float* __restrict__ x,
    float* __restrict__ r, float* __restrict__ p, double* __restrict__ partS, int set_direction) {
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        Nbhd<VEC> nb;
        nb.load(T, code, x, i, j, k0, c0);
        float bv[VEC], rv[VEC], zv[VEC];
        ldf<VEC>(b, c0, bv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (nb.cc[e] < kCodeFluid0) { rv[e] = 0.0f; zv[e] = 0.0f; continue; }
            // FP64: the true residual must not inherit FP32 cancellation error.
            const double r64 = (double)bv[e] - nb.apply64(e);
            rv[e] = (float)r64;
            zv[e] = precond(rv[e], nb.cc[e]);
            rr += r64 * r64;
            rz += (double)rv[e] * (double)zv[e];
        }
        if (r) stf<VEC>(r, c0, rv);
        if (set_direction) stf<VEC>(p, c0, zv);
    }
    double v[2] = {rr, rz};
    write_partials<2>(v, partS);
}

// One workgroup: totals of the set-up pass.  `cold`: x = 0, r = b.
__global__ __launch_bounds__(256) void pcg_setup_finalize(const double* __restrict__ partS, unsigned G,
                                                          PcgCtrl* ctrl, int cold) {
    double v[2];
    sum_partials<2>(partS, G, v);
    if (threadIdx.x != 0) return;
    ctrl->bsq = v[0];
    ctrl->stop = ctrl->stop * v[0];  // tol^2 was stored in stop
    if (cold) {
        ctrl->rz[0] = v[1];
        ctrl->rr = v[0];
        if (v[0] == 0.0 || ctrl->max_iter <= 0) ctrl->done = 1;
    }
}

// One workgroup: totals of a residual pass.  With `restart`, the residual
// becomes the recursive one, r.z goes to the slot of iteration `it_next`, and
// the solve is done if it already meets the tolerance.
__global__ __launch_bounds__(256) void pcg_residual_finalize(const double* __restrict__ partS, unsigned G,
                                                             PcgCtrl* ctrl, int restart, int it_next) {
    double v[2];
    sum_partials<2>(partS, G, v);
    if (threadIdx.x != 0) return;
    ctrl->rr_true = v[0];
    if (restart) {
        ctrl->rr = v[0];
        ctrl->rz[it_next & 1] = v[1];
        ctrl->iter = it_next;
        ctrl->done = (v[0] <= ctrl->stop || it_next >= ctrl->max_iter) ? 1 : 0;
        ctrl->converged = v[0] <= ctrl->stop ? 1 : 0;
    }
}

// Ap = A p; partials of p.Ap.
template <int VEC>
__global__ __launch_bounds__(256) void pcg_spmv_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                       const float* __restrict__ p, float* __restrict__ ap,
                                                       double* __restrict__ partA, const PcgCtrl* ctrl) {
    if (ctrl->done) return;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double pap = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        Nbhd<VEC> nb;
        nb.load(T, code, p, i, j, k0, c0);
        float out[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            out[e] = nb.apply(e);
            pap += (double)nb.vc[e] * (double)out[e];
        }
        stf<VEC>(ap, c0, out);
    }
    double v[1] = {pap};
    write_partials<1>(v, partA);
}

// Ap = A p with one 256-thread tile per workgroup (better latency hiding for
// a stencil than a grid-stride loop); one p.Ap partial per tile.
template <int VEC, typename S = float>
__global__ __launch_bounds__(256) void pcg_spmv1_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                        const S* __restrict__ p, S* __restrict__ ap,
                                                        double* __restrict__ partA, const PcgCtrl* ctrl) {
    if (ctrl->done) return;
    double pap = 0.0;
    int i, j, k0;
    int64_t c0;
    if (one_tile(T, VEC, i, j, k0, c0)) {
        Nbhd<VEC> nb;
        nb.load(T, code, p, i, j, k0, c0);
        float out[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            out[e] = through<S>(nb.apply(e));
            pap += (double)nb.vc[e] * (double)out[e];
        }
        stv<VEC>(ap, c0, out);
    }
    double v[1] = {pap};
    write_partials<1>(v, partA);
}

// Ap = A p and p.Ap for FP32 vectors on 4-cell chunks, marching down i: a
// workgroup owns a 14 x 56 (j,k) tile and an i segment; each of its 256
// threads owns one 4-cell chunk of the tile or of its one-chunk halo ring.
// i-neighbours are the thread's own registers (planes rolled down i), j and
// k neighbours come from two double-buffered LDS planes, so a cell costs
// one global load of p and one of its codes instead of seven of each (the
// one-tile kernel above is bound by those load requests).  Same float
// operations per cell as Nbhd::apply, so Ap is bit-identical; the FP64 p.Ap
// partials are grouped per workgroup.
constexpr int kRollKT = 14, kRollTJ = 14;

struct RollShape {
    int jtiles, ktiles, ni;
    unsigned grid;
    int tj = 0;  // j rows per tile (fcg_spmv_dir_kernel)
};

// At most 16 workgroups per CU (two rounds of the 8 a CU holds): tiles over
// (j,k), i split into segments.  440x440x280 in one run (workgroups, i planes
// each: us per launch): 4000, 18: 120.8; 5920, 12: 123.0; 1920, 37: 124.2;
// 4160, 17 (about 4096 rounded up, just past two rounds): 128.2; 2080, 34
// (just past one round): 145.3.
static RollShape roll_shape(int nx, int ny, int nz) {
    RollShape s;
    s.jtiles = (ny + kRollTJ - 1) / kRollTJ;
    s.ktiles = (nz + 4 * kRollKT - 1) / (4 * kRollKT);
    const int cols = s.jtiles * s.ktiles;
    int nseg = 16 * device_cus() / cols;
    nseg = nseg < 1 ? 1 : (nseg > nx ? nx : nseg);
    s.ni = (nx + nseg - 1) / nseg;
    nseg = (nx + s.ni - 1) / s.ni;
    s.grid = (unsigned)(nseg * cols);
    return s;
}

// The flexible CG's first pass (fcg_spmv_dir_kernel) takes the same stencil
// on a taller tile: up to 62 x 56 (j,k) cells and their one-chunk ring on 1024
// threads.  The pass is memory-bound and its ring is re-read from L2 or HBM
// by the neighbouring tiles; the taller tile cuts that overhead from 1.18x
// to 1.13x of the tile's cells (440x440x280: 339 -> 319 us per pass).  The
// rows are spread evenly over the j tiles (440: 8 x 55 rather than 7 x 62 +
// 6), and the grid is at most one workgroup per CU, each walking a long i
// segment.  Measured in one run at 440x440x280 (tile rows, workgroups, i
// planes each: us): 55, 240, 74: 289.5; 55, 480, 37: 297; 55, 1000, 18: 301.5;
// 62, 240, 74: 303; 62, 1040, 17 (the earlier rule, just past two rounds of
// the two workgroups a CU holds): 320; 55, 280, 63 (just past one round): 424;
// 256 equal shares of the (column, plane) pairs in column order, which
// leaves no CU idle but has the workgroups on 256 different planes at once:
// 524.  Plain PCG keeps the 14-row tile of pcg_spmv_roll_kernel.
constexpr int kDirTJ = 62, kDirThreads = (kDirTJ + 2) * (kRollKT + 2);
static RollShape dir_shape(int nx, int ny, int nz) {
    RollShape s;
    s.jtiles = (ny + kDirTJ - 1) / kDirTJ;
    s.tj = (ny + s.jtiles - 1) / s.jtiles;
    s.ktiles = (nz + 4 * kRollKT - 1) / (4 * kRollKT);
    const int cols = s.jtiles * s.ktiles;
    int nseg = device_cus() / cols;
    nseg = nseg < 1 ? 1 : (nseg > nx ? nx : nseg);
    s.ni = (nx + nseg - 1) / nseg;
    nseg = (nx + s.ni - 1) / s.ni;
    s.grid = (unsigned)(nseg * cols);
    return s;
}

__global__ __launch_bounds__(256) void pcg_spmv_roll_kernel(int nx, int ny, int nz, int ni_seg, int jtiles,
                                                            int ktiles, const uint8_t* __restrict__ code,
                                                            const float* __restrict__ p, float* __restrict__ ap,
                                                            double* __restrict__ part, const PcgCtrl* ctrl) {
    if (ctrl->done) return;
    constexpr int CH = kRollKT + 2, SR = kRollTJ + 2, W = 4 * CH;
    __shared__ __align__(16) float sV[2][SR * W + 8];
    __shared__ __align__(16) uint8_t sC[2][SR * W + 16];
    const int tid = threadIdx.x;
    const bool act = tid < CH * SR;
    const int r = tid / CH, c = tid - (tid / CH) * CH;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int cols = jtiles * ktiles;
    const int seg = (int)(b / cols), col = (int)(b % cols);
    const int jt = col / ktiles, kt = col - (col / ktiles) * ktiles;
    const int j0 = jt * kRollTJ, k0 = kt * 4 * kRollKT;
    const int i0 = seg * ni_seg, i1 = min(nx, i0 + ni_seg);
    const int j = j0 - 1 + r, k = k0 - 4 + 4 * c;
    const bool jk_in = act && j >= 0 && j < ny && k >= 0 && k < nz;
    const bool interior = act && r >= 1 && r <= kRollTJ && c >= 1 && c <= kRollKT && j < ny && k < nz;
    const int64_t pl = (int64_t)ny * nz;
    const int64_t off = (int64_t)j * nz + k;
    // planes outside the grid and chunks outside (j,k) read as air (value 0)
    auto ldv = [&](int q) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(p + (int64_t)q * pl + off);
        return v;
    };
    auto ldk = [&](int q) {
        uint32_t v = 0;
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const uint32_t*>(code + (int64_t)q * pl + off);
        return v;
    };
    float4 Vm = ldv(i0 - 1), Vc = ldv(i0), Vn = ldv(i0 + 1);
    uint32_t Km = ldk(i0 - 1), Kc = ldk(i0), Kn = ldk(i0 + 1);
    double pap = 0.0;
    for (int q = i0; q < i1; ++q) {
        float* V = sV[q & 1] + 4;
        uint8_t* C = sC[q & 1] + 8;
        if (act) {
            *reinterpret_cast<float4*>(V + r * W + 4 * c) = Vc;
            *reinterpret_cast<uint32_t*>(C + r * W + 4 * c) = Kc;
        }
        const float4 Vf = ldv(q + 2);
        const uint32_t Kf = ldk(q + 2);
        __syncthreads();
        if (interior) {
            const float4 d4 = *reinterpret_cast<const float4*>(V + (r - 1) * W + 4 * c);
            const float4 u4 = *reinterpret_cast<const float4*>(V + (r + 1) * W + 4 * c);
            const uint32_t kd = *reinterpret_cast<const uint32_t*>(C + (r - 1) * W + 4 * c);
            const uint32_t ku = *reinterpret_cast<const uint32_t*>(C + (r + 1) * W + 4 * c);
            const float vb = V[r * W + 4 * c - 1], vf = V[r * W + 4 * c + 4];
            const uint8_t cb = C[r * W + 4 * c - 1], cf = C[r * W + 4 * c + 4];
            const float vc[4] = {Vc.x, Vc.y, Vc.z, Vc.w}, vl[4] = {Vm.x, Vm.y, Vm.z, Vm.w};
            const float vr[4] = {Vn.x, Vn.y, Vn.z, Vn.w}, vd[4] = {d4.x, d4.y, d4.z, d4.w};
            const float vu[4] = {u4.x, u4.y, u4.z, u4.w};
            float out[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint8_t cc = (Kc >> (8 * e)) & 0xFF;
                if (cc < kCodeFluid0) { out[e] = 0.0f; continue; }
                const float ps = vc[e];
                const uint8_t c_b = e > 0 ? (uint8_t)((Kc >> (8 * (e - 1))) & 0xFF) : cb;
                const float v_b = e > 0 ? vc[e - 1] : vb;
                const uint8_t c_f = e < 3 ? (uint8_t)((Kc >> (8 * (e + 1))) & 0xFF) : cf;
                const float v_f = e < 3 ? vc[e + 1] : vf;
                float s = nbr_term((uint8_t)((Km >> (8 * e)) & 0xFF), vl[e], ps) +
                          nbr_term((uint8_t)((Kn >> (8 * e)) & 0xFF), vr[e], ps);
                s = s + nbr_term((uint8_t)((kd >> (8 * e)) & 0xFF), vd[e], ps);
                s = s + nbr_term((uint8_t)((ku >> (8 * e)) & 0xFF), vu[e], ps);
                s = s + nbr_term(c_b, v_b, ps);
                s = s + nbr_term(c_f, v_f, ps);
                out[e] = 6.0f * ps - s;
                pap += (double)ps * (double)out[e];
            }
            *reinterpret_cast<float4*>(ap + (int64_t)q * pl + off) = make_float4(out[0], out[1], out[2], out[3]);
        }
        Vm = Vc; Vc = Vn; Vn = Vf;
        Km = Kc; Kc = Kn; Kn = Kf;
    }
    double v[1] = {pap};
    write_partials<1>(v, part);
}

// One 1024-thread workgroup: deterministic FP64 total of G partials per value.
template <int NV>
__device__ __forceinline__ void reduce1024(const double* __restrict__ part, unsigned G, double (&out)[NV]) {
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        // Eight loads in flight per thread (a single running sum would wait
        // for each load in turn); fixed order, so the total is deterministic.
        const double* pk = part + (size_t)k * G;
        double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        unsigned w = threadIdx.x;
        for (; w + 7 * 1024 < G; w += 8 * 1024) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = pk[w + u * 1024];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] += v[u];
        }
        for (; w < G; w += 1024) a[0] += pk[w];
        double acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        acc = wave_sum64(acc);
        if (lane == 0) red[k][wave] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = red[k][0];
        for (int w = 1; w < 16; ++w) s += red[k][w];
        out[k] = s;
    }
}

// p.Ap total and alpha = r.z / p.Ap for iteration `it` (it < 0: read the
// device counter).
__global__ __launch_bounds__(1024) void pap_reduce_kernel(const double* __restrict__ partA, unsigned G,
                                                          PcgCtrl* ctrl, int it) {
    if (ctrl->done) return;
    double v[1];
    reduce1024<1>(partA, G, v);
    if (threadIdx.x != 0) return;
    const int itv = it >= 0 ? it : ctrl->iter;
    const double rz = ctrl->rz[itv & 1];
    ctrl->pAp = v[0];
    ctrl->alpha = v[0] != 0.0 ? rz / v[0] : 0.0;
}

// r.z_new, r.z_old totals and the flexible-CG direction's scalars, after the
// check of iteration it advanced ctrl->iter to it + 1: beta =
// (r.z_new - r.z_old) / r.z (Polak-Ribiere), or, with init, no beta and the
// next direction p = z.  A preconditioner that is not positive on this
// residual (r.z_new <= 0) ends the flexible CG for the plain-PCG fallback;
// so does reaching max_iter.
// apz_mode: r.z_old is not in the partials but follows from the iteration's
// quantities, r_{k+1}.z_k = r_k.z_k - alpha_k (A p_k).z_k (the fused first
// pass accumulates (A p_k).z_k while it holds both), which spares the last
// preconditioner pass the z_old stream.
__device__ void dots_step(PcgCtrl* ctrl, const double (&v)[2], int init, int apz_mode) {
    ctrl->dots[0] = v[0];
    ctrl->dots[1] = v[1];
    if (!(v[0] > 0.0)) {
        ctrl->fallback = 1;
        ctrl->done = 1;
        return;
    }
    const int it1 = ctrl->iter;  // iterations completed
    double beta = 0.0;
    if (!init) {
        const double rz_old = ctrl->rz[(it1 - 1) & 1];
        const double rzo = apz_mode ? rz_old - ctrl->alpha * ctrl->apz : v[1];
        if (apz_mode) ctrl->dots[1] = rzo;
        beta = rz_old != 0.0 ? (v[0] - rzo) / rz_old : 0.0;
    }
    ctrl->rz[it1 & 1] = v[0];
    ctrl->rz_new = v[0];
    ctrl->beta = beta;
    ctrl->fresh = init;
    if (!init && it1 >= ctrl->max_iter) ctrl->done = 1;
}

// The iteration's last kernel.  init (the first direction) restarts the
// count of posted iterations; with a ring, every iteration posts its state
// to it, a finished solve's no-op iterations included.
__global__ __launch_bounds__(1024) void dots_reduce_kernel(const double* __restrict__ partZ, unsigned G,
                                                           PcgCtrl* ctrl, int init, int apz_mode = 0,
                                                           PollRec* ring = nullptr) {
    if (!ctrl->done) {
        double v[2];
        reduce1024<2>(partZ, G, v);
        if (threadIdx.x == 0) dots_step(ctrl, v, init, apz_mode);
    }
    if (threadIdx.x != 0) return;
    if (init) ctrl->polls = 0;
    if (ring) post_poll(ctrl, ring);
}

// alpha = r.z / p.Ap (0 if p.Ap == 0, axpy_device.cu:82-91); x += alpha p,
// r -= alpha Ap (pressure_solve_pcg.cu:192-202); partials of r.r and r.z.
template <int VEC, typename S = float>
__global__ __launch_bounds__(256) void pcg_update_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                         const S* __restrict__ p,
                                                         const S* __restrict__ ap,
                                                         float* __restrict__ x, S* __restrict__ r,
                                                         const double* __restrict__ partA,
                                                         double* __restrict__ partB, PcgCtrl* ctrl, int it) {
    if (ctrl->done) return;
    (void)partA;
    (void)it;
    const double alpha = ctrl->alpha;  // formed by pap_reduce_kernel
    const float af = (float)alpha;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0, rzn = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC];
        float pv[VEC], av[VEC], xv[VEC], rv[VEC];
        ldc<VEC>(code, c0, cc);
        ldv<VEC>(p, c0, pv);
        ldv<VEC>(ap, c0, av);
        ldf<VEC>(x, c0, xv);
        ldv<VEC>(r, c0, rv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            xv[e] = xv[e] + af * pv[e];
            rv[e] = through<S>(rv[e] - af * av[e]);
            const float z = precond(rv[e], cc[e]);
            rr += (double)rv[e] * (double)rv[e];
            rzn += (double)rv[e] * (double)z;
        }
        stf<VEC>(x, c0, xv);
        stv<VEC>(r, c0, rv);
    }
    double v[2] = {rr, rzn};
    write_partials<2>(v, partB);
}

// Convergence test on the recursive r.r (mode 0: only when it >= 8 &&
// it % 5 == 0, reporting it + 1, pressure_solve_pcg.cu:205-216), then
// beta = r.z_new / r.z and p = M^-1 r + beta p.
template <int VEC, typename S = float>
__global__ __launch_bounds__(256) void pcg_direction_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                            const S* __restrict__ r,
                                                            S* __restrict__ p,
                                                            const double* __restrict__ partB,
                                                            PcgCtrl* ctrl, int it) {
    if (ctrl->done) return;
    double v[2];
    sum_partials<2>(partB, gridDim.x, v);
    const bool check = ctrl->mode != 0 || (it >= 8 && (it % 5) == 0);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    // mode 2 (inner solve of a refinement) also ends when the recursive
    // residual norm has grown 4x over its start: low-precision storage has
    // lost the process, and the caller replaces the residual.
    const bool diverged = ctrl->mode == 2 && v[0] > 16.0 * ctrl->bsq;
    if ((check && v[0] < ctrl->stop) || diverged) {
        if (lead) {
            ctrl->rr = v[0];
            ctrl->iter = it + 1;
            ctrl->converged = diverged ? 0 : 1;
            ctrl->done = 1;
        }
        return;
    }
    const double rz_old = ctrl->rz[it & 1];
    const double beta = rz_old != 0.0 ? v[1] / rz_old : 0.0;
    if (lead) {
        ctrl->rr = v[0];
        ctrl->rz_new = v[1];
        ctrl->rz[(it + 1) & 1] = v[1];
        ctrl->beta = beta;
        ctrl->iter = it + 1;
        if (it + 1 >= ctrl->max_iter) ctrl->done = 1;
    }
    const float bf = (float)beta;
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
        ldv<VEC>(r, c0, rv);
        ldv<VEC>(p, c0, pv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) pv[e] = precond(rv[e], cc[e]) + bf * pv[e];
        stv<VEC>(p, c0, pv);
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

// Float copies of the final scalars into the reference's 1-float slots.
__global__ void pcg_export_kernel(const PcgCtrl* ctrl, float* alpha, float* beta, float* rzold,
                                  float* rznew, float* bsq, float* rsq, float* pAp) {
    if (threadIdx.x != 0) return;
    if (alpha) *alpha = (float)ctrl->alpha;
    if (beta) *beta = (float)ctrl->beta;
    if (rzold) *rzold = (float)ctrl->rz[ctrl->iter & 1];
    if (rznew) *rznew = (float)ctrl->rz_new;
    if (bsq) *bsq = (float)ctrl->bsq;
    if (rsq) *rsq = (float)ctrl->rr;
    if (pAp) *pAp = (float)ctrl->pAp;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

constexpr unsigned kMaxWg = 1024;

// Pinned host mirror of the control block for asynchronous polling.
struct HostPoll {
    int* slots = nullptr;      // [2][2]: (done, iter) of the last two batches
    PcgCtrl* final = nullptr;  // full control block after a solve
    PollRec* ring = nullptr;   // [kPollRing]: the flexible CG's per-iteration posts (host view)
    PollRec* ring_dev = nullptr;  // the same memory as kernels address it
};
static HostPoll g_poll;

// The pinned polling block above, the flexible CG's graph cache and capture
// stream, and the pass timer are process-wide.  Every solver entry point
// holds this lock for the whole call, so solves issued from different host
// threads run one after another instead of racing on them (and a graph held
// by one solve cannot be evicted by another).
static std::recursive_mutex g_solve_mu;

static int ensure_poll() {
    if (g_poll.slots) return hipSuccess;
    void* h = nullptr;
    BS_RETURN_IF_ERR(hipHostMalloc(&h, 2048, hipHostMallocDefault));
    void* d = nullptr;
    if (const hipError_t e = hipHostGetDevicePointer(&d, (char*)h + 1024, 0); e != hipSuccess) {
        (void)hipHostFree(h);
        return (int)e;
    }
    static_assert(sizeof(PollRec) * kPollRing <= 1024, "poll ring must fit its block");
    g_poll.slots = (int*)h;
    g_poll.final = (PcgCtrl*)((char*)h + 256);
    g_poll.ring = (PollRec*)((char*)h + 1024);
    g_poll.ring_dev = (PollRec*)d;
    return hipSuccess;
}

#define BS_VEC_LAUNCH(KERNEL, VEC, GRID, ST, ...)                                   \
    do {                                                                           \
        if ((VEC) == 4) KERNEL<4><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);          \
        else if ((VEC) == 2) KERNEL<2><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);     \
        else KERNEL<1><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);                     \
    } while (0)

static uint32_t max_tiles(int nx, int ny, int nz);

// Whether the i-marching SpMV applies: nz % 4 == 0 with aligned vectors, and
// its partials fit the per-tile slots (any_size) or the grid has at least as
// many workgroups as vector tiles (the plain PCG's choice).
static bool roll_fits(const Tiles& T, int vec, const PcgBuffers& B, bool any_size) {
    const unsigned g = roll_shape(T.nx, T.ny, T.nz).grid;
    return B.partT && vec == 4 && g <= (any_size ? max_tiles(T.nx, T.ny, T.nz) : (unsigned)T.ntiles);
}

// Ap = A p and alpha = r.z / p.Ap into the control block (it < 0: the
// iteration number is the device counter).
static void enqueue_spmv(const Tiles& T, int vec, const PcgBuffers& B, int it, hipStream_t st,
                         bool any_size = false) {
    const RollShape rs = roll_shape(T.nx, T.ny, T.nz);
    if (roll_fits(T, vec, B, any_size)) {
        pcg_spmv_roll_kernel<<<rs.grid, 256, 0, st>>>(T.nx, T.ny, T.nz, rs.ni, rs.jtiles, rs.ktiles, B.code, B.p,
                                                      B.ap, B.partT, B.ctrl);
        pap_reduce_kernel<<<1, 1024, 0, st>>>(B.partT, rs.grid, B.ctrl, it);
    } else if (B.partT) {
        BS_VEC_LAUNCH(pcg_spmv1_kernel, vec, T.ntiles, st, T, B.code, B.p, B.ap, B.partT, B.ctrl);
        pap_reduce_kernel<<<1, 1024, 0, st>>>(B.partT, T.ntiles, B.ctrl, it);
    } else {
        BS_VEC_LAUNCH(pcg_spmv_kernel, vec, B.nwg, st, T, B.code, B.p, B.ap, B.partA, B.ctrl);
        pap_reduce_kernel<<<1, 1024, 0, st>>>(B.partA, B.nwg, B.ctrl, it);
    }
}

#define BS_VECS_LAUNCH(KERNEL, S, VEC, GRID, ST, ...)                                   \
    do {                                                                               \
        if ((VEC) == 8) KERNEL<8, S><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);           \
        else if ((VEC) == 4) KERNEL<4, S><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);      \
        else if ((VEC) == 2) KERNEL<2, S><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);      \
        else KERNEL<1, S><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);                      \
    } while (0)

// One PCG iteration whose vectors r, p, Ap are stored as S (x stays float).
template <typename S>
static void enqueue_pcg_iteration(const Tiles& T, int vec, const PcgBuffers& B, int it, hipStream_t st) {
    if constexpr (sizeof(S) == 4) {
        enqueue_spmv(T, vec, B, it, st);
        BS_VEC_LAUNCH(pcg_update_kernel, vec, B.nwg, st, T, B.code, B.p, B.ap, B.x, B.r, B.partA, B.partB,
                      B.ctrl, it);
        BS_VEC_LAUNCH(pcg_direction_kernel, vec, B.nwg, st, T, B.code, B.r, B.p, B.partB, B.ctrl, it);
    } else {
        S* p = reinterpret_cast<S*>(B.p);
        S* ap = reinterpret_cast<S*>(B.ap);
        S* r = reinterpret_cast<S*>(B.r);
        BS_VECS_LAUNCH(pcg_spmv1_kernel, S, vec, T.ntiles, st, T, B.code, p, ap, B.partT, B.ctrl);
        pap_reduce_kernel<<<1, 1024, 0, st>>>(B.partT, T.ntiles, B.ctrl, it);
        BS_VECS_LAUNCH(pcg_update_kernel, S, vec, B.nwg, st, T, B.code, p, ap, B.x, r, B.partA, B.partB, B.ctrl,
                       it);
        BS_VECS_LAUNCH(pcg_direction_kernel, S, vec, B.nwg, st, T, B.code, r, p, B.partB, B.ctrl, it);
    }
}

// Queues PCG iterations it0, it0+1, ... (< max_iter) in batches until the
// device reports done, polling the flag one batch behind; leaves the final
// control block in g_poll.final.
template <typename S>
static int pcg_iterate_t(const Tiles& T, int vec, const PcgBuffers& B, int it0, int max_iter, hipStream_t st) {
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    hipEvent_t ev[2];
    BS_RETURN_IF_ERR(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    if (const hipError_t e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming); e != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return (int)e;
    }
    const size_t done_off = offsetof(PcgCtrl, done);
    int it = it0, batch = 8, slot = 0;
    bool pending[2] = {false, false};
    hipError_t rc = hipSuccess;
    g_poll.slots[0] = g_poll.slots[2] = 0;
    while (it < max_iter) {
        const int nb = (max_iter - it) < batch ? (max_iter - it) : batch;
        for (int q = 0; q < nb; ++q, ++it) enqueue_pcg_iteration<S>(T, vec, B, it, st);
        rc = hipGetLastError();
        if (rc != hipSuccess) break;
        rc = hipMemcpyAsync(g_poll.slots + 2 * slot, (const char*)B.ctrl + done_off, 2 * sizeof(int),
                            hipMemcpyDeviceToHost, st);
        if (rc != hipSuccess) break;
        rc = hipEventRecord(ev[slot], st);
        if (rc != hipSuccess) break;
        pending[slot] = true;
        // Wait for the previous batch: one batch stays queued ahead of the host.
        const int prev = slot ^ 1;
        if (pending[prev]) {
            rc = hipEventSynchronize(ev[prev]);
            if (rc != hipSuccess) break;
            pending[prev] = false;
            if (g_poll.slots[2 * prev] != 0) break;
        }
        slot ^= 1;
        if (batch < 16) batch *= 2;
    }
    if (rc == hipSuccess) {
        rc = hipMemcpyAsync(g_poll.final, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st);
        if (rc == hipSuccess) rc = hipStreamSynchronize(st);
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return (int)rc;
}

static int pcg_iterate(const Tiles& T, int vec, const PcgBuffers& B, int it0, int max_iter, hipStream_t st) {
    return pcg_iterate_t<float>(T, vec, B, it0, max_iter, st);
}

static unsigned pick_nwg(const Tiles& T, unsigned cap) {
    unsigned nwg = T.ntiles < kMaxWg ? T.ntiles : kMaxWg;
    if (nwg > cap) nwg = cap;
    return nwg < 1 ? 1 : nwg;
}

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

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
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        if (iter_count_host) *iter_count_host = 0;
        return hipSuccess;
    }
    hipStream_t st = nullptr;
    const int64_t n = (int64_t)nx * ny * nz;
    const int vec = vec_for(nz, {p_out, div_u, mat, ws_b, ws_r, ws_z, ws_p_search, ws_Ap});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    PcgBuffers B;
    B.code = (uint8_t*)ws_z;
    B.b = ws_b; B.x = p_out; B.r = ws_r; B.p = ws_p_search; B.ap = ws_Ap;
    B.ctrl = (PcgCtrl*)dot_workspace;
    // 5 partials per workgroup: after the codes in ws_z when there is room,
    // else in dot_workspace after the control block.
    const uint64_t code_bytes = align256((size_t)n);
    const uint64_t zbytes = (uint64_t)n * 4;
    unsigned cap = zbytes > code_bytes ? (unsigned)((zbytes - code_bytes) / (5 * sizeof(double))) : 0;
    double* part;
    if (cap >= 64) {
        part = (double*)((char*)ws_z + code_bytes);
    } else {
        part = (double*)((char*)dot_workspace + 256);
        cap = (1024 * 4 - 256) / (5 * sizeof(double));
    }
    B.nwg = pick_nwg(T, cap);
    B.partA = part;
    B.partB = part + B.nwg;
    B.partS = part + 3 * B.nwg;
    // Per-tile p.Ap partials after those when ws_z has room for them too.
    B.partT = nullptr;
    if (cap >= 64 && zbytes - code_bytes >= (5 * (uint64_t)B.nwg + T.ntiles) * sizeof(double))
        B.partT = part + 5 * B.nwg;
    const double tol2 = (double)pressure_tol * (double)pressure_tol;
    pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, tol2, max_iter, 0);
    const float bscale = -rho * dx * dx / dt;  // pressure_solve_pcg.cu:137
    BS_VEC_LAUNCH(pcg_setup_kernel, vec, B.nwg, st, T, mat, div_u, bscale, mat_fluid, mat_air, 1, B.code,
                  B.b, B.x, B.r, B.p, B.partS);
    pcg_setup_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1);
    BS_RETURN_IF_ERR(hipGetLastError());
    int iters = 0;
    if (max_iter > 0) {
        const int rc = pcg_iterate(T, vec, B, 0, max_iter, st);
        if (rc != hipSuccess) return rc;
        iters = g_poll.final->iter;
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
// iteration convergence test, any stream.
// ---------------------------------------------------------------------------

namespace bs {
// Tile count of the widest vector width a grid can use (an upper bound for
// the per-tile partial arrays).
static uint32_t max_tiles(int nx, int ny, int nz) { return make_tiles(nx, ny, nz, 1).ntiles; }
}  // namespace bs

extern "C" size_t bs_pcg_workspace_bytes(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    return align256(n) + 4 * align256(n * 4) + align256(5 * kMaxWg * sizeof(double)) + 256 +
           align256((size_t)max_tiles(nx, ny, nz) * sizeof(double));
}

static void carve(void* ws, size_t n, PcgBuffers& B) {
    char* c = (char*)ws;
    B.code = (uint8_t*)c; c += align256(n);
    B.b = (float*)c; c += align256(n * 4);
    B.r = (float*)c; c += align256(n * 4);
    B.p = (float*)c; c += align256(n * 4);
    B.ap = (float*)c; c += align256(n * 4);
    double* part = (double*)c; c += align256(5 * kMaxWg * sizeof(double));
    B.partA = part;
    B.partB = part + kMaxWg;
    B.partS = part + 3 * kMaxWg;
    B.ctrl = (PcgCtrl*)c; c += 256;
    B.partT = (double*)c;
}

namespace bs {
// Queues PCG iterations like pcg_iterate, but waits for each batch with a
// plain copy of the done flag into host memory the caller owns (no pinned
// block, no events: nothing is allocated); the final control block lands in
// *h.
static int pcg_iterate_sync(const Tiles& T, int vec, const PcgBuffers& B, int it0, int max_iter,
                            hipStream_t st, PcgCtrl* h) {
    int it = it0, batch = 8;
    while (it < max_iter) {
        const int nb = (max_iter - it) < batch ? (max_iter - it) : batch;
        for (int q = 0; q < nb; ++q, ++it) enqueue_pcg_iteration<float>(T, vec, B, it, st);
        BS_RETURN_IF_ERR(hipGetLastError());
        int done = 0;
        BS_RETURN_IF_ERR(hipMemcpyAsync(&done, (const char*)B.ctrl + offsetof(PcgCtrl, done), sizeof(int),
                                        hipMemcpyDeviceToHost, st));
        BS_RETURN_IF_ERR(hipStreamSynchronize(st));
        if (done) break;
        if (batch < 32) batch *= 2;
    }
    BS_RETURN_IF_ERR(hipMemcpyAsync(h, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st));
    return (int)hipStreamSynchronize(st);
}

// The body of bs_pcg_solve over carved buffers (B.x set).  pinned: poll the
// done flag one batch behind through the pinned host block; otherwise through
// pcg_iterate_sync into host memory on the stack.
static int pcg_solve_body(PcgBuffers& B, const float* div_u, const int* mat, int nx, int ny, int nz,
                          float dx, float dt, float rho, int mat_fluid, int mat_air, double tol, int max_iter,
                          int warm_start, int max_restarts, hipStream_t st, bs_solve_stats* stats,
                          bool pinned) {
    float* x = B.x;
    const int vec = vec_for(nz, {x, div_u, mat, B.b, B.r, B.p, B.ap, B.code});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    B.nwg = pick_nwg(T, kMaxWg);
    // The partial slots are sized for kMaxWg workgroups.
    B.partB = B.partA + B.nwg;
    B.partS = B.partA + 3 * B.nwg;
    const float bscale = -rho * dx * dx / dt;
    pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, tol * tol, max_iter, 1);
    BS_VEC_LAUNCH(pcg_setup_kernel, vec, B.nwg, st, T, mat, div_u, bscale, mat_fluid, mat_air,
                  warm_start ? 0 : 1, B.code, B.b, B.x, B.r, B.p, B.partS);
    pcg_setup_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, warm_start ? 0 : 1);
    if (warm_start) {
        BS_VEC_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partS, 1);
        pcg_residual_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1, 0);
    }
    BS_RETURN_IF_ERR(hipGetLastError());
    int restarts = 0, conv = 0, it0 = 0;
    PcgCtrl h, hstage;
    PcgCtrl* hc = pinned ? g_poll.final : &hstage;
    for (;;) {
        int rc = pinned ? pcg_iterate(T, vec, B, it0, max_iter, st)
                        : pcg_iterate_sync(T, vec, B, it0, max_iter, st, &hstage);
        if (rc != hipSuccess) return rc;
        // True residual of the current x (r and p untouched).
        BS_VEC_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, nullptr, B.p, B.partS, 0);
        pcg_residual_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 0, 0);
        BS_RETURN_IF_ERR(hipGetLastError());
        BS_RETURN_IF_ERR(hipMemcpyAsync(hc, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st));
        BS_RETURN_IF_ERR(hipStreamSynchronize(st));
        h = *hc;
        const double rel_true = h.bsq > 0.0 ? sqrt(h.rr_true / h.bsq) : 0.0;
        if (stats) {
            stats->iterations = h.iter;
            stats->restarts = restarts;
            stats->relres_true = rel_true;
            stats->bnorm = sqrt(h.bsq);
        }
        if (h.bsq == 0.0 || rel_true <= tol) { conv = 1; break; }
        if (restarts >= max_restarts || h.iter >= max_iter) { conv = 0; break; }
        // Residual replacement: r = b - A x, p = z, continue the count.
        ++restarts;
        it0 = h.iter;
        BS_VEC_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partS, 1);
        pcg_residual_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1, it0);
        BS_RETURN_IF_ERR(hipGetLastError());
    }
    if (stats) {
        stats->converged = conv;
        stats->relres_recursive = h.bsq > 0.0 ? sqrt(h.rr / h.bsq) : 0.0;
    }
    return hipSuccess;
}
}  // namespace bs

// Solves A x = b (b from div_u as above) to ||b - A x||_2 <= tol ||b||_2 on
// the true residual: CG runs until the recursive residual meets tol, then the
// true residual is formed in one FP64 stencil pass; while it is above tol the
// solve restarts from x with r = b - A x (residual replacement), at most
// max_restarts times.  warm_start != 0 uses x's contents as the initial guess.
extern "C" int bs_pcg_solve(float* x, const float* div_u, const int* mat, int nx, int ny, int nz,
                            float dx, float dt, float rho, int mat_fluid, int mat_air, double tol,
                            int max_iter, int warm_start, int max_restarts, void* workspace,
                            size_t workspace_bytes, hipStream_t st, bs_solve_stats* stats) {
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (workspace_bytes < bs_pcg_workspace_bytes(nx, ny, nz)) return (int)hipErrorInvalidValue;
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    PcgBuffers B;
    carve(workspace, (size_t)nx * ny * nz, B);
    B.x = x;
    return pcg_solve_body(B, div_u, mat, nx, ny, nz, dx, dt, rho, mat_fluid, mat_air, tol, max_iter, warm_start,
                          max_restarts, st, stats, true);
}

// ---------------------------------------------------------------------------
// bs_pcg_solve for callers that cannot allocate: its workspace is `scratch`,
// a fixed device array of kScratchBytes that every device holds from the
// library's load until the process ends.  Host code never passes `scratch`
// itself (in host code the name is a host-side stand-in, not a device
// address): each device's address of it is resolved once with
// hipGetSymbolAddress, with that device current, and kept in g_scratch_ptr;
// the kernels get carved pointers into that address.  One array per device
// serves every call on it, so calls that overlap on one device race.  The
// kernels write every buffer they read before reading it (the control block
// by pcg_ctrl_init_kernel, codes/b/r/p by the set-up pass, partials by the
// kernel before their consumer), so the previous call's contents never leak.
// ---------------------------------------------------------------------------
namespace bs {
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
static void* g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

// The current device's address of `scratch`, resolved on first use.
int scratch_setup(void** out) {
    int dev = 0;
    BS_RETURN_IF_ERR(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        BS_RETURN_IF_ERR(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    if (out) *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace bs

extern "C" size_t bs_pcg_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" int bs_pcg_solve_noalloc(float* x, const float* div_u, const int* mat, int nx, int ny, int nz,
                                    float dx, float dt, float rho, int mat_fluid, int mat_air, double tol,
                                    int max_iter, int warm_start, int max_restarts, hipStream_t st,
                                    bs_solve_stats* stats) {
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (bs_pcg_workspace_bytes(nx, ny, nz) > kScratchBytes) return (int)hipErrorInvalidValue;
    void* ws = nullptr;
    BS_RETURN_IF_ERR((hipError_t)scratch_setup(&ws));
    PcgBuffers B;
    carve(ws, (size_t)nx * ny * nz, B);
    B.x = x;
    return pcg_solve_body(B, div_u, mat, nx, ny, nz, dx, dt, rho, mat_fluid, mat_air, tol, max_iter, warm_start,
                          max_restarts, st, stats, false);
}

// ===========================================================================
// Flexible CG with a nonlinear preconditioner (the learned one)
//   z_{k+1} = M(r_{k+1}),  beta = r_{k+1}.(z_{k+1} - z_k) / (r_k.z_k)
// (Polak-Ribiere form, which keeps CG's short recurrence usable when M varies
// between iterations), p = z + beta p.  The solve falls back to plain PCG when
// the residual has not dropped by 2 % for 20 iterations or r.z loses
// positivity, and always ends with the FP64 true-residual test.
// ===========================================================================
namespace bs {

// x += alpha p, r -= alpha Ap; partials of r.r.
template <int VEC>
__global__ __launch_bounds__(256) void fcg_update_kernel(Tiles T, const float* __restrict__ p,
                                                         const float* __restrict__ ap, float* __restrict__ x,
                                                         float* __restrict__ r, const double* __restrict__ partA,
                                                         double* __restrict__ partB, PcgCtrl* ctrl) {
    if (ctrl->done) return;
    (void)partA;
    const double alpha = ctrl->alpha;  // formed by pap_reduce_kernel
    const float af = (float)alpha;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        float pv[VEC], av[VEC], xv[VEC], rv[VEC];
        // Ap and x are not read again before the next iteration's update:
        // non-temporal, so that they do not displace r (read next by the
        // preconditioner) in the caches.
        ldf<VEC>(p, c0, pv);
        ldf_nt<VEC>(ap, c0, av);
        ldf_nt<VEC>(x, c0, xv);
        ldf<VEC>(r, c0, rv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            xv[e] = xv[e] + af * pv[e];
            rv[e] = rv[e] - af * av[e];
            rr += (double)rv[e] * (double)rv[e];
        }
        stf_nt<VEC>(x, c0, xv);
        stf<VEC>(r, c0, rv);
    }
    double v[1] = {rr};
    write_partials<1>(v, partB);
}

// One workgroup: convergence and stall test after the update of iteration
// it = ctrl->iter, which it advances to it + 1 (the only writer of the
// counter inside an iteration, so the other kernels can read it freely).
__global__ __launch_bounds__(256) void fcg_check_kernel(const double* __restrict__ partB, unsigned G, PcgCtrl* ctrl) {
    if (ctrl->done) return;
    double v[1];
    sum_partials<1>(partB, G, v);
    if (threadIdx.x != 0) return;
    const int it = ctrl->iter;
    const double rr = v[0];
    ctrl->rr = rr;
    ctrl->iter = it + 1;
    if (rr < ctrl->stop) {
        ctrl->converged = 1;
        ctrl->done = 1;
        return;
    }
    if (rr < 0.98 * ctrl->best_rr) {
        ctrl->best_rr = rr;
        ctrl->best_it = it;
    } else if (it - ctrl->best_it >= 20) {
        ctrl->fallback = 1;
        ctrl->done = 1;
    }
}

// Partials of r.z_new and r.z_old (z_old may be null).
__global__ __launch_bounds__(256) void fcg_dots_kernel(int64_t n, const float* __restrict__ r,
                                                       const float* __restrict__ z, const float* __restrict__ z_old,
                                                       double* __restrict__ partZ, const PcgCtrl* ctrl) {
    if (ctrl->done) return;
    double a = 0.0, b = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) {
        a += (double)r[t] * (double)z[t];
        if (z_old) b += (double)r[t] * (double)z_old[t];
    }
    double v[2] = {a, b};
    write_partials<2>(v, partZ);
}

// p = z_new + beta p with beta from dots_reduce_kernel; init: p = z.
template <int VEC>
__global__ __launch_bounds__(256) void fcg_direction_kernel(Tiles T, const float* __restrict__ z,
                                                            float* __restrict__ p, const PcgCtrl* ctrl, int init) {
    if (ctrl->done) return;
    const float bf = (float)ctrl->beta;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        float zv[VEC], pv[VEC];
        ldf_nt<VEC>(z, c0, zv);  // z is next read as z_old an iteration later
        if (init) {
            stf<VEC>(p, c0, zv);
        } else {
            ldf<VEC>(p, c0, pv);
#pragma unroll
            for (int e = 0; e < VEC; ++e) pv[e] = zv[e] + bf * pv[e];
            stf<VEC>(p, c0, pv);
        }
    }
}

// nbr_term where the value is exactly +0 on every non-fluid cell: an air
// neighbour's v is then the +0 nbr_term would return, so only the solid
// case needs a select.  The directions p = z + beta p hold +0 there: every
// preconditioner output z is +0 off the fluid (the V-cycle's last pass and
// the network store 0.0f there), p starts as z, and +0 + beta * (+0) is +0
// for either sign of beta; planes and chunks outside the grid read as +0.
__device__ __forceinline__ float nbr_term_z(uint8_t c, float v, float ps) { return c == kCodeSolid ? ps : v; }

// The flexible-CG iteration's first pass (nz % 4 == 0, 16-byte aligned
// vectors), pcg_spmv_roll_kernel's stencil on the taller dir_shape tile: the direction
// p_new = z + beta p_old (p_new = z when ctrl->fresh), Ap = A p_new with the
// partials of p_new.Ap, and the previous iteration's deferred x += alpha
// p_old (not when fresh).  Each float operation is the one of the direction,
// SpMV and update kernels, so the iterates are bit-identical to theirs.
// p_new goes to another buffer than p_old: the ring cells of a tile belong to
// its neighbours, whose old values are still read there.
__global__ __launch_bounds__(kDirThreads) void fcg_spmv_dir_kernel(int nx, int ny, int nz, int ni_seg, int jtiles,
                                                           int ktiles, int tj, const uint8_t* __restrict__ code,
                                                           const float* __restrict__ z,
                                                           const float* __restrict__ p_old,
                                                           float* __restrict__ p_new, float* __restrict__ x,
                                                           float* __restrict__ ap, double* __restrict__ part,
                                                           const PcgCtrl* ctrl) {
    if (ctrl->done) return;
    const bool fresh = ctrl->fresh != 0;
    const float bf = (float)ctrl->beta, af = (float)ctrl->alpha;
    constexpr int CH = kRollKT + 2, SR = kDirTJ + 2, W = 4 * CH;
    static_assert(CH * SR == kDirThreads, "one thread per chunk of the tile and its ring");
    __shared__ __align__(16) float sV[2][SR * W + 8];
    __shared__ __align__(16) uint8_t sC[2][SR * W + 16];
    const int tid = threadIdx.x;
    const int r = tid / CH, c = tid - (tid / CH) * CH;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int cols = jtiles * ktiles;
    const int seg = (int)(b / cols), col = (int)(b % cols);
    const int jt = col / ktiles, kt = col - (col / ktiles) * ktiles;
    const int j0 = jt * tj, k0 = kt * 4 * kRollKT;
    const int i0 = seg * ni_seg, i1 = min(nx, i0 + ni_seg);
    const int j = j0 - 1 + r, k = k0 - 4 + 4 * c;
    // rows past the tile's tj rows and its ring stay idle
    const bool jk_in = r <= tj + 1 && j >= 0 && j < ny && k >= 0 && k < nz;
    const bool interior = r >= 1 && r <= tj && c >= 1 && c <= kRollKT && j < ny && k < nz;
    const int64_t pl = (int64_t)ny * nz;
    const int64_t off = (int64_t)j * nz + k;
    // planes outside the grid and chunks outside (j,k) read as air (value 0)
    auto ldv = [&](const float* f, int q) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(f + (int64_t)q * pl + off);
        return v;
    };
    auto ldk = [&](int q) {
        uint32_t v = 0;
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const uint32_t*>(code + (int64_t)q * pl + off);
        return v;
    };
    auto ldx = [&](int q) {
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (interior && !fresh && q < i1) ldf_nt<4>(x, (int64_t)q * pl + off, v);
        return make_float4(v[0], v[1], v[2], v[3]);
    };
    auto dir = [&](float4 zv, float4 pv) {
        if (fresh) return zv;
        return make_float4(zv.x + bf * pv.x, zv.y + bf * pv.y, zv.z + bf * pv.z, zv.w + bf * pv.w);
    };
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // raw z and p_old two planes ahead; directions at q - 1, q (q + 1 is
    // formed at the top of iteration q, a plane after its loads were issued)
    float4 Pc = fresh ? zero4 : ldv(p_old, i0);
    float4 Zc = ldv(z, i0);
    float4 Vm = dir(ldv(z, i0 - 1), fresh ? zero4 : ldv(p_old, i0 - 1)), Vc = dir(Zc, Pc);
    float4 Zn = ldv(z, i0 + 1), Pn = fresh ? zero4 : ldv(p_old, i0 + 1);
    float4 Xc = ldx(i0);
    uint32_t Km = ldk(i0 - 1), Kc = ldk(i0), Kn = ldk(i0 + 1);
    double pap = 0.0, apz = 0.0;
    for (int q = i0; q < i1; ++q) {
        float* V = sV[q & 1] + 4;
        uint8_t* C = sC[q & 1] + 8;
        *reinterpret_cast<float4*>(V + r * W + 4 * c) = Vc;
        *reinterpret_cast<uint32_t*>(C + r * W + 4 * c) = Kc;
        const float4 Zf = ldv(z, q + 2);
        const float4 Pf = fresh ? zero4 : ldv(p_old, q + 2);
        const uint32_t Kf = ldk(q + 2);
        const float4 Xn = ldx(q + 1);
        const float4 Vn = dir(Zn, Pn);
        __syncthreads();
        if (interior) {
            const float4 d4 = *reinterpret_cast<const float4*>(V + (r - 1) * W + 4 * c);
            const float4 u4 = *reinterpret_cast<const float4*>(V + (r + 1) * W + 4 * c);
            const uint32_t kd = *reinterpret_cast<const uint32_t*>(C + (r - 1) * W + 4 * c);
            const uint32_t ku = *reinterpret_cast<const uint32_t*>(C + (r + 1) * W + 4 * c);
            const float vb = V[r * W + 4 * c - 1], vf = V[r * W + 4 * c + 4];
            const uint8_t cb = C[r * W + 4 * c - 1], cf = C[r * W + 4 * c + 4];
            const float vc[4] = {Vc.x, Vc.y, Vc.z, Vc.w}, vl[4] = {Vm.x, Vm.y, Vm.z, Vm.w};
            const float vr[4] = {Vn.x, Vn.y, Vn.z, Vn.w}, vd[4] = {d4.x, d4.y, d4.z, d4.w};
            const float vu[4] = {u4.x, u4.y, u4.z, u4.w};
            float out[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint8_t cc = (Kc >> (8 * e)) & 0xFF;
                if (cc < kCodeFluid0) { out[e] = 0.0f; continue; }
                const float ps = vc[e];
                const uint8_t c_b = e > 0 ? (uint8_t)((Kc >> (8 * (e - 1))) & 0xFF) : cb;
                const float v_b = e > 0 ? vc[e - 1] : vb;
                const uint8_t c_f = e < 3 ? (uint8_t)((Kc >> (8 * (e + 1))) & 0xFF) : cf;
                const float v_f = e < 3 ? vc[e + 1] : vf;
                float s = nbr_term_z((uint8_t)((Km >> (8 * e)) & 0xFF), vl[e], ps) +
                          nbr_term_z((uint8_t)((Kn >> (8 * e)) & 0xFF), vr[e], ps);
                s = s + nbr_term_z((uint8_t)((kd >> (8 * e)) & 0xFF), vd[e], ps);
                s = s + nbr_term_z((uint8_t)((ku >> (8 * e)) & 0xFF), vu[e], ps);
                s = s + nbr_term_z(c_b, v_b, ps);
                s = s + nbr_term_z(c_f, v_f, ps);
                out[e] = 6.0f * ps - s;
                pap += (double)ps * (double)out[e];
            }
            apz += (double)out[0] * (double)Zc.x;
            apz += (double)out[1] * (double)Zc.y;
            apz += (double)out[2] * (double)Zc.z;
            apz += (double)out[3] * (double)Zc.w;
            const int64_t o = (int64_t)q * pl + off;
            *reinterpret_cast<float4*>(ap + o) = make_float4(out[0], out[1], out[2], out[3]);
            *reinterpret_cast<float4*>(p_new + o) = Vc;
            if (!fresh) {
                const float xv[4] = {Xc.x + af * Pc.x, Xc.y + af * Pc.y, Xc.z + af * Pc.z, Xc.w + af * Pc.w};
                stf_nt<4>(x, o, xv);
            }
        }
        Vm = Vc; Vc = Vn; Pc = Pn;
        Zc = Zn; Zn = Zf; Pn = Pf; Xc = Xn;
        Km = Kc; Kc = Kn; Kn = Kf;
    }
    double v[2] = {pap, apz};
    write_partials<2, kDirThreads>(v, part);
}

// p.Ap, alpha = r.z / p.Ap and Ap.z totals after the fused first pass.
__global__ __launch_bounds__(1024) void fcg_pap_reduce_kernel(const double* __restrict__ part, unsigned G,
                                                              PcgCtrl* ctrl) {
    if (ctrl->done) return;
    double v[2];
    reduce1024<2>(part, G, v);
    if (threadIdx.x != 0) return;
    const double rz = ctrl->rz[ctrl->iter & 1];
    ctrl->pAp = v[0];
    ctrl->alpha = v[0] != 0.0 ? rz / v[0] : 0.0;
    ctrl->apz = v[1];
}

// r -= alpha Ap (the update's residual half; x follows in the next
// iteration's first pass); partials of r.r.
template <int VEC>
__global__ __launch_bounds__(256) void fcg_rupdate_kernel(Tiles T, const float* __restrict__ ap,
                                                          float* __restrict__ r, double* __restrict__ partB,
                                                          const PcgCtrl* ctrl) {
    if (ctrl->done) return;
    const float af = (float)ctrl->alpha;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        float av[VEC], rv[VEC];
        ldf_nt<VEC>(ap, c0, av);
        ldf<VEC>(r, c0, rv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            rv[e] = rv[e] - af * av[e];
            rr += (double)rv[e] * (double)rv[e];
        }
        stf<VEC>(r, c0, rv);
    }
    double v[1] = {rr};
    write_partials<1>(v, partB);
}

// After the flexible CG stops: the last update's deferred x += alpha p (p of
// iteration ctrl->iter - 1, in p_buf[ctrl->iter & 1]); nothing when no
// iteration ran.
template <int VEC>
__global__ __launch_bounds__(256) void fcg_finish_kernel(Tiles T, const float* __restrict__ p0,
                                                         const float* __restrict__ p1, float* __restrict__ x,
                                                         const PcgCtrl* ctrl) {
    const int it = ctrl->iter;
    if (it <= 0) return;
    const float* p = (it & 1) ? p1 : p0;
    const float af = (float)ctrl->alpha;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        float pv[VEC], xv[VEC];
        ldf<VEC>(p, c0, pv);
        ldf<VEC>(x, c0, xv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) xv[e] = xv[e] + af * pv[e];
        stf<VEC>(x, c0, xv);
    }
}

__global__ __launch_bounds__(256) void count_fluid_kernel(int64_t n, const uint8_t* __restrict__ code, int* out) {
    unsigned cnt = 0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) cnt += code[t] >= kCodeFluid0;
    double v[1] = {(double)cnt};  // one atomic per workgroup
    block_sums<1>(v);
    if (threadIdx.x == 0) atomicAdd((unsigned int*)out, (unsigned int)v[0]);
}

__global__ void fcg_prepare_kernel(PcgCtrl* ctrl, int it0) {
    if (threadIdx.x != 0) return;
    ctrl->best_rr = 1e300;
    ctrl->best_it = it0;
    ctrl->fallback = 0;
    ctrl->iter = it0;
}

struct FcgBuffers {
    PcgBuffers B;
    float* z[2];
    float* p2;  // second direction buffer of the fused first pass (p = {B.p, p2})
    float* r2;  // second residual buffer of the folded update (r = {B.r, r2})
    bool fused;  // fcg_spmv_dir_kernel + fcg_rupdate_kernel instead of SpMV, update, direction
    bool fold;   // (fused) the residual update runs in the V-cycle's first pass instead
    unsigned fold_G;  // r.r partials of that pass
    double* partZ;
    void* pws;  // preconditioner workspace
    int kind;
    const float* params;
    const void* packed;
    const unsigned* n_fluid = nullptr;  // kind 1: the fine grid's fluid count (device); the V-cycle keeps its own
    unsigned partZ_G;  // workgroups that wrote partZ
    int nx, ny, nz;
};

// The convergence check after the folded update (mg_apply calls it right
// after its first fine pass).
static void fcg_fold_check(hipStream_t st, void* arg) {
    const FcgBuffers& F = *static_cast<const FcgBuffers*>(arg);
    fcg_check_kernel<<<1, 256, 0, st>>>(F.B.partT, F.fold_G, F.B.ctrl);
}

// z = M(r) (+ partials of r.z and r.z_old into partZ).  With r_out (V-cycle
// only), r is the residual before the update r_out = r - alpha Ap, which the
// cycle's first pass performs (followed by the convergence check).
// ---------------------------------------------------------------------------
// Pass timing (diagnostic, bs_fcg_pass_timing): while it is on, flexible-CG
// solves launch eagerly (no graphs) and pass_mark brackets each fine-level
// pass with HIP events on the solve's stream.
// ---------------------------------------------------------------------------
constexpr int kTimedPasses = 3, kTimedMarks = 256;
struct PassTimer {
    bool on = false;
    bool created = false;
    hipEvent_t ev[kTimedPasses][2][kTimedMarks];
    int n[kTimedPasses] = {0, 0, 0};
};
static PassTimer g_timer;

void pass_mark(hipStream_t st, int pass, int edge) {
    if (!g_timer.on || pass < 0 || pass >= kTimedPasses) return;
    int& k = g_timer.n[pass];
    if (k >= kTimedMarks) return;
    (void)hipEventRecord(g_timer.ev[pass][edge][k], st);
    if (edge == 1) ++k;
}

// BOILSIM_FCG_EAGER=1 launches the kernels directly instead of replaying the
// captured graphs (same results; used to check the graph path); pass timing
// needs the direct launches too.
static bool fcg_eager() {
    const char* e = getenv("BOILSIM_FCG_EAGER");
    return g_timer.on || (e && e[0] == '1');
}

static int fcg_precond(hipStream_t st, const FcgBuffers& F, const float* r, float* z, const float* z_old,
                       float* r_out = nullptr) {
    const PcgBuffers& B = F.B;
    const int* done = &B.ctrl->done;
    const int64_t n = (int64_t)F.nx * F.ny * F.nz;
    if (F.kind == 1) {
        // z = gamma N(r) with the A-norm-optimal gamma, so r.z >= 0 always.
        const int rc = unet_apply(st, F.pws, F.nx, F.ny, F.nz, F.params, F.packed, r, B.code, F.n_fluid, nullptr, 0,
                                  z, done);
        if (rc != hipSuccess) return rc;
        scale_to_optimal(st, F.nx, F.ny, F.nz, B.code, r, z, F.partZ, done);
        fcg_dots_kernel<<<kUNetPartWg, 256, 0, st>>>(n, r, z, z_old, F.partZ, B.ctrl);
        return (int)hipGetLastError();
    }
    // coarse solve of the 4x level: kinds 2/4 the network, 3 one Jacobi sweep
    // + 20 Chebyshev steps, 5 + 60 steps, 6 two more geometric levels
    const int coarse = F.kind == 2 ? 1 : F.kind == 4 ? 2 : F.kind == 5 ? 3 : F.kind == 6 ? 4 : 0;
    (void)n;
    MgFold fold;
    if (r_out) {
        fold.ap = B.ap;
        fold.alpha = &B.ctrl->alpha;
        fold.r_out = r_out;
        fold.rr_part = B.partT;
        fold.after = fcg_fold_check;
        fold.after_arg = const_cast<FcgBuffers*>(&F);
    }
    // The V-cycle's last sweep also writes the r.z and r.z_old partials.
    return mg_apply(st, F.pws, F.nx, F.ny, F.nz, B.code, r, z, z_old, F.partZ, coarse, F.params, F.packed,
                    done, r_out ? &fold : nullptr);
}

// One flexible-CG iteration of parity `par` (z_new = z[1-par], z_old = z[par]).
// Fused: the direction formed from z[par] and p[par] goes to p[1-par], and x
// takes the previous iteration's step in the same pass (the last step is
// applied by fcg_finish_kernel).  Otherwise SpMV, update (x and r) and
// direction run as separate passes on one p.
static int fcg_enqueue_iteration(const Tiles& T, int vec, const FcgBuffers& F, int par, hipStream_t st) {
    const PcgBuffers& B = F.B;
    if (F.fused) {
        const RollShape rs = dir_shape(T.nx, T.ny, T.nz);
        float* pb[2] = {B.p, F.p2};
        pass_mark(st, 0, 0);
        fcg_spmv_dir_kernel<<<rs.grid, kDirThreads, 0, st>>>(T.nx, T.ny, T.nz, rs.ni, rs.jtiles, rs.ktiles, rs.tj, B.code, F.z[par],
                                                     pb[par], pb[par ^ 1], B.x, B.ap, B.partT, B.ctrl);
        pass_mark(st, 0, 1);
        fcg_pap_reduce_kernel<<<1, 1024, 0, st>>>(B.partT, rs.grid, B.ctrl);
        if (F.fold) {
            // r[par] -> r[1-par] in the V-cycle's first pass, then the check
            float* rb[2] = {B.r, F.r2};
            const int rc = fcg_precond(st, F, rb[par], F.z[par ^ 1], nullptr, rb[par ^ 1]);
            if (rc != hipSuccess) return rc;
            dots_reduce_kernel<<<1, 1024, 0, st>>>(F.partZ, F.partZ_G, B.ctrl, 0, 1, g_poll.ring_dev);
            return (int)hipGetLastError();
        }
        BS_VEC_LAUNCH(fcg_rupdate_kernel, vec, B.nwg, st, T, B.ap, B.r, B.partB, B.ctrl);
    } else {
        enqueue_spmv(T, vec, B, -1, st, true);
        BS_VEC_LAUNCH(fcg_update_kernel, vec, B.nwg, st, T, B.p, B.ap, B.x, B.r, B.partA, B.partB, B.ctrl);
    }
    fcg_check_kernel<<<1, 256, 0, st>>>(B.partB, B.nwg, B.ctrl);
    // fused: r.z_old from the first pass's Ap.z instead of a z_old stream
    const int rc = fcg_precond(st, F, B.r, F.z[par ^ 1], F.fused ? nullptr : F.z[par]);
    if (rc != hipSuccess) return rc;
    dots_reduce_kernel<<<1, 1024, 0, st>>>(F.partZ, F.partZ_G, B.ctrl, 0, F.fused ? 1 : 0, g_poll.ring_dev);
    if (!F.fused) BS_VEC_LAUNCH(fcg_direction_kernel, vec, B.nwg, st, T, F.z[par ^ 1], B.p, B.ctrl, 0);
    return (int)hipGetLastError();
}

// The first direction after the preconditioner's first application z[0]:
// the fused first pass forms it itself (ctrl->fresh), otherwise p = z[0].
static void fcg_start(const Tiles& T, int vec, const FcgBuffers& F, hipStream_t st) {
    const PcgBuffers& B = F.B;
    dots_reduce_kernel<<<1, 1024, 0, st>>>(F.partZ, F.partZ_G, B.ctrl, 1);
    if (!F.fused) BS_VEC_LAUNCH(fcg_direction_kernel, vec, B.nwg, st, T, F.z[0], B.p, B.ctrl, 1);
}

// The fused first pass applies where the i-marching SpMV does;
// BOILSIM_FCG_SPLIT=1 selects the separate passes (the same iterates up to
// the rounding of r.z_old, which the fused path forms from Ap.z).
static bool fcg_fused(const Tiles& T, int vec, const PcgBuffers& B) {
    const char* e = getenv("BOILSIM_FCG_SPLIT");
    if (e && e[0] == '1') return false;
    // two partials (p.Ap, Ap.z) per workgroup go to the per-tile slots
    return roll_fits(T, vec, B, true) && 2 * dir_shape(T.nx, T.ny, T.nz).grid <= max_tiles(T.nx, T.ny, T.nz);
}

// The folded residual update: fused first pass, V-cycle preconditioner whose
// first fine pass can take it, and room for its partials in partT.
static bool fcg_fold(const FcgBuffers& F) {
    if (!F.fused || F.kind < 2) return false;
    const unsigned g = mg_fold_groups(F.nx, F.ny, F.nz);
    return g > 0 && g <= max_tiles(F.nx, F.ny, F.nz);
}

// Instantiated graphs of the flexible CG's launch units (an (even, odd)
// iteration pair, a single iteration of either parity, the start), keyed by
// everything the captured launches depend on.  The ~55 launches of one
// iteration (most of them small coarse-level kernels) become one graph
// launch.
struct UnitGraphKey {
    const void* ws;
    const float* x;
    const float* params;
    const void* packed;
    const unsigned* n_fluid;
    int nx, ny, nz, kind, vec, fused, fold;
    int unit;  // 0: the (even, odd) pair; 1 + par: the single iteration of parity par; 3: the start
    unsigned nwg;
    bool operator==(const UnitGraphKey& o) const { return memcmp(this, &o, sizeof(*this)) == 0; }
};
struct UnitGraphEntry {
    UnitGraphKey key;
    hipGraphExec_t exec;
    bool used;
};
static UnitGraphEntry g_graphs[24];
static int g_graph_next = 0;
static hipStream_t g_capture_stream = nullptr;

static int unit_graph(const Tiles& T, int vec, const FcgBuffers& F, hipGraphExec_t* out, int unit = 0) {
    UnitGraphKey key;
    memset(&key, 0, sizeof(key));
    key.ws = F.B.code;
    key.x = F.B.x;
    key.params = F.params;
    key.packed = F.packed;
    key.n_fluid = F.n_fluid;
    key.nx = F.nx; key.ny = F.ny; key.nz = F.nz; key.kind = F.kind; key.vec = vec; key.nwg = F.B.nwg;
    key.fused = F.fused ? 1 : 0;
    key.fold = F.fold ? 1 : 0;
    key.unit = unit;
    for (auto& e : g_graphs)
        if (e.used && e.key == key) { *out = e.exec; return hipSuccess; }
    if (!g_capture_stream) BS_RETURN_IF_ERR(hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking));
    BS_RETURN_IF_ERR(hipStreamBeginCapture(g_capture_stream, hipStreamCaptureModeRelaxed));
    int rc = hipSuccess;
    if (unit == 3) {
        rc = fcg_precond(g_capture_stream, F, F.B.r, F.z[0], nullptr);
        if (rc == hipSuccess) {
            fcg_start(T, vec, F, g_capture_stream);
            rc = (int)hipGetLastError();
        }
    } else {
        rc = fcg_enqueue_iteration(T, vec, F, unit == 0 ? 0 : unit - 1, g_capture_stream);
        if (rc == hipSuccess && unit == 0) rc = fcg_enqueue_iteration(T, vec, F, 1, g_capture_stream);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(g_capture_stream, &graph);
    if (rc != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    BS_RETURN_IF_ERR(ec);
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    BS_RETURN_IF_ERR(ei);
    UnitGraphEntry& slot = g_graphs[g_graph_next];
    g_graph_next = (g_graph_next + 1) % 24;
    if (slot.used) (void)hipGraphExecDestroy(slot.exec);
    slot.key = key;
    slot.exec = exec;
    slot.used = true;
    *out = exec;
    return hipSuccess;
}

// The first preconditioner application z[0] = M(r) and the first direction's
// scalars (one graph launch instead of the cycle's ~55 kernel launches).
static int fcg_begin(const Tiles& T, int vec, const FcgBuffers& F, hipStream_t st) {
    if (fcg_eager()) {
        const int rc = fcg_precond(st, F, F.B.r, F.z[0], nullptr);
        if (rc != hipSuccess) return rc;
        fcg_start(T, vec, F, st);
        return (int)hipGetLastError();
    }
    hipGraphExec_t g = nullptr;
    const int rc = unit_graph(T, vec, F, &g, 3);
    if (rc != hipSuccess) return rc;
    return (int)hipGraphLaunch(g, st);
}

// Runs iterations from the device counter (even on entry) in graph-launched
// units until done, polling the flag one unit behind, then applies the
// pending step to x.  With `copy_final` the final control block is copied to
// g_poll.final and the stream synchronised; without, the caller does that
// after queueing more work.
static int fcg_iterate(const Tiles& T, int vec, const FcgBuffers& F, int max_iter, hipStream_t st,
                       bool copy_final = true) {
    const PcgBuffers& B = F.B;
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    const bool eager = fcg_eager();
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};  // pair, single of parity 0, single of parity 1
    int rc = hipSuccess;
    for (int u = 0; u < 3 && !eager && rc == hipSuccess; ++u) rc = unit_graph(T, vec, F, &exec[u], u);
    if (rc != hipSuccess) return rc;
    hipEvent_t ev[2];
    BS_RETURN_IF_ERR(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    if (const hipError_t e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming); e != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return (int)e;
    }
    // Units (an iteration pair, or a single iteration) are queued two deep,
    // each followed by an event, and polled oldest first: once a unit's
    // event has completed, its last iteration's post in the pinned ring
    // (written by that iteration's last kernel) is visible.  Posts count
    // from the first direction, so the unit that ends after `it` iterations
    // of this call posted sequence number `it`.  A unit is milliseconds of GPU work, so the host
    // stays ahead.  From the residual's reduction per iteration between the
    // last two polls: once the units in flight are predicted to reach the
    // stopping test, no further unit is queued until that is confirmed;
    // within two iterations of it the units are single iterations (so an odd
    // count does not run a pair's second half as no-ops), and a single
    // iteration also restores the even parity pairs start from.  A wrong
    // prediction costs one host round trip of GPU idle time or one more
    // unit.
    int it = 0, head = 0, inflight = 0;
    int unit_its[2] = {0, 0};
    int unit_end[2] = {0, 0};
    double rr_prev = -1.0, rr_last = -1.0, stop = -1.0;
    int it_prev = 0, it_last = 0;
    bool finished = false;
    while (!finished && rc == hipSuccess) {
        int want = 2;
        double q1 = 1.0;  // per iteration
        int its_inflight = 0;
        for (int k = 0; k < inflight; ++k) its_inflight += unit_its[(head + k) & 1];
        if (rr_prev > 0.0 && rr_last > 0.0 && it_last > it_prev) q1 = pow(rr_last / rr_prev, 1.0 / (it_last - it_prev));
        const bool rated = rr_prev > 0.0 && rr_last > 0.0 && it_last > it_prev && q1 < 1.0;
        if (rated && inflight > 0 && rr_last * pow(q1, its_inflight) < stop) want = inflight;
        while (inflight < want && it < max_iter && rc == hipSuccess) {
            const int slot = (head + inflight) & 1;
            int n = 2;
            if (it & 1) {
                n = 1;  // back to even parity
            } else if (rated) {
                // singles once two more iterations are predicted to suffice:
                // a pair could then end with a no-op half (an iteration's
                // launches), a superfluous single costs only its poll
                double pred = rr_last;
                for (int k = 0; k < its_inflight + 2; ++k) pred *= q1;
                if (pred < stop) n = 1;
            }
            if (it + n > max_iter) n = 1;
            if (eager) {
                rc = fcg_enqueue_iteration(T, vec, F, it & 1, st);
                if (rc == hipSuccess && n == 2) rc = fcg_enqueue_iteration(T, vec, F, 1, st);
            } else {
                rc = hipGraphLaunch(n == 2 ? exec[0] : exec[1 + (it & 1)], st);
            }
            if (rc == hipSuccess) rc = hipEventRecord(ev[slot], st);
            if (rc == hipSuccess) {
                unit_its[slot] = n;
                unit_end[slot] = it + n;
                ++inflight;
                its_inflight += n;
                it += n;
            }
        }
        if (rc != hipSuccess || inflight == 0) break;
        rc = hipEventSynchronize(ev[head]);
        if (rc != hipSuccess) break;
        std::atomic_thread_fence(std::memory_order_acquire);
        const volatile PollRec& post = g_poll.ring[(unit_end[head] - 1) & (kPollRing - 1)];
        PollRec h;
        h.rr = post.rr;
        h.stop = post.stop;
        h.iter = post.iter;
        h.done = post.done;
        h.seq = post.seq;
        if (h.seq != unit_end[head]) {  // every iteration posts once: a missing post is a bug
            rc = (int)hipErrorUnknown;
            break;
        }
        head ^= 1;
        --inflight;
        if (h.done) finished = true;
        stop = h.stop;
        rr_prev = rr_last;
        it_prev = it_last;
        rr_last = h.rr;
        it_last = h.iter;
    }
    if (rc == hipSuccess && F.fused) {
        BS_VEC_LAUNCH(fcg_finish_kernel, vec, B.nwg, st, T, B.p, F.p2, B.x, B.ctrl);
        rc = (int)hipGetLastError();
    }
    if (rc == hipSuccess && copy_final) {
        rc = hipMemcpyAsync(g_poll.final, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st);
        if (rc == hipSuccess) rc = hipStreamSynchronize(st);
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

// Two values (r.z_new, r.z_old) per preconditioner workgroup.
static size_t partz_slots(int nx, int ny, int nz) {
    const size_t g = mg_dots_groups(nx, ny, nz);
    return 2 * (g > kUNetPartWg ? g : (size_t)kUNetPartWg);
}

// Preconditioner kinds that run the network.
static bool needs_unet(int kind) { return kind == 1 || kind == 2 || kind == 4; }

static size_t precond_ws_bytes(int nx, int ny, int nz, int kind) {
    if (kind == 1) return unet_ws_bytes(nx, ny, nz);
    if (kind >= 2) return mg_ws_bytes(nx, ny, nz);
    return 0;
}

}  // namespace bs

extern "C" size_t bs_fcg_workspace_bytes(int nx, int ny, int nz, int precond) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    return bs_pcg_workspace_bytes(nx, ny, nz) + 4 * align256(n * 4) + align256(partz_slots(nx, ny, nz) * sizeof(double)) +
           align256(precond_ws_bytes(nx, ny, nz, precond));
}

// Flexible CG with preconditioner `precond` (1: U-Net on this grid, 2: V-cycle
// with the U-Net on the 4x-coarsened grid followed by a Chebyshev polish, 3:
// the same V-cycle with one Jacobi sweep there instead of the network, 4:
// Jacobi sweeps, the U-Net on their residual, Jacobi sweeps), to a true
// relative residual tol; on a stall it continues with plain Jacobi-PCG from x.
extern "C" int bs_fcg_solve(float* x, const float* div_u, const int* mat, int nx, int ny, int nz, float dx,
                            float dt, float rho, int mat_fluid, int mat_air, double tol, int max_iter,
                            int warm_start, int precond, const float* unet_params, const void* unet_packed,
                            void* workspace, size_t workspace_bytes, hipStream_t st, bs_fcg_stats* stats) {
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (precond < 1 || precond > 6) return (int)hipErrorInvalidValue;
    if (needs_unet(precond) && (!unet_params || !unet_packed)) return (int)hipErrorInvalidValue;
    if (workspace_bytes < bs_fcg_workspace_bytes(nx, ny, nz, precond)) return (int)hipErrorInvalidValue;
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    const size_t n = (size_t)nx * ny * nz;
    FcgBuffers F;
    carve(workspace, n, F.B);
    PcgBuffers& B = F.B;
    char* c = (char*)workspace + bs_pcg_workspace_bytes(nx, ny, nz);
    F.z[0] = (float*)c; c += align256(n * 4);
    F.z[1] = (float*)c; c += align256(n * 4);
    F.p2 = (float*)c; c += align256(n * 4);
    F.r2 = (float*)c; c += align256(n * 4);
    F.partZ = (double*)c; c += align256(partz_slots(nx, ny, nz) * sizeof(double));
    F.pws = c;
    F.kind = precond;
    F.params = unet_params;
    F.packed = unet_packed;
    F.nx = nx; F.ny = ny; F.nz = nz;
    F.partZ_G = precond == 1 ? kUNetPartWg : mg_dots_groups(nx, ny, nz);
    B.x = x;
    const int vec = vec_for(nz, {x, div_u, mat, B.b, B.r, B.p, B.ap, B.code, F.z[0], F.z[1], F.p2, F.r2});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    B.nwg = pick_nwg(T, kMaxWg);
    B.partB = B.partA + B.nwg;
    B.partS = B.partA + 3 * B.nwg;
    F.fused = fcg_fused(T, vec, B);
    F.fold = fcg_fold(F);
    F.fold_G = F.fold ? mg_fold_groups(nx, ny, nz) : 0u;
    const float bscale = -rho * dx * dx / dt;
    pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, tol * tol, max_iter, 1);
    BS_VEC_LAUNCH(pcg_setup_kernel, vec, B.nwg, st, T, mat, div_u, bscale, mat_fluid, mat_air,
                  warm_start ? 0 : 1, B.code, B.b, B.x, B.r, nullptr, B.partS);
    pcg_setup_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, warm_start ? 0 : 1);
    if (warm_start) {
        BS_VEC_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partS, 1);
        pcg_residual_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1, 0);
    }
    BS_RETURN_IF_ERR(hipGetLastError());
    if (precond == 1) {
        // The network normalises r by its RMS over the fluid cells of this grid.
        count_fluid_kernel<<<256, 256, 0, st>>>((int64_t)n, B.code, &B.ctrl->n_fluid);
        F.n_fluid = (const unsigned*)&B.ctrl->n_fluid;
    }
    if (precond >= 2) {
        const int rc = mg_setup(st, F.pws, nx, ny, nz, mat, mat_fluid, mat_air, B.code, precond == 6);
        if (rc != hipSuccess) return rc;
    }
    fcg_prepare_kernel<<<1, 64, 0, st>>>(B.ctrl, 0);
    int rc = fcg_begin(T, vec, F, st);
    if (rc != hipSuccess) return rc;
    // The true residual is queued behind the last iteration before the host
    // waits: the flexible CG's counters arrive with it.
    rc = fcg_iterate(T, vec, F, max_iter, st, false);
    if (rc != hipSuccess) return rc;
    PcgCtrl h;
    int fcg_iters = -1, fell_back = 0;
    // True residual; plain PCG (residual replacement) while it is above tol.
    int restarts = 0, it0 = 0;
    for (;;) {
        BS_VEC_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, nullptr, B.p, B.partS, 0);
        pcg_residual_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 0, 0);
        BS_RETURN_IF_ERR(hipGetLastError());
        BS_RETURN_IF_ERR(hipMemcpyAsync(g_poll.final, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st));
        BS_RETURN_IF_ERR(hipStreamSynchronize(st));
        h = *g_poll.final;
        if (fcg_iters < 0) {
            fcg_iters = h.iter;
            fell_back = h.fallback;
            it0 = h.iter;
        }
        const double rel_true = h.bsq > 0.0 ? sqrt(h.rr_true / h.bsq) : 0.0;
        if (stats) {
            stats->relres_true = rel_true;
            stats->bnorm = sqrt(h.bsq);
        }
        if (h.bsq == 0.0 || rel_true <= tol) {
            if (stats) stats->converged = 1;
            break;
        }
        if (restarts >= 8 || it0 >= max_iter) break;
        ++restarts;
        BS_VEC_LAUNCH(pcg_residual_kernel, vec, B.nwg, st, T, B.code, B.b, B.x, B.r, B.p, B.partS, 1);
        pcg_residual_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1, it0);
        BS_RETURN_IF_ERR(hipGetLastError());
        rc = pcg_iterate(T, vec, B, it0, max_iter, st);
        if (rc != hipSuccess) return rc;
        it0 = g_poll.final->iter;
    }
    if (stats) {
        stats->iterations = fcg_iters;
        stats->fallback = fell_back;
        stats->pcg_iterations = it0 - fcg_iters;
        stats->restarts = restarts;
        stats->relres_recursive = h.bsq > 0.0 ? sqrt(h.rr / h.bsq) : 0.0;
    }
    return hipSuccess;
}

// ===========================================================================
// Mixed-precision iterative refinement
//   x (FP64) <- x + A^-1 (b - A x),  the residual formed in FP64 from the FP64
//   iterate and A^-1 applied approximately by a cold-started Jacobi-PCG whose
//   vectors r, p, Ap are stored in FP32, FP16 or BF16 (arithmetic in FP32,
//   dots in FP64), on the residual scaled to unit RMS over the fluid cells.
// The FP64 iterate is what removes the FP32 floor on the true residual: for a
// smooth right-hand side |x| >> |b|, and rounding x to FP32 alone leaves a
// relative residual of order 1e-5.
// ===========================================================================
namespace bs {

struct RefCtrl {
    double bb;          // b.b of the outer problem
    double rr;          // |b - A x|^2 at the current FP64 iterate
    double scale;       // inner residual = scale * (b - A x), unit RMS over fluid
    double scale_prev;  // the scale before the last residual pass
    unsigned n_fluid;   // fluid cells
    int pad;
};

template <int VEC>
__device__ __forceinline__ void ldd(const double* __restrict__ p, int64_t off, double (&v)[VEC]) {
    if constexpr (VEC == 2) {
        const double2 t = *reinterpret_cast<const double2*>(p + off);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = p[off];
    }
}

__device__ __forceinline__ double nbr64(uint8_t c, double v, double xs) {
    return c >= kCodeFluid0 ? v : (c == kCodeSolid ? xs : 0.0);
}

// r = b - A x with x, the stencil and the subtraction in FP64 (stored as
// FP32); one partial of r.r per tile.
template <int VEC>
__global__ __launch_bounds__(256) void refine_residual_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                              const float* __restrict__ b,
                                                              const double* __restrict__ x, float* __restrict__ r,
                                                              double* __restrict__ part) {
    double rr = 0.0;
    int i, j, k0;
    int64_t c0;
    if (one_tile(T, VEC, i, j, k0, c0)) {
        const int64_t plane = (int64_t)T.ny * T.nz;
        uint8_t cc[VEC], cl[VEC], cr[VEC], cd[VEC], cu[VEC];
        double xc[VEC], xl[VEC], xr[VEC], xd[VEC], xu[VEC];
        ldc<VEC>(code, c0, cc);
        ldd<VEC>(x, c0, xc);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            cl[e] = cr[e] = cd[e] = cu[e] = kCodeAir;
            xl[e] = xr[e] = xd[e] = xu[e] = 0.0;
        }
        if (i > 0) { ldc<VEC>(code, c0 - plane, cl); ldd<VEC>(x, c0 - plane, xl); }
        if (i < T.nx - 1) { ldc<VEC>(code, c0 + plane, cr); ldd<VEC>(x, c0 + plane, xr); }
        if (j > 0) { ldc<VEC>(code, c0 - T.nz, cd); ldd<VEC>(x, c0 - T.nz, xd); }
        if (j < T.ny - 1) { ldc<VEC>(code, c0 + T.nz, cu); ldd<VEC>(x, c0 + T.nz, xu); }
        uint8_t cb = kCodeAir, cf = kCodeAir;
        double xb = 0.0, xf = 0.0;
        if (k0 > 0) { cb = code[c0 - 1]; xb = x[c0 - 1]; }
        if (k0 + VEC < T.nz) { cf = code[c0 + VEC]; xf = x[c0 + VEC]; }
        float bv[VEC], rv[VEC];
        ldf<VEC>(b, c0, bv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (cc[e] < kCodeFluid0) { rv[e] = 0.0f; continue; }
            const double xs = xc[e];
            const uint8_t c_b = e > 0 ? cc[e - 1] : cb;
            const double v_b = e > 0 ? xc[e - 1] : xb;
            const uint8_t c_f = e < VEC - 1 ? cc[e + 1] : cf;
            const double v_f = e < VEC - 1 ? xc[e + 1] : xf;
            double s = nbr64(cl[e], xl[e], xs) + nbr64(cr[e], xr[e], xs);
            s += nbr64(cd[e], xd[e], xs);
            s += nbr64(cu[e], xu[e], xs);
            s += nbr64(c_b, v_b, xs);
            s += nbr64(c_f, v_f, xs);
            const double r64 = (double)bv[e] - (6.0 * xs - s);
            rv[e] = (float)r64;
            rr += r64 * r64;
        }
        stf<VEC>(r, c0, rv);
    }
    double v[1] = {rr};
    write_partials<1>(v, part);
}

// what 0: b.b; 1: r.r and the inner scale sqrt(n_fluid / r.r).
__global__ __launch_bounds__(1024) void refine_reduce_kernel(const double* __restrict__ part, unsigned G,
                                                             RefCtrl* rc, int what) {
    double v[1];
    reduce1024<1>(part, G, v);
    if (threadIdx.x != 0) return;
    if (what == 0) {
        rc->bb = v[0];
    } else {
        rc->rr = v[0];
        rc->scale_prev = rc->scale;
        rc->scale = v[0] > 0.0 ? sqrt((double)rc->n_fluid / v[0]) : 1.0;
    }
}

// Residual replacement (reliable update): r = scale * rs with rs the
// FP64-formed true residual, p rescaled by scale / scale_prev so it stays the
// same direction in the new units, e = 0; partials of r.r and r.z laid out as
// pcg_update_kernel writes them, for the direction step that follows.
template <int VEC, typename S>
__global__ __launch_bounds__(256) void refine_replace_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                             float* __restrict__ rs_e, const RefCtrl* rc,
                                                             S* __restrict__ r, S* __restrict__ p,
                                                             double* __restrict__ partB) {
    const float sc = (float)rc->scale;
    const float ratio = (float)(rc->scale / rc->scale_prev);
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double rr = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC];
        float rv[VEC], pv[VEC], zero[VEC];
        ldc<VEC>(code, c0, cc);
        ldf<VEC>(rs_e, c0, rv);
        ldv<VEC>(p, c0, pv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            zero[e] = 0.0f;
            rv[e] = cc[e] >= kCodeFluid0 ? through<S>(sc * rv[e]) : 0.0f;
            pv[e] = through<S>(ratio * pv[e]);
            const float z = precond(rv[e], cc[e]);
            rr += (double)rv[e] * (double)rv[e];
            rz += (double)rv[e] * (double)z;
        }
        stf<VEC>(rs_e, c0, zero);
        stv<VEC>(r, c0, rv);
        stv<VEC>(p, c0, pv);
    }
    double v[2] = {rr, rz};
    write_partials<2>(v, partB);
}

// Resumes the inner PCG after a replacement at iteration it: r.z of the last
// direction in the new units, a fresh stop level (the replaced residual has
// r.r = n_fluid by construction of the scale).
__global__ void refine_resume_kernel(PcgCtrl* ctrl, const RefCtrl* rc, int it, double inner_tol2) {
    if (threadIdx.x != 0) return;
    const double ratio = rc->scale / rc->scale_prev;
    ctrl->rz[it & 1] *= ratio * ratio;
    ctrl->done = 0;
    ctrl->converged = 0;
    ctrl->bsq = (double)rc->n_fluid;
    ctrl->stop = inner_tol2 * (double)rc->n_fluid;
}

// Reduction the next inner segment aims at: inner_tol, or less when that
// would overshoot the outer tolerance (0.7 leaves room for the gap between
// the recursive and the true residual).
static double segment_target(double inner_tol, double tol, double rel) {
    const double need = rel > 0.0 ? 0.7 * tol / rel : inner_tol;
    return need > inner_tol ? (need < 1.0 ? need : 1.0) : inner_tol;
}

// Inner cold start from the stored residual rs (FP32): b_in = scale * rs
// rounded to S, e = 0 (in place of rs), r = b_in, p = M^-1 b_in; partials of
// b_in.b_in and r.z.
template <int VEC, typename S>
__global__ __launch_bounds__(256) void refine_inner_start_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                                 float* __restrict__ rs_e, const RefCtrl* rc,
                                                                 S* __restrict__ r, S* __restrict__ p,
                                                                 double* __restrict__ partS) {
    const float sc = (float)rc->scale;
    uint32_t t0, t1;
    tile_span(T.ntiles, t0, t1);
    double bb = 0.0, rz = 0.0;
    const int64_t plane = (int64_t)T.ny * T.nz;
    for (uint32_t t = t0; t < t1; ++t) {
        int i, j, k0;
        if (!tile_cell<VEC>(T, t, i, j, k0)) continue;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * T.nz + k0;
        uint8_t cc[VEC];
        float rv[VEC], zv[VEC], zero[VEC];
        ldc<VEC>(code, c0, cc);
        ldf<VEC>(rs_e, c0, rv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            zero[e] = 0.0f;
            rv[e] = cc[e] >= kCodeFluid0 ? through<S>(sc * rv[e]) : 0.0f;
            zv[e] = through<S>(precond(rv[e], cc[e]));
            bb += (double)rv[e] * (double)rv[e];
            rz += (double)rv[e] * (double)zv[e];
        }
        stf<VEC>(rs_e, c0, zero);
        stv<VEC>(r, c0, rv);
        stv<VEC>(p, c0, zv);
    }
    double v[2] = {bb, rz};
    write_partials<2>(v, partS);
}

// x += e / scale (FP64).
__global__ __launch_bounds__(256) void refine_accumulate_kernel(int64_t n, const float* __restrict__ e,
                                                                const RefCtrl* rc, double* __restrict__ x) {
    const double inv = 1.0 / rc->scale;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) x[t] += (double)e[t] * inv;
}

__global__ __launch_bounds__(256) void refine_zero_kernel(int64_t n, double* __restrict__ x) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) x[t] = 0.0;
}

__global__ __launch_bounds__(256) void refine_round_kernel(int64_t n, const double* __restrict__ x,
                                                           float* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) out[t] = (float)x[t];
}

__global__ void refine_ctrl_init_kernel(RefCtrl* rc) {
    if (threadIdx.x != 0) return;
    RefCtrl c;
    memset(&c, 0, sizeof(c));
    c.scale = 1.0;
    c.scale_prev = 1.0;
    *rc = c;
}

static int refine_vec(int nz, int storage) {
    if (storage != 0 && nz % 8 == 0) return 8;  // 16-byte accesses of 16-bit values
    return nz % 4 == 0 ? 4 : (nz % 2 == 0 ? 2 : 1);
}

static unsigned stream_grid(int64_t n) {
    const int64_t g = (n + 255) / 256;
    return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

// Outer state of a refinement: the FP32 residual (then the inner solution),
// the residual-pass partials and the control block, carved after the inner
// solver's own workspace.
struct RefineBufs {
    float* rs;
    double* rpart;
    RefCtrl* rc;
    Tiles T2;
    int v2;
};

static size_t refine_tail_bytes(int nx, int ny, int nz) {
    const size_t n = (size_t)nx * ny * nz;
    return align256(n * 4) + align256((size_t)make_tiles(nx, ny, nz, 1).ntiles * sizeof(double)) + 256;
}

// Codes and b (FP32, as the reference's compute_b) into B, b.b and the fluid
// count into the control block, x = 0 unless warm.
static int refine_begin(PcgBuffers& B, RefineBufs& R, char* tail, double* x, const float* div_u, const int* mat,
                        int nx, int ny, int nz, float dx, float dt, float rho, int mat_fluid, int mat_air,
                        int warm_start, hipStream_t st) {
    const int64_t n = (int64_t)nx * ny * nz;
    R.rs = (float*)tail; tail += align256((size_t)n * 4);
    R.rpart = (double*)tail; tail += align256((size_t)make_tiles(nx, ny, nz, 1).ntiles * sizeof(double));
    R.rc = (RefCtrl*)tail;
    R.v2 = (nz % 2 == 0 && ((uintptr_t)x % 16) == 0) ? 2 : 1;
    R.T2 = make_tiles(nx, ny, nz, R.v2);
    const int v4 = vec_for(nz, {div_u, mat, B.b, B.code});
    const Tiles T4 = make_tiles(nx, ny, nz, v4);
    const unsigned nwg4 = pick_nwg(T4, kMaxWg);
    const float bscale = -rho * dx * dx / dt;
    refine_ctrl_init_kernel<<<1, 64, 0, st>>>(R.rc);
    BS_VEC_LAUNCH(pcg_setup_kernel, v4, nwg4, st, T4, mat, div_u, bscale, mat_fluid, mat_air, 0, B.code, B.b,
                  nullptr, nullptr, nullptr, B.partS);
    refine_reduce_kernel<<<1, 1024, 0, st>>>(B.partS, nwg4, R.rc, 0);
    count_fluid_kernel<<<256, 256, 0, st>>>(n, B.code, (int*)&R.rc->n_fluid);
    if (!warm_start) refine_zero_kernel<<<stream_grid(n), 256, 0, st>>>(n, x);
    return (int)hipGetLastError();
}

// rs = b - A x (FP64-formed), its norm and the next inner scale; the control
// block is copied to h.
static int refine_residual(const PcgBuffers& B, const RefineBufs& R, const double* x, hipStream_t st, RefCtrl* h) {
    if (R.v2 == 2) refine_residual_kernel<2><<<R.T2.ntiles, 256, 0, st>>>(R.T2, B.code, B.b, x, R.rs, R.rpart);
    else refine_residual_kernel<1><<<R.T2.ntiles, 256, 0, st>>>(R.T2, B.code, B.b, x, R.rs, R.rpart);
    refine_reduce_kernel<<<1, 1024, 0, st>>>(R.rpart, R.T2.ntiles, R.rc, 1);
    BS_RETURN_IF_ERR(hipGetLastError());
    BS_RETURN_IF_ERR(hipMemcpyAsync(h, R.rc, sizeof(*h), hipMemcpyDeviceToHost, st));
    return (int)hipStreamSynchronize(st);
}

// One inner step of Jacobi-PCG in the refinement, vectors r, p, Ap stored as
// S: either a cold start on the scaled residual (cold), or the recursive
// residual replaced by the true one with the search direction kept and
// rescaled, so the Krylov process continues instead of restarting and the
// refinement costs no extra iterations.  Runs until the recursive residual
// has dropped by f or the iteration counter `it` (continued across steps)
// reaches max_inner.
template <typename S>
static int refine_pcg_step(const Tiles& Ti, int vi, const PcgBuffers& B, const RefineBufs& R, bool cold, double f,
                           int& it, int max_inner, hipStream_t st) {
    S* r = reinterpret_cast<S*>(B.r);
    S* p = reinterpret_cast<S*>(B.p);
    if (cold) {
        pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, f * f, max_inner, 2);
        BS_VECS_LAUNCH(refine_inner_start_kernel, S, vi, B.nwg, st, Ti, B.code, R.rs, R.rc, r, p, B.partS);
        pcg_setup_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1);
    } else {
        BS_VECS_LAUNCH(refine_replace_kernel, S, vi, B.nwg, st, Ti, B.code, R.rs, R.rc, r, p, B.partB);
        refine_resume_kernel<<<1, 64, 0, st>>>(B.ctrl, R.rc, it - 1, f * f);
        BS_VECS_LAUNCH(pcg_direction_kernel, S, vi, B.nwg, st, Ti, B.code, r, p, B.partB, B.ctrl, it - 1);
    }
    BS_RETURN_IF_ERR(hipGetLastError());
    const int rc = pcg_iterate_t<S>(Ti, vi, B, it, max_inner, st);
    if (rc != hipSuccess) return rc;
    it = g_poll.final->iter;
    return hipSuccess;
}

}  // namespace bs

extern "C" size_t bs_refine_workspace_bytes(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    const size_t n = (size_t)nx * ny * nz;
    (void)n;
    return bs_pcg_workspace_bytes(nx, ny, nz) + refine_tail_bytes(nx, ny, nz);
}

// Solves A x = b (b from div_u as bs_pcg_solve) for an FP64 x to a true
// relative residual tol (FP64 residual of the FP64 iterate) by iterative
// refinement.  storage: 0 FP32, 1 FP16, 2 BF16 vectors in the inner PCG, which
// runs to a recursive relative residual inner_tol (at most max_inner
// iterations) per refinement step; at most max_outer steps.  warm_start uses
// x's contents; x32 (nullable) receives x rounded to FP32.
extern "C" int bs_refine_solve(double* x, float* x32, const float* div_u, const int* mat, int nx, int ny, int nz,
                               float dx, float dt, float rho, int mat_fluid, int mat_air, double tol, int max_outer,
                               int storage, double inner_tol, int max_inner, int warm_start, void* workspace,
                               size_t workspace_bytes, hipStream_t st, bs_refine_stats* stats) {
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (storage < 0 || storage > 2 || max_outer < 0) return (int)hipErrorInvalidValue;
    if (workspace_bytes < bs_refine_workspace_bytes(nx, ny, nz)) return (int)hipErrorInvalidValue;
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    const int64_t n = (int64_t)nx * ny * nz;
    PcgBuffers B;
    carve(workspace, (size_t)n, B);
    RefineBufs R;
    int rcode = refine_begin(B, R, (char*)workspace + bs_pcg_workspace_bytes(nx, ny, nz), x, div_u, mat, nx, ny,
                             nz, dx, dt, rho, mat_fluid, mat_air, warm_start, st);
    if (rcode != hipSuccess) return rcode;
    float* rs = R.rs;
    RefCtrl* rc = R.rc;
    B.x = rs;
    const int vi = refine_vec(nz, storage);
    const Tiles Ti = make_tiles(nx, ny, nz, vi);
    B.nwg = pick_nwg(Ti, kMaxWg);
    B.partB = B.partA + B.nwg;
    B.partS = B.partA + 3 * B.nwg;
    int outer = 0, it = 0, conv = 0;
    double rel = 0.0, bb = 0.0, inner_rel = 0.0;
    for (;;) {
        // True residual of the FP64 iterate (and the scale of the next inner step).
        RefCtrl h;
        rcode = refine_residual(B, R, x, st, &h);
        if (rcode != hipSuccess) return rcode;
        bb = h.bb;
        rel = bb > 0.0 ? sqrt(h.rr / bb) : 0.0;
        if (bb == 0.0 || rel <= tol) { conv = 1; break; }
        if (outer >= max_outer || it >= max_inner) break;
        const double f = segment_target(inner_tol, tol, rel);
        const bool cold = outer == 0;
        ++outer;
        if (storage == 0) rcode = refine_pcg_step<float>(Ti, vi, B, R, cold, f, it, max_inner, st);
        else if (storage == 1) rcode = refine_pcg_step<_Float16>(Ti, vi, B, R, cold, f, it, max_inner, st);
        else rcode = refine_pcg_step<__bf16>(Ti, vi, B, R, cold, f, it, max_inner, st);
        if (rcode != hipSuccess) return rcode;
        inner_rel = g_poll.final->rr > 0.0 && h.n_fluid ? sqrt(g_poll.final->rr / (double)h.n_fluid) : 0.0;
        refine_accumulate_kernel<<<stream_grid(n), 256, 0, st>>>(n, rs, rc, x);
    }
    if (x32) refine_round_kernel<<<stream_grid(n), 256, 0, st>>>(n, x, x32);
    BS_RETURN_IF_ERR(hipGetLastError());
    BS_RETURN_IF_ERR(hipStreamSynchronize(st));
    if (stats) {
        stats->outer = outer;
        stats->inner_iterations = it;
        stats->converged = conv;
        stats->storage = storage;
        stats->relres_true = rel;
        stats->bnorm = sqrt(bb);
        stats->relres_inner_last = inner_rel;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// The same FP64 refinement with the learned flexible CG as the inner solver
// (cold-started on each scaled residual; FP32 vectors).
// ---------------------------------------------------------------------------

extern "C" size_t bs_fcg_refine_workspace_bytes(int nx, int ny, int nz, int precond) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    return bs_fcg_workspace_bytes(nx, ny, nz, precond) + refine_tail_bytes(nx, ny, nz);
}

// x (FP64) solves A x = b to a true relative residual tol; each refinement
// step runs bs_fcg_solve's flexible CG (preconditioner `precond`, 1-4) on the
// residual scaled to unit RMS until its recursive residual drops by inner_tol
// (at most max_inner iterations per step).  stats.inner_iterations counts the
// flexible-CG iterations of all steps.
extern "C" int bs_fcg_refine_solve(double* x, float* x32, const float* div_u, const int* mat, int nx, int ny,
                                   int nz, float dx, float dt, float rho, int mat_fluid, int mat_air, double tol,
                                   int max_outer, double inner_tol, int max_inner, int warm_start, int precond,
                                   const float* unet_params, const void* unet_packed, void* workspace,
                                   size_t workspace_bytes, hipStream_t st, bs_refine_stats* stats) {
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (precond < 1 || precond > 6 || max_outer < 0) return (int)hipErrorInvalidValue;
    if (needs_unet(precond) && (!unet_params || !unet_packed)) return (int)hipErrorInvalidValue;
    if (workspace_bytes < bs_fcg_refine_workspace_bytes(nx, ny, nz, precond)) return (int)hipErrorInvalidValue;
    BS_RETURN_IF_ERR((hipError_t)ensure_poll());
    const size_t n = (size_t)nx * ny * nz;
    FcgBuffers F;
    carve(workspace, n, F.B);
    PcgBuffers& B = F.B;
    char* c = (char*)workspace + bs_pcg_workspace_bytes(nx, ny, nz);
    F.z[0] = (float*)c; c += align256(n * 4);
    F.z[1] = (float*)c; c += align256(n * 4);
    F.p2 = (float*)c; c += align256(n * 4);
    F.r2 = (float*)c; c += align256(n * 4);
    F.partZ = (double*)c; c += align256(partz_slots(nx, ny, nz) * sizeof(double));
    F.pws = c;
    F.kind = precond;
    F.params = unet_params;
    F.packed = unet_packed;
    F.nx = nx; F.ny = ny; F.nz = nz;
    F.partZ_G = precond == 1 ? kUNetPartWg : mg_dots_groups(nx, ny, nz);
    RefineBufs R;
    int rc = refine_begin(B, R, (char*)workspace + bs_fcg_workspace_bytes(nx, ny, nz, precond), x, div_u, mat, nx,
                          ny, nz, dx, dt, rho, mat_fluid, mat_air, warm_start, st);
    if (rc != hipSuccess) return rc;
    B.x = R.rs;  // the inner solution e
    const int vec = vec_for(nz, {B.x, B.b, B.r, B.p, B.ap, B.code, F.z[0], F.z[1], F.p2, F.r2});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    B.nwg = pick_nwg(T, kMaxWg);
    B.partB = B.partA + B.nwg;
    B.partS = B.partA + 3 * B.nwg;
    F.fused = fcg_fused(T, vec, B);
    F.fold = fcg_fold(F);
    F.fold_G = F.fold ? mg_fold_groups(nx, ny, nz) : 0u;
    if (precond == 1) {
        F.n_fluid = &R.rc->n_fluid;
    } else {
        rc = mg_setup(st, F.pws, nx, ny, nz, mat, mat_fluid, mat_air, B.code, precond == 6);
        if (rc != hipSuccess) return rc;
    }
    // After an inner flexible CG stalls (fallback), the remaining steps run
    // the plain Jacobi-PCG inner solve of bs_refine_solve (FP32 vectors), as
    // bs_fcg_solve hands a stalled solve over to plain PCG.
    const int vp = refine_vec(nz, 0);
    const Tiles Tp = make_tiles(nx, ny, nz, vp);
    int outer = 0, its = 0, conv = 0, fell_back = 0, pcg_it = 0, pcg_steps = 0;
    double rel = 0.0, bb = 0.0, inner_rel = 0.0;
    for (;;) {
        RefCtrl h;
        rc = refine_residual(B, R, x, st, &h);
        if (rc != hipSuccess) return rc;
        bb = h.bb;
        rel = bb > 0.0 ? sqrt(h.rr / bb) : 0.0;
        if (bb == 0.0 || rel <= tol) { conv = 1; break; }
        if (outer >= max_outer || its >= max_inner) break;
        ++outer;
        const double f = segment_target(inner_tol, tol, rel);
        if (!fell_back) {
            pcg_ctrl_init_kernel<<<1, 64, 0, st>>>(B.ctrl, f * f, max_inner - its, 1);
            BS_VECS_LAUNCH(refine_inner_start_kernel, float, vec, B.nwg, st, T, B.code, R.rs, R.rc, B.r, B.p,
                           B.partS);
            pcg_setup_finalize<<<1, 256, 0, st>>>(B.partS, B.nwg, B.ctrl, 1);
            fcg_prepare_kernel<<<1, 64, 0, st>>>(B.ctrl, 0);
            rc = fcg_begin(T, vec, F, st);
            if (rc != hipSuccess) return rc;
            rc = fcg_iterate(T, vec, F, max_inner - its, st);
            if (rc != hipSuccess) return rc;
            const PcgCtrl& g = *g_poll.final;
            its += g.iter;
            inner_rel = g.bsq > 0.0 ? sqrt(g.rr / g.bsq) : 0.0;
            if (g.fallback && !g.converged) {
                fell_back = 1;
                B.nwg = pick_nwg(Tp, kMaxWg);
                B.partB = B.partA + B.nwg;
                B.partS = B.partA + 3 * B.nwg;
            }
        } else {
            // The plain PCG's own counter continues across its steps; its cap
            // is what the total budget leaves.
            const int before = pcg_it;
            rc = refine_pcg_step<float>(Tp, vp, B, R, pcg_steps == 0, f, pcg_it, pcg_it + (max_inner - its), st);
            if (rc != hipSuccess) return rc;
            ++pcg_steps;
            its += pcg_it - before;
            inner_rel = g_poll.final->rr > 0.0 && h.n_fluid ? sqrt(g_poll.final->rr / (double)h.n_fluid) : 0.0;
        }
        refine_accumulate_kernel<<<stream_grid((int64_t)n), 256, 0, st>>>((int64_t)n, R.rs, R.rc, x);
    }
    if (x32) refine_round_kernel<<<stream_grid((int64_t)n), 256, 0, st>>>((int64_t)n, x, x32);
    BS_RETURN_IF_ERR(hipGetLastError());
    BS_RETURN_IF_ERR(hipStreamSynchronize(st));
    if (stats) {
        stats->outer = outer;
        stats->inner_iterations = its;
        stats->converged = conv;
        stats->storage = 0;
        stats->relres_true = rel;
        stats->bnorm = sqrt(bb);
        stats->relres_inner_last = inner_rel;
        stats->fallback = fell_back;
        stats->pcg_iterations = pcg_it;
    }
    return hipSuccess;
}

// Pass timing of the learned flexible CG (diagnostic; no counterpart in the
// reference).  enable != 0 clears the records and turns timing on: the
// following bs_fcg_solve / bs_fcg_refine_solve calls launch their kernels
// directly and record HIP events on their stream around each fine-level pass
// (pass 0: fcg_spmv_dir_kernel, direction + SpMV + p.Ap + the deferred x
// step; 1: pre2_restrict_kernel<true>, the V-cycle's first fine pass with the
// folded residual update; 2: prolong_post_kernel, its last fine pass).
// enable == 0 waits for the records, turns timing off and writes, per pass,
// the median and mean launch time in microseconds and the number of launches
// timed (arrays of 3; null skips).
extern "C" int bs_fcg_pass_timing(int enable, double* us_median, double* us_mean, int* counts) {
    std::lock_guard<std::recursive_mutex> solve_lock(g_solve_mu);
    using namespace bs;
    if (enable) {
        if (!g_timer.created) {
            for (int p = 0; p < kTimedPasses; ++p)
                for (int e = 0; e < 2; ++e)
                    for (int k = 0; k < kTimedMarks; ++k)
                        BS_RETURN_IF_ERR(hipEventCreate(&g_timer.ev[p][e][k]));
            g_timer.created = true;
        }
        for (int p = 0; p < kTimedPasses; ++p) g_timer.n[p] = 0;
        g_timer.on = true;
        return hipSuccess;
    }
    g_timer.on = false;
    for (int p = 0; p < kTimedPasses; ++p) {
        const int m = g_timer.n[p];
        double d[kTimedMarks];
        double sum = 0.0;
        for (int k = 0; k < m; ++k) {
            BS_RETURN_IF_ERR(hipEventSynchronize(g_timer.ev[p][1][k]));
            float ms = 0.0f;
            BS_RETURN_IF_ERR(hipEventElapsedTime(&ms, g_timer.ev[p][0][k], g_timer.ev[p][1][k]));
            d[k] = 1e3 * (double)ms;
            sum += d[k];
        }
        for (int a = 1; a < m; ++a)  // insertion sort for the median
            for (int b = a; b > 0 && d[b - 1] > d[b]; --b) {
                const double t = d[b];
                d[b] = d[b - 1];
                d[b - 1] = t;
            }
        if (us_median) us_median[p] = m ? (m & 1 ? d[m / 2] : 0.5 * (d[m / 2 - 1] + d[m / 2])) : 0.0;
        if (us_mean) us_mean[p] = m ? sum / m : 0.0;
        if (counts) counts[p] = m;
    }
    return hipSuccess;
}
