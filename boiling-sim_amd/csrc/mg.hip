// Multigrid-augmented learned preconditioner for the masked pressure Poisson
// operator (A = 6I - S with the _pressure_neighbour rules, python/boilingsim/
// fluid.py:222-240): one V-cycle over the fine grid and its 2x coarsening
// with damped-Jacobi smoothing, the 4x-coarsened grid solved approximately by
// the trained U-Net (or, for comparison, by damped Jacobi sweeps).
//
// Coarse classes: a coarse cell is fluid if any child is fluid, else air if
// any child is air, else solid; the coarse operator is the same stencil on
// the coarse classes.  Restriction r_c = 4 * mean over the 8 children (the
// 4 = (2h/h)^2 keeps the unscaled stencil consistent across levels);
// prolongation is cell-centred trilinear (weights 3/4, 1/4 per axis).
// Smoothing (all on the Jacobi splitting z += c D^-1 (r - A z)): before, on
// both fine levels, the two-step Chebyshev polynomial over [0.3, 2] from
// zero; after, one undamped sweep on the fine level (a second costs more time
// than the iterations it saves) and 4 Chebyshev steps over [0.02, 2] on the
// 2x-coarsened one.  The network's correction on the coarsest grid is
// polished there by 20 Chebyshev steps over [0.02, 2], which also keeps the
// preconditioner positive where the network alone is not.
#include "boilsim_hip.h"
#include "grid.hpp"
#include "precond_api.hpp"

#include <string.h>

namespace bs {

struct MLvl {
    int nx, ny, nz;
    int64_t n;
    FastDiv fz, fyz;
};

static MLvl mk(int nx, int ny, int nz) {
    MLvl L;
    L.nx = nx; L.ny = ny; L.nz = nz;
    L.n = (int64_t)nx * ny * nz;
    L.fz = make_fastdiv((uint32_t)nz);
    L.fyz = make_fastdiv((uint32_t)(ny * nz));
    return L;
}

__device__ __forceinline__ void ijk(const MLvl& L, uint32_t c, int& i, int& j, int& k) {
    i = (int)fdiv(c, L.fyz);
    const uint32_t rem = c - (uint32_t)i * (uint32_t)(L.ny * L.nz);
    j = (int)fdiv(rem, L.fz);
    k = (int)(rem - (uint32_t)j * (uint32_t)L.nz);
}

constexpr float kOmega = 0.8f;
constexpr unsigned kStepWg = 512;
// Polish of the coarsest level's correction: Chebyshev steps on the Jacobi
// splitting over the eigenvalue interval [kChebLo, 2] of D^-1 A.  Measured
// at 440x440x280 (velocity seeds 11/21/31, flexible-CG iterations): 10
// damped-Jacobi sweeps 34/40/37, 40 sweeps 26/28/26, 20 Chebyshev steps over
// [0.02, 2] 21/23/21 (12 steps: 31; [0.0067, 2] needs 20-24 steps for the
// same; [0.2, 2] is worse); each step costs about one sweep (~5.6 us).
constexpr int kChebSteps = 20;
constexpr double kChebLo = 0.02;
// Post-smoothing of the 2x-coarsened level after the prolongation: 4
// Chebyshev steps over [0.02, 2] instead of two damped-Jacobi sweeps (seeds
// 11/21/31: 21/23/21 -> 17/20/19 iterations; 2 steps or [0.067, 2]: 20-22).
constexpr int kL1ChebSteps = 4;
constexpr double kL1ChebLo = 0.02;
// Pre-smoothing on both fine levels: the two-step Chebyshev polynomial over
// [0.3, 2] instead of two sweeps damped by 0.8; post-smoothing on the fine
// level: one undamped Jacobi sweep (seeds 11/21/31: 17/20/19 -> 15/15/15
// iterations; [0.1, 2] or [0.5, 2] and post weights 0.7 or 0.9 are worse).
constexpr double kPreLo = 0.3;
constexpr float kPostW = 1.0f;
// Non-learned coarse solves of the 4x level, for comparison with the network
// (coarse kinds 3 and 4, preconditioner kinds 5 and 6):
//   kind 5: one damped-Jacobi sweep and kAblChebSteps Chebyshev steps over
//           [kAblChebLo, 2].  Measured at 440x440x280 (velocity seeds 11, 21,
//           31, 41, 51; flexible-CG iterations / solve ms): 60 steps over
//           [0.001, 2] 13/14/13/13/13 and 19.4-20.8 ms, against 15 in all five
//           and 21.4-21.5 ms for the network with its 20-step polish; 40 steps
//           over [0.001, 2] 15-16; 80 over [0.0005, 2] 12-13 at 19.4-21.2 ms;
//           60 or 80 over [0.005, 2] 15-17.  The network followed by 40 steps
//           over [0.001, 2] takes 13-15 iterations (20.2-23.3 ms): given the
//           stronger polish, it no longer beats a single Jacobi sweep;
//   kind 6: two more geometric levels below the 4x one (each: two-step
//           Chebyshev pre-smoothing, restriction; prolongation and
//           kDeepPostSteps Chebyshev post-smoothing steps), a bottom solve
//           of kBotIters Jacobi-PCG iterations in one workgroup, then the
//           same 20-step polish as the network path.
constexpr int kAblChebSteps = 60;
constexpr double kAblChebLo = 0.001;
constexpr int kDeepPostSteps = 4;
constexpr int kBotThreads = 1024, kBotMax = 16384, kBotPer = kBotMax / kBotThreads;
constexpr int kBotIters = 24;

__device__ __forceinline__ float inv_diag(uint8_t code) {
    const int dg = 8 - (int)code;
    return dg > 0 ? 1.0f / (float)dg : 1.0f;
}

// (A v)[c] with codes, out of bounds = air (the multigrid form, lap_mg).
__device__ __forceinline__ float apply_a(const MLvl& L, const uint8_t* __restrict__ code,
                                         const float* __restrict__ v, int64_t c, int i, int j, int k) {
    const uint8_t cc = code[c];
    if (cc < kCodeFluid0) return 0.0f;
    const int64_t plane = (int64_t)L.ny * L.nz;
    float s = 0.0f;
    if (i > 0) s += fz(code[c - plane], v[c - plane]);
    if (i < L.nx - 1) s += fz(code[c + plane], v[c + plane]);
    if (j > 0) s += fz(code[c - L.nz], v[c - L.nz]);
    if (j < L.ny - 1) s += fz(code[c + L.nz], v[c + L.nz]);
    if (k > 0) s += fz(code[c - 1], v[c - 1]);
    if (k < L.nz - 1) s += fz(code[c + 1], v[c + 1]);
    return lap_mg(cc, v[c], s);
}

// Coarse class ids (0 fluid, 1 solid, 2 air) from the fine codes.
__global__ __launch_bounds__(256) void coarse_class_kernel(MLvl F, MLvl C, const uint8_t* __restrict__ fcode,
                                                           int* __restrict__ cmat) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t cc = (int64_t)blockIdx.x * 256 + threadIdx.x; cc < C.n; cc += stride) {
        int I, J, K;
        ijk(C, (uint32_t)cc, I, J, K);
        bool fl = false, air = false;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) {
                    const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + d;
                    if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
                    const uint8_t q = fcode[((int64_t)i * F.ny + j) * F.nz + k];
                    fl |= q >= kCodeFluid0;
                    air |= q == kCodeAir;
                }
        cmat[cc] = fl ? 0 : (air ? 2 : 1);
    }
}

// The same for fine nz % 4 == 0 (and 4-byte aligned codes): a thread forms
// two coarse cells along k from four 4-byte rows of fine codes.
__global__ __launch_bounds__(256) void coarse_class2_kernel(MLvl F, MLvl C, const uint8_t* __restrict__ fcode,
                                                            int* __restrict__ cmat) {
    const int64_t half = C.n >> 1;  // C.nz is even
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t h = (int64_t)blockIdx.x * 256 + threadIdx.x; h < half; h += stride) {
        const int64_t cc = 2 * h;
        int I, J, K;
        ijk(C, (uint32_t)cc, I, J, K);
        uint32_t any_fl = 0u, any_air = 0u;  // byte e: fine k = 2K + e
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int i = 2 * I + a, j = 2 * J + b;
                if (i >= F.nx || j >= F.ny) continue;
                const uint32_t w = *reinterpret_cast<const uint32_t*>(fcode + ((int64_t)i * F.ny + j) * F.nz + 2 * K);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t q = (w >> (8 * e)) & 0xFFu;
                    any_fl |= (q >= kCodeFluid0 ? 1u : 0u) << e;
                    any_air |= (q == kCodeAir ? 1u : 0u) << e;
                }
            }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const uint32_t m = 3u << (2 * o);  // fine k = 2(K+o), 2(K+o)+1
            cmat[cc + o] = (any_fl & m) ? 0 : ((any_air & m) ? 2 : 1);
        }
    }
}

__global__ __launch_bounds__(256) void codes_from_mat_kernel(MLvl L, const int* __restrict__ mat, int mf,
                                                             int ma, uint8_t* __restrict__ code,
                                                             unsigned int* __restrict__ nfluid) {
    const int64_t plane = (int64_t)L.ny * L.nz;
    const int64_t stride = (int64_t)gridDim.x * 256;
    unsigned cnt = 0;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < L.n; c += stride) {
        const int m = mat[c];
        if (m != mf) { code[c] = m == ma ? kCodeAir : kCodeSolid; continue; }
        ++cnt;
        int i, j, k;
        ijk(L, (uint32_t)c, i, j, k);
        int ns = 0;
        auto solid = [&](int64_t o) { const int q = mat[o]; return (q != mf && q != ma) ? 1 : 0; };
        if (i > 0) ns += solid(c - plane);
        if (i < L.nx - 1) ns += solid(c + plane);
        if (j > 0) ns += solid(c - L.nz);
        if (j < L.ny - 1) ns += solid(c + L.nz);
        if (k > 0) ns += solid(c - 1);
        if (k < L.nz - 1) ns += solid(c + 1);
        code[c] = (uint8_t)(kCodeFluid0 + ns);
    }
    if (nfluid) {  // one atomic per workgroup (per-thread ones on one address serialise)
        double v[1] = {(double)cnt};
        block_sums<1>(v);
        if (threadIdx.x == 0) atomicAdd(nfluid, (unsigned int)v[0]);
    }
}

// z = w D^-1 r (first sweep from z = 0).
__global__ __launch_bounds__(256) void jacobi0_kernel(MLvl L, const uint8_t* __restrict__ code,
                                                      const float* __restrict__ r, float* __restrict__ z,
                                                      const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < L.n; c += stride) {
        const uint8_t cc = code[c];
        z[c] = cc >= kCodeFluid0 ? kOmega * r[c] * inv_diag(cc) : 0.0f;
    }
}

// zo = zi + w D^-1 (r - A zi).
__global__ __launch_bounds__(256) void jacobi_kernel(MLvl L, const uint8_t* __restrict__ code,
                                                     const float* __restrict__ r, const float* __restrict__ zi,
                                                     float* __restrict__ zo, const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < L.n; c += stride) {
        const uint8_t cc = code[c];
        if (cc < kCodeFluid0) { zo[c] = 0.0f; continue; }
        int i, j, k;
        ijk(L, (uint32_t)c, i, j, k);
        const float az = apply_a(L, code, zi, c, i, j, k);
        zo[c] = zi[c] + kOmega * (r[c] - az) * inv_diag(cc);
    }
}

// rc[parent] = 4 * mean over its children of (r - A z) (fluid children; the
// others contribute 0).  Coarse-centred, so no atomics.
__global__ __launch_bounds__(256) void restrict_kernel(MLvl F, MLvl C, const uint8_t* __restrict__ fcode,
                                                       const float* __restrict__ r, const float* __restrict__ z,
                                                       float* __restrict__ rc, const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t cc = (int64_t)blockIdx.x * 256 + threadIdx.x; cc < C.n; cc += stride) {
        int I, J, K;
        ijk(C, (uint32_t)cc, I, J, K);
        float s = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) {
                    const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + d;
                    if (i >= F.nx || j >= F.ny || k >= F.nz) continue;
                    const int64_t f = ((int64_t)i * F.ny + j) * F.nz + k;
                    if (fcode[f] < kCodeFluid0) continue;
                    s += r[f] - (z ? apply_a(F, fcode, z, f, i, j, k) : 0.0f);
                }
        rc[cc] = 0.5f * s;
    }
}

// res = r - A z (fluid cells), 0 elsewhere.
__global__ __launch_bounds__(256) void residual_kernel(MLvl L, const uint8_t* __restrict__ code,
                                                      const float* __restrict__ r, const float* __restrict__ z,
                                                      float* __restrict__ res, const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < L.n; c += stride) {
        if (code[c] < kCodeFluid0) { res[c] = 0.0f; continue; }
        int i, j, k;
        ijk(L, (uint32_t)c, i, j, k);
        res[c] = r[c] - apply_a(L, code, z, c, i, j, k);
    }
}

// z += e on every cell.
__global__ __launch_bounds__(256) void add_kernel(int64_t n, const float* __restrict__ e, float* __restrict__ z,
                                                  const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < n; c += stride) z[c] += e[c];
}

// Partials of r.e and e.(A e) for the optimal step along e.
__global__ __launch_bounds__(256) void step_partials_kernel(MLvl L, const uint8_t* __restrict__ code,
                                                            const float* __restrict__ r,
                                                            const float* __restrict__ e,
                                                            double* __restrict__ part, const int* __restrict__ done) {
    if (done && *done) return;
    // Four cells per thread and round, every load of the round issued before
    // any arithmetic: indices past the grid or the edges are clamped to the
    // cell itself and their terms dropped, so the sums (over the thread's
    // cells c, c + stride, ... in order; A e as apply_a forms it) are those
    // of one cell at a time.
    constexpr int B = 4;
    double re = 0.0, eae = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    const int64_t plane = (int64_t)L.ny * L.nz;
    for (int64_t c0 = (int64_t)blockIdx.x * 256 + threadIdx.x; c0 < L.n; c0 += B * stride) {
        float rv[B], ev[B], av[B];
        bool use[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int64_t cb = c0 + b * stride;
            const bool in = cb < L.n;
            const int64_t c = in ? cb : c0;
            int i, j, k;
            ijk(L, (uint32_t)c, i, j, k);
            const bool ok[6] = {i > 0, i < L.nx - 1, j > 0, j < L.ny - 1, k > 0, k < L.nz - 1};
            const int64_t nb[6] = {c - plane, c + plane, c - L.nz, c + L.nz, c - 1, c + 1};
            const uint8_t cc = code[c];
            ev[b] = e[c];
            rv[b] = r[c];
            float s = 0.0f;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int64_t a = ok[q] ? nb[q] : c;
                const float t = fz(code[a], e[a]);
                s += ok[q] ? t : 0.0f;  // s is never -0, so adding +0 leaves it as skipping would
            }
            av[b] = lap_mg(cc, ev[b], s);
            use[b] = in && cc >= kCodeFluid0;
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (!use[b]) continue;
            const double evd = ev[b];
            re += (double)rv[b] * evd;
            eae += evd * (double)av[b];
        }
    }
    double v[2] = {re, eae};
    write_partials<2>(v, part);
}

// e *= gamma, gamma = (r.e)/(e.A e), the A-norm-optimal multiple of e as a
// correction for A x = r; gamma = 0 when either product is not positive.
__global__ __launch_bounds__(256) void step_scale_kernel(int64_t n, const double* __restrict__ part, unsigned G,
                                                         float* __restrict__ e, const int* __restrict__ done) {
    if (done && *done) return;
    double v[2];
    sum_partials<2>(part, G, v);
    const float g = (v[0] > 0.0 && v[1] > 0.0) ? (float)(v[0] / v[1]) : 0.0f;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < n; c += stride) e[c] *= g;
}

// z[fine] += trilinear(ec)[fine] on fluid cells (cell-centred weights 3/4, 1/4).
__global__ __launch_bounds__(256) void prolong_kernel(MLvl F, MLvl C, const uint8_t* __restrict__ fcode,
                                                      const float* __restrict__ ec, float* __restrict__ z,
                                                      const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    const int64_t cplane = (int64_t)C.ny * C.nz;
    for (int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x; f < F.n; f += stride) {
        if (fcode[f] < kCodeFluid0) continue;
        int i, j, k;
        ijk(F, (uint32_t)f, i, j, k);
        const int I = i >> 1, J = j >> 1, K = k >> 1;
        int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1), K2 = K + ((k & 1) ? 1 : -1);
        I2 = I2 < 0 ? 0 : (I2 >= C.nx ? C.nx - 1 : I2);
        J2 = J2 < 0 ? 0 : (J2 >= C.ny ? C.ny - 1 : J2);
        K2 = K2 < 0 ? 0 : (K2 >= C.nz ? C.nz - 1 : K2);
        const int ii[2] = {I, I2}, jj[2] = {J, J2}, kk[2] = {K, K2};
        const float w[2] = {0.75f, 0.25f};
        float v = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d)
                    v += w[a] * w[b] * w[d] * ec[ii[a] * cplane + (int64_t)jj[b] * C.nz + kk[d]];
        z[f] += v;
    }
}

// ---------------------------------------------------------------------------
// Vectorised kernels for the two fine levels (VEC consecutive k cells per
// thread, one 256-thread tile per workgroup, XCD-aware order).
// ---------------------------------------------------------------------------

// w / diag for the smoother; v_rcp_f32 (1 ulp) instead of an IEEE divide —
// the smoother needs no exact rounding and this runs ~26 times per thread.
__device__ __forceinline__ float winv(uint8_t code, float w = kOmega) {
    const int dg = 8 - (int)code;
    const float inv = dg > 0 ? __builtin_amdgcn_rcpf((float)dg) : 1.0f;
    return code >= kCodeFluid0 ? w * inv : 0.0f;
}

// Coefficients of the two-step smoother from zero: z1 = c1 D^-1 r,
// z = c2 (z1 + c3 D^-1 (r - A z1)); damped Jacobi is (w, 1, w).
struct Pre2Coef {
    float c1, c2, c3;
};

// The two-step Chebyshev polynomial over [lo, 2]: c1 = c3 = g = 2/(2 + lo),
// c2 = om_2 = 1/(1 - 1/(2 s^2)), s = (2 + lo)/(2 - lo).
static Pre2Coef pre2_chebyshev(double lo) {
    const double th = 0.5 * (2.0 + lo), sg = th / (0.5 * (2.0 - lo));
    const float g = (float)(1.0 / th);
    return Pre2Coef{g, (float)(1.0 / (1.0 - 1.0 / (2.0 * sg * sg))), g};
}

// Two smoothing steps from z = 0 in one pass: z1 = c1 D^-1 r is formed on the
// fly at the cell and its neighbours, z = c2 (z1 + c3 D^-1 (r - A z1)).
template <int VEC>
__global__ __launch_bounds__(256) void pre2_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                   const float* __restrict__ r, float* __restrict__ z,
                                                   const int* __restrict__ done, Pre2Coef K) {
    if (done && *done) return;
    int i, j, k0;
    int64_t c0;
    if (!one_tile(T, VEC, i, j, k0, c0)) return;
    Nbhd<VEC> nb;
    nb.load(T, code, r, i, j, k0, c0);
    float rc[VEC], out[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        rc[e] = nb.vc[e];
        nb.vc[e] *= winv(nb.cc[e], K.c1);
        nb.vl[e] *= winv(nb.cl[e], K.c1);
        nb.vr[e] *= winv(nb.cr[e], K.c1);
        nb.vd[e] *= winv(nb.cd[e], K.c1);
        nb.vu[e] *= winv(nb.cu[e], K.c1);
    }
    nb.vb *= winv(nb.cb, K.c1);
    nb.vf *= winv(nb.cf, K.c1);
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        const float az = nb.apply_mg(e);
        out[e] = nb.cc[e] >= kCodeFluid0 ? K.c2 * (nb.vc[e] + winv(nb.cc[e], K.c3) * (rc[e] - az)) : 0.0f;
    }
    stf<VEC>(z, c0, out);
}

// zo = zi + wD^-1 (r - A zi).
template <int VEC>
__global__ __launch_bounds__(256) void smooth_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                     const float* __restrict__ r, const float* __restrict__ zi,
                                                     float* __restrict__ zo, const int* __restrict__ done) {
    if (done && *done) return;
    int i, j, k0;
    int64_t c0;
    if (!one_tile(T, VEC, i, j, k0, c0)) return;
    Nbhd<VEC> nb;
    nb.load(T, code, zi, i, j, k0, c0);
    float rv[VEC], out[VEC];
    ldf<VEC>(r, c0, rv);
#pragma unroll
    for (int e = 0; e < VEC; ++e)
        out[e] = nb.cc[e] >= kCodeFluid0 ? nb.vc[e] + winv(nb.cc[e]) * (rv[e] - nb.apply_mg(e)) : 0.0f;
    stf<VEC>(zo, c0, out);
}

// Chebyshev semi-iteration step on A x = r with the Jacobi splitting:
// x_out = x_prev + om (x_cur - x_prev + ga D^-1 (r - A x_cur)) on fluid
// cells, 0 elsewhere (x_out may be x_prev: read and written per cell).
// znf: x_cur is exactly +0 on every non-fluid cell (Nbhd::load_v, no
// neighbour codes); the caller's arrays otherwise.
// gmode 1 / 2: the step reads x_cur and x_prev (1) or x_prev alone (2) as
// stored by step_scale_kernel, gamma * value, with gamma formed from the
// step partials (G per value) as that kernel forms it: the first two steps
// of a polish that starts from the network's A-norm-optimal correction, with
// the scaling pass folded in (the same float products, the same bits).
template <int VEC>
__global__ __launch_bounds__(256) void cheb_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                   const float* __restrict__ r, const float* __restrict__ x_cur,
                                                   const float* x_prev, float* x_out, float om, float ga,
                                                   const int* __restrict__ done, const double* __restrict__ gpart,
                                                   unsigned gG, int gmode, int znf) {
    if (done && *done) return;
    float gam = 1.0f;
    if (gmode) {  // every thread of the workgroup, before any returns
        double v[2];
        sum_partials<2>(gpart, gG, v);
        gam = (v[0] > 0.0 && v[1] > 0.0) ? (float)(v[0] / v[1]) : 0.0f;
    }
    int i, j, k0;
    int64_t c0;
    if (!one_tile(T, VEC, i, j, k0, c0)) return;
    Nbhd<VEC> nb;
    if (znf) nb.load_v(T, code, x_cur, i, j, k0, c0);
    else nb.load(T, code, x_cur, i, j, k0, c0);
    float rv[VEC], pv[VEC], out[VEC];
    ldf<VEC>(r, c0, rv);
    ldf<VEC>(x_prev, c0, pv);
    if (gmode == 1) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            nb.vc[e] *= gam; nb.vl[e] *= gam; nb.vr[e] *= gam; nb.vd[e] *= gam; nb.vu[e] *= gam;
        }
        nb.vb *= gam;
        nb.vf *= gam;
    }
    if (gmode) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) pv[e] *= gam;
    }
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        const uint8_t cc = nb.cc[e];
        const int dg = 8 - (int)cc;
        const float dinv = dg > 0 ? __builtin_amdgcn_rcpf((float)dg) : 1.0f;
        out[e] = cc >= kCodeFluid0 ? pv[e] + om * (nb.vc[e] - pv[e] + ga * dinv * (rv[e] - nb.apply_mg(e))) : 0.0f;
    }
    stf<VEC>(x_out, c0, out);
}

// `steps` Chebyshev steps on A x = r over the interval [lo, 2] of D^-1 A from
// x, ending in x (x_{k+1} overwrites x_{k-1}, so the iterates alternate
// between x and tmp; steps even): om_1 = 1, om_2 = 1/(1 - 1/(2 s^2)),
// om_{k+1} = 1/(1 - om_k/(4 s^2)), ga = 2/(2 + lo), s = (2 + lo)/(2 - lo) —
// the Chebyshev semi-iterative method on the Jacobi splitting (Golub and
// Varga).  Graph-capturable: the coefficients are kernel arguments.
static void chebyshev(hipStream_t st, const Tiles& T, int vec, const uint8_t* code, const float* r, float* x,
                      float* tmp, int steps, double lo, const int* done, const double* gpart = nullptr,
                      unsigned gG = 0, bool znf = true);

// Last smoothing sweep of the V-cycle fused with the flexible-CG products:
// zo = zi + w D^-1 (r - A zi) and one partial of r.zo and of r.z_old per tile.
template <int VEC>
__global__ __launch_bounds__(256) void smooth_dots_kernel(Tiles T, const uint8_t* __restrict__ code,
                                                          const float* __restrict__ r, const float* __restrict__ zi,
                                                          float* __restrict__ zo, const float* __restrict__ z_old,
                                                          double* __restrict__ part, const int* __restrict__ done,
                                                          float w) {
    if (done && *done) return;
    double a = 0.0, b = 0.0;
    int i, j, k0;
    int64_t c0;
    if (one_tile(T, VEC, i, j, k0, c0)) {
        Nbhd<VEC> nb;
        nb.load(T, code, zi, i, j, k0, c0);
        float rv[VEC], out[VEC], zo_[VEC];
        // r and z_old are next read an iteration later: non-temporal
        ldf_nt<VEC>(r, c0, rv);
        if (z_old) ldf_nt<VEC>(z_old, c0, zo_);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            out[e] = nb.cc[e] >= kCodeFluid0 ? nb.vc[e] + winv(nb.cc[e], w) * (rv[e] - nb.apply_mg(e)) : 0.0f;
            a += (double)rv[e] * (double)out[e];
            if (z_old) b += (double)rv[e] * (double)zo_[e];
        }
        stf<VEC>(zo, c0, out);
    }
    double v[2] = {a, b};
    write_partials<2>(v, part);
}

// rc[I,J,K] = 0.5 * sum over the 8 children of (r - A z), coarse-centred:
// a thread owns one coarse cell and reads its children as 2-cell k pairs
// (fine nz even).
__global__ __launch_bounds__(256) void restrict2_kernel(Tiles F, Tiles C, const uint8_t* __restrict__ fcode,
                                                        const float* __restrict__ r, const float* __restrict__ z,
                                                        float* __restrict__ rc, const int* __restrict__ done) {
    if (done && *done) return;
    int I, J, K;
    int64_t cc;
    if (!one_tile(C, 1, I, J, K, cc)) return;
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int i = 2 * I + a;
        if (i >= F.nx) continue;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = 2 * J + b;
            if (j >= F.ny) continue;
            const int k0 = 2 * K;
            const int64_t f0 = ((int64_t)i * F.ny + j) * F.nz + k0;
            Nbhd<2> nb;
            nb.load(F, fcode, z, i, j, k0, f0);
            float rv[2];
            ldf<2>(r, f0, rv);
#pragma unroll
            for (int e = 0; e < 2; ++e)
                if (nb.cc[e] >= kCodeFluid0) s += rv[e] - nb.apply_mg(e);
        }
    }
    rc[cc] = 0.5f * s;
}

// rc = 0.5 * sum over children of (r - A z) for two consecutive coarse cells
// per thread (fine nz % 4 == 0): the thread's 2x2 fine columns of 4 cells and
// the 8 columns around them (a 4x4 patch without corners) are loaded once as
// float4 values and 4-byte code words and shared by the four residuals.
__global__ __launch_bounds__(256) void restrict4_kernel(Tiles F, Tiles C, const uint8_t* __restrict__ fcode,
                                                        const float* __restrict__ r, const float* __restrict__ z,
                                                        float* __restrict__ rc, const int* __restrict__ done) {
    if (done && *done) return;
    int I, J, K0;
    int64_t cc0;
    if (!one_tile(C, 2, I, J, K0, cc0)) return;
    const int i0 = 2 * I, j0 = 2 * J, k0 = 2 * K0;
    const int64_t plane = (int64_t)F.ny * F.nz;
    float zv[4][4][4];
    uint8_t cv[4][4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if ((a == 0 || a == 3) && (b == 0 || b == 3)) continue;  // corners are not neighbours
            const int i = i0 + a - 1, j = j0 + b - 1;
#pragma unroll
            for (int e = 0; e < 4; ++e) { zv[a][b][e] = 0.0f; cv[a][b][e] = kCodeAir; }
            if (i >= 0 && i < F.nx && j >= 0 && j < F.ny) {
                const int64_t c = (int64_t)i * plane + (int64_t)j * F.nz + k0;
                ldf<4>(z, c, zv[a][b]);
                ldc<4>(fcode, c, cv[a][b]);
            }
        }
    float acc[2] = {0.0f, 0.0f};
#pragma unroll
    for (int a = 1; a < 3; ++a)
#pragma unroll
        for (int b = 1; b < 3; ++b) {
            const int i = i0 + a - 1, j = j0 + b - 1;
            if (i >= F.nx || j >= F.ny) continue;
            const int64_t c = (int64_t)i * plane + (int64_t)j * F.nz + k0;
            float rv[4];
            ldf<4>(r, c, rv);
            float zb = 0.0f, zf = 0.0f;
            uint8_t cb = kCodeAir, cf = kCodeAir;
            if (k0 > 0) { cb = fcode[c - 1]; zb = z[c - 1]; }
            if (k0 + 4 < F.nz) { cf = fcode[c + 4]; zf = z[c + 4]; }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (cv[a][b][e] < kCodeFluid0) continue;
                const float ps = zv[a][b][e];
                float s = fz(cv[a - 1][b][e], zv[a - 1][b][e]) + fz(cv[a + 1][b][e], zv[a + 1][b][e]);
                s += fz(cv[a][b - 1][e], zv[a][b - 1][e]);
                s += fz(cv[a][b + 1][e], zv[a][b + 1][e]);
                s += e > 0 ? fz(cv[a][b][e - 1], zv[a][b][e - 1]) : fz(cb, zb);
                s += e < 3 ? fz(cv[a][b][e + 1], zv[a][b][e + 1]) : fz(cf, zf);
                acc[e >> 1] += rv[e] - lap_mg(cv[a][b][e], ps, s);
            }
        }
    const float out[2] = {0.5f * acc[0], 0.5f * acc[1]};
    stf<2>(rc, cc0, out);
}

// z += trilinear(ec) on fluid cells, VEC fine cells per thread.
template <int VEC>
__global__ __launch_bounds__(256) void prolong_vec_kernel(Tiles F, Tiles C, const uint8_t* __restrict__ fcode,
                                                          const float* __restrict__ ec, float* __restrict__ z,
                                                          const int* __restrict__ done) {
    if (done && *done) return;
    int i, j, k0;
    int64_t c0;
    if (!one_tile(F, VEC, i, j, k0, c0)) return;
    uint8_t cc[VEC];
    float zv[VEC];
    ldc<VEC>(fcode, c0, cc);
    ldf<VEC>(z, c0, zv);
    const int I = i >> 1, J = j >> 1;
    int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
    I2 = I2 < 0 ? 0 : (I2 >= C.nx ? C.nx - 1 : I2);
    J2 = J2 < 0 ? 0 : (J2 >= C.ny ? C.ny - 1 : J2);
    const int64_t cp = (int64_t)C.ny * C.nz;
    const int64_t rows[4] = {I * cp + (int64_t)J * C.nz, I * cp + (int64_t)J2 * C.nz, I2 * cp + (int64_t)J * C.nz,
                             I2 * cp + (int64_t)J2 * C.nz};
    const float wr[4] = {0.5625f, 0.1875f, 0.1875f, 0.0625f};
    // The VEC fine cells k0..k0+VEC-1 use coarse k in [k0/2 - 1, k0/2 + W - 2]
    // (k0 is even for VEC >= 2; a single cell needs 3 slots for odd k):
    // gather that window once per row (clamped).
    constexpr int W = VEC == 1 ? 3 : VEC / 2 + 2;
    const int kb = (k0 >> 1) - 1;
    float cw[W];  // y-x weighted coarse values over the k window
#pragma unroll
    for (int w = 0; w < W; ++w) {
        int kk = kb + w;
        kk = kk < 0 ? 0 : (kk >= C.nz ? C.nz - 1 : kk);
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q) acc += wr[q] * ec[rows[q] + kk];
        cw[w] = acc;
    }
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        if (cc[e] < kCodeFluid0) continue;
        const int k = k0 + e;
        const int K = k >> 1;
        int K2 = K + ((k & 1) ? 1 : -1);
        K2 = K2 < 0 ? 0 : (K2 >= C.nz ? C.nz - 1 : K2);
        // window slots of K and K2 (clamping inside the window matches the
        // clamped index because both clamp to the same boundary value)
        const int sK = K - kb, sK2 = (K2 - kb) < 0 ? 0 : ((K2 - kb) >= W ? W - 1 : (K2 - kb));
        zv[e] += 0.75f * cw[sK] + 0.25f * cw[sK2];
    }
    stf<VEC>(z, c0, zv);
}

// z += trilinear(ec) on fluid cells for the 2x2 fine columns of one coarse
// column, 4 fine k cells each (fine nz % 4 == 0): the 3x3 coarse rows
// (clamped at the edges) over the 4-value k window are loaded once and
// shared; per fine cell the arithmetic is that of prolong_vec_kernel.
__global__ __launch_bounds__(256) void prolong4_kernel(Tiles F, Tiles CV, const uint8_t* __restrict__ fcode,
                                                       const float* __restrict__ ec, const float* __restrict__ z_in,
                                                       float* __restrict__ z, const int* __restrict__ done) {
    if (done && *done) return;
    // CV tiles the (coarse nx, coarse ny, fine nz) index space, 4 k per thread.
    int I, J, k0;
    int64_t unused;
    if (!one_tile(CV, 4, I, J, k0, unused)) return;
    (void)unused;
    const int cnx = CV.nx, cny = CV.ny, cnz = (F.nz + 1) / 2;
    const int64_t cp = (int64_t)cny * cnz;
    const int kb = (k0 >> 1) - 1;
    float E[3][3][4];
#pragma unroll
    for (int da = 0; da < 3; ++da)
#pragma unroll
        for (int db = 0; db < 3; ++db) {
            int Ir = I + da - 1, Jr = J + db - 1;
            Ir = Ir < 0 ? 0 : (Ir >= cnx ? cnx - 1 : Ir);
            Jr = Jr < 0 ? 0 : (Jr >= cny ? cny - 1 : Jr);
            const int64_t row = (int64_t)Ir * cp + (int64_t)Jr * cnz;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                int kk = kb + w;
                kk = kk < 0 ? 0 : (kk >= cnz ? cnz - 1 : kk);
                E[da][db][w] = ec[row + kk];
            }
        }
    const int64_t plane = (int64_t)F.ny * F.nz;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int i = 2 * I + a, j = 2 * J + b;
            if (i >= F.nx || j >= F.ny) continue;
            const int64_t c0 = (int64_t)i * plane + (int64_t)j * F.nz + k0;
            uint8_t cc[4];
            float zv[4];
            ldc<4>(fcode, c0, cc);
            ldf<4>(z_in, c0, zv);
            // rows (I,J), (I,J2), (I2,J), (I2,J2) with I2 = I -/+ 1 for even/odd i
            const int a2 = a ? 2 : 0, b2 = b ? 2 : 0;
            float cw[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                float acc = 0.0f;
                acc += 0.5625f * E[1][1][w];
                acc += 0.1875f * E[1][b2][w];
                acc += 0.1875f * E[a2][1][w];
                acc += 0.0625f * E[a2][b2][w];
                cw[w] = acc;
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (cc[e] < kCodeFluid0) continue;
                const int k = k0 + e;
                const int K = k >> 1;
                int K2 = K + ((k & 1) ? 1 : -1);
                K2 = K2 < 0 ? 0 : (K2 >= cnz ? cnz - 1 : K2);
                const int sK = K - kb, sK2 = (K2 - kb) < 0 ? 0 : ((K2 - kb) >= 4 ? 3 : (K2 - kb));
                zv[e] += 0.75f * cw[sK] + 0.25f * cw[sK2];
            }
            stf<4>(z, c0, zv);
        }
}

// ---------------------------------------------------------------------------
// Pre-smoothing and restriction of one level in a single pass (nx, ny even,
// nz % 4 == 0, 16-byte aligned vectors).  A workgroup owns a tj x 56 (j,k)
// tile (tj even, at most SR - 4) and walks down an even-length i segment;
// each of its SR x 16 threads holds one 4-cell k chunk of the tile or of its
// ring (two rows in j, one chunk in k).  At plane q it forms the two-step smoother z = c2 (z1 + c3 D^-1 (r - A
// z1)), z1 = c1 D^-1 r, over the tile and a one-cell ring (pre2_kernel's
// arithmetic, so z is bit-identical), then the residual r - A z at plane q-1 over the
// tile, which it sums over the 2x2x2 children of each coarse cell (times 0.5,
// the restriction of restrict4_kernel; only the order of the sum differs).
// i-neighbours are the thread's registers, j/k-neighbours come from LDS
// planes (z1 and z, each double-buffered), which leaves one barrier per
// plane; the global loads run two planes ahead.  z1 is formed once per cell
// when its plane arrives (in registers for the i-neighbours, in the LDS plane
// for the j/k ones) rather than at each of its seven uses.  z1 and z are held
// at exactly +0 on non-fluid cells, so A uses the multigrid form lap_mg with
// no selects and no neighbour codes: only the centre's code is read, from
// registers (the pass is issue-bound, not HBM-bound, so VALU work is what
// these save).  The march through a segment is a chain of per-plane
// latencies, so the pass pays off only where the level fills the chip with
// workgroups (the fine level and its 2x coarsening).  On the fine level the
// tile is 56 x 56 on 1024 threads (of which the ring holds a quarter,
// against a third of 256 threads on 12 x 56) and the grid one round of one
// workgroup per CU on long segments: 440x440x280, 240 workgroups of 74
// planes, 219 us per pass, against 257 us for 4070 workgroups of 256
// threads on 20 planes (and 242 / 243 / 248 us for 2 / 3 / 4 rounds of the
// taller tile, 234 us for 44 x 56 tiles on 768 threads).  The 2x level
// keeps the 256-thread tile (the taller one there: the same time).  All
// columns of a segment walk the same planes together, which keeps HBM reads
// within few planes at a time: equal shares of (column, plane) pairs in
// column order, which would use all 256 CUs, ran at 435 us.
// ---------------------------------------------------------------------------
// Workgroups of `fn` (nt threads each) that one CU holds at once, from the
// runtime's occupancy calculator (the built kernel's register and LDS use).
template <typename F>
static int resident_wgs(F fn, int nt) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, nt, 0) != hipSuccess || n < 1) n = 1;
    return n;
}

constexpr int kPrKT = 14;
// Thread rows (tile rows + the 4 ring rows) of the two workgroup shapes: 256
// threads (the 2x level) and kPrBigSR * 16 threads (the fine level, where the
// taller tile spends fewer threads and L2 reads on the ring).
constexpr int kPrSmallSR = 16, kPrBigSR = 64;

struct PrShape {
    int jtiles, ktiles, ni, tj;
    unsigned grid;
};

// At most `target` workgroups (when the columns allow), i segments of even
// length >= 8 (two planes of each segment's smoother are recomputed by its
// neighbours); the rows spread evenly over the j tiles, an even number per
// tile (up to sr - 4).
static PrShape pr_shape(int nx, int ny, int nz, int target, int sr) {
    PrShape s;
    const int tmax = sr - 4;
    s.jtiles = (ny + tmax - 1) / tmax;
    s.tj = (ny + s.jtiles - 1) / s.jtiles;
    s.tj += s.tj & 1;
    s.ktiles = (nz + 4 * kPrKT - 1) / (4 * kPrKT);
    const int cols = s.jtiles * s.ktiles;
    int nseg = target / cols;
    nseg = nseg < 1 ? 1 : nseg;
    int ni = (nx + nseg - 1) / nseg;
    ni = ni < 8 ? 8 : ni;
    ni += ni & 1;
    s.ni = ni;
    nseg = (nx + ni - 1) / ni;
    s.grid = (unsigned)(nseg * cols);
    return s;
}

static bool pr_fits(const MLvl& L, int vec) { return vec == 4 && (L.nx & 1) == 0 && (L.ny & 1) == 0 && (L.nz & 3) == 0; }

__device__ __forceinline__ uint8_t byte_of(uint32_t w, int e) { return (uint8_t)((w >> (8 * e)) & 0xFFu); }

// FOLD: the flexible CG's residual update runs in the same pass: the kernel
// reads r_old and Ap, works on r = r_old - alpha Ap (alpha = *alpha_p, the
// update kernel's float operations), writes r to r_out on the tile (r_out
// must not alias r_old: the ring cells belong to neighbouring tiles, which
// still read their old values) and one partial of r.r per workgroup.
template <bool FOLD, int SR>
__global__ __launch_bounds__(SR * (kPrKT + 2)) void pre2_restrict_kernel(int nx, int ny, int nz, int ni_seg,
                                                                         int jtiles, int ktiles, int tj,
                                                                         const uint8_t* __restrict__ code,
                                                            const float* __restrict__ r, float* __restrict__ z,
                                                            float* __restrict__ rc, const int* __restrict__ done,
                                                            Pre2Coef K, const float* __restrict__ ap,
                                                            const double* __restrict__ alpha_p,
                                                            float* __restrict__ r_out, double* __restrict__ rr_part) {
    if (done && *done) return;
    const float af = FOLD ? (float)*alpha_p : 0.0f;
    constexpr int CH = kPrKT + 2, W = 4 * CH, NT = SR * CH;
    static_assert(SR % 4 == 0, "a wave holds whole row pairs");
    __shared__ __align__(16) float sR[2][SR * W + 8];
    __shared__ __align__(16) float sZ[2][SR * W + 8];
    const int tid = threadIdx.x;
    const int row = tid / CH, ch = tid - (tid / CH) * CH;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int cols = jtiles * ktiles;
    const int seg = (int)(b / cols), col = (int)(b % cols);
    const int jt = col / ktiles, kt = col - (col / ktiles) * ktiles;
    const int j0 = jt * tj, k0 = kt * 4 * kPrKT;
    const int i0 = seg * ni_seg, i1 = min(nx, i0 + ni_seg);
    const int j = j0 - 2 + row, k = k0 - 4 + 4 * ch;
    // rows past the tile's tj rows and its ring stay idle (air)
    const bool jk_in = row < tj + 4 && j >= 0 && j < ny && k >= 0 && k < nz;
    const bool smooth_row = row >= 1 && row <= tj + 2;
    const bool inner = row >= 2 && row < tj + 2 && ch >= 1 && ch <= kPrKT && j < ny && k < nz;
    const int64_t pl = (int64_t)ny * nz;
    const int64_t off = (int64_t)j * nz + k;
    // planes outside the grid and chunks outside (j,k) read as air (code 0, value 0)
    auto ldv = [&](int q) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(r + (int64_t)q * pl + off);
        return v;
    };
    auto ldk = [&](int q) {
        uint32_t v = 0;
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const uint32_t*>(code + (int64_t)q * pl + off);
        return v;
    };
    auto lda = [&](int q) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (FOLD && jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(ap + (int64_t)q * pl + off);
        return v;
    };
    auto fold = [&](float4 rv, float4 av) {
        if (!FOLD) return rv;
        return make_float4(rv.x - af * av.x, rv.y - af * av.y, rv.z - af * av.z, rv.w - af * av.w);
    };
    // z1 = c1 D^-1 r of one 4-cell chunk (+0 at non-fluid cells)
    auto z1of = [&](float4 rv, uint32_t kw) {
        const float v[4] = {rv.x, rv.y, rv.z, rv.w};
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint8_t c = byte_of(kw, e);
            o[e] = c >= kCodeFluid0 ? v[e] * winv(c, K.c1) : 0.0f;
        }
        return make_float4(o[0], o[1], o[2], o[3]);
    };
    const int cnx = nx >> 1, cny = ny >> 1, cnz = nz >> 1;
    (void)cnx;
    // r at planes q-1, q, q+1 (updated when FOLD); raw r and Ap at q+2, q+3
    float4 Rm = fold(ldv(i0 - 2), lda(i0 - 2)), Rc = fold(ldv(i0 - 1), lda(i0 - 1)), Rn = fold(ldv(i0), lda(i0));
    float4 Rf = ldv(i0 + 1), Af = lda(i0 + 1);
    uint32_t Km = ldk(i0 - 2), Kc = ldk(i0 - 1), Kn = ldk(i0), Kf = ldk(i0 + 1);
    // z1 at planes q-1, q, q+1
    float4 Z1m = z1of(Rm, Km), Z1c = z1of(Rc, Kc), Z1n = z1of(Rn, Kn);
    float Zmm[4] = {0.f, 0.f, 0.f, 0.f}, Zm[4] = {0.f, 0.f, 0.f, 0.f};
    float acc[2] = {0.f, 0.f};
    double rr = 0.0;
    for (int q = i0 - 1; q <= i1; ++q) {
        const int par = (q - i0 + 1) & 1;
        float* VR = sR[par] + 4;
        float* VZ = sZ[par] + 4;
        *reinterpret_cast<float4*>(VR + row * W + 4 * ch) = Z1c;
        const float4 Rg = ldv(q + 3);  // two planes of loads in flight
        const float4 Ag = lda(q + 3);
        const uint32_t Kg = ldk(q + 3);
        if (FOLD && inner && q >= i0 && q < i1) {
            *reinterpret_cast<float4*>(r_out + (int64_t)q * pl + off) = Rc;
            rr += (double)Rc.x * (double)Rc.x;
            rr += (double)Rc.y * (double)Rc.y;
            rr += (double)Rc.z * (double)Rc.z;
            rr += (double)Rc.w * (double)Rc.w;
        }
        __syncthreads();
        // --- smoother at plane q (tile and its one-cell ring) ---------------
        float Zc[4] = {0.f, 0.f, 0.f, 0.f};
        if (smooth_row) {
            // z1 of the j-neighbours and across the chunk edges from LDS (air,
            // +0, outside the tile row)
            const float4 d4 = *reinterpret_cast<const float4*>(VR + (row - 1) * W + 4 * ch);
            const float4 u4 = *reinterpret_cast<const float4*>(VR + (row + 1) * W + 4 * ch);
            float vb = 0.0f, vf = 0.0f;
            if (ch > 0) vb = VR[row * W + 4 * ch - 1];
            if (ch < CH - 1) vf = VR[row * W + 4 * ch + 4];
            const float rcv[4] = {Rc.x, Rc.y, Rc.z, Rc.w};
            const float vc[4] = {Z1c.x, Z1c.y, Z1c.z, Z1c.w}, vl[4] = {Z1m.x, Z1m.y, Z1m.z, Z1m.w};
            const float vr[4] = {Z1n.x, Z1n.y, Z1n.z, Z1n.w};
            const float vd[4] = {d4.x, d4.y, d4.z, d4.w}, vu[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint8_t cc = byte_of(Kc, e);
                if (cc < kCodeFluid0) continue;
                const float ps = vc[e];
                const float v_b = e > 0 ? vc[e - 1] : vb;
                const float v_f = e < 3 ? vc[e + 1] : vf;
                float s = vl[e] + vr[e];
                s = s + vd[e];
                s = s + vu[e];
                s = s + v_b;
                s = s + v_f;
                const float az = lap_mg(cc, ps, s);
                Zc[e] = K.c2 * (ps + winv(cc, K.c3) * (rcv[e] - az));
            }
            if (inner && q >= i0 && q < i1)
                *reinterpret_cast<float4*>(z + (int64_t)q * pl + off) = make_float4(Zc[0], Zc[1], Zc[2], Zc[3]);
        }
        *reinterpret_cast<float4*>(VZ + row * W + 4 * ch) = make_float4(Zc[0], Zc[1], Zc[2], Zc[3]);
        // --- residual at plane p = q - 1, summed into the coarse cells ------
        const int p = q - 1;
        if (p >= i0) {
            if (inner) {
                const float* PZ = sZ[par ^ 1] + 4;  // plane p's smoothed values
                const float4 d4 = *reinterpret_cast<const float4*>(PZ + (row - 1) * W + 4 * ch);
                const float4 u4 = *reinterpret_cast<const float4*>(PZ + (row + 1) * W + 4 * ch);
                const float vb = PZ[row * W + 4 * ch - 1], vf = PZ[row * W + 4 * ch + 4];
                const float rp[4] = {Rm.x, Rm.y, Rm.z, Rm.w};
                const float vd[4] = {d4.x, d4.y, d4.z, d4.w}, vu[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint8_t cc = byte_of(Km, e);
                    if (cc < kCodeFluid0) continue;
                    const float ps = Zm[e];
                    const float v_b = e > 0 ? Zm[e - 1] : vb;
                    const float v_f = e < 3 ? Zm[e + 1] : vf;
                    float s = Zmm[e] + Zc[e];
                    s += vd[e];
                    s += vu[e];
                    s += v_b;
                    s += v_f;
                    acc[e >> 1] += rp[e] - lap_mg(cc, ps, s);
                }
            }
            if (p & 1) {  // both fine planes of coarse plane p/2 are in: add row pairs
                const float o0 = __shfl_down(acc[0], CH, 64);
                const float o1 = __shfl_down(acc[1], CH, 64);
                if (inner && (row & 1) == 0) {
                    const int64_t cc0 = ((int64_t)(p >> 1) * cny + (j >> 1)) * cnz + (k >> 1);
                    *reinterpret_cast<float2*>(rc + cc0) = make_float2(0.5f * (acc[0] + o0), 0.5f * (acc[1] + o1));
                }
                acc[0] = acc[1] = 0.0f;
            }
        }
        Rm = Rc; Rc = Rn; Rn = fold(Rf, Af); Rf = Rg; Af = Ag;
        Km = Kc; Kc = Kn; Kn = Kf; Kf = Kg;
        Z1m = Z1c; Z1c = Z1n; Z1n = z1of(Rn, Kn);
#pragma unroll
        for (int e = 0; e < 4; ++e) { Zmm[e] = Zm[e]; Zm[e] = Zc[e]; }
    }
    if (FOLD) {
        double v[1] = {rr};
        write_partials<1, NT>(v, rr_part);
    }
}

// ---------------------------------------------------------------------------
// Prolongation and post-smoothing of the fine level in one pass (nz % 4 == 0,
// 16-byte aligned vectors): zt = z_in + P ec on fluid cells (z_in elsewhere),
// then z_out = zt + w D^-1 (r - A zt) on fluid cells (0 elsewhere) with the
// partials of r.z_out and r.z_old.  A workgroup owns a tj x 56 (j,k) tile
// (tj even, at most SR - 2) and walks down an i segment; each of its SR x 16
// threads holds one 4-cell k chunk of the tile or of its one-cell ring.  At
// plane q it forms zt over the tile and ring (prolong4_kernel's float
// operations), at plane q-1 the sweep over the tile (smooth_dots_kernel's),
// so z_out is bit-identical to the two kernels'; only the dot products'
// summation order differs.  The coarse values come from LDS: each coarse
// plane's (tj/2 + 2) x 34 window under the tile is loaded once into a ring of
// three (a fine plane reads coarse planes q/2 and q/2 -/+ 1); zt planes are
// double-buffered.  zt is held at +0 on non-fluid cells (the sweep never
// reads it there), so the sweep's A is lap_mg with no neighbour codes.  On
// the fine level the tile is 44 x 56 on 768 threads, two workgroups per CU in
// one round (440x440x280: 500 workgroups of 44 planes, 162 us per pass,
// against 184 us for 7040 workgroups of 256 threads on 14 x 56 tiles; 172 us
// for two rounds of the taller tile, 197 us for 1.4 rounds, 188 / 168 us for
// one round of 62 x 56 on 1024 / 30 x 56 on 512 threads).
// ---------------------------------------------------------------------------
constexpr int kPpKT = 14;
constexpr int kPpEK = 2 * kPpKT + 6;  // coarse window columns
// Thread rows (tile rows + the 2 ring rows) of the two workgroup shapes: 256
// threads and kPpBigSR * 16 (the V-cycle's fine level).
constexpr int kPpSmallSR = 16, kPpBigSR = 48;
// coarse window rows under a tile of up to sr - 2 rows (an even number)
__host__ __device__ constexpr int pp_er(int sr) { return (sr - 2) / 2 + 2; }

struct PpShape {
    int jtiles, ktiles, ni, tj;
    unsigned grid;
};

// Rows spread evenly over the j tiles, an even number per tile (up to sr - 2);
// at most `target` workgroups of i segments.
static PpShape pp_shape(int nx, int ny, int nz, int sr, int target) {
    PpShape s;
    const int tmax = sr - 2;
    s.jtiles = (ny + tmax - 1) / tmax;
    s.tj = (ny + s.jtiles - 1) / s.jtiles;
    s.tj += s.tj & 1;
    s.ktiles = (nz + 4 * kPpKT - 1) / (4 * kPpKT);
    const int cols = s.jtiles * s.ktiles;
    int nseg = target / cols;
    nseg = nseg < 1 ? 1 : nseg;
    int ni = (nx + nseg - 1) / nseg;
    ni = ni < 4 ? 4 : ni;
    s.ni = ni;
    nseg = (nx + ni - 1) / ni;
    s.grid = (unsigned)(nseg * s.jtiles * s.ktiles);
    return s;
}

template <int SR>
__global__ __launch_bounds__(SR * (kPpKT + 2)) void prolong_post_kernel(int nx, int ny, int nz, int ni_seg,
                                                                        int jtiles, int ktiles, int tj,
                                                                        const uint8_t* __restrict__ code,
                                                           const float* __restrict__ z_in,
                                                           const float* __restrict__ ec, const float* __restrict__ r,
                                                           const float* __restrict__ z_old, float* __restrict__ z_out,
                                                           double* __restrict__ part, float w,
                                                           const int* __restrict__ done) {
    if (done && *done) return;
    constexpr int CH = kPpKT + 2, W = 4 * CH, NT = SR * CH;
    constexpr int EN = pp_er(SR) * kPpEK;
    __shared__ __align__(16) float sZ[2][SR * W + 8];
    __shared__ float sE[3][EN];
    const int tid = threadIdx.x;
    const int row = tid / CH, ch = tid - (tid / CH) * CH;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int cols = jtiles * ktiles;
    const int seg = (int)(b / cols), col = (int)(b % cols);
    const int jt = col / ktiles, kt = col - (col / ktiles) * ktiles;
    const int j0 = jt * tj, k0 = kt * 4 * kPpKT;
    const int i0 = seg * ni_seg, i1 = min(nx, i0 + ni_seg);
    const int j = j0 - 1 + row, k = k0 - 4 + 4 * ch;
    // rows past the tile's tj rows and its ring stay idle (air)
    const bool jk_in = row < tj + 2 && j >= 0 && j < ny && k >= 0 && k < nz;
    const bool inner = row >= 1 && row <= tj && ch >= 1 && ch <= kPpKT && j < ny && k < nz;
    const int cnx = (nx + 1) / 2, cny = (ny + 1) / 2, cnz = (nz + 1) / 2;
    const int Jb = j0 / 2 - 1, Kb = k0 / 2 - 3;  // coarse window origin
    const int64_t pl = (int64_t)ny * nz, cpl = (int64_t)cny * cnz;
    const int64_t off = (int64_t)j * nz + k;
    auto ldz = [&](int q) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(z_in + (int64_t)q * pl + off);
        return v;
    };
    auto ldk = [&](int q) {
        uint32_t v = 0;
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const uint32_t*>(code + (int64_t)q * pl + off);
        return v;
    };
    auto load_coarse = [&](int m) {  // coarse plane m's window into its ring slot
        if (m < 0 || m >= cnx) return;
        float* E = sE[m % 3];
        for (int t = tid; t < EN; t += NT) {
            const int a = t / kPpEK, c = t - a * kPpEK;
            const int J = Jb + a, K = Kb + c;
            E[t] = (J >= 0 && J < cny && K >= 0 && K < cnz) ? ec[(int64_t)m * cpl + (int64_t)J * cnz + K] : 0.0f;
        }
    };
    {
        const int m0 = (i0 - 1) >> 1;
        load_coarse(m0 - 1);
        load_coarse(m0);
        load_coarse(m0 + 1);
    }
    float4 Zc4 = ldz(i0 - 1), Zn4 = ldz(i0), Zf4 = ldz(i0 + 1);  // z_in at planes q, q+1, q+2
    uint32_t Km = 0, Kc = ldk(i0 - 1), Kn = ldk(i0), Kf = ldk(i0 + 1);
    // The thread's coarse rows J, J2 and its clamped 4-value k window (used
    // only where jk_in).  k is a multiple of 4, so fine cell k+e reads window
    // slots 1 + e/2 and {0, 2, 1, 3}[e]: where prolong4_kernel's clamped K2
    // differs from these (the chunks at k = 0 and k = nz - 4), the clamped
    // window holds the same value in both slots.
    const int Jr = j >> 1;
    int J2r = Jr + ((j & 1) ? 1 : -1);
    J2r = J2r < 0 ? 0 : (J2r >= cny ? cny - 1 : J2r);
    const int rJ = (Jr - Jb) * kPpEK, rJ2 = (J2r - Jb) * kPpEK;
    int cK[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int kk = (k >> 1) - 1 + u;
        kk = kk < 0 ? 0 : (kk >= cnz ? cnz - 1 : kk);
        cK[u] = kk - Kb;
    }
    float Tmm[4] = {0.f, 0.f, 0.f, 0.f}, Tm[4] = {0.f, 0.f, 0.f, 0.f};
    double da = 0.0, db = 0.0;
    for (int q = i0 - 1; q <= i1; ++q) {
        const int par = (q - i0 + 1) & 1;
        if (q & 1) load_coarse((q >> 1) + 1);
        const float4 Zg4 = ldz(q + 3);  // planes q .. q+2 are held: three planes of loads in flight
        const uint32_t Kg = ldk(q + 3);
        // the sweep's pointwise inputs at plane q - 1
        float rv[4] = {0.f, 0.f, 0.f, 0.f}, zo[4] = {0.f, 0.f, 0.f, 0.f};
        const int p = q - 1;
        if (inner && p >= i0) {
            ldf_nt<4>(r, (int64_t)p * pl + off, rv);
            if (z_old) ldf_nt<4>(z_old, (int64_t)p * pl + off, zo);
        }
        __syncthreads();
        // --- prolongation at plane q (tile and ring) ------------------------
        // zt on fluid cells; +0 elsewhere (the sweep never uses it there, so
        // a neighbour term is one select)
        float Tc[4] = {0.f, 0.f, 0.f, 0.f};
        if (jk_in && q >= 0 && q < nx && Kc != 0u) {  // Kc == 0: four air cells
            const int I = q >> 1;
            int I2 = I + ((q & 1) ? 1 : -1);
            I2 = I2 < 0 ? 0 : (I2 >= cnx ? cnx - 1 : I2);
            const float* EI = sE[I % 3];
            const float* EI2 = sE[I2 % 3];
            float cw[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float acc = 0.0f;
                acc += 0.5625f * EI[rJ + cK[u]];
                acc += 0.1875f * EI[rJ2 + cK[u]];
                acc += 0.1875f * EI2[rJ + cK[u]];
                acc += 0.0625f * EI2[rJ2 + cK[u]];
                cw[u] = acc;
            }
            const float zi[4] = {Zc4.x, Zc4.y, Zc4.z, Zc4.w};
            constexpr int sK[4] = {1, 1, 2, 2}, sK2[4] = {0, 2, 1, 3};
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (byte_of(Kc, e) >= kCodeFluid0) Tc[e] = zi[e] + (0.75f * cw[sK[e]] + 0.25f * cw[sK2[e]]);
        }
        float* VZ = sZ[par] + 4;
        *reinterpret_cast<float4*>(VZ + row * W + 4 * ch) = make_float4(Tc[0], Tc[1], Tc[2], Tc[3]);
        // --- post-sweep at plane p = q - 1 and the dot products ------------
        if (inner && p >= i0) {
            const float* PZ = sZ[par ^ 1] + 4;
            const float4 d4 = *reinterpret_cast<const float4*>(PZ + (row - 1) * W + 4 * ch);
            const float4 u4 = *reinterpret_cast<const float4*>(PZ + (row + 1) * W + 4 * ch);
            const float vb = PZ[row * W + 4 * ch - 1], vf = PZ[row * W + 4 * ch + 4];
            const float vd[4] = {d4.x, d4.y, d4.z, d4.w}, vu[4] = {u4.x, u4.y, u4.z, u4.w};
            float out[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint8_t cc = byte_of(Km, e);
                out[e] = 0.0f;
                if (cc >= kCodeFluid0) {
                    const float ps = Tm[e];
                    const float v_b = e > 0 ? Tm[e - 1] : vb;
                    const float v_f = e < 3 ? Tm[e + 1] : vf;
                    float sum = Tmm[e] + Tc[e];
                    sum = sum + vd[e];
                    sum = sum + vu[e];
                    sum = sum + v_b;
                    sum = sum + v_f;
                    out[e] = ps + winv(cc, w) * (rv[e] - lap_mg(cc, ps, sum));
                }
                da += (double)rv[e] * (double)out[e];
                if (z_old) db += (double)rv[e] * (double)zo[e];
            }
            *reinterpret_cast<float4*>(z_out + (int64_t)p * pl + off) = make_float4(out[0], out[1], out[2], out[3]);
        }
        Zc4 = Zn4; Zn4 = Zf4; Zf4 = Zg4;
        Km = Kc; Kc = Kn; Kn = Kf; Kf = Kg;
#pragma unroll
        for (int e = 0; e < 4; ++e) { Tmm[e] = Tm[e]; Tm[e] = Tc[e]; }
    }
    double v[2] = {da, db};
    write_partials<2, NT>(v, part);
}

// The fine level's taller tiles fill the chip in one round: as many
// workgroups as the CUs hold at once, each walking a long i segment (fewer,
// longer segments recompute fewer planes; a partial second round costs more
// than it hides).  The 256-thread shape keeps 30 per CU.
static int pp_target(int sr) {
    static const int occ = resident_wgs(prolong_post_kernel<kPpBigSR>, kPpBigSR * (kPpKT + 2));
    return (sr == kPpBigSR ? occ : 30) * device_cus();
}

// z_out = z_in + P ec from level LC to its refinement LF (tiles TF, vector
// width vf); z_out may be z_in.
// Sum over the workgroup of bottom_pcg_kernel (1024 threads, 16 waves): every
// thread gets the total; `slot` alternates so a reduction never overwrites
// partials another thread may still be reading.
__device__ __forceinline__ double bot_sum(double v, double* red, int slot) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[slot * 16 + (threadIdx.x >> 6)] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) s += red[slot * 16 + q];
    return s;
}

// Bottom solve of the deep coarse V-cycle: `its` Jacobi-PCG iterations on
// A e = r from e = 0 over a level of at most kBotMax cells, in one
// workgroup: the search direction and the codes live in LDS (the SpMV's
// neighbours), r, e and the iteration's Ap and z in registers (cells t,
// t + 1024, ... of thread t); dot products in FP64.  Stops early when p.Ap
// is no longer positive (r = 0).
__global__ __launch_bounds__(kBotThreads) void bottom_pcg_kernel(MLvl L, const uint8_t* __restrict__ code,
                                                                 const float* __restrict__ r_in,
                                                                 float* __restrict__ e_out, int its,
                                                                 const int* __restrict__ done) {
    if (done && *done) return;
    __shared__ float sp[kBotMax];
    __shared__ uint8_t sc[kBotMax];
    __shared__ double red[2 * 16];
    const int t = threadIdx.x;
    const int n = (int)L.n, nz = L.nz, pl = L.ny * L.nz;
    float r[kBotPer], e[kBotPer];
    double rz = 0.0;
#pragma unroll
    for (int q = 0; q < kBotPer; ++q) {
        const int c = t + q * kBotThreads;
        e[q] = 0.0f;
        r[q] = 0.0f;
        if (c < n) {
            const uint8_t cc = code[c];
            sc[c] = cc;
            r[q] = cc >= kCodeFluid0 ? r_in[c] : 0.0f;
            const float z = precond(r[q], cc);
            sp[c] = z;
            rz += (double)r[q] * (double)z;
        }
    }
    rz = bot_sum(rz, red, 1);  // its barrier also publishes sp and sc
    for (int it = 0; it < its; ++it) {
        float ap[kBotPer];
        double pap = 0.0;
#pragma unroll
        for (int q = 0; q < kBotPer; ++q) {
            const int c = t + q * kBotThreads;
            ap[q] = 0.0f;
            if (c >= n || sc[c] < kCodeFluid0) continue;
            const int i = c / pl, rem = c - i * pl, j = rem / nz, k = rem - j * nz;
            const float ps = sp[c];
            float s = (i > 0 ? nbr_term(sc[c - pl], sp[c - pl], ps) : 0.0f) +
                      (i < L.nx - 1 ? nbr_term(sc[c + pl], sp[c + pl], ps) : 0.0f);
            s = s + (j > 0 ? nbr_term(sc[c - nz], sp[c - nz], ps) : 0.0f);
            s = s + (j < L.ny - 1 ? nbr_term(sc[c + nz], sp[c + nz], ps) : 0.0f);
            s = s + (k > 0 ? nbr_term(sc[c - 1], sp[c - 1], ps) : 0.0f);
            s = s + (k < nz - 1 ? nbr_term(sc[c + 1], sp[c + 1], ps) : 0.0f);
            ap[q] = 6.0f * ps - s;
            pap += (double)ps * (double)ap[q];
        }
        pap = bot_sum(pap, red, 0);
        if (!(pap > 0.0)) break;  // the same decision in every thread
        const float al = (float)(rz / pap);
        float z[kBotPer];
        double rzn = 0.0;
#pragma unroll
        for (int q = 0; q < kBotPer; ++q) {
            const int c = t + q * kBotThreads;
            z[q] = 0.0f;
            if (c >= n) continue;
            e[q] = e[q] + al * sp[c];
            r[q] = r[q] - al * ap[q];
            z[q] = precond(r[q], sc[c]);
            rzn += (double)r[q] * (double)z[q];
        }
        rzn = bot_sum(rzn, red, 1);
        const float be = rz > 0.0 ? (float)(rzn / rz) : 0.0f;
        rz = rzn;
        // every thread has finished reading sp (the barrier above)
#pragma unroll
        for (int q = 0; q < kBotPer; ++q) {
            const int c = t + q * kBotThreads;
            if (c < n) sp[c] = z[q] + be * sp[c];
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kBotPer; ++q) {
        const int c = t + q * kBotThreads;
        if (c < n) e_out[c] = e[q];
    }
}

static int prolong_level(hipStream_t st, const Tiles& TF, int vf, const MLvl& LF, const MLvl& LC,
                         const Tiles& TC, const uint8_t* fcode, const float* ec, const float* z_in, float* z_out,
                         const int* done);

static inline size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

struct MgWs {
    MLvl L0, L1, L2, L3, L4;
    float* zt0;               // fine temporary
    uint8_t *code1, *code2;
    int *mat1, *mat2;
    float *r1, *z1, *zt1, *r2, *e2, *zt2, *rs2, *en2;
    uint8_t *code3, *code4;  // the two levels below the 4x one (coarse kind 4)
    int *mat3, *mat4;
    float *r3, *z3, *zt3, *r4, *e4;
    unsigned int* nfl;
    double* part;  // 2 x kStepWg partials of the optimal-step products
    void* unet;
    size_t bytes;
};

static MgWs carve_mg(void* base, int nx, int ny, int nz) {
    MgWs W;
    memset(&W, 0, sizeof(W));
    W.L0 = mk(nx, ny, nz);
    W.L1 = mk((nx + 1) / 2, (ny + 1) / 2, (nz + 1) / 2);
    W.L2 = mk((W.L1.nx + 1) / 2, (W.L1.ny + 1) / 2, (W.L1.nz + 1) / 2);
    W.L3 = mk((W.L2.nx + 1) / 2, (W.L2.ny + 1) / 2, (W.L2.nz + 1) / 2);
    W.L4 = mk((W.L3.nx + 1) / 2, (W.L3.ny + 1) / 2, (W.L3.nz + 1) / 2);
    size_t off = 0;
    auto take = [&](size_t b) -> char* {
        char* p = base ? (char*)base + off : nullptr;
        off += al(b);
        return p;
    };
    W.zt0 = (float*)take(W.L0.n * 4);
    W.code1 = (uint8_t*)take(W.L1.n);
    W.mat1 = (int*)take(W.L1.n * 4);
    W.r1 = (float*)take(W.L1.n * 4);
    W.z1 = (float*)take(W.L1.n * 4);
    W.zt1 = (float*)take(W.L1.n * 4);
    W.code2 = (uint8_t*)take(W.L2.n);
    W.mat2 = (int*)take(W.L2.n * 4);
    W.r2 = (float*)take(W.L2.n * 4);
    W.e2 = (float*)take(W.L2.n * 4);
    W.zt2 = (float*)take(W.L2.n * 4);
    W.rs2 = (float*)take(W.L2.n * 4);
    W.en2 = (float*)take(W.L2.n * 4);
    W.code3 = (uint8_t*)take(W.L3.n);
    W.mat3 = (int*)take(W.L3.n * 4);
    W.r3 = (float*)take(W.L3.n * 4);
    W.z3 = (float*)take(W.L3.n * 4);
    W.zt3 = (float*)take(W.L3.n * 4);
    W.code4 = (uint8_t*)take(W.L4.n);
    W.mat4 = (int*)take(W.L4.n * 4);
    W.r4 = (float*)take(W.L4.n * 4);
    W.e4 = (float*)take(W.L4.n * 4);
    W.nfl = (unsigned int*)take(64);
    W.part = (double*)take(2 * kStepWg * sizeof(double));
    W.unet = take(unet_ws_bytes(W.L2.nx, W.L2.ny, W.L2.nz));
    W.bytes = off;
    return W;
}

static unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

size_t mg_ws_bytes(int nx, int ny, int nz) { return carve_mg(nullptr, nx, ny, nz).bytes; }

static void optimal_step(hipStream_t st, const MLvl& L, const uint8_t* code, const float* r, float* e, double* part,
                         const int* done) {
    step_partials_kernel<<<kStepWg, 256, 0, st>>>(L, code, r, e, part, done);
    step_scale_kernel<<<grid_for(L.n), 256, 0, st>>>(L.n, part, kStepWg, e, done);
}

size_t step_ws_bytes() { return 2 * kStepWg * sizeof(double); }

void scale_to_optimal(hipStream_t st, int nx, int ny, int nz, const uint8_t* code, const float* r, float* e,
                      double* part, const int* done) {
    optimal_step(st, mk(nx, ny, nz), code, r, e, part, done);
}

__global__ void zero_u32(unsigned int* p) { if (threadIdx.x == 0) *p = 0u; }

int mg_setup(hipStream_t st, void* ws, int nx, int ny, int nz, const int* mat, int mf, int ma,
             const uint8_t* fcode_in, bool deep) {
    MgWs W = carve_mg(ws, nx, ny, nz);
    if (deep && W.L4.n > kBotMax) return (int)hipErrorInvalidValue;
    // The fine codes (given, or formed in a temporary: zt0 is free during
    // set-up), then the two coarse levels' classes and codes.
    const uint8_t* fcode = fcode_in;
    zero_u32<<<1, 64, 0, st>>>(W.nfl);
    if (!fcode) {
        codes_from_mat_kernel<<<grid_for(W.L0.n), 256, 0, st>>>(W.L0, mat, mf, ma, (uint8_t*)W.zt0, nullptr);
        fcode = (const uint8_t*)W.zt0;
    }
    if ((W.L0.nz & 3) == 0 && ((uintptr_t)fcode & 3) == 0)
        coarse_class2_kernel<<<grid_for(W.L1.n / 2), 256, 0, st>>>(W.L0, W.L1, fcode, W.mat1);
    else
        coarse_class_kernel<<<grid_for(W.L1.n), 256, 0, st>>>(W.L0, W.L1, fcode, W.mat1);
    codes_from_mat_kernel<<<grid_for(W.L1.n), 256, 0, st>>>(W.L1, W.mat1, 0, 2, W.code1, nullptr);
    coarse_class_kernel<<<grid_for(W.L2.n), 256, 0, st>>>(W.L1, W.L2, W.code1, W.mat2);
    // the fluid count's atomics: 256 workgroups, one each
    codes_from_mat_kernel<<<256, 256, 0, st>>>(W.L2, W.mat2, 0, 2, W.code2, W.nfl);
    if (deep) {
        coarse_class_kernel<<<grid_for(W.L3.n), 256, 0, st>>>(W.L2, W.L3, W.code2, W.mat3);
        codes_from_mat_kernel<<<grid_for(W.L3.n), 256, 0, st>>>(W.L3, W.mat3, 0, 2, W.code3, nullptr);
        coarse_class_kernel<<<grid_for(W.L4.n), 256, 0, st>>>(W.L3, W.L4, W.code3, W.mat4);
        codes_from_mat_kernel<<<grid_for(W.L4.n), 256, 0, st>>>(W.L4, W.mat4, 0, 2, W.code4, nullptr);
    }
    return (int)hipGetLastError();
}

// Partial count of the last pass: one per workgroup of the single-pass
// prolongation and post-sweep when nz % 4 == 0, else one per fine tile (the
// V-cycle's fine buffers all come from 256-byte aligned workspace, so the
// vector width depends on nz only).
unsigned mg_dots_groups(int nx, int ny, int nz) {
    return (nz & 3) == 0 ? pp_shape(nx, ny, nz, kPpBigSR, pp_target(kPpBigSR)).grid : make_tiles(nx, ny, nz, vec_for(nz, {})).ntiles;
}

#define MG_VEC(KERNEL, VEC, GRID, ST, ...)                                          \
    do {                                                                           \
        if ((VEC) == 4) KERNEL<4><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);          \
        else if ((VEC) == 2) KERNEL<2><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);     \
        else KERNEL<1><<<(GRID), 256, 0, (ST)>>>(__VA_ARGS__);                     \
    } while (0)

static int prolong_level(hipStream_t st, const Tiles& TF, int vf, const MLvl& LF, const MLvl& LC,
                         const Tiles& TC, const uint8_t* fcode, const float* ec, const float* z_in, float* z_out,
                         const int* done) {
    if (vf == 4) {
        const Tiles CV = make_tiles(LC.nx, LC.ny, LF.nz, 4);
        prolong4_kernel<<<CV.ntiles, 256, 0, st>>>(TF, CV, fcode, ec, z_in, z_out, done);
    } else {
        if (z_out != z_in)
            BS_RETURN_IF_ERR(hipMemcpyAsync(z_out, z_in, sizeof(float) * (size_t)LF.n, hipMemcpyDeviceToDevice, st));
        MG_VEC(prolong_vec_kernel, vf, TF.ntiles, st, TF, TC, fcode, ec, z_out, done);
    }
    return (int)hipGetLastError();
}

// Restriction of the residual r - A z from level F to its coarsening: the
// 2x2x4-block kernel when the fine vectors allow 4-wide access, the
// coarse-centred pair kernel for even nz, the one-cell kernel otherwise.
static void restrict_level(hipStream_t st, const Tiles& TF, int vf, const MLvl& LF, const MLvl& LC,
                           const Tiles& TC1, const uint8_t* fcode, const float* r, const float* z, float* rc,
                           const int* done) {
    if (vf == 4 && (LC.nz & 1) == 0) {
        const Tiles TC2 = make_tiles(LC.nx, LC.ny, LC.nz, 2);
        restrict4_kernel<<<TC2.ntiles, 256, 0, st>>>(TF, TC2, fcode, r, z, rc, done);
    } else if ((LF.nz & 1) == 0) {
        restrict2_kernel<<<TC1.ntiles, 256, 0, st>>>(TF, TC1, fcode, r, z, rc, done);
    } else {
        restrict_kernel<<<grid_for(LC.n), 256, 0, st>>>(LF, LC, fcode, r, z, rc, done);
    }
}

// Workgroups of the fine level's single pre-smoothing pass: at most 16 per
// CU, two rounds of the 8 each CU holds (440x440x280: 4070 workgroups of
// 20-plane segments, 258 us, against 6845 of 12 planes, 3.34 rounds, 268 us;
// 2035 of 40 planes 260 us, 8140 of 10 planes 271 us).
static int pr_target() {
    static const int occ = resident_wgs(pre2_restrict_kernel<true, kPrBigSR>, kPrBigSR * (kPrKT + 2));
    return occ * device_cus();
}

// Two smoothing sweeps from zero on level LF and the restricted residual on
// LC: one pass when the shape allows it (fused != 0), else pre2_kernel and
// the restriction kernels.
static int pre_restrict(hipStream_t st, const Tiles& TF, int vf, const MLvl& LF, const MLvl& LC, const Tiles& TC1,
                        const uint8_t* code, const float* r, float* z, float* rc, const int* done, int fused,
                        int target_wg, Pre2Coef K, const MgFold* fold = nullptr, bool big = false) {
    if (fused && pr_fits(LF, vf)) {
        const int sr = big ? kPrBigSR : kPrSmallSR;
        const PrShape S = pr_shape(LF.nx, LF.ny, LF.nz, target_wg, sr);
        const unsigned nt = (unsigned)sr * (kPrKT + 2);
        if (fold && big)
            pre2_restrict_kernel<true, kPrBigSR><<<S.grid, nt, 0, st>>>(
                LF.nx, LF.ny, LF.nz, S.ni, S.jtiles, S.ktiles, S.tj, code, r, z, rc, done, K, fold->ap, fold->alpha,
                fold->r_out, fold->rr_part);
        else if (fold)
            return (int)hipErrorInvalidValue;
        else if (big)
            pre2_restrict_kernel<false, kPrBigSR><<<S.grid, nt, 0, st>>>(
                LF.nx, LF.ny, LF.nz, S.ni, S.jtiles, S.ktiles, S.tj, code, r, z, rc, done, K, nullptr, nullptr,
                nullptr, nullptr);
        else
            pre2_restrict_kernel<false, kPrSmallSR><<<S.grid, nt, 0, st>>>(
                LF.nx, LF.ny, LF.nz, S.ni, S.jtiles, S.ktiles, S.tj, code, r, z, rc, done, K, nullptr, nullptr,
                nullptr, nullptr);
    } else if (fold) {
        return (int)hipErrorInvalidValue;
    } else {
        MG_VEC(pre2_kernel, vf, TF.ntiles, st, TF, code, r, z, done, K);
        restrict_level(st, TF, vf, LF, LC, TC1, code, r, z, rc, done);
    }
    return (int)hipGetLastError();
}

// gpart: x holds the network's correction before its A-norm-optimal scaling,
// whose step partials (gG per value) are given: the first two steps apply
// the scaling on load (gmode of cheb_kernel).  znf: x is exactly +0 on every
// non-fluid cell, as every iterate of the V-cycle is (its kernels store +0
// there), so the steps read no neighbour codes; the steps keep that
// property, so tmp need not have it.
static void chebyshev(hipStream_t st, const Tiles& T, int vec, const uint8_t* code, const float* r, float* x,
                      float* tmp, int steps, double lo, const int* done, const double* gpart, unsigned gG,
                      bool znf) {
    const double hi = 2.0;
    const double th = 0.5 * (hi + lo), sg = th / (0.5 * (hi - lo));
    const float ga = (float)(1.0 / th);
    double om = 1.0;
    float* cur = x;
    float* prv = tmp;
    for (int q = 0; q < steps; ++q) {
        if (q == 1) om = 1.0 / (1.0 - 1.0 / (2.0 * sg * sg));
        else if (q > 1) om = 1.0 / (1.0 - om / (4.0 * sg * sg));
        const float* xp = q == 0 ? cur : prv;  // step 0 has no previous iterate
        const int gm = gpart ? (q == 0 ? 1 : q == 1 ? 2 : 0) : 0;
        MG_VEC(cheb_kernel, vec, T.ntiles, st, T, code, r, cur, xp, prv, (float)om, ga, done, gpart, gG, gm,
               znf ? 1 : 0);
        float* t = cur;
        cur = prv;
        prv = t;
    }
}

unsigned mg_fold_groups(int nx, int ny, int nz) {
    const MLvl L = mk(nx, ny, nz);
    return pr_fits(L, 4) ? pr_shape(nx, ny, nz, pr_target(), kPrBigSR).grid : 0u;
}

int mg_apply(hipStream_t st, void* ws, int nx, int ny, int nz, const uint8_t* code0, const float* r, float* z,
             const float* z_old, double* dots_part, int coarse_kind, const float* params, const void* packed,
             const int* done, const MgFold* fold) {
    MgWs W = carve_mg(ws, nx, ny, nz);
    const int v0 = fold ? vec_for(W.L0.nz, {r, z, z_old, code0, W.zt0, fold->ap, fold->r_out})
                        : vec_for(W.L0.nz, {r, z, z_old, code0, W.zt0});
    const int v1 = vec_for(W.L1.nz, {W.r1, W.z1, W.zt1, W.code1});
    const Tiles T0 = make_tiles(W.L0.nx, W.L0.ny, W.L0.nz, v0);
    const Tiles T1 = make_tiles(W.L1.nx, W.L1.ny, W.L1.nz, v1);
    const Tiles C1 = make_tiles(W.L1.nx, W.L1.ny, W.L1.nz, 1);
    const Tiles C2 = make_tiles(W.L2.nx, W.L2.ny, W.L2.nz, 1);
    const unsigned g2 = grid_for(W.L2.n);
    // fine: two pre-smoothing sweeps from zero, restriction of the residual
    const Pre2Coef K0 = pre2_chebyshev(kPreLo), K1 = K0;
    // With the single-pass prolongation and post-sweep, the pre-smoothed
    // values go to the fine temporary (that pass reads them and writes z).
    const bool fused_post = v0 == 4;
    float* zpre = fused_post ? W.zt0 : z;
    if (fold) pass_mark(st, 1, 0);
    int rc = pre_restrict(st, T0, v0, W.L0, W.L1, C1, code0, r, zpre, W.r1, done, 1, pr_target(), K0, fold, true);
    if (rc != hipSuccess) return rc;
    if (fold) pass_mark(st, 1, 1);
    if (fold) {
        r = fold->r_out;  // the rest of the cycle works on the updated residual
        if (fold->after) fold->after(st, fold->after_arg);
    }
    // level 1: the same single pass on at most 4 workgroups per CU (220x220x140:
    // 912 of 14 planes; the learned projection 18.92 ms against 19.09 with
    // pre2_kernel + restrict4_kernel, 19.05 with 456 workgroups, 18.95 with
    // 1596; the restriction sums in another order, the iterations are the same)
    rc = pre_restrict(st, T1, v1, W.L1, W.L2, C2, W.code1, W.r1, W.z1, W.r2, done, 1, 4 * device_cus(), K1);
    if (rc != hipSuccess) return rc;
    // level 2: coarse solve.  coarse_kind 0: one damped-Jacobi sweep from
    // zero, then the polish; 1: the network (A-norm-optimal step), then the
    // polish; 2: 5 sweeps, the network on their residual, then 4 sweeps.
    // Coarsest-level sweeps: the vectorised one-tile smoother, in ping-pong
    // pairs ending in e2.
    const int v2 = vec_for(W.L2.nz, {W.r2, W.e2, W.zt2, W.code2});
    const Tiles T2 = make_tiles(W.L2.nx, W.L2.ny, W.L2.nz, v2);
    auto sweeps = [&](int pairs) {
        for (int q = 0; q < pairs; ++q) {
            MG_VEC(smooth_kernel, v2, T2.ntiles, st, T2, W.code2, W.r2, W.e2, W.zt2, done);
            MG_VEC(smooth_kernel, v2, T2.ntiles, st, T2, W.code2, W.r2, W.zt2, W.e2, done);
        }
    };
    static_assert(kChebSteps % 2 == 0 && kL1ChebSteps % 2 == 0, "Chebyshev polishes end in their input");
    auto polish = [&]() { chebyshev(st, T2, v2, W.code2, W.r2, W.e2, W.zt2, kChebSteps, kChebLo, done); };
    if (coarse_kind == 0) {
        jacobi0_kernel<<<g2, 256, 0, st>>>(W.L2, W.code2, W.r2, W.e2, done);
        polish();
    } else if (coarse_kind == 3) {
        jacobi0_kernel<<<g2, 256, 0, st>>>(W.L2, W.code2, W.r2, W.e2, done);
        chebyshev(st, T2, v2, W.code2, W.r2, W.e2, W.zt2, kAblChebSteps, kAblChebLo, done);
    } else if (coarse_kind == 4) {
        if (W.L4.n > kBotMax) return (int)hipErrorInvalidValue;
        const int v3 = vec_for(W.L3.nz, {W.r3, W.z3, W.zt3, W.code3});
        const Tiles T3 = make_tiles(W.L3.nx, W.L3.ny, W.L3.nz, v3);
        const Tiles C3 = make_tiles(W.L3.nx, W.L3.ny, W.L3.nz, 1);
        const Tiles C4 = make_tiles(W.L4.nx, W.L4.ny, W.L4.nz, 1);
        // down: 4x -> 8x -> 16x, two-step Chebyshev pre-smoothing on each
        rc = pre_restrict(st, T2, v2, W.L2, W.L3, C3, W.code2, W.r2, W.e2, W.r3, done, 0, 0, K1);
        if (rc != hipSuccess) return rc;
        rc = pre_restrict(st, T3, v3, W.L3, W.L4, C4, W.code3, W.r3, W.z3, W.r4, done, 0, 0, K1);
        if (rc != hipSuccess) return rc;
        bottom_pcg_kernel<<<1, kBotThreads, 0, st>>>(W.L4, W.code4, W.r4, W.e4, kBotIters, done);
        // up: prolongation and Chebyshev post-smoothing on the 8x level, then
        // the 4x level's prolongation and the same polish as the network path
        rc = prolong_level(st, T3, v3, W.L3, W.L4, C4, W.code3, W.e4, W.z3, W.z3, done);
        if (rc != hipSuccess) return rc;
        chebyshev(st, T3, v3, W.code3, W.r3, W.z3, W.zt3, kDeepPostSteps, kL1ChebLo, done);
        rc = prolong_level(st, T2, v2, W.L2, W.L3, C3, W.code2, W.z3, W.e2, W.e2, done);
        if (rc != hipSuccess) return rc;
        polish();
    } else if (coarse_kind == 1) {
        rc = unet_apply(st, W.unet, W.L2.nx, W.L2.ny, W.L2.nz, params, packed, W.r2, W.code2, W.nfl,
                        nullptr, 0, W.e2, done);
        if (rc != hipSuccess) return rc;
        // the A-norm-optimal step's products; the polish's first two steps
        // apply the scaling (optimal_step's second pass folded in)
        step_partials_kernel<<<kStepWg, 256, 0, st>>>(W.L2, W.code2, W.r2, W.e2, W.part, done);
        chebyshev(st, T2, v2, W.code2, W.r2, W.e2, W.zt2, kChebSteps, kChebLo, done, W.part, kStepWg);
    } else {
        jacobi0_kernel<<<g2, 256, 0, st>>>(W.L2, W.code2, W.r2, W.e2, done);
        jacobi_kernel<<<g2, 256, 0, st>>>(W.L2, W.code2, W.r2, W.e2, W.zt2, done);
        jacobi_kernel<<<g2, 256, 0, st>>>(W.L2, W.code2, W.r2, W.zt2, W.e2, done);
        residual_kernel<<<g2, 256, 0, st>>>(W.L2, W.code2, W.r2, W.e2, W.rs2, done);
        rc = unet_apply(st, W.unet, W.L2.nx, W.L2.ny, W.L2.nz, params, packed, W.rs2, W.code2, W.nfl,
                        nullptr, 0, W.en2, done);
        if (rc != hipSuccess) return rc;
        optimal_step(st, W.L2, W.code2, W.rs2, W.en2, W.part, done);
        add_kernel<<<g2, 256, 0, st>>>(W.L2.n, W.en2, W.e2, done);
        sweeps(2);
    }
    // back up: level 1
    rc = prolong_level(st, T1, v1, W.L1, W.L2, C2, W.code1, W.e2, W.z1, W.z1, done);
    if (rc != hipSuccess) return rc;
    chebyshev(st, T1, v1, W.code1, W.r1, W.z1, W.zt1, kL1ChebSteps, kL1ChebLo, done);
    // fine: prolongation, then one sweep fused with r.z, r.z_old
    if (fused_post) {
        const PpShape P = pp_shape(nx, ny, nz, kPpBigSR, pp_target(kPpBigSR));
        if (P.grid != mg_dots_groups(nx, ny, nz)) return (int)hipErrorInvalidValue;
        pass_mark(st, 2, 0);
        prolong_post_kernel<kPpBigSR><<<P.grid, kPpBigSR * (kPpKT + 2), 0, st>>>(
            nx, ny, nz, P.ni, P.jtiles, P.ktiles, P.tj, code0, W.zt0, W.z1, r, z_old, z, dots_part, kPostW, done);
        pass_mark(st, 2, 1);
        return (int)hipGetLastError();
    }
    rc = prolong_level(st, T0, v0, W.L0, W.L1, C1, code0, W.z1, z, W.zt0, done);
    if (rc != hipSuccess) return rc;
    if (T0.ntiles != mg_dots_groups(nx, ny, nz)) return (int)hipErrorInvalidValue;  // misaligned caller buffers
    MG_VEC(smooth_dots_kernel, v0, T0.ntiles, st, T0, code0, r, W.zt0, z, z_old, dots_part, done, kPostW);
    return (int)hipGetLastError();
}

}  // namespace bs

// The V-cycle's first pass on one level (diagnostic entry, no counterpart in
// the reference): z = two smoothing steps from zero on A z = r — damped
// Jacobi (w = 0.8) for lo <= 0, the two-step Chebyshev polynomial over
// [lo, 2] otherwise — and rc[I,J,K] = 0.5 x the sum of r - A z over the 8
// children of coarse cell (I,J,K) (coarse dims (n+1)/2).  fused = 1 runs the
// single-pass kernel on 256-thread tiles and fused = 2 on the fine level's
// taller tiles; both are an error unless nx, ny are even, nz % 4 == 0 and
// the vectors are 16-byte aligned; fused = 0 runs the two-kernel path.
// The V-cycle's last pass on the fine level (diagnostic entry, no counterpart
// in the reference): zt = z_in + P ec on fluid cells (cell-centred trilinear
// prolongation from the coarse grid ((n+1)/2 per axis); z_in elsewhere), then
// z_out = zt + w D^-1 (r - A zt) on fluid cells (0 elsewhere), and per
// workgroup the partials of r.z_out (part[0..G)) and r.z_old (part[G..2G),
// 0 if z_old is null); *groups = G.  fused = 1 runs the single-pass kernel
// on 256-thread tiles, fused = 2 on the fine level's taller tiles (both need
// nz % 4 == 0 and 16-byte aligned vectors, otherwise an error), fused = 0
// prolong + smooth_dots through tmp (nx*ny*nz floats).  z_out must not alias
// z_in.
extern "C" int bs_mg_prolong_post(const unsigned char* code, const float* z_in, const float* ec, const float* r,
                                  const float* z_old, float* z_out, float* tmp, double* part, size_t part_len,
                                  int* groups, int nx, int ny, int nz, int fused, double w, bs_stream_t stream) {
    using namespace bs;
    if (!groups) return (int)hipErrorInvalidValue;
    *groups = 0;
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    hipStream_t st = (hipStream_t)stream;
    const MLvl LF = mk(nx, ny, nz);
    const MLvl LC = mk((nx + 1) / 2, (ny + 1) / 2, (nz + 1) / 2);
    const int vf = vec_for(nz, {z_in, r, z_old, z_out, tmp, code});
    if (fused) {
        if (vf != 4) return (int)hipErrorInvalidValue;
        const int sr = fused == 2 ? kPpBigSR : kPpSmallSR;
        const PpShape P = pp_shape(nx, ny, nz, sr, pp_target(sr));
        if (part_len < 2 * (size_t)P.grid) return (int)hipErrorInvalidValue;
        *groups = (int)P.grid;
        if (fused == 2)
            prolong_post_kernel<kPpBigSR><<<P.grid, sr * (kPpKT + 2), 0, st>>>(
                nx, ny, nz, P.ni, P.jtiles, P.ktiles, P.tj, code, z_in, ec, r, z_old, z_out, part, (float)w, nullptr);
        else
            prolong_post_kernel<kPpSmallSR><<<P.grid, sr * (kPpKT + 2), 0, st>>>(
                nx, ny, nz, P.ni, P.jtiles, P.ktiles, P.tj, code, z_in, ec, r, z_old, z_out, part, (float)w, nullptr);
        return (int)hipGetLastError();
    }
    const Tiles TF = make_tiles(nx, ny, nz, vf);
    const Tiles TC = make_tiles(LC.nx, LC.ny, LC.nz, 1);
    if (part_len < 2 * (size_t)TF.ntiles) return (int)hipErrorInvalidValue;
    *groups = (int)TF.ntiles;
    const int rc = prolong_level(st, TF, vf, LF, LC, TC, code, ec, z_in, tmp, nullptr);
    if (rc != hipSuccess) return rc;
    MG_VEC(smooth_dots_kernel, vf, TF.ntiles, st, TF, code, r, tmp, z_out, z_old, part, nullptr, (float)w);
    return (int)hipGetLastError();
}

extern "C" int bs_mg_smooth_restrict(const unsigned char* code, const float* r, float* z, float* rc, int nx, int ny,
                                     int nz, int fused, double lo, bs_stream_t stream) {
    using namespace bs;
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    const MLvl LF = mk(nx, ny, nz);
    const MLvl LC = mk((nx + 1) / 2, (ny + 1) / 2, (nz + 1) / 2);
    const int vf = vec_for(nz, {r, z, code});
    if (fused && !pr_fits(LF, vf)) return (int)hipErrorInvalidValue;
    const Tiles TF = make_tiles(nx, ny, nz, vf);
    const Tiles TC1 = make_tiles(LC.nx, LC.ny, LC.nz, 1);
    if (lo >= 2.0) return (int)hipErrorInvalidValue;
    return pre_restrict((hipStream_t)stream, TF, vf, LF, LC, TC1, code, r, z, rc, nullptr, fused,
                        fused == 2 ? pr_target() : 4096, lo > 0.0 ? pre2_chebyshev(lo) : Pre2Coef{kOmega, 1.0f, kOmega},
                        nullptr, fused == 2);
}

// `steps` Chebyshev steps (even) over [lo, 2] of D^-1 A on A x = r from x,
// the result in x (tmp: device scratch of the same size) — the V-cycle's
// post-smoother and coarsest-level polish (diagnostic entry, no counterpart
// in the reference).
extern "C" int bs_mg_chebyshev(const unsigned char* code, const float* r, float* x, float* tmp, int nx, int ny,
                               int nz, int steps, double lo, bs_stream_t stream) {
    using namespace bs;
    if (nx <= 0 || ny <= 0 || nz <= 0 || steps == 0) return hipSuccess;
    if (steps < 0 || (steps & 1) || !(lo > 0.0 && lo < 2.0)) return (int)hipErrorInvalidValue;
    const int v = vec_for(nz, {r, x, tmp, code});
    const Tiles T = make_tiles(nx, ny, nz, v);
    chebyshev((hipStream_t)stream, T, v, code, r, x, tmp, steps, lo, nullptr, nullptr, 0, false);
    return (int)hipGetLastError();
}
