// Masked 7-point Jacobi pressure sweep on gfx950 and its fused n_iter driver.
//
// Arithmetic follows jacobi_pressure_step (python/boilingsim/fluid.py:243-276)
// exactly: s = ((((L+R)+D)+U)+B)+F over the neighbours i-1, i+1, j-1, j+1,
// k-1, k+1; rhs = (((rho*dx)*dx)*div)/dt; p_new = (s - rhs)/6, each operation
// rounded separately (the library is built with -ffp-contract=off and f32
// division is IEEE correctly rounded), so a sweep is bit-identical to the
// reference arithmetic.  Out-of-bounds neighbours are air, as in
// crates/cuda-kernels/src/jacobi_pressure.cu:121-124.
//
// Data movement.  One thread owns VEC consecutive k cells (dwordx4 loads when
// nz % 4 == 0), one workgroup of 256 threads owns a contiguous chunk of one
// (j,k) plane, and blockIdx.y walks i.  Neighbour rows/planes are re-read
// through L1/L2; the XCD remap keeps each XCD on a contiguous slab of planes
// so those re-reads hit its own L2.  HBM then sees each array once per sweep.
//
// Mask-free sweeps.  Inside the fused driver, sweeps after the first carry the
// class of each non-fluid cell in the pressure array itself (air and solid
// sentinels, see common.hpp), so those sweeps stream only p_in, div and p_out:
// 12 B/cell instead of the 16 B/cell of a sweep that also reads the int32
// mask.  The first sweep reads the mask and writes the encoded form; the last
// sweep writes the decoded form (0 at every non-fluid cell, as the reference
// does), and the other ping-pong buffer is decoded in place, so both caller
// buffers end exactly as the reference leaves them.
#include "boilsim_hip.h"
#include "common.hpp"

#include <stdlib.h>

namespace bs {

template <int VEC> struct VecF;
template <> struct VecF<1> { using T = float; };
template <> struct VecF<2> { using T = float2; };
template <> struct VecF<4> { using T = float4; };
template <int VEC> struct VecI;
template <> struct VecI<1> { using T = int; };
template <> struct VecI<2> { using T = int2; };
template <> struct VecI<4> { using T = int4; };

template <int VEC>
__device__ __forceinline__ void load_vec(const float* __restrict__ p, int64_t off, float (&v)[VEC]) {
    using T = typename VecF<VEC>::T;
    const T t = *reinterpret_cast<const T*>(p + off);
    const float* tv = reinterpret_cast<const float*>(&t);
#pragma unroll
    for (int e = 0; e < VEC; ++e) v[e] = tv[e];
}

template <int VEC>
__device__ __forceinline__ void load_vec(const int* __restrict__ p, int64_t off, int (&v)[VEC]) {
    using T = typename VecI<VEC>::T;
    const T t = *reinterpret_cast<const T*>(p + off);
    const int* tv = reinterpret_cast<const int*>(&t);
#pragma unroll
    for (int e = 0; e < VEC; ++e) v[e] = tv[e];
}

template <int VEC>
__device__ __forceinline__ void store_vec(float* __restrict__ p, int64_t off, const float (&v)[VEC]) {
    using T = typename VecF<VEC>::T;
    T t;
    float* tv = reinterpret_cast<float*>(&t);
#pragma unroll
    for (int e = 0; e < VEC; ++e) tv[e] = v[e];
    *reinterpret_cast<T*>(p + off) = t;
}

enum : int { kInMask = 0, kInEncoded = 1, kInZero = 2 };
enum : int { kOutDecoded = 0, kOutEncoded = 1 };

// Neighbour contribution in encoded form: air sentinel -> 0, solid sentinel ->
// p_self (Neumann ghost), anything else is a fluid value.
__device__ __forceinline__ float contrib_enc(float v, float ps) {
    const uint32_t b = __float_as_uint(v);
    float r = (b == kSentSolid) ? ps : v;
    return (b == kSentAir) ? 0.0f : r;
}

// Neighbour contribution from the class id (fluid.py:_pressure_neighbour).
__device__ __forceinline__ float contrib_mat(int m, float v, float ps, int mf, int ma) {
    if (m == mf) return v;
    if (m == ma) return 0.0f;
    return ps;
}

template <int VEC, int IN, int OUT>
__global__ __launch_bounds__(256) void jacobi_sweep_kernel(
    float* __restrict__ p_new, const float* __restrict__ p_old, const float* __restrict__ div,
    const int* __restrict__ mat, int nx, int ny, int nz, FastDiv nzv_div, uint32_t chunks,
    float c2, float dt, int mf, int ma) {
    const uint32_t nblk = gridDim.x * gridDim.y;
    const uint32_t lb = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nblk);
    const uint32_t chunk = lb % chunks;
    const int i = (int)(lb / chunks);
    const uint32_t nzv = nzv_div.d;
    const uint32_t q = chunk * 256u + threadIdx.x;
    if (q >= (uint32_t)ny * nzv) return;
    const int j = (int)fdiv(q, nzv_div);
    const int k0 = (int)(q - (uint32_t)j * nzv) * VEC;
    const int64_t plane = (int64_t)ny * nz;
    const int64_t c0 = (int64_t)i * plane + (int64_t)j * nz + k0;

    float pc[VEC], dv[VEC], out[VEC];
    if constexpr (IN != kInZero) load_vec<VEC>(p_old, c0, pc);
    load_vec<VEC>(div, c0, dv);

    if constexpr (IN == kInEncoded) {
        const float air = __uint_as_float(kSentAir);
        float pl[VEC], pr[VEC], pd[VEC], pu[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) { pl[e] = air; pr[e] = air; pd[e] = air; pu[e] = air; }
        if (i > 0) load_vec<VEC>(p_old, c0 - plane, pl);
        if (i < nx - 1) load_vec<VEC>(p_old, c0 + plane, pr);
        if (j > 0) load_vec<VEC>(p_old, c0 - nz, pd);
        if (j < ny - 1) load_vec<VEC>(p_old, c0 + nz, pu);
        const float pb_edge = (k0 > 0) ? p_old[c0 - 1] : air;
        const float pf_edge = (k0 + VEC < nz) ? p_old[c0 + VEC] : air;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const uint32_t cb = __float_as_uint(pc[e]);
            if (cb == kSentAir || cb == kSentSolid) {
                out[e] = (OUT == kOutEncoded) ? pc[e] : 0.0f;
                continue;
            }
            const float ps = pc[e];
            const float vb = (e > 0) ? pc[e - 1] : pb_edge;
            const float vf = (e < VEC - 1) ? pc[e + 1] : pf_edge;
            float s = contrib_enc(pl[e], ps) + contrib_enc(pr[e], ps);
            s = s + contrib_enc(pd[e], ps);
            s = s + contrib_enc(pu[e], ps);
            s = s + contrib_enc(vb, ps);
            s = s + contrib_enc(vf, ps);
            const float rhs = (c2 * dv[e]) / dt;
            out[e] = (s - rhs) / 6.0f;
        }
    } else if constexpr (IN == kInZero) {
        // p_old == 0 everywhere (the projection's cold start, fluid.py:627-629):
        // every neighbour term is +0, so s = +0 and only the class is needed.
        int mc[VEC];
        load_vec<VEC>(mat, c0, mc);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (mc[e] != mf) {
                if (OUT == kOutEncoded)
                    out[e] = __uint_as_float(mc[e] == ma ? kSentAir : kSentSolid);
                else
                    out[e] = 0.0f;
                continue;
            }
            const float rhs = (c2 * dv[e]) / dt;
            // s - rhs with s = +0: -rhs, except +0 (not -0) when rhs is a
            // zero.  Written as a select because the gfx950 instruction
            // selector folds (0.0f - x) into a negate modifier on the
            // division's operand, which turns 0 - (+0) into -0.
            const float num = (rhs == 0.0f) ? 0.0f : -rhs;
            out[e] = num / 6.0f;
        }
    } else {
        int mc[VEC], ml[VEC], mr[VEC], md[VEC], mu[VEC];
        float pl[VEC], pr[VEC], pd[VEC], pu[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            ml[e] = ma; mr[e] = ma; md[e] = ma; mu[e] = ma;
            pl[e] = 0.0f; pr[e] = 0.0f; pd[e] = 0.0f; pu[e] = 0.0f;
        }
        load_vec<VEC>(mat, c0, mc);
        if (i > 0) { load_vec<VEC>(p_old, c0 - plane, pl); load_vec<VEC>(mat, c0 - plane, ml); }
        if (i < nx - 1) { load_vec<VEC>(p_old, c0 + plane, pr); load_vec<VEC>(mat, c0 + plane, mr); }
        if (j > 0) { load_vec<VEC>(p_old, c0 - nz, pd); load_vec<VEC>(mat, c0 - nz, md); }
        if (j < ny - 1) { load_vec<VEC>(p_old, c0 + nz, pu); load_vec<VEC>(mat, c0 + nz, mu); }
        float pb_edge = 0.0f, pf_edge = 0.0f;
        int mb_edge = ma, mf_edge = ma;
        if (k0 > 0) { pb_edge = p_old[c0 - 1]; mb_edge = mat[c0 - 1]; }
        if (k0 + VEC < nz) { pf_edge = p_old[c0 + VEC]; mf_edge = mat[c0 + VEC]; }
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (mc[e] != mf) {
                if (OUT == kOutEncoded)
                    out[e] = __uint_as_float(mc[e] == ma ? kSentAir : kSentSolid);
                else
                    out[e] = 0.0f;
                continue;
            }
            const float ps = pc[e];
            const float vb = (e > 0) ? pc[e - 1] : pb_edge;
            const int mbv = (e > 0) ? mc[e - 1] : mb_edge;
            const float vf = (e < VEC - 1) ? pc[e + 1] : pf_edge;
            const int mfv = (e < VEC - 1) ? mc[e + 1] : mf_edge;
            float s = contrib_mat(ml[e], pl[e], ps, mf, ma) + contrib_mat(mr[e], pr[e], ps, mf, ma);
            s = s + contrib_mat(md[e], pd[e], ps, mf, ma);
            s = s + contrib_mat(mu[e], pu[e], ps, mf, ma);
            s = s + contrib_mat(mbv, vb, ps, mf, ma);
            s = s + contrib_mat(mfv, vf, ps, mf, ma);
            const float rhs = (c2 * dv[e]) / dt;
            out[e] = (s - rhs) / 6.0f;
        }
    }
    store_vec<VEC>(p_new, c0, out);
}

// ---------------------------------------------------------------------------
// Two sweeps per pass (temporal blocking).  A workgroup owns a (j,k) tile of
// TJ rows x 4*KT cells and walks down a range of i planes; each thread owns one
// 4-cell chunk of the tile plus its one-cell halo ring ((TJ+2) x (KT+2) chunks
// = 256 threads).  Per plane q it computes the first sweep at q over the tile
// and halo, and the second sweep at q-1 over the tile: i-neighbours come from
// the thread's own registers (planes rolled down i), j/k neighbours from two
// double-buffered LDS planes (p and first-sweep values), so one barrier per
// plane suffices.  HBM sees p_in and div once and p_out once for two sweeps.
// Both sweeps use exactly the single sweep's float operations: the neighbour
// sum in the same order (a plain sum when no neighbour carries a sentinel — a
// NaN pattern would make it NaN — the sentinel-aware form otherwise), rhs =
// (c2*div)/dt, and /6 as fma(x, zh, x*zl), which equals IEEE x/6 for every
// finite |x| >= 2^-96 (checked over all floats on the CPU and on gfx950); other
// x take the IEEE divide.  Requires nz % 4 == 0 and 16-byte aligned arrays.
// ---------------------------------------------------------------------------
constexpr int kPairKT = 14, kPairTJ = 14;

__device__ __forceinline__ bool is_sent(float v) {
    const uint32_t b = __float_as_uint(v);
    return b == kSentAir || b == kSentSolid;
}

__device__ __forceinline__ float div6(float x) {
    const float zh = 0x1.555556p-3f, zl = -0x1.555556p-28f;  // RN(1/6), RN(1/6 - zh)
    const float ax = fabsf(x);
    if (ax >= 0x1p-96f && ax <= 0x1.fffffep127f) return fmaf(x, zh, x * zl);
    return x / 6.0f;
}

__device__ __forceinline__ float sweep_enc(float ps, float L, float R, float D, float U, float B, float F,
                                           float rhs) {
    const float s = ((((L + R) + D) + U) + B) + F;
    if (!__builtin_isnan(s) && !__builtin_isnan(ps)) return div6(s - rhs);
    if (is_sent(ps)) return ps;
    float t = contrib_enc(L, ps) + contrib_enc(R, ps);
    t = t + contrib_enc(D, ps);
    t = t + contrib_enc(U, ps);
    t = t + contrib_enc(B, ps);
    t = t + contrib_enc(F, ps);
    return div6(t - rhs);
}

__global__ __launch_bounds__(256) void jacobi_pair_kernel(float* __restrict__ p_new, const float* __restrict__ p_old,
                                                          const float* __restrict__ div, int nx, int ny, int nz,
                                                          int ni_seg, int jtiles, int ktiles, float c2, float dt) {
    constexpr int KT = kPairKT, TJ = kPairTJ;
    constexpr int CH = KT + 2, SR = TJ + 2, PR = TJ + 4, W = 4 * CH, TK = 4 * KT;
    static_assert(CH * SR <= 256, "tile and halo must fit one workgroup");
    // +4 floats before and after each plane: the k-edge reads of the halo
    // chunks at the row ends land there (their results are never used).
    __shared__ __align__(16) float sP[2][PR * W + 8];
    __shared__ __align__(16) float sS[2][SR * W + 8];
    const int tid = threadIdx.x;
    const bool act = tid < CH * SR;
    const int r = tid / CH, c = tid - (tid / CH) * CH;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const int cols = jtiles * ktiles;
    const int seg = (int)(b / cols), col = (int)(b % cols);
    const int jt = col / ktiles, kt = col - (col / ktiles) * ktiles;
    const int j0 = jt * TJ, k0 = kt * TK;
    const int i0 = seg * ni_seg, i1 = min(nx, i0 + ni_seg);
    const int j = j0 - 1 + r, k = k0 - 4 + 4 * c;
    const bool jk_in = act && j >= 0 && j < ny && k >= 0 && k < nz;
    // The p plane also needs rows j0-2 and j0+TJ+1: the first and last thread
    // rows load them.
    const int jx = (r == 0) ? j0 - 2 : j0 + TJ + 1;
    const bool xrow = act && (r == 0 || r == SR - 1);
    const bool jx_in = xrow && jx >= 0 && jx < ny && k >= 0 && k < nz;
    const bool interior = act && r >= 1 && r <= TJ && c >= 1 && c <= KT && j < ny && k < nz;
    const int64_t pl = (int64_t)ny * nz;
    const int64_t off = (int64_t)j * nz + k, offx = (int64_t)jx * nz + k;
    const float air = __uint_as_float(kSentAir);
    // Out-of-domain positions read as air (value and class).
    auto ldp = [=](int q) {
        float4 v = make_float4(air, air, air, air);
        if (jk_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(p_old + (int64_t)q * pl + off);
        return v;
    };
    auto ldx = [=](int q) {
        float4 v = make_float4(air, air, air, air);
        if (jx_in && q >= 0 && q < nx) v = *reinterpret_cast<const float4*>(p_old + (int64_t)q * pl + offx);
        return v;
    };
    auto ldrhs = [=](int q) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (jk_in && q >= 0 && q < nx) {
            v = *reinterpret_cast<const float4*>(div + (int64_t)q * pl + off);
            v = make_float4((c2 * v.x) / dt, (c2 * v.y) / dt, (c2 * v.z) / dt, (c2 * v.w) / dt);
        }
        return v;
    };
    float4 Pm = ldp(i0 - 2), Pc = ldp(i0 - 1), Pn = ldp(i0), Xr = ldx(i0 - 1);
    float4 Hc = ldrhs(i0 - 1), Hm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 AIR4 = make_float4(air, air, air, air);
    float4 S1m = AIR4, S1c = AIR4;
    for (int q = i0 - 1; q <= i1; ++q) {
        float* P = sP[q & 1] + 4;
        float* S = sS[(q - 1) & 1] + 4;
        if (act) {
            *reinterpret_cast<float4*>(P + (r + 1) * W + 4 * c) = Pc;
            if (xrow) *reinterpret_cast<float4*>(P + (r == 0 ? 0 : PR - 1) * W + 4 * c) = Xr;
            *reinterpret_cast<float4*>(S + r * W + 4 * c) = S1c;
        }
        // Next plane's loads are in flight during this plane's arithmetic.
        const float4 Pf = ldp(q + 2), Xf = ldx(q + 1);
        const float4 Hf = ldrhs(q + 1);
        __syncthreads();
        // First sweep at plane q (tile and halo).
        float4 S1n = AIR4;
        if (jk_in && q >= 0 && q < nx) {
            const float4 d4 = *reinterpret_cast<const float4*>(P + r * W + 4 * c);
            const float4 u4 = *reinterpret_cast<const float4*>(P + (r + 2) * W + 4 * c);
            const float bk = P[(r + 1) * W + 4 * c - 1], fk = P[(r + 1) * W + 4 * c + 4];
            const float cc[4] = {Pc.x, Pc.y, Pc.z, Pc.w}, ll[4] = {Pm.x, Pm.y, Pm.z, Pm.w};
            const float rr[4] = {Pn.x, Pn.y, Pn.z, Pn.w}, dn[4] = {d4.x, d4.y, d4.z, d4.w};
            const float up[4] = {u4.x, u4.y, u4.z, u4.w}, hh[4] = {Hc.x, Hc.y, Hc.z, Hc.w};
            float o[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                o[e] = sweep_enc(cc[e], ll[e], rr[e], dn[e], up[e], e > 0 ? cc[e - 1] : bk, e < 3 ? cc[e + 1] : fk,
                                 hh[e]);
            S1n = make_float4(o[0], o[1], o[2], o[3]);
        }
        // Second sweep at plane q-1 (tile only).
        if (q - 1 >= i0 && interior) {
            const float4 d4 = *reinterpret_cast<const float4*>(S + (r - 1) * W + 4 * c);
            const float4 u4 = *reinterpret_cast<const float4*>(S + (r + 1) * W + 4 * c);
            const float bk = S[r * W + 4 * c - 1], fk = S[r * W + 4 * c + 4];
            const float cc[4] = {S1c.x, S1c.y, S1c.z, S1c.w}, ll[4] = {S1m.x, S1m.y, S1m.z, S1m.w};
            const float rr[4] = {S1n.x, S1n.y, S1n.z, S1n.w}, dn[4] = {d4.x, d4.y, d4.z, d4.w};
            const float up[4] = {u4.x, u4.y, u4.z, u4.w}, hh[4] = {Hm.x, Hm.y, Hm.z, Hm.w};
            float o[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                o[e] = sweep_enc(cc[e], ll[e], rr[e], dn[e], up[e], e > 0 ? cc[e - 1] : bk, e < 3 ? cc[e + 1] : fk,
                                 hh[e]);
            *reinterpret_cast<float4*>(p_new + (int64_t)(q - 1) * pl + off) = make_float4(o[0], o[1], o[2], o[3]);
        }
        Pm = Pc; Pc = Pn; Pn = Pf; Xr = Xf;
        S1m = S1c; S1c = S1n;
        Hm = Hc; Hc = Hf;
    }
}

struct PairShape {
    int jtiles, ktiles, ni;
    unsigned grid;
};

// At most 32 workgroups per CU (four rounds of the 8 a CU holds): tiles over
// (j,k), i split into segments.  440x440x280 in one run (workgroups, i planes
// each: us per launch): 7840, 9: 184.0; 10080, 7: 185.2; 11840, 6: 189.5;
// 17600, 4: 200.6; 4160, 17 (16 per CU rounded up, just past two rounds):
// 197.0.
static PairShape pair_shape(int nx, int ny, int nz) {
    PairShape s;
    s.jtiles = (ny + kPairTJ - 1) / kPairTJ;
    s.ktiles = (nz + 4 * kPairKT - 1) / (4 * kPairKT);
    const int cols = s.jtiles * s.ktiles;
    int nseg = 32 * device_cus() / cols;
    nseg = nseg < 1 ? 1 : (nseg > nx ? nx : nseg);
    s.ni = (nx + nseg - 1) / nseg;
    nseg = (nx + s.ni - 1) / s.ni;
    s.grid = (unsigned)(nseg * cols);
    return s;
}

// The pair kernel halves HBM traffic per sweep but re-reads two halo planes
// per i-segment and idles k lanes on ragged k tiles.  While one sweep's
// working set (p in, p out, div: 12 B/cell) fits the 256 MiB Infinity Cache,
// single sweeps are served from it and are faster (200-sweep solves,
// bit-identical results: 110x110x68 0.94 vs 1.52 ms, 300x300x192 6.26 vs
// 8.03 ms); above it the pair kernel wins (360x360x228 12.8 vs 15.8 ms,
// 440x440x280 20.0 vs 28.6 ms).  BOILSIM_JACOBI_PAIR=1 (=0) forces the pair
// kernel on (off) whatever the size; both give the same bits.
constexpr int64_t kInfinityCacheBytes = 256ll << 20;
static bool pair_pays(int64_t cells) {
    const char* e = getenv("BOILSIM_JACOBI_PAIR");
    if (e && e[0] == '1') return true;
    if (e && e[0] == '0') return false;
    return 12 * cells > kInfinityCacheBytes;
}

// Sentinel -> 0 in place (restores the reference's zero at non-fluid cells).
template <int VEC>
__global__ __launch_bounds__(256) void decode_kernel(float* __restrict__ p, int64_t nvec) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nvec) return;
    float v[VEC];
    load_vec<VEC>(p, t * VEC, v);
    bool any = false;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        const uint32_t b = __float_as_uint(v[e]);
        if (b == kSentAir || b == kSentSolid) { v[e] = 0.0f; any = true; }
    }
    if (any) store_vec<VEC>(p, t * VEC, v);
}

struct SweepShape {
    int vec;
    FastDiv nzv;
    uint32_t chunks;
    dim3 grid;
};

static int pick_vec(int nz, std::initializer_list<const void*> ptrs) {
    int vec = (nz % 4 == 0) ? 4 : (nz % 2 == 0) ? 2 : 1;
    for (const void* p : ptrs) {
        const uintptr_t a = (uintptr_t)p;
        while (vec > 1 && (a % (uintptr_t)(4 * vec)) != 0) vec >>= 1;
    }
    return vec;
}

static SweepShape sweep_shape(int nx, int ny, int nz, int vec) {
    SweepShape s;
    s.vec = vec;
    const uint32_t nzv = (uint32_t)(nz / vec);
    s.nzv = make_fastdiv(nzv);
    const uint64_t per_plane = (uint64_t)ny * nzv;
    s.chunks = (uint32_t)((per_plane + 255) / 256);
    s.grid = dim3(s.chunks, (unsigned)nx, 1);
    return s;
}

template <int IN, int OUT>
static void launch_sweep_t(const SweepShape& sh, hipStream_t st, float* p_new, const float* p_old,
                           const float* div, const int* mat, int nx, int ny, int nz, float c2,
                           float dt, int mf, int ma) {
    switch (sh.vec) {
        case 4:
            jacobi_sweep_kernel<4, IN, OUT><<<sh.grid, 256, 0, st>>>(p_new, p_old, div, mat, nx, ny, nz,
                                                                      sh.nzv, sh.chunks, c2, dt, mf, ma);
            break;
        case 2:
            jacobi_sweep_kernel<2, IN, OUT><<<sh.grid, 256, 0, st>>>(p_new, p_old, div, mat, nx, ny, nz,
                                                                      sh.nzv, sh.chunks, c2, dt, mf, ma);
            break;
        default:
            jacobi_sweep_kernel<1, IN, OUT><<<sh.grid, 256, 0, st>>>(p_new, p_old, div, mat, nx, ny, nz,
                                                                      sh.nzv, sh.chunks, c2, dt, mf, ma);
            break;
    }
}

static void launch_decode(hipStream_t st, float* p, int64_t n, int vec) {
    const int64_t nvec = n / vec;
    const unsigned blocks = (unsigned)((nvec + 255) / 256);
    switch (vec) {
        case 4: decode_kernel<4><<<blocks, 256, 0, st>>>(p, nvec); break;
        case 2: decode_kernel<2><<<blocks, 256, 0, st>>>(p, nvec); break;
        default: decode_kernel<1><<<blocks, 256, 0, st>>>(p, nvec); break;
    }
}

// (rho*dx)*dx in float32, the leading factor of the reference's rhs.
static inline float rhs_coef(float dx, float rho) {
    volatile float a = rho * dx;
    volatile float b = a * dx;
    return b;
}

// ---------------------------------------------------------------------------
// Diagnostic: div6 against the IEEE quotient x / 6.0f for all 2^32 float bit
// patterns on this device (262144 workgroups x 256 threads x 64 patterns).
// counts[0] += patterns where they differ (NaN = NaN), counts[1] += patterns
// on the fma path, counts[2] += patterns where the one-term x * RN(1/6)
// differs (the check's control).  Reads no memory but the counters.
// ---------------------------------------------------------------------------
constexpr uint32_t kDiv6PerThread = 64;
constexpr uint32_t kDiv6Groups = (uint32_t)((1ull << 32) / (256ull * kDiv6PerThread));

__device__ __forceinline__ bool same_float(float a, float b) {
    if (__builtin_isnan(a) && __builtin_isnan(b)) return true;
    return __float_as_uint(a) == __float_as_uint(b);
}

__global__ __launch_bounds__(256) void div6_check_kernel(unsigned long long* counts) {
    const float zh = 0x1.555556p-3f;
    const uint32_t base = (blockIdx.x * 256u + threadIdx.x) * kDiv6PerThread;
    unsigned c[3] = {0u, 0u, 0u};
    for (uint32_t k = 0; k < kDiv6PerThread; ++k) {
        const float x = __uint_as_float(base + k);
        const float q = x / 6.0f;
        const float ax = fabsf(x);
        c[0] += same_float(div6(x), q) ? 0u : 1u;
        c[1] += (ax >= 0x1p-96f && ax <= 0x1.fffffep127f) ? 1u : 0u;
        c[2] += same_float(x * zh, q) ? 0u : 1u;
    }
    __shared__ unsigned red[3][4];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        unsigned v = c[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const unsigned long long v = (unsigned long long)red[threadIdx.x][0] + red[threadIdx.x][1] +
                                     red[threadIdx.x][2] + red[threadIdx.x][3];
        if (v) atomicAdd(counts + threadIdx.x, v);
    }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

}  // namespace bs

using namespace bs;

extern "C" int bs_jacobi_pressure_launch_stream(float* p_new, const float* p_old, const float* div_u,
                                                const int* mat, int nx, int ny, int nz, float dx,
                                                float dt, float rho, int mat_fluid, int mat_air,
                                                hipStream_t stream) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    const int vec = pick_vec(nz, {p_new, p_old, div_u, mat});
    const SweepShape sh = sweep_shape(nx, ny, nz, vec);
    launch_sweep_t<kInMask, kOutDecoded>(sh, stream, p_new, p_old, div_u, mat, nx, ny, nz,
                                         rhs_coef(dx, rho), dt, mat_fluid, mat_air);
    return (int)hipGetLastError();
}

extern "C" int bs_jacobi_pressure_launch(float* p_new, const float* p_old, const float* div_u,
                                         const int* mat, int nx, int ny, int nz, float dx, float dt,
                                         float rho, int mat_fluid, int mat_air) {
    return bs_jacobi_pressure_launch_stream(p_new, p_old, div_u, mat, nx, ny, nz, dx, dt, rho,
                                            mat_fluid, mat_air, nullptr);
}

// Fused n_iter Jacobi solve (pressure_solve_launch, jacobi_pressure.cu:228-270).
// zero_init != 0 promises p == 0 on entry (the projection's cold start) and
// lets the first sweep skip reading p.
extern "C" int bs_jacobi_solve_launch(float* p, float* p_tmp, const float* div_u, const int* mat,
                                      int nx, int ny, int nz, float dx, float dt, float rho,
                                      int n_iter, int mat_fluid, int mat_air, int zero_init,
                                      hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || n_iter <= 0) return hipSuccess;
    const int vec = pick_vec(nz, {p, p_tmp, div_u, mat});
    const SweepShape sh = sweep_shape(nx, ny, nz, vec);
    const float c2 = rhs_coef(dx, rho);
    const int64_t n = (int64_t)nx * ny * nz;
    if (n_iter == 1) {
        if (zero_init)
            launch_sweep_t<kInZero, kOutDecoded>(sh, st, p_tmp, p, div_u, mat, nx, ny, nz, c2, dt,
                                                 mat_fluid, mat_air);
        else
            launch_sweep_t<kInMask, kOutDecoded>(sh, st, p_tmp, p, div_u, mat, nx, ny, nz, c2, dt,
                                                 mat_fluid, mat_air);
    } else {
        // The reference's ping-pong (fluid.py:688-700): sweep `it` reads p when
        // even and p_tmp when odd.  Here `cur` holds the latest result.
        float* cur = p;
        float* nxt = p_tmp;
        auto flip = [&] { float* t = cur; cur = nxt; nxt = t; };
        if (zero_init)
            launch_sweep_t<kInZero, kOutEncoded>(sh, st, nxt, cur, div_u, mat, nx, ny, nz, c2, dt,
                                                 mat_fluid, mat_air);
        else
            launch_sweep_t<kInMask, kOutEncoded>(sh, st, nxt, cur, div_u, mat, nx, ny, nz, c2, dt,
                                                 mat_fluid, mat_air);
        flip();
        // Sweeps 1 .. n-3 in pairs where the pair kernel applies.  A pair flips
        // the buffers once where the reference flips twice, so the number of
        // pairs is even and the buffer holding each later sweep's result is
        // the reference's.  The last two sweeps always run singly, so both
        // buffers end exactly as the reference leaves them.
        const int middle = n_iter >= 3 ? n_iter - 3 : 0;
        const int pairs = vec == 4 && pair_pays(n) ? (middle / 2) & ~1 : 0;
        if (pairs > 0) {
            const PairShape ps = pair_shape(nx, ny, nz);
            for (int q = 0; q < pairs; ++q) {
                jacobi_pair_kernel<<<ps.grid, 256, 0, st>>>(nxt, cur, div_u, nx, ny, nz, ps.ni, ps.jtiles,
                                                            ps.ktiles, c2, dt);
                flip();
            }
        }
        for (int q = 0; q < middle - 2 * pairs; ++q) {
            launch_sweep_t<kInEncoded, kOutEncoded>(sh, st, nxt, cur, div_u, mat, nx, ny, nz, c2, dt,
                                                    mat_fluid, mat_air);
            flip();
        }
        if (n_iter >= 3) {
            launch_sweep_t<kInEncoded, kOutEncoded>(sh, st, nxt, cur, div_u, mat, nx, ny, nz, c2, dt,
                                                    mat_fluid, mat_air);
            flip();
        }
        launch_sweep_t<kInEncoded, kOutDecoded>(sh, st, nxt, cur, div_u, mat, nx, ny, nz, c2, dt,
                                                mat_fluid, mat_air);
        // nxt holds the last result; cur the previous sweep's, encoded.  With
        // an odd count cur is p, which the copy below overwrites.
        if ((n_iter & 1) == 0) launch_decode(st, cur, n, vec);
    }
    BS_RETURN_IF_ERR(hipGetLastError());
    if (n_iter & 1) {
        // Odd count: the last sweep wrote p_tmp; the result must end in p
        // (jacobi_pressure.cu:262-268).
        BS_RETURN_IF_ERR(hipMemcpyAsync(p, p_tmp, sizeof(float) * (size_t)n,
                                        hipMemcpyDeviceToDevice, st));
    }
    return (int)hipGetLastError();
}

extern "C" int bs_pressure_solve_launch(float* p, float* p_tmp, const float* div_u, const int* mat,
                                        int nx, int ny, int nz, float dx, float dt, float rho,
                                        int n_iter, int mat_fluid, int mat_air) {
    return bs_jacobi_solve_launch(p, p_tmp, div_u, mat, nx, ny, nz, dx, dt, rho, n_iter, mat_fluid,
                                  mat_air, 0, nullptr);
}

// Device check of the exact division by 6 (no counterpart in the reference;
// see div6_check_kernel).  counts: three device uint64s, zeroed here.
extern "C" int bs_div6_check(unsigned long long* counts, hipStream_t stream) {
    bs::zero_u64_kernel<<<1, 64, 0, stream>>>(counts, 3);
    bs::div6_check_kernel<<<bs::kDiv6Groups, 256, 0, stream>>>(counts);
    return (int)hipGetLastError();
}
