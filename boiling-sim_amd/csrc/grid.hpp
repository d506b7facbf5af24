// Tile mapping, vector loads/stores, the one-byte cell code and the
// deterministic partial-sum helpers shared by the Krylov and multigrid kernels.
#pragma once

#include "common.hpp"

namespace bs {

// A tile is 256 threads x VEC consecutive cells of one (j,k) plane chunk at
// one i.  Grid-stride kernels give each workgroup a contiguous run of tiles
// ordered i-fastest, so a workgroup walks down i for a fixed plane chunk and
// the i-1 neighbour plane it needs was loaded by itself one tile earlier.
struct Tiles {
    int nx, ny, nz;
    FastDiv nzv;      // nz / VEC
    uint32_t chunks;  // 256-thread chunks per plane
    uint32_t ntiles;  // chunks * nx
};

inline Tiles make_tiles(int nx, int ny, int nz, int vec) {
    Tiles T;
    T.nx = nx; T.ny = ny; T.nz = nz;
    const uint32_t nzv = (uint32_t)(nz / vec);
    T.nzv = make_fastdiv(nzv);
    T.chunks = (uint32_t)(((uint64_t)ny * nzv + 255) / 256);
    T.ntiles = T.chunks * (uint32_t)nx;
    return T;
}

inline int vec_for(int nz, std::initializer_list<const void*> ptrs) {
    int vec = (nz % 4 == 0) ? 4 : (nz % 2 == 0) ? 2 : 1;
    for (const void* q : ptrs) {
        const uintptr_t a = (uintptr_t)q;
        while (vec > 1 && (a % (uintptr_t)(4 * vec)) != 0) vec >>= 1;
    }
    return vec;
}

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

// One tile per workgroup, tiles ordered chunk-fastest (a plane's chunks are
// consecutive blocks) behind the XCD remap; false when this thread is past
// the plane.  c0 is the linear index of the thread's first cell.
__device__ __forceinline__ bool one_tile(const Tiles& T, int VEC, int& i, int& j, int& k0, int64_t& c0) {
    const uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t chunk = t % T.chunks;
    i = (int)(t / T.chunks);
    const uint32_t q = chunk * 256u + threadIdx.x;
    if (q >= (uint32_t)T.ny * T.nzv.d) return false;
    j = (int)fdiv(q, T.nzv);
    k0 = (int)(q - (uint32_t)j * T.nzv.d) * VEC;
    c0 = ((int64_t)i * T.ny + j) * T.nz + k0;
    return true;
}

// VEC consecutive int32 values (material ids).
template <int VEC>
__device__ __forceinline__ void ldi(const int* __restrict__ p, int64_t off, int (&v)[VEC]) {
    if constexpr (VEC == 4) {
        const int4 t = *reinterpret_cast<const int4*>(p + off);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else if constexpr (VEC == 2) {
        const int2 t = *reinterpret_cast<const int2*>(p + off);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = p[off];
    }
}

template <int VEC>
__device__ __forceinline__ void ldf(const float* __restrict__ p, int64_t off, float (&v)[VEC]) {
    if constexpr (VEC == 8) {
        const float4 a = *reinterpret_cast<const float4*>(p + off);
        const float4 b = *reinterpret_cast<const float4*>(p + off + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else if constexpr (VEC == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p + off);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else if constexpr (VEC == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p + off);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = p[off];
    }
}

// Non-temporal loads and stores: for streams that the next kernels do not
// read again, so that they do not displace the fields that they do.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int VEC>
__device__ __forceinline__ void ldf_nt(const float* __restrict__ p, int64_t off, float (&v)[VEC]) {
    if constexpr (VEC == 8) {
        const f32x4_t a = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p + off));
        const f32x4_t b = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p + off + 4));
#pragma unroll
        for (int e = 0; e < 4; ++e) { v[e] = a[e]; v[4 + e] = b[e]; }
    } else if constexpr (VEC == 4) {
        const f32x4_t t = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p + off));
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = t[e];
    } else if constexpr (VEC == 2) {
        const f32x2_t t = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p + off));
        v[0] = t[0];
        v[1] = t[1];
    } else {
        v[0] = __builtin_nontemporal_load(p + off);
    }
}

template <int VEC>
__device__ __forceinline__ void stf_nt(float* __restrict__ p, int64_t off, const float (&v)[VEC]) {
    if constexpr (VEC == 8) {
        __builtin_nontemporal_store(f32x4_t{v[0], v[1], v[2], v[3]}, reinterpret_cast<f32x4_t*>(p + off));
        __builtin_nontemporal_store(f32x4_t{v[4], v[5], v[6], v[7]}, reinterpret_cast<f32x4_t*>(p + off + 4));
    } else if constexpr (VEC == 4) {
        __builtin_nontemporal_store(f32x4_t{v[0], v[1], v[2], v[3]}, reinterpret_cast<f32x4_t*>(p + off));
    } else if constexpr (VEC == 2) {
        __builtin_nontemporal_store(f32x2_t{v[0], v[1]}, reinterpret_cast<f32x2_t*>(p + off));
    } else {
        __builtin_nontemporal_store(v[0], p + off);
    }
}

template <int VEC>
__device__ __forceinline__ void stf(float* __restrict__ p, int64_t off, const float (&v)[VEC]) {
    if constexpr (VEC == 8) {
        *reinterpret_cast<float4*>(p + off) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4*>(p + off + 4) = make_float4(v[4], v[5], v[6], v[7]);
    } else if constexpr (VEC == 4) {
        *reinterpret_cast<float4*>(p + off) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (VEC == 2) {
        *reinterpret_cast<float2*>(p + off) = make_float2(v[0], v[1]);
    } else {
        p[off] = v[0];
    }
}

template <int VEC>
__device__ __forceinline__ void ldc(const uint8_t* __restrict__ p, int64_t off, uint8_t (&v)[VEC]) {
    if constexpr (VEC == 8) {
        const uint2 t = *reinterpret_cast<const uint2*>(p + off);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[e] = (t.x >> (8 * e)) & 0xFF;
            v[4 + e] = (t.y >> (8 * e)) & 0xFF;
        }
    } else if constexpr (VEC == 4) {
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
    if constexpr (VEC == 8) {
        uint2 t;
        t.x = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
        t.y = (uint32_t)v[4] | ((uint32_t)v[5] << 8) | ((uint32_t)v[6] << 16) | ((uint32_t)v[7] << 24);
        *reinterpret_cast<uint2*>(p + off) = t;
    } else if constexpr (VEC == 4) {
        *reinterpret_cast<uint32_t*>(p + off) =
            (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
    } else if constexpr (VEC == 2) {
        *reinterpret_cast<uint16_t*>(p + off) = (uint16_t)(v[0] | (v[1] << 8));
    } else {
        p[off] = v[0];
    }
}

// Vectors stored in 16-bit floats (the inner solve of the mixed-precision
// refinement): VEC elements of storage type S are loaded as one 2*VEC-byte
// access and widened to float; stores round to nearest even.
template <int BYTES> struct RawBytes;
template <> struct RawBytes<2> { using T = uint16_t; };
template <> struct RawBytes<4> { using T = uint32_t; };
template <> struct RawBytes<8> { using T = uint2; };
template <> struct RawBytes<16> { using T = uint4; };

template <int VEC, typename S>
__device__ __forceinline__ void ldv(const S* __restrict__ p, int64_t off, float (&v)[VEC]) {
    if constexpr (sizeof(S) == 4) {
        ldf<VEC>(reinterpret_cast<const float*>(p), off, v);
    } else {
        using R = typename RawBytes<2 * VEC>::T;
        const R raw = *reinterpret_cast<const R*>(p + off);
        S s[VEC];
        __builtin_memcpy(s, &raw, sizeof(raw));
#pragma unroll
        for (int e = 0; e < VEC; ++e) v[e] = (float)s[e];
    }
}

template <int VEC, typename S>
__device__ __forceinline__ void stv(S* __restrict__ p, int64_t off, const float (&v)[VEC]) {
    if constexpr (sizeof(S) == 4) {
        stf<VEC>(reinterpret_cast<float*>(p), off, v);
    } else {
        using R = typename RawBytes<2 * VEC>::T;
        S s[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) s[e] = (S)v[e];
        R raw;
        __builtin_memcpy(&raw, s, sizeof(raw));
        *reinterpret_cast<R*>(p + off) = raw;
    }
}

// Value of one stored element as float.
template <typename S>
__device__ __forceinline__ float ld1(const S* __restrict__ p, int64_t off) { return (float)p[off]; }

// A float rounded through storage type S (identity for float).
template <typename S>
__device__ __forceinline__ float through(float v) { return (float)(S)v; }

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

// The multigrid kernels' form of the same operator on a fluid cell of code c:
// (8 - c) v_c - s, s the sum of the fluid neighbours' values (a non-fluid
// neighbour adds +0; its effect is in 8 - c = 6 - #solid neighbours).  One
// select per neighbour (fz), or none where the values are +0 off the fluid by
// construction; the rounding differs from the reference's 6 v_c - S(v).
__device__ __forceinline__ float fz(uint8_t c, float v) { return c >= kCodeFluid0 ? v : 0.0f; }
__device__ __forceinline__ float lap_mg(uint8_t c, float ps, float s) { return (float)(8 - (int)c) * ps - s; }

// Loads of a 7-point neighbourhood for VEC cells (values and codes); out-of-
// bounds neighbours read as air.
template <int VEC>
struct Nbhd {
    uint8_t cc[VEC], cl[VEC], cr[VEC], cd[VEC], cu[VEC], cb, cf;
    float vc[VEC], vl[VEC], vr[VEC], vd[VEC], vu[VEC], vb, vf;

    template <typename S>
    __device__ __forceinline__ void load(const Tiles& T, const uint8_t* __restrict__ code,
                                         const S* __restrict__ v, int i, int j, int k0, int64_t c0) {
        const int64_t plane = (int64_t)T.ny * T.nz;
        ldc<VEC>(code, c0, cc);
        ldv<VEC>(v, c0, vc);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            cl[e] = cr[e] = cd[e] = cu[e] = kCodeAir;
            vl[e] = vr[e] = vd[e] = vu[e] = 0.0f;
        }
        if (i > 0) { ldc<VEC>(code, c0 - plane, cl); ldv<VEC>(v, c0 - plane, vl); }
        if (i < T.nx - 1) { ldc<VEC>(code, c0 + plane, cr); ldv<VEC>(v, c0 + plane, vr); }
        if (j > 0) { ldc<VEC>(code, c0 - T.nz, cd); ldv<VEC>(v, c0 - T.nz, vd); }
        if (j < T.ny - 1) { ldc<VEC>(code, c0 + T.nz, cu); ldv<VEC>(v, c0 + T.nz, vu); }
        cb = cf = kCodeAir;
        vb = vf = 0.0f;
        if (k0 > 0) { cb = code[c0 - 1]; vb = ld1(v, c0 - 1); }
        if (k0 + VEC < T.nz) { cf = code[c0 + VEC]; vf = ld1(v, c0 + VEC); }
    }

    // load() for values that are exactly +0 on every non-fluid cell (and
    // outside the grid): the neighbours' codes are not read (apply_mg then
    // needs only the centre's), they are set to a fluid code.
    template <typename S>
    __device__ __forceinline__ void load_v(const Tiles& T, const uint8_t* __restrict__ code,
                                           const S* __restrict__ v, int i, int j, int k0, int64_t c0) {
        const int64_t plane = (int64_t)T.ny * T.nz;
        ldc<VEC>(code, c0, cc);
        ldv<VEC>(v, c0, vc);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            cl[e] = cr[e] = cd[e] = cu[e] = kCodeFluid0;
            vl[e] = vr[e] = vd[e] = vu[e] = 0.0f;
        }
        if (i > 0) ldv<VEC>(v, c0 - plane, vl);
        if (i < T.nx - 1) ldv<VEC>(v, c0 + plane, vr);
        if (j > 0) ldv<VEC>(v, c0 - T.nz, vd);
        if (j < T.ny - 1) ldv<VEC>(v, c0 + T.nz, vu);
        cb = cf = kCodeFluid0;
        vb = vf = 0.0f;
        if (k0 > 0) vb = ld1(v, c0 - 1);
        if (k0 + VEC < T.nz) vf = ld1(v, c0 + VEC);
    }

    // 6*v - S(v) at element e in the reference's float order
    // (laplacian_spmv.cu:120-142); 0 for a non-fluid centre.
    __device__ __forceinline__ float apply(int e) const {
        if (cc[e] < kCodeFluid0) return 0.0f;
        const float ps = vc[e];
        const uint8_t c_b = e > 0 ? cc[e - 1] : cb;
        const float v_b = e > 0 ? vc[e - 1] : vb;
        const uint8_t c_f = e < VEC - 1 ? cc[e + 1] : cf;
        const float v_f = e < VEC - 1 ? vc[e + 1] : vf;
        float s = nbr_term(cl[e], vl[e], ps) + nbr_term(cr[e], vr[e], ps);
        s = s + nbr_term(cd[e], vd[e], ps);
        s = s + nbr_term(cu[e], vu[e], ps);
        s = s + nbr_term(c_b, v_b, ps);
        s = s + nbr_term(c_f, v_f, ps);
        return 6.0f * ps - s;
    }

    // The multigrid form (lap_mg): neighbour sum in the order -i, +i, -j, +j,
    // -k, +k over fluid neighbours only.
    __device__ __forceinline__ float apply_mg(int e) const {
        if (cc[e] < kCodeFluid0) return 0.0f;
        const uint8_t c_b = e > 0 ? cc[e - 1] : cb;
        const float v_b = e > 0 ? vc[e - 1] : vb;
        const uint8_t c_f = e < VEC - 1 ? cc[e + 1] : cf;
        const float v_f = e < VEC - 1 ? vc[e + 1] : vf;
        float s = fz(cl[e], vl[e]) + fz(cr[e], vr[e]);
        s = s + fz(cd[e], vd[e]);
        s = s + fz(cu[e], vu[e]);
        s = s + fz(c_b, v_b);
        s = s + fz(c_f, v_f);
        return lap_mg(cc[e], vc[e], s);
    }

    // Same operator with the neighbour sum and the product in FP64.
    __device__ __forceinline__ double apply64(int e) const {
        if (cc[e] < kCodeFluid0) return 0.0;
        const float ps = vc[e];
        const uint8_t c_b = e > 0 ? cc[e - 1] : cb;
        const float v_b = e > 0 ? vc[e - 1] : vb;
        const uint8_t c_f = e < VEC - 1 ? cc[e + 1] : cf;
        const float v_f = e < VEC - 1 ? vc[e + 1] : vf;
        double s = (double)nbr_term(cl[e], vl[e], ps) + (double)nbr_term(cr[e], vr[e], ps);
        s += (double)nbr_term(cd[e], vd[e], ps);
        s += (double)nbr_term(cu[e], vu[e], ps);
        s += (double)nbr_term(c_b, v_b, ps);
        s += (double)nbr_term(c_f, v_f, ps);
        return 6.0 * (double)ps - s;
    }
};

// ---------------------------------------------------------------------------
// Deterministic reductions without inter-workgroup fences: a kernel writes one
// FP64 partial per workgroup (value k of workgroup w at k*G + w); the kernel
// that needs the total sums the G partials itself, every workgroup in the same
// fixed order, so all agree bit for bit.  The stream order between the two
// kernels provides visibility, which avoids the per-workgroup L2 write-back an
// agent-scope fence costs on a multi-XCD device.
// ---------------------------------------------------------------------------
constexpr int kThreads = 256;

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sums of NV values over a workgroup of NT threads; results valid in every thread.
template <int NV, int NT = kThreads>
__device__ __forceinline__ void block_sums(double (&v)[NV]) {
    __shared__ double red[NV][NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum64(v[k]);
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = red[k][0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) s += red[k][w];
        v[k] = s;
    }
    __syncthreads();
}

// Writes this workgroup's NV partials.
template <int NV, int NT = kThreads>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* __restrict__ partials) {
    block_sums<NV, NT>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = v[k];
    }
}

// Sums G partials per value (written by an earlier kernel) into out[].  Each
// thread adds partials threadIdx.x, threadIdx.x + 256, ... in that order, with
// eight loads in flight at a time (a running sum over single loads waits for
// each in turn); past the end it adds +0.0, which leaves a sum that starts at
// +0.0 unchanged, so the totals are those of the plain strided loop.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, unsigned G,
                                             double (&out)[NV]) {
    constexpr unsigned kBatch = 8;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double* pk = partials + (size_t)k * G;
        double acc = 0.0;
        for (unsigned w = threadIdx.x; w < G; w += kBatch * kThreads) {
            double v[kBatch];
#pragma unroll
            for (unsigned u = 0; u < kBatch; ++u) {
                const unsigned q = w + u * kThreads;
                const double x = pk[q < G ? q : G - 1];
                v[u] = q < G ? x : 0.0;
            }
#pragma unroll
            for (unsigned u = 0; u < kBatch; ++u) acc += v[u];
        }
        out[k] = acc;
    }
    block_sums<NV>(out);
}

}  // namespace bs
