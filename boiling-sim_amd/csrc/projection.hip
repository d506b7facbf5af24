// Divergence, gradient subtraction and the on-device max|div| for the pressure
// projection (python/boilingsim/fluid.py:201-219, 279-330, 726-734), plus the
// fused projection entry point that chains them around the Jacobi solver.
//
// Each kernel is one HBM pass: a thread owns VEC consecutive k cells of one
// (i,j) row, neighbour faces/planes are re-read through L1/L2 only.
//   divergence: ux, uy, uz, mat read, div written   -> 16 + 4 B/cell
//   gradient:   p, mat read, 3 faces read+written   -> 28 + 4 B/cell
//   max|div|:   div read                            ->  4 B/cell
#include "boilsim_hip.h"
#include "common.hpp"

namespace bs {

template <int VEC>
__device__ __forceinline__ void ld(const float* __restrict__ p, int64_t off, float (&v)[VEC]) {
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
__device__ __forceinline__ void ld(const int* __restrict__ p, int64_t off, int (&v)[VEC]) {
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
__device__ __forceinline__ void st(float* __restrict__ p, int64_t off, const float (&v)[VEC]) {
    if constexpr (VEC == 4) {
        *reinterpret_cast<float4*>(p + off) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (VEC == 2) {
        *reinterpret_cast<float2*>(p + off) = make_float2(v[0], v[1]);
    } else {
        p[off] = v[0];
    }
}

struct RowMap {
    FastDiv nzv;     // nz / VEC
    uint32_t chunks; // 256-thread chunks per (j,k) plane
};

// Thread -> (i, j, k0) for the flat per-plane mapping shared by the kernels.
__device__ __forceinline__ bool row_coords(const RowMap& m, int ny, int VEC, int& i, int& j, int& k0) {
    const uint32_t nblk = gridDim.x * gridDim.y;
    const uint32_t lb = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nblk);
    const uint32_t chunk = lb % m.chunks;
    i = (int)(lb / m.chunks);
    const uint32_t q = chunk * 256u + threadIdx.x;
    if (q >= (uint32_t)ny * m.nzv.d) return false;
    j = (int)fdiv(q, m.nzv);
    k0 = (int)(q - (uint32_t)j * m.nzv.d) * VEC;
    return true;
}

// compute_divergence (fluid.py:201-219): dudx, dvdy, dwdz each divided by dx,
// summed (dudx + dvdy) + dwdz; non-fluid cells get 0.
template <int VEC>
__global__ __launch_bounds__(256) void divergence_kernel(float* __restrict__ div,
                                                         const float* __restrict__ ux,
                                                         const float* __restrict__ uy,
                                                         const float* __restrict__ uz,
                                                         const int* __restrict__ mat, int nx, int ny,
                                                         int nz, RowMap rm, float dx, int mf) {
    int i, j, k0;
    if (!row_coords(rm, ny, VEC, i, j, k0)) return;
    const int64_t c0 = ((int64_t)i * ny + j) * nz + k0;
    int m[VEC];
    ld<VEC>(mat, c0, m);
    float x0[VEC], x1[VEC], y0[VEC], y1[VEC], z[VEC + 1], out[VEC];
    ld<VEC>(ux, c0, x0);
    ld<VEC>(ux, c0 + (int64_t)ny * nz, x1);
    const int64_t yo = ((int64_t)i * (ny + 1) + j) * nz + k0;
    ld<VEC>(uy, yo, y0);
    ld<VEC>(uy, yo + nz, y1);
    const int64_t zo = ((int64_t)i * ny + j) * (nz + 1) + k0;
#pragma unroll
    for (int e = 0; e <= VEC; ++e) z[e] = uz[zo + e];
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        if (m[e] != mf) { out[e] = 0.0f; continue; }
        const float dudx = (x1[e] - x0[e]) / dx;
        const float dvdy = (y1[e] - y0[e]) / dx;
        const float dwdz = (z[e + 1] - z[e]) / dx;
        out[e] = dudx + dvdy + dwdz;
    }
    st<VEC>(div, c0, out);
}

// subtract_pressure_gradient_{x,y,z} (fluid.py:279-330), one pass: the thread
// of cell (i,j,k) updates that cell's low x-, y- and z-faces.  A face is
// corrected only when it is interior and both adjacent cells are fluid:
// u -= ((dt/rho)*(p_hi - p_lo))/dx.  Faces between fluid and air are left
// alone, exactly as the reference leaves them.
template <int VEC>
__global__ __launch_bounds__(256) void gradient_kernel(float* __restrict__ ux, float* __restrict__ uy,
                                                       float* __restrict__ uz,
                                                       const float* __restrict__ p,
                                                       const int* __restrict__ mat, int nx, int ny,
                                                       int nz, RowMap rm, float dtr, float dx, int mf) {
    int i, j, k0;
    if (!row_coords(rm, ny, VEC, i, j, k0)) return;
    const int64_t plane = (int64_t)ny * nz;
    const int64_t c0 = (int64_t)i * plane + (int64_t)j * nz + k0;
    int m[VEC];
    float pc[VEC];
    ld<VEC>(mat, c0, m);
    bool anyf = false;
#pragma unroll
    for (int e = 0; e < VEC; ++e) anyf |= (m[e] == mf);
    if (!anyf) return;
    ld<VEC>(p, c0, pc);
    // x faces (i >= 1)
    if (i > 0) {
        int ml[VEC];
        float pl[VEC], u[VEC];
        ld<VEC>(mat, c0 - plane, ml);
        ld<VEC>(p, c0 - plane, pl);
        ld<VEC>(ux, c0, u);
        bool w = false;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (m[e] == mf && ml[e] == mf) {
                float g = dtr * (pc[e] - pl[e]);
                g = g / dx;
                u[e] = u[e] - g;
                w = true;
            }
        }
        if (w) st<VEC>(ux, c0, u);
    }
    // y faces (j >= 1)
    if (j > 0) {
        int md[VEC];
        float pd[VEC], u[VEC];
        ld<VEC>(mat, c0 - nz, md);
        ld<VEC>(p, c0 - nz, pd);
        const int64_t yo = ((int64_t)i * (ny + 1) + j) * nz + k0;
        ld<VEC>(uy, yo, u);
        bool w = false;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (m[e] == mf && md[e] == mf) {
                float g = dtr * (pc[e] - pd[e]);
                g = g / dx;
                u[e] = u[e] - g;
                w = true;
            }
        }
        if (w) st<VEC>(uy, yo, u);
    }
    // z faces (k >= 1): rows of uz are nz+1 long, so these stay scalar.
    {
        const int64_t zo = ((int64_t)i * ny + j) * (nz + 1) + k0;
        float pb = 0.0f;
        int mb = mf + 1;
        if (k0 > 0) { pb = p[c0 - 1]; mb = mat[c0 - 1]; }
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const int mlo = (e > 0) ? m[e - 1] : mb;
            const float plo = (e > 0) ? pc[e - 1] : pb;
            if (k0 + e > 0 && m[e] == mf && mlo == mf) {
                float g = dtr * (pc[e] - plo);
                g = g / dx;
                uz[zo + e] = uz[zo + e] - g;
            }
        }
    }
}

// max |x| into *out_bits as the bit pattern of a non-negative float (the
// unsigned order of such patterns is the float order, so atomicMax on the
// bits is exact and order-independent).  *out_bits must start at 0.
__global__ __launch_bounds__(256) void max_abs_kernel(const float* __restrict__ x, int64_t n,
                                                      unsigned int* __restrict__ out_bits) {
    uint32_t best = 0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) {
        const uint32_t b = __float_as_uint(x[t]) & 0x7FFFFFFFu;
        best = b > best ? b : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)best, off, 64);
        best = o > best ? o : best;
    }
    __shared__ uint32_t sm[4];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = sm[0];
        for (int w = 1; w < 4; ++w) b = sm[w] > b ? sm[w] : b;
        atomicMax(out_bits, b);
    }
}

__global__ void zero_u32_kernel(unsigned int* p) { *p = 0u; }

__global__ __launch_bounds__(256) void fill_zero_kernel(float* __restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) p[t] = 0.0f;
}

// A fill kernel rather than hipMemsetAsync, so the sequence stays valid under
// stream capture.
static void fill_zero(float* p, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fill_zero_kernel<<<(unsigned)blocks, 256, 0, st>>>(p, n);
}

static int pick_vec2(int nz, std::initializer_list<const void*> ptrs) {
    int vec = (nz % 4 == 0) ? 4 : (nz % 2 == 0) ? 2 : 1;
    for (const void* p : ptrs) {
        const uintptr_t a = (uintptr_t)p;
        while (vec > 1 && (a % (uintptr_t)(4 * vec)) != 0) vec >>= 1;
    }
    return vec;
}

static RowMap row_map(int ny, int nz, int vec, dim3& grid, int nx) {
    RowMap rm;
    const uint32_t nzv = (uint32_t)(nz / vec);
    rm.nzv = make_fastdiv(nzv);
    rm.chunks = (uint32_t)(((uint64_t)ny * nzv + 255) / 256);
    grid = dim3(rm.chunks, (unsigned)nx, 1);
    return rm;
}

// ---------------------------------------------------------------------------
// No-slip faces, and the maxima the step reads back (fluid.py:111-162,
// 726-734, 796-844).
//
// enforce_no_slip_u{x,y,z} (fluid.py:111-150): a MAC face keeps its value only
// when it is interior and both cells it separates are fluid; every other face
// becomes +0.0.  In faces_kernel the thread of cell (i,j,k) owns that cell's
// low x-, y- and z-faces, and the cells on the high domain edges also own the
// edge faces ux[nx], uy[:,ny], uz[:,:,nz], so every face has one owner.
// A face is stored only when its bits change (a kept face, or one that is
// already +0.0, is not rewritten).
//
//   kNoSlip     the three no-slip launches in one pass
//   kDivNoSlip  no-slip, then compute_divergence of the result, in one pass:
//               a fluid cell reads its high faces (owned by its neighbours)
//               and applies the same keep rule to them, so the divergence
//               sees the post-no-slip values whether or not the owner has
//               stored them yet
//   kDivMax     compute_divergence, with max|div| over the cells and max|u|
//               over the faces reduced on the device (no face is written)
// ---------------------------------------------------------------------------
enum FaceMode { kNoSlip = 0, kDivNoSlip = 1, kDivMax = 2 };

__device__ __forceinline__ uint32_t abs_bits(float v) { return __float_as_uint(v) & 0x7FFFFFFFu; }
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// Both maxima of a workgroup into max_bits[0] (|div|) and max_bits[1] (|u|).
// The maxima only grow, so a block whose value does not exceed one it reads
// back (possibly stale, hence never too large) has nothing to add.
__device__ __forceinline__ void block_max2(uint32_t a, uint32_t b, unsigned int* __restrict__ max_bits) {
    for (int off = 32; off > 0; off >>= 1) {
        a = umax(a, (uint32_t)__shfl_xor((int)a, off, 64));
        b = umax(b, (uint32_t)__shfl_xor((int)b, off, 64));
    }
    __shared__ uint32_t sm[2][4];
    if ((threadIdx.x & 63) == 0) {
        sm[0][threadIdx.x >> 6] = a;
        sm[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = umax(umax(sm[0][0], sm[0][1]), umax(sm[0][2], sm[0][3]));
        b = umax(umax(sm[1][0], sm[1][1]), umax(sm[1][2], sm[1][3]));
        if (a > __hip_atomic_load(max_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(max_bits, a);
        if (b > __hip_atomic_load(max_bits + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(max_bits + 1, b);
    }
}

template <int VEC>
__device__ __forceinline__ bool changed(const float (&a)[VEC], const float (&b)[VEC]) {
    bool c = false;
#pragma unroll
    for (int e = 0; e < VEC; ++e) c |= __float_as_uint(a[e]) != __float_as_uint(b[e]);
    return c;
}

template <int VEC, int MODE>
__global__ __launch_bounds__(256) void faces_kernel(float* __restrict__ div, float* __restrict__ ux,
                                                    float* __restrict__ uy, float* __restrict__ uz,
                                                    const int* __restrict__ mat, int nx, int ny, int nz,
                                                    RowMap rm, float dx, int mf,
                                                    unsigned int* __restrict__ max_bits) {
    int i, j, k0;
    const bool live = row_coords(rm, ny, VEC, i, j, k0);
    uint32_t best_div = 0, best_u = 0;
    if (live) {
        const int64_t plane = (int64_t)ny * nz;
        const int64_t c0 = (int64_t)i * plane + (int64_t)j * nz + k0;
        const int64_t yo = ((int64_t)i * (ny + 1) + j) * nz + k0;
        const int64_t zo = ((int64_t)i * ny + j) * (nz + 1) + k0;
        const bool top_x = i == nx - 1, top_y = j == ny - 1, top_z = k0 + VEC == nz;
        int m[VEC];
        float x0[VEC], x1[VEC], y0[VEC], y1[VEC], z[VEC + 1];
        ld<VEC>(mat, c0, m);
        ld<VEC>(ux, c0, x0);
        ld<VEC>(ux, c0 + plane, x1);
        ld<VEC>(uy, yo, y0);
        ld<VEC>(uy, yo + nz, y1);
#pragma unroll
        for (int e = 0; e <= VEC; ++e) z[e] = uz[zo + e];
        if constexpr (MODE != kDivMax) {
            // Neighbour classes; a missing neighbour (domain edge) is ~mf,
            // which is never fluid, so edge faces are never kept.
            const int nf = ~mf;
            int ml[VEC], mr[VEC], md[VEC], mu[VEC];
#pragma unroll
            for (int e = 0; e < VEC; ++e) ml[e] = mr[e] = md[e] = mu[e] = nf;
            if (i > 0) ld<VEC>(mat, c0 - plane, ml);
            if (!top_x) ld<VEC>(mat, c0 + plane, mr);
            if (j > 0) ld<VEC>(mat, c0 - nz, md);
            if (!top_y) ld<VEC>(mat, c0 + nz, mu);
            const int mb = k0 > 0 ? mat[c0 - 1] : nf;
            const int mt = top_z ? nf : mat[c0 + VEC];
            float nx0[VEC], nx1[VEC], ny0[VEC], ny1[VEC], nz_[VEC + 1];
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const bool fl = m[e] == mf;
                nx0[e] = (fl && ml[e] == mf) ? x0[e] : 0.0f;
                nx1[e] = (fl && mr[e] == mf) ? x1[e] : 0.0f;
                ny0[e] = (fl && md[e] == mf) ? y0[e] : 0.0f;
                ny1[e] = (fl && mu[e] == mf) ? y1[e] : 0.0f;
            }
#pragma unroll
            for (int e = 0; e <= VEC; ++e) {
                const int lo = e > 0 ? m[e - 1] : mb;
                const int hi = e < VEC ? m[e] : mt;
                nz_[e] = (lo == mf && hi == mf) ? z[e] : 0.0f;
            }
            // the owned faces
            if (changed<VEC>(nx0, x0)) st<VEC>(ux, c0, nx0);
            if (top_x && changed<VEC>(nx1, x1)) st<VEC>(ux, c0 + plane, nx1);
            if (changed<VEC>(ny0, y0)) st<VEC>(uy, yo, ny0);
            if (top_y && changed<VEC>(ny1, y1)) st<VEC>(uy, yo + nz, ny1);
#pragma unroll
            for (int e = 0; e < VEC; ++e)
                if (__float_as_uint(nz_[e]) != __float_as_uint(z[e])) uz[zo + e] = nz_[e];
            if (top_z && __float_as_uint(nz_[VEC]) != __float_as_uint(z[VEC])) uz[zo + VEC] = nz_[VEC];
            if constexpr (MODE == kDivNoSlip) {
                float out[VEC];
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    if (m[e] != mf) { out[e] = 0.0f; continue; }
                    const float dudx = (nx1[e] - nx0[e]) / dx;
                    const float dvdy = (ny1[e] - ny0[e]) / dx;
                    const float dwdz = (nz_[e + 1] - nz_[e]) / dx;
                    out[e] = dudx + dvdy + dwdz;
                }
                st<VEC>(div, c0, out);
            }
        } else {
            float out[VEC];
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                if (m[e] != mf) {
                    out[e] = 0.0f;
                } else {
                    const float dudx = (x1[e] - x0[e]) / dx;
                    const float dvdy = (y1[e] - y0[e]) / dx;
                    const float dwdz = (z[e + 1] - z[e]) / dx;
                    out[e] = dudx + dvdy + dwdz;
                }
                best_div = umax(best_div, abs_bits(out[e]));
                best_u = umax(best_u, umax(abs_bits(x0[e]), umax(abs_bits(y0[e]), abs_bits(z[e]))));
                if (top_x) best_u = umax(best_u, abs_bits(x1[e]));
                if (top_y) best_u = umax(best_u, abs_bits(y1[e]));
            }
            if (top_z) best_u = umax(best_u, abs_bits(z[VEC]));
            st<VEC>(div, c0, out);
        }
    }
    if constexpr (MODE == kDivMax) block_max2(best_div, best_u, max_bits);
}

// max|u| over the three face arrays (compute_max_velocity, fluid.py:807-833):
// one grid-stride pass over each array, float4 loads on 16-byte-aligned
// arrays, one atomicMax per workgroup.
__global__ __launch_bounds__(256) void max_faces_kernel(const float* __restrict__ a, int64_t na,
                                                        const float* __restrict__ b, int64_t nb,
                                                        const float* __restrict__ c, int64_t nc,
                                                        unsigned int* __restrict__ out_bits) {
    uint32_t best = 0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    const int64_t t0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float* arr[3] = {a, b, c};
    const int64_t len[3] = {na, nb, nc};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float* x = arr[q];
        const int64_t n = len[q];
        int64_t head = 0;
        if (((uintptr_t)x & 15u) == 0) {
            const int64_t n4 = n / 4;
            const float4* x4 = reinterpret_cast<const float4*>(x);
            for (int64_t t = t0; t < n4; t += stride) {
                const float4 v = x4[t];
                best = umax(best, umax(umax(abs_bits(v.x), abs_bits(v.y)), umax(abs_bits(v.z), abs_bits(v.w))));
            }
            head = n4 * 4;
        }
        for (int64_t t = head + t0; t < n; t += stride) best = umax(best, abs_bits(x[t]));
    }
    for (int off = 32; off > 0; off >>= 1) best = umax(best, (uint32_t)__shfl_xor((int)best, off, 64));
    __shared__ uint32_t sm[4];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out_bits, umax(umax(sm[0], sm[1]), umax(sm[2], sm[3])));
}

__global__ void zero_u32x2_kernel(unsigned int* p) {
    p[0] = 0u;
    p[1] = 0u;
}

template <int MODE>
static int launch_faces(float* div, float* ux, float* uy, float* uz, const int* mat, int nx, int ny, int nz,
                        float dx, int mf, unsigned int* max_bits, hipStream_t st) {
    const int vec = pick_vec2(nz, {div, ux, uy, mat});
    dim3 grid;
    const RowMap rm = row_map(ny, nz, vec, grid, nx);
    switch (vec) {
        case 4: faces_kernel<4, MODE><<<grid, 256, 0, st>>>(div, ux, uy, uz, mat, nx, ny, nz, rm, dx, mf, max_bits); break;
        case 2: faces_kernel<2, MODE><<<grid, 256, 0, st>>>(div, ux, uy, uz, mat, nx, ny, nz, rm, dx, mf, max_bits); break;
        default: faces_kernel<1, MODE><<<grid, 256, 0, st>>>(div, ux, uy, uz, mat, nx, ny, nz, rm, dx, mf, max_bits); break;
    }
    return (int)hipGetLastError();
}

}  // namespace bs

using namespace bs;

extern "C" int bs_divergence_launch(float* div_u, const float* ux, const float* uy, const float* uz,
                                    const int* mat, int nx, int ny, int nz, float dx, int mat_fluid,
                                    hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    const int vec = pick_vec2(nz, {div_u, ux, uy, mat});
    dim3 grid;
    const RowMap rm = row_map(ny, nz, vec, grid, nx);
    switch (vec) {
        case 4: divergence_kernel<4><<<grid, 256, 0, st>>>(div_u, ux, uy, uz, mat, nx, ny, nz, rm, dx, mat_fluid); break;
        case 2: divergence_kernel<2><<<grid, 256, 0, st>>>(div_u, ux, uy, uz, mat, nx, ny, nz, rm, dx, mat_fluid); break;
        default: divergence_kernel<1><<<grid, 256, 0, st>>>(div_u, ux, uy, uz, mat, nx, ny, nz, rm, dx, mat_fluid); break;
    }
    return (int)hipGetLastError();
}

extern "C" int bs_subtract_gradient_launch(float* ux, float* uy, float* uz, const float* p,
                                           const int* mat, int nx, int ny, int nz, float dx,
                                           float dt, float rho, int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    const int vec = pick_vec2(nz, {ux, uy, p, mat});
    dim3 grid;
    const RowMap rm = row_map(ny, nz, vec, grid, nx);
    volatile float dtr = dt / rho;  // fluid.py:290 evaluates dt / rho in float32
    switch (vec) {
        case 4: gradient_kernel<4><<<grid, 256, 0, st>>>(ux, uy, uz, p, mat, nx, ny, nz, rm, dtr, dx, mat_fluid); break;
        case 2: gradient_kernel<2><<<grid, 256, 0, st>>>(ux, uy, uz, p, mat, nx, ny, nz, rm, dtr, dx, mat_fluid); break;
        default: gradient_kernel<1><<<grid, 256, 0, st>>>(ux, uy, uz, p, mat, nx, ny, nz, rm, dtr, dx, mat_fluid); break;
    }
    return (int)hipGetLastError();
}

extern "C" int bs_max_abs_launch(const float* x, long long n, unsigned int* out_bits, hipStream_t st) {
    zero_u32_kernel<<<1, 1, 0, st>>>(out_bits);
    if (n > 0) {
        long long blocks = (n + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        max_abs_kernel<<<(unsigned)blocks, 256, 0, st>>>(x, (int64_t)n, out_bits);
    }
    return (int)hipGetLastError();
}

// enforce_no_slip (fluid.py:111-162): the three face launches in one.
extern "C" int bs_enforce_no_slip_launch(float* ux, float* uy, float* uz, const int* mat, int nx, int ny, int nz,
                                         int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    return launch_faces<kNoSlip>(nullptr, ux, uy, uz, mat, nx, ny, nz, 1.0f, mat_fluid, nullptr, st);
}

// enforce_no_slip followed by compute_divergence (pipeline.py:606-608, then
// fluid.py:632-637), in one pass: the faces end as enforce_no_slip leaves
// them and div_u as compute_divergence computes it from those faces.
extern "C" int bs_divergence_no_slip_launch(float* div_u, float* ux, float* uy, float* uz, const int* mat, int nx,
                                            int ny, int nz, float dx, int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    return launch_faces<kDivNoSlip>(div_u, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid, nullptr, st);
}

// compute_divergence with max|div| (fluid.py:727-734) and max|u| over all
// faces (compute_max_velocity, fluid.py:807-833) from the same pass, left on
// the device as float bits: max_bits[0] = max|div|, max_bits[1] = max|u|.
extern "C" int bs_divergence_max_launch(float* div_u, const float* ux, const float* uy, const float* uz,
                                        const int* mat, int nx, int ny, int nz, float dx, int mat_fluid,
                                        unsigned int* max_bits, hipStream_t st) {
    zero_u32x2_kernel<<<1, 1, 0, st>>>(max_bits);
    if (nx <= 0 || ny <= 0 || nz <= 0) return (int)hipGetLastError();
    return launch_faces<kDivMax>(div_u, const_cast<float*>(ux), const_cast<float*>(uy), const_cast<float*>(uz),
                                 mat, nx, ny, nz, dx, mat_fluid, max_bits, st);
}

// compute_max_velocity's device reduction (fluid.py:807-833) in one launch
// over ux (nx+1,ny,nz), uy (nx,ny+1,nz) and uz (nx,ny,nz+1); max|u| is left
// on the device as float bits in *out_bits.
extern "C" int bs_max_velocity_launch(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz,
                                      unsigned int* out_bits, hipStream_t st) {
    zero_u32_kernel<<<1, 1, 0, st>>>(out_bits);
    if (nx <= 0 || ny <= 0 || nz <= 0) return (int)hipGetLastError();
    const int64_t na = (int64_t)(nx + 1) * ny * nz, nb = (int64_t)nx * (ny + 1) * nz,
                  nc = (int64_t)nx * ny * (nz + 1);
    int64_t blocks = (na / 4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    max_faces_kernel<<<(unsigned)blocks, 256, 0, st>>>(ux, na, uy, nb, uz, nc, out_bits);
    return (int)hipGetLastError();
}

extern "C" int bs_jacobi_solve_launch(float* p, float* p_tmp, const float* div_u, const int* mat,
                                      int nx, int ny, int nz, float dx, float dt, float rho,
                                      int n_iter, int mat_fluid, int mat_air, int zero_init,
                                      hipStream_t st);

// pressure_projection with the Jacobi solver (fluid.py:612-734): p = p_tmp = 0,
// div = ∇·u, n_iter sweeps, u -= grad p, div = ∇·u again, and the max|div|
// left on the device as float bits in *max_div_bits (the reference copies the
// whole field to the host for this, fluid.py:734).
extern "C" int bs_projection_jacobi_launch(float* p, float* p_tmp, float* div_u, float* ux, float* uy,
                                           float* uz, const int* mat, int nx, int ny, int nz,
                                           float dx, float dt, float rho, int n_iter, int mat_fluid,
                                           int mat_air, unsigned int* max_div_bits, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        if (max_div_bits) zero_u32_kernel<<<1, 1, 0, st>>>(max_div_bits);
        return (int)hipGetLastError();
    }
    const size_t n = (size_t)nx * ny * nz;
    if (n_iter <= 0) {
        // Nothing overwrites p or p_tmp: clear both as the reference does.
        fill_zero(p, (int64_t)n, st);
        fill_zero(p_tmp, (int64_t)n, st);
    }
    int rc = bs_divergence_launch(div_u, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid, st);
    if (rc) return rc;
    // zero_init: the reference zeroes p before the sweeps; the first sweep
    // here uses p == 0 without reading (or clearing) it.
    rc = bs_jacobi_solve_launch(p, p_tmp, div_u, mat, nx, ny, nz, dx, dt, rho, n_iter, mat_fluid,
                                mat_air, 1, st);
    if (rc) return rc;
    rc = bs_subtract_gradient_launch(ux, uy, uz, p, mat, nx, ny, nz, dx, dt, rho, mat_fluid, st);
    if (rc) return rc;
    rc = bs_divergence_launch(div_u, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid, st);
    if (rc) return rc;
    if (max_div_bits) rc = bs_max_abs_launch(div_u, (long long)n, max_div_bits, st);
    return rc;
}

// One fluid step's no-slip and projection (pipeline.py:606-620 around
// fluid.py:612-734, Jacobi solver): enforce_no_slip, pressure_projection,
// enforce_no_slip, with
//   - the first no-slip fused into the first divergence pass,
//   - the second no-slip left out: it would zero exactly the faces that are
//     on the domain edge or touch a non-fluid cell; the first one has already
//     set all of them to +0.0 and neither the solver nor the gradient
//     subtraction (which writes only interior fluid-fluid faces) touches them
//     afterwards, so it changes no bit,
//   - max|div| and max|u| (what compute_max_velocity returns once the step
//     has ended) from the diagnostic divergence pass, as float bits in
//     max_bits[0] and max_bits[1].
extern "C" int bs_projection_step_jacobi_launch(float* p, float* p_tmp, float* div_u, float* ux, float* uy,
                                                float* uz, const int* mat, int nx, int ny, int nz, float dx,
                                                float dt, float rho, int n_iter, int mat_fluid, int mat_air,
                                                unsigned int* max_bits, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        zero_u32x2_kernel<<<1, 1, 0, st>>>(max_bits);
        return (int)hipGetLastError();
    }
    const size_t n = (size_t)nx * ny * nz;
    if (n_iter <= 0) {
        fill_zero(p, (int64_t)n, st);
        fill_zero(p_tmp, (int64_t)n, st);
    }
    int rc = bs_divergence_no_slip_launch(div_u, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid, st);
    if (rc) return rc;
    rc = bs_jacobi_solve_launch(p, p_tmp, div_u, mat, nx, ny, nz, dx, dt, rho, n_iter, mat_fluid, mat_air, 1, st);
    if (rc) return rc;
    rc = bs_subtract_gradient_launch(ux, uy, uz, p, mat, nx, ny, nz, dx, dt, rho, mat_fluid, st);
    if (rc) return rc;
    return bs_divergence_max_launch(div_u, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid, max_bits, st);
}
