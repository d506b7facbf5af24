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
