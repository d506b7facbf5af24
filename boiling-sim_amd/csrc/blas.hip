// The reference PCG's building blocks as standalone launchers, with the
// reference's signatures, so each can be checked on its own.  The fused
// solver in pcg.hip does not call them; they exist for callers (and tests)
// that drive the pieces one at a time.
//   laplacian_spmv     crates/cuda-kernels/src/laplacian_spmv.cu:60-163
//   diag_inverse_apply crates/cuda-kernels/src/diag_inverse_apply.cu:43-111
//   dot                crates/cuda-kernels/src/dot_reduce.cu:37-115 (FP64 here)
//   axpy_device, scaled_axpy, divide_scalars, compute_b, zero_buffer
//                      crates/cuda-kernels/src/axpy_device.cu:19-145
#include "boilsim_hip.h"
#include "reduce.hpp"

namespace bs {

__device__ __forceinline__ int mat_or_air(const int* __restrict__ mat, int i, int j, int k, int nx,
                                          int ny, int nz, int ma) {
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return ma;
    return mat[((int64_t)i * ny + j) * nz + k];
}

__device__ __forceinline__ float nbr_val(const float* __restrict__ p, const int* __restrict__ mat, int i,
                                         int j, int k, int nx, int ny, int nz, float ps, int mf, int ma) {
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return 0.0f;
    const int64_t c = ((int64_t)i * ny + j) * nz + k;
    const int m = mat[c];
    if (m == mf) return p[c];
    if (m == ma) return 0.0f;
    return ps;
}

// One thread per cell, k fastest.
__device__ __forceinline__ bool cell_of(int nx, int ny, int nz, int& i, int& j, int& k) {
    k = blockIdx.x * 64 + (threadIdx.x & 63);
    j = blockIdx.y * 4 + (threadIdx.x >> 6);
    i = blockIdx.z;
    return k < nz && j < ny && i < nx;
}

__global__ __launch_bounds__(256) void spmv_kernel(float* __restrict__ ap, const float* __restrict__ p,
                                                   const int* __restrict__ mat, int nx, int ny, int nz,
                                                   int mf, int ma) {
    int i, j, k;
    if (!cell_of(nx, ny, nz, i, j, k)) return;
    const int64_t c = ((int64_t)i * ny + j) * nz + k;
    if (mat[c] != mf) { ap[c] = 0.0f; return; }
    const float ps = p[c];
    float s = nbr_val(p, mat, i - 1, j, k, nx, ny, nz, ps, mf, ma) + nbr_val(p, mat, i + 1, j, k, nx, ny, nz, ps, mf, ma);
    s = s + nbr_val(p, mat, i, j - 1, k, nx, ny, nz, ps, mf, ma);
    s = s + nbr_val(p, mat, i, j + 1, k, nx, ny, nz, ps, mf, ma);
    s = s + nbr_val(p, mat, i, j, k - 1, nx, ny, nz, ps, mf, ma);
    s = s + nbr_val(p, mat, i, j, k + 1, nx, ny, nz, ps, mf, ma);
    ap[c] = 6.0f * ps - s;
}

__global__ __launch_bounds__(256) void diag_apply_kernel(float* __restrict__ z, const float* __restrict__ r,
                                                         const int* __restrict__ mat, int nx, int ny,
                                                         int nz, int mf, int ma) {
    int i, j, k;
    if (!cell_of(nx, ny, nz, i, j, k)) return;
    const int64_t c = ((int64_t)i * ny + j) * nz + k;
    if (mat[c] != mf) { z[c] = 0.0f; return; }
    int ns = 0;
    const int nb[6] = {mat_or_air(mat, i - 1, j, k, nx, ny, nz, ma), mat_or_air(mat, i + 1, j, k, nx, ny, nz, ma),
                       mat_or_air(mat, i, j - 1, k, nx, ny, nz, ma), mat_or_air(mat, i, j + 1, k, nx, ny, nz, ma),
                       mat_or_air(mat, i, j, k - 1, nx, ny, nz, ma), mat_or_air(mat, i, j, k + 1, nx, ny, nz, ma)};
#pragma unroll
    for (int q = 0; q < 6; ++q) ns += (nb[q] != mf && nb[q] != ma) ? 1 : 0;
    const float dg = 6.0f - (float)ns;
    z[c] = (dg <= 0.0f) ? r[c] : r[c] / dg;
}

__global__ __launch_bounds__(256) void dot_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                  int64_t n, double* __restrict__ partials,
                                                  unsigned int* __restrict__ ticket,
                                                  double* __restrict__ out64, float* __restrict__ out32) {
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride)
        acc += (double)x[t] * (double)y[t];
    double v[1] = {acc};
    if (grid_reduce<1>(v, partials, ticket) && threadIdx.x == 0) {
        if (out64) *out64 = v[0];
        if (out32) *out32 = (float)v[0];
    }
}

__global__ void dot_zero_kernel(double* out64, float* out32, unsigned int* ticket) {
    if (threadIdx.x != 0) return;
    if (out64) *out64 = 0.0;
    if (out32) *out32 = 0.0f;
    *ticket = 0u;
}

__global__ void ticket_reset_kernel(unsigned int* ticket) {
    if (threadIdx.x == 0) *ticket = 0u;
}

__global__ __launch_bounds__(256) void axpy_kernel(float* __restrict__ y, const float* __restrict__ x,
                                                   const float* __restrict__ alpha_ptr, int n, int neg) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const float a = neg ? -(*alpha_ptr) : (*alpha_ptr);
    y[t] = a * x[t] + y[t];
}

__global__ __launch_bounds__(256) void scaled_axpy_kernel(float* p_out, const float* z, const float* p_old,
                                                          const float* beta_ptr, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const float bta = *beta_ptr;
    p_out[t] = z[t] + bta * p_old[t];
}

__global__ void divide_kernel(const float* num, const float* den, float* res) {
    if (threadIdx.x != 0) return;
    const float d = *den;
    *res = (d == 0.0f) ? 0.0f : (*num / d);
}

__global__ __launch_bounds__(256) void compute_b_kernel(float* __restrict__ b, const float* __restrict__ div,
                                                        const int* __restrict__ mat, int nx, int ny, int nz,
                                                        float scale, int mf) {
    int i, j, k;
    if (!cell_of(nx, ny, nz, i, j, k)) return;
    const int64_t c = ((int64_t)i * ny + j) * nz + k;
    b[c] = (mat[c] != mf) ? 0.0f : scale * div[c];
}

__global__ __launch_bounds__(256) void zero_kernel(float* __restrict__ p, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = 0.0f;
}

static dim3 cell_grid(int nx, int ny, int nz) {
    return dim3((unsigned)((nz + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nx);
}

}  // namespace bs

using namespace bs;

extern "C" int bs_laplacian_spmv_launch(float* Ap, const float* p, const int* mat, int nx, int ny, int nz,
                                        int mat_fluid, int mat_air) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    spmv_kernel<<<cell_grid(nx, ny, nz), 256>>>(Ap, p, mat, nx, ny, nz, mat_fluid, mat_air);
    return (int)hipGetLastError();
}

extern "C" int bs_diag_inverse_apply_launch(float* z, const float* r, const int* mat, int nx, int ny, int nz,
                                            int mat_fluid, int mat_air) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    diag_apply_kernel<<<cell_grid(nx, ny, nz), 256>>>(z, r, mat, nx, ny, nz, mat_fluid, mat_air);
    return (int)hipGetLastError();
}

// workspace: >= 1024 floats, as the reference's DOT_MAX_PARTIALS buffer.  The
// sum is accumulated in FP64; *result receives it rounded to float and, when
// result64 is not null, *result64 the FP64 value.
extern "C" int bs_dot_launch_f64(const float* x, const float* y, int n, float* workspace, float* result,
                                 double* result64) {
    unsigned int* ticket = (unsigned int*)workspace;
    double* partials = (double*)((char*)workspace + 256);
    if (n <= 0) {
        dot_zero_kernel<<<1, 64>>>(result64, result, ticket);
        return (int)hipGetLastError();
    }
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 240) blocks = 240;  // (4096 - 256) / 16 partial slots
    ticket_reset_kernel<<<1, 64>>>(ticket);
    dot_kernel<<<blocks, 256>>>(x, y, (int64_t)n, partials, ticket, result64, result);
    return (int)hipGetLastError();
}

extern "C" int bs_dot_launch(const float* x, const float* y, int n, float* workspace, float* result) {
    return bs_dot_launch_f64(x, y, n, workspace, result, nullptr);
}

extern "C" int bs_axpy_device_launch(float* y, const float* x, const float* alpha_ptr, int n, int negate_alpha) {
    if (n <= 0) return hipSuccess;
    axpy_kernel<<<(n + 255) / 256, 256>>>(y, x, alpha_ptr, n, negate_alpha);
    return (int)hipGetLastError();
}

extern "C" int bs_scaled_axpy_launch(float* p_out, const float* z, const float* p_old, const float* beta_ptr,
                                     int n) {
    if (n <= 0) return hipSuccess;
    scaled_axpy_kernel<<<(n + 255) / 256, 256>>>(p_out, z, p_old, beta_ptr, n);
    return (int)hipGetLastError();
}

extern "C" int bs_divide_scalars_launch(const float* num_ptr, const float* den_ptr, float* result_ptr) {
    divide_kernel<<<1, 64>>>(num_ptr, den_ptr, result_ptr);
    return (int)hipGetLastError();
}

extern "C" int bs_compute_b_launch(float* b, const float* div_u, const int* mat, int nx, int ny, int nz,
                                   float scale, int mat_fluid) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    compute_b_kernel<<<cell_grid(nx, ny, nz), 256>>>(b, div_u, mat, nx, ny, nz, scale, mat_fluid);
    return (int)hipGetLastError();
}

extern "C" int bs_zero_buffer_launch(float* buf, int n) {
    if (n <= 0) return hipSuccess;
    zero_kernel<<<(n + 255) / 256, 256>>>(buf, n);
    return (int)hipGetLastError();
}
