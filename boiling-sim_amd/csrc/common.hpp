// Shared device helpers for the gfx950 pressure-projection kernels.
//
// Layout contract (identical to the reference's Warp arrays,
// python/boilingsim/geometry.py:254-282): every cell field is C-contiguous
// (nx, ny, nz) with k fastest, cell (i,j,k) at (i*ny + j)*nz + k; the face
// fields are ux (nx+1,ny,nz), uy (nx,ny+1,nz), uz (nx,ny,nz+1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bs {

// Compute units of the current device (256 on the MI355X if the runtime
// cannot say).  The i-marching passes size their grids to whole rounds of
// workgroups per CU: a grid just past a round leaves the device idle but for
// a few workgroups while that last round runs.
inline int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 256;
    return cus;
}

// ---------------------------------------------------------------------------
// Cell classes.  The pressure path only distinguishes fluid (interior
// unknown), air (Dirichlet p = 0) and anything else (solid, Neumann: the
// ghost value is the cell's own pressure) — fluid.py:222-240.
// ---------------------------------------------------------------------------

// Sentinel encodings used by the Jacobi sweeps to carry the class of a
// non-fluid cell inside the pressure array itself, so a sweep streams only
// p_in, div and p_out (12 B/cell) and never reads the int32 mask.  Both are
// quiet NaNs with payloads no arithmetic result produces; they are only ever
// compared bitwise, never fed to arithmetic.
constexpr uint32_t kSentAir = 0xFFC0A11Au;
constexpr uint32_t kSentSolid = 0xFFC05D1Du;

// One-byte class code used by the PCG kernels:
//   0 = air, 1 = solid, 2 + n_solid = fluid (n_solid = solid neighbours, 0..6).
// The Jacobi diagonal of a fluid row is 6 - n_solid = 8 - code
// (diag_inverse_apply.cu:63-82); code 8 means diag 0, where the reference
// applies z = r (diag_inverse_apply.cu:84-88).
constexpr uint8_t kCodeAir = 0;
constexpr uint8_t kCodeSolid = 1;
constexpr uint8_t kCodeFluid0 = 2;

// ---------------------------------------------------------------------------
// Launch geometry
// ---------------------------------------------------------------------------

constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

// Maps the hardware dispatch order (round-robin over the 8 XCDs) to a logical
// block id such that each XCD works on one contiguous range of logical
// blocks, which keeps i/j neighbour planes inside that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t hw_id, uint32_t n_blocks) {
    const uint32_t per = n_blocks / kNumXcd;
    const uint32_t full = per * kNumXcd;
    if (hw_id >= full) return hw_id;  // ragged tail keeps its own id
    return (hw_id % kNumXcd) * per + hw_id / kNumXcd;
}

// Unsigned division by a runtime-constant divisor d >= 1 via a 32-bit
// multiply-high (Granlund–Montgomery round-up method): q = (mulhi(n, m) + n) >> s
// for n < 2^31.
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;  // s = ceil(log2 d)
    f.s = s;
    // m = floor(2^32 * (2^s - d) / d) + 1
    f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
    if (d == 1) { f.m = 0; f.s = 0; }
    return f;
}

__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
    return (t + n) >> f.s;
}

__device__ __forceinline__ bool bits_eq(float v, uint32_t pattern) {
    return __float_as_uint(v) == pattern;
}

}  // namespace bs

#define BS_RETURN_IF_ERR(expr)                         \
    do {                                               \
        hipError_t bs_err_ = (expr);                   \
        if (bs_err_ != hipSuccess) return (int)bs_err_; \
    } while (0)
