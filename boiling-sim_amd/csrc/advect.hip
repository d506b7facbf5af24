// Semi-Lagrangian advection of the MAC velocity and the cell temperature on
// gfx950: advect_all (python/boilingsim/fluid.py:737-775) with its kernels
// extend_temperature_into_solids (338-378), advect_temperature (381-401),
// advect_u{x,y,z} (404-492) and the sampler _tri_sample /
// _sample_velocity_at (33-97).
//
// Every float operation is the reference's, in its order, rounded on its own
// (-ffp-contract=off, IEEE division):
//   p      = origin + (a, b, c)*dx           (a, b, c: index or index + 0.5)
//   u      = (tri(ux, p; 0,.5,.5), tri(uy, p; .5,0,.5), tri(uz, p; .5,.5,0))
//   p_back = p - u*dt
//   new    = tri(field, p_back; field's offsets)
//   tri:   f = (p - origin)/dx - off, clamped to [0, float(n-1) - 1e-6],
//          i0 = int(f), t = f - float(i0), lerps along x, then y, then z.
// Two points the Warp code leaves to the hardware are fixed as the oracle
// fixes them: the clamp is min(max(0, f), hi) with IEEE maxNum/minNum (NaN
// -> 0), and where hi rounds to n-1 (n-1 >= 64) the index i0+1 = n the
// reference reads one past the array with weight 0 is clamped to n-1.
//
// One thread per cell writes that cell's low x-, y- and z-face and the cell
// itself (and the edge faces ux[nx], uy[:,ny], uz[:,:,nz] on the high
// edges), so one launch replaces the reference's four.  Each output gathers
// 32 values (three velocity components and the advected field, 8 corners
// each) around points less than a few cells away, so the gathers hit L1/L2
// and HBM sees each input about once; the division count is cut by sharing
// (p - origin)/dx between the three velocity samples at one point.
#include "boilsim_hip.h"
#include "common.hpp"

namespace bs {

struct AdvGeom {
    int nx, ny, nz;
    float ox, oy, oz, dx, dt;
    int mf;
};

struct Lattice {
    int n0, n1, n2;
    float h0, h1, h2;  // float(n-1) - 1e-6 per axis
};

__host__ __device__ inline Lattice lattice(int n0, int n1, int n2) {
    Lattice L;
    L.n0 = n0; L.n1 = n1; L.n2 = n2;
    L.h0 = (float)(n0 - 1) - 1.0e-6f;
    L.h1 = (float)(n1 - 1) - 1.0e-6f;
    L.h2 = (float)(n2 - 1) - 1.0e-6f;
    return L;
}

__device__ __forceinline__ float clampw(float x, float hi) { return fminf(fmaxf(0.0f, x), hi); }

// Trilinear sample given q = (p - origin)/dx per axis (the caller forms it
// once per point) and the field's lattice offsets.
__device__ __forceinline__ float tri_q(const float* __restrict__ f, const Lattice& L, float qx, float qy, float qz,
                                       float offx, float offy, float offz) {
    const float fx = clampw(qx - offx, L.h0);
    const float fy = clampw(qy - offy, L.h1);
    const float fz = clampw(qz - offz, L.h2);
    const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    const float tx = fx - (float)i0;
    const float ty = fy - (float)j0;
    const float tz = fz - (float)k0;
    const int i1 = i0 + 1 < L.n0 ? i0 + 1 : L.n0 - 1;
    const int j1 = j0 + 1 < L.n1 ? j0 + 1 : L.n1 - 1;
    const int k1 = k0 + 1 < L.n2 ? k0 + 1 : L.n2 - 1;
    const uint32_t r00 = ((uint32_t)i0 * L.n1 + j0) * L.n2, r10 = ((uint32_t)i1 * L.n1 + j0) * L.n2;
    const uint32_t r01 = ((uint32_t)i0 * L.n1 + j1) * L.n2, r11 = ((uint32_t)i1 * L.n1 + j1) * L.n2;
    const float c000 = f[r00 + k0], c100 = f[r10 + k0], c010 = f[r01 + k0], c110 = f[r11 + k0];
    const float c001 = f[r00 + k1], c101 = f[r10 + k1], c011 = f[r01 + k1], c111 = f[r11 + k1];
    const float sx = 1.0f - tx, sy = 1.0f - ty, sz = 1.0f - tz;
    const float c00 = c000 * sx + c100 * tx;
    const float c10 = c010 * sx + c110 * tx;
    const float c01 = c001 * sx + c101 * tx;
    const float c11 = c011 * sx + c111 * tx;
    const float c0 = c00 * sy + c10 * ty;
    const float c1 = c01 * sy + c11 * ty;
    return c0 * sz + c1 * tz;
}

struct Fields {
    const float* __restrict__ ux;
    const float* __restrict__ uy;
    const float* __restrict__ uz;
    Lattice Lx, Ly, Lz, Lc;
};

// The value at the output point (a, b, c) (index units, before *dx) of
// `field` advected along the velocity: new = tri(field, p - u(p)*dt).
__device__ __forceinline__ float advect_point(const Fields& F, const AdvGeom& G, const float* __restrict__ field,
                                              const Lattice& Lf, float a, float b, float c, float offx, float offy,
                                              float offz) {
    const float px = G.ox + a * G.dx, py = G.oy + b * G.dx, pz = G.oz + c * G.dx;
    const float qx = (px - G.ox) / G.dx, qy = (py - G.oy) / G.dx, qz = (pz - G.oz) / G.dx;
    const float u = tri_q(F.ux, F.Lx, qx, qy, qz, 0.0f, 0.5f, 0.5f);
    const float v = tri_q(F.uy, F.Ly, qx, qy, qz, 0.5f, 0.0f, 0.5f);
    const float w = tri_q(F.uz, F.Lz, qx, qy, qz, 0.5f, 0.5f, 0.0f);
    const float bx = px - u * G.dt, by = py - v * G.dt, bz = pz - w * G.dt;
    const float rx = (bx - G.ox) / G.dx, ry = (by - G.oy) / G.dx, rz = (bz - G.oz) / G.dx;
    return tri_q(field, Lf, rx, ry, rz, offx, offy, offz);
}

__device__ __forceinline__ bool cell_coords(uint32_t chunks, FastDiv nzd, int ny, int& i, int& j, int& k) {
    const uint32_t nblk = gridDim.x * gridDim.y;
    const uint32_t lb = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nblk);
    const uint32_t chunk = lb % chunks;
    i = (int)(lb / chunks);
    const uint32_t q = chunk * 256u + threadIdx.x;
    if (q >= (uint32_t)ny * nzd.d) return false;
    j = (int)fdiv(q, nzd);
    k = (int)(q - (uint32_t)j * nzd.d);
    return true;
}

// extend_temperature_into_solids (fluid.py:338-378): a non-fluid cell takes
// T of its first fluid neighbour in the order i-1, i+1, j-1, j+1, k-1, k+1.
__global__ __launch_bounds__(256) void extend_kernel(float* __restrict__ t_ext, const float* __restrict__ t,
                                                     const int* __restrict__ mat, int nx, int ny, int nz,
                                                     uint32_t chunks, FastDiv nzd, int mf) {
    int i, j, k;
    if (!cell_coords(chunks, nzd, ny, i, j, k)) return;
    const uint32_t plane = (uint32_t)ny * nz;
    const uint32_t c = (uint32_t)i * plane + (uint32_t)j * nz + k;
    uint32_t src = c;
    if (mat[c] != mf) {
        if (i > 0 && mat[c - plane] == mf) src = c - plane;
        else if (i < nx - 1 && mat[c + plane] == mf) src = c + plane;
        else if (j > 0 && mat[c - nz] == mf) src = c - nz;
        else if (j < ny - 1 && mat[c + nz] == mf) src = c + nz;
        else if (k > 0 && mat[c - 1] == mf) src = c - 1;
        else if (k < nz - 1 && mat[c + 1] == mf) src = c + 1;
    }
    t_ext[c] = t[src];
}

// advect_u{x,y,z} and advect_temperature for one cell's outputs.
__global__ __launch_bounds__(256) void advect_kernel(float* __restrict__ ux_new, float* __restrict__ uy_new,
                                                     float* __restrict__ uz_new, float* __restrict__ t_new,
                                                     Fields F, const float* __restrict__ t,
                                                     const float* __restrict__ t_ext, const int* __restrict__ mat,
                                                     AdvGeom G, uint32_t chunks, FastDiv nzd) {
    int i, j, k;
    if (!cell_coords(chunks, nzd, G.ny, i, j, k)) return;
    const int nx = G.nx, ny = G.ny, nz = G.nz, mf = G.mf;
    const uint32_t plane = (uint32_t)ny * nz;
    const uint32_t c = (uint32_t)i * plane + (uint32_t)j * nz + k;
    const bool fl = mat[c] == mf;
    const float fi = (float)i, fj = (float)j, fk = (float)k;
    const float hi_ = fi + 0.5f, hj = fj + 0.5f, hk = fk + 0.5f;
    // x-face i (between cells i-1 and i); face index = c in ux (nx+1,ny,nz)
    {
        float v = 0.0f;
        if (i > 0) {
            if (!fl && mat[c - plane] != mf) v = F.ux[c];
            else v = advect_point(F, G, F.ux, F.Lx, fi, hj, hk, 0.0f, 0.5f, 0.5f);
        }
        ux_new[c] = v;
        if (i == nx - 1) ux_new[c + plane] = 0.0f;
    }
    // y-face j; uy is (nx,ny+1,nz)
    {
        const uint32_t f = ((uint32_t)i * (ny + 1) + j) * nz + k;
        float v = 0.0f;
        if (j > 0) {
            if (!fl && mat[c - nz] != mf) v = F.uy[f];
            else v = advect_point(F, G, F.uy, F.Ly, hi_, fj, hk, 0.5f, 0.0f, 0.5f);
        }
        uy_new[f] = v;
        if (j == ny - 1) uy_new[f + nz] = 0.0f;
    }
    // z-face k; uz is (nx,ny,nz+1)
    {
        const uint32_t f = ((uint32_t)i * ny + j) * (nz + 1) + k;
        float v = 0.0f;
        if (k > 0) {
            if (!fl && mat[c - 1] != mf) v = F.uz[f];
            else v = advect_point(F, G, F.uz, F.Lz, hi_, hj, fk, 0.5f, 0.5f, 0.0f);
        }
        uz_new[f] = v;
        if (k == nz - 1) uz_new[f + 1] = 0.0f;
    }
    // the cell: T on fluid cells from the extended field, others keep T
    t_new[c] = fl ? advect_point(F, G, t_ext, F.Lc, hi_, hj, hk, 0.5f, 0.5f, 0.5f) : t[c];
}

struct CellGrid {
    dim3 grid;
    uint32_t chunks;
    FastDiv nzd;
};

static CellGrid cell_grid(int nx, int ny, int nz) {
    CellGrid g;
    g.nzd = make_fastdiv((uint32_t)nz);
    g.chunks = (uint32_t)(((uint64_t)ny * nz + 255) / 256);
    g.grid = dim3(g.chunks, (unsigned)nx, 1);
    return g;
}

// The face arrays are the largest: (nx+1)*ny*nz etc. must index in 32 bits.
static bool fits32(int nx, int ny, int nz) {
    const uint64_t m = (uint64_t)(nx + 1) * (ny + 1) * (nz + 1);
    return m < (1ull << 31);
}

}  // namespace bs

using namespace bs;

// extend_temperature_into_solids (fluid.py:338-378).
extern "C" int bs_extend_temperature_launch(float* t_ext, const float* t, const int* mat, int nx, int ny, int nz,
                                            int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (!fits32(nx, ny, nz)) return (int)hipErrorInvalidValue;
    const CellGrid g = cell_grid(nx, ny, nz);
    extend_kernel<<<g.grid, 256, 0, st>>>(t_ext, t, mat, nx, ny, nz, g.chunks, g.nzd, mat_fluid);
    return (int)hipGetLastError();
}

// advect_ux, advect_uy, advect_uz and advect_temperature (fluid.py:381-492)
// in one launch, into the *_new arrays (which must not overlap the inputs).
extern "C" int bs_advect_launch(float* ux_new, float* uy_new, float* uz_new, float* t_new, const float* ux,
                                const float* uy, const float* uz, const float* t, const float* t_ext, const int* mat,
                                int nx, int ny, int nz, float ox, float oy, float oz, float dx, float dt,
                                int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (!fits32(nx, ny, nz)) return (int)hipErrorInvalidValue;
    Fields F;
    F.ux = ux; F.uy = uy; F.uz = uz;
    F.Lx = lattice(nx + 1, ny, nz);
    F.Ly = lattice(nx, ny + 1, nz);
    F.Lz = lattice(nx, ny, nz + 1);
    F.Lc = lattice(nx, ny, nz);
    const AdvGeom G{nx, ny, nz, ox, oy, oz, dx, dt, mat_fluid};
    const CellGrid g = cell_grid(nx, ny, nz);
    advect_kernel<<<g.grid, 256, 0, st>>>(ux_new, uy_new, uz_new, t_new, F, t, t_ext, mat, G, g.chunks, g.nzd);
    return (int)hipGetLastError();
}

// advect_all (fluid.py:737-775): T_ext, the four advections into the
// workspace arrays, then the results copied back into ux, uy, uz, t.
extern "C" int bs_advect_all_launch(float* ux, float* uy, float* uz, float* t, float* ux_tmp, float* uy_tmp,
                                    float* uz_tmp, float* t_tmp, float* t_ext, const int* mat, int nx, int ny, int nz,
                                    float ox, float oy, float oz, float dx, float dt, int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    int rc = bs_extend_temperature_launch(t_ext, t, mat, nx, ny, nz, mat_fluid, st);
    if (rc) return rc;
    rc = bs_advect_launch(ux_tmp, uy_tmp, uz_tmp, t_tmp, ux, uy, uz, t, t_ext, mat, nx, ny, nz, ox, oy, oz, dx, dt,
                          mat_fluid, st);
    if (rc) return rc;
    const size_t c = sizeof(float);
    BS_RETURN_IF_ERR(hipMemcpyAsync(ux, ux_tmp, c * (size_t)(nx + 1) * ny * nz, hipMemcpyDeviceToDevice, st));
    BS_RETURN_IF_ERR(hipMemcpyAsync(uy, uy_tmp, c * (size_t)nx * (ny + 1) * nz, hipMemcpyDeviceToDevice, st));
    BS_RETURN_IF_ERR(hipMemcpyAsync(uz, uz_tmp, c * (size_t)nx * ny * (nz + 1), hipMemcpyDeviceToDevice, st));
    BS_RETURN_IF_ERR(hipMemcpyAsync(t, t_tmp, c * (size_t)nx * ny * nz, hipMemcpyDeviceToDevice, st));
    return (int)hipGetLastError();
}
