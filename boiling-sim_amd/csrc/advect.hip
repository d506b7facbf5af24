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
// (p - origin)/dx between the samples at one point and the points of a cell.
#include "boilsim_hip.h"
#include "common.hpp"
#include "grid.hpp"

#include <type_traits>

namespace bs {

struct AdvGeom {
    int nx, ny, nz;
    float ox, oy, oz, dx, dt;
    int mf;
};

// A sampled lattice: sizes, the upper clamps float(n-1) - 1e-6, and the row
// and plane strides in bytes.
struct Lattice {
    int n0, n1, n2;
    float h0, h1, h2;
    uint32_t row_b;    // n2 * 4
    uint64_t plane_b;  // n1 * n2 * 4
};

__host__ __device__ inline Lattice lattice(int n0, int n1, int n2) {
    Lattice L;
    L.n0 = n0; L.n1 = n1; L.n2 = n2;
    L.h0 = (float)(n0 - 1) - 1.0e-6f;
    L.h1 = (float)(n1 - 1) - 1.0e-6f;
    L.h2 = (float)(n2 - 1) - 1.0e-6f;
    L.row_b = (uint32_t)n2 * 4u;
    L.plane_b = (uint64_t)n1 * n2 * 4u;
    return L;
}

__device__ __forceinline__ float clampw(float x, float hi) { return fminf(fmaxf(0.0f, x), hi); }

// Byte offsets: 32-bit on the narrow path (the load is then the SGPR base
// plus one offset register, no 64-bit address arithmetic), 64-bit otherwise.
template <bool NARROW>
using Off = typename std::conditional<NARROW, uint32_t, uint64_t>::type;

template <bool NARROW>
__device__ __forceinline__ float at(const float* __restrict__ f, Off<NARROW> off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(f) + off);
}

// Byte offset of row i*n1 + j.  NARROW: every lattice has fewer than 2^24
// rows and rows of fewer than 2^22 floats, and every array fits in 4 GiB, so
// 24-bit multiplies and 32-bit byte offsets are exact.
template <bool NARROW>
__device__ __forceinline__ Off<NARROW> row_off(const Lattice& L, int i, int j) {
    if constexpr (NARROW) {
        const uint32_t row = __umul24((uint32_t)i, (uint32_t)L.n1) + (uint32_t)j;
        return __umul24(row, L.row_b);
    } else {
        return ((uint64_t)i * (uint64_t)L.n1 + (uint64_t)j) * (uint64_t)L.row_b;
    }
}

// Trilinear sample given q = (p - origin)/dx per axis (formed once per point
// by the caller) and the field's lattice offsets.
template <bool NARROW>
__device__ __forceinline__ float tri_q(const float* __restrict__ f, const Lattice& L, float qx, float qy, float qz,
                                       float offx, float offy, float offz) {
    using O = Off<NARROW>;
    const float fx = clampw(qx - offx, L.h0);
    const float fy = clampw(qy - offy, L.h1);
    const float fz = clampw(qz - offz, L.h2);
    const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    const float tx = fx - (float)i0;
    const float ty = fy - (float)j0;
    const float tz = fz - (float)k0;
    // steps to the +1 corners; 0 where the reference would read index n
    const O di = i0 + 1 < L.n0 ? (O)L.plane_b : (O)0;
    const O dj = j0 + 1 < L.n1 ? (O)L.row_b : (O)0;
    const O dk = k0 + 1 < L.n2 ? (O)4 : (O)0;
    const O r00 = row_off<NARROW>(L, i0, j0) + (O)4 * (O)(uint32_t)k0;
    const float c000 = at<NARROW>(f, r00), c100 = at<NARROW>(f, r00 + di);
    const float c010 = at<NARROW>(f, r00 + dj), c110 = at<NARROW>(f, r00 + di + dj);
    const float c001 = at<NARROW>(f, r00 + dk), c101 = at<NARROW>(f, r00 + di + dk);
    const float c011 = at<NARROW>(f, r00 + dj + dk), c111 = at<NARROW>(f, r00 + di + dj + dk);
    const float sx = 1.0f - tx, sy = 1.0f - ty, sz = 1.0f - tz;
    const float c00 = c000 * sx + c100 * tx;
    const float c10 = c010 * sx + c110 * tx;
    const float c01 = c001 * sx + c101 * tx;
    const float c11 = c011 * sx + c111 * tx;
    const float c0 = c00 * sy + c10 * ty;
    const float c1 = c01 * sy + c11 * ty;
    return c0 * sz + c1 * tz;
}

// One axis of a trilinear sample: f = clamp(q - off, [0, h]), i0 = int(f),
// t = f - float(i0), s = 1 - t, and whether index i0 + 1 is inside the
// lattice.  The velocity samples at a cell's four output points use 12
// distinct axis evaluations (of 36 uses), so they are formed once each.
struct Ax {
    int i0;
    bool more;
    float t, s;
};

__device__ __forceinline__ Ax axis(float q, float off, int n, float h) {
    Ax a;
    const float f = clampw(q - off, h);
    a.i0 = (int)f;
    a.t = f - (float)a.i0;
    a.s = 1.0f - a.t;
    a.more = a.i0 + 1 < n;
    return a;
}

// tri_q from three axis evaluations (the same float operations).
template <bool NARROW>
__device__ __forceinline__ float tri3(const float* __restrict__ f, const Lattice& L, const Ax& ax, const Ax& ay,
                                      const Ax& az) {
    using O = Off<NARROW>;
    const O di = ax.more ? (O)L.plane_b : (O)0;
    const O dj = ay.more ? (O)L.row_b : (O)0;
    const O dk = az.more ? (O)4 : (O)0;
    const O r00 = row_off<NARROW>(L, ax.i0, ay.i0) + (O)4 * (O)(uint32_t)az.i0;
    const float c000 = at<NARROW>(f, r00), c100 = at<NARROW>(f, r00 + di);
    const float c010 = at<NARROW>(f, r00 + dj), c110 = at<NARROW>(f, r00 + di + dj);
    const float c001 = at<NARROW>(f, r00 + dk), c101 = at<NARROW>(f, r00 + di + dk);
    const float c011 = at<NARROW>(f, r00 + dj + dk), c111 = at<NARROW>(f, r00 + di + dj + dk);
    const float c00 = c000 * ax.s + c100 * ax.t;
    const float c10 = c010 * ax.s + c110 * ax.t;
    const float c01 = c001 * ax.s + c101 * ax.t;
    const float c11 = c011 * ax.s + c111 * ax.t;
    const float c0 = c00 * ay.s + c10 * ay.t;
    const float c1 = c01 * ay.s + c11 * ay.t;
    return c0 * az.s + c1 * az.t;
}

struct Fields {
    const float* __restrict__ ux;
    const float* __restrict__ uy;
    const float* __restrict__ uz;
    Lattice Lx, Ly, Lz, Lc;
};

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
// Four consecutive k cells per thread; the neighbour classes are loaded
// together (not one after the other as the reference's early returns would
// suggest) and the first fluid one is selected.
template <int VEC>
__global__ __launch_bounds__(256) void extend_kernel(float* __restrict__ t_ext, const float* __restrict__ t,
                                                     const int* __restrict__ mat, int nx, int ny, int nz,
                                                     uint32_t chunks, FastDiv nzvd, int mf) {
    const uint32_t nblk = gridDim.x * gridDim.y;
    const uint32_t lb = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nblk);
    const uint32_t chunk = lb % chunks;
    const int i = (int)(lb / chunks);
    const uint32_t q = chunk * 256u + threadIdx.x;
    if (q >= (uint32_t)ny * nzvd.d) return;
    const int j = (int)fdiv(q, nzvd);
    const int k0 = (int)(q - (uint32_t)j * nzvd.d) * VEC;
    const int64_t plane = (int64_t)ny * nz;
    const int64_t c0 = (int64_t)i * plane + (int64_t)j * nz + k0;
    int m[VEC];
    float tv[VEC];
    ldi<VEC>(mat, c0, m);
    ldf<VEC>(t, c0, tv);
    bool any = false;
#pragma unroll
    for (int e = 0; e < VEC; ++e) any |= m[e] != mf;
    if (any) {
        const int nf = ~mf;
        int ml[VEC], mr[VEC], md[VEC], mu[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) ml[e] = mr[e] = md[e] = mu[e] = nf;
        if (i > 0) ldi<VEC>(mat, c0 - plane, ml);
        if (i < nx - 1) ldi<VEC>(mat, c0 + plane, mr);
        if (j > 0) ldi<VEC>(mat, c0 - nz, md);
        if (j < ny - 1) ldi<VEC>(mat, c0 + nz, mu);
        const int mb = k0 > 0 ? mat[c0 - 1] : nf;
        const int mt = k0 + VEC < nz ? mat[c0 + VEC] : nf;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if (m[e] == mf) continue;
            const int lo = e > 0 ? m[e - 1] : mb;
            const int hi = e < VEC - 1 ? m[e + 1] : mt;
            int64_t src = -1;
            if (ml[e] == mf) src = c0 + e - plane;
            else if (mr[e] == mf) src = c0 + e + plane;
            else if (md[e] == mf) src = c0 + e - nz;
            else if (mu[e] == mf) src = c0 + e + nz;
            else if (lo == mf) src = c0 + e - 1;
            else if (hi == mf) src = c0 + e + 1;
            if (src >= 0) tv[e] = t[src];
        }
    }
    stf<VEC>(t_ext, c0, tv);
}

// The departure sample: p_back = p - (u, v, w)*dt, then tri(field, p_back).
template <bool NARROW>
__device__ __forceinline__ float depart(const AdvGeom& G, const float* __restrict__ field, const Lattice& Lf,
                                        float px, float py, float pz, float u, float v, float w, float offx,
                                        float offy, float offz) {
    const float bx = px - u * G.dt, by = py - v * G.dt, bz = pz - w * G.dt;
    const float rx = (bx - G.ox) / G.dx, ry = (by - G.oy) / G.dx, rz = (bz - G.oz) / G.dx;
    return tri_q<NARROW>(field, Lf, rx, ry, rz, offx, offy, offz);
}

// advect_u{x,y,z} and advect_temperature for one cell's outputs, each traced
// only where the reference traces it.  Output points: X = (i, j+.5, k+.5)
// for the x-face, Y = (i+.5, j, k+.5), Z = (i+.5, j+.5, k), C = the centre.
template <bool NARROW>
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
    const uint32_t fy = ((uint32_t)i * (ny + 1) + j) * nz + k;  // y-face j in uy (nx,ny+1,nz)
    const uint32_t fz = ((uint32_t)i * ny + j) * (nz + 1) + k;  // z-face k in uz (nx,ny,nz+1)
    const bool fl = mat[c] == mf;
    const bool do_x = i > 0 && (fl || mat[c - plane] == mf);
    const bool do_y = j > 0 && (fl || mat[c - nz] == mf);
    const bool do_z = k > 0 && (fl || mat[c - 1] == mf);
    float vx = 0.0f, vy = 0.0f, vz = 0.0f, vt = 0.0f;
    if (do_x || do_y || do_z || fl) {
        // p = origin + a*dx and q = (p - origin)/dx per axis, for a = index
        // (face) and index + 0.5 (centre)
        const float fi = (float)i, fj = (float)j, fk = (float)k;
        const float pxi = G.ox + fi * G.dx, pxh = G.ox + (fi + 0.5f) * G.dx;
        const float pyi = G.oy + fj * G.dx, pyh = G.oy + (fj + 0.5f) * G.dx;
        const float pzi = G.oz + fk * G.dx, pzh = G.oz + (fk + 0.5f) * G.dx;
        const float qxi = (pxi - G.ox) / G.dx, qxh = (pxh - G.ox) / G.dx;
        const float qyi = (pyi - G.oy) / G.dx, qyh = (pyh - G.oy) / G.dx;
        const float qzi = (pzi - G.oz) / G.dx, qzh = (pzh - G.oz) / G.dx;
        // axis evaluations: [index or half][field lattice on that axis]
        const Ax x_i_u = axis(qxi, 0.0f, F.Lx.n0, F.Lx.h0), x_h_u = axis(qxh, 0.0f, F.Lx.n0, F.Lx.h0);
        const Ax x_i_c = axis(qxi, 0.5f, F.Ly.n0, F.Ly.h0), x_h_c = axis(qxh, 0.5f, F.Ly.n0, F.Ly.h0);
        const Ax y_i_v = axis(qyi, 0.0f, F.Ly.n1, F.Ly.h1), y_h_v = axis(qyh, 0.0f, F.Ly.n1, F.Ly.h1);
        const Ax y_i_c = axis(qyi, 0.5f, F.Lx.n1, F.Lx.h1), y_h_c = axis(qyh, 0.5f, F.Lx.n1, F.Lx.h1);
        const Ax z_i_w = axis(qzi, 0.0f, F.Lz.n2, F.Lz.h2), z_h_w = axis(qzh, 0.0f, F.Lz.n2, F.Lz.h2);
        const Ax z_i_c = axis(qzi, 0.5f, F.Lx.n2, F.Lx.h2), z_h_c = axis(qzh, 0.5f, F.Lx.n2, F.Lx.h2);
        if (do_x) {
            const float u = tri3<NARROW>(F.ux, F.Lx, x_i_u, y_h_c, z_h_c);
            const float v = tri3<NARROW>(F.uy, F.Ly, x_i_c, y_h_v, z_h_c);
            const float w = tri3<NARROW>(F.uz, F.Lz, x_i_c, y_h_c, z_h_w);
            vx = depart<NARROW>(G, F.ux, F.Lx, pxi, pyh, pzh, u, v, w, 0.0f, 0.5f, 0.5f);
        }
        if (do_y) {
            const float u = tri3<NARROW>(F.ux, F.Lx, x_h_u, y_i_c, z_h_c);
            const float v = tri3<NARROW>(F.uy, F.Ly, x_h_c, y_i_v, z_h_c);
            const float w = tri3<NARROW>(F.uz, F.Lz, x_h_c, y_i_c, z_h_w);
            vy = depart<NARROW>(G, F.uy, F.Ly, pxh, pyi, pzh, u, v, w, 0.5f, 0.0f, 0.5f);
        }
        if (do_z) {
            const float u = tri3<NARROW>(F.ux, F.Lx, x_h_u, y_h_c, z_i_c);
            const float v = tri3<NARROW>(F.uy, F.Ly, x_h_c, y_h_v, z_i_c);
            const float w = tri3<NARROW>(F.uz, F.Lz, x_h_c, y_h_c, z_i_w);
            vz = depart<NARROW>(G, F.uz, F.Lz, pxh, pyh, pzi, u, v, w, 0.5f, 0.5f, 0.0f);
        }
        if (fl) {
            const float u = tri3<NARROW>(F.ux, F.Lx, x_h_u, y_h_c, z_h_c);
            const float v = tri3<NARROW>(F.uy, F.Ly, x_h_c, y_h_v, z_h_c);
            const float w = tri3<NARROW>(F.uz, F.Lz, x_h_c, y_h_c, z_h_w);
            vt = depart<NARROW>(G, t_ext, F.Lc, pxh, pyh, pzh, u, v, w, 0.5f, 0.5f, 0.5f);
        }
    }
    // x-face i (between cells i-1 and i); its index in ux (nx+1,ny,nz) is c
    ux_new[c] = do_x ? vx : (i > 0 ? F.ux[c] : 0.0f);
    if (i == nx - 1) ux_new[c + plane] = 0.0f;
    uy_new[fy] = do_y ? vy : (j > 0 ? F.uy[fy] : 0.0f);
    if (j == ny - 1) uy_new[fy + nz] = 0.0f;
    uz_new[fz] = do_z ? vz : (k > 0 ? F.uz[fz] : 0.0f);
    if (k == nz - 1) uz_new[fz + 1] = 0.0f;
    // the cell: T on fluid cells from the extended field, others keep T
    t_new[c] = fl ? vt : t[c];
}

struct CellGrid {
    dim3 grid;
    uint32_t chunks;
    FastDiv nzd;
};

static CellGrid cell_grid(int nx, int ny, int nz, int vec) {
    CellGrid g;
    const uint32_t nzv = (uint32_t)(nz / vec);
    g.nzd = make_fastdiv(nzv);
    g.chunks = (uint32_t)(((uint64_t)ny * nzv + 255) / 256);
    g.grid = dim3(g.chunks, (unsigned)nx, 1);
    return g;
}

// The face arrays are the largest: (nx+1)*ny*nz etc. must index in 32 bits.
static bool fits32(int nx, int ny, int nz) {
    const uint64_t m = (uint64_t)(nx + 1) * (ny + 1) * (nz + 1);
    return m < (1ull << 31);
}

// The 24-bit / 32-bit-offset sampler: every lattice below 2^24 rows, rows
// below 2^22 floats, every array below 4 GiB.
static bool narrow(int nx, int ny, int nz) {
    const uint64_t rows = (uint64_t)(nx + 1) * (ny + 1);
    const uint64_t bytes = 4ull * (nx + 1) * (ny + 1) * (nz + 1);
    return rows < (1ull << 24) && (uint64_t)(nz + 1) < (1ull << 22) && bytes < (1ull << 32);
}

}  // namespace bs

using namespace bs;

// extend_temperature_into_solids (fluid.py:338-378).
extern "C" int bs_extend_temperature_launch(float* t_ext, const float* t, const int* mat, int nx, int ny, int nz,
                                            int mat_fluid, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (!fits32(nx, ny, nz)) return (int)hipErrorInvalidValue;
    const int vec = vec_for(nz, {t_ext, t, mat});
    const CellGrid g = cell_grid(nx, ny, nz, vec);
    switch (vec) {
        case 4: extend_kernel<4><<<g.grid, 256, 0, st>>>(t_ext, t, mat, nx, ny, nz, g.chunks, g.nzd, mat_fluid); break;
        case 2: extend_kernel<2><<<g.grid, 256, 0, st>>>(t_ext, t, mat, nx, ny, nz, g.chunks, g.nzd, mat_fluid); break;
        default: extend_kernel<1><<<g.grid, 256, 0, st>>>(t_ext, t, mat, nx, ny, nz, g.chunks, g.nzd, mat_fluid); break;
    }
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
    const CellGrid g = cell_grid(nx, ny, nz, 1);
    if (narrow(nx, ny, nz))
        advect_kernel<true><<<g.grid, 256, 0, st>>>(ux_new, uy_new, uz_new, t_new, F, t, t_ext, mat, G, g.chunks, g.nzd);
    else
        advect_kernel<false><<<g.grid, 256, 0, st>>>(ux_new, uy_new, uz_new, t_new, F, t, t_ext, mat, G, g.chunks, g.nzd);
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
