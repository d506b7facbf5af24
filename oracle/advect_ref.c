/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the reference's semi-Lagrangian advection
 * (python/boilingsim/fluid.py:33-103 _tri_sample / _sample_velocity_at,
 * 338-492 extend_temperature_into_solids / advect_temperature /
 * advect_u{x,y,z}, 737-775 advect_all).  Built with -ffp-contract=off: every
 * float32 operation is rounded on its own, in the order the Warp kernels
 * write them (the reference's bit-exact gate compiles its native kernels
 * with --fmad=false for the same reason, test_pressure_parity.py:1-20).
 *
 * Two points the Warp code leaves to the hardware are fixed here:
 *  - wp.clamp(x, a, b) is min(max(a, x), b) with IEEE maxNum/minNum, so a NaN
 *    coordinate clamps to 0;
 *  - the upper clamp float(n-1) - 1e-6 rounds to n-1 once n-1 >= 64, and the
 *    sampler then reads index n (one past the array) with weight tx = 0; the
 *    restatement reads index n-1 instead (the weight is 0 either way).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    float x, y, z;
} v3;

static inline float clampw(float x, float a, float b) { return fminf(fmaxf(a, x), b); }

/* _tri_sample (fluid.py:33-80): trilinear sample of an (n0, n1, n2) field at
 * world point p on the lattice offset by (ox, oy, oz) cells. */
static float tri_sample(const float* f, int n0, int n1, int n2, v3 p, v3 o, float dx, float ox, float oy,
                        float oz) {
    float fx = (p.x - o.x) / dx - ox;
    float fy = (p.y - o.y) / dx - oy;
    float fz = (p.z - o.z) / dx - oz;
    fx = clampw(fx, 0.0f, (float)(n0 - 1) - 1.0e-6f);
    fy = clampw(fy, 0.0f, (float)(n1 - 1) - 1.0e-6f);
    fz = clampw(fz, 0.0f, (float)(n2 - 1) - 1.0e-6f);
    const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    const float tx = fx - (float)i0;
    const float ty = fy - (float)j0;
    const float tz = fz - (float)k0;
    const int i1 = i0 + 1 < n0 ? i0 + 1 : n0 - 1;
    const int j1 = j0 + 1 < n1 ? j0 + 1 : n1 - 1;
    const int k1 = k0 + 1 < n2 ? k0 + 1 : n2 - 1;
#define F(a, b, c) f[((int64_t)(a) * n1 + (b)) * n2 + (c)]
    const float c000 = F(i0, j0, k0), c100 = F(i1, j0, k0), c010 = F(i0, j1, k0), c110 = F(i1, j1, k0);
    const float c001 = F(i0, j0, k1), c101 = F(i1, j0, k1), c011 = F(i0, j1, k1), c111 = F(i1, j1, k1);
#undef F
    const float c00 = c000 * (1.0f - tx) + c100 * tx;
    const float c10 = c010 * (1.0f - tx) + c110 * tx;
    const float c01 = c001 * (1.0f - tx) + c101 * tx;
    const float c11 = c011 * (1.0f - tx) + c111 * tx;
    const float c0 = c00 * (1.0f - ty) + c10 * ty;
    const float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}

typedef struct {
    const float *ux, *uy, *uz;
    int nx, ny, nz;
} Vel;

/* _sample_velocity_at (fluid.py:83-97). */
static v3 sample_velocity(const Vel* V, v3 p, v3 o, float dx) {
    v3 u;
    u.x = tri_sample(V->ux, V->nx + 1, V->ny, V->nz, p, o, dx, 0.0f, 0.5f, 0.5f);
    u.y = tri_sample(V->uy, V->nx, V->ny + 1, V->nz, p, o, dx, 0.5f, 0.0f, 0.5f);
    u.z = tri_sample(V->uz, V->nx, V->ny, V->nz + 1, p, o, dx, 0.5f, 0.5f, 0.0f);
    return u;
}

/* origin + vec3(a, b, c) * dx, then p - u * dt, component by component. */
static inline v3 point(v3 o, float a, float b, float c, float dx) {
    v3 p = {o.x + a * dx, o.y + b * dx, o.z + c * dx};
    return p;
}

static inline v3 back(v3 p, v3 u, float dt) {
    v3 q = {p.x - u.x * dt, p.y - u.y * dt, p.z - u.z * dt};
    return q;
}

#define CIDX(i, j, k) ((((int64_t)(i)) * ny + (j)) * (int64_t)nz + (k))

/* extend_temperature_into_solids (fluid.py:338-378). */
void orc_extend_temperature(float* t_ext, const float* t, const int32_t* mat, int nx, int ny, int nz,
                            int mat_fluid) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                float v = t[c];
                if (mat[c] != mat_fluid) {
                    if (i > 0 && mat[CIDX(i - 1, j, k)] == mat_fluid) v = t[CIDX(i - 1, j, k)];
                    else if (i < nx - 1 && mat[CIDX(i + 1, j, k)] == mat_fluid) v = t[CIDX(i + 1, j, k)];
                    else if (j > 0 && mat[CIDX(i, j - 1, k)] == mat_fluid) v = t[CIDX(i, j - 1, k)];
                    else if (j < ny - 1 && mat[CIDX(i, j + 1, k)] == mat_fluid) v = t[CIDX(i, j + 1, k)];
                    else if (k > 0 && mat[CIDX(i, j, k - 1)] == mat_fluid) v = t[CIDX(i, j, k - 1)];
                    else if (k < nz - 1 && mat[CIDX(i, j, k + 1)] == mat_fluid) v = t[CIDX(i, j, k + 1)];
                }
                t_ext[c] = v;
            }
}

/* advect_u{x,y,z} (fluid.py:404-492) and advect_temperature (381-401) into
 * the *_new arrays; the inputs are not modified. */
void orc_advect(float* ux_new, float* uy_new, float* uz_new, float* t_new, const float* ux, const float* uy,
                const float* uz, const float* t, const float* t_ext, const int32_t* mat, int nx, int ny, int nz,
                float ox, float oy, float oz, float dx, float dt, int mat_fluid) {
    const v3 o = {ox, oy, oz};
    const Vel V = {ux, uy, uz, nx, ny, nz};
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t f = ((int64_t)i * ny + j) * nz + k;
                if (i == 0 || i == nx) { ux_new[f] = 0.0f; continue; }
                if (mat[CIDX(i - 1, j, k)] != mat_fluid && mat[CIDX(i, j, k)] != mat_fluid) {
                    ux_new[f] = ux[f];
                    continue;
                }
                const v3 p = point(o, (float)i, (float)j + 0.5f, (float)k + 0.5f, dx);
                const v3 pb = back(p, sample_velocity(&V, p, o, dx), dt);
                ux_new[f] = tri_sample(ux, nx + 1, ny, nz, pb, o, dx, 0.0f, 0.5f, 0.5f);
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t f = ((int64_t)i * (ny + 1) + j) * nz + k;
                if (j == 0 || j == ny) { uy_new[f] = 0.0f; continue; }
                if (mat[CIDX(i, j - 1, k)] != mat_fluid && mat[CIDX(i, j, k)] != mat_fluid) {
                    uy_new[f] = uy[f];
                    continue;
                }
                const v3 p = point(o, (float)i + 0.5f, (float)j, (float)k + 0.5f, dx);
                const v3 pb = back(p, sample_velocity(&V, p, o, dx), dt);
                uy_new[f] = tri_sample(uy, nx, ny + 1, nz, pb, o, dx, 0.5f, 0.0f, 0.5f);
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k <= nz; ++k) {
                const int64_t f = ((int64_t)i * ny + j) * (nz + 1) + k;
                if (k == 0 || k == nz) { uz_new[f] = 0.0f; continue; }
                if (mat[CIDX(i, j, k - 1)] != mat_fluid && mat[CIDX(i, j, k)] != mat_fluid) {
                    uz_new[f] = uz[f];
                    continue;
                }
                const v3 p = point(o, (float)i + 0.5f, (float)j + 0.5f, (float)k, dx);
                const v3 pb = back(p, sample_velocity(&V, p, o, dx), dt);
                uz_new[f] = tri_sample(uz, nx, ny, nz + 1, pb, o, dx, 0.5f, 0.5f, 0.0f);
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mat_fluid) { t_new[c] = t[c]; continue; }
                const v3 p = point(o, (float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f, dx);
                const v3 pb = back(p, sample_velocity(&V, p, o, dx), dt);
                t_new[c] = tri_sample(t_ext, nx, ny, nz, pb, o, dx, 0.5f, 0.5f, 0.5f);
            }
}
