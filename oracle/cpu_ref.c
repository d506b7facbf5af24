/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of boiling-sim's pressure-projection arithmetic, used as the
 * checker for the HIP library and as the CPU baseline timed by bench.py.
 * Nothing in the product path may link or call this file; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Built with -O3 -ffp-contract=off (see oracle/Makefile) so that every
 * float32 operation is rounded separately, in the order the reference's
 * Warp kernels write it.  All arrays are C-contiguous with k fastest:
 *   cell (i,j,k)        -> (i*ny + j)*nz + k
 *   ux   (nx+1,ny,nz)   -> (i*ny + j)*nz + k
 *   uy   (nx,ny+1,nz)   -> (i*(ny+1) + j)*nz + k
 *   uz   (nx,ny,nz+1)   -> (i*ny + j)*(nz+1) + k
 * Out-of-bounds neighbours count as air, as in the reference's native
 * kernels (crates/cuda-kernels/src/jacobi_pressure.cu:121-124,
 * diag_inverse_apply.cu:31-39).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CIDX(i, j, k) ((((int64_t)(i)) * ny + (j)) * (int64_t)nz + (k))

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* compute_divergence — python/boilingsim/fluid.py:201-219.
 * Three divides, then (dudx + dvdy) + dwdz; non-fluid cells get 0. */
void orc_divergence(float* div, const float* ux, const float* uy, const float* uz,
                    const int32_t* mat, int nx, int ny, int nz, float dx, int mat_fluid) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mat_fluid) { div[c] = 0.0f; continue; }
                const float ux0 = ux[((int64_t)i * ny + j) * nz + k];
                const float ux1 = ux[((int64_t)(i + 1) * ny + j) * nz + k];
                const float uy0 = uy[((int64_t)i * (ny + 1) + j) * nz + k];
                const float uy1 = uy[((int64_t)i * (ny + 1) + j + 1) * nz + k];
                const float uz0 = uz[((int64_t)i * ny + j) * (nz + 1) + k];
                const float uz1 = uz[((int64_t)i * ny + j) * (nz + 1) + k + 1];
                const float dudx = (ux1 - ux0) / dx;
                const float dvdy = (uy1 - uy0) / dx;
                const float dwdz = (uz1 - uz0) / dx;
                div[c] = dudx + dvdy + dwdz;
            }
}

/* _pressure_neighbour — fluid.py:222-240. */
static inline float nbr_value(int m, float p_nbr, float p_self, int mat_fluid, int mat_air) {
    if (m == mat_fluid) return p_nbr;
    if (m == mat_air) return 0.0f;
    return p_self;
}

static inline float nbr_at(const float* p, const int32_t* mat, int i, int j, int k,
                           int nx, int ny, int nz, float p_self, int mat_fluid, int mat_air) {
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return 0.0f; /* air */
    const int64_t c = CIDX(i, j, k);
    return nbr_value(mat[c], p[c], p_self, mat_fluid, mat_air);
}

/* jacobi_pressure_step — fluid.py:243-276.
 * s = ((((L+R)+D)+U)+B)+F, rhs = (((rho*dx)*dx)*div)/dt, p_new = (s-rhs)/6. */
void orc_jacobi_step(float* p_new, const float* p_old, const float* div, const int32_t* mat,
                     int nx, int ny, int nz, float dx, float dt, float rho,
                     int mat_fluid, int mat_air) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mat_fluid) { p_new[c] = 0.0f; continue; }
                const float ps = p_old[c];
                const float L = nbr_at(p_old, mat, i - 1, j, k, nx, ny, nz, ps, mat_fluid, mat_air);
                const float R = nbr_at(p_old, mat, i + 1, j, k, nx, ny, nz, ps, mat_fluid, mat_air);
                const float D = nbr_at(p_old, mat, i, j - 1, k, nx, ny, nz, ps, mat_fluid, mat_air);
                const float U = nbr_at(p_old, mat, i, j + 1, k, nx, ny, nz, ps, mat_fluid, mat_air);
                const float B = nbr_at(p_old, mat, i, j, k - 1, nx, ny, nz, ps, mat_fluid, mat_air);
                const float F = nbr_at(p_old, mat, i, j, k + 1, nx, ny, nz, ps, mat_fluid, mat_air);
                float s = L + R;
                s = s + D;
                s = s + U;
                s = s + B;
                s = s + F;
                float rhs = rho * dx;
                rhs = rhs * dx;
                rhs = rhs * div[c];
                rhs = rhs / dt;
                p_new[c] = (s - rhs) / 6.0f;
            }
}

/* Ping-pong driver — fluid.py:688-704 (and pressure_solve_launch,
 * crates/cuda-kernels/src/jacobi_pressure.cu:228-270): iteration `it` reads p
 * when even, p_tmp when odd; the result always ends in p. */
void orc_pressure_solve(float* p, float* p_tmp, const float* div, const int32_t* mat,
                        int nx, int ny, int nz, float dx, float dt, float rho, int n_iter,
                        int mat_fluid, int mat_air) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || n_iter <= 0) return;
    for (int it = 0; it < n_iter; ++it) {
        float* dst = (it & 1) ? p : p_tmp;
        const float* src = (it & 1) ? p_tmp : p;
        orc_jacobi_step(dst, src, div, mat, nx, ny, nz, dx, dt, rho, mat_fluid, mat_air);
    }
    if (n_iter & 1) memcpy(p, p_tmp, sizeof(float) * (size_t)nx * ny * nz);
}

/* subtract_pressure_gradient_{x,y,z} — fluid.py:279-330.  Interior faces
 * whose two cells are both fluid: u -= (((dt/rho)*(p_hi-p_lo))/dx). */
void orc_gradient(float* ux, float* uy, float* uz, const float* p, const int32_t* mat,
                  int nx, int ny, int nz, float dx, float dt, float rho, int mat_fluid) {
    const float dtr = dt / rho;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (mat[CIDX(i - 1, j, k)] != mat_fluid || mat[CIDX(i, j, k)] != mat_fluid) continue;
                const int64_t f = ((int64_t)i * ny + j) * nz + k;
                float g = dtr * (p[CIDX(i, j, k)] - p[CIDX(i - 1, j, k)]);
                g = g / dx;
                ux[f] = ux[f] - g;
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (mat[CIDX(i, j - 1, k)] != mat_fluid || mat[CIDX(i, j, k)] != mat_fluid) continue;
                const int64_t f = ((int64_t)i * (ny + 1) + j) * nz + k;
                float g = dtr * (p[CIDX(i, j, k)] - p[CIDX(i, j - 1, k)]);
                g = g / dx;
                uy[f] = uy[f] - g;
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                if (mat[CIDX(i, j, k - 1)] != mat_fluid || mat[CIDX(i, j, k)] != mat_fluid) continue;
                const int64_t f = ((int64_t)i * ny + j) * (nz + 1) + k;
                float g = dtr * (p[CIDX(i, j, k)] - p[CIDX(i, j, k - 1)]);
                g = g / dx;
                uz[f] = uz[f] - g;
            }
}

double orc_max_abs(const float* a, int64_t n) {
    float m = 0.0f;
    for (int64_t t = 0; t < n; ++t) {
        const float v = fabsf(a[t]);
        if (v > m) m = v;
    }
    return (double)m;
}

/* pressure_projection — fluid.py:612-734 with the Jacobi solver: zero p and
 * p_tmp, divergence, n_iter sweeps, gradient, divergence again, return
 * max|div|. */
double orc_projection(float* p, float* p_tmp, float* div, float* ux, float* uy, float* uz,
                      const int32_t* mat, int nx, int ny, int nz, float dx, float dt, float rho,
                      int n_iter, int mat_fluid, int mat_air) {
    const int64_t n = (int64_t)nx * ny * nz;
    memset(p, 0, sizeof(float) * (size_t)n);
    memset(p_tmp, 0, sizeof(float) * (size_t)n);
    orc_divergence(div, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid);
    orc_pressure_solve(p, p_tmp, div, mat, nx, ny, nz, dx, dt, rho, n_iter, mat_fluid, mat_air);
    orc_gradient(ux, uy, uz, p, mat, nx, ny, nz, dx, dt, rho, mat_fluid);
    orc_divergence(div, ux, uy, uz, mat, nx, ny, nz, dx, mat_fluid);
    return orc_max_abs(div, n);
}

/* enforce_no_slip_u{x,y,z} — fluid.py:111-150.  Each face array in its own
 * loop nest, as the three launches of enforce_no_slip (fluid.py:153-162):
 * a face on the domain edge, or with a non-fluid cell on either side, is set
 * to 0.0; the others are left alone. */
void orc_enforce_no_slip(float* ux, float* uy, float* uz, const int32_t* mat, int nx, int ny, int nz,
                         int mat_fluid) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t f = ((int64_t)i * ny + j) * nz + k;
                if (i == 0 || i == nx) { ux[f] = 0.0f; continue; }
                if (mat[CIDX(i - 1, j, k)] != mat_fluid || mat[CIDX(i, j, k)] != mat_fluid) ux[f] = 0.0f;
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t f = ((int64_t)i * (ny + 1) + j) * nz + k;
                if (j == 0 || j == ny) { uy[f] = 0.0f; continue; }
                if (mat[CIDX(i, j - 1, k)] != mat_fluid || mat[CIDX(i, j, k)] != mat_fluid) uy[f] = 0.0f;
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k <= nz; ++k) {
                const int64_t f = ((int64_t)i * ny + j) * (nz + 1) + k;
                if (k == 0 || k == nz) { uz[f] = 0.0f; continue; }
                if (mat[CIDX(i, j, k - 1)] != mat_fluid || mat[CIDX(i, j, k)] != mat_fluid) uz[f] = 0.0f;
            }
}

/* compute_max_velocity — fluid.py:830-833 (the host path): the largest |u|
 * over the three face arrays. */
double orc_max_velocity(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz) {
    const double a = orc_max_abs(ux, (int64_t)(nx + 1) * ny * nz);
    const double b = orc_max_abs(uy, (int64_t)nx * (ny + 1) * nz);
    const double c = orc_max_abs(uz, (int64_t)nx * ny * (nz + 1));
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

/* ---------------------------------------------------------------------------
 * Linear-system view used by the PCG (crates/cuda-kernels/src/
 * laplacian_spmv.cu:60-143, diag_inverse_apply.cu:43-89,
 * axpy_device.cu:104-123, pressure_solve_pcg.cu:137):
 *   A = 6I - S with the _pressure_neighbour rules (non-fluid rows are 0),
 *   b = (-rho*dx*dx/dt) * div on fluid cells,
 *   M = diag(6 - #solid neighbours), diag <= 0 -> identity.
 * ------------------------------------------------------------------------- */

static inline double nbr_d(const double* x, const int32_t* mat, int i, int j, int k,
                           int nx, int ny, int nz, double xs, int mat_fluid, int mat_air) {
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return 0.0;
    const int64_t c = CIDX(i, j, k);
    if (mat[c] == mat_fluid) return x[c];
    if (mat[c] == mat_air) return 0.0;
    return xs;
}

void orc_laplacian_f64(double* y, const double* x, const int32_t* mat, int nx, int ny, int nz,
                       int mat_fluid, int mat_air) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mat_fluid) { y[c] = 0.0; continue; }
                const double xs = x[c];
                double s = nbr_d(x, mat, i - 1, j, k, nx, ny, nz, xs, mat_fluid, mat_air);
                s += nbr_d(x, mat, i + 1, j, k, nx, ny, nz, xs, mat_fluid, mat_air);
                s += nbr_d(x, mat, i, j - 1, k, nx, ny, nz, xs, mat_fluid, mat_air);
                s += nbr_d(x, mat, i, j + 1, k, nx, ny, nz, xs, mat_fluid, mat_air);
                s += nbr_d(x, mat, i, j, k - 1, nx, ny, nz, xs, mat_fluid, mat_air);
                s += nbr_d(x, mat, i, j, k + 1, nx, ny, nz, xs, mat_fluid, mat_air);
                y[c] = 6.0 * xs - s;
            }
}

static inline int is_solid(int m, int mat_fluid, int mat_air) {
    return m != mat_fluid && m != mat_air;
}

static inline int mat_or_air(const int32_t* mat, int i, int j, int k, int nx, int ny, int nz,
                             int mat_air) {
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return mat_air;
    return mat[CIDX(i, j, k)];
}

/* diag_inverse_apply.cu:63-89: the diagonal 6 - #solid neighbours; 0 marks a
 * non-fluid row and -1 marks a fluid row whose diagonal is <= 0 (z = r). */
void orc_diag(double* d, const int32_t* mat, int nx, int ny, int nz, int mat_fluid, int mat_air) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mat_fluid) { d[c] = 0.0; continue; }
                int ns = 0;
                ns += is_solid(mat_or_air(mat, i - 1, j, k, nx, ny, nz, mat_air), mat_fluid, mat_air);
                ns += is_solid(mat_or_air(mat, i + 1, j, k, nx, ny, nz, mat_air), mat_fluid, mat_air);
                ns += is_solid(mat_or_air(mat, i, j - 1, k, nx, ny, nz, mat_air), mat_fluid, mat_air);
                ns += is_solid(mat_or_air(mat, i, j + 1, k, nx, ny, nz, mat_air), mat_fluid, mat_air);
                ns += is_solid(mat_or_air(mat, i, j, k - 1, nx, ny, nz, mat_air), mat_fluid, mat_air);
                ns += is_solid(mat_or_air(mat, i, j, k + 1, nx, ny, nz, mat_air), mat_fluid, mat_air);
                const double dg = 6.0 - (double)ns;
                d[c] = dg > 0.0 ? dg : -1.0;
            }
}

/* b = (-rho*dx*dx/dt) * div on fluid cells (pressure_solve_pcg.cu:137,
 * axpy_device.cu:104-123).  The scale is formed in float32 as the reference
 * forms it, then applied in float64. */
void orc_rhs_f64(double* b, const float* div, const int32_t* mat, int nx, int ny, int nz,
                 float dx, float dt, float rho, int mat_fluid) {
    const float scale = -rho * dx * dx / dt;
    const int64_t n = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) b[c] = (mat[c] == mat_fluid) ? (double)scale * (double)div[c] : 0.0;
}

/* Deterministic FP64 dot: 256 fixed chunks summed in order, independent of
 * the thread count. */
#define ORC_DOT_CHUNKS 256
static double dot_d(const double* x, const double* y, int64_t n) {
    double part[ORC_DOT_CHUNKS];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ORC_DOT_CHUNKS; ++c) {
        const int64_t lo = n * c / ORC_DOT_CHUNKS, hi = n * (c + 1) / ORC_DOT_CHUNKS;
        double s = 0.0;
        for (int64_t t = lo; t < hi; ++t) s += x[t] * y[t];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < ORC_DOT_CHUNKS; ++c) s += part[c];
    return s;
}

static double dot_f(const float* x, const float* y, int64_t n) {
    double part[ORC_DOT_CHUNKS];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ORC_DOT_CHUNKS; ++c) {
        const int64_t lo = n * c / ORC_DOT_CHUNKS, hi = n * (c + 1) / ORC_DOT_CHUNKS;
        double s = 0.0;
        for (int64_t t = lo; t < hi; ++t) s += (double)x[t] * (double)y[t];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < ORC_DOT_CHUNKS; ++c) s += part[c];
    return s;
}

/* True relative residual ||b - A x||_2 / ||b||_2 in FP64 for an FP32 or FP64
 * solution.  The reference's CPU residual oracle is test_pressure_cg.py:295-325
 * (it reports the inf-norm; both norms are returned). */
void orc_residual_f64(const double* x, const double* b, const int32_t* mat, int nx, int ny, int nz,
                      int mat_fluid, int mat_air, double* rel2_out, double* relinf_out) {
    const int64_t n = (int64_t)nx * ny * nz;
    double* ax = (double*)malloc(sizeof(double) * (size_t)n);
    orc_laplacian_f64(ax, x, mat, nx, ny, nz, mat_fluid, mat_air);
    double rr = 0.0, bb = 0.0, rinf = 0.0, binf = 0.0;
    for (int64_t t = 0; t < n; ++t) {
        const double r = b[t] - ax[t];
        rr += r * r;
        bb += b[t] * b[t];
        if (fabs(r) > rinf) rinf = fabs(r);
        if (fabs(b[t]) > binf) binf = fabs(b[t]);
    }
    free(ax);
    if (rel2_out) *rel2_out = bb > 0.0 ? sqrt(rr / bb) : 0.0;
    if (relinf_out) *relinf_out = binf > 0.0 ? rinf / binf : 0.0;
}

/* FP64 Jacobi-preconditioned CG on A x = b from the initial guess in x.
 * Stops when the recursive residual satisfies ||r|| <= tol*||b|| (checked
 * every iteration); returns the iteration count.  This is the FP64 oracle
 * every solver's error is stated against. */
int orc_pcg_f64(double* x, const double* b, const int32_t* mat, int nx, int ny, int nz,
                double tol, int max_iter, int mat_fluid, int mat_air) {
    const int64_t n = (int64_t)nx * ny * nz;
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)n);
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    orc_diag(d, mat, nx, ny, nz, mat_fluid, mat_air);
    orc_laplacian_f64(q, x, mat, nx, ny, nz, mat_fluid, mat_air);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; ++t) {
        r[t] = (d[t] == 0.0) ? 0.0 : b[t] - q[t];
        z[t] = (d[t] == 0.0) ? 0.0 : (d[t] < 0.0 ? r[t] : r[t] / d[t]);
        p[t] = z[t];
    }
    const double bb = dot_d(b, b, n);
    int it = 0;
    if (bb == 0.0) {
        memset(x, 0, sizeof(double) * (size_t)n);
        goto done;
    }
    double rz = dot_d(r, z, n);
    const double stop = tol * tol * bb;
    if (dot_d(r, r, n) <= stop) goto done;
    for (it = 0; it < max_iter;) {
        orc_laplacian_f64(q, p, mat, nx, ny, nz, mat_fluid, mat_air);
        const double pq = dot_d(p, q, n);
        const double alpha = pq != 0.0 ? rz / pq : 0.0;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; ++t) {
            x[t] += alpha * p[t];
            r[t] -= alpha * q[t];
        }
        ++it;
        if (dot_d(r, r, n) <= stop) break;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; ++t) z[t] = (d[t] == 0.0) ? 0.0 : (d[t] < 0.0 ? r[t] : r[t] / d[t]);
        const double rzn = dot_d(r, z, n);
        const double beta = rz != 0.0 ? rzn / rz : 0.0;
        rz = rzn;
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; ++t) p[t] = z[t] + beta * p[t];
    }
done:
    free(r); free(z); free(p); free(q); free(d);
    return it;
}

/* Jacobi-PCG with the reference driver's control flow
 * (crates/cuda-kernels/src/pressure_solve_pcg.cu:89-243): cold start,
 * bsq == 0 -> zero output and 0 iterations, convergence tested on the
 * recursive residual only when iter >= 8 and iter % 5 == 0, and the count is
 * iter+1 on that break.  Vectors are FP32 and the scalars (dots, alpha, beta)
 * FP64 — the precision split of the HIP solver, which replaces the
 * reference's FP32 dots. Returns the iteration count; *bsq_out gets b.b. */
static inline float pcg_nb(const float* p, const int32_t* mat, int64_t c, float ps, int mf, int ma) {
    return nbr_value(mat[c], p[c], ps, mf, ma);
}

static inline float pcg_z(float r, float dg) {
    return dg == 0.0f ? 0.0f : (dg < 0.0f ? r : r / dg);
}

/* The PCG's state: FP32 vectors and per-plane FP64 partials (summed in
 * plane order, so every dot is deterministic for any thread count). */
typedef struct {
    int nx, ny, nz, mf, ma;
    int64_t n, plane;
    const int32_t* mat;
    float *b, *r, *p, *q, *dg;
    double* part; /* 3 x nx */
} PcgState;

static int pcg_state_init(PcgState* S, const float* div, const int32_t* mat, int nx, int ny, int nz, float dx,
                          float dt, float rho, int mf, int ma) {
    S->nx = nx; S->ny = ny; S->nz = nz; S->mf = mf; S->ma = ma; S->mat = mat;
    S->n = (int64_t)nx * ny * nz;
    S->plane = (int64_t)ny * nz;
    const size_t fb = sizeof(float) * (size_t)S->n;
    S->b = (float*)malloc(fb); S->r = (float*)malloc(fb); S->p = (float*)malloc(fb);
    S->q = (float*)malloc(fb); S->dg = (float*)malloc(fb);
    S->part = (double*)malloc(sizeof(double) * (size_t)nx * 3);
    double* d = (double*)malloc(sizeof(double) * (size_t)S->n);
    if (!S->b || !S->r || !S->p || !S->q || !S->dg || !S->part || !d) { free(d); return -1; }
    orc_diag(d, mat, nx, ny, nz, mf, ma);
    const float scale = -rho * dx * dx / dt;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < S->n; ++t) {
        S->b[t] = (mat[t] == mf) ? scale * div[t] : 0.0f;
        S->dg[t] = (float)d[t]; /* 0: non-fluid, -1: z = r, else 6 - #solid */
    }
    free(d);
    return 0;
}

static void pcg_state_free(PcgState* S) {
    free(S->b); free(S->r); free(S->p); free(S->q); free(S->dg); free(S->part);
}

static double pcg_sum(const double* v, int nx) {
    double s = 0.0;
    for (int i = 0; i < nx; ++i) s += v[i];
    return s;
}

/* r = b, p = z = M^-1 r; returns r.z. */
static double pcg_cold_start(PcgState* S) {
    const int64_t plane = S->plane;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < S->nx; ++i) {
        double s = 0.0;
        for (int64_t t = i * plane; t < (i + 1) * plane; ++t) {
            S->r[t] = S->b[t];
            S->p[t] = pcg_z(S->r[t], S->dg[t]);
            s += (double)S->r[t] * (double)S->p[t];
        }
        S->part[i] = s;
    }
    return pcg_sum(S->part, S->nx);
}

/* q = A p in float32 (6*p_self - s, s summed in the reference order);
 * returns p.q. */
static double pcg_spmv_pass(PcgState* S) {
    const int nx = S->nx, ny = S->ny, nz = S->nz, mf = S->mf, ma = S->ma;
    const int64_t plane = S->plane;
    const int32_t* mat = S->mat;
    const float* p = S->p;
    float* q = S->q;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        double s_pq = 0.0;
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mf) { q[c] = 0.0f; continue; }
                const float ps = p[c];
                float s = i > 0 ? pcg_nb(p, mat, c - plane, ps, mf, ma) : 0.0f;
                s = s + (i < nx - 1 ? pcg_nb(p, mat, c + plane, ps, mf, ma) : 0.0f);
                s = s + (j > 0 ? pcg_nb(p, mat, c - nz, ps, mf, ma) : 0.0f);
                s = s + (j < ny - 1 ? pcg_nb(p, mat, c + nz, ps, mf, ma) : 0.0f);
                s = s + (k > 0 ? pcg_nb(p, mat, c - 1, ps, mf, ma) : 0.0f);
                s = s + (k < nz - 1 ? pcg_nb(p, mat, c + 1, ps, mf, ma) : 0.0f);
                const float qc = 6.0f * ps - s;
                q[c] = qc;
                s_pq += (double)ps * (double)qc;
            }
        S->part[i] = s_pq;
    }
    return pcg_sum(S->part, nx);
}

/* x += af p, r -= af q; returns r.r in *rr and r.z (z = M^-1 r) in *rz. */
static void pcg_update_pass(PcgState* S, float* x, float af, double* rr, double* rz) {
    const int nx = S->nx;
    const int64_t plane = S->plane;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        double s_rr = 0.0, s_rz = 0.0;
        for (int64_t t = i * plane; t < (i + 1) * plane; ++t) {
            x[t] = x[t] + af * S->p[t];
            const float rt = S->r[t] - af * S->q[t];
            S->r[t] = rt;
            s_rr += (double)rt * (double)rt;
            s_rz += (double)rt * (double)pcg_z(rt, S->dg[t]);
        }
        S->part[nx + i] = s_rr;
        S->part[2 * nx + i] = s_rz;
    }
    *rr = pcg_sum(S->part + nx, nx);
    *rz = pcg_sum(S->part + 2 * nx, nx);
}

/* p = M^-1 r + bf p. */
static void pcg_direction_pass(PcgState* S, float bf) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < S->n; ++t) S->p[t] = pcg_z(S->r[t], S->dg[t]) + bf * S->p[t];
}

/* The true residual b - A x of the FP32 iterate, formed in FP64 (the
 * neighbour sum and the subtraction in double); returns its squared norm.
 * With restart, r takes it rounded to FP32 and p = M^-1 r, and *rz = r.z. */
static double pcg_true_residual(PcgState* S, const float* x, int restart, double* rz) {
    const int nx = S->nx, ny = S->ny, nz = S->nz, mf = S->mf, ma = S->ma;
    const int64_t plane = S->plane;
    const int32_t* mat = S->mat;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        double s_rr = 0.0, s_rz = 0.0;
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = CIDX(i, j, k);
                if (mat[c] != mf) {
                    if (restart) { S->r[c] = 0.0f; S->p[c] = 0.0f; }
                    continue;
                }
                const double xs = (double)x[c];
                double s = i > 0 ? (double)pcg_nb(x, mat, c - plane, x[c], mf, ma) : 0.0;
                s += i < nx - 1 ? (double)pcg_nb(x, mat, c + plane, x[c], mf, ma) : 0.0;
                s += j > 0 ? (double)pcg_nb(x, mat, c - nz, x[c], mf, ma) : 0.0;
                s += j < ny - 1 ? (double)pcg_nb(x, mat, c + nz, x[c], mf, ma) : 0.0;
                s += k > 0 ? (double)pcg_nb(x, mat, c - 1, x[c], mf, ma) : 0.0;
                s += k < nz - 1 ? (double)pcg_nb(x, mat, c + 1, x[c], mf, ma) : 0.0;
                const double r64 = (double)S->b[c] - (6.0 * xs - s);
                s_rr += r64 * r64;
                if (restart) {
                    const float rf = (float)r64;
                    const float zf = pcg_z(rf, S->dg[c]);
                    S->r[c] = rf;
                    S->p[c] = zf;
                    s_rz += (double)rf * (double)zf;
                }
            }
        S->part[nx + i] = s_rr;
        S->part[2 * nx + i] = s_rz;
    }
    if (rz) *rz = pcg_sum(S->part + 2 * nx, nx);
    return pcg_sum(S->part + nx, nx);
}

int orc_pcg_ref(float* p_out, const float* div, const int32_t* mat, int nx, int ny, int nz,
                float dx, float dt, float rho, float tol, int max_iter,
                int mat_fluid, int mat_air, double* bsq_out) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    PcgState S;
    if (pcg_state_init(&S, div, mat, nx, ny, nz, dx, dt, rho, mat_fluid, mat_air) != 0) {
        pcg_state_free(&S);
        return -1;
    }
    const double bsq = dot_f(S.b, S.b, S.n);
    if (bsq_out) *bsq_out = bsq;
    int iter = 0;
    memset(p_out, 0, sizeof(float) * (size_t)S.n);
    if (bsq == 0.0) goto done;
    const double stop = (double)tol * (double)tol * bsq;
    double rz = pcg_cold_start(&S);
    /* Three fused passes per iteration, the same float operations as the
     * reference's kernel sequence: (1) q = A p with p.q; (2) x += alpha p,
     * r -= alpha q with r.r and r.z; (3) p = z + beta p, z formed again from r. */
    for (iter = 0; iter < max_iter; ++iter) {
        const double pq = pcg_spmv_pass(&S);
        const double alpha = pq != 0.0 ? rz / pq : 0.0;
        double rr, rzn;
        pcg_update_pass(&S, p_out, (float)alpha, &rr, &rzn);
        if (iter >= 8 && iter % 5 == 0 && rr < stop) { ++iter; break; }
        const double beta = rz != 0.0 ? rzn / rz : 0.0;
        pcg_direction_pass(&S, (float)beta);
        rz = rzn;
    }
done:
    pcg_state_free(&S);
    return iter;
}

/* The same PCG to a TRUE relative residual tol, as bs_pcg_solve
 * (boiling-sim_amd/csrc/pcg.hip): cold start, the recursive residual tested
 * every iteration, then ||b - A x|| formed in FP64; while that is above tol
 * the solve restarts from x with r = b - A x (residual replacement), at most
 * max_restarts times, the iteration count continuing.  Returns the
 * iterations; *relres_out = ||b - A x||_2 / ||b||_2, *restarts_out the
 * restarts taken. */
int orc_pcg_true(float* x, const float* div, const int32_t* mat, int nx, int ny, int nz, float dx, float dt,
                 float rho, double tol, int max_iter, int max_restarts, int mat_fluid, int mat_air,
                 double* relres_out, int* restarts_out) {
    if (relres_out) *relres_out = 0.0;
    if (restarts_out) *restarts_out = 0;
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    PcgState S;
    if (pcg_state_init(&S, div, mat, nx, ny, nz, dx, dt, rho, mat_fluid, mat_air) != 0) {
        pcg_state_free(&S);
        return -1;
    }
    memset(x, 0, sizeof(float) * (size_t)S.n);
    const double bsq = dot_f(S.b, S.b, S.n);
    int iter = 0, restarts = 0;
    if (bsq == 0.0) goto done;
    const double stop = tol * tol * bsq;
    double rz = pcg_cold_start(&S);
    for (;;) {
        while (iter < max_iter) {
            const double pq = pcg_spmv_pass(&S);
            const double alpha = pq != 0.0 ? rz / pq : 0.0;
            double rr, rzn;
            pcg_update_pass(&S, x, (float)alpha, &rr, &rzn);
            ++iter;
            if (rr < stop) break;
            const double beta = rz != 0.0 ? rzn / rz : 0.0;
            pcg_direction_pass(&S, (float)beta);
            rz = rzn;
        }
        const double rr_true = pcg_true_residual(&S, x, 0, NULL);
        if (relres_out) *relres_out = sqrt(rr_true / bsq);
        if (rr_true <= stop || restarts >= max_restarts || iter >= max_iter) break;
        ++restarts;
        pcg_true_residual(&S, x, 1, &rz);
    }
done:
    if (restarts_out) *restarts_out = restarts;
    pcg_state_free(&S);
    return iter;
}
