/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of boiling-sim's backward-Euler Jacobi heat conduction
 * (python/boilingsim/thermal.py:259-388, apply_implicit_conduction_update,
 * and its driver in conduct_one_step, thermal.py:631-665), the checker for
 * the HIP conduction kernels.  Nothing in the product path links or calls it.
 *
 * Per non-air cell, in float32 with every operation rounded separately
 * (built with -ffp-contract=off), in the reference's order:
 *   inv_rho_cp = 1 / (rho[m] * cp[m])
 *   for the faces -x, +x, -y, +y, -z, +z that lie inside the grid:
 *     neighbour air:  g = ((dt * h) * inv_rho_cp) / dx,   term = g * T_amb
 *     otherwise:      kf = ((2 * k[m]) * k[n]) / ((k[m] + k[n]) + 1e-12)
 *                     g = ((dt * kf) * inv_rho_cp) / (dx * dx),   term = g * T_iter[n]
 *     rhs = rhs + term;  diag = diag + g            (both start at 0)
 *   T_dst = (T_old + rhs) / (1 + diag)
 * Air cells are not written.  Faces on the domain edge carry no flux.
 */
#include <stdint.h>
#include <string.h>

#define TIDX(i, j, k) ((((int64_t)(i)) * ny + (j)) * (int64_t)nz + (k))

/* One face's (g, term) pair, added to (rhs, diag). */
static inline void face(float* rhs, float* diag, int m_nbr, float t_nbr, float k_self, float inv_rho_cp,
                        const float* k_arr, float dx, float dt, float h_conv_air, float t_amb, int mat_air) {
    float g, term;
    if (m_nbr == mat_air) {
        g = ((dt * h_conv_air) * inv_rho_cp) / dx;
        term = g * t_amb;
    } else {
        const float k_nbr = k_arr[m_nbr];
        const float kf = ((2.0f * k_self) * k_nbr) / ((k_self + k_nbr) + 1.0e-12f);
        g = ((dt * kf) * inv_rho_cp) / (dx * dx);
        term = g * t_nbr;
    }
    *rhs = *rhs + term;
    *diag = *diag + g;
}

void orc_conduction_sweep(float* t_dst, const float* t_iter, const float* t_old, const int32_t* mat,
                          const float* k_arr, const float* rho_arr, const float* cp_arr, int nx, int ny, int nz,
                          float dx, float dt, float h_conv_air, float t_amb, int mat_air) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t c = TIDX(i, j, k);
                const int m = mat[c];
                if (m == mat_air) continue;
                const float k_self = k_arr[m];
                const float inv_rho_cp = 1.0f / (rho_arr[m] * cp_arr[m]);
                float rhs = 0.0f, diag = 0.0f;
#define FACE(ii, jj, kk)                                                                                 \
    face(&rhs, &diag, mat[TIDX(ii, jj, kk)], t_iter[TIDX(ii, jj, kk)], k_self, inv_rho_cp, k_arr, dx, dt, \
         h_conv_air, t_amb, mat_air)
                if (i > 0) FACE(i - 1, j, k);
                if (i < nx - 1) FACE(i + 1, j, k);
                if (j > 0) FACE(i, j - 1, k);
                if (j < ny - 1) FACE(i, j + 1, k);
                if (k > 0) FACE(i, j, k - 1);
                if (k < nz - 1) FACE(i, j, k + 1);
#undef FACE
                t_dst[c] = (t_old[c] + rhs) / (1.0f + diag);
            }
}

/* conduct_one_step's backward-Euler branch: t_old = t, then n_iter sweeps,
 * sweep `it` reading t when even and t_alt when odd; an odd count ends with
 * t = t_alt. */
void orc_conduction_solve(float* t, float* t_alt, float* t_old, const int32_t* mat, const float* k_arr,
                          const float* rho_arr, const float* cp_arr, int nx, int ny, int nz, float dx, float dt,
                          float h_conv_air, float t_amb, int n_iter, int mat_air) {
    const size_t bytes = sizeof(float) * (size_t)nx * ny * nz;
    memcpy(t_old, t, bytes);
    for (int it = 0; it < n_iter; ++it) {
        const float* src = (it % 2 == 0) ? t : t_alt;
        float* dst = (it % 2 == 0) ? t_alt : t;
        orc_conduction_sweep(dst, src, t_old, mat, k_arr, rho_arr, cp_arr, nx, ny, nz, dx, dt, h_conv_air, t_amb,
                             mat_air);
    }
    if (n_iter % 2 == 1) memcpy(t, t_alt, bytes);
}
