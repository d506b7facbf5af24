// Backward-Euler Jacobi heat conduction on gfx950: the implicit branch of the
// reference's conduct_one_step (python/boilingsim/thermal.py:631-665) and its
// sweep kernel apply_implicit_conduction_update (thermal.py:259-388).
//
// Per non-air cell, with every float operation rounded separately:
//   for the faces -x, +x, -y, +y, -z, +z inside the grid
//     neighbour air:  g = ((dt*h)*inv_rho_cp)/dx,  term = g*T_amb
//     otherwise:      kf = ((2*k_m)*k_n)/((k_m + k_n) + 1e-12)
//                     g = ((dt*kf)*inv_rho_cp)/(dx*dx),  term = g*T_iter[n]
//     rhs = rhs + term,  diag = diag + g
//   T_dst = (T_old + rhs)/(1 + diag);  air cells are not written.
// g depends only on the (cell material, neighbour material) pair, so each
// workgroup first evaluates those expressions for every pair into LDS (the
// same operations on the same values, hence the same bits), and the sweep
// itself does one table read, one multiply and two adds per face and one
// division per cell instead of two divisions per face.
#include "boilsim_hip.h"
#include "grid.hpp"

namespace bs {

constexpr int kMaxMat = 16;  // material ids 0..15 (the reference uses 0..3)

struct CondParams {
    float dx, dt, h, t_amb;
    int mat_air, n_mat;
};

// The pair table: g[m][n] for a non-air neighbour n, and in column kMaxMat
// the air neighbour's g_air[m]; t_air[m] = g_air[m]*T_amb is the air face's
// term.  One table read per face, the air case a select.
struct CondTable {
    float g[kMaxMat][kMaxMat + 1];
    float t_air[kMaxMat];
};

__device__ __forceinline__ void build_table(CondTable& tb, const float* __restrict__ k_arr,
                                            const float* __restrict__ rho_arr, const float* __restrict__ cp_arr,
                                            const CondParams& P) {
    const int nm = P.n_mat;
    for (int t = threadIdx.x; t < nm * nm; t += blockDim.x) {
        const int m = t / nm, n = t - (t / nm) * nm;
        const float k_self = k_arr[m];
        const float inv_rho_cp = 1.0f / (rho_arr[m] * cp_arr[m]);
        const float k_nbr = k_arr[n];
        const float kf = ((2.0f * k_self) * k_nbr) / ((k_self + k_nbr) + 1.0e-12f);
        tb.g[m][n] = ((P.dt * kf) * inv_rho_cp) / (P.dx * P.dx);
    }
    for (int m = threadIdx.x; m < nm; m += blockDim.x) {
        const float inv_rho_cp = 1.0f / (rho_arr[m] * cp_arr[m]);
        const float g = ((P.dt * P.h) * inv_rho_cp) / P.dx;
        tb.g[m][kMaxMat] = g;
        tb.t_air[m] = g * P.t_amb;
    }
}

// Material ids index the table; ids outside [0, n_mat) (an out-of-bounds read
// in the reference) are clamped to 0 to keep the lookup inside it.
__device__ __forceinline__ int mid(int m, int nm) { return (unsigned)m < (unsigned)nm ? m : 0; }

template <int VEC>
__global__ __launch_bounds__(256) void conduction_sweep_kernel(Tiles T, float* __restrict__ t_dst,
                                                               const float* __restrict__ t_iter,
                                                               const float* __restrict__ t_old,
                                                               const int* __restrict__ mat,
                                                               const float* __restrict__ k_arr,
                                                               const float* __restrict__ rho_arr,
                                                               const float* __restrict__ cp_arr, CondParams P) {
    __shared__ CondTable tb;
    build_table(tb, k_arr, rho_arr, cp_arr, P);
    // The sweep's own loads are issued before the table's barrier, so the two
    // latencies overlap; every thread reaches the barrier.
    int i = 0, j = 0, k0 = 0;
    int64_t c0 = 0;
    const bool valid = one_tile(T, VEC, i, j, k0, c0);
    const int64_t plane = (int64_t)T.ny * T.nz;
    const int ma = P.mat_air, nm = P.n_mat;
    int mc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) mc[e] = ma;
    if (valid) ldi<VEC>(mat, c0, mc);
    bool any = false;
#pragma unroll
    for (int e = 0; e < VEC; ++e) any |= mc[e] != ma;
    const bool work = valid && any;  // all air: nothing is written
    // neighbour columns (absent ones flagged by have_*)
    const bool have_l = i > 0, have_r = i < T.nx - 1, have_d = j > 0, have_u = j < T.ny - 1;
    float tc[VEC], to[VEC];
    int ml[VEC], mr[VEC], md[VEC], mu[VEC];
    float tl[VEC], tr[VEC], td[VEC], tu[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        tc[e] = to[e] = 0.0f;
        ml[e] = mr[e] = md[e] = mu[e] = 0;
        tl[e] = tr[e] = td[e] = tu[e] = 0.0f;
    }
    int mb = 0, mf = 0;
    float tb_ = 0.0f, tf = 0.0f;
    if (work) {
        ldf<VEC>(t_iter, c0, tc);
        ldf<VEC>(t_old, c0, to);
        if (have_l) { ldi<VEC>(mat, c0 - plane, ml); ldf<VEC>(t_iter, c0 - plane, tl); }
        if (have_r) { ldi<VEC>(mat, c0 + plane, mr); ldf<VEC>(t_iter, c0 + plane, tr); }
        if (have_d) { ldi<VEC>(mat, c0 - T.nz, md); ldf<VEC>(t_iter, c0 - T.nz, td); }
        if (have_u) { ldi<VEC>(mat, c0 + T.nz, mu); ldf<VEC>(t_iter, c0 + T.nz, tu); }
        if (k0 > 0) { mb = mat[c0 - 1]; tb_ = t_iter[c0 - 1]; }
        if (k0 + VEC < T.nz) { mf = mat[c0 + VEC]; tf = t_iter[c0 + VEC]; }
    }
    __syncthreads();
    if (!work) return;
    const bool inner = have_l && have_r && have_d && have_u && k0 > 0 && k0 + VEC < T.nz;
    float out[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        const int m = mc[e];
        out[e] = 0.0f;
        if (m == ma) continue;
        const int ms = mid(m, nm);
        const float* grow = tb.g[ms];
        const float tair = tb.t_air[ms];
        float rhs = 0.0f, diag = 0.0f;
        // one table read per face, the air case a select
        auto face = [&](int mn, float tn) {
            const bool air = mn == ma;
            const float g = grow[air ? kMaxMat : mid(mn, nm)];
            const float term = air ? tair : g * tn;
            rhs = rhs + term;
            diag = diag + g;
        };
        const int mb_e = e > 0 ? mc[e - 1] : mb, mf_e = e < VEC - 1 ? mc[e + 1] : mf;
        const float tb_e = e > 0 ? tc[e - 1] : tb_, tf_e = e < VEC - 1 ? tc[e + 1] : tf;
        if (inner) {  // no face of the chunk lies outside the grid
            face(ml[e], tl[e]);
            face(mr[e], tr[e]);
            face(md[e], td[e]);
            face(mu[e], tu[e]);
            face(mb_e, tb_e);
            face(mf_e, tf_e);
        } else {
            if (have_l) face(ml[e], tl[e]);
            if (have_r) face(mr[e], tr[e]);
            if (have_d) face(md[e], td[e]);
            if (have_u) face(mu[e], tu[e]);
            if (k0 + e > 0) face(mb_e, tb_e);
            if (k0 + e < T.nz - 1) face(mf_e, tf_e);
        }
        out[e] = (to[e] + rhs) / (1.0f + diag);
    }
    bool all = true;
#pragma unroll
    for (int e = 0; e < VEC; ++e) all &= mc[e] != ma;
    if (all) {
        stf<VEC>(t_dst, c0, out);
    } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e)
            if (mc[e] != ma) t_dst[c0 + e] = out[e];
    }
}

static int cond_vec(int nz, std::initializer_list<const void*> ptrs) { return vec_for(nz, ptrs); }

static int launch_cond_sweep(hipStream_t st, int vec, const Tiles& T, float* t_dst, const float* t_iter,
                             const float* t_old, const int* mat, const float* k_arr, const float* rho_arr,
                             const float* cp_arr, const CondParams& P) {
    if (vec == 4)
        conduction_sweep_kernel<4><<<T.ntiles, 256, 0, st>>>(T, t_dst, t_iter, t_old, mat, k_arr, rho_arr, cp_arr, P);
    else if (vec == 2)
        conduction_sweep_kernel<2><<<T.ntiles, 256, 0, st>>>(T, t_dst, t_iter, t_old, mat, k_arr, rho_arr, cp_arr, P);
    else
        conduction_sweep_kernel<1><<<T.ntiles, 256, 0, st>>>(T, t_dst, t_iter, t_old, mat, k_arr, rho_arr, cp_arr, P);
    return (int)hipGetLastError();
}

}  // namespace bs

using namespace bs;

// One sweep of apply_implicit_conduction_update (thermal.py:259-388).
// k_arr, rho_arr, cp_arr: n_mat floats on the device (material id -> k, rho,
// c_p); t_dst must not overlap t_iter.  Asynchronous on `stream`.
extern "C" int bs_conduction_sweep_launch(float* t_dst, const float* t_iter, const float* t_old, const int* mat,
                                          const float* k_arr, const float* rho_arr, const float* cp_arr, int n_mat,
                                          int nx, int ny, int nz, float dx, float dt, float h_conv_air,
                                          float t_amb_k, int mat_air, int mat_fluid, bs_stream_t stream) {
    (void)mat_fluid;  // a kernel argument of the reference that its arithmetic does not use
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (n_mat <= 0 || n_mat > kMaxMat) return (int)hipErrorInvalidValue;
    const int vec = cond_vec(nz, {t_dst, t_iter, t_old, mat});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    const CondParams P{dx, dt, h_conv_air, t_amb_k, mat_air, n_mat};
    return launch_cond_sweep(stream, vec, T, t_dst, t_iter, t_old, mat, k_arr, rho_arr, cp_arr, P);
}

// The backward-Euler branch of conduct_one_step (thermal.py:648-665): t_old =
// t, n_iter sweeps (sweep `it` reads t when even and t_alt when odd), and an
// odd count ends with t = t_alt (air cells included, as the reference's copy).
extern "C" int bs_conduction_solve_launch(float* t, float* t_alt, float* t_old, const int* mat, const float* k_arr,
                                          const float* rho_arr, const float* cp_arr, int n_mat, int nx, int ny,
                                          int nz, float dx, float dt, float h_conv_air, float t_amb_k, int n_iter,
                                          int mat_air, int mat_fluid, bs_stream_t stream) {
    (void)mat_fluid;
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    if (n_mat <= 0 || n_mat > kMaxMat) return (int)hipErrorInvalidValue;
    hipStream_t st = stream;
    const size_t bytes = sizeof(float) * (size_t)nx * ny * nz;
    BS_RETURN_IF_ERR(hipMemcpyAsync(t_old, t, bytes, hipMemcpyDeviceToDevice, st));
    const int vec = cond_vec(nz, {t, t_alt, t_old, mat});
    const Tiles T = make_tiles(nx, ny, nz, vec);
    const CondParams P{dx, dt, h_conv_air, t_amb_k, mat_air, n_mat};
    for (int it = 0; it < n_iter; ++it) {
        const float* src = (it % 2 == 0) ? t : t_alt;
        float* dst = (it % 2 == 0) ? t_alt : t;
        const int rc = launch_cond_sweep(st, vec, T, dst, src, t_old, mat, k_arr, rho_arr, cp_arr, P);
        if (rc != hipSuccess) return rc;
    }
    if (n_iter % 2 == 1) BS_RETURN_IF_ERR(hipMemcpyAsync(t, t_alt, bytes, hipMemcpyDeviceToDevice, st));
    return (int)hipGetLastError();
}
