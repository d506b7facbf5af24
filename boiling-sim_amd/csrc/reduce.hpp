// Deterministic single-pass FP64 grid reductions for 256-thread workgroups.
//
// Each workgroup reduces its values with a wave64 butterfly (__shfl_xor) and
// an LDS step, thread 0 publishes the workgroup partial and takes a ticket
// from an agent-scope atomic counter; the workgroup that draws the last ticket
// sums all partials in workgroup order (fixed, so the result does not depend
// on scheduling) and resets the counter for the next launch.  This replaces
// the reference's two-kernel FP32 dot (crates/cuda-kernels/src/dot_reduce.cu:
// 37-115) with one kernel and FP64 accumulation.
#pragma once

#include "common.hpp"

namespace bs {

constexpr int kRedThreads = 256;
constexpr int kRedWaves = kRedThreads / 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Workgroup sum, result valid in every thread.  `red` is LDS of kRedWaves doubles.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < kRedWaves; ++w) s += red[w];
    return s;
}

__device__ __forceinline__ void store_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// NV values per thread -> totals in the last workgroup.  partials holds
// NV * gridDim.x doubles (value k of workgroup w at k*gridDim.x + w).
// Returns true only in the last workgroup, where v[] then holds the totals
// (in every thread).  Must be reached by every thread of every workgroup.
template <int NV>
__device__ bool grid_reduce(double (&v)[NV], double* __restrict__ partials,
                            unsigned int* __restrict__ ticket) {
    __shared__ double red[NV][kRedWaves];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
    const unsigned G = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = red[k][0];
#pragma unroll
            for (int w = 1; w < kRedWaves; ++w) s += red[k][w];
            store_agent(partials + (size_t)k * G + blockIdx.x, s);
        }
        __threadfence();
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == G - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double acc = 0.0;
        for (unsigned w = threadIdx.x; w < G; w += kRedThreads) acc += load_agent(partials + (size_t)k * G + w);
        const double s = wave_sum(acc);
        __syncthreads();
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = red[k][0];
#pragma unroll
        for (int w = 1; w < kRedWaves; ++w) s += red[k][w];
        v[k] = s;
    }
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

}  // namespace bs
