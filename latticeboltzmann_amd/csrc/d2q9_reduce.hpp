// Block-level FP64 reduction helpers shared by the reduction kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_internal.hpp"

namespace lbm {

// ---------------------------------------------------------------------------
// Deterministic reductions: pass 1 with a fixed grid of kReduceBlocks blocks,
// each thread accumulating a grid-strided sequence with Neumaier compensation,
// then a fixed-shape tree in LDS; pass 2 folds the partials in one block.

__device__ __forceinline__ void neumaier_add(double &s, double &c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[kBlock];
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// max(m, v) for a convergence monitor: a NaN v (a diverged node) becomes
// +infinity, which every later fmax keeps, so the maximum cannot lose it.
__device__ __forceinline__ double monitor_max(double m, double v) {
    return fmax(m, v == v ? v : __builtin_inf());
}

__device__ __forceinline__ double block_max(double v) {
    __shared__ double red[kBlock];
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    return red[0];
}

}  // namespace lbm
