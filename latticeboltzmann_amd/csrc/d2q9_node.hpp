// Per-node D2Q9 arithmetic shared by the kernels.
//
// The opt1 functions reproduce the operation order of the reference's numpy
// opt1 path (cavity_opt1.py:100-160): with contraction off they give the same
// bits as numpy in FP32 and FP64.  Every translation unit that includes this
// header for the bit-exact path is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// Channel velocities (PyLB/Streaming.py:28-29):
// 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE.
__host__ __device__ constexpr int cx(int i) {
    return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int i) {
    return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0);
}

template <typename T>
struct Weights {
    static constexpr T w0 = T(4.0 / 9.0);
    static constexpr T w1 = T(1.0 / 9.0);
    static constexpr T w5 = T(1.0 / 36.0);
    // 6 * w_SE rounded in T, as numpy evaluates 6*w_i[D.SE] (cavity_opt1.py:212)
    static constexpr T lid = T(6) * T(1.0 / 36.0);
};

// f_eq of one node in the opt1 order (cavity_opt1.py:119-132).
template <typename T>
__host__ __device__ __forceinline__ void equilibrium_opt1(T rho, T ux, T uy, T feq[9]) {
    const T one = T(1), three = T(3), c45 = T(4.5), c15 = T(1.5);
    const T cu5 = ux + uy;
    const T cu6 = (-ux) + uy;
    const T cu7 = (-ux) - uy;
    const T cu8 = ux - uy;
    const T uu = ux * ux + uy * uy;
    const T uu15 = c15 * uu;
    const T r0 = Weights<T>::w0 * rho;
    const T r1 = Weights<T>::w1 * rho;
    const T r5 = Weights<T>::w5 * rho;
    feq[0] = r0 * (one - uu15);
    feq[1] = r1 * (((one + three * ux) + c45 * (ux * ux)) - uu15);
    feq[2] = r1 * (((one + three * uy) + c45 * (uy * uy)) - uu15);
    feq[3] = r1 * (((one - three * ux) + c45 * (ux * ux)) - uu15);
    feq[4] = r1 * (((one - three * uy) + c45 * (uy * uy)) - uu15);
    feq[5] = r5 * (((one + three * cu5) + c45 * (cu5 * cu5)) - uu15);
    feq[6] = r5 * (((one + three * cu6) + c45 * (cu6 * cu6)) - uu15);
    feq[7] = r5 * (((one + three * cu7) + c45 * (cu7 * cu7)) - uu15);
    feq[8] = r5 * (((one + three * cu8) + c45 * (cu8 * cu8)) - uu15);
}

// Moments in the opt1 order (cavity_opt1.py:156-158): rho summed f0..f8
// left to right, as np.sum over the outer axis does.
template <typename T>
__host__ __device__ __forceinline__ void moments_opt1(const T g[9], T &rho, T &ux, T &uy) {
    T r = g[0] + g[1];
    r = r + g[2];
    r = r + g[3];
    r = r + g[4];
    r = r + g[5];
    r = r + g[6];
    r = r + g[7];
    r = r + g[8];
    rho = r;
    ux = (((((g[1] - g[3]) + g[5]) - g[6]) - g[7]) + g[8]) / r;
    uy = (((((g[2] - g[4]) + g[5]) + g[6]) - g[7]) - g[8]) / r;
}

// BGK collide of one node in place (cavity_opt1.py:159): g += omega (feq - g).
template <typename T>
__host__ __device__ __forceinline__ void collide_opt1(T g[9], T omega, T &rho, T &ux, T &uy) {
    moments_opt1(g, rho, ux, uy);
    T feq[9];
    equilibrium_opt1(rho, ux, uy, feq);
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = g[i] + omega * (feq[i] - g[i]);
}

}  // namespace lbm
