// Grid addressing shared by the step kernels: population layouts, the nine
// pull addresses of a node, and cache-hinted loads and stores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_d2q9.h"

namespace lbm {

template <bool NT, typename T>
__device__ __forceinline__ T ld(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st(T *p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Where element (i, k, l) lives: i * cs + k * rs + l.
//   reference layout f[i][k][l] (LBM_LAYOUT_SOA):  cs = nx * ny, rs = ny
//   row-interleaved  f[k][i][l] (LBM_LAYOUT_ROWS): cs = ny,      rs = 9 * ny
// In the row-interleaved layout the nine channel segments of one x-row are
// adjacent, so the 9 reads and 9 writes of a block fall into a few MB of
// address space instead of nine planes hundreds of MB apart.
struct Layout {
    int64_t cs, rs;
};

__host__ __device__ inline Layout make_layout(int kind, int64_t nx, int64_t ny) {
    return kind == LBM_LAYOUT_ROWS ? Layout{ny, 9 * ny} : Layout{nx * ny, ny};
}

// Row base pointers of the nine pulls for row k.
template <typename T>
struct PullRows {
    const T *p[9];
    __device__ __forceinline__ PullRows(const T *src, Layout L, int k, int nx) {
        const int km = k == 0 ? nx - 1 : k - 1;
        const int kp = k == nx - 1 ? 0 : k + 1;
        const int64_t rk = int64_t(k) * L.rs, rm = int64_t(km) * L.rs, rp = int64_t(kp) * L.rs;
        p[0] = src + rk;
        p[1] = src + L.cs + rm;          // E  pulls from (k-1, l)
        p[2] = src + 2 * L.cs + rk;      // N  pulls from (k,   l-1)
        p[3] = src + 3 * L.cs + rp;      // W  pulls from (k+1, l)
        p[4] = src + 4 * L.cs + rk;      // S  pulls from (k,   l+1)
        p[5] = src + 5 * L.cs + rm;      // NE pulls from (k-1, l-1)
        p[6] = src + 6 * L.cs + rp;      // NW pulls from (k+1, l-1)
        p[7] = src + 7 * L.cs + rp;      // SW pulls from (k+1, l+1)
        p[8] = src + 8 * L.cs + rm;      // SE pulls from (k-1, l+1)
    }
};

template <bool NT, typename T>
__device__ __forceinline__ void pull9(const PullRows<T> &r, int l, int ny, T g[9]) {
    const int lm = l == 0 ? ny - 1 : l - 1;
    const int lp = l == ny - 1 ? 0 : l + 1;
    g[0] = ld<NT>(r.p[0] + l);
    g[1] = ld<NT>(r.p[1] + l);
    g[2] = ld<NT>(r.p[2] + lm);
    g[3] = ld<NT>(r.p[3] + l);
    g[4] = ld<NT>(r.p[4] + lp);
    g[5] = ld<NT>(r.p[5] + lm);
    g[6] = ld<NT>(r.p[6] + lm);
    g[7] = ld<NT>(r.p[7] + lp);
    g[8] = ld<NT>(r.p[8] + lp);
}

}  // namespace lbm
