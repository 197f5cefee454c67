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

// Pulls of node m of a tile held in LDS as planes of N values, rows W wide
// (the multi-step kernels): channel i from the neighbour at -c_i.
template <typename T>
__device__ __forceinline__ void pull_lds(const T *in, int N, int W, int m, T g[9]) {
    g[0] = in[m];
    g[1] = in[1 * N + m - W];      // E  from (k-1, l)
    g[2] = in[2 * N + m - 1];      // N  from (k,   l-1)
    g[3] = in[3 * N + m + W];      // W  from (k+1, l)
    g[4] = in[4 * N + m + 1];      // S  from (k,   l+1)
    g[5] = in[5 * N + m - W - 1];  // NE from (k-1, l-1)
    g[6] = in[6 * N + m + W - 1];  // NW from (k+1, l-1)
    g[7] = in[7 * N + m + W + 1];  // SW from (k+1, l+1)
    g[8] = in[8 * N + m - W + 1];  // SE from (k-1, l+1)
}

// Tile of this block for the multi-step kernels, numbered so that the blocks
// on one XCD (the hardware deals blocks to the 8 XCDs round-robin) take a
// contiguous band of tiles, whose halos then meet in that XCD's L2.  The
// grid holds ceil(ntiles / 8) * 8 blocks; a result >= ntiles means no tile.
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per = (ntiles + 7) >> 3;
    return int(blockIdx.x & 7) * per + int(blockIdx.x >> 3);
}

// Wrap a coordinate that is at most one period outside [0, n).
__device__ __forceinline__ int wrap1(int k, int n) { return k < 0 ? k + n : (k >= n ? k - n : k); }

}  // namespace lbm
