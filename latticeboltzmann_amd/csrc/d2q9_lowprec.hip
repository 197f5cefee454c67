// Reduced-precision population storage for gfx950.
//
// Populations are stored as scaled deviations from the rest state,
//     d_i = scale * (f_i - w_i),
// in FP32, BF16 or FP16.  The deviations are small (|f_i - w_i| << w_i), so a
// 16-bit float keeps ~3 significant digits of the deviation instead of ~3
// digits of f_i itself, and the scale lifts FP16 values out of its subnormal
// range.  The step computes in FP32 on the deviations:
//     rho = 1 + sum_i d_i,  j = sum_i c_i d_i   (sum_i w_i = 1, sum_i c_i w_i = 0)
//     d_eq_i = w_i (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u))
//     d_i += omega (d_eq_i - d_i)
// Bounce-back is unchanged on deviations (w_i = w_opp(i)); the lid density
// is 1 + the sum of the nine deviations the reference sums (their weights add
// to 1).  FP64 is used only for the mass and residual reductions and for
// encode/decode.  No reference counterpart: the reference stores f itself
// in FP32 or FP64 (cavity_opt1.py:59).

#include "d2q9_dev.hpp"
#include "d2q9_grid.hpp"
#include "d2q9_reduce.hpp"
#include "lbm_internal.hpp"

namespace lbm {

// CT is the compute type of the collide: float (the default form) or double.
template <int FMT, typename CT, int BC, bool MOM, int BS, int V, bool NTL, bool NTS>
__global__ __launch_bounds__(BS) void k_dev_step(const typename Codec<FMT>::raw *__restrict__ src,
                                                 typename Codec<FMT>::raw *__restrict__ dst, int nx,
                                                 int ny, Layout L, int nseg, CT omega, CT u0,
                                                 CT scale, CT inv_scale, int stochastic,
                                                 uint32_t seed, float *__restrict__ ux_o,
                                                 float *__restrict__ uy_o) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    const int b = blockIdx.x;
    const int k = b / nseg;
    const int seg = b - k * nseg;
    const PullRows<R> rows(src, L, k, nx);
    const int64_t rk = int64_t(k) * L.rs;
    const int l0 = seg * (V * BS) + threadIdx.x;

    R raw[V][9];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int l = l0 + v * BS;
        if (l < ny) pull9<NTL>(rows, l, ny, raw[v]);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int l = l0 + v * BS;
        if (l >= ny) continue;
        CT g[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = CT(C::dec(raw[v][i])) * inv_scale;
        if constexpr (BC == LBM_BC_CAVITY)
            cavity_walls_dev<FMT, CT>(g, src + rk + l, L.cs, k, l, nx, ny, u0, inv_scale);
        CT ux, uy;
        collide_dev<CT>(g, omega, ux, uy);
        R *d = dst + rk + l;
        R out[9];
        Dither dither(seed, uint64_t(k) * uint64_t(ny) + uint64_t(l));
        store_values<FMT, CT>(g, scale, stochastic, dither, out);
#pragma unroll
        for (int i = 0; i < 9; ++i) st<NTS>(d + i * L.cs, out[i]);
        if constexpr (MOM) {
            const int64_t j = int64_t(k) * ny + l;
            ux_o[j] = float(ux);
            uy_o[j] = float(uy);
        }
    }
}

// K fused steps per launch on stored deviations (temporal blocking, K = 2..4),
// the structure of k_stepk: phase 1 computes the first step of the tile and a
// halo ring K-1 deep from src into LDS, each later phase the next step on a
// ring one node shallower, the last phase writes the tile to dst.  LDS holds
// the stored (rounded) values, so every intermediate step is rounded, with the
// dither of its own step (seed + j * seed_stride for step j of the launch) and
// node, exactly as a single-step launch rounds it: the result is bitwise that
// of K k_dev_step launches.  Requires nx >= TR + 2K, ny >= TC + 2K.
template <int FMT, typename CT, int BC, int TR, int TC, int H, int BS>
__device__ __forceinline__ void dev_phase_global(const typename Codec<FMT>::raw *__restrict__ src,
                                                 typename Codec<FMT>::raw *out, int k0, int l0,
                                                 int nx, int ny, Layout L, CT omega, CT u0,
                                                 CT scale, CT inv_scale, int stochastic,
                                                 uint32_t seed) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    constexpr int W = TC + 2 * H, N = (TR + 2 * H) * W;
    constexpr int V = N > BS ? 2 : 1;  // both pulls of a thread before either collide
    for (int n0 = threadIdx.x; n0 < N; n0 += V * BS) {
        R raw[V][9];
        int kv[V], lv[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int n = n0 + v * BS;
            const int r = n / W, c = n - r * W;
            kv[v] = wrap1(k0 - H + r, nx);
            lv[v] = wrap1(l0 - H + c, ny);
            if (n < N) {
                const PullRows<R> rows(src, L, kv[v], nx);
                pull9<false>(rows, lv[v], ny, raw[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int n = n0 + v * BS;
            if (n >= N) continue;
            const int k = kv[v], l = lv[v];
            CT g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = CT(C::dec(raw[v][i])) * inv_scale;
            if constexpr (BC == LBM_BC_CAVITY)
                cavity_walls_dev<FMT, CT>(g, src + int64_t(k) * L.rs + l, L.cs, k, l, nx, ny, u0,
                                          inv_scale);
            CT ux, uy;
            collide_dev<CT>(g, omega, ux, uy);
            Dither dither(seed, uint64_t(k) * uint64_t(ny) + uint64_t(l));
            R q[9];
            store_values<FMT, CT>(g, scale, stochastic, dither, q);
#pragma unroll
            for (int i = 0; i < 9; ++i) out[i * N + n] = q[i];
        }
    }
}

// In place in the one LDS buffer: every thread pulls its nodes (at most two)
// and applies the walls into registers, waits at a barrier, then collides,
// rounds and writes the smaller ring over the old one.
template <int FMT, typename CT, int BC, int TR, int TC, int H, int BS>
__device__ __forceinline__ void dev_phase_lds(typename Codec<FMT>::raw *buf, int k0, int l0, int nx,
                                              int ny, CT omega, CT u0, CT scale, CT inv_scale,
                                              int stochastic, uint32_t seed) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    constexpr int Wi = TC + 2 * (H + 1), Ni = (TR + 2 * (H + 1)) * Wi;
    constexpr int W = TC + 2 * H, N = (TR + 2 * H) * W;
    constexpr int V = (N + BS - 1) / BS;
    static_assert(V <= 2, "an in-place LDS phase holds at most two nodes per thread");
    CT g[V][9];
    int kv[V], lv[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = threadIdx.x + v * BS;
        const int r = n / W, c = n - r * W;
        kv[v] = wrap1(k0 - H + r, nx);
        lv[v] = wrap1(l0 - H + c, ny);
        if (n >= N) continue;
        const int m = (r + 1) * Wi + (c + 1);
        R raw[9];
        pull_lds(buf, Ni, Wi, m, raw);
#pragma unroll
        for (int i = 0; i < 9; ++i) g[v][i] = CT(C::dec(raw[i])) * inv_scale;
        if constexpr (BC == LBM_BC_CAVITY)
            cavity_walls_dev<FMT, CT>(g[v], buf + m, int64_t(Ni), kv[v], lv[v], nx, ny, u0,
                                      inv_scale);
    }
    __syncthreads();  // every read of the old ring is done
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = threadIdx.x + v * BS;
        if (n >= N) continue;
        CT ux, uy;
        collide_dev<CT>(g[v], omega, ux, uy);
        Dither dither(seed, uint64_t(kv[v]) * uint64_t(ny) + uint64_t(lv[v]));
        R raw[9];
        store_values<FMT, CT>(g[v], scale, stochastic, dither, raw);
#pragma unroll
        for (int i = 0; i < 9; ++i) buf[i * N + n] = raw[i];
    }
}

template <int FMT, typename CT, int BC, bool MOM, int TR, int TC, int BS>
__device__ __forceinline__ void dev_phase_to_global(const typename Codec<FMT>::raw *in,
                                                    typename Codec<FMT>::raw *__restrict__ dst,
                                                    int k0, int l0, int nx, int ny, Layout L,
                                                    CT omega, CT u0, CT scale, CT inv_scale,
                                                    int stochastic, uint32_t seed,
                                                    float *__restrict__ ux_o,
                                                    float *__restrict__ uy_o) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    constexpr int Wi = TC + 2, Ni = (TR + 2) * Wi;
    for (int n = threadIdx.x; n < TR * TC; n += BS) {
        const int r = n / TC, c = n - r * TC;
        const int k = k0 + r, l = l0 + c;
        if (k >= nx || l >= ny) continue;
        const int m = (r + 1) * Wi + (c + 1);
        R raw[9];
        pull_lds(in, Ni, Wi, m, raw);
        CT g[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = CT(C::dec(raw[i])) * inv_scale;
        if constexpr (BC == LBM_BC_CAVITY)
            cavity_walls_dev<FMT, CT>(g, in + m, int64_t(Ni), k, l, nx, ny, u0, inv_scale);
        CT ux, uy;
        collide_dev<CT>(g, omega, ux, uy);
        Dither dither(seed, uint64_t(k) * uint64_t(ny) + uint64_t(l));
        store_values<FMT, CT>(g, scale, stochastic, dither, raw);
        R *d = dst + int64_t(k) * L.rs + l;
#pragma unroll
        for (int i = 0; i < 9; ++i) d[i * L.cs] = raw[i];
        if constexpr (MOM) {
            const int64_t j = int64_t(k) * ny + l;
            ux_o[j] = float(ux);
            uy_o[j] = float(uy);
        }
    }
}

// Phases that keep the unrounded compute-type values between steps (the
// store-every-K form: populations reach the storage format once per launch).
template <int FMT, typename CT, int BC, int TR, int TC, int H, int BS>
__device__ __forceinline__ void dev_phase_global_ct(const typename Codec<FMT>::raw *__restrict__ src,
                                                    CT *out, int k0, int l0, int nx, int ny,
                                                    Layout L, CT omega, CT u0, CT inv_scale,
                                                    bool edge) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    constexpr int W = TC + 2 * H, N = (TR + 2 * H) * W;
    constexpr int V = N > BS ? 2 : 1;  // both pulls of a thread before either collide
    for (int n0 = threadIdx.x; n0 < N; n0 += V * BS) {
        R raw[V][9];
        int kv[V], lv[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int n = n0 + v * BS;
            const int r = n / W, c = n - r * W;
            kv[v] = wrap1(k0 - H + r, nx);
            lv[v] = wrap1(l0 - H + c, ny);
            if (n < N) {
                const PullRows<R> rows(src, L, kv[v], nx);
                pull9<false>(rows, lv[v], ny, raw[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int n = n0 + v * BS;
            if (n >= N) continue;
            const int k = kv[v], l = lv[v];
            CT g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = CT(C::dec(raw[v][i])) * inv_scale;
            if constexpr (BC == LBM_BC_CAVITY)
                if (edge)
                    cavity_walls_dev<FMT, CT>(g, src + int64_t(k) * L.rs + l, L.cs, k, l, nx, ny,
                                              u0, inv_scale);
            CT ux, uy;
            collide_dev<CT>(g, omega, ux, uy);
#pragma unroll
            for (int i = 0; i < 9; ++i) out[i * N + n] = g[i];
        }
    }
}

template <typename CT, int BC, int TR, int TC, int H, int BS>
__device__ __forceinline__ void dev_phase_lds_ct(CT *buf, int k0, int l0, int nx, int ny, CT omega,
                                                 CT u0, bool edge) {
    constexpr int Wi = TC + 2 * (H + 1), Ni = (TR + 2 * (H + 1)) * Wi;
    constexpr int W = TC + 2 * H, N = (TR + 2 * H) * W;
    constexpr int V = (N + BS - 1) / BS;
    static_assert(V <= 2, "an in-place LDS phase holds at most two nodes per thread");
    CT g[V][9];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = threadIdx.x + v * BS;
        if (n >= N) continue;
        const int r = n / W, c = n - r * W;
        const int m = (r + 1) * Wi + (c + 1);
        pull_lds(buf, Ni, Wi, m, g[v]);
        if constexpr (BC == LBM_BC_CAVITY) {
            if (edge) {
                const int k = wrap1(k0 - H + r, nx), l = wrap1(l0 - H + c, ny);
                cavity_walls_ct<CT>(g[v], buf + m, int64_t(Ni), k, l, nx, ny, u0);
            }
        }
    }
    __syncthreads();  // every read of the old ring is done
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = threadIdx.x + v * BS;
        if (n >= N) continue;
        CT ux, uy;
        collide_dev<CT>(g[v], omega, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) buf[i * N + n] = g[v][i];
    }
}

template <int FMT, typename CT, int BC, bool MOM, int TR, int TC, int BS>
__device__ __forceinline__ void dev_phase_to_global_ct(const CT *in,
                                                       typename Codec<FMT>::raw *__restrict__ dst,
                                                       int k0, int l0, int nx, int ny, Layout L,
                                                       CT omega, CT u0, CT scale, int stochastic,
                                                       uint32_t seed, float *__restrict__ ux_o,
                                                       float *__restrict__ uy_o, bool edge) {
    using R = typename Codec<FMT>::raw;
    constexpr int Wi = TC + 2, Ni = (TR + 2) * Wi;
    for (int n = threadIdx.x; n < TR * TC; n += BS) {
        const int r = n / TC, c = n - r * TC;
        const int k = k0 + r, l = l0 + c;
        if (k >= nx || l >= ny) continue;
        const int m = (r + 1) * Wi + (c + 1);
        CT g[9];
        pull_lds(in, Ni, Wi, m, g);
        if constexpr (BC == LBM_BC_CAVITY)
            if (edge) cavity_walls_ct<CT>(g, in + m, int64_t(Ni), k, l, nx, ny, u0);
        CT ux, uy;
        collide_dev<CT>(g, omega, ux, uy);
        Dither dither(seed, uint64_t(k) * uint64_t(ny) + uint64_t(l));
        R raw[9];
        store_values<FMT, CT>(g, scale, stochastic, dither, raw);
        R *d = dst + int64_t(k) * L.rs + l;
#pragma unroll
        for (int i = 0; i < 9; ++i) d[i * L.cs] = raw[i];
        if constexpr (MOM) {
            const int64_t j = int64_t(k) * ny + l;
            ux_o[j] = float(ux);
            uy_o[j] = float(uy);
        }
    }
}

// Store-every-K form: the K - 1 intermediate steps stay unrounded in LDS (in
// the compute type) and only the last is rounded to the storage format, with
// the dither of the launch's last step (seed + (K-1) * seed_stride).  Fewer
// roundings than K single steps, hence different (smaller) rounding error.
template <int FMT, typename CT, int BC, bool MOM, int K, int TR, int TC, int BS>
__global__ __launch_bounds__(BS) void k_dev_stepk_once(
    const typename Codec<FMT>::raw *__restrict__ src, typename Codec<FMT>::raw *__restrict__ dst,
    int nx, int ny, Layout L, int ntc, int ntiles, CT omega, CT u0, CT scale, CT inv_scale,
    int stochastic, uint32_t seed, uint32_t seed_stride, float *__restrict__ ux_o,
    float *__restrict__ uy_o) {
    static_assert(K >= 2 && K <= 4, "2 to 4 steps per launch");
    constexpr int NA = (TR + 2 * (K - 1)) * (TC + 2 * (K - 1));  // the phase-1 ring, the largest
    __shared__ CT sa[9 * NA];
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;  // whole block: before any barrier
    const int tr = tile / ntc, tc = tile - tr * ntc;
    const int k0 = tr * TR, l0 = tc * TC;
    const uint32_t last = seed + uint32_t(K - 1) * seed_stride;
    // interior tiles (phase-1 ring reaching no wall and not wrapping) skip the walls
    const bool edge = k0 - (K - 1) <= 0 || k0 + TR + K - 2 >= nx - 1 || l0 - (K - 1) <= 0 ||
                      l0 + TC + K - 2 >= ny - 1;

    dev_phase_global_ct<FMT, CT, BC, TR, TC, K - 1, BS>(src, sa, k0, l0, nx, ny, L, omega, u0,
                                                        inv_scale, edge);
    __syncthreads();
    if constexpr (K >= 4) {
        dev_phase_lds_ct<CT, BC, TR, TC, 2, BS>(sa, k0, l0, nx, ny, omega, u0, edge);
        __syncthreads();
    }
    if constexpr (K >= 3) {
        dev_phase_lds_ct<CT, BC, TR, TC, 1, BS>(sa, k0, l0, nx, ny, omega, u0, edge);
        __syncthreads();
    }
    dev_phase_to_global_ct<FMT, CT, BC, MOM, TR, TC, BS>(sa, dst, k0, l0, nx, ny, L, omega, u0,
                                                         scale, stochastic, last, ux_o, uy_o, edge);
}

template <int FMT, typename CT, int BC, bool MOM, int K, int TR, int TC, int BS>
__global__ __launch_bounds__(BS) void k_dev_stepk(const typename Codec<FMT>::raw *__restrict__ src,
                                                  typename Codec<FMT>::raw *__restrict__ dst,
                                                  int nx, int ny, Layout L, int ntc, int ntiles,
                                                  CT omega, CT u0, CT scale, CT inv_scale,
                                                  int stochastic, uint32_t seed,
                                                  uint32_t seed_stride, float *__restrict__ ux_o,
                                                  float *__restrict__ uy_o) {
    using R = typename Codec<FMT>::raw;
    static_assert(K >= 2 && K <= 4, "2 to 4 steps per launch");
    constexpr int NA = (TR + 2 * (K - 1)) * (TC + 2 * (K - 1));  // the phase-1 ring, the largest
    __shared__ R sa[9 * NA];
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;  // whole block: before any barrier
    const int tr = tile / ntc, tc = tile - tr * ntc;
    const int k0 = tr * TR, l0 = tc * TC;
    const uint32_t s1 = seed + seed_stride, s2 = s1 + seed_stride, s3 = s2 + seed_stride;

    dev_phase_global<FMT, CT, BC, TR, TC, K - 1, BS>(src, sa, k0, l0, nx, ny, L, omega, u0, scale,
                                                     inv_scale, stochastic, seed);
    __syncthreads();
    if constexpr (K >= 4) {
        dev_phase_lds<FMT, CT, BC, TR, TC, 2, BS>(sa, k0, l0, nx, ny, omega, u0, scale, inv_scale,
                                                  stochastic, s1);
        __syncthreads();
    }
    if constexpr (K >= 3) {
        dev_phase_lds<FMT, CT, BC, TR, TC, 1, BS>(sa, k0, l0, nx, ny, omega, u0, scale, inv_scale,
                                                  stochastic, K == 4 ? s2 : s1);
        __syncthreads();
    }
    dev_phase_to_global<FMT, CT, BC, MOM, TR, TC, BS>(sa, dst, k0, l0, nx, ny, L, omega, u0, scale,
                                                      inv_scale, stochastic,
                                                      K == 2 ? s1 : (K == 3 ? s2 : s3), ux_o, uy_o);
}

// Grid and numbering as k_sweep.  Requires ny >= 64 and nx >= K.
template <int FMT, typename CT, int BC, bool MOM, int K, bool EACH>
__global__ __launch_bounds__(256) void k_dev_sweep(
    const typename Codec<FMT>::raw *__restrict__ src, typename Codec<FMT>::raw *__restrict__ dst,
    int nx, int ny, Layout L, int nstrips, int seg_rows, int nwaves, int nblocks, CT omega, CT u0,
    CT scale, CT inv_scale, int stochastic, uint32_t seed, uint32_t seed_stride,
    float *__restrict__ ux_o, float *__restrict__ uy_o) {
    const int b = xcd_tile(nblocks);
    const int wave = __builtin_amdgcn_readfirstlane(b * 4 + int(threadIdx.x >> 6));
    if (b >= nblocks || wave >= nwaves) return;  // whole waves; nothing waits for them
    const int seg = wave / nstrips, strip = wave - seg * nstrips;
    const int ks = seg * seg_rows;
    const int ke = min(ks + seg_rows, nx);
    const DevSweepLane ln = dev_sweep_lane<K>(strip, ny, int(threadIdx.x & 63));
    uint32_t seeds[K];  // the dither seed of each step of the launch
#pragma unroll
    for (int j = 0; j < K; ++j) seeds[j] = seed + uint32_t(j) * seed_stride;
    dev_sweep_run<FMT, CT, BC, MOM, K, EACH>(src, dst, ks, ke, nx, ny, L, ln, omega, u0, scale,
                                             inv_scale, stochastic, seeds, 0, ux_o, uy_o);
}

// FP64 populations (reference layout) -> stored deviations (any layout).
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_dev_encode(const double *__restrict__ f,
                                                       typename Codec<FMT>::raw *__restrict__ d,
                                                       int nx, int ny, Layout Ld, int nseg,
                                                       double scale) {
    const int64_t b = blockIdx.x;
    const int64_t row = b / nseg;  // row = i * nx + k
    const int l = int(b - row * nseg) * kBlock + threadIdx.x;
    if (l >= ny) return;
    const int i = int(row / nx);
    const int64_t k = row - int64_t(i) * nx;
    const double x = (f[(int64_t(i) * nx + k) * ny + l] - wd(i)) * scale;
    d[i * Ld.cs + k * Ld.rs + l] = Codec<FMT>::enc64(x);
}

// Stored deviations (any layout) -> FP64 populations (reference layout).
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_dev_decode(const typename Codec<FMT>::raw *__restrict__ d,
                                                       double *__restrict__ f, int nx, int ny,
                                                       Layout Ld, int nseg, double inv_scale) {
    const int64_t b = blockIdx.x;
    const int64_t row = b / nseg;
    const int l = int(b - row * nseg) * kBlock + threadIdx.x;
    if (l >= ny) return;
    const int i = int(row / nx);
    const int64_t k = row - int64_t(i) * nx;
    const double x = double(Codec<FMT>::dec(d[i * Ld.cs + k * Ld.rs + l]));
    f[(int64_t(i) * nx + k) * ny + l] = wd(i) + x * inv_scale;
}

// Sum of the n stored values / scale, in FP64 (pass 1; pass 2 is
// launch_sum_pass2, which adds the rest-state mass).
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_dev_sum_pass1(const typename Codec<FMT>::raw *__restrict__ d,
                                                          int64_t n, double inv_scale,
                                                          double *__restrict__ partial) {
    double s = 0.0, c = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride)
        neumaier_add(s, c, double(Codec<FMT>::dec(d[j])) * inv_scale);
    const double tot = block_sum(s + c);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Velocity residual on stored deviations, velocity in FP64.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_dev_residual_pass1(
    const typename Codec<FMT>::raw *__restrict__ d, int nx, int ny, Layout L, double inv_scale,
    double *__restrict__ uxp, double *__restrict__ uyp, double *__restrict__ partial) {
    double m = 0.0;
    const int64_t n = int64_t(nx) * ny;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride) {
        const int64_t k = j / ny, l = j - k * ny;
        const typename Codec<FMT>::raw *c = d + k * L.rs + l;
        double g[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = double(Codec<FMT>::dec(c[i * L.cs])) * inv_scale;
        const double rho = 1.0 + (((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + (g[7] + g[8]))));
        const double ux = (((g[1] - g[3]) + (g[5] - g[6])) + (g[8] - g[7])) / rho;
        const double uy = (((g[2] - g[4]) + (g[5] + g[6])) - (g[7] + g[8])) / rho;
        const double dx = ux - uxp[j], dy = uy - uyp[j];
        m = monitor_max(m, sqrt(dx * dx + dy * dy));
        uxp[j] = ux;
        uyp[j] = uy;
    }
    const double tot = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// Launch shapes of the reduced-precision step:
// (threads per block, nodes per thread, non-temporal loads, non-temporal stores).
#define LBM_DEV_SHAPES(X)        \
    X(0, 256, 1, false, true)    \
    X(1, 256, 2, false, true)    \
    X(2, 256, 4, false, true)    \
    X(3, 512, 1, false, true)    \
    X(4, 512, 2, false, true)    \
    X(5, 128, 4, false, true)    \
    X(6, 256, 4, false, false)   \
    X(7, 1024, 1, false, true)   \
    X(8, 256, 8, false, true)

#define LBM_DEV_COUNT(id, bs, v, ntl, nts) +1
constexpr int kNumDevShapes = 0 LBM_DEV_SHAPES(LBM_DEV_COUNT);
constexpr int kDefaultDevShape = 1;

inline bool valid_fmt(int fmt) {
    return fmt == LBM_FMT_F32 || fmt == LBM_FMT_BF16 || fmt == LBM_FMT_F16;
}
inline int64_t fmt_bytes(int fmt) { return fmt == LBM_FMT_F32 ? 4 : 2; }

// Scalar arguments of one reduced-precision step.
struct DevArgs {
    int nx, ny;
    Layout L;
    double omega, u0, scale;
    int stochastic;
    uint32_t seed;
    float *ux, *uy;
};

template <int FMT, typename CT, int BC, bool MOM, int BS, int V, bool NTL, bool NTS>
void launch_dev_v(const void *src, void *dst, const DevArgs &a, hipStream_t s) {
    using R = typename Codec<FMT>::raw;
    const int nseg = (a.ny + V * BS - 1) / (V * BS);
    hipLaunchKernelGGL((k_dev_step<FMT, CT, BC, MOM, BS, V, NTL, NTS>),
                       dim3(unsigned(int64_t(a.nx) * nseg)), dim3(BS), 0, s,
                       static_cast<const R *>(src), static_cast<R *>(dst), a.nx, a.ny, a.L, nseg,
                       CT(a.omega), CT(a.u0), CT(a.scale), CT(1) / CT(a.scale), a.stochastic,
                       a.seed, a.ux, a.uy);
}

template <int FMT, typename CT, int BC, bool MOM>
void launch_dev_bc(const void *src, void *dst, const DevArgs &a, int shape, hipStream_t s) {
#define LBM_DEV_LAUNCH(id, bs, v, ntl, nts)                              \
    case id:                                                            \
        launch_dev_v<FMT, CT, BC, MOM, bs, v, ntl, nts>(src, dst, a, s); \
        break;
    switch (shape) {
        LBM_DEV_SHAPES(LBM_DEV_LAUNCH)
        default: break;
    }
#undef LBM_DEV_LAUNCH
}

template <int FMT, typename CT>
void launch_dev_fmt(const void *src, void *dst, int bc, const DevArgs &a, int shape,
                    hipStream_t s) {
    const bool mom = a.ux != nullptr;
    if (bc == LBM_BC_CAVITY) {
        if (mom) launch_dev_bc<FMT, CT, LBM_BC_CAVITY, true>(src, dst, a, shape, s);
        else launch_dev_bc<FMT, CT, LBM_BC_CAVITY, false>(src, dst, a, shape, s);
    } else {
        if (mom) launch_dev_bc<FMT, CT, LBM_BC_PERIODIC, true>(src, dst, a, shape, s);
        else launch_dev_bc<FMT, CT, LBM_BC_PERIODIC, false>(src, dst, a, shape, s);
    }
}

// Every storage format with an FP32 or an FP64 collide.
void launch_dev_compute(const void *src, void *dst, int fmt, int compute, int bc,
                        const DevArgs &a, int shape, hipStream_t s) {
    if (compute == LBM_COMPUTE_F64) {
        switch (fmt) {
            case LBM_FMT_F32: launch_dev_fmt<LBM_FMT_F32, double>(src, dst, bc, a, shape, s); break;
            case LBM_FMT_BF16: launch_dev_fmt<LBM_FMT_BF16, double>(src, dst, bc, a, shape, s); break;
            default: launch_dev_fmt<LBM_FMT_F16, double>(src, dst, bc, a, shape, s); break;
        }
        return;
    }
    switch (fmt) {
        case LBM_FMT_F32: launch_dev_fmt<LBM_FMT_F32, float>(src, dst, bc, a, shape, s); break;
        case LBM_FMT_BF16: launch_dev_fmt<LBM_FMT_BF16, float>(src, dst, bc, a, shape, s); break;
        default: launch_dev_fmt<LBM_FMT_F16, float>(src, dst, bc, a, shape, s); break;
    }
}

// Shapes of the multi-step deviation kernel, indexed by `variant`: (steps per
// launch K, tile rows TR, tile columns TC, threads per block BS).
#define LBM_DEV_MULTI_SHAPES(X) \
    X(0, 2, 8, 64, 256)         \
    X(1, 2, 16, 64, 512)        \
    X(2, 3, 12, 32, 512)        \
    X(3, 3, 16, 32, 512)        \
    X(4, 4, 8, 64, 512)         \
    X(5, 4, 12, 32, 512)        \
    X(6, 4, 16, 32, 512)        \
    X(7, 4, 24, 32, 512)        \
    X(8, 4, 32, 32, 1024)       \
    X(9, 4, 40, 32, 1024)       \
    X(10, 4, 46, 32, 1024)
#define LBM_DEV_COUNT_MULTI(id, k, tr, tc, bs) +1
constexpr int kNumDevTileShapes = 0 LBM_DEV_MULTI_SHAPES(LBM_DEV_COUNT_MULTI);

// Shapes of the register sweep on deviations (k_dev_sweep), variants
// kNumDevTileShapes + id: (steps per launch K, rows per wave segment, 0 for
// balanced segments as balanced_seg_rows).  An FP64 collide keeps twice the
// registers per value: K > 4 takes the FP32 collide only.
#define LBM_DEV_SWEEP_SHAPES(X) \
    X(0, 4, 32)                 \
    X(1, 4, 64)                 \
    X(2, 6, 64)                 \
    X(3, 8, 64)                 \
    X(4, 4, 0)                  \
    X(5, 6, 0)
#define LBM_DEV_COUNT_SWEEP(id, k, seg) +1
constexpr int kNumDevSweepShapes = 0 LBM_DEV_SWEEP_SHAPES(LBM_DEV_COUNT_SWEEP);
constexpr int kNumDevMultiShapes = kNumDevTileShapes + kNumDevSweepShapes;
// Four steps on 46 x 32 tiles with 1024 threads: the fastest measured for FP32
// and FP16 storage (at 8192^2, FP16 stored every 4 steps 230 GLUPS and FP32
// storage 224 in a same-run comparison).  Its four rings (52x38 .. 46x32 nodes)
// take two passes of the block each, with few idle lanes in the second.
constexpr int kDefaultDevMultiShape = 10;

template <int FMT, typename CT, int BC, bool MOM, int K, int TR, int TC, int BS>
int launch_dev_multi_v(const void *src, void *dst, const DevArgs &a, uint32_t seed_stride,
                       bool round_each, hipStream_t s) {
    using R = typename Codec<FMT>::raw;
    if (a.nx < TR + 2 * K || a.ny < TC + 2 * K) return LBM_E_SIZE;
    const int64_t ntr = (a.nx + TR - 1) / TR, ntc = (a.ny + TC - 1) / TC;
    const int64_t ntiles = ntr * ntc;
    const int64_t nblocks = ((ntiles + 7) / 8) * 8;
    if (nblocks > 0x7fffffffLL) return LBM_E_SIZE;
    if (round_each)
        hipLaunchKernelGGL((k_dev_stepk<FMT, CT, BC, MOM, K, TR, TC, BS>), dim3(unsigned(nblocks)),
                           dim3(BS), 0, s, static_cast<const R *>(src), static_cast<R *>(dst),
                           a.nx, a.ny, a.L, int(ntc), int(ntiles), CT(a.omega), CT(a.u0),
                           CT(a.scale), CT(1) / CT(a.scale), a.stochastic, a.seed, seed_stride,
                           a.ux, a.uy);
    else {
        // unrounded values in the compute type: an FP64 tile can outgrow the
        // 160 KB of LDS of a gfx950 CU, and such shapes are refused
        constexpr int64_t NA = int64_t(TR + 2 * (K - 1)) * (TC + 2 * (K - 1));
        constexpr int64_t lds = 9 * NA * int64_t(sizeof(CT));
        if constexpr (lds <= 160 * 1024) {
            hipLaunchKernelGGL((k_dev_stepk_once<FMT, CT, BC, MOM, K, TR, TC, BS>),
                               dim3(unsigned(nblocks)), dim3(BS), 0, s, static_cast<const R *>(src),
                               static_cast<R *>(dst), a.nx, a.ny, a.L, int(ntc), int(ntiles),
                               CT(a.omega), CT(a.u0), CT(a.scale), CT(1) / CT(a.scale),
                               a.stochastic, a.seed, seed_stride, a.ux, a.uy);
        } else {
            return LBM_E_ARG;
        }
    }
    return LBM_OK;
}

template <int FMT, typename CT, int BC, bool MOM, int K>
int launch_dev_sweep_v(const void *src, void *dst, const DevArgs &a, uint32_t seed_stride,
                       bool round_each, int seg_rows, hipStream_t s) {
    using R = typename Codec<FMT>::raw;
    if (a.nx < 2 * K || a.ny < 64) return LBM_E_SIZE;
    if constexpr (sizeof(CT) == 8 && K > 4) {
        return LBM_E_ARG;
    } else {
        const int64_t nstrips = dev_sweep_strips<K>(a.ny);
        if (seg_rows == 0) {  // balanced segments (balanced_seg_rows)
            static std::atomic<int> per_cu_each[64], per_cu_once[64];
            const int64_t resident =
                round_each ? resident_waves(k_dev_sweep<FMT, CT, BC, MOM, K, true>, per_cu_each)
                           : resident_waves(k_dev_sweep<FMT, CT, BC, MOM, K, false>, per_cu_once);
            seg_rows = balanced_seg_rows(a.nx, nstrips, resident, 2 * (K - 1));
        }
        const int64_t nsegs = (a.nx + seg_rows - 1) / seg_rows;
        const int64_t nwaves = nstrips * nsegs;
        const int64_t nblocks = (nwaves + 3) / 4;
        const int64_t grid = ((nblocks + 7) / 8) * 8;
        if (grid > 0x7fffffffLL) return LBM_E_SIZE;
        if (round_each)
            hipLaunchKernelGGL((k_dev_sweep<FMT, CT, BC, MOM, K, true>), dim3(unsigned(grid)),
                               dim3(256), 0, s, static_cast<const R *>(src), static_cast<R *>(dst),
                               a.nx, a.ny, a.L, int(nstrips), seg_rows, int(nwaves), int(nblocks),
                               CT(a.omega), CT(a.u0), CT(a.scale), CT(1) / CT(a.scale),
                               a.stochastic, a.seed, seed_stride, a.ux, a.uy);
        else
            hipLaunchKernelGGL((k_dev_sweep<FMT, CT, BC, MOM, K, false>), dim3(unsigned(grid)),
                               dim3(256), 0, s, static_cast<const R *>(src), static_cast<R *>(dst),
                               a.nx, a.ny, a.L, int(nstrips), seg_rows, int(nwaves), int(nblocks),
                               CT(a.omega), CT(a.u0), CT(a.scale), CT(1) / CT(a.scale),
                               a.stochastic, a.seed, seed_stride, a.ux, a.uy);
        return LBM_OK;
    }
}

template <int FMT, typename CT, int BC, bool MOM>
int launch_dev_multi_bc(const void *src, void *dst, const DevArgs &a, uint32_t stride, bool each,
                        int shape, hipStream_t s) {
#define LBM_DEV_LAUNCH_MULTI(id, k, tr, tc, bs) \
    case id: return launch_dev_multi_v<FMT, CT, BC, MOM, k, tr, tc, bs>(src, dst, a, stride, each, s);
#define LBM_DEV_LAUNCH_SWEEP(id, k, seg) \
    case kNumDevTileShapes + id: return launch_dev_sweep_v<FMT, CT, BC, MOM, k>(src, dst, a, stride, each, seg, s);
    switch (shape) {
        LBM_DEV_MULTI_SHAPES(LBM_DEV_LAUNCH_MULTI)
        LBM_DEV_SWEEP_SHAPES(LBM_DEV_LAUNCH_SWEEP)
        default: return LBM_E_ARG;
    }
#undef LBM_DEV_LAUNCH_MULTI
#undef LBM_DEV_LAUNCH_SWEEP
}

template <int FMT, typename CT>
int launch_dev_multi_fmt(const void *src, void *dst, int bc, const DevArgs &a, uint32_t stride,
                         bool each, int shape, hipStream_t s) {
    const bool mom = a.ux != nullptr;
    if (bc == LBM_BC_CAVITY)
        return mom ? launch_dev_multi_bc<FMT, CT, LBM_BC_CAVITY, true>(src, dst, a, stride, each, shape, s)
                   : launch_dev_multi_bc<FMT, CT, LBM_BC_CAVITY, false>(src, dst, a, stride, each, shape, s);
    return mom ? launch_dev_multi_bc<FMT, CT, LBM_BC_PERIODIC, true>(src, dst, a, stride, each, shape, s)
               : launch_dev_multi_bc<FMT, CT, LBM_BC_PERIODIC, false>(src, dst, a, stride, each, shape, s);
}

int launch_dev_multi(const void *src, void *dst, int fmt, int compute, int bc, const DevArgs &a,
                     uint32_t stride, bool each, int shape, hipStream_t s) {
    if (compute == LBM_COMPUTE_F64) {
        switch (fmt) {
            case LBM_FMT_F32: return launch_dev_multi_fmt<LBM_FMT_F32, double>(src, dst, bc, a, stride, each, shape, s);
            case LBM_FMT_BF16: return launch_dev_multi_fmt<LBM_FMT_BF16, double>(src, dst, bc, a, stride, each, shape, s);
            default: return launch_dev_multi_fmt<LBM_FMT_F16, double>(src, dst, bc, a, stride, each, shape, s);
        }
    }
    switch (fmt) {
        case LBM_FMT_F32: return launch_dev_multi_fmt<LBM_FMT_F32, float>(src, dst, bc, a, stride, each, shape, s);
        case LBM_FMT_BF16: return launch_dev_multi_fmt<LBM_FMT_BF16, float>(src, dst, bc, a, stride, each, shape, s);
        default: return launch_dev_multi_fmt<LBM_FMT_F16, float>(src, dst, bc, a, stride, each, shape, s);
    }
}

inline int64_t dev_blocks(int64_t nx, int64_t ny) { return 9 * nx * ((ny + kBlock - 1) / kBlock); }

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_dev_variant_count(void) { return kNumDevShapes; }

int lbm_dev_encode(const double *f, void *d, int64_t nx, int64_t ny, int layout, int fmt,
                   double scale, void *stream) {
    if (!f || !d) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_layout(layout) || !valid_fmt(fmt) || !(scale > 0.0)) return LBM_E_ARG;
    if (overlaps(f, 9 * nx * ny * 8, d, 9 * nx * ny * fmt_bytes(fmt))) return LBM_E_ALIAS;
    const int64_t nb = dev_blocks(nx, ny);
    if (nb > 0x7fffffffLL) return LBM_E_SIZE;
    const Layout L = make_layout(layout, nx, ny);
    const int nseg = int(nb / (9 * nx));
    hipStream_t s = as_stream(stream);
    switch (fmt) {
        case LBM_FMT_F32:
            hipLaunchKernelGGL(k_dev_encode<LBM_FMT_F32>, dim3(unsigned(nb)), dim3(kBlock), 0, s, f,
                               static_cast<float *>(d), int(nx), int(ny), L, nseg, scale);
            break;
        case LBM_FMT_BF16:
            hipLaunchKernelGGL(k_dev_encode<LBM_FMT_BF16>, dim3(unsigned(nb)), dim3(kBlock), 0, s, f,
                               static_cast<uint16_t *>(d), int(nx), int(ny), L, nseg, scale);
            break;
        default:
            hipLaunchKernelGGL(k_dev_encode<LBM_FMT_F16>, dim3(unsigned(nb)), dim3(kBlock), 0, s, f,
                               static_cast<uint16_t *>(d), int(nx), int(ny), L, nseg, scale);
            break;
    }
    return launch_status();
}

int lbm_dev_decode(const void *d, double *f, int64_t nx, int64_t ny, int layout, int fmt,
                   double scale, void *stream) {
    if (!f || !d) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_layout(layout) || !valid_fmt(fmt) || !(scale > 0.0)) return LBM_E_ARG;
    if (overlaps(f, 9 * nx * ny * 8, d, 9 * nx * ny * fmt_bytes(fmt))) return LBM_E_ALIAS;
    const int64_t nb = dev_blocks(nx, ny);
    if (nb > 0x7fffffffLL) return LBM_E_SIZE;
    const Layout L = make_layout(layout, nx, ny);
    const int nseg = int(nb / (9 * nx));
    hipStream_t s = as_stream(stream);
    switch (fmt) {
        case LBM_FMT_F32:
            hipLaunchKernelGGL(k_dev_decode<LBM_FMT_F32>, dim3(unsigned(nb)), dim3(kBlock), 0, s,
                               static_cast<const float *>(d), f, int(nx), int(ny), L, nseg, 1.0 / scale);
            break;
        case LBM_FMT_BF16:
            hipLaunchKernelGGL(k_dev_decode<LBM_FMT_BF16>, dim3(unsigned(nb)), dim3(kBlock), 0, s,
                               static_cast<const uint16_t *>(d), f, int(nx), int(ny), L, nseg, 1.0 / scale);
            break;
        default:
            hipLaunchKernelGGL(k_dev_decode<LBM_FMT_F16>, dim3(unsigned(nb)), dim3(kBlock), 0, s,
                               static_cast<const uint16_t *>(d), f, int(nx), int(ny), L, nseg, 1.0 / scale);
            break;
    }
    return launch_status();
}

int lbm_dev_step(const void *src, void *dst, int fmt, int64_t nx, int64_t ny, int layout,
                 double omega, int bc, double u0, double scale, int compute, int rounding,
                 uint32_t seed, float *ux, float *uy, int variant, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if ((ux == nullptr) != (uy == nullptr)) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_fmt(fmt) || !valid_layout(layout) || !(scale > 0.0)) return LBM_E_ARG;
    if (bc != LBM_BC_PERIODIC && bc != LBM_BC_CAVITY) return LBM_E_ARG;
    if (compute != LBM_COMPUTE_F32 && compute != LBM_COMPUTE_F64) return LBM_E_ARG;
    if (rounding != LBM_ROUND_NEAREST && rounding != LBM_ROUND_STOCHASTIC) return LBM_E_ARG;
    if (variant < -1 || variant >= kNumDevShapes) return LBM_E_ARG;
    const int64_t bytes = 9 * nx * ny * fmt_bytes(fmt);
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    if (nx * ((ny + kBlock - 1) / kBlock) > 0x7fffffffLL) return LBM_E_SIZE;
    const int shape = variant < 0 ? kDefaultDevShape : variant;
    const DevArgs a{int(nx), int(ny), make_layout(layout, nx, ny), omega, u0, scale,
                    rounding == LBM_ROUND_STOCHASTIC ? 1 : 0, seed, ux, uy};
    hipStream_t s = as_stream(stream);
    launch_dev_compute(src, dst, fmt, compute, bc, a, shape, s);
    return launch_status();
}

int lbm_dev_multi_variant_count(void) { return kNumDevMultiShapes; }

int lbm_dev_multi_default_variant(void) { return kDefaultDevMultiShape; }

int lbm_dev_multi_info(int variant, int *steps, int64_t *min_nx, int64_t *min_ny) {
    if (!steps || !min_nx || !min_ny) return LBM_E_NULL;
    const int shape = variant < 0 ? kDefaultDevMultiShape : variant;
#define LBM_DEV_INFO_MULTI(id, k, tr, tc, bs) \
    case id: *steps = k; *min_nx = tr + 2 * k; *min_ny = tc + 2 * k; return LBM_OK;
#define LBM_DEV_INFO_SWEEP(id, k, seg) \
    case kNumDevTileShapes + id: *steps = k; *min_nx = 2 * k; *min_ny = 64; return LBM_OK;
    switch (shape) {
        LBM_DEV_MULTI_SHAPES(LBM_DEV_INFO_MULTI)
        LBM_DEV_SWEEP_SHAPES(LBM_DEV_INFO_SWEEP)
        default: return LBM_E_ARG;
    }
#undef LBM_DEV_INFO_MULTI
#undef LBM_DEV_INFO_SWEEP
}

int lbm_dev_step_multi(const void *src, void *dst, int fmt, int64_t nx, int64_t ny, int layout,
                       double omega, int bc, double u0, double scale, int compute, int rounding,
                       uint32_t seed, uint32_t seed_stride, int round_each, float *ux, float *uy,
                       int variant, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if ((ux == nullptr) != (uy == nullptr)) return LBM_E_NULL;
    if (!valid_grid(nx, ny) || nx > INT32_MAX || ny > INT32_MAX) return LBM_E_SIZE;
    if (!valid_fmt(fmt) || !valid_layout(layout) || !(scale > 0.0)) return LBM_E_ARG;
    if (bc != LBM_BC_PERIODIC && bc != LBM_BC_CAVITY) return LBM_E_ARG;
    if (compute != LBM_COMPUTE_F32 && compute != LBM_COMPUTE_F64) return LBM_E_ARG;
    if (rounding != LBM_ROUND_NEAREST && rounding != LBM_ROUND_STOCHASTIC) return LBM_E_ARG;
    if (variant < -1 || variant >= kNumDevMultiShapes) return LBM_E_ARG;
    const int64_t bytes = 9 * nx * ny * fmt_bytes(fmt);
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    const int shape = variant < 0 ? kDefaultDevMultiShape : variant;
    const DevArgs a{int(nx), int(ny), make_layout(layout, nx, ny), omega, u0, scale,
                    rounding == LBM_ROUND_STOCHASTIC ? 1 : 0, seed, ux, uy};
    if (round_each != 0 && round_each != 1) return LBM_E_ARG;
    const int rc = launch_dev_multi(src, dst, fmt, compute, bc, a, seed_stride, round_each != 0,
                                    shape, as_stream(stream));
    return rc ? rc : launch_status();
}

int lbm_dev_mass(const void *d, int64_t nx, int64_t ny, int fmt, double scale, double *work,
                 double *out, void *stream) {
    if (!d || !work || !out) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_fmt(fmt) || !(scale > 0.0)) return LBM_E_ARG;
    hipStream_t s = as_stream(stream);
    const int64_t n = 9 * nx * ny;
    switch (fmt) {
        case LBM_FMT_F32:
            hipLaunchKernelGGL(k_dev_sum_pass1<LBM_FMT_F32>, dim3(kReduceBlocks), dim3(kBlock), 0, s,
                               static_cast<const float *>(d), n, 1.0 / scale, work);
            break;
        case LBM_FMT_BF16:
            hipLaunchKernelGGL(k_dev_sum_pass1<LBM_FMT_BF16>, dim3(kReduceBlocks), dim3(kBlock), 0, s,
                               static_cast<const uint16_t *>(d), n, 1.0 / scale, work);
            break;
        default:
            hipLaunchKernelGGL(k_dev_sum_pass1<LBM_FMT_F16>, dim3(kReduceBlocks), dim3(kBlock), 0, s,
                               static_cast<const uint16_t *>(d), n, 1.0 / scale, work);
            break;
    }
    launch_sum_pass2(work, kReduceBlocks, double(nx * ny), out, s);
    return launch_status();
}

int lbm_dev_velocity_residual(const void *d, int64_t nx, int64_t ny, int layout, int fmt,
                              double scale, double *ux_prev, double *uy_prev, double *work,
                              double *out, void *stream) {
    if (!d || !ux_prev || !uy_prev || !work || !out) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_fmt(fmt) || !valid_layout(layout) || !(scale > 0.0)) return LBM_E_ARG;
    hipStream_t s = as_stream(stream);
    const Layout L = make_layout(layout, nx, ny);
    const double inv = 1.0 / scale;
    switch (fmt) {
        case LBM_FMT_F32:
            hipLaunchKernelGGL(k_dev_residual_pass1<LBM_FMT_F32>, dim3(kReduceBlocks), dim3(kBlock), 0,
                               s, static_cast<const float *>(d), int(nx), int(ny), L, inv, ux_prev,
                               uy_prev, work);
            break;
        case LBM_FMT_BF16:
            hipLaunchKernelGGL(k_dev_residual_pass1<LBM_FMT_BF16>, dim3(kReduceBlocks), dim3(kBlock), 0,
                               s, static_cast<const uint16_t *>(d), int(nx), int(ny), L, inv, ux_prev,
                               uy_prev, work);
            break;
        default:
            hipLaunchKernelGGL(k_dev_residual_pass1<LBM_FMT_F16>, dim3(kReduceBlocks), dim3(kBlock), 0,
                               s, static_cast<const uint16_t *>(d), int(nx), int(ny), L, inv, ux_prev,
                               uy_prev, work);
            break;
    }
    launch_max_pass2(work, kReduceBlocks, out, s);
    return launch_status();
}

}  // extern "C"
