// D2Q9 BGK kernels for gfx950 in FP64 and FP32: equilibrium, collide,
// periodic streaming, streaming with cavity walls, the fused pull
// stream-collide step, and deterministic FP64 reductions.
//
// Built with -ffp-contract=off: every product is rounded before it is added,
// as in numpy, so the FP64 results match the reference's numpy opt1 path bit
// for bit (cavity_opt1.py:100-245).
//
// Layout: f[i][k][l] (channel, x, y) with y contiguous.  One thread owns
// V nodes of one x-row, V*kBlock consecutive y values per block, so every
// load and store instruction of a wave touches 64 consecutive elements of one
// plane.  The pull of channel i reads plane i at (k - cx_i, l - cy_i): whole
// rows k-1, k, k+1 shifted by at most one element, so the reads stay
// coalesced and every population is read exactly once per step.

#include <mutex>

#include "d2q9_grid.hpp"
#include "d2q9_node.hpp"
#include "d2q9_reduce.hpp"
#include "lbm_internal.hpp"

#pragma clang fp contract(off)

namespace lbm {

// Cavity walls on the pulled values g of node (k, l), applied in the order of
// cavity_opt1.py:201-225: bottom, lid (density from the post-bottom values),
// left, right.  c points at the node's own pre-stream value of channel 0;
// channel i of the node is c[i * cs].
template <typename T>
__device__ __forceinline__ void cavity_walls(T g[9], const T *c, int64_t cs, int k, int l,
                                             int nx, int ny, T u0) {
    if (l == 0) {  // bottom: N <- S, NE <- SW, NW <- SE
        g[2] = c[4 * cs];
        g[5] = c[7 * cs];
        g[6] = c[8 * cs];
    }
    if (l == ny - 1) {  // moving lid
        const T c2 = c[2 * cs], c5 = c[5 * cs], c6 = c[6 * cs];
        T rl = c6 + c2;
        rl = rl + c5;
        rl = rl + g[6];
        rl = rl + g[2];
        rl = rl + g[5];
        rl = rl + g[3];
        rl = rl + g[0];
        rl = rl + g[1];
        const T t = (Weights<T>::lid * rl) * u0;
        g[4] = c2;
        g[8] = c6 + t;
        g[7] = c5 - t;
    }
    if (k == 0) {  // left: E <- W, NE <- SW, SE <- NW
        g[1] = c[3 * cs];
        g[5] = c[7 * cs];
        g[8] = c[6 * cs];
    }
    if (k == nx - 1) {  // right: W <- E, NW <- SE, SW <- NE
        g[3] = c[1 * cs];
        g[6] = c[8 * cs];
        g[7] = c[5 * cs];
    }
}

// ---------------------------------------------------------------------------
// Fused step: pull (+ walls) + collide, src -> dst.
// grid = nx * nseg blocks of BS threads, nseg = ceil(ny / (V * BS)); block b
// owns row k = b / nseg and V * BS consecutive y values of it.  NTL / NTS
// mark the loads / stores non-temporal (each population is touched once per
// step, so there is nothing to keep in cache beyond the one-element overlap
// of neighbouring pulls).
template <typename T, int BC, bool MOM, int BS, int V, bool NTL, bool NTS>
__global__ __launch_bounds__(BS) void k_step(const T *__restrict__ src, T *__restrict__ dst,
                                             int nx, int ny, Layout L, int nseg, T omega, T u0,
                                             T *__restrict__ rho_o, T *__restrict__ ux_o,
                                             T *__restrict__ uy_o) {
    const int b = blockIdx.x;
    const int k = b / nseg;
    const int seg = b - k * nseg;
    const PullRows<T> rows(src, L, k, nx);
    const int64_t rk = int64_t(k) * L.rs;
    const int l0 = seg * (V * BS) + threadIdx.x;

    T g[V][9];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int l = l0 + v * BS;
        if (l < ny) pull9<NTL>(rows, l, ny, g[v]);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int l = l0 + v * BS;
        if (l >= ny) continue;
        if constexpr (BC == LBM_BC_CAVITY) cavity_walls(g[v], src + rk + l, L.cs, k, l, nx, ny, u0);
        T rho, ux, uy;
        collide_opt1(g[v], omega, rho, ux, uy);
        T *d = dst + rk + l;
#pragma unroll
        for (int i = 0; i < 9; ++i) st<NTS>(d + i * L.cs, g[v][i]);
        if constexpr (MOM) {
            const int64_t j = int64_t(k) * ny + l;
            rho_o[j] = rho;
            ux_o[j] = ux;
            uy_o[j] = uy;
        }
    }
}

// K fused steps per launch (temporal blocking, K = 2..4), src -> dst.
// Block b owns a TR x TC tile of nodes.  Phase 1 computes step 1 on the tile
// and a halo ring K-1 nodes deep, at wrapped global coordinates, with the
// single-step node code (pull from src, walls, collide), into LDS; each later
// phase computes the next step on a ring one node shallower from the previous
// phase's LDS values, in place in the one LDS buffer; the last phase writes the
// tile to dst.  The wrapped halo also supplies the x-wrapped lid densities of
// the two top corners (cavity_opt1.py:206-213), so the result is bitwise that
// of K single-step launches, while HBM sees one read and one write per
// population per K steps plus the halo re-reads that L2 does not absorb.
// Tiles are numbered so that the blocks on one XCD (blockIdx % 8) take a
// contiguous band of tiles, whose halos then meet in that XCD's L2.
// Requires nx >= TR + 2K, ny >= TC + 2K.

// One phase from global memory: step of the nodes of a halo-H tile ring.
// When the ring has more nodes than the block has threads, a thread takes two
// nodes per pass and issues both nodes' pulls before either collide, so the
// threads that own a second node do not hold the block at the next barrier
// for a second, serial round of HBM latency.
template <typename T, int BC, int TR, int TC, int H, int BS>
__device__ __forceinline__ void phase_from_global(const T *__restrict__ src, T *out, int k0, int l0,
                                                  int nx, int ny, Layout L, T omega, T u0,
                                                  bool edge) {
    constexpr int W = TC + 2 * H, N = (TR + 2 * H) * W;
    constexpr int V = N > BS ? 2 : 1;
    for (int n0 = threadIdx.x; n0 < N; n0 += V * BS) {
        T g[V][9];
        int kv[V], lv[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int n = n0 + v * BS;
            const int r = n / W, c = n - r * W;
            int k = k0 - H + r, l = l0 - H + c;
            k = k < 0 ? k + nx : (k >= nx ? k - nx : k);
            l = l < 0 ? l + ny : (l >= ny ? l - ny : l);
            kv[v] = k;
            lv[v] = l;
            if (n < N) {
                const PullRows<T> rows(src, L, k, nx);
                pull9<false>(rows, l, ny, g[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int n = n0 + v * BS;
            if (n >= N) continue;
            const int k = kv[v], l = lv[v];
            if constexpr (BC == LBM_BC_CAVITY)
                if (edge) cavity_walls(g[v], src + int64_t(k) * L.rs + l, L.cs, k, l, nx, ny, u0);
            T rho, ux, uy;
            collide_opt1(g[v], omega, rho, ux, uy);
#pragma unroll
            for (int i = 0; i < 9; ++i) out[i * N + n] = g[v][i];
        }
    }
}

// One phase in place in LDS: halo-(H+1) values in, halo-H values out, in the
// same buffer.  Every thread first pulls its nodes (at most V = 2) and applies
// the walls, which read only the buffer, into registers; after a barrier it
// collides them and writes the smaller ring over the old one.  One buffer
// instead of two halves the LDS of a block.
template <typename T, int BC, int TR, int TC, int H, int BS>
__device__ __forceinline__ void phase_lds(T *buf, int k0, int l0, int nx, int ny, T omega, T u0,
                                          bool edge) {
    constexpr int Wi = TC + 2 * (H + 1), Ni = (TR + 2 * (H + 1)) * Wi;
    constexpr int W = TC + 2 * H, N = (TR + 2 * H) * W;
    constexpr int V = (N + BS - 1) / BS;
    static_assert(V <= 2, "an in-place LDS phase holds at most two nodes per thread");
    T g[V][9];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = threadIdx.x + v * BS;
        if (n >= N) continue;
        const int r = n / W, c = n - r * W;
        const int m = (r + 1) * Wi + (c + 1);
        pull_lds(buf, Ni, Wi, m, g[v]);
        if constexpr (BC == LBM_BC_CAVITY) {
            if (edge) {
                int k = k0 - H + r, l = l0 - H + c;
                k = k < 0 ? k + nx : (k >= nx ? k - nx : k);
                l = l < 0 ? l + ny : (l >= ny ? l - ny : l);
                cavity_walls(g[v], buf + m, int64_t(Ni), k, l, nx, ny, u0);
            }
        }
    }
    __syncthreads();  // every read of the old ring is done
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = threadIdx.x + v * BS;
        if (n >= N) continue;
        T rho, ux, uy;
        collide_opt1(g[v], omega, rho, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) buf[i * N + n] = g[v][i];
    }
}

// Last phase: the tile from halo-1 LDS values to dst.
template <typename T, int BC, bool MOM, int TR, int TC, int BS, bool NTS>
__device__ __forceinline__ void phase_to_global(const T *in, T *__restrict__ dst, int k0, int l0,
                                                int nx, int ny, Layout L, T omega, T u0,
                                                T *__restrict__ rho_o, T *__restrict__ ux_o,
                                                T *__restrict__ uy_o, bool edge) {
    constexpr int Wi = TC + 2, Ni = (TR + 2) * Wi;
    for (int n = threadIdx.x; n < TR * TC; n += BS) {
        const int r = n / TC, c = n - r * TC;
        const int k = k0 + r, l = l0 + c;
        if (k >= nx || l >= ny) continue;
        const int m = (r + 1) * Wi + (c + 1);
        T g[9];
        pull_lds(in, Ni, Wi, m, g);
        if constexpr (BC == LBM_BC_CAVITY)
            if (edge) cavity_walls(g, in + m, int64_t(Ni), k, l, nx, ny, u0);
        T rho, ux, uy;
        collide_opt1(g, omega, rho, ux, uy);
        T *d = dst + int64_t(k) * L.rs + l;
#pragma unroll
        for (int i = 0; i < 9; ++i) st<NTS>(d + i * L.cs, g[i]);
        if constexpr (MOM) {
            const int64_t j = int64_t(k) * ny + l;
            rho_o[j] = rho;
            ux_o[j] = ux;
            uy_o[j] = uy;
        }
    }
}

template <typename T, int BC, bool MOM, int K, int TR, int TC, int BS, bool NTS>
__global__ __launch_bounds__(BS) void k_stepk(const T *__restrict__ src, T *__restrict__ dst,
                                                  int nx, int ny, Layout L, int ntc, int ntiles,
                                                  T omega, T u0, T *__restrict__ rho_o,
                                                  T *__restrict__ ux_o, T *__restrict__ uy_o) {
    static_assert(K >= 2 && K <= 4, "2 to 4 steps per launch");
    constexpr int NA = (TR + 2 * (K - 1)) * (TC + 2 * (K - 1));  // the phase-1 ring, the largest
    __shared__ T sa[9 * NA];
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;  // whole block: before any barrier
    const int tr = tile / ntc, tc = tile - tr * ntc;
    const int k0 = tr * TR, l0 = tc * TC;
    // whether the phase-1 ring (the largest) reaches a wall row or column or
    // wraps across the domain's edge; the later rings lie inside it, so
    // interior tiles skip the wall code in every phase (a block-uniform branch)
    const bool edge = k0 - (K - 1) <= 0 || k0 + TR + K - 2 >= nx - 1 || l0 - (K - 1) <= 0 ||
                      l0 + TC + K - 2 >= ny - 1;

    phase_from_global<T, BC, TR, TC, K - 1, BS>(src, sa, k0, l0, nx, ny, L, omega, u0, edge);
    __syncthreads();
    if constexpr (K >= 4) {
        phase_lds<T, BC, TR, TC, 2, BS>(sa, k0, l0, nx, ny, omega, u0, edge);
        __syncthreads();
    }
    if constexpr (K >= 3) {
        phase_lds<T, BC, TR, TC, 1, BS>(sa, k0, l0, nx, ny, omega, u0, edge);
        __syncthreads();
    }
    phase_to_global<T, BC, MOM, TR, TC, BS, NTS>(sa, dst, k0, l0, nx, ny, L, omega, u0, rho_o, ux_o, uy_o,
                                            edge);
}

// ---------------------------------------------------------------------------
// K fused steps per launch by a sweep along x, the intermediate steps kept in
// registers (k_sweep).  A wave owns 64 consecutive columns (y) and a segment
// of rows (x) and walks the rows in order.  Iteration i computes step 1 of
// row i + K - 1 from src (pull, walls, collide: the single-step node code),
// then step j of row i + K - j from step j - 1's rows i + K - j - 1 .. + 1 for
// j = 2..K, and writes step K's row i to dst.  Each lane keeps, for every
// intermediate step, its column's values of the two newest rows that step has
// made, in two register buffers used in turn (the newest row goes over the
// oldest once the next step has pulled from it), so no values are copied
// between iterations.  Column neighbours (l - 1, l + 1) come from the
// adjacent lanes through ds_bpermute (the LDS crossbar: no LDS memory is
// allocated).  Step j is valid on lanes j - 1 .. 64 - j, so a wave writes the
// middle 64 - 2 (K - 1) columns; its first 2 (K - 1) iterations fill the
// pipeline.  Rows and columns are wrapped like the tile kernel's halo, so the
// walls, the lid and its x-wrapped corner densities come out bitwise as K
// single-step launches.  Waves never wait for each other: no barrier, no LDS.

// Value of `v` held by the lane whose byte address (4 * lane) is `addr`.
__device__ __forceinline__ double lane_fetch(double v, int addr) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_ds_bpermute(addr, int(unsigned(u))));
    const unsigned hi = unsigned(__builtin_amdgcn_ds_bpermute(addr, int(unsigned(u >> 32))));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ float lane_fetch(float v, int addr) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}

template <typename T>
struct SweepLane {
    int l;                 // the lane's column (wrapped)
    int from_lo, from_hi;  // ds_bpermute addresses of lanes - 1 and + 1
    bool writer;           // the lane's column is one the wave writes
    bool yedge;            // wave-uniform: the wave's columns include l = 0 or l = ny - 1
};

// cavity_walls for the sweep: the same rules in the same order, with the
// bottom and lid rules behind the wave-uniform test `yedge` (whether the
// wave's columns, halo lanes included, reach l = 0 or l = ny - 1), so
// interior waves branch past them instead of evaluating them for every lane.
template <typename T>
__device__ __forceinline__ void sweep_walls(T g[9], const T *c, int64_t cs, int k, int l, int nx,
                                            int ny, T u0, bool yedge) {
    if (yedge) {
        if (l == 0) {
            g[2] = c[4 * cs];
            g[5] = c[7 * cs];
            g[6] = c[8 * cs];
        }
        if (l == ny - 1) {
            const T c2 = c[2 * cs], c5 = c[5 * cs], c6 = c[6 * cs];
            T rl = c6 + c2;
            rl = rl + c5;
            rl = rl + g[6];
            rl = rl + g[2];
            rl = rl + g[5];
            rl = rl + g[3];
            rl = rl + g[0];
            rl = rl + g[1];
            const T t = (Weights<T>::lid * rl) * u0;
            g[4] = c2;
            g[8] = c6 + t;
            g[7] = c5 - t;
        }
    }
    if (k == 0) {
        g[1] = c[3 * cs];
        g[5] = c[7 * cs];
        g[8] = c[6 * cs];
    }
    if (k == nx - 1) {
        g[3] = c[1 * cs];
        g[6] = c[8 * cs];
        g[7] = c[5 * cs];
    }
}

// Pulls of the step-1 row i + K - 1 from src (the loads of one iteration).
template <typename T, int K>
__device__ __forceinline__ void sweep_pull(const T *__restrict__ src, int i, int nx, int ny,
                                           Layout L, const SweepLane<T> &ln, T g[9]) {
    const PullRows<T> rows(src, L, wrap1(i + K - 1, nx), nx);
    pull9<false>(rows, ln.l, ny, g);
}

// Register state of the sweep.  Step j (1 .. K-1) keeps its two newest rows.
// PP (ping-pong): two full buffers per step used in turn, S[PH][j] the older
// row and S[PH ^ 1][j] the newer; the row made next goes over the older one,
// so nothing is copied.  Otherwise the newer row is S[0][j] and the older is
// kept only in the three channels a pull takes from it (E, NE, SE: O[j]), and
// the rows move down by register copies: a third less state for 12 moves
// per step and iteration.  P holds step-1 pulls: with PF (prefetch) the
// next iteration's are issued into P[PH ^ 1] before this iteration's
// arithmetic, so their HBM latency overlaps it.
template <typename T, int K>
struct SweepState {
    T S[2][K - 1][9];
    T O[K - 1][3];
    T P[2][9];
};

template <typename T, int BC, bool MOM, int K, bool PP, bool PF, bool NTS, int PH>
__device__ __forceinline__ void sweep_iteration(const T *__restrict__ src, T *__restrict__ dst,
                                                int i, int ks, int ke, int nx, int ny, Layout L,
                                                const SweepLane<T> &ln, T omega, T u0,
                                                SweepState<T, K> &st, T *__restrict__ rho_o,
                                                T *__restrict__ ux_o, T *__restrict__ uy_o) {
    constexpr int cur = PF ? PH : 0;
    if constexpr (PF) {
        if (i + 1 < ke) sweep_pull<T, K>(src, i + 1, nx, ny, L, ln, st.P[PH ^ 1]);
    } else {
        sweep_pull<T, K>(src, i, nx, ny, L, ln, st.P[0]);
    }
    // step 1 of row i + K - 1, from its pulls
    T g[9];
    {
        const int k = wrap1(i + K - 1, nx);
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = st.P[cur][q];
        if constexpr (BC == LBM_BC_CAVITY)
            sweep_walls(g, src + int64_t(k) * L.rs + ln.l, L.cs, k, ln.l, nx, ny, u0, ln.yedge);
        T rho, ux, uy;
        collide_opt1(g, omega, rho, ux, uy);
    }
#pragma unroll
    for (int j = 2; j <= K; ++j) {
        // step j - 1's rows: older (r - 1), newer (r), and g (r + 1, made just now)
        T(&newer)[9] = PP ? st.S[PH ^ 1][j - 2] : st.S[0][j - 2];
        const T o1 = PP ? st.S[PH][j - 2][1] : st.O[j - 2][0];
        const T o5 = PP ? st.S[PH][j - 2][5] : st.O[j - 2][1];
        const T o8 = PP ? st.S[PH][j - 2][8] : st.O[j - 2][2];
        const bool active = i >= ks - 2 * (K - j);  // wave-uniform: the pipeline has filled
        T h[9];
        int k = 0;
        if (active) {
            k = wrap1(i + K - j, nx);
            h[0] = newer[0];
            h[1] = o1;                                 // E  from (k-1, l)
            h[2] = lane_fetch(newer[2], ln.from_lo);   // N  from (k,   l-1)
            h[3] = g[3];                               // W  from (k+1, l)
            h[4] = lane_fetch(newer[4], ln.from_hi);   // S  from (k,   l+1)
            h[5] = lane_fetch(o5, ln.from_lo);         // NE from (k-1, l-1)
            h[6] = lane_fetch(g[6], ln.from_lo);       // NW from (k+1, l-1)
            h[7] = lane_fetch(g[7], ln.from_hi);       // SW from (k+1, l+1)
            h[8] = lane_fetch(o8, ln.from_hi);         // SE from (k-1, l+1)
            if constexpr (BC == LBM_BC_CAVITY) sweep_walls(h, newer, 1, k, ln.l, nx, ny, u0, ln.yedge);
        }
        // the rows of step j - 1 move down: the new row replaces the older
        if constexpr (PP) {
#pragma unroll
            for (int q = 0; q < 9; ++q) st.S[PH][j - 2][q] = g[q];
        } else {
            st.O[j - 2][0] = newer[1];
            st.O[j - 2][1] = newer[5];
            st.O[j - 2][2] = newer[8];
#pragma unroll
            for (int q = 0; q < 9; ++q) newer[q] = g[q];
        }
        if (!active) return;  // later steps are not active either
        T rho, ux, uy;
        collide_opt1(h, omega, rho, ux, uy);
        if (j == K) {
            if (ln.writer && i >= ks) {
                T *d = dst + int64_t(k) * L.rs + ln.l;
#pragma unroll
                for (int q = 0; q < 9; ++q) lbm::st<NTS>(d + q * L.cs, h[q]);
                if constexpr (MOM) {
                    const int64_t m = int64_t(k) * ny + ln.l;
                    rho_o[m] = rho;
                    ux_o[m] = ux;
                    uy_o[m] = uy;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] = h[q];
        }
    }
}

// Grid: blocks of 4 waves, numbered XCD-contiguously (xcd_tile over blocks);
// wave w takes segment w / nstrips (rows [seg * seg_rows, + seg_rows)) and
// strip w % nstrips (columns [strip * (64 - 2(K-1)), + 64 - 2(K-1))).
// Requires ny >= 64 and nx >= K.
template <typename T, int BC, bool MOM, int K, bool PP, bool PF, bool NTS>
__global__ __launch_bounds__(256) void k_sweep(const T *__restrict__ src, T *__restrict__ dst,
                                               int nx, int ny, Layout L, int nstrips, int seg_rows,
                                               int nwaves, int nblocks, T omega, T u0,
                                               T *__restrict__ rho_o, T *__restrict__ ux_o,
                                               T *__restrict__ uy_o) {
    static_assert(K >= 2 && K <= 6, "2 to 6 steps per launch");
    constexpr int VW = 64 - 2 * (K - 1);
    const int b = xcd_tile(nblocks);
    // the wave index is the same in all lanes: in a scalar register, the row
    // loop, the wall tests of a row and the pipeline tests stay scalar
    const int wave = __builtin_amdgcn_readfirstlane(b * 4 + int(threadIdx.x >> 6));
    if (b >= nblocks || wave >= nwaves) return;  // whole waves; nothing waits for them
    const int lane = int(threadIdx.x & 63);
    const int seg = wave / nstrips, strip = wave - seg * nstrips;
    const int ks = seg * seg_rows;
    const int ke = min(ks + seg_rows, nx);
    const int c = strip * VW - (K - 1) + lane;  // in [-(K-1), ny + 63): one wrap suffices
    SweepLane<T> ln;
    ln.l = wrap1(c, ny);
    ln.from_lo = ((lane + 63) & 63) << 2;
    ln.from_hi = ((lane + 1) & 63) << 2;
    ln.writer = lane >= K - 1 && lane < 64 - (K - 1) && c < ny;
    // the wave's columns c0 .. c0 + 63 (before wrapping) include l = 0 or
    // l = ny - 1 when they reach 0 or ny - 1: the first strip, and every strip
    // whose halo lanes run past the last column
    const int c0 = strip * VW - (K - 1);
    ln.yedge = c0 <= 0 || c0 + 63 >= ny - 1;
    SweepState<T, K> st;
#pragma unroll
    for (int j = 0; j < K - 1; ++j) {
#pragma unroll
        for (int q = 0; q < 9; ++q) st.S[0][j][q] = st.S[1][j][q] = T(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) st.O[j][q] = T(0);
    }
    int i = ks - 2 * (K - 1);
    if constexpr (PF) sweep_pull<T, K>(src, i, nx, ny, L, ln, st.P[0]);
    for (; i + 1 < ke; i += 2) {
        sweep_iteration<T, BC, MOM, K, PP, PF, NTS, 0>(src, dst, i, ks, ke, nx, ny, L, ln, omega, u0,
                                                  st, rho_o, ux_o, uy_o);
        sweep_iteration<T, BC, MOM, K, PP, PF, NTS, 1>(src, dst, i + 1, ks, ke, nx, ny, L, ln, omega,
                                                  u0, st, rho_o, ux_o, uy_o);
    }
    if (i < ke)
        sweep_iteration<T, BC, MOM, K, PP, PF, NTS, 0>(src, dst, i, ks, ke, nx, ny, L, ln, omega, u0,
                                                  st, rho_o, ux_o, uy_o);
}

// Streaming with cavity walls, no collide (stream_and_bounce_back).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_stream_bounce(const T *__restrict__ src,
                                                          T *__restrict__ dst, int nx, int ny,
                                                          Layout L, int nseg, T u0) {
    const int b = blockIdx.x;
    const int k = b / nseg;
    const int l = (b - k * nseg) * kBlock + threadIdx.x;
    if (l >= ny) return;
    const PullRows<T> rows(src, L, k, nx);
    const int64_t rk = int64_t(k) * L.rs;
    T g[9];
    pull9<false>(rows, l, ny, g);
    cavity_walls(g, src + rk + l, L.cs, k, l, nx, ny, u0);
#pragma unroll
    for (int i = 0; i < 9; ++i) dst[i * L.cs + rk + l] = g[i];
}

// Periodic streaming of U-sized elements (a pure permutation, PyLB.stream).
template <typename U>
__global__ __launch_bounds__(kBlock) void k_stream_periodic(const U *__restrict__ src,
                                                            U *__restrict__ dst, int nx, int ny,
                                                            Layout L, int nseg) {
    const int b = blockIdx.x;
    const int k = b / nseg;
    const int l = (b - k * nseg) * kBlock + threadIdx.x;
    if (l >= ny) return;
    const PullRows<U> rows(src, L, k, nx);
    U g[9];
    pull9<false>(rows, l, ny, g);
    const int64_t rk = int64_t(k) * L.rs;
#pragma unroll
    for (int i = 0; i < 9; ++i) dst[i * L.cs + rk + l] = g[i];
}

// Copy between layouts: grid = 9 * nx * nseg blocks.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_relayout(const U *__restrict__ src,
                                                     U *__restrict__ dst, int nx, int ny,
                                                     Layout Ls, Layout Ld, int nseg) {
    const int64_t b = blockIdx.x;
    const int64_t row = b / nseg;  // row = i * nx + k
    const int l = int(b - row * nseg) * kBlock + threadIdx.x;
    if (l >= ny) return;
    const int64_t i = row / nx, k = row - i * nx;
    dst[i * Ld.cs + k * Ld.rs + l] = src[i * Ls.cs + k * Ls.rs + l];
}

// Moments of a grid in either layout; scalar fields are [k][l].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_moments_grid(const T *__restrict__ f, int nx, int ny,
                                                         Layout L, int nseg,
                                                         T *__restrict__ rho_o,
                                                         T *__restrict__ ux_o,
                                                         T *__restrict__ uy_o) {
    const int b = blockIdx.x;
    const int k = b / nseg;
    const int l = (b - k * nseg) * kBlock + threadIdx.x;
    if (l >= ny) return;
    const T *c = f + int64_t(k) * L.rs + l;
    T g[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = c[i * L.cs];
    T rho, ux, uy;
    moments_opt1(g, rho, ux, uy);
    const int64_t j = int64_t(k) * ny + l;
    rho_o[j] = rho;
    ux_o[j] = ux;
    uy_o[j] = uy;
}

// Collide in place on f[9][n] (PyLB.collide).
template <typename T, bool MOM>
__global__ __launch_bounds__(kBlock) void k_collide(T *__restrict__ f, int64_t n, T omega,
                                                    T *__restrict__ rho_o, T *__restrict__ ux_o,
                                                    T *__restrict__ uy_o) {
    const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    T g[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = f[i * n + j];
    T rho, ux, uy;
    collide_opt1(g, omega, rho, ux, uy);
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i * n + j] = g[i];
    if constexpr (MOM) {
        rho_o[j] = rho;
        ux_o[j] = ux;
        uy_o[j] = uy;
    }
}

// f[9][n] = f_eq(rho, ux, uy) (PyLB.equilibrium).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_equilibrium(const T *__restrict__ rho,
                                                        const T *__restrict__ ux,
                                                        const T *__restrict__ uy,
                                                        T *__restrict__ f, int64_t n) {
    const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (j >= n) return;
    T feq[9];
    equilibrium_opt1(rho[j], ux[j], uy[j], feq);
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i * n + j] = feq[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sum_pass1(const T *__restrict__ x, int64_t n,
                                                      double *__restrict__ partial) {
    double s = 0.0, c = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride)
        neumaier_add(s, c, double(x[j]));
    const double tot = block_sum(s + c);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// out = bias + sum of the np partials
__global__ __launch_bounds__(kBlock) void k_sum_pass2(const double *__restrict__ partial, int np,
                                                      double bias, double *__restrict__ out) {
    double s = 0.0, c = 0.0;
    for (int j = threadIdx.x; j < np; j += kBlock) neumaier_add(s, c, partial[j]);
    const double tot = block_sum(s + c);
    if (threadIdx.x == 0) out[0] = bias + tot;
}

__global__ __launch_bounds__(kBlock) void k_max_pass2(const double *__restrict__ partial, int np,
                                                      double *__restrict__ out) {
    double m = 0.0;
    for (int j = threadIdx.x; j < np; j += kBlock) m = fmax(m, partial[j]);
    const double tot = block_max(m);
    if (threadIdx.x == 0) out[0] = tot;
}

__global__ __launch_bounds__(kBlock) void k_residual_pass1(const double *__restrict__ f, int nx,
                                                           int ny, Layout L,
                                                           double *__restrict__ uxp,
                                                           double *__restrict__ uyp,
                                                           double *__restrict__ partial) {
    double m = 0.0;
    const int64_t n = int64_t(nx) * ny;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x; j < n; j += stride) {
        const int64_t k = j / ny, l = j - k * ny;
        const double *c = f + k * L.rs + l;
        double g[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = c[i * L.cs];
        double rho, ux, uy;
        moments_opt1(g, rho, ux, uy);
        const double dx = ux - uxp[j], dy = uy - uyp[j];
        m = monitor_max(m, sqrt(dx * dx + dy * dy));
        uxp[j] = ux;
        uyp[j] = uy;
    }
    const double tot = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Plain copy of n16 16-byte words with ordinary or non-temporal stores: the
// achievable HBM copy rate on this device, the practical ceiling the step
// kernels are compared against.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ __launch_bounds__(kBlock) void k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                   int64_t n16) {
    // One 16-byte word per lane: on gfx950 this outruns unrolled forms that keep
    // two or four words per lane in flight (6.2 vs 5.6-5.8 TB/s at 4.8 GB).
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n16) st<NT>(dst + i, src[i]);
}

// ---------------------------------------------------------------------------
// Launchers

// Launch shapes of the fused step, indexed by `variant`:
// (threads per block, nodes per thread, non-temporal loads, non-temporal stores).
#define LBM_STEP_SHAPES(X)        \
    X(0, 256, 1, false, false)    \
    X(1, 256, 2, false, false)    \
    X(2, 256, 4, false, false)    \
    X(3, 256, 1, false, true)     \
    X(4, 256, 2, false, true)     \
    X(5, 256, 4, false, true)     \
    X(6, 512, 1, false, true)     \
    X(7, 512, 2, false, true)     \
    X(8, 128, 4, false, true)     \
    X(9, 256, 1, true, true)      \
    X(10, 512, 1, true, true)     \
    X(11, 1024, 1, false, true)

#define LBM_COUNT_SHAPE(id, bs, v, ntl, nts) +1
constexpr int kNumStepShapes = 0 LBM_STEP_SHAPES(LBM_COUNT_SHAPE);
constexpr int kDefaultStepShape = 6;

// Blocks of kBlock threads covering ny (one thread per node) times nx rows.
inline int64_t row_blocks(int64_t nx, int64_t ny, int per_thread = 1) {
    const int64_t span = int64_t(per_thread) * kBlock;
    return nx * ((ny + span - 1) / span);
}

template <typename T, int BC, bool MOM, int BS, int V, bool NTL, bool NTS>
void launch_step_v(const T *src, T *dst, int nx, int ny, Layout L, T omega, T u0, T *rho, T *ux,
                   T *uy, hipStream_t s) {
    const int nseg = (ny + V * BS - 1) / (V * BS);
    hipLaunchKernelGGL((k_step<T, BC, MOM, BS, V, NTL, NTS>), dim3(unsigned(int64_t(nx) * nseg)),
                       dim3(BS), 0, s, src, dst, nx, ny, L, nseg, omega, u0, rho, ux, uy);
}

template <typename T, int BC, bool MOM>
void launch_step_bc(const T *src, T *dst, int nx, int ny, Layout L, T omega, T u0, T *rho, T *ux,
                    T *uy, int shape, hipStream_t s) {
#define LBM_LAUNCH_SHAPE(id, bs, v, ntl, nts)                                                  \
    case id:                                                                                  \
        launch_step_v<T, BC, MOM, bs, v, ntl, nts>(src, dst, nx, ny, L, omega, u0, rho, ux, uy, s); \
        break;
    switch (shape) {
        LBM_STEP_SHAPES(LBM_LAUNCH_SHAPE)
        default: break;
    }
#undef LBM_LAUNCH_SHAPE
}

template <typename T>
int step_impl(const T *src, T *dst, int64_t nx, int64_t ny, int layout, T omega, int bc, T u0,
              T *rho, T *ux, T *uy, int variant, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_layout(layout)) return LBM_E_ARG;
    if (bc != LBM_BC_PERIODIC && bc != LBM_BC_CAVITY) return LBM_E_ARG;
    if (variant < -1 || variant >= kNumStepShapes) return LBM_E_ARG;
    const bool mom = rho != nullptr;
    if (mom != (ux != nullptr) || mom != (uy != nullptr)) return LBM_E_NULL;
    const int64_t bytes = 9 * nx * ny * int64_t(sizeof(T));
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    if (row_blocks(nx, ny) > 0x7fffffffLL) return LBM_E_SIZE;
    const int shape = variant < 0 ? kDefaultStepShape : variant;
    hipStream_t s = as_stream(stream);
    const int inx = int(nx), iny = int(ny);
    const Layout L = make_layout(layout, nx, ny);
    if (bc == LBM_BC_CAVITY) {
        if (mom) launch_step_bc<T, LBM_BC_CAVITY, true>(src, dst, inx, iny, L, omega, u0, rho, ux, uy, shape, s);
        else launch_step_bc<T, LBM_BC_CAVITY, false>(src, dst, inx, iny, L, omega, u0, rho, ux, uy, shape, s);
    } else {
        if (mom) launch_step_bc<T, LBM_BC_PERIODIC, true>(src, dst, inx, iny, L, omega, u0, rho, ux, uy, shape, s);
        else launch_step_bc<T, LBM_BC_PERIODIC, false>(src, dst, inx, iny, L, omega, u0, rho, ux, uy, shape, s);
    }
    return launch_status();
}

// Shapes of the multi-step kernel, indexed by `variant`: (steps per launch K,
// tile rows TR, tile columns TC, threads per block BS, non-temporal stores of
// the finished tile).
#define LBM_MULTI_SHAPES(X)              \
    X(0, 2, 8, 64, 256, false)       \
    X(1, 2, 4, 128, 256, false)      \
    X(2, 2, 16, 32, 256, false)      \
    X(3, 2, 4, 64, 256, false)       \
    X(4, 3, 8, 32, 256, false)       \
    X(5, 3, 4, 64, 256, false)       \
    X(6, 4, 8, 32, 256, false)       \
    X(7, 2, 8, 64, 512, false)       \
    X(8, 2, 8, 128, 512, false)      \
    X(9, 3, 8, 32, 512, false)       \
    X(10, 3, 12, 32, 512, false)     \
    X(11, 3, 8, 64, 512, false)      \
    X(12, 4, 8, 32, 512, false)      \
    X(13, 3, 16, 32, 512, false)     \
    X(14, 4, 12, 32, 512, false)     \
    X(15, 4, 16, 32, 512, false)     \
    X(16, 4, 10, 32, 512, false)     \
    X(17, 3, 24, 32, 512, false)     \
    X(18, 4, 20, 32, 512, false)     \
    X(19, 4, 24, 28, 512, false)     \
    X(20, 4, 22, 32, 768, false)     \
    X(21, 4, 20, 32, 768, false)     \
    X(22, 4, 20, 32, 1024, false)    \
    X(23, 4, 22, 32, 768, true)
#define LBM_COUNT_MULTI(id, k, tr, tc, bs, nts) +1
constexpr int kNumTileShapes = 0 LBM_MULTI_SHAPES(LBM_COUNT_MULTI);

// Shapes of the register sweep (k_sweep), variants kNumTileShapes + id:
// (steps per launch K, rows per wave segment (0: balanced, balanced_seg_rows),
// ping-pong buffers, prefetch, non-temporal stores of the written rows).
#define LBM_SWEEP_SHAPES(X)              \
    X(0, 4, 16, true, true, false)    \
    X(1, 4, 32, true, true, false)    \
    X(2, 4, 32, false, true, false)   \
    X(3, 4, 64, false, false, false)  \
    X(4, 5, 32, false, false, false)  \
    X(5, 5, 64, true, false, false)   \
    X(6, 5, 64, false, false, false)  \
    X(7, 5, 64, false, true, false)   \
    X(8, 6, 64, false, false, false)  \
    X(9, 3, 32, false, true, false)   \
    X(10, 5, 0, true, false, false)   \
    X(11, 4, 0, true, true, false)    \
    X(12, 5, 0, true, false, true)
#define LBM_COUNT_SWEEP(id, k, seg, pp, pf, nts) +1
constexpr int kNumSweepShapes = 0 LBM_SWEEP_SHAPES(LBM_COUNT_SWEEP);
constexpr int kNumMultiShapes = kNumTileShapes + kNumSweepShapes;
// Four steps on 22 x 32 tiles with 768 threads: the fastest FP64 shape measured
// (115 GLUPS at 2048^2, 120 at 3072^2, 124.5 at 4096^2, against 105-113, 113-115
// and 120 for 20 x 32 tiles of 512 threads in the same runs).  76.6 KB of LDS,
// so two blocks per CU, 24 waves; its rings (28x38, 26x36, 24x34, 22x32 nodes)
// take 2, 2, 2 and 1 passes.  Blocks of 12 waves give each SIMD 3 waves of a
// block; blocks whose wave count is not a multiple of 4 (576, 640 threads)
// load the SIMDs unevenly and ran at 72-78 GLUPS.  The finished tile goes to
// HBM with non-temporal stores (variant 23; 119-120 against 116-117 GLUPS at
// 2048^2 in the same run): the next launch reads it, this one does not, and
// the stores then leave L2 to the halo rows neighbouring tiles still read.
constexpr int kDefaultMultiShape = 23;
// From 4096^2 nodes on, the five-step sweep with ping-pong row buffers,
// balanced segments and non-temporal stores: 124-127 GLUPS at 4096^2 (the
// 22 x 32 tile ties) and 141-144 at 8192^2 (139-141 with plain stores) against
// about 122 for the tiles.  At 2048^2 the sweep's waves are too few to hide
// memory latency (105 against 115 GLUPS; 64 against 94 at 1024^2).
constexpr int kLargeGridMultiShape = kNumTileShapes + 12;
constexpr int64_t kLargeGridNodes = int64_t(4096) * 4096;
// FP32 lattices below 4096^2 nodes: 20 x 32 tiles with 512 threads (172-190
// GLUPS at 2048^2 against 168-174 for the 22 x 32, 768-thread tile).  A block
// holds half the LDS of the FP64 one, so four 8-wave blocks share a CU (32
// waves) where two 12-wave blocks give 24.
constexpr int kDefaultMultiShapeF32 = 18;

template <typename T>
inline int default_multi_shape(int64_t nx, int64_t ny) {
    if (nx * ny >= kLargeGridNodes) return kLargeGridMultiShape;
    return sizeof(T) == sizeof(float) ? kDefaultMultiShapeF32 : kDefaultMultiShape;
}

template <typename T, int BC, bool MOM, int K, int TR, int TC, int BS, bool NTS>
int launch_multi_v(const T *src, T *dst, int64_t nx, int64_t ny, Layout L, T omega, T u0, T *rho,
                   T *ux, T *uy, hipStream_t s) {
    if (nx < TR + 2 * K || ny < TC + 2 * K) return LBM_E_SIZE;
    const int64_t ntr = (nx + TR - 1) / TR, ntc = (ny + TC - 1) / TC;
    const int64_t ntiles = ntr * ntc;
    const int64_t nblocks = ((ntiles + 7) / 8) * 8;
    if (nblocks > 0x7fffffffLL) return LBM_E_SIZE;
    hipLaunchKernelGGL((k_stepk<T, BC, MOM, K, TR, TC, BS, NTS>), dim3(unsigned(nblocks)), dim3(BS), 0,
                       s, src, dst, int(nx), int(ny), L, int(ntc), int(ntiles), omega, u0, rho, ux,
                       uy);
    return LBM_OK;
}

template <typename T, int BC, bool MOM, int K, bool PP, bool PF, bool NTS>
int launch_sweep_v(const T *src, T *dst, int64_t nx, int64_t ny, Layout L, int seg_rows, T omega,
                   T u0, T *rho, T *ux, T *uy, hipStream_t s) {
    constexpr int VW = 64 - 2 * (K - 1);
    if (nx < 2 * K || ny < 64) return LBM_E_SIZE;
    const int64_t nstrips = (ny + VW - 1) / VW;
    if (seg_rows == 0) {
        static std::atomic<int> per_cu[64];
        seg_rows = balanced_seg_rows(nx, nstrips, resident_waves(k_sweep<T, BC, MOM, K, PP, PF, NTS>, per_cu),
                                     2 * (K - 1));
    }
    const int64_t nsegs = (nx + seg_rows - 1) / seg_rows;
    const int64_t nwaves = nstrips * nsegs;
    const int64_t nblocks = (nwaves + 3) / 4;
    const int64_t grid = ((nblocks + 7) / 8) * 8;
    if (grid > 0x7fffffffLL) return LBM_E_SIZE;
    hipLaunchKernelGGL((k_sweep<T, BC, MOM, K, PP, PF, NTS>), dim3(unsigned(grid)), dim3(256), 0, s, src, dst,
                       int(nx), int(ny), L, int(nstrips), seg_rows, int(nwaves), int(nblocks),
                       omega, u0, rho, ux, uy);
    return LBM_OK;
}

template <typename T, int BC, bool MOM>
int launch_multi_bc(const T *src, T *dst, int64_t nx, int64_t ny, Layout L, T omega, T u0, T *rho,
                    T *ux, T *uy, int shape, hipStream_t s) {
#define LBM_LAUNCH_MULTI(id, k, tr, tc, bs, nts) \
    case id: return launch_multi_v<T, BC, MOM, k, tr, tc, bs, nts>(src, dst, nx, ny, L, omega, u0, rho, ux, uy, s);
#define LBM_LAUNCH_SWEEP(id, k, seg, pp, pf, nts) \
    case kNumTileShapes + id: return launch_sweep_v<T, BC, MOM, k, pp, pf, nts>(src, dst, nx, ny, L, seg, omega, u0, rho, ux, uy, s);
    switch (shape) {
        LBM_MULTI_SHAPES(LBM_LAUNCH_MULTI)
        LBM_SWEEP_SHAPES(LBM_LAUNCH_SWEEP)
        default: return LBM_E_ARG;
    }
#undef LBM_LAUNCH_MULTI
#undef LBM_LAUNCH_SWEEP
}

template <typename T>
int multi_impl(const T *src, T *dst, int64_t nx, int64_t ny, int layout, T omega, int bc, T u0,
               T *rho, T *ux, T *uy, int variant, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (!valid_grid(nx, ny) || nx > INT32_MAX || ny > INT32_MAX) return LBM_E_SIZE;
    if (!valid_layout(layout)) return LBM_E_ARG;
    if (bc != LBM_BC_PERIODIC && bc != LBM_BC_CAVITY) return LBM_E_ARG;
    if (variant < -1 || variant >= kNumMultiShapes) return LBM_E_ARG;
    const bool mom = rho != nullptr;
    if (mom != (ux != nullptr) || mom != (uy != nullptr)) return LBM_E_NULL;
    const int64_t bytes = 9 * nx * ny * int64_t(sizeof(T));
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    const int shape = variant < 0 ? default_multi_shape<T>(nx, ny) : variant;
    hipStream_t s = as_stream(stream);
    const Layout L = make_layout(layout, nx, ny);
    int rc;
    if (bc == LBM_BC_CAVITY)
        rc = mom ? launch_multi_bc<T, LBM_BC_CAVITY, true>(src, dst, nx, ny, L, omega, u0, rho, ux, uy, shape, s)
                 : launch_multi_bc<T, LBM_BC_CAVITY, false>(src, dst, nx, ny, L, omega, u0, rho, ux, uy, shape, s);
    else
        rc = mom ? launch_multi_bc<T, LBM_BC_PERIODIC, true>(src, dst, nx, ny, L, omega, u0, rho, ux, uy, shape, s)
                 : launch_multi_bc<T, LBM_BC_PERIODIC, false>(src, dst, nx, ny, L, omega, u0, rho, ux, uy, shape, s);
    return rc ? rc : launch_status();
}

inline int multi_shape_info(int variant, int *steps, int64_t *min_nx, int64_t *min_ny) {
    const int shape = variant < 0 ? kDefaultMultiShape : variant;
#define LBM_INFO_MULTI(id, k, tr, tc, bs, nts) \
    case id: *steps = k; *min_nx = tr + 2 * k; *min_ny = tc + 2 * k; return LBM_OK;
#define LBM_INFO_SWEEP(id, k, seg, pp, pf, nts) \
    case kNumTileShapes + id: *steps = k; *min_nx = 2 * k; *min_ny = 64; return LBM_OK;
    switch (shape) {
        LBM_MULTI_SHAPES(LBM_INFO_MULTI)
        LBM_SWEEP_SHAPES(LBM_INFO_SWEEP)
        default: return LBM_E_ARG;
    }
#undef LBM_INFO_MULTI
#undef LBM_INFO_SWEEP
}

template <typename T>
int stream_bounce_impl(const T *src, T *dst, int64_t nx, int64_t ny, T u0, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    const int64_t bytes = 9 * nx * ny * int64_t(sizeof(T));
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    const int64_t nb = row_blocks(nx, ny);
    if (nb > 0x7fffffffLL) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_stream_bounce<T>, dim3(unsigned(nb)), dim3(kBlock), 0, as_stream(stream),
                       src, dst, int(nx), int(ny), make_layout(LBM_LAYOUT_SOA, nx, ny),
                       int(nb / nx), u0);
    return launch_status();
}

template <typename U>
void launch_stream_periodic(const void *src, void *dst, int64_t nx, int64_t ny, hipStream_t s) {
    const int64_t nb = row_blocks(nx, ny);
    hipLaunchKernelGGL(k_stream_periodic<U>, dim3(unsigned(nb)), dim3(kBlock), 0, s,
                       static_cast<const U *>(src), static_cast<U *>(dst), int(nx), int(ny),
                       make_layout(LBM_LAYOUT_SOA, nx, ny), int(nb / nx));
}

template <typename U>
void launch_relayout(const void *src, void *dst, int64_t nx, int64_t ny, Layout Ls, Layout Ld,
                     hipStream_t s) {
    const int64_t nseg = (ny + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_relayout<U>, dim3(unsigned(9 * nx * nseg)), dim3(kBlock), 0, s,
                       static_cast<const U *>(src), static_cast<U *>(dst), int(nx), int(ny), Ls,
                       Ld, int(nseg));
}

template <typename T>
int collide_impl(T *f, int64_t n, T omega, T *rho, T *ux, T *uy, void *stream) {
    if (!f) return LBM_E_NULL;
    if (n <= 0 || n > (int64_t(1) << 40)) return LBM_E_SIZE;
    const bool mom = rho != nullptr;
    if (mom != (ux != nullptr) || mom != (uy != nullptr)) return LBM_E_NULL;
    const int64_t nb = (n + kBlock - 1) / kBlock;
    if (nb > 0x7fffffffLL) return LBM_E_SIZE;
    if (mom)
        hipLaunchKernelGGL((k_collide<T, true>), dim3(unsigned(nb)), dim3(kBlock), 0,
                           as_stream(stream), f, n, omega, rho, ux, uy);
    else
        hipLaunchKernelGGL((k_collide<T, false>), dim3(unsigned(nb)), dim3(kBlock), 0,
                           as_stream(stream), f, n, omega, rho, ux, uy);
    return launch_status();
}

template <typename T>
int moments_impl(const T *f, int64_t nx, int64_t ny, int layout, T *rho, T *ux, T *uy,
                 void *stream) {
    if (!f || !rho || !ux || !uy) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_layout(layout)) return LBM_E_ARG;
    const int64_t nb = row_blocks(nx, ny);
    if (nb > 0x7fffffffLL) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_moments_grid<T>, dim3(unsigned(nb)), dim3(kBlock), 0, as_stream(stream),
                       f, int(nx), int(ny), make_layout(layout, nx, ny), int(nb / nx), rho, ux,
                       uy);
    return launch_status();
}

template <typename T>
int equilibrium_impl(const T *rho, const T *ux, const T *uy, T *f, int64_t n, void *stream) {
    if (!rho || !ux || !uy || !f) return LBM_E_NULL;
    if (n <= 0 || n > (int64_t(1) << 40)) return LBM_E_SIZE;
    const int64_t nb = (n + kBlock - 1) / kBlock;
    if (nb > 0x7fffffffLL) return LBM_E_SIZE;
    const int64_t fb = 9 * n * int64_t(sizeof(T)), sb = n * int64_t(sizeof(T));
    if (overlaps(f, fb, rho, sb) || overlaps(f, fb, ux, sb) || overlaps(f, fb, uy, sb))
        return LBM_E_ALIAS;
    hipLaunchKernelGGL(k_equilibrium<T>, dim3(unsigned(nb)), dim3(kBlock), 0, as_stream(stream),
                       rho, ux, uy, f, n);
    return launch_status();
}

// Scratch of the allocation-free residual (lbm_velocity_residual_f64_noalloc):
// the kReduceBlocks per-block partial maxima, one copy per GPU for the life
// of the process.  Host code never passes `scratch` itself to a kernel (on
// the host it is only a stand-in for the symbol); it passes the device
// address that hipGetSymbolAddress returns for the current device.
constexpr int kScratchDoubles = kReduceBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
static std::mutex scratch_mutex;
static double *scratch_ptr[kMaxScratchDevices];  // per device, set once

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; nullptr on a HIP error.
double *scratch_for_current_device(int *err) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) {
        *err = e != hipSuccess ? int(e) : LBM_E_ARG;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = e != hipSuccess ? int(e) : LBM_E_NULL;
            return nullptr;
        }
        scratch_ptr[dev] = static_cast<double *>(p);
    }
    return scratch_ptr[dev];
}

void launch_sum_pass2(const double *partial, int np, double bias, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_pass2, dim3(1), dim3(kBlock), 0, s, partial, np, bias, out);
}
void launch_max_pass2(const double *partial, int np, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_max_pass2, dim3(1), dim3(kBlock), 0, s, partial, np, out);
}

template <typename T>
int sum_impl(const T *x, int64_t n, double *work, double *out, void *stream) {
    if (!x || !work || !out) return LBM_E_NULL;
    if (n <= 0) return LBM_E_SIZE;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_sum_pass1<T>, dim3(kReduceBlocks), dim3(kBlock), 0, s, x, n, work);
    launch_sum_pass2(work, kReduceBlocks, 0.0, out, s);
    return launch_status();
}

}  // namespace lbm

using namespace lbm;

extern "C" {

const char *lbm_version(void) { return "latticeboltzmann_amd 0.1.0 gfx950"; }
int lbm_step_variant_count(void) { return kNumStepShapes; }

int lbm_step_multi_variant_count(void) { return kNumMultiShapes; }

int lbm_step_multi_default_variant(void) { return kDefaultMultiShape; }

int lbm_step_multi_default_for(int64_t nx, int64_t ny) {
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    return default_multi_shape<double>(nx, ny);
}

int lbm_step_multi_default_for_f32(int64_t nx, int64_t ny) {
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    return default_multi_shape<float>(nx, ny);
}

int lbm_step_multi_info(int variant, int *steps, int64_t *min_nx, int64_t *min_ny) {
    if (!steps || !min_nx || !min_ny) return LBM_E_NULL;
    return multi_shape_info(variant, steps, min_nx, min_ny);
}

int lbm_equilibrium_f64(const double *rho, const double *ux, const double *uy, double *f,
                        int64_t n, void *stream) {
    return equilibrium_impl(rho, ux, uy, f, n, stream);
}
int lbm_equilibrium_f32(const float *rho, const float *ux, const float *uy, float *f, int64_t n,
                        void *stream) {
    return equilibrium_impl(rho, ux, uy, f, n, stream);
}

int lbm_collide_f64(double *f, int64_t n, double omega, double *rho, double *ux, double *uy,
                    void *stream) {
    return collide_impl(f, n, omega, rho, ux, uy, stream);
}
int lbm_collide_f32(float *f, int64_t n, float omega, float *rho, float *ux, float *uy,
                    void *stream) {
    return collide_impl(f, n, omega, rho, ux, uy, stream);
}

int lbm_moments_f64(const double *f, int64_t nx, int64_t ny, int layout, double *rho, double *ux,
                    double *uy, void *stream) {
    return moments_impl(f, nx, ny, layout, rho, ux, uy, stream);
}
int lbm_moments_f32(const float *f, int64_t nx, int64_t ny, int layout, float *rho, float *ux,
                    float *uy, void *stream) {
    return moments_impl(f, nx, ny, layout, rho, ux, uy, stream);
}

int lbm_stream_periodic(const void *src, void *dst, int64_t nx, int64_t ny, int elem_bytes,
                        void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return LBM_E_ARG;
    const int64_t bytes = 9 * nx * ny * elem_bytes;
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    if (row_blocks(nx, ny) > 0x7fffffffLL) return LBM_E_SIZE;
    hipStream_t s = as_stream(stream);
    switch (elem_bytes) {
        case 1: launch_stream_periodic<uint8_t>(src, dst, nx, ny, s); break;
        case 2: launch_stream_periodic<uint16_t>(src, dst, nx, ny, s); break;
        case 4: launch_stream_periodic<uint32_t>(src, dst, nx, ny, s); break;
        default: launch_stream_periodic<uint64_t>(src, dst, nx, ny, s); break;
    }
    return launch_status();
}

int lbm_copy(const void *src, void *dst, int64_t nbytes, int nontemporal, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (nbytes <= 0 || nbytes % 16) return LBM_E_SIZE;
    if (reinterpret_cast<uintptr_t>(src) % 16 || reinterpret_cast<uintptr_t>(dst) % 16) return LBM_E_ARG;
    if (overlaps(src, nbytes, dst, nbytes)) return LBM_E_ALIAS;
    const int64_t n16 = nbytes / 16;
    const int64_t blocks = (n16 + kBlock - 1) / kBlock;
    if (blocks > INT32_MAX) return LBM_E_SIZE;
    const dim3 grid{unsigned(blocks)};
    hipStream_t s = as_stream(stream);
    if (nontemporal)
        hipLaunchKernelGGL(k_copy16<true>, grid, dim3(kBlock), 0, s, static_cast<const u32x4 *>(src),
                           static_cast<u32x4 *>(dst), n16);
    else
        hipLaunchKernelGGL(k_copy16<false>, grid, dim3(kBlock), 0, s, static_cast<const u32x4 *>(src),
                           static_cast<u32x4 *>(dst), n16);
    return launch_status();
}

int lbm_step_multi_f64(const double *src, double *dst, int64_t nx, int64_t ny, int layout,
                       double omega, int bc, double u0, double *rho, double *ux, double *uy,
                       int variant, void *stream) {
    return multi_impl(src, dst, nx, ny, layout, omega, bc, u0, rho, ux, uy, variant, stream);
}

int lbm_step_multi_f32(const float *src, float *dst, int64_t nx, int64_t ny, int layout,
                       float omega, int bc, float u0, float *rho, float *ux, float *uy,
                       int variant, void *stream) {
    return multi_impl(src, dst, nx, ny, layout, omega, bc, u0, rho, ux, uy, variant, stream);
}

int lbm_relayout(const void *src, void *dst, int64_t nx, int64_t ny, int elem_bytes,
                 int src_layout, int dst_layout, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_layout(src_layout) || !valid_layout(dst_layout)) return LBM_E_ARG;
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return LBM_E_ARG;
    const int64_t bytes = 9 * nx * ny * elem_bytes;
    if (overlaps(src, bytes, dst, bytes)) return LBM_E_ALIAS;
    if (9 * row_blocks(nx, ny) > 0x7fffffffLL) return LBM_E_SIZE;
    const Layout Ls = make_layout(src_layout, nx, ny), Ld = make_layout(dst_layout, nx, ny);
    hipStream_t s = as_stream(stream);
    switch (elem_bytes) {
        case 1: launch_relayout<uint8_t>(src, dst, nx, ny, Ls, Ld, s); break;
        case 2: launch_relayout<uint16_t>(src, dst, nx, ny, Ls, Ld, s); break;
        case 4: launch_relayout<uint32_t>(src, dst, nx, ny, Ls, Ld, s); break;
        default: launch_relayout<uint64_t>(src, dst, nx, ny, Ls, Ld, s); break;
    }
    return launch_status();
}

int lbm_stream_bounce_back_f64(const double *src, double *dst, int64_t nx, int64_t ny, double u0,
                               void *stream) {
    return stream_bounce_impl(src, dst, nx, ny, u0, stream);
}
int lbm_stream_bounce_back_f32(const float *src, float *dst, int64_t nx, int64_t ny, float u0,
                               void *stream) {
    return stream_bounce_impl(src, dst, nx, ny, u0, stream);
}

int lbm_step_f64(const double *src, double *dst, int64_t nx, int64_t ny, double omega, int bc,
                 double u0, double *rho, double *ux, double *uy, int variant, void *stream) {
    return step_impl(src, dst, nx, ny, LBM_LAYOUT_SOA, omega, bc, u0, rho, ux, uy, variant, stream);
}
int lbm_step_f32(const float *src, float *dst, int64_t nx, int64_t ny, float omega, int bc,
                 float u0, float *rho, float *ux, float *uy, int variant, void *stream) {
    return step_impl(src, dst, nx, ny, LBM_LAYOUT_SOA, omega, bc, u0, rho, ux, uy, variant, stream);
}
int lbm_step_layout_f64(const double *src, double *dst, int64_t nx, int64_t ny, int layout,
                        double omega, int bc, double u0, double *rho, double *ux, double *uy,
                        int variant, void *stream) {
    return step_impl(src, dst, nx, ny, layout, omega, bc, u0, rho, ux, uy, variant, stream);
}
int lbm_step_layout_f32(const float *src, float *dst, int64_t nx, int64_t ny, int layout,
                        float omega, int bc, float u0, float *rho, float *ux, float *uy,
                        int variant, void *stream) {
    return step_impl(src, dst, nx, ny, layout, omega, bc, u0, rho, ux, uy, variant, stream);
}

int64_t lbm_reduce_workspace_bytes(void) { return int64_t(kReduceBlocks) * int64_t(sizeof(double)); }

int lbm_sum_f64(const double *x, int64_t n, double *work, double *out, void *stream) {
    return sum_impl(x, n, work, out, stream);
}
int lbm_sum_f32(const float *x, int64_t n, double *work, double *out, void *stream) {
    return sum_impl(x, n, work, out, stream);
}

int lbm_velocity_residual_f64(const double *f, int64_t nx, int64_t ny, int layout,
                              double *ux_prev, double *uy_prev, double *work, double *out,
                              void *stream) {
    if (!f || !ux_prev || !uy_prev || !work || !out) return LBM_E_NULL;
    if (!valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!valid_layout(layout)) return LBM_E_ARG;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_residual_pass1, dim3(kReduceBlocks), dim3(kBlock), 0, s, f, int(nx),
                       int(ny), make_layout(layout, nx, ny), ux_prev, uy_prev, work);
    hipLaunchKernelGGL(k_max_pass2, dim3(1), dim3(kBlock), 0, s, work, kReduceBlocks, out);
    return launch_status();
}

int64_t lbm_noalloc_scratch_bytes(void) { return int64_t(kScratchDoubles) * int64_t(sizeof(double)); }

int lbm_velocity_residual_f64_noalloc(const double *f, int64_t nx, int64_t ny, int layout,
                                      double *ux_prev, double *uy_prev, double *out,
                                      void *stream) {
    if (!f || !ux_prev || !uy_prev || !out) return LBM_E_NULL;
    if (!valid_grid(nx, ny) || nx > INT32_MAX || ny > INT32_MAX) return LBM_E_SIZE;
    if (!valid_layout(layout)) return LBM_E_ARG;
    // pass 1 writes every one of its kReduceBlocks partials before pass 2
    // reads them, so the scratch needs no clearing between calls
    static_assert(kReduceBlocks <= kScratchDoubles, "partials must fit the scratch");
    int err = LBM_OK;
    double *work = scratch_for_current_device(&err);
    if (!work) return err;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_residual_pass1, dim3(kReduceBlocks), dim3(kBlock), 0, s, f, int(nx),
                       int(ny), make_layout(layout, nx, ny), ux_prev, uy_prev, work);
    hipLaunchKernelGGL(k_max_pass2, dim3(1), dim3(kBlock), 0, s, work, kReduceBlocks, out);
    return launch_status();
}

}  // extern "C"
