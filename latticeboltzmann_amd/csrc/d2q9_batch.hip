// Batched stepping of many independent lattices of different sizes.
//
// All instances live in one population pool, each as scaled deviations in the
// row-interleaved layout f[k][i][l] at its own offset.  Work is described by
// node tiles: (instance, first node) with a fixed number of consecutive nodes
// k*ny + l per tile.  Instance sides are multiples of 64 in the sweeps this
// serves, so every wave of a tile reads and writes 64 consecutive y values of
// one channel row; tiles of small instances span several rows.  A launch runs
// a persistent grid (a few blocks per compute unit) that loops over its tile
// list, so one launch covers hundreds of uneven instances.  The per-node
// arithmetic is the single-lattice kernel's (d2q9_dev.hpp), so an instance
// stepped here has the same bits as the same lattice stepped alone.
//
// Also here: the per-instance shear-wave amplitude (a deterministic two-pass
// FP64 reduction), device counters used to index samples inside captured
// graphs, and the HIP-graph capture helpers.

#include "d2q9_dev.hpp"
#include "d2q9_grid.hpp"
#include "d2q9_reduce.hpp"
#include "lbm_internal.hpp"

struct lbm_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace lbm {

template <int FMT, typename CT, int BS, int V, bool NTS>
__global__ __launch_bounds__(BS) void k_batch_step(const typename Codec<FMT>::raw *__restrict__ src,
                                                   typename Codec<FMT>::raw *__restrict__ dst,
                                                   const lbm_batch_inst *__restrict__ insts,
                                                   const lbm_batch_tile *__restrict__ tiles,
                                                   int ntiles, CT scale, CT inv_scale,
                                                   int stochastic, uint32_t seed,
                                                   const uint32_t *__restrict__ epoch) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    const uint32_t seed_eff = epoch ? hash32(seed ^ (epoch[0] * 0x9e3779b9U)) : seed;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const lbm_batch_tile tile = tiles[t];
        const lbm_batch_inst I = insts[tile.inst];
        const int nx = I.nx, ny = I.ny;
        const int n = nx * ny;
        const Layout L{ny, 9 * int64_t(ny)};
        // row of node j: (j + 1/2) * fl(1/ny) has relative error below 2.4e-7, so for
        // n <= 2^21 it stays within 1/(2 ny) of the true quotient, whose fractional
        // part lies in [1/(2 ny), 1 - 1/(2 ny)]: the truncation is exact
        const bool fast_div = n <= (1 << 21);
        const float inv_ny = 1.0f / float(ny);
        const R *s = src + I.offset;
        R *d = dst + I.offset;

        R raw[V][9];
        int kk[V], ll[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int j = tile.first + v * BS + int(threadIdx.x);
            kk[v] = -1;
            ll[v] = 0;
            if (j < n) {
                const int k = fast_div ? int((float(j) + 0.5f) * inv_ny) : j / ny;
                kk[v] = k;
                ll[v] = j - k * ny;
                const PullRows<R> rows(s, L, k, nx);
                pull9<false>(rows, ll[v], ny, raw[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int k = kk[v], l = ll[v];
            if (k < 0) continue;
            CT g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = CT(C::dec(raw[v][i])) * inv_scale;
            if (I.bc == LBM_BC_CAVITY)
                cavity_walls_dev<FMT, CT>(g, s + k * L.rs + l, L.cs, k, l, nx, ny, CT(I.u0),
                                          inv_scale);
            CT ux, uy;
            collide_dev<CT>(g, CT(I.omega), ux, uy);
            R out[9];
            Dither dither(seed_eff, uint64_t(I.offset) / 9 + uint64_t(k) * ny + l);
            store_values<FMT, CT>(g, scale, stochastic, dither, out);
            R *o = d + k * L.rs + l;
#pragma unroll
            for (int i = 0; i < 9; ++i) st<NTS>(o + i * L.cs, out[i]);
        }
    }
}

// K steps of the instances of a work list per launch, by the register sweep of
// d2q9_dev.hpp: each wave takes one item (instance, strip of columns, range
// of rows), so one launch covers instances of every size.  Every step is
// rounded with the batch's dither (the step's seed combined on the device
// with *epoch; node = instance offset / 9 + k ny + l), so a launch is bitwise
// K lbm_batch_step launches.  Instances need ny >= 64 and nx >= 2K.
template <int FMT, typename CT, int K>
__global__ __launch_bounds__(256) void k_batch_sweep(const typename Codec<FMT>::raw *__restrict__ src,
                                                     typename Codec<FMT>::raw *__restrict__ dst,
                                                     const lbm_batch_inst *__restrict__ insts,
                                                     const lbm_batch_item *__restrict__ items,
                                                     int nitems, int nblocks, CT scale, CT inv_scale,
                                                     int stochastic, uint32_t seed,
                                                     uint32_t seed_stride,
                                                     const uint32_t *__restrict__ epoch) {
    const int b = xcd_tile(nblocks);
    const int w = __builtin_amdgcn_readfirstlane(b * 4 + int(threadIdx.x >> 6));
    if (b >= nblocks || w >= nitems) return;  // whole waves; nothing waits for them
    const lbm_batch_item it = items[w];
    const lbm_batch_inst I = insts[it.inst];
    const int nx = I.nx, ny = I.ny;
    const Layout L{ny, 9 * int64_t(ny)};
    const DevSweepLane ln = dev_sweep_lane<K>(it.strip, ny, int(threadIdx.x & 63));
    const uint32_t mix = epoch ? epoch[0] * 0x9e3779b9U : 0u;
    uint32_t seeds[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t sj = seed + uint32_t(j) * seed_stride;
        seeds[j] = epoch ? hash32(sj ^ mix) : sj;
    }
    const auto *s = src + I.offset;
    auto *d = dst + I.offset;
    const uint64_t base = uint64_t(I.offset) / 9;
    const int k1 = it.row0 + it.rows;
    if (I.bc == LBM_BC_CAVITY)
        dev_sweep_run<FMT, CT, LBM_BC_CAVITY, false, K, true>(
            s, d, it.row0, k1, nx, ny, L, ln, CT(I.omega), CT(I.u0), scale, inv_scale, stochastic,
            seeds, base, nullptr, nullptr);
    else
        dev_sweep_run<FMT, CT, LBM_BC_PERIODIC, false, K, true>(
            s, d, it.row0, k1, nx, ny, L, ln, CT(I.omega), CT(I.u0), scale, inv_scale, stochastic,
            seeds, base, nullptr, nullptr);
}

// Pass 1 of the amplitude: one kBlock-node tile per loop iteration, partial
// sum of u_dir * sin(2 pi pos / n) in FP64.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_batch_amp_partial(
    const typename Codec<FMT>::raw *__restrict__ pool, const lbm_batch_inst *__restrict__ insts,
    const lbm_batch_tile *__restrict__ tiles, int ntiles, double inv_scale,
    double *__restrict__ partial) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const lbm_batch_tile tile = tiles[t];
        const lbm_batch_inst I = insts[tile.inst];
        const int j = tile.first + int(threadIdx.x);
        double val = 0.0;
        if (j < I.nx * I.ny) {
            const int k = j / I.ny, l = j - k * I.ny;
            const typename Codec<FMT>::raw *c = pool + I.offset + int64_t(k) * 9 * I.ny + l;
            double g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = double(Codec<FMT>::dec(c[i * int64_t(I.ny)])) * inv_scale;
            const double rho = 1.0 + (((g[0] + g[1]) + (g[2] + g[3])) + ((g[4] + g[5]) + (g[6] + (g[7] + g[8]))));
            if (I.dir == 0) {
                const double jx = ((g[1] - g[3]) + (g[5] - g[6])) + (g[8] - g[7]);
                val = jx / rho * sinpi(2.0 * double(l) / double(I.ny));
            } else {
                const double jy = ((g[2] - g[4]) + (g[5] + g[6])) - (g[7] + g[8]);
                val = jy / rho * sinpi(2.0 * double(k) / double(I.nx));
            }
        }
        const double tot = block_sum(val);
        if (threadIdx.x == 0) partial[t] = tot;
        __syncthreads();  // the LDS slot of block_sum is reused next iteration
    }
}

// Pass 2: one thread per instance folds its tiles' partials in order.
__global__ __launch_bounds__(kBlock) void k_batch_amp_final(const double *__restrict__ partial,
                                                            const int32_t *__restrict__ inst_tiles,
                                                            int ninst,
                                                            const lbm_batch_inst *__restrict__ insts,
                                                            double *__restrict__ amp, int nslots,
                                                            const uint32_t *__restrict__ epoch) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ninst) return;
    const uint32_t slot = epoch ? epoch[0] : 0u;
    if (slot >= uint32_t(nslots)) return;
    double s = 0.0, c = 0.0;
    for (int t = inst_tiles[i]; t < inst_tiles[i + 1]; ++t) neumaier_add(s, c, partial[t]);
    const lbm_batch_inst I = insts[i];
    amp[int64_t(slot) * ninst + i] = (s + c) * 2.0 / (double(I.nx) * double(I.ny));
}

__global__ void k_counter_add(uint32_t *counter, uint32_t add, uint32_t set, int mode) {
    if (blockIdx.x == 0 && threadIdx.x == 0) counter[0] = mode ? set : counter[0] + add;
}

// ---------------------------------------------------------------------------
// Launch shapes of the batched step: (threads per block, nodes per thread,
// non-temporal stores).  A tile holds BS * V nodes.
#define LBM_BATCH_SHAPES(X)  \
    X(0, 256, 1, true)       \
    X(1, 256, 2, true)       \
    X(2, 256, 4, true)       \
    X(3, 128, 1, true)       \
    X(4, 128, 2, true)       \
    X(5, 512, 1, true)       \
    X(6, 64, 1, true)        \
    X(7, 256, 2, false)

#define LBM_BATCH_COUNT(id, bs, v, nts) +1
constexpr int kNumBatchShapes = 0 LBM_BATCH_SHAPES(LBM_BATCH_COUNT);

inline int batch_tile_nodes(int shape) {
#define LBM_BATCH_NODES(id, bs, v, nts) \
    if (shape == id) return bs * v;
    LBM_BATCH_SHAPES(LBM_BATCH_NODES)
#undef LBM_BATCH_NODES
    return 0;
}

struct BatchArgs {
    const void *src;
    void *dst;
    float scale;
    const lbm_batch_inst *insts;
    const lbm_batch_tile *tiles;
    int ntiles, stochastic;
    uint32_t seed;
    const uint32_t *epoch;
    int blocks;
};

template <int FMT, typename CT, int BS, int V, bool NTS>
void launch_batch_v(const BatchArgs &a, hipStream_t s) {
    using R = typename Codec<FMT>::raw;
    const int blocks = a.blocks > 0 ? a.blocks : device_cus() * (2048 / BS);
    const int grid = a.ntiles < blocks ? a.ntiles : blocks;
    hipLaunchKernelGGL((k_batch_step<FMT, CT, BS, V, NTS>), dim3(unsigned(grid)), dim3(BS), 0, s,
                       static_cast<const R *>(a.src), static_cast<R *>(a.dst), a.insts, a.tiles,
                       a.ntiles, CT(a.scale), CT(1) / CT(a.scale), a.stochastic, a.seed, a.epoch);
}

template <int FMT, typename CT>
void launch_batch_fmt(const BatchArgs &a, int shape, hipStream_t s) {
#define LBM_BATCH_LAUNCH(id, bs, v, nts)            \
    case id:                                       \
        launch_batch_v<FMT, CT, bs, v, nts>(a, s); \
        break;
    switch (shape) {
        LBM_BATCH_SHAPES(LBM_BATCH_LAUNCH)
        default: break;
    }
#undef LBM_BATCH_LAUNCH
}

inline bool valid_fmt_b(int fmt) {
    return fmt == LBM_FMT_F32 || fmt == LBM_FMT_BF16 || fmt == LBM_FMT_F16;
}

template <int FMT, typename CT, int K>
void launch_batch_sweep_v(const void *src, void *dst, float scale, const lbm_batch_inst *insts,
                          const lbm_batch_item *items, int nitems, int stochastic, uint32_t seed,
                          uint32_t stride, const uint32_t *epoch, hipStream_t s) {
    using R = typename Codec<FMT>::raw;
    const int nblocks = (nitems + 3) / 4;
    const int grid = ((nblocks + 7) / 8) * 8;
    hipLaunchKernelGGL((k_batch_sweep<FMT, CT, K>), dim3(unsigned(grid)), dim3(256), 0, s,
                       static_cast<const R *>(src), static_cast<R *>(dst), insts, items, nitems,
                       nblocks, CT(scale), CT(1) / CT(scale), stochastic, seed, stride, epoch);
}

template <typename CT, int K>
void launch_batch_sweep_k(int fmt, const void *src, void *dst, float scale,
                          const lbm_batch_inst *insts, const lbm_batch_item *items, int nitems,
                          int stochastic, uint32_t seed, uint32_t stride, const uint32_t *epoch,
                          hipStream_t s) {
    switch (fmt) {
        case LBM_FMT_F32:
            launch_batch_sweep_v<LBM_FMT_F32, CT, K>(src, dst, scale, insts, items, nitems,
                                                     stochastic, seed, stride, epoch, s);
            break;
        case LBM_FMT_BF16:
            launch_batch_sweep_v<LBM_FMT_BF16, CT, K>(src, dst, scale, insts, items, nitems,
                                                      stochastic, seed, stride, epoch, s);
            break;
        default:
            launch_batch_sweep_v<LBM_FMT_F16, CT, K>(src, dst, scale, insts, items, nitems,
                                                     stochastic, seed, stride, epoch, s);
            break;
    }
}

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_batch_variant_count(void) { return kNumBatchShapes; }
int lbm_batch_tile_nodes(int variant) {
    return (variant >= 0 && variant < kNumBatchShapes) ? batch_tile_nodes(variant) : 0;
}

int lbm_batch_step(const void *src, void *dst, int fmt, float scale, const lbm_batch_inst *insts,
                   const lbm_batch_tile *tiles, int ntiles, int compute, int rounding,
                   uint32_t seed, const uint32_t *epoch, int variant, int blocks, void *stream) {
    if (!src || !dst || !insts || !tiles) return LBM_E_NULL;
    if (src == dst) return LBM_E_ALIAS;
    if (ntiles <= 0 || blocks < 0) return LBM_E_SIZE;
    if (!valid_fmt_b(fmt) || !(scale > 0.0f)) return LBM_E_ARG;
    if (rounding != LBM_ROUND_NEAREST && rounding != LBM_ROUND_STOCHASTIC) return LBM_E_ARG;
    if (compute != LBM_COMPUTE_F32 && compute != LBM_COMPUTE_F64) return LBM_E_ARG;
    if (variant < 0 || variant >= kNumBatchShapes) return LBM_E_ARG;
    const BatchArgs a{src, dst, scale, insts, tiles, ntiles, rounding == LBM_ROUND_STOCHASTIC ? 1 : 0,
                      seed, epoch, blocks};
    hipStream_t s = as_stream(stream);
    if (compute == LBM_COMPUTE_F64) {
        switch (fmt) {
            case LBM_FMT_F32: launch_batch_fmt<LBM_FMT_F32, double>(a, variant, s); break;
            case LBM_FMT_BF16: launch_batch_fmt<LBM_FMT_BF16, double>(a, variant, s); break;
            default: launch_batch_fmt<LBM_FMT_F16, double>(a, variant, s); break;
        }
    } else {
        switch (fmt) {
            case LBM_FMT_F32: launch_batch_fmt<LBM_FMT_F32, float>(a, variant, s); break;
            case LBM_FMT_BF16: launch_batch_fmt<LBM_FMT_BF16, float>(a, variant, s); break;
            default: launch_batch_fmt<LBM_FMT_F16, float>(a, variant, s); break;
        }
    }
    return launch_status();
}

int lbm_batch_sweep_strips(int steps, int ny) {
    if (steps == 2) return dev_sweep_strips<2>(ny);
    if (steps == 4) return dev_sweep_strips<4>(ny);
    return 0;
}

int lbm_batch_step_multi(const void *src, void *dst, int fmt, float scale,
                         const lbm_batch_inst *insts, const lbm_batch_item *items, int nitems,
                         int steps, int compute, int rounding, uint32_t seed, uint32_t seed_stride,
                         const uint32_t *epoch, void *stream) {
    if (!src || !dst || !insts || !items) return LBM_E_NULL;
    if (src == dst) return LBM_E_ALIAS;
    if (nitems <= 0 || nitems > (INT32_MAX - 31) / 4 * 4) return LBM_E_SIZE;
    if (!valid_fmt_b(fmt) || !(scale > 0.0f)) return LBM_E_ARG;
    if (steps != 2 && steps != 4) return LBM_E_ARG;
    if (rounding != LBM_ROUND_NEAREST && rounding != LBM_ROUND_STOCHASTIC) return LBM_E_ARG;
    if (compute != LBM_COMPUTE_F32 && compute != LBM_COMPUTE_F64) return LBM_E_ARG;
    const int sr = rounding == LBM_ROUND_STOCHASTIC ? 1 : 0;
    hipStream_t s = as_stream(stream);
    if (compute == LBM_COMPUTE_F64) {
        if (steps == 2)
            launch_batch_sweep_k<double, 2>(fmt, src, dst, scale, insts, items, nitems, sr, seed,
                                            seed_stride, epoch, s);
        else
            launch_batch_sweep_k<double, 4>(fmt, src, dst, scale, insts, items, nitems, sr, seed,
                                            seed_stride, epoch, s);
    } else {
        if (steps == 2)
            launch_batch_sweep_k<float, 2>(fmt, src, dst, scale, insts, items, nitems, sr, seed,
                                           seed_stride, epoch, s);
        else
            launch_batch_sweep_k<float, 4>(fmt, src, dst, scale, insts, items, nitems, sr, seed,
                                           seed_stride, epoch, s);
    }
    return launch_status();
}

int lbm_batch_amplitude(const void *pool, int fmt, float scale, const lbm_batch_inst *insts,
                        int ninst, const lbm_batch_tile *tiles, int ntiles,
                        const int32_t *inst_tiles, double *partial, double *amp, int nslots,
                        const uint32_t *epoch, void *stream) {
    if (!pool || !insts || !tiles || !inst_tiles || !partial || !amp) return LBM_E_NULL;
    if (ninst <= 0 || ntiles <= 0 || nslots <= 0) return LBM_E_SIZE;
    if (!valid_fmt_b(fmt) || !(scale > 0.0f)) return LBM_E_ARG;
    hipStream_t s = as_stream(stream);
    const int grid = ntiles < device_cus() * 8 ? ntiles : device_cus() * 8;
    const double inv = 1.0 / double(scale);
    switch (fmt) {
        case LBM_FMT_F32:
            hipLaunchKernelGGL(k_batch_amp_partial<LBM_FMT_F32>, dim3(grid), dim3(kBlock), 0, s,
                               static_cast<const float *>(pool), insts, tiles, ntiles, inv, partial);
            break;
        case LBM_FMT_BF16:
            hipLaunchKernelGGL(k_batch_amp_partial<LBM_FMT_BF16>, dim3(grid), dim3(kBlock), 0, s,
                               static_cast<const uint16_t *>(pool), insts, tiles, ntiles, inv, partial);
            break;
        default:
            hipLaunchKernelGGL(k_batch_amp_partial<LBM_FMT_F16>, dim3(grid), dim3(kBlock), 0, s,
                               static_cast<const uint16_t *>(pool), insts, tiles, ntiles, inv, partial);
            break;
    }
    hipLaunchKernelGGL(k_batch_amp_final, dim3((ninst + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       partial, inst_tiles, ninst, insts, amp, nslots, epoch);
    return launch_status();
}

int lbm_counter_increment(uint32_t *counter, void *stream) {
    if (!counter) return LBM_E_NULL;
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, as_stream(stream), counter, 1u, 0u, 0);
    return launch_status();
}

int lbm_counter_set(uint32_t *counter, uint32_t value, void *stream) {
    if (!counter) return LBM_E_NULL;
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, as_stream(stream), counter, 0u, value, 1);
    return launch_status();
}

int lbm_graph_begin(void *stream) {
    if (!stream) return LBM_E_ARG;  // the legacy default stream cannot be captured
    device_cus();                   // query the device before capture starts
    hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal);
    return e == hipSuccess ? LBM_OK : int(e);
}

int lbm_graph_end(void *stream, lbm_graph **graph) {
    if (!stream || !graph) return LBM_E_NULL;
    *graph = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(as_stream(stream), &g);
    if (e != hipSuccess) return int(e);
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);  // the instantiation error is the one to report
        return int(e);
    }
    *graph = new lbm_graph{g, exec};
    return LBM_OK;
}

int lbm_graph_launch(lbm_graph *graph, void *stream) {
    if (!graph) return LBM_E_NULL;
    hipError_t e = hipGraphLaunch(graph->exec, as_stream(stream));
    return e == hipSuccess ? LBM_OK : int(e);
}

int lbm_graph_destroy(lbm_graph *graph) {
    if (!graph) return LBM_E_NULL;
    const hipError_t e1 = hipGraphExecDestroy(graph->exec);
    const hipError_t e2 = hipGraphDestroy(graph->graph);
    delete graph;
    if (e1 != hipSuccess) return int(e1);
    return e2 == hipSuccess ? LBM_OK : int(e2);
}

}  // extern "C"
