/*
 * lbm_d2q9.h -- C ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann kernels.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t passed
 * as `void *` (NULL = the default stream), enqueues its kernels on that stream
 * and returns without synchronising.  The return value is 0 on success, a
 * negative LBM_E* code for an argument error (nothing is launched), or a
 * positive hipError_t if a launch failed.
 *
 * Population arrays use the reference's layout f[i][k][l]: channel i = 0..8
 * (0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE; PyLB/Streaming.py:28-29),
 * x index k = 0..nx-1 (slow), y index l = 0..ny-1 (fast, contiguous); a plane
 * holds nx*ny values and the nine planes follow each other.  Scalar fields
 * (rho, ux, uy) are [k][l].
 *
 * Arithmetic of the *_f64 / *_f32 entry points follows the reference's numpy
 * opt1 operation order (cavity_opt1.py:100-245) without fused multiply-adds,
 * so FP64 results are bit-identical to that path.
 */
#ifndef LBM_D2Q9_H
#define LBM_D2Q9_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define LBM_OK 0
#define LBM_E_NULL (-1)      /* a required pointer is NULL                    */
#define LBM_E_SIZE (-2)      /* a size is zero, negative or too large         */
#define LBM_E_ARG (-3)       /* an enum or scalar argument is out of range    */
#define LBM_E_ALIAS (-4)     /* input and output buffers overlap              */
#define LBM_E_WORKSPACE (-5) /* workspace too small                           */

/* population layouts: element (i, k, l) lives at i*cs + k*rs + l
 *   LBM_LAYOUT_SOA : f[i][k][l], cs = nx*ny, rs = ny   (the reference's f_ikl)
 *   LBM_LAYOUT_ROWS: f[k][i][l], cs = ny, rs = 9*ny    (row-interleaved: the
 *                    nine channel segments of an x-row are adjacent in HBM) */
#define LBM_LAYOUT_SOA 0
#define LBM_LAYOUT_ROWS 1

/* boundary kinds for the fused step */
#define LBM_BC_PERIODIC 0 /* periodic in x and y (shear wave)                 */
#define LBM_BC_CAVITY 1   /* bounce-back walls, moving lid at l = ny-1        */

/* Library version string and the number of kernel variants of the fused step. */
const char *lbm_version(void);
int lbm_step_variant_count(void);

/* ---- equilibrium ------------------------------------------------------------
 * f[i][j] = f_eq_i(rho[j], ux[j], uy[j]) for j < n (f is 9 x n).
 * Replaces _lbkernels.equilibrium, array overload -> d2q9_equilibriumn<T>
 * (c/d2q9.h:98-108, bound at c/_lbkernels.cpp:38-39, 45-46), with the
 * arithmetic of cavity_opt1.py:100-132.  The scalar overload
 * d2q9_equilibrium1<T> (c/d2q9.h:59-81) is this call with n = 1.            */
int lbm_equilibrium_f64(const double *rho, const double *ux, const double *uy,
                        double *f, int64_t n, void *stream);
int lbm_equilibrium_f32(const float *rho, const float *ux, const float *uy,
                        float *f, int64_t n, void *stream);

/* ---- collide ------------------------------------------------------------------
 * BGK relaxation in place on f (9 x n): f += omega (f_eq(rho, u) - f).
 * rho/ux/uy (each n values, all NULL or all non-NULL) receive the
 * pre-collision moments, as the numpy collide returns them (cavity_opt1.py:160).
 * Replaces _lbkernels.collide -> d2q9_colliden<T> (c/d2q9.h:121-131, bound at
 * c/_lbkernels.cpp:40-41, 47-48); arithmetic of cavity_opt1.py:134-160.      */
int lbm_collide_f64(double *f, int64_t n, double omega, double *rho, double *ux,
                    double *uy, void *stream);
int lbm_collide_f32(float *f, int64_t n, float omega, float *rho, float *ux,
                    float *uy, void *stream);

/* Moments of an nx x ny grid in either layout, f unchanged; rho/ux/uy are
 * [k][l].  Order of the numpy collide (cavity_opt1.py:156-158); the reference
 * computes them inline in its drivers before each dump (cavity_opt2.py:280-281). */
int lbm_moments_f64(const double *f, int64_t nx, int64_t ny, int layout, double *rho,
                    double *ux, double *uy, void *stream);
int lbm_moments_f32(const float *f, int64_t nx, int64_t ny, int layout, float *rho,
                    float *ux, float *uy, void *stream);

/* ---- streaming ------------------------------------------------------------------
 * Periodic pull streaming src -> dst (9 x nx x ny each, no overlap):
 * dst_i[k][l] = src_i[(k - cx_i) mod nx][(l - cy_i) mod ny].  A pure
 * permutation of elements of elem_bytes in {1, 2, 4, 8}.
 * Replaces PyLB.stream (PyLB/Streaming.py:33-46).                              */
int lbm_stream_periodic(const void *src, void *dst, int64_t nx, int64_t ny,
                        int elem_bytes, void *stream);

/* K full time steps src -> dst in one launch (temporal blocking, K = 2..6
 * fixed by the variant).  Tile variants: each block computes the first K-1
 * steps of its tile and a shrinking halo in LDS and writes only the last.
 * Sweep variants: each wave walks a strip of 64 columns along x and keeps the
 * intermediate steps of the last rows in registers.  Either way HBM sees
 * about one read and one write of each population per K steps.  Bitwise equal to K
 * lbm_step_layout_* calls (same arithmetic, cavity_opt1.py:341-343 K times);
 * rho/ux/uy (nullable, all or none) receive the pre-collision moments of the
 * last step.  variant picks (K, tile shape), 0 .. lbm_step_multi_variant_count()-1
 * or -1 for the default of the grid size (lbm_step_multi_default_for);
 * lbm_step_multi_info gives its K and the smallest grid it takes (LBM_E_SIZE
 * for smaller grids).                                                        */
int lbm_step_multi_f64(const double *src, double *dst, int64_t nx, int64_t ny, int layout,
                       double omega, int bc, double u0, double *rho, double *ux, double *uy,
                       int variant, void *stream);
int lbm_step_multi_f32(const float *src, float *dst, int64_t nx, int64_t ny, int layout,
                       float omega, int bc, float u0, float *rho, float *ux, float *uy,
                       int variant, void *stream);
int lbm_step_multi_variant_count(void);
int lbm_step_multi_default_variant(void); /* the FP64 tile default (below 4096^2 nodes) */
/* The variant that -1 selects on an nx x ny grid: the default tile below
 * 4096^2 nodes, the five-step register sweep from there (LBM_E_SIZE for an
 * invalid grid); _f32 for lbm_step_multi_f32, whose tile default differs.
 * No counterpart in the reference (its driver steps one step per call).       */
int lbm_step_multi_default_for(int64_t nx, int64_t ny);
int lbm_step_multi_default_for_f32(int64_t nx, int64_t ny);
int lbm_step_multi_info(int variant, int *steps, int64_t *min_nx, int64_t *min_ny);

/* Copy nbytes (a multiple of 16, both pointers 16-byte aligned, disjoint) with
 * 16-byte loads and stores, non-temporal stores if nontemporal != 0.  No
 * reference counterpart: the achievable HBM copy rate, against which the
 * bandwidth-bound step kernels are measured.                                    */
int lbm_copy(const void *src, void *dst, int64_t nbytes, int nontemporal, void *stream);

/* Copy a population array between layouts (a permutation; src, dst disjoint).
 * No reference counterpart: the reference has one layout.                      */
int lbm_relayout(const void *src, void *dst, int64_t nx, int64_t ny, int elem_bytes,
                 int src_layout, int dst_layout, void *stream);

/* Streaming plus cavity walls: bounce-back at l = 0, k = 0, k = nx-1 and the
 * lid at l = ny-1 moving with u0 in +x, with the reference's override order.
 * Replaces stream_and_bounce_back (cavity_opt1.py:177-245, identical at
 * cavity_opt2.py:109-177).                                                      */
int lbm_stream_bounce_back_f64(const double *src, double *dst, int64_t nx,
                               int64_t ny, double u0, void *stream);
int lbm_stream_bounce_back_f32(const float *src, float *dst, int64_t nx,
                               int64_t ny, float u0, void *stream);

/* ---- fused step -------------------------------------------------------------------
 * One full time step src -> dst: pull streaming (+ walls and lid if
 * bc == LBM_BC_CAVITY), then BGK collide, reading and writing each population
 * once.  Equivalent to the driver loops cavity_opt1.py:341-343
 * (stream_and_bounce_back; collide) and shear_wave_opt1.py:171-172
 * (stream; collide).  rho/ux/uy (nullable, all or none) receive the
 * pre-collision moments.  variant selects a launch shape
 * (0 .. lbm_step_variant_count()-1, or -1 for the default); all variants give
 * identical results.                                                            */
int lbm_step_f64(const double *src, double *dst, int64_t nx, int64_t ny,
                 double omega, int bc, double u0, double *rho, double *ux,
                 double *uy, int variant, void *stream);
int lbm_step_f32(const float *src, float *dst, int64_t nx, int64_t ny,
                 float omega, int bc, float u0, float *rho, float *ux, float *uy,
                 int variant, void *stream);
/* The same step on either layout (lbm_step_* is layout LBM_LAYOUT_SOA); the
 * arithmetic, and therefore every bit of the result, does not depend on it. */
int lbm_step_layout_f64(const double *src, double *dst, int64_t nx, int64_t ny,
                        int layout, double omega, int bc, double u0, double *rho,
                        double *ux, double *uy, int variant, void *stream);
int lbm_step_layout_f32(const float *src, float *dst, int64_t nx, int64_t ny,
                        int layout, float omega, int bc, float u0, float *rho,
                        float *ux, float *uy, int variant, void *stream);

/* ---- reductions (FP64, deterministic) -------------------------------------------
 * Two-pass reductions with a fixed block count: identical inputs give
 * bit-identical results.  `work` is device scratch of at least
 * lbm_reduce_workspace_bytes() bytes; *out is a device double.
 * No reference counterpart: the reference prints no mass or residual.          */
int64_t lbm_reduce_workspace_bytes(void);
/* sum of n doubles (mass of an FP64 population array: n = 9 nx ny) */
int lbm_sum_f64(const double *x, int64_t n, double *work, double *out, void *stream);
/* sum of n floats accumulated in FP64 */
int lbm_sum_f32(const float *x, int64_t n, double *work, double *out, void *stream);
/* Velocity residual: out = max_j |(ux, uy)(f)_j - (ux_prev, uy_prev)_j| over
 * the nx*ny nodes of an FP64 population array f in the given layout;
 * ux_prev/uy_prev ([k][l]) are then overwritten with the current velocity.
 * A node whose difference is NaN (a diverged lattice) makes out = +infinity. */
int lbm_velocity_residual_f64(const double *f, int64_t nx, int64_t ny, int layout,
                              double *ux_prev, double *uy_prev, double *work,
                              double *out, void *stream);
/* The same residual without a workspace argument and without allocating:
 * its partial maxima go to a library-owned device array of
 * lbm_noalloc_scratch_bytes() bytes (8 KiB; one per GPU, held until the
 * process ends; every call writes all of it before reading it, so nothing
 * needs clearing).  Accepts every grid lbm_velocity_residual_f64 accepts
 * (LBM_E_SIZE beyond them) and gives bit-identical results.  Calls that
 * overlap on one GPU (other streams or threads) share that array and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight convergence loops).           */
int64_t lbm_noalloc_scratch_bytes(void);
int lbm_velocity_residual_f64_noalloc(const double *f, int64_t nx, int64_t ny, int layout,
                                      double *ux_prev, double *uy_prev, double *out,
                                      void *stream);

/* ---- reduced-precision storage ---------------------------------------------------
 * Populations stored as scaled deviations d_i = scale * (f_i - w_i) in FP32,
 * BF16 or FP16 (fmt), in either layout, with the collide computed in FP32.
 * FP64 is used only for encode/decode, the mass and the residual.
 * No reference counterpart (the reference stores f itself, cavity_opt1.py:59). */
#define LBM_FMT_F32 0
#define LBM_FMT_BF16 1
#define LBM_FMT_F16 2
int lbm_dev_variant_count(void);
/* FP64 populations f (reference layout, 9 x nx x ny) -> deviations d (layout) */
int lbm_dev_encode(const double *f, void *d, int64_t nx, int64_t ny, int layout, int fmt,
                   double scale, void *stream);
/* deviations d (layout) -> FP64 populations f (reference layout) */
int lbm_dev_decode(const void *d, double *f, int64_t nx, int64_t ny, int layout, int fmt,
                   double scale, void *stream);
/* Rounding of the 16-bit formats on store.  Both round the nine values of a
 * node so that its mass and momentum carry a single rounding error (the
 * momentum residual of the other channels moves into W/S, the mass residual
 * into the rest population).  NEAREST rounds to nearest even; STOCHASTIC
 * rounds up with probability equal to the distance from the value below,
 * driven by a hash of (seed, node): unbiased, and deterministic for a given
 * seed.  FP32 storage always rounds to nearest.                               */
#define LBM_ROUND_NEAREST 0
#define LBM_ROUND_STOCHASTIC 1
/* Compute type of the collide: FP32 (the reduced-precision form) or FP64,
 * for every storage format.  With FP64 compute the stored values are single
 * roundings of the FP64 results (FP32 storage then rounds like the 16-bit
 * formats: compensated, nearest or stochastic); with FP32 compute FP32
 * storage needs no rounding and `rounding` is ignored for it.  A flow that
 * changes by a few storage ulps per step (a long shear wave) needs FP64
 * compute with stochastic rounding: an FP32 collide's own rounding errors
 * repeat every step and bias it.                                            */
#define LBM_COMPUTE_F32 0
#define LBM_COMPUTE_F64 1
/* One fused step src -> dst on stored deviations; ux/uy (nullable, FP32
 * [k][l]) receive the pre-collision velocity.  variant in
 * 0 .. lbm_dev_variant_count()-1 or -1.  omega, u0 and scale are taken in
 * the compute type: the FP64 collide uses them as given, the FP32 collide
 * rounds them to FP32.                                                       */
int lbm_dev_step(const void *src, void *dst, int fmt, int64_t nx, int64_t ny, int layout,
                 double omega, int bc, double u0, double scale, int compute, int rounding,
                 uint32_t seed, float *ux, float *uy, int variant, void *stream);
/* K steps src -> dst in one launch (temporal blocking, K = 2..4 fixed by the
 * variant), as lbm_step_multi_* for stored deviations.  round_each = 1: LDS
 * holds the stored values, so step j of the launch (j = 0..K-1) rounds exactly
 * as a lbm_dev_step call with seed + j * seed_stride (mod 2^32), and the
 * result is bitwise that of those K calls.  round_each = 0 (store every K
 * steps): the intermediate steps stay unrounded in the compute type and only
 * the last is rounded, with seed + (K-1) * seed_stride; fewer roundings, so a
 * different and smaller rounding error (LBM_E_ARG for an FP64-compute shape
 * whose tile would not fit the 160 KB of LDS).  ux/uy receive the last step's
 * velocity.  lbm_dev_multi_info gives the variant's K and smallest grid.      */
int lbm_dev_step_multi(const void *src, void *dst, int fmt, int64_t nx, int64_t ny, int layout,
                       double omega, int bc, double u0, double scale, int compute, int rounding,
                       uint32_t seed, uint32_t seed_stride, int round_each, float *ux, float *uy,
                       int variant, void *stream);
int lbm_dev_multi_variant_count(void);
int lbm_dev_multi_default_variant(void); /* the variant that -1 selects */
int lbm_dev_multi_info(int variant, int *steps, int64_t *min_nx, int64_t *min_ny);
/* *out = sum_{i,k,l} f = nx*ny + (sum of stored values) / scale, in FP64     */
int lbm_dev_mass(const void *d, int64_t nx, int64_t ny, int fmt, double scale, double *work,
                 double *out, void *stream);
/* velocity residual of stored deviations (velocity evaluated in FP64); NaN
 * anywhere gives +infinity, as lbm_velocity_residual_f64                     */
int lbm_dev_velocity_residual(const void *d, int64_t nx, int64_t ny, int layout, int fmt,
                              double scale, double *ux_prev, double *uy_prev, double *work,
                              double *out, void *stream);

/* ---- batched instances ------------------------------------------------------------
 * Many independent lattices of different sizes in one population pool, each
 * stored as deviations in the row-interleaved layout at its own offset, and
 * stepped together: one launch per step covers every instance of a launch
 * group (a persistent grid looping over node tiles).  No reference
 * counterpart: the reference runs one lattice per process
 * (shear_wave_opt1.py:43-53, cavity_opt1.py:51-59).                           */
typedef struct lbm_batch_inst {
    int64_t offset; /* element offset of the instance's 9*nx*ny values in the pool */
    int32_t nx, ny;
    double omega;   /* in the compute type, as lbm_dev_step */
    double u0;      /* lid velocity when bc == LBM_BC_CAVITY */
    int32_t bc;     /* LBM_BC_PERIODIC or LBM_BC_CAVITY */
    int32_t dir;    /* observable: 0 = amplitude of u_x along y, 1 = of u_y along x */
} lbm_batch_inst;

typedef struct lbm_batch_tile {
    int32_t inst;  /* instance index */
    int32_t first; /* first node k*ny + l of the tile */
} lbm_batch_tile;

int lbm_batch_variant_count(void);
/* Nodes per tile of a batched launch shape (threads per block x nodes per thread). */
int lbm_batch_tile_nodes(int variant);
/* One step of every tile in tiles[0..ntiles) (tile size lbm_batch_tile_nodes(variant)),
 * src pool -> dst pool.  blocks = 0 picks the persistent grid size.  The dither
 * seed of stochastic rounding is seed combined with *epoch (device counter,
 * nullable), so graph replays get fresh streams.                               */
int lbm_batch_step(const void *src, void *dst, int fmt, float scale,
                   const lbm_batch_inst *insts, const lbm_batch_tile *tiles, int ntiles,
                   int compute, int rounding, uint32_t seed, const uint32_t *epoch, int variant,
                   int blocks, void *stream);
/* K steps (K = 2 or 4) of the instances of a work list in one launch, by a
 * sweep along x that keeps the intermediate steps in registers: item i is
 * worked by one wave of 64 lanes, which writes the rows [row0, row0 + rows)
 * of strip `strip` of instance `inst` (strip s covers columns
 * s (64 - 2(K-1)) .. + 64 - 2(K-1); an instance of 64 columns is one strip;
 * lbm_batch_sweep_strips gives the count).  Every step is rounded as
 * lbm_batch_step rounds it, step j (0 .. K-1) with seed + j * seed_stride
 * combined with *epoch, so the launch is bitwise K lbm_batch_step launches.
 * The items must cover each instance's rows and strips exactly once; every
 * instance needs ny >= 64 and nx >= 2K.                                      */
typedef struct lbm_batch_item {
    int32_t inst, strip, row0, rows;
} lbm_batch_item;
int lbm_batch_sweep_strips(int steps, int ny);
int lbm_batch_step_multi(const void *src, void *dst, int fmt, float scale,
                         const lbm_batch_inst *insts, const lbm_batch_item *items, int nitems,
                         int steps, int compute, int rounding, uint32_t seed, uint32_t seed_stride,
                         const uint32_t *epoch, void *stream);
/* Shear-wave amplitude of every instance by projection on its initial sine,
 * A = 2/(nx ny) sum u_dir sin(2 pi pos / n), FP64; written to
 * amp[*epoch * ninst + i] when *epoch < nslots.  tiles: 256-node tiles in
 * instance order; inst_tiles[i] .. inst_tiles[i+1] are instance i's tiles;
 * partial holds ntiles doubles.                                               */
int lbm_batch_amplitude(const void *pool, int fmt, float scale, const lbm_batch_inst *insts,
                        int ninst, const lbm_batch_tile *tiles, int ntiles,
                        const int32_t *inst_tiles, double *partial, double *amp, int nslots,
                        const uint32_t *epoch, void *stream);
/* *counter += 1 on the device (a one-thread kernel, safe inside a graph).     */
int lbm_counter_increment(uint32_t *counter, void *stream);
/* *counter = value on the device (a one-thread kernel).                      */
int lbm_counter_set(uint32_t *counter, uint32_t value, void *stream);

/* ---- HIP graphs ---------------------------------------------------------------------
 * Capture the work enqueued on a (non-default) stream between begin and end
 * into an executable graph, then replay it with one launch.                   */
typedef struct lbm_graph lbm_graph;
int lbm_graph_begin(void *stream);
int lbm_graph_end(void *stream, lbm_graph **graph);
int lbm_graph_launch(lbm_graph *graph, void *stream);
int lbm_graph_destroy(lbm_graph *graph);

/* ---- Warm-start surrogate --------------------------------------------------------------
 * No counterpart in the reference (it has no learned component).  Kernels of a
 * small convolutional encoder-decoder (3x3 same convolutions as im2col + GEMM on
 * BF16 MFMA with FP32 accumulation, 2x2 average pooling, nearest 2x
 * upsampling, channel concatenation) with their backward passes, an MSE loss
 * and Adam, plus the initializer that turns the network's prediction into
 * populations.  Activations are NHWC; BF16 buffers are raw 16-bit words.
 * Every reduction runs in a fixed order, so results are deterministic.        */

/* Bytes of the split-K workspace lbm_nn_gemm needs (0 when ksplit <= 1).     */
int64_t lbm_nn_gemm_workspace(int M, int N, int ksplit);
/* C[m*ldc + n] (+)= sum_k A[m*sam + k*sak] B[k*sbk + n*sbn]  (A, B BF16),
 * then + bias[n] (nullable), then max(., 0) if relu; written to C (FP32,
 * nullable) and/or Cb (BF16, nullable).  accumulate adds the old C.  ksplit > 1
 * splits K over that many block planes into work (lbm_nn_gemm_workspace bytes)
 * and sums them in order.                                                    */
int lbm_nn_gemm(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn,
                float *C, void *Cb, int64_t ldc, const float *bias, int M, int N, int K,
                int accumulate, int relu, int ksplit, float *work, void *stream);
/* 3x3 same convolutions without a materialised im2col matrix: the GEMM's
 * tile loader gathers the shifted activations itself (16-byte loads when the
 * gathered channel count is a multiple of 8).  x, dy BF16 NHWC; w BF16
 * [cout][9*cin] ((dy*3 + dx)*cin + c); the sums run in FP32 over the same k
 * order as im2col + lbm_nn_gemm, so the forward and weight gradient are
 * bitwise those of the materialised form.
 *   conv3x3:  y[p][co] = relu?(sum_k x(p, k) w[co][k] + bias[co])  (y FP32 and/or yb BF16)
 *   wgrad:    dw[co][k] = sum_p dy[p][co] x(p, k)   (split-K over pixels as lbm_nn_gemm)
 *   dgrad:    dx[p][c] (+)= sum_{tap, co} dy[p - tap][co] wf[tap][co][c], wf from
 *             flip_weights (wf[tap][co][c] = w[co][tap*cin + c]).               */
int lbm_nn_conv3x3(const void *x, int B, int H, int W, int cin, const void *w, int cout,
                   const float *bias, float *y, void *yb, int relu, void *stream);
int lbm_nn_conv3x3_wgrad(const void *x, const void *dy, int B, int H, int W, int cin, int cout,
                         float *dw, int ksplit, float *work, void *stream);
int lbm_nn_conv3x3_flip_weights(const void *w, int cin, int cout, void *wf, void *stream);
int lbm_nn_conv3x3_dgrad(const void *dy, const void *wf, int B, int H, int W, int cin, int cout,
                         float *dx, int accumulate, void *stream);
/* cols[((b*H + y)*W + x)*9*C + (dy*3 + dx)*C + c] = x[b][y+dy-1][x+dx-1][c] (0 outside). */
int lbm_nn_im2col3x3(const void *x, int B, int H, int W, int C, void *cols, void *stream);
/* Adjoint of im2col3x3 (a gather, FP32): dx (+)= col2im(dcols).              */
int lbm_nn_col2im3x3(const float *dcols, int B, int H, int W, int C, float *dx, int accumulate,
                     void *stream);
/* dy *= (y > 0), with an optional BF16 copy of the result.                   */
int lbm_nn_relu_backward(const float *y, float *dy, void *dy_bf16, int64_t n, void *stream);
/* db[n] = sum_m dy[m*N + n], FP64 accumulation in a fixed order; work holds
 * lbm_nn_bias_grad_workspace(N) bytes.                                       */
int64_t lbm_nn_bias_grad_workspace(int N);
int lbm_nn_bias_grad(const float *dy, int64_t M, int N, float *db, double *work, void *stream);
/* 2x2 average pool x [B][H][W][C] -> y [B][H/2][W/2][C] (+ optional BF16 copy). */
int lbm_nn_avgpool2(const float *x, int B, int H, int W, int C, float *y, void *yb, void *stream);
int lbm_nn_avgpool2_backward(const float *dy, int B, int H, int W, int C, float *dx,
                             int accumulate, void *stream);
/* Nearest 2x upsample x [B][H][W][C] into channels [coff, coff+C) of
 * y [B][2H][2W][Ct] (+ optional BF16 copy); backward sums the four children. */
int lbm_nn_upsample2(const float *x, int B, int H, int W, int C, float *y, void *yb, int Ct,
                     int coff, void *stream);
int lbm_nn_upsample2_backward(const float *dy, int B, int H, int W, int C, int Ct, int coff,
                              float *dx, void *stream);
/* dst[p*Cd + doff + c] (+)= src[p*Cs + soff + c], c < C, p < P (+ BF16 copy). */
int lbm_nn_copy_channels(const float *src, int Cs, int soff, float *dst, void *dst_bf16, int Cd,
                         int doff, int C, int64_t P, int accumulate, void *stream);
/* loss = mean((y - t)^2) (FP64, via work of lbm_reduce_workspace_bytes);
 * dy = 2 (y - t) / n.                                                        */
int lbm_nn_mse(const float *y, const float *t, int64_t n, float *dy, double *work, double *loss,
               void *stream);
/* Large-scale error term over y, t [B][H][W][C] (H, W multiples of f):
 * e_bar = mean of (y - t) over each f x f block; *loss = weight * mean(e_bar^2)
 * (FP64, via work of lbm_reduce_workspace_bytes); dy += its gradient.       */
int lbm_nn_mse_pooled(const float *y, const float *t, int B, int H, int W, int C, int f,
                      float weight, float *dy, double *work, double *loss, void *stream);
/* One Adam update (bias-corrected for step >= 1) of FP32 p, with BF16 copy.  */
int lbm_nn_adam(float *p, const float *g, float *m, float *v, void *p_bf16, int64_t n, float lr,
                float b1, float b2, float eps, int step, void *stream);
/* ema = decay * ema + (1 - decay) * p, 0 <= decay < 1 (weight averaging).    */
int lbm_nn_ema(float *ema, const float *p, int64_t n, float decay, void *stream);
int lbm_nn_to_bf16(const float *x, void *y, int64_t n, void *stream);
/* Warm-start populations (FP64, layout f[i][k][l]) from a prediction
 * q[m][m][3] = ((rho-1)/U^2, ux/U, uy/U), averages over the cells of an m x m
 * grid over the cavity, reconstructed at the nodes, then f = f_eq(rho, u).
 * Reconstruction (flags): default bilinear between cell centres, clamped at
 * the outermost centres; LBM_WARM_WALLS bilinear towards ghost cells mirrored
 * across the walls (u = 0 on the walls, (U, 0) on the lid, zero density
 * gradient); LBM_WARM_CUBIC (m >= 3, same ghosts) cell averages turned into
 * centre values, q - c lap(q) with c = (pool^2 - 1) / (24 pool^2) for averages
 * over pool x pool nodes (pool = 1: none), then separable Catmull-Rom cubics.
 * LBM_WARM_NEQ adds the first-order non-equilibrium part of a post-collision
 * BGK state, (1 - omega) * (-3 w_i rho / omega) (c_i c_i - 1/3) : S, S the
 * strain rate of the reconstructed u.                                       */
#define LBM_WARM_NEQ 1
#define LBM_WARM_WALLS 2
#define LBM_WARM_CUBIC 4
int lbm_warm_init_f64(const float *q, int m, int64_t nx, int64_t ny, double U, double omega,
                      int flags, int pool, double *f, void *stream);

/* ---- ghost-layer flows --------------------------------------------------------------
 * The reference's simple_flows / experimantal_flows scripts (Couette and
 * Poiseuille channels, PoiseuilleFlow.py:59-150; sliding lid, slidingLid.py:
 * 68-108; obstacle canal, obstacle_canal.py:265-458) keep a ghost layer inside
 * the (9, nx, ny) array and per step:
 *   1. optionally replace columns 0 and nx-1 by the pressure-periodic inlet /
 *      outlet values, f = f_eq(rho_in|out, u(column nx-2|1)) + f(column) -
 *      f_eq(column), before streaming (PoiseuilleFlow.py:76-89) or after it
 *      ((f_eq + f) - f_eq, obstacle_canal.py:362-373);
 *   2. stream periodically over the whole array (np.roll);
 *   3. overwrite populations by the rules, in order: target nodes x0 <= x < x1,
 *      y0 <= y < y1, channel ch, take the streamed value of channel src_ch at
 *      (x + dx, y + dy), plus coef (LBM_RULE_LID) or coef times the wall
 *      density 2 (g2 + g5 + g6) + g0 + g1 + g3 at the source node
 *      (LBM_RULE_LID_RHO, slidingLid.py:88-91), or 0 (LBM_RULE_ZERO);
 *   4. collide every node with the scripts' equilibrium arithmetic
 *      (PoiseuilleFlow.py:25-53): f -= omega (f - f_eq).
 * One launch reads each population once and writes it once (src -> dst), in
 * either layout (LBM_LAYOUT_SOA is the scripts' own; LBM_LAYOUT_ROWS streams
 * faster; the results are the same bits); rho/ux/uy are [k][l].  A
 * rule may not read a population an earlier rule writes (LBM_E_ARG), so every
 * rule reads the streamed values.  rules and pbc are host pointers, copied
 * into the launch.  flags: LBM_GHOST_NO_COLLIDE stops after step 3,
 * LBM_GHOST_COLLIDE_ONLY does step 4 alone.  rho/ux/uy (nullable, all or
 * none) receive the moments before the collision.  nx, ny >= 3, nx <= 65535.
 * Replaces the step functions of those scripts (no PyLB counterpart).         */
#define LBM_RULE_COPY 0
#define LBM_RULE_LID 1
#define LBM_RULE_LID_RHO 2
#define LBM_RULE_ZERO 3
#define LBM_GHOST_MAX_RULES 32
#define LBM_PBC_NONE 0
#define LBM_PBC_BEFORE_STREAM 1
#define LBM_PBC_AFTER_STREAM 2
#define LBM_GHOST_NO_COLLIDE 1
#define LBM_GHOST_COLLIDE_ONLY 2
typedef struct lbm_ghost_rule {
    int32_t x0, x1, y0, y1; /* target nodes                                    */
    int32_t ch, kind;       /* target channel, LBM_RULE_*                      */
    int32_t dx, dy, src_ch; /* source node offset and channel                  */
    int32_t pad;
    double coef; /* signed: -uw/6 or +uw/6 in the scripts                      */
} lbm_ghost_rule;
typedef struct lbm_ghost_pbc {
    int32_t mode; /* LBM_PBC_*                                                 */
    int32_t pad;
    double rho_in, rho_out;
} lbm_ghost_pbc;
/* f[i][k] = the scripts' equilibrium of (rho, ux, uy)[k], k < n
 * (PoiseuilleFlow.py:25-43, the initial state of every ghost-layer flow).       */
int lbm_ghost_equilibrium_f64(const double *rho, const double *ux, const double *uy, double *f,
                              int64_t n, void *stream);
int lbm_ghost_equilibrium_f32(const float *rho, const float *ux, const float *uy, float *f,
                              int64_t n, void *stream);
int lbm_ghost_step_f64(const double *src, double *dst, int64_t nx, int64_t ny, int layout,
                       double omega, const lbm_ghost_rule *rules, int nrules,
                       const lbm_ghost_pbc *pbc, int flags, double *rho, double *ux, double *uy,
                       void *stream);
int lbm_ghost_step_f32(const float *src, float *dst, int64_t nx, int64_t ny, int layout,
                       double omega, const lbm_ghost_rule *rules, int nrules,
                       const lbm_ghost_pbc *pbc, int flags, float *rho, float *ux, float *uy,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LBM_D2Q9_H */
