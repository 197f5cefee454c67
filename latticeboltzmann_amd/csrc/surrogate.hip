// Kernels of the warm-start surrogate: a small convolutional encoder-decoder
// trained and evaluated on the GPU, plus the initializer that turns its
// prediction into populations.
//
// Convolutions are 3x3, stride 1, zero padding, NHWC activations.  They run
// as im2col + GEMM: cols[pixel][(dy*3+dx)*C + c] (BF16) times the weights
// W[cout][(dy*3+dx)*C + c] (BF16), accumulated in FP32 by v_mfma_f32_16x16x32_bf16
// (gfx950).  Backward: dW = dY^T cols (split-K over pixels with a fixed-order
// reduction), dcols = dY W, dX = col2im(dcols) as a gather (no atomics), so
// every result is deterministic.  Activations are kept in FP32 with a BF16
// copy for the next GEMM; weights in FP32 (Adam) with a BF16 copy.
//
// No reference counterpart: the reference has no learned component.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d2q9_node.hpp"
#include "d2q9_reduce.hpp"
#include "lbm_internal.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace lbm {
namespace nn {

__device__ __forceinline__ float bf2f(uint16_t x) { return float(__builtin_bit_cast(__bf16, x)); }
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, __bf16(x)); }

// ---------------------------------------------------------------------------
// GEMM: C[m][n] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn] (BF16), FP32 accumulation.
// Block: 256 threads = 4 waves, tile 64 x 64, k-step 32 staged in LDS as
// As[m][k] and Bs[n][k]; each wave owns a 32 x 32 quarter = 2 x 2 MFMA
// 16x16x32 tiles, and reads each operand as 8 consecutive k of one row.
// MFMA operand layout (checked on gfx950): lane l holds A rows l%16,
// k = (l/16)*8 + i; B cols l%16, k = (l/16)*8 + i; result register r of lane l
// is row (l/16)*4 + r, col l%16.
// Tile loads (per operand, chosen on the host): LOAD_K: k contiguous, each
// thread loads 8 k of one row with one 16-byte load; LOAD_ROW: the row index
// contiguous, each thread loads 8 rows of one k; LOAD_SCALAR otherwise.  The
// vector forms need 16-byte alignment of the base and of the other stride
// and fall back to bounds-checked scalar loads at partial tiles.
// Split-K: blockIdx.z handles k in [z*kchunk, (z+1)*kchunk); with more than
// one split, results go to work[z][m][n] and k_gemm_reduce adds them in order.
constexpr int TM = 64, TN = 64, TK = 32, TP = TK + 8;
constexpr int LOAD_SCALAR = 0, LOAD_K = 1, LOAD_ROW = 2;

// A convolution operand gathered on the fly instead of read from a
// materialised im2col matrix: element (pixel p, index q) with q = tap * C + c,
// tap = (dy + 1) * 3 + (dx + 1), is X[pixel p + sign * (dy, dx)][c] of the NHWC
// BF16 activation X (H x W images, C channels), 0 outside the image.  sign +1
// is the im2col of a 3x3 same convolution; sign -1 gathers, for input pixel
// p, the output pixels that read p (the data gradient).  sign 0: no gather.
struct Gather {
    int H, W, C, sign;
};

struct GemmArgs {
    const uint16_t *A;
    int64_t sam, sak;
    const uint16_t *B;
    int64_t sbk, sbn;
    float *C;      // FP32 output (nullable)
    uint16_t *Cb;  // BF16 output (nullable)
    int64_t ldc;
    const float *bias;
    int M, N, K, kchunk, accumulate, relu;
    float *work;   // split-K partials
    int mode_a, mode_b;
    Gather ga;     // A(m, k) gathered with pixel m, index k (ga.sign != 0)
    Gather gb;     // B(k, n) gathered with pixel k, index n (gb.sign != 0)
};

// Index into X of gathered element (pixel p, index q), and whether it lies in
// the image.
__device__ __forceinline__ int64_t gather_index(const Gather &G, int p, int q, bool &inside) {
    const int tap = q / G.C, c = q - tap * G.C;
    const int dy = (tap / 3 - 1) * G.sign, dx = (tap % 3 - 1) * G.sign;
    const int x = p % G.W, y = (p / G.W) % G.H;
    inside = unsigned(y + dy) < unsigned(G.H) && unsigned(x + dx) < unsigned(G.W);
    return (int64_t(p) + int64_t(dy) * G.W + dx) * G.C + c;
}

__device__ __forceinline__ uint16_t gather_one(const uint16_t *__restrict__ X, const Gather &G,
                                               int p, int q) {
    bool in;
    const int64_t i = gather_index(G, p, q, in);
    return in ? X[i] : uint16_t(0);
}

// tile[row][k] <- X(row0 + row, k0 + k) for a gathered operand, zero outside
// rows < nrows, k < kend.  pix_row: the row is the pixel and k the (tap,
// channel) index (A of the forward and data-gradient GEMMs); otherwise k is
// the pixel and the row the index (B of the weight-gradient GEMM).  With C a
// multiple of 8, the 8 indices a thread loads lie in one tap: one 16-byte load.
__device__ __forceinline__ void load_tile_gather(uint16_t (*tile)[TP], const uint16_t *__restrict__ X,
                                                 const Gather &G, int row0, int nrows, int k0,
                                                 int kend, bool pix_row, int t) {
    const bool vec = G.C % 8 == 0;
    if (pix_row) {
        const int rr = t >> 2, kc = (t & 3) * 8;
        const int r = row0 + rr, k = k0 + kc;
        if (vec && r < nrows && k + 8 <= kend) {
            bool in;
            const int64_t i = gather_index(G, r, k, in);
            uint4 v = make_uint4(0, 0, 0, 0);
            if (in) v = *reinterpret_cast<const uint4 *>(X + i);
            *reinterpret_cast<uint4 *>(&tile[rr][kc]) = v;
            return;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
            tile[rr][kc + i] = (r < nrows && k + i < kend) ? gather_one(X, G, r, k + i) : uint16_t(0);
        return;
    }
    const int kr = t >> 3, rc = (t & 7) * 8;
    const int k = k0 + kr, r = row0 + rc;
    if (vec && k < kend && r + 8 <= nrows) {
        bool in;
        const int64_t i = gather_index(G, k, r, in);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (in) v = *reinterpret_cast<const uint4 *>(X + i);
        const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
        for (int i = 0; i < 8; ++i) tile[rc + i][kr] = e[i];
        return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
        tile[rc + i][kr] = (k < kend && r + i < nrows) ? gather_one(X, G, k, r + i) : uint16_t(0);
}

// tile[row][k] <- X(row0 + row, k0 + k) for row < 64, k < 32; X(r, k) =
// X[r*s_row + k*s_k], zero outside rows < nrows, k < kend.
__device__ __forceinline__ void load_tile(uint16_t (*tile)[TP], const uint16_t *__restrict__ X,
                                          int64_t s_row, int64_t s_k, int row0, int nrows, int k0,
                                          int kend, int mode, int t) {
    if (mode == LOAD_ROW) {
        const int kr = t >> 3, rc = (t & 7) * 8;
        const int k = k0 + kr, r = row0 + rc;
        if (k < kend && r + 8 <= nrows) {
            const uint4 v = *reinterpret_cast<const uint4 *>(X + int64_t(k) * s_k + r);
            const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
            for (int i = 0; i < 8; ++i) tile[rc + i][kr] = e[i];
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                tile[rc + i][kr] = (k < kend && r + i < nrows) ? X[int64_t(k) * s_k + r + i] : 0;
        }
        return;
    }
    const int rr = t >> 2, kc = (t & 3) * 8;
    const int r = row0 + rr, k = k0 + kc;
    if (mode == LOAD_K && r < nrows && k + 8 <= kend) {
        *reinterpret_cast<uint4 *>(&tile[rr][kc]) =
            *reinterpret_cast<const uint4 *>(X + int64_t(r) * s_row + k);
        return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
        tile[rr][kc + i] = (r < nrows && k + i < kend) ? X[int64_t(r) * s_row + int64_t(k + i) * s_k] : 0;
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __align__(16) uint16_t As[TM][TP];
    __shared__ __align__(16) uint16_t Bs[TN][TP];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        if (g.ga.sign) load_tile_gather(As, g.A, g.ga, m0, g.M, k0, kend, true, t);
        else load_tile(As, g.A, g.sam, g.sak, m0, g.M, k0, kend, g.mode_a, t);
        if (g.gb.sign) load_tile_gather(Bs, g.B, g.gb, n0, g.N, k0, kend, false, t);
        else load_tile(Bs, g.B, g.sbn, g.sbk, n0, g.N, k0, kend, g.mode_b, t);
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + i * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + j * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    const bool split = gridDim.z > 1;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][r];
                if (split) {
                    g.work[(int64_t(blockIdx.z) * g.M + m) * g.N + n] = v;
                    continue;
                }
                const int64_t o = int64_t(m) * g.ldc + n;
                if (g.accumulate) v += g.C[o];
                if (g.bias) v += g.bias[n];
                if (g.relu) v = fmaxf(v, 0.0f);
                if (g.C) g.C[o] = v;
                if (g.Cb) g.Cb[o] = f2bf(v);
            }
}

// Fixed-order sum of the split-K partials plus the epilogue.
__global__ __launch_bounds__(256) void k_gemm_reduce(GemmArgs g, int splits) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(g.M) * g.N) return;
    const int m = int(idx / g.N), n = int(idx - int64_t(m) * g.N);
    float v = 0.0f;
    for (int z = 0; z < splits; ++z) v += g.work[int64_t(z) * g.M * g.N + idx];
    const int64_t o = int64_t(m) * g.ldc + n;
    if (g.accumulate) v += g.C[o];
    if (g.bias) v += g.bias[n];
    if (g.relu) v = fmaxf(v, 0.0f);
    if (g.C) g.C[o] = v;
    if (g.Cb) g.Cb[o] = f2bf(v);
}

// im2col for 3x3 same convolution: x [B][H][W][C] BF16 -> cols [B*H*W][9*C].
// Tap (dy, dx) = (blockIdx.y / 3, blockIdx.y % 3); each thread moves VEC
// consecutive channels of one pixel (VEC = 8: one 16-byte load and store,
// used when C % 8 == 0).  Indices are 32-bit: the caller checks B*H*W*C < 2^31.
template <int VEC>
__global__ __launch_bounds__(256) void k_im2col(const uint16_t *__restrict__ x, int H, int W, int C,
                                                uint16_t *__restrict__ cols, int items) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= items) return;  // items = B*H*W*C/VEC
    const int cv = C / VEC;
    const int c = (idx % cv) * VEC;
    const int p = idx / cv;
    const int xw = p % W, r = p / W;
    const int yh = r % H, b = r / H;
    const int tap = blockIdx.y;
    const int yy = yh + tap / 3 - 1, xx = xw + tap % 3 - 1;
    const bool inside = yy >= 0 && yy < H && xx >= 0 && xx < W;
    const int64_t src = ((int64_t(b) * H + yy) * W + xx) * C + c;
    const int64_t dst = (int64_t(p) * 9 + tap) * C + c;
    if constexpr (VEC == 8) {
        uint4 v = make_uint4(0, 0, 0, 0);
        if (inside) v = *reinterpret_cast<const uint4 *>(x + src);
        *reinterpret_cast<uint4 *>(cols + dst) = v;
    } else {
        cols[dst] = inside ? x[src] : uint16_t(0);
    }
}

// col2im (gather): dx[b][y][x][c] (+)= sum over taps of dcols[pixel(y - dy + 1, x - dx + 1)][tap*C + c].
__global__ __launch_bounds__(256) void k_col2im(const float *__restrict__ dcols, int B, int H, int W,
                                                int C, float *__restrict__ dx, int accumulate) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t total = int64_t(B) * H * W * C;
    if (idx >= total) return;
    const int c = int(idx % C);
    int64_t r = idx / C;
    const int xw = int(r % W);
    const int64_t r2 = r / W;
    const int yh = int(r2 % H);
    const int b = int(r2 / H);
    float s = 0.0f;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int py = yh - (tap / 3 - 1), px = xw - (tap % 3 - 1);  // output pixel that read (yh, xw) at this tap
        if (py < 0 || py >= H || px < 0 || px >= W) continue;
        s += dcols[(((int64_t(b) * H + py) * W + px) * 9 + tap) * C + c];
    }
    dx[idx] = accumulate ? dx[idx] + s : s;
}

// dy *= (y > 0); writes a BF16 copy of the masked gradient.
__global__ __launch_bounds__(256) void k_relu_backward(const float *__restrict__ y, float *__restrict__ dy,
                                                       uint16_t *__restrict__ dyb, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = y[i] > 0.0f ? dy[i] : 0.0f;
    dy[i] = g;
    if (dyb) dyb[i] = f2bf(g);
}

// db[n] = sum_m dy[m][n] in two fixed-order passes.  Pass 1: block b sums
// rows [b*rows, (b+1)*rows) of columns [c0, c0 + cw), cw = min(N - c0, 256);
// thread t owns column c0 + t % cw and rows t / cw, + 256 / cw, ... (coalesced
// across a row), the threads of one column are then added in LDS in order.
constexpr int kBiasBlocks = 256;

__global__ __launch_bounds__(256) void k_bias_grad_partial(const float *__restrict__ dy, int64_t M,
                                                           int N, int c0, int64_t rows,
                                                           double *__restrict__ partial) {
    __shared__ double red[256];
    const int cw = min(N - c0, 256), t = threadIdx.x;
    const int lanes = 256 / cw;  // threads per column
    const int col = t % cw, r0 = t / cw;
    double s = 0.0;
    if (r0 < lanes) {
        const int64_t beg = int64_t(blockIdx.x) * rows, end = min(M, beg + rows);
        for (int64_t m = beg + r0; m < end; m += lanes) s += double(dy[m * N + c0 + col]);
    }
    red[t] = s;
    __syncthreads();
    if (t < cw) {
        double v = 0.0;
        for (int j = 0; j < lanes; ++j) v += red[j * cw + t];
        partial[int64_t(blockIdx.x) * N + c0 + t] = v;
    }
}

// Pass 2: block n adds the nblocks (<= 256) partials of column n, one per
// thread, in a fixed-shape tree, so all of them are loaded at once instead of
// one after another by a single thread.
__global__ __launch_bounds__(256) void k_bias_grad_final(const double *__restrict__ partial, int N,
                                                         int nblocks, float *__restrict__ db) {
    const int n = blockIdx.x, t = threadIdx.x;
    const double v = t < nblocks ? partial[int64_t(t) * N + n] : 0.0;
    const double tot = block_sum(v);
    if (t == 0) db[n] = float(tot);
}

// Weights W[cout][tap][cin] -> Wf[tap][cout][cin] (BF16): the B operand of the
// data-gradient GEMM, dX[p][c] = sum over (tap, co) of dY[p - tap][co] Wf[tap][co][c].
__global__ __launch_bounds__(256) void k_flip_weights(const uint16_t *__restrict__ w, int cin,
                                                      int cout, uint16_t *__restrict__ wf) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // over tap, cout, cin
    if (i >= 9 * cout * cin) return;
    const int c = i % cin, r = i / cin;
    const int co = r % cout, tap = r / cout;
    wf[i] = w[(int64_t(co) * 9 + tap) * cin + c];
}

// 2x2 average pool (NHWC): y [B][H/2][W/2][C].
__global__ __launch_bounds__(256) void k_avgpool2(const float *__restrict__ x, int B, int H, int W, int C,
                                                  float *__restrict__ y, uint16_t *__restrict__ yb) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(B) * Ho * Wo * C) return;
    const int c = int(idx % C);
    int64_t r = idx / C;
    const int xo = int(r % Wo);
    const int64_t r2 = r / Wo;
    const int yo = int(r2 % Ho);
    const int b = int(r2 / Ho);
    const float *p = x + ((int64_t(b) * H + 2 * yo) * W + 2 * xo) * C + c;
    const float v = 0.25f * ((p[0] + p[C]) + (p[int64_t(W) * C] + p[int64_t(W) * C + C]));
    y[idx] = v;
    if (yb) yb[idx] = f2bf(v);
}

__global__ __launch_bounds__(256) void k_avgpool2_backward(const float *__restrict__ dy, int B, int H,
                                                           int W, int C, float *__restrict__ dx,
                                                           int accumulate) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(B) * H * W * C) return;
    const int c = int(idx % C);
    int64_t r = idx / C;
    const int xw = int(r % W);
    const int64_t r2 = r / W;
    const int yh = int(r2 % H);
    const int b = int(r2 / H);
    const float g = 0.25f * dy[((int64_t(b) * (H / 2) + yh / 2) * (W / 2) + xw / 2) * C + c];
    dx[idx] = accumulate ? dx[idx] + g : g;
}

// Nearest 2x upsample of x [B][H][W][C] into channels [coff, coff + C) of
// y [B][2H][2W][Ct].
__global__ __launch_bounds__(256) void k_upsample2(const float *__restrict__ x, int B, int H, int W,
                                                   int C, float *__restrict__ y,
                                                   uint16_t *__restrict__ yb, int Ct, int coff) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int Ho = 2 * H, Wo = 2 * W;
    if (idx >= int64_t(B) * Ho * Wo * C) return;
    const int c = int(idx % C);
    int64_t r = idx / C;  // output pixel
    const int xo = int(r % Wo);
    const int64_t r2 = r / Wo;
    const int yo = int(r2 % Ho);
    const int b = int(r2 / Ho);
    const float v = x[((int64_t(b) * H + yo / 2) * W + xo / 2) * C + c];
    const int64_t o = r * Ct + coff + c;
    y[o] = v;
    if (yb) yb[o] = f2bf(v);
}

__global__ __launch_bounds__(256) void k_upsample2_backward(const float *__restrict__ dy, int B, int H,
                                                            int W, int C, int Ct, int coff,
                                                            float *__restrict__ dx) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(B) * H * W * C) return;
    const int c = int(idx % C);
    int64_t r = idx / C;
    const int xw = int(r % W);
    const int64_t r2 = r / W;
    const int yh = int(r2 % H);
    const int b = int(r2 / H);
    const int Wo = 2 * W;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int yo = 2 * yh + q / 2, xo = 2 * xw + q % 2;
        s += dy[((int64_t(b) * (2 * H) + yo) * Wo + xo) * Ct + coff + c];
    }
    dx[idx] = s;
}

// Channel-slice copy: dst[p][doff + c] (+)= src[p][soff + c] for c < C.
__global__ __launch_bounds__(256) void k_copy_channels(const float *__restrict__ src, int Cs, int soff,
                                                       float *__restrict__ dst, uint16_t *__restrict__ dstb,
                                                       int Cd, int doff, int C, int64_t P, int accumulate) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= P * C) return;
    const int c = int(idx % C);
    const int64_t p = idx / C;
    float v = src[p * Cs + soff + c];
    const int64_t o = p * Cd + doff + c;
    if (accumulate) v += dst[o];
    dst[o] = v;
    if (dstb) dstb[o] = f2bf(v);
}

// Mean squared error over n values: loss partials and dy = 2 (y - t) / n.
__global__ __launch_bounds__(kBlock) void k_mse(const float *__restrict__ y, const float *__restrict__ t,
                                                int64_t n, float *__restrict__ dy,
                                                double *__restrict__ partial) {
    double s = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        const float d = y[i] - t[i];
        s += double(d) * double(d);
        dy[i] = 2.0f * d / float(n);
    }
    const double tot = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot / double(n);
}

// Large-scale error term: e_bar = mean of (y - t) over each f x f block of
// cells (per image and channel); loss partials of weight * mean(e_bar^2) and
// dy += weight * 2 e_bar / (n_blocks f^2) on every cell of the block.  One
// thread per block: each cell is written by exactly one thread.
__global__ __launch_bounds__(kBlock) void k_mse_pooled(const float *__restrict__ y,
                                                       const float *__restrict__ t, int B, int H,
                                                       int W, int C, int f, float weight,
                                                       float *__restrict__ dy,
                                                       double *__restrict__ partial) {
    const int Hb = H / f, Wb = W / f;
    const int64_t nb = int64_t(B) * Hb * Wb * C;
    const double inv = 1.0 / (double(f) * f);
    double s = 0.0;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x; q < nb; q += stride) {
        const int c = int(q % C);
        int64_t r = q / C;
        const int xb = int(r % Wb);
        r /= Wb;
        const int yb = int(r % Hb);
        const int b = int(r / Hb);
        const int64_t base = ((int64_t(b) * H + int64_t(yb) * f) * W + int64_t(xb) * f) * C + c;
        double e = 0.0;
        for (int i = 0; i < f; ++i)
            for (int j = 0; j < f; ++j) {
                const int64_t o = base + (int64_t(i) * W + j) * C;
                e += double(y[o]) - double(t[o]);
            }
        e *= inv;
        s += e * e;
        const float g = float(weight * 2.0 * e * inv / double(nb));
        for (int i = 0; i < f; ++i)
            for (int j = 0; j < f; ++j) dy[base + (int64_t(i) * W + j) * C] += g;
    }
    const double tot = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot * double(weight) / double(nb);
}

// Adam on FP32 master weights, BF16 copy for the GEMMs.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v,
                                              uint16_t *__restrict__ pb, int64_t n, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = f2bf(pi);
}

// Exponential moving average of the parameters: ema = decay * ema + (1 - decay) * p.
__global__ __launch_bounds__(256) void k_ema(float *__restrict__ ema, const float *__restrict__ p,
                                             int64_t n, float decay) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) ema[i] = decay * ema[i] + (1.0f - decay) * p[i];
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y,
                                                 int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// Warm start: the surrogate's normalized fields q [m][m][3] =
// ((rho - 1) / U^2, ux / U, uy / U), averages over the cells of an m x m grid
// covering the cavity -> populations f (FP64, reference layout 9 x nx x ny) on
// nx x ny nodes.  Node (k, l) sits at cell coordinate (px, py) =
// ((k + 1/2) m / nx - 1/2, (l + 1/2) m / ny - 1/2); the walls lie on the faces
// px = -1/2, px = m - 1/2, py = -1/2 and the lid on py = m - 1/2.
// Reconstruction of the node values:
//   * default: bilinear between cell centres, clamped beyond the outermost;
//   * WALLS: bilinear, with ghost cells mirrored across each wall beyond the
//     outermost centres: velocity ghosts 2 u_wall - u_mirror (u_wall = 0 on the
//     walls, (1, 0) on the lid, in units of U), density ghosts equal to the
//     mirror cell, so profiles run into the wall values;
//   * CUBIC (with the same ghosts, any depth): the cell averages are first
//     turned into point values at the centres, u_c = q - c * lap(q) (5-point
//     Laplacian; an r x r node average of a smooth field is its centre value
//     plus c * lap with c = (r^2 - 1) / (24 r^2)), then interpolated with
//     separable Catmull-Rom cubics (exact for quadratics);
// then f = f_eq(rho, u) in the opt1 order; with NEQ, plus the first-order
// non-equilibrium part of the BGK post-collision state,
// (1 - omega) * (-3 w_i rho / omega) Q_i : S, with the strain rate S from
// central differences of the reconstructed velocity (one-sided at the walls).
constexpr int kWarmNeq = 1, kWarmWalls = 2, kWarmCubic = 4;

// q at cell (i, j); cells outside [0, m) are ghosts mirrored across the wall
// faces (requires i, j in [-m, 2m)).
__device__ __forceinline__ double cell(const float *__restrict__ q, int m, int ch, int i, int j) {
    const int ii = i < 0 ? -1 - i : (i >= m ? 2 * m - 1 - i : i);
    const int jj = j < 0 ? -1 - j : (j >= m ? 2 * m - 1 - j : j);
    double v = q[(int64_t(ii) * m + jj) * 3 + ch];
    if (ch == 0) return v;
    if (j != jj) v = ((j >= m && ch == 1) ? 2.0 : 0.0) - v;  // bottom wall / lid
    if (i != ii) v = -v;                                   // left / right wall
    return v;
}

__device__ __forceinline__ double bilerp(const float *__restrict__ q, int m, int ch, double px,
                                         double py, bool walls) {
    int i0, j0, i1, j1;
    if (walls) {
        px = fmin(fmax(px, -1.0), double(m));
        py = fmin(fmax(py, -1.0), double(m));
        i0 = min(int(floor(px)), m - 1);
        j0 = min(int(floor(py)), m - 1);
        i1 = i0 + 1;
        j1 = j0 + 1;
    } else {
        px = fmin(fmax(px, 0.0), double(m - 1));
        py = fmin(fmax(py, 0.0), double(m - 1));
        i0 = m < 2 ? 0 : min(int(px), m - 2);
        j0 = m < 2 ? 0 : min(int(py), m - 2);
        i1 = min(i0 + 1, m - 1);
        j1 = min(j0 + 1, m - 1);
    }
    const double tx = px - i0, ty = py - j0;
    const double a = cell(q, m, ch, i0, j0), b = cell(q, m, ch, i1, j0);
    const double c = cell(q, m, ch, i0, j1), d = cell(q, m, ch, i1, j1);
    return (1 - tx) * ((1 - ty) * a + ty * c) + tx * ((1 - ty) * b + ty * d);
}

__device__ __forceinline__ void catmull_rom(double t, double w[4]) {
    const double t2 = t * t, t3 = t2 * t;
    w[0] = 0.5 * (-t3 + 2.0 * t2 - t);
    w[1] = 0.5 * (3.0 * t3 - 5.0 * t2 + 2.0);
    w[2] = 0.5 * (-3.0 * t3 + 4.0 * t2 + t);
    w[3] = 0.5 * (t3 - t2);
}

// Point value at the centre of cell (i, j), i, j in [-2, m + 1].
__device__ __forceinline__ double point(const float *__restrict__ q, int m, int ch, int i, int j,
                                        double c) {
    const double v = cell(q, m, ch, i, j);
    const double lap = (cell(q, m, ch, i + 1, j) + cell(q, m, ch, i - 1, j)) +
                       (cell(q, m, ch, i, j + 1) + cell(q, m, ch, i, j - 1)) - 4.0 * v;
    return v - c * lap;
}

__device__ __forceinline__ double cubic(const float *__restrict__ q, int m, int ch, double px,
                                       double py, double c) {
    px = fmin(fmax(px, -0.5), double(m) - 0.5);
    py = fmin(fmax(py, -0.5), double(m) - 0.5);
    const int i0 = min(int(floor(px)), m - 1), j0 = min(int(floor(py)), m - 1);
    double wx[4], wy[4];
    catmull_rom(px - i0, wx);
    catmull_rom(py - j0, wy);
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double r = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) r += wy[b] * point(q, m, ch, i0 - 1 + a, j0 - 1 + b, c);
        s += wx[a] * r;
    }
    return s;
}

struct Recon {
    int mode;  // 0 clamped bilinear, 1 wall-ghost bilinear, 2 cubic
    double c;  // deconvolution coefficient (cubic)
};

__device__ __forceinline__ double reconstruct(const float *q, int m, int ch, double px, double py,
                                              Recon r) {
    return r.mode == 2 ? cubic(q, m, ch, px, py, r.c) : bilerp(q, m, ch, px, py, r.mode == 1);
}

__device__ __forceinline__ void sample(const float *q, int m, int nx, int ny, int k, int l, double U,
                                       Recon rc, double &rho, double &ux, double &uy) {
    const double px = (k + 0.5) * m / nx - 0.5, py = (l + 0.5) * m / ny - 0.5;
    rho = 1.0 + U * U * reconstruct(q, m, 0, px, py, rc);
    ux = U * reconstruct(q, m, 1, px, py, rc);
    uy = U * reconstruct(q, m, 2, px, py, rc);
}

__global__ __launch_bounds__(kBlock) void k_warm_init(const float *__restrict__ q, int m, int nx, int ny,
                                                      double U, double omega, int neq, Recon rc,
                                                      double *__restrict__ f) {
    const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t P = int64_t(nx) * ny;
    if (j >= P) return;
    const int k = int(j / ny), l = int(j - int64_t(k) * ny);
    double rho, ux, uy;
    sample(q, m, nx, ny, k, l, U, rc, rho, ux, uy);
    double feq[9];
    equilibrium_opt1(rho, ux, uy, feq);
    if (neq) {
        const int km = max(k - 1, 0), kp = min(k + 1, nx - 1), lm = max(l - 1, 0), lp = min(l + 1, ny - 1);
        double r_, uxm, uym, uxp, uyp, uxl, uyl, uxu, uyu;
        sample(q, m, nx, ny, km, l, U, rc, r_, uxm, uym);
        sample(q, m, nx, ny, kp, l, U, rc, r_, uxp, uyp);
        sample(q, m, nx, ny, k, lm, U, rc, r_, uxl, uyl);
        sample(q, m, nx, ny, k, lp, U, rc, r_, uxu, uyu);
        const double hx = double(max(kp - km, 1)), hy = double(max(lp - lm, 1));
        const double dxux = (uxp - uxm) / hx, dxuy = (uyp - uym) / hx;
        const double dyux = (uxu - uxl) / hy, dyuy = (uyu - uyl) / hy;
        const double sxx = dxux, syy = dyuy, sxy = 0.5 * (dyux + dxuy);
        const double fac = (1.0 - omega) * (-3.0 * rho / omega);
        const double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double cx_ = cx(i), cy_ = cy(i);
            const double qs = (cx_ * cx_ - 1.0 / 3) * sxx + 2.0 * cx_ * cy_ * sxy + (cy_ * cy_ - 1.0 / 3) * syy;
            feq[i] += fac * w[i] * qs;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i * P + j] = feq[i];
}

inline unsigned blocks_for(int64_t n, int per = 256) { return unsigned((n + per - 1) / per); }

}  // namespace nn
}  // namespace lbm

using namespace lbm;
using namespace lbm::nn;

extern "C" {

}  // extern "C"

// Tile-load form for an operand X(row, k) = X[row*s_row + k*s_k].
static int load_mode(const void *X, int64_t s_row, int64_t s_k) {
    const bool aligned = reinterpret_cast<uintptr_t>(X) % 16 == 0;
    if (aligned && s_k == 1 && s_row % 8 == 0) return LOAD_K;
    if (aligned && s_row == 1 && s_k % 8 == 0) return LOAD_ROW;
    return LOAD_SCALAR;
}

// Launch the GEMM (and the ordered split-K reduction when splits > 1).
static int run_gemm(const GemmArgs &g, int splits, hipStream_t s) {
    dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, splits);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, g);
    if (splits > 1)
        hipLaunchKernelGGL(k_gemm_reduce, dim3(blocks_for(int64_t(g.M) * g.N)), dim3(256), 0, s, g,
                           splits);
    return launch_status();
}

extern "C" {

int64_t lbm_nn_gemm_workspace(int M, int N, int ksplit) {
    return ksplit > 1 ? int64_t(ksplit) * M * N * int64_t(sizeof(float)) : 0;
}

int lbm_nn_gemm(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn,
                float *C, void *Cb, int64_t ldc, const float *bias, int M, int N, int K,
                int accumulate, int relu, int ksplit, float *work, void *stream) {
    if (!A || !B || (!C && !Cb)) return LBM_E_NULL;
    if (M <= 0 || N <= 0 || K <= 0 || ksplit <= 0) return LBM_E_SIZE;
    if (accumulate && !C) return LBM_E_ARG;
    if (ksplit > 1 && !work) return LBM_E_WORKSPACE;
    int kchunk = (K + ksplit - 1) / ksplit;
    kchunk = (kchunk + TK - 1) / TK * TK;
    const int splits = (K + kchunk - 1) / kchunk;
    GemmArgs g{static_cast<const uint16_t *>(A), sam, sak, static_cast<const uint16_t *>(B), sbk, sbn,
               C, static_cast<uint16_t *>(Cb), ldc, bias, M, N, K, kchunk, accumulate, relu, work,
               load_mode(A, sam, sak), load_mode(B, sbn, sbk), Gather{1, 1, 1, 0},
               Gather{1, 1, 1, 0}};
    return run_gemm(g, splits, as_stream(stream));
}

int lbm_nn_conv3x3(const void *x, int B, int H, int W, int cin, const void *w, int cout,
                   const float *bias, float *y, void *yb, int relu, void *stream) {
    if (!x || !w || (!y && !yb)) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || cin <= 0 || cout <= 0) return LBM_E_SIZE;
    const int64_t P = int64_t(B) * H * W;
    if (P * 9 * cin >= (int64_t(1) << 31) || P * cout >= (int64_t(1) << 31)) return LBM_E_SIZE;
    if (cin % 8 == 0 && reinterpret_cast<uintptr_t>(x) % 16) return LBM_E_ARG;
    const int K = 9 * cin;
    GemmArgs g{static_cast<const uint16_t *>(x), 0, 0, static_cast<const uint16_t *>(w), 1, K, y,
               static_cast<uint16_t *>(yb), cout, bias, int(P), cout, K, (K + TK - 1) / TK * TK, 0,
               relu, nullptr, LOAD_SCALAR, load_mode(w, K, 1), Gather{H, W, cin, 1},
               Gather{1, 1, 1, 0}};
    return run_gemm(g, 1, as_stream(stream));
}

int lbm_nn_conv3x3_wgrad(const void *x, const void *dy, int B, int H, int W, int cin, int cout,
                         float *dw, int ksplit, float *work, void *stream) {
    if (!x || !dy || !dw) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || cin <= 0 || cout <= 0 || ksplit <= 0) return LBM_E_SIZE;
    const int64_t P = int64_t(B) * H * W;
    if (P * 9 * cin >= (int64_t(1) << 31) || P * cout >= (int64_t(1) << 31)) return LBM_E_SIZE;
    if (cin % 8 == 0 && reinterpret_cast<uintptr_t>(x) % 16) return LBM_E_ARG;
    if (ksplit > 1 && !work) return LBM_E_WORKSPACE;
    const int K = int(P), N = 9 * cin;
    int kchunk = (K + ksplit - 1) / ksplit;
    kchunk = (kchunk + TK - 1) / TK * TK;
    const int splits = (K + kchunk - 1) / kchunk;
    GemmArgs g{static_cast<const uint16_t *>(dy), 1, cout, static_cast<const uint16_t *>(x), 0, 0,
               dw, nullptr, N, nullptr, cout, N, K, kchunk, 0, 0, work, load_mode(dy, 1, cout),
               LOAD_SCALAR, Gather{1, 1, 1, 0}, Gather{H, W, cin, 1}};
    return run_gemm(g, splits, as_stream(stream));
}

int lbm_nn_conv3x3_flip_weights(const void *w, int cin, int cout, void *wf, void *stream) {
    if (!w || !wf) return LBM_E_NULL;
    if (cin <= 0 || cout <= 0) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_flip_weights, dim3(blocks_for(int64_t(9) * cout * cin)), dim3(256), 0,
                       as_stream(stream), static_cast<const uint16_t *>(w), cin, cout,
                       static_cast<uint16_t *>(wf));
    return launch_status();
}

int lbm_nn_conv3x3_dgrad(const void *dy, const void *wf, int B, int H, int W, int cin, int cout,
                         float *dx, int accumulate, void *stream) {
    if (!dy || !wf || !dx) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || cin <= 0 || cout <= 0) return LBM_E_SIZE;
    const int64_t P = int64_t(B) * H * W;
    if (P * 9 * cout >= (int64_t(1) << 31) || P * cin >= (int64_t(1) << 31)) return LBM_E_SIZE;
    if (cout % 8 == 0 && reinterpret_cast<uintptr_t>(dy) % 16) return LBM_E_ARG;
    const int K = 9 * cout;
    GemmArgs g{static_cast<const uint16_t *>(dy), 0, 0, static_cast<const uint16_t *>(wf), cin, 1,
               dx, nullptr, cin, nullptr, int(P), cin, K, (K + TK - 1) / TK * TK, accumulate, 0,
               nullptr, LOAD_SCALAR, load_mode(wf, 1, cin), Gather{H, W, cout, -1},
               Gather{1, 1, 1, 0}};
    return run_gemm(g, 1, as_stream(stream));
}

int lbm_nn_im2col3x3(const void *x, int B, int H, int W, int C, void *cols, void *stream) {
    if (!x || !cols) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0) return LBM_E_SIZE;
    const int64_t n = int64_t(B) * H * W * C;
    if (n >= (int64_t(1) << 31)) return LBM_E_SIZE;
    const bool vec = (C % 8 == 0) && (reinterpret_cast<uintptr_t>(x) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(cols) % 16 == 0);
    const int items = int(vec ? n / 8 : n);
    dim3 grid(blocks_for(items), 9);
    hipStream_t s = as_stream(stream);
    if (vec)
        hipLaunchKernelGGL(k_im2col<8>, grid, dim3(256), 0, s, static_cast<const uint16_t *>(x), H, W, C,
                           static_cast<uint16_t *>(cols), items);
    else
        hipLaunchKernelGGL(k_im2col<1>, grid, dim3(256), 0, s, static_cast<const uint16_t *>(x), H, W, C,
                           static_cast<uint16_t *>(cols), items);
    return launch_status();
}

int lbm_nn_col2im3x3(const float *dcols, int B, int H, int W, int C, float *dx, int accumulate,
                     void *stream) {
    if (!dcols || !dx) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_col2im, dim3(blocks_for(int64_t(B) * H * W * C)), dim3(256), 0,
                       as_stream(stream), dcols, B, H, W, C, dx, accumulate);
    return launch_status();
}

int lbm_nn_relu_backward(const float *y, float *dy, void *dy_bf16, int64_t n, void *stream) {
    if (!y || !dy) return LBM_E_NULL;
    if (n <= 0) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_relu_backward, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), y, dy,
                       static_cast<uint16_t *>(dy_bf16), n);
    return launch_status();
}

int64_t lbm_nn_bias_grad_workspace(int N) { return int64_t(kBiasBlocks) * N * int64_t(sizeof(double)); }

int lbm_nn_bias_grad(const float *dy, int64_t M, int N, float *db, double *work, void *stream) {
    if (!dy || !db || !work) return LBM_E_NULL;
    if (M <= 0 || N <= 0) return LBM_E_SIZE;
    hipStream_t s = as_stream(stream);
    const int64_t rows = (M + kBiasBlocks - 1) / kBiasBlocks;
    const int nb = int((M + rows - 1) / rows);
    for (int c0 = 0; c0 < N; c0 += 256)
        hipLaunchKernelGGL(k_bias_grad_partial, dim3(nb), dim3(256), 0, s, dy, M, N, c0, rows, work);
    hipLaunchKernelGGL(k_bias_grad_final, dim3(N), dim3(256), 0, s, work, N, nb, db);
    return launch_status();
}

int lbm_nn_avgpool2(const float *x, int B, int H, int W, int C, float *y, void *yb, void *stream) {
    if (!x || !y) return LBM_E_NULL;
    if (B <= 0 || H < 2 || W < 2 || C <= 0 || (H & 1) || (W & 1)) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_avgpool2, dim3(blocks_for(int64_t(B) * (H / 2) * (W / 2) * C)), dim3(256), 0,
                       as_stream(stream), x, B, H, W, C, y, static_cast<uint16_t *>(yb));
    return launch_status();
}

int lbm_nn_avgpool2_backward(const float *dy, int B, int H, int W, int C, float *dx, int accumulate,
                             void *stream) {
    if (!dy || !dx) return LBM_E_NULL;
    if (B <= 0 || H < 2 || W < 2 || C <= 0 || (H & 1) || (W & 1)) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_avgpool2_backward, dim3(blocks_for(int64_t(B) * H * W * C)), dim3(256), 0,
                       as_stream(stream), dy, B, H, W, C, dx, accumulate);
    return launch_status();
}

int lbm_nn_upsample2(const float *x, int B, int H, int W, int C, float *y, void *yb, int Ct, int coff,
                     void *stream) {
    if (!x || !y) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || coff < 0 || coff + C > Ct) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_upsample2, dim3(blocks_for(int64_t(B) * 4 * H * W * C)), dim3(256), 0,
                       as_stream(stream), x, B, H, W, C, y, static_cast<uint16_t *>(yb), Ct, coff);
    return launch_status();
}

int lbm_nn_upsample2_backward(const float *dy, int B, int H, int W, int C, int Ct, int coff,
                              float *dx, void *stream) {
    if (!dy || !dx) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || coff < 0 || coff + C > Ct) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_upsample2_backward, dim3(blocks_for(int64_t(B) * H * W * C)), dim3(256), 0,
                       as_stream(stream), dy, B, H, W, C, Ct, coff, dx);
    return launch_status();
}

int lbm_nn_copy_channels(const float *src, int Cs, int soff, float *dst, void *dst_bf16, int Cd,
                         int doff, int C, int64_t P, int accumulate, void *stream) {
    if (!src || !dst) return LBM_E_NULL;
    if (C <= 0 || P <= 0 || soff < 0 || doff < 0 || soff + C > Cs || doff + C > Cd) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_copy_channels, dim3(blocks_for(P * C)), dim3(256), 0, as_stream(stream), src,
                       Cs, soff, dst, static_cast<uint16_t *>(dst_bf16), Cd, doff, C, P, accumulate);
    return launch_status();
}

int lbm_nn_mse(const float *y, const float *t, int64_t n, float *dy, double *work, double *loss,
               void *stream) {
    if (!y || !t || !dy || !work || !loss) return LBM_E_NULL;
    if (n <= 0) return LBM_E_SIZE;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_mse, dim3(kReduceBlocks), dim3(kBlock), 0, s, y, t, n, dy, work);
    launch_sum_pass2(work, kReduceBlocks, 0.0, loss, s);
    return launch_status();
}

int lbm_nn_mse_pooled(const float *y, const float *t, int B, int H, int W, int C, int f,
                      float weight, float *dy, double *work, double *loss, void *stream) {
    if (!y || !t || !dy || !work || !loss) return LBM_E_NULL;
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || f <= 0 || H % f || W % f) return LBM_E_SIZE;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_mse_pooled, dim3(kReduceBlocks), dim3(kBlock), 0, s, y, t, B, H, W, C, f,
                       weight, dy, work);
    launch_sum_pass2(work, kReduceBlocks, 0.0, loss, s);
    return launch_status();
}

int lbm_nn_adam(float *p, const float *g, float *m, float *v, void *p_bf16, int64_t n, float lr,
                float b1, float b2, float eps, int step, void *stream) {
    if (!p || !g || !m || !v) return LBM_E_NULL;
    if (n <= 0 || step <= 0) return LBM_E_SIZE;
    const float c1 = 1.0f - powf(b1, float(step)), c2 = 1.0f - powf(b2, float(step));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), p, g, m, v,
                       static_cast<uint16_t *>(p_bf16), n, lr, b1, b2, eps, c1, c2);
    return launch_status();
}

int lbm_nn_ema(float *ema, const float *p, int64_t n, float decay, void *stream) {
    if (!ema || !p) return LBM_E_NULL;
    if (n <= 0) return LBM_E_SIZE;
    if (!(decay >= 0.0f && decay < 1.0f)) return LBM_E_ARG;
    hipLaunchKernelGGL(k_ema, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), ema, p, n, decay);
    return launch_status();
}

int lbm_nn_to_bf16(const float *x, void *y, int64_t n, void *stream) {
    if (!x || !y) return LBM_E_NULL;
    if (n <= 0) return LBM_E_SIZE;
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), x,
                       static_cast<uint16_t *>(y), n);
    return launch_status();
}

int lbm_warm_init_f64(const float *q, int m, int64_t nx, int64_t ny, double U, double omega,
                      int flags, int pool, double *f, void *stream) {
    if (!q || !f) return LBM_E_NULL;
    if (m <= 0 || !valid_grid(nx, ny)) return LBM_E_SIZE;
    if (!(U > 0.0) || !(omega > 0.0 && omega < 2.0) ||
        (flags & ~(kWarmNeq | kWarmWalls | kWarmCubic)) || pool < 1)
        return LBM_E_ARG;
    if ((flags & kWarmCubic) && m < 3) return LBM_E_SIZE;
    Recon rc{(flags & kWarmCubic) ? 2 : ((flags & kWarmWalls) ? 1 : 0),
             (double(pool) * pool - 1.0) / (24.0 * double(pool) * pool)};
    const int64_t P = nx * ny;
    hipLaunchKernelGGL(k_warm_init, dim3(blocks_for(P, kBlock)), dim3(kBlock), 0, as_stream(stream), q,
                       m, int(nx), int(ny), U, omega, int((flags & kWarmNeq) != 0), rc, f);
    return launch_status();
}

}  // extern "C"
