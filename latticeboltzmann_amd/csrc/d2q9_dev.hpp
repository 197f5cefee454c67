// Per-node arithmetic of the reduced-precision (scaled deviation) storage:
// storage codecs, stochastic-rounding dither, cavity walls and BGK collide on
// deviations, and the conservation-compensated rounding of a node's nine
// values.  Shared by the single-lattice and the batched step kernels so all of
// them compute the same bits.  Units including it are built with
// -ffp-contract=off and the fused multiply-adds of the collide are written
// out (fma_ct): which products the compiler would fuse depends on the code
// around the collide (values in flight from memory, held in registers since an
// earlier step, or exchanged between lanes), so kernels that must agree bit for
// bit would otherwise round differently.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_d2q9.h"
#include "d2q9_grid.hpp"

namespace lbm {

__host__ __device__ constexpr float wf(int i) {
    return i == 0 ? 4.0f / 9.0f : (i <= 4 ? 1.0f / 9.0f : 1.0f / 36.0f);
}
__host__ __device__ constexpr double wd(int i) {
    return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0);
}

// Storage codecs: raw storage type and conversions.  enc rounds an FP32 value
// to the format, enc64 rounds an FP64 value to the format directly (one
// rounding), both to nearest with ties to even.
template <int FMT>
struct Codec;
template <>
struct Codec<LBM_FMT_F32> {
    using raw = float;
    __device__ static float dec(raw x) { return x; }
    __device__ static raw enc(float x) { return x; }
    __device__ static raw enc64(double x) { return float(x); }
};
template <>
struct Codec<LBM_FMT_BF16> {
    using raw = uint16_t;
    __device__ static float dec(raw x) { return float(__builtin_bit_cast(__bf16, x)); }
    __device__ static raw enc(float x) { return __builtin_bit_cast(uint16_t, __bf16(x)); }
    __device__ static raw enc64(double x) { return __builtin_bit_cast(uint16_t, __bf16(x)); }
};
template <>
struct Codec<LBM_FMT_F16> {
    using raw = uint16_t;
    __device__ static float dec(raw x) { return float(__builtin_bit_cast(_Float16, x)); }
    __device__ static raw enc(float x) { return __builtin_bit_cast(uint16_t, _Float16(x)); }
    __device__ static raw enc64(double x) { return __builtin_bit_cast(uint16_t, _Float16(x)); }
};

// Stochastic rounding: x is rounded up in magnitude with probability equal
// to its distance from the value below, so the rounding error has zero mean.
// Deterministic rounding of a slowly varying flow rounds every node the same
// way every step, which biases conserved quantities (a shear wave then decays
// at the wrong rate); stochastic rounding turns that bias into noise.
// Implemented on the FP32 bit pattern: add 16 (BF16) or 13 (FP16) random
// bits below the kept significand and truncate.  FP16 values outside the
// FP16 normal range fall back to nearest rounding.
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Per-node random stream: one hash of (seed, node) seeds a xorshift32
// generator (Marsaglia, 2003; shifts 13, 17, 5), whose states are used 16 bits
// at a time.  Shifts and xors run at full rate, unlike the 32-bit multiply of a
// linear congruential step (quarter rate), and the nine to eighteen dithers of
// a node are the store path's largest cost.  The state is forced odd (a
// xorshift state of 0 would stay 0).  Within one step (one seed) distinct
// nodes get distinct streams.
struct Dither {
    uint32_t s, spare;
    bool have;
    __device__ __forceinline__ Dither(uint32_t seed, uint64_t node)
        : s(hash32(seed ^ uint32_t(node) ^ (uint32_t(node >> 32) * 0x9e3779b9U)) | 1U), spare(0),
          have(false) {}
    __device__ __forceinline__ uint32_t next16() {
        if (have) {
            have = false;
            return spare;
        }
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        spare = s >> 16;
        have = true;
        return s & 0xffffU;
    }
};

template <int FMT>
__device__ __forceinline__ typename Codec<FMT>::raw enc_sr(float x, uint32_t r16);
template <>
__device__ __forceinline__ float enc_sr<LBM_FMT_F32>(float x, uint32_t) {
    return x;
}
template <>
__device__ __forceinline__ uint16_t enc_sr<LBM_FMT_BF16>(float x, uint32_t r16) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    if ((b & 0x7f800000U) == 0x7f800000U) return Codec<LBM_FMT_BF16>::enc(x);  // inf / nan
    return uint16_t((b + r16) >> 16);
}
template <>
__device__ __forceinline__ uint16_t enc_sr<LBM_FMT_F16>(float x, uint32_t r16) {
    const float a = fabsf(x);
    if (!(a >= 6.103515625e-05f && a < 65504.0f)) return Codec<LBM_FMT_F16>::enc(x);
    uint32_t b = __builtin_bit_cast(uint32_t, x);
    b = (b + (r16 & 0x1fffU)) & ~0x1fffU;  // now exactly representable in FP16
    return Codec<LBM_FMT_F16>::enc(__builtin_bit_cast(float, b));
}

// Stochastic rounding of an FP64 value to FP32: 29 random bits below the
// kept significand, then truncation (values outside the FP32 normal range
// round to nearest).
__device__ __forceinline__ float sr_f64_to_f32(double x, uint32_t r29) {
    const double a = fabs(x);
    if (!(a >= 1.1754943508222875e-38 && a < 3.0e38)) return float(x);
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    b = (b + uint64_t(r29 & 0x1fffffffU)) & ~uint64_t(0x1fffffff);  // now exact in FP32
    return float(__builtin_bit_cast(double, b));
}

// Stochastic rounding of an FP64 value straight to a 16-bit format (an FP64
// collide must not pass through FP32, whose nearest rounding would repeat the
// same error every step): 32 random bits are added just below the kept
// significand (BF16 keeps 7 fraction bits, FP16 10, of FP64's 52), then the
// dropped bits are cleared, which leaves a value the format holds exactly.
// Values outside the format's normal range round to nearest.
template <int FMT>
__device__ __forceinline__ uint16_t sr_f64_to_16(double x, uint32_t r32) {
    constexpr int drop = FMT == LBM_FMT_BF16 ? 45 : 42;
    const double a = fabs(x);
    const bool normal = FMT == LBM_FMT_BF16 ? (a >= 1.1754943508222875e-38 && a < 3.0e38)
                                            : (a >= 6.103515625e-05 && a < 65504.0);
    if (!normal) return Codec<FMT>::enc64(x);
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    b = (b + (uint64_t(r32) << (drop - 32))) & ~((uint64_t(1) << drop) - 1);
    return Codec<FMT>::enc64(__builtin_bit_cast(double, b));
}

// Weight of channel i in the compute type.
template <typename CT>
__host__ __device__ constexpr CT wt(int i) {
    return i == 0 ? CT(4) / CT(9) : (i <= 4 ? CT(1) / CT(9) : CT(1) / CT(36));
}

// Cavity walls on deviations (same rules and order as the FP64 kernel).
template <int FMT, typename CT>
__device__ __forceinline__ void cavity_walls_dev(CT g[9], const typename Codec<FMT>::raw *c,
                                                 int64_t cs, int k, int l, int nx, int ny,
                                                 CT u0, CT inv_scale) {
    using C = Codec<FMT>;
    auto dv = [&](int i) { return CT(C::dec(c[i * cs])) * inv_scale; };
    if (l == 0) {
        g[2] = dv(4);
        g[5] = dv(7);
        g[6] = dv(8);
    }
    if (l == ny - 1) {
        const CT c2 = dv(2), c5 = dv(5), c6 = dv(6);
        const CT rl = CT(1) + (((c6 + c2) + (c5 + g[6])) + ((g[2] + g[5]) + (g[3] + (g[0] + g[1]))));
        const CT t = (CT(6) / CT(36)) * rl * u0;
        g[4] = c2;
        g[8] = c6 + t;
        g[7] = c5 - t;
    }
    if (k == 0) {
        g[1] = dv(3);
        g[5] = dv(7);
        g[8] = dv(6);
    }
    if (k == nx - 1) {
        g[3] = dv(1);
        g[6] = dv(8);
        g[7] = dv(5);
    }
}

// The same walls on values already in the compute type (c[i * cs] is the
// node's own pre-stream value of channel i): the multi-step kernel's phases
// that keep unrounded values between steps.
template <typename CT>
__device__ __forceinline__ void cavity_walls_ct(CT g[9], const CT *c, int64_t cs, int k, int l,
                                                int nx, int ny, CT u0) {
    if (l == 0) {
        g[2] = c[4 * cs];
        g[5] = c[7 * cs];
        g[6] = c[8 * cs];
    }
    if (l == ny - 1) {
        const CT c2 = c[2 * cs], c5 = c[5 * cs], c6 = c[6 * cs];
        const CT rl = CT(1) + (((c6 + c2) + (c5 + g[6])) + ((g[2] + g[5]) + (g[3] + (g[0] + g[1]))));
        const CT t = (CT(6) / CT(36)) * rl * u0;
        g[4] = c2;
        g[8] = c6 + t;
        g[7] = c5 - t;
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

// a b + c with one rounding
__device__ __forceinline__ float fma_ct(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_ct(double a, double b, double c) { return __builtin_fma(a, b, c); }

// 1 / x: the hardware reciprocal in FP32 (within 1 ulp), the IEEE quotient in FP64.
__device__ __forceinline__ float recip_ct(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip_ct(double x) { return 1.0 / x; }

// x with the lowest significand bit cleared.
__device__ __forceinline__ float clear_lsb(float x) {
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) & ~1u);
}
__device__ __forceinline__ double clear_lsb(double x) {
    return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) & ~uint64_t(1));
}

// BGK collide on deviations in the compute type CT (float or double):
// d_i <- (1 - omega) d_i + omega deq_i with
//   omega deq_i = omega w_i (a + 4.5 rho cu_i^2) + 3 omega w_i (c_i . j),
//   a = drho - 1.5 rho u^2, cu_i = c_i . u.
// Opposite channels share the even part and take the odd part with opposite
// signs (about 60 operations per node).  The odd part uses the momentum j
// itself, with coefficients b1 (axes) and b5 (diagonals) built so that
// b1 + 2 b5 = omega / 2 holds exactly: with 3 w_i rounded independently
// (fl(1/3), fl(1/12)) each FP32 collide would scale the momentum by the same
// 1 + O(1e-8), a drift that biases the decay of a slow wave by several percent.
// u = j / rho is evaluated as j - j (drho / rho), so the reciprocal's error
// only touches the small correction.
template <typename CT>
__device__ __forceinline__ void collide_dev(CT d[9], CT omega, CT &ux, CT &uy) {
    const CT d78 = d[7] + d[8];
    const CT drho = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d78));
    const CT rho = CT(1) + drho;
    const CT jx = ((d[1] - d[3]) + (d[5] - d[6])) + (d[8] - d[7]);
    const CT jy = ((d[2] - d[4]) + (d[5] + d[6])) - d78;
    const CT q = drho * recip_ct(rho);
    ux = fma_ct(-jx, q, jx);
    uy = fma_ct(-jy, q, jy);
    const CT a = fma_ct(CT(-1.5) * rho, fma_ct(ux, ux, uy * uy), drho);
    const CT keep = CT(1) - omega;
    const CT r45 = CT(4.5) * rho;
    const CT ow0 = omega * wt<CT>(0), ow1 = omega * wt<CT>(1), ow5 = omega * wt<CT>(5);
    const CT e1 = ow1 * r45, a1 = ow1 * a;  // even part, axis channels
    const CT e5 = ow5 * r45, a5 = ow5 * a;  // even part, diagonals
    // b5 ~ omega / 12 on a significand of one bit less, so that omega / 2 - 2 b5 is exact
    const CT b5 = clear_lsb(omega * (CT(1) / CT(12)));
    const CT b1 = fma_ct(CT(-2), b5, CT(0.5) * omega);
    d[0] = fma_ct(keep, d[0], ow0 * a);
    auto pair = [&](int i, int j, CT cu, CT cj, CT e, CT a_, CT b) {
        const CT even = fma_ct(e, cu * cu, a_);
        d[i] = fma_ct(b, cj, fma_ct(keep, d[i], even));
        d[j] = fma_ct(-b, cj, fma_ct(keep, d[j], even));
    };
    pair(1, 3, ux, jx, e1, a1, b1);
    pair(2, 4, uy, jy, e1, a1, b1);
    pair(5, 7, ux + uy, jx + jy, e5, a5, b5);
    pair(8, 6, ux - uy, jx - jy, e5, a5, b5);
}

// Round one value of the compute type to the storage format: nearest, or
// stochastic with the dither's next random bits.
template <int FMT, bool SR, typename CT>
__device__ __forceinline__ typename Codec<FMT>::raw round_one(CT x, Dither &dither) {
    using C = Codec<FMT>;
    if constexpr (FMT == LBM_FMT_F32 && sizeof(CT) == 8) {
        if constexpr (SR) {
            const uint32_t hi = dither.next16(), lo = dither.next16();
            return sr_f64_to_f32(double(x), ((hi << 16) | lo) >> 3);
        } else {
            return C::enc64(double(x));
        }
    } else if constexpr (sizeof(CT) == 8) {
        // 16-bit formats from an FP64 collide: one rounding of the FP64 value
        if constexpr (SR) {
            const uint32_t hi = dither.next16(), lo = dither.next16();
            return sr_f64_to_16<FMT>(double(x), (hi << 16) | lo);
        } else {
            return C::enc64(double(x));
        }
    } else {
        if constexpr (SR) return enc_sr<FMT>(x, dither.next16());
        else return C::enc(x);
    }
}

// Round the nine scaled post-collision deviations x[] to the storage format so
// that the stored mass and momentum stay as close as possible to the exact
// ones.  Rounding each value independently leaves the conserved moments with
// the sum of nine rounding errors, which collisions never relax, so they
// random-walk from step to step.  Instead the diagonals (5-8) and E, N are
// rounded first; W and S are rounded at targets shifted by the x- and
// y-momentum residuals so far; the rest population, which does not stream,
// absorbs the mass residual.  Each conserved moment then carries the
// rounding error of a single value.  Residuals are exact in the compute type
// (x and the rounded value differ by less than one storage ulp).
template <int FMT, bool SR, typename CT>
__device__ __forceinline__ void round_conserving(const CT x[9], typename Codec<FMT>::raw r[9],
                                                 Dither &dither) {
    using C = Codec<FMT>;
    auto rnd = [&](int i, CT target) {
        r[i] = round_one<FMT, SR, CT>(target, dither);
        return target - CT(C::dec(r[i]));
    };
    const CT e5 = rnd(5, x[5]), e6 = rnd(6, x[6]), e7 = rnd(7, x[7]), e8 = rnd(8, x[8]);
    const CT e1 = rnd(1, x[1]), e2 = rnd(2, x[2]);
    const CT ex = ((e5 - e6) + (e8 - e7)) + e1;  // x-momentum stored short by ex
    const CT ey = ((e5 + e6) - (e7 + e8)) + e2;  // y-momentum stored short by ey
    const CT e3 = rnd(3, x[3] - ex);
    const CT e4 = rnd(4, x[4] - ey);
    // mass residual of channels 1..8: e_i for plain roundings, e3 + ex and e4 + ey
    // for the shifted ones
    const CT em = ((e5 + e6) + (e7 + e8)) + ((e1 + e2) + ((e3 + ex) + (e4 + ey)));
    rnd(0, x[0] + em);
}

// Store path shared by the step kernels: scale the post-collision deviations
// and round them to the storage format.  FP32 storage of an FP32 collide is
// exact (no rounding); every other combination uses the compensated rounding.
template <int FMT, typename CT>
__device__ __forceinline__ void store_values(const CT g[9], CT scale, int stochastic,
                                             Dither &dither, typename Codec<FMT>::raw out[9]) {
    if constexpr (FMT == LBM_FMT_F32 && sizeof(CT) == 4) {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = Codec<FMT>::enc(g[i] * scale);
    } else {
        CT x[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) x[i] = g[i] * scale;
        if (stochastic) round_conserving<FMT, true, CT>(x, out, dither);
        else round_conserving<FMT, false, CT>(x, out, dither);
    }
}

// ---------------------------------------------------------------------------
// K steps per launch on stored deviations by a sweep along x with the
// intermediate steps in registers (the single-lattice k_dev_sweep and the
// batched k_batch_sweep): a wave owns 64 columns and a segment of rows of one
// lattice; per iteration it computes step 1 of a new row from src, then step j
// of the row behind step j - 1's newest, j = 2..K, and writes step K's row;
// column neighbours come from the adjacent lanes through ds_bpermute.
// EACH: every step is rounded to the storage format with its own step's dither
// (seeds[j - 1], node node_base + k ny + l) and the next step reads the decoded
// stored value, so the launch is bitwise K single-step launches; otherwise
// (store every K) the intermediate steps stay unrounded in the compute type and
// only step K is rounded, with seeds[K - 1].
__device__ __forceinline__ float dev_lane_fetch(float v, int addr) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double dev_lane_fetch(double v, int addr) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_ds_bpermute(addr, int(unsigned(u))));
    const unsigned hi = unsigned(__builtin_amdgcn_ds_bpermute(addr, int(unsigned(u >> 32))));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// cavity_walls_ct with the bottom and lid rules behind the wave-uniform test
// yedge (whether the wave's columns, halo included, reach l = 0 or ny - 1).
template <typename CT>
__device__ __forceinline__ void dev_sweep_walls(CT g[9], const CT c[9], int k, int l, int nx,
                                                int ny, CT u0, bool yedge) {
    if (yedge) {
        if (l == 0) {
            g[2] = c[4];
            g[5] = c[7];
            g[6] = c[8];
        }
        if (l == ny - 1) {
            const CT c2 = c[2], c5 = c[5], c6 = c[6];
            const CT rl = CT(1) + (((c6 + c2) + (c5 + g[6])) + ((g[2] + g[5]) + (g[3] + (g[0] + g[1]))));
            const CT t = (CT(6) / CT(36)) * rl * u0;
            g[4] = c2;
            g[8] = c6 + t;
            g[7] = c5 - t;
        }
    }
    if (k == 0) {
        g[1] = c[3];
        g[5] = c[7];
        g[8] = c[6];
    }
    if (k == nx - 1) {
        g[3] = c[1];
        g[6] = c[8];
        g[7] = c[5];
    }
}

struct DevSweepLane {
    int l, from_lo, from_hi;
    bool writer, yedge;
};

template <int FMT, typename CT, int BC, bool MOM, int K, bool EACH, int PH>
__device__ __forceinline__ void dev_sweep_iteration(
    const typename Codec<FMT>::raw *__restrict__ src, typename Codec<FMT>::raw *__restrict__ dst,
    int i, int ks, int nx, int ny, Layout L, const DevSweepLane &ln, CT omega, CT u0, CT scale,
    CT inv_scale, int stochastic, const uint32_t (&seeds)[K], uint64_t node_base,
    CT (&A)[K - 1][9], CT (&B)[K - 1][9], float *__restrict__ ux_o, float *__restrict__ uy_o) {
    using C = Codec<FMT>;
    using R = typename C::raw;
    CT g[9];
    {  // step 1 of row i + K - 1, from src
        const int k = wrap1(i + K - 1, nx);
        const PullRows<R> rows(src, L, k, nx);
        R raw[9];
        pull9<false>(rows, ln.l, ny, raw);
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = CT(C::dec(raw[q])) * inv_scale;
        if constexpr (BC == LBM_BC_CAVITY) {
            if (ln.yedge || k == 0 || k == nx - 1)
                cavity_walls_dev<FMT, CT>(g, src + int64_t(k) * L.rs + ln.l, L.cs, k, ln.l, nx, ny,
                                          u0, inv_scale);
        }
        CT ux, uy;
        collide_dev<CT>(g, omega, ux, uy);
        if constexpr (EACH) {
            Dither dither(seeds[0], node_base + uint64_t(k) * uint64_t(ny) + uint64_t(ln.l));
            R q[9];
            store_values<FMT, CT>(g, scale, stochastic, dither, q);
#pragma unroll
            for (int c = 0; c < 9; ++c) g[c] = CT(C::dec(q[c])) * inv_scale;
        }
    }
#pragma unroll
    for (int j = 2; j <= K; ++j) {
        CT(&older)[9] = PH ? B[j - 2] : A[j - 2];
        CT(&newer)[9] = PH ? A[j - 2] : B[j - 2];
        const bool active = i >= ks - 2 * (K - j);  // wave-uniform
        CT h[9];
        int k = 0;
        if (active) {
            k = wrap1(i + K - j, nx);
            h[0] = newer[0];
            h[1] = older[1];
            h[2] = dev_lane_fetch(newer[2], ln.from_lo);
            h[3] = g[3];
            h[4] = dev_lane_fetch(newer[4], ln.from_hi);
            h[5] = dev_lane_fetch(older[5], ln.from_lo);
            h[6] = dev_lane_fetch(g[6], ln.from_lo);
            h[7] = dev_lane_fetch(g[7], ln.from_hi);
            h[8] = dev_lane_fetch(older[8], ln.from_hi);
            if constexpr (BC == LBM_BC_CAVITY) dev_sweep_walls<CT>(h, newer, k, ln.l, nx, ny, u0, ln.yedge);
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) older[q] = g[q];
        if (!active) return;
        CT ux, uy;
        collide_dev<CT>(h, omega, ux, uy);
        const uint32_t sj = seeds[j - 1];
        if (j == K) {
            if (ln.writer) {
                Dither dither(sj, node_base + uint64_t(k) * uint64_t(ny) + uint64_t(ln.l));
                R q[9];
                store_values<FMT, CT>(h, scale, stochastic, dither, q);
                R *d = dst + int64_t(k) * L.rs + ln.l;
#pragma unroll
                for (int c = 0; c < 9; ++c) d[c * L.cs] = q[c];
                if constexpr (MOM) {
                    const int64_t m = int64_t(k) * ny + ln.l;
                    ux_o[m] = float(ux);
                    uy_o[m] = float(uy);
                }
            }
        } else if constexpr (EACH) {
            Dither dither(sj, node_base + uint64_t(k) * uint64_t(ny) + uint64_t(ln.l));
            R q[9];
            store_values<FMT, CT>(h, scale, stochastic, dither, q);
#pragma unroll
            for (int c = 0; c < 9; ++c) g[c] = CT(C::dec(q[c])) * inv_scale;
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) g[c] = h[c];
        }
    }
}

// The rows [ks, ke) of one strip (described by ln) of one lattice.
template <int FMT, typename CT, int BC, bool MOM, int K, bool EACH>
__device__ __forceinline__ void dev_sweep_run(const typename Codec<FMT>::raw *__restrict__ src,
                                              typename Codec<FMT>::raw *__restrict__ dst, int ks,
                                              int ke, int nx, int ny, Layout L,
                                              const DevSweepLane &ln, CT omega, CT u0, CT scale,
                                              CT inv_scale, int stochastic,
                                              const uint32_t (&seeds)[K], uint64_t node_base,
                                              float *__restrict__ ux_o, float *__restrict__ uy_o) {
    CT A[K - 1][9], B[K - 1][9];
#pragma unroll
    for (int j = 0; j < K - 1; ++j)
#pragma unroll
        for (int q = 0; q < 9; ++q) A[j][q] = B[j][q] = CT(0);
    int i = ks - 2 * (K - 1);
    for (; i + 1 < ke; i += 2) {
        dev_sweep_iteration<FMT, CT, BC, MOM, K, EACH, 0>(src, dst, i, ks, nx, ny, L, ln, omega, u0,
                                                          scale, inv_scale, stochastic, seeds,
                                                          node_base, A, B, ux_o, uy_o);
        dev_sweep_iteration<FMT, CT, BC, MOM, K, EACH, 1>(src, dst, i + 1, ks, nx, ny, L, ln, omega,
                                                          u0, scale, inv_scale, stochastic, seeds,
                                                          node_base, A, B, ux_o, uy_o);
    }
    if (i < ke)
        dev_sweep_iteration<FMT, CT, BC, MOM, K, EACH, 0>(src, dst, i, ks, nx, ny, L, ln, omega, u0,
                                                          scale, inv_scale, stochastic, seeds,
                                                          node_base, A, B, ux_o, uy_o);
}

// Lane set-up of a wave on strip `strip` of an ny-column lattice (VW =
// 64 - 2(K-1) written columns per strip).  A lattice of exactly 64 columns is
// one strip with no halo lanes: its column neighbours wrap within the wave
// (lane 63's upper neighbour is lane 0, column 0), which is the periodic pull
// the walls then override, so all 64 lanes are written.
template <int K>
__device__ __forceinline__ DevSweepLane dev_sweep_lane(int strip, int ny, int lane) {
    constexpr int VW = 64 - 2 * (K - 1);
    DevSweepLane ln;
    ln.from_lo = ((lane + 63) & 63) << 2;
    ln.from_hi = ((lane + 1) & 63) << 2;
    if (ny == 64) {
        ln.l = lane;
        ln.writer = true;
        ln.yedge = true;
        return ln;
    }
    const int c = strip * VW - (K - 1) + lane;  // in [-(K-1), ny + 63): one wrap suffices
    ln.l = wrap1(c, ny);
    ln.writer = lane >= K - 1 && lane < 64 - (K - 1) && c < ny;
    const int c0 = strip * VW - (K - 1);  // the wave's columns before wrapping: c0 .. c0 + 63
    ln.yedge = c0 <= 0 || c0 + 63 >= ny - 1;
    return ln;
}

// Strips of an ny-column lattice.
template <int K>
__host__ __device__ constexpr int dev_sweep_strips(int ny) {
    return ny == 64 ? 1 : (ny + (64 - 2 * (K - 1)) - 1) / (64 - 2 * (K - 1));
}

}  // namespace lbm
