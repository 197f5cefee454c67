// Per-node arithmetic of the reduced-precision (scaled deviation) storage:
// storage codecs, stochastic-rounding dither, cavity walls and BGK collide on
// deviations, and the conservation-compensated rounding of a node's nine
// values.  Shared by the single-lattice and the batched step kernels so both
// compute the same bits; units including it are built with -ffp-contract=fast.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_d2q9.h"

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

// BGK collide on deviations in the compute type CT (float or double).
template <typename CT>
__device__ __forceinline__ void collide_dev(CT d[9], CT omega, CT &ux, CT &uy) {
    const CT drho = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + (d[7] + d[8])));
    const CT rho = CT(1) + drho;
    const CT jx = ((d[1] - d[3]) + (d[5] - d[6])) + (d[8] - d[7]);
    const CT jy = ((d[2] - d[4]) + (d[5] + d[6])) - (d[7] + d[8]);
    const CT inv = CT(1) / rho;
    ux = jx * inv;
    uy = jy * inv;
    const CT uu15 = CT(1.5) * (ux * ux + uy * uy);
    CT cu[9];
    cu[0] = CT(0);
    cu[1] = ux;
    cu[2] = uy;
    cu[3] = -ux;
    cu[4] = -uy;
    cu[5] = ux + uy;
    cu[6] = uy - ux;
    cu[7] = -(ux + uy);
    cu[8] = ux - uy;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const CT deq = wt<CT>(i) * (drho + rho * ((CT(3) * cu[i] + CT(4.5) * (cu[i] * cu[i])) - uu15));
        d[i] = d[i] + omega * (deq - d[i]);
    }
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

}  // namespace lbm
