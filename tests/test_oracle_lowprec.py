"""Pins the reduced-precision codec of the oracle (oracle/lowprec.py)."""

import numpy as np
import pytest

from oracle import d2q9_numpy as o
from oracle import lowprec as lp

RNG = np.random.default_rng(99)


def test_bf16_rounding_known_values():
    # 1 + 2^-8 lies exactly between two bf16 values (1 and 1 + 2^-7): ties to even -> 1
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    bits = lp._f32_to_bf16_bits(x)
    back = lp._bf16_bits_to_f32(bits)
    assert back[0] == 1.0
    assert back[1] == 1.0
    assert back[2] == 1 + 2 ** -6          # 1 + 3*2^-8 rounds up to 1 + 4*2^-8
    assert back[3] == -2.5
    assert np.isfinite(back[4]) and abs(back[4] / 3.0e38 - 1) < 2 ** -8
    assert bits[5] in (0, 1)               # subnormal stays tiny


def test_bf16_matches_truncation_plus_half_ulp_bound():
    x = (RNG.standard_normal(100000) * 10.0 ** RNG.uniform(-6, 3, 100000)).astype(np.float32)
    back = lp._bf16_bits_to_f32(lp._f32_to_bf16_bits(x)).astype(np.float64)
    err = np.abs(back - x.astype(np.float64))
    assert (err <= 0.5 * lp.ulp(x, "bf16") * (1 + 1e-6)).all()


@pytest.mark.parametrize("fmt", lp.FORMATS)
def test_round_trip_error_within_half_ulp(fmt):
    f = o.weights(np.float64)[:, None, None] * (1 + 0.3 * RNG.standard_normal((9, 12, 10)))
    scale = 1024.0 if fmt == "f16" else 1.0
    d = lp.encode(f, fmt, scale)
    back = lp.decode(d, fmt, scale)
    x = (f - o.weights()[:, None, None]) * scale
    # a single rounding from FP64 to the storage format
    bound = 0.5 * lp.ulp(x, fmt) / scale
    assert (np.abs(back - f) <= bound * (1 + 1e-9) + 1e-300).all()


def test_numpy_float16_cast_rounds_once_from_float64():
    # 1 + 2^-11 + 2^-40 is above the midpoint of 1 and 1 + 2^-10: a single
    # rounding gives 1 + 2^-10, a detour through float32 would give 1
    x = np.array([1 + 2 ** -11 + 2 ** -40])
    assert x.astype(np.float16)[0] == np.float16(1 + 2 ** -10)
    assert x.astype(np.float32).astype(np.float16)[0] == np.float16(1.0)


def test_round_to_odd_gives_single_rounding_bf16():
    # x = 1 + 2^-8 + 2^-30: double rounding via float32 gives 1 (tie to even),
    # a single rounding must give 1 + 2^-7 (x is above the midpoint)
    x = np.array([1 + 2 ** -8 + 2 ** -30, -(1 + 2 ** -8 + 2 ** -30), 1 + 2 ** -8, 0.75])
    bits = lp._f32_to_bf16_bits(lp._f64_to_f32_round_to_odd(x))
    back = lp._bf16_bits_to_f32(bits)
    assert back[0] == 1 + 2 ** -7 and back[1] == -(1 + 2 ** -7)
    assert back[2] == 1.0 and back[3] == 0.75
    # agrees with brute force: nearest of the two bf16 neighbours
    y = RNG.standard_normal(20000) * 10.0 ** RNG.uniform(-4, 4, 20000)
    got = lp._bf16_bits_to_f32(lp._f32_to_bf16_bits(lp._f64_to_f32_round_to_odd(y))).astype(float)
    lo = lp._bf16_bits_to_f32((y.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16))
    err = np.abs(got - y)
    assert (err <= 0.5 * lp.ulp(lo, "bf16") * (1 + 1e-12)).all()


def test_rows_layout_is_a_transpose():
    f = o.weights(np.float64)[:, None, None] * (1 + 0.1 * RNG.standard_normal((9, 5, 7)))
    a = lp.encode(f, "f16", 1024.0, "soa")
    b = lp.encode(f, "f16", 1024.0, "rows")
    assert np.array_equal(a.transpose(1, 0, 2), b)
    assert np.array_equal(lp.decode(b, "f16", 1024.0, "rows"), lp.decode(a, "f16", 1024.0))


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_conserving_rounding_keeps_moments_to_one_rounding(fmt):
    # post-collision deviations of a flowing state, scaled as stored
    scale = 1024.0 if fmt == "f16" else 1.0
    rho = 1 + 0.01 * RNG.standard_normal((40, 50))
    ux, uy = 0.08 * RNG.standard_normal((2, 40, 50))
    feq = o.equilibrium(rho, ux, uy)
    f = feq * (1 + 0.02 * RNG.standard_normal(feq.shape))
    x = ((f - o.weights()[:, None, None]) * scale).astype(np.float32)
    exact = lp.conserved(x)
    comp = lp.to_float(lp.round_conserving(x, fmt), fmt)
    plain = lp.round_values(x, fmt)[1]
    big = lp.ulp(np.abs(x).max(0), fmt)
    ratios = []
    for a, p, b in zip(lp.conserved(comp), lp.conserved(plain), exact):
        err_c = np.abs(a - b)
        assert (err_c <= 1.0 * big + 1e-6 * np.abs(x).max(0)).all()
        ratios.append(err_c.mean() / np.abs(p - b).mean())
    # versus independent rounding: mass error (one rounding of the small rest
    # deviation instead of nine) drops ~4x, momentum (one rounding of W or S
    # instead of six) ~1.6x
    assert ratios[0] < 0.35 and ratios[1] < 0.75 and ratios[2] < 0.75
    # each value stays within a few units in the last place
    assert (np.abs(comp - x) <= 4 * lp.ulp(np.abs(x).max(0), fmt)[None]).all()
