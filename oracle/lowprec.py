"""CPU oracle for the reduced-precision (scaled deviation) population storage.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The storage has no reference counterpart; what pins it is arithmetic that
numpy fixes on its own: IEEE round-to-nearest-even conversions to float32 and
float16 (numpy's casts) and to bfloat16 (restated below from the float32 bit
pattern), and the FP64 oracle step applied to the decoded populations.
"""

import numpy as np

from . import d2q9_numpy as o

FORMATS = ("f32", "bf16", "f16")


def _f32_to_bf16_bits(x):
    """Round float32 values to bfloat16 (nearest, ties to even); uint16 bits."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    rounded = (b + 0x7FFF + lsb) >> 16
    return rounded.astype(np.uint16)


def _bf16_bits_to_f32(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def _f64_to_f32_round_to_odd(x):
    """float32 nearest to x toward zero, with the last bit set when inexact.

    Rounding this intermediate to a format with at least two fewer
    significand bits gives the correctly rounded result of x itself."""
    x = np.asarray(x, dtype=np.float64)
    r = x.astype(np.float32)
    over = np.abs(r.astype(np.float64)) > np.abs(x)
    r = np.where(over, np.nextafter(r, np.float32(0)), r).astype(np.float32)
    inexact = r.astype(np.float64) != x
    bits = r.view(np.uint32) | inexact.astype(np.uint32)
    return bits.view(np.float32)


def encode(f, fmt, scale, layout="soa"):
    """Stored values for FP64 populations f (9, nx, ny): (f - w) * scale computed
    in FP64 and rounded once to the format (nearest, ties to even)."""
    w = o.weights(np.float64)[:, None, None]
    x = (f - w) * scale
    if fmt == "f32":
        d = x.astype(np.float32)
    elif fmt == "f16":
        d = x.astype(np.float16).view(np.uint16)   # numpy rounds float64 -> float16 once
    elif fmt == "bf16":
        d = _f32_to_bf16_bits(_f64_to_f32_round_to_odd(x))
    else:
        raise ValueError(fmt)
    return d.transpose(1, 0, 2).copy() if layout == "rows" else d


def to_float(d, fmt):
    if fmt == "f32":
        return np.asarray(d, dtype=np.float32).astype(np.float64)
    if fmt == "f16":
        return np.asarray(d, dtype=np.uint16).view(np.float16).astype(np.float64)
    return _bf16_bits_to_f32(d).astype(np.float64)


def decode(d, fmt, scale, layout="soa"):
    """FP64 populations w + value / scale, in the reference layout."""
    x = to_float(d, fmt)
    if layout == "rows":
        x = x.transpose(1, 0, 2)
    return o.weights(np.float64)[:, None, None] + x / scale


def round_values(x, fmt):
    """Round float32 values to the format; returns (raw, float64 value)."""
    x = np.asarray(x, dtype=np.float32)
    if fmt == "f32":
        raw = x.copy()
    elif fmt == "f16":
        raw = x.astype(np.float16).view(np.uint16)
    else:
        raw = _f32_to_bf16_bits(x)
    return raw, to_float(raw, fmt)


def round_conserving(x, fmt):
    """Conservation-compensated rounding of scaled deviations x (9, ...).

    Channels 5-8, 1 and 2 are rounded as they are; 3 (W) and 4 (S) at
    targets shifted by the x- and y-momentum residuals of those; the rest
    channel 0 at a target shifted by the mass residual of channels 1-8, so the
    stored mass and momentum each carry a single rounding error.  Returns the
    raw stored values.  Arithmetic in float32, as on the device.
    """
    x = np.asarray(x, dtype=np.float32)
    raw = np.empty(x.shape, dtype=np.float32 if fmt == "f32" else np.uint16)
    e = {}

    def rnd(i, target):
        target = np.asarray(target, dtype=np.float32)
        raw[i], val = round_values(target, fmt)
        return (target.astype(np.float64) - val).astype(np.float32)

    for i in (5, 6, 7, 8, 1, 2):
        e[i] = rnd(i, x[i])
    ex = ((e[5] - e[6]) + (e[8] - e[7])) + e[1]
    ey = ((e[5] + e[6]) - (e[7] + e[8])) + e[2]
    e[3] = rnd(3, x[3] - ex)
    e[4] = rnd(4, x[4] - ey)
    em = ((e[5] + e[6]) + (e[7] + e[8])) + ((e[1] + e[2]) + ((e[3] + ex) + (e[4] + ey)))
    rnd(0, x[0] + em)
    return raw


def conserved(v):
    """(mass, x-momentum, y-momentum) of deviation values v (9, ...), FP64."""
    c = o.C_IC.astype(np.float64)
    v = np.asarray(v, dtype=np.float64)
    return (v.sum(0), np.tensordot(c[:, 0], v, axes=(0, 0)), np.tensordot(c[:, 1], v, axes=(0, 0)))


def ulp(x, fmt):
    """Spacing of the storage format at |x| (x in stored, scaled units)."""
    x32 = np.abs(np.asarray(x, dtype=np.float32))
    if fmt == "f32":
        return np.spacing(x32).astype(np.float64)
    if fmt == "f16":
        return np.spacing(x32.astype(np.float16)).astype(np.float64)
    return np.spacing(x32).astype(np.float64) * 65536.0


def step_reference(d, fmt, scale, omega, bc, u0=0.1, layout="soa"):
    """Expected populations (FP64, reference layout) after one step of stored
    deviations: the FP64 oracle step applied to the decoded state."""
    f = decode(d, fmt, scale, layout)
    if bc == "cavity":
        o.cavity_step(f, omega, u0)
    else:
        o.periodic_step(f, omega)
    return f
