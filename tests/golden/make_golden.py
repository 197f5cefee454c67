"""Writes tests/golden/d2q9_vectors.npz: input/output vectors of the numpy
oracle, frozen so that the oracle and the HIP kernels are both checked
against fixed bits.

The reference holds no golden vectors for this path; these come from the
oracle (oracle/d2q9_numpy.py), which is pinned by the reference's KATs and by
closed forms in tests/test_oracle_numpy.py.  Regenerate with
``python tests/golden/make_golden.py`` (deterministic; the file only changes
if the oracle's arithmetic does).

Contents (for dt in f64, f32):
* collide_in_{5x7,8x8}_{dt}: random populations (seed 7);
  collide_{shape}_w{omega}_{dt}: after one BGK collide at omega in
  0.5, 1.0, 1.7, plus the returned moments (rho, ux, uy) as collide_mom_...;
* cavity_{8x8,16x16}_s{1,2,10}_{dt}: cavity from rest, omega 1.7, u0 0.1;
* shear_32_s10_{dt}: 32 x 32 shear wave (A 0.1, direction x), omega 1.0;
* stream_map_5x7, bounce_map_{5x7,8x8}: the flat source index of every
  output population of one periodic stream / one wall stream with u0 = 0
  (both are permutations).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import d2q9_numpy as o  # noqa: E402

OUT = os.path.join(HERE, "d2q9_vectors.npz")
OMEGAS = (0.5, 1.0, 1.7)


def vectors():
    v = {}
    rng = np.random.default_rng(7)
    base = {s: o.weights(np.float64)[:, None, None] * (1 + 0.2 * rng.standard_normal((9,) + s))
            for s in ((5, 7), (8, 8))}
    for dtn, dt in (("f64", np.float64), ("f32", np.float32)):
        cast = (lambda x: x) if dt == np.float64 else np.float32
        for s, f0 in base.items():
            tag = "%dx%d" % s
            f0 = f0.astype(dt)
            v["collide_in_%s_%s" % (tag, dtn)] = f0
            for w in OMEGAS:
                f = f0.copy()
                rho, ux, uy = o.collide(f, cast(w))
                v["collide_%s_w%s_%s" % (tag, w, dtn)] = f
                v["collide_mom_%s_w%s_%s" % (tag, w, dtn)] = np.stack([rho, ux, uy])
        for n in (8, 16):
            f = o.cavity_init(n, n, dt)
            done = 0
            for s in (1, 2, 10):
                for _ in range(s - done):
                    o.cavity_step(f, cast(1.7), 0.1)
                done = s
                v["cavity_%dx%d_s%d_%s" % (n, n, s, dtn)] = f.copy()
        f = o.shear_wave_init(32, 32, 0.1, "x", dt)
        for _ in range(10):
            o.periodic_step(f, cast(1.0))
        v["shear_32_s10_%s" % dtn] = f
    idx = np.arange(9 * 5 * 7, dtype=np.float64).reshape(9, 5, 7)
    g = idx.copy()
    o.stream(g)
    v["stream_map_5x7"] = g.astype(np.int64)
    for s in ((5, 7), (8, 8)):
        g = np.arange(9 * s[0] * s[1], dtype=np.float64).reshape((9,) + s)
        o.stream_and_bounce_back(g, 0.0)
        v["bounce_map_%dx%d" % s] = g.astype(np.int64)
    return v


if __name__ == "__main__":
    np.savez_compressed(OUT, **vectors())
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
