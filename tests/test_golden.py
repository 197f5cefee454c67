"""The committed oracle vectors (tests/golden/d2q9_vectors.npz): the oracle
still produces every bit of them, and their permutation parts are
permutations."""

import os

import numpy as np

from oracle import d2q9_numpy as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "d2q9_vectors.npz")


def load():
    return np.load(GOLDEN)  # allow_pickle is False by default


def test_oracle_reproduces_golden():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    fresh = mod.vectors()
    g = load()
    assert sorted(fresh) == sorted(g.files)
    for k in g.files:
        assert fresh[k].dtype == g[k].dtype and np.array_equal(fresh[k], g[k]), k


def test_maps_are_permutations():
    g = load()
    for k in ("stream_map_5x7", "bounce_map_5x7", "bounce_map_8x8"):
        m = g[k]
        assert np.array_equal(np.sort(m.ravel()), np.arange(m.size)), k
    # periodic stream: channel i of node (k, l) comes from node (k - cx, l - cy)
    m = g["stream_map_5x7"]
    for i, (cx, cy) in enumerate(o.C_IC):
        src = np.unravel_index(m[i], (9, 5, 7))
        assert np.all(src[0] == i)
        kk, ll = np.meshgrid(np.arange(5), np.arange(7), indexing="ij")
        assert np.array_equal(src[1], (kk - cx) % 5) and np.array_equal(src[2], (ll - cy) % 7)


def test_golden_states_conserve_what_they_must():
    g = load()
    for s in ("5x7", "8x8"):
        for w in (0.5, 1.0, 1.7):
            a, b = g["collide_in_%s_f64" % s], g["collide_%s_w%s_f64" % (s, w)]
            assert np.allclose(a.sum(0), b.sum(0), rtol=0, atol=1e-15)
    # shear wave: mass conserved to rounding, periodic
    f = g["shear_32_s10_f64"]
    assert abs(f.sum() - 32 * 32) < 1e-11
