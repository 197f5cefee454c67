"""HIP kernels against the committed oracle vectors, bit for bit."""

import os

import numpy as np
import pytest
import torch

from oracle import d2q9_numpy as o

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "d2q9_vectors.npz")
DT = {"f64": (np.float64, torch.float64), "f32": (np.float32, torch.float32)}


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


@pytest.mark.parametrize("dtn", ["f64", "f32"])
@pytest.mark.parametrize("shape", ["5x7", "8x8"])
@pytest.mark.parametrize("w", [0.5, 1.0, 1.7])
def test_collide_golden(gpu, g, dtn, shape, w):
    from latticeboltzmann_amd import kernels
    t = torch.from_numpy(g["collide_in_%s_%s" % (shape, dtn)]).cuda()
    mom = kernels.collide_(t.view(9, -1), w, moments=True)
    assert np.array_equal(t.cpu().numpy(), g["collide_%s_w%s_%s" % (shape, w, dtn)])
    want = g["collide_mom_%s_w%s_%s" % (shape, w, dtn)]
    for a, b in zip(mom, want):
        assert np.array_equal(a.cpu().numpy().reshape(b.shape), b)


@pytest.mark.parametrize("dtn", ["f64", "f32"])
@pytest.mark.parametrize("n", [8, 16])
@pytest.mark.parametrize("layout", ["rows", "soa"])
def test_cavity_golden(gpu, g, dtn, n, layout):
    from latticeboltzmann_amd.lattice import Lattice
    lat = Lattice(n, n, 1.7, "cavity", 0.1, DT[dtn][1], layout=layout)
    done = 0
    for s in (1, 2, 10):
        lat.step(s - done)
        done = s
        assert np.array_equal(lat.populations(), g["cavity_%dx%d_s%d_%s" % (n, n, s, dtn)]), s


@pytest.mark.parametrize("dtn", ["f64", "f32"])
def test_shear_golden(gpu, g, dtn):
    from latticeboltzmann_amd.lattice import Lattice
    lat = Lattice(32, 32, 1.0, "periodic", 0.0, DT[dtn][1])
    lat.load(o.shear_wave_init(32, 32, 0.1, "x", DT[dtn][0]))
    lat.step(10)
    assert np.array_equal(lat.populations(), g["shear_32_s10_%s" % dtn])


def test_permutation_maps_golden(gpu, g):
    from latticeboltzmann_amd import kernels
    idx = torch.arange(9 * 5 * 7, dtype=torch.int64, device="cuda").view(9, 5, 7)
    assert np.array_equal(kernels.stream(idx).cpu().numpy(), g["stream_map_5x7"])
    for s in ((5, 7), (8, 8)):
        f = torch.arange(9 * s[0] * s[1], dtype=torch.float64, device="cuda").view((9,) + s)
        out = kernels.stream_bounce_back(f, 0.0)
        assert np.array_equal(out.cpu().numpy().astype(np.int64), g["bounce_map_%dx%d" % s])
