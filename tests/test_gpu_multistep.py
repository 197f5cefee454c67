"""The multi-step (temporally blocked) kernel against single steps, bit for bit."""

import numpy as np
import pytest
import torch

from latticeboltzmann_amd import _native as nat

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(777)


def populations(nx, ny, dtype):
    from oracle import d2q9_numpy as o
    f = o.weights(np.float64)[:, None, None] * (1 + 0.2 * RNG.standard_normal((9, nx, ny)))
    return torch.from_numpy(f).to(dtype).cuda()


def to_layout(soa, layout):
    from latticeboltzmann_amd import kernels
    return soa.clone() if layout == nat.LAYOUT_SOA else kernels.relayout(soa, nat.LAYOUT_SOA, layout)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("bc", [nat.BC_CAVITY, nat.BC_PERIODIC])
@pytest.mark.parametrize("layout", [nat.LAYOUT_ROWS, nat.LAYOUT_SOA])
def test_multi_step_equals_single_steps(gpu, dtype, bc, layout):
    from latticeboltzmann_amd import kernels
    nvar = nat.load().lbm_step_multi_variant_count()
    for v in range(nvar):
        K, mnx, mny = kernels.step_multi_info(v)
        for nx, ny in [(mnx, mny), (mnx + 27, mny + 131), (130, 1030)]:
            print("multi", dtype, bc, layout, v, K, nx, ny, flush=True)
            f = to_layout(populations(nx, ny, dtype), layout)
            a = f.clone()
            b = torch.empty_like(f)
            m1 = [torch.empty((nx, ny), dtype=dtype, device="cuda") for _ in range(3)]
            m2 = [torch.empty((nx, ny), dtype=dtype, device="cuda") for _ in range(3)]
            for s in range(K):
                kernels.step(a, b, 1.7, bc, 0.1, moments=m1 if s == K - 1 else None, layout=layout)
                a, b = b, a
            c = torch.empty_like(f)
            kernels.step_multi(f, c, 1.7, bc, 0.1, moments=m2, variant=v, layout=layout)
            assert torch.equal(a, c), (v, K, nx, ny)
            for x, y in zip(m1, m2):
                assert torch.equal(x, y)


def test_multi_step_rejects_small_grids(gpu):
    from latticeboltzmann_amd import kernels
    K, mnx, mny = kernels.step_multi_info(0)
    f = populations(mnx - 1, mny, torch.float64)
    with pytest.raises(ValueError):
        kernels.step_multi(f, torch.empty_like(f), 1.0, variant=0)
    with pytest.raises(ValueError):
        kernels.step_multi(f, f, 1.0, variant=0)
    g = populations(64, 128, torch.float64)
    h = torch.empty_like(g)
    lib, sp = nat.load(), nat.stream_ptr()
    args = lambda **kw: dict(dict(layout=nat.LAYOUT_SOA, bc=nat.BC_CAVITY, variant=-1), **kw)  # noqa: E731
    for kw, code in ((args(layout=7), -3), (args(bc=5), -3), (args(variant=99), -3)):
        assert lib.lbm_step_multi_f64(g.data_ptr(), h.data_ptr(), 64, 128, kw["layout"], 1.0, kw["bc"],
                                      0.1, None, None, None, kw["variant"], sp) == code
    # moments all or none
    m = torch.empty((64, 128), dtype=torch.float64, device="cuda")
    assert lib.lbm_step_multi_f64(g.data_ptr(), h.data_ptr(), 64, 128, nat.LAYOUT_SOA, 1.0,
                                  nat.BC_CAVITY, 0.1, m.data_ptr(), None, None, -1, sp) == -1
    k = __import__("ctypes").c_int()
    a, b = __import__("ctypes").c_int64(), __import__("ctypes").c_int64()
    assert lib.lbm_step_multi_info(99, k, a, b) == -3
    torch.cuda.synchronize()


def test_lattice_multi_step_run(gpu):
    # a Lattice advancing 2..4 steps per launch reproduces single steps, with
    # moments and odd step counts
    from latticeboltzmann_amd.lattice import Lattice
    nvar = nat.load().lbm_step_multi_variant_count()
    ref = Lattice(96, 200, 1.3, "cavity", 0.1, multi_step=False)
    mr = [torch.empty((96, 200), dtype=torch.float64, device="cuda") for _ in range(3)]
    ref.step(37, moments=mr)
    for v in range(nvar):
        lat = Lattice(96, 200, 1.3, "cavity", 0.1)
        lat.set_multi_variant(v)
        m = [torch.empty_like(x) for x in mr]
        lat.step(37, moments=m)
        assert lat.launches(37) < 37
        assert torch.equal(lat.f, ref.f), v
        for x, y in zip(m, mr):
            assert torch.equal(x, y)


@pytest.mark.parametrize("fmt", ["f16", "bf16", "f32"])
@pytest.mark.parametrize("compute", ["f32", "f64"])
@pytest.mark.parametrize("rounding", ["stochastic", "nearest"])
def test_dev_multi_step_equals_single_steps(gpu, fmt, compute, rounding):
    # stored deviations: each intermediate step of a multi-step launch rounds
    # (and dithers) exactly as a single-step launch does
    from latticeboltzmann_amd.precision import DeviationLattice, multi_info
    nvar = nat.load().lbm_dev_multi_variant_count()
    for bc in ("cavity", "periodic"):
        for v in range(nvar):
            K, mnx, mny = multi_info(v)
            if compute == "f64" and K > 4:
                continue  # the library refuses FP64-collide shapes beyond four steps
            nx, ny = mnx + 5, mny + 37
            print("dev multi", fmt, compute, rounding, bc, v, K, nx, ny, flush=True)
            rho = 1 + 0.01 * RNG.standard_normal((nx, ny))
            ux, uy = 0.05 * RNG.standard_normal((nx, ny)), 0.05 * RNG.standard_normal((nx, ny))
            kw = dict(fmt=fmt, rounding=rounding, compute=compute, seed=11)
            ref = DeviationLattice(nx, ny, 1.3, bc, 0.1, multi_step=False, **kw)
            lat = DeviationLattice(nx, ny, 1.3, bc, 0.1, multi_step=v, **kw)
            assert lat.multi_k == K
            ref.init_equilibrium(rho, ux, uy)
            lat.init_equilibrium(rho, ux, uy)
            vr = [torch.empty((nx, ny), dtype=torch.float32, device="cuda") for _ in range(2)]
            vl = [torch.empty_like(t) for t in vr]
            ref.step(2 * K + 1, velocity=vr)
            lat.step(2 * K + 1, velocity=vl)
            assert lat.launches(2 * K + 1) == 3
            assert torch.equal(ref.d, lat.d), (bc, v)
            for a, b in zip(vr, vl):
                assert torch.equal(a, b)
            lat.step(K, velocity=vl)
            ref.step(K, velocity=vr)
            assert torch.equal(ref.d, lat.d)
            for a, b in zip(vr, vl):
                assert torch.equal(a, b)


@pytest.mark.parametrize("bc", [nat.BC_CAVITY, nat.BC_PERIODIC])
def test_multi_step_large_ragged_grid(gpu, bc):
    # 4099 x 8191 FP64: no side is a multiple of a tile, 2.4 GB per buffer
    from latticeboltzmann_amd import kernels
    nx, ny = 4099, 8191
    print("multi large", bc, nx, ny, flush=True)
    rows = 1 + 0.01 * torch.rand((nx, ny), dtype=torch.float64, device="cuda")
    ux = 0.05 * torch.rand((nx, ny), dtype=torch.float64, device="cuda")
    uy = 0.05 * torch.rand((nx, ny), dtype=torch.float64, device="cuda")
    f = kernels.relayout(kernels.equilibrium(rows, ux, uy), nat.LAYOUT_SOA, nat.LAYOUT_ROWS)
    del rows, ux, uy
    for v in (nat.load().lbm_step_multi_default_variant(), 0, 12):
        K = kernels.step_multi_info(v)[0]
        a, b = f.clone(), torch.empty_like(f)
        for _ in range(K):
            kernels.step(a, b, 0.9, bc, 0.1, layout=nat.LAYOUT_ROWS)
            a, b = b, a
        kernels.step_multi(f, b, 0.9, bc, 0.1, variant=v, layout=nat.LAYOUT_ROWS)
        assert torch.equal(a, b), v
        del a, b
        torch.cuda.empty_cache()


def test_step_graph_replay_after_odd_eager_steps(gpu):
    # A captured graph reads the buffer it was captured with; eager steps in
    # between that leave the state in the other buffer must not make the next
    # replay step a stale state.
    from latticeboltzmann_amd.lattice import Lattice, StepGraph
    omega = 1.529988
    lat = Lattice(64, 64, omega, "cavity", 0.1, torch.float64)
    ref = Lattice(64, 64, omega, "cavity", 0.1, torch.float64, multi_step=False)
    assert lat.launches(8) % 2 == 0
    with StepGraph(lat, 8) as advance:
        assert lat.launches(8) % 2 == 0
        advance()
        lat.step(1)  # one single-step launch: the state moves to the other buffer
        advance()
        lat.step(3)
        advance()
    ref.step(8 + 1 + 8 + 3 + 8)
    assert lat.time == ref.time == 28
    assert np.array_equal(lat.populations(), ref.populations())
