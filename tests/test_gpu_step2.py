"""The two-step (temporally blocked) kernel against two single steps, bit for bit."""

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
def test_step2_equals_two_steps(gpu, dtype, bc, layout):
    from latticeboltzmann_amd import kernels
    nvar = nat.load().lbm_step2_variant_count()
    for v in range(nvar):
        mnx, mny = kernels.step2_min_size(v)
        for nx, ny in [(mnx, mny), (mnx + 27, mny + 131), (130, 1030)]:
            print("step2", dtype, bc, layout, v, nx, ny, flush=True)
            f = to_layout(populations(nx, ny, dtype), layout)
            a, b = torch.empty_like(f), torch.empty_like(f)
            m1 = [torch.empty((nx, ny), dtype=dtype, device="cuda") for _ in range(3)]
            m2 = [torch.empty((nx, ny), dtype=dtype, device="cuda") for _ in range(3)]
            kernels.step(f, a, 1.7, bc, 0.1, layout=layout)
            kernels.step(a, b, 1.7, bc, 0.1, moments=m1, layout=layout)
            c = torch.empty_like(f)
            kernels.step2(f, c, 1.7, bc, 0.1, moments=m2, variant=v, layout=layout)
            assert torch.equal(b, c), (v, nx, ny)
            for x, y in zip(m1, m2):
                assert torch.equal(x, y)


def test_step2_rejects_small_grids(gpu):
    from latticeboltzmann_amd import kernels
    mnx, mny = kernels.step2_min_size(0)
    f = populations(mnx - 1, mny, torch.float64)
    with pytest.raises(ValueError):
        kernels.step2(f, torch.empty_like(f), 1.0, variant=0)
    with pytest.raises(ValueError):
        kernels.step2(f, f, 1.0, variant=0)
