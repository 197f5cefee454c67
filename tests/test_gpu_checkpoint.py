"""Population checkpoints: N + M steps in one run equal N steps, save, load
into a new lattice, M steps -- bit for bit, for the FP64 lattice and for
stored deviations (rounded every step, and stored every 4 steps)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_fp64_lattice_resumes_bitwise(gpu, tmp_path):
    from latticeboltzmann_amd import checkpoint
    from latticeboltzmann_amd.lattice import Lattice
    n_first, n_second = 37, 29  # odd counts: multi-step runs plus single steps
    kw = dict(omega=1.529988, bc="cavity", u0=0.1, dtype=torch.float64)
    whole = Lattice(96, 80, **kw).step(n_first + n_second)
    first = Lattice(96, 80, **kw).step(n_first)
    path = checkpoint.save(first, str(tmp_path / "cavity.npz"))
    first.step(5)  # the saved state must not follow the live lattice
    resumed = checkpoint.restore(Lattice(96, 80, **kw), path).step(n_second)
    assert resumed.time == whole.time == n_first + n_second
    assert np.array_equal(resumed.populations(), whole.populations())
    with pytest.raises(ValueError):
        checkpoint.restore(Lattice(96, 80, omega=1.5, bc="cavity", u0=0.1), path)


@pytest.mark.parametrize("store_every,n_first,n_second", [(1, 37, 29), (4, 36, 28)])
def test_deviation_lattice_resumes_bitwise(gpu, tmp_path, store_every, n_first, n_second):
    # the dither of stochastic rounding is seeded from the step counter, which
    # the checkpoint restores; storing every 4 steps, save at a multiple of 4
    from latticeboltzmann_amd import checkpoint
    from latticeboltzmann_amd.precision import DeviationLattice

    def make():
        return DeviationLattice(96, 80, 1.3, "cavity", 0.1, fmt="f16", rounding="stochastic",
                                seed=3, store_every=store_every)

    whole = make().step(n_first + n_second)
    first = make().step(n_first)
    path = checkpoint.save(first, str(tmp_path / "dev.npz"))
    resumed = checkpoint.restore(make(), path).step(n_second)
    assert resumed.time == whole.time
    assert torch.equal(resumed.d, whole.d)
