"""The C restatement (oracle/d2q9_ref.c) agrees bit for bit with the numpy one,
and its _lbkernels-formula collide agrees with both to the reference's 1e-7."""

import numpy as np
import pytest

from oracle import ctypes_ref as c
from oracle import d2q9_numpy as o

RNG = np.random.default_rng(7)


def populations(nx, ny, dtype):
    return (o.weights(np.float64)[:, None, None]
            * (1 + 0.2 * RNG.standard_normal((9, nx, ny)))).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(8, 8), (5, 7), (1, 6), (6, 1), (1, 1), (16, 12)])
@pytest.mark.parametrize("bc", ["cavity", "periodic"])
def test_c_step_bitwise_equals_numpy(dtype, shape, bc):
    f = populations(*shape, dtype)
    g = f.copy()
    omega = dtype(1.7)
    if bc == "cavity":
        mom = o.cavity_step(g, omega, 0.1)
    else:
        mom = o.periodic_step(g, omega)
    h, rho, ux, uy = c.step(f, omega, bc, 0.1)
    assert np.array_equal(g, h)
    for a, b in zip(mom, (rho, ux, uy)):
        assert np.array_equal(a, b)


def test_c_multi_step_run_equals_numpy():
    f = o.cavity_init(12, 10)
    g = f.copy()
    for _ in range(25):
        o.cavity_step(g, 1.53, 0.1)
    assert np.array_equal(c.run(f, 1.53, 25, "cavity", 0.1), g)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c_collide_bitwise_equals_numpy(dtype):
    f = populations(6, 5, dtype)
    a, b = f.copy(), f.copy()
    o.collide(a, dtype(0.8))
    c.collide(b, 0.8)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("omega", [0.5, 1.7])
def test_lbkernels_formula_within_reference_tolerance(omega):
    # tests/02-CollideTest.py:104-111 pins _lbkernels against numpy to 1e-7
    f = np.abs(RNG.random((9, 4, 4)))
    a, b = f.copy(), f.copy()
    o.collide(a, omega)
    c.lbk_collide(b, omega)
    assert np.abs(a - b).max() < 1e-7
    rho, ux, uy = RNG.random(4) + 0.5, RNG.random(4), RNG.random(4)
    e = np.empty((9, 4))
    c.lbk_equilibrium(rho, ux, uy, e)
    assert np.abs(e - o.equilibrium(rho, ux, uy)).max() < 1e-7
