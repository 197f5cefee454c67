"""Shear-wave decay through the product's ShearWave (256 x 256, omega 1, A 0.1,
1000 steps): initial state, observables and fitted viscosity."""

import numpy as np
import pytest

from oracle import d2q9_numpy as o

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("direction", ["x", "y"])
def test_shear_wave_matches_oracle_and_theory(gpu, direction):
    from latticeboltzmann_amd.shear import ShearWave, viscosity_theory
    n, omega, amp = 256, 1.0, 0.1
    sw = ShearWave(n, n, omega, amp, direction)
    f0 = sw.lat.populations()
    assert np.array_equal(f0, o.shear_wave_init(n, n, amp, direction))
    times, a_proj = sw.run(1000, 100)
    f = sw.lat.populations()
    assert abs(a_proj[-1] - o.shear_wave_amplitude(f, direction)) < 1e-15
    assert abs(sw.amplitude("fft") - o.shear_wave_amplitude_fft(f, direction)) < 1e-15
    nu_theory = viscosity_theory(omega)
    nu_lin = sw.viscosity(times[1:], a_proj[1:])
    nu_exp = sw.viscosity(times, a_proj, fit="exponential")
    print(direction, nu_lin / nu_theory - 1, nu_exp / nu_theory - 1)
    assert abs(nu_lin / nu_theory - 1) < 1e-3
    assert abs(nu_exp / nu_theory - 1) < 1e-2
