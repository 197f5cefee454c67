"""Shear-wave observables and fits of the product (host functions)."""

import math

import numpy as np

from latticeboltzmann_amd.shear import (amplitude_fft, amplitude_projection, fit_exponential,
                                        fit_loglinear, viscosity_theory)


def test_observables_of_a_sine():
    n = 96
    line = 0.03 * np.sin(2 * np.pi * np.arange(n) / n)
    assert abs(amplitude_projection(line) - 0.03) < 1e-16
    assert abs(amplitude_fft(line) - 0.03) < 1e-16
    # a phase shift leaves the FFT amplitude unchanged and shrinks the projection
    shifted = 0.03 * np.sin(2 * np.pi * np.arange(n) / n + 0.4)
    assert abs(amplitude_fft(shifted) - 0.03) < 1e-15
    assert abs(amplitude_projection(shifted) - 0.03 * math.cos(0.4)) < 1e-15


def test_fits_recover_nu_from_exact_decay():
    n, nu = 256, 0.05
    t = np.arange(0, 2000, 100, dtype=float)
    a = 0.1 * np.exp(-nu * (2 * np.pi / n) ** 2 * t)
    assert abs(fit_loglinear(t, a, n) / nu - 1) < 1e-12
    assert abs(fit_exponential(t, a, n, 0.1) / nu - 1) < 1e-12
    # with A0 fixed and a distorted first point the two fits differ, and the
    # exponential one minimizes the squared amplitude residual
    b = a.copy()
    b[0] *= 1.01
    nu_e = fit_exponential(t, b, n, 0.1)
    k2 = (2 * np.pi / n) ** 2
    r = lambda v: np.sum((0.1 * np.exp(-v * k2 * t) - b) ** 2)  # noqa: E731
    assert r(nu_e) <= r(nu_e * (1 + 1e-6)) and r(nu_e) <= r(nu_e * (1 - 1e-6))


def test_viscosity_theory():
    assert viscosity_theory(1.0) == 1.0 / 6.0
    assert abs(viscosity_theory(1.7) - (1 / 1.7 - 0.5) / 3) < 1e-16


def test_viscosity_log_ratio_exact():
    import math
    from latticeboltzmann_amd.shear import viscosity_log_ratio
    nu, n, t = 0.0731, 300, 1000
    k2 = (2 * math.pi / n) ** 2
    a_t = 0.1 * math.exp(-nu * k2 * t)
    assert abs(viscosity_log_ratio(0.1, a_t, t, n) - nu) < 1e-14
