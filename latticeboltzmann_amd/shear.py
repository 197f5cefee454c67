"""Shear-wave decay on the GPU: initial state, amplitude observables, viscosity fit.

The experiment of ``shear_wave_opt1.py:157-174`` and
``simple_flows/shear_wave_decay.py:136-167``: a periodic lattice starts at
f = f_eq(1, u) with a sine velocity profile, BGK relaxes it, and the fitted
decay rate of the amplitude, A(t) = A0 exp(-nu k^2 t), k = 2 pi / wavelength,
gives the viscosity, against the theory nu = (1/omega - 1/2)/3.

* ``direction="x"``: u_x = A sin(2 pi l / ny) (velocity along x, varying
  along y: the geometry of ``shear_wave_decay.py:136-143``);
  ``direction="y"``: u_y = A sin(2 pi k / nx) (``shear_wave_opt1.py:157-163``).
* Amplitude by projection on the initial sine over the line through the
  middle of the box, (2/n) sum u sin (``shear_wave_opt1.py:174``), or as the
  reference's ``shear_wave_decay.py:155-158`` does, the largest entry of
  (2/n) |FFT(u)| of that line.
* Fits in FP64: least squares of log A against t, the one-parameter form
  of ``shear_wave_decay.py:72-73, 167`` (A0 fixed, least squares in A) by
  Gauss-Newton, or the two-point log ratio of ``:526``.
* ``omega_viscosity_sweep``: the omega -> nu experiment of ``:479-538``, every
  omega in one batched launch per step.

The lattice is the FP64 (or FP32) ``Lattice``; the amplitude is read from the
device moments (one line per sample).
"""

import math

import numpy as np
import torch

from .lattice import Lattice


def viscosity_theory(omega):
    """nu = (1/omega - 1/2) / 3 (``shear_wave_decay.py:162``)."""
    return (1.0 / omega - 0.5) / 3.0


def amplitude_projection(line):
    """(2/n) sum_j u_j sin(2 pi j / n) of a velocity line (numpy, FP64)."""
    u = np.asarray(line, dtype=np.float64)
    n = u.shape[0]
    return float(np.dot(u, np.sin(2 * np.pi * np.arange(n) / n)) * 2.0 / n)


def amplitude_fft(line):
    """max_k (2/n) |FFT(u)_k| of a velocity line (``shear_wave_decay.py:155-158``)."""
    u = np.asarray(line, dtype=np.float64)
    return float(np.max(2.0 / u.shape[0] * np.abs(np.fft.fft(u))))


def fit_loglinear(t, amplitude, wavelength):
    """nu from the least-squares slope of log A(t), FP64."""
    t = np.asarray(t, dtype=np.float64)
    a = np.log(np.abs(np.asarray(amplitude, dtype=np.float64)))
    k = 2 * math.pi / wavelength
    return -np.polyfit(t, a, 1)[0] / (k * k)


def fit_exponential(t, amplitude, wavelength, a0, nu0=None, iters=50):
    """nu minimizing sum (a0 exp(-nu k^2 t) - A(t))^2 with a0 fixed
    (the model of ``shear_wave_decay.py:72-73`` fitted at ``:167``), by
    Gauss-Newton from the log-linear estimate."""
    t = np.asarray(t, dtype=np.float64)
    a = np.asarray(amplitude, dtype=np.float64)
    k2 = (2 * math.pi / wavelength) ** 2
    nu = fit_loglinear(t, a, wavelength) if nu0 is None else float(nu0)
    for _ in range(iters):
        m = a0 * np.exp(-nu * k2 * t)
        j = -k2 * t * m  # d model / d nu
        den = float(np.dot(j, j))
        if den == 0.0:
            break
        step = float(np.dot(j, a - m)) / den
        nu += step
        if abs(step) <= 1e-15 * max(abs(nu), 1e-300):
            break
    return nu


def viscosity_log_ratio(a0, a_t, t, wavelength):
    """nu = -ln(A(t) / A0) / (k^2 t), k = 2 pi / wavelength: the two-point
    estimate of ``shear_wave_decay.py:526``."""
    k = 2 * math.pi / wavelength
    return -math.log(a_t / a0) / (k * k * t)


def omega_viscosity_sweep(omegas, n=320, steps=1000, amplitude=0.1, fmt="f32"):
    """The omega -> nu experiment of ``shear_wave_decay.py:479-538``: one n x n
    shear wave (u_x = A sin(2 pi y / n)) per omega, ``steps`` steps, nu from
    the log ratio of the final to the initial amplitude (projection on the
    initial sine).  With n a multiple of 64 every omega runs in one batched
    lattice set (four steps of every instance per launch, ``fmt`` storage,
    FP64 collide); otherwise one FP64 lattice per omega.  Returns (nu, nu_theory)
    as arrays in the order of ``omegas``."""
    omegas = [float(w) for w in omegas]
    if n % 64 == 0 and steps % 2 == 0:
        from .batch import BatchedLattices
        inst = [dict(nx=n, ny=n, omega=w, amplitude=amplitude, direction="x", bc="periodic")
                for w in omegas]
        bat = BatchedLattices(inst, fmt=fmt)
        times, amps = bat.run(steps, sample_every=steps)
        nu = [viscosity_log_ratio(amps[0, i], amps[-1, i], times[-1] - times[0], n)
              for i in range(len(omegas))]
    else:
        nu = []
        for w in omegas:
            sw = ShearWave(n, n, w, amplitude, "x")
            a0 = sw.amplitude()
            sw.lat.step(steps)
            nu.append(viscosity_log_ratio(a0, sw.amplitude(), steps, n))
    return np.asarray(nu), viscosity_theory(np.asarray(omegas))


class ShearWave:
    """A periodic nx x ny lattice started as a shear wave."""

    def __init__(self, nx, ny, omega, amplitude=0.1, direction="x", dtype=torch.float64,
                 device=None, variant="auto"):
        if direction not in ("x", "y"):
            raise ValueError("direction must be 'x' or 'y'")
        self.nx, self.ny, self.omega = int(nx), int(ny), float(omega)
        self.amplitude0, self.direction = float(amplitude), direction
        self.lat = Lattice(self.nx, self.ny, self.omega, "periodic", 0.0, dtype, device=device,
                           variant=variant)
        self.init()

    @property
    def wavelength(self):
        return self.ny if self.direction == "x" else self.nx

    def init(self):
        """f = f_eq(1, u) with the sine profile, evaluated in FP64."""
        n = self.wavelength
        s = self.amplitude0 * np.sin(2 * np.pi * np.arange(n) / n)
        prof = torch.as_tensor(s, dtype=torch.float64)
        if self.direction == "x":
            self.lat.init_equilibrium(ux=prof[None, :])
        else:
            self.lat.init_equilibrium(uy=prof[:, None])
        return self

    def line(self):
        """Velocity component of the wave along the line through the middle (numpy)."""
        _, ux, uy = self.lat.moments()
        if self.direction == "x":
            return ux[self.nx // 2, :].double().cpu().numpy()
        return uy[:, self.ny // 2].double().cpu().numpy()

    def amplitude(self, method="projection"):
        line = self.line()
        return amplitude_fft(line) if method == "fft" else amplitude_projection(line)

    def run(self, steps, every, method="projection"):
        """Amplitude every ``every`` steps (and at the start): (times, amplitudes)."""
        times, amps = [self.lat.time], [self.amplitude(method)]
        for _ in range(steps // every):
            self.lat.step(every)
            times.append(self.lat.time)
            amps.append(self.amplitude(method))
        return np.array(times), np.array(amps)

    def viscosity(self, times, amps, fit="loglinear"):
        if fit == "exponential":
            return fit_exponential(times, amps, self.wavelength, self.amplitude0)
        return fit_loglinear(times, amps, self.wavelength)
