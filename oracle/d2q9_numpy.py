"""numpy restatement of the reference's D2Q9 BGK path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``; never by the product package.

Array convention (as in the reference): populations ``f[i, k, l]`` with
channel ``i`` in 0..8, ``k`` the x index (slow axis), ``l`` the y index (fast,
contiguous axis).  Channel velocities ``C_IC`` follow
``PyLB/Streaming.py:28-29`` (identical at ``cavity_opt1.py:88-89``).

Every arithmetic expression keeps the evaluation order of the reference's
numpy opt1 source, because the GPU kernels are required to reproduce it bit
for bit in FP64.  Python evaluates ``a*b*c`` as ``(a*b)*c`` and ``a+b-c`` as
``(a+b)-c``; numpy's ``np.sum(f, axis=0)`` over the outer axis accumulates
``f0+f1+...+f8`` sequentially (checked in ``tests/test_oracle_numpy.py``).
Python-float constants combine with float32 arrays in float32 (NEP 50 weak
scalars), so the same code is the float32 oracle.
"""

import numpy as np

# Channel velocities (cx, cy): 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE.
# Reference: PyLB/Streaming.py:28-29, cavity_opt1.py:72 (D enum), :88-89.
C_IC = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                 [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int64)
# Opposite channel of each channel (bounce-back partner).
OPPOSITE = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int64)
# Named channel indices (cavity_opt1.py:72).
E, N, W, S, NE, NW, SW, SE = 1, 2, 3, 4, 5, 6, 7, 8

W0, W1234, W5678 = 4 / 9, 1 / 9, 1 / 36          # cavity_opt1.py:92-94


def weights(dtype=np.float64):
    """Lattice weights in ``dtype`` (cavity_opt1.py:95-96)."""
    return np.array([W0, W1234, W1234, W1234, W1234,
                     W5678, W5678, W5678, W5678], dtype=dtype)


# --------------------------------------------------------------------------
# opt1 arithmetic
# --------------------------------------------------------------------------

def equilibrium(rho, ux, uy):
    """f_eq for every node; restates ``cavity_opt1.py:100-132``.

    Returns a new ``(9, *shape)`` array.  Broadcasting between the three
    inputs is allowed (``shear_wave_opt1.py:161-163`` passes ``(nx, 1)``).
    """
    cu5 = ux + uy                      # :119
    cu6 = -ux + uy                     # :120
    cu7 = -ux - uy                     # :121
    cu8 = ux - uy                      # :122
    uu = ux ** 2 + uy ** 2             # :123
    # Each plane: (w*rho) * ((((1 +/- 3*cu) + 4.5*cu**2) - 1.5*uu)), :124-132
    return np.array([
        W0 * rho * (1 - 3 / 2 * uu),
        W1234 * rho * (1 + 3 * ux + 9 / 2 * ux ** 2 - 3 / 2 * uu),
        W1234 * rho * (1 + 3 * uy + 9 / 2 * uy ** 2 - 3 / 2 * uu),
        W1234 * rho * (1 - 3 * ux + 9 / 2 * ux ** 2 - 3 / 2 * uu),
        W1234 * rho * (1 - 3 * uy + 9 / 2 * uy ** 2 - 3 / 2 * uu),
        W5678 * rho * (1 + 3 * cu5 + 9 / 2 * cu5 ** 2 - 3 / 2 * uu),
        W5678 * rho * (1 + 3 * cu6 + 9 / 2 * cu6 ** 2 - 3 / 2 * uu),
        W5678 * rho * (1 + 3 * cu7 + 9 / 2 * cu7 ** 2 - 3 / 2 * uu),
        W5678 * rho * (1 + 3 * cu8 + 9 / 2 * cu8 ** 2 - 3 / 2 * uu)])


def density(f):
    """rho = f0 + f1 + ... + f8, accumulated left to right.

    This is what ``np.sum(f, axis=0)`` (``cavity_opt1.py:156``) computes for
    every grid with more than one node.  For a single node numpy reduces the
    nine contiguous values with its pairwise kernel instead; the restatement
    keeps the sequential order there too.
    """
    rho = f[0] + f[1]
    for i in range(2, 9):
        rho = rho + f[i]
    return rho


def moments(f):
    """(rho, ux, uy) with the opt1 order (``cavity_opt1.py:156-158``)."""
    rho = density(f)
    ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) / rho
    uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) / rho
    return rho, ux, uy


def collide(f, omega):
    """BGK collision in place; restates ``cavity_opt1.py:134-160``.

    Returns ``(rho, ux, uy)`` of the pre-collision state, as the reference.
    ``omega`` may be a Python float or a 0-d array of f's dtype
    (``cavity_opt1.py:68``).
    """
    rho, ux, uy = moments(f)
    f += omega * (equilibrium(rho, ux, uy) - f)
    return rho, ux, uy


def stream(f):
    """Periodic pull streaming in place; restates ``PyLB/Streaming.py:33-46``.

    ``f_i[k, l] <- f_i[(k - cx_i) mod nx, (l - cy_i) mod ny]`` for i = 1..8.
    """
    for i in range(1, 9):
        f[i] = np.roll(f[i], tuple(C_IC[i]), axis=(0, 1))


def stream_and_bounce_back(f, u0=0.1):
    """Stream, then walls and moving lid, in place.

    Restates ``cavity_opt1.py:177-245`` (same as ``cavity_opt2.py:109-177``):
    half-way bounce-back on the bottom (l = 0), left (k = 0) and right
    (k = nx-1) walls; the top wall (l = ny-1) moves with velocity ``u0`` in +x.
    The overrides are applied in the reference's order; the corner block of
    the reference (``:227-245``) only rewrites values already set by the four
    edge rules and is therefore omitted.
    """
    w = weights(f.dtype)
    bottom = f[:, :, 0].copy()        # pre-stream values on each wall
    top = f[:, :, -1].copy()
    left = f[:, 0, :].copy()
    right = f[:, -1, :].copy()

    stream(f)

    # bottom wall (:202-204)
    f[N, :, 0] = bottom[S]
    f[NE, :, 0] = bottom[SW]
    f[NW, :, 0] = bottom[SE]

    # lid density from the post-bottom values, summed left to right (:208-210)
    rho_lid = (top[NW] + top[N] + top[NE]
               + f[NW, :, -1] + f[N, :, -1] + f[NE, :, -1]
               + f[W, :, -1] + f[0, :, -1] + f[E, :, -1])
    # moving lid (:211-213): ((6*w)*rho_lid)*u0 with w in f's dtype
    f[S, :, -1] = top[N]
    f[SE, :, -1] = top[NW] + 6 * w[SE] * rho_lid * u0
    f[SW, :, -1] = top[NE] - 6 * w[SW] * rho_lid * u0

    # left wall (:218-220)
    f[E, 0, :] = left[W]
    f[NE, 0, :] = left[SW]
    f[SE, 0, :] = left[NW]

    # right wall (:223-225)
    f[W, -1, :] = right[E]
    f[NW, -1, :] = right[SE]
    f[SW, -1, :] = right[NE]


def lid_density(f):
    """rho_lid(k) that ``stream_and_bounce_back`` uses for the pre-stream state f.

    Returns a length-nx array (same order of summation).  Used by the mass
    identity test: one wall step changes total mass by
    ``sum_k (6 w u0 rho_lid(k)) * (+1 at k = nx-1, -1 at k = 0)``.
    """
    g = f.copy()
    stream(g)
    g[N, :, 0] = f[S, :, 0]
    g[NE, :, 0] = f[SW, :, 0]
    g[NW, :, 0] = f[SE, :, 0]
    return (f[NW, :, -1] + f[N, :, -1] + f[NE, :, -1]
            + g[NW, :, -1] + g[N, :, -1] + g[NE, :, -1]
            + g[W, :, -1] + g[0, :, -1] + g[E, :, -1])


# --------------------------------------------------------------------------
# Driver-level steps and initial states
# --------------------------------------------------------------------------

def cavity_init(nx, ny, dtype=np.float64):
    """f = f_eq(rho=1, u=0) (``cavity_opt1.py:333-335``)."""
    return equilibrium(np.ones((nx, ny), dtype=dtype),
                       np.zeros((nx, ny), dtype=dtype),
                       np.zeros((nx, ny), dtype=dtype))


def cavity_step(f, omega, u0=0.1):
    """One cavity step in place: walls then collide (``cavity_opt1.py:341-343``)."""
    stream_and_bounce_back(f, u0)
    return collide(f, omega)


def periodic_step(f, omega):
    """One periodic step in place: stream then collide (``shear_wave_opt1.py:171-172``)."""
    stream(f)
    return collide(f, omega)


def shear_wave_init(nx, ny, amplitude=0.1, direction="x", dtype=np.float64):
    """Shear-wave initial state f = f_eq(1, u).

    ``direction='x'``: u_x = A sin(2 pi l / ny), u_y = 0 (velocity along x,
    varying along y; the geometry of ``shear_wave_decay.py:136-143``).
    ``direction='y'``: u_y = A sin(2 pi k / nx), u_x = 0 (the geometry of
    ``shear_wave_opt1.py:157-163``).
    The sine is evaluated in float64 and then cast to ``dtype``.
    """
    rho = np.ones((nx, ny), dtype=dtype)
    ux = np.zeros((nx, ny), dtype=dtype)
    uy = np.zeros((nx, ny), dtype=dtype)
    if direction == "x":
        ux[:, :] = (amplitude * np.sin(2 * np.pi * np.arange(ny) / ny)).astype(dtype)[None, :]
    elif direction == "y":
        uy[:, :] = (amplitude * np.sin(2 * np.pi * np.arange(nx) / nx)).astype(dtype)[:, None]
    else:
        raise ValueError("direction must be 'x' or 'y'")
    return equilibrium(rho, ux, uy)


def shear_wave_amplitude(f, direction="x"):
    """Amplitude of the shear wave by projection onto the initial sine.

    Follows ``shear_wave_opt1.py:174``: ``(u . sin) * 2 / n`` over the line
    through the middle of the box, in float64.
    """
    _, ux, uy = moments(f.astype(np.float64))
    nx, ny = ux.shape
    if direction == "x":
        line = ux[nx // 2, :]
        return float(np.dot(line, np.sin(2 * np.pi * np.arange(ny) / ny)) * 2 / ny)
    line = uy[:, ny // 2]
    return float(np.dot(line, np.sin(2 * np.pi * np.arange(nx) / nx)) * 2 / nx)


def shear_wave_amplitude_fft(f, direction="x"):
    """Amplitude as ``shear_wave_decay.py:155-158`` measures it: the largest
    entry of (2/n) |FFT(u)| over the line through the middle of the box."""
    _, ux, uy = moments(f.astype(np.float64))
    nx, ny = ux.shape
    line = ux[nx // 2, :] if direction == "x" else uy[:, ny // 2]
    return float(np.max(2.0 / line.shape[0] * np.abs(np.fft.fft(line))))


def viscosity_theory(omega):
    """nu = (1/omega - 1/2)/3 (``shear_wave_decay.py:162``)."""
    return (1.0 / omega - 0.5) / 3.0


def fit_viscosity(t, amplitude, wavelength):
    """Least-squares fit of A(t) = A0 exp(-nu k^2 t) in log space, FP64.

    The reference fits the same model with ``scipy.optimize.curve_fit``
    (``shear_wave_decay.py:167``) or by the two-point log ratio (``:526``).
    """
    t = np.asarray(t, dtype=np.float64)
    a = np.log(np.abs(np.asarray(amplitude, dtype=np.float64)))
    k = 2 * np.pi / wavelength
    slope = np.polyfit(t, a, 1)[0]
    return -slope / (k * k)


# --------------------------------------------------------------------------
# opt0 forms (np.dot), as used by the reference's unit test
# --------------------------------------------------------------------------

def equilibrium_opt0(rho, u_ckl):
    """``d2q9_equilibrium_ref`` (``tests/02-CollideTest.py:43-59``, ``cavity_opt0.py:93-109``)."""
    w = weights(np.float64)
    cu = np.dot(u_ckl.T, C_IC.T.astype(u_ckl.dtype)).T
    uu = np.sum(u_ckl ** 2, axis=0)
    return (w * (rho * (1 + 3 * cu + 9 / 2 * cu ** 2 - 3 / 2 * uu)).T).T


def collide_opt0(f, omega):
    """``d2q9_collide_ref`` (``tests/02-CollideTest.py:62-89``)."""
    rho = np.sum(f, axis=0)
    u = np.dot(f.T, C_IC.astype(f.dtype)).T / rho
    f += omega * (equilibrium_opt0(rho, u) - f)
    return rho, u
