"""Pins the numpy oracle (oracle/d2q9_numpy.py) without importing the reference.

Each check is chosen so that a plausible restatement mistake (wrong channel
table, wrong shift sign, swapped wall rule, missing lid term, dropped or
reordered term in the equilibrium) fails at least one of them.
"""

import numpy as np
import pytest

from oracle import d2q9_numpy as o

RNG = np.random.default_rng(20240601)


def random_populations(nx, ny, dtype=np.float64, spread=0.2, rng=RNG):
    f = o.weights(np.float64)[:, None, None] * (1 + spread * rng.standard_normal((9, nx, ny)))
    return f.astype(dtype)


# ---------------------------------------------------------------------------
# channel table and weights
# ---------------------------------------------------------------------------

def test_channel_table_matches_direction_names():
    # D = IntEnum('D', 'E N W S NE NW SW SE') (cavity_opt1.py:72) with the
    # velocities of PyLB/Streaming.py:28-29.
    names = {1: (1, 0), 2: (0, 1), 3: (-1, 0), 4: (0, -1),
             5: (1, 1), 6: (-1, 1), 7: (-1, -1), 8: (1, -1)}
    for i, c in names.items():
        assert tuple(o.C_IC[i]) == c
    assert tuple(o.C_IC[0]) == (0, 0)
    for i in range(9):
        assert tuple(o.C_IC[o.OPPOSITE[i]]) == tuple(-o.C_IC[i])


def test_weights_closed_form():
    w = o.weights()
    assert abs(w.sum() - 1.0) < 1e-15
    # second moment sum w c c = I/3 (isotropy of D2Q9)
    m2 = np.einsum("i,ia,ib->ab", w, o.C_IC.astype(float), o.C_IC.astype(float))
    assert np.allclose(m2, np.eye(2) / 3, atol=1e-15)


# ---------------------------------------------------------------------------
# streaming: the reference's own known-answer tests
# ---------------------------------------------------------------------------

def test_stream_one_step_per_channel_kat():
    # tests/simpleFlowsTest.py:260-297: a line of ones at x=1, y=1..7 moves by c_i.
    n = 9
    for i in range(1, 9):
        grid = np.zeros((9, n, n))
        grid[i, 1, 1:8] = 1
        before = grid[i].copy()
        o.stream(grid)
        cx, cy = o.C_IC[i]
        for y in range(1, 8):
            assert grid[i, (1 + cx) % n, (y + cy) % n] == before[1, y] == 1
        assert grid[i].sum() == 7


def test_stream_periodic_return_kat():
    # tests/simpleFlowsTest.py:176-196: a pulse at the centre of a 9x9 grid
    # in every channel is back after 9 steps.
    grid = np.zeros((9, 9, 9))
    grid[:, 4, 4] = 1
    for _ in range(9):
        o.stream(grid)
    assert (grid[:, 4, 4] == 1).all() and grid.sum() == 9


def test_stream_degenerate_axis_kat():
    # tests/simpleFlowsTest.py:198-230: a (9, 1, 9) grid; after 9 steps every
    # channel is back at y = 0; after 4 steps channels 5,6 sit at y=4 and 7,8 at y=5.
    grid = np.zeros((9, 1, 9))
    grid[:, :, 0] = 1
    for step in range(9):
        o.stream(grid)
        if step == 3:
            assert grid[0, 0, 0] == 1
            assert grid[5, 0, 4] == grid[6, 0, 4] == 1
            assert grid[7, 0, 5] == grid[8, 0, 5] == 1
    assert (grid[:, 0, 0] == 1).all()


def test_stream_is_the_index_permutation():
    f = RNG.random((9, 5, 7))
    g = f.copy()
    o.stream(g)
    for i in range(9):
        cx, cy = o.C_IC[i]
        for k in range(5):
            for l in range(7):
                assert g[i, k, l] == f[i, (k - cx) % 5, (l - cy) % 7]


# ---------------------------------------------------------------------------
# equilibrium and collide
# ---------------------------------------------------------------------------

def test_numpy_outer_axis_sum_is_sequential():
    # The oracle's density accumulates f0..f8 left to right; numpy's np.sum
    # over axis 0 (cavity_opt1.py:156) does the same for any grid of >1 node.
    for shape in [(9, 4, 4), (9, 64, 33), (9, 1, 9), (9, 7, 1)]:
        f = RNG.random(shape) * 10.0 ** RNG.uniform(-6, 6, size=shape)
        assert np.array_equal(np.sum(f, axis=0), o.density(f))
        f32 = f.astype(np.float32)
        assert np.array_equal(np.sum(f32, axis=0), o.density(f32))


def test_equilibrium_rest_state_is_weights():
    for dt in (np.float32, np.float64):
        feq = o.equilibrium(np.ones((3, 4), dt), np.zeros((3, 4), dt), np.zeros((3, 4), dt))
        assert feq.dtype == dt
        assert np.array_equal(feq, np.broadcast_to(o.weights(dt)[:, None, None], feq.shape))


def test_equilibrium_moments_closed_form():
    rho = 0.5 + RNG.random((6, 5))
    ux, uy = 0.2 * RNG.standard_normal((2, 6, 5))
    feq = o.equilibrium(rho, ux, uy)
    c = o.C_IC.astype(float)
    assert np.allclose(feq.sum(0), rho, rtol=1e-14, atol=0)
    assert np.allclose(np.einsum("ikl,i->kl", feq, c[:, 0]), rho * ux, rtol=1e-13, atol=1e-15)
    assert np.allclose(np.einsum("ikl,i->kl", feq, c[:, 1]), rho * uy, rtol=1e-13, atol=1e-15)
    # second moment: sum c_a c_b f_eq = rho (delta_ab / 3 + u_a u_b)
    pxx = np.einsum("ikl,i,i->kl", feq, c[:, 0], c[:, 0])
    pxy = np.einsum("ikl,i,i->kl", feq, c[:, 0], c[:, 1])
    pyy = np.einsum("ikl,i,i->kl", feq, c[:, 1], c[:, 1])
    assert np.allclose(pxx, rho * (1 / 3 + ux * ux), rtol=1e-13)
    assert np.allclose(pxy, rho * ux * uy, rtol=1e-12, atol=1e-15)
    assert np.allclose(pyy, rho * (1 / 3 + uy * uy), rtol=1e-13)


def test_equilibrium_matches_opt0_dot_form():
    # tests/02-CollideTest.py:94-102 (random 2x2 grid, 1e-7 absolute)
    rho = np.abs(RNG.random((2, 2)))
    ux, uy = RNG.random((2, 2)), RNG.random((2, 2))
    e1 = o.equilibrium(rho, ux, uy)
    e2 = o.equilibrium_opt0(rho, np.array([ux, uy]))
    assert np.abs(e1 - e2).max() < 1e-7
    assert np.abs(e1 - e2).max() < 1e-14  # both are FP64 forms of one polynomial


@pytest.mark.parametrize("omega", [0.5, 1.7])
def test_collide_matches_opt0_dot_form(omega):
    # tests/02-CollideTest.py:104-111 ((9,4,4) random, omega 0.5 and 1.7, 1e-7)
    f = np.abs(RNG.random((9, 4, 4)))
    c1, c2 = f.copy(), f.copy()
    o.collide(c1, omega)
    o.collide_opt0(c2, omega)
    assert np.abs(c1 - c2).max() < 1e-7


def test_collide_conserves_mass_and_momentum():
    f = random_populations(7, 6)
    g = f.copy()
    rho, ux, uy = o.collide(g, 1.3)
    c = o.C_IC.astype(float)
    assert np.allclose(g.sum(0), f.sum(0), rtol=1e-14)
    assert np.allclose(np.einsum("ikl,i->kl", g, c[:, 0]), np.einsum("ikl,i->kl", f, c[:, 0]),
                       rtol=1e-12, atol=1e-15)
    assert np.allclose(np.einsum("ikl,i->kl", g, c[:, 1]), np.einsum("ikl,i->kl", f, c[:, 1]),
                       rtol=1e-12, atol=1e-15)
    assert np.array_equal(rho, o.density(f))


def test_collide_omega_one_gives_equilibrium():
    f = random_populations(5, 5)
    g = f.copy()
    rho, ux, uy = o.collide(g, 1.0)
    assert np.allclose(g, o.equilibrium(rho, ux, uy), rtol=0, atol=1e-16)


def test_collide_fixed_point_is_equilibrium():
    rho = 1 + 0.1 * RNG.random((4, 4))
    ux, uy = 0.05 * RNG.standard_normal((2, 4, 4))
    feq = o.equilibrium(rho, ux, uy)
    g = feq.copy()
    o.collide(g, 1.7)
    assert np.allclose(g, feq, rtol=1e-14, atol=0)


# ---------------------------------------------------------------------------
# cavity walls: an independent brute force written from the wall physics
# ---------------------------------------------------------------------------

def brute_force_walls(f, u0):
    """Per node, per channel: pull from the source node (k - cx, l - cy); if the
    source lies outside the box the population is the node's own opposite
    population; crossing a side wall (left/right) gives plain reflection,
    crossing only the lid adds +-6 w rho_lid u0 by the sign of cx.  rho_lid
    sums the node's outgoing f6, f2, f5, then the incoming 6, 2, 5, 3, 0, 1
    (periodically pulled, bottom-reflected where they cross the bottom)."""
    _, nx, ny = f.shape
    w = o.weights(f.dtype)
    g = np.empty_like(f)

    def pulled(i, k, l):
        cx, cy = o.C_IC[i]
        if l - cy < 0:  # crosses the bottom wall
            return f[o.OPPOSITE[i], k, l]
        return f[i, (k - cx) % nx, (l - cy) % ny]

    for k in range(nx):
        l = ny - 1
        rho_lid = f[6, k, l] + f[2, k, l]
        for term in (f[5, k, l], pulled(6, k, l), pulled(2, k, l), pulled(5, k, l),
                     pulled(3, k, l), pulled(0, k, l), pulled(1, k, l)):
            rho_lid = rho_lid + term
        for l in range(ny):
            for i in range(9):
                cx, cy = o.C_IC[i]
                sx, sy = k - cx, l - cy
                side = sx < 0 or sx >= nx
                if side or sy < 0:
                    g[i, k, l] = f[o.OPPOSITE[i], k, l]
                elif sy >= ny:
                    t = 6 * w[i] * rho_lid * u0
                    if cx > 0:
                        g[i, k, l] = f[o.OPPOSITE[i], k, l] + t
                    elif cx < 0:
                        g[i, k, l] = f[o.OPPOSITE[i], k, l] - t
                    else:
                        g[i, k, l] = f[o.OPPOSITE[i], k, l]
                else:
                    g[i, k, l] = f[i, sx % nx, sy % ny]
    return g


@pytest.mark.parametrize("shape", [(4, 4), (5, 7), (8, 3), (1, 5), (6, 1), (2, 2), (1, 1)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_walls_match_brute_force(shape, dtype):
    f = random_populations(*shape, dtype=dtype)
    g = f.copy()
    o.stream_and_bounce_back(g, 0.1)
    assert np.array_equal(g, brute_force_walls(f, 0.1))


def test_walls_without_lid_are_a_permutation():
    f = RNG.random((9, 6, 5))
    g = f.copy()
    o.stream_and_bounce_back(g, 0.0)
    assert np.array_equal(np.sort(g.ravel()), np.sort(f.ravel()))


def test_walls_lid_mass_change_identity():
    # One wall step changes total mass by 6 w u0 (rho_lid(nx-1) - rho_lid(0)):
    # the SE lid term survives everywhere but the left corner, the SW term
    # everywhere but the right corner.
    f = random_populations(9, 7)
    g = f.copy()
    o.stream_and_bounce_back(g, 0.1)
    rl = o.lid_density(f)
    expect = 6 * (1 / 36) * 0.1 * (rl[-1] - rl[0])
    assert abs((g.sum() - f.sum()) - expect) < 1e-14


def test_walls_drive_flow_along_lid():
    # after a few cavity steps from rest the top row moves in +x
    f = o.cavity_init(16, 16)
    for _ in range(20):
        o.cavity_step(f, 1.2, 0.1)
    _, ux, uy = o.moments(f)
    assert ux[:, -1].mean() > 0.01
    assert abs(uy[:, 0]).max() < abs(ux[:, -1]).max()


# ---------------------------------------------------------------------------
# full steps: shear-wave decay against the analytic solution
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("direction", ["x", "y"])
def test_shear_wave_decay_matches_viscosity(direction):
    n, omega, steps = 48, 1.0, 300
    f = o.shear_wave_init(n, n, amplitude=0.05, direction=direction)
    amps = [o.shear_wave_amplitude(f, direction)]
    for _ in range(steps):
        o.periodic_step(f, omega)
        amps.append(o.shear_wave_amplitude(f, direction))
    nu = o.fit_viscosity(np.arange(steps + 1), amps, n)
    assert abs(nu / o.viscosity_theory(omega) - 1) < 2e-3
    assert abs(amps[0] - 0.05) < 1e-12


def test_fft_amplitude_of_a_pure_sine():
    # (2/n) |FFT| of A sin(2 pi j / n) peaks at A (k = +-1), for any phase of the line
    for direction, shape in (("x", (8, 64)), ("y", (48, 6))):
        f = o.shear_wave_init(*shape, amplitude=0.07, direction=direction)
        assert abs(o.shear_wave_amplitude_fft(f, direction) - 0.07) < 1e-15
        assert abs(o.shear_wave_amplitude(f, direction) - 0.07) < 1e-15
