"""Pins the ghost-layer oracle (oracle/ghost_ref.py) without the reference:
closed forms, the reference's bounce-back KATs, reflected pulses at the
obstacle faces, and the analytic Couette and Poiseuille profiles."""

import numpy as np
import pytest

from oracle import ghost_ref as g

RNG = np.random.default_rng(99)
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def test_equilibrium_is_the_textbook_form():
    rho = 1 + 0.1 * RNG.standard_normal(200)
    ux, uy = 0.1 * RNG.standard_normal(200), 0.1 * RNG.standard_normal(200)
    e = g.equilibrium(rho, ux, uy)
    cu = g.C[:, 0, None] * ux + g.C[:, 1, None] * uy
    tb = W[:, None] * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux ** 2 + uy ** 2))
    assert np.abs(e - tb).max() < 1e-15
    np.testing.assert_allclose(e.sum(0), rho, rtol=1e-15)
    np.testing.assert_allclose((g.C[:, 0, None] * e).sum(0), rho * ux, atol=1e-16)
    np.testing.assert_allclose((g.C[:, 1, None] * e).sum(0), rho * uy, atol=1e-16)


def test_moments_and_collide():
    f = g.at_rest(6, 5) * (1 + 0.1 * RNG.standard_normal((9, 6, 5)))
    rho, ux, uy = g.moments(f)
    assert np.allclose(rho, f.sum(0)) and np.allclose(ux * rho, (g.C[:, 0, None, None] * f).sum(0))
    h = f.copy()
    g.collide(h, 1.0)
    assert np.abs(h - g.equilibrium(rho, ux, uy)).max() < 1e-16  # omega = 1 -> f_eq


def test_stream_is_the_roll_of_the_channel_table():
    from oracle import d2q9_numpy as o
    f = RNG.standard_normal((9, 7, 5))
    a, b = f.copy(), f.copy()
    g.stream(a)
    o.stream(b)
    assert np.array_equal(a, b)
    assert np.array_equal(g.C, o.C_IC)


@pytest.mark.parametrize("vertical", [True, False])
def test_reference_pulse_kats(vertical):
    # tests/simpleFlowsTest.py:433-491: after 14 steps the pulse is back
    f = np.zeros((9, 9, 9))
    if vertical:
        f[1, 1, 1:8] = 1
    else:
        f[2, 1:8, 1] = 1
    g.pulse_walls(f, 14, vertical)
    got = f[1, 1, 1:7] if vertical else f[2, 1:7, 1]
    assert np.all(got == 1)


OB = (6, 6, 9, 9)  # sx, sy, ex, ey in a 16 x 16 grid


@pytest.mark.parametrize("start,end,steps", [
    ((1, 2, 7), (3, 2, 7), 7),    # E pulse hits the left face, comes back W
    ((3, 13, 7), (1, 13, 7), 7),  # W pulse hits the right face (ex), comes back E from ex + 1
    ((4, 7, 13), (2, 7, 13), 7),  # S pulse hits the top face, comes back N
    ((2, 7, 2), (4, 7, 2), 7),    # N pulse hits the bottom face, comes back S
    ((5, 3, 4), (7, 3, 5), 5),    # NE pulse hits the left face, comes back SW
])
def test_obstacle_faces_reflect(start, end, steps):
    f = np.zeros((9, 16, 16))
    f[start] = 1.0
    for _ in range(steps):
        g.stream(f)
        g.apply_obstacle(f, *OB)
    assert f[end] == 1.0


def test_lid_term_uses_the_wall_density():
    # at rest the wall density 2 (f2 + f5 + f6) + f0 + f1 + f3 is 1: the lid
    # rule subtracts uw / 6 from f5 (SW <- NE) and adds it to f6 (SE <- NW)
    f = g.at_rest(8, 8)
    h = f.copy()
    g.stream(h)
    g.bounce_back_cavity(h, 0.06)
    np.testing.assert_allclose(h[7, 1:-1, -2], f[5, 1:-1, -1] - 0.01, rtol=0, atol=1e-16)
    np.testing.assert_allclose(h[8, 1:-1, -2], f[6, 1:-1, -1] + 0.01, rtol=0, atol=1e-16)


def test_couette_profile():
    # 60 columns: the script leaves columns 0 and nx-1 without walls, which
    # disturbs a narrow channel; the middle column is the script's probe
    G, (rho, ux, uy) = g.couette(60, 22, 0.1, 0.5, 6000)
    exact = g.couette_profile(20, 0.1)
    assert np.abs(ux[30, 1:-1] - exact).max() < 2e-4


def test_poiseuille_profile():
    G, (rho, ux, uy) = g.poiseuille(42, 22, 1.001, 0.999, 0.5, 6000)
    ana = g.poiseuille_profile(40, 20, 0.001, 0.5)
    err = np.abs(ux[21, 1:-1] - ana).max() / ana.max()
    assert err < 0.02
    # the mirrored pressure drive: swapping inlet and outlet reverses the flow
    G, (rho, ux2, uy2) = g.poiseuille(42, 22, 0.999, 1.001, 0.5, 200)
    G, (rho, ux1, uy1) = g.poiseuille(42, 22, 1.001, 0.999, 0.5, 200)
    assert np.all(ux1[21, 1:-1] > 0) and np.allclose(ux2[21, 1:-1], -ux1[21, 1:-1], rtol=1e-2)


def test_obstacle_faces_index_encoded_kat():
    # Every population holds a code of its own (channel, x, y); after the
    # obstacle update each written cell must hold the code of the source cell
    # that the script's assignment names, and nothing else may change.  The
    # table is the script's twelve assignments as data (destination channel,
    # column, rows <- source channel, column, rows), read off
    # obstacle_canal.py:421-458, including the right and top faces writing NW
    # (6) from SE (8) where a bounce-back pair would take SW or NE.
    nx, ny, sx, sy, ex, ey = 12, 10, 4, 3, 7, 6
    c, x, y = np.meshgrid(np.arange(9), np.arange(nx), np.arange(ny), indexing="ij")
    g0 = (10000 * c + 100 * x + y).astype(np.float64)
    rows, cols = slice(sy, ey), slice(sx, ex)
    table = [
        ((3, sx - 1, rows), (1, sx, rows)),     # :421-422
        ((7, sx - 1, rows), (5, sx, rows)),     # :424-425
        ((6, sx - 1, rows), (8, sx, rows)),     # :427-428
        ((1, ex + 1, rows), (3, ex, rows)),     # :431-432
        ((5, ex + 1, rows), (7, ex, rows)),     # :434-435
        ((6, ex + 1, rows), (8, ex, rows)),     # :437-438
        ((2, cols, ey + 1), (4, cols, ey)),     # :441-442
        ((5, cols, ey + 1), (7, cols, ey)),     # :444-445
        ((6, cols, ey + 1), (8, cols, ey)),     # :447-448
        ((4, cols, sy - 1), (2, cols, sy)),     # :451-452
        ((7, cols, sy - 1), (5, cols, sy)),     # :454-455
        ((8, cols, sy - 1), (6, cols, sy)),     # :457-458
    ]
    expected = g0.copy()
    for dst, src in table:
        expected[dst] = g0[src]
    got = g0.copy()
    g.apply_obstacle(got, sx, sy, ex, ey)
    assert np.array_equal(got, expected)
    # the written cells are exactly the table's destinations
    changed = np.argwhere(got != g0)
    assert len(changed) == 3 * 2 * (ey - sy) + 3 * 2 * (ex - sx)
