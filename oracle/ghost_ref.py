"""numpy restatement of the reference's ghost-layer flows (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` and never by the product
package.

The reference's ``simulators/simple_flows`` and ``experimantal_flows``
scripts keep one ghost layer around the fluid, stream with ``np.roll`` over
the whole array, then overwrite selected populations next to the ghost layer
(bounce-back, moving wall), and collide every node, ghosts included, with
their own equilibrium arithmetic.  This module restates, keeping the
evaluation order of every expression (float64 and float32 alike):

* ``equilibrium``: ``PoiseuilleFlow.py:25-43`` (identical in
  ``slidingLid.py:37-55`` and ``obstacle_canal.py:204-221``);
* ``moments``: ``PoiseuilleFlow.py:49-53``;
* ``collide``: ``PoiseuilleFlow.py:45-47``;
* ``stream``: ``PoiseuilleFlow.py:55-57``;
* ``bounce_back_channel``: ``PoiseuilleFlow.py:59-74`` (bottom wall at rest,
  top wall moving with uw, x in 1..nx-2);
* ``pressure_periodic``: ``PoiseuilleFlow.py:76-89`` (before streaming);
* ``bounce_back_cavity``: ``slidingLid.py:68-91`` (four walls, lid term
  weighted by the wall density);
* ``bounce_back_walls``, ``pressure_periodic_after``, ``apply_obstacle``:
  ``obstacle_canal.py:307-329, 362-373, 413-458`` on one process;
* the main loops of ``PoiseuilleFlow.py:93-113`` (Couette),
  ``:129-150`` (Poiseuille), ``slidingLid.py:94-108`` and
  ``obstacle_canal.py:265-276``.

Pinned in ``tests/test_oracle_ghost.py`` by the equilibrium's closed-form
moments and its agreement with the textbook form; by the reference's own
bounce-back KATs (``tests/simpleFlowsTest.py:433-491``: a pulse returns to its
start after 14 steps); by pulses that each obstacle face sends back; by the
lid term of a state at rest (wall density 1); and by the analytic Couette and
Poiseuille profiles (``PoiseuilleFlow.py:118-119, 152-154``) and the reversal
of the Poiseuille flow when inlet and outlet densities swap.  The obstacle's
right face writes NW from SE (``obstacle_canal.py:437-438``) where a
bounce-back would write SE from NW; that rule is restated as written, no pulse
test covers it, and an index-encoded known-answer table of the script's twelve
assignments (``obstacle_canal.py:421-458``) checks every cell the obstacle
update writes.
"""

import numpy as np

C = np.array([[0, 1, 0, -1, 0, 1, -1, -1, 1],
              [0, 0, 1, 0, -1, 1, 1, -1, -1]]).T


def equilibrium(rho, ux, uy):
    p3 = 3 * (ux + uy)
    m3 = 3 * (ux - uy)
    uu = 3 * (ux * ux + uy * uy)
    x6 = 6 * ux
    y6 = 6 * uy
    xx9 = 9 * ux * ux
    yy9 = 9 * uy * uy
    xy9 = 9 * ux * uy
    return np.array([(2 * rho / 9) * (2 - uu),
                     (rho / 18) * (2 + x6 + xx9 - uu),
                     (rho / 18) * (2 + y6 + yy9 - uu),
                     (rho / 18) * (2 - x6 + xx9 - uu),
                     (rho / 18) * (2 - y6 + yy9 - uu),
                     (rho / 36) * (1 + p3 + xy9 + uu),
                     (rho / 36) * (1 - m3 - xy9 + uu),
                     (rho / 36) * (1 - p3 + xy9 + uu),
                     (rho / 36) * (1 + m3 - xy9 + uu)])


def moments(g):
    rho = np.sum(g, axis=0)
    ux = ((g[1] + g[5] + g[8]) - (g[3] + g[6] + g[7])) / rho
    uy = ((g[2] + g[5] + g[6]) - (g[4] + g[7] + g[8])) / rho
    return rho, ux, uy


def collide(g, omega):
    """In place; returns the pre-collision moments."""
    rho, ux, uy = moments(g)
    g -= omega * (g - equilibrium(rho, ux, uy))
    return rho, ux, uy


def stream(g):
    for i in range(1, 9):
        g[i] = np.roll(g[i], C[i], axis=(0, 1))


def bounce_back_channel(g, uw):
    g[2, 1:-1, 1] = g[4, 1:-1, 0]
    g[5, 1:-1, 1] = g[7, 1:-1, 0]
    g[6, 1:-1, 1] = g[8, 1:-1, 0]
    g[4, 1:-1, -2] = g[2, 1:-1, -1]
    g[7, 1:-1, -2] = g[5, 1:-1, -1] - 1 / 6 * uw
    g[8, 1:-1, -2] = g[6, 1:-1, -1] + 1 / 6 * uw


def pressure_periodic(g, rho_in, rho_out):
    rho, ux, uy = moments(g)
    eq = equilibrium(rho, ux, uy)
    eq_in = equilibrium(rho_in, ux[-2, :], uy[-2, :])
    g[:, 0, :] = eq_in + (g[:, -2, :] - eq[:, -2, :])
    eq_out = equilibrium(rho_out, ux[1, :], uy[1, :])
    g[:, -1, :] = eq_out + (g[:, 1, :] - eq[:, 1, :])


def bounce_back_cavity(g, uw):
    g[1, 1, 1:-1] = g[3, 0, 1:-1]
    g[5, 1, 1:-1] = g[7, 0, 1:-1]
    g[8, 1, 1:-1] = g[6, 0, 1:-1]
    g[3, -2, 1:-1] = g[1, -1, 1:-1]
    g[6, -2, 1:-1] = g[8, -1, 1:-1]
    g[7, -2, 1:-1] = g[5, -1, 1:-1]
    g[2, 1:-1, 1] = g[4, 1:-1, 0]
    g[5, 1:-1, 1] = g[7, 1:-1, 0]
    g[6, 1:-1, 1] = g[8, 1:-1, 0]
    rho_wall = 2.0 * (g[2, 1:-1, -1] + g[5, 1:-1, -1] + g[6, 1:-1, -1]) + \
        g[0, 1:-1, -1] + g[1, 1:-1, -1] + g[3, 1:-1, -1]
    g[4, 1:-1, -2] = g[2, 1:-1, -1]
    g[7, 1:-1, -2] = g[5, 1:-1, -1] - 1 / 6 * uw * rho_wall
    g[8, 1:-1, -2] = g[6, 1:-1, -1] + 1 / 6 * uw * rho_wall


def bounce_back_walls(g, uw, left=False, right=False, bottom=True, top=True):
    """``obstacle_canal.py:307-329``: full-length wall rules per side."""
    if right:
        g[3, -2, :] = g[1, -1, :]
        g[6, -2, :] = g[8, -1, :]
        g[7, -2, :] = g[5, -1, :]
    if left:
        g[1, 1, :] = g[3, 0, :]
        g[5, 1, :] = g[7, 0, :]
        g[8, 1, :] = g[6, 0, :]
    if bottom:
        g[2, :, 1] = g[4, :, 0]
        g[5, :, 1] = g[7, :, 0]
        g[6, :, 1] = g[8, :, 0]
    if top:
        g[4, :, -2] = g[2, :, -1]
        g[7, :, -2] = g[5, :, -1] - 1 / 6 * uw
        g[8, :, -2] = g[6, :, -1] + 1 / 6 * uw


def pressure_periodic_after(g, rho_in, rho_out):
    """``obstacle_canal.py:362-373`` (x direction), applied after streaming."""
    rho, ux, uy = moments(g)
    eq = equilibrium(rho, ux, uy)
    eq_in = equilibrium(rho_in, ux[-2, :], uy[-2, :])
    g[:, 0, :] = eq_in + g[:, -2, :] - eq[:, -2, :]
    eq_out = equilibrium(rho_out, ux[1, :], uy[1, :])
    g[:, -1, :] = eq_out + g[:, 1, :] - eq[:, 1, :]


def apply_obstacle(g, sx, sy, ex, ey):
    """``obstacle_canal.py:413-458`` with every face bouncing back."""
    g[3, sx - 1, sy:ey] = g[1, sx, sy:ey]
    g[7, sx - 1, sy:ey] = g[5, sx, sy:ey]
    g[6, sx - 1, sy:ey] = g[8, sx, sy:ey]
    g[1, ex + 1, sy:ey] = g[3, ex, sy:ey]
    g[5, ex + 1, sy:ey] = g[7, ex, sy:ey]
    g[6, ex + 1, sy:ey] = g[8, ex, sy:ey]
    g[2, sx:ex, ey + 1] = g[4, sx:ex, ey]
    g[5, sx:ex, ey + 1] = g[7, sx:ex, ey]
    g[6, sx:ex, ey + 1] = g[8, sx:ex, ey]
    g[4, sx:ex, sy - 1] = g[2, sx:ex, sy]
    g[7, sx:ex, sy - 1] = g[5, sx:ex, sy]
    g[8, sx:ex, sy - 1] = g[6, sx:ex, sy]


def at_rest(nx, ny, dtype=np.float64):
    return equilibrium(np.ones((nx, ny), dtype), np.zeros((nx, ny), dtype),
                       np.zeros((nx, ny), dtype))


# ---- main loops -------------------------------------------------------------

def couette(nx, ny, uw, omega, steps, dtype=np.float64):
    """``PoiseuilleFlow.py:99-113``: grid (nx, ny) including the two ghost
    rows; per step moments, collide, stream, walls.  Returns the grid and the
    last moments (taken before the last collide)."""
    g = at_rest(nx, ny, dtype)
    mom = None
    for _ in range(steps):
        mom = moments(g)
        g -= omega * (g - equilibrium(*mom))
        stream(g)
        bounce_back_channel(g, uw)
    return g, mom


def poiseuille(nx, ny, rho_in, rho_out, omega, steps, uw=0.0, dtype=np.float64):
    """``PoiseuilleFlow.py:136-150``: per step pressure-periodic columns,
    stream, walls, moments, collide."""
    g = at_rest(nx, ny, dtype)
    mom = None
    for _ in range(steps):
        pressure_periodic(g, rho_in, rho_out)
        stream(g)
        bounce_back_channel(g, uw)
        mom = collide(g, omega)
    return g, mom


def sliding_lid(nx, ny, uw, omega, steps, dtype=np.float64):
    """``slidingLid.py:98-108``: per step stream, walls and lid, collide."""
    g = at_rest(nx, ny, dtype)
    mom = None
    for _ in range(steps):
        stream(g)
        bounce_back_cavity(g, uw)
        mom = collide(g, omega)
    return g, mom


def obstacle_canal(nx, ny, rho_in, rho_out, omega, steps, obstacle, uw=0.0, dtype=np.float64):
    """``obstacle_canal.py:265-276`` on one process, left and right periodic
    with the pressure step, top and bottom bouncing back, every obstacle face
    bouncing back; ``obstacle = (sx, sy, ex, ey)``."""
    g = at_rest(nx, ny, dtype)
    sx, sy, ex, ey = obstacle
    mom = None
    for _ in range(steps):
        stream(g)
        pressure_periodic_after(g, rho_in, rho_out)
        bounce_back_walls(g, uw)
        apply_obstacle(g, sx, sy, ex, ey)
        mom = collide(g, omega)
    return g, mom


def pulse_walls(g, steps, vertical):
    """The bounce-back KATs of ``tests/simpleFlowsTest.py:433-491``: stream,
    copy the ghost values back in and clear them, no collision."""
    m = g.shape[1] - 1
    for _ in range(steps):
        stream(g)
        if vertical:
            g[1, 1, :] = g[3, 0, :]
            g[5, 1, :] = g[7, 0, :]
            g[8, 1, :] = g[6, 0, :]
            g[3, 0, :] = 0
            g[7, 0, :] = 0
            g[6, 0, :] = 0
            g[3, m - 1, :] = g[1, m, :]
            g[6, m - 1, :] = g[8, m, :]
            g[7, m - 1, :] = g[5, m, :]
            g[1, m, :] = 0
            g[8, m, :] = 0
            g[5, m, :] = 0
        else:
            g[2, :, 1] = g[4, :, 0]
            g[5, :, 1] = g[7, :, 0]
            g[6, :, 1] = g[8, :, 0]
            g[4, :, 0] = 0
            g[7, :, 0] = 0
            g[8, :, 0] = 0
            g[4, :, m - 1] = g[2, :, m]
            g[7, :, m - 1] = g[5, :, m]
            g[8, :, m - 1] = g[6, :, m]
            g[2, :, m] = 0
            g[5, :, m] = 0
            g[6, :, m] = 0
    return g


# ---- analytic profiles --------------------------------------------------------

def couette_profile(ny_fluid, uw):
    """Steady Couette profile between half-way walls, uw (j + 1/2) / ny_fluid
    at fluid row j (``PoiseuilleFlow.py:118-119`` plots uw j / ny_fluid, half a
    row off)."""
    return uw * (np.arange(ny_fluid) + 0.5) / ny_fluid


def poiseuille_profile(nx_fluid, ny_fluid, diff, omega):
    """Parabola of ``PoiseuilleFlow.py:152-154`` at the fluid rows:
    delta * y (L - y) / 3 with y = j + 1/2, delta = 2 diff / nx / nu / 2."""
    nu = (1 / omega - 0.5) / 3
    delta = 2.0 * diff / nx_fluid / nu / 2.0
    y = np.arange(ny_fluid) + 0.5
    return delta * y * (ny_fluid - y) / 3.0
