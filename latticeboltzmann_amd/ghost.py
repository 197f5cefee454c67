"""Ghost-layer flows on the GPU: the reference's Couette, Poiseuille,
sliding-lid and obstacle-canal scripts.

The reference's ``simulators/simple_flows`` and ``experimantal_flows``
scripts keep one ghost layer inside the ``(9, nx, ny)`` array, stream it with
``np.roll``, overwrite populations next to the ghost layer and collide every
node with their own equilibrium arithmetic (``PoiseuilleFlow.py:25-53``).
``GhostLattice`` runs one such step per launch (``lbm_ghost_step_*``): the
boundary handling of a script is a table of override rules (``Rule``) and an
optional pressure-periodic pair of columns, and the kernel reproduces the
script's arithmetic bit for bit in FP64 and FP32.

Rule tables of the scripts:

* ``channel_walls``: ``PoiseuilleFlow.py:59-74`` (rest wall at the bottom, wall
  moving with uw at the top, columns 1..nx-2);
* ``cavity_walls``: ``slidingLid.py:68-91`` (four walls, lid term weighted by
  the wall density);
* ``side_walls``: ``obstacle_canal.py:307-329`` (full-length walls per side);
* ``obstacle_faces``: ``obstacle_canal.py:413-458`` (a rectangle whose four
  faces bounce back);
* ``pulse_walls``: the bounce-back tests of ``tests/simpleFlowsTest.py:433-491``
  (walls that also clear the ghost values).

Drivers with the scripts' loops: ``couette`` (``PoiseuilleFlow.py:93-113``),
``poiseuille`` (``:129-150``), ``sliding_lid`` (``slidingLid.py:94-108``),
``obstacle_canal`` (``obstacle_canal.py:265-276``, one process).
"""

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from . import kernels

_SUFFIX = {torch.float64: "f64", torch.float32: "f32"}


@dataclass(frozen=True)
class Rule:
    """Populations of channel ``ch`` at nodes x0 <= x < x1, y0 <= y < y1 take
    the streamed value of channel ``src_ch`` at (x + dx, y + dy), plus ``coef``
    (kind LID) or ``coef`` times the wall density at the source (LID_RHO), or
    0 (ZERO)."""
    x0: int
    x1: int
    y0: int
    y1: int
    ch: int
    src_ch: int = 0
    dx: int = 0
    dy: int = 0
    kind: int = nat.RULE_COPY
    coef: float = 0.0


def _lid(uw):
    """The scripts' moving-wall term ``1 / 6 * uw`` (a Python float)."""
    return 1 / 6 * uw


def channel_walls(nx, ny, uw):
    """``PoiseuilleFlow.py:59-74``: bottom ghost row y = 0, top ghost row y = ny-1,
    columns 1..nx-2; the top wall moves with uw."""
    c = _lid(uw)
    return [Rule(1, nx - 1, 1, 2, 2, 4, 0, -1),
            Rule(1, nx - 1, 1, 2, 5, 7, 0, -1),
            Rule(1, nx - 1, 1, 2, 6, 8, 0, -1),
            Rule(1, nx - 1, ny - 2, ny - 1, 4, 2, 0, 1),
            Rule(1, nx - 1, ny - 2, ny - 1, 7, 5, 0, 1, nat.RULE_LID, -c),
            Rule(1, nx - 1, ny - 2, ny - 1, 8, 6, 0, 1, nat.RULE_LID, c)]


def cavity_walls(nx, ny, uw):
    """``slidingLid.py:68-91``: walls on all four ghost layers, the lid at
    y = ny-1 moving with uw, its term scaled by the wall density."""
    c = _lid(uw)
    return [Rule(1, 2, 1, ny - 1, 1, 3, -1, 0),
            Rule(1, 2, 1, ny - 1, 5, 7, -1, 0),
            Rule(1, 2, 1, ny - 1, 8, 6, -1, 0),
            Rule(nx - 2, nx - 1, 1, ny - 1, 3, 1, 1, 0),
            Rule(nx - 2, nx - 1, 1, ny - 1, 6, 8, 1, 0),
            Rule(nx - 2, nx - 1, 1, ny - 1, 7, 5, 1, 0),
            Rule(1, nx - 1, 1, 2, 2, 4, 0, -1),
            Rule(1, nx - 1, 1, 2, 5, 7, 0, -1),
            Rule(1, nx - 1, 1, 2, 6, 8, 0, -1),
            Rule(1, nx - 1, ny - 2, ny - 1, 4, 2, 0, 1),
            Rule(1, nx - 1, ny - 2, ny - 1, 7, 5, 0, 1, nat.RULE_LID_RHO, -c),
            Rule(1, nx - 1, ny - 2, ny - 1, 8, 6, 0, 1, nat.RULE_LID_RHO, c)]


def side_walls(nx, ny, uw=0.0, left=False, right=False, bottom=True, top=True):
    """``obstacle_canal.py:307-329``: full-length walls on the chosen sides
    (right, left, bottom, top in that order); the top wall moves with uw."""
    c = _lid(uw)
    rules = []
    if right:
        rules += [Rule(nx - 2, nx - 1, 0, ny, 3, 1, 1, 0), Rule(nx - 2, nx - 1, 0, ny, 6, 8, 1, 0),
                  Rule(nx - 2, nx - 1, 0, ny, 7, 5, 1, 0)]
    if left:
        rules += [Rule(1, 2, 0, ny, 1, 3, -1, 0), Rule(1, 2, 0, ny, 5, 7, -1, 0),
                  Rule(1, 2, 0, ny, 8, 6, -1, 0)]
    if bottom:
        rules += [Rule(0, nx, 1, 2, 2, 4, 0, -1), Rule(0, nx, 1, 2, 5, 7, 0, -1),
                  Rule(0, nx, 1, 2, 6, 8, 0, -1)]
    if top:
        rules += [Rule(0, nx, ny - 2, ny - 1, 4, 2, 0, 1),
                  Rule(0, nx, ny - 2, ny - 1, 7, 5, 0, 1, nat.RULE_LID, -c),
                  Rule(0, nx, ny - 2, ny - 1, 8, 6, 0, 1, nat.RULE_LID, c)]
    return rules


def obstacle_faces(sx, sy, ex, ey):
    """``obstacle_canal.py:413-458`` for the rectangle (sx, sy)-(ex, ey), every
    face bouncing back, in the script's order and with its index choices
    (the right face writes column ex + 1 from column ex)."""
    return [Rule(sx - 1, sx, sy, ey, 3, 1, 1, 0), Rule(sx - 1, sx, sy, ey, 7, 5, 1, 0),
            Rule(sx - 1, sx, sy, ey, 6, 8, 1, 0),
            Rule(ex + 1, ex + 2, sy, ey, 1, 3, -1, 0), Rule(ex + 1, ex + 2, sy, ey, 5, 7, -1, 0),
            Rule(ex + 1, ex + 2, sy, ey, 6, 8, -1, 0),
            Rule(sx, ex, ey + 1, ey + 2, 2, 4, 0, -1), Rule(sx, ex, ey + 1, ey + 2, 5, 7, 0, -1),
            Rule(sx, ex, ey + 1, ey + 2, 6, 8, 0, -1),
            Rule(sx, ex, sy - 1, sy, 4, 2, 0, 1), Rule(sx, ex, sy - 1, sy, 7, 5, 0, 1),
            Rule(sx, ex, sy - 1, sy, 8, 6, 0, 1)]


def pulse_walls(nx, ny, vertical):
    """The walls of ``tests/simpleFlowsTest.py:433-491``: copy the ghost
    values back in, then clear them (full length, no moving wall)."""
    m = (nx if vertical else ny) - 1
    z = nat.RULE_ZERO
    if vertical:
        return [Rule(1, 2, 0, ny, 1, 3, -1, 0), Rule(1, 2, 0, ny, 5, 7, -1, 0),
                Rule(1, 2, 0, ny, 8, 6, -1, 0),
                Rule(0, 1, 0, ny, 3, kind=z), Rule(0, 1, 0, ny, 7, kind=z),
                Rule(0, 1, 0, ny, 6, kind=z),
                Rule(m - 1, m, 0, ny, 3, 1, 1, 0), Rule(m - 1, m, 0, ny, 6, 8, 1, 0),
                Rule(m - 1, m, 0, ny, 7, 5, 1, 0),
                Rule(m, m + 1, 0, ny, 1, kind=z), Rule(m, m + 1, 0, ny, 8, kind=z),
                Rule(m, m + 1, 0, ny, 5, kind=z)]
    return [Rule(0, nx, 1, 2, 2, 4, 0, -1), Rule(0, nx, 1, 2, 5, 7, 0, -1),
            Rule(0, nx, 1, 2, 6, 8, 0, -1),
            Rule(0, nx, 0, 1, 4, kind=z), Rule(0, nx, 0, 1, 7, kind=z), Rule(0, nx, 0, 1, 8, kind=z),
            Rule(0, nx, m - 1, m, 4, 2, 0, 1), Rule(0, nx, m - 1, m, 7, 5, 0, 1),
            Rule(0, nx, m - 1, m, 8, 6, 0, 1),
            Rule(0, nx, m, m + 1, 2, kind=z), Rule(0, nx, m, m + 1, 5, kind=z),
            Rule(0, nx, m, m + 1, 6, kind=z)]


class GhostLattice:
    """A ghost-layer lattice of nx x ny nodes on one GPU.

    ``rules``: override rules in order (at most 32); ``pbc``: None, or
    ``("before" | "after", rho_in, rho_out)`` for the pressure-periodic inlet
    (column 0) and outlet (column nx-1), built before streaming
    (``PoiseuilleFlow.py:76-89``) or after it (``obstacle_canal.py:362-373``).
    The state starts at rest: the scripts' equilibrium of rho = 1, u = 0.
    ``layout``: ``"rows"`` (default; f[k][i][l], the nine channel segments of
    an x-row adjacent, faster) or ``"soa"`` (the scripts' f[i][k][l]); the
    results are the same bits.  ``populations()`` and ``load()`` use the
    scripts' ``(9, nx, ny)`` either way.
    """

    def __init__(self, nx, ny, omega, rules=(), pbc=None, dtype=torch.float64, device=None,
                 layout="rows"):
        nat.require_gpu()
        if dtype not in _SUFFIX:
            raise ValueError("dtype must be torch.float64 or torch.float32")
        self.nx, self.ny, self.omega, self.dtype = int(nx), int(ny), float(omega), dtype
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.rules = list(rules)
        if len(self.rules) > nat.GHOST_MAX_RULES:
            raise ValueError("at most %d rules" % nat.GHOST_MAX_RULES)
        arr = (nat.GhostRule * max(1, len(self.rules)))()
        for k, r in enumerate(self.rules):
            arr[k] = nat.GhostRule(r.x0, r.x1, r.y0, r.y1, r.ch, r.kind, r.dx, r.dy, r.src_ch, 0,
                                   float(r.coef))
        self._rules = arr
        self._pbc = None
        if pbc is not None:
            mode, rho_in, rho_out = pbc
            modes = {"before": nat.PBC_BEFORE_STREAM, "after": nat.PBC_AFTER_STREAM}
            if mode not in modes:
                raise ValueError("pbc mode must be 'before' or 'after'")
            self._pbc = nat.GhostPbc(modes[mode], 0, float(rho_in), float(rho_out))
        if layout not in ("rows", "soa"):
            raise ValueError("layout must be 'rows' or 'soa'")
        self.layout = nat.LAYOUT_ROWS if layout == "rows" else nat.LAYOUT_SOA
        shape = (self.nx, 9, self.ny) if self.layout == nat.LAYOUT_ROWS else (9, self.nx, self.ny)
        self.f = torch.empty(shape, dtype=dtype, device=self.device)
        self._g = torch.empty_like(self.f)
        self.time = 0
        self.init_equilibrium()

    def init_equilibrium(self, rho=None, ux=None, uy=None):
        """f = the scripts' equilibrium of (rho, ux, uy); defaults rho = 1, u = 0."""
        shp = (self.nx, self.ny)

        def field(x, default):
            if x is None:
                return torch.full(shp, default, dtype=self.dtype, device=self.device)
            return torch.as_tensor(x, dtype=self.dtype, device=self.device).expand(shp).contiguous()

        r, u, v = field(rho, 1.0), field(ux, 0.0), field(uy, 0.0)
        soa = torch.empty((9, self.nx, self.ny), dtype=self.dtype, device=self.device)
        nat.call("lbm_ghost_equilibrium_" + _SUFFIX[self.dtype], nat.ptr(r), nat.ptr(u),
                 nat.ptr(v), nat.ptr(soa), self.nx * self.ny, nat.stream_ptr())
        self._from_soa(soa)
        self.time = 0
        return self

    def _from_soa(self, soa):
        if self.layout == nat.LAYOUT_SOA:
            self.f.copy_(soa)
        else:
            kernels.relayout(soa, nat.LAYOUT_SOA, nat.LAYOUT_ROWS, out=self.f)

    def load(self, f):
        """Replace the populations by ``f`` (shape (9, nx, ny))."""
        t = torch.as_tensor(np.asarray(f) if not isinstance(f, torch.Tensor) else f)
        if tuple(t.shape) != (9, self.nx, self.ny):
            raise ValueError("populations must have shape (9, %d, %d)" % (self.nx, self.ny))
        self._from_soa(t.to(device=self.device, dtype=self.dtype).contiguous())
        return self

    def _launch(self, flags, moments):
        rho = ux = uy = None
        if moments is not None:
            rho, ux, uy = moments
        nat.call("lbm_ghost_step_" + _SUFFIX[self.dtype], nat.ptr(self.f), nat.ptr(self._g),
                 self.nx, self.ny, self.layout, self.omega, ctypes.addressof(self._rules),
                 len(self.rules),
                 ctypes.addressof(self._pbc) if self._pbc is not None else None, int(flags),
                 nat.ptr(rho), nat.ptr(ux), nat.ptr(uy), nat.stream_ptr())
        self.f, self._g = self._g, self.f

    def moment_buffers(self):
        return tuple(torch.empty((self.nx, self.ny), dtype=self.dtype, device=self.device)
                     for _ in range(3))

    def step(self, n=1, moments=None):
        """n full steps (walls, then collide); ``moments`` receives the
        pre-collision moments of the last one."""
        for s in range(n):
            self._launch(0, moments if s == n - 1 else None)
        self.time += n
        return self

    def collide_only(self, moments=None):
        self._launch(nat.GHOST_COLLIDE_ONLY, moments)
        return self

    def stream_and_walls(self):
        """Streaming and rules without the collision."""
        self._launch(nat.GHOST_NO_COLLIDE, None)
        return self

    def populations_device(self):
        """Populations (9, nx, ny) as a device tensor."""
        if self.layout == nat.LAYOUT_SOA:
            return self.f.clone()
        return kernels.relayout(self.f, nat.LAYOUT_ROWS, nat.LAYOUT_SOA)

    def populations(self):
        """Populations (9, nx, ny) as a numpy array."""
        return self.populations_device().cpu().numpy()


# ---- the scripts' flows ----------------------------------------------------------

def couette(nx=100, ny=52, uw=0.1, omega=0.5, steps=4000, dtype=torch.float64, layout="rows"):
    """``PoiseuilleFlow.py:93-113``: per step moments, collide, stream, walls
    (the state after the walls is the loop's state).  Returns the lattice and
    the moments of the last collision (what the script plots)."""
    lat = GhostLattice(nx, ny, omega, channel_walls(nx, ny, uw), dtype=dtype, layout=layout)
    mom = lat.moment_buffers()
    if steps <= 0:
        return lat, None
    lat.collide_only(mom if steps == 1 else None)
    if steps > 1:
        lat.step(steps - 1, moments=mom)
    lat.stream_and_walls()
    lat.time = steps
    return lat, mom


def poiseuille(nx=102, ny=52, rho_in=1.001, rho_out=0.999, omega=0.5, steps=4000, uw=0.0,
               dtype=torch.float64, layout="rows"):
    """``PoiseuilleFlow.py:129-150``: pressure-periodic columns before
    streaming, channel walls, collide."""
    lat = GhostLattice(nx, ny, omega, channel_walls(nx, ny, uw), ("before", rho_in, rho_out),
                       dtype=dtype, layout=layout)
    mom = lat.moment_buffers()
    lat.step(steps, moments=mom)
    return lat, mom


def sliding_lid_omega(n, re, uw):
    """``slidingLid.py:28``: omega = 2 Re / (6 n uw + Re)."""
    return (2 * re) / (6 * n * uw + re)


def sliding_lid(n=300, re=1000, uw=0.1, steps=10000, dtype=torch.float64, omega=None,
                layout="rows"):
    """``slidingLid.py:94-108``: (n + 2)^2 nodes, walls in the ghost layers, lid
    at the top moving with uw."""
    om = sliding_lid_omega(n, re, uw) if omega is None else omega
    lat = GhostLattice(n + 2, n + 2, om, cavity_walls(n + 2, n + 2, uw), dtype=dtype,
                       layout=layout)
    mom = lat.moment_buffers()
    lat.step(steps, moments=mom)
    return lat, mom


def obstacle_canal(nx=100, ny=50, obstacle=(50, 24, 53, 27), rho_diff=0.01, omega=0.5, steps=1000,
                   uw=0.0, dtype=torch.float64, layout="rows"):
    """``obstacle_canal.py:265-276`` on one process: (nx + 2) x (ny + 2)
    nodes, pressure-periodic left / right after streaming, walls at the top and
    bottom, an obstacle (sx, sy, ex, ey) with bouncing faces."""
    sx, sy, ex, ey = obstacle
    gx, gy = nx + 2, ny + 2
    if not (2 <= sx < ex and ex + 1 < gx - 1 and 3 <= sy < ey and ey + 1 < gy - 2):
        raise ValueError("the obstacle must lie inside the canal, clear of the walls")
    rules = side_walls(gx, gy, uw) + obstacle_faces(sx, sy, ex, ey)
    lat = GhostLattice(gx, gy, omega, rules, ("after", 1 + rho_diff, 1 - rho_diff), dtype=dtype,
                       layout=layout)
    mom = lat.moment_buffers()
    lat.step(steps, moments=mom)
    return lat, mom


def couette_profile(ny_fluid, uw):
    """Steady Couette velocity at the fluid rows between half-way walls:
    uw (j + 1/2) / ny_fluid (the script plots uw j / ny_fluid,
    ``PoiseuilleFlow.py:118-119``, half a row off)."""
    return uw * (np.arange(ny_fluid) + 0.5) / ny_fluid


def poiseuille_profile(nx_fluid, ny_fluid, diff, omega):
    """``PoiseuilleFlow.py:152-154``: delta y (L - y) / 3 at y = j + 1/2 with
    delta = 2 diff / nx / nu / 2 and nu = (1/omega - 1/2) / 3."""
    nu = (1 / omega - 0.5) / 3
    delta = 2.0 * diff / nx_fluid / nu / 2.0
    y = np.arange(ny_fluid) + 0.5
    return delta * y * (ny_fluid - y) / 3.0
