"""Device-resident D2Q9 solver: populations stay in HBM across steps.

This is the engine behind the reference's driver loops
(``cavity_opt1.py:338-347``, ``shear_wave_opt1.py:170-174``): two population
buffers ping-pong through the fused pull stream-collide kernel, one launch
per step, on one stream.  By default the populations are kept in the
row-interleaved layout ``f[k][i][l]`` (nine channel segments of an x-row
adjacent in HBM), which streams faster than the reference's ``f[i][k][l]``;
``populations()`` and ``load()`` convert to and from the reference layout,
and the arithmetic (hence every bit of the state) is the same in both.

The MPI decomposition and halo exchange of the reference
(``cavity_opt1.py:247-329``) have no counterpart: one MI355X holds the whole
domain.
"""

import math

import numpy as np
import torch

from . import _native as nat
from . import kernels

_TORCH_DTYPES = {"float64": torch.float64, "float32": torch.float32,
                 np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                 torch.float64: torch.float64, torch.float32: torch.float32}

BCS = {"periodic": nat.BC_PERIODIC, "cavity": nat.BC_CAVITY}


def omega_for_cavity(n, reynolds, u_lid):
    """omega = 1 / (3 nu + 1/2) with nu = U N / Re (lattice units)."""
    nu = u_lid * n / reynolds
    return 1.0 / (3.0 * nu + 0.5)


class Lattice:
    """A D2Q9 BGK lattice of ``nx x ny`` nodes resident on one GPU.

    Parameters
    ----------
    nx, ny : grid size (x slow, y fast, as the reference's ``f_ikl``).
    omega : relaxation parameter.
    bc : ``"cavity"`` (walls + moving lid at y = ny-1) or ``"periodic"``.
    u0 : lid velocity (cavity only).
    dtype : float64 or float32 storage and arithmetic.
    variant : launch shape of the fused step (-1 = library default, "auto" =
        autotuned; "auto" also tunes the multi-step kernel's shape).
    layout : ``"rows"`` (default) or ``"soa"`` (the reference's f[i][k][l]).
    multi_step : advance runs of K steps with the temporally blocked
        multi-step kernel (``lbm_step_multi_*``, K = 2..4 per launch, bitwise
        equal to K single steps) where the grid is large enough for its tiles.
    """

    def __init__(self, nx, ny, omega, bc="cavity", u0=0.1, dtype=torch.float64, device=None,
                 variant=-1, layout="rows", multi_step=True):
        nat.require_gpu()
        if bc not in BCS:
            raise ValueError("bc must be 'cavity' or 'periodic'")
        self.nx, self.ny = int(nx), int(ny)
        self.omega, self.bc, self.u0 = float(omega), bc, float(u0)
        self.dtype = _TORCH_DTYPES[dtype]
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if layout not in ("rows", "soa"):
            raise ValueError("layout must be 'rows' or 'soa'")
        self.layout = nat.LAYOUT_ROWS if layout == "rows" else nat.LAYOUT_SOA
        self.multi_variant, self.multi_k = None, 1
        if multi_step:
            if variant == "auto":
                from .autotune import multi_variant
                self.set_multi_variant(multi_variant(self.nx, self.ny, self.dtype, self.layout,
                                                     BCS[bc]))
            else:
                self.set_multi_variant(-1)
        if variant == "auto":
            from .autotune import step_variant
            variant = step_variant(self.nx, self.ny, self.dtype, self.layout, BCS[bc])
        self.variant = int(variant)
        if self.layout == nat.LAYOUT_ROWS:
            shape = (self.nx, 9, self.ny)
        else:
            shape = (9, self.nx, self.ny)
        self.f = torch.empty(shape, dtype=self.dtype, device=self.device)
        self._g = torch.empty(shape, dtype=self.dtype, device=self.device)
        self.time = 0
        self.init_equilibrium()

    # -- state ---------------------------------------------------------------
    def init_equilibrium(self, rho=None, ux=None, uy=None):
        """f = f_eq(rho, u); defaults rho = 1, u = 0 (``cavity_opt1.py:333-335``)."""
        shp = (self.nx, self.ny)

        def field(x, default):
            if x is None:
                return torch.full(shp, default, dtype=self.dtype, device=self.device)
            t = torch.as_tensor(x, dtype=self.dtype, device=self.device)
            return t.expand(shp).contiguous()

        soa = kernels.equilibrium(field(rho, 1.0), field(ux, 0.0), field(uy, 0.0))
        self._from_soa(soa)
        self.time = 0
        return self

    def _from_soa(self, soa):
        if self.layout == nat.LAYOUT_SOA:
            self.f.copy_(soa)
        else:
            kernels.relayout(soa, nat.LAYOUT_SOA, nat.LAYOUT_ROWS, out=self.f)

    def load(self, f):
        """Replace the populations by ``f`` (reference layout, shape (9, nx, ny))."""
        t = torch.as_tensor(np.asarray(f) if not isinstance(f, torch.Tensor) else f)
        if tuple(t.shape) != (9, self.nx, self.ny):
            raise ValueError("populations must have shape (9, %d, %d)" % (self.nx, self.ny))
        self._from_soa(t.to(device=self.device, dtype=self.dtype).contiguous())
        return self

    def populations_device(self):
        """Populations in the reference layout (9, nx, ny) as a device tensor."""
        if self.layout == nat.LAYOUT_SOA:
            return self.f.clone()
        return kernels.relayout(self.f, nat.LAYOUT_ROWS, nat.LAYOUT_SOA)

    def populations(self):
        """Populations in the reference layout (9, nx, ny) as a numpy array."""
        return self.populations_device().cpu().numpy()

    # -- stepping ------------------------------------------------------------
    def set_multi_variant(self, variant):
        """Use multi-step variant ``variant`` (None: single steps only); a
        variant whose tiles do not fit the grid is ignored."""
        self.multi_variant, self.multi_k = None, 1
        if variant is None:
            return
        if variant < 0:
            lib = nat.load()
            pick = (lib.lbm_step_multi_default_for_f32 if self.dtype == torch.float32
                    else lib.lbm_step_multi_default_for)
            variant = pick(self.nx, self.ny)
            if variant < 0:
                variant = nat.load().lbm_step_multi_default_variant()
        k, mnx, mny = kernels.step_multi_info(variant)
        if self.nx >= mnx and self.ny >= mny:
            self.multi_variant, self.multi_k = int(variant), k

    def _runs(self, n):
        return n // self.multi_k if self.multi_variant is not None else 0

    def launches(self, n):
        """Kernel launches (buffer swaps) that ``step(n)`` makes."""
        q = self._runs(n)
        return q + (n - q * self.multi_k)

    def step(self, n=1, moments=None):
        """Advance ``n`` steps.  ``moments=(rho, ux, uy)`` tensors receive the
        pre-collision moments of the last step.  Runs of K steps go through the
        multi-step kernel when enabled; the bits are those of n single steps."""
        bc = BCS[self.bc]
        runs = self._runs(n)
        single = n - runs * self.multi_k
        for p in range(runs):
            mom = moments if (moments is not None and p == runs - 1 and single == 0) else None
            kernels.step_multi(self.f, self._g, self.omega, bc, self.u0, moments=mom,
                               variant=self.multi_variant, layout=self.layout)
            self.f, self._g = self._g, self.f
        for s in range(single):
            mom = moments if (moments is not None and s == single - 1) else None
            kernels.step(self.f, self._g, self.omega, bc, self.u0, moments=mom,
                         variant=self.variant, layout=self.layout)
            self.f, self._g = self._g, self.f
        self.time += n
        return self

    # -- observables ---------------------------------------------------------
    def moments(self):
        """(rho, ux, uy) of the current populations, as device tensors."""
        return kernels.moments(self.f, layout=self.layout)

    def mass(self):
        """Total mass sum_i,k,l f, accumulated in FP64 on the device."""
        return float(kernels.total_sum(self.f).item())

    def velocity(self):
        """(ux, uy) as numpy arrays."""
        _, ux, uy = self.moments()
        return ux.cpu().numpy(), uy.cpu().numpy()

    def run_until_steady(self, tol=1e-6, every=1000, max_steps=10**7, u_ref=None, graph=True):
        """Step until max|u(t + every) - u(t)| / u_ref < tol (FP64 only).

        ``u_ref`` defaults to the lid velocity.  With ``graph`` (and an even
        number of launches per ``every`` steps) the steps between two residual
        checks are captured once into a HIP graph and replayed, which removes the
        per-launch overhead that dominates small grids.  Returns the number of
        steps taken and the last residual.
        """
        u_ref = u_ref or self.u0
        uxp = torch.zeros((self.nx, self.ny), dtype=torch.float64, device=self.device)
        uyp = torch.zeros_like(uxp)
        kernels.velocity_residual(self.f, uxp, uyp, layout=self.layout)
        taken, res = 0, float("inf")
        with StepGraph(self, every, enabled=graph) as advance:
            while taken < max_steps:
                advance()
                taken += every
                res = float(kernels.velocity_residual(self.f, uxp, uyp,
                                                      layout=self.layout).item()) / u_ref
                if res < tol:
                    break
                if math.isinf(res) or math.isnan(res):
                    raise FloatingPointError("lattice diverged: non-finite velocity residual "
                                             "after %d steps" % taken)
        return taken, res


class StepGraph:
    """Context manager returning a callable that advances ``lat`` by ``n`` steps.

    Enabled, the launches of n steps (an even number of them, so the state
    ends in the buffer it started in; otherwise the steps run eagerly) are
    captured once into a HIP graph on a side stream (capture needs a
    stream of its own) and each call replays it on the caller's current
    stream, in order with the work around it; disabled, each call launches
    the steps eagerly.  Replaying on the capture stream instead needs two
    cross-stream waits per replay; measured on a 256^2 grid, such replays ran
    at 4.6 us per step against 4.2 us on the caller's stream.
    """

    def __init__(self, lat, n, enabled=True):
        self.lat, self.n = lat, int(n)
        swaps = lat.launches(self.n) if hasattr(lat, "launches") else self.n
        self.enabled = bool(enabled) and swaps % 2 == 0
        self.graph = None
        self._state_ptr = None

    def _state(self):
        """(current buffer, other buffer) attribute names of the lattice."""
        return ("d", "_e") if hasattr(self.lat, "_e") else ("f", "_g")

    def _realign(self):
        """The graph reads and writes the buffers it was captured with.  An odd
        number of eager launches since the capture (``lat.step`` with a count
        that is not a multiple of the replayed one) leaves the state in the
        other buffer: copy it back and swap the lattice's references, so the
        replay steps the current state."""
        cur, other = self._state()
        a, b = getattr(self.lat, cur), getattr(self.lat, other)
        if a.data_ptr() == self._state_ptr:
            return
        if b.data_ptr() != self._state_ptr:
            raise RuntimeError("StepGraph: the lattice's buffers changed since the capture")
        b.copy_(a)
        setattr(self.lat, cur, b)
        setattr(self.lat, other, a)

    def __enter__(self):
        if not self.enabled:
            return lambda: self.lat.step(self.n)
        import ctypes
        lib = nat.load()
        self.side = torch.cuda.Stream(device=self.lat.device)
        self.side.wait_stream(torch.cuda.current_stream(self.lat.device))
        g = ctypes.c_void_p()
        self._state_ptr = getattr(self.lat, self._state()[0]).data_ptr()
        with torch.cuda.stream(self.side):
            nat.call("lbm_graph_begin", nat.stream_ptr())
            try:
                t0 = self.lat.time
                self.lat.step(self.n)
                self.lat.time = t0
            finally:
                status = lib.lbm_graph_end(nat.stream_ptr(), ctypes.byref(g))
            nat.check(status, "lbm_graph_end")
        self.graph = g

        def replay():
            self._realign()
            nat.call("lbm_graph_launch", self.graph, nat.stream_ptr())
            self.lat.time += self.n
        return replay

    def __exit__(self, *exc):
        if self.graph is not None:
            torch.cuda.synchronize(self.lat.device)
            nat.check(nat.load().lbm_graph_destroy(self.graph), "lbm_graph_destroy")
            self.graph = None
        return False
