"""Warm-started lid-driven cavity on the GPU.

``Predictor`` wraps a trained ``Surrogate`` (weights in an ``.npz``);
``warm_populations`` turns a prediction into FP64 populations with the HIP
initializer; ``CavityStepper`` is the stepper the ``Controller`` drives: a
FP64 ``Lattice`` whose ``every`` steps between residual checks are replayed
from one HIP graph.
"""

import math
import os
import time

import numpy as np
import torch

from .. import _native as nat
from .. import kernels
from ..lattice import Lattice, StepGraph, omega_for_cavity
from .controller import WarmStartUnavailable
from .features import GRID, RE_RANGE, U_RANGE, features, in_channels
from .network import Surrogate

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(_HERE, "surrogate_weights.npz")
# the network for the upper Reynolds range (trained on cavities with N of 512
# and 1024, where BGK stays stable up to Re 5000); load_default() averages the
# networks whose ranges cover a cavity
HIGH_RE_WEIGHTS = os.path.join(_HERE, "surrogate_weights_high_re.npz")
# a second network for the lower Reynolds range, fine-tuned from the default
# one on cavities with N of 512 and 1024; load_default() averages the two
LOW_RE_B_WEIGHTS = os.path.join(_HERE, "surrogate_weights_low_re_b.npz")
ANY_N = (1, 1 << 30)


class OutOfRange(WarmStartUnavailable, ValueError):
    """(Re, U) outside the ranges the surrogate was trained on."""


class Predictor:
    """A trained surrogate plus the normalization it was trained with."""

    def __init__(self, net, grid=GRID, re_range=RE_RANGE, u_range=U_RANGE, pool=2, octaves=0,
                 n_range=ANY_N):
        self.net, self.grid, self.pool, self.octaves = net, int(grid), int(pool), int(octaves)
        self.re_range, self.u_range = tuple(re_range), tuple(u_range)
        self.n_range = tuple(int(x) for x in n_range)
        if net.in_ch != in_channels(self.octaves):
            raise ValueError("network takes %d input channels, features give %d"
                             % (net.in_ch, in_channels(self.octaves)))

    @classmethod
    def from_state(cls, state, grid=GRID, re_range=RE_RANGE, u_range=U_RANGE, device=None, pool=2,
                   octaves=0, n_range=ANY_N):
        return cls(Surrogate.from_state(state, device=device), grid, re_range, u_range, pool, octaves,
                   n_range)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, device=None):
        with np.load(path, allow_pickle=False) as z:
            state = {k: z[k] for k in z.files}
        return cls.from_state(state, int(state["grid"]), tuple(state["re_range"]),
                              tuple(state["u_range"]), device, int(state.get("pool", 2)),
                              int(state.get("octaves", 0)), tuple(state.get("n_range", ANY_N)))

    def save(self, path):
        state = self.net.state_dict()
        state.update(grid=self.grid, re_range=np.array(self.re_range), u_range=np.array(self.u_range),
                     pool=self.pool, octaves=self.octaves, n_range=np.array(self.n_range))
        np.savez(path, **state)

    def inputs(self, params):
        """(B, m, m, 6) device tensor for a list of (Re, U)."""
        x = np.stack([features(re, u, self.grid, self.re_range, self.u_range, self.octaves)
                      for re, u in params])
        return torch.as_tensor(x).to(self.net.device)

    def covers(self, re, u, n=None):
        """True when (Re, U) lies in the ranges the network was trained on (and
        the lattice size n, when given, in the sizes it is meant for)."""
        return (self.re_range[0] <= re <= self.re_range[1]
                and self.u_range[0] <= u <= self.u_range[1]
                and (n is None or self.n_range[0] <= n <= self.n_range[1]))

    def predict(self, re, u, extrapolate=False, n=None):
        """(m, m, 3) FP32 device tensor: ((rho - 1)/U^2, ux/U, uy/U) on cell centres.

        Raises ``OutOfRange`` (a ``WarmStartUnavailable``, so the controller
        cold-starts) for (Re, U) outside the trained ranges, or a lattice size
        n outside the sizes the network is meant for, unless ``extrapolate``:
        the inputs are normalized against those ranges, and what the network
        answers outside them was never fitted."""
        if not extrapolate and not self.covers(re, u, n):
            raise OutOfRange("Re %.6g, U %.6g, N %s outside the trained ranges Re %s, U %s, N %s"
                             % (re, u, n, self.re_range, self.u_range, self.n_range))
        return self.net.forward(self.inputs([(re, u)]))[0].contiguous()


class PredictorSet:
    """Several trained surrogates, each for its own (Re, U, N) ranges.  With
    ``combine="first"`` a cavity goes to the first member whose ranges cover
    it; with ``combine="mean"`` (the shipped set) the predictions of every
    member covering it are averaged: networks trained apart err partly
    independently, and the average's large-scale error, the part the
    lattice's slowest modes must remove, is smaller than either's.  A cavity
    no member covers raises ``OutOfRange`` (the controller then cold-starts).
    Members share the cell grid."""

    COMBINE = ("first", "mean")

    def __init__(self, members, combine="first"):
        self.members = list(members)
        if not self.members:
            raise ValueError("a predictor set needs at least one member")
        if combine not in self.COMBINE:
            raise ValueError("combine must be one of %s" % (self.COMBINE,))
        self.combine = combine
        grids = {m.grid for m in self.members}
        if len(grids) != 1:
            raise ValueError("members must share one cell grid, got %s" % sorted(grids))
        self.grid = self.members[0].grid
        self.pool = self.members[0].pool
        self.re_range = (min(m.re_range[0] for m in self.members),
                         max(m.re_range[1] for m in self.members))
        self.u_range = (min(m.u_range[0] for m in self.members),
                        max(m.u_range[1] for m in self.members))

    def covering(self, re, u, n=None):
        """The members whose ranges cover (Re, U) and the lattice size n."""
        return [m for m in self.members if m.covers(re, u, n)]

    def member(self, re, u, n=None):
        c = self.covering(re, u, n)
        return c[0] if c else None

    def covers(self, re, u, n=None):
        return self.member(re, u, n) is not None

    def predict(self, re, u, extrapolate=False, n=None):
        ms = self.covering(re, u, n)
        if not ms:
            if not extrapolate:
                raise OutOfRange("Re %.6g, U %.6g, N %s outside every member's ranges" % (re, u, n))
            ms = [min(self.members, key=lambda p: abs(math.log(re) - math.log(
                min(max(re, p.re_range[0]), p.re_range[1]))))]
        if self.combine == "first" or len(ms) == 1:
            return ms[0].predict(re, u, extrapolate=extrapolate)
        q = ms[0].predict(re, u, extrapolate=extrapolate).clone()
        for m in ms[1:]:
            q += m.predict(re, u, extrapolate=extrapolate)
        return q / len(ms)


def shipped_weights():
    """The network files shipped with the package: the low-Reynolds default
    first, then every other ``surrogate_weights*.npz`` beside it, by name."""
    others = sorted(f for f in os.listdir(_HERE)
                    if f.startswith("surrogate_weights") and f.endswith(".npz")
                    and os.path.join(_HERE, f) != DEFAULT_WEIGHTS)
    return [DEFAULT_WEIGHTS] + [os.path.join(_HERE, f) for f in others]


def load_default(device=None):
    """The shipped networks as one ``PredictorSet`` averaging the members that
    cover a cavity (``combine="mean"``), or the single low-Reynolds
    ``Predictor`` when it is the only one."""
    members = [Predictor.load(p, device) for p in shipped_weights()]
    return members[0] if len(members) == 1 else PredictorSet(members, combine="mean")


RECONSTRUCTIONS = ("bilinear", "walls", "cubic")


def warm_populations(q, nx, ny, U, omega, neq=True, recon="cubic", pool=2):
    """FP64 populations (9, nx, ny) on q's device from a prediction q (m, m, 3)
    of cell averages over pool x pool nodes.

    ``recon``: "bilinear" (clamped at the outermost cell centres), "walls"
    (bilinear towards ghost cells mirrored across the walls, so boundary
    layers run into the wall and lid velocities) or "cubic" (the same ghosts;
    cell averages deconvolved to centre values, then Catmull-Rom cubics).
    ``neq`` adds the post-collision non-equilibrium part of the strain rate."""
    if recon not in RECONSTRUCTIONS:
        raise ValueError("recon must be one of %s" % (RECONSTRUCTIONS,))
    q = q.contiguous().to(torch.float32)
    m = q.shape[0]
    if q.shape != (m, m, 3):
        raise ValueError("prediction must have shape (m, m, 3)")
    f = torch.empty((9, nx, ny), dtype=torch.float64, device=q.device)
    flags = (nat.WARM_NEQ if neq else 0) | {"bilinear": 0, "walls": nat.WARM_WALLS,
                                            "cubic": nat.WARM_CUBIC}[recon]
    nat.call("lbm_warm_init_f64", q.data_ptr(), m, nx, ny, float(U), float(omega), flags, int(pool),
             f.data_ptr(), nat.stream_ptr())
    return f


class CavityStepper:
    """FP64 N x N cavity at (Re, U) for the ``Controller``; use as a context manager."""

    def __init__(self, n, re, u, predictor=None, neq=True, recon="cubic", every=1000, graph=True,
                 variant="auto", device=None):
        self.n, self.re, self.u = int(n), float(re), float(u)
        self.omega = omega_for_cavity(self.n, self.re, self.u)
        self.predictor, self.neq, self.recon = predictor, bool(neq), recon
        self.every = int(every)
        self.lat = Lattice(self.n, self.n, self.omega, "cavity", self.u, torch.float64,
                           device=device, variant=variant)
        self._graph = StepGraph(self.lat, self.every, enabled=graph)
        self._advance = None
        d = self.lat.device
        self.uxp = torch.zeros((self.n, self.n), dtype=torch.float64, device=d)
        self.uyp = torch.zeros_like(self.uxp)

    def __enter__(self):
        self._advance = self._graph.__enter__()
        return self

    def __exit__(self, *exc):
        self._graph.__exit__(*exc)
        self._advance = None
        return False

    def init_cold(self):
        self.lat.init_equilibrium()

    def init_warm(self):
        if self.predictor is None:
            raise RuntimeError("a warm start needs a predictor")
        if isinstance(self.predictor, (Predictor, PredictorSet)):
            q = self.predictor.predict(self.re, self.u, n=self.n)
        else:  # any object with predict(re, u)
            q = self.predictor.predict(self.re, self.u)
        # the prediction holds averages over cells of n/m x n/m nodes
        m = q.shape[0]
        pool = self.n // m if self.n % m == 0 else getattr(self.predictor, "pool", 2)
        self.lat.load(warm_populations(q, self.n, self.n, self.u, self.omega, self.neq, self.recon, pool))
        self.lat.time = 0

    def advance(self, n):
        if n == self.every and self._advance is not None:
            self._advance()
        else:
            self.lat.step(n)

    def residual(self):
        r = kernels.velocity_residual(self.lat.f, self.uxp, self.uyp, layout=self.lat.layout)
        return float(r.item()) / self.u

    def sync_time(self, t0):
        torch.cuda.synchronize(self.lat.device)
        return time.perf_counter() - t0

    def fields(self):
        """(rho, ux, uy) device tensors of the current state."""
        return self.lat.moments()
