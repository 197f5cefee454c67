"""Reduced-precision population storage: FP32, BF16 or FP16 deviations.

``DeviationLattice`` keeps ``d_i = scale * (f_i - w_i)`` in HBM in the storage
format and steps it with an FP32 collide (``lbm_dev_step``).  Halving or
quartering the bytes per population raises the lattice-update rate of the
bandwidth-bound step by the same factor.  The error this costs is measured
against an FP64 run of the same problem by ``velocity_errors``; FP64 is kept
for the mass and residual reductions.

Default scales: FP16 ``2**10`` (keeps deviations down to ~6e-8 out of the
FP16 subnormal range while the largest cavity deviations, ~0.1, stay far
below the FP16 maximum); BF16 and FP32 ``1`` (their exponent range is FP32's).

Rounding of the 16-bit formats (``rounding``): ``"stochastic"`` (default) or
``"nearest"``; both keep each node's mass and momentum to a single rounding
error (see ``lbm_dev_step`` in include/lbm_d2q9.h).  Stochastic rounding is
seeded by ``seed`` and the step counter, so a run is reproducible.

Storing every K steps (``store_every``): runs of K steps go through one
launch whose K - 1 intermediate steps stay unrounded on chip, and only the
K-th is rounded to the storage format.  The 16-bit formats default to K = 4
(when ``multi_step`` is left at ``"auto"`` and the grid takes a four-step
shape): it rounds a quarter as often, with about half the error of rounding
every step, and runs about twice as fast as the FP64 kernels.
``store_every=1`` rounds every step (every launch shape then computes the bits
of single steps).

``compute="f64"`` runs the collide in FP64 and rounds each result once to the
storage format with the same compensated rounding.  A flow that changes by
only a few FP32 ulps per step is sensitive to the FP32 collide's own
rounding: FP32 storage with an FP32 collide (not rounded at all) biases a
4096^2 shear wave's decay by ~2 %; FP64 compute with stochastic rounding
removes that at the same storage cost.
"""

import math

import torch

from . import _native as nat
from . import kernels

FORMATS = {"f32": nat.FMT_F32, "bf16": nat.FMT_BF16, "f16": nat.FMT_F16}
_RAW = {nat.FMT_F32: torch.float32, nat.FMT_BF16: torch.int16, nat.FMT_F16: torch.int16}
DEFAULT_SCALE = {"f32": 1.0, "bf16": 1.0, "f16": 1024.0}
BYTES = {"f32": 4, "bf16": 2, "f16": 2}
ROUNDING = {"nearest": nat.ROUND_NEAREST, "stochastic": nat.ROUND_STOCHASTIC}
COMPUTE = {"f32": nat.COMPUTE_F32, "f64": nat.COMPUTE_F64}
SEED_STRIDE = 0x85EBCA77  # dither seed increment from one step to the next
# store_every when left to the lattice (multi_step "auto"): the 16-bit formats
# round once per four steps; FP32 storage of an FP32 collide is exact anyway
DEFAULT_STORE_EVERY = {"f32": 1, "bf16": 4, "f16": 4}


def multi_info(variant=-1):
    """(steps per launch, smallest nx, smallest ny) of deviation multi-step
    variant ``variant``."""
    import ctypes
    k, a, b = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
    nat.call("lbm_dev_multi_info", int(variant), ctypes.byref(k), ctypes.byref(a),
             ctypes.byref(b))
    return k.value, a.value, b.value


class DeviationLattice:
    """A D2Q9 lattice whose populations are stored as scaled deviations.

    Same stepping interface as :class:`latticeboltzmann_amd.lattice.Lattice`.
    """

    def __init__(self, nx, ny, omega, bc="cavity", u0=0.1, fmt="f16", scale=None, device=None,
                 variant=-1, layout="rows", rounding="stochastic", seed=0, compute="f32",
                 multi_step="auto", store_every=None):
        nat.require_gpu()
        if fmt not in FORMATS:
            raise ValueError("fmt must be one of %s" % sorted(FORMATS))
        if compute not in COMPUTE:
            raise ValueError("compute must be 'f32' or 'f64'")
        self.compute = COMPUTE[compute]
        if rounding not in ROUNDING:
            raise ValueError("rounding must be 'nearest' or 'stochastic'")
        self.rounding = ROUNDING[rounding]
        self.seed = int(seed) & 0xFFFFFFFF
        if bc not in ("cavity", "periodic"):
            raise ValueError("bc must be 'cavity' or 'periodic'")
        self.nx, self.ny = int(nx), int(ny)
        self.omega, self.bc, self.u0 = float(omega), bc, float(u0)
        self.fmt_name, self.fmt = fmt, FORMATS[fmt]
        self.scale = float(scale if scale is not None else DEFAULT_SCALE[fmt])
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.layout = nat.LAYOUT_ROWS if layout == "rows" else nat.LAYOUT_SOA
        self._bc = nat.BC_CAVITY if bc == "cavity" else nat.BC_PERIODIC
        shape = (self.nx, 9, self.ny) if self.layout == nat.LAYOUT_ROWS else (9, self.nx, self.ny)
        self.d = torch.empty(shape, dtype=_RAW[self.fmt], device=self.device)
        self._e = torch.empty_like(self.d)
        self.multi_variant, self.multi_k, self.round_each = None, 1, True
        chosen = store_every is None
        if chosen:
            store_every = DEFAULT_STORE_EVERY[fmt] if multi_step == "auto" else 1
        self.store_every = int(store_every)
        if self.store_every not in (1, 2, 3, 4, 6, 8):
            raise ValueError("store_every must be 1, 2, 3, 4, 6 or 8")
        auto = variant == "auto"
        if auto:
            variant = self._tune()
        self.variant = int(variant)
        if self.store_every > 1:
            try:
                self._set_store_every(self.store_every, tune=multi_step == "auto" and auto)
            except ValueError:
                if not chosen:
                    raise
                self.store_every = 1  # the default on a grid too small for it: every step
        if self.store_every == 1:  # rounding every step; multi-step launches keep the bits
            if multi_step == "auto" and auto:
                self._tune_multi(self.variant)
            elif multi_step is True:
                self.set_multi_variant(-1)
            elif not isinstance(multi_step, (bool, str)) and multi_step is not None:
                self.set_multi_variant(int(multi_step))
        self.time = 0
        self.init_equilibrium()

    # shapes tried first for each K when storing every K steps (6 and 8: the
    # register sweeps, variants 13 and 14)
    _STORE_EVERY_SHAPES = {2: 1, 3: 3, 4: 10, 6: 13, 8: 14}

    def _set_store_every(self, K, tune):
        """Store every K steps: runs of K steps through one multi-step launch
        whose K - 1 intermediate steps stay unrounded on chip."""
        n_var = nat.load().lbm_dev_multi_variant_count()
        fits = [v for v in range(n_var) if multi_info(v)[0] == K
                and self.nx >= multi_info(v)[1] and self.ny >= multi_info(v)[2]]
        if not fits:
            raise ValueError("the grid is too small for %d steps per launch" % K)
        # FP64-compute tiles that do not fit LDS are refused by the library
        ok = []
        # the state buffers serve as scratch: the state is initialised after tuning
        a, b = self.d.zero_(), self._e
        for v in fits:
            try:
                self._call_multi(a, b, v, round_each=False)
                ok.append(v)
            except ValueError:
                pass
        if not ok:
            raise ValueError("no tile shape with %d steps per launch fits this format" % K)
        best = self._STORE_EVERY_SHAPES[K] if self._STORE_EVERY_SHAPES[K] in ok else ok[0]
        if tune and len(ok) > 1:
            from .autotune import best_variant

            def make_launch(nx, ny):
                def launch(v):
                    self._call_multi(a, b, v, round_each=False)
                return launch

            best = best_variant("dev_once%d_k%d_c%d_r%d" % (n_var, K, self.compute, self.rounding),
                                self.nx, self.ny, self.fmt_name, self.layout, self._bc,
                                make_launch, ok)
        self.multi_variant, self.multi_k, self.round_each = int(best), K, False

    def set_multi_variant(self, variant):
        """Advance runs of K steps with multi-step variant ``variant`` of
        ``lbm_dev_step_multi`` (None: single steps), rounding every step; a
        variant whose tiles do not fit the grid is ignored."""
        self.multi_variant, self.multi_k, self.round_each = None, 1, True
        if variant is None:
            return
        if variant < 0:
            variant = nat.load().lbm_dev_multi_default_variant()
        k, mnx, mny = multi_info(variant)
        if self.nx >= mnx and self.ny >= mny:
            self.multi_variant, self.multi_k = int(variant), k

    def _tune_multi(self, single):
        """Pick single steps or a multi-step shape, whichever is faster per step."""
        from .autotune import best_variant
        n_var = nat.load().lbm_dev_multi_variant_count()
        info = {v: multi_info(v) for v in range(n_var)}
        cands = ["s%d" % single]
        a, b = self.d.zero_(), self._e  # scratch: the state is initialised after tuning
        for v, (k, mnx, mny) in info.items():
            if self.nx >= mnx and self.ny >= mny:
                try:  # the library refuses shapes this compute type cannot hold
                    self._call_multi(a, b, v)
                    cands.append("m%d" % v)
                except ValueError:
                    pass

        def make_launch(nx, ny):

            def launch(c):
                if c[0] == "s":
                    self._call_step(a, b, int(c[1:]))
                else:
                    self._call_multi(a, b, int(c[1:]))
            return launch

        def per_step(times):
            return {c: t / (1 if c[0] == "s" else info[int(c[1:])][0]) for c, t in times.items()}

        best = best_variant("dev_multi%d_c%d_r%d" % (n_var, self.compute, self.rounding), self.nx,
                            self.ny, self.fmt_name, self.layout, self._bc, make_launch, cands,
                            normalize=per_step)
        self.set_multi_variant(int(best[1:]) if best[0] == "m" else None)

    def _tune(self):
        from .autotune import best_variant

        def make_launch(nx, ny):
            a, b = self.d.zero_(), self._e  # scratch: the state is initialised after tuning

            def launch(v):
                self._call_step(a, b, v)
            return launch

        n_var = nat.load().lbm_dev_variant_count()
        return best_variant("dev_step_c%d" % self.compute, self.nx, self.ny, self.fmt_name,
                            self.layout, self._bc, make_launch, range(n_var))

    # -- state -----------------------------------------------------------------
    def init_equilibrium(self, rho=None, ux=None, uy=None):
        shp = (self.nx, self.ny)

        def field(x, default):
            if x is None:
                return torch.full(shp, default, dtype=torch.float64, device=self.device)
            return torch.as_tensor(x, dtype=torch.float64, device=self.device).expand(shp).contiguous()

        self.load(kernels.equilibrium(field(rho, 1.0), field(ux, 0.0), field(uy, 0.0)))
        self.time = 0
        return self

    def load(self, f):
        """Encode FP64 populations (reference layout (9, nx, ny)) into storage."""
        f = torch.as_tensor(f, device=self.device, dtype=torch.float64).contiguous()
        if tuple(f.shape) != (9, self.nx, self.ny):
            raise ValueError("populations must have shape (9, %d, %d)" % (self.nx, self.ny))
        nat.call("lbm_dev_encode", nat.ptr(f), nat.ptr(self.d), self.nx, self.ny, self.layout,
                 self.fmt, self.scale, nat.stream_ptr())
        return self

    def populations_device(self):
        """Decoded FP64 populations in the reference layout, on the device."""
        f = torch.empty((9, self.nx, self.ny), dtype=torch.float64, device=self.device)
        nat.call("lbm_dev_decode", nat.ptr(self.d), nat.ptr(f), self.nx, self.ny, self.layout,
                 self.fmt, self.scale, nat.stream_ptr())
        return f

    def populations(self):
        return self.populations_device().cpu().numpy()

    # -- stepping ----------------------------------------------------------------
    def _seed(self, step_index):
        # a different dither stream every step, reproducible from (seed, time)
        return (self.seed * 0x9E3779B1 + step_index * SEED_STRIDE + 1) & 0xFFFFFFFF

    def _call_step(self, src, dst, variant, ux=None, uy=None, step_index=0):
        nat.call("lbm_dev_step", nat.ptr(src), nat.ptr(dst), self.fmt, self.nx, self.ny,
                 self.layout, self.omega, self._bc, self.u0, self.scale, self.compute,
                 self.rounding, self._seed(step_index), nat.ptr(ux), nat.ptr(uy), int(variant),
                 nat.stream_ptr())

    def _call_multi(self, src, dst, variant, ux=None, uy=None, step_index=0, round_each=True):
        nat.call("lbm_dev_step_multi", nat.ptr(src), nat.ptr(dst), self.fmt, self.nx, self.ny,
                 self.layout, self.omega, self._bc, self.u0, self.scale, self.compute,
                 self.rounding, self._seed(step_index), SEED_STRIDE, int(bool(round_each)),
                 nat.ptr(ux), nat.ptr(uy), int(variant), nat.stream_ptr())

    def launches(self, n):
        """Kernel launches (buffer swaps) that ``step(n)`` makes."""
        q = n // self.multi_k if self.multi_variant is not None else 0
        return q + (n - q * self.multi_k)

    def step(self, n=1, velocity=None):
        """Advance ``n`` steps (runs of K through the multi-step kernel when
        enabled: the same bits as single steps, or, with ``store_every`` K,
        rounded to the storage format once per K steps; a remainder of fewer
        than K steps runs as single steps).  ``velocity=(ux, uy)`` FP32 tensors
        receive the pre-collision velocity of the last step."""
        runs = n // self.multi_k if self.multi_variant is not None else 0
        single = n - runs * self.multi_k
        ux = uy = None
        for r in range(runs):
            if velocity is not None and r == runs - 1 and single == 0:
                ux, uy = velocity
            self._call_multi(self.d, self._e, self.multi_variant, ux, uy, step_index=self.time,
                             round_each=self.round_each)
            self.d, self._e = self._e, self.d
            self.time += self.multi_k
        for s_ in range(single):
            if velocity is not None and s_ == single - 1:
                ux, uy = velocity
            self._call_step(self.d, self._e, self.variant, ux, uy, step_index=self.time)
            self.d, self._e = self._e, self.d
            self.time += 1
        return self

    # -- observables ---------------------------------------------------------------
    def moments(self):
        """(rho, ux, uy) in FP64 from the decoded populations."""
        return kernels.moments(self.populations_device())

    def velocity(self):
        _, ux, uy = self.moments()
        return ux.cpu().numpy(), uy.cpu().numpy()

    def mass(self):
        ws = kernels._workspace(self.device)
        nat.call("lbm_dev_mass", nat.ptr(self.d), self.nx, self.ny, self.fmt, self.scale,
                 nat.ptr(ws.work), nat.ptr(ws.out), nat.stream_ptr())
        return float(ws.out.item())

    def run_until_steady(self, tol=1e-6, every=1000, max_steps=10**7, u_ref=None):
        u_ref = u_ref or self.u0
        ws = kernels._workspace(self.device)
        uxp = torch.zeros((self.nx, self.ny), dtype=torch.float64, device=self.device)
        uyp = torch.zeros_like(uxp)

        def residual():
            nat.call("lbm_dev_velocity_residual", nat.ptr(self.d), self.nx, self.ny, self.layout,
                     self.fmt, self.scale, nat.ptr(uxp), nat.ptr(uyp), nat.ptr(ws.work),
                     nat.ptr(ws.out), nat.stream_ptr())
            return float(ws.out.item())

        residual()
        taken, res = 0, float("inf")
        while taken < max_steps:
            self.step(every)
            taken += every
            res = residual() / u_ref
            if res < tol:
                break
            if math.isinf(res) or math.isnan(res):
                raise FloatingPointError("lattice diverged: non-finite velocity residual "
                                         "after %d steps" % taken)
        return taken, res


def velocity_errors(ux, uy, ux_ref, uy_ref, u_ref):
    """RMS and max of |u - u_ref| over the grid, divided by ``u_ref`` (FP64)."""
    dx = ux.double() - ux_ref.double()
    dy = uy.double() - uy_ref.double()
    mag = torch.sqrt(dx * dx + dy * dy)
    return {"l2_over_U": float(torch.sqrt(torch.mean(mag * mag)).item()) / u_ref,
            "linf_over_U": float(mag.max().item()) / u_ref}


def _bench(argv=None):
    """Time a deviation-storage cavity:
    ``python -m latticeboltzmann_amd.precision [--n 8192] [--fmt f16] [--store-every 4] [--steps 96]``
    (Re 1000, U 0.1; one JSON line with the MLUPS and the launch shape)."""
    import argparse
    import json
    from .lattice import omega_for_cavity
    ap = argparse.ArgumentParser(description=_bench.__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--fmt", default="f16", choices=sorted(FORMATS))
    ap.add_argument("--store-every", type=int, default=4)
    ap.add_argument("--rounding", default="stochastic", choices=sorted(ROUNDING))
    ap.add_argument("--steps", type=int, default=96)
    a = ap.parse_args(argv)
    om = omega_for_cavity(a.n, 1000.0, 0.1)
    lat = DeviationLattice(a.n, a.n, om, "cavity", 0.1, fmt=a.fmt, rounding=a.rounding,
                           store_every=a.store_every, variant="auto")
    k = lat.multi_k
    lat.step(2 * k)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    lat.step(a.steps)
    end.record()
    torch.cuda.synchronize()
    sec = start.elapsed_time(end) / 1e3
    print(json.dumps({"n": a.n, "fmt": a.fmt, "store_every": a.store_every, "steps": a.steps,
                      "variant": lat.variant, "multi_variant": lat.multi_variant, "steps_per_launch": k,
                      "mlups": round(a.n * a.n * a.steps / sec / 1e6, 1)}), flush=True)


if __name__ == "__main__":
    _bench()
