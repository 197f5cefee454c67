"""Many independent lattices stepped together: the batched parameter sweep.

``BatchedLattices`` keeps hundreds of shear-wave (or cavity) instances of
different sizes in one population pool (scaled deviations, row-interleaved
layout per instance) and steps all of them together.  By default one launch
advances every instance four steps (``lbm_batch_step_multi``: a work list of
(instance, strip of columns, rows) items, one wave each, with the
intermediate steps kept in registers), bitwise equal to four single steps;
steps left over run one per launch through the single-step kernel
(``lbm_batch_step``), one persistent launch per launch-shape group, whose
tile shape per size class (nx x ny) the autotuner picks.  A chunk of steps plus the amplitude sample that
ends it is captured once into a HIP graph and replayed, so a 2000-step sweep is
a handful of graph launches.

The observable is the shear-wave amplitude by projection on the initial sine
(as ``shear_wave_opt1.py:174`` does on one line, here over the whole field),
sampled every ``sample_every`` steps in FP64 on the device.  ``viscosities``
fits nu per instance in FP64 (A = A0 exp(-nu k^2 t), ``shear_wave_decay.py:162-167``).
"""

import numpy as np
import torch

from . import _native as nat
from . import kernels
from .autotune import best_variant
from .precision import COMPUTE, DEFAULT_SCALE, FORMATS, ROUNDING, _RAW

SIZES = (64, 128, 256, 512, 1024)

INST_DTYPE = np.dtype([("offset", "<i8"), ("nx", "<i4"), ("ny", "<i4"), ("omega", "<f8"),
                       ("u0", "<f8"), ("bc", "<i4"), ("dir", "<i4")])
TILE_DTYPE = np.dtype([("inst", "<i4"), ("first", "<i4")])
ITEM_DTYPE = np.dtype([("inst", "<i4"), ("strip", "<i4"), ("row0", "<i4"), ("rows", "<i4")])
SWEEP_ROWS = 64  # rows per work item of the K-step sweep
SWEEP_MIN_NY = 64  # a wave of the K-step sweep spans 64 columns
AMP_TILE = 256  # nodes per amplitude tile (the reduction block size)
SEED_STEP = 0x85EBCA77  # difference of the dither seeds of consecutive steps


def shear_wave_sweep(n_instances=512, seed=0):
    """The batched sweep's instances, drawn with ``numpy.random.default_rng(seed)``.

    Nx and Ny independently from {64, 128, 256, 512, 1024}; omega ~ U[0.3, 1.9];
    amplitude ~ U[0.01, 0.1]; direction 'x' (u_x = A sin(2 pi y / Ny)) or 'y'
    (u_y = A sin(2 pi x / Nx)) with probability 1/2, drawn in that order.
    """
    rng = np.random.default_rng(seed)
    n = int(n_instances)
    nx = rng.choice(SIZES, n)
    ny = rng.choice(SIZES, n)
    omega = rng.uniform(0.3, 1.9, n)
    amp = rng.uniform(0.01, 0.1, n)
    direction = np.where(rng.random(n) < 0.5, "x", "y")
    return [dict(nx=int(nx[i]), ny=int(ny[i]), omega=float(omega[i]), amplitude=float(amp[i]),
                 direction=str(direction[i]), bc="periodic") for i in range(n)]


def _to_device_bytes(arr, device):
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).to(device)


def sweep_items(table, steps, rows=SWEEP_ROWS):
    """Work list of the K-step batched sweep: for every instance, every strip
    of columns (``lbm_batch_sweep_strips``) times every range of ``rows``
    rows, each item worked by one wave."""
    lib = nat.load()
    out = []
    for i in range(len(table)):
        nx, ny = int(table["nx"][i]), int(table["ny"][i])
        for strip in range(lib.lbm_batch_sweep_strips(steps, ny)):
            for r0 in range(0, nx, rows):
                out.append((i, strip, r0, min(rows, nx - r0)))
    return np.array(out, dtype=np.int32).view(ITEM_DTYPE).reshape(-1)


def _tiles(inst_ids, sizes, tile_nodes):
    rows = []
    for i in inst_ids:
        n = sizes[i]
        first = np.arange(0, n, tile_nodes, dtype=np.int32)
        rows.append(np.stack([np.full_like(first, i), first], axis=1))
    arr = np.concatenate(rows).astype(np.int32)
    return arr.view(TILE_DTYPE).reshape(-1)


class BatchedLattices:
    """Independent D2Q9 lattices stepped together on one GPU."""

    def __init__(self, instances, fmt="f32", scale=None, rounding="stochastic", seed=0,
                 variant="auto", device=None, compute=None, steps_per_launch=4):
        nat.require_gpu()
        # FP32 storage defaults to an FP64 collide with unbiased rounding: a
        # slowly decaying wave changes by only a few FP32 ulps per step
        compute = compute or ("f64" if fmt == "f32" else "f32")
        if compute not in COMPUTE:
            raise ValueError("compute must be 'f32' or 'f64'")
        self.compute, self.compute_name = COMPUTE[compute], compute
        self.instances = [dict(x) for x in instances]
        self.n = len(self.instances)
        if self.n == 0:
            raise ValueError("no instances")
        self.fmt_name, self.fmt = fmt, FORMATS[fmt]
        self.scale = float(scale if scale is not None else DEFAULT_SCALE[fmt])
        self.rounding = ROUNDING[rounding]
        self.seed = int(seed) & 0xFFFFFFFF
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        lib = nat.load()

        table = np.zeros(self.n, dtype=INST_DTYPE)
        off = 0
        for i, p in enumerate(self.instances):
            nx, ny = int(p["nx"]), int(p["ny"])
            if nx % 64 or ny % 64:
                raise ValueError("instance sides must be multiples of 64")
            table[i] = (off, nx, ny, p["omega"], p.get("u0", 0.0),
                        nat.BC_CAVITY if p.get("bc", "periodic") == "cavity" else nat.BC_PERIODIC,
                        0 if p.get("direction", "x") == "x" else 1)
            off += 9 * nx * ny
        self.table = table
        self.total = off
        self.sizes = (table["nx"].astype(np.int64) * table["ny"]).tolist()
        self.a = torch.empty(off, dtype=_RAW[self.fmt], device=self.device)
        self.b = torch.empty_like(self.a)
        self._insts = _to_device_bytes(table, self.device)
        self.time = 0
        self.init_states()

        # amplitude tiles in instance order
        amp_tiles = _tiles(range(self.n), self.sizes, AMP_TILE)
        counts = np.array([(s + AMP_TILE - 1) // AMP_TILE for s in self.sizes], dtype=np.int32)
        inst_tiles = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self._amp_tiles = _to_device_bytes(amp_tiles, self.device)
        self._amp_ntiles = len(amp_tiles)
        self._inst_tiles = torch.from_numpy(inst_tiles).to(self.device)
        self._partial = torch.empty(self._amp_ntiles, dtype=torch.float64, device=self.device)
        self._epoch = torch.zeros(1, dtype=torch.int32, device=self.device)

        # launch groups: size class -> tile shape
        classes = {}
        for i, p in enumerate(self.instances):
            classes.setdefault((int(p["nx"]), int(p["ny"])), []).append(i)
        self.class_variant = {}
        n_var = lib.lbm_batch_variant_count()
        for (nx, ny), ids in sorted(classes.items()):
            if variant == "auto":
                v = self._tune_class(nx, ny, ids, n_var)
            else:
                v = int(variant)
            self.class_variant[(nx, ny)] = v
        groups = {}
        for key, ids in classes.items():
            groups.setdefault(self.class_variant[key], []).extend(ids)
        self.groups = []
        for v, ids in sorted(groups.items()):
            t = _tiles(sorted(ids), self.sizes, lib.lbm_batch_tile_nodes(v))
            self.groups.append((v, _to_device_bytes(t, self.device), len(t)))
        # K steps per launch: every instance in one work list, one launch
        self.multi_k = int(steps_per_launch)
        if self.multi_k not in (1, 2, 4):
            raise ValueError("steps_per_launch must be 1, 2 or 4")
        if self.multi_k > 1:
            if any(int(p["nx"]) < 2 * self.multi_k or int(p["ny"]) < SWEEP_MIN_NY
                   for p in self.instances):
                raise ValueError("steps_per_launch > 1 needs nx >= 2 * steps_per_launch and "
                                 "ny >= %d in every instance" % SWEEP_MIN_NY)
            items = sweep_items(table, self.multi_k)
            self._items = _to_device_bytes(items, self.device)
            self._nitems = len(items)

    # -- state -------------------------------------------------------------------
    def init_states(self):
        """f = f_eq(rho = 1, u = initial shear wave, or 0 for a cavity) for every instance."""
        esz = self.a.element_size()
        for i, p in enumerate(self.instances):
            nx, ny = int(p["nx"]), int(p["ny"])
            rho = torch.ones((nx, ny), dtype=torch.float64, device=self.device)
            ux = torch.zeros_like(rho)
            uy = torch.zeros_like(rho)
            if p.get("bc", "periodic") == "periodic":
                a = float(p["amplitude"])
                if p.get("direction", "x") == "x":
                    ux[:] = a * torch.sin(2 * np.pi * torch.arange(ny, dtype=torch.float64,
                                                                  device=self.device) / ny)[None, :]
                else:
                    uy[:] = a * torch.sin(2 * np.pi * torch.arange(nx, dtype=torch.float64,
                                                                  device=self.device) / nx)[:, None]
            f = kernels.equilibrium(rho, ux, uy)
            dst = self.a.data_ptr() + int(self.table["offset"][i]) * esz
            nat.call("lbm_dev_encode", nat.ptr(f), dst, nx, ny, nat.LAYOUT_ROWS, self.fmt,
                     self.scale, nat.stream_ptr())
        self.time = 0

    def populations(self, i):
        """Instance i decoded to FP64 in the reference layout (9, nx, ny), numpy."""
        nx, ny = int(self.table["nx"][i]), int(self.table["ny"][i])
        f = torch.empty((9, nx, ny), dtype=torch.float64, device=self.device)
        src = self.a.data_ptr() + int(self.table["offset"][i]) * self.a.element_size()
        nat.call("lbm_dev_decode", src, nat.ptr(f), nx, ny, nat.LAYOUT_ROWS, self.fmt, self.scale,
                 nat.stream_ptr())
        return f.cpu().numpy()

    # -- stepping ----------------------------------------------------------------------
    def _launch(self, src, dst, variant, tiles, ntiles, seed):
        nat.call("lbm_batch_step", nat.ptr(src), nat.ptr(dst), self.fmt, self.scale,
                 nat.ptr(self._insts), nat.ptr(tiles), ntiles, self.compute, self.rounding, seed,
                 nat.ptr(self._epoch), variant, 0, nat.stream_ptr())

    def _tune_class(self, nx, ny, ids, n_var):
        lib = nat.load()

        def make_launch(nx_, ny_):
            tiles = {v: _tiles(ids, self.sizes, lib.lbm_batch_tile_nodes(v)) for v in range(n_var)}
            dev = {v: (_to_device_bytes(t, self.device), len(t)) for v, t in tiles.items()}

            def launch(v):
                self._launch(self.a, self.b, v, dev[v][0], dev[v][1], 0)
            return launch

        return best_variant("batch_step_c%d" % self.compute, nx, ny, self.fmt_name, nat.LAYOUT_ROWS,
                            "x%d" % len(ids), make_launch, range(n_var), reps=5)

    def _step_seed(self, s):
        # consecutive steps differ by SEED_STEP, so a K-step launch takes the
        # seed of its first step and SEED_STEP as the stride
        return (self.seed * 0x9E3779B1 + s * SEED_STEP + 0x165667B1) & 0xFFFFFFFF

    def _launch_multi(self, src, dst, seed):
        nat.call("lbm_batch_step_multi", nat.ptr(src), nat.ptr(dst), self.fmt, self.scale,
                 nat.ptr(self._insts), nat.ptr(self._items), self._nitems, self.multi_k,
                 self.compute, self.rounding, seed, SEED_STEP, nat.ptr(self._epoch),
                 nat.stream_ptr())

    def _launches(self, n, first):
        """(kind, step index) of the launches that advance n steps from step
        ``first``: runs of multi_k steps through the sweep, the rest single."""
        k = self.multi_k
        runs = n // k if k > 1 else 0
        out = [("multi", first + r * k) for r in range(runs)]
        out += [("single", first + runs * k + s) for s in range(n - runs * k)]
        return out

    def step(self, n=1):
        """Advance every instance n steps (eager launches)."""
        for kind, s in self._launches(n, self.time):
            if kind == "multi":
                self._launch_multi(self.a, self.b, self._step_seed(s))
            else:
                for v, tiles, nt in self.groups:
                    self._launch(self.a, self.b, v, tiles, nt, self._step_seed(s))
            self.a, self.b = self.b, self.a
        self.time += n
        return self

    def _amplitude(self, amp, nslots, pool=None):
        pool = self.a if pool is None else pool
        nat.call("lbm_batch_amplitude", nat.ptr(pool), self.fmt, self.scale, nat.ptr(self._insts),
                 self.n, nat.ptr(self._amp_tiles), self._amp_ntiles, nat.ptr(self._inst_tiles),
                 nat.ptr(self._partial), nat.ptr(amp), nslots, nat.ptr(self._epoch),
                 nat.stream_ptr())

    def run(self, nsteps, sample_every=100, graph=True):
        """Run nsteps steps, sampling every instance's amplitude every
        ``sample_every`` steps (and at the start).  With ``graph`` the chunk
        of ``sample_every`` steps + sample is captured once and replayed (as
        two graphs used in turn when a chunk takes an odd number of launches,
        so that every replay starts from the buffer holding the state).

        Returns (times, amplitudes) with amplitudes of shape (samples, n).
        """
        if nsteps % sample_every or sample_every % 2:
            raise ValueError("nsteps must be a multiple of sample_every, which must be even")
        chunks = nsteps // sample_every
        nslots = chunks + 1
        amp = torch.full((nslots, self.n), float("nan"), dtype=torch.float64, device=self.device)
        nat.call("lbm_counter_set", nat.ptr(self._epoch), 0, nat.stream_ptr())
        self._amplitude(amp, nslots)
        nat.call("lbm_counter_increment", nat.ptr(self._epoch), nat.stream_ptr())
        t0 = self.time
        launches = self._launches(sample_every, 0)
        odd = len(launches) % 2 == 1

        def chunk(a, b):
            # sample_every steps, then the amplitude sample of the buffer that
            # holds the state; the dither seed of a step is its index in the
            # chunk combined on the device with the chunk counter
            for kind, st in launches:
                if kind == "multi":
                    self._launch_multi(a, b, self._step_seed(st))
                else:
                    for v, tiles, nt in self.groups:
                        self._launch(a, b, v, tiles, nt, self._step_seed(st))
                a, b = b, a
            self._amplitude(amp, nslots, a)
            nat.call("lbm_counter_increment", nat.ptr(self._epoch), nat.stream_ptr())

        starts = [(self.a, self.b), (self.b, self.a)]
        order = [starts[c % 2] if odd else starts[0] for c in range(chunks)]
        if not graph:
            for a, b in order:
                chunk(a, b)
        else:
            side = torch.cuda.Stream(device=self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            lib = nat.load()
            graphs = []
            with torch.cuda.stream(side):
                try:
                    for a, b in starts[:2 if odd and chunks > 1 else 1]:
                        g = nat.ctypes.c_void_p()
                        nat.call("lbm_graph_begin", nat.stream_ptr())
                        try:
                            chunk(a, b)
                        finally:
                            status = lib.lbm_graph_end(nat.stream_ptr(), nat.ctypes.byref(g))
                        nat.check(status, "lbm_graph_end")
                        graphs.append(g)
                    for c in range(chunks):
                        nat.call("lbm_graph_launch", graphs[c % len(graphs)], nat.stream_ptr())
                finally:
                    side.synchronize()
                    for g in graphs:
                        nat.check(lib.lbm_graph_destroy(g), "lbm_graph_destroy")
            torch.cuda.current_stream(self.device).wait_stream(side)
        if odd and chunks % 2 == 1:
            self.a, self.b = self.b, self.a
        self.time += nsteps
        times = t0 + sample_every * np.arange(nslots)
        return times, amp.cpu().numpy()

    # -- analysis ------------------------------------------------------------------------
    def wavelengths(self):
        return np.array([p["ny"] if p.get("direction", "x") == "x" else p["nx"]
                         for p in self.instances], dtype=np.float64)

    def viscosities(self, times, amplitudes, floor=1e-3, t_min=None):
        """nu per instance from a log-linear least-squares fit in FP64, using the
        samples at t >= t_min whose amplitude is above ``floor`` times the
        initial one.

        The start f = f_eq carries no non-equilibrium part, so the first steps
        decay at a different rate (an initial layer of a few relaxation
        times); by default the fit starts at the second sample.
        """
        nus = np.full(self.n, np.nan)
        k2 = (2 * np.pi / self.wavelengths()) ** 2
        times = np.asarray(times)
        if t_min is None:
            t_min = times[1] if len(times) > 2 else times[0]
        for i in range(self.n):
            a = amplitudes[:, i]
            ok = np.isfinite(a) & (a > floor * a[0]) & (times >= t_min)
            if ok.sum() >= 2:
                slope = np.polyfit(times[ok].astype(np.float64), np.log(a[ok]), 1)[0]
                nus[i] = -slope / k2[i]
        return nus

    def theory(self):
        omega = np.array([p["omega"] for p in self.instances])
        return (1.0 / omega - 0.5) / 3.0
