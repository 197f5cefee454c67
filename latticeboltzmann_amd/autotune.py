"""Launch-shape autotuner for the fused step kernels.

Every launch shape (``variant``) of a step kernel computes the same bits; they
differ only in threads per block, nodes per thread and cache hints, and the
fastest one depends on the grid and the storage type.  ``best_variant`` times
each shape on a scratch grid of the requested size class and caches the winner
per (library build, device, kernel, dtype, layout, boundary kind, size class)
in a JSON file (``$LBM_TUNE_CACHE`` or
``~/.cache/latticeboltzmann_amd/tune.json``).  The build is the content hash
of ``liblbm_d2q9.so``, so a rebuilt library (other kernels, another shape
table) is tuned afresh instead of reusing choices measured on another build.

Size classes round each side up to a power of two, so a 1000 x 1000 grid and a
1024 x 1024 grid share one entry.
"""

import json
import os
import threading

import torch

from . import _native as nat

_LOCK = threading.Lock()
_MEMO = {}


def cache_path():
    return os.environ.get("LBM_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "latticeboltzmann_amd",
                                       "tune.json"))


def size_class(nx, ny):
    def up(n):
        return 1 << max(0, int(n - 1).bit_length())
    return "%dx%d" % (up(nx), up(ny))


def _load_cache():
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(entries):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as fh:
            json.dump(entries, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs a re-tune next time


def time_variants(launch, variants, reps=10, warmup=2):
    """Mean seconds per launch of ``launch(variant)`` for each variant."""
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    for v in variants:
        for _ in range(warmup):
            launch(v)
        start.record()
        for _ in range(reps):
            launch(v)
        end.record()
        end.synchronize()
        out[v] = start.elapsed_time(end) / 1e3 / reps
    return out


def best_variant(kernel, nx, ny, dtype, layout, bc, make_launch, variants, reps=10,
                 use_cache=True, normalize=None):
    """Fastest variant of ``kernel`` for this problem class.

    ``make_launch(nx, ny)`` returns a callable ``launch(variant)`` that runs one
    launch on scratch buffers of that size (it owns them); ``normalize`` maps
    the launch times to the times compared (e.g. per step).
    """
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    key = "|".join([nat.build_id(), dev, kernel, str(dtype), str(layout), str(bc),
                    size_class(nx, ny)])
    with _LOCK:
        if key in _MEMO:
            return _MEMO[key]
        entries = _load_cache() if use_cache else {}
        if key in entries and entries[key].get("variant") in variants:
            _MEMO[key] = entries[key]["variant"]
            return _MEMO[key]
    launch = make_launch(nx, ny)
    times = time_variants(launch, list(variants), reps=reps)
    if normalize is not None:
        times = normalize(times)
    best = min(times, key=times.get)
    with _LOCK:
        _MEMO[key] = best
        if use_cache:
            entries = _load_cache()
            entries[key] = {"variant": best, "seconds": {str(k): v for k, v in times.items()}}
            _store_cache(entries)
    return best


def step_variant(nx, ny, dtype, layout, bc, reps=10, use_cache=True):
    """Autotuned variant of the FP64/FP32 fused step (``lbm_step_layout_*``)."""
    from . import kernels

    def make_launch(nx, ny):
        shape = (9, nx, ny) if layout == nat.LAYOUT_SOA else (nx, 9, ny)
        a = torch.empty(shape, dtype=dtype, device="cuda")
        rows = torch.ones((nx, ny), dtype=dtype, device="cuda")
        zero = torch.zeros_like(rows)
        soa = kernels.equilibrium(rows, zero, zero)
        if layout == nat.LAYOUT_SOA:
            a.copy_(soa)
        else:
            kernels.relayout(soa, nat.LAYOUT_SOA, nat.LAYOUT_ROWS, out=a)
        b = torch.empty_like(a)

        def launch(v):
            kernels.step(a, b, 1.0, bc, 0.1, variant=v, layout=layout)
        return launch

    n_var = nat.load().lbm_step_variant_count()
    return best_variant("step", nx, ny, dtype, layout, bc, make_launch, range(n_var),
                        reps=reps, use_cache=use_cache)


def multi_variant(nx, ny, dtype, layout, bc, reps=10, use_cache=True):
    """Autotuned variant of the multi-step kernel (``lbm_step_multi_*``): the
    (steps per launch, tile) shape with the least time per step, or None when
    the grid is smaller than every shape needs."""
    from . import kernels
    n_var = nat.load().lbm_step_multi_variant_count()
    info = {v: kernels.step_multi_info(v) for v in range(n_var)}
    fits = [v for v, (k, mnx, mny) in info.items() if nx >= mnx and ny >= mny]
    if not fits:
        return None

    def make_launch(nx, ny):
        shape = (9, nx, ny) if layout == nat.LAYOUT_SOA else (nx, 9, ny)
        a = torch.empty(shape, dtype=dtype, device="cuda")
        rows = torch.ones((nx, ny), dtype=dtype, device="cuda")
        zero = torch.zeros_like(rows)
        soa = kernels.equilibrium(rows, zero, zero)
        if layout == nat.LAYOUT_SOA:
            a.copy_(soa)
        else:
            kernels.relayout(soa, nat.LAYOUT_SOA, nat.LAYOUT_ROWS, out=a)
        b = torch.empty_like(a)

        def launch(v):
            kernels.step_multi(a, b, 1.0, bc, 0.1, variant=v, layout=layout)
        return launch

    def per_step(times):
        return {v: t / info[v][0] for v, t in times.items()}

    return best_variant("step_multi", nx, ny, dtype, layout, bc, make_launch, fits,
                        reps=reps, use_cache=use_cache, normalize=per_step)
