"""Population checkpoints: save a running lattice and resume it bit for bit.

The reference writes only velocity dumps (``cavity_opt1.py:345-347``); a long
run (the 50 000-step 8192^2 precision sweep, a cavity run to steady state)
cannot be resumed from them, because the populations carry more than the
velocity.  A checkpoint here holds the populations exactly as stored (the
FP64/FP32 values of a ``Lattice``, or the raw FP32/BF16/FP16 words of a
``DeviationLattice`` with their scale), the step counter (which also seeds a
deviation lattice's stochastic rounding), and the parameters the step
depends on.  Resuming from it continues the run with the same bits as a run
that was never interrupted.

Format: one ``.npz`` written with numpy (no pickled objects; ``numpy.load``
reads it with its default ``allow_pickle=False``), written to a temporary
name and renamed, so a crash during a save leaves the previous checkpoint.

A ``DeviationLattice`` storing every K steps rounds once per K-step launch,
so its state is that of the uninterrupted run only at multiples of K of its
own step counter: save it there.
"""

import os

import numpy as np
import torch

FORMAT_VERSION = 1

_LATTICE_KEYS = ("nx", "ny", "omega", "bc", "u0", "layout")
_DEVIATION_KEYS = _LATTICE_KEYS + ("fmt", "scale", "rounding", "seed", "compute", "store_every")


def _kind(lat):
    return "deviation" if hasattr(lat, "fmt") and hasattr(lat, "d") else "lattice"


def _params(lat):
    keys = _DEVIATION_KEYS if _kind(lat) == "deviation" else _LATTICE_KEYS
    return {k: getattr(lat, k) for k in keys}


def _state(lat):
    return lat.d if _kind(lat) == "deviation" else lat.f


def save(lat, path):
    """Write ``lat``'s populations, step counter and parameters to ``path``."""
    state = _state(lat)
    raw = state.detach().contiguous().cpu()
    arrays = {
        "version": np.int64(FORMAT_VERSION),
        "kind": np.str_(_kind(lat)),
        "time": np.int64(lat.time),
        "dtype": np.str_(str(state.dtype).replace("torch.", "")),
        "shape": np.asarray(tuple(state.shape), dtype=np.int64),
        # the raw bytes: BF16 has no numpy dtype, and bytes keep every bit
        "state": raw.view(torch.uint8).numpy() if raw.dtype == torch.bfloat16
        else raw.numpy().view(np.uint8),
    }
    for k, v in _params(lat).items():
        arrays["param_" + k] = np.asarray(v)
    tmp = "%s.tmp%d" % (path, os.getpid())
    with open(tmp, "wb") as fh:
        np.savez(fh, **arrays)
    os.replace(tmp, path)
    return path


def read(path):
    """The checkpoint's fields as a dict (numpy values), without a device."""
    with np.load(path) as z:
        out = {k: z[k] for k in z.files}
    if int(out["version"]) != FORMAT_VERSION:
        raise ValueError("checkpoint format version %d, expected %d"
                         % (int(out["version"]), FORMAT_VERSION))
    return out


def restore(lat, path):
    """Load a checkpoint into ``lat``, which must have been built with the same
    parameters (grid, omega, boundary, lid velocity, layout, and for a
    deviation lattice the format, scale, rounding, seed, compute type and
    store interval); raises ``ValueError`` otherwise."""
    ck = read(path)
    if str(ck["kind"]) != _kind(lat):
        raise ValueError("checkpoint of a %s, lattice is a %s" % (ck["kind"], _kind(lat)))
    for k, v in _params(lat).items():
        saved = ck["param_" + k].item()
        if saved != v:
            raise ValueError("checkpoint %s = %r, lattice has %r" % (k, saved, v))
    state = _state(lat)
    if tuple(int(x) for x in ck["shape"]) != tuple(state.shape) or \
            str(ck["dtype"]) != str(state.dtype).replace("torch.", ""):
        raise ValueError("checkpoint holds %s %s, lattice stores %s %s"
                         % (tuple(ck["shape"]), ck["dtype"], tuple(state.shape), state.dtype))
    raw = torch.from_numpy(np.ascontiguousarray(ck["state"])).reshape(-1)
    if raw.numel() != state.numel() * state.element_size():
        raise ValueError("checkpoint state has %d bytes, expected %d"
                         % (raw.numel(), state.numel() * state.element_size()))
    state.view(torch.uint8).view(-1).copy_(raw.to(state.device))
    lat.time = int(ck["time"])
    return lat
