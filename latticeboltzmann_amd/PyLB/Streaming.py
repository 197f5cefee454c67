"""``PyLB.stream`` and ``PyLB.stream_and_bounce_back`` on the GPU.

``stream(f_ikl)`` replaces ``PyLB/Streaming.py:33-46``: periodic streaming of
a ``(9, nx, ny)`` array in place, channel ``i`` moved by ``c_i``
(``PyLB/Streaming.py:28-29``).  As with the reference's ``np.roll`` loop the
values are only permuted, so any dtype of 1, 2, 4 or 8 bytes is accepted and
the result is bit-exact.

``stream_and_bounce_back(f_ikl, u0=0.1)`` replaces the cavity drivers' wall
function (``cavity_opt1.py:177-245``, ``cavity_opt2.py:109-177``): streaming,
half-way bounce-back on the bottom (l = 0), left (k = 0) and right
(k = nx-1) walls, and a lid at l = ny-1 moving with ``u0`` in +x; float32 or
float64.

numpy arguments are modified in place through a device round trip; torch
tensors on a ROCm device are processed on the device.
"""

import numpy as np
import torch

from latticeboltzmann_amd import kernels
from . import _staging

# Channel velocities, as PyLB/Streaming.py:28-29 exposes them.
c_ic = np.array([[0, 1, 0, -1, 0, 1, -1, -1, 1],
                 [0, 0, 1, 0, -1, 1, 1, -1, -1]]).T


def _check_populations(f, name):
    if isinstance(f, np.ndarray):
        if not f.flags.writeable:
            raise ValueError("%s(): array is read-only" % name)
    elif not _staging.is_device_tensor(f):
        raise TypeError("%s(): expected a numpy array or a ROCm device tensor" % name)
    if f.ndim != 3 or f.shape[0] != 9:
        raise ValueError("%s(): populations must have shape (9, nx, ny), got %s"
                         % (name, tuple(f.shape)))
    return f.shape[1] > 0 and f.shape[2] > 0


def stream(f_ikl):
    """Propagate channel occupations by one cell, periodic, in place."""
    if not _check_populations(f_ikl, "stream"):
        return
    itemsize = f_ikl.itemsize if isinstance(f_ikl, np.ndarray) else f_ikl.element_size()
    if itemsize not in (1, 2, 4, 8):
        raise TypeError("stream(): element size must be 1, 2, 4 or 8 bytes")
    if isinstance(f_ikl, np.ndarray):
        out = kernels.stream(_staging.to_device(f_ikl))
        _staging.write_back(f_ikl, out)
    else:
        src = f_ikl.contiguous()
        out = kernels.stream(src)
        f_ikl.copy_(out)


def stream_and_bounce_back(f_ikl, u0=0.1):
    """Stream and apply the cavity's no-slip walls and moving lid, in place."""
    if not _check_populations(f_ikl, "stream_and_bounce_back"):
        return
    if isinstance(f_ikl, np.ndarray):
        ok = f_ikl.dtype in (np.dtype(np.float32), np.dtype(np.float64))
    else:
        ok = f_ikl.dtype in (torch.float32, torch.float64)
    if not ok:
        raise TypeError("stream_and_bounce_back(): dtype must be float32 or float64")
    if isinstance(f_ikl, np.ndarray):
        out = kernels.stream_bounce_back(_staging.to_device(f_ikl), float(u0))
        _staging.write_back(f_ikl, out)
    else:
        out = kernels.stream_bounce_back(f_ikl.contiguous(), float(u0))
        f_ikl.copy_(out)
