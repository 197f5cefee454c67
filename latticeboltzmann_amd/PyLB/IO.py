"""npy writer of the reference's ``PyLB.IO`` (``PyLB/IO.py:27-80``), one process.

``save_mpiio(comm, fn, g_kl)`` writes a two-dimensional array to ``fn`` in
the npy format, version 1.0, with the header the reference composes
(``PyLB/IO.py:49-60``):

* the magic string ``\\x93NUMPY\\x01\\x00``;
* the header length as a little-endian 16-bit integer;
* the text of the Python dict ``{'descr': ..., 'fortran_order': False,
  'shape': (nx, ny)}``, padded with spaces and ended by a newline so that
  magic + length + text is a multiple of 16 bytes;
* the array in C order.

``numpy.load`` reads the file.  The header differs from ``numpy.save``'s
(which adds a trailing ``, `` to the dict and pads to 64 bytes), so the
bytes are the reference's, not numpy's.

The reference writes through an MPI-IO file view with one block per rank
(``PyLB/IO.py:62-80``).  Here one process holds the whole domain, so the view
is the whole file: ``comm`` may be ``None`` or a communicator of size 1.
``g_kl`` may be a numpy array or a torch tensor (a device tensor is copied to
the host first).

The reference calls ``np.asscalar`` (``PyLB/IO.py:55``), which numpy 1.23
removed; the shape is written from Python ints here, which gives the same
text.
"""

import struct

import numpy as np

MAGIC = b"\x93NUMPY\x01\x00"  # npy format version 1.0


def npy_header(dtype, shape):
    """Header bytes (magic, length, padded dict text) for a C-order array."""
    nx, ny = (int(s) for s in shape)
    text = str({"descr": np.lib.format.dtype_to_descr(np.dtype(dtype)),
                "fortran_order": False,
                "shape": (nx, ny)})
    pad = (15 - (len(text) + len(MAGIC) + 2)) % 16
    text = text + " " * pad + "\n"
    return MAGIC + struct.pack("<H", len(text)) + text.encode("latin-1")


def _host_array(g_kl):
    try:
        import torch
        if isinstance(g_kl, torch.Tensor):
            return g_kl.detach().cpu().contiguous().numpy()
    except ImportError:  # pragma: no cover - torch is part of the install
        pass
    return np.ascontiguousarray(g_kl)


def _single_rank(comm):
    if comm is None:
        return
    size = comm.Get_size() if hasattr(comm, "Get_size") else getattr(comm, "size", 1)
    if int(size) != 1:
        raise NotImplementedError("save_mpiio: one process holds the domain; got a communicator "
                                  "of size %d" % int(size))


def save_mpiio(comm, fn, g_kl):
    """Write the two-dimensional array ``g_kl`` to ``fn`` (npy, version 1.0).

    Same arguments as the reference (``PyLB/IO.py:27``); ``comm`` is ``None``
    or a communicator of size 1.
    """
    _single_rank(comm)
    a = _host_array(g_kl)
    if a.ndim != 2:
        raise ValueError("save_mpiio writes two-dimensional arrays, got shape %s" % (a.shape,))
    with open(fn, "wb") as fh:
        fh.write(npy_header(a.dtype, a.shape))
        fh.write(memoryview(a).cast("B"))
