"""Velocity dumps of a running lattice: ``ux_{i}.npy`` / ``uy_{i}.npy``.

The reference's cavity drivers write the velocity field every ``dump_freq``
steps with ``save_mpiio`` (``cavity_opt1.py:345-347``, ``cavity_opt2.py:279-283``;
writer ``PyLB/IO.py:27-80``): after step ``i`` (0-based) with
``i % dump_freq == 0``, the moments ``collide`` returned, i.e. those of the
streamed, pre-collision state of that step.

Here the fused step writes those moments to device buffers on the dump steps
only (24 extra bytes per node in FP64).  A copy stream moves them to pinned
host memory while the next steps run, and a writer thread puts the files on
disk, so the lattice does not wait for PCIe or the file system until every
buffer slot is in flight.  Files are byte-identical to the reference's
single-process output (same header, same values: the moments are bit-exact
with the reference's ``collide`` return).

``run_cavity`` is the reference's main loop (``cavity_opt1.py:338-347``).  The
reference's final dump (``cavity_opt1.py:351-352``, ``cavity_opt2.py:287-288``)
names ``u_ckl``, which opt1 never defines (a ``NameError``) and opt2 leaves at
the last periodic dump; here the final dump writes the last step's
velocities as ``ux_{nsteps-1}`` / ``uy_{nsteps-1}``, which is what those lines
mean to do.
"""

import os
import queue
import threading

import torch

from .PyLB.IO import npy_header


class VelocityDumps:
    """Asynchronous writer of a lattice's velocity dumps.

    ``slot()`` hands out the device moment buffers for the next dump step
    (blocking while every slot is still in flight), ``submit(i)`` queues the
    dump of step ``i`` after the step that filled them, ``close()`` waits for
    every file and re-raises a writer error.  Use as a context manager.
    """

    def __init__(self, lat, directory=".", depth=3):
        if depth < 1:
            raise ValueError("depth must be >= 1")
        self.lat, self.directory = lat, directory
        os.makedirs(directory, exist_ok=True)
        shp, dt, dev = (lat.nx, lat.ny), lat.dtype, lat.device
        self._dev = [tuple(torch.empty(shp, dtype=dt, device=dev) for _ in range(3))
                     for _ in range(depth)]
        self._host = [tuple(torch.empty(shp, dtype=dt, pin_memory=True) for _ in range(2))
                      for _ in range(depth)]
        self._copied = [torch.cuda.Event() for _ in range(depth)]
        self._free = queue.Queue()
        for s in range(depth):
            self._free.put(s)
        self._copy_stream = torch.cuda.Stream(device=dev)
        self._jobs = queue.Queue()
        self._error = None
        self._current = None
        self.files = []
        self.bytes_written = 0
        self._thread = threading.Thread(target=self._writer, daemon=True)
        self._thread.start()

    # -- producer side (the stepping thread) ---------------------------------
    def slot(self):
        """Device (rho, ux, uy) buffers for the next dump step."""
        if self._current is not None:
            raise RuntimeError("slot() called twice without submit()")
        self._raise_error()
        s = self._free.get()
        self._current = s
        return self._dev[s]

    def submit(self, index):
        """Dump the buffers of the last ``slot()`` as step ``index``."""
        s, self._current = self._current, None
        if s is None:
            raise RuntimeError("submit() without slot()")
        filled = torch.cuda.Event()
        filled.record()  # after the step that wrote the moments (current stream)
        with torch.cuda.stream(self._copy_stream):
            self._copy_stream.wait_event(filled)
            _, ux, uy = self._dev[s]
            hx, hy = self._host[s]
            hx.copy_(ux, non_blocking=True)
            hy.copy_(uy, non_blocking=True)
            self._copied[s].record(self._copy_stream)
        # The slot returns to the free list only after the writer thread has
        # waited for this copy, so the next step that fills the slot's device
        # buffers is enqueued after the copy has read them.
        self._jobs.put((int(index), s))

    def close(self):
        self._jobs.put(None)
        self._thread.join()
        self._raise_error()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- writer thread --------------------------------------------------------
    def _writer(self):
        while True:
            job = self._jobs.get()
            if job is None:
                return
            index, s = job
            try:
                self._copied[s].synchronize()
                for name, t in zip(("ux", "uy"), self._host[s]):
                    fn = os.path.join(self.directory, "{}_{}.npy".format(name, index))
                    a = t.numpy()
                    with open(fn, "wb") as fh:
                        fh.write(npy_header(a.dtype, a.shape))
                        fh.write(memoryview(a).cast("B"))
                    self.files.append(fn)
                    self.bytes_written += a.nbytes
            except BaseException as e:  # reported by the next slot() or close()
                if self._error is None:
                    self._error = e
            finally:
                self._free.put(s)

    def _raise_error(self):
        if self._error is not None:
            e, self._error = self._error, None
            raise e


def dump_steps(nsteps, dump_freq, final=True):
    """0-based step indices whose velocities are dumped."""
    if dump_freq < 1:
        raise ValueError("dump_freq must be >= 1")
    idx = list(range(0, nsteps, dump_freq))
    if final and nsteps > 0 and idx[-1] != nsteps - 1:
        idx.append(nsteps - 1)
    return idx


def run_cavity(lat, nsteps, dump_freq=10000, directory=".", final=True, depth=3):
    """Run ``nsteps`` steps of ``lat`` writing the reference's velocity dumps
    (``cavity_opt1.py:338-347``); returns the list of files written.

    The lattice is stepped in chunks between dump steps; a dump step writes
    its pre-collision moments into a ``VelocityDumps`` slot.
    """
    with VelocityDumps(lat, directory, depth) as d:
        done = 0
        for i in dump_steps(nsteps, dump_freq, final):
            if i > done:
                lat.step(i - done)
            lat.step(1, moments=d.slot())
            d.submit(i)
            done = i + 1
        if nsteps > done:
            lat.step(nsteps - done)
    return d.files
