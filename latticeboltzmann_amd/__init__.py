"""MI355X-native D2Q9 lattice Boltzmann (gfx950 HIP kernels behind a C ABI).

Layers
------
``csrc/`` + ``include/lbm_d2q9.h``
    hand-written HIP kernels and the C ABI (``liblbm_d2q9.so``).
``_native``
    ctypes binding of the C ABI; fails loudly if the library or GPU is absent.
``kernels``
    typed calls on torch device tensors (zero copy).
``PyLB``
    drop-in for the reference's ``PyLB`` package (numpy or device arrays).
``lattice``
    device-resident solver object (two ping-pong population buffers).
"""

__version__ = "0.1.0"


def lattice(*args, **kwargs):
    """Shortcut for :class:`latticeboltzmann_amd.lattice.Lattice`."""
    from .lattice import Lattice
    return Lattice(*args, **kwargs)
