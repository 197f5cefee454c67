"""Drop-in replacement of the reference's ``PyLB`` package (``PyLB/__init__.py:21-22``).

Same names and call semantics as the reference:

* ``equilibrium`` and ``collide`` (``PyLB/Collision.py:21``, the pybind11
  module ``_lbkernels``, ``c/_lbkernels.cpp:33-49``);
* ``stream`` (``PyLB/Streaming.py:33-46``);
* ``stream_and_bounce_back`` (added here: the reference keeps it in its
  cavity drivers, ``cavity_opt1.py:177-245``).

Arguments may be numpy arrays (modified in place: staged to the GPU, computed
by the HIP kernels, copied back into the caller's memory) or torch tensors on
a ROCm device (computed in place, zero copy).  There is no CPU code path.

Use as ``from latticeboltzmann_amd import PyLB`` or, for unchanged driver
scripts, put ``latticeboltzmann_amd`` on ``PYTHONPATH`` and ``import PyLB``.
"""

from .Collision import collide, equilibrium
from .Streaming import stream, stream_and_bounce_back

__all__ = ["collide", "equilibrium", "stream", "stream_and_bounce_back"]
