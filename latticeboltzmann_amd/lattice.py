"""Device-resident D2Q9 solver: populations stay in HBM across steps.

This is the engine behind the reference's driver loops
(``cavity_opt1.py:338-347``, ``shear_wave_opt1.py:170-174``): two population
buffers ping-pong through the fused pull stream-collide kernel, one launch
per step, on one stream.  The MPI decomposition and halo exchange of the
reference (``cavity_opt1.py:247-329``) have no counterpart: one MI355X holds
the whole domain.
"""

import numpy as np
import torch

from . import _native as nat
from . import kernels

_TORCH_DTYPES = {"float64": torch.float64, "float32": torch.float32,
                 np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                 torch.float64: torch.float64, torch.float32: torch.float32}

BCS = {"periodic": nat.BC_PERIODIC, "cavity": nat.BC_CAVITY}


def omega_for_cavity(n, reynolds, u_lid):
    """omega = 1 / (3 nu + 1/2) with nu = U N / Re (lattice units)."""
    nu = u_lid * n / reynolds
    return 1.0 / (3.0 * nu + 0.5)


class Lattice:
    """A D2Q9 BGK lattice of ``nx x ny`` nodes resident on one GPU.

    Parameters
    ----------
    nx, ny : grid size (x slow, y fast, as the reference's ``f_ikl``).
    omega : relaxation parameter.
    bc : ``"cavity"`` (walls + moving lid at y = ny-1) or ``"periodic"``.
    u0 : lid velocity (cavity only).
    dtype : float64 or float32 storage and arithmetic.
    variant : launch shape of the fused step (-1 = library default).
    """

    def __init__(self, nx, ny, omega, bc="cavity", u0=0.1, dtype=torch.float64, device=None,
                 variant=-1):
        nat.require_gpu()
        if bc not in BCS:
            raise ValueError("bc must be 'cavity' or 'periodic'")
        self.nx, self.ny = int(nx), int(ny)
        self.omega, self.bc, self.u0 = float(omega), bc, float(u0)
        self.dtype = _TORCH_DTYPES[dtype]
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.variant = int(variant)
        shape = (9, self.nx, self.ny)
        self.f = torch.empty(shape, dtype=self.dtype, device=self.device)
        self._g = torch.empty(shape, dtype=self.dtype, device=self.device)
        self.time = 0
        self.init_equilibrium()

    # -- state ---------------------------------------------------------------
    def init_equilibrium(self, rho=None, ux=None, uy=None):
        """f = f_eq(rho, u); defaults rho = 1, u = 0 (``cavity_opt1.py:333-335``)."""
        shp = (self.nx, self.ny)

        def field(x, default):
            if x is None:
                return torch.full(shp, default, dtype=self.dtype, device=self.device)
            t = torch.as_tensor(x, dtype=self.dtype, device=self.device)
            return t.expand(shp).contiguous()

        kernels.equilibrium(field(rho, 1.0), field(ux, 0.0), field(uy, 0.0), out=self.f)
        self.time = 0
        return self

    def load(self, f):
        """Replace the populations by ``f`` (numpy or tensor of shape (9, nx, ny))."""
        t = torch.as_tensor(np.asarray(f) if not isinstance(f, torch.Tensor) else f)
        if tuple(t.shape) != (9, self.nx, self.ny):
            raise ValueError("populations must have shape (9, %d, %d)" % (self.nx, self.ny))
        self.f.copy_(t.to(self.dtype))
        return self

    def populations(self):
        """Populations as a numpy array (a device-to-host copy)."""
        return self.f.cpu().numpy()

    # -- stepping ------------------------------------------------------------
    def step(self, n=1, moments=None):
        """Advance ``n`` steps.  ``moments=(rho, ux, uy)`` tensors receive the
        pre-collision moments of the last step."""
        bc = BCS[self.bc]
        for s in range(n):
            mom = moments if (moments is not None and s == n - 1) else None
            kernels.step(self.f, self._g, self.omega, bc, self.u0, moments=mom,
                         variant=self.variant)
            self.f, self._g = self._g, self.f
        self.time += n
        return self

    # -- observables ---------------------------------------------------------
    def moments(self):
        """(rho, ux, uy) of the current populations, as device tensors."""
        return kernels.moments(self.f)

    def mass(self):
        """Total mass sum_i,k,l f, accumulated in FP64 on the device."""
        return float(kernels.total_sum(self.f).item())

    def velocity(self):
        """(ux, uy) as numpy arrays."""
        _, ux, uy = self.moments()
        return ux.cpu().numpy(), uy.cpu().numpy()
