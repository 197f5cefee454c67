"""Inputs and targets of the warm-start surrogate on its m x m cell grid.

Grid axes follow the lattice: axis 0 is x (index k, slow), axis 1 is y (index
l, fast); the lid is the row y = m - 1 and moves in +x.  Cell (i, j) covers
the nodes k in [i N/m, (i+1) N/m) of an N x N cavity, so its centre is the
node-centre position (i + 1/2) N/m.

Input channels (6): normalized log Re, normalized U, no-slip wall mask
(cells touching the left, right or bottom wall), lid mask (top row), and the
cell-centre coordinates x, y in [-1, 1].
Output channels (3): (rho - 1) / U^2, ux / U, uy / U.
"""

import numpy as np

GRID = 128
IN_CHANNELS = 6
OUT_CHANNELS = 3
RE_RANGE = (100.0, 1000.0)
U_RANGE = (0.05, 0.1)


def features(re, u, m=GRID, re_range=RE_RANGE, u_range=U_RANGE):
    """(m, m, 6) float32 input of one cavity."""
    lo, hi = np.log(re_range[0]), np.log(re_range[1])
    x = np.empty((m, m, IN_CHANNELS), dtype=np.float32)
    x[..., 0] = (2.0 * (np.log(re) - lo) / (hi - lo)) - 1.0
    x[..., 1] = (2.0 * (u - u_range[0]) / (u_range[1] - u_range[0])) - 1.0
    wall = np.zeros((m, m), dtype=np.float32)
    wall[0, :] = wall[-1, :] = wall[:, 0] = 1.0
    lid = np.zeros((m, m), dtype=np.float32)
    lid[:, -1] = 1.0
    x[..., 2] = wall
    x[..., 3] = lid
    c = (np.arange(m, dtype=np.float32) + 0.5) / m * 2.0 - 1.0
    x[..., 4] = c[:, None]
    x[..., 5] = c[None, :]
    return x


def pooled_target(rho, ux, uy, U, m=GRID):
    """Average the node fields of an N x N cavity (N a multiple of m) over the
    m x m cells and normalize: (m, m, 3) float32 as a torch tensor on the
    fields' device."""
    import torch
    n = rho.shape[0]
    if rho.shape != (n, n) or n % m:
        raise ValueError("fields must be N x N with N a multiple of %d" % m)
    r = n // m
    t = torch.stack([(rho - 1.0) / (U * U), ux / U, uy / U], dim=-1)
    return t.reshape(m, r, m, r, 3).mean(dim=(1, 3)).to(torch.float32)
