"""Inputs and targets of the warm-start surrogate on its m x m cell grid.

Grid axes follow the lattice: axis 0 is x (index k, slow), axis 1 is y (index
l, fast); the lid is the row y = m - 1 and moves in +x.  Cell (i, j) covers
the nodes k in [i N/m, (i+1) N/m) of an N x N cavity, so its centre is the
node-centre position (i + 1/2) N/m.

Input channels (6 + 4 * octaves): normalized log Re and U in [-1, 1], the
no-slip wall mask (cells touching the left, right or bottom wall), the lid
mask (top row), the cell-centre coordinates x, y in [-1, 1], then per octave
j sin and cos of pi 2^j times each normalized scalar.  The network's GEMMs
take BF16 operands, which resolve a value near 1 only to 2^-8: the log Re
channel alone pins Re to ~0.3 %; the octaves carry the finer digits.
Output channels (3): (rho - 1) / U^2, ux / U, uy / U.
"""

import numpy as np

GRID = 128
IN_CHANNELS = 6
OUT_CHANNELS = 3
RE_RANGE = (100.0, 1000.0)
U_RANGE = (0.05, 0.1)


def in_channels(octaves=0):
    return IN_CHANNELS + 4 * int(octaves)


def features(re, u, m=GRID, re_range=RE_RANGE, u_range=U_RANGE, octaves=0):
    """(m, m, 6 + 4 octaves) float32 input of one cavity."""
    lo, hi = np.log(re_range[0]), np.log(re_range[1])
    a = (2.0 * (np.log(re) - lo) / (hi - lo)) - 1.0
    b = (2.0 * (u - u_range[0]) / (u_range[1] - u_range[0])) - 1.0
    x = np.empty((m, m, in_channels(octaves)), dtype=np.float32)
    x[..., 0] = a
    x[..., 1] = b
    wall = np.zeros((m, m), dtype=np.float32)
    wall[0, :] = wall[-1, :] = wall[:, 0] = 1.0
    lid = np.zeros((m, m), dtype=np.float32)
    lid[:, -1] = 1.0
    x[..., 2] = wall
    x[..., 3] = lid
    c = (np.arange(m, dtype=np.float32) + 0.5) / m * 2.0 - 1.0
    x[..., 4] = c[:, None]
    x[..., 5] = c[None, :]
    for j in range(int(octaves)):
        w = np.pi * 2.0 ** j
        x[..., IN_CHANNELS + 4 * j:IN_CHANNELS + 4 * j + 4] = [np.sin(w * a), np.cos(w * a),
                                                               np.sin(w * b), np.cos(w * b)]
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
