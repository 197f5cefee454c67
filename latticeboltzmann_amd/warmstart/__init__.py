"""Learned warm start for the lid-driven cavity.

A small convolutional encoder-decoder (``network.Surrogate``), trained on the
GPU by the project's own BF16-MFMA forward and backward kernels on steady
cavities the FP64 LBM itself produces (``data``, ``train``), predicts
(rho, u) on a 128 x 128 grid from (Re, U, wall mask); the HIP initializer
interpolates it to the N x N lattice and sets the populations
(``solve.warm_populations``); the ``controller.Controller`` runs the LBM to
steady state and falls back to a cold start when the warm start stalls,
diverges or runs over budget.  The cold-started LBM stays the yardstick for the
converged field (``evaluate``).  The reference repository has no learned
counterpart.
"""

from .controller import Controller, WarmStartUnavailable
from .features import GRID, RE_RANGE, U_RANGE, features, pooled_target

__all__ = ["Controller", "WarmStartUnavailable", "GRID", "RE_RANGE", "U_RANGE", "features",
           "pooled_target"]
