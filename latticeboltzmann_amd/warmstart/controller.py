"""Run a cavity to steady state from a warm start, with a cold-start fallback.

The controller only sees a *stepper*: an object with

* ``init_warm()`` / ``init_cold()``: set the populations (surrogate
  prediction, or rest at f_eq(1, 0) as ``cavity_opt1.py:333-335``);
* ``advance(n)``: n LBM steps;
* ``residual()``: max |u(now) - u(last call)| / U, FP64, and remember u(now).

Convergence: ``residual() < tol`` after a block of ``every`` steps, the
criterion of the cold start, so warm and cold runs stop at the same
threshold.  A warm start is abandoned for a cold start when

* the residual is not finite (the prediction drove the lattice unstable);
* it stalls: the best residual has not dropped below ``min_drop`` times its
  value for ``patience`` consecutive checks, or for half of the checks made
  so far when that is longer (on a large, high-Re cavity the residual
  oscillates while it falls, and new 10 % lows come 60-80 checks apart a
  few hundred checks into the run, so the window grows with the run);
* it runs over budget: more than ``budget`` times ``cold_estimate`` steps
  (when an estimate of the cold-start cost is given).

A stepper that cannot offer a warm start at all raises
``WarmStartUnavailable`` from ``init_warm``; the run is then a cold start
reported as a fallback.
"""

import math
import time


class WarmStartUnavailable(Exception):
    """Raised by a stepper's ``init_warm`` when it has no prediction to offer
    (for instance, Re or U outside the range the surrogate was trained on);
    the controller then runs the cold start."""


class Controller:
    def __init__(self, tol=1e-6, every=1000, patience=50, min_drop=0.9, budget=0.75,
                 max_steps=10 ** 7, patience_frac=0.5):
        if not (0 < min_drop < 1) or patience < 1 or every < 1 or not (0 <= patience_frac < 1):
            raise ValueError("need 0 < min_drop < 1, patience >= 1, every >= 1, "
                             "0 <= patience_frac < 1")
        self.tol, self.every, self.patience = float(tol), int(every), int(patience)
        self.patience_frac = float(patience_frac)
        self.min_drop, self.budget, self.max_steps = float(min_drop), float(budget), int(max_steps)

    def _converge(self, stepper, watch, limit):
        stepper.residual()
        best, best_at, k, steps, res = math.inf, 0, 0, 0, math.inf
        while steps < limit:
            stepper.advance(self.every)
            steps += self.every
            res = float(stepper.residual())
            if not math.isfinite(res):
                return "diverged", steps, res
            if res < self.tol:
                return "converged", steps, res
            if res < best * self.min_drop or best == math.inf:
                best, best_at = res, k
            elif watch and k - best_at >= max(self.patience, self.patience_frac * k):
                return "stalled", steps, res
            k += 1
        return ("over budget" if watch else "max steps"), steps, res

    def settle(self, stepper, max_steps=None):
        """Step an already initialised stepper until its residual is below
        ``tol``, with no stall test: (status, steps, residual)."""
        return self._converge(stepper, False, self.max_steps if max_steps is None else int(max_steps))

    def run(self, stepper, warm=True, cold_estimate=None):
        """Returns a dict: steps (total), warm_steps, cold_steps, fell_back,
        reason, residual, seconds (wall time including initialization)."""
        t0 = time.perf_counter()
        out = {"warm": bool(warm), "fell_back": False, "reason": None, "warm_steps": 0,
               "cold_steps": 0}
        if warm:
            try:
                stepper.init_warm()
            except WarmStartUnavailable:
                out["fell_back"], out["reason"] = True, "no warm start"
                warm = False
        if warm:
            limit = self.max_steps
            if cold_estimate:
                limit = min(limit, max(self.every, int(self.budget * cold_estimate)))
            status, steps, res = self._converge(stepper, True, limit)
            out["warm_steps"] = steps
            if status == "converged":
                out.update(steps=steps, residual=res, reason=status,
                           seconds=stepper.sync_time(t0))
                return out
            out["fell_back"], out["reason"] = True, status
        stepper.init_cold()
        status, steps, res = self._converge(stepper, False, self.max_steps)
        out["cold_steps"] = steps
        out.update(steps=out["warm_steps"] + steps, residual=res,
                   seconds=stepper.sync_time(t0))
        if out["reason"] is None:
            out["reason"] = status
        return out
