"""Time to steady state, warm versus cold start, on held-out cavities.

For each cavity (Re, U drawn from a seed the training data did not use) the
plain cold-started FP64 LBM runs to max|u(t+every) - u(t)|/U < tol and is the
yardstick; the warm start (surrogate -> populations -> the same LBM, with the
controller's cold-start fallback) runs to the same threshold.  Reported per
cavity: steps and wall time of both (the warm time includes the network's
inference and the initializer), whether it fell back, and the converged-field
difference max and RMS |u_warm - u_cold| / U.  With ``--pooled`` a third run
starts from the cold result itself averaged onto the surrogate's grid: the
best any predictor on that grid could do.

Usage::

    python -m latticeboltzmann_amd.warmstart.evaluate --weights w.npz --count 40 --seed 1000 \
        --n 256 --out eval.json
"""

import argparse
import json
import sys
import time

import numpy as np
import torch

from .controller import Controller
from .data import draw_cavities
from .features import pooled_target
from .solve import DEFAULT_WEIGHTS, CavityStepper, Predictor


class _Fixed:
    """Predictor stand-in returning a fixed field (the pooled cold result)."""

    def __init__(self, q, pool=2):
        self.q, self.pool = q, pool

    def predict(self, re, u):
        return self.q


def field_difference(a, b, U):
    d = torch.sqrt((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2)
    return float(d.max()) / U, float(torch.sqrt((d * d).mean())) / U


def evaluate_one(pred, n, re, u, ctl, neq=True, pooled=False, recon="cubic"):
    rec = {"n": n, "re": re, "u": u}
    with CavityStepper(n, re, u) as st:
        cold = ctl.run(st, warm=False)
        fc = tuple(t.clone() for t in st.fields())
    rec["cold"] = {k: cold[k] for k in ("steps", "seconds", "reason")}
    legs = [("warm", pred)]
    if pooled:
        m = pred.grid if pred else 128
        legs.append(("pooled", _Fixed(pooled_target(*fc, u, m), n // m)))
    for name, p in legs:
        with CavityStepper(n, re, u, predictor=p, neq=neq, recon=recon) as st:
            r = ctl.run(st, warm=True)
            fw = st.fields()
            dmax, drms = field_difference(fw, fc, u)
        rec[name] = {k: r[k] for k in ("steps", "seconds", "fell_back", "reason", "warm_steps")}
        rec[name].update(diff_max=dmax, diff_rms=drms,
                         step_speedup=cold["steps"] / r["steps"],
                         time_speedup=cold["seconds"] / r["seconds"])
    return rec


def summarize(records, legs=("warm",)):
    out = {"cavities": len(records)}
    for leg in legs:
        rs = [r[leg] for r in records if leg in r]
        if not rs:
            continue
        out[leg] = {
            "median_step_speedup": float(np.median([r["step_speedup"] for r in rs])),
            "median_time_speedup": float(np.median([r["time_speedup"] for r in rs])),
            "total_time_speedup": float(sum(r["cold"]["seconds"] for r in records if leg in r)
                                        / sum(r["seconds"] for r in rs)),
            "min_time_speedup": float(min(r["time_speedup"] for r in rs)),
            "fallbacks": int(sum(r["fell_back"] for r in rs)),
            "diff_max_median": float(np.median([r["diff_max"] for r in rs])),
            "diff_max_worst": float(max(r["diff_max"] for r in rs)),
            "frac_diff_below_1e-4": float(np.mean([r["diff_max"] < 1e-4 for r in rs])),
        }
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--count", type=int, default=40)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-neq", action="store_true")
    ap.add_argument("--pooled", action="store_true")
    ap.add_argument("--recon", default="cubic", choices=("bilinear", "walls", "cubic"))
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    pred = Predictor.load(a.weights)
    ctl = Controller(tol=a.tol)
    records, t0 = [], time.time()
    for i, (re, u) in enumerate(draw_cavities(a.count, a.seed, pred.re_range, pred.u_range)):
        if a.time_limit and time.time() - t0 > a.time_limit:
            break
        rec = evaluate_one(pred, a.n, re, u, ctl, neq=not a.no_neq, pooled=a.pooled, recon=a.recon)
        records.append(rec)
        print(json.dumps(rec), flush=True)
    summary = summarize(records, ("warm", "pooled") if a.pooled else ("warm",))
    print(json.dumps(summary), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump({"summary": summary, "records": records}, fh, indent=1)


if __name__ == "__main__":
    sys.exit(main())
