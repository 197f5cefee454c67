"""Time to steady state, warm versus cold start, on held-out cavities.

For each cavity (Re, U drawn from a seed the training data did not use) the
plain cold-started FP64 LBM runs to max|u(t+every) - u(t)|/U < tol and is the
yardstick; the warm start (surrogate -> populations -> the same LBM, with the
controller's cold-start fallback) runs to the same threshold.  Reported per
cavity: steps and wall time of both (the warm time includes the network's
inference and the initializer), whether it fell back, and the converged-field
difference max and RMS |u_warm - u_cold| / U.  With ``--pooled`` a third run
starts from the cold result itself averaged onto the surrogate's grid: the
best any predictor on that grid could do.  With ``--tight T`` the cold run
continues past its stop until the residual is below T (e.g. 1e-8), and the
cold and warm results are also compared with that field: how far each run's
stopping point sits from the steady state, which bounds what the warm-cold
difference can resolve.

With several sizes (``--n 256,512,1024``) each cavity's N is drawn with its
(Re, U) from the seed, uniformly over the sizes.  Cavities whose cold start
does not converge (the BGK lattice diverges or stalls above the threshold)
have no yardstick: they are reported and left out of the summary.

Usage::

    python -m latticeboltzmann_amd.warmstart.evaluate --weights w.npz --count 40 --seed 1000 \
        --n 256 --out eval.json
    python -m latticeboltzmann_amd.warmstart.evaluate --weights default --n 256,512,1024 \
        --re-range 100 5000 --count 60 --seed 3000
"""

import argparse
import json
import sys
import time

import numpy as np
import torch

from .controller import Controller
from .data import draw_cavities, draw_config_cavities
from .features import pooled_target
from .solve import DEFAULT_WEIGHTS, CavityStepper, Predictor, load_default


class _Fixed:
    """Predictor stand-in returning a fixed field (the pooled cold result)."""

    def __init__(self, q, pool=2):
        self.q, self.pool = q, pool

    def predict(self, re, u, **kw):
        return self.q


def field_difference(a, b, U):
    d = torch.sqrt((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2)
    return float(d.max()) / U, float(torch.sqrt((d * d).mean())) / U


def evaluate_one(pred, n, re, u, ctl, neq=True, pooled=False, recon="cubic", tight=None):
    rec = {"n": n, "re": re, "u": u}
    ft = None
    with CavityStepper(n, re, u) as st:
        cold = ctl.run(st, warm=False)
        fc = tuple(t.clone() for t in st.fields())
        if tight is not None and cold["reason"] == "converged":
            # run on from the cold stop to a much smaller residual: the steady state
            # (at most four times the cold run's steps more)
            deep = Controller(tol=tight, every=ctl.every)
            status, extra, res = deep.settle(st, min(ctl.max_steps, 4 * cold["steps"]))
            ft = st.fields()
            rec["tight"] = {"tol": tight, "reason": status, "extra_steps": extra, "residual": res}
            rec["tight"]["cold_err_max"] = field_difference(fc, ft, u)[0]
    rec["cold"] = {k: cold[k] for k in ("steps", "seconds", "reason")}
    if cold["reason"] != "converged":
        return rec  # no converged yardstick to measure a warm start against
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
        if ft is not None:
            rec[name]["err_max_vs_tight"] = field_difference(fw, ft, u)[0]
    return rec


def summarize(records, legs=("warm",)):
    out = {"cavities": len(records),
           "cold_not_converged": [{"n": r["n"], "re": r["re"], "u": r["u"],
                                   "reason": r["cold"]["reason"], "steps": r["cold"]["steps"]}
                                  for r in records if r["cold"]["reason"] != "converged"]}
    for leg in legs:
        rs = [r[leg] for r in records if leg in r]
        if not rs:
            continue
        out.setdefault("by_n", {})
        for n in sorted({r["n"] for r in records if leg in r}):
            sub = [r[leg] for r in records if leg in r and r["n"] == n]
            out["by_n"]["%s_%d" % (leg, n)] = {
                "cavities": len(sub),
                "median_time_speedup": float(np.median([r["time_speedup"] for r in sub])),
                "median_step_speedup": float(np.median([r["step_speedup"] for r in sub])),
                "diff_max_worst": float(max(r["diff_max"] for r in sub)),
                "fallbacks": int(sum(r["fell_back"] for r in sub))}
        out[leg] = {
            "median_step_speedup": float(np.median([r["step_speedup"] for r in rs])),
            "median_time_speedup": float(np.median([r["time_speedup"] for r in rs])),
            "total_time_speedup": float(sum(r["cold"]["seconds"] for r in records if leg in r)
                                        / sum(r["seconds"] for r in rs)),
            "p10_time_speedup": float(np.percentile([r["time_speedup"] for r in rs], 10)),
            "min_time_speedup": float(min(r["time_speedup"] for r in rs)),
            "fallbacks": int(sum(r["fell_back"] for r in rs)),
            "diff_max_median": float(np.median([r["diff_max"] for r in rs])),
            "diff_max_worst": float(max(r["diff_max"] for r in rs)),
            "frac_diff_below_1e-4": float(np.mean([r["diff_max"] < 1e-4 for r in rs])),
        }
        # only cavities whose continued run settled below the tightened residual
        # have a steady state to compare with
        with_ref = [r for r in records if leg in r and "err_max_vs_tight" in r[leg]]
        tight = [(r["tight"]["cold_err_max"], r[leg]["err_max_vs_tight"]) for r in with_ref
                 if r["tight"]["reason"] == "converged"]
        if with_ref:
            out[leg]["vs_tight"] = {"cavities": len(tight),
                                    "no_steady_state": len(with_ref) - len(tight)}
        if tight:
            c, w = np.array(tight).T
            out[leg]["vs_tight"].update({
                "cold_err_max_median": float(np.median(c)), "cold_err_max_worst": float(c.max()),
                "err_max_median": float(np.median(w)), "err_max_worst": float(w.max()),
                "frac_err_below_1e-4": float(np.mean(w < 1e-4)),
                "frac_cold_err_below_1e-4": float(np.mean(c < 1e-4)),
                "frac_not_farther_than_cold": float(np.mean(w <= c))})
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS,
                    help="an .npz of one network, or 'default' for the shipped set")
    ap.add_argument("--count", type=int, default=40)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--n", default="256", help="lattice size, or a comma list to draw from")
    ap.add_argument("--re-range", type=float, nargs=2, default=None,
                    help="Re range of the draw (default: the predictor's)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-neq", action="store_true")
    ap.add_argument("--pooled", action="store_true")
    ap.add_argument("--recon", default="cubic", choices=("bilinear", "walls", "cubic"))
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--tight", type=float, default=None,
                    help="also run the cold start on to this residual and compare both results with it")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    pred = load_default() if a.weights == "default" else Predictor.load(a.weights)
    ctl = Controller(tol=a.tol)
    sizes = [int(x) for x in str(a.n).split(",")]
    re_range = tuple(a.re_range) if a.re_range else pred.re_range
    if len(sizes) == 1:
        cases = [(sizes[0], re, u) for re, u in
                 draw_cavities(a.count, a.seed, re_range, pred.u_range)]
    else:
        cases = draw_config_cavities(a.count, a.seed, sizes, re_range, pred.u_range)
    records, t0 = [], time.time()
    for n, re, u in cases:
        if a.time_limit and time.time() - t0 > a.time_limit:
            break
        rec = evaluate_one(pred, n, re, u, ctl, neq=not a.no_neq, pooled=a.pooled, recon=a.recon,
                           tight=a.tight)
        records.append(rec)
        print(json.dumps(rec), flush=True)
    summary = summarize(records, ("warm", "pooled") if a.pooled else ("warm",))
    print(json.dumps(summary), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump({"summary": summary, "records": records}, fh, indent=1)


if __name__ == "__main__":
    sys.exit(main())
