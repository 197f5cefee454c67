"""Training data for the warm-start surrogate: steady cavities from the plain LBM.

Each sample is a lid-driven cavity (N x N, FP64, ``cavity_opt1.py`` physics)
run from rest to max|u(t+every) - u(t)|/U < tol, then averaged onto the
surrogate's m x m cells and normalized (``features.pooled_target``).  Re is
drawn log-uniformly and U uniformly from their ranges.

Cavities that do not converge (the BGK lattice diverges, or the residual
has not reached ``tol`` within ``max_steps``) are recorded with their reason
in the output (``skipped_*`` arrays) and left out of the targets: they are
the evidence of where the scheme stops settling.

Usage::

    python -m latticeboltzmann_amd.warmstart.data --count 400 --seed 0 --n 256 \
        --out cavities.npz [--re-range 100 5000] [--dtype float16] [--time-limit 900]
"""

import argparse
import json
import sys
import time

import numpy as np
import torch

from .controller import Controller
from .features import GRID, RE_RANGE, U_RANGE, pooled_target
from .solve import CavityStepper


def draw_cavities(count, seed, re_range=RE_RANGE, u_range=U_RANGE):
    """[(Re, U)] for ``count`` cavities from ``seed``."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(re_range[0]), np.log(re_range[1])
    re = np.exp(rng.uniform(lo, hi, size=count))
    u = rng.uniform(u_range[0], u_range[1], size=count)
    return [(float(a), float(b)) for a, b in zip(re, u)]


CONFIG_SIZES = (256, 512, 1024)
CONFIG_RE_RANGE = (100.0, 5000.0)


def draw_config_cavities(count, seed, sizes=CONFIG_SIZES, re_range=CONFIG_RE_RANGE,
                         u_range=U_RANGE):
    """[(N, Re, U)]: N uniform over ``sizes``, Re log-uniform, U uniform, from ``seed``."""
    rng = np.random.default_rng(seed)
    n = rng.choice(np.asarray(sizes), size=count)
    lo, hi = np.log(re_range[0]), np.log(re_range[1])
    re = np.exp(rng.uniform(lo, hi, size=count))
    u = rng.uniform(u_range[0], u_range[1], size=count)
    return [(int(a), float(b), float(c)) for a, b, c in zip(n, re, u)]


def steady_cavity(n, re, u, tol=1e-6, every=1000, max_steps=4_000_000):
    """Cold-start an N x N cavity to steady state; returns (run record, (rho, ux, uy))."""
    ctl = Controller(tol=tol, every=every, max_steps=max_steps)
    with CavityStepper(n, re, u) as st:
        rec = ctl.run(st, warm=False)
        fields = tuple(t.clone() for t in st.fields())
    return rec, fields


def save(path, params, n, records, targets, dtype, skipped=()):
    np.savez(path, re=np.array([p[0] for p in params]), u=np.array([p[1] for p in params]),
             n=np.full(len(params), n), steps=np.array([r["steps"] for r in records]),
             seconds=np.array([r["seconds"] for r in records]),
             residual=np.array([r["residual"] for r in records]),
             targets=(np.stack(targets) if targets else np.zeros((0, GRID, GRID, 3))).astype(dtype),
             skipped_re=np.array([s[0] for s in skipped], dtype=np.float64),
             skipped_u=np.array([s[1] for s in skipped], dtype=np.float64),
             skipped_steps=np.array([s[2] for s in skipped], dtype=np.int64),
             skipped_reason=np.array([s[3] for s in skipped], dtype="U16"))


def generate(count, seed=0, n=256, m=GRID, tol=1e-6, every=1000, max_steps=4_000_000, out=None,
             dtype="float32", time_limit=None, log=None, re_range=RE_RANGE, u_range=U_RANGE):
    params = draw_cavities(count, seed, re_range, u_range)
    done, records, targets, skipped = [], [], [], []
    t0 = time.time()
    for i, (re, u) in enumerate(params):
        if time_limit and time.time() - t0 > time_limit:
            break
        rec, (rho, ux, uy) = steady_cavity(n, re, u, tol, every, max_steps)
        if rec["reason"] != "converged":
            skipped.append((re, u, rec["steps"], rec["reason"]))
            if log:
                log(json.dumps({"sample": i, "re": re, "u": u, "skipped": rec["reason"],
                                "steps": rec["steps"], "seconds": round(rec["seconds"], 3)}))
            continue
        done.append((re, u))
        records.append(rec)
        targets.append(pooled_target(rho, ux, uy, u, m).cpu().numpy())
        if log:
            log(json.dumps({"sample": i, "re": round(re, 2), "u": round(u, 4), "steps": rec["steps"],
                            "seconds": round(rec["seconds"], 3)}))
        if out and len(done) % 20 == 0:
            save(out, done, n, records, targets, dtype, skipped)
    if out and (done or skipped):
        save(out, done, n, records, targets, dtype, skipped)
    return done, records, targets


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--count", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-steps", type=int, default=4_000_000)
    ap.add_argument("--dtype", default="float32", choices=("float32", "float16"))
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--re-range", type=float, nargs=2, default=RE_RANGE)
    ap.add_argument("--u-range", type=float, nargs=2, default=U_RANGE)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    torch.cuda.init()
    generate(a.count, a.seed, a.n, tol=a.tol, max_steps=a.max_steps, out=a.out, dtype=a.dtype,
             time_limit=a.time_limit, log=lambda s: print(s, flush=True),
             re_range=tuple(a.re_range), u_range=tuple(a.u_range))


if __name__ == "__main__":
    sys.exit(main())
