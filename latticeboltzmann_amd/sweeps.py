"""Precision sweep: the same problem in FP64 and with FP32/BF16/FP16 deviation storage.

* Cavity (default 8192 x 8192, Re 1000, U 0.1, 50000 steps): velocity error
  of each storage format against the FP64 run (RMS and max of |u - u_64| / U),
  and total-mass drift against the FP64 run and against the initial mass.
* Shear wave (default 4096 x 4096, omega 1, amplitude 0.01, 50000 steps):
  viscosity fitted from the amplitude's decay, A(t) ~ A0 exp(-nu k^2 t),
  against (1/omega - 1/2)/3, per format including FP64, each storage format
  with the FP32 collide and with an FP64 collide.
* The FP16 cavity is also run with the FP64 collide, and FP16/BF16 storing
  every 3 and every 4 steps (intermediate steps unrounded on chip).

Usage::

    python -m latticeboltzmann_amd.sweeps [--cavity-n 8192] [--shear-n 4096] [--steps 50000] \
        [--no-shear] [--out sweep.json]
"""

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .batch import BatchedLattices
from .lattice import Lattice, omega_for_cavity
from .precision import DeviationLattice, velocity_errors

FORMATS = ("f32", "bf16", "f16")


def _timed(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


def cavity_precision(n=8192, re=1000.0, u=0.1, steps=50000, formats=FORMATS, log=None,
                     store_every=(3, 4)):
    om = omega_for_cavity(n, re, u)
    ref = Lattice(n, n, om, "cavity", u, torch.float64, variant="auto")
    m0 = ref.mass()
    _, sec = _timed(lambda: ref.step(steps))
    _, uxr, uyr = ref.moments()
    m64 = ref.mass()
    out = {"n": n, "re": re, "u": u, "omega": om, "steps": steps,
           "f64": {"seconds": sec, "mlups": n * n * steps / sec / 1e6, "mass_drift": m64 / m0 - 1}}
    del ref
    torch.cuda.empty_cache()
    if log:
        log(json.dumps({"cavity": "f64", **out["f64"]}))
    runs = [(f, "f32", 1) for f in formats] + [("f16", "f64", 1)]
    runs += [(f, "f32", k) for k in store_every for f in ("bf16", "f16")]
    for fmt, compute, every in runs:
        lat = DeviationLattice(n, n, om, "cavity", u, fmt=fmt, variant="auto", compute=compute,
                               store_every=every)
        _, sec = _timed(lambda: lat.step(steps))
        _, ux, uy = lat.moments()
        rec = velocity_errors(ux, uy, uxr, uyr, u)
        m = lat.mass()
        rec.update(seconds=sec, mlups=n * n * steps / sec / 1e6, mass_drift=m / m0 - 1,
                   mass_vs_f64=m / m64 - 1, compute=compute, store_every=every)
        key = fmt if compute == "f32" else fmt + "_f64_collide"
        if every > 1:
            key += "_store_every_%d" % every
        out[key] = rec
        if log:
            log(json.dumps({"cavity": key, **rec}))
        del lat, ux, uy
        torch.cuda.empty_cache()
    return out


def _fit(times, amps, n):
    times, amps = np.asarray(times, dtype=np.float64), np.asarray(amps, dtype=np.float64)
    keep = times >= times[1]
    slope = np.polyfit(times[keep], np.log(amps[keep]), 1)[0]
    return -slope / (2 * math.pi / n) ** 2


def shear_precision(n=4096, omega=1.0, amp=0.01, steps=50000, every=2500, formats=FORMATS, log=None):
    theory = (1.0 / omega - 0.5) / 3.0
    out = {"n": n, "omega": omega, "amplitude": amp, "steps": steps, "nu_theory": theory}
    # FP64: the fused kernel, amplitude by projection on the initial sine
    lat = Lattice(n, n, omega, "periodic", 0.0, torch.float64, variant="auto")
    ly = torch.arange(n, dtype=torch.float64, device=lat.device)
    s = torch.sin(2 * math.pi * ly / n)
    lat.init_equilibrium(ux=(amp * s)[None, :])
    times, amps = [], []
    t0 = time.perf_counter()
    for c in range(steps // every + 1):
        _, ux, _ = lat.moments()
        times.append(lat.time)
        amps.append(float((ux * s[None, :]).sum().item()) * 2.0 / (n * n))
        if c < steps // every:
            lat.step(every)
    torch.cuda.synchronize()
    nu = _fit(times, amps, n)
    out["f64"] = {"nu": nu, "nu_rel_err": nu / theory - 1, "seconds": time.perf_counter() - t0}
    del lat
    torch.cuda.empty_cache()
    if log:
        log(json.dumps({"shear": "f64", **out["f64"]}))
    for fmt in formats:
        for compute in ("f32", "f64"):
            bat = BatchedLattices([dict(nx=n, ny=n, omega=omega, amplitude=amp, direction="x",
                                        bc="periodic")], fmt=fmt, compute=compute)
            (times, amps), sec = _timed(lambda: bat.run(steps, sample_every=every, graph=True))
            nu = _fit(times, amps[:, 0], n)
            key = fmt if compute == "f32" else fmt + "_f64_collide"
            out[key] = {"nu": nu, "nu_rel_err": nu / theory - 1, "seconds": sec,
                        "mlups": n * n * steps / sec / 1e6, "compute": compute}
            if log:
                log(json.dumps({"shear": key, **out[key]}))
            del bat
            torch.cuda.empty_cache()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--cavity-n", type=int, default=8192)
    ap.add_argument("--shear-n", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--shear-every", type=int, default=2500)
    ap.add_argument("--no-shear", action="store_true", help="run the cavity sweep only")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    res = {"cavity": cavity_precision(a.cavity_n, steps=a.steps, log=log)}
    if not a.no_shear:
        res["shear"] = shear_precision(a.shear_n, steps=a.steps, every=a.shear_every, log=log)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(res, fh, indent=1)


if __name__ == "__main__":
    sys.exit(main())
