"""Times the CPU oracle on one core: the CPU baseline that bench.py reports.

TEST / MEASUREMENT INFRASTRUCTURE ONLY (see oracle/__init__.py).

Runs in its own process pinned to one core, with BLAS/OpenMP threads set to
one, so the figure is a single-core rate.  Three legs:

* ``numpy``: the numpy opt1 restatement (cavity_opt1.py:100-245), the
  reference's optimised pure-numpy flavour;
* ``opt2``: the structure of cavity_opt2.py:272-277, numpy
  stream_and_bounce_back followed by the C restatement of _lbkernels.collide
  (d2q9_colliden, c/d2q9.h:121-131; single thread, -O2, as the reference's
  setup.py builds it);
* ``c``: the single-thread C restatement of the fused step (-O2);
* ``opt2_native``: the opt2 leg with the collide compiled -O3 -march=native
  on the host that runs this (the variant the measurement plan reports
  beside -O2), built into a temporary directory; skipped without a compiler.

Prints one JSON object: MLUPS per leg, steps timed, grid, core count used.
"""

import argparse
import json
import os
import platform
import sys
import time

for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[var] = "1"

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import ctypes_ref, d2q9_numpy  # noqa: E402


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def native_collide():
    """oracle_lbk_collide_f64 from d2q9_ref.c built -O3 -march=native for this
    host (in a temporary directory), or None when no compiler is available."""
    import ctypes
    import shutil
    import subprocess
    import tempfile
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        return None
    out = os.path.join(tempfile.mkdtemp(prefix="lbm_oracle_native_"), "liboracle_native.so")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "d2q9_ref.c")
    r = subprocess.run([cc, "-O3", "-march=native", "-ffp-contract=off", "-fPIC", "-shared",
                        "-std=c11", "-o", out, src], capture_output=True)
    if r.returncode:
        return None
    fn = ctypes.CDLL(out).oracle_lbk_collide_f64
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double]
    return fn


def time_leg(fn, f, budget_s, max_steps):
    """Run single steps until ``budget_s`` seconds or ``max_steps`` steps."""
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps:
        fn(f)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    return steps, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--omega", type=float, default=0.897343)
    ap.add_argument("--budget", type=float, default=8.0, help="seconds per leg")
    ap.add_argument("--max-steps", type=int, default=50)
    args = ap.parse_args()

    cores = 1
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass

    n = args.n
    f = d2q9_numpy.cavity_init(n, n)
    # warm the page cache / allocator with one step of each leg
    d2q9_numpy.cavity_step(f, args.omega, 0.1)
    steps_np, t_np = time_leg(lambda g: d2q9_numpy.cavity_step(g, args.omega, 0.1), f,
                              args.budget, args.max_steps)

    ctypes_ref.lib()
    g = np.ascontiguousarray(f)
    h = np.empty_like(g)

    def c_step(state):
        a, b = state
        ctypes_ref.lib().oracle_step_f64(ctypes_ref._ptr(a), ctypes_ref._ptr(b), n, n,
                                         args.omega, 1, 0.1, None, None, None)
        state.reverse()

    state = [g, h]
    c_step(state)
    steps_c, t_c = time_leg(c_step, state, args.budget, args.max_steps)

    def opt2_step(g2):
        d2q9_numpy.stream_and_bounce_back(g2, 0.1)
        ctypes_ref.lbk_collide(g2.reshape(9, -1), args.omega)

    f2 = d2q9_numpy.cavity_init(n, n)
    opt2_step(f2)
    steps_o2, t_o2 = time_leg(opt2_step, f2, args.budget, args.max_steps)

    native = native_collide()
    native_rec = None
    if native is not None:
        def opt2_native_step(g2):
            d2q9_numpy.stream_and_bounce_back(g2, 0.1)
            native(ctypes_ref._ptr(g2), g2[0].size, args.omega)

        f3 = d2q9_numpy.cavity_init(n, n)
        opt2_native_step(f3)
        steps_n, t_n = time_leg(opt2_native_step, f3, args.budget, args.max_steps)
        native_rec = {"steps": steps_n, "seconds": t_n, "mlups": n * n * steps_n / t_n / 1e6}

    print(json.dumps({
        "grid": [n, n],
        "numpy_opt1": {"steps": steps_np, "seconds": t_np,
                       "mlups": n * n * steps_np / t_np / 1e6},
        "c_port": {"steps": steps_c, "seconds": t_c, "mlups": n * n * steps_c / t_c / 1e6},
        "opt2": {"steps": steps_o2, "seconds": t_o2, "mlups": n * n * steps_o2 / t_o2 / 1e6},
        "opt2_native": native_rec,
        "cores": cores,
        "cpu_count": os.cpu_count(),
        "cpu_model": cpu_model(),
    }))


if __name__ == "__main__":
    main()
