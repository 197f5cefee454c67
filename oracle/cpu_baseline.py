"""Times the CPU oracle on one core: the CPU baseline that bench.py reports.

TEST / MEASUREMENT INFRASTRUCTURE ONLY (see oracle/__init__.py).

Runs in its own process pinned to one core, with BLAS/OpenMP threads set to
one, so the figure is a single-core rate.  Legs, per workload:

* ``numpy``: the numpy opt1 restatement (cavity_opt1.py:100-245;
  shear_wave_opt1.py:170-172 for the periodic wave), the reference's
  optimised pure-numpy flavour;
* ``opt2``: the structure of cavity_opt2.py:272-277 (shear_wave_opt2.py:
  95-97 for the wave), numpy streaming (with the walls for the cavity)
  followed by the C restatement of _lbkernels.collide (d2q9_colliden,
  c/d2q9.h:121-131; single thread, -O2, as the reference's setup.py builds
  it);
* ``c``: the single-thread C restatement of the fused step (-O2);
* ``opt2_native``: the opt2 leg with the collide compiled -O3 -march=native
  on the host that runs this (the variant the measurement plan reports
  beside -O2), built into a temporary directory; skipped without a compiler.

Workloads (``--workload``): ``cavity`` (an n x n lid-driven cavity from rest,
U = 0.1, at the given omega, each leg for up to ``--budget`` seconds or
``--max-steps`` steps), ``shear256`` (the 256 x 256 periodic shear wave,
omega 1, u_x = 0.1 sin(2 pi y / 256), 1000 steps) and ``crop512`` (the
512 x 512 cavity at Re 1000, omega 1.529988, 200 steps).

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


def legs(f0, step_np, c_bc, wall_stream, omega, u0, budget, max_steps, native):
    """Time the four legs on copies of the initial state f0 (9, nx, ny)."""
    _, nx, ny = f0.shape
    f = f0.copy()
    step_np(f)  # warm the page cache / allocator with one step of each leg
    steps_np, t_np = time_leg(step_np, f, budget, max_steps)

    ctypes_ref.lib()
    g = np.ascontiguousarray(f0.copy())
    h = np.empty_like(g)

    def c_step(state):
        a, b = state
        ctypes_ref.lib().oracle_step_f64(ctypes_ref._ptr(a), ctypes_ref._ptr(b), nx, ny, omega,
                                         c_bc, u0, None, None, None)
        state.reverse()

    state = [g, h]
    c_step(state)
    steps_c, t_c = time_leg(c_step, state, budget, max_steps)

    def opt2_step(g2):
        wall_stream(g2)
        ctypes_ref.lbk_collide(g2.reshape(9, -1), omega)

    f2 = f0.copy()
    opt2_step(f2)
    steps_o2, t_o2 = time_leg(opt2_step, f2, budget, max_steps)

    native_rec = None
    if native is not None:
        def opt2_native_step(g2):
            wall_stream(g2)
            native(ctypes_ref._ptr(g2), g2[0].size, omega)

        f3 = f0.copy()
        opt2_native_step(f3)
        steps_n, t_n = time_leg(opt2_native_step, f3, budget, max_steps)
        native_rec = {"steps": steps_n, "seconds": t_n, "mlups": nx * ny * steps_n / t_n / 1e6}

    return {
        "grid": [nx, ny],
        "numpy_opt1": {"steps": steps_np, "seconds": t_np,
                       "mlups": nx * ny * steps_np / t_np / 1e6},
        "c_port": {"steps": steps_c, "seconds": t_c, "mlups": nx * ny * steps_c / t_c / 1e6},
        "opt2": {"steps": steps_o2, "seconds": t_o2, "mlups": nx * ny * steps_o2 / t_o2 / 1e6},
        "opt2_native": native_rec,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cavity", choices=("cavity", "shear256", "crop512"))
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--omega", type=float, default=0.897343)
    ap.add_argument("--budget", type=float, default=8.0, help="seconds per leg (cavity)")
    ap.add_argument("--max-steps", type=int, default=50)
    args = ap.parse_args()

    cores = 1
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass

    native = native_collide()
    u0 = 0.1
    if args.workload == "shear256":
        omega = 1.0
        f0 = d2q9_numpy.shear_wave_init(256, 256, 0.1, "x")
        rec = legs(f0, lambda g: d2q9_numpy.periodic_step(g, omega), 0, d2q9_numpy.stream,
                   omega, 0.0, float("inf"), 1000, native)
        rec["workload"] = "shear wave 256x256, omega 1, A 0.1, 1000 steps"
    else:
        if args.workload == "crop512":
            n, omega, budget, max_steps = 512, 1.529988, float("inf"), 200
            rec_name = "cavity 512x512 (Re 1000 crop), omega 1.529988, 200 steps"
        else:
            n, omega, budget, max_steps = args.n, args.omega, args.budget, args.max_steps
            rec_name = "cavity %dx%d, omega %.6f, up to %d steps or %.0f s per leg" % (
                n, n, omega, max_steps, budget)
        f0 = d2q9_numpy.cavity_init(n, n)
        rec = legs(f0, lambda g: d2q9_numpy.cavity_step(g, omega, u0), 1,
                   lambda g: d2q9_numpy.stream_and_bounce_back(g, u0), omega, u0, budget,
                   max_steps, native)
        rec["workload"] = rec_name
    rec.update(cores=cores, cpu_count=os.cpu_count(), cpu_model=cpu_model())
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
