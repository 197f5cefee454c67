"""Benchmark of the fused D2Q9 stream-collide kernels on one MI355X.

Workload (the line's ``value``): the 2048 x 2048 lid-driven cavity in FP64,
Re = 1000, U = 0.1 (omega = 1/(3 U N / Re + 1/2) = 0.897343), started from
rho = 1, u = 0 at f_eq, K steps.  A step is one pull stream + walls + collide
pass over the whole grid.  Runs of four steps go through one launch of the
multi-step kernel (temporal blocking on 22 x 32 tiles: the earlier steps of a
tile and its halo stay in LDS), which computes the bits of as many single-step
launches.  Populations stay
in HBM; the timed region is K steps bracketed by device synchronisation,
timed with HIP events on the stream the kernels are launched on.

Also reported (``extra``): the single-step kernel on the same workload; an
8192 x 8192 cavity (past the 256 MB last-level cache) in FP64 and with
FP32/BF16/FP16 deviation storage (BF16 and FP16 in their default form, stored
every 4 steps, and rounded every step); a device-to-device copy of the 8192^2 FP64
byte count (torch's, and the library's own 16-byte copy, the practical
ceiling of a single-pass bandwidth-bound kernel); the ghost-layer sliding-lid
flow at 4098^2; the run again with the reference's velocity dumps; the
batched sweep of 512 shear-wave instances (seed 0, FP32 storage, 2000 steps,
four steps per launch, graph-captured) with its fitted-viscosity errors; and
the warm-start time to steady state against the cold start on held-out
256 x 256 cavities and on a sample drawn as the warm-start workload draws
them (N 256/512/1024, Re 100-5000).

``roofline``: the dominant kernel is the K-step launch (K steps per launch,
the earlier steps kept on chip), which moves about 1/K of a single pass's
bytes and is limited by its FP64 arithmetic instead.  Its hardware counters
(a committed summary under profiles/, taken with the same build and kernel
variant) show the VALU as the busiest unit, so the roofline is the VALU issue
rate: peak = 256 CUs x 64 lanes x 2.4 GHz = 39.3 T lane-ops/s (MI355X's
published 78.6 TFLOP/s FP64 vector rate, which counts an FMA as two flops);
achieved = the algorithmic FP64 work (128 FP64 VALU instructions per lattice
update: the opt1 node arithmetic without contraction, two IEEE divides,
counted in the single-step kernel's code) x lattice updates per launch /
average launch time, measured with HIP events on the launch stream.
``traffic`` is the HBM bytes per launch from FETCH_SIZE + WRITE_SIZE.
``extra.roofline_detail`` gives the counter view at the live launch time: the
fraction of VALU issue slots used (recomputed halo and index arithmetic
included) and the HBM rate.

``cpu_baseline``: the reference's CPU paths restated by the oracle, timed on
one host core on a bounded sample of the same 2048^2 cavity: ``value`` is the
opt2 structure (numpy walls + the C++ _lbkernels collide restated in C), the
fastest; the numpy opt1 restatement and a fused C step are in ``sample``;
``legs`` adds the 256^2 shear wave (1000 steps) and the 512^2 cavity crop
(200 steps), every leg timed in full.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

HBM_PEAK_GBS = 8000.0
BYTES_PER_LUP = {"f64": 144, "f32": 72, "bf16": 36, "f16": 36}
# one VALU operation per lane per clock: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
# FP64 VALU instructions of one lattice update of the bit-exact opt1 node
# arithmetic (pull, moments, two IEEE divides, nine equilibria, relaxation; no
# fused multiply-adds), counted in the compiled single-step kernel
FP64_OPS_PER_LUP = 128


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000, help="timed steps (default: the "
                    "20000 steps of the 2048^2 cavity workload)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--variant", default="auto", help="launch shape of the single-step kernel "
                    "(an integer, or 'auto' for the autotuner, which then also picks the "
                    "multi-step kernel's shape)")
    ap.add_argument("--multi-variant", default="default",
                    help="shape (steps per launch, tile) of the multi-step kernel: an integer, "
                         "'default' (the library's default shape, the one the committed "
                         "counter summary covers) or 'auto' (autotuned)")
    ap.add_argument("--single-step", action="store_true",
                    help="time the single-step kernel only (no multi-step launches)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the 8192^2 and copy legs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start leg")
    ap.add_argument("--no-dumps", action="store_true", help="skip the velocity-dump leg")
    ap.add_argument("--warm-count", type=int, default=6)
    ap.add_argument("--cpu-budget", type=float, default=8.0)
    ap.add_argument("--copy-only", action="store_true",
                    help="only time lbm_relayout (a plain copy of the 9*n*n FP64 populations) "
                         "--steps times: the known-size copy that scales the HBM counters")
    return ap.parse_args()


def copy_only(n, reps):
    """Time the relayout kernel on an n x n FP64 population array (read + write 144 B/node)."""
    import torch
    from latticeboltzmann_amd import _native, kernels
    a = torch.zeros((9, n, n), dtype=torch.float64, device="cuda")
    b = torch.empty((n, 9, n), dtype=torch.float64, device="cuda")
    kernels.relayout(a, _native.LAYOUT_SOA, _native.LAYOUT_ROWS, out=b)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        kernels.relayout(a, _native.LAYOUT_SOA, _native.LAYOUT_ROWS, out=b)
    end.record()
    torch.cuda.synchronize()
    sec = start.elapsed_time(end) / 1e3 / reps
    nbytes = 2 * a.numel() * 8
    print(json.dumps({"copy": "lbm_relayout", "n": n, "bytes_per_launch": nbytes,
                      "us_per_launch": round(sec * 1e6, 2), "gbs": round(nbytes / sec / 1e9, 1)}))


def timed_steps(lat, steps, warmup):
    import torch
    lat.step(warmup)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    lat.step(steps)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3  # seconds


def copy_ceiling(nbytes, reps=20):
    """Best HBM rate of lbm_copy (16-byte words, plain or non-temporal stores) on
    nbytes, in GB/s counting bytes read plus written: the practical ceiling of a
    bandwidth-bound kernel on this device."""
    import torch
    from latticeboltzmann_amd import _native
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    best = 0.0
    for nt in (0, 1):
        def run():
            _native.call("lbm_copy", a.data_ptr(), b.data_ptr(), nbytes, nt, _native.stream_ptr())
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        best = max(best, 2 * nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9)
    del a, b
    torch.cuda.empty_cache()
    return best


def copy_bandwidth(nbytes):
    """GB/s of a device-to-device copy moving ``nbytes`` read + ``nbytes`` written."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    start.record()
    for _ in range(reps):
        b.copy_(a)
    end.record()
    torch.cuda.synchronize()
    sec = start.elapsed_time(end) / 1e3 / reps
    del a, b
    return 2 * nbytes / sec / 1e9


def kernel_profile(name, variant=None):
    """A committed counter summary (profiles/<name>), per variant when it holds
    several, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            j = json.load(fh)
    except (OSError, ValueError):
        return None
    if variant is not None and "per_variant" in j:
        return j["per_variant"].get(str(variant))
    return j


def _cpu_leg(args):
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py")] + list(args)
    out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=900).stdout
    return json.loads(out.strip().splitlines()[-1])


def _leg_summary(rec):
    keep = ("numpy_opt1", "opt2", "opt2_native", "c_port")
    return {k: {"mlups": round(rec[k]["mlups"], 3), "steps": rec[k]["steps"],
                "seconds": round(rec[k]["seconds"], 2)} for k in keep if rec.get(k)}


def cpu_baseline(n, budget, workloads=True):
    """The reference's CPU paths (restated by the oracle) on one host core.

    ``value`` is the opt2 structure on the bench's own workload (the n x n
    cavity, a bounded sample of its steps); with ``workloads`` the 256^2
    shear wave (1000 steps) and the 512^2 cavity crop (200 steps) are timed
    in full beside it, every leg of each."""
    rec = _cpu_leg(["--n", str(n), "--budget", str(budget), "--max-steps", "40"])
    out = {
        # the reference's fastest CPU path, restated: cavity_opt2.py's numpy
        # walls plus the C++ collide of _lbkernels (d2q9_colliden), one core
        "value": round(rec["opt2"]["mlups"], 4),
        "unit": "MLUPS",
        "cores": rec["cores"],
        "kind": "port",
        "sample": "opt2 structure of the %dx%d FP64 cavity step (numpy stream_and_bounce_back + C "
                  "restatement of _lbkernels.collide, -O2), %d steps (%.1f s) on one core of %s "
                  "(%d visible); the same with the collide built -O3 -march=native: %s MLUPS; "
                  "numpy opt1 restatement: %.3f MLUPS; fused C restatement -O2: %.3f MLUPS"
                  % (n, n, rec["opt2"]["steps"], rec["opt2"]["seconds"], rec["cpu_model"],
                     rec["cpu_count"],
                     "%.3f" % rec["opt2_native"]["mlups"] if rec.get("opt2_native") else "n/a",
                     rec["numpy_opt1"]["mlups"], rec["c_port"]["mlups"]),
        "numpy_opt1_mlups": round(rec["numpy_opt1"]["mlups"], 4),
        "legs": {"cavity_%d" % n: _leg_summary(rec)},
    }
    if workloads:
        for name in ("shear256", "crop512"):
            r = _cpu_leg(["--workload", name])
            out["legs"][name] = dict(_leg_summary(r), workload=r["workload"])
    return out


def batch_sweep(steps=2000):
    """512 shear-wave instances (seed 0) in FP32 storage: throughput and nu errors."""
    import numpy as np
    import torch
    from latticeboltzmann_amd.batch import BatchedLattices, shear_wave_sweep
    bat = BatchedLattices(shear_wave_sweep(512, 0), fmt="f32")
    bat.run(20, sample_every=10, graph=True)  # tuning, capture and first-touch outside the timing
    bat = BatchedLattices(shear_wave_sweep(512, 0), fmt="f32")
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    times, amp = bat.run(steps, sample_every=50, graph=True)
    b.record()
    torch.cuda.synchronize()
    sec = a.elapsed_time(b) / 1e3
    nodes = int(sum(bat.sizes))
    rel = np.abs(bat.viscosities(times, amp) / bat.theory() - 1)
    out = {"instances": 512, "nodes": nodes, "steps": steps, "seconds": round(sec, 4),
           "mlups": round(nodes * steps / sec / 1e6, 1), "storage": "f32", "compute": bat.compute_name,
           "nu_rel_err_median": float(np.nanmedian(rel)), "nu_rel_err_p90": float(np.nanpercentile(rel, 90)),
           "nu_rel_err_max": float(np.nanmax(rel))}
    del bat
    torch.cuda.empty_cache()
    return out


def dumped_runs(n, omega, steps=20000, freqs=(10000, 1000)):
    """The cavity run again with the reference's velocity dumps (ux_i.npy,
    uy_i.npy every dump_freq steps and at the last step) written to a
    temporary directory; wall time against a plain run of the same steps
    timed the same way (the workload's 20000 steps, whatever --steps is)."""
    import shutil
    import tempfile
    import torch
    from latticeboltzmann_amd.dumps import run_cavity
    from latticeboltzmann_amd.lattice import Lattice
    lat = Lattice(n, n, omega, "cavity", 0.1, torch.float64, variant="auto")
    lat.step(20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat.step(steps)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    del lat
    out = {"steps": steps, "plain_seconds": round(plain_s, 4)}
    for freq in freqs:
        lat = Lattice(n, n, omega, "cavity", 0.1, torch.float64, variant="auto")
        lat.step(20)
        torch.cuda.synchronize()
        d = tempfile.mkdtemp(prefix="lbm_dumps_")
        try:
            t0 = time.perf_counter()
            files = run_cavity(lat, steps, freq, d)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
            nbytes = sum(os.path.getsize(f) for f in files)
        finally:
            shutil.rmtree(d, ignore_errors=True)
        out["every_%d" % freq] = {"files": len(files), "bytes": nbytes, "seconds": round(sec, 4),
                                  "mlups": round(n * n * steps / sec / 1e6, 1),
                                  "overhead_vs_plain": round(sec / plain_s - 1.0, 4)}
        del lat
        torch.cuda.empty_cache()
    return out


def ghost_flow(n=4096, steps=200):
    """The sliding-lid script's flow (ghost-layer walls, slidingLid.py) on
    (n + 2)^2 nodes in FP64: one fused ghost-layer step per launch."""
    import torch
    from latticeboltzmann_amd import ghost
    lat = ghost.GhostLattice(n + 2, n + 2, ghost.sliding_lid_omega(n, 1000, 0.1),
                             ghost.cavity_walls(n + 2, n + 2, 0.1))
    s = timed_steps(lat, steps, 20)
    nodes = (n + 2) ** 2
    m = nodes * steps / s / 1e6
    gbs = BYTES_PER_LUP["f64"] * m * 1e6 / 1e9
    ok = bool(torch.isfinite(lat.f).all())
    del lat
    torch.cuda.empty_cache()
    return {"nodes": nodes, "steps": steps, "mlups": round(m, 1), "hbm_gbs_algorithmic": round(gbs, 1),
            "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 4), "finite": ok}


def precision_check(n, omega, steps):
    """The headline cavity run again in FP16 storage, in its default form
    (stored every 4 steps, ``f16``), stored every 3 steps and rounded every
    step, against FP64: velocity errors over U, mass drift."""
    import torch
    from latticeboltzmann_amd.lattice import Lattice
    from latticeboltzmann_amd.precision import DeviationLattice, velocity_errors
    ref = Lattice(n, n, omega, "cavity", 0.1, torch.float64)
    m0 = ref.mass()
    ref.step(steps)
    _, uxr, uyr = ref.moments()
    del ref
    out = {"n": n, "steps": steps}
    for key, every in (("f16_each_step", 1), ("f16_store_every_3", 3), ("f16", 4)):
        lat = DeviationLattice(n, n, omega, "cavity", 0.1, fmt="f16", variant="auto",
                               store_every=every)
        sec = timed_steps(lat, steps, 0)
        _, ux, uy = lat.moments()
        rec = velocity_errors(ux, uy, uxr, uyr, 0.1)
        rec.update(mlups=round(n * n * steps / sec / 1e6, 1), mass_drift=lat.mass() / m0 - 1)
        out[key] = rec
        del lat, ux, uy
    torch.cuda.empty_cache()
    return out


def warm_start(count, n=256, seed=2000, mix_count=6, mix_seed=5000, mix_budget_s=150.0):
    """Held-out cavities: cold vs warm-started steps and wall time to steady
    state, warm starts from the shipped set of networks.  ``n256``: ``count``
    cavities at N = 256 (Re 100-1000).  ``config_mix``: up to ``mix_count``
    cavities drawn as the warm-start workload draws them (N uniform over
    256/512/1024, Re log-uniform in [100, 5000], U uniform in [0.05, 0.1]),
    stopping once ``mix_budget_s`` seconds have been spent; draws whose cold
    start diverges are listed and left out of the summary.  For the mix the
    cold run also continues to a residual of 1e-8 (at most 4 x its steps
    more), and both stopping points are compared with that field
    (``vs_tight``, per cavity ``*_err_vs_steady_over_U``): how far each
    run's stop sits from the steady state itself."""
    import time as _time
    from latticeboltzmann_amd.warmstart.controller import Controller
    from latticeboltzmann_amd.warmstart.data import draw_cavities, draw_config_cavities
    from latticeboltzmann_amd.warmstart.evaluate import evaluate_one, summarize
    from latticeboltzmann_amd.warmstart.solve import DEFAULT_WEIGHTS, Predictor, load_default
    if not os.path.exists(DEFAULT_WEIGHTS):
        return None
    ctl = Controller(tol=1e-6, every=1000)

    def table(recs):
        return [{"n": r["n"], "re": round(r["re"], 1), "u": round(r["u"], 4),
                 "cold_steps": r["cold"]["steps"], "cold_reason": r["cold"]["reason"],
                 "warm_steps": r["warm"]["steps"] if "warm" in r else None,
                 "cold_s": round(r["cold"]["seconds"], 3),
                 "warm_s": round(r["warm"]["seconds"], 3) if "warm" in r else None,
                 "diff_max_over_U": r["warm"]["diff_max"] if "warm" in r else None,
                 **({"cold_err_vs_steady_over_U": r["tight"]["cold_err_max"],
                     "warm_err_vs_steady_over_U": r["warm"].get("err_max_vs_tight"),
                     "steady_residual": r["tight"]["residual"]} if "tight" in r else {})}
                for r in recs]

    low = Predictor.load(DEFAULT_WEIGHTS)
    full = load_default()
    recs = [evaluate_one(full, n, re, u, ctl) for re, u in
            draw_cavities(count, seed, low.re_range, low.u_range)]
    out = {"n256": dict(summarize(recs)["warm"], cavities=table(recs),
                        criterion="max|u(t+1000)-u(t)|/U < 1e-6, FP64, N=%d" % n)}
    recs, t0 = [], _time.perf_counter()
    for nn, re, u in draw_config_cavities(mix_count, mix_seed, re_range=(100.0, 5000.0)):
        if _time.perf_counter() - t0 > mix_budget_s:
            break
        recs.append(evaluate_one(full, nn, re, u, ctl, tight=1e-8))
    s = summarize(recs)
    mix = dict(s.get("warm", {}), cavities=table(recs), cold_not_converged=s["cold_not_converged"])
    out["config_mix"] = mix
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the domain of every workload fits one MI355X); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from latticeboltzmann_amd import _native
    from latticeboltzmann_amd.lattice import Lattice, omega_for_cavity

    _native.require_gpu()
    _native.load()
    n = args.n
    if args.copy_only:
        copy_only(n, args.steps)
        return
    omega = omega_for_cavity(n, 1000.0, 0.1)
    if args.variant != "auto":
        args.variant = int(args.variant)
    lat = Lattice(n, n, omega, "cavity", 0.1, torch.float64, variant=args.variant,
                  multi_step=not args.single_step)
    if not args.single_step and args.multi_variant != "auto":
        lat.set_multi_variant(-1 if args.multi_variant == "default" else int(args.multi_variant))
    lat_variant, multi_var, multi_k = lat.variant, lat.multi_variant, lat.multi_k
    launches = lat.launches(args.steps)
    sec = timed_steps(lat, args.steps, args.warmup)
    mlups = n * n * args.steps / sec / 1e6
    launch_s = sec / launches
    lups_per_launch = n * n * args.steps / launches
    ok = bool(torch.isfinite(lat.f).all())
    del lat
    torch.cuda.empty_cache()
    if multi_var is not None:
        dominant = "%d steps per launch, multi-step variant %d" % (multi_k, multi_var)
        prof = kernel_profile("roofline_f64_cavity_%d.json" % n, "multi%d" % multi_var)
    else:
        dominant = "single step, variant %d" % lat_variant
        prof = kernel_profile("roofline_f64_cavity_%d.json" % n, "single%d" % lat_variant)
    alg_tops = FP64_OPS_PER_LUP * lups_per_launch / launch_s / 1e12
    traffic = prof.get("hbm_bytes_per_launch") if prof else None
    roofline = {"bound": "valu", "achieved": round(alg_tops, 3), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "T VALU lane-op/s", "frac": round(alg_tops / VALU_PEAK_TOPS, 4),
                "traffic": traffic}
    detail = {"kernel": prof.get("kernel") if prof else None,
              "fp64_ops_per_lattice_update": FP64_OPS_PER_LUP,
              "lattice_updates_per_launch": int(lups_per_launch),
              "launch_us": round(launch_s * 1e6, 2),
              # HBM bytes a launch must move: one read and one write of every
              # population per launch, however many steps it makes
              "compulsory_bytes_per_launch": BYTES_PER_LUP["f64"] * n * n,
              "compulsory_gbs": round(BYTES_PER_LUP["f64"] * n * n / launch_s / 1e9, 1),
              # what a single pass per step would have to move at this rate
              "single_pass_equivalent_gbs": round(
                  BYTES_PER_LUP["f64"] * lups_per_launch / launch_s / 1e9, 1)}
    if prof:
        detail["valu_issue_frac_from_counters"] = round(
            prof["valu_wave_insts_per_launch"] * 64 / launch_s / 1e12 / VALU_PEAK_TOPS, 4)
        detail["valu_lane_ops_per_lattice_update_from_counters"] = round(
            prof["valu_wave_insts_per_launch"] * 64 / lups_per_launch, 1)
        detail["hbm_gbs_from_counters"] = round(traffic / launch_s / 1e9, 1)
        detail["hbm_frac_from_counters"] = round(traffic / launch_s / 1e9 / HBM_PEAK_GBS, 4)
        detail["profiled_launch_us"] = prof.get("kernel_trace_avg_us")
        if "bench_command_trace" in prof:  # this command's own timed loop under the kernel trace
            detail["profiled_launch_us_this_command"] = prof["bench_command_trace"]["timed_region_avg_us"]

    extra = {"finite_state": ok, "omega": omega, "variant": lat_variant,
             "multi_variant": multi_var, "steps_per_launch": multi_k, "dominant_kernel": dominant,
             "launches": launches, "roofline_detail": detail}
    if multi_var is not None and not args.no_extra:
        lat = Lattice(n, n, omega, "cavity", 0.1, torch.float64, variant=args.variant,
                      multi_step=False)
        s1 = timed_steps(lat, 2000, 50)
        del lat
        torch.cuda.empty_cache()
        m1 = n * n * 2000 / s1 / 1e6
        extra["single_step_%d" % n] = {"mlups": round(m1, 1),
                                       "hbm_gbs_algorithmic": round(144 * m1 / 1e3, 1),
                                       "frac_of_8tbs": round(144 * m1 / 1e3 / HBM_PEAK_GBS, 4)}
    if not args.no_extra:
        from latticeboltzmann_amd.precision import DeviationLattice
        big = 8192
        om_big = omega_for_cavity(big, 1000.0, 0.1)
        legs = {}
        # the deviation formats in their default form (BF16 and FP16 stored
        # every 4 steps) and rounding every step
        for name, fmt, every in (("f64", None, None), ("f32", "f32", None), ("bf16", "bf16", None),
                                 ("f16", "f16", None), ("bf16_each_step", "bf16", 1),
                                 ("f16_each_step", "f16", 1)):
            if fmt is None:
                lat = Lattice(big, big, om_big, "cavity", 0.1, torch.float64, variant=args.variant,
                              multi_step=not args.single_step)
            else:
                lat = DeviationLattice(big, big, om_big, "cavity", 0.1, fmt=fmt, variant="auto",
                                       store_every=every)
            steps = 200
            s = timed_steps(lat, steps, 20)
            m = big * big * steps / s / 1e6
            k = lat.multi_k
            legs[name] = {"mlups": round(m, 1), "steps": steps, "variant": lat.variant,
                          "multi_variant": lat.multi_variant, "steps_per_launch": k,
                          "store_every": getattr(lat, "store_every", 1),
                          # one read and one write of every population per launch
                          "compulsory_gbs": round(BYTES_PER_LUP[fmt or "f64"] / k * m / 1e3, 1)}
            del lat
            torch.cuda.empty_cache()
        lat = Lattice(big, big, om_big, "cavity", 0.1, torch.float64, variant=args.variant,
                      multi_step=False)
        s = timed_steps(lat, 200, 20)
        del lat
        torch.cuda.empty_cache()
        m = big * big * 200 / s / 1e6
        legs["f64_single_step"] = {"mlups": round(m, 1), "hbm_gbs_algorithmic": round(144 * m / 1e3, 1),
                                   "frac_of_8tbs": round(144 * m / 1e3 / HBM_PEAK_GBS, 4)}
        # FP16 storage against FP64 with temporal blocking (the fastest FP64
        # form) and against the plain single-pass FP64 kernel; the default
        # FP16 form ("f16") stores every 4 steps, here also every 3
        for every in (3,):
            lat = DeviationLattice(big, big, om_big, "cavity", 0.1, fmt="f16", variant="auto",
                                   store_every=every)
            s = timed_steps(lat, 48 * every, 4 * every)
            del lat
            torch.cuda.empty_cache()
            me = big * big * 48 * every / s / 1e6
            key = "f16_store_every_%d" % every
            legs[key] = {"mlups": round(me, 1), "steps": 48 * every, "steps_per_launch": every,
                         "compulsory_gbs": round(36 / every * me / 1e3, 1)}
            legs[key + "_over_f64_mlups"] = round(me / legs["f64"]["mlups"], 3)
        legs["f16_over_f64_mlups"] = round(legs["f16"]["mlups"] / legs["f64"]["mlups"], 3)
        legs["f16_each_step_over_f64_mlups"] = round(legs["f16_each_step"]["mlups"]
                                                     / legs["f64"]["mlups"], 3)
        legs["f16_over_f64_single_step_mlups"] = round(legs["f16"]["mlups"] / m, 3)
        extra["cavity_8192"] = legs
        nbytes = 9 * big * big * 8
        extra["device_copy_gbs_same_bytes"] = round(copy_bandwidth(nbytes), 1)
        ceiling = copy_ceiling(nbytes)
        extra["copy_ceiling_gbs"] = round(ceiling, 1)
        extra["f64_8192_single_step_over_copy_ceiling"] = round(
            legs["f64_single_step"]["hbm_gbs_algorithmic"] / ceiling, 3)
        extra["ghost_sliding_lid_4098"] = ghost_flow()
    if not args.no_dumps:
        extra["dumps_%d" % n] = dumped_runs(n, omega)
    if not args.no_extra:
        extra["precision_%d" % n] = precision_check(n, omega, args.steps)
    if not args.no_batch:
        extra["batch_shear_512"] = batch_sweep()
    if not args.no_warm:
        extra["warm_start"] = warm_start(args.warm_count)

    line = {
        "metric": "MLUPS + % HBM roofline, FP64 vs FP16-store; time-to-steady-state cavity",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (cavity from rest)",
        "config": {"workload": "lid-driven cavity %dx%d FP64 D2Q9 BGK, Re=1000, U=0.1, "
                               "omega=%.6f, fused pull stream-collide" % (n, n, omega),
                   "nx": n, "ny": n, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": None,
        "extra": extra,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
