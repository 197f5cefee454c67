"""The precision sweep and the batched sweep at their full workload sizes.

Precision: the 8192 x 8192 lid-driven cavity (Re 1000, U 0.1) in FP64 and
with FP32, BF16 and FP16 deviation storage (FP32 collide; FP16 also stored
every 4 steps), 5000 steps, errors against the FP64 run of the same problem;
the 4096 x 4096 shear wave (omega 1, amplitude 0.01, 50000 steps), fitted
viscosity against (1/omega - 1/2)/3 (``shear_wave_decay.py:162,167``, the
projection of ``shear_wave_opt1.py:174``).

Batched: the 512 shear-wave instances drawn from seed 0 (FP32 storage, 2000
steps, graph-captured), per-instance fitted viscosity against the analytic
value; a few small batched instances, one step, against the FP64 oracle step
of their decoded stored state.

The bounds are stated per result; the values measured on MI355X are printed
(``-s``) and quoted beside each bound.
"""

import math

import numpy as np
import pytest
import torch

from oracle import d2q9_numpy as o
from oracle import lowprec as lp

pytestmark = pytest.mark.gpu

CAVITY_N, CAVITY_STEPS = 8192, 5000


@pytest.fixture(scope="module")
def cavity_fp64(gpu):
    from latticeboltzmann_amd.lattice import Lattice, omega_for_cavity
    om = omega_for_cavity(CAVITY_N, 1000.0, 0.1)
    print("cavity %d^2 FP64, %d steps" % (CAVITY_N, CAVITY_STEPS), flush=True)
    ref = Lattice(CAVITY_N, CAVITY_N, om, "cavity", 0.1, torch.float64)
    m0 = ref.mass()
    ref.step(CAVITY_STEPS)
    _, ux, uy = ref.moments()
    m64 = ref.mass()
    del ref
    torch.cuda.empty_cache()
    yield {"omega": om, "ux": ux, "uy": uy, "m0": m0, "m64": m64}


# (format, store_every): bounds on velocity L2 / U, L-inf / U, the L2 error
# relative to the L2 norm of the FP64 velocity (not diluted by the resting
# part of the cavity), and |mass / mass_64 - 1|; measured on MI355X (5000 steps):
#   f32 5.5e-8, 3.0e-7, 7.8e-7, 3e-10;   bf16 1.7e-4, 1.1e-2, 2.5e-3, 5e-9;
#   f16 2.3e-5, 1.3e-3, 3.3e-4, 3e-10;   f16 stored every 4 steps 1.1e-5, 5.5e-4, 1.5e-4, 1.6e-9;
#   bf16 stored every 4 steps 8.4e-5, 4.3e-3, 1.2e-3, 1.4e-9 (every 4 steps is the 16-bit default).
# FP16 is held to the target of a velocity L2 error below 1e-3 U with room to spare.
CAVITY_BOUNDS = {
    ("f32", 1): (5e-7, 5e-6, 5e-6, 1e-9),
    ("bf16", 1): (1e-3, 5e-2, 1.5e-2, 1e-7),
    ("f16", 1): (1e-4, 5e-3, 1.5e-3, 1e-8),
    ("f16", 4): (5e-5, 3e-3, 8e-4, 1e-8),
    ("bf16", 4): (5e-4, 2.5e-2, 7e-3, 3e-8),
}


@pytest.mark.parametrize("fmt,every", list(CAVITY_BOUNDS))
def test_cavity_8192_storage_against_fp64(cavity_fp64, fmt, every):
    from latticeboltzmann_amd.precision import DeviationLattice, velocity_errors
    ref = cavity_fp64
    print("cavity %d^2 %s store_every %d, %d steps" % (CAVITY_N, fmt, every, CAVITY_STEPS), flush=True)
    lat = DeviationLattice(CAVITY_N, CAVITY_N, ref["omega"], "cavity", 0.1, fmt=fmt,
                           store_every=every)
    lat.step(CAVITY_STEPS)
    _, ux, uy = lat.moments()
    err = velocity_errors(ux, uy, ref["ux"], ref["uy"], 0.1)
    norm = float(torch.sqrt(torch.mean(ref["ux"] ** 2 + ref["uy"] ** 2)).item())
    err["l2_relative"] = err["l2_over_U"] * 0.1 / norm
    err["mass_vs_f64"] = lat.mass() / ref["m64"] - 1
    print("  ", fmt, every, err, flush=True)
    del lat, ux, uy
    torch.cuda.empty_cache()
    l2, linf, rel, mass = CAVITY_BOUNDS[(fmt, every)]
    assert err["l2_over_U"] < l2
    assert err["linf_over_U"] < linf
    assert err["l2_relative"] < rel
    assert abs(err["mass_vs_f64"]) < mass
    # the FP64 run itself: the lid corners leak mass, per step
    # 6 w u0 (rho_lid(nx-1) - rho_lid(0)) (cavity_opt1.py:206-213), tiny here
    assert abs(ref["m64"] / ref["m0"] - 1) < 1e-6


SHEAR_N, SHEAR_STEPS, SHEAR_EVERY, SHEAR_AMP = 4096, 50000, 2500, 0.01


def _fit_nu(times, amps, n):
    # least squares of log A against t from the second sample on: the start
    # f = f_eq has no non-equilibrium part, so the first steps decay differently
    times, amps = np.asarray(times, dtype=np.float64)[1:], np.asarray(amps, dtype=np.float64)[1:]
    slope = np.polyfit(times, np.log(amps), 1)[0]
    return -slope / (2 * math.pi / n) ** 2


def test_shear_wave_4096_fp64_viscosity(gpu):
    from latticeboltzmann_amd.lattice import Lattice
    n, omega = SHEAR_N, 1.0
    print("shear wave %d^2 FP64, %d steps" % (n, SHEAR_STEPS), flush=True)
    lat = Lattice(n, n, omega, "periodic", 0.0, torch.float64)
    s = torch.sin(2 * math.pi * torch.arange(n, dtype=torch.float64, device=lat.device) / n)
    lat.init_equilibrium(ux=(SHEAR_AMP * s)[None, :])
    times, amps = [], []
    for c in range(SHEAR_STEPS // SHEAR_EVERY + 1):
        _, ux, _ = lat.moments()
        times.append(lat.time)
        amps.append(float((ux * s[None, :]).sum().item()) * 2.0 / (n * n))
        if c < SHEAR_STEPS // SHEAR_EVERY:
            lat.step(SHEAR_EVERY)
    nu = _fit_nu(times, amps, n)
    rel = nu / ((1 / omega - 0.5) / 3) - 1
    print("   FP64 nu relative error %.3e" % rel, flush=True)
    assert abs(rel) < 1e-6  # measured 9.6e-14


# (format, collide) -> bound on |nu / nu_theory - 1|; measured values in the comments
# (runs of stochastic rounding vary: bf16 with the FP64 collide has measured
# 6.4e-4 and 1.8e-3)
SHEAR_BOUNDS = {
    ("f32", "f64"): 1e-6,   # 6.9e-8
    ("bf16", "f64"): 5e-3,  # 6.4e-4
    ("f16", "f64"): 1e-3,   # 2.0e-5
    # the wave changes by a few FP32 ulps per step, so the FP32 collide must not
    # drift the momentum: its odd-part coefficients sum exactly (with fl(1/3)
    # and fl(1/12) rounded apart it biased nu by about 5 %)
    ("bf16", "f32"): 1e-2,  # 2.2e-3
    ("f16", "f32"): 3e-3,   # 3.1e-4
}


@pytest.mark.parametrize("fmt,compute", list(SHEAR_BOUNDS))
def test_shear_wave_4096_storage_viscosity(gpu, fmt, compute):
    from latticeboltzmann_amd.batch import BatchedLattices
    n = SHEAR_N
    print("shear wave %d^2 %s (%s collide), %d steps" % (n, fmt, compute, SHEAR_STEPS), flush=True)
    bat = BatchedLattices([dict(nx=n, ny=n, omega=1.0, amplitude=SHEAR_AMP, direction="x",
                                bc="periodic")], fmt=fmt, compute=compute)
    times, amps = bat.run(SHEAR_STEPS, sample_every=SHEAR_EVERY, graph=True)
    del bat
    torch.cuda.empty_cache()
    nu = _fit_nu(times, amps[:, 0], n)
    rel = nu / (1.0 / 6.0) - 1
    print("   %s %s nu relative error %.3e" % (fmt, compute, rel), flush=True)
    assert abs(rel) < SHEAR_BOUNDS[(fmt, compute)]


def test_batched_sweep_512_instances_seed0(gpu):
    from latticeboltzmann_amd.batch import BatchedLattices, shear_wave_sweep
    inst = shear_wave_sweep(512, 0)
    print("batched sweep: 512 instances, seed 0, 2000 steps, FP32 storage", flush=True)
    bat = BatchedLattices(inst, fmt="f32")
    assert bat.compute_name == "f64"  # FP32 storage: FP64 collide, unbiased rounding
    times, amp = bat.run(2000, sample_every=50, graph=True)
    assert times[-1] == 2000 and amp.shape == (len(times), 512)
    nu = bat.viscosities(times, amp)
    theory = bat.theory()
    assert np.allclose(theory, [(1 / p["omega"] - 0.5) / 3 for p in inst])
    rel = np.abs(nu / theory - 1)
    med, p90, worst = float(np.nanmedian(rel)), float(np.nanpercentile(rel, 90)), float(np.nanmax(rel))
    print("   nu relative error: median %.2e, p90 %.2e, max %.2e, fitted %d of 512"
          % (med, p90, worst, int(np.isfinite(rel).sum())), flush=True)
    assert np.isfinite(rel).sum() >= 500
    assert med < 5e-4     # measured 4.8e-5
    assert p90 < 5e-3     # measured 9.6e-4
    assert worst < 0.05   # measured 2.5e-2
    # the worst instances (wavelength 64, omega near 0.3, nu near 0.9) are the
    # BGK lattice's own departure from (1/omega - 1/2)/3 at that wavelength,
    # not storage error: the same instance in FP64, sampled and fitted the same
    # way, gives the same nu (measured: equal to 1e-8)
    from latticeboltzmann_amd.shear import ShearWave
    wl = bat.wavelengths()
    for i in np.argsort(-np.nan_to_num(rel))[:2]:
        p = inst[i]
        sw = ShearWave(p["nx"], p["ny"], p["omega"], amplitude=p["amplitude"],
                       direction=p["direction"])
        t64, a64 = sw.run(2000, 50)
        ok = (a64 > 1e-3 * a64[0]) & (t64 >= t64[1])
        nu64 = -np.polyfit(t64[ok].astype(np.float64), np.log(a64[ok]), 1)[0] / (2 * np.pi / wl[i]) ** 2
        print("   instance %d (%dx%d, omega %.3f): nu error %.3e batched FP32, %.3e FP64"
              % (i, p["nx"], p["ny"], p["omega"], nu[i] / theory[i] - 1, nu64 / theory[i] - 1), flush=True)
        assert abs(nu[i] / nu64 - 1) < 1e-6
        del sw
    # every instance stays finite and keeps its mass (periodic: no walls)
    for i in (0, 100, 511):
        f = bat.populations(i)
        assert np.isfinite(f).all()
        p = inst[i]
        assert abs(f.sum() / (p["nx"] * p["ny"]) - 1) < 1e-6


SMALL = [  # batched instance sides are multiples of 64
    dict(nx=64, ny=64, omega=1.3, amplitude=0.05, direction="x", bc="periodic"),
    dict(nx=64, ny=128, omega=0.7, amplitude=0.08, direction="y", bc="periodic"),
    dict(nx=128, ny=64, omega=1.6, amplitude=0.0, direction="x", bc="cavity", u0=0.1),
    dict(nx=64, ny=192, omega=1.9, amplitude=0.1, direction="y", bc="periodic"),
]


@pytest.mark.parametrize("compute", ["f32", "f64"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_batched_instances_against_fp64_oracle(gpu, compute, rounding):
    # one batched step of each instance against the FP64 oracle step of its
    # decoded stored state.  FP32 collide into FP32 storage: within one FP32
    # ulp of the value plus 4 ulps of the node's sum of magnitudes (the density
    # and momentum sums round at the scale of sum_i |x_i|, and every relaxed
    # population inherits omega w_i times that error); an FP64 collide rounded
    # once (compensated) within 4 (nearest) or 6 (stochastic) ulps of each
    # node's largest value
    from latticeboltzmann_amd.batch import BatchedLattices
    bat = BatchedLattices(SMALL, fmt="f32", compute=compute, rounding=rounding, variant=0)
    rng = np.random.default_rng(11)
    w = o.weights()[:, None, None]
    # perturb the stored state so the step is not a near-equilibrium no-op
    for i, p in enumerate(SMALL):
        off, n = int(bat.table["offset"][i]), 9 * p["nx"] * p["ny"]
        f = w * (1 + 0.05 * rng.standard_normal((9, p["nx"], p["ny"])))
        d = np.ascontiguousarray((f - w).transpose(1, 0, 2)).astype(np.float32).reshape(-1)
        bat.a[off:off + n].copy_(torch.from_numpy(d.view(np.float32)).to(bat.a.device)
                                 .view(bat.a.dtype))
    stored = [bat.a[int(bat.table["offset"][i]):int(bat.table["offset"][i]) + 9 * p["nx"] * p["ny"]]
              .cpu().numpy().view(np.float32).reshape(p["nx"], 9, p["ny"])
              for i, p in enumerate(SMALL)]
    bat.step(1)
    for i, p in enumerate(SMALL):
        # the instance table holds FP64 values; an FP32 collide rounds them to FP32
        rnd = (lambda x: float(np.float32(x))) if compute == "f32" else float
        om, u0 = rnd(p["omega"]), rnd(p.get("u0", 0.0))
        expect = lp.step_reference(stored[i], "f32", 1.0, om, p["bc"], u0, "rows")
        got = bat.populations(i)
        x_exp, x_got = expect - w, got - w
        if compute == "f32":
            tol = lp.ulp(x_exp, "f32") + 4 * lp.ulp(np.abs(x_exp).sum(0), "f32")[None]
        else:
            big = lp.ulp(np.abs(x_exp).max(0), "f32")
            tol = (4 if rounding == "nearest" else 6) * big[None] + 1e-12
        err = np.abs(x_got - x_exp)
        print("   instance %d %s: max err/tol %.3f" % (i, p["bc"], float((err / tol).max())), flush=True)
        assert (err <= tol).all(), (i, float((err / tol).max()))
