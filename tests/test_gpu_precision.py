"""Reduced-precision storage on the GPU against the oracle and the FP64 path.

Codec: bit-exact against numpy's round-to-nearest-even casts.  One step:
within one unit in the last place of the storage format (plus FP32 collide
rounding, 1e-8 absolute) of the FP64 oracle step of the decoded state.  Many
steps: velocity error against the FP64 GPU run, stated per format.
"""

import math

import numpy as np
import pytest
import torch

from oracle import d2q9_numpy as o
from oracle import lowprec as lp

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(4242)
SCALES = {"f32": 1.0, "bf16": 1.0, "f16": 1024.0}


def populations(nx, ny, spread=0.05):
    return o.weights(np.float64)[:, None, None] * (1 + spread * RNG.standard_normal((9, nx, ny)))


def lattice(nx, ny, fmt, bc="cavity", omega=1.3, layout="rows", variant=-1,
            rounding="stochastic", seed=0, compute="f32"):
    from latticeboltzmann_amd.precision import DeviationLattice
    # rounding every step: the per-step storage these tests bound
    return DeviationLattice(nx, ny, omega, bc, 0.1, fmt=fmt, scale=SCALES[fmt], layout=layout,
                            variant=variant, rounding=rounding, seed=seed, compute=compute,
                            store_every=1)


@pytest.mark.parametrize("fmt", lp.FORMATS)
@pytest.mark.parametrize("layout", ["rows", "soa"])
def test_encode_decode_bitexact(gpu, fmt, layout):
    f = populations(37, 300, spread=0.3)
    lat = lattice(37, 300, fmt, layout=layout)
    lat.load(f)
    raw = lat.d.cpu().numpy()
    expect = lp.encode(f, fmt, SCALES[fmt], layout)
    assert np.array_equal(raw.view(expect.dtype), expect)
    assert np.array_equal(lat.populations(), lp.decode(expect, fmt, SCALES[fmt], layout))


@pytest.mark.parametrize("fmt", lp.FORMATS)
@pytest.mark.parametrize("bc", ["cavity", "periodic"])
@pytest.mark.parametrize("shape", [(1, 1), (3, 1), (1, 5), (6, 7), (33, 300), (17, 1030)])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_one_step_against_fp64_oracle(gpu, fmt, bc, shape, rounding):
    # FP32 storage: every value within one FP32 ulp (+ collide rounding) of the
    # FP64 oracle step.  16-bit storage uses the conservation-compensated
    # rounding: every value within 4 ulps (6 with stochastic rounding, whose
    # single-value error reaches 1 ulp), and mass and momentum of each node
    # within one ulp of its largest value.
    from latticeboltzmann_amd import _native
    f = populations(*shape)
    s = SCALES[fmt]
    n_var = _native.load().lbm_dev_variant_count()
    w = o.weights()[:, None, None]
    for layout in ("rows", "soa"):
        for v in range(n_var):
            lat = lattice(*shape, fmt, bc=bc, layout=layout, variant=v, rounding=rounding)
            lat.load(f)
            stored = lat.d.cpu().numpy().view(np.float32 if fmt == "f32" else np.uint16)
            expect = lp.step_reference(stored, fmt, s, 1.3, bc, 0.1, layout)
            lat.step(1)
            got = lat.populations()
            x_exp, x_got = (expect - w) * s, (got - w) * s
            big = lp.ulp(np.abs(x_exp).max(0), fmt)
            slack = 1e-8 * s
            if fmt == "f32":
                tol = lp.ulp(x_exp, fmt) + slack
            else:
                tol = (4 if rounding == "nearest" else 6) * big[None] + slack
                # one rounding of the absorbing value: below one ulp of it with
                # nearest rounding, one full ulp (of a target that may lie in
                # the next binade) with stochastic rounding
                mom_tol = (1 if rounding == "nearest" else 2) * big + 3 * slack
                for a, b in zip(lp.conserved(x_got), lp.conserved(x_exp)):
                    assert (np.abs(a - b) <= mom_tol).all(), (layout, v, "moments")
            err = np.abs(x_got - x_exp)
            assert (err <= tol).all(), (layout, v, float((err / tol).max()))


@pytest.mark.parametrize("fmt", lp.FORMATS)
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_one_step_fp64_collide_against_oracle(gpu, fmt, rounding):
    # FP64 collide, one rounding to the storage format: within the
    # compensated-rounding bounds of the FP64 oracle step, in every launch shape
    # (omega and u0 are FP32-representable, so every compute type and the
    # oracle use the same values)
    from latticeboltzmann_amd import _native
    from latticeboltzmann_amd.precision import DeviationLattice
    w = o.weights()[:, None, None]
    s = SCALES[fmt]
    om, u0 = float(np.float32(1.3)), float(np.float32(0.1))
    for shape in ((6, 7), (33, 300)):
        f = populations(*shape)
        for v in range(_native.load().lbm_dev_variant_count()):
            lat = DeviationLattice(*shape, om, "cavity", u0, fmt=fmt, scale=s, variant=v,
                                   rounding=rounding, compute="f64", store_every=1)
            lat.load(f)
            stored = lat.d.cpu().numpy().view(np.float32 if fmt == "f32" else np.uint16)
            expect = lp.step_reference(stored, fmt, s, om, "cavity", u0, "rows")
            lat.step(1)
            x_exp, x_got = (expect - w) * s, (lat.populations() - w) * s
            big = lp.ulp(np.abs(x_exp).max(0), fmt)
            # FP32 storage is rounded with the compensation too under an FP64
            # collide, so every format gets the compensated-rounding bounds
            tol = (4 if rounding == "nearest" else 6) * big[None] + 1e-12
            err = np.abs(x_got - x_exp)
            assert (err <= tol).all(), (shape, v, float((err / tol).max()))


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("compute", ["f32", "f64"])
def test_stochastic_rounding_unbiased_and_deterministic(gpu, fmt, compute):
    # a uniform flow at equilibrium is a fixed point of the step; every node
    # stores the same exact value, so nearest rounding has one fixed error per
    # channel while stochastic rounding averages out over the nodes (with the
    # FP64 collide the values are rounded stochastically straight from FP64)
    n = 256
    rho, ux, uy = 1.0, 0.05, -0.02
    feq = o.equilibrium(np.full((n, n), rho), np.full((n, n), ux), np.full((n, n), uy))
    s = SCALES[fmt]
    outs = {}
    for rounding, seed in (("nearest", 0), ("stochastic", 1), ("stochastic", 1), ("stochastic", 2)):
        lat = lattice(n, n, fmt, bc="periodic", omega=1.0, rounding=rounding, seed=seed,
                      compute=compute)
        lat.load(feq)
        # the exact result of the step is the FP64 step of the (rounded) stored input
        stored = lat.d.cpu().numpy().view(np.uint16)[:2, :, :2].copy()
        ref = lp.step_reference(stored, fmt, s, 1.0, "periodic", layout="rows")
        x_exact = (ref[:, 0, 0] - o.weights()) * s
        big = lp.ulp(np.abs(x_exact).max(), fmt)
        lat.step(1)
        outs.setdefault((rounding, seed), []).append(lat.d.cpu().numpy().copy())
        x = (lat.populations() - o.weights()[:, None, None]) * s
        bias = np.abs(x.reshape(9, -1).mean(1) - x_exact)
        if rounding == "stochastic":
            assert (bias < 0.02 * big).all(), bias / big
            # the mass and momentum of the whole grid stay exact to far below one ulp
            for a, b in zip(lp.conserved(x.reshape(9, -1).mean(1)), lp.conserved(x_exact)):
                assert abs(a - b) < 0.02 * big
    a, b = outs[("stochastic", 1)]
    assert np.array_equal(a, b)                        # same seed, same bits
    assert not np.array_equal(a, outs[("stochastic", 2)][0])


def test_mass_and_residual_fp64(gpu):
    from latticeboltzmann_amd import _native, kernels
    for fmt in lp.FORMATS:
        f = populations(64, 80, spread=0.3)
        lat = lattice(64, 80, fmt)
        lat.load(f)
        dec = lat.populations()
        assert abs(lat.mass() - math.fsum(dec.ravel())) <= 1e-12 * dec.sum()
        # residual from zero equals max |u| of the decoded state (to FP64 rounding)
        ws = kernels._workspace(lat.device)
        uxp = torch.zeros((64, 80), dtype=torch.float64, device=lat.device)
        uyp = torch.zeros_like(uxp)
        _native.call("lbm_dev_velocity_residual", _native.ptr(lat.d), 64, 80, lat.layout, lat.fmt,
                     lat.scale, _native.ptr(uxp), _native.ptr(uyp), _native.ptr(ws.work),
                     _native.ptr(ws.out), _native.stream_ptr())
        _, ux, uy = o.moments(dec)
        assert abs(float(ws.out.item()) - np.sqrt(ux * ux + uy * uy).max()) < 1e-13
        assert np.allclose(uxp.cpu().numpy(), ux, rtol=0, atol=1e-14)


def test_cavity_velocity_error_against_fp64(gpu):
    # 256^2 cavity, Re = 100, U = 0.1, 3000 steps; velocity error of each
    # storage format against the FP64 run of the same problem
    from latticeboltzmann_amd.lattice import Lattice, omega_for_cavity
    from latticeboltzmann_amd.precision import velocity_errors
    n, steps = 256, 3000
    omega = omega_for_cavity(n, 100.0, 0.1)
    ref = Lattice(n, n, omega, "cavity", 0.1, torch.float64).step(steps)
    _, ux_r, uy_r = ref.moments()
    # measured on MI355X: f32 1.0e-6, f16 6.6e-5, bf16 5.1e-4 (stochastic rounding)
    bounds = {"f32": 1e-5, "f16": 2e-4, "bf16": 2e-3}
    for fmt, bound in bounds.items():
        lat = lattice(n, n, fmt, omega=omega).step(steps)
        _, ux, uy = lat.moments()
        err = velocity_errors(ux, uy, ux_r, uy_r, 0.1)
        print("cavity 256^2 Re=100 3000 steps", fmt, err)
        assert err["l2_over_U"] < bound, (fmt, err)
        assert abs(lat.mass() / ref.mass() - 1) < (1e-5 if fmt != "bf16" else 1e-3), fmt


def test_shear_wave_viscosity_fp16(gpu):
    n, omega, amp = 256, 1.0, 0.01
    from latticeboltzmann_amd.precision import DeviationLattice
    lat = DeviationLattice(n, n, omega, "periodic", fmt="f16", scale=1024.0, store_every=1)
    lat.load(o.shear_wave_init(n, n, amp, "x"))
    ts, amps = [], []
    for _ in range(11):
        ts.append(lat.time)
        amps.append(o.shear_wave_amplitude(lat.populations(), "x"))
        lat.step(200)
    nu = o.fit_viscosity(ts, amps, n)
    assert abs(nu / o.viscosity_theory(omega) - 1) < 1e-2


# storing every 6 or 8 steps runs on the register sweep, which takes the FP32
# collide only (checked below)
STORE_EVERY_CASES = [(c, k) for c in ("f32", "f64") for k in (2, 3, 4, 6, 8)
                     if not (c == "f64" and k > 4)]


@pytest.mark.parametrize("fmt", lp.FORMATS)
@pytest.mark.parametrize("compute,K", STORE_EVERY_CASES)
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_store_every_k_against_fp64_oracle(gpu, fmt, compute, rounding, K):
    # storing every K steps: K unrounded steps on chip, then one compensated
    # rounding, so the stored state lies within the one-step rounding bounds of
    # the FP64 oracle run for K steps (plus K steps of FP32 arithmetic drift)
    from latticeboltzmann_amd.precision import DeviationLattice
    w = o.weights()[:, None, None]
    s = SCALES[fmt]
    om, u0 = float(np.float32(1.3)), float(np.float32(0.1))
    shape = (24, 80)
    f = populations(*shape)
    for bc in ("cavity", "periodic"):
        print("store every", fmt, compute, rounding, K, bc, flush=True)
        lat = DeviationLattice(*shape, om, bc, u0, fmt=fmt, scale=s, rounding=rounding,
                               compute=compute, store_every=K)
        assert lat.multi_k == K and not lat.round_each
        lat.load(f)
        stored = lat.d.cpu().numpy().view(np.float32 if fmt == "f32" else np.uint16)
        expect = lp.decode(stored, fmt, s, "rows")
        for _ in range(K):
            (o.cavity_step if bc == "cavity" else o.periodic_step)(expect, *((om, u0) if bc == "cavity" else (om,)))
        lat.step(K)
        x_exp, x_got = (expect - w) * s, (lat.populations() - w) * s
        drift = (3 * K if compute == "f32" else 0) * lp.ulp(np.abs(x_exp).max(0), "f32")[None] + 1e-8 * s
        if fmt == "f32" and compute == "f32":
            tol = drift + lp.ulp(x_exp, "f32")
        else:
            big = lp.ulp(np.abs(x_exp).max(0), fmt)
            tol = (4 if rounding == "nearest" else 6) * big[None] + drift
        err = np.abs(x_got - x_exp)
        assert (err <= tol).all(), (bc, float((err / tol).max()))


@pytest.mark.parametrize("K", [6, 8])
def test_store_every_6_and_8_need_the_fp32_collide(gpu, K):
    # the register sweeps that store every 6 or 8 steps hold FP32 levels only:
    # asked for explicitly with the FP64 collide, the lattice refuses
    from latticeboltzmann_amd.precision import DeviationLattice
    with pytest.raises(ValueError):
        DeviationLattice(24, 80, 1.3, "cavity", 0.1, fmt="f16", compute="f64", store_every=K)
    lat = DeviationLattice(24, 80, 1.3, "cavity", 0.1, fmt="f16", compute="f32", store_every=K)
    assert lat.multi_k == K and not lat.round_each
