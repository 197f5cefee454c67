"""Host logic of the warm start: controller (with a mock stepper), features, layer table."""

import math

import numpy as np
import pytest

from latticeboltzmann_amd.warmstart.controller import Controller
from latticeboltzmann_amd.warmstart.features import features, GRID, IN_CHANNELS
from latticeboltzmann_amd.warmstart.network import unet_layers
from oracle import surrogate_ref as s


class MockStepper:
    """Residual histories scripted per start: ``warm(k)`` / ``cold(k)`` give the
    residual after the k-th block of steps."""

    def __init__(self, warm, cold):
        self.curves = {"warm": warm, "cold": cold}
        self.mode, self.k, self.calls, self.stepped = None, 0, [], 0

    def init_warm(self):
        self.mode, self.k = "warm", 0
        self.calls.append("warm")

    def init_cold(self):
        self.mode, self.k = "cold", 0
        self.calls.append("cold")

    def advance(self, n):
        self.k += 1
        self.stepped += n

    def residual(self):
        return self.curves[self.mode](self.k) if self.k else math.inf

    def sync_time(self, t0):
        return 1.0


def geometric(r0, rate):
    return lambda k: r0 * rate ** k


def test_warm_converges_without_fallback():
    st = MockStepper(geometric(1e-3, 0.8), geometric(1e-2, 0.97))
    r = Controller(tol=1e-6).run(st, warm=True)
    assert not r["fell_back"] and r["reason"] == "converged"
    assert st.calls == ["warm"]
    assert r["steps"] == r["warm_steps"] == 1000 * math.ceil(math.log(1e-3) / math.log(0.8))


def test_forced_stall_falls_back_to_cold():
    st = MockStepper(lambda k: 5e-5, geometric(1e-2, 0.9))
    ctl = Controller(tol=1e-6, patience=10)
    r = ctl.run(st, warm=True)
    assert r["fell_back"] and r["reason"] == "stalled"
    assert st.calls == ["warm", "cold"]
    assert r["warm_steps"] == 11 * 1000  # first check sets the best, then 10 without progress
    cold_blocks = math.ceil(math.log(1e-4) / math.log(0.9))
    assert r["cold_steps"] == 1000 * cold_blocks
    assert r["steps"] == r["warm_steps"] + r["cold_steps"]


def test_unavailable_warm_start_runs_cold():
    from latticeboltzmann_amd.warmstart.controller import WarmStartUnavailable

    class NoPrediction(MockStepper):
        def init_warm(self):
            self.calls.append("warm")
            raise WarmStartUnavailable("outside the trained range")

    st = NoPrediction(geometric(1e-3, 0.5), geometric(1e-2, 0.9))
    r = Controller(tol=1e-6).run(st, warm=True)
    assert r["fell_back"] and r["reason"] == "no warm start"
    assert st.calls == ["warm", "cold"] and r["warm_steps"] == 0
    assert r["steps"] == r["cold_steps"] == 1000 * math.ceil(math.log(1e-4) / math.log(0.9))


def test_predictor_refuses_inputs_outside_its_ranges():
    from latticeboltzmann_amd.warmstart.controller import WarmStartUnavailable
    from latticeboltzmann_amd.warmstart.solve import OutOfRange, Predictor

    class Net:  # no forward pass: the range check comes first
        in_ch = IN_CHANNELS

        def forward(self, x):
            raise AssertionError("forward must not run for out-of-range inputs")

    p = Predictor(Net(), re_range=(100.0, 1000.0), u_range=(0.05, 0.1))
    assert p.covers(100.0, 0.05) and p.covers(1000.0, 0.1) and p.covers(400.0, 0.07)
    for re, u in [(99.0, 0.07), (1001.0, 0.07), (400.0, 0.049), (400.0, 0.11), (5000.0, 0.1)]:
        assert not p.covers(re, u)
        with pytest.raises(OutOfRange):
            p.predict(re, u)
    assert issubclass(OutOfRange, WarmStartUnavailable)


def test_predictor_set_routes_by_ranges_and_lattice_size():
    from latticeboltzmann_amd.warmstart.solve import OutOfRange, Predictor, PredictorSet

    class Net:
        in_ch = IN_CHANNELS

        def __init__(self, tag):
            self.tag = tag

        def forward(self, x):
            return [self.tag]

    large = Predictor(Net("large"), re_range=(100.0, 1000.0), u_range=(0.05, 0.1), n_range=(384, 4096))
    low = Predictor(Net("low"), re_range=(100.0, 1000.0), u_range=(0.05, 0.1))
    high = Predictor(Net("high"), re_range=(1000.0, 5000.0), u_range=(0.05, 0.1))
    ps = PredictorSet([large, low, high])
    assert ps.member(500.0, 0.07, n=512) is large and ps.member(500.0, 0.07, n=256) is low
    assert ps.member(500.0, 0.07) is large  # no size given: the first member covering (Re, U)
    assert ps.member(3000.0, 0.07, n=1024) is high and ps.member(1000.0, 0.07, n=256) is low
    assert not large.covers(500.0, 0.07, n=8192) and not large.covers(500.0, 0.07, n=256)
    with pytest.raises(OutOfRange):
        large.predict(500.0, 0.07, n=256)
    with pytest.raises(OutOfRange):
        ps.predict(6000.0, 0.07, n=512)
    assert ps.re_range == (100.0, 5000.0)


def test_stall_window_grows_with_the_run():
    # fast decay for 400 checks, then 20 % drops every 80 checks: past check
    # 400 half of the run (200+ checks) is longer than the 50-check patience,
    # so this slow tail is not a stall; with the window fixed it is
    def warm(k):
        return 1e-2 * 0.99 ** min(k, 400) * 0.8 ** (max(0, k - 400) // 80)

    st = MockStepper(warm, geometric(1e-2, 0.9))
    r = Controller(tol=1e-6, patience=50).run(st, warm=True)
    assert not r["fell_back"] and r["reason"] == "converged"
    st = MockStepper(warm, geometric(1e-2, 0.9))
    r = Controller(tol=1e-6, patience=50, patience_frac=0.0).run(st, warm=True)
    assert r["fell_back"] and r["reason"] == "stalled"


def test_predictor_set_mean_averages_the_covering_members():
    import torch
    from latticeboltzmann_amd.warmstart.solve import Predictor, PredictorSet

    class Net:
        in_ch = IN_CHANNELS
        device = torch.device("cpu")

        def __init__(self, value):
            self.value = value

        def forward(self, x):
            return torch.full((x.shape[0], 2, 2, 3), self.value)

    a = Predictor(Net(1.0), re_range=(100.0, 1000.0), u_range=(0.05, 0.1), n_range=(384, 4096))
    b = Predictor(Net(3.0), re_range=(100.0, 1000.0), u_range=(0.05, 0.1))
    c = Predictor(Net(10.0), re_range=(1000.0, 5000.0), u_range=(0.05, 0.1))
    ps = PredictorSet([a, b, c], combine="mean")
    assert float(ps.predict(500.0, 0.07, n=512)[0, 0, 0]) == 2.0   # a and b
    assert float(ps.predict(500.0, 0.07, n=256)[0, 0, 0]) == 3.0   # b alone
    assert float(ps.predict(3000.0, 0.07, n=512)[0, 0, 0]) == 10.0  # c alone
    assert float(PredictorSet([a, b, c]).predict(500.0, 0.07, n=512)[0, 0, 0]) == 1.0  # first
    with pytest.raises(ValueError):
        PredictorSet([a], combine="median")


def test_nan_falls_back_immediately():
    st = MockStepper(lambda k: float("nan"), geometric(1e-3, 0.5))
    r = Controller(tol=1e-6).run(st, warm=True)
    assert r["fell_back"] and r["reason"] == "diverged" and r["warm_steps"] == 1000


def test_slow_but_steady_progress_is_not_a_stall():
    # halves every 30 checks: beats min_drop = 0.9 within patience = 50
    st = MockStepper(geometric(1e-4, 0.5 ** (1 / 30)), geometric(1e-2, 0.9))
    r = Controller(tol=1e-6, patience=50).run(st, warm=True)
    assert not r["fell_back"]


def test_oscillating_residual_stalls():
    st = MockStepper(lambda k: 1e-4 * (1.5 if k % 2 else 1.0), geometric(1e-2, 0.9))
    r = Controller(tol=1e-6, patience=20).run(st, warm=True)
    assert r["fell_back"] and r["reason"] == "stalled"


def test_budget_limits_warm_steps():
    st = MockStepper(geometric(1e-2, 0.99), geometric(1e-2, 0.9))
    r = Controller(tol=1e-6, budget=0.5).run(st, warm=True, cold_estimate=40_000)
    assert r["fell_back"] and r["reason"] == "over budget" and r["warm_steps"] == 20_000


def test_cold_run_reports_max_steps():
    st = MockStepper(None, lambda k: 1e-3)
    r = Controller(tol=1e-6, max_steps=5000).run(st, warm=False)
    assert r["reason"] == "max steps" and r["steps"] == 5000 and st.calls == ["cold"]


def test_controller_rejects_bad_settings():
    with pytest.raises(ValueError):
        Controller(min_drop=1.5)
    with pytest.raises(ValueError):
        Controller(patience=0)


def test_features_layout():
    x = features(100.0, 0.05, m=8)
    assert x.shape == (8, 8, IN_CHANNELS) and x.dtype == np.float32
    assert np.all(x[..., 0] == -1.0) and np.all(x[..., 1] == -1.0)
    assert np.all(features(1000.0, 0.1, m=8)[..., :2] == 1.0)
    # walls: left (x = 0), right (x = m-1), bottom (y = 0); lid: top row y = m-1
    assert x[0, 3, 2] == 1 and x[7, 3, 2] == 1 and x[3, 0, 2] == 1 and x[3, 3, 2] == 0
    assert x[3, 7, 3] == 1 and x[3, 6, 3] == 0
    assert x[0, 0, 4] == x[0, 5, 4] < 0 < x[7, 0, 4] and x[0, 0, 5] < 0 < x[0, 7, 5]
    assert GRID == 128


def test_feature_octaves():
    from latticeboltzmann_amd.warmstart.features import in_channels
    x = features(300.0, 0.07, m=4, octaves=2)
    assert x.shape == (4, 4, in_channels(2)) == (4, 4, 14)
    a, b = float(x[0, 0, 0]), float(x[0, 0, 1])
    for j in range(2):
        w = np.pi * 2 ** j
        np.testing.assert_allclose(x[1, 2, 6 + 4 * j:10 + 4 * j],
                                   [np.sin(w * a), np.cos(w * a), np.sin(w * b), np.cos(w * b)],
                                   rtol=1e-6, atol=1e-6)
    assert np.array_equal(x[..., :6], features(300.0, 0.07, m=4))


def test_dataset_density_is_mean_free(tmp_path):
    from latticeboltzmann_amd.warmstart.train import load_dataset
    rng = np.random.default_rng(0)
    t = rng.normal(size=(3, 4, 4, 3)).astype(np.float16)
    p = tmp_path / "d.npz"
    np.savez(p, re=np.array([100.0, 200.0, 300.0]), u=np.array([0.05, 0.06, 0.07]), targets=t)
    params, y = load_dataset([str(p)])
    assert params[1] == (200.0, 0.06) and y.dtype == np.float32
    assert np.abs(y[..., 0].mean(axis=(1, 2))).max() < 1e-6
    np.testing.assert_array_equal(y[..., 1:], t[..., 1:].astype(np.float32))
    _, raw = load_dataset([str(p)], mean_free_density=False)
    np.testing.assert_array_equal(raw, t.astype(np.float32))


def test_layer_table_matches_oracle():
    assert unet_layers(6, (32, 64, 128), 128, 3) == s.unet_layers(6, (32, 64, 128), 128, 3)


def test_evaluation_summary():
    from latticeboltzmann_amd.warmstart.evaluate import summarize
    recs = []
    for i, (cs, ws, cst, wst, d, fb) in enumerate([(100, 40, 1.0, 0.40, 2e-5, False),
                                                     (200, 50, 2.0, 0.55, 5e-5, False),
                                                     (300, 150, 3.0, 1.60, 2e-4, True)]):
        recs.append({"n": 256 * (1 + i % 2), "re": 100.0, "u": 0.1,
                     "cold": {"steps": cs, "seconds": cst, "reason": "converged"},
                     "warm": {"steps": ws, "seconds": wst, "fell_back": fb, "diff_max": d,
                              "step_speedup": cs / ws, "time_speedup": cst / wst}})
    # a cavity whose cold start never converged has no yardstick: listed, not summarized
    recs.append({"n": 1024, "re": 4900.0, "u": 0.05,
                 "cold": {"steps": 4000000, "seconds": 40.0, "reason": "diverged"}})
    out = summarize(recs)
    w = out["warm"]
    assert w["median_step_speedup"] == 2.5 and w["fallbacks"] == 1
    assert np.isclose(w["total_time_speedup"], 6.0 / 2.55)
    assert np.isclose(w["min_time_speedup"], 3.0 / 1.6)
    assert w["diff_max_worst"] == 2e-4 and np.isclose(w["frac_diff_below_1e-4"], 2 / 3)
    assert [c["n"] for c in out["cold_not_converged"]] == [1024]
    assert out["by_n"]["warm_256"]["cavities"] == 2 and out["by_n"]["warm_512"]["cavities"] == 1
    assert out["by_n"]["warm_512"]["median_time_speedup"] == 2.0 / 0.55


def test_evaluation_summary_against_tightened_steady_state():
    from latticeboltzmann_amd.warmstart.evaluate import summarize
    recs = []
    # (cold error, warm error) against the run continued to a smaller residual
    for i, (ce, we) in enumerate([(3e-5, 2e-5), (1.5e-4, 0.5e-4), (2e-4, 2.5e-4)]):
        recs.append({"n": 512, "re": 1000.0, "u": 0.1,
                     "cold": {"steps": 100, "seconds": 1.0, "reason": "converged"},
                     "tight": {"tol": 1e-8, "cold_err_max": ce, "reason": "converged"},
                     "warm": {"steps": 30, "seconds": 0.3, "fell_back": False, "diff_max": ce + we,
                              "step_speedup": 100 / 30, "time_speedup": 1 / 0.3,
                              "err_max_vs_tight": we}})
    # a cavity whose continued run never settles has no steady state to compare with
    drift = dict(recs[0], tight={"tol": 1e-8, "cold_err_max": 9.0, "reason": "max steps"})
    drift["warm"] = dict(recs[0]["warm"], err_max_vs_tight=9.0)
    t = summarize(recs + [drift])["warm"]["vs_tight"]
    assert t["cavities"] == 3 and t["no_steady_state"] == 1
    assert t["cold_err_max_median"] == 1.5e-4 and t["cold_err_max_worst"] == 2e-4
    assert t["err_max_median"] == 0.5e-4 and t["err_max_worst"] == 2.5e-4
    assert np.isclose(t["frac_err_below_1e-4"], 2 / 3)
    assert np.isclose(t["frac_cold_err_below_1e-4"], 1 / 3)
    assert np.isclose(t["frac_not_farther_than_cold"], 2 / 3)
    # without the tightened run there is no such entry
    for r in recs:
        del r["warm"]["err_max_vs_tight"]
    assert "vs_tight" not in summarize(recs)["warm"]


def test_settle_runs_on_without_a_stall_test():
    # a residual that stops improving for a long time still settles (no stall
    # test), and the step cap ends a run that never gets there
    st = MockStepper(None, lambda k: 1e-3 if k < 200 else 1e-3 * 0.5 ** (k - 200))
    st.init_cold()
    status, steps, res = Controller(tol=1e-8, patience=10).settle(st)
    assert status == "converged" and res < 1e-8 and steps == 1000 * 217
    st.init_cold()
    status, steps, _ = Controller(tol=1e-8).settle(st, max_steps=50_000)
    assert status == "max steps" and steps == 50_000
