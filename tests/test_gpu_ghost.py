"""Ghost-layer flows on the GPU against the oracle's restatement of the
reference's scripts, bit for bit, and against the analytic profiles."""

import numpy as np
import pytest
import torch

from oracle import ghost_ref as g

pytestmark = pytest.mark.gpu

DT = [(torch.float64, np.float64), (torch.float32, np.float32)]


def same(lat, mom, G, omom):
    assert np.array_equal(lat.populations(), G)
    for a, b in zip(mom, omom):
        assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("tdt,ndt", DT)
@pytest.mark.parametrize("steps", [1, 2, 37])
@pytest.mark.parametrize("layout", ["rows", "soa"])
def test_couette_matches_oracle(gpu, tdt, ndt, steps, layout):
    from latticeboltzmann_amd import ghost
    print("couette", tdt, steps, layout, flush=True)
    lat, mom = ghost.couette(20, 12, 0.1, 1.3, steps, dtype=tdt, layout=layout)
    G, omom = g.couette(20, 12, 0.1, 1.3, steps, dtype=ndt)
    same(lat, mom, G, omom)


@pytest.mark.parametrize("tdt,ndt", DT)
@pytest.mark.parametrize("steps", [1, 37])
def test_poiseuille_matches_oracle(gpu, tdt, ndt, steps):
    from latticeboltzmann_amd import ghost
    print("poiseuille", tdt, steps, flush=True)
    lat, mom = ghost.poiseuille(22, 12, 1.001, 0.999, 0.7, steps, uw=0.02, dtype=tdt)
    G, omom = g.poiseuille(22, 12, 1.001, 0.999, 0.7, steps, uw=0.02, dtype=ndt)
    same(lat, mom, G, omom)


@pytest.mark.parametrize("tdt,ndt", DT)
@pytest.mark.parametrize("layout", ["rows", "soa"])
def test_sliding_lid_matches_oracle(gpu, tdt, ndt, layout):
    from latticeboltzmann_amd import ghost
    print("sliding lid", tdt, layout, flush=True)
    lat, mom = ghost.sliding_lid(16, steps=37, uw=0.1, omega=1.7, dtype=tdt, layout=layout)
    G, omom = g.sliding_lid(18, 18, 0.1, 1.7, 37, dtype=ndt)
    same(lat, mom, G, omom)


@pytest.mark.parametrize("tdt,ndt", DT)
@pytest.mark.parametrize("layout", ["rows", "soa"])
def test_obstacle_canal_matches_oracle(gpu, tdt, ndt, layout):
    from latticeboltzmann_amd import ghost
    print("obstacle canal", tdt, layout, flush=True)
    ob = (12, 8, 15, 11)
    lat, mom = ghost.obstacle_canal(30, 20, ob, rho_diff=0.01, omega=0.8, steps=37, dtype=tdt,
                                    layout=layout)
    G, omom = g.obstacle_canal(32, 22, 1 + 0.01, 1 - 0.01, 0.8, 37, ob, dtype=ndt)
    same(lat, mom, G, omom)


@pytest.mark.parametrize("vertical", [True, False])
def test_pulse_kats(gpu, vertical):
    from latticeboltzmann_amd import ghost
    f = np.zeros((9, 9, 9))
    if vertical:
        f[1, 1, 1:8] = 1
    else:
        f[2, 1:8, 1] = 1
    lat = ghost.GhostLattice(9, 9, 1.0, ghost.pulse_walls(9, 9, vertical)).load(f)
    for _ in range(14):
        lat.stream_and_walls()
    got = lat.populations()
    assert np.array_equal(got, g.pulse_walls(f.copy(), 14, vertical))
    assert np.all((got[1, 1, 1:7] if vertical else got[2, 1:7, 1]) == 1)


def test_obstacle_pulses(gpu):
    from latticeboltzmann_amd import ghost
    rng = np.random.default_rng(3)
    f = rng.random((9, 16, 16))
    lat = ghost.GhostLattice(16, 16, 1.0, ghost.obstacle_faces(6, 6, 9, 9)).load(f)
    h = f.copy()
    for _ in range(9):
        lat.stream_and_walls()
        g.stream(h)
        g.apply_obstacle(h, 6, 6, 9, 9)
    assert np.array_equal(lat.populations(), h)


def test_collide_only_and_equilibrium(gpu):
    from latticeboltzmann_amd import ghost
    rng = np.random.default_rng(4)
    rho = 1 + 0.1 * rng.standard_normal((7, 5))
    ux, uy = 0.1 * rng.standard_normal((7, 5)), 0.1 * rng.standard_normal((7, 5))
    lat = ghost.GhostLattice(7, 5, 0.9).init_equilibrium(rho, ux, uy)
    assert np.array_equal(lat.populations(), g.equilibrium(rho, ux, uy))
    f = g.equilibrium(rho, ux, uy) * (1 + 0.05 * rng.standard_normal((9, 7, 5)))
    lat.load(f)
    mom = lat.moment_buffers()
    lat.collide_only(mom)
    h = f.copy()
    om = g.collide(h, 0.9)
    same(lat, mom, h, om)


def test_rule_table_rejections(gpu):
    from latticeboltzmann_amd import ghost, _native as nat
    R = ghost.Rule
    bad = [
        [R(1, 5, 1, 2, 2, 4, 0, -1), R(1, 5, 2, 3, 2, 2, 0, -1)],  # reads what rule 0 wrote
        [R(0, 5, 0, 1, 2, 4, 0, -1)],                              # source outside the grid
        [R(1, 9, 1, 2, 2, 4, 0, -1)],                              # target outside the grid
        [R(1, 5, 1, 2, 9, 4, 0, -1)],                              # channel 9
        [R(1, 5, 8, 9, 7, 5, 0, -1, nat.RULE_LID_RHO, 0.1), R(1, 5, 7, 8, 0, 0, 0, 0)][::-1],
    ]
    for rules in bad:
        lat = ghost.GhostLattice(8, 10, 1.0, rules)
        with pytest.raises(ValueError):
            lat.step()
    with pytest.raises(ValueError):
        ghost.GhostLattice(8, 10, 1.0, [R(1, 2, 1, 2, 0)] * 33)


def test_couette_analytic(gpu):
    from latticeboltzmann_amd import ghost
    lat, (rho, ux, uy) = ghost.couette(100, 52, 0.1, 0.5, 20000)
    prof = ux.cpu().numpy()[50, 1:-1]
    assert np.abs(prof - ghost.couette_profile(50, 0.1)).max() < 3e-4


def test_poiseuille_analytic(gpu):
    from latticeboltzmann_amd import ghost
    lat, (rho, ux, uy) = ghost.poiseuille(102, 52, 1.001, 0.999, 0.5, 15000)
    ana = ghost.poiseuille_profile(100, 50, 0.001, 0.5)
    prof = ux.cpu().numpy()[51, 1:-1]
    assert np.abs(prof - ana).max() / ana.max() < 0.01


def test_sliding_lid_reference_config(gpu):
    # slidingLid.py:20-28: 300 + 2 nodes, Re 1000, lid 0.1, 10000 steps
    from latticeboltzmann_amd import ghost
    lat, (rho, ux, uy) = ghost.sliding_lid()
    u = ux.cpu().numpy()
    assert np.isfinite(lat.populations()).all()
    assert u[151, -2] > 0.05 and u[151, 20:150].min() < 0  # lid-driven primary vortex
