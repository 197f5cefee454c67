"""Host-side logic that needs no GPU: sweep draws, tile tables, table layouts,
checkpoint files."""

import ctypes

import numpy as np
import pytest
import torch


def test_sweep_draws_are_reproducible():
    from latticeboltzmann_amd.batch import SIZES, shear_wave_sweep
    a, b = shear_wave_sweep(512, 0), shear_wave_sweep(512, 0)
    assert a == b and len(a) == 512
    assert {p["nx"] for p in a} <= set(SIZES) and {p["ny"] for p in a} <= set(SIZES)
    assert all(0.3 <= p["omega"] <= 1.9 and 0.01 <= p["amplitude"] <= 0.1 for p in a)
    assert {p["direction"] for p in a} == {"x", "y"}
    # the draws of numpy.random.default_rng(0), in the documented order
    rng = np.random.default_rng(0)
    nx = rng.choice(SIZES, 512)
    assert [p["nx"] for p in a] == nx.tolist()
    assert shear_wave_sweep(512, 1) != a


def test_instance_and_tile_records_match_the_c_structs():
    from latticeboltzmann_amd.batch import INST_DTYPE, TILE_DTYPE

    class Inst(ctypes.Structure):
        _fields_ = [("offset", ctypes.c_int64), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                    ("omega", ctypes.c_double), ("u0", ctypes.c_double), ("bc", ctypes.c_int32),
                    ("dir", ctypes.c_int32)]

    class Tile(ctypes.Structure):
        _fields_ = [("inst", ctypes.c_int32), ("first", ctypes.c_int32)]

    assert INST_DTYPE.itemsize == ctypes.sizeof(Inst) == 40
    assert TILE_DTYPE.itemsize == ctypes.sizeof(Tile) == 8
    for name, _ in Inst._fields_:
        assert INST_DTYPE.fields[name][1] == getattr(Inst, name).offset


def test_tiles_cover_every_node_once():
    from latticeboltzmann_amd.batch import _tiles
    sizes = [64 * 64, 64 * 1024, 128 * 192]
    for tile in (64, 256, 1024):
        t = _tiles(range(3), sizes, tile)
        for i, n in enumerate(sizes):
            firsts = np.sort(t["first"][t["inst"] == i])
            assert firsts[0] == 0 and np.all(np.diff(firsts) == tile)
            assert firsts[-1] < n <= firsts[-1] + tile


def test_size_class_rounds_up_to_powers_of_two():
    from latticeboltzmann_amd.autotune import size_class
    assert size_class(1000, 1000) == "1024x1024"
    assert size_class(64, 65) == "64x128"
    assert size_class(1, 1) == "1x1"


def _bare(cls, **attrs):
    # the host arithmetic of a lattice without touching a device
    obj = cls.__new__(cls)
    obj.__dict__.update(attrs)
    return obj


def test_lattice_launch_count():
    from latticeboltzmann_amd.lattice import Lattice, StepGraph
    lat = _bare(Lattice, multi_variant=10, multi_k=3)
    assert [lat.launches(n) for n in (0, 1, 2, 3, 4, 1000, 1001)] == [0, 1, 2, 1, 2, 334, 335]
    single = _bare(Lattice, multi_variant=None, multi_k=1)
    assert single.launches(1000) == 1000
    # a graph is captured only when the state ends in the buffer it started in
    assert StepGraph(lat, 1000).enabled and not StepGraph(lat, 999).enabled
    assert not StepGraph(_bare(Lattice, multi_variant=12, multi_k=4), 1001).enabled
    assert StepGraph(single, 2).enabled and not StepGraph(single, 3).enabled


def test_deviation_lattice_launch_count():
    from latticeboltzmann_amd.precision import DeviationLattice
    lat = _bare(DeviationLattice, multi_variant=2, multi_k=3)
    assert [lat.launches(n) for n in (1, 3, 7, 9)] == [1, 1, 3, 3]
    assert _bare(DeviationLattice, multi_variant=None, multi_k=1).launches(7) == 7


class _FakeLattice:
    """Host stand-in with a Lattice's checkpointed attributes (CPU tensors)."""

    def __init__(self, nx=4, ny=6, omega=1.2, dtype=torch.float64):
        self.nx, self.ny, self.omega, self.bc, self.u0, self.layout = nx, ny, omega, "cavity", 0.1, 1
        self.f = torch.zeros((nx, 9, ny), dtype=dtype)
        self.time = 0


class _FakeDeviationLattice(_FakeLattice):
    def __init__(self, fmt=2, **kw):
        super().__init__(**kw)
        del self.f
        self.fmt, self.scale, self.rounding, self.seed, self.compute, self.store_every = \
            fmt, 1024.0, 1, 7, 0, 4
        raw = torch.float32 if fmt == 0 else torch.int16
        self.d = torch.zeros((self.nx, 9, self.ny), dtype=raw)


def test_checkpoint_round_trip_and_parameter_checks(tmp_path):
    from latticeboltzmann_amd import checkpoint
    rng = np.random.default_rng(5)
    a = _FakeLattice()
    a.f.copy_(torch.from_numpy(rng.standard_normal(a.f.shape)))
    a.f[0, 0, 0] = float("nan")  # every bit, NaN payloads included
    a.time = 1234
    p = str(tmp_path / "lat.npz")
    checkpoint.save(a, p)
    b = _FakeLattice()
    checkpoint.restore(b, p)
    assert b.time == 1234
    assert np.array_equal(b.f.numpy().view(np.uint64), a.f.numpy().view(np.uint64))
    for attr, val in (("omega", 1.3), ("bc", "periodic"), ("layout", 0)):
        c = _FakeLattice()
        setattr(c, attr, val)
        with pytest.raises(ValueError):
            checkpoint.restore(c, p)
    with pytest.raises(ValueError):  # grid shape
        checkpoint.restore(_FakeLattice(nx=5), p)
    with pytest.raises(ValueError):  # storage type
        checkpoint.restore(_FakeLattice(dtype=torch.float32), p)
    with pytest.raises(ValueError):  # kind
        checkpoint.restore(_FakeDeviationLattice(), p)


def test_checkpoint_deviation_words(tmp_path):
    from latticeboltzmann_amd import checkpoint
    a = _FakeDeviationLattice()
    a.d.copy_(torch.arange(a.d.numel(), dtype=torch.int16).view(a.d.shape))
    a.time = 48
    p = str(tmp_path / "dev.npz")
    checkpoint.save(a, p)
    b = _FakeDeviationLattice()
    checkpoint.restore(b, p)
    assert b.time == 48 and torch.equal(a.d, b.d)
    c = _FakeDeviationLattice()
    c.seed = 8  # another dither stream: not the same run
    with pytest.raises(ValueError):
        checkpoint.restore(c, p)
    d = _FakeDeviationLattice()
    d.store_every = 1
    with pytest.raises(ValueError):
        checkpoint.restore(d, p)
