"""The npy writer: the oracle restatement of PyLB/IO.py pinned, and the
product writer (host arrays) against it, byte for byte."""

import ast

import numpy as np
import pytest

from oracle import npy_ref as r
from latticeboltzmann_amd.PyLB import IO
from latticeboltzmann_amd.dumps import dump_steps

CASES = [((3, 4), "<f8"), ((2048, 2048), "<f8"), ((1, 1), "<f4"), ((123, 45678), "<f4"),
         ((7, 5), "<i8"), ((10**6, 3), "<f2")]


def test_oracle_header_known_text():
    # written out by hand: dict text, 12 spaces of padding, newline -> 80 bytes
    text = "{'descr': '<f8', 'fortran_order': False, 'shape': (3, 4)}" + " " * 12 + "\n"
    assert r.header("<f8", (3, 4)) == b"\x93NUMPY\x01\x00" + bytes([len(text), 0]) + text.encode()
    assert len(r.header("<f8", (3, 4))) == 80


@pytest.mark.parametrize("shape,dt", CASES)
def test_oracle_header_closed_form(shape, dt):
    h = r.header(dt, shape)
    assert len(h) % 16 == 0 and h.endswith(b"\n")
    assert h[:8] == b"\x93NUMPY\x01\x00"
    assert int.from_bytes(h[8:10], "little") == len(h) - 10
    # no more than 15 spaces of padding before the newline
    body = h[10:-1].decode()
    assert len(body) - len(body.rstrip(" ")) < 16
    d = ast.literal_eval(body.rstrip(" "))  # the reference writes str(dict)
    assert d == {"descr": dt, "fortran_order": False, "shape": shape}


@pytest.mark.parametrize("shape,dt", [c for c in CASES if np.prod(c[0]) < 10**6])
def test_oracle_file_round_trip(tmp_path, shape, dt):
    a = (np.random.default_rng(1).standard_normal(shape) * 100).astype(dt)
    fn = tmp_path / "a.npy"
    r.save_single_rank(fn, a)
    b = np.load(fn)
    assert b.dtype == a.dtype and np.array_equal(a, b)
    raw = fn.read_bytes()
    assert raw[len(r.header(dt, shape)):] == a.tobytes()


@pytest.mark.parametrize("shape,dt", CASES)
def test_product_header_equals_oracle(shape, dt):
    assert IO.npy_header(dt, shape) == r.header(dt, shape)


def test_product_writer_equals_oracle(tmp_path):
    a = np.random.default_rng(2).standard_normal((33, 17))
    IO.save_mpiio(None, tmp_path / "p.npy", a)
    r.save_single_rank(tmp_path / "o.npy", a)
    assert (tmp_path / "p.npy").read_bytes() == (tmp_path / "o.npy").read_bytes()
    # a non-contiguous view is written in C order of its values
    IO.save_mpiio(None, tmp_path / "v.npy", a[::2, 1:])
    assert np.array_equal(np.load(tmp_path / "v.npy"), a[::2, 1:])


class _Comm:
    def __init__(self, n):
        self.n = n

    def Get_size(self):
        return self.n


def test_product_writer_rejects(tmp_path):
    IO.save_mpiio(_Comm(1), tmp_path / "ok.npy", np.zeros((2, 2)))
    with pytest.raises(NotImplementedError):
        IO.save_mpiio(_Comm(2), tmp_path / "x.npy", np.zeros((2, 2)))
    with pytest.raises(ValueError):
        IO.save_mpiio(None, tmp_path / "x.npy", np.zeros(4))


def test_dump_schedule():
    assert dump_steps(25, 10) == [0, 10, 20, 24]
    assert dump_steps(21, 10) == [0, 10, 20]
    assert dump_steps(25, 10, final=False) == [0, 10, 20]
    assert dump_steps(1, 10000) == [0]
    assert dump_steps(0, 5) == []
    with pytest.raises(ValueError):
        dump_steps(5, 0)


def test_oracle_cavity_dump_names():
    d = r.cavity_dumps(6, 5, 1.2, 7, 3)
    assert sorted(d) == sorted("u%s_%d.npy" % (c, i) for c in "xy" for i in (0, 3, 6))
    d = r.cavity_dumps(6, 5, 1.2, 8, 3)
    assert "ux_7.npy" in d and d["ux_0.npy"].shape == (6, 5)
