"""The C-ABI library builds, loads and exports every symbol include/lbm_d2q9.h
declares; host-side argument validation of the drop-in raises like the
reference's bindings.  No GPU compute is called here."""

import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_d2q9.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from latticeboltzmann_amd import _native
    lib = _native.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_ctypes_signatures_cover_header():
    from latticeboltzmann_amd import _native
    assert set(_native.SIGNATURES) <= set(declared_functions())


def test_version_and_variants_callable_without_gpu():
    from latticeboltzmann_amd import _native
    lib = _native.load()
    assert lib.lbm_version().startswith(b"latticeboltzmann_amd")
    assert lib.lbm_step_variant_count() >= 1
    assert lib.lbm_reduce_workspace_bytes() > 0


def test_null_and_size_errors_are_reported_before_any_launch():
    from latticeboltzmann_amd import _native
    lib = _native.load()
    assert lib.lbm_step_f64(None, None, 4, 4, 1.0, 0, 0.0, None, None, None, -1, None) == -1
    # nonzero fake pointers never reach a launch when a size is invalid
    assert lib.lbm_step_f64(16, 4096, 0, 4, 1.0, 0, 0.0, None, None, None, -1, None) == -2
    assert lib.lbm_step_f64(16, 4096, 4, 4, 1.0, 7, 0.0, None, None, None, -1, None) == -3
    assert lib.lbm_step_f64(16, 32, 4, 4, 1.0, 0, 0.0, None, None, None, -1, None) == -4
    assert lib.lbm_stream_periodic(16, 1 << 20, 4, 4, 3, None) == -3
    with pytest.raises(ValueError):
        _native.check(-2, "x")
    with pytest.raises(RuntimeError):
        _native.check(700, "x")


def test_dropin_rejects_like_the_reference_bindings():
    from latticeboltzmann_amd import PyLB
    f = np.zeros((9, 4))
    with pytest.raises(TypeError):
        PyLB.collide(f.astype(np.int64), 1.0)            # non-float dtype
    with pytest.raises(TypeError):
        PyLB.collide(np.zeros((9, 2, 2)), 1.0)           # rank 3 (Eigen::Ref is 2-D)
    with pytest.raises(TypeError):
        PyLB.collide(np.zeros((8, 4)), 1.0)              # not 9 rows
    with pytest.raises(TypeError):
        PyLB.collide(np.zeros((4, 9)).T, 1.0)            # not C-contiguous
    ro = np.zeros((9, 4))
    ro.setflags(write=False)
    with pytest.raises(TypeError):
        PyLB.collide(ro, 1.0)                            # read-only
    with pytest.raises(TypeError):
        PyLB.collide(f, "fast")                          # omega not a number
    r = np.ones(4)
    with pytest.raises(TypeError):
        PyLB.equilibrium(r, r, r.astype(np.float32), f)  # mixed dtypes
    with pytest.raises(ValueError):
        PyLB.equilibrium(r, r, np.ones(3), f)            # length mismatch (stricter)
    with pytest.raises(TypeError):
        PyLB.equilibrium(r, r)                           # no such overload
    with pytest.raises(TypeError):
        PyLB.equilibrium("a", 0.0, 0.0)
    with pytest.raises(ValueError):
        PyLB.stream(np.zeros((8, 3, 3)))
    with pytest.raises(TypeError):
        PyLB.stream_and_bounce_back(np.zeros((9, 3, 3), dtype=np.int32))


def test_dropin_imports_as_top_level_pylb():
    """The reference's env.sh:17 puts its package directory on PYTHONPATH and
    the drivers say `import PyLB` (cavity_opt2.py:43, tests/02-CollideTest.py:27).
    The same import, with this package's directory on the path, gives the
    GPU functions and the IO module."""
    import subprocess
    import sys
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join(
        [os.path.join(ROOT, "latticeboltzmann_amd"), ROOT, env.get("PYTHONPATH", "")])
    code = ("import PyLB, PyLB.IO as IO\n"
            "from latticeboltzmann_amd import kernels\n"
            "assert PyLB.__file__.endswith('latticeboltzmann_amd/PyLB/__init__.py'), PyLB.__file__\n"
            "names = ['equilibrium', 'collide', 'stream', 'stream_and_bounce_back']\n"
            "assert all(callable(getattr(PyLB, n)) for n in names)\n"
            "assert callable(IO.save_mpiio)\n"
            "print('ok')\n")
    out = subprocess.run([sys.executable, "-c", code], env=env, cwd="/tmp",
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


def test_multi_step_shape_tables_without_gpu():
    # every multi-step shape reports 2..8 steps and a smallest grid of at
    # least its tile plus the halo (the row sweeps: 2K rows, one wave of 64
    # columns); the defaults are shapes of the tables, and each store-every-K
    # shape of DeviationLattice steps K times per launch
    from latticeboltzmann_amd import kernels
    from latticeboltzmann_amd.precision import DeviationLattice, multi_info
    from latticeboltzmann_amd import _native
    lib = _native.load()
    for count, info, default in (
            (lib.lbm_step_multi_variant_count(), kernels.step_multi_info,
             lib.lbm_step_multi_default_variant()),
            (lib.lbm_dev_multi_variant_count(), multi_info, lib.lbm_dev_multi_default_variant())):
        assert 0 <= default < count
        assert info(-1) == info(default)
    # the size-dependent default: the tile below 4096^2 nodes, a sweep from there
    small, large = lib.lbm_step_multi_default_for(2048, 2048), lib.lbm_step_multi_default_for(8192, 8192)
    assert small == lib.lbm_step_multi_default_variant()
    assert 0 <= large < lib.lbm_step_multi_variant_count() and large != small
    assert kernels.step_multi_info(large)[2] == 64  # register sweeps need 64 columns
    assert lib.lbm_step_multi_default_for(0, 5) == -2
    # FP32 lattices: their own tile default below 4096^2, the same sweep from there
    small32 = lib.lbm_step_multi_default_for_f32(2048, 2048)
    assert 0 <= small32 < lib.lbm_step_multi_variant_count() and small32 != large
    assert kernels.step_multi_info(small32)[0] == 4
    assert lib.lbm_step_multi_default_for_f32(8192, 8192) == large
    assert lib.lbm_step_multi_default_for_f32(-1, 5) == -2
    for count, info, default in (
            (lib.lbm_step_multi_variant_count(), kernels.step_multi_info,
             lib.lbm_step_multi_default_variant()),
            (lib.lbm_dev_multi_variant_count(), multi_info, lib.lbm_dev_multi_default_variant())):
        for v in range(count):
            K, mnx, mny = info(v)
            assert 2 <= K <= 8 and mnx >= 2 * K and mny >= 2 * K + 1, (v, K, mnx, mny)
    for K, shape in DeviationLattice._STORE_EVERY_SHAPES.items():
        assert multi_info(shape)[0] == K, (K, shape)
