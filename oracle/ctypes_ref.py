"""ctypes loader for the C restatement ``oracle/d2q9_ref.c``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_d2q9.so")
_lib = None


def build():
    """Compile ``liboracle_d2q9.so`` with the recipe in ``oracle/Makefile``."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle_d2q9.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        i64, i32, vp = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
        for suf, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(_lib, "oracle_step_" + suf).argtypes = [vp, vp, i64, i64, ct, i32, ct, vp, vp, vp]
            getattr(_lib, "oracle_run_" + suf).argtypes = [vp, vp, i64, i64, ct, i32, ct, i64]
            getattr(_lib, "oracle_collide_" + suf).argtypes = [vp, i64, ct]
            getattr(_lib, "oracle_lbk_collide_" + suf).argtypes = [vp, i64, ct]
            getattr(_lib, "oracle_lbk_equilibrium_" + suf).argtypes = [vp, vp, vp, vp, i64]
    return _lib


def _suffix(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError("oracle supports float32 and float64 only")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run(f, omega, nsteps, bc="cavity", u0=0.1):
    """Return the state after ``nsteps`` full steps from ``f`` (not modified)."""
    suf = _suffix(f.dtype)
    a = np.ascontiguousarray(f).copy()
    b = np.empty_like(a)
    _, nx, ny = a.shape
    getattr(lib(), "oracle_run_" + suf)(_ptr(a), _ptr(b), nx, ny, omega,
                                          1 if bc == "cavity" else 0, u0, nsteps)
    return a


def step(f, omega, bc="cavity", u0=0.1):
    """One full step; returns (f_new, rho, ux, uy)."""
    suf = _suffix(f.dtype)
    a = np.ascontiguousarray(f)
    _, nx, ny = a.shape
    out = np.empty_like(a)
    rho, ux, uy = (np.empty((nx, ny), a.dtype) for _ in range(3))
    getattr(lib(), "oracle_step_" + suf)(_ptr(a), _ptr(out), nx, ny, omega,
                                           1 if bc == "cavity" else 0, u0,
                                           _ptr(rho), _ptr(ux), _ptr(uy))
    return out, rho, ux, uy


def collide(f, omega):
    """opt1-order collide in place on a C-contiguous (9, ...) array."""
    getattr(lib(), "oracle_collide_" + _suffix(f.dtype))(_ptr(f), f[0].size, omega)


def lbk_collide(f, omega):
    """_lbkernels-formula collide in place (c/d2q9.h:121-131)."""
    getattr(lib(), "oracle_lbk_collide_" + _suffix(f.dtype))(_ptr(f), f[0].size, omega)


def lbk_equilibrium(rho, ux, uy, f):
    """_lbkernels-formula equilibrium into f (c/d2q9.h:98-108)."""
    getattr(lib(), "oracle_lbk_equilibrium_" + _suffix(f.dtype))(
        _ptr(rho), _ptr(ux), _ptr(uy), _ptr(f), rho.size)
