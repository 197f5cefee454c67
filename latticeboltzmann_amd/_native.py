"""Loader for ``liblbm_d2q9.so``, the C ABI declared in ``include/lbm_d2q9.h``.

There is no fallback: if the library is missing or no GPU is visible, the
calls raise.  ``torch`` is imported before the library is opened so that the
library binds to the HIP runtime torch has already loaded (both are the
``libamdhip64.so.7`` soname); device pointers and streams are then shared.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_d2q9.so")

LBM_OK = 0
_ERRORS = {
    -1: (ValueError, "a required pointer is NULL"),
    -2: (ValueError, "size out of range"),
    -3: (ValueError, "argument out of range"),
    -4: (ValueError, "input and output buffers overlap"),
    -5: (ValueError, "workspace too small"),
}

BC_PERIODIC = 0
BC_CAVITY = 1
LAYOUT_SOA = 0   # f[i][k][l], the reference's f_ikl
LAYOUT_ROWS = 1  # f[k][i][l], row-interleaved
FMT_F32, FMT_BF16, FMT_F16 = 0, 1, 2
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
COMPUTE_F32, COMPUTE_F64 = 0, 1
WARM_NEQ, WARM_WALLS, WARM_CUBIC = 1, 2, 4

_vp, _i64, _i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
_f64, _f32 = ctypes.c_double, ctypes.c_float

RULE_COPY, RULE_LID, RULE_LID_RHO, RULE_ZERO = 0, 1, 2, 3
GHOST_MAX_RULES = 32
PBC_NONE, PBC_BEFORE_STREAM, PBC_AFTER_STREAM = 0, 1, 2
GHOST_NO_COLLIDE, GHOST_COLLIDE_ONLY = 1, 2


class GhostRule(ctypes.Structure):
    """lbm_ghost_rule (include/lbm_d2q9.h)."""
    _fields_ = [("x0", ctypes.c_int32), ("x1", ctypes.c_int32), ("y0", ctypes.c_int32),
                ("y1", ctypes.c_int32), ("ch", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("dx", ctypes.c_int32), ("dy", ctypes.c_int32), ("src_ch", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("coef", ctypes.c_double)]


class GhostPbc(ctypes.Structure):
    """lbm_ghost_pbc (include/lbm_d2q9.h)."""
    _fields_ = [("mode", ctypes.c_int32), ("pad", ctypes.c_int32), ("rho_in", ctypes.c_double),
                ("rho_out", ctypes.c_double)]


# name -> (restype, argtypes); keep in step with include/lbm_d2q9.h
SIGNATURES = {
    "lbm_version": (ctypes.c_char_p, []),
    "lbm_step_variant_count": (_i32, []),
    "lbm_equilibrium_f64": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp]),
    "lbm_equilibrium_f32": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp]),
    "lbm_collide_f64": (_i32, [_vp, _i64, _f64, _vp, _vp, _vp, _vp]),
    "lbm_collide_f32": (_i32, [_vp, _i64, _f32, _vp, _vp, _vp, _vp]),
    "lbm_moments_f64": (_i32, [_vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp]),
    "lbm_moments_f32": (_i32, [_vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp]),
    "lbm_copy": (_i32, [_vp, _vp, _i64, _i32, _vp]),
    "lbm_step_multi_f64": (_i32, [_vp, _vp, _i64, _i64, _i32, _f64, _i32, _f64, _vp, _vp, _vp,
                                  _i32, _vp]),
    "lbm_step_multi_f32": (_i32, [_vp, _vp, _i64, _i64, _i32, _f32, _i32, _f32, _vp, _vp, _vp,
                                  _i32, _vp]),
    "lbm_step_multi_variant_count": (_i32, []),
    "lbm_step_multi_default_variant": (_i32, []),
    "lbm_step_multi_default_for": (_i32, [_i64, _i64]),
    "lbm_step_multi_default_for_f32": (_i32, [_i64, _i64]),
    "lbm_step_multi_info": (_i32, [_i32, ctypes.POINTER(_i32), ctypes.POINTER(_i64),
                                   ctypes.POINTER(_i64)]),
    "lbm_relayout": (_i32, [_vp, _vp, _i64, _i64, _i32, _i32, _i32, _vp]),
    "lbm_stream_periodic": (_i32, [_vp, _vp, _i64, _i64, _i32, _vp]),
    "lbm_stream_bounce_back_f64": (_i32, [_vp, _vp, _i64, _i64, _f64, _vp]),
    "lbm_stream_bounce_back_f32": (_i32, [_vp, _vp, _i64, _i64, _f32, _vp]),
    "lbm_step_f64": (_i32, [_vp, _vp, _i64, _i64, _f64, _i32, _f64, _vp, _vp, _vp, _i32, _vp]),
    "lbm_step_f32": (_i32, [_vp, _vp, _i64, _i64, _f32, _i32, _f32, _vp, _vp, _vp, _i32, _vp]),
    "lbm_step_layout_f64": (_i32, [_vp, _vp, _i64, _i64, _i32, _f64, _i32, _f64, _vp, _vp, _vp,
                                   _i32, _vp]),
    "lbm_step_layout_f32": (_i32, [_vp, _vp, _i64, _i64, _i32, _f32, _i32, _f32, _vp, _vp, _vp,
                                   _i32, _vp]),
    "lbm_reduce_workspace_bytes": (_i64, []),
    "lbm_sum_f64": (_i32, [_vp, _i64, _vp, _vp, _vp]),
    "lbm_sum_f32": (_i32, [_vp, _i64, _vp, _vp, _vp]),
    "lbm_velocity_residual_f64": (_i32, [_vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    "lbm_noalloc_scratch_bytes": (_i64, []),
    "lbm_velocity_residual_f64_noalloc": (_i32, [_vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp]),
    # reduced-precision storage
    "lbm_dev_variant_count": (_i32, []),
    "lbm_dev_encode": (_i32, [_vp, _vp, _i64, _i64, _i32, _i32, _f64, _vp]),
    "lbm_dev_decode": (_i32, [_vp, _vp, _i64, _i64, _i32, _i32, _f64, _vp]),
    "lbm_dev_step": (_i32, [_vp, _vp, _i32, _i64, _i64, _i32, _f64, _i32, _f64, _f64, _i32,
                            _i32, ctypes.c_uint32, _vp, _vp, _i32, _vp]),
    "lbm_dev_step_multi": (_i32, [_vp, _vp, _i32, _i64, _i64, _i32, _f64, _i32, _f64, _f64,
                                  _i32, _i32, ctypes.c_uint32, ctypes.c_uint32, _i32, _vp, _vp,
                                  _i32, _vp]),
    "lbm_dev_multi_variant_count": (_i32, []),
    "lbm_dev_multi_default_variant": (_i32, []),
    "lbm_dev_multi_info": (_i32, [_i32, ctypes.POINTER(_i32), ctypes.POINTER(_i64),
                                  ctypes.POINTER(_i64)]),
    "lbm_dev_mass": (_i32, [_vp, _i64, _i64, _i32, _f64, _vp, _vp, _vp]),
    "lbm_dev_velocity_residual": (_i32, [_vp, _i64, _i64, _i32, _i32, _f64, _vp, _vp, _vp, _vp,
                                         _vp]),
    # batched instances
    "lbm_batch_variant_count": (_i32, []),
    "lbm_batch_tile_nodes": (_i32, [_i32]),
    "lbm_batch_step": (_i32, [_vp, _vp, _i32, _f32, _vp, _vp, _i32, _i32, _i32, ctypes.c_uint32,
                              _vp, _i32, _i32, _vp]),
    "lbm_batch_amplitude": (_i32, [_vp, _i32, _f32, _vp, _i32, _vp, _i32, _vp, _vp, _vp, _i32,
                                   _vp, _vp]),
    "lbm_batch_sweep_strips": (_i32, [_i32, _i32]),
    "lbm_batch_step_multi": (_i32, [_vp, _vp, _i32, _f32, _vp, _vp, _i32, _i32, _i32, _i32,
                                    ctypes.c_uint32, ctypes.c_uint32, _vp, _vp]),
    "lbm_counter_increment": (_i32, [_vp, _vp]),
    "lbm_counter_set": (_i32, [_vp, ctypes.c_uint32, _vp]),
    # HIP graphs
    "lbm_graph_begin": (_i32, [_vp]),
    "lbm_graph_end": (_i32, [_vp, ctypes.POINTER(ctypes.c_void_p)]),
    "lbm_graph_launch": (_i32, [_vp, _vp]),
    "lbm_graph_destroy": (_i32, [_vp]),
    # warm-start surrogate
    "lbm_nn_gemm_workspace": (_i64, [_i32, _i32, _i32]),
    "lbm_nn_gemm": (_i32, [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _vp, _i64, _vp, _i32, _i32,
                           _i32, _i32, _i32, _i32, _vp, _vp]),
    "lbm_nn_conv3x3": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _i32, _vp, _vp, _vp, _i32, _vp]),
    "lbm_nn_conv3x3_wgrad": (_i32, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp, _i32, _vp, _vp]),
    "lbm_nn_conv3x3_flip_weights": (_i32, [_vp, _i32, _i32, _vp, _vp]),
    "lbm_nn_conv3x3_dgrad": (_i32, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp, _i32, _vp]),
    "lbm_nn_im2col3x3": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _vp]),
    "lbm_nn_col2im3x3": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _i32, _vp]),
    "lbm_nn_relu_backward": (_i32, [_vp, _vp, _vp, _i64, _vp]),
    "lbm_nn_bias_grad_workspace": (_i64, [_i32]),
    "lbm_nn_bias_grad": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp]),
    "lbm_nn_avgpool2": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp]),
    "lbm_nn_avgpool2_backward": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _i32, _vp]),
    "lbm_nn_upsample2": (_i32, [_vp, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _i32, _vp]),
    "lbm_nn_upsample2_backward": (_i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp]),
    "lbm_nn_copy_channels": (_i32, [_vp, _i32, _i32, _vp, _vp, _i32, _i32, _i32, _i64, _i32,
                                    _vp]),
    "lbm_nn_mse": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "lbm_nn_mse_pooled": (_i32, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _f32, _vp, _vp, _vp, _vp]),
    "lbm_nn_adam": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _i32, _vp]),
    "lbm_nn_ema": (_i32, [_vp, _vp, _i64, _f32, _vp]),
    "lbm_nn_to_bf16": (_i32, [_vp, _vp, _i64, _vp]),
    "lbm_warm_init_f64": (_i32, [_vp, _i32, _i64, _i64, _f64, _f64, _i32, _i32, _vp, _vp]),
    "lbm_ghost_equilibrium_f64": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp]),
    "lbm_ghost_equilibrium_f32": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp]),
    "lbm_ghost_step_f64": (_i32, [_vp, _vp, _i64, _i64, _i32, _f64, _vp, _i32, _vp, _i32, _vp,
                                  _vp, _vp, _vp]),
    "lbm_ghost_step_f32": (_i32, [_vp, _vp, _i64, _i64, _i32, _f64, _vp, _i32, _vp, _i32, _vp,
                                  _vp, _vp, _vp]),
}

_lib = None


def load(path=LIB_PATH):
    """Open the library (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "liblbm_d2q9.so is not built; run `python -m latticeboltzmann_amd.build` "
            "(the HIP kernels are the only implementation)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        # a library built from older sources lacks the newer entry points:
        # those raise when called (AttributeError), the others still load
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


_BUILD_ID = None


def build_id(path=LIB_PATH):
    """Content hash of the library file (16 hex digits): names the build the
    autotuner's cached choices were measured with."""
    global _BUILD_ID
    if _BUILD_ID is None:
        import hashlib
        h = hashlib.sha1()
        with open(path, "rb") as fh:
            for chunk in iter(lambda: fh.read(1 << 20), b""):
                h.update(chunk)
        _BUILD_ID = h.hexdigest()[:16]
    return _BUILD_ID


def check(status, what):
    """Raise for a non-zero status returned by a C-ABI call."""
    if status == LBM_OK:
        return
    if status in _ERRORS:
        exc, msg = _ERRORS[status]
        raise exc("%s: %s" % (what, msg))
    raise RuntimeError("%s: HIP error %d" % (what, status))


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("latticeboltzmann_amd needs a ROCm GPU; none is visible")


def stream_ptr(stream=None):
    """hipStream_t of ``stream`` (default: torch's current stream) as an int."""
    s = torch.cuda.current_stream() if stream is None else stream
    return s.cuda_stream


def ptr(t):
    """Device address of a tensor (or None for None)."""
    return None if t is None else t.data_ptr()


def call(name, *args):
    """Call a C-ABI entry point and raise on a non-zero status."""
    status = getattr(load(), name)(*args)
    check(status, name)
