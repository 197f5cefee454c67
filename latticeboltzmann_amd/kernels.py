"""Device-level calls of the D2Q9 kernels on torch tensors (zero copy).

Each function validates shapes, dtypes, devices and contiguity on the host,
then calls the C ABI (``include/lbm_d2q9.h``) on torch's current stream.
Population tensors are C-contiguous, float32 or float64, on a ROCm device,
in one of two layouts:

* ``LAYOUT_SOA``: shape ``(9, nx, ny)``, the reference's ``f_ikl``;
* ``LAYOUT_ROWS``: shape ``(nx, 9, ny)``, the nine channel segments of each
  x-row adjacent (the layout the device-resident solver keeps in HBM).
"""

import torch

from . import _native as nat

_SUFFIX = {torch.float64: "f64", torch.float32: "f32"}


def _suffix(t, what):
    try:
        return _SUFFIX[t.dtype]
    except KeyError:
        raise TypeError("%s: dtype must be float32 or float64, got %s" % (what, t.dtype)) from None


def _require_device(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise TypeError("expected a torch tensor on a ROCm device")
        if not t.is_contiguous():
            raise TypeError("tensor must be C-contiguous")


def _populations(f, what, layout=nat.LAYOUT_SOA):
    """(nx, ny) of a population tensor in ``layout``."""
    _require_device(f)
    if layout == nat.LAYOUT_SOA:
        if f.dim() != 3 or f.shape[0] != 9:
            raise ValueError("%s: populations must have shape (9, nx, ny), got %s"
                             % (what, tuple(f.shape)))
        return f.shape[1], f.shape[2]
    if layout == nat.LAYOUT_ROWS:
        if f.dim() != 3 or f.shape[1] != 9:
            raise ValueError("%s: row-interleaved populations must have shape (nx, 9, ny), got %s"
                             % (what, tuple(f.shape)))
        return f.shape[0], f.shape[2]
    raise ValueError("%s: unknown layout %r" % (what, layout))


def _same(a, b, what):
    if b is None:
        return
    if b.dtype != a.dtype:
        raise TypeError("%s: all arrays must share one dtype" % what)


def equilibrium(rho, ux, uy, out=None):
    """f_eq for every node; returns (or fills) a ``(9, *rho.shape)`` tensor."""
    _require_device(rho, ux, uy, out)
    suf = _suffix(rho, "equilibrium")
    for t in (ux, uy, out):
        _same(rho, t, "equilibrium")
    if ux.shape != rho.shape or uy.shape != rho.shape:
        raise ValueError("equilibrium: rho, ux and uy must have the same shape")
    if out is None:
        out = torch.empty((9,) + tuple(rho.shape), dtype=rho.dtype, device=rho.device)
    elif out.numel() != 9 * rho.numel() or out.shape[0] != 9:
        raise ValueError("equilibrium: output must hold 9 x %d values" % rho.numel())
    nat.call("lbm_equilibrium_" + suf, nat.ptr(rho), nat.ptr(ux), nat.ptr(uy), nat.ptr(out),
             rho.numel(), nat.stream_ptr())
    return out


def collide_(f, omega, moments=False):
    """BGK collide in place on ``f`` (shape ``(9, ...)``).

    With ``moments=True`` returns the pre-collision ``(rho, ux, uy)``.
    """
    _require_device(f)
    suf = _suffix(f, "collide")
    if f.dim() < 2 or f.shape[0] != 9:
        raise ValueError("collide: populations must have shape (9, ...)")
    n = f.numel() // 9
    rho = ux = uy = None
    if moments:
        rho, ux, uy = (torch.empty(f.shape[1:], dtype=f.dtype, device=f.device) for _ in range(3))
    nat.call("lbm_collide_" + suf, nat.ptr(f), n, float(omega), nat.ptr(rho), nat.ptr(ux),
             nat.ptr(uy), nat.stream_ptr())
    return (rho, ux, uy) if moments else None


def moments(f, layout=nat.LAYOUT_SOA):
    """(rho, ux, uy) of populations ``f``, each ``(nx, ny)``; f unchanged."""
    nx, ny = _populations(f, "moments", layout)
    suf = _suffix(f, "moments")
    rho, ux, uy = (torch.empty((nx, ny), dtype=f.dtype, device=f.device) for _ in range(3))
    nat.call("lbm_moments_" + suf, nat.ptr(f), nx, ny, int(layout), nat.ptr(rho), nat.ptr(ux),
             nat.ptr(uy), nat.stream_ptr())
    return rho, ux, uy


def relayout(f, src_layout, dst_layout, out=None):
    """Copy populations between layouts (bit-exact permutation)."""
    nx, ny = _populations(f, "relayout", src_layout)
    shape = (9, nx, ny) if dst_layout == nat.LAYOUT_SOA else (nx, 9, ny)
    if out is None:
        out = torch.empty(shape, dtype=f.dtype, device=f.device)
    _require_device(out)
    if tuple(out.shape) != shape or out.dtype != f.dtype:
        raise ValueError("relayout: output must have shape %s and the input's dtype" % (shape,))
    nat.call("lbm_relayout", nat.ptr(f), nat.ptr(out), nx, ny, f.element_size(),
             int(src_layout), int(dst_layout), nat.stream_ptr())
    return out


def stream(f, out=None):
    """Periodic streaming ``f -> out`` (a permutation of elements of any size)."""
    nx, ny = _populations(f, "stream")
    if f.element_size() not in (1, 2, 4, 8):
        raise TypeError("stream: element size must be 1, 2, 4 or 8 bytes")
    if out is None:
        out = torch.empty_like(f)
    _require_device(out)
    if out.shape != f.shape or out.dtype != f.dtype:
        raise ValueError("stream: output must match the input")
    nat.call("lbm_stream_periodic", nat.ptr(f), nat.ptr(out), nx, ny, f.element_size(),
             nat.stream_ptr())
    return out


def stream_bounce_back(f, u0=0.1, out=None):
    """Streaming with the cavity walls and moving lid, ``f -> out``."""
    nx, ny = _populations(f, "stream_and_bounce_back")
    suf = _suffix(f, "stream_and_bounce_back")
    if out is None:
        out = torch.empty_like(f)
    _require_device(out)
    if out.shape != f.shape or out.dtype != f.dtype:
        raise ValueError("stream_and_bounce_back: output must match the input")
    nat.call("lbm_stream_bounce_back_" + suf, nat.ptr(f), nat.ptr(out), nx, ny, float(u0),
             nat.stream_ptr())
    return out


def step(src, dst, omega, bc=nat.BC_CAVITY, u0=0.1, moments=None, variant=-1,
         layout=nat.LAYOUT_SOA):
    """One fused step ``src -> dst`` (pull streaming, walls if cavity, collide).

    ``moments`` may be a tuple ``(rho, ux, uy)`` of ``(nx, ny)`` tensors that
    receive the pre-collision moments.
    """
    nx, ny = _populations(src, "step", layout)
    _require_device(dst)
    suf = _suffix(src, "step")
    _same(src, dst, "step")
    if dst.shape != src.shape:
        raise ValueError("step: output must match the input")
    rho = ux = uy = None
    if moments is not None:
        rho, ux, uy = moments
        _require_device(rho, ux, uy)
        for t in (rho, ux, uy):
            _same(src, t, "step")
            if t.shape != (nx, ny):
                raise ValueError("step: moment arrays must have shape (nx, ny)")
    nat.call("lbm_step_layout_" + suf, nat.ptr(src), nat.ptr(dst), nx, ny, int(layout),
             float(omega), int(bc), float(u0), nat.ptr(rho), nat.ptr(ux), nat.ptr(uy),
             int(variant), nat.stream_ptr())
    return dst


def step_multi_info(variant=-1):
    """(steps per launch, smallest nx, smallest ny) of multi-step variant ``variant``."""
    import ctypes
    k, a, b = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
    nat.call("lbm_step_multi_info", int(variant), ctypes.byref(k), ctypes.byref(a),
             ctypes.byref(b))
    return k.value, a.value, b.value


def step_multi(src, dst, omega, bc=nat.BC_CAVITY, u0=0.1, moments=None, variant=-1,
               layout=nat.LAYOUT_SOA):
    """K fused steps ``src -> dst`` in one launch (temporal blocking; K from
    ``step_multi_info(variant)``), bitwise equal to K ``step`` calls;
    ``moments`` receive the pre-collision moments of the last step."""
    nx, ny = _populations(src, "step_multi", layout)
    _require_device(dst)
    suf = _suffix(src, "step_multi")
    _same(src, dst, "step_multi")
    if dst.shape != src.shape:
        raise ValueError("step_multi: output must match the input")
    rho = ux = uy = None
    if moments is not None:
        rho, ux, uy = moments
        _require_device(rho, ux, uy)
        for t in (rho, ux, uy):
            _same(src, t, "step_multi")
            if t.shape != (nx, ny):
                raise ValueError("step_multi: moment arrays must have shape (nx, ny)")
    nat.call("lbm_step_multi_" + suf, nat.ptr(src), nat.ptr(dst), nx, ny, int(layout),
             float(omega), int(bc), float(u0), nat.ptr(rho), nat.ptr(ux), nat.ptr(uy),
             int(variant), nat.stream_ptr())
    return dst


class Workspace:
    """Device scratch of the two-pass reductions plus a device result slot."""

    def __init__(self, device):
        nbytes = nat.load().lbm_reduce_workspace_bytes()
        self.work = torch.empty(nbytes // 8, dtype=torch.float64, device=device)
        self.out = torch.empty(1, dtype=torch.float64, device=device)


_workspaces = {}


def _workspace(device):
    key = (device.type, device.index)
    if key not in _workspaces:
        _workspaces[key] = Workspace(device)
    return _workspaces[key]


def total_sum(x, ws=None):
    """FP64 sum of every element of ``x`` (float32 or float64); device scalar."""
    _require_device(x)
    suf = _suffix(x, "sum")
    ws = ws or _workspace(x.device)
    nat.call("lbm_sum_" + suf, nat.ptr(x), x.numel(), nat.ptr(ws.work), nat.ptr(ws.out),
             nat.stream_ptr())
    return ws.out.clone()


def velocity_residual(f, ux_prev, uy_prev, ws=None, layout=nat.LAYOUT_SOA):
    """max_j |u_j(f) - u_prev_j| for FP64 populations; u_prev is updated.

    Returns a device scalar (1-element FP64 tensor).
    """
    nx, ny = _populations(f, "velocity_residual", layout)
    if f.dtype != torch.float64:
        raise TypeError("velocity_residual: FP64 populations expected")
    _require_device(ux_prev, uy_prev)
    for t in (ux_prev, uy_prev):
        if t.dtype != torch.float64 or tuple(t.shape) != (nx, ny):
            raise ValueError("velocity_residual: previous velocity must be FP64 (nx, ny)")
    ws = ws or _workspace(f.device)
    nat.call("lbm_velocity_residual_f64", nat.ptr(f), nx, ny, int(layout), nat.ptr(ux_prev),
             nat.ptr(uy_prev), nat.ptr(ws.work), nat.ptr(ws.out), nat.stream_ptr())
    return ws.out.clone()
