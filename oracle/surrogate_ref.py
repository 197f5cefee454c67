"""CPU oracle for the warm-start surrogate and its initializer.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference has no learned component, so nothing here restates reference
code except the equilibrium (``cavity_opt1.py:100-132``, via
``d2q9_numpy.equilibrium``).  What pins each function:

* ``conv3x3`` / ``conv3x3_backward``: ``conv3x3`` against a direct four-loop
  sum over taps and channels; the backward against central finite differences
  of the whole network's forward, per layer and along random directions, and
  the adjoint identity
  <im2col(x), y> = <x, col2im(y)>.
* ``avgpool2`` / ``upsample2`` and their backwards: adjoint identities and
  hand-computed 2x2 cases.
* ``unet_forward`` / ``unet_backward``: central finite differences of the
  FP64 forward along random directions in parameter and input space.
* ``adam``: the published update (Kingma and Ba, 2015, Algorithm 1), checked
  on a quadratic whose minimum it must approach and on the first step, whose
  size is lr * sign(g).
* ``warm_init``: bilinear interpolation is exact for fields linear in x and
  y away from the clamped border, and, with wall ghosts, up to the walls for
  linear profiles that meet the wall and lid velocities (u = 0 on the walls,
  ux = U on the lid) with zero normal density gradient; the cubic form is
  exact for quadratics in the interior (the pool x pool node average of a
  quadratic is its centre value plus (pool^2 - 1)/(24 pool^2) times its
  Laplacian, which the 5-point stencil gives exactly, and Catmull-Rom
  reproduces quadratics), and up to the walls for fields linear in the
  wall-normal distance; constant fields give
  the equilibrium of
  ``d2q9_numpy``; the non-equilibrium part has zero mass and momentum and
  second moment -(2/3)(1 - omega) rho / omega * S (Chapman-Enskog, D2Q9
  fourth-order isotropy).

With ``bf16=True`` the forward and backward round every GEMM operand to
bfloat16 (nearest even), which is where the GPU path rounds; the arithmetic
is otherwise FP64.
"""

import numpy as np

from . import d2q9_numpy as o
from .lowprec import _bf16_bits_to_f32, _f32_to_bf16_bits


def bf16(x):
    """Round to bfloat16 (via float32, nearest even), returned as float64."""
    x = np.asarray(x, dtype=np.float64)
    return _bf16_bits_to_f32(_f32_to_bf16_bits(x.astype(np.float32))).astype(np.float64).reshape(x.shape)


def _maybe(x, on):
    return bf16(x) if on else x


# -- layers (NHWC) -------------------------------------------------------------

def im2col(x):
    """x (B, H, W, C) -> cols (B*H*W, 9*C), column index (dy*3 + dx)*C + c."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    taps = [xp[:, dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)]
    return np.stack(taps, axis=3).reshape(B * H * W, 9 * C)


def col2im(cols, shape):
    """Adjoint of im2col."""
    B, H, W, C = shape
    c = cols.reshape(B, H, W, 9, C)
    xp = np.zeros((B, H + 2, W + 2, C), dtype=cols.dtype)
    for dy in range(3):
        for dx in range(3):
            xp[:, dy:dy + H, dx:dx + W] += c[:, :, :, dy * 3 + dx]
    return xp[:, 1:-1, 1:-1].copy()


def conv3x3(x, w, b, relu, bf=False):
    """3x3 same convolution, w (Cout, 9*Cin), b (Cout,)."""
    B, H, W, _ = x.shape
    y = _maybe(im2col(x), bf) @ _maybe(w, bf).T + b
    if relu:
        y = np.maximum(y, 0.0)
    return y.reshape(B, H, W, -1)


def conv3x3_direct(x, w, b):
    """Four-loop restatement of conv3x3 (no ReLU), for pinning im2col."""
    B, H, W, C = x.shape
    cout = w.shape[0]
    wk = w.reshape(cout, 3, 3, C)
    y = np.zeros((B, H, W, cout))
    for dy in range(3):
        for dx in range(3):
            for yy in range(H):
                sy = yy + dy - 1
                if sy < 0 or sy >= H:
                    continue
                for xx in range(W):
                    sx = xx + dx - 1
                    if sx < 0 or sx >= W:
                        continue
                    y[:, yy, xx, :] += x[:, sy, sx, :] @ wk[:, dy, dx, :].T
    return y + b


def conv3x3_backward(x, w, y, dy, relu, need_dx, bf=False):
    """Gradients of conv3x3 given the layer output y and dL/dy."""
    B, H, W, _ = x.shape
    cout = w.shape[0]
    g = dy.reshape(-1, cout)
    if relu:
        g = np.where(y.reshape(-1, cout) > 0.0, g, 0.0)
    gb = _maybe(g, bf)
    cols = _maybe(im2col(x), bf)
    dw = gb.T @ cols
    db = g.sum(axis=0)
    dx = col2im(gb @ _maybe(w, bf), x.shape) if need_dx else None
    return dx, dw, db


def avgpool2(x):
    B, H, W, C = x.shape
    return 0.25 * x.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def avgpool2_backward(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def upsample2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample2_backward(dy):
    B, H, W, C = dy.shape
    return dy.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def mse(y, t):
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def mse_pooled(y, t, f, weight):
    """weight * mean over f x f blocks of (mean of y - t over the block)^2 and its
    gradient with respect to y (NHWC)."""
    B, H, W, C = y.shape
    e = (y - t).reshape(B, H // f, f, W // f, f, C).mean(axis=(2, 4))
    loss = weight * float(np.mean(e * e))
    g = weight * 2.0 * e / e.size / (f * f)
    return loss, np.repeat(np.repeat(g, f, axis=1), f, axis=2)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Kingma and Ba (2015), Algorithm 1; returns updated (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# -- the encoder-decoder -----------------------------------------------------------

def unet_layers(in_ch, widths, mid, out_ch):
    """Layer table [(name, cin, cout, relu)] of the encoder-decoder.

    Encoder level i: two convolutions to widths[i], then a 2x2 average pool
    (the second output is kept as the skip).  Middle: two convolutions to
    ``mid``.  Decoder level i (deepest first): nearest 2x upsample,
    concatenate [upsampled, skip_i], two convolutions to widths[i].  Head: a
    convolution to ``out_ch`` without ReLU."""
    layers, c = [], in_ch
    for i, w in enumerate(widths):
        layers += [("enc%da" % i, c, w, True), ("enc%db" % i, w, w, True)]
        c = w
    layers += [("mida", c, mid, True), ("midb", mid, mid, True)]
    c = mid
    for i in reversed(range(len(widths))):
        w = widths[i]
        layers += [("dec%da" % i, c + w, w, True), ("dec%db" % i, w, w, True)]
        c = w
    layers.append(("head", c, out_ch, False))
    return layers


def init_params(layers, seed=0, bf=False):
    rng = np.random.default_rng(seed)
    p = {}
    for name, cin, cout, relu in layers:
        std = np.sqrt(2.0 / (9 * cin)) if relu else np.sqrt(1.0 / (9 * cin))
        w = rng.normal(0.0, std, size=(cout, 9 * cin))
        b = rng.normal(0.0, 0.05, size=cout)
        p[name] = (bf16(w), bf16(b)) if bf else (w, b)
    return p


def unet_forward(params, x, widths, bf=False):
    """Returns (output, cache)."""
    cache = {}
    h, skips = x, []
    nlev = len(widths)
    for i in range(nlev):
        for s in ("a", "b"):
            name = "enc%d%s" % (i, s)
            cache[name] = (h, None)
            h = conv3x3(h, *params[name], relu=True, bf=bf)
            cache[name] = (cache[name][0], h)
        skips.append(h)
        h = avgpool2(h)
    for s in ("a", "b"):
        name = "mid" + s
        cache[name] = (h, None)
        h = conv3x3(h, *params[name], relu=True, bf=bf)
        cache[name] = (cache[name][0], h)
    for i in reversed(range(nlev)):
        h = np.concatenate([upsample2(h), skips[i]], axis=3)
        for s in ("a", "b"):
            name = "dec%d%s" % (i, s)
            cache[name] = (h, None)
            h = conv3x3(h, *params[name], relu=True, bf=bf)
            cache[name] = (cache[name][0], h)
    cache["head"] = (h, None)
    out = conv3x3(h, *params["head"], relu=False, bf=bf)
    cache["head"] = (h, out)
    return out, cache


def unet_backward(params, cache, dout, widths, bf=False):
    """Returns (grads {name: (dw, db)}, dx of the network input)."""
    grads = {}
    nlev = len(widths)

    def back(name, dy, relu=True, need_dx=True):
        x, y = cache[name]
        dx, dw, db = conv3x3_backward(x, params[name][0], y, dy, relu, need_dx, bf=bf)
        grads[name] = (dw, db)
        return dx

    g = back("head", dout, relu=False)
    dskips = [None] * nlev
    for i in range(nlev):
        g = back("dec%db" % i, g)
        g = back("dec%da" % i, g)
        cu = g.shape[3] - widths[i]
        dskips[i] = g[..., cu:]
        g = upsample2_backward(g[..., :cu])
    g = back("midb", g)
    g = back("mida", g)
    for i in reversed(range(nlev)):
        g = avgpool2_backward(g) + dskips[i]
        g = back("enc%db" % i, g)
        g = back("enc%da" % i, g)
    return grads, g


def network_loss(params, x, t, widths):
    out, _ = unet_forward(params, x, widths)
    return mse(out, t)[0]


# -- warm-start initializer ------------------------------------------------------------

def _bilinear(q, m, px, py):
    """Clamped bilinear interpolation of one channel q (m, m) at cell coordinates."""
    px = np.clip(px, 0.0, m - 1)
    py = np.clip(py, 0.0, m - 1)
    i0 = np.minimum(np.floor(px).astype(np.int64), max(m - 2, 0))
    j0 = np.minimum(np.floor(py).astype(np.int64), max(m - 2, 0))
    i1 = np.minimum(i0 + 1, m - 1)
    j1 = np.minimum(j0 + 1, m - 1)
    tx = px - i0
    ty = py - j0
    a, b, c, d = q[i0, j0], q[i1, j0], q[i0, j1], q[i1, j1]
    return (1 - tx) * ((1 - ty) * a + ty * c) + tx * ((1 - ty) * b + ty * d)


def ghost_padded(q, depth=1):
    """(m+2d, m+2d, 3) array with d rings of ghost cells mirrored across the
    walls: cell -1-j mirrors cell j, cell m+j mirrors m-1-j; velocity ghost =
    2 u_wall - u_mirror (u_wall = 0 on the left, right and bottom walls, (1, 0)
    on the lid y = m - 1/2, in units of U), density ghost = density of the
    mirror cell.  Corner ghosts apply the y rule, then the x rule."""
    q = np.asarray(q, dtype=np.float64)
    m = q.shape[0]
    idx = np.arange(-depth, m + depth)
    mir = np.where(idx < 0, -1 - idx, np.where(idx >= m, 2 * m - 1 - idx, idx))
    g = q[mir][:, mir].copy()
    bottom, top = idx < 0, idx >= m
    g[:, bottom, 1:] = -g[:, bottom, 1:]
    g[:, top, 1] = 2.0 - g[:, top, 1]
    g[:, top, 2] = -g[:, top, 2]
    out = (idx < 0) | (idx >= m)
    g[out, :, 1:] = -g[out, :, 1:]
    return g


def _bilinear_walls(gq, m, px, py):
    """Bilinear interpolation of one ghost-padded (depth 1) channel gq (m+2, m+2)."""
    px = np.clip(px, -1.0, m)
    py = np.clip(py, -1.0, m)
    i0 = np.minimum(np.floor(px).astype(np.int64), m - 1)
    j0 = np.minimum(np.floor(py).astype(np.int64), m - 1)
    tx = px - i0
    ty = py - j0
    a, b = gq[i0 + 1, j0 + 1], gq[i0 + 2, j0 + 1]
    c, d = gq[i0 + 1, j0 + 2], gq[i0 + 2, j0 + 2]
    return (1 - tx) * ((1 - ty) * a + ty * c) + tx * ((1 - ty) * b + ty * d)


def catmull_rom(t):
    """Weights of cells i0-1 .. i0+2 for a point at i0 + t (Catmull and Rom, 1974)."""
    t2, t3 = t * t, t * t * t
    return [0.5 * (-t3 + 2 * t2 - t), 0.5 * (3 * t3 - 5 * t2 + 2), 0.5 * (-3 * t3 + 4 * t2 + t),
            0.5 * (t3 - t2)]


def centre_values(q, pool):
    """Point values at the cell centres from averages over pool x pool nodes,
    q - c * lap(q), c = (pool^2 - 1) / (24 pool^2); ghost-padded by 2."""
    g = ghost_padded(q, 3)
    v = g[1:-1, 1:-1]
    lap = (g[2:, 1:-1] + g[:-2, 1:-1]) + (g[1:-1, 2:] + g[1:-1, :-2]) - 4.0 * v
    c = (pool * pool - 1.0) / (24.0 * pool * pool)
    return v - c * lap


def _cubic(pc, m, px, py):
    """Separable Catmull-Rom interpolation of centre values pc (padded by 2)."""
    px = np.clip(px, -0.5, m - 0.5)
    py = np.clip(py, -0.5, m - 0.5)
    i0 = np.minimum(np.floor(px).astype(np.int64), m - 1)
    j0 = np.minimum(np.floor(py).astype(np.int64), m - 1)
    wx, wy = catmull_rom(px - i0), catmull_rom(py - j0)
    s = 0.0
    for a in range(4):
        r = 0.0
        for b in range(4):
            r = r + wy[b] * pc[i0 - 1 + a + 2, j0 - 1 + b + 2]
        s = s + wx[a] * r
    return s


def sample_fields(q, nx, ny, U, recon="bilinear", pool=2):
    """(rho, ux, uy) at the nodes from cell averages q (m, m, 3).

    recon: "bilinear" (clamped), "walls" (bilinear with mirrored ghosts) or
    "cubic" (ghosts, deconvolution to centre values, Catmull-Rom)."""
    q = np.asarray(q, dtype=np.float64)
    m = q.shape[0]
    k = np.arange(nx, dtype=np.float64)[:, None]
    l = np.arange(ny, dtype=np.float64)[None, :]
    px = np.broadcast_to((k + 0.5) * m / nx - 0.5, (nx, ny))
    py = np.broadcast_to((l + 0.5) * m / ny - 0.5, (nx, ny))
    if recon == "walls":
        g = ghost_padded(q)
        ch = [_bilinear_walls(g[..., c], m, px, py) for c in range(3)]
    elif recon == "cubic":
        pc = centre_values(q, pool)
        ch = [_cubic(pc[..., c], m, px, py) for c in range(3)]
    else:
        ch = [_bilinear(q[..., c], m, px, py) for c in range(3)]
    return 1.0 + U * U * ch[0], U * ch[1], U * ch[2]


def strain_rate(ux, uy):
    """(Sxx, Sxy, Syy) by central differences, one-sided at the walls."""
    def grad(u, axis):
        g = np.empty_like(u)
        n = u.shape[axis]
        if n == 1:
            return np.zeros_like(u)
        sl = lambda a, b: tuple(slice(a, b) if ax == axis else slice(None) for ax in range(2))
        g[sl(1, -1)] = (u[sl(2, None)] - u[sl(None, -2)]) / 2.0
        g[sl(0, 1)] = u[sl(1, 2)] - u[sl(0, 1)]
        g[sl(n - 1, n)] = u[sl(n - 1, n)] - u[sl(n - 2, n - 1)]
        return g
    dxux, dyux = grad(ux, 0), grad(ux, 1)
    dxuy, dyuy = grad(uy, 0), grad(uy, 1)
    return dxux, 0.5 * (dyux + dxuy), dyuy


def neq_post_collision(rho, sxx, sxy, syy, omega):
    """(1 - omega) * (-3 w_i rho / omega) (c_i c_i - I/3) : S, shape (9, nx, ny)."""
    w = o.weights()
    out = np.empty((9,) + rho.shape)
    for i in range(9):
        cx, cy = o.C_IC[i]
        qs = (cx * cx - 1.0 / 3) * sxx + 2.0 * cx * cy * sxy + (cy * cy - 1.0 / 3) * syy
        out[i] = (1.0 - omega) * (-3.0 * rho / omega) * w[i] * qs
    return out


def warm_init(q, nx, ny, U, omega, neq, recon="bilinear", pool=2):
    rho, ux, uy = sample_fields(q, nx, ny, U, recon, pool)
    f = o.equilibrium(rho, ux, uy)
    if neq:
        f = f + neq_post_collision(rho, *strain_rate(ux, uy), omega)
    return f
