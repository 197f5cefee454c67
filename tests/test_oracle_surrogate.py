"""Pinning of the surrogate oracle (CPU only, FP64)."""

import numpy as np
import pytest

from oracle import d2q9_numpy as o
from oracle import surrogate_ref as s

WIDTHS = (4, 6)
LAYERS = s.unet_layers(3, WIDTHS, 8, 3)


def _net(seed=0, B=2, H=8, W=8):
    rng = np.random.default_rng(seed)
    params = s.init_params(LAYERS, seed=seed)
    x = rng.normal(size=(B, H, W, 3))
    t = rng.normal(size=(B, H, W, 3))
    return params, x, t


def test_im2col_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 5, 7, 3))
    w = rng.normal(size=(4, 27))
    b = rng.normal(size=4)
    np.testing.assert_allclose(s.conv3x3(x, w, b, relu=False), s.conv3x3_direct(x, w, b),
                               rtol=1e-13, atol=1e-13)


def test_tap_order_is_row_then_column():
    # a weight on tap (dy, dx) = (0, 1) alone reads the pixel above (row - 1)
    x = np.zeros((1, 3, 3, 1))
    x[0, 0, 1, 0] = 1.0
    w = np.zeros((1, 9))
    w[0, 0 * 3 + 1] = 1.0
    y = s.conv3x3(x, w, np.zeros(1), relu=False)
    assert y[0, 1, 1, 0] == 1.0 and np.count_nonzero(y) == 1


@pytest.mark.parametrize("fwd,bwd,shape", [
    (s.im2col, lambda c, shp: s.col2im(c, shp), (2, 4, 6, 3)),
    (s.avgpool2, lambda d, shp: s.avgpool2_backward(d), (2, 4, 6, 3)),
    (s.upsample2, lambda d, shp: s.upsample2_backward(d), (2, 3, 5, 2)),
])
def test_adjoint_identities(fwd, bwd, shape):
    rng = np.random.default_rng(2)
    x = rng.normal(size=shape)
    y = rng.normal(size=fwd(x).shape)
    assert np.isclose(np.vdot(fwd(x), y), np.vdot(x, bwd(y, shape)), rtol=1e-12)


def test_pool_and_upsample_by_hand():
    x = np.arange(16.0).reshape(1, 4, 4, 1)
    np.testing.assert_array_equal(s.avgpool2(x)[0, :, :, 0], [[2.5, 4.5], [10.5, 12.5]])
    u = s.upsample2(np.array([[1.0, 2.0]]).reshape(1, 1, 2, 1))[0, :, :, 0]
    np.testing.assert_array_equal(u, [[1, 1, 2, 2], [1, 1, 2, 2]])


def test_layer_table_parameter_count():
    # the production network: ~1.2 M weights
    layers = s.unet_layers(6, (32, 64, 128), 128, 3)
    n = sum(9 * cin * cout + cout for _, cin, cout, _ in layers)
    assert 1.0e6 < n < 1.4e6


def _directional(params, x, t, vp, vx, eps=1e-6):
    def shifted(sign):
        p2 = {k: (w + sign * eps * vp[k][0], b + sign * eps * vp[k][1]) for k, (w, b) in params.items()}
        return s.network_loss(p2, x + sign * eps * vx, t, WIDTHS)
    return (shifted(+1) - shifted(-1)) / (2 * eps)


def _analytic(params, x, t, vp, vx):
    out, cache = s.unet_forward(params, x, WIDTHS)
    _, dout = s.mse(out, t)
    grads, dx = s.unet_backward(params, cache, dout, WIDTHS)
    return sum(np.vdot(grads[k][0], vp[k][0]) + np.vdot(grads[k][1], vp[k][1]) for k in params) \
        + np.vdot(dx, vx)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_backward_matches_finite_differences(seed):
    params, x, t = _net(seed)
    rng = np.random.default_rng(100 + seed)
    vp = {k: (rng.normal(size=w.shape), rng.normal(size=b.shape)) for k, (w, b) in params.items()}
    vx = rng.normal(size=x.shape)
    fd = _directional(params, x, t, vp, vx)
    an = _analytic(params, x, t, vp, vx)
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_backward_per_layer_finite_differences():
    # one direction per layer, so a wrong gradient in any single layer shows
    params, x, t = _net(3)
    rng = np.random.default_rng(7)
    zero = {k: (np.zeros_like(w), np.zeros_like(b)) for k, (w, b) in params.items()}
    for name in params:
        vp = dict(zero)
        vp[name] = (rng.normal(size=params[name][0].shape), rng.normal(size=params[name][1].shape))
        fd = _directional(params, x, t, vp, np.zeros_like(x))
        an = _analytic(params, x, t, vp, np.zeros_like(x))
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-8), name


def test_planted_tap_flip_is_caught(monkeypatch):
    # a col2im that scatters every tap to the mirrored offset is a plausible
    # indexing mistake; the finite-difference check must reject it
    params, x, t = _net(4)
    rng = np.random.default_rng(8)
    vp = {k: (np.zeros_like(w), np.zeros_like(b)) for k, (w, b) in params.items()}
    vx = rng.normal(size=x.shape)
    good = s.col2im

    def flipped(cols, shape):
        B, H, W, C = shape
        c = cols.reshape(B, H, W, 9, C)[:, :, :, ::-1]
        return good(c.reshape(cols.shape), shape)

    fd = _directional(params, x, t, vp, vx)
    monkeypatch.setattr(s, "col2im", flipped)
    an = _analytic(params, x, t, vp, vx)
    assert abs(fd - an) > 1e-3 * abs(fd)


def test_bf16_emulation_is_close_to_fp64():
    params, x, t = _net(5)
    params = {k: (s.bf16(w), s.bf16(b)) for k, (w, b) in params.items()}
    out64, _ = s.unet_forward(params, x, WIDTHS)
    out16, _ = s.unet_forward(params, x, WIDTHS, bf=True)
    rel = np.linalg.norm(out16 - out64) / np.linalg.norm(out64)
    assert 1e-5 < rel < 3e-2


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1 + 2^-7: even is 1
    assert s.bf16(1.0 + 2.0 ** -8) == 1.0
    assert s.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert s.bf16(-2.5) == -2.5


def test_pooled_error_term():
    rng = np.random.default_rng(13)
    y, t = rng.normal(size=(2, 8, 12, 3)), rng.normal(size=(2, 8, 12, 3))
    loss, g = s.mse_pooled(y, t, 4, 2.5)
    e = s.avgpool2(s.avgpool2(y - t))  # 4 x 4 block means by two 2 x 2 pools
    assert np.isclose(loss, 2.5 * np.mean(e * e), rtol=1e-13)
    d = np.zeros_like(y)
    d[1, 5, 7, 2] = 1e-6
    fd = (s.mse_pooled(y + d, t, 4, 2.5)[0] - s.mse_pooled(y - d, t, 4, 2.5)[0]) / 2e-6
    assert np.isclose(fd, g[1, 5, 7, 2], rtol=1e-6)


def test_adam_first_step_and_quadratic():
    g = np.array([3.0, -0.2, 1e-3])
    p, m, v = s.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)
    x, m, v = np.array([5.0, -3.0]), np.zeros(2), np.zeros(2)
    for k in range(1, 3001):
        x, m, v = s.adam(x, 2.0 * (x - np.array([1.0, 2.0])), m, v, k, lr=0.01)
    np.testing.assert_allclose(x, [1.0, 2.0], atol=1e-3)


def test_mse_gradient():
    rng = np.random.default_rng(9)
    y, t = rng.normal(size=(3, 4)), rng.normal(size=(3, 4))
    loss, dy = s.mse(y, t)
    assert np.isclose(loss, np.mean((y - t) ** 2))
    e = np.zeros_like(y)
    e[1, 2] = 1e-6
    assert np.isclose((s.mse(y + e, t)[0] - s.mse(y - e, t)[0]) / 2e-6, dy[1, 2], rtol=1e-6)


# -- warm-start initializer -------------------------------------------------------

def test_warm_init_constant_field_is_equilibrium():
    q = np.zeros((8, 8, 3))
    q[..., 0], q[..., 1], q[..., 2] = 0.5, 0.3, -0.2
    U = 0.1
    f = s.warm_init(q, 20, 12, U, 1.2, neq=True)
    ref = o.equilibrium(np.full((20, 12), 1 + 0.5 * U * U), np.full((20, 12), 0.03),
                        np.full((20, 12), -0.02))
    np.testing.assert_allclose(f, ref, rtol=0, atol=1e-15)


def test_bilinear_exact_for_linear_fields():
    m, nx, ny, U = 16, 64, 48, 0.05
    i = np.arange(m, dtype=float)[:, None, None]
    j = np.arange(m, dtype=float)[None, :, None]
    coef = np.array([0.3, -1.1, 0.7])
    q = 0.2 + coef * (i / m) + 0.5 * coef * (j / m)
    rho, ux, uy = s.sample_fields(q, nx, ny, U)
    k = np.arange(nx)[:, None]
    l = np.arange(ny)[None, :]
    px = (k + 0.5) * m / nx - 0.5
    py = (l + 0.5) * m / ny - 0.5
    # nodes beyond the outermost cell centres take the value there (clamping)
    cx = np.clip(np.broadcast_to(px, (nx, ny)), 0, m - 1)[..., None]
    cy = np.clip(np.broadcast_to(py, (nx, ny)), 0, m - 1)[..., None]
    exact = 0.2 + coef * (cx / m) + 0.5 * coef * (cy / m)
    assert (px < 0).any() and (py > m - 1).any()
    np.testing.assert_allclose((rho - 1) / U ** 2, exact[..., 0], atol=1e-12)
    np.testing.assert_allclose(ux / U, exact[..., 1], atol=1e-12)
    np.testing.assert_allclose(uy / U, exact[..., 2], atol=1e-12)


def test_wall_ghosts_exact_for_profiles_meeting_the_walls():
    # Cell averages of fields linear in x and y that vanish on the walls
    # x = -1/2, y = -1/2 (cell units) are reproduced exactly at every node,
    # including the half cell next to each wall, where clamping is wrong.
    m, nx, ny, U = 8, 32, 24, 0.07
    i = np.arange(m, dtype=float)[:, None]
    j = np.arange(m, dtype=float)[None, :]
    a, b = 0.03, -0.02
    q = np.zeros((m, m, 3))
    q[..., 1] = a * (i + 0.5) * (j + 0.5)   # bilinear, zero on both walls
    q[..., 2] = b * (i + 0.5) + 0 * j       # zero on the left wall only ...
    q[..., 0] = 0.4 + 0 * i * j             # uniform density
    rho, ux, uy = s.sample_fields(q, nx, ny, U, recon="walls")
    px = (np.arange(nx)[:, None] + 0.5) * m / nx - 0.5
    py = (np.arange(ny)[None, :] + 0.5) * m / ny - 0.5
    # ux is exact next to the left and bottom walls (it vanishes on both)
    near = np.broadcast_to((px <= m - 1) & (py <= m - 1), (nx, ny))
    np.testing.assert_allclose((ux / U)[near], np.broadcast_to(a * (px + 0.5) * (py + 0.5), (nx, ny))[near],
                               atol=1e-13)
    # uy vanishes on the left wall only: exact next to it, away from the
    # half cells along the bottom wall and the lid (uy = 0 on both)
    left = np.broadcast_to((px <= m - 1) & (py >= 0) & (py <= m - 1), (nx, ny))
    np.testing.assert_allclose((uy / U)[left], np.broadcast_to(b * (px + 0.5), (nx, ny))[left], atol=1e-13)
    assert (px < 0).any() and (py < 0).any()
    np.testing.assert_allclose(rho, 1 + 0.4 * U * U, rtol=1e-15)
    # the clamped form misses the wall layer
    _, ux_c, _ = s.sample_fields(q, nx, ny, U, recon="bilinear")
    assert np.abs(ux_c[0, 0] / U - a * (px[0, 0] + 0.5) * (py[0, 0] + 0.5)) > 1e-3


def test_wall_ghosts_lid_profile():
    # ux = U (1 - c (m - 1/2 - y)) meets the lid value U at y = m - 1/2 and is
    # reproduced up to the lid; the mirror rule sets the ghost to 2 U - u
    m, nx, U, c = 8, 32, 0.1, 0.1
    j = np.arange(m, dtype=float)
    q = np.zeros((m, m, 3))
    q[..., 1] = (1.0 - c * (m - 0.5 - (j + 0.0)))[None, :]
    g = s.ghost_padded(q)
    np.testing.assert_allclose(g[1:-1, -1, 1], 2.0 - q[:, -1, 1])
    assert np.all(g[0, 1:-1, 1] == -q[0, :, 1]) and np.all(g[1:-1, 0, 2] == -q[:, 0, 2])
    _, ux, _ = s.sample_fields(q, nx, nx, U, recon="walls")
    py = (np.arange(nx) + 0.5) * m / nx - 0.5
    top = py >= m - 1
    np.testing.assert_allclose(ux[nx // 2, top] / U, 1.0 - c * (m - 0.5 - py[top]), atol=1e-13)


def _node_average(u, pool):
    n = u.shape[0]
    return u.reshape(n // pool, pool, n // pool, pool, *u.shape[2:]).mean(axis=(1, 3))


@pytest.mark.parametrize("pool", [1, 2, 4])
def test_cubic_exact_for_quadratics_in_the_interior(pool):
    m = 12
    n = m * pool
    x = (np.arange(n) + 0.5) / pool - 0.5  # node positions in cell units
    X, Y = np.meshgrid(x, x, indexing="ij")
    fields = [0.3 + 0.02 * X * X - 0.01 * X * Y + 0.005 * Y,
              0.1 * X - 0.03 * Y * Y + 0.02 * X * Y, -0.2 + 0.01 * X * X + 0.04 * Y * Y]
    true = np.stack(fields, -1)
    q = _node_average(true, pool)
    U = 0.1
    rho, ux, uy = s.sample_fields(q, n, n, U, recon="cubic", pool=pool)
    got = np.stack([(rho - 1) / U ** 2, ux / U, uy / U], -1)
    inner = (X >= 2) & (X <= m - 3) & (Y >= 2) & (Y <= m - 3)
    np.testing.assert_allclose(got[inner], true[inner], atol=1e-12)
    # without the deconvolution the same reconstruction is off for pool > 1
    if pool > 1:
        plain = s._cubic(s.ghost_padded(q, 2)[..., 1], m, X, Y)
        assert np.abs(plain[inner] - true[..., 1][inner]).max() > 1e-4


def test_cubic_exact_up_to_walls_for_linear_wall_profiles():
    # u vanishing linearly on the left and bottom walls; ux meeting the lid
    # linearly: ghosts continue these profiles exactly
    m, pool = 10, 2
    n = m * pool
    x = (np.arange(n) + 0.5) / pool - 0.5
    X, Y = np.meshgrid(x, x, indexing="ij")
    uy = 0.02 * (X + 0.5) * (Y + 0.5)
    ux = 1.0 - 0.05 * (m - 0.5 - Y)
    true = np.stack([0.7 + 0 * X, ux, uy], -1)
    q = _node_average(true, pool)
    rho, gx, gy = s.sample_fields(q, n, n, 1.0, recon="cubic", pool=pool)
    left_bottom = (X <= m / 2) & (Y <= m / 2)
    np.testing.assert_allclose(gy[left_bottom], uy[left_bottom], atol=1e-12)
    top = (Y >= m / 2) & (X >= 2) & (X <= m - 3)
    np.testing.assert_allclose(gx[top], ux[top], atol=1e-12)
    np.testing.assert_allclose(rho, 1.7, rtol=1e-14)


def test_ghost_mirror_indices_deep():
    q = np.random.default_rng(3).normal(size=(5, 5, 3))
    g = s.ghost_padded(q, 3)
    # cell -3 mirrors cell 2, cell m+2 mirrors cell m-3 (x ghosts flip velocity)
    np.testing.assert_array_equal(g[0, 3:-3, 0], q[2, :, 0])
    np.testing.assert_array_equal(g[0, 3:-3, 1:], -q[2, :, 1:])
    np.testing.assert_array_equal(g[-1, 3:-3, 2], -q[5 - 3, :, 2])
    np.testing.assert_array_equal(g[3:-3, -1, 1], 2.0 - q[:, 2, 1])


def test_neq_moments_match_chapman_enskog():
    rng = np.random.default_rng(11)
    rho = 1 + 0.01 * rng.normal(size=(5, 6))
    sxx, sxy, syy = (1e-3 * rng.normal(size=(5, 6)) for _ in range(3))
    omega = 1.3
    fn = s.neq_post_collision(rho, sxx, sxy, syy, omega)
    c = o.C_IC.astype(float)
    assert np.abs(fn.sum(axis=0)).max() < 1e-18
    assert np.abs(np.tensordot(c[:, 0], fn, 1)).max() < 1e-18
    assert np.abs(np.tensordot(c[:, 1], fn, 1)).max() < 1e-18
    fac = -(2.0 / 3.0) * (1 - omega) * rho / omega
    np.testing.assert_allclose(np.tensordot(c[:, 0] * c[:, 0], fn, 1), fac * sxx, rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(np.tensordot(c[:, 0] * c[:, 1], fn, 1), fac * sxy, rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(np.tensordot(c[:, 1] * c[:, 1], fn, 1), fac * syy, rtol=1e-12, atol=1e-20)


def test_strain_rate_of_linear_shear():
    k = np.arange(10, dtype=float)[:, None]
    l = np.arange(7, dtype=float)[None, :]
    ux = 0.002 * l + 0 * k
    uy = 0.001 * k + 0 * l
    sxx, sxy, syy = s.strain_rate(ux, uy)
    np.testing.assert_allclose(sxy, 0.5 * (0.002 + 0.001), rtol=1e-12)
    assert np.abs(sxx).max() == 0 and np.abs(syy).max() == 0
