"""Warm-start surrogate kernels and network against the FP64 oracle
(with BF16 rounding emulated where the GPU rounds), the initializer against its
restatement, and the controller driving the real lattice."""

import numpy as np
import pytest
import torch

from oracle import surrogate_ref as s

pytestmark = pytest.mark.gpu


def _nat():
    from latticeboltzmann_amd import _native as nat
    return nat


def _bf(rng, shape, scale=1.0):
    """Random BF16 device tensor and its exact values as float64."""
    t = torch.as_tensor(rng.normal(0.0, scale, size=shape).astype(np.float32)).to(torch.bfloat16)
    return t.cuda(), t.to(torch.float64).numpy()


def _p(t):
    return None if t is None else t.data_ptr()


GEMM_CASES = [
    # M, N, K, a_trans, b_trans, bias, relu, accumulate, ksplit
    (1, 1, 1, False, False, False, False, False, 1),
    (64, 64, 32, False, False, True, True, False, 1),
    (100, 37, 77, True, False, True, False, True, 1),
    (300, 130, 1000, False, True, False, True, False, 1),
    (33, 200, 5000, True, True, False, False, False, 7),
    (128, 96, 4096, False, False, True, True, True, 16),
]


@pytest.mark.parametrize("case", GEMM_CASES)
def test_gemm_matches_float64(gpu, case):
    M, N, K, at, bt, use_bias, relu, acc, ks = case
    nat = _nat()
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A, An = _bf(rng, (K, M) if at else (M, K))
    B, Bn = _bf(rng, (N, K) if bt else (K, N))
    a_ref = An.T if at else An
    b_ref = Bn.T if bt else Bn
    sam, sak = (1, M) if at else (K, 1)
    sbk, sbn = (1, K) if bt else (N, 1)
    ldc = N + 3
    C0 = torch.as_tensor(rng.normal(size=(M, ldc)).astype(np.float32)).cuda()
    C = C0.clone()
    Cb = torch.zeros((M, ldc), dtype=torch.bfloat16, device="cuda")
    bias = torch.as_tensor(rng.normal(size=N).astype(np.float32)).cuda() if use_bias else None
    work = None
    if ks > 1:
        work = torch.empty(nat.load().lbm_nn_gemm_workspace(M, N, ks) // 4, dtype=torch.float32,
                           device="cuda")
    nat.call("lbm_nn_gemm", _p(A), sam, sak, _p(B), sbk, sbn, _p(C), _p(Cb), ldc, _p(bias), M, N, K,
             int(acc), int(relu), ks, _p(work), nat.stream_ptr())
    ref = a_ref @ b_ref
    if acc:
        ref = ref + C0[:, :N].double().cpu().numpy()
    if use_bias:
        ref = ref + bias.double().cpu().numpy()
    if relu:
        ref = np.maximum(ref, 0.0)
    got = C.double().cpu().numpy()
    scale = np.abs(a_ref) @ np.abs(b_ref) + 1.0
    assert np.all(np.abs(got[:, :N] - ref) <= 2e-6 * scale)
    # columns beyond N are untouched; the BF16 copy is the nearest-even rounding
    assert torch.equal(C[:, N:], C0[:, N:])
    assert torch.equal(Cb[:, :N], C[:, :N].to(torch.bfloat16))


def test_gemm_split_k_is_deterministic(gpu):
    nat = _nat()
    rng = np.random.default_rng(5)
    M, N, K, ks = 64, 300, 20000, 24
    A, _ = _bf(rng, (M, K))
    B, _ = _bf(rng, (K, N))
    work = torch.empty(nat.load().lbm_nn_gemm_workspace(M, N, ks) // 4, dtype=torch.float32,
                       device="cuda")
    outs = []
    for _ in range(2):
        C = torch.empty((M, N), dtype=torch.float32, device="cuda")
        nat.call("lbm_nn_gemm", _p(A), K, 1, _p(B), N, 1, _p(C), None, N, None, M, N, K, 0, 0, ks,
                 _p(work), nat.stream_ptr())
        outs.append(C)
    assert torch.equal(outs[0], outs[1])


def test_gemm_rejects_bad_arguments(gpu):
    nat = _nat()
    lib = nat.load()
    x = torch.zeros(64, dtype=torch.bfloat16, device="cuda")
    c = torch.zeros(64, dtype=torch.float32, device="cuda")
    s_ = nat.stream_ptr()
    assert lib.lbm_nn_gemm(None, 1, 1, _p(x), 1, 1, _p(c), None, 1, None, 1, 1, 1, 0, 0, 1, None, s_) == -1
    assert lib.lbm_nn_gemm(_p(x), 1, 1, _p(x), 1, 1, _p(c), None, 1, None, 0, 1, 1, 0, 0, 1, None, s_) == -2
    assert lib.lbm_nn_gemm(_p(x), 1, 1, _p(x), 1, 1, None, _p(x), 1, None, 1, 1, 1, 1, 0, 1, None, s_) == -3
    assert lib.lbm_nn_gemm(_p(x), 1, 1, _p(x), 1, 1, _p(c), None, 1, None, 1, 1, 64, 0, 0, 2, None, s_) == -5


@pytest.mark.parametrize("C", [3, 16])
def test_im2col_col2im(gpu, C):
    # C = 16 takes the 16-byte vector path, C = 3 the scalar one
    nat = _nat()
    rng = np.random.default_rng(6)
    B, H, W = 2, 5, 7
    x, xn = _bf(rng, (B, H, W, C))
    cols = torch.empty((B * H * W, 9 * C), dtype=torch.bfloat16, device="cuda")
    nat.call("lbm_nn_im2col3x3", _p(x), B, H, W, C, _p(cols), nat.stream_ptr())
    assert np.array_equal(cols.double().cpu().numpy(), s.im2col(xn))
    d = torch.as_tensor(rng.normal(size=(B * H * W, 9 * C)).astype(np.float32)).cuda()
    dx = torch.full((B, H, W, C), 0.5, dtype=torch.float32, device="cuda")
    nat.call("lbm_nn_col2im3x3", _p(d), B, H, W, C, _p(dx), 1, nat.stream_ptr())
    ref = 0.5 + s.col2im(d.double().cpu().numpy(), (B, H, W, C))
    np.testing.assert_allclose(dx.double().cpu().numpy(), ref, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("cin,cout,ksplit", [(3, 16, 1), (16, 3, 3), (16, 32, 4), (6, 8, 1)])
def test_conv3x3_implicit_gemm(gpu, cin, cout, ksplit):
    # the gathering convolution GEMMs against im2col + GEMM: forward and weight
    # gradient load the same tiles in the same k order (bitwise equal), the data
    # gradient sums the nine taps inside the GEMM instead of in col2im
    nat = _nat()
    rng = np.random.default_rng(16)
    B, H, W = 2, 9, 7
    P, K, sp = B * H * W, 9 * cin, nat.stream_ptr()
    x, _ = _bf(rng, (B, H, W, cin))
    w, wn = _bf(rng, (cout, 3, 3, cin), 0.2)
    dy, _ = _bf(rng, (B, H, W, cout))
    bias = torch.as_tensor(rng.normal(size=cout).astype(np.float32)).cuda()
    cols = torch.empty((P, K), dtype=torch.bfloat16, device="cuda")
    nat.call("lbm_nn_im2col3x3", _p(x), B, H, W, cin, _p(cols), sp)
    # forward with bias and ReLU, both outputs
    ye, yeb = (torch.empty((P, cout), dtype=t, device="cuda") for t in (torch.float32, torch.bfloat16))
    nat.call("lbm_nn_gemm", _p(cols), K, 1, _p(w), 1, K, _p(ye), _p(yeb), cout, _p(bias),
             P, cout, K, 0, 1, 1, None, sp)
    y, yb = torch.empty_like(ye), torch.empty_like(yeb)
    nat.call("lbm_nn_conv3x3", _p(x), B, H, W, cin, _p(w), cout, _p(bias), _p(y), _p(yb), 1, sp)
    assert torch.equal(y, ye) and torch.equal(yb, yeb)
    # weight gradient, split over pixels
    lib = nat.load()
    work = torch.empty(max(1, int(lib.lbm_nn_gemm_workspace(cout, K, ksplit)) // 4),
                       dtype=torch.float32, device="cuda")
    dwe, dw = (torch.empty((cout, K), dtype=torch.float32, device="cuda") for _ in range(2))
    nat.call("lbm_nn_gemm", _p(dy), 1, cout, _p(cols), K, 1, _p(dwe), None, K, None,
             cout, K, P, 0, 0, ksplit, _p(work), sp)
    nat.call("lbm_nn_conv3x3_wgrad", _p(x), _p(dy), B, H, W, cin, cout, _p(dw), ksplit, _p(work), sp)
    assert torch.equal(dw, dwe)
    # data gradient: against GEMM + col2im, and against float64 on the CPU
    dcols = torch.empty((P, K), dtype=torch.float32, device="cuda")
    nat.call("lbm_nn_gemm", _p(dy), cout, 1, _p(w), K, 1, _p(dcols), None, K, None,
             P, K, cout, 0, 0, 1, None, sp)
    dxe = torch.empty((B, H, W, cin), dtype=torch.float32, device="cuda")
    nat.call("lbm_nn_col2im3x3", _p(dcols), B, H, W, cin, _p(dxe), 0, sp)
    wf = torch.empty(9 * cin * cout, dtype=torch.bfloat16, device="cuda")
    nat.call("lbm_nn_conv3x3_flip_weights", _p(w), cin, cout, _p(wf), sp)
    assert np.array_equal(wf.view(3, 3, cout, cin).double().cpu().numpy(), wn.transpose(1, 2, 0, 3))
    dx = torch.full((B, H, W, cin), 0.25, dtype=torch.float32, device="cuda")
    nat.call("lbm_nn_conv3x3_dgrad", _p(dy), _p(wf), B, H, W, cin, cout, _p(dx), 1, sp)
    ref = 0.25 + s.col2im(dy.double().cpu().numpy().reshape(P, cout) @ wn.reshape(cout, K),
                          (B, H, W, cin))
    np.testing.assert_allclose(dx.double().cpu().numpy(), ref, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(dx.cpu().numpy() - 0.25, dxe.cpu().numpy(), rtol=1e-5, atol=1e-5)
    # argument checks
    assert lib.lbm_nn_conv3x3(None, B, H, W, cin, _p(w), cout, None, _p(y), None, 0, sp) == -1
    assert lib.lbm_nn_conv3x3(_p(x), 0, H, W, cin, _p(w), cout, None, _p(y), None, 0, sp) == -2
    assert lib.lbm_nn_conv3x3_wgrad(_p(x), _p(dy), B, H, W, cin, cout, _p(dw), 2, None, sp) == -5
    assert lib.lbm_nn_conv3x3_dgrad(_p(dy), _p(wf), B, H, W, cin, -1, _p(dx), 0, sp) == -2
    torch.cuda.synchronize()


def test_pool_upsample_concat_kernels(gpu):
    nat = _nat()
    rng = np.random.default_rng(7)
    st = nat.stream_ptr()
    B, H, W, C = 2, 6, 4, 5
    x = torch.as_tensor(rng.normal(size=(B, H, W, C)).astype(np.float32)).cuda()
    xn = x.double().cpu().numpy()
    y = torch.empty((B, H // 2, W // 2, C), dtype=torch.float32, device="cuda")
    yb = torch.empty_like(y, dtype=torch.bfloat16)
    nat.call("lbm_nn_avgpool2", _p(x), B, H, W, C, _p(y), _p(yb), st)
    np.testing.assert_allclose(y.double().cpu().numpy(), s.avgpool2(xn), rtol=1e-6, atol=1e-7)
    assert torch.equal(yb, y.to(torch.bfloat16))
    dx = torch.empty_like(x)
    nat.call("lbm_nn_avgpool2_backward", _p(y), B, H, W, C, _p(dx), 0, st)
    np.testing.assert_allclose(dx.double().cpu().numpy(),
                               s.avgpool2_backward(y.double().cpu().numpy()), rtol=1e-7)
    # upsample into channels [2, 2+C) of a 9-channel buffer
    Ct, off = 9, 2
    up = torch.zeros((B, 2 * H, 2 * W, Ct), dtype=torch.float32, device="cuda")
    upb = torch.zeros_like(up, dtype=torch.bfloat16)
    nat.call("lbm_nn_upsample2", _p(x), B, H, W, C, _p(up), _p(upb), Ct, off, st)
    u = up.double().cpu().numpy()
    assert np.array_equal(u[..., off:off + C], s.upsample2(xn))
    assert not u[..., :off].any() and not u[..., off + C:].any()
    g = torch.as_tensor(rng.normal(size=(B, 2 * H, 2 * W, Ct)).astype(np.float32)).cuda()
    gx = torch.empty_like(x)
    nat.call("lbm_nn_upsample2_backward", _p(g), B, H, W, C, Ct, off, _p(gx), st)
    np.testing.assert_allclose(gx.double().cpu().numpy(),
                               s.upsample2_backward(g.double().cpu().numpy()[..., off:off + C]),
                               rtol=1e-6, atol=1e-6)
    # copy channels [1, 4) of x into [5, 8) of up, accumulating
    before = up.clone()
    nat.call("lbm_nn_copy_channels", _p(x), C, 1, _p(up), None, Ct, 5, 3, B * H * W, 1, st)
    flat_up = before.view(-1)[:B * H * W * Ct].view(B * H * W, Ct).clone()
    flat_up[:, 5:8] += x.view(B * H * W, C)[:, 1:4]
    assert torch.equal(up.view(-1)[:B * H * W * Ct].view(B * H * W, Ct), flat_up)


def test_relu_bias_mse_adam_kernels(gpu):
    nat = _nat()
    rng = np.random.default_rng(8)
    st = nat.stream_ptr()
    n = 5000
    y = torch.as_tensor(rng.normal(size=n).astype(np.float32)).cuda()
    dy = torch.as_tensor(rng.normal(size=n).astype(np.float32)).cuda()
    dy0 = dy.clone()
    dyb = torch.empty_like(dy, dtype=torch.bfloat16)
    nat.call("lbm_nn_relu_backward", _p(y), _p(dy), _p(dyb), n, st)
    assert torch.equal(dy, torch.where(y > 0, dy0, torch.zeros_like(dy0)))
    assert torch.equal(dyb, dy.to(torch.bfloat16))
    for M, N in ((1000, 5), (250, 20), (3, 300), (70000, 32)):
        g = torch.as_tensor(rng.normal(size=(M, N)).astype(np.float32)).cuda()
        db = torch.empty(N, dtype=torch.float32, device="cuda")
        bw = torch.empty(nat.load().lbm_nn_bias_grad_workspace(N) // 8, dtype=torch.float64,
                         device="cuda")
        nat.call("lbm_nn_bias_grad", _p(g), M, N, _p(db), _p(bw), st)
        ref = g.double().cpu().numpy().sum(0)
        assert np.all(np.abs(db.double().cpu().numpy() - ref) <= 1e-6 * (np.abs(ref) + 1.0)), (M, N)
    t = torch.as_tensor(rng.normal(size=n).astype(np.float32)).cuda()
    work = torch.empty(nat.load().lbm_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
    loss = torch.empty(1, dtype=torch.float64, device="cuda")
    d = torch.empty_like(y)
    nat.call("lbm_nn_mse", _p(y), _p(t), n, _p(d), _p(work), _p(loss), st)
    ref_loss, ref_d = s.mse(y.double().cpu().numpy(), t.double().cpu().numpy())
    # the kernel forms y - t in FP32 and squares and sums in FP64
    d32 = (y - t).double().cpu().numpy()
    assert abs(float(loss.item()) - np.mean(d32 * d32)) <= 1e-12 * ref_loss
    assert abs(float(loss.item()) - ref_loss) <= 1e-6 * ref_loss
    np.testing.assert_allclose(d.double().cpu().numpy(), ref_d, rtol=1e-6)
    # three Adam steps against the published update
    p = torch.as_tensor(rng.normal(size=n).astype(np.float32)).cuda()
    m_, v_ = torch.zeros_like(p), torch.zeros_like(p)
    pb = torch.empty_like(p, dtype=torch.bfloat16)
    pn, mn, vn = p.double().cpu().numpy(), np.zeros(n), np.zeros(n)
    for k in range(1, 4):
        gk = torch.as_tensor(rng.normal(size=n).astype(np.float32)).cuda()
        nat.call("lbm_nn_adam", _p(p), _p(gk), _p(m_), _p(v_), _p(pb), n, 1e-2, 0.9, 0.999, 1e-8, k, st)
        pn, mn, vn = s.adam(pn, gk.double().cpu().numpy(), mn, vn, k, 1e-2)
    np.testing.assert_allclose(p.double().cpu().numpy(), pn, rtol=1e-5, atol=1e-6)
    assert torch.equal(pb, p.to(torch.bfloat16))


WIDTHS, MID = (8, 16), 16


def _small_net(seed=0):
    from latticeboltzmann_amd.warmstart.network import Surrogate
    net = Surrogate(in_ch=6, widths=WIDTHS, mid=MID, out_ch=3, seed=seed)
    # BF16-representable master weights, so the oracle sees the same values
    net.load_flat(net.p.to(torch.bfloat16).float().cpu().numpy())
    rng = np.random.default_rng(seed + 1)
    for name, cin, cout, relu in net.layers:
        b = net.views(name)[1]
        b.copy_(torch.as_tensor(rng.normal(0, 0.05, size=cout).astype(np.float32)))
    net.sync_bf16()
    params = {}
    for name, *_ in net.layers:
        w, b = net.views(name)
        params[name] = (w.double().cpu().numpy(), b.double().cpu().numpy())
    return net, params


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_network_forward_backward_match_oracle(gpu):
    net, params = _small_net(3)
    rng = np.random.default_rng(9)
    x, xn = _bf(rng, (2, 16, 16, 6))
    x = x.float()
    t = torch.as_tensor(rng.normal(size=(2, 16, 16, 3)).astype(np.float32)).cuda()
    out, cache = net.forward(x, keep=True)
    ref, rcache = s.unet_forward(params, xn, WIDTHS, bf=True)
    assert _rel(out.double().cpu().numpy(), ref) < 1e-4
    loss, dout = net.mse(out, t)
    rloss, rdout = s.mse(ref, t.double().cpu().numpy())
    assert abs(float(loss.item()) - rloss) < 1e-4 * rloss
    dx = net.backward(cache, dout, need_input_grad=True)
    grads, rdx = s.unet_backward(params, rcache, rdout, WIDTHS, bf=True)
    for name, *_ in net.layers:
        gw, gb = net.views(name, net.g)
        assert _rel(gw.double().cpu().numpy(), grads[name][0]) < 1e-3, name
        assert _rel(gb.double().cpu().numpy(), grads[name][1]) < 1e-3, name
    assert _rel(dx.double().cpu().numpy(), rdx) < 1e-3
    # and the BF16 network stays close to the unrounded FP64 network
    ref64, _ = s.unet_forward(params, xn, WIDTHS)
    assert _rel(out.double().cpu().numpy(), ref64) < 3e-2


def test_training_reduces_loss_and_is_deterministic(gpu):
    finals = []
    for _ in range(2):
        net, _ = _small_net(4)
        rng = np.random.default_rng(10)
        x = torch.as_tensor(rng.normal(size=(4, 16, 16, 6)).astype(np.float32)).cuda()
        i = np.arange(16)[:, None] / 16.0
        j = np.arange(16)[None, :] / 16.0
        tgt = np.stack([np.sin(3 * i + j), np.cos(2 * j) * i, i * j], -1).astype(np.float32)
        t = torch.as_tensor(np.broadcast_to(tgt, (4, 16, 16, 3)).copy()).cuda()
        losses = [float(net.train_step(x, t, 3e-3).item()) for _ in range(150)]
        assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
        finals.append(net.p.clone())
    assert torch.equal(finals[0], finals[1])


@pytest.mark.parametrize("recon", ["bilinear", "walls", "cubic"])
@pytest.mark.parametrize("neq", [False, True])
def test_warm_init_matches_oracle(gpu, neq, recon):
    from latticeboltzmann_amd.warmstart.solve import warm_populations
    rng = np.random.default_rng(11)
    for m, nx, ny, pool in ((16, 40, 24, 2), (8, 64, 64, 4), (3, 5, 7, 1), (1, 5, 7, 2)):
        if recon == "cubic" and m < 3:
            continue
        U, om = 0.08, 1.3
        q = rng.normal(size=(m, m, 3)).astype(np.float32)
        f = warm_populations(torch.as_tensor(q).cuda(), nx, ny, U, om, neq=neq, recon=recon, pool=pool)
        ref = s.warm_init(q.astype(np.float64), nx, ny, U, om, neq, recon, pool)
        np.testing.assert_allclose(f.cpu().numpy(), ref, rtol=1e-13, atol=1e-15)


def test_warm_init_rejects_bad_arguments(gpu):
    from latticeboltzmann_amd.warmstart.solve import warm_populations
    q = torch.zeros((4, 4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        warm_populations(q, 8, 8, 0.0, 1.0)
    with pytest.raises(ValueError):
        warm_populations(q, 8, 8, 0.1, 2.0)
    with pytest.raises(ValueError):
        warm_populations(q[:, :2], 8, 8, 0.1, 1.0)
    with pytest.raises(ValueError):
        warm_populations(q, 8, 8, 0.1, 1.0, recon="spline")
    nat = _nat()
    lib = nat.load()
    f = torch.empty((9, 8, 8), dtype=torch.float64, device="cuda")
    assert lib.lbm_warm_init_f64(_p(q), 4, 8, 8, 0.1, 1.0, 8, 2, _p(f), nat.stream_ptr()) == -3
    assert lib.lbm_warm_init_f64(_p(q), 4, 8, 8, 0.1, 1.0, 0, 0, _p(f), nat.stream_ptr()) == -3
    assert lib.lbm_warm_init_f64(_p(q), 2, 8, 8, 0.1, 1.0, 4, 2, _p(f), nat.stream_ptr()) == -2


class _Fixed:
    def __init__(self, q):
        self.q = q

    def predict(self, re, u):
        return self.q


def test_controller_on_lattice_warm_from_pooled_steady_state(gpu):
    # a warm start from the steady field itself (averaged onto a 64 x 64 grid)
    # must converge in fewer steps than the cold start, to a field that agrees
    from latticeboltzmann_amd.warmstart.controller import Controller
    from latticeboltzmann_amd.warmstart.features import pooled_target
    from latticeboltzmann_amd.warmstart.solve import CavityStepper
    n, re, u = 128, 100.0, 0.1
    ctl = Controller(tol=1e-6, every=1000)
    with CavityStepper(n, re, u, every=1000) as st:
        cold = ctl.run(st, warm=False)
        fc = tuple(t.clone() for t in st.fields())
    assert cold["reason"] == "converged"
    q = pooled_target(*fc, u, 64)
    with CavityStepper(n, re, u, predictor=_Fixed(q), every=1000) as st:
        warm = ctl.run(st, warm=True)
        fw = st.fields()
    d = torch.sqrt((fw[1] - fc[1]) ** 2 + (fw[2] - fc[2]) ** 2).max().item() / u
    print("cold", cold["steps"], "warm", warm["steps"], "diff", d)
    assert not warm["fell_back"]
    assert warm["steps"] < cold["steps"]
    assert d < 1e-3


def test_controller_on_lattice_falls_back_from_zero_budget(gpu):
    # a warm start given no budget falls back and still converges cold
    from latticeboltzmann_amd.warmstart.controller import Controller
    from latticeboltzmann_amd.warmstart.solve import CavityStepper
    n, re, u = 128, 100.0, 0.1
    q = torch.zeros((64, 64, 3), dtype=torch.float32, device="cuda")
    ctl = Controller(tol=1e-6, every=1000, budget=0.0)
    with CavityStepper(n, re, u, predictor=_Fixed(q), every=1000) as st:
        r = ctl.run(st, warm=True, cold_estimate=1000)
    assert r["fell_back"] and r["reason"] == "over budget" and r["warm_steps"] == 1000
    assert r["cold_steps"] > 0 and r["residual"] < 1e-6


def test_monitor_flags_diverging_reference_cavity(gpu):
    # The reference's lid rule leaks mass at the lid corners (per step
    # 6 w u0 (rho_lid(nx-1) - rho_lid(0))); on a 64 x 64 cavity at Re 100,
    # U = 0.1 the mass grows ~0.3 % per 10^4 steps and the run turns unstable
    # near 1.5e5 steps (the CPU oracle does the same).  The residual monitor
    # must report it rather than a converged state.
    from latticeboltzmann_amd.lattice import Lattice, omega_for_cavity
    lat = Lattice(64, 64, omega_for_cavity(64, 100.0, 0.1), "cavity", 0.1, torch.float64)
    with pytest.raises(FloatingPointError):
        lat.run_until_steady(tol=1e-6, every=500, max_steps=300_000)
    assert torch.isnan(lat.f).any()


def test_residual_of_injected_nan_is_infinite(gpu):
    from latticeboltzmann_amd import kernels
    from latticeboltzmann_amd.lattice import Lattice
    lat = Lattice(32, 48, 1.2, "cavity", 0.1, torch.float64).step(10)
    ux = torch.zeros((32, 48), dtype=torch.float64, device="cuda")
    uy = torch.zeros_like(ux)
    assert np.isfinite(kernels.velocity_residual(lat.f, ux, uy, layout=lat.layout).item())
    lat.f.view(-1)[12345] = float("nan")
    assert kernels.velocity_residual(lat.f, ux, uy, layout=lat.layout).item() == float("inf")


def test_fold_and_unfold_output_standardization(gpu):
    # the saved network predicts std * y + mean through its head; unfolding
    # recovers the trained weights
    from latticeboltzmann_amd.warmstart.network import Surrogate
    from latticeboltzmann_amd.warmstart.train import folded_state, unfold_state
    net, _ = _small_net(6)
    mean, std = np.array([0.5, -0.1, 0.02]), np.array([2.0, 0.3, 0.1])
    x = torch.as_tensor(np.random.default_rng(12).normal(size=(2, 16, 16, 6)).astype(np.float32)).cuda()
    raw = net.forward(x).double()
    folded = Surrogate.from_state(folded_state(net, mean, std))
    out = folded.forward(x).double()
    expect = raw * torch.as_tensor(std, device="cuda") + torch.as_tensor(mean, device="cuda")
    # the folded head's BF16 copy rounds differently from the unfolded one
    assert float((out - expect).norm() / expect.norm()) < 1e-2
    back = Surrogate(6, WIDTHS, MID, 3)
    unfold_state(back, folded_state(net, mean, std), mean, std)
    assert torch.allclose(back.p, net.p, rtol=1e-6, atol=1e-7)


def test_pooled_error_term_matches_oracle(gpu):
    nat = _nat()
    rng = np.random.default_rng(14)
    B, H, W, C, f, wgt = 2, 16, 24, 3, 8, 3.0
    y = torch.as_tensor(rng.normal(size=(B, H, W, C)).astype(np.float32)).cuda()
    t = torch.as_tensor(rng.normal(size=(B, H, W, C)).astype(np.float32)).cuda()
    dy0 = torch.as_tensor(rng.normal(size=(B, H, W, C)).astype(np.float32)).cuda()
    dy = dy0.clone()
    work = torch.empty(nat.load().lbm_reduce_workspace_bytes() // 8, dtype=torch.float64, device="cuda")
    loss = torch.empty(1, dtype=torch.float64, device="cuda")
    nat.call("lbm_nn_mse_pooled", _p(y), _p(t), B, H, W, C, f, wgt, _p(dy), _p(work), _p(loss),
             nat.stream_ptr())
    rl, rg = s.mse_pooled(y.double().cpu().numpy(), t.double().cpu().numpy(), f, wgt)
    assert abs(float(loss.item()) - rl) <= 1e-12 * rl
    # dy0 + g is rounded once to FP32: allow one rounding of the sum
    want = dy0.double().cpu().numpy() + rg
    got = dy.double().cpu().numpy()
    assert np.all(np.abs(got - want) <= 2.0 ** -23 * np.abs(want) + 1e-6 * np.abs(rg) + 1e-12)
    lib = nat.load()
    assert lib.lbm_nn_mse_pooled(_p(y), _p(t), B, H, W, C, 5, wgt, _p(dy), _p(work), _p(loss),
                                 nat.stream_ptr()) == -2


def test_parameter_average_kernel(gpu):
    nat = _nat()
    rng = np.random.default_rng(15)
    p = torch.as_tensor(rng.normal(size=1000).astype(np.float32)).cuda()
    ema = torch.as_tensor(rng.normal(size=1000).astype(np.float32)).cuda()
    e0 = ema.clone()
    nat.call("lbm_nn_ema", _p(ema), _p(p), 1000, 0.99, nat.stream_ptr())
    want = 0.99 * e0.double() + 0.01 * p.double()
    assert float((ema.double() - want).abs().max()) < 1e-6
    assert nat.load().lbm_nn_ema(_p(ema), _p(p), 1000, 1.0, nat.stream_ptr()) == -3
