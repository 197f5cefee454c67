"""Convolutional encoder-decoder of the warm start, on the project's HIP kernels.

Every arithmetic step runs in ``liblbm_d2q9.so`` (``lbm_nn_*``): 3x3 same
convolutions as BF16 MFMA GEMMs with FP32 accumulation whose tile loader
gathers the shifted activations itself (no im2col matrix in memory: forward,
weight gradient and data gradient), a fused bias/ReLU epilogue, 2x2 average
pooling, nearest 2x upsampling, channel concatenation, their backward passes,
the MSE loss and Adam.  torch provides
device memory only (no autograd, no torch arithmetic on the hot path).

Layout: activations NHWC (FP32, plus a BF16 copy that feeds the next GEMM);
weights of a layer ``W[cout][(dy*3 + dx)*cin + c]`` in one flat FP32 buffer
(Adam's master copy) with a flat BF16 copy.  The architecture is a U-Net:
per level two ReLU convolutions and a 2x2 average pool on the way down, two
ReLU convolutions in the middle, and per level a nearest upsample,
concatenation with the level's skip, and two ReLU convolutions on the way up,
then a linear 3x3 head.
"""

import math

import numpy as np
import torch

from .. import _native as nat


def unet_layers(in_ch, widths, mid, out_ch):
    """Layer table [(name, cin, cout, relu)], encoder first, head last."""
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


def _p(t):
    return t.data_ptr()


class Surrogate:
    """The encoder-decoder with its parameters, gradients and Adam state on one GPU."""

    def __init__(self, in_ch=6, widths=(32, 64, 128), mid=128, out_ch=3, seed=0, device=None):
        nat.require_gpu()
        self.in_ch, self.widths, self.mid, self.out_ch = in_ch, tuple(widths), mid, out_ch
        self.layers = unet_layers(in_ch, self.widths, mid, out_ch)
        self.spec = {name: (cin, cout, relu) for name, cin, cout, relu in self.layers}
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.offsets, n = {}, 0
        for name, cin, cout, _ in self.layers:
            self.offsets[name] = (n, n + 9 * cin * cout)
            n += 9 * cin * cout + cout
        self.nparams = n
        f32 = dict(dtype=torch.float32, device=self.device)
        self.p = torch.empty(n, **f32)
        self.pb = torch.empty(n, dtype=torch.bfloat16, device=self.device)
        self.g = torch.zeros(n, **f32)
        self.m = torch.zeros(n, **f32)
        self.v = torch.zeros(n, **f32)
        self.step_count = 0
        self.ema = None
        self.work = torch.empty(int(nat.load().lbm_reduce_workspace_bytes()) // 8,
                                dtype=torch.float64, device=self.device)
        self.init_params(seed)

    # -- parameters -------------------------------------------------------------
    def init_params(self, seed=0):
        """He-normal weights (variance 2 / fan_in; 1 / fan_in for the head), zero bias."""
        rng = np.random.default_rng(seed)
        flat = np.zeros(self.nparams, dtype=np.float32)
        for name, cin, cout, relu in self.layers:
            a, b = self.offsets[name]
            std = math.sqrt((2.0 if relu else 1.0) / (9 * cin))
            flat[a:b] = rng.normal(0.0, std, size=b - a)
        self.load_flat(flat)

    def load_flat(self, flat):
        self.p.copy_(torch.as_tensor(np.asarray(flat, dtype=np.float32)))
        self.sync_bf16()
        self.m.zero_()
        self.v.zero_()
        self.step_count = 0

    def sync_bf16(self):
        nat.call("lbm_nn_to_bf16", _p(self.p), _p(self.pb), self.nparams, nat.stream_ptr())

    def views(self, name, buf=None):
        """(weight (cout, 9 cin), bias (cout,)) views of ``buf`` (default the parameters)."""
        buf = self.p if buf is None else buf
        cin, cout, _ = self.spec[name]
        a, b = self.offsets[name]
        return buf[a:b].view(cout, 9 * cin), buf[b:b + cout]

    def state_dict(self):
        return {"flat": self.p.cpu().numpy(), "in_ch": self.in_ch, "widths": np.array(self.widths),
                "mid": self.mid, "out_ch": self.out_ch}

    @classmethod
    def from_state(cls, state, device=None):
        net = cls(int(state["in_ch"]), tuple(int(w) for w in state["widths"]), int(state["mid"]),
                  int(state["out_ch"]), device=device)
        net.load_flat(state["flat"])
        return net

    # -- building blocks ----------------------------------------------------------
    def _conv(self, name, x, xb, relu=None):
        B, H, W, cin = x.shape
        cin_, cout, relu_ = self.spec[name]
        assert cin == cin_, (name, cin, cin_)
        relu = relu_ if relu is None else relu
        y = torch.empty((B, H, W, cout), dtype=torch.float32, device=self.device)
        yb = torch.empty((B, H, W, cout), dtype=torch.bfloat16, device=self.device)
        wb, bias = self.views(name, self.pb)[0], self.views(name)[1]
        nat.call("lbm_nn_conv3x3", _p(xb), B, H, W, cin, _p(wb), cout, _p(bias), _p(y), _p(yb),
                 int(relu), nat.stream_ptr())
        return y, yb

    def _pool(self, x):
        B, H, W, C = x.shape
        y = torch.empty((B, H // 2, W // 2, C), dtype=torch.float32, device=self.device)
        yb = torch.empty_like(y, dtype=torch.bfloat16)
        nat.call("lbm_nn_avgpool2", _p(x), B, H, W, C, _p(y), _p(yb), nat.stream_ptr())
        return y, yb

    def _concat_up(self, h, skip):
        B, H, W, Ch = h.shape
        Cs = skip.shape[3]
        cat = torch.empty((B, 2 * H, 2 * W, Ch + Cs), dtype=torch.float32, device=self.device)
        catb = torch.empty_like(cat, dtype=torch.bfloat16)
        s = nat.stream_ptr()
        nat.call("lbm_nn_upsample2", _p(h), B, H, W, Ch, _p(cat), _p(catb), Ch + Cs, 0, s)
        nat.call("lbm_nn_copy_channels", _p(skip), Cs, 0, _p(cat), _p(catb), Ch + Cs, Ch, Cs,
                 B * 4 * H * W, 0, s)
        return cat, catb

    # -- forward / backward ---------------------------------------------------------
    def forward(self, x, keep=False):
        """x (B, H, W, in_ch) FP32 -> output (B, H, W, out_ch) FP32; H, W divisible
        by 2^levels.  ``keep`` returns the activations needed by ``backward``."""
        x = x.contiguous().to(torch.float32)
        B, H, W, C = x.shape
        lv = len(self.widths)
        if C != self.in_ch or H % (1 << lv) or W % (1 << lv):
            raise ValueError("input must be (B, H, W, %d) with H, W divisible by %d"
                             % (self.in_ch, 1 << lv))
        xb = torch.empty_like(x, dtype=torch.bfloat16)
        nat.call("lbm_nn_to_bf16", _p(x), _p(xb), x.numel(), nat.stream_ptr())
        cache = {}

        def conv(name, h, hb):
            y, yb = self._conv(name, h, hb)
            if keep:
                cache[name] = (hb, y, tuple(h.shape))
            return y, yb

        h, hb, skips = x, xb, []
        for i in range(lv):
            h, hb = conv("enc%da" % i, h, hb)
            h, hb = conv("enc%db" % i, h, hb)
            skips.append(h)
            h, hb = self._pool(h)
        h, hb = conv("mida", h, hb)
        h, hb = conv("midb", h, hb)
        for i in reversed(range(lv)):
            h, hb = self._concat_up(h, skips[i])
            h, hb = conv("dec%da" % i, h, hb)
            h, hb = conv("dec%db" % i, h, hb)
        out, _ = conv("head", h, hb)
        return (out, cache) if keep else out

    def _ksplit(self, cout, K, P):
        base = -(-K // 64) * -(-cout // 64)
        return max(1, min(64, -(-1024 // base), P // 512))

    def _back(self, cache, name, dy, need_dx=True):
        xb, y, (B, H, W, cin) = cache[name]
        _, cout, relu = self.spec[name]
        P, K = B * H * W, 9 * cin
        s = nat.stream_ptr()
        dy = dy.contiguous()
        dyb = torch.empty_like(dy, dtype=torch.bfloat16)
        if relu:
            nat.call("lbm_nn_relu_backward", _p(y), _p(dy), _p(dyb), dy.numel(), s)
        else:
            nat.call("lbm_nn_to_bf16", _p(dy), _p(dyb), dy.numel(), s)
        gw, gbias = self.views(name, self.g)
        ks = self._ksplit(cout, K, P)
        work = None
        if ks > 1:
            work = torch.empty(int(nat.load().lbm_nn_gemm_workspace(cout, K, ks)) // 4,
                               dtype=torch.float32, device=self.device)
        # dW[co][k] = sum_p dy[p][co] x(p, k), x(p, k) the gathered 3x3 neighbourhood
        nat.call("lbm_nn_conv3x3_wgrad", _p(xb), _p(dyb), B, H, W, cin, cout, _p(gw), ks,
                 None if work is None else _p(work), s)
        bw = torch.empty(int(nat.load().lbm_nn_bias_grad_workspace(cout)) // 8, dtype=torch.float64,
                         device=self.device)
        nat.call("lbm_nn_bias_grad", _p(dy), P, cout, _p(gbias), _p(bw), s)
        if not need_dx:
            return None
        # dx[p][c] = sum over taps and co of dy[p - tap][co] W[co][tap][c]
        wb = self.views(name, self.pb)[0]
        wf = torch.empty(9 * cin * cout, dtype=torch.bfloat16, device=self.device)
        nat.call("lbm_nn_conv3x3_flip_weights", _p(wb), cin, cout, _p(wf), s)
        dx = torch.empty((B, H, W, cin), dtype=torch.float32, device=self.device)
        nat.call("lbm_nn_conv3x3_dgrad", _p(dyb), _p(wf), B, H, W, cin, cout, _p(dx), 0, s)
        return dx

    def backward(self, cache, dout, need_input_grad=False):
        """Fill the gradient buffer from dL/d(output); optionally return dL/d(input)."""
        s = nat.stream_ptr()
        lv = len(self.widths)
        g = self._back(cache, "head", dout)
        dskips = [None] * lv
        for i in range(lv):
            g = self._back(cache, "dec%db" % i, g)
            g = self._back(cache, "dec%da" % i, g)
            B, H, W, Ct = g.shape
            Cs = self.widths[i]
            Ch = Ct - Cs
            dskips[i] = g
            gu = torch.empty((B, H // 2, W // 2, Ch), dtype=torch.float32, device=self.device)
            nat.call("lbm_nn_upsample2_backward", _p(g), B, H // 2, W // 2, Ch, Ct, 0, _p(gu), s)
            g = gu
        g = self._back(cache, "midb", g)
        g = self._back(cache, "mida", g)
        for i in reversed(range(lv)):
            B, h2, w2, C = g.shape
            gs = torch.empty((B, 2 * h2, 2 * w2, C), dtype=torch.float32, device=self.device)
            nat.call("lbm_nn_avgpool2_backward", _p(g), B, 2 * h2, 2 * w2, C, _p(gs), 0, s)
            cat = dskips[i]
            Ct = cat.shape[3]
            nat.call("lbm_nn_copy_channels", _p(cat), Ct, Ct - C, _p(gs), None, C, 0, C,
                     B * 4 * h2 * w2, 1, s)
            g = self._back(cache, "enc%db" % i, gs)
            g = self._back(cache, "enc%da" % i, g, need_dx=need_input_grad or i > 0)
        return g

    def mse(self, out, target):
        """(loss as a device FP64 scalar, dL/d(out))."""
        dy = torch.empty_like(out)
        loss = torch.empty(1, dtype=torch.float64, device=self.device)
        nat.call("lbm_nn_mse", _p(out), _p(target.contiguous()), out.numel(), _p(dy), _p(self.work),
                 _p(loss), nat.stream_ptr())
        return loss, dy

    def loss(self, out, target, scales=()):
        """MSE plus, for each (f, weight) in ``scales``, weight times the MSE of
        the error averaged over f x f cell blocks (large-scale error).  Returns
        (loss as a device FP64 scalar, dL/d(out))."""
        loss, dy = self.mse(out, target)
        B, H, W, C = out.shape
        for f, wgt in scales:
            extra = torch.empty(1, dtype=torch.float64, device=self.device)
            nat.call("lbm_nn_mse_pooled", _p(out), _p(target.contiguous()), B, H, W, C, int(f),
                     float(wgt), _p(dy), _p(self.work), _p(extra), nat.stream_ptr())
            loss = loss + extra
        return loss, dy

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        nat.call("lbm_nn_adam", _p(self.p), _p(self.g), _p(self.m), _p(self.v), _p(self.pb),
                 self.nparams, float(lr), b1, b2, eps, self.step_count, nat.stream_ptr())

    def ema_update(self, decay):
        """Exponential moving average of the FP32 parameters (started at the
        current parameters on the first call)."""
        if self.ema is None:
            self.ema = self.p.clone()
        nat.call("lbm_nn_ema", _p(self.ema), _p(self.p), self.nparams, float(decay), nat.stream_ptr())

    def train_step(self, x, target, lr, scales=()):
        out, cache = self.forward(x, keep=True)
        loss, dout = self.loss(out, target, scales)
        self.backward(cache, dout)
        self.adam(lr)
        return loss


def _bench(argv=None):
    """Time forward and training steps on synthetic 128 x 128 inputs:
    ``python -m latticeboltzmann_amd.warmstart.network [--batch 16] [--steps 10]``."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description=_bench.__doc__.split("\n")[0])
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--in-ch", type=int, default=18)
    a = ap.parse_args(argv)
    net = Surrogate(in_ch=a.in_ch, seed=0)
    g = torch.Generator(device=net.device).manual_seed(0)
    x = torch.randn(a.batch, 128, 128, a.in_ch, device=net.device, generator=g)
    t = torch.randn(a.batch, 128, 128, 3, device=net.device, generator=g)
    flops = 0
    for name, cin, cout, _ in net.layers:
        lvl = int(name[3]) if name[:3] in ("enc", "dec") else (0 if name == "head" else len(net.widths))
        flops += 2 * 9 * cin * cout * (128 >> lvl) ** 2 * a.batch
    for _ in range(3):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.forward(x)
    torch.cuda.synchronize()
    tf = (time.perf_counter() - t0) / a.steps
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    tt = (time.perf_counter() - t0) / a.steps
    print(json.dumps({"params": net.nparams, "batch": a.batch, "forward_ms": round(tf * 1e3, 3),
                      "train_step_ms": round(tt * 1e3, 3), "forward_tflops": round(flops / tf / 1e12, 2),
                      "train_tflops": round(3 * flops / tt / 1e12, 2)}), flush=True)


if __name__ == "__main__":
    _bench()
