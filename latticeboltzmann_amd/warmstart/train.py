"""Train the warm-start surrogate on the GPU with the project's own kernels.

Adam with a linear warm-up and cosine decay of the learning rate, MSE on the
targets standardized per channel (the density channel (rho - 1)/U^2 peaks at
the lid corners and would otherwise dominate the velocity channels),
mini-batches drawn without replacement each epoch; a held-out fraction of the
samples gives the validation loss.  The saved network predicts the physical
normalized fields directly: the standardization is folded into the linear
head (rows of its weights times the channel's scale, bias times the scale plus
the channel's mean).

Usage::

    python -m latticeboltzmann_amd.warmstart.train data1.npz [data2.npz ...] \
        --epochs 300 --batch 16 --lr 2e-3 --out surrogate_weights.npz
"""

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .features import RE_RANGE, U_RANGE, in_channels
from .network import Surrogate
from .solve import ANY_N, Predictor


def load_dataset(paths, mean_free_density=True):
    """(params [(Re, U)], targets (S, m, m, 3) float32).

    With ``mean_free_density`` each sample's density channel loses its
    spatial mean.  Scaling every population by one factor maps a steady
    cavity of this scheme to another (BGK, bounce-back and the lid term
    6 w rho_lid u0 are all homogeneous of degree one in f), so the mean density
    carries no information about the flow; here it only records how much
    mass the lid corners leaked before the run converged."""
    params, targets = [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            params += list(zip(z["re"].tolist(), z["u"].tolist()))
            targets.append(z["targets"].astype(np.float32))
    t = np.concatenate(targets)
    if mean_free_density:
        t[..., 0] -= t[..., 0].mean(axis=(1, 2), keepdims=True)
    return params, t


def lr_at(step, total, lr, warmup=100, floor=0.02):
    if step < warmup:
        return lr * (step + 1) / warmup
    t = (step - warmup) / max(1, total - warmup)
    return lr * (floor + (1 - floor) * 0.5 * (1 + math.cos(math.pi * min(t, 1.0))))


def folded_state(net, mean, std):
    """State of ``net`` with y = std * y_standardized + mean folded into the head."""
    state = net.state_dict()
    flat = state["flat"].copy()
    a, b = net.offsets["head"]
    cin, cout, _ = net.spec["head"]
    flat[a:b] = (flat[a:b].reshape(cout, 9 * cin) * std[:, None]).ravel()
    flat[b:b + cout] = flat[b:b + cout] * std + mean
    state["flat"] = flat.astype(np.float32)
    return state


def unfold_state(net, state, mean, std):
    """Inverse of ``folded_state``: load a saved predictor's weights into ``net``
    as the network of standardized targets (to continue training it)."""
    flat = np.asarray(state["flat"], dtype=np.float64).copy()
    if flat.size != net.nparams:
        raise ValueError("saved network has %d parameters, this one %d" % (flat.size, net.nparams))
    a, b = net.offsets["head"]
    cin, cout, _ = net.spec["head"]
    flat[a:b] = (flat[a:b].reshape(cout, 9 * cin) / std[:, None]).ravel()
    flat[b:b + cout] = (flat[b:b + cout] - mean) / std
    net.load_flat(flat.astype(np.float32))


def channel_errors(pred, X, Y, batch):
    """RMS error per output channel of a (folded) predictor, in the targets' units."""
    se = torch.zeros(Y.shape[-1], dtype=torch.float64, device=X.device)
    for a in range(0, X.shape[0], batch):
        d = (pred.net.forward(X[a:a + batch]) - Y[a:a + batch]).double()
        se += (d * d).sum(dim=(0, 1, 2))
    return (se / (Y.numel() // Y.shape[-1])).sqrt().cpu().numpy()


def evaluate_loss(net, X, Y, batch):
    tot = torch.zeros(1, dtype=torch.float64, device=X.device)
    for a in range(0, X.shape[0], batch):
        out = net.forward(X[a:a + batch])
        loss, _ = net.mse(out, Y[a:a + batch])
        tot += loss * out.shape[0]
    return float(tot.item()) / X.shape[0]


def large_scale_error(pred, X, Y, batch, block=32):
    """RMS over validation samples of the velocity error averaged over
    block x block cells (units of U): the part of the error that only the
    slowest modes of the lattice remove."""
    se, n = 0.0, 0
    for a in range(0, X.shape[0], batch):
        d = (pred.net.forward(X[a:a + batch]) - Y[a:a + batch])[..., 1:].double()
        B, H, W, C = d.shape
        e = d.reshape(B, H // block, block, W // block, block, C).mean(dim=(2, 4))
        se += float((e * e).sum().item())
        n += e.numel()
    return math.sqrt(se / n)


def train(paths, epochs=300, batch=16, lr=2e-3, seed=0, val_frac=0.1, out=None, widths=(32, 64, 128),
          mid=128, time_limit=None, log=None, octaves=0, init=None, scales=(), ema=0.0,
          re_range=RE_RANGE, u_range=U_RANGE, n_range=None):
    """Train from scratch, or continue from the predictor saved at ``init``
    (same architecture and data split; a network trained for other input
    ranges is a starting point only).  ``re_range``/``u_range`` normalize the
    inputs and are the ranges the saved predictor accepts; ``n_range`` (lo,
    hi) limits the lattice sizes it is used for (default: any).

    With ``ema`` > 0 the saved network is the exponential moving average of
    the weights (decay ``ema`` per step).  At every log point (10 epochs) the
    saved candidate is scored on the validation split by its large-scale
    velocity error, and the best one is kept."""
    params, targets = load_dataset(paths)
    S = len(params)
    lo_re, hi_re = min(p[0] for p in params), max(p[0] for p in params)
    lo_u, hi_u = min(p[1] for p in params), max(p[1] for p in params)
    if lo_re < re_range[0] or hi_re > re_range[1] or lo_u < u_range[0] or hi_u > u_range[1]:
        raise ValueError("training data spans Re %.4g..%.4g, U %.4g..%.4g: outside the ranges %s, %s"
                         % (lo_re, hi_re, lo_u, hi_u, re_range, u_range))
    pred = Predictor(Surrogate(in_ch=in_channels(octaves), widths=widths, mid=mid, seed=seed),
                     re_range=re_range, u_range=u_range, octaves=octaves,
                     n_range=n_range if n_range else ANY_N)
    net = pred.net
    X = pred.inputs(params)
    rng = np.random.default_rng(seed)
    order = rng.permutation(S)
    nval = int(round(val_frac * S))
    val, tr = order[:nval], order[nval:]
    mean = targets[tr].mean(axis=(0, 1, 2)).astype(np.float64)
    std = targets[tr].std(axis=(0, 1, 2)).astype(np.float64)
    Yphys = torch.as_tensor(targets).to(net.device)
    Y = torch.as_tensor(((targets - mean) / std).astype(np.float32)).to(net.device)
    Xv, Yv = X[torch.as_tensor(val)], Y[torch.as_tensor(val)]
    Yv_phys = Yphys[torch.as_tensor(val)]
    if init:
        with np.load(init, allow_pickle=False) as z:
            if int(z["octaves"]) != octaves:
                raise ValueError("--init network was trained with %d octaves" % int(z["octaves"]))
            unfold_state(net, {"flat": z["flat"]}, mean, std)

    def candidate():
        """The folded predictor of the current (or averaged) weights."""
        state = folded_state(net, mean, std)
        if ema > 0 and net.ema is not None:
            avg = net.state_dict()
            avg["flat"] = net.ema.cpu().numpy()
            state = folded_state(Surrogate.from_state(avg, device=net.device), mean, std)
        return Predictor.from_state(state, pred.grid, pred.re_range, pred.u_range,
                                    device=net.device, octaves=octaves, n_range=pred.n_range)

    steps_per_epoch = max(1, len(tr) // batch)
    total = epochs * steps_per_epoch
    step, t0, hist = 0, time.time(), []
    best, best_err = None, math.inf
    for ep in range(epochs):
        perm = rng.permutation(tr)
        acc = torch.zeros(1, dtype=torch.float64, device=net.device)
        for b in range(steps_per_epoch):
            idx = torch.as_tensor(perm[b * batch:(b + 1) * batch])
            acc += net.train_step(X[idx], Y[idx], lr_at(step, total, lr), scales)
            if ema > 0:
                net.ema_update(ema)
            step += 1
        last = ep == epochs - 1 or (time_limit and time.time() - t0 > time_limit)
        if ep % 10 == 9 or last:
            rec = {"epoch": ep + 1, "train_loss": float(acc.item()) / steps_per_epoch,
                   "val_loss": evaluate_loss(net, Xv, Yv, batch) if nval else None,
                   "seconds": round(time.time() - t0, 1)}
            cand = candidate()
            if nval:
                err = large_scale_error(cand, Xv, Yv_phys, batch)
                rec["val_large_scale"] = err
                if err < best_err:
                    best, best_err = cand, err
                    rec["kept"] = True
            else:
                best = cand
            hist.append(rec)
            if log:
                log(json.dumps(rec))
            if out and rec.get("kept", not nval):
                best.save(out)
        if last:
            break
    final = best if best is not None else candidate()
    if out:
        final.save(out)
    if nval:
        rms = channel_errors(final, Xv, Yv_phys, batch)
        rec = {"val_rms_rho": float(rms[0]), "val_rms_ux": float(rms[1]), "val_rms_uy": float(rms[2]),
               "val_large_scale": large_scale_error(final, Xv, Yv_phys, batch)}
        hist.append(rec)
        if log:
            log(json.dumps(rec))
    return final, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("data", nargs="+")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--val-frac", type=float, default=0.1)
    ap.add_argument("--octaves", type=int, default=0, help="sin/cos octaves of the scalar inputs")
    ap.add_argument("--init", default=None, help="continue training from this saved predictor")
    ap.add_argument("--scale-loss", action="append", default=[], metavar="F:W",
                    help="add W times the MSE of the error averaged over F x F cells")
    ap.add_argument("--ema", type=float, default=0.0,
                    help="keep an exponential moving average of the weights with this decay per step")
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--re-range", type=float, nargs=2, default=RE_RANGE)
    ap.add_argument("--u-range", type=float, nargs=2, default=U_RANGE)
    ap.add_argument("--n-range", type=int, nargs=2, default=None,
                    help="lattice sizes the saved network is meant for (default: any)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    train(a.data, a.epochs, a.batch, a.lr, a.seed, a.val_frac, a.out, time_limit=a.time_limit,
          log=lambda s: print(s, flush=True), octaves=a.octaves, init=a.init,
          scales=[(int(v.split(":")[0]), float(v.split(":")[1])) for v in a.scale_loss], ema=a.ema,
          re_range=tuple(a.re_range), u_range=tuple(a.u_range),
          n_range=tuple(a.n_range) if a.n_range else None)


if __name__ == "__main__":
    sys.exit(main())
