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

from .network import Surrogate
from .solve import Predictor


def load_dataset(paths):
    params, targets = [], []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            params += list(zip(z["re"].tolist(), z["u"].tolist()))
            targets.append(z["targets"].astype(np.float32))
    return params, np.concatenate(targets)


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


def train(paths, epochs=300, batch=16, lr=2e-3, seed=0, val_frac=0.1, out=None, widths=(32, 64, 128),
          mid=128, time_limit=None, log=None):
    params, targets = load_dataset(paths)
    S = len(params)
    pred = Predictor(Surrogate(widths=widths, mid=mid, seed=seed))
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

    def save():
        if out:
            Predictor.from_state(folded_state(net, mean, std), pred.grid, pred.re_range,
                                 pred.u_range).save(out)
    steps_per_epoch = max(1, len(tr) // batch)
    total = epochs * steps_per_epoch
    step, t0, hist = 0, time.time(), []
    for ep in range(epochs):
        perm = rng.permutation(tr)
        acc = torch.zeros(1, dtype=torch.float64, device=net.device)
        for b in range(steps_per_epoch):
            idx = torch.as_tensor(perm[b * batch:(b + 1) * batch])
            acc += net.train_step(X[idx], Y[idx], lr_at(step, total, lr))
            step += 1
        if ep % 10 == 9 or ep == epochs - 1 or (time_limit and time.time() - t0 > time_limit):
            rec = {"epoch": ep + 1, "train_loss": float(acc.item()) / steps_per_epoch,
                   "val_loss": evaluate_loss(net, Xv, Yv, batch) if nval else None,
                   "seconds": round(time.time() - t0, 1)}
            hist.append(rec)
            if log:
                log(json.dumps(rec))
            save()
        if time_limit and time.time() - t0 > time_limit:
            break
    save()
    final = Predictor.from_state(folded_state(net, mean, std), pred.grid, pred.re_range,
                                 pred.u_range, device=net.device)
    if nval:
        rms = channel_errors(final, Xv, Yphys[torch.as_tensor(val)], batch)
        rec = {"val_rms_rho": float(rms[0]), "val_rms_ux": float(rms[1]), "val_rms_uy": float(rms[2])}
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
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    train(a.data, a.epochs, a.batch, a.lr, a.seed, a.val_frac, a.out, time_limit=a.time_limit,
          log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    sys.exit(main())
