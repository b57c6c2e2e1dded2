"""Trains the warm-start network on pairs that the solver itself generates.

    python -m primaldualtoolbox.train_warmstart --train 2000 --held-out 200 \
        --epochs 20 --out warmstart.npz --log train_log.json

Pairs: 320 x 320, 15-coil slices like the benchmark workload (training seeds
1000.., held-out seeds 3000..), each solved to relative gap 1e-4 by the
batched GPU solver; x = A^*f / s and t = ((u - A^*f) / s, v / s) with
s = RMS(A^*f).  Training: Adam, lr 1e-3, batch 16, on the GPU through the
network's own forward/backward kernels.  Evaluation: loss on the held-out
pairs, and iterations to the same gap from the predicted (u, v) versus the
cold start u = A^*f, v = 0 on the held-out slices.
"""
from __future__ import annotations

import argparse
import ctypes as ct
import json
import sys
import time

import numpy as np

from . import _capi as C
from . import pymrireconstruction as mri
from . import synthetic as syn
from . import warmstart as ws

GAP_TOL, GAP_EVERY = 1e-4, 50
SOLVE_STREAMS = 4  # slices solved at once (pdt_solve_concurrent)


def log(*a):
    print(*a, flush=True)


def build_solver(case, max_iter: int = 200000):
    op = mri.MriCartesianRemoveROOSOperator()
    op.setCoilSens(case["sens"])
    op.setMask(case["mask"])
    opt = mri.TgvMriOptimizer_2c3c()
    opt.setOperator(op)
    opt.setNoisyData(case["kspace"])
    a1 = 1e-5 * float(np.linalg.norm(case["kspace"]))
    p = opt.getParameters()
    p.alpha1, p.alpha0, p.reduction, p.max_iter = a1, 2 * a1, 1.0, max_iter
    p.gap_tol, p.gap_every = GAP_TOL, GAP_EVERY
    return op, opt


def solve_pairs(cases):
    """Solves the slices (several at a time, each as its own solve) and
    returns (x, t, iterations)."""
    built = [build_solver(c) for c in cases]
    mri.solve_concurrent([o for _, o in built], streams=SOLVE_STREAMS)
    xs, ts, its = [], [], []
    for (op, opt), c in zip(built, cases):
        st = opt.getStats()
        if not st["converged"]:
            raise RuntimeError(f"a training slice did not reach the gap: {st}")
        x, t = ws.training_pair(op.adjoint(c["kspace"]), opt.getResult(), opt.getV())
        xs.append(x)
        ts.append(t)
        its.append(int(st["iters"]))
    return np.stack(xs), np.stack(ts), its


def make_pairs(seeds, workers: int, chunk: int, tag: str):
    """Generates and solves the slices chunk by chunk; the next chunk's
    slices are generated on the CPU while the GPU solves the current one."""
    seeds = list(seeds)
    X, T, iters = [], [], []
    t0 = time.time()
    with syn.process_pool(workers) as ex:
        pending = [ex.submit(syn.mri320_case, s) for s in seeds[:chunk]]
        for k in range(0, len(seeds), chunk):
            cases = [f.result() for f in pending]
            pending = [ex.submit(syn.mri320_case, s) for s in seeds[k + chunk:k + 2 * chunk]]
            x, t, it = solve_pairs(cases)
            del cases
            X.append(x)
            T.append(t)
            iters += it
            log(f"[train] {tag}: {k + len(it)}/{len(seeds)} pairs, median {int(np.median(it))} iterations, "
                f"{time.time() - t0:.0f} s")
    return np.concatenate(X), np.concatenate(T), iters


class Batches:
    """Device buffers of one batch, refilled from host arrays."""

    def __init__(self, batch: int, h: int, w: int):
        self.shape = (batch, h, w)
        self.x = C.DeviceArray.empty(C.PDT_F32, [batch * h * w * ws.CIN])
        self.t = C.DeviceArray.empty(C.PDT_F32, [batch * h * w * ws.COUT])

    def fill(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        C.check(C.lib().pdt_array_upload(self.x.h, x.ctypes.data_as(ct.c_void_p), x.nbytes))
        C.check(C.lib().pdt_array_upload(self.t.h, t.ctypes.data_as(ct.c_void_p), t.nbytes))


def mean_loss(net, X, T, bufs: Batches) -> float:
    n = bufs.shape[0]
    tot, cnt = 0.0, 0
    for k in range(0, len(X) - n + 1, n):
        bufs.fill(X[k:k + n], T[k:k + n])
        loss = ct.c_double()
        C.check(C.lib().pdt_cnn_loss_grad(net._h, bufs.x.h, bufs.t.h, *bufs.shape, ct.byref(loss), None, 0))
        tot += loss.value
        cnt += 1
    return tot / max(cnt, 1)


def train(net, X, T, Xv, Tv, epochs: int, batch: int, lr: float, seed: int, lr_final: float | None = None):
    """Adam over shuffled batches; the step size follows a cosine from lr
    to lr_final over the epochs (constant when lr_final is None)."""
    rng = np.random.default_rng(seed)
    h, w = X.shape[1:3]
    bufs = Batches(batch, h, w)
    hist = []
    for ep in range(epochs):
        t0 = time.time()
        lr_ep = lr if lr_final is None else lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * ep / epochs))
        perm = rng.permutation(len(X))
        losses = []
        for k in range(0, len(X) - batch + 1, batch):
            idx = np.sort(perm[k:k + batch])
            bufs.fill(X[idx], T[idx])
            losses.append(net.train_step(bufs.x, bufs.t, lr=lr_ep, shape=bufs.shape))
        val = mean_loss(net, Xv, Tv, bufs)
        hist.append({"epoch": ep + 1, "lr": float(lr_ep), "train_loss": float(np.mean(losses)), "val_loss": val,
                     "seconds": time.time() - t0})
        if ep % 5 == 4 or ep + 1 == epochs or ep < 3:
            log(f"[train] epoch {ep + 1}/{epochs}: lr {lr_ep:.2e} train {np.mean(losses):.4e} val {val:.4e} "
                f"({time.time() - t0:.1f} s, {len(losses)} steps)")
    return hist


def warm_vs_cold(net, seeds, workers: int, chunk: int = 64):
    """Per held-out slice: iterations to the gap from the cold start and from
    the network's (u, v), several slices at a time."""
    cold, warm = [], []
    with syn.process_pool(workers) as ex:
        for k in range(0, len(seeds), chunk):
            cases = list(ex.map(syn.mri320_case, seeds[k:k + chunk]))
            for use_net in (False, True):
                built = [build_solver(c) for c in cases]
                if use_net:
                    for _, o in built:
                        net.warm_start(o)
                mri.solve_concurrent([o for _, o in built], streams=SOLVE_STREAMS)
                its = [int(o.getStats()["iters"]) for _, o in built]
                (warm if use_net else cold).extend(its)
            log(f"[train] held-out {k + len(cases)}/{len(seeds)}: median iterations cold {int(np.median(cold))} "
                f"warm {int(np.median(warm))}")
    return cold, warm


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--held-out", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=None, help="cosine decay of the step size to this value")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--train-seed0", type=int, default=1000)
    ap.add_argument("--held-out-seed0", type=int, default=3000)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--out", default=ws.DEFAULT_WEIGHTS)
    ap.add_argument("--log", default="")
    a = ap.parse_args(argv)
    t_all = time.time()
    net = ws.WarmStartCNN(seed=a.seed)
    log(f"[train] network {net.channels}, {net.n_params} parameters")
    X, T, it_tr = make_pairs(range(a.train_seed0, a.train_seed0 + a.train), a.workers, a.chunk, "training")
    Xv, Tv, it_v = make_pairs(range(a.held_out_seed0, a.held_out_seed0 + a.held_out), a.workers, a.chunk,
                              "held-out")
    t_data = time.time() - t_all
    hist = train(net, X, T, Xv, Tv, a.epochs, a.batch, a.lr, a.seed, a.lr_final)
    del X, T
    cold, warm = warm_vs_cold(net, list(range(a.held_out_seed0, a.held_out_seed0 + a.held_out)), a.workers)
    summary = {
        "train_pairs": a.train, "held_out": a.held_out, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
        "lr_final": a.lr_final,
        "train_seeds": [a.train_seed0, a.train_seed0 + a.train - 1],
        "held_out_seeds": [a.held_out_seed0, a.held_out_seed0 + a.held_out - 1],
        "data_seconds": t_data, "history": hist,
        "held_out_iterations_cold": {"median": float(np.median(cold)), "mean": float(np.mean(cold))},
        "held_out_iterations_warm": {"median": float(np.median(warm)), "mean": float(np.mean(warm))},
        "held_out_iteration_ratio": float(np.sum(cold) / max(np.sum(warm), 1)),
        "total_seconds": time.time() - t_all,
    }
    net.save(a.out, epochs=a.epochs, train_pairs=a.train, val_loss=hist[-1]["val_loss"] if hist else np.nan)
    if a.log:
        with open(a.log, "w") as fh:
            json.dump(summary, fh, indent=1)
    log(f"[train] saved {a.out}; held-out iterations cold {np.mean(cold):.0f} warm {np.mean(warm):.0f} "
        f"(ratio {summary['held_out_iteration_ratio']:.2f})")
    return summary


if __name__ == "__main__":
    main(sys.argv[1:])
