"""Train the near-separatrix classifier on the GPU and calibrate its threshold.

    python -m msceer_amd.train_sepnet [--volumes 64] [--size 128] [--epochs 8] [--out PATH]

Training data: volumes 5000+i of the two families (smoothed noise and the
Gaussian-mixture labelling family, `learned.training_volume`), labelled by
the full GPU integrator; targets from `learned.separatrix_target`.  The
threshold is the largest candidate whose learned labels differ from the full
integrator's on at most `--budget` of the vertices of held-out volumes and
whose mask leaves at least 1 - `--max-fraction` of them to the combinatorial
labels.  When no candidate does both, the mask would integrate (nearly) every
vertex and cost more than the plain integrator: the script reports that and
saves nothing (exit status 2).
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import learned, native


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--volumes", type=int, default=64)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pos-weight", type=float, default=4.0)
    ap.add_argument("--holdout", type=int, default=4)
    ap.add_argument("--budget", type=float, default=5e-4)
    ap.add_argument("--dilation", type=int, default=2)
    ap.add_argument("--max-fraction", type=float, default=0.5,
                    help="largest mask fraction (vertices integrated) worth using")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=learned.WEIGHTS)
    args = ap.parse_args(argv)

    ctx = native.Context(0)
    t0 = time.perf_counter()
    vols, targets = [], []
    for i in range(args.volumes + args.holdout):
        f = learned.training_volume(i, args.size).astype(np.float32)
        acc, comb = learned.label_pair(f, ctx)
        t = learned.separatrix_target(acc, comb)
        vols.append(f)
        targets.append(t)
        if i % 8 == 0:
            log(f"volume {i}: target fraction {t.mean():.3f}, combinatorial mismatch {(acc != comb).mean():.3f} "
                f"({time.perf_counter() - t0:.1f} s)")
    train_v, train_t = vols[:args.volumes], targets[:args.volumes]
    hold_v = vols[args.volumes:]

    net = learned.SeparatrixNet(ctx, seed=args.seed)
    net.dilation = args.dilation
    t1 = time.perf_counter()
    hist = []
    for ep in range(args.epochs):
        lr = args.lr * (0.5 ** (ep // 3))
        hist += learned.train(net, train_v, train_t, epochs=1, lr=lr, pos_weight=args.pos_weight)
        log(f"epoch {ep}: lr {lr:.1e}, mean loss {np.mean(hist[-len(train_v):]):.4f} "
            f"({time.perf_counter() - t1:.1f} s)")
    train_s = time.perf_counter() - t1

    # threshold: largest candidate within the mismatch budget on held-out volumes
    table = []
    chosen = None
    for thr in (0.7, 0.6, 0.5, 0.4, 0.3, 0.2, 0.1, 0.05, 0.02):
        net.threshold = thr
        mism, frac = [], []
        for f in hold_v:
            _, plain, _ = learned.labels.ondemand_asc(f, want_grad=False, ctx=ctx)
            _, lab, _, fr = learned.learned_labels(f, net, ctx, want_grad=False)
            mism.append(float((lab != plain).mean()))
            frac.append(fr)
        row = (thr, float(np.mean(mism)), float(np.max(mism)), float(np.mean(frac)))
        table.append(row)
        log(f"threshold {thr}: mismatch mean {row[1]:.5f} max {row[2]:.5f}, mask fraction {row[3]:.3f}")
        if row[2] <= args.budget:
            if row[3] <= args.max_fraction:
                chosen = thr
            break  # lower thresholds only flag more vertices
    if chosen is None:
        log(f"no threshold keeps the mismatch within {args.budget} with a mask fraction of at most "
            f"{args.max_fraction}: the mask would not save work; nothing saved")
        net.close()
        ctx.close()
        return 2
    net.threshold = chosen
    net.save(args.out, loss_history=np.array(hist, np.float32), calibration=np.array(table, np.float64),
             train_seconds=train_s, volumes=args.volumes, size=args.size, epochs=args.epochs)
    log(f"saved {args.out}: threshold {chosen}, dilation {net.dilation}, training {train_s:.1f} s")
    net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
