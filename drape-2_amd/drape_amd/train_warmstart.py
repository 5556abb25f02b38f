"""Train the learned Newton warm start on solver-generated frames and save its weights.

Data (state before a step; target: the velocity residual the step realised beyond free flight, dv - h g), all
from the plain solver: drops of the 32x32 sheet
(the 1,024-particle configuration) from many jitter seeds, the 65,536-particle drop and the 120,000-particle
tube. Training uses the network's own GPU forward/backward kernels (Adam, stacked batches of one mesh).

    python -m drape_amd.train_warmstart --out drape-2_amd/drape_amd/weights/warmstart.npy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from drape_amd import learn, scenes  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--sheet-seeds", type=int, default=300)
    ap.add_argument("--sheet-steps", type=int, default=60)
    ap.add_argument("--drop-steps", type=int, default=120)
    ap.add_argument("--drop-every", type=int, default=2)
    ap.add_argument("--drop-seeds", type=int, default=1, help="jitter seeds of the 65,536-particle drop")
    ap.add_argument("--tube-steps", type=int, default=200)
    ap.add_argument("--tube-every", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-4)
    ap.add_argument("--batch-vertices", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    t0 = time.time()

    def log(msg):
        print(f"[{time.time() - t0:7.1f}s] {msg}", flush=True)

    samples = []
    sheet_graph = learn.MeshGraph.of_scene(scenes.c1_sheet())
    for k in range(args.sheet_seeds):
        samples += learn.collect(scenes.c1_sheet(seed=1000 + k), args.sheet_steps, graph=sheet_graph)
        if k % 25 == 0:
            log(f"sheet seed {k}: {len(samples)} frames")
    n_sheet = len(samples)
    if args.drop_steps:
        drop_graph = learn.MeshGraph.of_scene(scenes.c2_sheet())
        for k in range(args.drop_seeds):
            samples += learn.collect(scenes.c2_sheet(seed=2 + 100 * k), args.drop_steps, every=args.drop_every,
                                     graph=drop_graph)
            log(f"65k drop seed {k}: {len(samples) - n_sheet} frames")
    n_drop = len(samples)
    if args.tube_steps:
        samples += learn.collect(scenes.c3_tube(), args.tube_steps, every=args.tube_every)
        log(f"120k tube: {len(samples) - n_drop} frames")
    vertices = sum(s.graph.n for s in samples)
    log(f"{len(samples)} (state, residual) pairs, {vertices} vertex samples")

    net = learn.WarmStartNet(seed=args.seed)
    hist = learn.train(net, samples, args.epochs, lr=args.lr, lr_final=args.lr_final, seed=args.seed, log=log,
                       batch_vertices=args.batch_vertices)
    # Held-out drop of the sheet: guess error of the network (u = y + h d) against the previous-velocity guess
    # (u = h v, a residual of -h g on the free rows) over the steps' realised residuals.
    held = learn.collect(scenes.c1_sheet(seed=999), 40, graph=sheet_graph)
    num = den = 0.0
    plain_res = learn.residual_target(scenes.c1_sheet(seed=999), np.zeros((1024, 3), np.float32), 1 / 60)
    for s in held[1:]:
        num += float(np.sum((net.predict(s.graph, s.frame) - s.target) ** 2))
        den += float(np.sum((plain_res.astype(np.float64) - s.target) ** 2))
    ratio = float(np.sqrt(num / den))
    log(f"held-out sheet drop: guess error network / previous velocity = {ratio:.3f}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    summary = {"pairs": len(samples), "sheet_frames": n_sheet, "drop_frames": n_drop - n_sheet,
               "tube_frames": len(samples) - n_drop, "vertex_samples": vertices, "epochs": args.epochs,
               "loss_per_epoch": hist, "held_out_guess_error_ratio": ratio, "seconds": time.time() - t0,
               "args": vars(args)}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
