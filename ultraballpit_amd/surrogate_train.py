#!/usr/bin/env python3
"""Train the dam-break surrogate on the GPU and save it.

Rollouts: seeded dam breaks of 262,144 and 1,048,576 particles (same tank
as the 1M benchmark scene; fill height and jitter seed varied, none equal
to the benchmark scene's) generated by the plain
engine at a fixed dt; after `--warm` settling cycles, rows are collected over
`--cycles` correction cycles (K = 8 landing states 4 dt apart plus the mark
state, a fraction `--frac` of the particles each).  The network is trained
with the project's BF16-MFMA kernels (Adam, cosine learning rate), then `--rounds` - 1
rounds add rows on the network's own states (corrective targets back to the plain
trajectory) and retrain; the result is saved
as an .npz of FP32 arrays (ubp_mlp_model fields).

    python -m ultraballpit_amd.surrogate_train --out ultraballpit_amd/models/dam_break_surrogate.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import ultraballpit_amd as ubp  # noqa: E402
from ultraballpit_amd import scenes, surrogate  # noqa: E402


def rollout_scenes(counts, seed: int = 7):
    """Dam breaks of counts[k] particles: water column 0.25 x 0.25 x H with
    H in [0.4, 0.5] m (lattice spacing from the particle count), seeds."""
    u = scenes.uniform(seed, 300, len(counts))
    out = []
    for k, n_keep in enumerate(counts):
        hgt = 0.4 + 0.1 * float(u[k])
        dx = (0.25 * 0.25 * hgt / n_keep) ** (1.0 / 3.0)
        lat = (int(round(0.25 / dx)), int(round(0.25 / dx)), int(np.ceil(hgt / (0.25 / round(0.25 / dx)))))
        out.append(scenes.dam_break(lat=lat, n_keep=min(n_keep, lat[0] * lat[1] * lat[2]),
                                    column=(0.25, 0.25, hgt), seed=1000 + k, sdf_nodes=(128, 32, 80)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "models", "dam_break_surrogate.npz"))
    ap.add_argument("--particles", default="262144,262144,262144,262144,1048576,1048576",
                    help="particle count of each rollout")
    ap.add_argument("--rollouts", type=int, default=0,
                    help="instead of --particles: this many rollouts of --rollout-particles each")
    ap.add_argument("--rollout-particles", type=int, default=262144)
    ap.add_argument("--warm", type=int, default=10, help="settling cycles before collection")
    ap.add_argument("--cycles", type=int, default=10)
    ap.add_argument("--frac", type=float, default=0.125)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--rounds", type=int, default=4, help="1 + rounds of rows on the network's own states")
    args = ap.parse_args()
    if not ubp.gpu_available():
        sys.exit("no gfx950 device")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    counts = ([args.rollout_particles] * args.rollouts if args.rollouts > 0
              else [int(x) for x in args.particles.split(",")])
    sc = rollout_scenes(counts)
    t0 = time.perf_counter()
    model, losses, info = surrogate.train(sc, cycles=args.cycles, steps=args.steps, batch=args.batch, lr=args.lr,
                                          frac=args.frac, seed=args.seed, warm_cycles=args.warm,
                                          rounds=args.rounds, log=log)
    info.update({"rounds": args.rounds, "rollouts": len(sc), "particles": [len(s["x"]) for s in sc], "warm_cycles": args.warm,
                 "cycles": args.cycles, "frac": args.frac, "steps": args.steps, "batch": args.batch,
                 "lr": args.lr, "wall_s": time.perf_counter() - t0,
                 "loss_curve": [float(x) for x in losses[:: max(1, len(losses) // 50)]]})
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    ubp.save_model(args.out, model)
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(info, f, indent=1)
    print(json.dumps({k: v for k, v in info.items() if k != "loss_curve"}), flush=True)


if __name__ == "__main__":
    main()
