"""Train the warm-start network on the GPU from 64-iteration solves of seeded drapes.

    python -m taremincloth_amd.train_warmstart --out taremincloth_amd/weights/warmstart_r3.npz

Scenes: 128 x 128 cloths at 1 cm (2 triangles per quad, areal density 0.15 kg/m^2,
compliances 1e-8 / 1e-6 / 1e-5, self-collision off) dropped from a seeded height over a
seeded sphere and, in half of the scenes, a capsule; friction seeded in [0.2, 0.5].  Each
scene runs `--frames` frames of 10 substeps at 64 plain iterations and records every
`--stride`-th substep (features, target correction relative to `--history-weight` times
the previous correction).  One extra scene is held out for the validation loss.  Training and data stay on the device (tc_sim_ws_record, tc_ws_train).
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

from taremincloth_amd import taremin_cloth_core as T  # noqa: E402
from taremincloth_amd.warmstart import WarmStartNet  # noqa: E402

DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "warmstart_r3.npz")


def grid(n, spacing, height, rng):
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    half = (n - 1) * spacing * 0.5
    pos = np.stack([(i * spacing - half).ravel(), (j * spacing - half).ravel(), np.full(n * n, height)], axis=1)
    pos = (pos + rng.uniform(-1e-3, 1e-3, pos.shape)).astype(np.float32)
    idx = np.arange(n * n, dtype=np.uint32).reshape(n, n)
    edges = np.concatenate([np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1),
                            np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], 1)]).astype(np.uint32)
    v00, v10, v01, v11 = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    faces = np.empty((2 * v00.size, 3), np.uint32)
    faces[0::2] = np.stack([v00, v10, v11], 1)
    faces[1::2] = np.stack([v00, v11, v01], 1)
    return pos, edges, faces


def make_scene(seed, n=128, iters=64):
    rng = np.random.default_rng(seed)
    pos, edges, faces = grid(n, 0.01, float(rng.uniform(0.35, 0.7)), rng)
    inv = T.compute_combined_inv_masses(pos, faces, areal_density=0.15)
    sim = T.ClothSimulator(pos, edges, faces, inv, thickness=0.005, stiffness=1000.0, bending_stiffness=10.0)
    sim.set_gravity(0.0, 0.0, -9.81)
    sim.set_compliances_raw(1e-8, 1e-6, 1e-5)
    sim.set_enable_self_collision(False)
    fr = float(rng.uniform(0.2, 0.5))
    c = rng.uniform(-0.25, 0.25, 2)
    sim.add_sphere_collider([float(c[0]), float(c[1]), 0.0], float(rng.uniform(0.15, 0.35)), fr, 0.0)
    if rng.random() < 0.5:
        y, z, hl = float(rng.uniform(-0.4, 0.4)), float(rng.uniform(0.0, 0.2)), float(rng.uniform(0.2, 0.4))
        sim.add_capsule_collider([-hl, y, z], [hl, y, z], float(rng.uniform(0.05, 0.12)), fr, 0.0)
    sim.set_solver_iterations(iters)
    return sim


def record(seeds, frames, stride, n, history_weight):
    sims = [make_scene(s, n) for s in seeds]
    for sim in sims:
        sim.ws_record(frames * 10 // stride + 1, stride, history_weight)
    for _ in range(frames):
        for sim in sims:  # the simulators' streams run side by side
            sim.step(1.0 / 60.0, 10, 64)
    for sim in sims:
        sim.synchronize()
    return sims


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--frames", type=int, default=62)
    ap.add_argument("--stride", type=int, default=4)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--schedule", default="1e-3:30000,3e-4:20000,1e-4:10000")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--history-weight", type=float, default=0.5,
                    help="weight of the previous correction in the warm start the network is trained for")
    args = ap.parse_args(argv)

    t0 = time.time()
    hw = args.history_weight
    train = record([1000 + k for k in range(args.scenes)], args.frames, args.stride, args.grid, hw)
    held = record([9000], args.frames, args.stride, args.grid, hw)
    substeps = args.scenes * args.frames * 10
    nrec = sum(s.ws_num_records() for s in train)
    print(f"recorded {nrec} samples from {substeps} training substeps ({time.time() - t0:.1f} s)", flush=True)

    net = WarmStartNet(args.rounds, hw)
    net.init(seed=args.seed)
    log = {"initial_train": net.eval(train), "initial_val": net.eval(held), "phases": []}
    print(f"initial loss train {log['initial_train']:.3f} val {log['initial_val']:.3f}", flush=True)
    step_seed = args.seed
    for part in args.schedule.split(","):
        lr, steps = part.split(":")
        t1 = time.time()
        losses = net.train(train, steps=int(steps), lr=float(lr), seed=step_seed)
        step_seed += 1
        el = time.time() - t1
        tr, va = net.eval(train), net.eval(held)
        log["phases"].append({"lr": float(lr), "steps": int(steps), "seconds": el, "last_batch_loss": float(losses[-200:].mean()),
                              "train": tr, "val": va})
        print(f"lr {lr} x {steps}: {int(steps) / el:.0f} steps/s, train {tr:.3f} val {va:.3f}", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    meta = {"grid": args.grid, "scenes": args.scenes, "frames": args.frames, "stride": args.stride,
            "training_substeps": substeps, "samples": nrec, "schedule": args.schedule, "history_weight": hw,
            "log": log}
    net.save(args.out, meta=np.array(json.dumps(meta)))
    print(f"saved {args.out} ({time.time() - t0:.1f} s total)", flush=True)


if __name__ == "__main__":
    main()
