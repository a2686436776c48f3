"""Train the learned warm start on the batched piles and compare solver
iterations to a residual tolerance for the zero, previous-step and learned
warm starts.

    python -m experiments_amd.train_warmstart [--scenes 4096] [--settle 240]
        [--log-steps 200] [--train-steps 20000] [--batch 65536]
        [--eval-steps 20] [--max-sweeps 4096] [--out w.npz]

Protocol:
  1. Training scenes settle for `settle` steps with the fixed-sweep plain
     solver.  The plain solver (zero warm start) then runs to the tolerance
     for `log-steps` steps, and every step's features and solved impulses are
     logged on the device.
  2. The network is trained with the library's own forward/backward/Adam
     kernels.
  3. Held-out scenes (another seed) settle the same way, once per warm start.
     Each is then stepped `eval-steps` times, every solve running to the same
     tolerance.  The sweeps used per scene and step (means per phase, also
     over the scene-steps every start solved to the tolerance), the
     scene-steps left above it and the final residuals are reported.
Prints one JSON line.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import _lib as L
from . import scenes
from .warmstart import Mlp, compare_starts, log_plain_solver


def settled_world(scene, settle, warm, tol, max_sweeps, settle_sweeps=10):
    """The scene after `settle` fixed-sweep plain steps, then switched to the
    given warm start and tolerance (same state for every warm start)."""
    w = scenes.build_world(scene, solver_iterations=settle_sweeps)
    w.use_graph(True)
    if settle:
        w.step(settle)
    w.set_solver(max_sweeps, tol, warm)
    return w


def train(X, T, steps, batch, lr=2e-3, lr_min=2e-5, chunk=1000, seed=1):
    """Adam with a cosine learning-rate schedule, `chunk` steps per call."""
    net = Mlp(seed=seed)
    hist = []
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        mid = (done + 0.5 * k) / max(steps, 1)
        lr_k = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * mid))
        hist.append(net.train(X, T, batch=batch, steps=k, lr=lr_k, seed=seed + done))
        done += k
        print("trained %d of %d steps: loss %.5f" % (done, steps, float(hist[-1][-1])),
              file=sys.stderr, flush=True)
    return net, np.concatenate(hist) if hist else np.zeros(0, np.float32)


def evaluate(scene, weights, settle, steps, tol, max_sweeps):
    out = compare_starts(scene, weights, settle, steps, tol, max_sweeps)
    for name in ("zero", "previous", "learned"):
        print("evaluated %s: %.1f normal / %.1f friction sweeps per scene" %
              (name, out[name]["normal_sweeps"], out[name]["friction_sweeps"]),
              file=sys.stderr, flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--settle", type=int, default=240)
    ap.add_argument("--log-steps", type=int, default=200)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--eval-steps", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-sweeps", type=int, default=4096)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    # torch provides the device buffers: initialise it before the library's
    # first HIP call so both use one runtime instance
    from .warmstart import _torch
    _torch()

    t0 = time.perf_counter()
    train_scene = scenes.batched_piles(n_scenes=args.scenes, iterations=args.max_sweeps)
    w = settled_world(train_scene, args.settle, L.WARM_ZERO, args.tol, args.max_sweeps)
    X, T = log_plain_solver(w, args.log_steps)
    w.close()
    print("logged %d contacts" % X.shape[0], file=sys.stderr, flush=True)
    t_log = time.perf_counter() - t0
    t1 = time.perf_counter()
    net, hist = train(X, T, args.train_steps, args.batch)
    print("trained: loss %.4f -> %.4f" % (float(hist[:10].mean()), float(hist[-100:].mean())),
          file=sys.stderr, flush=True)
    import torch
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    weights = net.get_weights()
    if args.out:
        np.savez(args.out, w_bf16=weights[0], bias=weights[1], loss=hist)
    eval_scene = scenes.batched_piles(n_scenes=args.scenes, seed=scenes.SEED + 1,
                                      iterations=args.max_sweeps)
    res = evaluate(eval_scene, weights, args.settle, args.eval_steps, args.tol, args.max_sweeps)
    print(json.dumps({
        "samples": int(X.shape[0]), "log_s": round(t_log, 2), "train_s": round(t_train, 2),
        "train_steps": args.train_steps, "batch": args.batch,
        "loss_first": float(hist[:10].mean()) if len(hist) else None,
        "loss_last": float(hist[-100:].mean()) if len(hist) else None,
        "target_var": float(T.var(0).mean()),
        "tol": args.tol, "max_sweeps": args.max_sweeps, "settle": args.settle,
        "eval_steps": args.eval_steps,
        "eval": res}))


if __name__ == "__main__":
    main()
