"""Train the learned warm start on the batched piles and compare solver
iterations to a residual tolerance for the zero, previous-step and learned
warm starts.

    python -m experiments_amd.train_warmstart [--scenes 4096] [--log-steps 200]
        [--train-steps 20000] [--batch 65536] [--eval-steps 100] [--out w.npz]

Protocol: the plain solver (zero warm start) runs to the tolerance for
`log-steps` steps on the training scenes and every step's features and solved
impulses are logged on the device; the network is trained with the library's
own forward/backward/Adam kernels; then held-out scenes (another seed) are
stepped `eval-steps` times from the same initial state with each warm start,
every solve running to the same tolerance, and the sweeps used and final
residuals are reported.  Prints one JSON line.
"""
import argparse
import json
import math
import time

import numpy as np

from . import _lib as L
from . import scenes
from .warmstart import Mlp, log_plain_solver, run_iterations


def make_world(scene, warm, tol, max_sweeps, graph=True):
    w = scenes.build_world(scene, warm_start=warm, residual_tolerance=tol,
                           solver_iterations=max_sweeps)
    w.use_graph(graph)
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
    return net, np.concatenate(hist) if hist else np.zeros(0, np.float32)


def evaluate(scene, weights, steps, tol, max_sweeps):
    out = {}
    for name, mode in (("zero", L.WARM_ZERO), ("previous", L.WARM_PREVIOUS),
                       ("learned", L.WARM_LEARNED)):
        w = make_world(scene, mode, tol, max_sweeps)
        if mode == L.WARM_LEARNED:
            w.set_warmstart_weights(*weights)
        r = run_iterations(w, steps)
        r.pop("per_step")
        r["mean_sweeps"] = r["mean_iters_normal"] + r["mean_iters_friction"]
        out[name] = r
        w.close()
    z = out["zero"]["mean_sweeps"]
    for name in ("previous", "learned"):
        out[name]["fewer_sweeps_than_zero"] = z / max(out[name]["mean_sweeps"], 1e-9)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=4096)
    ap.add_argument("--log-steps", type=int, default=200)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-sweeps", type=int, default=100)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    # torch provides the device buffers: initialise it before the library's
    # first HIP call so both use one runtime instance
    from .warmstart import _torch
    _torch()

    t0 = time.perf_counter()
    train_scene = scenes.batched_piles(n_scenes=args.scenes, iterations=args.max_sweeps)
    w = make_world(train_scene, L.WARM_ZERO, args.tol, args.max_sweeps)
    X, T = log_plain_solver(w, args.log_steps)
    w.close()
    t_log = time.perf_counter() - t0
    t1 = time.perf_counter()
    net, hist = train(X, T, args.train_steps, args.batch)
    import torch
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    weights = net.get_weights()
    if args.out:
        np.savez(args.out, w_bf16=weights[0], bias=weights[1], loss=hist)
    eval_scene = scenes.batched_piles(n_scenes=args.scenes, seed=scenes.SEED + 1,
                                      iterations=args.max_sweeps)
    res = evaluate(eval_scene, weights, args.eval_steps, args.tol, args.max_sweeps)
    print(json.dumps({
        "samples": int(X.shape[0]), "log_s": round(t_log, 2), "train_s": round(t_train, 2),
        "train_steps": args.train_steps, "batch": args.batch,
        "loss_first": float(hist[:10].mean()) if len(hist) else None,
        "loss_last": float(hist[-100:].mean()) if len(hist) else None,
        "target_var": float(T.var(0).mean()),
        "tol": args.tol, "max_sweeps": args.max_sweeps, "eval_steps": args.eval_steps,
        "eval": res}))


if __name__ == "__main__":
    main()
