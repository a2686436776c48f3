"""Learned warm start of the contact solver: the network handle, logging of
training pairs from the plain solver and the A/B evaluation of iterations to
a residual tolerance.

Every forward, backward and Adam step runs in libmi_rb.so's kernels (BF16
MFMA forward, FP32 MFMA backward); torch only provides device buffers.  The
reference has no counterpart: its solver starts every step from zero impulses
(dynamics/basic_demo/Stubs/solverSetup.cl:529-531).
"""
import ctypes
import os

import numpy as np

from . import _lib as L

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")

FEATURES = ["n.x", "n.y", "n.z", "10*min depth", "vr.n", "vr.t0", "vr.t1", "1/mA", "1/mB",
            "10*prev lambda_n", "10*prev friction0", "10*prev friction1", "npts/4", "has prev",
            "-mean jac_inv", "mean b"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned warm start trains on the GPU; no device is visible")
    return torch


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class Mlp:
    """16-64-64-3 network: bf16 weights, fp32 biases, Adam state on device."""

    def __init__(self, seed=1):
        self.h = ctypes.c_void_p()
        L.call("mi_rb_mlp_create", int(seed), ctypes.byref(self.h))

    def close(self):
        if self.h:
            L.lib().mi_rb_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        w = np.zeros(L.MLP_NW, np.uint16)
        b = np.zeros(L.MLP_NB, np.float32)
        L.call("mi_rb_mlp_get_weights", self.h, w.ctypes.data_as(ctypes.c_void_p),
               b.ctypes.data_as(ctypes.c_void_p))
        return w, b

    def set_weights(self, w_bf16, bias):
        w = np.ascontiguousarray(w_bf16, np.uint16).reshape(-1)
        b = np.ascontiguousarray(bias, np.float32).reshape(-1)
        if w.size != L.MLP_NW or b.size != L.MLP_NB:
            raise ValueError("expected %d bf16 weights and %d biases" % (L.MLP_NW, L.MLP_NB))
        L.call("mi_rb_mlp_set_weights", self.h, w.ctypes.data_as(ctypes.c_void_p),
               b.ctypes.data_as(ctypes.c_void_p))

    def forward(self, x):
        """x: float32 device tensor (n, 16) -> device tensor (n, 3)."""
        torch = _torch()
        x = x.contiguous()
        y = torch.empty((x.shape[0], 3), dtype=torch.float32, device=x.device)
        L.call("mi_rb_mlp_forward", self.h, _vp(x), int(x.shape[0]), _vp(y))
        return y

    def gradients(self, x, t):
        """Gradient of the MSE loss over all rows: (device tensor [5312+131], loss)."""
        torch = _torch()
        x, t = x.contiguous(), t.contiguous()
        g = torch.empty(L.MLP_NW + L.MLP_NB, dtype=torch.float32, device=x.device)
        loss = ctypes.c_float()
        L.call("mi_rb_mlp_gradients", self.h, _vp(x), _vp(t), int(x.shape[0]), _vp(g),
               ctypes.byref(loss))
        return g, loss.value

    def train(self, x, t, batch, steps, lr, seed=0):
        """Adam steps on minibatches drawn from (x, t); returns the loss history."""
        hist = np.zeros(max(steps, 1), np.float32)
        L.call("mi_rb_mlp_train", self.h, _vp(x), _vp(t), int(x.shape[0]), int(batch), int(steps),
               float(lr), int(seed), hist.ctypes.data_as(ctypes.c_void_p))
        return hist[:steps]


def log_plain_solver(world, steps, every=1):
    """Step `world` (configured with the zero warm start and a residual
    tolerance) and collect each logged step's features and solved impulses.
    Returns device tensors X (N, 16), T (N, 3)."""
    torch = _torch()
    cap = int(world.cfg.max_contacts)
    xs, ts = [], []
    bx = torch.empty((cap, 16), dtype=torch.float32, device="cuda")
    bt = torch.empty((cap, 3), dtype=torch.float32, device="cuda")
    for k in range(steps):
        world.step(1)
        if k % every:
            continue
        n = world.copy_training_batch(bx.data_ptr(), bt.data_ptr(), cap)
        xs.append(bx[:n].clone())
        ts.append(bt[:n].clone())
    return torch.cat(xs), torch.cat(ts)


def run_iterations(world, steps):
    """Step and record per-step solver sweeps (mean over scenes and maximum)
    and final residuals."""
    rows = []
    for _ in range(steps):
        world.step(1)
        st = world.stats()
        rows.append((st["mean_normal_iterations"], st["mean_friction_iterations"],
                     st["normal_iterations"], st["friction_iterations"], st["normal_residual"],
                     st["friction_residual"]))
    a = np.array(rows, np.float64).reshape(-1, 6)
    return dict(mean_iters_normal=float(a[:, 0].mean()), mean_iters_friction=float(a[:, 1].mean()),
                max_iters_normal=int(a[:, 2].max()), max_iters_friction=int(a[:, 3].max()),
                max_residual_normal=float(a[:, 4].max()), max_residual_friction=float(a[:, 5].max()),
                per_step=a)


def run_scene_iterations(world, steps):
    """Step a batched world solved to a tolerance and record, per step and
    scene, the normal and friction sweeps used and the final residuals (the
    solver's per-scene record).  Returns int arrays normal, friction (steps,
    scenes), float arrays res_normal, res_friction and a bool mask `active`
    of the scenes that had contacts."""
    nn, nf, rn, rf, act = [], [], [], [], []
    for _ in range(steps):
        world.step(1)
        st = world.read(L.BUF_TILE_STATS, np.int32, (4,))
        act.append(st[:, 0] >= 0)
        nn.append(st[:, 0].copy())
        nf.append(st[:, 1].copy())
        rn.append(st[:, 2].view(np.float32).copy())
        rf.append(st[:, 3].view(np.float32).copy())
    return dict(normal=np.array(nn), friction=np.array(nf), res_normal=np.array(rn),
                res_friction=np.array(rf), active=np.array(act))


def _start_summary(r, tol):
    a = r["active"]
    return {"normal_sweeps": round(float(r["normal"][a].mean()), 2),
            "friction_sweeps": round(float(r["friction"][a].mean()), 2),
            "sweeps": round(float((r["normal"] + r["friction"])[a].mean()), 2),
            "max_normal_sweeps": int(r["normal"][a].max()),
            "max_friction_sweeps": int(r["friction"][a].max()),
            "unconverged_scene_steps_normal": int((a & (r["res_normal"] >= tol)).sum()),
            "unconverged_scene_steps_friction": int((a & (r["res_friction"] >= tol)).sum()),
            "final_residual_max_normal": float(r["res_normal"][a].max()),
            "final_residual_max_friction": float(r["res_friction"][a].max())}


def compare_starts(scene, weights, settle, steps, tol=1e-4, max_sweeps=4096, settle_sweeps=10,
                   skip=5):
    """Sweeps to the residual tolerance for the zero, previous-step and (with
    weights) learned warm starts, each run from the same state: `scene`
    advanced `settle` steps with `settle_sweeps` fixed sweeps, then `skip`
    unrecorded and `steps` recorded steps solved to `tol` (at most
    `max_sweeps` per phase; the skipped steps keep the first warm starts, which
    read impulses of fixed-sweep solves, out of the record).  Sweeps are
    counted per scene and step from the solver's per-scene record: means over
    the scene-steps with contacts, the scene-steps that ended above the
    tolerance, and the means over the scene-steps that every start solved to
    the tolerance (`converged_in_every_start`), with the ratios between
    starts."""
    from . import scenes
    starts = [("zero", L.WARM_ZERO), ("previous", L.WARM_PREVIOUS)]
    if weights is not None:
        starts.append(("learned", L.WARM_LEARNED))
    runs, out = {}, {}
    for name, mode in starts:
        w = scenes.build_world(scene, solver_iterations=settle_sweeps)
        try:
            w.use_graph(True)
            w.step(settle)
            if mode == L.WARM_LEARNED:
                w.set_warmstart_weights(*weights)
            w.set_solver(max_sweeps, tol, mode)
            w.use_graph(True)
            w.step(skip)
            runs[name] = run_scene_iterations(w, steps)
            flags = w.stats()["overflow_flags"]
        finally:
            w.close()
        out[name] = _start_summary(runs[name], tol)
        out[name]["overflow_flags"] = flags
    common = np.ones_like(runs["zero"]["active"])
    for r in runs.values():
        common &= r["active"] & (r["res_normal"] < tol) & (r["res_friction"] < tol)
    conv = {"scene_steps": int(common.sum()), "of": int(runs["zero"]["active"].sum())}
    for name, r in runs.items():
        conv[name] = {"normal_sweeps": round(float(r["normal"][common].mean()), 2),
                      "sweeps": round(float((r["normal"] + r["friction"])[common].mean()), 2)}
    out["converged_in_every_start"] = conv

    def ratios(d, ref, tag):
        d["fewer_sweeps_than_" + tag] = round(ref["sweeps"] / d["sweeps"], 3)
        d["fewer_normal_sweeps_than_" + tag] = round(ref["normal_sweeps"] / d["normal_sweeps"], 3)

    for name in runs:
        if name != "zero":
            ratios(out[name], out["zero"], "zero")
            ratios(conv[name], conv["zero"], "zero")
    if "learned" in runs:
        ratios(out["learned"], out["previous"], "previous")
        ratios(conv["learned"], conv["previous"], "previous")
    return out


def load_weights(path=WEIGHTS_PATH):
    """The trained network shipped with the package (written by
    experiments_amd.train_warmstart --out): (bf16 weights, fp32 biases), or
    None when there is none."""
    if not os.path.exists(path):
        return None
    z = np.load(path)
    return z["w_bf16"].astype(np.uint16), z["bias"].astype(np.float32)
