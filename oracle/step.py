"""Oracle restatement of one step on a World's exact state (test infrastructure).

Chains the oracle stages in the reference's order (CLPhysicsDemo::stepSimulation,
opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.cpp:740-943) on the same inputs the
GPU sees, and reconstructs the GPU's constraint visiting order so the
Gauss-Seidel sweeps are compared on one schedule.
"""
import numpy as np

import oracle as O
from experiments_amd import _lib as L  # registered arrays and buffers of a World
from experiments_amd.world import shape_arrays


class Snapshot:
    def __init__(self, w):
        self.state = w.get_state()
        self.sh = shape_arrays(w)
        self.params = w.params()
        self.shapes = O.Shapes(self.sh["hulls"], self.sh["verts"], self.sh["planes"],
                               self.sh["face_index"], self.sh["indices"], self.sh["edges"])
        self.is_large = w.read(L.BUF_IS_LARGE, np.int32)
        self.cfg = w.cfg
        sc = self.sh["scene"]
        ns = int(sc.max()) + 1 if len(sc) else 1
        self.scene_start = np.array([int(np.argmax(sc == s)) for s in range(ns)], np.int32)
        self.scene_count = np.bincount(sc, minlength=ns).astype(np.int32)


def oracle_broadphase(s):
    st, sh, p = s.state, s.sh, s.params
    mn, mx = O.world_aabbs(st["pos"], st["quat"], sh["shape"], sh["hulls"])
    kv = O.calc_hash(mn, mx, np.float32(p["inv_cell"]), p["dims"], p["cell_bits"],
                     scene=sh["scene"], is_large=s.is_large)
    k, v = O.radix_sort_kv(kv[:, 0], kv[:, 1], p["key_bits"])
    skv = np.stack([k, v], 1)
    pairs = O.find_pairs_grid(skv, mn, mx, np.float32(p["inv_cell"]), p["dims"], p["cell_bits"],
                              scene=sh["scene"], is_large=s.is_large, scene_start=s.scene_start,
                              scene_count=s.scene_count)
    return mn, mx, skv, pairs


def oracle_setup(s):
    st, sh, c = s.state, s.sh, s.cfg
    lv, invI = O.setup_bodies(st["linvel"], st["angvel"], st["quat"], sh["inv_mass"],
                              sh["inv_inertia_local"], list(c.gravity), c.dt)
    av = st["angvel"].copy()
    av[sh["inv_mass"] == 0] = 0
    lv[sh["inv_mass"] == 0] = 0
    return lv, av, invI


def oracle_contacts(s, pairs):
    st, sh, c = s.state, s.sh, s.cfg
    slots, has = O.narrowphase(pairs, st["pos"], st["quat"], sh["inv_mass"], sh["shape"],
                               s.shapes, c.friction, c.contact_threshold)
    return slots[has == 1]


def gpu_order(w):
    """Sequential constraint order equivalent to the GPU's tile schedule."""
    order = w.read(L.BUF_SOLVE_ORDER, np.int32)
    tiles = w.read(L.BUF_PHASES, np.int32, (4,))
    seq = []
    for ph in range(4):
        for t in range(len(tiles)):
            first, end, phase, _ = tiles[t]
            if phase == ph and end > first:
                seq.append(order[first:end])
    return np.concatenate(seq) if seq else np.zeros(0, np.int32)


def gpu_tile_orders(w):
    """The GPU's constraint order split by solver tile (scene tiles solve
    independently), in tile order."""
    order = w.read(L.BUF_SOLVE_ORDER, np.int32)
    tiles = w.read(L.BUF_PHASES, np.int32, (4,))
    return [order[first:end] for first, end, _, _ in tiles if end > first]


def oracle_step(s, order, contacts=None, prev=None, warm=0, prediction=None, segments=None):
    """Full step on the oracle from the snapshot state, visiting constraints in
    `order`.  With `segments` (a list of orders, one per independent tile) each
    segment is solved on its own, stopping at the residual tolerance of its
    own rows.  prev: (keys int2[m], totals float4[m], points float4[m][4]) of
    the previous solve;
    warm: 0 zero, 1 previous impulses, 2 impulses from `prediction` (n, 3).
    Returns pos, quat, linvel, angvel, the solved constraints and the
    warm-start features."""
    st, sh, c = s.state, s.sh, s.cfg
    if contacts is None:
        _, _, _, pairs = oracle_broadphase(s)
        contacts = oracle_contacts(s, pairs)
    lv, av, invI = oracle_setup(s)
    cons = O.contact_to_constraint(contacts, st["pos"], lv, av, sh["inv_mass"], invI,
                                   c.solver_dt if c.solver_dt > 0 else c.dt, c.position_drift,
                                   c.baumgarte, c.friction)
    keys, imp, pts = prev if prev is not None else (None, None, None)
    feats, cons = O.contact_features(contacts, cons, st["pos"], lv, av, sh["inv_mass"], keys, imp,
                                     pts, apply_previous=(warm != 0))
    if warm == 2:
        cons = O.warm_from_prediction(cons, prediction, c.friction)
    if segments is None:
        cons, lv2, av2, it, res = O.solve(cons, order, lv, av, st["pos"], sh["inv_mass"], invI,
                                          c.solver_iterations, c.solver_iterations, c.friction,
                                          c.residual_tolerance)
    else:
        lv2, av2, it, res = lv, av, [], []
        for seg in segments:
            cons, lv2, av2, i_s, r_s = O.solve(cons, seg, lv2, av2, st["pos"], sh["inv_mass"],
                                               invI, c.solver_iterations, c.solver_iterations,
                                               c.friction, c.residual_tolerance)
            it.append(i_s)
            res.append(r_s)
    pos, quat, av3 = O.integrate(st["pos"], st["quat"], lv2, av2, sh["inv_mass"], c.dt,
                                 c.angular_damping, c.max_angular_motion)
    return dict(pos=pos, quat=quat, linvel=lv2, angvel=av3, cons=cons, iters=it, res=res,
                features=feats)
