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
        # compound children as proxy bodies after the n bodies (in body order,
        # children in order), as the library numbers them
        self.children = w.read(L.BUF_CHILDREN, np.uint8, (48,)).view(O.CHILD_DTYPE).reshape(-1)
        hulls = self.sh["hulls"]
        self.proxy_first = np.full(len(self.sh["shape"]), -1, np.int64)
        proxies = []
        for i, sid in enumerate(self.sh["shape"]):
            h = hulls[sid]
            if h["num_vertices"] == 0 and h["num_edges"] > 0:
                self.proxy_first[i] = len(proxies)
                proxies += [(i, int(h["edge_offset"]) + c) for c in range(int(h["num_edges"]))]
        self.proxy = np.array(proxies, np.int32).reshape(-1, 2)
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


def is_mesh(h):
    """A compound whose num_indices is nonzero is a concave triangle mesh."""
    return h["num_vertices"] == 0 and h["num_edges"] > 0 and h["num_indices"] != 0


def mesh_triangles_overlapping(s, m, b, mn, mx):
    """Triangles (child indices) of mesh body m whose world AABB overlaps
    body b's (sat.cl:985-1023, the triangle taken to world space by m's
    transform), in triangle order."""
    h = s.sh["hulls"][s.sh["shape"][m]]
    T, e0 = int(h["num_edges"]), int(h["edge_offset"])
    tri = s.shapes.hulls[s.children["shape"][e0:e0 + T]]
    vo = tri["vertex_offset"].astype(np.int64)
    local = s.shapes.verts[(vo[:, None] + np.arange(3)).reshape(-1)]
    pseudo = np.zeros(3 * T, O.CHILD_DTYPE)      # the vertices as children at the identity
    pseudo["pos"][:, :3] = local[:, :3]
    pseudo["orn"][:, 3] = 1.0
    proxy = np.stack([np.zeros(3 * T, np.int32), np.arange(3 * T, dtype=np.int32)], 1)
    wp, _ = O.compound_proxies(proxy, pseudo, s.state["pos"][m:m + 1], s.state["quat"][m:m + 1])
    wv = wp[1:, :3].reshape(T, 3, 3)
    lo, hi = wv.min(axis=1), wv.max(axis=1)
    bmn, bmx = mn[b, :3], mx[b, :3]
    sep = (lo > bmx) | (hi < bmn)
    return np.nonzero(~sep.any(axis=1))[0]


def expand_compound_pairs(s, pairs):
    """The pair list the narrowphase sees: the broadphase pairs, then for
    each pair with a compound (not both static) one pair per (child of A or
    A, child of B or B), child of A outer (findCompoundPairsKernel,
    sat.cl:763-925), in pair order.  A concave mesh pairs with hull bodies
    only, one pair per triangle whose world AABB overlaps the hull's."""
    if len(s.proxy) == 0:
        return pairs
    hulls, shape, im = s.sh["hulls"], s.sh["shape"], s.sh["inv_mass"]
    n = len(shape)
    mn, mx = O.world_aabbs(s.state["pos"], s.state["quat"], shape, hulls)

    def kids(b):
        h = hulls[shape[b]]
        return int(h["num_edges"]) if h["num_vertices"] == 0 and h["num_edges"] > 0 else 0

    extra = []
    for a, b in np.asarray(pairs).reshape(-1, 2):
        ha, hb = hulls[shape[a]], hulls[shape[b]]
        if is_mesh(ha) or is_mesh(hb):
            other = hb if is_mesh(ha) else ha
            if (is_mesh(ha) and is_mesh(hb)) or other["num_vertices"] == 0 or \
                    (im[a] == 0 and im[b] == 0):
                continue
            m, o = (a, b) if is_mesh(ha) else (b, a)
            for t in mesh_triangles_overlapping(s, m, o, mn, mx):
                p = n + s.proxy_first[m] + int(t)
                extra.append((p, b) if m == a else (a, p))
            continue
        ca, cb = kids(a), kids(b)
        if not (ca or cb) or (im[a] == 0 and im[b] == 0):
            continue
        for i in range(ca or 1):
            for j in range(cb or 1):
                extra.append((n + s.proxy_first[a] + i if ca else a,
                              n + s.proxy_first[b] + j if cb else b))
    return np.concatenate([np.asarray(pairs).reshape(-1, 2),
                           np.array(extra, np.int32).reshape(-1, 2)]).astype(np.int32)


def oracle_contacts(s, pairs):
    st, sh, c = s.state, s.sh, s.cfg
    pos, quat, im, shape = st["pos"], st["quat"], sh["inv_mass"], sh["shape"]
    owner = np.arange(len(shape))
    if len(s.proxy):
        pairs = expand_compound_pairs(s, pairs)
        pos, quat = O.compound_proxies(s.proxy, s.children, pos, quat)
        im = np.concatenate([im, im[s.proxy[:, 0]]]).astype(np.float32)
        shape = np.concatenate([shape, s.children["shape"][s.proxy[:, 1]]]).astype(np.int32)
        owner = np.concatenate([owner, s.proxy[:, 0]])
    slots, has = O.narrowphase(pairs, pos, quat, im, shape, s.shapes, c.friction,
                               c.contact_threshold, sat_mode=getattr(c, "sat_mode", 0),
                               reduction_mode=getattr(c, "reduction_mode", 0))
    out = slots[has == 1].copy()
    if len(s.proxy):
        # a compound child names its compound (sign: static)
        for k in ("body_a", "body_b"):
            b = owner[np.abs(out[k])]
            out[k] = np.where(im[b] == 0, -b, b)
    return out


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


def oracle_step(s, order, contacts=None, prev=None, warm=0, prediction=None, segments=None,
                pairs=None):
    """Full step on the oracle from the snapshot state, visiting constraints in
    `order`.  With `segments` (a list of orders, one per independent tile) each
    segment is solved on its own, stopping at the residual tolerance of its
    own rows.  prev: (keys int2[m], totals float4[m], points float4[m][4]) of
    the previous solve, keyed by pair entry; pairs: the step's pair list
    (body pairs, then compound child pairs) that `contacts` index, which
    gives each contact its key;
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
    ckeys = None
    if pairs is not None:
        ckeys = np.asarray(pairs, np.int32)[np.asarray(contacts["pair_index"], np.int64)]
    feats, cons = O.contact_features(contacts, cons, st["pos"], lv, av, sh["inv_mass"], keys, imp,
                                     pts, apply_previous=(warm != 0), contact_keys=ckeys)
    if warm == 2:
        cons = O.warm_from_prediction(cons, prediction, c.friction)
    if segments is None:
        cons, lv2, av2, it, res = O.solve(cons, order, lv, av, st["pos"], sh["inv_mass"], invI,
                                          c.solver_iterations, c.solver_iterations, c.friction,
                                          c.residual_tolerance)
    else:
        cons, lv2, av2, it, res = O.solve_segments(cons, segments, lv, av, st["pos"],
                                                   sh["inv_mass"], invI, c.solver_iterations,
                                                   c.solver_iterations, c.friction,
                                                   c.residual_tolerance)
        it, res = list(it), list(res)
    pos, quat, av3 = O.integrate(st["pos"], st["quat"], lv2, av2, sh["inv_mass"], c.dt,
                                 c.angular_damping, c.max_angular_motion)
    return dict(pos=pos, quat=quat, linvel=lv2, angvel=av3, cons=cons, iters=it, res=res,
                features=feats)
