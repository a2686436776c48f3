"""CPU oracle for the rigid-body step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker.  The product (experiments_amd)
never imports it and never falls back to it.

ctypes bindings over oracle/_build/librb_oracle.so, built from rb_oracle.c by
oracle/Makefile (or experiments_amd.build.build_oracle()).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librb_oracle.so")

_lib = None

HULL_DTYPE = np.dtype([
    ("face_offset", "<i4"), ("num_faces", "<i4"), ("vertex_offset", "<i4"),
    ("num_vertices", "<i4"), ("edge_offset", "<i4"), ("num_edges", "<i4"),
    ("index_offset", "<i4"), ("num_indices", "<i4"),
    ("center", "<f4", 4), ("aabb_min", "<f4", 4), ("aabb_max", "<f4", 4)])
assert HULL_DTYPE.itemsize == 80

CONTACT_DTYPE = np.dtype([
    ("pos", "<f4", (4, 4)), ("normal", "<f4", 4), ("coeffs", "<u4"),
    ("pair_index", "<i4"), ("body_a", "<i4"), ("body_b", "<i4")])
assert CONTACT_DTYPE.itemsize == 96

CONSTRAINT_DTYPE = np.dtype([
    ("linear", "<f4", 4), ("pos", "<f4", (4, 4)), ("center", "<f4", 4),
    ("jac_inv", "<f4", 4), ("b", "<f4", 4), ("lambda", "<f4", 4),
    ("f_jac_inv", "<f4", 2), ("f_lambda", "<f4", 2),
    ("body_a", "<u4"), ("body_b", "<u4"), ("batch", "<u4"), ("pad", "<u4")])
assert CONSTRAINT_DTYPE.itemsize == 176


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_exclusive_scan.restype = ctypes.c_uint32
        for name in ("orc_find_pairs_grid", "orc_find_pairs_brute", "orc_sap_pairs", "orc_sat",
                     "orc_clip_hull_hull", "orc_clip_face", "orc_reduce_manifold",
                     "orc_reduce_manifold_global",
                     "orc_narrowphase", "orc_color_contacts"):
            getattr(_lib, name).restype = ctypes.c_int
        _lib.orc_sweep_residual.restype = ctypes.c_float
        _lib.orc_sweep_residual.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.orc_bf16_to_f32.restype = ctypes.c_float
        _lib.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
        _lib.orc_f32_to_bf16.restype = ctypes.c_uint16
        _lib.orc_f32_to_bf16.argtypes = [ctypes.c_float]
    return _lib


def _p(a):
    """Pointer to a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _f(x):
    return ctypes.c_float(float(x))


def f4(a, n=None):
    """(n,3)/(n,4) float array -> contiguous (n,4) float32 with w = 0."""
    a = np.asarray(a, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    out = np.zeros((a.shape[0], 4), np.float32)
    out[:, :a.shape[1]] = a
    return out


def plane_space(n):
    """The friction tangents (t0, t1) of normal n (orc_plane_space)."""
    nn = np.ascontiguousarray(n[:3], np.float32)
    t0 = np.zeros(4, np.float32)
    t1 = np.zeros(4, np.float32)
    lib().orc_plane_space(_p(nn), _p(t0), _p(t1))
    return t0[:3], t1[:3]


# ---------------------------------------------------------------- registration
def register_inertia(mass, local_min, local_max, pos, quat, margin=0.01):
    """Inverse local inertia diagonal (x, y, z, 0) a body gets at registration
    (orc_register_inertia)."""
    out = np.zeros(4, np.float32)
    mn = np.ascontiguousarray(local_min[:3], np.float32)
    mx = np.ascontiguousarray(local_max[:3], np.float32)
    p = np.ascontiguousarray(pos[:3], np.float32)
    q = np.ascontiguousarray(quat[:4], np.float32)
    lib().orc_register_inertia(_f(mass), _p(mn), _p(mx), _p(p), _p(q), _f(margin), _p(out))
    return out


# ----------------------------------------------------------------- primitives
def radix_sort_kv(keys, values, sort_bits=32):
    kv = np.stack([np.asarray(keys, np.uint32), np.asarray(values, np.uint32)], 1).copy()
    lib().orc_radix_sort_kv(_p(kv), ctypes.c_int(len(kv)), ctypes.c_int(sort_bits))
    return kv[:, 0].copy(), kv[:, 1].copy()


def exclusive_scan(x):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.zeros_like(x)
    total = lib().orc_exclusive_scan(_p(x), _p(out), ctypes.c_int(len(x)))
    return out, int(total)


def bound_search(keys, n_dst, mode):
    kv = np.stack([np.asarray(keys, np.uint32), np.arange(len(keys), dtype=np.uint32)], 1).copy()
    dst = np.zeros(n_dst, np.uint32)
    lib().orc_bound_search(_p(kv), ctypes.c_int(len(kv)), _p(dst), ctypes.c_int(n_dst),
                           ctypes.c_int({"lower": 0, "upper": 1, "count": 2}[mode]))
    return dst


# ------------------------------------------------------------- broadphase
def world_aabbs(pos, quat, shape, hulls):
    pos, quat = f4(pos), np.ascontiguousarray(quat, np.float32)
    n = len(pos)
    mn = np.zeros((n, 4), np.float32)
    mx = np.zeros((n, 4), np.float32)
    lib().orc_world_aabbs(ctypes.c_int(n), _p(pos), _p(quat),
                          _p(np.ascontiguousarray(shape, np.int32)),
                          _p(np.ascontiguousarray(hulls)), _p(mn), _p(mx))
    return mn, mx


def calc_hash(aabb_min, aabb_max, inv_cell, dims, cell_bits=0, scene=None, is_large=None):
    mn, mx = f4(aabb_min), f4(aabb_max)
    n = len(mn)
    kv = np.zeros((n, 2), np.uint32)
    d = np.asarray(dims, np.int32)
    sc = None if scene is None else np.ascontiguousarray(scene, np.int32)
    lg = None if is_large is None else np.ascontiguousarray(is_large, np.int32)
    lib().orc_calc_hash(ctypes.c_int(n), _p(mn), _p(mx), _p(sc), _p(lg), _f(inv_cell), _p(d),
                        ctypes.c_int(cell_bits), _p(kv))
    return kv


def find_pairs_grid(sorted_kv, aabb_min, aabb_max, inv_cell, dims, cell_bits=0, scene=None,
                    is_large=None, scene_start=None, scene_count=None, max_pairs=1 << 22):
    mn, mx = f4(aabb_min), f4(aabb_max)
    kv = np.ascontiguousarray(sorted_kv, np.uint32)
    out = np.zeros((max_pairs, 2), np.int32)
    conv = lambda a: None if a is None else np.ascontiguousarray(a, np.int32)
    n = lib().orc_find_pairs_grid(ctypes.c_int(len(mn)), _p(kv), _p(mn), _p(mx), _p(conv(scene)),
                                  _p(conv(is_large)), _p(conv(scene_start)),
                                  _p(conv(scene_count)), _f(inv_cell),
                                  _p(np.asarray(dims, np.int32)), ctypes.c_int(cell_bits),
                                  ctypes.c_int(max_pairs), _p(out))
    return out[:min(n, max_pairs)].copy()


def find_pairs_brute(aabb_min, aabb_max, scene=None, max_pairs=1 << 22):
    mn, mx = f4(aabb_min), f4(aabb_max)
    out = np.zeros((max_pairs, 2), np.int32)
    sc = None if scene is None else np.ascontiguousarray(scene, np.int32)
    n = lib().orc_find_pairs_brute(ctypes.c_int(len(mn)), _p(mn), _p(mx), _p(sc),
                                   ctypes.c_int(max_pairs), _p(out))
    return out[:min(n, max_pairs)].copy()


def sap_pairs(aabb_min, aabb_max, axis=0, max_pairs=1 << 22):
    """Sweep-and-prune pair set (orc_sap_pairs): (x > y) sorted by (x, y)."""
    mn, mx = f4(aabb_min), f4(aabb_max)
    out = np.zeros((max_pairs, 2), np.int32)
    n = lib().orc_sap_pairs(ctypes.c_int(len(mn)), _p(mn), _p(mx), ctypes.c_int(axis),
                            ctypes.c_int(max_pairs), _p(out))
    return out[:min(n, max_pairs)].copy()


def setup_bodies(linvel, angvel, quat, inv_mass, inv_inertia_local, gravity, dt):
    linvel = f4(linvel)
    n = len(linvel)
    world = np.zeros((n, 3, 4), np.float32)
    lib().orc_setup_bodies(ctypes.c_int(n), _p(linvel), _p(f4(angvel)),
                           _p(np.ascontiguousarray(quat, np.float32)),
                           _p(np.ascontiguousarray(inv_mass, np.float32)),
                           _p(f4(inv_inertia_local)), _p(world),
                           _p(np.asarray(gravity, np.float32)), _f(dt))
    return linvel, world


# ------------------------------------------------------------ narrowphase
class Shapes:
    """Flat hull arrays in the library's layout (see rb_oracle.h)."""

    def __init__(self, hulls, verts, planes, face_index, indices, edges):
        self.hulls = np.ascontiguousarray(hulls)
        self.verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 4)
        self.planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
        self.face_index = np.ascontiguousarray(face_index, np.int32).reshape(-1, 2)
        self.indices = np.ascontiguousarray(indices, np.int32)
        self.edges = np.ascontiguousarray(edges, np.float32).reshape(-1, 4)


def sat(shapes, sa, sb, posA, ornA, posB, ornB, edge_reject=False):
    sep = np.zeros(4, np.float32)
    H = shapes.hulls
    r = lib().orc_sat(_p(H[sa:sa + 1]), _p(H[sb:sb + 1]), _p(shapes.verts), _p(shapes.planes),
                      _p(shapes.edges), _p(f4(posA)[0]), _p(np.asarray(ornA, np.float32).copy()),
                      _p(f4(posB)[0]), _p(np.asarray(ornB, np.float32).copy()),
                      ctypes.c_int(1 if edge_reject else 0), _p(sep))
    return r, sep


def clip_face(poly, normal, eq):
    poly = f4(poly)
    out = np.zeros((len(poly) * 2 + 2, 4), np.float32)
    n = lib().orc_clip_face(_p(poly), ctypes.c_int(len(poly)), _p(f4(normal)[0]), _f(eq), _p(out))
    return out[:n].copy()


def reduce_manifold(points, normal, split_path=False):
    """Indices of the kept points: extractManifoldSequential, or with
    split_path extractManifoldSequentialGlobal (indices may repeat)."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
    idx = np.full(4, -1, np.int32)
    fn = lib().orc_reduce_manifold_global if split_path else lib().orc_reduce_manifold
    n = fn(_p(pts), ctypes.c_int(len(pts)), _p(np.asarray(normal, np.float32).copy()), _p(idx))
    return idx[:n].copy()


CHILD_DTYPE = np.dtype([("pos", "<f4", 4), ("orn", "<f4", 4), ("shape", "<i4"), ("pad", "<i4", 3)])
assert CHILD_DTYPE.itemsize == 48


def compound_proxies(proxy, children, pos, quat):
    """pos/quat of the n bodies extended by the world transforms of the
    compound children proxy[k] = (body, child index)."""
    n, P = len(pos), len(proxy)
    pos4, quat4 = np.zeros((n + P, 4), np.float32), np.zeros((n + P, 4), np.float32)
    pos4[:n] = f4(pos)
    quat4[:n] = np.ascontiguousarray(quat, np.float32)
    if P:
        pr = np.ascontiguousarray(proxy, np.int32).reshape(-1, 2)
        ch = np.ascontiguousarray(children, CHILD_DTYPE)
        lib().orc_compound_proxies(ctypes.c_int(P), ctypes.c_int(n), _p(pr), _p(ch), _p(pos4),
                                   _p(quat4))
    return pos4, quat4


def narrowphase(pairs, pos, quat, inv_mass, shape, shapes, friction=0.7, max_dist=0.02,
                sat_mode=0, reduction_mode=0):
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    P = len(pairs)
    slots = np.zeros(max(P, 1), CONTACT_DTYPE)
    has = np.zeros(max(P, 1), np.int32)
    lib().orc_narrowphase(ctypes.c_int(P), _p(pairs), _p(f4(pos)),
                          _p(np.ascontiguousarray(quat, np.float32)),
                          _p(np.ascontiguousarray(inv_mass, np.float32)),
                          _p(np.ascontiguousarray(shape, np.int32)), _p(shapes.hulls),
                          _p(shapes.verts), _p(shapes.planes), _p(shapes.face_index),
                          _p(shapes.indices), _p(shapes.edges), _f(friction), _f(max_dist),
                          ctypes.c_int(sat_mode), ctypes.c_int(reduction_mode), _p(slots),
                          _p(has))
    return slots[:P], has[:P]


# ----------------------------------------------------------------- solver
def contact_to_constraint(contacts, pos, linvel, angvel, inv_mass, inv_inertia_world, dt,
                          drift=0.005, coeff=0.2, friction=0.7):
    contacts = np.ascontiguousarray(contacts, CONTACT_DTYPE)
    out = np.zeros(max(len(contacts), 1), CONSTRAINT_DTYPE)
    lib().orc_contact_to_constraint(
        ctypes.c_int(len(contacts)), _p(contacts), _p(f4(pos)), _p(f4(linvel)), _p(f4(angvel)),
        _p(np.ascontiguousarray(inv_mass, np.float32)),
        _p(np.ascontiguousarray(inv_inertia_world, np.float32)), _f(dt), _f(drift), _f(coeff),
        _f(friction), _p(out))
    return out[:len(contacts)]


def contact_features(contacts, cons, pos, linvel, angvel, inv_mass, prev_keys=None,
                     prev_imp=None, prev_pts=None, apply_previous=False, contact_keys=None):
    """16 warm-start features per contact; with apply_previous the previous
    impulses found for the contact's key (contact_keys: int2 per contact, the
    contact's pair entry; None: its body pair) are written into a copy of
    cons (per point: the nearest previous point's)."""
    contacts = np.ascontiguousarray(contacts, CONTACT_DTYPE)
    cons = np.array(cons, CONSTRAINT_DTYPE, copy=True)
    nc = len(contacts)
    feats = np.zeros((max(nc, 1), 16), np.float32)
    pk = np.ascontiguousarray(prev_keys if prev_keys is not None else np.zeros((0, 2)), np.int32)
    pi = np.ascontiguousarray(prev_imp if prev_imp is not None else np.zeros((0, 4)), np.float32)
    pp = np.ascontiguousarray(prev_pts if prev_pts is not None else np.zeros((len(pk), 4, 4)),
                              np.float32)
    if len(pk):
        pp = pp.reshape(len(pk), 4, 4)
    lib().orc_contact_features(
        ctypes.c_int(nc), _p(contacts), _p(cons), _p(f4(pos)), _p(f4(linvel)), _p(f4(angvel)),
        _p(np.ascontiguousarray(inv_mass, np.float32)), ctypes.c_int(len(pk)), _p(pk), _p(pi),
        _p(pp), ctypes.c_int(1 if apply_previous else 0),
        _p(np.ascontiguousarray(contact_keys, np.int32)) if contact_keys is not None else None,
        _p(feats))
    return feats[:nc], cons


def warm_from_prediction(cons, y, friction):
    cons = np.array(cons, CONSTRAINT_DTYPE, copy=True)
    y = np.ascontiguousarray(y, np.float32).reshape(-1, 3)
    lib().orc_warm_from_prediction(ctypes.c_int(len(cons)), _p(cons), _p(y), _f(friction))
    return cons


def color_contacts(cons, inv_mass):
    cons = np.ascontiguousarray(cons, CONSTRAINT_DTYPE)
    col = np.zeros(max(len(cons), 1), np.int32)
    im = np.ascontiguousarray(inv_mass, np.float32)
    nc = lib().orc_color_contacts(ctypes.c_int(len(cons)), _p(cons), _p(im), ctypes.c_int(len(im)),
                                  _p(col))
    return col[:len(cons)], nc


def solve(cons, order, linvel, angvel, pos, inv_mass, inv_inertia_world, normal_iters,
          friction_iters, friction=0.7, tol=0.0):
    cons = np.ascontiguousarray(cons, CONSTRAINT_DTYPE).copy()
    lv, av = f4(linvel), f4(angvel)
    iters = np.zeros(2, np.int32)
    res = np.zeros(2, np.float32)
    order = None if order is None else np.ascontiguousarray(order, np.int32)
    # with an order, only the rows it lists are visited (a tile's segment)
    n = len(cons) if order is None else len(order)
    lib().orc_solve(ctypes.c_int(n), _p(cons), _p(order), _p(lv), _p(av), _p(f4(pos)),
                    _p(np.ascontiguousarray(inv_mass, np.float32)),
                    _p(np.ascontiguousarray(inv_inertia_world, np.float32)),
                    ctypes.c_int(normal_iters), ctypes.c_int(friction_iters), _f(friction),
                    _f(tol), _p(iters), _p(res))
    return cons, lv, av, iters, res


def solve_segments(cons, segments, linvel, angvel, pos, inv_mass, inv_inertia_world,
                   normal_iters, friction_iters, friction=0.7, tol=0.0):
    """solve() on each segment (a list of row orders, one per independent
    tile) in turn, in one call; iterations and residuals per segment."""
    cons = np.ascontiguousarray(cons, CONSTRAINT_DTYPE).copy()
    lv, av = f4(linvel), f4(angvel)
    segs = [np.asarray(x, np.int32) for x in segments]
    off = np.zeros(len(segs) + 1, np.int32)
    off[1:] = np.cumsum([len(x) for x in segs])
    order = np.ascontiguousarray(np.concatenate(segs) if segs else np.zeros(1, np.int32), np.int32)
    iters = np.zeros((max(len(segs), 1), 2), np.int32)
    res = np.zeros((max(len(segs), 1), 2), np.float32)
    lib().orc_solve_segments(ctypes.c_int(len(segs)), _p(off), _p(cons), _p(order), _p(lv),
                             _p(av), _p(f4(pos)), _p(np.ascontiguousarray(inv_mass, np.float32)),
                             _p(np.ascontiguousarray(inv_inertia_world, np.float32)),
                             ctypes.c_int(normal_iters), ctypes.c_int(friction_iters),
                             _f(friction), _f(tol), _p(iters), _p(res))
    return cons, lv, av, iters[:len(segs)], res[:len(segs)]


def integrate(pos, quat, linvel, angvel, inv_mass, dt, damping=0.99,
              max_ang=0.25 * 3.14159254):
    pos = f4(pos)
    quat = np.ascontiguousarray(quat, np.float32).copy()
    av = f4(angvel)
    lib().orc_integrate(ctypes.c_int(len(pos)), _p(pos), _p(quat), _p(f4(linvel)), _p(av),
                        _p(np.ascontiguousarray(inv_mass, np.float32)), _f(dt), _f(damping),
                        _f(max_ang))
    return pos, quat, av


# -------------------------------------------------------------------- MLP
N_W = 64 * 16 + 64 * 64 + 3 * 64
N_B = 64 + 64 + 3


def mlp_forward(x, w_bf16, bias):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
    n = len(x)
    h1 = np.zeros((n, 64), np.float32)
    h2 = np.zeros((n, 64), np.float32)
    y = np.zeros((n, 3), np.float32)
    lib().orc_mlp_forward(ctypes.c_int(n), _p(x), _p(np.ascontiguousarray(w_bf16, np.uint16)),
                          _p(np.ascontiguousarray(bias, np.float32)), _p(h1), _p(h2), _p(y))
    return h1, h2, y


def mlp_backward(x, w_bf16, bias, target):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
    h1, h2, y = mlp_forward(x, w_bf16, bias)
    gw = np.zeros(N_W, np.float32)
    gb = np.zeros(N_B, np.float32)
    loss = np.zeros(1, np.float32)
    lib().orc_mlp_backward(ctypes.c_int(len(x)), _p(x), _p(np.ascontiguousarray(w_bf16, np.uint16)),
                           _p(np.ascontiguousarray(bias, np.float32)), _p(h1), _p(h2), _p(y),
                           _p(np.ascontiguousarray(target, np.float32)), _p(gw), _p(gb),
                           _p(loss))
    return gw, gb, float(loss[0]), y
