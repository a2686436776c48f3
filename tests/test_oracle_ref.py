"""The oracle's clipping (orc_clip_hull_hull) against the reference's own CPU
emulation of the clip kernel (opencl/gpu_rigidbody_pipeline2/satClipHullContacts.cpp,
built by `make -C oracle ref` into oracle/_ref/) on random hull pairs."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from experiments_amd import scenes
from test_abi import hull

REF = os.path.join(os.path.dirname(O.__file__), "_ref", "libref_clip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")


def table(point_sets):
    H = np.zeros(len(point_sets), O.HULL_DTYPE)
    V, P, FI, IND, E = [], [], [], [], []
    for k, pts in enumerate(point_sets):
        v, pl, fi, ind, ed = hull(pts)
        H[k]["face_offset"], H[k]["num_faces"] = len(P), len(pl)
        H[k]["vertex_offset"], H[k]["num_vertices"] = len(V), len(v)
        H[k]["edge_offset"], H[k]["num_edges"] = len(E), len(ed)
        H[k]["center"][:3] = v.mean(0)
        H[k]["aabb_min"][:3], H[k]["aabb_max"][:3] = v.min(0), v.max(0)
        base = len(IND)
        for f in range(len(pl)):
            FI.append([base + fi[f, 0], fi[f, 1]])
        IND += ind.tolist()
        V += [list(x) + [0] for x in v]
        P += pl.tolist()
        E += [list(x) + [0] for x in ed]
    return O.Shapes(H, np.array(V, np.float32), np.array(P, np.float32), np.array(FI),
                    np.array(IND), np.array(E, np.float32))


def test_clip_matches_reference_cpu_emulation():
    lib = ctypes.CDLL(REF)
    lib.ref_clip_hull_hull.restype = ctypes.c_int
    rng = scenes.PCG32(11, 1)
    pts = [scenes.cube_points(0.5)] + [scenes.random_hull_points(rng, 8) for _ in range(6)] + \
          [scenes.random_hull_points(rng, 32) for _ in range(6)]
    sh = table(pts)
    P = lambda a: np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)
    r = np.random.default_rng(4)
    checked = exact = 0
    for trial in range(600):
        a, b = r.integers(0, len(pts), 2)
        qa = r.normal(size=4); qa /= np.linalg.norm(qa)
        qb = r.normal(size=4); qb /= np.linalg.norm(qb)
        pa = r.uniform(-0.9, 0.9, 3).astype(np.float32)
        pb = np.zeros(3, np.float32)
        qa, qb = qa.astype(np.float32), qb.astype(np.float32)
        ok, sep = O.sat(sh, a, b, pa, qa, pb, qb)
        if not ok:
            continue
        out_o = np.zeros((64, 4), np.float32)
        out_r = np.zeros((64, 4), np.float32)
        hA, hB = sh.hulls[a:a + 1], sh.hulls[b:b + 1]
        n_o = O.lib().orc_clip_hull_hull(P(sep), P(hA), P(hB), P(sh.verts), P(sh.planes),
                                         P(sh.face_index), P(sh.indices), P(O.f4(pa)[0]), P(qa),
                                         P(O.f4(pb)[0]), P(qb), ctypes.c_float(-1e30),
                                         ctypes.c_float(0.02), P(out_o))
        n_r = lib.ref_clip_hull_hull(P(sep), P(hA), P(hB), P(sh.verts), len(sh.verts),
                                     P(sh.planes), P(sh.face_index), len(sh.planes),
                                     P(sh.indices), len(sh.indices), P(O.f4(pa)[0]), P(qa),
                                     P(O.f4(pb)[0]), P(qb), ctypes.c_float(-1e30),
                                     ctypes.c_float(0.02), P(out_r))
        assert n_o == n_r, trial
        assert np.allclose(out_o[:n_o], out_r[:n_r], rtol=1e-5, atol=2e-6), trial
        checked += 1
        exact += int(np.array_equal(out_o[:n_o], out_r[:n_r]))
    assert checked > 200
    print("checked %d pairs, %d bit-identical" % (checked, exact))
