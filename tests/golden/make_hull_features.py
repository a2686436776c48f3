"""Write tests/golden/hull_features.npz: the reference's own polyhedral
features (btConvexHullComputer + btConvexUtility::initializePolyhedralFeatures,
opencl/gpu_rigidbody_pipeline/btConvexUtility.cpp:31-510) for a few hull
inputs, in the reference's order.

Run in the development container only, after `make -C oracle ref` has built
oracle/_ref/libref_hull.so from the reference checkout:

    python tests/golden/make_hull_features.py

Inputs: the unit cube, and shapes 0-2 of the random 8- and 32-vertex hull
pools of the 65,536-hull scene (experiments_amd.scenes.hull_pool).  Only
data is written: inputs and the reference's outputs.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def ref_features(lib, pts):
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    pts = np.ascontiguousarray(pts, np.float32)
    v = np.zeros((64, 4), np.float32)
    pl = np.zeros((128, 4), np.float32)
    fi = np.zeros((128, 2), np.int32)
    ind = np.zeros(512, np.int32)
    ed = np.zeros((256, 4), np.float32)
    cnt = np.zeros(4, np.int32)
    one = np.ones(3, np.float32)
    r = lib.ref_hull_features(P(pts), len(pts), P(one), P(v), 64, P(pl), P(fi), 128, P(ind), 512,
                              P(ed), 256, P(cnt))
    assert r == 0
    nv, nf, ni, ne = (int(x) for x in cnt)
    return v[:nv, :3].copy(), pl[:nf].copy(), fi[:nf].copy(), ind[:ni].copy(), ed[:ne, :3].copy()


def main():
    from experiments_amd import scenes
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_hull.so"))
    p8, p32 = scenes.hull_pool()
    inputs = [scenes.cube_points(0.5)] + p8[:3] + p32[:3]
    out = {"n": len(inputs)}
    for k, pts in enumerate(inputs):
        v, pl, fi, ind, ed = ref_features(lib, pts)
        out["points_%d" % k] = np.asarray(pts, np.float32)
        out["verts_%d" % k] = v
        out["planes_%d" % k] = pl
        out["face_index_%d" % k] = fi
        out["indices_%d" % k] = ind
        out["edges_%d" % k] = ed
    np.savez_compressed(os.path.join(HERE, "hull_features.npz"), **out)
    print("wrote %d hulls" % len(inputs))


if __name__ == "__main__":
    main()
