"""Bullet's serialized scene files (".bullet") through the library's C++
reader (include/mi_rb.h: mi_rb_bullet_read, mi_rb_load_bullet).

The reference loads such files in its demo's createScene with bParse and
btBulletDataExtractor (opencl/gpu_rigidbody_pipeline2/BulletDataExtractor.cpp:
454-520, 594-1010); load_bullet registers the same shapes and bodies into a
world.
"""
import ctypes

import numpy as np

from . import _lib as L

SHAPE_BOX, SHAPE_CONVEX_HULL, SHAPE_SPHERE = 0, 4, 8
SHAPE_TRIANGLE_MESH, SHAPE_STATIC_PLANE, SHAPE_COMPOUND = 21, 28, 31


def read_bullet(path):
    """Parse a float .bullet file: (shapes, bodies, info).  shapes and bodies
    are lists of dicts; hull points are unscaled."""
    cnt = (ctypes.c_int * 4)()
    p = str(path).encode()
    L.call("mi_rb_bullet_read", p, None, 0, None, 0, None, 0, cnt)
    ns, nb, npts = cnt[0], cnt[1], cnt[2]
    sh = (L.BulletShape * max(ns, 1))()
    bo = (L.BulletBody * max(nb, 1))()
    pts = np.zeros((max(npts, 1), 3), np.float32)
    L.call("mi_rb_bullet_read", p, sh, ns, bo, nb, pts.ctypes.data_as(ctypes.c_void_p), npts, cnt)
    shapes = []
    for i in range(ns):
        s = sh[i]
        shapes.append({"type": s.type, "local_scaling": list(s.local_scaling),
                       "implicit_dims": list(s.implicit_dims), "margin": s.margin,
                       "plane_normal": list(s.plane_normal), "plane_constant": s.plane_constant,
                       "points": pts[s.first_point:s.first_point + s.num_points].copy()})
    bodies = []
    for i in range(nb):
        b = bo[i]
        bodies.append({"shape": b.shape, "inverse_mass": b.inverse_mass, "basis": list(b.basis),
                       "origin": list(b.origin), "quat": list(b.quat), "linvel": list(b.linvel),
                       "angvel": list(b.angvel), "friction": b.friction,
                       "restitution": b.restitution})
    info = {"pointer_size": cnt[3] & 0xff, "big_endian": bool(cnt[3] & 0x100), "points": npts}
    return shapes, bodies, info


def load_bullet(world, path, keep_static=False):
    """Register a .bullet file's scene into a world (before upload); returns
    the summary counts."""
    s = L.BulletSummary()
    L.call("mi_rb_load_bullet", world.h, str(path).encode(), 1 if keep_static else 0,
           ctypes.byref(s))
    return {k: getattr(s, k) for k, _ in L.BulletSummary._fields_}
