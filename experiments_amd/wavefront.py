"""Wavefront .obj meshes through the library's reader (mi_rb_obj_read):
the vertices and the triangles (first three vertices of each face, 0-based)
that mi_rb_register_concave_mesh_obj registers."""
import ctypes

import numpy as np

from . import _lib as L


def read_obj(path):
    """(vertices (V, 3) float32, triangles (T, 3) int32) of an .obj file."""
    p = str(path).encode()
    nv, nt = ctypes.c_int(), ctypes.c_int()
    L.call("mi_rb_obj_read", p, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(nt))
    xyz = np.zeros((max(nv.value, 1), 3), np.float32)
    tri = np.zeros((max(nt.value, 1), 3), np.int32)
    L.call("mi_rb_obj_read", p, xyz.ctypes.data_as(ctypes.c_void_p), nv.value,
           tri.ctypes.data_as(ctypes.c_void_p), nt.value, ctypes.byref(nv), ctypes.byref(nt))
    return xyz[:nv.value], tri[:nt.value]
