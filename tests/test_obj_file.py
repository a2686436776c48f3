"""Wavefront .obj meshes: the oracle reader (oracle/obj_file.py) pinned by
known answers read off the reference's own .obj files (tests/golden: the
demo's plane.obj ground, test.obj with every face-token form, and
cornell_box.obj with a quad) and by the geometry of the plane; the
library's reader (mi_rb_obj_read) equals the oracle on all three."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd.wavefront import read_obj
from oracle.obj_file import read_obj as oracle_read_obj, read_obj_text

GOLDEN = Path(__file__).parent / "golden"
FILES = ["plane.obj", "test.obj", "cornell_box.obj"]


def test_face_token_forms_and_indices():
    v, f = oracle_read_obj(GOLDEN / "test.obj")
    assert len(v) == 17
    # "f 1/1/1 3/2/2 4/3/3", "f 4//1 2//2 1//3", "f 5/1 6/2 8/3"
    assert np.array_equal(f[:3], [[0, 2, 3], [3, 1, 0], [4, 5, 7]])
    # negative indices count back; a quad keeps its first three vertices
    v, f = read_obj_text("v 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nf -4 -3 -2 -1\nvn 0 1 0\n")
    assert np.array_equal(f, [[0, 1, 2]]) and len(v) == 4
    with pytest.raises(ValueError):
        read_obj_text("v 0 0 0\nv 1 0 0\nv 1 0 1\nf 0 1 2\n")


def test_plane_geometry():
    v, f = oracle_read_obj(GOLDEN / "plane.obj")
    assert len(v) == 441 and len(f) == 800          # a 21 x 21 grid, two triangles per cell
    assert np.abs(v[:, 1]).max() < 1e-5             # flat: y = 0 up to the file's rounding
    t = v[f].astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    area = 0.5 * np.linalg.norm(n, axis=1)
    span = (v[:, 0].max() - v[:, 0].min()) * (v[:, 2].max() - v[:, 2].min())
    assert abs(area.sum() - span) < 1e-4 * span      # the triangles tile the plane
    assert np.all(np.abs(n[:, 1]) > (1 - 1e-9) * np.linalg.norm(n, axis=1))


def test_cornell_box_quad_keeps_first_triangle():
    _, f = oracle_read_obj(GOLDEN / "cornell_box.obj")
    assert len(f) == 31


@pytest.mark.parametrize("name", FILES)
def test_library_reader_equals_oracle(name):
    v, f = read_obj(GOLDEN / name)
    rv, rf = oracle_read_obj(GOLDEN / name)
    assert np.array_equal(v, rv) and np.array_equal(f, rf)


def test_reader_errors(tmp_path):
    nv, nt = ctypes.c_int(), ctypes.c_int()
    st = L.lib().mi_rb_obj_read(str(tmp_path / "missing.obj").encode(), None, 0, None, 0,
                                ctypes.byref(nv), ctypes.byref(nt))
    assert st == L.INVALID_ARG
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")       # index past the vertices
    st = L.lib().mi_rb_obj_read(str(bad).encode(), None, 0, None, 0, ctypes.byref(nv),
                                ctypes.byref(nt))
    assert st == L.INVALID_ARG
    xyz = np.zeros((1, 3), np.float32)
    tri = np.zeros((1, 3), np.int32)
    st = L.lib().mi_rb_obj_read(str(GOLDEN / "test.obj").encode(),
                                xyz.ctypes.data_as(ctypes.c_void_p), 1,
                                tri.ctypes.data_as(ctypes.c_void_p), 1, ctypes.byref(nv),
                                ctypes.byref(nt))
    assert st == L.CAPACITY and nv.value == 17
