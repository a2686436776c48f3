"""Snapshot files ("MIRB" v2, mi_rb_save_state's format) are checked before
they replace a world's arrays: a valid file passes mi_rb_snapshot_check
(host only), and truncated files, inconsistent array lengths and offsets or
indices outside the arrays are refused with MI_RB_INVALID_ARG, capacities
beyond the configuration with MI_RB_CAPACITY."""
import ctypes
import struct

import numpy as np
import pytest

import oracle as O
from experiments_amd import _lib as L
from experiments_amd import scenes
from experiments_amd.world import reference_config
from test_abi import hull

CHILD = np.dtype([("pos", "<f4", 4), ("orn", "<f4", 4), ("shape", "<i4"), ("pad", "<i4", 3)])


def arrays():
    """A cube and a ground box, three bodies (two cubes, the static ground)."""
    a = {k: [] for k in ("hulls", "verts", "planes", "edges", "face_index", "indices", "hext")}
    H = []
    for pts in (scenes.cube_points(0.5), scenes.box_points(10.0, 0.5, 10.0)):
        v, pl, fi, ind, ed = hull(pts)
        h = np.zeros(1, O.HULL_DTYPE)
        h["face_offset"], h["num_faces"] = len(a["planes"]), len(pl)
        h["vertex_offset"], h["num_vertices"] = len(a["verts"]), len(v)
        h["edge_offset"], h["num_edges"] = len(a["edges"]), len(ed)
        h["index_offset"], h["num_indices"] = len(a["indices"]), len(ind)
        h["center"][0, :3] = v.mean(0)
        h["center"][0, 3] = 0.4
        h["aabb_min"][0, :3], h["aabb_max"][0, :3] = v.min(0), v.max(0)
        H.append(h)
        base = len(a["indices"])
        a["face_index"] += [[base + f[0], f[1]] for f in fi]
        a["indices"] += ind.tolist()
        a["verts"] += [list(x) + [0] for x in v]
        a["planes"] += pl.tolist()
        a["edges"] += [list(x) + [0] for x in ed]
        a["hext"].append(float(np.linalg.norm((v.max(0) - v.min(0)) / 2)))
    f4 = lambda x: np.array(x, np.float32).reshape(-1, 4)
    return {
        "hulls": np.concatenate(H), "verts": f4(a["verts"]), "planes": f4(a["planes"]),
        "edges": f4(a["edges"]), "face_index": np.array(a["face_index"], np.int32),
        "indices": np.array(a["indices"], np.int32), "hext": np.array(a["hext"], np.float32),
        "pos": f4([[0, 1, 0, 0], [0, 2.1, 0, 0], [0, -0.5, 0, 0]]),
        "quat": f4([[0, 0, 0, 1]] * 3), "linvel": f4([[0] * 4] * 3), "angvel": f4([[0] * 4] * 3),
        "inv_inertia": f4([[6, 6, 6, 0], [6, 6, 6, 0], [0, 0, 0, 0]]),
        "inv_mass": np.array([1, 1, 0], np.float32), "shape": np.array([0, 0, 1], np.int32),
        "scene": np.zeros(3, np.int32), "children": np.zeros(0, CHILD)}


ORDER = ["hulls", "verts", "planes", "edges", "face_index", "indices", "hext", "pos", "quat",
         "linvel", "angvel", "inv_inertia", "inv_mass", "shape", "scene", "children"]


def write(path, a):
    b = bytearray(b"MIRB\x02\x00\x00\x00")
    for k in ORDER:
        x = np.ascontiguousarray(a[k])
        b += struct.pack("<Q", len(x)) + x.tobytes()
    path.write_bytes(bytes(b))
    return str(path).encode()


def check(path, **cfg):
    c = reference_config(max_bodies=cfg.get("max_bodies", 64), max_shapes=16,
                         max_hull_vertices=1024)
    counts = (ctypes.c_int * 4)()
    return L.lib().mi_rb_snapshot_check(path, ctypes.byref(c), counts), list(counts)


def test_valid_snapshot_passes(tmp_path):
    st, counts = check(write(tmp_path / "ok.mirb", arrays()))
    assert st == L.OK and counts == [3, 2, 16, 0]


def test_truncated_and_padded_files_are_refused(tmp_path):
    p = tmp_path / "ok.mirb"
    write(p, arrays())
    raw = p.read_bytes()
    for cut in (4, 8, 100, len(raw) // 2, len(raw) - 1):
        q = tmp_path / ("cut%d.mirb" % cut)
        q.write_bytes(raw[:cut])
        assert check(str(q).encode())[0] == L.INVALID_ARG, cut
    q = tmp_path / "extra.mirb"
    q.write_bytes(raw + b"\0" * 8)
    assert check(str(q).encode())[0] == L.INVALID_ARG


@pytest.mark.parametrize("corrupt", [
    "shape_out_of_range", "negative_shape", "vertex_offset", "face_offset", "edge_offset",
    "face_index", "vertex_index", "lengths_differ", "scene_order", "nan_position",
    "negative_mass", "child_table"])
def test_inconsistent_snapshots_are_refused(tmp_path, corrupt):
    a = arrays()
    if corrupt == "shape_out_of_range":
        a["shape"][1] = 2
    elif corrupt == "negative_shape":
        a["shape"][0] = -1
    elif corrupt == "vertex_offset":
        a["hulls"]["vertex_offset"][1] = len(a["verts"]) - 4
    elif corrupt == "face_offset":
        a["hulls"]["num_faces"][1] = len(a["planes"])
    elif corrupt == "edge_offset":
        a["hulls"]["edge_offset"][0] = 100
    elif corrupt == "face_index":
        a["face_index"][3, 0] = len(a["indices"]) - 1
    elif corrupt == "vertex_index":
        a["indices"][5] = 8
    elif corrupt == "lengths_differ":
        a["linvel"] = a["linvel"][:2]
    elif corrupt == "scene_order":
        a["scene"][:] = [1, 0, 1]
    elif corrupt == "nan_position":
        a["pos"][1, 1] = np.nan
    elif corrupt == "negative_mass":
        a["inv_mass"][0] = -1
    elif corrupt == "child_table":
        # a compound whose children lie beyond the child table
        h = a["hulls"][:1].copy()
        h["num_vertices"], h["num_faces"], h["num_indices"] = 0, 0, 0
        h["edge_offset"], h["num_edges"] = 0, 2
        a["hulls"] = np.concatenate([a["hulls"], h])
        a["hext"] = np.append(a["hext"], 1.0).astype(np.float32)
        a["children"] = np.zeros(1, CHILD)
    assert check(write(tmp_path / "bad.mirb", a))[0] == L.INVALID_ARG


def test_capacities_are_enforced(tmp_path):
    st, _ = check(write(tmp_path / "ok.mirb", arrays()), max_bodies=2)
    assert st == L.CAPACITY
