""".bullet scene files: the Python restatement of the format (oracle) checked
against real files, and the library's C++ reader checked against it.

* le32Bitf.bullet is the reference's own test file (Bullet 2.77, 32-bit,
  little endian): its static body and its triangle mesh, reached through two
  levels of pointers, are known answers (a 2 x 2 x 2 cube: 8 corners, 12
  triangles, surface area 24, volume 8).
* Trimmed Blender files (same container, written by other software) pin the
  64-bit-pointer and big-endian layouts: every struct's members add up to
  the DNA's type length and every chunk holds whole structs.
* The stack_* files hold one scene in all four layouts; the C++ reader and
  the oracle must agree on every field.  No device work: CPU tests.
"""
import os

import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd import bullet as BL
from oracle import bullet_file as B

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STACKS = ["stack_le32.bullet", "stack_le64.bullet", "stack_be32.bullet", "stack_be64.bullet"]


def gold(name):
    return os.path.join(GOLD, name)


@pytest.mark.parametrize("name,ptr,big", [("blend_be32_trim.blend", 4, True),
                                          ("blend_le64_trim.blend", 8, False)])
def test_oracle_layouts_on_blender_files(name, ptr, big):
    f = B.BulletFile(open(gold(name), "rb").read(), magic=b"BLENDER")
    assert (f.ptr_size, f.big) == (ptr, big)
    assert len(f.dna.structs) > 200
    for t, _ in f.dna.structs:
        tn = f.dna.types[t]
        assert f.dna.size(tn, ptr) == f.dna.tlen[t], tn
    data = [c for c in f.chunks if c[0] != b"DNA1" and c[3] > 0]
    assert len(data) > 20
    for c in data:
        assert c[1] == c[4] * f.dna.size(f.struct_name(c), ptr), c[:5]


def test_oracle_reference_fixture_known_answers():
    f, shapes, bodies = B.read_scene(gold("le32Bitf.bullet"))
    assert (f.ptr_size, f.big, f.doubles, f.version) == (4, False, False, 277)
    assert len(f.dna.structs) == 47
    assert [c[0] for c in f.chunks] == [b"RBDY", b"SHAP", b"ARAY", b"ARAY", b"ARAY", b"QBVH",
                                        b"ARAY", b"ARAY", b"DNA1"]
    assert len(shapes) == 1 and shapes[0]["type"] == B.TRIANGLE_MESH
    assert shapes[0]["margin"] == np.float32(0.06)
    (b,) = bodies
    assert b["shape"] == 0 and b["inverse_mass"] == 0.0
    assert b["basis"] == [1, 0, 0, 0, 1, 0, 0, 0, 1] and b["origin"] == [0, 0, 0]
    assert b["friction"] == 0.5 and b["restitution"] == 0.0
    # the mesh: shape -> mesh part -> vertex and index arrays
    sh = [c for c in f.chunks if c[0] == b"SHAP"][0]
    mi = "m_meshInterface."
    assert f.get(sh, 0, mi + "m_numMeshParts") == 1
    part = f.by_ptr[f.get(sh, 0, mi + "m_meshPartsPtr")]
    nt, nv = f.get(part, 0, "m_numTriangles"), f.get(part, 0, "m_numVertices")
    assert (nt, nv) == (12, 36)
    vc = f.by_ptr[f.get(part, 0, "m_vertices3f")]
    ic = f.by_ptr[f.get(part, 0, "m_indices32")]
    V = np.array([[f.get(vc, i, "m_floats[%d]" % k) for k in range(3)] for i in range(nv)])
    idx = np.array([f.get(ic, i, "m_value") for i in range(3 * nt)]).reshape(nt, 3)
    corners = np.unique(np.round(V, 4), axis=0)
    assert corners.shape == (8, 3) and np.allclose(np.abs(corners), 1.0)
    a, b_, c = V[idx[:, 0]], V[idx[:, 1]], V[idx[:, 2]]
    cr = np.cross(b_ - a, c - a)
    assert abs(0.5 * np.linalg.norm(cr, axis=1).sum() - 24.0) < 1e-4
    assert abs(abs(np.einsum("ij,ij->i", a, cr).sum() / 6.0) - 8.0) < 1e-4


def same_scene(cpp, orc):
    shapes, bodies, _ = cpp
    _, oshapes, obodies = orc
    assert len(shapes) == len(oshapes) and len(bodies) == len(obodies)
    for a, b in zip(shapes, oshapes):
        for k in ("type", "local_scaling", "implicit_dims", "margin", "plane_normal",
                  "plane_constant"):
            assert np.array_equal(np.float32(a[k]), np.float32(b[k])), k
        assert np.array_equal(a["points"], b["points"])
    for a, b in zip(bodies, obodies):
        for k in ("shape", "inverse_mass", "basis", "origin", "linvel", "angvel", "friction",
                  "restitution"):
            assert np.array_equal(np.float32(a[k]), np.float32(b[k])), k
        q, r = np.array(a["quat"]), np.array(B.quat_from_basis(b["basis"]))
        assert min(np.abs(q - r).max(), np.abs(q + r).max()) < 1e-6


@pytest.mark.parametrize("name", ["le32Bitf.bullet"] + STACKS)
def test_cpp_reader_equals_oracle(name):
    same_scene(BL.read_bullet(gold(name)), B.read_scene(gold(name)))


def test_stack_layouts_hold_one_scene():
    ref = BL.read_bullet(gold(STACKS[0]))
    for name, (ptr, big) in zip(STACKS, [(4, False), (8, False), (4, True), (8, True)]):
        shapes, bodies, info = BL.read_bullet(gold(name))
        assert (info["pointer_size"], info["big_endian"]) == (ptr, big)
        assert shapes[1]["points"].shape == (12, 3)
        for a, b in zip(bodies, ref[1]):
            assert a == b


def test_quaternions_rotate_like_the_basis():
    _, bodies, _ = BL.read_bullet(gold("stack_le32.bullet"))
    for b in bodies:
        x, y, z, w = b["quat"]
        assert abs(x * x + y * y + z * z + w * w - 1.0) < 1e-6
        m = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(m.reshape(-1), b["basis"], atol=1e-6)


def test_double_precision_refused():
    with pytest.raises(L.MiRbError) as e:
        BL.read_bullet(gold("stack_double.bullet"))
    assert e.value.status == 6
    with pytest.raises(ValueError):
        B.read_scene(gold("stack_double.bullet"))


def test_missing_file():
    with pytest.raises(L.MiRbError) as e:
        BL.read_bullet(gold("no_such_file.bullet"))
    assert e.value.status == 1


def test_damaged_files_are_refused_or_read(tmp_path):
    """Truncations and byte flips: the reader returns FORMAT or a result,
    never reads outside the file."""
    data = open(gold("stack_be64.bullet"), "rb").read()
    rng = np.random.default_rng(7)
    cases = [data[:n] for n in range(0, len(data), 211)]
    for _ in range(150):
        d = bytearray(data)
        for p in rng.integers(0, len(d), size=int(rng.integers(1, 6))):
            d[p] = int(rng.integers(0, 256))
        cases.append(bytes(d))
    path = tmp_path / "damaged.bullet"
    seen = set()
    for d in cases:
        path.write_bytes(d)
        try:
            BL.read_bullet(path)
            seen.add(0)
        except L.MiRbError as e:
            assert e.status == 6
            seen.add(6)
    assert seen == {0, 6}


def test_oracle_registrations_follow_the_extractor():
    _, shapes, bodies = B.read_scene(gold("stack_le32.bullet"))
    hulls, regs = B.registrations(shapes, bodies, keep_static=False)
    # box, convex hull, ground box; sphere and plane left out; the pyramid ground last
    assert len(hulls) == 4
    assert np.allclose(hulls[0][1], 0.5) and np.allclose(hulls[2][1], [19.96, 1.0, 19.96])
    assert np.allclose(hulls[1][0], shapes[1]["points"] * [1.0, 0.5, 1.0])
    dyn = [r for r in regs if r[0] > 0]
    assert len(dyn) == 36 + 8 and len(regs) == 44 + 2500
    assert [r[3] for r in dyn] == [0] * 36 + [1] * 8
    hulls_s, regs_s = B.registrations(shapes, bodies, keep_static=True)
    assert len(hulls_s) == 3 and len(regs_s) == 45 and regs_s[-1][3] == 2
