"""Writes the .bullet fixtures under tests/golden/ (run from the repository
root; the committed files are its output).

* le32Bitf.bullet is the reference's own test file, copied as data
  (skeletal_animation/animkit/UnitTests/FbtUnitTests/TestData/le32Bitf.bullet):
  Bullet 2.77, 4-byte pointers, little endian, one static body with a
  triangle-mesh shape, and the full 2.77 DNA.
* blend_be32_trim.blend and blend_le64_trim.blend: the header, the first
  chunks of under 2 KB and the DNA of two Blender files in the same test
  directory (be32bit.blend: 4-byte pointers, big endian; le64bitLink.blend:
  8-byte pointers, little endian).  Blender writes the same container, so
  these pin the 64-bit and big-endian cases of the format on files written
  by other software.  Needs /root/reference (this container only).
* stack_{le,be}{32,64}.bullet: one small scene (boxes, convex hulls, a
  sphere and a static plane that the conversion leaves out, a static ground
  box) written with the 2.77 DNA of le32Bitf.bullet in all four pointer
  sizes and byte orders; stack_double.bullet is le32 with the double-
  precision flag, which readers must refuse.
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import bullet_file as B  # noqa: E402

REF_TESTDATA = "/root/reference/skeletal_animation/animkit/UnitTests/FbtUnitTests/TestData"


def trim_blend(src, dst, max_chunks=48, max_len=2048):
    d = open(src, "rb").read()
    f = B.BulletFile(d, magic=b"BLENDER")
    hdr = 16 + f.ptr_size
    out = bytearray(d[:12])
    kept = 0
    for c in f.chunks:
        if c[0] == b"DNA1" or (kept < max_chunks and c[1] <= max_len):
            out += d[c[5] - hdr:c[5] + c[1]]
            kept += c[0] != b"DNA1"
    e = ">" if f.big else "<"
    out += b"ENDB" + struct.pack(e + "i", 0) + b"\0" * f.ptr_size + struct.pack(e + "ii", 0, 0)
    open(dst, "wb").write(bytes(out))


class Writer:
    def __init__(self, dna, ptr_size, big):
        self.dna, self.ps, self.big = dna, ptr_size, big
        self.e = ">" if big else "<"
        self.chunks = []
        self.next_ptr = 0x7f3a00001000 if ptr_size == 8 else 0x10001000

    def alloc(self):
        p = self.next_ptr
        self.next_ptr += 0x1000
        return p

    def _resolve(self, sname, path):
        off = 0
        parts = path.split(".")
        for i, part in enumerate(parts):
            idx = 0
            if "[" in part:
                part, rest = part.split("[")
                idx = int(rest.rstrip("]"))
            mem = {m[0]: m for m in self.dna.members(sname, self.ps)}[part]
            _, tname, ptr, cnt, moff, esz = mem
            assert idx < cnt, path
            off += moff + idx * esz
            sname = tname
        return off, tname, ptr

    def pack(self, sname, values):
        buf = bytearray(self.dna.size(sname, self.ps))
        for path, v in values.items():
            off, tname, ptr = self._resolve(sname, path)
            if ptr:
                struct.pack_into(self.e + ("Q" if self.ps == 8 else "I"), buf, off, v)
            else:
                struct.pack_into(self.e + {"float": "f", "int": "i"}[tname], buf, off, v)
        return bytes(buf)

    def chunk(self, code, sname, elems, ptr=None):
        ptr = ptr if ptr is not None else self.alloc()
        sdna = self.dna.index[sname]
        self.chunks.append((code, b"".join(elems), ptr, sdna, len(elems)))
        return ptr

    def write(self, path, doubles=False):
        e = self.e
        out = bytearray(b"BULLET" + (b"d" if doubles else b"f") + (b"-" if self.ps == 8 else b"_") +
                        (b"V" if self.big else b"v") + b"277")
        dna = self.dna.encode(self.big, self.ps)
        for code, data, ptr, sdna, nr in self.chunks + [(b"DNA1", dna, self.alloc(), 0, 1)]:
            out += code + struct.pack(e + "i", len(data))
            out += struct.pack(e + ("Q" if self.ps == 8 else "I"), ptr)
            out += struct.pack(e + "ii", sdna, nr) + data
        out += b"ENDB" + struct.pack(e + "i", 0) + b"\0" * self.ps + struct.pack(e + "ii", 0, 0)
        open(path, "wb").write(bytes(out))


def v3(prefix, xyz):
    return {prefix + ".m_floats[%d]" % k: float(xyz[k]) for k in range(3)}


def quat_basis(q):
    x, y, z, w = q
    return [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]


def stack_scene(w):
    rng = np.random.default_rng(2013)
    cs = "m_collisionShapeData."
    shapes = []
    # box: Bullet keeps half extents minus the margin
    box = {cs + "m_shapeType": B.BOX, "m_collisionMargin": 0.04, **v3("m_localScaling", (1, 1, 1)),
           **v3("m_implicitShapeDimensions", (0.46, 0.46, 0.46))}
    shapes.append(w.chunk(b"SHAP", "btConvexInternalShapeData", [w.pack("btConvexInternalShapeData", box)]))
    pts = rng.normal(size=(12, 3))
    pts = (pts / np.linalg.norm(pts, axis=1, keepdims=True) * 0.6).astype(np.float32)
    arr = w.chunk(b"ARAY", "btVector3FloatData",
                  [w.pack("btVector3FloatData", {"m_floats[%d]" % k: float(p[k]) for k in range(3)})
                   for p in pts])
    ci = "m_convexInternalShapeData."
    hull = {ci + cs + "m_shapeType": B.CONVEX_HULL, ci + "m_collisionMargin": 0.04,
            **v3(ci + "m_localScaling", (1.0, 0.5, 1.0)), "m_unscaledPointsFloatPtr": arr,
            "m_numUnscaledPoints": len(pts)}
    shapes.append(w.chunk(b"SHAP", "btConvexHullShapeData", [w.pack("btConvexHullShapeData", hull)]))
    sphere = {cs + "m_shapeType": B.SPHERE, "m_collisionMargin": 0.5, **v3("m_localScaling", (1, 1, 1)),
              **v3("m_implicitShapeDimensions", (0.5, 0.5, 0.5))}
    shapes.append(w.chunk(b"SHAP", "btConvexInternalShapeData", [w.pack("btConvexInternalShapeData", sphere)]))
    plane = {cs + "m_shapeType": B.STATIC_PLANE, **v3("m_localScaling", (1, 1, 1)),
             **v3("m_planeNormal", (0, 1, 0)), "m_planeConstant": 0.0}
    shapes.append(w.chunk(b"SHAP", "btStaticPlaneShapeData", [w.pack("btStaticPlaneShapeData", plane)]))
    ground = {cs + "m_shapeType": B.BOX, "m_collisionMargin": 0.04, **v3("m_localScaling", (2, 1, 2)),
              **v3("m_implicitShapeDimensions", (9.96, 0.96, 9.96))}
    shapes.append(w.chunk(b"SHAP", "btConvexInternalShapeData", [w.pack("btConvexInternalShapeData", ground)]))
    co = "m_collisionObjectData."

    def body(shape_ptr, inv_mass, pos, q, linvel=(0, 0, 0)):
        vals = {co + "m_collisionShape": shape_ptr, "m_inverseMass": float(inv_mass),
                co + "m_friction": 0.5, co + "m_restitution": 0.0}
        m = quat_basis(q)
        for r in range(3):
            for k in range(3):
                vals[co + "m_worldTransform.m_basis.m_el[%d].m_floats[%d]" % (r, k)] = float(m[3 * r + k])
        vals.update(v3(co + "m_worldTransform.m_origin", pos))
        vals.update(v3("m_linearVelocity", linvel))
        return w.pack("btRigidBodyFloatData", vals)

    rb = []
    for i in range(3):
        for j in range(3):
            for k in range(4):
                rb.append(body(shapes[0], 1.0 / (1 + (i + j + k) % 3), (i * 1.1 - 1.1, 0.55 + k * 1.02, j * 1.1 - 1.1),
                               (0, 0, 0, 1)))
    for n in range(8):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        rb.append(body(shapes[1], 0.25, (4.0 + (n % 4) * 1.5, 1.0 + (n // 4) * 1.5, 0.0), q, (0, -1, 0)))
    rb.append(body(shapes[2], 1.0, (-4.0, 2.0, 0.0), (0, 0, 0, 1)))
    rb.append(body(shapes[2], 1.0, (-4.0, 3.5, 0.0), (0, 0, 0, 1)))
    rb.append(body(shapes[3], 0.0, (0.0, 0.0, 0.0), (0, 0, 0, 1)))
    rb.append(body(shapes[4], 0.0, (0.0, -1.0, 0.0), (0, 0, 0, 1)))
    # bodies one chunk each, as btSerializer writes them
    for r in rb:
        w.chunk(b"RBDY", "btRigidBodyFloatData", [r])


def main():
    os.chdir(HERE)
    if os.path.isdir(REF_TESTDATA):
        trim_blend(os.path.join(REF_TESTDATA, "be32bit.blend"), "blend_be32_trim.blend")
        trim_blend(os.path.join(REF_TESTDATA, "le64bitLink.blend"), "blend_le64_trim.blend")
    dna = B.BulletFile(open("le32Bitf.bullet", "rb").read()).dna
    for big in (False, True):
        for ps in (4, 8):
            w = Writer(dna, ps, big)
            stack_scene(w)
            w.write("stack_%s%d.bullet" % ("be" if big else "le", 8 * ps))
    w = Writer(dna, 4, False)
    stack_scene(w)
    w.write("stack_double.bullet", doubles=True)


if __name__ == "__main__":
    main()
