"""Test infrastructure: a pure-Python restatement of Bullet's ".bullet" file
format and of the reference's scene conversion, the checker for the
library's C++ reader (experiments_amd/csrc/bullet_file.cpp).  Only tests and
the fixture scripts use it.

Format (bParse, the reader the reference uses through btBulletFile in
opencl/gpu_rigidbody_pipeline2/BulletDataExtractor.cpp:454-520): a 12-byte
header ("BULLET", 'f' or 'd' for float or double scalars, '_' or '-' for 4-
or 8-byte pointers, 'v' or 'V' for little or big endian, three version
digits), then chunks (code, length, the address the data had when written,
struct index, element count; 16 bytes plus a pointer), ended by "ENDB" or the
end of the file.  The "DNA1" chunk holds the type system: names, types, type
lengths and, per struct, its (type, name) members, so every layout is read
from the file.  Blender's ".blend" files use the same container with a
"BLENDER" magic, which is how the 64-bit and big-endian cases are checked
against files written by other software.

Conversion (btBulletDataExtractor::convertAllObjects and
convertCollisionShape, BulletDataExtractor.cpp:594-1010): shapes in file
order; a box becomes the unit cube scaled by half extent * local scaling +
margin (createBoxShape, :964-997), a convex hull its points times the local
scaling; spheres, planes, meshes and other types give no shape.  Bodies are
registered shape by shape, in file order within a shape, with mass 1 /
inverse mass; static bodies (inverse mass 0) only when keep_static, and
without keep_static a 50 x 50 grid of static square pyramids (scaled 2.5, 2,
2.5) at i * 5.01 - 120 is added (:660-690).

Here member paths are written out in full per struct type (no search by
name), unlike the C++ reader.
"""
import math
import struct

import numpy as np

BOX, CONVEX_HULL, SPHERE, TRIANGLE_MESH, STATIC_PLANE, COMPOUND = 0, 4, 8, 21, 28, 31


def _parse_name(name):
    """'*m_x[3][2]' -> ('m_x', is_pointer, count)."""
    ptr = name.startswith("(*") or name.startswith("*")
    base = name.lstrip("(*")
    count = 1
    if "[" in base:
        dims = base[base.index("["):]
        base = base[:base.index("[")]
        for part in dims.split("[")[1:]:
            count *= int(part.split("]")[0])
    base = base.split(")")[0]
    return base, ptr, count


class DNA:
    def __init__(self, names, types, tlen, structs):
        self.names, self.types, self.tlen = names, types, list(tlen)
        self.structs = structs          # list of (type index, [(type index, name index)])
        self.index = {types[t]: i for i, (t, _) in enumerate(structs)}
        self._members = {}

    @classmethod
    def parse(cls, blob, big):
        e = ">" if big else "<"
        p = 0

        def tag(t):
            nonlocal p
            assert blob[p:p + 4] == t, (blob[p:p + 4], t)
            p += 4

        def align():
            nonlocal p
            p = (p + 3) & ~3

        def strings():
            nonlocal p
            (n,) = struct.unpack_from(e + "i", blob, p)
            p += 4
            out = []
            for _ in range(n):
                q = blob.index(b"\0", p)
                out.append(blob[p:q].decode("latin-1"))
                p = q + 1
            align()
            return out

        tag(b"SDNA")
        tag(b"NAME")
        names = strings()
        tag(b"TYPE")
        types = strings()
        tag(b"TLEN")
        tlen = struct.unpack_from(e + "%dh" % len(types), blob, p)
        p += 2 * len(types)
        align()
        tag(b"STRC")
        (ns,) = struct.unpack_from(e + "i", blob, p)
        p += 4
        structs = []
        for _ in range(ns):
            t, nf = struct.unpack_from(e + "hh", blob, p)
            p += 4
            fl = struct.unpack_from(e + "%dh" % (2 * nf), blob, p)
            p += 4 * nf
            structs.append((t, [(fl[2 * k], fl[2 * k + 1]) for k in range(nf)]))
        return cls(names, types, tlen, structs)

    def members(self, sname, ptr_size):
        """[(name, type name, is_pointer, count, offset, element size)]."""
        key = (sname, ptr_size)
        if key not in self._members:
            self._members[key] = self._layout(sname, ptr_size)
        return self._members[key]

    def _layout(self, sname, ptr_size):
        out, off = [], 0
        for t, n in self.structs[self.index[sname]][1]:
            base, ptr, cnt = _parse_name(self.names[n])
            tname = self.types[t]
            esz = ptr_size if ptr else self.size(tname, ptr_size)
            out.append((base, tname, ptr, cnt, off, esz))
            off += esz * cnt
        return out

    def size(self, tname, ptr_size):
        if tname in self.index:
            return sum(m[3] * m[5] for m in self.members(tname, ptr_size))
        return self.tlen[self.types.index(tname)]

    def encode(self, big, ptr_size):
        """The DNA block for a file with this byte order and pointer size."""
        e = ">" if big else "<"
        out = bytearray(b"SDNA")

        def align():
            while len(out) % 4:
                out.append(0)

        def strings(tag, lst):
            out.extend(tag)
            out.extend(struct.pack(e + "i", len(lst)))
            for s in lst:
                out.extend(s.encode("latin-1") + b"\0")
            align()

        strings(b"NAME", self.names)
        strings(b"TYPE", self.types)
        out.extend(b"TLEN")
        out.extend(struct.pack(e + "%dh" % len(self.types),
                               *[self.size(t, ptr_size) for t in self.types]))
        align()
        out.extend(b"STRC")
        out.extend(struct.pack(e + "i", len(self.structs)))
        for t, fl in self.structs:
            out.extend(struct.pack(e + "hh", t, len(fl)))
            for ft, fn in fl:
                out.extend(struct.pack(e + "hh", ft, fn))
        return bytes(out)


class BulletFile:
    def __init__(self, data, magic=b"BULLET"):
        m = len(magic)
        if data[:m] != magic:
            raise ValueError("bad magic")
        flags = data[m:m + 3] if magic == b"BULLET" else b"f" + data[m:m + 2]
        self.doubles = flags[0:1] == b"d"
        self.ptr_size = 8 if flags[1:2] == b"-" else 4
        self.big = flags[2:3] == b"V"
        self.version = int(data[m + (3 if magic == b"BULLET" else 2):12])
        e = ">" if self.big else "<"
        self.data = data
        self.chunks = []
        off, hdr = 12, 16 + self.ptr_size
        dna_blob = None
        while off + hdr <= len(data):
            code = data[off:off + 4]
            ln = struct.unpack_from(e + "i", data, off + 4)[0]
            old = struct.unpack_from(e + ("Q" if self.ptr_size == 8 else "I"), data, off + 8)[0]
            sdna, nr = struct.unpack_from(e + "ii", data, off + 8 + self.ptr_size)
            if code == b"ENDB":
                break
            if ln < 0 or off + hdr + ln > len(data):
                raise ValueError("truncated chunk")
            self.chunks.append((code, ln, old, sdna, nr, off + hdr))
            if code == b"DNA1":
                dna_blob = data[off + hdr:off + hdr + ln]
            off += hdr + ln
        if dna_blob is None:
            raise ValueError("no DNA1")
        self.dna = DNA.parse(dna_blob, self.big)
        self.by_ptr = {c[2]: c for c in self.chunks if c[0] != b"DNA1" and c[2]}

    def struct_name(self, chunk):
        return self.dna.types[self.dna.structs[chunk[3]][0]]

    def get(self, chunk, elem, path):
        """Value at a full member path ('a.b[1].c[2]') of element elem."""
        e = ">" if self.big else "<"
        sname = self.struct_name(chunk)
        off = chunk[5] + elem * self.dna.size(sname, self.ptr_size)
        parts = path.split(".")
        for i, part in enumerate(parts):
            idx = 0
            if "[" in part:
                part, rest = part.split("[")
                idx = int(rest.rstrip("]"))
            mem = {m[0]: m for m in self.dna.members(sname, self.ptr_size)}[part]
            _, tname, ptr, cnt, moff, esz = mem
            assert idx < cnt
            off += moff + idx * esz
            if i + 1 < len(parts):
                sname = tname
                continue
            if ptr:
                return struct.unpack_from(e + ("Q" if self.ptr_size == 8 else "I"), self.data, off)[0]
            fmt = {"float": "f", "double": "d", "int": "i", "short": "h", "char": "b"}[tname]
            return struct.unpack_from(e + fmt, self.data, off)[0]


# full member paths per struct (Bullet 2.77 serialization structs)
_SHAPE_BASE = {
    "btConvexInternalShapeData": "m_collisionShapeData",
    "btConvexHullShapeData": "m_convexInternalShapeData.m_collisionShapeData",
    "btStaticPlaneShapeData": "m_collisionShapeData",
    "btTriangleMeshShapeData": "m_collisionShapeData",
    "btCompoundShapeData": "m_collisionShapeData",
    "btCollisionShapeData": None,
}
_CONVEX = {"btConvexInternalShapeData": "", "btConvexHullShapeData": "m_convexInternalShapeData."}


def read_scene(path_or_bytes):
    """Shapes and rigid bodies of a float .bullet file, as dicts."""
    data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else \
        open(path_or_bytes, "rb").read()
    f = BulletFile(bytes(data))
    if f.doubles:
        raise ValueError("double-precision file")
    shapes, ptr_of = [], {}
    for c in f.chunks:
        if c[0] != b"SHAP":
            continue
        sname = f.struct_name(c)
        for el in range(c[4]):
            base = _SHAPE_BASE.get(sname)
            pre = (base + ".") if base else ""
            s = {"type": f.get(c, el, pre + "m_shapeType"), "local_scaling": [1.0] * 3,
                 "implicit_dims": [0.0] * 3, "margin": 0.0, "plane_normal": [0.0] * 3,
                 "plane_constant": 0.0, "points": np.zeros((0, 3), np.float32)}
            if sname in _CONVEX:
                q = _CONVEX[sname]
                s["local_scaling"] = [f.get(c, el, q + "m_localScaling.m_floats[%d]" % k) for k in range(3)]
                s["implicit_dims"] = [f.get(c, el, q + "m_implicitShapeDimensions.m_floats[%d]" % k)
                                      for k in range(3)]
                s["margin"] = f.get(c, el, q + "m_collisionMargin")
            if sname == "btConvexHullShapeData":
                ptr = f.get(c, el, "m_unscaledPointsFloatPtr")
                n = f.get(c, el, "m_numUnscaledPoints")
                if ptr and n > 0:
                    pc = f.by_ptr[ptr]
                    s["points"] = np.array([[f.get(pc, i, "m_floats[%d]" % k) for k in range(3)]
                                            for i in range(n)], np.float32)
            if sname == "btStaticPlaneShapeData":
                s["local_scaling"] = [f.get(c, el, "m_localScaling.m_floats[%d]" % k) for k in range(3)]
                s["plane_normal"] = [f.get(c, el, "m_planeNormal.m_floats[%d]" % k) for k in range(3)]
                s["plane_constant"] = f.get(c, el, "m_planeConstant")
            if sname in ("btTriangleMeshShapeData", "btCompoundShapeData"):
                s["margin"] = f.get(c, el, "m_collisionMargin")
            ptr_of[c[2] + el * f.dna.size(sname, f.ptr_size)] = len(shapes)
            shapes.append(s)
    bodies = []
    for c in f.chunks:
        if c[0] != b"RBDY":
            continue
        for el in range(c[4]):
            co = "m_collisionObjectData."
            wt = co + "m_worldTransform."
            b = {"shape": ptr_of.get(f.get(c, el, co + "m_collisionShape"), -1),
                 "inverse_mass": f.get(c, el, "m_inverseMass"),
                 "basis": [f.get(c, el, wt + "m_basis.m_el[%d].m_floats[%d]" % (r, k))
                           for r in range(3) for k in range(3)],
                 "origin": [f.get(c, el, wt + "m_origin.m_floats[%d]" % k) for k in range(3)],
                 "linvel": [f.get(c, el, "m_linearVelocity.m_floats[%d]" % k) for k in range(3)],
                 "angvel": [f.get(c, el, "m_angularVelocity.m_floats[%d]" % k) for k in range(3)],
                 "friction": f.get(c, el, co + "m_friction"),
                 "restitution": f.get(c, el, co + "m_restitution")}
            bodies.append(b)
    return f, shapes, bodies


def quat_from_basis(m):
    """Rotation matrix (rows, 9 values) to (x, y, z, w): the component of
    largest magnitude from the diagonal (4w^2 = 1 + trace, 4x^2 = 1 + m00 -
    m11 - m22, ...), the others from the off-diagonal terms."""
    m = [float(v) for v in m]
    m00, m01, m02, m10, m11, m12, m20, m21, m22 = m
    cand = [1 + m00 - m11 - m22, 1 - m00 + m11 - m22, 1 - m00 - m11 + m22, 1 + m00 + m11 + m22]
    k = int(np.argmax(cand))
    r = math.sqrt(cand[k]) * 0.5
    f = 0.25 / r
    if k == 3:
        q = [(m21 - m12) * f, (m02 - m20) * f, (m10 - m01) * f, r]
    elif k == 0:
        q = [r, (m01 + m10) * f, (m02 + m20) * f, (m21 - m12) * f]
    elif k == 1:
        q = [(m01 + m10) * f, r, (m12 + m21) * f, (m02 - m20) * f]
    else:
        q = [(m02 + m20) * f, (m12 + m21) * f, r, (m10 - m01) * f]
    return q


def registrations(shapes, bodies, keep_static=False):
    """What the conversion registers: hulls [(points Nx3, scale3)] in order
    and bodies [(mass, pos3, quat4, hull index)] in order."""
    hulls, index = [], {}
    for i, s in enumerate(shapes):
        if s["type"] == BOX:
            pts = np.array([[x, y, z] for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)], np.float32)
            scale = [np.float32(s["implicit_dims"][k]) * np.float32(s["local_scaling"][k]) +
                     np.float32(s["margin"]) for k in range(3)]
        elif s["type"] == CONVEX_HULL and len(s["points"]):
            pts = (s["points"] * np.array(s["local_scaling"], np.float32)).astype(np.float32)
            scale = [1.0, 1.0, 1.0]
        else:
            continue
        index[i] = len(hulls)
        hulls.append((pts, np.array(scale, np.float32)))
    regs = []
    for i in range(len(shapes)):
        if i not in index:
            continue
        for b in bodies:
            if b["shape"] != i or (not keep_static and b["inverse_mass"] == 0.0):
                continue
            mass = 0.0 if b["inverse_mass"] == 0.0 else float(np.float32(1.0) / np.float32(b["inverse_mass"]))
            regs.append((mass, list(b["origin"]), quat_from_basis(b["basis"]), index[i]))
    if not keep_static:
        pyr = np.array([[-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1], [0, -1, 0]], np.float32)
        g = len(hulls)
        hulls.append((pyr, np.array([2.5, 2.0, 2.5], np.float32)))
        for i in range(50):
            for j in range(50):
                regs.append((0.0, [float(np.float32(i * 5.01 - 120)), 0.0,
                                   float(np.float32(j * 5.01 - 120))], [0.0, 0.0, 0.0, 1.0], g))
    return hulls, regs
