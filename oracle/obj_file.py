"""Oracle restatement of the reference's Wavefront .obj reading for concave
meshes (test infrastructure: only tests import it).  It follows
rendering/WavefrontObjLoader/obj_parser.cpp: obj_parse_vertex_index (55-92)
takes each face token's leading integer as the vertex index ("v", "v/t",
"v//n", "v/t/n"), obj_convert_to_list_index (19-28) makes it 0-based
(index - 1) or counts it back from the vertices read so far when negative;
registerConcaveMeshShape(objLoader*, ...) (btGpuNarrowPhaseAndSolver.cpp:
335-400) takes each face's first three vertices as its triangle.  Vertex
lines give x, y, z.  Every other line is skipped."""
import re

import numpy as np


def _leading_int(tok):
    m = re.match(r"[+-]?\d+", tok)
    return int(m.group(0)) if m else 0


def read_obj_text(text):
    verts, tris = [], []
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            verts.append([np.float32(float(x)) for x in t[1:4]])
        elif t[0] == "f":
            idx = [_leading_int(x) for x in t[1:]]
            if len(idx) < 3 or 0 in idx:
                raise ValueError("face with fewer than 3 vertices or a zero index")
            n = len(verts)
            tris.append([i - 1 if i > 0 else n + i for i in idx[:3]])
    v = np.array(verts, np.float32).reshape(-1, 3)
    f = np.array(tris, np.int32).reshape(-1, 3)
    if f.size and (f.min() < 0 or f.max() >= len(v)):
        raise ValueError("vertex index out of range")
    return v, f


def read_obj(path):
    with open(path, "r", encoding="latin-1") as fh:
        return read_obj_text(fh.read())
