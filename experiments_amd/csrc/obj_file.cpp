// Wavefront .obj meshes for concave shapes: the reference reads them with
// its objLoader (rendering/WavefrontObjLoader/obj_parser.cpp:19-105) and
// registers them with registerConcaveMeshShape(objLoader*, ...)
// (btGpuNarrowPhaseAndSolver.cpp:335-400), as the demo does for its ground
// mesh (BulletDataExtractor.cpp:262-310).  Only "v" and "f" lines matter:
//   - "v x y z": a vertex;
//   - "f a b c ...": a face whose vertex tokens are "v", "v/t", "v//n" or
//     "v/t/n"; the vertex index is the token's leading integer, 1-based, or
//     counted back from the vertices read so far when negative (0 is no
//     vertex and an error here);
//   - a face gives the triangle of its first three vertices (the reference
//     takes vertex_index[0..2]; a quad keeps only its first half).
// Every other line (vt, vn, mtllib, usemtl, o, g, s, comments) is skipped.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mi_rb.h"

namespace {

struct ObjMesh {
    std::vector<float> xyz;
    std::vector<int> tri;
};

mi_rb_status parse_obj(const char *path, ObjMesh &m) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return MI_RB_INVALID_ARG;
    std::string line;
    mi_rb_status st = MI_RB_OK;
    for (int c = 0; st == MI_RB_OK;) {
        line.clear();
        while ((c = std::fgetc(f)) != EOF && c != '\n') line.push_back((char)c);
        if (line.empty() && c == EOF) break;
        std::vector<char> buf(line.begin(), line.end());
        buf.push_back('\0');
        char *save = nullptr;
        const char *ws = " \t\r\n";
        char *tok = strtok_r(buf.data(), ws, &save);
        if (!tok) continue;
        if (std::strcmp(tok, "v") == 0) {
            float v[3];
            for (int k = 0; k < 3; k++) {
                char *t = strtok_r(nullptr, ws, &save);
                char *end = nullptr;
                v[k] = t ? std::strtof(t, &end) : 0.f;
                if (!t || end == t) st = MI_RB_INVALID_ARG;
            }
            m.xyz.insert(m.xyz.end(), v, v + 3);
        } else if (std::strcmp(tok, "f") == 0) {
            const int nv = (int)m.xyz.size() / 3;
            int idx[3], count = 0;
            for (char *t; (t = strtok_r(nullptr, ws, &save)) != nullptr;) {
                const int i = std::atoi(t);
                if (count < 3) idx[count] = i < 0 ? nv + i : i - 1;
                if (i == 0) st = MI_RB_INVALID_ARG;
                count++;
            }
            if (count < 3) st = MI_RB_INVALID_ARG;
            if (st == MI_RB_OK) m.tri.insert(m.tri.end(), idx, idx + 3);
        }
        if (c == EOF) break;
    }
    std::fclose(f);
    if (st != MI_RB_OK) return st;
    const int nv = (int)m.xyz.size() / 3;
    for (int i : m.tri)
        if (i < 0 || i >= nv) return MI_RB_INVALID_ARG;
    return MI_RB_OK;
}

}  // namespace

extern "C" mi_rb_status mi_rb_obj_read(const char *path, float *xyz, int max_vertices, int *indices,
                                       int max_triangles, int *num_vertices, int *num_triangles) {
    if (!path || !num_vertices || !num_triangles) return MI_RB_INVALID_ARG;
    ObjMesh m;
    const mi_rb_status st = parse_obj(path, m);
    if (st != MI_RB_OK) return st;
    const int nv = (int)m.xyz.size() / 3, nt = (int)m.tri.size() / 3;
    *num_vertices = nv;
    *num_triangles = nt;
    if (!xyz && !indices) return MI_RB_OK;  // sizes only
    if (!xyz || !indices || max_vertices < nv || max_triangles < nt) return MI_RB_CAPACITY;
    std::memcpy(xyz, m.xyz.data(), sizeof(float) * m.xyz.size());
    std::memcpy(indices, m.tri.data(), sizeof(int) * m.tri.size());
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_register_concave_mesh_obj(void *world, const char *path,
                                                        const float scale[3], int *out_shape) {
    if (!world || !path || !out_shape) return MI_RB_INVALID_ARG;
    ObjMesh m;
    const mi_rb_status st = parse_obj(path, m);
    if (st != MI_RB_OK) return st;
    if (m.tri.empty()) return MI_RB_INVALID_ARG;
    return mi_rb_register_concave_mesh(world, m.xyz.data(), 3 * (int)sizeof(float),
                                       (int)m.xyz.size() / 3, m.tri.data(), (int)m.tri.size() / 3,
                                       scale, out_shape);
}
