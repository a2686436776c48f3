// hull.cpp -- convex hull and polyhedral features at shape registration (host).
//
// Replaces btConvexHullComputer + btConvexUtility::initializePolyhedralFeatures /
// initialize (opencl/gpu_rigidbody_pipeline/btConvexUtility.cpp:31-510) as used by
// CLPhysicsDemo::registerConvexPolyhedron (opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.cpp:491-530):
//   * the hull of the points, as outward-wound triangles (incremental hull in double);
//   * the face plane of each triangle from its first two normalized edges, with
//     plane w = -min over the face's vertices of dot(v, n) (btConvexUtility.cpp:73-113);
//   * coplanar triangles (normal dot > 0.999) merged into one polygon ordered
//     counter-clockwise about the face normal, unless a dropped vertex is used
//     by another face (btConvexUtility.cpp:117-268);
//   * unique edge directions, normalized, one per +-direction within 1e-6
//     (btConvexUtility.cpp:339-390).
// Vertices keep the order in which they appear in the input.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mi_rb.h"
#include "hull.h"

namespace {

struct V3 {
    double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double len(V3 a) { return std::sqrt(dot(a, a)); }

struct Tri {
    int v[3];
    V3 n;
    double d;
    bool alive;
};

// Incremental convex hull; returns outward-wound triangles over point indices.
bool build_hull(const std::vector<V3> &P, std::vector<Tri> &tris) {
    int n = (int)P.size();
    if (n < 4) return false;
    double scale = 0;
    for (auto &p : P) scale = std::max(scale, std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z))));
    const double eps = 1e-9 * std::max(scale, 1.0);
    // initial tetrahedron
    int i0 = 0, i1 = -1, i2 = -1, i3 = -1;
    for (int i = 1; i < n; i++)
        if (len(sub(P[i], P[i0])) > eps * 10) { i1 = i; break; }
    if (i1 < 0) return false;
    for (int i = 1; i < n; i++)
        if (len(cross(sub(P[i1], P[i0]), sub(P[i], P[i0]))) > eps * 10 * std::max(scale, 1.0)) { i2 = i; break; }
    if (i2 < 0) return false;
    V3 nrm = cross(sub(P[i1], P[i0]), sub(P[i2], P[i0]));
    for (int i = 1; i < n; i++)
        if (std::fabs(dot(nrm, sub(P[i], P[i0]))) > eps * 10 * std::max(scale * scale, 1.0)) { i3 = i; break; }
    if (i3 < 0) return false;
    auto make = [&](int a, int b, int c) {
        Tri t;
        t.v[0] = a; t.v[1] = b; t.v[2] = c;
        t.n = cross(sub(P[b], P[a]), sub(P[c], P[a]));
        double l = len(t.n);
        t.n = {t.n.x / l, t.n.y / l, t.n.z / l};
        t.d = dot(t.n, P[a]);
        t.alive = true;
        return t;
    };
    V3 cen = {(P[i0].x + P[i1].x + P[i2].x + P[i3].x) / 4, (P[i0].y + P[i1].y + P[i2].y + P[i3].y) / 4,
              (P[i0].z + P[i1].z + P[i2].z + P[i3].z) / 4};
    int q[4] = {i0, i1, i2, i3};
    int faces[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    tris.clear();
    for (auto &f : faces) {
        Tri t = make(q[f[0]], q[f[1]], q[f[2]]);
        if (dot(t.n, cen) - t.d > 0) t = make(q[f[0]], q[f[2]], q[f[1]]);
        tris.push_back(t);
    }
    for (int p = 0; p < n; p++) {
        if (p == i0 || p == i1 || p == i2 || p == i3) continue;
        std::vector<int> vis;
        for (int t = 0; t < (int)tris.size(); t++)
            if (tris[t].alive && dot(tris[t].n, P[p]) - tris[t].d > eps) vis.push_back(t);
        if (vis.empty()) continue;
        // horizon: directed edges of visible faces whose twin is not visible
        std::vector<std::pair<int, int>> edges;
        for (int t : vis)
            for (int k = 0; k < 3; k++) edges.push_back({tris[t].v[k], tris[t].v[(k + 1) % 3]});
        std::vector<std::pair<int, int>> horizon;
        for (auto &e : edges) {
            bool twin = false;
            for (auto &f : edges)
                if (f.first == e.second && f.second == e.first) { twin = true; break; }
            if (!twin) horizon.push_back(e);
        }
        for (int t : vis) tris[t].alive = false;
        for (auto &e : horizon) tris.push_back(make(e.first, e.second, p));
    }
    std::vector<Tri> out;
    for (auto &t : tris)
        if (t.alive) out.push_back(t);
    tris.swap(out);
    return !tris.empty();
}

struct FaceF {
    std::vector<int> idx;  // indices into the hull vertex list
    float plane[4];
};

float fdot(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
void fnormalize(float *v) {
    float l = std::sqrt(fdot(v, v));
    if (l > 0) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

// face plane as btConvexUtility computes it: normalize(cross(e0, e1)) of the
// first two normalized edges, w = -min_v dot(v, n)
void face_plane(const std::vector<float> &V, FaceF &f) {
    float e[2][3];
    for (int k = 0; k < 2; k++) {
        const float *a = &V[3 * f.idx[k]], *b = &V[3 * f.idx[(k + 1) % f.idx.size()]];
        e[k][0] = b[0] - a[0]; e[k][1] = b[1] - a[1]; e[k][2] = b[2] - a[2];
        fnormalize(e[k]);
    }
    float nrm[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                    e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    fnormalize(nrm);
    float eq = 1e30f;
    for (int v : f.idx) {
        float d = fdot(&V[3 * v], nrm);
        if (eq > d) eq = d;
    }
    f.plane[0] = nrm[0]; f.plane[1] = nrm[1]; f.plane[2] = nrm[2]; f.plane[3] = -eq;
}

}  // namespace

int rb_build_hull(const float *xyz, int stride_bytes, int n, const float scale[3], HullData &out) {
    out = HullData();
    if (!xyz || n < 4 || stride_bytes < 12) return MI_RB_INVALID_ARG;
    std::vector<V3> P(n);
    const unsigned char *base = reinterpret_cast<const unsigned char *>(xyz);
    float s[3] = {1.f, 1.f, 1.f};
    if (scale) { s[0] = scale[0]; s[1] = scale[1]; s[2] = scale[2]; }
    std::vector<float> inF(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const float *v = reinterpret_cast<const float *>(base + (size_t)i * stride_bytes);
        inF[3 * i] = v[0] * s[0]; inF[3 * i + 1] = v[1] * s[1]; inF[3 * i + 2] = v[2] * s[2];
        P[i] = {inF[3 * i], inF[3 * i + 1], inF[3 * i + 2]};
    }
    std::vector<Tri> tris;
    if (!build_hull(P, tris)) return MI_RB_INVALID_ARG;
    // hull vertices in input order (duplicates of a used point are dropped)
    std::vector<int> used(n, 0), remap(n, -1);
    for (auto &t : tris)
        for (int k = 0; k < 3; k++) used[t.v[k]] = 1;
    std::vector<float> V;
    for (int i = 0; i < n; i++)
        if (used[i]) {
            remap[i] = (int)(V.size() / 3);
            V.push_back(inF[3 * i]); V.push_back(inF[3 * i + 1]); V.push_back(inF[3 * i + 2]);
        }
    std::vector<FaceF> tmp(tris.size());
    for (size_t t = 0; t < tris.size(); t++) {
        for (int k = 0; k < 3; k++) tmp[t].idx.push_back(remap[tris[t].v[k]]);
        face_plane(V, tmp[t]);
    }
    // merge coplanar groups (btConvexUtility.cpp:117-268)
    std::vector<FaceF> faces;
    std::vector<int> todo;
    for (int i = 0; i < (int)tmp.size(); i++) todo.push_back(i);
    while (!todo.empty()) {
        std::vector<int> group;
        int ref = todo.back();
        todo.pop_back();
        group.push_back(ref);
        for (int j = (int)todo.size() - 1; j >= 0; j--) {
            int i = todo[j];
            if (fdot(tmp[ref].plane, tmp[i].plane) > 0.999f) {
                group.push_back(i);
                todo.erase(todo.begin() + j);
            }
        }
        bool merged = false;
        if (group.size() > 1) {
            std::vector<int> pts;
            double an[3] = {0, 0, 0};
            for (int g : group) {
                for (int k = 0; k < 3; k++) an[k] += tmp[g].plane[k];
                for (int v : tmp[g].idx)
                    if (std::find(pts.begin(), pts.end(), v) == pts.end()) pts.push_back(v);
            }
            double al = std::sqrt(an[0] * an[0] + an[1] * an[1] + an[2] * an[2]);
            V3 nrm = {an[0] / al, an[1] / al, an[2] / al};
            // 2D convex hull (monotone chain) in the plane, counter-clockwise about nrm
            V3 ax = std::fabs(nrm.x) > 0.7 ? V3{0, 1, 0} : V3{1, 0, 0};
            V3 u = cross(ax, nrm);
            u = {u.x / len(u), u.y / len(u), u.z / len(u)};
            V3 w = cross(nrm, u);
            struct P2 { double a, b; int id; };
            std::vector<P2> q;
            for (int v : pts) {
                V3 p = {V[3 * v], V[3 * v + 1], V[3 * v + 2]};
                q.push_back({dot(p, u), dot(p, w), v});
            }
            std::sort(q.begin(), q.end(), [](const P2 &l, const P2 &r) {
                return l.a < r.a || (l.a == r.a && (l.b < r.b || (l.b == r.b && l.id < r.id)));
            });
            auto crs = [](const P2 &o, const P2 &a, const P2 &b) {
                return (a.a - o.a) * (b.b - o.b) - (a.b - o.b) * (b.a - o.a);
            };
            std::vector<P2> h(2 * q.size());
            int k = 0;
            for (size_t i = 0; i < q.size(); i++) {
                while (k >= 2 && crs(h[k - 2], h[k - 1], q[i]) <= 1e-12) k--;
                h[k++] = q[i];
            }
            for (int i = (int)q.size() - 2, t = k + 1; i >= 0; i--) {
                while (k >= t && crs(h[k - 2], h[k - 1], q[i]) <= 1e-12) k--;
                h[k++] = q[i];
            }
            h.resize(k > 1 ? k - 1 : k);
            std::vector<int> hull_ids;
            for (auto &p2 : h) hull_ids.push_back(p2.id);
            bool reject = false;
            for (int v : pts) {
                if (std::find(hull_ids.begin(), hull_ids.end(), v) != hull_ids.end()) continue;
                for (int j = 0; j < (int)tmp.size() && !reject; j++) {
                    if (std::find(group.begin(), group.end(), j) != group.end()) continue;
                    if (std::find(tmp[j].idx.begin(), tmp[j].idx.end(), v) != tmp[j].idx.end()) reject = true;
                }
                if (reject) break;
            }
            if (!reject && hull_ids.size() >= 3) {
                // the merged face keeps the plane of the group's first triangle
                // (btConvexUtility.cpp:181-183)
                FaceF f;
                f.idx = hull_ids;
                for (int k = 0; k < 4; k++) f.plane[k] = tmp[ref].plane[k];
                faces.push_back(f);
                merged = true;
            }
        }
        if (!merged)
            for (int g : group) faces.push_back(tmp[g]);
    }
    // unique edges (btConvexUtility.cpp:339-390)
    std::vector<float> E;
    for (auto &f : faces) {
        int nv = (int)f.idx.size();
        for (int j = 0; j < nv; j++) {
            int a = f.idx[j], b = f.idx[(j + 1) % nv];
            int v0 = a > b ? a : b, v1 = a > b ? b : a;
            float e[3] = {V[3 * v1] - V[3 * v0], V[3 * v1 + 1] - V[3 * v0 + 1], V[3 * v1 + 2] - V[3 * v0 + 2]};
            fnormalize(e);
            bool found = false;
            for (size_t p = 0; p < E.size() / 3 && !found; p++) {
                const float *u = &E[3 * p];
                bool same = std::fabs(u[0] - e[0]) <= 1e-6f && std::fabs(u[1] - e[1]) <= 1e-6f &&
                            std::fabs(u[2] - e[2]) <= 1e-6f;
                bool opp = std::fabs(u[0] + e[0]) <= 1e-6f && std::fabs(u[1] + e[1]) <= 1e-6f &&
                           std::fabs(u[2] + e[2]) <= 1e-6f;
                found = same || opp;
            }
            if (!found) { E.push_back(e[0]); E.push_back(e[1]); E.push_back(e[2]); }
        }
    }
    out.verts = V;
    out.edges = E;
    for (auto &f : faces) {
        out.face_index.push_back((int)out.indices.size());
        out.face_index.push_back((int)f.idx.size());
        for (int v : f.idx) out.indices.push_back(v);
        for (int k = 0; k < 4; k++) out.planes.push_back(f.plane[k]);
    }
    // vertex-average centre (CLPhysicsDemo.cpp:380-385) and local AABB (:397-406)
    float c[3] = {0, 0, 0};
    int nv = (int)(V.size() / 3);
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) c[k] += V[3 * i + k];
    float inv = 1.f / (float)nv;
    for (int k = 0; k < 3; k++) out.center[k] = c[k] * inv;
    // inscribed-ball radius about the centre from the hull's own triangles
    // (merged faces can lean up to ~2.6 degrees off their triangles, so their
    // planes may lie outside part of the hull)
    {
        const V3 cc = {out.center[0], out.center[1], out.center[2]};
        double rin = 1e300;
        for (auto &t : tris) {
            V3 a, b, d;
            const float *pa = &V[3 * remap[t.v[0]]], *pb = &V[3 * remap[t.v[1]]],
                        *pc = &V[3 * remap[t.v[2]]];
            a = {pa[0], pa[1], pa[2]}; b = {pb[0], pb[1], pb[2]}; d = {pc[0], pc[1], pc[2]};
            V3 nn = cross(sub(b, a), sub(d, a));
            const double l = len(nn);
            if (!(l > 0.0)) continue;
            double dist = dot(nn, sub(a, cc)) / l;
            rin = std::min(rin, std::fabs(dist));
        }
        out.inradius = rin < 1e300 ? rin : 0.0;
    }
    for (int k = 0; k < 3; k++) { out.aabb_min[k] = 1e30f; out.aabb_max[k] = -1e30f; }
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) {
            out.aabb_min[k] = std::min(out.aabb_min[k], V[3 * i + k]);
            out.aabb_max[k] = std::max(out.aabb_max[k], V[3 * i + k]);
        }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_hull_build(const float *xyz, int stride_bytes, int n,
                                         const float scale[3], float *verts4, int max_verts,
                                         float *planes4, int *face_index2, int max_faces,
                                         int *indices, int max_indices, float *edges4,
                                         int max_edges, int counts[4]) {
    HullData h;
    int st = rb_build_hull(xyz, stride_bytes, n, scale, h);
    if (st != MI_RB_OK) return (mi_rb_status)st;
    int nv = (int)h.verts.size() / 3, nf = (int)h.planes.size() / 4;
    int ni = (int)h.indices.size(), ne = (int)h.edges.size() / 3;
    if (counts) { counts[0] = nv; counts[1] = nf; counts[2] = ni; counts[3] = ne; }
    if (nv > max_verts || nf > max_faces || ni > max_indices || ne > max_edges) return MI_RB_CAPACITY;
    for (int i = 0; i < nv; i++) {
        for (int k = 0; k < 3; k++) if (verts4) verts4[4 * i + k] = h.verts[3 * i + k];
        if (verts4) verts4[4 * i + 3] = 0.f;
    }
    if (planes4) std::memcpy(planes4, h.planes.data(), sizeof(float) * 4 * nf);
    if (face_index2) std::memcpy(face_index2, h.face_index.data(), sizeof(int) * 2 * nf);
    if (indices) std::memcpy(indices, h.indices.data(), sizeof(int) * ni);
    for (int i = 0; i < ne; i++) {
        for (int k = 0; k < 3; k++) if (edges4) edges4[4 * i + k] = h.edges[3 * i + k];
        if (edges4) edges4[4 * i + 3] = 0.f;
    }
    return MI_RB_OK;
}
