// hull.cpp -- convex hull and polyhedral features at shape registration (host).
//
// Replaces btConvexHullComputer + btConvexUtility::initializePolyhedralFeatures /
// initialize (opencl/gpu_rigidbody_pipeline/btConvexUtility.cpp:31-510) as used by
// CLPhysicsDemo::registerConvexPolyhedron (opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.cpp:491-530):
//   * the points snapped to the integer grid btConvexHullComputer works on
//     (btConvexHullComputer.cpp:1940-2082), and the hull of the grid points as
//     outward-wound triangles (incremental hull with exact predicates);
//   * the face plane of each triangle from its first two normalized edges, with
//     plane w = -min over the face's vertices of dot(v, n) (btConvexUtility.cpp:73-113);
//   * coplanar triangles (normal dot > 0.999) merged into one polygon ordered
//     counter-clockwise about the face normal, unless a dropped vertex is used
//     by another face (btConvexUtility.cpp:117-268);
//   * unique edge directions, normalized, one per +-direction within 1e-6
//     (btConvexUtility.cpp:339-390).
// Vertices keep the order in which they appear in the input; the reference
// lists them in the order btConvexHullComputer walks its edge graph, so
// vertex, face and unique-edge order (and the sign of an edge direction)
// differ from the reference while the sets are the same.
// mi_rb_register_convex_polyhedron takes features built elsewhere (the
// reference's own, in its order) unchanged.
#include <algorithm>
#include <cmath>
#include <cstdint>
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

// Incremental convex hull of integer-valued points; returns outward-wound
// triangles over point indices.  Every predicate is an exact sign of an
// integer determinant (the coordinates are small enough for doubles to hold
// every product), so points on a face's plane are never taken as outside it.
bool build_hull(const std::vector<V3> &P, std::vector<Tri> &tris) {
    int n = (int)P.size();
    if (n < 4) return false;
    auto same = [](V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    auto nonzero = [](V3 a) { return a.x != 0.0 || a.y != 0.0 || a.z != 0.0; };
    // initial tetrahedron
    int i0 = 0, i1 = -1, i2 = -1, i3 = -1;
    for (int i = 1; i < n; i++)
        if (!same(P[i], P[i0])) { i1 = i; break; }
    if (i1 < 0) return false;
    for (int i = 1; i < n; i++)
        if (nonzero(cross(sub(P[i1], P[i0]), sub(P[i], P[i0])))) { i2 = i; break; }
    if (i2 < 0) return false;
    V3 nrm = cross(sub(P[i1], P[i0]), sub(P[i2], P[i0]));
    for (int i = 1; i < n; i++)
        if (dot(nrm, sub(P[i], P[i0])) != 0.0) { i3 = i; break; }
    if (i3 < 0) return false;
    auto make = [&](int a, int b, int c) {
        Tri t;
        t.v[0] = a; t.v[1] = b; t.v[2] = c;
        t.n = cross(sub(P[b], P[a]), sub(P[c], P[a]));
        t.d = dot(t.n, P[a]);
        t.alive = true;
        return t;
    };
    // four times the tetrahedron's centroid (integer)
    V3 cen4 = {P[i0].x + P[i1].x + P[i2].x + P[i3].x, P[i0].y + P[i1].y + P[i2].y + P[i3].y,
               P[i0].z + P[i1].z + P[i2].z + P[i3].z};
    int q[4] = {i0, i1, i2, i3};
    int faces[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    tris.clear();
    for (auto &f : faces) {
        Tri t = make(q[f[0]], q[f[1]], q[f[2]]);
        if (dot(t.n, cen4) - 4.0 * t.d > 0) t = make(q[f[0]], q[f[2]], q[f[1]]);
        tris.push_back(t);
    }
    for (int p = 0; p < n; p++) {
        if (p == i0 || p == i1 || p == i2 || p == i3) continue;
        std::vector<int> vis;
        for (int t = 0; t < (int)tris.size(); t++)
            if (tris[t].alive && dot(tris[t].n, P[p]) - tris[t].d > 0.0) vis.push_back(t);
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

// btConvexHullInternal::compute (bullet2/LinearMath/btConvexHullComputer.cpp:1940-2035)
// builds the hull of integer points: every axis of the points' AABB is cut
// into 10216 steps about the AABB centre (the step's sign flipped for one
// axis order) and each coordinate is truncated to a whole number of steps;
// the vertices it returns are those grid points mapped back as
// step * integer + centre (getCoordinates, :2075-2082).  The same grid is
// used here, in the same float operations, so registered vertices equal the
// reference's bit for bit and the hull topology is decided exactly on the
// integers.
struct HullGrid {
    float scaling[3], inv[3], center[3];
};
static HullGrid hull_grid(const std::vector<float> &p, int n) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            float v = p[3 * i + k];
            if (v < mn[k]) mn[k] = v;
            if (mx[k] < v) mx[k] = v;
        }
    float s[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    int max_axis = s[0] < s[1] ? (s[1] < s[2] ? 2 : 1) : (s[0] < s[2] ? 2 : 0);
    int min_axis = s[0] < s[1] ? (s[0] < s[2] ? 0 : 2) : (s[1] < s[2] ? 1 : 2);
    if (min_axis == max_axis) min_axis = (max_axis + 1) % 3;
    int med_axis = 3 - max_axis - min_axis;
    const float step = 1.0f / 10216.0f;
    HullGrid g;
    for (int k = 0; k < 3; k++) {
        s[k] *= step;
        if ((med_axis + 1) % 3 != max_axis) s[k] *= -1.0f;
        g.scaling[k] = s[k];
        g.inv[k] = s[k] != 0.f ? 1.0f / s[k] : s[k];
        g.center[k] = (mn[k] + mx[k]) * 0.5f;
    }
    return g;
}

int rb_build_hull(const float *xyz, int stride_bytes, int n, const float scale[3], HullData &out) {
    out = HullData();
    if (!xyz || n < 4 || stride_bytes < 12) return MI_RB_INVALID_ARG;
    const unsigned char *base = reinterpret_cast<const unsigned char *>(xyz);
    float s[3] = {1.f, 1.f, 1.f};
    if (scale) { s[0] = scale[0]; s[1] = scale[1]; s[2] = scale[2]; }
    std::vector<float> raw(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const float *v = reinterpret_cast<const float *>(base + (size_t)i * stride_bytes);
        raw[3 * i] = v[0] * s[0]; raw[3 * i + 1] = v[1] * s[1]; raw[3 * i + 2] = v[2] * s[2];
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(raw[3 * i + k])) return MI_RB_INVALID_ARG;
    }
    // grid points: P holds the integer coordinates (exact in double, |x| <=
    // 5108 per axis, so every orientation test below is exact), inF the
    // vertex each maps back to.  The integer axes are given the sign of their
    // step so that P has the orientation of the float space.
    const HullGrid g = hull_grid(raw, n);
    std::vector<V3> P(n);
    std::vector<float> inF(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        double q[3];
        for (int k = 0; k < 3; k++) {
            float t = (raw[3 * i + k] - g.center[k]) * g.inv[k];
            int32_t c = (int32_t)t;
            inF[3 * i + k] = (float)c * g.scaling[k] + g.center[k];
            q[k] = g.scaling[k] < 0.f ? -(double)c : (double)c;
        }
        P[i] = {q[0], q[1], q[2]};
    }
    std::vector<Tri> tris;
    if (!build_hull(P, tris)) return MI_RB_INVALID_ARG;
    // hull vertices in input order (later points on the same grid point are
    // dropped)
    std::vector<int> used(n, 0), remap(n, -1);
    for (auto &t : tris)
        for (int k = 0; k < 3; k++) used[t.v[k]] = 1;
    std::vector<float> V;
    for (int i = 0; i < n; i++)
        if (used[i]) {
            remap[i] = (int)(V.size() / 3);
            V.push_back(inF[3 * i]); V.push_back(inF[3 * i + 1]); V.push_back(inF[3 * i + 2]);
        }
    // btConvexHullComputer's faces are the hull's exact planar polygons:
    // triangles that share an edge and lie in one plane of the integer grid
    // are joined, and each polygon is the loop of its boundary edges
    const int nt = (int)tris.size();
    std::vector<int> group(nt);
    for (int t = 0; t < nt; t++) group[t] = t;
    auto root = [&](int t) {
        while (group[t] != t) t = group[t] = group[group[t]];
        return t;
    };
    for (int t = 0; t < nt; t++)
        for (int u = t + 1; u < nt; u++) {
            int shared = 0, other = -1;
            for (int k = 0; k < 3; k++) {
                bool in_t = tris[t].v[0] == tris[u].v[k] || tris[t].v[1] == tris[u].v[k] ||
                            tris[t].v[2] == tris[u].v[k];
                if (in_t) shared++; else other = tris[u].v[k];
            }
            if (shared == 2 && dot(tris[t].n, P[other]) - tris[t].d == 0.0)
                group[root(u)] = root(t);
        }
    std::vector<FaceF> tmp;
    std::vector<int> done(nt, 0);
    for (int t = 0; t < nt; t++) {
        if (done[t]) continue;
        const int r = root(t);
        std::vector<std::pair<int, int>> dir;
        for (int u = t; u < nt; u++)
            if (root(u) == r) {
                done[u] = 1;
                for (int k = 0; k < 3; k++) dir.push_back({tris[u].v[k], tris[u].v[(k + 1) % 3]});
            }
        std::vector<std::pair<int, int>> boundary;
        for (auto &e : dir) {
            bool twin = false;
            for (auto &f : dir)
                if (f.first == e.second && f.second == e.first) { twin = true; break; }
            if (!twin) boundary.push_back(e);
        }
        FaceF f;
        int v = tris[t].v[0];
        for (auto &e : boundary)
            if (e.first == v) break;
        // start at the first triangle's first vertex when it is on the
        // boundary, else at the first boundary edge
        bool on = false;
        for (auto &e : boundary) on = on || e.first == v;
        if (!on) v = boundary[0].first;
        for (size_t k = 0; k < boundary.size(); k++) {
            f.idx.push_back(remap[v]);
            int next = -1;
            for (auto &e : boundary)
                if (e.first == v) { next = e.second; break; }
            if (next < 0) return MI_RB_INVALID_ARG;
            v = next;
        }
        face_plane(V, f);
        tmp.push_back(f);
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

// Features built elsewhere (a btConvexUtility's vertices, faces and unique
// edges), kept as given; the centre is the vertex average and the AABB that
// of the vertices (CLPhysicsDemo.cpp:361-405), the inscribed radius the
// smallest centre-to-plane distance over the faces' fan triangles.
int rb_hull_from_features(const float *verts, int vert_stride, int nv, const float *planes4,
                          const int *face_index2, int nf, const int *indices, int ni,
                          const float *edges, int edge_stride, int ne, HullData &out) {
    out = HullData();
    if (!verts || !planes4 || !face_index2 || !indices || !edges || nv < 4 || nf < 4 ||
        ni < 3 * nf || ne < 1 || vert_stride < 12 || edge_stride < 12)
        return MI_RB_INVALID_ARG;
    auto at = [](const float *base, int stride, int i) {
        return reinterpret_cast<const float *>(reinterpret_cast<const unsigned char *>(base) +
                                               (size_t)i * stride);
    };
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) {
            float v = at(verts, vert_stride, i)[k];
            if (!std::isfinite(v)) return MI_RB_INVALID_ARG;
            out.verts.push_back(v);
        }
    for (int f = 0; f < nf; f++) {
        int first = face_index2[2 * f], cnt = face_index2[2 * f + 1];
        if (first < 0 || cnt < 3 || first > ni - cnt) return MI_RB_INVALID_ARG;
        for (int k = 0; k < 4; k++) {
            if (!std::isfinite(planes4[4 * f + k])) return MI_RB_INVALID_ARG;
            out.planes.push_back(planes4[4 * f + k]);
        }
        out.face_index.push_back(first);
        out.face_index.push_back(cnt);
    }
    for (int i = 0; i < ni; i++) {
        if (indices[i] < 0 || indices[i] >= nv) return MI_RB_INVALID_ARG;
        out.indices.push_back(indices[i]);
    }
    for (int i = 0; i < ne; i++)
        for (int k = 0; k < 3; k++) {
            float v = at(edges, edge_stride, i)[k];
            if (!std::isfinite(v)) return MI_RB_INVALID_ARG;
            out.edges.push_back(v);
        }
    const std::vector<float> &V = out.verts;
    float c[3] = {0, 0, 0};
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) c[k] += V[3 * i + k];
    float inv = 1.f / (float)nv;
    for (int k = 0; k < 3; k++) out.center[k] = c[k] * inv;
    const V3 cc = {out.center[0], out.center[1], out.center[2]};
    double rin = 1e300;
    for (int f = 0; f < nf; f++) {
        const int first = out.face_index[2 * f], cnt = out.face_index[2 * f + 1];
        const float *p0 = &V[3 * out.indices[first]];
        V3 a = {p0[0], p0[1], p0[2]};
        for (int j = 1; j + 1 < cnt; j++) {
            const float *p1 = &V[3 * out.indices[first + j]], *p2 = &V[3 * out.indices[first + j + 1]];
            V3 b = {p1[0], p1[1], p1[2]}, d = {p2[0], p2[1], p2[2]};
            V3 nn = cross(sub(b, a), sub(d, a));
            const double l = len(nn);
            if (!(l > 0.0)) continue;
            rin = std::min(rin, std::fabs(dot(nn, sub(a, cc)) / l));
        }
    }
    out.inradius = rin < 1e300 ? rin : 0.0;
    for (int k = 0; k < 3; k++) { out.aabb_min[k] = 1e30f; out.aabb_max[k] = -1e30f; }
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) {
            out.aabb_min[k] = std::min(out.aabb_min[k], V[3 * i + k]);
            out.aabb_max[k] = std::max(out.aabb_max[k], V[3 * i + k]);
        }
    return MI_RB_OK;
}
