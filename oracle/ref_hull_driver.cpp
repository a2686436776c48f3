// ref_hull_driver.cpp -- TEST INFRASTRUCTURE.  Builds the reference's own
// shape-feature code (opencl/gpu_rigidbody_pipeline/btConvexUtility.cpp with
// bullet2/LinearMath/btConvexHullComputer.cpp, compiled from the reference
// checkout by oracle/Makefile into oracle/_ref/) and exposes one C entry
// point with the arrays CLPhysicsDemo::registerConvexPolyhedron hands to
// registerConvexHullShape (CLPhysicsDemo.cpp:491-530,
// btGpuNarrowPhaseAndSolver.cpp:423-490), in the reference's own order, so
// tests can compare the library's hull features with them.  Never part of
// the product.
#include "btConvexUtility.h"

extern "C" int ref_hull_features(const float *xyz, int n, const float *scale, float *verts4,
                                 int max_verts, float *planes4, int *face_index2, int max_faces,
                                 int *indices, int max_indices, float *edges4, int max_edges,
                                 int counts[4]) {
    btAlignedObjectArray<btVector3> pts;
    for (int i = 0; i < n; i++)
        pts.push_back(btVector3(xyz[3 * i] * scale[0], xyz[3 * i + 1] * scale[1],
                                xyz[3 * i + 2] * scale[2]));
    btConvexUtility util;
    util.initializePolyhedralFeatures(&pts[0], pts.size(), true);
    int nv = util.m_vertices.size(), nf = util.m_faces.size(), ne = util.m_uniqueEdges.size();
    int ni = 0;
    for (int f = 0; f < nf; f++) ni += util.m_faces[f].m_indices.size();
    counts[0] = nv; counts[1] = nf; counts[2] = ni; counts[3] = ne;
    if (nv > max_verts || nf > max_faces || ni > max_indices || ne > max_edges) return 1;
    for (int i = 0; i < nv; i++) {
        verts4[4 * i] = util.m_vertices[i].getX();
        verts4[4 * i + 1] = util.m_vertices[i].getY();
        verts4[4 * i + 2] = util.m_vertices[i].getZ();
        verts4[4 * i + 3] = 0.f;
    }
    int at = 0;
    for (int f = 0; f < nf; f++) {
        for (int k = 0; k < 4; k++) planes4[4 * f + k] = util.m_faces[f].m_plane[k];
        face_index2[2 * f] = at;
        face_index2[2 * f + 1] = util.m_faces[f].m_indices.size();
        for (int k = 0; k < util.m_faces[f].m_indices.size(); k++)
            indices[at++] = util.m_faces[f].m_indices[k];
    }
    for (int i = 0; i < ne; i++) {
        edges4[4 * i] = util.m_uniqueEdges[i].getX();
        edges4[4 * i + 1] = util.m_uniqueEdges[i].getY();
        edges4[4 * i + 2] = util.m_uniqueEdges[i].getZ();
        edges4[4 * i + 3] = 0.f;
    }
    return 0;
}
