// ref_clip_driver.cpp -- TEST INFRASTRUCTURE.  Builds the reference's own CPU
// emulation of the hull clipping kernels
// (opencl/gpu_rigidbody_pipeline2/satClipHullContacts.cpp, compiled from the
// reference checkout by oracle/Makefile into oracle/_ref/) and exposes one C
// entry point, so tests can compare the oracle's restated clipping with the
// reference code on arbitrary hull pairs.  Never part of the product.
#include "satClipHullContacts.cpp"

#include "rb_oracle.h"

extern "C" int ref_clip_hull_hull(const float *sep, const orc_hull *hA, const orc_hull *hB,
                                  const float *verts, int nverts, const float *planes,
                                  const int *face_index, int nfaces, const int *indices,
                                  int nindices, const float *posA, const float *ornA,
                                  const float *posB, const float *ornB, float min_dist,
                                  float max_dist, float *out /* 64 float4 */) {
    btAlignedObjectArray<float4> V;
    btAlignedObjectArray<btGpuFace> F;
    btAlignedObjectArray<int> I;
    V.resize(nverts);
    for (int i = 0; i < nverts; i++) V[i] = make_float4(verts[4 * i], verts[4 * i + 1], verts[4 * i + 2], 0.f);
    F.resize(nfaces);
    for (int i = 0; i < nfaces; i++) {
        F[i].m_plane = make_float4(planes[4 * i], planes[4 * i + 1], planes[4 * i + 2], planes[4 * i + 3]);
        F[i].m_indexOffset = face_index[2 * i];
        F[i].m_numIndices = face_index[2 * i + 1];
    }
    I.resize(nindices);
    for (int i = 0; i < nindices; i++) I[i] = indices[i];
    ConvexPolyhedronCL cA, cB;
    memset(&cA, 0, sizeof(cA));
    memset(&cB, 0, sizeof(cB));
    cA.m_faceOffset = hA->face_offset; cA.m_numFaces = hA->num_faces;
    cA.m_vertexOffset = hA->vertex_offset; cA.m_numVertices = hA->num_vertices;
    cB.m_faceOffset = hB->face_offset; cB.m_numFaces = hB->num_faces;
    cB.m_vertexOffset = hB->vertex_offset; cB.m_numVertices = hB->num_vertices;
    float4 s = make_float4(sep[0], sep[1], sep[2], sep[3]);
    float4 pA = make_float4(posA[0], posA[1], posA[2], 0.f), pB = make_float4(posB[0], posB[1], posB[2], 0.f);
    float4 oA = make_float4(ornA[0], ornA[1], ornA[2], ornA[3]);
    float4 oB = make_float4(ornB[0], ornB[1], ornB[2], ornB[3]);
    float4 b1[64], b2[64], pts[64];
    int n = clipHullAgainstHull(s, &cA, &cB, pA, oA, pB, oB, b1, b2, 64, min_dist, max_dist, V, F,
                                I, pts, 64);
    for (int i = 0; i < n; i++) {
        out[4 * i] = pts[i].x; out[4 * i + 1] = pts[i].y;
        out[4 * i + 2] = pts[i].z; out[4 * i + 3] = pts[i].w;
    }
    return n;
}
