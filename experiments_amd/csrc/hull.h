// hull.h -- host convex hull builder used at shape registration.
#pragma once
#include <vector>

struct HullData {
    std::vector<float> verts;      // xyz per vertex
    std::vector<float> planes;     // nx ny nz w per face
    std::vector<int> face_index;   // (first index, count) per face, relative to this hull
    std::vector<int> indices;      // vertex indices per face, relative to this hull
    std::vector<float> edges;      // unique unit edge directions, xyz
    float center[3] = {0, 0, 0};   // vertex average
    double inradius = 0.0;         // smallest centre-to-facet distance over the hull triangles
    float aabb_min[3] = {0, 0, 0};
    float aabb_max[3] = {0, 0, 0};
};

// Returns an mi_rb_status value.
int rb_build_hull(const float *xyz, int stride_bytes, int n, const float scale[3], HullData &out);
// Features given as they are (validated: indices in range, finite values).
int rb_hull_from_features(const float *verts, int vert_stride, int nv, const float *planes4,
                          const int *face_index2, int nf, const int *indices, int ni,
                          const float *edges, int edge_stride, int ne, HullData &out);
