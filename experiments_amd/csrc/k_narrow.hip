// k_narrow.hip -- convex-hull narrowphase for gfx950: one wavefront per pair.
//
// Per pair, in one kernel (the reference needs findSeparatingAxisKernel,
// sat.cl:928-1261, then clipHullHullKernel, satClipHullContacts.cl:1019-1181,
// with a host sync and readbacks in between):
//   1. SAT: the lanes split the face axes of A and B, then the edge x edge
//      axes; each axis projects both hulls' local vertices (staged in LDS)
//      exactly as project() does (sat.cl:227-250).  Face axes can separate;
//      edge axes never reject (sat.cl:599-615).  The minimum-overlap axis is
//      the first minimum in the reference's order (faces A, faces B, edges),
//      found by a lexicographic (depth, axis index) wave reduction.
//   2. Clipping: the incident face of B (max dot) in world space is clipped
//      against the side planes of A's reference face (min dot); each clip
//      plane is one wave-parallel Sutherland-Hodgman step: lane v emits the
//      0-2 output vertices of input edge (v-1, v), placed by a ballot prefix,
//      which reproduces clipFace's serial output order (:303-346).
//   3. Points with depth <= threshold are kept and reduced to <= 4 with the
//      signed-integer extremes of extractManifoldSequential (:788-884).
// The manifold is written to the pair's own slot, so the contact list is
// ordered by pair index after compaction (no append atomics).
#include <float.h>
#include <hip/hip_runtime.h>

#include "dmath.h"
#include "rb_internal.h"

using namespace rb;

namespace {

// lexicographic (value, index) reductions over the wave
__device__ __forceinline__ void wave_argmin(float &d, int &idx) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        float od = __shfl_xor(d, o, 64);
        int oi = __shfl_xor(idx, o, 64);
        if (od < d || (od == d && oi < idx)) { d = od; idx = oi; }
    }
}
__device__ __forceinline__ void wave_argmax(float &d, int &idx) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        float od = __shfl_xor(d, o, 64);
        int oi = __shfl_xor(idx, o, 64);
        if (od > d || (od == d && oi < idx)) { d = od; idx = oi; }
    }
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        int ov = __shfl_xor(v, o, 64);
        v = ov > v ? ov : v;
    }
    return v;
}

// project(): min/max of the hull's local vertices along dir (world), plus dot(pos, dir)
// The vertex list in LDS is padded to a multiple of 8 with copies of its last
// vertex (which leaves min and max unchanged), so the loop runs 8 independent
// min/max chains and keeps 8 LDS reads in flight.
__device__ __forceinline__ void project(const float4 *sv, int nv8, float4 pos, float4 orn,
                                        float4 dir, float &mn, float &mx) {
    float lo[8], hi[8];
#pragma unroll
    for (int k = 0; k < 8; k++) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    float4 ld = qt_inv_rotate(orn, dir);
    float off = dot3(pos, dir);
    for (int i = 0; i < nv8; i += 8) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float dp = dot3(sv[i + k], ld);
            lo[k] = dp < lo[k] ? dp : lo[k];
            hi[k] = dp > hi[k] ? dp : hi[k];
        }
    }
#pragma unroll
    for (int k = 1; k < 8; k++) {
        lo[0] = lo[k] < lo[0] ? lo[k] : lo[0];
        hi[0] = hi[k] > hi[0] ? hi[k] : hi[0];
    }
    float l = lo[0], h = hi[0];
    if (l > h) { float t = l; l = h; h = t; }
    mn = l + off;
    mx = h + off;
}

__device__ __forceinline__ int pad8(int n) { return (n + 7) & ~7; }

constexpr int NP_BLOCK = 64;

__global__ __launch_bounds__(NP_BLOCK) void k_narrowphase(
    int max_pairs, const int *__restrict__ num_pairs_ptr, const int2 *__restrict__ pairs,
    const float4 *__restrict__ pos, const float4 *__restrict__ quat,
    const float *__restrict__ inv_mass, const int *__restrict__ shape,
    const HullInfo *__restrict__ hulls, const float4 *__restrict__ verts,
    const float4 *__restrict__ planes, const int2 *__restrict__ face_index,
    const int *__restrict__ indices, const float4 *__restrict__ edges, float friction,
    float max_dist, Contact4 *__restrict__ slots, uint32_t *__restrict__ has_contact,
    int *overflow) {
    __shared__ float4 sVA[RB_MAX_HULL_VERTS], sVB[RB_MAX_HULL_VERTS];
    __shared__ float4 sEA[RB_MAX_HULL_EDGES], sEB[RB_MAX_HULL_EDGES];
    __shared__ float4 sPoly[2][RB_MAX_POLY];
    __shared__ float4 sPts[RB_MAX_POLY];

    const int lane = threadIdx.x;
    int npairs = *num_pairs_ptr;
    if (npairs > max_pairs) npairs = max_pairs;

    for (int p = blockIdx.x; p < npairs; p += gridDim.x) {
        int2 pr = pairs[p];
        int a = pr.x, b = pr.y;
        float imA = inv_mass[a], imB = inv_mass[b];
        if (imA == 0.f && imB == 0.f) {
            if (lane == 0) has_contact[p] = 0;
            continue;
        }
        const HullInfo hA = hulls[shape[a]];
        const HullInfo hB = hulls[shape[b]];
        if (hA.num_vertices > RB_MAX_HULL_VERTS || hB.num_vertices > RB_MAX_HULL_VERTS ||
            hA.num_edges > RB_MAX_HULL_EDGES || hB.num_edges > RB_MAX_HULL_EDGES ||
            hA.num_faces > 2 * NP_BLOCK || hB.num_faces > 2 * NP_BLOCK) {
            if (lane == 0) {
                has_contact[p] = 0;
                atomicOr(overflow, (int)MI_RB_OVERFLOW_POLYGON);
            }
            continue;
        }
        float4 pA = pos[a], pB = pos[b], oA = quat[a], oB = quat[b];
        pA.w = 0.f;
        pB.w = 0.f;
        __syncthreads();
        for (int i = lane; i < pad8(hA.num_vertices); i += NP_BLOCK)
            sVA[i] = verts[hA.vertex_offset + (i < hA.num_vertices ? i : hA.num_vertices - 1)];
        for (int i = lane; i < pad8(hB.num_vertices); i += NP_BLOCK)
            sVB[i] = verts[hB.vertex_offset + (i < hB.num_vertices ? i : hB.num_vertices - 1)];
        for (int i = lane; i < hA.num_edges; i += NP_BLOCK)
            sEA[i] = qt_rotate(oA, edges[hA.edge_offset + i]);
        for (int i = lane; i < hB.num_edges; i += NP_BLOCK)
            sEB[i] = qt_rotate(oB, edges[hB.edge_offset + i]);
        __syncthreads();

        float4 dc = sub4(xform(hA.center, pA, oA), xform(hB.center, pB, oB));

        // ---- SAT, face axes ----
        float best = FLT_MAX;
        int best_ax = 0x7fffffff;
        int sep_found = 0;
        const int nf = hA.num_faces + hB.num_faces;
        for (int ax = lane; ax < nf; ax += NP_BLOCK) {
            bool onA = ax < hA.num_faces;
            float4 nrm = onA ? planes[hA.face_offset + ax] : planes[hB.face_offset + ax - hA.num_faces];
            float4 w = qt_rotate(onA ? oA : oB, nrm);
            if (dot3(dc, w) < 0) w = scl4(w, -1.f);
            float mn0, mx0, mn1, mx1;
            if (onA) {
                project(sVA, pad8(hA.num_vertices), pA, oA, w, mn0, mx0);
                project(sVB, pad8(hB.num_vertices), pB, oB, w, mn1, mx1);
            } else {
                project(sVB, pad8(hB.num_vertices), pB, oB, w, mn0, mx0);
                project(sVA, pad8(hA.num_vertices), pA, oA, w, mn1, mx1);
            }
            if (mx0 < mn1 || mx1 < mn0) {
                sep_found = 1;
            } else {
                float d0 = mx0 - mn1, d1 = mx1 - mn0;
                float d = d0 < d1 ? d0 : d1;
                if (d < best) { best = d; best_ax = ax; }
            }
        }
        if (__ballot(sep_found)) {
            if (lane == 0) has_contact[p] = 0;
            continue;
        }
        // ---- SAT, edge x edge axes (never reject) ----
        const int ne = hA.num_edges * hB.num_edges;
        for (int k = lane; k < ne; k += NP_BLOCK) {
            int e0 = k / hB.num_edges, e1 = k - e0 * hB.num_edges;
            float4 c = cross3(sEA[e0], sEB[e1]);
            if (fabsf(c.x) > 1e-6f || fabsf(c.y) > 1e-6f || fabsf(c.z) > 1e-6f) {
                c = normalize3(c);
                if (dot3(dc, c) < 0) c = scl4(c, -1.f);
                float mn0, mx0, mn1, mx1;
                project(sVA, pad8(hA.num_vertices), pA, oA, c, mn0, mx0);
                project(sVB, pad8(hB.num_vertices), pB, oB, c, mn1, mx1);
                float d0 = mx0 - mn1, d1 = mx1 - mn0;
                float dist = d0 < d1 ? d0 : d1;
                if (dist < best) { best = dist; best_ax = nf + k; }
            }
        }
        wave_argmin(best, best_ax);
        if (best_ax == 0x7fffffff) {
            if (lane == 0) has_contact[p] = 0;
            continue;
        }
        float4 sep;
        if (best_ax < hA.num_faces) {
            sep = qt_rotate(oA, planes[hA.face_offset + best_ax]);
            if (dot3(dc, sep) < 0) sep = scl4(sep, -1.f);
        } else if (best_ax < nf) {
            sep = qt_rotate(oB, planes[hB.face_offset + best_ax - hA.num_faces]);
            if (dot3(dc, sep) < 0) sep = scl4(sep, -1.f);
        } else {
            int k = best_ax - nf;
            int e0 = k / hB.num_edges, e1 = k - e0 * hB.num_edges;
            sep = normalize3(cross3(sEA[e0], sEB[e1]));
            if (dot3(dc, sep) < 0) sep = scl4(sep, -1.f);
        }
        if (dot3(neg4(dc), sep) > 0.0f) sep = neg4(sep);

        // ---- clipping: incident face of B ----
        float dB = -FLT_MAX;
        int fB = 0x7fffffff;
        for (int f = lane; f < hB.num_faces; f += NP_BLOCK) {
            float4 pl = planes[hB.face_offset + f];
            float d = dot3(qt_rotate(oB, mk4(pl.x, pl.y, pl.z, 0.f)), sep);
            if (d > dB) { dB = d; fB = f; }
        }
        wave_argmax(dB, fB);
        float dA = FLT_MAX;
        int fA = 0x7fffffff;
        for (int f = lane; f < hA.num_faces; f += NP_BLOCK) {
            float4 pl = planes[hA.face_offset + f];
            float d = dot3(qt_rotate(oA, mk4(pl.x, pl.y, pl.z, 0.f)), sep);
            if (d < dA) { dA = d; fA = f; }
        }
        wave_argmin(dA, fA);
        if (fB == 0x7fffffff || fA == 0x7fffffff) {
            if (lane == 0) has_contact[p] = 0;
            continue;
        }
        int2 fiB = face_index[hB.face_offset + fB];
        int nin = fiB.y < RB_MAX_POLY ? fiB.y : RB_MAX_POLY;
        for (int e = lane; e < nin; e += NP_BLOCK)
            sPoly[0][e] = xform(verts[hB.vertex_offset + indices[fiB.x + e]], pB, oB);
        int2 fiA = face_index[hA.face_offset + fA];
        float4 plA = planes[hA.face_offset + fA];
        float4 wpn = qt_rotate(oA, mk4(plA.x, plA.y, plA.z, 0.f));
        int cur = 0;
        bool poly_over = false;
        __syncthreads();
        for (int e0 = 0; e0 < fiA.y; e0++) {
            float4 va = verts[hA.vertex_offset + indices[fiA.x + e0]];
            float4 vb = verts[hA.vertex_offset + indices[fiA.x + (e0 + 1) % fiA.y]];
            float4 we = qt_rotate(oA, sub4(va, vb));
            float4 pn = neg4(cross3(we, wpn));
            float4 wa = xform(va, pA, oA);
            float eq = -dot3(wa, pn);
            const float4 *in = sPoly[cur];
            float4 *outp = sPoly[cur ^ 1];
            // one lane per input vertex (nin <= 64)
            int v = lane;
            int cnt = 0;
            float4 o0 = mk4(0, 0, 0, 0), o1 = o0;
            if (v < nin) {
                float4 first = in[v == 0 ? nin - 1 : v - 1];
                float4 end = in[v];
                float ds = dot3(pn, first) + eq;
                float de = dot3(pn, end) + eq;
                if (ds < 0) {
                    if (de < 0) {
                        o0 = end; cnt = 1;
                    } else {
                        float t = ds * 1.f / (ds - de);
                        o0 = mk4(first.x + (end.x - first.x) * t, first.y + (end.y - first.y) * t,
                                 first.z + (end.z - first.z) * t, 0.f);
                        cnt = 1;
                    }
                } else if (de < 0) {
                    float t = ds * 1.f / (ds - de);
                    o0 = mk4(first.x + (end.x - first.x) * t, first.y + (end.y - first.y) * t,
                             first.z + (end.z - first.z) * t, 0.f);
                    o1 = end;
                    cnt = 2;
                }
            }
            // exclusive prefix of cnt over lanes
            uint64_t b1 = __ballot(cnt >= 1), b2 = __ballot(cnt == 2);
            uint64_t lt = lanemask_lt();
            int off = __popcll(b1 & lt) + __popcll(b2 & lt);
            int total = __popcll(b1) + __popcll(b2);
            if (total > RB_MAX_POLY) poly_over = true;
            if (cnt >= 1 && off < RB_MAX_POLY) outp[off] = o0;
            if (cnt == 2 && off + 1 < RB_MAX_POLY) outp[off + 1] = o1;
            __syncthreads();
            nin = total < RB_MAX_POLY ? total : RB_MAX_POLY;
            cur ^= 1;
        }
        if (poly_over && lane == 0) atomicOr(overflow, (int)MI_RB_OVERFLOW_POLYGON);
        // keep points behind A's reference face
        float peq = plA.w - dot3(wpn, pA);
        int keep = 0;
        float4 kp = mk4(0, 0, 0, 0);
        if (lane < nin) {
            float4 q = sPoly[cur][lane];
            float depth = dot3(wpn, q) + peq;
            if (depth <= -1e30f) depth = -1e30f;
            if (depth <= max_dist) { keep = 1; kp = mk4(q.x, q.y, q.z, depth); }
        }
        uint64_t kb = __ballot(keep);
        int npts = __popcll(kb);
        if (keep) sPts[__popcll(kb & lanemask_lt())] = kp;
        __syncthreads();
        if (npts == 0) {
            if (lane == 0) has_contact[p] = 0;
            continue;
        }
        // ---- reduction to <= 4 points ----
        float4 nrm = neg4(sep);
        int idx[4] = {0, 1, 2, 3};
        int nr = npts;
        if (npts >= 4) {
            float4 c = sPts[0];
            for (int j = 1; j < npts; j++) c = add4(c, sPts[j]);
            float fn = (float)npts;
            c = mk4(c.x / fn, c.y / fn, c.z / fn, c.w / fn);
            float4 av = sub4(sPts[0], c);
            float4 u = cross3(nrm, av);
            float4 v = cross3(nrm, u);
            u = normalize3(u);
            v = normalize3(v);
            int key[4] = {INT_MIN, INT_MIN, INT_MIN, INT_MIN};
            if (lane < npts) {
                float4 r = sub4(sPts[lane], c);
                float f[4] = {dot3(u, r), dot3(neg4(u), r), dot3(v, r), dot3(neg4(v), r)};
                for (int k = 0; k < 4; k++)
                    key[k] = (int)((__float_as_uint(f[k]) & 0xffffff00u) | (0xffu & (uint32_t)lane));
            }
            for (int k = 0; k < 4; k++) idx[k] = wave_max_i32(key[k]) & 0xff;
            nr = 4;
        }
        if (lane == 0) {
            Contact4 ct;
            ct.normal = mk4(nrm.x, nrm.y, nrm.z, (float)nr);
            for (int i = 0; i < 4; i++) ct.pos[i] = i < nr ? sPts[idx[i]] : mk4(0, 0, 0, 0);
            ct.coeffs = (uint32_t)(0.f * 0xffff) | ((uint32_t)(friction * 0xffff) << 16);
            ct.pair_index = p;
            ct.body_a = imA == 0.f ? -a : a;
            ct.body_b = imB == 0.f ? -b : b;
            slots[p] = ct;
            has_contact[p] = 1;
        }
    }
}

}  // namespace

hipError_t launch_narrowphase(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    int grid = rb_num_cus() * 16;
    hipLaunchKernelGGL(k_narrowphase, dim3(grid), dim3(NP_BLOCK), 0, s, p.max_pairs,
                       d.counters + RB_CNT_PAIRS, d.pairs, d.pos, d.quat, d.inv_mass, d.shape,
                       d.hulls, d.verts, d.planes, d.face_index, d.indices, d.edges, p.friction,
                       p.threshold, d.contact_slots, d.has_contact, d.counters + RB_CNT_OVERFLOW);
    return hipGetLastError();
}
