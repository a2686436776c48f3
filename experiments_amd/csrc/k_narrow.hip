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

// Hull vertices are read through the scalar data cache: every lane needs
// every vertex for its own axis and the vertex address is the same on all
// lanes, so a scalar load hands the vertex to the whole wave as SGPR operands.
// (An LDS broadcast of a float4 still moves 1 KB per wave and made the axis
// loops LDS-bound.)  The vertex arrays are read-only for the whole step.
struct Vtx { float x, y, z, w; };
struct Vtx4 { Vtx v[4]; };  // four consecutive vertices: one 64-byte scalar load
typedef __attribute__((address_space(4))) const Vtx cfloat4;
typedef __attribute__((address_space(4))) const Vtx4 cVtx4;

// project(): min/max of the hull's local vertices along dir (world), plus
// dot(pos, dir) (sat.cl:227-250).  Four independent min/max chains over
// groups of four vertices; the last group repeats the last vertex, which
// leaves min and max unchanged.  fminf/fmaxf equal the reference's compare
// and select for every non-NaN input and ignore a NaN projection the same
// way; only the sign of an exact zero extreme may differ, which no overlap
// comparison can see.
__device__ __forceinline__ void project(const cfloat4 *sv, int nv, float4 pos, float4 orn,
                                        float4 dir, float &mn, float &mx) {
    float lo[4], hi[4];
#pragma unroll
    for (int k = 0; k < 4; k++) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    float4 ld = qt_inv_rotate(orn, dir);
    float off = dot3(pos, dir);
    int i = 0;
    for (; i + 4 <= nv; i += 4) {
        const cVtx4 *q = (const cVtx4 *)(sv + i);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float dp = (q->v[k].x * ld.x + q->v[k].y * ld.y) + q->v[k].z * ld.z;
            lo[k] = fminf(lo[k], dp);
            hi[k] = fmaxf(hi[k], dp);
        }
    }
    if (i < nv) {
        const int last = nv - 1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const cfloat4 *v = sv + (i + k < last ? i + k : last);
            float dp = (v->x * ld.x + v->y * ld.y) + v->z * ld.z;
            lo[k] = fminf(lo[k], dp);
            hi[k] = fmaxf(hi[k], dp);
        }
    }
#pragma unroll
    for (int k = 1; k < 4; k++) {
        lo[0] = lo[k] < lo[0] ? lo[k] : lo[0];
        hi[0] = hi[k] > hi[0] ? hi[k] : hi[0];
    }
    float l = lo[0], h = hi[0];
    if (l > h) { float t = l; l = h; h = t; }
    mn = l + off;
    mx = h + off;
}

// project() over a part of the vertices: this lane takes vertices part,
// part + 2^lg, ... (vector loads), and the 2^lg lanes of an aligned group
// combine their minima and maxima.  Min and max are exact, so the result
// equals project()'s for any split.  Every lane of the wave must call it.
__device__ __forceinline__ void project_part(const float4 *__restrict__ v, int nv, int part,
                                             int lg, float4 pos, float4 orn, float4 dir,
                                             float &mn, float &mx) {
    float4 ld = qt_inv_rotate(orn, dir);
    float off = dot3(pos, dir);
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (int i = part; i < nv; i += 1 << lg) {
        const float4 q = v[i];
        float dp = (q.x * ld.x + q.y * ld.y) + q.z * ld.z;
        lo = fminf(lo, dp);
        hi = fmaxf(hi, dp);
    }
    for (int o = 1; o < (1 << lg); o <<= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    if (lo > hi) { float t = lo; lo = hi; hi = t; }
    mn = lo + off;
    mx = hi + off;
}

// log2 of the lanes per axis for a batch of cnt <= 64 axes: 64 / the next
// power of two of cnt
__device__ __forceinline__ int lanes_log2(int cnt) {
    return cnt > 32 ? 0 : cnt > 16 ? 1 : cnt > 8 ? 2 : cnt > 4 ? 3 : cnt > 2 ? 4 : cnt > 1 ? 5 : 6;
}

// Lower bounds of the SAT overlap along a unit axis c oriented with
// dot(dc, c) >= 0, dc = centre A - centre B.  Each hull holds a ball of
// radius rho (HullInfo center.w) around its centre, so max_A(c) - min_B(c)
// >= rhoA + rhoB + dot(dc, c) and max_B(c) - min_A(c) >= rhoA + rhoB -
// dot(dc, c); also max_B(c) - min_A(c) >= dot(sB - sA, c) for any vertex sB of
// B and sA of A, which with the vertices furthest along dc (B) and -dc (A)
// is close to the overlap for axes near dc.  An axis whose bound exceeds a
// known overlap by the rounding margin can neither separate (when that
// overlap is >= 0) nor be the minimum.
__device__ __forceinline__ float sat_margin(float4 pA, float4 pB) {
    return 1e-3f + 4e-6f * ((fabsf(pA.x) + fabsf(pA.y) + fabsf(pA.z)) +
                            (fabsf(pB.x) + fabsf(pB.y) + fabsf(pB.z)));
}
__device__ __forceinline__ float sat_bound(float rho, float4 dc, bool use_dv, float4 dv,
                                           float4 c) {
    const float dd = dot3(dc, c);
    const float lb1 = use_dv ? fmaxf(rho - dd, dot3(dv, c)) : rho - dd;
    return fminf(rho + dd, lb1);
}
// sB - sA in world space (every lane of the wave must call it)
__device__ __forceinline__ float4 support_diff(const float4 *__restrict__ verts,
                                               const HullInfo &hA, const HullInfo &hB,
                                               float4 pA, float4 oA, float4 pB, float4 oB,
                                               float4 dc, int lane) {
    const float4 lb = qt_inv_rotate(oB, dc), la = qt_inv_rotate(oA, neg4(dc));
    float eb = -FLT_MAX, ea = -FLT_MAX;
    int ib = 0, ia = 0;
    for (int v = lane; v < hB.num_vertices; v += 64) {
        const float4 q = verts[hB.vertex_offset + v];
        const float e = (q.x * lb.x + q.y * lb.y) + q.z * lb.z;
        if (e > eb) { eb = e; ib = v; }
    }
    for (int v = lane; v < hA.num_vertices; v += 64) {
        const float4 q = verts[hA.vertex_offset + v];
        const float e = (q.x * la.x + q.y * la.y) + q.z * la.z;
        if (e > ea) { ea = e; ia = v; }
    }
    wave_argmax(eb, ib);
    wave_argmax(ea, ia);
    return sub4(xform(verts[hB.vertex_offset + ib], pB, oB),
                xform(verts[hA.vertex_offset + ia], pA, oA));
}

constexpr int NP_BLOCK = 64;
// wave-max key of a projection: float order as signed integers, the lane in
// the low 6 bits (ties and the dropped bits go to the higher lane); a NaN
// projection ranks below every number and a lane without a vertex below all
__device__ __forceinline__ int probe_key(float v, int lane, bool valid) {
    int f = __float_as_int(v);
    f = f >= 0 ? f : (f ^ 0x7fffffff);
    if (!(v >= -FLT_MAX)) f = INT_MIN + 64;
    if (f < INT_MIN + 64) f = INT_MIN + 64;
    if (!valid) f = INT_MIN;
    return (f & ~63) | lane;
}
// probe directions of the clip pass's second edge-axis bound: dc, four at
// 0.35 and four at 0.8 (rotated by 45 degrees) around it
constexpr int NPROBE = 9;
// the small-hull clip pass (k_narrow_clip_small): SG lanes per candidate for
// hulls of at most SG_VERTS vertices
constexpr int SG = 16;
constexpr int SG_VERTS = 8;
constexpr int SG_EDGES = 18;
constexpr int SG_FACES = 16;
constexpr int SG_POLY = 2 * SG_VERTS;
constexpr int SG_GROUPS = NP_BLOCK / SG;
static_assert(SG_POLY <= SG, "one lane per polygon vertex");
__device__ __forceinline__ bool small_hull(const HullInfo &h) {
    return h.num_vertices <= SG_VERTS && h.num_edges <= SG_EDGES && h.num_faces <= SG_FACES;
}
// LDS written by lanes of this wave, then read by others of it
__device__ __forceinline__ void group_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// the group's bits of a wave ballot, and the group lanes below this one
__device__ __forceinline__ uint32_t group_ballot(bool x, int gbase) {
    return (uint32_t)((__ballot(x) >> gbase) & ((1ull << SG) - 1ull));
}
__device__ __forceinline__ void group_argmin(float &d, int &idx) {
#pragma unroll
    for (int o = SG / 2; o >= 1; o >>= 1) {
        float od = __shfl_xor(d, o, 64);
        int oi = __shfl_xor(idx, o, 64);
        if (od < d || (od == d && oi < idx)) { d = od; idx = oi; }
    }
}
__device__ __forceinline__ void group_argmax(float &d, int &idx) {
#pragma unroll
    for (int o = SG / 2; o >= 1; o >>= 1) {
        float od = __shfl_xor(d, o, 64);
        int oi = __shfl_xor(idx, o, 64);
        if (od > d || (od == d && oi < idx)) { d = od; idx = oi; }
    }
}
__device__ __forceinline__ int group_max_i32(int v) {
#pragma unroll
    for (int o = SG / 2; o >= 1; o >>= 1) {
        int ov = __shfl_xor(v, o, 64);
        v = ov > v ? ov : v;
    }
    return v;
}
// log2 of the lanes per axis for a batch of cnt <= SG axes
__device__ __forceinline__ int group_lanes_log2(int cnt) {
    return cnt > 8 ? 0 : cnt > 4 ? 1 : cnt > 2 ? 2 : cnt > 1 ? 3 : 4;
}

static_assert(RB_MAX_HULL_VERTS <= RB_MAX_POLY, "a face's side planes fit the point buffer");
// face axes up to which pass 1 projects every axis instead of ordering them
// by an overlap bound first
constexpr int FACE_PLAIN_MAX = 32;

__device__ __forceinline__ bool hull_limits_ok(const HullInfo &hA, const HullInfo &hB) {
    return hA.num_vertices <= RB_MAX_HULL_VERTS && hB.num_vertices <= RB_MAX_HULL_VERTS &&
           hA.num_edges <= RB_MAX_HULL_EDGES && hB.num_edges <= RB_MAX_HULL_EDGES &&
           hA.num_faces <= 2 * NP_BLOCK && hB.num_faces <= 2 * NP_BLOCK;
}

// Shapes without vertices: a sphere (no faces; radius in center.w) or a
// plane (one face: normal and constant, n . x = constant).
__device__ __forceinline__ int shape_kind(const HullInfo &h) { return shape_kind_of(h); }

// The contact of a pair with a sphere or a plane: clipHullHullKernel's first
// branches (satClipHullContacts.cl:1055-1118).  Sphere-sphere and
// plane-sphere give one point (the plane body first in a plane-sphere
// contact, as computeContactPlaneConvex is called); a hull against a sphere
// or a plane and a plane against a plane give none, as in the reference.
__device__ bool special_contact(int p, int a, int b, float imA, float imB, const HullInfo &hA,
                                const HullInfo &hB, float4 pA, float4 oA, float4 pB, float4 oB,
                                const float4 *__restrict__ planes, float friction, Contact4 &ct) {
    const int kA = shape_kind(hA), kB = shape_kind(hB);
    ct.coeffs = (uint32_t)(0.f * 0xffff) | ((uint32_t)(friction * 0xffff) << 16);
    ct.pair_index = p;
    for (int i = 1; i < 4; i++) ct.pos[i] = mk4(0, 0, 0, 0);
    if (kA == SHAPE_KIND_SPHERE && kB == SHAPE_KIND_SPHERE) {
        const float rA = hA.center.w, rB = hB.center.w;
        const float4 diff = sub4(pA, pB);
        const float len = sqrtf(dot3(diff, diff));
        if (!(len <= rA + rB)) return false;
        const float dist = len - (rA + rB);
        float4 n = mk4(1.f, 0.f, 0.f, 0.f);
        if (len > 0.00001f) n = mk4(diff.x / len, diff.y / len, diff.z / len, diff.w / len);
        float4 c = add4(pB, scl4(n, rB));
        c.w = dist;
        const float4 nn = neg4(n);
        ct.normal = mk4(nn.x, nn.y, nn.z, 1.f);
        ct.pos[0] = c;
        ct.body_a = imA == 0.f ? -a : a;
        ct.body_b = imB == 0.f ? -b : b;
        return true;
    }
    const bool plane_sphere = kA == SHAPE_KIND_PLANE && kB == SHAPE_KIND_SPHERE;
    const bool sphere_plane = kA == SHAPE_KIND_SPHERE && kB == SHAPE_KIND_PLANE;
    if (!plane_sphere && !sphere_plane) return false;
    // computeContactPlaneConvex(plane body, sphere body)
    const int ip = plane_sphere ? a : b, is = plane_sphere ? b : a;
    const float imP = plane_sphere ? imA : imB, imS = plane_sphere ? imB : imA;
    const HullInfo &hP = plane_sphere ? hA : hB;
    const float radius = (plane_sphere ? hB : hA).center.w;
    const float4 posA1 = plane_sphere ? pA : pB, ornA1 = plane_sphere ? oA : oB;
    const float4 posB1 = plane_sphere ? pB : pA, ornB1 = plane_sphere ? oB : oA;
    const float4 eq = planes[hP.face_offset];
    const float4 n = mk4(eq.x, eq.y, eq.z, 0.f);
    const float pc = eq.w;
    // trInverse / trMul: the sphere in the plane's frame and the plane in the sphere's
    const float4 invOrnA = qt_inv(ornA1), invPosA = qt_rotate(invOrnA, neg4(posA1));
    const float4 cOrn = qt_mul(invOrnA, ornB1), cPos = xform(posB1, invPosA, invOrnA);
    const float4 pOrn = qt_mul(qt_inv(ornB1), ornA1);  // (only the rotation is used)
    const float4 vtx1 = scl4(qt_rotate(pOrn, neg4(n)), radius);
    const float4 vtxInPlane = xform(vtx1, cPos, cOrn);
    const float distance = dot3(n, vtxInPlane) - pc;
    if (!(distance < 0.f)) return false;
    const float4 proj = sub4(vtxInPlane, scl4(n, distance));
    const float4 world = xform(proj, posA1, ornA1);
    const float4 nB = qt_rotate(ornA1, n);
    float4 pOnB = add4(world, scl4(nB, distance));
    pOnB.w = distance;
    ct.normal = mk4(nB.x, nB.y, nB.z, 1.f);
    ct.pos[0] = pOnB;
    ct.body_a = imP == 0.f ? -ip : ip;
    ct.body_b = imS == 0.f ? -is : is;
    return true;
}

// Pairs with a sphere or a plane (launched only for worlds that have such
// shapes, after the face pass, which leaves these pairs without contact):
// one thread per pair.
__global__ __launch_bounds__(256) void k_narrow_special(
    int max_pairs, const int *__restrict__ num_pairs_ptr, const int2 *__restrict__ pairs,
    const float4 *__restrict__ pos, const float4 *__restrict__ quat,
    const float *__restrict__ inv_mass, const int *__restrict__ shape,
    const HullInfo *__restrict__ hulls, const float4 *__restrict__ planes, float friction,
    Contact4 *__restrict__ slots, uint32_t *__restrict__ has_contact) {
    int npairs = *num_pairs_ptr;
    if (npairs > max_pairs) npairs = max_pairs;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x) {
        const int2 pr = pairs[p];
        const HullInfo &hA = hulls[shape[pr.x]], &hB = hulls[shape[pr.y]];
        if (hA.num_vertices != 0 && hB.num_vertices != 0) continue;
        const float imA = inv_mass[pr.x], imB = inv_mass[pr.y];
        if (imA == 0.f && imB == 0.f) continue;
        float4 pA = pos[pr.x], pB = pos[pr.y];
        pA.w = 0.f;
        pB.w = 0.f;
        Contact4 ct;
        if (special_contact(p, pr.x, pr.y, imA, imB, hA, hB, pA, quat[pr.x], pB, quat[pr.y], planes,
                            friction, ct)) {
            slots[p] = ct;
            has_contact[p] = 1;
        }
    }
}

// Pass 0, QL lanes per pair: the face of A that B's centre lies furthest
// outside of, and the face of B that A's centre lies furthest outside of.
// Both axes are built and projected with pass 1's arithmetic, so an axis
// that separates here separates in pass 1 too, and pass 1 (which projects
// every face axis, or leaves out only axes that cannot separate) would give
// the pair no contact: it gets none here.  The other pairs (and every pair
// that is not hull-hull within the limits) are flagged for pass 1 (pairs of
// two small hulls in cand for its SG-lane kernel, the others in has_contact
// for its wave-per-pair kernel), which takes them in pair order from the
// compacted flags and sets cand and has_contact itself (pass 2 for its
// candidates).  Most broadphase pairs are apart, so pass 1 runs on a
// fraction of them.  With QL > 1
// the lanes of a pair split its faces and vertices.
// lanes per pair of pass 0 (4 timed faster than 1 and 8 on the 65,537-hull
// scene: fewer cache lines per load than 1, less repeated work than 8)
constexpr int QUICK_LANES = 4;
template <int QL>
__device__ __forceinline__ void group_argmax(float &d, int &idx) {
#pragma unroll
    for (int o = 1; o < QL; o <<= 1) {
        const float od = __shfl_xor(d, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (od > d || (od == d && oi < idx)) { d = od; idx = oi; }
    }
}
// project() with the group's own hull (vector loads), its vertices split
// over the QL lanes of the group (every lane of the group must call it)
template <int QL>
__device__ __forceinline__ void project_group(const float4 *__restrict__ v, int nv, int sub,
                                              float4 pos, float4 orn, float4 dir, float &mn,
                                              float &mx) {
    const float4 ld = qt_inv_rotate(orn, dir);
    const float off = dot3(pos, dir);
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (int i = sub; i < nv; i += QL) {
        const float4 q = v[i];
        const float dp = (q.x * ld.x + q.y * ld.y) + q.z * ld.z;
        lo = fminf(lo, dp);
        hi = fmaxf(hi, dp);
    }
#pragma unroll
    for (int o = 1; o < QL; o <<= 1) {
        lo = fminf(lo, __shfl_xor(lo, o, 64));
        hi = fmaxf(hi, __shfl_xor(hi, o, 64));
    }
    mn = lo + off;
    mx = hi + off;
}
template <int QL>
__global__ __launch_bounds__(256) void k_narrow_quick(
    int max_pairs, const int *__restrict__ num_pairs_ptr, const int2 *__restrict__ pairs,
    const float4 *__restrict__ pos, const float4 *__restrict__ quat,
    const float *__restrict__ inv_mass, const int *__restrict__ shape,
    const HullInfo *__restrict__ hulls, const float4 *__restrict__ verts,
    const float4 *__restrict__ planes, uint32_t *__restrict__ has_contact,
    uint32_t *__restrict__ cand) {
    int npairs = *num_pairs_ptr;
    if (npairs > max_pairs) npairs = max_pairs;
    const int sub = threadIdx.x & (QL - 1);
    const int p = (int)((blockIdx.x * blockDim.x + threadIdx.x) / QL);
    // a group's lanes share p, so they take the same branches
    if (p >= npairs) return;
    const int2 pr = pairs[p];
    bool sep = false, small = false;
    if (inv_mass[pr.x] != 0.f || inv_mass[pr.y] != 0.f) {
        const HullInfo hA = hulls[shape[pr.x]];
        const HullInfo hB = hulls[shape[pr.y]];
        if (hA.num_vertices > 0 && hB.num_vertices > 0 && hull_limits_ok(hA, hB)) {
            small = small_hull(hA) && small_hull(hB);
            float4 pA = pos[pr.x], pB = pos[pr.y];
            const float4 oA = quat[pr.x], oB = quat[pr.y];
            pA.w = 0.f;
            pB.w = 0.f;
            const float4 cA = xform(hA.center, pA, oA), cB = xform(hB.center, pB, oB);
            const float4 dc = sub4(cA, cB);
            const float4 *pa = planes + hA.face_offset, *pb = planes + hB.face_offset;
            // the faces: largest signed distance of the other centre (local
            // frame); any choice keeps the result exact
            const float4 lB = qt_inv_rotate(oA, sub4(cB, pA));
            const float4 lA = qt_inv_rotate(oB, sub4(cA, pB));
            float sa = -FLT_MAX, sb = -FLT_MAX;
            int fa = 0x7fffffff, fb = 0x7fffffff;
            for (int f = sub; f < hA.num_faces; f += QL) {
                const float4 pl = pa[f];
                const float d = dot3(pl, lB) + pl.w;
                if (d > sa) { sa = d; fa = f; }
            }
            for (int f = sub; f < hB.num_faces; f += QL) {
                const float4 pl = pb[f];
                const float d = dot3(pl, lA) + pl.w;
                if (d > sb) { sb = d; fb = f; }
            }
            group_argmax<QL>(sa, fa);
            group_argmax<QL>(sb, fb);
            if (fa >= hA.num_faces) fa = 0;  // NaN distances: any face will do
            if (fb >= hB.num_faces) fb = 0;
            const float4 *vA = verts + hA.vertex_offset, *vB = verts + hB.vertex_offset;
            float mn0, mx0, mn1, mx1;
            if (hA.num_faces > 0) {
                float4 w = qt_rotate(oA, pa[fa]);
                if (dot3(dc, w) < 0) w = scl4(w, -1.f);
                project_group<QL>(vA, hA.num_vertices, sub, pA, oA, w, mn0, mx0);
                project_group<QL>(vB, hB.num_vertices, sub, pB, oB, w, mn1, mx1);
                sep = mx0 < mn1 || mx1 < mn0;
            }
            if (!sep && hB.num_faces > 0) {
                float4 w = qt_rotate(oB, pb[fb]);
                if (dot3(dc, w) < 0) w = scl4(w, -1.f);
                project_group<QL>(vB, hB.num_vertices, sub, pB, oB, w, mn0, mx0);
                project_group<QL>(vA, hA.num_vertices, sub, pA, oA, w, mn1, mx1);
                sep = mx0 < mn1 || mx1 < mn0;
            }
        }
    }
    // listed pairs: has_contact flags the wave-per-pair face pass, cand the
    // small-hull one (each compacted into its part of the work list)
    if (sub == 0) {
        has_contact[p] = !sep && !small ? 1u : 0u;
        cand[p] = !sep && small ? 1u : 0u;
    }
}

// Pass 1, every pair: face axes of A and B.  A pair separated by a face axis
// (or static-static, or beyond the hull limits) has no contact; the others
// become candidates for pass 2 with their face minimum (depth, axis) and the
// number of edge-edge axes still to test, the cost key of pass 2's order.
__global__ __launch_bounds__(NP_BLOCK) void k_narrow_face(
    int max_pairs, const uint32_t *__restrict__ n_work_ptr, const uint32_t *__restrict__ work,
    const int2 *__restrict__ pairs,
    const float4 *__restrict__ pos, const float4 *__restrict__ quat,
    const float *__restrict__ inv_mass, const int *__restrict__ shape,
    const HullInfo *__restrict__ hulls, const float4 *__restrict__ verts,
    const float4 *__restrict__ planes, uint32_t *__restrict__ has_contact,
    uint32_t *__restrict__ cand, int4 *__restrict__ face_best, int *overflow) {
    // (hull_limits_ok: at most 2 * NP_BLOCK faces per hull)
    __shared__ float4 sW[4 * NP_BLOCK];  // face axes in world space
    __shared__ float sLB[4 * NP_BLOCK];  // their overlap bounds
    __shared__ int sQ[4 * NP_BLOCK];     // axes left to project
    const int lane = threadIdx.x;
    int nwork = (int)min(*n_work_ptr, (uint32_t)max_pairs);
    // The pair and body loads of the next pair go out at the top of each
    // iteration (and the pair of the one after), so their latency overlaps
    // the current pair's work.
    struct Head {
        int a, b, sA, sB;
        float imA, imB;
        float4 pA, pB, oA, oB;
    };
    auto load_head = [&](int2 pr, Head &h) {
        h.a = pr.x;
        h.b = pr.y;
        h.imA = inv_mass[pr.x];
        h.imB = inv_mass[pr.y];
        h.sA = shape[pr.x];
        h.sB = shape[pr.y];
        h.pA = pos[pr.x];
        h.pB = pos[pr.y];
        h.oA = quat[pr.x];
        h.oB = quat[pr.y];
    };
    const int G = gridDim.x;
    Head cur, nxt;
    int p_next = (int)blockIdx.x + G < nwork ? (int)work[blockIdx.x + G] : 0;
    int2 pr_next = (int)blockIdx.x + G < nwork ? pairs[p_next] : make_int2(0, 0);
    int p_after = 0;
    int2 pr_after = make_int2(0, 0);
    int p = (int)blockIdx.x < nwork ? (int)work[blockIdx.x] : 0;
    if ((int)blockIdx.x < nwork) load_head(pairs[p], cur);
    for (int i = blockIdx.x; i < nwork;
         i += G, cur = nxt, p = p_next, p_next = p_after, pr_next = pr_after) {
        if (i + G < nwork) load_head(pr_next, nxt);
        if (i + 2 * G < nwork) {
            p_after = (int)work[i + 2 * G];
            pr_after = pairs[p_after];
        }
        const int a = cur.a, b = cur.b;
        const float imA = cur.imA, imB = cur.imB;
        if (imA == 0.f && imB == 0.f) {
            if (lane == 0) { has_contact[p] = 0; cand[p] = 0; }
            continue;
        }
        const HullInfo hA = hulls[cur.sA];
        const HullInfo hB = hulls[cur.sB];
        if (hA.num_vertices == 0 || hB.num_vertices == 0) {  // a sphere or a plane: k_narrow_special
            if (lane == 0) { has_contact[p] = 0; cand[p] = 0; }
            continue;
        }
        if (!hull_limits_ok(hA, hB)) {
            if (lane == 0) {
                has_contact[p] = 0;
                cand[p] = 0;
                atomicOr(overflow, (int)MI_RB_OVERFLOW_POLYGON);
            }
            continue;
        }
        float4 pA = cur.pA, pB = cur.pB, oA = cur.oA, oB = cur.oB;
        pA.w = 0.f;
        pB.w = 0.f;
        // counts and offsets are the same on every lane: make them scalar
        const int nvA = __builtin_amdgcn_readfirstlane(hA.num_vertices);
        const int nvB = __builtin_amdgcn_readfirstlane(hB.num_vertices);
        const cfloat4 *sVA = (const cfloat4 *)(verts + __builtin_amdgcn_readfirstlane(hA.vertex_offset));
        const cfloat4 *sVB = (const cfloat4 *)(verts + __builtin_amdgcn_readfirstlane(hB.vertex_offset));
        float4 dc = sub4(xform(hA.center, pA, oA), xform(hB.center, pB, oB));

        // ---- SAT, face axes ----
        const int nf = hA.num_faces + hB.num_faces;
        const int nfA = hA.num_faces;
        const float4 *gVA = verts + hA.vertex_offset, *gVB = verts + hB.vertex_offset;
        int sep_found = 0;
        float best = FLT_MAX;
        int best_ax = 0x7fffffff;
        // Project every axis of a batch of up to 64; a batch of fewer axes
        // gives each axis 64 / next_pow2(count) lanes that split the vertices.
        // Lane results: separation flag, lexicographic (overlap, axis) minimum.
        auto project_batch = [&](int cnt, auto axis_at) {
            const int lg = lanes_log2(cnt);
            const int slot = lane >> lg, part = lane & ((1 << lg) - 1);
            const bool valid = slot < cnt;
            const int ax = axis_at(valid ? slot : 0);  // spare lanes repeat an axis
            const bool onA = ax < nfA;
            const float4 w = sW[ax];
            float mn0, mx0, mn1, mx1;
            if (lg == 0) {
                if (onA) {
                    project(sVA, nvA, pA, oA, w, mn0, mx0);
                    project(sVB, nvB, pB, oB, w, mn1, mx1);
                } else {
                    project(sVB, nvB, pB, oB, w, mn0, mx0);
                    project(sVA, nvA, pA, oA, w, mn1, mx1);
                }
            } else {
                project_part(onA ? gVA : gVB, onA ? nvA : nvB, part, lg, onA ? pA : pB,
                             onA ? oA : oB, w, mn0, mx0);
                project_part(onA ? gVB : gVA, onA ? nvB : nvA, part, lg, onA ? pB : pA,
                             onA ? oB : oA, w, mn1, mx1);
            }
            if (valid && part == 0) {
                if (mx0 < mn1 || mx1 < mn0) {
                    sep_found = 1;
                } else {
                    const float d0 = mx0 - mn1, d1 = mx1 - mn0;
                    const float d = d0 < d1 ? d0 : d1;
                    if (d < best || (d == best && ax < best_ax)) { best = d; best_ax = ax; }
                }
            }
        };
        __syncthreads();  // the previous pair's reads of sW, sLB and sQ are done
        if (nf <= FACE_PLAIN_MAX) {
            // few axes: all of them in one or two batches
            for (int ax = lane; ax < nf; ax += NP_BLOCK) {
                const bool onA = ax < nfA;
                const float4 nrm = onA ? planes[hA.face_offset + ax] : planes[hB.face_offset + ax - nfA];
                float4 w = qt_rotate(onA ? oA : oB, nrm);
                if (dot3(dc, w) < 0) w = scl4(w, -1.f);
                sW[ax] = w;
            }
            __syncthreads();
            for (int base = 0; base < nf; base += NP_BLOCK)
                project_batch(nf - base < NP_BLOCK ? nf - base : NP_BLOCK,
                              [&](int k) { return base + k; });
        } else {
            // many axes: 1. every face axis in world space (oriented along
            // dc) and its overlap bound; 2. the axis with the smallest bound
            // projected on all lanes: a separating axis ends the pair, else
            // its overlap is a known overlap; 3. the axes whose bound does
            // not exceed it, compacted and projected in batches.  Axes left
            // out cannot separate or be the minimum, so the result is that
            // of projecting every axis.
            const float margin = sat_margin(pA, pB);
            const float rho = hA.center.w + hB.center.w;
            const float4 dv = support_diff(verts, hA, hB, pA, oA, pB, oB, dc, lane);
            float lbmin = FLT_MAX;
            int lbax = 0x7fffffff;
            for (int ax = lane; ax < nf; ax += NP_BLOCK) {
                const bool onA = ax < nfA;
                const float4 nrm = onA ? planes[hA.face_offset + ax] : planes[hB.face_offset + ax - nfA];
                float4 w = qt_rotate(onA ? oA : oB, nrm);
                if (dot3(dc, w) < 0) w = scl4(w, -1.f);
                sW[ax] = w;
                const float lb = sat_bound(rho, dc, true, dv, w) - margin;
                sLB[ax] = lb;
                if (lb < lbmin) { lbmin = lb; lbax = ax; }
            }
            wave_argmin(lbmin, lbax);
            __syncthreads();
            project_batch(1, [&](int) { return lbax; });
            best = __shfl(best, 0, 64);
            best_ax = __shfl(best_ax, 0, 64);
            sep_found = __shfl(sep_found, 0, 64);
            if (!sep_found) {
                const float ref = best;
                int qn = 0;
                for (int base = 0; base < nf; base += NP_BLOCK) {
                    const int k = base + lane;
                    const bool keep = k < nf && k != lbax && !(sLB[k] > ref);
                    const uint64_t m = __ballot(keep);
                    if (keep) sQ[qn + __popcll(m & lanemask_lt())] = k;
                    qn += __popcll(m);
                }
                __syncthreads();
                for (int b0 = 0; b0 < qn; b0 += NP_BLOCK)
                    project_batch(qn - b0 < NP_BLOCK ? qn - b0 : NP_BLOCK,
                                  [&](int k) { return sQ[b0 + k]; });
            }
        }
        if (__ballot(sep_found)) {
            if (lane == 0) { has_contact[p] = 0; cand[p] = 0; }
            continue;
        }
        wave_argmin(best, best_ax);
        if (lane == 0) {
            cand[p] = 1;
            face_best[p] = make_int4(__float_as_int(best), best_ax, hA.num_edges * hB.num_edges,
                                     small_hull(hA) && small_hull(hB) ? 1 : 0);
        }
    }
}

// Pass 1 for pairs of two small hulls (listed by pass 0 after the others):
// SG lanes per pair, four pairs per wave.  The face axes (at most 32) are
// those of the wave kernel's plain path, built and projected with the same
// arithmetic (two axes per lane, every vertex of both hulls from the
// group's LDS), so the separation and the lexicographic (overlap, axis)
// minimum are the same.
__global__ __launch_bounds__(NP_BLOCK) void k_narrow_face_small(
    int max_pairs, const uint32_t *__restrict__ n_large_ptr,
    const uint32_t *__restrict__ n_small_ptr, const uint32_t *__restrict__ work,
    const int2 *__restrict__ pairs, const float4 *__restrict__ pos,
    const float4 *__restrict__ quat, const float *__restrict__ inv_mass,
    const int *__restrict__ shape, const HullInfo *__restrict__ hulls,
    const float4 *__restrict__ verts, const float4 *__restrict__ planes,
    uint32_t *__restrict__ has_contact, uint32_t *__restrict__ cand,
    int4 *__restrict__ face_best) {
    __shared__ float4 sVA[SG_GROUPS][SG_VERTS], sVB[SG_GROUPS][SG_VERTS];
    const int lane = threadIdx.x;
    const int grp = lane / SG, gl = lane & (SG - 1), gbase = grp * SG;
    const uint32_t n_large = *n_large_ptr;
    const uint32_t n_small = *n_small_ptr;
    int first = (int)min(n_large, (uint32_t)max_pairs);
    int end = (int)min(n_large + n_small, (uint32_t)max_pairs);
    const int W = (int)gridDim.x * SG_GROUPS;
    for (int i = first + (int)blockIdx.x * SG_GROUPS + grp; i < end; i += W) {
        const int p = (int)work[i];
        const int2 pr = pairs[p];
        const HullInfo hA = hulls[shape[pr.x]];
        const HullInfo hB = hulls[shape[pr.y]];
        float4 pA = pos[pr.x], pB = pos[pr.y];
        const float4 oA = quat[pr.x], oB = quat[pr.y];
        pA.w = 0.f;
        pB.w = 0.f;
        const int nvA = hA.num_vertices, nvB = hB.num_vertices;
        group_sync();  // the previous pair's reads of the group's vertices are done
        if (gl < nvA) sVA[grp][gl] = verts[hA.vertex_offset + gl];
        if (gl < nvB) sVB[grp][gl] = verts[hB.vertex_offset + gl];
        group_sync();
        const float4 dc = sub4(xform(hA.center, pA, oA), xform(hB.center, pB, oB));
        const int nfA = hA.num_faces, nf = nfA + hB.num_faces;
        bool sep_found = false;
        float best = FLT_MAX;
        int best_ax = 0x7fffffff;
        for (int ax = gl; ax < nf; ax += SG) {
            const bool onA = ax < nfA;
            const float4 nrm = onA ? planes[hA.face_offset + ax] : planes[hB.face_offset + ax - nfA];
            float4 w = qt_rotate(onA ? oA : oB, nrm);
            if (dot3(dc, w) < 0) w = scl4(w, -1.f);
            float mn0, mx0, mn1, mx1;
            if (onA) {
                project_part(sVA[grp], nvA, 0, 0, pA, oA, w, mn0, mx0);
                project_part(sVB[grp], nvB, 0, 0, pB, oB, w, mn1, mx1);
            } else {
                project_part(sVB[grp], nvB, 0, 0, pB, oB, w, mn0, mx0);
                project_part(sVA[grp], nvA, 0, 0, pA, oA, w, mn1, mx1);
            }
            if (mx0 < mn1 || mx1 < mn0) {
                sep_found = true;
            } else {
                const float d0 = mx0 - mn1, d1 = mx1 - mn0;
                const float d = d0 < d1 ? d0 : d1;
                if (d < best || (d == best && ax < best_ax)) { best = d; best_ax = ax; }
            }
        }
        if (group_ballot(sep_found, gbase)) {
            if (gl == 0) { has_contact[p] = 0; cand[p] = 0; }
            continue;
        }
        group_argmin(best, best_ax);
        if (gl == 0) {
            cand[p] = 1;
            face_best[p] = make_int4(__float_as_int(best), best_ax, hA.num_edges * hB.num_edges, 1);
        }
    }
}

// Candidates in descending order of edge-edge axis count (stable), pairs of
// small hulls (the SG-lane clip pass) after the others.
__global__ __launch_bounds__(256) void k_narrow_keys(int max_pairs, const int *__restrict__ num_pairs_ptr,
                                                     const uint32_t *__restrict__ cand,
                                                     const uint32_t *__restrict__ cand_off,
                                                     const int4 *__restrict__ face_best,
                                                     uint2 *__restrict__ kv) {
    int npairs = *num_pairs_ptr;
    if (npairs > max_pairs) npairs = max_pairs;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs || !cand[p]) return;
    // large pairs first, each class by descending edge-axis count (at most
    // 128 x 128 < 0x8000)
    const int4 fb = face_best[p];
    const uint32_t ne = (uint32_t)(fb.z < 0x7fff ? fb.z : 0x7fff);
    kv[cand_off[p]] = make_uint2((fb.w ? 0x8000u : 0u) | (0x7fffu - ne), (uint32_t)p);
}

// Pass 2, the candidates: edge x edge axes (never reject), then clipping of
// B's incident face against A's reference face and the reduction to <= 4
// points.  Workgroups take candidates in boustrophedon order over the
// cost-sorted list (round r: workgroup b takes r*G + b, or r*G + G-1-b on odd
// rounds), which spreads the expensive pairs over all workgroups.
__global__ __launch_bounds__(NP_BLOCK) void k_narrow_clip(
    const int *__restrict__ n_cand_ptr, const uint2 *__restrict__ order,
    const int4 *__restrict__ face_best, const int2 *__restrict__ pairs,
    const float4 *__restrict__ pos, const float4 *__restrict__ quat,
    const float *__restrict__ inv_mass, const int *__restrict__ shape,
    const HullInfo *__restrict__ hulls, const float4 *__restrict__ verts,
    const float4 *__restrict__ planes, const int2 *__restrict__ face_index,
    const int *__restrict__ indices, const float4 *__restrict__ edges, float friction,
    float max_dist, Contact4 *__restrict__ slots, uint32_t *__restrict__ has_contact,
    const int *__restrict__ owner, int *overflow, int edge_reject, int reduce_mode) {
    __shared__ float4 sEA[RB_MAX_HULL_EDGES], sEB[RB_MAX_HULL_EDGES];
    __shared__ float4 sPoly[2][RB_MAX_POLY];
    __shared__ float4 sPts[RB_MAX_POLY];
    __shared__ int sQ[2 * NP_BLOCK];  // edge axes that survived the bound test

    const int lane = threadIdx.x;
    const int n_cand = (int)*n_cand_ptr;
    const int G = gridDim.x, bid = (int)blockIdx.x;
    // candidate of round r (boustrophedon over the cost-sorted order); the
    // next round's pair and the index of the one after are loaded while
    // this round's candidate runs
    auto cand_at = [&](int r) { return r * G + ((r & 1) ? (G - 1 - bid) : bid); };
    int i = cand_at(0);
    int p = i < n_cand ? (int)order[i].y : 0;
    int2 pr = i < n_cand ? pairs[p] : make_int2(0, 0);
    int i_nx = cand_at(1);
    int p_nx = i_nx < n_cand ? (int)order[i_nx].y : 0;
    int2 pr_nx = make_int2(0, 0);
    int i_nx2 = 0, p_nx2 = 0;
    for (int r = 0; i < n_cand;
         r++, i = i_nx, p = p_nx, pr = pr_nx, i_nx = i_nx2, p_nx = p_nx2) {
        pr_nx = i_nx < n_cand ? pairs[p_nx] : make_int2(0, 0);
        i_nx2 = cand_at(r + 2);
        p_nx2 = i_nx2 < n_cand ? (int)order[i_nx2].y : 0;
        int a = pr.x, b = pr.y;
        float imA = inv_mass[a], imB = inv_mass[b];
        const HullInfo hA = hulls[shape[a]];
        const HullInfo hB = hulls[shape[b]];
        float4 pA = pos[a], pB = pos[b], oA = quat[a], oB = quat[b];
        pA.w = 0.f;
        pB.w = 0.f;
        __syncthreads();
        // counts and offsets are the same on every lane: make them scalar
        const int nvA = __builtin_amdgcn_readfirstlane(hA.num_vertices);
        const int nvB = __builtin_amdgcn_readfirstlane(hB.num_vertices);
        const cfloat4 *sVA = (const cfloat4 *)(verts + __builtin_amdgcn_readfirstlane(hA.vertex_offset));
        const cfloat4 *sVB = (const cfloat4 *)(verts + __builtin_amdgcn_readfirstlane(hB.vertex_offset));
        for (int i = lane; i < hA.num_edges; i += NP_BLOCK)
            sEA[i] = qt_rotate(oA, edges[hA.edge_offset + i]);
        for (int i = lane; i < hB.num_edges; i += NP_BLOCK)
            sEB[i] = qt_rotate(oB, edges[hB.edge_offset + i]);
        __syncthreads();

        float4 dc = sub4(xform(hA.center, pA, oA), xform(hB.center, pB, oB));

        // lane-local minimum over the edge x edge axes, then the wave's,
        // then against the face minimum of the first pass: face axes come
        // first in the axis order, so a tie keeps the face axis
        float best = FLT_MAX;
        int best_ax = 0x7fffffff;
        const int nf = hA.num_faces + hB.num_faces;
        const int ne = hA.num_edges * hB.num_edges;
        // An edge axis whose overlap bound (sat_bound) exceeds the face
        // minimum by the rounding margin cannot be the minimum and is not
        // projected; the minimum found is the one of the full search.
        const int4 fb0 = face_best[p];
        const float face_d = __int_as_float(fb0.x);
        const float margin = sat_margin(pA, pB);
        const float rho = hA.center.w + hB.center.w;
        // Pairs with many edge axes take a second bound from probe vertices:
        // for any vertex sB of B and sA of A, max_B(c) - min_A(c) >=
        // dot(sB - sA, c), so with the vertices furthest along NPROBE
        // directions around dc (B) and against them (A), max over the B
        // probes minus min over the A probes bounds that overlap term for
        // every axis near dc, where the competitive edge axes are.  Probe
        // positions are taken relative to B's origin (no large coordinates).
        const bool use_probe = ne > NP_BLOCK;
        float4 *sProbe = sPts;  // B's probes, then A's (sPts is free until the clipping)
        if (use_probe) {
            float4 n = dot3(dc, dc) > 0.f ? normalize3(dc) : mk4(0.f, 1.f, 0.f, 0.f);
            float4 e1 = normalize3(cross3(n, fabsf(n.x) < 0.9f ? mk4(1.f, 0.f, 0.f, 0.f)
                                                                 : mk4(0.f, 1.f, 0.f, 0.f)));
            float4 e2 = cross3(n, e1);
            const float4 dAB = sub4(pA, pB);
            const float4 wB4 = lane < hB.num_vertices
                                   ? qt_rotate(oB, verts[hB.vertex_offset + lane])
                                   : mk4(0.f, 0.f, 0.f, 0.f);
            const float4 wA4 = lane < hA.num_vertices
                                   ? add4(qt_rotate(oA, verts[hA.vertex_offset + lane]), dAB)
                                   : mk4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < NPROBE; k++) {
                const float t = k == 0 ? 0.f : (k <= 4 ? 0.35f : 0.8f * 0.70710678f);
                const float a = k <= 4 ? ((k & 1) ? 1.f : 0.f) * ((k & 2) ? -1.f : 1.f) : ((k & 1) ? 1.f : -1.f);
                const float b = k <= 4 ? ((k & 1) ? 0.f : 1.f) * ((k & 2) ? -1.f : 1.f) : ((k & 2) ? 1.f : -1.f);
                const float4 u = add4(n, add4(scl4(e1, t * a), scl4(e2, t * b)));
                // Any vertex gives a valid bound, so the furthest one is
                // found on one ordered integer key per lane (the projection's
                // ordered bits with the lane in the low 6 bits); a NaN
                // projection counts as none and lanes without a vertex lose.
                const int ib =
                    wave_max_i32(probe_key(dot3(wB4, u), lane, lane < hB.num_vertices)) & 63;
                const int ia =
                    wave_max_i32(probe_key(-dot3(wA4, u), lane, lane < hA.num_vertices)) & 63;
                if (lane == ib) sProbe[k] = wB4;
                if (lane == ia) sProbe[NPROBE + k] = wA4;
            }
            __syncthreads();
        }
        // (the vertex bound only pays for pairs with many edge axes)
        const bool use_dv = ne > 512;
        const float4 dv = !use_dv ? mk4(0.f, 0.f, 0.f, 0.f)
                          : use_probe ? sub4(sProbe[0], sProbe[NPROBE])
                                      : support_diff(verts, hA, hB, pA, oA, pB, oB, dc, lane);
        // The bound test runs on all lanes, one axis each; the surviving
        // axes are compacted into an LDS queue and projected 64 at a time, so
        // the projection loops run with full waves instead of with the few
        // lanes whose axis survived.  Lanes see axes out of index order, so
        // the lane-local minimum breaks ties by the axis index.
        const int nEB = hB.num_edges;
        auto axis_of = [&](int k) {
            int e0 = k / nEB, e1 = k - e0 * nEB;
            float4 c = cross3(sEA[e0], sEB[e1]);
            bool ok = fabsf(c.x) > 1e-6f || fabsf(c.y) > 1e-6f || fabsf(c.z) > 1e-6f;
            if (ok) {
                c = normalize3(c);
                if (dot3(dc, c) < 0) c = scl4(c, -1.f);
            }
            return ok ? c : mk4(0.f, 0.f, 0.f, -1.f);  // w = -1: degenerate axis
        };
        auto project_axis = [&](int k) {
            float4 c = axis_of(k);
            float mn0, mx0, mn1, mx1;
            project(sVA, nvA, pA, oA, c, mn0, mx0);
            project(sVB, nvB, pB, oB, c, mn1, mx1);
            float d0 = mx0 - mn1, d1 = mx1 - mn0;
            float dist = d0 < d1 ? d0 : d1;
            int ax = nf + k;
            if (dist < best || (dist == best && ax < best_ax)) { best = dist; best_ax = ax; }
        };
        // Stage 2 (pairs with probes): the probe bound over up to 64 queued
        // axes of sQ, one per lane; survivors go to sP, projected 64 at a
        // time.  Without probes every queued axis survives.
        int *sP = (int *)&sPoly[0][0];  // (sPoly is free until the clipping)
        int pn = 0;
        auto stage2 = [&](int cnt) {
            const int k = lane < cnt ? sQ[lane] : 0;
            bool keep = lane < cnt;
            if (keep && use_probe) {
                const int e0 = k / nEB, e1 = k - e0 * nEB;
                float4 c = cross3(sEA[e0], sEB[e1]);
                const float inv = __builtin_amdgcn_rsqf(dot3(c, c));
                c = scl4(c, inv);
                float dd = dot3(dc, c);
                if (dd < 0.f) {
                    c = neg4(c);
                    dd = -dd;
                }
                float hb = -FLT_MAX, la = FLT_MAX;
#pragma unroll
                for (int j = 0; j < NPROBE; j++) {
                    hb = fmaxf(hb, dot3(sProbe[j], c));
                    la = fminf(la, dot3(sProbe[NPROBE + j], c));
                }
                const float t2 = fmaxf(rho - dd, hb - la);
                keep = !(fminf(rho + dd, t2) - margin > face_d);
            }
            const uint64_t m = __ballot(keep);
            if (keep) sP[pn + __popcll(m & lanemask_lt())] = k;
            pn += __popcll(m);
            __syncthreads();
            if (pn >= NP_BLOCK) {
                project_axis(sP[lane]);
                const int rest = pn - NP_BLOCK;
                const int moved = lane < rest ? sP[NP_BLOCK + lane] : 0;
                __syncthreads();
                if (lane < rest) sP[lane] = moved;
                pn = rest;
                __syncthreads();
            }
        };
        int qn = 0;
        // The bound test takes the unit axis from the hardware's approximate
        // reciprocal square root (error ~1 ulp, far below the margin) and
        // orients it by the sign of dot(dc, c): the bound is that of
        // sat_bound up to rounding.  The degenerate-axis test is axis_of's.
        // Lane's axis k = base + lane as (e0, e1), stepped by 64 without a
        // division per axis.
        // Rows and columns of the edge-axis grid that the bound drops whole:
        // every axis c = ea x eb is perpendicular to ea, so |dot(dc, c)| is
        // at most dc's distance from ea's direction, |dc x ea| / |ea|.  When
        // that is below rho - margin - face_d (less a rounding allowance),
        // rho - |dot(dc, c)| - margin exceeds face_d on every axis of the row
        // and the bound drops them all.  The other rows and columns are
        // listed in increasing order, so the axes reaching the bound keep
        // their order and their index k.
        int *sRA = (int *)&sPoly[0][0] + 2 * NP_BLOCK, *sRB = sRA + RB_MAX_HULL_EDGES;
        const float T = rho - margin - face_d - 1e-5f * (1.f + sqrtf(dot3(dc, dc)));
        int naA = 0, naB = 0;
        for (int base = 0; base < hA.num_edges; base += NP_BLOCK) {
            const int i = base + lane;
            bool act = false;
            if (i < hA.num_edges) {
                const float4 e = sEA[i], x = cross3(dc, e);
                act = !(T > 0.f && dot3(x, x) < T * T * dot3(e, e));
            }
            const uint64_t m = __ballot(act);
            if (act) sRA[naA + __popcll(m & lanemask_lt())] = i;
            naA += __popcll(m);
        }
        for (int base = 0; base < nEB; base += NP_BLOCK) {
            const int i = base + lane;
            bool act = false;
            if (i < nEB) {
                const float4 e = sEB[i], x = cross3(dc, e);
                act = !(T > 0.f && dot3(x, x) < T * T * dot3(e, e));
            }
            const uint64_t m = __ballot(act);
            if (act) sRB[naB + __popcll(m & lanemask_lt())] = i;
            naB += __popcll(m);
        }
        __syncthreads();
        const int nact = naA * naB;
        const int q64 = naB > 0 ? NP_BLOCK / naB : 0, r64 = NP_BLOCK - q64 * naB;
        int i0 = naB > 0 ? lane / naB : 0, i1 = lane - i0 * naB;
        for (int base = 0; base < nact; base += NP_BLOCK) {
            bool keep = false;
            int k = 0;
            if (base + lane < nact) {
                const int e0 = sRA[i0], e1 = sRB[i1];
                k = e0 * nEB + e1;
                const float4 c = cross3(sEA[e0], sEB[e1]);
                if (fabsf(c.x) > 1e-6f || fabsf(c.y) > 1e-6f || fabsf(c.z) > 1e-6f) {
                    const float inv = __builtin_amdgcn_rsqf(dot3(c, c));
                    float dd = dot3(dc, c) * inv, dvc = dot3(dv, c) * inv;
                    if (dd < 0.f) {
                        dd = -dd;
                        dvc = -dvc;
                    }
                    const float lb1 = use_dv ? fmaxf(rho - dd, dvc) : rho - dd;
                    keep = !(fminf(rho + dd, lb1) - margin > face_d);
                }
            }
            i0 += q64;
            i1 += r64;
            if (i1 >= naB) {
                i1 -= naB;
                i0++;
            }
            const uint64_t m = __ballot(keep);
            if (keep) sQ[qn + __popcll(m & lanemask_lt())] = k;
            qn += __popcll(m);
            __syncthreads();
            if (qn >= NP_BLOCK) {
                stage2(NP_BLOCK);
                int rest = qn - NP_BLOCK;
                int moved = lane < rest ? sQ[NP_BLOCK + lane] : 0;
                __syncthreads();
                if (lane < rest) sQ[lane] = moved;
                qn = rest;
                __syncthreads();
            }
        }
        if (qn > 0) stage2(qn);
        if (pn > 0) {
            // the last partial batch: 64 / next_pow2(pn) lanes per axis
            const int lg = lanes_log2(pn);
            const int slot = lane >> lg, part = lane & ((1 << lg) - 1);
            const int k = sP[slot < pn ? slot : 0];
            if (lg == 0) {
                if (slot < pn) project_axis(k);
            } else {
                float4 c = axis_of(k);
                float mn0, mx0, mn1, mx1;
                project_part(verts + hA.vertex_offset, nvA, part, lg, pA, oA, c, mn0, mx0);
                project_part(verts + hB.vertex_offset, nvB, part, lg, pB, oB, c, mn1, mx1);
                float d0 = mx0 - mn1, d1 = mx1 - mn0;
                float dist = d0 < d1 ? d0 : d1;
                int ax = nf + k;
                if (slot < pn && part == 0 &&
                    (dist < best || (dist == best && ax < best_ax))) {
                    best = dist;
                    best_ax = ax;
                }
            }
        }
        __syncthreads();
        wave_argmin(best, best_ax);
        if (face_d < best || (face_d == best && fb0.y < best_ax)) { best = face_d; best_ax = fb0.y; }
        // the CPU twin rejects on an edge axis whose projections are disjoint:
        // every face axis overlaps here (pass 1 kept the pair), so a negative
        // minimum is such an edge axis
        if (best_ax == 0x7fffffff || (edge_reject && best < 0.f)) {
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
        // the side planes of A's reference face (n, w), one lane per edge (a
        // face has at most RB_MAX_HULL_VERTS <= RB_MAX_POLY vertices), kept in
        // sPts until the clipped points replace them
        float4 *sSide = sPts;
        for (int e = lane; e < fiA.y; e += NP_BLOCK) {
            float4 va = verts[hA.vertex_offset + indices[fiA.x + e]];
            float4 vb = verts[hA.vertex_offset + indices[fiA.x + (e + 1) % fiA.y]];
            float4 we = qt_rotate(oA, sub4(va, vb));
            float4 pn = neg4(cross3(we, wpn));
            float4 wa = xform(va, pA, oA);
            sSide[e] = mk4(pn.x, pn.y, pn.z, -dot3(wa, pn));
        }
        int cur = 0;
        bool poly_over = false;
        __syncthreads();
        for (int e0 = 0; e0 < fiA.y; e0++) {
            const float4 side = sSide[e0];
            const float4 pn = mk4(side.x, side.y, side.z, 0.f);
            const float eq = side.w;
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
        if (reduce_mode == MI_RB_REDUCE_SPLIT_PATH) {
            if (npts > 4) {
                // extractManifoldSequentialGlobal (satClipHullContacts.cl:678-767):
                // per direction the first point with the smallest projection
                // below FLT_MIN, the first deepest point
                float4 c = sPts[0];
                for (int j = 1; j < npts; j++) c = add4(c, sPts[j]);
                float fn = (float)npts;
                c = mk4(c.x / fn, c.y / fn, c.z / fn, c.w / fn);
                float4 av = sub4(sPts[0], c);
                float4 u = cross3(nrm, av);
                float4 v = cross3(nrm, u);
                u = normalize3(u);
                v = normalize3(v);
                float f[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
                float dw = FLT_MAX;
                int iw = 0x7fffffff;
                if (lane < npts) {
                    float4 q = sPts[lane];
                    float4 r = sub4(q, c);
                    f[0] = dot3(u, r); f[1] = dot3(neg4(u), r);
                    f[2] = dot3(v, r); f[3] = dot3(neg4(v), r);
                    for (int k = 0; k < 4; k++)
                        if (!(f[k] < FLT_MIN)) f[k] = FLT_MAX;
                    dw = q.w;
                    iw = lane;
                }
                for (int k = 0; k < 4; k++) {
                    int ik = f[k] < FLT_MAX ? lane : 0x7fffffff;
                    float fk = f[k];
                    wave_argmin(fk, ik);
                    if (ik != 0x7fffffff) idx[k] = ik;
                }
                wave_argmin(dw, iw);
                if (idx[0] != iw && idx[1] != iw && idx[2] != iw && idx[3] != iw) idx[0] = iw;
                nr = 4;
            }
        } else if (npts >= 4) {
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
            // a compound child (proxy body) names its compound
            const int oa = owner[a], ob = owner[b];
            ct.body_a = imA == 0.f ? -oa : oa;
            ct.body_b = imB == 0.f ? -ob : ob;
            slots[p] = ct;
            has_contact[p] = 1;
        }
    }
}

// ---- pass 2 for small hulls: SG lanes per candidate ----
// Candidates whose hulls both have at most 8 vertices (cubes, the 8-vertex
// hulls: at most 18 unique edges, 16 faces, 8 vertices in a face) run SG = 16
// lanes each, four per wave, instead of a wave each: their edge axes (at most
// 324), faces and clipped polygons (at most 8 + 8 points) fit 16 lanes, and a
// wave per pair left most lanes idle.  Every step is pass 2's with the same
// arithmetic and the same tie rules, so the manifold is identical; the lanes
// of a group take the same branches, reductions stay inside the group, and
// LDS is the group's own.

struct SmallClipLDS {
    float4 ea[SG_EDGES], eb[SG_EDGES];
    float4 va[SG_VERTS], vb[SG_VERTS];
    float4 pa[SG_FACES], pb[SG_FACES];  // face planes (local), loaded with the edges
    float4 poly[2][SG_POLY];
    float4 pts[2 * NPROBE];  // probes, then the side planes, then the kept points
    int q[2 * SG], pq[2 * SG];
};

__global__ __launch_bounds__(NP_BLOCK) void k_narrow_clip_small(
    const uint32_t *__restrict__ n_cand_ptr, const int *__restrict__ split_ptr,
    const uint2 *__restrict__ order, const int4 *__restrict__ face_best,
    const int2 *__restrict__ pairs, const float4 *__restrict__ pos,
    const float4 *__restrict__ quat, const float *__restrict__ inv_mass,
    const int *__restrict__ shape, const HullInfo *__restrict__ hulls,
    const float4 *__restrict__ verts, const float4 *__restrict__ planes,
    const int2 *__restrict__ face_index, const int *__restrict__ indices,
    const float4 *__restrict__ edges, float friction, float max_dist, Contact4 *__restrict__ slots,
    uint32_t *__restrict__ has_contact, const int *__restrict__ owner, int *overflow,
    int edge_reject, int reduce_mode) {
    __shared__ SmallClipLDS SL[SG_GROUPS];
    const int lane = threadIdx.x;
    const int grp = lane / SG, gl = lane & (SG - 1), gbase = grp * SG;
    const uint32_t lt = (1u << gl) - 1u;
    SmallClipLDS &S = SL[grp];
    const int first = *split_ptr;
    const int n_cand = (int)*n_cand_ptr;
    const int W = (int)gridDim.x * SG_GROUPS;
    const int wid = (int)blockIdx.x * SG_GROUPS + grp;
    for (int r = 0;; r++) {
        const int i = first + r * W + ((r & 1) ? (W - 1 - wid) : wid);
        if (i >= n_cand) break;
        const int p = (int)order[i].y;
        const int2 pr = pairs[p];
        const int a = pr.x, b = pr.y;
        const float imA = inv_mass[a], imB = inv_mass[b];
        const HullInfo hA = hulls[shape[a]];
        const HullInfo hB = hulls[shape[b]];
        float4 pA = pos[a], pB = pos[b];
        const float4 oA = quat[a], oB = quat[b];
        pA.w = 0.f;
        pB.w = 0.f;
        const int nvA = hA.num_vertices, nvB = hB.num_vertices;
        group_sync();  // the previous candidate's reads of S are done
        if (gl < hA.num_edges) S.ea[gl] = qt_rotate(oA, edges[hA.edge_offset + gl]);
        if (gl + SG < hA.num_edges) S.ea[gl + SG] = qt_rotate(oA, edges[hA.edge_offset + gl + SG]);
        if (gl < hB.num_edges) S.eb[gl] = qt_rotate(oB, edges[hB.edge_offset + gl]);
        if (gl + SG < hB.num_edges) S.eb[gl + SG] = qt_rotate(oB, edges[hB.edge_offset + gl + SG]);
        if (gl < nvA) S.va[gl] = verts[hA.vertex_offset + gl];
        if (gl < nvB) S.vb[gl] = verts[hB.vertex_offset + gl];
        // the faces' planes and (first index, count) records are read after
        // the SAT; loading them here overlaps their latency with it (lane gl
        // keeps face gl's records of both hulls)
        if (gl < hA.num_faces) S.pa[gl] = planes[hA.face_offset + gl];
        if (gl < hB.num_faces) S.pb[gl] = planes[hB.face_offset + gl];
        const int2 fiA_l = gl < hA.num_faces ? face_index[hA.face_offset + gl] : make_int2(0, 0);
        const int2 fiB_l = gl < hB.num_faces ? face_index[hB.face_offset + gl] : make_int2(0, 0);
        group_sync();

        const float4 dc = sub4(xform(hA.center, pA, oA), xform(hB.center, pB, oB));
        float best = FLT_MAX;
        int best_ax = 0x7fffffff;
        const int nf = hA.num_faces + hB.num_faces;
        const int ne = hA.num_edges * hB.num_edges;
        const int4 fb0 = face_best[p];
        const float face_d = __int_as_float(fb0.x);
        const float margin = sat_margin(pA, pB);
        const float rho = hA.center.w + hB.center.w;
        // probe vertices (as pass 2 takes them) for pairs with many edge axes
        const bool use_probe = ne > NP_BLOCK;
        if (use_probe) {
            float4 n = dot3(dc, dc) > 0.f ? normalize3(dc) : mk4(0.f, 1.f, 0.f, 0.f);
            float4 e1 = normalize3(cross3(n, fabsf(n.x) < 0.9f ? mk4(1.f, 0.f, 0.f, 0.f)
                                                                 : mk4(0.f, 1.f, 0.f, 0.f)));
            float4 e2 = cross3(n, e1);
            const float4 dAB = sub4(pA, pB);
            const float4 wB4 = gl < nvB ? qt_rotate(oB, S.vb[gl]) : mk4(0.f, 0.f, 0.f, 0.f);
            const float4 wA4 = gl < nvA ? add4(qt_rotate(oA, S.va[gl]), dAB) : mk4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < NPROBE; k++) {
                const float t = k == 0 ? 0.f : (k <= 4 ? 0.35f : 0.8f * 0.70710678f);
                const float a2 = k <= 4 ? ((k & 1) ? 1.f : 0.f) * ((k & 2) ? -1.f : 1.f) : ((k & 1) ? 1.f : -1.f);
                const float b2 = k <= 4 ? ((k & 1) ? 0.f : 1.f) * ((k & 2) ? -1.f : 1.f) : ((k & 2) ? 1.f : -1.f);
                const float4 u = add4(n, add4(scl4(e1, t * a2), scl4(e2, t * b2)));
                const int ib = group_max_i32(probe_key(dot3(wB4, u), gl, gl < nvB)) & (SG - 1);
                const int ia = group_max_i32(probe_key(-dot3(wA4, u), gl, gl < nvA)) & (SG - 1);
                if (gl == ib) S.pts[k] = wB4;
                if (gl == ia) S.pts[NPROBE + k] = wA4;
            }
            group_sync();
        }
        const int nEB = hB.num_edges;
        auto axis_of = [&](int k) {
            int e0 = k / nEB, e1 = k - e0 * nEB;
            float4 c = cross3(S.ea[e0], S.eb[e1]);
            bool ok = fabsf(c.x) > 1e-6f || fabsf(c.y) > 1e-6f || fabsf(c.z) > 1e-6f;
            if (ok) {
                c = normalize3(c);
                if (dot3(dc, c) < 0) c = scl4(c, -1.f);
            }
            return ok ? c : mk4(0.f, 0.f, 0.f, -1.f);
        };
        // axis k on 2^lg lanes (part of them), min into (best, best_ax)
        auto project_axis = [&](int k, int lg, int part, bool valid) {
            const float4 c = axis_of(k);
            float mn0, mx0, mn1, mx1;
            project_part(S.va, nvA, part, lg, pA, oA, c, mn0, mx0);
            project_part(S.vb, nvB, part, lg, pB, oB, c, mn1, mx1);
            const float d0 = mx0 - mn1, d1 = mx1 - mn0;
            const float dist = d0 < d1 ? d0 : d1;
            const int ax = nf + k;
            if (valid && part == 0 && (dist < best || (dist == best && ax < best_ax))) {
                best = dist;
                best_ax = ax;
            }
        };
        int pn = 0;
        auto stage2 = [&](int cnt) {
            const int k = gl < cnt ? S.q[gl] : 0;
            bool keep = gl < cnt;
            if (keep && use_probe) {
                const int e0 = k / nEB, e1 = k - e0 * nEB;
                float4 c = cross3(S.ea[e0], S.eb[e1]);
                const float inv = __builtin_amdgcn_rsqf(dot3(c, c));
                c = scl4(c, inv);
                float dd = dot3(dc, c);
                if (dd < 0.f) {
                    c = neg4(c);
                    dd = -dd;
                }
                float hb = -FLT_MAX, la = FLT_MAX;
#pragma unroll
                for (int j = 0; j < NPROBE; j++) {
                    hb = fmaxf(hb, dot3(S.pts[j], c));
                    la = fminf(la, dot3(S.pts[NPROBE + j], c));
                }
                const float t2 = fmaxf(rho - dd, hb - la);
                keep = !(fminf(rho + dd, t2) - margin > face_d);
            }
            const uint32_t m = group_ballot(keep, gbase);
            if (keep) S.pq[pn + __popc(m & lt)] = k;
            pn += __popc(m);
            group_sync();
            if (pn >= SG) {
                project_axis(S.pq[gl], 0, 0, true);
                const int rest = pn - SG;
                const int moved = gl < rest ? S.pq[SG + gl] : 0;
                group_sync();
                if (gl < rest) S.pq[gl] = moved;
                pn = rest;
                group_sync();
            }
        };
        int qn = 0;
        // the first bound, as pass 2 takes it (no vertex bound: ne <= 324)
        const int qs = nEB > 0 ? SG / nEB : 0, rs = SG - qs * nEB;
        int e0 = nEB > 0 ? gl / nEB : 0, e1 = gl - e0 * nEB;
        for (int base = 0; base < ne; base += SG) {
            const int k = base + gl;
            bool keep = false;
            if (k < ne) {
                const float4 c = cross3(S.ea[e0], S.eb[e1]);
                if (fabsf(c.x) > 1e-6f || fabsf(c.y) > 1e-6f || fabsf(c.z) > 1e-6f) {
                    const float inv = __builtin_amdgcn_rsqf(dot3(c, c));
                    float dd = dot3(dc, c) * inv;
                    if (dd < 0.f) dd = -dd;
                    keep = !(fminf(rho + dd, rho - dd) - margin > face_d);
                }
            }
            e0 += qs;
            e1 += rs;
            if (e1 >= nEB) {
                e1 -= nEB;
                e0++;
            }
            const uint32_t m = group_ballot(keep, gbase);
            if (keep) S.q[qn + __popc(m & lt)] = k;
            qn += __popc(m);
            group_sync();
            if (qn >= SG) {
                stage2(SG);
                const int rest = qn - SG;
                const int moved = gl < rest ? S.q[SG + gl] : 0;
                group_sync();
                if (gl < rest) S.q[gl] = moved;
                qn = rest;
                group_sync();
            }
        }
        if (qn > 0) stage2(qn);
        if (pn > 0) {
            const int lg = group_lanes_log2(pn);
            const int slot = gl >> lg, part = gl & ((1 << lg) - 1);
            project_axis(S.pq[slot < pn ? slot : 0], lg, part, slot < pn);
        }
        group_argmin(best, best_ax);
        if (face_d < best || (face_d == best && fb0.y < best_ax)) { best = face_d; best_ax = fb0.y; }
        if (best_ax == 0x7fffffff || (edge_reject && best < 0.f)) {
            if (gl == 0) has_contact[p] = 0;
            continue;
        }
        float4 sep;
        if (best_ax < hA.num_faces) {
            sep = qt_rotate(oA, S.pa[best_ax]);
            if (dot3(dc, sep) < 0) sep = scl4(sep, -1.f);
        } else if (best_ax < nf) {
            sep = qt_rotate(oB, S.pb[best_ax - hA.num_faces]);
            if (dot3(dc, sep) < 0) sep = scl4(sep, -1.f);
        } else {
            const int k = best_ax - nf;
            const int f0 = k / nEB, f1 = k - f0 * nEB;
            sep = normalize3(cross3(S.ea[f0], S.eb[f1]));
            if (dot3(dc, sep) < 0) sep = scl4(sep, -1.f);
        }
        if (dot3(neg4(dc), sep) > 0.0f) sep = neg4(sep);

        // ---- clipping: incident face of B, reference face of A ----
        float dB = -FLT_MAX;
        int fB = 0x7fffffff;
        if (gl < hB.num_faces) {
            const float4 pl = S.pb[gl];
            const float d = dot3(qt_rotate(oB, mk4(pl.x, pl.y, pl.z, 0.f)), sep);
            if (d > dB) { dB = d; fB = gl; }
        }
        group_argmax(dB, fB);
        float dA = FLT_MAX;
        int fA = 0x7fffffff;
        if (gl < hA.num_faces) {
            const float4 pl = S.pa[gl];
            const float d = dot3(qt_rotate(oA, mk4(pl.x, pl.y, pl.z, 0.f)), sep);
            if (d < dA) { dA = d; fA = gl; }
        }
        group_argmin(dA, fA);
        if (fB == 0x7fffffff || fA == 0x7fffffff) {
            if (gl == 0) has_contact[p] = 0;
            continue;
        }
        // fB, fA < SG_FACES: the group's lanes fB, fA hold those faces' records
        const int2 fiB = make_int2(__shfl(fiB_l.x, gbase + fB, 64), __shfl(fiB_l.y, gbase + fB, 64));
        int nin = fiB.y < SG_VERTS ? fiB.y : SG_VERTS;
        if (gl < nin) S.poly[0][gl] = xform(S.vb[indices[fiB.x + gl]], pB, oB);
        const int2 fiA = make_int2(__shfl(fiA_l.x, gbase + fA, 64), __shfl(fiA_l.y, gbase + fA, 64));
        const float4 plA = S.pa[fA];
        const float4 wpn = qt_rotate(oA, mk4(plA.x, plA.y, plA.z, 0.f));
        float4 *sSide = S.pts;
        const int nside = fiA.y < SG_VERTS ? fiA.y : SG_VERTS;
        if (gl < nside) {
            const float4 va = S.va[indices[fiA.x + gl]];
            const float4 vb = S.va[indices[fiA.x + (gl + 1) % fiA.y]];
            const float4 we = qt_rotate(oA, sub4(va, vb));
            const float4 pn4 = neg4(cross3(we, wpn));
            const float4 wa = xform(va, pA, oA);
            sSide[gl] = mk4(pn4.x, pn4.y, pn4.z, -dot3(wa, pn4));
        }
        int cur = 0;
        bool poly_over = fiB.y > SG_VERTS || fiA.y > SG_VERTS;
        group_sync();
        for (int s0 = 0; s0 < nside; s0++) {
            const float4 side = sSide[s0];
            const float4 pn4 = mk4(side.x, side.y, side.z, 0.f);
            const float eq = side.w;
            const float4 *in = S.poly[cur];
            float4 *outp = S.poly[cur ^ 1];
            int cnt = 0;
            float4 o0 = mk4(0, 0, 0, 0), o1 = o0;
            if (gl < nin) {
                const float4 f0 = in[gl == 0 ? nin - 1 : gl - 1];
                const float4 end = in[gl];
                const float ds = dot3(pn4, f0) + eq;
                const float de = dot3(pn4, end) + eq;
                if (ds < 0) {
                    if (de < 0) {
                        o0 = end; cnt = 1;
                    } else {
                        float t = ds * 1.f / (ds - de);
                        o0 = mk4(f0.x + (end.x - f0.x) * t, f0.y + (end.y - f0.y) * t,
                                 f0.z + (end.z - f0.z) * t, 0.f);
                        cnt = 1;
                    }
                } else if (de < 0) {
                    float t = ds * 1.f / (ds - de);
                    o0 = mk4(f0.x + (end.x - f0.x) * t, f0.y + (end.y - f0.y) * t,
                             f0.z + (end.z - f0.z) * t, 0.f);
                    o1 = end;
                    cnt = 2;
                }
            }
            const uint32_t b1 = group_ballot(cnt >= 1, gbase), b2 = group_ballot(cnt == 2, gbase);
            const int off = __popc(b1 & lt) + __popc(b2 & lt);
            const int total = __popc(b1) + __popc(b2);
            if (total > SG_POLY) poly_over = true;
            if (cnt >= 1 && off < SG_POLY) outp[off] = o0;
            if (cnt == 2 && off + 1 < SG_POLY) outp[off + 1] = o1;
            group_sync();
            nin = total < SG_POLY ? total : SG_POLY;
            cur ^= 1;
        }
        if (poly_over && gl == 0) atomicOr(overflow, (int)MI_RB_OVERFLOW_POLYGON);
        const float peq = plA.w - dot3(wpn, pA);
        int keep = 0;
        float4 kp = mk4(0, 0, 0, 0);
        if (gl < nin) {
            const float4 q = S.poly[cur][gl];
            float depth = dot3(wpn, q) + peq;
            if (depth <= -1e30f) depth = -1e30f;
            if (depth <= max_dist) { keep = 1; kp = mk4(q.x, q.y, q.z, depth); }
        }
        const uint32_t kb = group_ballot(keep, gbase);
        const int npts = __popc(kb);
        float4 *sPts = S.pts;
        group_sync();  // every lane has read the side planes
        if (keep) sPts[__popc(kb & lt)] = kp;
        group_sync();
        if (npts == 0) {
            if (gl == 0) has_contact[p] = 0;
            continue;
        }
        // ---- reduction to <= 4 points (pass 2's, on the group) ----
        const float4 nrm = neg4(sep);
        int idx[4] = {0, 1, 2, 3};
        int nr = npts;
        if (reduce_mode == MI_RB_REDUCE_SPLIT_PATH) {
            if (npts > 4) {
                float4 c = sPts[0];
                for (int j = 1; j < npts; j++) c = add4(c, sPts[j]);
                const float fn = (float)npts;
                c = mk4(c.x / fn, c.y / fn, c.z / fn, c.w / fn);
                const float4 av = sub4(sPts[0], c);
                float4 u = cross3(nrm, av);
                float4 v = cross3(nrm, u);
                u = normalize3(u);
                v = normalize3(v);
                float f[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
                float dw = FLT_MAX;
                int iw = 0x7fffffff;
                if (gl < npts) {
                    const float4 q = sPts[gl];
                    const float4 rr = sub4(q, c);
                    f[0] = dot3(u, rr); f[1] = dot3(neg4(u), rr);
                    f[2] = dot3(v, rr); f[3] = dot3(neg4(v), rr);
                    for (int k = 0; k < 4; k++)
                        if (!(f[k] < FLT_MIN)) f[k] = FLT_MAX;
                    dw = q.w;
                    iw = gl;
                }
                for (int k = 0; k < 4; k++) {
                    int ik = f[k] < FLT_MAX ? gl : 0x7fffffff;
                    float fk = f[k];
                    group_argmin(fk, ik);
                    if (ik != 0x7fffffff) idx[k] = ik;
                }
                group_argmin(dw, iw);
                if (idx[0] != iw && idx[1] != iw && idx[2] != iw && idx[3] != iw) idx[0] = iw;
                nr = 4;
            }
        } else if (npts >= 4) {
            float4 c = sPts[0];
            for (int j = 1; j < npts; j++) c = add4(c, sPts[j]);
            const float fn = (float)npts;
            c = mk4(c.x / fn, c.y / fn, c.z / fn, c.w / fn);
            const float4 av = sub4(sPts[0], c);
            float4 u = cross3(nrm, av);
            float4 v = cross3(nrm, u);
            u = normalize3(u);
            v = normalize3(v);
            int key[4] = {INT_MIN, INT_MIN, INT_MIN, INT_MIN};
            if (gl < npts) {
                const float4 rr = sub4(sPts[gl], c);
                const float f[4] = {dot3(u, rr), dot3(neg4(u), rr), dot3(v, rr), dot3(neg4(v), rr)};
                for (int k = 0; k < 4; k++)
                    key[k] = (int)((__float_as_uint(f[k]) & 0xffffff00u) | (0xffu & (uint32_t)gl));
            }
            for (int k = 0; k < 4; k++) idx[k] = group_max_i32(key[k]) & 0xff;
            nr = 4;
        }
        if (gl == 0) {
            Contact4 ct;
            ct.normal = mk4(nrm.x, nrm.y, nrm.z, (float)nr);
            for (int k = 0; k < 4; k++) ct.pos[k] = k < nr ? sPts[idx[k]] : mk4(0, 0, 0, 0);
            ct.coeffs = (uint32_t)(0.f * 0xffff) | ((uint32_t)(friction * 0xffff) << 16);
            ct.pair_index = p;
            const int oa = owner[a], ob = owner[b];
            ct.body_a = imA == 0.f ? -oa : oa;
            ct.body_b = imB == 0.f ? -ob : ob;
            slots[p] = ct;
            has_contact[p] = 1;
        }
    }
}

// First position of the small-hull candidates in the cost-sorted order (the
// sort key's bit 15; small candidates sort after the others).
__global__ __launch_bounds__(256) void k_clip_split(const uint32_t *__restrict__ n_cand_ptr,
                                                    const uint2 *__restrict__ order, int *split) {
    const int n = (int)*n_cand_ptr;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = (order[i].x >> 15) & 1u;
    const uint32_t prev = i > 0 ? (order[i - 1].x >> 15) & 1u : 0u;
    if (s && !prev) *split = i;
    if (i == n - 1 && !s) *split = n;
}

// ---- compound shapes (findCompoundPairsKernel, sat.cl:763-925) ----
// Proxy body n + k is child proxy[k].y of body proxy[k].x: its world
// transform is the body's times the child's (qtRotate(orn, childPos) + pos,
// qtMul(orn, childOrn)).
__global__ __launch_bounds__(256) void k_proxy_transforms(int P, int n, const int2 *__restrict__ proxy,
                                                          const ChildShape *__restrict__ children,
                                                          float4 *__restrict__ pos,
                                                          float4 *__restrict__ quat) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const int2 pr = proxy[k];
    const ChildShape c = children[pr.y];
    const float4 pb = pos[pr.x], qb = quat[pr.x];
    float4 np = add4(qt_rotate(qb, c.pos), pb);
    np.w = 0.f;
    pos[n + k] = np;
    quat[n + k] = qt_mul(qb, c.orn);
}

// Child pairs of pair p: (children of A or A) x (children of B or B) when A
// or B is a compound and not both are static; 0 otherwise.
__device__ __forceinline__ int compound_children(const HullInfo &h) {
    return shape_kind_of(h) == SHAPE_KIND_COMPOUND ? h.num_edges : 0;
}
// Triangle tri (a child hull of mesh body m at the identity) against body
// b's world AABB: sat.cl:985-1023 tests the triangle's mesh-frame AABB
// against b's world AABB; here its vertices are taken to world space first,
// which is the same test for a mesh body at the identity.
__device__ __forceinline__ bool mesh_tri_overlaps(const HullInfo &tri, const float4 *__restrict__ verts,
                                                  float4 pm, float4 qm, float4 bmin, float4 bmax) {
    float4 lo = mk4(1e30f, 1e30f, 1e30f, 0.f), hi = mk4(-1e30f, -1e30f, -1e30f, 0.f);
    for (int k = 0; k < 3; k++) {
        const float4 v = add4(qt_rotate(qm, verts[tri.vertex_offset + k]), pm);
        lo = mk4(fminf(lo.x, v.x), fminf(lo.y, v.y), fminf(lo.z, v.z), 0.f);
        hi = mk4(fmaxf(hi.x, v.x), fmaxf(hi.y, v.y), fmaxf(hi.z, v.z), 0.f);
    }
    return !(lo.x > bmax.x || hi.x < bmin.x || lo.z > bmax.z || hi.z < bmin.z || lo.y > bmax.y ||
             hi.y < bmin.y);
}
// What the triangle cull reads: transforms, the bodies' world AABBs from
// the broadphase, hull vertices and the child table.
struct MeshView {
    const float4 *pos, *quat, *aabb_min, *aabb_max, *verts;
    const ChildShape *children;
};
// A mesh pairs with convex hull bodies only: 0 when A is the mesh, 1 when B
// is, 2 for a mesh with anything else (no child pairs), -1 without a mesh.
__device__ __forceinline__ int mesh_side(const HullInfo &ha, const HullInfo &hb) {
    const bool ma = compound_is_mesh(ha), mb = compound_is_mesh(hb);
    if (!ma && !mb) return -1;
    if (ma && mb) return 2;  // mesh against mesh: nothing
    const HullInfo &other = ma ? hb : ha;
    return shape_kind_of(other) == SHAPE_KIND_HULL ? (ma ? 0 : 1) : 2;
}
// One wave per pair (the pair's type is uniform across the wave): a mesh
// pair tests 64 triangles at a time, counting with a ballot, and writes the
// overlapping ones in triangle order through the ballot's prefix counts.
__device__ __forceinline__ unsigned long long lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
__global__ __launch_bounds__(256) void k_compound_count(int max_pairs, const int *__restrict__ num_pairs_ptr,
                                                        const int2 *__restrict__ pairs,
                                                        const int *__restrict__ shape,
                                                        const float *__restrict__ inv_mass,
                                                        const HullInfo *__restrict__ hulls, MeshView mv,
                                                        uint32_t *__restrict__ cnt,
                                                        int *__restrict__ overflow) {
    int npairs = *num_pairs_ptr;
    if (npairs > max_pairs) npairs = max_pairs;
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < npairs; p += waves) {
        const int2 pr = pairs[p];
        const HullInfo ha = hulls[shape[pr.x]], hb = hulls[shape[pr.y]];
        const bool both_static = inv_mass[pr.x] == 0.f && inv_mass[pr.y] == 0.f;
        const int side = mesh_side(ha, hb);
        uint32_t c = 0;
        if (side >= 0) {
            // a mesh against a sphere, a plane, a compound or a mesh: no
            // contacts, reported
            if (side == 2 && !both_static && lane == 0)
                atomicOr(overflow, (int)MI_RB_OVERFLOW_UNSUPPORTED_PAIR);
            if (side < 2 && !both_static) {
                const int m = side == 0 ? pr.x : pr.y, b = side == 0 ? pr.y : pr.x;
                const HullInfo &hm = side == 0 ? ha : hb;
                const float4 pm = mv.pos[m], qm = mv.quat[m], bmin = mv.aabb_min[b],
                             bmax = mv.aabb_max[b];
                for (int t0 = 0; t0 < hm.num_edges; t0 += 64) {
                    const int t = t0 + lane;
                    const bool ov = t < hm.num_edges &&
                                    mesh_tri_overlaps(hulls[mv.children[hm.edge_offset + t].shape],
                                                      mv.verts, pm, qm, bmin, bmax);
                    c += (uint32_t)__popcll(__ballot(ov));
                }
            }
        } else {
            const int ca = compound_children(ha);
            const int cb = compound_children(hb);
            c = (ca || cb) && !both_static ? (uint32_t)((ca ? ca : 1) * (cb ? cb : 1)) : 0u;
        }
        if (lane == 0) cnt[p] = c;
    }
}
// Child pairs in pair order, child of A outer, child of B inner, after the
// npairs pairs (up to max_pairs; k_compound_total flags the rest).
__global__ __launch_bounds__(256) void k_compound_write(int max_pairs, const int *__restrict__ num_pairs_ptr,
                                                        int2 *__restrict__ pairs,
                                                        const int *__restrict__ shape,
                                                        const HullInfo *__restrict__ hulls, MeshView mv,
                                                        const int *__restrict__ proxy_first, int n,
                                                        const uint32_t *__restrict__ cnt,
                                                        const uint32_t *__restrict__ off) {
    int npairs = *num_pairs_ptr;
    if (npairs > max_pairs) npairs = max_pairs;
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < npairs; p += waves) {
        const uint32_t c = cnt[p];
        if (c == 0) continue;
        const int2 pr = pairs[p];
        const HullInfo ha = hulls[shape[pr.x]], hb = hulls[shape[pr.y]];
        const int fa = proxy_first[pr.x], fb = proxy_first[pr.y];
        const long long base = (long long)npairs + off[p];
        const int side = mesh_side(ha, hb);
        if (side >= 0) {  // overlapping triangles in triangle order (c > 0: side is 0 or 1)
            const int m = side == 0 ? pr.x : pr.y, b = side == 0 ? pr.y : pr.x;
            const HullInfo &hm = side == 0 ? ha : hb;
            const float4 pm = mv.pos[m], qm = mv.quat[m], bmin = mv.aabb_min[b], bmax = mv.aabb_max[b];
            uint32_t k0 = 0;
            for (int t0 = 0; t0 < hm.num_edges && k0 < c; t0 += 64) {
                const int t = t0 + lane;
                const bool ov = t < hm.num_edges &&
                                mesh_tri_overlaps(hulls[mv.children[hm.edge_offset + t].shape], mv.verts,
                                                  pm, qm, bmin, bmax);
                const unsigned long long mask = __ballot(ov);
                const uint32_t k = k0 + (uint32_t)__popcll(mask & lanes_below(lane));
                if (ov && k < c && base + k < max_pairs)
                    pairs[base + k] = side == 0 ? make_int2(n + fa + t, pr.y) : make_int2(pr.x, n + fb + t);
                k0 += (uint32_t)__popcll(mask);
            }
            continue;
        }
        const int cb = compound_children(hb);
        const int nb = cb ? cb : 1;
        for (uint32_t k = lane; k < c; k += 64) {
            const long long dst = base + k;
            if (dst >= max_pairs) break;
            const int i = (int)k / nb, j = (int)k % nb;
            pairs[dst] = make_int2(fa >= 0 ? n + fa + i : pr.x, fb >= 0 ? n + fb + j : pr.y);
        }
    }
}
__global__ void k_compound_total(int max_pairs, int *num_pairs, const uint32_t *total,
                                 int *overflow) {
    long long np = (long long)min(*num_pairs, max_pairs) + *total;
    if (np > max_pairs) {
        np = max_pairs;
        atomicOr(overflow, (int)MI_RB_OVERFLOW_PAIRS);
    }
    *num_pairs = (int)np;
}

}  // namespace

hipError_t launch_narrowphase(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                              const SortWork &sw, uint32_t *scan_status, size_t scan_stride) {
    const int grid = rb_num_cus() * 4 * (p.narrow_waves > 0 ? p.narrow_waves : 4);
    // pass 1's pairs: the flags of pass 0 compacted into cand_kv, which
    // k_narrow_keys fills only after pass 1
    uint32_t *work = (uint32_t *)d.cand_kv;
    uint32_t *n_work = (uint32_t *)(d.counters + RB_CNT_FACE_WORK);
    RB_LAUNCH(k_narrow_quick<QUICK_LANES>,
              dim3(((long long)p.max_pairs * QUICK_LANES + 255) / 256), dim3(256), 0, s,
              p.max_pairs, d.counters + RB_CNT_PAIRS, d.pairs, d.pos, d.quat, d.inv_mass, d.shape,
              d.hulls, d.verts, d.planes, d.has_contact, d.narrow_cand);
    // the wave kernel's pairs, then the small-hull kernel's after them
    uint32_t *n_small = (uint32_t *)(d.counters + RB_CNT_FACE_SMALL);
    hipError_t e0 = launch_compact_flags(d.has_contact, work, p.max_pairs, n_work,
                                         scan_status + RB_SCAN_WORK * scan_stride, s,
                                         d.counters + RB_CNT_PAIRS, d.counters + RB_CNT_OVERFLOW,
                                         nullptr, true);
    if (e0 != hipSuccess) return e0;
    e0 = launch_compact_flags(d.narrow_cand, work, p.max_pairs, n_small,
                              scan_status + RB_SCAN_SMALL * scan_stride, s,
                              d.counters + RB_CNT_PAIRS, d.counters + RB_CNT_OVERFLOW, n_work, true);
    if (e0 != hipSuccess) return e0;
    RB_LAUNCH(k_narrow_face, dim3(grid), dim3(NP_BLOCK), 0, s, p.max_pairs, n_work, work,
                       d.pairs, d.pos, d.quat, d.inv_mass, d.shape,
                       d.hulls, d.verts, d.planes, d.has_contact, d.narrow_cand, d.face_best,
                       d.counters + RB_CNT_OVERFLOW);
    RB_LAUNCH(k_narrow_face_small, dim3(grid), dim3(NP_BLOCK), 0, s, p.max_pairs, n_work, n_small,
              work, d.pairs, d.pos, d.quat, d.inv_mass, d.shape, d.hulls, d.verts, d.planes,
              d.has_contact, d.narrow_cand, d.face_best);
    if (p.has_vertexless)
        RB_LAUNCH(k_narrow_special, dim3(rb_num_cus() * 4), dim3(256), 0, s, p.max_pairs,
                           d.counters + RB_CNT_PAIRS, d.pairs, d.pos, d.quat, d.inv_mass, d.shape,
                           d.hulls, d.planes, p.friction, d.contact_slots, d.has_contact);
    hipError_t e = launch_exclusive_scan(d.narrow_cand, d.cand_off, p.max_pairs,
                                         (uint32_t *)(d.counters + RB_CNT_NARROW_CAND),
                                         scan_status + RB_SCAN_CAND * scan_stride, s,
                                         d.counters + RB_CNT_PAIRS, d.counters + RB_CNT_OVERFLOW,
                                         nullptr, nullptr, true);
    if (e != hipSuccess) return e;
    RB_LAUNCH(k_narrow_keys, dim3((p.max_pairs + 255) / 256), dim3(256), 0, s,
                       p.max_pairs, d.counters + RB_CNT_PAIRS, d.narrow_cand, d.cand_off,
                       d.face_best, d.cand_kv);
    e = launch_radix_sort(d.cand_kv, p.max_pairs, 16, 8, sw, s, d.counters + RB_CNT_NARROW_CAND);
    if (e != hipSuccess) return e;
    // large-hull candidates [0, split) a wave each, then the small-hull ones
    // [split, n) SG lanes each
    RB_LAUNCH(k_clip_split, dim3((p.max_pairs + 255) / 256), dim3(256), 0, s,
              (const uint32_t *)(d.counters + RB_CNT_NARROW_CAND), d.cand_kv,
              d.counters + RB_CNT_CLIP_SPLIT);
    RB_LAUNCH(k_narrow_clip, dim3(grid), dim3(NP_BLOCK), 0, s,
                       d.counters + RB_CNT_CLIP_SPLIT, d.cand_kv, d.face_best,
                       d.pairs, d.pos, d.quat, d.inv_mass, d.shape, d.hulls, d.verts, d.planes,
                       d.face_index, d.indices, d.edges, p.friction, p.threshold, d.contact_slots,
                       d.has_contact, d.owner, d.counters + RB_CNT_OVERFLOW, p.sat_edge_reject,
                       p.reduce_mode);
    RB_LAUNCH(k_narrow_clip_small, dim3(grid), dim3(NP_BLOCK), 0, s,
              (const uint32_t *)(d.counters + RB_CNT_NARROW_CAND), d.counters + RB_CNT_CLIP_SPLIT,
              d.cand_kv, d.face_best, d.pairs, d.pos, d.quat, d.inv_mass, d.shape, d.hulls,
              d.verts, d.planes, d.face_index, d.indices, d.edges, p.friction, p.threshold,
              d.contact_slots, d.has_contact, d.owner, d.counters + RB_CNT_OVERFLOW,
              p.sat_edge_reject, p.reduce_mode);
    return rb_launch_status();
}

hipError_t launch_compound_pairs(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                                 uint32_t *scan_status, size_t scan_stride) {
    RB_LAUNCH(k_proxy_transforms, dim3((p.n_proxies + 255) / 256), dim3(256), 0, s,
                       p.n_proxies, p.n_bodies, d.proxy, d.children, d.pos, d.quat);
    const int grid = rb_num_cus() * 4;
    const MeshView mv{d.pos, d.quat, d.aabb_min, d.aabb_max, d.verts, d.children};
    RB_LAUNCH(k_compound_count, dim3(grid), dim3(256), 0, s, p.max_pairs,
                       d.counters + RB_CNT_PAIRS, d.pairs, d.shape, d.inv_mass, d.hulls, mv,
                       d.narrow_cand, d.counters + RB_CNT_OVERFLOW);
    hipError_t e = launch_exclusive_scan(d.narrow_cand, d.cand_off, p.max_pairs,
                                         (uint32_t *)(d.counters + RB_CNT_NARROW_CAND),
                                         scan_status + RB_SCAN_COMPOUND * scan_stride, s,
                                         d.counters + RB_CNT_PAIRS, d.counters + RB_CNT_OVERFLOW,
                                         nullptr, nullptr, true);
    if (e != hipSuccess) return e;
    RB_LAUNCH(k_compound_write, dim3(grid), dim3(256), 0, s, p.max_pairs,
                       d.counters + RB_CNT_PAIRS, d.pairs, d.shape, d.hulls, mv, d.proxy_first,
                       p.n_bodies, d.narrow_cand, d.cand_off);
    RB_LAUNCH(k_compound_total, dim3(1), dim3(1), 0, s, p.max_pairs,
                       d.counters + RB_CNT_PAIRS, (const uint32_t *)(d.counters + RB_CNT_NARROW_CAND),
                       d.counters + RB_CNT_OVERFLOW);
    return rb_launch_status();
}
