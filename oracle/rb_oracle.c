/*
 * rb_oracle.c -- CPU oracle for the GPU rigid-body step.  TEST INFRASTRUCTURE.
 *
 * Restates, serially and in plain C99, the reference algorithms of the
 * OpenCL rigid-body pipeline (erwincoumans/experiments).  Each function names
 * the reference file:line it follows.  It is loaded only by the tests, the
 * smoke check and the CPU-baseline leg of the benchmark; the product never
 * links it.  Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include "rb_oracle.h"
#include <omp.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* float4 helpers following the reference's OpenCL helper definitions.      */
/* dot3F4 (sat.cl:118-124) is evaluated as ((x*x + y*y) + z*z).              */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z, w; } f4;

static f4 ld4(const float *p) { f4 r = {p[0], p[1], p[2], p[3]}; return r; }
static void st4(float *p, f4 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w; }
static f4 mk4(float x, float y, float z, float w) { f4 r = {x, y, z, w}; return r; }
static f4 add4(f4 a, f4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static f4 sub4(f4 a, f4 b) { return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static f4 neg4(f4 a) { return mk4(-a.x, -a.y, -a.z, -a.w); }
static f4 scl4(f4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }
static float dot3(f4 a, f4 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* OpenCL cross() on float4 returns w = 0. */
static f4 cross3(f4 a, f4 b)
{
    return mk4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f);
}
/* qtMul (sat.cl:158-166): ans = cross(a,b) + (a.w*b + b.w*a); ans.w = a.w*b.w - dot3(a,b) */
static f4 qt_mul(f4 a, f4 b)
{
    f4 c = cross3(a, b);
    f4 t = mk4(a.w * b.x + b.w * a.x, a.w * b.y + b.w * a.y, a.w * b.z + b.w * a.z, 0.f);
    f4 r = mk4(c.x + t.x, c.y + t.y, c.z + t.z, a.w * b.w - dot3(a, b));
    return r;
}
static f4 qt_inv(f4 q) { return mk4(-q.x, -q.y, -q.z, q.w); }
/* qtRotate (sat.cl:175-182): q * (v.xyz, 0) * q^-1 */
static f4 qt_rotate(f4 q, f4 v)
{
    f4 vc = mk4(v.x, v.y, v.z, 0.f);
    return qt_mul(qt_mul(q, vc), qt_inv(q));
}
static f4 qt_inv_rotate(f4 q, f4 v) { return qt_rotate(qt_inv(q), v); }
/* transform (sat.cl:197-201) */
static f4 xform(f4 p, f4 t, f4 q) { return add4(qt_rotate(q, p), t); }
/* normalize3 / fast_normalize (sat.cl:205-210): IEEE v * (1/sqrt(|v|^2)) */
static f4 normalize3(f4 a)
{
    float l2 = dot3(a, a);
    float inv = 1.0f / sqrtf(l2);
    return mk4(a.x * inv, a.y * inv, a.z * inv, 0.f);
}

/* qtGetRotationMatrix with fabs (broadphaseKernel.cl:92-113) */
static void abs_rot(f4 q, f4 m[3])
{
    float x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z;
    m[0] = mk4(fabsf(1 - 2 * y2 - 2 * z2), fabsf(2 * q.x * q.y - 2 * q.w * q.z),
               fabsf(2 * q.x * q.z + 2 * q.w * q.y), 0.f);
    m[1] = mk4(fabsf(2 * q.x * q.y + 2 * q.w * q.z), fabsf(1 - 2 * x2 - 2 * z2),
               fabsf(2 * q.y * q.z - 2 * q.w * q.x), 0.f);
    m[2] = mk4(fabsf(2 * q.x * q.z - 2 * q.w * q.y), fabsf(2 * q.y * q.z + 2 * q.w * q.x),
               fabsf(1 - 2 * x2 - 2 * y2), 0.f);
}

/* ======================================================================== */
/* Primitives                                                               */
/* ======================================================================== */

/* Stable LSD radix sort of (key,value) pairs by the low sort_bits of the
 * unsigned key, 8-bit digits.  Follows btRadixSort32CL::executeHost
 * (opencl/broadphase_benchmark/btRadixSort32CL.cpp:70-138). kv is n pairs. */
void orc_radix_sort_kv(uint32_t *kv, int n, int sort_bits)
{
    if (n <= 0) return;
    uint32_t *src = kv;
    uint32_t *work = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)n);
    uint32_t *dst = work;
    int count = 0;
    for (int start = 0; start < sort_bits; start += 8) {
        int tables[256], counter[256];
        for (int i = 0; i < 256; i++) tables[i] = 0;
        for (int i = 0; i < n; i++) tables[(src[2 * i] >> start) & 255]++;
        int sum = 0;
        for (int i = 0; i < 256; i++) {
            int c = tables[i];
            tables[i] = sum;
            sum += c;
            counter[i] = 0;
        }
        for (int i = 0; i < n; i++) {
            int t = (src[2 * i] >> start) & 255;
            int d = tables[t] + counter[t]++;
            dst[2 * d] = src[2 * i];
            dst[2 * d + 1] = src[2 * i + 1];
        }
        uint32_t *tmp = src; src = dst; dst = tmp;
        count++;
    }
    if (count & 1) memcpy(kv, src, sizeof(uint32_t) * 2 * (size_t)n);
    free(work);
}

/* Exclusive scan; returns the total.  btPrefixScanCL::executeHost
 * (opencl/broadphase_benchmark/btPrefixScanCL.cpp:101-126). */
uint32_t orc_exclusive_scan(const uint32_t *in, uint32_t *out, int n)
{
    uint32_t s = 0;
    for (int i = 0; i < n; i++) {
        uint32_t v = in[i];
        out[i] = s;
        s += v;
    }
    return s;
}

/* Bound search over sorted (key,value) pairs: mode 0 = lower, 1 = upper,
 * 2 = count (upper - lower, zero for absent keys).  Keys >= n_dst are skipped.
 * BoundSearch<TYPE_HOST>::execute
 * (opencl/primitives/AdlPrimitives/Search/BoundSearchHost.inl:44-105). */
void orc_bound_search(const uint32_t *kv, int n, uint32_t *dst, int n_dst, int mode)
{
    if (mode == 2) {
        uint32_t *lo = (uint32_t *)calloc((size_t)n_dst + 1, 4);
        uint32_t *hi = (uint32_t *)calloc((size_t)n_dst + 1, 4);
        orc_bound_search(kv, n, lo, n_dst, 0);
        orc_bound_search(kv, n, hi, n_dst, 1);
        for (int i = 0; i < n_dst; i++) dst[i] = hi[i] - lo[i];
        free(lo); free(hi);
        return;
    }
    if (mode == 0) {
        for (int i = 0; i < n; i++) {
            uint32_t ik = (i == 0) ? 0xffffffffu : kv[2 * (i - 1)];
            uint32_t jk = kv[2 * i];
            if (ik != jk && jk < (uint32_t)n_dst) dst[jk] = (uint32_t)i;
        }
    } else {
        for (int i = 0; i <= n; i++) {
            uint32_t ik = (i == 0) ? 0u : kv[2 * (i - 1)];
            uint32_t jk = (i == n) ? (uint32_t)n_dst : kv[2 * i];
            if (i > 0 && ik != jk && ik < (uint32_t)n_dst) dst[ik] = (uint32_t)i;
        }
    }
}

/* ======================================================================== */
/* Body setup and broadphase                                                */
/* ======================================================================== */

/* World AABB from the shape's local AABB: initializeGpuAabbsFull
 * (opencl/broadphase_benchmark/broadphaseKernel.cl:293-331). */
void orc_world_aabbs(int n, const float *pos, const float *quat, const int *shape,
                     const orc_hull *hulls, float *aabb_min, float *aabb_max)
{
    for (int i = 0; i < n; i++) {
        const orc_hull *h = &hulls[shape[i]];
        f4 p = ld4(pos + 4 * i), q = ld4(quat + 4 * i);
        f4 mn = ld4(h->aabb_min), mx = ld4(h->aabb_max);
        f4 he = mk4((mx.x - mn.x) * 0.5f, (mx.y - mn.y) * 0.5f, (mx.z - mn.z) * 0.5f, 0.f);
        f4 lc = mk4((mx.x + mn.x) * 0.5f, (mx.y + mn.y) * 0.5f, (mx.z + mn.z) * 0.5f, 0.f);
        f4 wc = xform(lc, p, q);
        f4 m[3];
        abs_rot(q, m);
        f4 ext = mk4(dot3(m[0], he), dot3(m[1], he), dot3(m[2], he), 0.f);
        st4(aabb_min + 4 * i, mk4(wc.x - ext.x, wc.y - ext.y, wc.z - ext.z, 0.f));
        st4(aabb_max + 4 * i, mk4(wc.x + ext.x, wc.y + ext.y, wc.z + ext.z, 0.f));
    }
}

/* getGridPos / getPosHash (opencl/3dGridBroadphase/Shared/bt3dGridBroadphaseOCL.cl:3-22) */
static void grid_pos(const float *mn, const float *mx, float inv_cell, const int dims[3], int gp[3])
{
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = (mn[k] + mx[k]) * 0.5f;
    for (int k = 0; k < 3; k++) gp[k] = ((int)floorf(c[k] * inv_cell)) & (dims[k] - 1);
}
static uint32_t pos_hash(int x, int y, int z, const int dims[3])
{
    x &= dims[0] - 1; y &= dims[1] - 1; z &= dims[2] - 1;
    return (uint32_t)(z * dims[1] * dims[0] + y * dims[0] + x);
}

/* Cell key per body: kCalcHashAABB (bt3dGridBroadphaseOCL.cl:26-47), with the
 * scene index in the bits above the cell hash, and large bodies (the overflow
 * list) keyed 0xffffffff so that they sort after every cell. */
void orc_calc_hash(int n, const float *aabb_min, const float *aabb_max, const int *scene,
                   const int *is_large, float inv_cell, const int dims[3], int cell_bits,
                   uint32_t *kv)
{
    for (int i = 0; i < n; i++) {
        uint32_t key;
        if (is_large && is_large[i]) {
            key = 0xffffffffu;
        } else {
            int gp[3];
            grid_pos(aabb_min + 4 * i, aabb_max + 4 * i, inv_cell, dims, gp);
            key = pos_hash(gp[0], gp[1], gp[2], dims);
            if (scene) key |= ((uint32_t)scene[i]) << cell_bits;
        }
        kv[2 * i] = key;
        kv[2 * i + 1] = (uint32_t)i;
    }
}

/* Inclusive AABB overlap: testAABBOverlap (bt3dGridBroadphaseOCL.cl:89-94) */
static int aabb_overlap(const float *mn0, const float *mx0, const float *mn1, const float *mx1)
{
    return (mn0[0] <= mx1[0]) && (mn1[0] <= mx0[0]) && (mn0[1] <= mx1[1]) &&
           (mn1[1] <= mx0[1]) && (mn0[2] <= mx1[2]) && (mn1[2] <= mx0[2]);
}

static int cmp_pair(const void *a, const void *b)
{
    const int32_t *p = (const int32_t *)a, *q = (const int32_t *)b;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    return 0;
}

typedef struct { int32_t *p; int n, cap, total; } pairbuf;
static void push_pair(pairbuf *b, int x, int y)
{
    if (b->n >= b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->p = (int32_t *)realloc(b->p, sizeof(int32_t) * 2 * (size_t)b->cap);
    }
    b->p[2 * b->n] = x; b->p[2 * b->n + 1] = y; b->n++;
}
static int finish_pairs(pairbuf *b, int max_pairs, int32_t *pairs)
{
    if (b->n) qsort(b->p, (size_t)b->n, 8, cmp_pair);
    int w = b->n < max_pairs ? b->n : max_pairs;
    if (w > 0) memcpy(pairs, b->p, (size_t)w * 8);
    free(b->p);
    return b->n;
}

/* Stateless restatement of the grid pair search kFindOverlappingPairs /
 * findPairsInCell (bt3dGridBroadphaseOCL.cl:100-213): every small body scans
 * the 27 cells around the cell of its AABB centre and keeps each other body of
 * lower index whose AABB overlaps (inclusive test).  Bodies on the overflow
 * list (large) are paired by brute force within their scene, as
 * kFindPairsLarge does (bt3dGridBroadphaseOCL.cl:216-264).  Pairs are emitted
 * as (x = higher index, y = lower index), as squeezePairCaches writes them
 * (opencl/broadphase_benchmark/computeAabbKernelOCL.cl:69-111), sorted by
 * (x, y).  Returns the number of pairs (may exceed max_pairs). */
int orc_find_pairs_grid(int n, const uint32_t *sorted_kv, const float *aabb_min,
                        const float *aabb_max, const int *scene, const int *is_large,
                        const int *scene_start, const int *scene_count, float inv_cell,
                        const int dims[3], int cell_bits, int max_pairs, int32_t *pairs)
{
    pairbuf b = {0, 0, 0, 0};
    int ncells = dims[0] * dims[1] * dims[2];
    (void)ncells;
    for (int s = 0; s < n; s++) {
        uint32_t key = sorted_kv[2 * s];
        if (key == 0xffffffffu) continue;
        int i = (int)sorted_kv[2 * s + 1];
        int sc = scene ? scene[i] : 0;
        int gp[3];
        grid_pos(aabb_min + 4 * i, aabb_max + 4 * i, inv_cell, dims, gp);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    uint32_t h = pos_hash(gp[0] + dx, gp[1] + dy, gp[2] + dz, dims);
                    uint32_t nk = h | (((uint32_t)sc) << cell_bits);
                    /* lower bound of nk in the sorted keys (cell start) */
                    int lo = 0, hi = n;
                    while (lo < hi) {
                        int mid = (lo + hi) >> 1;
                        if (sorted_kv[2 * mid] < nk) lo = mid + 1; else hi = mid;
                    }
                    for (int t = lo; t < n && sorted_kv[2 * t] == nk; t++) {
                        int j = (int)sorted_kv[2 * t + 1];
                        if (j < i && aabb_overlap(aabb_min + 4 * i, aabb_max + 4 * i,
                                                  aabb_min + 4 * j, aabb_max + 4 * j))
                            push_pair(&b, i, j);
                    }
                }
    }
    /* overflow list: every pair with at least one large body, within the scene */
    if (is_large) {
        for (int i = 0; i < n; i++) {
            if (!is_large[i]) continue;
            int sc = scene ? scene[i] : 0;
            int j0 = scene_start ? scene_start[sc] : 0;
            int j1 = scene_count ? j0 + scene_count[sc] : n;
            for (int j = j0; j < j1; j++) {
                if (j == i) continue;
                if (is_large[j] && j > i) continue; /* large-large pair once */
                if (aabb_overlap(aabb_min + 4 * i, aabb_max + 4 * i, aabb_min + 4 * j,
                                 aabb_max + 4 * j))
                    push_pair(&b, i > j ? i : j, i > j ? j : i);
            }
        }
    }
    return finish_pairs(&b, max_pairs, pairs);
}

/* Brute-force O(N^2) overlap set, the reference's own host check
 * (opencl/gpu_rigidbody_pipeline2/btGpuSapBroadphase.cpp:98-181), per scene. */
int orc_find_pairs_brute(int n, const float *aabb_min, const float *aabb_max, const int *scene,
                         int max_pairs, int32_t *pairs)
{
    pairbuf b = {0, 0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) {
            if (scene && scene[i] != scene[j]) continue;
            if (aabb_overlap(aabb_min + 4 * i, aabb_max + 4 * i, aabb_min + 4 * j,
                             aabb_max + 4 * j))
                push_pair(&b, i, j);
        }
    return finish_pairs(&b, max_pairs, pairs);
}

/* Sweep and prune (btGpuSapBroadphase::calculateOverlappingPairs,
 * opencl/gpu_rigidbody_pipeline2/btGpuSapBroadphase.cpp:91-434, with the
 * kernels flipFloatKernel / scatterKernel / computePairsKernelOriginal of
 * opencl/broadphase_benchmark/sap.cl:90-116, 290-320): the AABB mins along
 * `axis` as FloatFlip keys (negative floats: all bits flipped, others: the
 * sign bit), a stable radix sort, then each AABB in sorted order tests the
 * following ones until their min passes its max.  Every AABB takes part in
 * the sweep (the reference pairs its large AABBs by brute force instead).
 * Output as the grid: (x > y) sorted by (x, y).  Returns the pair count. */
int orc_sap_pairs(int n, const float *aabb_min, const float *aabb_max, int axis, int max_pairs,
                  int32_t *pairs)
{
    pairbuf b = {0, 0, 0, 0};
    uint32_t *kv = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        uint32_t u;
        memcpy(&u, aabb_min + 4 * i + axis, 4);
        kv[2 * i] = u ^ ((uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u);
        kv[2 * i + 1] = (uint32_t)i;
    }
    orc_radix_sort_kv(kv, n, 32);
    for (int s = 0; s < n; s++) {
        const int i = (int)kv[2 * s + 1];
        const float lim = aabb_max[4 * i + axis];
        for (int t = s + 1; t < n; t++) {
            const int j = (int)kv[2 * t + 1];
            if (lim < aabb_min[4 * j + axis]) break;
            if (aabb_overlap(aabb_min + 4 * i, aabb_max + 4 * i, aabb_min + 4 * j, aabb_max + 4 * j))
                push_pair(&b, i > j ? i : j, i > j ? j : i);
        }
    }
    free(kv);
    return finish_pairs(&b, max_pairs, pairs);
}

/* mtMul (broadphaseKernel.cl:130-148): ans[i][j] = dot3(a.row[i], b.col[j]) */
static void mt_mul(const f4 a[3], const f4 b[3], f4 out[3])
{
    for (int i = 0; i < 3; i++) {
        f4 c0 = mk4(b[0].x, b[1].x, b[2].x, 0.f);
        f4 c1 = mk4(b[0].y, b[1].y, b[2].y, 0.f);
        f4 c2 = mk4(b[0].z, b[1].z, b[2].z, 0.f);
        out[i] = mk4(dot3(a[i], c0), dot3(a[i], c1), dot3(a[i], c2), 0.f);
    }
}

/* setupBodiesKernel (broadphaseKernel.cl:163-212): gravity on dynamic bodies
 * and world inverse inertia |R| I^-1 |R|^T; static bodies get zero velocity. */
void orc_setup_bodies(int n, float *linvel, const float *angvel, const float *quat,
                      const float *inv_mass, const float *inv_inertia_local,
                      float *inv_inertia_world, const float gravity[3], float dt)
{
    (void)angvel;
    for (int i = 0; i < n; i++) {
        if (inv_mass[i] != 0.f) {
            float *v = linvel + 4 * i;
            v[0] = v[0] + gravity[0] * dt;
            v[1] = v[1] + gravity[1] * dt;
            v[2] = v[2] + gravity[2] * dt;
            f4 m[3], mt[3], loc[3], tmp[3], out[3];
            abs_rot(ld4(quat + 4 * i), m);
            mt[0] = mk4(m[0].x, m[1].x, m[2].x, 0.f);
            mt[1] = mk4(m[0].y, m[1].y, m[2].y, 0.f);
            mt[2] = mk4(m[0].z, m[1].z, m[2].z, 0.f);
            const float *d = inv_inertia_local + 4 * i;
            loc[0] = mk4(d[0], 0.f, 0.f, 0.f);
            loc[1] = mk4(0.f, d[1], 0.f, 0.f);
            loc[2] = mk4(0.f, 0.f, d[2], 0.f);
            mt_mul(m, loc, tmp);
            mt_mul(tmp, mt, out);
            for (int r = 0; r < 3; r++) st4(inv_inertia_world + 12 * i + 4 * r, out[r]);
        } else {
            for (int r = 0; r < 12; r++) inv_inertia_world[12 * i + r] = 0.f;
        }
    }
}

/* ======================================================================== */
/* Narrowphase: SAT (sat.cl), clipping and reduction (satClipHullContacts.cl) */
/* ======================================================================== */

/* project (sat.cl:227-250): min/max of the hull's vertices along dir, computed
 * in the hull's local frame, plus dot(pos, dir). */
static void project(const orc_hull *h, const float *verts, f4 pos, f4 orn, f4 dir, float *mn,
                    float *mx)
{
    float lo = FLT_MAX, hi = -FLT_MAX;
    f4 ld = qt_inv_rotate(orn, dir);
    float off = dot3(pos, dir);
    for (int i = 0; i < h->num_vertices; i++) {
        float dp = dot3(ld4(verts + 4 * (h->vertex_offset + i)), ld);
        if (dp < lo) lo = dp;
        if (dp > hi) hi = dp;
    }
    if (lo > hi) { float t = lo; lo = hi; hi = t; }
    *mn = lo + off;
    *mx = hi + off;
}

/* TestSepAxis (sat.cl:476-494): 0 if the projections are disjoint, else the
 * overlap depth. */
static int test_sep_axis(const orc_hull *hA, const orc_hull *hB, const float *verts, f4 pA,
                         f4 oA, f4 pB, f4 oB, f4 axis, float *depth)
{
    float mn0, mx0, mn1, mx1;
    project(hA, verts, pA, oA, axis, &mn0, &mx0);
    project(hB, verts, pB, oB, axis, &mn1, &mx1);
    if (mx0 < mn1 || mx1 < mn0) return 0;
    float d0 = mx0 - mn1, d1 = mx1 - mn0;
    *depth = d0 < d1 ? d0 : d1;
    return 1;
}

/* findSeparatingAxis over the faces of hA (sat.cl:497-553) */
static int sep_faces(const orc_hull *hA, const orc_hull *hB, const float *verts,
                     const float *planes, f4 pA, f4 oA, f4 pB, f4 oB, f4 dc, f4 *sep, float *dmin)
{
    for (int f = 0; f < hA->num_faces; f++) {
        f4 nrm = ld4(planes + 4 * (hA->face_offset + f));
        f4 w = qt_rotate(oA, nrm);
        if (dot3(dc, w) < 0) w = scl4(w, -1.f);
        float d;
        if (!test_sep_axis(hA, hB, verts, pA, oA, pB, oB, w, &d)) return 0;
        if (d < *dmin) { *dmin = d; *sep = w; }
    }
    if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
    return 1;
}

/* findSeparatingAxisEdgeEdge (sat.cl:556-635): edge axes never reject; the
 * CPU twin (ConvexHullContact.cpp:415-453) rejects on the first edge axis
 * whose projections are disjoint.  Returns 0 when edge_reject is set and an
 * edge axis separates. */
static int sep_edges(const orc_hull *hA, const orc_hull *hB, const float *verts,
                     const float *edges, f4 pA, f4 oA, f4 pB, f4 oB, f4 dc, f4 *sep, float *dmin,
                     int edge_reject)
{
    for (int e0 = 0; e0 < hA->num_edges; e0++) {
        f4 ew0 = qt_rotate(oA, ld4(edges + 4 * (hA->edge_offset + e0)));
        for (int e1 = 0; e1 < hB->num_edges; e1++) {
            f4 ew1 = qt_rotate(oB, ld4(edges + 4 * (hB->edge_offset + e1)));
            f4 c = cross3(ew0, ew1);
            if (fabsf(c.x) > 1e-6f || fabsf(c.y) > 1e-6f || fabsf(c.z) > 1e-6f) {
                c = normalize3(c);
                if (dot3(dc, c) < 0) c = scl4(c, -1.f);
                float mn0, mx0, mn1, mx1;
                project(hA, verts, pA, oA, c, &mn0, &mx0);
                project(hB, verts, pB, oB, c, &mn1, &mx1);
                if (edge_reject && (mx0 < mn1 || mx1 < mn0)) return 0;
                float d0 = mx0 - mn1, d1 = mx1 - mn0;
                float dist = d0 < d1 ? d0 : d1;
                if (dist < *dmin) { *dmin = dist; *sep = c; }
            }
        }
    }
    if (dot3(neg4(dc), *sep) > 0.0f) *sep = neg4(*sep);
    return 1;
}

/* Convex-convex branch of findSeparatingAxisKernel (sat.cl:1195-1258).
 * Returns 1 when no face axis separates (the pair goes on to clipping), and
 * the minimum-overlap axis oriented along c0 - c1. */
int orc_sat(const orc_hull *hA, const orc_hull *hB, const float *verts, const float *planes,
            const float *edges, const float *posA, const float *ornA, const float *posB,
            const float *ornB, int edge_reject, float sep_out[4])
{
    f4 pA = ld4(posA), pB = ld4(posB), oA = ld4(ornA), oB = ld4(ornB);
    pA.w = 0.f; pB.w = 0.f;
    float dmin = FLT_MAX;
    f4 c0 = xform(ld4(hA->center), pA, oA);
    f4 c1 = xform(ld4(hB->center), pB, oB);
    f4 dc = sub4(c0, c1);
    f4 sep = mk4(0, 0, 0, 0);
    if (!sep_faces(hA, hB, verts, planes, pA, oA, pB, oB, dc, &sep, &dmin)) return 0;
    if (!sep_faces(hB, hA, verts, planes, pB, oB, pA, oA, dc, &sep, &dmin)) return 0;
    if (!sep_edges(hA, hB, verts, edges, pA, oA, pB, oB, dc, &sep, &dmin, edge_reject)) return 0;
    st4(sep_out, sep);
    return 1;
}

/* lerp3 (satClipHullContacts.cl:242-248) */
static f4 lerp3(f4 a, f4 b, float t)
{
    return mk4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, 0.f);
}

/* clipFace: Sutherland-Hodgman against one plane, keeping the back side
 * (satClipHullContacts.cl:303-346). */
int orc_clip_face(const float *in, int n_in, const float nrm[4], float eq, float *out)
{
    int n_out = 0;
    if (n_in <= 0) return 0;
    f4 pn = ld4(nrm);
    f4 first = ld4(in + 4 * (n_in - 1));
    float ds = dot3(pn, first) + eq;
    for (int ve = 0; ve < n_in; ve++) {
        f4 end = ld4(in + 4 * ve);
        float de = dot3(pn, end) + eq;
        if (ds < 0) {
            if (de < 0) {
                st4(out + 4 * n_out++, end);
            } else {
                st4(out + 4 * n_out++, lerp3(first, end, ds * 1.f / (ds - de)));
            }
        } else if (de < 0) {
            st4(out + 4 * n_out++, lerp3(first, end, ds * 1.f / (ds - de)));
            st4(out + 4 * n_out++, end);
        }
        first = end;
        ds = de;
    }
    return n_out;
}

/* clipHullAgainstHull + clipFaceAgainstHull (satClipHullContacts.cl:353-458, 556-606):
 * reference face of B (max dot with the axis) in world space, clipped against
 * the side planes of A's face with minimum dot; points with depth <= max_dist
 * below A's face are kept (depth clamped at min_dist). */
int orc_clip_hull_hull(const float sep_in[4], const orc_hull *hA, const orc_hull *hB,
                       const float *verts, const float *planes, const int *face_index,
                       const int *indices, const float *posA, const float *ornA,
                       const float *posB, const float *ornB, float min_dist, float max_dist,
                       float *out)
{
    f4 sep = ld4(sep_in), pA = ld4(posA), oA = ld4(ornA), pB = ld4(posB), oB = ld4(ornB);
    float b1[64 * 4], b2[64 * 4];
    int nb1 = 0;
    int closestB = -1;
    float dmax = -FLT_MAX;
    for (int f = 0; f < hB->num_faces; f++) {
        const float *pl = planes + 4 * (hB->face_offset + f);
        f4 wn = qt_rotate(oB, mk4(pl[0], pl[1], pl[2], 0.f));
        float d = dot3(wn, sep);
        if (d > dmax) { dmax = d; closestB = f; }
    }
    if (closestB < 0) return 0;
    {
        int fi = hB->face_offset + closestB;
        int ioff = face_index[2 * fi], ni = face_index[2 * fi + 1];
        for (int e = 0; e < ni && nb1 < 64; e++) {
            f4 v = ld4(verts + 4 * (hB->vertex_offset + indices[ioff + e]));
            st4(b1 + 4 * nb1++, xform(v, pB, oB));
        }
    }
    /* clipFaceAgainstHull */
    int closestA = -1;
    float dmin = FLT_MAX;
    for (int f = 0; f < hA->num_faces; f++) {
        const float *pl = planes + 4 * (hA->face_offset + f);
        f4 wn = qt_rotate(oA, mk4(pl[0], pl[1], pl[2], 0.f));
        float d = dot3(wn, sep);
        if (d < dmin) { dmin = d; closestA = f; }
    }
    if (closestA < 0) return 0;
    int fa = hA->face_offset + closestA;
    const float *plA = planes + 4 * fa;
    int ioffA = face_index[2 * fa], nA = face_index[2 * fa + 1];
    float *vin = b1, *vout = b2;
    int nin = nb1;
    for (int e0 = 0; e0 < nA; e0++) {
        f4 a = ld4(verts + 4 * (hA->vertex_offset + indices[ioffA + e0]));
        f4 b = ld4(verts + 4 * (hA->vertex_offset + indices[ioffA + (e0 + 1) % nA]));
        f4 edge0 = sub4(a, b);
        f4 we = qt_rotate(oA, edge0);
        f4 wpn = qt_rotate(oA, mk4(plA[0], plA[1], plA[2], 0.f));
        f4 pn = neg4(cross3(we, wpn));
        f4 wa = xform(a, pA, oA);
        float eq = -dot3(wa, pn);
        float tmp[4];
        st4(tmp, pn);
        int nout = orc_clip_face(vin, nin, tmp, eq, vout);
        float *t = vout; vout = vin; vin = t;
        nin = nout;
    }
    f4 wpn = qt_rotate(oA, mk4(plA[0], plA[1], plA[2], 0.f));
    float peq = plA[3] - dot3(wpn, pA);
    int nc = 0;
    for (int i = 0; i < nin; i++) {
        f4 p = ld4(vin + 4 * i);
        float depth = dot3(wpn, p) + peq;
        if (depth <= min_dist) depth = min_dist;
        if (depth <= max_dist) st4(out + 4 * nc++, mk4(p.x, p.y, p.z, depth));
    }
    return nc;
}

/* extractManifoldSequential (satClipHullContacts.cl:788-884): keep at most 4
 * of the clipped points: extremes along +u, -u, +v, -v in the contact plane,
 * compared as the float bits with the low 8 bits replaced by the point index,
 * taking the signed-integer maximum. */
/* extractManifoldSequentialGlobal (satClipHullContacts.cl:678-767), the split
 * path's reduction: up to 4 points pass through; else the points with the
 * smallest projection on u, -u, v and -v about the centre (each only from
 * projections below FLT_MIN, the reference's initial "maximum"; first of
 * equals), the deepest point replacing the first if it is none of them; 4
 * indices, which may repeat.  idx keeps 0, 1, 2, 3 where nothing qualifies. */
int orc_reduce_manifold_global(const float *pts, int n, const float normal[4], int idx[4])
{
    for (int i = 0; i < 4; i++) idx[i] = i;
    if (n == 0) return 0;
    if (n <= 4) return n;
    if (n > 64) n = 64;
    f4 c = ld4(pts);
    for (int j = 1; j < n; j++) c = add4(c, ld4(pts + 4 * j));
    {
        float fn = (float)n;
        c = mk4(c.x / fn, c.y / fn, c.z / fn, c.w / fn);
    }
    f4 nn = ld4(normal);
    f4 av = sub4(ld4(pts), c);
    f4 u = cross3(nn, av);
    f4 v = cross3(nn, u);
    u = normalize3(u);
    v = normalize3(v);
    float minw = FLT_MAX;
    int min_index = -1;
    float lim[4] = {FLT_MIN, FLT_MIN, FLT_MIN, FLT_MIN};
    for (int ie = 0; ie < n; ie++) {
        f4 p = ld4(pts + 4 * ie);
        if (p.w < minw) { minw = p.w; min_index = ie; }
        f4 r = sub4(p, c);
        float f[4] = {dot3(u, r), dot3(neg4(u), r), dot3(v, r), dot3(neg4(v), r)};
        for (int k = 0; k < 4; k++)
            if (f[k] < lim[k]) { lim[k] = f[k]; idx[k] = ie; }
    }
    if (idx[0] != min_index && idx[1] != min_index && idx[2] != min_index && idx[3] != min_index)
        idx[0] = min_index;
    return 4;
}

int orc_reduce_manifold(const float *pts, int n, const float normal[4], int idx[4])
{
    if (n == 0) return 0;
    if (n > 64) n = 64;
    f4 c = ld4(pts);
    for (int j = 1; j < n; j++) c = add4(c, ld4(pts + 4 * j));
    {
        float fn = (float)n;
        c = mk4(c.x / fn, c.y / fn, c.z / fn, c.w / fn);
    }
    if (n < 4) {
        for (int i = 0; i < n; i++) idx[i] = i;
        return n;
    }
    f4 nn = ld4(normal);
    f4 av = sub4(ld4(pts), c);
    f4 u = cross3(nn, av);
    f4 v = cross3(nn, u);
    u = normalize3(u);
    v = normalize3(v);
    int32_t best[4] = {INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN};
    for (int ie = 0; ie < n; ie++) {
        f4 r = sub4(ld4(pts + 4 * ie), c);
        float f[4] = {dot3(u, r), dot3(neg4(u), r), dot3(v, r), dot3(neg4(v), r)};
        for (int k = 0; k < 4; k++) {
            uint32_t bits;
            memcpy(&bits, &f[k], 4);
            int32_t key = (int32_t)((bits & 0xffffff00u) | (0xffu & (uint32_t)ie));
            if (ie == 0 || key > best[k]) best[k] = key;
        }
    }
    for (int k = 0; k < 4; k++) idx[k] = best[k] & 0xff;
    return 4;
}

/* Shapes without vertices: a sphere (no faces, radius in center[3]) or a
 * plane (one face: normal and constant, n . x = constant).  clipHullHullKernel's
 * first branches (satClipHullContacts.cl:1055-1118): sphere-sphere gives one
 * point on B's surface with the normal from A to B; plane-sphere gives the
 * sphere's deepest point with the plane's world normal, the plane body first.
 * The plane-sphere case is written here in closed form (the signed distance
 * of the sphere's centre from the plane in world space); the reference forms
 * the same point through relative transforms.  A hull against a sphere or a
 * plane, and a plane against a plane, give no contact, as in the reference. */
static int special_contact(int p, int a, int b, const orc_hull *hA, const orc_hull *hB,
                           const float *pos, const float *quat, const float *planes,
                           const float *inv_mass, float friction, orc_contact *c)
{
    /* 0 hull, 1 sphere, 2 plane, 3 compound (its pair is expanded into child pairs) */
    int kA = hA->num_vertices > 0 ? 0 : (hA->num_edges > 0 ? 3 : (hA->num_faces > 0 ? 2 : 1));
    int kB = hB->num_vertices > 0 ? 0 : (hB->num_edges > 0 ? 3 : (hB->num_faces > 0 ? 2 : 1));
    memset(c, 0, sizeof(*c));
    c->coeffs = (uint32_t)(0.f * 0xffff) | ((uint32_t)(friction * 0xffff) << 16);
    c->pair_index = p;
    c->normal[3] = 1.f;
    if (kA == 1 && kB == 1) {
        f4 pa = mk4(pos[4 * a], pos[4 * a + 1], pos[4 * a + 2], 0.f);
        f4 pb = mk4(pos[4 * b], pos[4 * b + 1], pos[4 * b + 2], 0.f);
        float rA = hA->center[3], rB = hB->center[3];
        f4 d = sub4(pa, pb);
        float len = sqrtf(dot3(d, d));
        if (!(len <= rA + rB)) return 0;
        f4 n = len > 0.00001f ? mk4(d.x / len, d.y / len, d.z / len, 0.f) : mk4(1.f, 0.f, 0.f, 0.f);
        f4 q = add4(pb, scl4(n, rB));
        q.w = len - (rA + rB);
        st4(c->pos[0], q);
        c->normal[0] = -n.x; c->normal[1] = -n.y; c->normal[2] = -n.z;
        c->body_a = inv_mass[a] == 0 ? -a : a;
        c->body_b = inv_mass[b] == 0 ? -b : b;
        return 1;
    }
    if (!((kA == 2 && kB == 1) || (kA == 1 && kB == 2))) return 0;
    int ip = kA == 2 ? a : b, is = kA == 2 ? b : a;
    const orc_hull *hP = kA == 2 ? hA : hB, *hS = kA == 2 ? hB : hA;
    const float *eq = planes + 4 * hP->face_offset;
    f4 nw = qt_rotate(ld4(quat + 4 * ip), mk4(eq[0], eq[1], eq[2], 0.f));
    f4 pp = mk4(pos[4 * ip], pos[4 * ip + 1], pos[4 * ip + 2], 0.f);
    f4 ps = mk4(pos[4 * is], pos[4 * is + 1], pos[4 * is + 2], 0.f);
    float r = hS->center[3];
    float dist = dot3(nw, ps) - (eq[3] + dot3(nw, pp)) - r;
    if (!(dist < 0.f)) return 0;
    f4 q = sub4(ps, scl4(nw, r));
    q.w = dist;
    st4(c->pos[0], q);
    c->normal[0] = nw.x; c->normal[1] = nw.y; c->normal[2] = nw.z;
    c->body_a = inv_mass[ip] == 0 ? -ip : ip;
    c->body_b = inv_mass[is] == 0 ? -is : is;
    return 1;
}

/* World transforms of compound children (findCompoundPairsKernel and
 * clipCompoundsHullHullKernel, sat.cl:787-797, satClipHullContacts.cl:1228-1248):
 * proxy body n + k is child proxy[2k+1] of body proxy[2k]; position
 * qtRotate(orn, childPos) + pos (w = 0), orientation qtMul(orn, childOrn).
 * children: 12 floats per child (position xyzw, orientation xyzw, shape and
 * padding). */
void orc_compound_proxies(int P, int n, const int32_t *proxy, const float *children, float *pos,
                          float *quat)
{
    for (int k = 0; k < P; k++) {
        int b = proxy[2 * k], c = proxy[2 * k + 1];
        const float *ch = children + 12 * c;
        f4 pb = ld4(pos + 4 * b), qb = ld4(quat + 4 * b);
        f4 np = add4(qt_rotate(qb, mk4(ch[0], ch[1], ch[2], 0.f)), pb);
        np.w = 0.f;
        st4(pos + 4 * (n + k), np);
        st4(quat + 4 * (n + k), qt_mul(qb, ld4(ch + 4)));
    }
}

/* clipHullHullKernel convex-convex branch (satClipHullContacts.cl:1019-1181)
 * after findSeparatingAxisKernel (sat.cl:928-1261), one pair at a time.  The
 * manifold of pair p goes to slots[p]; has_contact[p] says whether it exists.
 * Returns the number of contacts. */
int orc_narrowphase(int num_pairs, const int32_t *pairs, const float *pos, const float *quat,
                    const float *inv_mass, const int *shape, const orc_hull *hulls,
                    const float *verts, const float *planes, const int *face_index,
                    const int *indices, const float *edges, float friction,
                    float max_dist, int sat_mode, int reduction_mode, orc_contact *slots,
                    int *has_contact)
{
    int total = 0;
    /* pairs are independent: spread them over the host cores (each writes
     * only its own slot, so the result does not depend on the thread count) */
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64)
    for (int p = 0; p < num_pairs; p++) {
        int a = pairs[2 * p], b = pairs[2 * p + 1];
        has_contact[p] = 0;
        memset(&slots[p], 0, sizeof(orc_contact));
        if (inv_mass[a] == 0 && inv_mass[b] == 0) continue;
        const orc_hull *hA = &hulls[shape[a]], *hB = &hulls[shape[b]];
        if (hA->num_vertices == 0 || hB->num_vertices == 0) {
            if (special_contact(p, a, b, hA, hB, pos, quat, planes, inv_mass, friction, &slots[p])) {
                has_contact[p] = 1;
                total++;
            }
            continue;
        }
        float sep[4];
        float pa[4], pb[4];
        memcpy(pa, pos + 4 * a, 16); memcpy(pb, pos + 4 * b, 16);
        pa[3] = 0.f; pb[3] = 0.f;
        if (!orc_sat(hA, hB, verts, planes, edges, pa, quat + 4 * a, pb, quat + 4 * b,
                     sat_mode == 1, sep))
            continue;
        float pts[64 * 4];
        int n = orc_clip_hull_hull(sep, hA, hB, verts, planes, face_index, indices, pos + 4 * a,
                                   quat + 4 * a, pos + 4 * b, quat + 4 * b, -1e30f, max_dist, pts);
        if (n <= 0) continue;
        float nrm[4] = {-sep[0], -sep[1], -sep[2], -sep[3]};
        int idx[4] = {-1, -1, -1, -1};
        int nr = reduction_mode == 1 ? orc_reduce_manifold_global(pts, n, nrm, idx)
                                     : orc_reduce_manifold(pts, n, nrm, idx);
        orc_contact *c = &slots[p];
        c->normal[0] = nrm[0]; c->normal[1] = nrm[1]; c->normal[2] = nrm[2];
        c->normal[3] = (float)nr;
        c->coeffs = (uint32_t)(0.f * 0xffff) | ((uint32_t)(friction * 0xffff) << 16);
        c->pair_index = p;
        c->body_a = inv_mass[a] == 0 ? -a : a;
        c->body_b = inv_mass[b] == 0 ? -b : b;
        for (int i = 0; i < nr; i++) memcpy(c->pos[i], pts + 4 * idx[i], 16);
        has_contact[p] = 1;
        total++;
    }
    return total;
}

/* ======================================================================== */
/* Solver (dynamics/basic_demo/Stubs)                                       */
/* Body registration inertia: CLPhysicsDemo::registerPhysicsInstance's world
 * AABB of the shape (CLPhysicsDemo.cpp:536-552: btTransformAabb with margin
 * 0.01, bullet2/LinearMath/btAabbUtil2.h:193-207, the rotation from
 * btMatrix3x3::setRotation's scalar form, btMatrix3x3.h:204-249) and
 * btGpuNarrowphaseAndSolver::registerRigidBody's box inertia of twice that
 * AABB's size (btGpuNarrowPhaseAndSolver.cpp:614-640).  Writes the inverse
 * diagonal (x, y, z, 0); a body without mass gets zeros. */
void orc_register_inertia(float mass, const float local_min[3], const float local_max[3],
                          const float pos[3], const float quat[4], float margin, float out[4])
{
    if (mass == 0.f) { out[0] = out[1] = out[2] = out[3] = 0.f; return; }
    float x = quat[0], y = quat[1], z = quat[2], w = quat[3];
    float d = ((x * x + y * y) + z * z) + w * w;
    float s = 2.0f / d;
    float xs = x * s, ys = y * s, zs = z * s;
    float wx = w * xs, wy = w * ys, wz = w * zs;
    float xx = x * xs, xy = x * ys, xz = x * zs;
    float yy = y * ys, yz = y * zs, zz = z * zs;
    float m[3][3] = {{1.0f - (yy + zz), xy - wz, xz + wy},
                     {xy + wz, 1.0f - (xx + zz), yz - wx},
                     {xz - wy, yz + wx, 1.0f - (xx + yy)}};
    float he[3], lc[3], mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
        he[k] = 0.5f * (local_max[k] - local_min[k]);
        he[k] = he[k] + margin;
        lc[k] = 0.5f * (local_max[k] + local_min[k]);
    }
    for (int r = 0; r < 3; r++) {
        float c = ((m[r][0] * lc[0] + m[r][1] * lc[1]) + m[r][2] * lc[2]) + pos[r];
        float e = (he[0] * fabsf(m[r][0]) + he[1] * fabsf(m[r][1])) + he[2] * fabsf(m[r][2]);
        mn[r] = c - e;
        mx[r] = c + e;
    }
    float lx = 2.f * (mx[0] - mn[0]), ly = 2.f * (mx[1] - mn[1]), lz = 2.f * (mx[2] - mn[2]);
    float ix = (mass / 12.0f) * (ly * ly + lz * lz);
    float iy = (mass / 12.0f) * (lx * lx + lz * lz);
    float iz = (mass / 12.0f) * (lx * lx + ly * ly);
    out[0] = 1.f / ix; out[1] = 1.f / iy; out[2] = 1.f / iz; out[3] = 0.f;
}

/* ======================================================================== */

/* mtMul1 / mtMul3 (solveContact.cl:130-153) */
static f4 mt_mul1(const float *m, f4 b)
{
    return mk4(dot3(ld4(m), b), dot3(ld4(m + 4), b), dot3(ld4(m + 8), b), 0.f);
}
static f4 mt_mul3(f4 a, const float *m)
{
    f4 cx = mk4(m[0], m[4], m[8], 0.f), cy = mk4(m[1], m[5], m[9], 0.f),
       cz = mk4(m[2], m[6], m[10], 0.f);
    return mk4(dot3(a, cx), dot3(a, cy), dot3(a, cz), 0.f);
}
/* calcJacCoeff (solverSetup.cl:450-459) */
static float jac_coeff(f4 a0, f4 a1, float imA, const float *IA, float imB, const float *IB)
{
    float j0 = imA, j1 = dot3(mt_mul3(a0, IA), a0), j2 = imB, j3 = dot3(mt_mul3(a1, IB), a1);
    return -1.f / (((j0 + j1) + j2) + j3);
}
/* calcRelVel (solverSetup.cl:444-447) */
static float rel_vel(f4 l0, f4 l1, f4 a0, f4 a1, f4 vA, f4 wA, f4 vB, f4 wB)
{
    return ((dot3(l0, vA) + dot3(a0, wA)) + dot3(l1, vB)) + dot3(a1, wB);
}
/* btPlaneSpace1 (solverSetup.cl:491-516) */
static void plane_space(f4 n, f4 *p, f4 *q)
{
    if (fabsf(n.z) > 0.70710678f) {
        float a = n.y * n.y + n.z * n.z;
        float k = 1.f / sqrtf(a);
        *p = mk4(0.f, -n.z * k, n.y * k, 0.f);
        *q = mk4(a * k, -n.x * p->z, n.x * p->y, 0.f);
    } else {
        float a = n.x * n.x + n.y * n.y;
        float k = 1.f / sqrtf(a);
        *p = mk4(-n.y * k, n.x * k, 0.f, 0.f);
        *q = mk4(-n.z * p->y, n.z * p->x, a * k, 0.f);
    }
}

void orc_plane_space(const float n[3], float t0[4], float t1[4])
{
    f4 p, q;
    plane_space(mk4(n[0], n[1], n[2], 0.f), &p, &q);
    st4(t0, p);
    st4(t1, q);
}

/* ContactToConstraintKernel / setConstraint4 (solverSetup.cl:520-660), with
 * restitution 0 and the friction coefficient as a parameter. */
void orc_contact_to_constraint(int nc, const orc_contact *contacts, const float *pos,
                               const float *linvel, const float *angvel, const float *inv_mass,
                               const float *inv_inertia_world, float dt, float drift,
                               float coeff, float friction, orc_constraint *out)
{
    for (int k = 0; k < nc; k++) {
        const orc_contact *src = &contacts[k];
        orc_constraint *c = &out[k];
        memset(c, 0, sizeof(*c));
        int a = abs(src->body_a), b = abs(src->body_b);
        f4 pA = ld4(pos + 4 * a), pB = ld4(pos + 4 * b);
        f4 vA = ld4(linvel + 4 * a), wA = ld4(angvel + 4 * a);
        f4 vB = ld4(linvel + 4 * b), wB = ld4(angvel + 4 * b);
        float imA = inv_mass[a], imB = inv_mass[b];
        const float *IA = inv_inertia_world + 12 * a, *IB = inv_inertia_world + 12 * b;
        c->body_a = (uint32_t)a;
        c->body_b = (uint32_t)b;
        float dt_inv = 1.f / dt;
        f4 n = ld4(src->normal);
        float npts = src->normal[3];
        c->linear[0] = -n.x; c->linear[1] = -n.y; c->linear[2] = -n.z; c->linear[3] = friction;
        for (int ic = 0; ic < 4; ic++) {
            f4 wp = ld4(src->pos[ic]);
            f4 r0 = sub4(wp, pA), r1 = sub4(wp, pB);
            if ((float)ic >= npts) { c->jac_inv[ic] = 0.f; continue; }
            f4 lin = neg4(n);
            f4 a0 = neg4(cross3(r0, n));
            f4 a1 = cross3(r1, n);
            c->jac_inv[ic] = jac_coeff(a0, a1, imA, IA, imB, IB);
            float rv = rel_vel(lin, neg4(lin), a0, a1, vA, wA, vB, wB);
            float e = 0.f;
            c->b[ic] = e * rv;
            c->b[ic] += ((wp.w + drift) * coeff) * dt_inv;
            c->lambda[ic] = 0.f;
        }
        if (npts > 0) {
            f4 cen = mk4(0, 0, 0, 0);
            for (int i = 0; (float)i < npts; i++) cen = add4(cen, ld4(src->pos[i]));
            cen = mk4(cen.x / npts, cen.y / npts, cen.z / npts, cen.w / npts);
            f4 t0, t1;
            plane_space(n, &t0, &t1);
            f4 r0 = sub4(cen, pA), r1 = sub4(cen, pB);
            f4 tg[2] = {t0, t1};
            for (int i = 0; i < 2; i++) {
                f4 a0 = neg4(cross3(r0, tg[i]));
                f4 a1 = cross3(r1, tg[i]);
                c->f_jac_inv[i] = jac_coeff(a0, a1, imA, IA, imB, IB);
                c->f_lambda[i] = 0.f;
            }
            st4(c->center, cen);
        }
        for (int i = 0; i < 4; i++) {
            if ((float)i < npts) memcpy(c->pos[i], src->pos[i], 16);
            else memset(c->pos[i], 0, 16);
        }
        c->batch = (uint32_t)src->pair_index;
    }
}

/* Warm-start inputs (no reference counterpart: the reference zeroes lambda
 * every step, solverSetup.cl:529-531).  For contact k with constraint cons[k]
 * (orc_contact_to_constraint on the same inputs) the body pair is looked up
 * in the previous step's (a, b)-sorted impulse cache: prev_imp holds the sum
 * of normal impulses, friction 0, friction 1 and the number of points;
 * prev_pts four (x, y, z, normal impulse) per entry.  16 features are written:
 *   n.xyz, 10 * min depth, vr.n, vr.t0, vr.t1, 1/mA, 1/mB,
 *   10 * previous impulses (3), npts / 4, found flag, -mean jac_inv, mean b
 * with vr the relative velocity of the manifold centre and t0, t1 from
 * plane_space(n).  With apply_previous a found pair's impulses become the
 * initial lambdas: each point with a row takes the normal impulse of the
 * nearest previous point (first of equal distances), friction the previous
 * friction impulses. */
void orc_contact_features(int nc, const orc_contact *contacts, orc_constraint *cons,
                          const float *pos, const float *linvel, const float *angvel,
                          const float *inv_mass, int nprev, const int *prev_keys,
                          const float *prev_imp, const float *prev_pts, int apply_previous,
                          const int *contact_keys, float *features)
{
    for (int k = 0; k < nc; k++) {
        const orc_contact *src = &contacts[k];
        orc_constraint *c = &cons[k];
        int a = abs(src->body_a), b = abs(src->body_b);
        f4 pA = ld4(pos + 4 * a), pB = ld4(pos + 4 * b);
        f4 z = mk4(0, 0, 0, 0);
        float imA = inv_mass[a], imB = inv_mass[b];
        f4 vA = imA != 0.f ? ld4(linvel + 4 * a) : z, wA = imA != 0.f ? ld4(angvel + 4 * a) : z;
        f4 vB = imB != 0.f ? ld4(linvel + 4 * b) : z, wB = imB != 0.f ? ld4(angvel + 4 * b) : z;
        f4 n = ld4(src->normal);
        float npts = src->normal[3];
        /* linear search: the cache is small in tests, order does not matter;
         * an entry is found by the contact's key (its pair entry) or, without
         * keys, by the body pair */
        f4 prev = z;
        int found = -1;
        const int ka = contact_keys ? contact_keys[2 * k] : a;
        const int kb = contact_keys ? contact_keys[2 * k + 1] : b;
        for (int i = 0; i < nprev; i++)
            if (prev_keys[2 * i] == ka && prev_keys[2 * i + 1] == kb) {
                prev = ld4(prev_imp + 4 * i);
                found = i;
                break;
            }
        f4 t0, t1;
        plane_space(n, &t0, &t1);
        f4 cen = npts > 0 ? ld4(c->center) : z;
        f4 rA = sub4(cen, pA), rB = sub4(cen, pB);
        f4 vr = sub4(add4(vA, cross3(wA, rA)), add4(vB, cross3(wB, rB)));
        float dmin = FLT_MAX, jsum = 0.f, bsum = 0.f;
        for (int i = 0; i < 4; i++)
            if ((float)i < npts) {
                dmin = fminf(dmin, src->pos[i][3]);
                jsum += c->jac_inv[i];
                bsum += c->b[i];
            }
        float inv_n = npts > 0 ? 1.f / npts : 0.f;
        float *f = features + 16 * (size_t)k;
        f[0] = n.x; f[1] = n.y; f[2] = n.z; f[3] = npts > 0 ? 10.f * dmin : 0.f;
        f[4] = dot3(vr, n); f[5] = dot3(vr, t0); f[6] = dot3(vr, t1); f[7] = imA;
        f[8] = imB; f[9] = prev.x * 10.f; f[10] = prev.y * 10.f; f[11] = prev.z * 10.f;
        f[12] = npts * 0.25f; f[13] = found >= 0 ? 1.f : 0.f; f[14] = -jsum * inv_n;
        f[15] = bsum * inv_n;
        if (apply_previous && found >= 0 && npts > 0) {
            const float *pp = prev_pts + 16 * (size_t)found;
            int pn = (int)prev.w;
            for (int i = 0; i < 4; i++) {
                float lbest = 0.f;
                if ((float)i < npts && c->jac_inv[i] != 0.f) {
                    float dbest = FLT_MAX;
                    for (int j = 0; j < pn; j++) {
                        float dx = src->pos[i][0] - pp[4 * j], dy = src->pos[i][1] - pp[4 * j + 1];
                        float dz = src->pos[i][2] - pp[4 * j + 2];
                        float d2 = (dx * dx + dy * dy) + dz * dz;
                        if (d2 < dbest) { dbest = d2; lbest = pp[4 * j + 3]; }
                    }
                }
                c->lambda[i] = lbest;
            }
            c->f_lambda[0] = prev.y;
            c->f_lambda[1] = prev.z;
        }
    }
}

/* Initial lambdas from network predictions y[k][3] (10 x impulse): the
 * normal impulse, clamped at zero, is shared over the points in proportion
 * to the lambdas cons already holds (the matched previous impulses), or
 * evenly over the points with a row when they sum to zero.  Friction: the
 * friction lambdas cons holds (the previous step's) where the matched
 * impulses sum above zero, else the predicted ones; clamped to friction x that
 * normal impulse, none without friction rows. */
void orc_warm_from_prediction(int nc, orc_constraint *cons, const float *y, float friction)
{
    for (int k = 0; k < nc; k++) {
        orc_constraint *c = &cons[k];
        int cnt = 0;
        for (int i = 0; i < 4; i++) cnt += c->jac_inv[i] != 0.f;
        float sum = ((c->lambda[0] + c->lambda[1]) + c->lambda[2]) + c->lambda[3];
        float ln = y[3 * k] / 10.f;
        ln = ln > 0.f ? ln : 0.f;
        float lv[4];
        for (int i = 0; i < 4; i++) {
            if (c->jac_inv[i] == 0.f) lv[i] = 0.f;
            else if (sum > 0.f) lv[i] = ln * (c->lambda[i] / sum);
            else lv[i] = cnt > 0 ? ln / (float)cnt : 0.f;
        }
        for (int i = 0; i < 4; i++) c->lambda[i] = lv[i];
        float lim = friction * ln;
        /* matched manifolds keep the previous step's friction impulses */
        float f0 = sum > 0.f ? c->f_lambda[0] : y[3 * k + 1] / 10.f;
        float f1 = sum > 0.f ? c->f_lambda[1] : y[3 * k + 2] / 10.f;
        f0 = fminf(fmaxf(f0, -lim), lim);
        f1 = fminf(fmaxf(f1, -lim), lim);
        if (c->f_jac_inv[0] == 0.f) { f0 = 0.f; f1 = 0.f; }
        c->f_lambda[0] = f0;
        c->f_lambda[1] = f1;
    }
}

/* Colour priority: a bijective 32-bit mix of the constraint index, so that the
 * colouring is a deterministic function of the constraint list. */
static uint32_t color_prio(uint32_t k)
{
    k ^= k >> 16; k *= 0x7feb352dU; k ^= k >> 15; k *= 0x846ca68bU; k ^= k >> 16;
    return k;
}
typedef struct { uint32_t prio; int idx; } prio_ent;
static int cmp_prio_desc(const void *a, const void *b)
{
    const prio_ent *p = (const prio_ent *)a, *q = (const prio_ent *)b;
    if (p->prio != q->prio) return p->prio > q->prio ? -1 : 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

/* Graph colouring of the constraints: no dynamic body appears twice in one
 * colour (the invariant the reference's CreateBatches establishes per cell,
 * dynamics/basic_demo/Stubs/batchingKernels.cl:130-341, with static bodies
 * free).  Greedy in decreasing priority order: each constraint takes the
 * smallest colour unused by the already-coloured constraints that share one of
 * its dynamic bodies.  Returns the number of colours, or -1 past 128. */
int orc_color_contacts(int nc, const orc_constraint *cs, const float *inv_mass, int n_bodies,
                       int *color)
{
    prio_ent *ord = (prio_ent *)malloc(sizeof(prio_ent) * (size_t)(nc ? nc : 1));
    for (int k = 0; k < nc; k++) { ord[k].prio = color_prio((uint32_t)k); ord[k].idx = k; }
    qsort(ord, (size_t)nc, sizeof(prio_ent), cmp_prio_desc);
    uint64_t *used = (uint64_t *)calloc((size_t)n_bodies * 2, 8);
    int ncol = 0;
    for (int t = 0; t < nc; t++) {
        int k = ord[t].idx;
        int a = (int)cs[k].body_a, b = (int)cs[k].body_b;
        uint64_t m0 = 0, m1 = 0;
        if (inv_mass[a] != 0.f) { m0 |= used[2 * a]; m1 |= used[2 * a + 1]; }
        if (inv_mass[b] != 0.f) { m0 |= used[2 * b]; m1 |= used[2 * b + 1]; }
        int c = -1;
        for (int i = 0; i < 128; i++) {
            uint64_t w = i < 64 ? m0 : m1;
            if (!((w >> (i & 63)) & 1u)) { c = i; break; }
        }
        if (c < 0) { free(ord); free(used); return -1; }
        color[k] = c;
        if (c + 1 > ncol) ncol = c + 1;
        if (inv_mass[a] != 0.f) used[2 * a + (c >> 6)] |= 1ull << (c & 63);
        if (inv_mass[b] != 0.f) used[2 * b + (c >> 6)] |= 1ull << (c & 63);
    }
    free(ord);
    free(used);
    return ncol;
}

float orc_sweep_residual(float max_dlambda_scaled, float max_lambda)
{
    float den = max_lambda > 1.f ? max_lambda : 1.f;
    return max_dlambda_scaled / den;
}

typedef struct { f4 v, w; } bvel;

/* solveContact (solveContact.cl:272-312) on one constraint; returns the row
 * residual contribution max |dlambda|*|1/jac_inv| and max |lambda|. */
static void solve_contact(orc_constraint *c, bvel *A, bvel *B, f4 pA, f4 pB, float imA,
                          const float *IA, float imB, const float *IB, float *rmax, float *lmax)
{
    f4 lin0 = ld4(c->linear);
    for (int ic = 0; ic < 4; ic++) {
        if (c->jac_inv[ic] == 0.f) continue;
        f4 wp = ld4(c->pos[ic]);
        f4 r0 = sub4(wp, pA), r1 = sub4(wp, pB);
        f4 n = neg4(lin0);
        f4 lin = neg4(n);
        f4 a0 = neg4(cross3(r0, n));
        f4 a1 = cross3(r1, n);
        float rl = rel_vel(lin0, neg4(lin0), a0, a1, A->v, A->w, B->v, B->w) + c->b[ic];
        rl *= c->jac_inv[ic];
        float prev = c->lambda[ic];
        float upd = prev + rl;
        upd = fmaxf(upd, 0.f);
        upd = fminf(upd, FLT_MAX);
        rl = upd - prev;
        c->lambda[ic] = upd;
        f4 li0 = scl4(scl4(lin, imA), rl);
        f4 li1 = scl4(scl4(neg4(lin), imB), rl);
        f4 ai0 = scl4(mt_mul1(IA, a0), rl);
        f4 ai1 = scl4(mt_mul1(IB, a1), rl);
        A->v = add4(A->v, li0); A->w = add4(A->w, ai0);
        B->v = add4(B->v, li1); B->w = add4(B->w, ai1);
        float sc = fabsf(rl) / fabsf(c->jac_inv[ic]);
        if (sc > *rmax) *rmax = sc;
        if (fabsf(upd) > *lmax) *lmax = fabsf(upd);
    }
}

/* solveFrictionConstraint (solveFriction.cl:297-430) on one constraint. */
static void solve_friction(orc_constraint *c, bvel *A, bvel *B, f4 pA, f4 pB, float imA,
                           const float *IA, float imB, const float *IB, float friction,
                           float *rmax, float *lmax)
{
    float sum = 0;
    for (int j = 0; j < 4; j++) sum += c->lambda[j];
    float mx = friction * sum, mn = -mx;
    if (c->f_jac_inv[0] == 0 && c->f_jac_inv[0] == 0) return;
    f4 cen = ld4(c->center);
    f4 n = neg4(ld4(c->linear));
    f4 t[2];
    plane_space(n, &t[0], &t[1]);
    f4 r0 = sub4(cen, pA), r1 = sub4(cen, pB);
    for (int i = 0; i < 2; i++) {
        f4 lin = neg4(t[i]);
        f4 a0 = neg4(cross3(r0, t[i]));
        f4 a1 = cross3(r1, t[i]);
        float rl = rel_vel(lin, neg4(lin), a0, a1, A->v, A->w, B->v, B->w);
        rl *= c->f_jac_inv[i];
        float prev = c->f_lambda[i];
        float upd = prev + rl;
        upd = fmaxf(upd, mn);
        upd = fminf(upd, mx);
        rl = upd - prev;
        c->f_lambda[i] = upd;
        f4 li0 = scl4(scl4(lin, imA), rl);
        f4 li1 = scl4(scl4(neg4(lin), imB), rl);
        f4 ai0 = scl4(mt_mul1(IA, a0), rl);
        f4 ai1 = scl4(mt_mul1(IB, a1), rl);
        A->v = add4(A->v, li0); A->w = add4(A->w, ai0);
        B->v = add4(B->v, li1); B->w = add4(B->w, ai1);
        float sc = fabsf(rl) / fabsf(c->f_jac_inv[i]);
        if (sc > *rmax) *rmax = sc;
        if (fabsf(upd) > *lmax) *lmax = fabsf(upd);
    }
    f4 ab = normalize3(sub4(pB, pA));
    f4 ac = normalize3(sub4(cen, pA));
    if (dot3(ab, ac) > 0.95f || (imA == 0.f || imB == 0.f)) {
        float na = dot3(n, A->w), nb = dot3(n, B->w);
        A->w = sub4(A->w, scl4(n, na * 0.1f));
        B->w = sub4(B->w, scl4(n, nb * 0.1f));
    }
}

/* Apply the impulses already held in the constraints (warm start). */
static void apply_warm(orc_constraint *c, bvel *A, bvel *B, f4 pA, f4 pB, float imA,
                       const float *IA, float imB, const float *IB)
{
    f4 lin0 = ld4(c->linear);
    for (int ic = 0; ic < 4; ic++) {
        if (c->jac_inv[ic] == 0.f || c->lambda[ic] == 0.f) continue;
        f4 wp = ld4(c->pos[ic]);
        f4 r0 = sub4(wp, pA), r1 = sub4(wp, pB);
        f4 n = neg4(lin0);
        f4 lin = neg4(n);
        f4 a0 = neg4(cross3(r0, n)), a1 = cross3(r1, n);
        float rl = c->lambda[ic];
        A->v = add4(A->v, scl4(scl4(lin, imA), rl));
        A->w = add4(A->w, scl4(mt_mul1(IA, a0), rl));
        B->v = add4(B->v, scl4(scl4(neg4(lin), imB), rl));
        B->w = add4(B->w, scl4(mt_mul1(IB, a1), rl));
    }
    if (c->f_jac_inv[0] == 0) return;
    f4 cen = ld4(c->center);
    f4 n = neg4(lin0);
    f4 t[2];
    plane_space(n, &t[0], &t[1]);
    f4 r0 = sub4(cen, pA), r1 = sub4(cen, pB);
    for (int i = 0; i < 2; i++) {
        float rl = c->f_lambda[i];
        if (rl == 0.f) continue;
        f4 lin = neg4(t[i]);
        f4 a0 = neg4(cross3(r0, t[i])), a1 = cross3(r1, t[i]);
        A->v = add4(A->v, scl4(scl4(lin, imA), rl));
        A->w = add4(A->w, scl4(mt_mul1(IA, a0), rl));
        B->v = add4(B->v, scl4(scl4(neg4(lin), imB), rl));
        B->w = add4(B->w, scl4(mt_mul1(IB, a1), rl));
    }
}

static void load_vel(bvel *bv, const float *linvel, const float *angvel, int i)
{
    bv->v = ld4(linvel + 4 * i);
    bv->w = ld4(angvel + 4 * i);
}
static void store_vel(const bvel *bv, float *linvel, float *angvel, int i, float im)
{
    if (im != 0.f) { st4(linvel + 4 * i, bv->v); st4(angvel + 4 * i, bv->w); }
    else { memset(linvel + 4 * i, 0, 16); memset(angvel + 4 * i, 0, 16); }
}

/* The solve schedule of Solver::solveContactConstraint
 * (dynamics/basic_demo/Stubs/Solver.cpp:561-711): all normal sweeps, then all
 * friction sweeps; within a sweep the constraints are visited in `order`
 * (colour-major, which is what the batched kernels' batch loop realises,
 * solveContact.cl:406-480).  With tol > 0 a sweep count is a maximum and the
 * loop stops once the sweep residual falls below tol.  Velocities are the
 * body arrays; static bodies read and write zero. */
void orc_solve(int nc, orc_constraint *cs, const int *order, float *linvel, float *angvel,
               const float *pos, const float *inv_mass, const float *inv_inertia_world,
               int normal_iters, int friction_iters, float friction, float tol,
               int *iters_used, float *residual)
{
    for (int t = 0; t < nc; t++) {
        orc_constraint *c = &cs[order ? order[t] : t];
        int a = (int)c->body_a, b = (int)c->body_b;
        bvel A, B;
        load_vel(&A, linvel, angvel, a);
        load_vel(&B, linvel, angvel, b);
        apply_warm(c, &A, &B, ld4(pos + 4 * a), ld4(pos + 4 * b), inv_mass[a],
                   inv_inertia_world + 12 * a, inv_mass[b], inv_inertia_world + 12 * b);
        store_vel(&A, linvel, angvel, a, inv_mass[a]);
        store_vel(&B, linvel, angvel, b, inv_mass[b]);
    }
    for (int phase = 0; phase < 2; phase++) {
        int iters = phase == 0 ? normal_iters : friction_iters;
        int done = 0;
        float r = 0.f;
        for (int it = 0; it < iters; it++) {
            float rmax = 0.f, lmax = 0.f;
            for (int t = 0; t < nc; t++) {
                orc_constraint *c = &cs[order ? order[t] : t];
                int a = (int)c->body_a, b = (int)c->body_b;
                bvel A, B;
                load_vel(&A, linvel, angvel, a);
                load_vel(&B, linvel, angvel, b);
                if (phase == 0)
                    solve_contact(c, &A, &B, ld4(pos + 4 * a), ld4(pos + 4 * b), inv_mass[a],
                                  inv_inertia_world + 12 * a, inv_mass[b],
                                  inv_inertia_world + 12 * b, &rmax, &lmax);
                else
                    solve_friction(c, &A, &B, ld4(pos + 4 * a), ld4(pos + 4 * b), inv_mass[a],
                                   inv_inertia_world + 12 * a, inv_mass[b],
                                   inv_inertia_world + 12 * b, friction, &rmax, &lmax);
                store_vel(&A, linvel, angvel, a, inv_mass[a]);
                store_vel(&B, linvel, angvel, b, inv_mass[b]);
            }
            done++;
            r = orc_sweep_residual(rmax, lmax);
            if (tol > 0.f && r < tol) break;
        }
        if (iters_used) iters_used[phase] = done;
        if (residual) residual[phase] = r;
    }
}

/* Independent segments of the visiting order (one per scene tile), each
 * solved on its own with its own tolerance stop: orc_solve on rows
 * order[seg_off[k] .. seg_off[k+1]).  iters/residual: 2 per segment. */
void orc_solve_segments(int nseg, const int *seg_off, orc_constraint *cs, const int *order,
                        float *linvel, float *angvel, const float *pos, const float *inv_mass,
                        const float *inv_inertia_world, int normal_iters, int friction_iters,
                        float friction, float tol, int *iters_used, float *residual)
{
    for (int k = 0; k < nseg; k++)
        orc_solve(seg_off[k + 1] - seg_off[k], cs, order + seg_off[k], linvel, angvel, pos,
                  inv_mass, inv_inertia_world, normal_iters, friction_iters, friction, tol,
                  iters_used + 2 * k, residual + 2 * k);
}

/* quatMult / quatNorm and integrateTransformsKernel2
 * (opencl/broadphase_benchmark/integrateKernel.cl:3-26, 102-158), together
 * with copyVelocitiesKernel's zeroing of static bodies (broadphaseKernel.cl:216-233). */
void orc_integrate(int n, float *pos, float *quat, const float *linvel, float *angvel,
                   const float *inv_mass, float dt, float damping, float max_ang)
{
    for (int i = 0; i < n; i++) {
        float *w = angvel + 4 * i;
        (void)inv_mass;
        w[0] *= damping; w[1] *= damping; w[2] *= damping;
        float l2 = (w[0] * w[0] + w[1] * w[1]) + w[2] * w[2];
        float ang = sqrtf(l2);
        if (ang * dt > max_ang) ang = max_ang / dt;
        float ax[3], s;
        if (ang < 0.001f)
            s = 0.5f * dt - (((dt * dt * dt) * 0.020833333333f) * ang) * ang;
        else
            s = sinf((0.5f * ang) * dt) / ang;
        ax[0] = w[0] * s; ax[1] = w[1] * s; ax[2] = w[2] * s;
        float dw = cosf((ang * dt) * 0.5f);
        float *q = quat + 4 * i;
        float q1[4] = {ax[0], ax[1], ax[2], dw};
        float r[4];
        r[0] = ((q1[3] * q[0] + q1[0] * q[3]) + q1[1] * q[2]) - q1[2] * q[1];
        r[1] = ((q1[3] * q[1] + q1[1] * q[3]) + q1[2] * q[0]) - q1[0] * q[2];
        r[2] = ((q1[3] * q[2] + q1[2] * q[3]) + q1[0] * q[1]) - q1[1] * q[0];
        r[3] = ((q1[3] * q[3] - q1[0] * q[0]) - q1[1] * q[1]) - q1[2] * q[2];
        float len = sqrtf(((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]) + r[3] * r[3]);
        if (len > 0.f) {
            float inv = 1.f / len;
            q[0] = r[0] * inv; q[1] = r[1] * inv; q[2] = r[2] * inv; q[3] = r[3] * inv;
        } else {
            q[0] = q[1] = q[2] = 0.f; q[3] = 1.f;
        }
        const float *v = linvel + 4 * i;
        pos[4 * i + 0] = pos[4 * i + 0] + v[0] * dt;
        pos[4 * i + 1] = pos[4 * i + 1] + v[1] * dt;
        pos[4 * i + 2] = pos[4 * i + 2] + v[2] * dt;
    }
}

/* ======================================================================== */
/* Learned warm start: MLP 16-64-64-3.  No reference counterpart (the       */
/* reference zeroes lambda each step, solverSetup.cl:529-531).  bf16 inputs */
/* and weights, fp32 accumulation, ReLU hidden layers, MSE loss.            */
/* ======================================================================== */
float orc_bf16_to_f32(uint16_t h)
{
    uint32_t u = ((uint32_t)h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
/* round to nearest even */
uint16_t orc_f32_to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
static float rbf(float x) { return orc_bf16_to_f32(orc_f32_to_bf16(x)); }

/* weights: W1[64][16], W2[64][64], W3[3][64] (row-major, out x in), biases
 * b1[64], b2[64], b3[3] in fp32. */
void orc_mlp_forward(int n, const float *x, const uint16_t *w, const float *bias, float *h1,
                     float *h2, float *y)
{
    const uint16_t *W1 = w, *W2 = w + 64 * 16, *W3 = w + 64 * 16 + 64 * 64;
    const float *b1 = bias, *b2 = bias + 64, *b3 = bias + 128;
    for (int s = 0; s < n; s++) {
        const float *xs = x + 16 * s;
        float *a1 = h1 + 64 * s, *a2 = h2 + 64 * s, *ys = y + 3 * s;
        for (int o = 0; o < 64; o++) {
            double acc = 0;
            for (int k = 0; k < 16; k++) acc += (double)orc_bf16_to_f32(W1[o * 16 + k]) * rbf(xs[k]);
            float v = (float)acc + b1[o];
            a1[o] = v > 0.f ? v : 0.f;
        }
        for (int o = 0; o < 64; o++) {
            double acc = 0;
            for (int k = 0; k < 64; k++) acc += (double)orc_bf16_to_f32(W2[o * 64 + k]) * rbf(a1[k]);
            float v = (float)acc + b2[o];
            a2[o] = v > 0.f ? v : 0.f;
        }
        for (int o = 0; o < 3; o++) {
            double acc = 0;
            for (int k = 0; k < 64; k++) acc += (double)orc_bf16_to_f32(W3[o * 64 + k]) * rbf(a2[k]);
            ys[o] = (float)acc + b3[o];
        }
    }
}

/* Gradients of loss = mean over samples of sum_o (y_o - t_o)^2 / 3 with
 * respect to the fp32 master weights (straight-through bf16 rounding), in the
 * same W/b layout as orc_mlp_forward. */
void orc_mlp_backward(int n, const float *x, const uint16_t *w, const float *bias,
                      const float *h1, const float *h2, const float *y, const float *target,
                      float *grad_w, float *grad_b, float *loss)
{
    (void)bias;
    const uint16_t *W2 = w + 64 * 16, *W3 = w + 64 * 16 + 64 * 64;
    int nw = 64 * 16 + 64 * 64 + 3 * 64;
    double *gw = (double *)calloc((size_t)nw, sizeof(double));
    double gb[131];
    memset(gb, 0, sizeof(gb));
    double L = 0;
    for (int s = 0; s < n; s++) {
        const float *xs = x + 16 * s, *a1 = h1 + 64 * s, *a2 = h2 + 64 * s;
        float dy[3];
        for (int o = 0; o < 3; o++) {
            float e = y[3 * s + o] - target[3 * s + o];
            L += (double)e * e;
            dy[o] = 2.f * e / (3.f * (float)n);
        }
        float d2[64], d1[64];
        for (int k = 0; k < 64; k++) {
            double acc = 0;
            for (int o = 0; o < 3; o++) acc += (double)dy[o] * orc_bf16_to_f32(W3[o * 64 + k]);
            d2[k] = a2[k] > 0.f ? (float)acc : 0.f;
        }
        for (int k = 0; k < 64; k++) {
            double acc = 0;
            for (int o = 0; o < 64; o++) acc += (double)d2[o] * orc_bf16_to_f32(W2[o * 64 + k]);
            d1[k] = a1[k] > 0.f ? (float)acc : 0.f;
        }
        for (int o = 0; o < 3; o++) {
            for (int k = 0; k < 64; k++) gw[64 * 16 + 64 * 64 + o * 64 + k] += (double)dy[o] * rbf(a2[k]);
            gb[128 + o] += dy[o];
        }
        for (int o = 0; o < 64; o++) {
            for (int k = 0; k < 64; k++) gw[64 * 16 + o * 64 + k] += (double)d2[o] * rbf(a1[k]);
            gb[64 + o] += d2[o];
        }
        for (int o = 0; o < 64; o++) {
            for (int k = 0; k < 16; k++) gw[o * 16 + k] += (double)d1[o] * rbf(xs[k]);
            gb[o] += d1[o];
        }
    }
    for (int i = 0; i < nw; i++) grad_w[i] = (float)gw[i];
    for (int i = 0; i < 131; i++) grad_b[i] = (float)gb[i];
    if (loss) *loss = (float)(L / (3.0 * (double)(n ? n : 1)));
    free(gw);
}

/* threads the parallel oracle loops use (OpenMP) */
int orc_num_threads(void) { return omp_get_max_threads(); }
