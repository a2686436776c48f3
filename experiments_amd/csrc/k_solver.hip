// k_solver.hip -- contact constraints, tile colouring, batched Gauss-Seidel
// solve and integration for gfx950.
//
// Constraint maths follows the reference solver:
//   setConstraint4      dynamics/basic_demo/Stubs/solverSetup.cl:520-610
//   solveContact        solveContact.cl:272-312 (normal rows, lambda >= 0)
//   solveFriction       solveFriction.cl:297-430 (2 tangents at the centre,
//                       |lambda| <= mu * sum lambda_n, angular damping)
//   integrateTransformsKernel2 + copyVelocitiesKernel
//                       opencl/broadphase_benchmark/integrateKernel.cl:102-158,
//                       broadphaseKernel.cl:216-233
// The schedule replaces the reference's 16x16 solver cells with 4 parity
// launches per iteration and the nondeterministic CreateBatches
// (batchingKernels.cl:130-341): constraints are grouped into tiles (one scene
// each, or XZ columns of a single large scene), each tile is coloured in LDS so
// that no dynamic body repeats within a colour, and one workgroup walks its
// tile's colours with barriers between them, keeping the tile's body velocities
// in LDS.  Spatial tiles are run in 4 phases (tile x and z parity) per sweep
// so that tiles sharing bodies never run together; scene tiles are independent
// and run every sweep in one launch.  An optional warm start applies the
// previous step's (or a learned prediction of the) impulses before the sweeps.
#include <float.h>
#include <hip/hip_runtime.h>

#include "dmath.h"
#include "rb_internal.h"

using namespace rb;

namespace {

constexpr int TB = 256;        // threads per tile workgroup (colouring, spatial solve tiles)
constexpr int TB_SCENE = 512;  // threads per scene tile in the solve
constexpr int HASH_SLOTS = 2 * RB_TILE_MAX_BODIES;

__device__ __forceinline__ uint32_t color_prio(uint32_t k) {
    k ^= k >> 16; k *= 0x7feb352dU; k ^= k >> 15; k *= 0x846ca68bU; k ^= k >> 16;
    return k;
}

__device__ __forceinline__ float jac_coeff(float4 a0, float4 a1, float imA, const float4 *IA,
                                           float imB, const float4 *IB) {
    float j0 = imA, j1 = dot3(mt_mul3(a0, IA), a0), j2 = imB, j3 = dot3(mt_mul3(a1, IB), a1);
    return -1.f / (((j0 + j1) + j2) + j3);
}
__device__ __forceinline__ float rel_vel(float4 l0, float4 l1, float4 a0, float4 a1, float4 vA,
                                         float4 wA, float4 vB, float4 wB) {
    return ((dot3(l0, vA) + dot3(a0, wA)) + dot3(l1, vB)) + dot3(a1, wB);
}
__device__ __forceinline__ void plane_space(float4 n, float4 &p, float4 &q) {
    if (fabsf(n.z) > 0.70710678f) {
        float a = n.y * n.y + n.z * n.z;
        float k = 1.f / sqrtf(a);
        p = mk4(0.f, -n.z * k, n.y * k, 0.f);
        q = mk4(a * k, -n.x * p.z, n.x * p.y, 0.f);
    } else {
        float a = n.x * n.x + n.y * n.y;
        float k = 1.f / sqrtf(a);
        p = mk4(-n.y * k, n.x * k, 0.f, 0.f);
        q = mk4(-n.z * p.y, n.z * p.x, a * k, 0.f);
    }
}

// Tile coordinates of a point (clamped, so neighbours stay neighbours).
__device__ __forceinline__ int2 tile_xz(float x, float z, const StepParams &sp) {
    int tx = (int)floorf((x - sp.tile_origin_x) * sp.tile_inv);
    int tz = (int)floorf((z - sp.tile_origin_z) * sp.tile_inv);
    tx = tx < 0 ? 0 : (tx >= sp.tiles_x ? sp.tiles_x - 1 : tx);
    tz = tz < 0 ? 0 : (tz >= sp.tiles_z ? sp.tiles_z - 1 : tz);
    return make_int2(tx, tz);
}

// ---- compaction of the pair slots + ContactToConstraintKernel ----
__global__ __launch_bounds__(256) void k_constraints(
    StepParams sp, const int *__restrict__ num_pairs_ptr, const uint32_t *__restrict__ has_contact,
    const uint32_t *__restrict__ contact_off, const Contact4 *__restrict__ slots,
    Contact4 *__restrict__ contacts, const float4 *__restrict__ pos,
    const float4 *__restrict__ linvel, const float4 *__restrict__ angvel,
    const float *__restrict__ inv_mass, const float4 *__restrict__ invI, const int *__restrict__ scene,
    const float4 *__restrict__ aabb_min, const float4 *__restrict__ aabb_max, ConstraintSoA cs,
    uint2 *__restrict__ tile_kv, int *counters, const int2 *__restrict__ prev_keys,
    const float4 *__restrict__ prev_imp, const float4 *__restrict__ prev_pts,
    const int *__restrict__ prev_count, float *__restrict__ features, int *__restrict__ solve_pos,
    const int2 *__restrict__ pairs) {
    int np = *num_pairs_ptr;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        // the contact count for the later kernels: the contact scan's total,
        // capped at the capacity (no kernel of this launch reads it)
        uint32_t t = (uint32_t)counters[RB_CNT_CONTACT_TOTAL];
        if (t > (uint32_t)sp.max_contacts) t = (uint32_t)sp.max_contacts;
        counters[RB_CNT_CONTACTS] = (int)t;
    }
    if (p >= np || !has_contact[p]) return;
    int k = (int)contact_off[p];
    if (k >= sp.max_contacts) {
        atomicOr(counters + RB_CNT_OVERFLOW, (int)MI_RB_OVERFLOW_CONTACTS);
        return;
    }
    Contact4 src = slots[p];
    contacts[k] = src;
    int a = abs(src.body_a), b = abs(src.body_b);
    float4 pA = pos[a], pB = pos[b];
    float4 vA = linvel[a], wA = angvel[a], vB = linvel[b], wB = angvel[b];
    float imA = inv_mass[a], imB = inv_mass[b];
    float4 IA[3] = {invI[3 * a], invI[3 * a + 1], invI[3 * a + 2]};
    float4 IB[3] = {invI[3 * b], invI[3 * b + 1], invI[3 * b + 2]};
    // static bodies: zero velocity and inverse inertia (setupBodiesKernel)
    if (imA == 0.f) { vA = mk4(0, 0, 0, 0); wA = vA; }
    if (imB == 0.f) { vB = mk4(0, 0, 0, 0); wB = vB; }
    float dt_inv = 1.f / sp.solver_dt;
    float4 n = src.normal;
    float npts = n.w;
    cs.linear[k] = mk4(-n.x, -n.y, -n.z, sp.friction);
    // (the loops over the 4 points are unrolled so that the point arrays and
    // the contact stay in registers)
    float jac[4], bb[4];
#pragma unroll
    for (int ic = 0; ic < 4; ic++) {
        float4 wp = src.pos[ic];
        float4 r0 = sub4(wp, pA), r1 = sub4(wp, pB);
        if ((float)ic >= npts) { jac[ic] = 0.f; bb[ic] = 0.f; continue; }
        float4 lin = neg4(n);
        float4 a0 = neg4(cross3(r0, n));
        float4 a1 = cross3(r1, n);
        jac[ic] = jac_coeff(a0, a1, imA, IA, imB, IB);
        float rv = rel_vel(lin, neg4(lin), a0, a1, vA, wA, vB, wB);
        float e = 0.f;
        float bv = e * rv;
        bv += ((wp.w + sp.drift) * sp.baumgarte) * dt_inv;
        bb[ic] = bv;
    }
    cs.jac_inv[k] = mk4(jac[0], jac[1], jac[2], jac[3]);
    cs.b[k] = mk4(bb[0], bb[1], bb[2], bb[3]);
    cs.lambda[k] = mk4(0, 0, 0, 0);
    float4 fr = mk4(0, 0, 0, 0);
    float4 cen = mk4(0, 0, 0, 0);
    if (npts > 0) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if ((float)i < npts) cen = add4(cen, src.pos[i]);
        cen = mk4(cen.x / npts, cen.y / npts, cen.z / npts, cen.w / npts);
        float4 t0, t1;
        plane_space(n, t0, t1);
        float4 r0 = sub4(cen, pA), r1 = sub4(cen, pB);
        fr.x = jac_coeff(neg4(cross3(r0, t0)), cross3(r1, t0), imA, IA, imB, IB);
        fr.y = jac_coeff(neg4(cross3(r0, t1)), cross3(r1, t1), imA, IA, imB, IB);
    }
    cs.center[k] = cen;
#pragma unroll
    for (int i = 0; i < 4; i++) cs.pos[4 * k + i] = (float)i < npts ? src.pos[i] : mk4(0, 0, 0, 0);
    cs.bodies[k] = make_int2(a, b);
    cs.pair[k] = p;
    solve_pos[k] = -1;  // set by the tile packing when the row is solved
    // ---- warm start: previous step's impulses of the same pair entry (the
    // body pair, or for a compound child or mesh triangle the child pair) ----
    int prev_n = *prev_count;
    float4 prev = mk4(0, 0, 0, 0);
    int has_prev = 0, lo_match = 0;
    {
        const int2 key = pairs[p];
        int lo = 0, hi = prev_n;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            int2 kk = prev_keys[mid];
            if (kk.x < key.x || (kk.x == key.x && kk.y < key.y)) lo = mid + 1; else hi = mid;
        }
        if (lo < prev_n && prev_keys[lo].x == key.x && prev_keys[lo].y == key.y) {
            prev = prev_imp[lo];
            has_prev = 1;
            lo_match = lo;
        }
    }
    // per-contact features of the learned warm start (16 floats)
    {
        float4 t0, t1;
        plane_space(n, t0, t1);
        float4 rA = sub4(cen, pA), rB = sub4(cen, pB);
        float4 va = add4(vA, cross3(wA, rA)), vb = add4(vB, cross3(wB, rB));
        float4 vr = sub4(va, vb);
        float dmin = FLT_MAX, jsum = 0.f, bsum = 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if ((float)i < npts) {
                dmin = fminf(dmin, src.pos[i].w);
                jsum += jac[i];
                bsum += bb[i];
            }
        float inv_n = npts > 0 ? 1.f / npts : 0.f;
        float f[16] = {n.x, n.y, n.z, npts > 0 ? 10.f * dmin : 0.f,
                       dot3(vr, n), dot3(vr, t0), dot3(vr, t1), imA,
                       imB, prev.x * RB_MLP_SCALE, prev.y * RB_MLP_SCALE, prev.z * RB_MLP_SCALE,
                       npts * 0.25f, (float)has_prev, -jsum * inv_n, bsum * inv_n};
        float4 *fo = reinterpret_cast<float4 *>(features + 16 * (size_t)k);
        fo[0] = mk4(f[0], f[1], f[2], f[3]);
        fo[1] = mk4(f[4], f[5], f[6], f[7]);
        fo[2] = mk4(f[8], f[9], f[10], f[11]);
        fo[3] = mk4(f[12], f[13], f[14], f[15]);
    }
    if (sp.warm_start != MI_RB_WARM_ZERO && has_prev && npts > 0) {
        // each point takes the normal impulse of the nearest point of the
        // pair's previous manifold (first of equals); the learned start
        // rescales these to its predicted total
        const int pn = (int)prev.w;
        float lv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if ((float)i >= npts || jac[i] == 0.f) continue;
            float4 q = src.pos[i];
            float dbest = FLT_MAX, lbest = 0.f;
            for (int j = 0; j < pn; j++) {
                float4 pp = prev_pts[4 * (size_t)lo_match + j];
                float dx = q.x - pp.x, dy = q.y - pp.y, dz = q.z - pp.z;
                float d2 = (dx * dx + dy * dy) + dz * dz;
                if (d2 < dbest) { dbest = d2; lbest = pp.w; }
            }
            lv[i] = lbest;
        }
        cs.lambda[k] = mk4(lv[0], lv[1], lv[2], lv[3]);
        fr.z = prev.y;
        fr.w = prev.z;
    }
    cs.fric[k] = fr;
    // Owner tile (the reference bins by body A's cell, SetSortDataKernel,
    // solverSetup2.cl:452-479).  Scene tiles: the scene.  Spatial tiles: the
    // minimum, per axis, of the tiles of the dynamic bodies' AABB centres; a
    // row then only touches bodies in its owner tile and the next tile in +x
    // and +z (AABB centres of touching bodies are at most one cell apart and a
    // tile is at least a cell), so tiles of equal (x, z) parity share no
    // dynamic body and 4 phases per sweep suffice.
    int t;
    if (sp.tile_mode == 1) {
        t = scene[imA != 0.f ? a : b];
    } else {
        float4 ca = mk4((aabb_min[a].x + aabb_max[a].x) * 0.5f, 0.f,
                        (aabb_min[a].z + aabb_max[a].z) * 0.5f, 0.f);
        float4 cb = mk4((aabb_min[b].x + aabb_max[b].x) * 0.5f, 0.f,
                        (aabb_min[b].z + aabb_max[b].z) * 0.5f, 0.f);
        int2 ta = tile_xz(ca.x, ca.z, sp), tb = tile_xz(cb.x, cb.z, sp);
        if (imA == 0.f) ta = tb;
        if (imB == 0.f) tb = ta;
        int tx = ta.x < tb.x ? ta.x : tb.x, tz = ta.y < tb.y ? ta.y : tb.y;
        t = tz * sp.tiles_x + tx;
    }
    tile_kv[k] = make_uint2((uint32_t)t, (uint32_t)k);
}

// tile_first/tile_cnt of every tile (zeroed by the step's first kernel):
// the run boundaries of the sorted tile keys
__global__ void k_tile_bounds2(const uint2 *__restrict__ tile_kv, const int *counters,
                               int *tile_first, int *tile_cnt) {
    int nc = counters[RB_CNT_CONTACTS];
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    uint32_t t = tile_kv[k].x;
    if (k == 0 || tile_kv[k - 1].x != t) tile_first[t] = k;
    if (k == nc - 1 || tile_kv[k + 1].x != t) tile_cnt[t] = k + 1;  // end of the run
}

// ---- per-tile colouring (Luby rounds in LDS) and colour-major order ----
struct TileLDS {
    int key[HASH_SLOTS];                 // body index per hash slot, -1 empty
    unsigned long long mark[HASH_SLOTS]; // (round << 32) | priority
    unsigned long long used[HASH_SLOTS]; // colours already taken by the body's rows
    short la[RB_TILE_MAX_ROWS], lb[RB_TILE_MAX_ROWS];  // local slots (-1 static)
    unsigned char color[RB_TILE_MAX_ROWS];
    int ccount[RB_MAX_LOCAL_COLORS + 1];
    int coff[RB_MAX_LOCAL_COLORS + 1];
    int remaining;
};

// open-addressing body -> LDS slot tables (SLOTS a power of two)
template <int SLOTS>
__device__ __forceinline__ int hash_insert(int *key, int body) {
    uint32_t h = ((uint32_t)body * 2654435761u) & (SLOTS - 1);
    for (int probe = 0; probe < SLOTS; probe++) {
        int old = atomicCAS(&key[h], -1, body);
        if (old == -1 || old == body) return (int)h;
        h = (h + 1) & (SLOTS - 1);
    }
    return -1;
}
template <int SLOTS>
__device__ __forceinline__ int hash_find(const int *key, int body) {
    uint32_t h = ((uint32_t)body * 2654435761u) & (SLOTS - 1);
    for (int probe = 0; probe < SLOTS; probe++) {
        int k = key[h];
        if (k == body) return (int)h;
        if (k == -1) return -1;
        h = (h + 1) & (SLOTS - 1);
    }
    return -1;
}

__global__ __launch_bounds__(TB) void k_tile_color(
    const uint2 *__restrict__ tile_kv, const int *__restrict__ tile_first,
    const int *__restrict__ tile_cnt, const int2 *__restrict__ bodies,
    const float *__restrict__ inv_mass, int *__restrict__ order, int *__restrict__ row_color,
    int *__restrict__ tile_color_off, int *__restrict__ tile_ncolors, int *counters) {
    __shared__ TileLDS L;
    const int t = blockIdx.x;
    const int first = tile_first[t];
    int R = tile_cnt[t] - first;
    if (R > RB_TILE_MAX_ROWS) {
        if (threadIdx.x == 0) atomicOr(counters + RB_CNT_OVERFLOW, (int)MI_RB_OVERFLOW_TILE);
        R = RB_TILE_MAX_ROWS;
    }
    for (int i = threadIdx.x; i < HASH_SLOTS; i += TB) {
        L.key[i] = -1;
        L.mark[i] = 0ull;
        L.used[i] = 0ull;
    }
    for (int i = threadIdx.x; i <= RB_MAX_LOCAL_COLORS; i += TB) L.ccount[i] = 0;
    if (threadIdx.x == 0) L.remaining = 0;
    __syncthreads();
    for (int r = threadIdx.x; r < R; r += TB) {
        int k = (int)tile_kv[first + r].y;
        int2 bd = bodies[k];
        int sa = inv_mass[bd.x] != 0.f ? hash_insert<HASH_SLOTS>(L.key, bd.x) : -1;
        int sb = inv_mass[bd.y] != 0.f ? hash_insert<HASH_SLOTS>(L.key, bd.y) : -1;
        L.la[r] = (short)sa;
        L.lb[r] = (short)sb;
        L.color[r] = 0xff;
    }
    __syncthreads();
    // Jones-Plassmann rounds: a row whose priority is the highest among the
    // uncoloured rows at both its dynamic bodies takes the lowest colour that
    // none of those bodies' coloured rows holds.  The rows taking colours in
    // one round share no body, so each body's colour mask has one writer.
    int ncol = 0;
    for (int round = 0; round < RB_MAX_LOCAL_COLORS; round++) {
        unsigned long long tag = ((unsigned long long)(round + 1) << 32);
        for (int r = threadIdx.x; r < R; r += TB) {
            if (L.color[r] != 0xff) continue;
            unsigned long long v = tag | color_prio(tile_kv[first + r].y);
            if (L.la[r] >= 0) atomicMax(&L.mark[L.la[r]], v);
            if (L.lb[r] >= 0) atomicMax(&L.mark[L.lb[r]], v);
        }
        __syncthreads();
        int left = 0;
        for (int r = threadIdx.x; r < R; r += TB) {
            if (L.color[r] != 0xff) continue;
            unsigned long long v = tag | color_prio(tile_kv[first + r].y);
            const int la = L.la[r], lb = L.lb[r];
            bool win = (la < 0 || L.mark[la] == v) && (lb < 0 || L.mark[lb] == v);
            if (win) {
                const unsigned long long u = (la >= 0 ? L.used[la] : 0ull) |
                                             (lb >= 0 ? L.used[lb] : 0ull);
                int c = RB_MAX_LOCAL_COLORS - 1;
                if (~u) c = __builtin_ctzll(~u);
                else atomicOr(counters + RB_CNT_OVERFLOW, (int)MI_RB_OVERFLOW_COLORS);
                L.color[r] = (unsigned char)c;
                if (la >= 0) L.used[la] |= 1ull << c;
                if (lb >= 0) L.used[lb] |= 1ull << c;
            } else {
                left++;
            }
        }
        if (left) atomicAdd(&L.remaining, left);
        __syncthreads();
        int rem = L.remaining;
        __syncthreads();
        if (threadIdx.x == 0) L.remaining = 0;
        if (rem == 0) break;
        __syncthreads();
    }
    __syncthreads();
    // rows still uncoloured: one colour each after the last (serial, rare)
    if (threadIdx.x == 0) {
        int extra = 0;
        for (int r = 0; r < R; r++)
            if (L.color[r] == 0xff) extra++;
        if (extra) atomicOr(counters + RB_CNT_OVERFLOW, (int)MI_RB_OVERFLOW_COLORS);
    }
    for (int r = threadIdx.x; r < R; r += TB) {
        if (L.color[r] == 0xff) L.color[r] = (unsigned char)(RB_MAX_LOCAL_COLORS - 1);
        atomicAdd(&L.ccount[L.color[r]], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int c = 0; c < RB_MAX_LOCAL_COLORS; c++) {
            L.coff[c] = run;
            tile_color_off[t * (RB_MAX_LOCAL_COLORS + 1) + c] = run;
            run += L.ccount[c];
            L.ccount[c] = 0;
        }
        tile_color_off[t * (RB_MAX_LOCAL_COLORS + 1) + RB_MAX_LOCAL_COLORS] = run;
        int nc = ncol;
        for (int c = 0; c < RB_MAX_LOCAL_COLORS; c++)
            if (L.coff[c] < run && c + 1 > nc) nc = c + 1;
        tile_ncolors[t] = nc;
        atomicMax(counters + RB_CNT_NCOLORS, nc);
    }
    __syncthreads();
    // colour-major, index order within a colour: rows are in index order here,
    // so a ballot-ranked stable scatter per colour keeps it
    for (int base = 0; base < R; base += TB) {
        int r = base + threadIdx.x;
        int c = r < R ? L.color[r] : -1;
        // stable rank among equal colours within this chunk: serial over waves
        for (int w = 0; w < TB / 64; w++) {
            if ((threadIdx.x >> 6) == w && r < R) {
                uint64_t peers = __ballot(true);
                for (int bit = 0; bit < 6; bit++) {
                    uint64_t bal = __ballot((c >> bit) & 1);
                    peers &= ((c >> bit) & 1) ? bal : ~bal;
                }
                int below = __popcll(peers & lanemask_lt());
                int before = L.ccount[c];
                wave_sync();
                if (below == 0) L.ccount[c] = before + __popcll(peers);
                wave_sync();
                int dst = L.coff[c] + before + below;
                int k = (int)tile_kv[first + r].y;
                order[first + dst] = k;
                row_color[k] = c;
            }
            __syncthreads();
        }
    }
}

// ---- solve ----
constexpr int SOLVE_SLOTS = 1024;  // LDS velocity cache of one tile (32 KB)
struct SolveLDS {
    float4 v[SOLVE_SLOTS];
    float4 w[SOLVE_SLOTS];
};
constexpr int SLOT_STATIC = -1;    // static body: zero velocity, no writes
constexpr int SLOT_GLOBAL = -2;    // dynamic body beyond the LDS cache: global memory

// A row body's velocities by its slot.  The LDS read is unconditional (slot
// clamped to 0) and the rare global and static cases overwrite it, so the
// compiler keeps an LDS read (ds_read) instead of merging the two sources
// into one generic-address load through a select of pointers (a flat load,
// whose latency is that of the vector memory path, plus a zero vector
// materialised in scratch).
__device__ __forceinline__ void fetch_vel(int h, int b, const SolveLDS &L,
                                          const float4 *__restrict__ linvel,
                                          const float4 *__restrict__ angvel, float4 &v,
                                          float4 &w) {
    const int s = h >= 0 ? h : 0;
    v = L.v[s];
    w = L.w[s];
    if (h < 0) {
        // (both slots read the body's global velocity and select by value)
        const float4 gv = linvel[b], gw = angvel[b];
        const bool g = h == SLOT_GLOBAL;
        v = g ? gv : mk4(0.f, 0.f, 0.f, 0.f);
        w = g ? gw : mk4(0.f, 0.f, 0.f, 0.f);
    }
}

// Constraint fields one row needs (Constraint4 minus the body indices).
struct RowData {
    float4 lin0, jac, bb, lam, fr, cen;
    float4 p[4];
};
// The row's two bodies: inverse mass, world inverse inertia, position and the
// LDS cache slot of the velocity (-1: static, or kept in global memory).
struct RowBodies {
    int a, b, ha, hb;
    float imA, imB;
    float4 IA[3], IB[3];
    float4 pA, pB;
};

// A row of the packed solve arrays (position `i` in solve order).
__device__ __forceinline__ void load_packed(int i, const PackedRows &pk, RowData &d, RowBodies &rb) {
    d.lin0 = pk.lin0[i];
    d.jac = pk.jac[i];
    d.bb = pk.b[i];
    d.lam = pk.lam[i];
    d.fr = pk.fric[i];
    d.cen = pk.cen[i];
    for (int q = 0; q < 4; q++) d.p[q] = pk.p[4 * (size_t)i + q];
    const int4 m = pk.meta[i];
    rb.a = m.x;
    rb.b = m.y;
    rb.ha = m.z;
    rb.hb = m.w;
    const float2 im = pk.im[i];
    rb.imA = im.x;
    rb.imB = im.y;
    for (int q = 0; q < 3; q++) {
        rb.IA[q] = pk.IA[3 * (size_t)i + q];
        rb.IB[q] = pk.IB[3 * (size_t)i + q];
    }
    rb.pA = pk.pA[i];
    rb.pB = pk.pB[i];
}

// solveContact (solveContact.cl:272-312): sequential impulses on the normal
// rows, lambda >= 0.
// RES: compute the sweep residual terms (the largest scaled impulse change
// and the largest impulse); launches whose sweep reports no residual skip
// them, which leaves every velocity and impulse unchanged.
template <bool RES = true>
__device__ __forceinline__ void row_normal(RowData &d, float4 &vA, float4 &wA, float4 &vB,
                                           float4 &wB, const RowBodies &rb, float &rmax,
                                           float &lmax) {
    const float4 lin0 = d.lin0;
    float jv[4] = {d.jac.x, d.jac.y, d.jac.z, d.jac.w};
    float bv[4] = {d.bb.x, d.bb.y, d.bb.z, d.bb.w};
    float lv[4] = {d.lam.x, d.lam.y, d.lam.z, d.lam.w};
    for (int ic = 0; ic < 4; ic++) {
        if (jv[ic] == 0.f) continue;
        float4 wp = d.p[ic];
        float4 r0 = sub4(wp, rb.pA), r1 = sub4(wp, rb.pB);
        float4 n = neg4(lin0);
        float4 lin = neg4(n);
        float4 a0 = neg4(cross3(r0, n));
        float4 a1 = cross3(r1, n);
        float rl = rel_vel(lin0, neg4(lin0), a0, a1, vA, wA, vB, wB) + bv[ic];
        rl *= jv[ic];
        float prev = lv[ic];
        float upd = prev + rl;
        upd = fmaxf(upd, 0.f);
        upd = fminf(upd, FLT_MAX);
        rl = upd - prev;
        lv[ic] = upd;
        float4 li0 = scl4(scl4(lin, rb.imA), rl);
        float4 li1 = scl4(scl4(neg4(lin), rb.imB), rl);
        float4 ai0 = scl4(mt_mul1(rb.IA, a0), rl);
        float4 ai1 = scl4(mt_mul1(rb.IB, a1), rl);
        vA = add4(vA, li0); wA = add4(wA, ai0);
        vB = add4(vB, li1); wB = add4(wB, ai1);
        if (RES) {
            float sc = fabsf(rl) / fabsf(jv[ic]);
            rmax = fmaxf(rmax, sc);
            lmax = fmaxf(lmax, fabsf(upd));
        }
    }
    d.lam = mk4(lv[0], lv[1], lv[2], lv[3]);
}

// solveFriction (solveFriction.cl:297-430): two tangents at the manifold
// centre, |lambda| <= mu * sum lambda_n, then the angular damping about n.
template <bool RES = true>
__device__ __forceinline__ void row_friction(RowData &d, float4 &vA, float4 &wA, float4 &vB,
                                             float4 &wB, const RowBodies &rb, float friction,
                                             float &rmax, float &lmax) {
    const float4 lam = d.lam;
    float sum = 0;
    sum += lam.x; sum += lam.y; sum += lam.z; sum += lam.w;
    float mx = friction * sum, mn = -mx;
    const float4 fr = d.fr;
    if (fr.x == 0 && fr.x == 0) return;
    const float4 cen = d.cen;
    float4 n = neg4(d.lin0);
    float4 t[2];
    plane_space(n, t[0], t[1]);
    float4 r0 = sub4(cen, rb.pA), r1 = sub4(cen, rb.pB);
    float fj[2] = {fr.x, fr.y};
    float fl[2] = {fr.z, fr.w};
    for (int i = 0; i < 2; i++) {
        float4 lin = neg4(t[i]);
        float4 a0 = neg4(cross3(r0, t[i]));
        float4 a1 = cross3(r1, t[i]);
        float rl = rel_vel(lin, neg4(lin), a0, a1, vA, wA, vB, wB);
        rl *= fj[i];
        float prev = fl[i];
        float upd = prev + rl;
        upd = fmaxf(upd, mn);
        upd = fminf(upd, mx);
        rl = upd - prev;
        fl[i] = upd;
        float4 li0 = scl4(scl4(lin, rb.imA), rl);
        float4 li1 = scl4(scl4(neg4(lin), rb.imB), rl);
        float4 ai0 = scl4(mt_mul1(rb.IA, a0), rl);
        float4 ai1 = scl4(mt_mul1(rb.IB, a1), rl);
        vA = add4(vA, li0); wA = add4(wA, ai0);
        vB = add4(vB, li1); wB = add4(wB, ai1);
        if (RES) {
            float sc = fabsf(rl) / fabsf(fj[i]);
            rmax = fmaxf(rmax, sc);
            lmax = fmaxf(lmax, fabsf(upd));
        }
    }
    d.fr = mk4(fr.x, fr.y, fl[0], fl[1]);
    float4 ab = normalize3(sub4(rb.pB, rb.pA));
    float4 ac = normalize3(sub4(cen, rb.pA));
    if (dot3(ab, ac) > 0.95f || (rb.imA == 0.f || rb.imB == 0.f)) {
        float na = dot3(n, wA), nb = dot3(n, wB);
        wA = sub4(wA, scl4(n, na * 0.1f));
        wB = sub4(wB, scl4(n, nb * 0.1f));
    }
}

// Apply the impulses a row already holds (warm start) to its bodies.
__device__ __forceinline__ void row_warm(const RowData &d, float4 &vA, float4 &wA, float4 &vB,
                                         float4 &wB, const RowBodies &rb) {
    const float4 lin0 = d.lin0;
    float jv[4] = {d.jac.x, d.jac.y, d.jac.z, d.jac.w};
    float lv[4] = {d.lam.x, d.lam.y, d.lam.z, d.lam.w};
    for (int ic = 0; ic < 4; ic++) {
        if (jv[ic] == 0.f || lv[ic] == 0.f) continue;
        float4 wp = d.p[ic];
        float4 r0 = sub4(wp, rb.pA), r1 = sub4(wp, rb.pB);
        float4 n = neg4(lin0);
        float4 lin = neg4(n);
        float4 a0 = neg4(cross3(r0, n)), a1 = cross3(r1, n);
        float rl = lv[ic];
        vA = add4(vA, scl4(scl4(lin, rb.imA), rl));
        wA = add4(wA, scl4(mt_mul1(rb.IA, a0), rl));
        vB = add4(vB, scl4(scl4(neg4(lin), rb.imB), rl));
        wB = add4(wB, scl4(mt_mul1(rb.IB, a1), rl));
    }
    const float4 fr = d.fr;
    if (fr.x == 0.f) return;
    float4 n = neg4(lin0);
    float4 t[2];
    plane_space(n, t[0], t[1]);
    float4 r0 = sub4(d.cen, rb.pA), r1 = sub4(d.cen, rb.pB);
    float fl[2] = {fr.z, fr.w};
    for (int i = 0; i < 2; i++) {
        float rl = fl[i];
        if (rl == 0.f) continue;
        float4 lin = neg4(t[i]);
        float4 a0 = neg4(cross3(r0, t[i])), a1 = cross3(r1, t[i]);
        vA = add4(vA, scl4(scl4(lin, rb.imA), rl));
        wA = add4(wA, scl4(mt_mul1(rb.IA, a0), rl));
        vB = add4(vB, scl4(scl4(neg4(lin), rb.imB), rl));
        wB = add4(wB, scl4(mt_mul1(rb.IB, a1), rl));
    }
}

// After the solve: the impulses go back from the packed solve arrays to the
// constraints, and into the previous-step impulse cache for the warm start,
// keyed by the constraint's pair entry (unique per manifold; constraints are
// ordered by pair, so without compound child pairs, which are appended after
// the body pairs, the keys come out sorted): the totals (sum of normal
// impulses, friction 0, friction 1, number of points) and each point's
// position with its normal impulse.
__global__ void k_save_prev(const int *counters, ConstraintSoA cs, const int *solve_pos,
                            const float4 *pk_lam, const float4 *pk_fric, int2 *prev_keys,
                            float4 *prev_imp, float4 *prev_pts, int *prev_count,
                            const int2 *pairs) {
    int nc = counters[RB_CNT_CONTACTS];
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *prev_count = nc;
    if (k >= nc) return;
    float4 l, f;
    const int pos = solve_pos[k];
    if (pos >= 0) {
        l = pk_lam[pos];
        f = pk_fric[pos];
        cs.lambda[k] = l;
        cs.fric[k] = f;
    } else {  // not solved (tile capacity exceeded): the initial impulses stay
        l = cs.lambda[k];
        f = cs.fric[k];
    }
    float4 jac = cs.jac_inv[k];
    int np = (jac.x != 0.f) + (jac.y != 0.f) + (jac.z != 0.f) + (jac.w != 0.f);
    prev_keys[k] = pairs[cs.pair[k]];
    prev_imp[k] = mk4(((l.x + l.y) + l.z) + l.w, f.z, f.w, (float)np);
    float lv[4] = {l.x, l.y, l.z, l.w};
    for (int i = 0; i < 4; i++) {
        float4 p = cs.pos[4 * (size_t)k + i];
        prev_pts[4 * (size_t)k + i] = mk4(p.x, p.y, p.z, lv[i]);
    }
}

// Per step, after the colouring: each tile's rows in solve order as packed
// arrays (constraint fields, both bodies' inverse mass, world inverse inertia
// and position, and their slots in a dense list of the tile's dynamic
// bodies), so that a solve launch reads them contiguously, with no hashing
// and no gathers through body indices.
struct PackLDS {
    int key[HASH_SLOTS];
    int dense[HASH_SLOTS];
    int count;
};
__global__ __launch_bounds__(TB) void k_tile_pack(
    const int *__restrict__ tile_first, const int *__restrict__ tile_cnt,
    const int *__restrict__ order, ConstraintSoA cs, const float4 *__restrict__ pos,
    const float *__restrict__ inv_mass, const float4 *__restrict__ invI, PackedRows pk,
    int *__restrict__ solve_pos, int *__restrict__ tile_bodies, int *__restrict__ tile_nb,
    const int *__restrict__ tile_ncolors, int4 *__restrict__ tile_hdr, int *counters,
    int precompute) {
    __shared__ PackLDS L;
    const int t = blockIdx.x, tid = threadIdx.x;
    const int first = tile_first[t];
    int R = tile_cnt[t] - first;
    if (R > RB_TILE_MAX_ROWS) R = RB_TILE_MAX_ROWS;
    if (R <= 0) {
        if (tid == 0) {
            tile_nb[t] = 0;
            tile_hdr[t] = make_int4(first, 0, 0, 0);
        }
        return;
    }
    for (int i = tid; i < HASH_SLOTS; i += TB) { L.key[i] = -1; L.dense[i] = -1; }
    if (tid == 0) L.count = 0;
    __syncthreads();
    for (int r = tid; r < R; r += TB) {
        int2 bd = cs.bodies[order[first + r]];
        bool fail = false;
        if (inv_mass[bd.x] != 0.f) fail |= hash_insert<HASH_SLOTS>(L.key, bd.x) < 0;
        if (inv_mass[bd.y] != 0.f) fail |= hash_insert<HASH_SLOTS>(L.key, bd.y) < 0;
        if (fail) atomicOr(counters + RB_CNT_OVERFLOW, (int)MI_RB_OVERFLOW_TILE);
    }
    __syncthreads();
    // dense slots: the occupied hash slots numbered in slot order
    for (int i = tid; i < HASH_SLOTS; i += TB)
        if (L.key[i] >= 0) L.dense[i] = atomicAdd(&L.count, 1);
    __syncthreads();
    const int nb = L.count;
    for (int i = tid; i < HASH_SLOTS; i += TB)
        if (L.key[i] >= 0) tile_bodies[2 * (size_t)first + L.dense[i]] = L.key[i];
    if (tid == 0) {
        tile_nb[t] = nb;
        tile_hdr[t] = make_int4(first, R, tile_ncolors[t], nb < SOLVE_SLOTS ? nb : SOLVE_SLOTS);
        atomicAdd(counters + RB_CNT_SOLVE_BODIES, nb);
    }
    for (int r = tid; r < R; r += TB) {
        const int i = first + r;
        const int k = order[i];
        const int2 bd = cs.bodies[k];
        const float imA = inv_mass[bd.x], imB = inv_mass[bd.y];
        int sa = SLOT_STATIC, sb = SLOT_STATIC;
        if (imA != 0.f) {
            int h = hash_find<HASH_SLOTS>(L.key, bd.x);
            sa = h >= 0 && L.dense[h] < SOLVE_SLOTS ? L.dense[h] : SLOT_GLOBAL;
        }
        if (imB != 0.f) {
            int h = hash_find<HASH_SLOTS>(L.key, bd.y);
            sb = h >= 0 && L.dense[h] < SOLVE_SLOTS ? L.dense[h] : SLOT_GLOBAL;
        }
        pk.lin0[i] = cs.linear[k];
        pk.jac[i] = cs.jac_inv[k];
        pk.b[i] = cs.b[k];
        pk.lam[i] = cs.lambda[k];
        pk.fric[i] = cs.fric[k];
        pk.cen[i] = cs.center[k];
        for (int q = 0; q < 4; q++) pk.p[4 * (size_t)i + q] = cs.pos[4 * (size_t)k + q];
        pk.meta[i] = make_int4(bd.x, bd.y, sa, sb);
        pk.im[i] = make_float2(imA, imB);
        for (int q = 0; q < 3; q++) {
            pk.IA[3 * (size_t)i + q] = invI[3 * (size_t)bd.x + q];
            pk.IB[3 * (size_t)i + q] = invI[3 * (size_t)bd.y + q];
        }
        const float4 pA = pos[bd.x], pB = pos[bd.y];
        pk.pA[i] = pA;
        pk.pB[i] = pB;
        solve_pos[k] = i;
        // velocity-independent terms of the normal row (as row_normal forms
        // them), read by the spatial tiles' normal sweeps only
        if (precompute) {
            const float4 lin0 = cs.linear[k];
            const float4 n = neg4(lin0), lin = neg4(n);
            float4 IA[3], IB[3];
            for (int q = 0; q < 3; q++) {
                IA[q] = invI[3 * (size_t)bd.x + q];
                IB[q] = invI[3 * (size_t)bd.y + q];
            }
            for (int q = 0; q < 4; q++) {
                const float4 wp = cs.pos[4 * (size_t)k + q];
                const float4 r0 = sub4(wp, pA), r1 = sub4(wp, pB);
                const float4 a0 = neg4(cross3(r0, n)), a1 = cross3(r1, n);
                pk.na0[4 * (size_t)i + q] = a0;
                pk.na1[4 * (size_t)i + q] = a1;
                pk.nia0[4 * (size_t)i + q] = mt_mul1(IA, a0);
                pk.nia1[4 * (size_t)i + q] = mt_mul1(IB, a1);
            }
            pk.nl[2 * (size_t)i] = scl4(lin, imA);
            pk.nl[2 * (size_t)i + 1] = scl4(neg4(lin), imB);
        }
    }
}

// One row of a sweep (mode 0 normal, 1 friction, 2 warm start), velocities
// from the LDS cache or, for bodies that did not fit it, global memory.
template <bool RES = true>
__device__ __forceinline__ void run_row(int mode, RowData &d, const RowBodies &rb, SolveLDS &L,
                                        float4 *linvel, float4 *angvel, float friction,
                                        float &rmax, float &lmax) {
    float4 vA, wA, vB, wB;
    fetch_vel(rb.ha, rb.a, L, linvel, angvel, vA, wA);
    fetch_vel(rb.hb, rb.b, L, linvel, angvel, vB, wB);
    if (mode == 0) row_normal<RES>(d, vA, wA, vB, wB, rb, rmax, lmax);
    else if (mode == 1) row_friction<RES>(d, vA, wA, vB, wB, rb, friction, rmax, lmax);
    else row_warm(d, vA, wA, vB, wB, rb);
    if (rb.ha >= 0) { L.v[rb.ha] = vA; L.w[rb.ha] = wA; }
    else if (rb.ha == SLOT_GLOBAL) { linvel[rb.a] = vA; angvel[rb.a] = wA; }
    if (rb.hb >= 0) { L.v[rb.hb] = vB; L.w[rb.hb] = wB; }
    else if (rb.hb == SLOT_GLOBAL) { linvel[rb.b] = vB; angvel[rb.b] = wB; }
}

// A normal row with its velocity-independent terms precomputed by
// k_tile_pack (the same operations as row_normal performs, so the results
// are identical): what a launch of normal sweeps needs.
struct NormalRow {
    float4 lin0, jac, bb, lam;
    float4 a0[4], a1[4], ia0[4], ia1[4];
    float4 l0, l1;  // lin0 * imA, -lin0 * imB
    int a, b, ha, hb;
};
__device__ __forceinline__ void load_normal(int i, const PackedRows &pk, NormalRow &r) {
    r.lin0 = pk.lin0[i];
    r.jac = pk.jac[i];
    r.bb = pk.b[i];
    r.lam = pk.lam[i];
    for (int q = 0; q < 4; q++) {
        r.a0[q] = pk.na0[4 * (size_t)i + q];
        r.a1[q] = pk.na1[4 * (size_t)i + q];
        r.ia0[q] = pk.nia0[4 * (size_t)i + q];
        r.ia1[q] = pk.nia1[4 * (size_t)i + q];
    }
    r.l0 = pk.nl[2 * (size_t)i];
    r.l1 = pk.nl[2 * (size_t)i + 1];
    const int4 m = pk.meta[i];
    r.a = m.x;
    r.b = m.y;
    r.ha = m.z;
    r.hb = m.w;
}
// row_normal on a NormalRow
template <bool RES = true>
__device__ __forceinline__ void row_normal_pre(NormalRow &d, float4 &vA, float4 &wA, float4 &vB,
                                               float4 &wB, float &rmax, float &lmax) {
    const float4 lin0 = d.lin0;
    float jv[4] = {d.jac.x, d.jac.y, d.jac.z, d.jac.w};
    float bv[4] = {d.bb.x, d.bb.y, d.bb.z, d.bb.w};
    float lv[4] = {d.lam.x, d.lam.y, d.lam.z, d.lam.w};
#pragma unroll
    for (int ic = 0; ic < 4; ic++) {
        if (jv[ic] == 0.f) continue;
        float rl = rel_vel(lin0, neg4(lin0), d.a0[ic], d.a1[ic], vA, wA, vB, wB) + bv[ic];
        rl *= jv[ic];
        float prev = lv[ic];
        float upd = prev + rl;
        upd = fmaxf(upd, 0.f);
        upd = fminf(upd, FLT_MAX);
        rl = upd - prev;
        lv[ic] = upd;
        vA = add4(vA, scl4(d.l0, rl)); wA = add4(wA, scl4(d.ia0[ic], rl));
        vB = add4(vB, scl4(d.l1, rl)); wB = add4(wB, scl4(d.ia1[ic], rl));
        if (RES) {
            float sc = fabsf(rl) / fabsf(jv[ic]);
            rmax = fmaxf(rmax, sc);
            lmax = fmaxf(lmax, fabsf(upd));
        }
    }
    d.lam = mk4(lv[0], lv[1], lv[2], lv[3]);
}
template <bool RES = true>
__device__ __forceinline__ void run_row_normal(NormalRow &d, SolveLDS &L, float4 *linvel,
                                               float4 *angvel, float &rmax, float &lmax) {
    float4 vA, wA, vB, wB;
    fetch_vel(d.ha, d.a, L, linvel, angvel, vA, wA);
    fetch_vel(d.hb, d.b, L, linvel, angvel, vB, wB);
    row_normal_pre<RES>(d, vA, wA, vB, wB, rmax, lmax);
    if (d.ha >= 0) { L.v[d.ha] = vA; L.w[d.ha] = wA; }
    else if (d.ha == SLOT_GLOBAL) { linvel[d.a] = vA; angvel[d.a] = wA; }
    if (d.hb >= 0) { L.v[d.hb] = vB; L.w[d.hb] = wB; }
    else if (d.hb == SLOT_GLOBAL) { linvel[d.b] = vB; angvel[d.b] = wB; }
}

// The sweep residual terms (largest scaled impulse change, largest impulse)
// go into per-sweep maxima: a max over the wave's lanes, then over the
// workgroup's waves in LDS, then one atomic per workgroup (on the float bits:
// the terms are >= 0, so the bits order like the values and the result
// equals per-lane atomics; one atomic per wave on two addresses had cost a
// spatial launch about 5 us).  They are needed after
// every sweep with a tolerance, otherwise only for the last normal and the
// last friction sweep, the residuals a step reports.
__host__ __device__ __forceinline__ bool sweep_residual_needed(const StepParams &sp, int sweep) {
    return sp.tol > 0.f || sweep == sp.normal_iters - 1 ||
           sweep == sp.normal_iters + sp.friction_iters - 1;
}
// Every thread of the workgroup calls it; gred (LDS, two words) is zero on
// entry and left zero.
__device__ __forceinline__ void add_sweep_residual(float *residual, int sweep, float rmax,
                                                   float lmax, unsigned int *gred) {
    unsigned r = __float_as_uint(rmax), l = __float_as_uint(lmax);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        r = max(r, (unsigned)__shfl_xor((int)r, o));
        l = max(l, (unsigned)__shfl_xor((int)l, o));
    }
    if ((threadIdx.x & 63) == 0 && (r != 0u || l != 0u)) {
        atomicMax(&gred[0], r);
        atomicMax(&gred[1], l);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned r0 = gred[0], l0 = gred[1];
        if (r0 != 0u || l0 != 0u) {
            atomicMax((unsigned int *)&residual[2 * sweep], r0);
            atomicMax((unsigned int *)&residual[2 * sweep + 1], l0);
        }
        gred[0] = 0u;
        gred[1] = 0u;
    }
    __syncthreads();
}

// Shared state of one tile's solve: the LDS velocity cache and the
// tile-wide residual reduction of the current sweep.
struct TileSolveLDS {
    SolveLDS c;
    unsigned int red[2];
    unsigned int gred[2];  // add_sweep_residual's workgroup maxima
};

// The sweeps of one tile: the optional warm-start pass, then normal sweeps,
// then friction sweeps, over the tile's packed rows (positions first..first+R
// in solve order).  REG: the tile has at most TBK rows and row r lives in the
// registers of thread r for the whole launch (rd, rb and its colour rc,
// loaded by the caller; stored once at the end);
// otherwise rows are read from the packed arrays on every pass.  Within a
// colour rows share no dynamic body, so the thread that runs a row does not
// change the result.  local_tol: a scene tile ends each phase as soon as its
// own residual is below sp.tol.
template <int TBK, bool REG, bool RES>
__device__ __forceinline__ void tile_sweeps(const StepParams &sp, int sweep0, int nsweeps_n,
                                            int nsweeps_f, int warm, bool local_tol, int first,
                                            int R, int nc, const int *coff, const PackedRows &pk,
                                            float4 *linvel, float4 *angvel, float *residual,
                                            TileSolveLDS &L, int used[2], float rfin[2],
                                            RowData &rd, RowBodies &rb, int rc) {
    const int tid = threadIdx.x;
    bool swept = false;
    for (int phase = warm ? -1 : 0; phase < 2; phase++) {
        const int nsw = phase < 0 ? 1 : (phase == 0 ? nsweeps_n : nsweeps_f);
        const int mode = phase < 0 ? 2 : phase;
        for (int it = 0; it < nsw; it++) {
            float rmax = 0.f, lmax = 0.f;
            if (local_tol) {
                if (tid == 0) { L.red[0] = 0u; L.red[1] = 0u; }
                __syncthreads();
            }
            for (int c = 0; c < nc; c++) {
                if (REG) {
                    if (rc == c)
                        run_row<RES>(mode, rd, rb, L.c, linvel, angvel, sp.friction, rmax, lmax);
                } else {
                    int c0 = coff[c], c1 = coff[c + 1];
                    if (c1 > R) c1 = R;
                    for (int r = c0 + tid; r < c1; r += TBK) {
                        RowData d;
                        RowBodies b;
                        load_packed(first + r, pk, d, b);
                        run_row<RES>(mode, d, b, L.c, linvel, angvel, sp.friction, rmax, lmax);
                        if (mode == 0) pk.lam[first + r] = d.lam;
                        else if (mode == 1) pk.fric[first + r] = d.fr;
                    }
                }
                __syncthreads();
            }
            if (phase < 0) continue;
            swept = true;
            const int sweep = sweep0 + (phase == 0 ? it : nsweeps_n + it);
            if (RES && sweep_residual_needed(sp, sweep)) add_sweep_residual(residual, sweep, rmax, lmax, L.gred);
            if (local_tol) {
                // rmax, lmax >= 0: their float bits order like the values
                if (rmax > 0.f) atomicMax(&L.red[0], __float_as_uint(rmax));
                if (lmax > 0.f) atomicMax(&L.red[1], __float_as_uint(lmax));
                __syncthreads();
                float rm = __uint_as_float(L.red[0]), lm = __uint_as_float(L.red[1]);
                float r = rm / (lm > 1.f ? lm : 1.f);
                used[phase] = it + 1;
                rfin[phase] = r;
                __syncthreads();  // every thread has read red before the next reset
                if (r < sp.tol) break;
            }
        }
    }
    if (REG && tid < R && swept) {
        pk.lam[first + tid] = rd.lam;
        pk.fric[first + tid] = rd.fr;
    }
}

// One normal sweep of a tile on precomputed rows (spatial tiles: a launch
// per sweep and parity phase).  REG: row tid is nr, colour rc.
template <int TBK, bool REG, bool RES>
__device__ __forceinline__ void tile_normal_sweep(const StepParams &sp, int sweep, int first,
                                                  int R, int nc, const int *coff,
                                                  const PackedRows &pk,
                                                  float4 *linvel, float4 *angvel,
                                                  float *residual, TileSolveLDS &L,
                                                  NormalRow &nr, int rc) {
    const int tid = threadIdx.x;
    float rmax = 0.f, lmax = 0.f;
    for (int c = 0; c < nc; c++) {
        if (REG) {
            if (rc == c) run_row_normal<RES>(nr, L.c, linvel, angvel, rmax, lmax);
        } else {
            int c0 = coff[c], c1 = coff[c + 1];
            if (c1 > R) c1 = R;
            for (int r = c0 + tid; r < c1; r += TBK) {
                NormalRow d;
                load_normal(first + r, pk, d);
                run_row_normal<RES>(d, L.c, linvel, angvel, rmax, lmax);
                pk.lam[first + r] = d.lam;
            }
        }
        __syncthreads();
    }
    if (RES && sweep_residual_needed(sp, sweep)) add_sweep_residual(residual, sweep, rmax, lmax, L.gred);
    if (REG && tid < R) pk.lam[first + tid] = nr.lam;
}

// The tile's dynamic bodies' velocities into LDS slots (slot s: body
// bodies[s]).  Thread tid keeps the body of slot tid (every slot when nb <=
// TBK) in a register for the store back, which then waits on no index load.
template <int TBK>
__device__ __forceinline__ int gather_velocities(const int *__restrict__ bodies, int nb,
                                                 const float4 *__restrict__ linvel,
                                                 const float4 *__restrict__ angvel, SolveLDS &L) {
    const int tid = threadIdx.x;
    const int b0 = tid < nb ? bodies[tid] : 0;
    if (tid < nb) {
        L.v[tid] = linvel[b0];
        L.w[tid] = angvel[b0];
    }
    for (int s = tid + TBK; s < nb; s += TBK) {
        const int b = bodies[s];
        L.v[s] = linvel[b];
        L.w[s] = angvel[b];
    }
    return b0;
}
template <int TBK>
__device__ __forceinline__ void store_velocities(const int *__restrict__ bodies, int nb, int b0,
                                                 float4 *__restrict__ linvel,
                                                 float4 *__restrict__ angvel, const SolveLDS &L) {
    const int tid = threadIdx.x;
    if (tid < nb) {
        linvel[b0] = L.v[tid];
        angvel[b0] = L.w[tid];
    }
    for (int s = tid + TBK; s < nb; s += TBK) {
        const int b = bodies[s];
        linvel[b] = L.v[s];
        angvel[b] = L.w[s];
    }
}

// One workgroup per tile (spatial: the tiles of one parity phase; scene: one
// scene); the tile's dynamic body velocities live in LDS for the launch.
// The solve of one tile (rows first..first+R of the packed arrays): its
// dynamic bodies' velocities gathered into LDS, the sweeps, the velocities
// stored back.  NORMAL: one normal sweep on the precomputed rows.
// RES: the launch's sweeps compute their residual terms (a launch of one
// sweep whose residual is not reported runs without them)
template <int TBK, bool NORMAL, bool RES>
__device__ __forceinline__ void solve_tile(
    const StepParams &sp, int t, int sweep0, int nsweeps_n, int nsweeps_f, int warm,
    const int4 *__restrict__ tile_hdr, const int *__restrict__ tile_color_off,
    const int *__restrict__ tile_bodies, const PackedRows &pk, float4 *__restrict__ linvel,
    float4 *__restrict__ angvel, float *__restrict__ residual, int *counters,
    int4 *__restrict__ tile_stats, TileSolveLDS &L) {
    const bool local_tol = sp.tile_mode == 1 && sp.tol > 0.f;
    const int tid = threadIdx.x;
    // (first, rows, colours, cached bodies) of the tile, written by k_tile_pack
    const int4 hdr = tile_hdr[t];
    if (sp.tol > 0.f && counters[RB_CNT_DONE] != 0) return;  // converged in an earlier sweep
    const int first = hdr.x, R = hdr.y, nc = hdr.z, nb = hdr.w;
    if (R <= 0) {
        if (local_tol && tid == 0) tile_stats[t] = make_int4(-1, -1, 0, 0);
        return;
    }
    const int *coff = tile_color_off + t * (RB_MAX_LOCAL_COLORS + 1);
    const int *bodies = tile_bodies + 2 * (size_t)first;
    // (the barrier after the velocity gather orders this before any use)
    if (tid == 0) {
        L.gred[0] = 0u;
        L.gred[1] = 0u;
    }
    if constexpr (NORMAL) {
        // one normal sweep (nsweeps_n == 1) on the precomputed rows
        NormalRow nr;
        int rc = -1;
        if (R <= TBK && tid < R) {
            load_normal(first + tid, pk, nr);
            for (int c = 0; c < nc; c++)
                if (coff[c] <= tid && tid < coff[c + 1]) rc = c;
        }
        const int b0 = gather_velocities<TBK>(bodies, nb, linvel, angvel, L.c);
        __syncthreads();
        if (R <= TBK)
            tile_normal_sweep<TBK, true, RES>(sp, sweep0, first, R, nc, coff, pk, linvel, angvel,
                                              residual, L, nr, rc);
        else
            tile_normal_sweep<TBK, false, RES>(sp, sweep0, first, R, nc, coff, pk, linvel,
                                               angvel, residual, L, nr, rc);
        __syncthreads();
        store_velocities<TBK>(bodies, nb, b0, linvel, angvel, L.c);
        return;
    }
    // a tile of at most TBK rows keeps row tid in thread tid's registers;
    // its loads go out together with the velocity gather below
    RowData rd;
    RowBodies rb;
    int rc = -1;
    if (R <= TBK && tid < R) {
        load_packed(first + tid, pk, rd, rb);
        for (int c = 0; c < nc; c++)
            if (coff[c] <= tid && tid < coff[c + 1]) rc = c;
    }
    const int b0 = gather_velocities<TBK>(bodies, nb, linvel, angvel, L.c);
    __syncthreads();
    int used[2] = {0, 0};
    float rfin[2] = {0.f, 0.f};
    if (R <= TBK)
        tile_sweeps<TBK, true, RES>(sp, sweep0, nsweeps_n, nsweeps_f, warm, local_tol, first, R,
                                    nc, coff, pk, linvel, angvel, residual, L, used, rfin, rd, rb,
                                    rc);
    else
        tile_sweeps<TBK, false, RES>(sp, sweep0, nsweeps_n, nsweeps_f, warm, local_tol, first, R,
                                     nc, coff, pk, linvel, angvel, residual, L, used, rfin, rd,
                                     rb, rc);
    if (local_tol && tid == 0)
        tile_stats[t] = make_int4(used[0], used[1], (int)__float_as_uint(rfin[0]),
                                  (int)__float_as_uint(rfin[1]));
    __syncthreads();
    store_velocities<TBK>(bodies, nb, b0, linvel, angvel, L.c);
}


template <int TBK, bool NORMAL = false, bool RES = true>
__global__ __launch_bounds__(TBK) void k_tile_solve(
    StepParams sp, int phase, int sweep0, int nsweeps_n, int nsweeps_f, int warm,
    const int4 *__restrict__ tile_hdr, const int *__restrict__ tile_color_off,
    const int *__restrict__ tile_bodies, PackedRows pk,
    float4 *__restrict__ linvel, float4 *__restrict__ angvel, float *__restrict__ residual,
    int *counters, int4 *__restrict__ tile_stats, const uint2 *__restrict__ tile_order) {
    __shared__ TileSolveLDS L;
    int t;
    if (sp.tile_mode == 1) {
        // largest scenes first: the workgroups dispatched last are the short ones
        t = (int)tile_order[blockIdx.x].y;
    } else {
        // tiles of this phase: (tx & 1, tz & 1) == (phase & 1, phase >> 1)
        int px = phase & 1, pz = phase >> 1;
        int nx = (sp.tiles_x - px + 1) / 2;
        int bx = blockIdx.x % nx, bz = blockIdx.x / nx;
        int tx = px + 2 * bx, tz = pz + 2 * bz;
        if (tx >= sp.tiles_x || tz >= sp.tiles_z) return;
        t = tz * sp.tiles_x + tx;
    }
    solve_tile<TBK, NORMAL, RES>(sp, t, sweep0, nsweeps_n, nsweeps_f, warm, tile_hdr, tile_color_off,
                            tile_bodies, pk, linvel, angvel, residual, counters, tile_stats, L);
}

// Scene tiles with a tolerance: sweeps used and final residuals over the
// tiles that had rows (maximum, and sums for the mean).
__global__ __launch_bounds__(256) void k_tile_summary(int n_tiles, const int4 *tile_stats,
                                                      int *counters) {
    __shared__ int smax[2], ssum[2], scnt;
    __shared__ unsigned int sres[2];
    if (threadIdx.x == 0) {
        smax[0] = smax[1] = 0; ssum[0] = ssum[1] = 0; scnt = 0; sres[0] = sres[1] = 0u;
    }
    __syncthreads();
    int mx0 = 0, mx1 = 0, sm0 = 0, sm1 = 0, cnt = 0;
    unsigned int r0 = 0u, r1 = 0u;
    for (int t = threadIdx.x; t < n_tiles; t += blockDim.x) {
        int4 st = tile_stats[t];
        if (st.x < 0) continue;
        mx0 = max(mx0, st.x); mx1 = max(mx1, st.y);
        sm0 += st.x; sm1 += st.y; cnt++;
        r0 = max(r0, (unsigned int)st.z); r1 = max(r1, (unsigned int)st.w);
    }
    atomicMax(&smax[0], mx0); atomicMax(&smax[1], mx1);
    atomicAdd(&ssum[0], sm0); atomicAdd(&ssum[1], sm1); atomicAdd(&scnt, cnt);
    atomicMax(&sres[0], r0); atomicMax(&sres[1], r1);
    __syncthreads();
    if (threadIdx.x == 0) {
        counters[RB_CNT_ITERS_N] = smax[0];
        counters[RB_CNT_ITERS_F] = smax[1];
        counters[RB_CNT_SUM_ITERS_N] = ssum[0];
        counters[RB_CNT_SUM_ITERS_F] = ssum[1];
        counters[RB_CNT_TILES_SOLVED] = scnt;
        counters[RB_CNT_RES_N] = (int)sres[0];
        counters[RB_CNT_RES_F] = (int)sres[1];
    }
}

// After a sweep (spatial tiles): decide convergence for the remaining sweeps.
__global__ void k_sweep_done(float *residual, int sweep, int is_friction, float tol,
                             int *counters) {
    if (counters[RB_CNT_DONE] != 0) return;
    float rmax = residual[2 * sweep], lmax = residual[2 * sweep + 1];
    float r = rmax / (lmax > 1.f ? lmax : 1.f);
    int slot = is_friction ? RB_CNT_ITERS_F : RB_CNT_ITERS_N;
    counters[slot] += 1;
    ((float *)counters)[is_friction ? RB_CNT_RES_F : RB_CNT_RES_N] = r;
    if (tol > 0.f && r < tol) counters[RB_CNT_DONE] = 1;
}
__global__ void k_clear_done(int *counters) { counters[RB_CNT_DONE] = 0; }

// fixed sweep counts: iterations used and the last sweep's residual per phase
__global__ void k_sweep_summary(const float *residual, int nn, int nf, int *counters) {
    counters[RB_CNT_ITERS_N] = nn;
    counters[RB_CNT_ITERS_F] = nf;
    float *f = (float *)counters;
    if (nn > 0) {
        float l = residual[2 * (nn - 1) + 1];
        f[RB_CNT_RES_N] = residual[2 * (nn - 1)] / (l > 1.f ? l : 1.f);
    }
    if (nf > 0) {
        int sw = nn + nf - 1;
        float l = residual[2 * sw + 1];
        f[RB_CNT_RES_F] = residual[2 * sw] / (l > 1.f ? l : 1.f);
    }
}

// ---- integration ----
__global__ __launch_bounds__(256) void k_integrate(int n, float4 *__restrict__ pos,
                                                   float4 *__restrict__ quat,
                                                   const float4 *__restrict__ linvel,
                                                   float4 *__restrict__ angvel,
                                                   const float *__restrict__ inv_mass, float dt,
                                                   float damping, float max_ang) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 w = angvel[i];
    w.x *= damping; w.y *= damping; w.z *= damping;
    angvel[i] = w;
    float l2 = (w.x * w.x + w.y * w.y) + w.z * w.z;
    float ang = sqrtf(l2);
    if (ang * dt > max_ang) ang = max_ang / dt;
    float s;
    if (ang < 0.001f)
        s = 0.5f * dt - (((dt * dt * dt) * 0.020833333333f) * ang) * ang;
    else
        s = sinf((0.5f * ang) * dt) / ang;
    float ax = w.x * s, ay = w.y * s, az = w.z * s;
    float dw = cosf((ang * dt) * 0.5f);
    float4 q = quat[i];
    float rx = ((dw * q.x + ax * q.w) + ay * q.z) - az * q.y;
    float ry = ((dw * q.y + ay * q.w) + az * q.x) - ax * q.z;
    float rz = ((dw * q.z + az * q.w) + ax * q.y) - ay * q.x;
    float rw = ((dw * q.w - ax * q.x) - ay * q.y) - az * q.z;
    float len = sqrtf(((rx * rx + ry * ry) + rz * rz) + rw * rw);
    if (len > 0.f) {
        float inv = 1.f / len;
        q = mk4(rx * inv, ry * inv, rz * inv, rw * inv);
    } else {
        q = mk4(0.f, 0.f, 0.f, 1.f);
    }
    quat[i] = q;
    float4 v = linvel[i];
    float4 p = pos[i];
    p.x = p.x + v.x * dt;
    p.y = p.y + v.y * dt;
    p.z = p.z + v.z * dt;
    pos[i] = p;
    (void)inv_mass;
}

}  // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_constraints(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    int grid = (p.max_pairs + 255) / 256;
    RB_LAUNCH(k_constraints, dim3(grid), dim3(256), 0, s, p, d.counters + RB_CNT_PAIRS,
                       d.has_contact, d.contact_off, d.contact_slots, d.contacts, d.pos, d.linvel,
                       d.angvel, d.inv_mass, d.inv_inertia_world, d.scene, d.aabb_min, d.aabb_max,
                       d.cs, d.tile_kv, d.counters, d.prev_keys, d.prev_imp, d.prev_pts, d.prev_count,
                       d.features, d.solve_pos, d.pairs);
    return rb_launch_status();
}

hipError_t launch_tile_bounds(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    RB_LAUNCH(k_tile_bounds2, dim3((p.max_contacts + 255) / 256), dim3(256), 0, s,
                       d.tile_kv, d.counters, d.tile_first, d.tile_cnt);
    return rb_launch_status();
}

hipError_t launch_tiles(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    RB_LAUNCH(k_tile_color, dim3(p.n_tiles), dim3(TB), 0, s, d.tile_kv, d.tile_first,
                       d.tile_cnt, d.cs.bodies, d.inv_mass, d.solve_order, d.row_color,
                       d.tile_color_off, d.tile_ncolors, d.counters);
    RB_LAUNCH(k_tile_pack, dim3(p.n_tiles), dim3(TB), 0, s, d.tile_first, d.tile_cnt,
                       d.solve_order, d.cs, d.pos, d.inv_mass, d.inv_inertia_world, d.pk,
                       d.solve_pos, d.tile_bodies, d.tile_nb, d.tile_ncolors, d.tile_hdr,
                       d.counters, p.tile_mode == 0 ? 1 : 0);
    return rb_launch_status();
}

namespace {
__global__ void k_tile_order_keys(int n_tiles, const int4 *__restrict__ tile_hdr,
                                  uint2 *__restrict__ kv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const int R = tile_hdr[t].y;
    const uint32_t r = (uint32_t)(R < 0 ? 0 : (R > 0xffff ? 0xffff : R));
    kv[t] = make_uint2(0xffffu - r, (uint32_t)t);
}
}  // namespace

hipError_t launch_tile_order(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                             const SortWork &sw) {
    if (p.tile_mode != 1) return hipSuccess;
    RB_LAUNCH(k_tile_order_keys, dim3((p.n_tiles + 255) / 256), dim3(256), 0, s, p.n_tiles,
              d.tile_hdr, d.tile_order);
    hipError_t e = rb_launch_status();
    if (e != hipSuccess) return e;
    // (the order is a schedule only: every scene is solved by its own
    // workgroup, so results do not depend on it)
    return launch_radix_sort(d.tile_order, p.n_tiles, 16, 8, sw, s);
}

hipError_t launch_solve(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                        hipEvent_t *kev, int *n_launches, hipEvent_t *span) {
    const int S = p.normal_iters + p.friction_iters;
    const int warm = p.warm_start != MI_RB_WARM_ZERO;
    int nl = 0;
    if (span) hipEventRecord(span[0], s);
    auto mark = [&](int end) {
        if (kev) hipEventRecord(kev[2 * nl + end], s);
        if (end) nl++;
    };
    // (the step's first kernel has cleared the converged flag and the residuals)
    // scene tiles: one workgroup per scene, 256 or 512 threads
#define RB_SOLVE_ARGS                                                                         \
    d.tile_hdr, d.tile_color_off, d.tile_bodies, d.pk, d.linvel, d.angvel, d.residual,        \
        d.counters, d.tile_stats, d.tile_order
    if (p.tile_mode == 1) {
        // independent scenes: every sweep of a scene in one workgroup, one
        // launch; with a tolerance each scene stops on its own residual
        mark(0);
        if (p.solve_block == 256)
            RB_LAUNCH(k_tile_solve<256>, dim3(p.n_tiles), dim3(256), 0, s, p, 0, 0,
                               p.normal_iters, p.friction_iters, warm, RB_SOLVE_ARGS);
        else
            RB_LAUNCH(k_tile_solve<TB_SCENE>, dim3(p.n_tiles), dim3(TB_SCENE), 0, s, p,
                               0, 0, p.normal_iters, p.friction_iters, warm, RB_SOLVE_ARGS);
        mark(1);
        if (p.tol > 0.f)
            RB_LAUNCH(k_tile_summary, dim3(1), dim3(256), 0, s, p.n_tiles, d.tile_stats,
                               d.counters);
        else
            RB_LAUNCH(k_sweep_summary, dim3(1), dim3(1), 0, s, d.residual, p.normal_iters,
                               p.friction_iters, d.counters);
    } else {
        for (int sw = warm ? -1 : 0; sw < S; sw++) {
            int fr = sw >= p.normal_iters;
            for (int ph = 0; ph < 4; ph++) {
                int nx = (p.tiles_x - (ph & 1) + 1) / 2, nz = (p.tiles_z - (ph >> 1) + 1) / 2;
                if (nx <= 0 || nz <= 0) continue;
                mark(0);
                if (sw < 0)
                    RB_LAUNCH(k_tile_solve<TB>, dim3(nx * nz), dim3(TB), 0, s, p, ph, 0, 0, 0,
                                       1, RB_SOLVE_ARGS);
                else if (!fr && sweep_residual_needed(p, sw))
                    RB_LAUNCH((k_tile_solve<TB, true>), dim3(nx * nz), dim3(TB), 0, s, p,
                                       ph, sw, 1, 0, 0, RB_SOLVE_ARGS);
                else if (!fr)
                    RB_LAUNCH((k_tile_solve<TB, true, false>), dim3(nx * nz), dim3(TB), 0, s, p,
                              ph, sw, 1, 0, 0, RB_SOLVE_ARGS);
                else if (sweep_residual_needed(p, sw))
                    RB_LAUNCH(k_tile_solve<TB>, dim3(nx * nz), dim3(TB), 0, s, p, ph, sw,
                                       0, 1, 0, RB_SOLVE_ARGS);
                else
                    RB_LAUNCH((k_tile_solve<TB, false, false>), dim3(nx * nz), dim3(TB), 0, s,
                              p, ph, sw, 0, 1, 0, RB_SOLVE_ARGS);
                mark(1);
            }
            if (sw >= 0 && p.tol > 0.f) {
                RB_LAUNCH(k_sweep_done, dim3(1), dim3(1), 0, s, d.residual, sw, fr, p.tol,
                                   d.counters);
                if (sw == p.normal_iters - 1)
                    RB_LAUNCH(k_clear_done, dim3(1), dim3(1), 0, s, d.counters);
            }
        }
        if (p.tol <= 0.f)
            RB_LAUNCH(k_sweep_summary, dim3(1), dim3(1), 0, s, d.residual, p.normal_iters,
                               p.friction_iters, d.counters);
    }
#undef RB_SOLVE_ARGS
    if (span) hipEventRecord(span[1], s);
    if (n_launches) *n_launches = nl;
    const bool staged = p.n_proxies > 0;
    RB_LAUNCH(k_save_prev, dim3((p.max_contacts + 255) / 256), dim3(256), 0, s,
                       d.counters, d.cs, d.solve_pos, d.pk.lam, d.pk.fric,
                       staged ? d.stage_keys : d.prev_keys, staged ? d.stage_imp : d.prev_imp,
                       staged ? d.stage_pts : d.prev_pts, d.prev_count, d.pairs);
    return rb_launch_status();
}

namespace {
// (key part, row) of every cache entry: y first, then x of the order so far
__global__ void k_cache_key_y(const int *counters, const int2 *keys, uint2 *kv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < counters[RB_CNT_CONTACTS]) kv[k] = make_uint2((uint32_t)keys[k].y, (uint32_t)k);
}
__global__ void k_cache_key_x(const int *counters, const int2 *keys, uint2 *kv) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < counters[RB_CNT_CONTACTS]) kv[k].x = (uint32_t)keys[kv[k].y].x;
}
__global__ void k_cache_gather(const int *counters, const uint2 *kv, const int2 *keys,
                               const float4 *imp, const float4 *pts, int2 *out_keys,
                               float4 *out_imp, float4 *out_pts) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= counters[RB_CNT_CONTACTS]) return;
    const int src = (int)kv[k].y;
    out_keys[k] = keys[src];
    out_imp[k] = imp[src];
    for (int i = 0; i < 4; i++) out_pts[4 * (size_t)k + i] = pts[4 * (size_t)src + i];
}
}  // namespace

hipError_t launch_cache_order(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                              const SortWork &sw) {
    if (p.n_proxies <= 0) return hipSuccess;
    // pair entries name bodies and compound/triangle proxies below n + proxies
    const uint32_t top = (uint32_t)(p.n_bodies + p.n_proxies);
    int bits = 1;
    while (bits < 32 && (top >> bits) != 0) bits++;
    const int grid = (p.max_contacts + 255) / 256;
    int *nc = d.counters + RB_CNT_CONTACTS;
    RB_LAUNCH(k_cache_key_y, dim3(grid), dim3(256), 0, s, d.counters, d.stage_keys,
                       d.tile_kv);
    hipError_t e = launch_radix_sort(d.tile_kv, p.max_contacts, bits, 8, sw, s, nc);
    if (e != hipSuccess) return e;
    RB_LAUNCH(k_cache_key_x, dim3(grid), dim3(256), 0, s, d.counters, d.stage_keys,
                       d.tile_kv);
    e = launch_radix_sort(d.tile_kv, p.max_contacts, bits, 8, sw, s, nc);
    if (e != hipSuccess) return e;
    RB_LAUNCH(k_cache_gather, dim3(grid), dim3(256), 0, s, d.counters, d.tile_kv,
                       d.stage_keys, d.stage_imp, d.stage_pts, d.prev_keys, d.prev_imp,
                       d.prev_pts);
    return rb_launch_status();
}

// Body transforms into a renderer's vertex buffer (copyTransformsToVBOKernel,
// opencl/broadphase_benchmark/broadphaseKernel.cl:236-247): float4 (x, y, z,
// 1) of body i at base + i, its quaternion at base + n + i.
__global__ __launch_bounds__(256) void k_copy_transforms(int n, const float4 *__restrict__ pos,
                                                         const float4 *__restrict__ quat,
                                                         float4 *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    dst[i] = make_float4(p.x, p.y, p.z, 1.f);
    dst[n + i] = quat[i];
}

hipError_t launch_copy_transforms(const DeviceWorld &d, int n, float4 *dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    RB_LAUNCH(k_copy_transforms, dim3((n + 255) / 256), dim3(256), 0, s, n, d.pos, d.quat,
                       dst);
    return rb_launch_status();
}

hipError_t launch_integrate(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    if (p.n_bodies == 0) return hipSuccess;
    RB_LAUNCH(k_integrate, dim3((p.n_bodies + 255) / 256), dim3(256), 0, s, p.n_bodies,
                       d.pos, d.quat, d.linvel, d.angvel, d.inv_mass, p.dt, p.damping, p.max_ang);
    return rb_launch_status();
}
