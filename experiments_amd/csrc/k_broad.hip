// k_broad.hip -- body setup and uniform-grid broadphase for gfx950.
//
//  k_setup      one pass over the body SoA: world AABB (initializeGpuAabbsFull,
//               opencl/broadphase_benchmark/broadphaseKernel.cl:293-331), cell key
//               (kCalcHashAABB, opencl/3dGridBroadphase/Shared/bt3dGridBroadphaseOCL.cl:26-47),
//               gravity and world inverse inertia (setupBodiesKernel,
//               broadphaseKernel.cl:163-212).  The reference runs these as three
//               launches with host syncs in between; here they share one read of
//               the body state.
//  k_cell_start first sorted index of every occupied cell (kFindCellStart,
//               bt3dGridBroadphaseOCL.cl:61-86).  The table is never cleared: a
//               lookup is valid only if it points at the first entry of that key
//               in the current sorted array, which a stale entry cannot.
//  k_find_pairs 27-cell stencil per small body (kFindOverlappingPairs /
//               findPairsInCell, bt3dGridBroadphaseOCL.cl:100-213), keeping the
//               overlapping bodies of lower index in a sorted per-body list in LDS.
//  k_large_pairs overflow list (kFindPairsLarge, :216-264): one workgroup per
//               large body tests its scene, in index order.
//  k_emit_pairs  int2 (x > y) pairs at scanned offsets (countOverlappingpairs /
//               squeezePairCaches, opencl/broadphase_benchmark/computeAabbKernelOCL.cl:40-111),
//               globally sorted by (x, y).
#include <hip/hip_runtime.h>

#include "dmath.h"
#include "rb_internal.h"

using namespace rb;

namespace {

constexpr uint32_t KEY_NONE = 0xffffffffu;

__device__ __forceinline__ void grid_pos(float4 mn, float4 mx, float inv_cell, const int *dims,
                                         int gp[3]) {
    float cx = (mn.x + mx.x) * 0.5f, cy = (mn.y + mx.y) * 0.5f, cz = (mn.z + mx.z) * 0.5f;
    gp[0] = ((int)floorf(cx * inv_cell)) & (dims[0] - 1);
    gp[1] = ((int)floorf(cy * inv_cell)) & (dims[1] - 1);
    gp[2] = ((int)floorf(cz * inv_cell)) & (dims[2] - 1);
}
__device__ __forceinline__ uint32_t pos_hash(int x, int y, int z, const int *dims) {
    x &= dims[0] - 1;
    y &= dims[1] - 1;
    z &= dims[2] - 1;
    return (uint32_t)(z * dims[1] * dims[0] + y * dims[0] + x);
}

// Stencil cells that can hold a member overlapping the AABB [mn, mx]: a
// member of cell k (unwrapped, along one axis) has its centre in
// [k c, (k+1) c) and a half-extent of at most c/2 (the grid's cell covers
// every member), so its AABB lies in [k c - c/2, (k+1) c + c/2].  Bit q of the
// result (q = 9 (dz+1) + 3 (dy+1) + (dx+1)) is set unless that cell is
// provably out of reach; wrapped aliases of a cell only hold members farther
// away, which the AABB test rejects anyway.
__device__ __forceinline__ uint32_t reachable_cells(float4 mn, float4 mx, float inv_cell) {
    const float c = 1.0f / inv_cell, h = 0.5f * c, eps = 1e-4f * c;
    const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    uint32_t ok[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float ctr = (lo[k] + hi[k]) * 0.5f;
        const float g = floorf(ctr * inv_cell);
        ok[k] = 0;
#pragma unroll
        for (int d = -1; d <= 1; d++) {
            const float c0 = (g + (float)d) * c - h - eps, c1 = (g + (float)(d + 1)) * c + h + eps;
            if (hi[k] >= c0 && lo[k] <= c1) ok[k] |= 1u << (d + 1);
        }
    }
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 27; q++)
        if (((ok[0] >> (q % 3)) & 1u) && ((ok[1] >> ((q / 3) % 3)) & 1u) && ((ok[2] >> (q / 9)) & 1u))
            m |= 1u << q;
    return m;
}

struct GridP {
    float inv_cell;
    int dims[3];
    int cell_bits;
};

__global__ __launch_bounds__(256) void k_setup(
    int n, const float4 *__restrict__ pos, const float4 *__restrict__ quat,
    float4 *__restrict__ linvel, float4 *__restrict__ angvel, const float *__restrict__ inv_mass,
    const int *__restrict__ shape, const int *__restrict__ scene, const int *__restrict__ is_large,
    const HullInfo *__restrict__ hulls, const float4 *__restrict__ inv_inertia_local,
    float4 *__restrict__ inv_inertia_world, float4 *__restrict__ aabb_min,
    float4 *__restrict__ aabb_max, uint2 *__restrict__ kv, GridP g, float gx, float gy, float gz,
    float dt, uint32_t *__restrict__ occ, int occ_words) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the cell occupancy bits of the previous step
    for (int w = i; w < occ_words; w += gridDim.x * blockDim.x) occ[w] = 0u;
    if (i >= n) return;
    float4 p = pos[i], q = quat[i];
    const HullInfo *h = hulls + shape[i];
    float4 hmn = h->aabb_min, hmx = h->aabb_max;
    float4 he = mk4((hmx.x - hmn.x) * 0.5f, (hmx.y - hmn.y) * 0.5f, (hmx.z - hmn.z) * 0.5f, 0.f);
    float4 lc = mk4((hmx.x + hmn.x) * 0.5f, (hmx.y + hmn.y) * 0.5f, (hmx.z + hmn.z) * 0.5f, 0.f);
    float4 wc = xform(lc, p, q);
    float4 m[3];
    abs_rot(q, m);
    float4 ext = mk4(dot3(m[0], he), dot3(m[1], he), dot3(m[2], he), 0.f);
    float4 mn = mk4(wc.x - ext.x, wc.y - ext.y, wc.z - ext.z, 0.f);
    float4 mx = mk4(wc.x + ext.x, wc.y + ext.y, wc.z + ext.z, 0.f);
    aabb_min[i] = mn;
    aabb_max[i] = mx;
    uint32_t key = KEY_NONE;
    if (!is_large[i]) {
        int gp[3];
        grid_pos(mn, mx, g.inv_cell, g.dims, gp);
        key = pos_hash(gp[0], gp[1], gp[2], g.dims) | ((uint32_t)scene[i] << g.cell_bits);
    }
    kv[i] = make_uint2(key, (uint32_t)i);
    float im = inv_mass[i];
    float4 *I = inv_inertia_world + 3 * i;
    if (im != 0.f) {
        float4 v = linvel[i];
        v.x = v.x + gx * dt;
        v.y = v.y + gy * dt;
        v.z = v.z + gz * dt;
        linvel[i] = v;
        float4 d = inv_inertia_local[i];
        float4 mt[3] = {mk4(m[0].x, m[1].x, m[2].x, 0.f), mk4(m[0].y, m[1].y, m[2].y, 0.f),
                        mk4(m[0].z, m[1].z, m[2].z, 0.f)};
        float4 loc[3] = {mk4(d.x, 0.f, 0.f, 0.f), mk4(0.f, d.y, 0.f, 0.f), mk4(0.f, 0.f, d.z, 0.f)};
        float4 tmp[3], out[3];
        // mtMul(a, b): ans[i][j] = dot3(a.row[i], b.col[j]) (broadphaseKernel.cl:130-148)
        for (int r = 0; r < 3; r++) {
            float4 c0 = mk4(loc[0].x, loc[1].x, loc[2].x, 0.f);
            float4 c1 = mk4(loc[0].y, loc[1].y, loc[2].y, 0.f);
            float4 c2 = mk4(loc[0].z, loc[1].z, loc[2].z, 0.f);
            tmp[r] = mk4(dot3(m[r], c0), dot3(m[r], c1), dot3(m[r], c2), 0.f);
        }
        for (int r = 0; r < 3; r++) {
            float4 c0 = mk4(mt[0].x, mt[1].x, mt[2].x, 0.f);
            float4 c1 = mk4(mt[0].y, mt[1].y, mt[2].y, 0.f);
            float4 c2 = mk4(mt[0].z, mt[1].z, mt[2].z, 0.f);
            out[r] = mk4(dot3(tmp[r], c0), dot3(tmp[r], c1), dot3(tmp[r], c2), 0.f);
        }
        I[0] = out[0];
        I[1] = out[1];
        I[2] = out[2];
    } else {
        float4 z = mk4(0.f, 0.f, 0.f, 0.f);
        linvel[i] = z;
        angvel[i] = z;
        I[0] = z;
        I[1] = z;
        I[2] = z;
    }
}

// First sorted index and occupancy bit of every occupied cell, and the AABBs
// in sorted order (the pair search then reads a cell member's key and box at
// one index).  occ[] was cleared by k_setup.
__global__ void k_cell_start(const uint2 *__restrict__ kv, int n,
                             const float4 *__restrict__ aabb_min,
                             const float4 *__restrict__ aabb_max, uint32_t *__restrict__ cell_start,
                             uint32_t *__restrict__ occ, float4 *__restrict__ smn,
                             float4 *__restrict__ smx) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint2 e = kv[s];
    if (e.x == KEY_NONE) return;
    smn[s] = aabb_min[e.y];
    smx[s] = aabb_max[e.y];
    if (s == 0 || kv[s - 1].x != e.x) {
        cell_start[e.x] = (uint32_t)s;
        atomicOr(&occ[e.x >> 5], 1u << (e.x & 31u));
    }
}

constexpr int FP_BLOCK = 256;

// LANES lanes per sorted entry; lane q walks the stencil cells
// [q*C, q*C + C), C = ceil(27 / LANES), cells numbered z, then y, then x
// from outermost to innermost, so a body's cells are searched in parallel
// (32 lanes for few dense bodies, 1 for many).  The overlaps are numbered in
// the order a serial search visits them (cells in stencil order, members in
// sorted order, then the scene's large bodies in index order): a count pass,
// a prefix over the lanes, and a store pass that keeps the first K in LDS.
// The kept indices are distinct (a body sits in one cell), so the list is
// ordered by rank.
template <int LANES>
__global__ __launch_bounds__(FP_BLOCK) void k_find_pairs(
    int n, const uint2 *__restrict__ kv, const uint32_t *__restrict__ cell_start,
    const uint32_t *__restrict__ occ, const float4 *__restrict__ smn,
    const float4 *__restrict__ smx, const float4 *__restrict__ aabb_min,
    const float4 *__restrict__ aabb_max, const int *__restrict__ large_list,
    const int *__restrict__ large_first,
    const int *__restrict__ large_num, GridP g, int K, int *__restrict__ body_slots,
    uint32_t *__restrict__ pair_count, int *overflow) {
    constexpr int ENTRIES = FP_BLOCK / LANES;
    constexpr int C = (27 + LANES - 1) / LANES;
    extern __shared__ int s_lists[];
    const int lane = threadIdx.x & (LANES - 1);
    const int slot = threadIdx.x / LANES;
    const int seg = slot & (64 / LANES - 1);  // which LANES-lane segment of the wave
    const int t = blockIdx.x * ENTRIES + slot;
    // every exit below is uniform over the entry's lanes
    if (t >= n) return;
    const uint2 e = kv[t];
    if (e.x == KEY_NONE) return;  // large body: k_large_count / k_large_write
    const int i = (int)e.y;
    int *dst = body_slots + (size_t)i * K;
    // the list: in LDS for the lanes' rank sort (odd stride: the lists start
    // in distinct banks); one lane keeps it in its own slots of body_slots,
    // so its workgroups need no LDS (260 B per lane limited them to two per
    // CU) and sort it there
    int *lst = LANES == 1 ? dst : s_lists + slot * (K | 1);
    const int sc = (int)(e.x >> g.cell_bits);
    const float4 mn0 = smn[t], mx0 = smx[t];
    int gp[3];
    grid_pos(mn0, mx0, g.inv_cell, g.dims, gp);
    const uint32_t sk = (uint32_t)sc << g.cell_bits;
    uint32_t c = 0;
    if constexpr (LANES == 1) {
        // one lane: the occupancy bits of the 27 cells fetched together,
        // then the occupied cells walked in order
        uint32_t bits = 0;
        const uint32_t reach = reachable_cells(mn0, mx0, g.inv_cell);
#pragma unroll
        for (int q = 0; q < 27; q++) {
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
            const uint32_t nkq = pos_hash(gp[0] + dx, gp[1] + dy, gp[2] + dz, g.dims) | sk;
            bits |= ((occ[nkq >> 5] >> (nkq & 31u)) & 1u) << q;
        }
        bits &= reach;
        while (bits) {
            const int q = __builtin_ctz(bits);
            bits &= bits - 1u;
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
            const uint32_t nkq = pos_hash(gp[0] + dx, gp[1] + dy, gp[2] + dz, g.dims) | sk;
            for (uint32_t u = cell_start[nkq]; u < (uint32_t)n; u++) {
                const uint2 f = kv[u];
                if (f.x != nkq) break;
                if ((int)f.y < i && aabb_overlap(mn0, mx0, smn[u], smx[u])) {
                    if (c < (uint32_t)K) lst[c] = (int)f.y;
                    c++;
                }
            }
        }
    } else {
        uint32_t nk[C], u0[C];
        const uint32_t reach = reachable_cells(mn0, mx0, g.inv_cell);
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int q = lane * C + k;
            nk[k] = 0;
            u0[k] = (uint32_t)n;
            if (q < 27 && ((reach >> q) & 1u)) {
                const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
                nk[k] = pos_hash(gp[0] + dx, gp[1] + dy, gp[2] + dz, g.dims) | sk;
                if ((occ[nk[k] >> 5] >> (nk[k] & 31u)) & 1u) u0[k] = cell_start[nk[k]];
            }
        }
        // one walk: a lane's first HITS overlaps stay in registers, so the
        // members are walked again only by a lane that found more
        constexpr int HITS = 4;
        int hit[HITS] = {0, 0, 0, 0};
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < C; k++)
            for (uint32_t u = u0[k]; u < (uint32_t)n; u++) {
                const uint2 f = kv[u];
                if (f.x != nk[k]) break;
                if ((int)f.y < i && aabb_overlap(mn0, mx0, smn[u], smx[u])) {
#pragma unroll
                    for (int h = 0; h < HITS; h++)
                        if (cnt == (uint32_t)h) hit[h] = (int)f.y;
                    cnt++;
                }
            }
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < LANES; o <<= 1) {
            uint32_t v = __shfl_up(incl, o, LANES);
            if (lane >= o) incl += v;
        }
        c = __shfl(incl, LANES - 1, LANES);
        uint32_t v = incl - cnt;
        if (cnt <= (uint32_t)HITS) {
#pragma unroll
            for (int h = 0; h < HITS; h++)
                if ((uint32_t)h < cnt && v + h < (uint32_t)K) lst[v + h] = hit[h];
        } else {
#pragma unroll
            for (int k = 0; k < C; k++)
                for (uint32_t u = u0[k]; u < (uint32_t)n && v < (uint32_t)K; u++) {
                    const uint2 f = kv[u];
                    if (f.x != nk[k]) break;
                    if ((int)f.y < i && aabb_overlap(mn0, mx0, smn[u], smx[u])) lst[v++] = (int)f.y;
                }
        }
    }
    if (large_num) {
        const int l0 = large_first[sc], nl = large_num[sc];
        const uint64_t segmask = LANES == 64 ? ~0ull : ((1ull << LANES) - 1ull);
        for (int q0 = 0; q0 < nl; q0 += LANES) {
            const int q = q0 + lane;
            bool hit = false, stop = false;
            int L = 0;
            if (q < nl) {
                L = large_list[l0 + q];
                stop = L >= i;  // list sorted by index
                hit = !stop && aabb_overlap(mn0, mx0, aabb_min[L], aabb_max[L]);
            }
            const uint64_t hm = (__ballot(hit) >> (LANES * seg)) & segmask;
            const uint64_t sm = (__ballot(stop) >> (LANES * seg)) & segmask;
            if (hit) {
                uint32_t w = c + (uint32_t)__popcll(hm & ((1ull << lane) - 1ull));
                if (w < (uint32_t)K) lst[w] = L;
            }
            c += (uint32_t)__popcll(hm);
            if (sm) break;
        }
    }
    if (c > (uint32_t)K) {
        if (lane == 0) atomicOr(overflow, (int)MI_RB_OVERFLOW_BODY_PAIRS);
        c = (uint32_t)K;
    }
    if constexpr (LANES == 1) {
        // insertion sort in place (distinct indices, a few per body)
        for (uint32_t k = 1; k < c; k++) {
            const int x = dst[k];
            int m = (int)k - 1;
            while (m >= 0 && dst[m] > x) {
                dst[m + 1] = dst[m];
                m--;
            }
            dst[m + 1] = x;
        }
    } else {
        wave_sync();  // the list's LDS stores come before the other lanes' loads
        for (uint32_t w = lane; w < c; w += LANES) {
            const int j = lst[w];
            uint32_t r = 0;
            for (uint32_t k = 0; k < c; k++) r += lst[k] < j ? 1u : 0u;
            dst[r] = j;
        }
    }
    if (lane == 0) pair_count[i] = c;
}

// Overflow list: pairs (x, j) of each large body x with every body j < x of
// its scene.  The scene range is split into chunks of LARGE_CHUNK bodies, one
// workgroup per (large body, chunk): a counting pass, then an ordered writing
// pass that places each chunk after the counts of the chunks before it.
constexpr int LARGE_CHUNK = RB_LARGE_CHUNK;

__device__ __forceinline__ int large_range(const int *large_list, const int *scene,
                                           const int *scene_start, int q, int c, int &x, int &j0,
                                           int &j1) {
    x = large_list[q];
    int s0 = scene ? scene_start[scene[x]] : 0;
    j0 = s0 + c * LARGE_CHUNK;
    j1 = j0 + LARGE_CHUNK < x ? j0 + LARGE_CHUNK : x;
    return j0 < j1;
}

__global__ __launch_bounds__(256) void k_large_count(
    const int *__restrict__ large_list, const int *__restrict__ scene,
    const int *__restrict__ scene_start, const float4 *__restrict__ aabb_min,
    const float4 *__restrict__ aabb_max, int nchunks, uint32_t *__restrict__ chunk_cnt) {
    __shared__ uint32_t s_n;
    const int q = blockIdx.x, c = blockIdx.y;
    int x, j0, j1;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    if (large_range(large_list, scene, scene_start, q, c, x, j0, j1)) {
        float4 mn0 = aabb_min[x], mx0 = aabb_max[x];
        uint32_t cnt = 0;
        for (int j = j0 + threadIdx.x; j < j1; j += 256)
            cnt += aabb_overlap(mn0, mx0, aabb_min[j], aabb_max[j]) ? 1u : 0u;
        if (cnt) atomicAdd(&s_n, cnt);
    }
    __syncthreads();
    if (threadIdx.x == 0) chunk_cnt[q * nchunks + c] = s_n;
}

__global__ __launch_bounds__(256) void k_large_write(
    const int *__restrict__ large_list, const int *__restrict__ large_off,
    const int *__restrict__ scene, const int *__restrict__ scene_start,
    const float4 *__restrict__ aabb_min, const float4 *__restrict__ aabb_max, int nchunks,
    const uint32_t *__restrict__ chunk_cnt, int *__restrict__ large_slots,
    uint32_t *__restrict__ pair_count) {
    __shared__ uint32_t s_w[4];
    __shared__ uint32_t s_base;
    const int q = blockIdx.x, c = blockIdx.y;
    int x, j0, j1;
    int have = large_range(large_list, scene, scene_start, q, c, x, j0, j1);
    if (threadIdx.x == 0) {
        uint32_t base = 0, tot = 0;
        for (int k = 0; k < nchunks; k++) {
            uint32_t v = chunk_cnt[q * nchunks + k];
            if (k < c) base += v;
            tot += v;
        }
        s_base = base;
        if (c == 0) pair_count[large_list[q]] = tot;
    }
    __syncthreads();
    if (!have) return;
    int *dst = large_slots + large_off[q];
    float4 mn0 = aabb_min[x], mx0 = aabb_max[x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int b0 = j0; b0 < j1; b0 += 256) {
        int j = b0 + threadIdx.x;
        bool hit = j < j1 && aabb_overlap(mn0, mx0, aabb_min[j], aabb_max[j]);
        uint64_t bal = __ballot(hit);
        uint32_t below = (uint32_t)__popcll(bal & lanemask_lt());
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t off = s_base, tot = 0;
        for (int w = 0; w < 4; w++) {
            if (w < wave) off += s_w[w];
            tot += s_w[w];
        }
        if (hit) dst[off + below] = j;
        __syncthreads();
        if (threadIdx.x == 0) s_base += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_emit_pairs(
    int n, const uint32_t *__restrict__ pair_count, const uint32_t *__restrict__ pair_off,
    const int *__restrict__ body_slots, int K, const int *__restrict__ is_large,
    const int *__restrict__ large_index, const int *__restrict__ large_off,
    const int *__restrict__ large_slots, int max_pairs, int2 *__restrict__ pairs,
    const uint32_t *total, int *num_pairs, int *overflow) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) {
        uint32_t tt = *total;
        if (tt > (uint32_t)max_pairs) {
            atomicOr(overflow, (int)MI_RB_OVERFLOW_PAIRS);
            tt = (uint32_t)max_pairs;
        }
        *num_pairs = (int)tt;
    }
    if (x >= n) return;
    if (is_large && is_large[x]) return;  // long lists: k_emit_large
    uint32_t c = pair_count[x], off = pair_off[x];
    const int *src = body_slots + (size_t)x * K;
    for (uint32_t k = 0; k < c; k++) {
        uint32_t o = off + k;
        if (o < (uint32_t)max_pairs) pairs[o] = make_int2(x, src[k]);
    }
}

// Pair lists of the overflow-list bodies can hold a whole scene's worth of
// partners: spread each one over a grid row of workgroups instead of one thread.
__global__ __launch_bounds__(256) void k_emit_large(
    const int *__restrict__ large_list, const uint32_t *__restrict__ pair_count,
    const uint32_t *__restrict__ pair_off, const int *__restrict__ large_off,
    const int *__restrict__ large_slots, int max_pairs, int2 *__restrict__ pairs) {
    const int q = blockIdx.x;
    const int x = large_list[q];
    const uint32_t c = pair_count[x], off = pair_off[x];
    const int *src = large_slots + large_off[q];
    for (uint32_t k = blockIdx.y * blockDim.x + threadIdx.x; k < c; k += gridDim.y * blockDim.x) {
        uint32_t o = off + k;
        if (o < (uint32_t)max_pairs) pairs[o] = make_int2(x, src[k]);
    }
}


// ---- hierarchical grid (stand-alone grid API) ----
// Level l has cells of cell0 * 2^l; an AABB lives on the lowest level whose
// cell is at least its largest extent, keyed (level << cell_bits) | cell of
// its centre.  Two overlapping AABBs on levels lx <= ly have centres within
// one level-ly cell per axis, so each pair is found exactly once, by the
// lower-level body: its 27 own-level neighbour cells (partners of lower index)
// and the 27 cells around its centre on every higher level (all partners).
struct HGridP {
    float inv_cell0;
    int dims[3];
    int cell_bits;
    int levels;
};

__device__ __forceinline__ int hg_level(float4 mn, float4 mx, float inv_cell0, int levels) {
    float ext = fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z) * inv_cell0;
    float c = 1.f;
    for (int l = 0; l < levels; l++, c *= 2.f)
        if (ext <= c) return l;
    return -1;
}
__device__ __forceinline__ void hg_cell(float4 mn, float4 mx, float inv_l, const int *dims,
                                        int gp[3]) {
    grid_pos(mn, mx, inv_l, dims, gp);
}

__global__ __launch_bounds__(256) void k_hg_hash(const float4 *__restrict__ mn,
                                                 const float4 *__restrict__ mx, int n, HGridP g,
                                                 uint2 *__restrict__ kv, uint32_t *__restrict__ counts,
                                                 uint32_t *__restrict__ cursor,
                                                 uint32_t *__restrict__ big_n, int *overflow) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { big_n[0] = 0u; big_n[1] = 0u; }
    if (i >= n) return;
    counts[i] = 0u;
    cursor[i] = 0u;
    float4 a = mn[i], b = mx[i];
    int l = hg_level(a, b, g.inv_cell0, g.levels);
    uint32_t key = KEY_NONE;
    if (l < 0) {
        atomicOr(overflow, (int)MI_RB_OVERFLOW_LEVEL);
    } else {
        int gp[3];
        hg_cell(a, b, g.inv_cell0 * exp2f(-(float)l), g.dims, gp);
        key = ((uint32_t)l << g.cell_bits) | pos_hash(gp[0], gp[1], gp[2], g.dims);
    }
    kv[i] = make_uint2(key, (uint32_t)i);
}

// Cell starts of the sorted keys, the occupancy bit of every occupied cell,
// a flag per occupied level, and the AABBs in sorted order (so that the pair
// search reads a cell member's key and box at one index).  occ[] and the
// level flags after it are cleared before this kernel.
__global__ __launch_bounds__(256) void k_hg_cells(const uint2 *__restrict__ kv, int n,
                                                  const float4 *__restrict__ aabb_min,
                                                  const float4 *__restrict__ aabb_max, HGridP g,
                                                  uint32_t *__restrict__ cell_start,
                                                  uint32_t *__restrict__ occ,
                                                  uint32_t *__restrict__ level_used,
                                                  float4 *__restrict__ smn,
                                                  float4 *__restrict__ smx) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint2 e = kv[s];
    smn[s] = aabb_min[e.y];
    smx[s] = aabb_max[e.y];
    if (e.x == KEY_NONE) return;
    if (s == 0 || kv[s - 1].x != e.x) {
        cell_start[e.x] = (uint32_t)s;
        atomicOr(&occ[e.x >> 5], 1u << (e.x & 31u));
        level_used[e.x >> g.cell_bits] = 1u;
    }
}

// HG_LANES lanes per sorted entry, lane q taking stencil cells q, q +
// HG_LANES, ...  A pair (x, y), x > y, belongs to body x.  WRITE = 0: count
// every found pair for its body; WRITE = 1: store y in the body's range of
// the scanned counts (slot order within a range is that of the atomics;
// k_pairs_group orders it).  The entry visits the 27 cells around it on its
// own level and on every coarser level that holds bodies, and walks only the
// occupied ones.  With lane_hits the count pass also records each lane's
// pair count and its first HG_KEEP partners, and the write pass stores
// those without walking again (only a lane that found more walks again).
constexpr int HG_LANES = RB_HG_LANES;
constexpr int HG_KEEP = RB_HG_KEEP;
template <int WRITE>
__global__ __launch_bounds__(256) void k_hg_pairs(int n, const uint2 *__restrict__ kv,
                                                  const uint32_t *__restrict__ cell_start,
                                                  const uint32_t *__restrict__ occ,
                                                  const uint32_t *__restrict__ level_used,
                                                  const float4 *__restrict__ smn,
                                                  const float4 *__restrict__ smx, HGridP g,
                                                  uint32_t *__restrict__ counts,
                                                  const uint32_t *__restrict__ offsets,
                                                  uint32_t *__restrict__ cursor, int max_pairs,
                                                  uint32_t *__restrict__ ys,
                                                  uint32_t *__restrict__ lane_hits) {
    constexpr int C = (27 + HG_LANES - 1) / HG_LANES;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = gid / HG_LANES, lane = gid % HG_LANES;
    if (t >= n) return;
    uint2 e = kv[t];
    if (e.x == KEY_NONE) return;
    const int x = (int)e.y;
    uint32_t *rec = lane_hits ? lane_hits + (size_t)gid * (1 + HG_KEEP) : nullptr;
    if (WRITE && rec) {
        const uint32_t found = rec[0];
        if (found <= (uint32_t)HG_KEEP) {
            for (uint32_t h = 0; h < found; h++) {
                const int y = (int)rec[1 + h];
                const uint32_t hi = (uint32_t)(x > y ? x : y), lo = (uint32_t)(x > y ? y : x);
                const uint32_t pos = offsets[hi] + atomicAdd(&cursor[hi], 1u);
                if (pos < (uint32_t)max_pairs) ys[pos] = lo;
            }
            return;
        }
    }
    uint32_t found = 0, keep[HG_KEEP] = {};
    const int lx = (int)(e.x >> g.cell_bits);
    const float4 mn0 = smn[t], mx0 = smx[t];
    for (int l = lx; l < g.levels; l++) {
        if (l != lx && !level_used[l]) continue;
        int gp[3];
        const float inv_l = g.inv_cell0 * exp2f(-(float)l);
        hg_cell(mn0, mx0, inv_l, g.dims, gp);
        const uint32_t reach = reachable_cells(mn0, mx0, inv_l);
        const uint32_t lk = (uint32_t)l << g.cell_bits;
        uint32_t nk[C], u0[C];
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int q = lane + k * HG_LANES;
            nk[k] = 0;
            u0[k] = (uint32_t)n;
            if (q < 27 && ((reach >> q) & 1u)) {
                const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
                nk[k] = lk | pos_hash(gp[0] + dx, gp[1] + dy, gp[2] + dz, g.dims);
                if ((occ[nk[k] >> 5] >> (nk[k] & 31u)) & 1u) u0[k] = cell_start[nk[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < C; k++)
            for (uint32_t u = u0[k]; u < (uint32_t)n; u++) {
                uint2 f = kv[u];
                if (f.x != nk[k]) break;
                int y = (int)f.y;
                if (l == lx && y >= x) continue;
                if (!aabb_overlap(mn0, mx0, smn[u], smx[u])) continue;
                const uint32_t hi = (uint32_t)(x > y ? x : y), lo = (uint32_t)(x > y ? y : x);
                if (!WRITE) {
                    atomicAdd(&counts[hi], 1u);
#pragma unroll
                    for (int h = 0; h < HG_KEEP; h++)
                        if (found == (uint32_t)h) keep[h] = (uint32_t)y;
                    found++;
                } else {
                    uint32_t pos = offsets[hi] + atomicAdd(&cursor[hi], 1u);
                    if (pos < (uint32_t)max_pairs) ys[pos] = lo;
                }
            }
    }
    if (!WRITE && rec) {
        rec[0] = found;
#pragma unroll
        for (int h = 0; h < HG_KEEP; h++) rec[1 + h] = keep[h];
    }
}

// The pair range of body x: [offsets[x], offsets[x] + counts[x]) clipped to
// the capacity.
__device__ __forceinline__ uint32_t pair_range(const uint32_t *counts, const uint32_t *offsets,
                                               int max_pairs, int x, uint32_t &off) {
    off = offsets[x];
    uint32_t len = counts[x];
    if (off >= (uint32_t)max_pairs) return 0u;
    if (off + len > (uint32_t)max_pairs) len = (uint32_t)max_pairs - off;
    return len;
}

constexpr int GROUP_SMALL = 8;

// One thread per body: its range sorted by y (an odd-even transposition
// network over GROUP_SMALL registers) and written as int2 (x, y).  Longer
// ranges go to a list for k_pairs_group_big (from the front of big[]), those
// longer than huge_min to one for k_pairs_group_huge (from the back).  Also
// the pair count and the capacity flag.
__global__ __launch_bounds__(256) void k_pairs_group(int n, const uint32_t *__restrict__ counts,
                                                     const uint32_t *__restrict__ offsets,
                                                     const uint32_t *__restrict__ total,
                                                     int max_pairs, const uint32_t *__restrict__ ys,
                                                     int2 *__restrict__ pairs, int *num_pairs,
                                                     int *overflow, uint32_t *__restrict__ big,
                                                     uint32_t *__restrict__ big_n, int huge_min) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) {
        uint32_t tt = *total;
        if (tt > (uint32_t)max_pairs) atomicOr(overflow, (int)MI_RB_OVERFLOW_PAIRS);
        *num_pairs = (int)(tt < (uint32_t)max_pairs ? tt : (uint32_t)max_pairs);
    }
    if (x >= n) return;
    uint32_t off;
    const uint32_t len = pair_range(counts, offsets, max_pairs, x, off);
    if (len == 0) return;
    if (len > (uint32_t)huge_min) {
        big[n - 1 - (int)atomicAdd(big_n + 1, 1u)] = (uint32_t)x;  // from the back
        return;
    }
    if (len > (uint32_t)GROUP_SMALL) {
        big[atomicAdd(big_n, 1u)] = (uint32_t)x;
        return;
    }
    uint32_t v[GROUP_SMALL];
#pragma unroll
    for (int k = 0; k < GROUP_SMALL; k++) v[k] = (uint32_t)k < len ? ys[off + k] : 0xffffffffu;
#pragma unroll
    for (int r = 0; r < GROUP_SMALL; r++)
#pragma unroll
        for (int k = r & 1; k + 1 < GROUP_SMALL; k += 2) {
            uint32_t a = v[k], b = v[k + 1];
            v[k] = a < b ? a : b;
            v[k + 1] = a < b ? b : a;
        }
#pragma unroll
    for (int k = 0; k < GROUP_SMALL; k++)
        if ((uint32_t)k < len) pairs[off + k] = make_int2(x, (int)v[k]);
}

// Bodies with more than GROUP_SMALL pairs, one workgroup each: the rank of
// every y among its range (the y of one body are distinct) in LDS chunks.
constexpr int GROUP_CHUNK = 1024;
__global__ __launch_bounds__(256) void k_pairs_group_big(const uint32_t *__restrict__ counts,
                                                         const uint32_t *__restrict__ offsets,
                                                         int max_pairs,
                                                         const uint32_t *__restrict__ ys,
                                                         int2 *__restrict__ pairs,
                                                         const uint32_t *__restrict__ big,
                                                         const uint32_t *__restrict__ big_n) {
    __shared__ uint32_t sy[GROUP_CHUNK];
    const uint32_t nb = *big_n;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int x = (int)big[b];
        uint32_t off;
        const uint32_t len = pair_range(counts, offsets, max_pairs, x, off);
        for (uint32_t i0 = 0; i0 < len; i0 += 256) {
            const uint32_t i = i0 + threadIdx.x;
            const uint32_t yi = i < len ? ys[off + i] : 0u;
            uint32_t rank = 0;
            for (uint32_t j0 = 0; j0 < len; j0 += GROUP_CHUNK) {
                const uint32_t m = len - j0 < (uint32_t)GROUP_CHUNK ? len - j0 : GROUP_CHUNK;
                __syncthreads();
                for (uint32_t j = threadIdx.x; j < m; j += 256) sy[j] = ys[off + j0 + j];
                __syncthreads();
                for (uint32_t j = 0; j < m; j++) rank += sy[j] < yi ? 1u : 0u;
            }
            if (i < len) pairs[off + rank] = make_int2(x, (int)yi);
        }
    }
}

// Bodies with more than huge_min pairs (a ground touching thousands of
// bodies), one workgroup each: the y of the range are distinct indices below
// x, so one bit per index in LDS (cleared, set), per-thread chunk counts and
// their prefix give every y its rank.  Dynamic LDS: (n + 31) / 32 words.
constexpr int HUGE_BLOCK = 1024;
__global__ __launch_bounds__(HUGE_BLOCK) void k_pairs_group_huge(
    int n, const uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
    int max_pairs, const uint32_t *__restrict__ ys, int2 *__restrict__ pairs,
    const uint32_t *__restrict__ big, const uint32_t *__restrict__ big_n) {
    extern __shared__ uint32_t bm[];
    __shared__ uint32_t cbase[HUGE_BLOCK];
    const int tid = threadIdx.x;
    const uint32_t nh = big_n[1];
    for (uint32_t b = blockIdx.x; b < nh; b += gridDim.x) {
        const int x = (int)big[n - 1 - (int)b];
        uint32_t off;
        const uint32_t len = pair_range(counts, offsets, max_pairs, x, off);
        const int W = (x + 31) >> 5;
        const int chunk = (W + HUGE_BLOCK - 1) / HUGE_BLOCK;
        __syncthreads();
        for (int w = tid; w < W; w += HUGE_BLOCK) bm[w] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < len; i += HUGE_BLOCK) {
            const uint32_t y = ys[off + i];
            atomicOr(&bm[y >> 5], 1u << (y & 31u));
        }
        __syncthreads();
        uint32_t c = 0;
        for (int w = tid * chunk; w < (tid + 1) * chunk && w < W; w++) c += __popc(bm[w]);
        cbase[tid] = c;
        __syncthreads();
        // exclusive prefix of the chunk counts (Hillis-Steele in place)
        for (int o = 1; o < HUGE_BLOCK; o <<= 1) {
            const uint32_t v = tid >= o ? cbase[tid - o] : 0u;
            __syncthreads();
            cbase[tid] += v;
            __syncthreads();
        }
        const uint32_t excl = cbase[tid] - c;
        __syncthreads();
        cbase[tid] = excl;
        __syncthreads();
        for (uint32_t i = tid; i < len; i += HUGE_BLOCK) {
            const uint32_t y = ys[off + i];
            const int w = (int)(y >> 5), ck = w / chunk;
            uint32_t r = cbase[ck];
            for (int k = ck * chunk; k < w; k++) r += __popc(bm[k]);
            r += __popc(bm[w] & ((1u << (y & 31u)) - 1u));
            pairs[off + r] = make_int2(x, (int)y);
        }
    }
}

// ---- sweep and prune (btGpuSapBroadphase::calculateOverlappingPairs,
// opencl/gpu_rigidbody_pipeline2/btGpuSapBroadphase.cpp:91-434; kernels
// opencl/broadphase_benchmark/sap.cl:90-116, 290-320) ----

// order-preserving map of a float to uint32 (sap.cl FloatFlip: negative
// values have every bit flipped, others only the sign bit)
__device__ __forceinline__ uint32_t float_flip(float f) {
    const uint32_t u = __float_as_uint(f);
    return u ^ ((uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u);
}
__device__ __forceinline__ float axis_of4(float4 v, int axis) {
    return axis == 0 ? v.x : (axis == 1 ? v.y : v.z);
}

// (FloatFlip(min[axis]), index) per AABB; large AABBs (is_large[i] != 0)
// sort last and go to a list of their own (lg, count lg_n, zero at launch).
// Zeroes the per-body pair counts.
__global__ __launch_bounds__(256) void k_sap_keys(const float4 *__restrict__ mn,
                                                  const int *__restrict__ is_large, int n, int axis,
                                                  uint2 *__restrict__ kv,
                                                  uint32_t *__restrict__ counts,
                                                  uint32_t *__restrict__ cursor,
                                                  uint32_t *__restrict__ big_n,
                                                  uint32_t *__restrict__ lg,
                                                  uint32_t *__restrict__ lg_n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { big_n[0] = 0u; big_n[1] = 0u; }
    if (i >= n) return;
    const bool large = is_large && is_large[i] != 0;
    kv[i] = make_uint2(large ? 0xffffffffu : float_flip(axis_of4(mn[i], axis)), (uint32_t)i);
    if (large) lg[atomicAdd(lg_n, 1u)] = (uint32_t)i;
    counts[i] = 0u;
    cursor[i] = 0u;
}

// Pairs of the large AABBs (sap2Kernel's role, btGpuSapBroadphase.cpp:311-333):
// one thread per sorted AABB j tests it against every large AABB L; a pair
// of two large AABBs is taken by the one of lower index only.
template <int WRITE>
__global__ __launch_bounds__(256) void k_sap_large(int n, const float4 *__restrict__ mn,
                                                   const float4 *__restrict__ mx,
                                                   const float4 *__restrict__ smn,
                                                   const float4 *__restrict__ smx,
                                                   const uint32_t *__restrict__ lg,
                                                   const uint32_t *__restrict__ lg_n,
                                                   uint32_t *__restrict__ counts,
                                                   const uint32_t *__restrict__ offsets,
                                                   uint32_t *__restrict__ cursor, int max_pairs,
                                                   uint32_t *__restrict__ ys) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int nl = (int)*lg_n;
    if (nl == 0) return;
    const int ns = n - nl;
    const float4 bmn = smn[j], bmx = smx[j];
    const int y = __float_as_int(bmn.w);
    for (int q = 0; q < nl; q++) {
        const int L = (int)lg[q];
        if (L == y || (j >= ns && y > L)) continue;
        if (!aabb_overlap(mn[L], mx[L], bmn, bmx)) continue;
        const uint32_t hi = (uint32_t)(L > y ? L : y), lo = (uint32_t)(L > y ? y : L);
        if (!WRITE) {
            atomicAdd(&counts[hi], 1u);
        } else {
            const uint32_t pos = offsets[hi] + atomicAdd(&cursor[hi], 1u);
            if (pos < (uint32_t)max_pairs) ys[pos] = lo;
        }
    }
}

// AABBs in sorted order, the original index in min.w (as int bits)
__global__ __launch_bounds__(256) void k_sap_scatter(const uint2 *__restrict__ kv, int n,
                                                     const float4 *__restrict__ mn,
                                                     const float4 *__restrict__ mx,
                                                     float4 *__restrict__ smn,
                                                     float4 *__restrict__ smx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = kv[i].y;
    float4 a = mn[k];
    a.w = __int_as_float((int)k);
    smn[i] = a;
    smx[i] = mx[k];
}

// One thread per sorted small AABB i: the small AABBs after it whose min
// along the axis does not pass its max (computePairsKernelOriginal), each
// tested on all three axes.  A pair (x, y), x > y, belongs to body x: WRITE = 0 counts it,
// WRITE = 1 stores y in x's range of the scanned counts.
template <int WRITE>
__global__ __launch_bounds__(256) void k_sap_pairs(int n, int axis, const float4 *__restrict__ smn,
                                                   const float4 *__restrict__ smx,
                                                   const uint32_t *__restrict__ lg_n,
                                                   uint32_t *__restrict__ counts,
                                                   const uint32_t *__restrict__ offsets,
                                                   uint32_t *__restrict__ cursor, int max_pairs,
                                                   uint32_t *__restrict__ ys) {
    n -= (int)*lg_n;  // the large AABBs sorted last do not sweep
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 amn = smn[i], amx = smx[i];
    const int x = __float_as_int(amn.w);
    const float lim = axis_of4(amx, axis);
    for (int j = i + 1; j < n; j++) {
        const float4 bmn = smn[j];
        if (lim < axis_of4(bmn, axis)) break;
        if (!aabb_overlap(amn, amx, bmn, smx[j])) continue;
        const int y = __float_as_int(bmn.w);
        const uint32_t hi = (uint32_t)(x > y ? x : y), lo = (uint32_t)(x > y ? y : x);
        if (!WRITE) {
            atomicAdd(&counts[hi], 1u);
        } else {
            const uint32_t pos = offsets[hi] + atomicAdd(&cursor[hi], 1u);
            if (pos < (uint32_t)max_pairs) ys[pos] = lo;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers
static GridP grid_of(const StepParams &p) {
    GridP g;
    g.inv_cell = p.inv_cell;
    g.dims[0] = p.dims[0];
    g.dims[1] = p.dims[1];
    g.dims[2] = p.dims[2];
    g.cell_bits = p.cell_bits;
    return g;
}

hipError_t launch_setup(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    if (p.n_bodies == 0) return hipSuccess;
    RB_LAUNCH(k_setup, dim3((p.n_bodies + 255) / 256), dim3(256), 0, s, p.n_bodies,
                       d.pos, d.quat, d.linvel, d.angvel, d.inv_mass, d.shape, d.scene, d.is_large,
                       d.hulls, d.inv_inertia_local, d.inv_inertia_world, d.aabb_min, d.aabb_max,
                       d.kv, grid_of(p), p.gravity[0], p.gravity[1], p.gravity[2], p.dt,
                       d.occ, d.occ_words);
    return rb_launch_status();
}

hipError_t launch_cell_start(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    const int n = p.n_bodies;
    if (n == 0) return hipSuccess;
    RB_LAUNCH(k_cell_start, dim3((n + 255) / 256), dim3(256), 0, s, d.kv, n, d.aabb_min,
                       d.aabb_max, d.cell_start, d.occ, d.smn, d.smx);
    return rb_launch_status();
}

hipError_t launch_find_pairs(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    int n = p.n_bodies;
    if (n == 0) return hipSuccess;
    int K = p.max_pairs_per_body;
    // 32 lanes per body while the bodies alone do not fill the chip, else
    // one lane per body (its cells walked in turn)
    const int lanes = n < (1 << 18) ? 32 : 1;
    const int entries = FP_BLOCK / lanes;
    size_t lds = lanes == 1 ? 0 : sizeof(int) * (size_t)entries * (K | 1);
    auto kern = lanes == 32 ? k_find_pairs<32> : k_find_pairs<1>;
    RB_LAUNCH(kern, dim3((n + entries - 1) / entries), dim3(FP_BLOCK), lds, s, n,
                       d.kv, d.cell_start, d.occ, d.smn, d.smx, d.aabb_min, d.aabb_max,
                       d.large_list,
                       p.n_large ? d.large_first : nullptr, p.n_large ? d.large_num : nullptr,
                       grid_of(p), K, d.body_slots, d.pair_count, d.counters + RB_CNT_OVERFLOW);
    if (p.n_large > 0) {
        dim3 g(p.n_large, p.large_chunks);
        RB_LAUNCH(k_large_count, g, dim3(256), 0, s, d.large_list, d.scene,
                           d.scene_start, d.aabb_min, d.aabb_max, p.large_chunks, d.large_chunk_cnt);
        RB_LAUNCH(k_large_write, g, dim3(256), 0, s, d.large_list, d.large_off, d.scene,
                           d.scene_start, d.aabb_min, d.aabb_max, p.large_chunks,
                           d.large_chunk_cnt, d.large_slots, d.pair_count);
    }
    return rb_launch_status();
}

hipError_t launch_emit_pairs(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    int n = p.n_bodies;
    if (n == 0) return hipSuccess;
    RB_LAUNCH(k_emit_pairs, dim3((n + 255) / 256), dim3(256), 0, s, n, d.pair_count,
                       d.pair_off, d.body_slots, p.max_pairs_per_body, d.is_large, d.large_index,
                       d.large_off, d.large_slots, p.max_pairs, d.pairs,
                       (const uint32_t *)(d.counters + RB_CNT_SORT_TILE),
                       d.counters + RB_CNT_PAIRS, d.counters + RB_CNT_OVERFLOW);
    if (p.n_large > 0)
        RB_LAUNCH(k_emit_large, dim3(p.n_large, 32), dim3(256), 0, s, d.large_list,
                           d.pair_count, d.pair_off, d.large_off, d.large_slots, p.max_pairs,
                           d.pairs);
    return rb_launch_status();
}

// Each body's pair range ordered by y: ranges of up to GROUP_SMALL pairs in
// registers, up to HUGE_MIN by rank in LDS chunks, longer ones (a ground)
// through an LDS bitmap of the indices below the body when it fits.
constexpr int HUGE_MIN = 4096;
constexpr size_t HUGE_LDS_MAX = 60 * 1024;
static void launch_pairs_group(int n, uint32_t *counts, uint32_t *offsets, uint32_t *total,
                               int max_pairs, uint32_t *ys, int2 *pairs, int *num_pairs,
                               int *overflow, uint32_t *big, uint32_t *big_n, int blocks,
                               hipStream_t s) {
    const size_t bm_bytes = sizeof(uint32_t) * (((size_t)n + 31) / 32 + 1);
    const bool huge = bm_bytes <= HUGE_LDS_MAX;
    RB_LAUNCH(k_pairs_group, dim3(blocks), dim3(256), 0, s, n, counts, offsets, total,
                       max_pairs, ys, pairs, num_pairs, overflow, big, big_n,
                       huge ? HUGE_MIN : INT_MAX);
    if (n > 0) {
        RB_LAUNCH(k_pairs_group_big, dim3(1024), dim3(256), 0, s, counts, offsets,
                           max_pairs, ys, pairs, big, big_n);
        if (huge)
            RB_LAUNCH(k_pairs_group_huge, dim3(64), dim3(HUGE_BLOCK), bm_bytes, s, n,
                               counts, offsets, max_pairs, ys, pairs, big, big_n);
    }
}

// Hierarchical-grid pair search of n AABBs: hash, sort, cell starts, count
// per body, scan, write, then each body's pairs ordered by y.
hipError_t launch_hgrid_pairs(const float4 *mn, const float4 *mx, int n, float inv_cell0,
                              const int dims[3], int cell_bits, int levels, uint2 *kv,
                              uint32_t *cell_start, uint32_t *occ, float4 *smn, float4 *smx,
                              uint32_t *counts, uint32_t *offsets, uint32_t *cursor,
                              uint32_t *big, uint32_t *total, uint32_t *ys, int max_pairs,
                              int2 *pairs, int *num_pairs, int *overflow, const SortWork &sw,
                              uint32_t *scan_status, hipStream_t s, hipEvent_t *ev,
                              uint32_t *lane_hits) {
    HGridP g;
    g.inv_cell0 = inv_cell0;
    g.dims[0] = dims[0]; g.dims[1] = dims[1]; g.dims[2] = dims[2];
    g.cell_bits = cell_bits;
    g.levels = levels;
    const int blocks = n > 0 ? (n + 255) / 256 : 1;
    const int hg_blocks = (int)(((int64_t)n * HG_LANES + 255) / 256);
    uint32_t *big_n = big + (n > 0 ? n : 0);
    if (ev) hipEventRecord(ev[0], s);
    RB_LAUNCH(k_hg_hash, dim3(blocks), dim3(256), 0, s, mn, mx, n, g, kv, counts, cursor,
                       big_n, overflow);
    if (ev) hipEventRecord(ev[1], s);
    hipError_t e = launch_radix_sort(kv, n, cell_bits + 2, 8, sw, s);
    if (e != hipSuccess) return e;
    if (ev) hipEventRecord(ev[2], s);
    // occupancy words of every level's cells, then one flag per level
    const size_t occ_words = ((size_t)1 << (cell_bits + 2)) / 32;
    uint32_t *level_used = occ + occ_words;
    if (n > 0) {
        e = launch_fill_u32(occ, 0u, occ_words + 4, s);
        if (e != hipSuccess) return e;
        RB_LAUNCH(k_hg_cells, dim3(blocks), dim3(256), 0, s, kv, n, mn, mx, g, cell_start,
                           occ, level_used, smn, smx);
        RB_LAUNCH(k_hg_pairs<0>, dim3(hg_blocks), dim3(256), 0, s, n, kv, cell_start, occ,
                           level_used, smn, smx, g, counts, (const uint32_t *)nullptr, cursor,
                           max_pairs, ys, lane_hits);
    }
    e = launch_exclusive_scan(counts, offsets, n, total, scan_status, s, nullptr, overflow);
    if (e != hipSuccess) return e;
    if (n > 0)
        RB_LAUNCH(k_hg_pairs<1>, dim3(hg_blocks), dim3(256), 0, s, n, kv, cell_start, occ,
                           level_used, smn, smx, g, counts, offsets, cursor, max_pairs, ys,
                           lane_hits);
    if (ev) hipEventRecord(ev[3], s);
    launch_pairs_group(n, counts, offsets, total, max_pairs, ys, pairs, num_pairs, overflow, big,
                       big_n, blocks, s);
    if (ev) hipEventRecord(ev[4], s);
    return rb_launch_status();
}

// Sweep and prune of n AABBs along `axis`: keys, 32-bit sort, scatter, count
// per body, scan, write, then each body's pairs ordered by y (as the grid).
hipError_t launch_sap_pairs(const float4 *mn, const float4 *mx, const int *is_large, int n,
                            int axis, uint2 *kv, float4 *smn, float4 *smx, uint32_t *counts,
                            uint32_t *offsets, uint32_t *cursor, uint32_t *big, uint32_t *lg,
                            uint32_t *total, uint32_t *ys, int max_pairs, int2 *pairs,
                            int *num_pairs, int *overflow, const SortWork &sw,
                            uint32_t *scan_status, hipStream_t s, hipEvent_t *ev) {
    const int blocks = n > 0 ? (n + 255) / 256 : 1;
    uint32_t *big_n = big + (n > 0 ? n : 0);
    uint32_t *lg_n = lg + (n > 0 ? n : 0);
    if (ev) hipEventRecord(ev[0], s);
    hipError_t e = launch_fill_u32(lg_n, 0u, 1, s);
    if (e != hipSuccess) return e;
    RB_LAUNCH(k_sap_keys, dim3(blocks), dim3(256), 0, s, mn, is_large, n, axis, kv, counts,
                       cursor, big_n, lg, lg_n);
    if (ev) hipEventRecord(ev[1], s);
    e = launch_radix_sort(kv, n, 32, 8, sw, s);
    if (e != hipSuccess) return e;
    if (ev) hipEventRecord(ev[2], s);
    if (n > 0) {
        RB_LAUNCH(k_sap_scatter, dim3(blocks), dim3(256), 0, s, kv, n, mn, mx, smn, smx);
        RB_LAUNCH(k_sap_pairs<0>, dim3(blocks), dim3(256), 0, s, n, axis, smn, smx, lg_n,
                           counts, (const uint32_t *)nullptr, cursor, max_pairs, ys);
        RB_LAUNCH(k_sap_large<0>, dim3(blocks), dim3(256), 0, s, n, mn, mx, smn, smx, lg,
                           lg_n, counts, (const uint32_t *)nullptr, cursor, max_pairs, ys);
    }
    e = launch_exclusive_scan(counts, offsets, n, total, scan_status, s, nullptr, overflow);
    if (e != hipSuccess) return e;
    if (n > 0) {
        RB_LAUNCH(k_sap_pairs<1>, dim3(blocks), dim3(256), 0, s, n, axis, smn, smx, lg_n,
                           counts, offsets, cursor, max_pairs, ys);
        RB_LAUNCH(k_sap_large<1>, dim3(blocks), dim3(256), 0, s, n, mn, mx, smn, smx, lg,
                           lg_n, counts, offsets, cursor, max_pairs, ys);
    }
    if (ev) hipEventRecord(ev[3], s);
    launch_pairs_group(n, counts, offsets, total, max_pairs, ys, pairs, num_pairs, overflow, big,
                       big_n, blocks, s);
    if (ev) hipEventRecord(ev[4], s);
    return rb_launch_status();
}
