// world.cpp -- host side of the rigid-body step: the C ABI of mi_rb.h.
//
// Mirrors CLPhysicsDemo (opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.cpp) and the
// registry of btGpuNarrowphaseAndSolver (opencl/gpu_rigidbody_pipeline/btGpuNarrowPhaseAndSolver.cpp:
// registerConvexHullShape :423-490, registerRigidBody :566-657): shapes and bodies are
// registered on the host, uploaded once, and each step is a fixed sequence of
// kernel launches on one HIP stream with no host synchronisation inside, so it
// can be captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hull.h"
#include "rb_internal.h"

mi_rb_status rb_device_error(hipError_t e, const char *what, const char *file, int line) {
    if (getenv("MI_RB_VERBOSE"))
        fprintf(stderr, "mi_rb: %s failed: %s (%s:%d)\n", what, hipGetErrorString(e), file, line);
    return MI_RB_DEVICE;
}

int rb_num_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

namespace {

template <typename T>
hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, sizeof(T) * n);
}

int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
int ilog2(uint64_t v) {
    int b = 0;
    while ((1ull << b) < v) b++;
    return b;
}

// Start of a step: the counters (the overflow bits are sticky), the
// solver's per-sweep residual maxima, the scans' status words and the sorts'
// digit and tile counters, and the solver tiles' row ranges.
__global__ void k_zero_counters(int *c, float *residual, int n_residual, uint32_t *scan_status,
                                int n_scan_words, int *tile_first, int *tile_cnt, int n_tiles) {
    const int i = threadIdx.x;
    if (blockIdx.x == 0) {
        if (i < RB_CNT_COUNT && i != RB_CNT_OVERFLOW) c[i] = 0;
        for (int k = i; k < n_residual; k += blockDim.x) residual[k] = 0.f;
    }
    // the status words of the step's scans (one region per scan), then the
    // step's sort heads
    for (int k = blockIdx.x * blockDim.x + i; k < n_scan_words; k += gridDim.x * blockDim.x)
        scan_status[k] = 0u;
    // the solver tiles' row ranges (k_tile_bounds2 sets the non-empty ones)
    for (int k = blockIdx.x * blockDim.x + i; k < n_tiles; k += gridDim.x * blockDim.x) {
        tile_first[k] = 0;
        tile_cnt[k] = 0;
    }
}

// Launch shapes used when mi_rb_tuning leaves a field 0 (the autotuner's
// choice on the 65,537-hull scene and on the 4,096 batched scenes).
constexpr int DEFAULT_SOLVE_BLOCK = 256;
constexpr int DEFAULT_NARROW_WAVES = 8;

struct World {
    mi_rb_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    bool uploaded = false;
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev[10] = {};
    hipEvent_t ev_order = nullptr;  // orders work on a caller's stream after the world's steps
    bool timing = true;
    bool events_full = false;  // every stage event of the last eager step was recorded
    std::vector<hipEvent_t> kev;  // eager timing: around every k_tile_solve launch
    int kev_launches = 0;         // launches timed by kev in the last eager step
    bool solve_timing = false;    // events around the solver launches of every step
    hipEvent_t span[2] = {};      // ... recorded before the first and after the last
    int span_launches = 0;        // solver launches between them
    // learned warm-start weights (host copy; uploaded with the world)
    std::vector<uint16_t> mlp_w_host;
    std::vector<float> mlp_b_host;

    // host shapes
    std::vector<HullInfo> hulls;
    std::vector<ChildShape> children;  // compound children (HullInfo edge_offset/num_edges)
    std::vector<float4> verts, planes, edges;
    std::vector<int2> face_index;
    std::vector<int> indices;
    std::vector<float> hull_hext_norm;  // |local AABB half extent| per shape

    // host bodies
    std::vector<float4> pos, quat, linvel, angvel, inv_inertia_local;
    std::vector<float> inv_mass;
    std::vector<int> shape, scene;

    // derived
    StepParams sp{};
    int key_bits = 0, tile_bits = 0;
    std::vector<int> is_large, large_list;
    DeviceWorld d{};
    SortWork sw{};
    SortWork sw_step[RB_STEP_SORTS];  // sw with the step's precleared heads
    uint32_t *scan_status = nullptr;  // scan regions, then the step's sort heads
    size_t scan_stride = 0;  // words per scan region (RB_SCAN_REGIONS regions)
    long long steps = 0;
    float last_ms = 0.f;
    mi_rb_tuning tuning{0, 0, 0, 0};  // launch shapes; 0 = default
    std::vector<void *> allocs;

    template <typename T>
    hipError_t alloc(T **p, size_t n) {
        hipError_t e = dalloc(p, n);
        if (e == hipSuccess) allocs.push_back((void *)*p);
        return e;
    }
    void free_device() {
        if (graph) { hipGraphExecDestroy(graph); graph = nullptr; }
        for (void *p : allocs) hipFree(p);
        allocs.clear();
        d = DeviceWorld();
        uploaded = false;
    }
};

World *W(void *w) { return static_cast<World *>(w); }

// inertia at registration (btGpuNarrowPhaseAndSolver.cpp:566-657 with the
// world AABB of CLPhysicsDemo::registerPhysicsInstance, CLPhysicsDemo.cpp:536-552)
float4 local_inv_inertia(const World &w, float mass, int shape, const float pos[3],
                         const float q[4]) {
    if (mass == 0.f) return make_float4(0, 0, 0, 0);
    const HullInfo &h = w.hulls[shape];
    const float lmin[3] = {h.aabb_min.x, h.aabb_min.y, h.aabb_min.z};
    const float lmax[3] = {h.aabb_max.x, h.aabb_max.y, h.aabb_max.z};
    float lx, ly, lz;
    if (w.cfg.inertia_mode == MI_RB_INERTIA_BOX) {
        lx = lmax[0] - lmin[0]; ly = lmax[1] - lmin[1]; lz = lmax[2] - lmin[2];
    } else {
        // btTransformAabb (bullet2/LinearMath/btAabbUtil2.h:193-207) with
        // margin 0.01 and the body's transform, the rotation as
        // btMatrix3x3::setRotation forms it; then "fake larger inertia":
        // l = 2 * (aabbMax - aabbMin)
        const float margin = 0.01f;
        float x = q[0], y = q[1], z = q[2], qw = q[3];
        float d = ((x * x + y * y) + z * z) + qw * qw;
        float s = 2.0f / d;
        float xs = x * s, ys = y * s, zs = z * s;
        float wx = qw * xs, wy = qw * ys, wz = qw * zs;
        float xx = x * xs, xy = x * ys, xz = x * zs;
        float yy = y * ys, yz = y * zs, zz = z * zs;
        float m[3][3] = {{1.0f - (yy + zz), xy - wz, xz + wy},
                         {xy + wz, 1.0f - (xx + zz), yz - wx},
                         {xz - wy, yz + wx, 1.0f - (xx + yy)}};
        float he[3], lc[3], l[3];
        for (int k = 0; k < 3; k++) {
            he[k] = 0.5f * (lmax[k] - lmin[k]);
            he[k] = he[k] + margin;
            lc[k] = 0.5f * (lmax[k] + lmin[k]);
        }
        for (int r = 0; r < 3; r++) {
            float c = ((m[r][0] * lc[0] + m[r][1] * lc[1]) + m[r][2] * lc[2]) + pos[r];
            float e = (he[0] * std::fabs(m[r][0]) + he[1] * std::fabs(m[r][1])) +
                      he[2] * std::fabs(m[r][2]);
            l[r] = 2.f * ((c + e) - (c - e));
        }
        lx = l[0]; ly = l[1]; lz = l[2];
    }
    float ix = (mass / 12.0f) * (ly * ly + lz * lz);
    float iy = (mass / 12.0f) * (lx * lx + lz * lz);
    float iz = (mass / 12.0f) * (lx * lx + ly * ly);
    return make_float4(1.f / ix, 1.f / iy, 1.f / iz, 0.f);
}

mi_rb_status derive_and_upload(World &w) {
    const int n = (int)w.pos.size();
    const int S = (int)w.hulls.size();
    if (n == 0 || S == 0) return MI_RB_STATE;
    w.free_device();
    StepParams &p = w.sp;
    p = StepParams();
    p.n_bodies = n;
    p.has_vertexless = 0;
    for (const HullInfo &h : w.hulls) p.has_vertexless |= h.num_vertices == 0;
    int n_scenes = 0;
    for (int i = 0; i < n; i++) n_scenes = std::max(n_scenes, w.scene[i] + 1);
    p.n_scenes = n_scenes;
    std::vector<int> scene_start(n_scenes, 0), scene_count(n_scenes, 0);
    for (int i = n - 1; i >= 0; i--) scene_start[w.scene[i]] = i;
    for (int i = 0; i < n; i++) scene_count[w.scene[i]]++;
    // cell size: 2 x the largest rotated-AABB half extent of a dynamic body
    float hmax = 0.f;
    for (int i = 0; i < n; i++)
        if (w.inv_mass[i] != 0.f) hmax = std::max(hmax, w.hull_hext_norm[w.shape[i]]);
    if (hmax == 0.f)
        for (int i = 0; i < n; i++) hmax = std::max(hmax, w.hull_hext_norm[w.shape[i]]);
    float cell = w.cfg.cell_size > 0.f ? w.cfg.cell_size : 2.f * hmax;
    p.inv_cell = 1.0f / cell;
    // large bodies: rotated AABB may exceed half a cell
    w.is_large.assign(n, 0);
    w.large_list.clear();
    for (int i = 0; i < n; i++)
        if (w.hull_hext_norm[w.shape[i]] > 0.5f * cell) { w.is_large[i] = 1; w.large_list.push_back(i); }
    p.n_large = (int)w.large_list.size();
    // grid: power-of-two dims covering the (per-scene) extent of the small bodies
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    std::vector<float> slo(3 * (size_t)n_scenes, 1e30f), shi(3 * (size_t)n_scenes, -1e30f);
    for (int i = 0; i < n; i++) {
        if (w.is_large[i]) continue;
        float c[3] = {w.pos[i].x, w.pos[i].y, w.pos[i].z};
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], c[k]); hi[k] = std::max(hi[k], c[k]);
            slo[3 * w.scene[i] + k] = std::min(slo[3 * w.scene[i] + k], c[k]);
            shi[3 * w.scene[i] + k] = std::max(shi[3 * w.scene[i] + k], c[k]);
        }
    }
    float ext[3] = {0, 0, 0};
    for (int s = 0; s < n_scenes; s++)
        for (int k = 0; k < 3; k++)
            if (shi[3 * s + k] >= slo[3 * s + k]) ext[k] = std::max(ext[k], shi[3 * s + k] - slo[3 * s + k]);
    int dims[3];
    for (int k = 0; k < 3; k++) {
        int want = (int)std::ceil(ext[k] / cell) + 3;
        dims[k] = w.cfg.grid_dims[k] > 0 ? w.cfg.grid_dims[k] : std::min(std::max(next_pow2(want), 4), 512);
        if (dims[k] & (dims[k] - 1)) return MI_RB_INVALID_ARG;
    }
    int scene_bits = ilog2((uint64_t)n_scenes);
    while (ilog2((uint64_t)dims[0] * dims[1] * dims[2]) + scene_bits > 30) {
        int k = dims[0] >= dims[1] && dims[0] >= dims[2] ? 0 : (dims[1] >= dims[2] ? 1 : 2);
        dims[k] >>= 1;
    }
    for (int k = 0; k < 3; k++) p.dims[k] = dims[k];
    p.cell_bits = ilog2((uint64_t)dims[0] * dims[1] * dims[2]);
    // one spare key bit: overflow-list bodies carry key 0xffffffff, whose low
    // key_bits bits then exceed every cell key, so they sort last
    w.key_bits = p.cell_bits + scene_bits + 1;
    // parameters
    const mi_rb_config &c = w.cfg;
    p.dt = c.dt;
    p.solver_dt = c.solver_dt > 0.f ? c.solver_dt : c.dt;
    for (int k = 0; k < 3; k++) p.gravity[k] = c.gravity[k];
    p.max_pairs = c.max_pairs;
    p.max_pairs_per_body = c.max_pairs_per_body;
    p.max_contacts = c.max_contacts;
    p.friction = c.friction;
    p.baumgarte = c.baumgarte;
    p.drift = c.position_drift;
    p.threshold = c.contact_threshold;
    p.damping = c.angular_damping;
    p.max_ang = c.max_angular_motion;
    p.sat_edge_reject = c.sat_mode == MI_RB_SAT_CPU_TWIN ? 1 : 0;
    p.reduce_mode = c.reduction_mode;
    p.normal_iters = c.solver_iterations;
    p.friction_iters = c.solver_iterations;
    p.tol = c.residual_tolerance;
    p.warm_start = c.warm_start;
    p.solve_block = w.tuning.solve_block ? w.tuning.solve_block : DEFAULT_SOLVE_BLOCK;
    p.narrow_waves = w.tuning.narrow_waves ? w.tuning.narrow_waves : DEFAULT_NARROW_WAVES;
    // solver tiles
    if (n_scenes > 1) {
        p.tile_mode = 1;
        p.n_tiles = n_scenes;
        p.tiles_x = n_scenes;
        p.tiles_z = 1;
        p.tile_inv = 0.f;
    } else {
        p.tile_mode = 0;
        float T = c.tile_size > 0.f ? std::max(c.tile_size, cell) : cell;
        p.tile_inv = 1.0f / T;
        p.tile_origin_x = lo[0] - T;
        p.tile_origin_z = lo[2] - T;
        p.tiles_x = std::max(1, (int)std::ceil((hi[0] - lo[0]) / T) + 3);
        p.tiles_z = std::max(1, (int)std::ceil((hi[2] - lo[2]) / T) + 3);
        p.n_tiles = p.tiles_x * p.tiles_z;
    }
    w.tile_bits = std::max(1, ilog2((uint64_t)p.n_tiles));

    // ---- device buffers ----
    DeviceWorld &d = w.d;
    const int maxP = c.max_pairs, maxC = c.max_contacts, K = c.max_pairs_per_body;
    const size_t ncells = (size_t)1 << (w.key_bits - 1);
    RB_CHECK(w.alloc(&d.hulls, S));
    RB_CHECK(w.alloc(&d.verts, w.verts.size()));
    RB_CHECK(w.alloc(&d.planes, w.planes.size()));
    RB_CHECK(w.alloc(&d.edges, w.edges.size()));
    RB_CHECK(w.alloc(&d.face_index, w.face_index.size()));
    RB_CHECK(w.alloc(&d.indices, w.indices.size()));
    // compound children become proxy bodies n .. n+P-1 of the narrowphase
    std::vector<int> proxy_first(n, -1), owner;
    std::vector<int2> proxies;
    for (int i = 0; i < n; i++) {
        const HullInfo &h = w.hulls[w.shape[i]];
        if (shape_kind_of(h) != SHAPE_KIND_COMPOUND) continue;
        proxy_first[i] = (int)proxies.size();
        for (int c = 0; c < h.num_edges; c++) proxies.push_back(make_int2(i, h.edge_offset + c));
    }
    const int P = (int)proxies.size();
    p.n_proxies = P;
    for (int i = 0; i < n; i++) owner.push_back(i);
    for (const int2 &pr : proxies) owner.push_back(pr.x);
    RB_CHECK(w.alloc(&d.pos, (size_t)n + P));
    RB_CHECK(w.alloc(&d.quat, (size_t)n + P));
    RB_CHECK(w.alloc(&d.linvel, n));
    RB_CHECK(w.alloc(&d.angvel, n));
    RB_CHECK(w.alloc(&d.inv_mass, (size_t)n + P));
    RB_CHECK(w.alloc(&d.shape, (size_t)n + P));
    RB_CHECK(w.alloc(&d.owner, (size_t)n + P));
    RB_CHECK(w.alloc(&d.proxy_first, n));
    RB_CHECK(w.alloc(&d.proxy, std::max(P, 1)));
    RB_CHECK(w.alloc(&d.children, std::max((int)w.children.size(), 1)));
    RB_CHECK(w.alloc(&d.scene, n));
    RB_CHECK(w.alloc(&d.is_large, n));
    RB_CHECK(w.alloc(&d.inv_inertia_local, n));
    RB_CHECK(w.alloc(&d.inv_inertia_world, 3 * (size_t)n));
    RB_CHECK(w.alloc(&d.scene_start, n_scenes));
    RB_CHECK(w.alloc(&d.scene_count, n_scenes));
    RB_CHECK(w.alloc(&d.large_list, p.n_large + 1));
    RB_CHECK(w.alloc(&d.large_off, p.n_large + 1));
    RB_CHECK(w.alloc(&d.large_first, n_scenes));
    RB_CHECK(w.alloc(&d.large_num, n_scenes));
    RB_CHECK(w.alloc(&d.large_index, n));
    RB_CHECK(w.alloc(&d.aabb_min, n));
    RB_CHECK(w.alloc(&d.aabb_max, n));
    RB_CHECK(w.alloc(&d.kv, n));
    RB_CHECK(w.alloc(&d.cell_start, ncells));
    d.occ_words = (int)(ncells / 32 + 1);
    RB_CHECK(w.alloc(&d.occ, (size_t)d.occ_words));
    RB_CHECK(w.alloc(&d.smn, n));
    RB_CHECK(w.alloc(&d.smx, n));
    RB_CHECK(w.alloc(&d.body_slots, (size_t)n * K));
    RB_CHECK(w.alloc(&d.pair_count, n));
    RB_CHECK(w.alloc(&d.pair_off, n));
    RB_CHECK(w.alloc(&d.pairs, maxP));
    RB_CHECK(w.alloc(&d.contact_slots, maxP));
    RB_CHECK(w.alloc(&d.has_contact, maxP));
    RB_CHECK(w.alloc(&d.contact_off, maxP));
    RB_CHECK(w.alloc(&d.contacts, maxC));
    RB_CHECK(w.alloc(&d.cs.linear, maxC));
    RB_CHECK(w.alloc(&d.cs.pos, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.cs.center, maxC));
    RB_CHECK(w.alloc(&d.cs.jac_inv, maxC));
    RB_CHECK(w.alloc(&d.cs.b, maxC));
    RB_CHECK(w.alloc(&d.cs.lambda, maxC));
    RB_CHECK(w.alloc(&d.cs.fric, maxC));
    RB_CHECK(w.alloc(&d.cs.bodies, maxC));
    RB_CHECK(w.alloc(&d.cs.pair, maxC));
    RB_CHECK(w.alloc(&d.tile_kv, maxC));
    RB_CHECK(w.alloc(&d.tile_first, p.n_tiles));
    RB_CHECK(w.alloc(&d.tile_cnt, p.n_tiles));
    RB_CHECK(w.alloc(&d.solve_order, maxC));
    RB_CHECK(w.alloc(&d.row_color, maxC));
    RB_CHECK(w.alloc(&d.tile_color_off, (size_t)p.n_tiles * (RB_MAX_LOCAL_COLORS + 1)));
    RB_CHECK(w.alloc(&d.tile_ncolors, p.n_tiles));
    RB_CHECK(w.alloc(&d.tile_stats, p.n_tiles));
    RB_CHECK(w.alloc(&d.pk.lin0, maxC));
    RB_CHECK(w.alloc(&d.pk.na0, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.na1, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.nia0, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.nia1, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.nl, 2 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.p, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.cen, maxC));
    RB_CHECK(w.alloc(&d.pk.jac, maxC));
    RB_CHECK(w.alloc(&d.pk.b, maxC));
    RB_CHECK(w.alloc(&d.pk.lam, maxC));
    RB_CHECK(w.alloc(&d.pk.fric, maxC));
    RB_CHECK(w.alloc(&d.pk.IA, 3 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.IB, 3 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.pk.pA, maxC));
    RB_CHECK(w.alloc(&d.pk.pB, maxC));
    RB_CHECK(w.alloc(&d.pk.meta, maxC));
    RB_CHECK(w.alloc(&d.pk.im, maxC));
    RB_CHECK(w.alloc(&d.solve_pos, maxC));
    RB_CHECK(w.alloc(&d.tile_bodies, 2 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.tile_nb, p.n_tiles));
    RB_CHECK(w.alloc(&d.tile_hdr, p.n_tiles));
    RB_CHECK(w.alloc(&d.tile_order, p.n_tiles));
    RB_CHECK(w.alloc(&d.narrow_cand, maxP));
    RB_CHECK(w.alloc(&d.cand_off, maxP));
    RB_CHECK(w.alloc(&d.face_best, maxP));
    RB_CHECK(w.alloc(&d.cand_kv, maxP));
    RB_CHECK(w.alloc(&d.residual, 4 * RB_MAX_SWEEPS + 2));  // 2 floats per sweep of both phases
    RB_CHECK(w.alloc(&d.counters, RB_CNT_COUNT));
    RB_CHECK(w.alloc(&d.prev_keys, maxC));
    RB_CHECK(w.alloc(&d.prev_imp, maxC));
    RB_CHECK(w.alloc(&d.prev_pts, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.prev_count, 1));
    const size_t stage_n = p.n_proxies > 0 ? (size_t)maxC : 1;
    RB_CHECK(w.alloc(&d.stage_keys, stage_n));
    RB_CHECK(w.alloc(&d.stage_imp, stage_n));
    RB_CHECK(w.alloc(&d.stage_pts, 4 * stage_n));
    RB_CHECK(w.alloc(&d.features, 16 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.mlp_out, 4 * (size_t)maxC));
    RB_CHECK(w.alloc(&d.mlp_w, RB_MLP_NW));
    RB_CHECK(w.alloc(&d.mlp_b, RB_MLP_NB));
    if (!w.mlp_w_host.empty()) {
        RB_CHECK(hipMemcpy(d.mlp_w, w.mlp_w_host.data(), 2 * (size_t)RB_MLP_NW, hipMemcpyHostToDevice));
        RB_CHECK(hipMemcpy(d.mlp_b, w.mlp_b_host.data(), 4 * (size_t)RB_MLP_NB, hipMemcpyHostToDevice));
    }
    // large-body pair lists: each may pair with every earlier body of its scene
    std::vector<int> large_off(p.n_large + 1, 0), large_index(n, -1);
    std::vector<int> large_first(n_scenes, 0), large_num(n_scenes, 0);
    for (int q = 0; q < p.n_large; q++) {
        int x = w.large_list[q];
        large_index[x] = q;
        large_off[q + 1] = large_off[q] + (x - scene_start[w.scene[x]]);
        large_num[w.scene[x]]++;
    }
    for (int s = 0, run = 0; s < n_scenes; s++) { large_first[s] = run; run += large_num[s]; }
    RB_CHECK(w.alloc(&d.large_slots, (size_t)large_off[p.n_large] + 1));
    int max_scene = 0;
    for (int sc = 0; sc < n_scenes; sc++) max_scene = std::max(max_scene, scene_count[sc]);
    p.large_chunks = std::max(1, (max_scene + RB_LARGE_CHUNK - 1) / RB_LARGE_CHUNK);
    RB_CHECK(w.alloc(&d.large_chunk_cnt, (size_t)std::max(p.n_large, 1) * p.large_chunks));
    // sort / scan work
    int sort_cap = std::max(std::max(n, maxC), maxP);
    w.sw.max_n = sort_cap;
    w.sw.max_tiles = rb_sort_tiles(sort_cap);
    RB_CHECK(w.alloc(&w.sw.tmp, sort_cap));
    RB_CHECK(w.alloc(&w.sw.hist, 8 * 256));
    RB_CHECK(w.alloc(&w.sw.status, sort_status_words(sort_cap, 8)));
    RB_CHECK(w.alloc(&w.sw.tile_ctr, 16));
    w.sw.err = d.counters + RB_CNT_OVERFLOW;
    int scan_cap = std::max(n, maxP);
    w.scan_stride = rb_scan_status_words(scan_cap);
    RB_CHECK(w.alloc(&w.scan_status,
                     RB_SCAN_REGIONS * w.scan_stride + RB_STEP_SORTS * RB_SORT_HEAD_WORDS));
    for (int k = 0; k < RB_STEP_SORTS; k++) {
        w.sw_step[k] = w.sw;
        w.sw_step[k].hist = w.scan_status + RB_SCAN_REGIONS * w.scan_stride + k * RB_SORT_HEAD_WORDS;
        w.sw_step[k].tile_ctr = w.sw_step[k].hist + 8 * 256;
        w.sw_step[k].precleared = true;
    }

    // ---- uploads ----
    hipStream_t s = w.stream;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    RB_CHECK(up(d.hulls, w.hulls.data(), sizeof(HullInfo) * S));
    RB_CHECK(up(d.verts, w.verts.data(), sizeof(float4) * w.verts.size()));
    RB_CHECK(up(d.planes, w.planes.data(), sizeof(float4) * w.planes.size()));
    if (!w.edges.empty()) RB_CHECK(up(d.edges, w.edges.data(), sizeof(float4) * w.edges.size()));
    RB_CHECK(up(d.face_index, w.face_index.data(), sizeof(int2) * w.face_index.size()));
    RB_CHECK(up(d.indices, w.indices.data(), sizeof(int) * w.indices.size()));
    RB_CHECK(up(d.pos, w.pos.data(), sizeof(float4) * n));
    RB_CHECK(up(d.quat, w.quat.data(), sizeof(float4) * n));
    RB_CHECK(up(d.linvel, w.linvel.data(), sizeof(float4) * n));
    RB_CHECK(up(d.angvel, w.angvel.data(), sizeof(float4) * n));
    RB_CHECK(up(d.inv_mass, w.inv_mass.data(), sizeof(float) * n));
    RB_CHECK(up(d.shape, w.shape.data(), sizeof(int) * n));
    RB_CHECK(up(d.owner, owner.data(), sizeof(int) * owner.size()));
    RB_CHECK(up(d.proxy_first, proxy_first.data(), sizeof(int) * n));
    if (P > 0) {
        std::vector<float> pim(P);
        std::vector<int> psh(P);
        for (int k = 0; k < P; k++) {
            pim[k] = w.inv_mass[proxies[k].x];
            psh[k] = w.children[proxies[k].y].shape;
        }
        RB_CHECK(up(d.inv_mass + n, pim.data(), sizeof(float) * P));
        RB_CHECK(up(d.shape + n, psh.data(), sizeof(int) * P));
        RB_CHECK(up(d.proxy, proxies.data(), sizeof(int2) * P));
        RB_CHECK(up(d.children, w.children.data(), sizeof(ChildShape) * w.children.size()));
    }
    RB_CHECK(up(d.scene, w.scene.data(), sizeof(int) * n));
    RB_CHECK(up(d.is_large, w.is_large.data(), sizeof(int) * n));
    RB_CHECK(up(d.inv_inertia_local, w.inv_inertia_local.data(), sizeof(float4) * n));
    RB_CHECK(up(d.scene_start, scene_start.data(), sizeof(int) * n_scenes));
    RB_CHECK(up(d.scene_count, scene_count.data(), sizeof(int) * n_scenes));
    if (p.n_large) RB_CHECK(up(d.large_list, w.large_list.data(), sizeof(int) * p.n_large));
    RB_CHECK(up(d.large_off, large_off.data(), sizeof(int) * (p.n_large + 1)));
    RB_CHECK(up(d.large_first, large_first.data(), sizeof(int) * n_scenes));
    RB_CHECK(up(d.large_num, large_num.data(), sizeof(int) * n_scenes));
    RB_CHECK(up(d.large_index, large_index.data(), sizeof(int) * n));
    RB_CHECK(hipMemsetAsync(d.cell_start, 0xff, sizeof(uint32_t) * ncells, s));
    RB_CHECK(hipMemsetAsync(d.occ, 0, sizeof(uint32_t) * d.occ_words, s));
    RB_CHECK(hipMemsetAsync(d.counters, 0, sizeof(int) * RB_CNT_COUNT, s));
    RB_CHECK(hipMemsetAsync(d.prev_count, 0, sizeof(int), s));
    RB_CHECK(hipMemsetAsync(d.inv_inertia_world, 0, sizeof(float4) * 3 * n, s));
    RB_CHECK(hipMemsetAsync(d.pair_count, 0, sizeof(uint32_t) * n, s));
    RB_CHECK(hipMemsetAsync(d.has_contact, 0, sizeof(uint32_t) * maxP, s));
    RB_CHECK(hipStreamSynchronize(s));
    w.uploaded = true;
    return MI_RB_OK;
}

// Enqueue one step, stopping after `last_stage`.
mi_rb_status enqueue_step(World &w, int last_stage) {
    DeviceWorld &d = w.d;
    const StepParams &p = w.sp;
    hipStream_t s = w.stream;
    const bool ev = w.timing && !w.use_graph;
    (void)hipGetLastError();  // start from a clean error state: launches below are checked one by one
    w.events_full = ev && last_stage >= MI_RB_STAGE_INTEGRATE;
    if (ev) RB_CHECK(hipEventRecord(w.ev[0], s));
    RB_LAUNCH(k_zero_counters, dim3(16), dim3(256), 0, s, d.counters, d.residual,
                       2 * (p.normal_iters + p.friction_iters + 1), w.scan_status,
                       (int)(RB_SCAN_REGIONS * w.scan_stride + RB_STEP_SORTS * RB_SORT_HEAD_WORDS),
                       d.tile_first, d.tile_cnt,
                       p.n_tiles);
    RB_CHECK(launch_setup(d, p, s));
    if (ev) RB_CHECK(hipEventRecord(w.ev[1], s));
    if (last_stage < MI_RB_STAGE_SORT) return MI_RB_OK;
    RB_CHECK(launch_radix_sort(d.kv, p.n_bodies, w.key_bits,
                               w.tuning.sort_digit_bits ? w.tuning.sort_digit_bits : 8,
                               w.sw_step[RB_STEP_SORT_BODY], s,
                               nullptr, w.tuning.sort_tile));
    if (ev) RB_CHECK(hipEventRecord(w.ev[2], s));
    if (last_stage < MI_RB_STAGE_PAIRS) return MI_RB_OK;
    RB_CHECK(launch_cell_start(d, p, s));
    RB_CHECK(launch_find_pairs(d, p, s));
    RB_CHECK(launch_exclusive_scan(d.pair_count, d.pair_off, p.n_bodies,
                                   (uint32_t *)(d.counters + RB_CNT_SORT_TILE),
                                   w.scan_status + RB_SCAN_PAIRS * w.scan_stride, s, nullptr,
                                   d.counters + RB_CNT_OVERFLOW, nullptr, nullptr, true));
    RB_CHECK(launch_emit_pairs(d, p, s));
    if (p.n_proxies > 0)
        RB_CHECK(launch_compound_pairs(d, p, s, w.scan_status, w.scan_stride));
    if (ev) RB_CHECK(hipEventRecord(w.ev[3], s));
    if (last_stage < MI_RB_STAGE_NARROWPHASE) return MI_RB_OK;
    RB_CHECK(launch_narrowphase(d, p, s, w.sw_step[RB_STEP_SORT_CAND], w.scan_status,
                                w.scan_stride));
    if (ev) RB_CHECK(hipEventRecord(w.ev[4], s));
    if (last_stage < MI_RB_STAGE_CONSTRAINTS) return MI_RB_OK;
    RB_CHECK(launch_exclusive_scan(d.has_contact, d.contact_off, p.max_pairs,
                                   (uint32_t *)(d.counters + RB_CNT_CONTACT_TOTAL),
                                   w.scan_status + RB_SCAN_CONTACTS * w.scan_stride, s,
                                   d.counters + RB_CNT_PAIRS, d.counters + RB_CNT_OVERFLOW,
                                   nullptr, nullptr, true));
    RB_CHECK(launch_constraints(d, p, s));
    if (p.warm_start == MI_RB_WARM_LEARNED) RB_CHECK(launch_mlp_warm(d, p, s));
    if (ev) RB_CHECK(hipEventRecord(w.ev[5], s));
    if (last_stage < MI_RB_STAGE_COLOR) return MI_RB_OK;
    RB_CHECK(launch_radix_sort(d.tile_kv, p.max_contacts, w.tile_bits, 8,
                               w.sw_step[RB_STEP_SORT_TILE], s,
                               d.counters + RB_CNT_CONTACTS));
    RB_CHECK(launch_tile_bounds(d, p, s));
    RB_CHECK(launch_tiles(d, p, s));
    RB_CHECK(launch_tile_order(d, p, s, w.sw_step[RB_STEP_SORT_ORDER]));
    if (ev) RB_CHECK(hipEventRecord(w.ev[6], s));
    if (last_stage < MI_RB_STAGE_SOLVE) return MI_RB_OK;
    if (ev) {
        const size_t need = 2 * (size_t)4 * (p.normal_iters + p.friction_iters + 1) + 2;
        while (w.kev.size() < need) {
            hipEvent_t e = nullptr;
            RB_CHECK(hipEventCreate(&e));
            w.kev.push_back(e);
        }
    }
    // (eager steps only: events recorded inside a capture do not time replays)
    const bool span = ev && w.solve_timing && w.span[0] && w.span[1];
    w.kev_launches = 0;
    // (per-launch events and the span are not recorded together: the span
    // serves the captured steps, the per-launch events the eager ones)
    RB_CHECK(launch_solve(d, p, s, ev && !span ? w.kev.data() : nullptr, &w.kev_launches,
                          span ? w.span : nullptr));
    w.span_launches = span ? w.kev_launches : 0;
    RB_CHECK(launch_cache_order(d, p, s, w.sw));
    if (ev) RB_CHECK(hipEventRecord(w.ev[7], s));
    if (last_stage < MI_RB_STAGE_INTEGRATE) return MI_RB_OK;
    RB_CHECK(launch_integrate(d, p, s));
    if (ev) RB_CHECK(hipEventRecord(w.ev[8], s));
    return MI_RB_OK;
}

}  // namespace

// ================================================================ C ABI
extern "C" void mi_rb_config_reference(mi_rb_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_bodies = 1 << 17;
    c->max_shapes = 8192;
    c->max_hull_vertices = 1 << 20;
    c->max_pairs = 1 << 21;
    c->max_contacts = 1 << 20;
    c->max_pairs_per_body = 31;
    c->dt = 0.0166666f;           // integrateKernel.cl:114, broadphaseKernel.cl:172
    c->solver_dt = 1.0f / 60.0f;  // btGpuNarrowPhaseAndSolver.cpp:1082
    c->gravity[1] = -9.8f;
    c->solver_iterations = 4;     // btGpuNarrowPhaseAndSolver.cpp:1330
    c->friction = 0.7f;
    c->baumgarte = 0.2f;
    c->position_drift = 0.005f;
    c->contact_threshold = 0.02f;
    c->angular_damping = 0.99f;
    c->max_angular_motion = 0.25f * 3.14159254f;
    c->warm_start = MI_RB_WARM_ZERO;
    c->inertia_mode = MI_RB_INERTIA_REFERENCE;
}

extern "C" mi_rb_status mi_rb_create(const mi_rb_config *cfg, void **out_world) {
    if (!cfg || !out_world) return MI_RB_INVALID_ARG;
    if (cfg->max_bodies <= 0 || cfg->max_pairs <= 0 || cfg->max_contacts <= 0 ||
        cfg->max_pairs_per_body <= 0 || cfg->max_pairs_per_body > 64 || cfg->dt <= 0.f ||
        cfg->solver_iterations < 0 || cfg->solver_iterations > RB_MAX_SWEEPS || cfg->sat_mode < 0 ||
        cfg->sat_mode > 1 || cfg->reduction_mode < 0 || cfg->reduction_mode > 1)
        return MI_RB_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device) return MI_RB_DEVICE;
    World *w = new World();
    w->cfg = *cfg;
    w->device = cfg->device;
    if (hipSetDevice(w->device) != hipSuccess || hipStreamCreate(&w->stream) != hipSuccess) {
        delete w;
        return MI_RB_DEVICE;
    }
    for (auto &e : w->ev) hipEventCreate(&e);
    hipEventCreateWithFlags(&w->ev_order, hipEventDisableTiming);
    *out_world = w;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_destroy(void *world) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    hipSetDevice(w->device);
    if (w->stream) hipStreamSynchronize(w->stream);
    w->free_device();
    for (auto &e : w->ev) if (e) hipEventDestroy(e);
    for (auto &e : w->kev) if (e) hipEventDestroy(e);
    for (auto &e : w->span) if (e) hipEventDestroy(e);
    if (w->ev_order) hipEventDestroy(w->ev_order);
    if (w->stream) hipStreamDestroy(w->stream);
    delete w;
    return MI_RB_OK;
}

// Appends one hull's features to the shape arrays (registerConvexHullShape,
// btGpuNarrowPhaseAndSolver.cpp:423-490).
static mi_rb_status add_hull(World *w, const HullData &h, int *out_shape) {
    int nv = (int)h.verts.size() / 3, nf = (int)h.planes.size() / 4;
    int ne = (int)h.edges.size() / 3, ni = (int)h.indices.size();
    if (nv > RB_MAX_HULL_VERTS || ne > RB_MAX_HULL_EDGES || nf > 2 * RB_MAX_HULL_FACES)
        return MI_RB_CAPACITY;
    for (int f = 0; f < nf; f++)
        if (h.face_index[2 * f + 1] > RB_MAX_FACE_VERTS) return MI_RB_CAPACITY;
    if ((int)w->verts.size() + nv > w->cfg.max_hull_vertices) return MI_RB_CAPACITY;
    HullInfo hi;
    hi.face_offset = (int)w->planes.size();
    hi.num_faces = nf;
    hi.vertex_offset = (int)w->verts.size();
    hi.num_vertices = nv;
    hi.edge_offset = (int)w->edges.size();
    hi.num_edges = ne;
    hi.index_offset = (int)w->indices.size();
    hi.num_indices = ni;
    // w: radius of a ball around the centre inside the hull (the smallest
    // centre-to-plane distance over the hull triangles and the merged faces,
    // shrunk by 1e-4 relative), a lower bound the narrowphase uses to skip
    // axes
    double rho = h.inradius;
    for (int i = 0; i < nf; i++) {
        const float *pl = &h.planes[4 * i];
        double d = -((double)pl[0] * h.center[0] + (double)pl[1] * h.center[1] +
                     (double)pl[2] * h.center[2] + (double)pl[3]);
        rho = std::min(rho, d);
    }
    rho = nf > 0 && rho > 0.0 ? rho * (1.0 - 1e-4) : 0.0;
    hi.center = make_float4(h.center[0], h.center[1], h.center[2], (float)rho);
    hi.aabb_min = make_float4(h.aabb_min[0], h.aabb_min[1], h.aabb_min[2], 0.f);
    hi.aabb_max = make_float4(h.aabb_max[0], h.aabb_max[1], h.aabb_max[2], 0.f);
    for (int i = 0; i < nv; i++)
        w->verts.push_back(make_float4(h.verts[3 * i], h.verts[3 * i + 1], h.verts[3 * i + 2], 0.f));
    for (int i = 0; i < nf; i++) {
        w->planes.push_back(make_float4(h.planes[4 * i], h.planes[4 * i + 1], h.planes[4 * i + 2],
                                        h.planes[4 * i + 3]));
        w->face_index.push_back(make_int2(hi.index_offset + h.face_index[2 * i], h.face_index[2 * i + 1]));
    }
    for (int v : h.indices) w->indices.push_back(v);
    for (int i = 0; i < ne; i++)
        w->edges.push_back(make_float4(h.edges[3 * i], h.edges[3 * i + 1], h.edges[3 * i + 2], 0.f));
    float hx = (hi.aabb_max.x - hi.aabb_min.x) * 0.5f, hy = (hi.aabb_max.y - hi.aabb_min.y) * 0.5f,
          hz = (hi.aabb_max.z - hi.aabb_min.z) * 0.5f;
    w->hull_hext_norm.push_back(std::sqrt(hx * hx + hy * hy + hz * hz));
    w->hulls.push_back(hi);
    *out_shape = (int)w->hulls.size() - 1;
    w->uploaded = false;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_register_convex_hull(void *world, const float *xyz, int stride_bytes,
                                                   int n, const float scale[3], int *out_shape) {
    World *w = W(world);
    if (!w || !xyz || n < 4 || !out_shape) return MI_RB_INVALID_ARG;
    if ((int)w->hulls.size() >= w->cfg.max_shapes) return MI_RB_CAPACITY;
    HullData h;
    int st = rb_build_hull(xyz, stride_bytes, n, scale, h);
    if (st != MI_RB_OK) return (mi_rb_status)st;
    return add_hull(w, h, out_shape);
}

extern "C" mi_rb_status mi_rb_register_convex_polyhedron(
    void *world, const float *verts, int vert_stride_bytes, int num_vertices, const float *planes4,
    const int *face_index2, int num_faces, const int *indices, int num_indices, const float *edges,
    int edge_stride_bytes, int num_edges, int *out_shape) {
    World *w = W(world);
    if (!w || !out_shape) return MI_RB_INVALID_ARG;
    if ((int)w->hulls.size() >= w->cfg.max_shapes) return MI_RB_CAPACITY;
    HullData h;
    int st = rb_hull_from_features(verts, vert_stride_bytes, num_vertices, planes4, face_index2,
                                   num_faces, indices, num_indices, edges, edge_stride_bytes,
                                   num_edges, h);
    if (st != MI_RB_OK) return (mi_rb_status)st;
    return add_hull(w, h, out_shape);
}

// A shape without vertices (a sphere or a plane): no faces for a sphere, the
// plane's equation as the one face of a plane; the local AABB is the
// reference's (CLPhysicsDemo.cpp:420-489).
static mi_rb_status add_vertexless_shape(World *w, float radius, const float *plane,
                                         int *out_shape) {
    if ((int)w->hulls.size() >= w->cfg.max_shapes) return MI_RB_CAPACITY;
    HullInfo hi{};
    hi.face_offset = (int)w->planes.size();
    hi.num_faces = plane ? 1 : 0;
    hi.vertex_offset = (int)w->verts.size();
    hi.edge_offset = (int)w->edges.size();
    hi.index_offset = (int)w->indices.size();
    const float e = plane ? 1e30f : radius;
    hi.center = make_float4(0.f, 0.f, 0.f, plane ? 0.f : radius);
    hi.aabb_min = make_float4(-e, -e, -e, 0.f);
    hi.aabb_max = make_float4(e, e, e, 0.f);
    if (plane) {
        w->planes.push_back(make_float4(plane[0], plane[1], plane[2], plane[3]));
        w->face_index.push_back(make_int2(hi.index_offset, 0));
    }
    w->hull_hext_norm.push_back(std::sqrt(3.f) * e);
    w->hulls.push_back(hi);
    *out_shape = (int)w->hulls.size() - 1;
    w->uploaded = false;
    return MI_RB_OK;
}

// A compound of n hull children; a mesh compound (num_indices 1) marks its
// children as triangles to be culled per pair by their world AABBs.
static mi_rb_status add_compound(World *w, int n, const int *child_shapes, const float *child_pos,
                                 const float *child_orn, bool mesh, int *out_shape) {
    if ((int)w->hulls.size() >= w->cfg.max_shapes) return MI_RB_CAPACITY;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    std::vector<ChildShape> kids(n);
    for (int i = 0; i < n; i++) {
        const int cs = child_shapes[i];
        if (cs < 0 || cs >= (int)w->hulls.size() || shape_kind_of(w->hulls[cs]) != SHAPE_KIND_HULL)
            return MI_RB_INVALID_ARG;
        const float *q = child_orn + 4 * i;
        const float qn = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        if (!(qn > 0.5f && qn < 2.f)) return MI_RB_INVALID_ARG;
        ChildShape &k = kids[i];
        k.pos = make_float4(child_pos[3 * i], child_pos[3 * i + 1], child_pos[3 * i + 2], 0.f);
        k.orn = make_float4(q[0], q[1], q[2], q[3]);
        k.shape = cs;
        k.pad[0] = k.pad[1] = k.pad[2] = 0;
        // the child's local AABB, rotated into the compound frame
        const HullInfo &h = w->hulls[cs];
        const float c[3] = {(h.aabb_max.x + h.aabb_min.x) * 0.5f, (h.aabb_max.y + h.aabb_min.y) * 0.5f,
                            (h.aabb_max.z + h.aabb_min.z) * 0.5f};
        const float e[3] = {(h.aabb_max.x - h.aabb_min.x) * 0.5f, (h.aabb_max.y - h.aabb_min.y) * 0.5f,
                            (h.aabb_max.z - h.aabb_min.z) * 0.5f};
        const float x = q[0], y = q[1], z = q[2], s0 = q[3];
        const float m[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - s0 * z), 2 * (x * z + s0 * y)},
                               {2 * (x * y + s0 * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s0 * x)},
                               {2 * (x * z - s0 * y), 2 * (y * z + s0 * x), 1 - 2 * (x * x + y * y)}};
        for (int r = 0; r < 3; r++) {
            const float cc = m[r][0] * c[0] + m[r][1] * c[1] + m[r][2] * c[2] + child_pos[3 * i + r];
            const float ee = std::fabs(m[r][0]) * e[0] + std::fabs(m[r][1]) * e[1] +
                             std::fabs(m[r][2]) * e[2];
            lo[r] = std::min(lo[r], cc - ee);
            hi[r] = std::max(hi[r], cc + ee);
        }
    }
    HullInfo hi_{};
    hi_.face_offset = (int)w->planes.size();
    hi_.vertex_offset = (int)w->verts.size();
    hi_.index_offset = (int)w->indices.size();
    hi_.edge_offset = (int)w->children.size();
    hi_.num_edges = n;
    hi_.num_indices = mesh ? 1 : 0;
    hi_.aabb_min = make_float4(lo[0], lo[1], lo[2], 0.f);
    hi_.aabb_max = make_float4(hi[0], hi[1], hi[2], 0.f);
    for (const ChildShape &k : kids) w->children.push_back(k);
    const float hx = (hi[0] - lo[0]) * 0.5f, hy = (hi[1] - lo[1]) * 0.5f, hz = (hi[2] - lo[2]) * 0.5f;
    w->hull_hext_norm.push_back(std::sqrt(hx * hx + hy * hy + hz * hz));
    w->hulls.push_back(hi_);
    *out_shape = (int)w->hulls.size() - 1;
    w->uploaded = false;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_register_compound(void *world, int n, const int *child_shapes,
                                                const float *child_pos, const float *child_orn,
                                                int *out_shape) {
    World *w = W(world);
    if (!w || n < 1 || n > 64 || !child_shapes || !child_pos || !child_orn || !out_shape)
        return MI_RB_INVALID_ARG;
    return add_compound(w, n, child_shapes, child_pos, child_orn, false, out_shape);
}

// One triangle of a concave mesh as a hull of 3 vertices, 5 faces and 3
// unique edges, the polyhedron sat.cl:1040-1116 builds per triangle: the
// front face (the triangle's plane, n = (v1-v0) x (v2-v0) normalised,
// c = -n.v0), the back face (-n, n.v0), and per edge (v_i, v_prev) a face of
// 2 indices whose plane holds the edge and is perpendicular to the triangle
// (normal (n x (v_prev - v_i)) normalised, c = -normal.v_i); unique edges
// v1-v0, v2-v1, v0-v2.  Returns false for a triangle without area.
static bool add_triangle_hull(World *w, const float v[3][3]) {
    auto sub = [](const float *a, const float *b, float *r) {
        r[0] = a[0] - b[0], r[1] = a[1] - b[1], r[2] = a[2] - b[2];
    };
    auto cross = [](const float *a, const float *b, float *r) {
        r[0] = a[1] * b[2] - a[2] * b[1], r[1] = a[2] * b[0] - a[0] * b[2], r[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto normalize = [&](float *a) {
        const float l = std::sqrt(dot(a, a));
        if (!(l > 0.f) || !std::isfinite(l)) return false;
        a[0] /= l, a[1] /= l, a[2] /= l;
        return true;
    };
    float e01[3], e02[3], nrm[3];
    sub(v[1], v[0], e01);
    sub(v[2], v[0], e02);
    cross(e01, e02, nrm);
    if (!normalize(nrm)) return false;
    float side[3][4];
    for (int i = 0, prev = 2; i < 3; prev = i++) {
        float d[3];
        sub(v[prev], v[i], d);
        cross(nrm, d, side[i]);
        if (!normalize(side[i])) return false;
        side[i][3] = -dot(side[i], v[i]);
    }
    HullInfo hi{};
    hi.face_offset = (int)w->planes.size();
    hi.num_faces = 5;
    hi.vertex_offset = (int)w->verts.size();
    hi.num_vertices = 3;
    hi.edge_offset = (int)w->edges.size();
    hi.num_edges = 3;
    hi.index_offset = (int)w->indices.size();
    hi.num_indices = 3 + 3 + 2 * 3;
    float lo[3], hi3[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = std::min(v[0][k], std::min(v[1][k], v[2][k]));
        hi3[k] = std::max(v[0][k], std::max(v[1][k], v[2][k]));
    }
    // zero thickness: no inner ball (w = 0)
    hi.center = make_float4((v[0][0] + v[1][0] + v[2][0]) * (1.f / 3.f),
                            (v[0][1] + v[1][1] + v[2][1]) * (1.f / 3.f),
                            (v[0][2] + v[1][2] + v[2][2]) * (1.f / 3.f), 0.f);
    hi.aabb_min = make_float4(lo[0], lo[1], lo[2], 0.f);
    hi.aabb_max = make_float4(hi3[0], hi3[1], hi3[2], 0.f);
    for (int k = 0; k < 3; k++) w->verts.push_back(make_float4(v[k][0], v[k][1], v[k][2], 0.f));
    const int io = hi.index_offset;
    w->planes.push_back(make_float4(nrm[0], nrm[1], nrm[2], -dot(nrm, v[0])));
    w->face_index.push_back(make_int2(io, 3));
    w->planes.push_back(make_float4(-nrm[0], -nrm[1], -nrm[2], dot(nrm, v[0])));
    w->face_index.push_back(make_int2(io + 3, 3));
    for (int x : {0, 1, 2, 2, 1, 0}) w->indices.push_back(x);
    for (int i = 0, prev = 2; i < 3; prev = i++) {
        w->planes.push_back(make_float4(side[i][0], side[i][1], side[i][2], side[i][3]));
        w->face_index.push_back(make_int2(io + 6 + 2 * i, 2));
        w->indices.push_back(i);
        w->indices.push_back(prev);
    }
    for (int i = 0; i < 3; i++) {
        float e[3];
        sub(v[(i + 1) % 3], v[i], e);
        w->edges.push_back(make_float4(e[0], e[1], e[2], 0.f));
    }
    const float hx = (hi3[0] - lo[0]) * 0.5f, hy = (hi3[1] - lo[1]) * 0.5f, hz = (hi3[2] - lo[2]) * 0.5f;
    w->hull_hext_norm.push_back(std::sqrt(hx * hx + hy * hy + hz * hz));
    w->hulls.push_back(hi);
    return true;
}

extern "C" mi_rb_status mi_rb_register_concave_mesh(void *world, const float *xyz, int stride_bytes,
                                                    int num_vertices, const int *indices,
                                                    int num_triangles, const float scale[3],
                                                    int *out_shape) {
    World *w = W(world);
    if (!w || !xyz || !indices || !out_shape || num_vertices < 3 || num_triangles < 1 ||
        num_triangles > MI_RB_MAX_MESH_TRIANGLES || stride_bytes < 3 * (int)sizeof(float))
        return MI_RB_INVALID_ARG;
    for (int i = 0; i < 3 * num_triangles; i++)
        if (indices[i] < 0 || indices[i] >= num_vertices) return MI_RB_INVALID_ARG;
    // room for every triangle and the mesh shape, checked before anything is added
    if ((long long)w->hulls.size() + num_triangles + 1 > w->cfg.max_shapes ||
        (long long)w->verts.size() + 3LL * num_triangles > w->cfg.max_hull_vertices)
        return MI_RB_CAPACITY;
    const float sx = scale ? scale[0] : 1.f, sy = scale ? scale[1] : 1.f, sz = scale ? scale[2] : 1.f;
    auto vert = [&](int i, float *r) {
        const float *p = (const float *)((const char *)xyz + (size_t)i * stride_bytes);
        r[0] = p[0] * sx, r[1] = p[1] * sy, r[2] = p[2] * sz;
    };
    std::vector<int> tri_shapes;
    for (int t = 0; t < num_triangles; t++) {
        float v[3][3];
        for (int k = 0; k < 3; k++) vert(indices[3 * t + k], v[k]);
        // a triangle without area has no plane and is left out
        if (add_triangle_hull(w, v)) tri_shapes.push_back((int)w->hulls.size() - 1);
    }
    if (tri_shapes.empty()) return MI_RB_INVALID_ARG;
    const int m = (int)tri_shapes.size();
    std::vector<float> zero(3 * (size_t)m, 0.f), ident(4 * (size_t)m, 0.f);
    for (int i = 0; i < m; i++) ident[4 * i + 3] = 1.f;
    return add_compound(w, m, tri_shapes.data(), zero.data(), ident.data(), true, out_shape);
}

extern "C" mi_rb_status mi_rb_register_sphere(void *world, float radius, int *out_shape) {
    World *w = W(world);
    if (!w || !out_shape || !(radius > 0.f) || !(radius < 1e18f)) return MI_RB_INVALID_ARG;
    return add_vertexless_shape(w, radius, nullptr, out_shape);
}

extern "C" mi_rb_status mi_rb_register_plane(void *world, const float normal[3], float constant,
                                             int *out_shape) {
    World *w = W(world);
    if (!w || !normal || !out_shape || !std::isfinite(constant)) return MI_RB_INVALID_ARG;
    const float len2 = normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2];
    if (!(len2 > 0.f) || !std::isfinite(len2)) return MI_RB_INVALID_ARG;
    const float eq[4] = {normal[0], normal[1], normal[2], constant};
    return add_vertexless_shape(w, 0.f, eq, out_shape);
}

extern "C" mi_rb_status mi_rb_register_unsupported_shape(void *world, int kind) {
    (void)kind;
    return world ? MI_RB_UNSUPPORTED_SHAPE : MI_RB_INVALID_ARG;
}

extern "C" mi_rb_status mi_rb_register_body(void *world, float mass, const float pos[3],
                                            const float quat_xyzw[4], int shape, int scene,
                                            int *out_body) {
    World *w = W(world);
    if (!w || !pos || !quat_xyzw || !out_body || mass < 0.f) return MI_RB_INVALID_ARG;
    if (shape < 0 || shape >= (int)w->hulls.size() || scene < 0) return MI_RB_INVALID_ARG;
    // a plane is infinite: it can only be static
    if (mass != 0.f && w->hulls[shape].num_vertices == 0 && w->hulls[shape].num_faces > 0)
        return MI_RB_INVALID_ARG;
    // so is a concave mesh (the reference gives meshes no inertia)
    if (mass != 0.f && compound_is_mesh(w->hulls[shape])) return MI_RB_INVALID_ARG;
    if (!w->scene.empty() && scene < w->scene.back()) return MI_RB_INVALID_ARG;
    if ((int)w->pos.size() >= w->cfg.max_bodies) return MI_RB_CAPACITY;
    w->pos.push_back(make_float4(pos[0], pos[1], pos[2], 0.f));
    w->quat.push_back(make_float4(quat_xyzw[0], quat_xyzw[1], quat_xyzw[2], quat_xyzw[3]));
    w->linvel.push_back(make_float4(0, 0, 0, 0));
    w->angvel.push_back(make_float4(0, 0, 0, 0));
    w->inv_mass.push_back(mass != 0.f ? 1.f / mass : 0.f);
    w->shape.push_back(shape);
    w->scene.push_back(scene);
    w->inv_inertia_local.push_back(local_inv_inertia(*w, mass, shape, pos, quat_xyzw));
    *out_body = (int)w->pos.size() - 1;
    w->uploaded = false;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_upload(void *world) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    hipSetDevice(w->device);
    return derive_and_upload(*w);
}

extern "C" mi_rb_status mi_rb_set_transform(void *world, int body, const float pos[3],
                                            const float quat[4]) {
    World *w = W(world);
    if (!w || !pos || !quat || body < 0 || body >= (int)w->pos.size()) return MI_RB_INVALID_ARG;
    w->pos[body] = make_float4(pos[0], pos[1], pos[2], 0.f);
    w->quat[body] = make_float4(quat[0], quat[1], quat[2], quat[3]);
    if (w->uploaded) {
        RB_CHECK(hipMemcpyAsync(w->d.pos + body, &w->pos[body], sizeof(float4), hipMemcpyHostToDevice, w->stream));
        RB_CHECK(hipMemcpyAsync(w->d.quat + body, &w->quat[body], sizeof(float4), hipMemcpyHostToDevice, w->stream));
        RB_CHECK(hipStreamSynchronize(w->stream));
    }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_set_linvel(void *world, int body, const float v[3]) {
    World *w = W(world);
    if (!w || !v || body < 0 || body >= (int)w->pos.size()) return MI_RB_INVALID_ARG;
    w->linvel[body] = make_float4(v[0], v[1], v[2], 0.f);
    if (w->uploaded) {
        RB_CHECK(hipMemcpyAsync(w->d.linvel + body, &w->linvel[body], sizeof(float4), hipMemcpyHostToDevice, w->stream));
        RB_CHECK(hipStreamSynchronize(w->stream));
    }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_set_body_state(void *world, const float *pos4, const float *quat4,
                                             const float *linvel4, const float *angvel4) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    size_t n = w->pos.size();
    const float *src[4] = {pos4, quat4, linvel4, angvel4};
    std::vector<float4> *dst[4] = {&w->pos, &w->quat, &w->linvel, &w->angvel};
    float4 *dev[4] = {w->d.pos, w->d.quat, w->d.linvel, w->d.angvel};
    for (int f = 0; f < 4; f++) {
        if (!src[f]) continue;
        std::memcpy(dst[f]->data(), src[f], sizeof(float4) * n);
        if (w->uploaded)
            RB_CHECK(hipMemcpyAsync(dev[f], src[f], sizeof(float4) * n, hipMemcpyHostToDevice, w->stream));
    }
    if (w->uploaded) RB_CHECK(hipStreamSynchronize(w->stream));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_get_body_state(void *world, float *pos4, float *quat4,
                                             float *linvel4, float *angvel4) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    size_t n = w->pos.size();
    float *dst[4] = {pos4, quat4, linvel4, angvel4};
    const std::vector<float4> *host[4] = {&w->pos, &w->quat, &w->linvel, &w->angvel};
    float4 *dev[4] = {w->d.pos, w->d.quat, w->d.linvel, w->d.angvel};
    for (int f = 0; f < 4; f++) {
        if (!dst[f]) continue;
        if (w->uploaded)
            RB_CHECK(hipMemcpyAsync(dst[f], dev[f], sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream));
        else
            std::memcpy(dst[f], host[f]->data(), sizeof(float4) * n);
    }
    if (w->uploaded) RB_CHECK(hipStreamSynchronize(w->stream));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_use_graph(void *world, int enable) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    if (w->graph) { hipGraphExecDestroy(w->graph); w->graph = nullptr; }
    w->use_graph = enable != 0;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_set_solve_timing(void *world, int enable) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    if (w->graph) { hipGraphExecDestroy(w->graph); w->graph = nullptr; }
    w->solve_timing = enable != 0;
    hipSetDevice(w->device);
    for (auto &e : w->span)
        if (!e && hipEventCreate(&e) != hipSuccess) return MI_RB_DEVICE;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_step(void *world, int nsteps) {
    World *w = W(world);
    if (!w || nsteps < 0) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    if (w->sp.warm_start == MI_RB_WARM_LEARNED && w->mlp_w_host.empty()) return MI_RB_STATE;
    hipSetDevice(w->device);

    if (w->use_graph && !w->graph) {
        hipGraph_t g;
        RB_CHECK(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
        mi_rb_status st = enqueue_step(*w, MI_RB_STAGE_INTEGRATE);
        hipError_t e = hipStreamEndCapture(w->stream, &g);
        if (st != MI_RB_OK) return st;
        RB_CHECK(e);
        RB_CHECK(hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0));
        hipGraphDestroy(g);
    }
    RB_CHECK(hipEventRecord(w->ev[9], w->stream));
    for (int k = 0; k < nsteps; k++) {
        if (w->use_graph) {
            RB_CHECK(hipGraphLaunch(w->graph, w->stream));
        } else {
            mi_rb_status st = enqueue_step(*w, MI_RB_STAGE_INTEGRATE);
            if (st != MI_RB_OK) return st;
        }
        w->steps++;
    }
    RB_CHECK(hipEventRecord(w->ev[8], w->stream));
    RB_CHECK(hipEventSynchronize(w->ev[8]));
    float ms = 0.f;
    hipEventElapsedTime(&ms, w->ev[9], w->ev[8]);
    w->last_ms = ms;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_step_partial(void *world, int stage) {
    World *w = W(world);
    if (!w || stage < 0 || stage > MI_RB_STAGE_INTEGRATE) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    if (w->sp.warm_start == MI_RB_WARM_LEARNED && w->mlp_w_host.empty()) return MI_RB_STATE;
    hipSetDevice(w->device);
    mi_rb_status st = enqueue_step(*w, stage);
    if (st != MI_RB_OK) return st;
    RB_CHECK(hipStreamSynchronize(w->stream));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_set_tuning(void *world, const mi_rb_tuning *t) {
    World *w = W(world);
    if (!w || !t) return MI_RB_INVALID_ARG;
    if (!(t->solve_block == 0 || t->solve_block == 256 || t->solve_block == 512)) return MI_RB_INVALID_ARG;
    if (t->narrow_waves < 0 || t->narrow_waves > 8) return MI_RB_INVALID_ARG;
    if (!(t->sort_tile == 0 || t->sort_tile == 1024 || t->sort_tile == 4096)) return MI_RB_INVALID_ARG;
    if (!(t->sort_digit_bits == 0 || t->sort_digit_bits == 4 || t->sort_digit_bits == 6 ||
          t->sort_digit_bits == 7 || t->sort_digit_bits == 8))
        return MI_RB_INVALID_ARG;
    w->tuning = *t;
    w->sp.solve_block = t->solve_block ? t->solve_block : DEFAULT_SOLVE_BLOCK;
    w->sp.narrow_waves = t->narrow_waves ? t->narrow_waves : DEFAULT_NARROW_WAVES;
    if (w->graph) { hipGraphExecDestroy(w->graph); w->graph = nullptr; }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_set_solver(void *world, int iterations, float tolerance,
                                         int warm_start) {
    World *w = W(world);
    if (!w || iterations < 0 || iterations > RB_MAX_SWEEPS || !(tolerance >= 0.f))
        return MI_RB_INVALID_ARG;
    if (warm_start < MI_RB_WARM_ZERO || warm_start > MI_RB_WARM_LEARNED) return MI_RB_INVALID_ARG;
    w->cfg.solver_iterations = iterations;
    w->cfg.residual_tolerance = tolerance;
    w->cfg.warm_start = warm_start;
    w->sp.normal_iters = iterations;
    w->sp.friction_iters = iterations;
    w->sp.tol = tolerance;
    w->sp.warm_start = warm_start;
    if (w->graph) { hipGraphExecDestroy(w->graph); w->graph = nullptr; }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_get_tuning(void *world, mi_rb_tuning *t) {
    World *w = W(world);
    if (!w || !t) return MI_RB_INVALID_ARG;
    t->solve_block = w->tuning.solve_block ? w->tuning.solve_block : DEFAULT_SOLVE_BLOCK;
    t->narrow_waves = w->tuning.narrow_waves ? w->tuning.narrow_waves : DEFAULT_NARROW_WAVES;
    t->sort_tile = w->tuning.sort_tile;
    t->sort_digit_bits = w->tuning.sort_digit_bits;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_autotune(void *world, int steps, mi_rb_tuning *best, float ms_out[8]) {
    World *w = W(world);
    if (!w || steps <= 0 || !best) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    hipSetDevice(w->device);
    DeviceWorld &d = w->d;
    const size_t n = w->pos.size(), maxC = (size_t)w->cfg.max_contacts;
    // snapshot of everything a step reads that a step changes
    std::vector<float4> st(4 * n), pk(maxC), pi(maxC), pp(4 * maxC);
    int pc = 0;
    float4 *dev[4] = {d.pos, d.quat, d.linvel, d.angvel};
    RB_CHECK(hipStreamSynchronize(w->stream));
    for (int f = 0; f < 4; f++)
        RB_CHECK(hipMemcpy(st.data() + f * n, dev[f], 16 * n, hipMemcpyDeviceToHost));
    RB_CHECK(hipMemcpy(pk.data(), d.prev_keys, 8 * maxC, hipMemcpyDeviceToHost));
    RB_CHECK(hipMemcpy(pi.data(), d.prev_imp, 16 * maxC, hipMemcpyDeviceToHost));
    RB_CHECK(hipMemcpy(pp.data(), d.prev_pts, 64 * maxC, hipMemcpyDeviceToHost));
    RB_CHECK(hipMemcpy(&pc, d.prev_count, sizeof(int), hipMemcpyDeviceToHost));
    auto restore = [&]() -> mi_rb_status {
        for (int f = 0; f < 4; f++)
            RB_CHECK(hipMemcpy(dev[f], st.data() + f * n, 16 * n, hipMemcpyHostToDevice));
        RB_CHECK(hipMemcpy(d.prev_keys, pk.data(), 8 * maxC, hipMemcpyHostToDevice));
        RB_CHECK(hipMemcpy(d.prev_imp, pi.data(), 16 * maxC, hipMemcpyHostToDevice));
        RB_CHECK(hipMemcpy(d.prev_pts, pp.data(), 64 * maxC, hipMemcpyHostToDevice));
        RB_CHECK(hipMemcpy(d.prev_count, &pc, sizeof(int), hipMemcpyHostToDevice));
        return MI_RB_OK;
    };
    const mi_rb_tuning keep = w->tuning;
    const bool keep_graph = w->use_graph;
    const long long keep_steps = w->steps;
    const int blocks[2] = {256, 512}, waves[4] = {2, 4, 6, 8};
    float best_ms = 0.f;
    mi_rb_tuning bt = keep;
    // one trial: the step time of tuning t from the saved state
    auto trial = [&](const mi_rb_tuning &t, float &ms) -> mi_rb_status {
        mi_rb_status stt = mi_rb_set_tuning(world, &t);
        if (stt == MI_RB_OK) stt = restore();
        w->use_graph = true;
        if (stt == MI_RB_OK) stt = mi_rb_step(world, 1);  // capture + warm
        if (stt == MI_RB_OK) stt = restore();
        if (stt == MI_RB_OK) stt = mi_rb_step(world, steps);
        if (stt != MI_RB_OK) {
            mi_rb_set_tuning(world, &keep);
            w->use_graph = keep_graph;
            return stt;
        }
        ms = w->last_ms / (float)steps;
        return MI_RB_OK;
    };
    int idx = 0;
    for (int bi = 0; bi < 2; bi++)
        for (int wi = 0; wi < 4; wi++, idx++) {
            const mi_rb_tuning t{blocks[bi], waves[wi], keep.sort_tile, keep.sort_digit_bits};
            float ms = 0.f;
            mi_rb_status stt = trial(t, ms);
            if (stt != MI_RB_OK) return stt;
            if (ms_out) ms_out[idx] = ms;
            if (idx == 0 || ms < best_ms) { best_ms = ms; bt = t; }
        }
    // the body-key sort's tile and digit width with the fastest launch shape
    const int tiles[2] = {1024, 4096}, digits[2] = {7, 8};
    for (int ti = 0; ti < 2; ti++)
        for (int di = 0; di < 2; di++) {
            const mi_rb_tuning t{bt.solve_block, bt.narrow_waves, tiles[ti], digits[di]};
            float ms = 0.f;
            mi_rb_status stt = trial(t, ms);
            if (stt != MI_RB_OK) return stt;
            if (ms < best_ms) { best_ms = ms; bt = t; }
        }
    mi_rb_status stt = mi_rb_set_tuning(world, &bt);
    if (stt != MI_RB_OK) return stt;
    w->use_graph = keep_graph;
    w->steps = keep_steps;
    stt = restore();
    if (stt != MI_RB_OK) return stt;
    *best = bt;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_last_step_ms(void *world, float *ms) {
    World *w = W(world);
    if (!w || !ms) return MI_RB_INVALID_ARG;
    *ms = w->last_ms;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_readback(void *world) {
    World *w = W(world);
    if (!w) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    size_t n = w->pos.size();
    RB_CHECK(hipMemcpyAsync(w->pos.data(), w->d.pos, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream));
    RB_CHECK(hipMemcpyAsync(w->quat.data(), w->d.quat, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream));
    RB_CHECK(hipMemcpyAsync(w->linvel.data(), w->d.linvel, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream));
    RB_CHECK(hipMemcpyAsync(w->angvel.data(), w->d.angvel, sizeof(float4) * n, hipMemcpyDeviceToHost, w->stream));
    RB_CHECK(hipStreamSynchronize(w->stream));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_get_transform(void *world, int body, float pos[4], float quat[4]) {
    World *w = W(world);
    if (!w || !pos || !quat || body < 0 || body >= (int)w->pos.size()) return MI_RB_INVALID_ARG;
    std::memcpy(pos, &w->pos[body], 16);
    std::memcpy(quat, &w->quat[body], 16);
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_get_stats(void *world, mi_rb_stats *out) {
    World *w = W(world);
    if (!w || !out) return MI_RB_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    out->num_bodies = (int)w->pos.size();
    out->steps = w->steps;
    if (!w->uploaded) return MI_RB_OK;
    int c[RB_CNT_COUNT];
    RB_CHECK(hipMemcpyAsync(c, w->d.counters, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    RB_CHECK(hipStreamSynchronize(w->stream));
    out->num_pairs = c[RB_CNT_PAIRS];
    out->num_contacts = c[RB_CNT_CONTACTS];
    out->num_colors = c[RB_CNT_NCOLORS];
    out->solve_bodies = c[RB_CNT_SOLVE_BODIES];
    out->face_pass_pairs = c[RB_CNT_FACE_WORK] + c[RB_CNT_FACE_SMALL];
    out->overflow_flags = c[RB_CNT_OVERFLOW];
    out->normal_iterations = c[RB_CNT_ITERS_N];
    out->friction_iterations = c[RB_CNT_ITERS_F];
    std::memcpy(&out->normal_residual, &c[RB_CNT_RES_N], 4);
    std::memcpy(&out->friction_residual, &c[RB_CNT_RES_F], 4);
    if (w->sp.tile_mode == 1 && w->sp.tol > 0.f) {
        int tiles = c[RB_CNT_TILES_SOLVED] > 0 ? c[RB_CNT_TILES_SOLVED] : 1;
        out->mean_normal_iterations = (float)c[RB_CNT_SUM_ITERS_N] / (float)tiles;
        out->mean_friction_iterations = (float)c[RB_CNT_SUM_ITERS_F] / (float)tiles;
    } else {
        out->mean_normal_iterations = (float)out->normal_iterations;
        out->mean_friction_iterations = (float)out->friction_iterations;
    }
    if (w->timing && !w->use_graph && w->events_full) {
        for (int k = 0; k < 8; k++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, w->ev[k], w->ev[k + 1]) == hipSuccess) out->stage_ms[k] = ms;
        }
        float sum = 0.f;
        for (int k = 0; k < w->kev_launches; k++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, w->kev[2 * k], w->kev[2 * k + 1]) == hipSuccess) sum += ms;
        }
        out->solve_events_ms = sum;
        out->solve_launches = w->kev_launches;
        (void)hipGetLastError();
    }
    if (w->solve_timing && w->span_launches > 0) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, w->span[0], w->span[1]) == hipSuccess) {
            out->solve_kernel_ms = ms;
            out->solve_launches = w->span_launches;
        }
        (void)hipGetLastError();  // an event query must not leave a sticky error behind
    }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_get_params(void *world, mi_rb_params *out) {
    World *w = W(world);
    if (!w || !out) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    const StepParams &p = w->sp;
    out->inv_cell = p.inv_cell;
    out->cell_size = 1.0f / p.inv_cell;
    for (int k = 0; k < 3; k++) out->dims[k] = p.dims[k];
    out->cell_bits = p.cell_bits;
    out->key_bits = w->key_bits;
    out->n_scenes = p.n_scenes;
    out->n_large = p.n_large;
    out->tile_mode = p.tile_mode;
    out->n_tiles = p.n_tiles;
    out->tiles_x = p.tiles_x;
    out->tiles_z = p.tiles_z;
    out->tile_inv = p.tile_inv;
    out->tile_origin_x = p.tile_origin_x;
    out->tile_origin_z = p.tile_origin_z;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_device_ptrs(void *world, mi_rb_device_view *out) {
    World *w = W(world);
    if (!w || !out) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    out->pos = w->d.pos;
    out->quat = w->d.quat;
    out->linvel = w->d.linvel;
    out->angvel = w->d.angvel;
    out->inv_mass = w->d.inv_mass;
    out->pairs = w->d.pairs;
    out->contacts = w->d.contacts;
    out->stream = w->stream;
    out->num_bodies = (int)w->pos.size();
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_copy_transforms(void *world, void *d_vertex_buffer,
                                              int start_offset, int capacity_float4,
                                              void *stream) {
    World *w = W(world);
    if (!w || !d_vertex_buffer || start_offset < 0 || (start_offset & 3) || capacity_float4 < 0)
        return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    const int n = (int)w->pos.size();
    if ((int64_t)start_offset / 4 + 2 * (int64_t)n > capacity_float4) return MI_RB_CAPACITY;
    hipSetDevice(w->device);
    (void)hipGetLastError();
    // the step's stream unless the caller names one (then ordered after the
    // world's pending steps)
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : w->stream;
    if (stream) {
        RB_CHECK(hipEventRecord(w->ev_order, w->stream));
        RB_CHECK(hipStreamWaitEvent(s, w->ev_order, 0));
    }
    RB_CHECK(launch_copy_transforms(w->d, n, static_cast<float4 *>(d_vertex_buffer) + start_offset / 4, s));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_read_buffer(void *world, int which, void *host, size_t max_bytes,
                                          int *count) {
    World *w = W(world);
    if (!w || !count) return MI_RB_INVALID_ARG;
    if (which == MI_RB_BUF_CHILDREN) {  // host-side: compound children
        *count = (int)w->children.size();
        if (host && max_bytes && *count)
            std::memcpy(host, w->children.data(),
                        std::min(max_bytes, sizeof(ChildShape) * w->children.size()));
        return MI_RB_OK;
    }
    if (which >= MI_RB_BUF_HULLS && which <= MI_RB_BUF_SHAPE) {  // host-side registration arrays
        const void *src = nullptr;
        size_t elem = 0;
        int cnt = 0;
        switch (which) {
            case MI_RB_BUF_HULLS: src = w->hulls.data(); elem = sizeof(HullInfo); cnt = (int)w->hulls.size(); break;
            case MI_RB_BUF_VERTS: src = w->verts.data(); elem = 16; cnt = (int)w->verts.size(); break;
            case MI_RB_BUF_PLANES: src = w->planes.data(); elem = 16; cnt = (int)w->planes.size(); break;
            case MI_RB_BUF_FACE_INDEX: src = w->face_index.data(); elem = 8; cnt = (int)w->face_index.size(); break;
            case MI_RB_BUF_INDICES: src = w->indices.data(); elem = 4; cnt = (int)w->indices.size(); break;
            case MI_RB_BUF_EDGES: src = w->edges.data(); elem = 16; cnt = (int)w->edges.size(); break;
            case MI_RB_BUF_INV_MASS: src = w->inv_mass.data(); elem = 4; cnt = (int)w->inv_mass.size(); break;
            case MI_RB_BUF_INV_INERTIA_LOCAL: src = w->inv_inertia_local.data(); elem = 16; cnt = (int)w->inv_inertia_local.size(); break;
            case MI_RB_BUF_SCENE: src = w->scene.data(); elem = 4; cnt = (int)w->scene.size(); break;
            case MI_RB_BUF_IS_LARGE: src = w->is_large.data(); elem = 4; cnt = (int)w->is_large.size(); break;
            case MI_RB_BUF_SHAPE: src = w->shape.data(); elem = 4; cnt = (int)w->shape.size(); break;
            default: return MI_RB_INVALID_ARG;
        }
        *count = cnt;
        if (host && max_bytes && cnt) std::memcpy(host, src, std::min(max_bytes, elem * (size_t)cnt));
        return MI_RB_OK;
    }
    if (!w->uploaded) return MI_RB_STATE;
    int c[RB_CNT_COUNT];
    RB_CHECK(hipMemcpyAsync(c, w->d.counters, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    RB_CHECK(hipStreamSynchronize(w->stream));
    const int n = (int)w->pos.size(), np = c[RB_CNT_PAIRS], nc = c[RB_CNT_CONTACTS];
    const void *src = nullptr;
    size_t elem = 0;
    int cnt = 0;
    std::vector<char> tmp;
    switch (which) {
        case MI_RB_BUF_SORTED_KV: src = w->d.kv; elem = 8; cnt = n; break;
        case MI_RB_BUF_AABB: {
            cnt = n;
            elem = 32;
            tmp.resize((size_t)32 * n);
            RB_CHECK(hipMemcpy(tmp.data(), w->d.aabb_min, 16 * (size_t)n, hipMemcpyDeviceToHost));
            RB_CHECK(hipMemcpy(tmp.data() + 16 * (size_t)n, w->d.aabb_max, 16 * (size_t)n, hipMemcpyDeviceToHost));
            break;
        }
        case MI_RB_BUF_PAIRS: src = w->d.pairs; elem = 8; cnt = np; break;
        case MI_RB_BUF_CONTACTS: src = w->d.contacts; elem = 96; cnt = nc; break;
        case MI_RB_BUF_CONSTRAINTS: {
            // pack Constraint4 records (176 bytes) from the SoA
            cnt = nc;
            elem = 176;
            std::vector<float4> lin(nc), cen(nc), jac(nc), bb(nc), lam(nc), fr(nc), ps(4 * (size_t)nc);
            std::vector<int2> bd(nc);
            std::vector<int> pr(nc);
            if (nc) {
                RB_CHECK(hipMemcpy(lin.data(), w->d.cs.linear, 16 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(cen.data(), w->d.cs.center, 16 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(jac.data(), w->d.cs.jac_inv, 16 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(bb.data(), w->d.cs.b, 16 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(lam.data(), w->d.cs.lambda, 16 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(fr.data(), w->d.cs.fric, 16 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(ps.data(), w->d.cs.pos, 64 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(bd.data(), w->d.cs.bodies, 8 * (size_t)nc, hipMemcpyDeviceToHost));
                RB_CHECK(hipMemcpy(pr.data(), w->d.cs.pair, 4 * (size_t)nc, hipMemcpyDeviceToHost));
            }
            tmp.assign((size_t)176 * nc, 0);
            for (int k = 0; k < nc; k++) {
                float *r = reinterpret_cast<float *>(tmp.data() + (size_t)176 * k);
                std::memcpy(r, &lin[k], 16);
                std::memcpy(r + 4, &ps[4 * (size_t)k], 64);
                std::memcpy(r + 20, &cen[k], 16);
                std::memcpy(r + 24, &jac[k], 16);
                std::memcpy(r + 28, &bb[k], 16);
                std::memcpy(r + 32, &lam[k], 16);
                r[36] = fr[k].x; r[37] = fr[k].y; r[38] = fr[k].z; r[39] = fr[k].w;
                uint32_t *u = reinterpret_cast<uint32_t *>(r + 40);
                u[0] = (uint32_t)bd[k].x; u[1] = (uint32_t)bd[k].y; u[2] = (uint32_t)pr[k]; u[3] = 0;
            }
            break;
        }
        case MI_RB_BUF_SOLVE_ORDER: src = w->d.solve_order; elem = 4; cnt = nc; break;
        case MI_RB_BUF_FEATURES: src = w->d.features; elem = 64; cnt = nc; break;
        case MI_RB_BUF_PREDICTIONS: src = w->d.mlp_out; elem = 12; cnt = nc; break;
        case MI_RB_BUF_PREV_KEYS:
        case MI_RB_BUF_PREV_IMPULSES:
        case MI_RB_BUF_PREV_POINTS: {
            int pc = 0;
            RB_CHECK(hipMemcpy(&pc, w->d.prev_count, sizeof(int), hipMemcpyDeviceToHost));
            cnt = pc;
            if (which == MI_RB_BUF_PREV_KEYS) { src = w->d.prev_keys; elem = 8; }
            else if (which == MI_RB_BUF_PREV_IMPULSES) { src = w->d.prev_imp; elem = 16; }
            else { src = w->d.prev_pts; elem = 64; }
            break;
        }
        case MI_RB_BUF_COLORS: src = w->d.row_color; elem = 4; cnt = nc; break;
        case MI_RB_BUF_TILE_STATS: src = w->d.tile_stats; elem = 16; cnt = w->sp.n_tiles; break;
        case MI_RB_BUF_INV_INERTIA: src = w->d.inv_inertia_world; elem = 48; cnt = n; break;
        case MI_RB_BUF_PHASES: {
            // per tile: first, end, phase (spatial: (tx%3)+3*(tz%3); scene: 0), ncolors
            const StepParams &p = w->sp;
            cnt = p.n_tiles;
            elem = 16;
            std::vector<int> tf(p.n_tiles), te(p.n_tiles), tn(p.n_tiles);
            RB_CHECK(hipMemcpy(tf.data(), w->d.tile_first, 4 * (size_t)p.n_tiles, hipMemcpyDeviceToHost));
            RB_CHECK(hipMemcpy(te.data(), w->d.tile_cnt, 4 * (size_t)p.n_tiles, hipMemcpyDeviceToHost));
            RB_CHECK(hipMemcpy(tn.data(), w->d.tile_ncolors, 4 * (size_t)p.n_tiles, hipMemcpyDeviceToHost));
            tmp.resize((size_t)16 * p.n_tiles);
            int *o = reinterpret_cast<int *>(tmp.data());
            for (int t = 0; t < p.n_tiles; t++) {
                int ph = p.tile_mode == 1 ? 0 : ((t % p.tiles_x) & 1) + 2 * ((t / p.tiles_x) & 1);
                o[4 * t] = tf[t]; o[4 * t + 1] = te[t]; o[4 * t + 2] = ph; o[4 * t + 3] = tn[t];
            }
            break;
        }
        default: return MI_RB_INVALID_ARG;
    }
    *count = cnt;
    size_t bytes = elem * (size_t)cnt;
    if (host && max_bytes) {
        size_t b = std::min(bytes, max_bytes);
        if (!tmp.empty()) std::memcpy(host, tmp.data(), b);
        else if (b) RB_CHECK(hipMemcpy(host, src, b, hipMemcpyDeviceToHost));
    }
    return MI_RB_OK;
}

// ---- snapshot: "MIRB" v1, counts, then host arrays ----
namespace {
template <typename T>
void put(std::vector<char> &b, const std::vector<T> &v) {
    uint64_t n = v.size();
    b.insert(b.end(), (char *)&n, (char *)&n + 8);
    if (n) b.insert(b.end(), (const char *)v.data(), (const char *)v.data() + sizeof(T) * n);
}
template <typename T>
bool get(const std::vector<char> &b, size_t &o, std::vector<T> &v) {
    if (o + 8 > b.size()) return false;
    uint64_t n;
    std::memcpy(&n, &b[o], 8);
    o += 8;
    if (o + sizeof(T) * n > b.size()) return false;
    v.resize(n);
    if (n) std::memcpy(v.data(), &b[o], sizeof(T) * n);
    o += sizeof(T) * n;
    return true;
}
}  // namespace

extern "C" mi_rb_status mi_rb_save_state(void *world, const char *path) {
    World *w = W(world);
    if (!w || !path) return MI_RB_INVALID_ARG;
    if (w->uploaded) {
        mi_rb_status st = mi_rb_readback(world);
        if (st != MI_RB_OK) return st;
    }
    std::vector<char> b = {'M', 'I', 'R', 'B', 2, 0, 0, 0};
    put(b, w->hulls); put(b, w->verts); put(b, w->planes); put(b, w->edges);
    put(b, w->face_index); put(b, w->indices); put(b, w->hull_hext_norm);
    put(b, w->pos); put(b, w->quat); put(b, w->linvel); put(b, w->angvel);
    put(b, w->inv_inertia_local); put(b, w->inv_mass); put(b, w->shape); put(b, w->scene);
    put(b, w->children);
    FILE *f = fopen(path, "wb");
    if (!f) return MI_RB_INVALID_ARG;
    size_t wr = fwrite(b.data(), 1, b.size(), f);
    fclose(f);
    return wr == b.size() ? MI_RB_OK : MI_RB_INVALID_ARG;
}

namespace {
// The arrays of a snapshot file, parsed and checked before any of them
// replaces a world's.
struct SnapshotArrays {
    std::vector<HullInfo> hulls;
    std::vector<ChildShape> children;
    std::vector<float4> verts, planes, edges;
    std::vector<int2> face_index;
    std::vector<int> indices;
    std::vector<float> hull_hext_norm;
    std::vector<float4> pos, quat, linvel, angvel, inv_inertia_local;
    std::vector<float> inv_mass;
    std::vector<int> shape, scene;
};

bool finite4(const std::vector<float4> &v) {
    for (const float4 &x : v)
        if (!std::isfinite(x.x) || !std::isfinite(x.y) || !std::isfinite(x.z) || !std::isfinite(x.w))
            return false;
    return true;
}

// Reads a "MIRB" v1/v2 file and checks that every count, offset and index
// stays inside the arrays and the configuration's capacities, so that the
// device kernels can trust the shapes.  Returns MI_RB_OK,
// MI_RB_INVALID_ARG (unreadable, truncated or inconsistent) or
// MI_RB_CAPACITY (beyond the configuration).
mi_rb_status parse_snapshot(const char *path, const mi_rb_config &cfg, SnapshotArrays &s) {
    FILE *f = fopen(path, "rb");
    if (!f) return MI_RB_INVALID_ARG;
    std::vector<char> b;
    char buf[65536];
    size_t r;
    while ((r = fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + r);
    fclose(f);
    // version 1: no compound children; version 2 adds them
    if (b.size() < 8 || std::memcmp(b.data(), "MIRB", 4) != 0 || (b[4] != 1 && b[4] != 2))
        return MI_RB_INVALID_ARG;
    size_t o = 8;
    bool ok = get(b, o, s.hulls) && get(b, o, s.verts) && get(b, o, s.planes) &&
              get(b, o, s.edges) && get(b, o, s.face_index) && get(b, o, s.indices) &&
              get(b, o, s.hull_hext_norm) && get(b, o, s.pos) && get(b, o, s.quat) &&
              get(b, o, s.linvel) && get(b, o, s.angvel) && get(b, o, s.inv_inertia_local) &&
              get(b, o, s.inv_mass) && get(b, o, s.shape) && get(b, o, s.scene);
    if (ok && b[4] == 2) ok = get(b, o, s.children);
    if (!ok || o != b.size()) return MI_RB_INVALID_ARG;
    const size_t n = s.pos.size(), S = s.hulls.size();
    if (s.quat.size() != n || s.linvel.size() != n || s.angvel.size() != n ||
        s.inv_inertia_local.size() != n || s.inv_mass.size() != n || s.shape.size() != n ||
        s.scene.size() != n || s.hull_hext_norm.size() != S ||
        s.planes.size() != s.face_index.size())
        return MI_RB_INVALID_ARG;
    if (n > (size_t)cfg.max_bodies || S > (size_t)cfg.max_shapes ||
        s.verts.size() > (size_t)cfg.max_hull_vertices)
        return MI_RB_CAPACITY;
    if (!finite4(s.pos) || !finite4(s.quat) || !finite4(s.linvel) || !finite4(s.angvel) ||
        !finite4(s.inv_inertia_local) || !finite4(s.verts) || !finite4(s.planes) ||
        !finite4(s.edges))
        return MI_RB_INVALID_ARG;
    auto in = [](long long off, long long cnt, size_t size) {
        return off >= 0 && cnt >= 0 && off + cnt <= (long long)size;
    };
    for (size_t k = 0; k < S; k++) {
        const HullInfo &h = s.hulls[k];
        const int kind = shape_kind_of(h);
        if (kind == SHAPE_KIND_HULL) {
            if (h.num_vertices > RB_MAX_HULL_VERTS || h.num_edges > RB_MAX_HULL_EDGES ||
                h.num_faces > 2 * RB_MAX_HULL_FACES || h.num_faces < 1 ||
                !in(h.vertex_offset, h.num_vertices, s.verts.size()) ||
                !in(h.face_offset, h.num_faces, s.planes.size()) ||
                !in(h.edge_offset, h.num_edges, s.edges.size()) ||
                !in(h.index_offset, h.num_indices, s.indices.size()))
                return MI_RB_INVALID_ARG;
            for (int fc = 0; fc < h.num_faces; fc++) {
                const int2 fi = s.face_index[h.face_offset + fc];
                if (fi.y < 1 || fi.y > RB_MAX_FACE_VERTS || !in(fi.x, fi.y, s.indices.size()))
                    return MI_RB_INVALID_ARG;
                for (int j = 0; j < fi.y; j++) {
                    const int v = s.indices[fi.x + j];
                    if (v < 0 || v >= h.num_vertices) return MI_RB_INVALID_ARG;
                }
            }
        } else if (kind == SHAPE_KIND_COMPOUND) {
            const int max_children = compound_is_mesh(h) ? MI_RB_MAX_MESH_TRIANGLES : 64;
            if (h.num_edges > max_children || !in(h.edge_offset, h.num_edges, s.children.size()))
                return MI_RB_INVALID_ARG;
            for (int c = 0; c < h.num_edges; c++) {
                const ChildShape &ch = s.children[h.edge_offset + c];
                if (ch.shape < 0 || ch.shape >= (int)S ||
                    shape_kind_of(s.hulls[ch.shape]) != SHAPE_KIND_HULL)
                    return MI_RB_INVALID_ARG;
            }
        } else if (kind == SHAPE_KIND_PLANE) {
            if (!in(h.face_offset, 1, s.planes.size())) return MI_RB_INVALID_ARG;
        }
    }
    for (size_t i = 0; i < n; i++) {
        const int sh = s.shape[i];
        if (sh < 0 || sh >= (int)S || s.scene[i] < 0 || (i > 0 && s.scene[i] < s.scene[i - 1]))
            return MI_RB_INVALID_ARG;
        if (!std::isfinite(s.inv_mass[i]) || s.inv_mass[i] < 0.f) return MI_RB_INVALID_ARG;
        const HullInfo &h = s.hulls[sh];
        const int kind = shape_kind_of(h);
        // planes and concave meshes are static only (as at registration)
        if (s.inv_mass[i] != 0.f && (kind == SHAPE_KIND_PLANE || compound_is_mesh(h)))
            return MI_RB_INVALID_ARG;
    }
    return MI_RB_OK;
}
}  // namespace

extern "C" mi_rb_status mi_rb_load_state(void *world, const char *path) {
    World *w = W(world);
    if (!w || !path) return MI_RB_INVALID_ARG;
    SnapshotArrays s;
    mi_rb_status st = parse_snapshot(path, w->cfg, s);
    if (st != MI_RB_OK) return st;  // the world is left as it was
    if (w->uploaded) {
        hipSetDevice(w->device);
        hipStreamSynchronize(w->stream);
    }
    w->free_device();
    w->hulls.swap(s.hulls); w->children.swap(s.children);
    w->verts.swap(s.verts); w->planes.swap(s.planes); w->edges.swap(s.edges);
    w->face_index.swap(s.face_index); w->indices.swap(s.indices);
    w->hull_hext_norm.swap(s.hull_hext_norm);
    w->pos.swap(s.pos); w->quat.swap(s.quat); w->linvel.swap(s.linvel); w->angvel.swap(s.angvel);
    w->inv_inertia_local.swap(s.inv_inertia_local); w->inv_mass.swap(s.inv_mass);
    w->shape.swap(s.shape); w->scene.swap(s.scene);
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_snapshot_check(const char *path, const mi_rb_config *cfg,
                                             int counts[4]) {
    if (!path || !cfg) return MI_RB_INVALID_ARG;
    SnapshotArrays s;
    mi_rb_status st = parse_snapshot(path, *cfg, s);
    if (st == MI_RB_OK && counts) {
        counts[0] = (int)s.pos.size();
        counts[1] = (int)s.hulls.size();
        counts[2] = (int)s.verts.size();
        counts[3] = (int)s.children.size();
    }
    return st;
}

extern "C" mi_rb_status mi_rb_set_warmstart_weights(void *world, const uint16_t *w_bf16,
                                                    const float *bias) {
    World *w = W(world);
    if (!w || !w_bf16 || !bias) return MI_RB_INVALID_ARG;
    w->mlp_w_host.assign(w_bf16, w_bf16 + RB_MLP_NW);
    w->mlp_b_host.assign(bias, bias + RB_MLP_NB);
    if (w->uploaded) {
        hipSetDevice(w->device);
        // ordered with the world's stream: steps already queued keep their weights
        RB_CHECK(hipStreamSynchronize(w->stream));
        RB_CHECK(hipMemcpy(w->d.mlp_w, w->mlp_w_host.data(), 2 * (size_t)RB_MLP_NW, hipMemcpyHostToDevice));
        RB_CHECK(hipMemcpy(w->d.mlp_b, w->mlp_b_host.data(), 4 * (size_t)RB_MLP_NB, hipMemcpyHostToDevice));
    }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_copy_training_batch(void *world, float *x_dev, float *t_dev,
                                                  int capacity, int *count) {
    World *w = W(world);
    if (!w || !x_dev || !t_dev || capacity < 0 || !count) return MI_RB_INVALID_ARG;
    if (!w->uploaded) return MI_RB_STATE;
    hipSetDevice(w->device);
    *count = 0;
    if (capacity == 0) return MI_RB_OK;
    RB_CHECK(launch_mlp_targets(w->d, capacity, x_dev, t_dev, w->stream));
    int nc = 0;
    RB_CHECK(hipMemcpyAsync(&nc, w->d.counters + RB_CNT_CONTACTS, sizeof(int), hipMemcpyDeviceToHost,
                            w->stream));
    RB_CHECK(hipStreamSynchronize(w->stream));
    *count = nc < capacity ? nc : capacity;
    return MI_RB_OK;
}

// ================================================================ grid API
namespace {
struct Grid {
    mi_rb_grid_desc desc;
    int max_n;
    uint2 *kv = nullptr;
    uint32_t *ys = nullptr;              // per pair slot: the smaller index, grouped by the larger
    uint32_t *cursor = nullptr, *big = nullptr;
    uint32_t *cell_start = nullptr, *counts = nullptr, *offsets = nullptr, *total = nullptr;
    uint32_t *occ = nullptr;             // occupancy bit per cell, then one flag per level
    float4 *smn = nullptr, *smx = nullptr;  // AABBs in sorted order
    int *overflow = nullptr;
    SortWork sw{};
    uint32_t *scan_status = nullptr;
    uint32_t *lane_hits = nullptr;       // per body and search lane: pair count, first partners
    int cell_bits = 0;
    hipEvent_t ev[5] = {};
    void release() {
        void *ps[] = {kv, ys, cursor, big, cell_start, counts, offsets, total, overflow, sw.tmp,
                      sw.hist, sw.status, sw.tile_ctr, scan_status, occ, smn, smx, lane_hits};
        for (void *p : ps) if (p) hipFree(p);
        for (auto &e : ev) if (e) hipEventDestroy(e);
    }
};
}  // namespace

extern "C" mi_rb_status mi_rb_grid_create(const mi_rb_grid_desc *desc, int max_aabbs, void **out_grid) {
    if (!desc || !out_grid || max_aabbs <= 0 || desc->cell_size <= 0.f || desc->max_pairs <= 0 ||
        desc->levels < 0 || desc->levels > 4)
        return MI_RB_INVALID_ARG;
    for (int k = 0; k < 3; k++)
        if (desc->dims[k] < 4 || (desc->dims[k] & (desc->dims[k] - 1))) return MI_RB_INVALID_ARG;
    Grid *g = new Grid();
    g->desc = *desc;
    if (g->desc.levels == 0) g->desc.levels = 1;
    g->max_n = max_aabbs;
    g->cell_bits = ilog2((uint64_t)desc->dims[0] * desc->dims[1] * desc->dims[2]);
    if (g->cell_bits > 28) { delete g; return MI_RB_INVALID_ARG; }
    size_t ncells = (size_t)1 << (g->cell_bits + 2);
    int n = max_aabbs, P = desc->max_pairs;
    bool ok = dalloc(&g->kv, n) == hipSuccess && dalloc(&g->ys, P) == hipSuccess &&
              dalloc(&g->cursor, n) == hipSuccess && dalloc(&g->big, (size_t)n + 2) == hipSuccess &&
              dalloc(&g->cell_start, ncells) == hipSuccess &&
              dalloc(&g->occ, ncells / 32 + 4) == hipSuccess && dalloc(&g->smn, n) == hipSuccess &&
              dalloc(&g->smx, n) == hipSuccess && dalloc(&g->counts, n) == hipSuccess &&
              dalloc(&g->offsets, n) == hipSuccess && dalloc(&g->total, 4) == hipSuccess &&
              dalloc(&g->overflow, 4) == hipSuccess && dalloc(&g->sw.tmp, n) == hipSuccess &&
              dalloc(&g->sw.hist, 8 * 256) == hipSuccess &&
              dalloc(&g->sw.status, sort_status_words(n, 8)) == hipSuccess &&
              dalloc(&g->sw.tile_ctr, 16) == hipSuccess &&
              dalloc(&g->scan_status, rb_scan_status_words(n)) == hipSuccess &&
              dalloc(&g->lane_hits, (size_t)n * RB_HG_LANES * (1 + RB_HG_KEEP)) == hipSuccess;
    if (!ok) { g->release(); delete g; return MI_RB_DEVICE; }
    g->sw.max_n = n;
    g->sw.max_tiles = rb_sort_tiles(n);
    g->sw.err = g->overflow;
    hipMemset(g->cell_start, 0xff, sizeof(uint32_t) * ncells);
    hipMemset(g->overflow, 0, 16);
    for (auto &e : g->ev) hipEventCreate(&e);
    if (hipDeviceSynchronize() != hipSuccess) { g->release(); delete g; return MI_RB_DEVICE; }
    *out_grid = g;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_grid_destroy(void *grid) {
    Grid *g = static_cast<Grid *>(grid);
    if (!g) return MI_RB_INVALID_ARG;
    hipDeviceSynchronize();
    g->release();
    delete g;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_grid_find_pairs(void *grid, const void *d_aabb_min,
                                              const void *d_aabb_max, const void *d_is_large,
                                              int n, void *d_pairs, void *d_num_pairs,
                                              void *stream) {
    Grid *g = static_cast<Grid *>(grid);
    if (!g || !d_aabb_min || !d_aabb_max || !d_pairs || !d_num_pairs || n < 0 || n > g->max_n)
        return MI_RB_INVALID_ARG;
    if (d_is_large) return MI_RB_INVALID_ARG;  // the levels replace the overflow list
    hipStream_t s = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    RB_CHECK(launch_hgrid_pairs(static_cast<const float4 *>(d_aabb_min),
                                static_cast<const float4 *>(d_aabb_max), n,
                                1.0f / g->desc.cell_size, g->desc.dims, g->cell_bits,
                                g->desc.levels, g->kv, g->cell_start, g->occ, g->smn, g->smx,
                                g->counts, g->offsets, g->cursor, g->big, g->total, g->ys,
                                g->desc.max_pairs,
                                static_cast<int2 *>(d_pairs), static_cast<int *>(d_num_pairs),
                                g->overflow, g->sw, g->scan_status, s, g->ev, g->lane_hits));
    return MI_RB_OK;
}

namespace {
struct Sap {
    mi_rb_sap_desc desc;
    int max_n;
    uint2 *kv = nullptr;
    float4 *smn = nullptr, *smx = nullptr;
    uint32_t *ys = nullptr, *counts = nullptr, *offsets = nullptr, *cursor = nullptr,
             *big = nullptr, *lg = nullptr, *total = nullptr;
    int *overflow = nullptr;
    SortWork sw{};
    uint32_t *scan_status = nullptr;
    hipEvent_t ev[5] = {};
    void release() {
        void *ps[] = {kv, smn, smx, ys, counts, offsets, cursor, big, lg, total, overflow, sw.tmp,
                      sw.hist, sw.status, sw.tile_ctr, scan_status};
        for (void *p : ps) if (p) hipFree(p);
        for (auto &e : ev) if (e) hipEventDestroy(e);
    }
};
}  // namespace

extern "C" mi_rb_status mi_rb_sap_create(const mi_rb_sap_desc *desc, int max_aabbs, void **out_sap) {
    if (!desc || !out_sap || max_aabbs <= 0 || desc->axis < 0 || desc->axis > 2 ||
        desc->max_pairs <= 0)
        return MI_RB_INVALID_ARG;
    Sap *g = new Sap();
    g->desc = *desc;
    g->max_n = max_aabbs;
    const int n = max_aabbs, P = desc->max_pairs;
    bool ok = dalloc(&g->kv, n) == hipSuccess && dalloc(&g->smn, n) == hipSuccess &&
              dalloc(&g->smx, n) == hipSuccess && dalloc(&g->ys, P) == hipSuccess &&
              dalloc(&g->counts, n) == hipSuccess && dalloc(&g->offsets, n) == hipSuccess &&
              dalloc(&g->cursor, n) == hipSuccess && dalloc(&g->big, (size_t)n + 2) == hipSuccess &&
              dalloc(&g->lg, (size_t)n + 1) == hipSuccess &&
              dalloc(&g->total, 4) == hipSuccess && dalloc(&g->overflow, 4) == hipSuccess &&
              dalloc(&g->sw.tmp, n) == hipSuccess && dalloc(&g->sw.hist, 8 * 256) == hipSuccess &&
              dalloc(&g->sw.status, sort_status_words(n, 8)) == hipSuccess &&
              dalloc(&g->sw.tile_ctr, 16) == hipSuccess &&
              dalloc(&g->scan_status, rb_scan_status_words(n)) == hipSuccess;
    if (!ok) { g->release(); delete g; return MI_RB_DEVICE; }
    g->sw.max_n = n;
    g->sw.max_tiles = rb_sort_tiles(n);
    g->sw.err = g->overflow;
    hipMemset(g->overflow, 0, 16);
    for (auto &e : g->ev) hipEventCreate(&e);
    if (hipDeviceSynchronize() != hipSuccess) { g->release(); delete g; return MI_RB_DEVICE; }
    *out_sap = g;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_sap_destroy(void *sap) {
    Sap *g = static_cast<Sap *>(sap);
    if (!g) return MI_RB_INVALID_ARG;
    hipDeviceSynchronize();
    g->release();
    delete g;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_sap_find_pairs(void *sap, const void *d_aabb_min,
                                             const void *d_aabb_max, const void *d_is_large,
                                             int n, void *d_pairs, void *d_num_pairs,
                                             void *stream) {
    Sap *g = static_cast<Sap *>(sap);
    if (!g || !d_aabb_min || !d_aabb_max || !d_pairs || !d_num_pairs || n < 0 || n > g->max_n)
        return MI_RB_INVALID_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    (void)hipGetLastError();
    RB_CHECK(launch_sap_pairs(static_cast<const float4 *>(d_aabb_min),
                              static_cast<const float4 *>(d_aabb_max),
                              static_cast<const int *>(d_is_large), n, g->desc.axis, g->kv,
                              g->smn, g->smx, g->counts, g->offsets, g->cursor, g->big, g->lg,
                              g->total, g->ys, g->desc.max_pairs, static_cast<int2 *>(d_pairs),
                              static_cast<int *>(d_num_pairs), g->overflow, g->sw,
                              g->scan_status, s, g->ev));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_sap_overflow(void *sap, int *flags) {
    Sap *g = static_cast<Sap *>(sap);
    if (!g || !flags) return MI_RB_INVALID_ARG;
    RB_CHECK(hipDeviceSynchronize());
    RB_CHECK(hipMemcpy(flags, g->overflow, sizeof(int), hipMemcpyDeviceToHost));
    RB_CHECK(hipMemset(g->overflow, 0, sizeof(int)));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_sap_last_ms(void *sap, float ms[4]) {
    Sap *g = static_cast<Sap *>(sap);
    if (!g || !ms) return MI_RB_INVALID_ARG;
    RB_CHECK(hipEventSynchronize(g->ev[4]));
    for (int k = 0; k < 4; k++) {
        ms[k] = 0.f;
        hipEventElapsedTime(&ms[k], g->ev[k], g->ev[k + 1]);
    }
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_grid_overflow(void *grid, int *flags) {
    Grid *g = static_cast<Grid *>(grid);
    if (!g || !flags) return MI_RB_INVALID_ARG;
    RB_CHECK(hipDeviceSynchronize());
    RB_CHECK(hipMemcpy(flags, g->overflow, sizeof(int), hipMemcpyDeviceToHost));
    RB_CHECK(hipMemset(g->overflow, 0, sizeof(int)));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_grid_last_ms(void *grid, float ms[4]) {
    Grid *g = static_cast<Grid *>(grid);
    if (!g || !ms) return MI_RB_INVALID_ARG;
    RB_CHECK(hipEventSynchronize(g->ev[4]));
    for (int k = 0; k < 4; k++) {
        ms[k] = 0.f;
        hipEventElapsedTime(&ms[k], g->ev[k], g->ev[k + 1]);
    }
    return MI_RB_OK;
}
