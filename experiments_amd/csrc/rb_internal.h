// rb_internal.h -- layouts and launch interfaces shared by the HIP kernels and
// the host orchestration of the rigid-body step.  gfx950, wave64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mi_rb.h"

#define RB_CHECK(expr)                                                          \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) return rb_device_error(e_, #expr, __FILE__, __LINE__); \
    } while (0)

mi_rb_status rb_device_error(hipError_t e, const char *what, const char *file, int line);

// Convex hull record, 80 bytes (ConvexPolyhedronCL, opencl/gpu_rigidbody_pipeline2/ConvexPolyhedronCL.h:15-30,
// plus the local AABB that CLPhysicsDemo keeps beside it).
struct HullInfo {
    int face_offset, num_faces, vertex_offset, num_vertices;
    int edge_offset, num_edges, index_offset, num_indices;
    float4 center;    // vertex average (CLPhysicsDemo.cpp:380-385)
    float4 aabb_min;  // local AABB of the vertices
    float4 aabb_max;
};
static_assert(sizeof(HullInfo) == 80, "HullInfo layout");
// A shape without vertices is a compound (num_edges children from
// edge_offset in the child table), a plane (one face at face_offset: normal
// xyz and constant w, n . x = w) or else a sphere (radius in center.w).
enum { SHAPE_KIND_HULL = 0, SHAPE_KIND_SPHERE = 1, SHAPE_KIND_PLANE = 2, SHAPE_KIND_COMPOUND = 3 };
inline __host__ __device__ int shape_kind_of(const HullInfo &h) {
    return h.num_vertices > 0 ? SHAPE_KIND_HULL
                              : (h.num_edges > 0 ? SHAPE_KIND_COMPOUND
                                                 : (h.num_faces > 0 ? SHAPE_KIND_PLANE
                                                                    : SHAPE_KIND_SPHERE));
}
// A compound with a nonzero num_indices is a concave triangle mesh: its
// children are triangle hulls at the identity, and a pair with a hull gets
// one child pair per triangle whose world AABB overlaps the hull's.
inline __host__ __device__ bool compound_is_mesh(const HullInfo &h) {
    return h.num_vertices == 0 && h.num_edges > 0 && h.num_indices != 0;
}
// A child of a compound shape: a hull shape at a position and orientation
// in the compound's frame (btGpuChildShape, 48 bytes).
struct ChildShape {
    float4 pos;   // w unused
    float4 orn;   // quaternion xyzw
    int shape;    // hull shape index
    int pad[3];
};
static_assert(sizeof(ChildShape) == 48, "ChildShape layout");

// Contact manifold in the Contact4 layout (dynamics/basic_demo/Stubs/AdlContact4.h:26-59).
struct Contact4 {
    float4 pos[4];    // xyz point, w depth
    float4 normal;    // xyz normal, w = number of points
    uint32_t coeffs;  // restitution | friction << 16, 16-bit fixed point
    int pair_index;   // m_batchIdx
    int body_a;       // negated when static
    int body_b;
};
static_assert(sizeof(Contact4) == 96, "Contact4 layout");

// Limits of the wave-per-pair narrowphase.
#define RB_MAX_HULL_VERTS 64
#define RB_MAX_HULL_FACES 64
#define RB_MAX_FACE_VERTS 64
#define RB_MAX_HULL_EDGES 128
#define RB_MAX_POLY 64

// Learned warm-start network 16-64-64-3: bf16 weights W1[64][16], W2[64][64],
// W3[3][64] (out-major, in that order), fp32 biases b1[64], b2[64], b3[3].
#define RB_MLP_IN 16
#define RB_MLP_H 64
#define RB_MLP_OUT 3
#define RB_MLP_NW (RB_MLP_H * RB_MLP_IN + RB_MLP_H * RB_MLP_H + RB_MLP_OUT * RB_MLP_H)
#define RB_MLP_NB (2 * RB_MLP_H + RB_MLP_OUT)
#define RB_MLP_SCALE 10.f  // network units per unit of impulse (inputs and outputs)

// Solver sweeps per phase (normal, then friction), fixed or the cap under a tolerance
#define RB_MAX_SWEEPS 4096

// Solver tiles
#define RB_TILE_MAX_ROWS 2048     // constraints per tile held by one workgroup
#define RB_TILE_MAX_BODIES 1024   // distinct bodies per tile cached in LDS
#define RB_MAX_LOCAL_COLORS 64    // colours per tile

// Constraint rows, structure of arrays (fields of Constraint4,
// dynamics/basic_demo/Stubs/AdlConstraint4.h:23-47).
struct ConstraintSoA {
    float4 *linear;     // -normal, w friction
    float4 *pos;        // 4 per constraint
    float4 *center;
    float4 *jac_inv;    // 4 rows
    float4 *b;
    float4 *lambda;
    float4 *fric;       // x,y f_jac_inv; z,w f_lambda
    int2 *bodies;       // (a, b) absolute indices
    int *pair;          // pair index (m_batchIdx)
};

// Device state of one world.
// A tile's rows in solve order, packed for the solve launches (position i).
struct PackedRows {
    float4 *lin0, *p /* 4 per row */, *cen, *jac, *b, *lam, *fric;
    float4 *IA, *IB;  // 3 rows each: world inverse inertia of body A, B
    float4 *pA, *pB;
    int4 *meta;       // body A, body B, their LDS slots (-1 static, -2 global)
    float2 *im;       // inverse masses
    // the velocity-independent terms of the normal rows, 4 per row (one per
    // point): angular Jacobians a0, a1 and their images under the inverse
    // inertias; 2 per row: the linear impulse directions scaled by 1/m
    float4 *na0, *na1, *nia0, *nia1, *nl;
};

struct DeviceWorld {
    // shapes
    HullInfo *hulls;
    float4 *verts, *planes, *edges;
    int2 *face_index;
    int *indices;
    // bodies
    float4 *pos, *quat, *linvel, *angvel;
    float *inv_mass;
    int *shape, *scene, *is_large;
    // compound children as proxy bodies n .. n+P-1 of pos/quat/inv_mass/shape
    ChildShape *children;
    int2 *proxy;          // per proxy: (parent body, child index)
    int *proxy_first;     // per body: its first proxy (-1: not a compound)
    int *owner;           // per body and proxy: the body a contact names
    float4 *inv_inertia_local;   // diagonal
    float4 *inv_inertia_world;   // 3 rows per body
    int *scene_start, *scene_count;
    int *large_list;             // indices of large bodies (overflow list), ascending
    int *large_off;              // [n_large+1] offsets of their pair lists in large_slots
    int *large_first, *large_num;// per scene: first position in large_list, count
    int *large_index;            // per body: position in large_list, or -1
    uint32_t *large_chunk_cnt;   // [n_large x large_chunks] overlap counts per scene chunk
    // broadphase
    float4 *aabb_min, *aabb_max;
    uint2 *kv, *kv_tmp;
    uint32_t *cell_start;
    uint32_t *occ;               // occupancy bit per cell key, cleared by k_setup
    int occ_words;
    float4 *smn, *smx;           // AABBs in sorted key order
    int *body_slots;             // max_pairs_per_body per body
    int *large_slots;
    uint32_t *pair_count, *pair_off;
    int2 *pairs;
    // narrowphase
    Contact4 *contact_slots;     // one per pair
    uint32_t *has_contact, *contact_off;
    Contact4 *contacts;          // compacted, ordered by pair
    // solver
    ConstraintSoA cs;
    uint2 *tile_kv;              // (owner tile, constraint), sorted by tile
    int *tile_first, *tile_cnt;  // per tile: first sorted position, end (one past last)
    int *solve_order;            // constraints by tile, colour-major within a tile
    int *row_color;              // local colour per constraint
    int *tile_color_off;         // per tile RB_MAX_LOCAL_COLORS+1 offsets (relative)
    int *tile_ncolors;
    int4 *tile_stats;            // per tile: normal / friction sweeps used, final residuals (bits)
    PackedRows pk;               // rows in solve order (built per step after the colouring)
    int *solve_pos;              // per constraint: its position in the packed rows, -1 if unsolved
    int *tile_bodies;            // per tile at 2 * first: its dynamic bodies (dense slots)
    int *tile_nb;                // per tile: number of those bodies
    int4 *tile_hdr;              // per tile: first row, rows, colours, cached bodies
    uint2 *tile_order;           // scene tiles by descending rows: (key, tile)
    // narrowphase pass 1 -> pass 2
    uint32_t *narrow_cand;       // per pair: 1 if not separated by a face axis
    uint32_t *cand_off;          // its exclusive scan
    int4 *face_best;             // per pair: face minimum (depth bits, axis), edge-axis count
    uint2 *cand_kv;              // (cost key, pair) of the candidates, sorted
    float *residual;             // per sweep: max scaled dlambda, max lambda (as uint bits)
    // warm start
    int2 *prev_keys;             // previous step's contacts keyed by their pair entry (x, y), sorted
    int2 *stage_keys;            // worlds with compound children: the cache before it is sorted
    float4 *stage_imp, *stage_pts;
    float4 *prev_imp;            // their impulses: (sum of normal lambdas, friction 0, friction 1, points)
    float4 *prev_pts;            // 4 per entry: point position (xyz) and its normal lambda (w)
    int *prev_count;             // number of entries (persists across steps)
    float *features;             // 16 floats per contact, learned warm-start inputs
    float *mlp_out;              // 3 floats per contact, learned warm-start predictions
    uint16_t *mlp_w;
    float *mlp_b;
    // counters (device scalars)
    int *counters;               // see RB_CNT_*
};

enum {
    RB_CNT_PAIRS = 0,
    RB_CNT_CONTACTS = 1,
    RB_CNT_OVERFLOW = 2,
    RB_CNT_NCOLORS = 3,
    RB_CNT_ITERS_N = 4,
    RB_CNT_ITERS_F = 5,
    RB_CNT_RES_N = 6,
    RB_CNT_RES_F = 7,
    RB_CNT_SORT_TILE = 8,     // uint32 total of the pair-count scan
    RB_CNT_CONTACT_TOTAL = 9, // uint32 total of the contact scan
    RB_CNT_NARROW_CAND = 10,  // uint32 number of narrowphase pass-2 candidates
    RB_CNT_SOLVE_BODIES = 11, // dynamic bodies of the solver tiles, summed over tiles
    RB_CNT_DONE = 12,         // solver converged flag for the remaining sweeps
    RB_CNT_SUM_ITERS_N = 13,  // scene tiles with a tolerance: sweeps summed over tiles
    RB_CNT_SUM_ITERS_F = 14,
    RB_CNT_TILES_SOLVED = 15, // ... and the number of tiles that had rows
    RB_CNT_FACE_WORK = 16,    // uint32 pairs the quick face test left for the face pass
    RB_CNT_CLIP_SPLIT = 17,   // first small-hull candidate in the clip order
    RB_CNT_FACE_SMALL = 18,   // uint32 small-hull pairs of the face pass (after FACE_WORK's)
    RB_CNT_COUNT = 19
};

// Step parameters passed by value to kernels.
struct StepParams {
    int n_bodies;
    int n_scenes;
    int n_large;
    int large_chunks;    // scene chunks of 4096 bodies scanned per large body
    float dt;
    float solver_dt;
    float gravity[3];
    float inv_cell;
    int dims[3];
    int cell_bits;
    int max_pairs;
    int max_pairs_per_body;
    int max_contacts;
    float friction;
    float baumgarte;
    float drift;
    float threshold;
    float damping;
    float max_ang;
    int sat_edge_reject;  // 1: an edge axis with disjoint projections separates (CPU twin)
    int reduce_mode;      // MI_RB_REDUCE_*
    // solver tiling
    float tile_inv;
    int tiles_x, tiles_z;
    float tile_origin_x, tile_origin_z;
    int n_tiles;
    int tile_mode;       // 0: spatial tiles (4 phases per sweep), 1: one tile per scene
    int normal_iters, friction_iters;
    float tol;
    int warm_start;
    // launch shapes (mi_rb_tuning)
    int solve_block;     // threads per scene-tile solve workgroup: 256 or 512
    int narrow_waves;    // narrowphase waves per SIMD (grid = CUs * 4 * narrow_waves)
    int has_vertexless;  // some shape is a sphere, a plane or a compound
    int n_proxies;       // compound children bodies (after the n bodies)
};

// ---- checked kernel launches ----
// Every kernel launch goes through RB_LAUNCH: a grid or workgroup outside
// what a dispatch may carry (an empty grid, a dimension above 2^31-1, more
// than 1024 work-items per group) is refused on the host and named on
// stderr instead of reaching the queue (inside a captured graph it would
// become a dispatch packet).  rb_launch_status() then reports it as a failed
// launch (hipErrorInvalidConfiguration) once, else hipGetLastError().
bool rb_launch_ok(dim3 grid, dim3 block, const char *kernel);
hipError_t rb_launch_status();
#define RB_LAUNCH(kernel, grid, block, shmem, stream, ...)                         \
    do {                                                                          \
        if (rb_launch_ok(dim3(grid), dim3(block), #kernel))                       \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);  \
    } while (0)

// ---- launch interfaces (implemented in the .hip files) ----
struct SortWork {
    uint2 *tmp;           // n elements
    uint32_t *hist;       // passes x radix counters
    uint32_t *status;     // look-back words: passes x tiles x radix
    uint32_t *tile_ctr;   // passes counters
    uint32_t *scan_status;// scan look-back words
    int *err = nullptr;   // overflow word that gets MI_RB_OVERFLOW_LOOKBACK on a look-back timeout
    int max_n;
    int max_tiles;
    // hist and tile_ctr are zero on entry (a step clears one head per sort
    // in its first kernel), so no clearing launch precedes the sort
    bool precleared = false;
};
// A sort's digit counters and tile counters (RB_SORT_HEAD_WORDS words:
// 8 passes x 256 digits, then 8 tile counters); a step keeps one head per
// sort after its scans' status regions.
#define RB_SORT_HEAD_WORDS (8 * 256 + 16)
enum {
    RB_STEP_SORT_BODY = 0, RB_STEP_SORT_CAND = 1, RB_STEP_SORT_TILE = 2, RB_STEP_SORT_ORDER = 3,
    RB_STEP_SORTS = 4
};
size_t sort_status_words(int max_n, int digit_bits);
// n is the capacity; when n_dev is given the kernels read the element count
// from device memory (min(n, *n_dev)), so the launch needs no host sync.
// tile_keys: 1024 or 4096 elements per tile, 0 = chosen by n.
hipError_t launch_radix_sort(uint2 *kv, int n, int sort_bits, int digit_bits, const SortWork &w,
                             hipStream_t s, const int *n_dev = nullptr, int tile_keys = 0);
// out (may be null): exclusive prefix sums of in; idx_out (may be null, in
// holding 0/1 flags): the indices of the set flags in increasing order, from
// idx_out[*idx_base] when idx_base (a device count) is given.  precleared:
// the status words are already zero (a step clears the regions of its scans
// in its first kernel), so no clearing launch precedes the scan.
hipError_t launch_exclusive_scan(const uint32_t *in, uint32_t *out, int n, uint32_t *total,
                                 uint32_t *status, hipStream_t s, const int *n_dev = nullptr,
                                 int *err = nullptr, uint32_t *idx_out = nullptr,
                                 const uint32_t *idx_base = nullptr, bool precleared = false);
hipError_t launch_compact_flags(const uint32_t *flags, uint32_t *idx_out, int n, uint32_t *total,
                                uint32_t *status, hipStream_t s, const int *n_dev, int *err,
                                const uint32_t *idx_base = nullptr, bool precleared = false);
// the step's scans, each on its own region of the scan status words
enum {
    RB_SCAN_PAIRS = 0, RB_SCAN_CONTACTS = 1, RB_SCAN_COMPOUND = 2,
    RB_SCAN_WORK = 3, RB_SCAN_SMALL = 4, RB_SCAN_CAND = 5, RB_SCAN_REGIONS = 6
};
hipError_t launch_bound_search(const uint2 *kv, int n, uint32_t *dst, int n_dst, int mode,
                               hipStream_t s);

hipError_t launch_setup(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_cell_start(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_find_pairs(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_emit_pairs(const DeviceWorld &d, const StepParams &p, hipStream_t s);
// compound pairs: proxy transforms, then one child pair per (child of A,
// child of B) appended after the pairs (the compound pair itself gets none)
// scan_status, scan_stride: the step's scan regions (zeroed by its first
// kernel), region r at scan_status + r * scan_stride
hipError_t launch_compound_pairs(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                                 uint32_t *scan_status, size_t scan_stride);
hipError_t launch_narrowphase(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                              const SortWork &sw, uint32_t *scan_status, size_t scan_stride);
hipError_t launch_constraints(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_tile_bounds(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_tiles(const DeviceWorld &d, const StepParams &p, hipStream_t s);
// scene tiles: the solve's workgroup order, largest scene first
hipError_t launch_tile_order(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                             const SortWork &sw);
// kev: with timing, events recorded before and after every k_tile_solve
// launch (2 per launch, at least 2 * (4 * (sweeps + 1)) of them), else null;
// *n_launches gets the number of launches
// span: two events recorded before the first and after the last solver
// launch (the launches run back to back), else null
hipError_t launch_solve(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                        hipEvent_t *kev = nullptr, int *n_launches = nullptr,
                        hipEvent_t *span = nullptr);
// worlds with compound children: the impulse cache written by the solve in
// contact order, sorted by key into prev_* (a stable sort by y, then by x)
hipError_t launch_cache_order(const DeviceWorld &d, const StepParams &p, hipStream_t s,
                              const SortWork &sw);
hipError_t launch_copy_transforms(const DeviceWorld &d, int n, float4 *dst, hipStream_t s);
hipError_t launch_integrate(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
// elements per tile of the radix sort (4096, or 1024 for small sorts: the
// status arrays are sized for the small tile) and of the scan (their
// look-back status arrays hold one entry per tile)
constexpr int RB_SORT_TILE = 4096;
constexpr int RB_SORT_TILE_SMALL = 1024;
constexpr int RB_SCAN_TILE = 4096;
constexpr int RB_LARGE_CHUNK = 4096;  // bodies per chunk of the large-body pair search
inline int rb_sort_tiles(int n) { return (n + RB_SORT_TILE_SMALL - 1) / RB_SORT_TILE_SMALL; }
inline size_t rb_scan_status_words(int n) { return 4 + 2 * (size_t)((n + RB_SCAN_TILE - 1) / RB_SCAN_TILE); }
hipError_t launch_mlp_warm(const DeviceWorld &d, const StepParams &p, hipStream_t s);
hipError_t launch_mlp_targets(const DeviceWorld &d, int capacity, float *x_out, float *t_out,
                              hipStream_t s);

// the stand-alone grid API.  lane_hits (may be null): RB_HG_LANES x (1 +
// RB_HG_KEEP) words per body, where the count pass leaves each lane's pairs
// for the write pass
#define RB_HG_LANES 8
#define RB_HG_KEEP 2
hipError_t launch_hgrid_pairs(const float4 *mn, const float4 *mx, int n, float inv_cell0,
                              const int dims[3], int cell_bits, int levels, uint2 *kv,
                              uint32_t *cell_start, uint32_t *occ, float4 *smn, float4 *smx,
                              uint32_t *counts, uint32_t *offsets, uint32_t *cursor,
                              uint32_t *big, uint32_t *total, uint32_t *ys, int max_pairs,
                              int2 *pairs, int *num_pairs, int *overflow, const SortWork &sw,
                              uint32_t *scan_status, hipStream_t s, hipEvent_t *ev,
                              uint32_t *lane_hits);
hipError_t launch_sap_pairs(const float4 *mn, const float4 *mx, const int *is_large, int n,
                            int axis, uint2 *kv, float4 *smn, float4 *smx, uint32_t *counts,
                            uint32_t *offsets, uint32_t *cursor, uint32_t *big, uint32_t *lg,
                            uint32_t *total, uint32_t *ys, int max_pairs, int2 *pairs,
                            int *num_pairs, int *overflow, const SortWork &sw,
                            uint32_t *scan_status, hipStream_t s, hipEvent_t *ev);

int rb_num_cus();
