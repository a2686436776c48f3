/*
 * mi_rb.h -- C ABI of the MI355X (gfx950) rigid-body step.
 *
 * Drop-in replacement for the per-step GPU path behind the reference's
 * CLPhysicsDemo host API (opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.h:48-80)
 * and its primitive C API (opencl/c_api/btbDeviceCL.h, btcFindPairs.h).
 * Plain C types only; every call returns mi_rb_status.  One world per host
 * thread; calls on one world are not thread-safe.  Indices are dense,
 * 0-based, in registration order (as in the reference).
 */
#ifndef MI_RB_H
#define MI_RB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_RB_ABI_VERSION 1

typedef enum mi_rb_status {
    MI_RB_OK = 0,
    MI_RB_INVALID_ARG = 1,       /* bad pointer, index or size */
    MI_RB_CAPACITY = 2,          /* a configured capacity would be exceeded */
    MI_RB_DEVICE = 3,            /* HIP runtime error or no usable gfx950 device */
    MI_RB_UNSUPPORTED_SHAPE = 4, /* shape kinds this library does not build */
    MI_RB_STATE = 5,             /* call out of order (e.g. step before upload) */
    MI_RB_FORMAT = 6             /* a file that is malformed or of an unsupported kind */
} mi_rb_status;

/* warm start of the contact solver */
enum { MI_RB_WARM_ZERO = 0, MI_RB_WARM_PREVIOUS = 1, MI_RB_WARM_LEARNED = 2 };
/* body inertia at registration */
enum { MI_RB_INERTIA_REFERENCE = 0, MI_RB_INERTIA_BOX = 1 };
/* sticky overflow bits reported in mi_rb_stats.overflow_flags */
enum {
    MI_RB_OVERFLOW_PAIRS = 1,          /* global pair capacity */
    MI_RB_OVERFLOW_BODY_PAIRS = 2,     /* per-body pair list (max_pairs_per_body) */
    MI_RB_OVERFLOW_CONTACTS = 4,       /* contact capacity */
    MI_RB_OVERFLOW_COLORS = 8,         /* solver colouring ran out of colours */
    MI_RB_OVERFLOW_TILE = 16,          /* solver tile capacity (bodies or rows) */
    MI_RB_OVERFLOW_POLYGON = 32,       /* clipped polygon > 64 vertices */
    MI_RB_OVERFLOW_LEVEL = 64,         /* grid API: an AABB larger than the top level's cell */
    MI_RB_OVERFLOW_LOOKBACK = 128,     /* a sort or scan look-back gave up waiting (results wrong) */
    MI_RB_OVERFLOW_UNSUPPORTED_PAIR = 256 /* a concave mesh met a body that is not a convex hull
                                           * (sphere, plane, compound, mesh): no contacts */
};

typedef struct mi_rb_config {
    int max_bodies;            /* body capacity */
    int max_shapes;            /* hull capacity */
    int max_hull_vertices;     /* pooled hull vertices (faces/indices/edges scale with it) */
    int max_pairs;             /* broadphase pair capacity */
    int max_contacts;          /* contact manifold capacity */
    int max_pairs_per_body;    /* per-body pair list length (reference: 32-1) */
    float dt;                  /* step for gravity and integration */
    float solver_dt;           /* step in the Baumgarte bias; 0 = dt */
    float gravity[3];
    int solver_iterations;     /* normal sweeps, then as many friction sweeps (<= 4096) */
    float friction;            /* Coulomb coefficient (reference: 0.7) */
    float baumgarte;           /* position constraint coefficient (0.2) */
    float position_drift;      /* allowed penetration in the bias (0.005) */
    float contact_threshold;   /* keep clipped points with depth <= threshold (0.02) */
    float angular_damping;     /* angular velocity scale per step (0.99) */
    float max_angular_motion;  /* radians per step (0.25*pi) */
    float cell_size;           /* broadphase grid cell; 0 = 2 x max bounding radius */
    int grid_dims[3];          /* power of two per axis; 0 = derived from the scene */
    float residual_tolerance;  /* > 0: sweep until the residual drops below it */
    int warm_start;            /* MI_RB_WARM_* */
    int inertia_mode;          /* MI_RB_INERTIA_* */
    float tile_size;           /* solver tile edge in XZ; 0 = derived */
    int device;                /* HIP device ordinal */
    int sat_mode;              /* MI_RB_SAT_*: the kernel's SAT or the CPU twin's */
    int reduction_mode;        /* MI_RB_REDUCE_*: contact reduction of the clip kernel */
} mi_rb_config;

/* SAT of the hull narrowphase: as the GPU kernel (findSeparatingAxisKernel,
 * sat.cl:928, edge x edge axes never reject a pair), or as the CPU twin
 * (findSeparatingAxis, gpu_rigidbody_pipeline2/ConvexHullContact.cpp:326-453:
 * an edge axis with disjoint projections separates the pair) */
enum { MI_RB_SAT_GPU = 0, MI_RB_SAT_CPU_TWIN = 1 };
/* contact reduction to <= 4 points: extractManifoldSequential of
 * clipHullHullKernel (satClipHullContacts.cl:788-884, the default), or
 * extractManifoldSequentialGlobal of the split path (:678-767: up to 4 points
 * pass through, else the extremes along +-u, +-v below FLT_MIN with the
 * deepest point replacing the first if missing; points may repeat) */
enum { MI_RB_REDUCE_SEQUENTIAL = 0, MI_RB_REDUCE_SPLIT_PATH = 1 };

typedef struct mi_rb_stats {
    int num_bodies;
    int num_pairs;
    int num_contacts;
    int num_colors;            /* solver phases in the last step */
    int overflow_flags;        /* sticky MI_RB_OVERFLOW_* bits */
    int normal_iterations;     /* sweeps used in the last step */
    int friction_iterations;
    float normal_residual;     /* last sweep's residual */
    float friction_residual;
    float stage_ms[8];         /* setup, sort, pairs, narrowphase, constraints, colouring, solve, integrate */
    long long steps;
    /* scene tiles with a residual tolerance stop one by one: the sweeps above
     * are the maximum over scenes, these the mean over scenes with contacts */
    float mean_normal_iterations;
    float mean_friction_iterations;
    /* dynamic bodies of the solver's tiles, summed over tiles (a body in
     * several tiles counts in each): what one sweep gathers */
    int solve_bodies;
    /* the solver kernel's (k_tile_solve) launches of the last eager step:
     * with mi_rb_set_solve_timing, their number and the HIP-event span from
     * before the first to after the last (they run back to back); otherwise
     * the sum of HIP-event spans around each launch (which add the launch
     * overhead) */
    float solve_kernel_ms;
    int solve_launches;
    float solve_events_ms;
    /* narrowphase pairs that the quick face test (one face of each hull) did
     * not separate: the pairs of the wave-per-pair face pass */
    int face_pass_pairs;
} mi_rb_stats;

/* parameters derived at upload (grid, overflow list, solver tiles) */
typedef struct mi_rb_params {
    float cell_size, inv_cell;
    int dims[3];
    int cell_bits;      /* bits of the cell hash */
    int key_bits;       /* bits the cell-key sort orders by */
    int n_scenes, n_large;
    int tile_mode;      /* 0 spatial XZ tiles, 1 one tile per scene */
    int n_tiles, tiles_x, tiles_z;
    float tile_inv, tile_origin_x, tile_origin_z;
} mi_rb_params;

typedef struct mi_rb_device_view {
    void *pos, *quat, *linvel, *angvel; /* float4[num_bodies] */
    void *inv_mass;                     /* float[num_bodies] */
    void *pairs;                        /* int2[num_pairs], (x > y) */
    void *contacts;                     /* Contact4[num_contacts] */
    void *stream;                       /* hipStream_t of the world */
    int num_bodies;
} mi_rb_device_view;

/* debug/parity buffers readable after a step (mi_rb_read_buffer) */
enum {
    MI_RB_BUF_SORTED_KV = 0,   /* uint2[num_bodies] (cell key, body) after the radix sort */
    MI_RB_BUF_AABB = 1,        /* float4[2*num_bodies] min,max */
    MI_RB_BUF_PAIRS = 2,       /* int2[num_pairs] */
    MI_RB_BUF_CONTACTS = 3,    /* Contact4[num_contacts], ordered by pair */
    MI_RB_BUF_CONSTRAINTS = 4, /* Constraint4[num_contacts] after the solve */
    MI_RB_BUF_SOLVE_ORDER = 5, /* int[num_contacts] constraint visiting order of one sweep */
    MI_RB_BUF_COLORS = 6,      /* int2[num_contacts] (phase, colour) of each constraint */
    MI_RB_BUF_INV_INERTIA = 7, /* float4[3*num_bodies] world inverse inertia rows */
    MI_RB_BUF_PHASES = 8,      /* int4[n_tiles] (first, end, phase, colours) of the solve order */
    /* registered data (host copies, readable before upload) */
    MI_RB_BUF_HULLS = 9,       /* 80-byte hull records */
    MI_RB_BUF_VERTS = 10,      /* float4 */
    MI_RB_BUF_PLANES = 11,     /* float4 */
    MI_RB_BUF_FACE_INDEX = 12, /* int2 (first index, count) */
    MI_RB_BUF_INDICES = 13,    /* int */
    MI_RB_BUF_EDGES = 14,      /* float4 */
    MI_RB_BUF_INV_MASS = 15,   /* float per body */
    MI_RB_BUF_INV_INERTIA_LOCAL = 16, /* float4 per body (diagonal) */
    MI_RB_BUF_SCENE = 17,      /* int per body */
    MI_RB_BUF_IS_LARGE = 18,   /* int per body (overflow list), after upload */
    MI_RB_BUF_SHAPE = 19,      /* int per body */
    MI_RB_BUF_FEATURES = 20,   /* float[16] per contact: warm-start network inputs */
    MI_RB_BUF_PREV_KEYS = 21,  /* int2 per cached manifold, its pair entry (x, y) (compound
                                * child pairs name proxies): impulse cache of the last solve,
                                * sorted */
    MI_RB_BUF_PREV_IMPULSES = 22, /* float4 per cached pair: sum lambda_n, friction 0, 1, points */
    MI_RB_BUF_PREDICTIONS = 23, /* float[3] per contact: network output (learned warm start) */
    MI_RB_BUF_PREV_POINTS = 24, /* float4[4] per cached pair: point xyz, its lambda_n in w */
    MI_RB_BUF_CHILDREN = 25,   /* 48 B per compound child: float4 pos, float4 orn, int shape, pad */
    MI_RB_BUF_TILE_STATS = 26  /* int4 per solver tile (per scene in a batched world): normal and
                                * friction sweeps used and the final normal and friction
                                * residuals (float bits); x = -1 for a tile without rows.
                                * Written by steps with a tolerance on scene tiles only */
};

/* stages at which mi_rb_step_partial stops */
enum {
    MI_RB_STAGE_SETUP = 0, MI_RB_STAGE_SORT = 1, MI_RB_STAGE_PAIRS = 2,
    MI_RB_STAGE_NARROWPHASE = 3, MI_RB_STAGE_CONSTRAINTS = 4, MI_RB_STAGE_COLOR = 5,
    MI_RB_STAGE_SOLVE = 6, MI_RB_STAGE_INTEGRATE = 7
};

/* ---- world: replaces CLPhysicsDemo (CLPhysicsDemo.h:48-80) ---- */

/* Fills the reference preset: constants hard-coded in the reference kernels
 * (dt 0.0166666, gravity -9.8, 4 iterations, friction 0.7, Baumgarte 0.2,
 * drift 0.005, threshold 0.02, damping 0.99, 0.25*pi). */
void mi_rb_config_reference(mi_rb_config *cfg);

/* CLPhysicsDemo::CLPhysicsDemo + init (CLPhysicsDemo.cpp:174-182,603-655) */
mi_rb_status mi_rb_create(const mi_rb_config *cfg, void **out_world);
/* CLPhysicsDemo::cleanup (CLPhysicsDemo.cpp:668-690) */
mi_rb_status mi_rb_destroy(void *world);

/* CLPhysicsDemo::registerConvexPolyhedron(const float*, int strideBytes, int n,
 * const float* scaling, bool) (CLPhysicsDemo.cpp:491-530): convex hull of the
 * points (scaled) on btConvexHullComputer's integer grid (so the vertices
 * are the reference's), coplanar faces merged as btConvexUtility does;
 * returns the shape index.  Vertices are listed in input order, where the
 * reference lists them in its hull walk's order. */
mi_rb_status mi_rb_register_convex_hull(void *world, const float *xyz, int stride_bytes, int n,
                                        const float scale[3], int *out_shape);
/* CLPhysicsDemo::registerConvexPolyhedron(btConvexUtility*, bool)
 * (CLPhysicsDemo.cpp:331-410) with btGpuNarrowphaseAndSolver::
 * registerConvexHullShape (btGpuNarrowPhaseAndSolver.cpp:423-490): registers
 * polyhedral features built elsewhere (a btConvexUtility's) unchanged and in
 * their order: num_vertices points (xyz, vert_stride_bytes apart), num_faces
 * faces (plane nx, ny, nz, w as float4 in planes4, and (first index, count)
 * into indices in face_index2), num_indices vertex indices, num_edges unique
 * edge directions (xyz, edge_stride_bytes apart).  As in the reference the
 * centre is the vertex average and the local AABB that of the vertices.
 * MI_RB_INVALID_ARG for indices out of range, faces of fewer than 3 vertices
 * or values that are not finite; MI_RB_CAPACITY beyond 64 vertices, 128
 * faces, 64 vertices in a face or 128 edges. */
mi_rb_status mi_rb_register_convex_polyhedron(void *world, const float *verts,
                                              int vert_stride_bytes, int num_vertices,
                                              const float *planes4, const int *face_index2,
                                              int num_faces, const int *indices, int num_indices,
                                              const float *edges, int edge_stride_bytes,
                                              int num_edges, int *out_shape);
/* CLPhysicsDemo::registerSphereShape(radius) (CLPhysicsDemo.cpp:420-456): a
 * sphere; local AABB +-radius.  Contacts: sphere-sphere and sphere-plane,
 * one point each (satClipHullContacts.cl:1055-1118); a sphere against a hull
 * gets none, as in the reference. */
mi_rb_status mi_rb_register_sphere(void *world, float radius, int *out_shape);
/* CLPhysicsDemo::registerPlaneShape(normal, constant) (CLPhysicsDemo.cpp:459-489):
 * the plane n . x = constant, for static bodies only (a body with mass on it
 * gives MI_RB_INVALID_ARG); local AABB +-1e30.  Contacts with spheres only. */
mi_rb_status mi_rb_register_plane(void *world, const float normal[3], float constant,
                                  int *out_shape);
/* CLPhysicsDemo::registerCompoundShape(childShapes) (CLPhysicsDemo.h:66;
 * btGpuNarrowPhaseAndSolver::registerCompoundShape): a rigid union of n hull
 * shapes, child i at child_pos[3i..] with orientation child_orn[4i..] (xyzw)
 * in the compound's frame.  Each child pair is found and clipped like a hull
 * pair with the child's world transform (sat.cl:639-925,
 * satClipHullContacts.cl:1184-1310); the contact names the compound body.
 * At most 64 children, each a hull shape. */
mi_rb_status mi_rb_register_compound(void *world, int n, const int *child_shapes,
                                     const float *child_pos, const float *child_orn,
                                     int *out_shape);
/* CLPhysicsDemo::registerConcaveMesh(vertices, indices, scaling)
 * (CLPhysicsDemo.cpp:287-300, btGpuNarrowPhaseAndSolver.cpp:264-333): a
 * triangle mesh for static bodies only (a body with mass on it gives
 * MI_RB_INVALID_ARG).  num_vertices points (scaled) at stride_bytes apart,
 * triangle t = indices[3t..3t+2].  Each triangle becomes the polyhedron
 * sat.cl:1040-1116 builds (3 vertices, front and back faces, 3 edge faces);
 * per pair with a convex hull body, each triangle whose world AABB overlaps
 * the hull's gives a triangle-hull pair, found and clipped like a hull pair
 * (sat.cl:970-1191, satClipHullContacts.cl:1387); contacts name the mesh
 * body.  Triangles without area are left out; a mesh against a sphere, a
 * plane, a compound or a mesh gets no contacts and sets
 * MI_RB_OVERFLOW_UNSUPPORTED_PAIR.  Uses one shape slot and 3
 * hull vertices per triangle plus one slot for the mesh. */
#define MI_RB_MAX_MESH_TRIANGLES 65536
mi_rb_status mi_rb_register_concave_mesh(void *world, const float *xyz, int stride_bytes,
                                         int num_vertices, const int *indices, int num_triangles,
                                         const float scale[3], int *out_shape);
/* Wavefront .obj meshes (the reference's objLoader, rendering/
 * WavefrontObjLoader/obj_parser.cpp:19-105, as registerConcaveMeshShape(
 * objLoader*, ...) uses it, btGpuNarrowPhaseAndSolver.cpp:335-400): "v x y z"
 * lines and "f" lines of "v", "v/t", "v//n" or "v/t/n" tokens (1-based,
 * negative = counted back); each face gives the triangle of its first three
 * vertices, as in the reference.  mi_rb_obj_read returns the counts, and the
 * vertices (xyz) and triangles (3 indices each, 0-based) when both arrays
 * are given and large enough (MI_RB_CAPACITY otherwise); a malformed or
 * missing file gives MI_RB_INVALID_ARG.  Host only. */
mi_rb_status mi_rb_obj_read(const char *path, float *xyz, int max_vertices, int *indices,
                            int max_triangles, int *num_vertices, int *num_triangles);
/* CLPhysicsDemo::registerConcaveMesh of an .obj file (the demo's ground mesh,
 * BulletDataExtractor.cpp:262-310): mi_rb_register_concave_mesh of what
 * mi_rb_obj_read gives. */
mi_rb_status mi_rb_register_concave_mesh_obj(void *world, const char *path, const float scale[3],
                                             int *out_shape);
/* Shape kinds this library does not build: returns MI_RB_UNSUPPORTED_SHAPE. */
mi_rb_status mi_rb_register_unsupported_shape(void *world, int kind);

/* CLPhysicsDemo::registerPhysicsInstance(mass, pos, orn xyzw, collidable, user)
 * (CLPhysicsDemo.cpp:533-600); bodies of one scene must be registered
 * contiguously, scenes in increasing order. */
mi_rb_status mi_rb_register_body(void *world, float mass, const float pos[3],
                                 const float quat_xyzw[4], int shape, int scene, int *out_body);

/* CLPhysicsDemo::writeBodiesToGpu (CLPhysicsDemo.cpp:205-215): uploads shapes
 * and bodies, allocates the step's device buffers. */
mi_rb_status mi_rb_upload(void *world);
/* CLPhysicsDemo::setObjectTransform / setObjectLinearVelocity (CLPhysicsDemo.cpp:692-701) */
mi_rb_status mi_rb_set_transform(void *world, int body, const float pos[3], const float quat[4]);
mi_rb_status mi_rb_set_linvel(void *world, int body, const float v[3]);
/* bulk state transfer (float4 arrays; NULL skips a field) */
mi_rb_status mi_rb_set_body_state(void *world, const float *pos4, const float *quat4,
                                  const float *linvel4, const float *angvel4);
mi_rb_status mi_rb_get_body_state(void *world, float *pos4, float *quat4, float *linvel4,
                                  float *angvel4);

/* CLPhysicsDemo::stepSimulation (CLPhysicsDemo.cpp:740-943), nsteps times;
 * enqueued on the world's stream (one captured graph per step when enabled). */
mi_rb_status mi_rb_step(void *world, int nsteps);
/* one step that stops after `stage` (MI_RB_STAGE_*), for stage parity tests */
mi_rb_status mi_rb_step_partial(void *world, int stage);
/* capture the step in a hipGraph and replay it (1) or launch eagerly (0) */
mi_rb_status mi_rb_use_graph(void *world, int enable);

/* CLPhysicsDemo::readbackBodiesToCpu + getObjectTransformFromCpu (CLPhysicsDemo.cpp:190-203) */
mi_rb_status mi_rb_readback(void *world);
mi_rb_status mi_rb_get_transform(void *world, int body, float pos[4], float quat[4]);

mi_rb_status mi_rb_get_stats(void *world, mi_rb_stats *out);
mi_rb_status mi_rb_get_params(void *world, mi_rb_params *out);
/* CLPhysicsDemo::getBodiesGpu (CLPhysicsDemo.cpp:945): non-owning device view */
mi_rb_status mi_rb_device_ptrs(void *world, mi_rb_device_view *out);
/* copyTransformsToBVO (opencl/broadphase_benchmark/findPairsOpenCL.cpp:140-160,
 * kernel broadphaseKernel.cl:236-247; called by gpu_rigidbody_pipeline2/main.cpp:274-278):
 * writes the bodies' transforms into a device-visible vertex buffer of float4
 * (a renderer's interop buffer): (x, y, z, 1) of body i at float4 index
 * start_offset/4 + i and its quaternion (x, y, z, w) at start_offset/4 + n + i.
 * start_offset counts floats (a multiple of 4); capacity_float4 is the
 * buffer's size.  Enqueued on `stream` (NULL: the world's stream) after the
 * world's pending steps; MI_RB_CAPACITY if the buffer is too small. */
mi_rb_status mi_rb_copy_transforms(void *world, void *d_vertex_buffer, int start_offset,
                                   int capacity_float4, void *stream);
/* copy a step buffer (MI_RB_BUF_*) to host; *count = elements available */
mi_rb_status mi_rb_read_buffer(void *world, int which, void *host, size_t max_bytes, int *count);
/* versioned binary snapshot of shapes + bodies + velocities.  load_state
 * checks the whole file first (array lengths, offsets and indices inside
 * the arrays, capacities of the world's configuration): a truncated or
 * inconsistent file gives MI_RB_INVALID_ARG (MI_RB_CAPACITY beyond the
 * configuration) and leaves the world as it was. */
mi_rb_status mi_rb_save_state(void *world, const char *path);
mi_rb_status mi_rb_load_state(void *world, const char *path);
/* The same check without a world (host only): counts = {bodies, shapes,
 * vertices, compound children} of a valid file. */
mi_rb_status mi_rb_snapshot_check(const char *path, const mi_rb_config *cfg, int counts[4]);

/* ---- .bullet scene files ----
 * The reference reads Bullet's serialized scenes with bParse::btBulletFile and
 * converts them with btBulletDataExtractor (gpu_rigidbody_pipeline2/
 * BulletDataExtractor.cpp:454-520 createScene, :594-700 convertAllObjects,
 * :703-1010 convertCollisionShape, createBoxShape).  Float files with 4- or
 * 8-byte pointers in either byte order; every struct layout is taken from
 * the file's own DNA.  Double-precision files give MI_RB_FORMAT, as the
 * reference's demo refuses them.  Host code: no device work. */
enum {
    MI_RB_SHAPE_BOX = 0, MI_RB_SHAPE_CONVEX_HULL = 4, MI_RB_SHAPE_SPHERE = 8,
    MI_RB_SHAPE_TRIANGLE_MESH = 21, MI_RB_SHAPE_STATIC_PLANE = 28, MI_RB_SHAPE_COMPOUND = 31
};
typedef struct mi_rb_bullet_shape {
    int type;                /* Bullet's shape type number (MI_RB_SHAPE_* and others) */
    int num_points;          /* convex hull: unscaled points, from first_point of the points */
    int first_point;
    float local_scaling[3];
    float implicit_dims[3];  /* box: half extents (without margin); sphere: radius in [0] */
    float margin;
    float plane_normal[3];   /* static plane */
    float plane_constant;
} mi_rb_bullet_shape;
typedef struct mi_rb_bullet_body {
    int shape;               /* index into the file's shapes, -1 if none */
    float inverse_mass;
    float basis[9];          /* world rotation, rows */
    float origin[3];
    float quat[4];           /* the rotation as x, y, z, w */
    float linvel[3], angvel[3];
    float friction, restitution;
} mi_rb_bullet_body;
typedef struct mi_rb_bullet_summary {
    int shapes, shapes_registered, unsupported_shapes;
    int bodies, bodies_registered, ground_bodies;
    int pointer_size, big_endian;
} mi_rb_bullet_summary;
/* Parse only: up to max_* shapes, bodies and points (xyz); counts = {shapes,
 * bodies, points, pointer size | 0x100 if big-endian}.  Any array may be NULL. */
mi_rb_status mi_rb_bullet_read(const char *path, mi_rb_bullet_shape *shapes, int max_shapes,
                               mi_rb_bullet_body *bodies, int max_bodies, float *points,
                               int max_points, int counts[4]);
/* btBulletDataExtractor::convertAllObjects into a world before mi_rb_upload:
 * boxes (the unit cube scaled by half extent * scaling + margin) and convex
 * hulls (points * scaling) are registered in file order, other shape types
 * are counted in unsupported_shapes; bodies are registered shape by shape
 * (mass 1/inverse mass, rotation from the basis), static ones only with
 * keep_static; without keep_static a 50 x 50 grid of static pyramids is
 * added as the ground, as the reference does. */
mi_rb_status mi_rb_load_bullet(void *world, const char *path, int keep_static,
                               mi_rb_bullet_summary *out);
/* Learned warm start (no reference counterpart: the reference zeroes lambda
 * every step, dynamics/basic_demo/Stubs/solverSetup.cl:529-531).  Weights of
 * the 16-64-64-3 network used when warm_start == MI_RB_WARM_LEARNED: bf16
 * W1[64][16], W2[64][64], W3[3][64] (out-major, 5312 values) and fp32 biases
 * b1[64], b2[64], b3[3] (131 values), host pointers. */
mi_rb_status mi_rb_set_warmstart_weights(void *world, const uint16_t *w_bf16, const float *bias);
/* Training pairs of the last step, into caller device buffers: per contact the
 * 16 features (x_dev, float[16]) and the solved impulses in network units
 * (t_dev, float[3]: 10 x (sum of normal lambdas, friction 0, friction 1)).
 * *count = contacts written (at most capacity). */
mi_rb_status mi_rb_copy_training_batch(void *world, float *x_dev, float *t_dev, int capacity,
                                       int *count);
/* Solver settings after creation (for comparing warm starts on one state):
 * sweeps per phase (max with a tolerance), residual tolerance (0 = fixed
 * sweeps), MI_RB_WARM_*.  Drops the captured graph. */
mi_rb_status mi_rb_set_solver(void *world, int iterations, float tolerance, int warm_start);
/* HIP-event time of the last `steps` steps, milliseconds */
mi_rb_status mi_rb_last_step_ms(void *world, float *ms);
/* Measurement (no reference counterpart): with enable, every eager step
 * records HIP events before the first and after the last launch of the
 * solver kernel, which run back to back, so mi_rb_get_stats reports their
 * span (solve_kernel_ms) and number (solve_launches) for the last step.
 * Drops the captured graph. */
mi_rb_status mi_rb_set_solve_timing(void *world, int enable);

/* Launch shapes of the step (no reference counterpart: the reference fixes
 * its work-group sizes in the host code).  0 = default.  Changing them drops
 * the captured graph; results do not depend on them. */
typedef struct mi_rb_tuning {
    int solve_block;     /* threads per scene-tile solver workgroup: 256 or 512 */
    int narrow_waves;    /* narrowphase waves per SIMD: 1..8 */
    int sort_tile;       /* body-key sort, keys per tile: 1024 or 4096; 0 = by key count */
    int sort_digit_bits; /* body-key sort, bits per radix digit: 4, 6, 7 or 8; 0 = 8 */
} mi_rb_tuning;
mi_rb_status mi_rb_set_tuning(void *world, const mi_rb_tuning *t);
mi_rb_status mi_rb_get_tuning(void *world, mi_rb_tuning *t);
/* Autotuner: times every candidate launch shape over `steps` graph steps from
 * the world's current state (restored after each trial, so the world ends
 * where it started), applies the fastest and returns it.  ms_out (may be
 * NULL) receives the per-candidate step times in the order solve_block in
 * {256, 512} x narrow_waves in {2, 4, 6, 8}.  With the fastest of those, the
 * body-key sort's sort_tile in {1024, 4096} x sort_digit_bits in {7, 8} is
 * timed the same way and the fastest kept (those times are not reported). */
mi_rb_status mi_rb_autotune(void *world, int steps, mi_rb_tuning *best, float ms_out[8]);

/* ---- host utility (no device work): convex hull + polyhedral features ---- */
/* Builds the hull record used by registration, for inspection and tests.
 * Arrays are written up to the given capacities; counts are returned. */
mi_rb_status mi_rb_hull_build(const float *xyz, int stride_bytes, int n, const float scale[3],
                              float *verts4, int max_verts, float *planes4, int *face_index2,
                              int max_faces, int *indices, int max_indices, float *edges4,
                              int max_edges, int counts[4]);

/* ---- primitives: replace btbDeviceCL.h / btRadixSort32CL / btPrefixScanCL / btBoundSearchCL ---- */
/* All take device pointers and run on `stream` (hipStream_t, NULL = default). */

/* btbCreateRadixSort (c_api/btbDeviceCL.h:37) */
mi_rb_status mi_rb_radix_sort_create(int max_elements, void **out_sort);
mi_rb_status mi_rb_radix_sort_destroy(void *sort);
/* btbSort / btRadixSort32CL::execute (btRadixSort32CL.cpp:161-535): stable
 * sort of (uint32 key, uint32 value) pairs (btSortData) by the low sort_bits
 * key bits, in place. */
mi_rb_status mi_rb_radix_sort_kv(void *sort, void *d_keyvalue, int n, int sort_bits,
                                 void *stream);
/* mi_rb_radix_sort_kv without a sort handle and without allocating: no
 * counterpart in the reference.  For callers that cannot call hipMalloc
 * (code under stream capture, tight serving loops, hosts that forbid device
 * allocation).  The workspace is a file-scope device array of about 9 MiB
 * that each GPU the library runs on holds until the process ends; its
 * address is looked up once per device.  n above MI_RB_SORT_NOALLOC_MAX
 * gives MI_RB_CAPACITY.  Every call on one GPU shares that GPU's workspace,
 * so calls that overlap on one GPU (concurrent streams or host threads)
 * race: serialise them.  Enqueued on `stream` of the current device. */
#define MI_RB_SORT_NOALLOC_MAX (1 << 20)
mi_rb_status mi_rb_radix_sort_kv_noalloc(void *d_keyvalue, int n, int sort_bits, void *stream);
/* btPrefixScanCL::execute (btPrefixScanCL.cpp:46-99): exclusive scan of uint32;
 * the total is written to d_total when non-NULL (device pointer). */
mi_rb_status mi_rb_exclusive_scan_u32(void *sort, const void *d_in, void *d_out, int n,
                                      void *d_total, void *stream);
/* btBoundSearchCL::execute (btBoundSearchCL.cpp:78-136): mode 0 lower, 1 upper,
 * 2 count over sorted (key,value) pairs into d_dst[n_dst]. */
mi_rb_status mi_rb_bound_search(const void *d_keyvalue, int n, void *d_dst, int n_dst, int mode,
                                void *stream);

/* ---- grid pair search: replaces btcFindPairs.h over device AABBs ---- */
/* Hierarchical uniform grid: level l has cells of cell_size * 2^l (levels
 * 1..4); an AABB lives on the lowest level whose cell is at least its largest
 * extent, so bodies of very different sizes need no overflow list. */
typedef struct mi_rb_grid_desc {
    float cell_size;   /* level-0 cell edge */
    int dims[3];       /* cells per axis (power of two; positions wrap) */
    int max_pairs;
    int max_pairs_per_body; /* unused by the hierarchical search (no per-body cap) */
    int levels;        /* 1..4 (0 = 1) */
} mi_rb_grid_desc;
mi_rb_status mi_rb_grid_create(const mi_rb_grid_desc *desc, int max_aabbs, void **out_grid);
mi_rb_status mi_rb_grid_destroy(void *grid);
/* Overlapping pairs of n AABBs (float4 min[n], max[n], device): int2 (x > y)
 * sorted by (x, y) to d_pairs, the count to d_num_pairs (device int).
 * d_is_large must be NULL (the levels replace the overflow list).  Overflow
 * bits (pairs beyond max_pairs, AABBs beyond the top level) accumulate in the
 * grid; mi_rb_grid_overflow reads and clears them. */
mi_rb_status mi_rb_grid_find_pairs(void *grid, const void *d_aabb_min, const void *d_aabb_max,
                                   const void *d_is_large, int n, void *d_pairs, void *d_num_pairs,
                                   void *stream);
mi_rb_status mi_rb_grid_overflow(void *grid, int *flags);
/* time of the last find_pairs per phase (hash, sort, cells+pairs, scan+emit), ms */
mi_rb_status mi_rb_grid_last_ms(void *grid, float ms[4]);

/* ---- sweep-and-prune pair search over device AABBs ----
 * Replaces btGpuSapBroadphase (opencl/gpu_rigidbody_pipeline2/btGpuSapBroadphase.cpp:91-434:
 * createProxy / calculateOverlappingPairs / getOverlappingPairBuffer; kernels
 * opencl/broadphase_benchmark/sap.cl:90-116, 290-320): AABB mins along `axis`
 * as FloatFlip keys, a stable 32-bit radix sort, then each AABB tests the
 * following ones until their min passes its max.  Unlike the reference
 * (atomic append) the output is the same as the grid's: every overlapping
 * pair once as int2 (x, y), x > y, sorted by (x, y).  As in the reference,
 * AABBs flagged large are tested against all others instead of sweeping.  Sticky
 * MI_RB_OVERFLOW_PAIRS when max_pairs is exceeded. */
typedef struct mi_rb_sap_desc {
    int axis;       /* 0, 1 or 2 (the reference sweeps along x) */
    int max_pairs;  /* capacity of the pair output */
} mi_rb_sap_desc;
mi_rb_status mi_rb_sap_create(const mi_rb_sap_desc *desc, int max_aabbs, void **out_sap);
mi_rb_status mi_rb_sap_destroy(void *sap);
/* d_aabb_min/max: float4[n] device arrays; d_is_large: int[n] device array or
 * NULL, nonzero for AABBs that do not sweep but are tested against every AABB
 * (the reference's large proxies, e.g. a ground); d_pairs int2[max_pairs];
 * d_num_pairs int[1]; enqueued on `stream` (NULL: the null stream) */
mi_rb_status mi_rb_sap_find_pairs(void *sap, const void *d_aabb_min, const void *d_aabb_max,
                                  const void *d_is_large, int n, void *d_pairs,
                                  void *d_num_pairs, void *stream);
mi_rb_status mi_rb_sap_overflow(void *sap, int *flags);
/* milliseconds of the last call: keys, sort, sweep (count, scan, write), pair order */
mi_rb_status mi_rb_sap_last_ms(void *sap, float ms[4]);

/* ---- learned warm-start network: training and inference (no reference counterpart) ---- */
/* Network handle with He-uniform weights from `seed`, Adam state and scratch. */
mi_rb_status mi_rb_mlp_create(uint64_t seed, void **out_mlp);
void mi_rb_mlp_destroy(void *mlp);
/* host copies of the bf16 weights / fp32 biases (layout as mi_rb_set_warmstart_weights);
 * set also resets the Adam moments */
mi_rb_status mi_rb_mlp_set_weights(void *mlp, const uint16_t *w_bf16, const float *bias);
mi_rb_status mi_rb_mlp_get_weights(void *mlp, uint16_t *w_bf16, float *bias);
/* y_dev[n][3] = network(x_dev[n][16]) (device pointers, synchronous) */
mi_rb_status mi_rb_mlp_forward(void *mlp, const float *x_dev, int n, float *y_dev);
/* gradients of loss = mean_n sum_o (y_o - t_o)^2 / 3 with respect to the
 * weights and biases (grad_dev: float[5312 + 131], device) and the loss (host) */
mi_rb_status mi_rb_mlp_gradients(void *mlp, const float *x_dev, const float *t_dev, int n,
                                 float *grad_dev, float *loss);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on minibatches of `batch`
 * rows drawn with replacement from n_samples (device x[16], t[3]); seed < 0
 * takes rows 0..batch-1 in order instead (batch <= n_samples).  The loss of
 * each step goes to loss_history[steps] (host, may be NULL). */
mi_rb_status mi_rb_mlp_train(void *mlp, const float *x_dev, const float *t_dev, int n_samples,
                             int batch, int steps, float lr, int seed, float *loss_history);

#ifdef __cplusplus
}
#endif
#endif
