/*
 * rb_oracle.h -- CPU oracle for the GPU rigid-body step (test infrastructure).
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it, and only as the
 * checker: the product path never calls it and never falls back to it.
 *
 * Every function restates the reference algorithm for one stage of
 * CLPhysicsDemo::stepSimulation (opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.cpp:740-943)
 * in plain C99, in a fixed serial order, and cites the reference file:line
 * it follows.  Floating point follows the reference's operation order with
 * IEEE sqrt/divide/sin/cos in place of the OpenCL native_* / fast_normalize
 * builtins; compiled with -ffp-contract=off so no multiply-add is fused.
 *
 * Array conventions (shared with the HIP library, plain pointers only):
 *   float4 arrays are float[4*n]; 3x3 matrices are 3 float4 rows (float[12]).
 *   quaternions are (x, y, z, w).
 */
#ifndef RB_ORACLE_H
#define RB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Convex hull record (80 bytes), same binary layout as the library's. */
typedef struct {
    int32_t face_offset, num_faces, vertex_offset, num_vertices;
    int32_t edge_offset, num_edges, index_offset, num_indices;
    float center[4];   /* local reference centre: average of the vertices */
    float aabb_min[4]; /* local AABB of the vertices (no margin) */
    float aabb_max[4];
} orc_hull;

/* Contact manifold, Contact4 layout (dynamics/basic_demo/Stubs/AdlContact4.h:26-59), 96 bytes. */
typedef struct {
    float pos[4][4];    /* xyz world point, w = depth (negative = penetration) */
    float normal[4];    /* xyz normal (from B towards A negated: -sepAxis), w = number of points */
    uint32_t coeffs;    /* restitution | friction<<16 in 16-bit fixed point */
    int32_t pair_index; /* m_batchIdx: index of the broadphase pair */
    int32_t body_a;     /* negated when body A is static */
    int32_t body_b;
} orc_contact;

/* Constraint4 (dynamics/basic_demo/Stubs/AdlConstraint4.h:23-47), 176 bytes. */
typedef struct {
    float linear[4];       /* -normal, w = friction coefficient */
    float pos[4][4];
    float center[4];
    float jac_inv[4];
    float b[4];
    float lambda[4];
    float f_jac_inv[2];
    float f_lambda[2];
    uint32_t body_a, body_b, batch, pad;
} orc_constraint;

/* ---- registration ---------------------------------------------------------- */
void orc_register_inertia(float mass, const float local_min[3], const float local_max[3],
                          const float pos[3], const float quat[4], float margin, float out[4]);

/* ---- primitives ---------------------------------------------------------- */
int orc_num_threads(void);
void orc_radix_sort_kv(uint32_t *kv, int n, int sort_bits);
uint32_t orc_exclusive_scan(const uint32_t *in, uint32_t *out, int n);
void orc_bound_search(const uint32_t *kv, int n, uint32_t *dst, int n_dst, int mode);

/* ---- body setup / broadphase --------------------------------------------- */
void orc_world_aabbs(int n, const float *pos, const float *quat, const int *shape,
                     const orc_hull *hulls, float *aabb_min, float *aabb_max);
void orc_calc_hash(int n, const float *aabb_min, const float *aabb_max, const int *scene,
                   const int *is_large, float inv_cell, const int dims[3], int cell_bits,
                   uint32_t *kv);
int orc_find_pairs_grid(int n, const uint32_t *sorted_kv, const float *aabb_min,
                        const float *aabb_max, const int *scene, const int *is_large,
                        const int *scene_start, const int *scene_count, float inv_cell,
                        const int dims[3], int cell_bits, int max_pairs, int32_t *pairs);
int orc_find_pairs_brute(int n, const float *aabb_min, const float *aabb_max, const int *scene,
                         int max_pairs, int32_t *pairs);
int orc_sap_pairs(int n, const float *aabb_min, const float *aabb_max, int axis, int max_pairs,
                  int32_t *pairs);
void orc_setup_bodies(int n, float *linvel, const float *angvel, const float *quat,
                      const float *inv_mass, const float *inv_inertia_local,
                      float *inv_inertia_world, const float gravity[3], float dt);

/* ---- narrowphase ---------------------------------------------------------- */
/* edge_reject: 0 as the GPU kernel (edge axes never reject), 1 as the CPU
 * twin (an edge axis with disjoint projections separates the pair) */
int orc_sat(const orc_hull *hA, const orc_hull *hB, const float *verts, const float *planes,
            const float *edges, const float *posA, const float *ornA, const float *posB,
            const float *ornB, int edge_reject, float sep[4]);
int orc_clip_hull_hull(const float sep[4], const orc_hull *hA, const orc_hull *hB,
                       const float *verts, const float *planes, const int *face_index,
                       const int *indices, const float *posA, const float *ornA,
                       const float *posB, const float *ornB, float min_dist, float max_dist,
                       float *out /* 64 float4 */);
int orc_clip_face(const float *in, int n_in, const float nrm[4], float eq, float *out);
int orc_reduce_manifold(const float *pts, int n, const float normal[4], int idx[4]);
int orc_reduce_manifold_global(const float *pts, int n, const float normal[4], int idx[4]);
void orc_compound_proxies(int P, int n, const int32_t *proxy, const float *children, float *pos,
                          float *quat);
int orc_narrowphase(int num_pairs, const int32_t *pairs, const float *pos, const float *quat,
                    const float *inv_mass, const int *shape, const orc_hull *hulls,
                    const float *verts, const float *planes, const int *face_index,
                    const int *indices, const float *edges, float friction,
                    float max_dist, int sat_mode, int reduction_mode, orc_contact *slots,
                    int *has_contact);

/* ---- solver ---------------------------------------------------------------- */
/* the two friction tangents of a contact normal (btPlaneSpace1) */
void orc_plane_space(const float n[3], float t0[4], float t1[4]);
void orc_contact_to_constraint(int nc, const orc_contact *contacts, const float *pos,
                               const float *linvel, const float *angvel, const float *inv_mass,
                               const float *inv_inertia_world, float dt, float drift,
                               float coeff, float friction, orc_constraint *out);
int orc_color_contacts(int nc, const orc_constraint *cs, const float *inv_mass, int n_bodies,
                       int *color);
void orc_contact_features(int nc, const orc_contact *contacts, orc_constraint *cons,
                          const float *pos, const float *linvel, const float *angvel,
                          const float *inv_mass, int nprev, const int *prev_keys,
                          const float *prev_imp, const float *prev_pts, int apply_previous,
                          const int *contact_keys /* int2 per contact or NULL */,
                          float *features);
void orc_warm_from_prediction(int nc, orc_constraint *cons, const float *y, float friction);
void orc_solve(int nc, orc_constraint *cs, const int *order, float *linvel, float *angvel,
               const float *pos, const float *inv_mass, const float *inv_inertia_world,
               int normal_iters, int friction_iters, float friction, float tol,
               int *iters_used, float *residual);
void orc_solve_segments(int nseg, const int *seg_off, orc_constraint *cs, const int *order,
                        float *linvel, float *angvel, const float *pos, const float *inv_mass,
                        const float *inv_inertia_world, int normal_iters, int friction_iters,
                        float friction, float tol, int *iters_used, float *residual);
void orc_integrate(int n, float *pos, float *quat, const float *linvel, float *angvel,
                   const float *inv_mass, float dt, float damping, float max_ang);

/* residual of one sweep: max_rows |dlambda| * |1/jac_inv| / max(1, max_rows |lambda|) */
float orc_sweep_residual(float max_dlambda_scaled, float max_lambda);

/* ---- learned warm start (MLP 16-64-64-3, bf16 weights, fp32 accumulation) ---- */
void orc_mlp_forward(int n, const float *x, const uint16_t *w_bf16, const float *bias,
                     float *h1, float *h2, float *y);
void orc_mlp_backward(int n, const float *x, const uint16_t *w_bf16, const float *bias,
                      const float *h1, const float *h2, const float *y, const float *target,
                      float *grad_w, float *grad_b, float *loss);
float orc_bf16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);

#ifdef __cplusplus
}
#endif
#endif
