// dmath.h -- float4 helpers for the step kernels.
//
// The arithmetic is written out operation by operation, in the order of the
// reference's OpenCL helpers (dot3F4 = ((x*x + y*y) + z*z), qtMul, qtRotate,
// sat.cl:118-201), and the kernels are compiled with -ffp-contract=off, so a
// kernel and a serial restatement of the same formulas round identically.
// sqrt and divide are the IEEE operations (correctly rounded on gfx950 by
// default); fast_normalize/native_* of the reference are replaced by them.
#pragma once
#include <hip/hip_runtime.h>

namespace rb {

__device__ __forceinline__ float4 mk4(float x, float y, float z, float w) {
    float4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r;
}
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 sub4(float4 a, float4 b) {
    return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
__device__ __forceinline__ float4 neg4(float4 a) { return mk4(-a.x, -a.y, -a.z, -a.w); }
__device__ __forceinline__ float4 scl4(float4 a, float s) {
    return mk4(a.x * s, a.y * s, a.z * s, a.w * s);
}
__device__ __forceinline__ float dot3(float4 a, float4 b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__device__ __forceinline__ float4 cross3(float4 a, float4 b) {
    return mk4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.f);
}
// quaternion product: cross(a,b) + (a.w*b + b.w*a), w = a.w*b.w - dot3(a,b)
__device__ __forceinline__ float4 qt_mul(float4 a, float4 b) {
    float4 c = cross3(a, b);
    float tx = a.w * b.x + b.w * a.x;
    float ty = a.w * b.y + b.w * a.y;
    float tz = a.w * b.z + b.w * a.z;
    return mk4(c.x + tx, c.y + ty, c.z + tz, a.w * b.w - dot3(a, b));
}
__device__ __forceinline__ float4 qt_inv(float4 q) { return mk4(-q.x, -q.y, -q.z, q.w); }
__device__ __forceinline__ float4 qt_rotate(float4 q, float4 v) {
    float4 vc = mk4(v.x, v.y, v.z, 0.f);
    return qt_mul(qt_mul(q, vc), qt_inv(q));
}
__device__ __forceinline__ float4 qt_inv_rotate(float4 q, float4 v) {
    return qt_rotate(qt_inv(q), v);
}
__device__ __forceinline__ float4 xform(float4 p, float4 t, float4 q) {
    return add4(qt_rotate(q, p), t);
}
__device__ __forceinline__ float4 normalize3(float4 a) {
    float l2 = dot3(a, a);
    float inv = 1.0f / sqrtf(l2);
    return mk4(a.x * inv, a.y * inv, a.z * inv, 0.f);
}
// rotation matrix with element-wise fabs (broadphaseKernel.cl:92-113)
__device__ __forceinline__ void abs_rot(float4 q, float4 m[3]) {
    float x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z;
    m[0] = mk4(fabsf(1 - 2 * y2 - 2 * z2), fabsf(2 * q.x * q.y - 2 * q.w * q.z),
               fabsf(2 * q.x * q.z + 2 * q.w * q.y), 0.f);
    m[1] = mk4(fabsf(2 * q.x * q.y + 2 * q.w * q.z), fabsf(1 - 2 * x2 - 2 * z2),
               fabsf(2 * q.y * q.z - 2 * q.w * q.x), 0.f);
    m[2] = mk4(fabsf(2 * q.x * q.z - 2 * q.w * q.y), fabsf(2 * q.y * q.z + 2 * q.w * q.x),
               fabsf(1 - 2 * x2 - 2 * y2), 0.f);
}
__device__ __forceinline__ float4 mt_mul1(const float4 m[3], float4 b) {
    return mk4(dot3(m[0], b), dot3(m[1], b), dot3(m[2], b), 0.f);
}
__device__ __forceinline__ float4 mt_mul3(float4 a, const float4 m[3]) {
    float4 cx = mk4(m[0].x, m[1].x, m[2].x, 0.f);
    float4 cy = mk4(m[0].y, m[1].y, m[2].y, 0.f);
    float4 cz = mk4(m[0].z, m[1].z, m[2].z, 0.f);
    return mk4(dot3(a, cx), dot3(a, cy), dot3(a, cz), 0.f);
}

__device__ __forceinline__ int aabb_overlap(float4 mn0, float4 mx0, float4 mn1, float4 mx1) {
    return (mn0.x <= mx1.x) && (mn1.x <= mx0.x) && (mn0.y <= mx1.y) && (mn1.y <= mx0.y) &&
           (mn0.z <= mx1.z) && (mn1.z <= mx0.z);
}

// wave64 helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << __lane_id()) - 1ull;
}
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

}  // namespace rb
