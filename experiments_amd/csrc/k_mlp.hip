// k_mlp.hip -- learned warm start of the contact solver: an MLP 16-64-64-3
// that maps per-contact features to the contact's initial normal and two
// friction impulses.  No reference counterpart: the reference solver zeroes
// lambda every step (dynamics/basic_demo/Stubs/solverSetup.cl:529-531).
//
// Forward: BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation.  A
// wavefront owns 16 contacts at a time and computes the layers transposed,
// H^T = W . X^T, so that the 16x16 accumulator of one layer (lane l holds rows
// 4*(l/16)+j, column l%16) is exactly the B operand layout of the next layer:
// activations never leave registers and the 5,312 weights stay resident in
// 48 VGPRs per lane for the whole grid-stride loop.
//
// Backward: FP32 MFMA (v_mfma_f32_16x16x4_f32) for the delta propagation
// (d1^T = W2^T . d2^T) and the three weight-gradient GEMMs, which reduce over
// the 64 contacts of a workgroup tile staged in LDS.  Each workgroup writes
// its partial gradients; one reduction kernel sums them in a fixed order
// (deterministic) and applies Adam to FP32 master weights, rounding the
// working copy to BF16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "rb_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NW1 = RB_MLP_H * RB_MLP_IN;   // 1024
constexpr int NW2 = RB_MLP_H * RB_MLP_H;    // 4096
constexpr int NW3 = RB_MLP_OUT * RB_MLP_H;  // 192
constexpr int NW = RB_MLP_NW;
constexpr int NB = RB_MLP_NB;
constexpr int PSTRIDE = NW + NB + 1;        // partial gradients + loss per workgroup
constexpr int MT = 64;                      // contacts per backward workgroup tile
constexpr int GRAD_BLOCKS = 256;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float rbf(float x) { return bf2f(f2bf(x)); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r.x = (short)f2bf(a); r.y = (short)f2bf(b); r.z = (short)f2bf(c); r.w = (short)f2bf(d);
    return r;
}
__device__ __forceinline__ bf16x4 load4(const uint16_t *p) {
    uint2 u = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r.x = (short)(u.x & 0xffffu); r.y = (short)(u.x >> 16);
    r.z = (short)(u.y & 0xffffu); r.w = (short)(u.y >> 16);
    return r;
}
__device__ __forceinline__ f32x4 mfma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 zero4() { f32x4 z = {0.f, 0.f, 0.f, 0.f}; return z; }

// Weights of one wavefront in MFMA A-operand layout: lane l holds rows
// 16*mb + l%16 and the 4 consecutive inputs 16*kb + 4*(l/16) .. +3.
struct MlpRegs {
    bf16x4 w1[4];
    bf16x4 w2[4][4];
    bf16x4 w3[4];
    float b1[4][4], b2[4][4], b3[4];
};

__device__ __forceinline__ void mlp_load(MlpRegs &m, const uint16_t *W, const float *B, int lane) {
    const int r = lane & 15, q = lane >> 4;
    for (int mb = 0; mb < 4; mb++) m.w1[mb] = load4(W + (16 * mb + r) * RB_MLP_IN + 4 * q);
    for (int mb = 0; mb < 4; mb++)
        for (int kb = 0; kb < 4; kb++)
            m.w2[mb][kb] = load4(W + NW1 + (16 * mb + r) * RB_MLP_H + 16 * kb + 4 * q);
    for (int kb = 0; kb < 4; kb++)
        m.w3[kb] = r < RB_MLP_OUT ? load4(W + NW1 + NW2 + r * RB_MLP_H + 16 * kb + 4 * q)
                                  : pack4(0.f, 0.f, 0.f, 0.f);
    for (int mb = 0; mb < 4; mb++)
        for (int j = 0; j < 4; j++) {
            m.b1[mb][j] = B[16 * mb + 4 * q + j];
            m.b2[mb][j] = B[RB_MLP_H + 16 * mb + 4 * q + j];
        }
    for (int j = 0; j < 4; j++) m.b3[j] = 4 * q + j < RB_MLP_OUT ? B[2 * RB_MLP_H + 4 * q + j] : 0.f;
}

// Forward pass of 16 contacts.  x: this lane's features 4*(l/16)..+3 of
// contact l%16.  a1/a2: post-ReLU activations, lane l holding hidden units
// 16*mb + 4*(l/16) + j of contact l%16.  y: lanes 0..15 hold outputs j = 0..2
// of contact l.
__device__ __forceinline__ void mlp_tile(const MlpRegs &m, float4 x, f32x4 a1[4], f32x4 a2[4],
                                         f32x4 &y) {
    const bf16x4 xb = pack4(x.x, x.y, x.z, x.w);
    bf16x4 h[4];
    for (int mb = 0; mb < 4; mb++) {
        f32x4 acc = mfma_bf16(m.w1[mb], xb, zero4());
        for (int j = 0; j < 4; j++) {
            float v = acc[j] + m.b1[mb][j];
            a1[mb][j] = v > 0.f ? v : 0.f;
        }
        h[mb] = pack4(a1[mb][0], a1[mb][1], a1[mb][2], a1[mb][3]);
    }
    for (int mb = 0; mb < 4; mb++) {
        f32x4 acc = zero4();
        for (int kb = 0; kb < 4; kb++) acc = mfma_bf16(m.w2[mb][kb], h[kb], acc);
        for (int j = 0; j < 4; j++) {
            float v = acc[j] + m.b2[mb][j];
            a2[mb][j] = v > 0.f ? v : 0.f;
        }
    }
    bf16x4 h2[4];
    for (int kb = 0; kb < 4; kb++) h2[kb] = pack4(a2[kb][0], a2[kb][1], a2[kb][2], a2[kb][3]);
    f32x4 acc = zero4();
    for (int kb = 0; kb < 4; kb++) acc = mfma_bf16(m.w3[kb], h2[kb], acc);
    for (int j = 0; j < 4; j++) y[j] = acc[j] + m.b3[j];
}

// Impulses from a prediction (network units are RB_MLP_SCALE x impulse):
// the predicted normal impulse, clamped at zero, is shared over the points in
// proportion to the impulses the constraint setup matched from the previous
// step (evenly over the points with a row when there are none).  Friction:
// a manifold with matched previous impulses keeps the previous step's
// friction impulses the setup gave it, others take the predicted ones; either
// is clamped to the cone of the predicted normal impulse.
__device__ __forceinline__ void apply_prediction(int k, float y0, float y1, float y2,
                                                 float friction, ConstraintSoA cs) {
    const float4 jac = cs.jac_inv[k], lp = cs.lambda[k];
    const float cnt = (float)((jac.x != 0.f) + (jac.y != 0.f) + (jac.z != 0.f) + (jac.w != 0.f));
    const float sum = ((lp.x + lp.y) + lp.z) + lp.w;
    float ln = y0 / RB_MLP_SCALE;
    ln = ln > 0.f ? ln : 0.f;
    const float jv[4] = {jac.x, jac.y, jac.z, jac.w}, pv[4] = {lp.x, lp.y, lp.z, lp.w};
    float lv[4];
    for (int i = 0; i < 4; i++) {
        if (jv[i] == 0.f) lv[i] = 0.f;
        else if (sum > 0.f) lv[i] = ln * (pv[i] / sum);
        else lv[i] = cnt > 0.f ? ln / cnt : 0.f;
    }
    cs.lambda[k] = make_float4(lv[0], lv[1], lv[2], lv[3]);
    float4 fr = cs.fric[k];
    float lim = friction * ln;
    float f0 = sum > 0.f ? fr.z : y1 / RB_MLP_SCALE, f1 = sum > 0.f ? fr.w : y2 / RB_MLP_SCALE;
    f0 = fminf(fmaxf(f0, -lim), lim);
    f1 = fminf(fmaxf(f1, -lim), lim);
    if (fr.x == 0.f) { f0 = 0.f; f1 = 0.f; }
    fr.z = f0;
    fr.w = f1;
    cs.fric[k] = fr;
}

// Forward over n contacts (n read from *n_dev when given).  APPLY: also set
// the constraints' initial impulses from the prediction.
template <bool APPLY>
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ x, int n,
                                                     const int *n_dev,
                                                     const uint16_t *__restrict__ W,
                                                     const float *__restrict__ B,
                                                     float *__restrict__ y, float friction,
                                                     ConstraintSoA cs) {
    if (n_dev) n = min(n, *n_dev);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    MlpRegs m;
    mlp_load(m, W, B, lane);
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int t = blockIdx.x * (blockDim.x >> 6) + wave; 16 * t < n; t += waves) {
        const int c = 16 * t + r;
        float4 xv = c < n ? reinterpret_cast<const float4 *>(x)[4 * (size_t)c + q]
                          : make_float4(0.f, 0.f, 0.f, 0.f);
        f32x4 a1[4], a2[4], yv;
        mlp_tile(m, xv, a1, a2, yv);
        if (q == 0 && c < n) {
            y[3 * (size_t)c] = yv[0];
            y[3 * (size_t)c + 1] = yv[1];
            y[3 * (size_t)c + 2] = yv[2];
            if (APPLY) apply_prediction(c, yv[0], yv[1], yv[2], friction, cs);
        }
    }
}

__device__ __forceinline__ uint32_t mix32(uint32_t k) {
    k ^= k >> 16; k *= 0x7feb352dU; k ^= k >> 15; k *= 0x846ca68bU; k ^= k >> 16;
    return k;
}

struct GradLDS {
    uint16_t x[MT][RB_MLP_IN];   // bf16 inputs
    uint16_t a1[MT][RB_MLP_H];   // bf16 activations (the values the forward multiplies)
    uint16_t a2[MT][RB_MLP_H];
    float d1[MT][RB_MLP_H];      // deltas at the hidden pre-activations
    float d2[MT][RB_MLP_H];
    float dy[MT][4];
    float w2[RB_MLP_H][RB_MLP_H];  // W2 as fp32 (exact bf16 values)
    float w3[RB_MLP_OUT][RB_MLP_H];
};

// Gradients of loss = mean_s sum_o (y_o - t_o)^2 / 3 over a batch of n rows.
// Row i is sample i (sample_seed < 0) or a hashed draw from n_samples.
// Writes PSTRIDE partial sums per workgroup (weights, biases, sum of e^2).
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ x,
                                                  const float *__restrict__ tgt, int n,
                                                  int n_samples, int sample_seed,
                                                  const uint16_t *__restrict__ W,
                                                  const float *__restrict__ B,
                                                  float *__restrict__ part) {
    __shared__ GradLDS L;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, tid = threadIdx.x;
    const int r = lane & 15, q = lane >> 4;
    MlpRegs m;
    mlp_load(m, W, B, lane);
    for (int i = tid; i < NW2; i += blockDim.x) L.w2[i / RB_MLP_H][i % RB_MLP_H] = bf2f(W[NW1 + i]);
    for (int i = tid; i < NW3; i += blockDim.x) L.w3[i / RB_MLP_H][i % RB_MLP_H] = bf2f(W[NW1 + NW2 + i]);
    f32x4 gw2[4], gw1 = zero4(), gw3 = zero4();
    for (int nb = 0; nb < 4; nb++) gw2[nb] = zero4();
    float gb = 0.f, lsum = 0.f;
    const float dscale = 3.f * (float)n;
    __syncthreads();
    const int ntiles = (n + MT - 1) / MT;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        // ---- forward of the wave's 16 rows, deltas in registers ----
        const int row = tile * MT + 16 * wave + r;  // this lane's row
        const int cl = 16 * wave + r;                // row within the tile
        const bool valid = row < n;
        int s = row;
        if (sample_seed >= 0)
            s = (int)(mix32((uint32_t)row * 0x9e3779b9u ^ mix32((uint32_t)sample_seed)) %
                      (uint32_t)n_samples);
        float4 xv = valid ? reinterpret_cast<const float4 *>(x)[4 * (size_t)s + q]
                          : make_float4(0.f, 0.f, 0.f, 0.f);
        f32x4 a1[4], a2[4], yv;
        mlp_tile(m, xv, a1, a2, yv);
        L.x[cl][4 * q] = f2bf(xv.x); L.x[cl][4 * q + 1] = f2bf(xv.y);
        L.x[cl][4 * q + 2] = f2bf(xv.z); L.x[cl][4 * q + 3] = f2bf(xv.w);
        for (int mb = 0; mb < 4; mb++)
            for (int j = 0; j < 4; j++) {
                L.a1[cl][16 * mb + 4 * q + j] = f2bf(a1[mb][j]);
                L.a2[cl][16 * mb + 4 * q + j] = f2bf(a2[mb][j]);
            }
        float dy[3] = {0.f, 0.f, 0.f};
        if (q == 0 && valid) {
            for (int o = 0; o < 3; o++) {
                float e = yv[o] - tgt[3 * (size_t)s + o];
                lsum += e * e;
                dy[o] = 2.f * e / dscale;
            }
        }
        if (q == 0) {
            L.dy[cl][0] = dy[0]; L.dy[cl][1] = dy[1]; L.dy[cl][2] = dy[2]; L.dy[cl][3] = 0.f;
        }
        // every lane of contact r needs its dy: broadcast from lane r
        for (int o = 0; o < 3; o++) dy[o] = __shfl(dy[o], r, 64);
        for (int mb = 0; mb < 4; mb++)
            for (int j = 0; j < 4; j++) {
                int k = 16 * mb + 4 * q + j;
                float acc = (dy[0] * L.w3[0][k] + dy[1] * L.w3[1][k]) + dy[2] * L.w3[2][k];
                L.d2[cl][k] = a2[mb][j] > 0.f ? acc : 0.f;
            }
        __syncthreads();
        // d1^T (64 x 16 rows) = W2^T . d2^T, masked by a1 > 0 (a1 in this lane's registers)
        for (int kb = 0; kb < 4; kb++) {
            f32x4 acc = zero4();
            for (int s4 = 0; s4 < 16; s4++)
                acc = mfma_f32(L.w2[4 * s4 + q][16 * kb + r], L.d2[16 * wave + r][4 * s4 + q], acc);
            for (int j = 0; j < 4; j++)
                L.d1[cl][16 * kb + 4 * q + j] = a1[kb][j] > 0.f ? acc[j] : 0.f;
        }
        __syncthreads();
        // ---- weight gradients: reductions over the tile's 64 rows ----
        for (int s4 = 0; s4 < 16; s4++) {
            const int c = 4 * s4 + q;
            const float ad2 = L.d2[c][16 * wave + r];
            for (int nb = 0; nb < 4; nb++) gw2[nb] = mfma_f32(ad2, bf2f(L.a1[c][16 * nb + r]), gw2[nb]);
            gw1 = mfma_f32(L.d1[c][16 * wave + r], bf2f(L.x[c][r]), gw1);
            gw3 = mfma_f32(r < 3 ? L.dy[c][r] : 0.f, bf2f(L.a2[c][16 * wave + r]), gw3);
        }
        if (tid < RB_MLP_H) {
            for (int c = 0; c < MT; c++) gb += L.d1[c][tid];
        } else if (tid < 2 * RB_MLP_H) {
            for (int c = 0; c < MT; c++) gb += L.d2[c][tid - RB_MLP_H];
        } else if (tid < 2 * RB_MLP_H + RB_MLP_OUT) {
            for (int c = 0; c < MT; c++) gb += L.dy[c][tid - 2 * RB_MLP_H];
        }
        __syncthreads();
    }
    float *P = part + (size_t)blockIdx.x * PSTRIDE;
    for (int j = 0; j < 4; j++) {
        const int o = 16 * wave + 4 * q + j;
        for (int nb = 0; nb < 4; nb++) P[NW1 + o * RB_MLP_H + 16 * nb + r] = gw2[nb][j];
        P[o * RB_MLP_IN + r] = gw1[j];
        if (4 * q + j < RB_MLP_OUT) P[NW1 + NW2 + (4 * q + j) * RB_MLP_H + 16 * wave + r] = gw3[j];
    }
    if (tid < NB) P[NW + tid] = gb;
    // loss: sum of e^2 over the workgroup
    __shared__ float red[4];
    float v = lsum;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (tid == 0) P[NW + NB] = ((red[0] + red[1]) + red[2]) + red[3];
}

struct AdamArgs {
    float lr_t;      // lr * sqrt(1 - beta2^t) / (1 - beta1^t)
    float beta1, beta2, eps;
    int step;        // index into the loss history
};

// Sum the workgroup partials (fixed order) and, with master != null, take one
// Adam step.  grad/loss outputs are optional.
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float *__restrict__ part, int nblk,
                                                         int n, float *grad, float *loss,
                                                         float *master, float *mom, float *vel,
                                                         uint16_t *W, float *B, AdamArgs a) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= PSTRIDE) return;
    float g = 0.f;
    for (int b = 0; b < nblk; b++) g += part[(size_t)b * PSTRIDE + p];
    if (p == NW + NB) {
        if (loss) loss[a.step] = g / (3.f * (float)n);
        return;
    }
    if (grad) grad[p] = g;
    if (!master) return;
    float mm = a.beta1 * mom[p] + (1.f - a.beta1) * g;
    float vv = a.beta2 * vel[p] + (1.f - a.beta2) * g * g;
    mom[p] = mm;
    vel[p] = vv;
    float th = master[p] - a.lr_t * mm / (sqrtf(vv) + a.eps);
    master[p] = th;
    if (p < NW) W[p] = f2bf(th);
    else B[p - NW] = th;
}

// Training pairs of the last step: features and the solved impulses in
// network units (sum of normal lambdas, friction 0, friction 1).
__global__ __launch_bounds__(256) void k_mlp_targets(const int *counters, ConstraintSoA cs,
                                                     const float *__restrict__ features,
                                                     int capacity, float *__restrict__ x_out,
                                                     float *__restrict__ t_out) {
    int n = counters[RB_CNT_CONTACTS];
    if (n > capacity) n = capacity;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 *fi = reinterpret_cast<const float4 *>(features) + 4 * (size_t)k;
    float4 *fo = reinterpret_cast<float4 *>(x_out) + 4 * (size_t)k;
    fo[0] = fi[0]; fo[1] = fi[1]; fo[2] = fi[2]; fo[3] = fi[3];
    float4 l = cs.lambda[k], f = cs.fric[k];
    t_out[3 * (size_t)k] = (((l.x + l.y) + l.z) + l.w) * RB_MLP_SCALE;
    t_out[3 * (size_t)k + 1] = f.z * RB_MLP_SCALE;
    t_out[3 * (size_t)k + 2] = f.w * RB_MLP_SCALE;
}

int forward_grid(int n) {
    int waves = (n + 15) / 16;
    int blocks = (waves + 3) / 4;
    int cap = rb_num_cus() * 4;
    return blocks < 1 ? 1 : (blocks > cap ? cap : blocks);
}

}  // namespace

hipError_t launch_mlp_warm(const DeviceWorld &d, const StepParams &p, hipStream_t s) {
    RB_LAUNCH(k_mlp_forward<true>, dim3(forward_grid(p.max_contacts)), dim3(256), 0, s,
                       d.features, p.max_contacts, d.counters + RB_CNT_CONTACTS, d.mlp_w, d.mlp_b,
                       d.mlp_out, p.friction, d.cs);
    return rb_launch_status();
}

hipError_t launch_mlp_targets(const DeviceWorld &d, int capacity, float *x_out, float *t_out,
                              hipStream_t s) {
    RB_LAUNCH(k_mlp_targets, dim3((capacity + 255) / 256), dim3(256), 0, s, d.counters,
                       d.cs, d.features, capacity, x_out, t_out);
    return rb_launch_status();
}

// ---- C ABI: a network handle (weights, Adam state, scratch) ----
struct MlpHandle {
    int device = 0;
    hipStream_t stream = nullptr;
    uint16_t *w = nullptr;   // bf16 working weights
    float *b = nullptr;      // fp32 biases
    float *master = nullptr; // fp32 master copy of weights and biases (NW + NB)
    float *mom = nullptr, *vel = nullptr;
    float *part = nullptr;   // GRAD_BLOCKS x PSTRIDE
    float *loss = nullptr;   // per-step loss history
    int loss_cap = 0;
    int adam_t = 0;
};

static mi_rb_status mlp_upload(MlpHandle *h, const uint16_t *w_bf16, const float *bias) {
    std::vector<float> master(NW + NB);
    for (int i = 0; i < NW; i++) {
        uint32_t u = (uint32_t)w_bf16[i] << 16;
        std::memcpy(&master[i], &u, 4);
    }
    for (int i = 0; i < NB; i++) master[NW + i] = bias[i];
    RB_CHECK(hipMemcpyAsync(h->w, w_bf16, 2 * (size_t)NW, hipMemcpyHostToDevice, h->stream));
    RB_CHECK(hipMemcpyAsync(h->b, bias, 4 * (size_t)NB, hipMemcpyHostToDevice, h->stream));
    RB_CHECK(hipMemcpyAsync(h->master, master.data(), 4 * (size_t)(NW + NB), hipMemcpyHostToDevice,
                            h->stream));
    RB_CHECK(hipMemsetAsync(h->mom, 0, 4 * (size_t)(NW + NB), h->stream));
    RB_CHECK(hipMemsetAsync(h->vel, 0, 4 * (size_t)(NW + NB), h->stream));
    RB_CHECK(hipStreamSynchronize(h->stream));
    h->adam_t = 0;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_mlp_create(uint64_t seed, void **out) {
    if (!out) return MI_RB_INVALID_ARG;
    *out = nullptr;
    MlpHandle *h = new MlpHandle();
    hipGetDevice(&h->device);
    mi_rb_status st = MI_RB_OK;
    auto fail = [&](hipError_t e) {
        if (e != hipSuccess && st == MI_RB_OK) st = rb_device_error(e, "mlp alloc", __FILE__, __LINE__);
    };
    fail(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    fail(hipMalloc(&h->w, 2 * (size_t)NW));
    fail(hipMalloc(&h->b, 4 * (size_t)NB));
    fail(hipMalloc(&h->master, 4 * (size_t)(NW + NB)));
    fail(hipMalloc(&h->mom, 4 * (size_t)(NW + NB)));
    fail(hipMalloc(&h->vel, 4 * (size_t)(NW + NB)));
    fail(hipMalloc(&h->part, 4 * (size_t)GRAD_BLOCKS * PSTRIDE));
    if (st != MI_RB_OK) {
        mi_rb_mlp_destroy(h);
        return st;
    }
    // He-uniform initial weights from a seeded 64-bit LCG, zero biases
    std::vector<uint16_t> w(NW);
    std::vector<float> bias(NB, 0.f);
    uint64_t x = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((x >> 40) * (1.0 / 16777216.0)) * 2.f - 1.f;
    };
    auto host_bf = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    };
    for (int i = 0; i < NW; i++) {
        int fan_in = i < NW1 ? RB_MLP_IN : RB_MLP_H;
        float lim = std::sqrt(6.f / (float)fan_in);
        if (i >= NW1 + NW2) lim = std::sqrt(3.f / (float)fan_in);
        w[i] = host_bf(uni() * lim);
    }
    st = mlp_upload(h, w.data(), bias.data());
    if (st != MI_RB_OK) {
        mi_rb_mlp_destroy(h);
        return st;
    }
    *out = h;
    return MI_RB_OK;
}

extern "C" void mi_rb_mlp_destroy(void *handle) {
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->w); hipFree(h->b); hipFree(h->master); hipFree(h->mom); hipFree(h->vel);
    hipFree(h->part); hipFree(h->loss);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

extern "C" mi_rb_status mi_rb_mlp_set_weights(void *handle, const uint16_t *w_bf16,
                                              const float *bias) {
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    if (!h || !w_bf16 || !bias) return MI_RB_INVALID_ARG;
    return mlp_upload(h, w_bf16, bias);
}

extern "C" mi_rb_status mi_rb_mlp_get_weights(void *handle, uint16_t *w_bf16, float *bias) {
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    if (!h || !w_bf16 || !bias) return MI_RB_INVALID_ARG;
    RB_CHECK(hipMemcpyAsync(w_bf16, h->w, 2 * (size_t)NW, hipMemcpyDeviceToHost, h->stream));
    RB_CHECK(hipMemcpyAsync(bias, h->b, 4 * (size_t)NB, hipMemcpyDeviceToHost, h->stream));
    RB_CHECK(hipStreamSynchronize(h->stream));
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_mlp_forward(void *handle, const float *x_dev, int n, float *y_dev) {
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    if (!h || n < 0 || (n > 0 && (!x_dev || !y_dev))) return MI_RB_INVALID_ARG;
    if (n == 0) return MI_RB_OK;
    ConstraintSoA none{};
    RB_LAUNCH(k_mlp_forward<false>, dim3(forward_grid(n)), dim3(256), 0, h->stream, x_dev,
                       n, (const int *)nullptr, h->w, h->b, y_dev, 0.f, none);
    RB_CHECK(rb_launch_status());
    RB_CHECK(hipStreamSynchronize(h->stream));
    return MI_RB_OK;
}

static int grad_blocks(int n) {
    int tiles = (n + MT - 1) / MT;
    return tiles < GRAD_BLOCKS ? (tiles < 1 ? 1 : tiles) : GRAD_BLOCKS;
}

extern "C" mi_rb_status mi_rb_mlp_gradients(void *handle, const float *x_dev, const float *t_dev,
                                            int n, float *grad_dev, float *loss) {
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    if (!h || n <= 0 || !x_dev || !t_dev || !grad_dev) return MI_RB_INVALID_ARG;
    if (h->loss_cap < 1) {
        hipFree(h->loss);
        h->loss = nullptr;
        RB_CHECK(hipMalloc(&h->loss, sizeof(float)));
        h->loss_cap = 1;
    }
    int nb = grad_blocks(n);
    RB_LAUNCH(k_mlp_grad, dim3(nb), dim3(256), 0, h->stream, x_dev, t_dev, n, n, -1, h->w,
                       h->b, h->part);
    AdamArgs a{0.f, 0.f, 0.f, 0.f, 0};
    RB_LAUNCH(k_mlp_reduce_adam, dim3((PSTRIDE + 255) / 256), dim3(256), 0, h->stream,
                       h->part, nb, n, grad_dev, h->loss, (float *)nullptr, (float *)nullptr,
                       (float *)nullptr, (uint16_t *)nullptr, (float *)nullptr, a);
    RB_CHECK(rb_launch_status());
    float l = 0.f;
    RB_CHECK(hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    RB_CHECK(hipStreamSynchronize(h->stream));
    if (loss) *loss = l;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_mlp_train(void *handle, const float *x_dev, const float *t_dev,
                                        int n_samples, int batch, int steps, float lr, int seed,
                                        float *loss_history) {
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    if (!h || n_samples <= 0 || batch <= 0 || steps < 0 || !x_dev || !t_dev || !(lr > 0.f))
        return MI_RB_INVALID_ARG;
    // seed < 0: every step takes rows 0..batch-1 in order (full-batch steps)
    if (seed < 0 && batch > n_samples) return MI_RB_INVALID_ARG;
    if (steps == 0) return MI_RB_OK;
    if (h->loss_cap < steps) {
        hipFree(h->loss);
        h->loss = nullptr;
        RB_CHECK(hipMalloc(&h->loss, sizeof(float) * (size_t)steps));
        h->loss_cap = steps;
    }
    const float b1 = 0.9f, b2 = 0.999f;
    int nb = grad_blocks(batch);
    for (int i = 0; i < steps; i++) {
        h->adam_t++;
        double c1 = 1.0 - std::pow((double)b1, h->adam_t), c2 = 1.0 - std::pow((double)b2, h->adam_t);
        AdamArgs a{(float)(lr * std::sqrt(c2) / c1), b1, b2, 1e-8f, i};
        int draw_seed =
            seed < 0 ? -1 : (int)(((uint32_t)seed * 0x9e3779b9u + (uint32_t)h->adam_t) & 0x7fffffffu);
        RB_LAUNCH(k_mlp_grad, dim3(nb), dim3(256), 0, h->stream, x_dev, t_dev, batch,
                           n_samples, draw_seed, h->w, h->b, h->part);
        RB_LAUNCH(k_mlp_reduce_adam, dim3((PSTRIDE + 255) / 256), dim3(256), 0, h->stream,
                           h->part, nb, batch, (float *)nullptr, h->loss, h->master, h->mom, h->vel,
                           h->w, h->b, a);
    }
    RB_CHECK(rb_launch_status());
    if (loss_history)
        RB_CHECK(hipMemcpyAsync(loss_history, h->loss, sizeof(float) * (size_t)steps,
                                hipMemcpyDeviceToHost, h->stream));
    RB_CHECK(hipStreamSynchronize(h->stream));
    return MI_RB_OK;
}
