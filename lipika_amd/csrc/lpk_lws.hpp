// lpk_lws.hpp -- the learned warm start's network: a 10 -> 64 -> 64 -> 64 -> 1
// multilayer perceptron with GELU hidden units, BF16 weights and FP32
// accumulation on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// It predicts the emission-time solve's light time tau* (the root of
// tau = |R(w tau) P(tu - tau) - O| / c) as a correction over the straight-line
// light time tau_sl = |P(tu) - O| / c:
//     tau_pred = tau_sl + 1e-6 s * net(features),
// from ten features of the geometry at the query time tu:
//     O / R_E (3),  (a - 1.7e7 m) / 1e7 m,  cos i, sin i,
//     cos, sin of the Earth-fixed node  raan - w_E (t - t_el),
//     cos, sin of the argument of latitude  nu0 + n (t - t_el),
// which place the circular orbit of src/geometry/kepler.rs:32-77 in the
// Earth-fixed frame: P = Rz(node) Rx(i) (a cos u, a sin u, 0).
//
// Parameters live in one f32 vector (the optimiser's master copy):
//     W1 [64][10], b1 [64], W2 [64][64], b2 [64], W3 [64][64], b3 [64], w4 [64], b4,
// with W[out][in]. The three hidden weight matrices are also kept as BF16 in
// the order the matrix-core operands are read (lws_pack_index).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lpk_device.hpp"

namespace lpk {

constexpr int kLwsIn = 10;
constexpr int kLwsHidden = 64;
constexpr int kOffW1 = 0, kOffB1 = 640, kOffW2 = 704, kOffB2 = 4800, kOffW3 = 4864, kOffB3 = 8960;
constexpr int kOffW4 = 9024, kOffB4 = 9088;
constexpr int kLwsParams = 9089;
// One training workgroup's partial row: the gradient, then the loss sum.
constexpr int kLwsRow = kLwsParams + 1;
// BF16 operand elements: W1 as 2 row blocks x 64 lanes x 8, W2 and W3 as
// 2 row blocks x 4 k-steps x 64 lanes x 8.
constexpr int kPackW1 = 2 * 64 * 8;
constexpr int kPackW23 = 2 * 4 * 64 * 8;
constexpr int kLwsPacked = kPackW1 + 2 * kPackW23;

constexpr double kLwsTargetScale = 1e-6;   // seconds per unit of network output
constexpr double kEarthRadius = 6378137.0; // feature scale of the receiver position
constexpr double kLwsAMid = 1.7e7, kLwsAScale = 1e7;
constexpr double kGmEarth = 3.986005e14;   // src/geometry/kepler.rs:12

// GELU, tanh form: 0.5 z (1 + tanh(k0 (z + k1 z^3))).
constexpr float kGeluK0 = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluK1 = 0.044715f;

// Training workgroup: 256 lanes, 64 queries (two waves of 32 run the forward).
constexpr int kTrainQ = 64;

// Packed index of W[r][c] of a hidden layer (layer 1: c < 10, K padded to 16).
//   mfma_f32_32x32x16_bf16 operand layout on gfx950, lane l and element i:
//     A: row l % 32, k = 8 (l / 32) + i;   B: k = 8 (l / 32) + i, column l % 32;
//     D: register i -> row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.
//   The forward computes H_out^T = W H_in^T with the 32 queries as columns, so
//   each lane's D registers (hidden units of its own query) are the next
//   layer's B operand without a transpose: k-step s, half h, element i takes
//   unit 32 (s / 2) + row(8 (s % 2) + i, h), with row(j, h) = 8 (j / 4) + 4 h + j % 4,
//   and the A operand is packed in that k order.
inline __host__ __device__ int lws_pack_index(int layer, int r, int c) {
    const int m = r >> 5;
    if (layer == 0) {
        const int h = c >> 3, i = c & 7;
        return (m * 64 + (r & 31) + 32 * h) * 8 + i;
    }
    const int t = c >> 5, cp = c & 31;
    const int h = (cp >> 2) & 1;
    const int ip = 4 * (cp >> 3) + (cp & 3);
    const int s = 2 * t + (ip >> 3), i = ip & 7;
    return kPackW1 + (layer - 1) * kPackW23 + ((m * 4 + s) * 64 + (r & 31) + 32 * h) * 8 + i;
}

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

struct LwsNetDev {
    const bf16x8 *packed;   // kLwsPacked BF16 elements, 16-byte aligned
    const float *theta;     // kLwsParams
};

__device__ __forceinline__ int lws_row(int j, int h) { return 8 * (j >> 2) + 4 * h + (j & 3); }

// tanh(u) = sign(u) (1 - e) / (1 + e) with e = 2^(-2 |u| log2 e), on the
// hardware exp2 and reciprocal (a few ulp; absolute error < 3e-7).
__device__ __forceinline__ float lws_tanh(float u) {
    const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * fabsf(u));
    return copysignf((1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e), u);
}

__device__ __forceinline__ float lws_gelu(float z) {
    const float u = kGeluK0 * (z + kGeluK1 * z * z * z);
    return 0.5f * z * (1.0f + lws_tanh(u));
}

__device__ __forceinline__ float lws_gelu_grad(float z) {
    const float u = kGeluK0 * (z + kGeluK1 * z * z * z);
    const float th = lws_tanh(u);
    return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * kGeluK0 * (1.0f + 3.0f * kGeluK1 * z * z);
}

__device__ __forceinline__ float bf16_round(float v) { return static_cast<float>(static_cast<__bf16>(v)); }

// Features of a query at receiver position obs (ECEF metres) for an orbit of
// semi-major axis a and inclination (cos_i, sin_i), Earth-fixed node `node`
// and argument of latitude `u` (radians) at the query time.
__device__ __forceinline__ void lws_features(const double *obs, double a, double cos_i, double sin_i, double node,
                                             double u, float *x) {
    x[0] = static_cast<float>(obs[0] / kEarthRadius);
    x[1] = static_cast<float>(obs[1] / kEarthRadius);
    x[2] = static_cast<float>(obs[2] / kEarthRadius);
    x[3] = static_cast<float>((a - kLwsAMid) / kLwsAScale);
    x[4] = static_cast<float>(cos_i);
    x[5] = static_cast<float>(sin_i);
    double sn, cn, su, cu;
    sincos(node, &sn, &cn);
    sincos(u, &su, &cu);
    x[6] = static_cast<float>(cn);
    x[7] = static_cast<float>(sn);
    x[8] = static_cast<float>(cu);
    x[9] = static_cast<float>(su);
}

// One wave evaluates the network for 32 queries. Lane l serves query l % 32 in
// half l / 32 of every operand; x holds that query's ten features on both
// halves. Every lane of the wave must call it (the matrix cores need the full
// wave); the output is returned on both lanes of a query. With zs set, the
// pre-activations of the three hidden layers are stored at
// zs[(layer * kTrainQ + qrow) * 64 + unit] for the backward pass.
__device__ __forceinline__ float lws_forward(const LwsNetDev &net, const float *x, float *zs, int qrow) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int h = lane >> 5;
    const float *th = net.theta;
    // layer 1: the inputs as a BF16 pair (hi + lo) over two k-steps with the
    // same weights, so the features keep about 16 significant bits
    bf16x8 bhi, blo;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float v = h ? (i < 2 ? x[8 + i] : 0.0f) : x[i];
        const __bf16 hi = static_cast<__bf16>(v);
        bhi[i] = hi;
        blo[i] = static_cast<__bf16>(v - static_cast<float>(hi));
    }
    float act[2][16];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const bf16x8 a = net.packed[m * 64 + lane];
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, blo, acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int u = 32 * m + lws_row(i, h);
            const float z = acc[i] + th[kOffB1 + u];
            if (zs) zs[qrow * 64 + u] = z;
            act[m][i] = lws_gelu(z);
        }
    }
    // layers 2 and 3
#pragma unroll
    for (int L = 0; L < 2; ++L) {
        bf16x8 b[4];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 8; ++i) b[s][i] = static_cast<__bf16>(act[s >> 1][8 * (s & 1) + i]);
        const bf16x8 *w = net.packed + (kPackW1 + L * kPackW23) / 8;
        const int boff = L ? kOffB3 : kOffB2;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            f32x16 acc = {};
#pragma unroll
            for (int s = 0; s < 4; ++s)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[(m * 4 + s) * 64 + lane], b[s], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int u = 32 * m + lws_row(i, h);
                const float z = acc[i] + th[boff + u];
                if (zs) zs[((L + 1) * kTrainQ + qrow) * 64 + u] = z;
                act[m][i] = lws_gelu(z);
            }
        }
    }
    // output unit on the vector ALU: each half sums its 32 units in FP32
    float p = 0.0f;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int i = 0; i < 16; ++i) p = p + th[kOffW4 + 32 * m + lws_row(i, h)] * act[m][i];
    p = p + __shfl_xor(p, 32);
    return p + th[kOffB4];
}

}  // namespace lpk
