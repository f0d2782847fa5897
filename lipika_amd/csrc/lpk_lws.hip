// lpk_lws.hip -- kernels of the learned warm start (network in lpk_lws.hpp):
//   lws_gen_kernel     training pairs: seeded geometry -> features and the
//                      converged f64 light time on the analytic circular orbit
//   lws_infer_kernel   batched forward pass, 32 queries per wave on the matrix cores
//   lws_train_kernel   one minibatch slice of 64 queries per workgroup: the
//                      matrix-core forward, then the backward pass in FP32 on
//                      the vector ALU, writing the slice's gradient and loss sum
//   lws_update_kernel  fixed-order sum of the slices' rows, then Adam on the
//                      FP32 master parameters and the BF16 operand copies
// Nothing here depends on the scene: the network learns the geometry of the
// light-time equation itself.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lpk_device.hpp"
#include "lpk_lws.hpp"

namespace lpk {

constexpr uint64_t kLwsDataTag = 0x6C77736461746131ULL;  // "lwsdata1"
constexpr double kPi = 3.141592653589793;
// WGS-84 (src/geometry/ecef.rs)
constexpr double kWgs84A = 6378137.0;
constexpr double kWgs84F = 1.0 / 298.257223563;
constexpr double kWgs84E2 = kWgs84F * (2.0 - kWgs84F);

__device__ __forceinline__ double unit_draw(uint64_t w) { return static_cast<double>(w >> 11) * 0x1.0p-53; }

// Earth-fixed position on a circular orbit: Rz(node) Rx(i) (a cos u, a sin u, 0),
// the operations of src/geometry/kepler.rs:41-48 with the Earth-fixed node.
__device__ __forceinline__ void orbit_position(double a, double ci, double si, double node, double u, double *p) {
    double su, cu, sn, cn;
    sincos(u, &su, &cu);
    sincos(node, &sn, &cn);
    const double xp = a * cu, yp = a * su;
    const double y1 = yp * ci, z1 = yp * si;
    p[0] = xp * cn - y1 * sn;
    p[1] = xp * sn + y1 * cn;
    p[2] = z1;
}

// Light time at query time tq for the orbit in state (node, u) at tq: the
// straight-line value |P(tq) - O| / c, and the root of
// tau = |R(w tau) P(tq - tau) - O| / c (the emission equation of
// src/channel/mod.rs:70-99 with the same linearised Sagnac rotation), iterated
// from the straight-line value until the update is below 1e-16 s.
__device__ void lws_light_time(const double *obs, double a, double ci, double si, double node, double u, double n,
                               double &tau_sl, double &tau) {
    double p[3];
    orbit_position(a, ci, si, node, u, p);
    const double dx = p[0] - obs[0], dy = p[1] - obs[1], dz = p[2] - obs[2];
    tau_sl = sqrt(dx * dx + dy * dy + dz * dz) * kInvC;
    tau = tau_sl;
    for (int it = 0; it < 12; ++it) {
        orbit_position(a, ci, si, node + kEarthRotation * tau, u - n * tau, p);
        const double w = kEarthRotation * tau;
        const double sx = p[0] + w * p[1], sy = p[1] - w * p[0];
        const double ex = sx - obs[0], ey = sy - obs[1], ez = p[2] - obs[2];
        const double tn = sqrt(ex * ex + ey * ey + ez * ez) * kInvC;
        const double d = tn - tau;
        tau = tn;
        if (fabs(d) < 1e-16) break;
    }
}

// Pair p from Philox blocks 2p and 2p + 1 of (seed, kLwsDataTag):
//   receiver  lat U[-90, 90] deg, lon U[-180, 180) deg, h U[0, 2000] m (WGS-84)
//   orbit     LEO a U[6.9e6, 7.5e6] m or MEO a U[2.6e7, 2.7e7] m (even odds),
//             i U[0, 180) deg, Earth-fixed node and argument of latitude U[0, 360) deg
// X[p] = features, Y[p] = (tau - tau_sl) / 1e-6 s; G[p] (optional) = obs xyz,
// a, i, node, u, tau.
__global__ void __launch_bounds__(256) lws_gen_kernel(uint64_t n, uint64_t seed, float *__restrict__ X,
                                                      float *__restrict__ Y, double *__restrict__ G) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint64_t w0[4] = {2 * p, 0, 0, 0}, w1[4] = {2 * p + 1, 0, 0, 0};
    philox4x64(w0, seed, kLwsDataTag);
    philox4x64(w1, seed, kLwsDataTag);
    const double lat = (unit_draw(w0[0]) - 0.5) * kPi;
    const double lon = (2.0 * unit_draw(w0[1]) - 1.0) * kPi;
    const double hgt = 2000.0 * unit_draw(w0[2]);
    const bool leo = unit_draw(w0[3]) < 0.5;
    const double a = leo ? 6.9e6 + 6.0e5 * unit_draw(w1[0]) : 2.6e7 + 1.0e6 * unit_draw(w1[0]);
    const double inc = kPi * unit_draw(w1[1]);
    const double node = 2.0 * kPi * unit_draw(w1[2]);
    const double u = 2.0 * kPi * unit_draw(w1[3]);
    double slat, clat, slon, clon;
    sincos(lat, &slat, &clat);
    sincos(lon, &slon, &clon);
    const double nn = kWgs84A / sqrt(1.0 - kWgs84E2 * slat * slat);
    const double obs[3] = {(nn + hgt) * clat * clon, (nn + hgt) * clat * slon, (nn * (1.0 - kWgs84E2) + hgt) * slat};
    double si, ci;
    sincos(inc, &si, &ci);
    const double mm = sqrt(kGmEarth / (a * a * a));
    double tau_sl, tau;
    lws_light_time(obs, a, ci, si, node, u, mm, tau_sl, tau);
    float x[kLwsIn];
    lws_features(obs, a, ci, si, node, u, x);
#pragma unroll
    for (int k = 0; k < kLwsIn; ++k) X[p * kLwsIn + k] = x[k];
    Y[p] = static_cast<float>((tau - tau_sl) / kLwsTargetScale);
    if (G) {
        double *g = G + p * 8;
        g[0] = obs[0];
        g[1] = obs[1];
        g[2] = obs[2];
        g[3] = a;
        g[4] = inc;
        g[5] = node;
        g[6] = u;
        g[7] = tau;
    }
}

// Y[q] = net(X[q]) for q < n; 128 queries per workgroup (4 waves of 32).
__global__ void __launch_bounds__(256) lws_infer_kernel(LwsNetDev net, const float *__restrict__ X, uint64_t n,
                                                        float *__restrict__ Y) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const uint64_t q = (uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 32 + (lane & 31);
    const bool valid = q < n;
    float x[kLwsIn];
#pragma unroll
    for (int k = 0; k < kLwsIn; ++k) x[k] = valid ? X[q * kLwsIn + k] : 0.0f;
    const float y = lws_forward(net, x, nullptr, 0);
    if (valid && lane < 32) Y[q] = y;
}

// Gradient of the slice [q0, q0 + 64) of a minibatch, q0 = base + 64 blockIdx.x
// (queries at or past `end` count zero), for the loss scale * sum (y - t)^2 / 2
// (scale = 2 / batch gives the batch mean squared error). The backward pass
// treats the BF16 roundings of the forward as the identity (straight-through):
//   dz3 = dy w4 gelu'(z3), dW3 = dz3 (x) bf16(h2), dh2 = bf16(W3)^T dz3,
//   dz2 = dh2 gelu'(z2), ... down to dW1 = dz1 (x) (x_hi + x_lo).
// Row blockIdx.x of `partials` receives the gradient in parameter order and,
// at kLwsParams, the slice's sum of squared errors.
__global__ void __launch_bounds__(256) lws_train_kernel(LwsNetDev net, const float *__restrict__ X,
                                                        const float *__restrict__ Y, uint64_t base, uint64_t end,
                                                        float scale, float *__restrict__ partials) {
    __shared__ float s_x[kTrainQ][kLwsIn];          // the inputs as the matrix cores saw them
    __shared__ float s_z[3 * kTrainQ * 64];         // pre-activations, then their gradients
    __shared__ float s_h[kTrainQ][64];              // BF16 activations of the layer below
    __shared__ float s_dh[kTrainQ][64];             // gradient of those activations
    __shared__ float s_w[64][64];                   // BF16 weights of the layer above
    __shared__ float s_dy[kTrainQ], s_sq[kTrainQ];
    __shared__ float s_red[4][2][64];

    const int t = static_cast<int>(threadIdx.x);
    const int wave = t >> 6, lane = t & 63;
    const uint64_t q0 = base + uint64_t(blockIdx.x) * kTrainQ;
    float *row = partials + size_t(blockIdx.x) * kLwsRow;
    const float *th = net.theta;

    if (wave < 2) {
        const int qrow = wave * 32 + (lane & 31);
        const uint64_t q = q0 + qrow;
        const bool valid = q < end;
        float x[kLwsIn];
#pragma unroll
        for (int k = 0; k < kLwsIn; ++k) x[k] = valid ? X[q * kLwsIn + k] : 0.0f;
        const float y = lws_forward(net, x, s_z, qrow);
        if (lane < 32) {
#pragma unroll
            for (int k = 0; k < kLwsIn; ++k) {
                const float hi = bf16_round(x[k]);
                s_x[qrow][k] = hi + bf16_round(x[k] - hi);
            }
            const float d = valid ? y - Y[q] : 0.0f;
            s_dy[qrow] = d * scale;
            s_sq[qrow] = d * d;
        }
    }
    __syncthreads();

    float *z1 = s_z, *z2 = s_z + kTrainQ * 64, *z3 = s_z + 2 * kTrainQ * 64;
    // ---- output unit and layer 3's pre-activation gradient
    {
        const int u = t & 63, part = t >> 6;
        const float w4 = th[kOffW4 + u];
        float gw4 = 0.0f, gb3 = 0.0f;
        for (int j = 0; j < kTrainQ / 4; ++j) {
            const int q = part * (kTrainQ / 4) + j;
            const float z = z3[q * 64 + u];
            const float dy = s_dy[q];
            gw4 += dy * lws_gelu(z);
            const float dz = dy * w4 * lws_gelu_grad(z);
            gb3 += dz;
            z3[q * 64 + u] = dz;
        }
        s_red[part][0][u] = gw4;
        s_red[part][1][u] = gb3;
        for (int e = t; e < kTrainQ * 64; e += 256) (&s_h[0][0])[e] = bf16_round(lws_gelu(z2[e]));
        for (int e = t; e < 64 * 64; e += 256) (&s_w[0][0])[e] = bf16_round(th[kOffW3 + e]);
    }
    __syncthreads();
    if (t < 64) {
        row[kOffW4 + t] = ((s_red[0][0][t] + s_red[1][0][t]) + s_red[2][0][t]) + s_red[3][0][t];
        row[kOffB3 + t] = ((s_red[0][1][t] + s_red[1][1][t]) + s_red[2][1][t]) + s_red[3][1][t];
    } else if (t == 64) {
        float sq = 0.0f, dsum = 0.0f;
        for (int q = 0; q < kTrainQ; ++q) {
            sq += s_sq[q];
            dsum += s_dy[q];
        }
        row[kLwsParams] = sq;
        row[kOffB4] = dsum;
    }

    // ---- layers 3 and 2: weight gradient and the gradient of the layer below
    for (int L = 2; L >= 1; --L) {
        float *dzL = L == 2 ? z3 : z2;
        const int woff = L == 2 ? kOffW3 : kOffW2;
        {
            // dW[u][v] = sum_q dz[q][u] h[q][v]: a 4 x 4 block per lane
            const int ub = (t >> 4) * 4, vb = (t & 15) * 4;
            float acc[4][4] = {};
            for (int q = 0; q < kTrainQ; ++q) {
                float dz[4], hv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    dz[a] = dzL[q * 64 + ub + a];
                    hv[a] = s_h[q][vb + a];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] += dz[a] * hv[b];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) row[woff + (ub + a) * 64 + vb + b] = acc[a][b];
        }
        {
            // dh[q][v] = sum_u W[u][v] dz[q][u]: 16 consecutive v per lane
            const int q = t >> 2, vb = (t & 3) * 16;
            float acc[16] = {};
            for (int u = 0; u < 64; ++u) {
                const float dz = dzL[q * 64 + u];
#pragma unroll
                for (int b = 0; b < 16; ++b) acc[b] += s_w[u][vb + b] * dz;
            }
#pragma unroll
            for (int b = 0; b < 16; ++b) s_dh[q][vb + b] = acc[b];
        }
        __syncthreads();
        {
            // dz of the layer below (in place), its bias gradient, and the
            // operands of the next step
            float *zb = L == 2 ? z2 : z1;
            const int u = t & 63, part = t >> 6;
            float gb = 0.0f;
            for (int j = 0; j < kTrainQ / 4; ++j) {
                const int q = part * (kTrainQ / 4) + j;
                const float dz = s_dh[q][u] * lws_gelu_grad(zb[q * 64 + u]);
                zb[q * 64 + u] = dz;
                gb += dz;
            }
            s_red[part][0][u] = gb;
            if (L == 2) {
                for (int e = t; e < kTrainQ * 64; e += 256) (&s_h[0][0])[e] = bf16_round(lws_gelu(z1[e]));
                for (int e = t; e < 64 * 64; e += 256) (&s_w[0][0])[e] = bf16_round(th[kOffW2 + e]);
            }
        }
        __syncthreads();
        if (t < 64)
            row[(L == 2 ? kOffB2 : kOffB1) + t] =
                ((s_red[0][0][t] + s_red[1][0][t]) + s_red[2][0][t]) + s_red[3][0][t];
    }

    // ---- layer 1's weight gradient against the inputs
    {
        const int u = t >> 2, k0 = t & 3;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int q = 0; q < kTrainQ; ++q) {
            const float dz = z1[q * 64 + u];
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (k0 + 4 * j < kLwsIn) acc[j] += dz * s_x[q][k0 + 4 * j];
        }
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (k0 + 4 * j < kLwsIn) row[kOffW1 + u * kLwsIn + k0 + 4 * j] = acc[j];
    }
}

struct AdamArgs {
    float lr, beta1, beta2, eps;
    float c1, c2;   // 1 / (1 - beta1^t), 1 / (1 - beta2^t)
};

// 64 parameters per workgroup: lane group q (of 4) sums rows q, q + 4, ...
// of its parameter in row order, and the four partial sums are added in
// group order, so g is the same on every run. With grad_out set, g is
// written there and nothing is updated; otherwise Adam updates the master
// value, its moments and the parameter's BF16 operand copy. The loss column
// goes to *loss_out when that is set.
__global__ void __launch_bounds__(256) lws_update_kernel(const float *__restrict__ partials, int n_rows,
                                                         float *__restrict__ theta, float *__restrict__ m1,
                                                         float *__restrict__ m2, __bf16 *__restrict__ packed,
                                                         AdamArgs a, float *__restrict__ grad_out,
                                                         float *__restrict__ loss_out) {
    __shared__ float s_part[4][64];
    const int pl = static_cast<int>(threadIdx.x & 63), q = static_cast<int>(threadIdx.x >> 6);
    const int p = static_cast<int>(blockIdx.x) * 64 + pl;
    float acc = 0.0f;
    if (p < kLwsRow)
        for (int r = q; r < n_rows; r += 4) acc += partials[size_t(r) * kLwsRow + p];
    s_part[q][pl] = acc;
    __syncthreads();
    if (q != 0 || p >= kLwsRow) return;
    const float g = ((s_part[0][pl] + s_part[1][pl]) + s_part[2][pl]) + s_part[3][pl];
    if (p == kLwsParams) {
        if (loss_out) *loss_out = g;
        return;
    }
    if (grad_out) {
        grad_out[p] = g;
        return;
    }
    const float m = a.beta1 * m1[p] + (1.0f - a.beta1) * g;
    const float v = a.beta2 * m2[p] + (1.0f - a.beta2) * (g * g);
    m1[p] = m;
    m2[p] = v;
    const float th = theta[p] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    theta[p] = th;
    int layer = -1, r = 0, c = 0;
    if (p < kOffB1) {
        layer = 0;
        r = p / kLwsIn;
        c = p % kLwsIn;
    } else if (p >= kOffW2 && p < kOffB2) {
        layer = 1;
        r = (p - kOffW2) >> 6;
        c = (p - kOffW2) & 63;
    } else if (p >= kOffW3 && p < kOffB3) {
        layer = 2;
        r = (p - kOffW3) >> 6;
        c = (p - kOffW3) & 63;
    }
    if (layer >= 0) packed[lws_pack_index(layer, r, c)] = static_cast<__bf16>(th);
}

// BF16 operand copies of all three hidden weight matrices (after a parameter
// upload); the padding of layer 1's operand (inputs 10..15) is zeroed.
__global__ void __launch_bounds__(256) lws_pack_kernel(const float *__restrict__ theta, __bf16 *__restrict__ packed) {
    const int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    // layer 1's padding: lanes 32..63 (inputs 8..15), elements 2..7
    if (e < kPackW1 && ((e >> 3) & 63) >= 32 && (e & 7) >= 2) packed[e] = static_cast<__bf16>(0.0f);
    if (e < 64 * kLwsIn) packed[lws_pack_index(0, e / kLwsIn, e % kLwsIn)] = static_cast<__bf16>(theta[kOffW1 + e]);
    if (e < 64 * 64) {
        packed[lws_pack_index(1, e >> 6, e & 63)] = static_cast<__bf16>(theta[kOffW2 + e]);
        packed[lws_pack_index(2, e >> 6, e & 63)] = static_cast<__bf16>(theta[kOffW3 + e]);
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_lws_gen(uint64_t n, uint64_t seed, float *X, float *Y, double *G, hipStream_t s) {
    if (!n) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lws_gen_kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, n, seed, X, Y, G);
    return hipGetLastError();
}

hipError_t launch_lws_infer(const LwsNetDev &net, const float *X, uint64_t n, float *Y, hipStream_t s) {
    if (!n) return hipSuccess;
    const uint64_t blocks = (n + 127) / 128;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lws_infer_kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, net, X, n, Y);
    return hipGetLastError();
}

hipError_t launch_lws_train(const LwsNetDev &net, const float *X, const float *Y, uint64_t base, uint64_t end,
                            uint32_t n_slices, float scale, float *partials, hipStream_t s) {
    if (!n_slices) return hipSuccess;
    hipLaunchKernelGGL(lws_train_kernel, dim3(n_slices), dim3(256), 0, s, net, X, Y, base, end, scale, partials);
    return hipGetLastError();
}

hipError_t launch_lws_update(const float *partials, int n_rows, float *theta, float *m1, float *m2, void *packed,
                             float lr, float beta1, float beta2, float eps, float c1, float c2, float *grad_out,
                             float *loss_out, hipStream_t s) {
    const AdamArgs a{lr, beta1, beta2, eps, c1, c2};
    hipLaunchKernelGGL(lws_update_kernel, dim3((kLwsRow + 63) / 64), dim3(256), 0, s, partials, n_rows, theta,
                       m1, m2, static_cast<__bf16 *>(packed), a, grad_out, loss_out);
    return hipGetLastError();
}

hipError_t launch_lws_pack(const float *theta, void *packed, hipStream_t s) {
    hipLaunchKernelGGL(lws_pack_kernel, dim3(16), dim3(256), 0, s, theta, static_cast<__bf16 *>(packed));
    return hipGetLastError();
}

}  // namespace lpk
