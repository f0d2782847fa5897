// lpk_kernels.hip -- gfx950 kernels of the IQ-generation backend.
//
// render_kernel<FORM>: one workgroup renders one tile of kTile consecutive
// output samples of one (receiver, band, chunk) job; each lane owns one
// sample and walks the job's in-band transmitters in summation order:
//   light-time + Sagnac solve (src/channel/mod.rs:70-99)
//   -> s(t_emit) (src/signal/*.rs) -> FSPL amplitude and f64 carrier phase
//   -> f32 mix (src/channel/mod.rs:251-268) -> ordered f32 sum
//   (src/pipeline.rs:965-976) -> AWGN (src/channel/mod.rs:323-334)
//   -> post-sum ops (src/rx/mod.rs:83-101) -> one coalesced complex64 store.
// Every tile of every job of a render call is one launch.
//
// walk_*: the receiver clock's white-FM walk (src/clock/mod.rs:299-307) as a
// fixed-shape block scan over the master grid, so every value depends only on
// (seed, index), never on how the render was chunked.
//
// All code is compiled with -ffp-contract=off: the plain form keeps the
// reference's rounding sequence (Rust never fuses to FMA).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lpk_device.hpp"

namespace lpk {

// ---------------------------------------------------------------- helpers

// Rust `x as i64` after floor(): saturating, NaN -> 0.
__device__ __forceinline__ int64_t sat_i64(double x) {
    if (!(x == x)) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return static_cast<int64_t>(x);
}

// i64::rem_euclid for n > 0; a 32-bit path covers every index a real scene
// produces.
__device__ __forceinline__ int64_t rem_euclid(int64_t a, int64_t n) {
    if (a >= INT32_MIN && a <= INT32_MAX && n <= INT32_MAX) {
        int32_t r = static_cast<int32_t>(a) % static_cast<int32_t>(n);
        return r < 0 ? r + n : r;
    }
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ double rem_euclid_f64(double a, double b) {
    double r = fmod(a, b);
    return r < 0.0 ? r + fabs(b) : r;
}

// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        const uint64_t m0 = 0xD2E7470EE14C6C93ULL, m1 = 0xCA5A826395121157ULL;
        const uint64_t hi0 = __umul64hi(m0, c[0]), lo0 = m0 * c[0];
        const uint64_t hi1 = __umul64hi(m1, c[2]), lo1 = m1 * c[2];
        const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// Standard normal for master-grid index k of a clock walk (f64 Box-Muller).
__device__ __forceinline__ double clock_normal(uint64_t seed, uint64_t k) {
    uint64_t w[4] = {k, 0, 0, 0};
    philox4x64(w, seed, kClockTag);
    const double u1 = (static_cast<double>(w[0] >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = static_cast<double>(w[1] >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(kTau * u2);
}

// Two standard normals for global sample i of a band's noise (f32 Box-Muller).
__device__ __forceinline__ void awgn_normals(uint64_t seed, uint64_t i, float &zr, float &zi) {
    uint64_t w[4] = {i, 0, 0, 0};
    philox4x64(w, seed, kAwgnTag);
    const float u1 = (static_cast<float>(w[0] >> 40) + 0.5f) * 0x1.0p-24f;
    const float u2 = static_cast<float>(w[1] >> 40) * 0x1.0p-24f;
    const float r = sqrtf(-2.0f * logf(u1));
    const float th = 6.28318530717958647692f * u2;
    float s, c;
    sincosf(th, &s, &c);
    zr = r * c;
    zi = r * s;
}

// ---------------------------------------------------------------- geometry

// PrecomputedTrajectory::position_ecef (src/geometry/precomputed.rs:111-121):
// the clamp of floor(f) to [0, n-2] is done in f64, which gives exactly
// Rust's saturating cast + clamp (NaN -> 0 included).
__device__ __forceinline__ void tx_position(const TxDev &T, double t_abs, double &x, double &y,
                                            double &z) {
    if (T.traj_kind == 0) {
        x = T.pos[0];
        y = T.pos[1];
        z = T.pos[2];
        return;
    }
    const double f = (t_abs - T.grid_t0) / T.grid_dt;
    const double flo = fmin(fmax(floor(f), 0.0), static_cast<double>(T.grid_n - 2));
    const int64_t lo = static_cast<int64_t>(flo);
    double frac = f - flo;
    frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
    const double omf = 1.0 - frac;
    const double *g = T.grid;
    const int64_t n = T.grid_n;
    x = g[lo] * omf + g[lo + 1] * frac;
    y = g[n + lo] * omf + g[n + lo + 1] * frac;
    z = g[2 * n + lo] * omf + g[2 * n + lo + 1] * frac;
}

// solve_emission_times_for, one sample (src/channel/mod.rs:70-99).
__device__ __forceinline__ void solve_emission(const TxDev &T, const double *obs, double tu,
                                               double epoch, double &t_emit, double &r) {
    double t = tu - kEmissionInitialLag;
    r = 0.0;
#pragma unroll
    for (int it = 0; it < kNewtonIterations; ++it) {
        double sx, sy, sz;
        tx_position(T, epoch + t, sx, sy, sz);
        const double wtau = kEarthRotation * (tu - t);
        const double sx_rot = sx + wtau * sy;
        const double sy_rot = sy - wtau * sx;
        const double dx = sx_rot - obs[0];
        const double dy = sy_rot - obs[1];
        const double dz = sz - obs[2];
        r = sqrt(dx * dx + dy * dy + dz * dz);
        t = tu - r * kInvC;
    }
    t_emit = t;
}

// ---------------------------------------------------------------- signals

__device__ __forceinline__ int64_t frame_index(int64_t raw, int64_t n, int repeating, int64_t k) {
    // src/signal/mod.rs:108-115; -1 means silent
    if (!repeating) return (raw >= 0 && raw < n) ? raw : -1;
    if (n <= 0) return -1;
    return rem_euclid(static_cast<int64_t>(static_cast<uint64_t>(raw) + static_cast<uint64_t>(k)), n);
}

__device__ __forceinline__ int nav_dummy_pm(uint64_t seed, int64_t bit_idx) {
    // src/signal/ca_code.rs:94-108
    const int64_t pos = rem_euclid(bit_idx, kNavSubframeBits);
    int bit;
    if (pos < 8) {
        bit = (0x8B >> (7 - pos)) & 1;  // preamble 1000 1011
    } else {
        uint64_t x = seed ^ (static_cast<uint64_t>(bit_idx) * 0x9E3779B97F4A7C15ULL);
        x ^= x >> 30;
        x *= 0xBF58476D1CE4E5B9ULL;
        x ^= x >> 27;
        x *= 0x94D049BB133111EBULL;
        x ^= x >> 31;
        bit = static_cast<int>(x & 1);
    }
    return 1 - 2 * bit;
}

__device__ __forceinline__ bool gate_on(const TxDev &T, double t) {
    // src/signal/gate.rs:67-77
    double u = rem_euclid_f64(t + T.gate_phase, T.gate_period);
    if (u >= T.gate_period) u = 0.0;
    bool on = false;
    for (int b = 0; b < T.n_bursts; ++b) on |= (u >= T.bursts[2 * b]) && (u < T.bursts[2 * b + 1]);
    return on;
}

// SignalModel::evaluate_baseband_batch, one sample.
__device__ __forceinline__ void baseband(const TxDev &T, double t, float &re, float &im) {
    re = 0.0f;
    im = 0.0f;
    if (T.gated && !gate_on(T, t)) return;
    switch (T.signal_kind) {
    case 0: {  // CW, src/signal/cw.rs:41-52
        const float ph = T.cw_anchored ? static_cast<float>(rem_euclid_f64(T.omega * t + T.anchor_phase, kTau))
                                       : static_cast<float>(T.omega * t);
        re = cosf(ph);
        im = sinf(ph);
        return;
    }
    case 1:    // BPSK, src/signal/bpsk.rs:58-71
    case 2:    // OOK, src/signal/ook.rs:60-74
    case 3: {  // QAM16, src/signal/qam16.rs:96-109
        const int64_t raw = sat_i64(floor((t - T.epoch_rel) * T.rate));
        const int64_t k = frame_index(raw, T.n_symbols, T.repeating, T.repeat_from);
        if (k < 0) return;
        if (T.signal_kind == 3) {
            const float2 s = reinterpret_cast<const float2 *>(T.symbols)[k];
            re = s.x;
            im = s.y;
        } else {
            re = static_cast<float>(T.bits[k]);
        }
        return;
    }
    case 4: {  // C/A, src/signal/ca_code.rs:167-193
        const double dt = t - T.epoch_rel;
        const int64_t idx = sat_i64(floor(dt * kCaChipRate));
        int chip = T.ca_code[rem_euclid(idx, kCaCodeLength)];
        if (T.nav_mode == 1) {
            chip *= nav_dummy_pm(T.nav_seed, sat_i64(floor(dt * kNavBitRate)));
        } else if (T.nav_mode == 2) {
            const int64_t bi = sat_i64(floor(dt * kNavBitRate));
            int pm = 1;
            if (T.n_nav_bits > 0) pm = T.nav_bits[bi < 0 ? 0 : (bi < T.n_nav_bits ? bi : T.n_nav_bits - 1)];
            chip *= pm;
        }
        re = static_cast<float>(chip);
        return;
    }
    default:
        return;
    }
}

// ---------------------------------------------------------------- clock

// ClockStream::delta_t_at (src/clock/mod.rs:329-346) over the materialised
// walk eps[k - eps_k0] = epsilon(k); clock start epoch 0.0.
__device__ __forceinline__ double clock_delta(const JobDev &J, double t) {
    const double elapsed = t - 0.0;
    const double x = elapsed * J.master;
    if (x <= 0.0) {
        const double e0 = J.walk ? J.eps[0 - J.eps_k0] : 0.0;
        return (e0 + 0.0) + (J.c0 + J.y0 * 0.0);
    }
    const int64_t k0 = static_cast<int64_t>(floor(x));
    const double frac = x - static_cast<double>(k0);
    double elo = 0.0, ehi = 0.0;
    if (J.walk) {
        elo = J.eps[k0 - J.eps_k0];
        ehi = J.eps[k0 + 1 - J.eps_k0];
    }
    const double eps = elo * (1.0 - frac) + ehi * frac;
    return (eps + 0.0) + (J.c0 + J.y0 * elapsed);
}

// ---------------------------------------------------------------- render

struct TapOut {
    double *t_emit, *range, *phase, *delta_t;
    uint32_t slot, on;
};

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// run of tiles so neighbouring tiles of a job share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, per = n >> 3, rem = n & 7u;
    return x * per + (x < rem ? x : rem) + (b >> 3);
}

template <int FORM>
__global__ void __launch_bounds__(kTile) render_kernel(const JobDev *__restrict__ jobs, int n_jobs,
                                                       const LinkDev *__restrict__ links,
                                                       const TxDev *__restrict__ txs,
                                                       uint32_t n_tiles, TapOut tap) {
    const uint32_t tile = xcd_tile(blockIdx.x, n_tiles);
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    const JobDev &J = jobs[lo];
    const uint64_t j = static_cast<uint64_t>(tile - J.tile0) * kTile + threadIdx.x;
    if (j >= J.clen) return;
    const uint64_t i = J.i0 + j;

    // src/pipeline.rs:918-926
    const double t_local = static_cast<double>(i) / J.fs;
    const double dlt = J.delta_t ? J.delta_t[j] : clock_delta(J, t_local);
    const double tu = t_local - dlt;

    float acc_re = 0.0f, acc_im = 0.0f;
    float cmp_re = 0.0f, cmp_im = 0.0f;  // Neumaier compensation (precision form)
    for (uint32_t s = 0; s < J.n_active; ++s) {
        const LinkDev L = links[J.link0 + s];
        const TxDev &T = txs[L.tx];
        double t_emit, r;
        solve_emission(T, J.obs, tu, J.epoch, t_emit, r);
        float bre, bim;
        baseband(T, t_emit, bre, bim);
        // src/channel/mod.rs:255-267
        const double delay = r * kInvC;
        const double phase = T.neg_tau_src * (delay + dlt) + L.tau_if * t_local;
        float cpf, spf;
        if (FORM == 0) {
            double sp, cp;
            sincos(phase, &sp, &cp);
            cpf = static_cast<float>(cp);
            spf = static_cast<float>(sp);
        } else {
            // Range-reduce the f64 phase to about [-pi, pi] with a three-part
            // Cody-Waite split of 2*pi (24 + 24 + 53 significant bits, so the
            // first two products are exact for |k| < 2^29), then f32 sincos.
            const double k = rint(phase * 0.15915494309189535);
            double rr = phase - k * 6.2831854820251465;
            rr = rr - k * -1.7484555314695172e-07;
            rr = rr - k * -6.8604979977715316e-15;
            sincosf(static_cast<float>(rr), &spf, &cpf);
        }
        const float amplitude = static_cast<float>(T.amp_factor / r);
        const float cre = amplitude * (bre * cpf - bim * spf);
        const float cim = amplitude * (bre * spf + bim * cpf);
        if (FORM == 0) {
            acc_re = acc_re + cre;
            acc_im = acc_im + cim;
        } else {
            // Neumaier TwoSum accumulation across transmitters.
            const float tr = acc_re + cre, ti = acc_im + cim;
            cmp_re += fabsf(acc_re) >= fabsf(cre) ? (acc_re - tr) + cre : (cre - tr) + acc_re;
            cmp_im += fabsf(acc_im) >= fabsf(cim) ? (acc_im - ti) + cim : (cim - ti) + acc_im;
            acc_re = tr;
            acc_im = ti;
        }
        if (tap.on && s == tap.slot && lo == 0) {
            tap.t_emit[j] = t_emit;
            tap.range[j] = r;
            tap.phase[j] = phase;
        }
    }
    if (FORM != 0) {
        acc_re = acc_re + cmp_re;
        acc_im = acc_im + cmp_im;
    }
    if (tap.on && lo == 0) tap.delta_t[j] = dlt;

    // AWGN, src/channel/mod.rs:328-333 (sigma = sqrt(P/2) as f32)
    if (J.sigma > 0.0f) {
        float zr, zi;
        awgn_normals(J.noise_seed, i, zr, zi);
        acc_re += J.sigma * zr;
        acc_im += J.sigma * zi;
    }
    // RxChain::apply_post_sum, src/rx/mod.rs:142-146 and :83-101
    int16_t code_re = 0, code_im = 0;
    for (int p = 0; p < J.n_post; ++p) {
        const PostDev &P = J.post[p];
        if (P.kind == 0) {
            acc_re *= P.a;
            acc_im *= P.a;
        } else if (P.kind == 1) {
            const float f = 1.0f - P.a * (acc_re * acc_re + acc_im * acc_im);
            acc_re *= f;
            acc_im *= f;
        } else {
            const float fs = P.a, lsb = P.b;
            const float xr = acc_re < -fs ? -fs : (acc_re > fs ? fs : acc_re);
            const float xi = acc_im < -fs ? -fs : (acc_im > fs ? fs : acc_im);
            const float qr = roundf(xr / lsb), qi = roundf(xi / lsb);
            acc_re = qr * lsb;
            acc_im = qi * lsb;
            code_re = static_cast<int16_t>(qr);
            code_im = static_cast<int16_t>(qi);
        }
    }
    reinterpret_cast<float2 *>(J.out)[j] = make_float2(acc_re, acc_im);
    if (J.codes) reinterpret_cast<short2 *>(J.codes)[j] = make_short2(code_re, code_im);
}

// ---------------------------------------------------------------- clock walk

// Sum of the kWalkPerThread increments a thread owns, sequentially.
__device__ __forceinline__ double walk_thread_sum(uint64_t seed, double sigma, int64_t k0) {
    double s = 0.0;
    for (int q = 0; q < kWalkPerThread; ++q) s = s + (0.0 + sigma * clock_normal(seed, k0 + q));
    return s;
}

// Fixed-tree inclusive scan of one double per thread over a 256-thread block
// (Hillis-Steele in LDS); the same tree on every call.
__device__ __forceinline__ double block_inclusive_scan(double v, double *lds) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kWalkThreads; off <<= 1) {
        const double add = t >= off ? lds[t - off] : 0.0;
        __syncthreads();
        if (t >= off) lds[t] = lds[t] + add;
        __syncthreads();
    }
    return lds[t];
}

__global__ void __launch_bounds__(kWalkThreads) walk_blocksum_kernel(uint64_t seed, double sigma,
                                                                     int64_t first_block,
                                                                     double *blocksum) {
    __shared__ double lds[kWalkThreads];
    const int64_t b = first_block + blockIdx.x;
    const double s = walk_thread_sum(seed, sigma, b * kWalkBlock + int64_t(threadIdx.x) * kWalkPerThread);
    const double inc = block_inclusive_scan(s, lds);
    if (threadIdx.x == kWalkThreads - 1) blocksum[blockIdx.x] = inc;
}

// base[b + 1] = base[b] + blocksum[b], left to right (one lane).
__global__ void walk_carry_kernel(double *base, int64_t first_block, int64_t n_blocks,
                                  const double *blocksum) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = base[first_block];
    for (int64_t q = 0; q < n_blocks; ++q) {
        acc = acc + blocksum[q];
        base[first_block + q + 1] = acc;
    }
}

// eps(k) for k in [k_lo, k_hi]: out[k - k_lo] = base[b] + (prefix of thread
// totals before this thread + running sum of this thread's increments up to k).
__global__ void __launch_bounds__(kWalkThreads) walk_materialize_kernel(uint64_t seed, double sigma,
                                                                        int64_t k_lo, int64_t k_hi,
                                                                        int64_t first_block,
                                                                        const double *base,
                                                                        double *out) {
    __shared__ double lds[kWalkThreads];
    // Increments staged in LDS, [q][thread] so a wavefront's accesses are
    // consecutive (64 KiB per workgroup).
    __shared__ double incr[kWalkPerThread * kWalkThreads];
    const int t = threadIdx.x;
    const int64_t b = first_block + blockIdx.x;
    const int64_t k0 = b * kWalkBlock + int64_t(t) * kWalkPerThread;
    double s = 0.0;
    for (int q = 0; q < kWalkPerThread; ++q) {
        const double v = 0.0 + sigma * clock_normal(seed, k0 + q);
        incr[q * kWalkThreads + t] = v;
        s = s + v;
    }
    block_inclusive_scan(s, lds);
    const double before = t ? lds[t - 1] : 0.0;
    const double bb = base[b];
    double p = 0.0;
    for (int q = 0; q < kWalkPerThread; ++q) {
        p = p + incr[q * kWalkThreads + t];
        const int64_t k = k0 + q;
        if (k >= k_lo && k <= k_hi) out[k - k_lo] = bb + (before + p);
    }
}

}  // namespace lpk

// ---------------------------------------------------------------- launchers
// Plain C++ entry points for lpk_api.hip (same library, not exported).

namespace lpk {

hipError_t launch_render(int form, const JobDev *jobs, int n_jobs, const LinkDev *links,
                         const TxDev *txs, uint32_t n_tiles, double *tap_t_emit, double *tap_range,
                         double *tap_phase, double *tap_delta_t, uint32_t tap_slot,
                         hipStream_t stream) {
    TapOut tap{tap_t_emit, tap_range, tap_phase, tap_delta_t, tap_slot, tap_t_emit ? 1u : 0u};
    if (form == 0)
        hipLaunchKernelGGL(render_kernel<0>, dim3(n_tiles), dim3(kTile), 0, stream, jobs, n_jobs, links,
                           txs, n_tiles, tap);
    else
        hipLaunchKernelGGL(render_kernel<1>, dim3(n_tiles), dim3(kTile), 0, stream, jobs, n_jobs, links,
                           txs, n_tiles, tap);
    return hipGetLastError();
}

hipError_t launch_walk_extend(uint64_t seed, double sigma, int64_t first_block, int64_t n_blocks,
                              double *base, double *blocksum_scratch, hipStream_t stream) {
    if (n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(walk_blocksum_kernel, dim3(uint32_t(n_blocks)), dim3(kWalkThreads), 0, stream,
                       seed, sigma, first_block, blocksum_scratch);
    hipLaunchKernelGGL(walk_carry_kernel, dim3(1), dim3(64), 0, stream, base, first_block, n_blocks,
                       blocksum_scratch);
    return hipGetLastError();
}

hipError_t launch_walk_materialize(uint64_t seed, double sigma, int64_t k_lo, int64_t k_hi,
                                   const double *base, double *out, hipStream_t stream) {
    const int64_t b0 = k_lo / kWalkBlock, b1 = k_hi / kWalkBlock;
    hipLaunchKernelGGL(walk_materialize_kernel, dim3(uint32_t(b1 - b0 + 1)), dim3(kWalkThreads), 0,
                       stream, seed, sigma, k_lo, k_hi, b0, base, out);
    return hipGetLastError();
}

}  // namespace lpk
