// lpk_device.hpp -- device-side descriptors shared by the kernels and the
// host launcher. Everything here lives in HBM once per scene (TxDev) or once
// per render call (JobDev, LinkDev).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lpk {

// Physical constants, as the reference defines them.
constexpr double kSpeedOfLight = 299792458.0;          // src/lib.rs:53
constexpr double kInvC = 1.0 / kSpeedOfLight;           // src/channel/mod.rs:66
constexpr double kEmissionInitialLag = 0.07;            // src/channel/mod.rs:23
constexpr double kEarthRotation = 7.2921151467e-5;      // src/channel/mod.rs:27
constexpr double kTau = 6.283185307179586;              // std::f64::consts::TAU
constexpr double kCaChipRate = 1023000.0;               // src/signal/ca_code.rs:23
constexpr int kCaCodeLength = 1023;                     // src/signal/ca_code.rs:24
constexpr double kNavBitRate = 50.0;                    // src/signal/ca_code.rs:27
constexpr double kChipsPerNavBit = kCaChipRate / kNavBitRate;   // 20 460, exact
constexpr int64_t kNavSubframeBits = 300;               // src/signal/ca_code.rs:29
constexpr int kNewtonIterations = 2;                    // src/pipeline.rs:56

// Philox4x64-10 stream tags: one key word separates the clock walk from the
// noise stream drawn with the same seed.
constexpr uint64_t kClockTag = 0x636C6F636B77616CULL;   // "clockwal"
constexpr uint64_t kAwgnTag = 0x6177676E6E6F6973ULL;    // "awgnnois"
constexpr uint64_t kFlickerTag = 0x666C69636B65726EULL; // "flickern"

// Flicker FM (src/clock/mod.rs:55-149): a sum of kFlickerPoles AR(1) processes
// log-spaced in frequency from kFlickerFLo to half of kFlickerRate, stepped on
// a kFlickerRate coarse grid; its integral D[k] is interpolated linearly.
constexpr double kFlickerRate = 128.0;                  // src/clock/mod.rs:67
constexpr int kFlickerPoles = 20;                       // src/clock/mod.rs:68
constexpr double kFlickerFLo = 1e-4;                    // src/clock/mod.rs:69
constexpr int kFlickerCalibM = 256;                     // tau = 2 s (src/clock/mod.rs:106)

// Clock walk: blocks of kWalkBlock master-grid increments; each of the
// kWalkThreads threads of a block owns kWalkPerThread consecutive ones.
// epsilon(k) = base[k >> 16] (f64) + toff[(k - k0) >> 5] (f64) + local[k - k0] (f32).
constexpr int kWalkThreads = 256;
constexpr int kWalkPerThread = 256;
constexpr int kWalkSubShift = 5;
constexpr int kWalkBlockShift = 16;
constexpr int64_t kWalkBlock = int64_t(1) << kWalkBlockShift;
static_assert(kWalkThreads * kWalkPerThread == (1 << kWalkBlockShift), "walk block shape");

// Render tile: one workgroup of 2^tile_shift lanes, one output sample per
// lane; tiles sit at absolute multiples of the tile size in a band's sample
// index. The tile size is a tuning choice (128, 256 or 512 samples).
constexpr int kDefaultTileShift = 8;
// The precision form without an antenna works on super-tiles of
// 2^kSuperShift tiles per workgroup: the anchor solves of all of them run as
// one batch, so their latency is paid once per super-tile.
constexpr int kSuperShift = 3;
// log2 of the samples one workgroup covers (the unit of a job's tile count).
__host__ __device__ constexpr int span_shift(int form, bool antenna, int tile_shift) {
    return tile_shift + (form == 1 && !antenna ? kSuperShift : 0);
}
// Transmitters per pass of the precision form. A 72-link band rendered in one
// pass (one tile per anchor batch, two record groups) made a device memory
// fault that no host-side check of the kernels' indexing has explained, so a
// pass holds at most the count rendered cleanly on the device (four tiles per
// anchor batch, one record group): a render call splits a larger band-chunk
// into passes of this many links, which add their sums in f64 through HBM
// (JobDev::part). Streaming plans take at most this many per band.
constexpr uint32_t kPrecisionMaxLinks = 32;
// A dispatch packet counts its grid in work-items with 32 bits, so one
// launch covers at most 2^32 / 512 - 1 tiles of the largest size; a render
// call is sliced into launches of kMaxLaunchTiles.
constexpr uint32_t kMaxLaunchTiles = 1u << 22;

// One receiver's share of a batched clock-walk launch: walk blocks
// [first_block, first_block + n_blocks); blocks from win_block on also store
// the window the render reads (local per index, toff per 32 indices, both
// indexed from win_block * kWalkBlock). blocksum is indexed from first_block.
struct WalkSegDev {
    uint64_t seed;
    double sigma;
    int64_t first_block, win_block;
    uint32_t block0, n_blocks;  // the segment's first workgroup and its count
    float *local;
    double *toff;
    double *blocksum;
};

// base[first + q + 1] = base[first + q] + blocksum[q] for q < n, in order.
struct CarrySegDev {
    double *base;
    const double *blocksum;
    int64_t first, n;
};

struct TxDev {
    int32_t traj_kind;       // 0 stationary, 1 grid
    int32_t signal_kind;     // lpk_signal_kind
    double pos[3];
    double grid_t0, grid_dt;
    double grid_inv_dt;      // 1 / grid_dt (the converged form's grid coordinate)
    int64_t grid_n;
    double grid_hi;          // grid_n - 2: the last cell's first node (grid_n < 2^31)
    const double *grid;      // (3, n)
    double carrier;
    double neg_tau_src;      // -TAU * f_src            (src/channel/mod.rs:239)
    double amp_factor;       // lambda/(2 TAU) * sqrt(P) (src/channel/mod.rs:243)
    // CW
    double omega;            // TAU * baseband offset  (src/signal/cw.rs:42)
    double anchor_phase;
    int32_t cw_anchored;
    // framed
    int32_t repeating;
    double rate, epoch_rel;
    int64_t repeat_from;
    int64_t n_symbols;
    const int8_t *bits;
    const float *symbols;    // interleaved
    // C/A
    int32_t nav_mode;
    int32_t gated;
    const int8_t *ca_code;
    const float *ca_chips2;  // the chips as f32, the code twice over (2046 values)
    uint64_t nav_seed;
    const int8_t *nav_bits;
    int64_t n_nav_bits;
    // gate
    int32_t n_bursts, _pad;
    double gate_period, gate_phase;
    const double *bursts;
    // circular Keplerian elements (src/geometry/kepler.rs:19-25), radians and
    // rad/s, for the learned warm start's features; kep_on = 0 when the
    // trajectory has none (the start is then the straight-line light time)
    int32_t kep_on, _kpad;
    double kep_a, kep_cos_i, kep_sin_i, kep_raan, kep_nu0, kep_n, kep_epoch;
    // horizon gate (src/channel/horizon.rs:164-170): speed and distance bounds
    // of the trajectory and the position slack; gate_ok = 0 when not finite
    int32_t gate_ok, _gpad2;
    double max_speed, max_radius, gate_slack;
};

// Horizon gate (src/channel/horizon.rs): a (tile, transmitter) pair whose
// arrival elevation is certainly below the antenna cutoff for every sample of
// the tile is skipped; the per-sample path would give each of them zero.
constexpr double kGateMarginDeg = 1.0e-3;     // src/channel/horizon.rs:64
constexpr double kGateMaxLightTime = 100.0;   // src/channel/horizon.rs:68
constexpr int kGateMaxLinks = 1024;           // links per tile the gate tests (more go per sample)
constexpr int kPredStageLinks = 256;          // converged forms: links per block of starts staged in LDS

// Converged emission-time solve (the converged and learned forms): the
// reference's iteration t <- tu - |R(w (tu - t)) P(t) - O| / c on the
// unrounded time axis, stopped at the first evaluation whose update is below
// kSolveTol seconds, from the reference's start tu - 0.07 s or from a learned
// start, for at most kColdEvals evaluations (the iteration contracts by about
// |dr/dt| / c per evaluation, so a poor start only costs evaluations). The
// learned start is the network's light time every kLwsSpacing seconds of
// reception time (whole tiles, at least one), converged there by the same
// iteration, and interpolated linearly in between. The light time's
// curvature bounds that interpolation: about 4e-16 s over 2 ms for MEO (one
// evaluation then passes the tolerance), 2e-13 s for a LEO overhead pass
// (|d2tau/dt2| < 4e-7 s/s^2; two there). A query the network's start needs
// more than kLearnedEvals evaluations to converge is counted as a fallback,
// and so is a sample's solve that needs more than that.
constexpr double kSolveTol = 1e-13;
constexpr double kLwsSpacing = 2e-3;
constexpr int kLearnedEvals = 3;
constexpr int kColdEvals = 12;

// One (band, active transmitter) pair of a render call, in summation order.
struct LinkDev {
    uint32_t tx;
    uint32_t _pad;
    double tau_if;           // TAU * (f_src - f_lo)   (src/channel/mod.rs:240)
};

struct PostDev {
    int32_t kind;            // lpk_post_kind
    int32_t _pad;
    float a, b;              // GAIN: g; NONLINEARITY: k; ADC: full scale, lsb
};

// Transmitter groups of the precision form: a job's links are ordered framed
// (BPSK, OOK, QAM16), CW, everything else (gated transmitters, anchored CW),
// then C/A, and each group has a sample loop of its own. C/A comes last
// because its chip reads are the only vector-memory loads of the loops' fast
// paths: loads complete in order, so a chip read waits for the clock-walk
// stage of the next tile issued before the loops, which the other groups'
// work covers. The precision form sums exactly, so the order does not change
// its result; the plain form keeps the reference's order.
constexpr int kGroupFramed = 0, kGroupCw = 1, kGroupGeneric = 2, kGroupCa = 3;

struct JobDev {
    uint32_t tile0;          // first tile of this job in the launch
    uint32_t n_tiles;
    uint32_t link0, n_active;
    uint32_t group_end[3];   // precision form: end of the framed, CW and other groups
    uint32_t _gpad;
    double fs, lo, epoch;
    uint64_t i0, clen;
    double obs[3];           // stationary receiver position
    // moving receiver (rx_kind 1, CircularKeplerian, src/geometry/kepler.rs:27-96):
    // radius, mean motion, true anomaly at its epoch (rad), sin and cos of the
    // inclination and of the node, and the elements' epoch (unix s)
    int32_t rx_kind, _rxpad;
    double rk_a, rk_n, rk_nu0, rk_si, rk_ci, rk_so, rk_co, rk_epoch;
    // the precision form's smooth receiver model at the exact time
    // E + tu, E = epoch - rk_epoch: E, rk_nu0 + rk_n E and omega E, each as
    // a double-double (hi, lo)
    double rk_e[2], rk_nue[2], rk_the[2];
    // bounds of the receiver's speed, acceleration and jerk in the Earth-fixed
    // frame (0 for a fixed receiver): a w, a w^2, a w^3 with w = n + omega
    double rx_v, rx_a, rx_j;
    // clock
    const double *delta_t;   // host-provided offsets (device copy) or null
    const double *walk_base; // per 65 536-index block, absolute block index
    const double *walk_toff; // per 32 indices of the window
    const float *walk_local; // per index of the window
    int64_t eps_k0;          // first master index of the window (block aligned)
    double c0, y0, master;
    const double *flicker;   // flicker integral D[k] on the 128 Hz grid, or null
    int32_t walk;
    // AWGN and post-sum
    int32_t n_post;
    float sigma;
    int32_t _pad;
    uint64_t noise_seed;
    PostDev post[4];
    float *out;              // interleaved complex64
    int16_t *codes;          // optional ADC codes (2 per sample)
    // receiver antenna (src/rx/mod.rs:23-58): gain sqrt(G) at the arrival elevation
    int32_t ant_on, ant_cut;
    double ant_peak, ant_horizon, ant_fb, ant_cutoff;
    uint32_t *unocc;         // per link of the job: set to 1 where some sample's gain is non-zero
    // learned form: the job's predicted light times start at pred[pred0], one
    // row of pred_nb values per link, at every pred_qs-th tile boundary
    // (the last clamped to the job's end)
    uint64_t pred0;
    uint32_t pred_qs, pred_nb;
    // 1 / fs: the precision form forms its sample times by a reciprocal and
    // one correction step instead of a correctly rounded division
    double inv_fs;
    // walk_rel = 1: the clock window spans fewer than 2^31 master indices, so
    // its lookups take 32-bit indices relative to eps_k0 (as a double, exact),
    // with walk_base_w = walk_base + eps_k0 / 65 536 (the window's first block)
    const double *walk_base_w;
    double eps_k0d;
    int32_t walk_rel;
    // precision form in several passes over the job's links (kPrecisionMaxLinks
    // each): 0 one pass; 1 the first (writes its f64 sums to part), 2 a middle
    // one (adds part to its sums and writes them back), 3 the last (adds part,
    // then noise, post-sum ops and the stores)
    int32_t pass;
    double2 *part;           // per sample of the chunk (re, im), for passes 1-3
};

// High 64 bits of a 64 x 64-bit product, and count of leading zeros, on the
// device's instructions or the host's.
__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) >> 64);
#endif
}

__host__ __device__ __forceinline__ int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clz(static_cast<int>(x));
#else
    return __builtin_clz(x);
#endif
}

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11; the multipliers and
// Weyl key increments of their Random123 library) on counter c with key
// (k0, k1): 32x32 -> 64-bit products, which the vector ALU forms directly.
__host__ __device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = static_cast<uint32_t>(p1);
        c[2] = n2;
        c[3] = static_cast<uint32_t>(p0);
    }
}

// Block `block` of the (seed, tag) stream of the clock walk and the AWGN:
// Philox4x32-10 on counter (block, tag) with key `seed`, as two 64-bit words
// (outputs 1:0 and 3:2, high half first).
__host__ __device__ __forceinline__ void philox_words(uint64_t block, uint64_t seed, uint64_t tag, uint64_t w[2]) {
    uint32_t c[4] = {static_cast<uint32_t>(block), static_cast<uint32_t>(block >> 32), static_cast<uint32_t>(tag),
                     static_cast<uint32_t>(tag >> 32)};
    philox4x32(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    w[0] = (static_cast<uint64_t>(c[1]) << 32) | c[0];
    w[1] = (static_cast<uint64_t>(c[3]) << 32) | c[2];
}

// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__host__ __device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        const uint64_t m0 = 0xD2E7470EE14C6C93ULL, m1 = 0xCA5A826395121157ULL;
        const uint64_t hi0 = mulhi64(m0, c[0]), lo0 = m0 * c[0];
        const uint64_t hi1 = mulhi64(m1, c[2]), lo1 = m1 * c[2];
        const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// Box-Muller pair from one 64-bit Philox word: u1 = n 2^-24 with n = 2 m1 + 1
// odd from bits 63..41 (24 bits, exact in f32), and the angle u2 = m2 2^-24
// turns from bits 39..16. The pair is evaluated in f32 with correctly rounded
// operations only (+ - * / sqrt and explicit fmaf, no contraction, no libm
// transcendental), so every IEEE-754 machine yields the same bits.
//   ln u1: n = f 2^e with f in [sqrt(1/2), sqrt(2)] split exactly (u1 near 1
//   keeps its relative accuracy); ln f = 2 atanh(s), s = (f-1)/(f+1),
//   |s| <= 0.1716, by its series through s^9 (truncation < 1e-9 relative).
//   sin/cos of 2 pi u2: the quadrant (top 2 bits) and the reflection about
//   pi/4 are exact integer steps; sin and cos of the remaining angle
//   (<= pi/4) come from their Taylor series through the 9th and 10th powers.
__host__ __device__ __forceinline__ void box_muller(uint64_t w, float &zc, float &zs) {
    const uint32_t n = (static_cast<uint32_t>(w >> 41) << 1) | 1u;
    int e = 31 - clz32(n);
    float f = ldexpf(static_cast<float>(n), -e);
    if (f > 1.41421356f) {
        f = f * 0.5f;
        e += 1;
    }
    const float s = (f - 1.0f) / (f + 1.0f);
    const float z = s * s;
    float p = fmaf(z, 1.0f / 9.0f, 1.0f / 7.0f);
    p = fmaf(p, z, 1.0f / 5.0f);
    p = fmaf(p, z, 1.0f / 3.0f);
    p = fmaf(p, z, 1.0f);
    const float ln_u1 = fmaf(static_cast<float>(e - 24), 0.693147182f, (2.0f * s) * p);
    const float r = sqrtf(-2.0f * ln_u1);

    const uint32_t m2 = static_cast<uint32_t>(w >> 16) & 0xFFFFFFu;
    const uint32_t quad = m2 >> 22, v = m2 & 0x3FFFFFu;
    const bool flip = v > 0x200000u;
    const float th = static_cast<float>(flip ? 0x400000u - v : v) * (1.57079632679489662f * 0x1.0p-22f);
    const float t2 = th * th;
    float ps = fmaf(t2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = fmaf(ps, t2, 1.0f / 120.0f);
    ps = fmaf(ps, t2, -1.0f / 6.0f);
    float pc = fmaf(t2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = fmaf(pc, t2, -1.0f / 720.0f);
    pc = fmaf(pc, t2, 1.0f / 24.0f);
    pc = fmaf(pc, t2, -0.5f);
    const float sn = fmaf(th * t2, ps, th);
    const float cs = fmaf(t2, pc, 1.0f);
    const float a = flip ? cs : sn, b = flip ? sn : cs;  // sin, cos within the quadrant
    float sv, cv;
    if (quad == 0) { sv = a; cv = b; }
    else if (quad == 1) { sv = b; cv = -a; }
    else if (quad == 2) { sv = -a; cv = -b; }
    else { sv = -b; cv = a; }
    zc = r * cv;
    zs = r * sv;
}

}  // namespace lpk
