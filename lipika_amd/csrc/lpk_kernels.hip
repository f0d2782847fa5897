// lpk_kernels.hip -- gfx950 kernels of the IQ-generation backend.
//
// render_kernel<FORM, ANT, TS>: one workgroup renders one tile of 2^TS
// consecutive output samples (tiles sit at absolute multiples of 2^TS in the band's
// sample index) of one (receiver, band, chunk) job. Every tile of every job of
// a render call is one launch. Each lane owns one sample and walks the job's
// in-band transmitters in summation order:
//   light-time + Sagnac solve (src/channel/mod.rs:70-99)
//   -> s(t_emit) (src/signal/*.rs) -> FSPL amplitude and carrier phase
//   -> f32 mix (src/channel/mod.rs:251-268) -> sum over transmitters
//   (src/pipeline.rs:965-976) -> AWGN (src/channel/mod.rs:323-334)
//   -> post-sum ops (src/rx/mod.rs:83-101) -> one coalesced complex64 store.
//
// FORM 0 (plain) is the reference's f64 path operation for operation. FORM 1
// (precision) keeps the solve in f64 but evaluates it per tile and
// transmitter: the smooth two-iteration map is fitted by a quadratic in the
// reception time through three anchor samples, and each lane adds the exact
// effect of the reference's rounding of its absolute query times (first
// order, through the range rates); a tile whose query times cross a
// trajectory grid node takes the exact per-sample solve instead. The phase
// is formed in f64 in the reference's operation order and reduced exactly;
// sine and cosine (hardware, in turns) and the mix run in f32, and the f32
// contributions are summed in f64 (stronger than a compensated f32 sum).
//
// FORM 2 (converged) is the plain form with the emission-time solve iterated
// to a residual tolerance (kSolveTol) on the unrounded time axis instead of
// the reference's two fixed iterations; it starts either from the
// reference's t - 0.07 s or, in the learned form, from the light times of
// lws_predict_kernel at the tile's query boundaries (the network's
// predictions, converged there by the same iteration), interpolated.
//
// walk_gen_kernel: the receiver clock's white-FM walk (src/clock/mod.rs:
// 299-307) on the master grid, as fixed 65 536-index blocks; values depend
// only on (seed, index), never on how the render is chunked.
//
// Everything is compiled with -ffp-contract=off so the plain form keeps the
// reference's rounding; the precision form asks for its FMAs explicitly.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "lpk_device.hpp"
#include "lpk_lws.hpp"

namespace lpk {

// ---------------------------------------------------------------- helpers

// Rust `x as i64` after floor(): saturating, NaN -> 0.
__device__ __forceinline__ int64_t sat_i64(double x) {
    if (!(x == x)) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return static_cast<int64_t>(x);
}

// i64::rem_euclid for n > 0, with a 32-bit path for every index a real scene
// produces.
__device__ __forceinline__ int64_t rem_euclid(int64_t a, int64_t n) {
    if (a >= INT32_MIN && a <= INT32_MAX && n <= INT32_MAX) {
        const int32_t r = static_cast<int32_t>(a) % static_cast<int32_t>(n);
        return r < 0 ? r + n : r;
    }
    const int64_t r = a % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ double rem_euclid_f64(double a, double b) {
    const double r = fmod(a, b);
    return r < 0.0 ? r + fabs(b) : r;
}

// ---------------------------------------------------------------- geometry

// PrecomputedTrajectory::position_ecef (src/geometry/precomputed.rs:111-121):
// the clamp of floor(f) to [0, n-2] is done in f64, which is exactly Rust's
// saturating cast + clamp (NaN -> 0 included). `cell` gets the node index.
__device__ __forceinline__ void tx_position(const TxDev &T, double t_abs, double &x, double &y,
                                            double &z, int64_t &cell) {
    if (T.traj_kind == 0) {
        x = T.pos[0];
        y = T.pos[1];
        z = T.pos[2];
        cell = 0;
        return;
    }
    const double f = (t_abs - T.grid_t0) / T.grid_dt;
    const double flo = fmin(fmax(floor(f), 0.0), T.grid_hi);
    const int32_t lo = static_cast<int32_t>(flo);
    double frac = f - flo;
    frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
    const double omf = 1.0 - frac;
    const double *g = T.grid;
    const int64_t n = T.grid_n;
    x = g[lo] * omf + g[lo + 1] * frac;
    y = g[n + lo] * omf + g[n + lo + 1] * frac;
    z = g[2 * n + lo] * omf + g[2 * n + lo + 1] * frac;
    cell = lo;
}

// The receiver's position at reception time tu (epoch-relative), queried at
// the absolute time epoch + tu as the chunk body queries position_ecef_batch
// (src/pipeline.rs:929-930): fixed (src/geometry/stationary.rs:21-24), or
// CircularKeplerian::position_ecef (src/geometry/kepler.rs:27-85, eci_at and
// eci_to_ecef in the reference's operation order; the device's f64 sincos is
// within an ulp of glibc's).
__device__ __forceinline__ void rx_position(const JobDev &J, double tu, double *o) {
    if (J.rx_kind == 0) {
        o[0] = J.obs[0];
        o[1] = J.obs[1];
        o[2] = J.obs[2];
        return;
    }
    const double t_unix = J.epoch + tu;
    const double dt = t_unix - J.rk_epoch;
    const double nu = J.rk_nu0 + J.rk_n * dt;
    double s_nu, c_nu;
    sincos(nu, &s_nu, &c_nu);
    const double xp = J.rk_a * c_nu;
    const double yp = J.rk_a * s_nu;
    const double y1 = yp * J.rk_ci;
    const double z1 = yp * J.rk_si;
    const double x_eci = xp * J.rk_co - y1 * J.rk_so;
    const double y_eci = xp * J.rk_so + y1 * J.rk_co;
    const double theta = kEarthRotation * (t_unix - J.rk_epoch);
    double s, c;
    sincos(theta, &s, &c);
    o[0] = x_eci * c + y_eci * s;
    o[1] = -x_eci * s + y_eci * c;
    o[2] = z1;
}

// The precision form's smooth receiver: the orbit of rx_position at the exact
// time E + tu, E = epoch - rk_epoch (true anomaly nu = nu0 + n (E + tu) and
// Earth angle theta = omega (E + tu), each formed as a double-double from the
// job's constants), with the position's derivatives along the two angles (pn
// along nu, pt along theta). The angles' low parts join to first order.
__device__ __forceinline__ void rx_smooth(const JobDev &J, double tu, double *o, double *pn, double *pt) {
    double hi[2], lo[2];
    const double *cst[2] = {J.rk_nue, J.rk_the};
    const double rate[2] = {J.rk_n, kEarthRotation};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double p = rate[k] * tu, pe = fma(rate[k], tu, -p);
        const double s = cst[k][0] + p, bb = s - cst[k][0];
        hi[k] = s;
        lo[k] = ((cst[k][0] - (s - bb)) + (p - bb)) + (pe + cst[k][1]);
    }
    double s_nu, c_nu, s, c;
    sincos(hi[0], &s_nu, &c_nu);
    sincos(hi[1], &s, &c);
    const double xp = J.rk_a * c_nu, yp = J.rk_a * s_nu;
    // ECI position and its derivative along nu (R_z(node) R_x(inclination))
    const double x = xp * J.rk_co - yp * J.rk_ci * J.rk_so, y = xp * J.rk_so + yp * J.rk_ci * J.rk_co;
    const double dx = -yp * J.rk_co - xp * J.rk_ci * J.rk_so, dy = -yp * J.rk_so + xp * J.rk_ci * J.rk_co;
    // to ECEF by R_z(-theta); d/d theta of (X, Y, Z) is (Y, -X, 0)
    const double ex = x * c + y * s, ey = -x * s + y * c;
    pn[0] = dx * c + dy * s;
    pn[1] = -dx * s + dy * c;
    pn[2] = xp * J.rk_si;
    pt[0] = ey;
    pt[1] = -ex;
    pt[2] = 0.0;
    o[0] = fma(pn[0], lo[0], fma(ey, lo[1], ex));
    o[1] = fma(pn[1], lo[0], fma(-ex, lo[1], ey));
    o[2] = fma(pn[2], lo[0], yp * J.rk_si);
}

// How rx_position's query at the rounded time epoch + tu departs from the
// smooth receiver at E + tu: dn in the true anomaly and dth in the Earth
// angle, each exact (error-free products and sums): the query time's own
// rounding d, times the angle's rate, less the roundings of n dt, nu0 + n dt
// and omega dt. Zero for a fixed receiver.
__device__ __forceinline__ void rx_rounding(const JobDev &J, double tu, double &dn, double &dth) {
    if (J.rx_kind == 0) {
        dn = dth = 0.0;
        return;
    }
    const double dt = (J.epoch + tu) - J.rk_epoch;                     // as rx_position forms it
    const double d = ((dt - J.rk_e[0]) - tu) - J.rk_e[1];              // dt less E + tu
    const double p = J.rk_n * dt, e1 = fma(J.rk_n, dt, -p);            // n dt = p + e1
    const double s = J.rk_nu0 + p, bb = s - J.rk_nu0;
    const double e2 = (J.rk_nu0 - (s - bb)) + (p - bb);                // nu0 + p = s + e2
    dn = fma(J.rk_n, d, -(e1 + e2));
    const double th = kEarthRotation * dt, e3 = fma(kEarthRotation, dt, -th);
    dth = fma(kEarthRotation, d, -e3);
}

// solve_emission_times_for, one sample (src/channel/mod.rs:70-99). cell0 and
// cell1 are the grid nodes the two iterations' queries fell in.
__device__ __forceinline__ void solve_emission(const TxDev &T, const double *obs, double tu,
                                               double epoch, double &t_emit, double &r,
                                               int64_t &cell0, int64_t &cell1) {
    double t = tu - kEmissionInitialLag;
    r = 0.0;
#pragma unroll
    for (int it = 0; it < kNewtonIterations; ++it) {
        double sx, sy, sz;
        int64_t cell;
        tx_position(T, epoch + t, sx, sy, sz, cell);
        if (it == 0) cell0 = cell;
        else cell1 = cell;
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

// solve_emission for the precision form's tiles that cross a trajectory grid
// node: the same iteration with the grid coordinate scaled by 1 / dt instead
// of divided by dt (one rounding more in the coordinate, under 1e-13 m of
// position; the lerp is continuous across a node, so a coordinate that lands
// on the other side of one moves nothing).
__device__ __forceinline__ double solve_emission_scaled(const TxDev &T, const double *obs, double tu, double epoch) {
    double t = tu - kEmissionInitialLag, r = 0.0;
#pragma unroll
    for (int it = 0; it < kNewtonIterations; ++it) {
        double sx, sy, sz;
        if (T.traj_kind == 0) {
            sx = T.pos[0];
            sy = T.pos[1];
            sz = T.pos[2];
        } else {
            const double f = ((epoch + t) - T.grid_t0) * T.grid_inv_dt;
            const double flo = fmin(fmax(floor(f), 0.0), T.grid_hi);
            const int32_t lo = static_cast<int32_t>(flo);
            double frac = f - flo;
            frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
            const double omf = 1.0 - frac;
            const double *g = T.grid;
            const int64_t n = T.grid_n;
            sx = g[lo] * omf + g[lo + 1] * frac;
            sy = g[n + lo] * omf + g[n + lo + 1] * frac;
            sz = g[2 * n + lo] * omf + g[2 * n + lo + 1] * frac;
        }
        const double wtau = kEarthRotation * (tu - t);
        const double dx = (sx + wtau * sy) - obs[0];
        const double dy = (sy - wtau * sx) - obs[1];
        const double dz = sz - obs[2];
        r = sqrt(dx * dx + dy * dy + dz * dz);
        t = tu - r * kInvC;
    }
    return r;
}

// PrecomputedTrajectory::position_ecef (src/geometry/precomputed.rs:108-121)
// at time trel after the grid's node 0, given directly instead of as an
// absolute time, so it is not rounded at the scale of the epoch, and scaled
// by 1 / dt instead of divided by dt (one rounding more in the grid
// coordinate: under 1e-10 m of position).
__device__ __forceinline__ void tx_position_rel(const TxDev &T, double trel, double &x, double &y, double &z) {
    if (T.traj_kind == 0) {
        x = T.pos[0];
        y = T.pos[1];
        z = T.pos[2];
        return;
    }
    const double f = trel * T.grid_inv_dt;
    const double flo = fmin(fmax(floor(f), 0.0), T.grid_hi);
    const int32_t lo = static_cast<int32_t>(flo);
    double frac = f - flo;
    frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
    const double omf = 1.0 - frac;
    const double *g = T.grid;
    const int64_t n = T.grid_n;
    x = g[lo] * omf + g[lo + 1] * frac;
    y = g[n + lo] * omf + g[n + lo + 1] * frac;
    z = g[2 * n + lo] * omf + g[2 * n + lo + 1] * frac;
}

// One evaluation of the reference's iteration (src/channel/mod.rs:75-96) at
// emission time t, with positions at eoff + t after the grid's node 0 (eoff =
// scene epoch - grid start): returns tu - r / c and sets r.
__device__ __forceinline__ double emission_step(const TxDev &T, const double *obs, double tu, double eoff, double t,
                                                double &r) {
    double sx, sy, sz;
    tx_position_rel(T, eoff + t, sx, sy, sz);
    const double wtau = kEarthRotation * (tu - t);
    const double sx_rot = sx + wtau * sy;
    const double sy_rot = sy - wtau * sx;
    const double dx = sx_rot - obs[0];
    const double dy = sy_rot - obs[1];
    const double dz = sz - obs[2];
    r = sqrt(dx * dx + dy * dy + dz * dz);
    return tu - r * kInvC;
}

// The converged solve: the iteration from emission time t0 until the first
// update |t_next - t| < kSolveTol. Callers run its first evaluation
// (t_emit = t1 from t0, with range r) and stop there when its update is
// small, the usual case from a learned start; this goes on from t1 for every
// other lane. Returns the number of evaluations, or -1 when max_evals did
// not converge; t_emit and r are then the last iterate and the range it came
// from (r / c = tu - t_emit).
__device__ __forceinline__ int solve_converged_tail(const TxDev &T, const double *obs, double tu, double eoff,
                                                    int max_evals, double &t_emit, double &r) {
    double t = t_emit;
    for (int e = 2; e <= max_evals; ++e) {
        const double tn = emission_step(T, obs, tu, eoff, t, r);
        const double step = tn - t;
        t = tn;
        if (fabs(step) < kSolveTol) {
            t_emit = t;
            return e;
        }
    }
    t_emit = t;
    return -1;
}

// certainly_below_cutoff (src/channel/horizon.rs:150-210) for a stationary
// receiver (ball radius 0 around O) over reception times [t_min, t_max]
// (epoch-relative). True only if every sample of that span would get zero gain.
__device__ __forceinline__ bool gate_below_cutoff(const JobDev &J, const TxDev &T, double t_min, double t_max,
                                                  const double *oc, double o_rad) {
    if (!T.gate_ok || !(t_min <= t_max)) return false;
    const double ox = oc[0], oy = oc[1], oz = oc[2];
    const double o_norm = sqrt(ox * ox + oy * oy + oz * oz);
    if (!isfinite(o_norm) || !isfinite(o_rad)) return false;
    const double lt = (1.01 * T.max_radius + o_norm + o_rad) * kInvC;
    const double light_time = lt > kEmissionInitialLag ? lt : kEmissionInitialLag;
    if (!(light_time <= kGateMaxLightTime)) return false;
    const double t_mag = fabs(t_min) > fabs(t_max) ? fabs(t_min) : fabs(t_max);
    const double lag = light_time * (1.0 + 1e-9) + 1e-6 + 1e-12 * t_mag;
    const double t_lo = J.epoch + (t_min - lag);
    const double t_hi = J.epoch + t_max;
    const double t_mid = t_lo + 0.5 * (t_hi - t_lo);
    const double h1 = t_mid - t_lo, h2 = t_hi - t_mid;
    const double half_span = h1 > h2 ? h1 : h2;
    double sx, sy, sz;
    int64_t cell;
    tx_position(T, t_mid, sx, sy, sz, cell);
    const double source_radius = T.max_speed * half_span * (1.0 + 1e-9) + T.gate_slack;
    const double dx = sx - ox, dy = sy - oy, dz = sz - oz;
    const double range = sqrt(dx * dx + dy * dy + dz * dz);
    const double rho = source_radius + o_rad;
    if (!(rho < range) || !(o_rad < o_norm)) return false;
    const double sin_el = (dx * ox + dy * oy + dz * oz) / (range * o_norm);
    const double c = sin_el < -1.0 ? -1.0 : (sin_el > 1.0 ? 1.0 : sin_el);
    const double deg = 180.0 / 3.141592653589793;
    const double el_centre = asin(c) * deg;
    const double spread = asin(rho / range) * deg + asin(o_rad / o_norm) * deg;
    return el_centre + spread + kGateMarginDeg < J.ant_cutoff;
}

// Antenna::amplitude_at (src/rx/mod.rs:48-58) at the arrival elevation of
// the source at the solved emission time (src/channel/mod.rs:190-207): the
// position is queried at epoch + t_emit (no Sagnac rotation), sin(el) =
// LOS . up with up = O / |O|, clamped, then asin and x * (180 / pi). Only the
// render kernels instantiated for antenna scenes contain it.
__device__ __forceinline__ float antenna_gain(const JobDev &J, const double *obs, const TxDev &T, double t_emit) {
    double sx, sy, sz;
    int64_t cell;
    tx_position(T, J.epoch + t_emit, sx, sy, sz, cell);
    const double ox = obs[0], oy = obs[1], oz = obs[2];
    const double on = sqrt(ox * ox + oy * oy + oz * oz);
    const double dx = sx - ox, dy = sy - oy, dz = sz - oz;
    const double dn = sqrt(dx * dx + dy * dy + dz * dz);
    const double sin_el = (dx * ox + dy * oy + dz * oz) / (dn * on);
    const double c = sin_el < -1.0 ? -1.0 : (sin_el > 1.0 ? 1.0 : sin_el);
    const double el = asin(c) * (180.0 / 3.141592653589793);
    if (J.ant_cut && el < J.ant_cutoff) return 0.0f;
    const double dbi = el >= 0.0 ? J.ant_horizon + (J.ant_peak - J.ant_horizon) * (el / 90.0) : J.ant_horizon - J.ant_fb;
    return static_cast<float>(pow(10.0, dbi / 20.0));
}

// ---------------------------------------------------------------- signals

// Rust's saturating `floor(x) as i64` for an already floored x, with a
// one-instruction path for every index a real scene produces (NaN fails the
// range test and takes sat_i64, which maps it to 0).
__device__ __forceinline__ int64_t floor_to_i64(double fl) {
    if (fabs(fl) < 2147483648.0) return static_cast<int64_t>(static_cast<int32_t>(fl));
    return sat_i64(fl);
}

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

// Value of symbol `u` (floor((t - epoch_rel) * rate), src/signal/mod.rs:108-115)
// of a framed signal: BPSK (src/signal/bpsk.rs:58-71), OOK
// (src/signal/ook.rs:60-74), QAM16 (src/signal/qam16.rs:96-109).
__device__ __forceinline__ void framed_value(const TxDev &T, double u, float &re, float &im) {
    re = 0.0f;
    im = 0.0f;
    const int64_t k = frame_index(floor_to_i64(u), T.n_symbols, T.repeating, T.repeat_from);
    if (k < 0) return;
    if (T.signal_kind == 3) {
        const float2 s = reinterpret_cast<const float2 *>(T.symbols)[k];
        re = s.x;
        im = s.y;
    } else {
        re = static_cast<float>(T.bits[k]);
    }
}

// NAV sign for bit index `b` (floor(dt * 50)), src/signal/ca_code.rs:94-108, :180-190.
__device__ __forceinline__ int nav_pm(const TxDev &T, double b) {
    if (T.nav_mode == 1) return nav_dummy_pm(T.nav_seed, floor_to_i64(b));
    if (T.nav_mode == 2 && T.n_nav_bits > 0) {
        const int64_t bi = floor_to_i64(b);
        return T.nav_bits[bi < 0 ? 0 : (bi < T.n_nav_bits ? bi : T.n_nav_bits - 1)];
    }
    return 1;
}

// C/A chip, src/signal/ca_code.rs:167-178.
__device__ __forceinline__ int ca_chip(const TxDev &T, double dt) {
    return T.ca_code[rem_euclid(floor_to_i64(floor(dt * kCaChipRate)), kCaCodeLength)];
}

// CW phase as the reference forms it in f32, src/signal/cw.rs:41-52.
__device__ __forceinline__ float cw_phase(const TxDev &T, double t) {
    return T.cw_anchored ? static_cast<float>(rem_euclid_f64(T.omega * t + T.anchor_phase, kTau))
                         : static_cast<float>(T.omega * t);
}

// SignalModel::evaluate_baseband_batch, one sample.
__device__ __forceinline__ void baseband(const TxDev &T, double t, float &re, float &im) {
    re = 0.0f;
    im = 0.0f;
    if (T.gated && !gate_on(T, t)) return;
    switch (T.signal_kind) {
    case 0:  // CW: cos, sin of one f32 phase (one sincosf)
        sincosf(cw_phase(T, t), &im, &re);
        return;
    case 1:
    case 2:
    case 3:
        framed_value(T, floor((t - T.epoch_rel) * T.rate), re, im);
        return;
    case 4: {
        const double dt = t - T.epoch_rel;
        re = static_cast<float>(ca_chip(T, dt) * nav_pm(T, floor(dt * kNavBitRate)));
        return;
    }
    default:
        return;
    }
}

// ---------------------------------------------------------------- clock

// epsilon(k) of the materialised walk window (k >= 0).
__device__ __forceinline__ double walk_eps(const JobDev &J, int64_t k) {
    const int64_t rel = k - J.eps_k0;
    return J.walk_base[k >> kWalkBlockShift] + J.walk_toff[rel >> kWalkSubShift] +
           static_cast<double>(J.walk_local[rel]);
}

// epsilon(eps_k0 + k) of the materialised walk window by its window-relative
// index k: the same three terms summed in the same order as walk_eps, with
// 32-bit index arithmetic.
__device__ __forceinline__ double walk_eps_rel(const JobDev &J, int32_t k) {
    return J.walk_base_w[k >> kWalkBlockShift] + J.walk_toff[k >> kWalkSubShift] +
           static_cast<double>(J.walk_local[k]);
}

// ClockStream::delta_t_at (src/clock/mod.rs:329-346); clock start epoch 0.0.
__device__ __forceinline__ double clock_delta(const JobDev &J, double t) {
    const double elapsed = t - 0.0;
    // Flicker::delta_at (src/clock/mod.rs:139-148): D interpolated on the 128 Hz grid
    double flick = 0.0;
    if (J.flicker) {
        const double xf = fmax(elapsed, 0.0) * kFlickerRate;
        const double mf = floor(xf);
        const int64_t m0 = static_cast<int64_t>(mf);
        const double fr = xf - mf;
        flick = J.flicker[m0] * (1.0 - fr) + J.flicker[m0 + 1] * fr;
    }
    const double x = elapsed * J.master;
    if (x <= 0.0) {
        const double e0 = J.walk ? walk_eps(J, 0) : 0.0;
        return (e0 + flick) + (J.c0 + J.y0 * 0.0);
    }
    double frac, elo = 0.0, ehi = 0.0;
    if (J.walk_rel) {
        // x - eps_k0 is exact (eps_k0 is an integer no greater than x), so the
        // floor and the fraction are those of x itself
        const double rx = x - J.eps_k0d;
        const double fl = floor(rx);
        frac = rx - fl;
        const int32_t k = static_cast<int32_t>(fl);
        elo = walk_eps_rel(J, k);
        ehi = walk_eps_rel(J, k + 1);
    } else {
        const double fl = floor(x);
        frac = x - fl;
        if (J.walk) {
            const int64_t k0 = static_cast<int64_t>(fl);
            elo = walk_eps(J, k0);
            ehi = walk_eps(J, k0 + 1);
        }
    }
    const double eps = elo * (1.0 - frac) + ehi * frac;
    return (eps + flick) + (J.c0 + J.y0 * elapsed);
}

// ---------------------------------------------------------------- render

// Sample time i / fs of the precision form: the product with the rounded
// reciprocal and one correction step, which rounds as the division does except
// in rare cases one ulp away (3.6e-15 s at 30 s).
__device__ __forceinline__ double sample_time(double i, double fs, double inv_fs) {
    const double q = i * inv_fs;
    return fma(fma(-q, fs, i), inv_fs, q);
}

// ClockStream::delta_t_at (src/clock/mod.rs:329-346) for the precision form
// of a band sampled at its receiver's master rate (fs == master): the master
// index x = (i / fs) * master lies within one ulp of the sample index i, so
// the interpolation weights are 1 and 0 (or within 1e-16 of them) and the
// walk value is epsilon(i) itself, read at one index instead of two. Where x
// rounds to just below i the reference weighs epsilon(i - 1) by about 1e-16;
// leaving that out moves delta_t by under 1e-27 s. The terms are summed in
// the reference's order.
__device__ __forceinline__ double clock_at_master_eps(const JobDev &J, double t, double eps) {
    const double elapsed = t - 0.0;
    double flick = 0.0;
    if (J.flicker) {
        const double xf = fmax(elapsed, 0.0) * kFlickerRate;
        const double mf = floor(xf);
        const int64_t m0 = static_cast<int64_t>(mf);
        const double fr = xf - mf;
        flick = J.flicker[m0] * (1.0 - fr) + J.flicker[m0 + 1] * fr;
    }
    return (eps + flick) + (J.c0 + J.y0 * elapsed);
}
__device__ __forceinline__ double clock_delta_at_master(const JobDev &J, double t, uint64_t i) {
    return clock_at_master_eps(J, t, walk_eps_rel(J, static_cast<int32_t>(static_cast<int64_t>(i) - J.eps_k0)));
}

struct TapOut {
    double *t_emit, *range, *phase, *delta_t;
    uint32_t slot, on;
};

// Per (tile, transmitter) record of the precision form, in LDS. The solve is
// fitted in two parts: the smooth map (trajectory queried at unrounded times),
// a quadratic in tu through three anchor samples, plus the first-order effect
// of the reference's rounding of the absolute query time epoch + t
// (src/channel/mod.rs:79), which the lanes compute exactly per sample. The
// carrier phase (src/channel/mod.rs:256-257) is linear in range, delta_t and
// t_local, so it is carried in turns from its value at the middle sample,
// reduced there once in double-double. The first part of the record holds
// everything the sample loop reads on its fast path, so the loop touches no
// transmitter table in global memory but the C/A chips.
struct alignas(16) Anchor {
    // read by every lane of the sample loop
    double r0m, q0d1, q0c2, rd0;   // r0(tu) = r0m + x (q0d1 + q0c2 x) + rd0 del0, x = tu - tu_mid
    double r1m, q1d1, q1c2, rd1q;  // r1(tu) = r1m + x (q1d1 + q1c2 x) + rd1q del1 - k01 del0
    double k01;                    // (d r1 / d t1, position and Sagnac) * rd0 / c
    double r1_mid;                 // range at the middle sample (model, or exact when not smooth)
    union {
        struct {                   // written by the fit, read by the sample loop
            double ph0;            // carrier phase at the middle sample, turns in [-1/2, 1/2]
            double ph_r, ph_x;     // its derivatives in turns per metre of range and per second of tu
            double ph_spare;
            double epoch_rel, rate;   // signal epoch; symbol rate (framed) or omega (CW)
            double hu, hu1;        // framed: the symbol index the tile starts in, and the next; C/A: the chip
                                   // offset from cbase where the tile's second NAV bit starts, and the first bit
            double cbase;          // C/A: a chip count just before the tile's first sample
            const float *chip0;    // C/A: the doubled f32 chip table from residue cmod on
        };
        struct {                   // the anchor pass's results, read by the fit
            double r0s[3], r1s[3]; // smooth ranges of iterations 1 and 2 at the anchors
            int64_t cell[2][2];    // grid nodes of both iterations at the first and last anchors
        };
    };
    float h0re, h0im;              // value of index hu: symbol (framed) or NAV sign (C/A, h0re)
    float h1re, h1im;              // value of index hu1 (the four are read as one 16-byte vector)
    float amp_m;                   // (float)(amp_factor / r1_mid)
    float amp_k;                   // amp_m / r1_mid: amplitude slope per metre of range
    uint32_t cmod;                 // C/A: cbase mod 1023
    uint32_t tx;                   // transmitter (slow paths read its table)
    double neg_tau_src, tau_if;    // -TAU f_src, TAU (f_src - f_lo), for taps
};
static_assert(sizeof(Anchor) == 208, "Anchor layout");

// A moving receiver's coefficients of a (tile, transmitter) record, beside
// its Anchor: r0 += n0 dn + t0 dth and r1 += n1 dn + t1 dth for the receiver
// query's departures dn, dth (rx_rounding) in its two angles.
struct RxCoef {
    double n0, t0, n1, t1;
};
static_assert(offsetof(Anchor, h0re) % 16 == 0, "hint values are one aligned 16-byte read");
constexpr int kAnchorGroup = 64;  // transmitters per anchor pass
constexpr int kSpanRecords = 128; // (tile, transmitter) records of one anchor batch of a span
constexpr int kSpanSamples = 2;    // adjacent samples per lane of the span kernels

// A wave-uniform 64-bit value from two 32-bit halves, in scalar registers.
__device__ __forceinline__ uint64_t uniform_u64(uint32_t lo, uint32_t hi) {
    const uint32_t l = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(lo)));
    const uint32_t h = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(hi)));
    return (static_cast<uint64_t>(h) << 32) | l;
}
constexpr double kInvTau = 0.15915494309189535;

// An exact Cody-Waite reduction of hi + lo radians (24 + 24 + 53-bit split of
// 2 pi, exact for |k| < 2^29) to turns in [-1/2, 1/2].
__device__ __forceinline__ double reduce_to_turns(double hi, double lo) {
    const double kk = rint(hi * kInvTau);
    const double rr =
        fma(-kk, -6.8604979977715316e-15, fma(-kk, -1.7484555314695172e-07, fma(-kk, 6.2831854820251465, hi)));
    return (rr + lo) * kInvTau;
}

// The carrier phase neg_tau_src (r / c + dlt) + tau_if t_local in turns,
// evaluated in double-double (error products and sums exact) before the
// reduction, so that it is the phase of the reference's f64 inputs to ~1e-16
// turns.
__device__ __forceinline__ double carrier_turns(double neg_tau_src, double tau_if, double r, double dlt, double t) {
    const double p = r * kInvC, pe = fma(r, kInvC, -p);
    const double d = p + dlt, dv = d - p, de = (p - (d - dv)) + (dlt - dv) + pe;
    const double a = neg_tau_src * d, ae = fma(neg_tau_src, d, -a) + neg_tau_src * de;
    const double b = tau_if * t, be = fma(tau_if, t, -b);
    const double s = a + b, sv = s - a, se = (a - (s - sv)) + (b - sv);
    return reduce_to_turns(s, se + ae + be);
}

// The two-iteration solve with the trajectory queried at unrounded times
// ((epoch - t0) + t instead of epoch + t); also returns the node cells and
// the range rates the rounding correction needs. With the receiver's
// derivatives along its two angles (pn, pt; null for none), cr receives each
// iteration's range derivatives along them: (nu, theta) of iteration 1, then
// of iteration 2.
__device__ __forceinline__ void solve_smooth(const TxDev &T, const double *obs, double tu, double eoff,
                                             double &r0, double &r1, int64_t &c0, int64_t &c1,
                                             double &rd0, double &rd1q, double &rd1t,
                                             const double *pn = nullptr, const double *pt = nullptr,
                                             double *cr = nullptr) {
    double t = tu - kEmissionInitialLag;
    double rr[2], rq[2], rt[2];
    int64_t cc[2];
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        double sx, sy, sz, vx = 0.0, vy = 0.0, vz = 0.0;
        int64_t cell = 0;
        if (T.traj_kind == 0) {
            sx = T.pos[0];
            sy = T.pos[1];
            sz = T.pos[2];
        } else {
            const double f = (eoff + t) * T.grid_inv_dt;
            const double flo = fmin(fmax(floor(f), 0.0), T.grid_hi);
            const int32_t lo = static_cast<int32_t>(flo);
            double frac = f - flo;
            frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
            const double *g = T.grid;
            const int64_t n = T.grid_n;
            const double ax = g[lo], ay = g[n + lo], az = g[2 * n + lo];
            const double bx = g[lo + 1], by = g[n + lo + 1], bz = g[2 * n + lo + 1];
            sx = fma(bx - ax, frac, ax);
            sy = fma(by - ay, frac, ay);
            sz = fma(bz - az, frac, az);
            const double idt = T.grid_inv_dt;
            vx = (bx - ax) * idt;
            vy = (by - ay) * idt;
            vz = (bz - az) * idt;
            cell = lo;
        }
        const double w = kEarthRotation * (tu - t);
        const double dx = fma(w, sy, sx) - obs[0];
        const double dy = fma(-w, sx, sy) - obs[1];
        const double dz = sz - obs[2];
        const double r = sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
        const double ir = __builtin_amdgcn_rcp(r);   // for the range rates only
        // d/d(query time) of the rotated position: V + w J V; of the Sagnac
        // angle: -omega J S, with J (x, y, z) = (y, -x, 0)
        const double pq = (dx * fma(w, vy, vx) + dy * fma(-w, vx, vy) + dz * vz) * ir;
        const double pw = (dx * sy - dy * sx) * ir * -kEarthRotation;
        if (pn) {
            // the receiver enters as -obs: d r / d angle = -(d . d obs / d angle) / r
            cr[2 * it] = -(dx * pn[0] + dy * pn[1] + dz * pn[2]) * ir;
            cr[2 * it + 1] = -(dx * pt[0] + dy * pt[1] + dz * pt[2]) * ir;
        }
        rr[it] = r;
        rq[it] = pq;
        rt[it] = pq + pw;
        cc[it] = cell;
        t = tu - r * kInvC;
    }
    r0 = rr[0];
    r1 = rr[1];
    c0 = cc[0];
    c1 = cc[1];
    rd0 = rq[0];
    rd1q = rq[1];
    rd1t = rt[1];
}

// N samples' contributions of one transmitter in the precision form, for the
// transmitter group G (kGroup*): the fitted range with the per-sample rounding
// corrections (`smooth`, wave-uniform) or the exact solve, s(t_emit), the
// carrier phase in turns from the middle sample's, the hardware sine and
// cosine, the f32 mix, and the exact f64 sum. Slow paths (a tile crossing a
// grid node, a symbol, NAV bit or chip outside the tile's hints, gates) read
// the transmitter's table. x is tu less its value at the middle sample, as
// (t_local - t_local_mid) - (delta_t - delta_t_mid), and gq = -f_lo (t_local -
// t_local_mid), the band's part of the phase. The N samples are independent chains that share the record's reads.
// MOV: the receiver may move; dn, dth are the samples' receiver departures
// (rx_rounding), and obs the sample's receiver position where the caller has
// it (antenna kernels; null otherwise).
template <int G, bool ANT, int N, bool MOV>
__device__ __forceinline__ void precision_link(const Anchor &A, const RxCoef *C, bool smooth,
                                               const TxDev *__restrict__ txs,
                                               const JobDev &J, uint32_t link, const double *tu, const double *x,
                                               const double *del0, const double *gq, const double *dn,
                                               const double *dth, const double *obs, double *sum_re,
                                               double *sum_im) {
    static_assert(!ANT || N == 1, "antenna kernels take one sample per lane");
    double r[N];
    if (smooth) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            double r0 = fma(A.rd0, del0[q], fma(x[q], fma(A.q0c2, x[q], A.q0d1), A.r0m));
            double r1 = fma(x[q], fma(A.q1c2, x[q], A.q1d1), A.r1m);
            if (MOV) {
                r0 = fma(C->n0, dn[q], fma(C->t0, dth[q], r0));
                r1 = fma(C->n1, dn[q], fma(C->t1, dth[q], r1));
            }
            const double t1 = fma(-r0, kInvC, tu[q]);
            const double del1 = ((J.epoch + t1) - J.epoch) - t1;
            r[q] = fma(A.rd1q, del1, fma(-A.k01, del0[q], r1));
        }
    } else {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            if (ANT || !MOV) {
                r[q] = solve_emission_scaled(txs[A.tx], ANT ? obs : J.obs, tu[q], J.epoch);
            } else {
                double o[3];
                rx_position(J, tu[q], o);
                r[q] = solve_emission_scaled(txs[A.tx], o, tu[q], J.epoch);
            }
        }
    }
    double t_emit[N];
#pragma unroll
    for (int q = 0; q < N; ++q) t_emit[q] = fma(-r[q], kInvC, tu[q]);
    // antenna gain; an occluded sample adds nothing (src/channel/mod.rs:253-255)
    float g = 1.0f;
    if (ANT && J.ant_on) {
        g = antenna_gain(J, obs, txs[A.tx], t_emit[0]);
        if (g != 0.0f && J.unocc) J.unocc[J.link0 + link] = 1u;
        if (g == 0.0f) return;
    }
    // s(t_emit): the exact index decides; the tile's hints supply the value
    // when the index is one they hold (read here, before the phase work, so
    // that their latency is covered), and the rare other indices are
    // evaluated after it. CW's own phase joins the carrier phase.
    typedef const __attribute__((address_space(1))) float gchip;
    float bre[N], bim[N], chip[N];
    double ph_cw[N], sig_b[N];
    bool hinted[N];
    // the hint values as one vector read, selected per lane
    float4 hv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (G == kGroupCa || G == kGroupFramed) hv = *reinterpret_cast<const float4 *>(&A.h0re);
#pragma unroll
    for (int q = 0; q < N; ++q) {
        bre[q] = bim[q] = chip[q] = 0.0f;
        ph_cw[q] = sig_b[q] = 0.0;
        hinted[q] = true;
        if (G == kGroupCa) {
            const double dt = t_emit[q] - A.epoch_rel;
            // chips past the tile's base: fl(dt R) - cbase, exact while cbase is at
            // least half the count (Sterbenz), so its floor is floor(dt R) - cbase
            const double d = dt * kCaChipRate - A.cbase;
            // within one code period of the base (NaN fails the test): residue
            // cmod + d of the doubled table, and the NAV bit from the chip count
            // (bit b spans chips [20 460 b, 20 460 (b + 1)); hu holds the offset
            // from cbase where the tile's second bit starts); else the count's
            // own residue and the bit index floor(dt 50)
            int32_t off;
            if (d >= 0.0 && d < static_cast<double>(kCaCodeLength))
                off = static_cast<int32_t>(d);
            else
                off = static_cast<int32_t>(rem_euclid(floor_to_i64(floor(dt * kCaChipRate)), kCaCodeLength)) -
                      static_cast<int32_t>(A.cmod);
            chip[q] = ((gchip *)A.chip0)[off];
            hinted[q] = d >= 0.0 && d < A.hu + kChipsPerNavBit;
            if (!hinted[q]) sig_b[q] = floor(dt * kNavBitRate);
            bre[q] = d >= A.hu ? hv.z : hv.x;
        } else if (G == kGroupFramed) {
            sig_b[q] = floor((t_emit[q] - A.epoch_rel) * A.rate);
            const bool h0 = sig_b[q] == A.hu;
            bre[q] = h0 ? hv.x : hv.z;
            bim[q] = h0 ? hv.y : hv.w;
            hinted[q] = h0 || sig_b[q] == A.hu1;
        } else if (G == kGroupCw) {
            if (A.rate != 0.0)   // rate holds omega (uniform): a tone at the carrier has no phase of its own
                ph_cw[q] = static_cast<double>(static_cast<float>(A.rate * t_emit[q]));
            bre[q] = 1.0f;
        } else {
            const TxDev &T = txs[A.tx];
            if (!T.gated || gate_on(T, t_emit[q])) {
                if (T.signal_kind == 0) {
                    ph_cw[q] = static_cast<double>(cw_phase(T, t_emit[q]));
                    bre[q] = 1.0f;
                } else {
                    baseband(T, t_emit[q], bre[q], bim[q]);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        // the carrier phase in turns, its fraction, and the hardware sine and
        // cosine of that fraction. The phase -f (r / c + delta_t) + (f - f_lo)
        // t_local is linear in range, delta_t and t_local; with the offsets
        // dtl of t_local and ddl of delta_t from the middle sample and
        // x = dtl - ddl, its change is ph_r dr + f x - f_lo dtl (gq = -f_lo dtl)
        const double dr = r[q] - A.r1_mid;
        double turns = fma(A.ph_r, dr, fma(A.ph_x, x[q], A.ph0 + gq[q]));
        if (G == kGroupGeneric || (G == kGroupCw && A.rate != 0.0)) turns = fma(ph_cw[q], kInvTau, turns);
        const float fr = static_cast<float>(__builtin_amdgcn_fract(turns));
        const float spf = __builtin_amdgcn_sinf(fr), cpf = __builtin_amdgcn_cosf(fr);
        // amp_factor / r to first order about the middle sample's range
        float amplitude = fmaf(-A.amp_k, static_cast<float>(dr), A.amp_m);
        if (ANT) amplitude *= g;
        if (G == kGroupCa) {
            if (!hinted[q]) bre[q] = static_cast<float>(nav_pm(txs[A.tx], sig_b[q]));
            bre[q] *= chip[q];
        } else if (G == kGroupFramed) {
            if (!hinted[q]) framed_value(txs[A.tx], sig_b[q], bre[q], bim[q]);
        }
        float cre, cim;
        if (G == kGroupCa || G == kGroupCw) {
            const float ab = amplitude * bre[q];
            cre = ab * cpf;
            cim = ab * spf;
        } else {
            const float ar = amplitude * bre[q], ai = amplitude * bim[q];
            cre = fmaf(ar, cpf, -(ai * spf));
            cim = fmaf(ar, spf, ai * cpf);
        }
        // the f32 contributions summed exactly in f64
        sum_re[q] += static_cast<double>(cre);
        sum_im[q] += static_cast<double>(cim);
    }
}

// The tapped link's emission time, range and carrier phase in the precision
// form: the range as the sample loop forms it, the phase in the reference's
// f64 operation order (src/channel/mod.rs:255-258) from that range.
template <bool MOV>
__device__ __forceinline__ void precision_tap(const Anchor &A, const RxCoef *C, bool smooth,
                                              const TxDev *__restrict__ txs,
                                              const JobDev &J, double tu, double t_local, double dlt, double x,
                                              double del0, double dn, double dth, double &te, double &r,
                                              double &ph) {
    if (smooth) {
        double r0 = fma(A.rd0, del0, fma(x, fma(A.q0c2, x, A.q0d1), A.r0m));
        double r1 = fma(x, fma(A.q1c2, x, A.q1d1), A.r1m);
        if (MOV) {
            r0 = fma(C->n0, dn, fma(C->t0, dth, r0));
            r1 = fma(C->n1, dn, fma(C->t1, dth, r1));
        }
        const double t1 = fma(-r0, kInvC, tu);
        const double del1 = ((J.epoch + t1) - J.epoch) - t1;
        r = fma(A.rd1q, del1, fma(-A.k01, del0, r1));
    } else {
        double o[3] = {J.obs[0], J.obs[1], J.obs[2]};
        if (MOV) rx_position(J, tu, o);
        r = solve_emission_scaled(txs[A.tx], o, tu, J.epoch);
    }
    te = fma(-r, kInvC, tu);
    ph = A.neg_tau_src * (r * kInvC + dlt) + A.tau_if * t_local;
}

// Anchor pass of one (tile, transmitter) record: the smooth solve at anchor a
// (0 first, 1 middle, 2 last valid sample of the tile) of reception time tu,
// the receiver at obs. A moving receiver's derivatives along its angles at
// the middle anchor (pn, pt; null for a fixed receiver) give the record's
// receiver coefficients, the effect on r1 including that of r0 through t1.
__device__ __forceinline__ void anchor_solve(Anchor &A, const TxDev &T, const JobDev &J, double tu, uint32_t a,
                                             const double *obs, const double *pn = nullptr,
                                             const double *pt = nullptr, RxCoef *C = nullptr) {
    const double eoff = T.traj_kind ? J.epoch - T.grid_t0 : 0.0;
    double r0, r1, rd0, rd1q, rd1t, cr[4];
    int64_t c0, c1;
    if (pn && a == 1) solve_smooth(T, obs, tu, eoff, r0, r1, c0, c1, rd0, rd1q, rd1t, pn, pt, cr);
    else solve_smooth(T, obs, tu, eoff, r0, r1, c0, c1, rd0, rd1q, rd1t);
    A.r0s[a] = r0;
    A.r1s[a] = r1;
    if (a != 1) {
        A.cell[a >> 1][0] = c0;
        A.cell[a >> 1][1] = c1;
    } else {
        A.rd0 = rd0;
        A.rd1q = rd1q;
        A.k01 = rd1t * rd0 * kInvC;
        if (pn) {
            C->n0 = cr[0];
            C->t0 = cr[1];
            C->n1 = cr[2] - rd1t * cr[0] * kInvC;
            C->t1 = cr[3] - rd1t * cr[1] * kInvC;
        }
    }
}

// The fit of one (tile, transmitter) record from its anchor pass; at[a] holds
// (tu, t_local, delta_t) of anchor a. Returns whether the tile is smooth for
// the transmitter (the lanes then use the fitted range, else the exact solve).
// MOV: the receiver may move (C then holds the record's coefficients).
template <bool MOV>
__device__ __forceinline__ bool anchor_fit(Anchor &A, const RxCoef *C, const LinkDev &L, const TxDev &T,
                                           const JobDev &J, const double (*at)[3], bool fit_ok) {
    const bool mov = MOV && J.rx_kind != 0;
    // the anchor pass's results share storage with fields written below
    const double r0s[3] = {A.r0s[0], A.r0s[1], A.r0s[2]};
    const double r1s[3] = {A.r1s[0], A.r1s[1], A.r1s[2]};
    const bool same = A.cell[0][0] == A.cell[1][0] && A.cell[0][1] == A.cell[1][1];
    // The quadratic's truncation over the tile, |r'''| h^3 / (9 sqrt 3) with
    // |r'''| <= j + 6 v a / r + 3 v^3 / r^2 for bounds v, a, j of the relative
    // motion (the transmitter's grid is linear within a cell; the receiver's
    // bounds come with the job), held to 1e-7 rad of carrier phase; and the
    // amplitude's first-order model in the range, (v h / r)^2 <= 1e-7. A tile
    // outside either bound (long tiles of low-rate bands seen from a moving
    // receiver, close passes) takes the exact per-sample solve; NaN bounds fail.
    const double h = 0.5 * (at[2][0] - at[0][0]);
    const double v = T.max_speed + J.rx_v;
    const double rmin = fmin(fmin(r0s[0], r0s[1]), fmin(r0s[2], fmin(r1s[0], fmin(r1s[1], r1s[2])))) - v * h;
    const double r3 = J.rx_j + (6.0 * v * J.rx_a + 3.0 * v * v * v / rmin) / rmin;
    const bool bounded = rmin > 0.0 && v * h <= 3.2e-4 * rmin &&
                         fabs(T.neg_tau_src) * r3 * (h * h * h) * 0.06415002990995841 <= 1e-7 * kSpeedOfLight;
    const bool smooth = fit_ok && (T.traj_kind == 0 || same) && bounded;
    double r1m;
    A.r0m = r0s[1];
    A.r1m = r1s[1];
    if (smooth) {
        const double xa = at[0][0] - at[1][0], xb = at[2][0] - at[1][0];
        const double a0 = (r0s[1] - r0s[0]) / (-xa), b0 = (r0s[2] - r0s[1]) / xb;
        const double a1 = (r1s[1] - r1s[0]) / (-xa), b1 = (r1s[2] - r1s[1]) / xb;
        // Newton form r_m + x (d1 + c2 (x - xa)), with the -c2 xa
        // term folded into the linear coefficient
        A.q0c2 = (b0 - a0) / (xb - xa);
        A.q0d1 = fma(-A.q0c2, xa, a0);
        A.q1c2 = (b1 - a1) / (xb - xa);
        A.q1d1 = fma(-A.q1c2, xa, a1);
        // the model at the middle sample (x = 0)
        const double t0m = at[1][0] - kEmissionInitialLag;
        const double d0m = ((J.epoch + t0m) - J.epoch) - t0m;
        double r0m = fma(A.rd0, d0m, r0s[1]), r1c = r1s[1];
        if (mov) {
            double dn, dth;
            rx_rounding(J, at[1][0], dn, dth);
            r0m = fma(C->n0, dn, fma(C->t0, dth, r0m));
            r1c = fma(C->n1, dn, fma(C->t1, dth, r1c));
        }
        const double t1m = fma(-r0m, kInvC, at[1][0]);
        const double d1m = ((J.epoch + t1m) - J.epoch) - t1m;
        r1m = fma(A.rd1q, d1m, fma(-A.k01, d0m, r1c));
    } else {
        // the exact solve at the middle sample, as the lanes solve
        double o[3] = {J.obs[0], J.obs[1], J.obs[2]};
        if (mov) rx_position(J, at[1][0], o);
        r1m = solve_emission_scaled(T, o, at[1][0], J.epoch);
    }
    A.r1_mid = r1m;
    const float amp_m = static_cast<float>(T.amp_factor / r1m);
    A.amp_m = amp_m;
    A.amp_k = amp_m * static_cast<float>(1.0 / r1m);
    A.neg_tau_src = T.neg_tau_src;
    A.tau_if = L.tau_if;
    A.ph0 = carrier_turns(T.neg_tau_src, L.tau_if, r1m, at[1][2], at[1][1]);
    A.ph_r = T.neg_tau_src * kInvTau * kInvC;
    A.ph_x = -T.neg_tau_src * kInvTau;
    A.ph_spare = 0.0;
    A.epoch_rel = T.epoch_rel;
    A.rate = T.signal_kind == 0 ? T.omega : T.rate;
    A.chip0 = T.ca_chips2;
    A.tx = L.tx;
    // Values of the symbol (or NAV bit) the tile starts in and of the next
    // one; lanes take them when their own exact index matches and evaluate
    // the signal otherwise. For C/A, a chip count one before the first
    // sample's and its residue.
    const double te0 = at[0][0] - r1s[0] * kInvC;
    A.hu = 0.0;
    A.hu1 = 1.0;
    A.cbase = 0.0;
    A.cmod = 0u;
    A.h0re = A.h0im = A.h1re = A.h1im = 0.0f;
    if (T.signal_kind >= 1 && T.signal_kind <= 3) {
        A.hu = floor((te0 - T.epoch_rel) * T.rate);
        A.hu1 = A.hu + 1.0;
        framed_value(T, A.hu, A.h0re, A.h0im);
        framed_value(T, A.hu1, A.h1re, A.h1im);
    } else if (T.signal_kind == 4) {
        // chip counts and NAV bits (src/signal/ca_code.rs:167-190): 1.023e6 / 50
        // = 20 460 chips per bit, so bit b spans chip counts [20 460 b, 20 460 (b + 1))
        const double cb = floor((te0 - T.epoch_rel) * kCaChipRate) - 1.0;
        if (fabs(cb) < 4503599627370496.0) {   // 2^52: counts stay exact
            A.cbase = cb;
            A.cmod = static_cast<uint32_t>(rem_euclid(floor_to_i64(cb), kCaCodeLength));
            A.chip0 = T.ca_chips2 + A.cmod;
            const double b_lo = floor(cb / kChipsPerNavBit);
            A.hu = fma(b_lo + 1.0, kChipsPerNavBit, -cb);   // in [1, 20 460]: exact
            A.hu1 = b_lo;
            A.h0re = static_cast<float>(nav_pm(T, b_lo));
            A.h1re = static_cast<float>(nav_pm(T, b_lo + 1.0));
        } else {
            A.cbase = __builtin_nan("");      // every lane takes the general index
            A.hu = 0.0;
        }
    }
    return smooth;
}

// N samples' contributions from the links [g0, g0 + ng) of their job, whose
// records are R[0..ng): the transmitter groups in order, links whose bit is
// set in gmask skipped (antenna kernels), the fitted range for those whose
// bit is set in smask.
template <bool ANT, int N, bool MOV>
__device__ __forceinline__ void precision_links(const Anchor *R, const RxCoef *RC, uint64_t smask, uint64_t gmask,
                                                const TxDev *__restrict__ txs, const JobDev &J, uint32_t g0,
                                                uint32_t ng, const double *tu, const double *x, const double *del0,
                                                const double *gq, const double *dn, const double *dth,
                                                const double *obs, double *sum_re, double *sum_im) {
    const uint32_t e_fr = J.group_end[kGroupFramed] > g0 ? min(J.group_end[kGroupFramed] - g0, ng) : 0u;
    const uint32_t e_cw = J.group_end[kGroupCw] > g0 ? min(J.group_end[kGroupCw] - g0, ng) : 0u;
    const uint32_t e_gn = J.group_end[kGroupGeneric] > g0 ? min(J.group_end[kGroupGeneric] - g0, ng) : 0u;
    uint32_t s = 0;
    for (; s < e_fr; ++s)
        if (!ANT || !((gmask >> s) & 1u))
            precision_link<kGroupFramed, ANT, N, MOV>(R[s], MOV ? RC + s : nullptr, (smask >> s) & 1u, txs, J, g0 + s, tu, x, del0, gq, dn, dth,
                                                 obs, sum_re, sum_im);
    for (; s < e_cw; ++s)
        if (!ANT || !((gmask >> s) & 1u))
            precision_link<kGroupCw, ANT, N, MOV>(R[s], MOV ? RC + s : nullptr, (smask >> s) & 1u, txs, J, g0 + s, tu, x, del0, gq, dn, dth, obs, sum_re,
                                             sum_im);
    for (; s < e_gn; ++s)
        if (!ANT || !((gmask >> s) & 1u))
            precision_link<kGroupGeneric, ANT, N, MOV>(R[s], MOV ? RC + s : nullptr, (smask >> s) & 1u, txs, J, g0 + s, tu, x, del0, gq, dn, dth,
                                                  obs, sum_re, sum_im);
    for (; s < ng; ++s)
        if (!ANT || !((gmask >> s) & 1u))
            precision_link<kGroupCa, ANT, N, MOV>(R[s], MOV ? RC + s : nullptr, (smask >> s) & 1u, txs, J, g0 + s, tu, x, del0, gq, dn, dth, obs, sum_re,
                                             sum_im);
}

// RxChain::apply_post_sum (src/rx/mod.rs:142-146 and :83-101) and the
// stores of output sample j.
__device__ __forceinline__ void finish_sample(const JobDev &J, float acc_re, float acc_im, uint64_t j) {
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

// The precision form without an antenna over one span of 2^kSuperShift
// tiles of 2^TS samples (`span` counts spans from the job's first), with
// 2^(TS-1) lanes, each owning two adjacent samples of a tile. The anchor
// solves and fits of several tiles run as one batch (as many (tile,
// transmitter) records as kSpanRecords holds), so the latency of their
// dependent global reads is paid once per batch; the tiles of the batch then
// run their sample loops in turn, each lane's two samples as independent
// chains that share every record read. Per tile the arithmetic is that of a
// one-tile workgroup. MOV: the receiver may move.
template <int TS, bool MOV>
__device__ __forceinline__ void precision_span(const JobDev &J, uint32_t span, const LinkDev *__restrict__ links,
                                               const TxDev *__restrict__ txs, const TapOut &tap) {
    constexpr int kTile = 1 << TS, kLanes = kTile / kSpanSamples, kSub = 1 << kSuperShift;
    constexpr int kWaves = kLanes / 64;
    static_assert(kWaves >= 1, "at least one wavefront per span");
    __shared__ Anchor s_rec[kSpanRecords];        // record kk * ng + s: tile kk of the batch, link g0 + s
    __shared__ double s_at[kSub][3][3];           // (tu, t_local, delta_t) at each tile's anchors
    __shared__ double s_rx[MOV ? kSub : 1][3][3];  // moving receiver: its smooth position at the anchors
    __shared__ double s_rxp[MOV ? kSub : 1][2][3]; // and its derivatives along nu and theta at the middle one
    __shared__ RxCoef s_rxc[MOV ? kSpanRecords : 1]; // and each record's coefficients
    __shared__ uint32_t s_mask[kSub][2];          // smooth links of each tile of the batch
    __shared__ uint8_t s_smooth[kSpanRecords];     // per record: the tile is smooth for the link
    const uint64_t span0 = ((J.i0 >> (TS + kSuperShift)) + span) << (TS + kSuperShift);
    const uint64_t jend = J.i0 + J.clen;
    const uint32_t na = J.n_active;
    const int kb_max = na == 0 ? kSub : (na >= uint32_t(kAnchorGroup) ? 1 : min(kSub, kSpanRecords / int(na)));
    const bool noisy = J.sigma > 0.0f;
    // a band sampled at its receiver's master rate reads one walk value per sample
    const bool at_master = J.walk && J.walk_rel && !J.delta_t && J.fs == J.master;
    for (int k0 = 0; k0 < kSub; k0 += kb_max) {
        const int kb = min(kb_max, kSub - k0);
        double sum_re[kSpanSamples] = {}, sum_im[kSpanSamples] = {};
        for (uint32_t g0 = 0; g0 == 0 || g0 < na; g0 += kAnchorGroup) {
            const uint32_t ng = na - g0 < uint32_t(kAnchorGroup) ? na - g0 : uint32_t(kAnchorGroup);
            const bool last_group = g0 + ng >= na;
            // the anchor samples of each tile of the batch (src/pipeline.rs:918-926)
            if (threadIdx.x < 3 * kb) {
                const int kk = threadIdx.x / 3, a = threadIdx.x - 3 * kk;
                const uint64_t t0 = span0 + uint64_t(k0 + kk) * kTile;
                const uint64_t first = J.i0 > t0 ? J.i0 : t0, end = jend < t0 + kTile ? jend : t0 + kTile;
                if (first < end) {
                    const uint64_t ia = a == 0 ? first : (a == 1 ? (first + end - 1) >> 1 : end - 1);
                    const double tl = sample_time(static_cast<double>(ia), J.fs, J.inv_fs);
                    const double dl = J.delta_t ? J.delta_t[ia - J.i0] : clock_delta(J, tl);
                    s_at[kk][a][0] = tl - dl;
                    s_at[kk][a][1] = tl;
                    s_at[kk][a][2] = dl;
                    if (MOV) {
                        double pn[3] = {0.0, 0.0, 0.0}, pt[3] = {0.0, 0.0, 0.0}, o[3] = {J.obs[0], J.obs[1], J.obs[2]};
                        if (J.rx_kind) rx_smooth(J, tl - dl, o, pn, pt);
                        for (int k = 0; k < 3; ++k) {
                            s_rx[kk][a][k] = o[k];
                            if (a == 1) {
                                s_rxp[kk][0][k] = pn[k];
                                s_rxp[kk][1][k] = pt[k];
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // smooth solves at the three anchors of every record
            for (uint32_t q = threadIdx.x; q < 3 * kb * ng; q += kLanes) {
                const uint32_t rec = q / 3, a = q - 3 * rec, kk = rec / ng;
                const uint64_t t0 = span0 + uint64_t(k0 + kk) * kTile;
                if ((J.i0 > t0 ? J.i0 : t0) >= (jend < t0 + kTile ? jend : t0 + kTile)) continue;
                const LinkDev L = links[J.link0 + g0 + (rec - kk * ng)];
                if (MOV)
                    anchor_solve(s_rec[rec], txs[L.tx], J, s_at[kk][a][0], a, s_rx[kk][a], s_rxp[kk][0], s_rxp[kk][1],
                                 &s_rxc[rec]);
                else anchor_solve(s_rec[rec], txs[L.tx], J, s_at[kk][a][0], a, J.obs);
            }
            __syncthreads();
            // the fits, one record a lane
            for (uint32_t rec = threadIdx.x; rec < kb * ng; rec += kLanes) {
                const uint32_t kk = rec / ng;
                const uint64_t t0 = span0 + uint64_t(k0 + kk) * kTile;
                const uint64_t first = J.i0 > t0 ? J.i0 : t0, end = jend < t0 + kTile ? jend : t0 + kTile;
                bool smooth = false;
                if (first < end) {
                    const LinkDev L = links[J.link0 + g0 + (rec - kk * ng)];
                    smooth = anchor_fit<MOV>(s_rec[rec], MOV ? &s_rxc[rec] : nullptr, L, txs[L.tx], J, s_at[kk],
                                             end - 1 - first >= 2);
                }
                s_smooth[rec] = smooth ? 1u : 0u;
            }
            __syncthreads();
            // each tile's smooth links as a mask: wave w takes tiles w, w + kWaves, ...
            for (int kr = 0; kr < kb; kr += kWaves) {
                const int kk = kr + static_cast<int>(threadIdx.x >> 6);
                const uint32_t s = threadIdx.x & 63u;
                const uint64_t m = __ballot(kk < kb && s < ng && s_smooth[kk * ng + s]);
                if (kk < kb && s == 0) {
                    s_mask[kk][0] = static_cast<uint32_t>(m);
                    s_mask[kk][1] = static_cast<uint32_t>(m >> 32);
                }
            }
            __syncthreads();
            // at the master rate, the walk values of the next tile's samples are read
            // during this tile's link loop (the same three terms walk_eps_rel sums)
            float pf_loc[kSpanSamples] = {0.0f, 0.0f};
            double pf_toff = 0.0, pf_base = 0.0;
            bool pf_same = false;   // the lane's two samples share a 32-index sub-block
            int pf_kk = -1;         // the tile the values are for
            for (int kk = 0; kk < kb; ++kk) {
                const uint64_t t0 = span0 + uint64_t(k0 + kk) * kTile;
                const uint64_t first = J.i0 > t0 ? J.i0 : t0, end = jend < t0 + kTile ? jend : t0 + kTile;
                if (first >= end) continue;   // a tile with no sample of its job (uniform)
                const bool use_pf = pf_kk == kk && pf_same;
                // the lane's two samples (src/pipeline.rs:918-926), clamped into the job
                uint64_t i[kSpanSamples];
                bool valid[kSpanSamples];
                double t_local[kSpanSamples], dlt[kSpanSamples], tu[kSpanSamples], x[kSpanSamples];
                double del0[kSpanSamples], gq[kSpanSamples], dn[kSpanSamples], dth[kSpanSamples];
#pragma unroll
                for (int q = 0; q < kSpanSamples; ++q) {
                    i[q] = t0 + kSpanSamples * threadIdx.x + q;
                    valid[q] = i[q] >= J.i0 && i[q] < jend;
                    const uint64_t ic = valid[q] ? i[q] : first;
                    t_local[q] = sample_time(static_cast<double>(ic), J.fs, J.inv_fs);
                    if (J.delta_t) dlt[q] = J.delta_t[ic - J.i0];
                    else if (at_master && use_pf)
                        dlt[q] = clock_at_master_eps(J, t_local[q], (pf_base + pf_toff) + static_cast<double>(pf_loc[q]));
                    else dlt[q] = at_master ? clock_delta_at_master(J, t_local[q], ic) : clock_delta(J, t_local[q]);
                    tu[q] = t_local[q] - dlt[q];
                    // tu less its value at the middle sample, formed as the difference of
                    // the exact offsets of t_local and delta_t (the phase must not carry the
                    // rounding of tu itself; the fitted range does not notice it)
                    x[q] = (t_local[q] - s_at[kk][1][1]) - (dlt[q] - s_at[kk][1][2]);
                    // rounding of the iteration-1 query time, exact (Sterbenz) per sample
                    const double t0q = tu[q] - kEmissionInitialLag;
                    del0[q] = ((J.epoch + t0q) - J.epoch) - t0q;
                    // the band's part of the carrier phase from the middle sample's
                    gq[q] = -J.lo * (t_local[q] - s_at[kk][1][1]);
                    dn[q] = dth[q] = 0.0;
                    if (MOV) rx_rounding(J, tu[q], dn[q], dth[q]);
                }
                pf_kk = -1;
                // (not in the moving-receiver kernels, which are at their register limit)
                if (!MOV && !J.delta_t && at_master && kk + 1 < kb) {
                    const uint64_t t1 = t0 + kTile;
                    const uint64_t first1 = J.i0 > t1 ? J.i0 : t1, end1 = jend < t1 + kTile ? jend : t1 + kTile;
                    if (first1 < end1) {
                        int32_t rr[kSpanSamples];
#pragma unroll
                        for (int q = 0; q < kSpanSamples; ++q) {
                            const uint64_t in = t1 + kSpanSamples * threadIdx.x + q;
                            const uint64_t icn = (in >= J.i0 && in < jend) ? in : first1;
                            rr[q] = static_cast<int32_t>(static_cast<int64_t>(icn) - J.eps_k0);
                            pf_loc[q] = J.walk_local[rr[q]];
                        }
                        pf_same = (rr[0] >> kWalkSubShift) == (rr[1] >> kWalkSubShift);
                        pf_base = J.walk_base_w[rr[0] >> kWalkBlockShift];
                        pf_toff = J.walk_toff[rr[0] >> kWalkSubShift];
                        pf_kk = kk + 1;
                    }
                }
                const uint64_t smask = uniform_u64(s_mask[kk][0], s_mask[kk][1]);
                const Anchor *R = s_rec + kk * ng;
                const RxCoef *RC = MOV ? s_rxc + kk * ng : nullptr;
                precision_links<false, kSpanSamples, MOV>(R, RC, smask, 0ull, txs, J, g0, ng, tu, x, del0, gq, dn, dth,
                                                          nullptr, sum_re, sum_im);
                if (tap.on) {
                    // the sample times and clock offsets again (as above), for the taps
#pragma unroll
                    for (int q = 0; q < kSpanSamples; ++q) {
                        if (!valid[q]) continue;
                        const double tl = sample_time(static_cast<double>(i[q]), J.fs, J.inv_fs);
                        const double dl = J.delta_t ? J.delta_t[i[q] - J.i0]
                                                    : (at_master ? clock_delta_at_master(J, tl, i[q])
                                                                 : clock_delta(J, tl));
                        if (last_group) tap.delta_t[i[q] - J.i0] = dl;
                        if (tap.slot >= g0 && tap.slot < g0 + ng) {
                            const uint32_t ts = tap.slot - g0;
                            double te, r, ph;
                            precision_tap<MOV>(R[ts], MOV ? RC + ts : nullptr, (smask >> ts) & 1u, txs, J, tu[q], tl, dl, x[q], del0[q], dn[q],
                                               dth[q], te, r, ph);
                            tap.t_emit[i[q] - J.i0] = te;
                            tap.range[i[q] - J.i0] = r;
                            tap.phase[i[q] - J.i0] = ph;
                        }
                    }
                }
                if (last_group && (J.pass == 1 || J.pass == 2)) {
                    // a pass before the last: the f64 sums so far go to HBM
#pragma unroll
                    for (int q = 0; q < kSpanSamples; ++q) {
                        if (valid[q]) {
                            double2 *pp = J.part + (i[q] - J.i0);
                            if (J.pass == 2) {
                                const double2 prev = *pp;
                                sum_re[q] += prev.x;
                                sum_im[q] += prev.y;
                            }
                            *pp = make_double2(sum_re[q], sum_im[q]);
                        }
                        sum_re[q] = 0.0;
                        sum_im[q] = 0.0;
                    }
                } else if (last_group) {
                    if (J.pass == 3) {
                        // the earlier passes' sums
#pragma unroll
                        for (int q = 0; q < kSpanSamples; ++q) {
                            if (valid[q]) {
                                const double2 prev = J.part[i[q] - J.i0];
                                sum_re[q] += prev.x;
                                sum_im[q] += prev.y;
                            }
                        }
                    }
                    // AWGN, src/channel/mod.rs:328-333 (sigma = sqrt(P/2) as f32): the
                    // lane's samples are words 0 and 1 of Philox block i >> 1
                    uint64_t w[2] = {0, 0};
                    if (noisy) philox_words((t0 >> 1) + threadIdx.x, J.noise_seed, kAwgnTag, w);
#pragma unroll
                    for (int q = 0; q < kSpanSamples; ++q) {
                        float acc_re = static_cast<float>(sum_re[q]), acc_im = static_cast<float>(sum_im[q]);
                        if (noisy) {
                            float zr, zi;
                            box_muller(w[q], zr, zi);
                            acc_re += J.sigma * zr;
                            acc_im += J.sigma * zi;
                        }
                        if (valid[q]) finish_sample(J, acc_re, acc_im, i[q] - J.i0);
                        sum_re[q] = 0.0;
                        sum_im[q] = 0.0;
                    }
                }
            }
            __syncthreads();   // the next batch or link group rewrites the records and anchors
        }
    }
}

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// run of tiles so neighbouring tiles of a job share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, per = n >> 3, rem = n & 7u;
    return x * per + (x < rem ? x : rem) + (b >> 3);
}

// Launch-wide choices of a render dispatch.
struct RenderLaunch {
    uint32_t tile_base, n_tiles;  // this dispatch's slice of the call's tiles
    const uint32_t *n_live;       // if set: tiles from *n_live on have no work (streaming plans)
    uint32_t xcd_map;             // deal contiguous tile runs to each XCD
    const double *pred;           // converged form: predicted light times (JobDev::pred0), or null
    unsigned long long *stats;    // converged form: evaluations, contributions, fallbacks, unconverged
    uint32_t gate;                // antenna kernels: apply the horizon gate
    unsigned long long *gate_stats;  // (tile, link) pairs the gate skipped, and pairs it tested
};

template <int FORM, bool ANT, int TS, bool MOV>
__global__ void __launch_bounds__(1 << TS, ANT ? 4 : (FORM == 1 ? (MOV ? 3 : 4) : 6))
    render_kernel(const JobDev *__restrict__ jobs, int n_jobs, const LinkDev *__restrict__ links,
                  const TxDev *__restrict__ txs, RenderLaunch lx, TapOut tap) {
    constexpr int kTile = 1 << TS;
    constexpr bool kSpan = FORM == 1 && !ANT;          // the super-tile path has LDS of its own
    static_assert(!MOV || kSpan, "only the precision form's span kernels take a moving receiver");
    __shared__ uint64_t s_noise[kSpan ? 1 : kTile];    // one Philox word per sample
    __shared__ double s_anc_t[3][3];                   // tu, t_local, delta_t at the anchors
    __shared__ Anchor s_anc[FORM == 1 && ANT ? kAnchorGroup : 1];

    const uint32_t tile = lx.tile_base + (lx.xcd_map ? xcd_tile(blockIdx.x, lx.n_tiles) : blockIdx.x);
    if (lx.n_live && tile >= *lx.n_live) return;
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    const JobDev &J = jobs[lo];
    if constexpr (kSpan) {
        precision_span<TS, MOV>(J, tile - J.tile0, links, txs, tap);
        return;
    }
    const uint64_t tile_start = ((J.i0 >> TS) + (tile - J.tile0)) << TS;
    const uint64_t i = tile_start + threadIdx.x;
    const bool valid = i >= J.i0 && i < J.i0 + J.clen;
    const uint64_t j = i - J.i0;

    // AWGN draws: lanes below kTile/2 compute the tile's Philox blocks
    // (sample i: block i >> 1, word i & 1) into LDS; consumed after the
    // transmitter loop.
    const bool noisy = J.sigma > 0.0f;
    if (noisy && threadIdx.x < kTile / 2) {
        uint64_t w[2];
        philox_words((tile_start >> 1) + threadIdx.x, J.noise_seed, kAwgnTag, w);
        s_noise[2 * threadIdx.x] = w[0];
        s_noise[2 * threadIdx.x + 1] = w[1];
    }

    // src/pipeline.rs:918-926 (clamped to a valid sample on the tile's edge lanes)
    const uint64_t first = J.i0 > tile_start ? J.i0 : tile_start;
    const uint64_t end = (J.i0 + J.clen) < tile_start + kTile ? (J.i0 + J.clen) : tile_start + kTile;
    if (first >= end) return;  // a tile with no sample of its job (uniform over the workgroup)
    const uint64_t ic = valid ? i : first;
    const double t_local = static_cast<double>(ic) / J.fs;
    const double dlt = J.delta_t ? J.delta_t[ic - J.i0] : clock_delta(J, t_local);
    const double tu = t_local - dlt;
    // the receiver at this sample's reception time (src/pipeline.rs:929-930)
    double obs[3];
    rx_position(J, tu, obs);

    // Horizon gate (antenna kernels), src/channel/horizon.rs:97-145 over the
    // tile: the span of its reception times and a ball holding every receiver
    // position (centre: the first sample's; radius: the largest distance of
    // any sample's from it; 0 for a fixed receiver); a non-finite time or
    // position anywhere turns the gate off for the tile.
    const bool gating = ANT && lx.gate && !tap.on && J.ant_on && J.ant_cut;
    double g_tmin = 0.0, g_tmax = 0.0, g_orad = 0.0;
    double g_oc[3] = {J.obs[0], J.obs[1], J.obs[2]};
    bool g_ok = false;
    if (ANT && gating) {
        __shared__ double s_tspan[3 * (kTile / 64)];
        __shared__ uint32_t s_tbad[kTile / 64];
        __shared__ double s_ocen[3];
        const bool moving = J.rx_kind != 0;   // uniform
        if (moving) {
            if (i == first) {
                s_ocen[0] = obs[0];
                s_ocen[1] = obs[1];
                s_ocen[2] = obs[2];
            }
            __syncthreads();
            g_oc[0] = s_ocen[0];
            g_oc[1] = s_ocen[1];
            g_oc[2] = s_ocen[2];
        }
        const double d0 = obs[0] - g_oc[0], d1 = obs[1] - g_oc[1], d2 = obs[2] - g_oc[2];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        double mn = tu, mx = tu, mr = r2;
        for (int o = 32; o >= 1; o >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, o));
            mx = fmax(mx, __shfl_xor(mx, o));
            mr = fmax(mr, __shfl_xor(mr, o));
        }
        const bool bad = __ballot(!isfinite(tu) || !isfinite(r2)) != 0;
        if ((threadIdx.x & 63) == 0) {
            s_tspan[3 * (threadIdx.x >> 6)] = mn;
            s_tspan[3 * (threadIdx.x >> 6) + 1] = mx;
            s_tspan[3 * (threadIdx.x >> 6) + 2] = mr;
            s_tbad[threadIdx.x >> 6] = bad ? 1u : 0u;
        }
        __syncthreads();
        g_tmin = s_tspan[0];
        g_tmax = s_tspan[1];
        double r2max = s_tspan[2];
        g_ok = s_tbad[0] == 0;
        for (int w = 1; w < kTile / 64; ++w) {
            g_tmin = fmin(g_tmin, s_tspan[3 * w]);
            g_tmax = fmax(g_tmax, s_tspan[3 * w + 1]);
            r2max = fmax(r2max, s_tspan[3 * w + 2]);
            g_ok = g_ok && s_tbad[w] == 0;
        }
        g_orad = sqrt(r2max);
    }

    float acc_re = 0.0f, acc_im = 0.0f;
    double sum_re = 0.0, sum_im = 0.0;  // exact sum of the f32 contributions (precision form)

    if (FORM != 1) {
        // FORM 2: where this lane sits between the tile's two predicted light
        // times (boundary samples clamped into the job, as the predictor
        // places them)
        const double *prow = nullptr;
        double frac = 0.0;
        // solve counters (converged forms). Every solve costs two evaluations
        // unless it ends elsewhere, so the wave counts the links it solved
        // (uniform) and each lane only the evaluations beyond two; fallbacks
        // (more than kLearnedEvals from a learned start) and unconverged solves
        // are counted per wave from ballots over the valid lanes
        uint32_t n_solved = 0, n_fb = 0, n_unc = 0;
        int32_t n_extra = 0;
        // the links the gate skips for this tile, as a bitmap
        __shared__ uint32_t s_gated[ANT ? kGateMaxLinks / 32 : 1];
        if (ANT && gating) {
            for (uint32_t w = threadIdx.x; w < kGateMaxLinks / 32; w += kTile) s_gated[w] = 0u;
            __syncthreads();
            const uint32_t ng = J.n_active < kGateMaxLinks ? J.n_active : kGateMaxLinks;
            if (g_ok)
                for (uint32_t s = threadIdx.x; s < ng; s += kTile)
                    if (gate_below_cutoff(J, txs[links[J.link0 + s].tx], g_tmin, g_tmax, g_oc, g_orad))
                        atomicOr(&s_gated[s >> 5], 1u << (s & 31));
            __syncthreads();
            if (threadIdx.x == 0 && lx.gate_stats) {
                uint32_t n = 0;
                for (uint32_t w = 0; w < (ng + 31) / 32; ++w) n += __popc(s_gated[w]);
                if (n) atomicAdd(&lx.gate_stats[0], static_cast<unsigned long long>(n));
                atomicAdd(&lx.gate_stats[1], static_cast<unsigned long long>(ng));
            }
        }
        __shared__ double2 s_pred[FORM == 2 ? kPredStageLinks : 1];
        if (FORM == 2 && lx.pred) {
            const uint32_t bq = (tile - J.tile0) / J.pred_qs;   // the query boundary at or before this tile
            const uint64_t last = J.i0 + J.clen - 1;
            const uint64_t base = J.i0 >> TS;
            uint64_t ib0 = (base + uint64_t(bq) * J.pred_qs) << TS;
            uint64_t ib1 = (base + uint64_t(bq + 1) * J.pred_qs) << TS;
            ib0 = ib0 < J.i0 ? J.i0 : ib0;
            ib1 = ib1 > last ? last : ib1;
            frac = ib1 > ib0 ? static_cast<double>(ic - ib0) / static_cast<double>(ib1 - ib0) : 0.0;
            prow = lx.pred + J.pred0 + bq;
        }
        // the converged forms take the links in blocks of kPredStageLinks and
        // stage each block's start as (lag at the tile's first query, its
        // change to the next), so that no solve waits on a global load: the
        // learned form's two predictions, or (0.07 s, 0) for the reference's
        // start; the plain form takes every link in one pass
        for (uint32_t s0 = 0, s1; s0 < J.n_active; s0 = s1) {
            const uint32_t blk = static_cast<uint32_t>(kPredStageLinks);
            s1 = FORM != 2 || J.n_active - s0 <= blk ? J.n_active : s0 + blk;
            if (FORM == 2) {
                if (s0) __syncthreads();   // every wave is done with the previous block
                for (uint32_t s = s0 + threadIdx.x; s < s1; s += kTile) {
                    double ta = kEmissionInitialLag, dab = 0.0;
                    if (prow) {
                        const double *pp = prow + static_cast<uint64_t>(s) * J.pred_nb;
                        ta = pp[0];
                        dab = pp[1] - ta;
                    }
                    s_pred[s - s0] = make_double2(ta, dab);
                }
                __syncthreads();
            }
            for (uint32_t s = s0; s < s1; ++s) {
                if (ANT && gating && s < kGateMaxLinks && ((s_gated[s >> 5] >> (s & 31)) & 1u)) continue;
                const LinkDev L = links[J.link0 + s];
                const TxDev &T = txs[L.tx];
                double t_emit, r;
                if (FORM == 0) {
                    int64_t c0, c1;
                    solve_emission(T, obs, tu, J.epoch, t_emit, r, c0, c1);
                } else {
                    const double eoff = T.traj_kind ? J.epoch - T.grid_t0 : 0.0;
                    // the learned start (the converged light times of the tile's
                    // two queries, interpolated) or the reference's, t - 0.07 s;
                    // then the iteration to the tolerance from there: one
                    // evaluation (enough from a learned start), the rest only
                    // in waves where some lane needs more
                    const double2 p = s_pred[s - s0];
                    const double t0 = tu - (p.x + p.y * frac);
                    t_emit = emission_step(T, obs, tu, eoff, t0, r);
                    const bool at_one = fabs(t_emit - t0) < kSolveTol;
                    if (__any(!at_one)) {
                        int e = 1;
                        if (!at_one) e = solve_converged_tail(T, obs, tu, eoff, kColdEvals, t_emit, r);
                        if (lx.stats) {
                            const int ev = e < 0 ? kColdEvals : e;
                            n_extra += valid ? ev - 1 : 0;
                            n_fb += prow ? __popcll(__ballot(valid && ev > kLearnedEvals)) : 0u;
                            n_unc += __popcll(__ballot(valid && e < 0));
                        }
                    }
                    ++n_solved;
                }
                // src/channel/mod.rs:255-267
                const double delay = r * kInvC;
                const double phase = T.neg_tau_src * (delay + dlt) + L.tau_if * t_local;
                if (tap.on && s == tap.slot && valid) {
                    tap.t_emit[j] = t_emit;
                    tap.range[j] = r;
                    tap.phase[j] = phase;
                }
                // antenna gain; an occluded sample adds nothing (src/channel/mod.rs:253-255)
                float g = 1.0f;
                if (ANT && J.ant_on) {
                    g = antenna_gain(J, obs, T, t_emit);
                    if (g != 0.0f && J.unocc) J.unocc[J.link0 + s] = 1u;
                    if (g == 0.0f) continue;
                }
                float bre, bim;
                baseband(T, t_emit, bre, bim);
                double sp, cp;
                sincos(phase, &sp, &cp);
                const float cpf = static_cast<float>(cp), spf = static_cast<float>(sp);
                const float amplitude = static_cast<float>(T.amp_factor / r) * g;
                acc_re = acc_re + amplitude * (bre * cpf - bim * spf);
                acc_im = acc_im + amplitude * (bre * spf + bim * cpf);
            }
        }
        if (FORM == 2 && lx.stats) {
            // every wave solved the same links: thread 0 counts the tile's
            // contributions and one evaluation for each; waves add what
            // differs from that, when anything does (evaluations beyond one,
            // signed, as a wrapping add), and their fallbacks and unconverged
            // solves
            if (threadIdx.x == 0) {
                const uint64_t lo = tile_start > J.i0 ? tile_start : J.i0;
                const uint64_t hi = tile_start + kTile < J.i0 + J.clen ? tile_start + kTile : J.i0 + J.clen;
                const unsigned long long ct = static_cast<unsigned long long>(hi > lo ? hi - lo : 0) * n_solved;
                if (ct) {
                    atomicAdd(&lx.stats[0], ct);
                    atomicAdd(&lx.stats[1], ct);
                }
            }
            if (__any(n_extra != 0)) {
                int32_t extra = n_extra;
                for (int o = 32; o >= 1; o >>= 1) extra += __shfl_xor(extra, o);
                if ((threadIdx.x & 63) == 0 && extra != 0)
                    atomicAdd(&lx.stats[0], static_cast<unsigned long long>(static_cast<int64_t>(extra)));
            }
            if ((threadIdx.x & 63) == 0) {
                if (n_fb) atomicAdd(&lx.stats[2], static_cast<unsigned long long>(n_fb));
                if (n_unc) atomicAdd(&lx.stats[3], static_cast<unsigned long long>(n_unc));
            }
        }
    } else {
        const int a_first = static_cast<int>(first - tile_start);
        const int a_last = static_cast<int>(end - tile_start) - 1;
        const int a_mid = (a_first + a_last) >> 1;
        // a lane may be several anchors when the tile holds fewer than 3 samples
        const int me = static_cast<int>(threadIdx.x);
        for (int a = 0; a < 3; ++a) {
            if (me == (a == 0 ? a_first : (a == 1 ? a_mid : a_last))) {
                s_anc_t[a][0] = tu;
                s_anc_t[a][1] = t_local;
                s_anc_t[a][2] = dlt;
            }
        }
        __syncthreads();
        // tu less its value at the middle sample, from the exact offsets of
        // t_local and delta_t
        const double x = (t_local - s_anc_t[1][1]) - (dlt - s_anc_t[1][2]);
        // rounding of the iteration-1 query time, exact (Sterbenz) per sample
        const double t0q = tu - kEmissionInitialLag;
        const double del0 = ((J.epoch + t0q) - J.epoch) - t0q;
        // the band's part of the carrier phase from the middle sample's
        const double gq = -J.lo * (t_local - s_anc_t[1][1]);
        const double dn = 0.0, dth = 0.0;   // a fixed receiver (the antenna kernels take no other)
        const bool fit_ok = a_last - a_first >= 2;
        // the group's smooth and gated links as bits (wave 0 holds the group)
        __shared__ uint32_t s_mask[4];
        for (uint32_t g0 = 0; g0 < J.n_active; g0 += kAnchorGroup) {
            const uint32_t ng = J.n_active - g0 < kAnchorGroup ? J.n_active - g0 : kAnchorGroup;
            // smooth solves at the three anchors of every transmitter of the group
            for (uint32_t q = threadIdx.x; q < 3 * ng; q += kTile) {
                const uint32_t s = q / 3, a = q - 3 * s;
                const LinkDev L = links[J.link0 + g0 + s];
                anchor_solve(s_anc[s], txs[L.tx], J, s_anc_t[a][0], a, J.obs);
            }
            __syncthreads();
            bool smooth = false, gated = false;
            for (uint32_t s = threadIdx.x; s < ng; s += kTile) {
                const LinkDev L = links[J.link0 + g0 + s];
                const TxDev &T = txs[L.tx];
                smooth = anchor_fit<false>(s_anc[s], nullptr, L, T, J, s_anc_t, fit_ok);
                gated = ANT && gating && g_ok && gate_below_cutoff(J, T, g_tmin, g_tmax, J.obs, 0.0);
                if (ANT && gating && lx.gate_stats) {
                    if (gated) atomicAdd(&lx.gate_stats[0], 1ull);
                    atomicAdd(&lx.gate_stats[1], 1ull);
                }
            }
            // wave 0 holds threads 0..63, hence every link of the group
            if (threadIdx.x < 64) {
                const uint64_t ms = __ballot(threadIdx.x < ng && smooth);
                const uint64_t mg = __ballot(threadIdx.x < ng && gated);
                if (threadIdx.x == 0) {
                    s_mask[0] = static_cast<uint32_t>(ms);
                    s_mask[1] = static_cast<uint32_t>(ms >> 32);
                    s_mask[2] = static_cast<uint32_t>(mg);
                    s_mask[3] = static_cast<uint32_t>(mg >> 32);
                }
            }
            __syncthreads();
            const uint64_t smask = uniform_u64(s_mask[0], s_mask[1]);
            const uint64_t gmask = uniform_u64(s_mask[2], s_mask[3]);
            precision_links<ANT, 1, false>(s_anc, nullptr, smask, gmask, txs, J, g0, ng, &tu, &x, &del0, &gq, &dn, &dth,
                                           J.obs, &sum_re, &sum_im);
            if (tap.on && valid && tap.slot >= g0 && tap.slot < g0 + ng) {
                const uint32_t ts = tap.slot - g0;
                double te, r, ph;
                precision_tap<false>(s_anc[ts], nullptr, (smask >> ts) & 1u, txs, J, tu, t_local, dlt, x, del0, dn, dth, te,
                                     r, ph);
                tap.t_emit[j] = te;
                tap.range[j] = r;
                tap.phase[j] = ph;
            }
            __syncthreads();
        }
        if (J.pass == 1 || J.pass == 2) {
            // a pass before the last: the f64 sums so far go to HBM (uniform return)
            if (tap.on && valid) tap.delta_t[j] = dlt;
            if (valid) {
                double2 *pp = J.part + j;
                if (J.pass == 2) {
                    const double2 prev = *pp;
                    sum_re += prev.x;
                    sum_im += prev.y;
                }
                *pp = make_double2(sum_re, sum_im);
            }
            return;
        }
        if (J.pass == 3 && valid) {
            const double2 prev = J.part[j];
            sum_re += prev.x;
            sum_im += prev.y;
        }
        acc_re = static_cast<float>(sum_re);
        acc_im = static_cast<float>(sum_im);
    }
    if (tap.on && valid) tap.delta_t[j] = dlt;

    // AWGN, src/channel/mod.rs:328-333 (sigma = sqrt(P/2) as f32)
    if (noisy) {
        __syncthreads();
        float zr, zi;
        box_muller(s_noise[threadIdx.x], zr, zi);
        acc_re += J.sigma * zr;
        acc_im += J.sigma * zi;
    }
    if (!valid) return;
    finish_sample(J, acc_re, acc_im, j);
}

// ---------------------------------------------------------------- learned start

// Light time at every pred_qs-th tile boundary of every link of a render
// call (query q = pred0 + link * pred_nb + j of its job); query j sits at
// sample ((i0 >> TS) + j * pred_qs) << TS clamped into the job. The network
// predicts it as the straight-line light time at that sample's reception
// time plus its correction (none for a trajectory without orbital elements),
// and the converged iteration refines the prediction there: the query keeps
// the converged light time, or the prediction when twelve evaluations do not
// converge. The refinement's evaluations are added to stats[0], and queries
// that needed more than kLearnedEvals to stats[2] (fallbacks). 128 queries
// per workgroup, 32 per wave on the matrix cores (lanes l and l + 32 of a
// wave share query l).
__global__ void __launch_bounds__(256) lws_predict_kernel(const JobDev *__restrict__ jobs, int n_jobs,
                                                          const LinkDev *__restrict__ links,
                                                          const TxDev *__restrict__ txs, LwsNetDev net,
                                                          int tile_shift, uint64_t n_queries,
                                                          double *__restrict__ pred,
                                                          unsigned long long *__restrict__ stats) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const uint64_t q = (uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 32 + (lane & 31);
    const bool valid = q < n_queries;
    const uint64_t qc = valid ? q : n_queries - 1;
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].pred0 <= qc) lo = mid;
        else hi = mid - 1;
    }
    const JobDev &J = jobs[lo];
    const uint64_t pitch = J.pred_nb;
    const uint64_t local = qc - J.pred0;
    const uint64_t s = local / pitch;
    const uint64_t b = (local - s * pitch) * J.pred_qs;
    const uint64_t last = J.i0 + J.clen - 1;
    uint64_t ib = ((J.i0 >> tile_shift) + b) << tile_shift;
    ib = ib < J.i0 ? J.i0 : (ib > last ? last : ib);
    const double t_local = static_cast<double>(ib) / J.fs;
    const double dlt = J.delta_t ? J.delta_t[ib - J.i0] : clock_delta(J, t_local);
    const double tu = t_local - dlt;
    const TxDev &T = txs[links[J.link0 + s].tx];
    const double eoff = T.traj_kind ? J.epoch - T.grid_t0 : 0.0;
    double sx, sy, sz;
    tx_position_rel(T, eoff + tu, sx, sy, sz);
    const double dx = sx - J.obs[0], dy = sy - J.obs[1], dz = sz - J.obs[2];
    const double tau_sl = sqrt(dx * dx + dy * dy + dz * dz) * kInvC;
    float x[kLwsIn] = {};
    if (T.kep_on) {
        const double dt = (J.epoch - T.kep_epoch) + tu;
        lws_features(J.obs, T.kep_a, T.kep_cos_i, T.kep_sin_i, T.kep_raan - kEarthRotation * dt,
                     T.kep_nu0 + T.kep_n * dt, x);
    }
    const float y = lws_forward(net, x, nullptr, 0);
    double tau = T.kep_on ? tau_sl + static_cast<double>(y) * kLwsTargetScale : tau_sl;
    double t = tu - tau, r;
    int e = -1;
    for (int k = 1; k <= kColdEvals; ++k) {
        const double tn = emission_step(T, J.obs, tu, eoff, t, r);
        const double step = tn - t;
        t = tn;
        if (fabs(step) < kSolveTol) {
            e = k;
            break;
        }
    }
    if (e > 0) tau = tu - t;
    const bool mine = valid && lane < 32;
    if (mine) pred[q] = tau;
    if (stats) {
        uint32_t ev = mine ? static_cast<uint32_t>(e < 0 ? kColdEvals : e) : 0u;
        for (int o = 16; o >= 1; o >>= 1) ev += __shfl_xor(ev, o);
        const uint32_t fb = __popcll(__ballot(mine && (e < 0 || e > kLearnedEvals)));
        if (lane == 0) {
            if (ev) atomicAdd(&stats[0], static_cast<unsigned long long>(ev));
            if (fb) atomicAdd(&stats[2], static_cast<unsigned long long>(fb));
        }
    }
}

// ---------------------------------------------------------------- truth export

// export_truth's geometry (src/pipeline.rs:548-611) for one receiver: for
// transmitter q and grid time t[k] (epoch-relative), the range of
// solve_emission_times with `iterations` iterations (the reference uses 4 on
// this coarse grid) and the source position at epoch + t[k]. One lane per
// (transmitter, time).
__global__ void __launch_bounds__(256) truth_kernel(const TxDev *__restrict__ txs, uint32_t n_tx,
                                                    const double *__restrict__ obs3, double epoch,
                                                    const double *__restrict__ t, uint64_t n, int iterations,
                                                    double *__restrict__ range, double *__restrict__ src) {
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= uint64_t(n_tx) * n) return;
    const uint32_t q = uint32_t(g / n);
    const uint64_t k = g - uint64_t(q) * n;
    const TxDev &T = txs[q];
    const double tu = t[k];
    const double obs[3] = {obs3[3 * k], obs3[3 * k + 1], obs3[3 * k + 2]};   // the receiver at t[k]
    double tt = tu - kEmissionInitialLag, r = 0.0;
    for (int it = 0; it < iterations; ++it) {
        double sx, sy, sz;
        int64_t cell;
        tx_position(T, epoch + tt, sx, sy, sz, cell);
        const double wtau = kEarthRotation * (tu - tt);
        const double sx_rot = sx + wtau * sy;
        const double sy_rot = sy - wtau * sx;
        const double dx = sx_rot - obs[0];
        const double dy = sy_rot - obs[1];
        const double dz = sz - obs[2];
        r = sqrt(dx * dx + dy * dy + dz * dz);
        tt = tu - r * kInvC;
    }
    range[g] = r;
    double sx, sy, sz;
    int64_t cell;
    tx_position(T, epoch + tu, sx, sy, sz, cell);
    src[(uint64_t(q) * 3 + 0) * n + k] = sx;
    src[(uint64_t(q) * 3 + 1) * n + k] = sy;
    src[(uint64_t(q) * 3 + 2) * n + k] = sz;
}

// ---------------------------------------------------------------- clock walk

// One 65 536-index block of the walk per workgroup, over the segments of
// every receiver of a render call. Each thread owns 256 consecutive
// increments (64 Philox4x32 blocks, 4 normals per block: master index k uses
// block k >> 2, word (k >> 1) & 1, and the cos (even k) or sin (odd k) branch
// of its Box-Muller pair) and sums them sequentially in f64. In window blocks
// it stores, per index, the prefix since the start of its 32-index sub-block
// as f32 (a value a few sigma*sqrt(32) in size; staged through LDS so each
// wavefront writes whole 128-byte rows), and per sub-block the f64 offset of
// that start within the walk block (toff); a fixed-tree scan of the 256
// thread totals places each thread in the block, and the block total goes to
// blocksum.
constexpr int kWalkSub = 1 << kWalkSubShift;
constexpr int kWalkStageRow = kWalkSub + 1;  // padded: conflict-free rows and columns

__global__ void __launch_bounds__(kWalkThreads) walk_gen_kernel(const WalkSegDev *__restrict__ segs, int n_segs,
                                                                const uint32_t *n_live) {
    __shared__ double lds[kWalkThreads];
    __shared__ float stage[kWalkThreads * kWalkStageRow];
    constexpr int kSubsPerThread = kWalkPerThread / kWalkSub;
    if (n_live && blockIdx.x >= *n_live) return;  // streaming plans launch the largest grid
    int lo = 0, hi = n_segs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].block0 <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const WalkSegDev &S = segs[lo];
    const int t = threadIdx.x;
    const int64_t rb = int64_t(blockIdx.x - S.block0);
    const int64_t b = S.first_block + rb;
    const bool win = b >= S.win_block;  // uniform over the workgroup
    const int64_t k0 = (b << kWalkBlockShift) + int64_t(t) * kWalkPerThread;
    float *out = win ? S.local + ((b - S.win_block) << kWalkBlockShift) : nullptr;
    const uint64_t seed = S.seed;
    const double sigma = S.sigma;
    double s = 0.0;
    double subs[kSubsPerThread];
#pragma unroll
    for (int m = 0; m < kSubsPerThread; ++m) {
        subs[m] = s;  // thread-local prefix before sub-block m
        const double sub0 = s;
        for (int q = 0; q < kWalkSub; q += 4) {
            uint64_t w[2];
            philox_words(static_cast<uint64_t>((k0 + m * kWalkSub + q) >> 2), seed, kClockTag, w);
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                float zc, zs;
                box_muller(w[p], zc, zs);
                s = s + (0.0 + sigma * static_cast<double>(zc));
                if (win) stage[t * kWalkStageRow + q + 2 * p] = static_cast<float>(s - sub0);
                s = s + (0.0 + sigma * static_cast<double>(zs));
                if (win) stage[t * kWalkStageRow + q + 2 * p + 1] = static_cast<float>(s - sub0);
            }
        }
        if (win) {
            // thread r's 32 values of sub-block m go to out[r * 256 + m * 32 + c]
            __syncthreads();
#pragma unroll 4
            for (int e = t; e < kWalkThreads * kWalkSub; e += kWalkThreads) {
                const int row = e >> kWalkSubShift, col = e & (kWalkSub - 1);
                out[row * kWalkPerThread + m * kWalkSub + col] = stage[row * kWalkStageRow + col];
            }
            __syncthreads();
        }
    }
    // fixed-tree (Hillis-Steele) inclusive scan of the thread totals
    lds[t] = s;
    __syncthreads();
    for (int off = 1; off < kWalkThreads; off <<= 1) {
        const double add = t >= off ? lds[t - off] : 0.0;
        __syncthreads();
        if (t >= off) lds[t] = lds[t] + add;
        __syncthreads();
    }
    if (win) {
        const double excl = t ? lds[t - 1] : 0.0;
        double *dst = S.toff + ((b - S.win_block) << (kWalkBlockShift - kWalkSubShift)) + int64_t(t) * kSubsPerThread;
#pragma unroll
        for (int m = 0; m < kSubsPerThread; ++m) dst[m] = excl + subs[m];
    }
    if (t == kWalkThreads - 1) S.blocksum[rb] = lds[t];
}

// One workgroup per receiver: base[first + q + 1] = base[first + q] +
// blocksum[q], strictly left to right, 64 block sums staged at a time.
__global__ void __launch_bounds__(64) walk_carry_kernel(const CarrySegDev *__restrict__ segs) {
    __shared__ double buf[64];
    const CarrySegDev &S = segs[blockIdx.x];
    const int t = threadIdx.x;
    double acc = S.base[S.first];
    for (int64_t q0 = 0; q0 < S.n; q0 += 64) {
        const int m = S.n - q0 < 64 ? static_cast<int>(S.n - q0) : 64;
        if (t < m) buf[t] = S.blocksum[q0 + t];
        __syncthreads();
        if (t == 0) {
            for (int i = 0; i < m; ++i) {
                acc = acc + buf[i];
                buf[i] = acc;
            }
        }
        __syncthreads();
        if (t < m) S.base[S.first + q0 + t + 1] = buf[t];
        __syncthreads();
    }
}

}  // namespace lpk

// ---------------------------------------------------------------- launchers
// Plain C++ entry points for lpk_api.hip (same library, not exported).

namespace lpk {

template <int FORM, bool ANT>
using RenderFn = void (*)(const JobDev *, int, const LinkDev *, const TxDev *, RenderLaunch, TapOut);

template <int FORM, bool ANT, bool MOV = false>
RenderFn<FORM, ANT> render_fn(int tile_shift) {
    return tile_shift == 7 ? render_kernel<FORM, ANT, 7, MOV>
                           : (tile_shift == 9 ? render_kernel<FORM, ANT, 9, MOV> : render_kernel<FORM, ANT, 8, MOV>);
}

hipError_t launch_render(int form, bool antenna, bool moving, int tile_shift, bool xcd_map, const JobDev *jobs,
                         int n_jobs, const LinkDev *links, const TxDev *txs, uint32_t n_tiles,
                         const uint32_t *n_live, double *tap_t_emit, double *tap_range, double *tap_phase,
                         double *tap_delta_t, uint32_t tap_slot, const double *pred, unsigned long long *stats,
                         bool gate, unsigned long long *gate_stats, hipStream_t stream) {
    TapOut tap{tap_t_emit, tap_range, tap_phase, tap_delta_t, tap_slot, tap_t_emit ? 1u : 0u};
    if (tile_shift < 7 || tile_shift > 9 || form < 0 || form > 2) return hipErrorInvalidValue;
    if (moving && antenna && form == 1) return hipErrorInvalidValue;   // no such kernel
    void (*kernel)(const JobDev *, int, const LinkDev *, const TxDev *, RenderLaunch, TapOut) =
        form == 0   ? (antenna ? render_fn<0, true>(tile_shift) : render_fn<0, false>(tile_shift))
        : form == 1 ? (antenna ? render_fn<1, true>(tile_shift)
                               : (moving ? render_fn<1, false, true>(tile_shift) : render_fn<1, false>(tile_shift)))
                    : (antenna ? render_fn<2, true>(tile_shift) : render_fn<2, false>(tile_shift));
    for (uint32_t base = 0; base < n_tiles; base += kMaxLaunchTiles) {
        const uint32_t n = n_tiles - base < kMaxLaunchTiles ? n_tiles - base : kMaxLaunchTiles;
        const RenderLaunch lx{base, n, n_live, xcd_map ? 1u : 0u, pred, stats, gate ? 1u : 0u, gate_stats};
        const uint32_t lanes = form == 1 && !antenna ? (1u << tile_shift) / kSpanSamples : (1u << tile_shift);
        hipLaunchKernelGGL(kernel, dim3(n), dim3(lanes), 0, stream, jobs, n_jobs, links, txs, lx, tap);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_lws_predict(const JobDev *jobs, int n_jobs, const LinkDev *links, const TxDev *txs,
                              const LwsNetDev &net, int tile_shift, uint64_t n_queries, double *pred,
                              unsigned long long *stats, hipStream_t stream) {
    if (!n_queries || n_jobs <= 0) return hipSuccess;
    const uint64_t blocks = (n_queries + 127) / 128;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lws_predict_kernel, dim3(uint32_t(blocks)), dim3(256), 0, stream, jobs, n_jobs, links, txs,
                       net, tile_shift, n_queries, pred, stats);
    return hipGetLastError();
}

hipError_t launch_truth(const TxDev *txs, uint32_t n_tx, const double *obs, double epoch, const double *t, uint64_t n,
                        int iterations, double *range, double *src, hipStream_t stream) {
    const uint64_t total = uint64_t(n_tx) * n;
    if (!total) return hipSuccess;
    const uint64_t blocks = (total + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(truth_kernel, dim3(uint32_t(blocks)), dim3(256), 0, stream, txs, n_tx, obs, epoch, t, n,
                       iterations, range, src);
    return hipGetLastError();
}

hipError_t launch_walk_gen(const WalkSegDev *segs, int n_segs, uint32_t n_blocks, const uint32_t *n_live,
                           hipStream_t stream) {
    if (n_blocks == 0 || n_segs <= 0) return hipSuccess;
    hipLaunchKernelGGL(walk_gen_kernel, dim3(n_blocks), dim3(kWalkThreads), 0, stream, segs, n_segs, n_live);
    return hipGetLastError();
}

hipError_t launch_walk_carry(const CarrySegDev *segs, int n_segs, hipStream_t stream) {
    if (n_segs <= 0) return hipSuccess;
    hipLaunchKernelGGL(walk_carry_kernel, dim3(uint32_t(n_segs)), dim3(64), 0, stream, segs);
    return hipGetLastError();
}

}  // namespace lpk
