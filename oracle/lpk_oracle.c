/*
 * lpk_oracle.c -- CPU FP64 oracle for the per-chunk IQ path.
 *
 * TEST INFRASTRUCTURE ONLY (see lpk_oracle.h). Never linked into the product.
 * Built with -O3 -ffp-contract=off and without fast-math so that every f64
 * and f32 operation rounds exactly as the reference's Rust (which never
 * contracts to FMA) does. Transcendentals (sin, cos, sqrt, pow, log) come from
 * glibc libm, which is also what the reference's f64::sin etc. call on Linux.
 */
#define _GNU_SOURCE
#include "lpk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/lib.rs:53 */
static const double SPEED_OF_LIGHT_M_S = 299792458.0;
/* src/channel/mod.rs:23 */
static const double EMISSION_INITIAL_LAG_S = 0.07;
/* src/channel/mod.rs:27, src/geometry/kepler.rs:13 */
static const double EARTH_ROTATION_RATE_RAD_S = 7.2921151467e-5;
/* src/geometry/kepler.rs:12 */
static const double GM_EARTH = 3.986005e14;
/* src/geometry/ecef.rs:3-5 */
static const double WGS84_A = 6378137.0;
#define WGS84_F (1.0 / 298.257223563)
#define WGS84_E2 (WGS84_F * (2.0 - WGS84_F))
/* std::f64::consts::TAU, PI */
static const double TAU = 6.283185307179586;
static const double PI = 3.141592653589793;
/* src/channel/mod.rs:301 */
static const double BOLTZMANN_J_PER_K = 1.380649e-23;
/* src/signal/ca_code.rs:23-31 */
static const double CA_CHIP_RATE_HZ = 1023000.0;
#define CA_CODE_LENGTH 1023
static const double NAV_BIT_RATE_HZ = 50.0;
static const int64_t NAV_SUBFRAME_BITS = 300;
static const uint8_t NAV_PREAMBLE_BITS[8] = {1, 0, 0, 0, 1, 0, 1, 1};
/* src/pipeline.rs:56 */
static const int NEWTON_ITERATIONS = 2;

/* Rust `x as i64`: saturating, NaN -> 0. */
static int64_t sat_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
/* i64::rem_euclid for n > 0 */
static int64_t rem_euclid_i64(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}
/* f64::rem_euclid: r = self % rhs; if r < 0 { r + |rhs| } */
static double rem_euclid_f64(double a, double b) {
    double r = fmod(a, b);
    return r < 0.0 ? r + fabs(b) : r;
}
/* f64::to_radians: self * (PI / 180) with the constant rounded once */
static double to_radians(double d) { return d * (PI / 180.0); }

/* ------------------------------------------------------------------ geometry */

/* src/geometry/ecef.rs:10-20 */
void orc_geodetic_to_ecef(double lat_deg, double lon_deg, double h_m, double out[3]) {
    double lat = to_radians(lat_deg);
    double lon = to_radians(lon_deg);
    double sin_lat, cos_lat;
    sincos(lat, &sin_lat, &cos_lat);
    double sin_lon, cos_lon;
    sincos(lon, &sin_lon, &cos_lon);
    double n = WGS84_A / sqrt(1.0 - WGS84_E2 * sin_lat * sin_lat);
    out[0] = (n + h_m) * cos_lat * cos_lon;
    out[1] = (n + h_m) * cos_lat * sin_lon;
    out[2] = (n * (1.0 - WGS84_E2) + h_m) * sin_lat;
}

/* src/geometry/kepler.rs:27-85 (eci_at + eci_to_ecef, position only) */
void orc_kepler_position(double a_m, double inc_deg, double raan_deg, double nu0_deg,
                         double epoch_unix, double t_unix, double out[3]) {
    double r = a_m;
    double n = sqrt(GM_EARTH / (a_m * a_m * a_m));   /* a.powi(3) = a*(a*a) = (a*a)*a */
    double dt = t_unix - epoch_unix;
    double nu = to_radians(nu0_deg) + n * dt;
    double s_nu, c_nu;
    sincos(nu, &s_nu, &c_nu);
    double inc = to_radians(inc_deg), raan = to_radians(raan_deg);
    double s_i, c_i, s_o, c_o;
    sincos(inc, &s_i, &c_i);
    sincos(raan, &s_o, &c_o);
    double xp = r * c_nu;
    double yp = r * s_nu;
    double y1 = yp * c_i;
    double z1 = yp * s_i;
    double x_eci = xp * c_o - y1 * s_o;
    double y_eci = xp * s_o + y1 * c_o;
    double z_eci = z1;
    double theta = EARTH_ROTATION_RATE_RAD_S * (t_unix - epoch_unix);
    double s, c;
    sincos(theta, &s, &c);
    out[0] = x_eci * c + y_eci * s;
    out[1] = -x_eci * s + y_eci * c;
    out[2] = z_eci;
}

/* src/pipeline.rs:352-364, :929-930: the receiver's trajectory at t_unix */
void orc_rx_position(const orc_rx *rx, double t_unix, double out[3]) {
    if (rx->traj_kind == 1) {
        orc_kepler_position(rx->kep_a_m, rx->kep_inc_deg, rx->kep_raan_deg, rx->kep_nu0_deg, rx->kep_epoch_unix,
                            t_unix, out);
    } else {
        out[0] = rx->pos_ecef[0];
        out[1] = rx->pos_ecef[1];
        out[2] = rx->pos_ecef[2];
    }
}

/* src/geometry/precomputed.rs:54-58 */
uint64_t orc_grid_size(double duration_s, double grid_dt_s) {
    double total = duration_s + 2.0 * 0.5;
    return (uint64_t)ceil(total / grid_dt_s) + 1;
}

/* src/geometry/precomputed.rs:47-70 with the circular Keplerian inner trajectory */
uint64_t orc_grid_sample_kepler(double a_m, double inc_deg, double raan_deg, double nu0_deg,
                                double orbit_epoch_unix, double t0_unix, double duration_s,
                                double grid_dt_s, double *t0_out, double *xyz_out) {
    double pad_s = 0.5;
    double start = t0_unix - pad_s;
    uint64_t n = orc_grid_size(duration_s, grid_dt_s);
    for (uint64_t i = 0; i < n; ++i) {
        double t = start + (double)i * grid_dt_s;
        double p[3];
        orc_kepler_position(a_m, inc_deg, raan_deg, nu0_deg, orbit_epoch_unix, t, p);
        xyz_out[i] = p[0];
        xyz_out[n + i] = p[1];
        xyz_out[2 * n + i] = p[2];
    }
    *t0_out = start;
    return n;
}

/* src/geometry/precomputed.rs:111-121 (position_ecef: division by dt) */
static void grid_lerp(double f, uint64_t n, const double *xyz, double out[3]) {
    double fl = floor(f);
    int64_t max_lo = (int64_t)n - 2;
    int64_t lo = sat_i64(fl);
    if (lo < 0) lo = 0;
    if (lo > max_lo) lo = max_lo;
    double frac = f - (double)lo;
    if (frac < 0.0) frac = 0.0;
    if (frac > 1.0) frac = 1.0;
    double omf = 1.0 - frac;
    out[0] = xyz[lo] * omf + xyz[lo + 1] * frac;
    out[1] = xyz[n + lo] * omf + xyz[n + lo + 1] * frac;
    out[2] = xyz[2 * n + lo] * omf + xyz[2 * n + lo + 1] * frac;
}

void orc_grid_position(double t0_unix, double dt_s, uint64_t n, const double *xyz,
                       double t_unix, double out[3]) {
    grid_lerp((t_unix - t0_unix) / dt_s, n, xyz, out);
}

static void tx_position(const orc_tx *tx, double t_unix, double out[3]) {
    if (tx->traj_kind == ORC_TRAJ_STATIONARY) {   /* src/geometry/stationary.rs:21-24 */
        out[0] = tx->pos_ecef[0];
        out[1] = tx->pos_ecef[1];
        out[2] = tx->pos_ecef[2];
    } else {
        orc_grid_position(tx->grid_t0_unix, tx->grid_dt_s, tx->grid_n, tx->grid_xyz, t_unix, out);
    }
}

/* The converged solve: the iteration of src/channel/mod.rs:75-96 from emission
 * time t_start, with the grid queried at (epoch - grid_t0) + t after its node 0
 * (src/geometry/precomputed.rs:108-121 without rounding the absolute time, and
 * with the grid coordinate scaled by 1 / dt rather than divided by dt),
 * until an update |t_next - t| < tol. Returns the evaluations used, or -1 if
 * max_evals did not converge; *t_emit = last iterate, *range = the range it
 * came from. */
int orc_solve_converged(const orc_tx *tx, const double obs[3], double tu, double epoch, double t_start,
                        double tol, int max_evals, double *t_emit, double *range) {
    const double inv_c = 1.0 / SPEED_OF_LIGHT_M_S;
    const double eoff = tx->traj_kind == ORC_TRAJ_STATIONARY ? 0.0 : epoch - tx->grid_t0_unix;
    const double inv_dt = tx->traj_kind == ORC_TRAJ_STATIONARY ? 0.0 : 1.0 / tx->grid_dt_s;
    double t = t_start, r = 0.0;
    for (int e = 1; e <= max_evals; ++e) {
        double s[3];
        if (tx->traj_kind == ORC_TRAJ_STATIONARY) {
            s[0] = tx->pos_ecef[0];
            s[1] = tx->pos_ecef[1];
            s[2] = tx->pos_ecef[2];
        } else {
            grid_lerp((eoff + t) * inv_dt, tx->grid_n, tx->grid_xyz, s);
        }
        double wtau = EARTH_ROTATION_RATE_RAD_S * (tu - t);
        double sx_rot = s[0] + wtau * s[1];
        double sy_rot = s[1] - wtau * s[0];
        double dx = sx_rot - obs[0];
        double dy = sy_rot - obs[1];
        double dz = s[2] - obs[2];
        r = sqrt(dx * dx + dy * dy + dz * dz);
        double tn = tu - r * inv_c;
        double step = tn - t;
        t = tn;
        if (fabs(step) < tol) {
            *t_emit = t;
            *range = r;
            return e;
        }
    }
    *t_emit = t;
    *range = r;
    return -1;
}

/* src/channel/mod.rs:70-99 (one sample of solve_emission_times_for) */
void orc_solve_emission(const orc_tx *tx, const double obs[3], double tu, double epoch,
                        int iterations, double *t_emit, double *range) {
    const double inv_c = 1.0 / SPEED_OF_LIGHT_M_S;
    double ox = obs[0], oy = obs[1], oz = obs[2];
    double t = tu - EMISSION_INITIAL_LAG_S;
    double r = 0.0;
    for (int it = 0; it < iterations; ++it) {
        double s[3];
        tx_position(tx, epoch + t, s);
        double wtau = EARTH_ROTATION_RATE_RAD_S * (tu - t);
        double sx_rot = s[0] + wtau * s[1];
        double sy_rot = s[1] - wtau * s[0];
        double dx = sx_rot - ox;
        double dy = sy_rot - oy;
        double dz = s[2] - oz;
        r = sqrt(dx * dx + dy * dy + dz * dz);
        t = tu - r * inv_c;
    }
    *t_emit = t;
    *range = r;
}

/* ------------------------------------------------------------------ signals */

/* src/signal/ca_code.rs:42-78 */
int orc_ca_code(int prn, int8_t out[1023]) {
    static const int taps[32][2] = {
        {2, 6}, {3, 7}, {4, 8}, {5, 9}, {1, 9}, {2, 10}, {1, 8}, {2, 9}, {3, 10},
        {2, 3}, {3, 4}, {5, 6}, {6, 7}, {7, 8}, {8, 9}, {9, 10}, {1, 4}, {2, 5},
        {3, 6}, {4, 7}, {5, 8}, {6, 9}, {1, 3}, {4, 6}, {5, 7}, {6, 8}, {7, 9},
        {8, 10}, {1, 6}, {2, 7}, {3, 8}, {4, 9}};
    if (prn < 1 || prn > 32) return -1;
    int a = taps[prn - 1][0], b = taps[prn - 1][1];
    int g1[10], g2[10];
    for (int i = 0; i < 10; ++i) g1[i] = g2[i] = 1;
    for (int k = 0; k < CA_CODE_LENGTH; ++k) {
        int o1 = g1[9];
        int o2 = g2[a - 1] ^ g2[b - 1];
        int fb1 = g1[2] ^ g1[9];
        int fb2 = g2[1] ^ g2[2] ^ g2[5] ^ g2[7] ^ g2[8] ^ g2[9];
        for (int i = 9; i >= 1; --i) { g1[i] = g1[i - 1]; g2[i] = g2[i - 1]; }
        g1[0] = fb1;
        g2[0] = fb2;
        out[k] = (int8_t)(1 - 2 * (o1 ^ o2));
    }
    return 0;
}

/* src/signal/ca_code.rs:94-108 */
int8_t orc_nav_dummy_bit_pm(uint64_t seed, int64_t bit_idx) {
    int64_t pos = rem_euclid_i64(bit_idx, NAV_SUBFRAME_BITS);
    int bit;
    if (pos < 8) {
        bit = NAV_PREAMBLE_BITS[pos];
    } else {
        uint64_t x = seed ^ ((uint64_t)bit_idx * 0x9E3779B97F4A7C15ULL);
        x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ULL;
        x ^= x >> 27; x *= 0x94D049BB133111EBULL;
        x ^= x >> 31;
        bit = (int)(x & 1);
    }
    return (int8_t)(1 - 2 * bit);
}

/* src/signal/ca_code.rs:146-151 */
uint64_t orc_nav_fold_seed(uint64_t nav_seed, uint64_t prn) {
    return nav_seed + prn * 0xD1B54A32D192ED03ULL;
}

/* src/signal/mod.rs:50-69 */
int orc_elapsed_periods(double elapsed_s, double rate_hz, int64_t *whole_out, double *frac_out) {
    double p = elapsed_s * rate_hz;
    if (!isfinite(p) || fabs(p) >= 9.0e18) return -1;
    double err = fma(elapsed_s, rate_hz, -p);
    double whole = floor(p);
    double frac = (p - whole) + err;
    if (frac < 0.0) {
        whole -= 1.0;
        frac += 1.0;
    } else if (frac >= 1.0) {
        whole += 1.0;
        frac -= 1.0;
    }
    *whole_out = (int64_t)whole;
    *frac_out = frac;
    return 0;
}

/* src/signal/mod.rs:87-101 */
int orc_frame_anchor(double render_epoch, double signal_epoch, double rate_hz, uint64_t n_symbols,
                     double *epoch_rel_s, int64_t *symbol_offset) {
    if (!(rate_hz > 0.0)) return -1;
    int64_t whole;
    double frac;
    if (orc_elapsed_periods(render_epoch - signal_epoch, rate_hz, &whole, &frac)) return -1;
    *epoch_rel_s = -frac / rate_hz;
    int64_t n = n_symbols > 0 ? (int64_t)n_symbols : 1;
    *symbol_offset = rem_euclid_i64(whole, n);
    return 0;
}

/* src/signal/cw.rs:26-34 */
int orc_cw_anchor_phase(double render_epoch, double signal_epoch, double offset_hz, double *phase) {
    int64_t whole;
    double frac;
    if (orc_elapsed_periods(render_epoch - signal_epoch, offset_hz, &whole, &frac)) return -1;
    *phase = TAU * frac;
    return 0;
}

/* src/signal/mod.rs:108-115: returns -1 for "silent" */
static int64_t frame_index(int64_t raw, int64_t n, int repeating, int64_t k) {
    if (!repeating) return (raw >= 0 && raw < n) ? raw : -1;
    if (n <= 0) return -1;
    return rem_euclid_i64((int64_t)((uint64_t)raw + (uint64_t)k), n);
}

/* src/signal/gate.rs:67-77 */
int orc_gate_is_on(const orc_tx *tx, double t) {
    double u = rem_euclid_f64(t + tx->gate_phase_s, tx->gate_period_s);
    if (u >= tx->gate_period_s) u = 0.0;
    for (uint32_t b = 0; b < tx->n_bursts; ++b) {
        if (u >= tx->bursts[2 * b] && u < tx->bursts[2 * b + 1]) return 1;
    }
    return 0;
}

/* One sample of evaluate_baseband_batch for every signal kind:
 *   CW     src/signal/cw.rs:41-52
 *   BPSK   src/signal/bpsk.rs:58-71
 *   OOK    src/signal/ook.rs:60-74
 *   QAM16  src/signal/qam16.rs:96-109
 *   C/A    src/signal/ca_code.rs:167-193
 *   gate   src/signal/gate.rs:91-104 */
static void baseband_one(const orc_tx *tx, double t, float *re, float *im) {
    if (tx->gated && !orc_gate_is_on(tx, t)) {
        *re = 0.0f;
        *im = 0.0f;
        return;
    }
    switch (tx->signal_kind) {
    case ORC_SIG_CW: {
        double omega = TAU * tx->baseband_offset_hz;
        float phase;
        if (!tx->cw_anchored) phase = (float)(omega * t);
        else phase = (float)rem_euclid_f64(omega * t + tx->anchor_phase_rad, TAU);
        sincosf(phase, im, re);
        return;
    }
    case ORC_SIG_BPSK:
    case ORC_SIG_OOK:
    case ORC_SIG_QAM16: {
        double idx_f = (t - tx->epoch_rel_s) * tx->rate_hz;
        int64_t raw = sat_i64(floor(idx_f));
        int64_t k = frame_index(raw, (int64_t)tx->n_symbols, tx->repeating, tx->repeat_from);
        if (k < 0) {
            *re = 0.0f;
            *im = 0.0f;
        } else if (tx->signal_kind == ORC_SIG_QAM16) {
            *re = tx->symbols[2 * k];
            *im = tx->symbols[2 * k + 1];
        } else {
            *re = (float)tx->bits[k];
            *im = 0.0f;
        }
        return;
    }
    case ORC_SIG_CA: {
        double dt = t - tx->epoch_rel_s;
        double chip_phase = dt * CA_CHIP_RATE_HZ;
        int64_t idx = sat_i64(floor(chip_phase));
        int64_t idx_mod = rem_euclid_i64(idx, CA_CODE_LENGTH);
        int chip = tx->ca_code[idx_mod];
        int nav_pm = 1;
        if (tx->nav_mode == ORC_NAV_DUMMY) {
            nav_pm = orc_nav_dummy_bit_pm(tx->nav_seed, sat_i64(floor(dt * NAV_BIT_RATE_HZ)));
        } else if (tx->nav_mode == ORC_NAV_FULL) {
            int64_t bit_idx = sat_i64(floor(dt * NAV_BIT_RATE_HZ));
            if (tx->n_nav_bits == 0) nav_pm = 1;
            else if (bit_idx < 0) nav_pm = tx->nav_bits[0];
            else if ((uint64_t)bit_idx < tx->n_nav_bits) nav_pm = tx->nav_bits[bit_idx];
            else nav_pm = tx->nav_bits[tx->n_nav_bits - 1];
        }
        chip *= nav_pm;
        *re = (float)chip;
        *im = 0.0f;
        return;
    }
    default:
        *re = 0.0f;
        *im = 0.0f;
    }
}

void orc_baseband(const orc_tx *tx, const double *t_emit, uint64_t n, float *out_iq) {
    for (uint64_t i = 0; i < n; ++i) baseband_one(tx, t_emit[i], &out_iq[2 * i], &out_iq[2 * i + 1]);
}

/* src/pipeline.rs:70-75 */
int orc_bands_overlap(double tx_carrier, double tx_bw, double band_lo, double band_bw) {
    return fabs(tx_carrier - band_lo) <= 0.5 * (tx_bw + band_bw);
}

/* ------------------------------------------- reference payload RNG (restated) */

static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* rand_xoshiro 0.6 Xoshiro256PlusPlus::seed_from_u64: four SplitMix64 outputs. */
void orc_xoshiro_seed(uint64_t seed, uint64_t s[4]) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
        x += 0x9E3779B97F4A7C15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s[i] = z ^ (z >> 31);
    }
}

/* xoshiro256++ (Blackman & Vigna) */
uint64_t orc_xoshiro_next(uint64_t s[4]) {
    uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* src/signal/bpsk.rs:27-41: gen_bool(0.5) is `next_u64 < 2^63`. */
void orc_payload_bpsk(const int8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out) {
    uint64_t s[4];
    orc_xoshiro_seed(seed, s);
    for (uint64_t i = 0; i < n_sync; ++i) out[i] = sync[i];
    for (uint64_t i = 0; i < n_payload; ++i)
        out[n_sync + i] = (orc_xoshiro_next(s) < 0x8000000000000000ULL) ? 1 : -1;
}

/* src/signal/ook.rs:26-44 */
int orc_payload_ook(const uint8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out) {
    for (uint64_t i = 0; i < n_sync; ++i)
        if (sync[i] > 1) return -1;
    uint64_t s[4];
    orc_xoshiro_seed(seed, s);
    for (uint64_t i = 0; i < n_sync; ++i) out[i] = (int8_t)sync[i];
    for (uint64_t i = 0; i < n_payload; ++i)
        out[n_sync + i] = (orc_xoshiro_next(s) < 0x8000000000000000ULL) ? 1 : 0;
    return 0;
}

/* src/signal/qam16.rs:20-34 */
static float gray_level(int b0, int b1) {
    if (b0 == 0 && b1 == 0) return -3.0f;
    if (b0 == 0 && b1 == 1) return -1.0f;
    if (b0 == 1 && b1 == 1) return 1.0f;
    return 3.0f;
}
static void map_symbol(int b0, int b1, int b2, int b3, float *re, float *im) {
    const float scale = 0.31622777f;   /* src/signal/qam16.rs:17 */
    *re = gray_level(b0, b1) * scale;
    *im = gray_level(b2, b3) * scale;
}

/* src/signal/qam16.rs:62-80: gen_range(0..16u8) is `next_u64 >> 60`. */
int orc_payload_qam16(const uint8_t *sync, uint64_t n_sync_bits, uint64_t n_payload, uint64_t seed, float *out) {
    for (uint64_t i = 0; i < n_sync_bits; ++i)
        if (sync[i] > 1) return -1;
    if (n_sync_bits % 4) return -1;
    uint64_t ns = n_sync_bits / 4;
    for (uint64_t k = 0; k < ns; ++k)
        map_symbol(sync[4 * k], sync[4 * k + 1], sync[4 * k + 2], sync[4 * k + 3], &out[2 * k], &out[2 * k + 1]);
    uint64_t s[4];
    orc_xoshiro_seed(seed, s);
    for (uint64_t i = 0; i < n_payload; ++i) {
        unsigned v = (unsigned)(orc_xoshiro_next(s) >> 60);
        map_symbol((v >> 3) & 1, (v >> 2) & 1, (v >> 1) & 1, v & 1, &out[2 * (ns + i)], &out[2 * (ns + i) + 1]);
    }
    return 0;
}

/* ------------------------------------------------ project RNG: Philox4x64-10 */

static void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) {
    unsigned __int128 p = (unsigned __int128)a * b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}

void orc_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11, with the multipliers
 * and key increments of their Random123 library). */
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Word q (0 or 1) of block `block` of a (seed, tag) stream: Philox4x32-10 on
 * counter (block, tag) with key seed, outputs 2q+1 : 2q as one 64-bit word. */
static uint64_t stream_word(uint64_t block, uint64_t seed, uint64_t tag, unsigned q) {
    uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32), (uint32_t)tag, (uint32_t)(tag >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, out[4];
    orc_philox4x32(ctr, key, out);
    return ((uint64_t)out[2 * q + 1] << 32) | out[2 * q];
}

/* Domain tags that separate the clock-walk and AWGN streams of one seed. */
#define LPK_CLOCK_TAG 0x636C6F636B77616CULL   /* "clockwal" */
#define LPK_AWGN_TAG 0x6177676E6E6F6973ULL    /* "awgnnois" */
#define LPK_FLICKER_TAG 0x666C69636B65726EULL /* "flickern" */
/* src/clock/mod.rs:67-69, and tau = 2 s of the calibration (:106) */
static const double FLICKER_RATE_HZ = 128.0;
static const double FLICKER_F_LO = 1e-4;
static const int FLICKER_CALIB_M = 256;

/* Box-Muller pair from one 64-bit Philox word: u1 = (m1 + 1/2) 2^-24 from bits
 * 63..40, angle u2 = m2 2^-24 turns from bits 39..16. The project defines the
 * pair with correctly rounded f64 operations only (+ - * / sqrt, no FMA, no
 * libm), so any IEEE-754 machine reproduces its bits; each value is rounded
 * once to f32. */
void orc_box_muller(uint64_t w, float *zc, float *zs) {
    /* ln of u1 = n 2^-24, n = 2 m1 + 1 from bits 63..41 (exact in f32): exact
     * split n = f 2^e with f in [sqrt(1/2), sqrt(2)], then ln f = 2
     * atanh((f-1)/(f+1)) by its series; every step in f32 with explicit fmaf. */
    uint32_t n = ((uint32_t)(w >> 41) << 1) | 1u;
    int e = 31 - __builtin_clz(n);
    float f = ldexpf((float)n, -e);
    if (f > 1.41421356f) {
        f = f * 0.5f;
        e += 1;
    }
    float s = (f - 1.0f) / (f + 1.0f), z = s * s;
    float p = fmaf(z, 1.0f / 9.0f, 1.0f / 7.0f);
    p = fmaf(p, z, 1.0f / 5.0f);
    p = fmaf(p, z, 1.0f / 3.0f);
    p = fmaf(p, z, 1.0f);
    float ln_u1 = fmaf((float)(e - 24), 0.693147182f, (2.0f * s) * p);
    float r = sqrtf(-2.0f * ln_u1);
    /* angle 2 pi m2 2^-24: quadrant and reflection about pi/4 exactly in
     * integers, then the Taylor series of sin and cos on [0, pi/4]. */
    uint32_t m2 = (uint32_t)(w >> 16) & 0xFFFFFFu;
    uint32_t quad = m2 >> 22, v = m2 & 0x3FFFFFu;
    int flip = v > 0x200000u;
    float th = (float)(flip ? 0x400000u - v : v) * (1.57079632679489662f * 0x1.0p-22f);
    float t2 = th * th;
    float ps = fmaf(t2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = fmaf(ps, t2, 1.0f / 120.0f);
    ps = fmaf(ps, t2, -1.0f / 6.0f);
    float pc = fmaf(t2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = fmaf(pc, t2, -1.0f / 720.0f);
    pc = fmaf(pc, t2, 1.0f / 24.0f);
    pc = fmaf(pc, t2, -0.5f);
    float sn = fmaf(th * t2, ps, th), cs = fmaf(t2, pc, 1.0f);
    float a = flip ? cs : sn, b = flip ? sn : cs, sv, cv;
    switch (quad) {
    case 0: sv = a; cv = b; break;
    case 1: sv = b; cv = -a; break;
    case 2: sv = -a; cv = -b; break;
    default: sv = -b; cv = a; break;
    }
    *zc = r * cv;
    *zs = r * sv;
}

/* Standard normal for master-grid index k of a receiver's clock walk: word
 * (k >> 1) & 1 of block k >> 2 of the (seed, "clockwal") stream, the cos
 * branch of its Box-Muller pair for even k and the sin branch for odd k. */
double orc_clock_normal(uint64_t clock_seed, uint64_t k) {
    float zc, zs;
    orc_box_muller(stream_word(k >> 2, clock_seed, LPK_CLOCK_TAG, (unsigned)((k >> 1) & 1)), &zc, &zs);
    return (double)((k & 1) ? zs : zc);
}

/* Two standard normals (re, im) for global sample i of a band's noise stream:
 * word i & 1 of block i >> 1 of the (seed, "awgnnois") stream, both branches
 * of its Box-Muller pair. */
void orc_awgn_normals(uint64_t noise_seed, uint64_t i, float *z_re, float *z_im) {
    orc_box_muller(stream_word(i >> 1, noise_seed, LPK_AWGN_TAG, (unsigned)(i & 1)), z_re, z_im);
}

/* Normal n of a receiver's flicker stream: block n >> 3 of Philox4x64-10
 * (ctr = (n >> 3, 0, 0, 0), key = (seed, "flickern")), word (n >> 1) & 3, the
 * cos branch for even n and the sin branch for odd n. Step k of pole j uses
 * n = (k + 1) * 20 + j; n < 20 start the poles. */
double orc_flicker_normal(uint64_t clock_seed, uint64_t n) {
    uint64_t ctr[4] = {n >> 3, 0, 0, 0}, key[2] = {clock_seed, LPK_FLICKER_TAG}, w[4];
    orc_philox4x64(ctr, key, w);
    float zc, zs;
    orc_box_muller(w[(n >> 1) & 3], &zc, &zs);
    return (double)((n & 1) ? zs : zc);
}

/* src/clock/mod.rs:71-80 */
void orc_flicker_poles(double *alpha, double *gain) {
    double f_hi = FLICKER_RATE_HZ / 2.0;
    for (int j = 0; j < ORC_FLICKER_POLES; ++j) {
        double f = FLICKER_F_LO * pow(f_hi / FLICKER_F_LO, (double)j / (double)(ORC_FLICKER_POLES - 1));
        double a = exp(-2.0 * PI * f / FLICKER_RATE_HZ);
        alpha[j] = a;
        gain[j] = sqrt(1.0 - a * a);
    }
}

/* The calibration of src/clock/mod.rs:98-108 (scale = floor / the Allan
 * deviation at tau = 2 s of the unit-scaled process), with the estimator's
 * expectation in place of one random realisation: for a unit AR(1) pole a,
 * (E[S1^2] - E[S1 S2]) / m^2 over two adjacent blocks of m steps (the
 * non-overlapping estimator of :150-163), summed over the independent poles. */
double orc_flicker_unit_adev(const double *alpha, int m) {
    double var = 0.0;
    for (int j = 0; j < ORC_FLICKER_POLES; ++j) {
        double a = alpha[j], s11 = (double)m, s12 = 0.0, p = 1.0;
        for (int d = 1; d < 2 * m; ++d) {
            p *= a;
            if (d < m) s11 += 2.0 * (double)(m - d) * p;
            s12 += (d <= m ? (double)d : (double)(2 * m - d)) * p;
        }
        var += (s11 - s12) / ((double)m * (double)m);
    }
    return sqrt(var);
}

/* src/clock/mod.rs:125-133 */
static void flicker_advance_to(orc_clock *c, int64_t m) {
    while (c->f_frontier < m) {
        c->f_lo = c->f_hi;
        double y = 0.0;
        for (int j = 0; j < ORC_FLICKER_POLES; ++j) {
            c->f_state[j] = c->f_alpha[j] * c->f_state[j] +
                            c->f_gain[j] * orc_flicker_normal(c->seed, (uint64_t)(c->f_frontier + 2) * ORC_FLICKER_POLES + j);
            y += c->f_state[j];
        }
        y = y * c->f_scale;
        c->f_hi += y / FLICKER_RATE_HZ;
        c->f_frontier += 1;
    }
}

/* src/clock/mod.rs:139-148 */
static double flicker_delta_at(orc_clock *c, double elapsed) {
    if (!c->flick) return 0.0;
    double x = fmax(elapsed, 0.0) * FLICKER_RATE_HZ;
    int64_t m0 = sat_i64(floor(x));
    double frac = x - (double)m0;
    flicker_advance_to(c, m0 + 1);
    return c->f_lo * (1.0 - frac) + c->f_hi * frac;
}

/* ---------------------------------------------------------------- the clock */

/* src/clock/mod.rs:268-295. The flicker process starts its poles from their
 * stationary distribution N(0, 1), which the reference's 200 000 warm-up steps
 * approach (:110-116). */
void orc_clock_init(orc_clock *c, const orc_rx *rx) {
    c->c0 = rx->initial_time_offset_s;
    c->y0 = rx->frequency_offset_y0;
    c->master = rx->master_rate_hz;
    c->walk = rx->allan_deviation_1s > 0.0;
    c->sigma_step = c->walk ? rx->allan_deviation_1s / sqrt(rx->master_rate_hz) : 0.0;
    c->seed = rx->clock_seed;
    c->frontier = -1;
    c->eps_hi = 0.0;
    c->eps_lo = 0.0;
    c->flick = rx->flicker_floor > 0.0;
    c->f_frontier = -1;
    c->f_hi = 0.0;
    c->f_lo = 0.0;
    if (c->flick) {
        orc_flicker_poles(c->f_alpha, c->f_gain);
        c->f_scale = rx->flicker_floor / orc_flicker_unit_adev(c->f_alpha, FLICKER_CALIB_M);
        for (int j = 0; j < ORC_FLICKER_POLES; ++j) c->f_state[j] = orc_flicker_normal(c->seed, (uint64_t)j);
    }
}

/* src/clock/mod.rs:299-307 */
static void clock_advance_to(orc_clock *c, int64_t k) {
    while (c->frontier < k) {
        c->eps_lo = c->eps_hi;
        if (c->walk) c->eps_hi += 0.0 + c->sigma_step * orc_clock_normal(c->seed, (uint64_t)(c->frontier + 1));
        c->frontier += 1;
    }
}

/* src/clock/mod.rs:318-324 */
double orc_clock_delta_t_at_master_index(orc_clock *c, int64_t k) {
    clock_advance_to(c, k > 0 ? k : 0);
    double eps = k < 0 ? 0.0 : c->eps_hi;
    double elapsed = (double)(k > 0 ? k : 0) / c->master;
    double flick = flicker_delta_at(c, elapsed);
    return (eps + flick) + (c->c0 + c->y0 * elapsed);
}

/* src/clock/mod.rs:329-346 (clock start epoch 0.0, src/pipeline.rs:869-872) */
double orc_clock_delta_t_at(orc_clock *c, double t) {
    double elapsed = t - 0.0;
    double x = elapsed * c->master;
    if (x <= 0.0) {
        clock_advance_to(c, 0);
        double flick = flicker_delta_at(c, 0.0);
        return (c->eps_hi + flick) + (c->c0 + c->y0 * 0.0);
    }
    int64_t k0 = sat_i64(floor(x));
    double frac = x - (double)k0;
    clock_advance_to(c, k0 + 1);
    double eps = c->eps_lo * (1.0 - frac) + c->eps_hi * frac;
    double flick = flicker_delta_at(c, elapsed);
    return (eps + flick) + (c->c0 + c->y0 * elapsed);
}

/* ---------------------------------------------------------- receiver ops */

/* src/channel/mod.rs:310-312 */
double orc_thermal_noise_power_w(double temperature_k, double bandwidth_hz) {
    return BOLTZMANN_J_PER_K * temperature_k * bandwidth_hz;
}

/* src/rx/mod.rs:83-101. Codes (2 per sample) are the ADC's integer levels
 * round(clamp(x)/lsb), written by the last ADC op of the chain. */
void orc_post_ops(const orc_post_op *ops, uint32_t n_ops, float *iq, uint64_t n, int16_t *codes) {
    for (uint32_t o = 0; o < n_ops; ++o) {
        const orc_post_op *op = &ops[o];
        if (op->kind == ORC_POST_GAIN) {
            float g = powf(10.0f, (float)(op->value / 20.0));
            for (uint64_t i = 0; i < 2 * n; ++i) iq[i] *= g;
        } else if (op->kind == ORC_POST_NONLINEARITY) {
            float k = (float)(4.0 / 3.0 / op->value);
            for (uint64_t i = 0; i < n; ++i) {
                float re = iq[2 * i], im = iq[2 * i + 1];
                float f = 1.0f - k * (re * re + im * im);
                iq[2 * i] = re * f;
                iq[2 * i + 1] = im * f;
            }
        } else if (op->kind == ORC_POST_ADC) {
            float fs = (float)op->value;
            int bits = op->bits < 2 ? 2 : op->bits;
            float levels = (float)((1ULL << (bits - 1)) - 1);
            float lsb = fs / levels;
            for (uint64_t i = 0; i < 2 * n; ++i) {
                float x = iq[i];
                if (x < -fs) x = -fs;
                else if (x > fs) x = fs;
                float q = roundf(x / lsb);
                iq[i] = q * lsb;
                if (codes) codes[i] = (int16_t)q;
            }
        }
    }
}

/* ------------------------------------------------------ per-tx contribution */

/* src/channel/mod.rs:164-271 with no antenna: solve, baseband, mix into `out`
 * (which the caller zeroed). Optional taps for one transmitter. */
/* Antenna::amplitude_at, src/rx/mod.rs:48-58: sqrt(G) at elevation elev_deg. */
float orc_antenna_amplitude(const orc_band *band, double elev_deg) {
    if (band->ant_has_cutoff && elev_deg < band->ant_cutoff_deg) return 0.0f;
    double dbi = elev_deg >= 0.0
        ? band->ant_horizon_dbi + (band->ant_peak_dbi - band->ant_horizon_dbi) * (elev_deg / 90.0)
        : band->ant_horizon_dbi - band->ant_fb_db;
    return (float)pow(10.0, dbi / 20.0);
}

/* Arrival elevation of a source at `src` seen from `obs` against obs's
 * geocentric up, in degrees, src/channel/mod.rs:193-201 (clamp, asin,
 * to_degrees = x * (180 / pi)). */
double orc_arrival_elevation_deg(const double obs[3], const double src[3]) {
    double ox = obs[0], oy = obs[1], oz = obs[2];
    double on = sqrt(ox * ox + oy * oy + oz * oz);
    double dx = src[0] - ox, dy = src[1] - oy, dz = src[2] - oz;
    double dn = sqrt(dx * dx + dy * dy + dz * dz);
    double sin_el = (dx * ox + dy * oy + dz * oz) / (dn * on);
    double c = sin_el < -1.0 ? -1.0 : (sin_el > 1.0 ? 1.0 : sin_el);
    return asin(c) * (180.0 / PI);
}

static void tx_contribution(const orc_tx *tx, const orc_band *band, const double *obs3,
                            const double *t_local, const double *t_utc, const double *delta_t,
                            uint64_t n, double lo, double epoch, double *t_emit, double *ranges,
                            float *bb, float *out, uint8_t *unoccluded, double *tap_t_emit,
                            double *tap_range, double *tap_phase) {
    /* obs3: the receiver position of every sample, [3 i + axis] (src/pipeline.rs:929-930) */
    for (uint64_t i = 0; i < n; ++i) {
        const double *obs = obs3 + 3 * i;
        if (band->converged)
            orc_solve_converged(tx, obs, t_utc[i], epoch, t_utc[i] - EMISSION_INITIAL_LAG_S, 1e-13, 12, &t_emit[i],
                                &ranges[i]);
        else
            orc_solve_emission(tx, obs, t_utc[i], epoch, NEWTON_ITERATIONS, &t_emit[i], &ranges[i]);
    }
    orc_baseband(tx, t_emit, n, bb);
    int any = 0;
    double src_carrier = tx->carrier_hz;
    double wavelength_m = SPEED_OF_LIGHT_M_S / src_carrier;
    double tx_amplitude_v = sqrt(tx->tx_power_w);
    double neg_tau_src = -TAU * src_carrier;
    double tau_if = TAU * (src_carrier - lo);
    double amp_factor = (wavelength_m / (2.0 * TAU)) * tx_amplitude_v;
    const double inv_c = 1.0 / SPEED_OF_LIGHT_M_S;
    for (uint64_t i = 0; i < n; ++i) {
        double r = ranges[i];
        double delay = r * inv_c;
        double phase = neg_tau_src * (delay + delta_t[i]) + tau_if * t_local[i];
        if (tap_t_emit) {
            tap_t_emit[i] = t_emit[i];
            tap_range[i] = r;
            tap_phase[i] = phase;
        }
        /* antenna gain at the solved emission time, src/channel/mod.rs:190-207;
         * an occluded sample leaves the contribution at zero (:253-255) */
        float g = 1.0f;
        if (band->ant_on) {
            double src[3];
            tx_position(tx, epoch + t_emit[i], src);
            g = orc_antenna_amplitude(band, orc_arrival_elevation_deg(obs3 + 3 * i, src));
            if (g == 0.0f) continue;
        }
        any = 1;
        double sp, cp;
        sincos(phase, &sp, &cp);
        float amplitude = (float)(amp_factor / r) * g;
        float bre = bb[2 * i], bim = bb[2 * i + 1];
        float cpf = (float)cp, spf = (float)sp;
        out[2 * i] = out[2 * i] + amplitude * (bre * cpf - bim * spf);
        out[2 * i + 1] = out[2 * i + 1] + amplitude * (bre * spf + bim * cpf);
    }
    if (unoccluded && any) *unoccluded = 1;
}

typedef struct {
    double *t_local, *delta_t, *t_utc, *t_emit, *ranges, *obs;
    float *bb, *contrib, *acc;
} chunk_scratch;

static int scratch_alloc(chunk_scratch *s, uint64_t clen, uint32_t n_active) {
    memset(s, 0, sizeof(*s));
    s->t_local = malloc(clen * sizeof(double));
    s->delta_t = malloc(clen * sizeof(double));
    s->t_utc = malloc(clen * sizeof(double));
    s->obs = malloc(3 * clen * sizeof(double));
    s->t_emit = malloc((n_active ? n_active : 1) * clen * sizeof(double));
    s->ranges = malloc((n_active ? n_active : 1) * clen * sizeof(double));
    s->bb = malloc((n_active ? n_active : 1) * 2 * clen * sizeof(float));
    s->contrib = malloc((n_active ? n_active : 1) * 2 * clen * sizeof(float));
    s->acc = malloc(2 * clen * sizeof(float));
    return s->t_local && s->delta_t && s->t_utc && s->obs && s->t_emit && s->ranges && s->bb && s->contrib &&
           s->acc;
}
static void scratch_free(chunk_scratch *s) {
    free(s->t_local); free(s->delta_t); free(s->t_utc); free(s->obs); free(s->t_emit);
    free(s->ranges); free(s->bb); free(s->contrib); free(s->acc);
}

/* One chunk [i0, i0+clen) of generate_band, src/pipeline.rs:911-983. The
 * transmitter fan-out runs as OpenMP tasks when `parallel_tx` is set (the
 * reference's nested rayon par_iter, :943-964); the sum is always in active
 * order (:965-976). */
static void render_chunk(const orc_tx *txs, const orc_band *band, const orc_rx *rx,
                         orc_clock *clk, float sigma, uint64_t i0, uint64_t clen,
                         chunk_scratch *s, int parallel_tx, uint32_t tap_slot,
                         double *tap_t_emit, double *tap_range, double *tap_phase,
                         double *tap_delta_t) {
    double fs = band->fs_hz;
    for (uint64_t j = 0; j < clen; ++j) s->t_local[j] = (double)(i0 + j) / fs;
    for (uint64_t j = 0; j < clen; ++j) s->delta_t[j] = orc_clock_delta_t_at(clk, s->t_local[j]);
    for (uint64_t j = 0; j < clen; ++j) s->t_utc[j] = s->t_local[j] - s->delta_t[j];
    /* src/pipeline.rs:929-930: positions at the absolute UTC grid epoch + t_utc */
    for (uint64_t j = 0; j < clen; ++j) orc_rx_position(rx, band->epoch_unix + s->t_utc[j], s->obs + 3 * j);
    if (tap_delta_t)
        for (uint64_t j = 0; j < clen; ++j) tap_delta_t[i0 + j] = s->delta_t[j];
    uint32_t na = band->n_active;
    if (parallel_tx && na > 1) {
#pragma omp taskloop grainsize(1)
        for (uint32_t k = 0; k < na; ++k) {
            float *c = s->contrib + (uint64_t)k * 2 * clen;
            memset(c, 0, 2 * clen * sizeof(float));
            tx_contribution(&txs[band->active[k]], band, s->obs, s->t_local, s->t_utc, s->delta_t, clen,
                            band->lo_hz, band->epoch_unix, s->t_emit + (uint64_t)k * clen,
                            s->ranges + (uint64_t)k * clen, s->bb + (uint64_t)k * 2 * clen, c,
                            band->unoccluded ? band->unoccluded + k : NULL, NULL, NULL, NULL);
        }
    } else {
        for (uint32_t k = 0; k < na; ++k) {
            float *c = s->contrib + (uint64_t)k * 2 * clen;
            memset(c, 0, 2 * clen * sizeof(float));
            int tap = tap_t_emit && k == tap_slot;
            tx_contribution(&txs[band->active[k]], band, s->obs, s->t_local, s->t_utc, s->delta_t, clen,
                            band->lo_hz, band->epoch_unix, s->t_emit + (uint64_t)k * clen,
                            s->ranges + (uint64_t)k * clen, s->bb + (uint64_t)k * 2 * clen, c,
                            band->unoccluded ? band->unoccluded + k : NULL,
                            tap ? tap_t_emit + i0 : NULL, tap ? tap_range + i0 : NULL,
                            tap ? tap_phase + i0 : NULL);
        }
    }
    memset(s->acc, 0, 2 * clen * sizeof(float));
    for (uint32_t k = 0; k < na; ++k) {
        const float *c = s->contrib + (uint64_t)k * 2 * clen;
        for (uint64_t j = 0; j < 2 * clen; ++j) s->acc[j] = s->acc[j] + c[j];
    }
    /* src/channel/mod.rs:323-334 */
    if (sigma > 0.0f) {
        for (uint64_t j = 0; j < clen; ++j) {
            float zr, zi;
            orc_awgn_normals(band->noise_seed, i0 + j, &zr, &zi);
            s->acc[2 * j] += 0.0f + sigma * zr;
            s->acc[2 * j + 1] += 0.0f + sigma * zi;
        }
    }
}

int orc_render_band(const orc_tx *txs, uint32_t n_tx, const orc_rx *rx, const orc_band *band,
                    uint64_t n, uint64_t chunk_samples, float *out_iq, int16_t *codes,
                    uint32_t tap_slot, double *tap_t_emit, double *tap_range, double *tap_phase,
                    double *tap_delta_t) {
    if (chunk_samples == 0) chunk_samples = 1;
    for (uint32_t k = 0; k < band->n_active; ++k)
        if (band->active[k] >= n_tx) return -1;
    orc_clock clk;
    orc_clock_init(&clk, rx);
    float sigma = (float)sqrt(band->noise_power_w / 2.0);
    uint64_t clen_max = chunk_samples < n ? chunk_samples : n;
    chunk_scratch s;
    if (!scratch_alloc(&s, clen_max ? clen_max : 1, band->n_active)) {
        scratch_free(&s);
        return -1;
    }
    for (uint64_t i0 = 0; i0 < n; i0 += chunk_samples) {
        uint64_t clen = n - i0 < chunk_samples ? n - i0 : chunk_samples;
        render_chunk(txs, band, rx, &clk, sigma, i0, clen, &s, 0, tap_slot,
                     tap_t_emit, tap_range, tap_phase, tap_delta_t);
        orc_post_ops(band->post, band->n_post, s.acc, clen, codes ? codes + 2 * i0 : NULL);
        memcpy(out_iq + 2 * i0, s.acc, 2 * clen * sizeof(float));
    }
    scratch_free(&s);
    return 0;
}

/* A window of one band, samples [i0, i0 + n), as one chunk of the chunk body
 * (src/pipeline.rs:915-983) with the receiver clock offsets delta_t[0..n)
 * supplied by the caller instead of walked from the band's start. Taps, when
 * given, hold every active slot: tap_*[k * n + j]. Returns 0, or -1 on bad
 * input. */
int orc_render_window(const orc_tx *txs, uint32_t n_tx, const orc_band *band, const orc_rx *rx,
                      uint64_t i0, uint64_t n, const double *delta_t, float *out_iq, int16_t *codes,
                      double *tap_t_emit, double *tap_range, double *tap_phase) {
    for (uint32_t k = 0; k < band->n_active; ++k)
        if (band->active[k] >= n_tx) return -1;
    if (!delta_t || n == 0) return n == 0 ? 0 : -1;
    chunk_scratch s;
    if (!scratch_alloc(&s, n, band->n_active)) {
        scratch_free(&s);
        return -1;
    }
    const double fs = band->fs_hz;
    for (uint64_t j = 0; j < n; ++j) s.t_local[j] = (double)(i0 + j) / fs;   /* src/pipeline.rs:918 */
    for (uint64_t j = 0; j < n; ++j) s.delta_t[j] = delta_t[j];
    for (uint64_t j = 0; j < n; ++j) s.t_utc[j] = s.t_local[j] - s.delta_t[j];   /* :926 */
    for (uint64_t j = 0; j < n; ++j) orc_rx_position(rx, band->epoch_unix + s.t_utc[j], s.obs + 3 * j);   /* :929-930 */
    const uint32_t na = band->n_active;
    for (uint32_t k = 0; k < na; ++k) {
        float *c = s.contrib + (uint64_t)k * 2 * n;
        memset(c, 0, 2 * n * sizeof(float));
        tx_contribution(&txs[band->active[k]], band, s.obs, s.t_local, s.t_utc, s.delta_t, n, band->lo_hz,
                        band->epoch_unix, s.t_emit + (uint64_t)k * n, s.ranges + (uint64_t)k * n,
                        s.bb + (uint64_t)k * 2 * n, c, band->unoccluded ? band->unoccluded + k : NULL,
                        tap_t_emit ? tap_t_emit + (uint64_t)k * n : NULL, tap_range ? tap_range + (uint64_t)k * n : NULL,
                        tap_phase ? tap_phase + (uint64_t)k * n : NULL);
    }
    /* src/pipeline.rs:965-976 */
    memset(s.acc, 0, 2 * n * sizeof(float));
    for (uint32_t k = 0; k < na; ++k) {
        const float *c = s.contrib + (uint64_t)k * 2 * n;
        for (uint64_t j = 0; j < 2 * n; ++j) s.acc[j] = s.acc[j] + c[j];
    }
    const float sigma = (float)sqrt(band->noise_power_w / 2.0);
    if (sigma > 0.0f) {
        for (uint64_t j = 0; j < n; ++j) {
            float zr, zi;
            orc_awgn_normals(band->noise_seed, i0 + j, &zr, &zi);
            s.acc[2 * j] += 0.0f + sigma * zr;
            s.acc[2 * j + 1] += 0.0f + sigma * zi;
        }
    }
    orc_post_ops(band->post, band->n_post, s.acc, n, codes);
    memcpy(out_iq, s.acc, 2 * n * sizeof(float));
    scratch_free(&s);
    return 0;
}

/* --------------------------------------------------- timed CPU baseline */

static uint64_t fnv1a(uint64_t h, const float *iq, uint64_t n_floats) {
    const unsigned char *b = (const unsigned char *)iq;
    for (uint64_t i = 0; i < n_floats * 4; ++i) h = (h ^ b[i]) * 0x100000001B3ULL;
    return h;
}

static void render_band_hashed(const orc_tx *txs, const orc_rx *rx, const orc_band *band,
                               uint64_t n, uint64_t chunk_samples, uint64_t *hash) {
    orc_clock clk;
    orc_clock_init(&clk, rx);
    float sigma = (float)sqrt(band->noise_power_w / 2.0);
    uint64_t clen_max = chunk_samples < n ? chunk_samples : n;
    chunk_scratch s;
    uint64_t h = 0xCBF29CE484222325ULL;
    if (scratch_alloc(&s, clen_max ? clen_max : 1, band->n_active)) {
        for (uint64_t i0 = 0; i0 < n; i0 += chunk_samples) {
            uint64_t clen = n - i0 < chunk_samples ? n - i0 : chunk_samples;
            render_chunk(txs, band, rx, &clk, sigma, i0, clen, &s, 1, 0, NULL, NULL,
                         NULL, NULL);
            orc_post_ops(band->post, band->n_post, s.acc, clen, NULL);
            h = fnv1a(h, s.acc, 2 * clen);
        }
    }
    scratch_free(&s);
    *hash = h;
}

uint64_t orc_render_scene_hashed(const orc_tx *txs, uint32_t n_tx, const orc_rx *rxs,
                                 const orc_band *bands, uint32_t n_bands, const uint64_t *n_limit,
                                 uint64_t chunk_samples, int n_threads, uint64_t *hash_out) {
    (void)n_tx;
    uint64_t contributions = 0;
    for (uint32_t b = 0; b < n_bands; ++b) contributions += n_limit[b] * bands[b].n_active;
    if (chunk_samples == 0) chunk_samples = 1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    /* src/pipeline.rs:513-535: (rx, band) units in parallel; nested tx tasks. */
#pragma omp parallel
#pragma omp single
    {
        for (uint32_t b = 0; b < n_bands; ++b) {
#pragma omp task firstprivate(b)
            render_band_hashed(txs, &rxs[bands[b].rx], &bands[b], n_limit[b], chunk_samples, &hash_out[b]);
        }
#pragma omp taskwait
    }
    return contributions;
}
