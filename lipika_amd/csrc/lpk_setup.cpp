// lpk_setup.cpp -- host-side scene setup: the tables and constants that
// Transmitter::from_config (src/pipeline.rs:312-341) builds once per scene.
// Pure host C++, compiled with -ffp-contract=off so each f64 operation rounds
// as the reference's Rust does; transcendentals are glibc's, as in the
// reference.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/lpk.h"

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kTau = 6.283185307179586;
constexpr double kGM = 3.986005e14;                  // src/geometry/kepler.rs:12
constexpr double kOmegaE = 7.2921151467e-5;          // src/geometry/kepler.rs:13
constexpr double kWgs84A = 6378137.0;                // src/geometry/ecef.rs:3
constexpr double kWgs84F = 1.0 / 298.257223563;      // src/geometry/ecef.rs:4
constexpr double kWgs84E2 = kWgs84F * (2.0 - kWgs84F);

// f64::to_radians multiplies by the once-rounded PI/180.
inline double radians(double deg) { return deg * (kPi / 180.0); }

int64_t floor_div_rem(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

// SplitMix64 seeding + xoshiro256++ (rand_xoshiro 0.6, seed_from_u64).
struct Xoshiro256pp {
    uint64_t s[4];
    explicit Xoshiro256pp(uint64_t seed) {
        uint64_t x = seed;
        for (auto &w : s) {
            x += 0x9E3779B97F4A7C15ULL;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t out = rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    // rand 0.8 Rng::gen_bool(0.5): Bernoulli with p_int = 2^63.
    bool coin() { return next() < 0x8000000000000000ULL; }
    // rand 0.8 Rng::gen_range(0..16u8): next_u32 (upper half) widened-multiplied by 16.
    unsigned nibble() { return unsigned(next() >> 60); }
};

// src/signal/qam16.rs:20-34
float gray(unsigned b0, unsigned b1) {
    if (!b0 && !b1) return -3.0f;
    if (!b0 && b1) return -1.0f;
    if (b0 && b1) return 1.0f;
    return 3.0f;
}
void qam_point(unsigned b0, unsigned b1, unsigned b2, unsigned b3, float *iq) {
    const float scale = 0.31622777f;  // src/signal/qam16.rs:17
    iq[0] = gray(b0, b1) * scale;
    iq[1] = gray(b2, b3) * scale;
}

}  // namespace

extern "C" {

void lpk_geodetic_to_ecef(double lat_deg, double lon_deg, double h_m, double out[3]) {
    const double lat = radians(lat_deg), lon = radians(lon_deg);
    double sl, cl, slon, clon;
    ::sincos(lat, &sl, &cl);
    ::sincos(lon, &slon, &clon);
    const double n = kWgs84A / std::sqrt(1.0 - kWgs84E2 * sl * sl);
    out[0] = (n + h_m) * cl * clon;
    out[1] = (n + h_m) * cl * slon;
    out[2] = (n * (1.0 - kWgs84E2) + h_m) * sl;
}

uint64_t lpk_grid_size(double duration_s, double grid_dt_s) {
    // src/geometry/precomputed.rs:55-58
    const double total = duration_s + 2.0 * 0.5;
    return uint64_t(std::ceil(total / grid_dt_s)) + 1;
}

int lpk_grid_fill_kepler(double a_m, double inc_deg, double raan_deg, double nu0_deg,
                         double orbit_epoch_unix, double scene_epoch_unix, double duration_s,
                         double grid_dt_s, double *t0_out, double *xyz, uint64_t n) {
    if (!(grid_dt_s > 0.0) || !t0_out || !xyz) return LPK_E_INVALID;
    if (n != lpk_grid_size(duration_s, grid_dt_s)) return LPK_E_INVALID;
    const double start = scene_epoch_unix - 0.5;
    // CircularKeplerian::eci_at + eci_to_ecef, src/geometry/kepler.rs:27-85
    const double mean_motion = std::sqrt(kGM / (a_m * a_m * a_m));
    const double nu0 = radians(nu0_deg);
    double si, ci, so, co;
    ::sincos(radians(inc_deg), &si, &ci);
    ::sincos(radians(raan_deg), &so, &co);
    for (uint64_t i = 0; i < n; ++i) {
        const double t = start + double(i) * grid_dt_s;
        const double dt = t - orbit_epoch_unix;
        const double nu = nu0 + mean_motion * dt;
        double snu, cnu;
        ::sincos(nu, &snu, &cnu);
        const double xp = a_m * cnu, yp = a_m * snu;
        const double y1 = yp * ci, z1 = yp * si;
        const double xe = xp * co - y1 * so;
        const double ye = xp * so + y1 * co;
        const double th = kOmegaE * (t - orbit_epoch_unix);
        double s, c;
        ::sincos(th, &s, &c);
        xyz[i] = xe * c + ye * s;
        xyz[n + i] = -xe * s + ye * c;
        xyz[2 * n + i] = z1;
    }
    *t0_out = start;
    return LPK_OK;
}

void lpk_kepler_position(double a_m, double inc_deg, double raan_deg, double nu0_deg, double orbit_epoch_unix,
                         double t_unix, double out[3]) {
    // CircularKeplerian::position_ecef, src/geometry/kepler.rs:27-85
    const double mean_motion = std::sqrt(kGM / (a_m * a_m * a_m));
    const double nu = radians(nu0_deg) + mean_motion * (t_unix - orbit_epoch_unix);
    double snu, cnu, si, ci, so, co, s, c;
    ::sincos(nu, &snu, &cnu);
    ::sincos(radians(inc_deg), &si, &ci);
    ::sincos(radians(raan_deg), &so, &co);
    const double xp = a_m * cnu, yp = a_m * snu;
    const double y1 = yp * ci, z1 = yp * si;
    const double xe = xp * co - y1 * so;
    const double ye = xp * so + y1 * co;
    ::sincos(kOmegaE * (t_unix - orbit_epoch_unix), &s, &c);
    out[0] = xe * c + ye * s;
    out[1] = -xe * s + ye * c;
    out[2] = z1;
}

int lpk_ca_code(int prn, int8_t out[1023]) {
    // IS-GPS-200 Table 3-Ia G2 tap pairs, src/signal/ca_code.rs:35-40
    static const uint8_t taps[32][2] = {
        {2, 6}, {3, 7}, {4, 8}, {5, 9}, {1, 9}, {2, 10}, {1, 8}, {2, 9}, {3, 10},
        {2, 3}, {3, 4}, {5, 6}, {6, 7}, {7, 8}, {8, 9}, {9, 10}, {1, 4}, {2, 5},
        {3, 6}, {4, 7}, {5, 8}, {6, 9}, {1, 3}, {4, 6}, {5, 7}, {6, 8}, {7, 9},
        {8, 10}, {1, 6}, {2, 7}, {3, 8}, {4, 9}};
    if (prn < 1 || prn > 32 || !out) return LPK_E_INVALID;
    // G1 = 1 + x^3 + x^10, G2 = 1 + x^2 + x^3 + x^6 + x^8 + x^9 + x^10, both
    // all-ones at the start; registers held as 10-bit words, stage k at bit k.
    unsigned g1 = 0x3FF, g2 = 0x3FF;
    const unsigned ta = taps[prn - 1][0] - 1, tb = taps[prn - 1][1] - 1;
    for (int k = 0; k < 1023; ++k) {
        const unsigned o1 = (g1 >> 9) & 1;
        const unsigned o2 = ((g2 >> ta) ^ (g2 >> tb)) & 1;
        const unsigned f1 = ((g1 >> 2) ^ (g1 >> 9)) & 1;
        const unsigned f2 = ((g2 >> 1) ^ (g2 >> 2) ^ (g2 >> 5) ^ (g2 >> 7) ^ (g2 >> 8) ^ (g2 >> 9)) & 1;
        g1 = ((g1 << 1) | f1) & 0x3FF;
        g2 = ((g2 << 1) | f2) & 0x3FF;
        out[k] = int8_t(1 - 2 * int(o1 ^ o2));
    }
    return LPK_OK;
}

uint64_t lpk_nav_fold_seed(uint64_t nav_seed, uint64_t prn) {
    return nav_seed + prn * 0xD1B54A32D192ED03ULL;
}

int lpk_payload_bpsk(const int8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out) {
    if ((n_sync && !sync) || !out) return LPK_E_INVALID;
    std::memcpy(out, sync, n_sync);
    Xoshiro256pp rng(seed);
    for (uint64_t i = 0; i < n_payload; ++i) out[n_sync + i] = rng.coin() ? 1 : -1;
    return LPK_OK;
}

int lpk_payload_ook(const uint8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out) {
    if ((n_sync && !sync) || !out) return LPK_E_INVALID;
    for (uint64_t i = 0; i < n_sync; ++i) {
        if (sync[i] > 1) return LPK_E_INVALID;  // src/signal/ook.rs:35-37
        out[i] = int8_t(sync[i]);
    }
    Xoshiro256pp rng(seed);
    for (uint64_t i = 0; i < n_payload; ++i) out[n_sync + i] = rng.coin() ? 1 : 0;
    return LPK_OK;
}

int lpk_payload_qam16(const uint8_t *sync_bits, uint64_t n_sync_bits, uint64_t n_payload,
                      uint64_t seed, float *out_iq) {
    if ((n_sync_bits && !sync_bits) || !out_iq) return LPK_E_INVALID;
    if (n_sync_bits % 4) return LPK_E_INVALID;  // src/signal/qam16.rs:70-76
    for (uint64_t i = 0; i < n_sync_bits; ++i)
        if (sync_bits[i] > 1) return LPK_E_INVALID;
    const uint64_t ns = n_sync_bits / 4;
    for (uint64_t k = 0; k < ns; ++k)
        qam_point(sync_bits[4 * k], sync_bits[4 * k + 1], sync_bits[4 * k + 2], sync_bits[4 * k + 3],
                  out_iq + 2 * k);
    Xoshiro256pp rng(seed);
    for (uint64_t i = 0; i < n_payload; ++i) {
        const unsigned v = rng.nibble();
        qam_point((v >> 3) & 1, (v >> 2) & 1, (v >> 1) & 1, v & 1, out_iq + 2 * (ns + i));
    }
    return LPK_OK;
}

int lpk_elapsed_periods(double elapsed_s, double rate_hz, int64_t *whole_out, double *frac_out) {
    // src/signal/mod.rs:50-69: the exact product is p + err (FMA residual).
    const double p = elapsed_s * rate_hz;
    if (!std::isfinite(p) || std::fabs(p) >= 9.0e18) return LPK_E_NONFINITE;
    const double err = std::fma(elapsed_s, rate_hz, -p);
    double whole = std::floor(p);
    double frac = (p - whole) + err;
    if (frac < 0.0) {
        whole -= 1.0;
        frac += 1.0;
    } else if (frac >= 1.0) {
        whole += 1.0;
        frac -= 1.0;
    }
    *whole_out = int64_t(whole);
    *frac_out = frac;
    return LPK_OK;
}

int lpk_frame_anchor(double render_epoch, double signal_epoch, double rate_hz, uint64_t n_symbols,
                     double *epoch_rel_s, int64_t *symbol_offset) {
    // src/signal/mod.rs:87-101
    if (!(rate_hz > 0.0)) return LPK_E_INVALID;
    int64_t whole;
    double frac;
    const int rc = lpk_elapsed_periods(render_epoch - signal_epoch, rate_hz, &whole, &frac);
    if (rc) return rc;
    *epoch_rel_s = -frac / rate_hz;
    *symbol_offset = floor_div_rem(whole, n_symbols ? int64_t(n_symbols) : 1);
    return LPK_OK;
}

int lpk_cw_anchor_phase(double render_epoch, double signal_epoch, double offset_hz, double *phase) {
    // src/signal/cw.rs:26-34
    int64_t whole;
    double frac;
    const int rc = lpk_elapsed_periods(render_epoch - signal_epoch, offset_hz, &whole, &frac);
    if (rc) return rc;
    *phase = kTau * frac;
    return LPK_OK;
}

int lpk_bands_overlap(double tx_carrier_hz, double tx_bw_hz, double band_lo_hz, double band_bw_hz) {
    return std::fabs(tx_carrier_hz - band_lo_hz) <= 0.5 * (tx_bw_hz + band_bw_hz);
}

double lpk_thermal_noise_power_w(double temperature_k, double bandwidth_hz) {
    return 1.380649e-23 * temperature_k * bandwidth_hz;  // src/channel/mod.rs:301-312
}

}  // extern "C"
