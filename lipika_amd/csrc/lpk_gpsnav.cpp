// lpk_gpsnav.cpp -- host-side GPS LNAV and broadcast-ephemeris builders
// (src/signal/gps_nav.rs): GPS time, the broadcast ephemeris of a circular
// Keplerian orbit or of an ECEF state, the IS-GPS-200 subframe 1-3 encoder
// with its parity, the SV-position algorithm, a trajectory grid sampled from
// a broadcast ephemeris, and the RINEX 3 GPS navigation record reader. They
// feed the NAV bit table and the position grid of lpk_tx_desc. Pure host C++,
// compiled with -ffp-contract=off so each f64 operation rounds as the
// reference's Rust does; transcendentals are glibc's.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lpk.h"

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kGpsEpochUnix = 315964800.0;   // src/signal/gps_nav.rs:15
constexpr double kSecondsPerWeek = 604800.0;    // :16
constexpr double kGpsMu = 3.986005e14;          // :18
constexpr double kOmegaEDot = 7.2921151467e-5;  // :20

inline double radians(double deg) { return deg * (kPi / 180.0); }

// f64::rem_euclid
inline double rem_euclid(double a, double b) {
    const double r = std::fmod(a, b);
    return r < 0.0 ? r + std::fabs(b) : r;
}

// src/signal/gps_nav.rs:44
inline double round_to_16s(double sow) { return std::round(sow / 16.0) * 16.0; }

// src/signal/gps_nav.rs:47-52
double wrap_pi(double x) {
    const double twopi = 2.0 * kPi;
    double r = rem_euclid(x + kPi, twopi) - kPi;
    if (r >= kPi) r -= twopi;
    return r;
}

// Rust `f as u32` / `f as u8` / `f as u16`: saturating, NaN -> 0.
uint32_t sat_u(double f, double max) {
    if (!(f == f) || f <= 0.0) return 0;
    if (f >= max) return static_cast<uint32_t>(max);
    return static_cast<uint32_t>(f);
}

// Rust `(phys / lsb).round() as i64` (saturating)
int64_t scaled_i(double phys, double lsb) {
    const double v = std::round(phys / lsb);
    if (!(v == v)) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return static_cast<int64_t>(v);
}
uint64_t scaled_u(double phys, double lsb) {
    const double v = std::round(phys / lsb);
    if (!(v == v) || v <= 0.0) return 0;
    if (v >= 18446744073709551615.0) return UINT64_MAX;
    return static_cast<uint64_t>(v);
}
inline double p2(int n) { return std::ldexp(1.0, n); }
inline double sc(double rad) { return rad / kPi; }  // radians -> semicircles

// The 24 source data bits of one word, most significant first.
struct Word {
    uint8_t b[24];
    int n = 0;
    void u(uint64_t v, int nbits) {
        for (int i = nbits - 1; i >= 0; --i) b[n++] = static_cast<uint8_t>((v >> i) & 1u);
    }
    void s(int64_t v, int nbits) {
        const uint64_t mask = nbits >= 64 ? UINT64_MAX : ((uint64_t(1) << nbits) - 1);
        u(static_cast<uint64_t>(v) & mask, nbits);
    }
};
Word zero_word() {
    Word w;
    w.u(0, 24);
    return w;
}

// IS-GPS-200 Table 20-XIV parity equations over the source bits, with the 24
// data bits at positions 29..6 of a 32-bit word and D29*, D30* at 31, 30
// (src/signal/gps_nav.rs:288-316). For words 2 and 10 (`nib`) data bits 23
// and 24 are chosen so that parity bits 29 and 30 come out zero.
constexpr uint32_t kParity[6] = {0x3B1F3480u, 0x1D8F9A40u, 0x2EC7CD00u, 0x1763E680u, 0x2BB1F340u, 0x0B7A89C0u};

inline uint32_t odd(uint32_t x) { return static_cast<uint32_t>(__builtin_popcount(x)) & 1u; }

uint32_t checksum(uint32_t source, bool nib) {
    uint32_t d = source & 0x3FFFFFC0u;
    const uint32_t d29 = (source >> 31) & 1u, d30 = (source >> 30) & 1u;
    if (nib) {
        if ((d30 + odd(kParity[4] & d)) & 1u) d ^= 1u << 6;
        if ((d29 + odd(kParity[5] & d)) & 1u) d ^= 1u << 7;
    }
    uint32_t w = d;
    if (d30) w ^= 0x3FFFFFC0u;
    w |= ((d29 + odd(kParity[0] & d)) & 1u) << 5;
    w |= ((d30 + odd(kParity[1] & d)) & 1u) << 4;
    w |= ((d29 + odd(kParity[2] & d)) & 1u) << 3;
    w |= ((d30 + odd(kParity[3] & d)) & 1u) << 2;
    w |= ((d30 + odd(kParity[4] & d)) & 1u) << 1;
    w |= (d29 + odd(kParity[5] & d)) & 1u;
    return w & 0x3FFFFFFFu;
}

// One 30-bit transmitted word (bits MSB first) from its 24 source bits and the
// previous word's last two transmitted bits (src/signal/gps_nav.rs:318-333).
void encode_word(const Word &src, uint8_t d29p, uint8_t d30p, bool solve, uint8_t out[30]) {
    uint32_t source = (uint32_t(d29p) << 31) | (uint32_t(d30p) << 30);
    for (int i = 0; i < 24; ++i) source |= uint32_t(src.b[i] & 1u) << (29 - i);
    const uint32_t w = checksum(source, solve);
    for (int i = 0; i < 30; ++i) out[i] = static_cast<uint8_t>((w >> (29 - i)) & 1u);
}

// (8 most significant, 24 least significant) bits of a 32-bit field.
void split_32(int64_t v, uint64_t &hi, uint64_t &lo) {
    const uint64_t u = static_cast<uint64_t>(v) & 0xFFFFFFFFull;
    hi = u >> 24;
    lo = u & 0xFFFFFFull;
}

// Words 3-10 of subframes 1, 2 and 3 (src/signal/gps_nav.rs:390-542).
void subframe_words(const lpk_ephemeris &e, int id, Word w[8]) {
    for (int k = 0; k < 8; ++k) w[k] = Word();
    if (id == 1) {
        const int64_t tgd = scaled_i(e.tgd, p2(-31)), af2 = scaled_i(e.af2, p2(-55)), af1 = scaled_i(e.af1, p2(-43)),
                      af0 = scaled_i(e.af0, p2(-31));
        const uint64_t toc = scaled_u(e.toc_s, 16.0);
        w[0].u(e.week & 0x3FFu, 10);
        w[0].u(0, 2);
        w[0].u(e.ura_index, 4);
        w[0].u(e.health, 6);
        w[0].u((uint64_t(e.iodc) >> 8) & 0x3u, 2);
        w[1] = w[2] = w[3] = zero_word();
        w[4].u(0, 16);
        w[4].s(tgd, 8);
        w[5].u(e.iodc & 0xFFu, 8);
        w[5].u(toc, 16);
        w[6].s(af2, 8);
        w[6].s(af1, 16);
        w[7].s(af0, 22);
        w[7].u(0, 2);
    } else if (id == 2) {
        const int64_t crs = scaled_i(e.crs, 1.0 / 32.0), dn = scaled_i(sc(e.delta_n_rad_s), p2(-43)),
                      m0 = scaled_i(sc(e.m0_rad), p2(-31)), cuc = scaled_i(e.cuc, p2(-29)),
                      ecc = static_cast<int64_t>(scaled_u(e.e, p2(-33))), cus = scaled_i(e.cus, p2(-29)),
                      sqa = static_cast<int64_t>(scaled_u(e.sqrt_a, p2(-19)));
        const uint64_t toe = scaled_u(e.toe_s, 16.0);
        uint64_t m0h, m0l, eh, el, ah, al;
        split_32(m0, m0h, m0l);
        split_32(ecc, eh, el);
        split_32(sqa, ah, al);
        w[0].u(e.iode, 8);
        w[0].s(crs, 16);
        w[1].s(dn, 16);
        w[1].u(m0h, 8);
        w[2].u(m0l, 24);
        w[3].s(cuc, 16);
        w[3].u(eh, 8);
        w[4].u(el, 24);
        w[5].s(cus, 16);
        w[5].u(ah, 8);
        w[6].u(al, 24);
        w[7].u(toe, 16);
        w[7].u(0, 1);
        w[7].u(0, 5);
        w[7].u(0, 2);
    } else if (id == 3) {
        const int64_t cic = scaled_i(e.cic, p2(-29)), omg0 = scaled_i(sc(e.omega0_rad), p2(-31)),
                      cis = scaled_i(e.cis, p2(-29)), i0 = scaled_i(sc(e.i0_rad), p2(-31)),
                      crc = scaled_i(e.crc, 1.0 / 32.0), argp = scaled_i(sc(e.argp_rad), p2(-31)),
                      omgdot = scaled_i(sc(e.omega_dot_rad_s), p2(-43)), idot = scaled_i(sc(e.idot_rad_s), p2(-43));
        uint64_t oh, ol, ih, il, wh, wl;
        split_32(omg0, oh, ol);
        split_32(i0, ih, il);
        split_32(argp, wh, wl);
        w[0].s(cic, 16);
        w[0].u(oh, 8);
        w[1].u(ol, 24);
        w[2].s(cis, 16);
        w[2].u(ih, 8);
        w[3].u(il, 24);
        w[4].s(crc, 16);
        w[4].u(wh, 8);
        w[5].u(wl, 24);
        w[6].s(omgdot, 24);
        w[7].u(e.iode, 8);
        w[7].s(idot, 14);
        w[7].u(0, 2);
    } else {
        for (int k = 0; k < 8; ++k) w[k] = zero_word();
    }
}

// One 300-bit subframe: TLM, HOW (TOW count, subframe id, solved t-bits), then
// words 3-10 with word 10's t-bits solved (src/signal/gps_nav.rs:335-383).
void build_subframe(int id, uint32_t tow_count, const Word data[8], uint8_t &d29p, uint8_t &d30p, uint8_t *out) {
    Word tlm;
    tlm.u(0x8B, 8);
    tlm.u(0, 14);
    tlm.u(0, 2);
    encode_word(tlm, d29p, d30p, false, out);
    Word how;
    how.u(tow_count & 0x1FFFFu, 17);
    how.u(0, 1);
    how.u(0, 1);
    how.u(static_cast<uint64_t>(id), 3);
    how.u(0, 2);
    encode_word(how, out[28], out[29], true, out + 30);
    for (int k = 0; k < 8; ++k) {
        uint8_t *prev = out + 30 * (k + 1);
        encode_word(data[k], prev[28], prev[29], k == 7, out + 30 * (k + 2));
    }
    d29p = out[298];
    d30p = out[299];
}

// Classical osculating elements of an inertial state (src/signal/gps_nav.rs:196-257).
struct Classical {
    double a, e, i, raan, argp, m;
};
Classical classical_elements(const double r[3], const double v[3], double mu) {
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto cross = [](const double *a, const double *b, double *o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    const double rmag = std::sqrt(dot(r, r));
    const double vmag2 = dot(v, v);
    double h[3];
    cross(r, v, h);
    const double hmag = std::sqrt(dot(h, h));
    const double nvec[3] = {-h[1], h[0], 0.0};
    const double nmag = std::sqrt(nvec[0] * nvec[0] + nvec[1] * nvec[1]);
    double vxh[3];
    cross(v, h, vxh);
    const double evec[3] = {vxh[0] / mu - r[0] / rmag, vxh[1] / mu - r[1] / rmag, vxh[2] / mu - r[2] / rmag};
    const double e = std::sqrt(dot(evec, evec));
    const double energy = vmag2 / 2.0 - mu / rmag;
    const double a = -mu / (2.0 * energy);
    const double inc = std::acos(h[2] / hmag);
    double raan = nmag > 1e-12 ? std::acos(nvec[0] / nmag) : 0.0;
    if (nvec[1] < 0.0) raan = 2.0 * kPi - raan;
    auto clamp1 = [](double x) { return x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x); };
    double argp = (nmag > 1e-12 && e > 1e-12) ? std::acos(clamp1(dot(nvec, evec) / (nmag * e))) : 0.0;
    if (evec[2] < 0.0) argp = 2.0 * kPi - argp;
    double nu;
    if (e > 1e-12) {
        nu = std::acos(clamp1(dot(evec, r) / (e * rmag)));
    } else {
        double u = nmag > 1e-12 ? std::acos(clamp1(dot(nvec, r) / (nmag * rmag))) : std::acos(r[0] / rmag);
        if (r[2] < 0.0) u = 2.0 * kPi - u;
        nu = u;
    }
    if (e > 1e-12 && dot(r, v) < 0.0) nu = 2.0 * kPi - nu;
    const double big_e = 2.0 * std::atan2(std::sqrt(1.0 - e) * std::sin(nu / 2.0), std::sqrt(1.0 + e) * std::cos(nu / 2.0));
    const double m = big_e - e * std::sin(big_e);
    return Classical{a, e, inc, raan, argp, m};
}

lpk_ephemeris blank(uint32_t prn, uint32_t week) {
    lpk_ephemeris e;
    std::memset(&e, 0, sizeof(e));
    e.prn = prn;
    e.week = week;
    return e;
}

// A RINEX float field: 19 columns from `start`, D exponents read as E, a
// missing, short or unparsable field as 0 (src/signal/gps_nav.rs:656-664).
double rinex_field(const std::string &line, size_t start) {
    if (start >= line.size()) return 0.0;
    const size_t end = start + 19 < line.size() ? start + 19 : line.size();
    std::string f = line.substr(start, end - start);
    size_t a = 0, b = f.size();
    while (a < b && std::isspace(static_cast<unsigned char>(f[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(f[b - 1]))) --b;
    f = f.substr(a, b - a);
    if (f.empty()) return 0.0;
    for (char &c : f) {
        if (c == 'D' || c == 'd') c = 'E';
        if (c == 'x' || c == 'X') return 0.0;  // no hexadecimal floats
    }
    char *endp = nullptr;
    const double v = std::strtod(f.c_str(), &endp);
    return (endp && *endp == '\0') ? v : 0.0;
}

}  // namespace

extern "C" {

void lpk_gps_time(double unix_s, double utc_offset_s, uint32_t *week, double *sow) {
    const double gps = unix_s - kGpsEpochUnix + utc_offset_s;
    const double w = std::floor(gps / kSecondsPerWeek);
    if (week) *week = sat_u(w, 4294967295.0);
    if (sow) *sow = gps - w * kSecondsPerWeek;
}

int lpk_toe_snap(double unix_s, double utc_offset_s, double *unix_at_toe, uint32_t *week, double *toe_sow) {
    if (!unix_at_toe || !week || !toe_sow) return LPK_E_INVALID;
    double sow;
    lpk_gps_time(unix_s, utc_offset_s, week, &sow);
    *toe_sow = round_to_16s(sow);
    *unix_at_toe = unix_s + (*toe_sow - sow);
    return LPK_OK;
}

int lpk_ephemeris_from_kepler(double a_m, double inc_deg, double raan_deg, double nu0_deg, double orbit_epoch_unix,
                              uint32_t prn, double scene_epoch_unix, double utc_offset_s, lpk_ephemeris *out) {
    if (!out) return LPK_E_INVALID;
    uint32_t week;
    double sow;
    lpk_gps_time(scene_epoch_unix, utc_offset_s, &week, &sow);
    const double n = std::sqrt(kGpsMu / (a_m * a_m * a_m));
    const double toe = round_to_16s(sow);
    const double dt0 = scene_epoch_unix - orbit_epoch_unix;
    const double dt_toe = dt0 + (toe - sow);
    lpk_ephemeris e = blank(prn, week);
    e.toe_s = toe;
    e.toc_s = toe;
    e.sqrt_a = std::sqrt(a_m);
    e.m0_rad = wrap_pi(radians(nu0_deg) + n * dt_toe);
    e.i0_rad = radians(inc_deg);
    e.omega0_rad = wrap_pi(radians(raan_deg) + kOmegaEDot * sow - kOmegaEDot * dt0);
    *out = e;
    return LPK_OK;
}

int lpk_ephemeris_from_state(const double r[3], const double v[3], uint32_t prn, uint32_t week, double toe_sow,
                             lpk_ephemeris *out) {
    if (!r || !v || !out) return LPK_E_INVALID;
    const double phi = kOmegaEDot * toe_sow;
    const double s = std::sin(phi), c = std::cos(phi);
    const double ri[3] = {r[0] * c - r[1] * s, r[0] * s + r[1] * c, r[2]};
    const double vr[3] = {v[0] - kOmegaEDot * r[1], v[1] + kOmegaEDot * r[0], v[2]};
    const double vi[3] = {vr[0] * c - vr[1] * s, vr[0] * s + vr[1] * c, vr[2]};
    const Classical el = classical_elements(ri, vi, kGpsMu);
    lpk_ephemeris e = blank(prn, week);
    e.toe_s = toe_sow;
    e.toc_s = toe_sow;
    e.sqrt_a = std::sqrt(el.a);
    e.e = el.e;
    e.m0_rad = wrap_pi(el.m);
    e.i0_rad = el.i;
    e.omega0_rad = wrap_pi(el.raan);
    e.argp_rad = wrap_pi(el.argp);
    *out = e;
    return LPK_OK;
}

uint64_t lpk_lnav_length(double duration_s) {
    if (!(duration_s >= 0.0)) return 0;
    return 300u * (uint64_t(std::ceil(duration_s / 6.0)) + 1u);
}

int lpk_lnav_bits(const lpk_ephemeris *eph, double sow, double duration_s, int8_t *out, uint64_t n) {
    if (!eph || !out || n != lpk_lnav_length(duration_s) || n == 0) return LPK_E_INVALID;
    const uint32_t base_z = static_cast<uint32_t>(std::round(sow / 6.0));
    const uint64_t n_sub = n / 300u;
    uint8_t d29p = 0, d30p = 0;
    uint8_t sf[300];
    Word words[8];
    for (uint64_t g = 0; g < n_sub; ++g) {
        const int id = static_cast<int>(g % 5) + 1;
        subframe_words(*eph, id, words);
        build_subframe(id, base_z + static_cast<uint32_t>(g) + 1u, words, d29p, d30p, sf);
        for (int k = 0; k < 300; ++k) out[300 * g + k] = static_cast<int8_t>(1 - 2 * int(sf[k]));
    }
    return LPK_OK;
}

void lpk_sv_position(const lpk_ephemeris *eph, double t_sow, double out[3]) {
    const lpk_ephemeris &e = *eph;
    const double a = e.sqrt_a * e.sqrt_a;
    const double n = std::sqrt(kGpsMu / (a * a * a)) + e.delta_n_rad_s;
    double tk = t_sow - e.toe_s;
    if (tk > 302400.0) tk -= kSecondsPerWeek;
    if (tk < -302400.0) tk += kSecondsPerWeek;
    const double mk = e.m0_rad + n * tk;
    double ek = mk;
    for (int k = 0; k < 10; ++k) ek = mk + e.e * std::sin(ek);
    const double nu = std::atan2(std::sqrt(1.0 - e.e * e.e) * std::sin(ek), std::cos(ek) - e.e);
    const double phi = nu + e.argp_rad;
    const double s2 = std::sin(2.0 * phi), c2 = std::cos(2.0 * phi);
    const double du = e.cus * s2 + e.cuc * c2;
    const double dr = e.crs * s2 + e.crc * c2;
    const double di = e.cis * s2 + e.cic * c2;
    const double u = phi + du;
    const double r = a * (1.0 - e.e * std::cos(ek)) + dr;
    const double inc = e.i0_rad + di + e.idot_rad_s * tk;
    const double xp = r * std::cos(u), yp = r * std::sin(u);
    const double om = e.omega0_rad + (e.omega_dot_rad_s - kOmegaEDot) * tk - kOmegaEDot * e.toe_s;
    const double so = std::sin(om), co = std::cos(om), si = std::sin(inc), ci = std::cos(inc);
    out[0] = xp * co - yp * ci * so;
    out[1] = xp * so + yp * ci * co;
    out[2] = yp * si;
}

int lpk_grid_fill_broadcast(const lpk_ephemeris *eph, double utc_offset_s, double scene_epoch_unix,
                            double duration_s, double grid_dt_s, double *t0_out, double *xyz, uint64_t n) {
    if (!eph || !(grid_dt_s > 0.0) || !t0_out || !xyz) return LPK_E_INVALID;
    if (n != lpk_grid_size(duration_s, grid_dt_s)) return LPK_E_INVALID;
    // GpsBroadcastTrajectory (src/signal/gps_nav.rs:630-654) on the grid of
    // PrecomputedTrajectory::sample (src/geometry/precomputed.rs:47-70)
    const double toe_unix = kGpsEpochUnix + double(eph->week) * kSecondsPerWeek + eph->toe_s - utc_offset_s;
    const double start = scene_epoch_unix - 0.5;
    for (uint64_t i = 0; i < n; ++i) {
        const double t = start + double(i) * grid_dt_s;
        double p[3];
        lpk_sv_position(eph, eph->toe_s + (t - toe_unix), p);
        xyz[i] = p[0];
        xyz[n + i] = p[1];
        xyz[2 * n + i] = p[2];
    }
    *t0_out = start;
    return LPK_OK;
}

int lpk_rinex3_gps_ephemeris(const char *text, uint32_t prn, double near_unix, double utc_offset_s,
                             lpk_ephemeris *out) {
    if (!text || !out) return LPK_E_INVALID;
    // str::lines: split at '\n', a trailing '\r' removed from each line
    std::vector<std::string> lines;
    const char *p = text;
    while (*p) {
        const char *q = std::strchr(p, '\n');
        std::string l = q ? std::string(p, q - p) : std::string(p);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        lines.push_back(l);
        if (!q) break;
        p = q + 1;
    }
    size_t i = 0;
    for (size_t k = 0; k < lines.size(); ++k)
        if (lines[k].find("END OF HEADER") != std::string::npos) {
            i = k + 1;
            break;
        }
    bool found = false;
    double best = 0.0;
    while (i + 7 < lines.size()) {
        const std::string &l0 = lines[i];
        if (l0.size() < 3 || l0[0] != 'G') {
            ++i;
            continue;
        }
        // "Gnn": the PRN as u8 (a bad field reads as 0)
        std::string ps = l0.substr(1, 2);
        size_t a = 0, b = ps.size();
        while (a < b && std::isspace(static_cast<unsigned char>(ps[a]))) ++a;
        while (b > a && std::isspace(static_cast<unsigned char>(ps[b - 1]))) --b;
        ps = ps.substr(a, b - a);
        if (!ps.empty() && ps[0] == '+') ps = ps.substr(1);
        uint32_t rec_prn = 0;
        bool ok = !ps.empty();
        for (char c : ps) ok = ok && c >= '0' && c <= '9';
        if (ok) rec_prn = static_cast<uint32_t>(std::strtoul(ps.c_str(), nullptr, 10));
        if (rec_prn > 255) rec_prn = 0;
        if (rec_prn != prn) {
            i += 8;
            continue;
        }
        const std::string &l1 = lines[i + 1], &l2 = lines[i + 2], &l3 = lines[i + 3], &l4 = lines[i + 4],
                          &l5 = lines[i + 5], &l6 = lines[i + 6];
        const double toe = rinex_field(l3, 4), week = rinex_field(l5, 42);
        lpk_ephemeris e = blank(prn, sat_u(week, 4294967295.0));
        e.toe_s = toe;
        e.toc_s = toe;
        e.af0 = rinex_field(l0, 23);
        e.af1 = rinex_field(l0, 42);
        e.af2 = rinex_field(l0, 61);
        e.iode = sat_u(rinex_field(l1, 4), 255.0);
        e.crs = rinex_field(l1, 23);
        e.delta_n_rad_s = rinex_field(l1, 42);
        e.m0_rad = rinex_field(l1, 61);
        e.cuc = rinex_field(l2, 4);
        e.e = rinex_field(l2, 23);
        e.cus = rinex_field(l2, 42);
        e.sqrt_a = rinex_field(l2, 61);
        e.cic = rinex_field(l3, 23);
        e.omega0_rad = rinex_field(l3, 42);
        e.cis = rinex_field(l3, 61);
        e.i0_rad = rinex_field(l4, 4);
        e.crc = rinex_field(l4, 23);
        e.argp_rad = rinex_field(l4, 42);
        e.omega_dot_rad_s = rinex_field(l4, 61);
        e.idot_rad_s = rinex_field(l5, 4);
        e.health = sat_u(rinex_field(l6, 23), 255.0);
        e.tgd = rinex_field(l6, 42);
        e.iodc = sat_u(rinex_field(l6, 61), 65535.0);
        const double toe_unix = kGpsEpochUnix + week * kSecondsPerWeek + toe - utc_offset_s;
        const double d = std::fabs(toe_unix - near_unix);
        if (!found || d < best) {
            found = true;
            best = d;
            *out = e;
        }
        i += 8;
    }
    return found ? LPK_OK : LPK_E_INVALID;
}

}  // extern "C"
