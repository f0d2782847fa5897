/*
 * lpk_oracle.h -- CPU FP64 oracle for the per-chunk IQ path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline. The product (lipika_amd/liblpk.so) never
 * links, loads or calls it.
 *
 * The oracle restates, in plain C with IEEE f64/f32 arithmetic and no FMA
 * contraction (built with -ffp-contract=off, no fast-math), the reference's
 * f64 chunk pipeline:
 *   src/pipeline.rs:836-1019        generate_band (chunk loop, fan-out, ordered sum)
 *   src/channel/mod.rs:49-101       emission-time solve with Sagnac rotation
 *   src/channel/mod.rs:164-271      per-transmitter mix (amplitude, phase, f32 mix)
 *   src/channel/mod.rs:301-334      kTB and AWGN
 *   src/geometry/{ecef,kepler,precomputed}.rs geodetic, circular Kepler, 10 ms grid lerp
 *   src/signal/{cw,bpsk,ook,qam16,ca_code,gate}.rs CW, BPSK, OOK, 16-QAM, C/A (+NAV), burst gate
 *   src/clock/mod.rs:247-346        ClockStream
 *   src/rx/mod.rs:83-101            gain, IP3, ADC
 * Each function in lpk_oracle.c cites the file:line it follows.
 *
 * Random streams: the reference draws payload bits through rand 0.8 /
 * rand_xoshiro 0.6 (restated here: xoshiro256++ seeded by SplitMix64), and the
 * clock walk and AWGN through rand_distr 0.4's ziggurat normals, which are not
 * under /root/reference and are not reproduced. The clock walk and AWGN use
 * Philox4x32-10 (Salmon et al., SC'11), the flicker stream Philox4x64-10, each
 * with Box-Muller normals, the same generators the product uses, and are
 * verified statistically against the reference's tests.
 *
 * The checks that pin each function are in tests/ (test_oracle_pins.py and the
 * per-subsystem tests); a function whose only check is self-consistency is
 * called "parity unpinned" in the project's design notes.
 */
#ifndef LPK_ORACLE_H
#define LPK_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- signal / trajectory kinds (same numbering as include/lpk.h) ---- */
enum { ORC_SIG_CW = 0, ORC_SIG_BPSK = 1, ORC_SIG_OOK = 2, ORC_SIG_QAM16 = 3, ORC_SIG_CA = 4 };
enum { ORC_NAV_OFF = 0, ORC_NAV_DUMMY = 1, ORC_NAV_FULL = 2 };
enum { ORC_TRAJ_STATIONARY = 0, ORC_TRAJ_GRID = 1 };
enum { ORC_POST_GAIN = 0, ORC_POST_NONLINEARITY = 1, ORC_POST_ADC = 2 };

typedef struct {
    int32_t traj_kind;
    int32_t signal_kind;
    double pos_ecef[3];          /* stationary */
    double grid_t0_unix, grid_dt_s;
    uint64_t grid_n;
    const double *grid_xyz;      /* (3, n) row-major */
    double carrier_hz, tx_power_w, bandwidth_hz;
    double baseband_offset_hz;   /* CW */
    int32_t cw_anchored, _p0;
    double anchor_phase_rad;
    double rate_hz;              /* framed: bit/symbol rate */
    double epoch_rel_s;          /* framed symbol 0 / C/A chip 0 tag, scene-epoch relative */
    int32_t repeating, _p1;
    int64_t repeat_from;
    uint64_t n_symbols;
    const int8_t *bits;          /* BPSK (+-1) / OOK (0/1) */
    const float *symbols;        /* QAM16 interleaved re, im */
    int32_t nav_mode, prn;
    const int8_t *ca_code;       /* 1023 chips, +-1 */
    uint64_t nav_seed;           /* folded dummy-NAV seed */
    const int8_t *nav_bits;      /* full NAV, +-1 */
    uint64_t n_nav_bits;
    int32_t gated;
    uint32_t n_bursts;
    double gate_period_s, gate_phase_s;
    const double *bursts;        /* (start, end) pairs */
} orc_tx;

typedef struct {
    int32_t kind, bits;
    double value;                /* gain dB / iip3 W / ADC full-scale V */
} orc_post_op;

typedef struct {
    double pos_ecef[3];
    double allan_deviation_1s, frequency_offset_y0, initial_time_offset_s;
    double master_rate_hz;
    uint64_t clock_seed;
    double flicker_floor;        /* flicker-FM Allan floor, src/clock/mod.rs:26-30 (0 = none) */
    /* trajectory, src/pipeline.rs:352-364: 0 fixed at pos_ecef, 1 CircularKeplerian
     * (src/geometry/kepler.rs:19-96) queried per sample at epoch + t_utc (:929-930) */
    int32_t traj_kind, _p;
    double kep_a_m, kep_inc_deg, kep_raan_deg, kep_nu0_deg, kep_epoch_unix;
} orc_rx;

/* Receiver position at absolute UTC time t_unix (Trajectory::position_ecef of
 * the receiver's trajectory). */
void orc_rx_position(const orc_rx *rx, double t_unix, double out[3]);

typedef struct {
    uint32_t rx;
    uint32_t n_active;
    const uint32_t *active;      /* Sigma order */
    double fs_hz, lo_hz, epoch_unix, duration_s;
    double noise_power_w;
    uint64_t noise_seed;
    uint32_t n_post, _p;
    orc_post_op post[4];
    /* receiver antenna, src/rx/mod.rs:23-58 (ant_on = 0: none) */
    int32_t ant_on, ant_has_cutoff;
    double ant_peak_dbi, ant_horizon_dbi, ant_fb_db, ant_cutoff_deg;
    uint8_t *unoccluded;         /* optional output [n_active]: ORed over the rendered chunks */
    /* 1: the converged solve (orc_solve_converged from t - 0.07 s, tolerance
     * 1e-13 s, at most 12 evaluations) in place of the reference's two iterations */
    int32_t converged, _p2;
} orc_band;

/* ---- antenna (src/rx/mod.rs:48-58, src/channel/mod.rs:190-207) ---- */
float orc_antenna_amplitude(const orc_band *band, double elev_deg);
double orc_arrival_elevation_deg(const double obs[3], const double src[3]);

/* ---- geometry ---- */
void orc_geodetic_to_ecef(double lat_deg, double lon_deg, double h_m, double out[3]);
void orc_kepler_position(double a_m, double inc_deg, double raan_deg, double nu0_deg,
                         double epoch_unix, double t_unix, double out[3]);
uint64_t orc_grid_size(double duration_s, double grid_dt_s);
/* Fills a (3, n) grid; returns n and writes the grid start time. */
uint64_t orc_grid_sample_kepler(double a_m, double inc_deg, double raan_deg, double nu0_deg,
                                double orbit_epoch_unix, double t0_unix, double duration_s,
                                double grid_dt_s, double *t0_out, double *xyz_out);
void orc_grid_position(double t0_unix, double dt_s, uint64_t n, const double *xyz,
                       double t_unix, double out[3]);
int orc_solve_converged(const orc_tx *tx, const double obs[3], double tu, double epoch, double t_start,
                        double tol, int max_evals, double *t_emit, double *range);
void orc_solve_emission(const orc_tx *tx, const double obs[3], double t_utc_rel, double epoch,
                        int iterations, double *t_emit, double *range);

/* ---- signals ---- */
int orc_ca_code(int prn, int8_t out[1023]);
int8_t orc_nav_dummy_bit_pm(uint64_t seed, int64_t bit_idx);
uint64_t orc_nav_fold_seed(uint64_t nav_seed, uint64_t prn);
int orc_elapsed_periods(double elapsed_s, double rate_hz, int64_t *whole, double *frac);
int orc_frame_anchor(double render_epoch, double signal_epoch, double rate_hz, uint64_t n_symbols,
                     double *epoch_rel_s, int64_t *symbol_offset);
int orc_cw_anchor_phase(double render_epoch, double signal_epoch, double offset_hz, double *phase);
int orc_gate_is_on(const orc_tx *tx, double t);
void orc_baseband(const orc_tx *tx, const double *t_emit, uint64_t n, float *out_iq);
int orc_bands_overlap(double tx_carrier, double tx_bw, double band_lo, double band_bw);

/* ---- reference RNG restatement (payload tables) ---- */
void orc_xoshiro_seed(uint64_t seed, uint64_t s[4]);
uint64_t orc_xoshiro_next(uint64_t s[4]);
void orc_payload_bpsk(const int8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out);
int orc_payload_ook(const uint8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out);
int orc_payload_qam16(const uint8_t *sync, uint64_t n_sync_bits, uint64_t n_payload, uint64_t seed, float *out_iq);

/* ---- project RNG: Philox4x64-10, Philox4x32-10 and the normals built on them ---- */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_box_muller(uint64_t w, float *zc, float *zs);
double orc_clock_normal(uint64_t clock_seed, uint64_t k);
void orc_awgn_normals(uint64_t noise_seed, uint64_t i, float *z_re, float *z_im);

/* ---- clock ---- */
#define ORC_FLICKER_POLES 20
typedef struct {
    double c0, y0, sigma_step, master;
    int walk;
    uint64_t seed;
    int64_t frontier;
    double eps_hi, eps_lo;
    /* flicker FM (src/clock/mod.rs:55-149) */
    int flick;
    double f_alpha[ORC_FLICKER_POLES], f_gain[ORC_FLICKER_POLES], f_state[ORC_FLICKER_POLES];
    double f_scale;
    int64_t f_frontier;
    double f_hi, f_lo;
} orc_clock;
/* Flicker pieces, exposed for the pins: the poles (alpha, gain), the Allan
 * deviation of the unit-scaled pole sum at tau = m / 128 s, and normal n of a
 * receiver's flicker stream. */
void orc_flicker_poles(double *alpha, double *gain);
double orc_flicker_unit_adev(const double *alpha, int m);
double orc_flicker_normal(uint64_t clock_seed, uint64_t n);
void orc_clock_init(orc_clock *c, const orc_rx *rx);
double orc_clock_delta_t_at(orc_clock *c, double t);
double orc_clock_delta_t_at_master_index(orc_clock *c, int64_t k);

/* ---- receiver ops ---- */
double orc_thermal_noise_power_w(double temperature_k, double bandwidth_hz);
void orc_post_ops(const orc_post_op *ops, uint32_t n_ops, float *iq, uint64_t n, int16_t *codes);

/* ---- the band render (generate_band) ----
 * Renders samples [0, n) of one band in chunks of chunk_samples. out_iq gets
 * 2*n floats. Optional taps (length n each, may be NULL) record, for active
 * slot tap_slot, the solved emission time, range and f64 carrier phase, and
 * the receiver clock offset per sample. Returns 0, or -1 on bad input. */
int orc_render_band(const orc_tx *txs, uint32_t n_tx, const orc_rx *rx, const orc_band *band,
                    uint64_t n, uint64_t chunk_samples, float *out_iq, int16_t *codes,
                    uint32_t tap_slot, double *tap_t_emit, double *tap_range, double *tap_phase,
                    double *tap_delta_t);

/* Samples [i0, i0 + n) of one band with caller-supplied clock offsets
 * delta_t[0..n) (a window late in a long band without walking its clock);
 * optional taps of every active slot k at tap_*[k * n + j]. */
int orc_render_window(const orc_tx *txs, uint32_t n_tx, const orc_band *band, const orc_rx *rx,
                      uint64_t i0, uint64_t n, const double *delta_t, float *out_iq, int16_t *codes,
                      double *tap_t_emit, double *tap_range, double *tap_phase);

/* The reference CPU path as a timed baseline: renders every band over samples
 * [0, n_limit[b]) with the reference's two-level parallelism (bands in
 * parallel, transmitters of a chunk in parallel, fixed-order sum) on n_threads
 * OpenMP threads, hashing each band's IQ bytes with FNV-1a into hash_out[b].
 * Returns the number of contributions computed. */
uint64_t orc_render_scene_hashed(const orc_tx *txs, uint32_t n_tx, const orc_rx *rxs,
                                 const orc_band *bands, uint32_t n_bands, const uint64_t *n_limit,
                                 uint64_t chunk_samples, int n_threads, uint64_t *hash_out);

#ifdef __cplusplus
}
#endif
#endif
