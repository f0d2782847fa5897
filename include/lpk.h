/*
 * lpk.h -- C ABI of the MI355X (gfx950) IQ-generation backend.
 *
 * Drop-in boundary for lipika's per-chunk IQ path. The reference has no FFI;
 * its named slot for a backend is `enum Backend { Cpu }`
 * (src/backend/mod.rs:14-18, selected by `--backend`, src/bin/lipika.rs:25-27),
 * and the chunk interface it would bind is internal Rust:
 *   - scene setup        Transmitter::from_config      src/pipeline.rs:312-341
 *                        Receiver::from_config         src/pipeline.rs:352-364
 *   - one chunk body     generate_band, loop body      src/pipeline.rs:915-983
 *                        (clock, fan-out via gated_contribution :1033-1079,
 *                         ordered sum :965-976, AWGN :979, post-sum :983)
 *   - per-transmitter    accumulate_transmitter_contribution
 *                                                      src/channel/mod.rs:138-156
 *
 * Conventions:
 *   - every function returns 0 (LPK_OK) or a negative lpk_status; nothing
 *     aborts and no C++ exception crosses the boundary;
 *   - all pointers in descriptors are host pointers unless named `*_dev`;
 *     table contents are copied by lpk_scene_upload, so the caller may free
 *     them on return;
 *   - a context owns one HIP stream and is not thread-safe: one host thread per
 *     context, any number of contexts per process;
 *   - lpk_render_chunk is asynchronous on the context's stream; lpk_sync waits.
 * Host language note: the reference is Rust and no Rust toolchain exists in
 * this build, so the host side above this ABI is C++ / Python; a Rust
 * `extern "C"` block over this header is the intended binding.
 */
#ifndef LPK_H
#define LPK_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPK_ABI_VERSION 4

typedef enum {
    LPK_OK = 0,
    LPK_E_INVALID = -1,    /* bad argument or descriptor */
    LPK_E_NOMEM = -2,      /* host or device allocation failed */
    LPK_E_HIP = -3,        /* a HIP runtime call failed */
    LPK_E_NONFINITE = -4,  /* a non-finite input where a finite one is required */
    LPK_E_STATE = -5,      /* call order (e.g. render before scene upload) */
    LPK_E_NODEVICE = -6    /* no usable gfx950 device */
} lpk_status;

/* src/signal/mod.rs:27-37 and the five implementations */
typedef enum {
    LPK_SIG_CW = 0,        /* src/signal/cw.rs:37-53 */
    LPK_SIG_BPSK = 1,      /* src/signal/bpsk.rs:58-71 */
    LPK_SIG_OOK = 2,       /* src/signal/ook.rs:60-74 */
    LPK_SIG_QAM16 = 3,     /* src/signal/qam16.rs:96-109 */
    LPK_SIG_CA = 4         /* src/signal/ca_code.rs:167-193 */
} lpk_signal_kind;

/* src/signal/ca_code.rs:112-121 */
typedef enum { LPK_NAV_OFF = 0, LPK_NAV_DUMMY = 1, LPK_NAV_FULL = 2 } lpk_nav_mode;

/* AnyTrajectory::{Stationary, Precomputed}, src/geometry/mod.rs:25-55 */
typedef enum { LPK_TRAJ_STATIONARY = 0, LPK_TRAJ_GRID = 1 } lpk_traj_kind;

/* PostOp, src/rx/mod.rs:62-79 */
typedef enum { LPK_POST_GAIN = 0, LPK_POST_NONLINEARITY = 1, LPK_POST_ADC = 2 } lpk_post_kind;

/* Arithmetic form of a render. */
typedef enum {
    LPK_FORM_PLAIN = 0,      /* the reference's f64 path, operation for operation */
    LPK_FORM_PRECISION = 1,  /* f64 geometry and phase, f32 signal path, f64 transmitter sum;
                              * a band-chunk of more than 32 transmitters renders in passes of
                              * 32 that add their sums through HBM (render calls and plans) */
    LPK_FORM_CONVERGED = 2,  /* the plain form with the emission-time solve (src/channel/mod.rs:70-99)
                              * iterated on the unrounded time axis until an update is below
                              * 1e-13 s, from the reference's start t - 0.07 s */
    LPK_FORM_LEARNED = 3     /* LPK_FORM_CONVERGED started from a warm-start network's predicted
                              * light time (lpk_set_warm_start); a start that has not converged
                              * after 3 evaluations falls back to the reference's start */
} lpk_form;

/* One transmitter, as Transmitter::from_config leaves it (src/pipeline.rs:297-341). */
typedef struct {
    /* trajectory: Stationary (src/geometry/stationary.rs:21-24) or the 10 ms
     * PrecomputedTrajectory grid (src/geometry/precomputed.rs:23-70) */
    int32_t traj_kind;           /* lpk_traj_kind */
    int32_t signal_kind;         /* lpk_signal_kind */
    double pos_ecef[3];          /* STATIONARY: ECEF metres */
    double grid_t0_unix;         /* GRID: absolute time of node 0 (scene epoch - 0.5 s) */
    double grid_dt_s;            /* GRID: node spacing */
    uint64_t grid_n;             /* GRID: node count, 2 .. 2^31 - 1 */
    const double *grid_xyz;      /* GRID: (3, n) row-major, as ndarray Array2 (3, N) */
    double carrier_hz;           /* SignalModel::carrier_hz */
    double tx_power_w;           /* 10^(dBW/10), src/pipeline.rs:338 */
    double bandwidth_hz;         /* effective occupied bandwidth, src/config/mod.rs:191-193 */
    /* CW */
    double baseband_offset_hz;
    int32_t cw_anchored, _pad0;
    double anchor_phase_rad;     /* ContinuousWave::anchor_phase */
    /* BPSK / OOK / QAM16 (framed) */
    double rate_hz;              /* bit or symbol rate */
    double epoch_rel_s;          /* symbol-0 (framed) or chip-0 (C/A) tag, scene-epoch relative */
    int32_t repeating, _pad1;    /* repeat_from: None (0) / Some (1) */
    int64_t repeat_from;
    uint64_t n_symbols;
    const int8_t *bits;          /* BPSK +-1 or OOK 0/1, n_symbols entries */
    const float *symbols;        /* QAM16: n_symbols (re, im) pairs */
    /* C/A */
    int32_t nav_mode;            /* lpk_nav_mode */
    int32_t prn;
    const int8_t *ca_code;       /* 1023 chips, +-1 (lpk_ca_code) */
    uint64_t nav_seed;           /* DUMMY: seed already folded with the PRN (lpk_nav_fold_seed) */
    const int8_t *nav_bits;      /* FULL: +-1 LNAV bits at 50 bps */
    uint64_t n_nav_bits;
    /* burst gate, src/signal/gate.rs:21-104 */
    int32_t gated;
    uint32_t n_bursts;
    double gate_period_s;
    double gate_phase_s;         /* in [0, period) */
    const double *bursts;        /* n_bursts (start, end) pairs, end exclusive */
    /* the orbit the grid was sampled from, CircularKeplerian (src/geometry/kepler.rs:19-25);
     * only the learned form reads it (its network's features). has_kepler = 0: none */
    int32_t has_kepler, _pad2;
    double kepler_a_m, kepler_inc_deg, kepler_raan_deg, kepler_nu0_deg, kepler_epoch_unix;
} lpk_tx_desc;

/* Receiver trajectory kinds (src/pipeline.rs:77-107 as Receiver::from_config
 * builds them, :352-364). */
enum lpk_rx_traj_kind {
    LPK_RX_STATIONARY = 0,       /* fixed at pos_ecef (src/geometry/stationary.rs:21-39) */
    LPK_RX_KEPLER = 1            /* CircularKeplerian (src/geometry/kepler.rs:19-96), evaluated on the
                                  * device at every sample's epoch + t_utc, as position_ecef_batch is
                                  * queried in the chunk body (src/pipeline.rs:929-930) */
};

/* One receiver with its clock (src/pipeline.rs:344-364, NodeClockSpec
 * src/clock/mod.rs:20-46). */
typedef struct {
    double pos_ecef[3];          /* LPK_RX_STATIONARY: the position */
    double allan_deviation_1s;   /* white-FM sigma_y(1 s); 0 = no walk */
    double frequency_offset_y0;
    double initial_time_offset_s;
    double master_rate_hz;       /* highest band rate of the receiver, src/pipeline.rs:864-866 */
    uint64_t clock_seed;         /* src/pipeline.rs:825-827 */
    double flicker_floor;        /* flicker-FM Allan floor (src/clock/mod.rs:26-30); 0 = none */
    int32_t traj_kind, _pad;     /* lpk_rx_traj_kind */
    /* LPK_RX_KEPLER: the elements, as TrajectoryConfig::CircularKeplerian holds them */
    double kepler_a_m, kepler_inc_deg, kepler_raan_deg, kepler_nu0_deg, kepler_epoch_unix;
} lpk_rx_desc;

typedef struct {
    int32_t kind;                /* lpk_post_kind */
    int32_t bits;                /* ADC bits, 0..64 (src/rx/mod.rs:93-97; 2 below 2); at most 16 with adc_codes_host */
    double value;                /* GAIN: dB; NONLINEARITY: iip3 W; ADC: full-scale V */
} lpk_post_op;

/* One chunk [i0, i0 + clen) of one receiver band (src/pipeline.rs:911-990). */
typedef struct {
    uint32_t rx;                 /* receiver index */
    uint32_t n_active;
    const uint32_t *active;      /* in-band transmitter indices, in summation order */
    double fs_hz;                /* band sample rate */
    double lo_hz;                /* band LO */
    double epoch_unix;           /* scene epoch */
    uint64_t i0;                 /* global index of the chunk's first sample */
    uint64_t clen;               /* samples in the chunk */
    const double *delta_t;       /* optional host array [clen]: receiver clock offset per
                                  * sample (the caller's own ClockStream); NULL = device clock */
    double noise_power_w;        /* kTB or override; 0 = no AWGN */
    uint64_t noise_seed;         /* src/pipeline.rs:828-830 */
    uint32_t n_post, _pad;
    lpk_post_op post[4];         /* RxChain::post_sum in order */
    void *out_dev;               /* device complex64 [clen]; NULL = internal staging */
    void *out_host;              /* optional host complex64 [clen], filled by the render */
    int16_t *adc_codes_host;     /* optional host int16 [2*clen]: ADC integer levels */
    /* Receiver antenna pattern (src/rx/mod.rs:23-58), applied per source in the
     * channel contribution (src/channel/mod.rs:190-207, :253-255): sqrt(G) at the
     * arrival elevation, 0 below the cutoff (the sample's contribution is then
     * skipped). has_antenna = 0: no antenna, gain 1, nothing skipped. */
    int32_t has_antenna, has_cutoff;
    double antenna_peak_dbi, antenna_horizon_dbi, antenna_fb_db, antenna_cutoff_deg;
    uint8_t *unoccluded;         /* optional host uint8 [n_active], in `active` order: 1 where
                                  * the transmitter's gain was non-zero for some sample of the
                                  * chunk (src/channel/mod.rs:208-211); the render call then
                                  * returns only once the flags are written */
} lpk_band_chunk;

typedef struct lpk_ctx lpk_ctx;
typedef struct lpk_lws lpk_lws;

/* ---- context (the Backend slot, src/backend/mod.rs:14-18) ---- */
int lpk_abi_version(void);
/* Creates a context on HIP device `hip_device` with its own stream. */
int lpk_ctx_create(int hip_device, lpk_ctx **out);
void lpk_ctx_destroy(lpk_ctx *ctx);
/* Text of the last error on this context (or of the last failed create when
 * ctx is NULL). Valid until the next call on the context. */
const char *lpk_last_error(const lpk_ctx *ctx);

/* ---- scene upload (Transmitter::from_config / Receiver::from_config,
 *      src/pipeline.rs:312-364): copies grids, tables and constants to HBM
 *      once; replaces any previous scene of the context. ---- */
int lpk_scene_upload(lpk_ctx *ctx, const lpk_tx_desc *tx, uint32_t n_tx,
                     const lpk_rx_desc *rx, uint32_t n_rx);

/* ---- chunk render (the chunk body of generate_band, src/pipeline.rs:915-983):
 *      every job of the call runs in one kernel launch. Asynchronous. ---- */
int lpk_render_chunk(lpk_ctx *ctx, const lpk_band_chunk *jobs, uint32_t n_jobs, int form);
int lpk_sync(lpk_ctx *ctx);
/* Forget every receiver's clock-walk progress, so the next render regenerates
 * the walk from master index 0, as a fresh ClockStream does
 * (src/clock/mod.rs:268-295, one per band in src/pipeline.rs:869-872). The
 * values produced are unchanged; only the work is redone. Asynchronous: it
 * changes host bookkeeping only, and the stream orders the work after it. */
int lpk_clock_rewind(lpk_ctx *ctx);

/* Diagnostic render of one job that also records, for active slot `slot`,
 * the solved emission time, range (src/channel/mod.rs:97-98), f64 carrier
 * phase (:257-258) and the clock offset per sample (host arrays of clen
 * doubles; any may be NULL). Synchronous. Same kernels as lpk_render_chunk. */
int lpk_render_taps(lpk_ctx *ctx, const lpk_band_chunk *job, int form, uint32_t slot,
                    double *t_emit, double *range, double *phase, double *delta_t);

/* ---- tuning of the render launch (tile size and work split). The reference
 *      has no counterpart: its chunk body runs on the CPU thread pool
 *      (src/pipeline.rs:943-964). The plain form's samples do not depend on
 *      the tuning; the precision form fits its model per tile, so its samples
 *      move within its error bound when the tile size changes. ---- */
typedef struct {
    int32_t tile_shift;          /* log2 of the samples per workgroup tile: 7, 8 (default) or 9 */
    int32_t xcd_map;             /* 1 (default): deal contiguous tile runs to each XCD; 0: round-robin */
} lpk_tuning;
int lpk_set_tuning(lpk_ctx *ctx, const lpk_tuning *tuning);
int lpk_get_tuning(const lpk_ctx *ctx, lpk_tuning *tuning);
/* Renders `jobs` (which must not ask for host outputs or flags) once per
 * candidate tuning after a warm-up, keeps the fastest on the context and
 * returns it with its render-kernel time. Synchronous. */
int lpk_autotune(lpk_ctx *ctx, const lpk_band_chunk *jobs, uint32_t n_jobs, int form,
                 lpk_tuning *best, double *best_ms);

/* ---- streaming plans: generate_band's chunk loop (src/pipeline.rs:905-990)
 *      for many bands at once. Step k is a time slice: chunk_samples
 *      samples of the band with the highest sample rate and, for every other
 *      band, its samples in the same interval (lpk_plan_chunk_range gives
 *      each band's [i0, i0 + clen) in step k; outputs do not depend on the
 *      chunking, src/pipeline.rs:905-990 and tests/streaming_chunk_invariance.rs).
 *      The job, clock walk and carry tables of the step go to HBM and the
 *      walk, carry and render kernels run: even steps on the context's
 *      stream, odd steps on a second stream of the plan, so that a step's
 *      walk and render start while the previous step's render finishes
 *      (the walk carries run in step order across the two). Every other
 *      entry point of the context, lpk_sync included, orders itself after
 *      the steps already made.
 *      Each band descriptor covers the whole band (i0 = 0, clen = its
 *      sample count) with the device clock (delta_t NULL) and no host
 *      outputs or flags; out_dev, if set, holds the whole band and step k
 *      writes its chunk there, otherwise the plan keeps two chunk buffers per
 *      band, one per stream, and lpk_plan_output gives the one the last
 *      step wrote. Steps run in order from 0; step 0 restarts
 *      the plan's receivers' clock walks. The tuning is the context's at
 *      creation. In the precision form a band of more than 32 transmitters
 *      renders in passes (one launch each, in order on the step's stream)
 *      through a staging buffer per stream of 16 B per sample of its chunk;
 *      a band may hold up to 48 x 32 transmitters. ---- */
typedef struct lpk_plan lpk_plan;
int lpk_plan_create(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, uint64_t chunk_samples,
                    int form, lpk_plan **out);
int lpk_plan_step(lpk_plan *plan, uint64_t chunk);     /* asynchronous, like lpk_render_chunk */
/* n_chunks: the plan's steps; max_tiles: the render workgroups one step
 * launches at most (with precision passes, the sum of each pass's largest) */
int lpk_plan_info(const lpk_plan *plan, uint64_t *n_chunks, uint32_t *max_tiles);
int lpk_plan_chunk_range(const lpk_plan *plan, uint32_t band, uint64_t chunk, uint64_t *i0, uint64_t *clen);
int lpk_plan_output(const lpk_plan *plan, uint32_t band, void **chunk_dev);
void lpk_plan_destroy(lpk_plan *plan);

/* Time of the last lpk_render_chunk kernel on the context's stream, in ms,
 * measured with HIP events around the launch (valid after lpk_sync). */
int lpk_last_kernel_ms(lpk_ctx *ctx, double *ms);
/* Kernel-time accounting for benchmarks (the reference's profile::time
 * sections, src/profile.rs:27-36, for the GPU path): lpk_timing_reset waits
 * for the stream and clears the record; every later render launch is
 * bracketed by HIP events; lpk_timing_read waits for the stream and returns
 * the summed render-kernel time and the number of launches since the reset. */
int lpk_timing_reset(lpk_ctx *ctx);
int lpk_timing_read(lpk_ctx *ctx, double *total_ms, uint64_t *n_launches);

/* ---- device memory helpers for callers that keep output in HBM ---- */
int lpk_device_alloc(lpk_ctx *ctx, size_t bytes, void **dev);
int lpk_device_free(lpk_ctx *ctx, void *dev);
int lpk_copy_to_host(lpk_ctx *ctx, void *host, const void *dev, size_t bytes);
/* FNV-1a 64 of `bytes` bytes of device memory (the hashing sink of
 * examples/sky_throughput.rs:73-82), computed on the host after a copy. */
int lpk_device_fnv1a(lpk_ctx *ctx, const void *dev, size_t bytes, uint64_t *hash);

/* ---- host-side table builders used by scene setup (pure host code) ---- */
/* src/geometry/ecef.rs:10-20 */
void lpk_geodetic_to_ecef(double lat_deg, double lon_deg, double h_m, double out[3]);
/* PrecomputedTrajectory::sample over a CircularKeplerian orbit
 * (src/geometry/precomputed.rs:47-70, src/geometry/kepler.rs:27-85).
 * lpk_grid_size gives n; xyz must hold 3*n doubles. */
uint64_t lpk_grid_size(double duration_s, double grid_dt_s);
int lpk_grid_fill_kepler(double a_m, double inc_deg, double raan_deg, double nu0_deg,
                         double orbit_epoch_unix, double scene_epoch_unix, double duration_s,
                         double grid_dt_s, double *t0_out, double *xyz, uint64_t n);
/* CircularKeplerian::position_ecef (src/geometry/kepler.rs:27-85) at t_unix:
 * the position a Keplerian receiver has at that time. */
void lpk_kepler_position(double a_m, double inc_deg, double raan_deg, double nu0_deg, double orbit_epoch_unix,
                         double t_unix, double out[3]);
/* src/signal/ca_code.rs:66-78 */
int lpk_ca_code(int prn, int8_t out[1023]);
/* src/signal/ca_code.rs:146-151 */
uint64_t lpk_nav_fold_seed(uint64_t nav_seed, uint64_t prn);
/* BpskFrame/OokFrame/Qam16Frame::random (src/signal/bpsk.rs:27-41,
 * ook.rs:26-44, qam16.rs:62-80): sync word then seeded payload. */
int lpk_payload_bpsk(const int8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out);
int lpk_payload_ook(const uint8_t *sync, uint64_t n_sync, uint64_t n_payload, uint64_t seed, int8_t *out);
int lpk_payload_qam16(const uint8_t *sync_bits, uint64_t n_sync_bits, uint64_t n_payload,
                      uint64_t seed, float *out_iq);
/* src/signal/mod.rs:50-101, src/signal/cw.rs:26-34 */
int lpk_elapsed_periods(double elapsed_s, double rate_hz, int64_t *whole, double *frac);
int lpk_frame_anchor(double render_epoch, double signal_epoch, double rate_hz, uint64_t n_symbols,
                     double *epoch_rel_s, int64_t *symbol_offset);
int lpk_cw_anchor_phase(double render_epoch, double signal_epoch, double offset_hz, double *phase);
/* src/pipeline.rs:70-75 */
int lpk_bands_overlap(double tx_carrier_hz, double tx_bw_hz, double band_lo_hz, double band_bw_hz);
/* src/channel/mod.rs:310-312 */
double lpk_thermal_noise_power_w(double temperature_k, double bandwidth_hz);

/* ---- GPS LNAV and broadcast ephemeris (host builders, src/signal/gps_nav.rs) ----
 * They produce the nav_bits table (LPK_NAV_FULL) and the position grid of an
 * lpk_tx_desc. Angles in radians, times in GPS seconds of week. */
typedef struct lpk_ephemeris {       /* Ephemeris, src/signal/gps_nav.rs:56-85 */
    uint32_t prn, week;
    double toe_s, toc_s, sqrt_a, e, m0_rad, delta_n_rad_s, i0_rad, idot_rad_s;
    double omega0_rad, omega_dot_rad_s, argp_rad;
    double cuc, cus, crc, crs, cic, cis;
    double af0, af1, af2, tgd;
    uint32_t iode, iodc, health, ura_index;
} lpk_ephemeris;
/* GpsTime::from_unix (src/signal/gps_nav.rs:31-41) */
void lpk_gps_time(double unix_s, double utc_offset_s, uint32_t *week, double *sow);
/* The 16 s toe grid of Ephemeris::from_trajectory (src/signal/gps_nav.rs:140-144):
 * the UTC time whose GPS seconds of week are round16(sow), that toe and its week. */
int lpk_toe_snap(double unix_s, double utc_offset_s, double *unix_at_toe, uint32_t *week, double *toe_sow);
/* Ephemeris::from_circular_keplerian (src/signal/gps_nav.rs:87-129) */
int lpk_ephemeris_from_kepler(double a_m, double inc_deg, double raan_deg, double nu0_deg, double orbit_epoch_unix,
                              uint32_t prn, double scene_epoch_unix, double utc_offset_s, lpk_ephemeris *out);
/* Ephemeris::from_trajectory (src/signal/gps_nav.rs:131-180) from the ECEF
 * state r (m), v (m/s) at the snapped toe (lpk_toe_snap) */
int lpk_ephemeris_from_state(const double r[3], const double v[3], uint32_t prn, uint32_t week, double toe_sow,
                             lpk_ephemeris *out);
/* lnav_bits (src/signal/gps_nav.rs:554-585): +-1 LNAV bits for duration_s from
 * GPS seconds of week sow; n must be lpk_lnav_length(duration_s). */
uint64_t lpk_lnav_length(double duration_s);
int lpk_lnav_bits(const lpk_ephemeris *eph, double sow, double duration_s, int8_t *out, uint64_t n);
/* sv_position_ecef (src/signal/gps_nav.rs:587-628) */
void lpk_sv_position(const lpk_ephemeris *eph, double t_sow, double out[3]);
/* PrecomputedTrajectory::sample (src/geometry/precomputed.rs:47-70) over a
 * GpsBroadcastTrajectory (src/signal/gps_nav.rs:630-654); xyz holds 3*n doubles,
 * n = lpk_grid_size(duration_s, grid_dt_s). */
int lpk_grid_fill_broadcast(const lpk_ephemeris *eph, double utc_offset_s, double scene_epoch_unix,
                            double duration_s, double grid_dt_s, double *t0_out, double *xyz, uint64_t n);
/* parse_rinex3_gps_nav (src/signal/gps_nav.rs:656-728): the GPS record of prn
 * whose toe is nearest near_unix; LPK_E_INVALID when the text holds none. */
int lpk_rinex3_gps_ephemeris(const char *text, uint32_t prn, double near_unix, double utc_offset_s,
                             lpk_ephemeris *out);

/* ---- TLE trajectories (host builders, src/geometry/tle.rs:15-85). The
 *      reference propagates with the third-party sgp4 crate (2.3, WGS72,
 *      improved mode); this is SGP4/SDP4 restated from Spacetrack Report #3
 *      as revised by Vallado et al. (AIAA 2006-6753). ---- */
typedef struct lpk_tle {
    uint32_t satnum;
    int32_t epoch_year;           /* four digits (two-digit years 57-99 are 1957-1999) */
    double epoch_day;             /* day of year, 1.0 = 1 January 0h UTC */
    double epoch_unix;            /* UTC unix seconds of the epoch, whole microseconds (tle.rs:37-39) */
    double jd_epoch;              /* Julian date of the epoch */
    double ndot, nddot, bstar;    /* as in the TLE: rev/day^2 / 2, rev/day^3 / 6, 1 / earth radii */
    double inclo_deg, nodeo_deg, ecco, argpo_deg, mo_deg, no_revday;
} lpk_tle;
/* Elements::from_tle: 69-column lines with valid checksums; LPK_E_INVALID otherwise */
int lpk_tle_parse(const char *line1, const char *line2, lpk_tle *out);
/* Constants::propagate: TEME position (km) and velocity (km/s) at `minutes` from epoch */
int lpk_sgp4_teme(const lpk_tle *tle, double minutes, double r_km[3], double v_kms[3]);
/* TleTrajectory::position_ecef / velocity_ecef (tle.rs:64-85); vel_ms may be NULL */
int lpk_tle_state_ecef(const lpk_tle *tle, double t_unix, double pos_m[3], double vel_ms[3]);
/* PrecomputedTrajectory::sample (src/geometry/precomputed.rs:47-70) over a
 * TleTrajectory; xyz holds 3*n doubles, n = lpk_grid_size(duration_s, grid_dt_s). */
int lpk_grid_fill_tle(const lpk_tle *tle, double scene_epoch_unix, double duration_s, double grid_dt_s,
                      double *t0_out, double *xyz, uint64_t n);


/* ---- learned warm start (no counterpart in the reference: its solve always
 * starts at t - 0.07 s, src/channel/mod.rs:75-96) ----
 * A BF16 multilayer perceptron, 10 inputs -> 3 x 64 GELU -> 1, FP32
 * accumulation on the matrix cores, predicting the light time tau of the
 * emission equation as tau_sl + 1e-6 s * net(features), tau_sl = |P(t) - O| / c.
 * Features: O / 6378137 m (3), (a - 1.7e7 m) / 1e7 m, cos i, sin i, cos and sin
 * of the Earth-fixed node raan - w_E (t - t_el), cos and sin of the argument of
 * latitude nu0 + n (t - t_el). Parameters (f32, LPK_LWS_PARAMS), W[out][in]:
 *   W1[64][10] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4.
 * A network belongs to the context it was created on; destroy it before the
 * context. */
#define LPK_LWS_PARAMS 9089
typedef struct {
    uint64_t n_pairs;            /* training pairs generated on the device */
    uint64_t data_seed;
    uint32_t epochs;
    uint32_t batch;              /* minibatch, a multiple of 64, <= 2^20 */
    float lr, lr_final;          /* Adam step size, cosine-decayed from lr to lr_final */
    uint64_t n_val;              /* validation pairs, drawn with an independent seed */
} lpk_lws_train_cfg;
typedef struct {
    double loss_first, loss_last;   /* mean squared error (network units) over the first / last epoch */
    double val_rms_s, val_max_s;    /* validation error of the predicted light time, seconds */
    double gen_ms, train_ms;        /* device time of pair generation and of the training steps */
    uint64_t steps;
} lpk_lws_train_report;
int lpk_lws_create(lpk_ctx *ctx, uint64_t init_seed, lpk_lws **out);   /* Glorot-uniform init */
void lpk_lws_destroy(lpk_lws *net);
int lpk_lws_get_params(lpk_lws *net, float *theta);        /* host [LPK_LWS_PARAMS] */
int lpk_lws_set_params(lpk_lws *net, const float *theta);  /* also resets the optimiser */
/* Training pairs: x_dev [n][10] features, y_dev [n] (tau - tau_sl) / 1e-6 s, and
 * optionally geom_dev [n][8] = receiver ECEF xyz, a, i, node, u (rad), tau (s).
 * Geometry: receiver lat U[-90, 90] deg, lon U[-180, 180) deg, h U[0, 2000] m; orbit
 * LEO a U[6.9e6, 7.5e6] m or MEO a U[2.6e7, 2.7e7] m, i U[0, 180) deg, node and u U[0, 360) deg.
 * tau is the converged root of the emission equation on the analytic orbit. Asynchronous. */
int lpk_lws_generate(lpk_ctx *ctx, uint64_t n, uint64_t seed, float *x_dev, float *y_dev, double *geom_dev);
int lpk_lws_infer(lpk_lws *net, const float *x_dev, uint64_t n, float *y_dev);   /* asynchronous */
/* Gradient of the mean squared error over the n pairs (synchronous; loss may be NULL). */
int lpk_lws_gradient(lpk_lws *net, const float *x_dev, const float *y_dev, uint64_t n, float *grad, double *loss);
/* lpk_lws_gradient without any device allocation: no counterpart in the
 * reference, for callers that cannot allocate (HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). Same kernels and results; the
 * workspace is a static device array of 4.7 MB per device, so n is at most
 * 8192 pairs (LPK_E_INVALID beyond). Calls that overlap on one device race
 * on that workspace. Synchronous; loss may be NULL. */
int lpk_lws_gradient_noalloc(lpk_lws *net, const float *x_dev, const float *y_dev, uint64_t n, float *grad,
                             double *loss);
/* Generates the pairs and trains with Adam (beta 0.9 / 0.999, eps 1e-8); synchronous. */
int lpk_lws_train(lpk_lws *net, const lpk_lws_train_cfg *cfg, lpk_lws_train_report *rep);
/* The network LPK_FORM_LEARNED predicts with (NULL: none). */
int lpk_set_warm_start(lpk_ctx *ctx, lpk_lws *net);
/* Converged-form counters since the last reset: emission-equation evaluations,
 * contributions, learned starts that fell back, solves that did not converge. */
int lpk_solve_stats(lpk_ctx *ctx, uint64_t stats[4]);
int lpk_solve_stats_reset(lpk_ctx *ctx);   /* also resets the horizon-gate counters */

/* ---- horizon gate (src/channel/horizon.rs) ----
 * In bands with an antenna cutoff, a (tile, transmitter) pair whose arrival
 * elevation is certainly below the cutoff for the whole tile
 * (certainly_below_cutoff, src/channel/horizon.rs:150-210) is skipped; every
 * sample would have had zero gain, so the output is byte-identical with the
 * gate off (the reference's *_without_horizon_gate path). On by default. */
int lpk_set_horizon_gate(lpk_ctx *ctx, int on);
/* (tile, transmitter) pairs skipped and tested since lpk_solve_stats_reset */
int lpk_horizon_gate_stats(lpk_ctx *ctx, uint64_t stats[2]);

/* ---- output ----
 * Streams whole bands to raw little-endian complex64 files, the FileSink /
 * StreamingC64Writer of src/pipeline.rs:398-406 and src/output/mod.rs:95-139
 * (the JSON sidecar is the caller's): chunk k of every band renders on the
 * device while the host writes chunk k - 1 from pinned, double-buffered
 * memory. bands: whole-band descriptors (i0 = 0, clen = the band's length, no
 * out_dev / out_host / adc_codes_host); paths[b]: band b's file, truncated.
 * The bytes equal those of chunked lpk_render_chunk calls of any chunk size.
 * Synchronous. */
int lpk_stream_to_files(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, const char *const *paths,
                        uint64_t chunk_samples, int form);
/* The IqSink of src/pipeline.rs:372-375 (sink.push at :984) as a C callback:
 * called on the calling thread with chunk [i0, i0 + n) of band `band` (an
 * index into the descriptors given), in chunk order per band; iq holds 2 * n
 * floats (interleaved complex64) valid for the call only. Return 0 to
 * continue, non-zero to stop that band (ControlFlow::Break). */
typedef int (*lpk_iq_sink)(void *user, uint32_t band, uint64_t i0, const float *iq, uint64_t n);
/* generate_band's chunk loop with its sink (src/pipeline.rs:905-990) for
 * whole bands at once: chunk k of every band renders on the device into
 * pinned, double-buffered host memory while the sink takes chunk k - 1 of
 * each band. bands: whole-band descriptors as for lpk_stream_to_files; a
 * band whose sink returned non-zero renders no further chunk. The bytes
 * pushed equal those of chunked lpk_render_chunk calls. Synchronous. */
int lpk_stream_to_sink(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, uint64_t chunk_samples, int form,
                       lpk_iq_sink sink, void *user);
/* The hashing sink of examples/sky_throughput.rs:73-82: user points to one
 * uint64 FNV-1a state per band (start each at 0xCBF29CE484222325). */
int lpk_sink_fnv1a(void *user, uint32_t band, uint64_t i0, const float *iq, uint64_t n);
/* FNV-1a 64 of `bytes` host bytes, continuing from state h. */
uint64_t lpk_fnv1a64(const void *data, size_t bytes, uint64_t h);
/* Pinned host memory, for out_host buffers that the device fills
 * asynchronously (no counterpart in the reference, which renders into
 * ordinary heap arrays). */
int lpk_host_alloc(lpk_ctx *ctx, size_t bytes, void **host);
int lpk_host_free(lpk_ctx *ctx, void *host);
/* export_truth's geometry (src/pipeline.rs:548-611) for receiver rx over the
 * epoch-relative grid t_rel[n]: range[q * n + k] of every uploaded transmitter
 * q from solve_emission_times with `iterations` iterations (the reference uses
 * 4) against the receiver's position at epoch + t_rel[k] (its trajectory's, as
 * :562-563 queries it), and src_xyz[(q * 3 + c) * n + k] the source position
 * at epoch + t_rel[k]. Synchronous; host arrays. */
int lpk_truth(lpk_ctx *ctx, uint32_t rx, const double *t_rel, uint64_t n, double epoch_unix, int iterations,
              double *range, double *src_xyz);

#ifdef __cplusplus
}
#endif
#endif
