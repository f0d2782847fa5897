// lpk_api.hip -- the C ABI of include/lpk.h: contexts, scene upload, the
// per-call job table and the launches. No exception leaves this file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/lpk.h"
#include "lpk_device.hpp"
#include "lpk_lws.hpp"

namespace lpk {
hipError_t launch_render(int form, bool antenna, bool moving, int tile_shift, bool xcd_map, const JobDev *jobs,
                         int n_jobs, const LinkDev *links, const TxDev *txs, uint32_t n_tiles,
                         const uint32_t *n_live, double *tap_t_emit, double *tap_range, double *tap_phase,
                         double *tap_delta_t, uint32_t tap_slot, const double *pred, unsigned long long *stats,
                         bool gate, unsigned long long *gate_stats, hipStream_t stream);
hipError_t launch_lws_predict(const JobDev *jobs, int n_jobs, const LinkDev *links, const TxDev *txs,
                              const LwsNetDev &net, int tile_shift, uint64_t n_queries, double *pred,
                              unsigned long long *stats, hipStream_t stream);
hipError_t launch_lws_gen(uint64_t n, uint64_t seed, float *X, float *Y, double *G, hipStream_t s);
hipError_t launch_lws_infer(const LwsNetDev &net, const float *X, uint64_t n, float *Y, hipStream_t s);
hipError_t launch_lws_train(const LwsNetDev &net, const float *X, const float *Y, uint64_t base, uint64_t end,
                            uint32_t n_slices, float scale, float *partials, hipStream_t s);
hipError_t launch_lws_update(const float *partials, int n_rows, float *theta, float *m1, float *m2, void *packed,
                             float lr, float beta1, float beta2, float eps, float c1, float c2, float *grad_out,
                             float *loss_out, hipStream_t s);
hipError_t launch_lws_pack(const float *theta, void *packed, hipStream_t s);
hipError_t launch_truth(const TxDev *txs, uint32_t n_tx, const double *obs, double epoch, const double *t, uint64_t n,
                        int iterations, double *range, double *src, hipStream_t stream);
hipError_t launch_walk_gen(const WalkSegDev *segs, int n_segs, uint32_t n_blocks, const uint32_t *n_live,
                           hipStream_t stream);
hipError_t launch_walk_carry(const CarrySegDev *segs, int n_segs, hipStream_t stream);
}  // namespace lpk

using namespace lpk;

// Workspace of lpk_lws_gradient_noalloc: kNoAllocRows rows of per-slice
// partial gradients (kTrainQ pairs each, so at most kNoAllocRows * kTrainQ
// pairs per call), then the reduced gradient and loss from kNoAllocGradOff
// (a 256-byte boundary). 4.7 MB, held on every device the library runs on
// until the process ends. Host code never passes `scratch` itself (a host
// stand-in): it uses the device address hipGetSymbolAddress gives for the
// current device.
constexpr int kNoAllocRows = 128;
constexpr size_t kNoAllocGradOff = size_t(kNoAllocRows) * kLwsRow;
constexpr size_t kNoAllocFloats = kNoAllocGradOff + size_t((kLwsRow + 63) / 64) * 64;
static_assert(kNoAllocGradOff % 64 == 0, "gradient block on a 256-byte boundary");
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {

constexpr int kMaxDevices = 64;
// A plan step's table header: word 1 the live walk blocks, word
// kPlanPassHdr + p the live tiles of pass p (of at most kPlanMaxPasses).
constexpr uint32_t kPlanPassHdr = 16;
constexpr uint32_t kPlanMaxPasses = 256 / 4 - kPlanPassHdr;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices] = {};
hipError_t g_scratch_rc[kMaxDevices] = {};
float *g_scratch[kMaxDevices] = {};

// Looks up `scratch`'s address on `device` (which must be current) once per
// device; later calls return what the first found.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void *p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_rc[device] = rc;
        g_scratch[device] = rc == hipSuccess ? static_cast<float *>(p) : nullptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_rc[device];
}

}  // namespace

namespace {

std::string g_create_error;

struct Fail {
    int code;
    std::string msg;
};

// Growable device buffer (never shrinks; reallocation waits for the stream).
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void *ensure(size_t bytes, hipStream_t s) {
        if (bytes <= cap && p) return p;
        size_t want = std::max<size_t>(bytes, 256);
        if (p) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            throw Fail{LPK_E_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed"};
        }
        cap = want;
        return p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned host staging slot, reusable once its event has completed.
struct PinSlot {
    void *p = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};

// Flicker FM of one receiver (src/clock/mod.rs:55-149). Its 128 Hz series is
// small (7 680 steps per minute), so it is stepped here on the host, with the
// device's Philox4x64-10 / Box-Muller normals, and the render kernels
// interpolate its integral D[k] (uploaded to `dev`).
struct FlickerGen {
    bool on = false;
    uint64_t seed = 0;
    double alpha[kFlickerPoles] = {}, gain[kFlickerPoles] = {}, state[kFlickerPoles] = {};
    double scale = 0.0;
    std::vector<double> D;      // D[k] = sum of the steps 0..k of y / 128
    DevBuf dev;
    size_t dev_n = 0;           // entries of D uploaded
};

struct RxState {
    lpk_rx_desc d{};
    FlickerGen flick;
    int walk = 0;
    double sigma_step = 0.0;
    double *base = nullptr;     // base[b] = eps(b*kWalkBlock - 1), base[0] = 0
    int64_t base_cap = 0;       // entries
    int64_t nb_known = 0;       // base[0..nb_known] valid
    DevBuf local, toff;         // the current call's window (f32 per index, f64 per 32)
};

}  // namespace

struct lpk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // the last launch's pair (owned by tpool)
    // event pairs around every render launch since the last timing reset
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tpool;
    size_t tused = 0;
    std::string err;
    // scene
    std::vector<void *> scene_allocs;
    std::vector<TxDev> tx_host;
    std::vector<double> tx_carrier;
    TxDev *tx_dev = nullptr;
    uint32_t n_tx = 0;
    std::vector<RxState> rx;
    bool have_scene = false;
    bool timed = false;
    // tuning of the render launch
    int tile_shift = kDefaultTileShift;
    int xcd_map = 1;
    // per-call buffers
    DevBuf jobs, links, dt_stage, out_stage, codes_stage, blocksum, walk_tables, part_stage;
    // converged and learned forms: the network the learned form predicts
    // with, its prediction table and the solve counters
    lpk_lws *lws = nullptr;
    DevBuf pred;
    // counters: [0..3] converged-form solves, [4..5] horizon gate (skipped, tested)
    unsigned long long *stats = nullptr;
    int gate = 1;   // apply the horizon gate in antenna scenes
    PinSlot pin[4];
    int pin_next = 0;
    // live streaming plans: their odd steps run on a stream of their own,
    // which every other entry point joins (the context's stream waits on the
    // plan's last odd step) before it enqueues work
    std::vector<lpk_plan *> plans;
};

// A streaming plan (lpk_plan_*): chunk k of every band per step as
// [table copy, clock walk, carry, render], the even steps on the context's
// stream and the odd ones on the plan's second stream, so that a step's
// clock walk and the start of its render overlap the previous step's render
// tail. The two slots have their own pinned and device tables, walk windows,
// block sums and chunk buffers; the carries, which extend each receiver's
// shared base table, run in step order across the streams (an event per
// slot), and a restart at chunk 0 waits for both slots.
struct lpk_plan {
    lpk_ctx *ctx = nullptr;
    int form = 0, tile_shift = kDefaultTileShift, xcd_map = 1;
    bool antenna = false;
    bool moving = false;    // some band's receiver moves
    uint64_t chunk = 0, n_chunks = 0, next = 0;
    std::vector<lpk_band_chunk> bands;
    std::vector<double> rate_ratio;       // per band: fs / the highest fs of the plan
    std::vector<uint64_t> chunk_cap;      // per band: its largest chunk
    std::vector<std::vector<uint32_t>> active;
    std::vector<uint32_t> order;          // bands in job-table order (heaviest first)
    std::vector<uint64_t> link0;          // per band: its first link
    uint64_t total_links = 0;
    // precision bands of more than kPrecisionMaxLinks links render in passes
    std::vector<uint32_t> n_pass;         // per band: its passes (1 for most)
    std::vector<uint64_t> part_off;       // per band in passes: its samples' offset in part
    std::vector<uint32_t> pass_jobs;      // per pass: the bands it renders (in `order`)
    std::vector<uint32_t> pass_tiles;     // per pass: its largest tile count over the steps
    DevBuf part[2];                       // per slot: the passes' f64 sums
    std::vector<uint32_t> walk_rx;        // receivers whose device clock the plan walks
    std::vector<int> walk_slot;           // per receiver: its index in walk_rx, or -1
    std::vector<DevBuf> local[2], toff[2];  // per slot and walked receiver: the step's window
    std::vector<int64_t> win_cap;         // per walked receiver: window capacity in blocks
    DevBuf blocksum[2];
    uint32_t max_blocks = 0;
    std::vector<DevBuf> outbuf[2];        // per slot and band: the chunk buffer when out_dev is NULL
    size_t off_jobs = 0, off_links = 0, off_flags = 0, off_segs = 0, off_carry = 0, bytes = 0;
    DevBuf tables[2];
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipEvent_t carry_ev[2] = {nullptr, nullptr};  // each slot's last walk carry
    hipEvent_t ctx_ev = nullptr;          // the context's stream, when other work came between steps
    hipStream_t aux = nullptr;            // the odd steps' stream
    bool aux_pending = false;             // odd steps not yet joined into the context's stream
    bool foreign = false;                 // other work was enqueued since this plan's last step
    bool used[2] = {false, false};
};

// A warm-start network on its context's device: FP32 master parameters, the
// Adam moments and the BF16 operand copies the matrix cores read.
struct lpk_lws {
    lpk_ctx *ctx = nullptr;
    float *theta = nullptr, *m1 = nullptr, *m2 = nullptr;   // kLwsParams each
    void *packed = nullptr;                                  // kLwsPacked BF16
    uint64_t step = 0;                                       // Adam steps taken
    DevBuf partials, grad, loss;
};

namespace {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Fail{LPK_E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

int fail(lpk_ctx *ctx, const Fail &f) {
    if (ctx) ctx->err = f.msg;
    return f.code;
}

// Orders the context's stream after every step enqueued on a plan's second
// stream (but those of `self`), and notes in those plans that other work
// follows their steps (their next odd step then waits for it).
void join_plans(lpk_ctx *ctx, const lpk_plan *self = nullptr) {
    for (lpk_plan *P : ctx->plans) {
        if (P == self) continue;
        if (P->aux_pending) {
            hip_check(hipStreamWaitEvent(ctx->stream, P->done[1], 0), "plan join");
            P->aux_pending = false;
        }
        P->foreign = true;
    }
}

template <class F>
int guarded(lpk_ctx *ctx, F &&body, bool join = true) {
    try {
        if (join && ctx) join_plans(ctx);
        body();
        return LPK_OK;
    } catch (const Fail &f) {
        return fail(ctx, f);
    } catch (const std::bad_alloc &) {
        return fail(ctx, Fail{LPK_E_NOMEM, "host allocation failed"});
    } catch (...) {
        return fail(ctx, Fail{LPK_E_INVALID, "unexpected internal error"});
    }
}

void *upload(lpk_ctx *ctx, const void *host, size_t bytes) {
    if (!bytes) return nullptr;
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) throw Fail{LPK_E_NOMEM, "scene table allocation failed"};
    ctx->scene_allocs.push_back(d);
    hip_check(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice), "scene table copy");
    return d;
}

void free_scene(lpk_ctx *ctx) {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void *p : ctx->scene_allocs) (void)hipFree(p);
    ctx->scene_allocs.clear();
    for (auto &r : ctx->rx) {
        if (r.base) (void)hipFree(r.base);
        r.base = nullptr;
        r.local.release();
        r.toff.release();
        r.flick.dev.release();
    }
    ctx->rx.clear();
    ctx->tx_host.clear();
    ctx->tx_carrier.clear();
    ctx->tx_dev = nullptr;
    ctx->n_tx = 0;
    ctx->have_scene = false;
}

bool is_fin(double x) { return std::isfinite(x); }

// Normal n of a receiver's flicker stream: Philox4x64-10 block n >> 3 of
// (clock seed, "flickern"), word (n >> 1) & 3, the cos branch of its
// Box-Muller pair for even n and the sin branch for odd n. Step k of pole j
// uses n = (k + 1) * 20 + j; n < 20 start the poles.
double flicker_normal(uint64_t seed, uint64_t n) {
    uint64_t w[4] = {n >> 3, 0, 0, 0};
    philox4x64(w, seed, kFlickerTag);
    float zc, zs;
    box_muller(w[(n >> 1) & 3], zc, zs);
    return static_cast<double>((n & 1) ? zs : zc);
}

// The AR(1) poles of src/clock/mod.rs:71-80 and the Allan deviation of their
// unit-scaled sum at tau = m / 128 s: for a pole a (autocorrelation a^|d|,
// unit variance) the non-overlapping estimator of src/clock/mod.rs:150-163 has
// expectation (E[S1^2] - E[S1 S2]) / m^2, where S1 and S2 are sums over two
// adjacent blocks of m steps:
//   E[S1^2]   = m + 2 sum_{d=1}^{m-1} (m - d) a^d,
//   E[S1 S2]  = sum_{d=1}^{m} d a^d + sum_{d=m+1}^{2m-1} (2m - d) a^d.
// The reference measures that deviation once on a random realisation; the
// expectation makes the calibration exact and the same on every run.
void flicker_poles(double *alpha, double *gain) {
    const double f_hi = kFlickerRate / 2.0;
    for (int j = 0; j < kFlickerPoles; ++j) {
        const double f = kFlickerFLo * std::pow(f_hi / kFlickerFLo, double(j) / double(kFlickerPoles - 1));
        const double a = std::exp(-2.0 * 3.141592653589793 * f / kFlickerRate);
        alpha[j] = a;
        gain[j] = std::sqrt(1.0 - a * a);
    }
}

double flicker_unit_adev(const double *alpha, int m) {
    double var = 0.0;
    for (int j = 0; j < kFlickerPoles; ++j) {
        const double a = alpha[j];
        double s11 = double(m), s12 = 0.0, p = 1.0;
        for (int d = 1; d < 2 * m; ++d) {
            p *= a;
            if (d < m) s11 += 2.0 * double(m - d) * p;
            s12 += (d <= m ? double(d) : double(2 * m - d)) * p;
        }
        var += (s11 - s12) / (double(m) * double(m));
    }
    return std::sqrt(var);
}

// Flicker::new without the burn-in: each pole starts from its stationary
// distribution N(0, 1), which the reference approaches by 200 000 warm-up steps.
void flicker_init(FlickerGen &F, double floor_adev, uint64_t seed) {
    F.on = floor_adev > 0.0;
    F.D.clear();
    F.dev_n = 0;
    if (!F.on) return;
    F.seed = seed;
    flicker_poles(F.alpha, F.gain);
    F.scale = floor_adev / flicker_unit_adev(F.alpha, kFlickerCalibM);
    for (int j = 0; j < kFlickerPoles; ++j) F.state[j] = flicker_normal(seed, uint64_t(j));
}

// Flicker::advance_to (src/clock/mod.rs:125-133) until D holds index k_hi.
void flicker_extend(FlickerGen &F, int64_t k_hi) {
    for (int64_t k = int64_t(F.D.size()); k <= k_hi; ++k) {
        double y = 0.0;
        for (int j = 0; j < kFlickerPoles; ++j) {
            F.state[j] = F.alpha[j] * F.state[j] + F.gain[j] * flicker_normal(F.seed, uint64_t(k + 1) * kFlickerPoles + j);
            y += F.state[j];
        }
        y = y * F.scale;
        F.D.push_back((k ? F.D[size_t(k - 1)] : 0.0) + y / kFlickerRate);
    }
}

LwsNetDev lws_net(const lpk_lws *net) {
    return LwsNetDev{static_cast<const bf16x8 *>(net->packed), net->theta};
}

// The solve and gate counters (zeroed on first use).
constexpr int kNumStats = 6;
void ensure_stats(lpk_ctx *ctx) {
    if (ctx->stats) return;
    void *p = nullptr;
    if (hipMalloc(&p, kNumStats * sizeof(unsigned long long)) != hipSuccess) throw Fail{LPK_E_NOMEM, "solve counters"};
    ctx->stats = static_cast<unsigned long long *>(p);
    hip_check(hipMemsetAsync(p, 0, kNumStats * sizeof(unsigned long long), ctx->stream), "solve counter reset");
}

TxDev make_tx(lpk_ctx *ctx, const lpk_tx_desc &d, uint32_t idx) {
    auto bad = [&](const char *why) {
        return Fail{LPK_E_INVALID, "transmitter " + std::to_string(idx) + ": " + why};
    };
    TxDev T{};
    T.traj_kind = d.traj_kind;
    T.signal_kind = d.signal_kind;
    if (d.traj_kind == LPK_TRAJ_STATIONARY) {
        for (int k = 0; k < 3; ++k) T.pos[k] = d.pos_ecef[k];
    } else if (d.traj_kind == LPK_TRAJ_GRID) {
        if (d.grid_n < 2 || !d.grid_xyz) throw bad("grid needs at least 2 nodes and a position table");
        if (d.grid_n > uint64_t(INT32_MAX)) throw bad("grid has more than 2^31 - 1 nodes");
        if (!(d.grid_dt_s > 0.0) || !is_fin(d.grid_dt_s) || !is_fin(d.grid_t0_unix))
            throw Fail{LPK_E_NONFINITE, "transmitter " + std::to_string(idx) + ": grid start/step"};
        T.grid_t0 = d.grid_t0_unix;
        T.grid_dt = d.grid_dt_s;
        T.grid_inv_dt = 1.0 / d.grid_dt_s;
        T.grid_n = int64_t(d.grid_n);
        T.grid_hi = static_cast<double>(T.grid_n - 2);
        T.grid = static_cast<const double *>(upload(ctx, d.grid_xyz, 3 * d.grid_n * sizeof(double)));
    } else {
        throw bad("unknown trajectory kind");
    }
    // horizon-gate bounds (src/geometry/precomputed.rs:86-106, src/channel/horizon.rs:164-170)
    if (d.traj_kind == LPK_TRAJ_STATIONARY) {
        T.max_speed = 0.0;
        T.max_radius = std::sqrt(d.pos_ecef[0] * d.pos_ecef[0] + d.pos_ecef[1] * d.pos_ecef[1] +
                                 d.pos_ecef[2] * d.pos_ecef[2]);
        T.gate_slack = 0.0;
    } else {
        const uint64_t n = d.grid_n;
        const double *g = d.grid_xyz;
        double seg2 = 0.0, r2 = 0.0;
        bool fin = true;
        for (uint64_t i = 0; i < n; ++i) {
            const double x = g[i], y = g[n + i], z = g[2 * n + i];
            const double q = x * x + y * y + z * z;
            fin = fin && std::isfinite(q);
            if (q > r2) r2 = q;
            if (i > 0) {
                const double dx = x - g[i - 1], dy = y - g[n + i - 1], dz = z - g[2 * n + i - 1];
                const double s2 = dx * dx + dy * dy + dz * dz;
                if (s2 > seg2) seg2 = s2;
            }
        }
        T.max_speed = fin ? std::sqrt(seg2) / d.grid_dt_s : NAN;
        T.max_radius = fin ? std::sqrt(r2) : NAN;
        T.gate_slack = 1.0;
    }
    T.gate_ok = std::isfinite(T.max_speed) && std::isfinite(T.max_radius) ? 1 : 0;
    if (!is_fin(d.carrier_hz) || !(d.carrier_hz > 0.0)) throw bad("carrier_hz must be positive and finite");
    if (!is_fin(d.tx_power_w) || d.tx_power_w < 0.0) throw bad("tx_power_w must be finite and >= 0");
    T.carrier = d.carrier_hz;
    // src/channel/mod.rs:208-210, :239-243
    const double wavelength_m = kSpeedOfLight / d.carrier_hz;
    const double tx_amplitude_v = std::sqrt(d.tx_power_w);
    T.neg_tau_src = -kTau * d.carrier_hz;
    T.amp_factor = (wavelength_m / (2.0 * kTau)) * tx_amplitude_v;
    switch (d.signal_kind) {
    case LPK_SIG_CW:
        T.omega = kTau * d.baseband_offset_hz;  // src/signal/cw.rs:42
        T.cw_anchored = d.cw_anchored ? 1 : 0;
        T.anchor_phase = d.anchor_phase_rad;
        break;
    case LPK_SIG_BPSK:
    case LPK_SIG_OOK:
    case LPK_SIG_QAM16:
        T.rate = d.rate_hz;
        T.epoch_rel = d.epoch_rel_s;
        T.repeating = d.repeating ? 1 : 0;
        T.repeat_from = d.repeat_from;
        T.n_symbols = int64_t(d.n_symbols);
        if (d.n_symbols) {
            if (d.signal_kind == LPK_SIG_QAM16) {
                if (!d.symbols) throw bad("qam16 needs its symbol table");
                T.symbols = static_cast<const float *>(upload(ctx, d.symbols, 2 * d.n_symbols * sizeof(float)));
            } else {
                if (!d.bits) throw bad("bpsk/ook needs its bit table");
                T.bits = static_cast<const int8_t *>(upload(ctx, d.bits, d.n_symbols));
            }
        }
        break;
    case LPK_SIG_CA:
        if (!d.ca_code) throw bad("ca_code needs its 1023-chip code");
        T.epoch_rel = d.epoch_rel_s;
        T.nav_mode = d.nav_mode;
        if (d.nav_mode < LPK_NAV_OFF || d.nav_mode > LPK_NAV_FULL) throw bad("unknown nav mode");
        T.ca_code = static_cast<const int8_t *>(upload(ctx, d.ca_code, 1023));
        {
            // the chips as f32, twice over: the precision form indexes it from
            // a chip residue plus an offset below one code period, unwrapped
            float chips2[2 * 1023];
            for (int c = 0; c < 2 * 1023; ++c) chips2[c] = static_cast<float>(d.ca_code[c % 1023]);
            T.ca_chips2 = static_cast<const float *>(upload(ctx, chips2, sizeof(chips2)));
        }
        T.nav_seed = d.nav_seed;
        if (d.nav_mode == LPK_NAV_FULL && d.n_nav_bits) {
            if (!d.nav_bits) throw bad("full nav needs its bit stream");
            T.nav_bits = static_cast<const int8_t *>(upload(ctx, d.nav_bits, d.n_nav_bits));
            T.n_nav_bits = int64_t(d.n_nav_bits);
        }
        break;
    default:
        throw bad("unknown signal kind");
    }
    if (d.gated) {
        // src/signal/gate.rs:30-52
        if (!(d.gate_period_s > 0.0) || !is_fin(d.gate_period_s)) throw bad("gate period must be positive");
        if (d.n_bursts && !d.bursts) throw bad("gate bursts missing");
        for (uint32_t b = 0; b < d.n_bursts; ++b) {
            const double s = d.bursts[2 * b], e = d.bursts[2 * b + 1];
            if (!(s >= 0.0) || !(e > s) || !(e <= d.gate_period_s)) throw bad("gate burst outside its period");
        }
        T.gated = 1;
        T.n_bursts = int32_t(d.n_bursts);
        T.gate_period = d.gate_period_s;
        T.gate_phase = d.gate_phase_s;
        T.bursts = static_cast<const double *>(upload(ctx, d.bursts, 2 * d.n_bursts * sizeof(double)));
    }
    if (d.has_kepler) {
        // CircularKeplerian (src/geometry/kepler.rs:19-36): radians and the mean motion
        if (!(d.kepler_a_m > 0.0) || !is_fin(d.kepler_a_m) || !is_fin(d.kepler_inc_deg) ||
            !is_fin(d.kepler_raan_deg) || !is_fin(d.kepler_nu0_deg) || !is_fin(d.kepler_epoch_unix))
            throw Fail{LPK_E_NONFINITE, "transmitter " + std::to_string(idx) + ": orbital elements"};
        const double rad = 3.141592653589793 / 180.0;
        T.kep_on = 1;
        T.kep_a = d.kepler_a_m;
        T.kep_cos_i = std::cos(d.kepler_inc_deg * rad);
        T.kep_sin_i = std::sin(d.kepler_inc_deg * rad);
        T.kep_raan = d.kepler_raan_deg * rad;
        T.kep_nu0 = d.kepler_nu0_deg * rad;
        T.kep_n = std::sqrt(kGmEarth / (d.kepler_a_m * d.kepler_a_m * d.kepler_a_m));
        T.kep_epoch = d.kepler_epoch_unix;
    }
    return T;
}

// Precision-form group of a transmitter (kGroup*).
int tx_group(const TxDev &T) {
    if (T.gated) return kGroupGeneric;
    if (T.signal_kind == LPK_SIG_CA) return kGroupCa;
    if (T.signal_kind == LPK_SIG_BPSK || T.signal_kind == LPK_SIG_OOK || T.signal_kind == LPK_SIG_QAM16)
        return kGroupFramed;
    if (T.signal_kind == LPK_SIG_CW && !T.cw_anchored) return kGroupCw;
    return kGroupGeneric;
}

void *pinned_slot(lpk_ctx *ctx, size_t bytes) {
    PinSlot &s = ctx->pin[ctx->pin_next];
    ctx->pin_next = (ctx->pin_next + 1) & 3;
    if (s.done) hip_check(hipEventSynchronize(s.done), "staging reuse");
    if (bytes > s.cap) {
        if (s.p) (void)hipHostFree(s.p);
        s.p = nullptr;
        s.cap = 0;
        size_t want = std::max<size_t>(bytes, 4096);
        if (hipHostMalloc(&s.p, want, 0) != hipSuccess) throw Fail{LPK_E_NOMEM, "pinned staging allocation failed"};
        s.cap = want;
    }
    if (!s.done) hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "event create");
    return s.p;
}
void pinned_mark(lpk_ctx *ctx, int slot_used) {
    hip_check(hipEventRecord(ctx->pin[slot_used].done, ctx->stream), "event record");
}

// Extend a receiver's flicker integral to cover samples up to (end) / fs of a
// band and upload it; returns the device table (null without flicker).
const double *ensure_flicker(lpk_ctx *ctx, RxState &R, uint64_t end, double fs) {
    FlickerGen &F = R.flick;
    if (!F.on) return nullptr;
    const int64_t k_hi = int64_t(std::floor(double(end) / fs * kFlickerRate)) + 2;
    if (int64_t(F.D.size()) <= k_hi || !F.dev.p) {
        // the previous upload may still read D; it is about to grow
        hip_check(hipStreamSynchronize(ctx->stream), "flicker table");
        flicker_extend(F, ((k_hi >> 10) + 1) << 10);
        void *d = F.dev.ensure(F.D.size() * sizeof(double), ctx->stream);
        hip_check(hipMemcpyAsync(d, F.D.data(), F.D.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
                  "flicker table upload");
        F.dev_n = F.D.size();
    }
    return static_cast<const double *>(F.dev.p);
}

// Grow a receiver's walk base table to hold base[0 .. b_hi + 1].
void ensure_base(lpk_ctx *ctx, RxState &R, int64_t b_hi) {
    if (b_hi + 2 <= R.base_cap) return;
    int64_t cap = std::max<int64_t>(R.base_cap ? R.base_cap : 4096, 4096);
    while (cap < b_hi + 2) cap *= 2;
    double *nb = nullptr;
    if (hipMalloc(&nb, size_t(cap) * sizeof(double)) != hipSuccess)
        throw Fail{LPK_E_NOMEM, "clock walk table allocation failed"};
    if (R.base) {
        hip_check(hipMemcpyAsync(nb, R.base, size_t(R.nb_known + 1) * sizeof(double), hipMemcpyDeviceToDevice,
                                 ctx->stream),
                  "walk table copy");
        hip_check(hipStreamSynchronize(ctx->stream), "walk table copy");
        (void)hipFree(R.base);
    } else {
        const double zero = 0.0;
        hip_check(hipMemcpy(nb, &zero, sizeof(double), hipMemcpyHostToDevice), "walk table init");
        R.nb_known = 0;
    }
    R.base = nb;
    R.base_cap = cap;
}

// Build, for every receiver with a window [wlo[r], whi[r]] of master indices,
// the walk blocks the render call needs, in two launches for all receivers:
// the block generator (block sums for every block from the known frontier or
// the window start, whichever is first, and the window's stored values) and
// the carry that extends each base table from its frontier.
void prepare_walks(lpk_ctx *ctx, const std::vector<int64_t> &wlo, const std::vector<int64_t> &whi) {
    std::vector<WalkSegDev> segs;
    std::vector<CarrySegDev> carries;
    std::vector<int64_t> bs_off, car_off;  // block-sum offsets of each segment and carry
    uint64_t blocks = 0;
    for (size_t r = 0; r < ctx->rx.size(); ++r) {
        if (whi[r] < 0) continue;
        RxState &R = ctx->rx[r];
        const int64_t b_lo = wlo[r] >> kWalkBlockShift, b_hi = whi[r] >> kWalkBlockShift;
        ensure_base(ctx, R, b_hi);
        const int64_t first = std::min(R.nb_known, b_lo);
        const int64_t n = b_hi + 1 - first, n_win = b_hi + 1 - b_lo;
        WalkSegDev S{};
        S.seed = R.d.clock_seed;
        S.sigma = R.sigma_step;
        S.first_block = first;
        S.win_block = b_lo;
        S.block0 = uint32_t(blocks);
        S.n_blocks = uint32_t(n);
        S.local = static_cast<float *>(R.local.ensure(size_t(n_win) * size_t(kWalkBlock) * sizeof(float), ctx->stream));
        S.toff = static_cast<double *>(
            R.toff.ensure(size_t(n_win) * size_t(kWalkBlock >> kWalkSubShift) * sizeof(double), ctx->stream));
        bs_off.push_back(int64_t(blocks));
        if (R.nb_known <= b_hi) {
            CarrySegDev Cs{};
            Cs.base = R.base;
            Cs.first = R.nb_known;
            Cs.n = b_hi + 1 - R.nb_known;
            carries.push_back(Cs);
            car_off.push_back(int64_t(blocks) + (R.nb_known - first));
            R.nb_known = b_hi + 1;
        }
        blocks += uint64_t(n);
        if (blocks > 0xFFFFFFull) throw Fail{LPK_E_INVALID, "clock walk windows too large for one launch"};
        segs.push_back(S);
    }
    if (segs.empty()) return;
    double *bs = static_cast<double *>(ctx->blocksum.ensure(size_t(blocks) * sizeof(double), ctx->stream));
    for (size_t q = 0; q < segs.size(); ++q) segs[q].blocksum = bs + bs_off[q];
    for (size_t q = 0; q < carries.size(); ++q) carries[q].blocksum = bs + car_off[q];
    const size_t seg_bytes = segs.size() * sizeof(WalkSegDev), car_bytes = carries.size() * sizeof(CarrySegDev);
    const int slot = ctx->pin_next;
    char *pin = static_cast<char *>(pinned_slot(ctx, seg_bytes + car_bytes));
    std::memcpy(pin, segs.data(), seg_bytes);
    if (car_bytes) std::memcpy(pin + seg_bytes, carries.data(), car_bytes);
    char *dev = static_cast<char *>(ctx->walk_tables.ensure(seg_bytes + car_bytes, ctx->stream));
    hip_check(hipMemcpyAsync(dev, pin, seg_bytes + car_bytes, hipMemcpyHostToDevice, ctx->stream), "walk table upload");
    pinned_mark(ctx, slot);
    hip_check(launch_walk_gen(reinterpret_cast<const WalkSegDev *>(dev), int(segs.size()), uint32_t(blocks), nullptr,
                              ctx->stream),
              "clock walk blocks");
    hip_check(launch_walk_carry(reinterpret_cast<const CarrySegDev *>(dev + seg_bytes), int(carries.size()),
                                ctx->stream),
              "clock walk carry");
}

// Master-index range [klo, khi] of the receiver clock that a job's samples
// interpolate (src/clock/mod.rs:329-346); false when the job uses no device walk.
bool job_window(const lpk_ctx *ctx, const lpk_band_chunk &B, uint64_t i0, uint64_t clen, int64_t &klo, int64_t &khi) {
    const RxState &R = ctx->rx[B.rx];
    if (!R.walk || B.delta_t || clen == 0) return false;
    const double M = R.d.master_rate_hz;
    const double xa = (double(i0) / B.fs_hz) * M;
    const double xb = (double(i0 + clen - 1) / B.fs_hz) * M;
    if (!(xb < 4.0e15)) throw Fail{LPK_E_INVALID, "clock index range too large"};
    // one index of margin on each side: the precision form's sample times may
    // sit an ulp from the division's
    klo = xa <= 1.0 ? 0 : int64_t(std::floor(xa)) - 1;
    khi = xb <= 0.0 ? 1 : int64_t(std::floor(xb)) + 2;
    return true;
}

// Checks one band-chunk descriptor against the scene and the descriptor
// contracts (src/pipeline.rs:759-767).
void validate_job(const lpk_ctx *ctx, const lpk_band_chunk &B, uint32_t q) {
    auto bad = [&](int code, const std::string &why) { return Fail{code, "job " + std::to_string(q) + ": " + why}; };
    if (B.rx >= ctx->rx.size()) throw bad(LPK_E_INVALID, "receiver index out of range");
    if (!(B.fs_hz > 0.0) || !is_fin(B.fs_hz)) throw bad(LPK_E_INVALID, "fs_hz must be positive and finite");
    if (!is_fin(B.lo_hz) || !is_fin(B.epoch_unix)) throw bad(LPK_E_NONFINITE, "lo_hz / epoch_unix");
    if (!is_fin(B.noise_power_w) || B.noise_power_w < 0.0) throw bad(LPK_E_INVALID, "noise_power_w");
    if (B.n_active && !B.active) throw bad(LPK_E_INVALID, "active list is NULL");
    if (B.clen > (uint64_t(1) << 40) || B.i0 > (uint64_t(1) << 52)) throw bad(LPK_E_INVALID, "chunk too large");
    for (uint32_t s = 0; s < B.n_active; ++s)
        if (B.active[s] >= ctx->n_tx) throw bad(LPK_E_INVALID, "active transmitter index out of range");
    if (B.n_post > 4) throw bad(LPK_E_INVALID, "at most 4 post-sum ops");
    for (uint32_t p = 0; p < B.n_post; ++p) {
        const lpk_post_op &P = B.post[p];
        if (P.kind < LPK_POST_GAIN || P.kind > LPK_POST_ADC) throw bad(LPK_E_INVALID, "unknown post op");
        // src/rx/mod.rs:93-97 takes any u32 bits with levels = (1u64 << (bits.max(2) - 1)) - 1, which
        // stays defined up to 64 bits; int16 codes need at most 16
        if (P.kind == LPK_POST_ADC && (P.bits < 0 || P.bits > 64)) throw bad(LPK_E_INVALID, "ADC bits must be 0..64");
        if (P.kind == LPK_POST_ADC && B.adc_codes_host && P.bits > 16)
            throw bad(LPK_E_INVALID, "ADC codes are int16: at most 16 bits when adc_codes_host is set");
        if (!is_fin(P.value)) throw bad(LPK_E_NONFINITE, "post op parameter");
    }
    if (B.delta_t)
        for (uint64_t j = 0; j < B.clen; ++j)
            if (!is_fin(B.delta_t[j])) throw bad(LPK_E_NONFINITE, "delta_t holds a non-finite value");
    if (B.has_antenna &&
        (!is_fin(B.antenna_peak_dbi) || !is_fin(B.antenna_horizon_dbi) || !is_fin(B.antenna_fb_db) ||
         (B.has_cutoff && !is_fin(B.antenna_cutoff_deg))))
        throw bad(LPK_E_NONFINITE, "antenna parameters");
}

// A streaming plan's precision form takes at most kPrecisionMaxLinks
// transmitters per band (a render call splits larger band-chunks into
// passes). The learned form's features assume a receiver fixed on the Earth;
// a moving receiver renders in the plain, converged or precision form.
void check_form_links(const lpk_ctx *ctx, int form, const lpk_band_chunk &B, uint32_t q) {
    if (form == LPK_FORM_LEARNED && B.rx < ctx->rx.size() && ctx->rx[B.rx].d.traj_kind != LPK_RX_STATIONARY)
        throw Fail{LPK_E_INVALID, "job " + std::to_string(q) +
                                      ": the learned form takes stationary receivers; use the plain, converged "
                                      "or precision form"};
}

// The precision form renders a band-chunk of more than kPrecisionMaxLinks
// transmitters in passes of at most that many, adding their f64 sums through
// JobDev::part: the number of passes of a job, and pass pp's job, which is
// the job's pp-th slice of links (group ends clipped to it).
uint32_t precision_passes(int form, uint32_t n_active) {
    return form == LPK_FORM_PRECISION && n_active > kPrecisionMaxLinks
               ? (n_active + kPrecisionMaxLinks - 1) / kPrecisionMaxLinks
               : 1u;
}
JobDev pass_slice(JobDev J, uint32_t pp, uint32_t np) {
    if (np <= 1) return J;
    const uint32_t lo = pp * kPrecisionMaxLinks;
    const uint32_t cnt = std::min(kPrecisionMaxLinks, J.n_active - lo);
    for (int g = 0; g < 3; ++g) J.group_end[g] = J.group_end[g] <= lo ? 0u : std::min(J.group_end[g] - lo, cnt);
    J.link0 += lo;
    J.n_active = cnt;
    J.pass = pp == 0 ? 1 : (pp + 1 == np ? 3 : 2);
    return J;
}

// Whether the receiver of a (validated) job moves: its render takes the
// kernels that evaluate the receiver per sample.
bool job_moving(const lpk_ctx *ctx, const lpk_band_chunk &B) {
    return ctx->rx[B.rx].d.traj_kind != LPK_RX_STATIONARY;
}

// The precision form takes a moving receiver in its antenna-free kernels
// only: a render call (or plan) with an antenna band renders every job in
// the antenna kernel, which takes fixed receivers.
void check_form_call(int form, bool any_moving, bool any_antenna) {
    if (form == LPK_FORM_PRECISION && any_moving && any_antenna)
        throw Fail{LPK_E_INVALID, "the precision form takes a moving receiver only in calls without an antenna "
                                  "band; use the plain or converged form"};
}

// The window-relative clock lookups of a job whose window starts at
// J.eps_k0 and reaches master index k_hi: used when every index fits 31 bits.
void set_walk_rel(JobDev &J, const double *base, int64_t k_hi) {
    J.eps_k0d = double(J.eps_k0);
    J.walk_base_w = base + (J.eps_k0 >> kWalkBlockShift);
    J.walk_rel = (k_hi - J.eps_k0 < (int64_t(1) << 31) - 2) ? 1 : 0;
}

// The job-table record of samples [i0, i0 + clen) of a band and its links
// (written at `lh`), apart from the tile range, link offset, output, clock
// window and flag pointers, which depend on the call. `perm` receives the
// order of the links: the reference's for the plain form, grouped by signal
// kind (stable within a group) for the precision form.
JobDev fill_job(const lpk_ctx *ctx, const lpk_band_chunk &B, uint64_t i0, uint64_t clen, int form, LinkDev *lh,
                std::vector<uint32_t> &perm) {
    const RxState &R = ctx->rx[B.rx];
    JobDev J{};
    J.n_active = B.n_active;
    perm.resize(B.n_active);
    for (uint32_t s = 0; s < B.n_active; ++s) perm[s] = s;
    if (form == LPK_FORM_PRECISION) {
        auto group = [&](uint32_t s) { return tx_group(ctx->tx_host[B.active[s]]); };
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return group(a) < group(b); });
        for (int g = 0; g < 3; ++g) {
            uint32_t e = 0;
            while (e < B.n_active && group(perm[e]) <= g) ++e;
            J.group_end[g] = e;
        }
    }
    for (uint32_t s = 0; s < B.n_active; ++s) {
        LinkDev L{};
        L.tx = B.active[perm[s]];
        L.tau_if = kTau * (ctx->tx_carrier[L.tx] - B.lo_hz);  // src/channel/mod.rs:240
        lh[s] = L;
    }
    if (B.has_antenna) {
        J.ant_on = 1;
        J.ant_cut = B.has_cutoff ? 1 : 0;
        J.ant_peak = B.antenna_peak_dbi;
        J.ant_horizon = B.antenna_horizon_dbi;
        J.ant_fb = B.antenna_fb_db;
        J.ant_cutoff = B.antenna_cutoff_deg;
    }
    J.fs = B.fs_hz;
    J.inv_fs = 1.0 / B.fs_hz;
    J.lo = B.lo_hz;
    J.epoch = B.epoch_unix;
    J.i0 = i0;
    J.clen = clen;
    for (int k = 0; k < 3; ++k) J.obs[k] = R.d.pos_ecef[k];
    if (R.d.traj_kind == LPK_RX_KEPLER) {
        // CircularKeplerian (src/geometry/kepler.rs:27-45): the per-receiver
        // constants in the reference's operation order, with glibc's sincos
        // for the inclination and node as its sin_cos pairs them
        const double deg = 3.141592653589793 / 180.0;   // f64::to_radians
        const double a = R.d.kepler_a_m;
        J.rx_kind = 1;
        J.rk_a = a;
        J.rk_n = std::sqrt(kGmEarth / (a * a * a));
        J.rk_nu0 = R.d.kepler_nu0_deg * deg;
        sincos(R.d.kepler_inc_deg * deg, &J.rk_si, &J.rk_ci);
        sincos(R.d.kepler_raan_deg * deg, &J.rk_so, &J.rk_co);
        J.rk_epoch = R.d.kepler_epoch_unix;
        // the precision form's smooth model: E = epoch - rk_epoch, nu0 + n E
        // and omega E exactly, as double-doubles (error-free sums and products)
        auto two_sum = [](double a, double b, double *o) {
            const double s = a + b, bb = s - a;
            o[0] = s;
            o[1] = (a - (s - bb)) + (b - bb);
        };
        auto scaled = [&](double k, double add, double *o) {   // add + k E
            const double p = k * J.rk_e[0], pe = std::fma(k, J.rk_e[0], -p);
            double s[2];
            two_sum(add, p, s);
            const double lo = s[1] + (pe + k * J.rk_e[1]);
            o[0] = s[0] + lo;
            o[1] = lo - (o[0] - s[0]);
        };
        const double w = J.rk_n + kEarthRotation;
        J.rx_v = a * w;
        J.rx_a = a * w * w;
        J.rx_j = a * w * w * w;
        two_sum(J.epoch, -J.rk_epoch, J.rk_e);
        scaled(J.rk_n, J.rk_nu0, J.rk_nue);
        scaled(kEarthRotation, 0.0, J.rk_the);
    }
    J.c0 = R.d.initial_time_offset_s;
    J.y0 = R.d.frequency_offset_y0;
    J.master = R.d.master_rate_hz;
    J.walk = (R.walk && !B.delta_t) ? 1 : 0;
    J.sigma = float(std::sqrt(B.noise_power_w / 2.0));  // src/channel/mod.rs:328
    J.noise_seed = B.noise_seed;
    J.n_post = int32_t(B.n_post);
    for (uint32_t p = 0; p < B.n_post; ++p) {
        const lpk_post_op &P = B.post[p];
        PostDev D{};
        D.kind = P.kind;
        if (P.kind == LPK_POST_GAIN) {
            D.a = std::pow(10.0f, float(P.value / 20.0));  // src/rx/mod.rs:86
        } else if (P.kind == LPK_POST_NONLINEARITY) {
            D.a = float(4.0 / 3.0 / P.value);  // src/rx/mod.rs:90
        } else {
            const float fsv = float(P.value);  // src/rx/mod.rs:94-96
            const int bits = P.bits < 2 ? 2 : P.bits;
            const float levels = float((1ULL << (bits - 1)) - 1);
            D.a = fsv;
            D.b = fsv / levels;
        }
        J.post[p] = D;
    }
    return J;
}

uint32_t tiles_of(uint64_t i0, uint64_t clen, int tile_shift) {
    return clen ? uint32_t(((i0 + clen - 1) >> tile_shift) - (i0 >> tile_shift) + 1) : 0u;
}

struct TapReq {
    uint32_t slot;
    double *t_emit, *range, *phase, *delta_t;
};

// Next timing event pair of the context (bench accounting, lpk_timing_*).
void timing_begin(lpk_ctx *ctx, hipStream_t s = nullptr) {
    if (ctx->tused == ctx->tpool.size() && ctx->tused < 65536) {
        std::pair<hipEvent_t, hipEvent_t> pr{nullptr, nullptr};
        hip_check(hipEventCreate(&pr.first), "event create");
        hip_check(hipEventCreate(&pr.second), "event create");
        ctx->tpool.push_back(pr);
    }
    if (ctx->tused < ctx->tpool.size()) {
        ctx->ev0 = ctx->tpool[ctx->tused].first;
        ctx->ev1 = ctx->tpool[ctx->tused].second;
        ++ctx->tused;
    }
    hip_check(hipEventRecord(ctx->ev0, s ? s : ctx->stream), "event record");
}
void timing_end(lpk_ctx *ctx, hipStream_t s = nullptr) {
    hip_check(hipEventRecord(ctx->ev1, s ? s : ctx->stream), "event record");
    ctx->timed = true;
}

void render(lpk_ctx *ctx, const lpk_band_chunk *jobs, uint32_t n_jobs, int form, const TapReq *tap) {
    if (!ctx->have_scene) throw Fail{LPK_E_STATE, "no scene uploaded"};
    if (form < LPK_FORM_PLAIN || form > LPK_FORM_LEARNED) throw Fail{LPK_E_INVALID, "unknown form"};
    if (form == LPK_FORM_LEARNED && !ctx->lws) throw Fail{LPK_E_STATE, "the learned form needs a network (lpk_set_warm_start)"};
    if (n_jobs && !jobs) throw Fail{LPK_E_INVALID, "jobs is NULL"};
    // ---- validation
    uint64_t total_links = 0, dt_total = 0, stage_total = 0, code_total = 0;
    bool want_flags = false, any_antenna = false, any_moving = false;
    for (uint32_t q = 0; q < n_jobs; ++q) {
        const lpk_band_chunk &B = jobs[q];
        validate_job(ctx, B, q);
        check_form_links(ctx, form, B, q);
        if (B.unoccluded && B.n_active) want_flags = true;
        if (B.has_antenna) any_antenna = true;
        if (job_moving(ctx, B)) any_moving = true;
        total_links += B.n_active;
        if (B.delta_t) dt_total += B.clen;
        if (!B.out_dev) stage_total += B.clen;
        if (B.adc_codes_host) code_total += B.clen;
    }
    check_form_call(form, any_moving, any_antenna);
    if (tap && n_jobs != 1) throw Fail{LPK_E_INVALID, "taps need exactly one job"};
    if (tap && tap->slot >= jobs[0].n_active) throw Fail{LPK_E_INVALID, "tap slot out of range"};
    const int ts = ctx->tile_shift;

    // ---- order: heavy jobs first so their tiles are dispatched first
    std::vector<uint32_t> order(n_jobs);
    for (uint32_t q = 0; q < n_jobs; ++q) order[q] = q;
    if (!tap)
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return jobs[a].n_active > jobs[b].n_active; });

    // ---- clock walk windows per receiver
    std::vector<int64_t> wlo(ctx->rx.size(), INT64_MAX), whi(ctx->rx.size(), -1);
    for (uint32_t q = 0; q < n_jobs; ++q) {
        const lpk_band_chunk &B = jobs[q];
        int64_t klo, khi;
        if (!job_window(ctx, B, B.i0, B.clen, klo, khi)) continue;
        wlo[B.rx] = std::min(wlo[B.rx], klo);
        whi[B.rx] = std::max(whi[B.rx], khi);
    }
    prepare_walks(ctx, wlo, whi);

    // ---- staging buffers
    double *dt_dev = dt_total ? static_cast<double *>(ctx->dt_stage.ensure(dt_total * 8, ctx->stream)) : nullptr;
    float *stage_dev = stage_total ? static_cast<float *>(ctx->out_stage.ensure(stage_total * 8, ctx->stream)) : nullptr;
    int16_t *codes_dev = code_total ? static_cast<int16_t *>(ctx->codes_stage.ensure(code_total * 4, ctx->stream)) : nullptr;

    // ---- passes: the precision form renders a band-chunk of more than
    // kPrecisionMaxLinks transmitters in passes of at most that many, the
    // passes adding their f64 sums through HBM (JobDev::part)
    std::vector<uint32_t> n_pass(n_jobs, 1);
    uint32_t max_pass = 1;
    uint64_t n_entries = 0, part_total = 0;
    for (uint32_t q = 0; q < n_jobs; ++q) {
        n_pass[q] = precision_passes(form, jobs[q].n_active);
        if (n_pass[q] > 1) part_total += jobs[q].clen;
        max_pass = std::max(max_pass, n_pass[q]);
        n_entries += n_pass[q];
    }
    double2 *part_dev =
        part_total ? static_cast<double2 *>(ctx->part_stage.ensure(part_total * sizeof(double2), ctx->stream)) : nullptr;

    // ---- job and link tables
    const size_t job_bytes = size_t(std::max<uint64_t>(n_entries, 1)) * sizeof(JobDev);
    const size_t link_bytes = size_t(std::max<uint64_t>(total_links, 1)) * sizeof(LinkDev);
    const size_t flag_bytes = size_t(std::max<uint64_t>(total_links, 1)) * sizeof(uint32_t);
    const int slot = ctx->pin_next;
    char *pin = static_cast<char *>(pinned_slot(ctx, job_bytes + link_bytes + flag_bytes));
    JobDev *jh = reinterpret_cast<JobDev *>(pin);
    LinkDev *lh = reinterpret_cast<LinkDev *>(pin + job_bytes);
    std::memset(pin + job_bytes + link_bytes, 0, flag_bytes);
    JobDev *jd = static_cast<JobDev *>(ctx->jobs.ensure(job_bytes, ctx->stream));
    LinkDev *ld = static_cast<LinkDev *>(ctx->links.ensure(link_bytes + flag_bytes, ctx->stream));
    uint32_t *fd = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(ld) + link_bytes);
    struct FlagJob {
        uint32_t q, link0;
        std::vector<uint32_t> perm;
    };
    std::vector<FlagJob> flag_jobs;

    uint64_t tiles = 0, link_off = 0, dt_off = 0, st_off = 0, cd_off = 0, n_pred = 0, part_off = 0;
    uint32_t tap_slot = tap ? tap->slot : 0;
    std::vector<uint32_t> perm;
    std::vector<JobDev> whole(n_jobs);   // in `order`
    std::vector<float *> out_ptr(n_jobs, nullptr);
    std::vector<int16_t *> code_ptr(n_jobs, nullptr);
    for (uint32_t o = 0; o < n_jobs; ++o) {
        const uint32_t q = order[o];
        const lpk_band_chunk &B = jobs[q];
        const RxState &R = ctx->rx[B.rx];
        JobDev J = fill_job(ctx, B, B.i0, B.clen, form, lh + link_off, perm);
        J.flicker = B.delta_t ? nullptr : ensure_flicker(ctx, ctx->rx[B.rx], B.i0 + B.clen, B.fs_hz);
        J.tile0 = uint32_t(tiles);
        J.n_tiles = tiles_of(B.i0, B.clen, span_shift(form, any_antenna, ts));
        tiles += J.n_tiles;
        if (tiles > 0xFFFFFFFFull) throw Fail{LPK_E_INVALID, "render call holds more than 2^32 tiles"};
        J.link0 = uint32_t(link_off);
        J.pred0 = n_pred;
        J.pred_qs = 1;
        J.pred_nb = 0;
        if (form == LPK_FORM_LEARNED && J.n_tiles) {
            const double per_tile = double(1u << ts) / B.fs_hz;
            J.pred_qs = uint32_t(std::max(1.0, std::floor(kLwsSpacing / per_tile)));
            J.pred_nb = (J.n_tiles + J.pred_qs - 1) / J.pred_qs + 1;
            n_pred += uint64_t(B.n_active) * J.pred_nb;
        }
        if (tap)
            for (uint32_t s = 0; s < B.n_active; ++s)
                if (perm[s] == tap->slot) tap_slot = s;
        if (B.has_antenna) J.unocc = fd;
        if (B.unoccluded && B.n_active) flag_jobs.push_back(FlagJob{q, uint32_t(link_off), perm});
        link_off += B.n_active;
        if (B.delta_t) {
            hip_check(hipMemcpyAsync(dt_dev + dt_off, B.delta_t, B.clen * 8, hipMemcpyHostToDevice, ctx->stream),
                      "delta_t upload");
            J.delta_t = dt_dev + dt_off;
            dt_off += B.clen;
        }
        if (J.walk) {
            J.walk_base = R.base;
            J.walk_toff = static_cast<const double *>(R.toff.p);
            J.walk_local = static_cast<const float *>(R.local.p);
            J.eps_k0 = (wlo[B.rx] >> kWalkBlockShift) << kWalkBlockShift;
            set_walk_rel(J, R.base, whi[B.rx]);
        }
        if (B.out_dev) {
            J.out = static_cast<float *>(B.out_dev);
        } else {
            J.out = stage_dev + 2 * st_off;
            st_off += B.clen;
        }
        out_ptr[q] = J.out;
        if (B.adc_codes_host) {
            J.codes = codes_dev + 2 * cd_off;
            cd_off += B.clen;
            code_ptr[q] = J.codes;
        }
        if (n_pass[q] > 1) {
            J.part = part_dev + part_off;
            part_off += B.clen;
        }
        whole[o] = J;
    }
    // pass p: the p-th slice of links of every job with more than p passes
    // (every job in pass 0), each pass's tiles numbered from 0
    std::vector<uint32_t> pass_jobs(max_pass, 0);
    std::vector<uint64_t> pass_tiles(max_pass, 0);
    {
        uint64_t e = 0;
        for (uint32_t pp = 0; pp < max_pass; ++pp) {
            for (uint32_t o = 0; o < n_jobs; ++o) {
                const uint32_t np = n_pass[order[o]];
                if (pp >= np) continue;
                JobDev J = pass_slice(whole[o], pp, np);
                J.tile0 = uint32_t(pass_tiles[pp]);
                pass_tiles[pp] += J.n_tiles;
                if (pass_tiles[pp] > 0xFFFFFFFFull) throw Fail{LPK_E_INVALID, "render call holds more than 2^32 tiles"};
                jh[e++] = J;
                ++pass_jobs[pp];
            }
        }
    }
    hip_check(hipMemcpyAsync(jd, jh, job_bytes, hipMemcpyHostToDevice, ctx->stream), "job table upload");
    (void)tiles;
    hip_check(hipMemcpyAsync(ld, lh, link_bytes + flag_bytes, hipMemcpyHostToDevice, ctx->stream),
              "link table upload");
    pinned_mark(ctx, slot);

    if (tiles) {
        const bool converged = form == LPK_FORM_CONVERGED || form == LPK_FORM_LEARNED;
        ensure_stats(ctx);
        double *pred = n_pred ? static_cast<double *>(ctx->pred.ensure(n_pred * 8, ctx->stream)) : nullptr;
        timing_begin(ctx);
        if (pred)
            hip_check(launch_lws_predict(jd, int(n_jobs), ld, ctx->tx_dev, lws_net(ctx->lws), ts, n_pred, pred,
                                         ctx->stats, ctx->stream),
                      "warm-start prediction launch");
        uint64_t e0 = 0;
        for (uint32_t pp = 0; pp < max_pass; ++pp) {
            // the tapped link is in pass tap_slot / kPrecisionMaxLinks of a job in passes
            uint32_t slot_p = tap_slot;
            if (max_pass > 1)
                slot_p = tap_slot / kPrecisionMaxLinks == pp ? tap_slot % kPrecisionMaxLinks : 0xFFFFFFFFu;
            if (pass_tiles[pp])
                hip_check(launch_render(converged ? 2 : form, any_antenna, any_moving, ts, ctx->xcd_map != 0, jd + e0,
                                        int(pass_jobs[pp]), ld, ctx->tx_dev, uint32_t(pass_tiles[pp]), nullptr,
                                        tap ? tap->t_emit : nullptr, tap ? tap->range : nullptr,
                                        tap ? tap->phase : nullptr, tap ? tap->delta_t : nullptr, slot_p, pred,
                                        converged ? ctx->stats : nullptr, ctx->gate != 0,
                                        any_antenna ? ctx->stats + 4 : nullptr, ctx->stream),
                          "render launch");
            e0 += pass_jobs[pp];
        }
        timing_end(ctx);
    }
    if (want_flags) {
        // src/channel/mod.rs:208-211: with no antenna, every sample counts
        std::vector<uint32_t> fl(total_links);
        hip_check(hipMemcpyAsync(fl.data(), fd, total_links * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
                  "occlusion flag download");
        hip_check(hipStreamSynchronize(ctx->stream), "occlusion flag download");
        for (const FlagJob &F : flag_jobs) {
            const lpk_band_chunk &B = jobs[F.q];
            for (uint32_t s = 0; s < B.n_active; ++s)
                B.unoccluded[F.perm[s]] = B.has_antenna ? (fl[F.link0 + s] ? 1 : 0) : (B.clen > 0 ? 1 : 0);
        }
    }
    for (uint32_t q = 0; q < n_jobs; ++q) {
        const lpk_band_chunk &B = jobs[q];
        if (B.out_host && B.clen)
            hip_check(hipMemcpyAsync(B.out_host, out_ptr[q], B.clen * 8, hipMemcpyDeviceToHost, ctx->stream),
                      "output download");
        if (B.adc_codes_host && B.clen)
            hip_check(hipMemcpyAsync(B.adc_codes_host, code_ptr[q], B.clen * 4, hipMemcpyDeviceToHost, ctx->stream),
                      "ADC code download");
    }
}

// --------------------------------------------------------------- plans

// First sample of band b in step k: steps are time slices, chunk_samples of
// the band with the highest sample rate and, for every other band, its
// samples in the same interval (sample j of band b sits at j / fs_b, so the
// slice [k T, (k + 1) T) with T = chunk / fs_max starts at the first j with
// j >= k chunk fs_b / fs_max). Each receiver's clock window per step is then
// that interval, so the walk is generated about once over a pass.
uint64_t plan_start(const lpk_plan *P, uint32_t b, uint64_t k) {
    const uint64_t n = P->bands[b].clen;
    const double s = std::ceil(double(k) * double(P->chunk) * P->rate_ratio[b]);
    return s >= double(n) ? n : uint64_t(s);
}

// Samples [i0, i0 + clen) of band b in step k.
void plan_chunk(const lpk_plan *P, uint32_t b, uint64_t k, uint64_t &i0, uint64_t &clen) {
    i0 = plan_start(P, b, k);
    clen = plan_start(P, b, k + 1) - i0;
}

// Union of the master-index windows of a receiver's bands in chunk k.
bool plan_window(const lpk_plan *P, uint32_t r, uint64_t k, int64_t &lo, int64_t &hi) {
    bool any = false;
    lo = INT64_MAX;
    hi = -1;
    for (uint32_t b = 0; b < P->bands.size(); ++b) {
        if (P->bands[b].rx != r) continue;
        uint64_t i0, clen;
        plan_chunk(P, b, k, i0, clen);
        int64_t klo, khi;
        if (!job_window(P->ctx, P->bands[b], i0, clen, klo, khi)) continue;
        lo = std::min(lo, klo);
        hi = std::max(hi, khi);
        any = true;
    }
    return any;
}

// Fills the pinned table of step k: header (live walk blocks, each pass's
// live tiles), jobs (pass by pass), links, zeroed flags, walk segments and
// carries.
void plan_fill(lpk_plan *P, uint64_t k, int slot, char *dst) {
    lpk_ctx *ctx = P->ctx;
    char *dev = static_cast<char *>(P->tables[slot].p);
    uint32_t *hdr = reinterpret_cast<uint32_t *>(dst);
    JobDev *jh = reinterpret_cast<JobDev *>(dst + P->off_jobs);
    LinkDev *lh = reinterpret_cast<LinkDev *>(dst + P->off_links);
    std::memset(dst + P->off_flags, 0, P->total_links * sizeof(uint32_t));
    uint32_t *flags_dev = reinterpret_cast<uint32_t *>(dev + P->off_flags);
    WalkSegDev *sh = reinterpret_cast<WalkSegDev *>(dst + P->off_segs);
    CarrySegDev *ch = reinterpret_cast<CarrySegDev *>(dst + P->off_carry);
    double *bs = static_cast<double *>(P->blocksum[slot].p);
    // clock walks
    std::vector<int64_t> k0(P->walk_rx.size(), 0), k1(P->walk_rx.size(), 0);
    uint64_t blocks = 0;
    for (size_t w = 0; w < P->walk_rx.size(); ++w) {
        RxState &R = ctx->rx[P->walk_rx[w]];
        WalkSegDev S{};
        S.seed = R.d.clock_seed;
        S.sigma = R.sigma_step;
        S.block0 = uint32_t(blocks);
        S.local = static_cast<float *>(P->local[slot][w].p);
        S.toff = static_cast<double *>(P->toff[slot][w].p);
        S.blocksum = bs + blocks;
        CarrySegDev C{};
        C.base = R.base;
        C.blocksum = bs;
        int64_t lo, hi;
        if (plan_window(P, P->walk_rx[w], k, lo, hi)) {
            const int64_t b_lo = lo >> kWalkBlockShift, b_hi = hi >> kWalkBlockShift;
            const int64_t first = std::min(R.nb_known, b_lo);
            if (b_hi + 1 - b_lo > P->win_cap[w] || b_hi + 2 > R.base_cap)
                throw Fail{LPK_E_STATE, "clock window beyond the plan's capacity"};
            S.first_block = first;
            S.win_block = b_lo;
            S.n_blocks = uint32_t(b_hi + 1 - first);
            k0[w] = b_lo << kWalkBlockShift;
            k1[w] = hi;
            if (R.nb_known <= b_hi) {
                C.first = R.nb_known;
                C.n = b_hi + 1 - R.nb_known;
                C.blocksum = bs + blocks + (R.nb_known - first);
                R.nb_known = b_hi + 1;
            }
            blocks += S.n_blocks;
        }
        sh[w] = S;
        ch[w] = C;
    }
    if (blocks > P->max_blocks)
        throw Fail{LPK_E_STATE, "clock walk of this step beyond the plan's capacity (clock rewound during the plan?)"};
    // jobs and links: every band's job, then the passes' slices of them
    std::vector<uint32_t> perm;
    std::vector<JobDev> whole(P->bands.size());
    for (uint32_t b : P->order) {
        const lpk_band_chunk &B = P->bands[b];
        uint64_t i0, clen;
        plan_chunk(P, b, k, i0, clen);
        JobDev J = fill_job(ctx, B, i0, clen, P->form, lh + P->link0[b], perm);
        J.flicker = B.delta_t ? nullptr : static_cast<const double *>(ctx->rx[B.rx].flick.dev.p);
        J.n_tiles = tiles_of(i0, clen, span_shift(P->form, P->antenna, P->tile_shift));
        J.link0 = uint32_t(P->link0[b]);
        if (B.has_antenna) J.unocc = flags_dev;
        if (J.walk) {
            const int w = P->walk_slot[B.rx];
            J.walk_base = ctx->rx[B.rx].base;
            J.walk_toff = static_cast<const double *>(P->toff[slot][w].p);
            J.walk_local = static_cast<const float *>(P->local[slot][w].p);
            J.eps_k0 = k0[w];
            set_walk_rel(J, ctx->rx[B.rx].base, k1[w]);
        }
        J.out = B.out_dev ? static_cast<float *>(B.out_dev) + 2 * i0 : static_cast<float *>(P->outbuf[slot][b].p);
        if (P->n_pass[b] > 1) J.part = static_cast<double2 *>(P->part[slot].p) + P->part_off[b];
        whole[b] = J;
    }
    for (uint32_t pp = 0; pp < P->pass_jobs.size(); ++pp) {
        uint64_t tiles = 0;
        for (uint32_t b : P->order) {
            if (pp >= P->n_pass[b]) continue;
            JobDev J = pass_slice(whole[b], pp, P->n_pass[b]);
            J.tile0 = uint32_t(tiles);
            tiles += J.n_tiles;
            *jh++ = J;
        }
        if (tiles > P->pass_tiles[pp]) throw Fail{LPK_E_STATE, "step beyond the plan's tile capacity"};
        hdr[kPlanPassHdr + pp] = uint32_t(tiles);
    }
    hdr[1] = uint32_t(blocks);
}

void plan_free(lpk_plan *P) {
    if (!P) return;
    if (P->ctx) {
        (void)hipStreamSynchronize(P->ctx->stream);
        auto &v = P->ctx->plans;
        v.erase(std::remove(v.begin(), v.end(), P), v.end());
    }
    if (P->aux) (void)hipStreamSynchronize(P->aux);
    for (int i = 0; i < 2; ++i) {
        if (P->done[i]) (void)hipEventDestroy(P->done[i]);
        if (P->carry_ev[i]) (void)hipEventDestroy(P->carry_ev[i]);
        if (P->pin[i]) (void)hipHostFree(P->pin[i]);
        for (auto &d : P->local[i]) d.release();
        for (auto &d : P->toff[i]) d.release();
        for (auto &d : P->outbuf[i]) d.release();
        P->part[i].release();
        P->blocksum[i].release();
        P->tables[i].release();
    }
    if (P->ctx_ev) (void)hipEventDestroy(P->ctx_ev);
    if (P->aux) (void)hipStreamDestroy(P->aux);
    delete P;
}

lpk_plan *plan_create(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, uint64_t chunk, int form) {
    if (!ctx->have_scene) throw Fail{LPK_E_STATE, "no scene uploaded"};
    if (form != LPK_FORM_PLAIN && form != LPK_FORM_PRECISION)
        throw Fail{LPK_E_INVALID, "plans take the plain or the precision form"};
    if (!n_bands || !bands) throw Fail{LPK_E_INVALID, "a plan needs bands"};
    if (chunk == 0 || chunk > (uint64_t(1) << 32)) throw Fail{LPK_E_INVALID, "chunk_samples out of range"};
    std::unique_ptr<lpk_plan, void (*)(lpk_plan *)> P(new lpk_plan, plan_free);
    P->ctx = ctx;
    P->form = form;
    P->tile_shift = ctx->tile_shift;
    P->xcd_map = ctx->xcd_map;
    P->chunk = chunk;
    P->bands.assign(bands, bands + n_bands);
    P->active.resize(n_bands);
    for (uint32_t b = 0; b < n_bands; ++b) {
        lpk_band_chunk &B = P->bands[b];
        validate_job(ctx, B, b);
        check_form_links(ctx, form, B, b);
        if (B.i0 != 0 || B.delta_t || B.out_host || B.adc_codes_host || B.unoccluded)
            throw Fail{LPK_E_INVALID, "band " + std::to_string(b) +
                                          ": a plan band starts at sample 0, uses the device clock and has "
                                          "device outputs only"};
        P->active[b].assign(B.active, B.active + B.n_active);
        B.active = P->active[b].data();
        P->antenna = P->antenna || B.has_antenna;
        P->moving = P->moving || job_moving(ctx, B);
    }
    check_form_call(form, P->moving, P->antenna);
    double fs_max = 0.0;
    for (const lpk_band_chunk &B : P->bands) fs_max = std::max(fs_max, B.fs_hz);
    P->rate_ratio.resize(n_bands);
    for (uint32_t b = 0; b < n_bands; ++b) P->rate_ratio[b] = P->bands[b].fs_hz / fs_max;   // 1 for the fastest
    // the steps until every band's last sample is rendered
    for (uint32_t b = 0; b < n_bands; ++b) {
        uint64_t k = uint64_t(std::ceil(double(P->bands[b].clen) / (double(chunk) * P->rate_ratio[b])));
        while (k > 0 && plan_start(P.get(), b, k - 1) >= P->bands[b].clen) --k;
        while (plan_start(P.get(), b, k) < P->bands[b].clen) ++k;
        P->n_chunks = std::max(P->n_chunks, k);
    }
    P->chunk_cap.assign(n_bands, 0);
    P->order.resize(n_bands);
    for (uint32_t b = 0; b < n_bands; ++b) P->order[b] = b;
    std::stable_sort(P->order.begin(), P->order.end(),
                     [&](uint32_t a, uint32_t b) { return P->bands[a].n_active > P->bands[b].n_active; });
    P->link0.assign(n_bands, 0);
    for (uint32_t b : P->order) {
        P->link0[b] = P->total_links;
        P->total_links += P->bands[b].n_active;
    }
    P->n_pass.assign(n_bands, 1);
    for (uint32_t b = 0; b < n_bands; ++b) {
        P->n_pass[b] = precision_passes(form, P->bands[b].n_active);
        if (P->n_pass[b] > P->pass_jobs.size()) P->pass_jobs.resize(P->n_pass[b], 0);
    }
    if (P->pass_jobs.size() > kPlanMaxPasses)
        throw Fail{LPK_E_INVALID, "a plan band holds more than " + std::to_string(kPlanMaxPasses * kPrecisionMaxLinks) +
                                      " transmitters in the precision form"};
    for (uint32_t b = 0; b < n_bands; ++b)
        for (uint32_t pp = 0; pp < P->n_pass[b]; ++pp) ++P->pass_jobs[pp];
    P->pass_tiles.assign(P->pass_jobs.size(), 0);
    // flicker tables over every band's whole extent (steps only read them)
    for (const lpk_band_chunk &B : P->bands)
        if (!B.delta_t) ensure_flicker(ctx, ctx->rx[B.rx], B.i0 + B.clen, B.fs_hz);
    // the receivers the plan walks, and a dry run of every step for capacities
    P->walk_slot.assign(ctx->rx.size(), -1);
    for (const auto &B : P->bands)
        if (ctx->rx[B.rx].walk && P->walk_slot[B.rx] < 0) {
            P->walk_slot[B.rx] = int(P->walk_rx.size());
            P->walk_rx.push_back(B.rx);
        }
    const size_t nw = P->walk_rx.size();
    std::vector<int64_t> nb(nw, 0), max_bhi(nw, -1);
    P->win_cap.assign(nw, 0);
    for (uint64_t k = 0; k < P->n_chunks; ++k) {
        uint64_t blocks = 0;
        std::vector<uint64_t> tiles(P->pass_jobs.size(), 0);
        for (size_t w = 0; w < nw; ++w) {
            int64_t lo, hi;
            if (!plan_window(P.get(), P->walk_rx[w], k, lo, hi)) continue;
            const int64_t b_lo = lo >> kWalkBlockShift, b_hi = hi >> kWalkBlockShift;
            blocks += uint64_t(b_hi + 1 - std::min(nb[w], b_lo));
            P->win_cap[w] = std::max(P->win_cap[w], b_hi + 1 - b_lo);
            max_bhi[w] = std::max(max_bhi[w], b_hi);
            if (nb[w] <= b_hi) nb[w] = b_hi + 1;
        }
        for (uint32_t b = 0; b < n_bands; ++b) {
            uint64_t i0, clen;
            plan_chunk(P.get(), b, k, i0, clen);
            const uint64_t nt = tiles_of(i0, clen, span_shift(P->form, P->antenna, P->tile_shift));
            for (uint32_t pp = 0; pp < P->n_pass[b]; ++pp) tiles[pp] += nt;
            P->chunk_cap[b] = std::max(P->chunk_cap[b], clen);
        }
        if (blocks > 0xFFFFFFull || tiles[0] > 0xFFFFFFFFull) throw Fail{LPK_E_INVALID, "plan step too large"};
        P->max_blocks = std::max<uint32_t>(P->max_blocks, uint32_t(blocks));
        for (size_t pp = 0; pp < tiles.size(); ++pp)
            P->pass_tiles[pp] = std::max<uint32_t>(P->pass_tiles[pp], uint32_t(tiles[pp]));
    }
    // the passes' staging: every band in passes, its largest chunk
    uint64_t part_total = 0;
    P->part_off.assign(n_bands, 0);
    for (uint32_t b = 0; b < n_bands; ++b)
        if (P->n_pass[b] > 1) {
            P->part_off[b] = part_total;
            part_total += P->chunk_cap[b];
        }
    // device buffers, one set per slot
    for (size_t w = 0; w < nw; ++w) ensure_base(ctx, ctx->rx[P->walk_rx[w]], max_bhi[w]);
    for (int i = 0; i < 2; ++i) {
        P->local[i].resize(nw);
        P->toff[i].resize(nw);
        for (size_t w = 0; w < nw; ++w) {
            const size_t cap = size_t(std::max<int64_t>(P->win_cap[w], 1));
            P->local[i][w].ensure(cap * size_t(kWalkBlock) * sizeof(float), ctx->stream);
            P->toff[i][w].ensure(cap * size_t(kWalkBlock >> kWalkSubShift) * sizeof(double), ctx->stream);
        }
        P->blocksum[i].ensure(size_t(std::max<uint32_t>(P->max_blocks, 1)) * sizeof(double), ctx->stream);
        if (part_total) P->part[i].ensure(size_t(part_total) * sizeof(double2), ctx->stream);
        P->outbuf[i].resize(n_bands);
        for (uint32_t b = 0; b < n_bands; ++b)
            if (!P->bands[b].out_dev)
                P->outbuf[i][b].ensure(size_t(std::max<uint64_t>(P->chunk_cap[b], 1)) * 8, ctx->stream);
    }
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    P->off_jobs = 256;
    uint64_t n_entries = 0;
    for (uint32_t b = 0; b < n_bands; ++b) n_entries += P->n_pass[b];
    P->off_links = align(P->off_jobs + n_entries * sizeof(JobDev));
    P->off_flags = align(P->off_links + std::max<uint64_t>(P->total_links, 1) * sizeof(LinkDev));
    P->off_segs = align(P->off_flags + std::max<uint64_t>(P->total_links, 1) * sizeof(uint32_t));
    P->off_carry = align(P->off_segs + std::max<size_t>(nw, 1) * sizeof(WalkSegDev));
    P->bytes = align(P->off_carry + std::max<size_t>(nw, 1) * sizeof(CarrySegDev));
    for (int i = 0; i < 2; ++i) {
        P->tables[i].ensure(P->bytes, ctx->stream);
        if (hipHostMalloc(&P->pin[i], P->bytes, 0) != hipSuccess) throw Fail{LPK_E_NOMEM, "plan staging allocation"};
        std::memset(P->pin[i], 0, P->bytes);
        hip_check(hipEventCreateWithFlags(&P->done[i], hipEventDisableTiming), "event create");
        hip_check(hipEventCreateWithFlags(&P->carry_ev[i], hipEventDisableTiming), "event create");
    }
    hip_check(hipEventCreateWithFlags(&P->ctx_ev, hipEventDisableTiming), "event create");
    hip_check(hipStreamCreateWithFlags(&P->aux, hipStreamNonBlocking), "plan stream");
    hip_check(hipStreamSynchronize(ctx->stream), "plan set-up");
    ctx->plans.push_back(P.get());
    return P.release();
}

void plan_step(lpk_plan *P, uint64_t k) {
    if (k >= P->n_chunks) throw Fail{LPK_E_INVALID, "chunk beyond the plan"};
    if (k != 0 && k != P->next) throw Fail{LPK_E_STATE, "plan steps run in order; expected chunk " +
                                                            std::to_string(P->next) + " or 0"};
    lpk_ctx *ctx = P->ctx;
    const int slot = int(k & 1);
    hipStream_t s = slot ? P->aux : ctx->stream;
    if (k == 0) {
        for (uint32_t r : P->walk_rx) ctx->rx[r].nb_known = 0;  // a fresh ClockStream per receiver
        // the restart rewrites the walk tables the previous pass may still read
        for (int i = 0; i < 2; ++i)
            if (P->used[i]) hip_check(hipStreamWaitEvent(s, P->done[i], 0), "plan restart");
    }
    join_plans(ctx, P);   // other plans' steps come before this one
    if (slot && P->foreign) {
        // other work on the context's stream since the last step (it may have
        // extended the walk tables): the odd step follows it
        hip_check(hipEventRecord(P->ctx_ev, ctx->stream), "event record");
        hip_check(hipStreamWaitEvent(s, P->ctx_ev, 0), "plan order");
    }
    P->foreign = false;
    if (P->used[slot]) hip_check(hipEventSynchronize(P->done[slot]), "plan staging reuse");
    plan_fill(P, k, slot, static_cast<char *>(P->pin[slot]));
    char *dev = static_cast<char *>(P->tables[slot].p);
    const uint32_t *live = reinterpret_cast<const uint32_t *>(dev);
    const int nw = int(P->walk_rx.size());
    timing_begin(ctx, s);
    hip_check(hipMemcpyAsync(dev, P->pin[slot], P->bytes, hipMemcpyHostToDevice, s), "plan table copy");
    if (nw) {
        hip_check(launch_walk_gen(reinterpret_cast<const WalkSegDev *>(dev + P->off_segs), nw, P->max_blocks, live + 1,
                                  s),
                  "clock walk blocks");
        // the carries extend the shared base tables in step order
        if (k != 0 && P->used[1 - slot]) hip_check(hipStreamWaitEvent(s, P->carry_ev[1 - slot], 0), "carry order");
        hip_check(launch_walk_carry(reinterpret_cast<const CarrySegDev *>(dev + P->off_carry), nw, s),
                  "clock walk carry");
        hip_check(hipEventRecord(P->carry_ev[slot], s), "event record");
    }
    // one launch per pass, in order on the slot's stream (pass p + 1 reads
    // the sums pass p wrote)
    const JobDev *jobs_dev = reinterpret_cast<const JobDev *>(dev + P->off_jobs);
    for (size_t pp = 0; pp < P->pass_jobs.size(); ++pp) {
        if (P->pass_tiles[pp])
            hip_check(launch_render(P->form, P->antenna, P->moving, P->tile_shift, P->xcd_map != 0, jobs_dev,
                                    int(P->pass_jobs[pp]), reinterpret_cast<const LinkDev *>(dev + P->off_links),
                                    ctx->tx_dev, P->pass_tiles[pp], live + kPlanPassHdr + pp, nullptr, nullptr,
                                    nullptr, nullptr, 0, nullptr, nullptr, ctx->gate != 0, nullptr, s),
                      "render launch");
        jobs_dev += P->pass_jobs[pp];
    }
    timing_end(ctx, s);
    hip_check(hipEventRecord(P->done[slot], s), "event record");
    P->used[slot] = true;
    P->next = k + 1;
    if (slot) P->aux_pending = true;
}

}  // namespace

extern "C" {

int lpk_abi_version(void) { return LPK_ABI_VERSION; }

int lpk_ctx_create(int hip_device, lpk_ctx **out) {
    if (!out) return LPK_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_create_error = "no HIP device";
        return LPK_E_NODEVICE;
    }
    if (hip_device < 0 || hip_device >= n) {
        g_create_error = "device index out of range";
        return LPK_E_INVALID;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_error = std::string("device is not gfx950: ") + prop.gcnArchName;
        return LPK_E_NODEVICE;
    }
    lpk_ctx *ctx = new (std::nothrow) lpk_ctx();
    if (!ctx) return LPK_E_NOMEM;
    ctx->device = hip_device;
    if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "HIP stream/event creation failed";
        delete ctx;
        return LPK_E_HIP;
    }
    // the allocation-free gradient's workspace on this device; a failure is
    // reported by that entry point, not here
    (void)scratch_setup(hip_device);
    *out = ctx;
    return LPK_OK;
}

void lpk_ctx_destroy(lpk_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    free_scene(ctx);
    ctx->pred.release();
    if (ctx->stats) (void)hipFree(ctx->stats);
    ctx->jobs.release();
    ctx->links.release();
    ctx->dt_stage.release();
    ctx->out_stage.release();
    ctx->codes_stage.release();
    ctx->blocksum.release();
    ctx->walk_tables.release();
    ctx->part_stage.release();
    for (auto &s : ctx->pin) {
        if (s.p) (void)hipHostFree(s.p);
        if (s.done) (void)hipEventDestroy(s.done);
    }
    for (auto &pr : ctx->tpool) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *lpk_last_error(const lpk_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int lpk_scene_upload(lpk_ctx *ctx, const lpk_tx_desc *tx, uint32_t n_tx, const lpk_rx_desc *rx, uint32_t n_rx) {
    if (!ctx) return LPK_E_INVALID;
    if ((n_tx && !tx) || (n_rx && !rx)) return fail(ctx, Fail{LPK_E_INVALID, "descriptor array is NULL"});
    return guarded(ctx, [&] {
        hip_check(hipSetDevice(ctx->device), "set device");
        free_scene(ctx);
        std::vector<TxDev> host;
        host.reserve(n_tx);
        for (uint32_t i = 0; i < n_tx; ++i) {
            host.push_back(make_tx(ctx, tx[i], i));
            ctx->tx_carrier.push_back(tx[i].carrier_hz);
        }
        ctx->tx_dev = static_cast<TxDev *>(upload(ctx, host.data(), host.size() * sizeof(TxDev)));
        ctx->tx_host = std::move(host);
        ctx->n_tx = n_tx;
        ctx->rx.resize(n_rx);
        for (uint32_t r = 0; r < n_rx; ++r) {
            const lpk_rx_desc &d = rx[r];
            if (d.traj_kind == LPK_RX_STATIONARY) {
                for (int k = 0; k < 3; ++k)
                    if (!is_fin(d.pos_ecef[k])) throw Fail{LPK_E_NONFINITE, "receiver position"};
            } else if (d.traj_kind == LPK_RX_KEPLER) {
                if (!is_fin(d.kepler_a_m) || !is_fin(d.kepler_inc_deg) || !is_fin(d.kepler_raan_deg) ||
                    !is_fin(d.kepler_nu0_deg) || !is_fin(d.kepler_epoch_unix))
                    throw Fail{LPK_E_NONFINITE, "receiver " + std::to_string(r) + ": Keplerian elements"};
                if (!(d.kepler_a_m > 0.0)) throw Fail{LPK_E_INVALID, "receiver semi-major axis must be positive"};
            } else {
                throw Fail{LPK_E_INVALID, "receiver " + std::to_string(r) + ": unknown trajectory kind"};
            }
            if (!(d.master_rate_hz > 0.0) || !is_fin(d.master_rate_hz))
                throw Fail{LPK_E_INVALID, "receiver master_rate_hz must be positive"};
            if (!is_fin(d.allan_deviation_1s) || d.allan_deviation_1s < 0.0 || !is_fin(d.frequency_offset_y0) ||
                !is_fin(d.initial_time_offset_s))
                throw Fail{LPK_E_NONFINITE, "receiver clock parameters"};
            if (!is_fin(d.flicker_floor) || d.flicker_floor < 0.0) throw Fail{LPK_E_NONFINITE, "receiver flicker floor"};
            RxState &R = ctx->rx[r];
            R.d = d;
            R.walk = d.allan_deviation_1s > 0.0;  // src/clock/mod.rs:274-279
            flicker_init(R.flick, d.flicker_floor, d.clock_seed);
            R.sigma_step = R.walk ? d.allan_deviation_1s / std::sqrt(d.master_rate_hz) : 0.0;
        }
        ctx->have_scene = true;
    });
}

int lpk_render_chunk(lpk_ctx *ctx, const lpk_band_chunk *jobs, uint32_t n_jobs, int form) {
    if (!ctx) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        hip_check(hipSetDevice(ctx->device), "set device");
        render(ctx, jobs, n_jobs, form, nullptr);
    });
}

int lpk_render_taps(lpk_ctx *ctx, const lpk_band_chunk *job, int form, uint32_t slot, double *t_emit,
                    double *range, double *phase, double *delta_t) {
    if (!ctx || !job) return LPK_E_INVALID;
    DevBuf taps;
    int rc = guarded(ctx, [&] {
        hip_check(hipSetDevice(ctx->device), "set device");
        const size_t n = job->clen;
        double *d = static_cast<double *>(taps.ensure(std::max<size_t>(n, 1) * 32, ctx->stream));
        TapReq req{slot, d, d + n, d + 2 * n, d + 3 * n};
        render(ctx, job, 1, form, &req);
        hip_check(hipStreamSynchronize(ctx->stream), "taps sync");
        double *outs[4] = {t_emit, range, phase, delta_t};
        for (int k = 0; k < 4; ++k)
            if (outs[k] && n) hip_check(hipMemcpy(outs[k], d + k * n, n * 8, hipMemcpyDeviceToHost), "taps download");
    });
    taps.release();
    return rc;
}

int lpk_clock_rewind(lpk_ctx *ctx) {
    if (!ctx) return LPK_E_INVALID;
    // Host bookkeeping only: later walk kernels rewrite the table in stream
    // order, after every earlier render has read it.
    for (auto &r : ctx->rx) r.nb_known = 0;
    return LPK_OK;
}

int lpk_sync(lpk_ctx *ctx) {
    if (!ctx) return LPK_E_INVALID;
    return guarded(ctx, [&] { hip_check(hipStreamSynchronize(ctx->stream), "stream sync"); });
}

int lpk_last_kernel_ms(lpk_ctx *ctx, double *ms) {
    if (!ctx || !ms) return LPK_E_INVALID;
    if (!ctx->timed) return fail(ctx, Fail{LPK_E_STATE, "no timed render yet"});
    return guarded(ctx, [&] {
        float f = 0.0f;
        hip_check(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1), "event elapsed");
        *ms = f;
    });
}

int lpk_timing_reset(lpk_ctx *ctx) {
    if (!ctx) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        hip_check(hipStreamSynchronize(ctx->stream), "stream sync");
        ctx->tused = 0;
    });
}

int lpk_timing_read(lpk_ctx *ctx, double *total_ms, uint64_t *n_launches) {
    if (!ctx || !total_ms || !n_launches) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        hip_check(hipStreamSynchronize(ctx->stream), "stream sync");
        double t = 0.0;
        for (size_t k = 0; k < ctx->tused; ++k) {
            float f = 0.0f;
            hip_check(hipEventElapsedTime(&f, ctx->tpool[k].first, ctx->tpool[k].second), "event elapsed");
            t += f;
        }
        *total_ms = t;
        *n_launches = ctx->tused;
    });
}

int lpk_device_alloc(lpk_ctx *ctx, size_t bytes, void **dev) {
    if (!ctx || !dev) return LPK_E_INVALID;
    *dev = nullptr;
    return guarded(ctx, [&] {
        hip_check(hipSetDevice(ctx->device), "set device");
        if (hipMalloc(dev, std::max<size_t>(bytes, 1)) != hipSuccess) {
            *dev = nullptr;
            throw Fail{LPK_E_NOMEM, "device allocation failed"};
        }
    });
}

int lpk_device_free(lpk_ctx *ctx, void *dev) {
    if (!ctx) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        hip_check(hipStreamSynchronize(ctx->stream), "stream sync");
        hip_check(hipFree(dev), "device free");
    });
}

int lpk_copy_to_host(lpk_ctx *ctx, void *host, const void *dev, size_t bytes) {
    if (!ctx || (bytes && (!host || !dev))) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        hip_check(hipStreamSynchronize(ctx->stream), "stream sync");
        hip_check(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost), "copy to host");
    });
}

int lpk_set_tuning(lpk_ctx *ctx, const lpk_tuning *tuning) {
    if (!ctx || !tuning) return LPK_E_INVALID;
    if (tuning->tile_shift < 7 || tuning->tile_shift > 9 || (tuning->xcd_map != 0 && tuning->xcd_map != 1))
        return fail(ctx, Fail{LPK_E_INVALID, "tile_shift must be 7, 8 or 9 and xcd_map 0 or 1"});
    ctx->tile_shift = tuning->tile_shift;
    ctx->xcd_map = tuning->xcd_map;
    return LPK_OK;
}

int lpk_get_tuning(const lpk_ctx *ctx, lpk_tuning *tuning) {
    if (!ctx || !tuning) return LPK_E_INVALID;
    tuning->tile_shift = ctx->tile_shift;
    tuning->xcd_map = ctx->xcd_map;
    return LPK_OK;
}

int lpk_autotune(lpk_ctx *ctx, const lpk_band_chunk *jobs, uint32_t n_jobs, int form, lpk_tuning *best,
                 double *best_ms) {
    if (!ctx || !best || !best_ms) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        if (n_jobs && !jobs) throw Fail{LPK_E_INVALID, "jobs is NULL"};
        if (!ctx->have_scene) throw Fail{LPK_E_STATE, "no scene uploaded"};
        if (form < LPK_FORM_PLAIN || form > LPK_FORM_LEARNED) throw Fail{LPK_E_INVALID, "unknown form"};
        // the job set is checked once, before any candidate runs
        uint64_t samples = 0;
        for (uint32_t q = 0; q < n_jobs; ++q) {
            if (jobs[q].out_host || jobs[q].adc_codes_host || jobs[q].unoccluded)
                throw Fail{LPK_E_INVALID, "autotune jobs keep their output in device memory"};
            validate_job(ctx, jobs[q], q);
            check_form_links(ctx, form, jobs[q], q);
            samples += jobs[q].clen;
        }
        if (!samples) throw Fail{LPK_E_INVALID, "autotune needs at least one job with samples"};
        // a failure part-way leaves the context's tuning as it was
        struct Restore {
            lpk_ctx *c;
            int ts, xcd;
            bool keep = false;
            ~Restore() {
                if (!keep) {
                    c->tile_shift = ts;
                    c->xcd_map = xcd;
                }
            }
        } restore{ctx, ctx->tile_shift, ctx->xcd_map};
        const int shifts[3] = {7, 8, 9};
        double best_t = 1e300;
        lpk_tuning pick{ctx->tile_shift, ctx->xcd_map};
        for (int xcd = 1; xcd >= 0; --xcd)
            for (int ts : shifts) {
                ctx->tile_shift = ts;
                ctx->xcd_map = xcd;
                render(ctx, jobs, n_jobs, form, nullptr);  // warm-up
                render(ctx, jobs, n_jobs, form, nullptr);
                hip_check(hipStreamSynchronize(ctx->stream), "autotune");
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event elapsed");
                if (ms < best_t) {
                    best_t = ms;
                    pick = lpk_tuning{ts, xcd};
                }
            }
        ctx->tile_shift = pick.tile_shift;
        ctx->xcd_map = pick.xcd_map;
        restore.keep = true;
        *best = pick;
        *best_ms = best_t;
    });
}

int lpk_plan_create(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, uint64_t chunk_samples, int form,
                    lpk_plan **out) {
    if (!ctx || !out) return LPK_E_INVALID;
    *out = nullptr;
    return guarded(ctx, [&] { *out = plan_create(ctx, bands, n_bands, chunk_samples, form); });
}

int lpk_plan_step(lpk_plan *plan, uint64_t chunk) {
    if (!plan) return LPK_E_INVALID;
    return guarded(plan->ctx, [&] { plan_step(plan, chunk); }, false);
}

int lpk_plan_info(const lpk_plan *plan, uint64_t *n_chunks, uint32_t *max_tiles) {
    if (!plan) return LPK_E_INVALID;
    if (n_chunks) *n_chunks = plan->n_chunks;
    if (max_tiles) {
        uint64_t t = 0;
        for (uint32_t x : plan->pass_tiles) t += x;
        *max_tiles = uint32_t(std::min<uint64_t>(t, 0xFFFFFFFFull));
    }
    return LPK_OK;
}

int lpk_plan_chunk_range(const lpk_plan *plan, uint32_t band, uint64_t chunk, uint64_t *i0, uint64_t *clen) {
    if (!plan || !i0 || !clen || band >= plan->bands.size()) return LPK_E_INVALID;
    if (chunk >= plan->n_chunks) {
        *i0 = plan->bands[band].clen;
        *clen = 0;
        return LPK_OK;
    }
    plan_chunk(plan, band, chunk, *i0, *clen);
    return LPK_OK;
}

int lpk_plan_output(const lpk_plan *plan, uint32_t band, void **chunk_dev) {
    if (!plan || !chunk_dev || band >= plan->bands.size()) return LPK_E_INVALID;
    // the buffer of the last step's slot
    *chunk_dev = plan->bands[band].out_dev ? plan->bands[band].out_dev : plan->outbuf[(plan->next + 1) & 1][band].p;
    return LPK_OK;
}

void lpk_plan_destroy(lpk_plan *plan) { plan_free(plan); }

// ---------------------------------------------------------------- learned warm start

int lpk_lws_create(lpk_ctx *ctx, uint64_t init_seed, lpk_lws **out) {
    if (!ctx || !out) return LPK_E_INVALID;
    *out = nullptr;
    return guarded(ctx, [&] {
        // Glorot-uniform weights from SplitMix64, zero biases
        std::vector<float> th(kLwsParams, 0.0f);
        uint64_t x = init_seed;
        auto next = [&]() {
            x += 0x9E3779B97F4A7C15ULL;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            return (double(z >> 11) * 0x1.0p-53) * 2.0 - 1.0;
        };
        auto fill = [&](int off, int n, int fan_in, int fan_out) {
            const double lim = std::sqrt(6.0 / double(fan_in + fan_out));
            for (int e = 0; e < n; ++e) th[off + e] = float(lim * next());
        };
        fill(kOffW1, 64 * kLwsIn, kLwsIn, 64);
        fill(kOffW2, 64 * 64, 64, 64);
        fill(kOffW3, 64 * 64, 64, 64);
        fill(kOffW4, 64, 64, 1);
        auto net = std::unique_ptr<lpk_lws>(new lpk_lws());
        net->ctx = ctx;
        void *p[4] = {nullptr, nullptr, nullptr, nullptr};
        const size_t pb = kLwsParams * sizeof(float);
        const size_t sizes[4] = {pb, pb, pb, kLwsPacked * 2};
        for (int i = 0; i < 4; ++i)
            if (hipMalloc(&p[i], sizes[i]) != hipSuccess) {
                for (int k = 0; k < i; ++k) (void)hipFree(p[k]);
                throw Fail{LPK_E_NOMEM, "network allocation"};
            }
        net->theta = static_cast<float *>(p[0]);
        net->m1 = static_cast<float *>(p[1]);
        net->m2 = static_cast<float *>(p[2]);
        net->packed = p[3];
        hip_check(hipMemcpyAsync(net->theta, th.data(), pb, hipMemcpyHostToDevice, ctx->stream), "network upload");
        hip_check(hipMemsetAsync(net->m1, 0, pb, ctx->stream), "network init");
        hip_check(hipMemsetAsync(net->m2, 0, pb, ctx->stream), "network init");
        hip_check(launch_lws_pack(net->theta, net->packed, ctx->stream), "network pack");
        hip_check(hipStreamSynchronize(ctx->stream), "network init");
        *out = net.release();
    });
}

void lpk_lws_destroy(lpk_lws *net) {
    if (!net) return;
    lpk_ctx *ctx = net->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->lws == net) ctx->lws = nullptr;
    (void)hipFree(net->theta);
    (void)hipFree(net->m1);
    (void)hipFree(net->m2);
    (void)hipFree(net->packed);
    net->partials.release();
    net->grad.release();
    net->loss.release();
    delete net;
}

int lpk_lws_get_params(lpk_lws *net, float *theta) {
    if (!net || !theta) return LPK_E_INVALID;
    return guarded(net->ctx, [&] {
        hip_check(hipMemcpyAsync(theta, net->theta, kLwsParams * sizeof(float), hipMemcpyDeviceToHost,
                                 net->ctx->stream),
                  "parameter download");
        hip_check(hipStreamSynchronize(net->ctx->stream), "parameter download");
    });
}

int lpk_lws_set_params(lpk_lws *net, const float *theta) {
    if (!net || !theta) return LPK_E_INVALID;
    return guarded(net->ctx, [&] {
        for (int p = 0; p < kLwsParams; ++p)
            if (!std::isfinite(theta[p])) throw Fail{LPK_E_NONFINITE, "parameter " + std::to_string(p)};
        lpk_ctx *ctx = net->ctx;
        const size_t pb = kLwsParams * sizeof(float);
        hip_check(hipMemcpyAsync(net->theta, theta, pb, hipMemcpyHostToDevice, ctx->stream), "parameter upload");
        hip_check(hipMemsetAsync(net->m1, 0, pb, ctx->stream), "optimiser reset");
        hip_check(hipMemsetAsync(net->m2, 0, pb, ctx->stream), "optimiser reset");
        hip_check(launch_lws_pack(net->theta, net->packed, ctx->stream), "network pack");
        hip_check(hipStreamSynchronize(ctx->stream), "parameter upload");
        net->step = 0;
    });
}

int lpk_lws_generate(lpk_ctx *ctx, uint64_t n, uint64_t seed, float *x_dev, float *y_dev, double *geom_dev) {
    if (!ctx || (n && (!x_dev || !y_dev))) return LPK_E_INVALID;
    return guarded(ctx, [&] { hip_check(launch_lws_gen(n, seed, x_dev, y_dev, geom_dev, ctx->stream), "pair generation"); });
}

int lpk_lws_infer(lpk_lws *net, const float *x_dev, uint64_t n, float *y_dev) {
    if (!net || (n && (!x_dev || !y_dev))) return LPK_E_INVALID;
    return guarded(net->ctx, [&] {
        hip_check(launch_lws_infer(lws_net(net), x_dev, n, y_dev, net->ctx->stream), "inference launch");
    });
}

int lpk_lws_gradient(lpk_lws *net, const float *x_dev, const float *y_dev, uint64_t n, float *grad, double *loss) {
    if (!net || !x_dev || !y_dev || !grad || n == 0) return LPK_E_INVALID;
    return guarded(net->ctx, [&] {
        lpk_ctx *ctx = net->ctx;
        const uint64_t rows = (n + kTrainQ - 1) / kTrainQ;
        if (rows > (1u << 20)) throw Fail{LPK_E_INVALID, "gradient batch too large"};
        float *part = static_cast<float *>(net->partials.ensure(rows * kLwsRow * sizeof(float), ctx->stream));
        float *g = static_cast<float *>(net->grad.ensure((kLwsParams + 1) * sizeof(float), ctx->stream));
        hip_check(launch_lws_train(lws_net(net), x_dev, y_dev, 0, n, uint32_t(rows), float(2.0 / double(n)), part,
                                   ctx->stream),
                  "training launch");
        hip_check(launch_lws_update(part, int(rows), net->theta, net->m1, net->m2, net->packed, 0.f, 0.f, 0.f, 0.f,
                                    0.f, 0.f, g, g + kLwsParams, ctx->stream),
                  "gradient reduction");
        std::vector<float> h(kLwsParams + 1);
        hip_check(hipMemcpyAsync(h.data(), g, h.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
                  "gradient download");
        hip_check(hipStreamSynchronize(ctx->stream), "gradient download");
        std::memcpy(grad, h.data(), kLwsParams * sizeof(float));
        if (loss) *loss = double(h[kLwsParams]) / double(n);
    });
}

int lpk_lws_gradient_noalloc(lpk_lws *net, const float *x_dev, const float *y_dev, uint64_t n, float *grad,
                             double *loss) {
    if (!net || !x_dev || !y_dev || !grad || n == 0) return LPK_E_INVALID;
    lpk_ctx *ctx = net->ctx;
    const uint64_t rows = (n + kTrainQ - 1) / kTrainQ;
    if (rows > uint64_t(kNoAllocRows))
        return fail(ctx, Fail{LPK_E_INVALID, "batch larger than the static gradient workspace (at most " +
                                                 std::to_string(kNoAllocRows * kTrainQ) + " pairs)"});
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, Fail{LPK_E_HIP, "hipSetDevice failed"});
    const hipError_t rc = scratch_setup(ctx->device);
    if (rc != hipSuccess)
        return fail(ctx, Fail{LPK_E_HIP, std::string("gradient workspace lookup: ") + hipGetErrorString(rc)});
    float *part = g_scratch[ctx->device];
    float *g = part + kNoAllocGradOff;
    float last = 0.0f;
    hipError_t e = launch_lws_train(lws_net(net), x_dev, y_dev, 0, n, uint32_t(rows), float(2.0 / double(n)), part,
                                    ctx->stream);
    if (e == hipSuccess)
        e = launch_lws_update(part, int(rows), net->theta, net->m1, net->m2, net->packed, 0.f, 0.f, 0.f, 0.f, 0.f,
                              0.f, g, g + kLwsParams, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(grad, g, kLwsParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&last, g + kLwsParams, sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, Fail{LPK_E_HIP, std::string("gradient: ") + hipGetErrorString(e)});
    if (loss) *loss = double(last) / double(n);
    return LPK_OK;
}

int lpk_lws_train(lpk_lws *net, const lpk_lws_train_cfg *cfg, lpk_lws_train_report *rep) {
    if (!net || !cfg || !rep) return LPK_E_INVALID;
    return guarded(net->ctx, [&] {
        lpk_ctx *ctx = net->ctx;
        const uint64_t n = cfg->n_pairs, batch = cfg->batch;
        if (batch == 0 || batch % kTrainQ || batch > (1u << 20) || n < batch || cfg->epochs == 0 ||
            !(cfg->lr > 0.0f) || !(cfg->lr_final >= 0.0f) || cfg->n_val == 0)
            throw Fail{LPK_E_INVALID, "training configuration"};
        *rep = lpk_lws_train_report{};
        struct Bufs {
            std::vector<void *> p;
            ~Bufs() {
                for (void *q : p) (void)hipFree(q);
            }
            float *get(size_t floats) {
                void *q = nullptr;
                if (hipMalloc(&q, std::max<size_t>(floats, 1) * sizeof(float)) != hipSuccess)
                    throw Fail{LPK_E_NOMEM, "training data allocation"};
                p.push_back(q);
                return static_cast<float *>(q);
            }
        } bufs;
        float *X = bufs.get(n * kLwsIn), *Y = bufs.get(n);
        float *Xv = bufs.get(cfg->n_val * kLwsIn), *Yv = bufs.get(cfg->n_val), *Pv = bufs.get(cfg->n_val);
        const uint64_t spe = n / batch, total = spe * cfg->epochs;
        float *losses = bufs.get(total);
        const uint32_t slices = uint32_t(batch / kTrainQ);
        float *part = static_cast<float *>(net->partials.ensure(size_t(slices) * kLwsRow * sizeof(float), ctx->stream));
        hipEvent_t ev[3];
        for (auto &e : ev) hip_check(hipEventCreate(&e), "event create");
        struct EvGuard {
            hipEvent_t *e;
            ~EvGuard() {
                for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
            }
        } eg{ev};
        hip_check(hipEventRecord(ev[0], ctx->stream), "event record");
        hip_check(launch_lws_gen(n, cfg->data_seed, X, Y, nullptr, ctx->stream), "pair generation");
        hip_check(launch_lws_gen(cfg->n_val, cfg->data_seed ^ 0x76616C6964617465ULL, Xv, Yv, nullptr, ctx->stream),
                  "pair generation");
        hip_check(hipEventRecord(ev[1], ctx->stream), "event record");
        const LwsNetDev nd = lws_net(net);
        const float b1 = 0.9f, b2 = 0.999f;
        for (uint64_t st = 0; st < total; ++st) {
            const uint64_t base = (st % spe) * batch;
            const double prog = double(st) / double(total);
            const float lr = float(cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(3.141592653589793 * prog)));
            const double t = double(net->step + 1);
            const float c1 = float(1.0 / (1.0 - std::pow(double(b1), t)));
            const float c2 = float(1.0 / (1.0 - std::pow(double(b2), t)));
            hip_check(launch_lws_train(nd, X, Y, base, n, slices, float(2.0 / double(batch)), part, ctx->stream),
                      "training launch");
            hip_check(launch_lws_update(part, int(slices), net->theta, net->m1, net->m2, net->packed, lr, b1, b2,
                                        1e-8f, c1, c2, nullptr, losses + st, ctx->stream),
                      "update launch");
            ++net->step;
        }
        hip_check(hipEventRecord(ev[2], ctx->stream), "event record");
        hip_check(launch_lws_infer(nd, Xv, cfg->n_val, Pv, ctx->stream), "validation inference");
        std::vector<float> hl(total), hy(cfg->n_val), hp(cfg->n_val);
        hip_check(hipMemcpyAsync(hl.data(), losses, total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "loss download");
        hip_check(hipMemcpyAsync(hy.data(), Yv, hy.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "validation download");
        hip_check(hipMemcpyAsync(hp.data(), Pv, hp.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "validation download");
        hip_check(hipStreamSynchronize(ctx->stream), "training");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, ev[0], ev[1]), "event elapsed");
        rep->gen_ms = ms;
        hip_check(hipEventElapsedTime(&ms, ev[1], ev[2]), "event elapsed");
        rep->train_ms = ms;
        auto epoch_mean = [&](uint64_t e) {
            double acc = 0.0;
            for (uint64_t k = e * spe; k < (e + 1) * spe; ++k) acc += hl[k];
            return acc / double(spe * batch);
        };
        rep->loss_first = epoch_mean(0);
        rep->loss_last = epoch_mean(cfg->epochs - 1);
        double se = 0.0, mx = 0.0;
        for (uint64_t k = 0; k < cfg->n_val; ++k) {
            const double d = (double(hp[k]) - double(hy[k])) * kLwsTargetScale;
            se += d * d;
            mx = std::max(mx, std::fabs(d));
        }
        rep->val_rms_s = std::sqrt(se / double(cfg->n_val));
        rep->val_max_s = mx;
        rep->steps = total;
    });
}

int lpk_set_warm_start(lpk_ctx *ctx, lpk_lws *net) {
    if (!ctx || (net && net->ctx != ctx)) return LPK_E_INVALID;
    ctx->lws = net;
    return LPK_OK;
}

int lpk_solve_stats(lpk_ctx *ctx, uint64_t stats[4]) {
    if (!ctx || !stats) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        ensure_stats(ctx);
        unsigned long long h[kNumStats];
        hip_check(hipMemcpyAsync(h, ctx->stats, sizeof h, hipMemcpyDeviceToHost, ctx->stream), "counter download");
        hip_check(hipStreamSynchronize(ctx->stream), "counter download");
        for (int k = 0; k < 4; ++k) stats[k] = h[k];
    });
}

int lpk_solve_stats_reset(lpk_ctx *ctx) {
    if (!ctx) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        ensure_stats(ctx);
        hip_check(hipMemsetAsync(ctx->stats, 0, kNumStats * sizeof(unsigned long long), ctx->stream), "counter reset");
    });
}

// ---------------------------------------------------------------- output

}  // extern "C"

namespace {

// generate_band's chunk loop (src/pipeline.rs:905-990) for every band at
// once, with the sink push of :984: chunk k of every band renders on the
// device into one of two pinned buffers while the host pushes chunk k - 1 of
// each band, in band order, from the other. push(b, i0, iq, n) returns false
// to stop band b (ControlFlow::Break, src/pipeline.rs:372-375); its later
// chunks are not rendered. Synchronous.
void stream_bands(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, uint64_t chunk_samples, int form,
                  const std::function<bool(uint32_t, uint64_t, const float *, uint64_t)> &push) {
    struct Pinned {
        void *p[2] = {nullptr, nullptr};
        ~Pinned() {
            for (void *q : p)
                if (q) (void)hipHostFree(q);
        }
    } pin;
    uint64_t slot_samples = 0, n_chunks = 0;
    for (uint32_t b = 0; b < n_bands; ++b) {
        const lpk_band_chunk &B = bands[b];
        if (B.i0 != 0 || B.out_dev || B.out_host || B.adc_codes_host)
            throw Fail{LPK_E_INVALID, "band " + std::to_string(b) + ": whole-band descriptor (i0 = 0, no outputs) expected"};
        slot_samples += std::min<uint64_t>(B.clen, chunk_samples);
        n_chunks = std::max<uint64_t>(n_chunks, (B.clen + chunk_samples - 1) / chunk_samples);
    }
    for (int i = 0; i < 2; ++i)
        if (slot_samples && hipHostMalloc(&pin.p[i], slot_samples * 8, hipHostMallocDefault) != hipSuccess)
            throw Fail{LPK_E_NOMEM, "pinned output buffers"};
    // chunk k of band b lands at off[b] of slot k % 2
    std::vector<uint64_t> off(n_bands, 0), len_prev(n_bands, 0), len_cur(n_bands, 0);
    std::vector<char> live(n_bands, 1);
    for (uint32_t b = 1; b < n_bands; ++b) off[b] = off[b - 1] + std::min<uint64_t>(bands[b - 1].clen, chunk_samples);
    auto push_slot = [&](int slot, uint64_t k, const std::vector<uint64_t> &len) {
        const float *base = static_cast<const float *>(pin.p[slot]);
        for (uint32_t b = 0; b < n_bands; ++b)
            if (len[b] && live[b] && !push(b, k * chunk_samples, base + 2 * off[b], len[b])) live[b] = 0;
    };
    std::vector<lpk_band_chunk> jobs;
    for (uint64_t k = 0; k < n_chunks; ++k) {
        const int slot = int(k & 1);
        jobs.clear();
        for (uint32_t b = 0; b < n_bands; ++b) {
            const uint64_t i0 = k * chunk_samples;
            len_cur[b] = live[b] && i0 < bands[b].clen ? std::min(chunk_samples, bands[b].clen - i0) : 0;
            if (!len_cur[b]) continue;
            lpk_band_chunk J = bands[b];
            J.i0 = i0;
            J.clen = len_cur[b];
            J.out_host = static_cast<char *>(pin.p[slot]) + off[b] * 8;
            if (J.delta_t) J.delta_t += i0;
            jobs.push_back(J);
        }
        if (!jobs.empty()) render(ctx, jobs.data(), uint32_t(jobs.size()), form, nullptr);
        if (k) push_slot(slot ^ 1, k - 1, len_prev);
        hip_check(hipStreamSynchronize(ctx->stream), "chunk render");
        std::swap(len_prev, len_cur);
    }
    if (n_chunks) push_slot(int((n_chunks - 1) & 1), n_chunks - 1, len_prev);
}

}  // namespace

extern "C" {

int lpk_stream_to_files(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, const char *const *paths,
                        uint64_t chunk_samples, int form) {
    if (!ctx || (n_bands && (!bands || !paths)) || chunk_samples == 0) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        struct Files {
            std::vector<FILE *> f;
            ~Files() {
                for (FILE *p : f)
                    if (p) std::fclose(p);
            }
        } files;
        for (uint32_t b = 0; b < n_bands; ++b) {
            if (!paths[b]) throw Fail{LPK_E_INVALID, "band " + std::to_string(b) + ": no path"};
            FILE *f = std::fopen(paths[b], "wb");
            if (!f) throw Fail{LPK_E_INVALID, std::string("cannot open ") + paths[b] + ": " + std::strerror(errno)};
            files.f.push_back(f);
        }
        stream_bands(ctx, bands, n_bands, chunk_samples, form,
                     [&](uint32_t b, uint64_t, const float *iq, uint64_t n) {
                         if (std::fwrite(iq, 8, n, files.f[b]) != n)
                             throw Fail{LPK_E_INVALID, std::string("write failed: ") + paths[b]};
                         return true;
                     });
        for (uint32_t b = 0; b < n_bands; ++b)
            if (std::fflush(files.f[b]) != 0) throw Fail{LPK_E_INVALID, std::string("write failed: ") + paths[b]};
    });
}

int lpk_stream_to_sink(lpk_ctx *ctx, const lpk_band_chunk *bands, uint32_t n_bands, uint64_t chunk_samples, int form,
                       lpk_iq_sink sink, void *user) {
    if (!ctx || (n_bands && !bands) || chunk_samples == 0 || !sink) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        stream_bands(ctx, bands, n_bands, chunk_samples, form,
                     [&](uint32_t b, uint64_t i0, const float *iq, uint64_t n) { return sink(user, b, i0, iq, n) == 0; });
    });
}

int lpk_sink_fnv1a(void *user, uint32_t band, uint64_t i0, const float *iq, uint64_t n) {
    (void)i0;
    uint64_t *h = static_cast<uint64_t *>(user) + band;
    *h = lpk_fnv1a64(iq, n * 8, *h);
    return 0;
}

uint64_t lpk_fnv1a64(const void *data, size_t bytes, uint64_t h) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < bytes; ++i) h = (h ^ p[i]) * 0x100000001B3ULL;
    return h;
}

int lpk_host_alloc(lpk_ctx *ctx, size_t bytes, void **host) {
    if (!ctx || !host) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        *host = nullptr;
        if (hipHostMalloc(host, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess)
            throw Fail{LPK_E_NOMEM, "pinned host allocation of " + std::to_string(bytes) + " bytes failed"};
    });
}

int lpk_host_free(lpk_ctx *ctx, void *host) {
    if (!ctx) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        if (host) hip_check(hipHostFree(host), "pinned host free");
    });
}

int lpk_truth(lpk_ctx *ctx, uint32_t rx, const double *t_rel, uint64_t n, double epoch_unix, int iterations,
              double *range, double *src_xyz) {
    if (!ctx || (n && (!t_rel || !range || !src_xyz)) || iterations < 0) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        if (!ctx->have_scene) throw Fail{LPK_E_STATE, "no scene uploaded"};
        if (rx >= ctx->rx.size()) throw Fail{LPK_E_INVALID, "receiver index out of range"};
        const uint64_t nt = ctx->n_tx;
        if (!n || !nt) return;
        DevBuf tb, rb, sb, ob;
        struct Rel {
            DevBuf *b[4];
            ~Rel() {
                for (DevBuf *q : b) q->release();
            }
        } rel{{&tb, &rb, &sb, &ob}};
        // the receiver at every grid time, as position_ecef_batch of its
        // trajectory at epoch + t (src/pipeline.rs:562-563)
        const lpk_rx_desc &D = ctx->rx[rx].d;
        std::vector<double> obs(3 * n);
        for (uint64_t k = 0; k < n; ++k) {
            if (D.traj_kind == LPK_RX_KEPLER)
                lpk_kepler_position(D.kepler_a_m, D.kepler_inc_deg, D.kepler_raan_deg, D.kepler_nu0_deg,
                                    D.kepler_epoch_unix, epoch_unix + t_rel[k], &obs[3 * k]);
            else
                for (int a = 0; a < 3; ++a) obs[3 * k + a] = D.pos_ecef[a];
        }
        double *td = static_cast<double *>(tb.ensure(n * 8, ctx->stream));
        double *rd = static_cast<double *>(rb.ensure(nt * n * 8, ctx->stream));
        double *sd = static_cast<double *>(sb.ensure(nt * 3 * n * 8, ctx->stream));
        double *od = static_cast<double *>(ob.ensure(3 * n * 8, ctx->stream));
        hip_check(hipMemcpyAsync(td, t_rel, n * 8, hipMemcpyHostToDevice, ctx->stream), "truth grid upload");
        hip_check(hipMemcpyAsync(od, obs.data(), 3 * n * 8, hipMemcpyHostToDevice, ctx->stream), "truth grid upload");
        hip_check(launch_truth(ctx->tx_dev, uint32_t(nt), od, epoch_unix, td, n, iterations, rd, sd, ctx->stream),
                  "truth launch");
        hip_check(hipMemcpyAsync(range, rd, nt * n * 8, hipMemcpyDeviceToHost, ctx->stream), "truth download");
        hip_check(hipMemcpyAsync(src_xyz, sd, nt * 3 * n * 8, hipMemcpyDeviceToHost, ctx->stream), "truth download");
        hip_check(hipStreamSynchronize(ctx->stream), "truth");
    });
}

int lpk_set_horizon_gate(lpk_ctx *ctx, int on) {
    if (!ctx) return LPK_E_INVALID;
    ctx->gate = on ? 1 : 0;
    return LPK_OK;
}

int lpk_horizon_gate_stats(lpk_ctx *ctx, uint64_t stats[2]) {
    if (!ctx || !stats) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        ensure_stats(ctx);
        unsigned long long h[kNumStats];
        hip_check(hipMemcpyAsync(h, ctx->stats, sizeof h, hipMemcpyDeviceToHost, ctx->stream), "counter download");
        hip_check(hipStreamSynchronize(ctx->stream), "counter download");
        stats[0] = h[4];
        stats[1] = h[5];
    });
}

int lpk_device_fnv1a(lpk_ctx *ctx, const void *dev, size_t bytes, uint64_t *hash) {
    if (!ctx || !hash || (bytes && !dev)) return LPK_E_INVALID;
    return guarded(ctx, [&] {
        std::vector<unsigned char> h(bytes);
        hip_check(hipStreamSynchronize(ctx->stream), "stream sync");
        if (bytes) hip_check(hipMemcpy(h.data(), dev, bytes, hipMemcpyDeviceToHost), "copy to host");
        uint64_t x = 0xCBF29CE484222325ULL;
        for (unsigned char c : h) x = (x ^ c) * 0x100000001B3ULL;
        *hash = x;
    });
}

}  // extern "C"
