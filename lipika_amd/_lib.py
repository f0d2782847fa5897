"""ctypes binding of include/lpk.h (the HIP backend's C ABI).

The library is ``lipika_amd/liblpk.so``, built in place by
``python -m lipika_amd.build``. There is no fallback: if the library is
missing or the device is not a gfx950, every render raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpk.so")

ABI_VERSION = 4  # LPK_ABI_VERSION of the header these structures mirror
LPK_OK = 0
STATUS = {
    0: "LPK_OK", -1: "LPK_E_INVALID", -2: "LPK_E_NOMEM", -3: "LPK_E_HIP",
    -4: "LPK_E_NONFINITE", -5: "LPK_E_STATE", -6: "LPK_E_NODEVICE",
}
SIG_CW, SIG_BPSK, SIG_OOK, SIG_QAM16, SIG_CA = range(5)
NAV_OFF, NAV_DUMMY, NAV_FULL = range(3)
TRAJ_STATIONARY, TRAJ_GRID = range(2)
RX_STATIONARY, RX_KEPLER = range(2)
POST_GAIN, POST_NONLINEARITY, POST_ADC = range(3)
FORM_PLAIN, FORM_PRECISION, FORM_CONVERGED, FORM_LEARNED = 0, 1, 2, 3
LWS_PARAMS = 9089

_d = C.c_double
_u64 = C.c_uint64
_i64 = C.c_int64
_i32 = C.c_int32
_u32 = C.c_uint32


class Ephemeris(C.Structure):
    """lpk_ephemeris (include/lpk.h), Ephemeris of src/signal/gps_nav.rs:56-85."""
    _fields_ = [("prn", _u32), ("week", _u32),
                ("toe_s", _d), ("toc_s", _d), ("sqrt_a", _d), ("e", _d), ("m0_rad", _d), ("delta_n_rad_s", _d),
                ("i0_rad", _d), ("idot_rad_s", _d), ("omega0_rad", _d), ("omega_dot_rad_s", _d), ("argp_rad", _d),
                ("cuc", _d), ("cus", _d), ("crc", _d), ("crs", _d), ("cic", _d), ("cis", _d),
                ("af0", _d), ("af1", _d), ("af2", _d), ("tgd", _d),
                ("iode", _u32), ("iodc", _u32), ("health", _u32), ("ura_index", _u32)]


class Tle(C.Structure):
    """lpk_tle (include/lpk.h): a parsed two-line element set."""
    _fields_ = [("satnum", _u32), ("epoch_year", _i32), ("epoch_day", _d), ("epoch_unix", _d), ("jd_epoch", _d),
                ("ndot", _d), ("nddot", _d), ("bstar", _d), ("inclo_deg", _d), ("nodeo_deg", _d), ("ecco", _d),
                ("argpo_deg", _d), ("mo_deg", _d), ("no_revday", _d)]


class TxDesc(C.Structure):
    _fields_ = [
        ("traj_kind", _i32), ("signal_kind", _i32),
        ("pos_ecef", _d * 3),
        ("grid_t0_unix", _d), ("grid_dt_s", _d), ("grid_n", _u64),
        ("grid_xyz", C.POINTER(_d)),
        ("carrier_hz", _d), ("tx_power_w", _d), ("bandwidth_hz", _d),
        ("baseband_offset_hz", _d), ("cw_anchored", _i32), ("_pad0", _i32),
        ("anchor_phase_rad", _d),
        ("rate_hz", _d), ("epoch_rel_s", _d),
        ("repeating", _i32), ("_pad1", _i32), ("repeat_from", _i64),
        ("n_symbols", _u64),
        ("bits", C.POINTER(C.c_int8)), ("symbols", C.POINTER(C.c_float)),
        ("nav_mode", _i32), ("prn", _i32),
        ("ca_code", C.POINTER(C.c_int8)), ("nav_seed", _u64),
        ("nav_bits", C.POINTER(C.c_int8)), ("n_nav_bits", _u64),
        ("gated", _i32), ("n_bursts", _u32),
        ("gate_period_s", _d), ("gate_phase_s", _d),
        ("bursts", C.POINTER(_d)),
        ("has_kepler", _i32), ("_pad2", _i32),
        ("kepler_a_m", _d), ("kepler_inc_deg", _d), ("kepler_raan_deg", _d), ("kepler_nu0_deg", _d),
        ("kepler_epoch_unix", _d),
    ]


class RxDesc(C.Structure):
    _fields_ = [
        ("pos_ecef", _d * 3),
        ("allan_deviation_1s", _d), ("frequency_offset_y0", _d),
        ("initial_time_offset_s", _d), ("master_rate_hz", _d),
        ("clock_seed", _u64), ("flicker_floor", _d),
        ("traj_kind", _i32), ("_pad", _i32),
        ("kepler_a_m", _d), ("kepler_inc_deg", _d), ("kepler_raan_deg", _d), ("kepler_nu0_deg", _d),
        ("kepler_epoch_unix", _d),
    ]


class PostOp(C.Structure):
    _fields_ = [("kind", _i32), ("bits", _i32), ("value", _d)]


class BandChunk(C.Structure):
    _fields_ = [
        ("rx", _u32), ("n_active", _u32), ("active", C.POINTER(_u32)),
        ("fs_hz", _d), ("lo_hz", _d), ("epoch_unix", _d),
        ("i0", _u64), ("clen", _u64),
        ("delta_t", C.POINTER(_d)),
        ("noise_power_w", _d), ("noise_seed", _u64),
        ("n_post", _u32), ("_pad", _u32),
        ("post", PostOp * 4),
        ("out_dev", C.c_void_p), ("out_host", C.c_void_p),
        ("adc_codes_host", C.POINTER(C.c_int16)),
        ("has_antenna", C.c_int32), ("has_cutoff", C.c_int32),
        ("antenna_peak_dbi", _d), ("antenna_horizon_dbi", _d), ("antenna_fb_db", _d),
        ("antenna_cutoff_deg", _d),
        ("unoccluded", C.POINTER(C.c_uint8)),
    ]


class Tuning(C.Structure):
    _fields_ = [("tile_shift", C.c_int32), ("xcd_map", C.c_int32)]


class LwsTrainCfg(C.Structure):
    _fields_ = [("n_pairs", _u64), ("data_seed", _u64), ("epochs", _u32), ("batch", _u32),
                ("lr", C.c_float), ("lr_final", C.c_float), ("n_val", _u64)]


class LwsTrainReport(C.Structure):
    _fields_ = [("loss_first", _d), ("loss_last", _d), ("val_rms_s", _d), ("val_max_s", _d),
                ("gen_ms", _d), ("train_ms", _d), ("steps", _u64)]


class LpkError(RuntimeError):
    """A non-zero lpk_status, with the library's error text."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_LIB = None

# lpk_iq_sink: int (*)(void *user, uint32_t band, uint64_t i0, const float *iq, uint64_t n)
IQ_SINK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.c_uint64, C.POINTER(C.c_float), C.c_uint64)
FNV_OFFSET = 0xCBF29CE484222325


def load() -> C.CDLL:
    """Load liblpk.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lipika_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "lpk_abi_version": (_i32, []),
        "lpk_ctx_create": (_i32, [C.c_int, P(C.c_void_p)]),
        "lpk_ctx_destroy": (None, [C.c_void_p]),
        "lpk_last_error": (C.c_char_p, [C.c_void_p]),
        "lpk_scene_upload": (_i32, [C.c_void_p, P(TxDesc), _u32, P(RxDesc), _u32]),
        "lpk_render_chunk": (_i32, [C.c_void_p, P(BandChunk), _u32, C.c_int]),
        "lpk_sync": (_i32, [C.c_void_p]),
        "lpk_clock_rewind": (_i32, [C.c_void_p]),
        "lpk_render_taps": (_i32, [C.c_void_p, P(BandChunk), C.c_int, _u32, P(_d), P(_d), P(_d), P(_d)]),
        "lpk_last_kernel_ms": (_i32, [C.c_void_p, P(_d)]),
        "lpk_set_tuning": (_i32, [C.c_void_p, P(Tuning)]),
        "lpk_get_tuning": (_i32, [C.c_void_p, P(Tuning)]),
        "lpk_autotune": (_i32, [C.c_void_p, P(BandChunk), _u32, C.c_int, P(Tuning), P(_d)]),
        "lpk_plan_create": (_i32, [C.c_void_p, P(BandChunk), _u32, _u64, C.c_int, P(C.c_void_p)]),
        "lpk_plan_step": (_i32, [C.c_void_p, _u64]),
        "lpk_plan_info": (_i32, [C.c_void_p, P(_u64), P(_u32)]),
        "lpk_plan_chunk_range": (_i32, [C.c_void_p, _u32, _u64, P(_u64), P(_u64)]),
        "lpk_plan_output": (_i32, [C.c_void_p, _u32, P(C.c_void_p)]),
        "lpk_plan_destroy": (None, [C.c_void_p]),
        "lpk_timing_reset": (_i32, [C.c_void_p]),
        "lpk_timing_read": (_i32, [C.c_void_p, P(_d), P(_u64)]),
        "lpk_device_alloc": (_i32, [C.c_void_p, C.c_size_t, P(C.c_void_p)]),
        "lpk_device_free": (_i32, [C.c_void_p, C.c_void_p]),
        "lpk_copy_to_host": (_i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
        "lpk_device_fnv1a": (_i32, [C.c_void_p, C.c_void_p, C.c_size_t, P(_u64)]),
        "lpk_geodetic_to_ecef": (None, [_d, _d, _d, P(_d)]),
        "lpk_grid_size": (_u64, [_d, _d]),
        "lpk_grid_fill_kepler": (_i32, [_d, _d, _d, _d, _d, _d, _d, _d, P(_d), P(_d), _u64]),
        "lpk_ca_code": (_i32, [C.c_int, P(C.c_int8)]),
        "lpk_kepler_position": (None, [_d, _d, _d, _d, _d, _d, P(_d)]),
        "lpk_nav_fold_seed": (_u64, [_u64, _u64]),
        "lpk_payload_bpsk": (_i32, [P(C.c_int8), _u64, _u64, _u64, P(C.c_int8)]),
        "lpk_payload_ook": (_i32, [P(C.c_uint8), _u64, _u64, _u64, P(C.c_int8)]),
        "lpk_payload_qam16": (_i32, [P(C.c_uint8), _u64, _u64, _u64, P(C.c_float)]),
        "lpk_elapsed_periods": (_i32, [_d, _d, P(_i64), P(_d)]),
        "lpk_frame_anchor": (_i32, [_d, _d, _d, _u64, P(_d), P(_i64)]),
        "lpk_cw_anchor_phase": (_i32, [_d, _d, _d, P(_d)]),
        "lpk_bands_overlap": (_i32, [_d, _d, _d, _d]),
        "lpk_thermal_noise_power_w": (_d, [_d, _d]),
        "lpk_lws_create": (_i32, [C.c_void_p, _u64, P(C.c_void_p)]),
        "lpk_lws_destroy": (None, [C.c_void_p]),
        "lpk_lws_get_params": (_i32, [C.c_void_p, P(C.c_float)]),
        "lpk_lws_set_params": (_i32, [C.c_void_p, P(C.c_float)]),
        "lpk_lws_generate": (_i32, [C.c_void_p, _u64, _u64, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lpk_lws_infer": (_i32, [C.c_void_p, C.c_void_p, _u64, C.c_void_p]),
        "lpk_lws_gradient": (_i32, [C.c_void_p, C.c_void_p, C.c_void_p, _u64, P(C.c_float), P(_d)]),
        "lpk_lws_gradient_noalloc": (_i32, [C.c_void_p, C.c_void_p, C.c_void_p, _u64, P(C.c_float), P(_d)]),
        "lpk_lws_train": (_i32, [C.c_void_p, P(LwsTrainCfg), P(LwsTrainReport)]),
        "lpk_set_warm_start": (_i32, [C.c_void_p, C.c_void_p]),
        "lpk_solve_stats": (_i32, [C.c_void_p, P(_u64)]),
        "lpk_solve_stats_reset": (_i32, [C.c_void_p]),
        "lpk_set_horizon_gate": (_i32, [C.c_void_p, C.c_int]),
        "lpk_horizon_gate_stats": (_i32, [C.c_void_p, P(_u64)]),
        "lpk_stream_to_files": (_i32, [C.c_void_p, P(BandChunk), _u32, P(C.c_char_p), _u64, C.c_int]),
        "lpk_stream_to_sink": (_i32, [C.c_void_p, P(BandChunk), _u32, _u64, C.c_int, IQ_SINK, C.c_void_p]),
        "lpk_fnv1a64": (_u64, [C.c_void_p, C.c_size_t, _u64]),
        "lpk_host_alloc": (_i32, [C.c_void_p, C.c_size_t, P(C.c_void_p)]),
        "lpk_host_free": (_i32, [C.c_void_p, C.c_void_p]),
        "lpk_truth": (_i32, [C.c_void_p, _u32, P(_d), _u64, _d, C.c_int, P(_d), P(_d)]),
        "lpk_gps_time": (None, [_d, _d, P(_u32), P(_d)]),
        "lpk_toe_snap": (_i32, [_d, _d, P(_d), P(_u32), P(_d)]),
        "lpk_ephemeris_from_kepler": (_i32, [_d, _d, _d, _d, _d, _u32, _d, _d, P(Ephemeris)]),
        "lpk_ephemeris_from_state": (_i32, [P(_d), P(_d), _u32, _u32, _d, P(Ephemeris)]),
        "lpk_lnav_length": (_u64, [_d]),
        "lpk_lnav_bits": (_i32, [P(Ephemeris), _d, _d, P(C.c_int8), _u64]),
        "lpk_sv_position": (None, [P(Ephemeris), _d, P(_d)]),
        "lpk_grid_fill_broadcast": (_i32, [P(Ephemeris), _d, _d, _d, _d, P(_d), P(_d), _u64]),
        "lpk_rinex3_gps_ephemeris": (_i32, [C.c_char_p, _u32, _d, _d, P(Ephemeris)]),
        "lpk_tle_parse": (_i32, [C.c_char_p, C.c_char_p, P(Tle)]),
        "lpk_sgp4_teme": (_i32, [P(Tle), _d, P(_d), P(_d)]),
        "lpk_tle_state_ecef": (_i32, [P(Tle), _d, P(_d), P(_d)]),
        "lpk_grid_fill_tle": (_i32, [P(Tle), _d, _d, _d, P(_d), P(_d), _u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIB = lib
    return lib


# Every symbol include/lpk.h declares (checked by the CPU tests).
EXPORTED = (
    "lpk_abi_version", "lpk_ctx_create", "lpk_ctx_destroy", "lpk_last_error",
    "lpk_scene_upload", "lpk_render_chunk", "lpk_sync", "lpk_clock_rewind", "lpk_render_taps",
    "lpk_last_kernel_ms", "lpk_timing_reset", "lpk_timing_read", "lpk_device_alloc", "lpk_device_free", "lpk_copy_to_host",
    "lpk_device_fnv1a", "lpk_geodetic_to_ecef", "lpk_grid_size", "lpk_grid_fill_kepler",
    "lpk_ca_code", "lpk_kepler_position", "lpk_nav_fold_seed", "lpk_payload_bpsk", "lpk_payload_ook",
    "lpk_payload_qam16", "lpk_elapsed_periods", "lpk_frame_anchor", "lpk_cw_anchor_phase",
    "lpk_bands_overlap", "lpk_thermal_noise_power_w",
    "lpk_set_tuning", "lpk_get_tuning", "lpk_autotune",
    "lpk_plan_create", "lpk_plan_step", "lpk_plan_info", "lpk_plan_chunk_range", "lpk_plan_output", "lpk_plan_destroy",
    "lpk_lws_create", "lpk_lws_destroy", "lpk_lws_get_params", "lpk_lws_set_params", "lpk_lws_generate",
    "lpk_lws_infer", "lpk_lws_gradient", "lpk_lws_gradient_noalloc", "lpk_lws_train", "lpk_set_warm_start", "lpk_solve_stats",
    "lpk_solve_stats_reset", "lpk_set_horizon_gate", "lpk_horizon_gate_stats",
    "lpk_stream_to_files", "lpk_stream_to_sink", "lpk_sink_fnv1a", "lpk_fnv1a64", "lpk_host_alloc", "lpk_host_free",
    "lpk_truth",
    "lpk_gps_time", "lpk_toe_snap", "lpk_ephemeris_from_kepler", "lpk_ephemeris_from_state", "lpk_lnav_length",
    "lpk_lnav_bits", "lpk_sv_position", "lpk_grid_fill_broadcast", "lpk_rinex3_gps_ephemeris",
    "lpk_tle_parse", "lpk_sgp4_teme", "lpk_tle_state_ecef", "lpk_grid_fill_tle",
)


def check(rc: int, ctx=None) -> None:
    if rc != LPK_OK:
        msg = load().lpk_last_error(ctx)
        raise LpkError(rc, msg.decode() if msg else "")
