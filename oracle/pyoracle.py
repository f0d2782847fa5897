"""ctypes binding of the CPU FP64 oracle (oracle/liblpk_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product never imports it.

It consumes the same prepared scene (lipika_amd.pipeline.PreparedScene) the
GPU backend uploads, so the two render identical inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpk_oracle.so")

_d, _u64, _i64, _i32, _u32 = C.c_double, C.c_uint64, C.c_int64, C.c_int32, C.c_uint32
P = C.POINTER


class OrcTx(C.Structure):
    _fields_ = [
        ("traj_kind", _i32), ("signal_kind", _i32), ("pos_ecef", _d * 3),
        ("grid_t0_unix", _d), ("grid_dt_s", _d), ("grid_n", _u64), ("grid_xyz", P(_d)),
        ("carrier_hz", _d), ("tx_power_w", _d), ("bandwidth_hz", _d),
        ("baseband_offset_hz", _d), ("cw_anchored", _i32), ("_p0", _i32), ("anchor_phase_rad", _d),
        ("rate_hz", _d), ("epoch_rel_s", _d), ("repeating", _i32), ("_p1", _i32), ("repeat_from", _i64),
        ("n_symbols", _u64), ("bits", P(C.c_int8)), ("symbols", P(C.c_float)),
        ("nav_mode", _i32), ("prn", _i32), ("ca_code", P(C.c_int8)), ("nav_seed", _u64),
        ("nav_bits", P(C.c_int8)), ("n_nav_bits", _u64),
        ("gated", _i32), ("n_bursts", _u32), ("gate_period_s", _d), ("gate_phase_s", _d), ("bursts", P(_d)),
    ]


class OrcPost(C.Structure):
    _fields_ = [("kind", _i32), ("bits", _i32), ("value", _d)]


class OrcRx(C.Structure):
    _fields_ = [("pos_ecef", _d * 3), ("allan_deviation_1s", _d), ("frequency_offset_y0", _d),
                ("initial_time_offset_s", _d), ("master_rate_hz", _d), ("clock_seed", _u64),
                ("flicker_floor", _d), ("traj_kind", _i32), ("_p", _i32),
                ("kep_a_m", _d), ("kep_inc_deg", _d), ("kep_raan_deg", _d), ("kep_nu0_deg", _d),
                ("kep_epoch_unix", _d)]


class OrcBand(C.Structure):
    _fields_ = [("rx", _u32), ("n_active", _u32), ("active", P(_u32)),
                ("fs_hz", _d), ("lo_hz", _d), ("epoch_unix", _d), ("duration_s", _d),
                ("noise_power_w", _d), ("noise_seed", _u64), ("n_post", _u32), ("_p", _u32),
                ("post", OrcPost * 4),
                ("ant_on", C.c_int32), ("ant_has_cutoff", C.c_int32),
                ("ant_peak_dbi", _d), ("ant_horizon_dbi", _d), ("ant_fb_db", _d), ("ant_cutoff_deg", _d),
                ("unoccluded", P(C.c_uint8)), ("converged", C.c_int32), ("_p2", C.c_int32)]


class OrcClock(C.Structure):
    _fields_ = [("c0", _d), ("y0", _d), ("sigma_step", _d), ("master", _d), ("walk", C.c_int),
                ("seed", _u64), ("frontier", _i64), ("eps_hi", _d), ("eps_lo", _d),
                ("flick", C.c_int), ("f_alpha", _d * 20), ("f_gain", _d * 20), ("f_state", _d * 20),
                ("f_scale", _d), ("f_frontier", _i64), ("f_hi", _d), ("f_lo", _d)]


_LIB = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_geodetic_to_ecef": (None, [_d, _d, _d, P(_d)]),
        "orc_kepler_position": (None, [_d, _d, _d, _d, _d, _d, P(_d)]),
        "orc_rx_position": (None, [P(OrcRx), _d, P(_d)]),
        "orc_grid_size": (_u64, [_d, _d]),
        "orc_grid_sample_kepler": (_u64, [_d, _d, _d, _d, _d, _d, _d, _d, P(_d), P(_d)]),
        "orc_grid_position": (None, [_d, _d, _u64, P(_d), _d, P(_d)]),
        "orc_solve_emission": (None, [P(OrcTx), P(_d), _d, _d, C.c_int, P(_d), P(_d)]),
        "orc_solve_converged": (C.c_int, [P(OrcTx), P(_d), _d, _d, _d, _d, C.c_int, P(_d), P(_d)]),
        "orc_ca_code": (C.c_int, [C.c_int, P(C.c_int8)]),
        "orc_nav_dummy_bit_pm": (C.c_int8, [_u64, _i64]),
        "orc_nav_fold_seed": (_u64, [_u64, _u64]),
        "orc_elapsed_periods": (C.c_int, [_d, _d, P(_i64), P(_d)]),
        "orc_frame_anchor": (C.c_int, [_d, _d, _d, _u64, P(_d), P(_i64)]),
        "orc_cw_anchor_phase": (C.c_int, [_d, _d, _d, P(_d)]),
        "orc_gate_is_on": (C.c_int, [P(OrcTx), _d]),
        "orc_baseband": (None, [P(OrcTx), P(_d), _u64, P(C.c_float)]),
        "orc_bands_overlap": (C.c_int, [_d, _d, _d, _d]),
        "orc_xoshiro_seed": (None, [_u64, P(_u64)]),
        "orc_xoshiro_next": (_u64, [P(_u64)]),
        "orc_payload_bpsk": (None, [P(C.c_int8), _u64, _u64, _u64, P(C.c_int8)]),
        "orc_payload_ook": (C.c_int, [P(C.c_uint8), _u64, _u64, _u64, P(C.c_int8)]),
        "orc_payload_qam16": (C.c_int, [P(C.c_uint8), _u64, _u64, _u64, P(C.c_float)]),
        "orc_philox4x64": (None, [P(_u64), P(_u64), P(_u64)]),
        "orc_philox4x32": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "orc_box_muller": (None, [_u64, P(C.c_float), P(C.c_float)]),
        "orc_clock_normal": (_d, [_u64, _u64]),
        "orc_flicker_normal": (_d, [_u64, _u64]),
        "orc_flicker_poles": (None, [P(_d), P(_d)]),
        "orc_flicker_unit_adev": (_d, [P(_d), C.c_int]),
        "orc_antenna_amplitude": (C.c_float, [P(OrcBand), _d]),
        "orc_arrival_elevation_deg": (_d, [P(_d), P(_d)]),
        "orc_awgn_normals": (None, [_u64, _u64, P(C.c_float), P(C.c_float)]),
        "orc_thermal_noise_power_w": (_d, [_d, _d]),
        "orc_clock_init": (None, [P(OrcClock), P(OrcRx)]),
        "orc_clock_delta_t_at": (_d, [P(OrcClock), _d]),
        "orc_clock_delta_t_at_master_index": (_d, [P(OrcClock), _i64]),
        "orc_post_ops": (None, [P(OrcPost), _u32, P(C.c_float), _u64, P(C.c_int16)]),
        "orc_render_band": (C.c_int, [P(OrcTx), _u32, P(OrcRx), P(OrcBand), _u64, _u64, P(C.c_float),
                                      P(C.c_int16), _u32, P(_d), P(_d), P(_d), P(_d)]),
        "orc_render_window": (C.c_int, [P(OrcTx), _u32, P(OrcBand), P(OrcRx), _u64, _u64, P(_d), P(C.c_float),
                                         P(C.c_int16), P(_d), P(_d), P(_d)]),
        "orc_render_scene_hashed": (_u64, [P(OrcTx), _u32, P(OrcRx), P(OrcBand), _u32, P(_u64), _u64,
                                           C.c_int, P(_u64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    _LIB = lib
    return lib


def _ptr(a, ct):
    return P(ct)() if a is None else a.ctypes.data_as(P(ct))


class OracleScene:
    """A PreparedScene marshalled into the oracle's structs (keeps arrays alive)."""

    def __init__(self, prepared, rng_seed_base: int):
        from lipika_amd.pipeline import clock_seed, noise_seed
        self.prepared = prepared
        self._keep = []
        txs = prepared.transmitters
        self.n_tx = len(txs)
        self.txs = (OrcTx * max(self.n_tx, 1))()
        for i, t in enumerate(txs):
            d = self.txs[i]
            d.traj_kind, d.signal_kind = t.traj_kind, t.signal_kind
            d.pos_ecef[:] = [float(v) for v in t.pos]
            if t.grid is not None:
                g = np.ascontiguousarray(t.grid, dtype=np.float64)
                self._keep.append(g)
                d.grid_t0_unix, d.grid_dt_s, d.grid_n, d.grid_xyz = t.grid_t0, t.grid_dt, g.shape[1], _ptr(g, _d)
            d.carrier_hz, d.tx_power_w, d.bandwidth_hz = t.carrier_hz, t.tx_power_w, t.bandwidth_hz
            d.baseband_offset_hz, d.cw_anchored, d.anchor_phase_rad = t.baseband_offset_hz, int(t.cw_anchored), t.anchor_phase
            d.rate_hz, d.epoch_rel_s, d.repeating, d.repeat_from = t.rate_hz, t.epoch_rel_s, int(t.repeating), t.repeat_from
            d.n_symbols = t.n_symbols
            for attr, ct, dt in (("bits", C.c_int8, np.int8), ("symbols", C.c_float, np.float32),
                                 ("ca_code", C.c_int8, np.int8)):
                v = getattr(t, attr)
                if v is not None:
                    a = np.ascontiguousarray(v, dtype=dt)
                    self._keep.append(a)
                    setattr(d, attr, _ptr(a, ct))
            d.nav_mode, d.prn, d.nav_seed = t.nav_mode, t.prn, t.nav_seed
            if getattr(t, "nav_bits", None) is not None:
                nb = np.ascontiguousarray(t.nav_bits, dtype=np.int8)
                self._keep.append(nb)
                d.nav_bits, d.n_nav_bits = _ptr(nb, C.c_int8), len(nb)
            if t.gated:
                b = np.ascontiguousarray(t.bursts, dtype=np.float64)
                self._keep.append(b)
                d.gated, d.n_bursts, d.gate_period_s, d.gate_phase_s, d.bursts = 1, len(b), t.gate_period, t.gate_phase, _ptr(b, _d)
        rxs = prepared.receivers
        self.rxs = (OrcRx * max(len(rxs), 1))()
        bands = []
        for r_idx, r in enumerate(rxs):
            d = self.rxs[r_idx]
            if r.kepler is None:
                d.pos_ecef[:] = [float(v) for v in r.pos]
            else:
                d.traj_kind = 1
                (d.kep_a_m, d.kep_inc_deg, d.kep_raan_deg, d.kep_nu0_deg, d.kep_epoch_unix) = r.kepler
            d.allan_deviation_1s = r.clock.allan_deviation_1s
            d.frequency_offset_y0 = r.clock.frequency_offset_y0
            d.initial_time_offset_s = r.clock.initial_time_offset_s
            d.master_rate_hz = r.master_rate_hz
            d.clock_seed = clock_seed(rng_seed_base, r_idx)
            d.flicker_floor = r.clock.flicker_floor
            for pb in r.bands:
                bands.append((r_idx, pb))
        self.units = bands
        self.bands = (OrcBand * max(len(bands), 1))()
        for w, (r_idx, pb) in enumerate(bands):
            b = self.bands[w]
            act = np.ascontiguousarray(pb.active, dtype=np.uint32)
            self._keep.append(act)
            b.rx, b.n_active, b.active = r_idx, len(act), _ptr(act, _u32)
            b.fs_hz, b.lo_hz, b.epoch_unix = pb.band.sample_rate_hz, pb.band.lo_carrier_hz, prepared.meta.epoch_unix
            b.duration_s = prepared.meta.duration_s
            b.noise_power_w, b.noise_seed = pb.noise_power_w, noise_seed(rng_seed_base, w)
            b.n_post = len(pb.post)
            for k, (kind, bits, value) in enumerate(pb.post):
                b.post[k].kind, b.post[k].bits, b.post[k].value = kind, bits, value
            if pb.antenna is not None:
                peak, horizon, fb, cutoff = pb.antenna
                b.ant_on, b.ant_peak_dbi, b.ant_horizon_dbi, b.ant_fb_db = 1, peak, horizon, fb
                if cutoff is not None:
                    b.ant_has_cutoff, b.ant_cutoff_deg = 1, cutoff

    def render_band(self, work_idx: int, n: int | None = None, chunk_samples: int = 1 << 20,
                    tap_slot: int | None = None, codes: bool = False, unoccluded: np.ndarray | None = None,
                    converged: bool = False):
        """Renders samples [0, n) of unit `work_idx`; returns (iq, codes, taps).
        `unoccluded` (uint8 [n_active], zeroed by the caller) receives the
        per-transmitter flags. `converged` swaps the reference's two-iteration
        solve for the converged one (orc_solve_converged)."""
        lib = load()
        r_idx, pb = self.units[work_idx]
        self.bands[work_idx].unoccluded = _ptr(unoccluded, C.c_uint8)
        self.bands[work_idx].converged = 1 if converged else 0
        n = pb.n if n is None else n
        out = np.empty(n, dtype=np.complex64)
        cd = np.empty(2 * n, dtype=np.int16) if codes else None
        taps = None
        args = [P(_d)()] * 4
        if tap_slot is not None:
            taps = [np.empty(n, dtype=np.float64) for _ in range(4)]
            args = [_ptr(a, _d) for a in taps]
        rc = lib.orc_render_band(self.txs, self.n_tx, C.byref(self.rxs[r_idx]), C.byref(self.bands[work_idx]),
                                 n, chunk_samples, out.ctypes.data_as(P(C.c_float)), _ptr(cd, C.c_int16),
                                 tap_slot or 0, *args)
        if rc != 0:
            raise ValueError("oracle rejected the band")
        return out, cd, taps

    def render_window(self, work_idx: int, i0: int, delta_t: np.ndarray, taps: bool = False,
                      converged: bool = False, codes: bool = False):
        """Samples [i0, i0 + len(delta_t)) of unit `work_idx` with the given
        receiver clock offsets (orc_render_window); returns (iq, codes, taps)
        with taps = (t_emit, range, phase), each [n_active, n], or None."""
        lib = load()
        r_idx, pb = self.units[work_idx]
        dt = np.ascontiguousarray(delta_t, dtype=np.float64)
        n = len(dt)
        band = self.bands[work_idx]
        band.unoccluded = P(C.c_uint8)()
        band.converged = 1 if converged else 0
        out = np.empty(n, dtype=np.complex64)
        cd = np.empty(2 * n, dtype=np.int16) if codes else None
        tp = [np.empty((len(pb.active), n), dtype=np.float64) for _ in range(3)] if taps else [None] * 3
        rc = lib.orc_render_window(self.txs, self.n_tx, C.byref(band), C.byref(self.rxs[r_idx]), i0, n, _ptr(dt, _d),
                                   out.ctypes.data_as(P(C.c_float)), _ptr(cd, C.c_int16), *[_ptr(a, _d) for a in tp])
        band.converged = 0
        if rc != 0:
            raise ValueError("oracle rejected the window")
        return out, cd, (tuple(tp) if taps else None)

    def render_scene_hashed(self, n_limit, chunk_samples: int, n_threads: int):
        lib = load()
        nl = (_u64 * len(self.units))(*n_limit)
        hashes = (_u64 * len(self.units))()
        contrib = lib.orc_render_scene_hashed(self.txs, self.n_tx, self.rxs, self.bands, len(self.units), nl,
                                              chunk_samples, n_threads, hashes)
        return int(contrib), list(hashes)
