"""Host-side mirror of the reference's scene orchestration
(src/pipeline.rs): scene preparation (Transmitter/Receiver::from_config),
per-(receiver, band) work units, seeds, band-overlap gating, and the chunk
loop -- with the chunk body running on the GPU through the C ABI
(include/lpk.h). Names, arguments and error behaviour follow the reference:
``run_scene`` (:443), ``render_window`` (:709), ``WindowOptions`` (:624).

There is no CPU fallback: without ``liblpk.so`` or a gfx950 device every
render raises.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from dataclasses import dataclass, field
from enum import Enum
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _lib
from .config import (
    Bpsk, CaCode, CircularKeplerian, Cw, Ook, Qam16, ReceiverBandConfig, ReceiverConfig, RinexNav,
    SceneConfig, SceneMeta, SkyConfig, StationaryEcef, StationaryGeodetic, Tle, TransmitterConfig,
    CA_CHIP_RATE_HZ,
)

NEWTON_ITERATIONS = 2                 # src/pipeline.rs:56
TX_TRAJECTORY_GRID_DT_S = 0.01        # src/pipeline.rs:62
DEFAULT_CHUNK_SAMPLES = 1 << 20       # src/pipeline.rs:367
DEFAULT_WINDOW_CHUNK_SAMPLES = 1 << 16  # src/pipeline.rs:637
DEFAULT_SIGNAL_EPOCH_UNIX = 0.0       # src/pipeline.rs:139
CA_CODE_LENGTH = 1023
U64 = (1 << 64) - 1


class TimeBase(Enum):
    """src/pipeline.rs:124-135"""
    SCENE = 0
    ABSOLUTE = 1


def rust_round(x: float) -> int:
    """f64::round (half away from zero) for the sample count, src/pipeline.rs:856."""
    r = math.floor(x)
    if x - r >= 0.5:
        r += 1
    return int(r)


def clock_seed(base: int, r_idx: int) -> int:
    """src/pipeline.rs:825-827"""
    return (base + 0xC10C_0000_0000_0000 + r_idx) & U64


def noise_seed(base: int, work_idx: int) -> int:
    """src/pipeline.rs:828-830"""
    return (base + work_idx) & U64


# ---------------------------------------------------------------- preparation
@dataclass
class PreparedTx:
    """A transmitter as Transmitter::from_config leaves it (src/pipeline.rs:297-341)."""
    name: str
    carrier_hz: float
    tx_power_w: float
    bandwidth_hz: float
    signal_kind: int
    traj_kind: int
    pos: np.ndarray = field(default_factory=lambda: np.zeros(3))
    grid_t0: float = 0.0
    grid_dt: float = 0.0
    grid: Optional[np.ndarray] = None          # (3, n) float64
    baseband_offset_hz: float = 0.0
    cw_anchored: bool = False
    anchor_phase: float = 0.0
    rate_hz: float = 0.0
    epoch_rel_s: float = 0.0
    repeating: bool = False
    repeat_from: int = 0
    bits: Optional[np.ndarray] = None          # int8
    symbols: Optional[np.ndarray] = None       # float32 (n, 2)
    nav_mode: int = _lib.NAV_OFF
    prn: int = 0
    ca_code: Optional[np.ndarray] = None       # int8 (1023)
    nav_seed: int = 0
    gated: bool = False
    gate_period: float = 0.0
    gate_phase: float = 0.0
    bursts: Optional[np.ndarray] = None        # float64 (k, 2): (start, end)
    # (a m, i deg, raan deg, nu0 deg, epoch unix) of a CircularKeplerian trajectory
    kepler: Optional[tuple] = None
    nav_bits: Optional[np.ndarray] = None      # int8 +-1 LNAV bits (nav = "full")

    @property
    def n_symbols(self) -> int:
        if self.signal_kind == _lib.SIG_QAM16:
            return 0 if self.symbols is None else len(self.symbols)
        return 0 if self.bits is None else len(self.bits)


@dataclass
class PreparedBand:
    band: ReceiverBandConfig
    n: int
    active: np.ndarray            # uint32 transmitter indices, summation order
    noise_power_w: float
    post: list                    # [(kind, bits, value)]
    antenna: Optional[tuple] = None  # (peak_dbi, horizon_dbi, fb_db, cutoff_deg or None)


@dataclass
class PreparedRx:
    name: str
    pos: Optional[np.ndarray]     # a stationary receiver's position, else None
    clock: object
    master_rate_hz: float
    bands: List[PreparedBand]
    # a CircularKeplerian receiver's elements (a, inclination, node, true
    # anomaly at the epoch, epoch), evaluated per sample on the device
    kepler: Optional[tuple] = None


@dataclass
class PreparedScene:
    meta: SceneMeta
    transmitters: List[PreparedTx]
    receivers: List[PreparedRx]


def _rem_euclid(a: float, b: float) -> float:
    r = math.fmod(a, b)
    return r + abs(b) if r < 0.0 else r


def station_ecef(traj) -> np.ndarray:
    lib = _lib.load()
    if isinstance(traj, StationaryGeodetic):
        out = (C.c_double * 3)()
        lib.lpk_geodetic_to_ecef(traj.lat_deg, traj.lon_deg, traj.h_m, out)
        return np.array(out[:], dtype=np.float64)
    if isinstance(traj, StationaryEcef):
        return np.array([traj.x_m, traj.y_m, traj.z_m], dtype=np.float64)
    raise NotImplementedError("only stationary receivers are on the GPU path")


def _check(rc: int) -> None:
    _lib.check(rc)


def load_rinex_ephemeris(traj: RinexNav, meta: SceneMeta) -> "_lib.Ephemeris":
    """load_rinex_ephemeris (src/pipeline.rs:109-121): the record of the PRN
    whose toe is nearest the scene epoch."""
    with open(traj.file, "r", encoding="utf-8") as f:
        text = f.read()
    eph = _lib.Ephemeris()
    rc = _lib.load().lpk_rinex3_gps_ephemeris(text.encode(), int(traj.prn), meta.epoch_unix, meta.gps_utc_offset_s,
                                              C.byref(eph))
    if rc != _lib.LPK_OK:
        raise ValueError(f"no GPS PRN {traj.prn} ephemeris found in {traj.file!r}")
    return eph


def parse_tle(traj: Tle) -> "_lib.Tle":
    """TleTrajectory::from_tle (src/geometry/tle.rs:33-44): the element set,
    or ValueError as the reference's "TLE parse" error."""
    e = _lib.Tle()
    if _lib.load().lpk_tle_parse(traj.line1.encode(), traj.line2.encode(), C.byref(e)) != _lib.LPK_OK:
        raise ValueError(f"TLE parse: {traj.line1!r} / {traj.line2!r} is not a valid element set")
    return e


def trajectory_state(traj, t_unix: float):
    """Trajectory::position_ecef and velocity_ecef (src/geometry/mod.rs:62-74,
    src/geometry/kepler.rs:27-96, src/geometry/tle.rs:64-85) of the
    trajectories this mirror builds."""
    if isinstance(traj, (StationaryGeodetic, StationaryEcef)):
        return station_ecef(traj), np.zeros(3)
    if isinstance(traj, CircularKeplerian):
        # CircularKeplerian::eci_at + eci_to_ecef in the reference's operation order
        a = traj.semi_major_axis_m
        n = math.sqrt(3.986005e14 / (a * a * a))
        nu = math.radians(traj.true_anomaly_deg_at_epoch) + n * (t_unix - traj.epoch_unix)
        s_nu, c_nu = math.sin(nu), math.cos(nu)
        s_i, c_i = math.sin(math.radians(traj.inclination_deg)), math.cos(math.radians(traj.inclination_deg))
        s_o, c_o = math.sin(math.radians(traj.raan_deg)), math.cos(math.radians(traj.raan_deg))
        xp, yp = a * c_nu, a * s_nu
        y1, z1 = yp * c_i, yp * s_i
        x, y = xp * c_o - y1 * s_o, xp * s_o + y1 * c_o
        v_mag = a * n
        vxp, vyp = -v_mag * s_nu, v_mag * c_nu
        vy1, vz1 = vyp * c_i, vyp * s_i
        vx, vy = vxp * c_o - vy1 * s_o, vxp * s_o + vy1 * c_o
        w = 7.2921151467e-5
        th = w * (t_unix - traj.epoch_unix)
        s_t, c_t = math.sin(th), math.cos(th)
        xe, ye = x * c_t + y * s_t, -x * s_t + y * c_t
        vxe = (vx * c_t + vy * s_t) + w * ye
        vye = (-vx * s_t + vy * c_t) - w * xe
        return np.array([xe, ye, z1]), np.array([vxe, vye, vz1])
    if isinstance(traj, Tle):
        pos, vel = (C.c_double * 3)(), (C.c_double * 3)()
        _check(_lib.load().lpk_tle_state_ecef(C.byref(parse_tle(traj)), t_unix, pos, vel))
        return np.array(pos[:]), np.array(vel[:])
    raise NotImplementedError(f"trajectory {traj!r} has no state here")


def build_lnav_bits(cfg: TransmitterConfig, meta: SceneMeta, prn: int) -> np.ndarray:
    """build_lnav_bits (src/pipeline.rs:255-295): the scene's LNAV stream from
    the ephemeris of the transmitter's own orbit."""
    lib = _lib.load()
    week, sow = C.c_uint32(), C.c_double()
    lib.lpk_gps_time(meta.epoch_unix, meta.gps_utc_offset_s, C.byref(week), C.byref(sow))
    if math.fmod(sow.value / 6.0, 1.0) != 0.0:
        warnings.warn(f"scene epoch is not on a 6 s GPS boundary (sow={sow.value}); full-NAV subframe edges won't "
                      "align to the TOW Z-count and SV-position time will be biased. Use an epoch_unix divisible by 6.")
    traj = cfg.trajectory
    eph = _lib.Ephemeris()
    if isinstance(traj, CircularKeplerian):
        _check(lib.lpk_ephemeris_from_kepler(traj.semi_major_axis_m, traj.inclination_deg, traj.raan_deg,
                                             traj.true_anomaly_deg_at_epoch, traj.epoch_unix, prn, meta.epoch_unix,
                                             meta.gps_utc_offset_s, C.byref(eph)))
    elif isinstance(traj, RinexNav):
        eph = load_rinex_ephemeris(traj, meta)
    else:
        t_toe, wk, toe = C.c_double(), C.c_uint32(), C.c_double()
        _check(lib.lpk_toe_snap(meta.epoch_unix, meta.gps_utc_offset_s, C.byref(t_toe), C.byref(wk), C.byref(toe)))
        r, v = trajectory_state(traj, t_toe.value)
        _check(lib.lpk_ephemeris_from_state((C.c_double * 3)(*r), (C.c_double * 3)(*v), prn, wk.value, toe.value,
                                            C.byref(eph)))
    n = int(lib.lpk_lnav_length(meta.duration_s))
    bits = np.empty(n, dtype=np.int8)
    _check(lib.lpk_lnav_bits(C.byref(eph), sow.value, meta.duration_s, bits.ctypes.data_as(C.POINTER(C.c_int8)), n))
    return bits


def build_transmitter(cfg: TransmitterConfig, meta: SceneMeta, time_base: TimeBase) -> PreparedTx:
    """Transmitter::from_config + build_signal (src/pipeline.rs:142-253, :312-341)."""
    lib = _lib.load()
    sig = cfg.kind
    kind = {Cw: _lib.SIG_CW, Bpsk: _lib.SIG_BPSK, Ook: _lib.SIG_OOK, Qam16: _lib.SIG_QAM16,
            CaCode: _lib.SIG_CA}[type(sig)]
    tx = PreparedTx(name=cfg.name, carrier_hz=float(cfg.carrier_hz),
                    tx_power_w=math.pow(10.0, cfg.tx_power_dbw / 10.0),   # src/pipeline.rs:338
                    bandwidth_hz=float(cfg.effective_bandwidth_hz()), signal_kind=kind,
                    traj_kind=_lib.TRAJ_STATIONARY)
    traj = cfg.trajectory
    if isinstance(traj, (StationaryGeodetic, StationaryEcef)):
        tx.pos = station_ecef(traj)
    elif isinstance(traj, CircularKeplerian):
        n = int(lib.lpk_grid_size(meta.duration_s, TX_TRAJECTORY_GRID_DT_S))
        grid = np.empty((3, n), dtype=np.float64)
        t0 = C.c_double()
        _check(lib.lpk_grid_fill_kepler(
            traj.semi_major_axis_m, traj.inclination_deg, traj.raan_deg, traj.true_anomaly_deg_at_epoch,
            traj.epoch_unix, meta.epoch_unix, meta.duration_s, TX_TRAJECTORY_GRID_DT_S, C.byref(t0),
            grid.ctypes.data_as(C.POINTER(C.c_double)), n))
        tx.traj_kind = _lib.TRAJ_GRID
        tx.grid, tx.grid_t0, tx.grid_dt = grid, t0.value, TX_TRAJECTORY_GRID_DT_S
        tx.kepler = (traj.semi_major_axis_m, traj.inclination_deg, traj.raan_deg, traj.true_anomaly_deg_at_epoch,
                     traj.epoch_unix)
    elif isinstance(traj, RinexNav):
        eph = load_rinex_ephemeris(traj, meta)
        n = int(lib.lpk_grid_size(meta.duration_s, TX_TRAJECTORY_GRID_DT_S))
        grid = np.empty((3, n), dtype=np.float64)
        t0 = C.c_double()
        _check(lib.lpk_grid_fill_broadcast(C.byref(eph), meta.gps_utc_offset_s, meta.epoch_unix, meta.duration_s,
                                           TX_TRAJECTORY_GRID_DT_S, C.byref(t0),
                                           grid.ctypes.data_as(C.POINTER(C.c_double)), n))
        tx.traj_kind = _lib.TRAJ_GRID
        tx.grid, tx.grid_t0, tx.grid_dt = grid, t0.value, TX_TRAJECTORY_GRID_DT_S
    elif isinstance(traj, Tle):
        tle = parse_tle(traj)
        n = int(lib.lpk_grid_size(meta.duration_s, TX_TRAJECTORY_GRID_DT_S))
        grid = np.empty((3, n), dtype=np.float64)
        t0 = C.c_double()
        if lib.lpk_grid_fill_tle(C.byref(tle), meta.epoch_unix, meta.duration_s, TX_TRAJECTORY_GRID_DT_S,
                                 C.byref(t0), grid.ctypes.data_as(C.POINTER(C.c_double)), n) != _lib.LPK_OK:
            raise ValueError(f"transmitter {cfg.name!r}: sgp4 propagation failed over the scene")
        tx.traj_kind = _lib.TRAJ_GRID
        tx.grid, tx.grid_t0, tx.grid_dt = grid, t0.value, TX_TRAJECTORY_GRID_DT_S
    else:
        raise NotImplementedError(f"trajectory {traj!r} is not on the GPU path")

    render_epoch = meta.epoch_unix
    anchor_epoch = cfg.signal_epoch_unix
    if anchor_epoch is None and time_base == TimeBase.ABSOLUTE:
        anchor_epoch = DEFAULT_SIGNAL_EPOCH_UNIX

    def frame_anchor(rate: float, n_sym: int):
        rel, off = C.c_double(), C.c_int64()
        _check(lib.lpk_frame_anchor(render_epoch, anchor_epoch, rate, n_sym, C.byref(rel), C.byref(off)))
        return rel.value, off.value

    if isinstance(sig, Cw):
        tx.baseband_offset_hz = float(sig.baseband_offset_hz)
        if anchor_epoch is not None:
            ph = C.c_double()
            _check(lib.lpk_cw_anchor_phase(render_epoch, anchor_epoch, sig.baseband_offset_hz, C.byref(ph)))
            tx.cw_anchored, tx.anchor_phase = True, ph.value
    elif isinstance(sig, (Bpsk, Ook, Qam16)):
        # payload lengths are usize in the reference (src/config/mod.rs:200-208);
        # the generators take them as uint32
        n_payload = sig.payload_symbols if isinstance(sig, Qam16) else sig.payload_bits
        if not 0 <= int(n_payload) < 2 ** 32:
            raise ValueError(f"payload length {n_payload} out of range (transmitter {cfg.name!r})")
        if isinstance(sig, Qam16):
            sync = np.asarray(sig.sync_word, dtype=np.int64)
            if np.any((sync < 0) | (sync > 1)):
                raise ValueError(f"qam16 sync_word entries must be 0 or 1 (transmitter {cfg.name!r})")
            if len(sync) % 4:
                raise ValueError(f"qam16 sync_word must hold a whole number of symbols (transmitter {cfg.name!r})")
            s8 = sync.astype(np.uint8)
            n_sym = len(s8) // 4 + sig.payload_symbols
            out = np.empty((n_sym, 2), dtype=np.float32)
            _check(lib.lpk_payload_qam16(s8.ctypes.data_as(C.POINTER(C.c_uint8)), len(s8), sig.payload_symbols,
                                         sig.payload_seed & U64, out.ctypes.data_as(C.POINTER(C.c_float))))
            tx.symbols, rate = out, float(sig.symbol_rate_baud)
        elif isinstance(sig, Ook):
            sync = np.asarray(sig.sync_word, dtype=np.int64)
            if np.any((sync < 0) | (sync > 1)):
                raise ValueError(f"ook sync_word entries must be 0 or 1 (transmitter {cfg.name!r})")
            s8 = sync.astype(np.uint8)
            out = np.empty(len(s8) + sig.payload_bits, dtype=np.int8)
            _check(lib.lpk_payload_ook(s8.ctypes.data_as(C.POINTER(C.c_uint8)), len(s8), sig.payload_bits,
                                       sig.payload_seed & U64, out.ctypes.data_as(C.POINTER(C.c_int8))))
            tx.bits, rate = out, float(sig.bit_rate_bps)
        else:
            s8 = np.asarray(sig.sync_word, dtype=np.int8)
            out = np.empty(len(s8) + sig.payload_bits, dtype=np.int8)
            _check(lib.lpk_payload_bpsk(s8.ctypes.data_as(C.POINTER(C.c_int8)), len(s8), sig.payload_bits,
                                        sig.payload_seed & U64, out.ctypes.data_as(C.POINTER(C.c_int8))))
            tx.bits, rate = out, float(sig.bit_rate_bps)
        tx.rate_hz = rate
        tx.epoch_rel_s = 0.0   # bit 0 at the scene epoch
        if anchor_epoch is not None:
            tx.epoch_rel_s, tx.repeat_from = frame_anchor(rate, tx.n_symbols)
            tx.repeating = True
    else:  # C/A
        code = np.empty(CA_CODE_LENGTH, dtype=np.int8)
        if not 1 <= sig.prn <= 32:
            raise ValueError(f"PRN must be 1..32, got {sig.prn}")
        _check(lib.lpk_ca_code(sig.prn, code.ctypes.data_as(C.POINTER(C.c_int8))))
        tx.ca_code, tx.prn = code, sig.prn
        if time_base == TimeBase.SCENE:
            tx.epoch_rel_s = sig.epoch_unix - meta.epoch_unix
            if sig.nav == "off":
                tx.nav_mode = _lib.NAV_OFF
            elif sig.nav == "dummy":
                tx.nav_mode = _lib.NAV_DUMMY
                tx.nav_seed = int(lib.lpk_nav_fold_seed(sig.nav_seed & U64, sig.prn))
            elif sig.nav == "full":
                tx.nav_mode = _lib.NAV_FULL
                tx.nav_bits = build_lnav_bits(cfg, meta, sig.prn)
            else:
                raise ValueError(f"transmitter {cfg.name!r}: unknown nav mode {sig.nav!r}")
        else:
            if sig.nav != "off":
                raise ValueError(f"transmitter '{cfg.name}': ca_code with nav = \"{sig.nav}\" is not supported "
                                 "in a window render; use nav = \"off\" (pure code)")
            code_rate = CA_CHIP_RATE_HZ / CA_CODE_LENGTH
            rel, off = C.c_double(), C.c_int64()
            _check(lib.lpk_frame_anchor(render_epoch, sig.epoch_unix, code_rate, 1, C.byref(rel), C.byref(off)))
            tx.epoch_rel_s = rel.value
    if cfg.gate is not None:
        g = cfg.gate
        if not (g.period_s > 0.0) or not math.isfinite(g.period_s):
            raise ValueError(f"transmitter '{cfg.name}': gate period_s must be positive and finite, got {g.period_s}")
        bursts = []
        for off_s, dur_s in g.bursts:
            if not (off_s >= 0.0) or not (dur_s > 0.0) or not (off_s + dur_s <= g.period_s):
                raise ValueError(f"transmitter '{cfg.name}': gate burst (offset_s = {off_s}, duration_s = {dur_s}) "
                                 "must lie within one period")
            bursts.append((off_s, off_s + dur_s))
        phase = render_epoch - g.epoch_unix                          # src/signal/gate.rs:56-66
        if not math.isfinite(phase):
            raise ValueError(f"transmitter '{cfg.name}': gate phase_s must be finite")
        tx.gated, tx.gate_period = True, float(g.period_s)
        tx.gate_phase = _rem_euclid(phase, g.period_s)
        tx.bursts = np.asarray(bursts, dtype=np.float64).reshape(-1, 2)
    return tx


def _post_ops(band: ReceiverBandConfig) -> list:
    """RxChain::from_config, src/rx/mod.rs:117-137."""
    ops = []
    fe = band.frontend
    if fe is None:
        return ops
    if fe.gain_db is not None:
        ops.append((_lib.POST_GAIN, 0, float(fe.gain_db)))
    if fe.nonlinearity is not None:
        ops.append((_lib.POST_NONLINEARITY, 0, math.pow(10.0, (fe.nonlinearity.iip3_dbm - 30.0) / 10.0)))
    if fe.adc is not None:
        ops.append((_lib.POST_ADC, int(fe.adc.bits), float(fe.adc.full_scale_v)))
    return ops


def _antenna(band: ReceiverBandConfig) -> Optional[tuple]:
    """The band's antenna pattern (src/rx/mod.rs:117-125), or None."""
    fe = band.frontend
    if fe is None or fe.antenna is None:
        return None
    a = fe.antenna
    vals = [float(a["peak_dbi"]), float(a["horizon_dbi"]), float(a["fb_db"])]
    cutoff = a.get("cutoff_deg")
    if not all(math.isfinite(v) for v in vals) or (cutoff is not None and not math.isfinite(float(cutoff))):
        raise ValueError(f"band '{band.name}': antenna parameters must be finite")
    return (*vals, None if cutoff is None else float(cutoff))


def prepare(scene: SceneConfig, time_base: TimeBase = TimeBase.SCENE) -> PreparedScene:
    """src/pipeline.rs:801-819, plus the per-band set-up of generate_band (:853-894)."""
    lib = _lib.load()
    meta = scene.scene
    txs = [build_transmitter(t, meta, time_base) for t in scene.transmitters]
    rxs = []
    for r in scene.receivers:
        if not r.bands:
            raise ValueError(f"receiver '{r.name}' has no bands; at least one [[receiver.band]] required")
        kepler = None
        if isinstance(r.trajectory, CircularKeplerian):
            k = r.trajectory
            kepler = tuple(float(v) for v in (k.semi_major_axis_m, k.inclination_deg, k.raan_deg,
                                              k.true_anomaly_deg_at_epoch, k.epoch_unix))
            if not all(math.isfinite(v) for v in kepler) or not kepler[0] > 0.0:
                raise ValueError(f"receiver '{r.name}': Keplerian elements must be finite with a positive radius")
            pos = None
        elif isinstance(r.trajectory, (StationaryGeodetic, StationaryEcef)):
            pos = station_ecef(r.trajectory)
        else:
            raise NotImplementedError(f"receiver '{r.name}': only stationary and circular Keplerian receivers "
                                      f"are on the GPU path")
        master = 0.0
        for b in r.bands:
            master = max(master, float(b.sample_rate_hz))
        bands = []
        for b in r.bands:
            n = rust_round(b.sample_rate_hz * meta.duration_s)
            bw = b.effective_bandwidth_hz()
            active = [i for i, t in enumerate(txs)
                      if lib.lpk_bands_overlap(t.carrier_hz, t.bandwidth_hz, b.lo_carrier_hz, bw)]
            p = b.noise_power_w if b.noise_power_w is not None else \
                float(lib.lpk_thermal_noise_power_w(b.noise_temperature_k, b.sample_rate_hz))
            bands.append(PreparedBand(b, n, np.asarray(active, dtype=np.uint32), float(p), _post_ops(b),
                                      _antenna(b)))
        rxs.append(PreparedRx(r.name, pos, r.clock, master, bands, kepler))
    return PreparedScene(meta, txs, rxs)


# ---------------------------------------------------------------- backend
def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.POINTER(ctype)()
    return a.ctypes.data_as(C.POINTER(ctype))


class Backend:
    """One HIP context (one device, one stream): the ``Backend`` slot of
    src/backend/mod.rs:14-18 filled with the gfx950 implementation."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.lpk_ctx_create(device, C.byref(h)), None)
        self.h = h
        self._keep: list = []
        self.scene: Optional[PreparedScene] = None

    def close(self) -> None:
        if self.h:
            self.lib.lpk_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc: int) -> None:
        _lib.check(rc, self.h)

    def upload(self, scene: PreparedScene, rng_seed_base: int) -> None:
        n_tx, n_rx = len(scene.transmitters), len(scene.receivers)
        tx_arr = (_lib.TxDesc * max(n_tx, 1))()
        keep = []
        for i, t in enumerate(scene.transmitters):
            d = tx_arr[i]
            d.traj_kind, d.signal_kind = t.traj_kind, t.signal_kind
            d.pos_ecef[:] = [float(v) for v in t.pos]
            if t.grid is not None:
                g = np.ascontiguousarray(t.grid, dtype=np.float64)
                keep.append(g)
                d.grid_t0_unix, d.grid_dt_s, d.grid_n = t.grid_t0, t.grid_dt, g.shape[1]
                d.grid_xyz = _ptr(g, C.c_double)
            d.carrier_hz, d.tx_power_w, d.bandwidth_hz = t.carrier_hz, t.tx_power_w, t.bandwidth_hz
            d.baseband_offset_hz, d.cw_anchored, d.anchor_phase_rad = t.baseband_offset_hz, int(t.cw_anchored), t.anchor_phase
            d.rate_hz, d.epoch_rel_s = t.rate_hz, t.epoch_rel_s
            d.repeating, d.repeat_from, d.n_symbols = int(t.repeating), t.repeat_from, t.n_symbols
            if t.bits is not None:
                b = np.ascontiguousarray(t.bits, dtype=np.int8)
                keep.append(b)
                d.bits = _ptr(b, C.c_int8)
            if t.symbols is not None:
                s = np.ascontiguousarray(t.symbols, dtype=np.float32)
                keep.append(s)
                d.symbols = _ptr(s, C.c_float)
            d.nav_mode, d.prn, d.nav_seed = t.nav_mode, t.prn, t.nav_seed & U64
            if t.nav_bits is not None:
                nb = np.ascontiguousarray(t.nav_bits, dtype=np.int8)
                keep.append(nb)
                d.nav_bits, d.n_nav_bits = _ptr(nb, C.c_int8), len(nb)
            if t.ca_code is not None:
                c = np.ascontiguousarray(t.ca_code, dtype=np.int8)
                keep.append(c)
                d.ca_code = _ptr(c, C.c_int8)
            if t.gated:
                bb = np.ascontiguousarray(t.bursts, dtype=np.float64)
                keep.append(bb)
                d.gated, d.n_bursts, d.gate_period_s, d.gate_phase_s = 1, len(bb), t.gate_period, t.gate_phase
                d.bursts = _ptr(bb, C.c_double)
            if t.kepler is not None:
                d.has_kepler = 1
                (d.kepler_a_m, d.kepler_inc_deg, d.kepler_raan_deg, d.kepler_nu0_deg,
                 d.kepler_epoch_unix) = (float(v) for v in t.kepler)
        rx_arr = (_lib.RxDesc * max(n_rx, 1))()
        for r_idx, r in enumerate(scene.receivers):
            d = rx_arr[r_idx]
            if r.kepler is None:
                d.traj_kind = _lib.RX_STATIONARY
                d.pos_ecef[:] = [float(v) for v in r.pos]
            else:
                d.traj_kind = _lib.RX_KEPLER
                (d.kepler_a_m, d.kepler_inc_deg, d.kepler_raan_deg, d.kepler_nu0_deg,
                 d.kepler_epoch_unix) = r.kepler
            d.allan_deviation_1s = r.clock.allan_deviation_1s
            d.frequency_offset_y0 = r.clock.frequency_offset_y0
            d.initial_time_offset_s = r.clock.initial_time_offset_s
            d.master_rate_hz = r.master_rate_hz
            d.clock_seed = clock_seed(rng_seed_base, r_idx)
            d.flicker_floor = r.clock.flicker_floor
        self.check(self.lib.lpk_scene_upload(self.h, tx_arr, n_tx, rx_arr, n_rx))
        del keep  # tables are copied by lpk_scene_upload
        self.scene = scene

    def job(self, r_idx: int, b_idx: int, work_idx: int, rng_seed_base: int, i0: int, clen: int,
            out_host: Optional[np.ndarray] = None, out_dev: Optional[int] = None,
            codes: Optional[np.ndarray] = None, delta_t: Optional[np.ndarray] = None,
            unoccluded: Optional[np.ndarray] = None) -> tuple:
        """One lpk_band_chunk for samples [i0, i0+clen) of band b_idx of receiver r_idx.
        Returns (descriptor, keep-alive list). `unoccluded` (uint8 [n_active])
        receives the per-transmitter flags of the chunk."""
        sc = self.scene
        pb = sc.receivers[r_idx].bands[b_idx]
        j = _lib.BandChunk()
        keep = [pb.active]
        j.rx, j.n_active = r_idx, len(pb.active)
        j.active = _ptr(pb.active, C.c_uint32)
        j.fs_hz, j.lo_hz, j.epoch_unix = pb.band.sample_rate_hz, pb.band.lo_carrier_hz, sc.meta.epoch_unix
        j.i0, j.clen = i0, clen
        if delta_t is not None:
            dt = np.ascontiguousarray(delta_t, dtype=np.float64)
            keep.append(dt)
            j.delta_t = _ptr(dt, C.c_double)
        j.noise_power_w, j.noise_seed = pb.noise_power_w, noise_seed(rng_seed_base, work_idx)
        j.n_post = len(pb.post)
        for k, (kind, bits, value) in enumerate(pb.post):
            j.post[k].kind, j.post[k].bits, j.post[k].value = kind, bits, value
        if out_dev is not None:
            j.out_dev = out_dev
        if out_host is not None:
            assert out_host.dtype == np.complex64 and out_host.flags.c_contiguous and len(out_host) >= clen
            j.out_host = out_host.ctypes.data
        if codes is not None:
            assert codes.dtype == np.int16 and codes.flags.c_contiguous and codes.size >= 2 * clen
            j.adc_codes_host = _ptr(codes, C.c_int16)
        if pb.antenna is not None:
            peak, horizon, fb, cutoff = pb.antenna
            j.has_antenna, j.antenna_peak_dbi, j.antenna_horizon_dbi, j.antenna_fb_db = 1, peak, horizon, fb
            if cutoff is not None:
                j.has_cutoff, j.antenna_cutoff_deg = 1, cutoff
        if unoccluded is not None:
            assert unoccluded.dtype == np.uint8 and unoccluded.flags.c_contiguous and len(unoccluded) >= len(pb.active)
            j.unoccluded = _ptr(unoccluded, C.c_uint8)
        return j, keep

    def render(self, jobs: Sequence, form: int = _lib.FORM_PLAIN) -> None:
        arr = (_lib.BandChunk * max(len(jobs), 1))(*jobs)
        self.check(self.lib.lpk_render_chunk(self.h, arr, len(jobs), form))

    def sync(self) -> None:
        self.check(self.lib.lpk_sync(self.h))

    def last_kernel_ms(self) -> float:
        ms = C.c_double()
        self.check(self.lib.lpk_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    def set_tuning(self, tile_shift: int = 8, xcd_map: int = 1) -> None:
        """The render launch's tile size (2^tile_shift samples) and work split."""
        self.check(self.lib.lpk_set_tuning(self.h, C.byref(_lib.Tuning(tile_shift, xcd_map))))

    def tuning(self) -> tuple:
        t = _lib.Tuning()
        self.check(self.lib.lpk_get_tuning(self.h, C.byref(t)))
        return t.tile_shift, t.xcd_map

    def autotune(self, jobs: Sequence, form: int = _lib.FORM_PLAIN) -> tuple:
        """Times every candidate tuning on `jobs` (device outputs only), keeps
        the fastest and returns ((tile_shift, xcd_map), render-kernel ms)."""
        arr = (_lib.BandChunk * max(len(jobs), 1))(*jobs)
        t, ms = _lib.Tuning(), C.c_double()
        self.check(self.lib.lpk_autotune(self.h, arr, len(jobs), form, C.byref(t), C.byref(ms)))
        return (t.tile_shift, t.xcd_map), ms.value

    def plan(self, bands: Sequence, chunk_samples: int, form: int = _lib.FORM_PLAIN) -> "Plan":
        """A streaming plan over whole-band descriptors (see Plan)."""
        return Plan(self, bands, chunk_samples, form)

    def warm_start(self, net: Optional["WarmStartNet"]) -> None:
        """The network FORM_LEARNED predicts light times with (None: none)."""
        self.check(self.lib.lpk_set_warm_start(self.h, net.h if net is not None else None))
        self._warm = net

    def solve_stats(self) -> dict:
        """Converged-form counters since the last reset."""
        a = (C.c_uint64 * 4)()
        self.check(self.lib.lpk_solve_stats(self.h, a))
        return dict(evaluations=a[0], contributions=a[1], fallbacks=a[2], unconverged=a[3])

    def solve_stats_reset(self) -> None:
        """Zeroes the converged-form and horizon-gate counters."""
        self.check(self.lib.lpk_solve_stats_reset(self.h))

    def set_horizon_gate(self, on: bool) -> None:
        """Skip (tile, transmitter) pairs certainly below the antenna cutoff
        (src/channel/horizon.rs); on by default, output identical either way."""
        self.check(self.lib.lpk_set_horizon_gate(self.h, 1 if on else 0))

    def horizon_gate_stats(self) -> dict:
        a = (C.c_uint64 * 2)()
        self.check(self.lib.lpk_horizon_gate_stats(self.h, a))
        return dict(skipped=a[0], tested=a[1])

    def stream_to_sink(self, bands: Sequence, chunk_samples: int, form: int = _lib.FORM_PLAIN,
                       sink: Optional[Callable[[int, int, np.ndarray], object]] = None) -> list:
        """generate_band's chunk loop with its sink (src/pipeline.rs:905-990,
        :984) for whole-band descriptors, through lpk_stream_to_sink: chunk k
        of every band renders on the device while the sink takes chunk k - 1.
        `sink(band, i0, iq)` gets a complex64 view valid for the call only and
        returns a truthy value to stop that band. With no sink, the library's
        FNV-1a sink hashes each band; returns the per-band hashes."""
        arr = (_lib.BandChunk * max(len(bands), 1))(*bands)
        if sink is None:
            hashes = (C.c_uint64 * max(len(bands), 1))(*([_lib.FNV_OFFSET] * len(bands)))
            fn = _lib.IQ_SINK(C.cast(self.lib.lpk_sink_fnv1a, C.c_void_p).value)
            self.check(self.lib.lpk_stream_to_sink(self.h, arr, len(bands), chunk_samples, form, fn, C.cast(hashes, C.c_void_p)))
            return list(hashes)[:len(bands)]
        err = []

        def trampoline(user, band, i0, iq, n):
            try:
                view = np.ctypeslib.as_array(iq, shape=(2 * n,)).view(np.complex64)
                return 1 if sink(int(band), int(i0), view) else 0
            except BaseException as e:  # no exception crosses the C ABI
                err.append(e)
                return 1

        fn = _lib.IQ_SINK(trampoline)
        self.check(self.lib.lpk_stream_to_sink(self.h, arr, len(bands), chunk_samples, form, fn, None))
        if err:
            raise err[0]
        return []

    def taps(self, job, slot: int, form: int = _lib.FORM_PLAIN):
        n = job.clen
        te, rg, ph, dt = (np.empty(n, dtype=np.float64) for _ in range(4))
        self.check(self.lib.lpk_render_taps(self.h, C.byref(job), form, slot, _ptr(te, C.c_double),
                                            _ptr(rg, C.c_double), _ptr(ph, C.c_double), _ptr(dt, C.c_double)))
        return te, rg, ph, dt


class Plan:
    """lpk_plan_*: chunk k of every band per step, the steps alternating
    between two streams so that each overlaps the previous one's end
    (generate_band's chunk loop, src/pipeline.rs:905-990, for all bands at
    once). `bands` are whole-band descriptors from Backend.job(..., i0=0,
    clen=n, out_dev=... or None)."""

    def __init__(self, backend: "Backend", bands: Sequence, chunk_samples: int, form: int):
        self.be = backend
        self._arr = (_lib.BandChunk * max(len(bands), 1))(*bands)
        self.h = C.c_void_p()
        backend.check(backend.lib.lpk_plan_create(backend.h, self._arr, len(bands), chunk_samples, form,
                                                  C.byref(self.h)))
        n, t = C.c_uint64(), C.c_uint32()
        backend.check(backend.lib.lpk_plan_info(self.h, C.byref(n), C.byref(t)))
        self.n_chunks, self.max_tiles = n.value, t.value

    def step(self, k: int) -> None:
        self.be.check(self.be.lib.lpk_plan_step(self.h, k))

    def chunk_range(self, band: int, k: int) -> tuple:
        """(i0, clen) of band `band` in step k: the steps are time slices."""
        i0, n = C.c_uint64(), C.c_uint64()
        self.be.check(self.be.lib.lpk_plan_chunk_range(self.h, band, k, C.byref(i0), C.byref(n)))
        return i0.value, n.value

    def output(self, band: int) -> int:
        """Device address of band `band`'s output for the last step."""
        p = C.c_void_p()
        self.be.check(self.be.lib.lpk_plan_output(self.h, band, C.byref(p)))
        return p.value

    def close(self) -> None:
        if self.h:
            self.be.lib.lpk_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceArray:
    """A device buffer of the backend's context (lpk_device_alloc)."""

    def __init__(self, backend: Backend, shape, dtype):
        self.be, self.shape, self.dtype = backend, tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        backend.check(backend.lib.lpk_device_alloc(backend.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        self.be.check(self.be.lib.lpk_copy_to_host(self.be.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            self.be.lib.lpk_device_free(self.be.h, self.ptr)
            self.ptr = None


class WarmStartNet:
    """lpk_lws_*: the learned warm start's network on the backend's device.
    Training pairs come from the device's own converged light-time solve;
    training runs the device's forward (matrix cores) and backward kernels."""

    def __init__(self, backend: Backend, init_seed: int = 1):
        self.be = backend
        h = C.c_void_p()
        backend.check(backend.lib.lpk_lws_create(backend.h, init_seed & U64, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            if getattr(self.be, "_warm", None) is self:
                self.be.warm_start(None)
            self.be.lib.lpk_lws_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self) -> np.ndarray:
        th = np.empty(_lib.LWS_PARAMS, dtype=np.float32)
        self.be.check(self.be.lib.lpk_lws_get_params(self.h, _ptr(th, C.c_float)))
        return th

    @params.setter
    def params(self, theta) -> None:
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.shape == (_lib.LWS_PARAMS,)
        self.be.check(self.be.lib.lpk_lws_set_params(self.h, _ptr(th, C.c_float)))

    def generate(self, n: int, seed: int, geometry: bool = False) -> tuple:
        """(X, Y, G) device arrays of n training pairs (G None unless asked)."""
        X, Y = DeviceArray(self.be, (n, 10), np.float32), DeviceArray(self.be, (n,), np.float32)
        G = DeviceArray(self.be, (n, 8), np.float64) if geometry else None
        self.be.check(self.be.lib.lpk_lws_generate(self.be.h, n, seed & U64, X.ptr, Y.ptr, G.ptr if G else None))
        return X, Y, G

    def infer(self, X: DeviceArray, n: int) -> np.ndarray:
        Y = DeviceArray(self.be, (n,), np.float32)
        try:
            self.be.check(self.be.lib.lpk_lws_infer(self.h, X.ptr, n, Y.ptr))
            return Y.numpy()
        finally:
            Y.free()

    def gradient(self, X: DeviceArray, Y: DeviceArray, n: int) -> tuple:
        g = np.empty(_lib.LWS_PARAMS, dtype=np.float32)
        loss = C.c_double()
        self.be.check(self.be.lib.lpk_lws_gradient(self.h, X.ptr, Y.ptr, n, _ptr(g, C.c_float), C.byref(loss)))
        return loss.value, g

    def train(self, n_pairs: int = 10_000_000, data_seed: int = 2026, epochs: int = 8, batch: int = 16384,
              lr: float = 2e-3, lr_final: float = 2e-5, n_val: int = 100_000) -> dict:
        cfg = _lib.LwsTrainCfg(n_pairs, data_seed & U64, epochs, batch, lr, lr_final, n_val)
        rep = _lib.LwsTrainReport()
        self.be.check(self.be.lib.lpk_lws_train(self.h, C.byref(cfg), C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_}


_FORMS = {"plain": _lib.FORM_PLAIN, "precision": _lib.FORM_PRECISION, "converged": _lib.FORM_CONVERGED,
          "learned": _lib.FORM_LEARNED}


def form_code(form) -> int:
    return _FORMS[form] if isinstance(form, str) else int(form)


# ---------------------------------------------------------------- run_scene
@dataclass
class ObserverOutput:
    """src/output/mod.rs (name, epoch, rate, LO, IQ, descriptor)."""
    name: str
    epoch_unix: float
    sample_rate_hz: float
    lo_carrier_hz: float
    iq: np.ndarray
    descriptor: dict


def _units(sc: PreparedScene):
    w = 0
    for r_idx, r in enumerate(sc.receivers):
        for b_idx, _ in enumerate(r.bands):
            yield w, r_idx, b_idx
            w += 1


def run_scene(scene: SceneConfig, rng_seed_base: int, chunk_samples: int = DEFAULT_CHUNK_SAMPLES,
              form="plain", backend: Optional[Backend] = None) -> List[ObserverOutput]:
    """Every (receiver, band) of the scene in memory (src/pipeline.rs:443-485).
    All bands' chunk c run in one launch."""
    chunk_samples = max(int(chunk_samples), 1)
    own = backend is None
    be = backend or Backend()
    try:
        sc = prepare(scene, TimeBase.SCENE)
        be.upload(sc, rng_seed_base)
        units = list(_units(sc))
        outs = {w: np.empty(sc.receivers[r].bands[b].n, dtype=np.complex64) for w, r, b in units}
        n_chunks = max([math.ceil(sc.receivers[r].bands[b].n / chunk_samples) for _, r, b in units] + [0])
        for c in range(n_chunks):
            jobs, keep = [], []
            for w, r, b in units:
                n = sc.receivers[r].bands[b].n
                i0 = c * chunk_samples
                if i0 >= n:
                    continue
                clen = min(chunk_samples, n - i0)
                j, k = be.job(r, b, w, rng_seed_base, i0, clen, out_host=outs[w][i0:i0 + clen])
                jobs.append(j)
                keep.append(k)
            be.render(jobs, form_code(form))
        be.sync()
        result = []
        for w, r, b in units:
            rx = sc.receivers[r]
            pb = rx.bands[b]
            desc = band_descriptor(scene, rx.name, pb.band, list(pb.active), pb.noise_power_w)
            result.append(ObserverOutput(f"{rx.name}__{pb.band.band_id}", sc.meta.epoch_unix,
                                         pb.band.sample_rate_hz, pb.band.lo_carrier_hz, outs[w], desc))
        return result
    finally:
        if own:
            be.close()


# ---------------------------------------------------------------- render_window
@dataclass
class WindowOptions:
    """src/pipeline.rs:623-643"""
    seed: int = 0
    chunk_samples: int = DEFAULT_WINDOW_CHUNK_SAMPLES
    n_threads: int = 1   # accepted for interface parity; the GPU render ignores it


@dataclass
class WindowSummary:
    """src/pipeline.rs:646-669"""
    n_samples_window: int
    n_samples_emitted: int
    stopped_early: bool
    contributing_transmitters: List[str]
    visible_transmitters: List[str]


BREAK = "break"   # a sink returns BREAK to stop the render (ControlFlow::Break)


def render_window(sky: SkyConfig, receiver: ReceiverConfig, band: ReceiverBandConfig, start_unix: float,
                  duration_s: float, options: WindowOptions, sink: Callable[[np.ndarray], object],
                  form="plain", backend: Optional[Backend] = None) -> WindowSummary:
    """One receiver, one band, one absolute-time window (src/pipeline.rs:709-794).
    ``sink`` is called on the calling thread with each chunk in order."""
    if not math.isfinite(start_unix):
        raise ValueError(f"window start_unix must be finite, got {start_unix}")
    if not (duration_s >= 0.0) or not math.isfinite(duration_s):
        raise ValueError(f"window duration_s must be finite and non-negative, got {duration_s}")
    if not (band.sample_rate_hz > 0.0) or not math.isfinite(band.sample_rate_hz):
        raise ValueError(f"band sample_rate_hz must be positive, got {band.sample_rate_hz}")
    scene = SceneConfig(SceneMeta(start_unix, duration_s, sky.gps_utc_offset_s), list(sky.transmitters),
                        [ReceiverConfig(receiver.name, receiver.trajectory, receiver.clock, [band])])
    chunk = max(int(options.chunk_samples), 1)
    own = backend is None
    be = backend or Backend()
    try:
        sc = prepare(scene, TimeBase.ABSOLUTE)
        be.upload(sc, options.seed)
        pb = sc.receivers[0].bands[0]
        n = pb.n
        i0, stopped = 0, False
        buf = np.empty(min(chunk, max(n, 1)), dtype=np.complex64)
        # per contributing transmitter: unoccluded for some emitted sample
        # (src/pipeline.rs:972-974, the summary's visible_transmitters)
        seen = np.zeros(len(pb.active), dtype=np.uint8)
        flags = np.zeros(max(len(pb.active), 1), dtype=np.uint8)
        while i0 < n:
            clen = min(chunk, n - i0)
            flags[:] = 0
            j, keep = be.job(0, 0, 0, options.seed, i0, clen, out_host=buf[:clen], unoccluded=flags)
            be.render([j], form_code(form))
            be.sync()
            seen |= flags[:len(pb.active)]
            flow = sink(buf[:clen].copy())
            i0 += clen
            if flow == BREAK and i0 < n:
                stopped = True
                break
        names = [sc.transmitters[i].name for i in pb.active]
        visible = [nm for nm, f in zip(names, seen) if f]
        return WindowSummary(rust_round(band.sample_rate_hz * duration_s), i0, stopped, names, visible)
    finally:
        if own:
            be.close()


# The writers (src/output/mod.rs, src/pipeline.rs:494-611), re-exported here
# beside run_scene.
from .output import (OutputManifest, band_descriptor, export_truth, run_scene_to_dir,  # noqa: E402
                     write_raw_with_sidecar)
