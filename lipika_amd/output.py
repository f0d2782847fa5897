"""Output of the IQ path: raw complex64 files with JSON sidecars
(src/output/mod.rs:1-139), the streamed scene writer run_scene_to_dir
(src/pipeline.rs:494-536) and the geometric truth export export_truth
(src/pipeline.rs:538-611).

The sidecar and truth files are written as serde_json writes them: object
keys sorted (serde_json's map without `preserve_order`, Cargo.toml:38), two
spaces of indentation for the pretty sidecar and none for the truth file, and
every f64 in the shortest round-trip form with ryu's layout (`437000000.0`,
`1.5e-7`, `0.00001`, `1e21`). Integer fields stay integers. The scene
configuration embedded in each sidecar follows the reference's Serialize
derive field for field (src/config/mod.rs:52-460): every field present,
`Option`s as null, except `signal_epoch_unix` and `gate`, which are skipped
when unset.
"""
from __future__ import annotations

import ctypes as C
import decimal
import math
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from .config import (Bpsk, CaCode, CircularKeplerian, Cw, Ook, Qam16, RinexNav, SceneConfig, StationaryEcef,
                     StationaryGeodetic, Tle, TransmitterConfig, default_bandwidth_hz)

SPEED_OF_LIGHT_M_S = 299792458.0           # src/lib.rs:53
TRUTH_ITERATIONS = 4                        # src/pipeline.rs:573 (NEWTON_ITERATIONS.max(4))


# ---------------------------------------------------------------- serde_json text

def format_f64(x: float) -> str:
    """An f64 as serde_json writes it (ryu's shortest round-trip digits and
    layout); non-finite values become null, as serde_json writes them."""
    x = float(x)
    if not math.isfinite(x):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    t = decimal.Decimal(repr(abs(x))).as_tuple()
    digits = "".join(str(d) for d in t.digits)
    exp = t.exponent
    stripped = digits.rstrip("0")
    exp += len(digits) - len(stripped)
    digits = stripped.lstrip("0")
    length, k = len(digits), exp
    kk = length + k                       # 10^(kk-1) <= |x| < 10^kk
    if 0 <= k and kk <= 16:
        return sign + digits + "0" * (kk - length) + ".0"
    if 0 < kk <= 16:
        return sign + digits[:kk] + "." + digits[kk:]
    if -5 < kk <= 0:
        return sign + "0." + "0" * (-kk) + digits
    if length == 1:
        return sign + digits + "e" + str(kk - 1)
    return sign + digits[0] + "." + digits[1:] + "e" + str(kk - 1)


_ESC = {'"': '\\"', "\\": "\\\\", "\b": "\\b", "\f": "\\f", "\n": "\\n", "\r": "\\r", "\t": "\\t"}


def _string(s: str) -> str:
    out = []
    for ch in s:
        if ch in _ESC:
            out.append(_ESC[ch])
        elif ord(ch) < 0x20:
            out.append("\\u%04x" % ord(ch))
        else:
            out.append(ch)
    return '"' + "".join(out) + '"'


def to_json(value, pretty: bool = False, _level: int = 0) -> str:
    """serde_json::to_string / to_string_pretty of a value built from dicts,
    lists, str, int, float, bool and None."""
    if value is None:
        return "null"
    if value is True:
        return "true"
    if value is False:
        return "false"
    if isinstance(value, (int, np.integer)) and not isinstance(value, bool):
        return str(int(value))
    if isinstance(value, (float, np.floating)):
        return format_f64(float(value))
    if isinstance(value, str):
        return _string(value)
    if isinstance(value, dict):
        if not value:
            return "{}"
        items = [(_string(str(k)), v) for k, v in sorted(value.items(), key=lambda kv: str(kv[0]))]
        if not pretty:
            return "{" + ",".join(f"{k}:{to_json(v)}" for k, v in items) + "}"
        ind, ind1 = "  " * _level, "  " * (_level + 1)
        return "{\n" + ",\n".join(f"{ind1}{k}: {to_json(v, True, _level + 1)}" for k, v in items) + "\n" + ind + "}"
    if isinstance(value, (list, tuple, np.ndarray)):
        seq = list(value)
        if not seq:
            return "[]"
        if not pretty:
            return "[" + ",".join(to_json(v) for v in seq) + "]"
        ind, ind1 = "  " * _level, "  " * (_level + 1)
        return "[\n" + ",\n".join(ind1 + to_json(v, True, _level + 1) for v in seq) + "\n" + ind + "]"
    raise TypeError(f"cannot write {type(value).__name__} as JSON")


# ---------------------------------------------------------------- the scene, as serde serialises it

def _f(x) -> Optional[float]:
    return None if x is None else float(x)


def _trajectory_value(t) -> dict:
    if isinstance(t, StationaryGeodetic):
        return {"kind": "stationary_geodetic", "lat_deg": _f(t.lat_deg), "lon_deg": _f(t.lon_deg), "h_m": _f(t.h_m)}
    if isinstance(t, StationaryEcef):
        return {"kind": "stationary_ecef", "x_m": _f(t.x_m), "y_m": _f(t.y_m), "z_m": _f(t.z_m)}
    if isinstance(t, CircularKeplerian):
        return {"kind": "circular_keplerian", "semi_major_axis_m": _f(t.semi_major_axis_m),
                "inclination_deg": _f(t.inclination_deg), "raan_deg": _f(t.raan_deg),
                "true_anomaly_deg_at_epoch": _f(t.true_anomaly_deg_at_epoch), "epoch_unix": _f(t.epoch_unix)}
    if isinstance(t, RinexNav):
        return {"kind": "rinex_nav", "file": t.file, "prn": int(t.prn)}
    if isinstance(t, Tle):
        return {"kind": "tle", "line1": t.line1, "line2": t.line2}
    raise TypeError(f"trajectory {t!r}")


def _signal_value(s) -> dict:
    if isinstance(s, Cw):
        return {"kind": "cw", "baseband_offset_hz": _f(s.baseband_offset_hz)}
    if isinstance(s, (Bpsk, Ook)):
        return {"kind": "bpsk" if isinstance(s, Bpsk) else "ook", "bit_rate_bps": _f(s.bit_rate_bps),
                "payload_bits": int(s.payload_bits), "payload_seed": int(s.payload_seed),
                "sync_word": [int(v) for v in s.sync_word]}
    if isinstance(s, Qam16):
        return {"kind": "qam16", "symbol_rate_baud": _f(s.symbol_rate_baud), "payload_symbols": int(s.payload_symbols),
                "payload_seed": int(s.payload_seed), "sync_word": [int(v) for v in s.sync_word]}
    if isinstance(s, CaCode):
        return {"kind": "ca_code", "prn": int(s.prn), "epoch_unix": _f(s.epoch_unix), "nav": str(s.nav),
                "nav_seed": int(s.nav_seed)}
    raise TypeError(f"signal {s!r}")


def _transmitter_value(t: TransmitterConfig) -> dict:
    v = {"name": t.name, "carrier_hz": _f(t.carrier_hz), "tx_power_dbw": _f(t.tx_power_dbw),
         "bandwidth_hz": _f(t.bandwidth_hz), "trajectory": _trajectory_value(t.trajectory)}
    if t.signal_epoch_unix is not None:
        v["signal_epoch_unix"] = _f(t.signal_epoch_unix)
    if t.gate is not None:
        v["gate"] = {"period_s": _f(t.gate.period_s), "epoch_unix": _f(t.gate.epoch_unix),
                     "bursts": [{"offset_s": _f(o), "duration_s": _f(d)} for o, d in t.gate.bursts]}
    v.update(_signal_value(t.kind))      # #[serde(flatten)] kind
    return v


def _band_value(b) -> dict:
    fe = None
    if b.frontend is not None:
        f = b.frontend
        ant = None
        if f.antenna is not None:
            a = f.antenna
            ant = {"peak_dbi": _f(a["peak_dbi"]), "horizon_dbi": _f(a["horizon_dbi"]), "fb_db": _f(a["fb_db"]),
                   "cutoff_deg": _f(a.get("cutoff_deg"))}
        fe = {"antenna": ant, "gain_db": _f(f.gain_db),
              "nonlinearity": None if f.nonlinearity is None else {"iip3_dbm": _f(f.nonlinearity.iip3_dbm)},
              "adc": None if f.adc is None else {"bits": int(f.adc.bits), "full_scale_v": _f(f.adc.full_scale_v)}}
    return {"band_id": b.band_id, "lo_carrier_hz": _f(b.lo_carrier_hz), "sample_rate_hz": _f(b.sample_rate_hz),
            "noise_temperature_k": _f(b.noise_temperature_k), "noise_power_w": _f(b.noise_power_w),
            "bandwidth_hz": _f(b.bandwidth_hz), "frontend": fe}


def scene_config_value(scene: SceneConfig) -> dict:
    """serde_json::to_value(scene) (src/pipeline.rs:1004)."""
    return {
        "scene": {"epoch_unix": _f(scene.scene.epoch_unix), "duration_s": _f(scene.scene.duration_s),
                  "gps_utc_offset_s": _f(scene.scene.gps_utc_offset_s)},
        "transmitter": [_transmitter_value(t) for t in scene.transmitters],
        "receiver": [{"name": r.name, "trajectory": _trajectory_value(r.trajectory),
                      "clock": {"allan_deviation_1s": _f(r.clock.allan_deviation_1s),
                                "flicker_floor": _f(r.clock.flicker_floor),
                                "frequency_offset_y0": _f(r.clock.frequency_offset_y0),
                                "initial_time_offset_s": _f(r.clock.initial_time_offset_s)},
                      "band": [_band_value(b) for b in r.bands]} for r in scene.receivers],
    }


def transmitter_descriptor(t: TransmitterConfig) -> dict:
    """src/pipeline.rs:1083-1097"""
    sig = _signal_value(t.kind)
    v = {"name": t.name, "kind": sig["kind"], "carrier_hz": _f(t.carrier_hz), "tx_power_dbw": _f(t.tx_power_dbw),
         "bandwidth_hz": float(t.bandwidth_hz if t.bandwidth_hz is not None else default_bandwidth_hz(t.kind)),
         "trajectory_kind": _trajectory_value(t.trajectory)["kind"]}
    if isinstance(t.kind, CaCode):
        v["prn"] = int(t.kind.prn)
        v["nav"] = str(t.kind.nav).lower()
    return v


def band_descriptor(scene: SceneConfig, receiver_name: str, band, active: List[int], noise_power_w: float) -> dict:
    """src/pipeline.rs:992-1005"""
    return {
        "receiver": receiver_name, "band_id": band.band_id, "duration_s": _f(scene.scene.duration_s),
        "band": {"lo_carrier_hz": _f(band.lo_carrier_hz), "sample_rate_hz": _f(band.sample_rate_hz),
                 "bandwidth_hz": float(band.effective_bandwidth_hz()),
                 "noise_temperature_k": _f(band.noise_temperature_k), "noise_power_w": float(noise_power_w)},
        "contributing_transmitters": [transmitter_descriptor(scene.transmitters[i]) for i in active],
        "scene_config": scene_config_value(scene),
    }


def sidecar_value(name: str, epoch_unix: float, sample_rate_hz: float, lo_carrier_hz: float, n_samples: int,
                  descriptor: dict) -> dict:
    """src/output/mod.rs:56-75"""
    return {"name": name, "epoch_unix": float(epoch_unix), "sample_rate_hz": float(sample_rate_hz),
            "lo_carrier_hz": float(lo_carrier_hz), "n_samples": int(n_samples),
            "format": "raw complex64 little-endian (interleaved f32 re, f32 im)",
            "numpy_load": "np.fromfile(path, dtype=np.complex64)", "descriptor": descriptor}


def write_sidecar(directory: str, name: str, epoch_unix: float, sample_rate_hz: float, lo_carrier_hz: float,
                  n_samples: int, descriptor: dict) -> str:
    path = os.path.join(directory, f"{name}.iq.json")
    with open(path, "w", encoding="utf-8") as f:
        f.write(to_json(sidecar_value(name, epoch_unix, sample_rate_hz, lo_carrier_hz, n_samples, descriptor),
                        pretty=True))
    return path


def write_raw_with_sidecar(out, directory: str) -> None:
    """ObserverOutput::write_raw_with_sidecar (src/output/mod.rs:36-50)."""
    os.makedirs(directory, exist_ok=True)
    np.ascontiguousarray(out.iq, dtype="<c8").tofile(os.path.join(directory, f"{out.name}.iq.c64"))
    write_sidecar(directory, out.name, out.epoch_unix, out.sample_rate_hz, out.lo_carrier_hz, len(out.iq),
                  out.descriptor)


# ---------------------------------------------------------------- streamed scene, truth

@dataclass
class OutputManifest:
    """src/pipeline.rs:431-437"""
    name: str
    path: str
    n_samples: int
    descriptor: dict


def run_scene_to_dir(scene: SceneConfig, rng_seed_base: int, directory: str, chunk_samples: int = 1 << 20,
                     form="plain", backend=None) -> List[OutputManifest]:
    """Every (receiver, band) streamed to <dir>/<receiver>__<band>.iq.c64 with
    its sidecar (src/pipeline.rs:494-536), through lpk_stream_to_files: chunk
    k of every band renders while chunk k - 1 is written."""
    from .pipeline import Backend, TimeBase, form_code, prepare
    chunk_samples = max(int(chunk_samples), 1)
    os.makedirs(directory, exist_ok=True)
    own = backend is None
    be = backend or Backend()
    try:
        sc = prepare(scene, TimeBase.SCENE)
        be.upload(sc, rng_seed_base)
        jobs, keep, names, metas = [], [], [], []
        w = 0
        for r, rx in enumerate(sc.receivers):
            for b, pb in enumerate(rx.bands):
                j, k = be.job(r, b, w, rng_seed_base, 0, pb.n)
                jobs.append(j)
                keep.append(k)
                name = f"{rx.name}__{pb.band.band_id}"
                names.append(name)
                metas.append((rx, pb))
                w += 1
        paths = [os.path.join(directory, f"{n}.iq.c64") for n in names]
        arr = (_lib.BandChunk * max(len(jobs), 1))(*jobs)
        parr = (C.c_char_p * max(len(paths), 1))(*[p.encode() for p in paths])
        be.check(be.lib.lpk_stream_to_files(be.h, arr, len(jobs), parr, chunk_samples, form_code(form)))
        out = []
        for name, path, (rx, pb) in zip(names, paths, metas):
            desc = band_descriptor(scene, rx.name, pb.band, list(pb.active), pb.noise_power_w)
            write_sidecar(directory, name, sc.meta.epoch_unix, pb.band.sample_rate_hz, pb.band.lo_carrier_hz, pb.n,
                          desc)
            out.append(OutputManifest(name, path, pb.n, desc))
        return out
    finally:
        if own:
            be.close()


def export_truth(scene: SceneConfig, directory: str, truth_hz: float, backend=None) -> None:
    """src/pipeline.rs:547-611: per receiver, <dir>/<receiver>__truth.json with
    every transmitter's range, delay, range rate and source position on a
    truth_hz grid (solved on the device with 4 iterations)."""
    from .pipeline import Backend, TimeBase, prepare, rust_round
    if truth_hz <= 0.0:
        return
    os.makedirs(directory, exist_ok=True)
    own = backend is None
    be = backend or Backend()
    try:
        sc = prepare(scene, TimeBase.SCENE)
        be.upload(sc, 0)
        epoch, dur = scene.scene.epoch_unix, scene.scene.duration_s
        ng = max(rust_round(dur * truth_hz), 1) + 1
        dt = 1.0 / truth_hz
        t_rel = np.array([k * dt for k in range(ng)], dtype=np.float64)
        t_abs = [epoch + x for x in t_rel]
        n_tx = len(sc.transmitters)
        for r, rx in enumerate(sc.receivers):
            ranges = np.empty((n_tx, ng), dtype=np.float64)
            src = np.empty((n_tx, 3, ng), dtype=np.float64)
            be.check(be.lib.lpk_truth(be.h, r, t_rel.ctypes.data_as(C.POINTER(C.c_double)), ng, epoch,
                                      TRUTH_ITERATIONS, ranges.ctypes.data_as(C.POINTER(C.c_double)),
                                      src.ctypes.data_as(C.POINTER(C.c_double))))
            txs = []
            for i, cfg in enumerate(scene.transmitters):
                rg = ranges[i]
                rr = [0.0] * ng
                if ng >= 2:
                    for k in range(ng):
                        if k == 0:
                            rr[k] = (rg[1] - rg[0]) / dt
                        elif k == ng - 1:
                            rr[k] = (rg[ng - 1] - rg[ng - 2]) / dt
                        else:
                            rr[k] = (rg[k + 1] - rg[k - 1]) / (2.0 * dt)
                d = transmitter_descriptor(cfg)
                d["range_m"] = [float(v) for v in rg]
                d["delay_s"] = [float(v) / SPEED_OF_LIGHT_M_S for v in rg]
                d["range_rate_m_s"] = [float(v) for v in rr]
                d["source_x_m"] = [float(v) for v in src[i, 0]]
                d["source_y_m"] = [float(v) for v in src[i, 1]]
                d["source_z_m"] = [float(v) for v in src[i, 2]]
                txs.append(d)
            truth = {"receiver": rx.name, "epoch_unix": float(epoch), "duration_s": float(dur),
                     "grid_hz": float(truth_hz), "t_utc": t_abs, "transmitters": txs}
            with open(os.path.join(directory, f"{rx.name}__truth.json"), "w", encoding="utf-8") as f:
                f.write(to_json(truth))
    finally:
        if own:
            be.close()
