"""Scene description mirroring the reference's serde schema
(src/config/mod.rs:52-467): the same names, defaults and TOML layout, for the
fields the IQ path consumes.

Parts of the schema the GPU path does not implement raise ``NotImplementedError``
when a scene uses them (TLE and RINEX trajectories, full LNAV, moving
receivers); nothing is silently dropped.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple, Union

try:  # Python 3.10 has no tomllib
    import tomli as _toml
except ImportError:  # pragma: no cover
    import tomllib as _toml  # type: ignore

CA_CHIP_RATE_HZ = 1_023_000.0  # src/signal/ca_code.rs:23


# ---------------------------------------------------------------- trajectories
@dataclass
class StationaryGeodetic:
    lat_deg: float
    lon_deg: float
    h_m: float
    kind: str = "stationary_geodetic"


@dataclass
class StationaryEcef:
    x_m: float
    y_m: float
    z_m: float
    kind: str = "stationary_ecef"


@dataclass
class CircularKeplerian:
    semi_major_axis_m: float
    inclination_deg: float
    raan_deg: float
    true_anomaly_deg_at_epoch: float
    epoch_unix: float
    kind: str = "circular_keplerian"


@dataclass
class RinexNav:
    """GPS broadcast ephemeris from a RINEX 3 nav file (src/config/mod.rs:440-444):
    the record of `prn` whose toe is nearest the scene epoch."""
    file: str
    prn: int
    kind: str = "rinex_nav"


@dataclass
class Tle:
    """An SGP4 trajectory from a TLE pair (src/config/mod.rs:432), propagated
    on the host (include/lpk.h, lpk_tle_*) into the transmitter's grid."""
    line1: str
    line2: str
    kind: str = "tle"


Trajectory = Union[StationaryGeodetic, StationaryEcef, CircularKeplerian, RinexNav, Tle]


# ---------------------------------------------------------------- signals
@dataclass
class Cw:
    baseband_offset_hz: float
    kind: str = "cw"


@dataclass
class Bpsk:
    bit_rate_bps: float
    payload_bits: int
    payload_seed: int
    sync_word: List[int]
    kind: str = "bpsk"


@dataclass
class Ook:
    bit_rate_bps: float
    payload_bits: int
    payload_seed: int
    sync_word: List[int]
    kind: str = "ook"


@dataclass
class Qam16:
    symbol_rate_baud: float
    payload_symbols: int
    payload_seed: int
    sync_word: List[int]
    kind: str = "qam16"


@dataclass
class CaCode:
    prn: int
    epoch_unix: float
    nav: str = "dummy"  # src/config/mod.rs:225-237 (default dummy)
    nav_seed: int = 0
    kind: str = "ca_code"


Signal = Union[Cw, Bpsk, Ook, Qam16, CaCode]


def default_bandwidth_hz(sig: Signal) -> float:
    """src/config/mod.rs:260-268"""
    if isinstance(sig, Cw):
        return 0.0
    if isinstance(sig, (Bpsk, Ook)):
        return 2.0 * sig.bit_rate_bps
    if isinstance(sig, Qam16):
        return 2.0 * sig.symbol_rate_baud
    return 2.0 * CA_CHIP_RATE_HZ


@dataclass
class GateConfig:
    period_s: float
    bursts: List[Tuple[float, float]]  # (offset_s, duration_s)
    epoch_unix: float = 0.0


@dataclass
class TransmitterConfig:
    name: str
    carrier_hz: float
    kind: Signal
    trajectory: Trajectory
    tx_power_dbw: float = 0.0
    bandwidth_hz: Optional[float] = None
    signal_epoch_unix: Optional[float] = None
    gate: Optional[GateConfig] = None

    def effective_bandwidth_hz(self) -> float:
        """src/config/mod.rs:191-193"""
        return self.bandwidth_hz if self.bandwidth_hz is not None else default_bandwidth_hz(self.kind)


# ---------------------------------------------------------------- receivers
@dataclass
class NodeClockSpec:
    """src/clock/mod.rs:20-46"""
    allan_deviation_1s: float = 0.0
    flicker_floor: float = 0.0
    frequency_offset_y0: float = 0.0
    initial_time_offset_s: float = 0.0


@dataclass
class AdcConfig:
    bits: int
    full_scale_v: float


@dataclass
class NonlinearityConfig:
    iip3_dbm: float


@dataclass
class FrontendConfig:
    antenna: Optional[dict] = None
    gain_db: Optional[float] = None
    nonlinearity: Optional[NonlinearityConfig] = None
    adc: Optional[AdcConfig] = None


@dataclass
class ReceiverBandConfig:
    band_id: str
    lo_carrier_hz: float
    sample_rate_hz: float
    noise_temperature_k: float = 290.0  # src/config/mod.rs:425
    noise_power_w: Optional[float] = None
    bandwidth_hz: Optional[float] = None
    frontend: Optional[FrontendConfig] = None

    def effective_bandwidth_hz(self) -> float:
        """src/config/mod.rs:420-422"""
        return self.bandwidth_hz if self.bandwidth_hz is not None else self.sample_rate_hz


@dataclass
class ReceiverConfig:
    name: str
    trajectory: Trajectory
    clock: NodeClockSpec = field(default_factory=NodeClockSpec)
    bands: List[ReceiverBandConfig] = field(default_factory=list)


@dataclass
class SceneMeta:
    epoch_unix: float
    duration_s: float
    gps_utc_offset_s: float = 18.0  # src/config/mod.rs:71


@dataclass
class SceneConfig:
    scene: SceneMeta
    transmitters: List[TransmitterConfig] = field(default_factory=list)
    receivers: List[ReceiverConfig] = field(default_factory=list)

    @staticmethod
    def from_toml_str(text: str) -> "SceneConfig":
        return _scene_from_dict(_toml.loads(text))


@dataclass
class SkyConfig:
    """src/config/mod.rs:150-185: transmitters without a scene or receivers."""
    transmitters: List[TransmitterConfig] = field(default_factory=list)
    gps_utc_offset_s: float = 18.0

    @staticmethod
    def from_toml_str(text: str) -> "SkyConfig":
        d = _toml.loads(text)
        unknown = set(d) - {"transmitter", "gps_utc_offset_s"}
        if unknown:   # serde(deny_unknown_fields)
            raise ValueError(f"unknown sky field(s): {sorted(unknown)}")
        return SkyConfig([_transmitter(t) for t in d.get("transmitter", [])],
                         float(d.get("gps_utc_offset_s", 18.0)))

    @staticmethod
    def from_toml_file(path: str) -> "SkyConfig":
        with open(path, "r", encoding="utf-8") as f:
            text = f.read()
        try:
            return SkyConfig.from_toml_str(text)
        except (ValueError, KeyError) as e:
            raise ValueError(f"parsing sky '{path}': {e}") from e


# ---------------------------------------------------------------- TOML
def _trajectory(d: dict) -> Trajectory:
    kind = d["kind"]
    if kind == "stationary_geodetic":
        return StationaryGeodetic(float(d["lat_deg"]), float(d["lon_deg"]), float(d["h_m"]))
    if kind == "stationary_ecef":
        return StationaryEcef(float(d["x_m"]), float(d["y_m"]), float(d["z_m"]))
    if kind == "circular_keplerian":
        return CircularKeplerian(float(d["semi_major_axis_m"]), float(d["inclination_deg"]),
                                 float(d["raan_deg"]), float(d["true_anomaly_deg_at_epoch"]),
                                 float(d["epoch_unix"]))
    if kind == "rinex_nav":
        return RinexNav(str(d["file"]), int(d["prn"]))
    if kind == "tle":
        return Tle(str(d["line1"]), str(d["line2"]))
    raise ValueError(f"unknown trajectory kind {kind!r}")


def _signal(d: dict) -> Signal:
    kind = d["kind"]
    if kind == "cw":
        return Cw(float(d["baseband_offset_hz"]))
    if kind == "bpsk":
        return Bpsk(float(d["bit_rate_bps"]), int(d["payload_bits"]), int(d["payload_seed"]),
                    [int(x) for x in d["sync_word"]])
    if kind == "ook":
        return Ook(float(d["bit_rate_bps"]), int(d["payload_bits"]), int(d["payload_seed"]),
                   [int(x) for x in d["sync_word"]])
    if kind == "qam16":
        return Qam16(float(d["symbol_rate_baud"]), int(d["payload_symbols"]), int(d["payload_seed"]),
                     [int(x) for x in d["sync_word"]])
    if kind == "ca_code":
        return CaCode(int(d["prn"]), float(d["epoch_unix"]), d.get("nav", "dummy"), int(d.get("nav_seed", 0)))
    raise ValueError(f"unknown signal kind {kind!r}")


def _transmitter(d: dict) -> TransmitterConfig:
    gate = None
    if "gate" in d:
        g = d["gate"]
        gate = GateConfig(float(g["period_s"]),
                          [(float(b["offset_s"]), float(b["duration_s"])) for b in g["bursts"]],
                          float(g.get("epoch_unix", 0.0)))
    return TransmitterConfig(
        name=d["name"], carrier_hz=float(d["carrier_hz"]), kind=_signal(d),
        trajectory=_trajectory(d["trajectory"]), tx_power_dbw=float(d.get("tx_power_dbw", 0.0)),
        bandwidth_hz=d.get("bandwidth_hz"), signal_epoch_unix=d.get("signal_epoch_unix"), gate=gate)


def _band(d: dict) -> ReceiverBandConfig:
    fe = None
    if "frontend" in d:
        f = d["frontend"]
        fe = FrontendConfig(
            antenna=f.get("antenna"), gain_db=f.get("gain_db"),
            nonlinearity=NonlinearityConfig(float(f["nonlinearity"]["iip3_dbm"])) if "nonlinearity" in f else None,
            adc=AdcConfig(int(f["adc"]["bits"]), float(f["adc"]["full_scale_v"])) if "adc" in f else None)
    return ReceiverBandConfig(d["band_id"], float(d["lo_carrier_hz"]), float(d["sample_rate_hz"]),
                              float(d.get("noise_temperature_k", 290.0)), d.get("noise_power_w"),
                              d.get("bandwidth_hz"), fe)


def _scene_from_dict(d: dict) -> SceneConfig:
    s = d["scene"]
    meta = SceneMeta(float(s["epoch_unix"]), float(s["duration_s"]), float(s.get("gps_utc_offset_s", 18.0)))
    txs = [_transmitter(t) for t in d.get("transmitter", [])]
    rxs = []
    for r in d.get("receiver", []):
        c = r.get("clock", {})
        clock = NodeClockSpec(float(c.get("allan_deviation_1s", 0.0)), float(c.get("flicker_floor", 0.0)),
                              float(c.get("frequency_offset_y0", 0.0)), float(c.get("initial_time_offset_s", 0.0)))
        rxs.append(ReceiverConfig(r["name"], _trajectory(r["trajectory"]), clock, [_band(b) for b in r.get("band", [])]))
    return SceneConfig(meta, txs, rxs)
