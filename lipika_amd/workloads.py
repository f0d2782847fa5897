"""Synthetic scenes for the five benchmark configurations (the reference has no
scene generator; geometry and payload draws here are this project's own,
seeded). Values the configuration text leaves open are marked ASSUMED.

    scene(1)  single CW LEO transmitter, one 48 kHz band, 10 s
    scene(2)  BPSK LEO + six C/A MEO, dual band 48 kHz + 2.048 MHz, 60 s
    scene(3)  32 C/A MEO x 16 receivers, 2.048 MHz, 120 s
    scene(4)  40 mixed transmitters x 64 receivers with 1-3 uneven bands, 30 s
    scene(5)  32 C/A MEO x 256 receivers, 2.048 MHz, 60 s
    micro()   the reference's criterion micro-benchmark chunk: one C/A SV, 5 s at 1.5 MHz
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from .config import (
    Bpsk, CaCode, CircularKeplerian, Cw, NodeClockSpec, Ook, Qam16, ReceiverBandConfig,
    ReceiverConfig, SceneConfig, SceneMeta, StationaryGeodetic, TransmitterConfig,
)

EPOCH = 1717260840.0
SEED = 2026
L1 = 1_575_420_000.0
UHF = 437_000_000.0
ISM = 2_400_000_000.0
GLASGOW = StationaryGeodetic(55.86, -4.25, 10.0)
LEO_A = 6_778_137.0
MEO_A = 26_578_137.0

# Sync words of examples/all_modulations.toml
SYNC_BPSK = [-1, 1, 1, -1, 1, -1, -1, 1, 1, 1, 1, -1, -1, 1, -1, 1,
             -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, -1, 1]
SYNC_OOK = [1, 0, 1, 0, 1, 0, 1, 0, 1, 1, 0, 0, 1, 0, 0, 1]
SYNC_QAM = [0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0]


def _kepler(a, inc, raan, nu):
    return CircularKeplerian(a, inc, raan, nu, EPOCH)


def scene1(duration_s: float = 10.0) -> SceneConfig:
    tx = TransmitterConfig("cw_leo", UHF, Cw(0.0), _kepler(LEO_A, 51.6, 0.0, 10.0), tx_power_dbw=0.0)
    rx = ReceiverConfig("glasgow", GLASGOW, NodeClockSpec(),
                        [ReceiverBandConfig("uhf", UHF, 48_000.0, 290.0)])
    return SceneConfig(SceneMeta(EPOCH, duration_s), [tx], [rx])


def scene2(duration_s: float = 60.0) -> SceneConfig:
    txs = [TransmitterConfig("leo_bpsk", UHF,
                             Bpsk(9600.0, 287_968, SEED, [-1, 1, 1, -1, 1, -1, -1, 1]),
                             _kepler(LEO_A, 51.6, 0.0, 10.0), tx_power_dbw=0.0)]  # RAAN 0, nu 10: ASSUMED
    for k in range(6):
        txs.append(TransmitterConfig(f"gps_prn{k + 1:02d}", L1, CaCode(k + 1, EPOCH),
                                     _kepler(MEO_A, 55.0, 60.0 * k, 30.0 + 50.0 * k), tx_power_dbw=27.0))
    rx = ReceiverConfig("glasgow", GLASGOW, NodeClockSpec(allan_deviation_1s=1e-11, frequency_offset_y0=0.5e-6),
                        [ReceiverBandConfig("uhf_obs", UHF, 48_000.0, 290.0),
                         ReceiverBandConfig("gnss_l1", L1, 2_048_000.0, 290.0)])
    return SceneConfig(SceneMeta(EPOCH, duration_s), txs, [rx])


def _meo_constellation(rng: np.random.Generator):
    planes = [6, 6, 5, 5, 5, 5]
    txs, prn = [], 1
    for p, n in enumerate(planes):
        for _ in range(n):
            nu = float(rng.uniform(0.0, 360.0))
            txs.append(TransmitterConfig(f"gps_prn{prn:02d}", L1, CaCode(prn, EPOCH),
                                         _kepler(MEO_A, 55.0, 60.0 * p, nu), tx_power_dbw=27.0))
            prn += 1
    return txs


def _random_receiver(rng: np.random.Generator, name: str, bands) -> ReceiverConfig:
    lat = float(rng.uniform(-90.0, 90.0))
    lon = float(rng.uniform(-180.0, 180.0))
    h = float(rng.uniform(0.0, 2000.0))
    adev = float(10.0 ** rng.uniform(-12.0, -9.0))
    y0 = float(rng.uniform(-1e-6, 1e-6))
    return ReceiverConfig(name, StationaryGeodetic(lat, lon, h),
                          NodeClockSpec(allan_deviation_1s=adev, frequency_offset_y0=y0), bands)


def scene3(duration_s: float = 120.0, n_rx: int = 16) -> SceneConfig:
    rng = np.random.default_rng(SEED)
    txs = _meo_constellation(rng)
    rxs = [_random_receiver(rng, f"rx{r:03d}", [ReceiverBandConfig("l1", L1, 2_048_000.0, 290.0)])
           for r in range(n_rx)]
    return SceneConfig(SceneMeta(EPOCH, duration_s), txs, rxs)


def scene5(duration_s: float = 60.0, n_rx: int = 256) -> SceneConfig:
    return scene3(duration_s, n_rx)


def scene4(duration_s: float = 30.0, n_rx: int = 64) -> SceneConfig:
    rng = np.random.default_rng(SEED + 4)
    carriers = [UHF, L1, ISM]
    txs = []
    kinds = ["cw", "ook", "bpsk", "qam16", "ca"]
    prn = 1
    for k in range(40):
        kind = kinds[k % 5]
        fc = carriers[int(rng.integers(0, 3))]  # ASSUMED: uniform over the three carriers
        if rng.uniform() < 0.5:
            traj = _kepler(float(rng.uniform(6.9e6, 7.5e6)), float(rng.uniform(0, 180)),
                           float(rng.uniform(0, 360)), float(rng.uniform(0, 360)))
        else:
            traj = _kepler(2.66e7, float(rng.uniform(0, 180)), float(rng.uniform(0, 360)), float(rng.uniform(0, 360)))
        seed = int(rng.integers(0, 2**63))
        if kind == "cw":
            sig = Cw(0.0)  # ASSUMED baseband offset 0
        elif kind == "ook":
            sig = Ook(1000.0, int(1000 * duration_s) - len(SYNC_OOK), seed, SYNC_OOK)
        elif kind == "bpsk":
            sig = Bpsk(9600.0, int(9600 * duration_s) - len(SYNC_BPSK), seed, SYNC_BPSK)
        elif kind == "qam16":
            sig = Qam16(24_000.0, int(24_000 * duration_s) - len(SYNC_QAM) // 4, seed, SYNC_QAM)
        else:
            sig = CaCode(prn, EPOCH)
            prn += 1
        txs.append(TransmitterConfig(f"tx{k:02d}_{kind}", fc, sig, traj, tx_power_dbw=10.0))  # ASSUMED EIRP
    rates = [48_000.0, 250_000.0, 2_048_000.0, 8_184_000.0]
    rxs = []
    for r in range(n_rx):
        nb = int(rng.integers(1, 4))
        bands = [ReceiverBandConfig(f"b{b}", carriers[int(rng.integers(0, 3))], rates[int(rng.integers(0, 4))], 290.0)
                 for b in range(nb)]
        rxs.append(_random_receiver(rng, f"rx{r:03d}", bands))
    return SceneConfig(SceneMeta(EPOCH, duration_s), txs, rxs)


def micro(duration_s: float = 5.0) -> SceneConfig:
    """The chunk of the reference's criterion micro-benchmark
    (benches/iq_generation.rs:15-60): C/A PRN 1 at 1575.42 MHz and 20 W from
    a circular orbit at GPS altitude (a = 6 378 137 + 20 200 000 m, i 55 deg,
    node 0, true anomaly 0 at the epoch), seen from Glasgow with a zero clock
    offset; one band at the carrier, 1.5 MHz, no noise; 5 s (7.5 M samples)."""
    tx = TransmitterConfig("sv01", L1, CaCode(1, EPOCH), _kepler(6_378_137.0 + 20_200_000.0, 55.0, 0.0, 0.0),
                           tx_power_dbw=10.0 * math.log10(20.0))
    rx = ReceiverConfig("glasgow", GLASGOW, NodeClockSpec(),
                        [ReceiverBandConfig("l1", L1, 1_500_000.0, noise_power_w=0.0)])
    return SceneConfig(SceneMeta(EPOCH, duration_s), [tx], [rx])


def scene3_orbiting(duration_s: float = 10.0, n_rx: int = 16) -> SceneConfig:
    """Configuration 3's sky and bands with every receiver on a circular LEO
    orbit (radius 6 900 to 7 200 km, inclination, node and anomaly uniform),
    its elements' epoch the scene's: the moving-receiver workload."""
    sc = scene3(duration_s, n_rx)
    rng = np.random.default_rng(SEED + 33)
    sc.receivers = [ReceiverConfig(rx.name, _kepler(float(rng.uniform(6.9e6, 7.2e6)), float(rng.uniform(0, 180)),
                                                    float(rng.uniform(0, 360)), float(rng.uniform(0, 360))),
                                   rx.clock, rx.bands) for rx in sc.receivers]
    return sc


def scene3_wide(duration_s: float = 10.0, n_rx: int = 16) -> SceneConfig:
    """Configuration 3's receivers and bands under a sky of 64 C/A MEOs: its
    32 and 32 more (PRNs 1-32 again, planes turned by 30 deg, names of their
    own), so that every band holds 64 transmitters: the precision form's
    two-pass workload."""
    sc = scene3(duration_s, n_rx)
    rng = np.random.default_rng(SEED + 64)
    extra = _meo_constellation(rng)
    for t in extra:
        t.name = t.name.replace("gps_prn", "gps_b_prn")
        t.trajectory = _kepler(MEO_A, 55.0, t.trajectory.raan_deg + 30.0, float(rng.uniform(0.0, 360.0)))
    sc.transmitters = sc.transmitters + extra
    return sc


def scene(n: int, duration_s: Optional[float] = None, **kw) -> SceneConfig:
    f = {1: scene1, 2: scene2, 3: scene3, 4: scene4, 5: scene5}[n]
    return f(**kw) if duration_s is None else f(duration_s, **kw)


def totals(prepared) -> tuple:
    """(output samples, contributions) of a prepared scene: contributions are
    the sum over (receiver band, in-band transmitter) of samples."""
    out = contrib = 0
    for r in prepared.receivers:
        for b in r.bands:
            out += b.n
            contrib += b.n * len(b.active)
    return out, contrib
