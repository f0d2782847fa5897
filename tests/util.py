"""Helpers shared by the tests: oracle transmitter structs for single-signal
checks, and small scenes restating the reference's inline test scenes."""
import ctypes as C

import numpy as np

from oracle import pyoracle
from lipika_amd.config import (
    Bpsk, CaCode, Cw, NodeClockSpec, Ook, Qam16, ReceiverBandConfig, ReceiverConfig, SceneConfig,
    SceneMeta, StationaryEcef, StationaryGeodetic, TransmitterConfig,
)

EPOCH = 1717260840.0
keep = []


def arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    keep.append(a)
    return a


def orc_tx(kind, **f):
    """An oracle transmitter struct for a single-signal check."""
    t = pyoracle.OrcTx()
    t.signal_kind = kind
    t.traj_kind = 0
    t.carrier_hz = f.get("carrier_hz", 1e9)
    for k in ("baseband_offset_hz", "anchor_phase_rad", "rate_hz", "epoch_rel_s", "gate_period_s", "gate_phase_s"):
        if k in f:
            setattr(t, k, f[k])
    for k in ("cw_anchored", "repeating", "repeat_from", "nav_mode", "prn", "nav_seed", "gated"):
        if k in f:
            setattr(t, k, f[k])
    if "bits" in f:
        b = arr(f["bits"], np.int8)
        t.bits, t.n_symbols = b.ctypes.data_as(C.POINTER(C.c_int8)), len(b)
    if "symbols" in f:
        s = arr(f["symbols"], np.float32)
        t.symbols, t.n_symbols = s.ctypes.data_as(C.POINTER(C.c_float)), len(s)
    if "ca_code" in f:
        c = arr(f["ca_code"], np.int8)
        t.ca_code = c.ctypes.data_as(C.POINTER(C.c_int8))
    if "bursts" in f:
        b = arr(f["bursts"], np.float64)
        t.bursts, t.n_bursts = b.ctypes.data_as(C.POINTER(C.c_double)), len(b)
    return t


def baseband(orc, tx, times):
    t = arr(times, np.float64)
    out = np.empty(2 * len(t), dtype=np.float32)
    orc.orc_baseband(C.byref(tx), t.ctypes.data_as(C.POINTER(C.c_double)), len(t),
                     out.ctypes.data_as(C.POINTER(C.c_float)))
    return out[0::2] + 1j * out[1::2]


def lab_receiver(bands, clock=None, pos=StationaryGeodetic(55.86, -4.25, 10.0), name="lab"):
    return ReceiverConfig(name, pos, clock or NodeClockSpec(), bands)


def two_tx_scene(which=("alpha", "beta"), duration_s=0.5):
    """tests/multi_transmitter_summation.rs:30-59"""
    txs = []
    if "alpha" in which:
        txs.append(TransmitterConfig("alpha", 1e9, Cw(0.0), StationaryEcef(6378137.0, 0.0, 1000.0), 30.0))
    if "beta" in which:
        txs.append(TransmitterConfig("beta", 1.005e9, Cw(0.0), StationaryEcef(6378137.0, 1000.0, 1000.0), 30.0))
    band = ReceiverBandConfig("test", 1e9, 1e6, noise_power_w=1e-15)
    return SceneConfig(SceneMeta(EPOCH, duration_s), txs, [lab_receiver([band])])


def cw_two_clocks_scene():
    """tests/multi_transmitter_summation.rs:101-132"""
    tx = TransmitterConfig("cw", 1e6, Cw(0.0), StationaryEcef(6478137.0, 0.0, 0.0), 0.0)
    rxs = []
    for name, off in (("rxa", 0.0), ("rxb", 1.3e-7)):
        rxs.append(ReceiverConfig(name, StationaryEcef(6378137.0, 0.0, 0.0), NodeClockSpec(initial_time_offset_s=off),
                                  [ReceiverBandConfig("b", 1e6, 100000.0, noise_power_w=1e-30)]))
    return SceneConfig(SceneMeta(EPOCH, 0.02), [tx], rxs)


def all_kinds_scene(duration_s=0.02, noise_power_w=0.0, clock=None, frontend=None):
    """Every signal kind in one 8 MHz band at 915 MHz, stationary
    (examples/all_modulations.toml, restated with a shorter duration)."""
    txs = [
        TransmitterConfig("tone", 912e6, Cw(0.0), StationaryGeodetic(55.90, -4.25, 30.0), -10.0),
        TransmitterConfig("keyfob_ook", 913e6, Ook(50000.0, 49984, 1, [1, 0, 1, 0, 1, 0, 1, 0, 1, 1, 0, 0, 1, 0, 0, 1]),
                          StationaryGeodetic(55.86, -4.18, 30.0), -10.0),
        TransmitterConfig("beacon_bpsk", 914e6,
                          Bpsk(100000.0, 99968, 2, [-1, 1, 1, -1, 1, -1, -1, 1, 1, 1, 1, -1, -1, 1, -1, 1, -1, 1,
                                                    -1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, -1, 1]),
                          StationaryGeodetic(55.82, -4.25, 30.0), -10.0),
        TransmitterConfig("link_qam16", 915e6, Qam16(100000.0, 99996, 3, [0, 0, 0, 0, 0, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0]),
                          StationaryGeodetic(55.86, -4.32, 30.0), -10.0),
        TransmitterConfig("pseudolite_prn07", 917.5e6, CaCode(7, EPOCH, "dummy"), StationaryGeodetic(55.89, -4.20, 30.0), -10.0),
    ]
    band = ReceiverBandConfig("ism_915", 915e6, 8e6, 290.0, noise_power_w=noise_power_w, frontend=frontend)
    return SceneConfig(SceneMeta(EPOCH, duration_s), txs,
                       [lab_receiver([band], clock or NodeClockSpec(allan_deviation_1s=1e-11))])


def pyoracle_clock(orc, rx):
    """An initialised oracle ClockStream for receiver struct `rx`."""
    c = pyoracle.OrcClock()
    keep.append(rx)
    orc.orc_clock_init(C.byref(c), C.byref(rx))
    return c
