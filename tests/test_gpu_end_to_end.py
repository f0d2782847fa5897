"""Whole-chain tests of the reference (TOML -> signal -> channel -> IQ) on the
GPU path: tests/ook_qam16_end_to_end.rs and tests/example_all_modulations.rs.
Each scene is a SceneConfig read from TOML and rendered by run_scene."""
import dataclasses
import os

import numpy as np
import pytest

from lipika_amd.config import SceneConfig
from lipika_amd.pipeline import run_scene

pytestmark = pytest.mark.gpu

TOML_TEMPLATE = """
[scene]
epoch_unix = 1717260840.0
duration_s = 0.05

[[transmitter]]
name = "src"
carrier_hz = 1_000_000_000.0
tx_power_dbw = 30.0
trajectory = { kind = "stationary_ecef", x_m = 6378137.0, y_m = 0.0, z_m = 0.0 }
@SIGNAL@

[[receiver]]
name = "lab"
trajectory = { kind = "stationary_ecef", x_m = 6378137.0, y_m = 0.0, z_m = 100000.0 }
clock = { allan_deviation_1s = 0.0 }

[[receiver.band]]
band_id = "test"
lo_carrier_hz = 1_000_000_000.0
sample_rate_hz = 1_000_000.0
noise_power_w = 1.0e-30
"""
CW = 'kind = "cw"\nbaseband_offset_hz = 0.0'
OOK = ('kind = "ook"\nbit_rate_bps = 100_000.0\npayload_bits = 6000\npayload_seed = 7\n'
       'sync_word = [1, 0, 1, 1, 0, 0, 1, 0]')
QAM16 = ('kind = "qam16"\nsymbol_rate_baud = 100_000.0\npayload_symbols = 6000\npayload_seed = 7\n'
         'sync_word = [0, 0, 0, 0, 1, 0, 1, 0]')


def envelope_power(backend, signal):
    """Per-sample envelope power once the signal has arrived (the first ~334
    samples precede the 100 km light time)."""
    scene = SceneConfig.from_toml_str(TOML_TEMPLATE.replace("@SIGNAL@", signal))
    iq = np.asarray(run_scene(scene, 1, backend=backend)[0].iq)[1000:].astype(np.complex128)
    return np.abs(iq) ** 2


def test_ook_keys_the_carrier_between_off_and_cw_power(backend):
    """ook_qam16_end_to_end.rs:62-79: every OOK sample is at CW power (1e-3)
    or off (below 1e-6 of it), on about half the time."""
    cw = float(np.mean(envelope_power(backend, CW)))
    r = envelope_power(backend, OOK) / cw
    on = np.abs(r - 1.0) < 1e-3
    assert np.all(on | (r < 1e-6))
    assert abs(float(on.mean()) - 0.5) < 0.05


def test_qam16_envelope_sits_on_the_three_rings_at_unit_mean_power(backend):
    """ook_qam16_end_to_end.rs:81-101: the envelope takes the rings 0.2, 1.0
    and 1.8 of CW power (to 1e-3) in proportions 1/4, 1/2, 1/4 (to 0.05), and
    its mean is the CW power (to 5 %)."""
    cw = float(np.mean(envelope_power(backend, CW)))
    r = envelope_power(backend, QAM16) / cw
    rings = np.array([0.2, 1.0, 1.8])
    k = np.argmin(np.abs(r[:, None] - rings[None, :]), axis=1)
    assert np.all(np.abs(r - rings[k]) < 1e-3)
    for ring, expect in enumerate((0.25, 0.5, 0.25)):
        assert abs(float(np.mean(k == ring)) - expect) < 0.05
    assert abs(float(np.mean(r)) - 1.0) < 0.05


def test_every_signal_kind_is_present_and_in_band(backend):
    """example_all_modulations.rs:9-33: examples/all_modulations.toml holds one
    transmitter of every kind; each alone for 10 ms over a negligible noise
    floor radiates into the band."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "all_modulations.toml")
    with open(path, encoding="utf-8") as f:
        scene = SceneConfig.from_toml_str(f.read())
    assert sorted(t.kind.kind for t in scene.transmitters) == ["bpsk", "ca_code", "cw", "ook", "qam16"]
    scene.scene.duration_s = 0.01
    scene.receivers[0].bands[0].noise_power_w = 1.0e-30
    for tx in scene.transmitters:
        solo = dataclasses.replace(scene, transmitters=[tx])
        iq = np.asarray(run_scene(solo, 1, backend=backend)[0].iq).astype(np.complex128)
        p = float(np.mean(np.abs(iq) ** 2))
        print(f"{tx.name}: {p:.3e} W in band", flush=True)
        assert p > 1e-15, tx.name
