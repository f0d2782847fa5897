"""The output format (src/output/mod.rs, docs/REQUIREMENTS.md fn-6), on the CPU:
serde_json's float text (ryu's shortest round-trip digits and layout), its
pretty and compact layouts with sorted keys, the sidecar's keys, and the
scene configuration as the reference's Serialize derive lays it out."""
import json
import math
import struct

import numpy as np

from lipika_amd import output, workloads
from lipika_amd.config import (AdcConfig, FrontendConfig, GateConfig, NonlinearityConfig, ReceiverBandConfig,
                               SceneConfig)


def test_f64_text_is_serde_json_text():
    # (value, serde_json::to_string) pairs: integral values keep ".0", the
    # fixed/scientific switch is at 1e16 and 1e-5, exponents have no '+' or
    # padding, and digits are the shortest that round-trip
    cases = [(1.0, "1.0"), (0.1, "0.1"), (-2.5, "-2.5"), (0.0, "0.0"), (-0.0, "-0.0"), (290.0, "290.0"),
             (437e6, "437000000.0"), (2.048e6, "2048000.0"), (1717260840.0, "1717260840.0"),
             (1e15, "1000000000000000.0"), (1e16, "1e16"), (1.5e16, "1.5e16"), (123456789012345680.0, "1.2345678901234568e17"),
             (1e21, "1e21"), (1e-4, "0.0001"), (1e-5, "0.00001"), (1.25e-5, "0.0000125"), (1e-6, "1e-6"),
             (1.5e-7, "1.5e-7"), (1e-11, "1e-11"), (5e-324, "5e-324"), (1.7976931348623157e308, "1.7976931348623157e308"),
             (0.3, "0.3"), (2.0 / 3.0, "0.6666666666666666"), (1e23, "1e23"), (float("nan"), "null")]
    for v, want in cases:
        assert output.format_f64(v) == want, (v, output.format_f64(v), want)


def test_f64_text_round_trips():
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2 ** 63 - 1, 20000, dtype=np.int64)
    for b in bits:
        x = struct.unpack("<d", struct.pack("<q", int(b)))[0]
        if not math.isfinite(x):
            continue
        s = output.format_f64(x)
        assert float(s) == x and "+" not in s
        assert ("e" in s) or ("." in s)      # a float never prints as an integer


def test_layouts_sort_keys_and_indent_like_serde_json():
    v = {"b": [1, 2.0, {"z": None, "a": "q\"\\\n\x01"}], "a": {}, "c": [], "d": True}
    assert output.to_json(v) == '{"a":{},"b":[1,2.0,{"a":"q\\"\\\\\\n\\u0001","z":null}],"c":[],"d":true}'
    assert output.to_json(v, pretty=True) == (
        '{\n  "a": {},\n  "b": [\n    1,\n    2.0,\n    {\n      "a": "q\\"\\\\\\n\\u0001",\n      "z": null\n'
        '    }\n  ],\n  "c": [],\n  "d": true\n}')
    assert json.loads(output.to_json(v, pretty=True)) == v


def test_sidecar_and_descriptor_layout():
    """fn-6: epoch, rate, LO and sample count at the top level, the descriptor
    naming the contributing transmitters (with PRNs) and embedding the scene."""
    scene = workloads.scene2(duration_s=1.0)
    scene.receivers[0].bands[0].frontend = FrontendConfig(gain_db=10.0, nonlinearity=NonlinearityConfig(-10.0),
                                                          adc=AdcConfig(8, 1.0))
    scene.transmitters[1].gate = GateConfig(10.0, [(1.0, 2.0)])
    desc = output.band_descriptor(scene, "rx", scene.receivers[0].bands[1], [1, 2], 1.5e-13)
    side = output.sidecar_value("rx__l1", scene.scene.epoch_unix, 2.048e6, 1575.42e6, 123, desc)
    assert sorted(side) == ["descriptor", "epoch_unix", "format", "lo_carrier_hz", "n_samples", "name", "numpy_load",
                            "sample_rate_hz"]
    assert side["format"] == "raw complex64 little-endian (interleaved f32 re, f32 im)"
    assert side["numpy_load"] == "np.fromfile(path, dtype=np.complex64)"
    assert sorted(desc) == ["band", "band_id", "contributing_transmitters", "duration_s", "receiver", "scene_config"]
    ct = desc["contributing_transmitters"]
    assert [c["prn"] for c in ct] == [1, 2] and all(c["kind"] == "ca_code" and c["nav"] == "dummy" for c in ct)
    assert ct[0]["bandwidth_hz"] == 2.046e6 and ct[0]["trajectory_kind"] == "circular_keplerian"
    sc = desc["scene_config"]
    assert sorted(sc) == ["receiver", "scene", "transmitter"]
    tx0, tx1 = sc["transmitter"][0], sc["transmitter"][1]
    # flattened signal kind, Option fields as null, signal_epoch_unix/gate skipped when unset
    assert tx0["kind"] == "bpsk" and tx0["bandwidth_hz"] is None and "gate" not in tx0 and "signal_epoch_unix" not in tx0
    assert tx1["gate"] == {"period_s": 10.0, "epoch_unix": 0.0, "bursts": [{"offset_s": 1.0, "duration_s": 2.0}]}
    band0 = sc["receiver"][0]["band"][0]
    assert band0["frontend"] == {"antenna": None, "gain_db": 10.0, "nonlinearity": {"iip3_dbm": -10.0},
                                 "adc": {"bits": 8, "full_scale_v": 1.0}}
    assert sc["receiver"][0]["band"][1]["frontend"] is None
    assert sc["scene"]["gps_utc_offset_s"] == 18.0
    text = output.to_json(side, pretty=True)
    assert json.loads(text) == json.loads(json.dumps(side))
    assert '"carrier_hz": 437000000.0' in text and '"payload_bits": 287968' in text


def test_scene_from_toml_serialises_back_to_its_fields():
    toml = """
[scene]
epoch_unix = 1717260840.0
duration_s = 2.0

[[transmitter]]
name = "tone"
carrier_hz = 437e6
kind = "cw"
baseband_offset_hz = 1500.0
trajectory = { kind = "stationary_ecef", x_m = 7e6, y_m = 0.0, z_m = 0.0 }

[[receiver]]
name = "rx"
trajectory = { kind = "stationary_geodetic", lat_deg = 55.86, lon_deg = -4.25, h_m = 10.0 }
[[receiver.band]]
band_id = "uhf"
lo_carrier_hz = 437e6
sample_rate_hz = 48000.0
"""
    scene = SceneConfig.from_toml_str(toml)
    v = output.scene_config_value(scene)
    assert v["transmitter"] == [{"name": "tone", "carrier_hz": 437e6, "tx_power_dbw": 0.0, "bandwidth_hz": None,
                                 "trajectory": {"kind": "stationary_ecef", "x_m": 7e6, "y_m": 0.0, "z_m": 0.0},
                                 "kind": "cw", "baseband_offset_hz": 1500.0}]
    assert v["receiver"][0]["clock"] == {"allan_deviation_1s": 0.0, "flicker_floor": 0.0, "frequency_offset_y0": 0.0,
                                         "initial_time_offset_s": 0.0}
    assert v["receiver"][0]["band"] == [{"band_id": "uhf", "lo_carrier_hz": 437e6, "sample_rate_hz": 48000.0,
                                         "noise_temperature_k": 290.0, "noise_power_w": None, "bandwidth_hz": None,
                                         "frontend": None}]
    assert '"x_m":7000000.0' in output.to_json(v)
