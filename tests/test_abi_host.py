"""CPU tests of the C-ABI library (loads, exports every symbol the header
declares, host-side table builders agree bit for bit with the oracle) and of
the host mirror of the reference's scene orchestration. No GPU needed."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from lipika_amd import _lib, workloads
from lipika_amd.config import (
    AdcConfig, Bpsk, CaCode, CircularKeplerian, Cw, FrontendConfig, GateConfig, NodeClockSpec, Ook, Qam16,
    ReceiverBandConfig, ReceiverConfig, SceneConfig, SceneMeta, SkyConfig, StationaryEcef, StationaryGeodetic,
    TransmitterConfig,
)
from lipika_amd.pipeline import (
    Backend, TimeBase, WindowOptions, clock_seed, noise_seed, prepare, render_window, rust_round,
)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = C.POINTER


def header_functions():
    text = open(os.path.join(ROOT, "include", "lpk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lpk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lpk):
    names = header_functions()
    assert len(names) >= 25
    assert sorted(_lib.EXPORTED) == names
    for n in names:
        assert hasattr(lpk, n), n
    assert lpk.lpk_abi_version() == _lib.ABI_VERSION


def test_ctypes_layout_matches_header(tmp_path):
    """Every ctypes structure of the binding has the size and field offsets the
    C header gives (compiled here with gcc against include/lpk.h)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc is not available")
    structs = {"lpk_tx_desc": _lib.TxDesc, "lpk_rx_desc": _lib.RxDesc, "lpk_post_op": _lib.PostOp,
               "lpk_band_chunk": _lib.BandChunk, "lpk_tuning": _lib.Tuning,
               "lpk_lws_train_cfg": _lib.LwsTrainCfg, "lpk_lws_train_report": _lib.LwsTrainReport,
               "lpk_ephemeris": _lib.Ephemeris, "lpk_tle": _lib.Tle}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "lpk.h"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in ct._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, ct in structs.items():
        assert got[(cname, "size")] == C.sizeof(ct), cname
        for field, _ in ct._fields_:
            assert got[(cname, field)] == getattr(ct, field).offset, (cname, field)


def test_context_creation_fails_cleanly_without_gpu(lpk):
    """Where no gfx950 is visible the ABI returns an error code instead of
    aborting (on a GPU box this succeeds)."""
    h = C.c_void_p()
    rc = lpk.lpk_ctx_create(0, C.byref(h))
    if rc == 0:
        lpk.lpk_ctx_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc in (-6, -1, -3) and not h.value
    assert lpk.lpk_last_error(None)
    assert lpk.lpk_render_chunk(None, None, 0, 0) == -1
    assert lpk.lpk_scene_upload(None, None, 0, None, 0) == -1


# ---------------------------------------------------------------- host builders vs oracle

def test_host_builders_match_oracle(lpk, orc):
    a, b = (C.c_double * 3)(), (C.c_double * 3)()
    for lat, lon, h in ((55.86, -4.25, 10.0), (-33.9, 151.2, 50.0), (0.0, 0.0, 0.0), (89.9, 10.0, 2000.0)):
        lpk.lpk_geodetic_to_ecef(lat, lon, h, a)
        orc.orc_geodetic_to_ecef(lat, lon, h, b)
        assert a[:] == b[:]
    for prn in range(1, 33):
        x, y = np.empty(1023, np.int8), np.empty(1023, np.int8)
        lpk.lpk_ca_code(prn, x.ctypes.data_as(P(C.c_int8)))
        orc.orc_ca_code(prn, y.ctypes.data_as(P(C.c_int8)))
        assert np.array_equal(x, y)
    for orbit in ((6778137.0, 51.6, 0.0, 10.0, 1717260840.0), (26578137.0, 55.0, 240.0, 230.0, 1717260840.0)):
        n = lpk.lpk_grid_size(60.0, 0.01)
        assert n == orc.orc_grid_size(60.0, 0.01) == 6101
        g1, g2 = np.empty(3 * n), np.empty(3 * n)
        t1, t2 = C.c_double(), C.c_double()
        assert lpk.lpk_grid_fill_kepler(*orbit, 1717260840.0, 60.0, 0.01, C.byref(t1), g1.ctypes.data_as(P(C.c_double)), n) == 0
        orc.orc_grid_sample_kepler(*orbit, 1717260840.0, 60.0, 0.01, C.byref(t2), g2.ctypes.data_as(P(C.c_double)))
        assert t1.value == t2.value and g1.tobytes() == g2.tobytes()
    sync = np.array([-1, 1, 1, -1], np.int8)
    p1, p2 = np.empty(1004, np.int8), np.empty(1004, np.int8)
    lpk.lpk_payload_bpsk(sync.ctypes.data_as(P(C.c_int8)), 4, 1000, 2026, p1.ctypes.data_as(P(C.c_int8)))
    orc.orc_payload_bpsk(sync.ctypes.data_as(P(C.c_int8)), 4, 1000, 2026, p2.ctypes.data_as(P(C.c_int8)))
    assert np.array_equal(p1, p2)
    s8 = np.array([1, 0, 1, 1], np.uint8)
    lpk.lpk_payload_ook(s8.ctypes.data_as(P(C.c_uint8)), 4, 1000, 5, p1.ctypes.data_as(P(C.c_int8)))
    orc.orc_payload_ook(s8.ctypes.data_as(P(C.c_uint8)), 4, 1000, 5, p2.ctypes.data_as(P(C.c_int8)))
    assert np.array_equal(p1, p2)
    q1, q2 = np.empty(2002, np.float32), np.empty(2002, np.float32)
    lpk.lpk_payload_qam16(s8.ctypes.data_as(P(C.c_uint8)), 4, 1000, 9, q1.ctypes.data_as(P(C.c_float)))
    orc.orc_payload_qam16(s8.ctypes.data_as(P(C.c_uint8)), 4, 1000, 9, q2.ctypes.data_as(P(C.c_float)))
    assert q1.tobytes() == q2.tobytes()
    for e, r in ((1717260840.0, 4800.0), (-0.25, 2.0), (1e9, 9600 / 7), (123.456, 1023000.0 / 1023)):
        w1, f1, w2, f2 = C.c_int64(), C.c_double(), C.c_int64(), C.c_double()
        assert lpk.lpk_elapsed_periods(e, r, C.byref(w1), C.byref(f1)) == orc.orc_elapsed_periods(e, r, C.byref(w2), C.byref(f2)) == 0
        assert (w1.value, f1.value) == (w2.value, f2.value)
    assert lpk.lpk_nav_fold_seed(7, 3) == orc.orc_nav_fold_seed(7, 3)
    assert lpk.lpk_thermal_noise_power_w(290.0, 2.048e6) == orc.orc_thermal_noise_power_w(290.0, 2.048e6)


def test_builders_reject_bad_input(lpk):
    out = np.empty(1023, np.int8)
    assert lpk.lpk_ca_code(0, out.ctypes.data_as(P(C.c_int8))) == -1
    bad = np.array([0, 2], np.uint8)
    assert lpk.lpk_payload_ook(bad.ctypes.data_as(P(C.c_uint8)), 2, 1, 0, out.ctypes.data_as(P(C.c_int8))) == -1
    q = np.empty(8, np.float32)
    assert lpk.lpk_payload_qam16(bad.ctypes.data_as(P(C.c_uint8)), 2, 1, 0, q.ctypes.data_as(P(C.c_float))) == -1
    w, f = C.c_int64(), C.c_double()
    assert lpk.lpk_elapsed_periods(float("nan"), 1.0, C.byref(w), C.byref(f)) == -4
    assert lpk.lpk_frame_anchor(1.0, 0.0, -1.0, 1, C.byref(f), C.byref(w)) == -1


def test_bands_overlap_cases(lpk, kats):
    for txc, txb, lo, bw, want in kats["bands_overlap"]["cases"]:
        assert lpk.lpk_bands_overlap(txc, txb, lo, bw) == want


# ---------------------------------------------------------------- scene preparation

DUAL_TOML = """
[scene]
epoch_unix = 1717260840.0
duration_s = 0.5

[[transmitter]]
name = "beacon"
carrier_hz = 437_000_000.0
tx_power_dbw = 0.0
kind = "bpsk"
bit_rate_bps = 9600.0
payload_bits = 100
payload_seed = 3
sync_word = [-1, 1, 1, -1]
trajectory = { kind = "circular_keplerian", semi_major_axis_m = 6778137.0, inclination_deg = 51.6, raan_deg = 0.0, true_anomaly_deg_at_epoch = 10.0, epoch_unix = 1717260840.0 }

[[transmitter]]
name = "gps_prn01"
carrier_hz = 1_575_420_000.0
tx_power_dbw = 27.0
kind = "ca_code"
prn = 1
epoch_unix = 1717260840.0
trajectory = { kind = "circular_keplerian", semi_major_axis_m = 26578137.0, inclination_deg = 55.0, raan_deg = 0.0, true_anomaly_deg_at_epoch = 30.0, epoch_unix = 1717260840.0 }

[[receiver]]
name = "glasgow"
trajectory = { kind = "stationary_geodetic", lat_deg = 55.86, lon_deg = -4.25, h_m = 10.0 }
clock = { allan_deviation_1s = 1e-11, frequency_offset_y0 = 0.5e-6 }
[[receiver.band]]
band_id = "uhf"
lo_carrier_hz = 437_000_000.0
sample_rate_hz = 48_000.0
[[receiver.band]]
band_id = "l1"
lo_carrier_hz = 1_575_420_000.0
sample_rate_hz = 2_048_000.0
noise_temperature_k = 150.0
[receiver.band.frontend]
gain_db = 20.0
adc = { bits = 4, full_scale_v = 1.0 }
"""


def test_toml_scene_prepares_like_the_reference():
    sc = prepare(SceneConfig.from_toml_str(DUAL_TOML))
    rx = sc.receivers[0]
    assert rx.master_rate_hz == 2.048e6
    uhf, l1 = rx.bands
    assert list(uhf.active) == [0] and list(l1.active) == [1]          # band-overlap gate, in tx order
    assert uhf.n == 24000 and l1.n == 1024000
    assert l1.noise_power_w == 1.380649e-23 * 150.0 * 2.048e6
    assert l1.post == [(_lib.POST_GAIN, 0, 20.0), (_lib.POST_ADC, 4, 1.0)]
    bpsk, ca = sc.transmitters
    assert bpsk.traj_kind == _lib.TRAJ_GRID and bpsk.grid.shape == (3, int(math.ceil(1.5 / 0.01)) + 1) and bpsk.grid_t0 == 1717260840.0 - 0.5
    assert list(bpsk.bits[:4]) == [-1, 1, 1, -1] and len(bpsk.bits) == 104 and not bpsk.repeating
    assert bpsk.bandwidth_hz == 19200.0 and ca.bandwidth_hz == 2046000.0
    assert ca.nav_mode == _lib.NAV_DUMMY and ca.nav_seed == _lib.load().lpk_nav_fold_seed(0, 1)
    assert ca.tx_power_w == math.pow(10.0, 2.7)


def test_seeds_and_rounding():
    assert clock_seed(5, 2) == (5 + 0xC10C_0000_0000_0000 + 2) % 2**64
    assert clock_seed(2**64 - 1, 1) == 0xC10C_0000_0000_0000
    assert noise_seed(2**64 - 1, 3) == 2
    assert rust_round(2.5) == 3 and rust_round(0.49999999999999994) == 0 and rust_round(48000.0 * 10.0) == 480000


def test_window_time_base_anchors_every_signal():
    """render_window's absolute time base (src/pipeline.rs:142-253): frames
    repeat from their signal epoch (default 0), C/A folds to its code period."""
    sky = SkyConfig([
        TransmitterConfig("b", 1e9, Bpsk(1000.0, 60, 1, [1, -1, 1, -1]), StationaryEcef(7e6, 0, 0)),
        TransmitterConfig("c", 1e9, Cw(123.0), StationaryEcef(7e6, 0, 0)),
        TransmitterConfig("g", 1e9, CaCode(3, 1717260840.0, "off"), StationaryEcef(7e6, 0, 0),
                          gate=GateConfig(15.0, [(0.0, 0.5)], 0.0)),
    ])
    scene = SceneConfig(SceneMeta(1717260844.0, 1.0), sky.transmitters,
                        [ReceiverConfig("r", StationaryEcef(6.4e6, 0, 0), NodeClockSpec(),
                                        [ReceiverBandConfig("x", 1e9, 2e6)])])
    sc = prepare(scene, TimeBase.ABSOLUTE)
    b, c, g = sc.transmitters
    assert b.repeating and -1e-3 < b.epoch_rel_s <= 0.0 and 0 <= b.repeat_from < 64
    assert c.cw_anchored and 0.0 <= c.anchor_phase < 2 * math.pi
    assert g.gated and g.gate_phase == math.fmod(1717260844.0, 15.0) and -1e-3 < g.epoch_rel_s <= 0
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(0.0, 1.0), [TransmitterConfig("d", 1e9, CaCode(1, 0.0, "dummy"),
                                                                     StationaryEcef(7e6, 0, 0))], []), TimeBase.ABSOLUTE)


def test_invalid_scenes_raise():
    tx = TransmitterConfig("o", 1e9, Ook(1000.0, 8, 0, [1, 0, 2]), StationaryEcef(7e6, 0, 0))
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(0.0, 1.0), [tx], []))
    tx = TransmitterConfig("q", 1e9, Qam16(1000.0, 8, 0, [1, 0, 1]), StationaryEcef(7e6, 0, 0))
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(0.0, 1.0), [tx], []))
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(0.0, 1.0), [], [ReceiverConfig("r", StationaryEcef(1, 0, 0), NodeClockSpec(), [])]))
    tx = TransmitterConfig("g", 1e9, Cw(0.0), StationaryEcef(7e6, 0, 0), gate=GateConfig(15.0, [(14.8, 0.5)]))
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(0.0, 1.0), [tx], []))
    # full LNAV is on the path (its builders are tested in test_gpsnav.py); an unknown mode is not
    sc = prepare(SceneConfig(SceneMeta(0.0, 1.0), [TransmitterConfig("f", 1e9, CaCode(1, 0.0, "full"),
                                                                      StationaryEcef(7e6, 0, 0))], []))
    assert sc.transmitters[0].nav_mode == _lib.NAV_FULL and len(sc.transmitters[0].nav_bits) == 600
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(0.0, 1.0), [TransmitterConfig("f", 1e9, CaCode(1, 0.0, "lnav"),
                                                                     StationaryEcef(7e6, 0, 0))], []))
    # flicker FM is on the path: a receiver with a flicker floor prepares
    sc = prepare(SceneConfig(SceneMeta(0.0, 1.0), [], [ReceiverConfig(
        "r", StationaryEcef(1, 0, 0), NodeClockSpec(flicker_floor=1e-10), [ReceiverBandConfig("x", 1e9, 1e6)])]))
    assert sc.receivers[0].clock.flicker_floor == 1e-10


def test_window_argument_validation():
    """src/pipeline.rs:759-767 -- checked before any device work."""
    rx = ReceiverConfig("r", StationaryEcef(6.4e6, 0, 0))
    band = ReceiverBandConfig("b", 1e9, 1e6)
    for start, dur, fs in ((float("nan"), 1.0, 1e6), (0.0, -1.0, 1e6), (0.0, float("inf"), 1e6), (0.0, 1.0, 0.0)):
        b = ReceiverBandConfig("b", 1e9, fs)
        with pytest.raises(ValueError):
            render_window(SkyConfig(), rx, b, start, dur, WindowOptions(), lambda c: None)


def test_a_sky_file_holds_transmitters_only():
    """tests/window_render.rs:475-484: scene-only tables are rejected, so a
    scene file is not silently read as a sky; the example sky file loads; a
    missing file is an error."""
    import os
    with pytest.raises(ValueError):
        SkyConfig.from_toml_str("[scene]\nepoch_unix = 0.0\nduration_s = 1.0\n")
    with pytest.raises(ValueError):
        SkyConfig.from_toml_str('[[receiver]]\nname = "rx"\ntrajectory = { kind = "stationary_ecef", x_m = 1.0, '
                                'y_m = 0.0, z_m = 0.0 }\nband = []\n')
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hiber_beacon_sky.toml")
    sky = SkyConfig.from_toml_file(path)
    assert sky.transmitters[0].name == "hiber_beacon" and sky.transmitters[0].trajectory.kind == "tle"
    assert SkyConfig.from_toml_str("").transmitters == []
    with pytest.raises(OSError):
        SkyConfig.from_toml_file("/nonexistent/sky.toml")


def test_workload_totals():
    p1 = prepare(workloads.scene1())
    assert workloads.totals(p1) == (480_000, 480_000)
    p2 = prepare(workloads.scene2())
    assert workloads.totals(p2) == (125_760_000, 740_160_000)
    assert [list(b.active) for b in p2.receivers[0].bands] == [[0], [1, 2, 3, 4, 5, 6]]
    p3 = prepare(workloads.scene3(duration_s=1.0))
    assert len(p3.transmitters) == 32 and len(p3.receivers) == 16
    assert all(len(b.active) == 32 for r in p3.receivers for b in r.bands)


def test_fnv1a64_known_answers(lpk):
    """FNV-1a 64 (Fowler, Noll, Vo; the hashing sink of
    examples/sky_throughput.rs:73-82) on its published test strings."""
    from lipika_amd import _lib
    for text, want in ((b"", 0xCBF29CE484222325), (b"a", 0xAF63DC4C8601EC8C), (b"foobar", 0x85944171F73967E8)):
        buf = C.create_string_buffer(text, len(text))
        assert lpk.lpk_fnv1a64(buf, len(text), _lib.FNV_OFFSET) == want, text
    # continuing from a state is hashing the concatenation
    a, b = C.create_string_buffer(b"foo", 3), C.create_string_buffer(b"bar", 3)
    assert lpk.lpk_fnv1a64(b, 3, lpk.lpk_fnv1a64(a, 3, _lib.FNV_OFFSET)) == 0x85944171F73967E8
