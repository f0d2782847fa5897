"""render_window on the GPU: the reference's window tests
(tests/window_render.rs) against this path. One receiver, one band, one
absolute-time window of a sky, streamed chunk by chunk to a sink:
  - it is the same render as the equivalent one-receiver scene;
  - chunk size does not change it, and the sink can stop it;
  - two windows that overlap in absolute time see the same emitted signal;
  - a burst gate keys the received signal at the light-delayed edges;
  - an empty sky is the receiver's own noise.
"""
import math
import os

import numpy as np
import pytest

from lipika_amd.config import (NodeClockSpec, ReceiverBandConfig, ReceiverConfig, SceneConfig, SkyConfig,
                               StationaryEcef, StationaryGeodetic)
from lipika_amd.pipeline import BREAK, WindowOptions, render_window, run_scene

pytestmark = pytest.mark.gpu

FS = 48_000.0
HIBER_HZ = 401_150_000.0
# a whole multiple of 15 s, 300 s before the example pass culminates over Glasgow
PASS_START_UNIX = 1_717_260_690.0
C_LIGHT = 299_792_458.0
SKY_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hiber_beacon_sky.toml")


def hiber_sky_text():
    with open(SKY_PATH, encoding="utf-8") as f:
        return f.read()


def hiber_sky():
    return SkyConfig.from_toml_str(hiber_sky_text())


def glasgow(clock=None):
    return ReceiverConfig("glasgow", StationaryGeodetic(55.86, -4.25, 10.0), clock or NodeClockSpec())


def collect(backend, sky, rx, band, start, duration, seed, chunk):
    iq = []
    s = render_window(sky, rx, band, start, duration, WindowOptions(seed, chunk, 1), iq.append, backend=backend)
    iq = np.concatenate(iq) if iq else np.zeros(0, dtype=np.complex64)
    assert not s.stopped_early and s.n_samples_emitted == len(iq) == s.n_samples_window
    return iq


def mean_power(iq):
    return float(np.mean(np.abs(iq.astype(np.complex128)) ** 2))


def test_window_equals_the_equivalent_scene_render(backend):
    """window_render.rs:75-121: the example sky seen by a receiver with a
    non-trivial clock and a thermal floor, byte-identical to the scene file
    that says the same thing; the window spans a burst edge."""
    start, duration = PASS_START_UNIX - 0.3, 1.0
    scene_text = f"""
[scene]
epoch_unix = {start!r}
duration_s = {duration!r}
{hiber_sky_text()}
[[receiver]]
name = "glasgow"
trajectory = {{ kind = "stationary_geodetic", lat_deg = 55.86, lon_deg = -4.25, h_m = 10.0 }}
clock = {{ allan_deviation_1s = 1e-9, frequency_offset_y0 = 0.5e-6, initial_time_offset_s = 1e-7 }}
[[receiver.band]]
band_id = "uhf"
lo_carrier_hz = {HIBER_HZ!r}
sample_rate_hz = {FS!r}
noise_temperature_k = 290.0
"""
    scene = SceneConfig.from_toml_str(scene_text)
    from_scene = run_scene(scene, 99, backend=backend)[0].iq
    clock = NodeClockSpec(allan_deviation_1s=1e-9, frequency_offset_y0=0.5e-6, initial_time_offset_s=1e-7)
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS, 290.0)
    from_window = collect(backend, hiber_sky(), glasgow(clock), band, start, duration, 99, 5000)
    assert len(from_window) == 48_000
    assert from_window.tobytes() == np.asarray(from_scene, dtype=np.complex64).tobytes()
    assert mean_power(from_window[24_000:]) > 3.0 * mean_power(from_window[:12_000])


def test_chunk_size_does_not_change_the_window(backend):
    """window_render.rs:123-146: every chunk size (1 sample included) gives the
    same bytes; the same seed the same bytes, another seed other noise."""
    sky = hiber_sky()
    rx = glasgow(NodeClockSpec(allan_deviation_1s=1e-9, flicker_floor=1e-10, frequency_offset_y0=-0.3e-6))
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS)
    start, duration = PASS_START_UNIX - 0.2, 0.8
    whole = collect(backend, sky, rx, band, start, duration, 7, 1 << 20)
    assert len(whole) == 38_400
    for chunk in (1, 997, 4096, 38_400, 0):
        other = collect(backend, sky, rx, band, start, duration, 7, chunk)
        assert whole.tobytes() == other.tobytes(), chunk
    assert whole.tobytes() == collect(backend, sky, rx, band, start, duration, 7, 1 << 20).tobytes()
    assert whole.tobytes() != collect(backend, sky, rx, band, start, duration, 8, 1 << 20).tobytes()


def test_chunks_arrive_in_order_and_the_sink_can_stop_the_render(backend):
    """window_render.rs:148-192"""
    sky = hiber_sky()
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS)
    whole = collect(backend, sky, glasgow(), band, PASS_START_UNIX, 0.5, 3, 1000)
    assert len(whole) == 24_000
    sizes = []
    render_window(sky, glasgow(), band, PASS_START_UNIX, 0.5, WindowOptions(3, 7000, 1), lambda c: sizes.append(len(c)),
                  backend=backend)
    assert sizes == [7000, 7000, 7000, 3000]
    got, calls = [], [0]

    def sink(c):
        got.append(c)
        calls[0] += 1
        return BREAK if calls[0] == 3 else None

    s = render_window(sky, glasgow(), band, PASS_START_UNIX, 0.5, WindowOptions(3, 1000, 1), sink, backend=backend)
    assert calls[0] == 3 and s.stopped_early and s.n_samples_emitted == 3000 and s.n_samples_window == 24_000
    assert s.contributing_transmitters == ["hiber_beacon"]
    assert np.concatenate(got).tobytes() == whole[:3000].tobytes()
    s = render_window(sky, glasgow(), band, PASS_START_UNIX, 0.5, WindowOptions(3, 24_000, 1), lambda c: BREAK,
                      backend=backend)
    assert not s.stopped_early and s.n_samples_emitted == 24_000


def max_relative_difference(a, b):
    """Largest per-sample difference relative to the RMS of the keyed-on signal."""
    p = np.abs(a.astype(np.complex128)) ** 2
    on = p[p > 0.0]
    assert len(on)
    rms = math.sqrt(float(np.mean(on)))
    return float(np.max(np.abs(a.astype(np.complex128) - b.astype(np.complex128)))) / rms


def test_overlapping_windows_agree_on_a_leo_beacon(backend):
    """window_render.rs:205-244: windows [T - 0.5, T + 2.5) and [T + 0.75,
    T + 3) agree over their 1.75 s overlap to the f64 step of absolute time in
    the orbit query; same on/off pattern sample for sample (burst 2: 24 000
    samples); the BPSK sign flips line up."""
    sky = hiber_sky()
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS, noise_power_w=0.0)
    a = collect(backend, sky, glasgow(), band, PASS_START_UNIX - 0.5, 3.0, 1, 4096)
    b = collect(backend, sky, glasgow(), band, PASS_START_UNIX + 0.75, 2.25, 2, 9000)
    lead = int(1.25 * FS)
    overlap = len(a) - lead
    assert overlap == 84_000
    a_ov, b_ov = a[lead:], b[:overlap]
    assert max_relative_difference(a_ov, b_ov) < 0.03
    on_a, on_b = np.abs(a_ov) > 0.0, np.abs(b_ov) > 0.0
    assert np.array_equal(on_a, on_b)
    assert 23_999 <= int(on_a.sum()) <= 24_001

    def flips(x):
        x = x.astype(np.complex128)
        ok = (np.abs(x[1:]) > 0.0) & (np.abs(x[:-1]) > 0.0) & ((x[1:] * np.conj(x[:-1])).real < 0.0)
        return np.flatnonzero(ok) + 1

    fa, fb = flips(a_ov), flips(b_ov)
    assert len(fa) > 500
    assert np.array_equal(fa, fb)


ANCHORED_SKY = """
[[transmitter]]
name = "tone"
carrier_hz = 433_000_000.0
kind = "cw"
baseband_offset_hz = 1234.5
trajectory = { kind = "stationary_ecef", x_m = 3600000.0, y_m = -270000.0, z_m = 5300000.0 }

[[transmitter]]
name = "bpsk"
carrier_hz = 434_000_000.0
kind = "bpsk"
bit_rate_bps = 1200.0
payload_bits = 95
payload_seed = 4
sync_word = [-1, 1, 1, -1, 1]
signal_epoch_unix = 1000000007.25
trajectory = { kind = "stationary_ecef", x_m = 3600000.0, y_m = -270000.0, z_m = 5300000.0 }

[[transmitter]]
name = "ook"
carrier_hz = 435_000_000.0
kind = "ook"
bit_rate_bps = 2400.0
payload_bits = 61
payload_seed = 5
sync_word = [1, 0, 1]
trajectory = { kind = "stationary_ecef", x_m = 3600000.0, y_m = -270000.0, z_m = 5300000.0 }

[[transmitter]]
name = "qam"
carrier_hz = 436_000_000.0
kind = "qam16"
symbol_rate_baud = 4800.0
payload_symbols = 50
payload_seed = 6
sync_word = [1, 0, 1, 1]
trajectory = { kind = "stationary_ecef", x_m = 3600000.0, y_m = -270000.0, z_m = 5300000.0 }

[[transmitter]]
name = "code"
carrier_hz = 437_000_000.0
kind = "ca_code"
prn = 7
epoch_unix = 1000000000.0
nav = "off"
trajectory = { kind = "stationary_ecef", x_m = 3600000.0, y_m = -270000.0, z_m = 5300000.0 }
"""


def test_overlapping_windows_agree_for_every_anchored_signal_kind(backend):
    """window_render.rs:246-332: stationary emitters of every anchored kind,
    each alone in a narrow band: two windows 0.75 s apart agree over their
    overlap to 1e-3 of the signal RMS. A C/A source with a NAV overlay is
    refused in a window render."""
    import dataclasses
    sky = SkyConfig.from_toml_str(ANCHORED_SKY)
    start_a, lead_s = 1_717_260_840.0, 0.75
    for name, lo, fs in (("tone", 433.0e6, 48_000.0), ("bpsk", 434.0e6, 48_000.0), ("ook", 435.0e6, 48_000.0),
                         ("qam", 436.0e6, 48_000.0), ("code", 437.0e6, 4_000_000.0)):
        band = ReceiverBandConfig(name, lo, fs, noise_power_w=0.0, bandwidth_hz=min(100_000.0, fs))
        dur_a = 0.8 if fs > 1.0e6 else 1.5
        a = collect(backend, sky, glasgow(), band, start_a, dur_a, 1, 1 << 16)
        b = collect(backend, sky, glasgow(), band, start_a + lead_s, dur_a - lead_s + 0.1, 1, 1 << 16)
        lead = int(lead_s * fs)
        overlap = len(a) - lead
        diff = max_relative_difference(a[lead:], b[:overlap])
        print(f"{name}: overlap differs by {diff:.2e} of the signal RMS", flush=True)
        assert diff < 1e-3, name
    with_nav = SkyConfig.from_toml_str(ANCHORED_SKY)
    for tx in with_nav.transmitters:
        if tx.kind.__class__.__name__ == "CaCode":
            tx.kind = dataclasses.replace(tx.kind, nav="dummy")
    band = ReceiverBandConfig("code", 437.0e6, 4_000_000.0)
    with pytest.raises((ValueError, NotImplementedError)):
        render_window(with_nav, glasgow(), band, start_a, 0.001, WindowOptions(), lambda c: None, backend=backend)


GATED_TONE_SKY = """
[[transmitter]]
name = "keyed_tone"
carrier_hz = 401_150_000.0
tx_power_dbw = 0.0
kind = "cw"
baseband_offset_hz = 0.0
gate = { period_s = 20, epoch_unix = 5.0, bursts = [
    { offset_s = 2.0, duration_s = 0.5 },
    { offset_s = 7.0, duration_s = 0.25 },
] }
trajectory = { kind = "stationary_ecef", x_m = 6678137.0, y_m = 0.0, z_m = 0.0 }
"""


def test_gate_keys_the_received_signal_at_the_light_delayed_edges(backend):
    """window_render.rs:334-393: a gated tone 300 km above a receiver on the
    equator (one-way delay 48.03 samples at 48 kHz): on exactly from the
    light-delayed burst edges, at the free-space amplitude, silent otherwise;
    ungated, on throughout; a burst that does not fit its period is an error."""
    import dataclasses
    sky = SkyConfig.from_toml_str(GATED_TONE_SKY)
    rx = ReceiverConfig("equator", StationaryEcef(6_378_137.0, 0.0, 0.0))
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS, noise_power_w=0.0)
    period_start = 1_717_260_840.0 + 5.0
    iq = collect(backend, sky, rx, band, period_start + 1.9, 6.0, 0, 10_000)
    delay = 300_000.0 / C_LIGHT * FS

    def edge(t):
        return int(math.ceil(t * FS + delay))

    on1, off1, on2, off2 = edge(0.1), edge(0.6), edge(5.1), edge(5.35)
    assert (on1, off1, on2, off2) == (4849, 28_849, 244_849, 256_849)
    amp = (C_LIGHT / HIBER_HZ) / (4.0 * math.pi * 300_000.0)
    n = np.arange(len(iq))
    on = ((n >= on1) & (n < off1)) | ((n >= on2) & (n < off2))
    a = np.abs(iq.astype(np.complex128))
    assert np.all(np.abs(a[on] / amp - 1.0) < 1e-4)
    assert np.all(iq[~on] == 0)
    ungated = SkyConfig.from_toml_str(GATED_TONE_SKY)
    ungated.transmitters[0].gate = None
    assert np.all(np.abs(collect(backend, ungated, rx, band, period_start + 1.9, 0.2, 0, 10_000)) > 0.0)
    bad = SkyConfig.from_toml_str(GATED_TONE_SKY)
    g = bad.transmitters[0].gate
    bad.transmitters[0].gate = dataclasses.replace(g, bursts=[g.bursts[0], (19.9, g.bursts[1][1])])
    with pytest.raises(ValueError):
        render_window(bad, rx, band, period_start, 0.01, WindowOptions(), lambda c: None, backend=backend)


def test_example_sky_is_three_bursts_every_fifteen_seconds(backend):
    """window_render.rs:395-429: one gate period of the example beacon, opened
    0.2 s early: exactly three keyed-on runs of 24 000 +- 1 samples, 1.5 s
    apart, each 2-8 ms of light time late."""
    sky = hiber_sky()
    assert len(sky.transmitters) == 1
    tx = sky.transmitters[0]
    assert tx.carrier_hz == 401_150_000.0 and tx.kind.bit_rate_bps == 4800.0
    assert tx.trajectory.kind == "tle"
    assert tx.gate.period_s == 15.0 and len(tx.gate.bursts) == 3
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS, noise_power_w=0.0)
    iq = collect(backend, sky, glasgow(), band, PASS_START_UNIX - 0.2, 15.0, 0, 1 << 16)
    on = np.abs(iq) > 0.0
    edges = np.flatnonzero(np.diff(np.concatenate(([0], on.astype(np.int8), [0]))))
    runs = list(zip(edges[0::2], edges[1::2] - edges[0::2]))
    assert len(runs) == 3, runs
    for k, (start, length) in enumerate(runs):
        assert 23_999 <= length <= 24_001, (k, length)
        late = start - (0.2 + 1.5 * k) * FS
        assert 0.002 * FS <= late < 0.008 * FS, (k, late)


def test_empty_sky_is_receiver_noise_only(backend):
    """window_render.rs:431-474: an empty document is a valid sky; its window
    is white noise at k T B; a sky whose only transmitter is out of band is the
    same recording; with the noise off, silence; a zero-length window is
    valid and empty."""
    sky = SkyConfig.from_toml_str("")
    assert sky.transmitters == []
    band = ReceiverBandConfig("uhf", HIBER_HZ, FS, 290.0)
    iq = []
    s = render_window(sky, glasgow(), band, PASS_START_UNIX, 2.0, WindowOptions(5, 8192, 1), iq.append,
                      backend=backend)
    iq = np.concatenate(iq)
    assert s.contributing_transmitters == [] and len(iq) == 96_000
    floor = 1.380649e-23 * 290.0 * FS
    x = iq.astype(np.complex128)
    assert abs(mean_power(iq) / floor - 1.0) < 0.02
    assert abs(np.mean(x)) < 0.02 * math.sqrt(floor)
    assert abs(float(np.mean((x[:-1] * np.conj(x[1:])).real))) < 0.02 * floor
    far = SkyConfig.from_toml_str("""
[[transmitter]]
name = "elsewhere"
carrier_hz = 1_575_420_000.0
kind = "cw"
baseband_offset_hz = 0.0
trajectory = { kind = "stationary_ecef", x_m = 7000000.0, y_m = 0.0, z_m = 0.0 }
""")
    assert iq.tobytes() == collect(backend, far, glasgow(), band, PASS_START_UNIX, 2.0, 5, 8192).tobytes()
    quiet = ReceiverBandConfig("uhf", HIBER_HZ, FS, 290.0, noise_power_w=0.0)
    assert np.all(collect(backend, sky, glasgow(), quiet, PASS_START_UNIX, 0.1, 5, 8192) == 0)
    assert len(collect(backend, sky, glasgow(), band, PASS_START_UNIX, 0.0, 5, 8192)) == 0
