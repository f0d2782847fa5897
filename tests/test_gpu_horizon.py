"""The horizon gate on the GPU path: the reference's gate tests
(tests/horizon_gate.rs) that a stationary receiver covers. With the gate on,
(tile, transmitter) pairs that cannot be above the antenna cutoff are
skipped; the IQ bytes, the window summaries and the written files must not
change. The gate is a context switch here (lpk_set_horizon_gate), where the
reference has *_without_horizon_gate twins of its entry points.
"""
import math
import os

import numpy as np
import pytest

from lipika_amd import _lib
from lipika_amd.config import (AdcConfig, Bpsk, CircularKeplerian, Cw, FrontendConfig, GateConfig, ReceiverBandConfig,
                               ReceiverConfig, SceneConfig, SceneMeta, SkyConfig, StationaryGeodetic,
                               TransmitterConfig)
from lipika_amd.pipeline import WindowOptions, render_window, run_scene, run_scene_to_dir, trajectory_state

pytestmark = pytest.mark.gpu

CARRIER_HZ = 401_150_000.0
FS = 1_000.0
START_UNIX = 1_717_260_690.0
DURATION_S = 120.0
GLASGOW = (55.86, -4.25, 20.0)
SYNC32 = [-1, 1, 1, -1, 1, -1, -1, 1, 1, 1, 1, -1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1,
          -1, 1]


def glasgow():
    return ReceiverConfig("glasgow", StationaryGeodetic(*GLASGOW))


def band(cutoff_deg):
    """horizon_gate.rs:46-57: the test carrier, a 2 dBi antenna and this cutoff."""
    fe = FrontendConfig(antenna={"peak_dbi": 2.0, "horizon_dbi": -1.0, "fb_db": 6.0, "cutoff_deg": cutoff_deg})
    return ReceiverBandConfig("uhf", CARRIER_HZ, FS, 290.0, frontend=fe)


def station(lat, lon, h):
    return trajectory_state(StationaryGeodetic(lat, lon, h), 0.0)[0]


def elevation_deg(o, p):
    d = p - o
    return math.degrees(math.asin(float(np.dot(d, o)) / (np.linalg.norm(d) * np.linalg.norm(o))))


def classify(orbit, o):
    """horizon_gate.rs:83-102: what the orbit does over Glasgow in the window."""
    def el(t):
        return elevation_deg(o, trajectory_state(orbit, START_UNIX + t)[0])
    first, last = el(0.0), el(DURATION_S)
    if first < -25.0 and last < -25.0:
        return "stays_down" if first < -40.0 else None
    track = [el(k * 0.25) for k in range(int(DURATION_S) * 4 + 1)]
    mx, mn = max(track), min(track)
    if first < -2.0 and last > 2.0:
        return "rises"
    if first > 2.0 and last < -2.0:
        return "sets"
    if mn > 5.0:
        return "stays_up"
    if 0.05 < mx < 1.0:
        return "grazes_above"
    if -1.0 < mx < -0.05:
        return "grazes_below"
    return None


_SKY = None


def leo_sky():
    """horizon_gate.rs:108-176: 19 circular LEOs chosen by what they do in the
    window (three of each kind of pass, four that stay down); every third a
    burst-gated BPSK beacon, the rest CW tones."""
    global _SKY
    if _SKY is not None:
        return _SKY
    o = station(*GLASGOW)
    have, txs, passes, k = {}, [], [], 0
    while len(txs) < 19:
        k += 1
        assert k < 2_000_000
        a = 6_878_137.0 + 30_000.0 * (k % 5)
        inc = [53.0, 65.0, 87.9, 97.5][k % 4]
        raan, nu = (k * 137.508) % 360.0, (k * 61.803) % 360.0
        orbit = CircularKeplerian(a, inc, raan, nu, START_UNIX)
        p = classify(orbit, o)
        if p is None or have.get(p, 0) >= (4 if p == "stays_down" else 3):
            continue
        have[p] = have.get(p, 0) + 1
        i = len(txs)
        if i % 3 == 2:
            kind = Bpsk(200.0, 268, 40 + i, SYNC32)
            gate = GateConfig(5.0, [(0.3 * (i % 5), 1.5), (3.0, 0.7)], 0.0)
        else:
            kind, gate = Cw(-400.0 + 43.0 * i), None
        txs.append(TransmitterConfig(f"leo-{i:02d}", CARRIER_HZ, kind, orbit, 0.0, signal_epoch_unix=0.0, gate=gate))
        passes.append((f"leo-{i:02d}", p))
    _SKY = (SkyConfig(txs), passes)
    return _SKY


def collect(backend, gate, sky, rx, b, start, duration, chunk, form="plain"):
    backend.set_horizon_gate(gate)
    try:
        iq = []
        s = render_window(sky, rx, b, start, duration, WindowOptions(11, chunk, 1), iq.append, form=form,
                          backend=backend)
    finally:
        backend.set_horizon_gate(True)
    iq = np.concatenate(iq) if iq else np.zeros(0, dtype=np.complex64)
    assert s.n_samples_emitted == len(iq)
    return iq, s


def test_gate_respects_the_configured_cutoff_and_its_absence(backend):
    """horizon_gate.rs:256-289: for cutoffs of -3, 0 and 10 deg and none, the
    gated window equals the ungated one byte for byte with the same summary;
    a higher cutoff hides more; no cutoff, or no antenna, sees all 19."""
    sky, passes = leo_sky()
    assert sorted(p for _, p in passes).count("grazes_above") == 3
    visible = []
    for cutoff in (-3.0, 0.0, 10.0, None):
        b = band(cutoff)
        ref, rs = collect(backend, False, sky, glasgow(), b, START_UNIX, DURATION_S, 1 << 20)
        got, gs = collect(backend, True, sky, glasgow(), b, START_UNIX, DURATION_S, 777)
        assert got.tobytes() == ref.tobytes(), cutoff
        assert gs == rs, cutoff
        visible.append(len(gs.visible_transmitters))
        if cutoff is None:
            assert gs.visible_transmitters == gs.contributing_transmitters
    print(f"visible per cutoff (-3, 0, 10, none): {visible}", flush=True)
    assert visible[0] > visible[1] > visible[2] and visible[3] == 19
    bare = ReceiverBandConfig("uhf", CARRIER_HZ, FS, 290.0)
    ref, _ = collect(backend, False, sky, glasgow(), bare, START_UNIX, 20.0, 1 << 20)
    got, gs = collect(backend, True, sky, glasgow(), bare, START_UNIX, 20.0, 999)
    assert got.tobytes() == ref.tobytes() and len(gs.visible_transmitters) == 19


def test_gate_does_not_change_a_scene_in_memory_or_streamed(backend, tmp_path):
    """horizon_gate.rs:291-343: a two-receiver scene (a noiseless band behind
    a 12-bit ADC, a band with no cutoff, a 5 deg site with thermal noise): the
    gated run_scene equals the ungated one byte for byte, descriptors
    included, and so do the bands streamed to files in small chunks."""
    sky, _ = leo_sky()
    base = band(0.0)
    fe = FrontendConfig(antenna=base.frontend.antenna, adc=AdcConfig(12, 1.0e-6))
    noiseless = ReceiverBandConfig("uhf", CARRIER_HZ, FS, 290.0, noise_power_w=0.0, frontend=fe)
    open_ = band(None)
    open_.band_id = "open"
    scene = SceneConfig(SceneMeta(START_UNIX, 60.0, 18.0), list(sky.transmitters),
                        [ReceiverConfig("glasgow", StationaryGeodetic(*GLASGOW), bands=[noiseless, open_]),
                         ReceiverConfig("svalbard", StationaryGeodetic(78.23, 15.39, 450.0), bands=[band(5.0)])])
    backend.set_horizon_gate(False)
    try:
        ref = run_scene(scene, 5, backend=backend)
    finally:
        backend.set_horizon_gate(True)
    got = run_scene(scene, 5, backend=backend)
    assert len(ref) == 3
    for r, g in zip(ref, got):
        assert r.name == g.name
        assert np.asarray(g.iq).tobytes() == np.asarray(r.iq).tobytes(), r.name
        assert r.descriptor == g.descriptor
    assert np.any(np.abs(got[0].iq) > 0.0)
    manifests = run_scene_to_dir(scene, 5, str(tmp_path), 1234, backend=backend)
    assert len(manifests) == 3
    for m, r in zip(manifests, ref):
        assert m.name == r.name
        with open(m.path, "rb") as f:
            assert f.read() == np.asarray(r.iq, dtype=np.complex64).tobytes(), r.name


def test_gate_handles_earth_fixed_transmitters(backend):
    """horizon_gate.rs:452-483: a mast 30 km away and a balloon overhead are
    seen; a site at the antipode and one beyond the horizon are not; gated
    equals ungated."""
    def fixed(name, lat, lon, h, off):
        return TransmitterConfig(name, CARRIER_HZ, Cw(off), StationaryGeodetic(lat, lon, h), -60.0,
                                 signal_epoch_unix=0.0)
    sky = SkyConfig([fixed("mast", 55.86, -3.77, 3_000.0, 100.0), fixed("balloon", 55.86, -4.25, 30_000.0, -200.0),
                     fixed("antipode", -55.86, 175.75, 100.0, 300.0),
                     fixed("over-the-horizon", 50.0, -4.25, 100.0, -350.0)])
    b = band(0.0)
    ref, rs = collect(backend, False, sky, glasgow(), b, START_UNIX, 5.0, 1 << 20)
    got, gs = collect(backend, True, sky, glasgow(), b, START_UNIX, 5.0, 700)
    assert got.tobytes() == ref.tobytes() and gs == rs
    assert len(gs.contributing_transmitters) == 4
    assert gs.visible_transmitters == ["mast", "balloon"]
