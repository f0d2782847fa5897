"""Moving receivers on the GPU path: a CircularKeplerian receiver whose
position the kernels evaluate at every sample's epoch + t_utc, as the chunk
body queries position_ecef_batch (src/pipeline.rs:929-930). Against the FP64
oracle, which takes the same per-sample positions (glibc's sincos there, the
device's f64 sincos here: positions within an ulp, so the IQ is ulp-bounded
rather than bit-identical); the orbiting-receiver horizon-gate test of the
reference (tests/horizon_gate.rs:345-370) on the GPU path; the forms that
need a fixed receiver refuse one that moves; and the truth export.
"""
import ctypes as C
import json

import numpy as np
import pytest

from lipika_amd import _lib, output
from lipika_amd.config import (Bpsk, CaCode, CircularKeplerian, Cw, NodeClockSpec, ReceiverBandConfig, ReceiverConfig,
                               SceneConfig, SceneMeta, StationaryGeodetic, TransmitterConfig)
from lipika_amd.pipeline import prepare
from oracle.pyoracle import OracleScene

import util
from test_gpu_horizon import DURATION_S, START_UNIX, band, collect, leo_sky
from test_gpu_parity import compare_iq, compare_taps, gpu_band

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _say(*a):
    print(*a, flush=True)


def orbiter(epoch):
    """horizon_gate.rs:349-356: a receiver 400 km up."""
    return CircularKeplerian(6_778_137.0, 70.0, 320.0, 50.0, epoch)


def moving_scene(duration_s=0.05, noise_power_w=0.0):
    """Two C/A MEOs, a BPSK LEO and a ground CW tone seen from a LEO receiver
    with a walking clock, one L1 band at 2.048 MHz."""
    txs = [TransmitterConfig("sv05", 1575.42e6, CaCode(5, util.EPOCH), CircularKeplerian(26.56e6, 55.0, 40.0, 20.0, util.EPOCH), 27.0),
           TransmitterConfig("sv12", 1575.42e6, CaCode(12, util.EPOCH), CircularKeplerian(26.56e6, 55.0, 100.0, 250.0, util.EPOCH), 27.0),
           TransmitterConfig("leo", 1575.42e6 - 5e3, Bpsk(9600.0, 480, 7, [1, -1, 1, 1]),
                             CircularKeplerian(7.1e6, 60.0, 330.0, 45.0, util.EPOCH), -10.0),
           TransmitterConfig("tone", 1575.42e6 + 3e3, Cw(250.0), StationaryGeodetic(40.0, 20.0, 100.0), -20.0)]
    band_ = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=noise_power_w)
    rx = ReceiverConfig("orbiter", orbiter(util.EPOCH), NodeClockSpec(allan_deviation_1s=1e-11, frequency_offset_y0=2e-7),
                        [band_])
    return SceneConfig(SceneMeta(util.EPOCH, duration_s), txs, [rx])


def test_moving_receiver_plain_form_against_oracle(backend):
    sc = prepare(moving_scene())
    seed = 31
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, taps = orc.render_band(0, tap_slot=0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3], chunk=7000)
    compare_iq(g, o, "moving receiver, plain, oracle clock offsets", strict=False)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed)
    compare_iq(g, o, "moving receiver, plain, device clock", strict=False)
    for slot in range(4):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 20_000, label="moving receiver")


def test_moving_receiver_with_noise_and_converged_form(backend):
    sc = prepare(moving_scene(noise_power_w=1e-15))
    seed = 32
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    # AWGN is the same Philox draw on both sides given the same clock offsets
    o, _, taps = orc.render_band(0, tap_slot=0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3])
    compare_iq(g, o, "moving receiver with noise, oracle clock offsets", strict=False)
    # the converged solve with the receiver moving within each iteration's range
    oc, _, _ = orc.render_band(0, converged=True)
    gc, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3], form=_lib.FORM_CONVERGED)
    compare_iq(gc, oc, "moving receiver, converged form, oracle clock offsets", strict=False)


@pytest.mark.parametrize("form", [_lib.FORM_PRECISION, _lib.FORM_LEARNED])
def test_forms_needing_a_fixed_receiver_refuse_a_moving_one(backend, form):
    sc = prepare(moving_scene())
    backend.upload(sc, 3)
    out = np.empty(1000, dtype=np.complex64)
    j, keep = backend.job(0, 0, 0, 3, 0, 1000, out_host=out)
    with pytest.raises(_lib.LpkError) as e:
        backend.render([j], form)
    assert _lib.STATUS[e.value.code] in ("LPK_E_INVALID", "LPK_E_STATE")
    backend.render([j], _lib.FORM_PLAIN)   # the context is still usable
    backend.sync()
    assert np.all(np.isfinite(out))


def test_gate_does_not_change_what_an_orbiting_receiver_records(backend):
    """horizon_gate.rs:345-370: a receiver in orbit below the 19 LEOs, its
    antenna blind more than 15 deg under its local horizontal; the gated
    window equals the ungated one byte for byte, with the same summary, at
    three chunkings; it sees some of the transmitters and not all."""
    sky, _ = leo_sky()
    rx = ReceiverConfig("orbiter", orbiter(START_UNIX), NodeClockSpec())
    b = band(-15.0)
    ref, rs = collect(backend, False, sky, rx, b, START_UNIX, DURATION_S, 1 << 20)
    for chunk in (1 << 20, 300, 5000):
        before = backend.horizon_gate_stats()
        got, gs = collect(backend, True, sky, rx, b, START_UNIX, DURATION_S, chunk)
        after = backend.horizon_gate_stats()
        skipped, tested = after["skipped"] - before["skipped"], after["tested"] - before["tested"]
        _say(f"orbiting receiver, chunks of {chunk}: the gate skipped {skipped} of {tested} (tile, link) pairs")
        assert got.tobytes() == ref.tobytes(), chunk
        assert gs == rs, chunk
        assert tested > 0 and skipped > 0
    visible = len(rs.visible_transmitters)
    _say(f"orbiting receiver: {visible} of 19 visible")
    assert 0 < visible < 19


def test_orbiting_receiver_with_antenna_against_oracle(backend):
    """The same orbiter, antenna and sky over 5 s, noise-free, rendered as a
    scene: the IQ against the oracle's (per-sample positions, arrival
    elevations and cutoff), and the transmitters each says are unoccluded."""
    sky, _ = leo_sky()
    b = band(-15.0)
    b.noise_power_w = 0.0
    scene = SceneConfig(SceneMeta(START_UNIX, 5.0), list(sky.transmitters),
                        [ReceiverConfig("orbiter", orbiter(START_UNIX), NodeClockSpec(), [b])])
    sc = prepare(scene)
    seed = 5
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    na = len(sc.receivers[0].bands[0].active)
    flags = np.zeros(na, dtype=np.uint8)
    o, _, taps = orc.render_band(0, tap_slot=0, unoccluded=flags)
    n = sc.receivers[0].bands[0].n
    gflags = np.zeros(na, dtype=np.uint8)
    g = np.empty(n, dtype=np.complex64)
    j, keep = backend.job(0, 0, 0, seed, 0, n, out_host=g, delta_t=taps[3], unoccluded=gflags)
    backend.render([j], _lib.FORM_PLAIN)
    backend.sync()
    compare_iq(g, o, "orbiting receiver with antenna, oracle clock offsets", strict=False)
    assert list(gflags != 0) == list(flags != 0)
    assert 0 < int(flags.sum()) < na


def test_truth_export_of_a_moving_receiver(backend, tmp_path, orc):
    """export_truth (src/pipeline.rs:547-611) for the orbiter: every range is
    the oracle's 4-iteration solve against the receiver's position at that
    grid time."""
    scene = moving_scene(duration_s=2.0)
    output.export_truth(scene, str(tmp_path), 10.0, backend=backend)
    truth = json.load(open(str(tmp_path / "orbiter__truth.json")))
    sc = prepare(scene)
    osc = OracleScene(sc, 0)
    epoch = scene.scene.epoch_unix
    t_rel = [k * 0.1 for k in range(21)]
    for i, tx in enumerate(truth["transmitters"]):
        want = []
        for t in t_rel:
            obs = (C.c_double * 3)()
            orc.orc_rx_position(C.byref(osc.rxs[0]), epoch + t, obs)
            te, r = C.c_double(), C.c_double()
            orc.orc_solve_emission(C.byref(osc.txs[i]), obs, t, epoch, 4, C.byref(te), C.byref(r))
            want.append(r.value)
        assert tx["range_m"] == want, tx["name"]
    _say(f"moving-receiver truth: {len(truth['transmitters'])} transmitters x 21 grid times equal the oracle's")
