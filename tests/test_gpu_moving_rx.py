"""Moving receivers on the GPU path: a CircularKeplerian receiver whose
position the kernels evaluate at every sample's epoch + t_utc, as the chunk
body queries position_ecef_batch (src/pipeline.rs:929-930). Against the FP64
oracle, which takes the same per-sample positions (glibc's sincos there, the
device's f64 sincos here: positions within an ulp, so the IQ is ulp-bounded
rather than bit-identical); the precision form, whose fitted solve adds the
receiver query's rounding per sample, within its bounds of the oracle for
orbit epochs at, 40 days before and 1.7e9 s before the scene's (where the
reference's own rounding of the anomaly and Earth angle moves the receiver by
up to millimetres from sample to sample); the orbiting-receiver horizon-gate
test of the reference (tests/horizon_gate.rs:345-370) on the GPU path in the
plain and precision forms; the learned form, which needs a fixed receiver,
refuses one that moves; and the truth export.
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

# The precision form's antenna kernel with a moving receiver ran for minutes
# without finishing on its first GPU run (this orbiter under the 19-LEO sky,
# 1 kHz); no cause was found by reading its code, so the library refuses the
# combination and these cases stay off the GPU.
ANT_PRECISION_MOVING = pytest.mark.skip(
    reason="the precision form refuses an antenna band with a moving receiver: its first GPU run did not "
           "finish and no cause has been found")


def _say(*a):
    print(*a, flush=True)


def orbiter(epoch):
    """horizon_gate.rs:349-356: a receiver 400 km up."""
    return CircularKeplerian(6_778_137.0, 70.0, 320.0, 50.0, epoch)


def moving_scene(duration_s=0.05, noise_power_w=0.0, orbit_epoch=None):
    """Two C/A MEOs, a BPSK LEO and a ground CW tone seen from a LEO receiver
    with a walking clock, one L1 band at 2.048 MHz. orbit_epoch: the
    receiver's element epoch (the scene's by default)."""
    txs = [TransmitterConfig("sv05", 1575.42e6, CaCode(5, util.EPOCH), CircularKeplerian(26.56e6, 55.0, 40.0, 20.0, util.EPOCH), 27.0),
           TransmitterConfig("sv12", 1575.42e6, CaCode(12, util.EPOCH), CircularKeplerian(26.56e6, 55.0, 100.0, 250.0, util.EPOCH), 27.0),
           TransmitterConfig("leo", 1575.42e6 - 5e3, Bpsk(9600.0, 480, 7, [1, -1, 1, 1]),
                             CircularKeplerian(7.1e6, 60.0, 330.0, 45.0, util.EPOCH), -10.0),
           TransmitterConfig("tone", 1575.42e6 + 3e3, Cw(250.0), StationaryGeodetic(40.0, 20.0, 100.0), -20.0)]
    band_ = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=noise_power_w)
    rx = ReceiverConfig("orbiter", orbiter(util.EPOCH if orbit_epoch is None else orbit_epoch),
                        NodeClockSpec(allan_deviation_1s=1e-11, frequency_offset_y0=2e-7), [band_])
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


# orbit element epochs: the scene's; 40 days earlier; the Unix epoch (dt ~ 1.7e9 s)
ORBIT_EPOCHS = {"scene epoch": lambda: util.EPOCH, "40 days before": lambda: util.EPOCH - 40 * 86400.0,
                "unix epoch": lambda: 0.0}


@pytest.mark.parametrize("which", list(ORBIT_EPOCHS))
def test_moving_receiver_precision_form_against_oracle(backend, which):
    """The precision form for the orbiter: IQ within 2e-6 of the band's peak
    of the oracle's, with the oracle's clock offsets and with the device's;
    every transmitter's tapped phase within 1e-6 rad and delay within 1 ps.
    The plain form (per-sample positions) is checked beside it."""
    sc = prepare(moving_scene(orbit_epoch=ORBIT_EPOCHS[which]()))
    seed = 41
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, taps = orc.render_band(0, tap_slot=0)
    peak = float(np.max(np.abs(o)))
    for label, dt in (("oracle clock offsets", taps[3]), ("device clock", None)):
        g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=dt, chunk=30_000, form=_lib.FORM_PRECISION)
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        _say(f"moving receiver ({which}), precision, {label}: n={len(o)} max|d|/peak={d.max() / peak:.3e}")
        assert np.all(np.isfinite(g))
        assert d.max() <= 2e-6 * peak
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3], chunk=30_000)
    compare_iq(g, o, f"moving receiver ({which}), plain, oracle clock offsets", strict=False)
    for slot in range(4):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 40_000, form=_lib.FORM_PRECISION,
                     label=f"moving receiver ({which}), precision")


@pytest.mark.parametrize("fs", [1_000.0, 48_000.0])
def test_moving_receiver_low_rate_band_precision(backend, fs):
    """A ground tone and a LEO BPSK seen from the orbiter in a UHF band at
    1 kHz (tiles of half a second, over which the quadratic would miss the
    receiver's orbit by tens of microradians of phase, so the fit's
    truncation bound sends them to the exact solve) and at 48 kHz: IQ within
    2e-6 of the peak, taps within 1e-6 rad and 1 ps."""
    txs = [TransmitterConfig("tone", 437e6 + 100.0, Cw(25.0), StationaryGeodetic(40.0, 20.0, 100.0), 10.0),
           TransmitterConfig("leo", 437e6 - 200.0, Bpsk(50.0, 480, 7, [1, -1, 1, 1]),
                             CircularKeplerian(7.1e6, 60.0, 330.0, 45.0, util.EPOCH), 0.0)]
    band_ = ReceiverBandConfig("uhf", 437e6, fs, noise_power_w=0.0)
    rx = ReceiverConfig("orbiter", orbiter(util.EPOCH), NodeClockSpec(allan_deviation_1s=1e-11), [band_])
    sc = prepare(SceneConfig(SceneMeta(util.EPOCH, 20.0 if fs < 10e3 else 1.0), txs, [rx]))
    seed = 9
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, taps = orc.render_band(0, tap_slot=0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3], form=_lib.FORM_PRECISION)
    peak = float(np.max(np.abs(o)))
    d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
    _say(f"orbiter, {fs:.0f} Hz band, precision: n={len(o)} max|d|/peak={d.max() / peak:.3e}")
    assert d.max() <= 2e-6 * peak
    for slot in range(2):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, sc.receivers[0].bands[0].n, form=_lib.FORM_PRECISION,
                     label=f"orbiter, {fs:.0f} Hz band")


def test_precision_call_with_moving_and_fixed_receivers(backend):
    """One precision render call holding the orbiter's band and a ground
    receiver's band (the moving-receiver kernels take the fixed receiver with
    zero departures), and a streaming plan of both; each within 2e-6 of its
    band's peak of the oracle's."""
    scene = moving_scene(duration_s=0.03)
    ground = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    scene.receivers.append(util.lab_receiver([ground], NodeClockSpec(allan_deviation_1s=1e-11)))
    sc = prepare(scene)
    seed = 43
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    ref = [orc.render_band(w)[0] for w in range(2)]
    n = [sc.receivers[r].bands[0].n for r in range(2)]
    out = [np.empty(n[r], dtype=np.complex64) for r in range(2)]
    chunk = 20_000
    for i0 in range(0, max(n), chunk):
        jobs, keep = [], []
        for r in range(2):
            if i0 < n[r]:
                clen = min(chunk, n[r] - i0)
                j, k = backend.job(r, 0, r, seed, i0, clen, out_host=out[r][i0:i0 + clen])
                jobs.append(j)
                keep.append(k)
        backend.render(jobs, _lib.FORM_PRECISION)
    backend.sync()
    for r in range(2):
        peak = float(np.max(np.abs(ref[r])))
        d = np.abs(out[r].astype(np.complex128) - ref[r].astype(np.complex128))
        _say(f"mixed call, receiver {r}: n={n[r]} max|d|/peak={d.max() / peak:.3e}")
        assert d.max() <= 2e-6 * peak
    # a streaming plan of both bands (the moving-receiver kernels in its steps)
    lib = _lib.load()
    bands, hold = [], []
    for r in range(2):
        j, k = backend.job(r, 0, r, seed, 0, n[r])
        bands.append(j)
        hold.append(k)
    with backend.plan(bands, 8192, _lib.FORM_PRECISION) as plan:
        got = [np.empty(n[r], dtype=np.complex64) for r in range(2)]
        for kk in range(plan.n_chunks):
            plan.step(kk)
            backend.sync()
            for r in range(2):
                i0, clen = plan.chunk_range(r, kk)
                if clen:
                    backend.check(lib.lpk_copy_to_host(backend.h, got[r][i0:].ctypes.data, plan.output(r), clen * 8))
    for r in range(2):
        peak = float(np.max(np.abs(ref[r])))
        d = np.abs(got[r].astype(np.complex128) - ref[r].astype(np.complex128))
        _say(f"plan, receiver {r}: max|d|/peak={d.max() / peak:.3e}")
        assert d.max() <= 2e-6 * peak


def test_precision_form_refuses_an_antenna_band_with_a_moving_receiver(backend):
    """The refusal comes from the call's validation, before any kernel runs;
    the plain form then renders the same job."""
    sky, _ = leo_sky()
    b = band(-15.0)
    scene = SceneConfig(SceneMeta(START_UNIX, 1.0), list(sky.transmitters),
                        [ReceiverConfig("orbiter", orbiter(START_UNIX), NodeClockSpec(), [b])])
    sc = prepare(scene)
    backend.upload(sc, 3)
    n = sc.receivers[0].bands[0].n
    out = np.empty(n, dtype=np.complex64)
    j, keep = backend.job(0, 0, 0, 3, 0, n, out_host=out)
    with pytest.raises(_lib.LpkError) as e:
        backend.render([j], _lib.FORM_PRECISION)
    assert _lib.STATUS[e.value.code] == "LPK_E_INVALID"
    backend.render([j], _lib.FORM_PLAIN)
    backend.sync()
    assert np.all(np.isfinite(out))


def test_learned_form_refuses_a_moving_receiver(backend):
    form = _lib.FORM_LEARNED
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


@pytest.mark.parametrize("form", ["plain", pytest.param("precision", marks=ANT_PRECISION_MOVING)])
def test_gate_does_not_change_what_an_orbiting_receiver_records(backend, form):
    """horizon_gate.rs:345-370: a receiver in orbit below the 19 LEOs, its
    antenna blind more than 15 deg under its local horizontal; the gated
    window equals the ungated one byte for byte, with the same summary, at
    three chunkings; it sees some of the transmitters and not all."""
    sky, _ = leo_sky()
    rx = ReceiverConfig("orbiter", orbiter(START_UNIX), NodeClockSpec())
    b = band(-15.0)
    ref, rs = collect(backend, False, sky, rx, b, START_UNIX, DURATION_S, 1 << 20, form)
    for chunk in (1 << 20, 300, 5000):
        before = backend.horizon_gate_stats()
        got, gs = collect(backend, True, sky, rx, b, START_UNIX, DURATION_S, chunk, form)
        after = backend.horizon_gate_stats()
        skipped, tested = after["skipped"] - before["skipped"], after["tested"] - before["tested"]
        _say(f"orbiting receiver, {form}, chunks of {chunk}: the gate skipped {skipped} of {tested} (tile, link) pairs")
        assert got.tobytes() == ref.tobytes(), chunk
        assert gs == rs, chunk
        assert tested > 0 and skipped > 0
    visible = len(rs.visible_transmitters)
    _say(f"orbiting receiver: {visible} of 19 visible")
    assert 0 < visible < 19


@pytest.mark.parametrize("form", [_lib.FORM_PLAIN, pytest.param(_lib.FORM_PRECISION, marks=ANT_PRECISION_MOVING)])
def test_orbiting_receiver_with_antenna_against_oracle(backend, form):
    """The same orbiter, antenna and sky over 5 s, noise-free, rendered as a
    scene: the IQ against the oracle's (per-sample positions, arrival
    elevations and cutoff; the precision form within 2e-6 of the peak), and
    the transmitters each says are unoccluded."""
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
    backend.render([j], form)
    backend.sync()
    if form == _lib.FORM_PLAIN:
        compare_iq(g, o, "orbiting receiver with antenna, oracle clock offsets", strict=False)
    else:
        peak = float(np.max(np.abs(o)))
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        _say(f"orbiting receiver with antenna, precision: n={n} max|d|/peak={d.max() / peak:.3e}")
        assert np.all(np.isfinite(g)) and d.max() <= 2e-6 * peak
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
