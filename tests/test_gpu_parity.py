"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle on
identical prepared scenes, plus the reference's end-to-end invariants run on
the GPU path. Tolerances:
  - solved range / emission time / f64 carrier phase (taps): phase within
    1e-6 rad and delay within 1 ps of the oracle (the plain form matches to
    the last bit in practice);
  - noise-free IQ: f32 ulp-bounded (|d| <= 1e-6 of the band's peak), and
    >= 99 % of samples bit-identical;
  - ADC codes: bit-exact with the oracle's clock offsets; on the device clock
    different only where the oracle's pre-quantisation value is within
    1e-4 LSB of a rounding boundary;
  - AWGN: same Philox draws and the same libm-free Box-Muller on both sides,
    so bit-identical given the same clock offsets; statistics against the
    reference's bounds.
"""
import ctypes as C
import math

import numpy as np
import pytest

from lipika_amd import _lib, workloads
from lipika_amd.config import AdcConfig, FrontendConfig, NonlinearityConfig, NodeClockSpec
from lipika_amd.pipeline import DEFAULT_CHUNK_SAMPLES, prepare, run_scene
from oracle.pyoracle import OracleScene

import util

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _say(*a):
    print(*a, flush=True)


def gpu_band(be, sc, w, r, b, seed, chunk=DEFAULT_CHUNK_SAMPLES, n=None, form=_lib.FORM_PLAIN, codes=False,
             delta_t=None):
    """Render unit w (receiver r, band b) of an uploaded scene chunk by chunk;
    `delta_t` (optional) supplies the receiver clock offsets per sample."""
    pb = sc.receivers[r].bands[b]
    n = pb.n if n is None else n
    out = np.empty(n, dtype=np.complex64)
    cd = np.empty(2 * n, dtype=np.int16) if codes else None
    for i0 in range(0, n, chunk):
        clen = min(chunk, n - i0)
        j, keep = be.job(r, b, w, seed, i0, clen, out_host=out[i0:i0 + clen],
                         codes=cd[2 * i0:2 * (i0 + clen)] if codes else None,
                         delta_t=None if delta_t is None else delta_t[i0:i0 + clen])
        be.render([j], form)
    be.sync()
    return out, cd


def check_band_both_clocks(be, orc, sc, w, r, b, seed, label, chunk=DEFAULT_CHUNK_SAMPLES):
    """With the oracle's own clock offsets fed through the ABI the IQ must be
    bit-identical; with the device clock walk (the same normals, summed in a
    fixed tree instead of left to right: offsets equal to ~1e-22 s) it must
    be ulp-bounded and almost everywhere bit-identical."""
    o, _, taps = orc.render_band(w, tap_slot=0)
    g, _ = gpu_band(be, sc, w, r, b, seed, chunk=chunk, delta_t=taps[3])
    assert compare_iq(g, o, f"{label}, oracle clock offsets") == 1.0
    g, _ = gpu_band(be, sc, w, r, b, seed, chunk=chunk)
    compare_iq(g, o, f"{label}, device clock")


def unit_index(sc, r, b):
    w = 0
    for ri, rx in enumerate(sc.receivers):
        for bi, _ in enumerate(rx.bands):
            if (ri, bi) == (r, b):
                return w
            w += 1
    raise KeyError


def noise_free(scene):
    for rx in scene.receivers:
        for b in rx.bands:
            b.noise_power_w = 0.0
    return scene


def compare_iq(gpu, ref, label, strict=True):
    peak = float(np.max(np.abs(ref))) if len(ref) else 0.0
    d = np.abs(gpu.astype(np.complex128) - ref.astype(np.complex128))
    same = np.mean((gpu.view(np.uint32) == ref.view(np.uint32)).reshape(-1, 2).all(axis=1)) if len(ref) else 1.0
    _say(f"{label}: n={len(ref)} peak={peak:.3e} max|d|={d.max() if len(d) else 0:.3e} bit-identical={same:.6f}")
    assert np.all(np.isfinite(gpu))
    assert d.max() <= 1e-6 * peak + 1e-30
    if strict:
        assert same >= 0.99
    return same


def compare_taps(be, orc_scene, sc, w, r, b, seed, slot, n, form=_lib.FORM_PLAIN, label=""):
    _, _, taps = orc_scene.render_band(w, n=n, tap_slot=slot)
    j, keep = be.job(r, b, w, seed, 0, n)
    te, rg, ph, dt = be.taps(j, slot, form)
    dphi = np.max(np.abs(ph - taps[2]))
    ddel = np.max(np.abs(rg - taps[1])) / C_LIGHT
    dte = np.max(np.abs(te - taps[0]))
    ddt = np.max(np.abs(dt - taps[3]))
    _say(f"{label} taps slot {slot}: max|dphi|={dphi:.3e} rad  max|ddelay|={ddel:.3e} s  "
         f"max|dt_emit|={dte:.3e} s  max|d delta_t|={ddt:.3e} s")
    assert dphi <= 1e-6
    assert ddel <= 1e-12
    assert dte <= 1e-12
    # The device walk sums the same normals in a fixed tree (the oracle left to
    # right) and stores each value as an f64 offset per 32 indices plus an f32
    # prefix within those 32: a few sigma_step * sqrt(32) * 2^-24 of error.
    rx = sc.receivers[r]
    sigma_step = rx.clock.allan_deviation_1s / math.sqrt(rx.master_rate_hz) if rx.master_rate_hz > 0 else 0.0
    assert ddt <= 1e-20 + 8.0 * sigma_step * math.sqrt(32.0) * 2.0 ** -24
    return dphi, ddel


# ---------------------------------------------------------------- scenes vs oracle

def test_config1_full_scene_noise_free(backend):
    seed = 2026
    scene = noise_free(workloads.scene1())
    sc = prepare(scene)
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed)
    o, _, _ = orc.render_band(0)
    compare_iq(g, o, "config 1 (10 s CW LEO)")
    compare_taps(backend, orc, sc, 0, 0, 0, seed, 0, 200_000, label="config 1")


def test_config2_prefix_both_bands(backend):
    seed = 2026
    scene = noise_free(workloads.scene2(duration_s=0.3))
    sc = prepare(scene)
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    for b, name in ((0, "UHF BPSK 48 kHz"), (1, "L1 6xC/A 2.048 MHz")):
        check_band_both_clocks(backend, orc, sc, b, 0, b, seed, f"config 2 prefix, {name}", chunk=100_000)
    for slot in range(6):
        compare_taps(backend, orc, sc, 1, 0, 1, seed, slot, 150_000, label="config 2 L1")


def test_all_signal_kinds_with_clock_walk(backend):
    seed = 7
    scene = util.all_kinds_scene(duration_s=0.01)
    sc = prepare(scene)
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    check_band_both_clocks(backend, orc, sc, 0, 0, 0, seed, "all kinds, 8 MHz", chunk=30_000)
    for slot in range(5):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 40_000, label="all kinds")


def test_awgn_matches_oracle_draws(backend):
    seed = 99
    scene = util.all_kinds_scene(duration_s=0.005, noise_power_w=None)
    sc = prepare(scene)
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, taps = orc.render_band(0, tap_slot=0)
    sigma = math.sqrt(sc.receivers[0].bands[0].noise_power_w / 2.0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3])
    same = compare_iq(g, o, f"AWGN sigma={sigma:.3e}, oracle clock offsets")
    assert same == 1.0
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed)
    d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
    _say(f"AWGN, device clock: max|d|={d.max():.3e} ({d.max() / sigma:.2e} sigma)")
    assert d.max() <= 1e-5 * sigma


def test_front_end_ops_and_adc_codes(backend):
    """src/rx/mod.rs:83-101 after the sum: gain, IP3 and an 8-bit ADC. With the
    oracle's clock offsets every code and IQ value is bit-exact; with the
    device clock a code may differ only where the oracle's value before the
    ADC lies at a rounding boundary (its distance to the nearest half-LSB is
    below the clock walk's effect on the sample, 1e-4 LSB)."""
    seed = 3
    fe = FrontendConfig(gain_db=110.0, nonlinearity=NonlinearityConfig(30.0), adc=AdcConfig(8, 1.0))
    scene = util.all_kinds_scene(duration_s=0.004, noise_power_w=None, frontend=fe)
    sc = prepare(scene)
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, oc, taps = orc.render_band(0, codes=True, tap_slot=0)
    lsb = 1.0 / 127.0
    assert oc.min() < -20 and oc.max() > 20          # the ADC is exercised, not idle
    g, gc = gpu_band(backend, sc, 0, 0, 0, seed, codes=True, delta_t=taps[3])
    assert np.array_equal(gc, oc), "ADC codes with the oracle's clock offsets"
    assert g.tobytes() == o.tobytes()
    g, gc = gpu_band(backend, sc, 0, 0, 0, seed, codes=True)
    mism = np.nonzero(gc != oc)[0]
    _say(f"ADC: {len(mism)} of {len(oc)} codes differ on the device clock; codes span [{oc.min()}, {oc.max()}]")
    if len(mism):
        # the oracle's value before the ADC: the same scene with gain and IP3 only
        pre_sc = prepare(util.all_kinds_scene(duration_s=0.004, noise_power_w=None,
                                              frontend=FrontendConfig(gain_db=110.0,
                                                                      nonlinearity=NonlinearityConfig(30.0))))
        pre, _, _ = OracleScene(pre_sc, seed).render_band(0)
        x = np.clip(pre.view(np.float32)[mism].astype(np.float64), -1.0, 1.0) / lsb
        dist = np.abs(np.abs(x - np.floor(x)) - 0.5)
        _say(f"ADC mismatches: distance of the pre-ADC value to the rounding boundary <= {dist.max():.2e} LSB")
        assert np.all(dist < 1e-4)
        assert np.all(np.abs(gc[mism].astype(int) - oc[mism].astype(int)) == 1)
    flat_o = o.view(np.float32)
    assert np.all(np.abs(flat_o) <= 1.0 + 1e-6)
    assert np.all(np.abs(np.round(flat_o / lsb) * lsb - flat_o) < 1e-6)


def test_adc_wider_than_int16(backend):
    """A 24-bit ADC (src/rx/mod.rs:93-97 takes any width): the quantised IQ is
    bit-identical to the oracle's; int16 codes are refused for it, and the
    refusal leaves the context usable."""
    seed = 8
    fe = FrontendConfig(gain_db=110.0, adc=AdcConfig(24, 1.0))
    sc = prepare(util.all_kinds_scene(duration_s=0.002, noise_power_w=None, frontend=fe))
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, taps = orc.render_band(0, tap_slot=0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3])
    assert g.tobytes() == o.tobytes()
    lsb = 1.0 / float((1 << 23) - 1)
    flat = o.view(np.float32).astype(np.float64)
    assert np.all(np.abs(np.round(flat / lsb) * lsb - flat) <= 1e-7 * np.maximum(np.abs(flat), lsb))
    with pytest.raises(_lib.LpkError) as e:
        gpu_band(backend, sc, 0, 0, 0, seed, codes=True, n=1000)
    assert _lib.STATUS[e.value.code] == "LPK_E_INVALID" and "16 bits" in str(e.value)
    g2, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3])
    assert g2.tobytes() == o.tobytes()


# ---------------------------------------------------------------- precision form

def _ism_meo_scene(duration_s=0.05):
    """Four MEO C/A transmitters at 2.4 GHz: carrier phases near 1.4e9 rad,
    beyond the 2^30 small-argument range of the f64 sincos."""
    from lipika_amd.config import CaCode, CircularKeplerian, ReceiverBandConfig, SceneConfig, SceneMeta, TransmitterConfig
    txs = [TransmitterConfig(f"m{k}", 2.4e9, CaCode(k + 1, util.EPOCH),
                             CircularKeplerian(2.66e7, 55.0, 90.0 * k, 20.0 + 70.0 * k, util.EPOCH), 27.0)
           for k in range(4)]
    band = ReceiverBandConfig("ism", 2.4e9, 8.184e6, noise_power_w=0.0)
    return SceneConfig(SceneMeta(util.EPOCH, duration_s), txs,
                       [util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-9, frequency_offset_y0=-8e-7))])


PRECISION_SCENES = {
    "config1 LEO 48 kHz": lambda: noise_free(workloads.scene1(duration_s=2.0)),
    "config2 both bands": lambda: noise_free(workloads.scene2(duration_s=0.1)),
    "all kinds 8 MHz": lambda: util.all_kinds_scene(duration_s=0.005),
    "MEO 2.4 GHz 8.184 MHz": _ism_meo_scene,
}


@pytest.mark.parametrize("name", list(PRECISION_SCENES))
def test_precision_form_error_vs_fp64_oracle(backend, name):
    """The precision form against the FP64 oracle: carrier phase within 1e-6
    rad and delay within 1 ps for every transmitter; IQ within the phasor
    displacement those bounds allow (2e-6 of the band's peak)."""
    seed = 5
    sc = prepare(PRECISION_SCENES[name]())
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    w = 0
    for r, rx in enumerate(sc.receivers):
        for b, pb in enumerate(rx.bands):
            g, _ = gpu_band(backend, sc, w, r, b, seed, form=_lib.FORM_PRECISION, chunk=300_000)
            o, _, _ = orc.render_band(w)
            peak = float(np.max(np.abs(o)))
            d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
            _say(f"precision form, {name}, band {b}: n={len(o)} max|d|/peak={d.max() / peak:.3e}")
            assert np.all(np.isfinite(g))
            assert d.max() <= 2e-6 * peak
            if len(pb.active) == 1:
                # one transmitter: the output's own carrier phase against the
                # oracle's, sample by sample (both rounded to complex64)
                on = np.abs(o) >= 0.1 * peak
                dphi = np.abs(np.angle(g[on].astype(np.complex128) * np.conj(o[on].astype(np.complex128))))
                _say(f"precision form, {name}, band {b}: output phase max|dphi|={dphi.max():.3e} rad "
                     f"over {int(on.sum())} samples")
                assert on.sum() > 0.5 * len(o) and dphi.max() <= 1e-6
            n = min(pb.n, 120_000)
            for slot in range(len(pb.active)):
                compare_taps(backend, orc, sc, w, r, b, seed, slot, n, form=_lib.FORM_PRECISION,
                             label=f"precision {name} band {b}")
            w += 1


def _antenna_scene():
    """A Glasgow band with an antenna whose cutoff a LEO crosses mid-window,
    a source 500 km overhead and one on the far side of the Earth."""
    from lipika_amd.config import (Bpsk, CircularKeplerian, Cw, FrontendConfig, ReceiverBandConfig, SceneConfig,
                                   SceneMeta, StationaryGeodetic, TransmitterConfig)
    from oracle import pyoracle
    orc = pyoracle.load()
    obs = (C.c_double * 3)()
    orc.orc_geodetic_to_ecef(55.86, -4.25, 10.0, obs)

    def elevation(orbit, t):
        p = (C.c_double * 3)()
        orc.orc_kepler_position(*orbit, util.EPOCH, t, p)
        return orc.orc_arrival_elevation_deg(obs, p)

    best = None
    for raan in range(0, 360, 15):
        for nu in range(0, 360, 3):
            orbit = (6.9e6, 60.0, float(raan), float(nu))
            e0, e1 = elevation(orbit, util.EPOCH), elevation(orbit, util.EPOCH + 0.2)
            if 10.0 < e0 < 60.0 and abs(e1 - e0) > 0.05 and (best is None or abs(e1 - e0) > best[1]):
                best = (orbit, abs(e1 - e0))
    orbit = best[0]
    cutoff = elevation(orbit, util.EPOCH + 0.1)
    e0, e1 = elevation(orbit, util.EPOCH), elevation(orbit, util.EPOCH + 0.2)
    assert min(e0, e1) < cutoff < max(e0, e1)  # the LEO crosses the cutoff inside the window
    sync = [-1, 1, 1, -1, 1, -1, -1, 1]
    txs = [TransmitterConfig("leo", 437e6, Bpsk(9600.0, 2000, 3, sync), CircularKeplerian(*orbit, util.EPOCH), 0.0),
           TransmitterConfig("overhead", 437.001e6, Cw(0.0), StationaryGeodetic(55.86, -4.25, 5.0e5), -10.0),
           TransmitterConfig("antipode", 436.999e6, Cw(0.0), StationaryGeodetic(-55.86, 175.75, 5.0e5), 30.0)]
    fe = FrontendConfig(antenna={"peak_dbi": 3.9, "horizon_dbi": -3.0, "fb_db": 8.5, "cutoff_deg": cutoff})
    band = ReceiverBandConfig("uhf", 437e6, 250e3, 290.0, noise_power_w=0.0, frontend=fe)
    return SceneConfig(SceneMeta(util.EPOCH, 0.2), txs, [util.lab_receiver([band])])


@pytest.mark.parametrize("form", [_lib.FORM_PLAIN, _lib.FORM_PRECISION])
def test_antenna_gain_and_occlusion(backend, form):
    """src/channel/mod.rs:190-211 on the GPU: the gain at each sample's arrival
    elevation, samples below the cutoff dropped, and the per-transmitter
    unoccluded flags, against the oracle (src/rx/mod.rs:48-58)."""
    seed = 12
    sc = prepare(_antenna_scene())
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    want_flags = np.zeros(3, dtype=np.uint8)
    o, _, _ = orc.render_band(0, unoccluded=want_flags)
    assert list(want_flags) == [1, 1, 0]
    n = sc.receivers[0].bands[0].n
    g = np.empty(n, dtype=np.complex64)
    flags = np.zeros(3, dtype=np.uint8)
    chunk = 20_000
    for i0 in range(0, n, chunk):
        f = np.zeros(3, dtype=np.uint8)
        j, keep = backend.job(0, 0, 0, seed, i0, min(chunk, n - i0), out_host=g[i0:i0 + chunk], unoccluded=f)
        backend.render([j], form)
        flags |= f
    backend.sync()
    assert list(flags) == [1, 1, 0]
    label = f"antenna, {'plain' if form == _lib.FORM_PLAIN else 'precision'}"
    if form == _lib.FORM_PLAIN:
        compare_iq(g, o, label)
    else:
        peak = float(np.max(np.abs(o)))
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        _say(f"{label}: n={n} max|d|/peak={d.max() / peak:.3e}")
        assert d.max() <= 2e-6 * peak
    for slot in range(3):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 20_000, form=form, label=label)


def test_plain_form_large_phase_scene(backend):
    """Phases beyond 2^30 rad take the f64 sincos's large-argument path."""
    seed = 8
    sc = prepare(_ism_meo_scene(0.01))
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    check_band_both_clocks(backend, orc, sc, 0, 0, 0, seed, "plain, MEO 2.4 GHz")
    for slot in range(4):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 40_000, label="plain MEO 2.4 GHz")


# ---------------------------------------------------------------- reference invariants on the GPU

def test_chunk_size_invariance_bytes(backend):
    """tests/streaming_chunk_invariance.rs:78-149 on the GPU path."""
    scene = workloads.scene2(duration_s=0.25)
    a = run_scene(scene, 11, chunk_samples=DEFAULT_CHUNK_SAMPLES, backend=backend)
    for chunk in (65_536, 12_345):
        b = run_scene(scene, 11, chunk_samples=chunk, backend=backend)
        for x, y in zip(a, b):
            assert x.iq.tobytes() == y.iq.tobytes(), (x.name, chunk)


def test_additivity_both_equals_a_plus_b_minus_noise(backend, kats):
    """tests/multi_transmitter_summation.rs:188-236"""
    seed = 99
    both = run_scene(util.two_tx_scene(("alpha", "beta")), seed, backend=backend)[0].iq
    a = run_scene(util.two_tx_scene(("alpha",)), seed, backend=backend)[0].iq
    b = run_scene(util.two_tx_scene(("beta",)), seed, backend=backend)[0].iq
    n = run_scene(util.two_tx_scene(()), seed, backend=backend)[0].iq
    err = np.max(np.abs(both - (a + b - n)))
    _say(f"additivity: max|both - (a+b-n)| = {err:.3e}, mean|both| = {np.mean(np.abs(both)):.3e}")
    assert err < kats["additivity"]["max_abs_err"]


def test_clock_offset_130ns_recovered(backend, kats):
    """tests/multi_transmitter_summation.rs:134-156"""
    outs = {o.name: o.iq for o in run_scene(util.cw_two_clocks_scene(), 7, backend=backend)}
    a = outs["rxb__b"].astype(np.complex128)
    b = outs["rxa__b"].astype(np.complex128)
    acc = np.sum(a * np.conj(b))
    rec = -math.atan2(acc.imag, acc.real) / (2 * math.pi * 1e6)
    _say(f"recovered clock offset {rec * 1e9:.4f} ns")
    assert abs(rec - 1.3e-7) < kats["clock_offset_recovery"]["tolerance_s"]


def test_noise_only_band_statistics(backend, kats):
    """tests/multi_transmitter_summation.rs:175-186, tests/window_render.rs:431-458"""
    k = kats["noise_power"]
    iq = run_scene(util.two_tx_scene(()), 12345, backend=backend)[0].iq.astype(np.complex128)
    p = np.mean(np.abs(iq) ** 2)
    sigma = math.sqrt(k["power_w"] / 2)
    lag1 = np.real(np.mean(iq[1:] * np.conj(iq[:-1]))) / p
    _say(f"noise: power/1e-15={p / k['power_w']:.4f} |mean|/sigma={abs(np.mean(iq.real)) / sigma:.4f} lag1={lag1:.4f}")
    assert abs(p / k["power_w"] - 1) < k["tolerance"]
    assert abs(np.mean(iq.real)) < k["mean_sigma"] * sigma and abs(np.mean(iq.imag)) < k["mean_sigma"] * sigma
    assert abs(lag1) < k["lag1"]


def test_full_config2_scale_properties(backend):
    """Config 2 at its full size (60 s, 125.76 M output samples) on the GPU:
    determinism, chunk invariance through the FNV-1a sink, per-band power."""
    from lipika_amd import workloads as wl
    seed = 2026
    scene = wl.scene2()
    sc = prepare(scene)
    backend.upload(sc, seed)
    lib = _lib.load()
    import ctypes as C
    hashes = []
    for chunk in (DEFAULT_CHUNK_SAMPLES, 3 << 18, DEFAULT_CHUNK_SAMPLES):
        hs = []
        for w, (r, b) in enumerate([(0, 0), (0, 1)]):
            n = sc.receivers[r].bands[b].n
            dev = C.c_void_p()
            backend.check(lib.lpk_device_alloc(backend.h, n * 8, C.byref(dev)))
            jobs, keep = [], []
            for i0 in range(0, n, chunk):
                j, k = backend.job(r, b, w, seed, i0, min(chunk, n - i0), out_dev=dev.value + i0 * 8)
                jobs.append(j)
                keep.append(k)
            backend.render(jobs)
            h = C.c_uint64()
            backend.check(lib.lpk_device_fnv1a(backend.h, dev, n * 8, C.byref(h)))
            if chunk == DEFAULT_CHUNK_SAMPLES and len(hashes) == 0:
                tail = np.empty(1 << 16, dtype=np.complex64)
                backend.check(lib.lpk_copy_to_host(backend.h, tail.ctypes.data, dev.value + (n - len(tail)) * 8,
                                                   len(tail) * 8))
                p = np.mean(np.abs(tail.astype(np.complex128)) ** 2)
                pn = sc.receivers[r].bands[b].noise_power_w
                _say(f"band {b}: tail power / kTB = {p / pn:.4f}")
                assert np.all(np.isfinite(tail)) and 0.9 < p / pn < 3.0
            backend.check(lib.lpk_device_free(backend.h, dev))
            hs.append(h.value)
        hashes.append(hs)
        _say(f"chunk {chunk}: fnv1a {[hex(x) for x in hs]}")
    assert hashes[0] == hashes[1] == hashes[2]


def _units(sc):
    return [(r, b) for r, rx in enumerate(sc.receivers) for b, _ in enumerate(rx.bands)]


@pytest.mark.parametrize("form", [_lib.FORM_PLAIN, _lib.FORM_PRECISION])
def test_streaming_plan_replays_match_eager_chunks(backend, form):
    """Every step of a plan (alternating between its two streams) equals an
    eager render of the same chunk of every band (clock walk and kTB noise
    on), with eager renders and synchronisations between the steps;
    restarting at chunk 0 renders the same bytes again. Steps are time slices: the fastest band's
    chunks hold chunk_samples samples, the others' the samples of the same
    interval, and together they tile every band exactly once."""
    seed = 21
    sc = prepare(workloads.scene2(duration_s=0.2))
    backend.upload(sc, seed)
    lib = _lib.load()
    chunk = 100_000
    units = _units(sc)
    bands, keep = [], []
    for w, (r, b) in enumerate(units):
        j, k = backend.job(r, b, w, seed, 0, sc.receivers[r].bands[b].n)
        bands.append(j)
        keep.append(k)
    checked = 0
    with backend.plan(bands, chunk, form) as plan:
        _say(f"plan: {plan.n_chunks} chunks of {chunk}, at most {plan.max_tiles} tiles per step")
        fs = [sc.receivers[r].bands[b].band.sample_rate_hz for r, b in units]
        for w, (r, b) in enumerate(units):
            n = sc.receivers[r].bands[b].n
            ranges = [plan.chunk_range(w, kk) for kk in range(plan.n_chunks + 1)]
            assert ranges[0][0] == 0 and ranges[-1] == (n, 0)
            assert all(ranges[q][0] + ranges[q][1] == ranges[q + 1][0] for q in range(plan.n_chunks))
            if fs[w] == max(fs):
                assert all(c == chunk for _, c in ranges[:-2])
            for q, (i0, c) in enumerate(ranges[:-1]):   # samples j / fs of step q lie in [q T, (q + 1) T)
                if c:
                    assert i0 / fs[w] >= q * chunk / max(fs) - 1e-12 and (i0 + c - 1) / fs[w] < (q + 1) * chunk / max(fs)
        for rnd, ks in enumerate((range(plan.n_chunks), range(2))):
            for kk in ks:
                plan.step(kk)
                backend.sync()
                got, jobs, want = [], [], []
                for w, (r, b) in enumerate(units):
                    i0, clen = plan.chunk_range(w, kk)
                    if clen == 0:
                        continue
                    g = np.empty(clen, dtype=np.complex64)
                    backend.check(lib.lpk_copy_to_host(backend.h, g.ctypes.data, plan.output(w), clen * 8))
                    o = np.empty(clen, dtype=np.complex64)
                    j, k = backend.job(r, b, w, seed, i0, clen, out_host=o)
                    keep.append(k)
                    jobs.append(j)
                    got.append(g)
                    want.append(o)
                backend.render(jobs, form)  # the eager render of the same chunk, between replays
                backend.sync()
                for g, o in zip(got, want):
                    assert g.tobytes() == o.tobytes(), (rnd, kk)
                    checked += 1
    _say(f"plan replays: {checked} band-chunks bit-identical to eager renders")


def test_two_plans_stepped_in_turn(backend):
    """Two live plans on one context (the plain form over one band, the
    precision form over the other, both on the same receiver's clock walk)
    stepped in turn without synchronising in between: each plan's chunks
    equal eager renders of the same chunks made afterwards, so the plans'
    second streams and the shared walk tables stay ordered."""
    seed = 23
    sc = prepare(workloads.scene2(duration_s=0.1))
    backend.upload(sc, seed)
    lib = _lib.load()
    units = _units(sc)
    assert len(units) == 2
    chunk = 40_000
    forms = (_lib.FORM_PLAIN, _lib.FORM_PRECISION)
    plans, outs, keep = [], [], []
    for w, (r, b) in enumerate(units):
        n = sc.receivers[r].bands[b].n
        buf = C.c_void_p()
        backend.check(lib.lpk_device_alloc(backend.h, n * 8, C.byref(buf)))
        outs.append((buf, n))
        j, k = backend.job(r, b, w, seed, 0, n, out_dev=buf.value)
        keep.append(k)
        plans.append(backend.plan([j], chunk, forms[w]))
    try:
        steps = max(p.n_chunks for p in plans)
        for rnd in range(2):   # the second pass restarts both plans at chunk 0
            for kk in range(steps):
                for p in plans:
                    if kk < p.n_chunks:
                        p.step(kk)
            backend.sync()
        for w, ((r, b), (buf, n)) in enumerate(zip(units, outs)):
            got = np.empty(n, dtype=np.complex64)
            backend.check(lib.lpk_copy_to_host(backend.h, got.ctypes.data, buf.value, n * 8))
            ranges = [plans[w].chunk_range(0, kk) for kk in range(plans[w].n_chunks)]
            want = np.empty(n, dtype=np.complex64)
            for i0, clen in ranges:
                if clen:
                    j, k = backend.job(r, b, w, seed, i0, clen, out_host=want[i0:i0 + clen])
                    backend.render([j], forms[w])
            backend.sync()
            assert got.tobytes() == want.tobytes(), w
        _say(f"two plans in turn: {steps} steps each, both bit-identical to eager renders")
    finally:
        for p in plans:
            p.close()
        for buf, _ in outs:
            backend.check(lib.lpk_device_free(backend.h, buf))


def test_tuning_and_autotune(backend):
    """The plain form's samples do not depend on the tile size or the XCD
    mapping; the precision form's stay within its bound of the plain form's.
    The autotuner keeps one of its candidates on the context."""
    seed = 5
    sc = prepare(util.all_kinds_scene(duration_s=0.004))
    backend.upload(sc, seed)
    try:
        backend.set_tuning(8, 1)
        ref, _ = gpu_band(backend, sc, 0, 0, 0, seed)
        peak = float(np.max(np.abs(ref)))
        for ts in (7, 8, 9):
            for xcd in (0, 1):
                backend.set_tuning(ts, xcd)
                assert backend.tuning() == (ts, xcd)
                g, _ = gpu_band(backend, sc, 0, 0, 0, seed)
                assert g.tobytes() == ref.tobytes(), (ts, xcd)
                p, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION)
                d = np.max(np.abs(p.astype(np.complex128) - ref.astype(np.complex128)))
                _say(f"tile 2^{ts}, xcd map {xcd}: precision max|d|/peak = {d / peak:.3e}")
                assert d <= 2e-6 * peak
        lib = _lib.load()
        n = sc.receivers[0].bands[0].n
        dev = C.c_void_p()
        backend.check(lib.lpk_device_alloc(backend.h, n * 8, C.byref(dev)))
        try:
            j, keep = backend.job(0, 0, 0, seed, 0, n, out_dev=dev.value)
            best, ms = backend.autotune([j], _lib.FORM_PRECISION)
            _say(f"autotune: tile 2^{best[0]}, xcd map {best[1]}, {ms:.3f} ms")
            assert best[0] in (7, 8, 9) and best[1] in (0, 1) and ms > 0.0
            assert backend.tuning() == best
            # job sets the render refuses, or that hold no sample, are refused
            # before any candidate runs, and the tuning stays as it was
            backend.set_tuning(7, 1)
            empty, keep_e = backend.job(0, 0, 0, seed, 0, 0, out_dev=dev.value)
            with pytest.raises(_lib.LpkError) as e:
                backend.autotune([empty], _lib.FORM_PRECISION)
            assert _lib.STATUS[e.value.code] == "LPK_E_INVALID"
            assert backend.tuning() == (7, 1)
        finally:
            backend.check(lib.lpk_device_free(backend.h, dev))
        # a job whose output goes to host memory is refused the same way
        host_out = np.empty(1000, dtype=np.complex64)
        j, keep = backend.job(0, 0, 0, seed, 0, 1000, out_host=host_out)
        with pytest.raises(_lib.LpkError) as e:
            backend.autotune([j], _lib.FORM_PRECISION)
        assert _lib.STATUS[e.value.code] == "LPK_E_INVALID" and "device memory" in str(e.value)
        assert backend.tuning() == (7, 1)
    finally:
        backend.set_tuning(8, 1)


@pytest.mark.parametrize("form", [_lib.FORM_PLAIN, _lib.FORM_PRECISION])
def test_render_call_larger_than_one_launch(backend, form):
    """One render call of 2.3e9 samples (8.98 M tiles of 256, three
    dispatches): chunks at the start, across both slice boundaries and at the
    very end equal the same chunks rendered on their own."""
    import ctypes as C
    from lipika_amd.config import ReceiverBandConfig, SceneConfig, SceneMeta, StationaryEcef, TransmitterConfig, Cw
    seed = 4
    tx = TransmitterConfig("cw", 1.0e9, Cw(1.25e6), StationaryEcef(6378137.0, 0.0, 300000.0), 30.0)
    band = ReceiverBandConfig("wide", 1.0e9, 10.0e6, noise_power_w=0.0)
    sc = prepare(SceneConfig(SceneMeta(util.EPOCH, 230.0), [tx], [util.lab_receiver([band])]))
    n = sc.receivers[0].bands[0].n
    assert n // 256 > 1 << 23
    backend.upload(sc, seed)
    lib = _lib.load()
    chunk = DEFAULT_CHUNK_SAMPLES
    dev = C.c_void_p()
    backend.check(lib.lpk_device_alloc(backend.h, n * 8, C.byref(dev)))
    try:
        jobs, keep = [], []
        for i0 in range(0, n, chunk):
            j, k = backend.job(0, 0, 0, seed, i0, min(chunk, n - i0), out_dev=dev.value + i0 * 8)
            jobs.append(j)
            keep.append(k)
        _say(f"one call: {len(jobs)} jobs, {n} samples, {(n + 255) // 256} tiles")
        backend.render(jobs, form)
        backend.sync()
        # dispatches hold at most 2^22 tiles: check the chunks around each slice edge
        edges = [(m * (1 << 22) * 256) // chunk for m in (1, 2)]
        for q in sorted({0, len(jobs) - 1, *[e + d for e in edges for d in (-1, 0, 1)]}):
            i0 = q * chunk
            clen = min(chunk, n - i0)
            got = np.empty(clen, dtype=np.complex64)
            backend.check(lib.lpk_copy_to_host(backend.h, got.ctypes.data, dev.value + i0 * 8, clen * 8))
            want = np.empty(clen, dtype=np.complex64)
            j, k = backend.job(0, 0, 0, seed, i0, clen, out_host=want)
            backend.render([j], form)
            backend.sync()
            _say(f"chunk {q}: max|want| = {np.max(np.abs(want)):.3e}")
            assert np.max(np.abs(want)) > 0.0
            assert got.tobytes() == want.tobytes(), q
    finally:
        backend.check(lib.lpk_device_free(backend.h, dev))


# ---------------------------------------------------------------- error paths (forced from the host)

def test_errors_are_codes_not_crashes(backend):
    import ctypes as C
    scene = workloads.scene1(duration_s=0.01)
    sc = prepare(scene)
    backend.upload(sc, 1)
    j, keep = backend.job(0, 0, 0, 1, 0, 100, out_host=np.empty(100, dtype=np.complex64))
    bad = np.array([5], dtype=np.uint32)
    j.active = bad.ctypes.data_as(C.POINTER(C.c_uint32))
    with pytest.raises(_lib.LpkError) as e:
        backend.render([j])
    assert e.value.code == -1
    j2, keep2 = backend.job(0, 0, 0, 1, 0, 4, out_host=np.empty(4, dtype=np.complex64),
                            delta_t=np.array([0.0, np.nan, 0.0, 0.0]))
    with pytest.raises(_lib.LpkError) as e:
        backend.render([j2])
    assert e.value.code == -4
    j3, keep3 = backend.job(0, 0, 0, 1, 0, 4, out_host=np.empty(4, dtype=np.complex64))
    j3.fs_hz = 0.0
    with pytest.raises(_lib.LpkError):
        backend.render([j3])
    # the context still works afterwards
    j4, keep4 = backend.job(0, 0, 0, 1, 0, 100, out_host=np.empty(100, dtype=np.complex64))
    backend.render([j4])
    backend.sync()


def test_host_delta_t_path_equals_device_clock(backend):
    """A caller-supplied clock offset array (the reference's own ClockStream)
    gives the same IQ as the device walk when it holds the same values."""
    seed = 21
    scene = workloads.scene2(duration_s=0.02)
    sc = prepare(scene)
    backend.upload(sc, seed)
    n = sc.receivers[0].bands[1].n
    j, keep = backend.job(0, 1, 1, seed, 0, n)
    _, _, _, dt = backend.taps(j, 0)
    a = np.empty(n, dtype=np.complex64)
    b = np.empty(n, dtype=np.complex64)
    ja, ka = backend.job(0, 1, 1, seed, 0, n, out_host=a)
    jb, kb = backend.job(0, 1, 1, seed, 0, n, out_host=b, delta_t=dt)
    backend.render([ja])
    backend.render([jb])
    backend.sync()
    assert a.tobytes() == b.tobytes()


# ---------------------------------------------------------------- flicker FM

def test_flicker_clock_matches_oracle(backend):
    """Flicker FM (src/clock/mod.rs:55-149) on top of the white walk and a
    drift: the host-stepped flicker integral interpolated on the device
    against the oracle's clock stream. With the walk off the clock offsets
    are bit-identical; IQ and taps are compared as for the walk alone, in the
    plain and precision forms."""
    import dataclasses
    seed = 17
    scene = noise_free(workloads.scene1(duration_s=2.0))
    rx = scene.receivers[0]
    for clock, label in ((NodeClockSpec(flicker_floor=1e-9, frequency_offset_y0=2e-7), "flicker only"),
                         (NodeClockSpec(allan_deviation_1s=1e-11, flicker_floor=1e-9, frequency_offset_y0=2e-7),
                          "flicker + white FM")):
        scene.receivers[0] = dataclasses.replace(rx, clock=clock)
        sc = prepare(scene)
        backend.upload(sc, seed)
        orc = OracleScene(sc, seed)
        check_band_both_clocks(backend, orc, sc, 0, 0, 0, seed, f"config 1 LEO, {label}", chunk=30_000)
        _, ddel = compare_taps(backend, orc, sc, 0, 0, 0, seed, 0, 96_000, label=label)
        _, _, taps = orc.render_band(0, n=96_000, tap_slot=0)
        j, keep = backend.job(0, 0, 0, seed, 0, 96_000)
        dt = backend.taps(j, 0, _lib.FORM_PLAIN)[3]
        if clock.allan_deviation_1s == 0.0:
            np.testing.assert_array_equal(dt, taps[3])
        # the flicker moved the clock: offsets minus drift span well over a ps
        resid = taps[3] - 2e-7 * (np.arange(96_000) / 48_000.0)
        assert np.ptp(resid) > 1e-12
        g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=40_000)
        o, _, _ = orc.render_band(0)
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128)).max()
        _say(f"flicker, {label}, precision form: max|d|/peak = {d / np.abs(o).max():.3e}")
        assert d <= 2e-6 * np.abs(o).max()


# ---------------------------------------------------------------- horizon gate

def _sky_scene(n_leo=24, duration_s=0.5):
    """A receiver with a 5 deg antenna cutoff under a sky of seeded LEO
    sources (src/channel/horizon.rs's case: most of them below the cutoff at
    any moment), 250 kHz."""
    from lipika_amd.config import (CircularKeplerian, Cw, FrontendConfig, ReceiverBandConfig, SceneConfig, SceneMeta,
                                   TransmitterConfig)
    rng = np.random.default_rng(91)
    txs = [TransmitterConfig(f"leo{k}", 437e6 + 2e3 * k, Cw(0.0),
                             CircularKeplerian(float(rng.uniform(6.9e6, 7.5e6)), float(rng.uniform(0, 180)),
                                               float(rng.uniform(0, 360)), float(rng.uniform(0, 360)), util.EPOCH), 0.0)
           for k in range(n_leo)]
    fe = FrontendConfig(antenna={"peak_dbi": 3.0, "horizon_dbi": -3.0, "fb_db": 10.0, "cutoff_deg": 5.0})
    band = ReceiverBandConfig("uhf", 437e6, 250e3, 290.0, noise_power_w=0.0, frontend=fe)
    return SceneConfig(SceneMeta(util.EPOCH, duration_s), txs, [util.lab_receiver([band])])


@pytest.mark.parametrize("form", [_lib.FORM_PLAIN, _lib.FORM_PRECISION, _lib.FORM_CONVERGED])
def test_horizon_gate_is_byte_identical(backend, form):
    """tests/horizon_gate.rs's contract on the GPU: with the gate on, the
    (tile, transmitter) pairs below the cutoff are skipped, and the IQ and the
    unoccluded flags are byte-identical to the ungated render; the plain form
    also matches the oracle (which has no gate)."""
    seed = 13
    sc = prepare(_sky_scene())
    backend.upload(sc, seed)
    n, na = sc.receivers[0].bands[0].n, len(sc.receivers[0].bands[0].active)
    out = {}
    for gate in (True, False):
        backend.set_horizon_gate(gate)
        backend.solve_stats_reset()
        g = np.empty(n, dtype=np.complex64)
        flags = np.zeros(na, dtype=np.uint8)
        for i0 in range(0, n, 40_000):
            f = np.zeros(na, dtype=np.uint8)
            j, keep = backend.job(0, 0, 0, seed, i0, min(40_000, n - i0), out_host=g[i0:i0 + 40_000], unoccluded=f)
            backend.render([j], form)
            flags |= f
        backend.sync()
        out[gate] = (g, flags, backend.horizon_gate_stats())
    backend.set_horizon_gate(True)
    (gon, fon, son), (goff, foff, soff) = out[True], out[False]
    _say(f"horizon gate, form {form}: {son} with the gate on; {int(fon.sum())} of {na} sources in view")
    assert son["tested"] > 0 and son["skipped"] > 0.3 * son["tested"]
    assert soff["skipped"] == 0
    assert 0 < fon.sum() < na
    np.testing.assert_array_equal(fon, foff)
    np.testing.assert_array_equal(gon.view(np.uint32), goff.view(np.uint32))
    if form == _lib.FORM_PLAIN:
        orc = OracleScene(sc, seed)
        o, _, _ = orc.render_band(0)
        compare_iq(gon, o, "gated plain form vs oracle")


# ---------------------------------------------------------------- full NAV and RINEX sources

def test_full_nav_and_rinex_sources_match_oracle(backend):
    """C/A with fully formed LNAV (src/signal/gps_nav.rs:554-585) on a
    circular orbit and on the reference's RINEX record (its grid from the
    broadcast ephemeris): the plain form bit-identical to the oracle with the
    oracle's clock offsets, the precision form within its bounds, across the
    NAV bit edges of 0.3 s at 2.048 MHz."""
    import math
    import os
    from lipika_amd.config import CaCode, CircularKeplerian, ReceiverBandConfig, RinexNav, SceneConfig, SceneMeta, \
        TransmitterConfig
    rinex = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rinex3_gps_prn01.rnx")
    toe_unix = 315964800.0 + 2317.0 * 604800.0 + 518400.0 - 18.0
    epoch = math.floor(toe_unix / 6.0) * 6.0 + 12.0
    txs = [TransmitterConfig("kep", 1575.42e6, CaCode(5, epoch, "full"),
                             CircularKeplerian(26.56e6, 55.0, 40.0, 120.0, epoch), 27.0),
           TransmitterConfig("brdc", 1575.42e6, CaCode(1, epoch, "full"), RinexNav(rinex, 1), 27.0)]
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    scene = SceneConfig(SceneMeta(epoch, 0.3), txs,
                        [util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-11, frequency_offset_y0=2e-7))])
    sc = prepare(scene)
    assert all(t.nav_mode == _lib.NAV_FULL for t in sc.transmitters)
    seed = 31
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    check_band_both_clocks(backend, orc, sc, 0, 0, 0, seed, "full NAV + RINEX, 2.048 MHz", chunk=150_000)
    o, _, _ = orc.render_band(0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=150_000)
    peak = float(np.max(np.abs(o)))
    d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
    _say(f"full NAV + RINEX, precision form: max|d|/peak={d.max() / peak:.3e}")
    assert d.max() <= 2e-6 * peak
    for slot in range(2):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 200_000, form=_lib.FORM_PRECISION, label="full NAV")


def test_tle_sources_match_oracle(backend):
    """Two GPS satellites from TLEs (examples/gps_real_tle.toml), propagated
    by SGP4 into their grids, with full LNAV from the ephemeris fitted to each
    TLE orbit: the plain form bit-identical to the oracle with the oracle's
    clock offsets, the precision form within its bounds."""
    from lipika_amd.config import CaCode, ReceiverBandConfig, SceneConfig, SceneMeta, Tle, TransmitterConfig
    epoch = 1780268400.0
    tles = {24: ("1 38833U 12053A   26150.86639939  .00000051  00000+0  00000+0 0  9997",
                 "2 38833  53.5827 144.3088 0179371  65.5174 296.3427  2.00567269 99098"),
            22: ("1 26407U 00040A   26150.58242625  .00000046  00000+0  00000+0 0  9999",
                 "2 26407  54.8538 215.8617 0121278 302.3883 230.2243  2.00557636189594")}
    txs = [TransmitterConfig(f"gps_prn{p}", 1575.42e6, CaCode(p, epoch, "full"), Tle(*l), 27.0)
           for p, l in tles.items()]
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    sc = prepare(SceneConfig(SceneMeta(epoch, 0.3), txs,
                             [util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-11))]))
    assert all(t.traj_kind == _lib.TRAJ_GRID and t.nav_mode == _lib.NAV_FULL for t in sc.transmitters)
    seed = 41
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    check_band_both_clocks(backend, orc, sc, 0, 0, 0, seed, "GPS from TLEs, 2.048 MHz", chunk=150_000)
    o, _, _ = orc.render_band(0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=150_000)
    peak = float(np.max(np.abs(o)))
    d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
    _say(f"GPS from TLEs, precision form: max|d|/peak={d.max() / peak:.3e}")
    assert d.max() <= 2e-6 * peak
    for slot in range(2):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 200_000, form=_lib.FORM_PRECISION, label="TLE")


def test_horizon_gate_tle_rise_is_byte_identical(backend):
    """tests/horizon_gate.rs:378-440 on the GPU: the example sky
    (examples/hiber_beacon_sky.toml: an ISS-like TLE carrying a burst-gated
    BPSK beacon) plus a CW tone on the same orbit, over Glasgow with a 0 deg
    antenna cutoff, 6 s at 48 kHz around the rise before the pass that
    culminates at Unix 1717260990: silent, then never silent; both sources
    visible; the gated render byte-identical to the ungated one at four chunk
    sizes, one of them a single gate stretch and one odd."""
    import math
    import os
    from lipika_amd.config import (Cw, FrontendConfig, ReceiverBandConfig, ReceiverConfig, SkyConfig,
                                   StationaryGeodetic, TransmitterConfig)
    from lipika_amd.pipeline import WindowOptions, parse_tle, render_window
    sky = SkyConfig.from_toml_file(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                                "hiber_beacon_sky.toml"))
    orbit = sky.transmitters[0].trajectory
    tle = parse_tle(orbit)
    lib = _lib.load()
    o = (C.c_double * 3)()
    lib.lpk_geodetic_to_ecef(55.86, -4.25, 20.0, o)
    o = np.array(o[:])

    def el(t):
        p = (C.c_double * 3)()
        assert lib.lpk_tle_state_ecef(C.byref(tle), t, p, None) == _lib.LPK_OK
        d = np.array(p[:]) - o
        return math.degrees(math.asin(np.dot(d, o / np.linalg.norm(o)) / np.linalg.norm(d)))

    rise = 1_717_260_990.0
    while el(rise) > 0.0:
        rise -= 1.0
        assert 1_717_260_990.0 - rise < 900.0
    start = rise - 3.0
    sky.transmitters.append(TransmitterConfig("tone", 401.15e6, Cw(1000.0), orbit, 0.0))
    rx = ReceiverConfig("glasgow", StationaryGeodetic(55.86, -4.25, 20.0))
    fe = FrontendConfig(antenna={"peak_dbi": 0.0, "horizon_dbi": 0.0, "fb_db": 0.0, "cutoff_deg": 0.0})
    band = ReceiverBandConfig("uhf", 401.15e6, 48_000.0, noise_power_w=0.0, frontend=fe)

    def collect(gate, chunk):
        backend.set_horizon_gate(gate)
        iq = []
        summary = render_window(sky, rx, band, start, 6.0, WindowOptions(11, chunk, 1), iq.append, backend=backend)
        iq = np.concatenate(iq)
        assert summary.n_samples_emitted == len(iq)
        return iq, summary

    try:
        ref, ref_summary = collect(False, 1 << 20)
        assert len(ref) == 288_000
        nz = np.flatnonzero(np.abs(ref) > 0.0)
        first = int(nz[0])
        _say(f"TLE rise: first signal at sample {first}")
        assert 48_000 < first < 240_000
        assert np.all(np.abs(ref[first:]) > 0.0)
        assert ref_summary.visible_transmitters == ["hiber_beacon", "tone"]
        for chunk in (1 << 16, 4800, 4799, 100_000):
            g, summary = collect(True, chunk)
            assert g.tobytes() == ref.tobytes(), chunk
            assert summary == ref_summary
        # a window that ends before the rise: in band, never visible, all zeros
        backend.set_horizon_gate(True)
        iq = []
        summary = render_window(sky, rx, band, start - 30.0, 2.0, WindowOptions(11, 1 << 16, 1), iq.append,
                                backend=backend)
        assert not np.any(np.concatenate(iq).view(np.uint32))
        assert summary.contributing_transmitters == ["hiber_beacon", "tone"] and summary.visible_transmitters == []
    finally:
        backend.set_horizon_gate(True)


def _many_links_scene():
    """72 in-band transmitters: 30 C/A and 30 BPSK on circular orbits, 12 CW
    tones on the ground, one 2.048 MHz L1 band of 0.01 s."""
    from lipika_amd.config import (Bpsk, CaCode, CircularKeplerian, Cw, ReceiverBandConfig, SceneConfig, SceneMeta,
                                   StationaryGeodetic, TransmitterConfig)
    txs = []
    for k in range(30):
        txs.append(TransmitterConfig(f"ca{k}", 1575.42e6 + 1e3 * k, CaCode(k % 32 + 1, util.EPOCH),
                                     CircularKeplerian(26.56e6, 55.0, 12.0 * k, 7.0 * k, util.EPOCH), 27.0))
    for k in range(30):
        txs.append(TransmitterConfig(f"bp{k}", 1575.42e6 - 2e3 * k, Bpsk(9600.0, 512, 40 + k, [1, -1, 1, 1]),
                                     CircularKeplerian(7.0e6 + 1e4 * k, 60.0, 11.0 * k, 13.0 * k, util.EPOCH), -10.0))
    for k in range(12):
        txs.append(TransmitterConfig(f"cw{k}", 1575.42e6 + 5e3 * k, Cw(200.0 * k),
                                     StationaryGeodetic(50.0 + k, -4.0 + k, 100.0), -20.0))
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    sc = prepare(SceneConfig(SceneMeta(util.EPOCH, 0.01), txs,
                             [util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-11))]))
    assert len(sc.receivers[0].bands[0].active) == 72
    return sc


def test_precision_form_at_its_32_link_limit(backend):
    """The precision form at its largest single pass: 32 in-band transmitters
    of every transmitter group (10 C/A and 10 BPSK on circular orbits, 8 CW
    and 4 burst-gated CW on the ground), against the oracle within its
    bounds, with taps on the first and last link of each group."""
    from lipika_amd.config import (Bpsk, CaCode, CircularKeplerian, Cw, GateConfig, ReceiverBandConfig, SceneConfig,
                                   SceneMeta, StationaryGeodetic, TransmitterConfig)
    txs = []
    for k in range(10):
        txs.append(TransmitterConfig(f"ca{k}", 1575.42e6 + 1e3 * k, CaCode(k + 1, util.EPOCH),
                                     CircularKeplerian(26.56e6, 55.0, 36.0 * k, 17.0 * k, util.EPOCH), 27.0))
    for k in range(10):
        txs.append(TransmitterConfig(f"bp{k}", 1575.42e6 - 2e3 * k, Bpsk(9600.0, 512, 40 + k, [1, -1, 1, 1]),
                                     CircularKeplerian(7.0e6 + 1e4 * k, 60.0, 33.0 * k, 13.0 * k, util.EPOCH), -10.0))
    for k in range(8):
        txs.append(TransmitterConfig(f"cw{k}", 1575.42e6 + 5e3 * k, Cw(200.0 * k),
                                     StationaryGeodetic(50.0 + k, -4.0 + k, 100.0), -20.0))
    for k in range(4):
        txs.append(TransmitterConfig(f"gated{k}", 1575.42e6 - 7e3 * k, Cw(150.0 * k),
                                     StationaryGeodetic(52.0 + k, -2.0 + k, 300.0), -20.0,
                                     gate=GateConfig(0.004, [(0.0005 * k, 0.0015)], 0.0)))
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    sc = prepare(SceneConfig(SceneMeta(util.EPOCH, 0.01), txs,
                             [util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-11))]))
    assert len(sc.receivers[0].bands[0].active) == 32
    seed = 23
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, _ = orc.render_band(0)
    peak = float(np.max(np.abs(o)))
    for chunk in (20_480, 7_000):
        g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=chunk)
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        _say(f"32 links, precision form, chunks of {chunk}: max|d|/peak={d.max() / peak:.3e}")
        assert np.all(np.isfinite(g)) and d.max() <= 2e-6 * peak
    for slot in (0, 9, 10, 19, 20, 27, 28, 31):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 20_480, form=_lib.FORM_PRECISION, label="32 links")


def test_plain_form_with_72_links(backend):
    """72 in-band transmitters in the plain form, with the oracle's clock
    offsets, against the oracle."""
    sc = _many_links_scene()
    seed = 17
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, taps = orc.render_band(0, tap_slot=0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, delta_t=taps[3])
    # CW tones with a baseband offset take an f32 sincos of a non-zero phase
    # (src/signal/cw.rs:41-52): the device's and glibc's differ by one ulp in
    # about 2.5 % of values, so a sample summing 12 of them matches to the
    # ulp, not always to the bit
    assert compare_iq(g, o, "72 links, plain, oracle clock offsets", strict=False) > 0.5


def _link_pass_scene(n_ca, n_bpsk, n_cw, n_gated, duration_s=0.01, rx_orbit=None, noise_power_w=0.0, frontend=None):
    """One 2.048 MHz L1 band with the given numbers of C/A and BPSK
    transmitters on circular orbits and CW and burst-gated CW tones on the
    ground, all in band; the receiver in the lab, or on `rx_orbit`."""
    from lipika_amd.config import (Bpsk, CaCode, CircularKeplerian, Cw, GateConfig, ReceiverBandConfig, SceneConfig,
                                   SceneMeta, StationaryGeodetic, TransmitterConfig)
    txs = []
    for k in range(n_ca):
        txs.append(TransmitterConfig(f"ca{k}", 1575.42e6 + 700.0 * k, CaCode(k % 32 + 1, util.EPOCH),
                                     CircularKeplerian(26.56e6, 55.0, 9.0 * k, 11.0 * k, util.EPOCH), 27.0))
    for k in range(n_bpsk):
        txs.append(TransmitterConfig(f"bp{k}", 1575.42e6 - 900.0 * k, Bpsk(9600.0, 512, 60 + k, [1, -1, 1, 1]),
                                     CircularKeplerian(7.05e6 + 7e3 * k, 62.0, 29.0 * k, 17.0 * k, util.EPOCH), -10.0))
    for k in range(n_cw):
        txs.append(TransmitterConfig(f"cw{k}", 1575.42e6 + 3e3 * k, Cw(170.0 * k),
                                     StationaryGeodetic(48.0 + 0.5 * k, -6.0 + k, 100.0), -20.0))
    for k in range(n_gated):
        txs.append(TransmitterConfig(f"gated{k}", 1575.42e6 - 5e3 * k, Cw(130.0 * k),
                                     StationaryGeodetic(53.0 + 0.3 * k, -3.0 + k, 300.0), -20.0,
                                     gate=GateConfig(0.004, [(0.0002 * k, 0.0015)], 0.0)))
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=noise_power_w,
                              **({"frontend": frontend} if frontend is not None else {}))
    rx = util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-11))
    if rx_orbit is not None:
        rx = util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-11), pos=rx_orbit)
    sc = prepare(SceneConfig(SceneMeta(util.EPOCH, duration_s), txs, [rx]))
    assert len(sc.receivers[0].bands[0].active) == n_ca + n_bpsk + n_cw + n_gated
    return sc


@pytest.mark.parametrize("counts", [(16, 16, 8, 8), (40, 30, 20, 10)])
def test_precision_form_in_link_passes(backend, counts):
    """Band-chunks of 48 and 100 in-band transmitters of every group in the
    precision form, rendered in passes of at most 32 links that add their f64
    sums through HBM: against the oracle within the form's bounds at two
    chunkings, with taps on links either side of every pass boundary."""
    sc = _link_pass_scene(*counts)
    n = sum(counts)
    seed = 29
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, _ = orc.render_band(0)
    peak = float(np.max(np.abs(o)))
    for chunk in (20_480, 7_000):
        g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=chunk)
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        _say(f"{n} links in passes, precision form, chunks of {chunk}: max|d|/peak={d.max() / peak:.3e}")
        assert np.all(np.isfinite(g)) and d.max() <= 2e-6 * peak
    slots = sorted({0, n - 1} | {b + e for b in range(32, n, 32) for e in (-1, 0)})
    for slot in slots:
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 20_480, form=_lib.FORM_PRECISION,
                     label=f"{n} links in passes")


def test_link_passes_with_noise_and_front_end(backend):
    """48 links in passes with kTB noise, gain, IP3 and a 12-bit ADC, which
    only the last pass applies (src/channel/mod.rs:310-334, src/rx/mod.rs:83-101):
    IQ within one ADC step plus the form's bound of the oracle's, ADC codes
    within one step of the oracle's and equal for at least 99 % of them."""
    fe = FrontendConfig(gain_db=130.0, nonlinearity=NonlinearityConfig(30.0), adc=AdcConfig(12, 1.0))
    sc = _link_pass_scene(16, 16, 8, 8, noise_power_w=None, frontend=fe)
    seed = 43
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, oc, _ = orc.render_band(0, codes=True)
    lsb = 1.0 / 2047.0
    assert oc.min() < -1000 and oc.max() > 1000     # the ADC is exercised
    for chunk in (20_480, 7_000):
        g, gc = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=chunk, codes=True)
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        diff = np.abs(gc.astype(int) - oc.astype(int))
        _say(f"48 links in passes with noise and front end, chunks of {chunk}: max|d|={d.max():.3e} "
             f"(lsb {lsb:.3e}), codes equal {np.mean(diff == 0):.5f}, max code step {diff.max()}")
        assert np.all(np.isfinite(g)) and d.max() <= 1.5 * lsb
        assert diff.max() <= 1 and np.mean(diff == 0) >= 0.99


def test_orbiting_receiver_in_link_passes(backend):
    """A receiver on a 400 km orbit with 48 in-band transmitters in the
    precision form's moving-receiver kernel, in two passes: a render call at
    two chunkings and a streaming plan, against the oracle within the form's
    bounds, and taps either side of the pass boundary."""
    from lipika_amd.config import CircularKeplerian
    sc = _link_pass_scene(16, 16, 8, 8, rx_orbit=CircularKeplerian(6_778_137.0, 70.0, 320.0, 50.0, util.EPOCH))
    seed = 37
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, _ = orc.render_band(0)
    peak = float(np.max(np.abs(o)))
    n = sc.receivers[0].bands[0].n
    for chunk in (20_480, 7_000):
        g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=chunk)
        d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
        _say(f"orbiting receiver, 48 links in passes, chunks of {chunk}: max|d|/peak={d.max() / peak:.3e}")
        assert np.all(np.isfinite(g)) and d.max() <= 2e-6 * peak
    j, keep = backend.job(0, 0, 0, seed, 0, n)
    got = np.zeros(n, dtype=np.complex64)
    with backend.plan([j], 6000, _lib.FORM_PRECISION) as plan:
        for kk in range(plan.n_chunks):
            plan.step(kk)
            backend.sync()
            i0, clen = plan.chunk_range(0, kk)
            if clen:
                backend.check(backend.lib.lpk_copy_to_host(backend.h, got[i0:].ctypes.data, plan.output(0), clen * 8))
    d = np.abs(got.astype(np.complex128) - o.astype(np.complex128))
    _say(f"orbiting receiver, plan of 48 links in passes: max|d|/peak={d.max() / peak:.3e}")
    assert np.all(np.isfinite(got)) and d.max() <= 2e-6 * peak
    for slot in (0, 31, 32, 47):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 20_480, form=_lib.FORM_PRECISION,
                     label="orbiting receiver, 48 links in passes")


def _antenna_pass_scene(duration_s=0.01):
    """40 C/A MEOs in six planes seen through a 5 deg cutoff antenna (about a
    third of them above it) in one 2.048 MHz L1 band."""
    from lipika_amd.config import (CaCode, CircularKeplerian, FrontendConfig, ReceiverBandConfig, SceneConfig,
                                   SceneMeta, TransmitterConfig)
    txs = [TransmitterConfig(f"sv{k}", 1575.42e6 + 400.0 * k, CaCode(k % 32 + 1, util.EPOCH),
                             CircularKeplerian(26.56e6, 55.0, 60.0 * (k % 6), 360.0 * k / 40.0 + 7.0 * (k % 6),
                                               util.EPOCH), 27.0)
           for k in range(40)]
    fe = FrontendConfig(antenna={"peak_dbi": 3.0, "horizon_dbi": -3.0, "fb_db": 10.0, "cutoff_deg": 5.0})
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0, frontend=fe)
    return SceneConfig(SceneMeta(util.EPOCH, duration_s), txs, [util.lab_receiver([band])])


def test_plan_of_an_antenna_band_in_link_passes(backend):
    """The 40-source antenna band as a streaming plan in the precision form:
    each step's two passes through the antenna kernel, against the oracle
    within the form's bounds, with the horizon gate on and off giving the same
    bytes."""
    sc = prepare(_antenna_pass_scene())
    seed = 31
    backend.upload(sc, seed)
    o, _, _ = OracleScene(sc, seed).render_band(0)
    peak = float(np.max(np.abs(o)))
    n = sc.receivers[0].bands[0].n
    got = {}
    for gate in (True, False):
        backend.set_horizon_gate(gate)
        g = np.zeros(n, dtype=np.complex64)
        j, keep = backend.job(0, 0, 0, seed, 0, n)
        with backend.plan([j], 6000, _lib.FORM_PRECISION) as plan:
            for kk in range(plan.n_chunks):
                plan.step(kk)
                backend.sync()
                i0, clen = plan.chunk_range(0, kk)
                if clen:
                    backend.check(backend.lib.lpk_copy_to_host(backend.h, g[i0:].ctypes.data, plan.output(0),
                                                               clen * 8))
        got[gate] = g
    backend.set_horizon_gate(True)
    d = np.abs(got[True].astype(np.complex128) - o.astype(np.complex128))
    _say(f"plan of the antenna band in passes: max|d|/peak={d.max() / peak:.3e}")
    assert np.all(np.isfinite(got[True])) and d.max() <= 2e-6 * peak
    assert got[True].tobytes() == got[False].tobytes()


def test_precision_antenna_band_in_link_passes(backend):
    """An antenna band of 40 MEO sources, about a third above its cutoff, in
    the precision form's antenna kernel in two passes: IQ against the oracle
    and the unoccluded flags, with the horizon gate on and off giving the same
    bytes."""
    sc = prepare(_antenna_pass_scene())
    seed = 31
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    na = len(sc.receivers[0].bands[0].active)
    assert na > 32
    flags = np.zeros(na, dtype=np.uint8)
    o, _, _ = orc.render_band(0, unoccluded=flags)
    n = sc.receivers[0].bands[0].n
    peak = float(np.max(np.abs(o)))
    got = {}
    for gate in (True, False):
        backend.set_horizon_gate(gate)
        g = np.empty(n, dtype=np.complex64)
        gflags = np.zeros(na, dtype=np.uint8)
        for i0 in range(0, n, 7_000):
            f = np.zeros(na, dtype=np.uint8)
            j, keep = backend.job(0, 0, 0, seed, i0, min(7_000, n - i0), out_host=g[i0:i0 + 7_000], unoccluded=f)
            backend.render([j], _lib.FORM_PRECISION)
            backend.sync()
            gflags |= f
        got[gate] = (g, gflags)
    backend.set_horizon_gate(True)
    g, gflags = got[True]
    d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
    _say(f"{na}-link antenna band in passes, precision: max|d|/peak={d.max() / peak:.3e}, "
         f"{int(flags.sum())} of {na} in view")
    assert np.all(np.isfinite(g)) and d.max() <= 2e-6 * peak
    assert list(gflags != 0) == list(flags != 0)
    assert got[False][0].tobytes() == g.tobytes() and list(got[False][1]) == list(gflags)


@pytest.mark.parametrize("counts", [(16, 16, 8, 8), (40, 30, 20, 10)])
def test_plan_in_precision_link_passes(backend, counts):
    """A streaming plan of a 48- and a 100-link band in the precision form
    renders each step in passes of at most 32 links: every step's chunk
    against the oracle within the form's bounds, the plan run twice (the
    second restart reuses both slots' staging), with a 12-link band of the
    same receiver in the plan that renders in one pass."""
    sc = _link_pass_scene(*counts)
    n_links = sum(counts)
    seed = 23
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, _ = orc.render_band(0)
    peak = float(np.max(np.abs(o)))
    n = sc.receivers[0].bands[0].n
    j, keep = backend.job(0, 0, 0, seed, 0, n)
    j12, keep12 = backend.job(0, 0, 0, seed, 0, n)
    j12.n_active = 12      # the first 12 active links: its own band in the plan
    chunk = 6000
    ref12 = np.empty(n, dtype=np.complex64)   # eager renders of the same chunks
    for i0 in range(0, n, chunk):
        jr, keep_r = backend.job(0, 0, 0, seed, i0, min(chunk, n - i0), out_host=ref12[i0:])
        jr.n_active = 12
        backend.render([jr], _lib.FORM_PRECISION)
        backend.sync()
    lib = backend.lib
    with backend.plan([j, j12], chunk, _lib.FORM_PRECISION) as plan:
        for rnd in range(2):
            got, got12 = np.zeros(n, dtype=np.complex64), np.zeros(n, dtype=np.complex64)
            for kk in range(plan.n_chunks):
                plan.step(kk)
                backend.sync()
                for band, dst in ((0, got), (1, got12)):
                    i0, clen = plan.chunk_range(band, kk)
                    if clen:
                        backend.check(lib.lpk_copy_to_host(backend.h, dst[i0:].ctypes.data, plan.output(band),
                                                           clen * 8))
            d = np.abs(got.astype(np.complex128) - o.astype(np.complex128))
            _say(f"plan of {n_links} links in passes, run {rnd}: max|d|/peak={d.max() / peak:.3e}")
            assert np.all(np.isfinite(got)) and d.max() <= 2e-6 * peak
            assert np.array_equal(got12, ref12)
    assert plan.max_tiles > 0


@pytest.mark.skip(reason="the first precision render of this 72-link band in one pass (one tile per anchor batch, "
                         "two record groups) ended in a device memory fault; host emulations of the kernels find every "
                         "access in bounds and no race, so the cause is not shown; render calls now split bands of "
                         "more than 32 links into passes of at most 32, which the link-pass tests cover")
def test_precision_form_with_more_links_than_one_record_group(backend):
    """72 in-band transmitters in the precision form: link groups of 64
    records, one tile per anchor batch, transmitter groups that straddle the
    group boundary; against the oracle at its bounds."""
    sc = _many_links_scene()
    seed = 17
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    o, _, _ = orc.render_band(0, tap_slot=0)
    g, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_PRECISION, chunk=7_000)
    peak = float(np.max(np.abs(o)))
    d = np.abs(g.astype(np.complex128) - o.astype(np.complex128))
    _say(f"72 links, precision form: max|d|/peak={d.max() / peak:.3e}")
    assert np.all(np.isfinite(g)) and d.max() <= 2e-6 * peak
    for slot in (0, 63, 64, 71):
        compare_taps(backend, orc, sc, 0, 0, 0, seed, slot, 20_480, form=_lib.FORM_PRECISION, label="72 links")
