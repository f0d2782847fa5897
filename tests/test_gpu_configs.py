"""GPU parity on the benchmark workloads themselves: configurations 3
(32 C/A MEO x 16 receivers, 120 s), 4 (40 mixed transmitters x 64 receivers
with 130 uneven bands, 30 s) and 5 (32 C/A MEO x 256 receivers, 60 s), each
rendered noise-free in windows at the start and at the end of the scene and
compared with the CPU FP64 oracle (orc_render_window) given the device's own
clock offsets for the window (the clock walk is pinned separately, in
test_gpu_parity.py and test_oracle_pins.py).

Bars, per window:
  - plain form: bit-identical to the oracle (src/pipeline.rs:915-983);
  - precision form: every link's solved range within 1 ps of delay and its
    f64 carrier phase (src/channel/mod.rs:256-258, formed from the solved
    range) within 1e-6 rad; IQ within 2e-6 of the band's peak;
  - batching form (a streaming plan, steps on two alternating streams): as the
    form it runs, against the oracle;
  - converged and learned forms (configuration 5): against the oracle's
    converged solve (orc_solve_converged), bit-identical and ulp-bounded.
Configuration 4's units are chosen to cover every (sample rate, LO carrier)
pairing, each with its LEO and MEO links.
"""
import collections
import math

import numpy as np
import pytest

from lipika_amd import _lib, workloads
from lipika_amd.pipeline import WarmStartNet, prepare
from oracle.pyoracle import OracleScene

from test_gpu_parity import noise_free

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0
WORST = collections.defaultdict(float)


def _say(*a):
    print(*a, flush=True)


def _note(key, v):
    WORST[key] = max(WORST[key], float(v))


def _units(sc):
    out, w = [], 0
    for r, rx in enumerate(sc.receivers):
        for b, pb in enumerate(rx.bands):
            out.append((w, r, b, pb))
            w += 1
    return out


def _device_offsets(be, r, b, w, seed, i0, n):
    """The device clock's offsets for samples [i0, i0 + n) (tap of slot 0)."""
    j, keep = be.job(r, b, w, seed, i0, n)
    return be.taps(j, 0, _lib.FORM_PLAIN)[3]


def _render(be, r, b, w, seed, i0, n, form):
    out = np.empty(n, dtype=np.complex64)
    j, keep = be.job(r, b, w, seed, i0, n, out_host=out)
    be.render([j], form)
    be.sync()
    return out


def _iq_error(g, o):
    peak = float(np.max(np.abs(o)))
    d = float(np.max(np.abs(g.astype(np.complex128) - o.astype(np.complex128))))
    same = float(np.mean((g.view(np.uint32) == o.view(np.uint32)).reshape(-1, 2).all(axis=1)))
    assert np.all(np.isfinite(g))
    return d / peak if peak > 0 else d, same


def _link_taps(be, orc_taps, r, b, w, seed, i0, n, form, carriers, label, tag):
    """Every link's range, emission time and phase against the oracle's."""
    ote, org, oph = orc_taps
    j, keep = be.job(r, b, w, seed, i0, n)
    worst = (0.0, 0.0)
    for slot in range(ote.shape[0]):
        te, rg, ph, _ = be.taps(j, slot, form)
        dphi = float(np.max(np.abs(ph - oph[slot])))
        ddel = float(np.max(np.abs(rg - org[slot]))) / C_LIGHT
        dte = float(np.max(np.abs(te - ote[slot])))
        # the same delay error as carrier phase, without the f64 rounding of
        # the phase expression itself
        dmod = 2.0 * math.pi * carriers[slot] * ddel
        _note(f"{tag} phase", dphi)
        _note(f"{tag} delay", ddel)
        _note(f"{tag} modelled phase", dmod)
        if dphi > 1e-6 or ddel > 1e-12 or dte > 1e-12:
            _say(f"{label} slot {slot} ({carriers[slot] / 1e6:.2f} MHz): |dphi|={dphi:.3e} rad "
                 f"|ddelay|={ddel:.3e} s |dt_emit|={dte:.3e} s")
        assert dphi <= 1e-6 and ddel <= 1e-12 and dte <= 1e-12
        worst = (max(worst[0], dphi), max(worst[1], ddel))
    return worst


def _check_window(be, orc, sc, unit, seed, i0, n, tag, label, plain_taps=False):
    w, r, b, pb = unit
    dt = _device_offsets(be, r, b, w, seed, i0, n)
    o, _, ot = orc.render_window(w, i0, dt, taps=True)
    carriers = [sc.transmitters[t].carrier_hz for t in pb.active]
    g = _render(be, r, b, w, seed, i0, n, _lib.FORM_PLAIN)
    ep, same = _iq_error(g, o)
    assert same == 1.0, f"{label}: plain form not bit-identical ({same:.6f}, |d|/peak={ep:.3e})"
    if plain_taps:
        _link_taps(be, ot, r, b, w, seed, i0, n, _lib.FORM_PLAIN, carriers, label + " plain", tag + " plain")
    g = _render(be, r, b, w, seed, i0, n, _lib.FORM_PRECISION)
    eq, _ = _iq_error(g, o)
    _note(f"{tag} precision IQ/peak", eq)
    assert eq <= 2e-6, f"{label}: precision IQ error {eq:.3e} of peak"
    ph, dl = _link_taps(be, ot, r, b, w, seed, i0, n, _lib.FORM_PRECISION, carriers, label, tag + " precision")
    _say(f"{label}: {len(pb.active)} links, plain bit-identical; precision |dphi|<={ph:.2e} rad "
         f"|ddelay|<={dl:.2e} s IQ<={eq:.2e} of peak")


@pytest.fixture(scope="module")
def cfg4():
    sc = prepare(noise_free(workloads.scene4()))
    best = {}
    for u in _units(sc):
        k = (u[3].band.sample_rate_hz, u[3].band.lo_carrier_hz)
        if k not in best or len(u[3].active) > len(best[k][3].active):
            best[k] = u
    return sc, [best[k] for k in sorted(best)]


def test_config4_units_cover_every_rate_and_carrier(cfg4):
    sc, units = cfg4
    assert len(units) == 12
    for w, r, b, pb in units:
        radii = {("LEO" if np.linalg.norm(sc.transmitters[t].grid[:, 0]) < 1e7 else "MEO") for t in pb.active}
        assert radii == {"LEO", "MEO"}, (w, radii)


def test_config4_windows_against_oracle(backend, cfg4):
    """Configuration 4: the first and the last 8 192 samples of a band of
    every (sample rate, LO) pairing, in the plain and precision forms."""
    sc, units = cfg4
    seed = workloads.SEED
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    for u in units:
        pb = u[3]
        n = 8192
        for i0 in (0, pb.n - n):
            _check_window(backend, orc, sc, u, seed, i0, n, "config 4",
                          f"config 4 unit {u[0]} ({pb.band.sample_rate_hz / 1e3:g} kHz, LO "
                          f"{pb.band.lo_carrier_hz / 1e6:g} MHz) at {i0}", plain_taps=(i0 == 0))
    _say("config 4 worst: " + ", ".join(f"{k}={v:.3e}" for k, v in WORST.items() if k.startswith("config 4")))


@pytest.mark.parametrize("form", [_lib.FORM_PLAIN, _lib.FORM_PRECISION])
def test_config4_plan_against_oracle(backend, cfg4, form):
    """The batching form on configuration 4's units: a streaming plan over
    the first 16 384 samples of each band, steps of 4 096 samples of the
    fastest band and the same time slice of the others, the steps
    alternating between the plan's two streams, against the oracle."""
    sc, units = cfg4
    seed = workloads.SEED
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    lib = _lib.load()
    n, chunk = 16384, 4096
    bands, keep = [], []
    for w, r, b, pb in units:
        j, k = backend.job(r, b, w, seed, 0, n)
        bands.append(j)
        keep.append(k)
    got = [np.empty(n, dtype=np.complex64) for _ in units]
    with backend.plan(bands, chunk, form) as plan:
        for kk in range(plan.n_chunks):
            plan.step(kk)
            backend.sync()
            for q in range(len(units)):
                i0, clen = plan.chunk_range(q, kk)
                if clen:
                    backend.check(lib.lpk_copy_to_host(backend.h, got[q][i0:].ctypes.data, plan.output(q), clen * 8))
        assert all(plan.chunk_range(q, plan.n_chunks - 1)[0] + plan.chunk_range(q, plan.n_chunks - 1)[1] == n
                   for q in range(len(units)))
    for q, (w, r, b, pb) in enumerate(units):
        dt = _device_offsets(backend, r, b, w, seed, 0, n)
        o, _, _ = orc.render_window(w, 0, dt)
        e, same = _iq_error(got[q], o)
        if form == _lib.FORM_PLAIN:
            assert same == 1.0, (w, same)
        else:
            _note("config 4 plan precision IQ/peak", e)
            assert e <= 2e-6, (w, e)
    _say(f"config 4 plan, form {form}: {len(units)} bands x {n // chunk} replays against the oracle")


@pytest.fixture(scope="module")
def cfg3():
    return prepare(noise_free(workloads.scene3()))


def test_config3_every_receiver_against_oracle(backend, cfg3):
    """Configuration 3 (the precision form's configuration): every receiver,
    32 C/A links each, windows of 4 096 samples at 0 s and at the end of the
    120 s scene."""
    sc = cfg3
    seed = workloads.SEED
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    for u in _units(sc):
        pb = u[3]
        assert len(pb.active) == 32
        n = 4096
        for i0 in (0, pb.n - n):
            _check_window(backend, orc, sc, u, seed, i0, n, "config 3", f"config 3 receiver {u[1]} at {i0}")
    _say("config 3 worst: " + ", ".join(f"{k}={v:.3e}" for k, v in WORST.items() if k.startswith("config 3")))


@pytest.fixture(scope="module")
def cfg5():
    return prepare(noise_free(workloads.scene5()))


def test_config5_forms_against_oracle(backend, cfg5):
    """Configuration 5: eight receivers spread over the 256, windows at 0 s
    and at the end of the 60 s scene. Plain against the reference's solve;
    converged (bit-identical) and learned (ulp-bounded: the same root from
    another start) against the oracle's converged solve; taps of the
    converged form within 1 ps and 1e-6 rad."""
    sc = cfg5
    seed = workloads.SEED
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    net = WarmStartNet(backend, init_seed=5)
    try:
        rep = net.train(n_pairs=1 << 21, data_seed=55, epochs=4, batch=8192, lr=2e-3, lr_final=2e-5, n_val=20_000)
        _say(f"config 5 network: validation rms {rep['val_rms_s']:.3e} s")
        backend.warm_start(net)
        units = _units(sc)
        for u in units[::32]:
            w, r, b, pb = u
            n = 8192
            for i0 in (0, pb.n - n):
                dt = _device_offsets(backend, r, b, w, seed, i0, n)
                o, _, _ = orc.render_window(w, i0, dt)
                e, same = _iq_error(_render(backend, r, b, w, seed, i0, n, _lib.FORM_PLAIN), o)
                assert same == 1.0, (w, i0, same)
                oc, _, ot = orc.render_window(w, i0, dt, taps=True, converged=True)
                e, same = _iq_error(_render(backend, r, b, w, seed, i0, n, _lib.FORM_CONVERGED), oc)
                assert same == 1.0, (w, i0, same)
                el, _ = _iq_error(_render(backend, r, b, w, seed, i0, n, _lib.FORM_LEARNED), oc)
                _note("config 5 learned IQ/peak", el)
                assert el <= 1e-6, (w, i0, el)
                j, keep = backend.job(r, b, w, seed, i0, n)
                for slot in range(0, len(pb.active), 4):
                    te, rg, ph, _ = backend.taps(j, slot, _lib.FORM_CONVERGED)
                    dphi = float(np.max(np.abs(ph - ot[2][slot])))
                    ddel = float(np.max(np.abs(rg - ot[1][slot]))) / C_LIGHT
                    _note("config 5 converged phase", dphi)
                    _note("config 5 converged delay", ddel)
                    assert dphi <= 1e-6 and ddel <= 1e-12, (w, i0, slot, dphi, ddel)
                _say(f"config 5 receiver {r} at {i0}: plain and converged bit-identical, learned "
                     f"{el:.2e} of peak")
    finally:
        backend.warm_start(None)
        net.close()


def test_zz_config_parity_summary():
    """Prints the worst errors met on the configuration workloads."""
    for k in sorted(WORST):
        _say(f"worst {k}: {WORST[k]:.3e}")
