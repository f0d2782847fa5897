"""GPU checks of the learned warm start and the converged forms, through the
C ABI, against the CPU oracles:
  - training pairs: features bit-for-bit (<= 2 f32 ulp) and targets within
    1e-12 s of oracle/lws_oracle.py's light time on the same geometry;
  - the matrix-core forward and the training kernel's gradient against the
    FP64 back-propagation with the same BF16 roundings (relative L2 <= 1e-3);
  - training on the device lowers the validation error of the predicted
    light time below the straight-line start by orders of magnitude;
  - the converged form is bit-identical to the oracle's converged solve (same
    clock offsets), and the learned form gives the same IQ (to the converged
    solve's tolerance) with one evaluation of the emission equation per
    contribution; a useless network only costs evaluations at the query
    points.
"""
import math

import numpy as np
import pytest

from lipika_amd import _lib, workloads
from lipika_amd.config import (CaCode, CircularKeplerian, NodeClockSpec, ReceiverBandConfig, SceneConfig, SceneMeta,
                               StationaryGeodetic, TransmitterConfig)
from lipika_amd.pipeline import WarmStartNet, prepare
from oracle import lws_oracle as L
from oracle.pyoracle import OracleScene

import util
from test_gpu_parity import compare_iq, gpu_band, noise_free

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def trained(backend):
    net = WarmStartNet(backend, init_seed=4)
    rep = net.train(n_pairs=1 << 22, data_seed=77, epochs=8, batch=8192, lr=2e-3, lr_final=2e-5, n_val=50_000)
    _say(f"training: {rep}")
    yield net, rep
    net.close()


def test_pair_generation_matches_oracle(backend):
    net = WarmStartNet(backend, init_seed=1)
    n = 8192
    X, Y, G = net.generate(n, seed=11, geometry=True)
    try:
        x, y, g = X.numpy(), Y.numpy(), G.numpy()
    finally:
        for a in (X, Y, G):
            a.free()
        net.close()
    obs, a, inc, node, u = g[:, 0:3], g[:, 3], g[:, 4], g[:, 5], g[:, 6]
    xo = L.features(obs, a, inc, node, u)
    np.testing.assert_array_max_ulp(x, xo, maxulp=2)
    tau_sl, tau = L.light_time(obs, a, inc, node, u)
    dtau = np.abs(g[:, 7] - tau).max()
    dy = np.abs(y.astype(np.float64) - (tau - tau_sl) / L.TARGET_SCALE).max()
    _say(f"pairs: max|dtau|={dtau:.3e} s  max|dy|={dy:.3e} (x 1e-6 s)  y in [{y.min():.3f}, {y.max():.3f}]")
    assert dtau <= 1e-15
    assert dy <= 1e-6
    leo = a < 1e7
    assert 0.45 < leo.mean() < 0.55
    assert np.all((a > 6.9e6) & (a < 2.7e7))


def test_forward_matches_oracle(backend):
    net = WarmStartNet(backend, init_seed=3)
    n = 4096
    X, Y, _ = net.generate(n, seed=12)
    try:
        yd = net.infer(X, n)
        x = X.numpy()
        th = net.params
    finally:
        X.free()
        Y.free()
        net.close()
    yo, _ = L.forward(th, x, quantize=True)
    ye, _ = L.forward(th, x, quantize=False)
    d = np.abs(yd - yo)
    _say(f"forward: max|d|={d.max():.3e} p99={np.percentile(d, 99):.3e} (vs unrounded: {np.abs(yd - ye).max():.3e});"
         f" |y| <= {np.abs(yo).max():.3f}")
    assert np.percentile(d, 99) <= 1e-4
    assert d.max() <= 5e-3


def test_gradient_matches_oracle(backend):
    net = WarmStartNet(backend, init_seed=5)
    n = 2048
    X, Y, _ = net.generate(n, seed=13)
    try:
        loss_d, gd = net.gradient(X, Y, n)
        x, y = X.numpy(), Y.numpy()
        th = net.params
    finally:
        X.free()
        Y.free()
        net.close()
    loss_o, go = L.loss_and_grad(th, x, y, quantize=True)
    rel = np.linalg.norm(gd - go) / np.linalg.norm(go)
    _say(f"gradient: loss device {loss_d:.6e} oracle {loss_o:.6e}; relative L2 error {rel:.3e}")
    assert abs(loss_d - loss_o) <= 1e-4 * loss_o
    assert rel <= 1e-3
    for lo, hi, name in ((L.OFF_W1, L.OFF_B1, "W1"), (L.OFF_B1, L.OFF_W2, "b1"), (L.OFF_W2, L.OFF_B2, "W2"),
                         (L.OFF_B2, L.OFF_W3, "b2"), (L.OFF_W3, L.OFF_B3, "W3"), (L.OFF_B3, L.OFF_W4, "b3"),
                         (L.OFF_W4, L.OFF_B4, "w4"), (L.OFF_B4, L.N_PARAMS, "b4")):
        r = np.linalg.norm(gd[lo:hi] - go[lo:hi]) / max(np.linalg.norm(go[lo:hi]), 1e-30)
        assert r <= 1e-2, (name, r)


def test_training_learns_the_light_time(trained):
    net, rep = trained
    # the target's own spread (the straight-line start's error) is ~2e-7 s;
    # 4M pairs x 8 epochs bring the light time within ~1e-8 s
    assert rep["loss_last"] < rep["loss_first"] / 50
    assert rep["val_rms_s"] < 1.5e-8
    assert rep["steps"] == 8 * ((1 << 22) // 8192)


def _meo_scene(n_rx=2, n_tx=8, duration_s=0.05):
    """Config-5 geometry, scaled down: C/A transmitters on MEO circles
    (i = 55 deg, six planes) seen by receivers at seeded positions, 2.048 MHz."""
    rng = np.random.default_rng(505)
    txs = [TransmitterConfig(f"ca{k}", 1575.42e6, CaCode(k + 1, util.EPOCH),
                             CircularKeplerian(26_578_137.0, 55.0, 60.0 * (k % 6), float(rng.uniform(0, 360)),
                                               util.EPOCH), 27.0)
           for k in range(n_tx)]
    rxs = []
    for r in range(n_rx):
        band = ReceiverBandConfig("L1", 1575.42e6, 2.048e6, noise_power_w=0.0)
        rxs.append(util.lab_receiver([band], NodeClockSpec(allan_deviation_1s=1e-10, frequency_offset_y0=3e-7),
                                     pos=StationaryGeodetic(float(rng.uniform(-60, 60)), float(rng.uniform(-180, 180)),
                                                            100.0),
                                     name=f"rx{r}"))
    return SceneConfig(SceneMeta(util.EPOCH, duration_s), txs, rxs)


def test_converged_form_matches_oracle(backend):
    """FORM_CONVERGED against the oracle's converged solve: bit-identical IQ
    with the oracle's clock offsets; solved range and phase taps within
    1e-12 s and 1e-6 rad. It departs from the reference's two-iteration map
    by that map's own truncation, which is reported, not bounded."""
    seed = 21
    sc = prepare(noise_free(workloads.scene2(duration_s=0.2)))
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    for w, (b, name) in enumerate(((0, "UHF BPSK LEO"), (1, "L1 6xC/A MEO"))):
        o, _, taps = orc.render_band(w, tap_slot=0, converged=True)
        g, _ = gpu_band(backend, sc, w, 0, b, seed, chunk=150_000, delta_t=taps[3], form=_lib.FORM_CONVERGED)
        assert compare_iq(g, o, f"converged form, {name}, oracle clock offsets") == 1.0
        n = min(len(o), 100_000)
        for slot in range(len(sc.receivers[0].bands[b].active)):
            _, _, to = orc.render_band(w, n=n, tap_slot=slot, converged=True)
            _, _, tr = orc.render_band(w, n=n, tap_slot=slot)
            j, keep = backend.job(0, b, w, seed, 0, n)
            te, rg, ph, dt = backend.taps(j, slot, _lib.FORM_CONVERGED)
            dphi = np.abs(ph - to[2]).max()
            ddel = np.abs(rg - to[1]).max() / C_LIGHT
            dref = np.abs(to[2] - tr[2]).max()
            _say(f"converged taps {name} slot {slot}: |dphi|={dphi:.3e} rad |ddelay|={ddel:.3e} s; "
                 f"two-iteration map departs by {dref:.3e} rad")
            assert dphi <= 1e-6 and ddel <= 1e-12


def test_learned_form_saves_evaluations(backend, trained):
    """Same IQ as the converged form from the reference's start (within the
    f32 ulp bound), with one evaluation of the emission equation per
    contribution: the converged light times at the queries, interpolated,
    start every MEO sample within about 1e-15 s of its root, so the first
    update is already below the tolerance. The count includes the queries'
    own evaluations."""
    net, _ = trained
    seed = 31
    sc = prepare(_meo_scene())
    backend.upload(sc, seed)
    backend.warm_start(net)
    res = {}
    for form in (_lib.FORM_CONVERGED, _lib.FORM_LEARNED):
        backend.solve_stats_reset()
        outs = [gpu_band(backend, sc, w, w, 0, seed, chunk=50_000, form=form)[0] for w in range(len(sc.receivers))]
        st = backend.solve_stats()
        res[form] = (outs, st)
        _say(f"form {form}: {st} -> {st['evaluations'] / st['contributions']:.4f} evaluations per contribution")
    (oc, sc_), (ol, sl) = res[_lib.FORM_CONVERGED], res[_lib.FORM_LEARNED]
    assert sc_["contributions"] == sl["contributions"] > 0
    assert sc_["unconverged"] == 0 and sl["unconverged"] == 0
    cold, learned = sc_["evaluations"] / sc_["contributions"], sl["evaluations"] / sl["contributions"]
    assert cold >= 2.9
    assert learned <= 1.05
    assert sl["fallbacks"] <= 1e-4 * sl["contributions"]
    for w in range(len(sc.receivers)):
        compare_iq(ol[w], oc[w], f"learned vs converged, receiver {w}", strict=False)
    orc = OracleScene(sc, seed)
    o, _, _ = orc.render_band(0, converged=True)
    compare_iq(ol[0], o, "learned form vs oracle converged solve", strict=False)


def test_useless_network_falls_back(backend):
    """A network that predicts light times 100 s off cannot converge within
    three evaluations (the first lands on the wrong part of the orbit, the
    update after it is still ~1e-9 s): every query's refinement is counted as
    a fallback and keeps iterating from where it landed (the iteration
    contracts by about |dr/dt| / c per evaluation). The converged light times
    at the queries then start every sample as well as a good network would:
    one evaluation each, fewer in all than the reference's start, and the
    output matches the converged form's to its tolerance. Both solves stop on
    an update below 1e-13 s, so they agree to about 1e-13 s * |dr/dt| / c
    (1e-18 s, an ulp or two of the emission time): close, not bit-identical."""
    net = WarmStartNet(backend, init_seed=6)
    th = np.zeros(_lib.LWS_PARAMS, dtype=np.float32)
    th[L.OFF_B4] = 1e8
    net.params = th
    seed = 41
    sc = prepare(_meo_scene(n_rx=1, n_tx=4, duration_s=0.01))
    backend.upload(sc, seed)
    backend.warm_start(net)
    try:
        backend.solve_stats_reset()
        gl, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_LEARNED)
        st = backend.solve_stats()
        backend.solve_stats_reset()
        gc, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_CONVERGED)
        sc_ = backend.solve_stats()
    finally:
        net.close()
    _say(f"useless network: {st}; converged form: {sc_}")
    assert st["contributions"] == sc_["contributions"] > 0
    assert st["unconverged"] == 0
    # every query falls back: at least one per link and 2 ms of the band
    n_links = len(sc.receivers[0].bands[0].active)
    assert st["fallbacks"] >= n_links * math.floor(0.01 / 2e-3)
    assert st["fallbacks"] < 1e-2 * st["contributions"]
    assert st["evaluations"] < 1.05 * st["contributions"] < sc_["evaluations"]
    compare_iq(gl, gc, "fallback vs converged", strict=False)


def test_learned_form_without_elements_uses_straight_line(backend, trained):
    """Stationary transmitters carry no orbital elements: the learned form
    starts them from the straight-line light time, converges, and matches the
    converged form."""
    net, _ = trained
    seed = 51
    sc = prepare(util.all_kinds_scene(duration_s=0.003))
    backend.upload(sc, seed)
    backend.warm_start(net)
    gl, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_LEARNED)
    gc, _ = gpu_band(backend, sc, 0, 0, 0, seed, form=_lib.FORM_CONVERGED)
    compare_iq(gl, gc, "learned (no elements) vs converged", strict=False)
