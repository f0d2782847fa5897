"""Flicker FM of the receiver clock (src/clock/mod.rs:55-149), on the CPU.

The reference draws its flicker stream from rand_distr's normals, which are
not reproduced (the same reason as the white-FM walk), so the oracle's
flicker process is pinned by mathematics and by the reference's own
statistical test:
  - the exact Allan deviation used for the calibration against a Monte-Carlo
    simulation of the same pole sum with numpy's normals;
  - the reference's flat-floor test (src/clock/mod.rs:355-372) with its own
    bounds, then a tighter level check the exact calibration allows;
  - the clock composition eps + flicker + (c + y0 t) of src/clock/mod.rs:344.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle

P = C.POINTER


@pytest.fixture(scope="module")
def orc():
    return pyoracle.load()


def poles(orc):
    a, g = (C.c_double * 20)(), (C.c_double * 20)()
    orc.orc_flicker_poles(a, g)
    return np.array(a[:]), np.array(g[:])


def allan(delta, m, fc):
    """src/clock/mod.rs:150-163, non-overlapping second differences."""
    idx = np.arange(0, len(delta) - 2 * m, m)
    d = delta[idx + 2 * m] - 2 * delta[idx + m] + delta[idx]
    tau = m / fc
    return math.sqrt(np.mean(d * d) / (2 * tau * tau))


def test_poles_span_the_band(orc):
    a, g = poles(orc)
    # log-spaced from 1e-4 Hz to 64 Hz: a = exp(-2 pi f / 128); the lowest
    # pole's a is within 5e-6 of 1, so f comes back to ~1e-11 only
    f = -np.log1p(a - 1.0) * 128 / (2 * math.pi)
    np.testing.assert_allclose(f[0], 1e-4, rtol=1e-9)
    np.testing.assert_allclose(f[-1], 64.0, rtol=1e-12)
    np.testing.assert_allclose(np.diff(np.log(f)), np.log(64 / 1e-4) / 19, rtol=1e-8)
    np.testing.assert_allclose(a * a + g * g, 1.0, rtol=1e-15)   # unit-variance drive


def test_exact_calibration_matches_simulation(orc):
    """orc_flicker_unit_adev (the estimator's expectation) against the same
    estimator on a long simulated realisation of the pole sum."""
    a, g = poles(orc)
    exact = orc.orc_flicker_unit_adev(a.ctypes.data_as(P(C.c_double)), 256)
    from scipy.signal import lfilter
    rng = np.random.default_rng(3)
    n = 1 << 20
    y = np.zeros(n)
    for j in range(20):
        # s_k = a s_{k-1} + g z_k from a stationary start s_{-1} ~ N(0, 1)
        z = rng.standard_normal(n)
        s, _ = lfilter([g[j]], [1.0, -a[j]], z, zi=[a[j] * rng.standard_normal()])
        y += s
    delta = np.concatenate([[0.0], np.cumsum(y) / 128.0])
    sim = allan(delta, 256, 128.0)
    # 4095 non-overlapping pairs: ~1.1 % standard error on the deviation
    assert abs(sim / exact - 1.0) < 0.05, (sim, exact)


def _flicker_track(orc, floor, n, seed=42, adev=0.0, y0=0.0, c0=0.0):
    rx = pyoracle.OrcRx()
    rx.pos_ecef[:] = [6378137.0, 0.0, 0.0]
    rx.allan_deviation_1s, rx.frequency_offset_y0, rx.initial_time_offset_s = adev, y0, c0
    rx.master_rate_hz, rx.clock_seed, rx.flicker_floor = 128.0, seed, floor
    c = pyoracle.OrcClock()
    orc.orc_clock_init(C.byref(c), C.byref(rx))
    return np.array([orc.orc_clock_delta_t_at_master_index(C.byref(c), k) for k in range(n)])


def test_flicker_has_the_reference_flat_floor(orc):
    """src/clock/mod.rs:355-372: pure flicker at 1e-10, master rate = 128 Hz,
    2^18 samples: sigma_y(2 s) in [0.3, 3] e-10 and sigma_y(8)/sigma_y(1) in
    (0.45, 2.2); with the exact calibration the level is also within 10 %."""
    delta = _flicker_track(orc, 1e-10, 1 << 18)
    a1, a2, a8 = (allan(delta, int(t * 128), 128.0) for t in (1.0, 2.0, 8.0))
    assert 0.3e-10 <= a2 <= 3.0e-10, a2
    assert 0.45 < a8 / a1 < 2.2, (a1, a8)
    assert abs(a2 / 1e-10 - 1.0) < 0.1, a2
    # white FM would fall as tau^-1/2: flicker stays flat over 1..16 s
    a16 = allan(delta, 16 * 128, 128.0)
    assert 0.6 < a16 / a1 < 1.6


def test_flicker_adds_to_the_clock_composition(orc):
    """eps + flicker + (c + y0 t): with the walk off, the track minus the
    deterministic part is the flicker integral, independent of c and y0;
    other seeds give other tracks; floor 0 adds nothing."""
    n = 4096
    base = _flicker_track(orc, 1e-10, n, seed=7)
    shifted = _flicker_track(orc, 1e-10, n, seed=7, y0=2e-7, c0=1e-3)
    t = np.arange(n) / 128.0
    np.testing.assert_allclose(shifted - (1e-3 + 2e-7 * t), base, atol=1e-18)
    other = _flicker_track(orc, 1e-10, n, seed=8)
    assert np.abs(other - base).max() > 1e-12
    off = _flicker_track(orc, 0.0, n, seed=7, y0=2e-7, c0=1e-3)
    np.testing.assert_array_equal(off, (1e-3 + 2e-7 * (np.arange(n) / 128.0)) + 0.0)
