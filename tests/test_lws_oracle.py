"""Pins of the learned warm start's CPU oracle (oracle/lws_oracle.py), on the
CPU. The reference has no network, so each function is pinned by mathematics
or by the reference's own restated propagator:
  - back-propagation against central finite differences of the unrounded
    network (a dropped term, a wrong sign or a transposed operand fails);
  - the light time against an independent bisection of the emission
    equation with positions from orc_kepler_position (src/geometry/kepler.rs);
  - the (node, argument of latitude) parametrisation against that propagator;
  - BF16 rounding against exact binary cases.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import lws_oracle as L
from oracle import pyoracle

RAD = math.pi / 180.0


@pytest.fixture(scope="module")
def orc():
    return pyoracle.load()


def kepler_ref(orc, a, inc_deg, raan_deg, nu0_deg, epoch, t):
    out = (C.c_double * 3)()
    orc.orc_kepler_position(a, inc_deg, raan_deg, nu0_deg, epoch, t, out)
    return np.array(out[:])


def random_geometry(rng, n):
    lat = rng.uniform(-90, 90, n) * RAD
    lon = rng.uniform(-180, 180, n) * RAD
    h = rng.uniform(0, 2000, n)
    f = 1 / 298.257223563
    e2 = f * (2 - f)
    nn = 6378137.0 / np.sqrt(1 - e2 * np.sin(lat) ** 2)
    obs = np.stack([(nn + h) * np.cos(lat) * np.cos(lon), (nn + h) * np.cos(lat) * np.sin(lon),
                    (nn * (1 - e2) + h) * np.sin(lat)], axis=-1)
    leo = rng.uniform(size=n) < 0.5
    a = np.where(leo, rng.uniform(6.9e6, 7.5e6, n), rng.uniform(2.6e7, 2.7e7, n))
    return obs, a, rng.uniform(0, math.pi, n), rng.uniform(0, 2 * math.pi, n), rng.uniform(0, 2 * math.pi, n)


def test_bf16_rounding_exact_cases():
    one = 1.0
    ulp = 2.0 ** -7                          # BF16 spacing just above 1
    cases = {one: one, one + ulp: one + ulp,
             one + ulp / 2: one,             # tie, to even (1.0)
             one + 3 * ulp / 2: one + 2 * ulp,   # tie, to even (1 + 2 ulp)
             one + ulp / 2 + 2.0 ** -20: one + ulp,
             -3.0: -3.0, 0.0: 0.0}
    for x, want in cases.items():
        assert L.bf16(np.float32(x)) == want, x


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(5)
    th = rng.normal(0, 0.3, L.N_PARAMS)
    X = rng.uniform(-1, 1, (6, L.N_IN)).astype(np.float32)
    Y = rng.normal(0, 0.5, 6)
    loss, g = L.loss_and_grad(th, X, Y, quantize=False)
    h = 1e-6
    fd = np.empty(L.N_PARAMS)
    for p in range(L.N_PARAMS):
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (L.loss_and_grad(tp, X, Y, quantize=False)[0] - L.loss_and_grad(tm, X, Y, quantize=False)[0]) / (2 * h)
    err = np.abs(fd - g)
    assert err.max() <= 1e-7 + 1e-6 * np.abs(g).max(), err.max()
    # every block carries gradient (no block silently zero)
    for lo, hi in ((L.OFF_W1, L.OFF_B1), (L.OFF_B1, L.OFF_W2), (L.OFF_W2, L.OFF_B2), (L.OFF_B2, L.OFF_W3),
                   (L.OFF_W3, L.OFF_B3), (L.OFF_B3, L.OFF_W4), (L.OFF_W4, L.OFF_B4), (L.OFF_B4, L.N_PARAMS)):
        assert np.abs(g[lo:hi]).max() > 0


def test_gelu_grad_is_the_derivative():
    z = np.linspace(-6, 6, 2001)
    fd = (L.gelu(z + 1e-6) - L.gelu(z - 1e-6)) / 2e-6
    assert np.abs(fd - L.gelu_grad(z)).max() < 1e-8
    # and the tanh form is within 1e-3 of the erf definition z Phi(z)
    from math import erf
    exact = np.array([0.5 * v * (1 + erf(v / math.sqrt(2))) for v in z])
    assert np.abs(L.gelu(z) - exact).max() < 1e-3


def test_quantised_network_stays_near_the_exact_one():
    rng = np.random.default_rng(9)
    th = rng.normal(0, 0.2, L.N_PARAMS)
    X = rng.uniform(-1, 1, (64, L.N_IN)).astype(np.float32)
    yq, _ = L.forward(th, X, quantize=True)
    ye, _ = L.forward(th, X, quantize=False)
    scale = np.abs(ye).max()
    assert np.abs(yq - ye).max() < 0.05 * scale
    assert np.abs(yq - ye).max() > 0     # the roundings are there


def test_parametrisation_matches_reference_propagator(orc):
    """P(node, u) with node = raan - w_E (t - t_el), u = nu0 + n (t - t_el)
    equals the restated src/geometry/kepler.rs position."""
    rng = np.random.default_rng(11)
    for _ in range(40):
        a = rng.choice([rng.uniform(6.9e6, 7.5e6), rng.uniform(2.6e7, 2.7e7)])
        inc, raan, nu0 = rng.uniform(0, 180), rng.uniform(0, 360), rng.uniform(0, 360)
        t_el = 1717260840.0 + rng.uniform(-1e5, 1e5)
        t = 1717260840.0 + rng.uniform(0, 120)
        n = math.sqrt(L.GM_EARTH / a ** 3)
        node = raan * RAD - L.OMEGA_E * (t - t_el)
        u = nu0 * RAD + n * (t - t_el)
        p = L.orbit_position(a, inc * RAD, node, u)
        np.testing.assert_allclose(p, kepler_ref(orc, a, inc, raan, nu0, t_el, t), atol=2e-6)


def test_light_time_is_the_root_found_by_bisection(orc):
    """The converged light time solves tau = |R(w tau) P(t - tau) - O| / c with
    P from the reference propagator, found here by bisection instead."""
    rng = np.random.default_rng(13)
    obs, a, inc, node, u = random_geometry(rng, 24)
    tau_sl, tau = L.light_time(obs, a, inc, node, u)
    # element epoch and query time 0, so the propagator's time argument -tau is
    # exact (at a Unix epoch its rounding alone moves a LEO orbit by ~2 mm)
    T = 0.0
    for k in range(24):
        # elements that put the orbit in state (node, u) at time T (t_el = T)
        args = (a[k], inc[k] / RAD, node[k] / RAD, u[k] / RAD, T)

        def f(tt):
            p = kepler_ref(orc, *args, T - tt)
            w = L.OMEGA_E * tt
            d = np.array([p[0] + w * p[1], p[1] - w * p[0], p[2]]) - obs[k]
            return tt - math.sqrt(float(d @ d)) / L.C_LIGHT

        lo, hi = 0.0, 0.2
        assert f(lo) < 0 < f(hi)
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if f(mid) < 0 else (lo, mid)
        assert abs(tau[k] - 0.5 * (lo + hi)) < 2e-15, (k, tau[k], lo)
        # the straight-line value differs by the motion during the flight
        assert 1e-10 < abs(tau[k] - tau_sl[k]) < 3e-6


def test_features_are_unit_scale():
    rng = np.random.default_rng(17)
    obs, a, inc, node, u = random_geometry(rng, 1000)
    X = L.features(obs, a, inc, node, u)
    assert X.dtype == np.float32 and X.shape == (1000, 10)
    assert np.abs(X).max() <= 1.01
