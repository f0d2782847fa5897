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


# ---------------------------------------------------------------- converged solve (orc_solve_converged)

def _grid_tx(orc, orbit, t0, dur, dt=0.01):
    n = orc.orc_grid_size(dur, dt)
    xyz = np.empty(3 * n)
    t0o = C.c_double()
    orc.orc_grid_sample_kepler(*orbit, t0, dur, dt, C.byref(t0o), xyz.ctypes.data_as(C.POINTER(C.c_double)))
    tx = pyoracle.OrcTx()
    tx.traj_kind, tx.grid_t0_unix, tx.grid_dt_s, tx.grid_n = 1, t0o.value, dt, n
    tx.grid_xyz = xyz.ctypes.data_as(C.POINTER(C.c_double))
    return tx, xyz


def _solve(orc, tx, obs, tu, epoch, start, max_evals):
    te, r = C.c_double(), C.c_double()
    o = (C.c_double * 3)(*obs)
    e = orc.orc_solve_converged(C.byref(tx), o, tu, epoch, start, 1e-13, max_evals, C.byref(te), C.byref(r))
    return e, te.value, r.value


def test_converged_solve_is_the_reference_iteration(orc):
    """Cut at two evaluations from t - 0.07 s, the converged solve is the
    reference's two-iteration map bit for bit (a stationary source, where the
    absolute-time rounding cannot differ)."""
    tx = pyoracle.OrcTx()
    tx.traj_kind = 0
    tx.pos_ecef[:] = [6378137.0, 1000.0, 1000.0]
    obs = [6378137.0, 0.0, 0.0]
    for tu in (0.0, 0.37, 12.5):
        e, te, r = _solve(orc, tx, obs, tu, 1717260840.0, tu - 0.07, 2)
        te_ref, r_ref = C.c_double(), C.c_double()
        orc.orc_solve_emission(C.byref(tx), (C.c_double * 3)(*obs), tu, 1717260840.0, 2, C.byref(te_ref),
                               C.byref(r_ref))
        assert e == -1                      # two evaluations do not reach 1e-13 s from 0.07 s
        assert te == te_ref.value and r == r_ref.value


def test_converged_solve_finds_the_root_found_by_bisection(orc):
    """The converged emission time solves t + |R(w (tu - t)) P(t) - O| / c = tu
    with P the grid lerp, against a bisection of that equation; scene epoch 0
    so the absolute query times are exact."""
    rng = np.random.default_rng(23)
    for k in range(12):
        a = 6.9e6 if k % 2 == 0 else 2.66e7
        orbit = (a, float(rng.uniform(0, 180)), float(rng.uniform(0, 360)), float(rng.uniform(0, 360)), 0.0)
        tx, keep = _grid_tx(orc, orbit, 0.0, 20.0)
        lat, lon = rng.uniform(-60, 60), rng.uniform(-180, 180)
        o = (C.c_double * 3)()
        orc.orc_geodetic_to_ecef(lat, lon, 50.0, o)
        obs = list(o[:])
        tu = float(rng.uniform(1.0, 19.0))
        e, te, r = _solve(orc, tx, obs, tu, 0.0, tu - 0.07, 12)
        assert 2 <= e <= 5
        assert abs((tu - te) - r / L.C_LIGHT) <= 2 * math.ulp(tu)   # t_emit = tu - r / c, rounded

        def g(t):
            p = (C.c_double * 3)()
            orc.orc_grid_position(tx.grid_t0_unix, tx.grid_dt_s, tx.grid_n, keep.ctypes.data_as(C.POINTER(C.c_double)),
                                  t, p)
            w = L.OMEGA_E * (tu - t)
            d = np.array([p[0] + w * p[1], p[1] - w * p[0], p[2]]) - np.array(obs)
            return t + math.sqrt(float(d @ d)) / L.C_LIGHT - tu

        lo, hi = tu - 0.2, tu
        assert g(lo) < 0 < g(hi)
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if g(mid) < 0 else (lo, mid)
        assert abs(te - 0.5 * (lo + hi)) <= 4e-15, (k, te - 0.5 * (lo + hi))
        # the straight two-iteration map stops short of it
        te2 = C.c_double()
        r2 = C.c_double()
        orc.orc_solve_emission(C.byref(tx), o, tu, 0.0, 2, C.byref(te2), C.byref(r2))
        assert te2.value != te or k % 2 == 1
