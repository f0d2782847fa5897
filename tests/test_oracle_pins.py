"""Pins of the CPU oracle: the reference's own known-answer values
(tests/golden/reference_kats.json, transcribed from its tests), public
specification tables, executable third-party implementations present here
(numpy's Philox), and what the mathematics fixes (closed forms, invariants,
brute force). Each check is chosen so that a plausible slip -- a sign, an
index, a transposed rotation, a dropped term -- fails it."""
import ctypes as C
import math

import numpy as np
import pytest
from numpy.random import Philox

import util

P = C.POINTER
D3 = C.c_double * 3


def geod(orc, lat, lon, h):
    out = D3()
    orc.orc_geodetic_to_ecef(lat, lon, h, out)
    return np.array(out[:])


def kepler(orc, orbit, t):
    out = D3()
    orc.orc_kepler_position(*orbit, t, out)
    return np.array(out[:])


def grid(orc, orbit, t0, dur, dt):
    n = orc.orc_grid_size(dur, dt)
    xyz = np.empty(3 * n)
    t0o = C.c_double()
    orc.orc_grid_sample_kepler(*orbit, t0, dur, dt, C.byref(t0o), xyz.ctypes.data_as(P(C.c_double)))
    return t0o.value, n, xyz


def grid_pos(orc, g, dt, t):
    t0, n, xyz = g
    out = D3()
    orc.orc_grid_position(t0, dt, n, xyz.ctypes.data_as(P(C.c_double)), t, out)
    return np.array(out[:])


# ---------------------------------------------------------------- geometry

def test_geodetic_known_answers(orc, kats):
    k = kats["geodetic_glasgow"]
    np.testing.assert_allclose(geod(orc, *k["input"]), k["ecef"], rtol=k["max_relative"])
    e = kats["geodetic_equator"]
    p = geod(orc, *e["input"])
    assert abs(p[0] / 6378137.0 - 1) < 1e-12 and abs(p[1]) < 1e-6 and abs(p[2]) < 1e-6


def test_geodetic_poles_and_ellipsoid():
    """The pole sits at the semi-minor axis b = a(1-f); any latitude's point lies on the ellipsoid."""
    from oracle import pyoracle
    orc = pyoracle.load()
    a, f = 6378137.0, 1 / 298.257223563
    b = a * (1 - f)
    p = geod(orc, 90.0, 0.0, 0.0)
    assert abs(p[2] - b) < 1e-6 and abs(math.hypot(p[0], p[1])) < 1e-6
    for lat, lon in ((12.3, 45.6), (-70.1, -120.0), (55.86, -4.25)):
        q = geod(orc, lat, lon, 0.0)
        assert abs((q[0] ** 2 + q[1] ** 2) / a ** 2 + q[2] ** 2 / b ** 2 - 1) < 1e-12
        up = geod(orc, lat, lon, 1000.0) - q  # height runs along the geodetic normal
        n = np.array([math.cos(math.radians(lat)) * math.cos(math.radians(lon)),
                      math.cos(math.radians(lat)) * math.sin(math.radians(lon)), math.sin(math.radians(lat))])
        np.testing.assert_allclose(up, 1000.0 * n, atol=1e-6)


def test_kepler_radius_and_closed_form(orc, kats):
    k = kats["kepler_gps_radius"]
    p = kepler(orc, k["orbit"], k["t_unix"])
    assert abs(np.linalg.norm(p) / k["orbit"][0] - 1) < k["max_relative"]
    # Textbook two-body circular orbit: ECI = Rz(RAAN) Rx(i) [a cos nu, a sin nu, 0],
    # ECEF = Rz(-omega_E (t - t_epoch)) ECI, nu = nu0 + sqrt(GM/a^3)(t - t_epoch).
    gm, we = 3.986005e14, 7.2921151467e-5

    def rz(t):
        return np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])

    def rx(t):
        return np.array([[1, 0, 0], [0, math.cos(t), -math.sin(t)], [0, math.sin(t), math.cos(t)]])

    for orbit, dt in (((26578137.0, 55.0, 60.0, 80.0, 1.7e9), 1234.5), ((6778137.0, 51.6, 200.0, 10.0, 1.7e9), 37.0)):
        a, inc, raan, nu0, ep = orbit
        nu = math.radians(nu0) + math.sqrt(gm / a ** 3) * dt
        eci = rz(math.radians(raan)) @ rx(math.radians(inc)) @ np.array([a * math.cos(nu), a * math.sin(nu), 0.0])
        ecef = rz(-we * dt) @ eci
        np.testing.assert_allclose(kepler(orc, orbit, ep + dt), ecef, atol=1e-6)


def test_grid_matches_orbit_and_lerp_identities(orc, kats):
    k = kats["grid_vs_orbit"]
    g = grid(orc, k["orbit"], k["t0"], k["duration_s"], k["dt"])
    t0, n, xyz = g
    assert t0 == k["t0"] - 0.5 and n == math.ceil((k["duration_s"] + 1.0) / k["dt"]) + 1
    for j in range(50):
        t = k["t0"] + 60.0 * j / 50
        np.testing.assert_allclose(grid_pos(orc, g, k["dt"], t), kepler(orc, k["orbit"], t),
                                   rtol=k["max_relative"], atol=k["epsilon_m"])
    # nodes hold the orbit; the midpoint is the average of its two nodes;
    # queries beyond both ends hold the end nodes.
    for i in (0, 17, n // 2, n - 2):
        node = np.array([xyz[i], xyz[n + i], xyz[2 * n + i]])
        nxt = np.array([xyz[i + 1], xyz[n + i + 1], xyz[2 * n + i + 1]])
        np.testing.assert_allclose(node, kepler(orc, k["orbit"], t0 + i * k["dt"]), atol=1e-7)
        np.testing.assert_allclose(grid_pos(orc, g, k["dt"], t0 + (i + 0.5) * k["dt"]), 0.5 * (node + nxt), atol=1e-6)
    first = np.array([xyz[0], xyz[n], xyz[2 * n]])
    last = np.array([xyz[n - 1], xyz[2 * n - 1], xyz[3 * n - 1]])
    np.testing.assert_array_equal(grid_pos(orc, g, k["dt"], t0 - 100.0), first)
    np.testing.assert_array_equal(grid_pos(orc, g, k["dt"], t0 + 1e6), last)


def test_grid_lipschitz_bound(orc, kats):
    k = kats["grid_path_bounds"]
    g = grid(orc, k["orbit"], k["t0"], k["duration_s"], k["dt"])
    t0, n, xyz = g
    pts = xyz.reshape(3, n)
    seg = np.max(np.linalg.norm(np.diff(pts, axis=1), axis=0)) / k["dt"]
    r = np.max(np.linalg.norm(pts, axis=0))
    lo, hi = k["speed_range"]
    assert lo < seg < hi and abs(r / 6928137.0 - 1) < 1e-9
    prev = None
    for j in range(4000):
        t = k["t0"] - 2.0 + j * 0.0087
        p = grid_pos(orc, g, k["dt"], t)
        assert np.linalg.norm(p) <= r * (1 + 1e-12)
        if prev is not None:
            assert np.linalg.norm(p - prev[1]) <= seg * (t - prev[0]) * (1 + 1e-9) + 1e-6
        prev = (t, p)


# ---------------------------------------------------------------- emission solve

def _stationary(pos):
    t = util.orc_tx(0)
    t.pos_ecef[:] = pos
    return t


def solve(orc, tx, obs, tu, epoch, iters):
    te, r = C.c_double(), C.c_double()
    orc.orc_solve_emission(C.byref(tx), D3(*obs), tu, epoch, iters, C.byref(te), C.byref(r))
    return te.value, r.value


def test_sagnac_known_answer(orc, kats):
    k = kats["sagnac"]
    _, r = solve(orc, _stationary(k["sat"]), k["obs"], 0.0, k["epoch"], k["iterations"])
    s, o = np.array(k["sat"]), np.array(k["obs"])
    naive = np.linalg.norm(s - o)
    th = 7.2921151467e-5 * r / 299792458.0
    sr = np.array([s[0] * math.cos(th) + s[1] * math.sin(th), -s[0] * math.sin(th) + s[1] * math.cos(th), s[2]])
    assert abs(r - naive) > k["min_vs_naive_m"]
    assert abs(r - np.linalg.norm(sr - o)) < k["max_vs_exact_m"]


def test_stationary_pair_delay(orc, kats):
    k = kats["stationary_pair_delay"]
    te, r = solve(orc, _stationary(k["tx"]), k["rx"], 0.0, k["epoch"], k["iterations"])
    assert abs(r / 299792458.0 - 300000.0 / 299792458.0) < k["delay_tolerance_s"]
    assert abs((0.0 - te) - r / 299792458.0) < 1e-15


def test_solve_on_rotation_axis_is_exact(orc):
    """On the Earth's axis the Sagnac rotation leaves both points fixed, so every
    iterate has r = |dz| exactly and t_emit = tu - r/c."""
    for iters in (1, 2, 4):
        te, r = solve(orc, _stationary([0.0, 0.0, 2.0e7]), [0.0, 0.0, 6.4e6], 3.25, 1.7e9, iters)
        assert r == 2.0e7 - 6.4e6
        assert te == 3.25 - r * (1.0 / 299792458.0)


def test_two_iteration_map_brute_force(orc):
    """Two fixed-point iterations from tu - 0.07 s for a source moving on a
    straight line (exactly representable by a 2-node grid), restated with
    exact rational arithmetic: agrees to f64 rounding."""
    from fractions import Fraction as F
    c, w = F(299792458), F(7.2921151467e-5)
    p0, v = [2.0e7, -1.5e7, 9.0e6], [1200.0, 3100.0, -800.0]
    t0g, dtg = -5.0, 10.0
    xyz = np.array([[p0[0], p0[0] + v[0] * dtg], [p0[1], p0[1] + v[1] * dtg], [p0[2], p0[2] + v[2] * dtg]]).ravel()
    tx = util.orc_tx(0)
    tx.traj_kind, tx.grid_t0_unix, tx.grid_dt_s, tx.grid_n = 1, t0g, dtg, 2
    arr = util.arr(xyz, np.float64)
    tx.grid_xyz = arr.ctypes.data_as(P(C.c_double))
    obs, epoch, tu = [3.9e6, 3.3e5, 5.0e6], 0.0, 0.4
    te, r = solve(orc, tx, obs, tu, epoch, 2)
    t = F(tu) - F(7, 100)
    rr = None
    for _ in range(2):
        frac = (F(epoch) + t - F(t0g)) / F(dtg)
        s = [F(p0[i]) + F(v[i]) * F(dtg) * frac for i in range(3)]
        wt = w * (F(tu) - t)
        sx, sy = s[0] + wt * s[1], s[1] - wt * s[0]
        d2 = (sx - F(obs[0])) ** 2 + (sy - F(obs[1])) ** 2 + (s[2] - F(obs[2])) ** 2
        rr = math.sqrt(float(d2))
        t = F(tu) - F(rr) / c
    assert abs(r - rr) < 1e-7            # metres: f64 rounding of 2e7 m
    assert abs(te - float(t)) < 1e-15


# ---------------------------------------------------------------- anchors

def test_elapsed_periods_known_answers(orc, kats):
    k = kats["elapsed_periods"]
    w, f = C.c_int64(), C.c_double()
    for e, r, whole, frac in k["exact"]:
        assert orc.orc_elapsed_periods(e, r, C.byref(w), C.byref(f)) == 0
        assert (w.value, f.value) == (whole, frac)
    e, r, whole, frac = k["fraction"][0]
    assert orc.orc_elapsed_periods(e + k["fraction_elapsed_offset"], r, C.byref(w), C.byref(f)) == 0
    assert w.value == whole and abs(f.value - frac) < k["fraction_tolerance"]
    c = k["consistency"]
    w1, f1 = C.c_int64(), C.c_double()
    orc.orc_elapsed_periods(c["e0"], 9600.0 / 7.0, C.byref(w), C.byref(f))
    orc.orc_elapsed_periods(c["e1"], 9600.0 / 7.0, C.byref(w1), C.byref(f1))
    assert w1.value - w.value == c["whole_step"] and abs(f1.value - f.value) < c["frac_tolerance"]
    assert orc.orc_elapsed_periods(float("inf"), 1.0, C.byref(w), C.byref(f)) != 0
    assert orc.orc_elapsed_periods(1.0e12, 1.0e12, C.byref(w), C.byref(f)) != 0


def test_frame_anchor_and_cw_anchor(orc, kats):
    k = kats["frame_anchor_bpsk"]
    rel, off = C.c_double(), C.c_int64()
    assert orc.orc_frame_anchor(k["render_epoch"], 1.0e9 - 0.64 - 0.0025, k["rate"], k["n_bits"],
                                C.byref(rel), C.byref(off)) == 0
    assert off.value == k["repeat_from"] and abs(rel.value - k["epoch_rel_s"]) < k["epoch_rel_tolerance"]
    assert orc.orc_frame_anchor(1.0, 0.0, 0.0, 4, C.byref(rel), C.byref(off)) != 0
    z = kats["cw_anchor_zero"]
    ph = C.c_double()
    assert orc.orc_cw_anchor_phase(z["epoch"], z["epoch"], z["offset_hz"], C.byref(ph)) == 0 and ph.value == 0.0


# ---------------------------------------------------------------- C/A and NAV

def ca(orc, prn):
    out = np.empty(1023, dtype=np.int8)
    assert orc.orc_ca_code(prn, out.ctypes.data_as(P(C.c_int8))) == 0
    return out


def test_ca_first_chips_against_icd(orc, kats):
    assert list(ca(orc, 1)[:10]) == kats["ca_prn1_first10"]["chips"]
    for prn, octal in enumerate(kats["ca_first10_octal"]["octal"], start=1):
        bits = [int(b) for b in format(int(str(octal), 8), "010b")]
        assert list(ca(orc, prn)[:10]) == [1 - 2 * b for b in bits], prn
    bad = np.empty(1023, dtype=np.int8)
    assert orc.orc_ca_code(0, bad.ctypes.data_as(P(C.c_int8))) != 0
    assert orc.orc_ca_code(33, bad.ctypes.data_as(P(C.c_int8))) != 0


def test_ca_gold_correlation_property(orc):
    """Periodic auto- and cross-correlation of 1023-chip Gold codes take only
    the values {-65, -1, 63} off the main peak."""
    codes = np.array([ca(orc, p) for p in range(1, 33)], dtype=np.int64)
    f = np.fft.fft(codes, axis=1)
    allowed = {-65, -1, 63}
    for i in range(32):
        auto = np.rint(np.fft.ifft(f[i] * np.conj(f[i])).real).astype(int)
        assert auto[0] == 1023 and set(auto[1:]) <= allowed
        for j in range(i + 1, 32, 5):
            cross = np.rint(np.fft.ifft(f[i] * np.conj(f[j])).real).astype(int)
            assert set(cross) <= allowed


def test_nav_preamble_and_payload(orc, kats):
    k = kats["nav_preamble"]
    for sf in range(3):
        got = [orc.orc_nav_dummy_bit_pm(k["seed"], sf * k["subframe_bits"] + i) for i in range(8)]
        assert got == k["pm"]
    payload = [orc.orc_nav_dummy_bit_pm(kats["nav_payload_flips"]["seed"], i) for i in range(8, 300)]
    assert 1 in payload and -1 in payload
    # negative bit indices wrap into the subframe (rem_euclid)
    assert [orc.orc_nav_dummy_bit_pm(5, -300 + i) for i in range(8)] == k["pm"]


# ---------------------------------------------------------------- framed signals

def test_qam16_constellation(orc, kats):
    k = kats["qam16"]
    sync = np.array([b for v in range(16) for b in ((v >> 3) & 1, (v >> 2) & 1, (v >> 1) & 1, v & 1)], dtype=np.uint8)
    out = np.empty(32, dtype=np.float32)
    assert orc.orc_payload_qam16(sync.ctypes.data_as(P(C.c_uint8)), 64, 0, 0, out.ctypes.data_as(P(C.c_float))) == 0
    pts = out[0::2] + 1j * out[1::2]
    levels = np.array(k["levels"]) / math.sqrt(k["scale_sqrt"])
    for p in pts:
        assert np.min(np.abs(levels - p.real)) < k["tolerance"] and np.min(np.abs(levels - p.imag)) < k["tolerance"]
    assert len({(round(p.real, 5), round(p.imag, 5)) for p in pts}) == 16
    assert abs(np.mean(np.abs(pts) ** 2) - k["mean_power"]) < k["tolerance"]
    step = 2 / math.sqrt(10)
    for a in range(16):
        for b in range(16):
            if abs(abs(pts[a] - pts[b]) - step) < 1e-5:
                assert bin(a ^ b).count("1") == 1


def test_qam16_payload_and_frame_edges(orc, kats):
    k = kats["qam16_payload"]
    sync = np.array(k["sync"], dtype=np.uint8)
    n_sym = len(sync) // 4 + k["n_payload"]
    out = np.empty(2 * n_sym, dtype=np.float32)
    assert orc.orc_payload_qam16(sync.ctypes.data_as(P(C.c_uint8)), len(sync), k["n_payload"], k["seed"],
                                 out.ctypes.data_as(P(C.c_float))) == 0
    pay = (out[0::2] + 1j * out[1::2])[2:]
    assert abs(np.mean(np.abs(pay) ** 2) - 1) < k["power_tolerance"]
    assert len({(round(p.real, 4), round(p.imag, 4)) for p in pay}) == 16
    tx = util.orc_tx(3, rate_hz=k["rate"], symbols=out.reshape(-1, 2))
    s = util.baseband(orc, tx, k["times"])
    a = 1 / math.sqrt(10)
    for v, lev in zip(s, k["expect_level"]):
        if lev is None:
            assert v == 0
        else:
            assert abs(v.real - lev * a) < 1e-6 and abs(v.imag - lev * a) < 1e-6
    bad = np.array([1, 0, 2, 0], dtype=np.uint8)
    assert orc.orc_payload_qam16(bad.ctypes.data_as(P(C.c_uint8)), 4, 1, 0, out.ctypes.data_as(P(C.c_float))) != 0
    assert orc.orc_payload_qam16(bad.ctypes.data_as(P(C.c_uint8)), 3, 1, 0, out.ctypes.data_as(P(C.c_float))) != 0


def test_ook_mid_bit_and_silence(orc, kats):
    k = kats["ook_frame"]
    sync = np.array(k["sync"], dtype=np.uint8)
    bits = np.empty(len(sync) + k["n_payload"], dtype=np.int8)
    orc.orc_payload_ook(sync.ctypes.data_as(P(C.c_uint8)), len(sync), k["n_payload"], k["seed"],
                        bits.ctypes.data_as(P(C.c_int8)))
    assert list(bits[:len(sync)]) == k["sync"]
    ones = int(np.sum(bits[len(sync):]))
    assert k["ones_range"][0] < ones < k["ones_range"][1]
    tx = util.orc_tx(2, rate_hz=k["rate"], bits=bits)
    s = util.baseband(orc, tx, (np.arange(len(bits)) + 0.5) * 1e-3)
    np.testing.assert_array_equal(s, bits.astype(np.float32))
    q = k["silent"]
    sync2 = np.array(q["sync"], dtype=np.uint8)
    bits2 = np.empty(len(sync2) + q["n_payload"], dtype=np.int8)
    orc.orc_payload_ook(sync2.ctypes.data_as(P(C.c_uint8)), len(sync2), q["n_payload"], q["seed"],
                        bits2.ctypes.data_as(P(C.c_int8)))
    s2 = util.baseband(orc, util.orc_tx(2, rate_hz=1000.0, bits=bits2), q["times"])
    assert list(s2.real) == [0.0, 1.0, float(bits2[7]), 0.0, 0.0]
    assert orc.orc_payload_ook(np.array([1, 0, 2], dtype=np.uint8).ctypes.data_as(P(C.c_uint8)), 3, 1, 0,
                               bits2.ctypes.data_as(P(C.c_int8))) != 0


def test_anchored_bpsk_repeats_for_all_time(orc):
    """src/signal/mod.rs:170-198: a repeating frame anchored at the render epoch."""
    sync = np.array([-1, 1, 1, -1, 1], dtype=np.int8)
    bits = np.empty(64, dtype=np.int8)
    orc.orc_payload_bpsk(sync.ctypes.data_as(P(C.c_int8)), 5, 59, 11, bits.ctypes.data_as(P(C.c_int8)))
    rel, off = C.c_double(), C.c_int64()
    orc.orc_frame_anchor(1.0e9, 1.0e9, 1000.0, 64, C.byref(rel), C.byref(off))
    tx = util.orc_tx(1, rate_hz=1000.0, bits=bits, repeating=1, repeat_from=off.value, epoch_rel_s=rel.value)
    ks = [-200, -65, -64, -1, 0, 1, 63, 64, 65, 1_000_003]
    s = util.baseband(orc, tx, [(k + 0.5) * 1e-3 for k in ks])
    assert list(s.real) == [float(bits[k % 64]) for k in ks]
    orc.orc_frame_anchor(1.0e9, 1.0e9 - 0.64 - 0.0025, 1000.0, 64, C.byref(rel), C.byref(off))
    tx2 = util.orc_tx(1, rate_hz=1000.0, bits=bits, repeating=1, repeat_from=off.value, epoch_rel_s=rel.value)
    s2 = util.baseband(orc, tx2, [0.0, 0.0003, 0.0007])
    assert list(s2.real) == [float(bits[2]), float(bits[2]), float(bits[3])]


def test_cw_tone(orc, kats):
    k = kats["cw_tone"]
    s = util.baseband(orc, util.orc_tx(0, baseband_offset_hz=k["offset_hz"]), k["times"])
    for t, v in zip(k["times"], s):
        assert abs(v.real - math.cos(2 * math.pi * k["offset_hz"] * t)) < k["epsilon"]
        assert abs(v.imag - math.sin(2 * math.pi * k["offset_hz"] * t)) < k["epsilon"]


def test_gate_truth_tables(orc, kats):
    k = kats["gate_truth_table"]
    bursts = [(o, o + d) for o, d in k["bursts"]]
    tx = util.orc_tx(0, gated=1, gate_period_s=k["period"], gate_phase_s=0.0, bursts=bursts)
    for t, on in k["cases"]:
        assert orc.orc_gate_is_on(C.byref(tx), t) == on, t
    for per in (-3.0, -1.0, 1.0, 2.0, 1.0e6):
        assert orc.orc_gate_is_on(C.byref(tx), 15 * per + 0.25) and not orc.orc_gate_is_on(C.byref(tx), 15 * per + 0.75)
    a = kats["gate_absolute"]
    phase = math.fmod(a["epoch"] + a["render_offset"] - 0.0, 15.0)
    tx2 = util.orc_tx(0, gated=1, gate_period_s=15.0, gate_phase_s=phase, bursts=bursts)
    for t, on in a["cases"]:
        assert orc.orc_gate_is_on(C.byref(tx2), t) == on, t


# ---------------------------------------------------------------- payload RNG restatement

def test_xoshiro_reference_vector(orc, kats):
    k = kats["xoshiro256pp_reference"]
    s = (C.c_uint64 * 4)(*k["state"])
    assert [orc.orc_xoshiro_next(s) for _ in range(10)] == k["outputs"]


def test_payload_statistics(orc):
    """gen_bool(0.5) and gen_range(0..16) restated: unbiased over a long draw."""
    n = 200_000
    bits = np.empty(n, dtype=np.int8)
    orc.orc_payload_bpsk(None, 0, n, 2026, bits.ctypes.data_as(P(C.c_int8)))
    assert set(np.unique(bits)) == {-1, 1} and abs(np.mean(bits)) < 0.01
    sym = np.empty(2 * n, dtype=np.float32)
    orc.orc_payload_qam16(None, 0, n, 7, sym.ctypes.data_as(P(C.c_float)))
    _, counts = np.unique(np.round(sym[0::2] * 1000) * 100 + np.round(sym[1::2] * 1000), return_counts=True)
    assert len(counts) == 16 and np.max(np.abs(counts / (n / 16) - 1)) < 0.05


# ---------------------------------------------------------------- project RNG

def test_philox_matches_numpy(orc):
    for key, ctr in ((0, 0), (2026, 41), ((5 << 64) | 12345, 7), (2**128 - 1, 2**64 + 3)):
        ref = Philox(key=key, counter=ctr).random_raw(4)
        c = ctr + 1  # numpy increments the counter before each block
        cw = (C.c_uint64 * 4)(*[(c >> (64 * i)) & (2**64 - 1) for i in range(4)])
        kw = (C.c_uint64 * 2)(key & (2**64 - 1), (key >> 64) & (2**64 - 1))
        out = (C.c_uint64 * 4)()
        orc.orc_philox4x64(cw, kw, out)
        assert list(out) == [int(x) for x in ref]


def test_philox4x32_known_answers(orc):
    """Philox4x32-10 against the known-answer vectors published with the
    Random123 library (kat_vectors: zero, all-ones and pi-digit counters and
    keys)."""
    kats = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
            ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
            ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
             (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kats:
        out = (C.c_uint32 * 4)()
        orc.orc_philox4x32((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
        assert tuple(out) == want


def test_stream_words_follow_the_stated_layout(orc):
    """AWGN sample i is word i & 1 of block i >> 1, clock index k the branch
    k & 1 of word (k >> 1) & 1 of block k >> 2: counter (block, tag), key seed,
    word q = outputs 2q+1 : 2q."""
    awgn_tag, clock_tag = 0x6177676E6E6F6973, 0x636C6F636B77616C
    zc, zs, zr, zi = C.c_float(), C.c_float(), C.c_float(), C.c_float()

    def word(block, seed, tag, q):
        ctr = (C.c_uint32 * 4)(block & 0xFFFFFFFF, block >> 32, tag & 0xFFFFFFFF, tag >> 32)
        out = (C.c_uint32 * 4)()
        orc.orc_philox4x32(ctr, (C.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32), out)
        return (out[2 * q + 1] << 32) | out[2 * q]

    for seed, i in ((5, 0), (5, 1), (2**63 + 11, 12345), (1, 2**33 + 7)):
        orc.orc_awgn_normals(seed, i, C.byref(zr), C.byref(zi))
        orc.orc_box_muller(word(i >> 1, seed, awgn_tag, i & 1), C.byref(zc), C.byref(zs))
        assert (zr.value, zi.value) == (zc.value, zs.value)
        orc.orc_box_muller(word(i >> 2, seed, clock_tag, (i >> 1) & 1), C.byref(zc), C.byref(zs))
        assert orc.orc_clock_normal(seed, i) == (zs.value if i & 1 else zc.value)


def test_box_muller_against_libm(orc):
    """The project's libm-free f32 Box-Muller pair against numpy's f64 log,
    sqrt, cos and sin of the same uniforms (u1 = (m1 + 1/2) 2^-23 from bits
    63..41, angle m2 2^-24 turns from bits 39..16): every normal within
    3e-7 of the radius everywhere, at every quadrant and reflection edge of
    the angle and both ends of u1; bits 40 and 15..0 of the word are ignored;
    the quadrant points are exact zeros and radii."""
    rng = np.random.default_rng(3)
    m1s = [0, 1, 2, (1 << 22) - 1, 1 << 22, (1 << 23) - 2, (1 << 23) - 1] + list(rng.integers(0, 1 << 23, 400))
    edges = [0, 1, (1 << 21) - 1, 1 << 21, (1 << 21) + 1, (1 << 22) - 1, 1 << 22, (1 << 22) + 1,
             3 << 21, (1 << 23) - 1, 1 << 23, 5 << 21, 3 << 22, 7 << 21, (1 << 24) - 1]
    m2s = edges + list(rng.integers(0, 1 << 24, 400))
    zc, zs = C.c_float(), C.c_float()
    got, want, radius = [], [], []
    for a in m1s:
        for b in m2s[:: 1 if a < 8 else 37]:
            w = (int(a) << 41) | (int(rng.integers(0, 2)) << 40) | (int(b) << 16) | int(rng.integers(0, 1 << 16))
            orc.orc_box_muller(w, C.byref(zc), C.byref(zs))
            u1 = (int(a) + 0.5) * 2.0 ** -23
            ang = 2.0 * math.pi * (int(b) * 2.0 ** -24)
            r = math.sqrt(-2.0 * math.log(u1))
            got.append((zc.value, zs.value))
            want.append((r * math.cos(ang), r * math.sin(ang)))
            radius.append(r)
    got = np.array(got, dtype=np.float64)
    want = np.array(want)
    err = np.abs(got - want) / np.array(radius)[:, None]
    assert np.all(err <= 3e-7), err.max()
    assert np.mean(got.astype(np.float32) == want.astype(np.float32)) > 0.5
    for b, (c, s) in {0: (1, 0), 1 << 22: (0, 1), 1 << 23: (-1, 0), 3 << 22: (0, -1)}.items():
        orc.orc_box_muller((5 << 41) | (b << 16), C.byref(zc), C.byref(zs))
        r = math.sqrt(-2.0 * math.log(5.5 * 2.0 ** -23))
        assert (zc.value == 0.0) == (c == 0) and (zs.value == 0.0) == (s == 0)
        assert abs(zc.value - r * c) <= 3e-7 * r and abs(zs.value - r * s) <= 3e-7 * r


def test_normals_are_standard_normal(orc):
    from scipy import stats
    z = np.array([orc.orc_clock_normal(77, k) for k in range(40_000)])
    assert stats.kstest(z, "norm").pvalue > 1e-3 and abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.02
    zr, zi = C.c_float(), C.c_float()
    pairs = []
    for i in range(20_000):
        orc.orc_awgn_normals(5, i, C.byref(zr), C.byref(zi))
        pairs.append((zr.value, zi.value))
    pr = np.array(pairs)
    for col in (pr[:, 0], pr[:, 1]):
        assert stats.kstest(col, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(pr[:, 0], pr[:, 1])[0, 1]) < 0.03


# ---------------------------------------------------------------- antenna

def _band_with_antenna(peak, horizon, fb, cutoff=None):
    from oracle.pyoracle import OrcBand
    b = OrcBand()
    b.ant_on, b.ant_peak_dbi, b.ant_horizon_dbi, b.ant_fb_db = 1, peak, horizon, fb
    if cutoff is not None:
        b.ant_has_cutoff, b.ant_cutoff_deg = 1, cutoff
    return b


def test_antenna_pattern_known_answers(orc, kats):
    k = kats["antenna_pattern"]
    a = k["antenna"]
    b = _band_with_antenna(a["peak_dbi"], a["horizon_dbi"], a["fb_db"])
    dbi = lambda e: 20.0 * math.log10(orc.orc_antenna_amplitude(C.byref(b), e))
    for e, want in k["cases_dbi"]:
        assert abs(dbi(e) - want) < k["tolerance_db"], e
    lo, mid, hi = k["monotone"]
    assert dbi(lo) < dbi(mid) < dbi(hi)
    # linear in dBi above the horizon: the midpoint gain is the mean of the ends
    assert abs(dbi(mid) - 0.5 * (dbi(lo) + dbi(hi))) < 1e-5
    bc = _band_with_antenna(a["peak_dbi"], a["horizon_dbi"], a["fb_db"], k["cutoff_deg"])
    assert orc.orc_antenna_amplitude(C.byref(bc), k["occluded_at"]) == 0.0
    assert orc.orc_antenna_amplitude(C.byref(bc), k["present_at"]) > 0.0
    assert orc.orc_antenna_amplitude(C.byref(bc), k["unchanged_at"]) == orc.orc_antenna_amplitude(
        C.byref(b), k["unchanged_at"])


def test_arrival_elevation_closed_form(orc):
    """Elevation against geocentric up: overhead 90, tangent 0, far side
    negative, and atan2(radial, tangential) for sources in the obs meridian plane."""
    R = 6378137.0
    obs = (C.c_double * 3)(R, 0.0, 0.0)
    el = lambda p: orc.orc_arrival_elevation_deg(obs, (C.c_double * 3)(*p))
    assert el((R + 5e5, 0.0, 0.0)) == 90.0
    assert el((R, 1e6, 0.0)) == 0.0
    assert el((-7e6, 0.0, 0.0)) == -90.0
    rng = np.random.default_rng(9)
    for _ in range(200):
        a, t = rng.uniform(-3e6, 3e6), rng.uniform(1e3, 3e6)
        ang = rng.uniform(0, 2 * math.pi)
        want = math.degrees(math.atan2(a, t))
        assert abs(el((R + a, t * math.cos(ang), t * math.sin(ang))) - want) < 1e-9


def test_antenna_cutoff_occludes_below_horizon(kats):
    """src/channel/mod.rs:384-420 through the oracle's band render: a 0 deg
    cutoff drops a source on the far side of the Earth and keeps an overhead
    one; an open antenna (and no antenna) lets the far source through."""
    from lipika_amd.config import (CaCode, FrontendConfig, ReceiverBandConfig, ReceiverConfig, SceneConfig,
                                   SceneMeta, StationaryEcef, TransmitterConfig)
    from lipika_amd.pipeline import prepare
    from oracle.pyoracle import OracleScene
    from lipika_amd.config import NodeClockSpec
    k = kats["antenna_occlusion"]
    ua = k["unit_antenna"]

    def power(src, cutoff):
        tx = TransmitterConfig("s", k["carrier_hz"], CaCode(k["prn"], k["epoch"]), StationaryEcef(*src),
                               10.0 * math.log10(k["tx_power_w"]))
        fe = None if cutoff == "none" else FrontendConfig(antenna=dict(ua, **({} if cutoff is None else
                                                                             {"cutoff_deg": cutoff})))
        band = ReceiverBandConfig("l1", k["carrier_hz"], 2.048e6, 290.0, noise_power_w=0.0, frontend=fe)
        rx = ReceiverConfig("eq", StationaryEcef(*k["obs"]), NodeClockSpec(), [band])
        sc = prepare(SceneConfig(SceneMeta(k["epoch"], k["n"] / 2.048e6), [tx], [rx]))
        flags = np.zeros(1, dtype=np.uint8)
        iq, _, _ = OracleScene(sc, 1).render_band(0, unoccluded=flags)
        return float(np.sum(np.abs(iq.astype(np.complex128)) ** 2)), int(flags[0])

    assert power(k["below"], "none")[0] > 0.0
    assert power(k["below"], None)[0] > 0.0
    assert power(k["below"], 0.0) == (0.0, 0)
    p, f = power(k["above"], 0.0)
    assert p > 0.0 and f == 1


# ---------------------------------------------------------------- receiver ops

def test_ktb(orc, kats):
    k = kats["ktb"]
    p = orc.orc_thermal_noise_power_w(k["temperature_k"], k["bandwidth_hz"])
    assert abs(10 * math.log10(p * 1e3) - k["dbm"]) < k["dbm_tolerance"]
    assert orc.orc_thermal_noise_power_w(580.0, 2.048e6) == 2 * p
    assert orc.orc_thermal_noise_power_w(290.0, 4.096e6) == 2 * p


def post(orc, ops, iq):
    arr = (util.pyoracle.OrcPost * 4)()
    for i, (kind, bits, v) in enumerate(ops):
        arr[i].kind, arr[i].bits, arr[i].value = kind, bits, v
    x = np.ascontiguousarray(np.array(iq, dtype=np.complex64))
    codes = np.empty(2 * len(x), dtype=np.int16)
    orc.orc_post_ops(arr, len(ops), x.view(np.float32).ctypes.data_as(P(C.c_float)), len(x),
                     codes.ctypes.data_as(P(C.c_int16)))
    return x, codes


def test_adc_gain_ip3(orc, kats):
    k = kats["adc"]
    x, codes = post(orc, [(2, k["bits"], k["full_scale_v"])], [complex(*v) for v in k["input"]])
    lsb = np.float32(1.0) / np.float32(7.0)
    for v in x.view(np.float32):
        assert abs(v) <= 1 + 1e-6 and abs(np.round(v / lsb) * lsb - v) < 1e-6
    assert (x[1].real, x[1].imag) == tuple(k["clipped"])
    np.testing.assert_array_equal(codes.astype(np.float32) * lsb, x.view(np.float32))
    assert list(codes[2:]) == [7, -7]
    g = kats["gain"]
    y, _ = post(orc, [(0, 0, g["db"])], [complex(*g["input"])])
    assert abs(y[0].real - g["output"][0]) < g["tolerance"] and abs(y[0].imag - g["output"][1]) < g["tolerance"]
    q = kats["ip3"]
    iip3 = 10 ** ((q["iip3_dbm"] - 30) / 10)
    z, _ = post(orc, [(1, 0, iip3)], [q["small"], q["large"]])
    assert abs(z[0] - q["small"]) / q["small"] < 1e-3 and abs(z[1]) < q["large"]
    expect = np.float32(q["large"]) * np.float32(1 - (4 / 3) * q["large"] ** 2 / iip3)
    assert abs(z[1].real - expect) < q["tolerance"]
    e, _ = post(orc, [], [0.3 - 0.7j, 1.1 + 0.2j])
    np.testing.assert_array_equal(e, np.array([0.3 - 0.7j, 1.1 + 0.2j], dtype=np.complex64))


# ---------------------------------------------------------------- clock

def rx_struct(adev=0.0, y0=0.0, c0=0.0, master=1e6, seed=42):
    r = util.pyoracle.OrcRx()
    r.allan_deviation_1s, r.frequency_offset_y0, r.initial_time_offset_s = adev, y0, c0
    r.master_rate_hz, r.clock_seed = master, seed
    return r


def test_clock_drift_exact(orc, kats):
    k = kats["clock_drift"]
    c = util.pyoracle_clock(orc, rx_struct(y0=k["y0"], master=k["fs"]))
    last = int(k["fs"] * k["duration_s"]) - 1
    got = orc.orc_clock_delta_t_at_master_index(C.byref(c), last)
    assert abs(got / k["expected_last"] - 1) < k["max_relative"]


def test_clock_allan_deviation(orc, kats):
    k = kats["clock_allan"]
    fs, dur = 1e5, 60
    c = util.pyoracle_clock(orc, rx_struct(adev=k["adev"], master=fs, seed=7))
    track = np.array([orc.orc_clock_delta_t_at_master_index(C.byref(c), i) for i in range(int(fs * dur))])
    d = track[int(fs)::int(fs)] - track[:-int(fs):int(fs)]
    assert abs(np.std(d) / k["adev"] - 1) < k["tolerance"]


def test_clock_two_band_coherence_and_replay(orc, kats):
    k = kats["clock_two_band"]
    hi = util.pyoracle_clock(orc, rx_struct(adev=k["adev"], y0=k["y0"], master=k["master"], seed=99))
    lo = util.pyoracle_clock(orc, rx_struct(adev=k["adev"], y0=k["y0"], master=k["master"], seed=99))
    n_hi = int(k["n_low"] / k["low"] * k["master"]) + 2
    hv = np.array([orc.orc_clock_delta_t_at(C.byref(hi), i / k["master"]) for i in range(n_hi)])
    for j in range(k["n_low"]):
        t = j / k["low"]
        got = orc.orc_clock_delta_t_at(C.byref(lo), t)
        x = t * k["master"]
        k0 = int(math.floor(x))
        ref = hv[k0] * (1 - (x - k0)) + hv[k0 + 1] * (x - k0)
        assert abs(got - ref) <= max(k["max_relative"] * abs(ref), k["epsilon"])
    again = util.pyoracle_clock(orc, rx_struct(adev=k["adev"], y0=k["y0"], master=k["master"], seed=99))
    assert all(orc.orc_clock_delta_t_at(C.byref(again), i / k["master"]) == hv[i] for i in range(2000))


# ---------------------------------------------------------------- end to end through the oracle

def oracle_scene(scene, seed):
    from lipika_amd.pipeline import prepare
    return util.pyoracle.OracleScene(prepare(scene), seed)


def test_oracle_clock_offset_recovered(kats):
    o = oracle_scene(util.cw_two_clocks_scene(), 7)
    a, _, _ = o.render_band(1)
    b, _, _ = o.render_band(0)
    acc = np.sum(a.astype(np.complex128) * np.conj(b.astype(np.complex128)))
    rec = -math.atan2(acc.imag, acc.real) / (2 * math.pi * 1e6)
    assert abs(rec - 1.3e-7) < kats["clock_offset_recovery"]["tolerance_s"]


def test_oracle_additivity_and_noise(kats):
    seed = 99
    both = oracle_scene(util.two_tx_scene(("alpha", "beta"), 0.1), seed).render_band(0)[0]
    a = oracle_scene(util.two_tx_scene(("alpha",), 0.1), seed).render_band(0)[0]
    b = oracle_scene(util.two_tx_scene(("beta",), 0.1), seed).render_band(0)[0]
    n = oracle_scene(util.two_tx_scene((), 0.1), seed).render_band(0)[0]
    assert np.max(np.abs(both - (a + b - n))) < kats["additivity"]["max_abs_err"]
    k = kats["noise_power"]
    iq = n.astype(np.complex128)
    assert abs(np.mean(np.abs(iq) ** 2) / k["power_w"] - 1) < k["tolerance"]


def test_oracle_chunk_invariance():
    o = oracle_scene(util.all_kinds_scene(duration_s=0.002, noise_power_w=None), 3)
    ref = o.render_band(0, chunk_samples=1 << 20)[0]
    for chunk in (4096, 777):
        assert o.render_band(0, chunk_samples=chunk)[0].tobytes() == ref.tobytes()


def test_oracle_window_equals_band_render():
    """orc_render_window given the band render's own clock offsets is the
    band render, bit for bit: the whole band, a window late in it (noise
    included: the AWGN index is absolute), and the taps of every link at once
    equal the single-slot taps of orc_render_band."""
    from lipika_amd import workloads
    sc = workloads.scene2(duration_s=0.05)
    o = oracle_scene(sc, 9)
    w = 1
    n = o.units[w][1].n
    ref, _, taps = o.render_band(w, tap_slot=0, chunk_samples=30_000)
    dt = taps[3]
    full, _, tp = o.render_window(w, 0, dt, taps=True)
    assert full.tobytes() == ref.tobytes()
    i0, m = n - 5000, 4000
    win, _, _ = o.render_window(w, i0, dt[i0:i0 + m])
    assert win.tobytes() == ref[i0:i0 + m].tobytes()
    for slot in range(tp[0].shape[0]):
        _, _, ts = o.render_band(w, tap_slot=slot, chunk_samples=30_000)
        for k in range(3):
            assert tp[k][slot].tobytes() == ts[k].tobytes(), (slot, k)
