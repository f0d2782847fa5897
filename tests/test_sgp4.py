"""TLE trajectories: SGP4/SDP4 on the host (lpk_tle_*, include/lpk.h), the
source the reference builds with the third-party sgp4 crate
(src/geometry/tle.rs:15-85).

Pins: the published SGP4 verification vectors of Vallado, Crawford, Hujsak
and Kelso (AIAA 2006-6753, test set SGP4-VER, improved mode, WGS72) for a
near-Earth orbit with drag (satellite 00005) and a deep-space 12 h resonant
Molniya orbit (08195); the reference's own TLE tests (tle.rs:92-113,
gps_nav.rs:977-1004); and frame invariants. Bit parity with the crate
itself is not pinned: the crate is not in the reference tree.
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from lipika_amd import _lib  # noqa: E402

V5 = ("1 00005U 58002B   00179.78495062  .00000023  00000-0  28098-4 0  4753",
      "2 00005  34.2682 348.7242 1859667 331.7664  19.3264 10.82419157413667")
V8195 = ("1 08195U 75081A   06176.33215444  .00000099  00000-0  11873-3 0   813",
         "2 08195  64.1586 279.0717 6877146 264.7651  20.2257  2.00491383225656")
# src/geometry/tle.rs:92-93 (GPS PRN 22)
GPS22 = ("1 26407U 00040A   26150.58242625  .00000046  00000+0  00000+0 0  9999",
         "2 26407  54.8538 215.8617 0121278 302.3883 230.2243  2.00557636189594")
# examples/iss_uhf_pass.toml:16
ISS = ("1 25544U 98067A   24001.50000000  .00010000  00000+0  18000-3 0  9994",
       "2 25544  51.6400 100.0000 0001000  90.0000 270.0000 15.50000000000009")

# examples/gps_real_tle.toml: eight live GPS satellites (Celestrak, epoch ~2026-150)
GPS_EXAMPLE = [
    ("1 38833U 12053A   26150.86639939  .00000051  00000+0  00000+0 0  9997",
     "2 38833  53.5827 144.3088 0179371  65.5174 296.3427  2.00567269 99098"),
    ("1 32260U 07047A   26150.77386177 -.00000055  00000+0  00000+0 0  9990",
     "2 32260  54.0827  80.7833 0166076  87.0458 274.8256  2.00556201136463"),
    ("1 26407U 00040A   26150.58242625  .00000046  00000+0  00000+0 0  9999",
     "2 26407  54.8538 215.8617 0121278 302.3883 230.2243  2.00557636189594"),
    ("1 45854U 20041A   26151.52781884 -.00000085  00000+0  00000+0 0  9993",
     "2 45854  56.6592  29.5897 0063281 204.6776 332.6203  2.00559374 43701"),
    ("1 67588U 26017A   26150.24909249 -.00000043  00000+0  00000+0 0  9999",
     "2 67588  55.0776  89.4749 0022608 256.9258 102.7871  2.00571628431254"),
    ("1 41019U 15062A   26150.82698463 -.00000088  00000+0  00000+0 0  9994",
     "2 41019  56.9480  31.4799 0111519 231.2324 127.8139  2.00567167 77488"),
    ("1 46826U 20078A   26150.43676341  .00000051  00000+0  00000+0 0  9995",
     "2 46826  53.9818 211.4986 0072702 206.7371 238.5468  2.00565060 41158"),
    ("1 28874U 05038A   26150.41242926 -.00000056  00000+0  00000+0 0  9999",
     "2 28874  54.8931 273.8947 0127762 296.0439  71.0693  2.00567025151481"),
]


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def with_checksum(line68):
    """A TLE line with its modulo-10 checksum (digits, and 1 per minus sign)."""
    return line68 + str(sum(int(c) if c.isdigit() else c == "-" for c in line68) % 10)


def parse(lib, lines):
    e = _lib.Tle()
    assert lib.lpk_tle_parse(lines[0].encode(), lines[1].encode(), C.byref(e)) == _lib.LPK_OK
    return e


def teme(lib, e, minutes):
    r, v = (C.c_double * 3)(), (C.c_double * 3)()
    assert lib.lpk_sgp4_teme(C.byref(e), minutes, r, v) == _lib.LPK_OK
    return np.array(r[:]), np.array(v[:])


def ecef(lib, e, t):
    r, v = (C.c_double * 3)(), (C.c_double * 3)()
    assert lib.lpk_tle_state_ecef(C.byref(e), t, r, v) == _lib.LPK_OK
    return np.array(r[:]), np.array(v[:])


def test_near_earth_verification_vectors(lib):
    """Satellite 00005 (perigee 1 058 km, period 133 min, B* 2.8e-5): the
    published TEME states at 0, 360, 720 and 1080 min, to their printed
    digits (1e-8 km, 1e-9 km/s)."""
    e = parse(lib, V5)
    assert abs(e.jd_epoch - 2451723.28495062) < 1e-8
    want = {
        0.0: ([7022.46529266, -1400.08296755, 0.03995155], [1.893841015, 6.405893759, 4.534807250]),
        360.0: ([-7154.03120202, -3783.17682504, -3536.19412294], [4.741887409, -4.151817765, -2.093935425]),
        720.0: ([-7134.59340119, 6531.68641334, 3260.27186483], [-4.113793027, -2.911922039, -2.557327851]),
        1080.0: ([5568.53901181, 4492.06992591, 3863.87641983], [-4.209106476, 5.159719888, 2.744852980]),
    }
    for t, (rw, vw) in want.items():
        r, v = teme(lib, e, t)
        assert np.max(np.abs(r - rw)) < 1e-8 * 1.5, (t, r)
        assert np.max(np.abs(v - vw)) < 1e-9 * 1.5, (t, v)


def test_deep_space_resonant_verification_vector(lib):
    """Satellite 08195 (Molniya, e 0.69, 12 h: the deep-space model with the
    half-day geopotential resonance): the published state at epoch, and the
    x coordinate at 120 min, after the resonance terms and lunar-solar
    periodics have acted."""
    e = parse(lib, V8195)
    r, v = teme(lib, e, 0.0)
    assert np.max(np.abs(r - [2349.89483350, -14785.93811562, 0.02119378])) < 1.5e-8
    assert np.max(np.abs(v - [2.721488096, -3.256811655, 4.498416672])) < 1.5e-9
    r, _ = teme(lib, e, 120.0)
    assert abs(r[0] - 15223.91713658) < 1.5e-8


def test_gps_tle_gives_gps_altitude_and_consistent_velocity(lib):
    """src/geometry/tle.rs:95-113: GPS-like radius at epoch, and the analytic
    ECEF velocity within 1 m/s of a central difference over +-0.1 s."""
    e = parse(lib, GPS22)
    t0 = e.epoch_unix
    p, v = ecef(lib, e, t0)
    assert abs(np.linalg.norm(p) - 26_560_000.0) < 1_500_000.0
    h = 0.1
    pp, _ = ecef(lib, e, t0 + h)
    pm, _ = ecef(lib, e, t0 - h)
    assert np.max(np.abs(v - (pp - pm) / (2 * h))) < 1.0


def test_teme_to_ecef_is_the_references_rotation(lib):
    """TleTrajectory::position_ecef (tle.rs:64-71): a rotation about z by the
    IAU 1982 GMST of the sample time; norms and z are kept, and the angle
    between TEME and ECEF x-y vectors is that GMST."""
    e = parse(lib, ISS)
    for dt in (0.0, 1234.5, 86400.0):
        t = e.epoch_unix + dt
        rt, _ = teme(lib, e, dt / 60.0)
        re_, _ = ecef(lib, e, t)
        assert abs(np.linalg.norm(re_) - 1000.0 * np.linalg.norm(rt)) < 1e-6
        assert abs(re_[2] - 1000.0 * rt[2]) < 1e-6
        jd = t / 86400.0 + 2440587.5
        tc = (jd - 2451545.0) / 36525.0
        gmst = ((67310.54841 + (876600.0 * 3600.0 + 8640184.812866) * tc + 0.093104 * tc * tc
                 - 6.2e-6 * tc ** 3) / 86400.0) % 1.0 * 2 * math.pi
        ang = math.atan2(rt[1], rt[0]) - math.atan2(re_[1], re_[0])
        assert abs(math.remainder(ang - gmst, 2 * math.pi)) < 1e-9


def test_ephemeris_from_tle_is_consistent_over_window(lib):
    """src/signal/gps_nav.rs:977-1004: the osculating broadcast ephemeris of
    the TLE trajectory at a 6 s-aligned toe reproduces it to under a metre
    at toe and within 5 km 60 s later."""
    e = parse(lib, GPS22)
    toe_unix = round(e.epoch_unix / 6.0) * 6.0
    t_toe, wk, toe = C.c_double(), C.c_uint32(), C.c_double()
    assert lib.lpk_toe_snap(toe_unix, 18.0, C.byref(t_toe), C.byref(wk), C.byref(toe)) == _lib.LPK_OK
    r, v = ecef(lib, e, t_toe.value)
    eph = _lib.Ephemeris()
    assert lib.lpk_ephemeris_from_state((C.c_double * 3)(*r), (C.c_double * 3)(*v), 22, wk.value, toe.value,
                                        C.byref(eph)) == _lib.LPK_OK
    week, sow = C.c_uint32(), C.c_double()
    lib.lpk_gps_time(toe_unix, 18.0, C.byref(week), C.byref(sow))
    s = (C.c_double * 3)()
    lib.lpk_sv_position(C.byref(eph), sow.value, s)
    p0, _ = ecef(lib, e, toe_unix)
    assert np.max(np.abs(p0 - s[:])) < 1.0
    lib.lpk_sv_position(C.byref(eph), sow.value + 60.0, s)
    pt, _ = ecef(lib, e, toe_unix + 60.0)
    assert np.linalg.norm(pt - s[:]) < 5000.0


def test_tle_parse_fields_and_errors(lib):
    """Elements::from_tle: columns, implied decimals and exponents, two-digit
    years (57-99 are the 1900s), the epoch as UTC unix time in whole
    microseconds; bad checksums and malformed lines are errors."""
    e = parse(lib, V5)
    assert e.satnum == 5 and e.epoch_year == 2000 and abs(e.epoch_day - 179.78495062) < 1e-12
    assert abs(e.bstar - 2.8098e-5) < 1e-18 and e.nddot == 0.0 and abs(e.ndot - 2.3e-7) < 1e-18
    assert abs(e.inclo_deg - 34.2682) < 1e-12 and abs(e.ecco - 0.1859667) < 1e-15
    assert abs(e.no_revday - 10.82419157) < 1e-12
    # 2000-06-27 (day 179) 18:50:19.733568 UTC, kept to the microsecond
    assert abs(e.epoch_unix - 962131819.733567) < 2e-6
    assert abs(e.epoch_unix * 1e6 - round(e.epoch_unix * 1e6)) < 0.01
    e = parse(lib, V8195)
    assert e.epoch_year == 2006 and abs(e.bstar - 1.1873e-4) < 1e-16
    old = (with_checksum("1 00005U 58002B   58179.78495062  .00000023  00000-0  28098-4 0  475"), V5[1])
    e = _lib.Tle()
    assert lib.lpk_tle_parse(old[0].encode(), old[1].encode(), C.byref(e)) == _lib.LPK_OK
    assert e.epoch_year == 1958
    bad = [(V5[0][:-1] + "0", V5[1]), (V5[0], V5[1][:-1] + "0"), (V5[0][:60], V5[1]), (V5[1], V5[0]),
           ("1 26407U", "2 26407")]
    for l1, l2 in bad:
        assert lib.lpk_tle_parse(l1.encode(), l2.encode(), C.byref(e)) == -1, (l1, l2)


def test_grid_and_scene_from_tle(lib):
    """The grid a TLE transmitter uploads holds the propagated ECEF states at
    the nodes of PrecomputedTrajectory::sample; a GPS TLE with full NAV
    builds an LNAV stream whose ephemeris is the TLE's GPS orbit."""
    from lipika_amd.config import CaCode, ReceiverBandConfig, SceneConfig, SceneMeta, Tle, TransmitterConfig
    from lipika_amd.pipeline import prepare
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from util import lab_receiver
    from test_gpsnav import decode_ephemeris, decode_words
    e = parse(lib, GPS22)
    epoch = 1780268400.0   # examples/gps_real_tle.toml:12, divisible by 6
    n = int(lib.lpk_grid_size(2.0, 0.01))
    xyz = np.empty(3 * n)
    t0 = C.c_double()
    assert lib.lpk_grid_fill_tle(C.byref(e), epoch, 2.0, 0.01, C.byref(t0),
                                 xyz.ctypes.data_as(C.POINTER(C.c_double)), n) == _lib.LPK_OK
    assert t0.value == epoch - 0.5
    for i in (0, n // 2, n - 1):
        p, _ = ecef(lib, e, t0.value + i * 0.01)
        assert np.array_equal(p, xyz[[i, n + i, 2 * n + i]])
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    txs = [TransmitterConfig("gps22", 1575.42e6, CaCode(22, epoch, "full"), Tle(*GPS22), 27.0)]
    sc = prepare(SceneConfig(SceneMeta(epoch, 18.0), txs, [lab_receiver([band])]))
    t = sc.transmitters[0]
    assert t.traj_kind == _lib.TRAJ_GRID and np.array_equal(t.grid[:, 0], ecef(lib, e, t.grid_t0)[0])
    d = decode_ephemeris(decode_words(t.nav_bits))
    assert abs(d["sqrt_a"] ** 2 - 26_560_000.0) < 1_500_000.0 and abs(math.degrees(d["i0"]) - 54.85) < 0.5


def test_example_sky_pass_over_glasgow(lib):
    """tests/horizon_gate.rs:380-396: the example sky's ISS-like TLE,
    propagated 152 days past its epoch, passes over Glasgow culminating at
    Unix 1717260990 and rises less than 900 s before. SGP4's along-track
    drift over that span would move the pass by minutes for a wrong drag or
    secular term; here the elevation peaks within 10 s of that time."""
    e = parse(lib, ISS)
    o = (C.c_double * 3)()
    lib.lpk_geodetic_to_ecef(55.86, -4.25, 20.0, o)
    o = np.array(o[:])
    up = o / np.linalg.norm(o)

    def el(t):
        p, _ = ecef(lib, e, t)
        d = p - o
        return math.degrees(math.asin(np.dot(d, up) / np.linalg.norm(d)))

    cul = 1_717_260_990.0
    assert el(cul) > 30.0
    ts = np.arange(cul - 60.0, cul + 60.0, 1.0)
    assert abs(ts[int(np.argmax([el(t) for t in ts]))] - cul) <= 10.0
    rise = cul
    while el(rise) > 0.0:
        rise -= 1.0
        assert cul - rise < 900.0


def test_example_constellation_is_above_fifteen_degrees(lib):
    """examples/gps_real_tle.toml:3-5: the eight GPS satellites of the example
    are all above 15 deg elevation over Glasgow at 2026-05-31 23:00 UTC; each
    propagated about half a day from its own epoch through the deep-space
    model and rotated by that instant's GMST."""
    o = (C.c_double * 3)()
    lib.lpk_geodetic_to_ecef(55.86, -4.25, 10.0, o)
    o = np.array(o[:])
    t = 1780268400.0
    for l1, l2 in GPS_EXAMPLE:
        p, _ = ecef(lib, parse(lib, (l1, l2)), t)
        d = p - o
        el = math.degrees(math.asin(np.dot(d, o) / (np.linalg.norm(d) * np.linalg.norm(o))))
        assert el > 15.0, (l1[2:7], el)


def test_a_decayed_orbit_is_an_error(lib):
    """The reference's propagate panics where SGP4 reports an error
    (tle.rs:66-67, `expect`); here the state and the grid calls return
    LPK_E_INVALID, and scene preparation raises ValueError. The example's
    ISS-like orbit (B* 1.8e-4) has decayed 3 000 days after its epoch."""
    from lipika_amd.config import Cw, ReceiverBandConfig, SceneConfig, SceneMeta, Tle, TransmitterConfig
    from lipika_amd.pipeline import prepare
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from util import lab_receiver
    e = parse(lib, ISS)
    r, v = (C.c_double * 3)(), (C.c_double * 3)()
    assert lib.lpk_sgp4_teme(C.byref(e), 100 * 1440.0, r, v) == _lib.LPK_OK
    assert lib.lpk_sgp4_teme(C.byref(e), 3000 * 1440.0, r, v) == -1
    late = e.epoch_unix + 3000 * 86400.0
    assert lib.lpk_tle_state_ecef(C.byref(e), late, r, v) == -1
    band = ReceiverBandConfig("uhf", 437e6, 48e3, noise_power_w=0.0)
    txs = [TransmitterConfig("iss", 437e6, Cw(0.0), Tle(*ISS), 0.0)]
    with pytest.raises(ValueError, match="sgp4"):
        prepare(SceneConfig(SceneMeta(late, 1.0), txs, [lab_receiver([band])]))


def test_geostationary_orbit_stays_over_one_longitude(lib):
    """The 24 h synchronous-resonance branch with Lyddane's low-inclination
    periodics (inclination 0.02 deg): a geostationary element set stays at
    the GEO radius and, Earth-fixed, within 0.2 deg of one longitude and of
    the equator over two days (its mean motion of 1.0027 rev/day is one
    sidereal day)."""
    geo = (with_checksum("1 28626U 05004A   06176.43865743 -.00000282  00000-0  10000-3 0  122"),
           with_checksum("2 28626   0.0177 257.7591 0001835  40.4209 214.8001  1.00271000  503"))
    e = parse(lib, geo)
    lons, lats = [], []
    for h in range(0, 49, 3):
        p, _ = ecef(lib, e, e.epoch_unix + h * 3600.0)
        assert abs(np.linalg.norm(p) - 42_164e3) < 50e3, h
        lons.append(math.degrees(math.atan2(p[1], p[0])))
        lats.append(math.degrees(math.asin(p[2] / np.linalg.norm(p))))
    lons = np.unwrap(np.radians(lons))
    assert np.degrees(np.ptp(lons)) < 0.2 and max(abs(x) for x in lats) < 0.2
