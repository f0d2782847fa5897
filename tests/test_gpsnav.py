"""GPS LNAV and broadcast-ephemeris builders (src/signal/gps_nav.rs), on the CPU.

The library's builders are checked the way the reference's own tests check
them (src/signal/gps_nav.rs:731-1031), with a receiver-side decoder written
here from the IS-GPS-200 parity equations (section 20.3.5.2, Table 20-XIV) and
field layouts (20.3.3.3-20.3.3.4), not from the encoder:
  - every word's parity, the TLM preamble, the solved t-bits of words 2 and 10;
  - subframe ids, the TOW count, and the ephemeris fields decoded back;
  - the decoded ephemeris placing the satellite within metres of the orbit it
    was made from, and the exact ephemeris within a millimetre over 10 min;
  - the RINEX 3 record of the reference's test (tests/golden/rinex3_gps_prn01.rnx,
    the text of src/signal/gps_nav.rs:838-849) and its parsed values.
"""
import ctypes as C
import math
import os
import warnings

import numpy as np
import pytest

from lipika_amd import _lib
from lipika_amd.config import (CaCode, CircularKeplerian, ReceiverBandConfig, RinexNav, SceneConfig, SceneMeta,
                               StationaryEcef, Tle, TransmitterConfig)
from lipika_amd.pipeline import prepare, trajectory_state

HERE = os.path.dirname(os.path.abspath(__file__))
RINEX = os.path.join(HERE, "golden", "rinex3_gps_prn01.rnx")
EPOCH = 1717260840.0
ORBIT = CircularKeplerian(26_560_000.0, 55.0, 40.0, 120.0, EPOCH)
GPS_EPOCH_UNIX, WEEK_S = 315964800.0, 604800.0


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def gps_time(lib, t):
    w, s = C.c_uint32(), C.c_double()
    lib.lpk_gps_time(t, 18.0, C.byref(w), C.byref(s))
    return w.value, s.value


def kepler_eph(lib, orbit=ORBIT, prn=1, epoch=EPOCH):
    e = _lib.Ephemeris()
    assert lib.lpk_ephemeris_from_kepler(orbit.semi_major_axis_m, orbit.inclination_deg, orbit.raan_deg,
                                         orbit.true_anomaly_deg_at_epoch, orbit.epoch_unix, prn, epoch, 18.0,
                                         C.byref(e)) == 0
    return e


def lnav(lib, eph, sow, duration):
    n = lib.lpk_lnav_length(duration)
    out = np.empty(n, dtype=np.int8)
    assert lib.lpk_lnav_bits(C.byref(eph), sow, duration, out.ctypes.data_as(C.POINTER(C.c_int8)), n) == 0
    return out


def sv(lib, eph, t_sow):
    p = (C.c_double * 3)()
    lib.lpk_sv_position(C.byref(eph), t_sow, p)
    return np.array(p[:])


# IS-GPS-200 Table 20-XIV: source bits d1..d24 in each parity bit, with the
# previous word's D29* or D30*
PARITY_EQ = [
    ("29", [1, 2, 3, 5, 6, 10, 11, 12, 13, 14, 17, 18, 20, 23]),
    ("30", [2, 3, 4, 6, 7, 11, 12, 13, 14, 15, 18, 19, 21, 24]),
    ("29", [1, 3, 4, 5, 7, 8, 12, 13, 14, 15, 16, 19, 20, 22]),
    ("30", [2, 4, 5, 6, 8, 9, 13, 14, 15, 16, 17, 20, 21, 23]),
    ("30", [1, 3, 5, 6, 7, 9, 10, 14, 15, 16, 17, 18, 21, 22, 24]),
    ("29", [3, 5, 6, 8, 9, 10, 11, 13, 15, 19, 22, 23, 24]),
]


def decode_words(pm):
    """Receiver side: the +-1 stream (bit 1 -> -1) to per-word source bits,
    checking each word's six parity bits against the equations above."""
    tx = (np.asarray(pm) < 0).astype(int)
    d29p = d30p = 0
    words = []
    for w in range(len(tx) // 30):
        word = tx[30 * w: 30 * w + 30]
        src = [int(word[n]) ^ d30p for n in range(24)]
        for k, (star, idx) in enumerate(PARITY_EQ):
            want = (d29p if star == "29" else d30p)
            for i in idx:
                want ^= src[i - 1]
            assert int(word[24 + k]) == want, f"parity bit {25 + k} of word {w}"
        words.append(src)
        d29p, d30p = int(word[28]), int(word[29])
    return words


def read_u(src, start, n):
    v = 0
    for i in range(n):
        v = (v << 1) | src[start + i]
    return v


def sext(v, n):
    return v - (1 << n) if v & (1 << (n - 1)) else v


def decode_ephemeris(words):
    """Subframe 2 and 3 fields (IS-GPS-200 20.3.3.4) back to SI units."""
    out = {}
    for sf in range(len(words) // 10):
        w = words[10 * sf: 10 * sf + 10]
        sfid = read_u(w[1], 19, 3)
        if sfid == 1:
            out["week"] = read_u(w[2], 0, 10)
        elif sfid == 2:
            out["m0"] = sext((read_u(w[3], 16, 8) << 24) | read_u(w[4], 0, 24), 32) * 2.0 ** -31 * math.pi
            out["e"] = ((read_u(w[5], 16, 8) << 24) | read_u(w[6], 0, 24)) * 2.0 ** -33
            out["sqrt_a"] = ((read_u(w[7], 16, 8) << 24) | read_u(w[8], 0, 24)) * 2.0 ** -19
            out["toe"] = read_u(w[9], 0, 16) * 16.0
        elif sfid == 3:
            out["omega0"] = sext((read_u(w[2], 16, 8) << 24) | read_u(w[3], 0, 24), 32) * 2.0 ** -31 * math.pi
            out["i0"] = sext((read_u(w[4], 16, 8) << 24) | read_u(w[5], 0, 24), 32) * 2.0 ** -31 * math.pi
            out["argp"] = sext((read_u(w[6], 16, 8) << 24) | read_u(w[7], 0, 24), 32) * 2.0 ** -31 * math.pi
    return out


def test_gps_time_of_a_six_second_epoch(lib):
    """src/signal/gps_nav.rs:778-785, plus the closed form of GPS time."""
    week, sow = gps_time(lib, EPOCH)
    assert 0.0 <= sow < WEEK_S and sow % 6.0 == 0.0
    assert week * WEEK_S + sow == EPOCH - GPS_EPOCH_UNIX + 18.0
    assert gps_time(lib, GPS_EPOCH_UNIX - 18.0) == (0, 0.0)


def test_lnav_parity_preamble_and_solved_tbits(lib):
    """src/signal/gps_nav.rs:740-776: every word's parity, the 0x8B preamble,
    and parity bits 29 and 30 forced to zero in words 2 and 10."""
    _, sow = gps_time(lib, EPOCH)
    bits = lnav(lib, kepler_eph(lib), sow, 36.0)
    assert len(bits) == 300 * 7 and set(np.unique(bits)) == {-1, 1}
    decode_words(bits)
    tx = (bits < 0).astype(int)
    for sf in range(7):
        word = tx[300 * sf: 300 * sf + 300]
        # TLM: the preamble, complemented when the previous word's D30 is 1
        d30p = 0 if sf == 0 else int(tx[300 * sf - 1])
        assert [b ^ d30p for b in word[:8]] == [1, 0, 0, 0, 1, 0, 1, 1]
        assert list(word[58:60]) == [0, 0] and list(word[298:300]) == [0, 0]


def test_lnav_decodes_back_to_the_ephemeris(lib):
    """src/signal/gps_nav.rs:825-895: ids, TOW count and ephemeris fields."""
    week, sow = gps_time(lib, EPOCH)
    eph = kepler_eph(lib)
    words = decode_words(lnav(lib, eph, sow, 30.0))
    base_z = round(sow / 6.0)
    for sf in range(len(words) // 10):
        how = words[10 * sf + 1]
        assert read_u(how, 19, 3) == sf % 5 + 1
        assert read_u(how, 0, 17) == base_z + sf + 1
    dec = decode_ephemeris(words)
    assert dec["week"] == week & 0x3FF
    assert abs(dec["m0"] - eph.m0_rad) < 1e-8 and abs(dec["sqrt_a"] - eph.sqrt_a) < 1e-3
    assert dec["toe"] == eph.toe_s and eph.toe_s % 16.0 == 0.0
    assert abs(dec["omega0"] - eph.omega0_rad) < 1e-8 and abs(dec["i0"] - eph.i0_rad) < 1e-8
    assert dec["e"] == 0.0 and dec["argp"] == 0.0


def test_exact_ephemeris_reproduces_the_circular_orbit(lib):
    """src/signal/gps_nav.rs:1004-1030: the IS-GPS-200 SV position of the
    closed-form ephemeris within 1 mm of the orbit over 10 minutes, and the
    decoded (quantised) one within 5 m (:896-938)."""
    _, sow = gps_time(lib, EPOCH)
    eph = kepler_eph(lib)
    dec = decode_ephemeris(decode_words(lnav(lib, eph, sow, 30.0)))
    q = _lib.Ephemeris()
    C.memmove(C.byref(q), C.byref(eph), C.sizeof(q))
    q.m0_rad, q.sqrt_a, q.toe_s, q.omega0_rad, q.i0_rad = dec["m0"], dec["sqrt_a"], dec["toe"], dec["omega0"], dec["i0"]
    for dt in (0.0, 1.0, 30.0, 120.0, 600.0):
        truth, _ = trajectory_state(ORBIT, EPOCH + dt)
        assert np.max(np.abs(sv(lib, eph, sow + dt) - truth)) < 1e-3, dt
        if dt <= 300.0:
            assert np.linalg.norm(sv(lib, q, sow + dt) - truth) < 5.0, dt


def test_ephemeris_from_a_state_reproduces_it_at_toe(lib):
    """Ephemeris::from_trajectory (src/signal/gps_nav.rs:131-180): the
    osculating elements of a circular orbit's state equal its closed-form
    elements, and any state (a fixed ground point here) is reproduced at toe."""
    t_toe, wk, toe = C.c_double(), C.c_uint32(), C.c_double()
    assert lib.lpk_toe_snap(EPOCH + 5.0, 18.0, C.byref(t_toe), C.byref(wk), C.byref(toe)) == 0
    assert toe.value % 16.0 == 0.0 and abs(t_toe.value - (EPOCH + 5.0)) <= 8.0
    r, v = trajectory_state(ORBIT, t_toe.value)
    e = _lib.Ephemeris()
    assert lib.lpk_ephemeris_from_state((C.c_double * 3)(*r), (C.c_double * 3)(*v), 3, wk.value, toe.value,
                                        C.byref(e)) == 0
    assert abs(e.sqrt_a ** 2 - ORBIT.semi_major_axis_m) < 1e-3 and e.e < 1e-9
    assert abs(e.i0_rad - math.radians(55.0)) < 1e-12
    exact = kepler_eph(lib, epoch=EPOCH + 5.0)
    assert np.linalg.norm(sv(lib, e, toe.value) - sv(lib, exact, toe.value)) < 1e-3
    assert np.linalg.norm(sv(lib, e, toe.value) - r) < 1e-3
    ground = np.array([3.5e6, -2.1e6, 4.9e6])
    assert lib.lpk_ephemeris_from_state((C.c_double * 3)(*ground), (C.c_double * 3)(0, 0, 0), 3, wk.value, toe.value,
                                        C.byref(e)) == 0
    assert np.linalg.norm(sv(lib, e, toe.value) - ground) < 1.0


def rinex_eph(lib, prn=1, near=None):
    text = open(RINEX).read().encode()
    near = GPS_EPOCH_UNIX + 2317.0 * WEEK_S + 518400.0 - 18.0 if near is None else near
    e = _lib.Ephemeris()
    return lib.lpk_rinex3_gps_ephemeris(text, prn, near, 18.0, C.byref(e)), e


def test_rinex3_record_parses(lib):
    """src/signal/gps_nav.rs:851-868 on the reference's own record."""
    rc, e = rinex_eph(lib)
    assert rc == 0
    assert (e.week, e.toe_s, e.iode, e.iodc, e.health) == (2317, 518400.0, 61, 61, 0)
    assert e.sqrt_a == 5.153664415359e3 and e.e == 9.470977378078e-3 and e.i0_rad == 9.609727413912e-1
    assert e.delta_n_rad_s == 4.916182694377e-09 and e.af0 == 4.629502538592e-04 and e.af1 == -1.022364418051e-11
    assert e.crs == -1.435937500000e+02 and e.m0_rad == -2.838821577537 and e.tgd == -1.117587089539e-08
    assert e.omega0_rad == 1.665292991674 and e.argp_rad == 6.967118599040e-01
    assert e.omega_dot_rad_s == -8.108285197566e-09 and e.idot_rad_s == -2.500170875643e-10
    r = np.linalg.norm(sv(lib, e, e.toe_s))
    assert 26.2e6 < r < 26.9e6
    assert rinex_eph(lib, prn=7)[0] != 0       # PRN 7 absent


def test_broadcast_grid_is_the_sv_position(lib):
    """GpsBroadcastTrajectory on the trajectory grid (src/signal/gps_nav.rs:
    630-654, :870-880): node i is sv_position at toe + (t_i - toe_unix)."""
    _, e = rinex_eph(lib)
    toe_unix = GPS_EPOCH_UNIX + 2317.0 * WEEK_S + 518400.0 - 18.0
    n = lib.lpk_grid_size(2.0, 0.01)
    xyz = np.empty((3, n))
    t0 = C.c_double()
    assert lib.lpk_grid_fill_broadcast(C.byref(e), 18.0, toe_unix, 2.0, 0.01, C.byref(t0),
                                       xyz.ctypes.data_as(C.POINTER(C.c_double)), n) == 0
    assert t0.value == toe_unix - 0.5
    for i in (0, 50, n - 1):
        t = t0.value + i * 0.01
        np.testing.assert_array_equal(xyz[:, i], sv(lib, e, e.toe_s + (t - toe_unix)))
    assert abs(np.linalg.norm(xyz[:, 50]) - np.linalg.norm(sv(lib, e, e.toe_s))) < 1e-6


def test_scene_with_full_nav_and_rinex_prepares(tmp_path):
    """build_lnav_bits (src/pipeline.rs:255-295) through prepare(): a full-NAV
    C/A source on a circular orbit and one on the RINEX record, the latter's
    grid from the broadcast ephemeris; a TLE source is refused."""
    toe_unix = GPS_EPOCH_UNIX + 2317.0 * WEEK_S + 518400.0 - 18.0
    epoch = math.floor(toe_unix / 6.0) * 6.0
    txs = [TransmitterConfig("k", 1575.42e6, CaCode(1, epoch, "full"), CircularKeplerian(26.56e6, 55.0, 40.0, 120.0, epoch), 27.0),
           TransmitterConfig("r", 1575.42e6, CaCode(2, epoch, "full"), RinexNav(RINEX, 1), 27.0)]
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    from util import lab_receiver
    sc = prepare(SceneConfig(SceneMeta(epoch, 12.0), txs, [lab_receiver([band])]))
    k, r = sc.transmitters
    assert k.nav_mode == r.nav_mode == _lib.NAV_FULL
    assert len(k.nav_bits) == len(r.nav_bits) == 300 * 3
    decode_words(k.nav_bits)
    dr = decode_ephemeris(decode_words(r.nav_bits))
    assert abs(dr["sqrt_a"] - 5153.664415359) < 1e-5 and abs(dr["i0"] - 0.9609727413912) < 1e-8
    assert r.traj_kind == _lib.TRAJ_GRID and r.grid.shape[0] == 3
    assert 26.2e6 < np.linalg.norm(r.grid[:, 0]) < 26.9e6
    bad = [TransmitterConfig("t", 1575.42e6, CaCode(3, epoch), Tle("1 26407U", "2 26407"), 27.0)]
    with pytest.raises(ValueError, match="TLE parse"):
        prepare(SceneConfig(SceneMeta(epoch, 1.0), bad, [lab_receiver([band])]))
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        txs2 = [TransmitterConfig("s", 1575.42e6, CaCode(4, epoch + 1.0, "full"), StationaryEcef(2e7, 0.0, 1e7), 27.0)]
        prepare(SceneConfig(SceneMeta(epoch + 1.0, 6.0), txs2, [lab_receiver([band])]))
        assert any("6 s GPS boundary" in str(x.message) for x in w)
