"""Moving receivers on the host side: the position a CircularKeplerian
receiver has at a sample (src/geometry/kepler.rs:27-85, queried at epoch +
t_utc by the chunk body, src/pipeline.rs:929-930) from the oracle and from the
library's host builder, pinned by the orbit's closed form; the mirror's scene
preparation and descriptor fill for such a receiver."""
import ctypes as C
import math

import numpy as np
import pytest

from lipika_amd import _lib
from lipika_amd.config import (CaCode, CircularKeplerian, NodeClockSpec, ReceiverBandConfig, ReceiverConfig,
                               SceneConfig, SceneMeta, TransmitterConfig, Tle)
from lipika_amd.pipeline import prepare

import util

ELEMENTS = (6_778_137.0, 70.0, 320.0, 50.0, util.EPOCH)


def _orc_rx(orc, elements):
    from oracle.pyoracle import OrcRx
    rx = OrcRx()
    rx.traj_kind = 1
    rx.kep_a_m, rx.kep_inc_deg, rx.kep_raan_deg, rx.kep_nu0_deg, rx.kep_epoch_unix = elements
    return rx


def closed_form(a, inc_deg, raan_deg, nu0_deg, epoch, t):
    """Two-body circular orbit rotated into the Earth-fixed frame: Rz(-w t)
    Rz(node) Rx(i) a (cos u, sin u, 0), u = nu0 + n t."""
    n = math.sqrt(3.986005e14 / a ** 3)
    u = math.radians(nu0_deg) + n * (t - epoch)
    i, o = math.radians(inc_deg), math.radians(raan_deg)
    p = np.array([a * math.cos(u), a * math.sin(u), 0.0])
    rx = np.array([[1, 0, 0], [0, math.cos(i), -math.sin(i)], [0, math.sin(i), math.cos(i)]])
    rz = np.array([[math.cos(o), -math.sin(o), 0], [math.sin(o), math.cos(o), 0], [0, 0, 1]])
    th = 7.2921151467e-5 * (t - epoch)
    re = np.array([[math.cos(th), math.sin(th), 0], [-math.sin(th), math.cos(th), 0], [0, 0, 1]])
    return re @ (rz @ (rx @ p))


def test_oracle_and_host_receiver_positions_agree_and_follow_the_orbit(orc, lpk):
    rx = _orc_rx(orc, ELEMENTS)
    a = ELEMENTS[0]
    for t in (util.EPOCH - 3.0, util.EPOCH, util.EPOCH + 0.123456789, util.EPOCH + 5400.0, util.EPOCH + 86400.5):
        o, h = (C.c_double * 3)(), (C.c_double * 3)()
        orc.orc_rx_position(C.byref(rx), t, o)
        lpk.lpk_kepler_position(*ELEMENTS, t, h)
        o, h = np.array(o[:]), np.array(h[:])
        # both restate kepler.rs in its operation order with glibc's sincos: bit for bit
        assert np.array_equal(o, h), t
        assert abs(np.linalg.norm(o) - a) <= 1e-9 * a          # kepler.rs:110-124
        assert np.max(np.abs(o - closed_form(*ELEMENTS, t))) <= 1e-6
    # a stationary receiver is where it was put
    from oracle.pyoracle import OrcRx
    fixed = OrcRx()
    fixed.pos_ecef[:] = [1.0, 2.0, 3.0]
    p = (C.c_double * 3)()
    orc.orc_rx_position(C.byref(fixed), util.EPOCH + 1.0, p)
    assert list(p) == [1.0, 2.0, 3.0]


def test_prepare_takes_keplerian_receivers_and_refuses_others():
    band = ReceiverBandConfig("l1", 1575.42e6, 2.048e6, noise_power_w=0.0)
    tx = TransmitterConfig("sv", 1575.42e6, CaCode(3, util.EPOCH), CircularKeplerian(26.56e6, 55.0, 10.0, 20.0, util.EPOCH),
                           27.0)
    sc = prepare(SceneConfig(SceneMeta(util.EPOCH, 0.01), [tx],
                             [ReceiverConfig("orbiter", CircularKeplerian(*ELEMENTS), NodeClockSpec(), [band])]))
    r = sc.receivers[0]
    assert r.pos is None and r.kepler == tuple(float(v) for v in ELEMENTS)
    tle = Tle("1 25544U 98067A   24156.50000000  .00016717  00000-0  10270-3 0  9005",
              "2 25544  51.6416 247.4627 0006703 130.5360 325.0288 15.49815350 45234")
    with pytest.raises(NotImplementedError):
        prepare(SceneConfig(SceneMeta(util.EPOCH, 0.01), [tx], [ReceiverConfig("iss", tle, NodeClockSpec(), [band])]))
    bad = CircularKeplerian(-1.0, 70.0, 320.0, 50.0, util.EPOCH)
    with pytest.raises(ValueError):
        prepare(SceneConfig(SceneMeta(util.EPOCH, 0.01), [tx], [ReceiverConfig("bad", bad, NodeClockSpec(), [band])]))


def test_receiver_descriptor_carries_the_trajectory():
    d = _lib.RxDesc()
    assert d.traj_kind == _lib.RX_STATIONARY
    d.traj_kind = _lib.RX_KEPLER
    d.kepler_a_m = 7e6
    assert C.sizeof(_lib.RxDesc) == 120 and _lib.RxDesc.kepler_epoch_unix.offset == 112
