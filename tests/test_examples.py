"""The reference's example scenes (examples/*.toml, kept as data fixtures in
tests/golden/) read and prepare on this path: every transmitter kind and
trajectory source they use (CW, BPSK, OOK, 16-QAM, C/A with full LNAV;
circular orbits, TLEs, a RINEX broadcast record, Earth-fixed sites)."""
import os

import numpy as np
import pytest

from lipika_amd import _lib
from lipika_amd.config import RinexNav, SceneConfig
from lipika_amd.pipeline import prepare

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EXAMPLES = ["all_modulations", "dual_band", "gps_broadcast", "gps_fix_synthetic", "gps_real_tle", "iss_uhf_pass"]


def load_example(name, duration_s=None):
    """The example scene, its RINEX file taken from the fixtures."""
    with open(os.path.join(GOLDEN, name + ".toml"), encoding="utf-8") as f:
        scene = SceneConfig.from_toml_str(f.read())
    for tx in scene.transmitters:
        if isinstance(tx.trajectory, RinexNav):
            tx.trajectory = RinexNav(os.path.join(GOLDEN, os.path.basename(tx.trajectory.file)), tx.trajectory.prn)
    if duration_s is not None:
        scene.scene.duration_s = duration_s
    return scene


@pytest.mark.parametrize("name", EXAMPLES)
def test_example_scene_prepares(name):
    scene = load_example(name)
    sc = prepare(scene)
    assert len(sc.transmitters) == len(scene.transmitters)
    for t in sc.transmitters:
        if t.traj_kind == _lib.TRAJ_GRID:
            assert np.all(np.isfinite(t.grid))
            r = np.linalg.norm(t.grid, axis=0)
            assert np.all((r > 6.3e6) & (r < 4.3e7)), t.name
        if t.nav_mode == _lib.NAV_FULL:
            assert len(t.nav_bits) >= 300 and set(np.unique(t.nav_bits)) <= {-1, 1}
    n_active = [len(b.active) for r in sc.receivers for b in r.bands]
    print(f"{name}: {len(sc.transmitters)} transmitters, in-band per band {n_active}", flush=True)
    assert sum(n_active) > 0
