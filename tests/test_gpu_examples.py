"""The reference's example scenes on the GPU path: the first 50 ms of every
band of each example (tests/test_examples.py reads them), in the plain form
against the oracle with the oracle's clock offsets, and in the precision
form within its bounds."""
import numpy as np
import pytest

from lipika_amd import _lib
from lipika_amd.pipeline import prepare
from oracle.pyoracle import OracleScene

from test_examples import EXAMPLES, load_example
from test_gpu_parity import compare_iq, gpu_band, unit_index

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", EXAMPLES)
def test_example_scene_renders_like_the_oracle(backend, name):
    sc = prepare(load_example(name, duration_s=0.05))
    seed = 12
    backend.upload(sc, seed)
    orc = OracleScene(sc, seed)
    for r, rx in enumerate(sc.receivers):
        for b, pb in enumerate(rx.bands):
            if not len(pb.active):
                continue
            w = unit_index(sc, r, b)
            o, _, taps = orc.render_band(w, tap_slot=0)
            g, _ = gpu_band(backend, sc, w, r, b, seed, delta_t=taps[3])
            # f32 sincos of CW baseband offsets: device and glibc may differ by one ulp
            assert compare_iq(g, o, f"{name} {rx.name}/{pb.band.band_id}, plain", strict=False) > 0.5
            p, _ = gpu_band(backend, sc, w, r, b, seed, form=_lib.FORM_PRECISION, delta_t=taps[3])
            peak = float(np.max(np.abs(o)))
            d = np.abs(p.astype(np.complex128) - o.astype(np.complex128))
            print(f"{name} {rx.name}/{pb.band.band_id}, precision: max|d|/peak={d.max() / max(peak, 1e-300):.3e}",
                  flush=True)
            assert np.all(np.isfinite(p)) and d.max() <= 2e-6 * peak + 1e-30
