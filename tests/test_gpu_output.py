"""The writers on the GPU (tests/streaming_chunk_invariance.rs's contract):
bands streamed to .iq.c64 files through lpk_stream_to_files hold the same
bytes as the in-memory run_scene, whatever the chunk size; the sidecars are
the same whether the IQ was streamed or kept in memory (fn-6); the truth
export's ranges are the oracle's four-iteration solve, bit for bit."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from lipika_amd import output, workloads
from lipika_amd.pipeline import prepare, run_scene
from oracle.pyoracle import OracleScene

import util

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_streamed_bytes_equal_in_memory_and_any_chunk(backend, tmp_path):
    scene = util.all_kinds_scene(duration_s=0.02, noise_power_w=None)
    seed = 4242
    mem = run_scene(scene, seed, chunk_samples=1 << 20, backend=backend)
    d1, d2 = str(tmp_path / "a"), str(tmp_path / "b")
    m1 = output.run_scene_to_dir(scene, seed, d1, chunk_samples=30_000, backend=backend)
    m2 = output.run_scene_to_dir(scene, seed, d2, chunk_samples=160_001, backend=backend)
    assert [m.name for m in m1] == [o.name for o in mem] == [m.name for m in m2]
    for o, a, b in zip(mem, m1, m2):
        fa, fb = open(a.path, "rb").read(), open(b.path, "rb").read()
        want = np.ascontiguousarray(o.iq, dtype="<c8").tobytes()
        _say(f"{o.name}: {len(want)} bytes; streamed (30 000-sample chunks) equal: {fa == want}; "
             f"(160 001-sample chunks) equal: {fb == want}")
        assert len(want) == 8 * a.n_samples > 0
        assert fa == want and fb == want
        # the sidecar of the streamed band equals the in-memory one's
        output.write_raw_with_sidecar(o, str(tmp_path / "mem"))
        side_stream = open(os.path.join(d1, f"{o.name}.iq.json"), "rb").read()
        side_mem = open(str(tmp_path / "mem" / f"{o.name}.iq.json"), "rb").read()
        assert side_stream == side_mem
        assert open(str(tmp_path / "mem" / f"{o.name}.iq.c64"), "rb").read() == want
        side = json.loads(side_stream)
        assert side["n_samples"] == len(o.iq) and side["sample_rate_hz"] == 8e6


def test_truth_export_is_the_four_iteration_solve(backend, tmp_path, orc):
    scene = workloads.scene2(duration_s=2.0)
    output.export_truth(scene, str(tmp_path), 10.0, backend=backend)
    truth = json.load(open(str(tmp_path / "glasgow__truth.json")))
    assert truth["grid_hz"] == 10.0 and len(truth["t_utc"]) == 21
    sc = prepare(scene)
    osc = OracleScene(sc, 0)
    obs = (C.c_double * 3)(*[float(v) for v in sc.receivers[0].pos])
    epoch = scene.scene.epoch_unix
    t_rel = [k * 0.1 for k in range(21)]
    assert truth["t_utc"] == [epoch + t for t in t_rel]
    for i, tx in enumerate(truth["transmitters"]):
        want = []
        for t in t_rel:
            te, r = C.c_double(), C.c_double()
            orc.orc_solve_emission(C.byref(osc.txs[i]), obs, t, epoch, 4, C.byref(te), C.byref(r))
            want.append(r.value)
        assert tx["range_m"] == want, tx["name"]
        assert tx["delay_s"] == [r / 299792458.0 for r in want]
        rr = tx["range_rate_m_s"]
        assert abs(rr[5] - (want[6] - want[4]) / 0.2) < 1e-9 and abs(rr[0] - (want[1] - want[0]) / 0.1) < 1e-9
        src = (C.c_double * 3)()
        orc.orc_grid_position(osc.txs[i].grid_t0_unix, osc.txs[i].grid_dt_s, osc.txs[i].grid_n, osc.txs[i].grid_xyz,
                              epoch + t_rel[7], src)
        assert [tx["source_x_m"][7], tx["source_y_m"][7], tx["source_z_m"][7]] == list(src[:])
    _say(f"truth: {len(truth['transmitters'])} transmitters x 21 grid times equal the oracle's 4-iteration solve")


def test_sink_stream_equals_chunked_renders_and_stops_per_band(backend):
    """lpk_stream_to_sink (generate_band's chunk loop with IqSink::push,
    src/pipeline.rs:905-990, :984): every band's chunks arrive in order and
    equal eager renders of the same chunks; the library's FNV-1a sink hashes
    exactly those bytes; a sink that returns Break stops its own band only."""
    from lipika_amd import _lib
    scene = workloads.scene2(duration_s=0.3)
    sc = prepare(scene)
    seed = 77
    backend.upload(sc, seed)
    units = [(r, b) for r, rx in enumerate(sc.receivers) for b, _ in enumerate(rx.bands)]
    chunk = 100_000
    whole, keep = [], []
    for w, (r, b) in enumerate(units):
        j, k = backend.job(r, b, w, seed, 0, sc.receivers[r].bands[b].n)
        whole.append(j)
        keep.append(k)
    got = {w: [] for w in range(len(units))}

    def sink(band, i0, iq):
        got[band].append((i0, iq.copy()))
        return False

    backend.stream_to_sink(whole, chunk, _lib.FORM_PRECISION, sink)
    lib = _lib.load()
    for w, (r, b) in enumerate(units):
        n = sc.receivers[r].bands[b].n
        want = np.empty(n, dtype=np.complex64)
        for i0 in range(0, n, chunk):
            j, k = backend.job(r, b, w, seed, i0, min(chunk, n - i0), out_host=want[i0:])
            backend.render([j], _lib.FORM_PRECISION)
        backend.sync()
        assert [i0 for i0, _ in got[w]] == list(range(0, n, chunk))
        assert np.concatenate([c for _, c in got[w]]).tobytes() == want.tobytes()
        got[w] = want
    hashes = backend.stream_to_sink(whole, chunk, _lib.FORM_PRECISION)
    for w in range(len(units)):
        assert hashes[w] == lib.lpk_fnv1a64(got[w].ctypes.data, got[w].nbytes, _lib.FNV_OFFSET)
    seen = {w: 0 for w in range(len(units))}

    def stopper(band, i0, iq):
        seen[band] += 1
        return band == 1 and seen[band] == 2   # Break after the L1 band's second chunk

    backend.stream_to_sink(whole, chunk, _lib.FORM_PLAIN, stopper)
    assert seen[1] == 2
    assert seen[0] == -(-sc.receivers[0].bands[0].n // chunk)
    _say(f"sink stream: {[len(v) for v in got.values()]} samples per band; hashes {[hex(h) for h in hashes]}")
