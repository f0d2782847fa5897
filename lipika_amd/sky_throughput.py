"""Throughput of render_window on a sky of many transmitters, with and without
the horizon gate (examples/sky_throughput.rs of the reference).

    python -m lipika_amd.sky_throughput [--sky SKY.toml | --leos N] [--duration S] [--form plain|precision]
                                        [--chunk N] [--repeat K]

One receiver (Glasgow: 55.86 N, 4.25 W, 20 m; 290 K) with an isotropic
antenna that sees nothing below the horizon (cutoff 0 deg), tuned to
401.15 MHz at 48 kHz, rendering `duration` seconds (default 60) from Unix
1717260690, once with the gate off and once with it on, `repeat` times
(default 2; the first pass includes warm-up). The samples are hashed, not
kept: both settings must print the same hash. Without --sky the sky is N
seeded circular LEOs (default 200) carrying CW tones 20 Hz apart.

Printed per pass: wall time, the render kernels' device time (HIP events),
emulated seconds per wall second, the transmitters in band and visible, and
the gate's tested / skipped (tile, transmitter) pairs.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import time

import numpy as np

from . import _lib
from .config import (CircularKeplerian, Cw, FrontendConfig, ReceiverBandConfig, ReceiverConfig, SkyConfig,
                     StationaryGeodetic, TransmitterConfig)
from .pipeline import Backend, WindowOptions, render_window

CARRIER_HZ = 401_150_000.0
FS = 48_000.0
START_UNIX = 1_717_260_690.0


def leo_sky(n: int, seed: int = 5) -> SkyConfig:
    """n circular LEOs at seeded radii 6 900-7 500 km and orientations."""
    rng = np.random.default_rng(seed)
    return SkyConfig([TransmitterConfig(f"leo{k}", CARRIER_HZ + 20.0 * k, Cw(0.0),
                                        CircularKeplerian(float(rng.uniform(6.9e6, 7.5e6)), float(rng.uniform(0, 180)),
                                                          float(rng.uniform(0, 360)), float(rng.uniform(0, 360)),
                                                          START_UNIX), 0.0)
                      for k in range(n)])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sky", help="sky TOML (SkyConfig)")
    ap.add_argument("--leos", type=int, default=200)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--form", default="plain", choices=["plain", "precision"])
    ap.add_argument("--chunk", type=int, default=1 << 20)
    ap.add_argument("--repeat", type=int, default=2)
    args = ap.parse_args(argv)
    sky = SkyConfig.from_toml_file(args.sky) if args.sky else leo_sky(args.leos)
    rx = ReceiverConfig("glasgow", StationaryGeodetic(55.86, -4.25, 20.0))
    fe = FrontendConfig(antenna={"peak_dbi": 0.0, "horizon_dbi": 0.0, "fb_db": 0.0, "cutoff_deg": 0.0})
    band = ReceiverBandConfig("uhf", CARRIER_HZ, FS, 290.0, frontend=fe)
    lib = _lib.load()
    be = Backend(0)
    try:
        digests = {}
        for rep in range(args.repeat):
            for gate in (False, True):
                be.set_horizon_gate(gate)
                before = be.horizon_gate_stats()
                h = hashlib.blake2b(digest_size=8)
                be.check(lib.lpk_timing_reset(be.h))
                t0 = time.perf_counter()
                s = render_window(sky, rx, band, START_UNIX, args.duration, WindowOptions(3, args.chunk, 1),
                                  lambda c: h.update(c.tobytes()), form=args.form, backend=be)
                wall = time.perf_counter() - t0
                km, nl = C.c_double(), C.c_uint64()
                be.check(lib.lpk_timing_read(be.h, C.byref(km), C.byref(nl)))
                after = be.horizon_gate_stats()
                digests.setdefault(gate, set()).add(h.hexdigest())
                print(f"pass {rep} gate {'on ' if gate else 'off'} ({args.form}): {wall * 1e3:8.1f} ms wall, "
                      f"render kernels {km.value:7.2f} ms over {nl.value} calls, "
                      f"{args.duration / wall:8.1f} emulated s per s; "
                      f"{len(s.contributing_transmitters)} in band, {len(s.visible_transmitters)} visible; "
                      f"pairs tested {after['tested'] - before['tested']}, skipped {after['skipped'] - before['skipped']}; "
                      f"hash {h.hexdigest()}", flush=True)
        same = digests[False] == digests[True] and len(digests[True]) == 1
        print("the gate does not change the samples" if same else "HASH MISMATCH between gate on and off", flush=True)
        return 0 if same else 1
    finally:
        be.set_horizon_gate(True)
        be.close()


if __name__ == "__main__":
    raise SystemExit(main())
