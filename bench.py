"""Benchmark of the per-chunk IQ generation path on one MI355X.

Workload (one "step"): the whole dual-band scene of benchmark configuration 2
rendered once -- BPSK LEO at 437 MHz (48 kHz band) plus six C/A-code MEO
transmitters at 1575.42 MHz (2.048 MHz band), receiver clock walk on, kTB
noise on, 60 s of emulated time, 1 Mi-sample chunks. Every chunk job of the
scene is submitted in one render call (one kernel launch); the IQ stays in
HBM. Synthetic geometry and payloads; inputs are resident in HBM before the
timed region starts.

Metric: contributions/s = sum over (receiver band, in-band transmitter) of
samples / wall time, in Msamples/s. Output samples/s is printed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form plain|precision]
Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s (rx-band×TX summed) at ≤1e-6 rad phase err; plain vs chosen form"
FLOPS_PER_CONTRIBUTION = 75        # FP64 work per contribution, counted from the reference source
PEAK_FP64_TFLOPS = 78.6            # MI355X published FP64 peak
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_leg(args, scene_n):
    from lipika_amd import _lib, workloads
    from lipika_amd.pipeline import Backend, form_code, prepare

    seed = 2026
    sc = prepare(workloads.scene(scene_n))
    out_samples, contributions = workloads.totals(sc)
    be = Backend(0)
    lib = _lib.load()
    be.upload(sc, seed)
    chunk = 1 << 20
    bufs, jobs, keep = [], [], []
    w = 0
    for r, rx in enumerate(sc.receivers):
        for b, pb in enumerate(rx.bands):
            dev = C.c_void_p()
            be.check(lib.lpk_device_alloc(be.h, max(pb.n, 1) * 8, C.byref(dev)))
            bufs.append(dev)
            for i0 in range(0, pb.n, chunk):
                j, k = be.job(r, b, w, seed, i0, min(chunk, pb.n - i0), out_dev=dev.value + i0 * 8)
                jobs.append(j)
                keep.append(k)
            w += 1
    form = form_code(args.form)
    log(f"[bench] config {scene_n}: {len(jobs)} chunk jobs, {out_samples} output samples, "
        f"{contributions} contributions per step; form={args.form}")
    for s in range(args.warmup):
        be.render(jobs, form)
    be.sync()
    log(f"[bench] {args.warmup} warmup steps done")
    be.check(lib.lpk_timing_reset(be.h))
    be.sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        be.render(jobs, form)
    be.sync()
    t1 = time.perf_counter()
    kern_ms, n_launch = C.c_double(), C.c_uint64()
    be.check(lib.lpk_timing_read(be.h, C.byref(kern_ms), C.byref(n_launch)))
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kernel_ms = kern_ms.value / max(n_launch.value, 1)
    log(f"[bench] {args.steps} timed steps: {ms_per_step:.3f} ms/step; render kernel {kernel_ms:.3f} ms/launch "
        f"over {n_launch.value} launches")
    for d in bufs:
        be.check(lib.lpk_device_free(be.h, d))
    be.close()
    return sc, out_samples, contributions, ms_per_step, kernel_ms


def cpu_leg(sc, seed, target_s):
    """The oracle (the project's C restatement of the reference CPU path,
    with its two-level parallelism) on a bounded prefix of the same scene."""
    from oracle.pyoracle import OracleScene
    threads = min(os.cpu_count() or 1, 16)
    orc = OracleScene(sc, seed)
    dur = sc.meta.duration_s

    def run(frac):
        nl = [max(1, int(pb.n * frac)) for _, pb in orc.units]
        t = time.perf_counter()
        contrib, _ = orc.render_scene_hashed(nl, 1 << 20, threads)
        return contrib, time.perf_counter() - t

    c0, t0 = run(0.01)
    frac = min(1.0, 0.01 * target_s / max(t0, 1e-3))
    c, t = run(frac)
    log(f"[bench] cpu baseline: {c} contributions ({frac * dur:.2f} s of {dur} s emulated) in {t:.2f} s "
        f"on {threads} threads")
    return {"value": c / t / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"first {frac * dur:.2f} s of every band of the same scene ({c} contributions, "
                      f"{t:.1f} s wall); the oracle restatement of the reference CPU path, bands and "
                      f"transmitters in parallel, FNV-1a sink"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", default="plain", choices=["plain", "precision"])
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the path shards by receiver with no collective; see the design notes).", file=sys.stderr)
        sys.exit(2)
    sc, out_samples, contributions, ms, kernel_ms = gpu_leg(args, args.config)
    value = contributions / (ms / 1e3) / 1e6
    achieved = contributions * FLOPS_PER_CONTRIBUTION / (kernel_ms / 1e3) / 1e12
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"config {args.config}: " + {
            2: "BPSK LEO 437 MHz + 6 C/A MEO 1575.42 MHz; bands 48 kHz and 2.048 MHz; clock walk and kTB noise on; 60 s",
        }.get(args.config, "see lipika_amd/workloads.py"),
            "form": args.form, "chunk_samples": 1 << 20, "contributions_per_step": contributions,
            "output_samples_per_step": out_samples,
            "output_msamples_per_s": round(out_samples / (ms / 1e3) / 1e6, 3)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "kernel": "render_kernel", "kernel_ms": round(kernel_ms, 4),
                     "flops_per_contribution": FLOPS_PER_CONTRIBUTION,
                     "output_gb_s": round(out_samples * 8 / (kernel_ms / 1e3) / 1e9, 2)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_leg(sc, 2026, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
