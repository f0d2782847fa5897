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

Configurations 4 (uneven bands, the batching form) and 5 (the learned warm
start) are measured after the headline and reported under config.

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
# HBM bytes per configuration-2 render dispatch (reads + writes), from the rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a copy kernel of known size: profiles/render_config2_hbm_traffic.md
TRAFFIC_CONFIG2 = {"plain": 2.063e9, "precision": 2.073e9}
WORKLOADS = {
    2: "BPSK LEO 437 MHz + 6 C/A MEO 1575.42 MHz; bands 48 kHz and 2.048 MHz; clock walk and kTB noise on; 60 s",
    4: "40 transmitters (CW, OOK, BPSK, 16-QAM, C/A) x 64 receivers with 130 bands at 48 kHz to 8.184 MHz; "
       "clock walks and kTB noise on; 30 s; every chunk of every band in one launch",
    5: "32 C/A MEO transmitters x 256 receivers, one L1 band each at 2.048 MHz, seeded clocks, kTB noise on; "
       "the first {chunks} Mi-sample chunks of every band ({secs:.3f} s of the 60 s scene) per step",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plan_leg(be, lib, sc, bufs, seed, form, steps, warmup, chunk=1 << 24):
    """The batching form: a streaming plan renders chunk k of every band per
    step as one graph replay, with the context's (autotuned) tuning; a pass is
    every step of the scene. Chunks of 16 Mi samples per band: the replays
    then hold enough tiles to fill the device (configuration 4: 15 replays;
    1 Mi-sample chunks take 235 replays whose tails show, and 1.4x the time),
    at about 4 GB of clock-walk windows for its 64 receivers."""
    from lipika_amd.pipeline import form_code
    units = [(r, b) for r, rx in enumerate(sc.receivers) for b, _ in enumerate(rx.bands)]
    tuning = be.tuning()
    bands, keep = [], []
    for w, (r, b) in enumerate(units):
        j, k = be.job(r, b, w, seed, 0, sc.receivers[r].bands[b].n, out_dev=bufs[w].value)
        bands.append(j)
        keep.append(k)
    with be.plan(bands, chunk, form_code(form)) as plan:
        def one_pass():
            for kk in range(plan.n_chunks):
                plan.step(kk)
        for _ in range(warmup):
            one_pass()
        be.sync()
        be.check(lib.lpk_timing_reset(be.h))
        t0 = time.perf_counter()
        for _ in range(steps):
            one_pass()
        be.sync()
        t1 = time.perf_counter()
        kern_ms, n_launch = C.c_double(), C.c_uint64()
        be.check(lib.lpk_timing_read(be.h, C.byref(kern_ms), C.byref(n_launch)))
        ms = (t1 - t0) * 1e3 / steps
        log(f"[bench] plan ({form}): {plan.n_chunks} graph replays per pass, {steps} timed passes: "
            f"{ms:.3f} ms per pass; graphs {kern_ms.value / steps:.3f} ms per pass")
        return {"form": form, "tile_shift": tuning[0], "xcd_map": tuning[1], "chunk_samples": chunk,
                "graph_replays_per_step": plan.n_chunks, "max_tiles_per_replay": plan.max_tiles,
                "ms_per_step": round(ms, 4), "graph_ms_per_step": round(kern_ms.value / steps, 4)}


def gpu_leg(args, scene_n, forms, steps, warmup, batching=False):
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
    log(f"[bench] config {scene_n}: {len(jobs)} chunk jobs, {out_samples} output samples, "
        f"{contributions} contributions per step")
    # the autotuner's sample: chunk 0 of every band
    first = [jobs[i] for i in range(len(jobs)) if jobs[i].i0 == 0]

    def step(form):
        # a step is a whole fresh render: the clock walk is regenerated from index 0
        be.check(lib.lpk_clock_rewind(be.h))
        be.render(jobs, form)

    results = {}
    for name in forms:
        form = form_code(name)
        tuning, tune_ms = be.autotune(first, form)
        log(f"[bench] config {scene_n} form {name}: autotune on chunk 0 of every band picks tile 2^{tuning[0]}, "
            f"xcd map {tuning[1]} ({tune_ms:.3f} ms)")
        for s in range(warmup):
            step(form)
        be.sync()
        be.check(lib.lpk_timing_reset(be.h))
        t0 = time.perf_counter()
        for s in range(steps):
            step(form)
        be.sync()
        t1 = time.perf_counter()
        kern_ms, n_launch = C.c_double(), C.c_uint64()
        be.check(lib.lpk_timing_read(be.h, C.byref(kern_ms), C.byref(n_launch)))
        ms_per_step = (t1 - t0) * 1e3 / steps
        kernel_ms = kern_ms.value / max(n_launch.value, 1)
        log(f"[bench] config {scene_n} form {name}: {steps} timed steps after {warmup} warmup: {ms_per_step:.3f} ms/step; "
            f"render kernels {kernel_ms:.3f} ms per render call over {n_launch.value} calls")
        results[name] = (ms_per_step, kernel_ms, tuning)
    plan = plan_leg(be, lib, sc, bufs, seed, "precision", steps, warmup) if batching else None
    for d in bufs:
        be.check(lib.lpk_device_free(be.h, d))
    be.close()
    return sc, out_samples, contributions, results, plan


def learned_leg(steps, warmup, chunks=2):
    """Configuration 5: the learned warm start. The network is trained on the
    device (1e7 pairs from the device's converged light-time solve), then the
    scene prefix is rendered in the plain form (the reference's two fixed
    iterations), the converged form (iterated to a 1e-13 s update from the
    reference's start) and the learned form (the same, from the network's
    prediction), reporting emission-equation evaluations per contribution."""
    from lipika_amd import _lib, workloads
    from lipika_amd.pipeline import Backend, WarmStartNet, form_code, prepare

    seed = 2026
    sc = prepare(workloads.scene5())
    be = Backend(0)
    lib = _lib.load()
    be.upload(sc, seed)
    net = WarmStartNet(be, init_seed=1)
    train = net.train(n_pairs=10_000_000, data_seed=seed, epochs=16, batch=16384, lr=3e-3, lr_final=3e-5,
                      n_val=200_000)
    log(f"[bench] config 5: trained the warm-start network: {train}")
    be.warm_start(net)
    chunk = 1 << 20
    bufs, jobs, keep = [], [], []
    contributions = 0
    w = 0
    for r, rx in enumerate(sc.receivers):
        for b, pb in enumerate(rx.bands):
            n = min(pb.n, chunks * chunk)
            dev = C.c_void_p()
            be.check(lib.lpk_device_alloc(be.h, max(n, 1) * 8, C.byref(dev)))
            bufs.append(dev)
            for i0 in range(0, n, chunk):
                j, k = be.job(r, b, w, seed, i0, min(chunk, n - i0), out_dev=dev.value + i0 * 8)
                jobs.append(j)
                keep.append(k)
            contributions += n * len(pb.active)
            w += 1
    first = [j for j in jobs if j.i0 == 0]
    forms = {}
    for name in ("plain", "converged", "learned"):
        form = form_code(name)
        tuning, _ = be.autotune(first, form)

        def step():
            be.check(lib.lpk_clock_rewind(be.h))
            be.render(jobs, form)

        for _ in range(warmup):
            step()
        be.sync()
        be.solve_stats_reset()
        be.check(lib.lpk_timing_reset(be.h))
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        be.sync()
        t1 = time.perf_counter()
        kern_ms, n_launch = C.c_double(), C.c_uint64()
        be.check(lib.lpk_timing_read(be.h, C.byref(kern_ms), C.byref(n_launch)))
        ms = (t1 - t0) * 1e3 / steps
        entry = {"value": round(contributions / (ms / 1e3) / 1e6, 3), "ms_per_step": round(ms, 4),
                 "kernel_ms": round(kern_ms.value / max(n_launch.value, 1), 4), "tile_shift": tuning[0],
                 "xcd_map": tuning[1]}
        if name == "plain":
            entry["evaluations_per_contribution"] = 2.0    # the reference's fixed two iterations
        else:
            st = be.solve_stats()
            entry["evaluations_per_contribution"] = round(st["evaluations"] / max(st["contributions"], 1), 4)
            entry["fallbacks"] = st["fallbacks"]
            entry["unconverged"] = st["unconverged"]
        log(f"[bench] config 5 form {name}: {ms:.3f} ms/step, {entry}")
        forms[name] = entry
    net.close()
    for d in bufs:
        be.check(lib.lpk_device_free(be.h, d))
    be.close()
    secs = chunks * chunk / 2.048e6
    lf, cf, pf = forms["learned"], forms["converged"], forms["plain"]
    return {"workload": "config 5: " + WORKLOADS[5].format(chunks=chunks, secs=secs),
            "contributions_per_step": contributions, "steps": steps, "forms": forms,
            "evaluations_saved_per_contribution": round(cf["evaluations_per_contribution"]
                                                        - lf["evaluations_per_contribution"], 4),
            "learned_speedup_vs_converged": round(cf["ms_per_step"] / lf["ms_per_step"], 3),
            "learned_speedup_vs_plain": round(pf["ms_per_step"] / lf["ms_per_step"], 3),
            "training": {"pairs": 10_000_000, "epochs": 16, "batch": 16384, "steps": train["steps"],
                         "train_ms": round(train["train_ms"], 1), "pair_generation_ms": round(train["gen_ms"], 2),
                         "validation_rms_s": train["val_rms_s"], "validation_max_s": train["val_max_s"]}}


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


def copy_probe():
    """A plain device copy of a known size (3 x 256 MiB), for calibrating the
    FETCH_SIZE / WRITE_SIZE counters against the render kernel's."""
    import torch
    n = 256 << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: 3 copies of {n} bytes")


def main():
    if "--copy-probe" in sys.argv:
        copy_probe()
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", default="precision", choices=["plain", "precision"],
                    help="the form the headline value is quoted for (both are timed)")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config4", dest="config4", action="store_false",
                    help="skip the configuration-4 (batching) measurement")
    ap.add_argument("--no-config5", dest="config5", action="store_false",
                    help="skip the configuration-5 (learned warm start) measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the path shards by receiver with no collective; see the design notes).", file=sys.stderr)
        sys.exit(2)
    forms = ["plain", "precision"] if args.form == "precision" else ["plain"]
    sc, out_samples, contributions, results, _ = gpu_leg(args, args.config, forms, args.steps, args.warmup)
    ms, kernel_ms, tuned = results[args.form]
    value = contributions / (ms / 1e3) / 1e6
    achieved = contributions * FLOPS_PER_CONTRIBUTION / (kernel_ms / 1e3) / 1e12

    def summary(results, contributions):
        out = {n: {"value": round(contributions / (m / 1e3) / 1e6, 3), "ms_per_step": round(m, 4),
                   "kernel_ms": round(k, 4), "tile_shift": t[0], "xcd_map": t[1]}
               for n, (m, k, t) in results.items()}
        if "plain" in results and "precision" in results:
            out["precision"]["speedup_vs_plain"] = round(results["plain"][0] / results["precision"][0], 3)
        return out

    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"config {args.config}: " + WORKLOADS.get(args.config, "see lipika_amd/workloads.py"),
                   "form": args.form, "chunk_samples": 1 << 20, "contributions_per_step": contributions,
                   "output_samples_per_step": out_samples,
                   "output_msamples_per_s": round(out_samples / (ms / 1e3) / 1e6, 3),
                   "forms": summary(results, contributions)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": TRAFFIC_CONFIG2[args.form] if args.config == 2 else None,
                     "traffic_source": "profiles/render_config2_hbm_traffic.md",
                     "compute_unit": "FP64 vector ALU (the path issues no matrix ops)",
                     "kernel": f"render_kernel<{0 if args.form == 'plain' else 1}, false, {tuned[0]}>",
                     "kernel_ms": round(kernel_ms, 4),
                     "flops_per_contribution": FLOPS_PER_CONTRIBUTION,
                     "output_gb_s": round(out_samples * 8 / (kernel_ms / 1e3) / 1e9, 2)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_leg(sc, 2026, args.cpu_seconds)
    if args.config4 and args.config != 4:
        # the speed-up targets (against the plain form and the CPU path) are stated on
        # configuration 4 (uneven bands)
        sc4, out4, contrib4, res4, plan4 = gpu_leg(args, 4, ["plain", "precision"], max(1, args.steps // 5), 1,
                                                    batching=True)
        c4 = {"workload": "config 4: " + WORKLOADS[4], "contributions_per_step": contrib4,
              "output_samples_per_step": out4, "forms": summary(res4, contrib4)}
        plan4["value"] = round(contrib4 / (plan4["ms_per_step"] / 1e3) / 1e6, 3)
        plan4["speedup_vs_plain"] = round(res4["plain"][0] / plan4["ms_per_step"], 3)
        c4["batching"] = plan4
        if not args.no_cpu:
            c4["cpu_baseline"] = cpu_leg(sc4, 2026, args.cpu_seconds)
            cpu = c4["cpu_baseline"]["value"]
            c4["precision_vs_cpu"] = round(c4["forms"]["precision"]["value"] / cpu, 1)
        res["config"]["config4"] = c4
    if args.config5:
        res["config"]["config5"] = learned_leg(max(1, args.steps // 10), 1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
