"""Benchmark of the per-chunk IQ generation path on one MI355X.

Headline workload (one "step"): benchmark configuration 4 rendered once --
40 transmitters (8 each of CW, OOK 1 kbps, BPSK 9600 bps, 16-QAM 24 kBd and
C/A code, on seeded LEO and MEO circular orbits, carriers 437 MHz,
1575.42 MHz and 2.4 GHz) seen by 64 receivers with 130 bands of 48 kHz to
8.184 MHz, receiver clock walks and kTB noise on, 30 s of emulated time,
1 Mi-sample chunks: 9 879 chunk jobs, all of them in one render call, the IQ
kept in HBM. The headline form is the precision form (the north-star target
form); the plain form (the reference's f64 path, operation for operation) and
the batching form (a streaming plan: one step per chunk, alternating between
two streams) are timed beside it. Synthetic geometry and payloads; inputs are resident in HBM
before the timed region starts.

Metric: contributions/s = sum over (receiver band, in-band transmitter) of
samples / wall time, in Msamples/s. Output samples/s is printed beside it.

Also measured, under config: configuration 3 (the precision form's own
configuration, with its phase, delay and IQ errors against the FP64 oracle),
configuration 2, configuration 5 (the learned warm start) and the drop-in
call pattern (one lpk_render_chunk per band chunk into pinned host memory, and
lpk_stream_to_sink, with the reference's FNV-1a hashing sink).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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
PEAK_FP64_TFLOPS = 78.6            # MI355X published FP64 vector peak
C_LIGHT = 299792458.0
# HBM bytes per render dispatch (reads + writes) of the precision form, from the rocprofv3
# FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel of known size.
TRAFFIC = {4: (159.94e9, "profiles/render_config4_hbm_traffic.md"),
           2: (2.073e9, "profiles/render_config2_hbm_traffic.md")}
WORKLOADS = {
    2: "BPSK LEO 437 MHz + 6 C/A MEO 1575.42 MHz; bands 48 kHz and 2.048 MHz; clock walk and kTB noise on; 60 s",
    3: "32 C/A MEO transmitters (PRNs 1-32, 6 planes) x 16 receivers, one L1 band each at 2.048 MHz, seeded "
       "clocks, kTB noise on; 120 s",
    4: "40 transmitters (CW, OOK, BPSK, 16-QAM, C/A) x 64 receivers with 130 bands at 48 kHz to 8.184 MHz; "
       "clock walks and kTB noise on; 30 s; every chunk of every band in one launch",
    5: "32 C/A MEO transmitters x 256 receivers, one L1 band each at 2.048 MHz, seeded clocks, kTB noise on; "
       "the first {chunks} Mi-sample chunks of every band ({secs:.3f} s of the 60 s scene) per step",
}
SEED = 2026


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Threads for the CPU leg: every CPU this process may run on, capped at
    the GPU host's CPU share for one GPU (OMP_NUM_THREADS there, 16)."""
    visible = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return min(visible, share), visible


class Scene:
    """A prepared configuration uploaded to a fresh context, with device
    output buffers and its 1 Mi-sample chunk jobs."""

    def __init__(self, n, chunk=1 << 20, max_chunks=None, noise=True):
        from lipika_amd import _lib, workloads
        from lipika_amd.pipeline import Backend, prepare
        sc = (workloads.micro() if n == "micro" else workloads.scene3_orbiting() if n == "3 orbiting"
              else workloads.scene(3, 10.0) if n == "3 10 s" else workloads.scene3_wide() if n == "3 wide"
              else workloads.scene(n))
        if not noise:
            for rx in sc.receivers:
                for b in rx.bands:
                    b.noise_power_w = 0.0
        self.n, self.sc = n, prepare(sc)
        self.be = Backend(0)
        self.lib = _lib.load()
        self.be.upload(self.sc, SEED)
        self.units, self.bufs, self.jobs, self.keep = [], [], [], []
        self.contributions = self.out_samples = 0
        w = 0
        for r, rx in enumerate(self.sc.receivers):
            for b, pb in enumerate(rx.bands):
                n_s = pb.n if max_chunks is None else min(pb.n, max_chunks * chunk)
                dev = C.c_void_p()
                self.be.check(self.lib.lpk_device_alloc(self.be.h, max(n_s, 1) * 8, C.byref(dev)))
                self.bufs.append(dev)
                for i0 in range(0, n_s, chunk):
                    j, k = self.be.job(r, b, w, SEED, i0, min(chunk, n_s - i0), out_dev=dev.value + i0 * 8)
                    self.jobs.append(j)
                    self.keep.append(k)
                self.units.append((w, r, b, pb, n_s))
                self.contributions += n_s * len(pb.active)
                self.out_samples += n_s
                w += 1

    def close(self):
        for d in self.bufs:
            self.be.check(self.lib.lpk_device_free(self.be.h, d))
        self.be.close()


FIXED_TUNING = None   # (tile_shift, xcd_map) from --tuning, else autotuned per form


def time_form(S, name, steps, warmup):
    """Autotunes the form on chunk 0 of every band, then times `steps` whole
    renders (clock walks regenerated from index 0 each step). Returns wall ms
    per step, render-kernel ms per render call (HIP events on the library's
    stream around every render launch) and the tuning."""
    from lipika_amd.pipeline import form_code
    be, lib = S.be, S.lib
    form = form_code(name)
    first = [j for j in S.jobs if j.i0 == 0]
    if FIXED_TUNING:
        be.set_tuning(*FIXED_TUNING)
        tuning, tune_ms = FIXED_TUNING, 0.0
    else:
        tuning, tune_ms = be.autotune(first, form)

    def step():
        be.check(lib.lpk_clock_rewind(be.h))
        be.render(S.jobs, form)

    for _ in range(warmup):
        step()
    be.sync()
    be.check(lib.lpk_timing_reset(be.h))
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    be.sync()
    t1 = time.perf_counter()
    kern_ms, n_launch = C.c_double(), C.c_uint64()
    be.check(lib.lpk_timing_read(be.h, C.byref(kern_ms), C.byref(n_launch)))
    ms = (t1 - t0) * 1e3 / steps
    kms = kern_ms.value / max(n_launch.value, 1)
    log(f"[bench] config {S.n} form {name}: tile 2^{tuning[0]}, xcd map {tuning[1]} (autotune {tune_ms:.2f} ms); "
        f"{steps} steps after {warmup} warm-up: {ms:.3f} ms/step, render kernels {kms:.3f} ms per call")
    return {"value": round(S.contributions / (ms / 1e3) / 1e6, 3), "ms_per_step": round(ms, 4),
            "kernel_ms": round(kms, 4), "tile_shift": tuning[0], "xcd_map": tuning[1]}


def plan_leg(S, form, steps, warmup, chunk):
    """The batching form: a streaming plan renders chunk k of every band per
    step (steps alternate between two streams), with the context's
    (autotuned) tuning; a pass is every step of the scene."""
    from lipika_amd.pipeline import form_code
    be, lib = S.be, S.lib
    tuning = be.tuning()
    bands, keep = [], []
    for (w, r, b, pb, n_s), buf in zip(S.units, S.bufs):
        j, k = be.job(r, b, w, SEED, 0, n_s, out_dev=buf.value)
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
        ms = (t1 - t0) * 1e3 / steps
        log(f"[bench] config {S.n} plan ({form}, {chunk}-sample chunks): {plan.n_chunks} steps per pass: "
            f"{ms:.3f} ms per pass")
        return {"form": form, "tile_shift": tuning[0], "xcd_map": tuning[1], "chunk_samples": chunk,
                "plan_steps_per_pass": plan.n_chunks, "max_workgroups_per_step": plan.max_tiles,
                "value": round(S.contributions / (ms / 1e3) / 1e6, 3),
                "ms_per_step": round(ms, 4)}


def dropin_leg(steps):
    """The drop-in as the Rust side would call it (generate_band's chunk loop
    with sink.push, src/pipeline.rs:905-990, :984), on configuration 2 whole:
    (a) one lpk_render_chunk per (band, 1 Mi-sample chunk) into one of two
    pinned host buffers, lpk_sync, then the FNV-1a sink on that chunk while
    the next one renders; (b) lpk_stream_to_sink, which renders chunk k of
    every band while the sink takes chunk k - 1, with the same FNV-1a sink;
    (c) the same with a sink that does nothing (device render + PCIe only)."""
    from lipika_amd import _lib
    S = Scene(2)
    be, lib = S.be, S.lib
    chunk = 1 << 20
    form = _lib.FORM_PRECISION
    slot_bytes = chunk * 8
    pin = []
    for _ in range(2):
        p = C.c_void_p()
        be.check(lib.lpk_host_alloc(be.h, slot_bytes, C.byref(p)))
        pin.append(p)
    order = []   # (w, r, b, i0, clen) in the reference's order: band after band, chunk after chunk
    for (w, r, b, pb, n_s) in S.units:
        for i0 in range(0, n_s, chunk):
            order.append((w, r, b, i0, min(chunk, n_s - i0)))

    def per_chunk_pass():
        hashes = [_lib.FNV_OFFSET] * len(S.units)
        prev = None
        keep = []
        for k, (w, r, b, i0, clen) in enumerate(order):
            j, kk = be.job(r, b, w, SEED, i0, clen)
            j.out_host = pin[k & 1].value
            keep.append(kk)
            be.render([j], form)
            if prev is not None:   # the sink takes the previous chunk while this one renders
                pw, pbuf, pn = prev
                hashes[pw] = lib.lpk_fnv1a64(pbuf, pn * 8, hashes[pw])
            be.sync()
            prev = (w, pin[k & 1].value, clen)
        if prev is not None:
            hashes[prev[0]] = lib.lpk_fnv1a64(prev[1], prev[2] * 8, hashes[prev[0]])
        return hashes

    whole, keep = [], []
    for (w, r, b, pb, n_s) in S.units:
        j, k = be.job(r, b, w, SEED, 0, n_s)
        whole.append(j)
        keep.append(k)

    def timed(fn):
        be.check(lib.lpk_clock_rewind(be.h))
        fn()
        t = time.perf_counter()
        for _ in range(steps):
            be.check(lib.lpk_clock_rewind(be.h))
            out = fn()
        return (time.perf_counter() - t) * 1e3 / steps, out

    ms_a, ha = timed(per_chunk_pass)
    ms_b, hb = timed(lambda: be.stream_to_sink(whole, chunk, form))
    ms_c, _ = timed(lambda: be.stream_to_sink(whole, chunk, form, lambda band, i0, iq: False))
    for p in pin:
        be.check(lib.lpk_host_free(be.h, p))
    res = {"workload": "config 2: " + WORKLOADS[2] + "; precision form; 1 Mi-sample chunks; output to pinned host "
                       "memory, FNV-1a sink (examples/sky_throughput.rs:73-82)",
           "output_bytes_per_step": S.out_samples * 8, "steps": steps,
           "per_chunk_calls": {"calls_per_step": len(order), "ms_per_step": round(ms_a, 3),
                               "value": round(S.contributions / (ms_a / 1e3) / 1e6, 3)},
           "stream_to_sink_fnv": {"ms_per_step": round(ms_b, 3),
                                  "value": round(S.contributions / (ms_b / 1e3) / 1e6, 3)},
           "stream_to_sink_null": {"ms_per_step": round(ms_c, 3),
                                   "value": round(S.contributions / (ms_c / 1e3) / 1e6, 3),
                                   "host_gb_s": round(S.out_samples * 8 / (ms_c / 1e3) / 1e9, 2)},
           "hashes_agree": ha == hb}
    log(f"[bench] drop-in (config 2 to host): per-chunk calls {ms_a:.1f} ms, stream+FNV {ms_b:.1f} ms, "
        f"stream+null sink {ms_c:.1f} ms per pass; hashes agree: {ha == hb}")
    S.close()
    return res


def learned_leg(steps, warmup, chunks=2):
    """Configuration 5: the learned warm start. The network is trained on the
    device (1e7 pairs from the device's converged light-time solve), then the
    first `chunks` chunks of every band are rendered in the plain form (the
    reference's two fixed iterations), the converged form (iterated to a
    1e-13 s update from the reference's start) and the learned form (the
    same, from the network's light times at 2 ms query points, converged
    there and interpolated). `straight_line_start` is the learned form with
    an all-zero network (the queries start from the straight-line light
    time): the network's own gain is the query evaluations it saves against
    that. Evaluation counts include the queries'. The whole 60 s scene is
    252 GB of output, more than one GPU's HBM beside its clock walks;
    per-contribution work does not depend on the time within the scene (MEO
    geometry, no horizon gate), so the full scene's time is the step's time
    scaled by `full_scene_factor`."""
    from lipika_amd.pipeline import WarmStartNet, form_code
    S = Scene(5, max_chunks=chunks)
    be, lib = S.be, S.lib
    net = WarmStartNet(be, init_seed=1)
    train = net.train(n_pairs=10_000_000, data_seed=SEED, epochs=16, batch=16384, lr=3e-3, lr_final=3e-5,
                      n_val=200_000)
    log(f"[bench] config 5: trained the warm-start network: {train}")
    be.warm_start(net)
    full = sum(pb.n * len(pb.active) for (_, _, _, pb, _) in S.units)
    forms = {}
    trained_params = net.params
    for name in ("plain", "converged", "learned", "straight_line_start"):
        if name == "straight_line_start":
            net.params = 0.0 * trained_params
            be.warm_start(net)
        be.solve_stats_reset()
        entry = time_form(S, "learned" if name == "straight_line_start" else name, steps, warmup)
        if name == "plain":
            entry["evaluations_per_contribution"] = 2.0    # the reference's fixed two iterations
        else:
            st = be.solve_stats()
            entry["evaluations_per_contribution"] = round(st["evaluations"] / max(st["contributions"], 1), 6)
            entry["fallbacks"] = st["fallbacks"]
            entry["unconverged"] = st["unconverged"]
        forms[name] = entry
    be.warm_start(None)
    net.close()
    secs = chunks * (1 << 20) / 2.048e6
    lf, cf, pf, sf = forms["learned"], forms["converged"], forms["plain"], forms["straight_line_start"]
    factor = full / S.contributions
    out = {"workload": "config 5: " + WORKLOADS[5].format(chunks=chunks, secs=secs),
           "contributions_per_step": S.contributions, "full_scene_contributions": full,
           "full_scene_factor": round(factor, 4),
           "full_scene_s": {k: round(v["ms_per_step"] * factor / 1e3, 3) for k, v in forms.items()},
           "steps": steps, "forms": forms,
           "evaluations_saved_per_contribution": round(cf["evaluations_per_contribution"]
                                                       - lf["evaluations_per_contribution"], 4),
           "learned_speedup_vs_converged": round(cf["ms_per_step"] / lf["ms_per_step"], 3),
           "learned_speedup_vs_plain": round(pf["ms_per_step"] / lf["ms_per_step"], 3),
           "network_evaluations_saved_per_contribution": round(sf["evaluations_per_contribution"]
                                                               - lf["evaluations_per_contribution"], 6),
           "training": {"pairs": 10_000_000, "epochs": 16, "batch": 16384, "steps": train["steps"],
                        "train_ms": round(train["train_ms"], 1), "pair_generation_ms": round(train["gen_ms"], 2),
                        "validation_rms_s": train["val_rms_s"], "validation_max_s": train["val_max_s"]}}
    S.close()
    return out


def cpu_leg(S, target_s):
    """The oracle (the project's C restatement of the reference CPU path,
    with its two-level parallelism: bands in parallel, transmitters of a
    chunk in parallel, fixed-order sum; FNV-1a sink) on a bounded prefix of
    the same scene, on the host's threads."""
    from oracle.pyoracle import OracleScene
    threads, visible = host_threads()
    orc = OracleScene(S.sc, SEED)
    dur = S.sc.meta.duration_s

    def run(frac):
        nl = [max(1, int(pb.n * frac)) for _, pb in orc.units]
        t = time.perf_counter()
        contrib, _ = orc.render_scene_hashed(nl, 1 << 20, threads)
        return contrib, time.perf_counter() - t

    c0, t0 = run(0.002)
    frac = min(1.0, 0.002 * target_s / max(t0, 1e-3))
    c, t = run(frac)
    log(f"[bench] config {S.n} cpu baseline: {c} contributions ({frac * dur:.3f} s of {dur} s emulated) in "
        f"{t:.2f} s on {threads} threads ({visible} visible)")
    return {"value": c / t / 1e6, "unit": "Msamples/s", "cores": threads, "host_cpus_visible": visible,
            "kind": "port",
            "sample": f"first {frac * dur:.3f} s of every band of the same scene ({c} contributions, {t:.1f} s "
                      f"wall; the rate does not depend on the time within the scene); the oracle restatement of "
                      f"the reference CPU path, bands and transmitters in parallel, FNV-1a sink; threads = the "
                      f"host CPUs this process may use, capped at the GPU host's per-GPU share (OMP_NUM_THREADS)"}


def oracle_errors(S, units, n):
    """Errors of the precision form against the FP64 oracle on windows of n
    samples at the start and at the end of each given unit, with the
    device's clock offsets for the window on both sides (orc_render_window):
    every link's carrier phase (the reference's f64 expression from the solved
    range, src/channel/mod.rs:256-258) and delay, and the noise-free IQ
    against the band's peak. The plain form must be bit-identical."""
    from lipika_amd import _lib
    from oracle.pyoracle import OracleScene
    import numpy as np
    be = S.be
    orc = OracleScene(S.sc, SEED)
    worst = {"phase_rad": 0.0, "delay_s": 0.0, "iq_rel_peak": 0.0}
    plain_same, windows = 1.0, 0
    for (w, r, b, pb, _) in units:
        for i0 in (0, pb.n - n):
            j, keep = be.job(r, b, w, SEED, i0, n)
            dt = be.taps(j, 0, _lib.FORM_PLAIN)[3]
            o, _, (ote, org, oph) = orc.render_window(w, i0, dt, taps=True)
            peak = float(np.max(np.abs(o)))
            for form in (_lib.FORM_PLAIN, _lib.FORM_PRECISION):
                g = np.empty(n, dtype=np.complex64)
                jj, kk = be.job(r, b, w, SEED, i0, n, out_host=g)
                be.render([jj], form)
                be.sync()
                if form == _lib.FORM_PLAIN:
                    same = float(np.mean((g.view(np.uint32) == o.view(np.uint32)).reshape(-1, 2).all(axis=1)))
                    plain_same = min(plain_same, same)
                else:
                    d = float(np.max(np.abs(g.astype(np.complex128) - o.astype(np.complex128))))
                    worst["iq_rel_peak"] = max(worst["iq_rel_peak"], d / peak)
            for slot in range(len(pb.active)):
                te, rg, ph, _ = be.taps(j, slot, _lib.FORM_PRECISION)
                worst["phase_rad"] = max(worst["phase_rad"], float(np.max(np.abs(ph - oph[slot]))))
                worst["delay_s"] = max(worst["delay_s"], float(np.max(np.abs(rg - org[slot]))) / C_LIGHT)
            windows += 1
    return {"precision": {k: float(f"{v:.4g}") for k, v in worst.items()},
            "plain_bit_identical_fraction": plain_same, "windows": windows, "samples_per_window": n,
            "method": "device clock offsets on both sides; noise-free; every link tapped"}


def micro_leg(steps, warmup, cpu):
    """The reference's criterion micro-benchmark chunk
    (benches/iq_generation.rs:15-60): one C/A SV at GPS altitude, 5 s at
    1.5 MHz, one transmitter_contribution_at_receiver per step. Each form is
    timed as a whole-chunk render call; the oracle renders the same chunk on
    one thread, as criterion runs it."""
    S = Scene("micro")
    out = {"workload": "the reference's criterion micro-benchmark chunk: C/A PRN 1 from a circular orbit at GPS "
                       "altitude, seen from Glasgow, 5 s at 1.5 MHz, no noise, zero clock offset "
                       "(benches/iq_generation.rs:15-60)",
           "contributions_per_step": S.contributions,
           "forms": {f: time_form(S, f, steps, warmup) for f in ("plain", "precision")}}
    if cpu:
        from oracle.pyoracle import OracleScene
        orc = OracleScene(S.sc, SEED)
        t = time.perf_counter()
        c, _ = orc.render_scene_hashed([pb.n for _, pb in orc.units], 1 << 30, 1)
        dt = time.perf_counter() - t
        out["cpu_one_thread"] = {"value": round(c / dt / 1e6, 3), "unit": "Msamples/s", "seconds": round(dt, 3),
                                 "kind": "port"}
        out["plain_vs_cpu_one_thread"] = round(out["forms"]["plain"]["value"] / out["cpu_one_thread"]["value"], 1)
    log(f"[bench] micro-benchmark chunk: {json.dumps(out)}")
    S.close()
    return out


def moving_leg(steps, warmup):
    """Moving receivers: configuration 3's sky and bands for 10 s with the 16
    receivers fixed (as configuration 3) and with each on a LEO orbit, both
    forms, every chunk in one render call."""
    out = {"workload": "configuration 3's 32 C/A MEOs and L1 bands at 2.048 MHz, 10 s, 16 receivers fixed and "
                       "on circular LEO orbits (6 900-7 200 km)"}
    for key, n in (("fixed", "3 10 s"), ("orbiting", "3 orbiting")):
        S = Scene(n)
        out[key] = {"contributions_per_step": S.contributions,
                    "forms": {f: time_form(S, f, steps, warmup) for f in ("plain", "precision")}}
        S.close()
    out["precision_vs_plain_orbiting"] = round(out["orbiting"]["forms"]["precision"]["value"] /
                                               out["orbiting"]["forms"]["plain"]["value"], 3)
    log(f"[bench] moving receivers: {json.dumps(out)}")
    return out


def passes_leg(steps, warmup, moving):
    """The precision form in two passes per band: configuration 3's receivers
    and bands for 10 s under 64 C/A MEOs (workloads.scene3_wide), every chunk
    in one render call, beside the same bands under the 32 of configuration 3
    (one pass, the moving leg's fixed receivers): the passes' cost per
    contribution."""
    S = Scene("3 wide")
    r = time_form(S, "precision", steps, warmup)
    one = moving["fixed"]["forms"]["precision"]
    out = {"workload": "configuration 3's 16 receivers and L1 bands at 2.048 MHz, 10 s, under 64 C/A MEOs "
                       "(two passes of 32 per band)",
           "contributions_per_step": S.contributions, "precision": r,
           "one_pass_value": one["value"],
           "per_contribution_vs_one_pass": round(r["value"] / one["value"], 4)}
    S.close()
    log(f"[bench] link passes: {json.dumps(out)}")
    return out


def roofline(contributions, kernel_ms, n):
    achieved = contributions * FLOPS_PER_CONTRIBUTION / (kernel_ms / 1e3) / 1e12
    traffic, source = TRAFFIC.get(n, (None, None))
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic, "traffic_source": source,
            "compute_unit": "FP64 vector ALU: the path issues no matrix op; achieved = contributions x 75 FP64 "
                            "flops of the reference's arithmetic / render-kernel time",
            "kernel": "render_kernel<1, false, TS> (precision form)", "kernel_ms": round(kernel_ms, 4),
            "flops_per_contribution": FLOPS_PER_CONTRIBUTION}


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
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and the oracle error checks")
    ap.add_argument("--tuning", type=str, default="",
                    help="tile_shift,xcd_map for every form instead of the autotuner's pick")
    ap.add_argument("--headline-only", action="store_true",
                    help="configuration 4's precision form only (for profiler runs)")
    ap.add_argument("--plan-only", action="store_true",
                    help="configuration 4's precision form tuned, then --steps plan passes in 1 Mi-sample "
                         "chunks (for profiler runs)")
    args = ap.parse_args()
    if args.tuning:
        global FIXED_TUNING
        FIXED_TUNING = tuple(int(v) for v in args.tuning.split(","))
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the path shards by receiver with no collective; see the design notes).", file=sys.stderr)
        sys.exit(2)
    K, W = args.steps, args.warmup
    few = max(2, K // 5)

    S4 = Scene(4)
    log(f"[bench] config 4: {len(S4.jobs)} chunk jobs, {S4.out_samples} output samples, "
        f"{S4.contributions} contributions per step")
    if args.plan_only:
        time_form(S4, "precision", 1, 0)
        print(json.dumps(plan_leg(S4, "precision", K, W, 1 << 20)), flush=True)
        S4.close()
        return
    prec = time_form(S4, "precision", K, W)
    forms4 = {"precision": prec}
    res = {"metric": METRIC, "value": prec["value"], "unit": "Msamples/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": prec["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "config 4: " + WORKLOADS[4], "form": "precision", "chunk_samples": 1 << 20,
                      "contributions_per_step": S4.contributions, "output_samples_per_step": S4.out_samples,
                      "output_msamples_per_s": round(S4.out_samples / (prec["ms_per_step"] / 1e3) / 1e6, 3),
                      "forms": forms4},
           "roofline": roofline(S4.contributions, prec["kernel_ms"], 4), "cpu_baseline": None}
    if args.headline_only:
        S4.close()
        print(json.dumps(res), flush=True)
        return
    forms4["plain"] = time_form(S4, "plain", few, 1)
    prec["speedup_vs_plain"] = round(forms4["plain"]["ms_per_step"] / prec["ms_per_step"], 3)
    if FIXED_TUNING:
        S4.be.set_tuning(*FIXED_TUNING)
    else:
        S4.be.autotune([j for j in S4.jobs if j.i0 == 0], 1)
    plans = {c: plan_leg(S4, "precision", few, 1, c) for c in (1 << 20, 1 << 24)}
    for p in plans.values():
        p["speedup_vs_plain"] = round(forms4["plain"]["ms_per_step"] / p["ms_per_step"], 3)
    res["config"]["batching"] = plans[1 << 20]
    res["config"]["batching_16mi_chunks"] = plans[1 << 24]
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_leg(S4, args.cpu_seconds)
        res["config"]["precision_vs_cpu"] = round(prec["value"] / res["cpu_baseline"]["value"], 1)
    S4.close()
    if not args.no_cpu:
        # errors of the headline form on configuration 4 (noise-free, the same geometry and clocks)
        S4n = Scene(4, max_chunks=0, noise=False)
        best = {}
        for u in S4n.units:
            k = (u[3].band.sample_rate_hz, u[3].band.lo_carrier_hz)
            if k not in best or len(u[3].active) > len(best[k][3].active):
                best[k] = u
        res["config"]["errors_vs_oracle"] = oracle_errors(S4n, [best[k] for k in sorted(best)], 8192)
        log(f"[bench] config 4 errors vs oracle: {res['config']['errors_vs_oracle']}")
        S4n.close()

    S3 = Scene(3)
    c3 = {"workload": "config 3: " + WORKLOADS[3], "contributions_per_step": S3.contributions,
          "output_samples_per_step": S3.out_samples,
          "forms": {"precision": time_form(S3, "precision", few, 1), "plain": time_form(S3, "plain", few, 1)}}
    c3["forms"]["precision"]["speedup_vs_plain"] = round(c3["forms"]["plain"]["ms_per_step"]
                                                         / c3["forms"]["precision"]["ms_per_step"], 3)
    c3["roofline_frac"] = roofline(S3.contributions, c3["forms"]["precision"]["kernel_ms"], 3)["frac"]
    if not args.no_cpu:
        c3["cpu_baseline"] = cpu_leg(S3, args.cpu_seconds / 3)
    S3.close()
    if not args.no_cpu:
        S3n = Scene(3, max_chunks=0, noise=False)
        c3["errors_vs_oracle"] = oracle_errors(S3n, S3n.units, 4096)
        log(f"[bench] config 3 errors vs oracle: {c3['errors_vs_oracle']}")
        S3n.close()
    res["config"]["config3"] = c3

    S2 = Scene(2)
    c2 = {"workload": "config 2: " + WORKLOADS[2], "contributions_per_step": S2.contributions,
          "output_samples_per_step": S2.out_samples,
          "forms": {"precision": time_form(S2, "precision", K, W), "plain": time_form(S2, "plain", K, W)}}
    c2["forms"]["precision"]["speedup_vs_plain"] = round(c2["forms"]["plain"]["ms_per_step"]
                                                         / c2["forms"]["precision"]["ms_per_step"], 3)
    S2.close()
    res["config"]["config2"] = c2
    res["config"]["dropin"] = dropin_leg(2)
    res["config"]["micro"] = micro_leg(K, W, not args.no_cpu)
    res["config"]["moving"] = moving_leg(K, W)
    res["config"]["link_passes"] = passes_leg(K, W, res["config"]["moving"])
    res["config"]["config5"] = learned_leg(max(1, K // 10), 1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
