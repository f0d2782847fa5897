"""Writes tests/golden/reference_kats.json: the known-answer values the
reference's own tests hold for the IQ path, transcribed from its test sources
(read as text; nothing of the reference is run or imported), plus two public
tables those tests lean on. Each entry names where its values come from.

Run: python tests/golden/make_reference_kats.py
"""
import json
import os

KATS = {
    "geodetic_glasgow": {
        "source": "src/geometry/ecef.rs:27-36",
        "input": [55.86, -4.25, 10.0],
        "ecef": [3577894.5228608, -265883.49155448, 5255718.21468305],
        "max_relative": 1e-9,
    },
    "geodetic_equator": {
        "source": "src/geometry/ecef.rs:38-45",
        "input": [0.0, 0.0, 0.0],
        "ecef": [6378137.0, 0.0, 0.0],
    },
    "kepler_gps_radius": {
        "source": "src/geometry/kepler.rs:110-124",
        "orbit": [26578137.0, 55.0, 0.0, 0.0, 0.0],
        "t_unix": 0.0,
        "max_relative": 1e-9,
    },
    "grid_vs_orbit": {
        "source": "src/geometry/precomputed.rs:204-225",
        "orbit": [26578137.0, 55.0, 0.0, 30.0, 1700000000.0],
        "t0": 1700000000.0, "duration_s": 60.0, "dt": 0.01,
        "max_relative": 1e-6, "epsilon_m": 1e-3,
    },
    "grid_path_bounds": {
        "source": "src/geometry/precomputed.rs:162-192",
        "orbit": [6928137.0, 53.0, 10.0, 0.0, 1700000000.0],
        "t0": 1700000000.0, "duration_s": 30.0, "dt": 0.01,
        "speed_range": [6900.0, 7700.0],
    },
    "sagnac": {
        "source": "src/channel/mod.rs:353-383",
        "sat": [25.0e6, 10.0e6, 0.0], "obs": [6378137.0, 0.0, 0.0],
        "epoch": 1717260840.0, "iterations": 4,
        "min_vs_naive_m": 1.0, "max_vs_exact_m": 0.01,
    },
    "stationary_pair_delay": {
        "source": "src/pipeline.rs:1143-1181",
        "tx": [7000000.0, 0.0, 0.0], "rx": [7000000.0, 0.0, 300000.0],
        "epoch": 1717260840.0, "iterations": 4, "delay_tolerance_s": 1e-9,
    },
    "elapsed_periods": {
        "source": "src/signal/mod.rs:127-150",
        "exact": [[1717260840.0, 4800.0, 8242852032000, 0.0], [-0.25, 2.0, -1, 0.5]],
        "fraction": [[1717260840.0, 4800.0, 8242852032000, 0.00457763671875]],
        "fraction_elapsed_offset": 9.5367431640625e-07,
        "fraction_tolerance": 1e-15,
        "consistency": {"rate": 1371.4285714285713, "e0": 1.0e9, "e1": 1.0e9 + 7.0, "whole_step": 9600,
                        "frac_tolerance": 1e-9},
        "errors": [["inf", 1.0], [1.0e12, 1.0e12]],
    },
    "frame_anchor_bpsk": {
        "source": "src/signal/mod.rs:170-198",
        "render_epoch": 1.0e9, "signal_epoch": 999999999.3575, "rate": 1000.0, "n_bits": 64,
        "repeat_from": 2, "epoch_rel_s": -0.0005, "epoch_rel_tolerance": 1.2e-7,
    },
    "cw_anchor_zero": {
        "source": "src/signal/mod.rs:221-223",
        "epoch": 1000000007.25, "offset_hz": 1234.5, "phase": 0.0,
    },
    "ca_prn1_first10": {
        "source": "src/signal/ca_code.rs:200-208",
        "chips": [-1, -1, 1, 1, -1, 1, 1, 1, 1, 1],
    },
    "ca_first10_octal": {
        "source": "IS-GPS-200 Table 3-Ia, 'First 10 Chips C/A' (octal, leading chip first); public spec",
        "octal": [1440, 1620, 1710, 1744, 1133, 1455, 1131, 1454, 1626, 1504, 1642, 1750, 1764, 1772,
                  1775, 1776, 1156, 1467, 1633, 1715, 1746, 1763, 1063, 1706, 1743, 1761, 1770, 1774,
                  1127, 1453, 1625, 1712],
    },
    "nav_preamble": {
        "source": "src/signal/ca_code.rs:233-246",
        "seed": 12345, "pm": [-1, 1, 1, 1, -1, 1, -1, -1], "subframe_bits": 300,
    },
    "nav_payload_flips": {"source": "src/signal/ca_code.rs:263-272", "seed": 999},
    "qam16": {
        "source": "src/signal/qam16.rs:253-287",
        "levels": [-3.0, -1.0, 1.0, 3.0], "scale_sqrt": 10.0, "mean_power": 1.0, "tolerance": 1e-6,
    },
    "qam16_payload": {
        "source": "src/signal/qam16.rs:289-325",
        "sync": [0, 0, 0, 0, 1, 0, 1, 0], "n_payload": 4000, "seed": 11, "rate": 2000.0,
        "times": [-1e-9, 0.0, 0.499e-3, 0.5e-3, 0.999e-3, 2001.0, 2001.5],
        "expect_level": [None, -3, -3, 3, 3, None, None], "power_tolerance": 0.05,
    },
    "ook_frame": {
        "source": "src/signal/ook.rs:87-119",
        "sync": [1, 0, 1, 1, 0, 0, 1, 0], "n_payload": 64, "seed": 3, "rate": 1000.0,
        "ones_range": [16, 48],
        "silent": {"sync": [1, 1], "n_payload": 6, "seed": 9,
                   "times": [-1e-6, 0.0, 7.999e-3, 8.0e-3, 1.0], "expect": ["0", "1", "bit7", "0", "0"]},
    },
    "cw_tone": {
        "source": "src/signal/cw.rs:60-75",
        "offset_hz": 1000.0, "times": [0.0, 0.25e-3, 0.5e-3, 0.7e-3, 1.0e-3], "epsilon": 1e-5,
    },
    "gate_truth_table": {
        "source": "src/signal/gate.rs:116-135",
        "period": 15.0, "bursts": [[0.0, 0.5], [1.5, 0.5], [3.0, 0.5]],
        "cases": [[0.0, 1], [0.25, 1], [0.499999, 1], [0.5, 0], [1.0, 0], [1.499999, 0], [1.5, 1],
                  [1.999999, 1], [2.0, 0], [3.0, 1], [3.25, 1], [3.5, 0], [10.0, 0], [14.999999, 0]],
    },
    "gate_absolute": {
        "source": "src/signal/gate.rs:137-157",
        "epoch": 1717260840.0, "render_offset": 4.0,
        "cases": [[0.0, 0], [10.9, 0], [11.0, 1], [11.4, 1], [11.5, 0], [12.5, 1]],
    },
    "ktb": {
        "source": "src/channel/mod.rs:340-350",
        "temperature_k": 290.0, "bandwidth_hz": 2.048e6, "dbm": -110.9, "dbm_tolerance": 0.1,
    },
    "adc": {
        "source": "src/rx/mod.rs:169-193",
        "bits": 4, "full_scale_v": 1.0, "input": [[0.5, -0.5], [2.0, -3.0]], "clipped": [1.0, -1.0],
    },
    "gain": {"source": "src/rx/mod.rs:211-223", "db": 20.0, "input": [0.1, -0.2], "output": [1.0, -2.0],
             "tolerance": 1e-5},
    "ip3": {"source": "src/rx/mod.rs:225-245", "iip3_dbm": 0.0, "small": 1e-4, "large": 0.02,
            "tolerance": 1e-6},
    "clock_drift": {
        "source": "src/clock/mod.rs:375-392",
        "y0": 5e-6, "fs": 1e6, "duration_s": 1.0, "expected_last": 4.999995e-06, "max_relative": 1e-12,
    },
    "clock_allan": {"source": "src/clock/mod.rs:394-422", "adev": 1e-11, "tolerance": 0.4},
    "clock_two_band": {
        "source": "src/clock/mod.rs:470-509",
        "adev": 1e-11, "y0": 0.3e-6, "master": 2.048e6, "low": 48000.0, "n_low": 500,
        "max_relative": 1e-9, "epsilon": 1e-18,
    },
    "clock_offset_recovery": {
        "source": "tests/multi_transmitter_summation.rs:101-156",
        "offset_s": 1.3e-7, "tolerance_s": 1e-9, "carrier_hz": 1.0e6, "fs": 100000.0, "duration_s": 0.02,
    },
    "additivity": {"source": "tests/multi_transmitter_summation.rs:188-236", "max_abs_err": 1e-10},
    "noise_power": {"source": "tests/multi_transmitter_summation.rs:175-186; tests/window_render.rs:431-458",
                    "power_w": 1e-15, "tolerance": 0.02, "mean_sigma": 0.02, "lag1": 0.02},
    "bands_overlap": {
        "source": "src/pipeline.rs:1103-1141",
        "cases": [[437000000.0, 19200.0, 1575420000.0, 2048000.0, 0],
                  [1575420000.0, 2046000.0, 1575420000.0, 2048000.0, 1],
                  [1576444000.0, 0.0, 1575420000.0, 2048000.0, 1],
                  [1576444001.0, 0.0, 1575420000.0, 2048000.0, 0],
                  [1580420000.0, 20000000.0, 1575420000.0, 2048000.0, 1]],
    },
    "antenna_pattern": {
        "source": "src/rx/mod.rs:196-209",
        "antenna": {"peak_dbi": 3.9, "horizon_dbi": -3.0, "fb_db": 8.5},
        "cases_dbi": [[90.0, 3.9], [0.0, -3.0], [-10.0, -11.5]],
        "tolerance_db": 1e-3,
        "monotone": [0.0, 45.0, 90.0],
        "cutoff_deg": 0.0, "occluded_at": -0.1, "present_at": 0.0, "unchanged_at": 45.0,
    },
    "antenna_occlusion": {
        "source": "src/channel/mod.rs:384-420",
        "obs": [6378137.0, 0.0, 0.0], "above": [6.9e6, 0.0, 0.0], "below": [-7.0e6, 0.0, 0.0],
        "carrier_hz": 1575420000.0, "prn": 1, "tx_power_w": 20.0, "epoch": 1717260840.0,
        "unit_antenna": {"peak_dbi": 0.0, "horizon_dbi": 0.0, "fb_db": 0.0},
        "n": 8,
    },
    "xoshiro256pp_reference": {
        "source": "xoshiro256++ reference vector (Blackman & Vigna), state {1,2,3,4}",
        "state": [1, 2, 3, 4],
        "outputs": [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
                    9973669472204895162, 14011001112246962877, 12406186145184390807,
                    15849039046786891736, 10450023813501588000],
    },
}


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(path, "w") as f:
        json.dump(KATS, f, indent=1)
    print(path)
