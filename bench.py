#!/usr/bin/env python3
"""Benchmark: 2-D Euler cell-updates/s of WENO5-SR vs WENO5-JS at 400^2 (one MI355X).

Workload: the four-quadrant Riemann problem C.4 at 400x400 (15x15 Gauss-Legendre
cell averages, transmissive boundaries), characteristic WENO5 + HLLC + SSPRK3,
CFL 0.4, FP64 state; closures WENO5-SR with the FP32 MLP (headline), WENO5-JS
(classical baseline), WENO5-SR FP64 / BF16.  One "step" is one SSPRK3 step over
the 400x400 grid (3 stages); one cell-update is one interior cell advanced by one
step.  `--steps N` times N steps of the developed flow from t = 0.125 (reached
untimed from the initial data; early steps from the piecewise-constant data run
the closure on few stencils and are not representative); `--steps 0` (default)
times the whole trajectory to t = 0.25, the protocol of the reference's timing
harness (for_paper_results/run_weno5_timing.py:111-140).  Either way the line also
carries the full-trajectory rates of both closures (`full_trajectory`).

Prints ONE JSON line (rank 0).  Extra keys: `js` (WENO5-JS throughput on the same
trajectory), `sr_vs_js_wall`, `accuracy` (L1 density error of each 400^2 run
against an 800^2 FP64 WENO5-JS solution block-averaged to 400^2), `roofline`
(sweep kernels, HIP-event timed inside the timed region) and `cpu_baseline`
(the CPU oracle on the host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "2D Euler cell-updates/s, WENO5-SR vs WENO-JS at 400^2, same L1 error"
T_END = 0.25
T_MID = 0.125          # start of the timed window with --steps N (the developed flow)
NX = NY = 400
# MI355X peaks: FP64 vector 78.6 TFLOP/s; FP32 157.3 TFLOP/s (vector and matrix core: a
# non-packed v_fma_f32 issues in 2.85 cycles per wave64 against 5.14 for v_fma_f64, and a
# v_mfma_f32_4x4x1f32 moves 256 multiply-adds in 10.4 cycles, measured).  A kernel's
# roofline peak is the blend of its algorithmic op mix (blended_peak).
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_VALU_TFLOPS = PEAK_FP64_TFLOPS     # all-FP64 kernels (WENO-JS, the FP64 closure)
PEAK_HBM_GBS = 8000.0
REFERENCE_SR32_WALL_S = 14.253   # experiments/14_mixed_precision_timing/tables/weno5_precision_timing.tex:11
REFERENCE_JS_WALL_S = 6.605      # same table, line 9

# ---------------------------------------------------------------- flop model
# Operations per call of each building block of the WENO5 step, as tallied by the
# op-counting build of the oracle (a counting scalar over the unchanged oracle source:
# add/sub, mul, div, sqrt, exp, log and tanh count one each; comparisons, min/max, fabs
# and sign flips zero; tests/test_opcount.py checks each value against the tally):
OPS = {
    "roe": 51, "jacobian": 17,              # Roe average; its (u, v, a, H)
    "project": 53, "back_project": 30,      # one cell onto the 4 fields; one face back
    "candidates_normal": 15, "candidates_gauss": 14,
    "js_beta": 51, "js_combine": 22, "z_combine": 27,
    "raw_sensors": 42, "plateau_test": 1,   # SR: deltas and gamma_s; the plateau compare
    "features": 9, "mlp_raw": 415,          # SR: features beyond the sensors; one 5-10-6-6-3 pass
    "reflection_average": 6, "sr_weights": 21, "weighted_sum": 5,
    "hllc": 104,                            # subsonic (star-state) path
}
MLP_RAW = OPS["mlp_raw"]
# The kernel's multiplicities: what the algorithm needs once, counted once.  A normal
# reconstruction is one head of one stencil; a Gauss pair is one transverse stencil
# with both Gauss-point heads, sharing its smoothness indicators (JS, Z) or its
# plateau test, features and closure (SR); the reflected features are a permutation
# of the direct ones.
_SR_SHARED = OPS["raw_sensors"] + OPS["plateau_test"]
_SR_CLOSURE = OPS["features"] + 2 * OPS["mlp_raw"] + OPS["reflection_average"]
NORMAL = {
    "js": OPS["candidates_normal"] + OPS["js_beta"] + OPS["js_combine"],
    "z": OPS["candidates_normal"] + OPS["js_beta"] + OPS["z_combine"],
    "sr": (_SR_SHARED + OPS["candidates_normal"] + OPS["weighted_sum"],                       # plateau
           _SR_SHARED + OPS["candidates_normal"] + OPS["weighted_sum"] + _SR_CLOSURE + OPS["sr_weights"]),
}
GAUSS = {
    "js": OPS["js_beta"] + 2 * (OPS["candidates_gauss"] + OPS["js_combine"]),
    "z": OPS["js_beta"] + 2 * (OPS["candidates_gauss"] + OPS["z_combine"]),
    "sr": (_SR_SHARED + 2 * (OPS["candidates_gauss"] + OPS["weighted_sum"]),
           _SR_SHARED + 2 * (OPS["candidates_gauss"] + OPS["weighted_sum"]) + _SR_CLOSURE + 2 * OPS["sr_weights"]),
}
# per face and transverse line: one Roe average and its Jacobian values, the six cells
# of the two stencils projected, the two face states projected back
FACE_LINE_FIXED = OPS["roe"] + OPS["jacobian"] + 6 * OPS["project"] + 2 * OPS["back_project"]
# per face: the two Gauss-point HLLC fluxes and their dt/h scaling
FLUX_FACE_FIXED = 2 * OPS["hllc"] + 8
BYTES_PER_SWEEP_CELL = 32 + 32               # state tile read + flux write (FP64 x 4)
# HBM bytes per SR32 sweep launch at 400^2 (one launch = the x and the y tiles), from
# the two rocprofv3 counter passes (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000 after the
# copy-kernel calibration) recorded in profiles/pmc_traffic_k_sweep_xy_c4_400x400.md
SWEEP_TRAFFIC_BYTES_400 = 37.00e6
# the same for the SR32 3-D flux launch at 224x88x88 (mean of the x, y, z face
# passes; profiles/pmc_traffic_k3_shockbubble_224x88x88.md) and the SR32 L+L~ WENO7
# sweep launch at 1200x300 (x and y tiles; profiles/pmc_traffic_k7_sweep_xy_doublemach_1200x300.md)
FLUX3_TRAFFIC_BYTES_224 = 814.1e6
SWEEP7_TRAFFIC_BYTES_1200 = 139.9e6


def blended_peak(total_ops: float, fp32_ops: float) -> float:
    """Peak TFLOP/s of an op mix: the FP64 part at the FP64 rate, the FP32 part (the FP32
    closure's features and MLP passes) at the FP32 rate."""
    if total_ops <= 0.0:
        return PEAK_FP64_TFLOPS
    fp32 = min(max(fp32_ops, 0.0), total_ops)
    return total_ops / ((total_ops - fp32) / PEAK_FP64_TFLOPS + fp32 / PEAK_FP32_TFLOPS)


# FP32 operations of one closure evaluation of the FP32 closure (features beyond the FP64
# sensors, the direct and the reflected 5-10-6-6-3 pass); zero for the FP64 closure
CLOSURE_FP32 = {"sr32": OPS["features"] + 2 * OPS["mlp_raw"], "srbf16": OPS["features"] + 2 * OPS["mlp_raw"]}


def sweep_closures_per_launch(frac_normal: float, frac_gauss: float) -> float:
    """Closure evaluations of one k_sweep_xy launch (normal heads and Gauss pairs)."""
    total = 0.0
    for nfaces, ntrans in ((NX + 1, NY), (NY + 1, NX)):
        total += nfaces * (ntrans + 4) * 8 * frac_normal + nfaces * ntrans * 8 * frac_gauss
    return total


def sweep_flops_per_launch(scheme: str, frac_normal: float, frac_gauss: float) -> float:
    """One k_sweep_xy launch: the x tiles ((nx+1) faces x ny lines) and the y tiles."""
    if scheme.startswith("sr"):
        n = NORMAL["sr"][0] + frac_normal * (NORMAL["sr"][1] - NORMAL["sr"][0])
        g = GAUSS["sr"][0] + frac_gauss * (GAUSS["sr"][1] - GAUSS["sr"][0])
    else:
        n, g = NORMAL[scheme], GAUSS[scheme]
    total = 0.0
    for nfaces, ntrans in ((NX + 1, NY), (NY + 1, NX)):
        total += nfaces * (ntrans + 4) * (FACE_LINE_FIXED + 8 * n) + nfaces * ntrans * (FLUX_FACE_FIXED + 8 * g)
    return total


def copy_calibration(args, out: dict, torch, dev) -> None:
    """One device copy of known size after the run: the reference kernel for the
    hardware-counter byte calibration."""
    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = {"read": n * 8, "write": n * 8}


def cpu_model_name() -> str:
    name = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return name


# ------------------------------------------------------------- 3-D workload
# 224 x 88 x 88 Mach-3 shock striking a light-gas bubble (rho = 1, p = 1 ahead of
# the shock, bubble radius 0.25 ny cells, density ratio 0.138), unit cell size,
# HLLC, SSPRK3, CFL 0.4, 200 fixed steps.
NX3, NY3, NZ3, STEPS3 = 224, 88, 88, 200
# algorithmic operations of the 3-D passes (same counting rules as above)
ROE5, EIG5 = 45, 22               # Roe average of two 5-variable states; eigen-system scalars
ROW5, DOT5, BACK5 = 15, 9, 10     # one scaled left row; one 5-term projection; one back-projection column
ROUNDTRIP5 = 5 * (5 * (ROW5 + DOT5) + 5 * BACK5)   # R (L q) of the five stencil cells
HLLC5 = 125                        # five-variable HLLC flux
FLUX_FACE3 = 4 * HLLC5 + 4 * 5 + 3 * 5             # four Gauss-point fluxes, dt/h scaling, flag sum


def gauss_cost(scheme: str, frac: float) -> float:
    if scheme.startswith("sr"):
        return GAUSS["sr"][0] + frac * (GAUSS["sr"][1] - GAUSS["sr"][0])
    return GAUSS[scheme]


def normal_cost(scheme: str, frac: float) -> float:
    if scheme.startswith("sr"):
        return NORMAL["sr"][0] + frac * (NORMAL["sr"][1] - NORMAL["sr"][0])
    return NORMAL[scheme]


def flux3_flops_per_launch(scheme: str, frac_g: float) -> float:
    """k3_flux, averaged over the x / y / z launches: per face 20 Gauss-pair
    reconstructions (2 first-pass variants x 2 sides x 5 components) + 4 fluxes."""
    nf = ((NX3 + 1) * NY3 * NZ3 + NX3 * (NY3 + 1) * NZ3 + NX3 * NY3 * (NZ3 + 1)) / 3.0
    return nf * (20 * gauss_cost(scheme, frac_g) + FLUX_FACE3)


def trans13_flops_per_launch(scheme: str, frac_g: float) -> float:
    """k3_trans1: 10 Gauss-pair reconstructions (2 sides x 5 components) per position."""
    npos = ((NX3 + 2) * NY3 * (NZ3 + 4) + NX3 * (NY3 + 2) * (NZ3 + 4) + (NX3 + 4) * NY3 * (NZ3 + 2)) / 3.0
    return npos * 10 * gauss_cost(scheme, frac_g)


def normal3_flops_per_launch(scheme: str, frac_n: float) -> float:
    """k3_normal: LR = 2 and LR = 1 (on R (L q)) characteristic reconstructions per position."""
    npos = ((NX3 + 2) * (NY3 + 4) * (NZ3 + 4) + (NX3 + 4) * (NY3 + 2) * (NZ3 + 4)
            + (NX3 + 4) * (NY3 + 4) * (NZ3 + 2)) / 3.0
    per_field = ROW5 + 5 * DOT5 + normal_cost(scheme, frac_n) + BACK5
    return npos * (2 * (ROE5 + EIG5 + 5 * per_field) + ROUNDTRIP5 + ROE5 + EIG5)


def bench_shockbubble3d(args) -> int:
    import torch
    from weno_sr_publishing_amd.checkpoints import load_f32, load_f64
    from weno_sr_publishing_amd.euler3d import Euler3Solver, ShockBubbleConfig, primitive3
    from weno_sr_publishing_amd.initial_device import shockbubble_bench_state

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1
    dev = torch.device("cuda:0")
    steps = args.steps if args.steps > 0 else STEPS3
    q_init0 = shockbubble_bench_state(NX3, NY3, NZ3, device=dev)   # built on the device
    cfg = ShockBubbleConfig(nx=NX3, ny=NY3, nz=NZ3, x_end=float(NX3), y_end=float(NY3), z_end=float(NZ3),
                            cfl=0.4, t_end=float("inf"))
    mlp32 = load_f32()
    mlps = {"sr32": mlp32, "sr64": load_f64(), "srbf16": mlp32, "js": None}
    ncell = NX3 * NY3 * NZ3

    def make(scheme):
        return Euler3Solver(cfg, scheme=scheme, riemann="hllc", mlp=mlps[scheme], device=dev)

    def timed(scheme, instrument=False):
        s = make(scheme)
        q_init = s.new_state(q_init0)
        q = torch.empty_like(q_init)
        if args.warmup > 0:
            q.copy_(q_init)
            s.run(q, float("inf"), max_steps=args.warmup, chunk=args.warmup)
        q.copy_(q_init)
        torch.cuda.synchronize()
        ms, nl = [0.0, 0.0, 0.0], [0, 0, 0]
        t0 = time.perf_counter()
        s.loop_init(q, float("inf"), steps)
        while True:
            if instrument:
                kms, kn = s.loop_steps_timed(q, 25)
                ms = [a + b for a, b in zip(ms, kms)]
                nl = [a + b for a, b in zip(nl, kn)]
            else:
                s.loop_steps(q, 50)
            n, t, _, done = s.loop_status()
            if done or n >= steps:
                break
        torch.cuda.synchronize()
        return s, q, n, time.perf_counter() - t0, ms, nl

    s_sr, q_sr, n_sr, wall_sr, _, _ = timed(args.closure)
    _, q_js, n_js, wall_js, _, _ = timed("js")
    pri = primitive3(q_sr.cpu().numpy()[3:-3, 3:-3, 3:-3])
    healthy = bool(np.isfinite(pri).all() and pri[..., 0].min() > 0 and pri[..., 4].min() > 0)
    _, _, _, wall_i, kms, kn = timed(args.closure, instrument=True)
    _, _, _, wall_js_i, kms_js, kn_js = timed("js", instrument=True)

    s_cnt = make(args.closure)
    s_cnt.set_counting(True)
    q_cnt = s_cnt.new_state(q_init0)
    s_cnt.run(q_cnt, float("inf"), max_steps=steps, chunk=50)
    c = s_cnt.counters()
    frac_n, frac_g = c[1] / max(1, c[0]), c[3] / max(1, c[2])
    flux_ms = kms[2] / max(1, kn[2])
    flops = flux3_flops_per_launch(args.closure, frac_g)
    achieved = flops / (flux_ms * 1e-3) / 1e12
    nf3 = ((NX3 + 1) * NY3 * NZ3 + NX3 * (NY3 + 1) * NZ3 + NX3 * NY3 * (NZ3 + 1)) / 3.0
    fp32_3 = nf3 * 20 * frac_g * CLOSURE_FP32.get(args.closure, 0)
    peak3 = blended_peak(flops, fp32_3)
    js_flux_ms = kms_js[2] / max(1, kn_js[2])
    js_achieved = flux3_flops_per_launch("js", 0.0) / (js_flux_ms * 1e-3) / 1e12
    per_kernel = {}
    for k, name in enumerate(("k3_normal", "k3_trans1", "k3_flux")):
        avg = kms[k] / max(1, kn[k])
        fl = (normal3_flops_per_launch(args.closure, frac_n), trans13_flops_per_launch(args.closure, frac_g),
              flops)[k]
        per_kernel[name] = {"avg_launch_ms": avg, "launches": kn[k], "algorithmic_flops_per_launch": fl,
                            "achieved_TFLOPs": fl / (avg * 1e-3) / 1e12}
    # algorithmic HBM bytes of one flux launch: 20 first-pass stencil values read per lane pair x 5 ... counted
    # as the unique data: four first-pass planes (2 x 5 x 8 B per position) + the flux planes (5 x 8 B per face)
    nf = ((NX3 + 1) * NY3 * NZ3 + NX3 * (NY3 + 1) * NZ3 + NX3 * NY3 * (NZ3 + 1)) / 3.0
    flux_bytes = nf * (4 * 5 * 8 + 5 * 8)
    value = ncell * n_sr / wall_sr
    out = {
        "metric": "3D Euler cell-updates/s, WENO5-SR vs WENO-JS, 224x88x88 Mach-3 shock-bubble",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": n_sr,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_sr / n_sr,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: planar Mach-3 shock (rho=1, p=1 ahead) and a light-gas bubble (radius 0.25 ny, "
                "density ratio 0.138), 3x3x3 Gauss-Legendre cell averages; trained WENO5-SR checkpoints",
        "config": {"workload": f"3D Euler 224x88x88 shock-bubble, WENO5-SR ({args.closure} closure) + HLLC + "
                               f"SSPRK3, CFL 0.4, {n_sr} fixed steps", "nx": NX3, "ny": NY3, "nz": NZ3,
                   "closure": args.closure, "state": "fp64", "riemann": "hllc", "boundary": "transmissive"},
        "wall_s": wall_sr,
        "healthy": healthy,
        "wall_s_event_instrumented": {"sr": wall_i, "js": wall_js_i},
        "js": {"value": ncell * n_js / wall_js, "steps": n_js, "wall_s": wall_js,
               "ms_per_step": 1e3 * wall_js / n_js},
        "sr_vs_js_wall": wall_sr / wall_js,
        "roofline": {
            "bound": "valu",
            "compute": "VALU (no MFMA): FP32 closure MLP + FP64 reconstruction/flux; peak is the op mix's blend "
                       "of the MI355X FP64 (78.6) and FP32 (157.3 TFLOP/s) vector rates",
            "kernel": f"k3_flux<x|y|z,{args.closure},hllc>",
            "achieved": achieved,
            "peak": peak3,
            "unit": "TFLOP/s",
            "frac": achieved / peak3,
            "peak_fp32_share": fp32_3 / flops,
            "traffic": FLUX3_TRAFFIC_BYTES_224 if args.closure == "sr32" else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
            "avg_launch_ms": flux_ms,
            "launches": kn[2],
            "algorithmic_flops_per_launch": flops,
            "closure_fraction": {"normal": frac_n, "gauss": frac_g},
            "closure_wave_fraction": {"normal": c[4] / max(1, c[0]), "gauss": c[5] / max(1, c[2])},
            "hbm_algorithmic_GBs": flux_bytes / (flux_ms * 1e-3) / 1e9,
            "hbm_frac": flux_bytes / (flux_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "per_kernel": per_kernel,
            "js_flux": {"avg_launch_ms": js_flux_ms, "achieved": js_achieved,
                        "frac": js_achieved / PEAK_VALU_TFLOPS},
        },
    }
    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        s_mid = make(args.closure)
        q_mid = s_mid.new_state(q_init0)
        s_mid.run(q_mid, float("inf"), max_steps=max(1, n_sr // 2), chunk=50)
        qm = q_mid.cpu().numpy()
        done_steps, t0 = 0, time.perf_counter()
        while True:
            qm, _, k, _ = oracle.run3d(qm, NX3, NY3, NZ3, 1.0, 1.0, 1.0, 0.4, float("inf"), max_steps=1,
                                       scheme=args.closure, riemann="hllc", mlp=mlps[args.closure])
            done_steps += k
            el = time.perf_counter() - t0
            if k == 0 or el >= args.cpu_seconds or done_steps >= 20:
                break
        cpu_rate = ncell * done_steps / el
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{done_steps} SSPRK3 steps of the 224x88x88 case ({args.closure} closure, "
                                         f"HLLC) from the step-{max(1, n_sr // 2)} state, OpenMP C oracle, {el:.1f} s",
                               "cpu_model": cpu_model_name(), "host_cpus": os.cpu_count()}
        out["gpu_vs_cpu"] = value / cpu_rate
    copy_calibration(args, out, torch, dev)
    print(json.dumps(out), flush=True)
    return 0


# ------------------------------------------------------------- WENO7 workload
# Double Mach reflection on [0,4]x[0,1] at 1200x300 (Mach-10 shock at 60 degrees,
# pre-shock rho=1.4, p=1), WENO7-SR with the FP32 closure and the state-level
# WENO7-Z fallback, HLLC, Shu RK4 with the downwind operator, CFL 0.4, t=0.2.
NX7, NY7, T7 = 1200, 300, 0.2
# Operations per call of the building blocks of a WENO7 scalar reconstruction, tallied
# by the op-counting build of the oracle (tests/test_opcount.py checks each value):
OPS7 = {
    "candidates_normal": 28, "candidates_gauss": 64,
    "beta_poly": 116, "js_combine": 35,      # JS: smoothness indicators; weights and sum of one head
    "beta_sos": 96, "z_combine": 43,         # Z: the same
    "plateau_test": 85, "features": 132,     # SR: plateau test (always), features (closure stencils)
    "mlp_raw": 2008,                         # one 6-24-16-16-4 pass
    "reflection_average": 8, "beta_from_ratios": 4, "sr_combine": 30,
    "guard": 121,                            # Z-fallback check of one head: beta_sos + the trip test
}
MLP7_RAW = OPS7["mlp_raw"]
_TRIP7 = OPS7["guard"] - OPS7["beta_sos"]
_SR7_CLOSURE = OPS7["features"] + 2 * OPS7["mlp_raw"] + OPS7["reflection_average"] + OPS7["beta_from_ratios"]
NORMAL7 = {
    "js": OPS7["candidates_normal"] + OPS7["beta_poly"] + OPS7["js_combine"],
    "z": OPS7["candidates_normal"] + OPS7["beta_sos"] + OPS7["z_combine"],
    "sr": (OPS7["plateau_test"] + OPS7["candidates_normal"] + OPS7["sr_combine"],
           OPS7["plateau_test"] + OPS7["candidates_normal"] + OPS7["sr_combine"] + _SR7_CLOSURE),
}
# a Gauss pair: one transverse stencil, both Gauss-point heads, sharing the smoothness
# indicators or the plateau test and closure (and the guard's beta_sos)
GAUSS7 = {
    "js": OPS7["beta_poly"] + 2 * (OPS7["candidates_gauss"] + OPS7["js_combine"]),
    "z": OPS7["beta_sos"] + 2 * (OPS7["candidates_gauss"] + OPS7["z_combine"]),
    "sr": (OPS7["plateau_test"] + 2 * (OPS7["candidates_gauss"] + OPS7["sr_combine"]),
           OPS7["plateau_test"] + 2 * (OPS7["candidates_gauss"] + OPS7["sr_combine"]) + _SR7_CLOSURE),
}
GUARD7 = {"normal": OPS7["guard"], "gauss": OPS7["beta_sos"] + 2 * _TRIP7}
# per face and transverse line: Roe average and Jacobian values, the eight cells of the
# two 7-cell stencils projected, the two face states projected back
FACE_LINE7_FIXED = OPS["roe"] + OPS["jacobian"] + 8 * OPS["project"] + 2 * OPS["back_project"]


def sweep7_flops_per_launch(scheme: str, frac_n: float, frac_g: float, both: bool, guard: bool = False) -> float:
    """One k7_sweep_xy launch at 1200x300: the x tiles and the y tiles; ``both`` adds
    the downwind operator's fluxes, ``guard`` the Z-fallback check of closure stencils."""
    if scheme.startswith("sr"):
        n = NORMAL7["sr"][0] + frac_n * (NORMAL7["sr"][1] - NORMAL7["sr"][0] + (GUARD7["normal"] if guard else 0))
        g = GAUSS7["sr"][0] + frac_g * (GAUSS7["sr"][1] - GAUSS7["sr"][0] + (GUARD7["gauss"] if guard else 0))
    else:
        n, g = NORMAL7[scheme], GAUSS7[scheme]
    flux = (4 if both else 2) * OPS["hllc"] + (8 if both else 4)
    total = 0.0
    for nfaces, ntrans in ((NX7 + 1, NY7), (NY7 + 1, NX7)):
        total += nfaces * (ntrans + 6) * (FACE_LINE7_FIXED + 8 * n) + nfaces * ntrans * (8 * g + flux)
    return total


# the multiply-adds of one 6-24-16-16-4 pass (6x24 + 24x16 + 16x16 + 16x4, 2 ops each): in
# the fused FP32 sweep they run on the matrix cores, the rest of the pass on the VALU
MLP7_MADD_OPS = 2 * (6 * 24 + 24 * 16 + 16 * 16 + 16 * 4)


def sweep7_mfma_ops_per_launch(frac_n: float, frac_g: float) -> float:
    """Closure multiply-add ops of one k7_sweep_xy launch at 1200x300 (both passes of every
    closure evaluation; the closure fractions as in sweep7_flops_per_launch)."""
    total = 0.0
    for nfaces, ntrans in ((NX7 + 1, NY7), (NY7 + 1, NX7)):
        total += nfaces * (ntrans + 6) * 8 * frac_n + nfaces * ntrans * 8 * frac_g
    return total * 2 * MLP7_MADD_OPS


def bench_doublemach7(args) -> int:
    import torch
    from weno_sr_publishing_amd.checkpoints import WENO7_SR_FP64, load_weno7_checkpoint
    from weno_sr_publishing_amd.common import interior_stats
    from weno_sr_publishing_amd.euler7 import Euler7Solver, double_mach_params7
    from weno_sr_publishing_amd.initial_device import double_mach_state

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1
    dev = torch.device("cuda:0")
    params = double_mach_params7(NX7, NY7, cfl=0.4, t_end=T7)
    u_init0 = double_mach_state(NX7, NY7, device=dev)   # built on the device
    mlp64 = load_weno7_checkpoint(WENO7_SR_FP64)
    closure = {"sr32": "sr32", "sr64": "sr64", "srbf16": "sr32"}[args.closure]
    mlps = {"sr32": mlp64.as_precision("fp32"), "sr64": mlp64, "js": None}
    max_steps = args.steps if args.steps > 0 else 10_000_000

    def make(scheme):
        return Euler7Solver(params, scheme, "hllc", "double-mach", mlps[scheme], loop="double-mach",
                            z_fallback=scheme != "js", device=dev)

    def timed(scheme, instrument=False):
        s = make(scheme)
        u_init = s.new_state(u_init0)
        u = torch.empty_like(u_init)
        if args.warmup > 0:
            u.copy_(u_init)
            s.run(u, T7, max_steps=args.warmup, chunk=args.warmup)
        u.copy_(u_init)
        torch.cuda.synchronize()
        ms, nl = [0.0, 0.0], [0, 0]
        t0 = time.perf_counter()
        s.loop_init(u, T7, max_steps)
        while True:
            if instrument:
                kms, kn = s.loop_steps_timed(u, 32)
                ms = [a + b for a, b in zip(ms, kms)]
                nl = [a + b for a, b in zip(nl, kn)]
            else:
                s.loop_steps(u, 64)
            n, t, _, done = s.loop_status()
            if done or n >= max_steps:
                break
        torch.cuda.synchronize()
        return s, u, n, t, time.perf_counter() - t0, ms, nl

    s_sr, u_sr, n_sr, t_sr, wall_sr, _, _ = timed(closure)
    fallbacks = s_sr.fallbacks()
    # reconstructed states per step that the fallback counter can count: four
    # right-hand sides (L~ shares L's reconstructions); normal face states (face, side)
    # on the Gauss stencils' lines and Gauss-point states (face, side, point)
    states_rhs = (2 * ((NX7 + 1) * (NY7 + 6) + (NY7 + 1) * (NX7 + 6))
                  + 4 * ((NX7 + 1) * NY7 + (NY7 + 1) * NX7))
    _, _, n_js, t_js, wall_js, _, _ = timed("js")
    stats = interior_stats(u_sr.cpu().numpy(), params)
    _, _, _, _, wall_i, kms, kn = timed(closure, instrument=True)
    _, _, _, _, wall_js_i, kms_js, kn_js = timed("js", instrument=True)
    s_cnt = make(closure)
    s_cnt.set_counting(True)
    u_cnt = s_cnt.new_state(u_init0)
    s_cnt.run(u_cnt, T7, max_steps=max_steps)
    c = s_cnt.counters()
    frac_n, frac_g = c[1] / max(1, c[0]), c[3] / max(1, c[2])
    both_ms = kms[0] / max(1, kn[0])
    flops = sweep7_flops_per_launch(closure, frac_n, frac_g, True, guard=True)
    achieved = flops / (both_ms * 1e-3) / 1e12
    closures7 = sweep7_mfma_ops_per_launch(frac_n, frac_g) / (2 * MLP7_MADD_OPS)
    fp32_7 = closures7 * (OPS7["features"] + 2 * OPS7["mlp_raw"]) if closure == "sr32" else 0.0
    peak7 = blended_peak(flops, fp32_7)
    js_both_ms = kms_js[0] / max(1, kn_js[0])
    js_achieved = sweep7_flops_per_launch("js", 0.0, 0.0, True) / (js_both_ms * 1e-3) / 1e12
    ncell = NX7 * NY7
    value = ncell * n_sr / wall_sr
    hbm = ((NX7 + 8) * (NY7 + 8) * 32 + ((NX7 + 1) * NY7 + NX7 * (NY7 + 1)) * 64) / (both_ms * 1e-3) / 1e9
    out = {
        "metric": "2D Euler cell-updates/s, WENO7-SR vs WENO7-JS, 1200x300 double Mach reflection",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": n_sr,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_sr / n_sr,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: double Mach reflection initial state (15x15 Gauss-Legendre cell averages); trained "
                "WENO7-SR checkpoint (FP64 parameters rounded to FP32 for the FP32 closure)",
        "config": {"workload": f"2D Euler double Mach 1200x300, WENO7-SR ({closure} closure, WENO7-Z fallback) + "
                               f"HLLC + Shu RK4, CFL 0.4, t=0..{T7 if args.steps <= 0 else round(t_sr, 6)}",
                   "nx": NX7, "ny": NY7, "closure": closure, "state": "fp64", "riemann": "hllc",
                   "boundary": "double-mach", "t_end": t_sr},
        "wall_s": wall_sr,
        "health": stats,
        "z_fallback": {"states": fallbacks, "fraction_of_states": fallbacks / max(1, 4 * states_rhs * n_sr),
                       "criterion": "the closure's dominant candidate has a WENO7-Z indicator above 50x the "
                                    "smallest + 240 h^3, or a ratio is not finite / outside [0, 1]; the whole "
                                    "state (4 components) is then reconstructed with WENO7-Z"},
        "wall_s_event_instrumented": {"sr": wall_i, "js": wall_js_i},
        "js": {"value": ncell * n_js / wall_js, "steps": n_js, "wall_s": wall_js,
               "ms_per_step": 1e3 * wall_js / n_js},
        "sr_vs_js_wall": wall_sr / wall_js,
        "roofline": {
            "bound": "valu",
            "compute": ("VALU (FP64 reconstruction, Riemann flux, closure activations) + MFMA "
                        "v_mfma_f32_4x4x1f32 (the FP32 closure's multiply-adds; share of the algorithmic ops "
                        "in mfma_ops_share); peak is the op mix's blend of the MI355X FP64 (78.6) and FP32 "
                        "(157.3 TFLOP/s, vector and matrix core) rates"
                        if closure == "sr32" else
                        "VALU (no MFMA): FP64 closure MLP + FP64 reconstruction/flux; peak is the MI355X "
                        "FP64-vector rate"),
            "mfma_ops_share": (sweep7_mfma_ops_per_launch(frac_n, frac_g) / flops if closure == "sr32" else 0.0),
            "kernel": f"k7_sweep_xy<{closure},hllc,L+L~> (x and y tiles of a right-hand side in one launch)",
            "achieved": achieved,
            "peak": peak7,
            "unit": "TFLOP/s",
            "frac": achieved / peak7,
            "peak_fp32_share": fp32_7 / flops,
            "traffic": SWEEP7_TRAFFIC_BYTES_1200 if closure == "sr32" else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
            "avg_launch_ms": both_ms,
            "launches": kn[0],
            "l_only_sweep": {"avg_launch_ms": kms[1] / max(1, kn[1]), "launches": kn[1]},
            "algorithmic_flops_per_launch": flops,
            "closure_fraction": {"normal": frac_n, "gauss": frac_g},
            "closure_wave_fraction": {"normal": c[4] / max(1, c[0]), "gauss": c[5] / max(1, c[2])},
            "hbm_algorithmic_GBs": hbm,
            "hbm_frac": hbm / PEAK_HBM_GBS,
            "js_sweep": {"avg_launch_ms": js_both_ms, "achieved": js_achieved,
                         "frac": js_achieved / PEAK_VALU_TFLOPS},
        },
    }
    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        s_mid = make(closure)
        u_mid = s_mid.new_state(u_init0)
        s_mid.run(u_mid, T7, max_steps=max(1, n_sr // 2))
        um = u_mid.cpu().numpy()
        done_steps, t0 = 0, time.perf_counter()
        while True:
            um, _, k, _ = oracle.run7_2d(um, NX7, NY7, params.dx, params.dy, params.gamma, params.cfl, T7,
                                         scheme=closure, riemann="hllc", boundary="double-mach",
                                         mlp=mlps[closure], mode="double-mach", max_steps=1, z_fallback=True)
            done_steps += k
            el = time.perf_counter() - t0
            if k == 0 or el >= args.cpu_seconds or done_steps >= 50:
                break
        cpu_rate = ncell * done_steps / el
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{done_steps} Shu RK4 steps of the 1200x300 double Mach case ({closure} "
                                         f"closure, WENO7-Z fallback, HLLC) from the step-{max(1, n_sr // 2)} state, "
                                         f"OpenMP C oracle, {el:.1f} s",
                               "cpu_model": cpu_model_name(), "host_cpus": os.cpu_count()}
        out["gpu_vs_cpu"] = value / cpu_rate
    copy_calibration(args, out, torch, dev)
    print(json.dumps(out), flush=True)
    return 0


# ------------------------------------------------------------- 1-D workloads
# Operation counts per cell and SSPRK3 stage of the 1-D signed advection step: one
# face reconstruction (NORMAL above) + flux a*face, difference, stage combination.
ADV_STAGE_FIXED = 8
# Reverse pass of one SR reconstruction in the adjoint (recomputed forward MLP pair
# 2 x 415, backward through the pair 2 x 810: both products of every dense layer plus
# the swish / tanh / softmax derivatives, feature and weight derivatives 80).
SR_RECON_BWD = 2 * MLP_RAW + 2 * 810 + 80
PLATEAU_RECON_BWD = 60


def plateau_fraction_1d(u: np.ndarray) -> float:
    """Share of periodic 5-cell stencils that are plateaus (delta_max <= 1e-13 q_scale)."""
    q = [np.roll(u, 2 - m, axis=-1) for m in range(5)]
    d0 = 13.0 / 12.0 * np.abs(q[0] - 2 * q[1] + q[2]) + 0.25 * np.abs(q[0] - 4 * q[1] + 3 * q[2])
    d1 = 13.0 / 12.0 * np.abs(q[1] - 2 * q[2] + q[3]) + 0.25 * np.abs(q[1] - q[3])
    d2 = 13.0 / 12.0 * np.abs(q[2] - 2 * q[3] + q[4]) + 0.25 * np.abs(3 * q[2] - 4 * q[3] + q[4])
    scale = np.maximum(np.max(np.abs(np.stack(q)), axis=0), 1.0)
    return float(np.mean(np.maximum(np.maximum(d0, d1), d2) <= 1e-13 * scale))


def bench_train1d(args) -> int:
    import torch
    from weno_sr_publishing_amd.profiles import make_batch, shifted_target
    from weno_sr_publishing_amd.training import ClosureTrainer, distance_schedule

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1
    dev = torch.device("cuda:0")
    B, N, DIST, CFL = 512, 256, 64, 0.6
    u0, vel = make_batch(B, N, seed=83)
    target = shifted_target(u0, vel, DIST)
    dts = distance_schedule(N, DIST, CFL)
    nsteps = len(dts)
    u0_t = torch.tensor(u0, device=dev)
    vel_t = torch.tensor(vel, device=dev)
    tgt_t = torch.tensor(target, device=dev)
    trainer = ClosureTrainer(precision="fp32", lr=1e-3, device=dev)
    iters = args.steps if args.steps > 0 else 50
    for _ in range(args.warmup):
        trainer.step(u0_t, vel_t, tgt_t, dts)
    torch.cuda.synchronize()
    losses = []
    t0 = time.perf_counter()
    for _ in range(iters):
        losses.append(trainer.step(u0_t, vel_t, tgt_t, dts))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the same iterations with CUDA events around the forward and adjoint launches
    events: dict = {}
    for _ in range(iters):
        loss, grad, _ = trainer.loss_and_grad(u0_t, vel_t, tgt_t, dts, events=events)
        trainer.param.grad.copy_(grad.to(trainer.dtype))
        trainer.opt.step()
    torch.cuda.synchronize()

    def avg_ms(name):
        ev = events[name]
        return sum(ev[i].elapsed_time(ev[i + 1]) for i in range(0, len(ev), 2)) / (len(ev) // 2)

    fwd_ms, adj_ms = avg_ms("forward"), avg_ms("adjoint")
    frac = 1.0 - plateau_fraction_1d(np.concatenate([u0, target]))
    recon = NORMAL["sr"][0] + frac * (NORMAL["sr"][1] - NORMAL["sr"][0])
    fwd_flops = B * N * nsteps * 3 * (recon + ADV_STAGE_FIXED)
    bwd_recon = PLATEAU_RECON_BWD + frac * (SR_RECON_BWD - PLATEAU_RECON_BWD)
    adj_flops = B * N * nsteps * (2 * (recon + ADV_STAGE_FIXED) + 3 * (bwd_recon + 2 * ADV_STAGE_FIXED))
    achieved = adj_flops / (adj_ms * 1e-3) / 1e12
    # FP32: the MLP pair of every closure stencil in the two recomputed stages, and its
    # forward recompute and backward in each of the three reverse stages
    adj_fp32 = B * N * nsteps * frac * (2 * 2 * MLP_RAW + 3 * (2 * MLP_RAW + 2 * 810))
    peak_adj = blended_peak(adj_flops, adj_fp32)
    cell_steps = B * N * nsteps
    value = cell_steps * iters / wall
    out = {
        "metric": "closure training throughput: trajectory cell-updates/s (forward rollout + hand-written adjoint "
                  "+ Adam)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": iters,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / iters,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: 512 random periodic profiles (Fourier, Gaussian, square, triangle, semi-ellipse, "
                "composites; exact cell averages, seed 83), a = +1 / -1 balanced, targets = exact shifted averages",
        "config": {"workload": f"1D closure training, batch {B}, N={N}, 64-cell rollout at CFL 0.6 ({nsteps} "
                               "SSPRK3 steps), MSE loss, FP32 closure gradients", "batch": B, "n": N,
                   "rollout_steps": nsteps, "closure": "sr32", "state": "fp64"},
        "loss_first_last": [losses[0], losses[-1]],
        "roofline": {
            "bound": "valu",
            "compute": "VALU: FP32 closure MLP forward/backward + FP64 state; the parameter-gradient sums "
                       "run as 16x16x4 MFMA outer products; peak is the op mix's blend of the MI355X FP64 (78.6) "
                       "and FP32 (157.3 TFLOP/s) vector rates",
            "kernel": "k_train_adjoint<sr32>",
            "achieved": achieved,
            "peak": peak_adj,
            "unit": "TFLOP/s",
            "frac": achieved / peak_adj,
            "peak_fp32_share": adj_fp32 / adj_flops,
            "traffic": None,
            "avg_launch_ms": adj_ms,
            "launches": iters,
            "algorithmic_flops_per_launch": adj_flops,
            "closure_fraction_estimate": frac,
            "forward_kernel": {"avg_launch_ms": fwd_ms, "algorithmic_flops_per_launch": fwd_flops,
                               "achieved": fwd_flops / (fwd_ms * 1e-3) / 1e12},
            "hbm_algorithmic_GBs": (B * (nsteps + 1) * N * 8 * 2) / (adj_ms * 1e-3) / 1e9,
        },
    }
    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        done, t0 = 0, time.perf_counter()
        while True:
            lo = done % B
            oracle.adv1d_run(u0[lo:lo + 32], vel[lo:lo + 32], dts, float(N), scheme="sr32",
                             mlp=trainer.mlp())
            done += 32
            el = time.perf_counter() - t0
            if el >= args.cpu_seconds:
                break
        cpu_rate = done * N * nsteps / el
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "cell-updates/s (forward rollout only)", "cores": threads,
                               "kind": "port",
                               "sample": f"{done} trajectories (the batch of {B}, cycled in chunks of 32) rolled "
                                         f"forward 64 cells ({nsteps} steps, sr32) by the OpenMP C oracle in "
                                         f"{el:.1f} s; the oracle has no adjoint: a gradient by central "
                                         "differences costs 2 x 189 + 1 such rollouts",
                               "cpu_model": cpu_model_name(), "host_cpus": os.cpu_count()}
        out["gpu_vs_cpu"] = value / cpu_rate
    print(json.dumps(out), flush=True)
    return 0


def bench_gste1d(args) -> int:
    import torch
    from weno_sr_publishing_amd.advection1d import Adv1DSolver
    from weno_sr_publishing_amd.checkpoints import load_f32, load_f64
    from weno_sr_publishing_amd.initial import GSTE_LENGTH, gste_cell_averages

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1
    dev = torch.device("cuda:0")
    n, t_end, cfl = 200, 10.0, 0.6
    dx = GSTE_LENGTH / n
    steps = int(np.ceil(t_end / (cfl * dx)))
    dts = [t_end / steps] * steps
    _, u0 = gste_cell_averages(n, 0.0, 15)
    _, exact = gste_cell_averages(n, t_end, 15)
    closure = {"sr32": "sr32", "sr64": "sr64", "srbf16": "sr32"}[args.closure]
    models = {"sr32": load_f32(), "sr64": load_f64(), "js": None}
    reps = args.steps if args.steps > 0 else 20

    def run(scheme, batch):
        solver = Adv1DSolver(n, 1.0 / dx, models[scheme], dev)
        u_init = torch.tensor(np.tile(u0, (batch, 1)), device=dev)
        u = u_init.clone()
        for _ in range(max(1, args.warmup // 10)):
            u.copy_(u_init)
            solver.rollout(u, [1.0], dts)
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            u.copy_(u_init)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            solver.rollout(u, [1.0], dts)
            b.record()
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return u, float(np.mean(ms))

    u_sr, ms_sr = run(closure, 1)
    _, ms_js = run("js", 1)
    ub, ms_b = run(closure, 4096)
    l1 = float(np.mean(np.abs(u_sr[0].cpu().numpy() - exact)))
    assert np.allclose(ub[0].cpu().numpy(), u_sr[0].cpu().numpy(), rtol=0, atol=1e-13)
    frac = 1.0 - plateau_fraction_1d(u0)
    recon = NORMAL["sr"][0] + frac * (NORMAL["sr"][1] - NORMAL["sr"][0])
    flops_b = 4096 * n * steps * 3 * (recon + ADV_STAGE_FIXED)
    fp32_b = 4096 * n * steps * 3 * frac * 2 * MLP_RAW if closure == "sr32" else 0.0
    peak_b = blended_peak(flops_b, fp32_b)
    value = n * steps / (ms_sr * 1e-3)
    out = {
        "metric": "1D advection cell-updates/s, GSTE (N=200, t=10, CFL 0.6), WENO5-SR",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_sr / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: GSTE profile (Gaussian, square, triangle, semi-ellipse), 15-point Gauss-Legendre cell "
                "averages",
        "config": {"workload": f"1D periodic advection GSTE N={n}, t={t_end}, CFL {cfl} ({steps} SSPRK3 steps), "
                               f"WENO5-SR ({closure} closure); one trajectory per launch", "n": n,
                   "closure": closure},
        "rollout_ms": ms_sr,
        "l1_vs_exact": l1,
        "js": {"rollout_ms": ms_js, "value": n * steps / (ms_js * 1e-3)},
        "batched_4096": {"rollout_ms": ms_b, "value": 4096 * n * steps / (ms_b * 1e-3)},
        "roofline": {
            "bound": "valu",
            "compute": "VALU; the single trajectory occupies one workgroup on one CU (latency-bound): the roofline "
                       "figure is that of 4096 copies of the run in one launch",
            "kernel": f"k_adv1d_rollout<{closure}> (batch 4096)",
            "achieved": flops_b / (ms_b * 1e-3) / 1e12,
            "peak": peak_b,
            "unit": "TFLOP/s",
            "frac": flops_b / (ms_b * 1e-3) / 1e12 / peak_b,
            "traffic": None,
            "avg_launch_ms": ms_b,
            "algorithmic_flops_per_launch": flops_b,
            "closure_fraction_estimate": frac,
        },
    }
    if not args.no_cpu:
        from oracle import oracle
        oracle.set_threads(1)
        runs, t0 = 0, time.perf_counter()
        while True:
            oracle.adv1d_run(u0[None, :], [1.0], dts, 1.0 / dx, scheme=closure, mlp=models[closure])
            runs += 1
            el = time.perf_counter() - t0
            if el >= min(args.cpu_seconds, 5.0):
                break
        cpu_rate = runs * n * steps / el
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                               "sample": f"the whole GSTE run ({steps} steps, {closure}) by the C oracle on one "
                                         f"thread, repeated {runs} times in {el:.2f} s",
                               "cpu_model": cpu_model_name(), "host_cpus": os.cpu_count()}
        out["gpu_vs_cpu"] = value / cpu_rate
    print(json.dumps(out), flush=True)
    return 0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed SSPRK3 steps from the developed flow at t=0.125 (0 = the full trajectory t=0..0.25)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--closure", default="sr32", choices=["sr32", "sr64", "srbf16"])
    ap.add_argument("--workload", default="c4-400",
                    choices=["c4-400", "shockbubble3d", "doublemach7", "train1d", "gste1d"],
                    help="c4-400: the headline 2-D line; shockbubble3d: the 3-D 224x88x88 Mach-3 "
                         "shock-bubble leg (200 fixed steps); doublemach7: WENO7 + Shu RK4 double Mach "
                         "reflection at 1200x300 to t=0.2; train1d: closure training iterations (512 "
                         "trajectories x 256 cells, 64-cell rollout, hand-written adjoint); gste1d: the 1-D "
                         "GSTE advection run (N=200, t=10); each prints its own JSON line")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-full", action="store_true",
                    help="with --steps N: skip the full-trajectory runs (full_trajectory = null)")
    ap.add_argument("--save-start", metavar="NPZ",
                    help="with --steps N: write the developed start states (t=0.125) of both closures and exit")
    ap.add_argument("--start-from", metavar="NPZ",
                    help="with --steps N: load the start states written by --save-start instead of advancing "
                         "from the initial data (a profiled run then holds only the window's kernels)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample length")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy one MIB-sized FP64 buffer into another (a reference kernel "
                         "of known traffic for hardware-counter byte calibration)")
    return ap.parse_args()


def main() -> int:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the WENO5 step is measured on "
                                   "one MI355X (multi-GPU domain decomposition is design-only)",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    if args.workload == "shockbubble3d":
        return bench_shockbubble3d(args)
    if args.workload == "doublemach7":
        return bench_doublemach7(args)
    if args.workload == "train1d":
        return bench_train1d(args)
    if args.workload == "gste1d":
        return bench_gste1d(args)

    import torch
    from weno_sr_publishing_amd.checkpoints import load_f32, load_f64
    from weno_sr_publishing_amd.common import conservative_block_average, rho_errors, state_health
    from weno_sr_publishing_amd.euler2d import Euler2DSolver
    from weno_sr_publishing_amd.initial_device import quadrant_initial
    from weno_sr_publishing_amd.params import Params

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        return 1
    dev = torch.device("cuda:0")
    params = Params(nx=NX, ny=NY, x_length=1.0, y_length=1.0, cfl=0.4, t_end=T_END)
    u0_dev = quadrant_initial(NX, NY, "c4", device=dev)   # built on the device
    mlp32 = load_f32()
    mlps = {"sr32": mlp32, "sr64": load_f64(), "srbf16": mlp32.as_precision("bf16"), "js": None}

    def make(scheme, p=params):
        return Euler2DSolver(p, scheme, "hllc", "transmissive", mlps.get(scheme), False, dev)

    def loop(s, u, t_span: float, max_steps: int, instrument: bool):
        """The device loop over ``t_span`` (at most ``max_steps`` steps) from the resident
        state ``u``; returns (steps, t advanced, wall s, sweep ms, sweep launches)."""
        sweep_ms, sweep_n = 0.0, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.loop_init(u, t_span, max_steps)
        steps = 0
        while True:
            chunk = min(64, max_steps - steps)   # no launches past the last step
            if instrument:
                ms, n = s.loop_steps_timed(u, chunk)
                sweep_ms += ms
                sweep_n += n
            else:
                s.loop_steps(u, chunk)
            steps, t, _, done = s.loop_status()
            if done or steps >= max_steps:
                break
        torch.cuda.synchronize()
        return steps, t, time.perf_counter() - t0, sweep_ms, sweep_n

    # Window start.  Transmissive boundaries do not depend on time, so a window from
    # the state at t_mid is the device loop run over T_END - t_mid from that state.
    windowed = args.steps > 0
    t_mid = T_MID if windowed else 0.0
    starts = {}
    loaded = None
    if windowed and args.start_from:
        with np.load(args.start_from, allow_pickle=False) as z:
            loaded = {k: z[k] for k in z.files}
        if float(loaded["t_mid"]) != t_mid:
            raise SystemExit(f"{args.start_from}: start states are at t={float(loaded['t_mid'])}, not {t_mid}")
    for scheme in (args.closure, "js"):
        s = make(scheme)
        if loaded is not None:
            u = s.new_state(loaded[scheme])
        else:
            u = s.new_state(u0_dev)
            if windowed:
                s.run(u, t_mid)
        starts[scheme] = u
    if windowed and args.save_start:
        np.savez(args.save_start, t_mid=np.float64(t_mid),
                 **{k: v.cpu().numpy() for k, v in starts.items()})
        print(json.dumps({"saved_start_states": args.save_start, "t_mid": t_mid, "closures": list(starts)}),
              flush=True)
        return 0
    max_steps = args.steps if windowed else 10_000_000

    def timed(scheme: str, instrument: bool):
        s = make(scheme)
        u = torch.empty_like(starts[scheme])
        if args.warmup > 0:
            u.copy_(starts[scheme])
            loop(s, u, T_END - t_mid, args.warmup, False)
        u.copy_(starts[scheme])
        return (s, u) + loop(s, u, T_END - t_mid, max_steps, instrument)

    # timed region: the plain device loop (no events between kernels) -> `value`
    s_sr, u_sr, steps_sr, dt_sr, wall_sr, _, _ = timed(args.closure, False)
    _, u_js, steps_js, dt_js, wall_js, _, _ = timed("js", False)
    health = state_health(u_sr.cpu().numpy(), 3, NX, NY)
    if not health["complete"]:
        print(json.dumps({"error": "WENO5-SR run did not complete", "health": health}), flush=True)
        return 1
    # the same windows again with HIP events around every sweep launch (on the
    # solver's stream) -> average sweep-kernel duration for the roofline
    _, _, _, _, wall_i, sweep_ms, sweep_n = timed(args.closure, True)
    _, _, _, _, wall_js_i, js_sweep_ms, js_sweep_n = timed("js", True)

    value = NX * NY * steps_sr / wall_sr
    js_value = NX * NY * steps_js / wall_js
    t_from, t_to = t_mid, t_mid + dt_sr

    # closure-evaluation counts over the same (deterministic) window, untimed
    s_cnt = make(args.closure)
    s_cnt.set_counting(True)
    u_cnt = starts[args.closure].clone()
    loop(s_cnt, u_cnt, T_END - t_mid, max_steps, False)
    c = s_cnt.counters()
    frac_n = c[1] / max(1, c[0])
    frac_g = c[3] / max(1, c[2])
    # lanes whose wave ran the closure: the VALU work actually issued for it
    wave_n = c[4] / max(1, c[0])
    wave_g = c[5] / max(1, c[2])
    flops = sweep_flops_per_launch(args.closure, frac_n, frac_g)
    peak = blended_peak(flops, sweep_closures_per_launch(frac_n, frac_g) * CLOSURE_FP32.get(args.closure, 0))
    avg_ms = sweep_ms / max(1, sweep_n)
    achieved = flops / (avg_ms * 1e-3) / 1e12
    js_avg_ms = js_sweep_ms / max(1, js_sweep_n)
    js_achieved = sweep_flops_per_launch("js", 1.0, 1.0) / (js_avg_ms * 1e-3) / 1e12
    hbm_gbs = ((NX + 1) * (NY + 4) + (NY + 1) * (NX + 4)) * BYTES_PER_SWEEP_CELL / (avg_ms * 1e-3) / 1e9

    # the whole trajectory t = 0 .. T_END from the IC for both closures (extra keys;
    # with --steps 0 it is the timed region itself)
    if windowed and args.no_full:
        full = {}
    elif windowed:
        full = {}
        for scheme in (args.closure, "js"):
            s = make(scheme)
            u = s.new_state(u0_dev)
            n, _, w, _, _ = loop(s, u, T_END, 10_000_000, False)
            full[scheme] = (u, n, w)
    else:
        full = {args.closure: (u_sr, steps_sr, wall_sr), "js": (u_js, steps_js, wall_js)}
    f_sr, f_js = full.get(args.closure), full.get("js")
    window = (f"t={t_from:.4f}..{t_to:.4f} ({steps_sr} steps from the developed flow)" if windowed
              else f"t=0..{T_END} (full trajectory, {steps_sr} steps)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps_sr,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_sr / steps_sr,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: C.4 four-quadrant Riemann initial data (15x15 Gauss-Legendre cell averages); "
                "trained WENO5-SR checkpoints",
        "config": {"workload": f"2D Euler C.4 400x400, WENO5-SR ({args.closure} closure) + HLLC + SSPRK3, "
                               f"CFL 0.4, {window}",
                   "nx": NX, "ny": NY, "closure": args.closure, "state": "fp64", "riemann": "hllc",
                   "boundary": "transmissive", "t_start": t_from, "t_stop": t_to},
        "wall_s": wall_sr,
        "wall_s_event_instrumented": {"sr": wall_i, "js": wall_js_i},
        "js": {"value": js_value, "steps": steps_js, "wall_s": wall_js, "ms_per_step": 1e3 * wall_js / steps_js},
        "sr_vs_js_wall": wall_sr / wall_js,
        "full_trajectory": None if f_sr is None else {
            "t": [0.0, T_END],
            "sr": {"value": NX * NY * f_sr[1] / f_sr[2], "steps": f_sr[1], "wall_s": f_sr[2]},
            "js": {"value": NX * NY * f_js[1] / f_js[2], "steps": f_js[1], "wall_s": f_js[2]},
            "sr_vs_js_wall": f_sr[2] / f_js[2],
        },
        "reference_rtx5080_wall_s": {"sr32": REFERENCE_SR32_WALL_S, "js": REFERENCE_JS_WALL_S},
        "speedup_vs_reference_wall": ({"sr": REFERENCE_SR32_WALL_S / f_sr[2], "js": REFERENCE_JS_WALL_S / f_js[2]}
                                      if args.closure == "sr32" and f_sr is not None else None),
        "roofline": {
            "bound": "valu",
            "compute": "VALU (no MFMA): FP32 closure MLP + FP64 reconstruction/flux; peak is the op mix's blend "
                       "of the MI355X FP64 (78.6) and FP32 (157.3 TFLOP/s) vector rates",
            "kernel": f"k_sweep_xy<{args.closure},hllc> (x and y tiles of a stage in one launch)",
            "window": window,
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "peak_fp32_share": sweep_closures_per_launch(frac_n, frac_g) * CLOSURE_FP32.get(args.closure, 0) / flops,
            "traffic": SWEEP_TRAFFIC_BYTES_400 if args.closure == "sr32" else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
            "avg_launch_ms": avg_ms,
            "launches": sweep_n,
            "algorithmic_flops_per_launch": flops,
            "closure_fraction": {"normal": frac_n, "gauss": frac_g},
            "closure_wave_fraction": {"normal": wave_n, "gauss": wave_g},
            "hbm_algorithmic_GBs": hbm_gbs,
            "hbm_frac": hbm_gbs / PEAK_HBM_GBS,
            "js_sweep": {"avg_launch_ms": js_avg_ms, "achieved": js_achieved, "frac": js_achieved / PEAK_VALU_TFLOPS},
        },
    }

    if not args.no_accuracy:
        p8 = Params(nx=2 * NX, ny=2 * NY, x_length=1.0, y_length=1.0, cfl=0.4, t_end=T_END)
        s8 = make("js", p8)
        u8 = quadrant_initial(2 * NX, 2 * NY, "c4", device=dev)
        s8.run(u8, T_END)
        ref = conservative_block_average(u8.cpu().numpy()[3:-3, 3:-3], NX)[..., 0]
        acc = {}
        for scheme in ("js", "sr32", "sr64", "srbf16"):
            if scheme in full:
                field = full[scheme][0]
            else:
                s = make(scheme)
                field = s.new_state(u0_dev)
                s.run(field, T_END)
            acc[scheme] = rho_errors(field.cpu().numpy()[3:-3, 3:-3, 0], ref)["rho_l1"]
        out["accuracy"] = {"rho_l1_vs_800sq_fp64_js": acc,
                           "sr_equal_or_lower_l1_than_js": acc[args.closure] <= acc["js"]}

    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        # the same window on the host cores: from the same start state, the oracle's
        # SSPRK3 steps over the window's steps, repeated until the time budget is spent
        start = starts[args.closure].cpu().numpy()
        mlp = mlps[args.closure]
        limit = steps_sr if windowed else 200
        done_steps, passes, t0 = 0, 0, time.perf_counter()
        while True:
            u, n_pass = start, 0
            while n_pass < limit:
                u, _, n, _ = oracle.run2d(u, NX, NY, params.dx, params.dy, params.gamma, params.cfl,
                                          T_END - t_mid, scheme=args.closure, mlp=mlp, max_steps=1)
                if n == 0:
                    break
                n_pass += n
                if not windowed and time.perf_counter() - t0 >= args.cpu_seconds:
                    break
            done_steps += n_pass
            passes += 1
            el = time.perf_counter() - t0
            if n_pass == 0 or el >= args.cpu_seconds or not windowed:
                break
        cpu_rate = NX * NY * done_steps / el
        where = (f"the GPU window's {steps_sr} steps from t={t_from:.4f}, {passes} pass(es)" if windowed
                 else f"the first {done_steps} steps from t=0")
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                               "sample": f"{done_steps} SSPRK3 steps of C.4 400x400 ({args.closure} closure): "
                                         f"{where}, OpenMP C oracle, {el:.1f} s",
                               "cpu_model": cpu_model_name(), "host_cpus": os.cpu_count()}
        out["gpu_vs_cpu"] = value / cpu_rate

    copy_calibration(args, out, torch, dev)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
