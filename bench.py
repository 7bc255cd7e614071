#!/usr/bin/env python3
"""Throughput of the spectral radiation integral on one MI355X.

Default workload: the betatron beam (2000 electrons x 8000 steps, far field
'total', 512 omega x 64 theta x 64 phi = 3.355e13 point-step evaluations per
spectrum) in split precision.  One "step" is one full spectrum of that batch,
with tracks and grid already resident in HBM.  The FP64 path is timed on a
200-particle share of the same beam (evals/s is intensive in the particle
count), and the split result of that share is compared with the FP64 result
(relative L2).

Prints one JSON line.  ``roofline`` prices the tile kernel at the reference
integrand's algorithmic cost (49 flop per far-field evaluation, 54 near) against
MI355X's FP32 vector peak; ``cpu_baseline`` times the CPU FP64 oracle (a
restatement of the reference kernels, not the reference itself) on a bounded
share of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "G spectral-pt×step evals/s: split-FP32 vs FP64 @≤1e-4 relL2; % vector peak"
FLOP_PER_EVAL = {"far": 49.0, "near": 54.0}
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, help="workload number (1-5), default 2")
    ap.add_argument("--particles", type=int, default=None, help="override particle count")
    ap.add_argument("--f64-particles", type=int, default=200)
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=4)
    ap.add_argument("--cpu-grid", type=str, default="128,32,16")
    return ap.parse_args()


def build_workload(cfg: int, n_particles=None):
    from synchrad_amd import workloads
    if cfg == 2:
        return workloads.config2(n_particles=n_particles)
    fn = workloads.CONFIGS[cfg]
    return fn(n_particles=n_particles) if n_particles else fn()


def device_objects(wl):
    from synchrad_amd import engine
    ax = wl.axes
    ax1 = ax["theta"] if wl.mode == "far" else ax["radius"]
    grid = engine.DeviceGrid(wl.mode, ax["omega"], ax1, ax["phi"], L_screen=wl.L_screen)
    batch = engine.DeviceBatch(wl.batch)
    return batch, grid


def time_spectra(batch, grid, comp, precision, steps, warmup):
    """Mean wall ms per spectrum (HIP events around `steps` launches on the
    launch stream), outputs of the last spectrum."""
    import torch
    from synchrad_amd import engine
    st = torch.cuda.current_stream()
    outs = None
    for _ in range(warmup):
        outs = engine.spectrum(batch, grid, comp=comp, precision=precision)
    torch.cuda.synchronize()
    if outs is None:
        outs = engine.spectrum(batch, grid, comp=comp, precision=precision)
        torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(st)
    for _ in range(steps):
        for o in outs:
            o.zero_()
        engine.spectrum(batch, grid, comp=comp, precision=precision, outs=outs)
    t1.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - wall0) * 1e3 / steps
    ms = t0.elapsed_time(t1) / steps
    return ms, wall, outs


def cpu_baseline(wl_full, n_particles, grid_str):
    """FP64 oracle on this host: `n_particles` of the workload on a reduced grid."""
    from oracle import oracle
    from synchrad_amd import workloads
    no, nt, nph = (int(v) for v in grid_str.split(","))
    if wl_full.name == "betatron":
        wl = workloads.config2(n_particles=n_particles, grid=(no, nt, nph))
    else:
        wl = wl_full
    tr = wl.track_list()[:n_particles]
    hb_int = sum(max(min(t[0].size - 1, t[0].size - 1), 0) for t in tr)
    ax = wl.axes
    evals = ax["omega"].size * ax["phi"].size * (ax.get("theta", ax.get("radius"))).size * hb_int
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    oracle.spectrum(wl.mode, wl.comp, tr, ax, wl.batch.dt, L_screen=wl.L_screen, threads=threads)
    sec = time.perf_counter() - t0
    return {"value": evals / sec / 1e9, "unit": "G evals/s", "cores": threads, "kind": "port",
            "sample": f"{len(tr)} particles x {tr[0][0].size} steps on a {no}x{nt}x{nph} grid "
                      f"({evals:.3g} evals, {sec:.1f} s), FP64 C/OpenMP restatement of the "
                      f"reference kernels"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU "
                                   "sharding is designed but not built)"}))
        sys.exit(2)
    import torch
    from synchrad_amd import engine, _abi
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    _abi.load()

    log(f"building workload config {args.config}")
    wl = build_workload(args.config, args.particles)
    evals = wl.evals
    log(f"workload {wl.name}: {wl.batch.n_particles} particles, {wl.n_grid} points, {evals:.4g} evals")
    batch, grid = device_objects(wl)
    torch.cuda.synchronize()

    split_prec = "float"
    log("timing split precision")
    ms, wall, outs = time_spectra(batch, grid, wl.comp, split_prec, args.steps, args.warmup)
    value = evals / (ms * 1e-3) / 1e9
    flop = FLOP_PER_EVAL[wl.mode]
    achieved_tf = evals * flop / (ms * 1e-3) / 1e12
    log(f"split: {ms:.1f} ms/spectrum ({wall:.1f} ms wall), {value:.1f} G evals/s, "
        f"{achieved_tf:.1f} TFLOP/s algorithmic")

    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "G evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 phase)",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {wl.batch.n_particles} e- x "
                               f"{int(np.diff(wl.batch.offsets).max())} steps, {wl.mode} field "
                               f"'{wl.comp}', grid {'x'.join(str(v) for v in wl.calc_input['grid'][-1])}",
                   "evals_per_spectrum": evals, "precision": "split-FP32"},
        "roofline": {"bound": "valu", "achieved": round(achieved_tf, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tf / PEAK_FP32_TFLOPS, 4),
                     "traffic": None,
                     "basis": f"{flop:g} algorithmic flop per point-step evaluation x "
                              f"{evals:.4g} evaluations per launch / mean launch time"},
    }

    if not args.no_f64:
        n64 = min(args.f64_particles, wl.batch.n_particles)
        wl64 = build_workload(args.config, n64) if n64 < wl.batch.n_particles else wl
        b64, g64 = (batch, grid) if wl64 is wl else device_objects(wl64)
        log(f"timing FP64 on {n64} particles")
        ms64, _, outs64 = time_spectra(b64, g64, wl.comp, "double", 1, 0)
        ev64 = wl64.evals
        v64 = ev64 / (ms64 * 1e-3) / 1e9
        _, _, outs_s = time_spectra(b64, g64, wl.comp, "float", 1, 0)
        num = sum(float(torch.sum((a - b) ** 2)) for a, b in zip(outs_s, outs64))
        den = sum(float(torch.sum(b ** 2)) for b in outs64)
        rel = (num / den) ** 0.5 if den > 0 else float("nan")
        log(f"FP64: {ms64:.1f} ms ({v64:.1f} G evals/s); split vs FP64 rel-L2 {rel:.3e}")
        res["f64"] = {"value": round(v64, 3), "unit": "G evals/s", "particles": n64,
                      "frac_fp64_peak": round(v64 * 1e9 * flop / 1e12 / PEAK_FP64_TFLOPS, 4)}
        res["split_over_f64"] = round(value / v64, 3)
        res["rel_l2_split_vs_f64"] = rel

    if not args.no_cpu:
        log("timing CPU oracle sample")
        res["cpu_baseline"] = cpu_baseline(wl, args.cpu_particles, args.cpu_grid)
        log(f"cpu: {res['cpu_baseline']}")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
