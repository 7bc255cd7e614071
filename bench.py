#!/usr/bin/env python3
"""Throughput of the spectral radiation integral on one MI355X.

Default workload (configuration 2 of the benchmark set): the betatron beam,
2000 electrons x 8000 steps, far field 'total', 512 omega x 64 theta x 64 phi
= 3.355e13 point-step evaluations per spectrum, split precision.  One "step"
is one full spectrum of that batch with tracks and grid resident in HBM.  The
FP64 path is timed on the same full batch (one untimed and one timed spectrum)
and the split spectrum is compared with the FP64 one (relative L2, largest
error).

Prints one JSON line.  ``roofline`` sets the tile kernel's executed vector
work (issued FP32 and FP64 lane-flops per launch, from the rocprofv3 counter
totals committed in profiles/) against MI355X's vector peaks, and states the
reference integrand's algorithmic rate (49 flop per far-field evaluation, 54
near) beside it; ``cpu_baseline`` times the CPU FP64 oracle
(a C restatement of the reference kernels, not the reference) on a bounded
share of the same workload on this host, and ``accuracy_vs_oracle`` compares
both GPU paths with that oracle result (relative L2 and largest error).
``end_to_end`` times one drop-in ``SynchRad.calculate_spectrum`` from host
track lists (packing, uploads, kernel, download); ``value`` is always the
HBM-resident rate.

Other configurations (--config 1..5) print their own line: 1 undulator (FP64),
3 uneven CSR batch (cartesian; also one launch per particle on a share),
4 near field, 5 forward + adjoint of a spectral loss.
``--copy-calibration BYTES`` only runs a device-to-device copy of BYTES (for
calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "G spectral-pt×step evals/s: split-FP32 vs FP64 @≤1e-4 relL2; % vector peak"
FLOP_PER_EVAL = {"far": 49.0, "near": 54.0}
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(REPO, "profiles", "tile_kernel_betatron_traffic.md")
COUNTERS_FILE = os.path.join(REPO, "profiles", "tile_kernel_betatron_counters.md")
ADJOINT_COUNTERS_FILE = os.path.join(REPO, "profiles", "adjoint_betatron_counters.md")
# CPU oracle sample per configuration: ~10-25 s on 16 host cores
CPU_SAMPLE_PARTICLES = {1: 1, 2: 24, 3: 24, 4: 4}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, help="workload number (1-5), default 2")
    ap.add_argument("--particles", type=int, default=None, help="override the particle count")
    ap.add_argument("--f64-particles", type=int, default=None,
                    help="particles of the FP64 leg (default: the whole workload)")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="always the library's own plan")
    ap.add_argument("--no-e2e", action="store_true", help="skip the drop-in call from host arrays")
    ap.add_argument("--cpu-particles", type=int, default=None,
                    help="particles of the CPU oracle sample (default: about 10-25 s of CPU work: "
                         "24 for configurations 2 and 3, 4 for 4)")
    ap.add_argument("--cpu-grid", type=str, default="128,32,16")
    ap.add_argument("--per-particle-share", type=int, default=400,
                    help="config 3: particles timed with one launch each")
    ap.add_argument("--polar-screen", action="store_true",
                    help="config 4: the reference's polar screen instead of the 64 x 64 Cartesian one")
    ap.add_argument("--copy-calibration", type=int, default=0)
    ap.add_argument("--active-particles", type=int, default=100,
                    help="particles over which the filter active fraction is counted")
    return ap.parse_args()


SCREEN = "cartesian"


def build_workload(cfg: int, n_particles=None):
    from synchrad_amd import workloads
    fn = workloads.CONFIGS[cfg]
    if cfg == 1:
        return fn()
    if cfg == 4:
        return fn(n_particles=n_particles, screen=SCREEN)
    return fn(n_particles=n_particles)


def device_objects(wl):
    from synchrad_amd import engine
    return engine.DeviceBatch(wl.batch), wl.device_grid()


def time_spectra(batch, grid, comp, precision, steps, warmup, tuning=None):
    """Mean ms per spectrum (HIP events around `steps` launches on the launch
    stream, outputs cleared inside the timed region), mean ms of the library
    call alone (events around each call, clearing excluded: the tile kernel
    plus, when particle slices are used, the slice reduction), wall ms, outputs."""
    import torch
    from synchrad_amd import engine
    st = torch.cuda.current_stream()
    outs = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=tuning)
    for _ in range(warmup):
        for o in outs:
            o.zero_()
        engine.spectrum(batch, grid, comp=comp, precision=precision, outs=outs, tuning=tuning)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps + 2)]
    wall0 = time.perf_counter()
    ev[0].record(st)
    for i in range(steps):
        for o in outs:
            o.zero_()
        ev[2 + 2 * i].record(st)
        engine.spectrum(batch, grid, comp=comp, precision=precision, outs=outs, tuning=tuning)
        ev[3 + 2 * i].record(st)
    ev[1].record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - wall0) * 1e3 / steps
    kernel = sum(ev[2 + 2 * i].elapsed_time(ev[3 + 2 * i]) for i in range(steps)) / steps
    return ev[0].elapsed_time(ev[1]) / steps, kernel, wall, outs


def rel_l2_dev(a_list, b_list):
    import torch
    num = sum(float(torch.sum((a - b) ** 2)) for a, b in zip(a_list, b_list))
    den = sum(float(torch.sum(b ** 2)) for b in b_list)
    return (num / den) ** 0.5 if den > 0 else float("nan")


def max_err_dev(a_list, b_list):
    """Largest elementwise |a - b| relative to the largest |b|."""
    import torch
    num = max(float(torch.max(torch.abs(a - b))) for a, b in zip(a_list, b_list))
    den = max(float(torch.max(torch.abs(b))) for b in b_list)
    return num / den if den > 0 else float("nan")


def active_fraction(wl, n_particles):
    """Share of point-step evaluations that pass the reference's filter
    |omega (tau_s - tau_{s-1})| < pi (far field, phase_{-1} = 0), counted on
    the device for the first `n_particles` particles in the threshold form
    |omega| < pi / |tau_s - tau_{s-1}|."""
    import torch
    if wl.mode != "far":
        return None
    hb, ax = wl.batch, wl.axes
    dev = torch.device("cuda")
    om = torch.sort(torch.abs(torch.as_tensor(2 * np.pi * ax["omega"], device=dev)))[0]
    th, ph = np.meshgrid(ax["theta"], ax["phi"])
    n = torch.as_tensor(np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]).reshape(3, -1),
                        device=dev)
    act = tot = 0
    for p in range(min(n_particles, hb.n_particles)):
        a, b = int(hb.offsets[p]), int(hb.offsets[p + 1])
        if b - a < 2:
            continue
        X = torch.as_tensor(np.stack([hb.x[a:b - 1], hb.y[a:b - 1], hb.z[a:b - 1]]), device=dev)
        t = (int(hb.it_start[p]) + torch.arange(b - a - 1, device=dev, dtype=torch.float64)) * hb.dt
        tau = t[:, None] - X.T @ n
        d = torch.abs(torch.diff(tau, dim=0, prepend=torch.zeros_like(tau[:1])))
        thr = torch.where(d > 0, np.pi / d, torch.full_like(d, float("inf")))
        act += int(torch.searchsorted(om, thr.reshape(-1).contiguous()).sum())
        tot += thr.numel() * om.numel()
    return act / tot if tot else None


def traffic_from_profile():
    """HBM bytes per launch of the tile kernel on the default workload, as
    measured by rocprofv3 FETCH_SIZE/WRITE_SIZE and recorded in profiles/."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def counters_from_profile(kind, path=COUNTERS_FILE):
    """Per-dispatch counter totals of a profiled kernel (the betatron tile
    kernel: 'split' or 'f64'; the adjoint's transpose sweep: 'adjoint_split'
    or 'adjoint_f64'), as rocprofv3 --pmc recorded them in profiles/: a line
    ``counters[kind]: NAME=value NAME=value ...``."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"counters\[" + kind + r"\]:([^\n]*)", text)
    if not m:
        return None
    return {k: float(v) for k, v in re.findall(r"(\w+)=([0-9.eE+-]+)", m.group(1))}


def executed_roofline(kind, kernel_ms, path=COUNTERS_FILE):
    """The tile kernel's executed vector work against the MI355X vector peaks.

    Issued lane-flops per launch from the counter totals (a wave instruction
    is 64 lanes; FMA = 2, add, mul and transcendental = 1; FP32 and FP64 kept
    apart), divided by the launch time measured now.  ``frac`` is the share of
    the launch the vector ALUs would need at their peaks for that work, FP32
    at 157.3 TFLOP/s plus FP64 at 78.6 -- achieved is stated in FP32-equivalent
    TFLOP/s (an FP64 flop counted as two, the ratio of the peaks)."""
    c = counters_from_profile(kind, path)
    if not c:
        return None
    t = kernel_ms * 1e-3
    f32 = 64.0 * (2 * c.get("SQ_INSTS_VALU_FMA_F32", 0) + c.get("SQ_INSTS_VALU_ADD_F32", 0)
                  + c.get("SQ_INSTS_VALU_MUL_F32", 0) + c.get("SQ_INSTS_VALU_TRANS_F32", 0))
    f64 = 64.0 * (2 * c.get("SQ_INSTS_VALU_FMA_F64", 0) + c.get("SQ_INSTS_VALU_ADD_F64", 0)
                  + c.get("SQ_INSTS_VALU_MUL_F64", 0) + c.get("SQ_INSTS_VALU_TRANS_F64", 0))
    a32, a64 = f32 / t / 1e12, f64 / t / 1e12
    out = {"fp32_tflops": round(a32, 3), "fp64_tflops": round(a64, 3),
           "frac_fp32": round(a32 / PEAK_FP32_TFLOPS, 4), "frac_fp64": round(a64 / PEAK_FP64_TFLOPS, 4),
           "achieved_fp32_equiv": round(a32 + 2 * a64, 3),
           "frac": round(a32 / PEAK_FP32_TFLOPS + a64 / PEAK_FP64_TFLOPS, 4),
           "flop_fp32_per_launch": f32, "flop_fp64_per_launch": f64}
    if "VALUBusy" in c:
        out["valu_busy_rocprof_pct"] = c["VALUBusy"]
    return out


def cpu_baseline(wl_full, cfg, n_particles, grid_str):
    """FP64 oracle on this host: `n_particles` of the workload on a reduced grid."""
    from oracle import oracle
    from synchrad_amd import workloads
    no, nt, nph = (int(v) for v in grid_str.split(","))
    if cfg == 2:
        wl = workloads.config2(n_particles=n_particles, grid=(no, nt, nph))
        sample_grid = f"{no}x{nt}x{nph}"
    else:
        wl = wl_full
        sample_grid = "x".join(str(v) for v in wl.calc_input["grid"][-1])
    tr = wl.track_list()[:n_particles]
    ax = wl.axes
    evals = wl.n_grid * sum(max(t[0].size - 1, 0) for t in tr)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    ref = oracle.spectrum(wl.mode, wl.comp, tr, ax, wl.batch.dt, L_screen=wl.L_screen, threads=threads)
    sec = time.perf_counter() - t0
    desc = f"{len(tr)} particles x {tr[0][0].size} steps on a {sample_grid} grid"
    res = {"value": round(evals / sec / 1e9, 4), "unit": "G evals/s", "cores": threads, "kind": "port",
           "sample": f"{desc} ({evals:.3g} evals, {sec:.1f} s), FP64 C/OpenMP restatement of the "
                     f"reference kernels"}
    return res, (wl, tr, ref, desc)


OUT_KEYS = {"total": ["total"], "cartesian": ["x", "y", "z"]}


def accuracy_vs_oracle(sample):
    """Both GPU paths against the CPU oracle's FP64 result on the cpu_baseline
    sample: relative L2 and the largest elementwise error relative to the
    largest reference value."""
    import torch
    from synchrad_amd import engine
    wl, tr, ref, desc = sample
    ax = wl.axes
    grid = wl.device_grid()
    batch = engine.DeviceBatch(engine.pack_tracks(tr, dt=wl.batch.dt))
    keys = OUT_KEYS[wl.comp]
    r = np.stack([ref[k] for k in keys])
    out = {"sample": desc}
    for prec in ("double", "float"):
        outs = engine.spectrum(batch, grid, comp=wl.comp, precision=prec)
        torch.cuda.synchronize()
        g = np.stack([o.cpu().numpy().swapaxes(-1, -3) for o in outs])
        out["f64" if prec == "double" else "split"] = {
            "rel_l2": float(np.linalg.norm(g - r) / np.linalg.norm(r)),
            "max_err_rel_to_max": float(np.abs(g - r).max() / np.abs(r).max())}
    return out


def end_to_end(wl, precision, calls=3):
    """Drop-in SynchRad.calculate_spectrum calls from host track lists: CSR
    packing, one upload, the library call and the download of the spectrum.
    The first call launches directly; a repeat with the same sizes refills the
    device batch and (from the second call) replays a captured HIP graph."""
    import torch
    from synchrad_amd.calc import SynchRad
    if wl.screen != "polar":
        return {"skipped": "the drop-in SynchRad API keeps the reference's polar screen only"}
    calc = SynchRad(dict(wl.calc_input, dtype=precision))
    tracks = wl.track_list()
    times = []
    for _ in range(calls):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        calc.calculate_spectrum(particleTracks=tracks, timeStep=wl.batch.dt, comp=wl.comp, L_screen=wl.L_screen,
                                verbose=False)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    steady = times[-1]
    return {"value": round(wl.evals / (steady * 1e-3) / 1e9, 3), "unit": "G evals/s", "ms": round(steady, 2),
            "ms_first_call": round(times[0], 2), "ms_second_call_with_capture": round(times[1], 2),
            "path": "SynchRad.calculate_spectrum from host track lists (CSR packing, one H2D copy, library "
                    "call, D2H); 'ms' is a repeated call of the same sizes, which replays the captured graph"}


def plan_for(batch, grid, comp, precision, args):
    from dataclasses import asdict
    from synchrad_amd import autotune
    if args.no_autotune:
        return None, None
    t0 = time.perf_counter()
    plan, rows = autotune.tune(batch, grid, comp=comp, precision=precision)
    log(f"autotune ({time.perf_counter() - t0:.1f} s): {asdict(plan)}")
    return plan, asdict(plan)


def bench_forward(args):
    import torch
    from synchrad_amd import engine
    cfg = args.config
    log(f"building workload config {cfg}")
    wl = build_workload(cfg, args.particles)
    evals = wl.evals
    log(f"workload {wl.name}: {wl.batch.n_particles} particles, {wl.n_grid} points, {evals:.4g} evals")
    batch, grid = device_objects(wl)
    torch.cuda.synchronize()
    flop = FLOP_PER_EVAL[wl.mode]
    main_prec = "double" if cfg == 1 else "float"
    plan, plan_d = plan_for(batch, grid, wl.comp, main_prec, args)

    log(f"timing {main_prec}")
    ms, ms_kernel, wall, outs = time_spectra(batch, grid, wl.comp, main_prec, args.steps, args.warmup, plan)
    value = evals / (ms * 1e-3) / 1e9
    achieved_tf = evals * flop / (ms_kernel * 1e-3) / 1e12
    log(f"{main_prec}: {ms:.1f} ms/spectrum ({ms_kernel:.1f} ms in the library call, {wall:.1f} ms wall), "
        f"{value:.1f} G evals/s, {achieved_tf:.1f} TFLOP/s algorithmic")
    nmax = int(np.diff(wl.batch.offsets).max())
    f_act = active_fraction(wl, args.active_particles)
    if f_act is not None:
        log(f"filter active fraction {f_act:.4f} (first {args.active_particles} particles)")
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
        "dtype": "f64" if main_prec == "double" else "f32 (f64 phase)",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {wl.batch.n_particles} e- x up to {nmax} steps, {wl.mode} "
                               f"field '{wl.comp}', grid "
                               f"{'x'.join(str(v) for v in wl.calc_input['grid'][-1])}",
                   "evals_per_spectrum": evals,
                   "precision": "FP64" if main_prec == "double" else "split-FP32",
                   "plan": plan_d},
        "roofline": roofline_entry(cfg, main_prec, evals, flop, ms_kernel, f_act),
    }

    if main_prec == "float" and not args.no_f64:
        n64 = wl.batch.n_particles if args.f64_particles is None else min(args.f64_particles,
                                                                          wl.batch.n_particles)
        wl64 = build_workload(cfg, n64) if n64 < wl.batch.n_particles else wl
        b64, g64 = (batch, grid) if wl64 is wl else device_objects(wl64)
        log(f"timing FP64 on {n64} particles (1 untimed + 1 timed spectrum)")
        ms64, ms64_kernel, _, outs64 = time_spectra(b64, g64, wl.comp, "double", 1, 1)
        ev64 = wl64.evals
        v64 = ev64 / (ms64 * 1e-3) / 1e9
        outs_s = outs if wl64 is wl else time_spectra(b64, g64, wl.comp, "float", 1, 0, plan)[3]
        rel = rel_l2_dev(outs_s, outs64)
        mx = max_err_dev(outs_s, outs64)
        log(f"FP64: {ms64:.1f} ms ({v64:.1f} G evals/s); split vs FP64 rel-L2 {rel:.3e}")
        res["f64"] = {"value": round(v64, 3), "unit": "G evals/s", "particles": n64, "ms": round(ms64, 3),
                      "roofline": roofline_entry(cfg if wl64 is wl else None, "double", ev64, flop, ms64_kernel,
                                                 None),
                      "form": "this library's FP64 path: FP64 phase, exact filter, FP64 table sin/cos, "
                              "FP64 sums (the same staged tiles as the split path)"}
        res["split_over_f64"] = round(value / v64, 3)
        res["rel_l2_split_vs_f64"] = rel
        res["max_err_split_vs_f64"] = mx

    if cfg == 3 and args.per_particle_share > 0:
        res["per_particle_launches"] = per_particle_compare(wl, grid, plan, args.per_particle_share)

    if not args.no_e2e:
        log("timing the drop-in call from host arrays")
        res["end_to_end"] = end_to_end(wl, main_prec)
        log(f"end to end: {res['end_to_end']}")

    if not args.no_cpu:
        log("timing CPU oracle sample")
        n_cpu = args.cpu_particles if args.cpu_particles is not None else CPU_SAMPLE_PARTICLES.get(cfg, 4)
        res["cpu_baseline"], sample = cpu_baseline(wl, cfg, n_cpu, args.cpu_grid)
        log(f"cpu: {res['cpu_baseline']}")
        res["accuracy_vs_oracle"] = accuracy_vs_oracle(sample)
        log(f"accuracy: {res['accuracy_vs_oracle']}")
    return res


def roofline_entry(cfg, precision, evals, flop, kernel_ms, f_act):
    """The dominant kernel against the vector ALU roofline.  ``achieved`` /
    ``frac``: the executed vector work from the counter totals in profiles/
    (betatron workload only; see executed_roofline).  ``algorithmic``: the
    reference integrand's flop count per evaluation times the evaluations,
    which exceeds the peak because this kernel shares the omega-independent
    work, rotates instead of taking every sin/cos and skips filtered steps."""
    kind = "f64" if precision == "double" else "split"
    ex = executed_roofline(kind, kernel_ms) if cfg == 2 else None
    alg = evals * flop / (kernel_ms * 1e-3) / 1e12
    peak_alg = PEAK_FP64_TFLOPS if precision == "double" else PEAK_FP32_TFLOPS
    out = {"bound": "valu",
           "achieved": ex["achieved_fp32_equiv"] if ex else None,
           "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
           "frac": ex["frac"] if ex else None,
           "traffic": traffic_from_profile() if (cfg == 2 and kind == "split") else None,
           "kernel_ms": round(kernel_ms, 3),
           "executed": ex,
           "algorithmic": {"flop_per_eval": flop, "tflops": round(alg, 3), "peak": peak_alg,
                           "rate_over_peak": round(alg / peak_alg, 4),
                           "filter_active_fraction": round(f_act, 4) if f_act is not None else None},
           "basis": "achieved/frac: issued lane-flops per launch from rocprofv3 counter totals "
                    "(profiles/tile_kernel_betatron_counters.md; FMA 2, add/mul/transcendental 1) over "
                    "this run's launch time (HIP events on the launch stream), FP64 counted as two "
                    "FP32 flops; algorithmic: the reference integrand's flop count per point-step "
                    "evaluation, a rate, not a hardware fraction"}
    return out


def per_particle_compare(wl, grid, plan, n_share):
    """One launch per particle (the reference's pattern, through the batch entry
    point with one particle per call) vs one launch for the same share."""
    import torch
    from synchrad_amd import engine
    hb = wl.batch
    n = min(n_share, hb.n_particles)
    tr = wl.track_list()[:n]
    share = engine.pack_tracks(tr, dt=hb.dt)
    b_all = engine.DeviceBatch(share)
    singles = [engine.DeviceBatch(engine.pack_tracks([t], dt=hb.dt)) for t in tr]
    evals = grid.n_grid * share.n_intervals()
    ms_one, _, _, outs = time_spectra(b_all, grid, wl.comp, "float", 1, 1, plan)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    outs2 = [torch.zeros_like(o) for o in outs]
    engine.spectrum(singles[0], grid, comp=wl.comp, precision="float", outs=outs2, tuning=plan)
    for o in outs2:
        o.zero_()
    torch.cuda.synchronize()
    e0.record(st)
    for b in singles:
        engine.spectrum(b, grid, comp=wl.comp, precision="float", outs=outs2, tuning=plan)
    e1.record(st)
    torch.cuda.synchronize()
    ms_many = e0.elapsed_time(e1)
    return {"particles": n, "evals": evals,
            "one_launch_G_evals_s": round(evals / (ms_one * 1e-3) / 1e9, 3),
            "per_particle_G_evals_s": round(evals / (ms_many * 1e-3) / 1e9, 3),
            "speedup": round(ms_many / ms_one, 3),
            "rel_l2": rel_l2_dev(outs2, outs)}


def bench_adjoint(args):
    """Configuration 5: forward + adjoint of a weighted L2 spectral loss."""
    import torch
    from synchrad_amd import autograd, engine, workloads
    wl = workloads.config5(n_particles=args.particles)
    tgt = workloads.config5(n_particles=args.particles, target=True)
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    dev = grid.device
    hb = wl.batch
    evals = wl.evals
    T = engine.spectrum(engine.DeviceBatch(tgt.batch), grid, comp="total", precision="double")[0]
    th = torch.as_tensor(ax["theta"], device=dev)
    W = (torch.sin(th)[None, None, :, None] * torch.ones_like(T))
    W = W / W.sum()
    leaves = [torch.tensor(getattr(hb, c), device=dev, requires_grad=True)
              for c in ("x", "y", "z", "ux", "uy", "uz", "weight")]
    res = {"metric": "forward+backward ms, spectral-loss adjoint", "unit": "ms", "n_gpus": 1,
           "higher_is_better": False, "data": "synthetic",
           "config": {"workload": f"adjoint-betatron: {hb.n_particles} e- x 4000 steps (fp32 tracks), "
                                  "far 'total', grid 128x32x32", "evals_forward": evals}}
    st = torch.cuda.current_stream()
    for prec in ("float", "double"):
        def step(ev=None):
            for t in leaves:
                t.grad = None
            if ev:
                ev[0].record(st)
            S = autograd.far_spectrum(*leaves, hb, grid, comp="total", precision=prec)[0]
            loss = torch.sum(W * (S - T) ** 2)
            if ev:
                ev[1].record(st)
            loss.backward()
            if ev:
                ev[2].record(st)
            return loss
        step()
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
        for ev in evs:
            step(ev)
        torch.cuda.synchronize()
        ms = sum(ev[0].elapsed_time(ev[2]) for ev in evs) / args.steps
        ms_bwd = sum(ev[1].elapsed_time(ev[2]) for ev in evs) / args.steps
        kind = "adjoint_split" if prec == "float" else "adjoint_f64"
        ex = executed_roofline(kind, ms_bwd, ADJOINT_COUNTERS_FILE)
        res[prec] = {"ms_fwd_bwd": round(ms, 2), "ms_forward": round(ms - ms_bwd, 2), "ms_backward": round(ms_bwd, 2),
                     "G_evals_per_s_fwd_equiv": round(evals / (ms * 1e-3) / 1e9, 2),
                     "roofline_backward": {"bound": "valu", "kernel": "k_adjoint (transpose sweep)",
                                           "achieved": ex["achieved_fp32_equiv"] if ex else None,
                                           "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                           "frac": ex["frac"] if ex else None, "executed": ex,
                                           "basis": "issued lane-flops of k_adjoint per backward from the counter "
                                                    "totals in profiles/adjoint_betatron_counters.md over this "
                                                    "run's backward time (HIP events around loss.backward(); "
                                                    "includes the slab reduction and the loss's own gradient)"}}
        log(f"adjoint {prec}: {ms:.1f} ms forward+backward ({ms_bwd:.1f} ms backward)")
    res["value"] = res["float"]["ms_fwd_bwd"]
    res["ms_per_step"] = res["value"]
    res["roofline"] = res["float"]["roofline_backward"]
    res["steps"], res["warmup"] = args.steps, 1
    res["dtype"] = "f32 (f64 phase)"
    return res


def copy_calibration(nbytes):
    """A plain device copy of `nbytes` bytes, for counter calibration."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)


def main():
    global SCREEN
    args = parse()
    if args.polar_screen:
        SCREEN = "polar"
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding "
                                   "is designed but not built)"}))
        sys.exit(2)
    import torch
    from synchrad_amd import _abi
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    _abi.load()
    res = bench_adjoint(args) if args.config == 5 else bench_forward(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
