"""Autotuner for the tile kernel's launch shape.

For a (mode, comp, precision, grid shape) it times the candidate plans --
sweep (direct / recurrence), omegas per thread M and waves per block (which is
also the number of steps per LDS tile, 64 per wave) -- through the library's
own ``sr_autotune`` (HIP events around the batch call, one untimed run per
candidate), on a sample of the request chosen so that each tile does the work
it does in the full launch:

* small launches: up to 16 of the batch's particles, cut to a common step
  budget (about ``target_evals`` point-step evaluations per run), on the real
  grid (so the filter activity, which sets each tile's work, is the real one);
  when the fastest candidate runs under SHORT_LAUNCH_MS, every candidate is
  timed again over about SHORT_TIMING_MS of launches (one 0.5 ms launch per
  candidate ranked the undulator's plans within noise of each other);
* large far-field launches (>= LARGE_EVALS evaluations), which the library
  slices into particle groups and launches heaviest tile first, in two
  stages: every particle at full length on a strided subset of about
  1/DIRECTION_SHARE of the directions, each candidate held at the
  particle-slice count the library plans for the full launch (``sr_plan``),
  so that a tile -- one direction, one omega chunk, one particle slice --
  carries exactly its full-launch work; then the subset's pick and the
  library's own plan are timed once each on the full launch, and the faster
  is kept.  The second stage is there because the subset ranked the plans
  differently from the full launch on the betatron beam (it preferred
  4 omegas x 2 waves, which ran the full launch 9 % slower than the
  library's 2 x 2).

The fastest plan is cached in memory and, if ``SYNCHRAD_AMD_TUNE_CACHE`` names
a file, in that JSON file.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import asdict, replace
from typing import Dict, List, Tuple

import numpy as np
import torch

from . import _abi, engine

_cache: Dict[str, dict] = {}

# far-field launches from this size on are ordered and sliced by the library
# (its kOrderSliceEvals)
LARGE_EVALS = 1e12
DIRECTION_SHARE = 8
# sample launches shorter than this are timed again, SHORT_TIMING_MS per
# candidate (at most MAX_REPS launches each)
SHORT_LAUNCH_MS = 5.0
SHORT_TIMING_MS = 20.0
MAX_REPS = 50


def candidates(mode: str, precision: str, uniform: bool, n_omega: int) -> List[engine.Tuning]:
    """Plans worth timing: FP64 sweeps directly with 1 or 2 omegas per thread;
    split precision uses the recurrence on linear grids (measured faster there
    than every direct plan) and the direct sweep otherwise.  Blocks whose
    omega run would leave more than a quarter of their lanes idle are left out."""
    T = engine.Tuning
    if engine.PRECISIONS[precision] == _abi.SR_PREC_F64:
        pairs = [("direct", m, w) for m in (1, 2) for w in (1, 2, 4)]
    elif uniform:
        pairs = [("recurrence", m, w) for m in ((2, 4) if mode == "far" else (2,)) for w in (1, 2, 4)]
    else:
        pairs = [("direct", m, w) for m in ((1, 2, 4) if mode == "far" else (1, 2)) for w in (1, 2, 4)]
    out = []
    for sweep, m, w in pairs:
        run = 64 * w * m
        chunks = -(-n_omega // run)
        if n_omega / (chunks * run) >= 0.75 or run == 64:
            out.append(T(sweep=sweep, omega_per_lane=m, waves_per_block=w))
    return out


def _key(batch: engine.DeviceBatch, grid: engine.DeviceGrid, comp: str, precision: str) -> str:
    hb = batch.host
    size = "large" if is_large(batch, grid) else int(min(hb.n_particles, 16))
    return json.dumps([grid.mode, comp, engine.PRECISIONS[precision], grid.n_omega, grid.n1, grid.n_phi,
                       bool(grid.omega_step), int(hb.n_snaps), size])


def is_large(batch: engine.DeviceBatch, grid: engine.DeviceGrid) -> bool:
    return grid.mode == "far" and grid.n_grid * batch.host.n_intervals() >= LARGE_EVALS


def sample_batch(hb: engine.HostBatch, n_grid: int, target_evals: float,
                 max_particles: int = 16) -> engine.HostBatch:
    """The first <= max_particles particles, each cut to the same step budget."""
    Np = min(hb.n_particles, max_particles)
    lengths = np.diff(hb.offsets)[:Np]
    budget = max(8, int(target_evals / max(n_grid, 1) / max(Np, 1)))
    keep = np.minimum(lengths, budget)
    off = np.zeros(Np + 1, dtype=np.int64)
    np.cumsum(keep, out=off[1:])
    idx = np.concatenate([np.arange(hb.offsets[p], hb.offsets[p] + keep[p]) for p in range(Np)]) \
        if Np else np.zeros(0, dtype=np.int64)
    cols = [getattr(hb, c)[idx] for c in ("x", "y", "z", "ux", "uy", "uz")]
    its = hb.it_start[:Np]
    if hb.snaps_stride:
        # per-track mode: each sample particle records at its own (new) end
        return engine.pack_csr(off, *cols, hb.weight[:Np], dt=hb.dt, nSnaps=1)
    return engine.pack_csr(off, *cols, hb.weight[:Np], its, dt=hb.dt,
                           it_range=(0, int(keep.max()) if Np else 1), nSnaps=1)


def direction_subset(grid: engine.DeviceGrid, share: int = DIRECTION_SHARE) -> engine.DeviceGrid:
    """The same omegas on every s_theta-th theta and s_phi-th phi, with
    s_theta * s_phi ~ share (theta strided first, so the polar spread of the
    filter activity is kept)."""
    th = np.arcsin(np.clip(grid.sin_theta.cpu().numpy(), -1, 1))
    th = np.where(grid.cos_theta.cpu().numpy() < 0, np.pi - th, th)
    ph = np.arctan2(grid.sin_phi.cpu().numpy(), grid.cos_phi.cpu().numpy())
    s_th = 2 if (share >= 2 and th.size >= 16) else 1
    s_ph = max(1, min(ph.size // 4, share // s_th))
    sub = engine.DeviceGrid("far", grid.omega_host, th[::s_th], ph[::s_ph], device=grid.device)
    # the subset grid's own sin/cos must be the full grid's values, not recomputed ones
    sub.sin_theta = grid.sin_theta[::s_th].contiguous()
    sub.cos_theta = grid.cos_theta[::s_th].contiguous()
    sub.sin_phi = grid.sin_phi[::s_ph].contiguous()
    sub.cos_phi = grid.cos_phi[::s_ph].contiguous()
    return sub


def _opts(comp: str, precision: str, tun: engine.Tuning, form_factor=None) -> _abi.SrOpts:
    o = _abi.SrOpts()
    o.comp = _abi.COMPS[comp]
    o.precision = engine.PRECISIONS[precision]
    o.omega_per_lane = tun.omega_per_lane
    o.waves_per_block = tun.waves_per_block
    o.slices = tun.slices
    o.persistent = tun.persistent
    o.tile_order = tun.tile_order
    o.sweep = engine.SWEEPS[tun.sweep]
    if form_factor is not None:
        o.form_factor = form_factor.data_ptr()
    return o


def planned(batch: engine.DeviceBatch, grid: engine.DeviceGrid, comp: str, precision: str,
            tun: engine.Tuning) -> _abi.SrPlanInfo:
    """The plan the library resolves for this request (sr_plan)."""
    L = _abi.load()
    info = _abi.SrPlanInfo()
    ff = grid.form_factor(0.0) if (comp == "cartesian_complex" and grid.mode == "far") else None
    tr, gr, o = batch.abi(), grid.abi(), _opts(comp, precision, tun, ff)
    _abi.check(L.sr_plan(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(o), ctypes.byref(info)))
    return info


def time_candidates(batch: engine.DeviceBatch, grid: engine.DeviceGrid, comp: str, precision: str,
                    cands: List[engine.Tuning], reps: int = 1, warmup: int = 1) -> Tuple[List[float], int]:
    """sr_autotune over the candidates: mean ms per candidate and the best index."""
    L = _abi.load()
    ff = grid.form_factor(0.0) if (comp == "cartesian_complex" and grid.mode == "far") else None
    base = _opts(comp, precision, engine.Tuning(), ff)
    arr = (_abi.SrOpts * len(cands))(*[_opts(comp, precision, c, ff) for c in cands])
    tr, gr = batch.abi(), grid.abi()
    nbytes = ctypes.c_size_t(0)
    _abi.check(L.sr_autotune_workspace_bytes(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(base), arr,
                                             len(cands), ctypes.byref(nbytes)))
    ws = torch.empty(max(int(nbytes.value), 1), dtype=torch.uint8, device=grid.device)
    n_out = L.sr_num_outputs(_abi.MODES[grid.mode], _abi.COMPS[comp])
    shape = (batch.host.n_snaps, grid.n_phi, grid.n1, grid.n_omega)
    outs = [torch.zeros(shape, dtype=torch.float64, device=grid.device) for _ in range(n_out)]
    ptrs = (ctypes.c_void_p * n_out)(*[o.data_ptr() for o in outs])
    ms = (ctypes.c_float * len(cands))()
    best = ctypes.c_int(-1)
    st = torch.cuda.current_stream(grid.device)
    _abi.check(L.sr_autotune(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(base), arr, len(cands), ptrs,
                             ctypes.c_void_p(ws.data_ptr()), nbytes.value, int(warmup), int(reps),
                             ctypes.c_void_p(st.cuda_stream), ms, ctypes.byref(best)))
    return [float(v) for v in ms], int(best.value)


def tune(batch: engine.DeviceBatch, grid: engine.DeviceGrid, comp: str = "total",
         precision: str = "float", target_evals: float = 3e10, verbose: bool = False
         ) -> Tuple[engine.Tuning, List[dict]]:
    """Best plan for this batch and grid, and the timings it was chosen from."""
    key = _key(batch, grid, comp, precision)
    path = os.environ.get("SYNCHRAD_AMD_TUNE_CACHE")
    if key not in _cache and path and os.path.exists(path):
        try:
            with open(path) as f:
                _cache.update(json.load(f))
        except (OSError, ValueError):
            pass
    if key in _cache:
        return engine.Tuning(**_cache[key]["plan"]), _cache[key]["timings"]
    cands = candidates(grid.mode, precision, bool(grid.omega_step), grid.n_omega)
    rows = []

    def record(stage, sb, sg, timed, ms):
        evals = sg.n_grid * sb.host.n_intervals()
        for tun, t in zip(timed, ms):
            rows.append({"stage": stage, "plan": asdict(tun), "ms": t,
                         "Gevals_per_s": evals / (t * 1e-3) / 1e9 if np.isfinite(t) and t > 0 else 0.0})
            if verbose:
                print(f"autotune {stage} {asdict(tun)}: {t:.3f} ms", flush=True)

    if is_large(batch, grid):
        # stage 1: full-length tracks on a subset of the directions, each plan
        # at the particle-slice count of the full launch (every tile carries
        # its full-launch work, only fewer tiles run)
        sg = direction_subset(grid)
        timed = [replace(c, slices=planned(batch, grid, comp, precision, c).slices) for c in cands]
        ms, best = time_candidates(batch, sg, comp, precision, timed)
        record("directions subset", batch, sg, timed, ms)
        # stage 2: the subset's pick against the library's own plan, once each
        # on the full launch (the subset measured a different ranking than the
        # full launch on the betatron beam, so it only nominates)
        lib = planned(batch, grid, comp, precision, engine.Tuning())
        lib_plan = engine.Tuning(sweep="recurrence" if lib.sweep == _abi.SR_SWEEP_RECURRENCE else "direct",
                                 omega_per_lane=lib.omega_per_lane, waves_per_block=lib.waves_per_block)
        finals = [lib_plan]
        pick = replace(cands[best], slices=0)
        if (pick.sweep, pick.omega_per_lane, pick.waves_per_block) != \
                (lib_plan.sweep, lib_plan.omega_per_lane, lib_plan.waves_per_block):
            finals.append(pick)
        if len(finals) > 1:
            ms2, best2 = time_candidates(batch, grid, comp, precision, finals, reps=1, warmup=0)
            record("full launch", batch, grid, finals, ms2)
            plan = finals[best2]
        else:
            plan = lib_plan
    else:
        sb = engine.DeviceBatch(sample_batch(batch.host, grid.n_grid, target_evals), grid.device)
        ms, best = time_candidates(sb, grid, comp, precision, cands)
        record("sample", sb, grid, cands, ms)
        # short launches (the 0.5 ms undulator spectrum): one launch per
        # candidate is within noise of the next, so time each ~20 ms again
        fastest = min((t for t in ms if np.isfinite(t) and t > 0), default=0.0)
        if len(cands) > 1 and 0.0 < fastest < SHORT_LAUNCH_MS:
            reps = int(min(MAX_REPS, np.ceil(SHORT_TIMING_MS / fastest)))
            ms, best = time_candidates(sb, grid, comp, precision, cands, reps=reps, warmup=1)
            record(f"sample x{reps}", sb, grid, cands, ms)
        # the launch gets the winning shape; its particle slices are the
        # library's choice for the launch
        plan = replace(cands[best], slices=0)
    _cache[key] = {"plan": asdict(plan), "timings": rows}
    if path:
        try:
            with open(path, "w") as f:
                json.dump(_cache, f)
        except OSError:
            pass
    return plan, rows
