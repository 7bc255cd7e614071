"""Autotuner for the tile kernel's launch shape.

For a (mode, comp, precision, grid shape) it times a set of candidate plans --
sweep (direct / recurrence), omegas per thread M and waves per block (which is
also the number of steps per LDS tile, 64 per wave) -- on a sample of the real
batch: up to 16 of its particles, truncated so that one candidate run costs
about ``target_evals`` point-step evaluations.  The grid is the real grid, so
the filter activity (which decides how much work each tile does) is the real
one.  Large far-field launches (>= LARGE_EVALS evaluations) launch their tiles
heaviest first and are planned with more particle slices; a sample that small
would be timed in another regime, so for them the sample keeps full-length
tracks of enough particles to be a large launch itself.  The fastest plan is
cached in memory and, if ``SYNCHRAD_AMD_TUNE_CACHE`` names a file, in that JSON
file.
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import engine

_cache: Dict[str, dict] = {}

# far-field launches from this size on are ordered and sliced by the library
# (its kOrderSliceEvals); their samples are at least LARGE_SAMPLE evaluations
LARGE_EVALS = 1e12
LARGE_SAMPLE = 1.2e12


def candidates(mode: str, precision: str, uniform: bool) -> List[engine.Tuning]:
    T = engine.Tuning
    if engine.PRECISIONS[precision] == 0:  # FP64: one or two omegas per thread, direct sweep
        return [T(sweep="direct", omega_per_lane=m, waves_per_block=w) for m in (1, 2) for w in (1, 2, 4)]
    ms = (1, 2, 4) if mode == "far" else (1, 2)
    out = [T(sweep="direct", omega_per_lane=m, waves_per_block=w) for m in ms for w in (1, 2, 4)]
    if uniform:
        # on linear grids the recurrence measured faster than every direct plan
        rms = (2, 4) if mode == "far" else (2,)
        out = [T(sweep="recurrence", omega_per_lane=m, waves_per_block=w) for m in rms for w in (1, 2, 4)]
    return out


def _key(batch: engine.DeviceBatch, grid: engine.DeviceGrid, comp: str, precision: str) -> str:
    hb = batch.host
    return json.dumps([grid.mode, comp, engine.PRECISIONS[precision], grid.n_omega, grid.n1, grid.n_phi,
                       bool(grid.omega_step), int(hb.n_snaps), int(min(hb.n_particles, 16))])


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


def time_plan(batch, grid, comp, precision, tun: engine.Tuning, reps: int = 2) -> float:
    outs = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=tun)
    st = torch.cuda.current_stream(grid.device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=tun, outs=outs)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def representative(batch: engine.DeviceBatch, grid: engine.DeviceGrid) -> bool:
    """Whether a tuning sample ranks the plans as the full launch would.  Not
    for large far-field launches: their tiles are ordered and sliced, and the
    per-tile work of the full launch is many times the sample's (on the
    betatron beam the sample's pick runs 8 % slower than the library's plan)."""
    return not (grid.mode == "far" and grid.n_grid * batch.host.n_intervals() >= LARGE_EVALS)


def tune(batch: engine.DeviceBatch, grid: engine.DeviceGrid, comp: str = "total",
         precision: str = "float", target_evals: float = 3e10, verbose: bool = False
         ) -> Tuple[engine.Tuning, List[dict]]:
    """Best plan for this batch and grid, and the timings it was chosen from."""
    key = _key(batch, grid, comp, precision)
    path = os.environ.get("SYNCHRAD_AMD_TUNE_CACHE")
    if key not in _cache and path and os.path.exists(path):
        try:
            _cache.update(json.load(open(path)))
        except (OSError, ValueError):
            pass
    if key in _cache:
        return engine.Tuning(**_cache[key]["plan"]), _cache[key]["timings"]
    full = grid.n_grid * batch.host.n_intervals()
    if grid.mode == "far" and full >= LARGE_EVALS:
        # full-length tracks of as many particles as a large launch needs
        lengths = np.diff(batch.host.offsets)
        mean = max(float(lengths.mean()), 1.0) if lengths.size else 1.0
        need = int(np.ceil(LARGE_SAMPLE / (grid.n_grid * mean)))
        hb = sample_batch(batch.host, grid.n_grid, max(target_evals, LARGE_SAMPLE * 4),
                          max_particles=max(16, need))
    else:
        hb = sample_batch(batch.host, grid.n_grid, target_evals)
    sb = engine.DeviceBatch(hb, grid.device)
    evals = grid.n_grid * hb.n_intervals()
    rows = []
    best, best_ms = None, np.inf
    for tun in candidates(grid.mode, precision, bool(grid.omega_step)):
        ms = time_plan(sb, grid, comp, precision, tun)
        rows.append({"plan": asdict(tun), "ms": ms, "Gevals_per_s": evals / (ms * 1e-3) / 1e9})
        if verbose:
            print(f"autotune {asdict(tun)}: {ms:.3f} ms", flush=True)
        if ms < best_ms:
            best, best_ms = tun, ms
    _cache[key] = {"plan": asdict(best), "timings": rows}
    if path:
        try:
            with open(path, "w") as f:
                json.dump(_cache, f)
        except OSError:
            pass
    return best, rows
