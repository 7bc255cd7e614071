"""The library's plan query and autotuner through the C ABI (sr_plan,
sr_autotune), and the Python autotuner's samples built on them."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle
from synchrad_amd import _abi, autotune, engine, workloads

pytestmark = pytest.mark.gpu


def _objects(n_particles=3, n_steps=2000, grid=(128, 8, 8)):
    wl = workloads.config2(n_particles=n_particles, n_steps=n_steps, grid=grid)
    ax = wl.axes
    return wl, engine.DeviceBatch(wl.batch), engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])


def test_plan_query_matches_the_launch():
    """sr_plan reports the shape sr_spectrum_batch runs: a forced shape comes
    back as given, and the workspace of the reported slices is what
    sr_workspace_bytes asks for."""
    wl, batch, grid = _objects()
    info = autotune.planned(batch, grid, "total", "float", engine.Tuning())
    assert info.sweep == _abi.SR_SWEEP_RECURRENCE and info.tiles > 0
    forced = autotune.planned(batch, grid, "total", "float",
                              engine.Tuning(sweep="direct", omega_per_lane=1, waves_per_block=2, slices=3))
    assert (forced.sweep, forced.omega_per_lane, forced.waves_per_block, forced.slices) == \
        (_abi.SR_SWEEP_DIRECT, 1, 2, 3)
    assert forced.tiles == 1 * 64 * 3            # 128 omega / (128 threads x 1) chunks x 64 directions x 3
    n_grid = grid.n_grid
    assert engine.workspace_bytes(batch, grid, "total", "float", engine.Tuning(slices=3, sweep="direct",
                                                                              omega_per_lane=1,
                                                                              waves_per_block=2)) == \
        3 * n_grid * 8 + 6 * batch.host.n_points * 8


def test_autotune_entry_point_times_every_runnable_candidate():
    wl, batch, grid = _objects()
    cands = [engine.Tuning(sweep="recurrence", omega_per_lane=2, waves_per_block=1),
             engine.Tuning(sweep="direct", omega_per_lane=2, waves_per_block=2),
             engine.Tuning(sweep="recurrence", omega_per_lane=4, waves_per_block=2)]
    ms, best = autotune.time_candidates(batch, grid, "total", "float", cands, reps=2)
    assert all(np.isfinite(m) and m > 0 for m in ms)
    assert ms[best] == min(ms)
    # a log grid has no constant step: the recurrence candidates cannot run it
    ax = dict(wl.axes)
    ax["omega"] = np.geomspace(ax["omega"][0], ax["omega"][-1], ax["omega"].size)
    log_grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    ms, best = autotune.time_candidates(batch, log_grid, "total", "float", cands, reps=1)
    assert not np.isfinite(ms[0]) and not np.isfinite(ms[2]) and best == 1


def test_large_launch_is_tuned_on_a_direction_subset():
    """A far-field launch above 1e12 evaluations is tuned on full-length
    tracks over a subset of the directions; the subset keeps the full grid's
    own sin/cos values, and the winning plan runs the full launch correctly."""
    wl = workloads.config2(n_particles=256, n_steps=8000, grid=(512, 32, 32))
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(wl.batch)
    assert autotune.is_large(batch, grid)
    sub = autotune.direction_subset(grid)
    assert sub.n1 * sub.n_phi * 6 <= grid.n1 * grid.n_phi <= sub.n1 * sub.n_phi * 10
    assert torch.equal(sub.sin_theta, grid.sin_theta[::2]) and torch.equal(sub.cos_phi, grid.cos_phi[::4])
    plan, rows = autotune.tune(batch, grid, comp="total", precision="float")
    assert plan.sweep == "recurrence" and plan.slices == 0
    first = [r for r in rows if r["stage"] == "directions subset"]
    assert len(first) >= 4 and all(r["plan"]["slices"] >= 1 for r in first)
    final = [r for r in rows if r["stage"] == "full launch"]
    if final:   # the subset's pick against the library's plan on the full launch
        assert len(final) == 2 and (plan.omega_per_lane, plan.waves_per_block) == min(
            ((r["plan"]["omega_per_lane"], r["plan"]["waves_per_block"]) for r in final),
            key=lambda k: next(r["ms"] for r in final if (r["plan"]["omega_per_lane"],
                                                         r["plan"]["waves_per_block"]) == k))
    print(f"large-launch autotune: {plan} from {[(r['plan']['omega_per_lane'], r['plan']['waves_per_block'], round(r['ms'], 1)) for r in rows]}",
          flush=True)
    tuned = engine.spectrum(batch, grid, comp="total", precision="float", tuning=plan)[0]
    lib = engine.spectrum(batch, grid, comp="total", precision="float")[0]
    rel = float(torch.linalg.norm(tuned - lib) / torch.linalg.norm(lib))
    assert rel < 1e-6
