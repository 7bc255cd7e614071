"""sr_spectrum_batch_noalloc against sr_spectrum_batch, through the C ABI.

The no-allocation entry point keeps its workspace in a static device array and
computes the step data inside the tile kernel; with the same plan its spectra
equal the workspace entry point's.  Two calls in a row, the second with a
smaller batch and grid, so that the scratch array's leftovers from the first
call must not leak into the second.
"""
import ctypes

import numpy as np
import pytest
import torch

from synchrad_amd import _abi, engine, workloads

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
SPLIT_TOL = 1e-4


def _objects(n_particles, n_steps, grid_shape, mode="far"):
    if mode == "far":
        wl = workloads.config2(n_particles=n_particles, n_steps=n_steps, grid=grid_shape)
        ax1 = wl.axes["theta"]
    else:
        wl = workloads.config4(scale=0.125, n_particles=n_particles)
        ax1 = wl.axes["radius"]
    grid = engine.DeviceGrid(wl.mode, wl.axes["omega"], ax1, wl.axes["phi"], L_screen=wl.L_screen)
    return engine.DeviceBatch(wl.batch), grid


def _noalloc(batch, grid, comp, precision, slices=0):
    L = _abi.load()
    n_out = L.sr_num_outputs(_abi.MODES[grid.mode], _abi.COMPS[comp])
    shape = (batch.host.n_snaps, grid.n_phi, grid.n1, grid.n_omega)
    # stale values in the outputs' memory must not matter: the call adds to
    # zeroed outputs, as every entry point does
    outs = [torch.zeros(shape, dtype=torch.float64, device=grid.device) for _ in range(n_out)]
    opts = _abi.SrOpts()
    opts.comp = _abi.COMPS[comp]
    opts.precision = engine.PRECISIONS[precision]
    opts.slices = slices
    tr, gr = batch.abi(), grid.abi()
    ptrs = (ctypes.c_void_p * n_out)(*[o.data_ptr() for o in outs])
    st = torch.cuda.current_stream(grid.device)
    rc = L.sr_spectrum_batch_noalloc(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts), ptrs,
                                     ctypes.c_void_p(st.cuda_stream))
    return rc, outs


def _rel(a, b):
    num = sum(float(torch.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(torch.sum(y ** 2)) for y in b)
    return (num / den) ** 0.5


@pytest.mark.parametrize("precision", ["double", "float"])
def test_noalloc_matches_workspace_entry_point_twice(precision):
    tol = F64_TOL if precision == "double" else SPLIT_TOL
    cases = [(_objects(6, 3000, (128, 16, 16)), "total", 2),      # particle slices: partials in scratch
             (_objects(3, 1200, (64, 8, 8)), "cartesian", 0)]      # smaller, tuned plan
    for (batch, grid), comp, slices in cases:
        print(f"noalloc {precision} {comp}: {batch.host.n_particles} particles, grid {grid.n_grid}",
              flush=True)
        ref = engine.spectrum(batch, grid, comp=comp, precision=precision,
                              tuning=engine.Tuning(slices=slices))
        rc, got = _noalloc(batch, grid, comp, precision, slices=slices)
        _abi.check(rc)
        torch.cuda.synchronize()
        r = _rel(got, ref)
        print(f"  rel-L2 vs sr_spectrum_batch {r:.3e}", flush=True)
        assert r <= tol
        # same plan, same kernels: the step data computed in the stage is the
        # pre-pass's own expressions
        for a, b in zip(got, ref):
            assert torch.equal(a, b)


def test_noalloc_near_field():
    batch, grid = _objects(2, 0, None, mode="near")
    ref = engine.spectrum(batch, grid, comp="total", precision="float")
    rc, got = _noalloc(batch, grid, "total", "float")
    _abi.check(rc)
    torch.cuda.synchronize()
    assert _rel(got, ref) <= SPLIT_TOL


def test_noalloc_rejects_requests_beyond_its_scratch():
    """32 explicit slices of a 512 x 64 x 64 grid need 32 x 16.8 MB of partial
    spectra, more than the 64 MiB array: a clear error, nothing launched."""
    batch, grid = _objects(40, 200, (512, 64, 64))
    rc, outs = _noalloc(batch, grid, "total", "float", slices=32)
    assert rc == _abi.SR_ERR_WORKSPACE
    msg = _abi.load().sr_last_error().decode()
    assert "scratch" in msg and str(_abi.SR_NOALLOC_SCRATCH_BYTES) in msg
    torch.cuda.synchronize()
    assert all(float(o.abs().max()) == 0.0 for o in outs)
