"""The C ABI library loads, exports what include/synchrad_amd.h declares, and
validates its arguments -- all without a GPU (no call here reaches a launch)."""
import ctypes
import os
import re

import pytest

from synchrad_amd import _abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "synchrad_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _abi.load()
    names = declared_functions()
    assert "sr_spectrum_batch" in names and "sr_far_total" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert set(names) == set(_abi.EXPORTS)


def test_version_and_output_counts():
    L = _abi.load()
    assert L.sr_abi_version() == _abi.SR_ABI_VERSION
    far, near = _abi.SR_MODE_FAR, _abi.SR_MODE_NEAR
    assert L.sr_num_outputs(far, _abi.SR_COMP_TOTAL) == 1
    assert L.sr_num_outputs(far, _abi.SR_COMP_CARTESIAN) == 3
    assert L.sr_num_outputs(far, _abi.SR_COMP_CARTESIAN_COMPLEX) == 6
    assert L.sr_num_outputs(far, _abi.SR_COMP_SPHERIC) == 3
    assert L.sr_num_outputs(far, _abi.SR_COMP_SPHERIC_COMPLEX) == 6
    assert L.sr_num_outputs(near, _abi.SR_COMP_TOTAL) == 1
    assert L.sr_num_outputs(near, _abi.SR_COMP_CARTESIAN) == 3
    assert L.sr_num_outputs(near, _abi.SR_COMP_CARTESIAN_COMPLEX) == 6
    assert L.sr_num_outputs(near, _abi.SR_COMP_SPHERIC) == _abi.SR_ERR_INVALID
    assert b"not available" in L.sr_last_error()


FAKE = 0x1000  # never dereferenced: these calls stop at validation or planning


def _grid(**kw):
    g = _abi.SrGrid()
    g.mode = _abi.SR_MODE_FAR
    g.n_omega, g.n_theta, g.n_phi = 512, 64, 64
    for f in ("omega", "sin_theta", "cos_theta", "sin_phi", "cos_phi"):
        setattr(g, f, FAKE)
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def _tracks(n_particles=4, **kw):
    t = _abi.SrTracks()
    t.n_particles = n_particles
    t.n_points = 100 * n_particles
    for f in ("offsets", "x", "y", "z", "ux", "uy", "uz", "weight", "it_start", "it_end", "it_snaps"):
        setattr(t, f, FAKE)
    t.n_snaps = 1
    t.snaps_stride = 1
    t.dt = 1e-3
    for k, v in kw.items():
        setattr(t, k, v)
    return t


def _opts(**kw):
    o = _abi.SrOpts()
    o.comp = _abi.SR_COMP_TOTAL
    o.precision = _abi.SR_PREC_SPLIT
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _ws(t, g, o):
    L = _abi.load()
    n = ctypes.c_size_t(123)
    rc = L.sr_workspace_bytes(ctypes.byref(t), ctypes.byref(g), ctypes.byref(o), ctypes.byref(n))
    return rc, n.value


def _err():
    return _abi.load().sr_last_error().decode()


@pytest.mark.parametrize("case,msg", [
    (dict(grid=dict(n_omega=0)), "sizes"),
    (dict(grid=dict(mode=7)), "mode"),
    (dict(grid=dict(sin_theta=None)), "sin_theta"),
    (dict(opts=dict(comp=9)), "comp"),
    (dict(opts=dict(precision=5)), "precision"),
    (dict(opts=dict(waves_per_block=3)), "waves_per_block"),
    (dict(opts=dict(omega_per_lane=3)), "omega_per_lane"),
    (dict(opts=dict(omega_per_lane=8)), "omega_per_lane"),
    (dict(opts=dict(sweep=5)), "sweep"),
    (dict(opts=dict(sweep=2)), "recurrence"),
    (dict(grid=dict(omega_step=float("inf"))), "omega_step"),
    (dict(opts=dict(slices=99)), "slices"),
    (dict(opts=dict(comp=_abi.SR_COMP_CARTESIAN_COMPLEX)), "form_factor"),
    (dict(tracks=dict(n_snaps=0)), "n_snaps"),
    (dict(tracks=dict(snaps_stride=5)), "snaps_stride"),
    (dict(tracks=dict(dt=0.0)), "dt"),
    (dict(tracks=dict(x=None)), "NULL"),
    (dict(opts=dict(tile_order=2)), "tile_order"),
])
def test_validation_errors(case, msg):
    g = _grid(**case.get("grid", {}))
    t = _tracks(**case.get("tracks", {}))
    o = _opts(**case.get("opts", {}))
    rc, n = _ws(t, g, o)
    assert rc == _abi.SR_ERR_INVALID and n == 0
    assert msg in _err()
    L = _abi.load()
    outs = (ctypes.c_void_p * 6)(*([FAKE] * 6))
    rc = L.sr_spectrum_batch(ctypes.byref(t), ctypes.byref(g), ctypes.byref(o), outs, None, 0, None)
    assert rc == _abi.SR_ERR_INVALID


def test_empty_batch_is_a_no_op():
    L = _abi.load()
    t = _tracks(n_particles=0)
    outs = (ctypes.c_void_p * 1)(FAKE)
    rc = L.sr_spectrum_batch(ctypes.byref(t), ctypes.byref(_grid()), ctypes.byref(_opts()), outs,
                             None, 0, None)
    assert rc == _abi.SR_OK


def test_workspace_plan():
    # every batch holds the direction-independent step data: 6 doubles per
    # sample in the far field (a, bbar), 3 in the near field (b)
    t = _tracks(n_particles=2000)
    pre = 6 * 8 * t.n_points
    # large grid, one snapshot: no particle slices; the step data, and a cost
    # and a launch index per tile (4096 tiles, launched heaviest first)
    rc, n = _ws(t, _grid(), _opts())
    assert rc == 0 and n == pre + 8 * 4096
    # index order: only the step data
    rc, n = _ws(t, _grid(), _opts(tile_order=1))
    assert rc == 0 and n == pre
    # the near field is not reordered (no filter-dependent work per tile)
    rc, n = _ws(t, _grid(mode=_abi.SR_MODE_NEAR, radius=FAKE, l_screen=1.0), _opts())
    assert rc == 0 and n == 3 * 8 * t.n_points
    # small grid: particle slices add n_slices * n_out * n_grid doubles
    g = _grid(n_omega=128, n_theta=32, n_phi=32)
    t256 = _tracks(n_particles=256)
    pre256 = 6 * 8 * t256.n_points
    rc, n = _ws(t256, g, _opts())
    assert rc == 0 and n > pre256 and (n - pre256) % (8 * 128 * 32 * 32) == 0
    # several snapshots: the slices hold every snapshot
    rc, n1 = _ws(t256, g, _opts())
    rc, n3 = _ws(_tracks(n_particles=256, n_snaps=3, snaps_stride=3), g, _opts())
    assert rc == 0 and n3 - pre256 == 3 * (n1 - pre256)
    # explicit slices
    rc, n = _ws(t256, g, _opts(slices=3, comp=_abi.SR_COMP_CARTESIAN))
    assert rc == 0 and n == 3 * 3 * 128 * 32 * 32 * 8 + pre256
    # empty batch: nothing
    rc, n = _ws(_tracks(n_particles=0), g, _opts())
    assert rc == 0 and n == 0


def test_large_far_launch_plan():
    """A far-field batch of >= 1e12 evaluations is planned with particle slices
    up to ~64K waves, as far as its tiles stay orderable, and launches them
    heaviest first: 1e6 samples on 512 x 64 x 64 (2.1e12 evaluations), direct
    sweep M = 2 x 4 waves (4096 tiles per slice) -> 4 slices, 16384 tiles."""
    t = _tracks(n_particles=2000)
    t.n_points = 1_000_000
    n_grid = 512 * 64 * 64
    pre = 6 * 8 * t.n_points
    rc, n = _ws(t, _grid(), _opts())
    assert rc == 0 and n == pre + 4 * n_grid * 8 + 8 * 16384
    # index order: the small-launch plan (no slices on this grid), no order
    rc, n = _ws(t, _grid(), _opts(tile_order=1))
    assert rc == 0 and n == pre
    # more evaluations would allow more slices, but the waves target stops at 4
    t.n_points = 10_000_000
    rc, n = _ws(t, _grid(), _opts())
    assert rc == 0 and n == 6 * 8 * t.n_points + 4 * n_grid * 8 + 8 * 16384


def test_per_particle_entry_points_validate():
    L = _abi.load()
    P = FAKE
    rc = L.sr_far_total(P, P, P, P, P, P, P, 1.0, 0, 0, 10, P, P, P, P, P, 8, 4, 4, 1e-3, 1, P,
                        _abi.SR_PREC_F64, None)
    assert rc == _abi.SR_ERR_INVALID and "itEnd" in _err()
    rc = L.sr_near_cartesian(P, P, P, P, P, P, P, P, P, 1.0, 0, 10, 0, P, P, P, P, 5.0, 8, 4, 4,
                             1e-3, 1, P, _abi.SR_PREC_F64, None)
    assert rc == _abi.SR_ERR_INVALID and "nSteps" in _err()
    rc = L.sr_far_cartesian(P, P, P, P, P, P, P, P, P, 1.0, 0, 10, 10, P, P, P, P, P, 8, 4, 4, 1e-3,
                            1, P, 7, None)
    assert rc == _abi.SR_ERR_INVALID and "precision" in _err()
