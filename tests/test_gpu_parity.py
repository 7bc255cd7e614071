"""Parity of the HIP kernels (through the C ABI) with the CPU FP64 oracle.

Bars: FP64 path <= 1e-10 relative L2 (and a max-element check), split path
<= 1e-4 relative L2, both against the FP64 oracle on the same inputs.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle
from synchrad_amd import _abi, engine, workloads

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
SPLIT_TOL = 1e-4


def rel_l2(got, ref):
    d = np.linalg.norm((got - ref).ravel())
    n = np.linalg.norm(ref.ravel())
    return d / n if n > 0 else d


def max_rel(got, ref, floor=1e-6):
    m = np.abs(ref).max()
    sel = np.abs(ref) > floor * m
    return float(np.max(np.abs(got[sel] - ref[sel]) / np.abs(ref[sel]))) if sel.any() else 0.0


def gpu_spectrum(tracks, axes, dt, mode="far", comp="total", precision="double", it_range=None,
                 nSnaps=1, L_screen=None, sigma=0.0, tuning=None):
    hb = engine.pack_tracks(tracks, dt=dt, it_range=it_range, nSnaps=nSnaps)
    batch = engine.DeviceBatch(hb)
    ax1 = axes["theta"] if mode == "far" else axes["radius"]
    grid = engine.DeviceGrid(mode, axes["omega"], ax1, axes["phi"], L_screen=L_screen)
    outs = engine.spectrum(batch, grid, comp=comp, precision=precision, sigma_particle=sigma,
                           tuning=tuning)
    torch.cuda.synchronize()
    keys = engine.COMP_KEYS[comp]
    return {k: np.ascontiguousarray(o.cpu().numpy().swapaxes(-1, -3)) for k, o in zip(keys, outs)}


def check(got, ref, precision, label=""):
    for k in ref:
        r = rel_l2(got[k], ref[k])
        print(f"{label} {k} {precision}: rel-L2 {r:.3e} max-rel {max_rel(got[k], ref[k]):.3e}", flush=True)
        if precision == "double":
            assert r <= F64_TOL, (k, r)
        else:
            assert r <= SPLIT_TOL, (k, r)


@pytest.fixture(scope="module")
def undulator():
    wl = workloads.config1()
    tr = wl.track_list()
    ref = oracle.spectrum("far", "total", tr, wl.axes, wl.batch.dt)
    return wl, tr, ref


@pytest.mark.parametrize("precision", ["double", "float"])
def test_far_total_undulator_full_config(undulator, precision):
    """Single-electron undulator, 256 x 32 x 32 grid, 1600 steps."""
    wl, tr, ref = undulator
    got = gpu_spectrum(tr, wl.axes, wl.batch.dt, precision=precision)
    check(got, ref, precision, "config1")
    if precision == "double":
        assert max_rel(got["total"], ref["total"]) < 1e-8


@pytest.fixture(scope="module")
def betatron4():
    wl = workloads.config2(n_particles=4, grid=(64, 16, 16))
    tr = wl.track_list()
    ref = oracle.spectrum("far", "total", tr, wl.axes, wl.batch.dt)
    return wl, tr, ref


@pytest.mark.parametrize("precision,sweep,M", [("double", "auto", 0), ("float", "direct", 1),
                                               ("float", "direct", 2), ("float", "recurrence", 2),
                                               ("float", "direct", 4), ("float", "recurrence", 4)])
def test_far_total_betatron_subset(betatron4, precision, sweep, M):
    wl, tr, ref = betatron4
    got = gpu_spectrum(tr, wl.axes, wl.batch.dt, precision=precision,
                       tuning=engine.Tuning(sweep=sweep, omega_per_lane=M))
    check(got, ref, precision, f"betatron-4e {sweep} M={M}")


@pytest.mark.parametrize("order", ["descending", "through_zero", "log"])
@pytest.mark.parametrize("sweep,M", [("direct", 2), ("direct", 4), ("recurrence", 2), ("recurrence", 4)])
def test_omega_order(order, sweep, M):
    """The filter compaction of the direct sweep assumes nothing of the omega
    order; the recurrence sweep sorts steps only for runs of non-negative,
    increasing omegas and keeps its masked form otherwise, and is refused
    for grids without a constant step (log grids)."""
    wl = workloads.config2(n_particles=2, n_steps=1500, grid=(96, 8, 8))
    ax = dict(wl.axes)
    om = ax["omega"]
    if order == "descending":
        ax["omega"] = om[::-1].copy()
    elif order == "through_zero":
        ax["omega"] = np.linspace(-0.3 * om.max(), om.max(), om.size)
    else:
        ax["omega"] = np.geomspace(om.min(), om.max(), om.size)
    tr = wl.track_list()
    tun = engine.Tuning(sweep=sweep, omega_per_lane=M)
    if order == "log" and sweep == "recurrence":
        with pytest.raises(_abi.SynchradError, match="constant omega_step"):
            gpu_spectrum(tr, ax, wl.batch.dt, precision="float", tuning=tun)
        return
    ref = oracle.spectrum("far", "total", tr, ax, wl.batch.dt)
    got = gpu_spectrum(tr, ax, wl.batch.dt, precision="float", tuning=tun)
    check(got, ref, "float", f"omega {order} {sweep} M={M}")


@pytest.mark.parametrize("comp", ["cartesian", "spheric", "cartesian_complex", "spheric_complex"])
@pytest.mark.parametrize("precision", ["double", "float"])
def test_far_components(comp, precision):
    wl = workloads.config2(n_particles=3, n_steps=2000, grid=(48, 8, 8))
    tr = wl.track_list()
    sigma = 2e-11 if comp == "cartesian_complex" else 0.0
    ref = oracle.spectrum("far", comp, tr, wl.axes, wl.batch.dt, sigma_particle=sigma)
    got = gpu_spectrum(tr, wl.axes, wl.batch.dt, comp=comp, precision=precision, sigma=sigma)
    if comp.endswith("complex") or comp == "spheric":
        # components that are ~0 everywhere (the far-field radial part) are
        # checked against the scale of the whole result instead of their own
        scale = max(np.linalg.norm(ref[k]) for k in ref)
        for k in ref:
            r = np.linalg.norm(got[k] - ref[k]) / scale
            print(f"{comp} {k} {precision}: err/scale {r:.3e}", flush=True)
            assert r <= (F64_TOL if precision == "double" else SPLIT_TOL)
    else:
        check(got, ref, precision, comp)


@pytest.mark.parametrize("comp", ["total", "cartesian", "cartesian_complex"])
@pytest.mark.parametrize("precision,sweep", [("double", "auto"), ("float", "direct"),
                                             ("float", "recurrence")])
def test_near_field(comp, precision, sweep):
    wl = workloads.config4(scale=0.25, n_particles=3)
    tr = wl.track_list()
    ref = oracle.spectrum("near", comp, tr, wl.axes, wl.batch.dt, L_screen=wl.L_screen)
    got = gpu_spectrum(tr, wl.axes, wl.batch.dt, mode="near", comp=comp, precision=precision,
                       L_screen=wl.L_screen, tuning=engine.Tuning(sweep=sweep))
    if comp == "cartesian_complex":
        scale = max(np.linalg.norm(ref[k]) for k in ref)
        for k in ref:
            r = np.linalg.norm(got[k] - ref[k]) / scale
            print(f"near {comp} {k} {precision}: err/scale {r:.3e}", flush=True)
            assert r <= (F64_TOL if precision == "double" else SPLIT_TOL)
    else:
        check(got, ref, precision, "near " + comp)


def _ragged_tracks(n=7, seed=3):
    """Short undulator tracks of different lengths, it_start offsets, one
    single-sample track and one empty track."""
    rng = np.random.default_rng(seed)
    tr = []
    lengths = [300, 1, 257, 0, 64, 513, 128][:n]
    for i, L in enumerate(lengths):
        g = rng.uniform(40, 80)
        arrs = workloads.planar_undulator(g, rng.uniform(0.5, 1.5), 0.01, max(L, 0), 0.01 / 24,
                                          psi=rng.uniform(0, 6.28))
        tr.append(list(arrs) + [float(rng.uniform(0.5, 1.5)), int(rng.integers(0, 200))])
    return tr


@pytest.mark.parametrize("precision,sweep", [("double", "auto"), ("float", "direct"),
                                             ("float", "recurrence")])
@pytest.mark.parametrize("it_range,nSnaps", [(None, 1), (None, 3), ((0, 700), 4), ((100, 650), 3)])
@pytest.mark.parametrize("slices", [0, 3])
def test_ragged_batch_snapshots(precision, sweep, it_range, nSnaps, slices):
    tr = _ragged_tracks()
    axes = oracle.make_axes([(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (40, 6, 5)], "far")
    dt = 0.01 / 24
    ref = oracle.spectrum("far", "total", tr, axes, dt, it_range=it_range, nSnaps=nSnaps)
    got = gpu_spectrum(tr, axes, dt, precision=precision, it_range=it_range, nSnaps=nSnaps,
                       tuning=engine.Tuning(sweep=sweep, slices=slices))
    check(got, ref, precision, f"ragged it_range={it_range} nSnaps={nSnaps} slices={slices}")


def test_f64_phases_beyond_table_range():
    """FP64 phases of 1.26e13..2.51e13 rad (tracks starting at iteration 2e9 of
    a global window, so t is ~1.25e6 m), on both sides of 2^44 = 1.76e13 rad, where
    the FP64 sin/cos switches from its direct table reduction to a 2 pi
    pre-reduction; the steps' phase increments stay below pi, so the filter
    keeps them.  Against the oracle's libm sin/cos of the same FP64 phases."""
    dt = 0.01 / 16
    tr = []
    for i, (g, K) in enumerate([(100.0, 1.0), (120.0, 0.8)]):
        arrs = workloads.planar_undulator(g, K, 0.01, 400, dt, psi=0.3 * i)
        tr.append(list(arrs) + [1.0 + 0.5 * i, 2_000_000_000 + 37 * i])
    axes = oracle.make_axes([(1.6e6, 3.2e6), (0.0, 0.01), (0.0, 2 * np.pi), (48, 4, 4)], "far")
    it_range = (2_000_000_000, 2_000_000_500)
    w = 2 * np.pi * axes["omega"]
    t0 = it_range[0] * dt
    assert w.min() * t0 < 2.0 ** 44 < w.max() * t0
    ref = oracle.spectrum("far", "total", tr, axes, dt, it_range=it_range)
    assert (ref["total"] > 0).all()
    got = gpu_spectrum(tr, axes, dt, precision="double", it_range=it_range)
    check(got, ref, "double", "phases beyond 2^44")


@pytest.mark.parametrize("precision", ["double", "float"])
def test_uneven_csr_slices_agree(precision):
    wl = workloads.config3(scale=0.25, n_particles=40)
    tr = wl.track_list()
    ref = oracle.spectrum("far", "cartesian", tr, wl.axes, wl.batch.dt)
    one = gpu_spectrum(tr, wl.axes, wl.batch.dt, comp="cartesian", precision=precision,
                       tuning=engine.Tuning(slices=1))
    check(one, ref, precision, "uneven slices=1")
    # 4 and the largest count, 32 (up to 2 particles per slice): the slices'
    # partial spectra are summed in slice order, so only FP64 regrouping differs
    for n in (4, 32):
        many = gpu_spectrum(tr, wl.axes, wl.batch.dt, comp="cartesian", precision=precision,
                            tuning=engine.Tuning(slices=n))
        for k in one:
            assert rel_l2(many[k], one[k]) < 1e-13, (n, k)


@pytest.mark.parametrize("tun", [engine.Tuning(waves_per_block=1), engine.Tuning(waves_per_block=2),
                                 engine.Tuning(omega_per_lane=1, waves_per_block=4),
                                 engine.Tuning(omega_per_lane=2, waves_per_block=2),
                                 engine.Tuning(omega_per_lane=4, waves_per_block=1),
                                 engine.Tuning(sweep="recurrence", omega_per_lane=4, waves_per_block=2),
                                 engine.Tuning(persistent=37)])
def test_tile_shapes_agree(tun):
    wl = workloads.config2(n_particles=2, n_steps=1500, grid=(200, 6, 4))
    tr = wl.track_list()
    ref = oracle.spectrum("far", "total", tr, wl.axes, wl.batch.dt)
    for prec in ("double", "float"):
        if prec == "double" and tun.sweep == "recurrence":
            continue    # split-only sweep: FP64 requests for it are refused
        got = gpu_spectrum(tr, wl.axes, wl.batch.dt, precision=prec, tuning=tun)
        check(got, ref, prec, f"tile {tun}")


def test_per_particle_entry_points_match_batch():
    """sr_far_total / sr_far_cartesian / sr_near_total (one call per particle,
    the reference kernel ABI) add up to the batched result."""
    L = _abi.load()
    wl = workloads.config1(scale=0.125)
    tr = _ragged_tracks(5, seed=9)
    axes = oracle.make_axes([(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (40, 6, 5)], "far")
    dt = 0.01 / 24
    dev = torch.device("cuda")
    f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    om, st, ct = f64(2 * np.pi * axes["omega"]), f64(np.sin(axes["theta"])), f64(np.cos(axes["theta"]))
    sp, cp = f64(np.sin(axes["phi"])), f64(np.cos(axes["phi"]))
    nO, nT, nP = axes["omega"].size, axes["theta"].size, axes["phi"].size
    for prec_name, prec in (("double", _abi.SR_PREC_F64), ("float", _abi.SR_PREC_SPLIT)):
        tot = torch.zeros((1, nP, nT, nO), dtype=torch.float64, device=dev)
        cart = [torch.zeros((1, nP, nT, nO), dtype=torch.float64, device=dev) for _ in range(3)]
        for t in tr:
            n = t[0].size
            if n == 0:
                continue
            d = [f64(a) for a in t[:6]]
            snaps = torch.as_tensor(oracle.snap_iterations((0, n), 1).view(np.int32), device=dev)
            args = [c.data_ptr() for c in d] + [float(t[6]), 0, n, n]
            ax = [om.data_ptr(), st.data_ptr(), ct.data_ptr(), sp.data_ptr(), cp.data_ptr(), nO, nT, nP,
                  dt, 1, snaps.data_ptr(), prec, None]
            _abi.check(L.sr_far_total(tot.data_ptr(), *args, *ax))
            _abi.check(L.sr_far_cartesian(*[c.data_ptr() for c in cart], *args, *ax))
        torch.cuda.synchronize()
        ref_t = gpu_spectrum(tr, axes, dt, precision=prec_name)["total"]
        ref_c = gpu_spectrum(tr, axes, dt, comp="cartesian", precision=prec_name)
        assert rel_l2(tot.cpu().numpy().swapaxes(-1, -3), ref_t) < 1e-13
        for k, c in zip("xyz", cart):
            assert rel_l2(c.cpu().numpy().swapaxes(-1, -3), ref_c[k]) < 1e-13


def test_near_per_particle_entry_point():
    L = _abi.load()
    wl = workloads.config4(scale=0.125, n_particles=2)
    tr = wl.track_list()
    axes = wl.axes
    dev = torch.device("cuda")
    f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    nO, nR, nP = axes["omega"].size, axes["radius"].size, axes["phi"].size
    ax = [f64(2 * np.pi * axes["omega"]), f64(axes["radius"]), f64(np.sin(axes["phi"])),
          f64(np.cos(axes["phi"]))]
    out = torch.zeros((1, nP, nR, nO), dtype=torch.float64, device=dev)
    for t in tr:
        n = t[0].size
        d = [f64(a) for a in t[:6]]
        snaps = torch.as_tensor(oracle.snap_iterations((0, n), 1).view(np.int32), device=dev)
        _abi.check(L.sr_near_total(out.data_ptr(), *[c.data_ptr() for c in d], float(t[6]), 0, n, n,
                                   *[a.data_ptr() for a in ax], wl.L_screen, nO, nR, nP, wl.batch.dt, 1,
                                   snaps.data_ptr(), _abi.SR_PREC_F64, None))
    torch.cuda.synchronize()
    ref = oracle.spectrum("near", "total", tr, axes, wl.batch.dt, L_screen=wl.L_screen)["total"]
    assert rel_l2(out.cpu().numpy().swapaxes(-1, -3), ref) <= F64_TOL


def test_empty_and_degenerate_batches():
    axes = oracle.make_axes([(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (16, 4, 4)], "far")
    dt = 0.01 / 24
    got = gpu_spectrum([], axes, dt)
    assert np.all(got["total"] == 0)
    z = np.zeros(1)
    got = gpu_spectrum([[z, z, z, z, z, z + 10.0, 1.0, 0]], axes, dt)   # one sample: no interval
    assert np.all(got["total"] == 0)


@pytest.mark.parametrize("comp,grid_shape", [("total", (256, 16, 8)), ("cartesian", (32, 4, 4)),
                                             ("cartesian_complex", (32, 4, 4))])
def test_graph_replays_match_eager(comp, grid_shape):
    """A captured spectrum replayed several times, with eager work and
    synchronisations in between, equals the eager result every time (the
    small grid makes the plan use particle slices: workspace clear + reduce)."""
    wl = workloads.config2(n_particles=6, n_steps=1200, grid=grid_shape)
    batch = engine.DeviceBatch(wl.batch)
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    eager = [o.clone() for o in engine.spectrum(batch, grid, comp=comp, precision="float",
                                                 sigma_particle=1e-11)]
    g = engine.SpectrumGraph(batch, grid, comp=comp, precision="float", sigma_particle=1e-11)
    for _ in range(4):
        outs = g.replay()
        torch.cuda.synchronize()
        for a, b in zip(outs, eager):
            assert torch.equal(a, b)
        engine.spectrum(batch, grid, comp=comp, precision="double")
        torch.cuda.synchronize()


def test_autotuned_plan_matches_oracle():
    from synchrad_amd import autotune
    wl = workloads.config2(n_particles=3, n_steps=2000, grid=(128, 8, 8))
    tr = wl.track_list()
    ref = oracle.spectrum("far", "total", tr, wl.axes, wl.batch.dt)
    batch = engine.DeviceBatch(wl.batch)
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    plan, rows = autotune.tune(batch, grid, comp="total", precision="float", target_evals=1e8)
    first = [r for r in rows if r["stage"] == "sample"]
    assert len(first) == len(autotune.candidates("far", "float", True, 128)) and all(r["ms"] > 0 for r in rows)
    out = engine.spectrum(batch, grid, comp="total", precision="float", tuning=plan)[0]
    got = out.cpu().numpy().swapaxes(-1, -3)
    assert rel_l2(got, ref["total"]) < SPLIT_TOL
    print(f"autotuned plan {plan}", flush=True)


def test_reference_test_script_runs_unchanged():
    """tests/test_undulator_analytic.py of the reference, through the drop-in
    import path, in double and then (after the reference's own re-init calls)
    in float; both energies within 3% of the analytic estimate."""
    from synchrad.calc import SynchRad
    from synchrad.utils import J_in_um
    K0, Periods, Np, g0 = 0.1, 50, 24, 100.0
    dg = 1e-4 * g0
    gg = g0 / (1.0 + K0 ** 2 / 2) ** 0.5
    k_res = 2 * gg ** 2
    dt = 1.0 / 32
    steps = int((Periods + 2) / dt) + 1

    def uxf(z):
        val = K0 * np.sin(2 * np.pi * z)
        val *= (z > 0) * (z < 1.5) * z / 1.5 + (z > 1.5)
        val *= (z > Periods - 1.5) * (z < Periods) * (Periods - z) / 1.5 + (z < Periods - 1.5)
        return val

    t = np.r_[-1:Periods + 1:steps * 1j]
    rng = np.random.default_rng(11)
    tracks = []
    for g0p in g0 + dg * rng.standard_normal(Np):
        ggp = g0p / (1.0 + K0 ** 2 / 2) ** 0.5
        vb = (1.0 - ggp ** -2) ** 0.5
        z = vb * t
        ux = uxf(z - 0.5 * dt)
        uz = (g0 ** 2 - 1 - ux ** 2) ** 0.5
        x = ux[0] / g0p * dt / 2 + np.cumsum(ux / g0p) * dt
        tracks.append([x, np.zeros_like(t), z, ux, np.zeros_like(t), uz, 1.0, 0])
    calc = SynchRad({"grid": [(0.02 * k_res, 1.1 * k_res), (0, 2.0 / g0), (0.0, 2 * np.pi), (128, 32, 32)]})
    e_th = Np * k_res * J_in_um * (7 * np.pi / 24) / 137.0 * K0 ** 2 * (1 + K0 ** 2 / 2) * Periods
    calc.calculate_spectrum(tracks.copy(), timeStep=dt, comp="total", Np_max=Np)
    e64 = calc.get_energy(lambda0_um=1)
    ref = oracle.spectrum("far", "total", tracks, {"omega": calc.Args["omega"], "theta": calc.Args["theta"],
                                                   "phi": calc.Args["phi"]}, dt)["total"]
    assert rel_l2(calc.Data["radiation"]["total"], ref) < F64_TOL
    calc.Args["dtype"] = "float"
    calc.Args["native"] = True
    calc._init_args(calc.Args)
    calc._init_data()
    calc._compile_kernels()
    calc.calculate_spectrum(tracks.copy(), timeStep=dt, comp="total", Np_max=Np)
    e32 = calc.get_energy(lambda0_um=1)
    for e in (e64, e32):
        print(f"deviation from the analytic estimate: {abs(e - e_th) / e_th:.4f}", flush=True)
        assert abs(e - e_th) / e_th < 0.03
    assert abs(e32 - e64) / e64 < 1e-6


@pytest.mark.parametrize("precision", ["float", "double"])
@pytest.mark.parametrize("comp,slices,n_particles", [("total", 1, 20), ("cartesian", 2, 80)])
def test_heaviest_tiles_first_changes_no_result(precision, comp, slices, n_particles):
    """Large far-field launches run their tiles in an estimated heaviest-first
    order; every tile writes only its own outputs, so the spectra equal the
    index-order launch bit for bit (2048 directions x 1 omega chunk x 1 or 2
    particle slices, 20 or 80 x 8000 steps: 2.1e10 and 8.4e10 evaluations,
    above the sizes where the order is used)."""
    wl = workloads.config2(n_particles=n_particles, n_steps=8000, grid=(64, 32, 64))
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(wl.batch)
    on = engine.Tuning(slices=slices)
    off = engine.Tuning(slices=slices, tile_order=1)
    assert engine.workspace_bytes(batch, grid, comp, precision, on) > \
        engine.workspace_bytes(batch, grid, comp, precision, off)
    a = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=on)
    b = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=off)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.parametrize("descending", [False, True])
def test_heaviest_tiles_first_partial_chunk_and_descending_grid(descending):
    """The order also covers grids whose last omega chunk is partial (300
    omegas over 256-omega chunks: 2048 directions x 2 chunks) and descending
    grids (the cost estimate uses the chunk's |omega| range); results equal
    the index-order launch bit for bit."""
    wl = workloads.config2(n_particles=20, n_steps=8000, grid=(300, 32, 64))
    ax = wl.axes
    om = ax["omega"][::-1].copy() if descending else ax["omega"]
    grid = engine.DeviceGrid("far", om, ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(wl.batch)
    on = engine.Tuning(slices=1, omega_per_lane=2, waves_per_block=2)
    off = engine.Tuning(slices=1, omega_per_lane=2, waves_per_block=2, tile_order=1)
    assert engine.workspace_bytes(batch, grid, "total", "float", on) > \
        engine.workspace_bytes(batch, grid, "total", "float", off)
    a = engine.spectrum(batch, grid, comp="total", precision="float", tuning=on)[0]
    b = engine.spectrum(batch, grid, comp="total", precision="float", tuning=off)[0]
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("sweep", ["direct", "recurrence"])
def test_zero_phase_step_is_kept(sweep):
    """A step whose retarded time does not advance (tau_s = tau_{s-1}: the
    track moves exactly dt along z per step, seen from theta = 0) passes the
    filter |phase_s - phase_{s-1}| = 0 < pi, so its threshold is +inf in the
    split path; the reciprocal there must not turn it into NaN and drop it."""
    dt = 1e-6
    n = 400
    s = np.arange(n, dtype=np.float64)
    x = 1e-7 * np.sin(0.05 * s)
    y = np.zeros(n)
    z = s * dt                                   # tau = t - z = 0 exactly on the axis
    ux = 0.5 * np.cos(0.05 * s)
    uy = np.zeros(n)
    uz = np.full(n, 50.0)
    tr = [[x, y, z, ux, uy, uz, 1.0, 0]]
    axes = {"omega": np.linspace(1e5, 4e5, 64), "theta": np.linspace(0.0, 0.01, 4),
            "phi": np.linspace(0.0, 2 * np.pi, 4, endpoint=False)}
    ref = oracle.spectrum("far", "total", tr, axes, dt)["total"]
    batch = engine.DeviceBatch(engine.pack_tracks(tr, dt=dt))
    grid = engine.DeviceGrid("far", axes["omega"], axes["theta"], axes["phi"])
    out = engine.spectrum(batch, grid, comp="total", precision="float",
                          tuning=engine.Tuning(sweep=sweep))[0]
    got = out.cpu().numpy().swapaxes(-1, -3)
    assert np.isfinite(got).all()
    # the on-axis column, where every step has tau_s = tau_{s-1}
    r_axis = rel_l2(got[..., 0, :], ref[..., 0, :])
    r = rel_l2(got, ref)
    print(f"zero-phase steps, {sweep}: on-axis rel-L2 {r_axis:.3e}, all {r:.3e}", flush=True)
    assert r_axis <= SPLIT_TOL and r <= SPLIT_TOL


@pytest.mark.parametrize("comp", ["total", "cartesian"])
@pytest.mark.parametrize("precision,sweep", [("double", "auto"), ("float", "direct"), ("float", "recurrence")])
def test_near_field_cartesian_screen(comp, precision, sweep):
    """The 64 x 64 Cartesian screen of the near-field workload (scaled): HIP
    against the oracle's Cartesian-screen restatement."""
    wl = workloads.config4(scale=0.25, n_particles=3, screen="cartesian")
    tr = wl.track_list()
    ref = oracle.spectrum("near", comp, tr, wl.axes, wl.batch.dt, L_screen=wl.L_screen)
    batch = engine.DeviceBatch(engine.pack_tracks(tr, dt=wl.batch.dt))
    outs = engine.spectrum(batch, wl.device_grid(), comp=comp, precision=precision,
                           tuning=engine.Tuning(sweep=sweep))
    torch.cuda.synchronize()
    got = {k: o.cpu().numpy().swapaxes(-1, -3) for k, o in zip(engine.COMP_KEYS[comp], outs)}
    check(got, ref, precision, "near xy " + comp)
