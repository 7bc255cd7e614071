"""Every per-particle C-ABI entry point against the CPU FP64 oracle, the
reference's near-field analytic script on the HIP path, and stream ordering.

The per-particle entry points keep the argument order of the reference's
OpenCL kernels (kernel_farfield.cl:6-28, 111-135; kernel_nearfield.cl:5-27,
106-130): one call adds one particle, as one reference launch does.  They are
driven here exactly as the reference's host loop drives its kernels
(synchrad/calc.py:257-267, 296-322) and compared with the oracle directly,
in both precisions.
"""
import numpy as np
import pytest
import torch

from oracle import oracle
from synchrad_amd import _abi, engine, workloads

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
SPLIT_TOL = 1e-4
PRECS = {"double": _abi.SR_PREC_F64, "float": _abi.SR_PREC_SPLIT}


def rel_l2(got, ref):
    n = np.linalg.norm(ref.ravel())
    return np.linalg.norm((got - ref).ravel()) / n if n > 0 else np.linalg.norm(got.ravel())


def _dev_f64(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def _ragged(seed):
    """Undulator tracks of different lengths (one single-sample, one empty)
    with it_start offsets."""
    rng = np.random.default_rng(seed)
    tr = []
    for L in (300, 1, 257, 0, 64, 513):
        arrs = workloads.planar_undulator(rng.uniform(40, 80), rng.uniform(0.5, 1.5), 0.01, L, 0.01 / 24,
                                          psi=rng.uniform(0, 6.28))
        tr.append(list(arrs) + [float(rng.uniform(0.5, 1.5)), int(rng.integers(0, 200))])
    return tr


def _per_particle(mode, comp, tracks, axes, dt, precision, it_range=None, nSnaps=1, L_screen=None):
    """The reference's particle loop over the per-particle entry points:
    per-track mode sets it_start = 0, itEnd = n and snapshots over [0, n];
    a global it_range keeps each track's it_start and one snapshot list."""
    L = _abi.load()
    nO, nP = axes["omega"].size, axes["phi"].size
    if mode == "far":
        n1 = axes["theta"].size
        ax = [_dev_f64(2 * np.pi * axes["omega"]), _dev_f64(np.sin(axes["theta"])),
              _dev_f64(np.cos(axes["theta"])), _dev_f64(np.sin(axes["phi"])), _dev_f64(np.cos(axes["phi"]))]
        ax_args = [a.data_ptr() for a in ax]
    else:
        n1 = axes["radius"].size
        ax = [_dev_f64(2 * np.pi * axes["omega"]), _dev_f64(axes["radius"]), _dev_f64(np.sin(axes["phi"])),
              _dev_f64(np.cos(axes["phi"]))]
        ax_args = [a.data_ptr() for a in ax] + [float(L_screen)]
    keys = engine.COMP_KEYS[comp]
    outs = [torch.zeros((nSnaps, nP, n1, nO), dtype=torch.float64, device="cuda") for _ in keys]
    fn = {("far", "total"): L.sr_far_total, ("far", "cartesian"): L.sr_far_cartesian,
          ("near", "total"): L.sr_near_total, ("near", "cartesian"): L.sr_near_cartesian}[(mode, comp)]
    keep = []
    for t in tracks:
        n = np.asarray(t[0]).size
        if n == 0:
            continue          # the reference would read out of bounds
        if it_range is None:
            its, ite, snaps = 0, n, oracle.snap_iterations((0, n), nSnaps)
        else:
            its, ite, snaps = int(t[7]), int(it_range[-1]), oracle.snap_iterations(it_range, nSnaps)
        d = [_dev_f64(a) for a in t[:6]]
        s = torch.as_tensor(np.ascontiguousarray(snaps, dtype=np.uint32).view(np.int32), device="cuda")
        keep += d + [s]
        rc = fn(*[o.data_ptr() for o in outs], *[a.data_ptr() for a in d], float(t[6]), its, ite, n,
                *ax_args, nO, n1, nP, dt, nSnaps, s.data_ptr(), PRECS[precision], None)
        _abi.check(rc)
    torch.cuda.synchronize()
    return {k: np.ascontiguousarray(o.cpu().numpy().swapaxes(-1, -3)) for k, o in zip(keys, outs)}


def _check(got, ref, precision, label):
    tol = F64_TOL if precision == "double" else SPLIT_TOL
    for k in ref:
        r = rel_l2(got[k], ref[k])
        print(f"{label} {k} {precision}: rel-L2 vs oracle {r:.3e}", flush=True)
        assert r <= tol, (label, k, r)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("comp", ["total", "cartesian"])
def test_far_entry_points_vs_oracle(comp, precision):
    """sr_far_total / sr_far_cartesian, per-track mode, ragged tracks."""
    tr = _ragged(21)
    axes = oracle.make_axes([(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (40, 6, 5)], "far")
    dt = 0.01 / 24
    ref = oracle.spectrum("far", comp, tr, axes, dt)
    got = _per_particle("far", comp, tr, axes, dt, precision)
    _check(got, ref, precision, f"sr_far_{comp}")


@pytest.mark.parametrize("precision", ["double", "float"])
def test_far_entry_point_global_window_snapshots_vs_oracle(precision):
    """sr_far_total with a shared it_range, it_start offsets and 3 snapshots."""
    tr = _ragged(22)
    axes = oracle.make_axes([(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (40, 6, 5)], "far")
    dt = 0.01 / 24
    it_range = (50, 650)
    ref = oracle.spectrum("far", "total", tr, axes, dt, it_range=it_range, nSnaps=3)
    got = _per_particle("far", "total", tr, axes, dt, precision, it_range=it_range, nSnaps=3)
    _check(got, ref, precision, "sr_far_total window")


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("comp", ["total", "cartesian"])
def test_near_entry_points_vs_oracle(comp, precision):
    """sr_near_total / sr_near_cartesian on a share of the near-field screen."""
    wl = workloads.config4(scale=0.125, n_particles=3)
    tr = wl.track_list()
    ref = oracle.spectrum("near", comp, tr, wl.axes, wl.batch.dt, L_screen=wl.L_screen)
    got = _per_particle("near", comp, tr, wl.axes, wl.batch.dt, precision, L_screen=wl.L_screen)
    _check(got, ref, precision, f"sr_near_{comp}")


def _reference_near_script_tracks(seed=12):
    """The tracks of tests/test_undulator_analytic_near.py:7-50 (seeded)."""
    K0, Periods, Np, g0 = 0.1, 50, 24, 100.0
    dg = 1e-4 * g0
    dt = 1.0 / 64
    steps = int((Periods + 2) / dt) + 1

    def uxf(z):
        val = K0 * np.sin(2 * np.pi * z)
        val *= (z > 0) * (z < 1.5) * z / 1.5 + (z > 1.5)
        val *= (z > Periods - 1.5) * (z < Periods) * (Periods - z) / 1.5 + (z < Periods - 1.5)
        return val

    t = np.r_[-1:Periods + 1:steps * 1j]
    rng = np.random.default_rng(seed)
    tracks = []
    for g0p in g0 + dg * rng.standard_normal(Np):
        ggp = g0p / (1.0 + K0 ** 2 / 2) ** 0.5
        vb = (1.0 - ggp ** -2) ** 0.5
        z = vb * t
        ux = uxf(z - 0.5 * dt)
        uz = (g0 ** 2 - 1 - ux ** 2) ** 0.5
        x = ux[0] / g0p * dt / 2 + np.cumsum(ux / g0p) * dt
        tracks.append([x, np.zeros_like(t), z, ux, np.zeros_like(t), uz, 1.0, 0])
    return tracks, dict(K0=K0, Periods=Periods, Np=Np, g0=g0, dt=dt)


def test_reference_near_field_script_on_gpu():
    """tests/test_undulator_analytic_near.py of the reference through the
    drop-in import path at its own 128 x 256 x 32 screen at L = 1e5: double,
    then float after the reference's own re-init calls.  The photon energy
    lies within 15 % of the analytic estimate in both (the oracle on the CPU
    gives 10.9 % on this screen), and the two precisions agree to 1e-4."""
    from synchrad.calc import SynchRad
    from synchrad.utils import J_in_um
    tracks, p = _reference_near_script_tracks()
    g0 = p["g0"]
    gg = g0 / (1.0 + p["K0"] ** 2 / 2) ** 0.5
    k_res = 2 * gg ** 2
    L_scr = 1e5
    calc = SynchRad({"grid": [(0.02 * k_res, 1.1 * k_res), (0, L_scr * 1 / g0), (0.0, 2 * np.pi),
                              (128, 256, 32)], "mode": "near"})
    calc.calculate_spectrum([list(t) for t in tracks], timeStep=p["dt"], comp="total", Np_max=p["Np"],
                            L_screen=L_scr)
    s64 = calc.Data["radiation"]["total"].copy()
    e64 = calc.get_energy(lambda0_um=1)
    e_th = (p["Np"] * k_res * J_in_um * (7 * np.pi / 24) / 137.0 * p["K0"] ** 2 * (1 + p["K0"] ** 2 / 2)
            * p["Periods"])
    calc.Args["dtype"] = "float"
    calc.Args["native"] = True
    calc._init_args(calc.Args)
    calc._init_data()
    calc._compile_kernels()
    calc.calculate_spectrum([list(t) for t in tracks], timeStep=p["dt"], comp="total", Np_max=p["Np"],
                            L_screen=L_scr)
    s32 = calc.Data["radiation"]["total"]
    e32 = calc.get_energy(lambda0_um=1)
    for name, e in (("double", e64), ("float", e32)):
        dev = abs(e - e_th) / e_th
        print(f"near-field script, {name}: deviation from the analytic estimate {100 * dev:.2f} %", flush=True)
        assert dev < 0.15
    r = rel_l2(s32, s64)
    print(f"near-field script: float vs double rel-L2 {r:.3e}", flush=True)
    assert r <= SPLIT_TOL


@pytest.mark.parametrize("fn", ["spectrum", "fields", "adjoint"])
def test_side_stream_matches_current_stream(fn):
    """Calls given a stream other than the current one order themselves after
    the current stream's work (output clears, uploads) and give the same
    result as calls on the current stream."""
    wl = workloads.config2(n_particles=4, n_steps=2000, grid=(128, 16, 8))
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(wl.batch)
    side = torch.cuda.Stream()

    def run(stream):
        if fn == "spectrum":
            return engine.spectrum(batch, grid, comp="total", precision="float", stream=stream)
        if fn == "fields":
            return [engine.far_fields(batch, grid, comp="total", precision="float", stream=stream)[1][0]]
        dS = [torch.ones((1, grid.n_phi, grid.n1, grid.n_omega), dtype=torch.float64, device="cuda")]
        return list(engine.far_adjoint(batch, grid, dS, comp="total", precision="float", stream=stream))

    ref = [t.clone() for t in run(None)]
    torch.cuda.synchronize()
    for _ in range(2):
        got = run(side)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
