"""Host-side logic that needs no GPU: axes, CSR packing, snapshot schedules,
weight normalisation, analysis helpers and the benchmark workloads."""
import numpy as np
import pytest

from oracle import oracle
from synchrad.calc import SynchRad
from synchrad_amd import engine, workloads
from synchrad_amd.utils import J_in_um, Utilities, alpha_fs


def _axes_of(source, grid, mode, features):
    if source == "product":
        a = SynchRad({"grid": grid, "mode": mode, "ctx": False, "Features": list(features)}).Args
        return {k: a[k] for k in ("omega", "phi", "theta" if mode == "far" else "radius")}, a
    return oracle.make_axes(grid, mode, features), None


@pytest.mark.parametrize("source", ["product", "oracle"])
@pytest.mark.parametrize("mode", ["far", "near"])
@pytest.mark.parametrize("features", [[], ["logGrid"], ["wavelengthGrid"]])
def test_axes_closed_forms(source, mode, features):
    """Each axis against what defines it (synchrad/calc.py:389-451), for the
    calculator and the oracle separately: linear omega with both end points;
    'logGrid' from omega_min to omega_max with a constant ratio
    (omega_max/omega_min)^(1/(N-1)); 'wavelengthGrid' with 1/omega evenly
    spaced from 1/omega_max to 1/omega_min; theta or radius inclusive of both
    ends; phi from phi_min in steps of (phi_max - phi_min)/N_phi, the end
    excluded."""
    w0, w1, N = 1e3, 7e5, 37
    grid = [(w0, w1), (0.0, 0.05), (0.3, 2 * np.pi), (N, 11, 9)]
    ax, args = _axes_of(source, grid, mode, features)
    om = ax["omega"]
    assert om.dtype == np.float64 and om.shape == (N,)
    i = np.arange(N)
    if not features:
        assert om[0] == w0 and om[-1] == w1
        np.testing.assert_allclose(om, w0 + (w1 - w0) * i / (N - 1), rtol=1e-15, atol=0)
    elif features == ["logGrid"]:
        assert om[0] == w0
        assert om[-1] == pytest.approx(w1, rel=1e-14)
        ratio = (w1 / w0) ** (1.0 / (N - 1))
        np.testing.assert_allclose(om[1:] / om[:-1], ratio, rtol=1e-13)
    else:
        lam = 1.0 / om
        assert lam[0] == pytest.approx(1.0 / w1, rel=1e-15) and lam[-1] == pytest.approx(1.0 / w0, rel=1e-15)
        np.testing.assert_allclose(lam, 1.0 / w1 + (1.0 / w0 - 1.0 / w1) * i / (N - 1), rtol=1e-14)
        if args is not None:
            np.testing.assert_allclose(args["wavelengths"], lam, rtol=1e-15)
    key = "theta" if mode == "far" else "radius"
    a1 = ax[key]
    assert a1[0] == 0.0 and a1[-1] == 0.05
    np.testing.assert_allclose(a1, 0.05 * np.arange(11) / 10, rtol=1e-15, atol=1e-18)
    ph = ax["phi"]
    np.testing.assert_allclose(ph, 0.3 + (2 * np.pi - 0.3) * np.arange(9) / 9, rtol=1e-15)
    assert ph[-1] < 2 * np.pi
    if args is not None:
        # integration elements: |d omega| per interval, the first theta/radius and phi steps
        np.testing.assert_allclose(args["dw"], np.abs(np.diff(om)), rtol=1e-15)
        d1 = args["dth" if mode == "far" else "dr"]
        assert d1 == pytest.approx(0.005, rel=1e-13)
        assert args["dph"] == pytest.approx((2 * np.pi - 0.3) / 9, rel=1e-14)
        np.testing.assert_allclose(args["dV"], args["dw"] * d1 * args["dph"], rtol=1e-15)


def test_single_node_axes_have_unit_elements():
    """One node on an axis: omega is its lower bound and the element is 1
    (synchrad/calc.py:403-406, 416-424, 439-447)."""
    a = SynchRad({"grid": [(5.0, 9.0), (0.1, 0.2), (0.0, 1.0), (1, 1, 1)], "ctx": False}).Args
    assert a["omega"].tolist() == [5.0] and a["theta"].tolist() == [0.1] and a["phi"].tolist() == [0.0]
    assert a["dw"].tolist() == [1.0] and a["dth"] == 1.0 and a["dph"] == 1.0
    b = SynchRad({"grid": [(5.0, 9.0), (0.1, 0.2), (0.0, 1.0), (1, 1, 1)], "mode": "near", "ctx": False}).Args
    assert b["dr"] == 1.0 and b["dph"] == 1.0 and b["radius"].tolist() == [0.1]


def test_dtype_selects_precision():
    grid = [(1.0, 2.0), (0.0, 0.1), (0.0, 6.28), (4, 4, 4)]
    assert SynchRad({"grid": grid, "ctx": False}).precision == "double"
    c = SynchRad({"grid": grid, "ctx": False, "dtype": "float"})
    assert c.precision == "float" and c.Args["omega"].dtype == np.float64
    with pytest.raises(ValueError):
        SynchRad({"grid": grid, "ctx": False, "dtype": "half"})


def test_per_track_snapshots_equal_per_particle_linspace():
    rng = np.random.default_rng(0)
    lengths = rng.integers(1, 60, 25)
    tracks = [[np.zeros(n)] * 6 + [1.0, 0] for n in lengths]
    for nSnaps in (1, 2, 5, 7):
        hb = engine.pack_tracks(tracks, dt=0.1, nSnaps=nSnaps)
        assert hb.snaps_stride == nSnaps
        got = hb.it_snaps.reshape(len(tracks), nSnaps)
        for p, n in enumerate(lengths):
            want = np.linspace(0, int(n), nSnaps + 1, dtype=np.uint32)[1:]
            assert np.array_equal(got[p], want)
        assert np.array_equal(hb.it_end, lengths.astype(np.uint32))
        assert np.all(hb.it_start == 0)


def test_global_it_range_packing():
    tracks = [[np.zeros(10)] * 6 + [2.0, 7], [np.zeros(4)] * 6 + [0.5, 3], [np.zeros(3)] * 6 + [1.0]]
    hb = engine.pack_tracks(tracks, dt=0.1, it_range=(2, 20), nSnaps=3)
    assert hb.snaps_stride == 0
    assert np.array_equal(hb.it_snaps, np.linspace(2, 20, 4, dtype=np.uint32)[1:])
    assert list(hb.it_start) == [7, 3, 0]
    assert list(hb.it_end) == [20, 20, 20]
    assert list(hb.offsets) == [0, 10, 14, 17]
    assert list(hb.weight) == [2.0, 0.5, 1.0]
    with pytest.raises(ValueError):
        engine.pack_tracks(tracks, dt=0.1, it_range=(0, 0))
    with pytest.raises(ValueError):
        engine.pack_tracks(tracks, dt=0.0)


def test_n_intervals_counts_active_steps():
    tracks = [[np.zeros(n)] * 6 + [1.0, 0] for n in (0, 1, 2, 50)]
    assert engine.pack_tracks(tracks, dt=1.0).n_intervals() == 0 + 0 + 1 + 49
    # a global window shorter than a track stops its loop (kernel_farfield.cl:59)
    assert engine.pack_tracks(tracks, dt=1.0, it_range=(0, 10)).n_intervals() == 0 + 0 + 1 + 9


def test_weight_normalisation_mutates_caller_lists(monkeypatch):
    """synchrad/calc.py:241-265: 'mean'/'max' divide and 'ones' overwrite the
    weights in the caller's own track lists; total_weight sums them."""
    grid = [(1.0, 2.0), (0.0, 0.1), (0.0, 6.28), (4, 4, 4)]
    calc = SynchRad({"grid": grid, "ctx": False})
    calc.device = "cpu-test"
    seen = {}

    def fake_integrate(self, batch, comp, sigma):
        seen["w"] = batch.weight.copy()

    monkeypatch.setattr(SynchRad, "_integrate", fake_integrate)
    monkeypatch.setattr(SynchRad, "_init_raditaion", lambda self, comp, n: setattr(self, "_rad_keys", []))
    monkeypatch.setattr(SynchRad, "_spectr_from_device", lambda self, n: None)
    calc._rad_device = {}
    calc._grid = None
    tracks = [[np.zeros(3)] * 6 + [w, 0] for w in (1.0, 2.0, 6.0)]
    calc.calculate_spectrum(tracks, timeStep=0.1, weights_normalize="mean", verbose=False)
    assert [t[6] for t in tracks] == [1.0 / 3, 2.0 / 3, 2.0]
    assert calc.total_weight == pytest.approx(3.0)
    assert np.allclose(seen["w"], [1.0 / 3, 2.0 / 3, 2.0])
    calc.calculate_spectrum(tracks, timeStep=0.1, weights_normalize="ones", verbose=False)
    assert [t[6] for t in tracks] == [1.0, 1.0, 1.0]
    tracks = [[np.zeros(3)] * 6 + [w, 0] for w in (1.0, 4.0)]
    calc.calculate_spectrum(tracks, timeStep=0.1, weights_normalize="max", Np_max=1, verbose=False)
    assert tracks[0][6] == 1.0 and tracks[1][6] == 4.0   # Np_max cuts before normalising
    assert calc.total_weight == 1.0


class _Spec(Utilities):
    def __init__(self, rad, Args):
        self.Data = {"radiation": rad}
        self.Args = Args


def test_energy_integrals_exact_for_linear_integrands():
    """get_energy_spectrum averages neighbouring theta samples and integrates
    val*sin(theta_mid) with the trapezoid rule (synchrad/utils.py:81-86); a
    spectrum built so that val_mid*sin(theta_mid) is linear in theta_mid is
    integrated exactly.  get_energy then integrates over omega."""
    theta = np.linspace(0.01, 0.2, 9)
    th_mid = 0.5 * (theta[1:] + theta[:-1])
    # choose val on the nodes so that midpoint averages equal c*th_mid/sin(th_mid)
    target = th_mid / np.sin(th_mid)
    val_nodes = np.zeros_like(theta)
    val_nodes[0] = target[0]
    for i in range(1, theta.size):
        val_nodes[i] = 2 * target[i - 1] - val_nodes[i - 1]
    omega = np.linspace(1.0, 3.0, 5)
    phi = np.linspace(0, 2 * np.pi, 6, endpoint=False)
    raw = omega[:, None, None] * val_nodes[None, :, None] * np.ones(phi.size)[None, None, :]
    Args = {"mode": "far", "comp": "total", "omega": omega, "theta": theta, "phi": phi,
            "dph": phi[1] - phi[0]}
    s = _Spec({"total": raw[None]}, Args)
    norm = alpha_fs / (4 * np.pi ** 2)
    es = s.get_energy_spectrum()
    want = norm * omega * (2 * np.pi) * 0.5 * (th_mid[-1] ** 2 - th_mid[0] ** 2)
    assert np.allclose(es, want, rtol=1e-12)
    assert s.get_energy() == pytest.approx(norm * 2 * np.pi * 0.5 * (th_mid[-1] ** 2 - th_mid[0] ** 2) * 4.0,
                                           rel=1e-12)
    assert np.allclose(s.get_full_spectrum(phot_num=True)[0], norm * raw[0] / omega[0])
    assert np.allclose(s.get_full_spectrum(lambda0_um=2.0), norm * raw * J_in_um / 2.0)
    Args["mode"] = "near"
    Args["radius"] = theta
    near = s.get_full_spectrum()
    assert np.allclose(near, alpha_fs * np.pi / 4 / (2 * np.pi) ** 2 * raw)


def test_workload_sizes_match_the_benchmark_definitions():
    assert workloads.config1().evals == 256 * 32 * 32 * 1599
    w2 = workloads.config2(n_particles=3)
    assert w2.n_grid == 512 * 64 * 64 and w2.batch.n_intervals() == 3 * 7999
    w4 = workloads.config4(n_particles=2)
    assert w4.n_grid == 128 * 64 * 64 and w4.batch.n_intervals() == 2 * 3999
    w5 = workloads.config5(n_particles=2)
    assert w5.n_grid == 128 * 32 * 32 and w5.batch.n_intervals() == 2 * 3999
    # fp32-rounded tracks, used as exact doubles
    assert np.array_equal(w5.batch.x, w5.batch.x.astype(np.float32).astype(np.float64))


def test_betatron_orbit_is_a_focusing_oscillation():
    wl = workloads.config2(n_particles=5, n_steps=4000, grid=(8, 4, 4))
    hb = wl.batch
    kp = wl.extra["kp"]
    x = hb.x.reshape(5, -1)
    ux = hb.ux.reshape(5, -1)
    g = np.sqrt(1 + hb.ux ** 2 + hb.uy ** 2 + hb.uz ** 2).reshape(5, -1)
    # betatron wavenumber k_b = k_p / sqrt(2 gamma): count zero crossings of u_x
    for p in range(5):
        crossings = np.sum(np.diff(np.sign(ux[p])) != 0)
        kb = kp / np.sqrt(2 * g[p].mean())
        length = 4000 * hb.dt
        assert abs(crossings - kb * length / np.pi) <= 2
    # energy is conserved by a transverse force only to second order
    assert np.all(np.abs(g[:, -1] / g[:, 0] - 1) < 1e-3)


def test_config3_lengths_are_log_uniform():
    wl = workloads.config3(scale=0.05, n_particles=400)
    n = np.diff(wl.batch.offsets)
    assert n.min() >= 200 and n.max() <= 20000
    assert abs(np.mean(np.log(n)) - 0.5 * (np.log(200) + np.log(20000))) < 0.25


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message before
    touching a device."""
    import json
    import os
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(repo, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_autotune_candidates_and_direction_subset_rule():
    """Candidates keep at least three quarters of a block's lanes busy; large
    far-field launches (>= 1e12 evaluations) are sampled on a subset of the
    directions, smaller ones on truncated tracks."""
    from types import SimpleNamespace
    from synchrad_amd import autotune

    split = autotune.candidates("far", "float", True, 512)
    assert {(c.sweep, c.omega_per_lane, c.waves_per_block) for c in split} == \
        {("recurrence", 2, 1), ("recurrence", 2, 2), ("recurrence", 2, 4), ("recurrence", 4, 1),
         ("recurrence", 4, 2)}
    assert all(c.sweep == "direct" for c in autotune.candidates("far", "float", False, 256))
    assert {c.omega_per_lane for c in autotune.candidates("far", "double", True, 512)} == {1, 2}

    def large(mode, n_grid, intervals):
        grid = SimpleNamespace(mode=mode, n_grid=n_grid)
        batch = SimpleNamespace(host=SimpleNamespace(n_intervals=lambda: intervals))
        return autotune.is_large(batch, grid)

    assert large("far", 512 * 64 * 64, 2000 * 7999)         # the betatron beam, 3.4e13
    assert not large("far", 256 * 32 * 32, 1599)            # one undulator electron, 4.2e8
    assert not large("near", 128 * 64 * 64, 500 * 3999)     # near field: tiles are not reordered


def _polar_spec(mode, comp="total", n=(6, 9, 8), seed=5):
    rng = np.random.default_rng(seed)
    omega = np.linspace(1.0, 4.0, n[0])
    ax1 = np.linspace(0.0, 0.05, n[1])
    phi = np.linspace(0, 2 * np.pi, n[2], endpoint=False)
    keys = engine.COMP_KEYS[comp]
    rad = {k: rng.random((2,) + n) for k in keys}
    if comp == "total":
        rad["total"][:, :, 0, :] = rad["total"][:, :, 0, :1]   # one value at r = 0 for every phi
    Args = {"mode": mode, "comp": comp, "omega": omega, "phi": phi, "dph": phi[1] - phi[0],
            "dw": np.abs(np.diff(omega)), ("theta" if mode == "far" else "radius"): ax1}
    return _Spec(rad, Args)


def test_radial_integral_exact_for_constant_spectra():
    """Near field: the screen integral of a constant c is c * pi * r_max^2
    (trapezoid rule of the linear r c, times 2 pi from the phi sum)."""
    s = _polar_spec("near")
    s.Data["radiation"]["total"][:] = 3.0
    pref = alpha_fs * np.pi / 4 / (2 * np.pi) ** 2
    es = s.get_energy_spectrum()
    np.testing.assert_allclose(es, pref * 3.0 * np.pi * 0.05 ** 2, rtol=1e-13)
    assert s.get_energy() == pytest.approx(pref * 3.0 * np.pi * 0.05 ** 2 * 3.0, rel=1e-13)


def test_quadrature_weights_are_the_trapezoid_rule():
    from synchrad_amd import utils
    rng = np.random.default_rng(1)
    x = np.sort(rng.random(17))
    f = rng.random(17)
    trap = getattr(np, "trapezoid", None) or np.trapz
    assert utils.trapezoid_weights(x) @ f == pytest.approx(trap(f, x), rel=1e-14)
    assert utils.trapezoid_weights(x[::-1]) @ f[::-1] == pytest.approx(-trap(f, x), rel=1e-14)
    assert utils.trapezoid_weights(x[:1]).tolist() == [0.0]
    # the polar rule on theta midpoints of midpoint averages times sin
    th = np.sort(rng.random(9)) * 0.1
    mid = 0.5 * (th[1:] + th[:-1])
    want = trap(0.5 * (f[1:9] + f[:8]) * np.sin(mid), mid)
    assert utils.polar_angle_weights(th) @ f[:9] == pytest.approx(want, rel=1e-13)


def test_spot_frequency_node_rule():
    """get_spot(k0) reads the first node not below k0, or the next one when
    that is closer (synchrad/utils.py:119-124)."""
    from synchrad_amd import utils
    om = np.array([1.0, 2.0, 3.0, 4.0])
    assert utils.nearest_node_rule(om, 1.9) == 1       # first node >= k0
    assert utils.nearest_node_rule(om, 2.0) == 1
    assert utils.nearest_node_rule(om, 0.2) == 0
    assert utils.nearest_node_rule(om, 2.7) == 2       # node 2 (3.0) is above and nearest
    assert utils.nearest_node_rule(np.array([1.0, 1.1, 5.0]), 0.5) == 0
    assert utils.nearest_node_rule(np.array([1.0, 1.1, 5.0]), 1.05) == 1
    assert utils.nearest_node_rule(om, 9.0) == 3
    s = _polar_spec("far")
    full = s.get_full_spectrum()
    # omega = 1, 1.6, 2.2, 2.8, 3.4, 4: k0 = 2.45 reads 2.8 (the first node above
    # it), although 2.2 lies closer -- the reference compares 2.8 with 3.4 only
    assert np.array_equal(s.get_spot(k0=2.45), full[3])
    trap = getattr(np, "trapezoid", None) or np.trapz
    np.testing.assert_allclose(s.get_spot(), trap(full, s.Args["omega"], axis=0), rtol=1e-14)


@pytest.mark.parametrize("mode", ["far", "near"])
def test_spot_cartesian_equals_scipy_griddata(mode):
    """The Cartesian resampling is griddata's linear interpolation over the
    polar nodes, zero outside their hull (synchrad/utils.py:140-151)."""
    from scipy.interpolate import griddata
    s = _polar_spec(mode)
    spot = s.get_spot()
    r = s.Args["theta" if mode == "far" else "radius"]
    ph, th = np.meshgrid(s.Args["phi"], r)
    pts = ((th * np.cos(ph)).ravel(), (th * np.sin(ph)).ravel())
    m = 0.7 * r.max()
    mesh = np.mgrid[-m:m:40j, -m:m:30j]
    want = griddata(pts, spot.ravel(), (mesh[0].ravel(), mesh[1].ravel()), fill_value=0.0,
                    method="linear").reshape(mesh[0].shape)
    got, ext = s.get_spot_cartesian(th_part=0.7, bins=(40, 30))
    assert got.shape == (40, 30) and ext.tolist() == [-m, m, -m, m]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15 * np.abs(want).max())


def test_coherent_components():
    """Coherent runs: 'total' sums the squares of the six outputs, a named
    component is the complex sum re + i im."""
    s = _polar_spec("far", comp="cartesian_complex")
    rad = s.Data["radiation"]
    pref = alpha_fs / (4 * np.pi ** 2)
    tot = s.get_full_spectrum()
    np.testing.assert_allclose(tot, pref * sum(rad[k][-1] ** 2 for k in rad), rtol=1e-14)
    y = s.get_full_spectrum(comp="y", iteration=0)
    assert np.iscomplexobj(y)
    np.testing.assert_allclose(y, pref * (rad["yre"][0] + 1j * rad["yim"][0]), rtol=1e-14)


@pytest.mark.parametrize("screen", ["polar", "cartesian"])
def test_bench_cpu_baseline_counts_the_sample(screen):
    """bench.py's CPU leg counts evaluations as grid points x step intervals of
    its sample, on either near-field screen."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "bench_mod", os.path.join(os.path.dirname(os.path.dirname(__file__)), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    wl = workloads.config4(scale=0.125, n_particles=1, screen=screen)
    res, sample = bench.cpu_baseline(wl, 4, 1, "0,0,0")
    assert wl.n_grid == 16 * 8 * 8
    assert res["value"] > 0 and f"{wl.n_grid * 3999:.3g} evals" in res["sample"]
    assert sample[2]["total"].shape == (1, 16, 8, 8)


def test_pack_tracks_writes_the_upload_buffer_once():
    """pack_tracks concatenates the columns straight into the batch's upload
    buffer; the upload packing then reuses that buffer and its bytes equal a
    fresh packing of the same arrays."""
    rng = np.random.default_rng(5)
    tracks = [[rng.standard_normal(n) for _ in range(6)] + [float(rng.uniform(0.5, 2)), int(i)]
              for i, n in enumerate((5, 0, 17, 1, 9))]
    for kw in ({}, {"it_range": (0, 20), "nSnaps": 3}):
        hb = engine.pack_tracks(tracks, dt=0.1, **kw)
        layout, total = engine._packed_layout(hb)
        buf = engine._pack_host(hb, layout, total)
        assert buf is hb.packed
        fresh = engine.pack_csr(hb.offsets.copy(), *(np.array(getattr(hb, c)) for c in ("x", "y", "z", "ux", "uy", "uz")),
                                hb.weight.copy(), hb.it_start.copy(), dt=0.1, **kw)
        assert fresh.packed is None
        other = engine._pack_host(fresh, layout, total)
        for name, (at, nbytes) in layout.items():
            assert bytes(buf[at:at + nbytes]) == bytes(other[at:at + nbytes]), name
        for i, c in enumerate(("x", "y", "z", "ux", "uy", "uz")):
            np.testing.assert_array_equal(getattr(hb, c), np.concatenate([t[i] for t in tracks]))
