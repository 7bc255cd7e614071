"""The drop-in's file paths on the GPU: a track file in the converters' layout
(synchrad/converters.py:101-127) ingested by calculate_spectrum(file_tracks=...)
(synchrad/calc.py:186-219), with and without a global misc/it_range, and the
spectrum file it writes (synchrad/calc.py:274-290) reloaded by
SynchRad(file_spectrum=...) (synchrad/calc.py:648-666)."""
import numpy as np
import pytest

from oracle import oracle
from synchrad_amd import h5io, workloads

pytestmark = pytest.mark.gpu

try:
    h5io.lib()
    _have_h5 = True
except ImportError:  # pragma: no cover - depends on the image
    _have_h5 = False


def _write_tracks(path, tracks, cdt, it_range=None):
    with h5io.File(path, "w") as f:
        for i, t in enumerate(tracks):
            for c, name in enumerate(("x", "y", "z", "ux", "uy", "uz")):
                f[f"tracks/{i}/{name}"] = np.asarray(t[c])
            f[f"tracks/{i}/w"] = float(t[6])
            f[f"tracks/{i}/it_start"] = int(t[7])
        f["misc/cdt"] = cdt
        f["misc/N_particles"] = len(tracks)
        if it_range is not None:
            f["misc/it_range"] = np.array(it_range)
        f["misc/propagation_direction"] = "z"


@pytest.mark.skipif(not _have_h5, reason="no HDF5 C library")
@pytest.mark.parametrize("global_range", [False, True])
@pytest.mark.parametrize("dtype,tol", [("double", 1e-10), ("float", 1e-4)])
def test_track_file_to_spectrum_file(tmp_path, global_range, dtype, tol):
    from synchrad.calc import SynchRad
    wl = workloads.config2(n_particles=3, n_steps=1200, grid=(48, 8, 8))
    tracks = wl.track_list()
    tracks = [[np.array(a) for a in t[:6]] + [float(t[6]), 5 * i] for i, t in enumerate(tracks)]
    it_range = (0, max(5 * i + t[0].size for i, t in enumerate(tracks))) if global_range else None
    ftr, fsp = str(tmp_path / "tracks.h5"), str(tmp_path / "spectrum.h5")
    _write_tracks(ftr, tracks, wl.batch.dt, it_range)

    grid = wl.calc_input["grid"]
    calc = SynchRad({"grid": grid, "dtype": dtype})
    calc.calculate_spectrum(file_tracks=ftr, comp="total", nSnaps=2, file_spectrum=fsp, verbose=False)
    got = calc.Data["radiation"]["total"]

    # the same tracks as a list (it_start per track, as the file holds them)
    ref_calc = SynchRad({"grid": grid, "dtype": dtype})
    ref_calc.calculate_spectrum(particleTracks=[list(t) for t in tracks], timeStep=wl.batch.dt, comp="total",
                                nSnaps=2, it_range=it_range, verbose=False)
    assert np.array_equal(got, ref_calc.Data["radiation"]["total"])

    ref = oracle.spectrum("far", "total", [list(t) for t in tracks], wl.axes, wl.batch.dt, it_range=it_range,
                          nSnaps=2)["total"]
    r = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"file round trip {dtype} global_range={global_range}: rel-L2 vs oracle {r:.3e}", flush=True)
    assert r <= tol

    back = SynchRad(file_spectrum=fsp)
    assert np.array_equal(back.Data["radiation"]["total"], got)
    assert back.Args["dtype"] == dtype and back.Args["mode"] == "far"
    assert np.array_equal(back.Args["omega"], calc.Args["omega"])
    assert np.array_equal(np.asarray(back.snap_iterations), np.asarray(calc.snap_iterations))
    assert back.total_weight == pytest.approx(calc.total_weight)
    # the analysis helpers work on the reloaded file
    assert np.isclose(back.get_energy(lambda0_um=1.0), calc.get_energy(lambda0_um=1.0), rtol=1e-12)


@pytest.mark.skipif(not _have_h5, reason="no HDF5 C library")
@pytest.mark.parametrize("dtype,tol", [("double", 1e-10), ("float", 1e-4)])
def test_nan_gapped_histories_to_spectrum(tmp_path, dtype, tol):
    """Undulator histories with NaN gaps (particles entering and leaving the
    tracker) cut into pieces and written as a track file by the converter;
    the file's spectrum (global misc/it_range, per-piece it_start) matches the
    oracle on the same pieces, and the CSR pieces give the same spectrum
    directly."""
    from synchrad.calc import SynchRad
    from synchrad_amd import converters, engine
    rng = np.random.default_rng(31)
    n_p, n_it, dt = 4, 900, 0.01 / 24
    hist = np.zeros((7, n_p, n_it))
    for p in range(n_p):
        arrs = workloads.planar_undulator(rng.uniform(40, 80), rng.uniform(0.5, 1.5), 0.01, n_it, dt,
                                          psi=rng.uniform(0, 6.28))
        hist[:6, p] = np.stack(arrs)
        hist[6, p] = rng.uniform(0.5, 1.5)
    hist[6, 0, 100:130] = np.nan
    hist[6, 1, :250] = np.nan
    hist[6, 2, 600:] = np.nan
    hist[6, 3, 40:45] = np.nan
    path = str(tmp_path / "pieces.h5")
    pieces = converters.tracks_from_arrays(*hist, cdt=dt, fname=path, shortest_track=8)
    assert pieces.n_tracks == 6            # 2 + 1 + 1 + 2 runs of present samples
    axes = oracle.make_axes([(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (40, 6, 5)], "far")
    grid = [(100.0, 4e4), (0.0, 0.03), (0.0, 2 * np.pi), (40, 6, 5)]
    calc = SynchRad({"grid": grid, "dtype": dtype})
    calc.calculate_spectrum(file_tracks=path, comp="total", verbose=False)
    got = calc.Data["radiation"]["total"]
    it_range = tuple(pieces.it_range())
    ref = oracle.spectrum("far", "total", pieces.track_list(), axes, dt, it_range=it_range)["total"]
    r = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"converted pieces {dtype}: rel-L2 vs oracle {r:.3e}", flush=True)
    assert r <= tol
    ax1 = axes["theta"]
    dgrid = engine.DeviceGrid("far", axes["omega"], ax1, axes["phi"])
    direct = engine.spectrum(engine.DeviceBatch(pieces.host_batch(dt, it_range=it_range)), dgrid, comp="total",
                             precision=dtype)[0]
    assert np.array_equal(direct.cpu().numpy().swapaxes(-1, -3), got)
