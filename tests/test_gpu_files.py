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
