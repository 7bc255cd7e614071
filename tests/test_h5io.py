"""HDF5 files through the libhdf5 binding (synchrad_amd.h5io): round trips of
the value kinds the reference's files hold, the layout checked independently
with the HDF5 tools' own h5dump, and a spectrum file in the reference's layout
(synchrad/calc.py:274-290) reloaded by SynchRad(file_spectrum=...)
(synchrad/calc.py:648-666).  CPU only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from synchrad_amd import h5io

try:
    h5io.lib()
except ImportError:  # pragma: no cover - depends on the image
    pytest.skip("no HDF5 C library", allow_module_level=True)


def _h5dump():
    for cand in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if cand and os.path.exists(cand):
            return cand
    return None


def test_round_trip_of_the_value_kinds(tmp_path):
    path = str(tmp_path / "kinds.h5")
    rng = np.random.default_rng(0)
    spec = rng.random((2, 5, 4, 3))
    with h5io.File(path, "w") as f:
        f["radiation/total"] = spec
        f["a/f64"] = 0.25
        f["a/f32"] = np.float32(1.5)
        f["a/i64"] = 7
        f["a/u32"] = np.arange(4, dtype=np.uint32)
        f["a/str"] = "double"
        f["a/strs"] = ["logGrid", "wavelengthGrid"]
        f["a/empty"] = []
        f["a/tuple"] = (64, 32, 16)
        f["a/flag"] = np.bool_(True)
    with h5io.File(path, "r") as f:
        assert f.keys() == ["a", "radiation"]
        assert f["a"].keys() == sorted(["f64", "f32", "i64", "u32", "str", "strs", "empty", "tuple", "flag"])
        got = f["radiation/total"][()]
        assert got.dtype == np.float64 and np.array_equal(got, spec)
        assert f["a/f64"][()] == 0.25 and isinstance(f["a/f64"][()], np.float64)
        assert f["a/f32"][()] == np.float32(1.5) and f["a/f32"][()].dtype == np.float32
        assert f["a/i64"][()] == 7 and f["a/i64"][()].dtype == np.int64
        u = f["a/u32"][()]
        assert u.dtype == np.uint32 and list(u) == [0, 1, 2, 3]
        assert f["a/str"][()] == b"double"                       # bytes, as h5py >= 3 returns
        assert list(f["a/strs"][()]) == [b"logGrid", b"wavelengthGrid"]
        assert f["a/empty"][()].shape == (0,)
        assert list(f["a/tuple"][()]) == [64, 32, 16]
        assert f["a/flag"][()] == 1
        assert "f64" in f["a"] and "nope" not in f["a"]
        with pytest.raises(KeyError):
            f["a/nope"]


def test_existing_names_are_not_overwritten(tmp_path):
    path = str(tmp_path / "dup.h5")
    with h5io.File(path, "w") as f:
        f["x"] = 1.0
        with pytest.raises(ValueError):
            f["x"] = 2.0


@pytest.mark.skipif(_h5dump() is None, reason="h5dump not available")
def test_layout_seen_by_h5dump(tmp_path):
    """A track file in the converters' layout (synchrad/converters.py:101-127),
    read back by the HDF5 project's own h5dump."""
    path = str(tmp_path / "tracks.h5")
    with h5io.File(path, "w") as f:
        for i in range(2):
            for c in ("x", "y", "z", "ux", "uy", "uz"):
                f[f"tracks/{i}/{c}"] = np.linspace(0.0, 1.0, 10 + i)
            f[f"tracks/{i}/w"] = 0.5
            f[f"tracks/{i}/it_start"] = 3
        f["misc/cdt"] = 0.01
        f["misc/N_particles"] = 2
        f["misc/it_range"] = np.array([3, 14])
        f["misc/propagation_direction"] = "z"
    out = subprocess.run([_h5dump(), "-H", path], capture_output=True, text=True, check=True).stdout
    assert 'GROUP "tracks"' in out and 'GROUP "misc"' in out
    block = out[out.index('DATASET "cdt"'):]
    assert "H5T_IEEE_F64LE" in block.split("DATASET")[1] and "SCALAR" in block.split("DATASET")[1]
    assert "DATASPACE  SIMPLE { ( 11 ) / ( 11 ) }" in out           # track 1 has 11 samples
    assert "H5T_STD_I64LE" in out[out.index('DATASET "N_particles"'):].split("DATASET")[1]
    data = subprocess.run([_h5dump(), "-d", "/misc/it_range", path], capture_output=True, text=True,
                          check=True).stdout
    assert "3, 14" in data
    sdata = subprocess.run([_h5dump(), "-d", "/misc/propagation_direction", path], capture_output=True,
                           text=True, check=True).stdout
    assert '"z"' in sdata and "H5T_CSET_UTF8" in sdata


def test_spectrum_file_reloads_in_synchrad(tmp_path):
    """The reference's spectrum-file layout, reloaded the reference's way."""
    from synchrad_amd.calc import SynchRad
    path = str(tmp_path / "spec.h5")
    rng = np.random.default_rng(1)
    rad = {k: rng.random((1, 6, 5, 4)) for k in ("x", "y", "z")}
    omega = np.linspace(1.0, 2.0, 6)
    with h5io.File(path, "w") as f:
        for k, v in rad.items():
            f["radiation/" + k] = v
        f["Args/mode"] = "far"
        f["Args/dtype"] = "double"
        f["Args/omega"] = omega
        f["Args/theta"] = np.linspace(0.0, 0.1, 5)
        f["Args/phi"] = np.linspace(0.0, 2 * np.pi, 4, endpoint=False)
        f["Args/dV"] = np.full(5, 0.01)
        f["Args/gridNodeNums"] = (6, 5, 4)
        f["Args/Features"] = []
        f["snap_iterations"] = np.array([100], dtype=np.uint32)
        f["total_weight"] = 3.0
    calc = SynchRad(file_spectrum=path)
    assert sorted(calc.Data["radiation"].keys()) == ["x", "y", "z"]
    for k in rad:
        assert np.array_equal(calc.Data["radiation"][k], rad[k])
    assert calc.Args["mode"] == "far" and calc.Args["dtype"] == "double"
    assert np.array_equal(calc.Args["omega"], omega)
    assert list(calc.snap_iterations) == [100] and calc.total_weight == 3.0
