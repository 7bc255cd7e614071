"""NaN-gapped particle histories to track pieces, CSR batches and track files
(synchrad/converters.py:65-127, 295-351)."""
import numpy as np
import pytest

from synchrad_amd import converters, engine, h5io

try:
    h5io.lib()
    _have_h5 = True
except ImportError:  # pragma: no cover - depends on the image
    _have_h5 = False


def _pieces_by_walking(w_row):
    """Runs of non-NaN weights of one history, walked sample by sample:
    (first iteration, length)."""
    out, start = [], None
    for it, v in enumerate(w_row):
        if not np.isnan(v) and start is None:
            start = it
        if np.isnan(v) and start is not None:
            out.append((start, it - start))
            start = None
    if start is not None:
        out.append((start, len(w_row) - start))
    return out


def _histories(n_p=7, n_it=40, seed=3):
    rng = np.random.default_rng(seed)
    cols = [rng.standard_normal((n_p, n_it)) for _ in range(6)]
    w = rng.uniform(0.5, 2.0, (n_p, n_it))
    w[rng.random((n_p, n_it)) < 0.2] = np.nan
    w[0, :] = np.nan                 # a particle never present
    w[1, :] = 1.5                    # always present
    w[2, :5] = np.nan                # leading gap
    w[3, -3:] = np.nan               # trailing gap
    w[4, ::2] = np.nan               # alternating: one-sample pieces
    return cols, w


def test_pieces_are_the_runs_of_present_samples():
    cols, w = _histories()
    pieces = converters.split_histories(*cols, w)
    got = list(zip(pieces.source.tolist(), pieces.it_start.tolist(), pieces.lengths().tolist()))
    want = [(p, s, n) for p in range(w.shape[0]) for s, n in _pieces_by_walking(w[p])]
    assert got == want
    for i, (p, s, n) in enumerate(want):
        a, b = pieces.offsets[i], pieces.offsets[i + 1]
        for c, name in enumerate(converters.COORDS):
            assert np.array_equal(getattr(pieces, name)[a:b], cols[c][p, s:s + n])
        assert pieces.weight[i] == w[p, s]
    assert (0, 0, 0) not in got and not any(p == 0 for p, _, _ in got)


def test_shortest_track_drops_short_pieces():
    cols, w = _histories()
    pieces = converters.split_histories(*cols, w, shortest_track=3)
    assert (pieces.lengths() > 3).all()
    all_pieces = converters.split_histories(*cols, w)
    assert pieces.n_tracks == int((all_pieces.lengths() > 3).sum())


def test_single_history_as_the_reference_returns_it():
    """[x, y, z, ux, uy, uz, w0, it_start] per non-empty piece; NaNs at both
    ends and in a row leave no empty pieces."""
    n = 12
    w = np.ones(n)
    w[[0, 4, 5, 11]] = np.nan
    w[6] = 3.0
    xs = [np.arange(n, dtype=float) + 10 * k for k in range(6)]
    out = converters.split_track_by_nans(*xs, w)
    assert [(int(t[7]), t[0].size, float(t[6])) for t in out] == [(1, 3, 1.0), (6, 5, 3.0)]
    assert np.array_equal(out[1][3], xs[3][6:11])
    assert converters.split_track_by_nans(*xs, np.full(n, np.nan)) == []
    full = converters.split_track_by_nans(*xs, np.ones(n))
    assert len(full) == 1 and int(full[0][7]) == 0 and full[0][0].size == n


def test_csr_pieces_feed_the_engine_like_lists():
    cols, w = _histories()
    pieces = converters.split_histories(*cols, w, shortest_track=1)
    a = pieces.host_batch(dt=0.01, it_range=tuple(pieces.it_range()), nSnaps=2)
    b = engine.pack_tracks(pieces.track_list(), dt=0.01, it_range=tuple(pieces.it_range()), nSnaps=2)
    for name in ("offsets", "x", "y", "z", "ux", "uy", "uz", "weight", "it_start", "it_end", "it_snaps"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    lo, hi = pieces.it_range()
    assert lo == pieces.it_start.min() and hi == max(s + n for s, n in zip(pieces.it_start, pieces.lengths()))


@pytest.mark.skipif(not _have_h5, reason="no HDF5 C library")
def test_track_file_written_in_the_reference_layout(tmp_path):
    from synchrad_amd.calc import read_track_file
    cols, w = _histories(n_p=9, n_it=60, seed=8)
    t = 1e-15 * np.arange(60) * (1 + 1e-3 * np.arange(60))
    path = str(tmp_path / "tracks.h5")
    pieces = converters.tracks_from_arrays(*cols, w, t=t, fname=path, shortest_track=8)
    assert (pieces.lengths() > 8).all() and pieces.n_tracks > 0
    with h5io.File(path, "r") as f:
        assert int(f["misc/N_particles"][()]) == pieces.n_tracks
        assert np.array_equal(f["misc/it_range"][()], pieces.it_range())
        assert f["misc/propagation_direction"][()] in (b"z", "z")
        np.testing.assert_allclose(f["misc/cdt_array"][()], np.diff(t) * 299792458.0, rtol=1e-15)
        assert sorted(f["tracks/0"].keys()) == sorted(["x", "y", "z", "ux", "uy", "uz", "w", "it_start"])
    tracks, cdt, it_range = read_track_file(path, None)
    assert cdt == pytest.approx((t[1] - t[0]) * 299792458.0, rel=1e-15)
    assert tuple(it_range) == tuple(pieces.it_range())
    for got, want in zip(tracks, pieces.track_list()):
        for k in range(6):
            assert np.array_equal(got[k], want[k])
        assert float(got[6]) == want[6] and int(got[7]) == want[7]


def test_comoving_z_gets_c_t_added():
    cols, w = _histories(n_p=5, n_it=10)
    w[:] = 1.0
    t = np.arange(10) * 1e-15
    pieces = converters.tracks_from_arrays(*cols, w, t=t, shortest_track=0, z_is_xi=True)
    np.testing.assert_allclose(pieces.z[:10], cols[2][0] + 299792458.0 * t, rtol=1e-15)
