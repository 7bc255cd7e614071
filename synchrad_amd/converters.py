"""Track preparation: NaN-gapped particle histories to CSR batches and track files.

Particle trackers (openPMD particle tracking, VSIM) hand over each particle's
history as arrays over the output iterations, with NaN where the particle is
absent (synchrad/converters.py:74-82 replaces missing iterations by NaN).  The
reference cuts every history at its NaNs into pieces -- each maximal run of
iterations whose weight is not NaN, starting at iteration ``it_start`` with
the weight of its first sample (``split_track_by_nans``,
synchrad/converters.py:295-351) -- keeps the pieces longer than
``shortest_track`` samples and writes them to its track-file layout
(synchrad/converters.py:88-127).

Here the cut is one vectorised pass over a whole (particles x iterations)
block, and its result is a CSR batch -- the layout the spectrum engine
consumes directly -- which can also be written to a track file that
``SynchRad.calculate_spectrum(file_tracks=...)`` (or the reference) reads.
The openPMD / VSIM readers themselves need ``openpmd-viewer``, which this
build does not ship; :func:`tracks_from_arrays` is what follows their reads.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import engine

COORDS = ("x", "y", "z", "ux", "uy", "uz")


@dataclass
class CSRTracks:
    """Track pieces, CSR: piece i owns samples [offsets[i], offsets[i+1])."""
    offsets: np.ndarray     # int64 [n + 1]
    x: np.ndarray           # float64 [offsets[-1]] (and y, z, ux, uy, uz)
    y: np.ndarray
    z: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray
    weight: np.ndarray      # float64 [n]: the weight at the piece's first sample
    it_start: np.ndarray    # int64 [n]: the iteration of the piece's first sample
    source: np.ndarray      # int64 [n]: the row (particle) the piece came from

    @property
    def n_tracks(self) -> int:
        return int(self.offsets.size - 1)

    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)

    def it_range(self):
        """[first it_start, last it_start + length] over the pieces, as the
        converters write misc/it_range (synchrad/converters.py:114-126)."""
        if self.n_tracks == 0:
            return np.array([0, 0], dtype=np.int64)
        return np.array([self.it_start.min(), (self.it_start + self.lengths()).max()], dtype=np.int64)

    def track_list(self) -> List[list]:
        """Reference-style lists [x, y, z, ux, uy, uz, w, it_start]."""
        out = []
        for i in range(self.n_tracks):
            a, b = self.offsets[i], self.offsets[i + 1]
            out.append([getattr(self, c)[a:b] for c in COORDS] + [float(self.weight[i]), int(self.it_start[i])])
        return out

    def host_batch(self, dt: float, it_range=None, nSnaps: int = 1) -> engine.HostBatch:
        """The pieces as an engine batch, without a detour through lists."""
        return engine.pack_csr(self.offsets, self.x, self.y, self.z, self.ux, self.uy, self.uz, self.weight,
                               self.it_start.astype(np.uint32), dt=dt, it_range=it_range, nSnaps=nSnaps)


def split_histories(x, y, z, ux, uy, uz, w, shortest_track: int = 0) -> CSRTracks:
    """Cut particle histories at their NaN weights.

    Each argument is (n_particles, n_iterations) (or one history of
    n_iterations).  A piece is a maximal run of iterations of one particle
    whose weight is not NaN; pieces of at most ``shortest_track`` samples are
    dropped.  Pieces are ordered by particle, then by iteration."""
    cols = [np.atleast_2d(np.asarray(a, dtype=np.float64)) for a in (x, y, z, ux, uy, uz, w)]
    shape = cols[-1].shape
    if any(c.shape != shape for c in cols):
        raise ValueError("all history arrays must have the same shape")
    n_p, n_it = shape
    present = ~np.isnan(cols[-1])
    # run boundaries along each row: a sample opens a run when the one before
    # it (or the row start) is absent, and closes it when the one after is
    padded = np.zeros((n_p, n_it + 2), dtype=bool)
    padded[:, 1:-1] = present
    opens = padded[:, 1:-1] & ~padded[:, :-2]
    closes = padded[:, 1:-1] & ~padded[:, 2:]
    row_o, it_o = np.nonzero(opens)           # row-major: by particle, then iteration
    row_c, it_c = np.nonzero(closes)
    lengths = it_c - it_o + 1
    keep = lengths > shortest_track
    row, first, lengths = row_o[keep], it_o[keep], lengths[keep]
    offsets = np.zeros(row.size + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    # sample k of piece i sits at (row[i], first[i] + k)
    piece = np.repeat(np.arange(row.size), lengths)
    step = np.arange(offsets[-1], dtype=np.int64) - offsets[piece]
    r_idx, c_idx = row[piece], first[piece] + step
    data = [c[r_idx, c_idx] for c in cols[:6]]
    weight = cols[-1][row, first]
    return CSRTracks(offsets, *data, weight=weight, it_start=first.astype(np.int64), source=row.astype(np.int64))


def split_track_by_nans(x, y, z, ux, uy, uz, w) -> List[list]:
    """One history cut at its NaN weights, as the reference returns it
    (synchrad/converters.py:295-351): [x, y, z, ux, uy, uz, w0, it_start]
    per non-empty piece."""
    pieces = split_histories(x, y, z, ux, uy, uz, w)
    return [[*t[:6], np.float64(t[6]), np.int64(t[7])] for t in pieces.track_list()]


def write_track_file(fname: str, tracks: CSRTracks, cdt: float, cdt_array=None,
                     propagation_direction: str = "z") -> None:
    """The reference's track-file layout (synchrad/converters.py:101-127):
    tracks/<i>/{x, y, z, ux, uy, uz} (1-D), tracks/<i>/{w, it_start}
    (scalars), misc/cdt, misc/cdt_array (when given), misc/N_particles,
    misc/it_range and misc/propagation_direction."""
    from .calc import _h5
    with _h5(fname, "w") as f:
        for i in range(tracks.n_tracks):
            a, b = int(tracks.offsets[i]), int(tracks.offsets[i + 1])
            for c in COORDS:
                f[f"tracks/{i:d}/{c}"] = getattr(tracks, c)[a:b]
            f[f"tracks/{i:d}/w"] = np.float64(tracks.weight[i])
            f[f"tracks/{i:d}/it_start"] = np.int64(tracks.it_start[i])
        f["misc/cdt"] = float(cdt)
        if cdt_array is not None:
            f["misc/cdt_array"] = np.asarray(cdt_array, dtype=np.float64)
        f["misc/N_particles"] = int(tracks.n_tracks)
        f["misc/it_range"] = tracks.it_range()
        f["misc/propagation_direction"] = propagation_direction


def tracks_from_arrays(x, y, z, ux, uy, uz, w, t=None, cdt: Optional[float] = None,
                       fname: Optional[str] = None, shortest_track: int = 8,
                       z_is_xi: bool = False) -> CSRTracks:
    """The part of the openPMD converter that follows its reads
    (synchrad/converters.py:65-127): histories (n_particles, n_iterations)
    with NaN where a particle is absent, the output times ``t`` (seconds;
    c*dt is taken from its first step, c*dt per step goes to misc/cdt_array),
    cut into pieces of more than ``shortest_track`` samples.  With
    ``z_is_xi`` the z histories are co-moving coordinates and c*t of each
    sample's iteration is added back.  Writes ``fname`` when given and
    returns the pieces."""
    from scipy.constants import c as c_light
    zz = np.atleast_2d(np.asarray(z, dtype=np.float64))
    if z_is_xi:
        if t is None:
            raise ValueError("z_is_xi needs the output times t")
        zz = zz + c_light * np.asarray(t, dtype=np.float64)[None, :]
    pieces = split_histories(x, y, zz, ux, uy, uz, w, shortest_track=shortest_track)
    cdt_array = None
    if t is not None:
        t = np.asarray(t, dtype=np.float64)
        cdt_array = np.diff(t) * c_light
        if cdt is None:
            cdt = float(cdt_array[0])
    if fname is not None:
        if cdt is None:
            raise ValueError("writing a track file needs cdt (or t)")
        write_track_file(fname, pieces, cdt, cdt_array)
    return pieces
