"""Device-side engine: CSR track batches, spectral grids and the batched call.

This is the layer ``SynchRad.calculate_spectrum`` sits on.  It replaces the
reference's per-particle host loop (synchrad/calc.py:257-267), its per-track
uploads (``_track_to_device``, synchrad/calc.py:579-603) and its per-particle
launches (``_process_track``, synchrad/calc.py:292-353) with one upload of a
CSR-packed batch and one stream-ordered library call per spectrum.

Device buffers are torch tensors on the current HIP device; the library only
borrows their pointers.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _abi

PRECISIONS = {"double": _abi.SR_PREC_F64, "f64": _abi.SR_PREC_F64,
              "float": _abi.SR_PREC_SPLIT, "split": _abi.SR_PREC_SPLIT}

# synchrad/calc.py:458-465
COMP_KEYS = {
    "cartesian": ["x", "y", "z"],
    "cartesian_complex": ["xre", "xim", "yre", "yim", "zre", "zim"],
    "spheric": ["r", "theta", "phi"],
    "spheric_complex": ["rre", "rim", "thetare", "thetaim", "phire", "phiim"],
    "total": ["total"],
}


def _dev(device=None) -> torch.device:
    if device is None:
        if not torch.cuda.is_available():
            raise _abi.SynchradError("no HIP device is visible; synchrad_amd runs on MI355X only")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def snap_iterations(it_range, nSnaps: int) -> np.ndarray:
    """``np.linspace(it0, it1, nSnaps+1, dtype=uint32)[1:]``, synchrad/calc.py:626-630."""
    return np.ascontiguousarray(
        np.linspace(*(tuple(it_range) + (int(nSnaps) + 1,)), dtype=np.uint32)[1:])


# ---------------------------------------------------------------------------
# host packing
# ---------------------------------------------------------------------------
@dataclass
class HostBatch:
    """CSR-packed tracks on the host (numpy)."""
    offsets: np.ndarray       # int64 [Np+1]
    x: np.ndarray             # float64 [Npts] (also y, z, ux, uy, uz)
    y: np.ndarray
    z: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray
    weight: np.ndarray        # float64 [Np]
    it_start: np.ndarray      # uint32 [Np]
    it_end: np.ndarray        # uint32 [Np]
    it_snaps: np.ndarray      # uint32 [Np*nSnaps] or [nSnaps]
    n_snaps: int
    snaps_stride: int
    dt: float
    # the upload buffer (uint8) when the arrays were written straight into it
    # (pack_tracks): the device upload then copies the tracks only once
    packed: Optional[np.ndarray] = field(default=None, repr=False, compare=False)

    @property
    def n_particles(self) -> int:
        return int(self.offsets.size - 1)

    @property
    def n_points(self) -> int:
        return int(self.offsets[-1])

    def n_intervals(self) -> int:
        """Sum over particles of the active step count min(n_p-1, it_end-1)."""
        n = np.diff(self.offsets)
        act = np.minimum(n - 1, self.it_end.astype(np.int64) - 1)
        return int(np.clip(act, 0, None).sum())


def pack_csr(offsets, x, y, z, ux, uy, uz, weight, it_start=None, *, dt: float,
             it_range=None, nSnaps: int = 1) -> HostBatch:
    """Pack already-concatenated tracks.  Time alignment as _process_track
    (synchrad/calc.py:296-307): without ``it_range`` every particle runs its own
    range (it_start = 0, it_end = n_p, snapshots over [0, n_p]); with it, all
    particles share it_end = it_range[-1] and one snapshot list."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    Np = offsets.size - 1
    if Np < 0 or (Np > 0 and (offsets[0] != 0 or np.any(np.diff(offsets) < 0))):
        raise ValueError("offsets must start at 0 and be non-decreasing")
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, ux, uy, uz)]
    for a in arrs:
        if a.size != offsets[-1]:
            raise ValueError("track arrays must have offsets[-1] samples")
    weight = np.array(np.broadcast_to(np.asarray(weight, dtype=np.float64), (Np,)))
    nSnaps = int(nSnaps)
    if nSnaps < 1:
        raise ValueError("nSnaps must be >= 1")
    lengths = np.diff(offsets)
    if it_range is None:
        its = np.zeros(Np, dtype=np.uint32)
        ite = lengths.astype(np.uint32)
        snaps = np.zeros((Np, nSnaps), dtype=np.uint32)
        for n in np.unique(lengths):
            snaps[lengths == n] = snap_iterations((0, int(n)), nSnaps)
        snaps = snaps.reshape(-1)
        stride = nSnaps
    else:
        it_range = tuple(int(v) for v in it_range)
        if it_range[-1] <= 0:
            raise ValueError("it_range end must be >= 1 (the reference loops 2^32 times on 0)")
        its = (np.zeros(Np, dtype=np.uint32) if it_start is None
               else np.array(np.broadcast_to(np.asarray(it_start).astype(np.uint32), (Np,))))
        ite = np.full(Np, it_range[-1], dtype=np.uint32)
        snaps = snap_iterations(it_range, nSnaps)
        stride = 0
    if not np.isfinite(dt) or dt == 0:
        raise ValueError("timeStep must be finite and non-zero")
    return HostBatch(offsets, *arrs, weight, its, ite, np.ascontiguousarray(snaps, dtype=np.uint32),
                     nSnaps, stride, float(dt))


def pack_tracks(tracks: Sequence[Sequence], *, dt: float, it_range=None, nSnaps: int = 1) -> HostBatch:
    """Pack reference-style track lists ``[x, y, z, ux, uy, uz, w, it_start]`` (or
    the 7-item form, synchrad/calc.py:580-584).  The six coordinate and
    momentum columns are concatenated straight into the batch's upload buffer
    (``HostBatch.packed``), so the tracks are copied once on the host."""
    Np = len(tracks)
    lengths = np.array([np.asarray(t[0]).size for t in tracks], dtype=np.int64)
    offsets = np.zeros(Np + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    n_snap_entries = Np * int(nSnaps) if it_range is None else int(nSnaps)
    layout, total = _layout(Np, int(offsets[-1]), n_snap_entries, True)
    buf = np.empty(total, dtype=np.uint8)
    cols = []
    for c, name in enumerate(("x", "y", "z", "ux", "uy", "uz")):
        at, nbytes = layout[name]
        col = buf[at:at + nbytes].view(np.float64)
        if Np:
            np.concatenate([np.asarray(t[c], dtype=np.float64).ravel() for t in tracks], out=col)
        cols.append(col)
    w = np.array([float(t[6]) for t in tracks], dtype=np.float64)
    its = np.array([int(np.uint32(t[7])) if len(t) == 8 else 0 for t in tracks], dtype=np.uint32)
    hb = pack_csr(offsets, *cols, w, its, dt=dt, it_range=it_range, nSnaps=nSnaps)
    hb.packed = buf
    return hb


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------
_PACK_ALIGN = 256


def _layout(n_particles: int, n_points: int, n_snap_entries: int, with_tracks: bool):
    """Byte offsets of a batch's arrays in one contiguous buffer: offsets
    (int64), the six coordinate and momentum arrays and the weights (float64;
    left out when the caller supplies them as device tensors), it_start,
    it_end and it_snaps (uint32), each 256-byte aligned."""
    tracks = [(c, n_points * 8) for c in ("x", "y", "z", "ux", "uy", "uz")] + [("weight", n_particles * 8)]
    parts = [("offsets", (n_particles + 1) * 8)] + (tracks if with_tracks else []) + \
        [("it_start", n_particles * 4), ("it_end", n_particles * 4), ("it_snaps", n_snap_entries * 4)]
    layout, at = {}, 0
    for name, nbytes in parts:
        layout[name] = (at, nbytes)
        at += (nbytes + _PACK_ALIGN - 1) // _PACK_ALIGN * _PACK_ALIGN
    return layout, max(at, _PACK_ALIGN)


def _packed_layout(hb: HostBatch, with_tracks: bool = True):
    return _layout(hb.n_particles, hb.n_points, int(hb.it_snaps.size), with_tracks)


def _pack_host(hb: HostBatch, layout, total) -> np.ndarray:
    """The upload buffer of ``hb``: its own ``packed`` buffer when it has the
    layout's size (arrays already in place are not copied again), else a new one."""
    own = hb.packed is not None and hb.packed.size == total
    buf = hb.packed if own else np.empty(total, dtype=np.uint8)
    base = buf.ctypes.data
    for name, (at, nbytes) in layout.items():
        if own and nbytes and getattr(hb, name).ctypes.data == base + at:
            continue
        if name in ("it_start", "it_end", "it_snaps"):
            arr = np.ascontiguousarray(getattr(hb, name), dtype=np.uint32)
        elif name == "offsets":
            arr = np.ascontiguousarray(hb.offsets, dtype=np.int64)
        else:
            arr = np.ascontiguousarray(getattr(hb, name), dtype=np.float64)
        buf[at:at + nbytes] = arr.view(np.uint8)
    return buf


class DeviceBatch:
    """A HostBatch on the HIP device: every array in one device buffer, filled
    by one host-to-device copy.

    ``tensors`` may supply the coordinate, momentum and weight arrays as
    existing device tensors (float64, contiguous; x, y, z, ux, uy, uz, weight),
    which are then used in place -- the autograd path does this.  ``refill``
    loads another batch of the same sizes into the same device memory (so a
    captured launch sequence sees the new tracks)."""

    def __init__(self, hb: HostBatch, device=None, tensors: Optional[Sequence[torch.Tensor]] = None):
        dev = _dev(device)
        self.host = hb
        self.device = dev
        self._with_tracks = tensors is None
        self._layout, total = _packed_layout(hb, self._with_tracks)
        self._buf = torch.from_numpy(_pack_host(hb, self._layout, total)).to(dev)
        views = {}
        for name, (at, nbytes) in self._layout.items():
            raw = self._buf[at:at + nbytes]
            views[name] = raw.view(torch.int64 if name == "offsets" else
                                   torch.int32 if name.startswith("it_") else torch.float64)
        self.offsets = views["offsets"]
        self.it_start, self.it_end, self.it_snaps = views["it_start"], views["it_end"], views["it_snaps"]
        if tensors is not None:
            ts = list(tensors)
            for t, n in zip(ts, [hb.n_points] * 6 + [hb.n_particles]):
                if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() != n or t.device != dev:
                    raise ValueError("track tensors must be contiguous float64 on the batch device")
            self.x, self.y, self.z, self.ux, self.uy, self.uz, self.weight = ts
        else:
            self.x, self.y, self.z = views["x"], views["y"], views["z"]
            self.ux, self.uy, self.uz = views["ux"], views["uy"], views["uz"]
            self.weight = views["weight"]
        self._offsets_host = np.ascontiguousarray(hb.offsets, dtype=np.int64)

    def shape_key(self):
        """What a refill must keep: sample and particle counts, snapshot list
        form and the time step (the library plans and captures on these)."""
        hb = self.host
        return (hb.n_particles, hb.n_points, hb.n_snaps, hb.snaps_stride, int(hb.it_snaps.size), hb.dt)

    def refill(self, hb: HostBatch) -> None:
        """Load ``hb`` (same shape_key) into this batch's device memory."""
        old = self.host
        self.host = hb
        if self.shape_key() != (old.n_particles, old.n_points, old.n_snaps, old.snaps_stride,
                                int(old.it_snaps.size), old.dt):
            self.host = old
            raise ValueError("refill needs a batch of the same sizes, snapshot form and time step")
        layout, total = _packed_layout(hb, self._with_tracks)
        self._buf.copy_(torch.from_numpy(_pack_host(hb, layout, total)))
        self._offsets_host = np.ascontiguousarray(hb.offsets, dtype=np.int64)

    @property
    def n_particles(self) -> int:
        return self.host.n_particles

    def abi(self) -> _abi.SrTracks:
        hb = self.host
        t = _abi.SrTracks()
        t.n_particles = hb.n_particles
        t.n_points = hb.n_points
        t.offsets = self.offsets.data_ptr()
        t.offsets_host = self._offsets_host.ctypes.data
        for name in ("x", "y", "z", "ux", "uy", "uz"):
            setattr(t, name, getattr(self, name).data_ptr())
        t.weight = self.weight.data_ptr()
        t.it_start = self.it_start.data_ptr()
        t.it_end = self.it_end.data_ptr()
        t.it_snaps = self.it_snaps.data_ptr()
        t.n_snaps = hb.n_snaps
        t.snaps_stride = hb.snaps_stride
        t.dt = hb.dt
        return t


class DeviceGrid:
    """Spectral axes on the device, formed as SynchRad._init_data forms them
    (synchrad/calc.py:494-512): omega times 2*pi, and numpy sin/cos of the
    double-precision angle axes.

    ``screen='cartesian'`` (near field only; the reference's screen is polar)
    takes ``ax1`` as the screen's x positions and ``phi`` as its y positions:
    the observer of (x_i, y_j) is (x_i, y_j, L_screen), and the spectra are
    laid out (nSnaps, n_y, n_x, nOmega)."""

    def __init__(self, mode: str, omega: np.ndarray, ax1: np.ndarray, phi: np.ndarray,
                 L_screen: Optional[float] = None, device=None, screen: str = "polar"):
        dev = _dev(device)
        if mode not in _abi.MODES:
            raise ValueError(f"mode must be 'far' or 'near', got {mode!r}")
        if screen not in ("polar", "cartesian") or (screen == "cartesian" and mode != "near"):
            raise ValueError("screen must be 'polar', or 'cartesian' in the near field")
        self.mode = mode
        self.screen = screen
        self.device = dev
        omega = np.asarray(omega, dtype=np.float64)
        ax1 = np.asarray(ax1, dtype=np.float64)
        phi = np.asarray(phi, dtype=np.float64)
        self.n_omega, self.n1, self.n_phi = omega.size, ax1.size, phi.size

        def f64(a):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)

        self.omega_host = omega
        w = np.double(2 * np.pi) * omega
        self.omega = f64(w)
        self.omega_step = uniform_step(w)
        self.screen_y = None
        if screen == "cartesian":
            self.sin_phi = self.cos_phi = None
            self.screen_y = f64(phi)
        else:
            self.sin_phi = f64(np.sin(phi))
            self.cos_phi = f64(np.cos(phi))
        self.sin_theta = self.cos_theta = self.radius = None
        if mode == "far":
            self.sin_theta = f64(np.sin(ax1))
            self.cos_theta = f64(np.cos(ax1))
            self.l_screen = 0.0
        else:
            if L_screen is None:
                raise ValueError("near field needs L_screen")
            self.radius = f64(ax1)
            self.l_screen = float(L_screen)

    @property
    def n_grid(self) -> int:
        return self.n_omega * self.n1 * self.n_phi

    def abi(self) -> _abi.SrGrid:
        g = _abi.SrGrid()
        g.mode = _abi.MODES[self.mode]
        g.n_omega, g.n_theta, g.n_phi = self.n_omega, self.n1, self.n_phi
        g.omega = self.omega.data_ptr()
        if self.screen_y is not None:
            g.screen_y = self.screen_y.data_ptr()
        else:
            g.sin_phi = self.sin_phi.data_ptr()
            g.cos_phi = self.cos_phi.data_ptr()
        if self.mode == "far":
            g.sin_theta = self.sin_theta.data_ptr()
            g.cos_theta = self.cos_theta.data_ptr()
        else:
            g.radius = self.radius.data_ptr()
        g.l_screen = self.l_screen
        g.omega_step = self.omega_step
        return g

    def form_factor(self, sigma_particle: float) -> torch.Tensor:
        """exp(-(2 pi omega sigma)^2 / 2), synchrad/calc.py:475-480 (made once
        per sigma and kept, so a graph capture finds it already on the device)."""
        key = float(sigma_particle)
        cache = self.__dict__.setdefault("_ff", {})
        if key not in cache:
            ff = np.exp(np.double(-0.5) * (np.double(2 * np.pi) * self.omega_host * np.double(key)) ** 2)
            cache[key] = torch.from_numpy(np.ascontiguousarray(ff)).to(self.device)
        return cache[key]


def uniform_step(w: np.ndarray, rtol: float = 1e-14) -> float:
    """The constant step of ``w`` if every sample lies within ``rtol*max|w|``
    of w[0] + i*step (linear grids built by np.r_/linspace do), else 0.  The
    recurrence sweep uses it; the bound keeps its phase error below
    rtol*max|omega*tau|."""
    n = w.size
    if n < 2:
        return 0.0
    step = (w[-1] - w[0]) / (n - 1)
    if step == 0.0:
        return 0.0
    dev = np.max(np.abs(w - (w[0] + step * np.arange(n))))
    return float(step) if dev <= rtol * np.max(np.abs(w)) else 0.0


def _launch_stream(stream, device):
    """The stream a library call is ordered on: ``stream``, or the current
    one.  A stream other than the current one first waits for the current
    stream's work, which made or cleared the call's inputs, outputs and
    workspace (torch allocates and fills them there); the caller orders its
    own later use of the outputs after ``stream``."""
    cur = torch.cuda.current_stream(device)
    if stream is None or stream == cur:
        return cur
    stream.wait_stream(cur)
    return stream


SWEEPS = {"auto": _abi.SR_SWEEP_AUTO, "direct": _abi.SR_SWEEP_DIRECT,
          "recurrence": _abi.SR_SWEEP_RECURRENCE}


@dataclass
class Tuning:
    omega_per_lane: int = 0
    waves_per_block: int = 0
    slices: int = 0
    persistent: int = 0
    sweep: str = "auto"
    tile_order: int = 0    # 0: large far-field launches run their heaviest tiles first; 1: index order


def spectrum(batch: DeviceBatch, grid: DeviceGrid, comp: str = "total", precision="double",
             outs: Optional[List[torch.Tensor]] = None, sigma_particle: float = 0.0,
             tuning=None, stream=None, workspace: Optional[torch.Tensor] = None) -> List[torch.Tensor]:
    """Add one batch's spectrum to ``outs`` (allocated zero if None) and return
    them.  Each output is float64 (nSnaps, nPhi, nTheta|nR, nOmega) on the
    device -- the reference's device layout (kernel_farfield.cl:102).
    Stream-ordered on the current (or given) torch stream; no synchronisation.

    ``tuning``: None (the library's own plan), a :class:`Tuning`, or "auto"
    (the autotuner's plan for this batch and grid, timed once and cached).
    ``workspace``: a preallocated uint8 device tensor to use instead of
    allocating one (needed under graph capture)."""
    L = _abi.load()
    if isinstance(tuning, str):
        if tuning != "auto":
            raise ValueError("tuning must be None, a Tuning or 'auto'")
        from . import autotune
        prec_name = precision if isinstance(precision, str) else ("double" if precision == 0 else "float")
        tuning = autotune.tune(batch, grid, comp=comp, precision=prec_name)[0]
    if comp not in _abi.COMPS:
        raise ValueError(f"unknown comp {comp!r}")
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    n_out = L.sr_num_outputs(_abi.MODES[grid.mode], _abi.COMPS[comp])
    if n_out < 0:
        _abi.check(n_out)
    shape = (batch.host.n_snaps, grid.n_phi, grid.n1, grid.n_omega)
    if outs is None:
        outs = [torch.zeros(shape, dtype=torch.float64, device=grid.device) for _ in range(n_out)]
    if len(outs) != n_out:
        raise ValueError(f"{comp} needs {n_out} output arrays")
    for o in outs:
        if tuple(o.shape) != shape or o.dtype != torch.float64 or not o.is_contiguous():
            raise ValueError(f"outputs must be contiguous float64 of shape {shape}")
    tun = tuning or Tuning()
    opts = _abi.SrOpts()
    opts.comp = _abi.COMPS[comp]
    opts.precision = prec
    opts.omega_per_lane = tun.omega_per_lane
    opts.waves_per_block = tun.waves_per_block
    opts.slices = tun.slices
    opts.persistent = tun.persistent
    opts.tile_order = tun.tile_order
    opts.sweep = SWEEPS[tun.sweep]
    ff = None
    if comp == "cartesian_complex" and grid.mode == "far":
        ff = grid.form_factor(sigma_particle)
        opts.form_factor = ff.data_ptr()
    tr = batch.abi()
    gr = grid.abi()
    nbytes = ctypes.c_size_t(0)
    _abi.check(L.sr_workspace_bytes(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts),
                                    ctypes.byref(nbytes)))
    ws = None
    if nbytes.value:
        if workspace is not None:
            if workspace.numel() < nbytes.value or workspace.dtype != torch.uint8:
                raise ValueError(f"workspace needs {nbytes.value} uint8 elements")
            ws = workspace
        else:
            ws = torch.empty(nbytes.value, dtype=torch.uint8, device=grid.device)
    out_ptrs = (ctypes.c_void_p * n_out)(*[o.data_ptr() for o in outs])
    st = _launch_stream(stream, grid.device)
    _abi.check(L.sr_spectrum_batch(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts), out_ptrs,
                                   None if ws is None else ctypes.c_void_p(ws.data_ptr()),
                                   nbytes.value, ctypes.c_void_p(st.cuda_stream)))
    # temporaries are freed to torch's stream-ordered caching allocator; tell
    # it about the launch stream when that is not the one they were made on
    if st != torch.cuda.current_stream(grid.device):
        for t in (ff, ws):
            if t is not None:
                t.record_stream(st)
    return outs


def workspace_bytes(batch: DeviceBatch, grid: DeviceGrid, comp: str = "total", precision="double",
                    tuning: Optional[Tuning] = None) -> int:
    """Bytes of workspace :func:`spectrum` needs for this request."""
    L = _abi.load()
    tun = tuning or Tuning()
    opts = _abi.SrOpts()
    opts.comp = _abi.COMPS[comp]
    opts.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    opts.omega_per_lane = tun.omega_per_lane
    opts.waves_per_block = tun.waves_per_block
    opts.slices = tun.slices
    opts.persistent = tun.persistent
    opts.tile_order = tun.tile_order
    opts.sweep = SWEEPS[tun.sweep]
    opts.form_factor = 1 if (comp == "cartesian_complex" and grid.mode == "far") else None
    tr, gr = batch.abi(), grid.abi()
    nbytes = ctypes.c_size_t(0)
    _abi.check(L.sr_workspace_bytes(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts),
                                    ctypes.byref(nbytes)))
    return int(nbytes.value)


class SpectrumGraph:
    """One spectrum captured as a HIP graph and replayed.

    Outputs and workspace are allocated once; each replay clears the outputs
    with a fill kernel (no memset node) and launches the library's kernels
    exactly as :func:`spectrum` does, so repeated spectra of the same batch
    (or of a batch whose device arrays were refilled in place) cost one graph
    launch."""

    def __init__(self, batch: DeviceBatch, grid: DeviceGrid, comp: str = "total", precision="float",
                 tuning: Optional[Tuning] = None, sigma_particle: float = 0.0, warm: bool = True):
        self.batch, self.grid, self.comp, self.precision = batch, grid, comp, precision
        if isinstance(tuning, str):
            from . import autotune
            tuning = autotune.tune(batch, grid, comp=comp, precision=precision)[0]
        self.tuning = tuning or Tuning()
        L = _abi.load()
        n_out = L.sr_num_outputs(_abi.MODES[grid.mode], _abi.COMPS[comp])
        shape = (batch.host.n_snaps, grid.n_phi, grid.n1, grid.n_omega)
        self.outs = [torch.zeros(shape, dtype=torch.float64, device=grid.device) for _ in range(n_out)]
        nb = workspace_bytes(batch, grid, comp, precision, self.tuning)
        self.workspace = torch.empty(max(nb, 1), dtype=torch.uint8, device=grid.device)
        self.ff = grid.form_factor(sigma_particle) if (comp == "cartesian_complex" and grid.mode == "far") else None
        self.sigma = sigma_particle
        # warm up on a side stream (loads the code objects), then capture;
        # warm=False when this process already ran the same launch
        if warm:
            s = torch.cuda.Stream(device=grid.device)
            s.wait_stream(torch.cuda.current_stream(grid.device))
            with torch.cuda.stream(s):
                self._body()
            torch.cuda.current_stream(grid.device).wait_stream(s)
        torch.cuda.synchronize(grid.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body()

    def _body(self):
        for o in self.outs:
            o.zero_()
        spectrum(self.batch, self.grid, comp=self.comp, precision=self.precision, outs=self.outs,
                 sigma_particle=self.sigma, tuning=self.tuning, workspace=self.workspace)

    def replay(self) -> List[torch.Tensor]:
        self.graph.replay()
        return self.outs


def _adjoint_opts(comp, precision, tuning):
    tun = tuning or Tuning()
    opts = _abi.SrOpts()
    opts.comp = _abi.COMPS[comp]
    opts.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    opts.waves_per_block = tun.waves_per_block
    opts.slices = tun.slices
    return opts


def field_bytes(batch: DeviceBatch, grid: DeviceGrid) -> int:
    """Bytes of every particle's complex far field (6 doubles per particle and
    spectral point), as :func:`far_fields` stores them."""
    return 48 * batch.host.n_particles * grid.n_grid


def far_fields(batch: DeviceBatch, grid: DeviceGrid, comp: str = "total", precision="double",
               outs: Optional[List[torch.Tensor]] = None, stream=None):
    """Every particle's complex far field (float64 [6, n_particles, n_grid]:
    x, y, z as re/im pairs) and the spectra of ``comp`` ('total' or
    'cartesian', one snapshot) formed from them.  Returns ``(fields, outs)``;
    ``outs`` are added to when given.  The backward of a spectrum passes the
    fields to :func:`far_adjoint` instead of recomputing them."""
    L = _abi.load()
    if grid.mode != "far":
        raise ValueError("fields cover the far field")
    n_out = {"total": 1, "cartesian": 3}.get(comp)
    if n_out is None:
        raise ValueError("fields cover comp 'total' and 'cartesian'")
    dev = grid.device
    shape = (1, grid.n_phi, grid.n1, grid.n_omega)
    if outs is None:
        outs = [torch.zeros(shape, dtype=torch.float64, device=dev) for _ in range(n_out)]
    elif len(outs) != n_out or any(tuple(o.shape) != shape or o.dtype != torch.float64 or not o.is_contiguous()
                                   for o in outs):
        raise ValueError(f"{comp} needs {n_out} contiguous float64 outputs of shape {shape}")
    fields = torch.empty((6, batch.host.n_particles, grid.n_grid), dtype=torch.float64, device=dev)
    opts = _adjoint_opts(comp, precision, None)
    tr, gr = batch.abi(), grid.abi()
    nbytes = ctypes.c_size_t(0)
    _abi.check(L.sr_far_fields_workspace_bytes(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts),
                                               ctypes.byref(nbytes)))
    ws = torch.empty(max(int(nbytes.value), 1), dtype=torch.uint8, device=dev)
    optr = (ctypes.c_void_p * n_out)(*[o.data_ptr() for o in outs])
    st = _launch_stream(stream, dev)
    _abi.check(L.sr_far_fields_batch(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts),
                                     ctypes.c_void_p(fields.data_ptr()), optr, ctypes.c_void_p(ws.data_ptr()),
                                     nbytes.value, ctypes.c_void_p(st.cuda_stream)))
    if st != torch.cuda.current_stream(dev):
        ws.record_stream(st)
    return fields, outs


def far_adjoint(batch: DeviceBatch, grid: DeviceGrid, dS: List[torch.Tensor], comp: str = "total",
                precision="double", tuning: Optional[Tuning] = None, grads=None, stream=None,
                fields: Optional[torch.Tensor] = None):
    """dL/d(tracks) of the far-field spectrum given ``dS`` = dL/dS (device
    float64, one array for 'total', three for 'cartesian', each shaped like the
    spectrum: (1, nPhi, nTheta, nOmega)).  Returns ``(gx, gy, gz, gux, guy, guz,
    gw)`` device tensors (added to ``grads`` if given).  ``fields`` from
    :func:`far_fields` (same batch, grid and precision) skips recomputing them."""
    L = _abi.load()
    if grid.mode != "far":
        raise ValueError("the adjoint covers the far field")
    n_in = {"total": 1, "cartesian": 3}.get(comp)
    if n_in is None:
        raise ValueError("the adjoint covers comp 'total' and 'cartesian'")
    if len(dS) != n_in:
        raise ValueError(f"{comp} needs {n_in} dS arrays")
    shape = (1, grid.n_phi, grid.n1, grid.n_omega)
    dS = [d.to(dtype=torch.float64).contiguous() for d in dS]
    for d in dS:
        if tuple(d.shape) != shape:
            raise ValueError(f"dS arrays must have shape {shape}")
    dev = grid.device
    npts, npart = batch.host.n_points, batch.host.n_particles
    if fields is not None and (tuple(fields.shape) != (6, npart, grid.n_grid) or fields.dtype != torch.float64
                               or not fields.is_contiguous()):
        raise ValueError(f"fields must be contiguous float64 of shape {(6, npart, grid.n_grid)}")
    if grads is None:
        grads = [torch.zeros(npts, dtype=torch.float64, device=dev) for _ in range(6)] + \
                [torch.zeros(npart, dtype=torch.float64, device=dev)]
    opts = _adjoint_opts(comp, precision, tuning)
    tr, gr = batch.abi(), grid.abi()
    nbytes = ctypes.c_size_t(0)
    size_fn = L.sr_far_adjoint_workspace_bytes if fields is None else L.sr_far_adjoint_fields_workspace_bytes
    _abi.check(size_fn(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts), ctypes.byref(nbytes)))
    ws = torch.empty(max(int(nbytes.value), 1), dtype=torch.uint8, device=dev)
    dptr = (ctypes.c_void_p * n_in)(*[d.data_ptr() for d in dS])
    st = _launch_stream(stream, dev)
    tail = [*[g.data_ptr() for g in grads], ctypes.c_void_p(ws.data_ptr()), nbytes.value,
            ctypes.c_void_p(st.cuda_stream)]
    if fields is None:
        _abi.check(L.sr_far_adjoint_batch(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts), dptr, *tail))
    else:
        _abi.check(L.sr_far_adjoint_from_fields(ctypes.byref(tr), ctypes.byref(gr), ctypes.byref(opts), dptr,
                                                ctypes.c_void_p(fields.data_ptr()), *tail))
    if st != torch.cuda.current_stream(dev):
        ws.record_stream(st)
    return tuple(grads)
