"""``SynchRad``: the reference's calculator API on the MI355X engine.

Drop-in for ``synchrad.calc.SynchRad`` (synchrad/calc.py:21-666): the same
constructor, ``calculate_spectrum`` signature, ``Args`` keys, ``Data['radiation']``
layout (float64 ``(nSnaps, nOmega, nTheta|nR, nPhi)``), track and spectrum file
layouts, and the private re-initialisation methods that the reference's test
scripts call (``_init_args``, ``_init_data``, ``_compile_kernels``).

What the calculator must produce is fixed by the reference; this module is
organised around the MI355X engine instead of the reference's particle loop:
the axes come from small node-construction helpers, every track (list or
file) is packed into one CSR batch, and one stream-ordered library call adds
the whole batch to the device spectra, which stay resident for the analysis
helpers (``Utilities``) until the next call.

Differences, all deliberate:

* ``dtype: 'double'`` runs the FP64 kernels.  ``dtype: 'float'`` runs the
  split-precision kernels: the phase and its reduction in FP64, amplitudes,
  sin/cos and compensated sums in FP32.  Host axes, tracks and ``timeStep``
  stay float64 in both modes (the reference casts them to float32 in its float
  mode, which costs several radians of phase on the betatron beam's 8000-step
  tracks).
* String options are compared with ``==`` (the reference uses ``is``).
* ``ctx`` and ``native`` are accepted and ignored (one HIP device per
  process); ``ctx=False`` still starts without a device for reloading files.
* Zero-length tracks contribute nothing (the reference reads out of bounds).
* A near-field call without ``L_screen`` raises ValueError (the reference
  prints a message and fails later).
* MPI is not used: ``rank`` is 0 and ``size`` is 1.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .utils import Utilities
from . import engine

try:  # h5 track/spectrum files: h5py when present, else the HDF5 C library
    import h5py  # type: ignore
    h5py_installed = True
except ImportError:  # pragma: no cover - depends on the image
    h5py = None
    h5py_installed = False

mpi_installed = False

TRACK_FIELDS = ("x", "y", "z", "ux", "uy", "uz", "w", "it_start")
# Args entries a spectrum file does not keep (synchrad/calc.py:280-282)
_NOT_SAVED = ("grid", "ctx")
_OMEGA_FEATURES = ("wavelengthGrid", "logGrid")


def _h5(path: str, mode: str):
    """An h5py.File, or the ctypes binding of libhdf5 used the same way."""
    if h5py_installed:
        return h5py.File(path, mode)
    from . import h5io
    return h5io.File(path, mode)


# ---------------------------------------------------------------------------
# axes (synchrad/calc.py:380-451 fixes them bit for bit)
# ---------------------------------------------------------------------------
def frequency_nodes(lo: float, hi: float, n: int, features: Sequence[str] = ()):
    """The omega nodes and, for 'wavelengthGrid', the wavelength nodes.

    The first recognised feature decides the spacing: none -- n nodes from lo
    to hi inclusive, evenly spaced; 'wavelengthGrid' -- wavelengths evenly
    spaced from 1/hi to 1/lo, omega their inverses (so omega descends);
    'logGrid' -- lo * exp(i * log(hi/lo)/(n-1)), i = 0..n-1."""
    spacing = next((f for f in features if f in _OMEGA_FEATURES), None)
    if spacing == "wavelengthGrid":
        lam = np.linspace(1.0 / hi, 1.0 / lo, n)
        return 1.0 / lam, lam
    if spacing == "logGrid":
        rate = np.log(hi / lo) / (n - 1.0)
        return lo * np.exp(rate * np.arange(n)), None
    return np.linspace(lo, hi, n), None


def inclusive_nodes(lo: float, hi: float, n: int) -> np.ndarray:
    """theta (far) or radius (near): n nodes, both ends included."""
    return np.linspace(lo, hi, n)


def periodic_nodes(lo: float, hi: float, n: int) -> np.ndarray:
    """phi: n nodes from lo in steps of (hi - lo)/n, hi excluded."""
    return lo + (hi - lo) / n * np.arange(n)


def first_step(nodes: np.ndarray):
    """The integration element of an axis: its first step, 1 for one node."""
    return nodes[1] - nodes[0] if nodes.size > 1 else np.double(1.0)


# ---------------------------------------------------------------------------
# files
# ---------------------------------------------------------------------------
def read_track_file(path: str, Np_max: Optional[int]):
    """Tracks in the converters' layout (synchrad/converters.py:101-127):
    tracks/<i>/<x|y|z|ux|uy|uz|w|it_start>, misc/cdt, misc/N_particles and an
    optional misc/it_range.  Returns (track lists, cdt, it_range or None)."""
    with _h5(path, "r") as f:
        cdt = f["misc/cdt"][()]
        file_range = tuple(f["misc/it_range"][()]) if "it_range" in f["misc"].keys() else None
        count = int(f["misc/N_particles"][()])
        if Np_max is not None:
            count = min(count, Np_max)
        tracks = [[f["tracks/%d/%s" % (ip, name)][()] for name in TRACK_FIELDS] for ip in range(count)]
    return tracks, cdt, file_range


def write_spectrum_file(path: str, radiation: Dict[str, np.ndarray], args: dict, snaps, total_weight) -> None:
    """radiation/<key>, Args/<key> (all but grid and ctx), snap_iterations and
    total_weight (synchrad/calc.py:274-290)."""
    with _h5(path, "w") as f:
        for name, arr in radiation.items():
            f["radiation/" + name] = arr
        for name, value in args.items():
            if name not in _NOT_SAVED:
                f["Args/" + name] = value
        f["snap_iterations"] = np.asarray(snaps)
        f["total_weight"] = total_weight


def read_spectrum_file(path: str):
    """(radiation, Args, snap_iterations, total_weight) of a spectrum file;
    strings come back as str (h5py 3 returns bytes, synchrad/calc.py:661-663)."""
    with _h5(path, "r") as f:
        radiation = {name: f["radiation/" + name][()] for name in f["radiation"].keys()}
        args = {}
        for name in f["Args"].keys():
            value = f["Args/" + name][()]
            args[name] = value.decode() if isinstance(value, bytes) else value
        return radiation, args, f["snap_iterations"][()], f["total_weight"][()]


def normalise_weights(tracks: List[list], how: Optional[str]) -> float:
    """Rescale the weight (item 6) of every track list in place -- 'mean' and
    'max' divide by that statistic of the weights, 'ones' sets 1 -- and return
    their sum (synchrad/calc.py:241-265: the caller's own lists change)."""
    if how in ("mean", "max"):
        stat = np.mean if how == "mean" else np.max
        scale = stat([t[6] for t in tracks])
        for t in tracks:
            t[6] /= scale
    elif how == "ones":
        for t in tracks:
            t[6] = 1.0
    total = 0.0
    for t in tracks:
        total += t[6]
    return total


class SynchRad(Utilities):
    """Spectral-angular radiation of point charges from their tracks."""

    def __init__(self, Args={}, file_spectrum=None):
        """See synchrad/calc.py:29-99 for the ``Args`` keys: ``grid``, ``mode``
        ('far'/'near'), ``dtype`` ('double'/'float'), ``ctx``, ``Features``
        ('wavelengthGrid'/'logGrid')."""
        self.rank, self.size = 0, 1
        if file_spectrum is not None:
            self._read_args(file_spectrum)
            return
        self._init_args(Args)
        self._init_comm()
        self._init_data()
        self._compile_kernels()

    # ------------------------------------------------------------------ API
    def calculate_spectrum(self, particleTracks=[], file_tracks=None, timeStep=None, comp="total",
                           L_screen=None, Np_max=None, it_range=None, nSnaps=1, sigma_particle=0,
                           weights_normalize=None, file_spectrum=None, verbose=True):
        """Compute the spectrum of ``particleTracks`` (or of ``file_tracks``),
        synchrad/calc.py:101-290.  Results land in ``self.Data['radiation']``."""
        if self.device is None:
            raise RuntimeError("this SynchRad was started without a device (ctx=False)")
        mode = self.Args["mode"]
        self.Args["sigma_particle"] = self.dtype(sigma_particle)
        if mode == "near":
            if L_screen is not None:
                self.Args["L_screen"] = L_screen
            if self.Args.get("L_screen") is None:
                raise ValueError("Define L_screen argument for near-field calculation")
        nSnaps = np.uint32(nSnaps)
        self._init_raditaion(comp, nSnaps)
        if timeStep is not None:
            self.Args["timeStep"] = self.dtype(timeStep)
        window = None if it_range is None else tuple(it_range)

        say = print if verbose else (lambda *a, **k: None)
        if file_tracks is not None:
            tracks, cdt, file_range = read_track_file(file_tracks, Np_max)
            self.Args["timeStep"] = self.dtype(cdt)
            if window is None and file_range is not None:
                window = file_range
                say("it_range from the input file will be used")
            elif window is None:
                say("Separate it_range for each track will be used")
            say("Tracks are loaded")
        else:
            if window is None:
                say("Separate it_range for each track will be used")
            tracks = particleTracks[:len(particleTracks) if Np_max is None else Np_max]
        if "timeStep" not in self.Args:
            raise ValueError("timeStep is required")

        self.total_weight = normalise_weights(tracks, weights_normalize)
        batch = engine.pack_tracks(tracks, dt=float(self.Args["timeStep"]), it_range=window,
                                   nSnaps=int(nSnaps))
        # the snapshot list the reference leaves behind: the shared window's,
        # or in per-track mode the last track's (synchrad/calc.py:297-301)
        if window is not None:
            self._set_snap_iterations(window, nSnaps)
        elif batch.n_particles:
            self._set_snap_iterations((0, int(np.diff(batch.offsets)[-1])), nSnaps)
        if batch.n_particles:
            self._integrate(batch, comp, float(self.Args["sigma_particle"]))
        self._spectr_from_device(nSnaps)

        if file_spectrum is not None:
            write_spectrum_file(file_spectrum, self.Data["radiation"], self.Args,
                                getattr(self, "snap_iterations", np.zeros(0, np.uint32)), self.total_weight)
            print(f"Spectrum is saved to {file_spectrum}")

    # ------------------------------------------------------------ internals
    def _integrate(self, batch, comp, sigma):
        """One library call adds the whole batch to the device spectra.

        A call that repeats the previous one's sizes -- particle and sample
        counts, snapshot form, time step, component, precision, particle size
        and device grid -- reuses its device batch (refilled in place by one
        copy) and, from then on, replays the launch sequence captured as a HIP
        graph on the first repeat: the library's planning, workspace and
        launches are then one graph launch (engine.SpectrumGraph)."""
        shape = (batch.n_particles, batch.n_points, batch.n_snaps, batch.snaps_stride, int(batch.it_snaps.size),
                 batch.dt)
        key = (id(self._grid), comp, self.precision, sigma if comp == "cartesian_complex" else 0.0) + shape
        prev = getattr(self, "_replay", None)
        if prev is not None and prev["key"] == key:
            prev["batch"].refill(batch)
            if prev["graph"] is None:
                # the first call loaded the kernels: capture without a warm-up spectrum
                prev["graph"] = engine.SpectrumGraph(prev["batch"], self._grid, comp=comp, precision=self.precision,
                                                     sigma_particle=sigma, warm=False)
            for k, out in zip(self._rad_keys, prev["graph"].replay()):
                self._rad_device[k] = out
            return
        self._replay = None          # a new shape releases the previous batch and graph
        dev_batch = engine.DeviceBatch(batch, self.device)
        engine.spectrum(dev_batch, self._grid, comp=comp, precision=self.precision,
                        outs=[self._rad_device[k] for k in self._rad_keys], sigma_particle=sigma)
        self._replay = {"key": key, "batch": dev_batch, "graph": None}

    def _init_args(self, Args):
        """Axes and integration elements, synchrad/calc.py:355-451."""
        self.Args = Args
        a = self.Args
        a.setdefault("mode", "far")
        a.setdefault("dtype", "double")
        a.setdefault("ctx", None)
        a.setdefault("Features", [])
        if a["mode"] not in ("far", "near"):
            raise ValueError(f"mode must be 'far' or 'near', got {a['mode']!r}")
        precisions = {"double": "double", "float": "float", "single": "float"}
        if a["dtype"] not in precisions:
            raise ValueError(f"dtype must be 'double' or 'float', got {a['dtype']!r}")
        self.precision = precisions[a["dtype"]]
        self.dtype = np.double          # host arrays stay float64 in both precisions

        (w_lo, w_hi), (lo1, hi1), (p_lo, p_hi), counts = a["grid"]
        n_w, n1, n_p = counts
        a["gridNodeNums"] = counts
        a["numGridNodes"] = int(np.prod(counts))

        omega, lam = frequency_nodes(w_lo, w_hi, n_w, a["Features"])
        if lam is not None:
            a["wavelengths"] = lam
        a["omega"] = omega.astype(self.dtype)
        a["dw"] = np.abs(np.diff(omega)) if n_w > 1 else np.array([1.0], dtype=self.dtype)

        ax1 = inclusive_nodes(lo1, hi1, n1)
        phi = periodic_nodes(p_lo, p_hi, n_p)
        d1 = first_step(ax1)
        name1, step1 = ("theta", "dth") if a["mode"] == "far" else ("radius", "dr")
        a[step1] = d1
        a["dph"] = first_step(phi)
        a[name1] = ax1.astype(self.dtype)
        a["phi"] = phi.astype(self.dtype)
        a["dV"] = a["dw"] * d1 * a["dph"]

    def _init_comm(self):
        """Device selection (synchrad/calc.py:513-558): the current HIP device,
        or none when ``ctx`` is False."""
        if self.Args.get("ctx") is False:
            self.device = None
            self.dev_type, self.dev_name, self.plat_name = "Starting without", "", "None"
            return
        import torch
        self.device = engine._dev()
        self.dev_type, self.plat_name = "GPU", "HIP"
        self.dev_name = torch.cuda.get_device_name(self.device)

    def _init_data(self):
        """Device axes, synchrad/calc.py:486-512 (a near-field screen waits
        for its L_screen)."""
        self.Data = {}
        self._grid = None
        if getattr(self, "device", None) is None or self.Args["mode"] == "near" and \
                self.Args.get("L_screen") is None:
            return
        second = self.Args["theta"] if self.Args["mode"] == "far" else self.Args["radius"]
        self._grid = engine.DeviceGrid(self.Args["mode"], self.Args["omega"], second, self.Args["phi"],
                                       L_screen=self.Args.get("L_screen"), device=self.device)

    def _compile_kernels(self):
        """The kernels are compiled ahead of time; this checks the library loads
        (the reference JIT-builds its OpenCL here, synchrad/calc.py:605-624)."""
        if getattr(self, "device", None) is not None:
            from . import _abi
            _abi.load()

    def _init_raditaion(self, comp, nSnaps):
        """Zeroed device spectra per output key, synchrad/calc.py:453-484 (the
        particle-size form factor is made with the device grid)."""
        import torch
        if comp not in engine.COMP_KEYS:
            raise KeyError(comp)
        mode = self.Args["mode"]
        if mode == "near" and comp not in ("total", "cartesian", "cartesian_complex"):
            raise KeyError(f"{comp} is not available in the near field")
        self.Args["comp"] = comp
        if mode == "near":
            self.Args["theta"] = np.arctan2(self.Args["radius"], self.Args["L_screen"])
            if self._grid is None or self._grid.l_screen != float(self.Args["L_screen"]):
                self._grid = engine.DeviceGrid("near", self.Args["omega"], self.Args["radius"], self.Args["phi"],
                                               L_screen=self.Args["L_screen"], device=self.device)
        n_w, n1, n_p = self.Args["gridNodeNums"]
        shape = (int(nSnaps), int(n_p), int(n1), int(n_w))
        self._rad_keys = list(engine.COMP_KEYS[comp])
        self._rad_device = {k: torch.zeros(shape, dtype=torch.float64, device=self.device)
                            for k in self._rad_keys}
        self._rad_device_mirrors = {}
        self.Data["radiation"] = {}

    def _spectr_from_device(self, nSnaps):
        """Device (nSnaps, nPhi, n1, nOmega) -> host (nSnaps, nOmega, n1, nPhi)
        float64, synchrad/calc.py:573-577.  The device copies stay for the
        analysis helpers while the host arrays are these very objects."""
        for key in self._rad_keys:
            # the axis swap runs on the device; one contiguous copy comes back
            host = self._rad_device[key].transpose(-1, -3).contiguous().cpu().numpy()
            self.Data["radiation"][key] = host
            self._rad_device_mirrors[key] = host

    def _set_snap_iterations(self, it_range, nSnaps):
        """synchrad/calc.py:626-630."""
        self.snap_iterations = engine.snap_iterations(it_range, int(nSnaps))

    def _read_args(self, file_spectrum):
        """Reload a spectrum file, synchrad/calc.py:648-666 (no device)."""
        self.device = None
        radiation, self.Args, self.snap_iterations, self.total_weight = read_spectrum_file(file_spectrum)
        self.Data = {"radiation": radiation}
