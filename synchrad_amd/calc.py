"""``SynchRad``: the reference's calculator API on the MI355X engine.

Drop-in for ``synchrad.calc.SynchRad`` (synchrad/calc.py:21-666): same
constructor, same ``calculate_spectrum`` signature, same ``Args`` keys, same
``Data['radiation']`` layout (float64 ``(nSnaps, nOmega, nTheta|nR, nPhi)``)
and the same private re-initialisation methods that the reference's test
scripts call (``_init_args``, ``_init_data``, ``_compile_kernels``).

Differences, all deliberate:

* ``dtype: 'double'`` runs the FP64 kernels.  ``dtype: 'float'`` runs the
  split-precision kernels: the phase and its reduction in FP64, amplitudes,
  sin/cos and compensated sums in FP32.  Host axes, tracks and ``timeStep``
  stay float64 in both modes (the reference casts them to float32 in its float
  mode, which costs several radians of phase on the betatron beam's 8000-step
  tracks).
* All particles go to the device as one CSR batch and are integrated by one
  stream-ordered call, instead of one launch per particle.
* String options are compared with ``==`` (the reference uses ``is``).
* ``ctx`` and ``native`` are accepted and ignored (one HIP device per
  process); ``ctx=False`` still starts without a device for reloading files.
* Zero-length tracks contribute nothing (the reference reads out of bounds).
* MPI is not used: ``rank`` is 0 and ``size`` is 1.
"""
from __future__ import annotations

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


def _h5_file(path: str, mode: str):
    """An h5py.File, or the ctypes binding of libhdf5 with the same use."""
    if h5py_installed:
        return h5py.File(path, mode)
    from . import h5io
    return h5io.File(path, mode)


class SynchRad(Utilities):
    """Spectral-angular radiation of point charges from their tracks."""

    def __init__(self, Args={}, file_spectrum=None):
        """See synchrad/calc.py:29-99 for the ``Args`` keys: ``grid``, ``mode``
        ('far'/'near'), ``dtype`` ('double'/'float'), ``ctx``, ``Features``
        ('wavelengthGrid'/'logGrid')."""
        self.rank = 0
        self.size = 1
        if file_spectrum is None:
            self._init_args(Args)
            self._init_comm()
            self._init_data()
            self._compile_kernels()
        else:
            self._read_args(file_spectrum)

    # ------------------------------------------------------------------ API
    def calculate_spectrum(self, particleTracks=[], file_tracks=None, timeStep=None, comp="total",
                           L_screen=None, Np_max=None, it_range=None, nSnaps=1, sigma_particle=0,
                           weights_normalize=None, file_spectrum=None, verbose=True):
        """Compute the spectrum of ``particleTracks`` (or of ``file_tracks``),
        synchrad/calc.py:101-290.  Results land in ``self.Data['radiation']``."""
        if self.device is None:
            raise RuntimeError("this SynchRad was started without a device (ctx=False)")
        self.Args["sigma_particle"] = self.dtype(sigma_particle)
        if self.Args["mode"] == "near":
            if L_screen is not None:
                self.Args["L_screen"] = L_screen
            elif "L_screen" not in self.Args:
                raise ValueError("Define L_screen argument for near-field calculation")

        nSnaps = np.uint32(nSnaps)
        self._init_raditaion(comp, nSnaps)

        if timeStep is not None:
            self.Args["timeStep"] = self.dtype(timeStep)
        if it_range is not None:
            it_range = tuple(it_range)

        if file_tracks is not None:
            # synchrad/calc.py:186-219
            with _h5_file(file_tracks, "r") as f_tracks:
                self.Args["timeStep"] = self.dtype(f_tracks["misc/cdt"][()])
                if it_range is None:
                    if "it_range" in f_tracks["misc"].keys():
                        it_range = tuple(f_tracks["misc/it_range"][()])
                        if verbose:
                            print("it_range from the input file will be used")
                    elif verbose:
                        print("Separate it_range for each track will be used")
                Np = int(f_tracks["misc/N_particles"][()])
                if Np_max is not None:
                    Np = min(Np_max, Np)
                cmps = ("x", "y", "z", "ux", "uy", "uz", "w", "it_start")
                particleTracks = [[f_tracks[f"tracks/{ip:d}/{cmp}"][()] for cmp in cmps]
                                  for ip in range(Np)]
                if verbose:
                    print("Tracks are loaded")
        else:
            if it_range is None and verbose:
                print("Separate it_range for each track will be used")
            Np = len(particleTracks)
            if Np_max is not None:
                Np = min(Np_max, Np)
            particleTracks = particleTracks[:Np]

        if "timeStep" not in self.Args:
            raise ValueError("timeStep is required")

        # weights, synchrad/calc.py:239-265 (normalisation mutates the caller's lists)
        self.total_weight = 0.0
        if weights_normalize in ("mean", "max"):
            weights = [track[6] for track in particleTracks]
            weight_norm = np.mean(weights) if weights_normalize == "mean" else np.max(weights)
        for track in particleTracks:
            if weights_normalize in ("mean", "max"):
                track[6] /= weight_norm
            elif weights_normalize == "ones":
                track[6] = 1.0
            self.total_weight += track[6]

        if it_range is not None:
            self._set_snap_iterations(it_range, nSnaps)
        hb = engine.pack_tracks(particleTracks, dt=float(self.Args["timeStep"]), it_range=it_range,
                                nSnaps=int(nSnaps))
        if it_range is None and hb.n_particles:
            # per-track mode leaves the last particle's snapshot list behind
            # (synchrad/calc.py:301)
            n_last = int(hb.offsets[-1] - hb.offsets[-2])
            self._set_snap_iterations((0, n_last), nSnaps)

        if hb.n_particles:
            batch = engine.DeviceBatch(hb, self.device)
            outs = [self._rad_device[k] for k in self._rad_keys]
            engine.spectrum(batch, self._grid, comp=comp, precision=self.precision, outs=outs,
                            sigma_particle=float(self.Args["sigma_particle"]))
        self._spectr_from_device(nSnaps)

        if file_spectrum is not None:
            # synchrad/calc.py:274-290
            with _h5_file(file_spectrum, "w") as f_out:
                for key in self.Data["radiation"].keys():
                    f_out["radiation/" + key] = self.Data["radiation"][key]
                for key in self.Args.keys():
                    if key in ("grid", "ctx"):
                        continue
                    f_out["Args/" + key] = self.Args[key]
                f_out["snap_iterations"] = self.snap_iterations
                f_out["total_weight"] = self.total_weight
            print(f"Spectrum is saved to {file_spectrum}")

    # ------------------------------------------------------------ internals
    def _init_args(self, Args):
        """Axes and integration elements, synchrad/calc.py:355-451."""
        self.Args = Args
        if "mode" not in self.Args.keys():
            self.Args["mode"] = "far"
        if self.Args["mode"] not in ("far", "near"):
            raise ValueError(f"mode must be 'far' or 'near', got {self.Args['mode']!r}")
        if "dtype" not in self.Args:
            self.Args["dtype"] = "double"
        if self.Args["dtype"] == "double":
            self.precision = "double"
        elif self.Args["dtype"] in ("float", "single"):
            self.precision = "float"
        else:
            raise ValueError(f"dtype must be 'double' or 'float', got {self.Args['dtype']!r}")
        # host arrays stay float64 in both precisions
        self.dtype = np.double
        if "ctx" not in self.Args.keys():
            self.Args["ctx"] = None

        self.Args["gridNodeNums"] = self.Args["grid"][-1]
        self.Args["numGridNodes"] = int(np.prod(self.Args["gridNodeNums"]))
        No = self.Args["gridNodeNums"][0]
        omega_min, omega_max = self.Args["grid"][0]
        if "Features" not in self.Args.keys():
            self.Args["Features"] = []

        omega = np.r_[omega_min:omega_max:No * 1j]
        for feature in self.Args["Features"]:
            if feature == "wavelengthGrid":
                self.Args["wavelengths"] = np.r_[1.0 / omega_max:1.0 / omega_min:No * 1j]
                omega = 1.0 / self.Args["wavelengths"]
                break
            elif feature == "logGrid":
                d_log_w = np.log(omega_max / omega_min) / (No - 1.0)
                omega = omega_min * np.exp(d_log_w * np.arange(No))
                break
        self.Args["omega"] = omega.astype(self.dtype)
        if No > 1:
            self.Args["dw"] = np.abs(omega[1:] - omega[:-1])
        else:
            self.Args["dw"] = np.array([1.0], dtype=self.dtype)

        if self.Args["mode"] == "far":
            Nt, Np = self.Args["gridNodeNums"][1:]
            theta_min, theta_max = self.Args["grid"][1]
            phi_min, phi_max = self.Args["grid"][2]
            theta = np.r_[theta_min:theta_max:Nt * 1j]
            phi = phi_min + (phi_max - phi_min) / Np * np.arange(Np)
            self.Args["dth"] = theta[1] - theta[0] if Nt > 1 else self.dtype(1.0)
            self.Args["dph"] = phi[1] - phi[0] if Np > 1 else self.dtype(1.0)
            self.Args["theta"] = theta.astype(self.dtype)
            self.Args["phi"] = phi.astype(self.dtype)
            self.Args["dV"] = self.Args["dw"] * self.Args["dth"] * self.Args["dph"]
        else:
            Nr, Np = self.Args["gridNodeNums"][1:]
            r_min, r_max = self.Args["grid"][1]
            phi_min, phi_max = self.Args["grid"][2]
            radius = np.r_[r_min:r_max:Nr * 1j]
            phi = phi_min + (phi_max - phi_min) / Np * np.arange(Np)
            self.Args["dr"] = radius[1] - radius[0] if Nr > 1 else 1.0
            self.Args["dph"] = phi[1] - phi[0] if Np > 1 else 1.0
            self.Args["phi"] = phi.astype(self.dtype)
            self.Args["radius"] = radius.astype(self.dtype)
            self.Args["dV"] = self.Args["dw"] * self.Args["dr"] * self.Args["dph"]

    def _init_comm(self):
        """Device selection (synchrad/calc.py:513-558): the current HIP device,
        or none when ``ctx`` is False."""
        if self.Args.get("ctx") is False:
            self.device = None
            self.dev_type, self.dev_name, self.plat_name = "Starting without", "", "None"
            return
        self.device = engine._dev()
        import torch
        self.dev_type = "GPU"
        self.dev_name = torch.cuda.get_device_name(self.device)
        self.plat_name = "HIP"

    def _init_data(self):
        """Device axes, synchrad/calc.py:486-512."""
        self.Data = {}
        if getattr(self, "device", None) is None:
            self._grid = None
            return
        if self.Args["mode"] == "near" and self.Args.get("L_screen") is None:
            self._grid = None  # made once calculate_spectrum gives L_screen
            return
        ax1 = self.Args["theta"] if self.Args["mode"] == "far" else self.Args["radius"]
        self._grid = engine.DeviceGrid(self.Args["mode"], self.Args["omega"], ax1, self.Args["phi"],
                                       L_screen=self.Args.get("L_screen"), device=self.device)

    def _compile_kernels(self):
        """The kernels are compiled ahead of time; this checks the library loads
        (the reference JIT-builds its OpenCL here, synchrad/calc.py:605-624)."""
        if getattr(self, "device", None) is None:
            return
        from . import _abi
        _abi.load()

    def _init_raditaion(self, comp, nSnaps):
        """Zeroed device spectra per component, synchrad/calc.py:453-484."""
        import torch
        if comp not in engine.COMP_KEYS:
            raise KeyError(comp)
        if self.Args["mode"] == "near" and comp not in ("total", "cartesian", "cartesian_complex"):
            raise KeyError(f"{comp} is not available in the near field")
        self.Args["comp"] = comp
        if self.Args["mode"] == "near":
            self.Args["theta"] = np.arctan2(self.Args["radius"], self.Args["L_screen"])
            if self._grid is None or self._grid.l_screen != float(self.Args["L_screen"]):
                self._grid = engine.DeviceGrid("near", self.Args["omega"], self.Args["radius"],
                                               self.Args["phi"], L_screen=self.Args["L_screen"],
                                               device=self.device)
        self.Data["radiation"] = {}
        self._rad_keys = list(engine.COMP_KEYS[comp])
        g = self.Args["gridNodeNums"]
        shape = (int(nSnaps), int(g[2]), int(g[1]), int(g[0]))
        self._rad_device = {k: torch.zeros(shape, dtype=torch.float64, device=self.device)
                            for k in self._rad_keys}

    def _spectr_from_device(self, nSnaps):
        """D2H and (nSnaps, nPhi, n1, nOmega) -> (nSnaps, nOmega, n1, nPhi),
        synchrad/calc.py:573-577."""
        for key in self._rad_keys:
            buff = self._rad_device[key].cpu().numpy().swapaxes(-1, -3)
            self.Data["radiation"][key] = np.ascontiguousarray(buff, dtype=np.double)

    def _set_snap_iterations(self, it_range, nSnaps):
        """synchrad/calc.py:626-630."""
        self.snap_iterations = engine.snap_iterations(it_range, int(nSnaps))

    def _read_args(self, file_spectrum):
        """Reload a spectrum file, synchrad/calc.py:648-666."""
        self.device = None
        self.Data = {"radiation": {}}
        self.Args = {}
        with _h5_file(file_spectrum, "r") as f:
            for key in f["radiation"].keys():
                self.Data["radiation"][key] = f["radiation/" + key][()]
            for key in f["Args"].keys():
                self.Args[key] = f["Args/" + key][()]
                if isinstance(self.Args[key], bytes):
                    self.Args[key] = self.Args[key].decode()
            self.snap_iterations = f["snap_iterations"][()]
            self.total_weight = f["total_weight"][()]
