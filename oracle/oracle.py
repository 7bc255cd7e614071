"""CPU FP64 oracle for synchrad's spectral radiation integral.

TEST INFRASTRUCTURE ONLY.  The tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module as the checker; the
product package ``synchrad_amd`` never imports it.

The arithmetic lives in ``oracle.c`` (built into ``liboracle.so`` by the
Makefile next to it).  This module restates the reference host logic that
surrounds the kernels, one kernel call per particle as the reference does:

* axes                 -- ``SynchRad._init_args``   synchrad/calc.py:380-451
* device axis arrays   -- ``SynchRad._init_data``   synchrad/calc.py:494-512
* snapshot iterations  -- ``_set_snap_iterations``  synchrad/calc.py:626-630
* per-track defaults   -- ``_process_track``        synchrad/calc.py:296-301
* output keys, FormFactor -- ``_init_raditaion``    synchrad/calc.py:455-484
* result layout        -- ``_spectr_from_device``   synchrad/calc.py:573-577

Parity status of each function is recorded in the tests that pin it
(``tests/test_oracle_pins.py``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

COMP_CODES = {
    "total": 0,
    "cartesian": 1,
    "cartesian_complex": 2,
    "spheric": 3,
    "spheric_complex": 4,
}

# synchrad/calc.py:458-465
COMP_KEYS = {
    "cartesian": ["x", "y", "z"],
    "cartesian_complex": ["xre", "xim", "yre", "yim", "zre", "zim"],
    "spheric": ["r", "theta", "phi"],
    "spheric_complex": ["rre", "rim", "thetare", "thetaim", "phire", "phiim"],
    "total": ["total"],
}

_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint32)


def lib(path: Optional[str] = None):
    """Load liboracle.so (built by ``make -C oracle``), or another build of
    oracle.c at ``path`` (the tests load deliberately broken builds to show the
    pins reject them)."""
    global _lib
    if path is not None or _lib is None:
        p = path or _LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(p)
        L.oracle_far.restype = ctypes.c_int
        L.oracle_far.argtypes = (
            [ctypes.c_int, ctypes.POINTER(_dp)] + [_dp] * 6
            + [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
            + [_dp] * 5 + [ctypes.c_uint32] * 3
            + [ctypes.c_double, ctypes.c_uint32, _up, _dp, ctypes.c_int])
        L.oracle_near_xy.restype = ctypes.c_int
        L.oracle_near_xy.argtypes = (
            [ctypes.c_int, ctypes.POINTER(_dp)] + [_dp] * 6
            + [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
            + [_dp] * 3 + [ctypes.c_double] + [ctypes.c_uint32] * 3
            + [ctypes.c_double, ctypes.c_uint32, _up, ctypes.c_int])
        L.oracle_near.restype = ctypes.c_int
        L.oracle_near.argtypes = (
            [ctypes.c_int, ctypes.POINTER(_dp)] + [_dp] * 6
            + [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
            + [_dp] * 4 + [ctypes.c_double] + [ctypes.c_uint32] * 3
            + [ctypes.c_double, ctypes.c_uint32, _up, ctypes.c_int])
        if path is not None:
            return L
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


# --------------------------------------------------------------------------
# host restatements
# --------------------------------------------------------------------------
def make_axes(grid, mode: str = "far", features: Sequence[str] = ()) -> Dict[str, np.ndarray]:
    """Spectral axes exactly as ``SynchRad._init_args`` builds them in double
    precision (synchrad/calc.py:380-451): inclusive omega (or the
    'wavelengthGrid'/'logGrid' variants), inclusive theta or radius, and an
    endpoint-exclusive phi."""
    No, N1, Nph = grid[-1]
    omega_min, omega_max = grid[0]
    omega = np.r_[omega_min:omega_max:No * 1j]
    for feature in features:
        if feature == "wavelengthGrid":
            omega = 1.0 / np.r_[1.0 / omega_max:1.0 / omega_min:No * 1j]
            break
        elif feature == "logGrid":
            d_log_w = np.log(omega_max / omega_min) / (No - 1.0)
            omega = omega_min * np.exp(d_log_w * np.arange(No))
            break
    a1_min, a1_max = grid[1]
    phi_min, phi_max = grid[2]
    ax1 = np.r_[a1_min:a1_max:N1 * 1j]
    phi = phi_min + (phi_max - phi_min) / Nph * np.arange(Nph)
    out = {"omega": omega.astype(np.double), "phi": phi.astype(np.double)}
    out["theta" if mode == "far" else "radius"] = ax1.astype(np.double)
    return out


def snap_iterations(it_range, nSnaps: int) -> np.ndarray:
    """synchrad/calc.py:626-630."""
    return np.ascontiguousarray(
        np.linspace(*(tuple(it_range) + (int(nSnaps) + 1,)), dtype=np.uint32)[1:])


def spectrum(mode: str, comp: str, tracks: List[list], axes: Dict[str, np.ndarray],
             timeStep: float, it_range=None, nSnaps: int = 1, L_screen: Optional[float] = None,
             sigma_particle: float = 0.0, threads: int = 0, library=None) -> Dict[str, np.ndarray]:
    """Incoherent/coherent spectra of ``tracks`` on ``axes`` (FP64).

    ``tracks`` items are ``[x, y, z, ux, uy, uz, w, it_start]`` or the 7-item
    form without ``it_start`` (synchrad/calc.py:580-584).  Near-field
    ``axes`` with "x" and "y" (and no "radius") describe a Cartesian screen,
    an extension the reference does not have (``oracle_near_xy``).  Returns
    ``{key: array(nSnaps, nOmega, nTheta|nR, nPhi)}`` as
    ``SynchRad.Data['radiation']`` holds it (synchrad/calc.py:573-577).
    """
    L = library if library is not None else lib()
    omega = axes["omega"].astype(np.double)
    omega_dev = np.ascontiguousarray(np.double(2 * np.pi) * omega)
    cartesian = mode == "near" and "x" in axes and "y" in axes
    if cartesian:
        # Cartesian near-field screen (no counterpart in the reference): x in
        # place of the radius, y in place of phi
        scr_x = np.ascontiguousarray(axes["x"], dtype=np.double)
        scr_y = np.ascontiguousarray(axes["y"], dtype=np.double)
        phi = scr_y
    else:
        phi = axes["phi"].astype(np.double)
        sinP = np.ascontiguousarray(np.sin(phi))
        cosP = np.ascontiguousarray(np.cos(phi))
    nO, nP = omega.size, phi.size
    if cartesian:
        if L_screen is None:
            raise ValueError("near field needs L_screen")
        n1 = scr_x.size
    elif mode == "far":
        theta = axes["theta"].astype(np.double)
        sinT = np.ascontiguousarray(np.sin(theta))
        cosT = np.ascontiguousarray(np.cos(theta))
        n1 = theta.size
    elif mode == "near":
        if L_screen is None:
            raise ValueError("near field needs L_screen")
        radius = np.ascontiguousarray(axes["radius"].astype(np.double))
        n1 = radius.size
    else:
        raise ValueError(f"unknown mode {mode!r}")
    keys = COMP_KEYS[comp]
    code = COMP_CODES[comp]
    nSnaps = int(nSnaps)
    shape = (nSnaps, nP, n1, nO)
    outs = [np.zeros(shape, dtype=np.double) for _ in keys]
    out_ptrs = (_dp * len(outs))(*[_p(o) for o in outs])
    ff = np.ascontiguousarray(
        np.exp(np.double(-0.5) * (np.double(2 * np.pi) * omega * np.double(sigma_particle)) ** 2))

    global_snaps = None
    if it_range is not None:
        global_snaps = snap_iterations(it_range, nSnaps)

    for track in tracks:
        if len(track) == 8:
            x, y, z, ux, uy, uz, wp, it_start = track
        else:
            x, y, z, ux, uy, uz, wp = track
            it_start = 0
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.double)) for a in (x, y, z, ux, uy, uz)]
        n = arrs[0].size
        if it_range is None:
            it_start = 0
            itEnd = n
            snaps = snap_iterations((0, n), nSnaps)
        else:
            itEnd = int(it_range[-1])
            snaps = global_snaps
        if n == 0 or itEnd == 0:
            continue  # the reference reads out of bounds / loops 2^32 times here
        snaps = np.ascontiguousarray(snaps, dtype=np.uint32)
        if mode == "far":
            rc = L.oracle_far(code, out_ptrs, *[_p(a) for a in arrs],
                              float(wp), int(np.uint32(it_start)), int(itEnd), int(n),
                              _p(omega_dev), _p(sinT), _p(cosT), _p(sinP), _p(cosP),
                              nO, n1, nP, float(timeStep), nSnaps,
                              snaps.ctypes.data_as(_up), _p(ff), int(threads))
        elif cartesian:
            rc = L.oracle_near_xy(code, out_ptrs, *[_p(a) for a in arrs],
                                  float(wp), int(np.uint32(it_start)), int(itEnd), int(n),
                                  _p(omega_dev), _p(scr_x), _p(scr_y), float(L_screen),
                                  nO, n1, nP, float(timeStep), nSnaps,
                                  snaps.ctypes.data_as(_up), int(threads))
        else:
            rc = L.oracle_near(code, out_ptrs, *[_p(a) for a in arrs],
                               float(wp), int(np.uint32(it_start)), int(itEnd), int(n),
                               _p(omega_dev), _p(radius), _p(sinP), _p(cosP), float(L_screen),
                               nO, n1, nP, float(timeStep), nSnaps,
                               snaps.ctypes.data_as(_up), int(threads))
        if rc != 0:
            raise RuntimeError(f"oracle returned {rc}")
    return {k: np.ascontiguousarray(o.swapaxes(-1, -3), dtype=np.double)
            for k, o in zip(keys, outs)}


def far_adjoint(comp: str, tracks: List[list], axes: Dict[str, np.ndarray], timeStep: float,
                G, it_range=None, threads: int = 0, library=None):
    """dL/d(track) for L with dL/dS = G, far field, one snapshot (FP64).

    ``G`` holds one array per output key of ``comp`` ('total' or 'cartesian') in
    the host layout ``(1, nOmega, nTheta, nPhi)``.  Returns ``(grads, gw)``:
    ``grads[p]`` is ``[gx, gy, gz, gux, guy, guz]`` for particle p and ``gw[p]``
    its weight derivative.
    """
    L = library if library is not None else lib()
    if not hasattr(L, "_adj_declared"):
        L.oracle_far_adjoint.restype = ctypes.c_int
        L.oracle_far_adjoint.argtypes = (
            [ctypes.c_int, ctypes.POINTER(_dp)] + [_dp] * 6
            + [ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
            + [_dp] * 5 + [ctypes.c_uint32] * 3
            + [ctypes.c_double, ctypes.c_uint32, _up] + [_dp] * 7 + [ctypes.c_int])
        L._adj_declared = True
    omega = axes["omega"].astype(np.double)
    theta = axes["theta"].astype(np.double)
    phi = axes["phi"].astype(np.double)
    omega_dev = np.ascontiguousarray(np.double(2 * np.pi) * omega)
    sinT, cosT = np.ascontiguousarray(np.sin(theta)), np.ascontiguousarray(np.cos(theta))
    sinP, cosP = np.ascontiguousarray(np.sin(phi)), np.ascontiguousarray(np.cos(phi))
    keys = COMP_KEYS[comp]
    Gdev = [np.ascontiguousarray(np.asarray(G[k], dtype=np.double).swapaxes(-1, -3)) for k in keys]
    g_ptrs = (_dp * len(Gdev))(*[_p(g) for g in Gdev])
    global_snaps = snap_iterations(it_range, 1) if it_range is not None else None
    grads, gws = [], []
    for track in tracks:
        if len(track) == 8:
            x, y, z, ux, uy, uz, wp, it_start = track
        else:
            x, y, z, ux, uy, uz, wp = track
            it_start = 0
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.double)) for a in (x, y, z, ux, uy, uz)]
        n = arrs[0].size
        out = [np.zeros(n) for _ in range(6)]
        gw = np.zeros(1)
        if it_range is None:
            it_start, itEnd, snaps = 0, n, snap_iterations((0, n), 1)
        else:
            itEnd, snaps = int(it_range[-1]), global_snaps
        if n > 0 and itEnd > 0:
            rc = L.oracle_far_adjoint(COMP_CODES[comp], g_ptrs, *[_p(a) for a in arrs], float(wp),
                                      int(np.uint32(it_start)), int(itEnd), int(n), _p(omega_dev),
                                      _p(sinT), _p(cosT), _p(sinP), _p(cosP), omega.size, theta.size,
                                      phi.size, float(timeStep), 1,
                                      np.ascontiguousarray(snaps, dtype=np.uint32).ctypes.data_as(_up),
                                      *[_p(o) for o in out], _p(gw), int(threads))
            if rc != 0:
                raise RuntimeError(f"oracle adjoint returned {rc}")
        grads.append(out)
        gws.append(gw[0])
    return grads, np.array(gws)
