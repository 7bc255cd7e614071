"""ctypes binding of ``include/synchrad_amd.h``.

The structs below mirror the header field for field.  Loading fails loudly when
``libsynchrad_amd.so`` is missing: there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes
import os

from ._build import LIB

SR_ABI_VERSION = 3

SR_OK = 0
SR_ERR_INVALID = -1
SR_ERR_UNSUPPORTED = -2
SR_ERR_HIP = -3
SR_ERR_WORKSPACE = -4

SR_MODE_FAR = 0
SR_MODE_NEAR = 1

SR_COMP_TOTAL = 0
SR_COMP_CARTESIAN = 1
SR_COMP_CARTESIAN_COMPLEX = 2
SR_COMP_SPHERIC = 3
SR_COMP_SPHERIC_COMPLEX = 4

SR_PREC_F64 = 0
SR_PREC_SPLIT = 1

SR_SWEEP_AUTO = 0
SR_SWEEP_DIRECT = 1
SR_SWEEP_RECURRENCE = 2

MODES = {"far": SR_MODE_FAR, "near": SR_MODE_NEAR}
COMPS = {
    "total": SR_COMP_TOTAL,
    "cartesian": SR_COMP_CARTESIAN,
    "cartesian_complex": SR_COMP_CARTESIAN_COMPLEX,
    "spheric": SR_COMP_SPHERIC,
    "spheric_complex": SR_COMP_SPHERIC_COMPLEX,
}

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32
_f64 = ctypes.c_double


class SrTracks(ctypes.Structure):
    _fields_ = [
        ("n_particles", _i64), ("n_points", _i64),
        ("offsets", _vp), ("offsets_host", _vp),
        ("x", _vp), ("y", _vp), ("z", _vp),
        ("ux", _vp), ("uy", _vp), ("uz", _vp),
        ("weight", _vp), ("it_start", _vp), ("it_end", _vp), ("it_snaps", _vp),
        ("n_snaps", _u32), ("snaps_stride", _u32),
        ("dt", _f64),
    ]


class SrGrid(ctypes.Structure):
    _fields_ = [
        ("mode", _i32), ("n_omega", _u32), ("n_theta", _u32), ("n_phi", _u32),
        ("omega", _vp), ("sin_theta", _vp), ("cos_theta", _vp), ("radius", _vp),
        ("sin_phi", _vp), ("cos_phi", _vp), ("l_screen", _f64), ("omega_step", _f64),
        ("screen_y", _vp),
    ]


class SrPlanInfo(ctypes.Structure):
    _fields_ = [("sweep", _i32), ("omega_per_lane", _i32), ("waves_per_block", _i32), ("slices", _i32),
                ("tiles", _i64), ("ordered", _i32)]


class SrOpts(ctypes.Structure):
    _fields_ = [
        ("comp", _i32), ("precision", _i32), ("omega_per_lane", _i32),
        ("waves_per_block", _i32), ("slices", _i32), ("persistent", _i32),
        ("sweep", _i32), ("form_factor", _vp), ("tile_order", _i32),
    ]


EXPORTS = (
    "sr_abi_version", "sr_last_error", "sr_num_outputs", "sr_workspace_bytes",
    "sr_spectrum_batch", "sr_far_total", "sr_far_cartesian", "sr_near_total",
    "sr_near_cartesian", "sr_far_adjoint_workspace_bytes", "sr_far_adjoint_batch",
    "sr_far_fields_workspace_bytes", "sr_far_fields_batch", "sr_far_adjoint_fields_workspace_bytes",
    "sr_far_adjoint_from_fields", "sr_spectrum_batch_noalloc", "sr_plan", "sr_autotune_workspace_bytes",
    "sr_autotune",
)

SR_NOALLOC_SCRATCH_BYTES = 64 << 20

_lib = None


class SynchradError(RuntimeError):
    pass


def load(path: str = None):
    """Load the library (once) and declare every exported signature.
    ``SYNCHRAD_AMD_LIB`` names another build of the library (for comparing
    compile-time variants); the default is the in-tree build."""
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        path = os.environ.get("SYNCHRAD_AMD_LIB") or LIB
    if not os.path.exists(path):
        raise SynchradError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(path)
    L.sr_abi_version.restype = ctypes.c_int
    L.sr_abi_version.argtypes = []
    L.sr_last_error.restype = ctypes.c_char_p
    L.sr_last_error.argtypes = []
    L.sr_num_outputs.restype = ctypes.c_int
    L.sr_num_outputs.argtypes = [_i32, _i32]
    L.sr_workspace_bytes.restype = ctypes.c_int
    L.sr_workspace_bytes.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                     ctypes.POINTER(SrOpts), ctypes.POINTER(ctypes.c_size_t)]
    L.sr_spectrum_batch.restype = ctypes.c_int
    L.sr_spectrum_batch.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                    ctypes.POINTER(SrOpts), ctypes.POINTER(_vp), _vp,
                                    ctypes.c_size_t, _vp]
    L.sr_plan.restype = ctypes.c_int
    L.sr_plan.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid), ctypes.POINTER(SrOpts),
                          ctypes.POINTER(SrPlanInfo)]
    L.sr_autotune_workspace_bytes.restype = ctypes.c_int
    L.sr_autotune_workspace_bytes.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                              ctypes.POINTER(SrOpts), ctypes.POINTER(SrOpts), ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_size_t)]
    L.sr_autotune.restype = ctypes.c_int
    L.sr_autotune.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid), ctypes.POINTER(SrOpts),
                              ctypes.POINTER(SrOpts), ctypes.c_int, ctypes.POINTER(_vp), _vp, ctypes.c_size_t,
                              ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_float),
                              ctypes.POINTER(ctypes.c_int)]
    L.sr_spectrum_batch_noalloc.restype = ctypes.c_int
    L.sr_spectrum_batch_noalloc.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                            ctypes.POINTER(SrOpts), ctypes.POINTER(_vp), _vp]
    track_args = [_vp] * 6 + [_f64, _u32, _u32, _u32]
    tail = [_u32] * 3 + [_f64, _u32, _vp, _i32, _vp]
    L.sr_far_total.restype = ctypes.c_int
    L.sr_far_total.argtypes = [_vp] + track_args + [_vp] * 5 + tail
    L.sr_far_cartesian.restype = ctypes.c_int
    L.sr_far_cartesian.argtypes = [_vp] * 3 + track_args + [_vp] * 5 + tail
    L.sr_near_total.restype = ctypes.c_int
    L.sr_near_total.argtypes = [_vp] + track_args + [_vp] * 4 + [_f64] + tail
    L.sr_near_cartesian.restype = ctypes.c_int
    L.sr_near_cartesian.argtypes = [_vp] * 3 + track_args + [_vp] * 4 + [_f64] + tail
    L.sr_far_adjoint_workspace_bytes.restype = ctypes.c_int
    L.sr_far_adjoint_workspace_bytes.argtypes = L.sr_workspace_bytes.argtypes
    L.sr_far_adjoint_batch.restype = ctypes.c_int
    L.sr_far_adjoint_batch.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                       ctypes.POINTER(SrOpts), ctypes.POINTER(_vp)] + [_vp] * 7 + [
                                       _vp, ctypes.c_size_t, _vp]
    L.sr_far_fields_workspace_bytes.restype = ctypes.c_int
    L.sr_far_fields_workspace_bytes.argtypes = L.sr_workspace_bytes.argtypes
    L.sr_far_fields_batch.restype = ctypes.c_int
    L.sr_far_fields_batch.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                      ctypes.POINTER(SrOpts), _vp, ctypes.POINTER(_vp), _vp,
                                      ctypes.c_size_t, _vp]
    L.sr_far_adjoint_fields_workspace_bytes.restype = ctypes.c_int
    L.sr_far_adjoint_fields_workspace_bytes.argtypes = L.sr_workspace_bytes.argtypes
    L.sr_far_adjoint_from_fields.restype = ctypes.c_int
    L.sr_far_adjoint_from_fields.argtypes = [ctypes.POINTER(SrTracks), ctypes.POINTER(SrGrid),
                                             ctypes.POINTER(SrOpts), ctypes.POINTER(_vp), _vp] + \
        [_vp] * 7 + [_vp, ctypes.c_size_t, _vp]
    v = L.sr_abi_version()
    if v != SR_ABI_VERSION:
        raise SynchradError(f"ABI version mismatch: library {v}, binding {SR_ABI_VERSION}")
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != SR_OK:
        msg = load().sr_last_error().decode(errors="replace")
        raise SynchradError(f"synchrad_amd error {rc}: {msg}")
