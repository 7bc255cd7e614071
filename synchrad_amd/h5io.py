"""HDF5 track and spectrum files through the HDF5 C library (ctypes).

The reference reads and writes its track files (``tracks/{i}/{x,y,z,ux,uy,uz,
w,it_start}``, ``misc/{cdt,N_particles,it_range,...}``, synchrad/converters.py:
101-127) and spectrum files (``radiation/<key>``, ``Args/<key>``,
``snap_iterations``, ``total_weight``, synchrad/calc.py:274-290, 648-666) with
h5py.  When h5py is not importable, :class:`File` offers the part of its
interface those call sites use, on top of ``libhdf5`` found at run time:

    f = File(path, "r")            # or "w" (truncate)
    f["misc/cdt"][()]              # dataset -> numpy scalar / array (strings -> bytes)
    "it_range" in f["misc"].keys() # group member names, in name order
    f["radiation/total"] = array   # creates missing groups, like h5py
    f.close()                      # or use it as a context manager

Values are stored as h5py stores them: numpy arrays and scalars with their own
type (bool as uint8), Python floats and ints as float64 / int64, str as a
variable-length UTF-8 string, lists and tuples through numpy (a list of str as
an array of variable-length strings).

``SYNCHRAD_AMD_HDF5_LIB`` names the library explicitly; otherwise the usual
names and the conda prefix are tried.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import glob
import os
import sys
from typing import List

import numpy as np

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_uint64
_size = ctypes.c_size_t

H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0x0000, 0x0001, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5I_GROUP, H5I_DATASET = 2, 5
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0
H5T_CSET_UTF8 = 1
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5_INDEX_NAME, H5_ITER_INC = 0, 0


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", _hsize), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


_lib = None


def _candidates() -> List[str]:
    env = os.environ.get("SYNCHRAD_AMD_HDF5_LIB")
    if env:
        return [env]
    out = []
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda", sys.prefix, "/usr"):
        if prefix:
            out += sorted(glob.glob(os.path.join(prefix, "lib", "libhdf5.so*")))
            out += sorted(glob.glob(os.path.join(prefix, "lib", "x86_64-linux-gnu", "libhdf5*.so*")))
    return [c for c in out if "_hl" not in os.path.basename(c) and "fortran" not in os.path.basename(c)]


def lib():
    """Load libhdf5 (once) and declare the functions used here."""
    global _lib
    if _lib is not None:
        return _lib
    err = None
    for cand in _candidates():
        try:
            L = ctypes.CDLL(cand)
            break
        except OSError as e:  # try the next name
            err = e
    else:
        raise ImportError(f"h5 files need h5py or the HDF5 C library (libhdf5), neither was found ({err})")
    sig = {
        "H5open": (_herr, []),
        "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Gcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid]),
        "H5Gclose": (_herr, [_hid]),
        "H5Gget_info_by_name": (_herr, [_hid, ctypes.c_char_p, ctypes.POINTER(_GInfo), _hid]),
        "H5Lexists": (ctypes.c_int, [_hid, ctypes.c_char_p, _hid]),
        "H5Lget_name_by_idx": (ctypes.c_ssize_t, [_hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _hsize,
                                                  ctypes.c_char_p, _size, _hid]),
        "H5Oopen": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Oclose": (_herr, [_hid]),
        "H5Iget_type": (ctypes.c_int, [_hid]),
        "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Dclose": (_herr, [_hid]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dget_type": (_hid, [_hid]),
        "H5Dvlen_reclaim": (_herr, [_hid, _hid, _hid, ctypes.c_void_p]),
        "H5Screate": (_hid, [ctypes.c_int]),
        "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
        "H5Sclose": (_herr, [_hid]),
        "H5Tcopy": (_hid, [_hid]),
        "H5Tset_size": (_herr, [_hid, _size]),
        "H5Tset_cset": (_herr, [_hid, ctypes.c_int]),
        "H5Tget_class": (ctypes.c_int, [_hid]),
        "H5Tget_size": (_size, [_hid]),
        "H5Tget_sign": (ctypes.c_int, [_hid]),
        "H5Tis_variable_str": (ctypes.c_int, [_hid]),
        "H5Tclose": (_herr, [_hid]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    if L.H5open() < 0:
        raise OSError("H5open failed")
    _lib = L
    return L


def _native(name: str) -> int:
    return _hid.in_dll(lib(), f"H5T_NATIVE_{name}_g").value


_NATIVE = {np.dtype(np.float64): "DOUBLE", np.dtype(np.float32): "FLOAT", np.dtype(np.int64): "INT64",
           np.dtype(np.uint64): "UINT64", np.dtype(np.int32): "INT32", np.dtype(np.uint32): "UINT32",
           np.dtype(np.int8): "INT8", np.dtype(np.uint8): "UINT8", np.dtype(np.int16): "INT16",
           np.dtype(np.uint16): "UINT16"}


def _check(rc, what):
    if rc < 0:
        raise OSError(f"HDF5: {what} failed")
    return rc


def _vlen_str_type() -> int:
    L = lib()
    t = _check(L.H5Tcopy(_hid.in_dll(L, "H5T_C_S1_g").value), "H5Tcopy")
    _check(L.H5Tset_size(t, H5T_VARIABLE), "H5Tset_size")
    _check(L.H5Tset_cset(t, H5T_CSET_UTF8), "H5Tset_cset")
    return t


def _as_array(value) -> np.ndarray:
    """The numpy array h5py would store for ``value``."""
    if isinstance(value, str):
        return np.array(value, dtype=object)
    arr = np.asarray(value)
    if arr.dtype.kind == "U" or (arr.dtype.kind == "O" and arr.size and isinstance(arr.flat[0], str)):
        return arr.astype(object)
    if arr.dtype.kind == "b":
        return arr.astype(np.uint8)
    return arr


class Dataset:
    def __init__(self, file: "File", path: str):
        self.file, self.path = file, path

    def __getitem__(self, key):
        if key != () and key is not Ellipsis:
            return self[()][key]
        return self.file._read(self.path)

    @property
    def shape(self):
        return np.shape(self[()])


class Group:
    def __init__(self, file: "File", path: str):
        self.file, self.path = file, path

    def keys(self) -> List[str]:
        return self.file._members(self.path)

    def __contains__(self, name) -> bool:
        return name in self.keys()

    def __iter__(self):
        return iter(self.keys())

    def __getitem__(self, name):
        return self.file[f"{self.path.rstrip('/')}/{name}"]

    def __setitem__(self, name, value):
        self.file[f"{self.path.rstrip('/')}/{name}"] = value


class File(Group):
    """A small h5py.File stand-in: modes "r" and "w" (create, truncate)."""

    def __init__(self, path: str, mode: str = "r"):
        L = lib()
        name = os.fsencode(path)
        if mode == "r":
            self.fid = L.H5Fopen(name, H5F_ACC_RDONLY, H5P_DEFAULT)
        elif mode == "w":
            self.fid = L.H5Fcreate(name, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        else:
            raise ValueError("mode must be 'r' or 'w'")
        if self.fid < 0:
            raise OSError(f"cannot open {path!r} (mode {mode!r})")
        super().__init__(self, "/")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if getattr(self, "fid", -1) >= 0:
            lib().H5Fclose(self.fid)
            self.fid = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- navigation ---------------------------------------------------------
    def _exists(self, path: str) -> bool:
        L = lib()
        parts = [p for p in path.split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if L.H5Lexists(self.fid, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def _kind(self, path: str) -> int:
        L = lib()
        oid = L.H5Oopen(self.fid, path.encode(), H5P_DEFAULT)
        if oid < 0:
            raise KeyError(path)
        kind = L.H5Iget_type(oid)
        L.H5Oclose(oid)
        return kind

    def __getitem__(self, path):
        path = "/" + path.strip("/")
        if path == "/":
            return self
        if not self._exists(path):
            raise KeyError(f"{path!r} is not in the file")
        kind = self._kind(path)
        if kind == H5I_GROUP:
            return Group(self, path)
        if kind == H5I_DATASET:
            return Dataset(self, path)
        raise KeyError(f"{path!r} is neither a group nor a dataset")

    def _members(self, path: str) -> List[str]:
        L = lib()
        info = _GInfo()
        _check(L.H5Gget_info_by_name(self.fid, path.encode(), ctypes.byref(info), H5P_DEFAULT), "H5Gget_info")
        names = []
        buf = ctypes.create_string_buffer(1024)
        for i in range(info.nlinks):
            n = _check(L.H5Lget_name_by_idx(self.fid, path.encode(), H5_INDEX_NAME, H5_ITER_INC, i, buf,
                                            len(buf), H5P_DEFAULT), "H5Lget_name_by_idx")
            if n >= len(buf):
                buf = ctypes.create_string_buffer(n + 1)
                L.H5Lget_name_by_idx(self.fid, path.encode(), H5_INDEX_NAME, H5_ITER_INC, i, buf, len(buf),
                                     H5P_DEFAULT)
            names.append(buf.value.decode())
        return names

    # -- datasets -----------------------------------------------------------
    def _read(self, path: str):
        L = lib()
        did = _check(L.H5Dopen2(self.fid, path.encode(), H5P_DEFAULT), f"open {path}")
        sid = tid = -1
        try:
            sid = L.H5Dget_space(did)
            nd = L.H5Sget_simple_extent_ndims(sid)
            dims = (_hsize * max(nd, 1))()
            if nd > 0:
                L.H5Sget_simple_extent_dims(sid, dims, None)
            shape = tuple(int(dims[i]) for i in range(nd))
            tid = L.H5Dget_type(did)
            cls, size = L.H5Tget_class(tid), int(L.H5Tget_size(tid))
            if cls == H5T_STRING:
                n = int(np.prod(shape)) if shape else 1
                if L.H5Tis_variable_str(tid) > 0:
                    mt = _vlen_str_type()
                    ptrs = (ctypes.c_char_p * n)()
                    _check(L.H5Dread(did, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, ptrs), f"read {path}")
                    vals = [p if p is not None else b"" for p in ptrs]
                    L.H5Dvlen_reclaim(mt, sid, H5P_DEFAULT, ptrs)
                    L.H5Tclose(mt)
                else:
                    raw = ctypes.create_string_buffer(n * size)
                    _check(L.H5Dread(did, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, raw), f"read {path}")
                    vals = [raw.raw[i * size:(i + 1) * size].rstrip(b"\0") for i in range(n)]
                if not shape:
                    return vals[0]
                return np.array(vals, dtype=object).reshape(shape)
            if cls == H5T_FLOAT:
                dt = np.dtype({4: np.float32, 8: np.float64}[size])
            elif cls == H5T_INTEGER:
                signed = L.H5Tget_sign(tid) != H5T_SGN_NONE
                dt = np.dtype(f"{'i' if signed else 'u'}{size}")
            else:
                raise TypeError(f"{path}: HDF5 type class {cls} is not supported here")
            out = np.empty(shape, dtype=dt)
            if out.size:
                _check(L.H5Dread(did, _native(_NATIVE[dt]), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                 out.ctypes.data), f"read {path}")
            return out[()] if not shape else out
        finally:
            if tid >= 0:
                L.H5Tclose(tid)
            if sid >= 0:
                L.H5Sclose(sid)
            L.H5Dclose(did)

    def __setitem__(self, path, value):
        L = lib()
        path = "/" + path.strip("/")
        if self._exists(path):
            raise ValueError(f"{path!r} already exists")
        parts = [p for p in path.split("/") if p]
        cur = ""
        for p in parts[:-1]:   # missing parent groups, as h5py creates them
            cur += "/" + p
            if not self._exists(cur):
                _check(L.H5Gclose(_check(L.H5Gcreate2(self.fid, cur.encode(), H5P_DEFAULT, H5P_DEFAULT,
                                                      H5P_DEFAULT), f"create group {cur}")), "H5Gclose")
        arr = _as_array(value)
        if arr.ndim == 0:
            sid = _check(L.H5Screate(H5S_SCALAR), "H5Screate")
        else:
            dims = (_hsize * arr.ndim)(*arr.shape)
            sid = _check(L.H5Screate_simple(arr.ndim, dims, None), "H5Screate_simple")
        own_type = arr.dtype.kind == "O"
        try:
            if own_type:
                tid = _vlen_str_type()
                enc = [str(v).encode() for v in arr.ravel()]
                buf = (ctypes.c_char_p * max(len(enc), 1))(*enc)
            else:
                if arr.dtype not in _NATIVE:
                    arr = arr.astype(np.float64 if arr.dtype.kind == "f" else np.int64)
                arr = np.ascontiguousarray(arr)
                tid = _native(_NATIVE[arr.dtype])
                buf = ctypes.c_void_p(arr.ctypes.data)
            did = _check(L.H5Dcreate2(self.fid, path.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                         f"create {path}")
            try:
                if arr.size:
                    _check(L.H5Dwrite(did, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), f"write {path}")
            finally:
                L.H5Dclose(did)
                if own_type:
                    L.H5Tclose(tid)
        finally:
            L.H5Sclose(sid)
