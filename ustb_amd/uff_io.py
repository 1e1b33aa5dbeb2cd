"""UFF files: uff.write_object / uff.read_object (+uff/write_object.m,
+uff/read_object.m) for the objects midprocess.das reads, over the HDF5 C
library (libhdf5, through ctypes).

The file format is the reference's:
- root attribute ``version`` = uff.version (version.m);
- an object is a group with attributes ``class``, ``name``, ``size`` and
  ``array`` (0), holding one entry per non-empty property;
- an array of objects is a group (``array`` 1, ``size`` [1 n]) of entries
  ``<name>_0001`` ...;
- numbers are single-precision datasets (attributes class, name, complex,
  imaginary); complex numbers are groups with ``real`` and ``imag`` datasets;
- char is a single dataset of character codes; enumerations
  (uff.window, uff.wavefront) are single datasets.
MATLAB writes an array of size [d1 ... dk] (trailing singleton dimensions
dropped) as an HDF5 dataset of dimensions (dk ... d1): channel data [L N W F]
is stored (F, W, N, L), the layout this package keeps in device memory, so
frames are read straight into it.
"""
from __future__ import annotations

import ctypes
import glob
import math
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int64, c_size_t, c_uint, c_uint64, c_void_p

import numpy as np

from .enums import Wavefront, Window
from .uff import ChannelData, LinearArray, MatrixArray, Point, Probe, Wave

UFF_VERSION = "v1.2.0"

hid_t = c_int64
hsize_t = c_uint64
herr_t = c_int
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0, 1, 2
H5I_GROUP, H5I_DATASET = 2, 5
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5S_SELECT_SET = 0
H5_INDEX_NAME, H5_ITER_INC = 0, 0


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", c_int), ("nlinks", hsize_t), ("max_corder", c_int64), ("mounted", ctypes.c_bool)]


_SIGS = {
    "H5open": (herr_t, []),
    "H5Fcreate": (hid_t, [c_char_p, c_uint, hid_t, hid_t]),
    "H5Fopen": (hid_t, [c_char_p, c_uint, hid_t]),
    "H5Fclose": (herr_t, [hid_t]),
    "H5Gcreate2": (hid_t, [hid_t, c_char_p, hid_t, hid_t, hid_t]),
    "H5Gopen2": (hid_t, [hid_t, c_char_p, hid_t]),
    "H5Gclose": (herr_t, [hid_t]),
    "H5Gget_info": (herr_t, [hid_t, POINTER(_GInfo)]),
    "H5Lexists": (c_int, [hid_t, c_char_p, hid_t]),
    "H5Lget_name_by_idx": (ctypes.c_ssize_t, [hid_t, c_char_p, c_int, c_int, hsize_t, c_char_p, c_size_t, hid_t]),
    "H5Oopen": (hid_t, [hid_t, c_char_p, hid_t]),
    "H5Oclose": (herr_t, [hid_t]),
    "H5Iget_type": (c_int, [hid_t]),
    "H5Dcreate2": (hid_t, [hid_t, c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
    "H5Dopen2": (hid_t, [hid_t, c_char_p, hid_t]),
    "H5Dget_space": (hid_t, [hid_t]),
    "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, c_void_p]),
    "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, c_void_p]),
    "H5Dclose": (herr_t, [hid_t]),
    "H5Screate_simple": (hid_t, [c_int, POINTER(hsize_t), POINTER(hsize_t)]),
    "H5Sget_simple_extent_ndims": (c_int, [hid_t]),
    "H5Sget_simple_extent_dims": (c_int, [hid_t, POINTER(hsize_t), POINTER(hsize_t)]),
    "H5Sselect_hyperslab": (herr_t, [hid_t, c_int, POINTER(hsize_t), POINTER(hsize_t), POINTER(hsize_t),
                                     POINTER(hsize_t)]),
    "H5Sclose": (herr_t, [hid_t]),
    "H5Acreate2": (hid_t, [hid_t, c_char_p, hid_t, hid_t, hid_t, hid_t]),
    "H5Aopen": (hid_t, [hid_t, c_char_p, hid_t]),
    "H5Aexists": (c_int, [hid_t, c_char_p]),
    "H5Awrite": (herr_t, [hid_t, hid_t, c_void_p]),
    "H5Aread": (herr_t, [hid_t, hid_t, c_void_p]),
    "H5Aget_type": (hid_t, [hid_t]),
    "H5Aget_space": (hid_t, [hid_t]),
    "H5Aclose": (herr_t, [hid_t]),
    "H5Tcopy": (hid_t, [hid_t]),
    "H5Tset_size": (herr_t, [hid_t, c_size_t]),
    "H5Tget_size": (c_size_t, [hid_t]),
    "H5Tget_class": (c_int, [hid_t]),
    "H5Tis_variable_str": (c_int, [hid_t]),
    "H5Tclose": (herr_t, [hid_t]),
    "H5free_memory": (herr_t, [c_void_p]),
}


class _H5:
    """The HDF5 C library (1.10 API, hid_t = int64_t)."""

    def __init__(self):
        path = os.environ.get("USTBX_HDF5_LIB")
        cands = [path] if path else []
        cands += ["libhdf5.so"] + sorted(glob.glob("/opt/conda/lib/libhdf5.so*")) + \
            sorted(glob.glob("/usr/lib/x86_64-linux-gnu/libhdf5*.so*"))
        err = None
        for c in cands:
            try:
                self.lib = ctypes.CDLL(c)
                break
            except OSError as e:
                err = e
        else:
            raise OSError(f"the HDF5 library was not found ({err}); set USTBX_HDF5_LIB")
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
        if self.lib.H5open() < 0:
            raise OSError("H5open failed")
        g = lambda n: hid_t.in_dll(self.lib, n).value
        self.FLOAT, self.DOUBLE, self.C_S1 = g("H5T_NATIVE_FLOAT_g"), g("H5T_NATIVE_DOUBLE_g"), g("H5T_C_S1_g")

    def __getattr__(self, name):
        return getattr(self.lib, name)


_h5 = None


def h5() -> _H5:
    global _h5
    if _h5 is None:
        _h5 = _H5()
    return _h5


def _ok(rc, what):
    if rc < 0:
        raise OSError(f"HDF5: {what} failed")
    return rc


def _dims(a):
    return (hsize_t * len(a))(*[int(x) for x in a])


# ---------------------------------------------------------------------------
# attributes and datasets
# ---------------------------------------------------------------------------

def _attr_str(loc: int, name: str, value: str):
    L = h5()
    t = L.H5Tcopy(L.C_S1)
    data = value.encode()
    L.H5Tset_size(t, max(1, len(data)))
    sp = L.H5Screate_simple(1, _dims([1]), None)
    a = _ok(L.H5Acreate2(loc, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
    buf = ctypes.create_string_buffer(data, max(1, len(data)))
    _ok(L.H5Awrite(a, t, buf), f"attribute {name}")
    L.H5Aclose(a)
    L.H5Sclose(sp)
    L.H5Tclose(t)


def _attr_num(loc: int, name: str, values):
    L = h5()
    v = np.ascontiguousarray(np.atleast_1d(np.asarray(values, dtype=np.float64)))
    sp = L.H5Screate_simple(1, _dims([v.size]), None)
    a = _ok(L.H5Acreate2(loc, name.encode(), L.DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
    _ok(L.H5Awrite(a, L.DOUBLE, v.ctypes.data), f"attribute {name}")
    L.H5Aclose(a)
    L.H5Sclose(sp)


def _read_attr(loc: int, name: str):
    """A string attribute as str (trailing NULs removed, read_object.m:19),
    a numeric one as a float64 array; None when absent."""
    L = h5()
    if L.H5Aexists(loc, name.encode()) <= 0:
        return None
    a = L.H5Aopen(loc, name.encode(), H5P_DEFAULT)
    t = L.H5Aget_type(a)
    sp = L.H5Aget_space(a)
    nd = L.H5Sget_simple_extent_ndims(sp)
    d = (hsize_t * max(nd, 1))()
    L.H5Sget_simple_extent_dims(sp, d, None)
    n = int(np.prod([d[i] for i in range(nd)])) if nd > 0 else 1
    try:
        if L.H5Tget_class(t) == H5T_STRING:
            if L.H5Tis_variable_str(t) > 0:
                ptrs = (c_char_p * n)()
                mt = L.H5Tcopy(L.C_S1)
                L.H5Tset_size(mt, ctypes.c_size_t(-1).value)  # H5T_VARIABLE
                _ok(L.H5Aread(a, mt, ptrs), f"attribute {name}")
                L.H5Tclose(mt)
                return ptrs[0].decode().rstrip("\x00") if ptrs[0] else ""
            size = L.H5Tget_size(t)
            buf = ctypes.create_string_buffer(size * n)
            _ok(L.H5Aread(a, t, buf), f"attribute {name}")
            return buf.raw[:size].split(b"\x00")[0].decode()
        out = np.zeros(n, dtype=np.float64)
        _ok(L.H5Aread(a, L.DOUBLE, out.ctypes.data), f"attribute {name}")
        return out
    finally:
        L.H5Sclose(sp)
        L.H5Tclose(t)
        L.H5Aclose(a)


def _hdf5_dims(matlab_shape) -> list[int]:
    """MATLAB size with trailing singletons dropped (at least two
    dimensions), reversed: the dataset's dimensions."""
    s = list(matlab_shape) if len(matlab_shape) >= 2 else list(matlab_shape) + [1] * (2 - len(matlab_shape))
    while len(s) > 2 and s[-1] == 1:
        s.pop()
    return s[::-1]


def _write_dataset(loc: int, name: str, c_order: np.ndarray, dims: list[int]):
    L = h5()
    sp = L.H5Screate_simple(len(dims), _dims(dims), None)
    ds = _ok(L.H5Dcreate2(loc, name.encode(), L.FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"dataset {name}")
    buf = np.ascontiguousarray(c_order, dtype=np.float32)
    _ok(L.H5Dwrite(ds, L.FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data), f"dataset {name}")
    L.H5Sclose(sp)
    return ds


def _read_dataset(loc: int, name: str, first: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
    """The dataset in its stored (C-order) dimensions as float32; with
    `first`, only index `first` of the leading dimension (a hyperslab). With
    `out` (C-contiguous float32 holding exactly that many values), HDF5 reads
    into it and it is returned reshaped to those dimensions."""
    L = h5()
    ds = _ok(L.H5Dopen2(loc, name.encode(), H5P_DEFAULT), f"dataset {name}")
    sp = L.H5Dget_space(ds)
    nd = L.H5Sget_simple_extent_ndims(sp)
    d = (hsize_t * nd)()
    L.H5Sget_simple_extent_dims(sp, d, None)
    dims = [int(d[i]) for i in range(nd)]
    try:
        count = dims if first is None else [1] + dims[1:]
        if out is None:
            out = np.empty(count, dtype=np.float32)
        elif out.dtype != np.float32 or not out.flags.c_contiguous or out.size != int(np.prod(count)):
            raise ValueError(f"out must be C-contiguous float32 with {int(np.prod(count))} values for {name}")
        else:
            out = out.reshape(count)
        if first is None:
            _ok(L.H5Dread(ds, L.FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), f"dataset {name}")
            return out
        if not 0 <= first < dims[0]:
            raise IndexError(f"index {first} outside the leading dimension {dims[0]} of {name}")
        _ok(L.H5Sselect_hyperslab(sp, H5S_SELECT_SET, _dims([first] + [0] * (nd - 1)), None, _dims(count), None),
            "hyperslab")
        mem = L.H5Screate_simple(nd, _dims(count), None)
        try:
            _ok(L.H5Dread(ds, L.FLOAT, mem, sp, H5P_DEFAULT, out.ctypes.data), f"dataset {name}")
        finally:
            L.H5Sclose(mem)
        return out
    finally:
        L.H5Sclose(sp)
        L.H5Dclose(ds)


def _children(gid: int) -> list[tuple[str, int]]:
    """(name, H5I type) of a group's members in name order (h5info's order)."""
    L = h5()
    info = _GInfo()
    _ok(L.H5Gget_info(gid, byref(info)), "group info")
    out = []
    for i in range(info.nlinks):
        n = L.H5Lget_name_by_idx(gid, b".", H5_INDEX_NAME, H5_ITER_INC, i, None, 0, H5P_DEFAULT)
        buf = ctypes.create_string_buffer(n + 1)
        L.H5Lget_name_by_idx(gid, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, n + 1, H5P_DEFAULT)
        name = buf.value.decode()
        o = L.H5Oopen(gid, name.encode(), H5P_DEFAULT)
        out.append((name, L.H5Iget_type(o)))
        L.H5Oclose(o)
    return out


class _File:
    def __init__(self, path: str, mode: str = "r"):
        L = h5()
        p = os.fsencode(path)
        if mode == "w" or (mode == "a" and not os.path.exists(path)):
            self.id = _ok(L.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {path}")
            _attr_str(self.id, "version", UFF_VERSION)
        else:
            if not os.path.exists(path):
                raise FileNotFoundError(f"UFF: The file {path} does not exist.")
            self.id = _ok(L.H5Fopen(p, H5F_ACC_RDWR if mode == "a" else H5F_ACC_RDONLY, H5P_DEFAULT), f"open {path}")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        h5().H5Fclose(self.id)


# ---------------------------------------------------------------------------
# write_object
# ---------------------------------------------------------------------------

def _geometry(p: Probe) -> np.ndarray:
    """probe.geometry [N 7]: x y z theta phi width height (probe.m:42)."""
    g = np.zeros((p.N_elements, 7))
    g[:, 0], g[:, 1], g[:, 2] = p.x, p.y, p.z
    return g


def _properties(obj) -> tuple[str, list[tuple[str, object]]]:
    """(class, [(property, value)]) of the objects this module writes, in
    the reference classes' property order."""
    if isinstance(obj, Point):
        return "uff.point", [("distance", obj.distance), ("azimuth", obj.azimuth), ("elevation", obj.elevation)]
    if isinstance(obj, Wave):
        return "uff.wave", [("wavefront", Wavefront(int(obj.wavefront))), ("source", obj.source),
                            ("origin", obj.origin), ("delay", obj.delay), ("sound_speed", obj.sound_speed)]
    if isinstance(obj, LinearArray):
        return "uff.linear_array", [("geometry", _geometry(obj)), ("N", obj.N_elements), ("pitch", obj.pitch)]
    if isinstance(obj, MatrixArray):
        return "uff.matrix_array", [("geometry", _geometry(obj)), ("pitch_x", obj.pitch_x),
                                    ("pitch_y", obj.pitch_y), ("N_x", obj.N_x), ("N_y", obj.N_y)]
    if isinstance(obj, Probe):
        return "uff.probe", [("geometry", _geometry(obj))]
    if isinstance(obj, ChannelData):
        data = obj.data
        if data is None and obj.device_data is not None:
            data = obj.device_data.cpu().numpy().transpose(3, 2, 1, 0)
        return "uff.channel_data", [("sampling_frequency", obj.sampling_frequency),
                                    ("initial_time", obj.initial_time), ("sound_speed", obj.sound_speed),
                                    ("modulation_frequency", obj.modulation_frequency),
                                    ("sequence", list(obj.sequence)), ("probe", obj.probe), ("data", data)]
    raise TypeError(f"Class {type(obj).__name__} not supported by UFF; skipping write.")


def _write(gid: int, obj, name: str) -> int:
    """write_object.m:50-135 under the open group gid; returns the number of
    objects written."""
    L = h5()
    if obj is None:
        return 0
    if isinstance(obj, (Window, Wavefront)):
        ds = _write_dataset(gid, name, np.array([[float(int(obj))]]), [1, 1])
        _attr_str(ds, "class", "uff.window" if isinstance(obj, Window) else "uff.wavefront")
        _attr_str(ds, "name", name)
        L.H5Dclose(ds)
        return 1
    if isinstance(obj, str):
        codes = np.array([[float(ord(c)) for c in obj]]) if obj else np.zeros((1, 0))
        ds = _write_dataset(gid, name, codes, _hdf5_dims((1, len(obj))))
        _attr_str(ds, "class", "char")
        _attr_str(ds, "name", name)
        L.H5Dclose(ds)
        return 1
    if isinstance(obj, (int, float, np.integer, np.floating, np.ndarray)):
        a = np.asarray(obj)
        if a.size == 0:
            return 0
        cls = "single" if a.dtype in (np.float32, np.complex64) else "double"
        mshape = a.shape if a.ndim >= 2 else (1, a.size) if a.ndim == 1 else (1, 1)
        dims = _hdf5_dims(mshape)
        c_order = np.asarray(a).reshape(mshape).transpose(tuple(range(len(mshape)))[::-1]) if a.ndim else a.reshape(1, 1)
        if not np.iscomplexobj(a):
            ds = _write_dataset(gid, name, c_order, dims)
            for k, v in (("class", cls), ("name", name)):
                _attr_str(ds, k, v)
            _attr_num(ds, "imaginary", 0)
            _attr_num(ds, "complex", 0)
            L.H5Dclose(ds)
            return 1
        g = _ok(L.H5Gcreate2(gid, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {name}")
        for part, vals, imag in (("real", c_order.real, 0), ("imag", c_order.imag, 1)):
            ds = _write_dataset(g, part, vals, dims)
            _attr_str(ds, "class", cls)
            _attr_str(ds, "name", name)
            _attr_num(ds, "imaginary", imag)
            L.H5Dclose(ds)
        _attr_str(g, "class", cls)
        _attr_str(g, "name", name)
        _attr_num(g, "complex", 1)
        L.H5Gclose(g)
        return 1
    if isinstance(obj, (list, tuple)):
        items = [o for o in obj]
        if not items:
            return 0
        if len(items) == 1:  # numel(object) == 1: written as the object itself (write_object.m:104)
            return _write(gid, items[0], name)
        g = _ok(L.H5Gcreate2(gid, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {name}")
        n = 0
        for i, o in enumerate(items):
            n += _write(g, o, f"{name}_{i + 1:04d}")
        if n:
            _attr_str(g, "class", _properties(items[0])[0])
            _attr_str(g, "name", name)
            _attr_num(g, "array", 1)
            _attr_num(g, "size", [1, len(items)])
        L.H5Gclose(g)
        return n
    cls, props = _properties(obj)
    g = _ok(L.H5Gcreate2(gid, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {name}")
    n = 0
    for pname, value in props:
        if value is not None and not (isinstance(value, np.ndarray) and value.size == 0):
            n += _write(g, value, pname) > 0
    if n:
        _attr_str(g, "class", cls)
        _attr_str(g, "name", name)
        _attr_num(g, "size", [1, 1])
        _attr_num(g, "array", 0)
    L.H5Gclose(g)
    return n


def write_object(filename: str, obj, name: str, location: str = ""):
    """uff.write_object(filename, object, name, location): creates the file
    with its version attribute when absent (write_object.m:33-39)."""
    with _File(filename, "a") as f:
        v = _read_attr(f.id, "version")
        if v != UFF_VERSION:
            raise ValueError(f"UFF: Unsupported file version ({v}). Current UFF version ({UFF_VERSION}).")
        L = h5()
        loc = f.id
        opened = []
        for part in [p for p in location.split("/") if p]:
            if L.H5Lexists(loc, part.encode(), H5P_DEFAULT) > 0:
                loc = L.H5Gopen2(loc, part.encode(), H5P_DEFAULT)
            else:
                loc = L.H5Gcreate2(loc, part.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
            opened.append(loc)
        if L.H5Lexists(loc, name.encode(), H5P_DEFAULT) > 0:
            raise FileExistsError(f"UFF: {location}/{name} already exists in the file")
        try:
            return _write(loc, obj, name)
        finally:
            for g in reversed(opened):
                L.H5Gclose(g)


# ---------------------------------------------------------------------------
# read_object
# ---------------------------------------------------------------------------

class Node(dict):
    """A UFF object read back: its properties, with ``cls`` and ``name``."""

    def __init__(self, cls: str, name: str):
        super().__init__()
        self.cls, self.name = cls, name


def _read(gid: int, name: str, kind: int):
    """read_object.m:48-150 for one entry of an open group. Numbers come back
    in the dataset's stored (C) order, i.e. the MATLAB array with its
    dimensions reversed."""
    L = h5()
    if kind == H5I_DATASET:
        ds = L.H5Dopen2(gid, name.encode(), H5P_DEFAULT)
        cls = _read_attr(ds, "class")
        L.H5Dclose(ds)
        data = _read_dataset(gid, name)
        if cls == "char":
            return "".join(chr(int(c)) for c in data.ravel())
        if cls == "uff.window":
            return Window(int(data.ravel()[0]))
        if cls == "uff.wavefront":
            return Wavefront(int(data.ravel()[0]))
        return data
    g = L.H5Gopen2(gid, name.encode(), H5P_DEFAULT)
    try:
        cls = _read_attr(g, "class")
        nm = _read_attr(g, "name")
        if cls is None or nm is None:
            return None  # unnamed locations are skipped (read_object.m:44-47)
        if cls in ("double", "single", "int16"):
            c = _read_attr(g, "complex")
            if c is not None and c[0]:
                re, im = _read_dataset(g, "real"), _read_dataset(g, "imag")
                out = np.empty(re.shape, dtype=np.complex64)
                out.real, out.imag = re, im
                return out
            return None
        kids = _children(g)
        size = _read_attr(g, "size")
        n = int(np.prod(size)) if size is not None else 1
        if n > 1 or (_read_attr(g, "array") is not None and _read_attr(g, "array")[0]):
            if len(kids) != n:
                raise ValueError("Size attribute does not match number of subgroups")
            return [_read(g, k, t) for k, t in kids]
        node = Node(cls, nm)
        for k, t in kids:
            node[k] = _read(g, k, t)
        return node
    finally:
        L.H5Gclose(g)


def read_object(filename: str, location: str):
    """uff.read_object(filename, location) for one location (read_object.m);
    objects come back as Node (a dict of properties)."""
    with _File(filename, "r") as f:
        v = _read_attr(f.id, "version")
        if v not in ("v1.0.0", "v1.0.1", "v1.1.0", "v1.1.1", "v1.2.0"):
            raise ValueError(f"UFF: Unsupported file version ({v}). Current UFF version ({UFF_VERSION}).")
        L = h5()
        path = location.strip("/")
        parent, _, leaf = path.rpartition("/")
        g = f.id if not parent else L.H5Gopen2(f.id, parent.encode(), H5P_DEFAULT)
        try:
            if L.H5Lexists(g, leaf.encode(), H5P_DEFAULT) <= 0:
                raise KeyError("UFF: The location does not exists")
            o = L.H5Oopen(g, leaf.encode(), H5P_DEFAULT)
            kind = L.H5Iget_type(o)
            L.H5Oclose(o)
            return _read(g, leaf, kind)
        finally:
            if parent:
                L.H5Gclose(g)


def _scalar(node, key, default):
    v = node.get(key) if node is not None else None
    return float(np.asarray(v).ravel()[0]) if v is not None and np.asarray(v).size else default


def _point(node) -> Point:
    return Point(_scalar(node, "distance", 0.0), _scalar(node, "azimuth", 0.0), _scalar(node, "elevation", 0.0))


def _wave(node) -> Wave:
    w = Wave(node.get("wavefront", Wavefront.spherical), _point(node.get("source")), _point(node.get("origin")),
             _scalar(node, "delay", 0.0), _scalar(node, "sound_speed", 1540.0))
    # wave.fix_origin_from_source (wave.m:89-114), as read_object.m:146-148 calls it
    s = w.source
    if (w.wavefront == Wavefront.spherical and math.isfinite(s.distance) and all(v == 0 for v in w.origin.xyz)
            and any(v != 0 for v in s.xyz)):
        w.origin = Point.from_xyz(s.x, s.y, 0.0 if s.z > 0 else s.z)
    return w


def _probe(node) -> Probe:
    geo = node.get("geometry")
    if node.cls == "uff.linear_array" and "N" in node and "pitch" in node:
        p = LinearArray(int(_scalar(node, "N", 0)), _scalar(node, "pitch", 0.0))
    elif node.cls == "uff.matrix_array" and "N_x" in node:
        p = MatrixArray(int(_scalar(node, "N_x", 0)), int(_scalar(node, "N_y", 0)), _scalar(node, "pitch_x", 0.0),
                        _scalar(node, "pitch_y", 0.0))
    elif geo is not None:
        g = np.asarray(geo, dtype=np.float64)  # stored (7, N): rows x, y, z, ...
        p = Probe(g[0], g[1], g[2])
    else:
        raise ValueError(f"probe of class {node.cls} without geometry")
    if geo is not None:
        g = np.asarray(geo, dtype=np.float64)
        if g.shape[-1] == p.N_elements and not np.allclose(g[0], p.x, rtol=0, atol=1e-9):
            p = Probe(g[0], g[1], g[2])  # the stored element positions win
    return p


def channel_data_from(node: Node) -> ChannelData:
    """A uff.channel_data Node as ChannelData; data keeps the stored layout
    (F, W, N, L), presented in the reference layout [L, N, W, F]."""
    data = node.get("data")
    if data is not None:
        a = np.asarray(data)
        a = a.reshape((1,) * (4 - a.ndim) + a.shape)  # dropped trailing singletons of [L N W F]
        data = a.transpose(3, 2, 1, 0)
    seq = node.get("sequence")
    seq = [seq] if isinstance(seq, Node) else (seq or [])
    return ChannelData(probe=_probe(node["probe"]), sequence=[_wave(w) for w in seq],
                       sampling_frequency=_scalar(node, "sampling_frequency", 0.0),
                       initial_time=_scalar(node, "initial_time", 0.0), sound_speed=_scalar(node, "sound_speed", 1540.0),
                       modulation_frequency=_scalar(node, "modulation_frequency", 0.0), data=data)


def read_channel_data(filename: str, location: str = "/channel_data") -> ChannelData:
    node = read_object(filename, location)
    if not isinstance(node, Node) or node.cls != "uff.channel_data":
        raise ValueError(f"{location} is not a uff.channel_data object")
    return channel_data_from(node)


class FrameReader:
    """Frames of a UFF channel data set one at a time (a hyperslab of the
    stored (F, W, N, L) datasets), for streaming into the serving loop."""

    def __init__(self, filename: str, location: str = "/channel_data"):
        self.filename = filename
        self.location = location.strip("/") + "/data"
        with _File(filename, "r") as f:
            L = h5()
            if L.H5Lexists(f.id, location.strip("/").encode(), H5P_DEFAULT) <= 0 or \
                    L.H5Lexists(f.id, self.location.encode(), H5P_DEFAULT) <= 0:
                raise KeyError(f"UFF: {location} has no data")
            o = L.H5Oopen(f.id, self.location.encode(), H5P_DEFAULT)
            self.complex = L.H5Iget_type(o) == H5I_GROUP  # real/imag pair
            L.H5Oclose(o)
            ds = L.H5Dopen2(f.id, (self.location + ("/real" if self.complex else "")).encode(), H5P_DEFAULT)
            sp = L.H5Dget_space(ds)
            nd = L.H5Sget_simple_extent_ndims(sp)
            d = (hsize_t * nd)()
            L.H5Sget_simple_extent_dims(sp, d, None)
            L.H5Sclose(sp)
            L.H5Dclose(ds)
        dims = [int(d[i]) for i in range(nd)]
        # [L N W F] with trailing singletons dropped, reversed
        self.shape = tuple([1] * (4 - len(dims)) + dims)
        self.frames = self.shape[0]
        self._parts = None  # IQ: the real and imag staging buffers, made on first use

    def frame(self, f: int, out: np.ndarray | None = None) -> np.ndarray:
        """Frame f as (W, N, L): complex64 for IQ data, float32 for RF.

        With `out` (a C-contiguous array of that shape and type, e.g. the
        NumPy view of a pinned staging buffer) the frame is written there:
        RF is read straight into it; IQ's real and imag datasets are read
        into two float32 buffers the reader keeps and then interleaved into
        it, so a stream of frames allocates nothing per frame."""
        if not 0 <= f < self.frames:
            raise IndexError(f"frame {f} of {self.frames}")
        shape = self.shape[1:]
        dtype = np.complex64 if self.complex else np.float32
        if out is None:
            out = np.empty(shape, dtype=dtype)
        elif out.dtype != dtype or tuple(out.shape) != shape or not out.flags.c_contiguous:
            raise ValueError(f"out must be a C-contiguous {np.dtype(dtype).name} array of shape {shape}")
        group, _, leaf = self.location.rpartition("/")
        if self.complex:
            group, names = self.location, ("real", "imag")
            if self._parts is None:
                self._parts = (np.empty(shape, np.float32), np.empty(shape, np.float32))
            dests = self._parts
        else:
            names, dests = (leaf,), (out,)
        with _File(self.filename, "r") as fh:
            L = h5()
            g = L.H5Gopen2(fh.id, group.encode(), H5P_DEFAULT)
            try:
                for name, dst in zip(names, dests):
                    # a stored 4-D dataset is read one frame at a time; with
                    # the frame dimension dropped (one frame) it is the frame
                    four = self._stored_ndim(g, name) == 4
                    _read_dataset(g, name, f if four else None, out=dst)
            finally:
                L.H5Gclose(g)
        if self.complex:
            pair = out.view(np.float32).reshape(shape + (2,))
            pair[..., 0] = self._parts[0]
            pair[..., 1] = self._parts[1]
        return out

    def planes(self, f: int, out: np.ndarray | None = None) -> np.ndarray:
        """IQ frame f as its stored real and imag planes, (2, W, N, L)
        float32, each read by HDF5 straight into its half of `out` (or of a
        new array): no host-side interleave. A caller that copies the
        planes to the GPU builds the complex frame there (the interleave at
        device-memory speed instead of a strided host pass)."""
        if not self.complex:
            raise ValueError("planes() is for IQ (real/imag) channel data; RF frames come from frame()")
        if not 0 <= f < self.frames:
            raise IndexError(f"frame {f} of {self.frames}")
        shape = (2,) + self.shape[1:]
        if out is None:
            out = np.empty(shape, dtype=np.float32)
        elif out.dtype != np.float32 or tuple(out.shape) != shape or not out.flags.c_contiguous:
            raise ValueError(f"out must be a C-contiguous float32 array of shape {shape}")
        with _File(self.filename, "r") as fh:
            L = h5()
            g = L.H5Gopen2(fh.id, self.location.encode(), H5P_DEFAULT)
            try:
                for k, name in enumerate(("real", "imag")):
                    four = self._stored_ndim(g, name) == 4
                    _read_dataset(g, name, f if four else None, out=out[k])
            finally:
                L.H5Gclose(g)
        return out

    @staticmethod
    def _stored_ndim(g: int, name: str) -> int:
        L = h5()
        ds = L.H5Dopen2(g, name.encode(), H5P_DEFAULT)
        sp = L.H5Dget_space(ds)
        nd = L.H5Sget_simple_extent_ndims(sp)
        L.H5Sclose(sp)
        L.H5Dclose(ds)
        return nd


def serving_frames(filename: str, location: str = "/channel_data", frames=None) -> list[np.ndarray]:
    """A UFF channel data set's frames as the serving loop's per-request
    payloads: one (1, W, N, L) array per frame, read frame by frame."""
    fr = FrameReader(filename, location)
    sel = range(fr.frames) if frames is None else frames
    return [fr.frame(f)[None] for f in sel]
