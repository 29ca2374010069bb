"""HDF5 snapshot output (src/periodicflow/io.py:15-196).

The reference writes snapshots through shenfun's ``ShenfunFile`` (h5py) and,
when the run is over, ``reconfigure_dataset`` rewrites the file into a flat
layout.  ``HDF5Writer`` keeps the constructor, the methods and both
layouts, and writes through libhdf5's C API over ctypes (h5py is not
needed):

* while the run goes (``write``, io.py:42-43): one group per variable
  holding ``<d>D/<step>`` float64 datasets of the physical component and
  ``mesh/x0 .. x{d-1}``, the coordinates of each axis (shenfun's HDF5File
  layout, which io.py:56-95 reads back);
* after ``reconfigure_dataset`` (io.py:97-196): root datasets ``x, y[, z]``
  (the mesh coordinates, with the end point appended when ``periodic``) and
  ``t = linspace(t_start, t_end, number of snapshots)`` (io.py:121-127); one
  group per variable with a float64 dataset per snapshot, named by the step
  zero-padded to the digits of the last step (io.py:129-134), of shape N,
  or N + 1 on every axis with the first plane repeated at the end when
  ``periodic`` (utils.py:7-14).

One process writes the whole box (no MPI): ``comm`` is accepted and unused.
The library is looked up as $PF_HDF5_LIB, then ``ctypes.util`` ("hdf5"),
then the usual install locations; a writer made without one raises
OSError.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
from pathlib import Path

import numpy as np

from . import logger

_hid = ctypes.c_int64
_hsize = ctypes.c_ulonglong
_P_DEFAULT = 0
_S_ALL = 0
_ACC_RDONLY = 0x0000
_ACC_RDWR = 0x0001
_ACC_TRUNC = 0x0002
_INDEX_NAME = 0
_ITER_INC = 0

_CANDIDATES = (
    "/opt/conda/lib/libhdf5.so",
    "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
    "/usr/lib/x86_64-linux-gnu/libhdf5.so",
    "/usr/local/lib/libhdf5.so",
)


class _GroupInfo(ctypes.Structure):
    # H5G_info_t (H5Gpublic.h): storage type, link count, max creation order, mounted
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", _hsize), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


class _H5:
    """The few libhdf5 (>= 1.10: 64-bit hid_t) calls the writer needs."""

    def __init__(self, path):
        lib = ctypes.CDLL(path)
        self.lib = lib
        self.path = path
        sig = {
            "H5open": (ctypes.c_int, []),
            "H5get_libversion": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint)] * 3),
            "H5Eset_auto2": (ctypes.c_int, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
            "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
            "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
            "H5Fflush": (ctypes.c_int, [_hid, ctypes.c_int]),
            "H5Fclose": (ctypes.c_int, [_hid]),
            "H5Gcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid]),
            "H5Gopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Gclose": (ctypes.c_int, [_hid]),
            "H5Gget_info_by_name": (ctypes.c_int, [_hid, ctypes.c_char_p, ctypes.POINTER(_GroupInfo), _hid]),
            "H5Lget_name_by_idx": (ctypes.c_ssize_t, [_hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _hsize,
                                                      ctypes.c_char_p, ctypes.c_size_t, _hid]),
            "H5Lexists": (ctypes.c_int, [_hid, ctypes.c_char_p, _hid]),
            "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.c_void_p]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.c_void_p]),
            "H5Sclose": (ctypes.c_int, [_hid]),
            "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Dwrite": (ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            "H5Dread": (ctypes.c_int, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            "H5Dclose": (ctypes.c_int, [_hid]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        self._ok("H5open", lib.H5open())
        v = [ctypes.c_uint() for _ in range(3)]
        lib.H5get_libversion(*[ctypes.byref(x) for x in v])
        self.version = tuple(x.value for x in v)
        if self.version < (1, 10, 0):
            raise OSError(f"libhdf5 {self.version} at {path}: version 1.10 or later is needed")
        # errors come back as return codes, turned into OSError below
        lib.H5Eset_auto2(0, None, None)
        self.native_double = _hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
        self.f64le = _hid.in_dll(lib, "H5T_IEEE_F64LE_g").value

    @staticmethod
    def _ok(what, rc, name=""):
        if rc < 0:
            raise OSError(f"HDF5 {what} failed" + (f" for '{name}'" if name else ""))
        return rc

    # ---- files and groups
    def create(self, path):
        return self._ok("H5Fcreate", self.lib.H5Fcreate(str(path).encode(), _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT),
                        str(path))

    def open(self, path, write=False):
        return self._ok("H5Fopen", self.lib.H5Fopen(str(path).encode(), _ACC_RDWR if write else _ACC_RDONLY,
                                                    _P_DEFAULT), str(path))

    def flush(self, fid):
        self._ok("H5Fflush", self.lib.H5Fflush(fid, 1))

    def close(self, fid):
        self._ok("H5Fclose", self.lib.H5Fclose(fid))

    def exists(self, loc, path):
        """Whether every link on the '/'-separated path exists."""
        parts = [p for p in str(path).split("/") if p]
        for i in range(len(parts)):
            rc = self._ok("H5Lexists", self.lib.H5Lexists(loc, "/".join(parts[: i + 1]).encode(), _P_DEFAULT), path)
            if rc == 0:
                return False
        return True

    def require_group(self, loc, path):
        """Create the groups along ``path`` that do not exist yet."""
        parts = [p for p in str(path).split("/") if p]
        for i in range(len(parts)):
            sub = "/".join(parts[: i + 1])
            if not self.exists(loc, sub):
                g = self._ok("H5Gcreate2", self.lib.H5Gcreate2(loc, sub.encode(), _P_DEFAULT, _P_DEFAULT,
                                                               _P_DEFAULT), sub)
                self._ok("H5Gclose", self.lib.H5Gclose(g))

    def members(self, loc, path="."):
        """Link names of a group in increasing name order (h5py's order)."""
        info = _GroupInfo()
        self._ok("H5Gget_info_by_name", self.lib.H5Gget_info_by_name(loc, str(path).encode(), ctypes.byref(info),
                                                                      _P_DEFAULT), path)
        out = []
        for i in range(info.nlinks):
            n = self._ok("H5Lget_name_by_idx", self.lib.H5Lget_name_by_idx(
                loc, str(path).encode(), _INDEX_NAME, _ITER_INC, i, None, 0, _P_DEFAULT), path)
            buf = ctypes.create_string_buffer(n + 1)
            self._ok("H5Lget_name_by_idx", self.lib.H5Lget_name_by_idx(
                loc, str(path).encode(), _INDEX_NAME, _ITER_INC, i, buf, n + 1, _P_DEFAULT), path)
            out.append(buf.value.decode())
        return out

    # ---- float64 datasets
    def create_dataset(self, loc, path, shape, data=None):
        """A float64 (IEEE little-endian) dataset of ``shape`` at ``path``,
        filled from ``data`` when given."""
        shape = tuple(int(s) for s in shape)
        parent = str(path).rsplit("/", 1)[0] if "/" in str(path) else ""
        if parent:
            self.require_group(loc, parent)
        dims = (_hsize * max(1, len(shape)))(*shape)
        sp = self._ok("H5Screate_simple", self.lib.H5Screate_simple(len(shape), dims, None), path)
        try:
            ds = self._ok("H5Dcreate2", self.lib.H5Dcreate2(loc, str(path).encode(), self.f64le, sp, _P_DEFAULT,
                                                            _P_DEFAULT, _P_DEFAULT), path)
        finally:
            self.lib.H5Sclose(sp)
        try:
            if data is not None:
                self._write(ds, data, shape, path)
        finally:
            self._ok("H5Dclose", self.lib.H5Dclose(ds), path)

    def write_dataset(self, loc, path, data):
        """Overwrite the whole existing dataset at ``path``."""
        ds = self._ok("H5Dopen2", self.lib.H5Dopen2(loc, str(path).encode(), _P_DEFAULT), path)
        try:
            self._write(ds, data, self._shape(ds, path), path)
        finally:
            self._ok("H5Dclose", self.lib.H5Dclose(ds), path)

    def _write(self, ds, data, shape, path):
        a = np.ascontiguousarray(data, dtype=np.float64)
        if a.shape != tuple(shape):
            raise ValueError(f"HDF5 dataset '{path}' has shape {tuple(shape)}, data has {a.shape}")
        self._ok("H5Dwrite", self.lib.H5Dwrite(ds, self.native_double, _S_ALL, _S_ALL, _P_DEFAULT,
                                               a.ctypes.data_as(ctypes.c_void_p)), path)

    def _shape(self, ds, path):
        sp = self._ok("H5Dget_space", self.lib.H5Dget_space(ds), path)
        try:
            nd = self._ok("H5Sget_simple_extent_ndims", self.lib.H5Sget_simple_extent_ndims(sp), path)
            dims = (_hsize * max(1, nd))()
            self._ok("H5Sget_simple_extent_dims", self.lib.H5Sget_simple_extent_dims(sp, dims, None), path)
            return tuple(int(dims[i]) for i in range(nd))
        finally:
            self.lib.H5Sclose(sp)

    def read(self, loc, path):
        """A dataset as a float64 NumPy array."""
        ds = self._ok("H5Dopen2", self.lib.H5Dopen2(loc, str(path).encode(), _P_DEFAULT), path)
        try:
            out = np.empty(self._shape(ds, path), dtype=np.float64)
            self._ok("H5Dread", self.lib.H5Dread(ds, self.native_double, _S_ALL, _S_ALL, _P_DEFAULT,
                                                 out.ctypes.data_as(ctypes.c_void_p)), path)
            return out
        finally:
            self._ok("H5Dclose", self.lib.H5Dclose(ds), path)


_LIB = None


def hdf5():
    """The process's libhdf5 binding (loaded on first use)."""
    global _LIB
    if _LIB is None:
        tried = []
        env = os.environ.get("PF_HDF5_LIB")
        names = ([env] if env else []) + [ctypes.util.find_library("hdf5")] + list(_CANDIDATES)
        for p in names:
            if not p or p in tried:
                continue
            tried.append(p)
            try:
                _LIB = _H5(p)
                break
            except OSError:
                continue
        if _LIB is None:
            raise OSError("HDF5 output needs libhdf5 (1.10 or later); none could be loaded from: " + ", ".join(tried)
                          + " (set PF_HDF5_LIB to its path)")
    return _LIB


def periodic_bc(u):
    """utils.py:7-14: append the first plane at the end of every axis."""
    u = np.asarray(u)
    for axis in range(u.ndim):
        u = np.concatenate((u, np.take(u, [0], axis=axis)), axis=axis)
    return u


def _host(a):
    """A host float64 copy of an array or device tensor."""
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def h5_path(file_name):
    """The file a run named ``file_name`` writes: '.h5' appended unless present."""
    s = str(file_name)
    return Path(s if s.endswith(".h5") else s + ".h5")


class HDF5Writer:
    """io.py:15-196.

    ``solution`` = {'space': an object with ``dim`` and ``x`` (the mesh: one
    broadcastable coordinate array per axis), 'data': {name: [array]}};
    each array (NumPy or a device tensor) is read at every ``write``, so the
    caller refreshes it in place beforehand, as the reference's solver does
    with ``backward()`` (simulation.py:215-217)."""

    def __init__(self, comm, rank, size, file_name, solution={}, time_points=[0.0, 1.0],
                 end_points=[2 * np.pi, 2 * np.pi], periodic=False):
        self.comm = comm
        self.mpi_rank = rank
        self.mpi_size = size
        self.file_name = file_name
        self.solution = solution
        self.time_points = time_points
        self.end_points = end_points
        self.periodic = periodic
        self.path = h5_path(file_name)
        self._h5 = hdf5()
        space = solution["space"]
        self._dim = int(space.dim)
        self._mesh = [np.asarray(xi, dtype=np.float64).reshape(-1) for xi in space.x]
        # mode 'w' (io.py:36-40): the file is created (or emptied) now
        self._fid = self._h5.create(self.path)

    def write(self, step: int):
        """io.py:42-43: each variable's current data as ``<var>/<d>D/<step>``.
        After ``close`` (the end of a solve) the file is reopened and
        appended to, so a later solve keeps writing into the same file."""
        h5 = self._h5
        if self._fid is None:
            self._fid = h5.open(self.path, write=True)
        sd = f"{self._dim}D"
        for name, fields in self.solution["data"].items():
            for field in fields:
                a = _host(field)
                if not h5.exists(self._fid, f"{name}/mesh"):
                    for i, xi in enumerate(self._mesh):
                        h5.create_dataset(self._fid, f"{name}/mesh/x{i}", xi.shape, xi)
                h5.create_dataset(self._fid, f"{name}/{sd}/{int(step)}", a.shape, a)
        h5.flush(self._fid)

    def close(self):
        """io.py:45-47."""
        if self._fid is not None:
            self._h5.close(self._fid)
            self._fid = None

    def _scan(self, fid):
        """What the run's file holds: its variables (sorted), the group of
        step datasets ('2D' or '3D'), its dimension and the steps in
        numeric order."""
        h5 = self._h5
        names = h5.members(fid)
        if not names:
            raise ValueError(f"{self.path}: no variables to reconfigure")
        kids = [k for k in h5.members(fid, names[0]) if k != "mesh"]
        key = kids[0] if kids else ""
        dim = {"2d": 2, "3d": 3}.get(key.lower())
        if dim is None:
            raise ValueError(f"{self.path}: '{names[0]}' holds no 2D or 3D step group (found {kids})")
        steps = sorted((int(k) for k in h5.members(fid, f"{names[0]}/{key}")))
        return names, key, dim, steps

    def reconfigure_dataset(self):
        """io.py:142-196: rewrite the run's file into the flat layout.

        One pass from the run's file into ``<stem>_temp.h5`` beside it, which
        then replaces it: per axis i the coordinates as root datasets 'x',
        'y'[, 'z'] (with the end point ``end_points[i]`` appended when
        ``periodic``), 't' = linspace(first, last time point, number of
        snapshots), and per variable one float64 dataset per snapshot named
        by its step padded with zeros to the width of the last step (with the
        first plane repeated at the end of every axis when ``periodic``)."""
        self.close()
        src_path = self.path
        if not src_path.exists():
            raise FileNotFoundError(f"Solution file '{src_path}' does not exist.")
        tmp_path = src_path.with_name(src_path.stem + "_temp.h5")
        logger.info(f"reconfiguring {src_path} into the flat snapshot layout")
        h5 = self._h5
        src = h5.open(src_path)
        try:
            names, key, dim, steps = self._scan(src)
            width = len(str(steps[-1])) if steps else 1
            dst = h5.create(tmp_path)
            try:
                for i, label in enumerate("xyz"[:dim]):
                    c = h5.read(src, f"{names[0]}/mesh/x{i}")
                    if self.periodic:
                        c = np.append(c, self.end_points[i])
                    h5.create_dataset(dst, label, c.shape, c)
                t = np.linspace(self.time_points[0], self.time_points[-1], len(steps))
                h5.create_dataset(dst, "t", t.shape, t)
                for name in names:
                    h5.require_group(dst, name)
                    for st in steps:
                        a = h5.read(src, f"{name}/{key}/{st}")
                        if self.periodic:
                            a = periodic_bc(a)
                        h5.create_dataset(dst, f"{name}/{st:0{width}d}", a.shape, a)
            finally:
                h5.close(dst)
        finally:
            h5.close(src)
        src_path.unlink()
        tmp_path.rename(src_path)
        logger.info(f"{src_path}: {len(names)} variable(s), {len(steps)} snapshot(s) written in the flat layout")
