"""HDF5 files through the HDF5 C library (libhdf5 1.10 or later, loaded with
ctypes), for the persistence layer (`luminis_mc.manager`).

A small h5py-like surface -- `File`, `Group`, `Dataset`, an `attrs` mapping --
that writes objects the way h5py writes them, so files made here are read by
h5py-based code and the other way round:

* float and integer arrays as native little-endian types, complex128 as the
  compound {r: f8, i: f8}, numpy bool as an int8 enum {FALSE: 0, TRUE: 1},
  `str` as variable-length UTF-8 strings;
* datasets above 256 elements chunked, shuffled and gzip-compressed when the
  caller asks for it (the reference's `_write_dataset` does for every array
  above 256 elements).

The library is found from $LMC_HDF5_LIB, the loader's search path
(`ctypes.util.find_library("hdf5")`) or the usual install prefixes; without it
`library()` raises RuntimeError naming what it looked for.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import glob
import os
import threading

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
htri_t = ctypes.c_int
hsize_t = ctypes.c_uint64

H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0x0000, 0x0001, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5T_CSET_UTF8 = 1
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_COMPOUND, H5T_ENUM = 0, 1, 3, 6, 8
H5T_SGN_NONE = 0
H5T_DIR_DEFAULT = 0
H5I_GROUP, H5I_DATASET = 2, 5
H5_INDEX_NAME, H5_ITER_INC = 0, 0
H5F_SCOPE_LOCAL = 0
H5F_CLOSE_STRONG = 3

_lock = threading.Lock()
_lib = None
_types: dict = {}

_CANDIDATES = ("/opt/conda/lib/libhdf5.so*", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*",
               "/usr/lib/x86_64-linux-gnu/libhdf5.so*", "/usr/local/lib/libhdf5.so*", "/usr/lib64/libhdf5.so*")


def _find() -> list[str]:
    names = []
    if os.environ.get("LMC_HDF5_LIB"):
        names.append(os.environ["LMC_HDF5_LIB"])
    found = ctypes.util.find_library("hdf5")
    if found:
        names.append(found)
    for pat in _CANDIDATES:
        names += sorted(p for p in glob.glob(pat) if "_hl" not in p and "_cpp" not in p and "fortran" not in p)
    return names


def _proto(L) -> None:
    P = ctypes.POINTER
    sigs = {
        "H5open": (herr_t, []), "H5Eset_auto2": (herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fclose": (herr_t, [hid_t]), "H5Fflush": (herr_t, [hid_t, ctypes.c_int]),
        "H5Gcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Gclose": (herr_t, [hid_t]),
        "H5Lexists": (htri_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Ldelete": (herr_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Literate": (herr_t, [hid_t, ctypes.c_int, ctypes.c_int, P(hsize_t), ctypes.c_void_p, ctypes.c_void_p]),
        "H5Oopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Oclose": (herr_t, [hid_t]),
        "H5Iget_type": (ctypes.c_int, [hid_t]),
        "H5Screate": (hid_t, [ctypes.c_int]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, P(hsize_t), P(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]), "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, P(hsize_t), P(hsize_t)]),
        "H5Tcopy": (hid_t, [hid_t]), "H5Tclose": (herr_t, [hid_t]),
        "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]), "H5Tget_size": (ctypes.c_size_t, [hid_t]),
        "H5Tset_cset": (herr_t, [hid_t, ctypes.c_int]), "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tget_sign": (ctypes.c_int, [hid_t]), "H5Tis_variable_str": (htri_t, [hid_t]),
        "H5Tcreate": (hid_t, [ctypes.c_int, ctypes.c_size_t]),
        "H5Tinsert": (herr_t, [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Tget_nmembers": (ctypes.c_int, [hid_t]),
        "H5Tenum_create": (hid_t, [hid_t]), "H5Tenum_insert": (herr_t, [hid_t, ctypes.c_char_p, ctypes.c_void_p]),
        "H5Tget_super": (hid_t, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]), "H5Pclose": (herr_t, [hid_t]),
        "H5Pset_fclose_degree": (herr_t, [hid_t, ctypes.c_int]),
        "H5Pset_chunk": (herr_t, [hid_t, ctypes.c_int, P(hsize_t)]),
        "H5Pset_deflate": (herr_t, [hid_t, ctypes.c_uint]), "H5Pset_shuffle": (herr_t, [hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Dclose": (herr_t, [hid_t]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dget_space": (hid_t, [hid_t]), "H5Dget_type": (hid_t, [hid_t]),
        "H5Dvlen_reclaim": (herr_t, [hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Acreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        "H5Aopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]), "H5Aclose": (herr_t, [hid_t]),
        "H5Awrite": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aread": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aexists": (htri_t, [hid_t, ctypes.c_char_p]), "H5Adelete": (herr_t, [hid_t, ctypes.c_char_p]),
        "H5Aget_space": (hid_t, [hid_t]), "H5Aget_type": (hid_t, [hid_t]),
        "H5Aiterate2": (herr_t, [hid_t, ctypes.c_int, ctypes.c_int, P(hsize_t), ctypes.c_void_p, ctypes.c_void_p]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args


def library():
    """The loaded libhdf5 (loaded once per process), with its error printing off."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        tried = _find()
        for path in tried:
            try:
                L = ctypes.CDLL(path)
                getattr(L, "H5Literate")
            except (OSError, AttributeError):
                continue
            _proto(L)
            if L.H5open() < 0:
                continue
            L.H5Eset_auto2(0, None, None)
            for n in ("DOUBLE", "FLOAT", "SCHAR", "SHORT", "INT", "LLONG", "UCHAR", "USHORT", "UINT", "ULLONG"):
                _types[n] = hid_t.in_dll(L, f"H5T_NATIVE_{n}_g").value
            _types["C_S1"] = hid_t.in_dll(L, "H5T_C_S1_g").value
            _types["DCPL"] = hid_t.in_dll(L, "H5P_CLS_DATASET_CREATE_ID_g").value
            _types["FAPL"] = hid_t.in_dll(L, "H5P_CLS_FILE_ACCESS_ID_g").value
            _lib = L
            return L
        raise RuntimeError("HDF5 persistence needs the HDF5 C library (libhdf5 >= 1.10); looked in: "
                           + (", ".join(tried) if tried else "$LMC_HDF5_LIB, ldconfig, " + ", ".join(_CANDIDATES)))


def _ok(v, what: str):
    if v < 0:
        raise RuntimeError(f"HDF5 call failed: {what}")
    return v


_INT = {(1, True): "SCHAR", (2, True): "SHORT", (4, True): "INT", (8, True): "LLONG",
        (1, False): "UCHAR", (2, False): "USHORT", (4, False): "UINT", (8, False): "ULLONG"}


class _Type:
    """An HDF5 datatype id for a numpy dtype (closed on exit when it is a copy)."""

    def __init__(self, dt: np.dtype):
        L, self.own = library(), True
        dt = np.dtype(dt)
        if dt == np.bool_:
            self.id = _ok(L.H5Tenum_create(_types["SCHAR"]), "H5Tenum_create")
            for name, val in ((b"FALSE", 0), (b"TRUE", 1)):
                v = ctypes.c_int8(val)
                _ok(L.H5Tenum_insert(self.id, name, ctypes.byref(v)), "H5Tenum_insert")
        elif dt.kind == "c":
            half = _types["DOUBLE"] if dt.itemsize == 16 else _types["FLOAT"]
            self.id = _ok(L.H5Tcreate(H5T_COMPOUND, dt.itemsize), "H5Tcreate")
            _ok(L.H5Tinsert(self.id, b"r", 0, half), "H5Tinsert")
            _ok(L.H5Tinsert(self.id, b"i", dt.itemsize // 2, half), "H5Tinsert")
        elif dt.kind == "f":
            self.id, self.own = _types["DOUBLE" if dt.itemsize == 8 else "FLOAT"], False
        elif dt.kind in "iu":
            self.id, self.own = _types[_INT[(dt.itemsize, dt.kind == "i")]], False
        else:
            raise TypeError(f"no HDF5 type for numpy dtype {dt}")

    def __enter__(self):
        return self.id

    def __exit__(self, *a):
        if self.own:
            library().H5Tclose(self.id)


def _vlen_str_type():
    L = library()
    t = _ok(L.H5Tcopy(_types["C_S1"]), "H5Tcopy")
    L.H5Tset_size(t, H5T_VARIABLE)
    L.H5Tset_cset(t, H5T_CSET_UTF8)
    return t


def _space(shape):
    L = library()
    if len(shape) == 0:
        return _ok(L.H5Screate(H5S_SCALAR), "H5Screate")
    dims = (hsize_t * len(shape))(*shape)
    return _ok(L.H5Screate_simple(len(shape), dims, None), "H5Screate_simple")


def _shape_of(space) -> tuple:
    L = library()
    n = L.H5Sget_simple_extent_ndims(space)
    if n <= 0:
        return ()
    dims = (hsize_t * n)()
    L.H5Sget_simple_extent_dims(space, dims, None)
    return tuple(int(d) for d in dims)


def _read_obj(oid, is_attr: bool):
    """Value of an attribute or dataset as numpy (strings as str)."""
    L = library()
    space = L.H5Aget_space(oid) if is_attr else L.H5Dget_space(oid)
    ftype = L.H5Aget_type(oid) if is_attr else L.H5Dget_type(oid)
    try:
        shape = _shape_of(space)
        cls, size = L.H5Tget_class(ftype), L.H5Tget_size(ftype)
        read = (lambda mt, buf: L.H5Aread(oid, mt, buf)) if is_attr else \
            (lambda mt, buf: L.H5Dread(oid, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf))
        n = int(np.prod(shape)) if shape else 1
        if cls == H5T_STRING:
            if L.H5Tis_variable_str(ftype) > 0:
                mt = _vlen_str_type()
                buf = (ctypes.c_char_p * n)()
                try:
                    _ok(read(mt, buf), "read string")
                    vals = [(b or b"").decode("utf-8", errors="replace") for b in buf]
                    L.H5Dvlen_reclaim(mt, space, H5P_DEFAULT, buf)
                finally:
                    L.H5Tclose(mt)
            else:
                mt = _ok(L.H5Tcopy(_types["C_S1"]), "H5Tcopy")
                L.H5Tset_size(mt, size)
                raw = ctypes.create_string_buffer(size * n)
                try:
                    _ok(read(mt, raw), "read string")
                finally:
                    L.H5Tclose(mt)
                vals = [raw.raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode("utf-8", errors="replace")
                        for i in range(n)]
            return vals[0] if not shape else np.array(vals, dtype=object).reshape(shape)
        if cls == H5T_COMPOUND and L.H5Tget_nmembers(ftype) == 2 and size in (8, 16):
            dt = np.dtype(np.complex128 if size == 16 else np.complex64)
        elif cls == H5T_ENUM and size == 1:
            dt = np.dtype(np.bool_)
        elif cls == H5T_FLOAT:
            dt = np.dtype(np.float64 if size == 8 else np.float32)
        elif cls == H5T_INTEGER:
            dt = np.dtype(("i" if L.H5Tget_sign(ftype) != H5T_SGN_NONE else "u") + str(size))
        else:
            raise TypeError(f"unsupported HDF5 type class {cls} (size {size})")
        out = np.empty(shape, dtype=dt)
        with _Type(np.dtype(np.int8) if dt == np.bool_ else dt) as mt:
            if dt == np.bool_:
                tmp = np.empty(shape, dtype=np.int8)
                _ok(read(mt, tmp.ctypes.data_as(ctypes.c_void_p)), "read")
                out = tmp.astype(bool)
            else:
                _ok(read(mt, out.ctypes.data_as(ctypes.c_void_p)), "read")
        return out[()] if not shape else out  # a scalar dataspace reads as a numpy scalar
    finally:
        L.H5Tclose(ftype)
        L.H5Sclose(space)


def _prepare(val):
    """(numpy array, is_str) of a value to store."""
    if isinstance(val, str):
        return np.array(val, dtype=object), True
    a = np.asarray(val)
    if a.dtype.kind in "USO":
        return np.array(a.astype(str).tolist() if a.ndim else str(a), dtype=object), True
    return np.array(a, order="C", copy=True), False  # keeps a 0-d value 0-d (a scalar dataspace)


def _write_obj(write, arr, is_str):
    L = library()
    if is_str:
        flat = [s.encode("utf-8") for s in np.asarray(arr, dtype=object).reshape(-1).tolist()]
        buf = (ctypes.c_char_p * max(1, len(flat)))(*flat)
        mt = _vlen_str_type()
        try:
            _ok(write(mt, buf), "write string")
        finally:
            L.H5Tclose(mt)
        return
    if arr.dtype == np.bool_:
        tmp = arr.astype(np.int8)
        with _Type(np.bool_) as mt:
            _ok(write(mt, tmp.ctypes.data_as(ctypes.c_void_p)), "write")
        return
    with _Type(arr.dtype) as mt:
        _ok(write(mt, arr.ctypes.data_as(ctypes.c_void_p)), "write")


class Attrs:
    """Attributes of a group or dataset, as a mutable mapping (it keeps its
    object, and so the object's open id, alive)."""

    def __init__(self, owner):
        self._owner = owner

    @property
    def _id(self) -> int:
        return self._owner.id

    def __setitem__(self, key: str, val) -> None:
        L = library()
        k = key.encode()
        if L.H5Aexists(self._id, k) > 0:
            _ok(L.H5Adelete(self._id, k), f"H5Adelete {key}")
        arr, is_str = _prepare(val)
        space = _space(arr.shape)
        try:
            ft = _vlen_str_type() if is_str else None
            with (_Type(arr.dtype) if not is_str else _Holder(ft)) as t:
                aid = _ok(L.H5Acreate2(self._id, k, t, space, H5P_DEFAULT, H5P_DEFAULT), f"H5Acreate2 {key}")
            try:
                _write_obj(lambda mt, buf: L.H5Awrite(aid, mt, buf), arr, is_str)
            finally:
                L.H5Aclose(aid)
        finally:
            L.H5Sclose(space)

    def __getitem__(self, key: str):
        L = library()
        aid = L.H5Aopen(self._id, key.encode(), H5P_DEFAULT)
        if aid < 0:
            raise KeyError(key)
        try:
            return _read_obj(aid, True)
        finally:
            L.H5Aclose(aid)

    def __contains__(self, key: str) -> bool:
        return library().H5Aexists(self._id, key.encode()) > 0

    def __delitem__(self, key: str) -> None:
        if _ok(library().H5Adelete(self._id, key.encode()), f"H5Adelete {key}") < 0:
            raise KeyError(key)

    def keys(self) -> list[str]:
        names = []
        cb_t = ctypes.CFUNCTYPE(herr_t, hid_t, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)
        cb = cb_t(lambda loc, name, info, data: (names.append(name.decode()), 0)[1])
        idx = hsize_t(0)
        _ok(library().H5Aiterate2(self._id, H5_INDEX_NAME, H5_ITER_INC, ctypes.byref(idx), cb, None), "H5Aiterate2")
        return names

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def __iter__(self):
        return iter(self.keys())

    def __len__(self) -> int:
        return len(self.keys())

    def get(self, key, default=None):
        return self[key] if key in self else default


class _Holder:
    def __init__(self, tid):
        self.tid = tid

    def __enter__(self):
        return self.tid

    def __exit__(self, *a):
        library().H5Tclose(self.tid)


def _chunks(shape, itemsize: int, target: int = 1 << 18) -> tuple:
    """Chunk shape for a compressed dataset: the whole array, halved along its
    largest axis until a chunk holds at most `target` bytes."""
    c = [max(1, int(s)) for s in shape]
    while int(np.prod(c)) * itemsize > target and max(c) > 1:
        i = int(np.argmax(c))
        c[i] = (c[i] + 1) // 2
    return tuple(c)


class Dataset:
    def __init__(self, did: int, name: str):
        self.id, self.name = did, name
        self.attrs = Attrs(self)

    @property
    def shape(self) -> tuple:
        L = library()
        sp = L.H5Dget_space(self.id)
        try:
            return _shape_of(sp)
        finally:
            L.H5Sclose(sp)

    def read(self):
        return _read_obj(self.id, False)

    def __getitem__(self, key):
        v = self.read()
        return v if key == () or key is Ellipsis else v[key]

    def __array__(self, dtype=None, copy=None):
        v = np.asarray(self.read())
        return v.astype(dtype) if dtype is not None else v

    def close(self) -> None:
        if self.id >= 0:
            library().H5Dclose(self.id)
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    def __init__(self, gid: int, name: str, owner=None):
        self.id, self.name = gid, name
        self._owner = owner  # keeps the file open while a group is in use
        self.attrs = Attrs(self)

    def _child_type(self, name: str) -> int:
        L = library()
        oid = L.H5Oopen(self.id, name.encode(), H5P_DEFAULT)
        if oid < 0:
            raise KeyError(name)
        try:
            return L.H5Iget_type(oid)
        finally:
            L.H5Oclose(oid)

    def __contains__(self, path: str) -> bool:
        L = library()
        parts = [p for p in path.split("/") if p]
        cur = ""
        for p in parts:  # H5Lexists needs each intermediate link to exist
            cur = f"{cur}/{p}" if cur else p
            if L.H5Lexists(self.id, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return bool(parts)

    def __getitem__(self, path: str):
        if path not in self:
            raise KeyError(path)
        L, p = library(), path.encode()
        if self._child_type(path) == H5I_GROUP:
            return Group(_ok(L.H5Gopen2(self.id, p, H5P_DEFAULT), f"H5Gopen2 {path}"), path, self._owner or self)
        return Dataset(_ok(L.H5Dopen2(self.id, p, H5P_DEFAULT), f"H5Dopen2 {path}"), path)

    def get(self, path: str, default=None):
        return self[path] if path in self else default

    def __delitem__(self, name: str) -> None:
        _ok(library().H5Ldelete(self.id, name.encode(), H5P_DEFAULT), f"H5Ldelete {name}")

    def keys(self) -> list[str]:
        names = []
        cb_t = ctypes.CFUNCTYPE(herr_t, hid_t, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)
        cb = cb_t(lambda g, name, info, data: (names.append(name.decode()), 0)[1])
        idx = hsize_t(0)
        _ok(library().H5Literate(self.id, H5_INDEX_NAME, H5_ITER_INC, ctypes.byref(idx), cb, None), "H5Literate")
        return names

    def __iter__(self):
        return iter(self.keys())

    def is_group(self, name: str) -> bool:
        return self._child_type(name) == H5I_GROUP

    def require_group(self, name: str) -> "Group":
        if name in self:
            g = self[name]
            if not isinstance(g, Group):
                raise TypeError(f"'{name}' exists and is not a group")
            return g
        L = library()
        return Group(_ok(L.H5Gcreate2(self.id, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                         f"H5Gcreate2 {name}"), name, self._owner or self)

    def create_dataset(self, name: str, data, compression=None, compression_opts=None, shuffle=False) -> Dataset:
        """A new dataset holding `data`; compression "gzip" (level
        compression_opts) chunks the array and filters it (shuffle first)."""
        L = library()
        arr, is_str = _prepare(data)
        space = _space(arr.shape)
        dcpl = H5P_DEFAULT
        try:
            if compression and arr.ndim > 0 and arr.size > 0 and not is_str:
                if compression != "gzip":
                    raise ValueError(f"unsupported compression '{compression}'")
                dcpl = _ok(L.H5Pcreate(_types["DCPL"]), "H5Pcreate")
                ch = _chunks(arr.shape, arr.dtype.itemsize)
                _ok(L.H5Pset_chunk(dcpl, len(ch), (hsize_t * len(ch))(*ch)), "H5Pset_chunk")
                if shuffle:
                    _ok(L.H5Pset_shuffle(dcpl), "H5Pset_shuffle")
                _ok(L.H5Pset_deflate(dcpl, int(4 if compression_opts is None else compression_opts)), "H5Pset_deflate")
            ft = _vlen_str_type() if is_str else None
            with (_Type(arr.dtype) if not is_str else _Holder(ft)) as t:
                did = _ok(L.H5Dcreate2(self.id, name.encode(), t, space, H5P_DEFAULT, dcpl, H5P_DEFAULT),
                          f"H5Dcreate2 {name}")
            ds = Dataset(did, name)
            if arr.size > 0 or is_str:
                _write_obj(lambda mt, buf: L.H5Dwrite(did, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), arr, is_str)
            return ds
        finally:
            if dcpl != H5P_DEFAULT:
                L.H5Pclose(dcpl)
            L.H5Sclose(space)

    def close(self) -> None:
        if self.id >= 0 and not isinstance(self, File):
            library().H5Gclose(self.id)
            self.id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class File(Group):
    """An HDF5 file: mode "w" (create, truncate), "r" (read) or "a"/"r+" (read-write)."""

    def __init__(self, path, mode: str = "r"):
        L, p = library(), os.fspath(path).encode()
        if mode not in ("w", "r", "a", "r+"):
            raise ValueError(f"unknown mode '{mode}'")
        # closing the file closes every group and dataset still open in it
        # (the library's default keeps the file open until the last one closes)
        fapl = _ok(L.H5Pcreate(_types["FAPL"]), "H5Pcreate")
        try:
            _ok(L.H5Pset_fclose_degree(fapl, H5F_CLOSE_STRONG), "H5Pset_fclose_degree")
            if mode == "w" or (mode == "a" and not os.path.exists(path)):
                fid = L.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, fapl)
            else:
                fid = L.H5Fopen(p, H5F_ACC_RDONLY if mode == "r" else H5F_ACC_RDWR, fapl)
        finally:
            L.H5Pclose(fapl)
        if fid < 0:
            raise OSError(f"cannot open HDF5 file {os.fspath(path)} (mode '{mode}')")
        self.fid, self.filename = fid, os.fspath(path)
        # the file's root group: groups and attributes under "/"
        Group.__init__(self, _ok(L.H5Gopen2(fid, b"/", H5P_DEFAULT), "H5Gopen2 /"), "/", None)

    def flush(self) -> None:
        if self.fid >= 0:
            library().H5Fflush(self.fid, H5F_SCOPE_LOCAL)

    def close(self) -> None:
        L = library()
        if self.id >= 0:
            L.H5Gclose(self.id)
            self.id = -1
        if self.fid >= 0:
            L.H5Fclose(self.fid)
            self.fid = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
