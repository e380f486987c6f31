"""Spectrogram ingest from filterbank HDF5 files.

The reference's integration test reads its input this way: a Breakthrough
Listen filterbank HDF5 file whose dataset ``data`` holds Float32 powers with
HDF5 dimensions (time, IF, channel) in C order, which Julia's HDF5.jl sees as
``data[chan, if, time]``; the test takes ``h5["data"][freq_range, 1, :]``, an
Nf x Nt spectrogram (test/heavytests.jl:10-14).  Header values (fch1, foff,
tsamp, ...) are attributes of the dataset (or of the file).

This module reads such files through the HDF5 C library (libhdf5, loaded with
ctypes; ``ZDT_HDF5_LIB`` names it when it is not on the loader's path).  A
hyperslab of channels x times of one IF is read straight into a column-major
Nf x Nt buffer -- a numpy array, or a (pinned) torch tensor such as
``fdrt_hip.pinned_spectrogram`` returns, which ``StreamSearch`` copies to the
GPU asynchronously.  Compressed datasets go through HDF5's own filter
pipeline: deflate and shuffle are built into libhdf5, and bitshuffle (filter
32008, what rawspec writes) is used when its plugin is on
``HDF5_PLUGIN_PATH``.  Without the plugin, a dataset whose only filter is
bitshuffle (plain or with LZ4, the Breakthrough Listen files' compression) is
read chunk by chunk with HDF5's direct chunk reads and decoded by the
library's own decoder (``zdt_bshuf_decompress``).  A dataset with any other
filter libhdf5 cannot run is refused before any read, naming the filter.

Indices are Python's: channel and time ranges are 0-based, half-open, and the
IF index is 0-based (the reference's ``data[freq_range, 1, :]`` is
``read(channels=range(freq_range.start - 1, ...), ifnum=0)``).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
from pathlib import Path

import numpy as np

__all__ = ["FilterbankError", "FilterbankFile", "read_spectrogram", "write_filterbank"]

_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_ulonglong

# header attributes of the filterbank format, read when present
HEADER_KEYS = ("fch1", "foff", "tsamp", "tstart", "nchans", "nifs", "nbits", "data_type", "machine_id",
               "telescope_id", "src_raj", "src_dej", "az_start", "za_start", "ibeam", "nbeams", "source_name",
               "rawdatafile")

_H5F_ACC_RDONLY = 0
_H5F_ACC_TRUNC = 2
_H5S_SELECT_SET = 0
_H5T_INTEGER, _H5T_FLOAT, _H5T_STRING = 0, 1, 3
_FILTER_NAMES = {1: "deflate", 2: "shuffle", 3: "fletcher32", 4: "szip", 5: "nbit", 6: "scaleoffset",
                 32001: "blosc", 32004: "lz4", 32008: "bitshuffle"}


class FilterbankError(RuntimeError):
    pass


class _Lib:
    """The subset of the HDF5 1.10+ C API this module calls."""

    def __init__(self):
        cands = [os.environ.get("ZDT_HDF5_LIB"), ctypes.util.find_library("hdf5"),
                 "/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
                 "/usr/lib/x86_64-linux-gnu/libhdf5.so", "libhdf5.so"]
        self.L = None
        errs = []
        for c in cands:
            if not c:
                continue
            try:
                self.L = C.CDLL(c)
                self.path = c
                break
            except OSError as e:
                errs.append(f"{c}: {e}")
        if self.L is None:
            raise FilterbankError("libhdf5 not found (set ZDT_HDF5_LIB to its path): " + "; ".join(errs))
        L = self.L
        sig = {
            "H5open": (_herr, []),
            "H5get_libversion": (_herr, [C.POINTER(C.c_uint)] * 3),
            "H5Eset_auto2": (_herr, [_hid, C.c_void_p, C.c_void_p]),
            "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
            "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
            "H5Fclose": (_herr, [_hid]),
            "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dclose": (_herr, [_hid]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Dget_create_plist": (_hid, [_hid]),
            "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sget_simple_extent_npoints": (C.c_int64, [_hid]),
            "H5Sselect_hyperslab": (_herr, [_hid, C.c_int, C.POINTER(_hsize), C.POINTER(_hsize),
                                            C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Screate": (_hid, [C.c_int]),
            "H5Sclose": (_herr, [_hid]),
            "H5Pcreate": (_hid, [_hid]),
            "H5Pclose": (_herr, [_hid]),
            "H5Pset_chunk": (_herr, [_hid, C.c_int, C.POINTER(_hsize)]),
            "H5Pset_deflate": (_herr, [_hid, C.c_uint]),
            "H5Pset_shuffle": (_herr, [_hid]),
            "H5Pset_filter": (_herr, [_hid, C.c_int, C.c_uint, C.c_size_t, C.POINTER(C.c_uint)]),
            "H5Pget_nfilters": (C.c_int, [_hid]),
            "H5Pget_filter2": (C.c_int, [_hid, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_uint), C.c_size_t, C.c_char_p, C.POINTER(C.c_uint)]),
            "H5Zfilter_avail": (C.c_int, [C.c_int]),
            "H5Pget_chunk": (C.c_int, [_hid, C.c_int, C.POINTER(_hsize)]),
            "H5Pget_filter_by_id2": (_herr, [_hid, C.c_int, C.POINTER(C.c_uint), C.POINTER(C.c_size_t),
                                             C.POINTER(C.c_uint), C.c_size_t, C.c_char_p, C.POINTER(C.c_uint)]),
            "H5Dget_type": (_hid, [_hid]),
            "H5Tget_order": (C.c_int, [_hid]),
            "H5Dget_chunk_storage_size": (_herr, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Dread_chunk": (_herr, [_hid, _hid, C.POINTER(_hsize), C.POINTER(C.c_uint32), C.c_void_p]),
            "H5Dwrite_chunk": (_herr, [_hid, _hid, C.c_uint32, C.POINTER(_hsize), C.c_size_t, C.c_void_p]),
            "H5Aexists": (C.c_int, [_hid, C.c_char_p]),
            "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
            "H5Aclose": (_herr, [_hid]),
            "H5Aget_type": (_hid, [_hid]),
            "H5Aget_space": (_hid, [_hid]),
            "H5Aread": (_herr, [_hid, _hid, C.c_void_p]),
            "H5Awrite": (_herr, [_hid, _hid, C.c_void_p]),
            "H5Tget_class": (C.c_int, [_hid]),
            "H5Tget_size": (C.c_size_t, [_hid]),
            "H5Tis_variable_str": (C.c_int, [_hid]),
            "H5Tcopy": (_hid, [_hid]),
            "H5Tset_size": (_herr, [_hid, C.c_size_t]),
            "H5Tclose": (_herr, [_hid]),
            "H5free_memory": (_herr, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.H5open() < 0:
            raise FilterbankError("H5open failed")
        L.H5Eset_auto2(0, None, None)  # errors come back as return codes, not printed stacks
        v = [C.c_uint() for _ in range(3)]
        L.H5get_libversion(*[C.byref(x) for x in v])
        self.version = tuple(x.value for x in v)
        if self.version < (1, 10, 0):
            raise FilterbankError(f"libhdf5 {self.version} is older than 1.10")

        def glob(name):
            return _hid.in_dll(L, name).value

        self.NATIVE_FLOAT = glob("H5T_NATIVE_FLOAT_g")
        self.NATIVE_DOUBLE = glob("H5T_NATIVE_DOUBLE_g")
        self.NATIVE_LLONG = glob("H5T_NATIVE_LLONG_g")
        self.IEEE_F32LE = glob("H5T_IEEE_F32LE_g")
        self.C_S1 = glob("H5T_C_S1_g")
        self.DCPL = glob("H5P_CLS_DATASET_CREATE_ID_g")

    def ok(self, rc, what):
        if rc < 0:
            raise FilterbankError(f"HDF5: {what} failed")
        return rc


_LIB = None


def _lib() -> _Lib:
    global _LIB
    if _LIB is None:
        _LIB = _Lib()
    return _LIB


def _dims(h, space) -> tuple:
    n = h.L.H5Sget_simple_extent_ndims(space)
    h.ok(n, "H5Sget_simple_extent_ndims")
    d = (_hsize * max(n, 1))()
    h.ok(h.L.H5Sget_simple_extent_dims(space, d, None), "H5Sget_simple_extent_dims")
    return tuple(int(d[i]) for i in range(n))


def _read_attr(h, obj, name: str):
    attr = h.ok(h.L.H5Aopen(obj, name.encode(), 0), f"H5Aopen({name})")
    try:
        t = h.ok(h.L.H5Aget_type(attr), "H5Aget_type")
        sp = h.ok(h.L.H5Aget_space(attr), "H5Aget_space")
        try:
            cls = h.L.H5Tget_class(t)
            n = max(int(h.L.H5Sget_simple_extent_npoints(sp)), 1)
            if cls == _H5T_STRING:
                if h.L.H5Tis_variable_str(t) > 0:
                    ptrs = (C.c_char_p * n)()
                    h.ok(h.L.H5Aread(attr, t, ptrs), f"H5Aread({name})")
                    vals = [p.decode(errors="replace") if p else "" for p in ptrs]
                    for i in range(n):  # the library allocated the strings
                        h.L.H5free_memory(C.cast(ptrs, C.POINTER(C.c_void_p))[i])
                else:
                    size = int(h.L.H5Tget_size(t))
                    buf = C.create_string_buffer(size * n)
                    h.ok(h.L.H5Aread(attr, t, buf), f"H5Aread({name})")
                    raw = buf.raw
                    vals = [raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode(errors="replace")
                            for i in range(n)]
                return vals[0] if n == 1 else vals
            if cls == _H5T_INTEGER:
                arr = np.zeros(n, np.int64)
                h.ok(h.L.H5Aread(attr, h.NATIVE_LLONG, arr.ctypes.data_as(C.c_void_p)), f"H5Aread({name})")
                return int(arr[0]) if n == 1 else arr.tolist()
            if cls == _H5T_FLOAT:
                arr = np.zeros(n, np.float64)
                h.ok(h.L.H5Aread(attr, h.NATIVE_DOUBLE, arr.ctypes.data_as(C.c_void_p)), f"H5Aread({name})")
                return float(arr[0]) if n == 1 else arr.tolist()
            return None
        finally:
            h.L.H5Sclose(sp)
            h.L.H5Tclose(t)
    finally:
        h.L.H5Aclose(attr)


def _span(r, n: int, what: str) -> tuple:
    """(start, count) of a 0-based half-open range (None: all) within [0, n)."""
    if r is None:
        return 0, n
    if isinstance(r, slice):
        if r.step not in (None, 1):
            raise ValueError(f"{what}: only unit-step ranges")
        r = range(*r.indices(n))
    if not isinstance(r, range) or r.step != 1:
        raise ValueError(f"{what}: a contiguous range is required")
    if len(r) == 0 or r.start < 0 or r.stop > n:
        raise ValueError(f"{what}: {r} outside [0, {n})")
    return r.start, len(r)


class FilterbankFile:
    """An open filterbank HDF5 file (read-only).

    ``shape`` is (nchans, nifs, ntime) -- the reference's data[chan, if, time]
    order; ``header`` holds the filterbank attributes found on the dataset or
    the file; ``filters`` lists the dataset's filter pipeline as (id, name)."""

    def __init__(self, path, dataset: str = "data"):
        h = _lib()
        self._h = h
        self.path = str(path)
        self._f = h.L.H5Fopen(self.path.encode(), _H5F_ACC_RDONLY, 0)
        if self._f < 0:
            raise FilterbankError(f"cannot open {self.path} as HDF5")
        self._d = h.L.H5Dopen2(self._f, dataset.encode(), 0)
        if self._d < 0:
            h.L.H5Fclose(self._f)
            raise FilterbankError(f"{self.path}: no dataset {dataset!r}")
        try:
            sp = h.ok(h.L.H5Dget_space(self._d), "H5Dget_space")
            try:
                dims = _dims(h, sp)
            finally:
                h.L.H5Sclose(sp)
            if len(dims) != 3:
                raise FilterbankError(f"{self.path}: {dataset} has {len(dims)} dimensions, expected (time, if, chan)")
            self.ntime, self.nifs, self.nchans = dims
            self.shape = (self.nchans, self.nifs, self.ntime)
            self.filters = self._filters()
            self._bshuf = self._bitshuffle_params()
            self.header = {}
            for key in HEADER_KEYS:
                for obj in (self._d, self._f):
                    if h.L.H5Aexists(obj, key.encode()) > 0:
                        self.header[key] = _read_attr(h, obj, key)
                        break
        except Exception:
            self.close()
            raise

    def _filters(self) -> list:
        h = self._h
        pl = h.ok(h.L.H5Dget_create_plist(self._d), "H5Dget_create_plist")
        out = []
        try:
            for i in range(max(h.L.H5Pget_nfilters(pl), 0)):
                flags, nelm, cfg = C.c_uint(), C.c_size_t(0), C.c_uint()
                name = C.create_string_buffer(64)
                fid = h.L.H5Pget_filter2(pl, i, C.byref(flags), C.byref(nelm), None, 64, name, C.byref(cfg))
                h.ok(fid, "H5Pget_filter2")
                out.append((int(fid), name.value.decode(errors="replace") or _FILTER_NAMES.get(int(fid), "?")))
        finally:
            h.L.H5Pclose(pl)
        return out

    def _bitshuffle_params(self):
        """(element size, block size, compression, chunk dims) when the chunks
        are to be decoded here: bitshuffle is the only filter, libhdf5 has no
        plugin for it, the data are little-endian Float32 and the compression
        is none (0) or LZ4 (2).  None otherwise."""
        h = self._h
        if [f for f, _ in self.filters] != [32008] or h.L.H5Zfilter_avail(32008) > 0:
            return None
        pl = h.ok(h.L.H5Dget_create_plist(self._d), "H5Dget_create_plist")
        ty = h.ok(h.L.H5Dget_type(self._d), "H5Dget_type")
        try:
            flags, nelm, cfg = C.c_uint(), C.c_size_t(8), C.c_uint()
            cd = (C.c_uint * 8)()
            h.ok(h.L.H5Pget_filter_by_id2(pl, 32008, C.byref(flags), C.byref(nelm), cd, 0, None, C.byref(cfg)),
                 "H5Pget_filter_by_id2")
            cd = list(cd)[:nelm.value]
            dims = (_hsize * 3)()
            if h.L.H5Pget_chunk(pl, 3, dims) != 3:
                return None
            if h.L.H5Tget_class(ty) != _H5T_FLOAT or h.L.H5Tget_size(ty) != 4 or h.L.H5Tget_order(ty) != 0:
                return None
        finally:
            h.L.H5Tclose(ty)
            h.L.H5Pclose(pl)
        elem = cd[2] if len(cd) > 2 else 4
        block = cd[3] if len(cd) > 3 else 0
        comp = cd[4] if len(cd) > 4 else 0
        if elem != 4 or comp not in (0, 2):
            return None
        return elem, block, comp, tuple(int(x) for x in dims)

    def check_filters(self) -> None:
        """Raises FilterbankError when libhdf5 cannot run one of the dataset's
        filters and this module cannot decode it (bitshuffle combined with
        other filters, or with a compression other than LZ4, without its
        plugin, for one)."""
        if self._bshuf is not None:
            return
        for fid, name in self.filters:
            if self._h.L.H5Zfilter_avail(fid) <= 0:
                label = _FILTER_NAMES.get(fid, name)
                raise FilterbankError(f"{self.path}: HDF5 filter {fid} ({label}) is not available to libhdf5 "
                                      f"{'.'.join(map(str, self._h.version))}; put its plugin on HDF5_PLUGIN_PATH")

    def frequencies(self, channels=None) -> np.ndarray:
        """Channel centre frequencies fch1 + foff * i (header units) of a channel range."""
        if "fch1" not in self.header or "foff" not in self.header:
            raise FilterbankError(f"{self.path}: no fch1/foff attributes")
        c0, nc = _span(channels, self.nchans, "channels")
        return self.header["fch1"] + self.header["foff"] * np.arange(c0, c0 + nc, dtype=np.float64)

    def read(self, channels=None, ifnum: int = 0, times=None, out=None):
        """The spectrogram data[channels, ifnum, times] as an Nf x Nt Float32
        matrix, column-major (channels contiguous).  ``out``: a numpy array or
        torch tensor of that shape and layout (strides (1, Nf)) to fill,
        pinned host memory included; otherwise a new Fortran-ordered array."""
        h = self._h
        c0, nc = _span(channels, self.nchans, "channels")
        t0, nt = _span(times, self.ntime, "times")
        if not 0 <= ifnum < self.nifs:
            raise ValueError(f"ifnum {ifnum} outside [0, {self.nifs})")
        self.check_filters()
        if out is None:
            out = np.empty((nc, nt), np.float32, order="F")
        ptr = _column_major_ptr(out, nc, nt)
        if self._bshuf is not None:
            dst = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(nt, nc)).T  # (nc, nt) view
            self._read_chunks(c0, nc, ifnum, t0, nt, dst)
            return out
        fs = h.ok(h.L.H5Dget_space(self._d), "H5Dget_space")
        ms = -1
        try:
            start = (_hsize * 3)(t0, ifnum, c0)
            count = (_hsize * 3)(nt, 1, nc)
            h.ok(h.L.H5Sselect_hyperslab(fs, _H5S_SELECT_SET, start, None, count, None), "H5Sselect_hyperslab")
            # memory: nt rows of nc channels, C order == the Nf x Nt column-major matrix
            mdims = (_hsize * 2)(nt, nc)
            ms = h.ok(h.L.H5Screate_simple(2, mdims, None), "H5Screate_simple")
            h.ok(h.L.H5Dread(self._d, h.NATIVE_FLOAT, ms, fs, 0, ptr), "H5Dread")
        finally:
            if ms >= 0:
                h.L.H5Sclose(ms)
            h.L.H5Sclose(fs)
        return out

    def _read_chunks(self, c0: int, nc: int, ifnum: int, t0: int, nt: int, dst: np.ndarray) -> None:
        """The hyperslab by direct chunk reads, bitshuffle chunks decoded by
        zdt_bshuf_decompress; dst: the (nc, nt) destination view."""
        from . import _lib as Z
        h = self._h
        elem, block, comp, (ct, ci, cc) = self._bshuf
        L = Z.lib()
        raw = np.empty(ct * ci * cc, np.float32)
        ic = ifnum // ci
        for tc in range(t0 // ct, (t0 + nt - 1) // ct + 1):
            for ccx in range(c0 // cc, (c0 + nc - 1) // cc + 1):
                off = (_hsize * 3)(tc * ct, ic * ci, ccx * cc)
                size = _hsize(0)
                h.ok(h.L.H5Dget_chunk_storage_size(self._d, off, C.byref(size)), "H5Dget_chunk_storage_size")
                ta, tb = max(t0, tc * ct), min(t0 + nt, (tc + 1) * ct)
                ca, cb = max(c0, ccx * cc), min(c0 + nc, (ccx + 1) * cc)
                if size.value == 0:  # never written: the fill value
                    dst[ca - c0:cb - c0, ta - t0:tb - t0] = 0.0
                    continue
                buf = np.empty(size.value, np.uint8)
                mask = C.c_uint32(0)
                h.ok(h.L.H5Dread_chunk(self._d, 0, off, C.byref(mask), buf.ctypes.data_as(C.c_void_p)), "H5Dread_chunk")
                if mask.value & 1:  # the filter was skipped for this chunk: stored as is
                    if size.value != raw.nbytes:
                        raise FilterbankError(f"{self.path}: unfiltered chunk of {size.value} bytes, {raw.nbytes} expected")
                    raw[:] = buf.view(np.float32)
                else:
                    Z.check(L.zdt_bshuf_decompress(buf.ctypes.data_as(C.c_void_p), size.value,
                                                   raw.ctypes.data_as(C.c_void_p), raw.nbytes, elem, block, comp))
                chunk = raw.reshape(ct, ci, cc)
                dst[ca - c0:cb - c0, ta - t0:tb - t0] = \
                    chunk[ta - tc * ct:tb - tc * ct, ifnum - ic * ci, ca - ccx * cc:cb - ccx * cc].T

    def close(self) -> None:
        h = self._h
        if getattr(self, "_d", -1) >= 0:
            h.L.H5Dclose(self._d)
            self._d = -1
        if getattr(self, "_f", -1) >= 0:
            h.L.H5Fclose(self._f)
            self._f = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _column_major_ptr(out, nc: int, nt: int) -> C.c_void_p:
    """Data pointer of an Nf x Nt float32 column-major buffer (numpy or torch,
    host memory), after checking its shape and strides."""
    if isinstance(out, np.ndarray):
        if out.dtype != np.float32 or out.shape != (nc, nt) or (nt > 1 and out.strides != (4, 4 * nc)) \
                or (nt == 1 and out.strides[0] != 4):
            raise ValueError(f"out must be a float32 ({nc}, {nt}) column-major array")
        return C.c_void_p(out.ctypes.data)
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is not None and isinstance(out, torch.Tensor):
        if out.dtype != torch.float32 or tuple(out.shape) != (nc, nt) or out.is_cuda \
                or (nt > 1 and out.stride() != (1, nc)) or (nt == 1 and out.stride()[0] != 1):
            raise ValueError(f"out must be a host float32 ({nc}, {nt}) tensor with strides (1, {nc})")
        return C.c_void_p(out.data_ptr())
    raise TypeError("out must be a numpy array or a torch tensor")


def read_spectrogram(path, channels=None, ifnum: int = 0, times=None, out=None, dataset: str = "data"):
    """One-call form of FilterbankFile(path).read(...): the reference's
    ``h5["data"][freq_range, 1, :]`` as an Nf x Nt Float32 matrix."""
    with FilterbankFile(path, dataset) as f:
        return f.read(channels, ifnum, times, out)


def write_filterbank(path, data, header: dict | None = None, chunks=None, deflate: int | None = None,
                     shuffle: bool = False, extra_filter: tuple | None = None, dataset: str = "data") -> None:
    """Write an Nf x Nt (or Nf x nifs x Nt) Float32 spectrogram as a
    filterbank HDF5 file: dataset (time, if, chan), header values as dataset
    attributes, CLASS/VERSION on the file.  ``chunks``: the chunk shape in
    (time, if, chan) order; ``deflate``/``shuffle``: HDF5's built-in filters;
    ``extra_filter``: (filter id, optional) added to the pipeline as is."""
    h = _lib()
    a = np.asarray(data, dtype=np.float32)
    if a.ndim == 2:
        a = a[:, None, :]
    if a.ndim != 3:
        raise ValueError("data must be Nf x Nt or Nf x nifs x Nt")
    nchans, nifs, ntime = a.shape
    buf = np.ascontiguousarray(np.transpose(a, (2, 1, 0)))  # (time, if, chan), C order
    f = h.L.H5Fcreate(str(path).encode(), _H5F_ACC_TRUNC, 0, 0)
    if f < 0:
        raise FilterbankError(f"cannot create {path}")
    dims = (_hsize * 3)(ntime, nifs, nchans)
    sp = dcpl = d = -1
    try:
        sp = h.ok(h.L.H5Screate_simple(3, dims, None), "H5Screate_simple")
        dcpl = h.ok(h.L.H5Pcreate(h.DCPL), "H5Pcreate")
        if chunks is not None or deflate is not None or shuffle or extra_filter is not None:
            ch = chunks or (min(ntime, 16), 1, min(nchans, 1 << 16))
            h.ok(h.L.H5Pset_chunk(dcpl, 3, (_hsize * 3)(*ch)), "H5Pset_chunk")
        if shuffle:
            h.ok(h.L.H5Pset_shuffle(dcpl), "H5Pset_shuffle")
        if extra_filter is not None:
            fid, optional = extra_filter
            h.ok(h.L.H5Pset_filter(dcpl, int(fid), 1 if optional else 0, 0, None), "H5Pset_filter")
        if deflate is not None:
            h.ok(h.L.H5Pset_deflate(dcpl, int(deflate)), "H5Pset_deflate")
        d = h.ok(h.L.H5Dcreate2(f, dataset.encode(), h.IEEE_F32LE, sp, 0, dcpl, 0), "H5Dcreate2")
        h.ok(h.L.H5Dwrite(d, h.NATIVE_FLOAT, 0, 0, 0, buf.ctypes.data_as(C.c_void_p)), "H5Dwrite")
        hdr = {"nchans": nchans, "nifs": nifs, "nbits": 32}
        hdr.update(header or {})
        for k, v in hdr.items():
            _write_attr(h, d, k, v)
        _write_attr(h, f, "CLASS", "FILTERBANK")
        _write_attr(h, f, "VERSION", "1.0")
    finally:
        if d >= 0:
            h.L.H5Dclose(d)
        if dcpl >= 0:
            h.L.H5Pclose(dcpl)
        if sp >= 0:
            h.L.H5Sclose(sp)
        h.L.H5Fclose(f)


def _write_attr(h, obj, name: str, value) -> None:
    scalar = h.ok(h.L.H5Screate(0), "H5Screate")  # H5S_SCALAR
    t = -1
    try:
        if isinstance(value, str):
            raw = value.encode()
            t = h.ok(h.L.H5Tcopy(h.C_S1), "H5Tcopy")
            h.ok(h.L.H5Tset_size(t, max(len(raw), 1)), "H5Tset_size")
            ftype, mtype, buf = t, t, C.create_string_buffer(raw, max(len(raw), 1))
        elif isinstance(value, (int, np.integer)) and not isinstance(value, bool):
            ftype = mtype = h.NATIVE_LLONG
            buf = (C.c_longlong * 1)(int(value))
        else:
            ftype = mtype = h.NATIVE_DOUBLE
            buf = (C.c_double * 1)(float(value))
        attr = h.ok(h.L.H5Acreate2(obj, name.encode(), ftype, scalar, 0, 0), f"H5Acreate2({name})")
        try:
            h.ok(h.L.H5Awrite(attr, mtype, buf), f"H5Awrite({name})")
        finally:
            h.L.H5Aclose(attr)
    finally:
        if t >= 0:
            h.L.H5Tclose(t)
        h.L.H5Sclose(scalar)
