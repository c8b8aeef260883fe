"""skelcl_amd -- SkelCL's skeleton execution layer, MI355X-native.

Python mirror of the reference's C++ API (include/SkelCL/SkelCL.h,
Map.h, Zip.h, Reduce.h, Scan.h, MapOverlap.h, AllPairs.h, Vector.h,
Matrix.h, Out.h, Local.h), over the C-ABI of libskelcl_amd.so:

    import skelcl_amd as skelcl
    skelcl.init()
    mult = skelcl.Zip("float(float, float)", "float func(float x, float y){ return x*y; }")
    sum_ = skelcl.Reduce("float(float)", "float func(float x, float y){ return x+y; }", "0")
    A = skelcl.Vector(np.arange(100, dtype=np.float32))
    C = sum_(mult(A, A))
    print(C[0])

Containers keep the reference's lazy coherence (VectorDef.h:550-626): data
moves to the GPU when a skeleton needs it and back when the host reads it.
The compute always runs in the HIP kernels of libskelcl_amd.so; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import re
import struct
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import (SkelError, check, lib, take_string, SKL_PADDING_NEAREST,
                   SKL_PADDING_NEUTRAL, SKL_PREC_BF16X3, SKL_PREC_DEFAULT, SKL_PREC_FP32,
                   SKL_PREC_FP8, SKL_PREC_FP64)

NEUTRAL = SKL_PADDING_NEUTRAL
NEAREST = SKL_PADDING_NEAREST
PREC_DEFAULT, PREC_FP32, PREC_BF16X3, PREC_FP8, PREC_FP64 = (
    SKL_PREC_DEFAULT, SKL_PREC_FP32, SKL_PREC_BF16X3, SKL_PREC_FP8, SKL_PREC_FP64)

__all__ = [
    "init", "terminate", "device_info", "synchronize", "Vector", "Matrix", "IndexVector",
    "IndexMatrix", "out", "Local", "Map", "Zip", "Reduce", "Scan", "MapOverlap", "AllPairs",
    "DeviceBuffer", "Stream", "Event", "NEUTRAL", "NEAREST", "SkelError", "sha1",
]

# ---------------------------------------------------------------------------
# element types: kernel-source name -> numpy dtype
# ---------------------------------------------------------------------------
_SCALARS = {
    "float": np.float32, "double": np.float64, "int": np.int32, "uint": np.uint32,
    "unsigned int": np.uint32, "unsigned": np.uint32, "long": np.int64, "ulong": np.uint64,
    "unsigned long": np.uint64, "size_t": np.uint64, "Index": np.uint64, "char": np.int8,
    "uchar": np.uint8, "unsigned char": np.uint8, "short": np.int16, "ushort": np.uint16,
    "unsigned short": np.uint16,
}
_VEC_BASES = {"char": np.int8, "uchar": np.uint8, "short": np.int16, "ushort": np.uint16,
              "int": np.int32, "uint": np.uint32, "long": np.int64, "ulong": np.uint64,
              "float": np.float32, "double": np.float64}


def dtype_of(name: str) -> np.dtype:
    name = " ".join(name.replace("skelcl::", "").split())
    if name in _SCALARS:
        return np.dtype(_SCALARS[name])
    if name == "IndexPoint":
        return np.dtype([("x", np.uint64), ("y", np.uint64)])
    m = re.fullmatch(r"([a-z]+)(2|3|4|8|16)", name)
    if m and m.group(1) in _VEC_BASES:
        n = int(m.group(2))
        n = 4 if n == 3 else n
        return np.dtype((_VEC_BASES[m.group(1)], (n,)))
    raise SkelError(_lib.SKL_ERR_INVALID, f"unknown element type '{name}'")


def _split_signature(sig: str):
    m = re.fullmatch(r"\s*([\w:\s]+?)\s*\((.*)\)\s*", sig)
    if not m:
        raise SkelError(_lib.SKL_ERR_INVALID, f"bad skeleton signature '{sig}'")
    out = " ".join(m.group(1).split()).replace("skelcl::", "")
    ins = [" ".join(a.split()).replace("skelcl::", "") for a in m.group(2).split(",")
           if a.strip()]
    return out, ins


def _type(name: str):
    if name == "void":
        return _lib.skl_type(b"void", 0)
    return _lib.skl_type(name.encode(), dtype_of(name).itemsize)


# ---------------------------------------------------------------------------
# runtime
# ---------------------------------------------------------------------------
def init(ndevices: int = 1) -> None:
    """skelcl::init(nDevices(n)) -- collapses to GPU 0 (src/SkelCL.cpp:51-54)."""
    check(lib().skl_init(ndevices, 0))


def terminate() -> None:
    check(lib().skl_terminate())


def is_initialized() -> bool:
    return bool(lib().skl_is_initialized())


def synchronize() -> None:
    check(lib().skl_device_synchronize())


def device_info() -> dict:
    info = _lib.skl_device_info()
    check(lib().skl_device_get_info(0, C.byref(info)))
    return {f: (getattr(info, f).decode() if isinstance(getattr(info, f), bytes)
                else getattr(info, f)) for f, _ in info._fields_}


def sha1(data: bytes) -> str:
    out = C.create_string_buffer(41)
    check(lib().skl_sha1_hex(data, len(data), out))
    return out.value.decode()


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        check(lib().skl_stream_create(C.byref(self.handle)))

    def synchronize(self):
        check(lib().skl_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().skl_stream_destroy(self.handle)
        except Exception:
            pass


def _stream(s) -> C.c_void_p:
    if s is None:
        return None
    return s.handle if isinstance(s, Stream) else s


class Event:
    def __init__(self):
        self.handle = C.c_void_p()
        check(lib().skl_event_create(C.byref(self.handle)))

    def record(self, stream=None):
        check(lib().skl_event_record(self.handle, _stream(stream)))

    def synchronize(self):
        check(lib().skl_event_synchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().skl_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().skl_event_destroy(self.handle)
        except Exception:
            pass


class DeviceBuffer:
    """DeviceBuffer (src/DeviceBuffer.cpp:84-174): raw device memory."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib().skl_malloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 1))
        b.upload(a, stream)
        return b

    @property
    def value(self) -> int:
        return self.ptr.value or 0

    def upload(self, a: np.ndarray, stream=None, offset: int = 0, blocking=True):
        a = np.ascontiguousarray(a)
        if a.nbytes + offset > self.nbytes:
            raise SkelError(_lib.SKL_ERR_INVALID, "upload larger than the buffer")
        check(lib().skl_memcpy_h2d(self.ptr, offset, a.ctypes.data_as(C.c_void_p), a.nbytes,
                                   _stream(stream), int(blocking)))

    def download(self, a: np.ndarray, stream=None, offset: int = 0, blocking=True):
        if not a.flags.c_contiguous:
            raise SkelError(_lib.SKL_ERR_INVALID, "download target must be contiguous")
        if a.nbytes + offset > self.nbytes:
            raise SkelError(_lib.SKL_ERR_INVALID, "download larger than the buffer")
        check(lib().skl_memcpy_d2h(a.ctypes.data_as(C.c_void_p), self.ptr, offset, a.nbytes,
                                   _stream(stream), int(blocking)))
        return a

    def to_numpy(self, dtype, count=None, stream=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        a = np.empty(n, dt)
        return self.download(a, stream)

    def fill(self, value: bytes, elem_size: int, count: int, stream=None):
        buf = C.create_string_buffer(bytes(value), elem_size)
        check(lib().skl_fill(self.ptr, buf, elem_size, count, _stream(stream)))

    def __del__(self):
        try:
            if self.ptr:
                lib().skl_free(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# containers (VectorDef.h / MatrixDef.h lazy coherence)
# ---------------------------------------------------------------------------
class _Container:
    def __init__(self, host: np.ndarray, etype: Optional[str]):
        self._host = host
        self.etype = etype
        self._dev: Optional[DeviceBuffer] = None
        self._host_valid = True
        self._dev_valid = False

    # --- coherence (VectorDef.h:550-626)
    def device_ptr(self) -> int:
        """Upload when the host copy is newer (startUpload)."""
        if self._dev is None or self._dev.nbytes < max(self._host.nbytes, 1):
            self._dev = DeviceBuffer(max(self._host.nbytes, 1))
            self._dev_valid = False
        if not self._dev_valid:
            self._dev.upload(self._host)
            self._dev_valid = True
        return self._dev.value

    def device_buffer_for_output(self) -> int:
        """Device storage that a skeleton will overwrite (no upload)."""
        if self._dev is None or self._dev.nbytes < max(self._host.nbytes, 1):
            self._dev = DeviceBuffer(max(self._host.nbytes, 1))
        return self._dev.value

    def data_on_device_modified(self):
        self._dev_valid = True
        self._host_valid = False

    def data_on_host_modified(self):
        self._host_valid = True
        self._dev_valid = False

    def _sync_host(self):
        if not self._host_valid:
            check(lib().skl_device_synchronize())
            self._dev.download(self._host)
            self._host_valid = True

    def numpy(self) -> np.ndarray:
        self._sync_host()
        return self._host

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    @property
    def dtype(self):
        return self._host.dtype


class Vector(_Container):
    """skelcl::Vector<T> (Vector.h): 1-D container."""

    def __init__(self, data=None, size: Optional[int] = None, etype: Optional[str] = None,
                 value=None):
        if data is not None:
            host = np.ascontiguousarray(np.array(data, copy=True))
            etype = _etype_for_array(host, etype)
        else:
            etype = etype or "float"
            dt = dtype_of(etype)
            host = np.zeros(size or 0, dt) if value is None else np.full(size or 0, value, dt)
        super().__init__(host, etype)

    def __len__(self):
        return self._host.shape[0]

    def size(self) -> int:
        return len(self)

    def __getitem__(self, i):
        return self.numpy()[i]

    def __setitem__(self, i, v):
        self._sync_host()
        self._host[i] = v
        self.data_on_host_modified()

    def resize(self, n: int):
        self._sync_host()
        h = np.zeros(n, self._host.dtype)
        k = min(n, len(self))
        h[:k] = self._host[:k]
        self._host = h
        self._dev = None
        self.data_on_host_modified()


class Matrix(_Container):
    """skelcl::Matrix<T> (Matrix.h): row-major 2-D container."""

    def __init__(self, data=None, shape=None, etype: Optional[str] = None, value=None):
        if data is not None:
            host = np.ascontiguousarray(np.array(data, copy=True))
            if host.ndim == 1 and shape is not None:
                host = host.reshape(shape)
            etype = _etype_for_array(host, etype)
        else:
            etype = etype or "float"
            dt = dtype_of(etype)
            host = np.zeros(shape, dt) if value is None else np.full(shape, value, dt)
        super().__init__(host, etype)

    def row_count(self) -> int:
        return self._host.shape[0]

    def column_count(self) -> int:
        return self._host.shape[1]

    @property
    def shape(self):
        return self._host.shape[:2]

    def __getitem__(self, idx):
        return self.numpy()[idx]

    def __setitem__(self, idx, v):
        self._sync_host()
        self._host[idx] = v
        self.data_on_host_modified()


class IndexVector:
    """skelcl::IndexVector (IndexVector.h): the indices 0..n-1, no storage."""

    def __init__(self, n: int):
        self.n = int(n)

    def __len__(self):
        return self.n


class IndexMatrix:
    """skelcl::IndexMatrix (IndexMatrix.h): the points (x, y), no storage."""

    def __init__(self, shape):
        self.shape = (int(shape[0]), int(shape[1]))


def _segment_offsets(offsets, n: int) -> "Vector":
    """Segment offsets as a device Vector of int64.  Host offsets are checked
    here (the device check only flags them); a Vector is passed as it is."""
    if isinstance(offsets, Vector):
        if len(offsets) < 1:
            raise SkelError(_lib.SKL_ERR_INVALID, "offsets need at least one entry")
        return offsets
    off = np.ascontiguousarray(offsets, dtype=np.int64).ravel()
    if off.size < 1:
        raise SkelError(_lib.SKL_ERR_INVALID, "offsets need at least one entry")
    if off[0] < 0 or off[-1] > n or np.any(np.diff(off) < 0):
        raise SkelError(_lib.SKL_ERR_INVALID,
                        f"offsets must be non-decreasing and inside [0, {n}]")
    return Vector(off)


def _etype_for_array(host: np.ndarray, hint: Optional[str]) -> str:
    """Element type of a host array: the given name when its layout matches
    (a vector type such as float4 is an array whose last axis holds the
    components), else the scalar name of the dtype."""
    dt = host.dtype
    if hint is not None:
        want = dtype_of(hint)
        if want.subdtype is not None:
            base, shp = want.subdtype
            if dt == base and host.ndim >= 2 and host.shape[-1] == shp[0]:
                return hint
        elif want == dt:
            return hint
        raise SkelError(_lib.SKL_ERR_INVALID,
                        f"array of {dt} with shape {host.shape} is not a container of {hint}")
    for name, t in (("float", np.float32), ("double", np.float64), ("int", np.int32),
                    ("uint", np.uint32), ("long", np.int64), ("ulong", np.uint64),
                    ("char", np.int8), ("uchar", np.uint8), ("short", np.int16),
                    ("ushort", np.uint16)):
        if dt == np.dtype(t):
            return name
    raise SkelError(_lib.SKL_ERR_INVALID, f"no element type for numpy dtype {dt}")


class out:
    """skelcl::out(container) (Out.h:65-97): marks an argument as written."""

    def __init__(self, container):
        self.container = container


class Local:
    """skelcl::Local(bytes) (Local.h:58-87): dynamic __local memory."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)


# ---------------------------------------------------------------------------
# skeletons
# ---------------------------------------------------------------------------
_PACK = {1: "b", 2: "h", 4: "i", 8: "q"}
_FMT = {"float": "f", "double": "d", "int": "i", "uint": "I", "unsigned int": "I",
        "unsigned": "I", "long": "q", "ulong": "Q", "unsigned long": "Q", "size_t": "Q",
        "Index": "Q", "char": "b", "uchar": "B", "unsigned char": "B", "short": "h",
        "ushort": "H", "unsigned short": "H", "bool": "?"}


class _Skeleton:
    kind = -1

    def __init__(self, sig_types: Sequence[str], source: str, func: str = "func", **kw):
        d = _lib.skl_skeleton_desc()
        d.kind = self.kind
        self._keep = [source.encode(), func.encode()]
        d.source, d.func = self._keep[0], self._keep[1]
        for key in ("source2", "func2", "identity", "neutral", "deriv_source", "deriv_func",
                    "deriv_func2", "build_options"):
            v = kw.get(key)
            if v is not None:
                b = str(v).encode()
                self._keep.append(b)
                setattr(d, key, b)
        d.ntypes = len(sig_types)
        for i, t in enumerate(sig_types):
            d.types[i] = _type(t)
            self._keep.append(d.types[i].name)
        d.overlap_range = kw.get("overlap_range", 0)
        d.padding = kw.get("padding", SKL_PADDING_NEAREST)
        d.work_group_size = kw.get("work_group_size", 0)
        d.precision = kw.get("precision", SKL_PREC_DEFAULT)
        self.types = list(sig_types)
        self.handle = C.c_void_p()
        check(lib().skl_skeleton_create(C.byref(d), C.byref(self.handle)))
        n = C.c_int()
        check(lib().skl_skeleton_extra_count(self.handle, C.byref(n)))
        self.extras = []
        for i in range(n.value):
            k, sz, tb = C.c_int(), C.c_size_t(), C.create_string_buffer(256)
            check(lib().skl_skeleton_extra_info(self.handle, i, C.byref(k), C.byref(sz), tb,
                                                256))
            self.extras.append((k.value, sz.value, tb.value.decode()))

    def __del__(self):
        try:
            if self.handle:
                lib().skl_skeleton_release(self.handle)
        except Exception:
            pass

    def set_work_group_size(self, wg: int):
        check(lib().skl_skeleton_set_work_group_size(self.handle, wg))

    def set_precision(self, p: int):
        check(lib().skl_skeleton_set_precision(self.handle, p))

    def source(self) -> str:
        p = C.c_void_p()
        check(lib().skl_skeleton_source(self.handle, C.byref(p)))
        return take_string(p)

    # --- extra arguments (KernelUtil.h:134-245)
    def _marshal(self, args):
        if len(args) != len(self.extras):
            raise SkelError(_lib.SKL_ERR_INVALID,
                            f"the user function takes {len(self.extras)} extra argument(s), "
                            f"{len(args)} given")
        arr = (_lib.skl_arg * max(1, len(args)))()
        keep, written = [], []
        for i, (a, (kind, size, tname)) in enumerate(zip(args, self.extras)):
            if isinstance(a, out):
                written.append(a.container)
                a = a.container
            if isinstance(a, Local):
                arr[i].kind, arr[i].size = _lib.SKL_ARG_LOCAL, a.nbytes
            elif isinstance(a, Matrix):
                arr[i].kind = _lib.SKL_ARG_MATRIX
                arr[i].value = a.device_ptr()
                arr[i].cols = a.column_count()
            elif isinstance(a, (Vector, DeviceBuffer)):
                arr[i].kind = _lib.SKL_ARG_BUFFER
                arr[i].value = a.device_ptr() if isinstance(a, Vector) else a.value
            elif isinstance(a, int) and kind == _lib.SKL_ARG_BUFFER:
                arr[i].kind, arr[i].value = _lib.SKL_ARG_BUFFER, a
            else:
                raw = self._pack_value(a, size, tname)
                b = C.create_string_buffer(raw, len(raw))
                keep.append(b)
                arr[i].kind = _lib.SKL_ARG_VALUE
                arr[i].value = C.cast(b, C.c_void_p)
                arr[i].size = len(raw)
        return arr, len(args), keep, written

    @staticmethod
    def _pack_value(a, size, tname) -> bytes:
        if isinstance(a, (bytes, bytearray)):
            return bytes(a)
        if isinstance(a, np.ndarray) or isinstance(a, np.generic):
            return np.ascontiguousarray(a).tobytes()
        t = " ".join(tname.replace("const", " ").split())
        if t in _FMT:
            return struct.pack("<" + _FMT[t], a)
        if isinstance(a, (tuple, list)):
            dt = dtype_of(t)
            return np.asarray(a, dtype=dt.base).tobytes()
        raise SkelError(_lib.SKL_ERR_INVALID, f"cannot pass {a!r} as '{tname}'")

    @staticmethod
    def _written(written):
        for c in written:
            c.data_on_device_modified()


def _out_vector(out_arg, n, etype):
    if isinstance(out_arg, out):
        out_arg = out_arg.container
    if out_arg is None:
        return Vector(size=n, etype=etype)
    if isinstance(out_arg, Vector) and len(out_arg) < n:
        out_arg.resize(n)
    return out_arg


class Map(_Skeleton):
    """skelcl::Map<Tout(Tin)> / Map<void(Tin)> / Map<Tout(Index)> /
    Map<Tout(IndexPoint)> (MapDef.h:75-752)."""

    def __init__(self, signature: str, source: str, func: str = "func", **kw):
        tout, tins = _split_signature(signature)
        tin = tins[0]
        if tin == "Index":
            self.kind = _lib.SKL_MAP_INDEX
            types = [tout]
        elif tin == "IndexPoint":
            self.kind = _lib.SKL_MAP_INDEXPOINT
            types = [tout]
        else:
            self.kind = _lib.SKL_MAP
            types = [tout, tin]
        super().__init__(types, source, func, **kw)

    def __call__(self, inp, *args, out=None, stream=None):
        arr, n_args, keep, written = self._marshal(args)
        tout = self.types[0]
        if self.kind == _lib.SKL_MAP_INDEX:
            n = len(inp)
            o = None if tout == "void" else _out_vector(out, n, tout)
            check(lib().skl_map_index_run(self.handle, o.device_buffer_for_output() if o else None,
                                          n, 0, arr, n_args, _stream(stream)))
        elif self.kind == _lib.SKL_MAP_INDEXPOINT:
            rows, cols = inp.shape
            o = None
            if tout != "void":
                o = out.container if isinstance(out, globals()["out"]) else out
                if o is None:
                    o = Matrix(shape=(rows, cols), etype=tout)
            check(lib().skl_map_indexpoint_run(
                self.handle, o.device_buffer_for_output() if o else None, rows, cols, 0, arr,
                n_args, _stream(stream)))
        else:
            n = inp._host.shape[0] if isinstance(inp, Vector) else inp._host.size
            src = inp.device_ptr()
            if tout == "void":
                o = None
            elif isinstance(inp, Matrix):
                o = out.container if isinstance(out, globals()["out"]) else out
                if o is None:
                    o = Matrix(shape=inp.shape, etype=tout)
            else:
                o = _out_vector(out, n, tout)
            dst = o.device_buffer_for_output() if o is not None else None
            check(lib().skl_map_run(self.handle, dst, src, n, arr, n_args, _stream(stream)))
        if o is not None:
            o.data_on_device_modified()
        self._written(written)
        return o

    def backward(self, inp, d_out, *args, stream=None):
        """Adjoint (needs deriv_source / deriv_func = f'): d_in = d_out * f'(in)."""
        arr, n_args, keep, _ = self._marshal(args)
        n = len(inp)
        d = Vector(size=n, etype=self.types[1])
        check(lib().skl_map_backward(self.handle, d.device_buffer_for_output(), inp.device_ptr(),
                                     d_out.device_ptr(), n, arr, n_args, _stream(stream)))
        d.data_on_device_modified()
        return d


class Zip(_Skeleton):
    """skelcl::Zip<Tout(Tl, Tr)> (ZipDef.h:70-442)."""
    kind = _lib.SKL_ZIP

    def __init__(self, signature: str, source: str, func: str = "func", **kw):
        tout, tins = _split_signature(signature)
        super().__init__([tout, tins[0], tins[1]], source, func, **kw)

    def __call__(self, left, right, *args, out=None, stream=None):
        n = left._host.size if isinstance(left, Matrix) else len(left)
        nr = right._host.size if isinstance(right, Matrix) else len(right)
        if n != nr:  # ZipDef.h:101,126
            raise SkelError(_lib.SKL_ERR_INVALID, f"Zip inputs differ in size ({n} vs {nr})")
        arr, n_args, keep, written = self._marshal(args)
        lp, rp = left.device_ptr(), right.device_ptr()
        tout = self.types[0]
        if tout == "void":
            o = None
        elif isinstance(left, Matrix):
            o = out.container if isinstance(out, globals()["out"]) else out
            if o is None:
                o = Matrix(shape=left.shape, etype=tout)
        else:
            o = _out_vector(out, n, tout)
        dst = None if o is None else (o.device_ptr() if o is left or o is right
                                      else o.device_buffer_for_output())
        check(lib().skl_zip_run(self.handle, dst, lp, rp, n, arr, n_args, _stream(stream)))
        if o is not None:
            o.data_on_device_modified()
        self._written(written)
        return o

    def backward(self, left, right, d_out, *args, stream=None):
        """Adjoint (needs deriv_source with deriv_func = df/dl, deriv_func2 =
        df/dr): returns (d_left, d_right)."""
        arr, n_args, keep, _ = self._marshal(args)
        n = len(left)
        dl = Vector(size=n, etype=self.types[1])
        dr = Vector(size=n, etype=self.types[2])
        check(lib().skl_zip_backward(self.handle, dl.device_buffer_for_output(),
                                     dr.device_buffer_for_output(), left.device_ptr(),
                                     right.device_ptr(), d_out.device_ptr(), n, arr, n_args,
                                     _stream(stream)))
        dl.data_on_device_modified()
        dr.data_on_device_modified()
        return dl, dr


class Reduce(_Skeleton):
    """skelcl::Reduce<T(T)> (ReduceDef.h:79-293): returns a Vector of size 1."""
    kind = _lib.SKL_REDUCE

    def __init__(self, signature: str, source: str, identity: str = "0",
                 func: str = "func", **kw):
        tout, _ = _split_signature(signature)
        self.identity = identity
        super().__init__([tout], source, func, identity=identity, **kw)

    def __call__(self, inp, *args, out=None, stream=None):
        arr, n_args, keep, written = self._marshal(args)
        n = len(inp) if isinstance(inp, Vector) else inp._host.size
        o = _out_vector(out, 1, self.types[0])
        check(lib().skl_reduce_run(self.handle, o.device_buffer_for_output(), inp.device_ptr(),
                                   n, arr, n_args, _stream(stream)))
        o.data_on_device_modified()
        self._written(written)
        return o

    def fused(self, zip_: "Zip", left, right, stream=None):
        """Zip -> Reduce in one kernel (extension): reduce_i zip(l[i], r[i])."""
        o = Vector(size=1, etype=self.types[0])
        check(lib().skl_zip_reduce_run(zip_.handle, self.handle, o.device_buffer_for_output(),
                                       left.device_ptr(), right.device_ptr(), len(left),
                                       _stream(stream)))
        o.data_on_device_modified()
        return o

    def backward(self, d_out, n, stream=None):
        """Adjoint of Reduce(+): every input gets d_out[0]."""
        d = Vector(size=n, etype=self.types[0])
        check(lib().skl_reduce_backward(self.handle, d.device_buffer_for_output(),
                                        d_out.device_ptr(), n, _stream(stream)))
        d.data_on_device_modified()
        return d

    def segmented(self, inp, offsets, stream=None):
        """Batched segmented Reduce(+) (extension): out[s] = sum of
        inp[offsets[s]:offsets[s+1]], all segments in one main launch after
        small prologue launches (batch_run does it in one launch)."""
        off = _segment_offsets(offsets, len(inp))
        nseg = len(off) - 1
        o = Vector(size=nseg, etype=self.types[0])
        check(lib().skl_reduce_segmented_run(self.handle, o.device_buffer_for_output(),
                                             inp.device_ptr(), len(inp), off.device_ptr(), nseg,
                                             _stream(stream)))
        o.data_on_device_modified()
        return o


class Scan(_Skeleton):
    """skelcl::Scan<T(T)> (ScanDef.h:70-327): exclusive by default, as the
    reference (ScanKernel.cl:142-153); inclusive=True gives the inclusive form."""
    kind = _lib.SKL_SCAN

    def __init__(self, signature: str, source: str, identity: str = "0",
                 func: str = "func", **kw):
        tout, _ = _split_signature(signature)
        super().__init__([tout], source, func, identity=identity, **kw)

    def __call__(self, inp, *args, out=None, inclusive=False, stream=None):
        arr, n_args, keep, written = self._marshal(args)
        n = len(inp)
        o = _out_vector(out, n, self.types[0])
        src = inp.device_ptr()
        dst = o.device_ptr() if o is inp else o.device_buffer_for_output()
        check(lib().skl_scan_run(self.handle, dst, src, n, int(bool(inclusive)), arr, n_args,
                                 _stream(stream)))
        o.data_on_device_modified()
        self._written(written)
        return o

    def backward(self, d_out, inclusive=False, stream=None):
        """Adjoint of Scan(+): the scan of d_out read from the end."""
        n = len(d_out)
        d = Vector(size=n, etype=self.types[0])
        check(lib().skl_scan_backward(self.handle, d.device_buffer_for_output(),
                                      d_out.device_ptr(), n, int(bool(inclusive)),
                                      _stream(stream)))
        d.data_on_device_modified()
        return d

    def segmented(self, inp, offsets, inclusive=False, stream=None):
        """Batched segmented Scan(+) (extension): every segment scanned on its
        own, all segments in one main launch after a small prologue launch
        (batch_run does it in one launch)."""
        off = _segment_offsets(offsets, len(inp))
        nseg = len(off) - 1
        o = Vector(size=len(inp), etype=self.types[0])
        check(lib().skl_scan_segmented_run(self.handle, o.device_buffer_for_output(),
                                           inp.device_ptr(), len(inp), off.device_ptr(), nseg,
                                           int(bool(inclusive)), _stream(stream)))
        o.data_on_device_modified()
        return o


def batch_run(inp, offsets, reduce: "Reduce" = None, scan: "Scan" = None, inclusive=False,
              descs=None, count: int = 0, precision: int = PREC_FP32, workers_per_cu: int = 0,
              stream=None):
    """One launch for a batch (extension, skl_batch_run): the segment sums
    (with `reduce`) and the segmented scan (with `scan`) of the float or int
    Vector `inp` over `offsets`, and `count` small products described by the device
    buffer `descs` (skl_gemm_desc records).  Returns (sums, scan) Vectors,
    None for a part left out.  Host offsets are checked here; device
    offsets are checked by the launch (skl_batch_check)."""
    n = len(inp) if inp is not None else 0
    off = _segment_offsets(offsets, n) if offsets is not None else None
    nseg = len(off) - 1 if off is not None else 0
    sums = Vector(size=nseg, etype=reduce.types[0]) if reduce is not None and nseg > 0 else None
    sc = Vector(size=n, etype=scan.types[0]) if scan is not None and nseg > 0 else None
    check(lib().skl_batch_run(
        reduce.handle if sums is not None else None,
        sums.device_buffer_for_output() if sums is not None else None,
        scan.handle if sc is not None else None,
        sc.device_buffer_for_output() if sc is not None else None,
        int(bool(inclusive)), inp.device_ptr() if n else None, n,
        off.device_ptr() if off is not None else None, nseg,
        descs.ptr if descs is not None else None, int(count), int(precision),
        int(workers_per_cu), _stream(stream)))
    for v in (sums, sc):
        if v is not None:
            v.data_on_device_modified()
    return sums, sc


class MapOverlap(_Skeleton):
    """skelcl::MapOverlap<Tout(Tin)> (MapOverlapDef.h:71-279); defaults R=1,
    NEAREST, neutral 0 as MapOverlap.h:119-122."""
    kind = _lib.SKL_MAPOVERLAP

    def __init__(self, signature: str, source: str, overlap_range: int = 1,
                 padding: int = NEAREST, neutral=0, func: str = "func", **kw):
        tout, tins = _split_signature(signature)
        super().__init__([tout, tins[0]], source, func, overlap_range=overlap_range,
                         padding=padding, neutral=neutral, **kw)

    def __call__(self, inp: Matrix, *args, out=None, stream=None):
        arr, n_args, keep, written = self._marshal(args)
        rows, cols = inp.shape
        o = out.container if isinstance(out, globals()["out"]) else out
        if o is None:
            o = Matrix(shape=(rows, cols), etype=self.types[0])
        check(lib().skl_mapoverlap_run(self.handle, o.device_buffer_for_output(),
                                       inp.device_ptr(), rows, cols, arr, n_args,
                                       _stream(stream)))
        o.data_on_device_modified()
        self._written(written)
        return o

    def stencil_backward(self, d_out: Matrix, weights: Vector, stream=None):
        """Adjoint of a weighted-stencil user function: the transposed stencil
        of d_out under the same padding."""
        rows, cols = d_out.shape
        d = Matrix(shape=(rows, cols), etype=self.types[1])
        check(lib().skl_mapoverlap_stencil_backward(self.handle, d.device_buffer_for_output(),
                                                    d_out.device_ptr(), weights.device_ptr(),
                                                    rows, cols, _stream(stream)))
        d.data_on_device_modified()
        return d


class AllPairs(_Skeleton):
    """skelcl::AllPairs<Tout(Tl, Tr)> (AllPairsDef.h:77-426):
    AllPairs(sig, reduce=Reduce, zip=Zip) or AllPairs(sig, source, func)."""

    def __init__(self, signature: str, source: Optional[str] = None, func: str = "func",
                 reduce: Optional[Reduce] = None, zip: Optional[Zip] = None, **kw):
        tout, tins = _split_signature(signature)
        types = [tout, tins[0], tins[1]]
        if reduce is not None and zip is not None:
            self.kind = _lib.SKL_ALLPAIRS
            super().__init__(types, reduce._keep[0].decode(), reduce._keep[1].decode(),
                             source2=zip._keep[0].decode(), func2=zip._keep[1].decode(),
                             identity=getattr(reduce, "identity", "0"), **kw)
        else:
            self.kind = _lib.SKL_ALLPAIRS_GENERAL
            super().__init__(types, source, func, **kw)

    def __call__(self, left: Matrix, right: Matrix, *args, out=None, stream=None):
        if left.column_count() != right.row_count():  # AllPairsDef.h:127-129
            raise SkelError(_lib.SKL_ERR_INVALID, "left.columnCount() != right.rowCount()")
        arr, n_args, keep, written = self._marshal(args)
        M, K = left.shape
        N = right.column_count()
        o = out.container if isinstance(out, globals()["out"]) else out
        if o is None:
            o = Matrix(shape=(M, N), etype=self.types[0])
        check(lib().skl_allpairs_run(self.handle, o.device_buffer_for_output(),
                                     left.device_ptr(), right.device_ptr(), M, N, K, arr,
                                     n_args, _stream(stream)))
        o.data_on_device_modified()
        self._written(written)
        return o

    def backward(self, left: Matrix, right: Matrix, d_out: Matrix, stream=None):
        """Adjoint of the (+,*) contraction: (dA, dB) = (dC.B^T, A^T.dC)."""
        M, K = left.shape
        N = right.column_count()
        dA = Matrix(shape=(M, K), etype=self.types[1])
        dB = Matrix(shape=(K, N), etype=self.types[2])
        check(lib().skl_allpairs_backward(self.handle, dA.device_buffer_for_output(),
                                          dB.device_buffer_for_output(), left.device_ptr(),
                                          right.device_ptr(), d_out.device_ptr(), M, N, K,
                                          _stream(stream)))
        dA.data_on_device_modified()
        dB.data_on_device_modified()
        return dA, dB


def check_source(kind: int, types: Sequence[str], source: str, func: str = "func",
                 **kw) -> str:
    """Generate and hiprtc-compile a skeleton's fused kernel without a GPU;
    returns the generated source (raises on a compile error)."""
    d = _lib.skl_skeleton_desc()
    d.kind = kind
    keep = [source.encode(), func.encode()]
    d.source, d.func = keep[0], keep[1]
    for key in ("source2", "func2", "identity", "neutral", "deriv_source", "deriv_func",
                "deriv_func2", "build_options"):
        v = kw.get(key)
        if v is not None:
            b = str(v).encode()
            keep.append(b)
            setattr(d, key, b)
    d.ntypes = len(types)
    for i, t in enumerate(types):
        d.types[i] = _type(t)
        keep.append(d.types[i].name)
    d.overlap_range = kw.get("overlap_range", 1)
    d.padding = kw.get("padding", SKL_PADDING_NEAREST)
    d.work_group_size = kw.get("work_group_size", 0)
    d.precision = kw.get("precision", SKL_PREC_DEFAULT)
    src, log = C.c_void_p(), C.c_void_p()
    st = lib().skl_skeleton_check(C.byref(d), C.byref(src), C.byref(log))
    text, lg = take_string(src), take_string(log)
    if st != 0:
        msg = lib().skl_last_error()
        raise SkelError(st, (msg.decode(errors="replace") if msg else "") + "\n" + lg)
    return text
