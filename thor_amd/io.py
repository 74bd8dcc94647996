"""ctypes binding of include/thor_io.h (libthor_io.so): THOR's .thr
configuration files and the HDF5 layout of its output files
(storage.h: 1-D datasets with "Variable" and "units" string attributes)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libthor_io.so")
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise capi.ThorError(capi.THOR_ERR_IO, f"{LIB_PATH} not built; run python -m thor_amd.build")
        capi.lib()  # the core library first (libthor_io.so links it)
        L = C.CDLL(LIB_PATH)
        L.thor_config_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.thor_config_free.argtypes = [C.c_void_p]
        L.thor_config_free.restype = None
        L.thor_config_get.argtypes = [C.c_void_p, C.c_char_p]
        L.thor_config_get.restype = C.c_char_p
        L.thor_config_unknown_count.argtypes = [C.c_void_p]
        L.thor_config_unknown.argtypes = [C.c_void_p, C.c_int]
        L.thor_config_unknown.restype = C.c_char_p
        L.thor_config_run.argtypes = [C.c_void_p, C.c_void_p]
        L.thor_h5_write.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.POINTER(C.c_double), C.c_long, C.c_char_p,
                                    C.c_char_p]
        L.thor_h5_size.argtypes = [C.c_char_p, C.c_char_p]
        L.thor_h5_size.restype = C.c_long
        L.thor_h5_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_double), C.c_long]
        L.thor_h5_names.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.thor_h5_is_int.argtypes = [C.c_char_p, C.c_char_p]
        L.thor_h5_attr.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != capi.THOR_OK:
        raise capi.ThorError(rc, capi.lib().thor_last_error().decode())


def h5_names(path: str) -> list:
    buf = C.create_string_buffer(1 << 16)
    _check(lib().thor_h5_names(path.encode(), buf, len(buf)))
    return [n for n in buf.value.decode().split("\n") if n]


def h5_read(path: str, name: str) -> np.ndarray:
    n = lib().thor_h5_size(path.encode(), name.encode())
    if n < 0:
        raise KeyError(name)
    a = np.zeros(n)
    _check(lib().thor_h5_read(path.encode(), name.encode(), a.ctypes.data_as(C.POINTER(C.c_double)), n))
    return a


def h5_attr(path: str, name: str, attr: str) -> str:
    buf = C.create_string_buffer(4096)
    _check(lib().thor_h5_attr(path.encode(), name.encode(), attr.encode(), buf, len(buf)))
    return buf.value.decode()


def h5_is_int(path: str, name: str) -> bool:
    r = lib().thor_h5_is_int(path.encode(), name.encode())
    if r < 0:
        raise KeyError(name)
    return bool(r)


def h5_layout(path: str) -> dict:
    """{name: {"int": bool, "size": n, "Variable": str, "units": str, "value": x (scalars)}}"""
    out = {}
    for n in h5_names(path):
        a = h5_read(path, n)
        d = {"int": h5_is_int(path, n), "size": int(a.size), "Variable": h5_attr(path, n, "Variable"),
             "units": h5_attr(path, n, "units")}
        if a.size == 1:
            d["value"] = float(a[0])
        out[n] = d
    return out


class Config:
    """A parsed .thr file."""

    def __init__(self, path: str):
        h = C.c_void_p()
        _check(lib().thor_config_load(path.encode(), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().thor_config_free(self._h)
            self._h = None

    def get(self, key: str):
        v = lib().thor_config_get(self._h, key.encode())
        return None if v is None else v.decode()

    def unknown(self) -> list:
        L = lib()
        return [L.thor_config_unknown(self._h, i).decode() for i in range(L.thor_config_unknown_count(self._h))]
