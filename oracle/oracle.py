"""ORACLE — test infrastructure only.

ctypes binding of the serial FP64 restatement in thor_oracle.cpp. Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
LIB_PM = os.path.join(_HERE, "liboracle_pm.so")  # portable elementary functions
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compiles liboracle.so and liboracle_pm.so with the recipe in oracle/Makefile."""
    if force or not (os.path.exists(LIB) and os.path.exists(LIB_PM)):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s", "all"], check=True)
    return LIB


class _Grid(C.Structure):
    _fields_ = [("point_num", C.c_int), ("nv", C.c_int), ("nr", C.c_int), ("nl_region", C.c_int),
                ("point_local", _ip), ("maps", _ip), ("div", _dp), ("grad", _dp), ("areas", _dp),
                ("areasTr", _dp), ("areasT", _dp), ("nvecoa", _dp), ("nvecti", _dp), ("nvecte", _dp),
                ("func_r", _dp), ("lonlat", _dp), ("Altitude", _dp), ("Altitudeh", _dp)]


class _Params(C.Structure):
    _fields_ = [("A", C.c_double), ("Omega", C.c_double), ("Gravit", C.c_double), ("P_Ref", C.c_double),
                ("Rd", C.c_double), ("Cp", C.c_double), ("Diffc", C.c_double), ("DivDampc", C.c_double),
                ("timestep", C.c_double), ("glevel", C.c_int), ("HyDiffOrder", C.c_int),
                ("n_substeps", C.c_int), ("DeepModel", C.c_int), ("NonHydro", C.c_int), ("HyDiff", C.c_int),
                ("DivDampP", C.c_int), ("GravHeightVar", C.c_int), ("VertHyDiff", C.c_int),
                ("VertHyDiffOrder", C.c_int), ("Diffc_v", C.c_double), ("DiffSponge", C.c_int),
                ("order_diff_sponge", C.c_int), ("Dv_sponge", C.c_double), ("ns_diff_sponge", C.c_double),
                ("energy_equation", C.c_int)]


_libs: dict = {}


def lib(variant: str = ""):
    if variant not in _libs:
        build()
        L = C.CDLL(LIB_PM if variant == "pm" else LIB)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(_Grid), C.POINTER(_Params)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_destroy.restype = None
        L.oracle_array_size.argtypes = [C.c_void_p, C.c_char_p]
        L.oracle_array_size.restype = C.c_long
        L.oracle_set.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_long]
        L.oracle_get.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_long]
        for f in ("oracle_thor", "oracle_profx_held_suarez"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = None
        L.oracle_steps.argtypes = [C.c_void_p, C.c_int]
        L.oracle_steps.restype = None
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_get_threads.argtypes = []
        L.oracle_get_threads.restype = C.c_int
        if os.environ.get("THOR_ORACLE_THREADS"):
            L.oracle_set_threads(int(os.environ["THOR_ORACLE_THREADS"]))
        L.oracle_globdiag.argtypes = [C.c_void_p, _dp]
        L.oracle_globdiag.restype = None
        L.oracle_tree_sum.argtypes = [_dp, C.c_long]
        L.oracle_tree_sum.restype = C.c_double
        L.oracle_math.argtypes = [C.c_int, C.c_double, C.c_double]
        L.oracle_math.restype = C.c_double
        L.oracle_kernel.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
        L.oracle_kernel.restype = C.c_int
        for f in (L.oracle_set_newtonian, L.oracle_set_rt):
            f.argtypes = [C.c_void_p, _dp]
            f.restype = None
        L.oracle_set_sponge.argtypes = [C.c_void_p, _dp]
        L.oracle_set_sponge.restype = None
        L.oracle_set_boundary_layer.argtypes = [C.c_void_p, _dp]
        L.oracle_set_boundary_layer.restype = None
        L.oracle_zonal_table.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_long, C.POINTER(C.c_int)]
        L.oracle_zonal_table.restype = C.c_int
        L.oracle_profx.argtypes = [C.c_void_p, C.c_int]
        L.oracle_profx.restype = C.c_int
        L.oracle_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oracle_run.restype = C.c_int
        L.oracle_set_tracers.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_tracers.restype = None
        _libs[variant] = L
    return _libs[variant]


def set_threads(n: int, variant: str = "") -> int:
    """Host threads of the oracle's block and column loops (n < 1: the OpenMP
    default). The results do not depend on it. Returns the count in force."""
    L = lib(variant)
    L.oracle_set_threads(int(n))
    return L.oracle_get_threads()


def tree_sum(a: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    return lib().oracle_tree_sum(a.ctypes.data_as(_dp), a.size)


def math(fn: str, x: float, y: float = 0.0, variant: str = "") -> float:
    """Elementary function as the oracle evaluates it (pow, log, sin, cos, exp)."""
    return lib(variant).oracle_math({"pow": 0, "log": 1, "sin": 2, "cos": 3, "exp": 4}[fn], x, y)


def _structs(grid, params):
    g = _Grid(grid.P, grid.nv, grid.nr, 16, *(
        np.ascontiguousarray(getattr(grid, n)).ctypes.data_as(_ip if n in ("point_local", "maps") else _dp)
        for n in ("point_local", "maps", "div", "grad", "areas", "areasTr", "areasT", "nvecoa",
                  "nvecti", "nvecte", "func_r", "lonlat", "Altitude", "Altitudeh")))
    p = _Params(params.A, params.Omega, params.Gravit, params.P_Ref, params.Rd, params.Cp,
                params.Diffc, params.DivDampc, params.timestep, params.glevel, params.HyDiffOrder,
                params.n_substeps, params.DeepModel, params.NonHydro, params.HyDiff, params.DivDampP,
                params.GravHeightVar, params.VertHyDiff, params.VertHyDiffOrder, params.Diffc_v,
                params.DiffSponge, params.order_diff_sponge, params.Dv_sponge, params.ns_diff_sponge,
                params.thermo_equation)
    return g, p


class Oracle:
    """Serial host loop of THOR's kernels on a grid given in THOR's layout."""

    def __init__(self, grid, params, variant: str = ""):
        self.grid = grid  # keeps the geometry arrays alive
        self.L = lib(variant)
        g, p = _structs(grid, params)
        self._g = g
        self._h = C.c_void_p(self.L.oracle_create(C.byref(g), C.byref(p)))

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.oracle_destroy(self._h)
            self._h = None

    def set(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if self.L.oracle_set(self._h, name.encode(), a.ctypes.data_as(_dp), a.size) != 0:
            raise KeyError(name)

    def get(self, name: str) -> np.ndarray:
        n = self.L.oracle_array_size(self._h, name.encode())
        if n < 0:
            raise KeyError(name)
        a = np.zeros(n)
        self.L.oracle_get(self._h, name.encode(), a.ctypes.data_as(_dp), n)
        return a

    def set_state(self, st: dict):
        for k in ("Mh", "Wh", "W", "Rho", "pressure"):
            self.set(k, st[k])

    def get_state(self) -> dict:
        P, nv = self.grid.P, self.grid.nv
        shp = {"Mh": (P, nv, 3), "Wh": (P, nv + 1), "W": (P, nv), "Rho": (P, nv), "pressure": (P, nv)}
        return {k: self.get(k).reshape(s) for k, s in shp.items()}

    def thor(self):
        self.L.oracle_thor(self._h)

    def profx(self):
        self.L.oracle_profx_held_suarez(self._h)

    def steps(self, n: int):
        self.L.oracle_steps(self._h, n)

    def globdiag(self) -> np.ndarray:
        out = np.zeros(5)
        self.L.oracle_globdiag(self._h, out.ctypes.data_as(_dp))
        return out

    def set_tracers(self, ntr: int):
        """Chemistry tracers driven through the physics hooks (0: none)."""
        self.L.oracle_set_tracers(self._h, ntr)

    def set_newtonian(self, q):
        """q: a ctypes NewtonianParams (or 8 numbers in its field order)."""
        a = np.array([getattr(q, n) for n, _ in q._fields_] if hasattr(q, "_fields_") else q, dtype=np.float64)
        self.L.oracle_set_newtonian(self._h, a.ctypes.data_as(_dp))

    def set_rt(self, q):
        """q: a ctypes RTParams (or 11 numbers in its field order)."""
        a = np.array([getattr(q, n) for n, _ in q._fields_] if hasattr(q, "_fields_") else q, dtype=np.float64)
        self.L.oracle_set_rt(self._h, a.ctypes.data_as(_dp))

    def set_sponge(self, q):
        """q: a ctypes SpongeParams (or its first 6 to 11 numbers in field order); None: off."""
        if q is None:
            self.L.oracle_set_sponge(self._h, None)
            return
        a = np.array([float(getattr(q, n)) for n, _ in q._fields_] if hasattr(q, "_fields_") else q,
                     dtype=np.float64)
        a = np.concatenate([a, np.zeros(11 - a.size)])
        self.L.oracle_set_sponge(self._h, a.ctypes.data_as(_dp))

    def set_boundary_layer(self, q):
        """q: a ctypes BLParams (or surf_drag, bl_sigma); None: off."""
        if q is None:
            self.L.oracle_set_boundary_layer(self._h, None)
            return
        a = np.array([q.surf_drag, q.bl_sigma] if hasattr(q, "_fields_") else q, dtype=np.float64)
        self.L.oracle_set_boundary_layer(self._h, a.ctypes.data_as(_dp))

    def zonal_table(self):
        """(table [P][3] = bin, points in bin, index in bin; max_count)."""
        t = np.zeros((self.grid.P, 3), dtype=np.int32)
        mc = C.c_int()
        if self.L.oracle_zonal_table(self._h, t.ctypes.data_as(C.POINTER(C.c_int)), t.size, C.byref(mc)) != 0:
            raise ValueError("sponge not set")
        return t, mc.value

    def profx_forcing(self, forcing: int):
        if self.L.oracle_profx(self._h, forcing) != 0:
            raise ValueError(forcing)

    def run(self, n: int, forcing: int):
        if self.L.oracle_run(self._h, n, forcing) != 0:
            raise ValueError(forcing)

    def kernel(self, name: str, a0: int = 0, a1: int = 0):
        """One restated kernel (see oracle_kernel in thor_oracle.h)."""
        if self.L.oracle_kernel(self._h, name.encode(), a0, a1) != 0:
            raise KeyError(name)


REF_LIB = os.path.join(_HERE, "_ref", "libthor_ref.so")


def ref_available() -> bool:
    """True when THOR's own kernels have been compiled into oracle/_ref/."""
    return os.path.exists(REF_LIB)


class RefKernels:
    """THOR's own kernels (compiled from the reference tree by `make -C oracle
    ref`) on device copies of named arrays in THOR's layout; same names and
    kernel arguments as Oracle.kernel. Needs a GPU."""

    def __init__(self, grid, params):
        L = C.CDLL(REF_LIB)
        L.ref_create.restype = C.c_void_p
        L.ref_create.argtypes = [C.POINTER(_Grid), C.POINTER(_Params)]
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_destroy.restype = None
        L.ref_array_size.argtypes = [C.c_void_p, C.c_char_p]
        L.ref_array_size.restype = C.c_long
        for f in (L.ref_set, L.ref_get):
            f.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_long]
            f.restype = C.c_int
        L.ref_kernel.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
        L.ref_kernel.restype = C.c_int
        self.L, self.grid = L, grid
        g, p = _structs(grid, params)
        self._g = g
        h = L.ref_create(C.byref(g), C.byref(p))
        if not h:
            raise RuntimeError("ref_create failed (no GPU or out of memory)")
        self._h = C.c_void_p(h)

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.ref_destroy(self._h)
            self._h = None

    def has(self, name: str) -> bool:
        return self.L.ref_array_size(self._h, name.encode()) >= 0

    def set(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if self.L.ref_set(self._h, name.encode(), a.ctypes.data_as(_dp), a.size) != 0:
            raise KeyError(name)

    def get(self, name: str) -> np.ndarray:
        n = self.L.ref_array_size(self._h, name.encode())
        if n < 0:
            raise KeyError(name)
        a = np.zeros(n)
        if self.L.ref_get(self._h, name.encode(), a.ctypes.data_as(_dp), n) != 0:
            raise RuntimeError(f"ref_get {name}")
        return a

    def kernel(self, name: str, a0: int = 0, a1: int = 0):
        rc = self.L.ref_kernel(self._h, name.encode(), a0, a1)
        if rc != 0:
            raise RuntimeError(f"ref_kernel {name}: {rc}")
