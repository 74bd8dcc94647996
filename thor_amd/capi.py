"""ctypes binding of the C ABI declared in include/thor_dyn.h.

The host-side mirror of THOR's driver objects: ``Grid`` (Icogrid,
src/ESP/grid.cu:85-302), ``Params`` (SimulationSetup,
src/headers/simulation_setup.h:54-119) and ``Dyn`` (the ESP object's dynamics
state and its Thor/ProfX/globdiag methods, src/headers/esp.h:466-470).
Errors raise ``ThorError`` carrying the library's status code and message,
where the reference would log and exit (thor_driver.cu:1277-1286).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libthor_amd.so")
# Parity build of the same sources with portable elementary functions.
LIB_PATH_PM = os.path.join(_HERE, "libthor_amd_pm.so")

THOR_OK = 0
THOR_ERR_BAD_ARG = -1
THOR_ERR_HIP = -2
THOR_ERR_NAN = -3
THOR_ERR_NO_DEVICE = -4
THOR_ERR_IO = -5
THOR_ERR_UNSUPPORTED = -6
THOR_ERR_HOOK = -7

PRECISION_FP64 = 0
PRECISION_MIXED = 1
FORCING_NONE = 0
FORCING_HELD_SUAREZ = 1
FORCING_NEWTONIAN = 2
FORCING_DOUBLE_GRAY = 3
THERMO_ENTROPY = 0
THERMO_ENERGY = 1


class NewtonianParams(C.Structure):
    """thor_newtonian_params: hot-Jupiter day-night Newtonian relaxation."""
    _fields_ = [(n, C.c_double) for n in ("T_mean", "dT", "lon_sub", "p_top", "p_bottom", "p_contrast",
                                          "tau_top", "tau_bottom")]


_fp = C.POINTER(C.c_float)


class SurrogateParams(C.Structure):
    """thor_rt_surrogate_params: normalisation, envelope and judge interval."""
    _fields_ = [("in_mean", _fp), ("in_std", _fp), ("in_lo", _fp), ("in_hi", _fp), ("out_mean", _fp),
                ("out_std", _fp), ("judge_every", C.c_int)]


FORCING_RT_SURROGATE = 4


class RTParams(C.Structure):
    """thor_rt_params: double-gray two-stream radiative transfer."""
    _fields_ = [(n, C.c_double) for n in ("kappa_sw", "kappa_lw", "n_sw", "n_lw", "f_lw", "diff_ang", "Tint",
                                          "albedo", "incflx", "ps0", "lon_sub")]


class SpongeParams(C.Structure):
    """thor_sponge_params: Rayleigh sponge (THOR's RayleighSponge, implicit mode)."""
    _fields_ = [("ns", C.c_double), ("Ruv", C.c_double), ("Rw", C.c_double), ("nlat_bins", C.c_int),
                ("damp_uv_to_mean", C.c_int), ("damp_w_to_mean", C.c_int), ("temp_sponge", C.c_int),
                ("RT_sponge", C.c_double), ("raysp_calc_mode", C.c_int), ("shrink_sponge", C.c_int),
                ("t_shrink", C.c_int)]


class BLParams(C.Structure):
    """thor_bl_params: boundary-layer module, RayleighHS type."""
    _fields_ = [("surf_drag", C.c_double), ("bl_sigma", C.c_double)]


class StepView(C.Structure):
    """thor_step_view: device pointers of the step handed to physics hooks."""
    _fields_ = [("plane", C.c_long), ("point_num", C.c_int), ("nv", C.c_int), ("mhs_fp32", C.c_int)] + \
        [(n, C.c_void_p) for n in ("Mhk", "Whk", "Wk", "Rhok", "pressurek", "Mhs", "Whs", "Rhos", "pressures",
                                    "temperature", "h", "pt", "pth")]


HOOK_INIT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(StepView))
HOOK_MODES = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.POINTER(StepView))


class Hooks(C.Structure):
    """thor_hooks: physics-module callbacks of the step (phy_modules.h:76-86)."""
    _fields_ = [("user", C.c_void_p), ("dyn_core_loop_init", HOOK_INIT), ("dyn_core_loop_slow_modes", HOOK_MODES),
                ("dyn_core_loop_fast_modes", HOOK_MODES), ("dyn_core_loop_end", HOOK_INIT)]


MAX_RK_ARRAYS = 10


class ThorError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"thor status {code}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class GridDesc(C.Structure):
    _fields_ = [
        ("point_num", C.c_int), ("nv", C.c_int), ("nr", C.c_int), ("nl_region", C.c_int),
        ("glevel", C.c_int), ("pent_ind", C.c_int * 12),
        ("point_local", _ip), ("maps", _ip),
        ("div", _dp), ("grad", _dp), ("areas", _dp), ("areasTr", _dp), ("areasT", _dp),
        ("nvecoa", _dp), ("nvecti", _dp), ("nvecte", _dp), ("func_r", _dp), ("lonlat", _dp),
        ("Altitude", _dp), ("Altitudeh", _dp), ("xyz", _dp), ("xyzq", _dp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("A", C.c_double), ("Omega", C.c_double), ("Gravit", C.c_double), ("P_Ref", C.c_double),
        ("Rd", C.c_double), ("Cp", C.c_double), ("Tmean", C.c_double), ("Diffc", C.c_double),
        ("DivDampc", C.c_double), ("timestep", C.c_double),
        ("glevel", C.c_int), ("HyDiff", C.c_int), ("HyDiffOrder", C.c_int), ("DivDampP", C.c_int),
        ("DeepModel", C.c_int), ("NonHydro", C.c_int), ("GravHeightVar", C.c_int),
        ("n_substeps", C.c_int), ("precision", C.c_int),
        ("VertHyDiff", C.c_int), ("VertHyDiffOrder", C.c_int), ("Diffc_v", C.c_double),
        ("DiffSponge", C.c_int), ("order_diff_sponge", C.c_int), ("Dv_sponge", C.c_double),
        ("ns_diff_sponge", C.c_double), ("thermo_equation", C.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class State(C.Structure):
    _fields_ = [("on_device", C.c_int), ("Mh", _dp), ("Wh", _dp), ("W", _dp), ("Rho", _dp),
                ("pressure", _dp)]


class Forcing(C.Structure):
    """thor_forcing: THOR's profx_Qheat, profx_dMh, profx_dWh as device pointers."""
    _fields_ = [("Qheat", _dp), ("dMh", _dp), ("dWh", _dp)]


_libs: dict = {}


def lib(variant: str = "") -> C.CDLL:
    """Loads libthor_amd.so (variant "") or the parity build libthor_amd_pm.so
    (variant "pm"), both built by thor_amd.build. Raises if it is missing."""
    if variant not in ("", "pm"):
        raise ValueError(f"unknown library variant {variant!r} (\"\" or \"pm\")")
    if variant not in _libs:
        path = LIB_PATH_PM if variant == "pm" else LIB_PATH
        if not os.path.exists(path):
            raise ThorError(THOR_ERR_IO, f"{path} not built; run python -m thor_amd.build")
        L = C.CDLL(path)
        L.thor_last_error.restype = C.c_char_p
        L.thor_grid_build.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                      C.c_char_p, C.POINTER(C.c_void_p)]
        L.thor_grid_describe.argtypes = [C.c_void_p, C.POINTER(GridDesc)]
        L.thor_grid_spring_iterations.argtypes = [C.c_void_p]
        L.thor_grid_destroy.argtypes = [C.c_void_p]
        L.thor_grid_destroy.restype = None
        L.thor_params_default_earth.argtypes = [C.POINTER(Params)]
        L.thor_params_default_earth.restype = None
        L.thor_ic_rest.argtypes = [C.POINTER(GridDesc), C.POINTER(Params), C.c_ulonglong, C.c_double,
                                   C.c_double, _dp, _dp, _dp, _dp, _dp]
        if hasattr(L, "thor_dyn_create"):
            L.thor_dyn_create.argtypes = [C.POINTER(GridDesc), C.POINTER(Params), C.POINTER(C.c_void_p)]
            L.thor_dyn_destroy.argtypes = [C.c_void_p]
            L.thor_dyn_destroy.restype = None
            for f in ("thor_dyn_set_state", "thor_dyn_get_state"):
                getattr(L, f).argtypes = [C.c_void_p, C.POINTER(State)]
            L.thor_dyn_thor.argtypes = [C.c_void_p]
            L.thor_dyn_step.argtypes = [C.c_void_p, C.POINTER(State), C.POINTER(Forcing), C.c_double, C.c_int,
                                        C.c_void_p]
            L.thor_dyn_profx.argtypes = [C.c_void_p, C.c_int]
            L.thor_dyn_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
            L.thor_dyn_sync.argtypes = [C.c_void_p]
            L.thor_dyn_globdiag.argtypes = [C.c_void_p, _dp]
            L.thor_dyn_globdiag_noalloc.argtypes = [C.c_void_p, _dp]
            L.thor_dyn_get_field.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_long]
            L.thor_dyn_set_field.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_long]
            L.thor_dyn_kernel_timing.argtypes = [C.c_void_p, C.c_int]
            L.thor_dyn_kernel_stat_count.argtypes = [C.c_void_p]
            L.thor_dyn_kernel_stat.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p),
                                               C.POINTER(C.c_long), _dp, _dp]
            L.thor_dyn_time_run.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
            L.thor_dyn_set_sponge.argtypes = [C.c_void_p, C.POINTER(SpongeParams)]
            L.thor_dyn_set_boundary_layer.argtypes = [C.c_void_p, C.POINTER(BLParams)]
            L.thor_dyn_set_hooks.argtypes = [C.c_void_p, C.POINTER(Hooks)]
            L.thor_tracers_create.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
            L.thor_tracers_destroy.argtypes = [C.c_void_p]
            L.thor_tracers_destroy.restype = None
            for f in (L.thor_tracers_set, L.thor_tracers_get):
                f.argtypes = [C.c_void_p, _dp, C.c_long]
            L.thor_tracers_hooks.argtypes = [C.c_void_p, C.POINTER(Hooks)]
            L.thor_dyn_register_rk_array.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
            L.thor_dyn_clear_rk_arrays.argtypes = [C.c_void_p]
            L.thor_dyn_set_newtonian.argtypes = [C.c_void_p, C.POINTER(NewtonianParams)]
            L.thor_dyn_set_rt.argtypes = [C.c_void_p, C.POINTER(RTParams)]
            L.thor_dyn_get_rt_column.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_long]
            L.thor_dyn_set_rt_surrogate.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(SurrogateParams)]
            L.thor_dyn_surrogate_stats.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long),
                                                   C.POINTER(C.c_long), _dp, _dp]
            L.thor_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]
            L.thor_mlp_destroy.argtypes = [C.c_void_p]
            L.thor_mlp_destroy.restype = None
            L.thor_mlp_param_count.argtypes = [C.c_void_p]
            L.thor_mlp_param_count.restype = C.c_long
            for f in (L.thor_mlp_get_params, L.thor_mlp_set_params):
                f.argtypes = [C.c_void_p, _fp, C.c_long]
            L.thor_mlp_set_data.argtypes = [C.c_void_p, _fp, _fp, C.c_long]
            L.thor_mlp_train_epoch.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_ulonglong, _dp]
            L.thor_mlp_forward.argtypes = [C.c_void_p, _fp, _fp, C.c_long]
            L.thor_mlp_gradients.argtypes = [C.c_void_p, _fp, _fp, C.c_long, _fp, _dp]
            L.thor_mlp_time_forward.argtypes = [C.c_void_p, C.c_long, C.c_int, _dp]
        _libs[variant] = L
    return _libs[variant]


def _check(code: int, variant: str = ""):
    if code != THOR_OK:
        raise ThorError(code, lib(variant).thor_last_error().decode())


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def default_params(**overrides) -> Params:
    p = Params()
    lib().thor_params_default_earth(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Grid:
    """Icosahedral grid + operators in THOR's per-point layout."""

    def __init__(self, glevel: int, nv: int, radius: float, top: float, spring: bool = True,
                 beta: float = 1.15, cache_dir: str | None = None):
        h = C.c_void_p()
        cd = cache_dir.encode() if cache_dir else None
        _check(lib().thor_grid_build(glevel, nv, radius, top, int(spring), beta, cd, C.byref(h)))
        self._h = h
        self.desc = GridDesc()
        _check(lib().thor_grid_describe(h, C.byref(self.desc)))
        d = self.desc
        self.P, self.nv, self.nr, self.glevel = d.point_num, d.nv, d.nr, d.glevel
        self.spring_iterations = lib().thor_grid_spring_iterations(h)
        P = self.P
        shapes = {
            "point_local": (P, 6), "maps": (self.nr, 18, 18), "div": (P, 7, 3), "grad": (P, 7, 3),
            "areas": (P, 6, 3), "areasTr": (P, 6), "areasT": (P,), "nvecoa": (P, 6, 3),
            "nvecti": (P, 6, 3), "nvecte": (P, 6, 3), "func_r": (P, 3), "lonlat": (P, 2),
            "Altitude": (self.nv,), "Altitudeh": (self.nv + 1,), "xyz": (P, 3), "xyzq": (P, 6, 3),
        }
        for name, shp in shapes.items():
            ptr = getattr(d, name)
            n = int(np.prod(shp))
            arr = np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shp)
            setattr(self, name, arr)
        self.pent_ind = np.array(list(d.pent_ind), dtype=np.int64)

    def __del__(self):
        if getattr(self, "_h", None) and "" in _libs:
            _libs[""].thor_grid_destroy(self._h)
            self._h = None


def ic_rest(grid: Grid, params: Params, seed: int = 1000, temperature_noise: float = 0.0,
            wind_noise: float = 0.0) -> dict:
    """Rest initial condition in THOR's layout (host numpy arrays)."""
    P, nv = grid.P, grid.nv
    st = {
        "Mh": np.zeros((P, nv, 3)), "Wh": np.zeros((P, nv + 1)), "W": np.zeros((P, nv)),
        "Rho": np.zeros((P, nv)), "pressure": np.zeros((P, nv)),
    }
    _check(lib().thor_ic_rest(C.byref(grid.desc), C.byref(params), seed, temperature_noise, wind_noise,
                              _ptr(st["Mh"]), _ptr(st["Wh"]), _ptr(st["W"]), _ptr(st["Rho"]),
                              _ptr(st["pressure"])))
    return st


FIELD_SHAPES = {
    "Mh": "v", "W": "s", "Rho": "s", "pressure": "s", "Wh": "i", "temperature": "s", "pt": "s",
    "h": "s", "hh": "i", "pth": "i", "gtil": "s", "gtilh": "i", "profx_Qheat": "s",
    "profx_dMh": "v", "profx_dWh": "i", "Etotal": "s",
    # mixed precision only: operator weights per point (see Geo in dyn_common.h)
    "wlap": 7, "wdd": 63, "udd": 21,
}


class Dyn:
    """Dynamical-core handle on the current HIP device (ESP::Thor drop-in)."""

    def __init__(self, grid: Grid, params: Params, variant: str = ""):
        L = lib(variant)
        self.variant = variant
        self.L = L
        self.grid = grid
        self.params = params
        h = C.c_void_p()
        self._chk(L.thor_dyn_create(C.byref(grid.desc), C.byref(params), C.byref(h)))
        self._h = h

    def _chk(self, code):
        _check(code, self.variant)

    def close(self):
        if getattr(self, "_h", None):
            self.L.thor_dyn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, kind):
        P, nv = self.grid.P, self.grid.nv
        if isinstance(kind, int):
            return (P, kind)
        return {"s": (P, nv), "i": (P, nv + 1), "v": (P, nv, 3)}[kind]

    def set_state(self, st: dict):
        arrs = {k: np.ascontiguousarray(st[k], dtype=np.float64) for k in ("Mh", "Wh", "W", "Rho", "pressure")}
        s = State(0, *(_ptr(arrs[k]) for k in ("Mh", "Wh", "W", "Rho", "pressure")))
        self._chk(self.L.thor_dyn_set_state(self._h, C.byref(s)))

    def get_state(self) -> dict:
        out = {k: np.zeros(self._shape(FIELD_SHAPES[k])) for k in ("Mh", "Wh", "W", "Rho", "pressure")}
        s = State(0, *(_ptr(out[k]) for k in ("Mh", "Wh", "W", "Rho", "pressure")))
        self._chk(self.L.thor_dyn_get_state(self._h, C.byref(s)))
        return out

    def get_field(self, name: str) -> np.ndarray:
        a = np.zeros(self._shape(FIELD_SHAPES[name]))
        self._chk(self.L.thor_dyn_get_field(self._h, name.encode(), _ptr(a), a.size))
        return a

    def set_field(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.L.thor_dyn_set_field(self._h, name.encode(), _ptr(a), a.size))

    def thor(self):
        self._chk(self.L.thor_dyn_thor(self._h))

    def step(self, state: dict, forcing: "dict | None" = None, dt: float = 0.0, n_substeps: int = 0,
             stream: int = 0):
        """thor_dyn_step on caller-owned device arrays: state and forcing map
        the field names to device addresses (THOR layout); the new state
        overwrites the state arrays in stream order."""
        cast = lambda a: C.cast(C.c_void_p(a), _dp)
        s = State(1, *(cast(state[k]) for k in ("Mh", "Wh", "W", "Rho", "pressure")))
        f = Forcing(*(cast(forcing[k]) for k in ("Qheat", "dMh", "dWh"))) if forcing else None
        self._chk(self.L.thor_dyn_step(self._h, C.byref(s), C.byref(f) if f else None, dt, n_substeps,
                                       C.c_void_p(stream) if stream else None))

    def set_newtonian(self, q: NewtonianParams):
        self._chk(self.L.thor_dyn_set_newtonian(self._h, C.byref(q)))

    def set_rt(self, q: RTParams):
        self._chk(self.L.thor_dyn_set_rt(self._h, C.byref(q)))

    def rt_column(self, name: str) -> np.ndarray:
        a = np.zeros(self.grid.P)
        self._chk(self.L.thor_dyn_get_rt_column(self._h, name.encode(), _ptr(a), a.size))
        return a

    def set_rt_surrogate(self, mlp: "MLP", norm: dict, judge_every: int = 10):
        """norm: float32 arrays in_mean, in_std, in_lo, in_hi [2 nv + 1], out_mean, out_std [nv]."""
        self._sur_norm = {k: np.ascontiguousarray(norm[k], dtype=np.float32)
                          for k in ("in_mean", "in_std", "in_lo", "in_hi", "out_mean", "out_std")}
        q = SurrogateParams(*(self._sur_norm[k].ctypes.data_as(_fp) for k in
                              ("in_mean", "in_std", "in_lo", "in_hi", "out_mean", "out_std")), judge_every)
        self._sur_mlp = mlp  # keeps the network alive while attached
        self._chk(self.L.thor_dyn_set_rt_surrogate(self._h, mlp._h, C.byref(q)))

    def surrogate_stats(self) -> dict:
        calls, judged, fb = C.c_long(), C.c_long(), C.c_long()
        rms, mx = C.c_double(), C.c_double()
        self._chk(self.L.thor_dyn_surrogate_stats(self._h, C.byref(calls), C.byref(judged), C.byref(fb),
                                                  C.byref(rms), C.byref(mx)))
        return {"calls": calls.value, "judged": judged.value, "fallback_columns": fb.value,
                "rms_err": rms.value, "max_err": mx.value}

    def profx(self, forcing=FORCING_HELD_SUAREZ):
        self._chk(self.L.thor_dyn_profx(self._h, forcing))

    def run(self, nsteps: int, forcing=FORCING_HELD_SUAREZ):
        self._chk(self.L.thor_dyn_run(self._h, nsteps, forcing))

    def sync(self):
        self._chk(self.L.thor_dyn_sync(self._h))

    def globdiag(self) -> np.ndarray:
        out = np.zeros(5)
        self._chk(self.L.thor_dyn_globdiag(self._h, _ptr(out)))
        return out

    def globdiag_noalloc(self) -> np.ndarray:
        """globdiag through the allocation-free entry point (device scratch)."""
        out = np.zeros(5)
        self._chk(self.L.thor_dyn_globdiag_noalloc(self._h, _ptr(out)))
        return out

    def kernel_timing(self, enable: bool):
        self._chk(self.L.thor_dyn_kernel_timing(self._h, int(enable)))

    def kernel_stats(self) -> list:
        L = self.L
        res = []
        for i in range(L.thor_dyn_kernel_stat_count(self._h)):
            name = C.c_char_p()
            n = C.c_long()
            ms = C.c_double()
            by = C.c_double()
            self._chk(L.thor_dyn_kernel_stat(self._h, i, C.byref(name), C.byref(n), C.byref(ms), C.byref(by)))
            res.append({"name": name.value.decode(), "launches": n.value, "total_ms": ms.value,
                        "bytes_per_launch": by.value})
        return res

    def time_run(self, nsteps: int, forcing=FORCING_HELD_SUAREZ) -> float:
        ms = C.c_double()
        self._chk(self.L.thor_dyn_time_run(self._h, nsteps, forcing, C.byref(ms)))
        return ms.value

    def set_sponge(self, q: "SpongeParams | None"):
        self._chk(self.L.thor_dyn_set_sponge(self._h, C.byref(q) if q is not None else None))

    def set_boundary_layer(self, q: "BLParams | None"):
        """THOR's boundary_layer module (RayleighHS) in every ProfX; None: off."""
        self._chk(self.L.thor_dyn_set_boundary_layer(self._h, C.byref(q) if q is not None else None))

    def set_hooks(self, init=None, slow_modes=None, fast_modes=None, end=None):
        """Installs Python physics hooks: init(stream, nstep, view), slow_modes /
        fast_modes(stream, nstep, times, view), end(stream, nstep, view), where
        stream is the handle's hipStream_t as an int and view a StepView. A
        hook returning a true value (or raising) stops the step. None: no hook."""
        def wrap(fn, typ, modes):
            if fn is None:
                return typ()

            def cb(user, stream, nstep, *rest):
                try:
                    if modes:
                        return 1 if fn(stream or 0, nstep, rest[0], rest[1].contents) else 0
                    return 1 if fn(stream or 0, nstep, rest[0].contents) else 0
                except Exception:
                    import traceback
                    traceback.print_exc()
                    return 1
            return typ(cb)
        hk = Hooks(None, wrap(init, HOOK_INIT, False), wrap(slow_modes, HOOK_MODES, True),
                   wrap(fast_modes, HOOK_MODES, True), wrap(end, HOOK_INIT, False))
        self._hooks = hk  # the callbacks must outlive the handle's use of them
        self._chk(self.L.thor_dyn_set_hooks(self._h, C.byref(hk)))

    def clear_hooks(self):
        self._chk(self.L.thor_dyn_set_hooks(self._h, None))
        self._hooks = None

    def register_rk_array(self, array_d: int, arrayk_d: int, arrayi_d: int, dimensions: int):
        """Registers device arrays (addresses, THOR layout [(p*nv + lev)*dims + i])
        that the RK scheme advances like the state."""
        self._chk(self.L.thor_dyn_register_rk_array(self._h, array_d, arrayk_d, arrayi_d, dimensions))

    def clear_rk_arrays(self):
        self._chk(self.L.thor_dyn_clear_rk_arrays(self._h))


class Tracers:
    """Passive tracers of THOR's chemistry module (thor_tracers_*): created on a
    dynamical core, whose hooks it then occupies."""

    def __init__(self, dyn: "Dyn", ntr: int):
        self.dyn, self.ntr, self.L = dyn, ntr, dyn.L
        h = C.c_void_p()
        dyn._chk(self.L.thor_tracers_create(dyn._h, ntr, C.byref(h)))
        self._h = h
        self._hooks = Hooks()
        dyn._chk(self.L.thor_tracers_hooks(h, C.byref(self._hooks)))
        dyn._chk(self.L.thor_dyn_set_hooks(dyn._h, C.byref(self._hooks)))

    def set(self, tracer: np.ndarray):
        """tracer mass [P][nv][ntr] (THOR's tracer_d layout)"""
        a = np.ascontiguousarray(tracer, dtype=np.float64)
        self.dyn._chk(self.L.thor_tracers_set(self._h, _ptr(a), a.size))

    def get(self) -> np.ndarray:
        g = self.dyn.grid
        a = np.zeros((g.P, g.nv, self.ntr))
        self.dyn._chk(self.L.thor_tracers_get(self._h, _ptr(a), a.size))
        return a

    def close(self):
        if getattr(self, "_h", None):
            self.dyn._chk(self.L.thor_dyn_set_hooks(self.dyn._h, None))
            self.L.thor_tracers_destroy(self._h)
            self._h = None


class MLP:
    """Per-column surrogate network (thor_mlp.h): n_in -> n_hidden GELU ->
    n_hidden GELU -> n_out, BF16 MFMA kernels with FP32 accumulation."""

    def __init__(self, n_in: int, n_hidden: int, n_out: int, seed: int = 0, variant: str = ""):
        self.L = lib(variant)
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        h = C.c_void_p()
        _check(self.L.thor_mlp_create(n_in, n_hidden, n_out, seed, C.byref(h)), variant)
        self._h = h
        self.variant = variant

    def _chk(self, code):
        _check(code, self.variant)

    def close(self):
        if getattr(self, "_h", None):
            self.L.thor_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def n_params(self) -> int:
        return self.L.thor_mlp_param_count(self._h)

    def params(self) -> np.ndarray:
        a = np.zeros(self.n_params, dtype=np.float32)
        self._chk(self.L.thor_mlp_get_params(self._h, a.ctypes.data_as(_fp), a.size))
        return a

    def set_params(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self._chk(self.L.thor_mlp_set_params(self._h, a.ctypes.data_as(_fp), a.size))

    def set_data(self, X: np.ndarray, Y: np.ndarray):
        self._X = np.ascontiguousarray(X, dtype=np.float32)
        self._Y = np.ascontiguousarray(Y, dtype=np.float32)
        assert self._X.shape[1] == self.n_in and self._Y.shape[1] == self.n_out and len(self._X) == len(self._Y)
        self._chk(self.L.thor_mlp_set_data(self._h, self._X.ctypes.data_as(_fp), self._Y.ctypes.data_as(_fp),
                                           len(self._X)))

    def train_epoch(self, batch: int, lr: float, seed: int) -> float:
        loss = C.c_double()
        self._chk(self.L.thor_mlp_train_epoch(self._h, batch, lr, seed, C.byref(loss)))
        return loss.value

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), self.n_out), dtype=np.float32)
        self._chk(self.L.thor_mlp_forward(self._h, X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), len(X)))
        return Y

    def gradients(self, X: np.ndarray, Y: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        self._chk(self.L.thor_mlp_gradients(self._h, X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), len(X),
                                            g.ctypes.data_as(_fp), C.byref(loss)))
        return g, loss.value

    def time_forward(self, n: int, repeats: int = 20) -> float:
        ms = C.c_double()
        self._chk(self.L.thor_mlp_time_forward(self._h, n, repeats, C.byref(ms)))
        return ms.value
