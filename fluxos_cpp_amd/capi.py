"""ctypes binding of libfluxos_hip.so (the C ABI of include/fluxos_hip.h).

This is the Python face of the library for tests, benchmarks and scripting;
the C++ driver (``fluxos_hip``) uses the same ABI.  Loading fails loudly when
the shared library is missing -- there is no CPU fallback.

Host arrays use the reference's layout: Fortran-ordered numpy arrays of shape
(NROWS+2, NCOLS+2) (``arma::Mat`` column-major memory), float64 for state
grids and float32 for flag grids.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBPATH = PKG / "libfluxos_hip.so"

MAX_CHEM = 8
FX_OK = 0
FX_ERRORS = {-1: "FX_ERR_ARG", -2: "FX_ERR_HIP", -3: "FX_ERR_STATE", -4: "FX_ERR_UNSUPPORTED",
             -5: "FX_ERR_NONFINITE", -6: "FX_ERR_IO", -7: "FX_ERR_NOMEM", -8: "FX_ERR_NOGPU"}
FP64_PARITY, FP32_MIXED = 0, 1
ADE_OFF, ADE_EXACT, ADE_FAST = 0, 1, 2

F_Z, F_H, F_QX, F_QY, F_US, F_LDRY, F_LDRY_PREV = 0, 1, 2, 3, 4, 5, 6
F_FE1, F_FE2, F_FE3, F_FN1, F_FN2, F_FN3, F_TWET, F_H0, F_ZB = 7, 8, 9, 10, 11, 12, 13, 14, 15
F_SOIL = 16
F_CONC0 = 32
FIELDS = {"z": F_Z, "h": F_H, "qx": F_QX, "qy": F_QY, "us": F_US, "ldry": F_LDRY,
          "ldry_prev": F_LDRY_PREV, "fe1": F_FE1, "fe2": F_FE2, "fe3": F_FE3, "fn1": F_FN1,
          "fn2": F_FN2, "fn3": F_FN3, "twet": F_TWET, "h0": F_H0, "zb": F_ZB, "soil": F_SOIL}
FLOAT_FIELDS = {F_LDRY, F_LDRY_PREV}

EXPORTS = ("fx_domain_create", "fx_domain_destroy", "fx_domain_bytes", "fx_set_static",
           "fx_set_field", "fx_get_field", "fx_set_clock", "fx_get_clock", "fx_set_meteo",
           "fx_set_inflow", "fx_courant_condition", "fx_add_meteo", "fx_add_inflow",
           "fx_hydrodynamics_calc", "fx_adesolver_calc", "fx_advance", "fx_mass_balance",
           "fx_synchronize", "fx_read_geo", "fx_initiation", "fx_last_error", "fx_version",
           "fx_profile", "fx_profile_read", "fx_batch_create", "fx_batch_destroy", "fx_batch_admit",
           "fx_batch_resume", "fx_batch_evict", "fx_batch_run", "fx_batch_poll", "fx_batch_info",
           "fx_read_dem", "fx_read_forcing", "fx_write_results", "fx_set_tiling", "fx_get_tiling",
           "fx_set_wintra", "fx_wintrasolver_calc", "fx_initiation_restart", "fx_read_results",
           "fx_find_last_step", "fx_solute_balance", "fx_mass_balance_noalloc")
EV_PRINT, EV_DONE, EV_ERROR = 1, 2, 3
K_CFL, K_MELT, K_FLUX, K_UPDATE, K_ADE_PASS1, K_ADE_SWEEP, K_STEP32, K_ADE32, K_CFL32 = range(9)
KERNEL_NAMES = {K_CFL: "k64_cfl", K_MELT: "k64_melt", K_FLUX: "k64_flux", K_UPDATE: "k64_update",
                K_ADE_PASS1: "k_ade_pass1", K_ADE_SWEEP: "k_ade_sweep", K_STEP32: "k32_step",
                K_ADE32: "k32_ade", K_CFL32: "k32_cfl"}


class FxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{FX_ERRORS.get(code, code)}: {msg}")
        self.code = code


class DomainDesc(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("dxy", C.c_int32), ("nchem", C.c_int32),
                ("hdry", C.c_double), ("nodata", C.c_double),
                ("gacc", C.c_float), ("cfl", C.c_float), ("cvdef", C.c_float), ("nuem", C.c_float),
                ("d_coef", C.c_double * MAX_CHEM), ("ade_solver", C.c_int32), ("precision", C.c_int32),
                ("ade_mode", C.c_int32), ("device", C.c_int32)]


class BatchEvent(C.Structure):
    _fields_ = [("slot", C.c_int32), ("kind", C.c_int32), ("tim", C.c_double), ("it", C.c_int32),
                ("steps", C.c_int32), ("solver_steps", C.c_int32), ("pad", C.c_int32)]


class DemHeader(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("cellsize", C.c_int32), ("pad", C.c_int32),
                ("xllcorner", C.c_double), ("yllcorner", C.c_double), ("nodata", C.c_double)]


class ResultsGrids(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("dxy", C.c_int32), ("pad", C.c_int32),
                ("xllcorner", C.c_double), ("yllcorner", C.c_double), ("h_min_print", C.c_double)] + \
               [(k, C.POINTER(C.c_double)) for k in ("zb", "z", "h", "qx", "qy", "us", "conc0", "soil_mass",
                                                      "fe1", "fn1", "twet", "ux", "uy")]


INIT_GRIDS = ("zb", "z", "h", "qx", "qy", "us", "ux", "uy", "conc0", "soil_mass", "twet")


class InitGrids(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_double)) for k in INIT_GRIDS] + [("ldry", C.POINTER(C.c_float))]


class StepInfo(C.Structure):
    _fields_ = [("tim", C.c_double), ("dtfl", C.c_double), ("hpall", C.c_double),
                ("print_next", C.c_double), ("it", C.c_int32), ("steps", C.c_int32),
                ("solver_steps", C.c_int32), ("status", C.c_int32)]


_lib: C.CDLL | None = None
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)


def load(path: Path | None = None) -> C.CDLL:
    """Load libfluxos_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(path or LIBPATH)
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -m fluxos_cpp_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(str(path))
    vp = C.c_void_p
    sig = {
        "fx_domain_create": ([C.POINTER(DomainDesc), C.POINTER(vp)], C.c_int),
        "fx_domain_destroy": ([vp], C.c_int),
        "fx_domain_bytes": ([C.POINTER(DomainDesc)], C.c_int64),
        "fx_set_static": ([vp, _D, _D, _F], C.c_int),
        "fx_set_field": ([vp, C.c_int32, vp], C.c_int),
        "fx_get_field": ([vp, C.c_int32, vp], C.c_int),
        "fx_set_clock": ([vp, C.c_double, C.c_double, C.c_int32], C.c_int),
        "fx_get_clock": ([vp, _D, _D, C.POINTER(C.c_int32)], C.c_int),
        "fx_set_meteo": ([vp, C.c_int32, C.c_int32, _F, _F], C.c_int),
        "fx_set_inflow": ([vp, C.c_int32, C.c_int32, _F, C.c_int64, C.c_int64], C.c_int),
        "fx_courant_condition": ([vp, C.c_double, _D, _D], C.c_int),
        "fx_add_meteo": ([vp], C.c_int),
        "fx_add_inflow": ([vp], C.c_int),
        "fx_hydrodynamics_calc": ([vp], C.c_int),
        "fx_adesolver_calc": ([vp, C.c_int32, C.c_int32], C.c_int),
        "fx_advance": ([vp, C.c_double, C.c_double, C.c_int32, C.POINTER(StepInfo)], C.c_int),
        "fx_mass_balance": ([vp, _D], C.c_int),
        "fx_mass_balance_noalloc": ([vp, _D], C.c_int),
        "fx_solute_balance": ([vp, C.c_int32, _D], C.c_int),
        "fx_set_tiling": ([vp, C.c_int32, C.c_int32], C.c_int),
        "fx_get_tiling": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "fx_synchronize": ([vp], C.c_int),
        "fx_read_geo": ([C.c_int32, C.c_int32, C.c_double, _F, C.c_double, _D, _D, _F], C.c_int),
        "fx_initiation": ([C.c_int32, C.c_int32, C.c_int32, _D, _D, _D, _F, _D], C.c_int),
        "fx_profile": ([vp, C.c_int32], C.c_int),
        "fx_profile_read": ([vp, C.c_int32, _D, C.POINTER(C.c_int64)], C.c_int),
        "fx_batch_create": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)], C.c_int),
        "fx_batch_destroy": ([vp], C.c_int),
        "fx_batch_admit": ([vp, vp, C.c_double, C.c_double, C.POINTER(C.c_int32)], C.c_int),
        "fx_batch_resume": ([vp, C.c_int32, C.c_double], C.c_int),
        "fx_batch_evict": ([vp, C.c_int32], C.c_int),
        "fx_batch_run": ([vp, C.c_int32], C.c_int),
        "fx_batch_poll": ([vp, C.POINTER(BatchEvent), C.c_int32, C.POINTER(C.c_int32)], C.c_int),
        "fx_batch_info": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
        "fx_read_dem": ([C.c_char_p, C.POINTER(DemHeader), _F], C.c_int),
        "fx_read_forcing": ([C.c_char_p, C.c_int32, C.c_int32, _F, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
                            C.c_int),
        "fx_write_results": ([C.c_char_p, C.POINTER(ResultsGrids)], C.c_int),
        "fx_set_wintra": ([vp, C.c_int32, C.c_double], C.c_int),
        "fx_wintrasolver_calc": ([vp, C.c_int32], C.c_int),
        "fx_initiation_restart": ([C.c_int32, C.c_int32, C.c_int32, C.c_double, _D, C.c_int64,
                                   C.POINTER(InitGrids)], C.c_int),
        "fx_read_results": ([C.c_char_p, C.c_int64, _D, C.POINTER(C.c_int64)], C.c_int),
        "fx_find_last_step": ([C.c_char_p, C.POINTER(C.c_int32)], C.c_int),
        "fx_last_error": ([], C.c_char_p),
        "fx_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != FX_OK:
        raise FxError(rc, load().fx_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F)


# ---------------------------------------------------------------------------
# host input layer
# ---------------------------------------------------------------------------
def read_geo(dem: np.ndarray, nodata: float, ks_input: float):
    """read_functions.cpp:145-225 on a parsed DEM (file order)."""
    lib = load()
    nr, nc = dem.shape
    shape = (nr + 2, nc + 2)
    zb = np.zeros(shape, order="F")
    ks = np.zeros(shape, order="F")
    weir = np.zeros(shape, dtype=np.float32, order="F")
    dem = np.ascontiguousarray(dem, dtype=np.float32)
    _check(lib.fx_read_geo(nr, nc, float(nodata), _fp(dem), float(ks_input), _dp(zb), _dp(ks), _fp(weir)))
    return zb, ks, weir


def initiation(zb: np.ndarray, ade_solver: bool = True):
    """initiate.cpp:27-231 cold start: returns (zb, z, h, ldry, conc0)."""
    lib = load()
    nr, nc = zb.shape[0] - 2, zb.shape[1] - 2
    zb = np.array(zb, order="F", dtype=np.float64, copy=True)
    z = np.zeros_like(zb, order="F")
    h = np.zeros_like(zb, order="F")
    ldry = np.zeros(zb.shape, dtype=np.float32, order="F")
    c0 = np.zeros_like(zb, order="F")
    _check(lib.fx_initiation(nr, nc, 1 if ade_solver else 0, _dp(zb), _dp(z), _dp(h), _fp(ldry), _dp(c0)))
    return zb, z, h, ldry, c0


def initiation_restart(zb: np.ndarray, table: np.ndarray, ade_solver: bool = True,
                       soil_conc_bckgrd: float = 0.0) -> dict:
    """initiate.cpp:27-231 with RESTART true: zb from read_geo and a Results
    file's data rows; returns every initial grid by name (zb included)."""
    lib = load()
    nr, nc = zb.shape[0] - 2, zb.shape[1] - 2
    out = {k: np.zeros(zb.shape, order="F") for k in INIT_GRIDS}
    out["zb"][...] = zb
    out["ldry"] = np.zeros(zb.shape, dtype=np.float32, order="F")
    g = InitGrids(*[_dp(out[k]) for k in INIT_GRIDS], _fp(out["ldry"]))
    t = np.ascontiguousarray(table, dtype=np.float64).reshape(-1, 16)
    _check(lib.fx_initiation_restart(nr, nc, 1 if ade_solver else 0, float(soil_conc_bckgrd), _dp(t),
                                     t.shape[0], C.byref(g)))
    return out


def read_results_file(path) -> np.ndarray:
    """The data rows of a Results/<t>.txt file (n x 16 float64)."""
    lib = load()
    n = C.c_int64()
    _check(lib.fx_read_results(str(path).encode(), 0, None, C.byref(n)))
    tab = np.zeros((n.value, 16))
    _check(lib.fx_read_results(str(path).encode(), n.value, _dp(tab), C.byref(n)))
    return tab


def find_last_step(folder) -> int:
    """findLastStep (common.cpp:69-106) on a Results folder."""
    t = C.c_int32()
    _check(load().fx_find_last_step(str(folder).encode(), C.byref(t)))
    return t.value


@dataclass
class InitialState:
    zb: np.ndarray
    ks: np.ndarray
    weir: np.ndarray
    z: np.ndarray
    h: np.ndarray
    ldry: np.ndarray
    conc0: np.ndarray
    melt_mult: np.ndarray | None


def read_dem_file(path) -> tuple[DemHeader, np.ndarray]:
    """ESRI-ASCII DEM: header and values in file order (float32)."""
    lib = load()
    h = DemHeader()
    _check(lib.fx_read_dem(str(path).encode(), C.byref(h), None))
    dem = np.zeros((h.nrows, h.ncols), dtype=np.float32)
    _check(lib.fx_read_dem(str(path).encode(), C.byref(h), _fp(dem)))
    return h, dem


def read_forcing_file(path, ncols: int | None = None) -> np.ndarray:
    """Rows of a forcing CSV after its header line (float32)."""
    lib = load()
    n, w = C.c_int32(), C.c_int32()
    _check(lib.fx_read_forcing(str(path).encode(), 0, 0, None, C.byref(n), C.byref(w)))
    cols = ncols or w.value
    tab = np.zeros((n.value, cols), dtype=np.float32)
    _check(lib.fx_read_forcing(str(path).encode(), n.value, cols, _fp(tab), C.byref(n), C.byref(w)))
    return tab


def write_results_file(path, dxy: int, xll: float, yll: float, h_min_print: float, zb, z, h, qx=None,
                       qy=None, us=None, conc0=None, soil_mass=None, fe1=None, fn1=None, twet=None,
                       ux=None, uy=None) -> None:
    """Results/<t>.txt from padded Fortran-ordered float64 grids."""
    keep = [None if a is None else np.asfortranarray(a, dtype=np.float64)
            for a in (zb, z, h, qx, qy, us, conc0, soil_mass, fe1, fn1, twet, ux, uy)]
    ptr = [None if a is None else _dp(a) for a in keep]
    nr, nc = zb.shape[0] - 2, zb.shape[1] - 2
    g = ResultsGrids(nr, nc, int(dxy), 0, float(xll), float(yll), float(h_min_print), *ptr)
    _check(load().fx_write_results(str(path).encode(), C.byref(g)))


def prepare(sc, nodata: float = -99999.0, ade: bool = True) -> InitialState:
    """The grids a run starts from (read_geo + initiation) for a gen.Scenario."""
    zb, ks, weir = read_geo(sc.dem, nodata, sc.hdry)
    zb, z, h, ldry, c0 = initiation(zb, ade)
    mult = None
    if sc.melt_mult is not None:
        mult = np.zeros(zb.shape, dtype=np.float32, order="F")
        mult[1:-1, 1:-1] = np.flipud(sc.melt_mult)
    return InitialState(zb, ks, weir, z, h, ldry, c0, mult)


# ---------------------------------------------------------------------------
# device domain
# ---------------------------------------------------------------------------
class HipDomain:
    """One FLUXOS domain resident on a GPU."""

    def __init__(self, nrows: int, ncols: int, dxy: int, hdry: float, nchem: int = 1,
                 d_coef=(0.01,), ade: bool = True, precision: int = FP64_PARITY,
                 ade_mode: int | None = None, nodata: float = -99999.0, device: int = 0):
        self.lib = load()
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.shape = (self.nrows + 2, self.ncols + 2)
        self.nchem = int(nchem)
        self.precision = precision
        self.wintra = False
        d = DomainDesc()
        d.nrows, d.ncols, d.dxy, d.nchem = self.nrows, self.ncols, int(dxy), self.nchem
        d.hdry, d.nodata = float(hdry), float(nodata)
        d.gacc, d.cfl, d.cvdef, d.nuem = 9.80665, 1.0, 0.07, 1.793e-6
        for k, v in enumerate(d_coef):
            d.d_coef[k] = float(v)
        d.ade_solver = 1 if ade else 0
        d.precision = precision
        if ade_mode is None:
            ade_mode = (ADE_EXACT if precision == FP64_PARITY else ADE_FAST) if ade else ADE_OFF
        d.ade_mode = ade_mode
        d.device = device
        self.desc = d
        h = C.c_void_p()
        _check(self.lib.fx_domain_create(C.byref(d), C.byref(h)))
        self.h = h

    def close(self) -> None:
        """Free the domain; a batch member is refused (FX_ERR_STATE) and kept."""
        if self.h:
            _check(self.lib.fx_domain_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return int(self.lib.fx_domain_bytes(C.byref(self.desc)))

    # -- transfer -------------------------------------------------------------------
    def set_static(self, zb, ks, weir) -> None:
        zb = np.asfortranarray(zb, dtype=np.float64)
        ks = np.asfortranarray(ks, dtype=np.float64)
        weir = np.asfortranarray(weir, dtype=np.float32)
        _check(self.lib.fx_set_static(self.h, _dp(zb), _dp(ks), _fp(weir)))

    def set(self, name: str | int, arr: np.ndarray) -> None:
        f = FIELDS[name] if isinstance(name, str) else name
        dt = np.float32 if f in FLOAT_FIELDS else np.float64
        a = np.asfortranarray(arr, dtype=dt)
        assert a.shape == self.shape
        _check(self.lib.fx_set_field(self.h, f, a.ctypes.data_as(C.c_void_p)))

    def get(self, name: str | int) -> np.ndarray:
        f = FIELDS[name] if isinstance(name, str) else name
        dt = np.float32 if f in FLOAT_FIELDS else np.float64
        a = np.zeros(self.shape, dtype=dt, order="F")
        _check(self.lib.fx_get_field(self.h, f, a.ctypes.data_as(C.c_void_p)))
        return a

    def set_conc(self, k: int, arr: np.ndarray) -> None:
        self.set(F_CONC0 + k, arr)

    def get_conc(self, k: int) -> np.ndarray:
        return self.get(F_CONC0 + k)

    def set_clock(self, tim: float, dtfl: float = 0.0, it: int = 0) -> None:
        _check(self.lib.fx_set_clock(self.h, float(tim), float(dtfl), int(it)))

    def clock(self):
        t, dt, it = C.c_double(), C.c_double(), C.c_int32()
        _check(self.lib.fx_get_clock(self.h, C.byref(t), C.byref(dt), C.byref(it)))
        return t.value, dt.value, it.value

    def set_meteo(self, table, melt_mult=None) -> None:
        if table is None:
            _check(self.lib.fx_set_meteo(self.h, 0, 0, _F(), _F()))
            return
        t = np.ascontiguousarray(table, dtype=np.float32)
        m = None if melt_mult is None else np.asfortranarray(melt_mult, dtype=np.float32)
        _check(self.lib.fx_set_meteo(self.h, t.shape[0], t.shape[1], _fp(t), _fp(m) if m is not None else _F()))

    def set_inflow(self, table, irow: int, icol: int) -> None:
        t = np.ascontiguousarray(table, dtype=np.float32)
        _check(self.lib.fx_set_inflow(self.h, t.shape[0], t.shape[1], _fp(t), int(irow), int(icol)))

    # -- granular seams ------------------------------------------------------------------
    def courant(self, print_next: float):
        dt, hp = C.c_double(), C.c_double()
        _check(self.lib.fx_courant_condition(self.h, float(print_next), C.byref(dt), C.byref(hp)))
        return dt.value, hp.value

    def add_meteo(self) -> None:
        _check(self.lib.fx_add_meteo(self.h))

    def add_inflow(self) -> None:
        _check(self.lib.fx_add_inflow(self.h))

    def hydrodynamics(self) -> None:
        _check(self.lib.fx_hydrodynamics_calc(self.h))

    def adesolver(self, it: int, ichem: int) -> None:
        _check(self.lib.fx_adesolver_calc(self.h, int(it), int(ichem)))

    # -- device-resident loop -----------------------------------------------------------------
    def advance(self, t_end: float, print_next: float, max_steps: int = 2 ** 30) -> StepInfo:
        info = StepInfo()
        _check(self.lib.fx_advance(self.h, float(t_end), float(print_next), int(max_steps), C.byref(info)))
        return info

    def mass_balance(self):
        out = (C.c_double * 3)()
        _check(self.lib.fx_mass_balance(self.h, out))
        return out[0], out[1], out[2]

    def mass_balance_noalloc(self):
        """mass_balance() computed in the library's fixed device array, with no
        allocation (bit-identical results)."""
        out = (C.c_double * 3)()
        _check(self.lib.fx_mass_balance_noalloc(self.h, out))
        return out[0], out[1], out[2]

    def solute_balance(self, k: int = 0):
        """(solute mass in the non-weir cells, solute brought in by melt and inflow), FP64."""
        out = (C.c_double * 2)()
        _check(self.lib.fx_solute_balance(self.h, int(k), out))
        return out[0], out[1]

    def profile(self, on: bool, every: int = 1) -> None:
        """Kernel timing with HIP events on the domain's stream: every launch,
        or one launch in `every` of each kernel class."""
        _check(self.lib.fx_profile(self.h, max(1, int(every)) if on else 0))

    def profile_read(self, kernel: int):
        ms, n = C.c_double(), C.c_int64()
        _check(self.lib.fx_profile_read(self.h, int(kernel), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_tiling(self, ch: int, ach: int) -> None:
        """Column-block lengths of the FP32 sweeps (results do not depend on it)."""
        _check(self.lib.fx_set_tiling(self.h, ch, ach))

    def tiling(self) -> tuple[int, int]:
        ch, ach = C.c_int32(), C.c_int32()
        _check(self.lib.fx_get_tiling(self.h, C.byref(ch), C.byref(ach)))
        return ch.value, ach.value

    def synchronize(self) -> None:
        _check(self.lib.fx_synchronize(self.h))

    # -- convenience -------------------------------------------------------------------------------
    @classmethod
    def from_scenario(cls, sc, precision: int = FP64_PARITY, ade: bool = True, ade_mode=None,
                      init: InitialState | None = None, device: int = 0, tiling="auto") -> "HipDomain":
        """A device domain for a gen.Scenario.  FP32-mixed domains take the
        tile shape the autotuner recorded for the nearest grid size
        (tiling="auto"), an explicit (ch, ach), or the defaults (None)."""
        st = init or prepare(sc, ade=ade)
        d = cls(sc.nrows, sc.ncols, sc.dxy, sc.hdry, sc.nchem, sc.d_coef, ade, precision, ade_mode,
                device=device)
        d.set_static(st.zb, st.ks, st.weir)
        d.set("z", st.z)
        d.set("h", st.h)
        d.set("ldry", st.ldry)
        d.set_conc(0, st.conc0)
        d.set_meteo(sc.meteo, st.melt_mult)
        d.set_clock(0.0, 0.0, 0)
        if precision == FP32_MIXED and tiling is not None:
            if tiling == "auto":
                from . import autotune
                tiling = autotune.choose(sc.nrows, sc.ncols)
            if tiling:
                d.set_tiling(*tiling)
        return d

    SNAPSHOT_FIELDS = ("z", "h", "qx", "qy", "fe1", "fe2", "fe3", "fn1", "fn2", "fn3", "twet", "ldry",
                       "ldry_prev", "h0")

    def set_wintra(self, on: bool, soil_release_rate: float) -> None:
        """EXTERNAL_MODULES.WINTRA.STATUS and SOIL_RELEASE_RATE [1/h]."""
        _check(self.lib.fx_set_wintra(self.h, 1 if on else 0, float(soil_release_rate)))
        self.wintra = bool(on)

    def wintrasolver_calc(self, ichem: int = 0) -> None:
        _check(self.lib.fx_wintrasolver_calc(self.h, int(ichem)))

    def snapshot(self, solutes: bool = True) -> dict:
        """Every grid a run restarts from (reference layout) and the clock."""
        out = {k: self.get(k) for k in self.SNAPSHOT_FIELDS}
        if getattr(self, "wintra", False):
            out["soil"] = self.get("soil")
        out["conc"] = [self.get_conc(k) for k in range(self.nchem)] if solutes else []
        out["clock"] = self.clock()
        return out

    def restore(self, snap: dict) -> None:
        for k in self.SNAPSHOT_FIELDS + ("soil",):
            if k in snap:
                self.set(k, snap[k])
        for k, c in enumerate(snap.get("conc", [])):
            self.set_conc(k, c)
        t, dt, it = snap["clock"]
        self.set_clock(t, dt, it)

    def state(self) -> dict:
        out = {k: self.get(k) for k in ("z", "h", "qx", "qy", "ldry")}
        for k in range(self.nchem):
            out[f"conc{k}"] = self.get_conc(k)
        return out


def run_loop(dom: HipDomain, t_end: float, print_step: int, timstart: int = 0, on_print=None,
             max_steps_per_call: int = 2 ** 30):
    """The reference's time loop (main.cpp:271-470) driven through fx_advance:
    one call per print interval."""
    print_next = timstart + print_step
    total = 0
    solver = 0
    while True:
        t, _, _ = dom.clock()
        if t > t_end:
            break
        info = dom.advance(t_end, print_next, max_steps_per_call)
        total += info.steps
        solver += info.solver_steps
        if info.tim >= print_next:
            if on_print:
                on_print(print_next, dom, info)
            print_next = print_next + print_step
        if info.steps == 0:
            break
    return total, solver


class Batch:
    """Many FP32-mixed domains advanced together (fx_batch_*).  Members keep
    their own clock and CFL step; ``run`` enqueues batched steps and ``poll``
    returns (slot, kind, tim, it, steps) for members that halted."""

    def __init__(self, nchem: int = 1, ade: bool = False, max_members: int = 256, device: int = 0):
        lib = load()
        h = C.c_void_p()
        _check(lib.fx_batch_create(device, nchem, 1 if ade else 0, max_members, C.byref(h)))
        self.h = h
        self.max_members = max_members
        self._ev = (BatchEvent * max_members)()
        self.members: dict[int, HipDomain] = {}

    def close(self) -> None:
        """Free the batch; members still in it leave it with their state."""
        if self.h:
            load().fx_batch_destroy(self.h)
            self.h = None
            self.members.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, dom: "HipDomain", t_end: float, print_next: float) -> int:
        slot = C.c_int32()
        _check(load().fx_batch_admit(self.h, dom.h, t_end, print_next, C.byref(slot)))
        self.members[slot.value] = dom
        return slot.value

    def resume(self, slot: int, print_next: float) -> None:
        _check(load().fx_batch_resume(self.h, slot, print_next))

    def evict(self, slot: int) -> "HipDomain":
        _check(load().fx_batch_evict(self.h, slot))
        return self.members.pop(slot)

    def run(self, nsteps: int) -> None:
        _check(load().fx_batch_run(self.h, nsteps))

    def poll(self) -> list[tuple[int, int, float, int, int, int]]:
        """(slot, kind, tim, it, steps, solver_steps) of members that halted."""
        n = C.c_int32()
        _check(load().fx_batch_poll(self.h, self._ev, self.max_members, C.byref(n)))
        return [(e.slot, e.kind, e.tim, e.it, e.steps, e.solver_steps) for e in self._ev[:n.value]]

    def info(self) -> tuple[int, int, int]:
        n, c, u = C.c_int32(), C.c_int64(), C.c_int64()
        _check(load().fx_batch_info(self.h, C.byref(n), C.byref(c), C.byref(u)))
        return n.value, c.value, u.value
